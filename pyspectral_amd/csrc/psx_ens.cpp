// Host side of the ensemble context (include/psx.h, psx_ens_*): plans, item lists, the
// per-phase persistent launches, the captured step and the device time loop for every flow
// of the ensemble at once.  The reference has no ensembles; each flow follows the same
// single-flow semantics (src/pyspectral/backends/cupy_cuda.py:175-195, simulate.py:40-96).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/psx.h"
#include "psx_ens.h"
#include "psx_geo.h"
#include "psx_host.h"
#include "psx_kernels.h"

using namespace psx;

namespace {

constexpr int kSlots = 8;

// Default spectral variant for size n: the first compiled one.
int default_variant(int n) {
  for (int v = 0; v < kEnsNVar; ++v)
    if (ens_compiled(n, v)) return v;
  return -1;
}

Geo2D ens_geo(int n, int var) {
  switch (ens_code(n, var)) {
#define PSX_X(NV, VV)                   \
  case ens_code(NV, VV): {              \
    constexpr Geo2D g = geo2d_ens(NV, VV); \
    return g;                           \
  }
    PSX_ENS_VARIANTS(PSX_X)
#undef PSX_X
    default:
      return Geo2D{};
  }
}

int ens_plan(const psx_flow_params& p, int var, Geo2D* g, psx_layout* lay) {
  if (p.dim != 2 || p.form != 0 || p.fp64 || p.dealias != 0)
    return fail("psx_ens: members are 2D basdevant fp32 flows with the 2/3 rule");
  if (var < 0) var = default_variant(p.n);
  if (var < 0 || !ens_compiled(p.n, var))
    return fail("psx_ens: no compiled ensemble kernels for n=%d variant %d", p.n, var);
  *g = ens_geo(p.n, var);
  if (!g->ok) return fail("psx_ens: no geometry for n=%d variant %d", p.n, var);
  std::memset(lay, 0, sizeof *lay);
  lay->state_pitch = g->sp;
  lay->state_elems = (long long)p.n * g->sp;
  lay->nfields = 2;
  lay->ldw = g->ldw;
  lay->jm = g->jm;
  lay->work_elems = 2LL * p.n * g->ldw;
  lay->four_step = g->n1 > 1;
  lay->e_line = g->lineE;
  lay->e_spec = g->specE;
  lay->e_outer = g->n1 > 1 ? g->outE : 0;
  lay->t_line = g->lineNT;
  lay->t_spec = g->specNT;
  lay->t_outer = g->n1 > 1 ? kEnsNT : 0;
  return 0;
}

void conv_plan(const CtPlan& cp, FftPlan* fp) {
  fp->L = cp.L;
  fp->np = cp.np;
  for (int i = 0; i < kMaxPasses; ++i) fp->rad[i] = i < cp.np ? cp.rad[i] : 1;
}

}  // namespace

struct psx_ens {
  int count = 0;
  std::vector<psx_flow_params> p;
  std::vector<Geo2D> g;
  std::vector<int> var;
  std::vector<EnsFlow> hflows;
  std::vector<void*> tw, twp;
  EnsFlow* dflows = nullptr;
  EnsItem* ditems[3] = {nullptr, nullptr, nullptr};  // line, spectral, outer
  int nitems[3] = {0, 0, 0};
  size_t lds[3] = {0, 0, 0};
  // resident work-groups per CU of each phase kernel: line, spectral stages (pre, 0-3),
  // outer forward, outer inverse (0: unknown)
  int res[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double* prm = nullptr;
  std::vector<long long> prm_off;
  unsigned long long* stats = nullptr;
  int ncu = 256, wg_per_cu = 2;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // device time loop
  DrvMember* dmem = nullptr;
  DrvState* dstate = nullptr;
  double* dring = nullptr;
  double* dpartial = nullptr;
  std::vector<DrvMember> hmem;
  int cap_rec = 0, cap_step = 0;
  int run_mode = 0, run_key = 0;
  hipGraph_t rgraph = nullptr;
  hipGraphExec_t rgexec = nullptr;
  DrvState* hslots = nullptr;
  hipEvent_t slot_ev[kSlots] = {};
  long long nchunk = 0;
  // per-launch event timing
  bool timing = false;
  std::vector<hipEvent_t> events;
  std::vector<int> kinds;
  int nrec = 0;
};

namespace {

int ens_mark(psx_ens* e, int kind, hipStream_t st) {
  if (!e->timing) return 0;
  if (e->nrec == (int)e->events.size()) {
    hipEvent_t ev;
    PSX_HIP(hipEventCreate(&ev));
    e->events.push_back(ev);
    e->kinds.push_back(-1);
  }
  PSX_HIP(hipEventRecord(e->events[e->nrec], st));
  e->kinds[e->nrec] = kind;
  ++e->nrec;
  return 0;
}

#define ENS_MARK(kind)                 \
  do {                                 \
    int r_ = ens_mark(e, (kind), st);  \
    if (r_) return r_;                 \
  } while (0)

int res_slot(int kind, int arg) {
  return kind == 0 ? 0 : (kind == 1 ? 2 + arg : (kind == 2 ? 6 : 7));  // kStagePre = -1
}

// Persistent grid of a phase: wg_per_cu work-groups per CU, at most as many as can be resident
// at once (a work-group that cannot start until another ends would run its items at the end,
// on an otherwise idle GPU), and no more than there are items.
int grid_of(const psx_ens* e, int kind, int arg) {
  const int ph = kind == 0 ? 0 : (kind == 1 ? 1 : 2);
  const int res = e->res[res_slot(kind, arg)];
  const int per_cu = res > 0 ? std::min(e->wg_per_cu, res) : e->wg_per_cu;
  return std::max(1, std::min(e->nitems[ph], per_cu * e->ncu));
}

// Item order of a column phase: in each block of 64 items, the eight runs of eight consecutive
// items (neighbouring column tiles of one member) are interleaved, so that the items of a run
// take positions equal mod 8.  Work-group b of a persistent launch starts on XCD b mod 8 and
// its items are b, b + grid, ... with the grid a multiple of 8 (or the whole list), so a run's
// tiles are processed at about the same time on one XCD and the 128-byte lines that
// neighbouring tiles share (tiles of 4 or 8 columns, rows not aligned to 128 bytes) are
// fetched into that XCD's L2 once instead of once per tile.
void xcd_runs(std::vector<EnsItem>& v) {
  const size_t full = v.size() / 64 * 64;
  std::vector<EnsItem> out(v);
  for (size_t b = 0; b < full; b += 64)
    for (int r = 0; r < 8; ++r)
      for (int s = 0; s < 8; ++s) out[b + 8 * s + r] = v[b + 8 * r + s];
  v.swap(out);
}

int ens_phase(psx_ens* e, int kind, int arg, hipStream_t st) {
  const int ph = kind == 0 ? 0 : (kind == 1 ? 1 : 2);
  if (e->nitems[ph] == 0) return 0;
  PSX_HIP(launch_ens(kind, arg, e->dflows, e->ditems[ph], e->nitems[ph], grid_of(e, kind, arg),
                     e->lds[ph], st));
  return 0;
}

int ens_prepare(psx_ens* e, hipStream_t st) {
  int r = ens_phase(e, 1, kStagePre, st);
  if (!r) r = ens_phase(e, 3, 0, st);
  return r;
}

// One IF-RK4 step of every flow; drv as enqueue_step of the single-flow context.
int ens_step(psx_ens* e, hipStream_t st, int drv) {
  for (int s = 0; s < 4; ++s) {
    int r;
    ENS_MARK(PSX_KIND_LINE);
    if ((r = ens_phase(e, 0, (s == 0 && drv == 2) ? kLineWantMax : 0, st))) return r;
    if (s == 0 && drv) {
      ENS_MARK(PSX_KIND_CTRL);
      PSX_HIP(launch_drv(0, e->dmem, e->count, st));
    }
    if (e->nitems[2]) {
      ENS_MARK(PSX_KIND_OUTER_FWD);
      if ((r = ens_phase(e, 2, 0, st))) return r;
    }
    ENS_MARK(PSX_KIND_SPEC);
    if ((r = ens_phase(e, 1, s, st))) return r;
    if (e->nitems[2]) {
      ENS_MARK(PSX_KIND_OUTER_INV);
      if ((r = ens_phase(e, 3, 0, st))) return r;
    }
  }
  ENS_MARK(PSX_KIND_CTRL);
  {
    int nmax = 0;
    for (const psx_flow_params& p : e->p) nmax = std::max(nmax, p.n);
    PSX_HIP(launch_ens_hermitian(e->dflows, e->count, nmax, st));
  }
  if (drv) {
    PSX_HIP(launch_drv(1, e->dmem, e->count, st));
    PSX_HIP(launch_drv(2, e->dmem, e->count, st));
  }
  ENS_MARK(-1);
  return 0;
}

// Capture nsteps steps of mode drv on the context's capture stream.
int ens_capture(psx_ens* e, int nsteps, int drv, hipGraph_t* graph, hipGraphExec_t* exec) {
  if (!e->cap) PSX_HIP(hipStreamCreateWithFlags(&e->cap, hipStreamNonBlocking));
  const bool was_timing = e->timing;
  e->timing = false;
  PSX_HIP(hipStreamBeginCapture(e->cap, hipStreamCaptureModeRelaxed));
  int r = 0;
  for (int k = 0; k < nsteps && !r; ++k) r = ens_step(e, e->cap, drv);
  hipGraph_t g = nullptr;
  hipError_t ce = hipStreamEndCapture(e->cap, &g);
  e->timing = was_timing;
  if (r) {
    if (g) (void)hipGraphDestroy(g);
    return r;
  }
  if (ce != hipSuccess) return fail("psx_ens: capture failed: %s", hipGetErrorString(ce));
  *graph = g;
  PSX_HIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  return 0;
}

void drop_graph(hipGraph_t* g, hipGraphExec_t* x) {
  if (*x) (void)hipGraphExecDestroy(*x);
  if (*g) (void)hipGraphDestroy(*g);
  *x = nullptr;
  *g = nullptr;
}

}  // namespace

extern "C" {

int psx_ens_variants(int n, int* out, int cap) {
  int k = 0;
  for (int v = 0; v < kEnsNVar; ++v)
    if (ens_compiled(n, v)) {
      if (out && k < cap) out[k] = v;
      ++k;
    }
  return k;
}

int psx_ens_query_layout(const psx_flow_params* p, int variant, psx_layout* out) {
  if (!p || !out) return fail("psx_ens_query_layout: null argument");
  Geo2D g;
  return ens_plan(*p, variant, &g, out);
}

int psx_ens_create(int count, const psx_flow_params* params, const int* variants,
                   void* const* u0, void* const* acc, void* const* work, psx_ens** out) {
  if (count < 1 || !params || !u0 || !acc || !work || !out)
    return fail("psx_ens_create: bad arguments");
  psx_ens* e = new psx_ens;
  e->count = count;
  e->p.assign(params, params + count);
  e->g.resize(count);
  e->var.resize(count);
  e->hflows.resize(count);
  e->tw.assign(count, nullptr);
  e->twp.assign(count, nullptr);
  e->prm_off.resize(count);
  auto bail = [&](int r) {
    psx_ens_destroy(e);
    return r;
  };
  long long nprm = 0;
  for (int m = 0; m < count; ++m) {
    psx_layout lay;
    const int v = variants && variants[m] >= 0 ? variants[m] : default_variant(params[m].n);
    if (int r = ens_plan(params[m], v, &e->g[m], &lay)) return bail(r);
    if (!u0[m] || !acc[m] || !work[m]) return bail(fail("psx_ens_create: null buffer %d", m));
    e->var[m] = v;
    e->prm_off[m] = nprm;
    nprm += 2 * (params[m].n / 2 + 1) + PrmLayout::kNScalars;
  }
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return bail(fail("psx_ens_create: no device"));
  e->ncu = prop.multiProcessorCount;
  if (hipMalloc(&e->prm, nprm * sizeof(double)) != hipSuccess ||
      hipMalloc(&e->stats, (size_t)count * 4 * sizeof(unsigned long long)) != hipSuccess)
    return bail(fail("psx_ens_create: hipMalloc failed"));
  std::vector<EnsItem> items[3];
  for (int m = 0; m < count; ++m) {
    const Geo2D& g = e->g[m];
    const int n = g.n, nh = n / 2 + 1;
    Flow<float>& f = e->hflows[m].f;
    std::memset(&f, 0, sizeof f);
    f.n = n;
    f.nh = nh;
    f.jm = g.jm;
    f.dim = 2;
    f.rot = 0;
    f.scale = (float)(1.0 / ((double)n * n));
    f.colT = f.yT = g.colT;
    f.ldw = g.ldw;
    f.sp = g.sp;
    f.fstride = (long long)n * g.ldw;
    f.sstride = (long long)n * g.sp;
    f.rowT = g.lineT;
    f.n1 = g.n1;
    f.n2 = g.n2;
    conv_plan(g.line, &f.prow);
    conv_plan(g.spec, &f.pcol);
    if (g.n1 > 1) conv_plan(g.outer, &f.pa);
    if (int r = make_twiddles<float>(n, &e->tw[m])) return bail(r);
    if (int r = make_pass_twiddles<float>(&f, &e->twp[m])) return bail(r);
    f.W = static_cast<cx<float>*>(work[m]);
    f.u0 = static_cast<cx<float>*>(u0[m]);
    f.acc = static_cast<cx<float>*>(acc[m]);
    f.tw = static_cast<const cx<float>*>(e->tw[m]);
    f.twp = static_cast<const cx<float>*>(e->twp[m]);
    f.prm = e->prm + e->prm_off[m];
    f.stats = e->stats + 4 * m;
    e->hflows[m].code = ens_code(n, e->var[m]);
    const int njt = g.ldw / g.colT;
    for (int t = 0; t < n / g.lineT; ++t) items[0].push_back({m, t});
    for (int t = 0; t < g.n1 * njt; ++t) items[1].push_back({m, t});
    if (g.n1 > 1)
      for (int t = 0; t < g.n2 * njt / g.outS; ++t) items[2].push_back({m, t});
    // row or column tiles plus the pass-twiddle copies the bodies keep in LDS
    const size_t line_tw = (size_t)ct_twoff(g.line, false, g.line.np - 1) +
                           ct_twoff(g.line, true, g.line.np - 1);
    const size_t line_lds = std::max(((size_t)g.lineT * ct_rows_pitch(n) + line_tw) * 8,
                                     (size_t)2 * kEnsNT * sizeof(double));
    e->lds[0] = std::max(e->lds[0], line_lds);
    e->lds[1] = std::max(e->lds[1], ((size_t)g.n2 * g.colT + ct_twsize(g.spec, false) +
                                     ct_twsize(g.spec, true) + (g.n1 > 1 ? g.n2 : 0)) * 8 +
                                        (size_t)2 * (g.n2 + g.colT) * sizeof(double));
    if (g.n1 > 1) e->lds[2] = std::max(e->lds[2], (size_t)g.outS * g.n1 * g.colT * 8);
    if (launch_params(e->prm + e->prm_off[m], nh, params[m].nu, 0.0, params[m].alpha, 0.0, params[m].klo2,
                      params[m].khi2, params[m].ka2, nullptr) != hipSuccess)
      return bail(fail("psx_ens_create: parameter kernel failed"));
  }
  if (hipMalloc(&e->dflows, count * sizeof(EnsFlow)) != hipSuccess)
    return bail(fail("psx_ens_create: hipMalloc failed"));
  if (hipMemcpy(e->dflows, e->hflows.data(), count * sizeof(EnsFlow), hipMemcpyHostToDevice) !=
      hipSuccess)
    return bail(fail("psx_ens_create: copy failed"));
  xcd_runs(items[1]);
  xcd_runs(items[2]);
  {
    e->res[0] = ens_resident(0, 0, e->lds[0]);
    for (int st = kStagePre; st <= 3; ++st)
      e->res[res_slot(1, st)] = ens_resident(1, st, e->lds[1]);
    if (!items[2].empty()) {
      e->res[6] = ens_resident(2, 0, e->lds[2]);
      e->res[7] = ens_resident(3, 0, e->lds[2]);
    }
  }
  for (int ph = 0; ph < 3; ++ph) {
    e->nitems[ph] = (int)items[ph].size();
    if (items[ph].empty()) continue;
    if (hipMalloc(&e->ditems[ph], items[ph].size() * sizeof(EnsItem)) != hipSuccess ||
        hipMemcpy(e->ditems[ph], items[ph].data(), items[ph].size() * sizeof(EnsItem),
                  hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail("psx_ens_create: item lists"));
  }
  if (hipDeviceSynchronize() != hipSuccess) return bail(fail("psx_ens_create: sync failed"));
  *out = e;
  return 0;
}

int psx_ens_destroy(psx_ens* e) {
  if (!e) return 0;
  for (hipEvent_t ev : e->events) (void)hipEventDestroy(ev);
  for (hipEvent_t ev : e->slot_ev)
    if (ev) (void)hipEventDestroy(ev);
  drop_graph(&e->graph, &e->gexec);
  drop_graph(&e->rgraph, &e->rgexec);
  if (e->cap) (void)hipStreamDestroy(e->cap);
  for (void* ptr : e->tw)
    if (ptr) (void)hipFree(ptr);
  for (void* ptr : e->twp)
    if (ptr) (void)hipFree(ptr);
  for (EnsItem* it : e->ditems)
    if (it) (void)hipFree(it);
  if (e->dflows) (void)hipFree(e->dflows);
  if (e->prm) (void)hipFree(e->prm);
  if (e->stats) (void)hipFree(e->stats);
  if (e->dmem) (void)hipFree(e->dmem);
  if (e->dstate) (void)hipFree(e->dstate);
  if (e->dring) (void)hipFree(e->dring);
  if (e->dpartial) (void)hipFree(e->dpartial);
  if (e->hslots) (void)hipHostFree(e->hslots);
  delete e;
  return 0;
}

int psx_ens_set_grid(psx_ens* e, int wg_per_cu) {
  if (!e) return fail("psx_ens_set_grid: null context");
  if (wg_per_cu < 1 || wg_per_cu > 16) return fail("psx_ens_set_grid: %d outside [1, 16]", wg_per_cu);
  if (wg_per_cu != e->wg_per_cu) {
    e->wg_per_cu = wg_per_cu;
    drop_graph(&e->graph, &e->gexec);
    drop_graph(&e->rgraph, &e->rgexec);
  }
  return 0;
}

int psx_ens_prepare(psx_ens* e, void* stream) {
  if (!e) return fail("psx_ens_prepare: null context");
  return ens_prepare(e, static_cast<hipStream_t>(stream));
}

int psx_ens_set_step(psx_ens* e, const double* dt, const double* amp, void* stream) {
  if (!e || !dt || !amp) return fail("psx_ens_set_step: null argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  for (int m = 0; m < e->count; ++m) {
    if (!std::isfinite(dt[m]) || !std::isfinite(amp[m]))
      return fail("psx_ens_set_step: non-finite dt/amp for member %d", m);
    const psx_flow_params& p = e->p[m];
    PSX_HIP(launch_params(e->prm + e->prm_off[m], p.n / 2 + 1, p.nu, dt[m], p.alpha, amp[m],
                          p.klo2, p.khi2, p.ka2, st));
  }
  return 0;
}

int psx_ens_step(psx_ens* e, void* stream) {
  if (!e) return fail("psx_ens_step: null context");
  return ens_step(e, static_cast<hipStream_t>(stream), 0);
}

int psx_ens_step_graph(psx_ens* e, void* stream) {
  if (!e) return fail("psx_ens_step_graph: null context");
  if (!e->gexec)
    if (int r = ens_capture(e, 1, 0, &e->graph, &e->gexec)) return r;
  PSX_HIP(hipGraphLaunch(e->gexec, static_cast<hipStream_t>(stream)));
  return 0;
}

int psx_ens_timing_begin(psx_ens* e) {
  if (!e) return fail("psx_ens_timing_begin: null context");
  e->timing = true;
  e->nrec = 0;
  return 0;
}

int psx_ens_timing_end(psx_ens* e, void* stream, double* ms, int* count) {
  if (!e || !ms || !count) return fail("psx_ens_timing_end: null argument");
  e->timing = false;
  PSX_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  for (int k = 0; k < PSX_NKINDS; ++k) {
    ms[k] = 0.0;
    count[k] = 0;
  }
  for (int i = 0; i + 1 < e->nrec; ++i) {
    const int k = e->kinds[i];
    if (k < 0 || k >= PSX_NKINDS) continue;
    float t = 0.f;
    PSX_HIP(hipEventElapsedTime(&t, e->events[i], e->events[i + 1]));
    ms[k] += t;
    count[k] += 1;
  }
  e->nrec = 0;
  return 0;
}

// ---------------------------------------------------------------------- device time loop
int psx_ens_run_begin(psx_ens* e, const psx_run_params* rp, void* stream) {
  if (!e || !rp) return fail("psx_ens_run_begin: null argument");
  const int c = e->count;
  const psx_run_params& r0 = rp[0];
  if (r0.cap_rec < 1 || r0.cap_step < 1) return fail("psx_ens_run_begin: ring capacities < 1");
  for (int m = 0; m < c; ++m) {
    if (rp[m].cfl_every < 1) return fail("psx_ens_run_begin: cfl_every < 1 (member %d)", m);
    if (rp[m].cap_rec != r0.cap_rec || rp[m].cap_step != r0.cap_step)
      return fail("psx_ens_run_begin: ring capacities differ between members");
    if (!(rp[m].dt_fixed > 0.0) && !(rp[m].cfl > 0.0 && rp[m].dt_max > 0.0))
      return fail("psx_ens_run_begin: member %d needs dt_fixed, or cfl and dt_max", m);
    if ((rp[m].dt_fixed > 0.0) != (r0.dt_fixed > 0.0))
      return fail("psx_ens_run_begin: all members CFL-adaptive or all fixed-dt");
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (e->cap_rec != r0.cap_rec || e->cap_step != r0.cap_step) {
    if (e->dring) (void)hipFree(e->dring);
    e->dring = nullptr;
    PSX_HIP(hipMalloc(&e->dring, (size_t)c * (3 * r0.cap_rec + 2 * r0.cap_step) * sizeof(double)));
    e->cap_rec = r0.cap_rec;
    e->cap_step = r0.cap_step;
  }
  if (!e->dmem) {
    PSX_HIP(hipMalloc(&e->dmem, c * sizeof(DrvMember)));
    PSX_HIP(hipMalloc(&e->dstate, c * sizeof(DrvState)));
    PSX_HIP(hipMalloc(&e->dpartial, (size_t)c * 2 * kDrvRedBlocks * sizeof(double)));
    PSX_HIP(hipHostMalloc(&e->hslots, (size_t)kSlots * c * sizeof(DrvState)));
    for (hipEvent_t& ev : e->slot_ev) PSX_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  e->nchunk = 0;
  e->hmem.resize(c);
  std::vector<DrvState> hs(c);
  for (int m = 0; m < c; ++m) {
    const psx_flow_params& p = e->p[m];
    const Flow<float>& f = e->hflows[m].f;
    DrvMember& dm = e->hmem[m];
    std::memset(&dm, 0, sizeof dm);
    dm.u0 = f.u0;
    dm.prm = e->prm + e->prm_off[m];
    dm.stats = f.stats;
    dm.s = e->dstate + m;
    dm.rec = e->dring + (size_t)m * 3 * e->cap_rec;
    dm.steps = e->dring + (size_t)c * 3 * e->cap_rec + (size_t)m * 2 * e->cap_step;
    dm.partial = e->dpartial + (size_t)m * 2 * kDrvRedBlocks;
    dm.nu = p.nu;
    dm.alpha = p.alpha;
    dm.klo2 = p.klo2;
    dm.khi2 = p.khi2;
    dm.ka2 = p.ka2;
    dm.n = p.n;
    dm.dim = 2;
    dm.sp = f.sp;
    dm.fp64 = 0;
    dm.cap_rec = e->cap_rec;
    dm.cap_step = e->cap_step;
    hs[m] = drv_initial_state(rp[m], p.n);
  }
  e->run_mode = r0.dt_fixed > 0.0 ? 1 : 2;
  // members over before their first step are marked inactive by the set-active kernel
  std::vector<DrvMember> idle;
  for (int m = 0; m < c; ++m)
    if (!hs[m].active) idle.push_back(e->hmem[m]);
  DrvMember* didle = nullptr;
  if (!idle.empty()) PSX_HIP(hipMalloc(&didle, idle.size() * sizeof(DrvMember)));
  // The copies read hs and idle in this frame asynchronously: the stream is drained before
  // returning, on the error paths too.
  auto enqueue = [&]() -> int {
    PSX_HIP(hipMemcpyAsync(e->dmem, e->hmem.data(), c * sizeof(DrvMember),
                           hipMemcpyHostToDevice, st));
    PSX_HIP(hipMemcpyAsync(e->dstate, hs.data(), c * sizeof(DrvState), hipMemcpyHostToDevice,
                           st));
    PSX_HIP(launch_fill_u32(reinterpret_cast<unsigned*>(e->stats), 0u, 8 * c, st));
    for (int m = 0; m < c; ++m) {
      const psx_flow_params& p = e->p[m];
      PSX_HIP(launch_params(e->prm + e->prm_off[m], p.n / 2 + 1, p.nu, 0.0, p.alpha,
                            rp[m].amp0, p.klo2, p.khi2, p.ka2, st));
    }
    if (didle) {
      PSX_HIP(hipMemcpyAsync(didle, idle.data(), idle.size() * sizeof(DrvMember),
                             hipMemcpyHostToDevice, st));
      PSX_HIP(launch_drv(4, didle, (int)idle.size(), st));
    }
    PSX_HIP(launch_drv(2, e->dmem, c, st));
    return 0;
  };
  const int r = enqueue();
  const hipError_t es = hipStreamSynchronize(st);
  if (didle) (void)hipFree(didle);
  if (r) {
    e->run_mode = 0;
    return r;
  }
  if (es != hipSuccess) {
    e->run_mode = 0;
    return fail("psx_ens_run_begin: %s", hipGetErrorString(es));
  }
  return 0;
}

int psx_ens_run_chunk(psx_ens* e, int nsteps, int use_graph, void* stream) {
  if (!e) return fail("psx_ens_run_chunk: null context");
  if (!e->dmem || !e->run_mode) return fail("psx_ens_run_chunk: no run begun");
  if (nsteps < 1) return fail("psx_ens_run_chunk: nsteps must be >= 1");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int slot = (int)(e->nchunk % kSlots);
  PSX_HIP(hipEventSynchronize(e->slot_ev[slot]));
  if (use_graph) {
    const int key = nsteps * 4 + e->run_mode;
    if (!e->rgexec || e->run_key != key) {
      drop_graph(&e->rgraph, &e->rgexec);
      if (int r = ens_capture(e, nsteps, e->run_mode, &e->rgraph, &e->rgexec)) return r;
      e->run_key = key;
    }
    PSX_HIP(hipGraphLaunch(e->rgexec, st));
  } else {
    for (int k = 0; k < nsteps; ++k)
      if (int r = ens_step(e, st, e->run_mode)) return r;
  }
  PSX_HIP(hipMemcpyAsync(e->hslots + (size_t)slot * e->count, e->dstate,
                         e->count * sizeof(DrvState), hipMemcpyDeviceToHost, st));
  PSX_HIP(hipEventRecord(e->slot_ev[slot], st));
  ++e->nchunk;
  return 0;
}

int psx_ens_run_peek(psx_ens* e, int lag, psx_run_state* out) {
  if (!e || !out) return fail("psx_ens_run_peek: null argument");
  if (lag < 0 || lag >= kSlots || lag >= e->nchunk)
    return fail("psx_ens_run_peek: lag %d outside [0, min(%d, %lld))", lag, kSlots, e->nchunk);
  const int slot = (int)((e->nchunk - 1 - lag) % kSlots);
  PSX_HIP(hipEventSynchronize(e->slot_ev[slot]));
  for (int m = 0; m < e->count; ++m) drv_export(e->hslots[(size_t)slot * e->count + m], out + m);
  return 0;
}

int psx_ens_run_poll(psx_ens* e, psx_run_state* out, double* rec, double* steps, void* stream) {
  if (!e || !out) return fail("psx_ens_run_poll: null argument");
  if (!e->dmem) return fail("psx_ens_run_poll: no run begun");
  hipStream_t st = static_cast<hipStream_t>(stream);
  std::vector<DrvState> hs(e->count);
  PSX_HIP(hipMemcpyAsync(hs.data(), e->dstate, e->count * sizeof(DrvState), hipMemcpyDeviceToHost, st));
  const size_t nrecd = (size_t)e->count * 3 * e->cap_rec;
  if (rec) PSX_HIP(hipMemcpyAsync(rec, e->dring, nrecd * sizeof(double), hipMemcpyDeviceToHost, st));
  if (steps)
    PSX_HIP(hipMemcpyAsync(steps, e->dring + nrecd,
                           (size_t)e->count * 2 * e->cap_step * sizeof(double),
                           hipMemcpyDeviceToHost, st));
  PSX_HIP(hipStreamSynchronize(st));
  for (int m = 0; m < e->count; ++m) drv_export(hs[m], out + m);
  return 0;
}

int psx_ens_run_end(psx_ens* e, void* stream) {
  if (!e) return fail("psx_ens_run_end: null context");
  if (!e->dmem) return 0;
  PSX_HIP(launch_drv(3, e->dmem, e->count, static_cast<hipStream_t>(stream)));
  e->run_mode = 0;
  return 0;
}

}  // extern "C"
