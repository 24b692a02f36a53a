// Host side of libpsx: planner, context, C ABI (include/psx.h).
#include "../../include/psx.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "psx_geo.h"
#include "psx_host.h"
#include "psx_les.h"
#include "psx_kernels.h"

using namespace psx;

namespace {

// ------------------------------------------------------------------------------- planning
struct Factors {
  int a = 0, b3 = 0, b5 = 0;
  bool ok = false;
};

Factors factor(int L) {
  Factors f;
  int l = L;
  while (l % 2 == 0) { l /= 2; ++f.a; }
  while (l % 3 == 0) { l /= 3; ++f.b3; }
  while (l % 5 == 0) { l /= 5; ++f.b5; }
  f.ok = (l == 1);
  return f;
}

// Radix list for length L with every radix dividing E; balanced power-of-two radices (<= 32)
// first, then 3s and 5s.
bool plan_fft(int L, int E, FftPlan* p) {
  Factors f = factor(L);
  if (!f.ok) return false;
  if (f.b3 && E % 3) return false;
  if (f.b5 && E % 5) return false;
  int lmax = 0;
  while (lmax < 5 && E % (1 << (lmax + 1)) == 0) ++lmax;
  if (f.a > 0 && lmax == 0) return false;
  const int np2 = f.a == 0 ? 0 : (f.a + lmax - 1) / lmax;
  const int np = np2 + f.b3 + f.b5;
  if (np > kMaxPasses || np == 0) return false;
  p->L = L;
  p->np = np;
  int k = 0;
  for (int i = 0; i < np2; ++i) {
    const int bits = f.a / np2 + (i < f.a % np2 ? 1 : 0);
    p->rad[k++] = 1 << bits;
  }
  for (int i = 0; i < f.b3; ++i) p->rad[k++] = 3;
  for (int i = 0; i < f.b5; ++i) p->rad[k++] = 5;
  for (; k < kMaxPasses; ++k) p->rad[k] = 1;
  return true;
}

int pow2_floor(int x) {
  int p = 1;
  while (p * 2 <= x) p *= 2;
  return p;
}

// Elements per thread for a transform kernel over length L in the given precision.
int pick_E(int L, bool fp64, bool heavy) {
  if (L % 5 == 0) return 20;  // lengths 5 * 2^a: one radix-5 pass
  const bool r3 = (L % 3 == 0);
  if (r3) return (fp64 || heavy) ? 12 : 24;
  if (fp64) return heavy ? 8 : 16;
  return 16;
}

// Row kernels: T whole rows of length L per work-group.
bool plan_rows(int L, int nrows, bool fp64, bool heavy, int nmax_vals, Launch* lc, int* T,
               FftPlan* p) {
  int E = pick_E(L, fp64, heavy);
  if (!plan_fft(L, E, p)) return false;
  int t = 1;
  while (L * t / E < 128 && nrows % (t * 2) == 0) t *= 2;  // t divides nrows
  int nt = L * t / E;
  if (nt > 512) {
    if (E == 16 && !fp64 && plan_fft(L, 32, p)) {
      E = 32;
      nt = L * t / E;
    }
    if (nt > 512) return false;
  }
  if ((L * t) % E) return false;
  *T = t;
  lc->E = E;
  lc->threads = nt;
  lc->blocks = (nrows + t - 1) / t;
  const size_t csz = fp64 ? 16 : 8;
  size_t lds = (size_t)t * rows_pitch(L) * csz;
  const size_t red = (size_t)nmax_vals * nt * sizeof(double);
  lc->lds = lds > red ? lds : red;
  return true;
}

// Column kernels: tiles of L rows x T columns.
bool plan_cols(int L, int T, int ngroups, int ldw, bool fp64, bool heavy, Launch* lc,
               FftPlan* p) {
  int E = pick_E(L, fp64, heavy);
  while (E > 4 && (L * T) % E) E /= 2;
  if (!plan_fft(L, E, p)) return false;
  const int nt = L * T / E;
  if (nt > 512 || nt < 1 || (L * T) % E) return false;
  lc->E = E;
  lc->threads = nt;
  lc->blocks = ngroups * (ldw / T);
  lc->lds = (size_t)L * T * (fp64 ? 16 : 8);
  return true;
}

}  // namespace

struct psx_ctx {
  psx_flow_params p{};
  psx_layout lay{};
  bool fp64 = false;
  Flow<float> f32{};
  Flow<double> f64{};
  Launch line{}, spec{}, outerF{}, outerI{}, yF{}, yI{};
  void* tw = nullptr;
  void* twp = nullptr;
  double* prm = nullptr;
  unsigned long long* stats = nullptr;
  double* partial = nullptr;
  double* dout = nullptr;
  int nred = 1024;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int nshell_max = 0;
  // device controller (psx_run_*): one member, its state, rings and chunk graphs
  DrvMember* dmem = nullptr;
  DrvState* dstate = nullptr;
  double* dring = nullptr;       // records (3 cap_rec) then steps (2 cap_step)
  double* dpartial = nullptr;
  DrvMember hmem{};
  int run_chunk = 0, run_mode = 0;
  // controller state copied to pinned host slots after each chunk (psx_run_peek)
  static constexpr int kSlots = 8;
  DrvState* hslots = nullptr;
  hipEvent_t slot_ev[kSlots] = {};
  long long nchunk = 0;
  // LES subgrid-stress model evaluated at the start of every step (psx_set_sgs)
  psx_sgs_desc sgs{};
  bool has_sgs = false;
  hipGraph_t rgraph = nullptr;
  hipGraphExec_t rgexec = nullptr;
  bool ct = false;  // compile-time specialised fp32 kernels for this size (2D or 3D)
  // per-launch event timing of eager steps (psx_timing_*)
  bool timing = false;
  std::vector<hipEvent_t> events;
  std::vector<int> kinds;
  int nrec = 0;
};

namespace {

bool supported_n(int n) {
  if (n < 8 || n % 2) return false;
  Factors f = factor(n);
  return f.ok && f.b3 + f.b5 <= 1 && f.a >= 2;
}

// Fill the flow geometry and launch shapes (no device allocation).
template <typename T>
int plan_flow(const psx_flow_params& p, psx_ctx* c, Flow<T>* f) {
  const bool fp64 = sizeof(T) == 8;
  const int n = p.n;
  if (!supported_n(n))
    return fail("psx: n=%d unsupported (even, >= 8, of the form 2^a, 3*2^a or 5*2^a, a >= 2)", n);
  if (p.dim != 2 && p.dim != 3) return fail("psx: dim must be 2 or 3");
  if (p.dim == 2 && p.form != 0) return fail("psx: 2D supports the basdevant form only");
  if (p.form != 0 && p.form != 1) return fail("psx: form must be 0 or 1");
  if (p.dealias != 0 && p.dealias != 1)
    return fail("psx: dealias must be 0 (2/3 rule) or 1 (phase shift)");
  std::memset(f, 0, sizeof *f);
  f->n = n;
  f->nh = n / 2 + 1;
  f->ps = p.dealias == 1;
  // retained half-axis columns: 3 j < n, or 9 j^2 < 2 n^2 for the phase shift's disc/ball
  f->jm = f->ps ? ct_retained_ps(n) : ct_retained(n);
  f->dim = p.dim;
  f->rot = p.dim == 3 && p.form == 1;
  f->scale = (T)(1.0 / std::pow((double)n, p.dim));
  const int colT = fp64 ? 8 : 16;
  const int nrows = p.dim == 2 ? n : n * n;
  const size_t csz = sizeof(T) * 2;
  int T_ = colT;
  while (T_ > 1 && T_ / 2 >= f->jm) T_ /= 2;  // tiny grids: do not pad far beyond jm
  f->colT = T_;
  f->yT = T_;
  f->ldw = (f->jm + T_ - 1) / T_ * T_;
  int nfields = 2;
  if (p.dim == 3) nfields = f->rot ? 6 : 5;
  f->fstride = (long long)nrows * f->ldw;
  f->sp = ct_state_pitch(n, fp64);
  f->sstride = (long long)nrows * f->sp;

  const int nmax = p.dim == 2 ? 2 : 3;
  if (p.dim == 2) {
    // one planner for the run-time and the compile-time specialised kernels
    const Geo2D g = geo2d(n, fp64, f->jm);
    if (!g.ok) return fail("psx: no 2D plan for n=%d", n);
    auto conv = [](const CtPlan& cp, FftPlan* fp) {
      fp->L = cp.L;
      fp->np = cp.np;
      for (int i = 0; i < kMaxPasses; ++i) fp->rad[i] = i < cp.np ? cp.rad[i] : 1;
    };
    f->colT = f->yT = g.colT;
    f->ldw = g.ldw;
    f->fstride = (long long)nrows * f->ldw;
    f->rowT = g.lineT;
    f->n1 = g.n1;
    f->n2 = g.n2;
    conv(g.line, &f->prow);
    conv(g.spec, &f->pcol);
    if (g.n1 > 1) conv(g.outer, &f->pa);
    if (n % g.lineT) return fail("psx: line tile of %d rows does not divide n=%d", g.lineT, n);
    c->line.E = g.lineE;
    c->line.threads = g.lineNT;
    c->line.blocks = n / g.lineT;
    const size_t lds = (size_t)g.lineT * rows_pitch(n) * csz;
    const size_t red = (size_t)nmax * g.lineNT * sizeof(double);
    c->line.lds = lds > red ? lds : red;
    const int njt = g.ldw / g.colT;
    c->spec.E = g.specE;
    c->spec.threads = g.specNT;
    c->spec.blocks = g.n1 * njt;
    c->spec.lds = (size_t)g.n2 * g.colT * csz;
    if (g.n1 > 1) {
      c->outerF.E = g.outE;
      c->outerF.threads = g.outNT;
      c->outerF.blocks = g.n2 * njt;
      c->outerF.lds = (size_t)g.n1 * g.colT * csz;
      c->outerI = c->outerF;
    }
    // the phase shift runs on the run-time kernels (the specialised ones assume the 2/3 rule)
    c->ct = !fp64 && !f->ps && ct2d_supported(n);
  } else if (!fp64 && !f->ps && ct3d_supported(n)) {
    // compile-time specialised 3D kernels: plans from the shared planner
    const Geo3D g = geo3d(n);
    if (!g.ok) return fail("psx: no 3D plan for n=%d", n);
    auto conv = [](const CtPlan& cp, FftPlan* fp) {
      fp->L = cp.L;
      fp->np = cp.np;
      for (int i = 0; i < kMaxPasses; ++i) fp->rad[i] = i < cp.np ? cp.rad[i] : 1;
    };
    f->colT = f->yT = g.T;
    f->ldw = g.ldw;
    f->fstride = (long long)nrows * f->ldw;
    f->rowT = g.lineT;
    f->n1 = 1;
    f->n2 = n;
    conv(g.line, &f->prow);
    conv(g.yp, &f->py);
    conv(g.xp, &f->pcol);
    const int njt = g.ldw / g.T;
    if (nrows % g.lineT) return fail("psx: line tile of %d rows does not divide n^2", g.lineT);
    c->line.E = g.lineE;
    c->line.threads = g.lineNT;
    c->line.blocks = nrows / g.lineT;
    c->line.lds = (size_t)g.lineT * rows_pitch(n) * csz;
    c->yF.E = g.yE;
    c->yF.threads = g.yNT;
    c->yF.blocks = n * njt;
    c->yF.lds = (size_t)n * g.T * csz;
    c->yI = c->yF;
    c->spec.E = g.xE;
    c->spec.threads = g.xNT;
    c->spec.blocks = n * njt;
    c->spec.lds = (size_t)n * g.T * csz;
    c->ct = true;
  } else {
    if (!plan_rows(n, nrows, fp64, true, nmax, &c->line, &f->rowT, &f->prow))
      return fail("psx: no line plan for n=%d", n);
    if ((size_t)n * T_ * csz > 98304) return fail("psx: 3D n=%d too large for one tile", n);
    f->n1 = 1;
    f->n2 = n;
    if (!plan_cols(n, T_, n, f->ldw, fp64, true, &c->spec, &f->pcol))
      return fail("psx: no x plan for n=%d", n);
    if (!plan_cols(n, T_, n, f->ldw, fp64, true, &c->yF, &f->py))
      return fail("psx: no y plan for n=%d", n);
    c->yI = c->yF;
  }
  c->lay.state_elems = (long long)(p.dim == 2 ? 1 : 3) * f->sstride;
  c->lay.state_pitch = f->sp;
  c->lay.work_elems = (long long)nfields * f->fstride;
  c->lay.nfields = nfields;
  c->lay.ldw = f->ldw;
  c->lay.jm = f->jm;
  c->lay.four_step = f->n1 > 1;
  c->lay.e_line = c->line.E;
  c->lay.e_spec = c->spec.E;
  c->lay.e_outer = f->n1 > 1 ? c->outerF.E : 0;
  c->lay.e_mid = p.dim == 3 ? c->yF.E : 0;
  c->lay.t_line = c->line.threads;
  c->lay.t_spec = c->spec.threads;
  c->lay.t_outer = f->n1 > 1 ? c->outerF.threads : 0;
  c->lay.t_mid = p.dim == 3 ? c->yF.threads : 0;
  return 0;
}

template <typename T>
Flow<T>& flow(psx_ctx* c);
template <>
Flow<float>& flow<float>(psx_ctx* c) { return c->f32; }
template <>
Flow<double>& flow<double>(psx_ctx* c) { return c->f64; }

// Record a timing event before a launch of the given kind (-1: end of a step).
int mark(psx_ctx* c, int kind, hipStream_t st) {
  if (!c->timing) return 0;
  if (c->nrec == (int)c->events.size()) {
    hipEvent_t e;
    PSX_HIP(hipEventCreate(&e));
    c->events.push_back(e);
    c->kinds.push_back(-1);
  }
  PSX_HIP(hipEventRecord(c->events[c->nrec], st));
  c->kinds[c->nrec] = kind;
  ++c->nrec;
  return 0;
}

#define PSX_MARK(kind) \
  do {                 \
    int r_ = mark(c, (kind), st); \
    if (r_) return r_; \
  } while (0)

template <typename T>
int enqueue_prepare(psx_ctx* c, hipStream_t st) {
  Flow<T>& f = flow<T>(c);
  if (f.dim == 2) {
    if constexpr (sizeof(T) == 4) {
      if (c->ct) {
        PSX_HIP(launch_ct2d(f, 1, kStagePre, st));
        if (f.n1 > 1) PSX_HIP(launch_ct2d(f, 3, 0, st));
        return 0;
      }
    }
    if (f.n1 > 1) {
      PSX_HIP(launch_spec2d<T>(f, c->spec, true, kStagePre, st));
      PSX_HIP(launch_outer2d<T>(f, c->outerI, true, st));
    } else {
      PSX_HIP(launch_spec2d<T>(f, c->spec, false, kStagePre, st));
    }
  } else {
    if constexpr (sizeof(T) == 4) {
      if (c->ct) {
        PSX_HIP(launch_ct3d(f, 2, kStagePre, st));
        PSX_HIP(launch_ct3d(f, 3, 0, st));
        return 0;
      }
    }
    PSX_HIP(launch_spec3d<T>(f, c->spec, kStagePre, st));
    PSX_HIP(launch_ypass3d<T>(f, c->yI, true, f.rot ? 6 : 3, st));
  }
  return 0;
}

// One IF-RK4 step.  drv: 0 plain step (parameters from psx_set_step); 1 controller step
// (device controller around the stages, psx_drv.h); 2 controller step with CFL maxima taken
// by the first stage's line kernel.
template <typename T>
int enqueue_step(psx_ctx* c, hipStream_t st, int drv = 0) {
  Flow<T>& f = flow<T>(c);
  if constexpr (sizeof(T) == 4) {
    if (c->has_sgs) {
      PSX_MARK(PSX_KIND_SGS);
      const int r = enqueue_sgs(c->sgs, f.u0, f.sp, f.n, st);
      if (r) return r;
    }
  }
  for (int s = 0; s < 4; ++s) {
    const int lflags = (s == 0 && drv == 2) ? kLineWantMax : 0;
    bool ct = false;
    if constexpr (sizeof(T) == 4) ct = c->ct;
    // line kernel (real axis): inverse transform, products, forward transform
    PSX_MARK(PSX_KIND_LINE);
    if constexpr (sizeof(T) == 4) {
      if (ct) PSX_HIP(f.dim == 3 ? launch_ct3d(f, 0, lflags, st) : launch_ct2d(f, 0, lflags, st));
    }
    if (!ct) {
      if (f.dim == 2) PSX_HIP(launch_line2d<T>(f, c->line, lflags, st));
      else PSX_HIP(launch_line3d<T>(f, c->line, lflags, st));
    }
    if (s == 0 && drv) {
      PSX_MARK(PSX_KIND_CTRL);
      PSX_HIP(launch_drv(0, c->dmem, 1, st));
    }
    if constexpr (sizeof(T) == 4) {
      if (ct && f.dim == 3) {
        PSX_MARK(PSX_KIND_MID_FWD);
        PSX_HIP(launch_ct3d(f, 1, 0, st));
        PSX_MARK(PSX_KIND_SPEC);
        PSX_HIP(launch_ct3d(f, 2, s, st));
        PSX_MARK(PSX_KIND_MID_INV);
        PSX_HIP(launch_ct3d(f, 3, 0, st));
      } else if (ct) {
        if (f.n1 > 1) {
          PSX_MARK(PSX_KIND_OUTER_FWD);
          PSX_HIP(launch_ct2d(f, 2, 0, st));
        }
        PSX_MARK(PSX_KIND_SPEC);
        PSX_HIP(launch_ct2d(f, 1, s, st));
        if (f.n1 > 1) {
          PSX_MARK(PSX_KIND_OUTER_INV);
          PSX_HIP(launch_ct2d(f, 3, 0, st));
        }
      }
    }
    if (ct) {
    } else if (f.dim == 2) {
      if (f.n1 > 1) {
        PSX_MARK(PSX_KIND_OUTER_FWD);
        PSX_HIP(launch_outer2d<T>(f, c->outerF, false, st));
        PSX_MARK(PSX_KIND_SPEC);
        PSX_HIP(launch_spec2d<T>(f, c->spec, true, s, st));
        PSX_MARK(PSX_KIND_OUTER_INV);
        PSX_HIP(launch_outer2d<T>(f, c->outerI, true, st));
      } else {
        PSX_MARK(PSX_KIND_SPEC);
        PSX_HIP(launch_spec2d<T>(f, c->spec, false, s, st));
      }
    } else {
      PSX_MARK(PSX_KIND_MID_FWD);
      PSX_HIP(launch_ypass3d<T>(f, c->yF, false, f.rot ? 3 : 5, st));
      PSX_MARK(PSX_KIND_SPEC);
      PSX_HIP(launch_spec3d<T>(f, c->spec, s, st));
      PSX_MARK(PSX_KIND_MID_INV);
      PSX_HIP(launch_ypass3d<T>(f, c->yI, true, f.rot ? 6 : 3, st));
    }
  }
  PSX_MARK(PSX_KIND_CTRL);
  PSX_HIP(launch_hermitian<T>(f, st));
  if (drv) {
    PSX_HIP(launch_drv(1, c->dmem, 1, st));
    PSX_HIP(launch_drv(2, c->dmem, 1, st));
  }
  PSX_MARK(-1);
  return 0;
}

template <typename T>
int velocity_max(psx_ctx* c, double* out, hipStream_t st) {
  Flow<T>& f = flow<T>(c);
  PSX_HIP(launch_fill_u32(reinterpret_cast<unsigned*>(c->stats), 0u, 6, st));
  bool done = false;
  if constexpr (sizeof(T) == 4) {
    if (c->ct) {
      if (f.dim == 3)
        PSX_HIP(launch_ct3d(f, 0, kLineWantMax | kLineMaxOnly, st));
      else
        PSX_HIP(launch_ct2d(f, 0, kLineWantMax | kLineMaxOnly, st));
      done = true;
    }
  }
  if (done) {
  } else if (f.dim == 2) {
    PSX_HIP(launch_line2d<T>(f, c->line, kLineWantMax | kLineMaxOnly, st));
  } else {
    PSX_HIP(launch_line3d<T>(f, c->line, kLineWantMax | kLineMaxOnly, st));
  }
  unsigned long long h[3] = {0, 0, 0};
  PSX_HIP(hipMemcpyAsync(h, c->stats, sizeof h, hipMemcpyDeviceToHost, st));
  PSX_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < 3; ++i) {
    double v;
    std::memcpy(&v, &h[i], sizeof v);
    out[i] = v;
  }
  if (f.dim == 2) out[2] = 0.0;
  return 0;
}

}  // namespace

extern "C" {

const char* psx_last_error(void) { return g_err.c_str(); }

int psx_device_info(char* buf, int len) {
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev < 1) {
    if (buf && len > 0) snprintf(buf, len, "no HIP device");
    return fail("psx: no HIP device");
  }
  int dev = 0;
  PSX_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  PSX_HIP(hipGetDeviceProperties(&prop, dev));
  if (buf && len > 0)
    snprintf(buf, len, "%s %s lds/block=%zu", prop.name, prop.gcnArchName,
             (size_t)prop.sharedMemPerBlock);
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail("psx: device %s is not gfx950", prop.gcnArchName);
  return 0;
}

int psx_query_layout(const psx_flow_params* p, psx_layout* out) {
  if (!p || !out) return fail("psx_query_layout: null argument");
  psx_ctx tmp;
  tmp.p = *p;
  int r = p->fp64 ? plan_flow<double>(*p, &tmp, &tmp.f64) : plan_flow<float>(*p, &tmp, &tmp.f32);
  if (r) return r;
  *out = tmp.lay;
  return 0;
}

int psx_create(const psx_flow_params* p, void* u0, void* acc, void* work, psx_ctx** out) {
  if (!p || !u0 || !acc || !work || !out) return fail("psx_create: null argument");
  psx_ctx* c = new psx_ctx;
  c->p = *p;
  c->fp64 = p->fp64 != 0;
  int r = c->fp64 ? plan_flow<double>(*p, c, &c->f64) : plan_flow<float>(*p, c, &c->f32);
  if (r) {
    delete c;
    return r;
  }
  const int n = p->n, nh = n / 2 + 1;
  r = c->fp64 ? make_twiddles<double>(n, &c->tw) : make_twiddles<float>(n, &c->tw);
  if (!r)
    r = c->fp64 ? make_pass_twiddles<double>(&c->f64, &c->twp)
                : make_pass_twiddles<float>(&c->f32, &c->twp);
  if (r) {
    psx_destroy(c);
    return r;
  }
  const int nprm = 2 * nh + PrmLayout::kNScalars;
  c->nshell_max = (int)std::floor(std::sqrt((double)p->dim) * (n / 2) + 0.5) + 1;
  if (hipMalloc(&c->prm, nprm * sizeof(double)) != hipSuccess ||
      hipMalloc(&c->stats, 4 * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&c->partial, (size_t)c->nred * 4 * sizeof(double)) != hipSuccess ||
      hipMalloc(&c->dout, (size_t)(c->nshell_max + 8) * sizeof(double)) != hipSuccess) {
    psx_destroy(c);
    return fail("psx_create: hipMalloc failed");
  }
  // default parameters: dt = 0, no forcing (the caller sets them before stepping)
  hipStream_t st = nullptr;
  const double zero9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // phase shifts: unshifted grids
  if (launch_params(c->prm, nh, p->nu, 0.0, p->alpha, 0.0, p->klo2, p->khi2, p->ka2, st) !=
          hipSuccess ||
      launch_set_f64(c->prm + 2 * nh + PrmLayout::kPsA, zero9, 9, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    psx_destroy(c);
    return fail("psx_create: parameter kernel failed");
  }
  auto fill = [&](auto& f) {
    using T = std::remove_reference_t<decltype(f.scale)>;
    f.W = static_cast<cx<T>*>(work);
    f.u0 = static_cast<cx<T>*>(u0);
    f.acc = static_cast<cx<T>*>(acc);
    f.tw = static_cast<const cx<T>*>(c->tw);
    f.twp = static_cast<const cx<T>*>(c->twp);
    f.prm = c->prm;
    f.stats = c->stats;
  };
  if (c->fp64) fill(c->f64);
  else fill(c->f32);
  *out = c;
  return 0;
}

int psx_destroy(psx_ctx* c) {
  if (!c) return 0;
  for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->cap) (void)hipStreamDestroy(c->cap);
  if (c->tw) (void)hipFree(c->tw);
  if (c->twp) (void)hipFree(c->twp);
  if (c->prm) (void)hipFree(c->prm);
  if (c->stats) (void)hipFree(c->stats);
  if (c->partial) (void)hipFree(c->partial);
  if (c->dout) (void)hipFree(c->dout);
  if (c->rgexec) (void)hipGraphExecDestroy(c->rgexec);
  if (c->rgraph) (void)hipGraphDestroy(c->rgraph);
  if (c->dmem) (void)hipFree(c->dmem);
  if (c->dstate) (void)hipFree(c->dstate);
  if (c->dring) (void)hipFree(c->dring);
  if (c->dpartial) (void)hipFree(c->dpartial);
  if (c->hslots) (void)hipHostFree(c->hslots);
  for (hipEvent_t e : c->slot_ev)
    if (e) (void)hipEventDestroy(e);
  delete c;
  return 0;
}

int psx_prepare(psx_ctx* c, void* stream) {
  if (!c) return fail("psx_prepare: null context");
  hipStream_t st = static_cast<hipStream_t>(stream);
  return c->fp64 ? enqueue_prepare<double>(c, st) : enqueue_prepare<float>(c, st);
}

int psx_set_step(psx_ctx* c, double dt, double amp, void* stream) {
  if (!c) return fail("psx_set_step: null context");
  if (!std::isfinite(dt) || !std::isfinite(amp)) return fail("psx_set_step: non-finite dt/amp");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int nh = c->p.n / 2 + 1;
  PSX_HIP(launch_params(c->prm, nh, c->p.nu, dt, c->p.alpha, amp, c->p.klo2, c->p.khi2,
                        c->p.ka2, st));
  return 0;
}

int psx_set_phase_shift(psx_ctx* c, const double* shift, const double* next, void* stream) {
  if (!c || !shift || !next) return fail("psx_set_phase_shift: null argument");
  if (c->p.dealias != 1) return fail("psx_set_phase_shift: the context uses the 2/3 rule");
  if (c->run_mode) return fail("psx_set_phase_shift: a device-loop run is active");
  const int d = c->p.dim;
  const double half = M_PI / c->p.n;  // the oracle's shift + pi / n, in FP64
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < d; ++i) {
    if (!std::isfinite(shift[i]) || !std::isfinite(next[i]))
      return fail("psx_set_phase_shift: non-finite shift");
    v[i] = shift[i];
    v[3 + i] = shift[i] + half;
    v[6 + i] = next[i];
  }
  const int nh = c->p.n / 2 + 1;
  static_assert(PrmLayout::kPsB == PrmLayout::kPsA + 3 && PrmLayout::kPsN == PrmLayout::kPsA + 6,
                "shift entries are contiguous");
  PSX_HIP(launch_set_f64(c->prm + 2 * nh + PrmLayout::kPsA, v, 9,
                         static_cast<hipStream_t>(stream)));
  return 0;
}

int psx_step(psx_ctx* c, void* stream) {
  if (!c) return fail("psx_step: null context");
  hipStream_t st = static_cast<hipStream_t>(stream);
  return c->fp64 ? enqueue_step<double>(c, st) : enqueue_step<float>(c, st);
}

int psx_step_graph(psx_ctx* c, void* stream) {
  if (!c) return fail("psx_step_graph: null context");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (!c->gexec) {
    if (!c->cap) PSX_HIP(hipStreamCreateWithFlags(&c->cap, hipStreamNonBlocking));
    const bool was_timing = c->timing;
    c->timing = false;  // no event records inside the captured graph
    PSX_HIP(hipStreamBeginCapture(c->cap, hipStreamCaptureModeRelaxed));
    int r = c->fp64 ? enqueue_step<double>(c, c->cap) : enqueue_step<float>(c, c->cap);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->cap, &g);
    c->timing = was_timing;
    if (r) {
      if (g) (void)hipGraphDestroy(g);
      return r;
    }
    if (e != hipSuccess) return fail("psx_step_graph: capture failed: %s", hipGetErrorString(e));
    c->graph = g;
    PSX_HIP(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
  }
  PSX_HIP(hipGraphLaunch(c->gexec, st));
  return 0;
}

int psx_timing_begin(psx_ctx* c) {
  if (!c) return fail("psx_timing_begin: null context");
  c->timing = true;
  c->nrec = 0;
  return 0;
}

int psx_timing_end(psx_ctx* c, void* stream, double* ms, int* count) {
  if (!c || !ms || !count) return fail("psx_timing_end: null argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  c->timing = false;
  PSX_HIP(hipStreamSynchronize(st));
  for (int k = 0; k < PSX_NKINDS; ++k) {
    ms[k] = 0.0;
    count[k] = 0;
  }
  for (int i = 0; i + 1 < c->nrec; ++i) {
    const int k = c->kinds[i];
    if (k < 0 || k >= PSX_NKINDS) continue;
    float t = 0.f;
    PSX_HIP(hipEventElapsedTime(&t, c->events[i], c->events[i + 1]));
    ms[k] += t;
    count[k] += 1;
  }
  c->nrec = 0;
  return 0;
}

int psx_velocity_max(psx_ctx* c, double* out3, void* stream) {
  if (!c || !out3) return fail("psx_velocity_max: null argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  return c->fp64 ? velocity_max<double>(c, out3, st) : velocity_max<float>(c, out3, st);
}

int psx_reduce(psx_ctx* c, const void* state, int pitch, double klo2, double khi2,
               double* out4, void* stream) {
  if (!c || !state || !out4) return fail("psx_reduce: null argument");
  const int n = c->p.n, dim = c->p.dim;
  if (pitch < n / 2 + 1) return fail("psx_reduce: pitch %d < n/2+1", pitch);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const long long sst = (long long)(dim == 2 ? n : n * n) * pitch;
  if (c->fp64)
    PSX_HIP(launch_reduce<double>(static_cast<const cx<double>*>(state), sst, pitch, n, dim,
                                  klo2, khi2, c->partial, c->nred, c->dout, st));
  else
    PSX_HIP(launch_reduce<float>(static_cast<const cx<float>*>(state), sst, pitch, n, dim, klo2,
                                 khi2, c->partial, c->nred, c->dout, st));
  PSX_HIP(hipMemcpyAsync(out4, c->dout, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
  PSX_HIP(hipStreamSynchronize(st));
  return 0;
}

int psx_spectrum(psx_ctx* c, const void* state, int pitch, double* out, int nshell,
                 void* stream) {
  if (!c || !state || !out) return fail("psx_spectrum: null argument");
  if (nshell < 1 || nshell > c->nshell_max)
    return fail("psx_spectrum: nshell=%d outside [1, %d]", nshell, c->nshell_max);
  const int n = c->p.n, dim = c->p.dim;
  if (pitch < n / 2 + 1) return fail("psx_spectrum: pitch %d < n/2+1", pitch);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const long long sst = (long long)(dim == 2 ? n : n * n) * pitch;
  if (c->fp64)
    PSX_HIP(launch_spectrum<double>(static_cast<const cx<double>*>(state), sst, pitch, n, dim,
                                    c->dout, nshell, st));
  else
    PSX_HIP(launch_spectrum<float>(static_cast<const cx<float>*>(state), sst, pitch, n, dim,
                                   c->dout, nshell, st));
  PSX_HIP(hipMemcpyAsync(out, c->dout, nshell * sizeof(double), hipMemcpyDeviceToHost, st));
  PSX_HIP(hipStreamSynchronize(st));
  return 0;
}

int psx_set_source(psx_ctx* c, void* src) {
  if (!c) return fail("psx_set_source: null context");
  if (c->p.dim != 3) return fail("psx_set_source: 3D flows only");
  if (src && c->p.dealias != 0) return fail("psx_set_source: 2/3-rule flows only");
  if (c->fp64) c->f64.src = static_cast<const cx<double>*>(src);
  else c->f32.src = static_cast<const cx<float>*>(src);
  // the flow descriptor is a kernel argument: captured graphs hold the old one
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->rgexec) (void)hipGraphExecDestroy(c->rgexec);
  if (c->rgraph) (void)hipGraphDestroy(c->rgraph);
  c->gexec = c->rgexec = nullptr;
  c->graph = c->rgraph = nullptr;
  c->run_chunk = 0;
  return 0;
}

int psx_set_sgs(psx_ctx* c, const psx_sgs_desc* d) {
  if (!c) return fail("psx_set_sgs: null context");
  if (!d) {
    c->has_sgs = false;
    return psx_set_source(c, nullptr);
  }
  if (c->p.dim != 3 || c->fp64) return fail("psx_set_sgs: 3D fp32 flows only");
  if (!c2c_supported(c->p.n)) return fail("psx_set_sgs: n=%d unsupported (64, 128, 256)", c->p.n);
  if (d->model != PSX_SGS_SMAGORINSKY && d->model != PSX_SGS_NET)
    return fail("psx_set_sgs: unknown model %d", d->model);
  if (!(d->backscatter >= 0.0 && d->backscatter <= 2.0))
    return fail("psx_set_sgs: backscatter factor %g outside [0, 2]", d->backscatter);
  if (!(d->smag >= 0.0) || (d->model == PSX_SGS_SMAGORINSKY && d->smag == 0.0))
    return fail("psx_set_sgs: Smagorinsky (Cs Delta)^2 = %g", d->smag);
  if (!d->src || !d->full || !d->grad || !d->tau) return fail("psx_set_sgs: null buffer");
  if (d->model == PSX_SGS_NET) {
    if (!d->x || !d->out) return fail("psx_set_sgs: null network buffer");
    for (int i = 0; i < 6; ++i)
      if (!d->w[i] || !d->b[i] || (i < 5 && !d->act[i]))
        return fail("psx_set_sgs: null network tensor (layer %d)", i);
  }
  // tables the evaluation needs are made here, never inside a graph capture of a step
  if (sgs_prepare(c->p.n)) return 1;
  const int r = psx_set_source(c, d->src);  // also drops the captured graphs
  if (r) return r;
  c->sgs = *d;
  c->has_sgs = true;
  return 0;
}

int psx_sgs_update(psx_ctx* c, void* stream) {
  if (!c) return fail("psx_sgs_update: null context");
  if (!c->has_sgs) return fail("psx_sgs_update: no subgrid-stress model attached");
  return enqueue_sgs(c->sgs, c->f32.u0, c->f32.sp, c->f32.n, static_cast<hipStream_t>(stream));
}

// ------------------------------------------------------------------ device controller
int psx_run_begin(psx_ctx* c, const psx_run_params* rp, void* stream) {
  if (!c || !rp) return fail("psx_run_begin: null argument");
  if (rp->cfl_every < 1) return fail("psx_run_begin: cfl_every must be >= 1");
  if (c->p.dealias != 0)
    return fail("psx_run_begin: phase-shift flows draw their shifts on the host; step them "
                "with psx_step / psx_step_graph");
  if (rp->cap_rec < 1 || rp->cap_step < 1) return fail("psx_run_begin: ring capacities < 1");
  if (!(rp->dt_fixed > 0.0) && !(rp->cfl > 0.0 && rp->dt_max > 0.0))
    return fail("psx_run_begin: need dt_fixed > 0, or cfl > 0 and dt_max > 0");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (c->hmem.cap_rec != rp->cap_rec || c->hmem.cap_step != rp->cap_step) {
    if (c->dring) (void)hipFree(c->dring);
    c->dring = nullptr;
    PSX_HIP(hipMalloc(&c->dring, (size_t)(3 * rp->cap_rec + 2 * rp->cap_step) * sizeof(double)));
  }
  if (!c->dmem) {
    PSX_HIP(hipMalloc(&c->dmem, sizeof(DrvMember)));
    PSX_HIP(hipMalloc(&c->dstate, sizeof(DrvState)));
    PSX_HIP(hipMalloc(&c->dpartial, 2 * kDrvRedBlocks * sizeof(double)));
    PSX_HIP(hipHostMalloc(&c->hslots, psx_ctx::kSlots * sizeof(DrvState)));
    for (hipEvent_t& e : c->slot_ev) PSX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  c->nchunk = 0;
  DrvMember m{};
  m.u0 = c->fp64 ? (const void*)c->f64.u0 : (const void*)c->f32.u0;
  m.prm = c->prm;
  m.stats = c->stats;
  m.s = c->dstate;
  m.rec = c->dring;
  m.steps = c->dring + 3 * rp->cap_rec;
  m.partial = c->dpartial;
  m.nu = c->p.nu;
  m.alpha = c->p.alpha;
  m.klo2 = c->p.klo2;
  m.khi2 = c->p.khi2;
  m.ka2 = c->p.ka2;
  m.n = c->p.n;
  m.dim = c->p.dim;
  m.sp = c->fp64 ? c->f64.sp : c->f32.sp;
  m.fp64 = c->fp64;
  m.cap_rec = rp->cap_rec;
  m.cap_step = rp->cap_step;
  const DrvState s = drv_initial_state(*rp, c->p.n);
  c->hmem = m;
  c->run_mode = s.dt_fixed > 0.0 ? 1 : 2;
  // The copies read `s` in this frame asynchronously: the stream is drained before
  // returning, on the error paths too.
  auto enqueue = [&]() -> int {
    PSX_HIP(hipMemcpyAsync(c->dmem, &c->hmem, sizeof(DrvMember), hipMemcpyHostToDevice, st));
    PSX_HIP(hipMemcpyAsync(c->dstate, &s, sizeof(DrvState), hipMemcpyHostToDevice, st));
    PSX_HIP(launch_fill_u32(reinterpret_cast<unsigned*>(c->stats), 0u, 6, st));
    if (s.active) {
      // the parameter block marks the flow active; the controller rewrites it every step
      const int nh = c->p.n / 2 + 1;
      PSX_HIP(launch_params(c->prm, nh, c->p.nu, 0.0, c->p.alpha, rp->amp0, c->p.klo2,
                            c->p.khi2, c->p.ka2, st));
    } else {
      PSX_HIP(launch_drv(4, c->dmem, 1, st));
    }
    PSX_HIP(launch_drv(2, c->dmem, 1, st));  // initial (or end) record when due
    return 0;
  };
  const int r = enqueue();
  const hipError_t es = hipStreamSynchronize(st);
  if (r) {
    c->run_mode = 0;
    return r;
  }
  if (es != hipSuccess) {
    c->run_mode = 0;
    return fail("psx_run_begin: %s", hipGetErrorString(es));
  }
  return 0;
}

int psx_run_chunk(psx_ctx* c, int nsteps, int use_graph, void* stream) {
  if (!c) return fail("psx_run_chunk: null context");
  if (!c->dmem || !c->run_mode) return fail("psx_run_chunk: no run begun");
  if (nsteps < 1) return fail("psx_run_chunk: nsteps must be >= 1");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int drv = c->run_mode;
  // a slot is reused only after the host no longer waits on it
  const int slot = (int)(c->nchunk % psx_ctx::kSlots);
  PSX_HIP(hipEventSynchronize(c->slot_ev[slot]));
  if (!use_graph) {
    for (int k = 0; k < nsteps; ++k) {
      int r = c->fp64 ? enqueue_step<double>(c, st, drv) : enqueue_step<float>(c, st, drv);
      if (r) return r;
    }
    PSX_HIP(hipMemcpyAsync(c->hslots + slot, c->dstate, sizeof(DrvState),
                           hipMemcpyDeviceToHost, st));
    PSX_HIP(hipEventRecord(c->slot_ev[slot], st));
    ++c->nchunk;
    return 0;
  }
  const int key = nsteps * 4 + drv;
  if (!c->rgexec || c->run_chunk != key) {
    if (c->rgexec) (void)hipGraphExecDestroy(c->rgexec);
    if (c->rgraph) (void)hipGraphDestroy(c->rgraph);
    c->rgexec = nullptr;
    c->rgraph = nullptr;
    if (!c->cap) PSX_HIP(hipStreamCreateWithFlags(&c->cap, hipStreamNonBlocking));
    const bool was_timing = c->timing;
    c->timing = false;
    PSX_HIP(hipStreamBeginCapture(c->cap, hipStreamCaptureModeRelaxed));
    int r = 0;
    for (int k = 0; k < nsteps && !r; ++k)
      r = c->fp64 ? enqueue_step<double>(c, c->cap, drv) : enqueue_step<float>(c, c->cap, drv);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->cap, &g);
    c->timing = was_timing;
    if (r) {
      if (g) (void)hipGraphDestroy(g);
      return r;
    }
    if (e != hipSuccess) return fail("psx_run_chunk: capture failed: %s", hipGetErrorString(e));
    c->rgraph = g;
    PSX_HIP(hipGraphInstantiate(&c->rgexec, c->rgraph, nullptr, nullptr, 0));
    c->run_chunk = key;
  }
  PSX_HIP(hipGraphLaunch(c->rgexec, st));
  PSX_HIP(hipMemcpyAsync(c->hslots + slot, c->dstate, sizeof(DrvState), hipMemcpyDeviceToHost,
                         st));
  PSX_HIP(hipEventRecord(c->slot_ev[slot], st));
  ++c->nchunk;
  return 0;
}

int psx_run_peek(psx_ctx* c, int lag, psx_run_state* out) {
  if (!c || !out) return fail("psx_run_peek: null argument");
  if (lag < 0 || lag >= psx_ctx::kSlots || lag >= c->nchunk)
    return fail("psx_run_peek: lag %d outside [0, min(%d, %lld))", lag, psx_ctx::kSlots,
                c->nchunk);
  const int slot = (int)((c->nchunk - 1 - lag) % psx_ctx::kSlots);
  PSX_HIP(hipEventSynchronize(c->slot_ev[slot]));
  drv_export(c->hslots[slot], out);
  return 0;
}

int psx_run_poll(psx_ctx* c, psx_run_state* out, double* rec, double* steps, void* stream) {
  if (!c || !out) return fail("psx_run_poll: null argument");
  if (!c->dmem) return fail("psx_run_poll: no run begun");
  hipStream_t st = static_cast<hipStream_t>(stream);
  DrvState s{};
  PSX_HIP(hipMemcpyAsync(&s, c->dstate, sizeof s, hipMemcpyDeviceToHost, st));
  if (rec)
    PSX_HIP(hipMemcpyAsync(rec, c->hmem.rec, 3 * (size_t)c->hmem.cap_rec * sizeof(double),
                           hipMemcpyDeviceToHost, st));
  if (steps)
    PSX_HIP(hipMemcpyAsync(steps, c->hmem.steps, 2 * (size_t)c->hmem.cap_step * sizeof(double),
                           hipMemcpyDeviceToHost, st));
  PSX_HIP(hipStreamSynchronize(st));
  drv_export(s, out);
  return 0;
}

int psx_run_end(psx_ctx* c, void* stream) {
  if (!c) return fail("psx_run_end: null context");
  if (!c->dmem) return 0;
  hipStream_t st = static_cast<hipStream_t>(stream);
  PSX_HIP(launch_drv(3, c->dmem, 1, st));
  c->run_mode = 0;
  return 0;
}

}  // extern "C"
