// Test hooks of the FFT engine: batched in-place complex fp32 transforms of rows of length L
// with exactly the compile-time plans, work-group shapes and twiddle tables the step kernels
// use for that length (the 2D line kernel's for 64..8192, the 3D line kernel's for 96, 192 and
// 384), and the twiddle tables themselves.  They let the engine be checked against numpy.fft
// directly, outside the fused kernels.  No counterpart in the reference (its FFTs are cuFFT).
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "../../include/psx.h"
#include "psx_fft_ct.h"
#include "psx_host.h"

namespace psx {
namespace {

struct LinePlan {
  int E, NT, RT;
  CtPlan P;
};

// The line-kernel plan for length L: 2D geometry for the 2D compile-time sizes, 3D geometry
// for the radix-3 3D sizes.
constexpr LinePlan line_plan(int L) {
  if (L == 96 || L == 192 || L == 384) {
    const Geo3D g = geo3d(L);
    return LinePlan{g.lineE, g.lineNT, g.lineT, g.line};
  }
  const Geo2D g = geo2d(L, false);
  return LinePlan{g.lineE, g.lineNT, g.lineT, g.line};
}

template <int L, bool INV>
__global__ void __launch_bounds__(line_plan(L).NT) k_fft_rows(cx<float>* __restrict__ d,
                                                             const cx<float>* __restrict__ tw,
                                                             int nrows) {
  constexpr LinePlan lp = line_plan(L);
  constexpr int E = lp.E, NT = lp.NT, RT = lp.RT;
  using Lay = RowsCT<L, RT>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<float>* lds = reinterpret_cast<cx<float>*>(smem_raw);
  const int tid = threadIdx.x;
  const int row0 = blockIdx.x * RT;
  cx<float> v[E];
  constexpr int R0 = ct_radix(lp.P, INV, 0);
#pragma unroll
  for (int i = 0; i < E; ++i) {
    int e, t;
    ct_elem<R0, L, NT, Lay>(i, tid, e, t);
    const int row = row0 + t;
    v[i] = row < nrows ? d[(long long)row * L + e] : czero<float>();
  }
  fft_ct<E, INV, NT, Lay, lp.P>(v, lds, tw, tid);
  constexpr int RL = ct_radix(lp.P, INV, lp.P.np - 1);
#pragma unroll
  for (int i = 0; i < E; ++i) {
    int e, t;
    ct_elem<RL, L, NT, Lay>(i, tid, e, t);
    const int row = row0 + t;
    if (row < nrows) d[(long long)row * L + e] = v[i];
  }
}

FftPlan to_fft_plan(const CtPlan& cp) {
  FftPlan fp{};
  fp.L = cp.L;
  fp.np = cp.np;
  for (int i = 0; i < kMaxPasses; ++i) fp.rad[i] = i < cp.np ? cp.rad[i] : 1;
  return fp;
}

// Both orders' twiddle tables of the line plan for L, on the device, made once per length.
const cx<float>* line_twiddles(int L, const CtPlan& cp, FftPlan* fp) {
  static std::mutex mu;
  static std::vector<std::pair<int, void*>> cache;
  std::lock_guard<std::mutex> lock(mu);
  *fp = to_fft_plan(cp);
  std::vector<cx<float>> buf;
  append_pass_twiddles<float>(fp, buf);
  for (auto& e : cache)
    if (e.first == L) return static_cast<const cx<float>*>(e.second);
  if (buf.empty()) buf.push_back(cx<float>{1.f, 0.f});
  void* dptr = nullptr;
  if (hipMalloc(&dptr, buf.size() * sizeof(cx<float>)) != hipSuccess) return nullptr;
  if (hipMemcpy(dptr, buf.data(), buf.size() * sizeof(cx<float>), hipMemcpyHostToDevice) !=
      hipSuccess) {
    (void)hipFree(dptr);
    return nullptr;
  }
  cache.emplace_back(L, dptr);
  return static_cast<const cx<float>*>(dptr);
}

template <int L>
int run_rows(cx<float>* d, int nrows, bool inv, hipStream_t st) {
  constexpr LinePlan lp = line_plan(L);
  static_assert(lp.E > 0 && lp.P.np > 0, "no line plan for this length");
  FftPlan fp{};
  const cx<float>* tw = line_twiddles(L, lp.P, &fp);
  if (!tw) return fail("psx_fft_lines: twiddle table allocation failed");
  const cx<float>* two = tw + fp.twoff[inv ? 1 : 0][0];
  const unsigned blocks = (unsigned)((nrows + lp.RT - 1) / lp.RT);
  const size_t lds = (size_t)lp.RT * ct_rows_pitch(L) * sizeof(cx<float>);
  if (inv) {
    PSX_HIP(hipFuncSetAttribute((const void*)k_fft_rows<L, true>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_fft_rows<L, true>), dim3(blocks), dim3(lp.NT), lds, st, d, two, nrows);
  } else {
    PSX_HIP(hipFuncSetAttribute((const void*)k_fft_rows<L, false>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_fft_rows<L, false>), dim3(blocks), dim3(lp.NT), lds, st, d, two,
                       nrows);
  }
  PSX_HIP(hipGetLastError());
  return 0;
}

#define PSX_FFT_LENGTHS(X) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(1024) X(2048) \
  X(4096) X(8192)

}  // namespace
}  // namespace psx

using namespace psx;

extern "C" {

int psx_fft_lines(void* data, int L, int nrows, int inverse, void* stream) {
  if (!data || nrows < 1) return fail("psx_fft_lines: bad arguments");
  auto* d = static_cast<cx<float>*>(data);
  const hipStream_t st = static_cast<hipStream_t>(stream);
  switch (L) {
#define PSX_X(LV) \
  case LV: return run_rows<LV>(d, nrows, inverse != 0, st);
    PSX_FFT_LENGTHS(PSX_X)
#undef PSX_X
    default:
      return fail("psx_fft_lines: L=%d has no line plan (64..8192, 96, 192, 384)", L);
  }
}

int psx_fft_twiddles(int L, int inverse, float* out, int cap, int* radices, int* npasses) {
  if (!out || !radices || !npasses) return fail("psx_fft_twiddles: null argument");
  CtPlan cp{};
  switch (L) {
#define PSX_X(LV) \
  case LV: cp = line_plan(LV).P; break;
    PSX_FFT_LENGTHS(PSX_X)
#undef PSX_X
    default:
      return fail("psx_fft_twiddles: L=%d has no line plan", L);
  }
  FftPlan fp = to_fft_plan(cp);
  std::vector<cx<float>> buf;
  append_pass_twiddles<float>(&fp, buf);
  const int order = inverse ? 1 : 0;
  const int begin = fp.twoff[order][0];
  const int end = order == 0 ? fp.twoff[1][0] : (int)buf.size();
  if (end - begin > cap) return fail("psx_fft_twiddles: %d entries > cap %d", end - begin, cap);
  for (int i = begin; i < end; ++i) {
    out[2 * (i - begin)] = buf[i].x;
    out[2 * (i - begin) + 1] = buf[i].y;
  }
  *npasses = fp.np;
  for (int q = 0; q < fp.np; ++q) radices[q] = inverse ? fp.rad[fp.np - 1 - q] : fp.rad[q];
  return end - begin >= 0 ? 0 : fail("psx_fft_twiddles: bad table layout");
}

}  // extern "C"
