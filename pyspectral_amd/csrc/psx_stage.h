// Pointwise physics shared by the 2D and 3D stage kernels: integrating-factor values and the
// Lawson IF-RK4 stage algebra.
//
// Reference semantics: src/pyspectral/kernels/spectral.cu:43-80 (if_factors, rk_stage) and
// src/pyspectral/backends/numpy_ref.py:145-153 (the Lawson combination).  Unlike the
// reference, whose table is rounded to the working precision on the host
// (cupy_cuda.py:151-161), the factors are formed here in FP64 from an FP64 table that the
// parameter kernel computes on the device, and rounded once.
#pragma once

#include "psx_flow.h"

namespace psx {

enum { kStagePre = -1 };

struct IfScalars {
  double ea1, eah, eb1, ebh, klo2, khi2, ka2;
  bool band;
};

__device__ __forceinline__ IfScalars load_if_scalars(const double* __restrict__ prm, int nh) {
  const double* s = prm + 2 * nh;
  IfScalars r;
  r.ea1 = s[PrmLayout::kEa1];
  r.eah = s[PrmLayout::kEah];
  r.eb1 = s[PrmLayout::kEb1];
  r.ebh = s[PrmLayout::kEbh];
  r.klo2 = s[PrmLayout::kKlo2];
  r.khi2 = s[PrmLayout::kKhi2];
  r.ka2 = s[PrmLayout::kKa2];
  r.band = s[PrmLayout::kBand] != 0.0;
  return r;
}

// exp(L dt) and exp(L dt/2) for one mode: separable viscous part from the tables, times the
// drag factor when k^2 <= ka^2 and the band factor when klo^2 <= k^2 <= khi^2.
// az < 0 for 2D.
__device__ __forceinline__ void if_factors(const double* __restrict__ prm, int nh,
                                           const IfScalars& s, int ax, int ay, int az,
                                           long long k2, double& E, double& Eh) {
  E = prm[ax] * prm[ay];
  Eh = prm[nh + ax] * prm[nh + ay];
  if (az >= 0) {
    E *= prm[az];
    Eh *= prm[nh + az];
  }
  const double dk2 = (double)k2;
  if (dk2 <= s.ka2) {
    E *= s.ea1;
    Eh *= s.eah;
  }
  if (s.band && dk2 >= s.klo2 && dk2 <= s.khi2) {
    E *= s.eb1;
    Eh *= s.ebh;
  }
}

// Phase-shift dealiasing (numpy_ref.py:68-72): exp(sign i k.s) for the grid shift s (three
// FP64 entries of the parameter block, zero beyond the dimension), argument and phase in FP64
// as the oracle forms them, rounded once.
template <typename T>
__device__ __forceinline__ cx<T> shift_phase(const double* __restrict__ s, int kx, int ky,
                                             int kz, double sign) {
  const double arg = ((double)kx * s[0] + (double)ky * s[1]) + (double)kz * s[2];
  double sn, cs;
  sincos(arg, &sn, &cs);
  return mkc((T)cs, (T)(sign * sn));
}
// Shift entries of stage `stage`'s nonlinear term (stages 0-1: the step's shift, 2-3: that
// plus pi/n) and of the inputs it prepares for the next stage (kStagePre and stage 0: the
// step's shift; 1, 2: plus pi/n; 3: the next step's shift).
__device__ __forceinline__ const double* shift_of_term(const double* prm, int nh, int stage) {
  return prm + 2 * nh + (stage < 2 ? PrmLayout::kPsA : PrmLayout::kPsB);
}
__device__ __forceinline__ const double* shift_of_inputs(const double* prm, int nh, int stage) {
  return prm + 2 * nh + (stage <= 0 ? PrmLayout::kPsA
                         : stage < 3 ? PrmLayout::kPsB : PrmLayout::kPsN);
}

// One Lawson stage for one mode/component.  With k1..k4 the stage nonlinear terms:
//   stage 0: acc = E k1,            next = Eh (u + dt/2 k1)
//   stage 1: acc += 2 Eh k2,        next = Eh u + dt/2 k2
//   stage 2: acc += 2 Eh k3,        next = E u + dt Eh k3
//   stage 3: u = E u + dt/6 (acc + k4), next = u
// The factors come in as D = 1 - E and Dh = 1 - Eh (formed in FP64, rounded once) and act
// as E x = x - D x: the factors of the energy-carrying modes lie within 1e-6 of 1, where the
// working precision would round E itself to one of a few neighbouring values (in fp32 a
// relative error of up to 3e-8 against a decay of 1e-7 per step, the same every step), while
// D keeps its full relative precision.
template <typename T>
__device__ __forceinline__ cx<T> emul(T D, cx<T> x) {
  return mkc(x.x - D * x.x, x.y - D * x.y);
}

template <typename T>
__device__ __forceinline__ cx<T> rk_update(int stage, T dt, T D, T Dh, cx<T> N,
                                           cx<T>* __restrict__ up, cx<T>* __restrict__ ap) {
  const cx<T> u = *up;
  if (stage == 0) {
    *ap = emul(D, N);
    return emul(Dh, cadd(u, cscl((T)0.5 * dt, N)));
  }
  if (stage == 1) {
    *ap = cadd(*ap, cscl((T)2, emul(Dh, N)));
    return cadd(emul(Dh, u), cscl((T)0.5 * dt, N));
  }
  if (stage == 2) {
    *ap = cadd(*ap, cscl((T)2, emul(Dh, N)));
    return cadd(emul(D, u), cscl(dt, emul(Dh, N)));
  }
  const cx<T> r = cadd(emul(D, u), cscl(dt / (T)6, cadd(*ap, N)));
  *up = r;
  return r;
}

// Value form of rk_update: u is the stage's state value, acc the accumulator (updated in
// place); returns the next stage input (stage 3: the new state).
template <int stage, typename T>
__device__ __forceinline__ cx<T> rk_value(T dt, T D, T Dh, cx<T> N, cx<T> u, cx<T>& acc) {
  if constexpr (stage == 0) {
    acc = emul(D, N);
    return emul(Dh, cadd(u, cscl((T)0.5 * dt, N)));
  } else if constexpr (stage == 1) {
    acc = cadd(acc, cscl((T)2, emul(Dh, N)));
    return cadd(emul(Dh, u), cscl((T)0.5 * dt, N));
  } else if constexpr (stage == 2) {
    acc = cadd(acc, cscl((T)2, emul(Dh, N)));
    return cadd(emul(D, u), cscl(dt, emul(Dh, N)));
  } else {
    return cadd(emul(D, u), cscl(dt / (T)6, cadd(acc, N)));
  }
}

// Raw buffer access to fp32 complex data: loads at out-of-range byte offsets return zero and
// stores there are dropped, so masked accesses need no branches.  The region must be smaller
// than 2 GiB (kOob lies beyond it).  Each complex is moved as two 32-bit halves, which the
// compiler merges into one 64-bit access (the 64-bit raw buffer load builtin of this toolchain
// loads only the first dword).
constexpr unsigned kOob = 0x80000000u;
// Cache policy of these accesses (template argument POL, the buffer instruction's cache-policy
// bits): non-temporal (kNt) by default.  The state, the accumulator and the packed line
// spectra are each read and written once per stage, so keeping them out of the caches leaves
// room for the data that is reused: measured 1.5 % on the 8192^2 step, 2.6 % on 384^3 and
// 6.8 % on the forced ensemble, results bit-identical.  Grids whose fields fit the caches
// (3D below 384^3: the 64^3 LES, the 256^3 DNS) were slower with it and use kCached.
constexpr int kNt = 2;
constexpr int kCached = 0;

template <int POL = kNt>
struct BufC32T {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ BufC32T(const void* base, unsigned long long bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0,
                                            (int)(bytes < kOob ? bytes : kOob - 8), 0x00020000)) {}
  __device__ __forceinline__ cx<float> load(unsigned off) const {
    const unsigned a = __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, POL);
    const unsigned b = __builtin_amdgcn_raw_buffer_load_b32(r, (int)(off + 4u), 0, POL);
    return mkc(__builtin_bit_cast(float, a), __builtin_bit_cast(float, b));
  }
  __device__ __forceinline__ void store(unsigned off, cx<float> x) const {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x.x), r, (int)off, 0,
                                          POL);
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x.y), r, (int)(off + 4u),
                                          0, POL);
  }
};
using BufC32 = BufC32T<kNt>;

// Block-wide max of NV non-negative values, folded into stats[] with one atomic per
// work-group (bit patterns of non-negative doubles order like unsigned 64-bit integers).
// Works for any block size up to 512, including partial wavefronts.  `red` is LDS scratch of
// at least NV * blockDim.x doubles that no thread still reads (caller synchronises before).
template <int NV>
__device__ __forceinline__ void block_max_to(double (&m)[NV],
                                             unsigned long long* __restrict__ stats,
                                             double* red) {
  const int tid = threadIdx.x, nt = blockDim.x;
  __syncthreads();
#pragma unroll
  for (int c = 0; c < NV; ++c) red[c * nt + tid] = m[c];
  __syncthreads();
  for (int s = 256; s > 0; s >>= 1) {
    if (tid < s && tid + s < nt) {
#pragma unroll
      for (int c = 0; c < NV; ++c) red[c * nt + tid] = fmax(red[c * nt + tid], red[c * nt + tid + s]);
    }
    __syncthreads();
  }
  if (tid == 0) {
#pragma unroll
    for (int c = 0; c < NV; ++c)
      atomicMax(stats + c, (unsigned long long)__double_as_longlong(red[c * nt]));
  }
  __syncthreads();
}

// Leray projection of one mode of a 3D state: u <- u - k (k.u) / k^2 (ik2 = 1/k^2, 0 at k = 0).
// Applied to the new state after every step and to a loaded state: the nonlinear term is
// projected in every stage, but the linear factors would keep (and, inside a forcing band,
// amplify) any compressive part the state carries, such as round-off of the projection itself,
// and the product forms conserve energy only for divergence-free velocities.
template <typename T>
__device__ __forceinline__ void leray(T fx, T fy, T fz, T ik2, cx<T>& a, cx<T>& b, cx<T>& c) {
  const cx<T> d = cscl(ik2, cadd(cadd(cscl(fx, a), cscl(fy, b)), cscl(fz, c)));
  a = csub(a, cscl(fx, d));
  b = csub(b, cscl(fy, d));
  c = csub(c, cscl(fz, d));
}

__device__ __forceinline__ double abs_or_inf(double a) {
  return __builtin_isnan(a) ? __builtin_inf() : fabs(a);
}
__device__ __forceinline__ float abs_or_inf(float a) {
  return __builtin_isnan(a) ? __builtin_inff() : fabsf(a);
}

// Fill a flow's FP64 parameter block (PrmLayout) for step dt and band growth rate amp, with
// the threads [tid, nthreads) of the caller: t_1[m] = exp(-nu m^2 dt), t_1/2[m] =
// exp(-nu m^2 dt/2) (src/pyspectral/backends/_fused_common.py:10-19) and the drag/band
// scalars.  The flow is marked active.
__device__ __forceinline__ void write_params(double* prm, int nh, double nu, double dt,
                                             double alpha, double amp, double klo2,
                                             double khi2, double ka2, int tid, int nthreads) {
  for (int m = tid; m < nh; m += nthreads) {
    const double dm = (double)m;
    prm[m] = exp(-nu * dm * dm * dt);
    prm[nh + m] = exp(-nu * dm * dm * 0.5 * dt);
  }
  if (tid == 0) {
    double* s = prm + 2 * nh;
    s[PrmLayout::kDt] = dt;
    s[PrmLayout::kEa1] = exp(-alpha * dt);
    s[PrmLayout::kEah] = exp(-alpha * 0.5 * dt);
    s[PrmLayout::kEb1] = exp(amp * dt);
    s[PrmLayout::kEbh] = exp(amp * 0.5 * dt);
    s[PrmLayout::kKlo2] = klo2;
    s[PrmLayout::kKhi2] = khi2;
    s[PrmLayout::kKa2] = ka2;
    s[PrmLayout::kBand] = amp != 0.0 ? 1.0 : 0.0;
    s[PrmLayout::kActive] = 1.0;
  }
}

// Stage kernels return at once for a flow whose parameter block marks it inactive: a run of
// the device controller that has reached its end turns the graphs queued behind it into
// no-ops.
__device__ __forceinline__ bool flow_inactive(const double* prm, int nh) {
  return prm[2 * nh + PrmLayout::kActive] == 0.0;
}

}  // namespace psx
