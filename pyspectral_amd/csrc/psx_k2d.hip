// 2D IF-RK4 stage kernels (vorticity form, Basdevant nonlinear term: 4 real FFTs per stage).
//
// One RK stage = line kernel + spectral kernel (one kernel when the full axis fits one
// work-group tile, otherwise outer-forward + inner-spectral + outer-inverse four-step
// kernels).  The reference runs cuFFT C2R -> phys2d -> cuFFT R2C -> post2d per stage
// (src/pyspectral/backends/cupy_cuda.py:163-173, src/pyspectral/kernels/spectral.cu:93-134);
// here the transforms are the project's own and every pointwise phase lives inside them:
//
//   line kernel (one HBM pass): per x row, c2r of (u_hat, v_hat) packed as one complex
//     sequence Z = U + iV, products P = (v-u)(v+u), Q = u v in registers, r2c of P + iQ,
//     split into P_hat, Q_hat, write only the retained half-axis columns j < jm.
//   spectral kernel: per column tile, forward transform along x of (P_hat, Q_hat), the
//     nonlinear term N = kx ky P + (kx^2 - ky^2) Q with the 2/3 mask, the FP64 integrating
//     factor, the Lawson stage update of u0/acc, the next stage's (u_hat, v_hat) =
//     (i ky, -i kx) omega/k^2 * 1/n^2, and its inverse transform along x.
#include "psx_kernels.h"

namespace psx {

// ------------------------------------------------------------------------------ line kernel
template <typename T, int E>
__global__ void __launch_bounds__(512, 1) k_line2d(Flow<T> f, int flags) {
  if (flow_inactive(f.prm, f.nh)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x, NT = blockDim.x;
  const int L = f.n;
  const RowsLayout lay{rows_pitch(L)};
  const int row0 = blockIdx.x * f.rowT;
  const int nrows = f.n;
  const int jm = f.jm;
  cx<T>* U = f.W;
  cx<T>* V = f.W + f.fstride;

  cx<T> v[E];
  {
    const int R = first_radix(f.prow, true);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      int e, t;
      reg_elem(i, R, L, tid, NT, lay, e, t);
      const int row = row0 + t;
      cx<T> z = czero<T>();
      if (row < nrows) {
        const long long base = (long long)row * f.ldw;
        if (e == 0) {
          z = mkc(U[base].x, V[base].x);  // imaginary parts of the k=0 column are dropped (c2r)
        } else if (e < jm) {
          const cx<T> a = U[base + e], b = V[base + e];
          z = mkc(a.x - b.y, a.y + b.x);  // U + iV
        } else if (e > L - jm) {
          const cx<T> a = U[base + (L - e)], b = V[base + (L - e)];
          z = mkc(a.x + b.y, b.x - a.y);  // conj(U) + i conj(V)
        }
      }
      v[i] = z;
    }
  }
  fft_exec<E, true>(v, lds, f.prow, f.twp, lay, tid, NT);

  if (flags & kLineWantMax) {
    double m[2] = {0.0, 0.0};
    const int R = last_radix(f.prow, true);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      int e, t;
      reg_elem(i, R, L, tid, NT, lay, e, t);
      if (row0 + t < nrows) {
        m[0] = fmax(m[0], abs_or_inf((double)v[i].x));
        m[1] = fmax(m[1], abs_or_inf((double)v[i].y));
      }
    }
    block_max_to<2>(m, f.stats, reinterpret_cast<double*>(smem_raw));
  }
  if (flags & kLineMaxOnly) return;

#pragma unroll
  for (int i = 0; i < E; ++i) {
    const T u = v[i].x, w = v[i].y;
    v[i] = mkc((w - u) * (w + u), u * w);
  }
  fft_exec<E, false>(v, lds, f.prow, f.twp, lay, tid, NT);

  __syncthreads();
  {
    const int R = last_radix(f.prow, false);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      int e, t;
      reg_elem(i, R, L, tid, NT, lay, e, t);
      lds[lay.addr(e, t)] = v[i];
    }
  }
  __syncthreads();
  const int total = f.rowT * jm;
  const T h = (T)0.5;
  for (int idx = tid; idx < total; idx += NT) {
    const int t = idx / jm;
    const int k = idx - t * jm;
    const int row = row0 + t;
    if (row >= nrows) continue;
    const cx<T> z = lds[lay.addr(k, t)];
    const cx<T> zc = conjg(lds[lay.addr(k == 0 ? 0 : L - k, t)]);
    const long long a = (long long)row * f.ldw + k;
    U[a] = cscl(h, cadd(z, zc));           // FFT of the real part
    V[a] = cscl(h, mul_mi(csub(z, zc)));   // FFT of the imaginary part
  }
}

// -------------------------------------------------------------------------- spectral kernel
// FOUR = false: one work-group holds whole columns (length n).  FOUR = true: the inner step
// of the four-step split n = n1 n2; the group of rows g*n2 .. g*n2 + n2-1 holds, for the
// outer index k1 = g, the modes k = k1 + n1 k2 after the forward transform.
template <typename T, int E, bool FOUR>
__global__ void __launch_bounds__(512, 1) k_spec2d(Flow<T> f, int stage) {
  if (flow_inactive(f.prm, f.nh)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x, NT = blockDim.x;
  const ColsLayout lay{f.colT, __ffs(f.colT) - 1};
  const int njt = f.ldw / f.colT;
  const int grp = blockIdx.x / njt;
  const int j0 = (blockIdx.x - grp * njt) * f.colT;
  const int L = FOUR ? f.n2 : f.n;
  const int eo = FOUR ? grp * f.n2 : 0;
  const int n = f.n, nh = f.nh;
  cx<T>* W0 = f.W;
  cx<T>* W1 = f.W + f.fstride;

  cx<T> va[E], vb[E];
  if (stage != kStagePre) {
    const int R = first_radix(f.pcol, false);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      int e, t;
      reg_elem(i, R, L, tid, NT, lay, e, t);
      const long long a = (long long)(eo + e) * f.ldw + j0 + t;
      va[i] = W0[a];
      vb[i] = W1[a];
    }
    fft_exec<E, false>(va, lds, f.pcol, f.twp, lay, tid, NT);
    fft_exec<E, false>(vb, lds, f.pcol, f.twp, lay, tid, NT);
  }
  {
    const int R = last_radix(f.pcol, false);
    const IfScalars s = load_if_scalars(f.prm, nh);
    const T dt = (T)f.prm[2 * nh + PrmLayout::kDt];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      int e, t;
      reg_elem(i, R, L, tid, NT, lay, e, t);
      const int k = FOUR ? grp + f.n1 * e : e;
      const int kx = k < n / 2 ? k : k - n;
      const int ky = j0 + t;
      const int ax = kx < 0 ? -kx : kx;
      cx<T> ou = czero<T>(), ov = czero<T>();
      const long long k2 = (long long)kx * kx + (long long)ky * ky;
      // retained modes: the 2/3 rule's square, or the phase shift's disc 9 |k|^2 < 2 n^2
      const bool keep = f.ps ? 9 * k2 < 2LL * n * n : (3 * ax < n && ky < f.jm);
      if (keep) {
        const long long si = (long long)k * f.sp + ky;
        cx<T> nxt;
        if (stage == kStagePre) {
          nxt = f.u0[si];
        } else {
          const T fx = (T)kx, fy = (T)ky;
          cx<T> N = cadd(cscl(fx * fy, va[i]), cscl(fx * fx - fy * fy, vb[i]));
          if (f.ps) N = cmul(N, shift_phase<T>(shift_of_term(f.prm, nh, stage), kx, ky, 0, -1.0));
          double E1, Eh;
          if_factors(f.prm, nh, s, ax, ky, -1, k2, E1, Eh);
          nxt = rk_update<T>(stage, dt, (T)(1.0 - E1), (T)(1.0 - Eh), N, f.u0 + si, f.acc + si);
        }
        const T ik2 = k2 > 0 ? f.scale / (T)k2 : (T)0;
        cx<T> psi = cscl(ik2, nxt);
        if (f.ps) psi = cmul(psi, shift_phase<T>(shift_of_inputs(f.prm, nh, stage), kx, ky, 0, 1.0));
        ou = cscl((T)ky, mul_i(psi));
        ov = cscl((T)kx, mul_mi(psi));
      }
      va[i] = ou;
      vb[i] = ov;
    }
  }
  fft_exec<E, true>(va, lds, f.pcol, f.twp, lay, tid, NT);
  {
    const int R = last_radix(f.pcol, true);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      int e, t;
      reg_elem(i, R, L, tid, NT, lay, e, t);
      cx<T> z = va[i];
      if (FOUR) z = cmulc(z, f.tw[e * grp]);
      W0[(long long)(eo + e) * f.ldw + j0 + t] = z;
    }
  }
  fft_exec<E, true>(vb, lds, f.pcol, f.twp, lay, tid, NT);
  {
    const int R = last_radix(f.pcol, true);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      int e, t;
      reg_elem(i, R, L, tid, NT, lay, e, t);
      cx<T> z = vb[i];
      if (FOUR) z = cmulc(z, f.tw[e * grp]);
      W1[(long long)(eo + e) * f.ldw + j0 + t] = z;
    }
  }
}

// ---------------------------------------------------------------------- outer four-step pass
// Rows x = n2 x1 + x2 for fixed x2 (= group).  Forward: transform over x1 -> k1, multiply by
// w_n^{x2 k1}, store at row k1 n2 + x2.  Inverse: transform over k1 -> x1 (the conjugate
// twiddle was applied by the inner kernel), store at row n2 x1 + x2.  Both work fields.
template <typename T, int E, bool INV>
__global__ void __launch_bounds__(512, 1) k_outer2d(Flow<T> f) {
  if (flow_inactive(f.prm, f.nh)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x, NT = blockDim.x;
  const ColsLayout lay{f.colT, __ffs(f.colT) - 1};
  const int njt = f.ldw / f.colT;
  const int x2 = blockIdx.x / njt;
  const int j0 = (blockIdx.x - x2 * njt) * f.colT;
  const int L = f.n1;
  for (int fld = 0; fld < 2; ++fld) {
    cx<T>* Wf = f.W + fld * f.fstride;
    cx<T> v[E];
    {
      const int R = first_radix(f.pa, INV);
#pragma unroll
      for (int i = 0; i < E; ++i) {
        int e, t;
        reg_elem(i, R, L, tid, NT, lay, e, t);
        v[i] = Wf[(long long)(e * f.n2 + x2) * f.ldw + j0 + t];
      }
    }
    fft_exec<E, INV>(v, lds, f.pa, f.twp, lay, tid, NT);
    {
      const int R = last_radix(f.pa, INV);
#pragma unroll
      for (int i = 0; i < E; ++i) {
        int e, t;
        reg_elem(i, R, L, tid, NT, lay, e, t);
        cx<T> z = v[i];
        if (!INV) z = cmul(z, f.tw[e * x2]);
        Wf[(long long)(e * f.n2 + x2) * f.ldw + j0 + t] = z;
      }
    }
  }
}

// ---------------------------------------------------------------------------------- launchers
template <typename T>
hipError_t launch_line2d(const Flow<T>& f, const Launch& lc, int flags, hipStream_t st) {
  switch (lc.E) {
#define PSX_CASE(EV)                                                                      \
  case EV: if constexpr (e_ok_2d<T>(EV)) {                                                                              \
    auto fn = k_line2d<T, EV>;                                                            \
    static const bool attr_ = set_lds_max((const void*)fn); (void)attr_;                                         \
    hipLaunchKernelGGL(fn, dim3(lc.blocks), dim3(lc.threads), lc.lds, st, f, flags);      \
    break;                                                                                \
  } else { return hipErrorInvalidValue; }
    PSX_E_CASES
#undef PSX_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_spec2d(const Flow<T>& f, const Launch& lc, bool four, int stage,
                         hipStream_t st) {
  switch (lc.E) {
#define PSX_CASE(EV)                                                                      \
  case EV: if constexpr (e_ok_2d<T>(EV)) {                                                                              \
    if (four) {                                                                           \
      auto fn = k_spec2d<T, EV, true>;                                                    \
      static const bool attr_ = set_lds_max((const void*)fn); (void)attr_;                                       \
      hipLaunchKernelGGL(fn, dim3(lc.blocks), dim3(lc.threads), lc.lds, st, f, stage);    \
    } else {                                                                              \
      auto fn = k_spec2d<T, EV, false>;                                                   \
      static const bool attr_ = set_lds_max((const void*)fn); (void)attr_;                                       \
      hipLaunchKernelGGL(fn, dim3(lc.blocks), dim3(lc.threads), lc.lds, st, f, stage);    \
    }                                                                                     \
    break;                                                                                \
  } else { return hipErrorInvalidValue; }
    PSX_E_CASES
#undef PSX_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_outer2d(const Flow<T>& f, const Launch& lc, bool inv, hipStream_t st) {
  switch (lc.E) {
#define PSX_CASE(EV)                                                                      \
  case EV: if constexpr (e_ok_2d<T>(EV)) {                                                                              \
    if (inv) {                                                                            \
      auto fn = k_outer2d<T, EV, true>;                                                   \
      static const bool attr_ = set_lds_max((const void*)fn); (void)attr_;                                       \
      hipLaunchKernelGGL(fn, dim3(lc.blocks), dim3(lc.threads), lc.lds, st, f);           \
    } else {                                                                              \
      auto fn = k_outer2d<T, EV, false>;                                                  \
      static const bool attr_ = set_lds_max((const void*)fn); (void)attr_;                                       \
      hipLaunchKernelGGL(fn, dim3(lc.blocks), dim3(lc.threads), lc.lds, st, f);           \
    }                                                                                     \
    break;                                                                                \
  } else { return hipErrorInvalidValue; }
    PSX_E_CASES
#undef PSX_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template hipError_t launch_line2d<float>(const Flow<float>&, const Launch&, int, hipStream_t);
template hipError_t launch_line2d<double>(const Flow<double>&, const Launch&, int, hipStream_t);
template hipError_t launch_spec2d<float>(const Flow<float>&, const Launch&, bool, int,
                                         hipStream_t);
template hipError_t launch_spec2d<double>(const Flow<double>&, const Launch&, bool, int,
                                          hipStream_t);
template hipError_t launch_outer2d<float>(const Flow<float>&, const Launch&, bool, hipStream_t);
template hipError_t launch_outer2d<double>(const Flow<double>&, const Launch&, bool,
                                           hipStream_t);

}  // namespace psx
