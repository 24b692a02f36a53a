// 3D IF-RK4 stage kernels (velocity form with Leray projection).
//
// Forms (reference: src/pyspectral/kernels/spectral.cu:137-238, numpy_ref.py:110-119):
//   basdevant  3 inverse fields (u,v,w), 5 forward fields T11=uu-q, T22=vv-q, T12=uv, T13=uw,
//              T23=vw with q = |u|^2/3 (T33 = -T11-T22); N_i = -i k_j T_ij.
//   rotational 6 inverse fields (u, omega = i k x u), 3 forward fields u x omega.
// Then N is masked (2/3 rule) and projected, N - k (k.N)/k^2, and the Lawson stage update
// runs per component.
//
// One RK stage = line kernel (z: c2r, products, r2c; real pairs packed as one complex line)
// -> y-pass forward -> x-pass (forward, physics, inverse) -> y-pass inverse.
#include "psx_kernels.h"

namespace psx {

template <typename T, int E>
__device__ __forceinline__ void line_load_pair(cx<T> (&v)[E], const Flow<T>& f, int a, int b,
                                               int row0, int nrows, int tid, int NT,
                                               const RowsLayout& lay) {
  const int L = f.n, jm = f.jm;
  const cx<T>* A = f.W + a * f.fstride;
  const cx<T>* B = f.W + (b >= 0 ? b : 0) * f.fstride;
  const int R = first_radix(f.prow, true);
#pragma unroll
  for (int i = 0; i < E; ++i) {
    int e, t;
    reg_elem(i, R, L, tid, NT, lay, e, t);
    const int row = row0 + t;
    cx<T> z = czero<T>();
    if (row < nrows) {
      const long long base = (long long)row * f.ldw;
      int src = -1;
      bool cj = false;
      if (e < jm) {
        src = e;
      } else if (e > L - jm) {
        src = L - e;
        cj = true;
      }
      if (src >= 0) {
        cx<T> p = A[base + src];
        cx<T> q = b >= 0 ? B[base + src] : czero<T>();
        if (src == 0) {
          p.y = (T)0;
          q.y = (T)0;
        }
        if (cj) {
          p = conjg(p);
          q = conjg(q);
        }
        z = mkc(p.x - q.y, p.y + q.x);  // P + iQ
      }
    }
    v[i] = z;
  }
}

// Split the forward transform of P + iQ (in registers, last-radix order) into the half
// spectra of P and Q (columns j < jm) through LDS.
template <typename T, int E>
__device__ __forceinline__ void line_store_pair(cx<T> (&v)[E], const Flow<T>& f, cx<T>* lds,
                                                int a, int b, int row0, int nrows, int tid,
                                                int NT, const RowsLayout& lay) {
  const int L = f.n, jm = f.jm;
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
  cx<T>* A = f.W + a * f.fstride;
  cx<T>* B = f.W + b * f.fstride;
  const int total = f.rowT * jm;
  const T h = (T)0.5;
  for (int idx = tid; idx < total; idx += NT) {
    const int t = idx / jm;
    const int k = idx - t * jm;
    const int row = row0 + t;
    if (row >= nrows) continue;
    const cx<T> z = lds[lay.addr(k, t)];
    const cx<T> zc = conjg(lds[lay.addr(k == 0 ? 0 : L - k, t)]);
    const long long o = (long long)row * f.ldw + k;
    A[o] = cscl(h, cadd(z, zc));
    B[o] = cscl(h, mul_mi(csub(z, zc)));
  }
}

// The forward transform of a real line (imaginary part zero) is its r2c spectrum directly.
template <typename T, int E>
__device__ __forceinline__ void line_store_real(const cx<T> (&v)[E], const Flow<T>& f, int a,
                                                int row0, int nrows, int tid, int NT,
                                                const RowsLayout& lay) {
  const int L = f.n, jm = f.jm;
  cx<T>* A = f.W + a * f.fstride;
  const int R = last_radix(f.prow, false);
#pragma unroll
  for (int i = 0; i < E; ++i) {
    int e, t;
    reg_elem(i, R, L, tid, NT, lay, e, t);
    if (e < jm && row0 + t < nrows) A[(long long)(row0 + t) * f.ldw + e] = v[i];
  }
}

template <typename T, int E>
__global__ void __launch_bounds__(512, 1) k_line3d(Flow<T> f, int flags) {
  if (flow_inactive(f.prm, f.nh)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x, NT = blockDim.x;
  const RowsLayout lay{rows_pitch(f.n)};
  const int row0 = blockIdx.x * f.rowT;
  const int nrows = f.n * f.n;
  const bool want_max = flags & kLineWantMax;
  const int R = last_radix(f.prow, true);

  if (!f.rot) {
    cx<T> z1[E], z2[E];
    line_load_pair<T, E>(z1, f, 0, 1, row0, nrows, tid, NT, lay);
    fft_exec<E, true>(z1, lds, f.prow, f.twp, lay, tid, NT);
    line_load_pair<T, E>(z2, f, 2, -1, row0, nrows, tid, NT, lay);
    fft_exec<E, true>(z2, lds, f.prow, f.twp, lay, tid, NT);
    if (want_max) {
      double m[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int i = 0; i < E; ++i) {
        int e, t;
        reg_elem(i, R, f.n, tid, NT, lay, e, t);
        if (row0 + t < nrows) {
          m[0] = fmax(m[0], abs_or_inf((double)z1[i].x));
          m[1] = fmax(m[1], abs_or_inf((double)z1[i].y));
          m[2] = fmax(m[2], abs_or_inf((double)z2[i].x));
        }
      }
      block_max_to<3>(m, f.stats, reinterpret_cast<double*>(smem_raw));
    }
    if (flags & kLineMaxOnly) return;
    cx<T> z3[E];
    const T third = (T)1 / (T)3;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const T u = z1[i].x, v = z1[i].y, w = z2[i].x;
      const T q = (u * u + v * v + w * w) * third;
      z1[i] = mkc(u * u - q, v * v - q);
      z2[i] = mkc(u * v, u * w);
      z3[i] = mkc(v * w, (T)0);
    }
    fft_exec<E, false>(z1, lds, f.prow, f.twp, lay, tid, NT);
    line_store_pair<T, E>(z1, f, lds, 0, 1, row0, nrows, tid, NT, lay);
    fft_exec<E, false>(z2, lds, f.prow, f.twp, lay, tid, NT);
    line_store_pair<T, E>(z2, f, lds, 2, 3, row0, nrows, tid, NT, lay);
    fft_exec<E, false>(z3, lds, f.prow, f.twp, lay, tid, NT);
    line_store_real<T, E>(z3, f, 4, row0, nrows, tid, NT, lay);
  } else {
    cx<T> z1[E], z2[E], z3[E];
    line_load_pair<T, E>(z1, f, 0, 1, row0, nrows, tid, NT, lay);
    fft_exec<E, true>(z1, lds, f.prow, f.twp, lay, tid, NT);
    line_load_pair<T, E>(z2, f, 2, 3, row0, nrows, tid, NT, lay);
    fft_exec<E, true>(z2, lds, f.prow, f.twp, lay, tid, NT);
    line_load_pair<T, E>(z3, f, 4, 5, row0, nrows, tid, NT, lay);
    fft_exec<E, true>(z3, lds, f.prow, f.twp, lay, tid, NT);
    if (want_max) {
      double m[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int i = 0; i < E; ++i) {
        int e, t;
        reg_elem(i, R, f.n, tid, NT, lay, e, t);
        if (row0 + t < nrows) {
          m[0] = fmax(m[0], abs_or_inf((double)z1[i].x));
          m[1] = fmax(m[1], abs_or_inf((double)z1[i].y));
          m[2] = fmax(m[2], abs_or_inf((double)z2[i].x));
        }
      }
      block_max_to<3>(m, f.stats, reinterpret_cast<double*>(smem_raw));
    }
    if (flags & kLineMaxOnly) return;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const T u = z1[i].x, v = z1[i].y, w = z2[i].x;
      const T ox = z2[i].y, oy = z3[i].x, oz = z3[i].y;
      z1[i] = mkc(v * oz - w * oy, w * ox - u * oz);
      z2[i] = mkc(u * oy - v * ox, (T)0);
    }
    fft_exec<E, false>(z1, lds, f.prow, f.twp, lay, tid, NT);
    line_store_pair<T, E>(z1, f, lds, 0, 1, row0, nrows, tid, NT, lay);
    fft_exec<E, false>(z2, lds, f.prow, f.twp, lay, tid, NT);
    line_store_real<T, E>(z2, f, 2, row0, nrows, tid, NT, lay);
  }
}

// Middle axis (y) of nfld fields: for fixed x, column tile j0..j0+yT, rows x n + y.
template <typename T, int E, bool INV>
__global__ void __launch_bounds__(512, 1) k_ypass3d(Flow<T> f, int nfld) {
  if (flow_inactive(f.prm, f.nh)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x, NT = blockDim.x;
  const ColsLayout lay{f.yT, __ffs(f.yT) - 1};
  const int njt = f.ldw / f.yT;
  const int x = blockIdx.x / njt;
  const int j0 = (blockIdx.x - x * njt) * f.yT;
  const int L = f.n;
  const long long rb = (long long)x * f.n;
  for (int fld = 0; fld < nfld; ++fld) {
    cx<T>* Wf = f.W + fld * f.fstride;
    cx<T> v[E];
    {
      const int R = first_radix(f.py, INV);
#pragma unroll
      for (int i = 0; i < E; ++i) {
        int e, t;
        reg_elem(i, R, L, tid, NT, lay, e, t);
        v[i] = Wf[(rb + e) * f.ldw + j0 + t];
      }
    }
    fft_exec<E, INV>(v, lds, f.py, f.twp, lay, tid, NT);
    {
      const int R = last_radix(f.py, INV);
#pragma unroll
      for (int i = 0; i < E; ++i) {
        int e, t;
        reg_elem(i, R, L, tid, NT, lay, e, t);
        Wf[(rb + e) * f.ldw + j0 + t] = v[i];
      }
    }
  }
}

// Outer axis (x) with the spectral physics: for fixed y-wavenumber index, column tile
// j0..j0+colT, rows x n + ky_idx.
template <typename T, int E>
__global__ void __launch_bounds__(512, 1) k_spec3d(Flow<T> f, int stage) {
  if (flow_inactive(f.prm, f.nh)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x, NT = blockDim.x;
  const ColsLayout lay{f.colT, __ffs(f.colT) - 1};
  const int njt = f.ldw / f.colT;
  const int iy = blockIdx.x / njt;
  const int j0 = (blockIdx.x - iy * njt) * f.colT;
  const int n = f.n, nh = f.nh, L = f.n;
  const int ky = iy < n / 2 ? iy : iy - n;
  const int ay = ky < 0 ? -ky : ky;
  const long long sst = f.sstride;

  cx<T> a0[E], a1[E], a2[E];
  const int R = last_radix(f.pcol, false);  // register order between the transforms
  if (stage != kStagePre) {
    const int nf = f.rot ? 3 : 5;
#pragma unroll
    for (int i = 0; i < E; ++i) a0[i] = a1[i] = a2[i] = czero<T>();
    for (int c = 0; c < nf; ++c) {
      const cx<T>* Wf = f.W + c * f.fstride;
      cx<T> v[E];
      {
        const int R0 = first_radix(f.pcol, false);
#pragma unroll
        for (int i = 0; i < E; ++i) {
          int e, t;
          reg_elem(i, R0, L, tid, NT, lay, e, t);
          v[i] = Wf[((long long)e * n + iy) * f.ldw + j0 + t];
        }
      }
      fft_exec<E, false>(v, lds, f.pcol, f.twp, lay, tid, NT);
#pragma unroll
      for (int i = 0; i < E; ++i) {
        if (f.rot) {
          if (c == 0) a0[i] = v[i];
          else if (c == 1) a1[i] = v[i];
          else a2[i] = v[i];
          continue;
        }
        int e, t;
        reg_elem(i, R, L, tid, NT, lay, e, t);
        const T kx = (T)(e < n / 2 ? e : e - n), fy = (T)ky, kz = (T)(j0 + t);
        const cx<T> mt = mul_mi(v[i]);  // -i T_c
        // N_i = -i k_j T_ij with T33 = -T11 - T22
        if (c == 0) {         // T11
          a0[i] = cadd(a0[i], cscl(kx, mt));
          a2[i] = csub(a2[i], cscl(kz, mt));
        } else if (c == 1) {  // T22
          a1[i] = cadd(a1[i], cscl(fy, mt));
          a2[i] = csub(a2[i], cscl(kz, mt));
        } else if (c == 2) {  // T12
          a0[i] = cadd(a0[i], cscl(fy, mt));
          a1[i] = cadd(a1[i], cscl(kx, mt));
        } else if (c == 3) {  // T13
          a0[i] = cadd(a0[i], cscl(kz, mt));
          a2[i] = cadd(a2[i], cscl(kx, mt));
        } else {              // T23
          a1[i] = cadd(a1[i], cscl(kz, mt));
          a2[i] = cadd(a2[i], cscl(fy, mt));
        }
      }
    }
  }
  {
    const IfScalars s = load_if_scalars(f.prm, nh);
    const T dt = (T)f.prm[2 * nh + PrmLayout::kDt];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      int e, t;
      reg_elem(i, R, L, tid, NT, lay, e, t);
      const int kx = e < n / 2 ? e : e - n;
      const int ax = kx < 0 ? -kx : kx;
      const int kz = j0 + t;
      cx<T> ua = czero<T>(), ub = czero<T>(), uc = czero<T>();
      const long long k2 = (long long)kx * kx + (long long)ky * ky + (long long)kz * kz;
      // retained modes: the 2/3 rule's cube, or the phase shift's ball 9 |k|^2 < 2 n^2
      const bool keep = f.ps ? 9 * k2 < 2LL * n * n : (3 * ax < n && 3 * ay < n && kz < f.jm);
      if (keep) {
        const long long si = ((long long)e * n + iy) * f.sp + kz;
        if (stage == kStagePre) {
          ua = f.u0[si];
          ub = f.u0[sst + si];
          uc = f.u0[2 * sst + si];
          const T ik2 = k2 > 0 ? (T)1 / (T)k2 : (T)0;
          leray<T>((T)kx, (T)ky, (T)kz, ik2, ua, ub, uc);
          f.u0[si] = ua;
          f.u0[sst + si] = ub;
          f.u0[2 * sst + si] = uc;
        } else {
          const T fx = (T)kx, fy = (T)ky, fz = (T)kz;
          const T ik2 = k2 > 0 ? (T)1 / (T)k2 : (T)0;
          if (f.src) {
            a0[i] = cadd(a0[i], f.src[si]);
            a1[i] = cadd(a1[i], f.src[sst + si]);
            a2[i] = cadd(a2[i], f.src[2 * sst + si]);
          }
          if (f.ps) {  // back from the shifted grid (the projection below commutes with it)
            const cx<T> ph = shift_phase<T>(shift_of_term(f.prm, nh, stage), kx, ky, kz, -1.0);
            a0[i] = cmul(a0[i], ph);
            a1[i] = cmul(a1[i], ph);
            a2[i] = cmul(a2[i], ph);
          }
          const cx<T> d = cscl(ik2, cadd(cadd(cscl(fx, a0[i]), cscl(fy, a1[i])), cscl(fz, a2[i])));
          const cx<T> n0 = csub(a0[i], cscl(fx, d));
          const cx<T> n1 = csub(a1[i], cscl(fy, d));
          const cx<T> n2 = csub(a2[i], cscl(fz, d));
          double E1, Eh;
          if_factors(f.prm, nh, s, ax, ay, kz, k2, E1, Eh);
          ua = rk_update<T>(stage, dt, (T)(1.0 - E1), (T)(1.0 - Eh), n0, f.u0 + si, f.acc + si);
          ub = rk_update<T>(stage, dt, (T)(1.0 - E1), (T)(1.0 - Eh), n1, f.u0 + sst + si, f.acc + sst + si);
          uc = rk_update<T>(stage, dt, (T)(1.0 - E1), (T)(1.0 - Eh), n2, f.u0 + 2 * sst + si,
                            f.acc + 2 * sst + si);
          if (stage == 3) {  // the new state, projected (rk_update stored it unprojected)
            leray<T>(fx, fy, fz, ik2, ua, ub, uc);
            f.u0[si] = ua;
            f.u0[sst + si] = ub;
            f.u0[2 * sst + si] = uc;
          }
        }
        ua = cscl(f.scale, ua);
        ub = cscl(f.scale, ub);
        uc = cscl(f.scale, uc);
        if (f.ps) {  // the next stage's inputs on its shifted grid
          const cx<T> ph = shift_phase<T>(shift_of_inputs(f.prm, nh, stage), kx, ky, kz, 1.0);
          ua = cmul(ua, ph);
          ub = cmul(ub, ph);
          uc = cmul(uc, ph);
        }
      }
      a0[i] = ua;
      a1[i] = ub;
      a2[i] = uc;
    }
  }
  const int nin = f.rot ? 6 : 3;
  for (int c = 0; c < nin; ++c) {
    cx<T> v[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      if (c == 0) {
        v[i] = a0[i];
      } else if (c == 1) {
        v[i] = a1[i];
      } else if (c == 2) {
        v[i] = a2[i];
      } else {
        int e, t;
        reg_elem(i, R, L, tid, NT, lay, e, t);
        const T fx = (T)(e < n / 2 ? e : e - n), fy = (T)ky, fz = (T)(j0 + t);
        // omega = i k x u
        if (c == 3) v[i] = mul_i(csub(cscl(fy, a2[i]), cscl(fz, a1[i])));
        else if (c == 4) v[i] = mul_i(csub(cscl(fz, a0[i]), cscl(fx, a2[i])));
        else v[i] = mul_i(csub(cscl(fx, a1[i]), cscl(fy, a0[i])));
      }
    }
    fft_exec<E, true>(v, lds, f.pcol, f.twp, lay, tid, NT);
    cx<T>* Wf = f.W + c * f.fstride;
    const int Ri = last_radix(f.pcol, true);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      int e, t;
      reg_elem(i, Ri, L, tid, NT, lay, e, t);
      Wf[((long long)e * n + iy) * f.ldw + j0 + t] = v[i];
    }
  }
}

// ---------------------------------------------------------------------------------- launchers
template <typename T>
hipError_t launch_line3d(const Flow<T>& f, const Launch& lc, int flags, hipStream_t st) {
  switch (lc.E) {
#define PSX_CASE(EV)                                                                 \
  case EV: if constexpr (e_ok_3d(EV)) {                                                                         \
    auto fn = k_line3d<T, EV>;                                                       \
    static const bool attr_ = set_lds_max((const void*)fn); (void)attr_;                                          \
    hipLaunchKernelGGL(fn, dim3(lc.blocks), dim3(lc.threads), lc.lds, st, f, flags); \
    break;                                                                           \
  } else { return hipErrorInvalidValue; }
    PSX_E_CASES
#undef PSX_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_ypass3d(const Flow<T>& f, const Launch& lc, bool inv, int nfld,
                          hipStream_t st) {
  switch (lc.E) {
#define PSX_CASE(EV)                                                                  \
  case EV: if constexpr (e_ok_3d(EV)) {                                                                          \
    if (inv) {                                                                        \
      auto fn = k_ypass3d<T, EV, true>;                                               \
      static const bool attr_ = set_lds_max((const void*)fn); (void)attr_;                                         \
      hipLaunchKernelGGL(fn, dim3(lc.blocks), dim3(lc.threads), lc.lds, st, f, nfld); \
    } else {                                                                          \
      auto fn = k_ypass3d<T, EV, false>;                                              \
      static const bool attr_ = set_lds_max((const void*)fn); (void)attr_;                                         \
      hipLaunchKernelGGL(fn, dim3(lc.blocks), dim3(lc.threads), lc.lds, st, f, nfld); \
    }                                                                                 \
    break;                                                                            \
  } else { return hipErrorInvalidValue; }
    PSX_E_CASES
#undef PSX_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_spec3d(const Flow<T>& f, const Launch& lc, int stage, hipStream_t st) {
  switch (lc.E) {
#define PSX_CASE(EV)                                                                 \
  case EV: if constexpr (e_ok_3d(EV)) {                                                                         \
    auto fn = k_spec3d<T, EV>;                                                       \
    static const bool attr_ = set_lds_max((const void*)fn); (void)attr_;                                          \
    hipLaunchKernelGGL(fn, dim3(lc.blocks), dim3(lc.threads), lc.lds, st, f, stage); \
    break;                                                                           \
  } else { return hipErrorInvalidValue; }
    PSX_E_CASES
#undef PSX_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template hipError_t launch_line3d<float>(const Flow<float>&, const Launch&, int, hipStream_t);
template hipError_t launch_line3d<double>(const Flow<double>&, const Launch&, int, hipStream_t);
template hipError_t launch_ypass3d<float>(const Flow<float>&, const Launch&, bool, int,
                                          hipStream_t);
template hipError_t launch_ypass3d<double>(const Flow<double>&, const Launch&, bool, int,
                                           hipStream_t);
template hipError_t launch_spec3d<float>(const Flow<float>&, const Launch&, int, hipStream_t);
template hipError_t launch_spec3d<double>(const Flow<double>&, const Launch&, int, hipStream_t);

}  // namespace psx
