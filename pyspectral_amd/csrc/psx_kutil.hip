// Parameter, fill and FP64 diagnostic kernels.
//
// Diagnostics restate src/pyspectral/diagnostics.py:51-97 on the device, accumulating in FP64
// whatever the working precision: density |u_hat|^2 * w_h / n^(2d) with the Hermitian weight
// w_h = 1 on the first and last half-axis plane and 2 elsewhere; 2D energy divides by k^2.
// The four scalar sums are reduced deterministically (fixed per-block partials, then one
// block), the shell spectrum with FP64 atomics per shell.
#include "psx_kernels.h"

namespace psx {

__global__ void k_params(double* prm, int nh, double nu, double dt, double alpha, double amp,
                         double klo2, double khi2, double ka2) {
  write_params(prm, nh, nu, dt, alpha, amp, klo2, khi2, ka2, threadIdx.x, blockDim.x);
}

hipError_t launch_params(double* prm, int nh, double nu, double dt, double alpha, double amp,
                         double klo2, double khi2, double ka2, hipStream_t st) {
  hipLaunchKernelGGL(k_params, dim3(1), dim3(256), 0, st, prm, nh, nu, dt, alpha, amp, klo2,
                     khi2, ka2);
  return hipGetLastError();
}

__global__ void k_fill_u32(unsigned* p, unsigned v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f64(double* p, double v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

struct F64x16 {
  double v[16];
};
__global__ void k_set_f64(double* p, F64x16 v, int n) {
  const int i = threadIdx.x;
  if (i < n) p[i] = v.v[i];
}
hipError_t launch_set_f64(double* p, const double* v, int n, hipStream_t st) {
  if (n < 1 || n > 16) return hipErrorInvalidValue;
  F64x16 a{};
  for (int i = 0; i < n; ++i) a.v[i] = v[i];
  hipLaunchKernelGGL(k_set_f64, dim3(1), dim3(16), 0, st, p, a, n);
  return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned* p, unsigned v, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}
hipError_t launch_fill_f64(double* p, double v, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_f64, dim3((n + 255) / 256), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}

// Hermitian projection of the self-mirrored half-axis plane (3D: kz = 0 plane of every
// component; 2D: the ky = 0 column): u(k) <- (u(k) + conj u(-k)) / 2 and u(-k) <- conj of
// that, by the one thread that owns the pair.  A real field has no anti-Hermitian part
// there; the transforms drop it (a c2r uses the real part of that plane), so it takes no part
// in the nonlinear term and only the linear factors act on it.  Round-off of the forward
// transforms feeds it a little every step, and inside a forcing band whose growth rate
// exceeds the viscous decay it would grow without bound; the projection removes it after
// every step.  The state is left unchanged up to that round-off.
template <typename T>
__global__ void k_hermitian(cx<T>* __restrict__ u0, int n, int sp, int dim, long long sstride,
                            const double* __restrict__ prm) {
  if (flow_inactive(prm, n / 2 + 1)) return;
  const long long rows = dim == 2 ? (long long)n : (long long)n * n;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int ncomp = dim == 2 ? 1 : 3;
  if (idx >= ncomp * rows) return;
  const int c = (int)(idx / rows);
  const long long r = idx - c * rows;
  long long rm;
  if (dim == 2) {
    rm = r ? n - r : 0;
  } else {
    const int x = (int)(r / n), y = (int)(r % n);
    rm = (long long)(x ? n - x : 0) * n + (y ? n - y : 0);
  }
  if (rm < r) return;  // the mirror's owner handles the pair
  cx<T>* a = u0 + c * sstride + r * sp;
  if (rm == r) {
    a->y = (T)0;
    return;
  }
  cx<T>* b = u0 + c * sstride + rm * sp;
  const cx<T> ua = *a, ub = *b;
  const cx<T> h = mkc((T)0.5 * (ua.x + ub.x), (T)0.5 * (ua.y - ub.y));
  *a = h;
  *b = conjg(h);
}

template <typename T>
hipError_t launch_hermitian(const Flow<T>& f, hipStream_t st) {
  const long long rows = f.dim == 2 ? (long long)f.n : (long long)f.n * f.n;
  const long long nt = (f.dim == 2 ? 1 : 3) * rows;
  hipLaunchKernelGGL(k_hermitian<T>, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st,
                     f.u0, f.n, f.sp, f.dim, f.sstride, f.prm);
  return hipGetLastError();
}
template hipError_t launch_hermitian<float>(const Flow<float>&, hipStream_t);
template hipError_t launch_hermitian<double>(const Flow<double>&, hipStream_t);

// Mode idx of a state component in NumPy layout -> integer wavenumbers and Hermitian weight.
__device__ __forceinline__ void decode_mode(long long idx, int n, int nh, int dim, int& kx,
                                            int& ky, int& kz) {
  const int j = (int)(idx % nh);
  const long long r = idx / nh;
  if (dim == 2) {
    const int i = (int)r;
    kx = i < n / 2 ? i : i - n;
    ky = j;
    kz = 0;
  } else {
    const int i = (int)(r / n), l = (int)(r % n);
    kx = i < n / 2 ? i : i - n;
    ky = l < n / 2 ? l : l - n;
    kz = j;
  }
}

constexpr int kRedThreads = 256;

template <typename T>
__global__ void __launch_bounds__(kRedThreads)
    k_reduce(const cx<T>* __restrict__ st, long long sstride, int pitch, int n, int dim,
             double klo2, double khi2, double* __restrict__ partial) {
  const int nh = n / 2 + 1;
  const long long nm = (dim == 2 ? (long long)n : (long long)n * n) * nh;
  const double inv_norm = 1.0 / pow((double)n, 2.0 * dim);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};  // E, Z, P, E_band
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < nm;
       idx += (long long)gridDim.x * blockDim.x) {
    int kx, ky, kz;
    decode_mode(idx, n, nh, dim, kx, ky, kz);
    const int j = (int)(idx % nh);
    double a = 0.0;
    for (int c = 0; c < (dim == 2 ? 1 : 3); ++c) {
      const cx<T> z = st[c * sstride + (idx / nh) * pitch + j];
      a += (double)z.x * (double)z.x + (double)z.y * (double)z.y;
    }
    const double w = (j == 0 || j == nh - 1) ? 1.0 : 2.0;
    const double d = a * w * inv_norm;
    const double k2 = (double)kx * kx + (double)ky * ky + (double)kz * kz;
    const bool inband = k2 >= klo2 && k2 <= khi2;
    if (dim == 2) {
      const double e = k2 > 0 ? d / k2 : 0.0;
      acc[0] += 0.5 * e;
      acc[1] += 0.5 * d;
      acc[2] += 0.5 * d * k2;
      if (inband) acc[3] += 0.5 * e;
    } else {
      acc[0] += 0.5 * d;
      acc[1] += 0.5 * d * k2;
      acc[2] += 0.5 * d * k2 * k2;
      if (inband) acc[3] += 0.5 * d;
    }
  }
  __shared__ double red[4][kRedThreads / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int c = 0; c < 4; ++c) {
    double x = acc[c];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if (lane == 0) red[c][wv] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int c = 0; c < 4; ++c) {
      double x = 0.0;
      for (int w = 0; w < kRedThreads / 64; ++w) x += red[c][w];
      partial[(long long)blockIdx.x * 4 + c] = x;
    }
  }
}

__global__ void k_reduce_final(const double* __restrict__ partial, int nblocks, double* out) {
  __shared__ double red[4][kRedThreads];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
    for (int c = 0; c < 4; ++c) acc[c] += partial[(long long)b * 4 + c];
  for (int c = 0; c < 4; ++c) red[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 4; ++c) out[c] = red[c][0];
}

template <typename T>
hipError_t launch_reduce(const cx<T>* state, long long sstride, int pitch, int n, int dim,
                         double klo2, double khi2, double* partial, int nblocks, double* out,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_reduce<T>, dim3(nblocks), dim3(kRedThreads), 0, st, state, sstride,
                     pitch, n, dim, klo2, khi2, partial);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRedThreads), 0, st, partial, nblocks, out);
  return hipGetLastError();
}

template <typename T>
__global__ void __launch_bounds__(kRedThreads)
    k_spectrum(const cx<T>* __restrict__ st, long long sstride, int pitch, int n, int dim,
               double* __restrict__ out, int nshell) {
  extern __shared__ double hist[];
  for (int s = threadIdx.x; s < nshell; s += blockDim.x) hist[s] = 0.0;
  __syncthreads();
  const int nh = n / 2 + 1;
  const long long nm = (dim == 2 ? (long long)n : (long long)n * n) * nh;
  const double inv_norm = 1.0 / pow((double)n, 2.0 * dim);
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < nm;
       idx += (long long)gridDim.x * blockDim.x) {
    int kx, ky, kz;
    decode_mode(idx, n, nh, dim, kx, ky, kz);
    const int j = (int)(idx % nh);
    double a = 0.0;
    for (int c = 0; c < (dim == 2 ? 1 : 3); ++c) {
      const cx<T> z = st[c * sstride + (idx / nh) * pitch + j];
      a += (double)z.x * (double)z.x + (double)z.y * (double)z.y;
    }
    const double w = (j == 0 || j == nh - 1) ? 1.0 : 2.0;
    double d = a * w * inv_norm;
    const double k2 = (double)kx * kx + (double)ky * ky + (double)kz * kz;
    if (dim == 2) d = k2 > 0 ? d / k2 : 0.0;
    const int shell = (int)(sqrt(k2) + 0.5);
    if (shell < nshell && d != 0.0) atomicAdd(&hist[shell], 0.5 * d);
  }
  __syncthreads();
  for (int s = threadIdx.x; s < nshell; s += blockDim.x)
    if (hist[s] != 0.0) atomicAdd(&out[s], hist[s]);
}

template <typename T>
hipError_t launch_spectrum(const cx<T>* state, long long sstride, int pitch, int n, int dim,
                           double* out, int nshell, hipStream_t st) {
  hipError_t e = launch_fill_f64(out, 0.0, nshell, st);
  if (e != hipSuccess) return e;
  const size_t lds = (size_t)nshell * sizeof(double);
  if (lds > 65536)
    (void)hipFuncSetAttribute((const void*)k_spectrum<T>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(k_spectrum<T>, dim3(1024), dim3(kRedThreads), lds, st, state, sstride,
                     pitch, n, dim, out, nshell);
  return hipGetLastError();
}

template hipError_t launch_reduce<float>(const cx<float>*, long long, int, int, int, double,
                                         double, double*, int, double*, hipStream_t);
template hipError_t launch_reduce<double>(const cx<double>*, long long, int, int, int, double,
                                          double, double*, int, double*, hipStream_t);
template hipError_t launch_spectrum<float>(const cx<float>*, long long, int, int, int, double*,
                                           int, hipStream_t);
template hipError_t launch_spectrum<double>(const cx<double>*, long long, int, int, int,
                                            double*, int, hipStream_t);

}  // namespace psx
