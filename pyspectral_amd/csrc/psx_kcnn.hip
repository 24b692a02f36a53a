// Convolutional subgrid-stress network on MFMA (gfx950, bf16 inputs, fp32 accumulation):
// periodic 3x3x3 convolutions with 32 channels in and out, their input and weight
// gradients, a mean-square loss and Adam.  No counterpart in the reference.
//
// MFMA v_mfma_f32_16x16x32_bf16 computes D[16 x 16] += A[16 x 32] B[32 x 16].  Lane l holds
// row (l & 15) of A and column (l & 15) of B, each with 8 consecutive k of k-group (l >> 4),
// and D[4 (l >> 4) + j][l & 15], j < 4.  Which k a lane's element stands for only has to
// agree between A and B, so each kernel chooses its own k order.
#include "psx_kernels.h"
#include "psx_les.h"

namespace psx {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Periodic neighbour of point p (cubes of n^3, n = 1 << ln) at offset (dx, dy, dz).
__device__ __forceinline__ long long nbr(long long p, int ln, int dx, int dy, int dz) {
  const int m = (1 << ln) - 1;
  const int z = (int)p & m, y = (int)(p >> ln) & m, x = (int)(p >> (2 * ln)) & m;
  const long long cube = p >> (3 * ln);
  return (((cube << ln | ((x + dx) & m)) << ln | ((y + dy) & m)) << ln) | ((z + dz) & m);
}

// ------------------------------------------------------------------ forward / input grad
// One wave: 64 consecutive points (4 row tiles of 16) x 32 output channels (2 tiles); the
// k loop runs over the 27 taps, 32 input channels each (one MFMA per tile pair and tap).
__global__ void __launch_bounds__(256) k_conv(ConvArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ln = __builtin_ctz(a.n);
  const long long P = (long long)a.nb << (3 * ln);
  const long long p0 = ((long long)blockIdx.x * 4 + wave) * 64;
  if (p0 >= P) return;
  const int r = lane & 15, kg = lane >> 4;
  f32x4 acc[4][2];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t][0] = acc[t][1] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16x8* W = reinterpret_cast<const bf16x8*>(a.w);
  const bf16x8* X = reinterpret_cast<const bf16x8*>(a.x);
  for (int tap = 0; tap < kCnnTaps; ++tap) {
    const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
    // B: weights [tap][out][in]; column = output channel, k = input channel
    const bf16x8 b0 = W[(tap * kCnnC + r) * 4 + kg];
    const bf16x8 b1 = W[(tap * kCnnC + 16 + r) * 4 + kg];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const long long q = nbr(p0 + 16 * t + r, ln, dx, dy, dz);
      const bf16x8 av = X[q * 4 + kg];  // row = point, k = input channel
      acc[t][0] = mfma(av, b0, acc[t][0]);
      acc[t][1] = mfma(av, b1, acc[t][1]);
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int o = 16 * u + r;
      const float b = a.bias ? a.bias[o] : 0.f;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const long long p = p0 + 16 * t + 4 * kg + j;
        float v = acc[t][u][j] + b;
        if (a.relu) v = fmaxf(v, 0.f);
        if (a.mask && !((float)a.mask[p * kCnnC + o] > 0.f)) v = 0.f;
        if (a.y) a.y[p * kCnnC + o] = (__bf16)v;
        if (a.yf) a.yf[p * kCnnC + o] = v;
      }
    }
}

hipError_t launch_conv(const ConvArgs& a, hipStream_t st) {
  if (a.n < 4 || (a.n & (a.n - 1)) || a.nb < 1 || !a.x || !a.w || (!a.y && !a.yf))
    return hipErrorInvalidValue;
  const long long P = (long long)a.nb * a.n * a.n * a.n;
  const unsigned blocks = (unsigned)((P + 255) / 256);
  hipLaunchKernelGGL(k_conv, dim3(blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

// ----------------------------------------------------------------------- weight gradient
// Work-group g accumulates rows (cube, x, y) g, g + G, ... (each a z line of n points): the
// output gradient line and the 9 neighbouring input lines (dx, dy) are staged channel-major
// in LDS with a one-point periodic halo, so that an MFMA operand (8 consecutive points of one
// channel) is one contiguous read.  Wave w owns taps w, w + 4, ...; D[out 16][in 16] tiles.
// Partial sums go to part[g][27 * 32 * 32 + 32] (the last 32: bias gradient).
constexpr int kWgLine = 64;                  // points per line (n = 64)
constexpr int kWgPitch = kWgLine + 8;        // halo of 1 each side, padded
constexpr int kWgTapsPerWave = 7;

__global__ void __launch_bounds__(256) k_conv_wgrad(const __bf16* __restrict__ x,
                                                    const __bf16* __restrict__ dz, int nb,
                                                    float* __restrict__ part) {
  constexpr int n = kWgLine, ln = 6;
  __shared__ __bf16 xs[9][kCnnC][kWgPitch];  // [dx,dy][channel][1 + z]
  __shared__ __bf16 gs[kCnnC][kWgPitch];     // [out channel][z]
  __shared__ float bsum[4][kCnnC];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r = lane & 15, kg = lane >> 4;
  const long long nlines = (long long)nb * n * n;
  f32x4 acc[kWgTapsPerWave][2][2];
#pragma unroll
  for (int i = 0; i < kWgTapsPerWave; ++i)
#pragma unroll
    for (int u = 0; u < 2; ++u) acc[i][u][0] = acc[i][u][1] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bacc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const int sp = tid >> 2, sc = (tid & 3) * 8;  // staging: point sp, channels sc..sc+7
  for (long long line = blockIdx.x; line < nlines; line += gridDim.x) {
    const long long base = line * n;  // point index of z = 0
    __syncthreads();
    {
      const bf16x8 gv = *reinterpret_cast<const bf16x8*>(dz + (base + sp) * kCnnC + sc);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        gs[sc + j][sp] = gv[j];
        bacc[j] += (float)gv[j];
      }
    }
#pragma unroll
    for (int s = 0; s < 9; ++s) {
      const int dx = s / 3 - 1, dy = s % 3 - 1;
      const long long q = nbr(base + sp, ln, dx, dy, 0);
      const bf16x8 xv = *reinterpret_cast<const bf16x8*>(x + q * kCnnC + sc);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        xs[s][sc + j][1 + sp] = xv[j];
        if (sp == 0) xs[s][sc + j][1 + n] = xv[j];       // halo z = n  -> z = 0
        if (sp == n - 1) xs[s][sc + j][0] = xv[j];       // halo z = -1 -> z = n - 1
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < n / 32; ++kk) {
      const int z0 = 32 * kk + 8 * kg;  // this lane's 8 points: z0 .. z0 + 7
      bf16x8 ga[2];
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int j = 0; j < 8; ++j) ga[u][j] = gs[16 * u + r][z0 + j];
#pragma unroll
      for (int i = 0; i < kWgTapsPerWave; ++i) {
        const int tap = wave + 4 * i;
        if (tap < kCnnTaps) {
          const int s = tap / 3, dzz = tap % 3 - 1;
#pragma unroll
          for (int v = 0; v < 2; ++v) {
            bf16x8 xb;
#pragma unroll
            for (int j = 0; j < 8; ++j) xb[j] = xs[s][16 * v + r][1 + z0 + j + dzz];
            acc[i][0][v] = mfma(ga[0], xb, acc[i][0][v]);
            acc[i][1][v] = mfma(ga[1], xb, acc[i][1][v]);
          }
        }
      }
    }
  }
  // partial sums of this work-group
  float* out = part + (long long)blockIdx.x * (kCnnTaps * kCnnC * kCnnC + kCnnC);
#pragma unroll
  for (int i = 0; i < kWgTapsPerWave; ++i) {
    const int tap = wave + 4 * i;
    if (tap < kCnnTaps) {
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int o = 16 * u + 4 * kg + j, ic = 16 * v + r;
            out[(tap * kCnnC + o) * kCnnC + ic] = acc[i][u][v][j];
          }
    }
  }
  // bias gradient: threads with the same channel group (tid & 3) hold channels sc..sc+7
  __syncthreads();
  float* red = reinterpret_cast<float*>(&xs[0][0][0]);  // reuse: 256 x 8 floats
#pragma unroll
  for (int j = 0; j < 8; ++j) red[tid * 8 + j] = bacc[j];
  __syncthreads();
  if (tid < kCnnC) {
    const int grp = tid / 8, j = tid % 8;
    float sacc = 0.f;
    for (int t = grp; t < 256; t += 4) sacc += red[t * 8 + j];
    out[kCnnTaps * kCnnC * kCnnC + tid] = sacc;
  }
  (void)bsum;
}

// Sum of the per-work-group partials (fixed order), added into dw / db.
__global__ void k_wgrad_reduce(const float* __restrict__ part, int ngroups, float* dw,
                               float* db) {
  constexpr int NW = kCnnTaps * kCnnC * kCnnC;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NW + kCnnC) return;
  float s = 0.f;
  for (int g = 0; g < ngroups; ++g) s += part[(long long)g * (NW + kCnnC) + i];
  if (i < NW) dw[i] += s;
  else db[i - NW] += s;
}

constexpr int kWgGroups = 512;

hipError_t launch_conv_wgrad(const __bf16* x, const __bf16* dz, float* dw, float* db, int n,
                             int nb, hipStream_t st) {
  if (n != kWgLine || nb < 1) return hipErrorInvalidValue;
  static float* part = nullptr;  // per-process partial buffer (512 x 27.7k floats)
  if (!part) {
    hipError_t e = hipMalloc(&part, (size_t)kWgGroups * (kCnnTaps * kCnnC * kCnnC + kCnnC) *
                                        sizeof(float));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_conv_wgrad, dim3(kWgGroups), dim3(256), 0, st, x, dz, nb, part);
  const int nw = kCnnTaps * kCnnC * kCnnC + kCnnC;
  hipLaunchKernelGGL(k_wgrad_reduce, dim3((nw + 255) / 256), dim3(256), 0, st, part,
                     kWgGroups, dw, db);
  return hipGetLastError();
}

// -------------------------------------------------------------------------------- loss
__global__ void k_mse(const float* __restrict__ y, const float* __restrict__ t, int tstride,
                      int nch, long long npts, float inv_count, __bf16* __restrict__ g,
                      double* loss) {
  __shared__ double red[256];
  double a = 0.0;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < npts;
       p += (long long)gridDim.x * blockDim.x) {
    for (int c = 0; c < kCnnC; ++c) {
      float gv = 0.f;
      if (c < nch) {
        const float d = y[p * kCnnC + c] - t[p * tstride + c];
        a += (double)d * d;
        gv = 2.f * d * inv_count;
      }
      g[p * kCnnC + c] = (__bf16)gv;
    }
  }
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0] * (double)inv_count);
}

hipError_t launch_mse(const float* y, const float* t, int tstride, int nch, long long npts,
                      __bf16* g, double* loss, hipStream_t st) {
  const float inv = (float)(1.0 / ((double)npts * nch));
  hipLaunchKernelGGL(k_mse, dim3(1024), dim3(256), 0, st, y, t, tstride, nch, npts, inv, g,
                     loss);
  return hipGetLastError();
}

// -------------------------------------------------------------------------------- Adam
__global__ void k_adam(AdamArgs a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const float g = a.g[i];
  const float m = a.b1 * a.m[i] + (1.f - a.b1) * g;
  const float v = a.b2 * a.v[i] + (1.f - a.b2) * g * g;
  a.m[i] = m;
  a.v[i] = v;
  float w = a.w[i] - a.lr * (m * a.c1) / (sqrtf(v * a.c2) + a.eps);
  if (a.live && !a.live[i]) w = 0.f;
  a.w[i] = w;
  const __bf16 wb = (__bf16)w;
  if (a.wb) a.wb[i] = wb;
  if (a.is_w && a.wt) {
    // i = (tap * 32 + o) * 32 + in  ->  wt[(26 - tap) * 32 + in][o]
    const int in = (int)(i % kCnnC), o = (int)((i / kCnnC) % kCnnC), tap = (int)(i / (kCnnC * kCnnC));
    a.wt[((kCnnTaps - 1 - tap) * kCnnC + in) * kCnnC + o] = wb;
  }
}

hipError_t launch_adam(const AdamArgs& a, hipStream_t st) {
  if (a.n <= 0 || !a.w || !a.m || !a.v || !a.g) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace psx
