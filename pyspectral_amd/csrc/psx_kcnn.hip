// Convolutional subgrid-stress network on MFMA (gfx950, bf16 inputs, fp32 accumulation):
// periodic 3x3x3 convolutions with 32 channels in and out, their input and weight
// gradients, a mean-square loss and Adam.  No counterpart in the reference.
//
// MFMA v_mfma_f32_16x16x32_bf16 computes D[16 x 16] += A[16 x 32] B[32 x 16].  Lane l holds
// row (l & 15) of A and column (l & 15) of B, each with 8 consecutive k of k-group (l >> 4),
// and D[4 (l >> 4) + j][l & 15], j < 4.  Which k a lane's element stands for only has to
// agree between A and B, so each kernel chooses its own k order.
#include <mutex>

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

// ------------------------------------------------- forward / input grad, n = 64, LDS-staged
// Work-group: a brick of 4 (x) x 4 (y) x 32 (z) output points of one cube, all 32 output
// channels.  The brick and its one-point periodic halo (6 x 6 x 34 points of 32 bf16
// channels, 64 B each) are staged in LDS once, so each input point is read from memory
// about 2.4 times instead of 27; every tap's A operand is then an LDS read.  The four 16-byte
// channel groups of halo point q are rotated by q >> 1: the 8 consecutive points of one MFMA
// row group then fall in distinct banks for any tap offset.  Wave w owns the brick lines
// (x = w, y = 0..3): 8 row tiles of 16 z points.  The result goes through LDS once more
// (fp32, padded rows) so that the epilogue (bias, ReLU, mask) reads and writes whole 16-byte
// channel groups of consecutive points.
constexpr int kCvBx = 4, kCvBy = 4, kCvBz = 32;
constexpr int kCvHx = kCvBx + 2, kCvHy = kCvBy + 2, kCvHz = kCvBz + 2;
constexpr int kCvHalo = kCvHx * kCvHy * kCvHz;       // 1224 points
constexpr int kCvPts = kCvBx * kCvBy * kCvBz;        // 512 points
constexpr int kCvLds = kCvHalo * 64;                 // 78336 bytes: two work-groups per CU
constexpr int kCvOutPitch = kCnnC + 4;               // fp32 staging row, 144 bytes
static_assert(kCvPts * kCvOutPitch * 4 <= kCvLds, "output staging fits the halo buffer");

__device__ __forceinline__ int cv_slot(int q, int g) {
  return q * 64 + (((g + (q >> 1)) & 3) << 4);
}

__global__ void __launch_bounds__(256) k_conv64(ConvArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int n = 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r = lane & 15, kg = lane >> 4;
  // bricks are dispatched round-robin over the 8 XCDs: give each XCD a contiguous range of
  // bricks, so that neighbouring bricks (which share halo points) share an L2
  const int per = (int)gridDim.x >> 3;
  const int brick = ((int)blockIdx.x & 7) * per + ((int)blockIdx.x >> 3);
  const int cube = brick >> 9;
  const int x0 = ((brick >> 5) & 15) * kCvBx, y0 = ((brick >> 1) & 15) * kCvBy,
            z0 = (brick & 1) * kCvBz;
  const unsigned char* X =
      reinterpret_cast<const unsigned char*>(a.x) + (size_t)cube * (n * n * n * 64);
  for (int c = tid; c < kCvHalo * 4; c += 256) {
    const int q = c >> 2, g = c & 3;
    const int hz = q % kCvHz, hy = (q / kCvHz) % kCvHy, hx = q / (kCvHz * kCvHy);
    const int gx = (x0 - 1 + hx) & (n - 1), gy = (y0 - 1 + hy) & (n - 1),
              gz = (z0 - 1 + hz) & (n - 1);
    const uint4 v = *reinterpret_cast<const uint4*>(X + ((((gx << 6) | gy) << 6 | gz) << 6) +
                                                    (g << 4));
    *reinterpret_cast<uint4*>(smem_raw + cv_slot(q, g)) = v;
  }
  __syncthreads();

  f32x4 acc[8][2];
#pragma unroll
  for (int t = 0; t < 8; ++t) acc[t][0] = acc[t][1] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16x8* W = reinterpret_cast<const bf16x8*>(a.w);
  // halo point of row r of tile t = 2 ly + zt at tap (dx, dy, dz):
  // ((wave + 1 + dx) * 6 + ly + 1 + dy) * 34 + 16 zt + r + 1 + dz
  const int qlane = ((wave + 1) * kCvHy + 1) * kCvHz + r + 1;
#pragma unroll
  for (int tap = 0; tap < kCnnTaps; ++tap) {
    const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
    const bf16x8 b0 = W[(tap * kCnnC + r) * 4 + kg];
    const bf16x8 b1 = W[(tap * kCnnC + 16 + r) * 4 + kg];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int q = qlane + (dx * kCvHy + (t >> 1) + dy) * kCvHz + 16 * (t & 1) + dz;
      const bf16x8 av = *reinterpret_cast<const bf16x8*>(smem_raw + cv_slot(q, kg));
      acc[t][0] = mfma(av, b0, acc[t][0]);
      acc[t][1] = mfma(av, b1, acc[t][1]);
    }
  }
  __syncthreads();  // every wave is done with the halo

  // accumulators -> LDS rows of local point lp = (lx * 4 + ly) * 32 + z
  float* so = reinterpret_cast<float*>(smem_raw);
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int lp0 = (wave * kCvBy + (t >> 1)) * kCvBz + 16 * (t & 1) + 4 * kg;
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int j = 0; j < 4; ++j) so[(lp0 + j) * kCvOutPitch + 16 * u + r] = acc[t][u][j];
  }
  __syncthreads();
  // epilogue: (local point, 8-channel group) items, consecutive items contiguous in memory
  for (int it = tid; it < kCvPts * 4; it += 256) {
    const int lp = it >> 2, g = it & 3;
    const int z = lp & (kCvBz - 1), ly = (lp >> 5) & 3, lx = lp >> 7;
    const size_t p = (((size_t)cube * n + x0 + lx) * n + y0 + ly) * n + z0 + z;
    const float4 lo = *reinterpret_cast<const float4*>(so + lp * kCvOutPitch + 8 * g);
    const float4 hi = *reinterpret_cast<const float4*>(so + lp * kCvOutPitch + 8 * g + 4);
    float v[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
    if (a.bias) {
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] += a.bias[8 * g + i];
    }
    if (a.relu) {
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = fmaxf(v[i], 0.f);
    }
    if (a.mask) {
      const bf16x8 m = *reinterpret_cast<const bf16x8*>(a.mask + p * kCnnC + 8 * g);
#pragma unroll
      for (int i = 0; i < 8; ++i)
        if (!((float)m[i] > 0.f)) v[i] = 0.f;
    }
    if (a.y) {
      bf16x8 o;
#pragma unroll
      for (int i = 0; i < 8; ++i) o[i] = (__bf16)v[i];
      *reinterpret_cast<bf16x8*>(a.y + p * kCnnC + 8 * g) = o;
    }
    if (a.yf) {
      *reinterpret_cast<float4*>(a.yf + p * kCnnC + 8 * g) = float4{v[0], v[1], v[2], v[3]};
      *reinterpret_cast<float4*>(a.yf + p * kCnnC + 8 * g + 4) =
          float4{v[4], v[5], v[6], v[7]};
    }
  }
}

hipError_t launch_conv(const ConvArgs& a, hipStream_t st) {
  if (a.n < 4 || (a.n & (a.n - 1)) || a.nb < 1 || !a.x || !a.w || (!a.y && !a.yf))
    return hipErrorInvalidValue;
  const long long P = (long long)a.nb * a.n * a.n * a.n;
  if (a.n == 64) {
    // 512 bricks per cube (a multiple of the 8 XCDs)
    hipLaunchKernelGGL(k_conv64, dim3((unsigned)(P / kCvPts)), dim3(256), kCvLds, st, a);
    return hipGetLastError();
  }
  const unsigned blocks = (unsigned)((P + 255) / 256);
  hipLaunchKernelGGL(k_conv, dim3(blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

// Sum of the per-work-group partials, added into dw / db.  A work-group owns 64 consecutive
// entries; its 4 waves sum the partials g = wave, wave + 4, ... in order, and the four wave
// sums are added in a fixed order, so the result does not depend on scheduling.
constexpr int kWgRedCols = 64;

__global__ void __launch_bounds__(256) k_wgrad_reduce(const float* __restrict__ part,
                                                      int ngroups, float* dw, float* db) {
  constexpr int NW = kCnnTaps * kCnnC * kCnnC;
  constexpr int NT = NW + kCnnC;
  __shared__ float red[4][kWgRedCols];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = blockIdx.x * kWgRedCols + lane;
  float s = 0.f;
  if (i < NT)
    for (int g = wave; g < ngroups; g += 4) s += part[(long long)g * NT + i];
  red[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && i < NT) {
    const float t = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
    if (i < NW) dw[i] += t;
    else db[i - NW] += t;
  }
}

// ------------------------------------------------------ weight gradient, n = 64, MFMA-only
// dW[tap][o][i] = sum_p dz[p][o] x[p + tap][i] is a contraction over points, so both MFMA
// operands need points along k, while the tensors are point-major (channels contiguous).
// The transposition is itself an MFMA: a 16-point x 32-channel row tile (the natural
// point-major operand) times the 32 x 16 identity slice I_u (I_u[k][n] = [k == 16 u + n])
// leaves lane (r, kg) holding the tile's points 4 kg + j, j < 4, of channel 16 u + r -- one
// product by 1.0 each, so exact.  Two tiles (a 32-point line segment) give each lane 8
// points of one channel: an operand with k = points, in the same point order for dz (A,
// rows = output channels) and for x (B, columns = input channels), which is all the
// contraction needs.  The shifted x tiles come from the LDS halo brick of the forward kernel.
// Work-groups persist over bricks (4 x 4 x 32 points); wave w accumulates taps w, w + 4, ...
// (7 or 6 taps, 32 x 32 each) over all points of its bricks, and writes per-work-group
// partial sums (fixed-order reduction afterwards, as above).
constexpr int kWg64Grid = 512;

__global__ void __launch_bounds__(256) k_wgrad64(const __bf16* __restrict__ x,
                                                 const __bf16* __restrict__ dz, int nbricks,
                                                 float* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int n = 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r = lane & 15, kg = lane >> 4;
  bf16x8 id[2];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int j = 0; j < 8; ++j) id[u][j] = (__bf16)((8 * kg + j == 16 * u + r) ? 1.f : 0.f);
  const f32x4 zero4 = f32x4{0.f, 0.f, 0.f, 0.f};
  f32x4 acc[7][2][2];
#pragma unroll
  for (int i = 0; i < 7; ++i)
#pragma unroll
    for (int u = 0; u < 2; ++u) acc[i][u][0] = acc[i][u][1] = zero4;
  float bsum[2] = {0.f, 0.f};
  const bf16x8* DZ = reinterpret_cast<const bf16x8*>(dz);

  for (int brick = blockIdx.x; brick < nbricks; brick += gridDim.x) {
    const int cube = brick >> 9;
    const int x0 = ((brick >> 5) & 15) * kCvBx, y0 = ((brick >> 1) & 15) * kCvBy,
              z0 = (brick & 1) * kCvBz;
    const unsigned char* X =
        reinterpret_cast<const unsigned char*>(x) + (size_t)cube * (n * n * n * 64);
    __syncthreads();  // the previous brick's halo is no longer read
    for (int c = tid; c < kCvHalo * 4; c += 256) {
      const int q = c >> 2, g = c & 3;
      const int hz = q % kCvHz, hy = (q / kCvHz) % kCvHy, hx = q / (kCvHz * kCvHy);
      const int gx = (x0 - 1 + hx) & (n - 1), gy = (y0 - 1 + hy) & (n - 1),
                gz = (z0 - 1 + hz) & (n - 1);
      const uint4 v = *reinterpret_cast<const uint4*>(
          X + ((((gx << 6) | gy) << 6 | gz) << 6) + (g << 4));
      *reinterpret_cast<uint4*>(smem_raw + cv_slot(q, g)) = v;
    }
    __syncthreads();
    for (int line = 0; line < kCvBx * kCvBy; ++line) {
      const int lx = line >> 2, ly = line & 3;
      // dz of the 32-point segment, transposed: A[u] = channel 16 u + r, 8 points
      bf16x8 A[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const size_t p =
            (((size_t)cube * n + x0 + lx) * n + y0 + ly) * n + z0 + 16 * h + r;
        const bf16x8 rows = DZ[p * 4 + kg];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const f32x4 t = mfma(rows, id[u], zero4);
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            A[u][4 * h + j] = (__bf16)t[j];
            bsum[u] += t[j];
          }
        }
      }
      const int qline = ((lx + 1) * kCvHy + ly + 1) * kCvHz + r + 1;
#pragma unroll
      for (int i = 0; i < 7; ++i) {
        const int tap = wave + 4 * i;
        if (tap < kCnnTaps) {
          const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dzz = tap % 3 - 1;
          bf16x8 B[2];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int q = qline + (dx * kCvHy + dy) * kCvHz + 16 * h + dzz;
            const bf16x8 rows = *reinterpret_cast<const bf16x8*>(smem_raw + cv_slot(q, kg));
#pragma unroll
            for (int v = 0; v < 2; ++v) {
              const f32x4 t = mfma(rows, id[v], zero4);
#pragma unroll
              for (int j = 0; j < 4; ++j) B[v][4 * h + j] = (__bf16)t[j];
            }
          }
#pragma unroll
          for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int v = 0; v < 2; ++v) acc[i][u][v] = mfma(A[u], B[v], acc[i][u][v]);
        }
      }
    }
  }

  // partial sums of this work-group: acc[i][u][v][j] = dW[tap][16 u + 4 kg + j][16 v + r]
  float* out = part + (long long)blockIdx.x * (kCnnTaps * kCnnC * kCnnC + kCnnC);
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    const int tap = wave + 4 * i;
    if (tap < kCnnTaps) {
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            out[(tap * kCnnC + 16 * u + 4 * kg + j) * kCnnC + 16 * v + r] = acc[i][u][v][j];
    }
  }
  // bias gradient: wave 0's lanes (r, kg) hold channel 16 u + r summed over their points
  __syncthreads();
  float* red = reinterpret_cast<float*>(smem_raw);
  if (wave == 0) {
    red[lane * 2] = bsum[0];
    red[lane * 2 + 1] = bsum[1];
  }
  __syncthreads();
  if (tid < kCnnC) {
    const int u = tid >> 4, rr = tid & 15;
    float s = 0.f;
#pragma unroll
    for (int g = 0; g < 4; ++g) s += red[(16 * g + rr) * 2 + u];
    out[kCnnTaps * kCnnC * kCnnC + tid] = s;
  }
}

constexpr int kWgGroups = 512;
static_assert(kWg64Grid == kWgGroups, "one partial row per work-group of the gradient kernel");
constexpr size_t kWgPartFloats = (size_t)kWgGroups * (kCnnTaps * kCnnC * kCnnC + kCnnC);

// Both kernels of one weight gradient on the partial-sum buffer `part` (kWgPartFloats floats).
// Every work-group of k_wgrad64 writes its whole partial row before k_wgrad_reduce reads it,
// so the buffer needs no clearing between calls.
static hipError_t enqueue_wgrad(const __bf16* x, const __bf16* dz, float* dw, float* db, int n,
                                int nb, float* part, hipStream_t st) {
  const int nbricks = nb * (n * n * n / kCvPts);
  hipLaunchKernelGGL(k_wgrad64, dim3(kWg64Grid), dim3(256), kCvLds, st, x, dz, nbricks, part);
  const int nw = kCnnTaps * kCnnC * kCnnC + kCnnC;
  hipLaunchKernelGGL(k_wgrad_reduce, dim3((nw + kWgRedCols - 1) / kWgRedCols), dim3(256), 0,
                     st, part, kWgGroups, dw, db);
  return hipGetLastError();
}

hipError_t launch_conv_wgrad(const __bf16* x, const __bf16* dz, float* dw, float* db, int n,
                             int nb, hipStream_t st) {
  if (n != 64 || nb < 1) return hipErrorInvalidValue;
  // a partial buffer of this call's own, allocated and freed in stream order: calls on other
  // streams or threads never share it, and a captured call records it as graph nodes
  float* part = nullptr;
  hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&part), kWgPartFloats * sizeof(float),
                                st);
  if (e != hipSuccess) return e;
  e = enqueue_wgrad(x, dz, dw, db, n, nb, part, st);
  const hipError_t ef = hipFreeAsync(part, st);
  return e != hipSuccess ? e : ef;
}

// Partial sums of psx_cnn_wgrad_noalloc: one array per GPU, held until the process ends.  Host
// code never passes the array's name to a kernel (on the host it is only a stand-in): the
// device address comes from hipGetSymbolAddress once per device, with that device current.
static __device__ __align__(256) float scratch[kWgPartFloats];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static float* g_scratch_ptr[kMaxDevices] = {};

static hipError_t scratch_ptr(float** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

size_t wgrad_scratch_bytes() { return sizeof(float) * kWgPartFloats; }

hipError_t launch_conv_wgrad_noalloc(const __bf16* x, const __bf16* dz, float* dw, float* db,
                                     int n, int nb, hipStream_t st) {
  if (n != 64 || nb < 1) return hipErrorInvalidValue;
  float* part = nullptr;
  const hipError_t e = scratch_ptr(&part);
  if (e != hipSuccess) return e;
  return enqueue_wgrad(x, dz, dw, db, n, nb, part, st);
}

// -------------------------------------------------------------------------------- loss
// One thread per point: the point's 32 outputs (128 B) and targets come in as float4 loads,
// the 32 bf16 gradients go out as four 16-byte stores.
__global__ void __launch_bounds__(256) k_mse(const float* __restrict__ y,
                                             const float* __restrict__ t, int tstride, int nch,
                                             long long npts, float inv_count,
                                             __bf16* __restrict__ g, double* loss) {
  __shared__ double red[256];
  double a = 0.0;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < npts;
       p += (long long)gridDim.x * blockDim.x) {
    const float4* yp = reinterpret_cast<const float4*>(y + p * kCnnC);
    const float* tp = t + p * tstride;
#pragma unroll
    for (int q = 0; q < kCnnC / 8; ++q) {
      const float4 lo = yp[2 * q], hi = yp[2 * q + 1];
      const float yv[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
      bf16x8 o;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int c = 8 * q + j;
        float gv = 0.f;
        if (c < nch) {
          const float d = yv[j] - tp[c];
          a += (double)d * d;
          gv = 2.f * d * inv_count;
        }
        o[j] = (__bf16)gv;
      }
      *reinterpret_cast<bf16x8*>(g + p * kCnnC + 8 * q) = o;
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
