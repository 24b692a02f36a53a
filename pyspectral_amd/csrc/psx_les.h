// Large-eddy-simulation support: small-cube complex FFTs, spectral helpers (psx_kles.hip) and
// the convolutional subgrid-stress network (psx_kcnn.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/psx.h"
#include "psx_fft.h"

namespace psx {

constexpr int kCnnC = 32;   // channels of every network tensor (inputs and outputs padded)
constexpr int kTauC = 8;    // channels of a stored stress target (6 used)
constexpr int kCnnTaps = 27;

enum : int {
  kLesExpand = 0,      // cin: state (pitch), n = components -> cout: full spectra
  kLesGradPack = 1,    // cin: 3 full spectra, n = pairs (6 or 1), scale -> cout: pairs
  kLesProducts = 2,    // cin: velocity pairs -> cout: 3 product pairs
  kLesTruncate = 3,    // cin: fine (L) spectra, n = arrays, pitch = packed flag, scale = k_c
                       //   -> cout: coarse (L2) full spectra
  kLesPairUp = 4,      // cin: 2 n full spectra, scale -> cout: n pairs
  kLesFeatures = 5,    // cin: gradient pairs (physical), scale = 1/g -> out: bf16 [L^3][32]
  kLesSmagorinsky = 6, // cin: gradient pairs, scale = (Cs D)^2 -> cout: 3 stress pairs
  kLesTauFromNet = 7,  // in2: fp32 [L^3][32] network output, scale -> cout: 3 stress pairs;
                       //   cin (or null): physical gradient pairs -> backscatter times scale2,
                       //   plus the Smagorinsky stress for scale3 = (Cs D)^2 > 0
  kLesTarget = 8,      // cin: filtered product pairs, in2: velocity pairs -> out: fp32 [L^3][8]
  kLesSource = 9,      // cin: stress pair spectra, pitch, in2: null or per-shell factors
                       // -> cout: source in the state layout
  kLesSumSq = 10,      // in2: fp32 values, n = count, pitch = width, L2 = channels counted
                       //   -> out: double sum of squares (accumulated)
};

struct LesArgs {
  const cx<float>* cin = nullptr;
  const void* in2 = nullptr;
  cx<float>* cout = nullptr;
  void* out = nullptr;
  long long n = 0;
  int L = 0, L2 = 0, pitch = 0;
  float scale = 1.f;
  float scale2 = 1.f;
  float scale3 = 0.f;
};

hipError_t launch_les(int which, const LesArgs& a, hipStream_t st);
// In-place unnormalised 3D complex FFT of nbatch contiguous L^3 arrays (L = 64, 128, 256);
// tw: both orders of the pass twiddles of c2c_plan(L).
hipError_t launch_c2c3(cx<float>* d, const cx<float>* tw, int L, int nbatch, bool inv,
                       hipStream_t st);
bool c2c_supported(int L);
void c2c_plan(int L, FftPlan* p);

// Network kernels (psx_kcnn.hip).  Tensors are point-major with kCnnC bf16 channels over
// nb cubes of n^3 points (n a power of two, periodic).  Weights: [27 taps][32 out][32 in] bf16,
// tap = (dx+1)*9 + (dy+1)*3 + (dz+1).
struct ConvArgs {
  const __bf16* x = nullptr;      // input activations
  const __bf16* w = nullptr;      // weights (forward) or flipped-transposed (input gradient)
  const float* bias = nullptr;    // forward only (null: none)
  const __bf16* mask = nullptr;   // input gradient: multiply by (mask > 0) (null: none)
  __bf16* y = nullptr;            // bf16 output (or null)
  float* yf = nullptr;            // fp32 output (or null)
  int n = 0, nb = 0;
  int relu = 0;
};
hipError_t launch_conv(const ConvArgs& a, hipStream_t st);

// dW[tap][o][i] += sum_p dz[p][o] x[p + tap][i], db[o] += sum_p dz[p][o] (fp32, accumulated).
// The per-work-group partial sums live in a buffer allocated and freed in stream order.
hipError_t launch_conv_wgrad(const __bf16* x, const __bf16* dz, float* dw, float* db, int n,
                             int nb, hipStream_t st);
// The same kernels on the library's static per-GPU partial-sum array (no allocation; calls
// that overlap on one GPU race on it).
hipError_t launch_conv_wgrad_noalloc(const __bf16* x, const __bf16* dz, float* dw, float* db,
                                     int n, int nb, hipStream_t st);
size_t wgrad_scratch_bytes();
// Cubes per weight-gradient call: nb * 512 bricks must fit the kernels' int brick index.
constexpr int kWgradMaxCubes = (1 << 22) - 1;

// Mean-square loss over the first `nch` channels: g = 2 (y - t) / count (bf16 [P][32],
// zero elsewhere), loss (double, accumulated) = sum (y - t)^2 / count.
hipError_t launch_mse(const float* y, const float* t, int tstride, int nch, long long npts,
                      __bf16* g, double* loss, hipStream_t st);

// Adam on fp32 master weights (n values) with the bias-corrected step, then bf16 copies in
// the forward layout and, for weight tensors (is_w), the flipped-transposed layout.
struct AdamArgs {
  float* w = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  const float* g = nullptr;
  __bf16* wb = nullptr;
  __bf16* wt = nullptr;
  const unsigned char* live = nullptr;  // 0: padded entry, kept at zero
  long long n = 0;
  float lr = 1e-3f, b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float c1 = 1.f, c2 = 1.f;  // 1 / (1 - b1^t), 1 / (1 - b2^t)
  int is_w = 0;
};
hipError_t launch_adam(const AdamArgs& a, hipStream_t st);

// Subgrid-stress source of the LES model d (psx_set_sgs) from the state u0 (3 components of
// L^2 rows, row pitch `pitch`) into d.src; 0 or a psx_last_error code.
int enqueue_sgs(const psx_sgs_desc& d, const cx<float>* u0, int pitch, int L, hipStream_t st);
// Allocate what enqueue_sgs uses for grid L (the c2c twiddle tables) now, so that a step
// captured into a graph later allocates nothing; 0 or a psx_last_error code.
int sgs_prepare(int L);

}  // namespace psx
