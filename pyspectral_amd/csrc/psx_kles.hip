// Large-eddy-simulation support kernels (fp32): 3D complex FFTs of small cubes, spectral
// packing of velocity and velocity-gradient fields, the box filter with truncation from a fine
// to a coarse grid, the Smagorinsky stress, the subgrid-stress training target, and the
// divergence of a subgrid stress as a Leray-projected source in the state layout.
//
// Conventions: a "full" spectrum is an L^3 complex array [x][y][z] (z fastest) holding the
// unnormalised DFT X(k) = sum_x x(x) e^{-i k.x}; index i maps to wavenumber i < L/2 ? i : i-L.
// Two real fields a, b travel in one complex field as a pair z = a + i b (spectrum A + i B);
// the pair spectrum Z splits back as A(k) = (Z(k) + conj Z(-k)) / 2,
// B(k) = (Z(k) - conj Z(-k)) / (2 i).  The reference has no LES; the state layout and
// normalisation are those of its rfftn convention (src/pyspectral/grid.py:13-67), and the
// filtered-field resampling follows the sharp spectral truncation of
// src/pyspectral/validation.py:25-35 (pad_spectrum) multiplied by the box-filter transfer
// function G(k) = prod_i sin(k_i D / 2) / (k_i D / 2), D = pi / k_c.
#include "psx_fft_ct.h"
#include "psx_kernels.h"
#include "psx_les.h"

namespace psx {

// ---------------------------------------------------------------------------- c2c FFTs
template <int L>
struct C3 {
  static constexpr CtPlan P = ct_plan(L, 16);
  static_assert(P.np > 0, "no plan");
  static constexpr int NT = L;  // 16 rows (or 16 columns) of L points, 16 per thread
};

// 16 contiguous rows of length L per work-group.
template <int L, bool INV>
__global__ void __launch_bounds__(L) k_c2c_rows(cx<float>* __restrict__ d,
                                                const cx<float>* __restrict__ tw) {
  constexpr CtPlan P = C3<L>::P;
  constexpr int NT = C3<L>::NT;
  using Lay = RowsCT<L, 16>;
  __shared__ cx<float> lds[16 * ct_rows_pitch(L)];
  cx<float>* base = d + (long long)blockIdx.x * 16 * L;
  const int tid = threadIdx.x;
  cx<float> v[16];
  constexpr int R0 = ct_radix(P, INV, 0), RL = ct_radix(P, INV, P.np - 1);
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    int e, t;
    ct_elem<R0, L, NT, Lay>(i, tid, e, t);
    v[i] = base[t * L + e];
  }
  fft_ct<16, INV, NT, Lay, P>(v, lds, tw + (INV ? ct_twsize(P, false) : 0), tid);
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    int e, t;
    ct_elem<RL, L, NT, Lay>(i, tid, e, t);
    base[t * L + e] = v[i];
  }
}

// Transforms along a strided axis: group g (offset g * gstride) holds ncols columns side by
// side, element e of column c at e * rstride + c; 16 columns per work-group.
template <int L, bool INV>
__global__ void __launch_bounds__(L) k_c2c_cols(cx<float>* __restrict__ d,
                                                const cx<float>* __restrict__ tw,
                                                long long rstride, int ncols, long long gstride) {
  constexpr CtPlan P = C3<L>::P;
  constexpr int NT = C3<L>::NT;
  using Lay = ColsCT<16>;
  __shared__ cx<float> lds[16 * L];
  const int ntile = ncols / 16;
  const long long g = blockIdx.x / ntile;
  const int j0 = (int)(blockIdx.x - g * ntile) * 16;
  cx<float>* base = d + g * gstride + j0;
  const int tid = threadIdx.x;
  cx<float> v[16];
  constexpr int R0 = ct_radix(P, INV, 0), RL = ct_radix(P, INV, P.np - 1);
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    int e, t;
    ct_elem<R0, L, NT, Lay>(i, tid, e, t);
    v[i] = base[e * rstride + t];
  }
  fft_ct<16, INV, NT, Lay, P>(v, lds, tw + (INV ? ct_twsize(P, false) : 0), tid);
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    int e, t;
    ct_elem<RL, L, NT, Lay>(i, tid, e, t);
    base[e * rstride + t] = v[i];
  }
}

bool c2c_supported(int L) { return L == 64 || L == 128 || L == 256; }

void c2c_plan(int L, FftPlan* p) {
  CtPlan c{};
  if (L == 64) c = C3<64>::P;
  if (L == 128) c = C3<128>::P;
  if (L == 256) c = C3<256>::P;
  p->L = c.L;
  p->np = c.np;
  for (int i = 0; i < kMaxPasses; ++i) p->rad[i] = i < c.np ? c.rad[i] : 1;
}

template <int L>
static hipError_t c2c_go(cx<float>* d, const cx<float>* tw, int nbatch, bool inv,
                         hipStream_t st) {
  const long long L2 = (long long)L * L, L3 = L2 * L;
  const dim3 blk(L);
  // z (contiguous), then y, then x; the order is immaterial for a separable transform
  const unsigned rows_blocks = (unsigned)(nbatch * L2 / 16);
  const unsigned y_blocks = (unsigned)(nbatch * L * (L / 16));
  const unsigned x_blocks = (unsigned)(nbatch * (L2 / 16));
  if (inv) {
    hipLaunchKernelGGL((k_c2c_rows<L, true>), dim3(rows_blocks), blk, 0, st, d, tw);
    hipLaunchKernelGGL((k_c2c_cols<L, true>), dim3(y_blocks), blk, 0, st, d, tw, (long long)L,
                       L, L2);
    hipLaunchKernelGGL((k_c2c_cols<L, true>), dim3(x_blocks), blk, 0, st, d, tw, L2, (int)L2,
                       L3);
  } else {
    hipLaunchKernelGGL((k_c2c_rows<L, false>), dim3(rows_blocks), blk, 0, st, d, tw);
    hipLaunchKernelGGL((k_c2c_cols<L, false>), dim3(y_blocks), blk, 0, st, d, tw, (long long)L,
                       L, L2);
    hipLaunchKernelGGL((k_c2c_cols<L, false>), dim3(x_blocks), blk, 0, st, d, tw, L2, (int)L2,
                       L3);
  }
  return hipGetLastError();
}

hipError_t launch_c2c3(cx<float>* d, const cx<float>* tw, int L, int nbatch, bool inv,
                       hipStream_t st) {
  switch (L) {
    case 64: return c2c_go<64>(d, tw, nbatch, inv, st);
    case 128: return c2c_go<128>(d, tw, nbatch, inv, st);
    case 256: return c2c_go<256>(d, tw, nbatch, inv, st);
    default: return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------------- spectral helpers
__device__ __forceinline__ int wavenum(int i, int L) { return i < L / 2 ? i : i - L; }
__device__ __forceinline__ int wrap(int i, int L) { return i < 0 ? i + L : i; }
// Retained by the 2/3 rule on an L grid (3 |k| < L on every axis).
__device__ __forceinline__ bool kept(int kx, int ky, int kz, int L) {
  return 3 * abs(kx) < L && 3 * abs(ky) < L && 3 * abs(kz) < L;
}

// State (rfftn layout, components `ncomp`, rows (x, y), pitch) -> full spectra.
__global__ void k_les_expand(const cx<float>* __restrict__ half, int pitch, int L, int ncomp,
                             cx<float>* __restrict__ full) {
  const long long L3 = (long long)L * L * L;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncomp * L3) return;
  const int c = (int)(idx / L3);
  const long long r = idx - c * L3;
  const int z = (int)(r % L), y = (int)((r / L) % L), x = (int)(r / ((long long)L * L));
  const long long cst = (long long)L * L * pitch;
  cx<float> v;
  if (z <= L / 2) {
    v = half[c * cst + ((long long)x * L + y) * pitch + z];
  } else {
    const int xm = x ? L - x : 0, ym = y ? L - y : 0;
    v = conjg(half[c * cst + ((long long)xm * L + ym) * pitch + (L - z)]);
  }
  full[idx] = v;
}

// Pairs for one inverse transform of the velocity and its gradient (scale 1/L^3 applied):
// (u, v), (w, du/dx), (du/dy, du/dz), (dv/dx, dv/dy), (dv/dz, dw/dx), (dw/dy, dw/dz).
// npairs = 1 keeps only (u, v) and writes (w, 0) as a second pair: velocity only.
__global__ void k_les_gradpack(const cx<float>* __restrict__ full, int L, int npairs,
                               float scale, cx<float>* __restrict__ out) {
  const long long L3 = (long long)L * L * L;
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= L3) return;
  const int z = (int)(r % L), y = (int)((r / L) % L), x = (int)(r / ((long long)L * L));
  const float k[3] = {(float)wavenum(x, L), (float)wavenum(y, L), (float)wavenum(z, L)};
  const bool nyq = x == L / 2 || y == L / 2 || z == L / 2;
  cx<float> u[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) u[c] = cscl(scale, full[c * L3 + r]);
  auto pair = [](cx<float> a, cx<float> b) { return cadd(a, mul_i(b)); };
  if (npairs == 1) {
    out[r] = pair(u[0], u[1]);
    out[L3 + r] = u[2];
    return;
  }
  // g[i][j] = d u_i / d x_j = i k_j u_i (zero on the Nyquist planes: odd derivative)
  cx<float> g[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) g[i][j] = nyq ? czero<float>() : cscl(k[j], mul_i(u[i]));
  out[r] = pair(u[0], u[1]);
  out[L3 + r] = pair(u[2], g[0][0]);
  out[2 * L3 + r] = pair(g[0][1], g[0][2]);
  out[3 * L3 + r] = pair(g[1][0], g[1][1]);
  out[4 * L3 + r] = pair(g[1][2], g[2][0]);
  out[5 * L3 + r] = pair(g[2][1], g[2][2]);
}

// Physical (u, v), (w, .) pairs -> product pairs (uu, vv), (ww, uv), (uw, vw).
__global__ void k_les_products(const cx<float>* __restrict__ vel, long long L3,
                               cx<float>* __restrict__ out) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= L3) return;
  const float u = vel[r].x, v = vel[r].y, w = vel[L3 + r].x;
  out[r] = mkc(u * u, v * v);
  out[L3 + r] = mkc(w * w, u * v);
  out[2 * L3 + r] = mkc(u * w, v * w);
}

// Fine pair spectra (npairs, full, Lf^3) -> coarse full spectra of the 2 npairs real fields
// (Lc^3 each): modes retained on the coarse grid only, times the box-filter transfer
// function (kc > 0) and (Lc / Lf)^3.  packed = 0: the input already holds one real field's
// spectrum per array (no pair split; nfields = npairs).
__global__ void k_les_truncate(const cx<float>* __restrict__ fine, int Lf, int npairs,
                               int packed, float kc, cx<float>* __restrict__ coarse, int Lc) {
  const long long Lc3 = (long long)Lc * Lc * Lc, Lf3 = (long long)Lf * Lf * Lf;
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= Lc3) return;
  const int z = (int)(r % Lc), y = (int)((r / Lc) % Lc), x = (int)(r / ((long long)Lc * Lc));
  const int kx = wavenum(x, Lc), ky = wavenum(y, Lc), kz = wavenum(z, Lc);
  const int nout = packed ? 2 * npairs : npairs;
  if (!kept(kx, ky, kz, Lc)) {
    for (int c = 0; c < nout; ++c) coarse[c * Lc3 + r] = czero<float>();
    return;
  }
  float g = (float)((double)Lc * Lc * Lc / ((double)Lf * Lf * Lf));
  if (kc > 0.f) {
    const float hd = 0.5f * 3.14159265358979f / kc;  // D / 2
    const int ks[3] = {kx, ky, kz};
#pragma unroll
    for (int i = 0; i < 3; ++i)
      if (ks[i] != 0) g *= __sinf(ks[i] * hd) / (ks[i] * hd);
  }
  const long long fi = ((long long)wrap(kx, Lf) * Lf + wrap(ky, Lf)) * Lf + wrap(kz, Lf);
  const long long fm = ((long long)wrap(-kx, Lf) * Lf + wrap(-ky, Lf)) * Lf + wrap(-kz, Lf);
  for (int p = 0; p < npairs; ++p) {
    const cx<float> zk = fine[p * Lf3 + fi];
    if (!packed) {
      coarse[p * Lc3 + r] = cscl(g, zk);
      continue;
    }
    const cx<float> zm = conjg(fine[p * Lf3 + fm]);
    const cx<float> a = cscl(0.5f * g, cadd(zk, zm));
    const cx<float> b = cscl(0.5f * g, mul_mi(csub(zk, zm)));
    coarse[2 * p * Lc3 + r] = a;
    coarse[(2 * p + 1) * Lc3 + r] = b;
  }
}

// Pair up full spectra of real fields (2 npairs of them, Lc^3 each) for one inverse
// transform, times scale.
__global__ void k_les_pairup(const cx<float>* __restrict__ f, long long L3, int npairs,
                             float scale, cx<float>* __restrict__ out) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= L3) return;
  for (int p = 0; p < npairs; ++p)
    out[p * L3 + r] = cscl(scale, cadd(f[2 * p * L3 + r], mul_i(f[(2 * p + 1) * L3 + r])));
}

// Model input: the nine gradient components (pairs 1..5 of k_les_gradpack, physical) times
// inv_g, as bf16 channels 0..8 of a 32-channel point-major tensor (channels 9..31 zero).
__global__ void k_les_features(const cx<float>* __restrict__ ph, long long L3, float inv_g,
                               __bf16* __restrict__ X) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= L3) return;
  const float gv[9] = {ph[L3 + r].y,     ph[2 * L3 + r].x, ph[2 * L3 + r].y,
                       ph[3 * L3 + r].x, ph[3 * L3 + r].y, ph[4 * L3 + r].x,
                       ph[4 * L3 + r].y, ph[5 * L3 + r].x, ph[5 * L3 + r].y};
  __bf16* o = X + r * kCnnC;
#pragma unroll
  for (int c = 0; c < kCnnC; ++c) o[c] = (__bf16)(c < 9 ? gv[c] * inv_g : 0.f);
}

// Smagorinsky stress tau_ij = -2 (Cs D)^2 |S| S_ij from the physical gradient pairs, as the
// pairs (t11, t22), (t33, t12), (t13, t23).
__global__ void k_les_smagorinsky(const cx<float>* __restrict__ ph, long long L3, float csd2,
                                  cx<float>* __restrict__ tau) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= L3) return;
  const float g00 = ph[L3 + r].y, g01 = ph[2 * L3 + r].x, g02 = ph[2 * L3 + r].y;
  const float g10 = ph[3 * L3 + r].x, g11 = ph[3 * L3 + r].y, g12 = ph[4 * L3 + r].x;
  const float g20 = ph[4 * L3 + r].y, g21 = ph[5 * L3 + r].x, g22 = ph[5 * L3 + r].y;
  const float s11 = g00, s22 = g11, s33 = g22;
  const float s12 = 0.5f * (g01 + g10), s13 = 0.5f * (g02 + g20), s23 = 0.5f * (g12 + g21);
  const float ss = s11 * s11 + s22 * s22 + s33 * s33 + 2.f * (s12 * s12 + s13 * s13 + s23 * s23);
  const float c = -2.f * csd2 * sqrtf(2.f * ss);
  tau[r] = mkc(c * s11, c * s22);
  tau[L3 + r] = mkc(c * s33, c * s12);
  tau[2 * L3 + r] = mkc(c * s13, c * s23);
}

// Network output (fp32, point-major, channels t11 t22 t33 t12 t13 t23 first) times scale as
// stress pairs.  With the physical gradient pairs ph (else null): backscatter damping, where
// the stress would move energy from the subgrid to the resolved scales, -tau_ij S_ij < 0
// (S the resolved strain rate), it is multiplied by bs (0: clipped); and, for csd2 > 0, the
// Smagorinsky stress -2 csd2 |S| S_ij added to it (a mixed model; csd2 = (Cs Delta)^2).
__global__ void k_les_tau_from_net(const float* __restrict__ Y, const cx<float>* __restrict__ ph,
                                   long long L3, float scale, float bs, float csd2,
                                   cx<float>* __restrict__ tau) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= L3) return;
  const float* y = Y + r * kCnnC;
  float t[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) t[c] = scale * y[c];
  if (ph) {
    const float g00 = ph[L3 + r].y, g01 = ph[2 * L3 + r].x, g02 = ph[2 * L3 + r].y;
    const float g10 = ph[3 * L3 + r].x, g11 = ph[3 * L3 + r].y, g12 = ph[4 * L3 + r].x;
    const float g20 = ph[4 * L3 + r].y, g21 = ph[5 * L3 + r].x, g22 = ph[5 * L3 + r].y;
    const float ts = t[0] * g00 + t[1] * g11 + t[2] * g22 +
                     t[3] * (g01 + g10) + t[4] * (g02 + g20) + t[5] * (g12 + g21);
    if (ts > 0.f) {
#pragma unroll
      for (int c = 0; c < 6; ++c) t[c] *= bs;
    }
    if (csd2 > 0.f) {
      const float s12 = 0.5f * (g01 + g10), s13 = 0.5f * (g02 + g20), s23 = 0.5f * (g12 + g21);
      const float ss = g00 * g00 + g11 * g11 + g22 * g22 +
                       2.f * (s12 * s12 + s13 * s13 + s23 * s23);
      const float c = -2.f * csd2 * sqrtf(2.f * ss);
      t[0] += c * g00;
      t[1] += c * g11;
      t[2] += c * g22;
      t[3] += c * s12;
      t[4] += c * s13;
      t[5] += c * s23;
    }
  }
  tau[r] = mkc(t[0], t[1]);
  tau[L3 + r] = mkc(t[2], t[3]);
  tau[2 * L3 + r] = mkc(t[4], t[5]);
}

// Subgrid-stress target tau_ij = filt(u_i u_j) - filt(u_i) filt(u_j) in physical space from
// the filtered-product pairs (uu, vv), (ww, uv), (uw, vw) and the filtered velocity pairs
// (u, v), (w, .), times scale; written point-major with kTauC channels (6 used).
__global__ void k_les_target(const cx<float>* __restrict__ fp, const cx<float>* __restrict__ vel,
                             long long L3, float scale, float* __restrict__ tau) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= L3) return;
  const float u = vel[r].x, v = vel[r].y, w = vel[L3 + r].x;
  float* o = tau + r * kTauC;
  o[0] = scale * (fp[r].x - u * u);
  o[1] = scale * (fp[r].y - v * v);
  o[2] = scale * (fp[L3 + r].x - w * w);
  o[3] = scale * (fp[L3 + r].y - u * v);
  o[4] = scale * (fp[2 * L3 + r].x - u * w);
  o[5] = scale * (fp[2 * L3 + r].y - v * w);
  o[6] = 0.f;
  o[7] = 0.f;
}

// Forward-transformed stress pairs -> S_i = -i k_j tau_ij, Leray-projected, 2/3-masked, in the
// state layout (3 components of L rows x L rows x pitch; columns kz <= L/2); with gain, times
// gain[(int)(|k| + 0.5)] (the shell index of the spectrum).
__global__ void k_les_source(const cx<float>* __restrict__ th, int L, int pitch,
                             const float* __restrict__ gain, cx<float>* __restrict__ src) {
  const int nh = L / 2 + 1;
  const long long n = (long long)L * L * nh;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int z = (int)(idx % nh);
  const long long row = idx / nh;
  const int y = (int)(row % L), x = (int)(row / L);
  const int kx = wavenum(x, L), ky = wavenum(y, L), kz = z;
  const long long cst = (long long)L * L * pitch;
  const long long so = row * pitch + z;
  if (!kept(kx, ky, kz, L)) {
    for (int c = 0; c < 3; ++c) src[c * cst + so] = czero<float>();
    return;
  }
  const long long L3 = (long long)L * L * L;
  const long long fi = ((long long)x * L + y) * L + z;
  const long long fm = ((long long)wrap(-kx, L) * L + wrap(-ky, L)) * L + wrap(-kz, L);
  cx<float> t[6];
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const cx<float> zk = th[p * L3 + fi], zm = conjg(th[p * L3 + fm]);
    t[2 * p] = cscl(0.5f, cadd(zk, zm));
    t[2 * p + 1] = cscl(0.5f, mul_mi(csub(zk, zm)));
  }
  // t: 11, 22, 33, 12, 13, 23
  const float fx = (float)kx, fy = (float)ky, fz = (float)kz;
  const cx<float> d0 = cadd(cadd(cscl(fx, t[0]), cscl(fy, t[3])), cscl(fz, t[4]));
  const cx<float> d1 = cadd(cadd(cscl(fx, t[3]), cscl(fy, t[1])), cscl(fz, t[5]));
  const cx<float> d2 = cadd(cadd(cscl(fx, t[4]), cscl(fy, t[5])), cscl(fz, t[2]));
  cx<float> s0 = mul_mi(d0), s1 = mul_mi(d1), s2 = mul_mi(d2);
  const int k2 = kx * kx + ky * ky + kz * kz;
  const float ik2 = k2 > 0 ? 1.f / (float)k2 : 0.f;
  const cx<float> kd = cscl(ik2, cadd(cadd(cscl(fx, s0), cscl(fy, s1)), cscl(fz, s2)));
  s0 = csub(s0, cscl(fx, kd));
  s1 = csub(s1, cscl(fy, kd));
  s2 = csub(s2, cscl(fz, kd));
  if (gain) {
    // |k|^2 <= 3 (L/2)^2 is an integer, so sqrt(k2) is never within fp32 rounding of n + 1/2
    const float g = gain[(int)(sqrtf((float)k2) + 0.5f)];
    s0 = cscl(g, s0);
    s1 = cscl(g, s1);
    s2 = cscl(g, s2);
  }
  src[so] = s0;
  src[cst + so] = s1;
  src[2 * cst + so] = s2;
}

// FP64 sums of squares of n values (strided: every `stride`-th, `count` per group of width
// `width`), one atomic per work-group; out must start at zero.
__global__ void k_les_sumsq(const float* __restrict__ x, long long n, int width, int count,
                            double* out) {
  __shared__ double red[256];
  double a = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    if ((int)(i % width) < count) a += (double)x[i] * (double)x[i];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

static inline dim3 grid1(long long n) { return dim3((unsigned)((n + 255) / 256)); }

hipError_t launch_les(int which, const LesArgs& a, hipStream_t st) {
  const long long L3 = (long long)a.L * a.L * a.L;
  switch (which) {
    case kLesExpand:
      hipLaunchKernelGGL(k_les_expand, grid1(a.n * L3), dim3(256), 0, st, a.cin, a.pitch, a.L,
                         (int)a.n, a.cout);
      break;
    case kLesGradPack:
      hipLaunchKernelGGL(k_les_gradpack, grid1(L3), dim3(256), 0, st, a.cin, a.L, (int)a.n,
                         a.scale, a.cout);
      break;
    case kLesProducts:
      hipLaunchKernelGGL(k_les_products, grid1(L3), dim3(256), 0, st, a.cin, L3, a.cout);
      break;
    case kLesTruncate: {
      const long long Lc3 = (long long)a.L2 * a.L2 * a.L2;
      hipLaunchKernelGGL(k_les_truncate, grid1(Lc3), dim3(256), 0, st, a.cin, a.L, (int)a.n,
                         a.pitch, a.scale, a.cout, a.L2);
      break;
    }
    case kLesPairUp:
      hipLaunchKernelGGL(k_les_pairup, grid1(L3), dim3(256), 0, st, a.cin, L3, (int)a.n,
                         a.scale, a.cout);
      break;
    case kLesFeatures:
      hipLaunchKernelGGL(k_les_features, grid1(L3), dim3(256), 0, st, a.cin, L3, a.scale,
                         static_cast<__bf16*>(a.out));
      break;
    case kLesSmagorinsky:
      hipLaunchKernelGGL(k_les_smagorinsky, grid1(L3), dim3(256), 0, st, a.cin, L3, a.scale,
                         a.cout);
      break;
    case kLesTauFromNet:
      hipLaunchKernelGGL(k_les_tau_from_net, grid1(L3), dim3(256), 0, st,
                         static_cast<const float*>(a.in2), a.cin, L3, a.scale, a.scale2,
                         a.scale3, a.cout);
      break;
    case kLesTarget:
      hipLaunchKernelGGL(k_les_target, grid1(L3), dim3(256), 0, st, a.cin,
                         static_cast<const cx<float>*>(a.in2), L3, a.scale,
                         static_cast<float*>(a.out));
      break;
    case kLesSource: {
      const long long nm = (long long)a.L * a.L * (a.L / 2 + 1);
      hipLaunchKernelGGL(k_les_source, grid1(nm), dim3(256), 0, st, a.cin, a.L, a.pitch,
                         static_cast<const float*>(a.in2), a.cout);
      break;
    }
    case kLesSumSq:
      hipLaunchKernelGGL(k_les_sumsq, dim3(1024), dim3(256), 0, st,
                         static_cast<const float*>(a.in2), a.n, a.pitch, a.L2,
                         static_cast<double*>(a.out));
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace psx
