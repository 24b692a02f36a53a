// 2D fp32 stage-kernel bodies, specialised at compile time on a launch geometry (psx_geo.h):
// the single-flow kernels (psx_k2d_ct.hip) run them with one work-group per tile, the
// ensemble kernels (psx_kens.hip) with persistent work-groups walking a list of (flow, tile)
// items of flows of several sizes.
//
// Work-field convention of this path at the real-axis boundary: the line kernel reads and
// writes each row's complex line spectrum Z directly, Z(e) = U(e) + i V(e) for the packed pair
// of real fields, stored as two half spectra: field 0 holds Z(j) and field 1 holds Z(n - j)
// for the retained j < jm (field 1, column 0 unused).  The column kernels pack (U, V) into Z
// after their inverse transform (Z(j) = U + iV, Z(n-j) = conj(U) + i conj(V), imaginary parts
// of the j = 0 column dropped as a c2r does) and split Z back into the two real fields'
// spectra before their forward transform (P = (Z(j) + conj Z(n-j))/2,
// Q = -i (Z(j) - conj Z(n-j))/2).  The line kernel then needs one load and one store per
// element and no LDS pass of its own.
#pragma once

#include "psx_fft_ct.h"
#include "psx_kernels.h"

namespace psx {

// Split of the two real fields' spectra from a packed line spectrum at column j.
__device__ __forceinline__ void z_split(cx<float> z, cx<float> zm, cx<float>& a, cx<float>& b) {
  const cx<float> zc = conjg(zm);
  a = cscl(0.5f, cadd(z, zc));
  b = cscl(0.5f, mul_mi(csub(z, zc)));
}

// Packed line spectrum from the two real fields' spectra at column j.
__device__ __forceinline__ void z_pack(int j, cx<float> u, cx<float> v, cx<float>& z,
                                       cx<float>& zm) {
  if (j == 0) {
    z = mkc(u.x, v.x);
    zm = z;
  } else {
    z = mkc(u.x - v.y, u.y + v.x);
    zm = mkc(u.x + v.y, v.x - u.y);
  }
}

// --------------------------------------------------------------------------- line kernel
// Rows [tile * lineT, (tile + 1) * lineT) of flow f; lineNT threads, tid in [0, lineNT).
// TWL: the twiddles of every pass but the last of each direction are read from an LDS copy
// after the row tiles (the caller allocates twl_line2d_elems<G>() more elements); the last
// passes' tables (60 KB each at 8192) stay in global memory.
template <Geo2D G>
constexpr int twl_line2d_elems() {
  return ct_twoff(G.line, false, G.line.np - 1) + ct_twoff(G.line, true, G.line.np - 1);
}

template <Geo2D G, bool TWL = false>
__device__ __forceinline__ void body_line2d(const Flow<float>& f, int tile, int tid, int flags,
                                            unsigned char* smem_raw) {
  using T = float;
  constexpr int N = G.n;
  constexpr int E = G.lineE, NT = G.lineNT, RT = G.lineT, L = N, JM = G.jm, LDW = G.ldw;
  constexpr long long FS = (long long)N * LDW;
  using Lay = RowsCT<L, RT>;
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int row0 = tile * RT;
  const BufC32 zb(f.W, (unsigned long long)(2 * FS) * 8);
  const cx<T>* tw_fwd = f.twp + f.prow.twoff[0][0];
  const cx<T>* tw_rev = f.twp + f.prow.twoff[1][0];
  constexpr int QG = G.line.np - 1;
  constexpr int TWNF = ct_twoff(G.line, false, QG);
  cx<T>* twl = lds + RT * Lay::LP;
  if constexpr (TWL) {
    constexpr int TWN = twl_line2d_elems<G>();
    for (int i = tid; i < TWN; i += NT) twl[i] = i < TWNF ? tw_fwd[i] : tw_rev[i - TWNF];
    __syncthreads();
  }
  // byte offset of element e of row-in-tile t: field 0 for e < JM, field 1 (mirrored) for
  // e > L - JM, out of range (zero / dropped) in between
  auto zoff = [&](int e, int t) -> unsigned {
    const unsigned base = (unsigned)(row0 + t) * LDW;
    return e < JM ? (base + e) * 8u
                  : (e > L - JM ? ((unsigned)FS + base + (L - e)) * 8u : kOob);
  };

  // slots in the dealiased gap JM <= e <= L - JM for every thread hold zeros: no load, and the
  // first inverse butterflies fold them; the matching outputs of the last forward pass are
  // never stored, so their arithmetic is dropped
  cx<T> v[E];
  {
    constexpr int R = ct_radix(G.line, true, 0);
    static_for<0, E>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if constexpr (ct_slot_within<R, L, NT, Lay>(i, JM, L - JM)) {
        v[i] = czero<T>();
      } else {
        int e, t;
        ct_elem<R, L, NT, Lay>(i, tid, e, t);
        v[i] = zb.load(zoff(e, t));
      }
    });
  }
  if constexpr (TWL) {
    // 8192-point rows load the last pass's twiddles one pass early (line kernel 263 -> 259.5
    // us); at 4096 the extra registers would drop it to one wave per SIMD
    if constexpr (L == 8192)
      fft_ct_tw2_pre<E, true, NT, Lay, G.line>(v, lds, twl + TWNF, tw_rev, tid);
    else
      fft_ct_tw2<E, true, NT, Lay, G.line, QG>(v, lds, twl + TWNF, tw_rev, tid);
  } else {
    fft_ct<E, true, NT, Lay, G.line>(v, lds, tw_rev, tid);
  }

  if (flags & kLineWantMax) {
    float m[2] = {0.f, 0.f};  // |u_i| maxima (exact in fp32)
#pragma unroll
    for (int i = 0; i < E; ++i) {
      m[0] = fmaxf(m[0], abs_or_inf(v[i].x));
      m[1] = fmaxf(m[1], abs_or_inf(v[i].y));
    }
    double md[2] = {(double)m[0], (double)m[1]};
      block_max_to<2>(md, f.stats, reinterpret_cast<double*>(smem_raw));
  }
  if (flags & kLineMaxOnly) return;

#pragma unroll
  for (int i = 0; i < E; ++i) {
    const T u = v[i].x, w = v[i].y;
    v[i] = mkc((w - u) * (w + u), u * w);
  }
  if constexpr (TWL) {
    if constexpr (L == 8192)
      fft_ct_tw2_pre<E, false, NT, Lay, G.line>(v, lds, twl, tw_fwd, tid);
    else
      fft_ct_tw2<E, false, NT, Lay, G.line, QG>(v, lds, twl, tw_fwd, tid);
  } else {
    fft_ct<E, false, NT, Lay, G.line>(v, lds, tw_fwd, tid);
  }
  {
    constexpr int R = ct_radix(G.line, false, G.line.np - 1);
    static_for<0, E>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if constexpr (!ct_slot_within<R, L, NT, Lay>(i, JM, L - JM)) {
        int e, t;
        ct_elem<R, L, NT, Lay>(i, tid, e, t);
        zb.store(zoff(e, t), v[i]);
      }
    });
  }
}

// ----------------------------------------------------------------------- spectral kernel
// Column tile `tile` (group tile / NJT of the outer split, columns (tile % NJT) * colT ..)
// of flow f; specNT threads, tid in [0, specNT).  MS: the two fields' transforms of one
// direction run together (fft_ct_multi, two LDS regions of n2 * colT elements) instead of one
// after the other in one region.
// TWL: the pass twiddles are read from an LDS copy after the transform regions (the caller
// allocates it; twl_spec2d_elems elements).
template <Geo2D G, int stage, bool MS = false, bool TWL = false>
__device__ __forceinline__ void body_spec2d(const Flow<float>& f, int tile, int tid,
                                            unsigned char* smem_raw) {
  using T = float;
  constexpr int N = G.n;
  constexpr bool FOUR = G.n1 > 1;
  constexpr int E = G.specE, NT = G.specNT, CT = G.colT, L = G.n2, N1 = G.n1, JM = G.jm;
  constexpr int LDW = G.ldw, NH = N / 2 + 1, SP = G.sp;
  constexpr long long FS = (long long)N * LDW;
  constexpr int NJT = LDW / CT;
  using Lay = ColsCT<CT>;
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int grp = tile / NJT;
  const int j0 = (tile - grp * NJT) * CT;
  const int eo = FOUR ? grp * L : 0;
  cx<T>* W0 = f.W;
  cx<T>* W1 = f.W + FS;
  const cx<T>* tw_fwd = f.twp + f.pcol.twoff[0][0];
  const cx<T>* tw_rev = f.twp + f.pcol.twoff[1][0];
  // TWL: LDS copies of the pass twiddles, of the FP64 integrating-factor table entries of this
  // tile's rows (|kx|) and columns (ky) (ifl[0, L) = t1(|kx|), [L, L + CT) = t1(ky), then the
  // same for the half-step table t_1/2) and, four-step, of the tile's w^(e grp) twiddles.
  // Read from global memory after the transforms, their latency sat on the critical path
  // (8192^2 spectral kernel 246 -> 209 us with the pass twiddles and factors).
  [[maybe_unused]] double* ifl = nullptr;
  if constexpr (TWL) {
    constexpr int TWF = ct_twsize(G.spec, false), TWR = ct_twsize(G.spec, true);
    cx<T>* twl = lds + (MS ? 2 : 1) * L * CT;
    for (int i = tid; i < TWF + TWR; i += NT) twl[i] = i < TWF ? tw_fwd[i] : tw_rev[i - TWF];
    ifl = reinterpret_cast<double*>(twl + TWF + TWR);
    for (int i = tid; i < 2 * (L + CT); i += NT) {
      const int h = i < L + CT ? 0 : 1, r = i - h * (L + CT);
      int idx;
      if (r < L) {
        const int k = FOUR ? grp + N1 * r : r;
        const int kx = k < N / 2 ? k : k - N;
        idx = kx < 0 ? -kx : kx;
      } else {
        idx = j0 + (r - L);
      }
      ifl[i] = f.prm[h * NH + (idx < NH ? idx : NH - 1)];
    }
    if constexpr (FOUR) {  // the four-step twiddles w^(e grp) of the tile's rows
      cx<T>* twg = reinterpret_cast<cx<T>*>(ifl + 2 * (L + CT));
      for (int e = tid; e < L; e += NT) twg[e] = f.tw[e * grp];
    }
    __syncthreads();
    tw_fwd = twl;
    tw_rev = twl + TWF;
  }

  cx<T> vv[2][E];
  cx<T>(&va)[E] = vv[0];
  cx<T>(&vb)[E] = vv[1];
  if constexpr (stage != kStagePre) {
    constexpr int R = ct_radix(G.spec, false, 0);
#pragma unroll
    for (int i = 0; i < E; ++i) {
      int e, t;
      ct_elem<R, L, NT, Lay>(i, tid, e, t);
      const long long a = (long long)(eo + e) * LDW + j0 + t;
      if constexpr (FOUR) {
        va[i] = W0[a];
        vb[i] = W1[a];
      } else {
        const cx<T> z = W0[a];
        const cx<T> zm = W1[a];
        z_split(z, j0 + t == 0 ? z : zm, va[i], vb[i]);
      }
    }
  }
  // State and accumulator of this thread's modes, loaded before the transforms so their
  // latency hides behind them; modes outside the 2/3 mask read zero and are never written.
  // With 16 elements per thread the accumulator is loaded after the transforms instead: held
  // across them it pushed the kernel past 256 VGPRs into scratch.
  constexpr bool kEarlyAcc = E <= 8;
  constexpr int RP = ct_radix(G.spec, false, G.spec.np - 1);
  const BufC32 ub(f.u0, (unsigned long long)N * SP * 8);
  const BufC32 ab(f.acc, (unsigned long long)N * SP * 8);
  unsigned soff[E];
  cx<T> uu[E], aa[kEarlyAcc ? E : 1];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    int e, t;
    ct_elem<RP, L, NT, Lay>(i, tid, e, t);
    const int k = FOUR ? grp + N1 * e : e;
    const int kx = k < N / 2 ? k : k - N;
    const int ax = kx < 0 ? -kx : kx;
    const int ky = j0 + t;
    soff[i] = (3 * ax < N && ky < JM) ? (unsigned)(k * SP + ky) * 8u : kOob;
    uu[i] = ub.load(soff[i]);
  }
  if constexpr (kEarlyAcc) {
    if constexpr (stage >= 1) {
#pragma unroll
      for (int i = 0; i < E; ++i) aa[i] = ab.load(soff[i]);
    } else {
#pragma unroll
      for (int i = 0; i < E; ++i) aa[i] = czero<T>();
    }
  }
  if constexpr (stage != kStagePre) {
    if constexpr (MS) {
      fft_ct_multi<2, E, false, NT, Lay, G.spec, L * CT>(vv, lds, tw_fwd, tid);
    } else {
      fft_ct<E, false, NT, Lay, G.spec>(va, lds, tw_fwd, tid);
      fft_ct<E, false, NT, Lay, G.spec>(vb, lds, tw_fwd, tid);
    }
  }
  {
    const IfScalars s = load_if_scalars(f.prm, NH);
    const T dt = (T)f.prm[2 * NH + PrmLayout::kDt];
    const T scale = f.scale;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      int e, t;
      ct_elem<RP, L, NT, Lay>(i, tid, e, t);
      const int k = FOUR ? grp + N1 * e : e;
      const int kx = k < N / 2 ? k : k - N;
      const int ky = j0 + t;
      const int ax = kx < 0 ? -kx : kx;
      const int k2 = kx * kx + ky * ky;
      cx<T> nxt = uu[i];
      if constexpr (stage != kStagePre) {
        const T fx = (T)kx, fy = (T)ky;
        const cx<T> Nl = cadd(cscl(fx * fy, va[i]), cscl(fx * fx - fy * fy, vb[i]));
        double E1, Eh;
        if constexpr (TWL) {
          // the operations of if_factors on the LDS copies of the table entries
          E1 = ifl[e] * ifl[L + t];
          Eh = ifl[L + CT + e] * ifl[2 * L + CT + t];
          const double dk2 = (double)k2;
          if (dk2 <= s.ka2) {
            E1 *= s.ea1;
            Eh *= s.eah;
          }
          if (s.band && dk2 >= s.klo2 && dk2 <= s.khi2) {
            E1 *= s.eb1;
            Eh *= s.ebh;
          }
        } else {
          if_factors(f.prm, NH, s, ax, ky, -1, k2, E1, Eh);
        }
        cx<T> acc = czero<T>();
        if constexpr (kEarlyAcc) {
          acc = aa[i];
        } else if constexpr (stage >= 1) {
          acc = ab.load(soff[i]);
        }
        nxt = rk_value<stage, T>(dt, (T)(1.0 - E1), (T)(1.0 - Eh), Nl, uu[i], acc);
        if constexpr (stage < 3) ab.store(soff[i], acc);
        else ub.store(soff[i], nxt);
      }
      if (soff[i] == kOob) nxt = czero<T>();  // 2/3 mask: no nonlinear term outside it
      const T ik2 = k2 > 0 ? scale * __builtin_amdgcn_rcpf((T)k2) : (T)0;
      const cx<T> psi = cscl(ik2, nxt);
      va[i] = cscl((T)ky, mul_i(psi));
      vb[i] = cscl((T)kx, mul_mi(psi));
    }
  }
  constexpr int RS = ct_radix(G.spec, true, G.spec.np - 1);
  if constexpr (MS) {
    fft_ct_multi<2, E, true, NT, Lay, G.spec, L * CT>(vv, lds, tw_rev, tid);
  } else {
    fft_ct<E, true, NT, Lay, G.spec>(va, lds, tw_rev, tid);
    fft_ct<E, true, NT, Lay, G.spec>(vb, lds, tw_rev, tid);
  }
#pragma unroll
  for (int i = 0; i < E; ++i) {
    int e, t;
    ct_elem<RS, L, NT, Lay>(i, tid, e, t);
    const long long a = (long long)(eo + e) * LDW + j0 + t;
    if constexpr (FOUR) {
      cx<T> w;
      if constexpr (TWL) w = reinterpret_cast<const cx<T>*>(ifl + 2 * (L + CT))[e];
      else w = f.tw[e * grp];
      W0[a] = cmulc(va[i], w);
      W1[a] = cmulc(vb[i], w);
    } else {
      cx<T> z, zm;
      z_pack(j0 + t, va[i], vb[i], z, zm);
      W0[a] = z;
      W1[a] = zm;
    }
  }
}

// ------------------------------------------------------------------- outer four-step pass
// Outer-axis transforms of column tile `tile` (inner index tile / NJT, columns
// (tile % NJT) * colT ..) of flow f; outNT threads, tid in [0, outNT).  MS as in
// body_spec2d (two LDS regions of n1 * colT elements).
// WS: the tile's threads lie within one wave, so its exchanges use wave-level syncs (several
// tiles share a work-group without waiting on each other).
// TWL: tw_lds holds this direction's pass-twiddle table in LDS (copied by the caller).
template <Geo2D G, bool INV, bool MS = false, bool WS = false, bool TWL = false>
__device__ __forceinline__ void body_outer2d(const Flow<float>& f, int tile, int tid,
                                             unsigned char* smem_raw,
                                             const cx<float>* tw_lds = nullptr) {
  using T = float;
  constexpr int N = G.n;
  if constexpr (G.n1 > 1) {
    constexpr int E = G.outE, NT = G.outNT, CT = G.colT, L = G.n1, N2 = G.n2;
    constexpr int LDW = G.ldw, NJT = LDW / CT;
    constexpr long long FS = (long long)N * LDW;
    using Lay = ColsCT<CT>;
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const int x2 = tile / NJT;
    const int j0 = (tile - x2 * NJT) * CT;
    const cx<T>* tw_o = TWL ? tw_lds : f.twp + f.pa.twoff[INV ? 1 : 0][0];
    constexpr int R0 = ct_radix(G.outer, INV, 0);
    constexpr int RL = ct_radix(G.outer, INV, G.outer.np - 1);
    cx<T> vv[2][E];
    cx<T>(&v0)[E] = vv[0];
    cx<T>(&v1)[E] = vv[1];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      int e, t;
      ct_elem<R0, L, NT, Lay>(i, tid, e, t);
      const long long a = (long long)(e * N2 + x2) * LDW + j0 + t;
      if constexpr (INV) {
        v0[i] = f.W[a];
        v1[i] = f.W[FS + a];
      } else {
        const cx<T> z = f.W[a];
        const cx<T> zm = f.W[FS + a];
        z_split(z, j0 + t == 0 ? z : zm, v0[i], v1[i]);
      }
    }
    if constexpr (MS) {
      fft_ct_multi<2, E, INV, NT, Lay, G.outer, L * CT, WS>(vv, lds, tw_o, tid);
    } else {
      fft_ct<E, INV, NT, Lay, G.outer, WS>(v0, lds, tw_o, tid);
      fft_ct<E, INV, NT, Lay, G.outer, WS>(v1, lds, tw_o, tid);
    }
#pragma unroll
    for (int i = 0; i < E; ++i) {
      int e, t;
      ct_elem<RL, L, NT, Lay>(i, tid, e, t);
      cx<T> z0 = v0[i], z1 = v1[i];
      if constexpr (!INV) {
        // the index is made opaque here, after both transforms: the compiler otherwise issues
        // the twiddle loads at the start and keeps them live across the transforms (48 more
        // VGPRs, three waves per SIMD fewer)
        int ti = e * x2;
        asm volatile("" : "+v"(ti));
        const cx<T> w = f.tw[ti];
        z0 = cmul(z0, w);
        z1 = cmul(z1, w);
      } else {
        z_pack(j0 + t, v0[i], v1[i], z0, z1);
      }
      const long long a = (long long)(e * N2 + x2) * LDW + j0 + t;
      f.W[a] = z0;
      f.W[FS + a] = z1;
    }
  }
}

}  // namespace psx
