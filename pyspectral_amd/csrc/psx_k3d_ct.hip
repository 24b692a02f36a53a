// 3D fp32 stage kernels specialised per grid size (PSX_CT3D_SIZES) and form: the algorithms of
// psx_k3d.hip with plans, launch shapes and strides fixed at compile time by geo3d()
// (psx_geo.h), branch-free masked state access, and packed line spectra at the z-axis
// boundary (as in psx_k2d_ct.hip):
//
//   z-line input (written by the y inverse pass):
//     basdevant   fields (0,1) = packed u + i v, field 2 = w_hat (one real field)
//     rotational  fields (0,1) = packed u + i v, (2,3) = w + i omega_x, (4,5) = omega_y + i omega_z
//   z-line output (read by the y forward pass):
//     basdevant   (0,1) = packed T11 + i T22, (2,3) = packed T12 + i T13, field 4 = T23
//     rotational  (0,1) = packed (u x omega)_x + i (u x omega)_y, field 2 = (u x omega)_z
// A packed pair stores Z(j) in its first field and Z(n - j) in its second (j < jm).
#include "psx_fft_ct.h"
#include "psx_kernels.h"

namespace psx {

// Cache policy of the 3D kernels' single-use accesses: non-temporal from 384^3 up; smaller
// grids' fields fit the caches (the 64^3 LES step and the 256^3 DNS were slower with it).
constexpr int kc3_pol(int n) { return n >= 384 ? kNt : kCached; }


template <int N>
struct G3 {
  static constexpr Geo3D g = geo3d(N);
  static_assert(g.ok, "no 3D plan for this size");
  static constexpr long long NR = (long long)N * N;      // rows per field
  static constexpr long long FS = NR * g.ldw;            // complex elements per field
  // element offsets of all work fields stay below the in-kernel out-of-range marker, and the
  // byte offsets below 2 GiB (raw buffer addressing)
  static_assert(6 * FS < 0x10000000LL, "work fields too large for 32-bit buffer offsets");
  static_assert(3 * NR * g.sp * 8 < 0x80000000LL, "state too large for 32-bit offsets");
};

__device__ __forceinline__ void z_split3(cx<float> z, cx<float> zm, cx<float>& a, cx<float>& b) {
  const cx<float> zc = conjg(zm);
  a = cscl(0.5f, cadd(z, zc));
  b = cscl(0.5f, mul_mi(csub(z, zc)));
}

__device__ __forceinline__ void z_pack3(int j, cx<float> u, cx<float> v, cx<float>& z,
                                        cx<float>& zm) {
  if (j == 0) {
    z = mkc(u.x, v.x);
    zm = z;
  } else {
    z = mkc(u.x - v.y, u.y + v.x);
    zm = mkc(u.x + v.y, v.x - u.y);
  }
}

// ----------------------------------------------------------------------------- z lines
// Row tiles (G.lineT rows) per work-group of the z-line kernel: a work-group walks TPW
// consecutive tiles and loads tile k + 1's packed spectra while it transforms tile k, so the
// loads of one tile overlap the transforms of the previous one (384^3: 729 -> 699 us with four
// tiles, 716 us with eight).  Grids with fewer tiles keep one per work-group (256^3 gained
// nothing), and so does the rotational form, whose three input pairs in flight beside its
// transforms would exceed 256 VGPRs.
template <int N, bool ROT>
constexpr int line3d_tpw() {
  constexpr int tiles = N * N / G3<N>::g.lineT;
  return ROT || tiles < 16384 ? 1 : 4;
}


template <int N, bool ROT>
__global__ void __launch_bounds__(G3<N>::g.lineNT, 1) kc_line3d(Flow<float> f, int flags) {
  if (flow_inactive(f.prm, f.nh)) return;
  using T = float;
  constexpr Geo3D G = G3<N>::g;
  constexpr int E = G.lineE, NT = G.lineNT, RT = G.lineT, L = N, JM = G.jm, LDW = G.ldw;
  constexpr long long FS = G3<N>::FS;
  constexpr int NF = ROT ? 6 : 5;
  constexpr int TPW = line3d_tpw<N, ROT>();
  using Lay = RowsCT<L, RT>;
  static_assert(ct_xaddr_ok<Lay, false, G.line.np - 1>() && ct_xaddr_ok<Lay, true, G.line.np - 1>(),
                "exchange layouts");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const BufC32T<kc3_pol(N)> wb(f.W, (unsigned long long)NF * FS * 8);
  constexpr int RI = ct_radix(G.line, true, 0);
  constexpr int RO = ct_radix(G.line, false, G.line.np - 1);

  // LDS: NREG transform regions of ST elements (the transforms of one direction run together,
  // fft_ct_multi), then the plan's forward and reversed pass-twiddle tables.  The twiddles are
  // read from LDS because vector-memory counters retire in order: a twiddle load from global
  // memory issued after the next tile's loads would wait for them.
  constexpr int ST = RT * Lay::LP;
  constexpr int NREG = ROT ? 3 : 2;
  constexpr int TWF = ct_twsize(G.line, false), TWR = ct_twsize(G.line, true);
  cx<T>* twl = lds + NREG * ST;
  {
    const cx<T>* tw_fwd = f.twp + f.prow.twoff[0][0];
    const cx<T>* tw_rev = f.twp + f.prow.twoff[1][0];
    for (int i = tid; i < TWF + TWR; i += NT) twl[i] = i < TWF ? tw_fwd[i] : tw_rev[i - TWF];
  }
  __syncthreads();

  // packed spectra of field pair (fa, fb) (or single field fa, raw: the Hermitian extension is
  // applied when the tile is taken) for rows r0 .. r0 + RT - 1
  auto pair_off = [&](long long r0, int fa, int fb, int e, int t) -> unsigned {
    const unsigned base = (unsigned)((r0 + t) * LDW);
    return e < JM ? (unsigned)(fa * FS) + base + e
                  : (e > L - JM ? (unsigned)(fb * FS) + base + (L - e) : kOob / 8u);
  };
  auto load_pair = [&](cx<T>(&v)[E], long long r0, int fa, int fb, int tx) {
    static_for<0, E>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if constexpr (ct_slot_within<RI, L, NT, Lay>(i, JM, L - JM)) {
        v[i] = czero<T>();
      } else {
        int e, t;
        ct_elem<RI, L, NT, Lay>(i, tx, e, t);
        const unsigned o = pair_off(r0, fa, fb, e, t);
        v[i] = wb.load(o == kOob / 8u ? kOob : o * 8u);
      }
    });
  };
  auto load_single_raw = [&](cx<T>(&v)[E], long long r0, int fa, int tx) {
    static_for<0, E>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if constexpr (ct_slot_within<RI, L, NT, Lay>(i, JM, L - JM)) {
        v[i] = czero<T>();
      } else {
        int e, t;
        ct_elem<RI, L, NT, Lay>(i, tx, e, t);
        const unsigned base = (unsigned)((r0 + t) * LDW) + (unsigned)(fa * FS);
        const int src = e < JM ? e : (e > L - JM ? L - e : -1);
        v[i] = wb.load(src >= 0 ? (base + src) * 8u : kOob);
      }
    });
  };
  // Hermitian extension of a real line's spectrum; c2r ignores the imaginary part of the mean
  auto take_single = [&](cx<T>(&v)[E], int tx) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      int e, t;
      ct_elem<RI, L, NT, Lay>(i, tx, e, t);
      if (e > L - JM) v[i] = conjg(v[i]);
      if (e == 0) v[i].y = (T)0;
    }
  };
  auto store_pair = [&](const cx<T>(&v)[E], long long r0, int fa, int fb, int tx) {
    static_for<0, E>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if constexpr (!ct_slot_within<RO, L, NT, Lay>(i, JM, L - JM)) {
        int e, t;
        ct_elem<RO, L, NT, Lay>(i, tx, e, t);
        const unsigned o = pair_off(r0, fa, fb, e, t);
        wb.store(o == kOob / 8u ? kOob : o * 8u, v[i]);
      }
    });
  };
  // (a real field's spectrum is stored for e < JM only)
  auto store_single = [&](const cx<T>(&v)[E], long long r0, int fa, int tx) {
    static_for<0, E>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if constexpr (!ct_slot_within<RO, L, NT, Lay>(i, JM, L)) {
        int e, t;
        ct_elem<RO, L, NT, Lay>(i, tx, e, t);
        const unsigned base = (unsigned)((r0 + t) * LDW) + (unsigned)(fa * FS);
        wb.store(e < JM ? (base + e) * 8u : kOob, v[i]);
      }
    });
  };
  constexpr int NPF = ROT ? 3 : 2;  // raw input sequences per tile
  auto fetch = [&](cx<T>(&pf)[NPF][E], long long r0, int tx) {
    load_pair(pf[0], r0, 0, 1, tx);
    if constexpr (!ROT) {
      load_single_raw(pf[1], r0, 2, tx);
    } else {
      load_pair(pf[1], r0, 2, 3, tx);
      load_pair(pf[2], r0, 4, 5, tx);
    }
  };

  const long long rw0 = (long long)blockIdx.x * TPW * RT;
  cx<T> pf[NPF][E];
  fetch(pf, rw0, tid);
#pragma nounroll
  for (int k = 0; k < TPW; ++k) {
    const long long r0 = rw0 + (long long)k * RT;
    // the thread index and the twiddle base are made opaque per tile: the index arithmetic and
    // twiddle reads are the same in every tile, and the compiler would otherwise hoist them out
    // of the loop and keep them in registers across it
    int tk = tid, twz = 0;
    asm volatile("" : "+v"(tk), "+v"(twz));
    const cx<T>* tw_fwd_l = twl + twz;
    const cx<T>* tw_rev_l = twl + TWF + twz;
    cx<T> z[3][E];
#pragma unroll
    for (int s = 0; s < NPF; ++s) {
#pragma unroll
      for (int i = 0; i < E; ++i) z[s][i] = pf[s][i];
    }
    if constexpr (!ROT) {
      if (k + 1 < TPW) fetch(pf, r0 + RT, tk);
      take_single(z[1], tk);
      fft_ct_multi<2, E, true, NT, Lay, G.line, ST>(z, lds, tw_rev_l, tk);
      if (flags & kLineWantMax) {
        float m[3] = {0.f, 0.f, 0.f};  // |u_i| maxima (exact in fp32)
#pragma unroll
        for (int i = 0; i < E; ++i) {
          m[0] = fmaxf(m[0], abs_or_inf(z[0][i].x));
          m[1] = fmaxf(m[1], abs_or_inf(z[0][i].y));
          m[2] = fmaxf(m[2], abs_or_inf(z[1][i].x));
        }
        double md[3] = {(double)m[0], (double)m[1], (double)m[2]};
        block_max_to<3>(md, f.stats, reinterpret_cast<double*>(smem_raw));
      }
      if (flags & kLineMaxOnly) continue;
      // the fifth (real) product v w first, alone in region 0, while only u, v, w are live;
      // then the two packed pairs together
      cx<T> z2[1][E];
#pragma unroll
      for (int i = 0; i < E; ++i) z2[0][i] = mkc(z[0][i].y * z[1][i].x, (T)0);
      fft_ct_multi<1, E, false, NT, Lay, G.line, ST>(z2, lds, tw_fwd_l, tk);
      store_single(z2[0], r0, 4, tk);
      const T third = (T)1 / (T)3;
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const T u = z[0][i].x, v = z[0][i].y, w = z[1][i].x;
        const T q = (u * u + v * v + w * w) * third;
        z[0][i] = mkc(u * u - q, v * v - q);
        z[1][i] = mkc(u * v, u * w);
      }
      fft_ct_multi<2, E, false, NT, Lay, G.line, ST>(z, lds, tw_fwd_l, tk);
      store_pair(z[0], r0, 0, 1, tk);
      store_pair(z[1], r0, 2, 3, tk);
    } else {
      fft_ct_multi<3, E, true, NT, Lay, G.line, ST>(z, lds, tw_rev_l, tk);
      if (flags & kLineWantMax) {
        float m[3] = {0.f, 0.f, 0.f};  // |u_i| maxima (exact in fp32)
#pragma unroll
        for (int i = 0; i < E; ++i) {
          m[0] = fmaxf(m[0], abs_or_inf(z[0][i].x));
          m[1] = fmaxf(m[1], abs_or_inf(z[0][i].y));
          m[2] = fmaxf(m[2], abs_or_inf(z[1][i].x));
        }
        double md[3] = {(double)m[0], (double)m[1], (double)m[2]};
        block_max_to<3>(md, f.stats, reinterpret_cast<double*>(smem_raw));
      }
      if (flags & kLineMaxOnly) continue;
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const T u = z[0][i].x, v = z[0][i].y, w = z[1][i].x;
        const T ox = z[1][i].y, oy = z[2][i].x, oz = z[2][i].y;
        z[0][i] = mkc(v * oz - w * oy, w * ox - u * oz);
        z[1][i] = mkc(u * oy - v * ox, (T)0);
      }
      // the next tile's three pairs are fetched once the inputs are consumed
      if (k + 1 < TPW) fetch(pf, r0 + RT, tk);
      fft_ct_multi<2, E, false, NT, Lay, G.line, ST>(z, lds, tw_fwd_l, tk);
      store_pair(z[0], r0, 0, 1, tk);
      store_single(z[1], r0, 2, tk);
    }
  }
}

// 8-byte load of a work-field element, non-temporal where the grid's policy says so: the y and
// x passes read each element once (384^3: forward y pass 405 -> 356 us, x pass 748 -> 724 us,
// step 8.65 -> 8.43 ms; their stores stay ordinary, as the next kernel reads them).
typedef float psx_f2 __attribute__((ext_vector_type(2)));
template <int N>
__device__ __forceinline__ cx<float> load_w(const cx<float>* p) {
  if constexpr (kc3_pol(N) == kNt) {
    const psx_f2 v = __builtin_nontemporal_load(reinterpret_cast<const psx_f2*>(p));
    return mkc(v.x, v.y);
  } else {
    return *p;
  }
}

// --------------------------------------------------------------------------- y passes
// Tile: fixed x (group), columns j0..j0+T, rows x N + y.  Forward: split packed pairs into the
// two real fields' spectra, transform along y.  Inverse: transform along y, pack pairs.
template <int N, bool ROT, bool INV>
__global__ void __launch_bounds__(G3<N>::g.yNT, 1) kc_ypass3d(Flow<float> f) {
  if (flow_inactive(f.prm, f.nh)) return;
  using T = float;
  constexpr Geo3D G = G3<N>::g;
  constexpr int E = G.yE, NT = G.yNT, CT = G.T, L = N, LDW = G.ldw, NJT = LDW / CT;
  constexpr long long FS = G3<N>::FS;
  using Lay = ColsCT<CT>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int x = blockIdx.x / NJT;
  const int j0 = (blockIdx.x - x * NJT) * CT;
  const long long rb = (long long)x * N;
  constexpr int R0 = ct_radix(G.yp, INV, 0);
  constexpr int RL = ct_radix(G.yp, INV, G.yp.np - 1);
  // packed pairs and single fields of this direction, transformed group by group (a pair's two
  // transforms together in two LDS regions); the next group's loads are issued before the
  // current group's transforms, so they overlap.  The pass twiddles are read from an LDS copy:
  // vector-memory counters retire in order, and a twiddle load from global memory issued after
  // the next group's loads would wait for them.
  constexpr int NPAIR = INV ? (ROT ? 3 : 1) : (ROT ? 1 : 2);
  constexpr int SINGLE = INV ? (ROT ? -1 : 2) : (ROT ? 2 : 4);
  constexpr int NG = NPAIR + (SINGLE >= 0 ? 1 : 0);
  constexpr int TW = ct_twsize(G.yp, INV);
  cx<T>* twl = lds + 2 * L * CT;
  {
    const cx<T>* tw_o = f.twp + f.py.twoff[INV ? 1 : 0][0];
    for (int i = tid; i < TW; i += NT) twl[i] = tw_o[i];
  }
  __syncthreads();

  // raw loads of group g (pair fields 2g, 2g + 1, or the single field)
  auto load_group = [&](auto gc, cx<T>(&g)[2][E]) {
    constexpr int gi = decltype(gc)::value;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      int e, t;
      ct_elem<R0, L, NT, Lay>(i, tid, e, t);
      const long long a = (rb + e) * LDW + j0 + t;
      if constexpr (gi < NPAIR) {
        g[0][i] = load_w<N>(f.W + 2 * gi * FS + a);
        g[1][i] = load_w<N>(f.W + (2 * gi + 1) * FS + a);
      } else {
        g[0][i] = load_w<N>(f.W + SINGLE * FS + a);
      }
    }
  };
  auto run_group = [&](auto gc, cx<T>(&g)[2][E]) {
    constexpr int gi = decltype(gc)::value;
    if constexpr (gi < NPAIR) {
      constexpr int fa = 2 * gi, fb = fa + 1;
      if constexpr (!INV) {
#pragma unroll
        for (int i = 0; i < E; ++i) {
          int e, t;
          ct_elem<R0, L, NT, Lay>(i, tid, e, t);
          const cx<T> z = g[0][i], zm = g[1][i];
          z_split3(z, j0 + t == 0 ? z : zm, g[0][i], g[1][i]);
        }
      }
      fft_ct_multi<2, E, INV, NT, Lay, G.yp, L * CT>(g, lds, twl, tid);
#pragma unroll
      for (int i = 0; i < E; ++i) {
        int e, t;
        ct_elem<RL, L, NT, Lay>(i, tid, e, t);
        const long long a = (rb + e) * LDW + j0 + t;
        if constexpr (INV) {
          cx<T> z, zm;
          z_pack3(j0 + t, g[0][i], g[1][i], z, zm);
          f.W[fa * FS + a] = z;
          f.W[fb * FS + a] = zm;
        } else {
          f.W[fa * FS + a] = g[0][i];
          f.W[fb * FS + a] = g[1][i];
        }
      }
    } else {
      fft_ct_multi<1, E, INV, NT, Lay, G.yp, L * CT>(g, lds, twl, tid);
#pragma unroll
      for (int i = 0; i < E; ++i) {
        int e, t;
        ct_elem<RL, L, NT, Lay>(i, tid, e, t);
        f.W[SINGLE * FS + (rb + e) * LDW + j0 + t] = g[0][i];
      }
    }
  };
  cx<T> cur[2][E];
  load_group(std::integral_constant<int, 0>{}, cur);
  static_for<0, NG>([&](auto gc) {
    constexpr int gi = decltype(gc)::value;
    cx<T> nxt[2][E];
    if constexpr (gi + 1 < NG) load_group(std::integral_constant<int, gi + 1>{}, nxt);
    run_group(gc, cur);
    if constexpr (gi + 1 < NG) {
#pragma unroll
      for (int i = 0; i < E; ++i) {
        cur[0][i] = nxt[0][i];
        cur[1][i] = nxt[1][i];
      }
    }
  });
}

// ------------------------------------------------------------------ x pass with physics
// Grid sizes whose x pass keeps its third accumulated component in LDS (see kc_spec3d).
constexpr bool kc_spec3d_a2_lds(int n) { return n == 384; }

template <int N, bool ROT, int stage>
__global__ void __launch_bounds__(G3<N>::g.xNT, 1) kc_spec3d(Flow<float> f) {
  if (flow_inactive(f.prm, f.nh)) return;
  using T = float;
  constexpr Geo3D G = G3<N>::g;
  constexpr int E = G.xE, NT = G.xNT, CT = G.T, L = N, JM = G.jm, LDW = G.ldw, SP = G.sp;
  constexpr int NJT = LDW / CT, NH = N / 2 + 1;
  constexpr long long FS = G3<N>::FS;
  constexpr long long SST = G3<N>::NR * SP;  // complex elements per state component
  using Lay = ColsCT<CT>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int iy = blockIdx.x / NJT;
  const int j0 = (blockIdx.x - iy * NJT) * CT;
  const int ky = iy < N / 2 ? iy : iy - N;
  const int ay = ky < 0 ? -ky : ky;
  const cx<T>* tw_fwd = f.twp + f.pcol.twoff[0][0];
  const cx<T>* tw_rev = f.twp + f.pcol.twoff[1][0];
  constexpr int R0 = ct_radix(G.xp, false, 0);
  constexpr int RP = ct_radix(G.xp, false, G.xp.np - 1);
  constexpr int RS = ct_radix(G.xp, true, G.xp.np - 1);
  const BufC32T<kc3_pol(N)> ub(f.u0, (unsigned long long)3 * SST * 8);
  const BufC32T<kc3_pol(N)> ab(f.acc, (unsigned long long)3 * SST * 8);
  // At 384 the third component lives in LDS slots of its own (thread-private, after the
  // transform buffer): in registers beside the other two, the transform and the state it
  // spilled to scratch.
  constexpr bool kA2Lds = kc_spec3d_a2_lds(N);
  cx<T>* a2l = lds + 2 * L * CT;
  // pass twiddles of both directions, copied into LDS after the other regions: read from
  // global memory inside the transforms, their latency sat on the critical path of every pass
  // (384^3: 710 -> 658 us)
  constexpr int TWF = ct_twsize(G.xp, false), TWR = ct_twsize(G.xp, true);
  cx<T>* twl = a2l + (kA2Lds ? E * NT : 0);
  for (int i = tid; i < TWF + TWR; i += NT) twl[i] = i < TWF ? tw_fwd[i] : tw_rev[i - TWF];
  // and the FP64 integrating-factor tables (2 (n/2 + 1) entries), read after the transforms
  // (384^3: 659 -> 590 us)
  double* ifl = reinterpret_cast<double*>(twl + TWF + TWR);
  for (int i = tid; i < 2 * NH; i += NT) ifl[i] = f.prm[i];
  __syncthreads();

  cx<T> a0[E], a1[E], a2reg[kA2Lds ? 1 : E];
  auto A2 = [&](int i) -> cx<T>& {
    if constexpr (kA2Lds) {
      return a2l[i * NT + tid];
    } else {
      return a2reg[i];
    }
  };
  if constexpr (stage != kStagePre) {
    constexpr int nf = ROT ? 3 : 5;
#pragma unroll
    for (int i = 0; i < E; ++i) a0[i] = a1[i] = A2(i) = czero<T>();
    // field c's x transform, added into the accumulated components
    auto load_x = [&](int c, cx<T>(&v)[E]) {
#pragma unroll
      for (int i = 0; i < E; ++i) {
        int e, t;
        ct_elem<R0, L, NT, Lay>(i, tid, e, t);
        v[i] = load_w<N>(f.W + c * FS + ((long long)e * N + iy) * LDW + j0 + t);
      }
    };
    auto accum = [&](auto cc, const cx<T>(&v)[E]) {
      constexpr int c = decltype(cc)::value;
#pragma unroll
      for (int i = 0; i < E; ++i) {
        if constexpr (ROT) {
          if constexpr (c == 0) a0[i] = v[i];
          else if constexpr (c == 1) a1[i] = v[i];
          else A2(i) = v[i];
        } else {
          int e, t;
          ct_elem<RP, L, NT, Lay>(i, tid, e, t);
          const T kx = (T)(e < N / 2 ? e : e - N), fy = (T)ky, kz = (T)(j0 + t);
          const cx<T> mt = mul_mi(v[i]);  // -i T_c;  N_i = -i k_j T_ij, T33 = -T11 - T22
          if constexpr (c == 0) {
            a0[i] = cadd(a0[i], cscl(kx, mt));
            A2(i) = csub(A2(i), cscl(kz, mt));
          } else if constexpr (c == 1) {
            a1[i] = cadd(a1[i], cscl(fy, mt));
            A2(i) = csub(A2(i), cscl(kz, mt));
          } else if constexpr (c == 2) {
            a0[i] = cadd(a0[i], cscl(fy, mt));
            a1[i] = cadd(a1[i], cscl(kx, mt));
          } else if constexpr (c == 3) {
            a0[i] = cadd(a0[i], cscl(kz, mt));
            A2(i) = cadd(A2(i), cscl(kx, mt));
          } else {
            a1[i] = cadd(a1[i], cscl(kz, mt));
            A2(i) = cadd(A2(i), cscl(fy, mt));
          }
        }
      }
    };
    // the fields' transforms two at a time (fft_ct_multi), the odd one alone
    static_for<0, nf / 2>([&](auto pc) {
      constexpr int c0 = 2 * decltype(pc)::value;
      cx<T> vv[2][E];
      load_x(c0, vv[0]);
      load_x(c0 + 1, vv[1]);
      fft_ct_multi<2, E, false, NT, Lay, G.xp, L * CT>(vv, lds, twl, tid);
      accum(std::integral_constant<int, c0>{}, vv[0]);
      accum(std::integral_constant<int, c0 + 1>{}, vv[1]);
    });
    if constexpr (nf % 2) {
      cx<T> v[E];
      load_x(nf - 1, v);
      fft_ct<E, false, NT, Lay, G.xp>(v, lds, twl, tid);
      accum(std::integral_constant<int, nf - 1>{}, v);
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
      const int kx = e < N / 2 ? e : e - N;
      const int ax = kx < 0 ? -kx : kx;
      const int kz = j0 + t;
      // state offset of this mode (out of range outside the 2/3 mask), formed here rather than
      // before the transforms, where its E registers were held across them and spilled
      const bool in_mask = 3 * ax < N && 3 * ay < N && kz < JM;
      const unsigned soffi = (unsigned)(((long long)e * N + iy) * SP + kz);
      auto so = [&](int c) -> unsigned {
        return in_mask ? (soffi + (unsigned)(c * SST)) * 8u : kOob;
      };
      const cx<T> u0a = ub.load(so(0)), u0b = ub.load(so(1)), u0c = ub.load(so(2));
      cx<T> ua = u0a, ubv = u0b, uc = u0c;
      if constexpr (stage == kStagePre) {
        const int k2 = kx * kx + ky * ky + kz * kz;
        const T ik2 = k2 > 0 ? __builtin_amdgcn_rcpf((T)k2) : (T)0;
        leray<T>((T)kx, (T)ky, (T)kz, ik2, ua, ubv, uc);
        ub.store(so(0), ua);
        ub.store(so(1), ubv);
        ub.store(so(2), uc);
      }
      if constexpr (stage != kStagePre) {
        const int k2 = kx * kx + ky * ky + kz * kz;
        const T fx = (T)kx, fy = (T)ky, fz = (T)kz;
        const T ik2 = k2 > 0 ? __builtin_amdgcn_rcpf((T)k2) : (T)0;
        if (f.src) {  // uniform: an additive source (state layout, zero outside the mask)
          const BufC32T<kc3_pol(N)> sb(const_cast<cx<T>*>(f.src), (unsigned long long)3 * SST * 8);
          a0[i] = cadd(a0[i], sb.load(so(0)));
          a1[i] = cadd(a1[i], sb.load(so(1)));
          A2(i) = cadd(A2(i), sb.load(so(2)));
        }
        const cx<T> d = cscl(ik2, cadd(cadd(cscl(fx, a0[i]), cscl(fy, a1[i])), cscl(fz, A2(i))));
        const cx<T> n0 = csub(a0[i], cscl(fx, d));
        const cx<T> n1 = csub(a1[i], cscl(fy, d));
        const cx<T> n2 = csub(A2(i), cscl(fz, d));
        double E1, Eh;
        if_factors(ifl, NH, s, ax, ay, kz, k2, E1, Eh);
        cx<T> c0 = czero<T>(), c1 = czero<T>(), c2 = czero<T>();
        if constexpr (stage >= 1) {
          c0 = ab.load(so(0));
          c1 = ab.load(so(1));
          c2 = ab.load(so(2));
        }
        ua = rk_value<stage, T>(dt, (T)(1.0 - E1), (T)(1.0 - Eh), n0, u0a, c0);
        ubv = rk_value<stage, T>(dt, (T)(1.0 - E1), (T)(1.0 - Eh), n1, u0b, c1);
        uc = rk_value<stage, T>(dt, (T)(1.0 - E1), (T)(1.0 - Eh), n2, u0c, c2);
        if constexpr (stage < 3) {
          ab.store(so(0), c0);
          ab.store(so(1), c1);
          ab.store(so(2), c2);
        } else {
          leray<T>(fx, fy, fz, ik2, ua, ubv, uc);
          ub.store(so(0), ua);
          ub.store(so(1), ubv);
          ub.store(so(2), uc);
        }
      }
      if (!in_mask) {  // 2/3 mask: no nonlinear term outside it
        ua = ubv = uc = czero<T>();
      }
      a0[i] = cscl(scale, ua);
      a1[i] = cscl(scale, ubv);
      A2(i) = cscl(scale, uc);
    }
  }
  constexpr int nin = ROT ? 6 : 3;
  // next inputs: the velocity components (and, rotational, omega = i k x u)
  auto make_in = [&](auto cc, cx<T>(&v)[E]) {
    constexpr int c = decltype(cc)::value;
#pragma unroll
    for (int i = 0; i < E; ++i) {
      if constexpr (c == 0) {
        v[i] = a0[i];
      } else if constexpr (c == 1) {
        v[i] = a1[i];
      } else if constexpr (c == 2) {
        v[i] = A2(i);
      } else {
        int e, t;
        ct_elem<RP, L, NT, Lay>(i, tid, e, t);
        const T fx = (T)(e < N / 2 ? e : e - N), fy = (T)ky, fz = (T)(j0 + t);
        // omega = i k x u
        if constexpr (c == 3) v[i] = mul_i(csub(cscl(fy, A2(i)), cscl(fz, a1[i])));
        else if constexpr (c == 4) v[i] = mul_i(csub(cscl(fz, a0[i]), cscl(fx, A2(i))));
        else v[i] = mul_i(csub(cscl(fx, a1[i]), cscl(fy, a0[i])));
      }
    }
  };
  auto store_x = [&](int c, const cx<T>(&v)[E]) {
#pragma unroll
    for (int i = 0; i < E; ++i) {
      int e, t;
      ct_elem<RS, L, NT, Lay>(i, tid, e, t);
      f.W[c * FS + ((long long)e * N + iy) * LDW + j0 + t] = v[i];
    }
  };
  static_for<0, nin / 2>([&](auto pc) {
    constexpr int c0 = 2 * decltype(pc)::value;
    cx<T> vv[2][E];
    make_in(std::integral_constant<int, c0>{}, vv[0]);
    make_in(std::integral_constant<int, c0 + 1>{}, vv[1]);
    fft_ct_multi<2, E, true, NT, Lay, G.xp, L * CT>(vv, lds, twl + TWF, tid);
    store_x(c0, vv[0]);
    store_x(c0 + 1, vv[1]);
  });
  if constexpr (nin % 2) {
    cx<T> v[E];
    make_in(std::integral_constant<int, nin - 1>{}, v);
    fft_ct<E, true, NT, Lay, G.xp>(v, lds, twl + TWF, tid);
    store_x(nin - 1, v);
  }
}

// ------------------------------------------------------------------------------ launchers
bool ct3d_supported(int n) {
  switch (n) {
#define PSX_X(NV) case NV:
    PSX_CT3D_SIZES(PSX_X)
#undef PSX_X
    return true;
    default:
      return false;
  }
}

template <int N, bool ROT>
static hipError_t ct3_launch(const Flow<float>& f, int which, int arg, hipStream_t st) {
  constexpr Geo3D G = G3<N>::g;
  constexpr int NJT = G.ldw / G.T;
  switch (which) {
    case 0: {
      auto fn = kc_line3d<N, ROT>;
      static const bool attr_ = set_lds_max((const void*)fn);
      (void)attr_;
      // transform regions (three rotational, two Basdevant) and the two twiddle tables
      size_t lds = ((size_t)(ROT ? 3 : 2) * G.lineT * RowsCT<N, G.lineT>::LP +
                    ct_twsize(G.line, false) + ct_twsize(G.line, true)) * 8;
      const size_t red = (size_t)3 * G.lineNT * sizeof(double);
      if (lds < red) lds = red;
      static_assert((N * N / G.lineT) % line3d_tpw<N, ROT>() == 0, "tiles per work-group");
      hipLaunchKernelGGL(fn, dim3(N * N / G.lineT / line3d_tpw<N, ROT>()), dim3(G.lineNT), lds, st, f,
                         arg);
      break;
    }
    case 1:
    case 3: {
      // two sequence regions and the pass twiddles of the direction
      const size_t lds = ((size_t)2 * N * G.T + ct_twsize(G.yp, which == 3)) * 8;
      static const bool attr_ = set_lds_max((const void*)kc_ypass3d<N, ROT, false>) &&
                                set_lds_max((const void*)kc_ypass3d<N, ROT, true>);
      (void)attr_;
      if (which == 1)
        hipLaunchKernelGGL((kc_ypass3d<N, ROT, false>), dim3(N * NJT), dim3(G.yNT), lds, st, f);
      else
        hipLaunchKernelGGL((kc_ypass3d<N, ROT, true>), dim3(N * NJT), dim3(G.yNT), lds, st, f);
      break;
    }
    case 2: {
      // two transform regions, the third component's slots where it lives in LDS, and the
      // pass twiddles of both directions
      const size_t lds = (size_t)2 * N * G.T * 8 +
                         (kc_spec3d_a2_lds(N) ? (size_t)G.xE * G.xNT * 8 : 0) +
                         (size_t)(ct_twsize(G.xp, false) + ct_twsize(G.xp, true)) * 8 +
                         (size_t)2 * (N / 2 + 1) * sizeof(double);
      const dim3 grid(N * NJT), block(G.xNT);
      static const bool attr_ = set_lds_max((const void*)kc_spec3d<N, ROT, kStagePre>) &&
                                set_lds_max((const void*)kc_spec3d<N, ROT, 0>) &&
                                set_lds_max((const void*)kc_spec3d<N, ROT, 1>) &&
                                set_lds_max((const void*)kc_spec3d<N, ROT, 2>) &&
                                set_lds_max((const void*)kc_spec3d<N, ROT, 3>);
      (void)attr_;
      switch (arg) {
        case kStagePre: hipLaunchKernelGGL((kc_spec3d<N, ROT, kStagePre>), grid, block, lds, st, f); break;
        case 0: hipLaunchKernelGGL((kc_spec3d<N, ROT, 0>), grid, block, lds, st, f); break;
        case 1: hipLaunchKernelGGL((kc_spec3d<N, ROT, 1>), grid, block, lds, st, f); break;
        case 2: hipLaunchKernelGGL((kc_spec3d<N, ROT, 2>), grid, block, lds, st, f); break;
        case 3: hipLaunchKernelGGL((kc_spec3d<N, ROT, 3>), grid, block, lds, st, f); break;
        default: return hipErrorInvalidValue;
      }
      break;
    }
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// which: 0 z lines (arg = flags), 1 y forward, 2 x pass with physics (arg = stage),
// 3 y inverse.
hipError_t launch_ct3d(const Flow<float>& f, int which, int arg, hipStream_t st) {
  switch (f.n) {
#define PSX_X(NV)                                                            \
  case NV:                                                                   \
    return f.rot ? ct3_launch<NV, true>(f, which, arg, st)                   \
                 : ct3_launch<NV, false>(f, which, arg, st);
    PSX_CT3D_SIZES(PSX_X)
#undef PSX_X
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace psx
