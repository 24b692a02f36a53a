// Compile-time planned variant of the FFT engine in psx_fft.h.
//
// Same pass formulation, register order and twiddle-table layout, but the plan (length,
// radices), the work-group size and the tile shape are template parameters, so every index
// in a pass folds to shifts, masks and constants; only the thread index is a run-time value.
// The kernels built on it are specialised per grid size (psx_geo.h, PSX_CT2D_SIZES).
#pragma once

#include "psx_fft.h"
#include "psx_geo.h"

namespace psx {

// Row tiles: T_ sequences of length L_ stored contiguously in LDS, one pad element per 16.
// A single 8192-element row (passes of radix 32, 16, 16 and the reverse) has one pad element
// per 32 instead: the radix-32 pass writes elements 32 apart, which one pad per 16 puts two
// lanes of every 32 on the same bank pair (64 four-byte banks, 32 lanes of 8-byte accesses
// at a time); measured 2 % faster on that kernel, where one pad per 8 is 3 % slower.  The
// 384-point rows of the 3D line kernel (radices 4 4 2 12) also use one per 32: 1.5 % faster.
template <int L_, int T_>
struct RowsCT {
  static constexpr bool kPad32 = (L_ == 8192 && T_ == 1) || L_ == 384;
  static constexpr int L = L_, T = T_;
  static constexpr int LP = kPad32 ? L_ + (L_ >> 5) + 1 : ct_rows_pitch(L_);  // <= the latter
  static constexpr int kShift = kPad32 ? 5 : 4;
  static constexpr PSX_HD int addr(int e, int t) { return t * LP + e + (e >> kShift); }
  // True when, for every butterfly of a radix-R pass with stride Ns, the addresses of its R
  // Stockham outputs (elements base + q Ns, base = (j - k) R + k, k = j % Ns) lie at
  // addr(base) + wdelta(q Ns): the writes then take one address computation per butterfly and
  // constant offsets instead of a shift and add per element.
  template <int R, int Ns>
  static constexpr bool write_folds() {
    for (int j = 0; j < L / R; ++j) {
      const int k = j % Ns, base = (j - k) * R + k;
      for (int q = 1; q < R; ++q)
        if (addr(base + q * Ns, 0) - addr(base, 0) != wdelta(q * Ns)) return false;
    }
    return true;
  }
  static constexpr PSX_HD int wdelta(int d) { return d + (d >> kShift); }
  // Address in the exchange after pass X of the forward (INV = false) or reversed plan, as the
  // multi-sequence transforms use it.  Four 384-point rows per tile (plan 4 4 2 12, 12
  // elements per thread, 32-lane half-waves within one row): three of the six exchanges take
  // an XOR swizzle of the element index within aligned 32- or 64-element blocks in place of
  // the pad, so that the Stockham writes of the pass (a thread's outputs Ns elements apart)
  // meet no bank conflicts while the next pass's reads (32 consecutive elements per
  // half-wave) stay conflict-free.  The other three keep the pad: the swizzles that halve
  // their remaining two-way write conflicts cost more address arithmetic than they save.
  static constexpr bool kSwz384 = L_ == 384 && T_ == 4;
  template <bool INV, int X>
  static constexpr PSX_HD int xaddr(int e, int t) {
    if constexpr (kSwz384 && INV && X == 0) return t * LP + (e ^ ((e >> 5) & 3));
    else if constexpr (kSwz384 && !INV && X == 1) return t * LP + (e ^ ((e >> 3) & 15));
    else if constexpr (kSwz384 && !INV && X == 2) return t * LP + (e ^ ((e >> 1) & 16));
    else return addr(e, t);
  }
  template <int Lr>
  static constexpr PSX_HD void decomp(int g, int& j, int& t) {
    t = g / Lr;
    j = g - t * Lr;
  }
};

// True when the addresses of every exchange X < NX of a row layout map the L elements of row 0
// one-to-one into [0, LP) (a swizzle must stay a permutation within the padded row).
template <class Lay, bool INV, int NX>
constexpr bool ct_xaddr_ok() {
  if constexpr (NX == 0) {
    return true;
  } else {
    bool seen[Lay::LP] = {};
    for (int e = 0; e < Lay::L; ++e) {
      const int a = Lay::template xaddr<INV, NX - 1>(e, 0);
      if (a < 0 || a >= Lay::LP || seen[a]) return false;
      seen[a] = true;
    }
    return ct_xaddr_ok<Lay, INV, NX - 1>();
  }
}

// Column tiles: T_ columns (power of two) side by side; lanes walk the columns.
template <int T_>
struct ColsCT {
  static constexpr int T = T_;
  static constexpr PSX_HD int addr(int e, int t) { return e * T + t; }
  template <bool INV, int X>
  static constexpr PSX_HD int xaddr(int e, int t) { return addr(e, t); }
  template <int R, int Ns>
  static constexpr bool write_folds() { return true; }
  static constexpr PSX_HD int wdelta(int d) { return d * T; }
  template <int Lr>
  static constexpr PSX_HD void decomp(int g, int& j, int& t) {
    t = g % T;
    j = g / T;
  }
};

// Barrier of an LDS exchange.  WS: the exchange stays within one wave (a tile of at most 64
// threads inside a 64-thread group): the workgroup-scope fences of a block barrier without the
// s_barrier, so the other waves of the work-group (other tiles) are not waited for.
template <bool WS>
__device__ __forceinline__ void ct_sync() {
  if constexpr (WS) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  } else {
    __syncthreads();
  }
}

template <int E, int R, bool INV, int L, int Ns, int NT, class Lay, bool WS = false, typename T>
__device__ __forceinline__ void ct_pass(cx<T> (&v)[E], cx<T>* lds, bool from_lds, bool to_lds,
                                        const cx<T>* __restrict__ tw, int tid) {
  static_assert(E % R == 0, "radix must divide the per-thread element count");
  constexpr int NB = E / R;
  constexpr int Lr = L / R;
  if (from_lds) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      int j, t;
      Lay::template decomp<Lr>(tid + b * NT, j, t);
#pragma unroll
      for (int m = 0; m < R; ++m) v[b * R + m] = lds[Lay::addr(j + m * Lr, t)];
    }
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    int j, t;
    Lay::template decomp<Lr>(tid + b * NT, j, t);
    if constexpr (Ns > 1) {
      const int k = j % Ns;
#pragma unroll
      for (int m = 1; m < R; ++m) {
        const cx<T> w = tw[(m - 1) * Ns + k];
        v[b * R + m] = INV ? cmulc(v[b * R + m], w) : cmul(v[b * R + m], w);
      }
    }
    dft<R, INV>(&v[b * R]);
  }
  if (to_lds) {
    ct_sync<WS>();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      int j, t;
      Lay::template decomp<Lr>(tid + b * NT, j, t);
      const int k = j % Ns;
      const int base = (j - k) * R + k;
      if constexpr (Lay::template write_folds<R, Ns>()) {
        const int a0 = Lay::addr(base, t);
#pragma unroll
        for (int q = 0; q < R; ++q) lds[a0 + Lay::wdelta(q * Ns)] = v[b * R + q];
      } else {
#pragma unroll
        for (int q = 0; q < R; ++q) lds[Lay::addr(base + q * Ns, t)] = v[b * R + q];
      }
    }
    ct_sync<WS>();
  }
}

// Whole transform; tw_order points at the start of this plan's forward (INV = false) or
// reversed-order (INV = true) twiddle table.
template <int E, bool INV, int NT, class Lay, CtPlan P, bool WS = false, typename T>
__device__ __forceinline__ void fft_ct(cx<T> (&v)[E], cx<T>* lds,
                                       const cx<T>* __restrict__ tw_order, int tid) {
  static_for<0, P.np>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int R = ct_radix(P, INV, q);
    constexpr int Ns = ct_ns(P, INV, q);
    constexpr int off = ct_twoff(P, INV, q);
    ct_pass<E, R, INV, P.L, Ns, NT, Lay, WS>(v, lds, q > 0, q < P.np - 1, tw_order + off, tid);
  });
}

// Radix-R pass whose twiddles were loaded beforehand into w[b][m - 1] (as ct_pass otherwise).
template <int E, int R, bool INV, int L, int Ns, int NT, class Lay, bool FROM, typename T>
__device__ __forceinline__ void ct_pass_w(cx<T> (&v)[E], cx<T>* lds,
                                          const cx<T> (&w)[E / R][R - 1], int tid) {
  constexpr int NB = E / R;
  constexpr int Lr = L / R;
  if constexpr (FROM) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      int j, t;
      Lay::template decomp<Lr>(tid + b * NT, j, t);
#pragma unroll
      for (int m = 0; m < R; ++m) v[b * R + m] = lds[Lay::addr(j + m * Lr, t)];
    }
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
#pragma unroll
    for (int m = 1; m < R; ++m)
      v[b * R + m] = INV ? cmulc(v[b * R + m], w[b][m - 1]) : cmul(v[b * R + m], w[b][m - 1]);
    dft<R, INV>(&v[b * R]);
  }
}

// As fft_ct_tw2 with QG = np - 1, the last pass's twiddles (from tw_far) loaded at the start of
// the pass before it, so that their latency overlaps that pass and its exchange.
template <int E, bool INV, int NT, class Lay, CtPlan P, typename T>
__device__ __forceinline__ void fft_ct_tw2_pre(cx<T> (&v)[E], cx<T>* lds,
                                               const cx<T>* __restrict__ tw_near,
                                               const cx<T>* __restrict__ tw_far, int tid) {
  constexpr int QL = P.np - 1;
  constexpr int RL = ct_radix(P, INV, QL);
  constexpr int NsL = ct_ns(P, INV, QL);
  constexpr int NBL = E / RL;
  static_assert(QL >= 1 && NsL > 1, "a last pass with twiddles");
  cx<T> wl[NBL][RL - 1];
  static_for<0, P.np>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int R = ct_radix(P, INV, q);
    constexpr int Ns = ct_ns(P, INV, q);
    constexpr int off = ct_twoff(P, INV, q);
    if constexpr (q == QL - 1) {
      const cx<T>* twl_g = tw_far + ct_twoff(P, INV, QL);
      if constexpr (sizeof(T) == 4) {
        // buffer loads: the table row (m - 1) NsL goes into the instruction's scalar offset,
        // where a pointer load needs a 64-bit address add per twiddle
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<cx<T>*>(twl_g), 0, (RL - 1) * NsL * 8, 0x00020000);
#pragma unroll
        for (int b = 0; b < NBL; ++b) {
          int j, t;
          Lay::template decomp<P.L / RL>(tid + b * NT, j, t);
          const int vo = (j % NsL) * 8;
#pragma unroll
          for (int m = 1; m < RL; ++m) {
            const int so = (m - 1) * NsL * 8;
            const unsigned re = __builtin_amdgcn_raw_buffer_load_b32(rs, vo, so, 0);
            const unsigned im = __builtin_amdgcn_raw_buffer_load_b32(rs, vo + 4, so, 0);
            wl[b][m - 1] = mkc(__builtin_bit_cast(T, re), __builtin_bit_cast(T, im));
          }
        }
      } else {
#pragma unroll
        for (int b = 0; b < NBL; ++b) {
          int j, t;
          Lay::template decomp<P.L / RL>(tid + b * NT, j, t);
          const int k = j % NsL;
#pragma unroll
          for (int m = 1; m < RL; ++m) wl[b][m - 1] = twl_g[(m - 1) * NsL + k];
        }
      }
    }
    if constexpr (q < QL)
      ct_pass<E, R, INV, P.L, Ns, NT, Lay>(v, lds, q > 0, true, tw_near + off, tid);
    else
      ct_pass_w<E, R, INV, P.L, Ns, NT, Lay, (q > 0)>(v, lds, wl, tid);
  });
}

// As fft_ct, with the twiddles of passes q < QG read from tw_near (a copy of the start of the
// order's table, e.g. in LDS) and those of the later passes from tw_far.
template <int E, bool INV, int NT, class Lay, CtPlan P, int QG, typename T>
__device__ __forceinline__ void fft_ct_tw2(cx<T> (&v)[E], cx<T>* lds,
                                           const cx<T>* __restrict__ tw_near,
                                           const cx<T>* __restrict__ tw_far, int tid) {
  static_for<0, P.np>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int R = ct_radix(P, INV, q);
    constexpr int Ns = ct_ns(P, INV, q);
    constexpr int off = ct_twoff(P, INV, q);
    if constexpr (q < QG)
      ct_pass<E, R, INV, P.L, Ns, NT, Lay>(v, lds, q > 0, q < P.np - 1, tw_near + off, tid);
    else
      ct_pass<E, R, INV, P.L, Ns, NT, Lay>(v, lds, q > 0, q < P.np - 1, tw_far + off, tid);
  });
}

// NS independent transforms of the same plan advanced pass by pass together: sequence s
// exchanges through its own LDS region at lds + s * STRIDE (elements), each pass loads a
// twiddle once for all NS sequences, and the NS exchanges of a pass share one pair of
// barriers.  Run one after another, each transform pays its own twiddle-load and barrier
// latency; together, the loads and LDS accesses of the NS sequences are in flight at once.
// v[0 .. NS-1] are transformed (NS <= NA).  Q is the pass's index in the plan: the pass reads
// the exchange after pass Q - 1 and writes the one after pass Q (the layout's xaddr).
template <int NS, int NA, int E, int R, bool INV, int L, int Ns, int NT, class Lay, int STRIDE,
          bool FROM, bool TO, bool WS = false, int Q = 0, typename T>
__device__ __forceinline__ void ct_pass_multi(cx<T> (&v)[NA][E], cx<T>* lds,
                                              const cx<T>* __restrict__ tw, int tid) {
  static_assert(E % R == 0, "radix must divide the per-thread element count");
  static_assert(NS >= 1 && NS <= NA, "sequence count");
  constexpr int NB = E / R;
  constexpr int Lr = L / R;
  if constexpr (FROM) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      int j, t;
      Lay::template decomp<Lr>(tid + b * NT, j, t);
#pragma unroll
      for (int s = 0; s < NS; ++s) {
#pragma unroll
        for (int m = 0; m < R; ++m)
          v[s][b * R + m] = lds[s * STRIDE + Lay::template xaddr<INV, Q - 1>(j + m * Lr, t)];
      }
    }
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    int j, t;
    Lay::template decomp<Lr>(tid + b * NT, j, t);
    if constexpr (Ns > 1) {
      const int k = j % Ns;
#pragma unroll
      for (int m = 1; m < R; ++m) {
        const cx<T> w = tw[(m - 1) * Ns + k];
#pragma unroll
        for (int s = 0; s < NS; ++s)
          v[s][b * R + m] = INV ? cmulc(v[s][b * R + m], w) : cmul(v[s][b * R + m], w);
      }
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) dft<R, INV>(&v[s][b * R]);
  }
  if constexpr (TO) {
    ct_sync<WS>();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      int j, t;
      Lay::template decomp<Lr>(tid + b * NT, j, t);
      const int k = j % Ns;
      const int base = (j - k) * R + k;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
#pragma unroll
        for (int q = 0; q < R; ++q)
          lds[s * STRIDE + Lay::template xaddr<INV, Q>(base + q * Ns, t)] = v[s][b * R + q];
      }
    }
    ct_sync<WS>();
  }
}

template <int NS, int E, bool INV, int NT, class Lay, CtPlan P, int STRIDE, bool WS = false,
          int NA, typename T>
__device__ __forceinline__ void fft_ct_multi(cx<T> (&v)[NA][E], cx<T>* lds,
                                             const cx<T>* __restrict__ tw_order, int tid) {
  static_for<0, P.np>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int R = ct_radix(P, INV, q);
    constexpr int Ns = ct_ns(P, INV, q);
    constexpr int off = ct_twoff(P, INV, q);
    ct_pass_multi<NS, NA, E, R, INV, P.L, Ns, NT, Lay, STRIDE, (q > 0), (q < P.np - 1), WS, q>(
        v, lds, tw_order + off, tid);
  });
}

// True when register slot i of a pass of radix R (NT threads, layout Lay) holds, in every
// thread, an element e with lo <= e <= hi: the line kernels use it to drop, at compile time,
// the loads and stores of slots that fall in the dealiased gap of the packed spectra.
template <int R, int L, int NT, class Lay>
constexpr bool ct_slot_within(int i, int lo, int hi) {
  const int b = i / R, m = i - b * R;
  constexpr int lr = L / R;
  for (int tid = 0; tid < NT; ++tid) {
    int j = 0, t = 0;
    Lay::template decomp<lr>(tid + b * NT, j, t);
    const int e = j + m * lr;
    if (e < lo || e > hi) return false;
  }
  return true;
}

// Register slot i of a pass of radix R holds element e of sequence t.
template <int R, int L, int NT, class Lay>
PSX_HD void ct_elem(int i, int tid, int& e, int& t) {
  const int b = i / R;
  const int m = i - b * R;
  int j;
  Lay::template decomp<L / R>(tid + b * NT, j, t);
  e = j + m * (L / R);
}

}  // namespace psx
