// Batched complex FFT engine for gfx950 work-groups (Stockham autosort, mixed radix).
//
// A work-group transforms T independent sequences of length L at once.  Each of its NT
// threads keeps E = L*T/NT elements in registers; a pass of radix R runs E/R butterflies
// per thread entirely in registers and exchanges data through LDS between passes, so a
// plan with P passes costs P-1 LDS round trips.  The first pass reads its inputs from
// registers (the caller loads them from HBM or produces them in place) and the last pass
// leaves its outputs in registers, which is what lets the stage kernels fuse pointwise
// physics between a forward and an inverse transform without touching LDS or HBM.
//
// Pass formulation (decimation in time, self-sorting): with Ns = product of the radices
// already applied, butterfly j reads x_m = a[j + m L/R], multiplies x_m by
// w_{Ns R}^{m (j mod Ns)}, takes an R-point DFT and writes y_q to
// a[(j - j mod Ns) R + (j mod Ns) + q Ns].  The inverse transform uses the reversed radix
// list, so the radix that ends a forward transform starts the inverse one: both see the
// same butterfly-to-register assignment and can be chained in registers.
//
// Twiddles: per-pass tables (see FftPlan::twoff) generated in FP64 on the host and rounded
// once to the working precision; the inverse uses their conjugates.  Constant twiddles inside
// the radix-8/16/32 butterflies are correctly rounded literals.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace psx {

#define PSX_HD __host__ __device__ __forceinline__

template <typename T>
struct cx {
  T x, y;
};

template <typename T> PSX_HD cx<T> mkc(T a, T b) { cx<T> r; r.x = a; r.y = b; return r; }
template <typename T> PSX_HD cx<T> cadd(cx<T> a, cx<T> b) { return mkc(a.x + b.x, a.y + b.y); }
template <typename T> PSX_HD cx<T> csub(cx<T> a, cx<T> b) { return mkc(a.x - b.x, a.y - b.y); }
template <typename T> PSX_HD cx<T> cscl(T s, cx<T> a) { return mkc(s * a.x, s * a.y); }
template <typename T> PSX_HD cx<T> cmul(cx<T> a, cx<T> b) {
  return mkc(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename T> PSX_HD cx<T> cmulc(cx<T> a, cx<T> b) {
  return mkc(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T> PSX_HD cx<T> conjg(cx<T> a) { return mkc(a.x, -a.y); }
template <typename T> PSX_HD cx<T> mul_i(cx<T> a) { return mkc(-a.y, a.x); }   // i a
template <typename T> PSX_HD cx<T> mul_mi(cx<T> a) { return mkc(a.y, -a.x); }  // -i a
template <typename T> PSX_HD cx<T> czero() { return mkc((T)0, (T)0); }

// exp(-2 pi i p / 32), p < 32, correctly rounded FP64.
static constexpr double kW32C[32] = {
    1.0, 0.9807852804032304, 0.9238795325112867, 0.8314696123025452, 0.7071067811865476,
    0.5555702330196022, 0.3826834323650898, 0.19509032201612828, 0.0, -0.19509032201612828,
    -0.3826834323650898, -0.5555702330196022, -0.7071067811865476, -0.8314696123025452,
    -0.9238795325112867, -0.9807852804032304, -1.0, -0.9807852804032304, -0.9238795325112867,
    -0.8314696123025452, -0.7071067811865476, -0.5555702330196022, -0.3826834323650898,
    -0.19509032201612828, 0.0, 0.19509032201612828, 0.3826834323650898, 0.5555702330196022,
    0.7071067811865476, 0.8314696123025452, 0.9238795325112867, 0.9807852804032304};
static constexpr double kW32S[32] = {
    0.0, -0.19509032201612828, -0.3826834323650898, -0.5555702330196022, -0.7071067811865476,
    -0.8314696123025452, -0.9238795325112867, -0.9807852804032304, -1.0, -0.9807852804032304,
    -0.9238795325112867, -0.8314696123025452, -0.7071067811865476, -0.5555702330196022,
    -0.3826834323650898, -0.19509032201612828, 0.0, 0.19509032201612828, 0.3826834323650898,
    0.5555702330196022, 0.7071067811865476, 0.8314696123025452, 0.9238795325112867,
    0.9807852804032304, 1.0, 0.9807852804032304, 0.9238795325112867, 0.8314696123025452,
    0.7071067811865476, 0.5555702330196022, 0.3826834323650898, 0.19509032201612828};

template <int I, int N, class F>
PSX_HD void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

// a * w_R^P with w_R = exp(-+ 2 pi i / R) (forward / inverse), R | 32, P compile-time.
template <int R, int P, bool INV, typename T>
PSX_HD cx<T> rot(cx<T> a) {
  constexpr int idx = ((P % R) * (32 / R)) % 32;
  if constexpr (idx == 0) {
    return a;
  } else if constexpr (idx == 8) {
    return INV ? mul_i(a) : mul_mi(a);
  } else if constexpr (idx == 16) {
    return mkc(-a.x, -a.y);
  } else if constexpr (idx == 24) {
    return INV ? mul_mi(a) : mul_i(a);
  } else {
    const T c = (T)kW32C[idx];
    const T s = INV ? (T)(-kW32S[idx]) : (T)kW32S[idx];
    return cmul(a, mkc(c, s));
  }
}

template <int R, bool INV, typename T> PSX_HD void dft(cx<T>* x);

template <bool INV, typename T>
PSX_HD void dft2(cx<T>* x) {
  cx<T> a = x[0];
  x[0] = cadd(a, x[1]);
  x[1] = csub(a, x[1]);
}

template <bool INV, typename T>
PSX_HD void dft4(cx<T>* x) {
  cx<T> a = cadd(x[0], x[2]);
  cx<T> b = csub(x[0], x[2]);
  cx<T> c = cadd(x[1], x[3]);
  cx<T> d = csub(x[1], x[3]);
  d = INV ? mul_i(d) : mul_mi(d);
  x[0] = cadd(a, c);
  x[2] = csub(a, c);
  x[1] = cadd(b, d);
  x[3] = csub(b, d);
}

template <bool INV, typename T>
PSX_HD void dft3(cx<T>* x) {
  const T c = (T)-0.5;
  const T s = INV ? (T)0.86602540378443865 : (T)-0.86602540378443865;  // sin(-+2pi/3)
  cx<T> t = cadd(x[1], x[2]);
  cx<T> d = csub(x[1], x[2]);
  cx<T> m = cadd(x[0], cscl(c, t));
  cx<T> r = mul_i(cscl(s, d));  // i s (x1 - x2)
  x[0] = cadd(x[0], t);
  x[1] = cadd(m, r);
  x[2] = csub(m, r);
}

template <bool INV, typename T>
PSX_HD void dft5(cx<T>* x) {
  const T c1 = (T)0.30901699437494742, c2 = (T)-0.80901699437494742;
  const T s1 = (T)0.95105651629515357, s2 = (T)0.58778525229247313;
  cx<T> t1 = cadd(x[1], x[4]), t3 = csub(x[1], x[4]);
  cx<T> t2 = cadd(x[2], x[3]), t4 = csub(x[2], x[3]);
  cx<T> a1 = cadd(x[0], cadd(cscl(c1, t1), cscl(c2, t2)));
  cx<T> a2 = cadd(x[0], cadd(cscl(c2, t1), cscl(c1, t2)));
  cx<T> b1 = cadd(cscl(s1, t3), cscl(s2, t4));
  cx<T> b2 = csub(cscl(s2, t3), cscl(s1, t4));
  // forward: y1 = a1 - i b1, y4 = a1 + i b1, y2 = a2 - i b2, y3 = a2 + i b2
  b1 = INV ? mul_i(b1) : mul_mi(b1);
  b2 = INV ? mul_i(b2) : mul_mi(b2);
  x[0] = cadd(x[0], cadd(t1, t2));
  x[1] = cadd(a1, b1);
  x[4] = csub(a1, b1);
  x[2] = cadd(a2, b2);
  x[3] = csub(a2, b2);
}

// 12 = 3 * 4 by the prime-factor (Good-Thomas) mapping, gcd(3, 4) = 1: 3-point DFTs over n1
// for each n2 of input n = (4 n1 + 3 n2) mod 12, then 4-point DFTs over n2 for each k1 into
// output k = (4 k1 + 9 k2) mod 12 (4 = 4 (4^-1 mod 3), 9 = 3 (3^-1 mod 4)); no twiddles
// between the two stages.
template <bool INV, typename T>
PSX_HD void dft12(cx<T>* x) {
  cx<T> a[3][4];
  static_for<0, 4>([&](auto n2c) {
    constexpr int n2 = decltype(n2c)::value;
    cx<T> t[3] = {x[(3 * n2) % 12], x[(4 + 3 * n2) % 12], x[(8 + 3 * n2) % 12]};
    dft3<INV>(t);
    a[0][n2] = t[0];
    a[1][n2] = t[1];
    a[2][n2] = t[2];
  });
  static_for<0, 3>([&](auto k1c) {
    constexpr int k1 = decltype(k1c)::value;
    cx<T> t[4] = {a[k1][0], a[k1][1], a[k1][2], a[k1][3]};
    dft4<INV>(t);
    static_for<0, 4>([&](auto k2c) {
      constexpr int k2 = decltype(k2c)::value;
      x[(4 * k1 + 9 * k2) % 12] = t[k2];
    });
  });
}

// R = R1 * R2: inner R2-point DFTs over m1 (input m = m0 + R1 m1), twiddle w_R^{m0 q0},
// outer R1-point DFTs over m0; output q = q0 + R2 q1.
template <int R1, int R2, bool INV, typename T>
PSX_HD void dft_split(cx<T>* x) {
  constexpr int R = R1 * R2;
  cx<T> b[R];
  static_for<0, R1>([&](auto m0c) {
    constexpr int m0 = decltype(m0c)::value;
    cx<T> t[R2];
    static_for<0, R2>([&](auto m1c) { t[m1c.value] = x[m0 + R1 * m1c.value]; });
    dft<R2, INV>(t);
    static_for<0, R2>([&](auto q0c) {
      constexpr int q0 = decltype(q0c)::value;
      b[m0 * R2 + q0] = rot<R, m0 * q0, INV>(t[q0]);
    });
  });
  static_for<0, R2>([&](auto q0c) {
    constexpr int q0 = decltype(q0c)::value;
    cx<T> t[R1];
    static_for<0, R1>([&](auto m0c) { t[m0c.value] = b[m0c.value * R2 + q0]; });
    dft<R1, INV>(t);
    static_for<0, R1>([&](auto q1c) { x[q0 + R2 * q1c.value] = t[q1c.value]; });
  });
}

template <int R, bool INV, typename T>
PSX_HD void dft(cx<T>* x) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    dft2<INV>(x);
  } else if constexpr (R == 3) {
    dft3<INV>(x);
  } else if constexpr (R == 4) {
    dft4<INV>(x);
  } else if constexpr (R == 5) {
    dft5<INV>(x);
  } else if constexpr (R == 8) {
    dft_split<2, 4, INV>(x);
  } else if constexpr (R == 12) {
    dft12<INV>(x);
  } else if constexpr (R == 16) {
    dft_split<4, 4, INV>(x);
  } else if constexpr (R == 32) {
    dft_split<4, 8, INV>(x);
  } else {
    static_assert(R == 1, "unsupported radix");
  }
}

// ------------------------------------------------------------------------------------
// Plans and LDS layouts
// ------------------------------------------------------------------------------------
constexpr int kMaxPasses = 8;

struct FftPlan {
  int L;                 // transform length
  int np;                // number of passes
  int rad[kMaxPasses];   // radices, forward order
  // Offsets into the context's pass-twiddle buffer: twoff[0][p] for pass p of the forward
  // order, twoff[1][q] for pass q of the reversed order (inverse transforms).  Pass p with
  // radix R after Ns = (product of earlier radices) owns (R-1)*Ns entries
  // tw[off + (m-1)*Ns + k] = exp(-2 pi i m k / (Ns R)), m in [1, R), k in [0, Ns), so the
  // lanes of a wave (consecutive k) read consecutive entries.
  int twoff[2][kMaxPasses];
};

// Column tiles: sequence t is a column, element e a row; LDS address e*T + t.  Lanes map to
// consecutive t, so both the HBM rows and the LDS rows are accessed contiguously.
struct ColsLayout {
  int T;       // power of two
  int logT;
  PSX_HD int addr(int e, int t) const { return e * T + t; }
  PSX_HD void decomp(int g, int /*Lr*/, int& j, int& t) const {
    t = g & (T - 1);
    j = g >> logT;
  }
};

// Row tiles: sequence t is a row of length L stored contiguously in LDS with one padding
// element every 16 (breaks the power-of-two strides of the first passes).  Lanes map to
// consecutive butterflies j of one row.
struct RowsLayout {
  int LP;  // padded row pitch
  PSX_HD int addr(int e, int t) const { return t * LP + e + (e >> 4); }
  PSX_HD void decomp(int g, int Lr, int& j, int& t) const {
    if ((Lr & (Lr - 1)) == 0) {
      t = g >> (__builtin_ctz(Lr));
      j = g & (Lr - 1);
    } else {
      t = g / Lr;
      j = g - t * Lr;
    }
  }
};

PSX_HD int rows_pitch(int L) { return L + (L >> 4) + 1; }

// One pass of radix R.  v[b*R + m] is input m of this thread's b-th butterfly on entry (for
// p > 0 it is re-read from LDS) and output q = m on exit.
template <int E, int R, bool INV, typename T, class Lay>
__device__ __forceinline__ void fft_pass(cx<T> (&v)[E], cx<T>* lds, bool from_lds,
                                         bool to_lds, int L, int Ns,
                                         const cx<T>* __restrict__ tw, const Lay& lay, int tid,
                                         int NT) {
  static_assert(E % R == 0, "radix must divide the per-thread element count");
  constexpr int NB = E / R;
  const int Lr = L / R;
  const bool ns_pow2 = (Ns & (Ns - 1)) == 0;
  if (from_lds) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      int j, t;
      lay.decomp(tid + b * NT, Lr, j, t);
#pragma unroll
      for (int m = 0; m < R; ++m) v[b * R + m] = lds[lay.addr(j + m * Lr, t)];
    }
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    int j, t;
    lay.decomp(tid + b * NT, Lr, j, t);
    if (Ns > 1) {
      const int k = ns_pow2 ? (j & (Ns - 1)) : j % Ns;
      const cx<T>* __restrict__ w = tw + k;
#pragma unroll
      for (int m = 1; m < R; ++m) {
        const cx<T> wm = w[(m - 1) * Ns];
        v[b * R + m] = INV ? cmulc(v[b * R + m], wm) : cmul(v[b * R + m], wm);
      }
    }
    dft<R, INV>(&v[b * R]);
  }
  if (to_lds) {
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      int j, t;
      lay.decomp(tid + b * NT, Lr, j, t);
      const int k = ns_pow2 ? (j & (Ns - 1)) : j % Ns;
      const int base = (j - k) * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) lds[lay.addr(base + q * Ns, t)] = v[b * R + q];
    }
    __syncthreads();
  }
}

// Full transform.  On entry v holds the first pass's inputs in register order (see
// reg_elem); on exit it holds the last pass's outputs in register order.  twp is the
// context's pass-twiddle buffer (FftPlan::twoff indexes into it).
template <int E, bool INV, typename T, class Lay>
__device__ __forceinline__ void fft_exec(cx<T> (&v)[E], cx<T>* lds, const FftPlan& pl,
                                         const cx<T>* __restrict__ twp, const Lay& lay, int tid,
                                         int NT) {
  int Ns = 1;
  for (int p = 0; p < pl.np; ++p) {
    const int R = INV ? pl.rad[pl.np - 1 - p] : pl.rad[p];
    const cx<T>* tw = twp + pl.twoff[INV ? 1 : 0][p];
    const bool from_lds = p > 0;
    const bool to_lds = p < pl.np - 1;
    switch (R) {
#define PSX_RADIX(RV)                                                                        \
  case RV:                                                                                  \
    if constexpr (E % RV == 0)                                                              \
      fft_pass<E, RV, INV>(v, lds, from_lds, to_lds, pl.L, Ns, tw, lay, tid, NT);           \
    break;
      PSX_RADIX(2)
      PSX_RADIX(3)
      PSX_RADIX(4)
      PSX_RADIX(5)
      PSX_RADIX(8)
      PSX_RADIX(16)
      PSX_RADIX(32)
#undef PSX_RADIX
      default:
        break;
    }
    Ns *= R;
  }
}

// Register slot i of a pass with radix R holds element e of sequence t.
template <class Lay>
PSX_HD void reg_elem(int i, int R, int L, int tid, int NT, const Lay& lay, int& e, int& t) {
  const int b = i / R;
  const int m = i - b * R;
  const int Lr = L / R;
  int j;
  lay.decomp(tid + b * NT, Lr, j, t);
  e = j + m * Lr;
}

// Radix that the first pass of a transform uses (its inputs' register order) and the radix
// of the last pass (its outputs' register order).
PSX_HD int first_radix(const FftPlan& p, bool inv) { return inv ? p.rad[p.np - 1] : p.rad[0]; }
PSX_HD int last_radix(const FftPlan& p, bool inv) { return inv ? p.rad[0] : p.rad[p.np - 1]; }

}  // namespace psx
