// Compile-time and run-time planning shared by the host planner and the kernels.
//
// One constexpr planner decides, for a transform length and an elements-per-thread count,
// the radix list; and, for a 2D grid size, every launch shape (work-group size, tile width,
// four-step split).  The host fills the run-time Flow/Launch structures from it and the
// specialised kernels evaluate the same functions at compile time, so both always agree on
// register order and on the layout of the pass-twiddle tables.
#pragma once

namespace psx {

constexpr int kMaxRad = 8;

struct CtPlan {
  int L = 0;
  int np = 0;
  int rad[kMaxRad] = {1, 1, 1, 1, 1, 1, 1, 1};
};

// Radix list for length L with every radix dividing E; balanced power-of-two radices
// (<= 32) first, then 3s and 5s.  np == 0 if impossible.
constexpr CtPlan ct_plan(int L, int E) {
  CtPlan p{};
  int l = L, a = 0, b3 = 0, b5 = 0;
  if (L <= 0 || E <= 0) return p;
  while (l % 2 == 0) { l /= 2; ++a; }
  while (l % 3 == 0) { l /= 3; ++b3; }
  while (l % 5 == 0) { l /= 5; ++b5; }
  if (l != 1) return p;
  if (b3 && E % 3) return p;
  if (b5 && E % 5) return p;
  int lmax = 0;
  while (lmax < 5 && E % (1 << (lmax + 1)) == 0) ++lmax;
  if (a > 0 && lmax == 0) return p;
  const int np2 = a == 0 ? 0 : (a + lmax - 1) / lmax;
  const int np = np2 + b3 + b5;
  if (np > kMaxRad || np == 0) return p;
  p.L = L;
  p.np = np;
  int k = 0;
  for (int i = 0; i < np2; ++i) {
    const int bits = a / np2 + (i < a % np2 ? 1 : 0);
    p.rad[k++] = 1 << bits;
  }
  for (int i = 0; i < b3; ++i) p.rad[k++] = 3;
  for (int i = 0; i < b5; ++i) p.rad[k++] = 5;
  return p;
}

// As ct_plan, with one radix-12 pass (prime-factor 3 x 4 butterfly) in place of a radix-3
// pass and two factors of 2, when 12 divides both L and E: one pass fewer for the 3 * 2^a
// lengths (384 = 4 4 2 12 instead of 4 4 4 2 3).  Used by the compile-time 3D kernels.
constexpr CtPlan ct_plan_r12(int L, int E) {
  if (L % 12 || E % 12 || L == 12) return ct_plan(L, E);
  CtPlan p = ct_plan(L / 12, E);
  if (p.np == 0 || p.np >= kMaxRad) return ct_plan(L, E);
  p.rad[p.np++] = 12;
  p.L = L;
  return p;
}

// Radix of pass q in the forward (inv = false) or reversed (inv = true) order.
constexpr int ct_radix(const CtPlan& p, bool inv, int q) {
  return inv ? p.rad[p.np - 1 - q] : p.rad[q];
}
// Product of the radices before pass q.
constexpr int ct_ns(const CtPlan& p, bool inv, int q) {
  int ns = 1;
  for (int i = 0; i < q; ++i) ns *= ct_radix(p, inv, i);
  return ns;
}
// Offset of pass q's twiddles inside one order's table (pass 0 owns none).
constexpr int ct_twoff(const CtPlan& p, bool inv, int q) {
  int off = 0;
  for (int i = 1; i < q; ++i) off += (ct_radix(p, inv, i) - 1) * ct_ns(p, inv, i);
  return off;
}
// Entries of one order's table.
constexpr int ct_twsize(const CtPlan& p, bool inv) { return ct_twoff(p, inv, p.np); }

constexpr int ct_rows_pitch(int L) { return L + (L >> 4) + 1; }
// Row pitch of the state (complex elements): n/2+1 rounded up to whole 128-byte lines.
constexpr int ct_state_pitch(int n, bool fp64) {
  const int q = fp64 ? 8 : 16;
  return (n / 2 + 1 + q - 1) / q * q;
}
constexpr int ct_retained(int n) { return (n - 1) / 3 + 1; }
// Retained half-axis width of the spherical phase-shift mask: 9 j^2 < 2 n^2 for j < jm.
constexpr int ct_retained_ps(int n) {
  int j = 0;
  while (9LL * j * j < 2LL * n * n) ++j;
  return j;
}

// Launch geometry of the 2D kernels for grid size n in fp32 (csz = 8) or fp64 (csz = 16).
struct Geo2D {
  bool ok = false;
  int n = 0, jm = 0, colT = 0, ldw = 0, n1 = 1, n2 = 0;
  int sp = 0;  // state row pitch: n/2+1 rounded up to whole 128-byte lines
  int lineE = 0, lineNT = 0, lineT = 0;
  int specE = 0, specNT = 0;
  int outE = 0, outNT = 0;
  int outS = 1;  // outer-pass tiles per work-group (ensemble kernels)
  CtPlan line{}, spec{}, outer{};
};

constexpr int ct_pick_E(int L, bool fp64) {
  if (L % 5 == 0) return 20;
  if (L % 3 == 0) return fp64 ? 12 : 24;
  return 16;
}

constexpr Geo2D geo2d(int n, bool fp64, int jm = 0) {
  Geo2D g{};
  const int csz = fp64 ? 16 : 8;
  g.n = n;
  g.jm = jm > 0 ? jm : ct_retained(n);
  g.sp = ct_state_pitch(n, fp64);
  int T = fp64 ? 8 : 16;
  while (T > 1 && T / 2 >= g.jm) T /= 2;
  g.colT = T;
  g.ldw = (g.jm + T - 1) / T * T;
  // line kernel: whole rows (long fp32 rows: 32 elements per thread, 3 passes at 8192); the
  // rows per work-group divide n, so every launch covers the rows exactly (n = 3 * 2^a
  // included)
  {
    int E = (!fp64 && n % 3 != 0 && n >= 4096) ? 32 : ct_pick_E(n, fp64);
    int t = 1;
    while (n * t / E < 128 && n % (t * 2) == 0) t *= 2;
    int nt = n * t / E;
    if (nt > 512 && !fp64 && n % 3 != 0) {
      E = 32;
      nt = n * t / E;
    }
    g.line = ct_plan(n, E);
    if (g.line.np == 0 || nt > 512 || (n * t) % E) return g;
    g.lineE = E;
    g.lineNT = nt;
    g.lineT = t;
  }
  // column kernels: one pass if a whole-column tile fits 64 KB, else the four-step split
  const int budget = 65536;
  auto cols = [&](int L, int& E, int& nt, CtPlan& pl) {
    E = ct_pick_E(L, fp64);
    while (E > 4 && (L * T) % E) E /= 2;
    pl = ct_plan(L, E);
    nt = L * T / E;
    return pl.np > 0 && nt >= 1 && nt <= 512 && (L * T) % E == 0;
  };
  if (n * T * csz <= budget || n <= 16) {
    g.n1 = 1;
    g.n2 = n;
    if (!cols(n, g.specE, g.specNT, g.spec)) return g;
  } else {
    int best = 0;
    for (int d = 2; d < n; ++d) {
      if (n % d) continue;
      if (d * d < n) continue;
      if (d * T * csz > budget) continue;
      best = d;
      break;
    }
    if (!best) return g;
    g.n2 = best;
    g.n1 = n / best;
    if (!cols(g.n2, g.specE, g.specNT, g.spec)) return g;
    // fp32: 8 elements per thread in the inner kernel, which holds both fields' transforms,
    // the state update and the next inputs (16 needed all 256 VGPRs: two waves per SIMD)
    if (!fp64 && g.specE == 16) {
      g.specE = 8;
      g.spec = ct_plan(g.n2, 8);
      g.specNT = g.n2 * T / 8;
      if (g.spec.np == 0 || g.specNT > 512) return g;
    }
    if (!cols(g.n1, g.outE, g.outNT, g.outer)) return g;
  }
  g.ok = true;
  return g;
}

// Grid sizes with compile-time specialised 2D fp32 kernels.
#define PSX_CT2D_SIZES(X) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) X(8192)

// ---------------------------------------------------------------------------- ensembles
// Ensemble kernels run every phase as one launch of kEnsNT-thread work-groups over the tiles
// of flows of several sizes, so every geometry below gives each kernel kind kEnsNT threads
// per work-group (the outer four-step pass as outS independent tiles of outNT threads).
// Spectral-pass variants (the autotuner's choices):
//   kEnsS16, kEnsS8, kEnsS4   one pass over whole columns, column tiles of 16, 8 or 4;
//   kEnsF256, kEnsF128        four-step split with inner length 256 or 128 (tiles of 16
//                             columns) and an outer pass of n/256 or n/128 points held in
//                             the registers of one thread.
constexpr int kEnsNT = 256;
enum : int { kEnsS16 = 0, kEnsS8 = 1, kEnsS4 = 2, kEnsF256 = 3, kEnsF128 = 4, kEnsNVar = 5 };

constexpr Geo2D geo2d_ens(int n, int var) {
  Geo2D g{};
  if (n < 64 || (n & (n - 1)) || var < 0 || var >= kEnsNVar) return g;
  g.n = n;
  g.jm = ct_retained(n);
  g.sp = ct_state_pitch(n, false);
  // line kernel: 16 elements per thread, kEnsNT * 16 / n whole rows per work-group
  g.lineE = 16;
  g.lineNT = kEnsNT;
  g.lineT = kEnsNT * 16 / n;
  g.line = ct_plan(n, 16);
  if (g.lineT < 1 || g.line.np == 0) return g;
  int T = 16, L = n;
  if (var == kEnsS8) T = 8;
  if (var == kEnsS4) T = 4;
  if (var == kEnsF256) L = 256;
  if (var == kEnsF128) L = 128;
  if (L > n) return g;
  const int E = L * T / kEnsNT;
  if (E < 4 || E > 16) return g;
  g.colT = T;
  g.ldw = (g.jm + T - 1) / T * T;
  g.n2 = L;
  g.n1 = n / L;
  g.specE = E;
  g.specNT = kEnsNT;
  g.spec = ct_plan(L, E);
  if (g.spec.np == 0 || L * T * 8 > 65536) return g;
  if (g.n1 > 1) {
    if (g.n1 > 16) return g;
    g.outE = g.n1;
    g.outNT = T;  // one thread per column, all n1 points in registers
    g.outS = kEnsNT / g.outNT;
    g.outer = ct_plan(g.n1, g.n1);
    if (g.outer.np != 1) return g;
    if ((g.n2 * (g.ldw / T)) % g.outS) return g;
  }
  g.ok = true;
  return g;
}

// Ensemble member sizes and the spectral variants compiled for each (first = default).  Every
// variant keeps at most 8 elements per thread: all bodies share the persistent spectral-phase
// kernel, whose registers the largest body sets, and with the 16-element variants (256 S16,
// 512 S8, 1024 S4, 2048 F256) it spilled to scratch; without them the forced ensemble's
// spectral phase takes 305 instead of 347 µs.
#define PSX_ENS_VARIANTS(X)                                                       \
  X(64, kEnsS16) X(128, kEnsS16) X(128, kEnsS8) X(256, kEnsS8) X(256, kEnsF128) \
  X(512, kEnsS4) X(512, kEnsF128) X(1024, kEnsF128) X(2048, kEnsF128)

// Launch geometry of the 3D kernels (fp32): z lines, y and x column passes over tiles of
// T columns of the retained half axis.
struct Geo3D {
  bool ok = false;
  int n = 0, jm = 0, T = 0, ldw = 0, sp = 0;
  int lineE = 0, lineNT = 0, lineT = 0;
  int yE = 0, yNT = 0;
  int xE = 0, xNT = 0;
  CtPlan line{}, yp{}, xp{};
};

constexpr Geo3D geo3d(int n) {
  Geo3D g{};
  g.n = n;
  g.jm = ct_retained(n);
  g.sp = ct_state_pitch(n, false);
  // small power-of-two grids: 4-column tiles and short rows per thread, so that the column
  // passes launch more work-groups than the GPU has CUs (64^3 with 16-column tiles: 128)
  const bool small = n <= 64 && (n & (n - 1)) == 0;
  int T = small ? 4 : 16;
  while (T > 1 && T / 2 >= g.jm) T /= 2;
  g.T = T;
  g.ldw = (g.jm + T - 1) / T * T;
  const bool r3 = n % 3 == 0;
  // radix-3 sizes: 12 elements per thread (24 would need more than 256 VGPRs for the
  // line kernel's several live sequences and drop it to one wave per SIMD)
  {
    const int E = r3 ? 12 : (small ? 8 : 16);
    int t = 1;
    while (n * t / E < 128 && (n * n) % (t * 2) == 0) t *= 2;  // t divides the n^2 rows
    const int nt = n * t / E;
    g.line = ct_plan_r12(n, E);
    if (g.line.np == 0 || nt > 512 || (n * t) % E) return g;
    g.lineE = E;
    g.lineNT = nt;
    g.lineT = t;
  }
  auto cols = [&](int E0, int& E, int& nt, CtPlan& pl, bool r12) {
    E = E0;
    while (E > 4 && ((n * T) % E || n * T / E > 512)) E = (E % 3 == 0 && E > 12) ? E / 2 : E;
    pl = r12 ? ct_plan_r12(n, E) : ct_plan(n, E);
    nt = n * T / E;
    return pl.np > 0 && nt >= 1 && nt <= 512 && (n * T) % E == 0;
  };
  if (n * T * 8 > 65536) return g;
  if (!cols(r3 ? 12 : (small ? 4 : 16), g.yE, g.yNT, g.yp, true)) return g;
  // the x pass keeps three accumulated components, the state and the transform in registers:
  // 8 elements per thread from 128 up (16 spilled past 256 VGPRs); radix-3 plans without the
  // radix-12 pass, whose butterfly spills next to the accumulators at 384
  if (!cols(r3 ? 12 : (small ? 4 : (n >= 128 ? 8 : 16)), g.xE, g.xNT, g.xp, false)) return g;
  g.ok = true;
  return g;
}

// Grid sizes with compile-time specialised 3D fp32 kernels.
#define PSX_CT3D_SIZES(X) X(64) X(96) X(128) X(192) X(256) X(384)

}  // namespace psx
