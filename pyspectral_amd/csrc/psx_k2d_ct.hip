// 2D fp32 stage kernels specialised per grid size (PSX_CT2D_SIZES): the same algorithms as
// psx_k2d.hip with the plans, launch shapes and HBM strides fixed at compile time by the
// shared planner geo2d() (psx_geo.h); one work-group per tile.  The bodies and the packed
// line-spectrum convention are in psx_k2d_body.h.
#include "psx_k2d_body.h"


namespace psx {

template <int N>
struct G2 {
  static constexpr Geo2D g = geo2d(N, false);
  static_assert(g.ok, "no 2D plan for this size");
};

// The spectral kernel of the four-step sizes transforms both fields of a direction together
// (two LDS regions; body_spec2d): 8192^2 spectral 278 -> 245 us.  The outer passes keep one
// region: their forward work-groups are held below four waves per SIMD by the LDS with two
// (154 -> 161 us), and with several one-wave tiles per work-group the inverse gained from two
// regions only before (158 -> 148 us with one tile per work-group, 124.5 against 121.0 us
// with eight).  One-pass sizes (whole columns of up to 64 KB per work-group) keep one region:
// two would halve their work-groups per CU.
template <int N>
constexpr bool kMs2 = G2<N>::g.n1 > 1;

template <int N>
__global__ void __launch_bounds__(G2<N>::g.lineNT, 1) kc_line2d(Flow<float> f, int flags) {
  if (flow_inactive(f.prm, f.nh)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  body_line2d<G2<N>::g, true>(f, blockIdx.x, threadIdx.x, flags, smem_raw);
}

template <int N, int stage>
__global__ void __launch_bounds__(G2<N>::g.specNT, 2) kc_spec2d(Flow<float> f) {
  if (flow_inactive(f.prm, f.nh)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  body_spec2d<G2<N>::g, stage, kMs2<N>, true>(f, blockIdx.x, threadIdx.x, smem_raw);
}

// Outer-pass tiles per work-group: tiles of at most one wave (n1 * colT / outE <= 64 threads,
// dividing 64) are packed kOuterS to a work-group, each exchanging through its own LDS with
// wave-level syncs; one-wave work-groups had averaged 1.3 resident waves per SIMD.  8192^2:
// forward 152.7 -> 147.2 us with 4 tiles (158 with 8), inverse 147.5 -> 128.0 us with 8
// (131.9 with 4).
template <int N, bool INV>
constexpr int kOuterS = (G2<N>::g.n1 > 1 && G2<N>::g.outNT <= 64 && 64 % G2<N>::g.outNT == 0 &&
                         (G2<N>::g.n2 * (G2<N>::g.ldw / G2<N>::g.colT)) % (INV ? 8 : 4) == 0)
                            ? (INV ? 8 : 4)
                            : 1;

template <int N, bool INV>
__global__ void __launch_bounds__((G2<N>::g.outNT > 0 ? G2<N>::g.outNT : 64) * (kOuterS<N, INV>), 1)
    kc_outer2d(Flow<float> f) {
  if (flow_inactive(f.prm, f.nh)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr Geo2D G = G2<N>::g;
  constexpr int S = kOuterS<N, INV>;
  // the inverse pass reads its pass twiddles from one LDS copy for the work-group's tiles
  // (8192^2: 128 -> 123 us); the forward pass keeps them in global memory (with the copy it
  // needed 171 VGPRs and took 257 us instead of 148).  Both directions run a tile's two
  // transforms one after the other through one LDS region (8192^2 inverse: 124.5 -> 121.0 us
  // against two regions; 125 us with 6 tiles per work-group, 137 with 12)
  constexpr int TW = ct_twsize(G.outer, INV);
  cx<float>* twl = reinterpret_cast<cx<float>*>(smem_raw) + S * G.n1 * G.colT;
  if constexpr (INV) {
    const cx<float>* tw_o = f.twp + f.pa.twoff[1][0];
    for (int i = threadIdx.x; i < TW; i += blockDim.x) twl[i] = tw_o[i];
    __syncthreads();
  }
  if constexpr (S > 1) {
    const int sub = threadIdx.x / G.outNT;
    body_outer2d<G, INV, false, true, INV>(
        f, blockIdx.x * S + sub, threadIdx.x - sub * G.outNT,
        smem_raw + (size_t)sub * G.n1 * G.colT * 8, twl);
  } else {
    body_outer2d<G, INV, INV, false, INV>(f, blockIdx.x, threadIdx.x, smem_raw, twl);
  }
}

// ------------------------------------------------------------------------------ launchers
bool ct2d_supported(int n) {
  switch (n) {
#define PSX_X(NV) case NV:
    PSX_CT2D_SIZES(PSX_X)
#undef PSX_X
    return true;
    default:
      return false;
  }
}

template <int N>
static hipError_t ct_launch(const Flow<float>& f, int which, int arg, hipStream_t st) {
  constexpr Geo2D G = G2<N>::g;
  constexpr int NJT = G.ldw / G.colT;
  switch (which) {
    case 0: {
      auto fn = kc_line2d<N>;
      static const bool attr_ = set_lds_max((const void*)fn);
      (void)attr_;
      // row tiles and the near passes' twiddle tables
      size_t lds = ((size_t)G.lineT * RowsCT<N, G.lineT>::LP + twl_line2d_elems<G>()) * 8;
      const size_t red = (size_t)2 * G.lineNT * sizeof(double);
      if (lds < red) lds = red;
      hipLaunchKernelGGL(fn, dim3(N / G.lineT), dim3(G.lineNT), lds, st, f, arg);
      break;
    }
    case 1: {
      // transform regions, the pass twiddles of both directions, the tile's four-step twiddles
      // and its FP64 integrating-factor table entries
      const size_t lds = ((size_t)(kMs2<N> ? 2 : 1) * G.n2 * G.colT + ct_twsize(G.spec, false) +
                          ct_twsize(G.spec, true) + (G.n1 > 1 ? G.n2 : 0)) * 8 +
                         (size_t)2 * (G.n2 + G.colT) * sizeof(double);
      const dim3 grid(G.n1 * NJT), block(G.specNT);
      static const bool attr_ = set_lds_max((const void*)kc_spec2d<N, kStagePre>) &&
                                set_lds_max((const void*)kc_spec2d<N, 0>) &&
                                set_lds_max((const void*)kc_spec2d<N, 1>) &&
                                set_lds_max((const void*)kc_spec2d<N, 2>) &&
                                set_lds_max((const void*)kc_spec2d<N, 3>);
      (void)attr_;
      switch (arg) {
        case kStagePre: hipLaunchKernelGGL((kc_spec2d<N, kStagePre>), grid, block, lds, st, f); break;
        case 0: hipLaunchKernelGGL((kc_spec2d<N, 0>), grid, block, lds, st, f); break;
        case 1: hipLaunchKernelGGL((kc_spec2d<N, 1>), grid, block, lds, st, f); break;
        case 2: hipLaunchKernelGGL((kc_spec2d<N, 2>), grid, block, lds, st, f); break;
        case 3: hipLaunchKernelGGL((kc_spec2d<N, 3>), grid, block, lds, st, f); break;
        default: return hipErrorInvalidValue;
      }
      break;
    }
    case 2:
    case 3: {
      if constexpr (G.n1 > 1) {
        if (which == 2) {
          constexpr int S = kOuterS<N, false>;
          const size_t lds = (size_t)S * G.n1 * G.colT * 8;
          auto fn = kc_outer2d<N, false>;
          hipLaunchKernelGGL(fn, dim3(G.n2 * NJT / S), dim3(G.outNT * S), lds, st, f);
        } else {
          constexpr int S = kOuterS<N, true>;
          // one transform region per tile and the pass twiddles
          const size_t lds = ((size_t)S * G.n1 * G.colT + ct_twsize(G.outer, true)) * 8;
          auto fn = kc_outer2d<N, true>;
          static const bool attr_ = set_lds_max((const void*)fn);
          (void)attr_;
          hipLaunchKernelGGL(fn, dim3(G.n2 * NJT / S), dim3(G.outNT * S), lds, st, f);
        }
      } else {
        return hipErrorInvalidValue;
      }
      break;
    }
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// which: 0 line (arg = flags), 1 spectral (arg = stage), 2 outer forward, 3 outer inverse.
hipError_t launch_ct2d(const Flow<float>& f, int which, int arg, hipStream_t st) {
  switch (f.n) {
#define PSX_X(NV) \
  case NV:        \
    return ct_launch<NV>(f, which, arg, st);
    PSX_CT2D_SIZES(PSX_X)
#undef PSX_X
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace psx
