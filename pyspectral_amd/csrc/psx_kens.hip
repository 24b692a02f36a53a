// Persistent multi-size ensemble kernels (psx_ens.h): one launch per phase for every flow
// of an ensemble.  The line and spectral bodies read their pass twiddles from LDS copies made
// per item (forced ensemble: spectral phase 285 -> 257 us).  Work-group b handles items b, b + gridDim.x, ...; each item runs the body
// of psx_k2d_body.h compiled for its flow's (size, spectral variant), so flows of different
// sizes share one launch and the launch never waits on a smaller grid's tail.
#include "psx_ens.h"
#include "psx_k2d_body.h"

namespace psx {

template <int N, int V>
struct GE {
  static constexpr Geo2D g = geo2d_ens(N, V);
  static_assert(g.ok, "no ensemble geometry for this size and variant");
};

template <int KIND, int ARG>
__device__ __forceinline__ void ens_item(const EnsFlow& ef, int tile, int tid, int flags,
                                         unsigned char* smem) {
  switch (ef.code) {
#define PSX_X(NV, VV)                                                                  \
  case ens_code(NV, VV): {                                                             \
    constexpr Geo2D G = GE<NV, VV>::g;                                                 \
    if constexpr (KIND == 0) {                                                         \
      body_line2d<G, true>(ef.f, tile, tid, flags, smem);                                    \
    } else if constexpr (KIND == 1) {                                                  \
      body_spec2d<G, ARG, false, true>(ef.f, tile, tid, smem);                                      \
    } else if constexpr (G.n1 > 1) {                                                   \
      const int sub = tid / G.outNT;                                                   \
      body_outer2d<G, KIND == 3>(ef.f, tile * G.outS + sub, tid % G.outNT,             \
                                 smem + (size_t)sub * G.n1 * G.colT * 8);              \
    }                                                                                  \
    break;                                                                             \
  }
    PSX_ENS_VARIANTS(PSX_X)
#undef PSX_X
    default:
      break;
  }
}

// The spectral phase is held to 168 VGPRs, three work-groups per CU (forced ensemble: 202 ->
// 178 us; 170-182 VGPRs and two work-groups without the bound); the inverse outer pass would
// spill at that bound, and the line phase at 129 VGPRs already holds three.
template <int KIND, int ARG>
__global__ void __launch_bounds__(kEnsNT, KIND == 1 ? 3 : 2)
    ke_phase(const EnsFlow* __restrict__ flows, const EnsItem* __restrict__ items, int nitems,
             int flags) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const EnsItem item = items[it];
    const EnsFlow& ef = flows[item.member];
    // the thread index is made opaque per item so that no body's index arithmetic is hoisted
    // out of the loop (it would stay live across every variant's code)
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    if (!flow_inactive(ef.f.prm, ef.f.nh))
      ens_item<KIND, ARG>(ef, item.tile, tid, flags, smem_raw);
    __syncthreads();  // the next item reuses the LDS
  }
}

// Hermitian projection of every member's ky = 0 column (k_hermitian, one member per grid row).
__global__ void ke_hermitian(const EnsFlow* __restrict__ flows) {
  const Flow<float>& f = flows[blockIdx.y].f;
  if (flow_inactive(f.prm, f.nh)) return;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = f.n;
  if (r >= n) return;
  const int rm = r ? n - r : 0;
  if (rm < r) return;
  cx<float>* a = f.u0 + (long long)r * f.sp;
  if (rm == r) {
    a->y = 0.f;
    return;
  }
  cx<float>* b = f.u0 + (long long)rm * f.sp;
  const cx<float> ua = *a, ub = *b;
  const cx<float> h = mkc(0.5f * (ua.x + ub.x), 0.5f * (ua.y - ub.y));
  *a = h;
  *b = conjg(h);
}

hipError_t launch_ens_hermitian(const EnsFlow* flows, int count, int nmax, hipStream_t st) {
  hipLaunchKernelGGL(ke_hermitian, dim3((nmax + 255) / 256, count), dim3(256), 0, st, flows);
  return hipGetLastError();
}

bool ens_compiled(int n, int var) {
  switch (ens_code(n, var)) {
#define PSX_X(NV, VV) case ens_code(NV, VV):
    PSX_ENS_VARIANTS(PSX_X)
#undef PSX_X
    return true;
    default:
      return false;
  }
}

template <int KIND, int ARG>
static hipError_t ens_go(const EnsFlow* flows, const EnsItem* items, int nitems, int grid,
                         size_t lds, int flags, hipStream_t st) {
  auto fn = ke_phase<KIND, ARG>;
  static const bool attr_ = set_lds_max((const void*)fn);
  (void)attr_;
  hipLaunchKernelGGL(fn, dim3(grid), dim3(kEnsNT), lds, st, flows, items, nitems, flags);
  return hipGetLastError();
}

template <int KIND, int ARG>
static int ens_res(size_t lds) {
  auto fn = ke_phase<KIND, ARG>;
  static const bool attr_ = set_lds_max((const void*)fn);
  (void)attr_;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kEnsNT, lds) != hipSuccess) return 0;
  return nb;
}

int ens_resident(int kind, int arg, size_t lds) {
  switch (kind) {
    case 0: return ens_res<0, 0>(lds);
    case 1:
      switch (arg) {
        case kStagePre: return ens_res<1, kStagePre>(lds);
        case 0: return ens_res<1, 0>(lds);
        case 1: return ens_res<1, 1>(lds);
        case 2: return ens_res<1, 2>(lds);
        case 3: return ens_res<1, 3>(lds);
        default: return 0;
      }
    case 2: return ens_res<2, 0>(lds);
    case 3: return ens_res<3, 0>(lds);
    default: return 0;
  }
}

hipError_t launch_ens(int kind, int arg, const EnsFlow* flows, const EnsItem* items, int nitems,
                      int grid, size_t lds, hipStream_t st) {
  if (nitems <= 0 || grid <= 0) return hipErrorInvalidValue;
  switch (kind) {
    case 0: return ens_go<0, 0>(flows, items, nitems, grid, lds, arg, st);
    case 1:
      switch (arg) {
        case kStagePre: return ens_go<1, kStagePre>(flows, items, nitems, grid, lds, 0, st);
        case 0: return ens_go<1, 0>(flows, items, nitems, grid, lds, 0, st);
        case 1: return ens_go<1, 1>(flows, items, nitems, grid, lds, 0, st);
        case 2: return ens_go<1, 2>(flows, items, nitems, grid, lds, 0, st);
        case 3: return ens_go<1, 3>(flows, items, nitems, grid, lds, 0, st);
        default: return hipErrorInvalidValue;
      }
    case 2: return ens_go<2, 0>(flows, items, nitems, grid, lds, 0, st);
    case 3: return ens_go<3, 0>(flows, items, nitems, grid, lds, 0, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace psx
