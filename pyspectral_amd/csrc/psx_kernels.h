// Declarations shared by the kernel translation units and the host planner.
#pragma once

#include <hip/hip_runtime.h>

#include "psx_drv.h"
#include "psx_stage.h"

namespace psx {

// Launch shape chosen by the planner for one kernel.
struct Launch {
  int E;        // elements per thread (compile-time variant)
  int threads;  // work-group size
  int blocks;   // grid size
  size_t lds;   // dynamic LDS bytes
};

enum : int { kLineWantMax = 1, kLineMaxOnly = 2 };

// Allow a kernel the whole LDS of a gfx950 CU (160 KB) as dynamic shared memory; called once
// per kernel from the launchers.
inline bool set_lds_max(const void* fn) {
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) ==
         hipSuccess;
}

// Elements-per-thread variants compiled for the transform kernels (the planner in psx_api.cpp
// only picks these): 2D fp32 {8,16,20,24,32}, fp64 {8,12,16,20}; 3D {8,12,16,20}.  20 serves
// the lengths 5 * 2^a (radix-5 pass).
template <typename T>
constexpr bool e_ok_2d(int E) {
  return sizeof(T) == 4 ? (E == 8 || E == 16 || E == 20 || E == 24 || E == 32)
                        : (E == 8 || E == 12 || E == 16 || E == 20);
}
constexpr bool e_ok_3d(int E) { return E == 8 || E == 12 || E == 16 || E == 20; }
#define PSX_E_CASES PSX_CASE(8) PSX_CASE(12) PSX_CASE(16) PSX_CASE(20) PSX_CASE(24) PSX_CASE(32)

template <typename T>
hipError_t launch_line2d(const Flow<T>& f, const Launch& lc, int flags, hipStream_t st);
template <typename T>
hipError_t launch_spec2d(const Flow<T>& f, const Launch& lc, bool four, int stage,
                         hipStream_t st);
template <typename T>
hipError_t launch_outer2d(const Flow<T>& f, const Launch& lc, bool inv, hipStream_t st);

template <typename T>
hipError_t launch_line3d(const Flow<T>& f, const Launch& lc, int flags, hipStream_t st);
template <typename T>
hipError_t launch_ypass3d(const Flow<T>& f, const Launch& lc, bool inv, int nfld,
                          hipStream_t st);
template <typename T>
hipError_t launch_spec3d(const Flow<T>& f, const Launch& lc, int stage, hipStream_t st);

// Compile-time specialised 2D fp32 kernels (psx_k2d_ct.hip).  which: 0 line (arg = flags),
// 1 spectral (arg = stage), 2 outer forward, 3 outer inverse.
bool ct2d_supported(int n);
hipError_t launch_ct2d(const Flow<float>& f, int which, int arg, hipStream_t st);

// Compile-time specialised 3D fp32 kernels (psx_k3d_ct.hip).  which: 0 z lines (arg = flags),
// 1 y forward, 2 x pass with physics (arg = stage), 3 y inverse.
bool ct3d_supported(int n);
hipError_t launch_ct3d(const Flow<float>& f, int which, int arg, hipStream_t st);

// Device controller (psx_kdrv.hip).  which: 0 before the first stage's transforms, 1 after
// the last stage, 2 record reduction, 3 / 4 mark the flows' parameter blocks active /
// inactive.
hipError_t launch_drv(int which, const DrvMember* mem, int count, hipStream_t st);

// Utility / diagnostics kernels (psx_kutil.hip).
// Hermitian projection of the state's self-mirrored plane after a step (k_hermitian).
template <typename T>
hipError_t launch_hermitian(const Flow<T>& f, hipStream_t st);
hipError_t launch_params(double* prm, int nh, double nu, double dt, double alpha, double amp,
                         double klo2, double khi2, double ka2, hipStream_t st);
hipError_t launch_fill_u32(unsigned* p, unsigned v, int n, hipStream_t st);
// Write n <= 16 doubles (kernel arguments) to device memory: a kernel, so captured graphs
// and the parameter block's other writers stay ordered with it on the stream.
hipError_t launch_set_f64(double* p, const double* v, int n, hipStream_t st);
hipError_t launch_fill_f64(double* p, double v, int n, hipStream_t st);
template <typename T>
hipError_t launch_reduce(const cx<T>* state, long long sstride, int pitch, int n, int dim,
                         double klo2, double khi2, double* partial, int nblocks, double* out,
                         hipStream_t st);
template <typename T>
hipError_t launch_spectrum(const cx<T>* state, long long sstride, int pitch, int n, int dim,
                           double* out, int nshell, hipStream_t st);

}  // namespace psx
