// Flow descriptor shared by the host planner and the stage kernels.
//
// HBM layout of one flow (all complex fields in the working precision T):
//   u0, acc   spectral state and RK accumulator in r2c layout with a padded row pitch sp >= nh
//             (nh = n/2 + 1): element (.., row, j) at row * sp + j, 2D rows n, 3D rows 3 n^2;
//             the first nh columns of each row are the NumPy rfftn coefficients.  Only the
//             retained 2/3-rule modes are ever read or written.
//   W         work fields, each of nrows x ldw complex, nrows = n^(d-1).  Column index j is the
//             half-axis wavenumber and only j < jm (the retained ones, 3 j < n) are stored;
//             ldw pads jm to a whole number of column tiles.  The same field holds, per stage,
//             (a) the inverse-transformed (in every full axis) velocity components, read by the
//             line kernel, and (b) the forward-transformed products the line kernel writes back.
//   tw        tw[i] = exp(-2 pi i * i / n), i < n.
//   prm       FP64 parameters: integrating-factor tables and scalars (see PrmLayout).
//   stats     maxima |u_i| (CFL), stored as FP64 bit patterns of non-negative numbers.
#pragma once

#include "psx_fft.h"

namespace psx {

// FP64 parameter block: [0, nh) exp(-nu m^2 dt), [nh, 2 nh) exp(-nu m^2 dt/2), then scalars.
// Phase shift (Flow::ps): kPsA the step's grid shift (3 entries: stages 0-1 and the inputs of
// stages 0-1), kPsB that shift + pi/n (stages 2-3 and their inputs), kPsN the next step's shift
// (inputs prepared by stage 3).
struct PrmLayout {
  static constexpr int kDt = 0, kEa1 = 1, kEah = 2, kEb1 = 3, kEbh = 4, kKlo2 = 5, kKhi2 = 6,
                       kKa2 = 7, kBand = 8, kActive = 9, kPsA = 16, kPsB = 19, kPsN = 22,
                       kNScalars = 32;
};

template <typename T>
struct Flow {
  cx<T>* W;          // work fields
  cx<T>* u0;         // state
  cx<T>* acc;        // RK accumulator
  const cx<T>* tw;   // twiddles exp(-2 pi i m / n), m < n (four-step and diagnostics)
  const cx<T>* twp;  // per-pass twiddle tables of all plans (FftPlan::twoff)
  const double* prm; // 2*nh table entries followed by PrmLayout scalars
  unsigned long long* stats;  // CFL maxima max|u_i| as FP64 bit patterns, 3 entries
  const cx<T>* src;  // 3D: optional source added to the nonlinear term of every stage, in the
                     // state layout (null: none); e.g. a subgrid-stress divergence
  long long fstride; // complex elements per work field (nrows * ldw)
  long long sstride; // complex elements per state component (3D)
  int n, nh, jm, ldw;
  int sp;            // state row pitch (complex elements)
  int dim, rot;      // rot: 3D rotational form
  int ps;            // phase-shift dealiasing: spherical mask 9 |k|^2 < 2 n^2, shifted grids
  int n1, n2;        // four-step split of the full axes (n1 = 1: single pass)
  int rowT;          // rows per work-group in the line kernel
  int colT;          // columns per work-group in the column kernels
  int yT;            // columns per work-group in the 3D y kernel
  FftPlan prow;      // line (real axis) transform, length n
  FftPlan pcol;      // full-axis transform of the fused spectral kernel (length n or n2)
  FftPlan pa;        // outer four-step transform (length n1)
  FftPlan py;        // 3D middle-axis transform (length n)
  T scale;           // 1 / n^dim
};

}  // namespace psx
