// Ensembles of independent 2D fp32 flows of several grid sizes, advanced together: every
// phase of a stage (line, outer forward, spectral, outer inverse) is one launch of
// persistent kEnsNT-thread work-groups that walk a list of (flow, tile) items, each run with
// the body compiled for that flow's size and spectral variant (psx_geo.h, geo2d_ens).
#pragma once

#include "psx_geo.h"
#include "psx_kernels.h"

namespace psx {

// One flow of the ensemble as the kernels see it.
struct EnsFlow {
  Flow<float> f;
  int code;  // n * 8 + spectral variant (selects the compiled body)
  int pad;
};

// One work-group's worth of a phase: tile `tile` of flow `member` (outer pass: tiles
// tile * outS .. tile * outS + outS - 1).
struct EnsItem {
  int member, tile;
};

constexpr int ens_code(int n, int var) { return n * 8 + var; }

// Phase kinds of launch_ens: 0 line (arg = flags), 1 spectral (arg = stage), 2 outer
// forward, 3 outer inverse.
hipError_t launch_ens(int kind, int arg, const EnsFlow* flows, const EnsItem* items, int nitems,
                      int grid, size_t lds, hipStream_t st);
// Work-groups of phase (kind, arg) that can be resident on one CU with lds bytes of LDS each
// (0 if the runtime cannot say).
int ens_resident(int kind, int arg, size_t lds);
// Hermitian projection of every member's ky = 0 column after a step (nmax: largest n).
hipError_t launch_ens_hermitian(const EnsFlow* flows, int count, int nmax, hipStream_t st);
// True if (n, var) has compiled ensemble bodies.
bool ens_compiled(int n, int var);

}  // namespace psx
