// Coarse-grid correction for the warm starts of the mixed forms (o3d_mg.hip).
//
// The Poisson problem L u = f on a cell-centred grid (the block averages of
// the fine grid) is solved approximately by multigrid V-cycles: red-black
// Gauss-Seidel smoothing, restriction by 2x2x2 averages, trilinear
// prolongation, and the coarsest level (at most 64 cells where the axes
// halve that far) solved by one block in LDS.  Periodic
// axes wrap; the others use the cell-centred mirror (ghost = edge cell).
#pragma once
#include <hip/hip_runtime.h>

#include "o3d_common.h"

namespace o3d {

constexpr int kMgMaxLevels = 10;
constexpr int kMgCoarsestCells = 4096;  // single-block solve (LDS)

struct MgHierarchy {
  int levels = 0;
  int m[kMgMaxLevels][3] = {};
  float *u[kMgMaxLevels] = {}, *f[kMgMaxLevels] = {}, *r[kMgMaxLevels] = {};
  double c[kMgMaxLevels][3] = {};  // 1 / h^2 per axis on each level
  int per[3] = {0, 0, 0};
};

// Levels for a level-0 grid of m0 cells with spacings h; returns the float
// count the hierarchy needs (bind with mg_bind).
size_t mg_plan(MgHierarchy &H, const int m0[3], const double h[3], const int per[3]);
void mg_bind(MgHierarchy &H, float *base);
// `cycles` V(nu1, nu2) cycles on L u0 = f0 (f0 = H.f[0] filled by the
// caller) from u0 = 0, or from the u0 = H.u[0] the caller wrote
// (from_zero = false).
hipError_t mg_solve(const MgHierarchy &H, int cycles, int nu1, int nu2, hipStream_t s,
                    bool from_zero = true);
// The coarse-grid correction of a correction solve held in colour-split
// storage (per row, the red nodes then the black ones, nx/2 entries each;
// the red half-array first): f0 = 2^3 block averages of the split residual
// r, and e = the trilinear interpolation of u0, written split.  The level-0
// grid is (nx/2, ny/2, nz/2).
hipError_t mg_pool_split(const float *r, float *f0, int nx, int ny, int nz, hipStream_t s);
hipError_t mg_prolong_split(float *e, const float *u0, int nx, int ny, int nz, const int per[3],
                            hipStream_t s);
// u0 = the trilinear interpolation of the level-1 field H.u[1] (a starting
// iterate given on the next coarser level).  Needs two levels at least.
hipError_t mg_seed_from_level1(const MgHierarchy &H, hipStream_t s);

}  // namespace o3d
