// Host-side launch interface of the stencil kernels (o3d_stencil.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "o3d_common.h"

namespace o3d {

// Finite-difference coefficients exactly as the reference precomputes them.
struct D1Coef { double a, b, c; };  // 1/(60h), 9/(60h), 45/(60h): src/derivation.f90:26-30
struct D2Coef { double a, b, c; };  // 1/(12h^2), 16/(12h^2), 30/(12h^2): :517-521
inline D1Coef d1coef(double h) {
  const double s = 60.0 * h;
  return {1.0 / s, 9.0 / s, 45.0 / s};
}
inline D2Coef d2coef(double h) {
  const double t = 12.0 * h * h;
  return {1.0 / t, 16.0 / t, 30.0 / t};
}

// Tile geometry of the z-marching stencil kernels.
struct Tiling {
  int ntx, nty, ntz, zc;
  int blocks() const { return ntx * nty * ntz; }
};
constexpr int kTX = 64;  // one wave per x-row
constexpr int kTY = 8;
Tiling make_tiling(int nx, int ny, int nz);
// Tiling whose grid fills whole rounds of `slots` resident blocks.
Tiling make_tiling_slots(int nx, int ny, int nz, int slots);
// The tilings the divergence and correction kernels launch with (from
// their occupancy on the current device).
template <typename T>
Tiling div_tiling(int nx, int ny, int nz);
template <typename T>
Tiling predict_tiling(int nx, int ny, int nz, bool les);
template <typename T, typename TP>
Tiling correct_tiling(int nx, int ny, int nz);

template <typename T>
struct PredictArgs {
  const T *u[3];
  T *up[3];
  T *f1[3];        // new history slot (written)
  const T *f2[3];  // n-1 (read when nh >= 2)
  const T *f3[3];  // n-2 (read when nh == 3)
  T *nut;          // nu_t out (may be null)
  FieldPartial *nut_part;  // per-block nu_t statistics (LES)
  int nx, ny, nz, bc[3];
  Tiling tl;
  T d1[3][3], d2[3][3];
  T adu, bdu, cdu, onere, csd2;
};

template <typename T>
struct DivArgs {
  const T *f[3];
  T *out;            // may be null (statistics only)
  FieldPartial *part;
  int nx, ny, nz, bc[3], odd;
  Tiling tl;
  T d1[3][3];
  T divisor;         // rhs = div / dt (src/integration.f90:239); 1 for a plain divergence
};

template <typename T, typename TP>
struct CorrectArgs {
  T *u[3];
  const T *up[3];
  const TP *p;
  FieldPartial *part;  // 3 records per block
  int nx, ny, nz, bc[3];
  Tiling tl;
  TP d1[3][3];
  TP dt;
  // Optional, after a solve of the mixed forms.  mean: p - mean[0] is the
  // pressure (every loaded value shifted, so the gradient is the shifted
  // pressure's) and pout receives it at the owned nodes (another buffer:
  // neighbouring tiles still load p).  p0: the FP32 increment
  // e = (p - mean) - p0 of the owned nodes into e, with the block's sum of
  // e^2 in epart[block].
  const double *mean = nullptr;
  TP *pout = nullptr;
  const TP *p0 = nullptr;
  float *e = nullptr;
  double *epart = nullptr;
};

template <typename T>
struct StatsArgs {
  const T *u[3];
  double *part;  // 17 sums per block
  int nx, ny, nz, bc[3];
  Tiling tl;
  T d1[3][3], d2[3][3];
  double xnu;
};

template <typename T>
struct DiagArgs {
  const T *u[3];
  T *out[3];     // mode 0: curl components; mode 1: out[0] = Q
  int mode;
  int nx, ny, nz, bc[3];
  Tiling tl;
  T d1[3][3];
};

template <typename T>
struct TranseqArgs {
  const T *phi;      // phi before the step (not modified)
  T *phi_new;        // updated phi before clipping (scratch)
  T *phi_out;        // clipped phi (may alias phi)
  const T *u[3];
  const T *src;
  const T *nu_t;     // read when les
  T *f1;             // F of this step
  const T *f2, *f3;  // F of the two previous steps (read when nh >= 2 / 3)
  double *part;      // 3 sums per block
  int nh, les;
  T adu, bdu, cdu, alpha0, sc;
  int nx, ny, nz, bc[3];
  Tiling tl;
  T d1[3][3], d2[3][3];
};

template <typename T>
hipError_t launch_transeq(const TranseqArgs<T> &a, double *sums, hipStream_t s);

template <typename T>
hipError_t launch_der(int which, T *df, const T *f, int nx, int ny, int nz, double h,
                      hipStream_t s);
template <typename T>
hipError_t launch_diag(const DiagArgs<T> &a, hipStream_t s);
template <typename T>
hipError_t launch_predict(const PredictArgs<T> &a, int nh, bool les, hipStream_t s);
template <typename T>
hipError_t launch_divergence(const DivArgs<T> &a, hipStream_t s);
template <typename T, typename TP>
hipError_t launch_correct(const CorrectArgs<T, TP> &a, hipStream_t s);
template <typename T>
hipError_t launch_statistics(const StatsArgs<T> &a, hipStream_t s);
template <typename T>
hipError_t launch_field_stats(const T *f, int n, FieldPartial *part, int nblocks, hipStream_t s);

// Reduce `nparts` FieldPartial records with stride `stride` (records per
// block) starting at `first` into out[0].
hipError_t launch_finalize_partials(const FieldPartial *part, int nparts, int stride, int first,
                                    FieldPartial *out, hipStream_t s);
// Several record sets finalized in one launch (one block each), each as
// launch_finalize_partials does.
struct FinalizeSet {
  const FieldPartial *part;
  int nparts, stride, first;
  FieldPartial *out;
};
constexpr int kMaxFinalizeSets = 8;
hipError_t launch_finalize_sets(const FinalizeSet *sets, int n, hipStream_t s);
hipError_t launch_finalize_sums(const double *part, int nparts, int width, double *out,
                                hipStream_t s);

}  // namespace o3d
