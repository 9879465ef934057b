// Host-side interface of the red-black SOR kernels (o3d_sor.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "o3d_common.h"

namespace o3d {

// Device-resident solver state.  Iteration `it` (1-based) is one launch; it
// decides from the state written by earlier launches whether the reference's
// stop rule has fired (src/poisson.f90:110-114), so the host can queue many
// launches without reading anything back.
struct SorState {
  unsigned long long dmax_bits[8];  // IEEE bits of dmax of iteration it, ring by it & 7
  double omega[8];                  // omega used by iteration it, ring by it & 7
  double omega_final;               // omega after the solve (persists, idyn = 1)
  double dmax_final;                // dmax of the last sweep
  int done;                         // stop rule fired
  int exit_iter;                    // the reference's printed exit iteration
  int reason;                       // 0 eps, 1 stagnation, 2 kmax
  int sweeps;                       // sweeps applied to the field
  int base;                         // iteration offset for graph-replayed batches
  int pad;
};

// The three neighbour rules of src/poisson.f90 (chosen from x and y
// boundary types only, src/initialization.f90:283-301), and the x-Dirichlet
// kinds SOR_XDIR + (y mirrored) + 2 (z mirrored), which have no counterpart
// in the reference: with the velocity imposed on the x ends the pressure is
// mirrored there (homogeneous Neumann), y and z follow their own types.
enum SolverKind : int { SOR_0000 = 0, SOR_0011 = 1, SOR_111111 = 2, SOR_XDIR = 3,
                        SOR_KINDS = 7 };
inline void solver_periodic(int kind, int per[3]) {
  if (kind >= SOR_XDIR) {
    per[0] = 0;
    per[1] = !((kind - SOR_XDIR) & 1);
    per[2] = !((kind - SOR_XDIR) & 2);
    return;
  }
  per[0] = kind != SOR_111111;
  per[1] = kind == SOR_0000;
  per[2] = kind != SOR_111111;
}

struct SorTiling {
  int ntx, nty, ntz, zc;
  int blocks() const { return ntx * nty * ntz; }
};
SorTiling make_sor_tiling(int nx, int ny, int nz, int slots, int zmult);
// Concurrent blocks of the sweep kernel on this device (occupancy x CUs).
template <typename T, typename TR>
int sor_slots();
// The tiling sor_kernel<T, TR> needs: z chunks a multiple of its ring depth.
template <typename T, typename TR>
SorTiling sor_tiling(int nx, int ny, int nz);

template <typename T, typename TR>
struct SorArgs {
  const T *pin;  // iterate before the sweep (not modified)
  T *pout;       // iterate after the sweep
  const TR *rhs;
  SorState *st;
  double *log;   // dmax per iteration (may be null), log[it-1]
  int log_len;
  int nx, ny, nz, per[3];
  SorTiling tl;
  double ox, oy, oz, A;  // 1/dx^2, 1/dy^2, 1/dz^2, -(2/dx^2+2/dy^2+2/dz^2)
  double eps, factor;
  int kmax, idyn;
  int it;        // iteration number of this launch (added to st->base)
  int gs;        // iterations 1..gs use omega = 1 (Gauss-Seidel smoothing) and
                 // end the solve only on kmax
  // Optional (fused kernel, T = double): a float32 correction in colour-split
  // storage added to pin as it is loaded (pin + e is the iterate the sweep
  // starts from; pin itself is not modified), nx even.
  const float *efold = nullptr;
  // Optional (T = double): psum[tile] = the FP64 sum of the new iterate over
  // the nodes the tile owns, in a fixed order (tile = the XCD-remapped block
  // index, tl.blocks() entries).
  double *psum = nullptr;
};

template <typename T, typename TR>
hipError_t launch_sor_iteration(const SorArgs<T, TR> &a, hipStream_t s);
// The fused SOR kernel addresses the iterate and the rhs through raw buffer
// descriptors, whose record counts and offsets are 32-bit byte counts: every
// field must fit in INT_MAX bytes.
template <typename T, typename TR>
inline bool sor_buffer_fits(int nx, int ny, int nz) {
  const long long bytes = (long long)nx * ny * nz *
                          (long long)(sizeof(T) > sizeof(TR) ? sizeof(T) : sizeof(TR));
  return bytes <= 0x7fffffffLL;
}
hipError_t launch_sor_init(SorState *st, double omega, int base, hipStream_t s);

// Colour-split sweeps: the iterate and the rhs stored per row as the red
// nodes (i + j + k even) then the black ones, each compressed to nx / 2
// entries; an iteration is a red and a black in-place launch with the point
// update of sor_kernel (identical iterates).  dmax partials go to dslot
// (sor_split_slot_words() words, zeroed by launch_sor_split_init).
template <typename T>
struct SorSplitArgs {
  T *red, *black;                   // iterate, updated in place
  const T *rhs_red, *rhs_black;
  SorState *st;
  unsigned long long *dslot;
  double *log;  // dmax per iteration (may be null), log[it-1]
  int log_len;
  int nx, ny, nz, per[3];
  double ox, oy, oz, A;
  double eps, factor;
  int kmax, idyn;
  int it;       // iteration number of this launch pair (added to st->base)
  int gs;       // iterations 1..gs use omega = 1 (Gauss-Seidel smoothing) and
                // end the solve only on kmax
};
// nx a multiple of 8 (FP32) or 4 (FP64): 16-byte vectors of one colour.
template <typename T>
bool sor_split_supported(int nx, int ny, int nz);
size_t sor_split_slot_words();
hipError_t launch_sor_split_init(unsigned long long *dslot, hipStream_t s);
template <typename T>
hipError_t launch_sor_split_iteration(const SorSplitArgs<T> &a, hipStream_t s);
template <typename T>
hipError_t launch_split(const T *f, T *red, T *black, int nx, int ny, int nz, hipStream_t s);
template <typename T>
hipError_t launch_unsplit(T *f, const T *red, const T *black, int nx, int ny, int nz,
                          hipStream_t s);

// Mixed precision helpers.  r = rhs - L p (FP64 arithmetic) stored in FP32,
// and e = 0; and p += e.
template <typename TR>
hipError_t launch_residual_f32(const double *p, const TR *rhs, float *r, float *e, int nx,
                               int ny, int nz, const int per[3], double ox, double oy,
                               double oz, double A, hipStream_t s);
hipError_t launch_fold(double *p, const float *e, int n, hipStream_t s);
// q = p + e and r = rhs - L q (FP32) in one pass (the extrapolated guess and
// its residual), equal to forming q and then calling launch_residual_f32.
// With part0 (2 * residual_blocks doubles), also the per-block (max |r0|,
// sum r0^2) of r0 = rhs - L p, for launch_residual_norm_pair without p0.
// q = p + e1 + beta (e1 - e2) and r = rhs - L q (FP32) in one pass, with
// part0 as launch_guess_residual; sums[0..1] (device) receive
// (sum (r + beta Ld) Ld, sum Ld Ld), d = e1 - e2, whose ratio is the
// least-squares beta of this rhs; partb holds 2 * residual_blocks doubles.
template <typename TR>
hipError_t launch_guess_adapt(const double *p, const float *e1, const float *e2, const TR *rhs,
                              double beta, double *q, float *r, int nx, int ny, int nz,
                              const int per[3], double ox, double oy, double oz, double A,
                              hipStream_t s, double *part0, double *partb, double *sums);
template <typename TR>
hipError_t launch_guess_residual(const double *p, const float *e, const TR *rhs, double *q,
                                 float *r, int nx, int ny, int nz, const int per[3], double ox,
                                 double oy, double oz, double A, hipStream_t s,
                                 double *part0 = nullptr);
// The same with r and e colour-split (red half-array, then black).
template <typename TR>
hipError_t launch_residual_split(const double *p, const TR *rhs, float *r, float *e, int nx,
                                 int ny, int nz, const int per[3], double ox, double oy,
                                 double oz, double A, hipStream_t s);
hipError_t launch_fold_split(double *p, const float *e, int nx, int ny, int nz, hipStream_t s);
// FP64 residual norms of two iterates p0, p1 (r = rhs - L p) in one launch:
// out (device, 4 doubles) = (max |r0|, sum r0^2, max |r1|, sum r1^2).  part
// holds 4 * residual_blocks(nx, ny, nz) doubles.  With r1s (nx even), the
// residual of p1 is also stored colour-split in FP32 and e1s is zeroed, as
// launch_residual_split would do for p1.  With p0 null, the first pair
// comes from the part0 records launch_guess_residual wrote for the same p0
// and rhs (bit-identical to computing them here).
template <typename TR>
hipError_t launch_residual_norm_pair(const double *p0, const double *p1, const TR *rhs,
                                     double *part, double *out, int nx, int ny, int nz,
                                     const int per[3], double ox, double oy, double oz, double A,
                                     hipStream_t s, float *r1s = nullptr,
                                     float *e1s = nullptr, const double *part0 = nullptr);
int residual_blocks(int nx, int ny, int nz);

template <typename T>
hipError_t launch_copy(T *dst, const T *src, int n, hipStream_t s);
template <typename TD, typename TS>
hipError_t launch_convert(TD *dst, const TS *src, int n, hipStream_t s);
hipError_t launch_fill_f64(double *dst, double v, int n, hipStream_t s);
hipError_t launch_fill_f32(float *dst, float v, int n, hipStream_t s);

}  // namespace o3d
