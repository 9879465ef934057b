// Shared device helpers for the osinco3d MI355X kernels (gfx950, wave64).
//
// Field layout everywhere: the reference's column-major (nx, ny, nz) array,
// x fastest, 0-based here: f[i + nx*(j + ny*k)].  32-bit cell indices (the
// host entry points reject grids of 2^31 cells or more).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define O3D_WAVE 64

namespace o3d {

// DIRICHLET: x-Dirichlet inflow / outflow (x axis only; no counterpart in
// the reference, which stops on it, src/initialization.f90:238-241).
enum Bc : int { PERIODIC = 0, FREE_SLIP = 1, DIRICHLET = 2 };

struct Grid {
  int nx, ny, nz;
  int bc[3];  // per axis, PERIODIC or FREE_SLIP
};

// Ghost index along one axis of n nodes for q in [-n+1, 2n-2].
// Periodic: the reference wraps node n+1 to node 1 (src/derivation.f90:34-56),
// i.e. period n.  Free-slip: mirror about the end node (ghost f(1-m) = f(1+m),
// f(n+m) = f(n-m)); `flip` reports that the mirrored value changes sign for a
// field that is odd along this axis (the "i" variants, src/derivation.f90:111-163).
__device__ __forceinline__ int ghost_index(int q, int n, int bc, bool &flip) {
  flip = false;
  if (bc == PERIODIC) {
    if (q < 0) q += n;
    else if (q >= n) q -= n;
    return q;
  }
  if (q < 0) { flip = true; return -q; }
  if (q > n - 1) { flip = true; return 2 * (n - 1) - q; }
  return q;
}

// x-Dirichlet ends: the ghost m nodes beyond an end (m = 1, 2, 3) is the
// cubic through the four nodes nearest that end (f0 at the end, f3 three
// nodes in), evaluated there and summed in this order, so every stencil
// that reaches past the end is exact on cubics.
template <typename T>
__device__ __forceinline__ T xdir_extrap(int m, T f0, T f1, T f2, T f3) {
  if (m == 1) return ((T(4) * f0 + T(-6) * f1) + T(4) * f2) + T(-1) * f3;
  if (m == 2) return ((T(10) * f0 + T(-20) * f1) + T(15) * f2) + T(-4) * f3;
  return ((T(20) * f0 + T(-45) * f1) + T(36) * f2) + T(-10) * f3;
}

// Fill column c of an LDS row whose column c holds x = x0 + c - 3, when that
// x lies 1..3 nodes beyond an x-Dirichlet end, from the row's nodes nearest
// the end (interior columns, which the fill never writes).
template <typename T>
__device__ __forceinline__ void xdir_fill(T *row, int c, int x0, int nx) {
  const int x = x0 + c - 3;
  if (x < 0 && x >= -3) {
    const int e = 3 - x0;
    row[c] = xdir_extrap<T>(-x, row[e], row[e + 1], row[e + 2], row[e + 3]);
  } else if (x >= nx && x <= nx + 2) {
    const int e = nx + 2 - x0;
    row[c] = xdir_extrap<T>(x - (nx - 1), row[e], row[e - 1], row[e - 2], row[e - 3]);
  }
}

template <typename T>
__device__ __forceinline__ T load_ghost(const T *__restrict__ f, int idx, bool neg) {
  T v = f[idx];
  return neg ? -v : v;
}

// XCD-aware tile order: the hardware hands consecutive workgroups to the 8
// XCDs in turn, so block b runs on XCD b % 8.  Give each XCD a contiguous run
// of tiles (x fastest, then y, then z) so that tiles sharing halo rows share
// that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  if ((nb & 7) != 0) return b;
  return (b & 7) * (nb >> 3) + (b >> 3);
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, O3D_WAVE);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, O3D_WAVE));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, O3D_WAVE));
  return v;
}

// Reduction ops for (value, index) argmax with the reference's tie rule:
// function_stats keeps the first maximum in column-major order (strict >,
// src/functions.f90:47-52), so equal values resolve to the smaller index.
__device__ __forceinline__ void argmax_merge(double &v, int &ix, double v2, int ix2) {
  if (v2 > v || (v2 == v && ix2 < ix)) { v = v2; ix = ix2; }
}
__device__ __forceinline__ void wave_argmax(double &v, int &ix) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double v2 = __shfl_xor(v, o, O3D_WAVE);
    int i2 = __shfl_xor(ix, o, O3D_WAVE);
    argmax_merge(v, ix, v2, i2);
  }
}

// Block-wide reductions through LDS; every thread gets the result.
// `scratch` must hold (blockDim/64) doubles (+ ints for argmax).
template <int NWAVES>
__device__ __forceinline__ double block_sum(double v, double *scratch) {
  v = wave_sum(v);
  const int w = threadIdx.x / O3D_WAVE, l = threadIdx.x % O3D_WAVE;
  __syncthreads();
  if (l == 0) scratch[w] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int q = 0; q < NWAVES; ++q) r += scratch[q];
  return r;
}
template <int NWAVES>
__device__ __forceinline__ double block_max(double v, double *scratch) {
  v = wave_max(v);
  const int w = threadIdx.x / O3D_WAVE, l = threadIdx.x % O3D_WAVE;
  __syncthreads();
  if (l == 0) scratch[w] = v;
  __syncthreads();
  double r = scratch[0];
#pragma unroll
  for (int q = 1; q < NWAVES; ++q) r = fmax(r, scratch[q]);
  return r;
}
template <int NWAVES>
__device__ __forceinline__ double block_min(double v, double *scratch) {
  v = wave_min(v);
  const int w = threadIdx.x / O3D_WAVE, l = threadIdx.x % O3D_WAVE;
  __syncthreads();
  if (l == 0) scratch[w] = v;
  __syncthreads();
  double r = scratch[0];
#pragma unroll
  for (int q = 1; q < NWAVES; ++q) r = fmin(r, scratch[q]);
  return r;
}
template <int NWAVES>
__device__ __forceinline__ void block_argmax(double &v, int &ix, double *sv, int *si) {
  wave_argmax(v, ix);
  const int w = threadIdx.x / O3D_WAVE, l = threadIdx.x % O3D_WAVE;
  __syncthreads();
  if (l == 0) { sv[w] = v; si[w] = ix; }
  __syncthreads();
  v = sv[0]; ix = si[0];
#pragma unroll
  for (int q = 1; q < NWAVES; ++q) argmax_merge(v, ix, sv[q], si[q]);
}

// Per-block partial record for field statistics (function_stats, the
// velocity min/max prints, CFL and the NaN guard).
struct FieldPartial {
  double vmin, vmax, sum, amax;  // amax = max |f|
  int imax;                      // linear index of the first maximum
  int nan;                       // any NaN seen
};

}  // namespace o3d
