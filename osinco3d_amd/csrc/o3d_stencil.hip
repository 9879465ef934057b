// Stencil kernels of the osinco3d time-step path for MI355X (gfx950).
//
// Design (memory-bound work, ~1.7 flop/B in FP64):
//  * z-marching 2.5D tiles: a 64x8 block (one wave per x-row) owns an x-y tile
//    and walks a chunk of z planes.  Each thread keeps a 7-deep register queue
//    of its own column per field (z-derivatives, 6th order needs k-3..k+3) and
//    the current plane is staged once in LDS with a 3-wide cross-shaped halo
//    for the x- and y-derivatives.  Each field value is read from HBM once per
//    chunk; halo re-reads come from L2 of the same XCD (xcd_tile).
//  * Boundaries are folded into the halo fill: periodic wrap and free-slip
//    mirror (with the sign flip of an odd field) are applied when a value is
//    loaded, so every node, boundary rows included, uses the interior formula.
//    With the reference's operation order this reproduces all 18 boundary
//    variants of src/derivation.f90 bit for bit (FP contraction is off).
//  * Several derivatives are fused per field read: the predict kernel computes
//    9 first and 9 second derivatives, Smagorinsky nu_t, convection, diffusion
//    and the Euler/AB2/AB3 update from one read of (u, v, w).
#include <float.h>
#include <limits.h>

#include <algorithm>
#include <type_traits>

#include "o3d_stencil.h"

namespace o3d {

Tiling make_tiling(int nx, int ny, int nz) {
  Tiling t;
  t.ntx = (nx + kTX - 1) / kTX;
  t.nty = (ny + kTY - 1) / kTY;
  const int layer = t.ntx * t.nty;
  int ntz = (2048 + layer - 1) / layer;  // aim at >= 2048 workgroups
  ntz = std::max(1, std::min(ntz, std::max(1, nz / 32)));  // chunks of >= 32 planes
  t.zc = (nz + ntz - 1) / ntz;
  t.ntz = (nz + t.zc - 1) / t.zc;
  return t;
}

// Tiling for a kernel with `slots` concurrently resident blocks on the
// device: the z chunk count that minimises rounds * (zc + 8) (a chunk
// re-reads 6 planes for its z stencil and prefetches 2 past its end), so the
// grid fills whole rounds of resident blocks.
Tiling make_tiling_slots(int nx, int ny, int nz, int slots) {
  Tiling t;
  t.ntx = (nx + kTX - 1) / kTX;
  t.nty = (ny + kTY - 1) / kTY;
  const int layer = t.ntx * t.nty;
  if (slots <= 0) return make_tiling(nx, ny, nz);
  double best = 1e30;
  int best_ntz = 1;
  for (int ntz = 1; ntz <= std::max(1, nz / 8); ++ntz) {
    const int zc = (nz + ntz - 1) / ntz;
    const int n_t = (nz + zc - 1) / zc;
    const int rounds = (layer * n_t + slots - 1) / slots;
    const double cost = (double)rounds * (zc + 8.0);
    if (cost < best - 1e-9) { best = cost; best_ntz = ntz; }
  }
  t.zc = (nz + best_ntz - 1) / best_ntz;
  t.ntz = (nz + t.zc - 1) / t.zc;
  return t;
}

// Resident blocks of `kernel` (kTX*kTY threads) over the device, or 0.
static int stencil_slots(const void *kernel) {
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kTX * kTY, 0) !=
          hipSuccess ||
      hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return per_cu * cus;
}

// Ghost mapping that stays in bounds for threads of a partial tile that lie
// more than 3 nodes beyond the domain (their values are never used).
__device__ __forceinline__ int ghost_safe(int q, int n, int bc, bool &flip) {
  if (q > n + 2) { flip = false; return n - 1; }
  return ghost_index(q, n, bc, flip);
}

struct TilePos {
  int x0, y0, z0, z1, tx, ty, i, j, gi, gj;
  bool fi, fj, valid;
};

__device__ __forceinline__ TilePos tile_pos(const Tiling &tl, int nx, int ny, int nz,
                                            const int *bc) {
  TilePos p;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tix = t % tl.ntx;
  const int r = t / tl.ntx;
  const int tiy = r % tl.nty;
  const int tiz = r / tl.nty;
  p.tx = threadIdx.x % kTX;
  p.ty = threadIdx.x / kTX;
  p.x0 = tix * kTX;
  p.y0 = tiy * kTY;
  p.z0 = tiz * tl.zc;
  p.z1 = min(p.z0 + tl.zc, nz);
  p.i = p.x0 + p.tx;
  p.j = p.y0 + p.ty;
  p.valid = p.i < nx && p.j < ny;
  p.gi = ghost_safe(p.i, nx, bc[0], p.fi);
  p.gj = ghost_safe(p.j, ny, bc[1], p.fj);
  return p;
}

// 6th-order first derivative from 7 consecutive values v(-3..3), in the
// reference's order a*(f+3 - f-3) - b*(f+2 - f-2) + c*(f+1 - f-1).
template <typename T>
__device__ __forceinline__ T d1(T m3, T m2, T m1, T p1, T p2, T p3, const T *c) {
  return c[0] * (p3 - m3) - c[1] * (p2 - m2) + c[2] * (p1 - m1);
}
// 4th-order second derivative: -a*(f-2 + f+2) + b*(f-1 + f+1) - c*f.
// FP64 keeps the reference's expression (bit-identical results).  In FP32
// the rounded coefficients no longer satisfy c = 2b - 2a exactly, so that
// form would give a constant field a second derivative of ~1e-7 c f, a
// bias that accumulates over thousands of steps; the FP32 form uses the
// equivalent differences -a((f-2 - f) + (f+2 - f)) + b((f-1 - f) + (f+1 - f)),
// which vanish exactly on constants and do not cancel catastrophically.
template <typename T>
__device__ __forceinline__ T d2(T m2, T m1, T z, T p1, T p2, const T *c) {
  if constexpr (sizeof(T) == 4)
    return -c[0] * ((m2 - z) + (p2 - z)) + c[1] * ((m1 - z) + (p1 - z));
  else
    return -c[0] * (m2 + p2) + c[1] * (m1 + p1) - c[2] * z;
}

// ------------------------------------------------------------ generic D1/D2
// One routine of src/derivation.f90 per launch (the procedure-pointer
// drop-ins).  which = axis*6 + (order-1)*3 + variant, variant 0 = _00,
// 1 = p_11 (even), 2 = i_11 (odd); 18/19 = the 2D stubs derz_2dsim and
// derzz_2dsim (src/derivation.f90:481-495, 934-948).
template <typename T, int AXIS, int ORDER>
__global__ void __launch_bounds__(256) der_kernel(T *__restrict__ df, const T *__restrict__ f,
                                                  int nx, int ny, int nz, int variant, T ca,
                                                  T cb, T cc) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = nx * ny * nz;
  if (idx >= N) return;
  const int i = idx % nx, j = (idx / nx) % ny, k = idx / (nx * ny);
  const int n = AXIS == 0 ? nx : (AXIS == 1 ? ny : nz);
  const int p = AXIS == 0 ? i : (AXIS == 1 ? j : k);
  const int stride = AXIS == 0 ? 1 : (AXIS == 1 ? nx : nx * ny);
  const int base = idx - p * stride;
  const int bc = variant == 0 ? PERIODIC : FREE_SLIP;
  const bool odd = variant == 2;
  auto F = [&](int s) -> T {
    if (variant == 3) {  // x-Dirichlet: extrapolated ends
      const int q = p + s;
      if (q < 0)
        return xdir_extrap<T>(-q, f[base], f[base + stride], f[base + 2 * stride],
                              f[base + 3 * stride]);
      if (q > n - 1) {
        const int e = base + (n - 1) * stride;
        return xdir_extrap<T>(q - (n - 1), f[e], f[e - stride], f[e - 2 * stride],
                              f[e - 3 * stride]);
      }
      return f[base + q * stride];
    }
    bool fl;
    const int q = ghost_index(p + s, n, bc, fl);
    const T v = f[base + q * stride];
    return (fl && odd) ? -v : v;
  };
  const T c[3] = {ca, cb, cc};
  if (ORDER == 1) {
    if (variant == 1 && (p == 0 || p == n - 1)) { df[idx] = T(0); return; }
    df[idx] = d1<T>(F(-3), F(-2), F(-1), F(1), F(2), F(3), c);
  } else {
    df[idx] = d2<T>(F(-2), F(-1), f[idx], F(1), F(2), c);
  }
}

template <typename T>
__global__ void der2dsim_kernel(T *__restrict__ df, const T *__restrict__ f, int N, T h) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < N) df[idx] = T(0) * f[idx] * h;
}

template <typename T>
hipError_t launch_der(int which, T *df, const T *f, int nx, int ny, int nz, double h,
                      hipStream_t s) {
  const int N = nx * ny * nz;
  const dim3 blk(256), grd((N + 255) / 256);
  if (which == 18 || which == 19) {
    der2dsim_kernel<T><<<grd, blk, 0, s>>>(df, f, N, T(h));
    return hipGetLastError();
  }
  if (which < 0 || which > 21) return hipErrorInvalidValue;
  // 20 / 21: the x first / second derivative with x-Dirichlet ends.
  const int axis = which >= 20 ? 0 : which / 6;
  const int order = which >= 20 ? which - 19 : (which % 6) / 3 + 1;
  const int variant = which >= 20 ? 3 : which % 3;
  double a, b, c;
  if (order == 1) { D1Coef k = d1coef(h); a = k.a; b = k.b; c = k.c; }
  else { D2Coef k = d2coef(h); a = k.a; b = k.b; c = k.c; }
#define O3D_DER(AX, OR)                                                                   \
  der_kernel<T, AX, OR><<<grd, blk, 0, s>>>(df, f, nx, ny, nz, variant, T(a), T(b), T(c))
  if (axis == 0) { if (order == 1) O3D_DER(0, 1); else O3D_DER(0, 2); }
  else if (axis == 1) { if (order == 1) O3D_DER(1, 1); else O3D_DER(1, 2); }
  else { if (order == 1) O3D_DER(2, 1); else O3D_DER(2, 2); }
#undef O3D_DER
  return hipGetLastError();
}

// ------------------------------------------------------------ block partials
__device__ __forceinline__ void partial_init(FieldPartial &p) {
  p.vmin = DBL_MAX;
  p.vmax = -DBL_MAX;
  p.sum = 0.0;
  p.amax = 0.0;
  p.imax = INT_MAX;
  p.nan = 0;
}
__device__ __forceinline__ void partial_add(FieldPartial &p, double v, int idx) {
  if (v != v) { p.nan = 1; return; }
  p.sum += v;
  if (v < p.vmin) p.vmin = v;
  if (v > p.vmax) { p.vmax = v; p.imax = idx; }
  p.amax = fmax(p.amax, fabs(v));
}

constexpr int kNW = kTX * kTY / O3D_WAVE;  // waves per stencil block

struct ReduceScratch {
  double v[kNW];
  int ix[kNW];
};

__device__ __forceinline__ void partial_store(FieldPartial p, FieldPartial *out,
                                              ReduceScratch &sc) {
  const double vmin = block_min<kNW>(p.vmin, sc.v);
  const double vmax0 = p.vmax;
  double vmax = vmax0;
  int imax = p.imax;
  block_argmax<kNW>(vmax, imax, sc.v, sc.ix);
  const double sum = block_sum<kNW>(p.sum, sc.v);
  const double amax = block_max<kNW>(p.amax, sc.v);
  const double nan = block_max<kNW>((double)p.nan, sc.v);
  if (threadIdx.x == 0) {
    FieldPartial r;
    r.vmin = vmin; r.vmax = vmax; r.sum = sum; r.amax = amax; r.imax = imax;
    r.nan = nan > 0.0 ? 1 : 0;
    *out = r;
  }
}

// ------------------------------------------------------------------ predict
// src/integration.f90:14-197 (+ src/les_turbulence.f90:10-97 when LES).
template <typename T, int NH, bool LES>
__global__ void __launch_bounds__(kTX *kTY) predict_kernel(const PredictArgs<T> A) {
  __shared__ T pl[3][kTY + 6][kTX + 6];
  __shared__ ReduceScratch sc;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const TilePos P = tile_pos(A.tl, nx, ny, nz, A.bc);
  const int plane = nx * ny;
  const int col = P.gi + nx * P.gj;
  const int tx = P.tx, ty = P.ty;
  // u_c is odd along axis c (parity table of src/integration.f90:118-165).
  const bool own_neg[2] = {P.fi, P.fj};

  // Halo duties of this thread: at most one x-halo and one y-halo cell.
  int hx_col = -1, hx_g = 0;
  bool hx_f = false;
  if (tx < 3) { hx_col = tx; hx_g = ghost_safe(P.x0 + tx - 3, nx, A.bc[0], hx_f); }
  else if (tx >= kTX - 3) { hx_col = tx + 6; hx_g = ghost_safe(P.x0 + tx + 3, nx, A.bc[0], hx_f); }
  int hy_row = -1, hy_g = 0;
  bool hy_f = false;
  if (ty < 3) { hy_row = ty; hy_g = ghost_safe(P.y0 + ty - 3, ny, A.bc[1], hy_f); }
  else if (ty >= kTY - 3) { hy_row = ty + 6; hy_g = ghost_safe(P.y0 + ty + 3, ny, A.bc[1], hy_f); }
  const int hx_idx = hx_g + nx * P.gj;
  const int hy_idx = P.gi + nx * hy_g;

  // Forced zeros of the even free-slip first derivative (src/derivation.f90:87,105).
  const bool edge_x = A.bc[0] == FREE_SLIP && (P.i == 0 || P.i == nx - 1);
  const bool edge_y = A.bc[1] == FREE_SLIP && (P.j == 0 || P.j == ny - 1);
  // Tiles whose plane reaches past an x-Dirichlet end (block-uniform).
  const bool xdir = A.bc[0] == DIRICHLET && (P.x0 == 0 || P.x0 + kTX + 2 >= nx);

  FieldPartial nut_p;
  partial_init(nut_p);

  // Column queue of planes k-3..k+3 per field (signs applied), and raw
  // loads in flight in two-slot rings whose slots are compile-time
  // constants of the unrolled pair of steps: plane k+5 of the queue is
  // loaded at step k and joins it at the end of step k+1; the x/y halo
  // values of plane k+2 are loaded at step k (every lane loads, lanes
  // without a halo cell their own column) and staged, with their signs, at
  // step k+2.  No register that a load is still writing is read early.
  auto zidx = [&](int qz, bool &fz) { return col + plane * ghost_index(qz, nz, A.bc[2], fz); };
  auto sgn = [](T v, bool neg) { return neg ? -v : v; };
  T q[3][7];
#pragma unroll
  for (int s = 0; s < 7; ++s) {
    bool fz;
    const int ix = zidx(P.z0 - 3 + s, fz);
    q[0][s] = load_ghost(A.u[0], ix, own_neg[0]);
    q[1][s] = load_ghost(A.u[1], ix, own_neg[1]);
    q[2][s] = load_ghost(A.u[2], ix, fz);
  }
  T zr[3][2];
  {
    bool fz;
    const int ix = zidx(P.z0 + 4, fz);
#pragma unroll
    for (int c = 0; c < 3; ++c) zr[c][1] = A.u[c][ix];
  }
  const int hxi = hx_col >= 0 ? hx_idx : col, hyi = hy_row >= 0 ? hy_idx : col;
  // Signs of the halo values: x halo u0 flips with the x mirror, u1 with
  // this column's y mirror; y halo u0 with this column's x mirror, u1 with
  // the y mirror; u2 never (z parity only).
  const bool hneg_x[3] = {hx_col >= 0 && hx_f, own_neg[1], false};
  const bool hneg_y[3] = {own_neg[0], hy_row >= 0 && hy_f, false};
  T hxr[3][2], hyr[3][2];
  auto load_halo = [&](int kp, int sl) {
    const int pk = plane * min(kp, nz - 1);  // planes after the last are not used
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      hxr[c][sl] = A.u[c][hxi + pk];
      hyr[c][sl] = A.u[c][hyi + pk];
    }
  };
  load_halo(P.z0, 0);
  load_halo(P.z0 + 1, 1);
  const T *c1x = A.d1[0], *c1y = A.d1[1], *c1z = A.d1[2];
  const T *c2x = A.d2[0], *c2y = A.d2[1], *c2z = A.d2[2];

  auto step = [&](int k, auto SL) {
    constexpr int sl = decltype(SL)::value;
    const int pk = plane * k;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      pl[c][ty + 3][tx + 3] = q[c][3];
      if (hx_col >= 0) pl[c][ty + 3][hx_col] = sgn(hxr[c][sl], hneg_x[c]);
      if (hy_row >= 0) pl[c][hy_row][tx + 3] = sgn(hyr[c][sl], hneg_y[c]);
    }
    __syncthreads();
    if (xdir) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        xdir_fill<T>(&pl[c][ty + 3][0], tx + 3, P.x0, nx);
        if (hx_col >= 0) xdir_fill<T>(&pl[c][ty + 3][0], hx_col, P.x0, nx);
      }
      __syncthreads();
    }
    if (P.valid) {
      const int idx = P.i + nx * P.j + pk;
      const int r = ty + 3, cc = tx + 3;
      // History loads issued before the stencil arithmetic that hides them.
      T h2[3], h3[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (NH >= 2) h2[c] = A.f2[c][idx];
        if (NH >= 3) h3[c] = A.f3[c][idx];
      }
      T g[3][3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        g[c][0] = d1<T>(pl[c][r][cc - 3], pl[c][r][cc - 2], pl[c][r][cc - 1], pl[c][r][cc + 1],
                        pl[c][r][cc + 2], pl[c][r][cc + 3], c1x);
        g[c][1] = d1<T>(pl[c][r - 3][cc], pl[c][r - 2][cc], pl[c][r - 1][cc], pl[c][r + 1][cc],
                        pl[c][r + 2][cc], pl[c][r + 3][cc], c1y);
        g[c][2] = d1<T>(q[c][0], q[c][1], q[c][2], q[c][4], q[c][5], q[c][6], c1z);
      }
      if (edge_x) { g[1][0] = T(0); g[2][0] = T(0); }
      if (edge_y) { g[0][1] = T(0); g[2][1] = T(0); }
      if (A.bc[2] == FREE_SLIP && (k == 0 || k == nz - 1)) { g[0][2] = T(0); g[1][2] = T(0); }
      T nu_eff = A.onere;
      if (LES) {
        // src/les_turbulence.f90:74-85
        const T s11 = g[0][0], s22 = g[1][1], s33 = g[2][2];
        const T s12 = T(0.5) * (g[0][1] + g[1][0]);
        const T s13 = T(0.5) * (g[0][2] + g[2][0]);
        const T s23 = T(0.5) * (g[1][2] + g[2][1]);
        const T smag = sqrt(T(2) * (s11 * s11 + s22 * s22 + s33 * s33 +
                                    T(2) * (s12 * s12 + s13 * s13 + s23 * s23)));
        const T nut = A.csd2 * smag;
        nu_eff = A.onere + nut;
        if (A.nut) A.nut[idx] = nut;
        partial_add(nut_p, (double)nut, idx);
      }
      const T ux = q[0][3], uy = q[1][3], uz = q[2][3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T lx = d2<T>(pl[c][r][cc - 2], pl[c][r][cc - 1], pl[c][r][cc], pl[c][r][cc + 1],
                           pl[c][r][cc + 2], c2x);
        const T ly = d2<T>(pl[c][r - 2][cc], pl[c][r - 1][cc], pl[c][r][cc], pl[c][r + 1][cc],
                           pl[c][r + 2][cc], c2y);
        const T lz = d2<T>(q[c][1], q[c][2], q[c][3], q[c][4], q[c][5], c2z);
        // src/integration.f90:129-134 (and :149-154, :169-174)
        const T F = nu_eff * (lx + ly + lz) - (ux * g[c][0] + uy * g[c][1] + uz * g[c][2]);
        T un = q[c][3] + A.adu * F;
        if (NH >= 2) un = un + A.bdu * h2[c];
        if (NH >= 3) un = un + A.cdu * h3[c];
        A.f1[c][idx] = F;
        A.up[c][idx] = un;
      }
    }
    {
      // k + 5 <= nz + 4 stays a valid mirror / wrap index for nz >= 6.
      bool fz;
      const int ix = zidx(k + 5, fz);
#pragma unroll
      for (int c = 0; c < 3; ++c) zr[c][sl] = A.u[c][ix];
    }
    load_halo(k + 2, sl);
    __syncthreads();
    bool fz4;
    (void)zidx(k + 4, fz4);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int s = 0; s < 6; ++s) q[c][s] = q[c][s + 1];
    }
    q[0][6] = sgn(zr[0][1 - sl], own_neg[0]);
    q[1][6] = sgn(zr[1][1 - sl], own_neg[1]);
    q[2][6] = sgn(zr[2][1 - sl], fz4);
  };
  // Whole pairs in the loop (every back edge leaves slot 1's loads last in
  // flight), the odd last plane after it.
  int k = P.z0;
  for (; k + 1 < P.z1; k += 2) {
    step(k, std::integral_constant<int, 0>{});
    step(k + 1, std::integral_constant<int, 1>{});
  }
  if (k < P.z1) step(k, std::integral_constant<int, 0>{});
  if (LES) partial_store(nut_p, A.nut_part + blockIdx.x, sc);
}

template <typename T>
Tiling predict_tiling(int nx, int ny, int nz, bool les) {
  static const int slots[2] = {stencil_slots((const void *)predict_kernel<T, 3, false>),
                               stencil_slots((const void *)predict_kernel<T, 3, true>)};
  return make_tiling_slots(nx, ny, nz, slots[les ? 1 : 0]);
}

template <typename T>
hipError_t launch_predict(const PredictArgs<T> &a, int nh, bool les, hipStream_t s) {
  const dim3 blk(kTX * kTY), grd(a.tl.blocks());
#define O3D_PRED(NH, LES) predict_kernel<T, NH, LES><<<grd, blk, 0, s>>>(a)
  if (les) {
    if (nh == 1) O3D_PRED(1, true); else if (nh == 2) O3D_PRED(2, true); else O3D_PRED(3, true);
  } else {
    if (nh == 1) O3D_PRED(1, false); else if (nh == 2) O3D_PRED(2, false); else O3D_PRED(3, false);
  }
#undef O3D_PRED
  return hipGetLastError();
}

// --------------------------------------------------------------- divergence
// src/differential_operators.f90:7-38; with divisor = dt this is the
// Poisson right-hand side of src/integration.f90:235-239.  Statistics of the
// (undivided) divergence are reduced per block (function_stats).
template <typename T>
__global__ void __launch_bounds__(kTX *kTY) div_kernel(const DivArgs<T> A) {
  __shared__ T su[kTY][kTX + 6];
  __shared__ T sv[kTY + 6][kTX];
  __shared__ ReduceScratch sc;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const TilePos P = tile_pos(A.tl, nx, ny, nz, A.bc);
  const int plane = nx * ny;
  const int col = P.gi + nx * P.gj;
  const int tx = P.tx, ty = P.ty;
  const bool odd = A.odd != 0;
  int hx_col = -1, hx_g = 0;
  bool hx_f = false;
  if (tx < 3) { hx_col = tx; hx_g = ghost_safe(P.x0 + tx - 3, nx, A.bc[0], hx_f); }
  else if (tx >= kTX - 3) { hx_col = tx + 6; hx_g = ghost_safe(P.x0 + tx + 3, nx, A.bc[0], hx_f); }
  int hy_row = -1, hy_g = 0;
  bool hy_f = false;
  if (ty < 3) { hy_row = ty; hy_g = ghost_safe(P.y0 + ty - 3, ny, A.bc[1], hy_f); }
  else if (ty >= kTY - 3) { hy_row = ty + 6; hy_g = ghost_safe(P.y0 + ty + 3, ny, A.bc[1], hy_f); }
  const int hx_idx = hx_g + nx * P.gj;
  const int hy_idx = P.gi + nx * hy_g;
  const bool zero_x = !odd && A.bc[0] == FREE_SLIP && (P.i == 0 || P.i == nx - 1);
  const bool zero_y = !odd && A.bc[1] == FREE_SLIP && (P.j == 0 || P.j == ny - 1);
  const bool xdir = A.bc[0] == DIRICHLET && (P.x0 == 0 || P.x0 + kTX + 2 >= nx);

  FieldPartial st;
  partial_init(st);
  // w column queue (planes k-3..k+3, signs applied) and raw loads in flight
  // in a two-slot ring: the load of plane k+5 is issued at step k and moved
  // into the queue at the end of step k+1.  u / v values of plane k+2 are
  // loaded at step k and staged (with their sign) at step k+2.  Slots are
  // compile-time constants of the unrolled pair of steps, and no register
  // that a load is still writing is read before its use, so the loads of
  // two steps stay in flight across the barriers.
  const bool neg_u = odd && P.fi, neg_v = odd && P.fj;
  const bool neg_hx = odd && hx_f, neg_hy = odd && hy_f;
  auto sgn = [](T v, bool neg) { return neg ? -v : v; };
  auto wz = [&](int q, bool &neg) {
    bool fz;
    const int kk = ghost_index(q, nz, A.bc[2], fz);
    neg = odd && fz;
    return col + plane * kk;
  };
  T qw[7];
#pragma unroll
  for (int s = 0; s < 7; ++s) {
    bool ng;
    const int ix = wz(P.z0 - 3 + s, ng);
    qw[s] = load_ghost(A.f[2], ix, ng);
  }
  T wr[2];
  {
    bool ng;
    wr[1] = A.f[2][wz(P.z0 + 4, ng)];
  }
  T nu[2], nv[2], nhx[2] = {T(0), T(0)}, nhy[2] = {T(0), T(0)};
  // Every lane issues the halo loads (lanes without a halo cell load their
  // own column), so the number of loads in flight is the same on every path.
  const int hxi = hx_col >= 0 ? hx_idx : col, hyi = hy_row >= 0 ? hy_idx : col;
  auto load_uv = [&](int kp, int sl) {
    const int pk = plane * min(kp, nz - 1);  // planes after the last are not used
    nu[sl] = A.f[0][col + pk];
    nv[sl] = A.f[1][col + pk];
    nhx[sl] = A.f[0][hxi + pk];
    nhy[sl] = A.f[1][hyi + pk];
  };
  load_uv(P.z0, 0);
  load_uv(P.z0 + 1, 1);
  auto step = [&](int k, auto SL) {
    constexpr int sl = decltype(SL)::value;
    const int pk = plane * k;
    su[ty][tx + 3] = sgn(nu[sl], neg_u);
    sv[ty + 3][tx] = sgn(nv[sl], neg_v);
    if (hx_col >= 0) su[ty][hx_col] = sgn(nhx[sl], neg_hx);
    if (hy_row >= 0) sv[hy_row][tx] = sgn(nhy[sl], neg_hy);
    __syncthreads();
    if (xdir) {
      xdir_fill<T>(&su[ty][0], tx + 3, P.x0, nx);
      if (hx_col >= 0) xdir_fill<T>(&su[ty][0], hx_col, P.x0, nx);
      __syncthreads();
    }
    if (P.valid) {
      const int idx = P.i + nx * P.j + pk;
      const int cc = tx + 3, r = ty + 3;
      T dfx = d1<T>(su[ty][cc - 3], su[ty][cc - 2], su[ty][cc - 1], su[ty][cc + 1],
                    su[ty][cc + 2], su[ty][cc + 3], A.d1[0]);
      T dfy = d1<T>(sv[r - 3][tx], sv[r - 2][tx], sv[r - 1][tx], sv[r + 1][tx], sv[r + 2][tx],
                    sv[r + 3][tx], A.d1[1]);
      T dfz = d1<T>(qw[0], qw[1], qw[2], qw[4], qw[5], qw[6], A.d1[2]);
      if (zero_x) dfx = T(0);
      if (zero_y) dfy = T(0);
      if (!odd && A.bc[2] == FREE_SLIP && (k == 0 || k == nz - 1)) dfz = T(0);
      const T div = dfx + dfy + dfz;
      if (A.out) A.out[idx] = div / A.divisor;
      partial_add(st, (double)div, idx);
    }
    {
      // k + 5 <= nz + 4 stays a valid mirror / wrap index for nz >= 6.
      bool ng;
      wr[sl] = A.f[2][wz(k + 5, ng)];
    }
    load_uv(k + 2, sl);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 6; ++s) qw[s] = qw[s + 1];
    bool ng;
    (void)wz(k + 4, ng);
    qw[6] = sgn(wr[1 - sl], ng);
  };
  // Whole pairs in the loop (every back edge leaves slot 1's loads last in
  // flight), the odd last plane after it.
  int k = P.z0;
  for (; k + 1 < P.z1; k += 2) {
    step(k, std::integral_constant<int, 0>{});
    step(k + 1, std::integral_constant<int, 1>{});
  }
  if (k < P.z1) step(k, std::integral_constant<int, 0>{});
  partial_store(st, A.part + blockIdx.x, sc);
}

template <typename T>
Tiling div_tiling(int nx, int ny, int nz) {
  static const int slots = stencil_slots((const void *)div_kernel<T>);
  return make_tiling_slots(nx, ny, nz, slots);
}

template <typename T>
hipError_t launch_divergence(const DivArgs<T> &a, hipStream_t s) {
  div_kernel<T><<<dim3(a.tl.blocks()), dim3(kTX * kTY), 0, s>>>(a);
  return hipGetLastError();
}

// ------------------------------------------------------------------ correct
// src/integration.f90:257-330: grad p with the even ("p") variants, then
// u = u* - dt*grad p; the NaN / max>1000 guard and the driver's velocity
// min/max and CFL scans (src/IOfunctions.f90:322-330, src/utils.f90:178-208)
// are reduced in the same pass.
template <typename T, typename TP>
__global__ void __launch_bounds__(kTX *kTY) correct_kernel(const CorrectArgs<T, TP> A) {
  __shared__ TP pl[kTY + 6][kTX + 6];
  __shared__ ReduceScratch sc;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const TilePos P = tile_pos(A.tl, nx, ny, nz, A.bc);
  const int plane = nx * ny;
  const int col = P.gi + nx * P.gj;
  const int tx = P.tx, ty = P.ty;
  int hx_col = -1, hx_g = 0;
  bool hx_f = false;
  if (tx < 3) { hx_col = tx; hx_g = ghost_safe(P.x0 + tx - 3, nx, A.bc[0], hx_f); }
  else if (tx >= kTX - 3) { hx_col = tx + 6; hx_g = ghost_safe(P.x0 + tx + 3, nx, A.bc[0], hx_f); }
  int hy_row = -1, hy_g = 0;
  bool hy_f = false;
  if (ty < 3) { hy_row = ty; hy_g = ghost_safe(P.y0 + ty - 3, ny, A.bc[1], hy_f); }
  else if (ty >= kTY - 3) { hy_row = ty + 6; hy_g = ghost_safe(P.y0 + ty + 3, ny, A.bc[1], hy_f); }
  const int hx_idx = hx_g + nx * P.gj;
  const int hy_idx = P.gi + nx * hy_g;
  // The pressure is mirrored on an x-Dirichlet axis too (homogeneous
  // Neumann), and the inflow / outflow planes keep u* (no correction).
  const bool zero_x = A.bc[0] != PERIODIC && (P.i == 0 || P.i == nx - 1);
  const bool zero_y = A.bc[1] == FREE_SLIP && (P.j == 0 || P.j == ny - 1);
  const bool keep = A.bc[0] == DIRICHLET && (P.i == 0 || P.i == nx - 1);

  FieldPartial st[3];
  partial_init(st[0]); partial_init(st[1]); partial_init(st[2]);
  // The shift of every loaded pressure value (p - 0 = p exactly without one).
  const TP mv = A.mean ? (TP)A.mean[0] : TP(0);
  double es = 0.0;  // sum of e^2 over the owned nodes
  // p column queue (planes k-3..k+3) and raw loads in flight in two-slot
  // rings (slots fixed at compile time by the unrolled pair of steps): the
  // load of plane k+5 is issued at step k and joins the queue at the end of
  // step k+1; the halo and u* values of plane k+2 are loaded at step k and
  // used at step k+2.  No register that a load is still writing is read
  // before its use.
  auto pz = [&](int qz) {
    bool fz;
    return col + plane * ghost_index(qz, nz, A.bc[2], fz);
  };
  TP q[7];
#pragma unroll
  for (int s = 0; s < 7; ++s) q[s] = A.p[pz(P.z0 - 3 + s)] - mv;
  TP pr[2];
  pr[1] = A.p[pz(P.z0 + 4)];
  TP nhx[2] = {TP(0), TP(0)}, nhy[2] = {TP(0), TP(0)};
  // Every lane issues the halo loads (lanes without a halo cell load their
  // own column), so the number of loads in flight is the same on every path.
  const int hxi = hx_col >= 0 ? hx_idx : col, hyi = hy_row >= 0 ? hy_idx : col;
  auto load_halo = [&](int kp, int sl) {
    const int pk = plane * min(kp, nz - 1);  // planes after the last are not used
    nhx[sl] = A.p[hxi + pk];
    nhy[sl] = A.p[hyi + pk];
  };
  T upv[2][3];
  TP p0v[2] = {TP(0), TP(0)};
  const int own = P.valid ? P.i + nx * P.j : 0;
  auto load_up = [&](int kp, int sl) {
    const int idx = own + plane * min(kp, nz - 1);
#pragma unroll
    for (int c = 0; c < 3; ++c) upv[sl][c] = A.up[c][idx];
    if (A.p0) p0v[sl] = A.p0[idx];
  };
  load_halo(P.z0, 0);
  load_halo(P.z0 + 1, 1);
  load_up(P.z0, 0);
  load_up(P.z0 + 1, 1);
  auto step = [&](int k, auto SL) {
    constexpr int sl = decltype(SL)::value;
    const int pk = plane * k;
    pl[ty + 3][tx + 3] = q[3];
    if (hx_col >= 0) pl[ty + 3][hx_col] = nhx[sl] - mv;
    if (hy_row >= 0) pl[hy_row][tx + 3] = nhy[sl] - mv;
    __syncthreads();
    if (P.valid) {
      const int idx = P.i + nx * P.j + pk;
      const int cc = tx + 3, r = ty + 3;
      TP g[3];
      g[0] = d1<TP>(pl[r][cc - 3], pl[r][cc - 2], pl[r][cc - 1], pl[r][cc + 1], pl[r][cc + 2],
                    pl[r][cc + 3], A.d1[0]);
      g[1] = d1<TP>(pl[r - 3][cc], pl[r - 2][cc], pl[r - 1][cc], pl[r + 1][cc], pl[r + 2][cc],
                    pl[r + 3][cc], A.d1[1]);
      g[2] = d1<TP>(q[0], q[1], q[2], q[4], q[5], q[6], A.d1[2]);
      if (zero_x) g[0] = TP(0);
      if (zero_y) g[1] = TP(0);
      if (A.bc[2] == FREE_SLIP && (k == 0 || k == nz - 1)) g[2] = TP(0);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T un = keep ? upv[sl][c] : T(TP(upv[sl][c]) - A.dt * g[c]);
        A.u[c][idx] = un;
        partial_add(st[c], (double)un, idx);
      }
      if (A.pout) A.pout[idx] = q[3];
      if (A.p0) {
        const float ev = (float)(q[3] - p0v[sl]);
        A.e[idx] = ev;
        es += (double)ev * ev;
      }
    }
    // k + 5 <= nz + 4 stays a valid mirror / wrap index for nz >= 6.
    pr[sl] = A.p[pz(k + 5)];
    load_halo(k + 2, sl);
    load_up(k + 2, sl);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 6; ++s) q[s] = q[s + 1];
    q[6] = pr[1 - sl] - mv;
  };
  // Whole pairs in the loop (every back edge leaves slot 1's loads last in
  // flight), the odd last plane after it.
  int k = P.z0;
  for (; k + 1 < P.z1; k += 2) {
    step(k, std::integral_constant<int, 0>{});
    step(k + 1, std::integral_constant<int, 1>{});
  }
  if (k < P.z1) step(k, std::integral_constant<int, 0>{});
#pragma unroll
  for (int c = 0; c < 3; ++c) partial_store(st[c], A.part + 3 * blockIdx.x + c, sc);
  if (A.epart) {
    es = block_sum<kNW>(es, sc.v);
    if (threadIdx.x == 0) A.epart[blockIdx.x] = es;
  }
}

template <typename T, typename TP>
Tiling correct_tiling(int nx, int ny, int nz) {
  static const int slots = stencil_slots((const void *)correct_kernel<T, TP>);
  return make_tiling_slots(nx, ny, nz, slots);
}

template <typename T, typename TP>
hipError_t launch_correct(const CorrectArgs<T, TP> &a, hipStream_t s) {
  correct_kernel<T, TP><<<dim3(a.tl.blocks()), dim3(kTX * kTY), 0, s>>>(a);
  return hipGetLastError();
}

// --------------------------------------------------------------- statistics
// src/utils.f90:243-375: per block, the 16 sums behind E_k, eps, eps2,
// enstrophy and the mean squares (FP64 accumulation).
// Slot order: 0 e_k, 1 eps, 2 eps2, 3 dzeta, 4..6 u^2 v^2 w^2, 7..15 (du_c/dx_a)^2.
template <typename T>
__global__ void __launch_bounds__(kTX *kTY) stats_kernel(const StatsArgs<T> A) {
  __shared__ T pl[3][kTY + 6][kTX + 6];
  __shared__ ReduceScratch sc;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const TilePos P = tile_pos(A.tl, nx, ny, nz, A.bc);
  const int plane = nx * ny;
  const int col = P.gi + nx * P.gj;
  const int tx = P.tx, ty = P.ty;
  const bool own_neg[2] = {P.fi, P.fj};
  int hx_col = -1, hx_g = 0;
  bool hx_f = false;
  if (tx < 3) { hx_col = tx; hx_g = ghost_safe(P.x0 + tx - 3, nx, A.bc[0], hx_f); }
  else if (tx >= kTX - 3) { hx_col = tx + 6; hx_g = ghost_safe(P.x0 + tx + 3, nx, A.bc[0], hx_f); }
  int hy_row = -1, hy_g = 0;
  bool hy_f = false;
  if (ty < 3) { hy_row = ty; hy_g = ghost_safe(P.y0 + ty - 3, ny, A.bc[1], hy_f); }
  else if (ty >= kTY - 3) { hy_row = ty + 6; hy_g = ghost_safe(P.y0 + ty + 3, ny, A.bc[1], hy_f); }
  const int hx_idx = hx_g + nx * P.gj;
  const int hy_idx = P.gi + nx * hy_g;
  const bool edge_x = A.bc[0] == FREE_SLIP && (P.i == 0 || P.i == nx - 1);
  const bool edge_y = A.bc[1] == FREE_SLIP && (P.j == 0 || P.j == ny - 1);
  const bool xdir = A.bc[0] == DIRICHLET && (P.x0 == 0 || P.x0 + kTX + 2 >= nx);
  double acc[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) acc[m] = 0.0;
  T q[3][7];
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    bool fz;
    const int kk = ghost_index(P.z0 - 3 + s, nz, A.bc[2], fz);
    q[0][s] = load_ghost(A.u[0], col + plane * kk, own_neg[0]);
    q[1][s] = load_ghost(A.u[1], col + plane * kk, own_neg[1]);
    q[2][s] = load_ghost(A.u[2], col + plane * kk, fz);
  }
  for (int k = P.z0; k < P.z1; ++k) {
    {
      bool fz;
      const int kk = ghost_index(k + 3, nz, A.bc[2], fz);
      q[0][6] = load_ghost(A.u[0], col + plane * kk, own_neg[0]);
      q[1][6] = load_ghost(A.u[1], col + plane * kk, own_neg[1]);
      q[2][6] = load_ghost(A.u[2], col + plane * kk, fz);
    }
    const int pk = plane * k;
    T hx[3], hy[3];
    if (hx_col >= 0) {
      hx[0] = load_ghost(A.u[0], hx_idx + pk, hx_f);
      hx[1] = load_ghost(A.u[1], hx_idx + pk, own_neg[1]);
      hx[2] = A.u[2][hx_idx + pk];
    }
    if (hy_row >= 0) {
      hy[0] = load_ghost(A.u[0], hy_idx + pk, own_neg[0]);
      hy[1] = load_ghost(A.u[1], hy_idx + pk, hy_f);
      hy[2] = A.u[2][hy_idx + pk];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      pl[c][ty + 3][tx + 3] = q[c][3];
      if (hx_col >= 0) pl[c][ty + 3][hx_col] = hx[c];
      if (hy_row >= 0) pl[c][hy_row][tx + 3] = hy[c];
    }
    __syncthreads();
    if (xdir) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        xdir_fill<T>(&pl[c][ty + 3][0], tx + 3, P.x0, nx);
        if (hx_col >= 0) xdir_fill<T>(&pl[c][ty + 3][0], hx_col, P.x0, nx);
      }
      __syncthreads();
    }
    if (P.valid) {
      const int r = ty + 3, cc = tx + 3;
      double g[3][3], l[3][3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        g[c][0] = d1<T>(pl[c][r][cc - 3], pl[c][r][cc - 2], pl[c][r][cc - 1], pl[c][r][cc + 1],
                        pl[c][r][cc + 2], pl[c][r][cc + 3], A.d1[0]);
        g[c][1] = d1<T>(pl[c][r - 3][cc], pl[c][r - 2][cc], pl[c][r - 1][cc], pl[c][r + 1][cc],
                        pl[c][r + 2][cc], pl[c][r + 3][cc], A.d1[1]);
        g[c][2] = d1<T>(q[c][0], q[c][1], q[c][2], q[c][4], q[c][5], q[c][6], A.d1[2]);
        l[c][0] = d2<T>(pl[c][r][cc - 2], pl[c][r][cc - 1], pl[c][r][cc], pl[c][r][cc + 1],
                        pl[c][r][cc + 2], A.d2[0]);
        l[c][1] = d2<T>(pl[c][r - 2][cc], pl[c][r - 1][cc], pl[c][r][cc], pl[c][r + 1][cc],
                        pl[c][r + 2][cc], A.d2[1]);
        l[c][2] = d2<T>(q[c][1], q[c][2], q[c][3], q[c][4], q[c][5], A.d2[2]);
      }
      if (edge_x) { g[1][0] = 0.0; g[2][0] = 0.0; }
      if (edge_y) { g[0][1] = 0.0; g[2][1] = 0.0; }
      if (A.bc[2] == FREE_SLIP && (k == 0 || k == nz - 1)) { g[0][2] = 0.0; g[1][2] = 0.0; }
      const double ux = q[0][3], uy = q[1][3], uz = q[2][3];
      const double xnu = A.xnu;
      acc[0] += 0.5 * (ux * ux + uy * uy + uz * uz);
      const double a2 = 2.0 * g[0][0], b2 = 2.0 * g[1][1], c2 = 2.0 * g[2][2];
      const double s12 = g[0][1] + g[1][0], s13 = g[0][2] + g[2][0], s23 = g[1][2] + g[2][1];
      acc[1] += 0.5 * xnu * (a2 * a2 + b2 * b2 + c2 * c2 + 2.0 * (s12 * s12) +
                             2.0 * (s13 * s13) + 2.0 * (s23 * s23));
      acc[2] += (-xnu) * (ux * (l[0][0] + l[0][1] + l[0][2]) + uy * (l[1][0] + l[1][1] + l[1][2]) +
                          uz * (l[2][0] + l[2][1] + l[2][2]));
      const double w1 = g[2][1] - g[1][2], w2 = g[0][2] - g[2][0], w3 = g[1][0] - g[0][1];
      acc[3] += 0.5 * (w1 * w1 + w2 * w2 + w3 * w3);
      acc[4] += ux * ux;
      acc[5] += uy * uy;
      acc[6] += uz * uz;
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int a = 0; a < 3; ++a) acc[7 + 3 * c + a] += g[c][a] * g[c][a];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int s = 0; s < 6; ++s) q[c][s] = q[c][s + 1];
  }
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const double v = block_sum<kNW>(acc[m], sc.v);
    if (threadIdx.x == 0) A.part[16 * blockIdx.x + m] = v;
  }
}

__global__ void __launch_bounds__(1024) finalize_sums_kernel(const double *part, int nparts,
                                                             int width, double *out);

// ------------------------------------------------------- scalar transport
// transeq, src/integration.f90:332-468: one z-march computes the p-variant
// first and second derivatives of phi, the right-hand side F1 = alpha_eff
// lap(phi) - u.grad(phi) + src, the AB update (into phi_new, before
// clipping) and per-block FP64 sums of phi_new, of clip(phi_new) and of
// min(c, 1 - c) with c = clip(phi_new) (the three global sums of the
// conservative clipping, :428-450); transeq_clip_kernel then redistributes
// the excess and clips again.
template <typename T>
__global__ void __launch_bounds__(kTX *kTY) transeq_kernel(const TranseqArgs<T> A) {
  __shared__ T pl[kTY + 6][kTX + 6];
  __shared__ double sv[kNW];
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const TilePos P = tile_pos(A.tl, nx, ny, nz, A.bc);
  const int plane = nx * ny;
  const int col = P.gi + nx * P.gj;
  const int tx = P.tx, ty = P.ty;
  int hx_col = -1, hx_g = 0;
  bool hx_f = false;
  if (tx < 3) { hx_col = tx; hx_g = ghost_safe(P.x0 + tx - 3, nx, A.bc[0], hx_f); }
  else if (tx >= kTX - 3) { hx_col = tx + 6; hx_g = ghost_safe(P.x0 + tx + 3, nx, A.bc[0], hx_f); }
  int hy_row = -1, hy_g = 0;
  bool hy_f = false;
  if (ty < 3) { hy_row = ty; hy_g = ghost_safe(P.y0 + ty - 3, ny, A.bc[1], hy_f); }
  else if (ty >= kTY - 3) { hy_row = ty + 6; hy_g = ghost_safe(P.y0 + ty + 3, ny, A.bc[1], hy_f); }
  const int hx_idx = hx_g + nx * P.gj;
  const int hy_idx = P.gi + nx * hy_g;
  // phi is even about free-slip walls (p variants): no sign flips; the even
  // first derivative is forced to zero on the wall nodes.
  const bool edge_x = A.bc[0] == FREE_SLIP && (P.i == 0 || P.i == nx - 1);
  const bool edge_y = A.bc[1] == FREE_SLIP && (P.j == 0 || P.j == ny - 1);
  double s_old = 0.0, s_clip = 0.0, s_w = 0.0;
  T q[8];
#pragma unroll
  for (int s = 0; s < 7; ++s) {
    bool fz;
    const int kk = ghost_index(P.z0 - 3 + s, nz, A.bc[2], fz);
    q[s] = A.phi[col + plane * kk];
  }
  for (int k = P.z0; k < P.z1; ++k) {
    {
      bool fz;  // k + 4 <= nz + 3 stays a valid mirror / wrap index for nz >= 6
      const int kk = ghost_index(k + 4, nz, A.bc[2], fz);
      q[7] = A.phi[col + plane * kk];
    }
    const int pk = plane * k;
    pl[ty + 3][tx + 3] = q[3];
    if (hx_col >= 0) pl[ty + 3][hx_col] = A.phi[hx_idx + pk];
    if (hy_row >= 0) pl[hy_row][tx + 3] = A.phi[hy_idx + pk];
    __syncthreads();
    if (P.valid) {
      const int idx = P.i + nx * P.j + pk;
      const int r = ty + 3, cc = tx + 3;
      const T ux = A.u[0][idx], uy = A.u[1][idx], uz = A.u[2][idx], sr = A.src[idx];
      T gx = d1<T>(pl[r][cc - 3], pl[r][cc - 2], pl[r][cc - 1], pl[r][cc + 1], pl[r][cc + 2],
                   pl[r][cc + 3], A.d1[0]);
      T gy = d1<T>(pl[r - 3][cc], pl[r - 2][cc], pl[r - 1][cc], pl[r + 1][cc], pl[r + 2][cc],
                   pl[r + 3][cc], A.d1[1]);
      T gz = d1<T>(q[0], q[1], q[2], q[4], q[5], q[6], A.d1[2]);
      if (edge_x) gx = T(0);
      if (edge_y) gy = T(0);
      if (A.bc[2] == FREE_SLIP && (k == 0 || k == nz - 1)) gz = T(0);
      const T lx = d2<T>(pl[r][cc - 2], pl[r][cc - 1], pl[r][cc], pl[r][cc + 1], pl[r][cc + 2],
                         A.d2[0]);
      const T ly = d2<T>(pl[r - 2][cc], pl[r - 1][cc], pl[r][cc], pl[r + 1][cc], pl[r + 2][cc],
                         A.d2[1]);
      const T lz = d2<T>(q[1], q[2], q[3], q[4], q[5], A.d2[2]);
      // src/integration.f90:405-411 (alpha_eff), :422-426 (F1, update)
      const T alpha = A.les ? A.alpha0 + A.nu_t[idx] / A.sc : A.alpha0;
      const T F = alpha * (lx + ly + lz) - (ux * gx + uy * gy + uz * gz) + sr;
      T pn = q[3] + A.adu * F;
      if (A.nh >= 2) pn = pn + A.bdu * A.f2[idx];
      if (A.nh >= 3) pn = pn + A.cdu * A.f3[idx];
      A.f1[idx] = F;
      A.phi_new[idx] = pn;
      const T c = fmax(T(0), fmin(T(1), pn));
      s_old += (double)pn;
      s_clip += (double)c;
      s_w += (double)fmin(c, T(1) - c);
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 7; ++s) q[s] = q[s + 1];
  }
  const double a = block_sum<kNW>(s_old, sv);
  const double b = block_sum<kNW>(s_clip, sv);
  const double c = block_sum<kNW>(s_w, sv);
  if (threadIdx.x == 0) {
    A.part[3 * blockIdx.x + 0] = a;
    A.part[3 * blockIdx.x + 1] = b;
    A.part[3 * blockIdx.x + 2] = c;
  }
}

// phi = clip(c + excess * (w / sum w)), c = clip(phi_new), w = min(c, 1 - c),
// excess = mean(phi_new) - mean(c) (src/integration.f90:436-450).  A zero
// weight sum (every value at a bound) leaves c unchanged.
template <typename T>
__global__ void __launch_bounds__(256) transeq_clip_kernel(T *phi, const T *phi_new,
                                                           const double *sums, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T c = fmax(T(0), fmin(T(1), phi_new[i]));
  const double N = (double)n;
  const T excess = T(sums[0] / N - sums[1] / N);
  const T sw = T(sums[2]);
  T v = c;
  if (sw != T(0)) {
    const T w = fmin(c, T(1) - c) / sw;
    v = c + excess * w;
  }
  phi[i] = fmax(T(0), fmin(T(1), v));
}

template <typename T>
hipError_t launch_transeq(const TranseqArgs<T> &a, double *sums, hipStream_t s) {
  transeq_kernel<T><<<dim3(a.tl.blocks()), dim3(kTX * kTY), 0, s>>>(a);
  finalize_sums_kernel<<<1, 1024, 0, s>>>(a.part, a.tl.blocks(), 3, sums);
  const int n = a.nx * a.ny * a.nz;
  transeq_clip_kernel<T><<<(n + 255) / 256, 256, 0, s>>>(a.phi_out, a.phi_new, sums, n);
  return hipGetLastError();
}

// ------------------------------------------------------- vortical diagnostics
// rotational (mode 0: out[0..2] = curl u, src/differential_operators.f90:40-76)
// and calculate_Q_criterion (mode 1: out[0] = Q, :78-108) in one pass over
// (u, v, w).  The nine first derivatives carry the parity of the reference's
// variant choice (u_c odd along axis c: derxi/deryi/derzi on the diagonal,
// the p variants with their wall zeros across), as in the predict kernel, and
// each output keeps the reference's operation order.
template <typename T>
__global__ void __launch_bounds__(kTX *kTY) diag_kernel(const DiagArgs<T> A) {
  __shared__ T pl[3][kTY + 6][kTX + 6];
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const TilePos P = tile_pos(A.tl, nx, ny, nz, A.bc);
  const int plane = nx * ny;
  const int col = P.gi + nx * P.gj;
  const int tx = P.tx, ty = P.ty;
  const bool own_neg[2] = {P.fi, P.fj};
  int hx_col = -1, hx_g = 0;
  bool hx_f = false;
  if (tx < 3) { hx_col = tx; hx_g = ghost_safe(P.x0 + tx - 3, nx, A.bc[0], hx_f); }
  else if (tx >= kTX - 3) { hx_col = tx + 6; hx_g = ghost_safe(P.x0 + tx + 3, nx, A.bc[0], hx_f); }
  int hy_row = -1, hy_g = 0;
  bool hy_f = false;
  if (ty < 3) { hy_row = ty; hy_g = ghost_safe(P.y0 + ty - 3, ny, A.bc[1], hy_f); }
  else if (ty >= kTY - 3) { hy_row = ty + 6; hy_g = ghost_safe(P.y0 + ty + 3, ny, A.bc[1], hy_f); }
  const int hx_idx = hx_g + nx * P.gj;
  const int hy_idx = P.gi + nx * hy_g;
  const bool edge_x = A.bc[0] == FREE_SLIP && (P.i == 0 || P.i == nx - 1);
  const bool edge_y = A.bc[1] == FREE_SLIP && (P.j == 0 || P.j == ny - 1);
  T q[3][8];
#pragma unroll
  for (int s = 0; s < 7; ++s) {
    bool fz;
    const int kk = ghost_index(P.z0 - 3 + s, nz, A.bc[2], fz);
    q[0][s] = load_ghost(A.u[0], col + plane * kk, own_neg[0]);
    q[1][s] = load_ghost(A.u[1], col + plane * kk, own_neg[1]);
    q[2][s] = load_ghost(A.u[2], col + plane * kk, fz);
  }
  for (int k = P.z0; k < P.z1; ++k) {
    {
      // k + 4 <= nz + 3 stays a valid mirror / wrap index for nz >= 6.
      bool fz;
      const int kk = ghost_index(k + 4, nz, A.bc[2], fz);
      q[0][7] = load_ghost(A.u[0], col + plane * kk, own_neg[0]);
      q[1][7] = load_ghost(A.u[1], col + plane * kk, own_neg[1]);
      q[2][7] = load_ghost(A.u[2], col + plane * kk, fz);
    }
    const int pk = plane * k;
    T hx[3], hy[3];
    if (hx_col >= 0) {
      hx[0] = load_ghost(A.u[0], hx_idx + pk, hx_f);
      hx[1] = load_ghost(A.u[1], hx_idx + pk, own_neg[1]);
      hx[2] = A.u[2][hx_idx + pk];
    }
    if (hy_row >= 0) {
      hy[0] = load_ghost(A.u[0], hy_idx + pk, own_neg[0]);
      hy[1] = load_ghost(A.u[1], hy_idx + pk, hy_f);
      hy[2] = A.u[2][hy_idx + pk];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      pl[c][ty + 3][tx + 3] = q[c][3];
      if (hx_col >= 0) pl[c][ty + 3][hx_col] = hx[c];
      if (hy_row >= 0) pl[c][hy_row][tx + 3] = hy[c];
    }
    __syncthreads();
    if (P.valid) {
      const int idx = P.i + nx * P.j + pk;
      const int r = ty + 3, cc = tx + 3;
      T g[3][3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        g[c][0] = d1<T>(pl[c][r][cc - 3], pl[c][r][cc - 2], pl[c][r][cc - 1], pl[c][r][cc + 1],
                        pl[c][r][cc + 2], pl[c][r][cc + 3], A.d1[0]);
        g[c][1] = d1<T>(pl[c][r - 3][cc], pl[c][r - 2][cc], pl[c][r - 1][cc], pl[c][r + 1][cc],
                        pl[c][r + 2][cc], pl[c][r + 3][cc], A.d1[1]);
        g[c][2] = d1<T>(q[c][0], q[c][1], q[c][2], q[c][4], q[c][5], q[c][6], A.d1[2]);
      }
      if (edge_x) { g[1][0] = T(0); g[2][0] = T(0); }
      if (edge_y) { g[0][1] = T(0); g[2][1] = T(0); }
      if (A.bc[2] == FREE_SLIP && (k == 0 || k == nz - 1)) { g[0][2] = T(0); g[1][2] = T(0); }
      if (A.mode == 0) {
        A.out[0][idx] = g[2][1] - g[1][2];  // duz/dy - duy/dz
        A.out[1][idx] = g[0][2] - g[2][0];  // dux/dz - duz/dx
        A.out[2][idx] = g[1][0] - g[0][1];  // duy/dx - dux/dy
      } else {
        A.out[0][idx] = -(T(0.5) * (g[0][0] * g[0][0] + g[1][1] * g[1][1] + g[2][2] * g[2][2])) -
                        g[0][1] * g[1][0] - g[0][2] * g[2][0] - g[1][2] * g[2][1];
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int s = 0; s < 7; ++s) q[c][s] = q[c][s + 1];
  }
}

template <typename T>
hipError_t launch_diag(const DiagArgs<T> &a, hipStream_t s) {
  diag_kernel<T><<<dim3(a.tl.blocks()), dim3(kTX * kTY), 0, s>>>(a);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_statistics(const StatsArgs<T> &a, hipStream_t s) {
  stats_kernel<T><<<dim3(a.tl.blocks()), dim3(kTX * kTY), 0, s>>>(a);
  return hipGetLastError();
}

// --------------------------------------------------------- plain field scan
template <typename T>
__global__ void __launch_bounds__(kTX *kTY) field_stats_kernel(const T *__restrict__ f, int n,
                                                               FieldPartial *part) {
  __shared__ ReduceScratch sc;
  FieldPartial st;
  partial_init(st);
  // Contiguous chunk per block so that per-thread sums run in index order.
  const int per = (n + gridDim.x - 1) / gridDim.x;
  const int lo = blockIdx.x * per, hi = min(n, lo + per);
  for (int idx = lo + threadIdx.x; idx < hi; idx += blockDim.x) partial_add(st, (double)f[idx], idx);
  partial_store(st, part + blockIdx.x, sc);
}

template <typename T>
hipError_t launch_field_stats(const T *f, int n, FieldPartial *part, int nblocks, hipStream_t s) {
  field_stats_kernel<T><<<dim3(nblocks), dim3(kTX * kTY), 0, s>>>(f, n, part);
  return hipGetLastError();
}

// ---------------------------------------------------------------- finalize
__device__ __forceinline__ void finalize_one(const FieldPartial *part, int nparts, int stride,
                                             int first, FieldPartial *out) {
  __shared__ double sv[16];
  __shared__ int si[16];
  FieldPartial st;
  partial_init(st);
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
    const FieldPartial p = part[(size_t)b * stride + first];
    st.sum += p.sum;
    st.vmin = fmin(st.vmin, p.vmin);
    argmax_merge(st.vmax, st.imax, p.vmax, p.imax);
    st.amax = fmax(st.amax, p.amax);
    st.nan |= p.nan;
  }
  const double vmin = block_min<16>(st.vmin, sv);
  double vmax = st.vmax;
  int imax = st.imax;
  block_argmax<16>(vmax, imax, sv, si);
  const double sum = block_sum<16>(st.sum, sv);
  const double amax = block_max<16>(st.amax, sv);
  const double nan = block_max<16>((double)st.nan, sv);
  if (threadIdx.x == 0) {
    FieldPartial r;
    r.vmin = vmin; r.vmax = vmax; r.imax = imax; r.sum = sum; r.amax = amax;
    r.nan = nan > 0.0 ? 1 : 0;
    *out = r;
  }
}

__global__ void __launch_bounds__(1024) finalize_partials_kernel(const FieldPartial *part,
                                                                 int nparts, int stride,
                                                                 int first, FieldPartial *out) {
  finalize_one(part, nparts, stride, first, out);
}

hipError_t launch_finalize_partials(const FieldPartial *part, int nparts, int stride, int first,
                                    FieldPartial *out, hipStream_t s) {
  finalize_partials_kernel<<<1, 1024, 0, s>>>(part, nparts, stride, first, out);
  return hipGetLastError();
}

struct FinalizeSets {
  FinalizeSet s[kMaxFinalizeSets];
};

// One block per set.
__global__ void __launch_bounds__(1024) finalize_sets_kernel(FinalizeSets f) {
  const FinalizeSet &t = f.s[blockIdx.x];
  finalize_one(t.part, t.nparts, t.stride, t.first, t.out);
}

hipError_t launch_finalize_sets(const FinalizeSet *sets, int n, hipStream_t s) {
  if (n < 1 || n > kMaxFinalizeSets) return hipErrorInvalidValue;
  FinalizeSets f{};
  for (int q = 0; q < n; ++q) f.s[q] = sets[q];
  finalize_sets_kernel<<<n, 1024, 0, s>>>(f);
  return hipGetLastError();
}

__global__ void __launch_bounds__(1024) finalize_sums_kernel(const double *part, int nparts,
                                                             int width, double *out) {
  __shared__ double sv[16];
  for (int m = 0; m < width; ++m) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) s += part[(size_t)b * width + m];
    s = block_sum<16>(s, sv);
    if (threadIdx.x == 0) out[m] = s;
  }
}

hipError_t launch_finalize_sums(const double *part, int nparts, int width, double *out,
                                hipStream_t s) {
  finalize_sums_kernel<<<1, 1024, 0, s>>>(part, nparts, width, out);
  return hipGetLastError();
}

// --------------------------------------------------------- instantiations
template hipError_t launch_der<double>(int, double *, const double *, int, int, int, double,
                                       hipStream_t);
template hipError_t launch_der<float>(int, float *, const float *, int, int, int, double,
                                      hipStream_t);
template hipError_t launch_predict<double>(const PredictArgs<double> &, int, bool, hipStream_t);
template hipError_t launch_predict<float>(const PredictArgs<float> &, int, bool, hipStream_t);
template Tiling predict_tiling<double>(int, int, int, bool);
template Tiling predict_tiling<float>(int, int, int, bool);
template Tiling div_tiling<double>(int, int, int);
template Tiling div_tiling<float>(int, int, int);
template Tiling correct_tiling<double, double>(int, int, int);
template Tiling correct_tiling<float, double>(int, int, int);
template hipError_t launch_divergence<double>(const DivArgs<double> &, hipStream_t);
template hipError_t launch_divergence<float>(const DivArgs<float> &, hipStream_t);
template hipError_t launch_correct<double, double>(const CorrectArgs<double, double> &,
                                                   hipStream_t);
template hipError_t launch_correct<float, double>(const CorrectArgs<float, double> &,
                                                  hipStream_t);
template hipError_t launch_statistics<double>(const StatsArgs<double> &, hipStream_t);
template hipError_t launch_diag<double>(const DiagArgs<double> &, hipStream_t);
template hipError_t launch_transeq<double>(const TranseqArgs<double> &, double *, hipStream_t);
template hipError_t launch_statistics<float>(const StatsArgs<float> &, hipStream_t);
template hipError_t launch_field_stats<double>(const double *, int, FieldPartial *, int,
                                               hipStream_t);
template hipError_t launch_field_stats<float>(const float *, int, FieldPartial *, int,
                                              hipStream_t);

}  // namespace o3d
