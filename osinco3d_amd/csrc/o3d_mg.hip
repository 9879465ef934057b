// Multigrid V-cycles for the coarse-grid correction of the warm starts.
//
// No counterpart in the reference (its pressure solve is SOR from the
// previous pressure).  In the mixed forms the residual of the warm-start
// guess is averaged over 2^3 blocks, this module solves the 7-point Poisson
// problem of those averages (spacing 2h) approximately, and the solver adds
// the trilinear interpolation of the result to the guess.  The SOR solve
// that follows is still the judge of convergence (same stop rule), so the
// correction only changes where SOR starts.  The grids are small (128^3 and
// below at 256^3), so the kernels are plain one-cell-per-thread kernels;
// their cost is mostly the launch count.
#include <algorithm>

#include "o3d_mg.h"

namespace o3d {

size_t mg_plan(MgHierarchy &H, const int m0[3], const double h[3], const int per[3]) {
  H = MgHierarchy();
  for (int a = 0; a < 3; ++a) H.per[a] = per[a];
  int m[3] = {m0[0], m0[1], m0[2]};
  double hh[3] = {h[0], h[1], h[2]};
  size_t floats = 0;
  for (int l = 0; l < kMgMaxLevels; ++l) {
    for (int a = 0; a < 3; ++a) {
      H.m[l][a] = m[a];
      H.c[l][a] = 1.0 / (hh[a] * hh[a]);
    }
    const size_t cells = (size_t)m[0] * m[1] * m[2];
    floats += 3 * cells;
    H.levels = l + 1;
    const bool halvable = m[0] % 2 == 0 && m[1] % 2 == 0 && m[2] % 2 == 0 && m[0] >= 4 &&
                          m[1] >= 4 && m[2] >= 4;
    if (cells <= 64 || !halvable) break;
    for (int a = 0; a < 3; ++a) {
      m[a] /= 2;
      hh[a] *= 2.0;
    }
  }
  return floats;
}

void mg_bind(MgHierarchy &H, float *base) {
  for (int l = 0; l < H.levels; ++l) {
    const size_t cells = (size_t)H.m[l][0] * H.m[l][1] * H.m[l][2];
    H.u[l] = base;
    H.f[l] = base + cells;
    H.r[l] = base + 2 * cells;
    base += 3 * cells;
  }
}

namespace {

// Neighbour along an axis of m cells: periodic wrap, or the cell-centred
// mirror (the ghost cell equals the edge cell).
__device__ __forceinline__ int mg_nb(int q, int m, int per) {
  if (q < 0) return per ? q + m : 0;
  if (q >= m) return per ? q - m : m - 1;
  return q;
}

struct MgGrid {
  int mx, my, mz, px, py, pz;
  float cx, cy, cz;
};

__device__ __forceinline__ float mg_lap_off(const float *u, const MgGrid &g, int i, int j, int k) {
  const size_t pl = (size_t)g.mx * g.my;
  const size_t row = (size_t)g.mx * j + pl * k;
  const float sx = u[row + mg_nb(i - 1, g.mx, g.px)] + u[row + mg_nb(i + 1, g.mx, g.px)];
  const float sy = u[(size_t)g.mx * mg_nb(j - 1, g.my, g.py) + pl * k + i] +
                   u[(size_t)g.mx * mg_nb(j + 1, g.my, g.py) + pl * k + i];
  const float sz = u[row - pl * k + pl * mg_nb(k - 1, g.mz, g.pz) + i] +
                   u[row - pl * k + pl * mg_nb(k + 1, g.mz, g.pz) + i];
  return g.cx * sx + g.cy * sy + g.cz * sz;
}

// One colour of a red-black Gauss-Seidel sweep.
__global__ void __launch_bounds__(256) mg_smooth_kernel(float *__restrict__ u,
                                                        const float *__restrict__ f, MgGrid g,
                                                        int colour) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = g.mx * g.my * g.mz;
  if (idx >= n) return;
  const int i = idx % g.mx, j = (idx / g.mx) % g.my, k = idx / (g.mx * g.my);
  if (((i + j + k) & 1) != colour) return;
  const float diag = -2.f * (g.cx + g.cy + g.cz);
  u[idx] = (f[idx] - mg_lap_off(u, g, i, j, k)) / diag;
}

__global__ void __launch_bounds__(256) mg_residual_kernel(float *__restrict__ r,
                                                          const float *__restrict__ u,
                                                          const float *__restrict__ f, MgGrid g) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = g.mx * g.my * g.mz;
  if (idx >= n) return;
  const int i = idx % g.mx, j = (idx / g.mx) % g.my, k = idx / (g.mx * g.my);
  const float diag = -2.f * (g.cx + g.cy + g.cz);
  r[idx] = f[idx] - (mg_lap_off(u, g, i, j, k) + diag * u[idx]);
}

// fc = average of the 8 children of each coarse cell; uc = 0.
__global__ void __launch_bounds__(256) mg_restrict_kernel(float *__restrict__ fc,
                                                          float *__restrict__ uc,
                                                          const float *__restrict__ r, int mx,
                                                          int my, int mz) {
  const int cx = mx / 2, cy = my / 2, cz = mz / 2;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= cx * cy * cz) return;
  const int I = idx % cx, J = (idx / cx) % cy, K = idx / (cx * cy);
  float s = 0.f;
  for (int c = 0; c < 8; ++c) {
    const int i = 2 * I + (c & 1), j = 2 * J + ((c >> 1) & 1), k = 2 * K + (c >> 2);
    s += r[(size_t)i + (size_t)mx * (j + (size_t)my * k)];
  }
  fc[idx] = 0.125f * s;
  uc[idx] = 0.f;
}

// Position of fine cell i on the grid coarser by 2: lower coarse cell and
// the weight of the upper one (cell-centred; periodic wraps, others clamp).
__device__ __forceinline__ void mg_axis(int i, int m, int per, int &c0, int &c1, float &w) {
  const float X = 0.5f * (float)i - 0.25f;
  int I0 = (int)floorf(X);
  w = X - (float)I0;
  int I1 = I0 + 1;
  if (per) {
    I0 = I0 < 0 ? I0 + m : I0;
    I1 = I1 >= m ? I1 - m : I1;
  } else {
    if (I0 < 0) { I0 = 0; w = 0.f; }
    if (I1 >= m) { I1 = m - 1; w = I0 == m - 1 ? 0.f : w; }
  }
  c0 = I0;
  c1 = I1;
}

// u += trilinear interpolation of the coarse correction uc.
__global__ void __launch_bounds__(256) mg_prolong_kernel(float *__restrict__ u,
                                                         const float *__restrict__ uc, int mx,
                                                         int my, int mz, int px, int py, int pz) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= mx * my * mz) return;
  const int i = idx % mx, j = (idx / mx) % my, k = idx / (mx * my);
  const int cx = mx / 2, cy = my / 2, cz = mz / 2;
  int x0, x1, y0, y1, z0, z1;
  float wx, wy, wz;
  mg_axis(i, cx, px, x0, x1, wx);
  mg_axis(j, cy, py, y0, y1, wy);
  mg_axis(k, cz, pz, z0, z1, wz);
  auto c = [&](int a, int b, int d) {
    return uc[(size_t)a + (size_t)cx * (b + (size_t)cy * d)];
  };
  const float v =
      (1.f - wz) * ((1.f - wy) * ((1.f - wx) * c(x0, y0, z0) + wx * c(x1, y0, z0)) +
                    wy * ((1.f - wx) * c(x0, y1, z0) + wx * c(x1, y1, z0))) +
      wz * ((1.f - wy) * ((1.f - wx) * c(x0, y0, z1) + wx * c(x1, y0, z1)) +
            wy * ((1.f - wx) * c(x0, y1, z1) + wx * c(x1, y1, z1)));
  u[idx] += v;
}

// The coarsest level: `sweeps` red-black Gauss-Seidel sweeps from u = 0 by
// one block, the level staged in LDS.
__global__ void __launch_bounds__(1024) mg_coarsest_kernel(float *__restrict__ u,
                                                           const float *__restrict__ f, MgGrid g,
                                                           int sweeps) {
  __shared__ float su[kMgCoarsestCells], sf[kMgCoarsestCells];
  const int n = g.mx * g.my * g.mz;
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    su[q] = 0.f;
    sf[q] = f[q];
  }
  __syncthreads();
  const float diag = -2.f * (g.cx + g.cy + g.cz);
  for (int s = 0; s < sweeps; ++s)
    for (int colour = 0; colour < 2; ++colour) {
      for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int i = q % g.mx, j = (q / g.mx) % g.my, k = q / (g.mx * g.my);
        if (((i + j + k) & 1) == colour) su[q] = (sf[q] - mg_lap_off(su, g, i, j, k)) / diag;
      }
      __syncthreads();
    }
  for (int q = threadIdx.x; q < n; q += blockDim.x) u[q] = su[q];
}

// The levels from `first` down, all of at most kMgCoarsestCells cells, as
// one V-cycle by one block in LDS: f of the first level in, u of the first
// level out (from u = 0).  Same smoother, residual, restriction and
// prolongation as the per-level kernels.
constexpr int kMgSmallLevels = 4;
struct MgSmall {
  int levels;
  MgGrid g[kMgSmallLevels];
  int off[kMgSmallLevels];  // offsets of each level in the LDS arrays
  int nu1, nu2, coarse_sweeps;
};

__device__ __forceinline__ void lds_smooth(float *u, const float *f, const MgGrid &g,
                                           int sweeps) {
  const int n = g.mx * g.my * g.mz;
  const float diag = -2.f * (g.cx + g.cy + g.cz);
  for (int s = 0; s < sweeps; ++s)
    for (int colour = 0; colour < 2; ++colour) {
      for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int i = q % g.mx, j = (q / g.mx) % g.my, k = q / (g.mx * g.my);
        if (((i + j + k) & 1) == colour) u[q] = (f[q] - mg_lap_off(u, g, i, j, k)) / diag;
      }
      __syncthreads();
    }
}

__global__ void __launch_bounds__(1024) mg_small_vcycle_kernel(float *__restrict__ u_out,
                                                               const float *__restrict__ f_in,
                                                               MgSmall S) {
  __shared__ float su[4096 + 512 + 64 + 8],
      sf[4096 + 512 + 64 + 8], sr[4096 + 512 + 64 + 8];
  const int n0 = S.g[0].mx * S.g[0].my * S.g[0].mz;
  for (int q = threadIdx.x; q < n0; q += blockDim.x) {
    sf[q] = f_in[q];
    su[q] = 0.f;
  }
  __syncthreads();
  // Down: smooth, residual, restrict.
  for (int l = 0; l < S.levels - 1; ++l) {
    const MgGrid &g = S.g[l];
    float *u = su + S.off[l], *f = sf + S.off[l], *r = sr + S.off[l];
    lds_smooth(u, f, g, S.nu1);
    const int n = g.mx * g.my * g.mz;
    const float diag = -2.f * (g.cx + g.cy + g.cz);
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
      const int i = q % g.mx, j = (q / g.mx) % g.my, k = q / (g.mx * g.my);
      r[q] = f[q] - (mg_lap_off(u, g, i, j, k) + diag * u[q]);
    }
    __syncthreads();
    const MgGrid &c = S.g[l + 1];
    float *uc = su + S.off[l + 1], *fc = sf + S.off[l + 1];
    const int nc = c.mx * c.my * c.mz;
    for (int q = threadIdx.x; q < nc; q += blockDim.x) {
      const int I = q % c.mx, J = (q / c.mx) % c.my, K = q / (c.mx * c.my);
      float t = 0.f;
      for (int e = 0; e < 8; ++e) {
        const int i = 2 * I + (e & 1), j = 2 * J + ((e >> 1) & 1), k = 2 * K + (e >> 2);
        t += r[i + g.mx * (j + g.my * k)];
      }
      fc[q] = 0.125f * t;
      uc[q] = 0.f;
    }
    __syncthreads();
  }
  // Coarsest.
  {
    const int l = S.levels - 1;
    lds_smooth(su + S.off[l], sf + S.off[l], S.g[l], S.coarse_sweeps);
  }
  // Up: prolong, smooth.
  for (int l = S.levels - 2; l >= 0; --l) {
    const MgGrid &g = S.g[l];
    const MgGrid &c = S.g[l + 1];
    float *u = su + S.off[l];
    const float *uc = su + S.off[l + 1];
    const int n = g.mx * g.my * g.mz;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
      const int i = q % g.mx, j = (q / g.mx) % g.my, k = q / (g.mx * g.my);
      int x0, x1, y0, y1, z0, z1;
      float wx, wy, wz;
      mg_axis(i, c.mx, g.px, x0, x1, wx);
      mg_axis(j, c.my, g.py, y0, y1, wy);
      mg_axis(k, c.mz, g.pz, z0, z1, wz);
      auto cv = [&](int a, int b, int d) { return uc[a + c.mx * (b + c.my * d)]; };
      u[q] += (1.f - wz) * ((1.f - wy) * ((1.f - wx) * cv(x0, y0, z0) + wx * cv(x1, y0, z0)) +
                            wy * ((1.f - wx) * cv(x0, y1, z0) + wx * cv(x1, y1, z0))) +
              wz * ((1.f - wy) * ((1.f - wx) * cv(x0, y0, z1) + wx * cv(x1, y0, z1)) +
                    wy * ((1.f - wx) * cv(x0, y1, z1) + wx * cv(x1, y1, z1)));
    }
    __syncthreads();
    lds_smooth(u, sf + S.off[l], g, S.nu2);
  }
  for (int q = threadIdx.x; q < n0; q += blockDim.x) u_out[q] = su[q];
}

MgGrid grid_of(const MgHierarchy &H, int l) {
  MgGrid g;
  g.mx = H.m[l][0]; g.my = H.m[l][1]; g.mz = H.m[l][2];
  g.px = H.per[0]; g.py = H.per[1]; g.pz = H.per[2];
  g.cx = (float)H.c[l][0]; g.cy = (float)H.c[l][1]; g.cz = (float)H.c[l][2];
  return g;
}

int blocks_of(const MgGrid &g) { return (g.mx * g.my * g.mz + 255) / 256; }

hipError_t smooth(const MgHierarchy &H, int l, int sweeps, hipStream_t s) {
  const MgGrid g = grid_of(H, l);
  for (int q = 0; q < sweeps; ++q)
    for (int colour = 0; colour < 2; ++colour)
      mg_smooth_kernel<<<blocks_of(g), 256, 0, s>>>(H.u[l], H.f[l], g, colour);
  return hipGetLastError();
}

// The levels from l down fit one block's LDS (every level at most
// kMgCoarsestCells cells, the sum at most 4096 + 512 + 64 + 8).
bool small_from(const MgHierarchy &H, int l) {
  if (H.levels - l > kMgSmallLevels) return false;
  size_t tot = 0;
  for (int q = l; q < H.levels; ++q) {
    const size_t c = (size_t)H.m[q][0] * H.m[q][1] * H.m[q][2];
    if (c > (size_t)kMgCoarsestCells) return false;
    tot += c;
  }
  return tot <= 4096 + 512 + 64 + 8;
}

hipError_t vcycle(const MgHierarchy &H, int l, int nu1, int nu2, hipStream_t s) {
  const MgGrid g = grid_of(H, l);
  const int n = g.mx * g.my * g.mz;
  if (l > 0 && H.levels - l >= 2 && small_from(H, l)) {
    MgSmall S;
    S.levels = H.levels - l;
    int off = 0;
    for (int q = 0; q < S.levels; ++q) {
      S.g[q] = grid_of(H, l + q);
      S.off[q] = off;
      off += H.m[l + q][0] * H.m[l + q][1] * H.m[l + q][2];
    }
    S.nu1 = nu1;
    S.nu2 = nu2;
    S.coarse_sweeps = 30;
    mg_small_vcycle_kernel<<<1, 1024, 0, s>>>(H.u[l], H.f[l], S);
    return hipGetLastError();
  }
  if (l == H.levels - 1) {
    // At most 64 cells when the axes halve that far: one wave, 30 sweeps.
    if (n <= kMgCoarsestCells) {
      const int threads = n <= 64 ? 64 : 1024;
      mg_coarsest_kernel<<<1, threads, 0, s>>>(H.u[l], H.f[l], g, n <= 64 ? 30 : 60);
      return hipGetLastError();
    }
    return smooth(H, l, 60, s);
  }
  hipError_t e = smooth(H, l, nu1, s);
  if (e != hipSuccess) return e;
  mg_residual_kernel<<<blocks_of(g), 256, 0, s>>>(H.r[l], H.u[l], H.f[l], g);
  const int nc = n / 8;
  mg_restrict_kernel<<<(nc + 255) / 256, 256, 0, s>>>(H.f[l + 1], H.u[l + 1], H.r[l], g.mx,
                                                      g.my, g.mz);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = vcycle(H, l + 1, nu1, nu2, s)) != hipSuccess) return e;
  mg_prolong_kernel<<<blocks_of(g), 256, 0, s>>>(H.u[l], H.u[l + 1], g.mx, g.my, g.mz, g.px,
                                                 g.py, g.pz);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return smooth(H, l, nu2, s);
}

__global__ void mg_zero_kernel(float *u, int n) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) u[idx] = 0.f;
}

__global__ void __launch_bounds__(256) mg_pool_split_kernel(const float *__restrict__ r,
                                                            float *__restrict__ f0, int nx,
                                                            int ny, int nz) {
  const int mx = nx / 2, my = ny / 2, mz = nz / 2;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= mx * my * mz) return;
  const int I = idx % mx, J = (idx / mx) % my, K = idx / (mx * my);
  const size_t hx = (size_t)(nx / 2), half = (size_t)nx * ny * nz / 2;
  float t = 0.f;
  // Cells 2I and 2I+1 of a row are one red and one black node, both entry I
  // of their colour's row.
  for (int q = 0; q < 4; ++q) {
    const size_t row = (size_t)(2 * J + (q & 1)) + (size_t)ny * (2 * K + (q >> 1));
    t += r[hx * row + I] + r[half + hx * row + I];
  }
  f0[idx] = 0.125f * t;
}

__global__ void __launch_bounds__(256) mg_prolong_split_kernel(float *__restrict__ e,
                                                               const float *__restrict__ u0,
                                                               int nx, int ny, int nz, int px,
                                                               int py, int pz) {
  const size_t half = (size_t)nx * ny * nz / 2;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 2 * half) return;
  const int black = idx >= half ? 1 : 0;
  const size_t o = idx - (black ? half : 0);
  const int hx = nx / 2;
  const int c = (int)(o % hx);
  const size_t row = o / hx;
  const int j = (int)(row % ny), k = (int)(row / ny);
  const int q = (j + k) & 1;
  const int i = 2 * c + (black ? 1 - q : q);
  const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
  int x0, x1, y0, y1, z0, z1;
  float wx, wy, wz;
  mg_axis(i, cx, px, x0, x1, wx);
  mg_axis(j, cy, py, y0, y1, wy);
  mg_axis(k, cz, pz, z0, z1, wz);
  auto cv = [&](int a, int b, int d) {
    return u0[(size_t)a + (size_t)cx * (b + (size_t)cy * d)];
  };
  e[idx] = (1.f - wz) * ((1.f - wy) * ((1.f - wx) * cv(x0, y0, z0) + wx * cv(x1, y0, z0)) +
                         wy * ((1.f - wx) * cv(x0, y1, z0) + wx * cv(x1, y1, z0))) +
           wz * ((1.f - wy) * ((1.f - wx) * cv(x0, y0, z1) + wx * cv(x1, y0, z1)) +
                 wy * ((1.f - wx) * cv(x0, y1, z1) + wx * cv(x1, y1, z1)));
}

}  // namespace

hipError_t mg_pool_split(const float *r, float *f0, int nx, int ny, int nz, hipStream_t s) {
  if (nx % 2 || ny % 2 || nz % 2) return hipErrorInvalidValue;
  const int nc = (nx / 2) * (ny / 2) * (nz / 2);
  mg_pool_split_kernel<<<(nc + 255) / 256, 256, 0, s>>>(r, f0, nx, ny, nz);
  return hipGetLastError();
}

hipError_t mg_prolong_split(float *e, const float *u0, int nx, int ny, int nz, const int per[3],
                            hipStream_t s) {
  if (nx % 2 || ny % 2 || nz % 2) return hipErrorInvalidValue;
  const size_t n = (size_t)nx * ny * nz;
  mg_prolong_split_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(e, u0, nx, ny, nz, per[0],
                                                                       per[1], per[2]);
  return hipGetLastError();
}

hipError_t mg_seed_from_level1(const MgHierarchy &H, hipStream_t s) {
  if (H.levels < 2) return hipErrorInvalidValue;
  const MgGrid g = grid_of(H, 0);
  const int n0 = g.mx * g.my * g.mz;
  mg_zero_kernel<<<(n0 + 255) / 256, 256, 0, s>>>(H.u[0], n0);
  mg_prolong_kernel<<<blocks_of(g), 256, 0, s>>>(H.u[0], H.u[1], g.mx, g.my, g.mz, g.px, g.py,
                                                 g.pz);
  return hipGetLastError();
}

hipError_t mg_solve(const MgHierarchy &H, int cycles, int nu1, int nu2, hipStream_t s,
                    bool from_zero) {
  if (H.levels < 1) return hipErrorInvalidValue;
  const int n0 = H.m[0][0] * H.m[0][1] * H.m[0][2];
  if (from_zero) mg_zero_kernel<<<(n0 + 255) / 256, 256, 0, s>>>(H.u[0], n0);
  hipError_t e = hipGetLastError();
  for (int c = 0; c < cycles && e == hipSuccess; ++c) e = vcycle(H, 0, nu1, nu2, s);
  return e;
}

}  // namespace o3d
