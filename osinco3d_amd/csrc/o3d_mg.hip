// Multigrid V-cycles for the coarse-grid correction of the warm starts.
//
// No counterpart in the reference (its pressure solve is SOR from the
// previous pressure).  In the mixed forms the residual of the warm-start
// guess is averaged over 2^3 blocks, this module solves the 7-point Poisson
// problem of those averages (spacing 2h) approximately, and the solver adds
// the trilinear interpolation of the result to the guess.  The SOR solve
// that follows still decides convergence (same stop rule), so the
// correction only changes where SOR starts.  The grids are small (128^3 and
// below at 256^3): the per-level kernels give each thread one cell of the
// colour being relaxed (or a pair of cells), and the levels of at most 16^3
// cells run as one V-cycle in one block's LDS; the launch count dominates.
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
  float diag, invd;  // -2 (cx + cy + cz) and its reciprocal
};

// Weighted sum of the 6 neighbours (32-bit offsets: mg_plan keeps every
// level below 2^31 cells).
__device__ __forceinline__ float mg_lap_off(const float *u, const MgGrid &g, int i, int j, int k) {
  const int pl = g.mx * g.my;
  const int jr = g.mx * j, kp = pl * k;
  const float sx = u[kp + jr + mg_nb(i - 1, g.mx, g.px)] + u[kp + jr + mg_nb(i + 1, g.mx, g.px)];
  const float sy = u[kp + g.mx * mg_nb(j - 1, g.my, g.py) + i] +
                   u[kp + g.mx * mg_nb(j + 1, g.my, g.py) + i];
  const float sz = u[pl * mg_nb(k - 1, g.mz, g.pz) + jr + i] +
                   u[pl * mg_nb(k + 1, g.mz, g.pz) + jr + i];
  return g.cx * sx + g.cy * sy + g.cz * sz;
}

// Gauss-Seidel update of cell (i, j, k) at linear index q.
__device__ __forceinline__ void mg_relax(float *u, const float *f, const MgGrid &g, int i, int j,
                                         int k, int q) {
  u[q] = (f[q] - mg_lap_off(u, g, i, j, k)) * g.invd;
}

// The same update from the neighbour sums (the operation order of
// mg_lap_off and mg_relax, so both give the same value).
__device__ __forceinline__ float mg_relax_value(float fq, float sx, float sy, float sz,
                                                const MgGrid &g) {
  return (fq - (g.cx * sx + g.cy * sy + g.cz * sz)) * g.invd;
}

// Residual f - L u at cell (i, j, k).
__device__ __forceinline__ float mg_res(const float *u, const float *f, const MgGrid &g, int i,
                                        int j, int k) {
  const int q = i + g.mx * (j + g.my * k);
  return f[q] - (mg_lap_off(u, g, i, j, k) + g.diag * u[q]);
}

// The residual averaged over the 8 children of coarse cell (I, J, K).
__device__ __forceinline__ float mg_child_mean(const float *u, const float *f, const MgGrid &g,
                                               int I, int J, int K) {
  float s = 0.f;
#pragma unroll
  for (int c = 0; c < 8; ++c)
    s += mg_res(u, f, g, 2 * I + (c & 1), 2 * J + ((c >> 1) & 1), 2 * K + (c >> 2));
  return 0.125f * s;
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

// Trilinear interpolation of the coarse field uc (cx, cy, cz cells) at the
// fine cells (2I, j, k) and (2I + 1, j, k).
__device__ __forceinline__ float2 mg_interp_pair(const float *uc, int I, int j, int k, int cx,
                                                 int cy, int cz, int px, int py, int pz) {
  int y0, y1, z0, z1;
  float wy, wz;
  mg_axis(j, cy, py, y0, y1, wy);
  mg_axis(k, cz, pz, z0, z1, wz);
  const int r00 = cx * (y0 + cy * z0), r10 = cx * (y1 + cy * z0);
  const int r01 = cx * (y0 + cy * z1), r11 = cx * (y1 + cy * z1);
  float v[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    int x0, x1;
    float wx;
    mg_axis(2 * I + h, cx, px, x0, x1, wx);
    v[h] = (1.f - wz) * ((1.f - wy) * ((1.f - wx) * uc[r00 + x0] + wx * uc[r00 + x1]) +
                         wy * ((1.f - wx) * uc[r10 + x0] + wx * uc[r10 + x1])) +
           wz * ((1.f - wy) * ((1.f - wx) * uc[r01 + x0] + wx * uc[r01 + x1]) +
                 wy * ((1.f - wx) * uc[r11 + x0] + wx * uc[r11 + x1]));
  }
  return make_float2(v[0], v[1]);
}

// One colour of a red-black Gauss-Seidel sweep, one thread per cell of that
// colour (mx even: every row holds mx / 2 cells of each colour).  The
// smoothing, restriction and prolongation kernels take their blocks in
// XCD-aware order (xcd_tile): each XCD works on a contiguous run of planes
// and finds most neighbour planes in its own L2 (at 256^3: the 128^3
// prolongation 10.2 -> 8.0 us, restriction 10.1 -> 9.7 us, smoothing 7.5 ->
// 7.2 us).  The pooling kernels only stream and keep the plain order.
__global__ void __launch_bounds__(256) mg_smooth_kernel(float *__restrict__ u,
                                                        const float *__restrict__ f, MgGrid g,
                                                        int colour) {
  const int hx = g.mx >> 1;
  const int p = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (p >= hx * g.my * g.mz) return;
  const int row = p / hx, t = p - row * hx;
  const int k = row / g.my, j = row - k * g.my;
  const int i = 2 * t + ((j + k + colour) & 1);
  mg_relax(u, f, g, i, j, k, i + g.mx * row);
}

// The same for mx odd: one thread per cell, the other colour idle.
__global__ void __launch_bounds__(256) mg_smooth_any_kernel(float *__restrict__ u,
                                                            const float *__restrict__ f,
                                                            MgGrid g, int colour) {
  const int idx = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (idx >= g.mx * g.my * g.mz) return;
  const int row = idx / g.mx, i = idx - row * g.mx;
  const int k = row / g.my, j = row - k * g.my;
  if (((i + j + k) & 1) != colour) return;
  mg_relax(u, f, g, i, j, k, idx);
}

// fc = the residual of the fine level averaged over the 8 children of each
// coarse cell; uc = 0.
__global__ void __launch_bounds__(256) mg_restrict_kernel(float *__restrict__ fc,
                                                          float *__restrict__ uc,
                                                          const float *__restrict__ u,
                                                          const float *__restrict__ f, MgGrid g) {
  const int cx = g.mx >> 1, cy = g.my >> 1, cz = g.mz >> 1;
  const int idx = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (idx >= cx * cy * cz) return;
  const int crow = idx / cx, I = idx - crow * cx;
  const int K = crow / cy, J = crow - K * cy;
  fc[idx] = mg_child_mean(u, f, g, I, J, K);
  uc[idx] = 0.f;
}

// u += trilinear interpolation of the coarse correction uc, one thread per
// pair of fine cells (2I, 2I + 1) of a row.
__global__ void __launch_bounds__(256) mg_prolong_kernel(float *__restrict__ u,
                                                         const float *__restrict__ uc, int mx,
                                                         int my, int mz, int px, int py, int pz) {
  const int cx = mx >> 1;
  const int p = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (p >= cx * my * mz) return;
  const int row = p / cx, I = p - row * cx;
  const int k = row / my, j = row - k * my;
  const float2 v = mg_interp_pair(uc, I, j, k, cx, my >> 1, mz >> 1, px, py, pz);
  float2 *o = reinterpret_cast<float2 *>(u + (size_t)mx * row) + I;
  float2 w = *o;
  w.x += v.x;
  w.y += v.y;
  *o = w;
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
  for (int s = 0; s < sweeps; ++s)
    for (int colour = 0; colour < 2; ++colour) {
      for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int row = q / g.mx, i = q - row * g.mx;
        const int k = row / g.my, j = row - k * g.my;
        if (((i + j + k) & 1) == colour) mg_relax(su, sf, g, i, j, k, q);
      }
      __syncthreads();
    }
  for (int q = threadIdx.x; q < n; q += blockDim.x) u[q] = su[q];
}

// The levels from `first` down, all of at most kMgCoarsestCells cells, as
// one V-cycle by one block in LDS: f of the first level in, u of the first
// level out (from u = 0).  Same smoother, residual, restriction and
// prolongation as the per-level kernels.  Each thread keeps the coordinates
// of its cells of one colour in registers for the sweeps of a level.
constexpr int kMgSmallLevels = 4;
constexpr int kMgSmallCells = 4096 + 512 + 64 + 8;
constexpr int kMgSmallThreads = 1024;
constexpr int kMgSmallPer = kMgCoarsestCells / 2 / kMgSmallThreads;
struct MgSmall {
  int levels;
  MgGrid g[kMgSmallLevels];
  int off[kMgSmallLevels];  // offsets of each level in the LDS arrays
  int nu1, nu2, coarse_sweeps;
};

__device__ __forceinline__ void lds_smooth(float *u, const float *f, const MgGrid &g,
                                           int sweeps) {
  if (g.mx & 1) {
    const int n = g.mx * g.my * g.mz;
    for (int s = 0; s < sweeps; ++s)
      for (int colour = 0; colour < 2; ++colour) {
        for (int q = threadIdx.x; q < n; q += kMgSmallThreads) {
          const int row = q / g.mx, i = q - row * g.mx;
          const int k = row / g.my, j = row - k * g.my;
          if (((i + j + k) & 1) == colour) mg_relax(u, f, g, i, j, k, q);
        }
        __syncthreads();
      }
    return;
  }
  const int hx = g.mx >> 1, nh = hx * g.my * g.mz;
  int t[kMgSmallPer], jj[kMgSmallPer], kk[kMgSmallPer];
#pragma unroll
  for (int s = 0; s < kMgSmallPer; ++s) {
    const int p = threadIdx.x + s * kMgSmallThreads;
    const int row = p / hx;
    t[s] = p - row * hx;
    kk[s] = row / g.my;
    jj[s] = row - kk[s] * g.my;
  }
  for (int sw = 0; sw < sweeps; ++sw)
    for (int colour = 0; colour < 2; ++colour) {
#pragma unroll
      for (int s = 0; s < kMgSmallPer; ++s)
        if (threadIdx.x + s * kMgSmallThreads < nh) {
          const int i = 2 * t[s] + ((jj[s] + kk[s] + colour) & 1);
          mg_relax(u, f, g, i, jj[s], kk[s], i + g.mx * (jj[s] + g.my * kk[s]));
        }
      __syncthreads();
    }
}

// `sweeps` red-black sweeps of a level of at most 64 cells by wave 0, one
// cell per lane in registers, the neighbours fetched across lanes: no block
// barrier per colour.  The same updates in the same order as lds_smooth.
// Called by every thread of the block (it ends with the block barrier).
__device__ __forceinline__ void wave_smooth(float *u, const float *f, const MgGrid &g,
                                            int sweeps) {
  if (threadIdx.x < O3D_WAVE) {
    const int n = g.mx * g.my * g.mz;
    const int q = threadIdx.x;
    const bool act = q < n;
    int i = 0, j = 0, k = 0;
    if (act) {
      const int row = q / g.mx;
      i = q - row * g.mx;
      k = row / g.my;
      j = row - k * g.my;
    }
    // Neighbour lanes (inactive lanes read their own).
    const int rj = g.mx * (j + g.my * k);
    const int lxm = act ? rj + mg_nb(i - 1, g.mx, g.px) : q;
    const int lxp = act ? rj + mg_nb(i + 1, g.mx, g.px) : q;
    const int lym = act ? i + g.mx * (mg_nb(j - 1, g.my, g.py) + g.my * k) : q;
    const int lyp = act ? i + g.mx * (mg_nb(j + 1, g.my, g.py) + g.my * k) : q;
    const int lzm = act ? i + g.mx * (j + g.my * mg_nb(k - 1, g.mz, g.pz)) : q;
    const int lzp = act ? i + g.mx * (j + g.my * mg_nb(k + 1, g.mz, g.pz)) : q;
    const int col = (i + j + k) & 1;
    float uq = act ? u[q] : 0.f;
    const float fq = act ? f[q] : 0.f;
    for (int sw = 0; sw < sweeps; ++sw)
      for (int colour = 0; colour < 2; ++colour) {
        const float sx = __shfl(uq, lxm) + __shfl(uq, lxp);
        const float sy = __shfl(uq, lym) + __shfl(uq, lyp);
        const float sz = __shfl(uq, lzm) + __shfl(uq, lzp);
        if (act && col == colour) uq = mg_relax_value(fq, sx, sy, sz, g);
      }
    if (act) u[q] = uq;
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kMgSmallThreads) mg_small_vcycle_kernel(
    float *__restrict__ u_out, const float *__restrict__ f_in, MgSmall S) {
  __shared__ float su[kMgSmallCells], sf[kMgSmallCells];
  const int n0 = S.g[0].mx * S.g[0].my * S.g[0].mz;
  for (int q = threadIdx.x; q < n0; q += kMgSmallThreads) {
    sf[q] = f_in[q];
    su[q] = 0.f;
  }
  __syncthreads();
  // Down: smooth, then the averaged residual as the next level's f.
#pragma unroll
  for (int l = 0; l < kMgSmallLevels - 1; ++l) {
    if (l >= S.levels - 1) break;
    const MgGrid &g = S.g[l], &c = S.g[l + 1];
    float *u = su + S.off[l];
    const float *f = sf + S.off[l];
    lds_smooth(u, f, g, S.nu1);
    const int nc = c.mx * c.my * c.mz;
    for (int q = threadIdx.x; q < nc; q += kMgSmallThreads) {
      const int crow = q / c.mx, I = q - crow * c.mx;
      const int K = crow / c.my, J = crow - K * c.my;
      sf[S.off[l + 1] + q] = mg_child_mean(u, f, g, I, J, K);
      su[S.off[l + 1] + q] = 0.f;
    }
    __syncthreads();
  }
  // Coarsest.
#pragma unroll
  for (int l = 1; l < kMgSmallLevels; ++l)
    if (l == S.levels - 1) {
      const MgGrid &g = S.g[l];
      if (g.mx * g.my * g.mz <= O3D_WAVE)
        wave_smooth(su + S.off[l], sf + S.off[l], g, S.coarse_sweeps);
      else
        lds_smooth(su + S.off[l], sf + S.off[l], g, S.coarse_sweeps);
    }
  // Up: prolong, smooth.
#pragma unroll
  for (int l = kMgSmallLevels - 2; l >= 0; --l) {
    if (l >= S.levels - 1) continue;
    const MgGrid &g = S.g[l], &c = S.g[l + 1];
    float *u = su + S.off[l];
    const float *uc = su + S.off[l + 1];
    const int hx = g.mx >> 1, np = hx * g.my * g.mz;
    for (int q = threadIdx.x; q < np; q += kMgSmallThreads) {
      const int row = q / hx, I = q - row * hx;
      const int k = row / g.my, j = row - k * g.my;
      const float2 v = mg_interp_pair(uc, I, j, k, c.mx, c.my, c.mz, g.px, g.py, g.pz);
      u[g.mx * row + 2 * I] += v.x;
      u[g.mx * row + 2 * I + 1] += v.y;
    }
    __syncthreads();
    lds_smooth(u, sf + S.off[l], g, S.nu2);
  }
  for (int q = threadIdx.x; q < n0; q += kMgSmallThreads) u_out[q] = su[q];
}

MgGrid grid_of(const MgHierarchy &H, int l) {
  MgGrid g;
  g.mx = H.m[l][0]; g.my = H.m[l][1]; g.mz = H.m[l][2];
  g.px = H.per[0]; g.py = H.per[1]; g.pz = H.per[2];
  g.cx = (float)H.c[l][0]; g.cy = (float)H.c[l][1]; g.cz = (float)H.c[l][2];
  g.diag = -2.f * (g.cx + g.cy + g.cz);
  g.invd = 1.f / g.diag;
  return g;
}

unsigned blocks_for(long long threads) { return (unsigned)((threads + 255) / 256); }

hipError_t smooth(const MgHierarchy &H, int l, int sweeps, hipStream_t s) {
  const MgGrid g = grid_of(H, l);
  const long long n = (long long)g.mx * g.my * g.mz;
  for (int q = 0; q < sweeps; ++q)
    for (int colour = 0; colour < 2; ++colour) {
      if (g.mx & 1)
        mg_smooth_any_kernel<<<blocks_for(n), 256, 0, s>>>(H.u[l], H.f[l], g, colour);
      else
        mg_smooth_kernel<<<blocks_for(n / 2), 256, 0, s>>>(H.u[l], H.f[l], g, colour);
    }
  return hipGetLastError();
}

// The levels from l down fit one block's LDS (every level at most
// kMgCoarsestCells cells, the sum at most kMgSmallCells), and every level
// above the last is halvable (mx even).
bool small_from(const MgHierarchy &H, int l) {
  if (H.levels - l > kMgSmallLevels) return false;
  size_t tot = 0;
  for (int q = l; q < H.levels; ++q) {
    const size_t c = (size_t)H.m[q][0] * H.m[q][1] * H.m[q][2];
    if (c > (size_t)kMgCoarsestCells) return false;
    tot += c;
  }
  return tot <= (size_t)kMgSmallCells;
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
    S.coarse_sweeps = 10;
    mg_small_vcycle_kernel<<<1, kMgSmallThreads, 0, s>>>(H.u[l], H.f[l], S);
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
  mg_restrict_kernel<<<blocks_for(n / 8), 256, 0, s>>>(H.f[l + 1], H.u[l + 1], H.u[l], H.f[l], g);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = vcycle(H, l + 1, nu1, nu2, s)) != hipSuccess) return e;
  mg_prolong_kernel<<<blocks_for(n / 2), 256, 0, s>>>(H.u[l], H.u[l + 1], g.mx, g.my, g.mz, g.px,
                                                      g.py, g.pz);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return smooth(H, l, nu2, s);
}

// Per-block sums of f (at most kMgMeanBlocks blocks), then f -= their total / n:
// the coarse problem made compatible with its null space (the constants).
constexpr int kMgMeanBlocks = 256;
__global__ void __launch_bounds__(256) mg_sum_kernel(const float *__restrict__ f, int n,
                                                     double *__restrict__ part) {
  __shared__ double sc[4];
  double s = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s += (double)f[i];
  s = block_sum<4>(s, sc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(256) mg_center_kernel(float *__restrict__ f, int n,
                                                        const double *__restrict__ part,
                                                        int nparts) {
  __shared__ double sc[4];
  double s = threadIdx.x < nparts ? part[threadIdx.x] : 0.0;
  s = block_sum<4>(s, sc);
  const float m = (float)(s / n);
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) f[i] -= m;
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

// The trilinear interpolation of the level-0 solution onto the fine grid,
// stored colour-split (the first FP32 correction solve's starting iterate):
// one thread per pair of cells (2c, 2c + 1) of a row, which are one red and
// one black cell, so both halves of the split array are written coalesced;
// 2D block (64 x 4) over (c, j), blockIdx.z = k: no 64-bit index division.
__global__ void __launch_bounds__(256) mg_prolong_split_kernel(float *__restrict__ e,
                                                               const float *__restrict__ u0,
                                                               int nx, int ny, int nz, int px,
                                                               int py, int pz) {
  const int hx = nx / 2;
  const int c = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (c >= hx || j >= ny) return;
  const size_t half = (size_t)nx * ny * nz / 2;
  const size_t o = (size_t)hx * (j + (size_t)ny * k) + c;
  const int q = (j + k) & 1;
  const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
  int y0, y1, z0, z1;
  float wy, wz;
  mg_axis(j, cy, py, y0, y1, wy);
  mg_axis(k, cz, pz, z0, z1, wz);
  auto cv = [&](int a, int b, int d) {
    return u0[(size_t)a + (size_t)cx * (b + (size_t)cy * d)];
  };
  auto at = [&](int i) {
    int x0, x1;
    float wx;
    mg_axis(i, cx, px, x0, x1, wx);
    return (1.f - wz) * ((1.f - wy) * ((1.f - wx) * cv(x0, y0, z0) + wx * cv(x1, y0, z0)) +
                         wy * ((1.f - wx) * cv(x0, y1, z0) + wx * cv(x1, y1, z0))) +
           wz * ((1.f - wy) * ((1.f - wx) * cv(x0, y0, z1) + wx * cv(x1, y0, z1)) +
                 wy * ((1.f - wx) * cv(x0, y1, z1) + wx * cv(x1, y1, z1)));
  };
  const float va = at(2 * c), vb = at(2 * c + 1);
  // Red cell: i = 2c + q; black: i = 2c + 1 - q (entry c of its colour's row).
  e[o] = q ? vb : va;
  e[half + o] = q ? va : vb;
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
  dim3 grid((nx / 2 + 63) / 64, (ny + 3) / 4, nz);
  mg_prolong_split_kernel<<<grid, dim3(64, 4), 0, s>>>(e, u0, nx, ny, nz, per[0], per[1],
                                                       per[2]);
  return hipGetLastError();
}

hipError_t mg_seed_from_level1(const MgHierarchy &H, hipStream_t s) {
  if (H.levels < 2) return hipErrorInvalidValue;
  const MgGrid g = grid_of(H, 0);
  const int n0 = g.mx * g.my * g.mz;
  mg_zero_kernel<<<(n0 + 255) / 256, 256, 0, s>>>(H.u[0], n0);
  mg_prolong_kernel<<<blocks_for(n0 / 2), 256, 0, s>>>(H.u[0], H.u[1], g.mx, g.my, g.mz, g.px,
                                                       g.py, g.pz);
  return hipGetLastError();
}

hipError_t mg_solve(const MgHierarchy &H, int cycles, int nu1, int nu2, hipStream_t s,
                    bool from_zero) {
  if (H.levels < 1) return hipErrorInvalidValue;
  if ((long long)H.m[0][0] * H.m[0][1] * H.m[0][2] >= (1LL << 31)) return hipErrorInvalidValue;
  const int n0 = H.m[0][0] * H.m[0][1] * H.m[0][2];
  if (!(H.per[0] && H.per[1] && H.per[2]) && n0 >= 2) {
    // With a mirrored axis the fine residual need not have zero mean (an
    // rhs the mirror cannot satisfy); only its zero-mean part is solvable
    // here.  On all-periodic grids the mean is zero up to rounding.  The
    // level-0 residual array is free: restriction forms residuals on the fly.
    // At most n0 / 2 partials: they fit the n0 floats of that array.
    double *part = reinterpret_cast<double *>(H.r[0]);
    const int nb = std::max(1, std::min(kMgMeanBlocks, n0 / 2));
    mg_sum_kernel<<<nb, 256, 0, s>>>(H.f[0], n0, part);
    mg_center_kernel<<<nb, 256, 0, s>>>(H.f[0], n0, part, nb);
  }
  if (from_zero) mg_zero_kernel<<<(n0 + 255) / 256, 256, 0, s>>>(H.u[0], n0);
  hipError_t e = hipGetLastError();
  for (int c = 0; c < cycles && e == hipSuccess; ++c) e = vcycle(H, 0, nu1, nu2, s);
  return e;
}

}  // namespace o3d
