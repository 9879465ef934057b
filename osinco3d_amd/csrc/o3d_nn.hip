// Learned warm start for the Poisson solve: a 3D U-Net-lite on MI355X.
//
// Layout: activations are channels-last [voxel][channel] BF16 (voxel = x +
// nx*(y + ny*z)), so the 8 consecutive channels an MFMA lane needs are one
// 16-byte load.  Gradients of activations are FP32 (accumulated), weight
// gradients and master weights FP32.
//
// 3x3x3 convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16:
//   M = 16 consecutive voxels (one wave), N = 16 output channels per tile,
//   K = 27 taps x C_in, flattened as k = tap*C_in + c and zero-padded to a
//   multiple of 32; lane l supplies A[m = l%16][k = 8(l/16)..+7] (one voxel,
//   8 channels of one tap) and B[k][n = l%16] (8 weights of one output
//   channel), and receives C[m = 4(l/16)+i][n = l%16].
// The data gradient is the same kernel with the weights transposed and the
// taps mirrored (periodic padding makes the adjoint of the convolution the
// convolution with the flipped kernel).  The weight gradient reduces over
// voxels in per-slab partial sums.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/o3d.h"
#include "../../include/o3d_nn.h"
#include "o3d_common.h"
#include "o3d_nn.h"
#include "o3d_sor.h"

namespace o3d {
namespace nn {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr float kSlope = 0.1f;  // leaky-ReLU slope

__device__ __forceinline__ int wrapi(int q, int n) { return q < 0 ? q + n : (q >= n ? q - n : q); }
__device__ __forceinline__ float act_f(float v) { return v > 0.f ? v : kSlope * v; }
__device__ __forceinline__ float dact(float y) { return y > 0.f ? 1.f : kSlope; }

// ----------------------------------------------------------------- convolution
template <typename TO>
struct ConvArgs {
  const bf16 *x;      // [V][ci]
  const bf16 *w;      // [co][kp]  k = tap*ci + c, zero padded
  const float *bias;  // [co] or null
  const bf16 *skip;   // [V][co] or null, added before the activation
  TO *y;              // [V][co]
  int nx, ny, nz;
  int lci, ci, co, kp;
  int act;            // 1: leaky ReLU
  int accumulate;     // y += result (FP32 gradient buffers)
};

// One wave computes 16*MV consecutive voxels (linear order, so coarse levels
// with nx < 16 work too) for 16*NT output channels: MV x NT MFMA tiles share
// each loaded A fragment across NT output tiles and each B fragment across MV
// voxel tiles.
template <typename TO, int NT, int MV>
__global__ void __launch_bounds__(256) conv_kernel(const ConvArgs<TO> a) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int V = a.nx * a.ny * a.nz;
  const int vbase = (blockIdx.x * 4 + wv) * 16 * MV;
  if (vbase >= V) return;  // no barriers in this kernel
  const int co0 = blockIdx.y * 16 * NT;
  const int g = lane >> 4, l16 = lane & 15;
  f32x4 acc[MV][NT];
#pragma unroll
  for (int m = 0; m < MV; ++m)
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int ktot = 27 * a.ci;
  const int nchunk = a.kp >> 5;
  // A-operand voxels of this lane, one per voxel tile.
  int xm[MV], ym[MV], zm[MV];
  bool vin[MV];
#pragma unroll
  for (int m = 0; m < MV; ++m) {
    const int vm = vbase + 16 * m + l16;
    vin[m] = vm < V;
    xm[m] = vm % a.nx;
    ym[m] = (vm / a.nx) % a.ny;
    zm[m] = vm / (a.nx * a.ny);
  }
  const bf16 *wrow[NT];
#pragma unroll
  for (int n = 0; n < NT; ++n) wrow[n] = a.w + (size_t)(co0 + 16 * n + l16) * a.kp;
  for (int kc = 0; kc < nchunk; ++kc) {
    const int k0 = (kc << 5) + (g << 3);
    bf16x8 bv[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) bv[n] = *(const bf16x8 *)(wrow[n] + k0);
    const bool kin = k0 < ktot;
    const int t = k0 >> a.lci;
    const int c0 = k0 & (a.ci - 1);
    const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
#pragma unroll
    for (int m = 0; m < MV; ++m) {
      bf16x8 av;
      if (kin && vin[m]) {
        const int vx = wrapi(xm[m] + dx, a.nx), vy = wrapi(ym[m] + dy, a.ny),
                  vz = wrapi(zm[m] + dz, a.nz);
        av = *(const bf16x8 *)(a.x + (size_t)(vx + a.nx * (vy + a.ny * vz)) * a.ci + c0);
      } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) av[q] = (bf16)0.f;
      }
#pragma unroll
      for (int n = 0; n < NT; ++n)
        acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv[n], acc[m][n], 0, 0, 0);
    }
  }
#pragma unroll
  for (int m = 0; m < MV; ++m)
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      const int c = co0 + 16 * n + l16;
      const float b = a.bias ? a.bias[c] : 0.f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int vox = vbase + 16 * m + 4 * g + i;
        if (vox >= V) continue;
        const size_t o = (size_t)vox * a.co + c;
        float v = acc[m][n][i] + b;
        if (a.skip) v += (float)a.skip[o];
        if (a.act) v = act_f(v);
        if (a.accumulate) v += (float)a.y[o];
        a.y[o] = (TO)v;
      }
    }
}

// The same implicit GEMM for the coarse grids, whose few voxel tiles would
// leave most of the machine idle: a block of KW waves owns one tile of 16
// consecutive voxels and 16*NT output channels and splits the K chunks among
// its waves (wave w takes chunks w, w + KW, ...); the waves' FP32
// accumulators are added in LDS, in wave order, and the block's threads
// apply bias, skip, activation and accumulation to the sums.
template <typename TO, int NT, int KW>
__global__ void __launch_bounds__(64 * KW) conv_kw_kernel(const ConvArgs<TO> a) {
  __shared__ float red[KW][NT][4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int V = a.nx * a.ny * a.nz;
  const int vbase = blockIdx.x * 16;
  const int co0 = blockIdx.y * 16 * NT;
  const int g = lane >> 4, l16 = lane & 15;
  f32x4 acc[NT];
#pragma unroll
  for (int n = 0; n < NT; ++n) acc[n] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int ktot = 27 * a.ci;
  const int nchunk = a.kp >> 5;
  const int vm = vbase + l16;
  const bool vin = vm < V;
  const int xm = vm % a.nx, ym = (vm / a.nx) % a.ny, zm = vm / (a.nx * a.ny);
  const bf16 *wrow[NT];
#pragma unroll
  for (int n = 0; n < NT; ++n) wrow[n] = a.w + (size_t)(co0 + 16 * n + l16) * a.kp;
  for (int kc = wv; kc < nchunk; kc += KW) {
    const int k0 = (kc << 5) + (g << 3);
    bf16x8 bv[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) bv[n] = *(const bf16x8 *)(wrow[n] + k0);
    const int t = k0 >> a.lci;
    const int c0 = k0 & (a.ci - 1);
    bf16x8 av;
    if (k0 < ktot && vin) {
      const int vx = wrapi(xm + t % 3 - 1, a.nx), vy = wrapi(ym + (t / 3) % 3 - 1, a.ny),
                vz = wrapi(zm + t / 9 - 1, a.nz);
      av = *(const bf16x8 *)(a.x + (size_t)(vx + a.nx * (vy + a.ny * vz)) * a.ci + c0);
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) av[q] = (bf16)0.f;
    }
#pragma unroll
    for (int n = 0; n < NT; ++n)
      acc[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv[n], acc[n], 0, 0, 0);
  }
#pragma unroll
  for (int n = 0; n < NT; ++n)
#pragma unroll
    for (int i = 0; i < 4; ++i) red[wv][n][i][lane] = acc[n][i];
  __syncthreads();
  for (int e = threadIdx.x; e < NT * 4 * 64; e += 64 * KW) {
    const int n = e >> 8, i = (e >> 6) & 3, ln = e & 63;
    const int vox = vbase + 4 * (ln >> 4) + i;
    if (vox >= V) continue;
    float v = red[0][n][i][ln];
#pragma unroll
    for (int w = 1; w < KW; ++w) v += red[w][n][i][ln];
    const int c = co0 + 16 * n + (ln & 15);
    const size_t o = (size_t)vox * a.co + c;
    if (a.bias) v += a.bias[c];
    if (a.skip) v += (float)a.skip[o];
    if (a.act) v = act_f(v);
    if (a.accumulate) v += (float)a.y[o];
    a.y[o] = (TO)v;
  }
}

// The same implicit GEMM with its operands staged through LDS, for the fine
// levels.  A block owns a brick of 16 (x) x TY (y) x 4 (z) voxels and all
// 16*NT output channels of its blockIdx.y group; wave w computes the z plane
// w of the brick, TY voxel tiles of 16 (one per x row).
//  * Input: the brick and its one-voxel periodic halo (18 x (TY+2) x 6
//    voxels, every input channel) are loaded once, coalesced, as [channel
//    group of 8][voxel] 16-byte entries, so the 16 lanes that supply one
//    MFMA A fragment (16 consecutive x voxels, 8 channels of one tap) read
//    256 contiguous bytes: one pass over the 64 banks.  Each input value is
//    read once from L2 instead of once per tap.  With UP the input is the
//    grid of half the size, upsampled 2x (nearest) as it is staged, which is
//    the decoder's up-sampling fused into the convolution that consumes it.
//  * Weights: slices of SK K-chunks of the block's 16*NT rows (as many as
//    fit in 48 KB of LDS next to the input tile, at least 8: one slice for
//    16 input channels), loaded cooperatively; the first slice's loads are
//    issued with the input tile's, before one barrier.  Rows are padded by
//    16 bytes so the 16 lanes reading one B fragment (16 rows, same k) hit
//    distinct banks.  The K loop then waits on LDS only, not on a global
//    load per chunk.
// Needs nx % 16 == 0, ny % TY == 0, nz % 4 == 0 and ci == CI (UP: even
// dimensions, x of (nx/2, ny/2, nz/2) voxels).
template <typename TO, int NT, int TY, int CI, bool UP>
__global__ void __launch_bounds__(256) conv_lds_kernel(const ConvArgs<TO> a) {
  constexpr int HX = 18, HY = TY + 2, HZ = 6, NV = HX * HY * HZ, NG = CI / 8;
  // K chunks per weight slice: as many as fit in 48 KB of LDS with the
  // input tile (three blocks per CU), at least 8, at most all of them.
  constexpr int KCH = (27 * CI + 31) / 32;
  constexpr int SKF = ((48 * 1024 - NG * NV * 16) / (NT * 16 * 16) - 1) / 4;
  constexpr int SK = (SKF > 8 ? SKF : 8) < KCH ? (SKF > 8 ? SKF : 8) : KCH;
  static_assert(SK >= 1, "no room for a weight slice");
  constexpr int WROW = SK * 4 + 1;   // 16-byte entries per padded weight row
  __shared__ bf16x8 tile[NG * NV];
  __shared__ bf16x8 wts[NT * 16 * WROW];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int g = lane >> 4, l16 = lane & 15;
  const int ntx = a.nx >> 4, nty = a.ny / TY;
  const int bx = (int)blockIdx.x % ntx, byz = (int)blockIdx.x / ntx;
  const int x0 = bx * 16, y0 = (byz % nty) * TY, z0 = (byz / nty) * 4;
  const int sx = UP ? a.nx >> 1 : a.nx, sy = UP ? a.ny >> 1 : a.ny;
  for (int e = threadIdx.x; e < NG * NV; e += 256) {
    const int cg = e / NV, v = e - cg * NV;
    const int hx = v % HX, hy = (v / HX) % HY, hz = v / (HX * HY);
    int gx = wrapi(x0 + hx - 1, a.nx), gy = wrapi(y0 + hy - 1, a.ny),
        gz = wrapi(z0 + hz - 1, a.nz);
    if (UP) {
      gx >>= 1;
      gy >>= 1;
      gz >>= 1;
    }
    tile[e] = *(const bf16x8 *)(a.x + (size_t)(gx + sx * (gy + (size_t)sy * gz)) * CI + cg * 8);
  }
  const int co0 = blockIdx.y * 16 * NT;
  f32x4 acc[TY][NT];
#pragma unroll
  for (int m = 0; m < TY; ++m)
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
  constexpr int ktot = 27 * CI;
  const int nchunk = a.kp >> 5;
  for (int s0 = 0; s0 < nchunk; s0 += SK) {
    if (s0 > 0) __syncthreads();  // every wave is done with the previous slice
    for (int v = threadIdx.x; v < NT * 16 * SK * 4; v += 256) {
      const int row = v / (SK * 4), c8 = v - row * (SK * 4);
      const int k = s0 * 32 + c8 * 8;
      bf16x8 w8;
      if (k < a.kp) {
        w8 = *(const bf16x8 *)(a.w + (size_t)(co0 + row) * a.kp + k);
      } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) w8[q] = (bf16)0.f;
      }
      wts[row * WROW + c8] = w8;
    }
    __syncthreads();  // the slice (and, the first time, the input tile) is in LDS
    const int ce = min(SK, nchunk - s0);
    for (int c = 0; c < ce; ++c) {
      const int k0 = ((s0 + c) << 5) + (g << 3);
      bf16x8 bv[NT];
#pragma unroll
      for (int n = 0; n < NT; ++n) bv[n] = wts[(16 * n + l16) * WROW + c * 4 + g];
      bf16x8 av[TY];
      if (k0 < ktot) {
        const int t = k0 / CI, cg = (k0 % CI) >> 3;
        const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
        const int base = cg * NV + (l16 + 1 + dx) + HX * ((1 + dy) + HY * (wv + 1 + dz));
#pragma unroll
        for (int m = 0; m < TY; ++m) av[m] = tile[base + HX * m];
      } else {
#pragma unroll
        for (int m = 0; m < TY; ++m)
#pragma unroll
          for (int q = 0; q < 8; ++q) av[m][q] = (bf16)0.f;
      }
#pragma unroll
      for (int m = 0; m < TY; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n)
          acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bv[n], acc[m][n], 0, 0, 0);
    }
  }
  // Epilogue through LDS, one 16-channel output tile at a time: each wave
  // writes its accumulators as [row m][voxel][channel] floats over the input
  // tile (no longer read), then every lane forms 8 consecutive channels of
  // one voxel, so the skip is read and the result written as 16-byte
  // vectors instead of one 2- or 4-byte access per value.  Same operations
  // per value, in the same order, as a per-value epilogue.
  const int z = z0 + wv;
  static_assert(TY * 16 * 16 * 4 * 4 <= NG * NV * 16, "epilogue does not fit the tile");
  float *ob = reinterpret_cast<float *>(tile) + wv * (TY * 16 * 16);
#pragma unroll
  for (int n = 0; n < NT; ++n) {
    __syncthreads();  // the tile (first pass) or the previous pass is done with
#pragma unroll
    for (int m = 0; m < TY; ++m)
#pragma unroll
      for (int i = 0; i < 4; ++i) ob[(m * 16 + 4 * g + i) * 16 + l16] = acc[m][n][i];
    __syncthreads();
    for (int t = lane; t < TY * 16 * 2; t += 64) {
      const int m = t >> 5, v16 = (t >> 1) & 15, c8 = t & 1;
      const int c = co0 + 16 * n + 8 * c8;
      const int vox = x0 + v16 + a.nx * (y0 + m + a.ny * z);
      const size_t o = (size_t)vox * a.co + c;
      const float *src = ob + (m * 16 + v16) * 16 + 8 * c8;
      float v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = src[q] + (a.bias ? a.bias[c + q] : 0.f);
      if (a.skip) {
        const bf16x8 sk = *(const bf16x8 *)(a.skip + o);
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] += (float)sk[q];
      }
      if (a.act) {
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = act_f(v[q]);
      }
      if constexpr (std::is_same<TO, float>::value) {
        f32x4 *yo = (f32x4 *)(a.y + o);
        if (a.accumulate) {
          const f32x4 y0v = yo[0], y1v = yo[1];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            v[q] += y0v[q];
            v[4 + q] += y1v[q];
          }
        }
        yo[0] = f32x4{v[0], v[1], v[2], v[3]};
        yo[1] = f32x4{v[4], v[5], v[6], v[7]};
      } else {
        bf16x8 *yo = (bf16x8 *)(a.y + o);
        if (a.accumulate) {
          const bf16x8 yv = *yo;
#pragma unroll
          for (int q = 0; q < 8; ++q) v[q] += (float)yv[q];
        }
        bf16x8 w;
#pragma unroll
        for (int q = 0; q < 8; ++q) w[q] = (bf16)v[q];
        *yo = w;
      }
    }
  }
}

// LDS bricks of a grid (TY rows per brick) and whether the grid tiles.
inline long long lds_bricks(int nx, int ny, int nz, int ty) {
  if (nx % 16 || ny % ty || nz % 4) return 0;
  return (long long)(nx / 16) * (ny / ty) * (nz / 4);
}

// Shape of the LDS-staged kernel for a convolution: rows per brick (4 while
// that leaves >= 512 blocks, else 1 while >= 256) and output tiles per block
// (the natural 16*NT channels, at most 2 tiles for LDS); ty = 0 when it does
// not apply (64 input channels: one-row bricks only; 128: never; coarser
// grids split the K chunks among the waves of a block, conv_kw_kernel,
// which measured faster there: the 64^3 network's forward pass took 0.246 ms
// with its 16^3 layers staged through LDS one output tile per block, 0.228
// ms with them on conv_kw_kernel).
struct LdsPlan {
  int ty = 0, nt = 0;
};
inline LdsPlan lds_plan(int nx, int ny, int nz, int ci, int co) {
  LdsPlan p;
  if (ci != 16 && ci != 32 && ci != 64) return p;
  const int nat = std::min(2, co >= 64 ? 4 : (co >= 32 ? 2 : 1));
  const long long b4 = ci <= 32 ? lds_bricks(nx, ny, nz, 4) : 0;
  const long long b1 = lds_bricks(nx, ny, nz, 1);
  if (b4 * (co / (16 * nat)) >= 512) {
    p.ty = 4;
    p.nt = nat;
  } else if (b1 * (co / (16 * nat)) >= 256) {
    p.ty = 1;
    p.nt = nat;
  }
  return p;
}

template <typename TO, int NT, int CI, bool UP>
void launch_conv_lds(const ConvArgs<TO> &a, int ty, hipStream_t s) {
  const int gy = a.co / (16 * NT);
  if constexpr (CI <= 32) {
    if (ty == 4) {
      conv_lds_kernel<TO, NT, 4, CI, UP>
          <<<dim3((unsigned)lds_bricks(a.nx, a.ny, a.nz, 4), gy), 256, 0, s>>>(a);
      return;
    }
  }
  conv_lds_kernel<TO, NT, 1, CI, UP>
      <<<dim3((unsigned)lds_bricks(a.nx, a.ny, a.nz, 1), gy), 256, 0, s>>>(a);
}

template <typename TO, bool UP>
void launch_conv_lds_any(const ConvArgs<TO> &a, LdsPlan p, hipStream_t s) {
  if (a.ci == 16 && p.nt == 1) launch_conv_lds<TO, 1, 16, UP>(a, p.ty, s);
  else if (a.ci == 16) launch_conv_lds<TO, 2, 16, UP>(a, p.ty, s);
  else if (a.ci == 32 && p.nt == 1) launch_conv_lds<TO, 1, 32, UP>(a, p.ty, s);
  else if (a.ci == 32) launch_conv_lds<TO, 2, 32, UP>(a, p.ty, s);
  else if (p.nt == 1) launch_conv_lds<TO, 1, 64, UP>(a, p.ty, s);
  else launch_conv_lds<TO, 2, 64, UP>(a, p.ty, s);
}

// Blocks a convolution launches with one 16-voxel tile per wave.
inline int conv_blocks(int V, int co) {
  const int nt = co >= 64 ? 4 : (co >= 32 ? 2 : 1);
  return ((V + 63) / 64) * (co / (16 * nt));
}

// Waves per block of the coarse-grid kernel.
constexpr int kConvKW = 8;

template <typename TO, int NT>
hipError_t launch_conv_nt(const ConvArgs<TO> &a, int V, hipStream_t s) {
  // Four voxel tiles per wave (four independent gathers per K chunk) once
  // the grid still has >= 1024 blocks (4096 waves); one tile per wave while
  // it has >= 1024 blocks that way; on coarser grids, the K chunks split
  // among the waves of a block.
  constexpr int kMV = 4;
  const int gy = a.co / (16 * NT);
  if ((long long)((V + 64 * kMV - 1) / (64 * kMV)) * gy >= 1024) {
    dim3 grid((V + 64 * kMV - 1) / (64 * kMV), gy);
    conv_kernel<TO, NT, kMV><<<grid, 256, 0, s>>>(a);
  } else if (conv_blocks(V, a.co) >= 1024) {
    dim3 grid((V + 63) / 64, gy);
    conv_kernel<TO, NT, 1><<<grid, 256, 0, s>>>(a);
  } else {
    dim3 grid((V + 15) / 16, gy);
    conv_kw_kernel<TO, NT, kConvKW><<<grid, 64 * kConvKW, 0, s>>>(a);
  }
  return hipGetLastError();
}

template <typename TO>
hipError_t launch_conv(const ConvArgs<TO> &a, hipStream_t s) {
  const int V = a.nx * a.ny * a.nz;
  const int nt = a.co >= 64 ? 4 : (a.co >= 32 ? 2 : 1);
  // Fine levels (16, 32 or 64 input channels, enough blocks): the LDS-staged
  // kernel.
  const LdsPlan lp = lds_plan(a.nx, a.ny, a.nz, a.ci, a.co);
  if (lp.ty) {
    launch_conv_lds_any<TO, false>(a, lp, s);
    return hipGetLastError();
  }
  if (nt == 4) return launch_conv_nt<TO, 4>(a, V, s);
  if (nt == 2) return launch_conv_nt<TO, 2>(a, V, s);
  return launch_conv_nt<TO, 1>(a, V, s);
}

// Weight gradient: dW[co][tap*ci + c] = sum_v dz[v][co] * x[v + off(tap)][c],
// one partial sum per (slab of voxels) written to part[slab][entry].
__global__ void __launch_bounds__(256) wgrad_kernel(const bf16 *__restrict__ dz,
                                                    const bf16 *__restrict__ x, float *part,
                                                    int nx, int ny, int nz, int ci, int co,
                                                    int lci, int vper) {
  const int entries = co * 27 * ci;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= entries) return;
  const int c = e & (ci - 1);
  const int t = (e >> lci) % 27;
  const int o = e / (27 * ci);
  const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dzz = t / 9 - 1;
  const int V = nx * ny * nz;
  const int v0 = blockIdx.y * vper, v1 = min(V, v0 + vper);
  float s = 0.f;
  for (int v = v0; v < v1; ++v) {
    const int ix = v % nx, iy = (v / nx) % ny, iz = v / (nx * ny);
    const int u = wrapi(ix + dx, nx) + nx * (wrapi(iy + dy, ny) + ny * wrapi(iz + dzz, nz));
    s += (float)dz[(size_t)v * co + o] * (float)x[(size_t)u * ci + c];
  }
  part[(size_t)blockIdx.y * entries + e] = s;
}

// dst[e] += sum_s part[s][e]
__global__ void __launch_bounds__(256) reduce_parts_kernel(const float *part, int nparts,
                                                           int entries, float *dst) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= entries) return;
  float s = 0.f;
  for (int p = 0; p < nparts; ++p) s += part[(size_t)p * entries + e];
  dst[e] += s;
}

// ------------------------------------------------------- elementwise pieces
// x[v] = (in[0][v], in[1][v], 0, ..., 0), C a multiple of 8 (16-byte stores).
__global__ void pack_input_kernel(const float *__restrict__ in, bf16 *__restrict__ x, int V,
                                  int C) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  bf16x8 *row = (bf16x8 *)(x + (size_t)v * C);
  bf16x8 z;
#pragma unroll
  for (int q = 0; q < 8; ++q) z[q] = (bf16)0.f;
  bf16x8 h = z;
  h[0] = (bf16)in[v];
  h[1] = (bf16)in[V + v];
  row[0] = h;
  for (int g = 1; g < C / 8; ++g) row[g] = z;
}

// The first convolution in the forward pass: only 2 of its 16 (padded)
// input channels are non-zero, so its K is made dense: k = 2*tap + channel
// over the 54 real pairs, zero-padded to 64 (two MFMA K-chunks instead of
// the padded layout's 14).  A wave computes kFirstTiles tiles of 16
// consecutive voxels x 16 output channels (the block's weight staging
// serves 64 * kFirstTiles voxels); lane (l16, g) gathers, per chunk, the
// two channels of taps 16*chunk + 4g .. +3 of voxel l16 straight from the
// network's FP32 input planes (the 16 lanes of a tap read 16 consecutive
// floats), rounded to BF16 as the packed input holds them.  The dense
// weights are staged in LDS from the padded layout (rows padded by 16
// bytes: the 16 lanes of a B fragment hit distinct banks); bias and leaky
// ReLU follow.
constexpr int kFirstTiles = 4;
__global__ void __launch_bounds__(256) conv_first_kernel(const float *__restrict__ in,
                                                         const bf16 *__restrict__ w,
                                                         const float *__restrict__ bias,
                                                         bf16 *__restrict__ y, int nx, int ny,
                                                         int nz, int kp) {
  __shared__ bf16 wd[16][64 + 8];
  for (int e = threadIdx.x; e < 16 * 64; e += blockDim.x) {
    const int co = e >> 6, k = e & 63;
    wd[co][k] = k < 54 ? w[(size_t)co * kp + (k >> 1) * 16 + (k & 1)] : (bf16)0.f;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int g = lane >> 4, l16 = lane & 15;
  const int V = nx * ny * nz;
  const float b = bias[l16];
  for (int tt = 0; tt < kFirstTiles; ++tt) {
    const int vbase = ((blockIdx.x * 4 + wv) * kFirstTiles + tt) * 16;
    if (vbase >= V) return;  // after the kernel's only barrier
    const int vm = vbase + l16;
    const bool vin = vm < V;
    const int i = vm % nx, j = (vm / nx) % ny, k = vm / (nx * ny);
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) {
      bf16x8 av;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int t = 16 * kc + 4 * g + q;
        bf16x2 pr;
        pr[0] = pr[1] = (bf16)0.f;
        if (t < 27 && vin) {
          const int vx = wrapi(i + t % 3 - 1, nx), vy = wrapi(j + (t / 3) % 3 - 1, ny),
                    vz = wrapi(k + t / 9 - 1, nz);
          const size_t u = (size_t)vx + (size_t)nx * (vy + (size_t)ny * vz);
          pr[0] = (bf16)in[u];
          pr[1] = (bf16)in[(size_t)V + u];
        }
        av[2 * q] = pr[0];
        av[2 * q + 1] = pr[1];
      }
      const bf16x8 bv = *(const bf16x8 *)&wd[l16][32 * kc + 8 * g];
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
    // acc[q] = C[voxel vbase + 4g + q][channel l16]
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int vox = vbase + 4 * g + q;
      if (vox < V) y[(size_t)vox * 16 + l16] = (bf16)act_f(acc[q] + b);
    }
  }
}

// 2x average pooling onto the grid (nx/2, ny/2, nz/2): one thread per
// coarse voxel and group of 8 channels (16-byte loads and stores), each
// channel's 8 children summed in the same order.
__global__ void pool_fwd_kernel(const bf16 *__restrict__ in, bf16 *__restrict__ out, int nx,
                                int ny, int nz, int C) {
  const int mx = nx / 2, my = ny / 2, mz = nz / 2;
  const int G = C / 8;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)mx * my * mz * G) return;
  const int cg = idx % G;
  const int v = idx / G;
  const int i = v % mx, j = (v / mx) % my, k = v / (mx * my);
  bf16x8 ch[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int u = (2 * i + (q & 1)) + nx * ((2 * j + ((q >> 1) & 1)) + ny * (2 * k + (q >> 2)));
    ch[q] = *(const bf16x8 *)(in + (size_t)u * C + 8 * cg);
  }
  bf16x8 o;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    float s = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) s += (float)ch[q][c];
    o[c] = (bf16)(0.125f * s);
  }
  *(bf16x8 *)(out + (size_t)v * C + 8 * cg) = o;
}

// g_in (fine, FP32) += g_out(coarse)/8
__global__ void pool_bwd_kernel(const float *__restrict__ gout, float *__restrict__ gin, int nx,
                                int ny, int nz, int C) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)nx * ny * nz * C) return;
  const int c = idx % C;
  const int v = idx / C;
  const int i = v % nx, j = (v / nx) % ny, k = v / (nx * ny);
  const int p = i / 2 + (nx / 2) * (j / 2 + (ny / 2) * (k / 2));
  gin[idx] += 0.125f * gout[(size_t)p * C + c];
}

// nearest up-sampling: fine (nx, ny, nz) from coarse (nx/2, ...), one thread
// per fine voxel and group of 8 channels.
__global__ void up_fwd_kernel(const bf16 *__restrict__ in, bf16 *__restrict__ out, int nx,
                              int ny, int nz, int C) {
  const int G = C / 8;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)nx * ny * nz * G) return;
  const int cg = idx % G;
  const int v = idx / G;
  const int i = v % nx, j = (v / nx) % ny, k = v / (nx * ny);
  const int p = i / 2 + (nx / 2) * (j / 2 + (ny / 2) * (k / 2));
  *(bf16x8 *)(out + (size_t)v * C + 8 * cg) = *(const bf16x8 *)(in + (size_t)p * C + 8 * cg);
}

// g_in (coarse) += sum of the 8 fine children
__global__ void up_bwd_kernel(const float *__restrict__ gout, float *__restrict__ gin, int nx,
                              int ny, int nz, int C) {
  const int mx = nx / 2, my = ny / 2, mz = nz / 2;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)mx * my * mz * C) return;
  const int c = idx % C;
  const int v = idx / C;
  const int i = v % mx, j = (v / mx) % my, k = v / (mx * my);
  float s = 0.f;
  for (int q = 0; q < 8; ++q) {
    const int u = (2 * i + (q & 1)) + nx * ((2 * j + ((q >> 1) & 1)) + ny * (2 * k + (q >> 2)));
    s += gout[(size_t)u * C + c];
  }
  gin[idx] += s;
}

// dz = bf16(g * act'(y)); gskip += g * act'(y); db[c] += sum_v dz (block
// partial per channel, atomics on the FP32 bias gradient).
__global__ void __launch_bounds__(256) act_bwd_kernel(const float *__restrict__ g,
                                                      const bf16 *__restrict__ y,
                                                      bf16 *__restrict__ dz, float *gskip,
                                                      float *db, int V, int C) {
  __shared__ float sb[128];
  for (int c = threadIdx.x; c < C; c += blockDim.x) sb[c] = 0.f;
  __syncthreads();
  const size_t n = (size_t)V * C;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (size_t)gridDim.x * blockDim.x) {
    const float d = g[idx] * dact((float)y[idx]);
    dz[idx] = (bf16)d;
    if (gskip) gskip[idx] += d;
    atomicAdd(&sb[idx % C], d);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x) atomicAdd(&db[c], sb[c]);
}

// out[v] = sum_c wh[c] d0[v][c] + bh + id[v]: the network learns a
// correction to its second input plane (the previous increment).
// C a multiple of 8: 16-byte loads of d0, the channels summed in order.
__global__ void head_fwd_kernel(const bf16 *__restrict__ d0, const float *__restrict__ wh,
                                const float *__restrict__ id, float *__restrict__ out, int V,
                                int C) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  float s = wh[C];
  const bf16x8 *row = (const bf16x8 *)(d0 + (size_t)v * C);
  for (int gq = 0; gq < C / 8; ++gq) {
    const bf16x8 h = row[gq];
#pragma unroll
    for (int q = 0; q < 8; ++q) s += wh[8 * gq + q] * (float)h[q];
  }
  out[v] = s + id[v];
}

__global__ void __launch_bounds__(256) head_bwd_kernel(const float *__restrict__ dout,
                                                       const bf16 *__restrict__ d0,
                                                       const float *__restrict__ wh,
                                                       float *__restrict__ gd0,
                                                       float *gwh, int V, int C) {
  __shared__ float sw[17];
  if (threadIdx.x <= C) sw[threadIdx.x] = 0.f;
  __syncthreads();
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
    const float d = dout[v];
    for (int c = 0; c < C; ++c) {
      gd0[(size_t)v * C + c] += d * wh[c];
      atomicAdd(&sw[c], d * (float)d0[(size_t)v * C + c]);
    }
    atomicAdd(&sw[C], d);
  }
  __syncthreads();
  if (threadIdx.x <= C) atomicAdd(&gwh[threadIdx.x], sw[threadIdx.x]);
}

__global__ void __launch_bounds__(256) mse_kernel(const float *__restrict__ out,
                                                  const float *__restrict__ tgt, float scale,
                                                  float *__restrict__ dout, double *part, int V) {
  __shared__ double sc[4];
  double s = 0.0;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
    const float d = out[v] - tgt[v];
    s += (double)d * d;
    if (dout) dout[v] = scale * 2.f * d / (float)V;
  }
  s = block_sum<4>(s, sc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void adam_kernel(float *w, const float *g, float *m, float *v, size_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 copies of one conv layer's master weights W[co][27*ci]:
// forward layout wf[co][kpf] and backward layout wb[ci][kpb] with
// wb[c][t'*co + o] = W[o][(26 - t')*ci + c].
__global__ void sync_weights_kernel(const float *W, bf16 *wf, bf16 *wb, int ci, int co, int kpf,
                                    int kpb) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int nf = co * kpf, nb = ci * kpb;
  if (e < nf) {
    const int o = e / kpf, k = e % kpf;
    wf[e] = (bf16)(k < 27 * ci ? W[(size_t)o * 27 * ci + k] : 0.f);
  }
  if (e < nb) {
    const int c = e / kpb, k = e % kpb;
    float v = 0.f;
    if (k < 27 * co) {
      const int tp = k / co, o = k % co;
      v = W[(size_t)o * 27 * ci + (26 - tp) * ci + c];
    }
    wb[e] = (bf16)v;
  }
}

// ------------------------------------------------------------------ network
struct Conv {
  int ci, co, level;
  size_t woff, boff;  // offsets into the flat parameter array
  int kpf, kpb;
  bf16 *wf = nullptr, *wb = nullptr;
};

int ilog2(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

struct Net {
  int nx[4], ny[4], nz[4];
  size_t V[4];
  std::vector<Conv> conv;
  size_t nparam = 0, head_off = 0;
  float *W = nullptr, *G = nullptr, *M = nullptr, *Vv = nullptr;  // FP32 master, grad, Adam
  bf16 *wbuf = nullptr;  // BF16 copies of all conv weights
  // activations (bf16)
  bf16 *xin, *e0a, *e0, *p1, *e1a, *e1, *p2, *e2a, *e2, *p3, *e3a, *e3, *u2, *d2a, *d2, *u1,
      *d1a, *d1, *u0, *d0a, *d0;
  bf16 *dz = nullptr;      // scratch for act gradients (largest level)
  float *gbuf = nullptr;   // FP32 activation gradients (one arena)
  float *g_d0, *g_d0a, *g_u0, *g_d1, *g_d1a, *g_u1, *g_d2, *g_d2a, *g_u2, *g_e3, *g_e3a, *g_p3,
      *g_e2, *g_e2a, *g_p2, *g_e1, *g_e1a, *g_p1, *g_e0, *g_e0a;
  size_t gbytes = 0;
  float *part = nullptr;   // wgrad partial sums
  size_t part_n = 0;
  double *mse_part = nullptr;
  double *hpart = nullptr;
  void *act_arena = nullptr;
  hipEvent_t ev[2];
  double last_ms = 0.0, last_flops = 0.0;
  // u0 / u1 (the up-sampled d1 / d2) not formed by the last forward pass:
  // its convolutions 12 / 10 read d1 / d2 up-sampled as they staged them.
  bool up_stale[2] = {false, false};
  // The input of the last forward pass (its caller keeps it for the
  // backward pass), not yet packed into xin.
  const float *fwd_in = nullptr;
  bool xin_stale = false;
  std::vector<void *> allocs;

  ~Net() {
    for (void *p : allocs) (void)hipFree(p);
    if (act_arena) (void)hipFree(act_arena);
    if (gbuf) (void)hipFree(gbuf);
    if (part) (void)hipFree(part);
    if (hpart) (void)hipHostFree(hpart);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
  }
};

#define NNCHK(...)                                                                       \
  do {                                                                                   \
    hipError_t e_ = (__VA_ARGS__);                                                       \
    if (e_ != hipSuccess) return o3d::set_error(hipGetErrorString(e_));                  \
  } while (0)
#define NCHK(x)              \
  do {                       \
    int r_ = (x);            \
    if (r_ != 0) return r_;  \
  } while (0)


void build_layers(Net &n) {
  // (ci, co, level): encoder, bottleneck, decoder (14 convolutions).
  const int spec[14][3] = {{16, 16, 0},  {16, 16, 0},  {16, 32, 1},  {32, 32, 1},  {32, 64, 2},
                           {64, 64, 2},  {64, 128, 3}, {128, 128, 3}, {128, 64, 2}, {64, 64, 2},
                           {64, 32, 1},  {32, 32, 1},  {32, 16, 0},  {16, 16, 0}};
  size_t off = 0;
  for (auto &s : spec) {
    Conv c;
    c.ci = s[0];
    c.co = s[1];
    c.level = s[2];
    c.woff = off;
    off += (size_t)c.co * 27 * c.ci;
    c.boff = off;
    off += c.co;
    c.kpf = ((27 * c.ci + 31) / 32) * 32;
    c.kpb = ((27 * c.co + 31) / 32) * 32;
    n.conv.push_back(c);
  }
  n.head_off = off;
  off += 17;
  n.nparam = off;
}

int alloc_activations(Net &n, int nx, int ny, int nz) {
  if (nx % 8 || ny % 8 || nz % 8)
    return o3d::set_error("network grid: every axis must be a multiple of 8");
  for (int l = 0; l < 4; ++l) {
    n.nx[l] = nx >> l;
    n.ny[l] = ny >> l;
    n.nz[l] = nz >> l;
    n.V[l] = (size_t)n.nx[l] * n.ny[l] * n.nz[l];
  }
  if (n.act_arena) { (void)hipFree(n.act_arena); n.act_arena = nullptr; }
  n.fwd_in = nullptr;  // a backward pass needs a forward pass on the new grid
  n.xin_stale = false;
  n.up_stale[0] = n.up_stale[1] = false;
  if (n.gbuf) { (void)hipFree(n.gbuf); n.gbuf = nullptr; }
  if (n.part) { (void)hipFree(n.part); n.part = nullptr; }
  // bf16 activations: channels per level.
  struct A { bf16 **p; int level, ch; };
  const A acts[] = {{&n.xin, 0, 16}, {&n.e0a, 0, 16}, {&n.e0, 0, 16}, {&n.p1, 1, 16},
                    {&n.e1a, 1, 32}, {&n.e1, 1, 32},  {&n.p2, 2, 32}, {&n.e2a, 2, 64},
                    {&n.e2, 2, 64},  {&n.p3, 3, 64},  {&n.e3a, 3, 128}, {&n.e3, 3, 128},
                    {&n.u2, 2, 128}, {&n.d2a, 2, 64}, {&n.d2, 2, 64}, {&n.u1, 1, 64},
                    {&n.d1a, 1, 32}, {&n.d1, 1, 32},  {&n.u0, 0, 32}, {&n.d0a, 0, 16},
                    {&n.d0, 0, 16}, {&n.dz, 0, 32}};
  size_t total = 0;
  for (const A &a : acts) total += ((n.V[a.level] * a.ch * sizeof(bf16) + 255) / 256) * 256;
  NNCHK(hipMalloc(&n.act_arena, total));
  char *p = (char *)n.act_arena;
  for (const A &a : acts) {
    *a.p = (bf16 *)p;
    p += ((n.V[a.level] * a.ch * sizeof(bf16) + 255) / 256) * 256;
  }
  struct G { float **p; int level, ch; };
  const G grads[] = {{&n.g_d0, 0, 16}, {&n.g_d0a, 0, 16}, {&n.g_u0, 0, 32}, {&n.g_d1, 1, 32},
                     {&n.g_d1a, 1, 32}, {&n.g_u1, 1, 64}, {&n.g_d2, 2, 64}, {&n.g_d2a, 2, 64},
                     {&n.g_u2, 2, 128}, {&n.g_e3, 3, 128}, {&n.g_e3a, 3, 128}, {&n.g_p3, 3, 64},
                     {&n.g_e2, 2, 64}, {&n.g_e2a, 2, 64}, {&n.g_p2, 2, 32}, {&n.g_e1, 1, 32},
                     {&n.g_e1a, 1, 32}, {&n.g_p1, 1, 16}, {&n.g_e0, 0, 16}, {&n.g_e0a, 0, 16}};
  total = 0;
  for (const G &g : grads) total += ((n.V[g.level] * g.ch * 4 + 255) / 256) * 256;
  NNCHK(hipMalloc((void **)&n.gbuf, total));
  n.gbytes = total;
  p = (char *)n.gbuf;
  for (const G &g : grads) {
    *g.p = (float *)p;
    p += ((n.V[g.level] * g.ch * 4 + 255) / 256) * 256;
  }
  // wgrad partials: entries * slabs, slabs ~ V / 4096 at level 0.
  size_t maxe = 0;
  for (const Conv &c : n.conv) maxe = std::max(maxe, (size_t)c.co * 27 * c.ci);
  const size_t slabs = std::max<size_t>(1, std::min<size_t>(512, n.V[0] / 4096));
  n.part_n = maxe * slabs;
  NNCHK(hipMalloc((void **)&n.part, n.part_n * 4));
  return 0;
}

hipStream_t st() { return o3d::current_stream(); }

int sync_weights(Net &n) {
  for (const Conv &c : n.conv) {
    const int e = std::max(c.co * c.kpf, c.ci * c.kpb);
    sync_weights_kernel<<<(e + 255) / 256, 256, 0, st()>>>(n.W + c.woff, c.wf, c.wb, c.ci, c.co,
                                                           c.kpf, c.kpb);
  }
  NNCHK(hipGetLastError());
  return 0;
}

// Deterministic He-style initialisation (splitmix64 -> uniform); the head
// starts small so the untrained network returns about its identity input.
void init_params(std::vector<float> &w, const Net &n, uint64_t seed) {
  uint64_t s = seed;
  auto next = [&]() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) / 9007199254740992.0;
  };
  std::fill(w.begin(), w.end(), 0.f);
  for (size_t li = 0; li < n.conv.size(); ++li) {
    const Conv &c = n.conv[li];
    const int fan_in = 27 * (li == 0 ? 2 : c.ci);
    const double bound = sqrt(6.0 / fan_in) * 0.8;
    for (int o = 0; o < c.co; ++o)
      for (int t = 0; t < 27; ++t)
        for (int ci = 0; ci < c.ci; ++ci) {
          if (li == 0 && ci >= 2) continue;  // padded input channels stay zero
          w[c.woff + (size_t)o * 27 * c.ci + t * c.ci + ci] = (float)((2 * next() - 1) * bound);
        }
  }
  for (int c = 0; c < 16; ++c) w[n.head_off + c] = (float)((2 * next() - 1) * 0.02);
}

// The first convolution from the network's FP32 input planes.
int conv_first_fwd(Net &n, const float *in, bf16 *y) {
  const Conv &c = n.conv[0];
  if (c.ci != 16 || c.co != 16) return o3d::set_error("first convolution: 16 -> 16 expected");
  const int V = (int)n.V[c.level];
  conv_first_kernel<<<(V + 64 * kFirstTiles - 1) / (64 * kFirstTiles), 256, 0, st()>>>(
      in, c.wf, n.W + c.boff, y,
                                                       n.nx[c.level], n.ny[c.level],
                                                       n.nz[c.level], c.kpf);
  NNCHK(hipGetLastError());
  n.last_flops += 54.0 * 2 * c.co * (double)V;  // the two real input channels
  return 0;
}

int conv_fwd(Net &n, int li, const bf16 *x, const bf16 *skip, bf16 *y) {
  const Conv &c = n.conv[li];
  ConvArgs<bf16> a;
  a.x = x; a.w = c.wf; a.bias = n.W + c.boff; a.skip = skip; a.y = y;
  a.nx = n.nx[c.level]; a.ny = n.ny[c.level]; a.nz = n.nz[c.level];
  a.lci = ilog2(c.ci); a.ci = c.ci; a.co = c.co; a.kp = c.kpf;
  a.act = 1; a.accumulate = 0;
  NNCHK(launch_conv<bf16>(a, st()));
  n.last_flops += 54.0 * c.ci * c.co * (double)n.V[c.level];
  return 0;
}

// Convolution li of the decoder on xc up-sampled 2x (nearest), fused into
// the LDS-staged kernel's staging; false (nothing launched) when that kernel
// does not apply to the layer's grid.
bool conv_fwd_up(Net &n, int li, const bf16 *xc, const bf16 *skip, bf16 *y, int *rc) {
  const Conv &c = n.conv[li];
  const int L = c.level;
  const LdsPlan lp = lds_plan(n.nx[L], n.ny[L], n.nz[L], c.ci, c.co);
  *rc = 0;
  if (!lp.ty) return false;
  ConvArgs<bf16> a;
  a.x = xc; a.w = c.wf; a.bias = n.W + c.boff; a.skip = skip; a.y = y;
  a.nx = n.nx[L]; a.ny = n.ny[L]; a.nz = n.nz[L];
  a.lci = ilog2(c.ci); a.ci = c.ci; a.co = c.co; a.kp = c.kpf;
  a.act = 1; a.accumulate = 0;
  launch_conv_lds_any<bf16, true>(a, lp, st());
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) *rc = o3d::set_error(hipGetErrorString(e));
  n.last_flops += 54.0 * c.ci * c.co * (double)n.V[L];
  return true;
}

// Backward of y = act(conv(x) + b + skip): gy (FP32) -> dz; gskip += dz;
// db, dW += ...; gx += conv^T(dz) when gx is given.
int conv_bwd(Net &n, int li, const float *gy, const bf16 *y, const bf16 *x, float *gskip,
             float *gx) {
  const Conv &c = n.conv[li];
  const int L = c.level;
  const int V = (int)n.V[L];
  const size_t nel = (size_t)V * c.co;
  const int blocks = (int)std::min<size_t>(2048, (nel + 255) / 256);
  act_bwd_kernel<<<blocks, 256, 0, st()>>>(gy, y, n.dz, gskip, n.G + c.boff, V, c.co);
  NNCHK(hipGetLastError());
  // Weight gradient.
  const int entries = c.co * 27 * c.ci;
  int slabs = (int)std::max<size_t>(1, std::min<size_t>(n.part_n / entries, (size_t)V / 2048));
  slabs = std::max(1, std::min(slabs, 512));
  const int vper = (V + slabs - 1) / slabs;
  dim3 g1((entries + 255) / 256, slabs);
  wgrad_kernel<<<g1, 256, 0, st()>>>(n.dz, x, n.part, n.nx[L], n.ny[L], n.nz[L], c.ci, c.co,
                                      ilog2(c.ci), vper);
  NNCHK(hipGetLastError());
  reduce_parts_kernel<<<(entries + 255) / 256, 256, 0, st()>>>(n.part, slabs, entries,
                                                                n.G + c.woff);
  NNCHK(hipGetLastError());
  if (gx) {
    ConvArgs<float> a;
    a.x = n.dz; a.w = c.wb; a.bias = nullptr; a.skip = nullptr; a.y = gx;
    a.nx = n.nx[L]; a.ny = n.ny[L]; a.nz = n.nz[L];
    a.lci = ilog2(c.co); a.ci = c.co; a.co = c.ci; a.kp = c.kpb;
    a.act = 0; a.accumulate = 1;
    NNCHK(launch_conv<float>(a, st()));
  }
  return 0;
}

int elem_blocks(size_t n) { return (int)((n + 255) / 256); }

int forward(Net &n, const float *in, float *out) {
  n.last_flops = 0.0;
  NNCHK(hipEventRecord(n.ev[0], st()));
  // The packed BF16 input (the first layer's weight gradient reads it) is
  // formed only when a backward pass follows.
  n.fwd_in = in;
  n.xin_stale = true;
  NCHK(conv_first_fwd(n, in, n.e0a));
  NCHK(conv_fwd(n, 1, n.e0a, nullptr, n.e0));
  pool_fwd_kernel<<<elem_blocks(n.V[1] * 2), 256, 0, st()>>>(n.e0, n.p1, n.nx[0], n.ny[0],
                                                               n.nz[0], 16);
  NCHK(conv_fwd(n, 2, n.p1, nullptr, n.e1a));
  NCHK(conv_fwd(n, 3, n.e1a, nullptr, n.e1));
  pool_fwd_kernel<<<elem_blocks(n.V[2] * 4), 256, 0, st()>>>(n.e1, n.p2, n.nx[1], n.ny[1],
                                                               n.nz[1], 32);
  NCHK(conv_fwd(n, 4, n.p2, nullptr, n.e2a));
  NCHK(conv_fwd(n, 5, n.e2a, nullptr, n.e2));
  pool_fwd_kernel<<<elem_blocks(n.V[3] * 8), 256, 0, st()>>>(n.e2, n.p3, n.nx[2], n.ny[2],
                                                               n.nz[2], 64);
  NCHK(conv_fwd(n, 6, n.p3, nullptr, n.e3a));
  NCHK(conv_fwd(n, 7, n.e3a, nullptr, n.e3));
  up_fwd_kernel<<<elem_blocks(n.V[2] * 16), 256, 0, st()>>>(n.e3, n.u2, n.nx[2], n.ny[2],
                                                              n.nz[2], 128);
  NCHK(conv_fwd(n, 8, n.u2, n.e2, n.d2a));
  NCHK(conv_fwd(n, 9, n.d2a, nullptr, n.d2));
  int rc = 0;
  n.up_stale[1] = conv_fwd_up(n, 10, n.d2, n.e1, n.d1a, &rc);
  NCHK(rc);
  if (!n.up_stale[1]) {
    up_fwd_kernel<<<elem_blocks(n.V[1] * 8), 256, 0, st()>>>(n.d2, n.u1, n.nx[1], n.ny[1],
                                                               n.nz[1], 64);
    NCHK(conv_fwd(n, 10, n.u1, n.e1, n.d1a));
  }
  NCHK(conv_fwd(n, 11, n.d1a, nullptr, n.d1));
  n.up_stale[0] = conv_fwd_up(n, 12, n.d1, n.e0, n.d0a, &rc);
  NCHK(rc);
  if (!n.up_stale[0]) {
    up_fwd_kernel<<<elem_blocks(n.V[0] * 4), 256, 0, st()>>>(n.d1, n.u0, n.nx[0], n.ny[0],
                                                               n.nz[0], 32);
    NCHK(conv_fwd(n, 12, n.u0, n.e0, n.d0a));
  }
  NCHK(conv_fwd(n, 13, n.d0a, nullptr, n.d0));
  head_fwd_kernel<<<elem_blocks(n.V[0]), 256, 0, st()>>>(n.d0, n.W + n.head_off, in + n.V[0],
                                                          out, (int)n.V[0], 16);
  NNCHK(hipGetLastError());
  NNCHK(hipEventRecord(n.ev[1], st()));
  return 0;
}

int backward(Net &n, const float *dout) {
  NNCHK(hipMemsetAsync(n.gbuf, 0, n.gbytes, st()));
  head_bwd_kernel<<<(int)std::min<size_t>(1024, (n.V[0] + 255) / 256), 256, 0, st()>>>(
      dout, n.d0, n.W + n.head_off, n.g_d0, n.G + n.head_off, (int)n.V[0], 16);
  NNCHK(hipGetLastError());
  NCHK(conv_bwd(n, 13, n.g_d0, n.d0, n.d0a, nullptr, n.g_d0a));
  if (n.up_stale[0]) {  // the forward pass read d1 up-sampled without forming u0
    up_fwd_kernel<<<elem_blocks(n.V[0] * 4), 256, 0, st()>>>(n.d1, n.u0, n.nx[0], n.ny[0],
                                                               n.nz[0], 32);
    NNCHK(hipGetLastError());
    n.up_stale[0] = false;
  }
  NCHK(conv_bwd(n, 12, n.g_d0a, n.d0a, n.u0, n.g_e0, n.g_u0));
  up_bwd_kernel<<<elem_blocks(n.V[1] * 32), 256, 0, st()>>>(n.g_u0, n.g_d1, n.nx[0], n.ny[0],
                                                             n.nz[0], 32);
  NCHK(conv_bwd(n, 11, n.g_d1, n.d1, n.d1a, nullptr, n.g_d1a));
  if (n.up_stale[1]) {
    up_fwd_kernel<<<elem_blocks(n.V[1] * 8), 256, 0, st()>>>(n.d2, n.u1, n.nx[1], n.ny[1],
                                                               n.nz[1], 64);
    NNCHK(hipGetLastError());
    n.up_stale[1] = false;
  }
  NCHK(conv_bwd(n, 10, n.g_d1a, n.d1a, n.u1, n.g_e1, n.g_u1));
  up_bwd_kernel<<<elem_blocks(n.V[2] * 64), 256, 0, st()>>>(n.g_u1, n.g_d2, n.nx[1], n.ny[1],
                                                             n.nz[1], 64);
  NCHK(conv_bwd(n, 9, n.g_d2, n.d2, n.d2a, nullptr, n.g_d2a));
  NCHK(conv_bwd(n, 8, n.g_d2a, n.d2a, n.u2, n.g_e2, n.g_u2));
  up_bwd_kernel<<<elem_blocks(n.V[3] * 128), 256, 0, st()>>>(n.g_u2, n.g_e3, n.nx[2], n.ny[2],
                                                              n.nz[2], 128);
  NCHK(conv_bwd(n, 7, n.g_e3, n.e3, n.e3a, nullptr, n.g_e3a));
  NCHK(conv_bwd(n, 6, n.g_e3a, n.e3a, n.p3, nullptr, n.g_p3));
  pool_bwd_kernel<<<elem_blocks(n.V[2] * 64), 256, 0, st()>>>(n.g_p3, n.g_e2, n.nx[2], n.ny[2],
                                                               n.nz[2], 64);
  NCHK(conv_bwd(n, 5, n.g_e2, n.e2, n.e2a, nullptr, n.g_e2a));
  NCHK(conv_bwd(n, 4, n.g_e2a, n.e2a, n.p2, nullptr, n.g_p2));
  pool_bwd_kernel<<<elem_blocks(n.V[1] * 32), 256, 0, st()>>>(n.g_p2, n.g_e1, n.nx[1], n.ny[1],
                                                               n.nz[1], 32);
  NCHK(conv_bwd(n, 3, n.g_e1, n.e1, n.e1a, nullptr, n.g_e1a));
  NCHK(conv_bwd(n, 2, n.g_e1a, n.e1a, n.p1, nullptr, n.g_p1));
  pool_bwd_kernel<<<elem_blocks(n.V[0] * 16), 256, 0, st()>>>(n.g_p1, n.g_e0, n.nx[0], n.ny[0],
                                                               n.nz[0], 16);
  NCHK(conv_bwd(n, 1, n.g_e0, n.e0, n.e0a, nullptr, n.g_e0a));
  if (n.xin_stale) {
    pack_input_kernel<<<elem_blocks(n.V[0]), 256, 0, st()>>>(n.fwd_in, n.xin, (int)n.V[0], 16);
    NNCHK(hipGetLastError());
    n.xin_stale = false;
  }
  NCHK(conv_bwd(n, 0, n.g_e0a, n.e0a, n.xin, nullptr, nullptr));
  NNCHK(hipGetLastError());
  return 0;
}

}  // namespace nn
}  // namespace o3d

using namespace o3d::nn;

extern "C" int o3d_nn_create(int nx, int ny, int nz, unsigned long long seed, void **out) {
  if (!out) return o3d::set_error("NULL argument");
  Net *n = new Net();
  build_layers(*n);
  int rc = 0;
  auto al = [&](void **p, size_t bytes) {
    if (rc) return;
    if (hipMalloc(p, bytes) != hipSuccess) { rc = o3d::set_error("hipMalloc failed"); return; }
    n->allocs.push_back(*p);
    if (hipMemset(*p, 0, bytes) != hipSuccess) rc = o3d::set_error("hipMemset failed");
  };
  al((void **)&n->W, n->nparam * 4);
  al((void **)&n->G, n->nparam * 4);
  al((void **)&n->M, n->nparam * 4);
  al((void **)&n->Vv, n->nparam * 4);
  size_t wb = 0;
  for (const Conv &c : n->conv) wb += (size_t)c.co * c.kpf + (size_t)c.ci * c.kpb;
  al((void **)&n->wbuf, wb * sizeof(bf16));
  al((void **)&n->mse_part, 1024 * sizeof(double));
  if (!rc && hipHostMalloc((void **)&n->hpart, 1024 * sizeof(double)) != hipSuccess)
    rc = o3d::set_error("hipHostMalloc failed");
  if (!rc && (hipEventCreate(&n->ev[0]) != hipSuccess || hipEventCreate(&n->ev[1]) != hipSuccess))
    rc = o3d::set_error("hipEventCreate failed");
  if (!rc) {
    bf16 *p = n->wbuf;
    for (Conv &c : n->conv) {
      c.wf = p; p += (size_t)c.co * c.kpf;
      c.wb = p; p += (size_t)c.ci * c.kpb;
    }
    std::vector<float> h(n->nparam);
    init_params(h, *n, seed);
    if (hipMemcpy(n->W, h.data(), n->nparam * 4, hipMemcpyHostToDevice) != hipSuccess)
      rc = o3d::set_error("hipMemcpy failed");
  }
  if (!rc) rc = alloc_activations(*n, nx, ny, nz);
  if (!rc) rc = sync_weights(*n);
  if (!rc && hipStreamSynchronize(st()) != hipSuccess) rc = o3d::set_error("sync failed");
  if (rc) { delete n; return rc; }
  *out = n;
  return 0;
}

extern "C" int o3d_nn_destroy(void *nn) {
  if (nn) {
    (void)hipStreamSynchronize(st());
    delete (Net *)nn;
  }
  return 0;
}

extern "C" int o3d_nn_resize(void *nn, int nx, int ny, int nz) {
  if (!nn) return o3d::set_error("NULL network");
  Net &n = *(Net *)nn;
  if (n.nx[0] == nx && n.ny[0] == ny && n.nz[0] == nz) return 0;
  NNCHK(hipStreamSynchronize(st()));
  return alloc_activations(n, nx, ny, nz);
}

extern "C" int o3d_nn_num_params(void *nn, long long *np) {
  if (!nn || !np) return o3d::set_error("NULL argument");
  *np = (long long)((Net *)nn)->nparam;
  return 0;
}

extern "C" int o3d_nn_get_params(void *nn, float *out) {
  if (!nn || !out) return o3d::set_error("NULL argument");
  Net &n = *(Net *)nn;
  NNCHK(hipMemcpyAsync(out, n.W, n.nparam * 4, hipMemcpyDefault, st()));
  NNCHK(hipStreamSynchronize(st()));
  return 0;
}

extern "C" int o3d_nn_set_params(void *nn, const float *in) {
  if (!nn || !in) return o3d::set_error("NULL argument");
  Net &n = *(Net *)nn;
  NNCHK(hipMemcpyAsync(n.W, in, n.nparam * 4, hipMemcpyDefault, st()));
  NCHK(sync_weights(n));
  NNCHK(hipStreamSynchronize(st()));
  return 0;
}

extern "C" int o3d_nn_get_grads(void *nn, float *out) {
  if (!nn || !out) return o3d::set_error("NULL argument");
  Net &n = *(Net *)nn;
  NNCHK(hipMemcpyAsync(out, n.G, n.nparam * 4, hipMemcpyDefault, st()));
  NNCHK(hipStreamSynchronize(st()));
  return 0;
}

extern "C" int o3d_nn_zero_grad(void *nn) {
  if (!nn) return o3d::set_error("NULL network");
  Net &n = *(Net *)nn;
  NNCHK(hipMemsetAsync(n.G, 0, n.nparam * 4, st()));
  return 0;
}

extern "C" int o3d_nn_forward(void *nn, const float *in, float *out) {
  if (!nn || !in || !out) return o3d::set_error("NULL argument");
  return forward(*(Net *)nn, in, out);
}

extern "C" int o3d_nn_backward(void *nn, const float *dout) {
  if (!nn || !dout) return o3d::set_error("NULL argument");
  return backward(*(Net *)nn, dout);
}

extern "C" int o3d_nn_mse(void *nn, const float *out, const float *target, float scale,
                          float *dout, double *loss) {
  if (!nn || !out || !target) return o3d::set_error("NULL argument");
  Net &n = *(Net *)nn;
  const int V = (int)n.V[0];
  const int blocks = std::min(1024, (V + 255) / 256);
  mse_kernel<<<blocks, 256, 0, st()>>>(out, target, scale, dout, n.mse_part, V);
  NNCHK(hipGetLastError());
  if (loss) {
    NNCHK(hipMemcpyAsync(n.hpart, n.mse_part, blocks * sizeof(double), hipMemcpyDeviceToHost,
                         st()));
    NNCHK(hipStreamSynchronize(st()));
    double s = 0.0;
    for (int b = 0; b < blocks; ++b) s += n.hpart[b];
    *loss = s / V;
  }
  return 0;
}

extern "C" int o3d_nn_adam_step(void *nn, float lr, float beta1, float beta2, float eps, int t) {
  if (!nn) return o3d::set_error("NULL network");
  Net &n = *(Net *)nn;
  const float c1 = 1.f - powf(beta1, (float)t), c2 = 1.f - powf(beta2, (float)t);
  adam_kernel<<<elem_blocks(n.nparam), 256, 0, st()>>>(n.W, n.G, n.M, n.Vv, n.nparam, lr, beta1,
                                                        beta2, eps, c1, c2);
  NNCHK(hipGetLastError());
  return sync_weights(n);
}

extern "C" int o3d_nn_last_timing(void *nn, double *ms, double *flops) {
  if (!nn) return o3d::set_error("NULL network");
  Net &n = *(Net *)nn;
  float t = 0.f;
  NNCHK(hipEventSynchronize(n.ev[1]));
  NNCHK(hipEventElapsedTime(&t, n.ev[0], n.ev[1]));
  if (ms) *ms = t;
  if (flops) *flops = n.last_flops;
  return 0;
}

