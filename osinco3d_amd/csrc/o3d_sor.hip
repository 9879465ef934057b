// Red-black SOR for the pressure Poisson equation (src/poisson.f90), MI355X.
//
// One launch performs one full iteration (red sweep, then black sweep) and
// touches HBM once per iteration: it reads the old iterate and the rhs and
// writes the new iterate (24 B/cell in FP64, 12 B/cell in FP32) to a second
// buffer (ping-pong).  Inside a block:
//  * the block owns a 64x8 x-y tile and marches a chunk of z planes;
//  * red values are computed on the tile plus a 1-wide halo from the old
//    iterate (which no block modifies), so the black update of the tile needs
//    no exchange with neighbouring blocks: the halo red values are recomputed
//    here bit-identically to the owning block (same inputs, same order);
//  * the march lags black by one plane: at step s the block computes red on
//    plane s, then black on plane s-1, whose red neighbours (s-2, s-1, s) are
//    all new.  Old planes, new red planes and rhs planes live in small LDS
//    rings (4, 4 and 3 planes).
// Colour of a node is (i+j+k) & 1 (0-based); colour 0 is swept first.  This is
// a valid two-colouring with the reference's neighbour rules when every
// periodic axis has an even node count (checked on the host).
// The per-iteration maximum update dmax (un-relaxed Gauss-Seidel value minus
// old value, src/poisson.f90:100-101) is reduced with wavefront shuffles and
// one 64-bit atomicMax per block on the IEEE bit pattern (dmax >= 0).
#include <algorithm>
#include <utility>

#include "o3d_sor.h"

namespace o3d {

constexpr int SX = 64, SY = 8;
constexpr int R2W = SX + 4, R2H = SY + 4;  // old iterate, halo 2
constexpr int R1W = SX + 2, R1H = SY + 2;  // new red values and rhs, halo 1
constexpr int R2N = R2W * R2H, R1N = R1W * R1H;
constexpr int SNT = SX * SY;
constexpr int SNW = SNT / O3D_WAVE;
static_assert(R2N <= 2 * SNT && R1N <= 2 * SNT, "two cells per thread at most");

// z-chunking: a block marches zc planes and re-reads 3 extra planes (red on
// z0-1 and z1, old on z0-2 and z1+1), so long chunks save traffic, but the
// grid should fill the machine in whole rounds of `slots` concurrent blocks
// (a partial last round leaves CUs idle).  Pick the chunk count with the
// least rounds * (zc + 3).
SorTiling make_sor_tiling(int nx, int ny, int nz, int slots, int zmult) {
  SorTiling t;
  t.ntx = (nx + SX - 1) / SX;
  t.nty = (ny + SY - 1) / SY;
  const int layer = t.ntx * t.nty;
  if (slots <= 0) slots = 1024;
  double best = 1e30;
  int best_ntz = 1;
  for (int ntz = 1; ntz <= std::max(1, nz / 8); ++ntz) {
    const int zc = ((nz + ntz - 1) / ntz + zmult - 1) / zmult * zmult;
    const int n_t = (nz + zc - 1) / zc;
    const int blocks = layer * n_t;
    const int rounds = (blocks + slots - 1) / slots;
    // Time ~ rounds * (zc + 3) plane steps; normalised by the ideal
    // nz * layer / slots.
    const double cost = (double)rounds * (zc + 3.0) * slots / ((double)nz * layer);
    if (cost < best - 1e-9) { best = cost; best_ntz = ntz; }
  }
  t.zc = ((nz + best_ntz - 1) / best_ntz + zmult - 1) / zmult * zmult;
  t.ntz = (nz + t.zc - 1) / t.zc;
  return t;
}

template <typename T, typename TR>
__global__ void sor_kernel(const SorArgs<T, TR> a);

template <typename T, typename TR>
int sor_slots() {
  static int slots = 0;
  if (slots == 0) {
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sor_kernel<T, TR>, SNT, 0) !=
            hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
      (void)hipGetLastError();
      return 1024;
    }
    slots = std::max(1, per_cu * cus);
  }
  return slots;
}

template <typename T>
struct SorShape;

template <typename T, typename TR>
SorTiling sor_tiling(int nx, int ny, int nz) {
  return make_sor_tiling(nx, ny, nz, sor_slots<T, TR>(), SorShape<T>::NO);
}

// Neighbour index for the SOR rules: periodic wrap, or the free-slip mirror
// m1 = p1 = 2 at node 1 and m1 = p1 = n-1 at node n (src/poisson.f90:196-206).
// Deeper halo positions use the same mirror; positions more than 2 beyond the
// domain belong to idle lanes of a partial tile and are clamped.
__device__ __forceinline__ int sor_ghost(int q, int n, int periodic) {
  if (q > n + 1) return n - 1;
  if (q < -2) return 0;
  if (periodic) {
    if (q < 0) return q + n;
    if (q >= n) return q - n;
    return q;
  }
  if (q < 0) return -q;
  if (q > n - 1) return 2 * (n - 1) - q;
  return q;
}

__device__ __forceinline__ double bits_to_double(unsigned long long b) {
  return __longlong_as_double((long long)b);
}

// The reference's loop tail for iteration it-1 (src/poisson.f90:110-122).
// Returns true when the solve is over; sets om to omega for iteration it.
// dm(q) is dmax of iteration q (q >= 1).
template <typename DM>
__device__ __forceinline__ bool sor_decide_with(const SorState *st, int it, double eps,
                                                double factor, int kmax, int idyn, double &om,
                                                int &reason, double &dlast, DM &&dm) {
  reason = -1;
  dlast = 0.0;
  if (it <= 1) { om = st->omega[0]; return false; }
  const double d1 = dm(it - 1);
  const double d0 = (it - 1 == 1) ? 1609.0 : dm(it - 2);
  dlast = d1;
  om = st->omega[(it - 1) & 7];
  if (d1 < eps) { reason = 0; return true; }
  if (fabs(d0 - d1) < eps / 1000.0) { reason = 1; return true; }
  if (it - 1 > 1 && idyn == 1) {
    if (d1 > d0) om = om * (2.0 - factor);
    else if (d1 < 0.1 * d0) om = fmin(om * factor, 2.0);
  }
  if (it > kmax) { reason = 2; return true; }
  return false;
}

__device__ __forceinline__ bool sor_decide(const SorState *st, int it, double eps, double factor,
                                           int kmax, int idyn, double &om, int &reason,
                                           double &dlast) {
  return sor_decide_with(st, it, eps, factor, kmax, idyn, om, reason, dlast,
                         [&](int q) { return bits_to_double(st->dmax_bits[q & 7]); });
}

// Point update of src/poisson.f90:95-98.  The FP64 form divides by A as the
// reference does (bit-identical results); the FP32 correction sweeps of the
// mixed form multiply by the precomputed 1/A instead.
template <typename T, typename TR>
__device__ __forceinline__ T sor_gs(T w_, T e_, T s_, T n_, T b_, T t_, TR r, T ox, T oy, T oz,
                                    T A, T invA) {
  const T num = -ox * (w_ + e_) - oy * (s_ + n_) - oz * (b_ + t_) + T(r);
  if constexpr (sizeof(T) == 8) return num / A;
  else return num * invA;
}

template <int K>
struct IC { static constexpr int value = K; };

template <int... K, typename F>
__device__ __forceinline__ void unroll_steps(std::integer_sequence<int, K...>, F &&f) {
  (f(IC<K>{}), ...);
}

// Planes handled per march step (PPS: one barrier pair per step) and the
// prefetch depth in steps (PD): loads of the old iterate and rhs are issued
// PD steps before their planes are published to LDS, so each block keeps
// PD*PPS planes of loads in flight across the barriers.
// Ring depths: phase 2 of a step reads old planes s-1..s+PPS and rhs planes
// s-1..s+PPS-1, phase 3 reads red planes s-2..s+PPS-1.  Each ring's depth
// divides the planes per loop iteration (PD*PPS), so a plane's slot depends
// only on its position inside the unrolled iteration: slots are loop
// invariants, not per-step arithmetic.
template <typename T>
struct SorShape {
  static constexpr int PPS = sizeof(T) == 4 ? 2 : 1;
  static constexpr int PD = sizeof(T) == 4 ? 2 : 3;
  static constexpr int NO = sizeof(T) == 4 ? 4 : 3;
  static constexpr int NRD = NO, NRH = NO;
  static_assert(NO >= PPS + 2 && NRH >= PPS + 1, "ring too shallow");
  static_assert((PD * PPS) % NO == 0, "ring depth must divide the planes per iteration");
};

// Ring slot of plane q (q >= -2).
template <int N>
__device__ __forceinline__ int slot(int q) {
  if constexpr ((N & (N - 1)) == 0) return q & (N - 1);
  else return (q + 4 * N) % N;
}

// Raw buffer loads: resource in SGPRs, a 32-bit per-lane column offset and a
// uniform plane offset, so the march needs no per-lane 64-bit address
// arithmetic.  Word 3 of the gfx9 buffer descriptor: raw 32-bit data.
constexpr int kBufWord3 = 0x00020000;

template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const T *p, int n) {
  // The launcher admits only fields of at most INT_MAX bytes
  // (sor_buffer_fits); the count is formed in 64 bits all the same.
  const long long bytes = (long long)n * (long long)sizeof(T);
  return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0,
                                           (int)(bytes < 0x7fffffffLL ? bytes : 0x7fffffffLL),
                                           kBufWord3);
}

template <typename T>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// FP32: at most 64 VGPRs (8 waves per SIMD: four 512-thread blocks per CU).
// FP64 needs more registers and runs at 5 waves per SIMD.
template <typename T>
constexpr int sor_waves() { return sizeof(T) == 4 ? 8 : 5; }
template <typename T, typename TR>
__global__ void __launch_bounds__(SNT, sor_waves<T>()) sor_kernel(const SorArgs<T, TR> a) {
  constexpr int PPS = SorShape<T>::PPS;
  constexpr int PD = SorShape<T>::PD;
  constexpr int NO = SorShape<T>::NO, NRD = SorShape<T>::NRD, NRH = SorShape<T>::NRH;
  // LDS rings: old iterate on R2, new red values and rhs on R1.  In FP32 the
  // old and rhs planes are padded to 2*SNT entries so every thread stores
  // its second cell unconditionally (no divergent branch per plane).
  constexpr bool PAD = false;
  constexpr int R2S = PAD ? 2 * SNT : R2N, R1S = PAD ? 2 * SNT : R1N;
  __shared__ T oldr[NO][R2S];
  __shared__ T redr[NRD][R1N];
  __shared__ TR rhsr[NRH][R1S];
  __shared__ double scr[SNW];
  SorState *st = a.st;
  if (st->done) return;
  const int it = a.it + st->base;
  double om, dlast;
  int reason;
  // No eps or stagnation exit inside the Gauss-Seidel opening (iterations
  // 1..a.gs): its small updates say little about the smooth error left.
  if (sor_decide(st, it, a.eps, a.factor, a.kmax, a.idyn, om, reason, dlast) &&
      !(reason != 2 && it - 1 <= a.gs)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done = 1;
      st->reason = reason;
      st->exit_iter = reason == 2 ? a.kmax + 1 : it - 1;
      st->sweeps = it - 1;
      st->dmax_final = dlast;
      st->omega_final = om;
      if (a.log && it - 2 < a.log_len) a.log[it - 2] = dlast;
    }
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->omega[it & 7] = om;
    st->dmax_bits[(it + 1) & 7] = 0ull;
    if (a.log && it >= 2 && it - 2 < a.log_len) a.log[it - 2] = dlast;
  }

  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int plane = nx * ny;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tix = t % a.tl.ntx;
  const int rr = t / a.tl.ntx;
  const int tiy = rr % a.tl.nty;
  const int tiz = rr / a.tl.nty;
  const int x0 = tix * SX, y0 = tiy * SY, z0 = tiz * a.tl.zc;
  const int z1 = min(z0 + a.tl.zc, nz);
  const int tid = threadIdx.x;
  const T ox = T(a.ox), oy = T(a.oy), oz = T(a.oz), A = T(a.A), invA = T(1.0 / a.A);
  // The first a.gs iterations are Gauss-Seidel sweeps (omega = 1); the
  // omega ring keeps the solve's omega.
  const double omu = it <= a.gs ? 1.0 : om;
  const T w = T(omu);
  const T w1 = T(1.0 - omu);

  // Global column offsets of this thread's R2 and R1 cells (the second of
  // each exists for the first R2N-SNT / R1N-SNT threads; the others load a
  // valid dummy address and skip the LDS store).
  const int c2b = tid + SNT;
  const bool v2b = c2b < R2N, v1b = c2b < R1N;
  const int col2a = sor_ghost(x0 - 2 + tid % R2W, nx, a.per[0]) +
                    nx * sor_ghost(y0 - 2 + tid / R2W, ny, a.per[1]);
  const int col2b = v2b ? sor_ghost(x0 - 2 + c2b % R2W, nx, a.per[0]) +
                              nx * sor_ghost(y0 - 2 + c2b / R2W, ny, a.per[1])
                        : col2a;
  const int col1a = sor_ghost(x0 - 1 + tid % R1W, nx, a.per[0]) +
                    nx * sor_ghost(y0 - 1 + tid / R1W, ny, a.per[1]);
  const int col1b = v1b ? sor_ghost(x0 - 1 + c2b % R1W, nx, a.per[0]) +
                              nx * sor_ghost(y0 - 1 + c2b / R1W, ny, a.per[1])
                        : col1a;
  // Red cells of R1: thread tid < 330 owns the (tid % 33)-th red node of R1
  // row tid / 33.
  constexpr int RPR = R1W / 2;
  const bool red_lane = tid < RPR * R1H;
  const int rb1 = tid / RPR, rk = tid % RPR;
  const int rpar = (x0 - 1 + y0 - 1 + rb1) & 1;
  // Black pass: this thread's tile node.
  const int tx = tid % SX, ty = tid / SX;
  const int bi = x0 + tx, bj = y0 + ty;
  const bool bown = bi < nx && bj < ny;
  const int b1i = (ty + 1) * R1W + tx + 1;
  const int b2i = (ty + 2) * R2W + tx + 2;
  const int bpar = (bi + bj) & 1;

  // Stored plane of plane zp >= 0 (the march never goes below z0 - 2; only
  // the upper ghost rule can apply inside the loop), for old and rhs.
  const int perz = a.per[2];
  auto zq = [&](int zp) {
    const int hi = zp > nz + 1 ? nz - 1 : (perz ? zp - nz : 2 * (nz - 1) - zp);
    return (unsigned)(zp < nz ? zp : hi);
  };
  auto zq_lo = [&](int zp) { return (unsigned)sor_ghost(zp, nz, perz); };
  const __amdgpu_buffer_rsrc_t rin = buf_rsrc(a.pin, plane * nz);
  const __amdgpu_buffer_rsrc_t rrh = buf_rsrc(a.rhs, plane * nz);
  const unsigned o2a = (unsigned)col2a * sizeof(T), o2b = (unsigned)col2b * sizeof(T);
  const unsigned o1a = (unsigned)col1a * sizeof(TR), o1b = (unsigned)col1b * sizeof(TR);
  // With a.efold, old values are pin + e, e in colour-split storage: node
  // (x, y, z) at entry (nx/2)(y + ny z) + x/2 of the red (x + y + z even) or
  // black half.  Per R2 column: its entry in a plane and its (x + y) parity;
  // the plane's parity picks the half.
  // (Folding applies to the FP64 judge sweep only; for T = float it is
  // compiled out, keeping that kernel within its register budget.)
  const bool fold = sizeof(T) == 8 && a.efold != nullptr;
  const __amdgpu_buffer_rsrc_t rfe = buf_rsrc(fold ? a.efold : (const float *)a.pin, plane * nz);
  const int hxn = nx >> 1;
  auto ecol = [&](int c) {
    const int gx = sor_ghost(x0 - 2 + c % R2W, nx, a.per[0]);
    const int gy = sor_ghost(y0 - 2 + c / R2W, ny, a.per[1]);
    return 2u * (unsigned)(hxn * gy + (gx >> 1)) + (unsigned)((gx + gy) & 1);
  };
  const unsigned eca = ecol(tid), ecb = v2b ? ecol(c2b) : eca;
  const unsigned halfB = (unsigned)(plane * nz / 2) * 4u;
  // The raw loads; the sum p + e is formed only when a plane is published
  // to LDS, so prefetched loads stay in flight across the march steps.
  auto ld_old = [&](unsigned o, unsigned gz) {
    return bload<T>(rin, o, (unsigned)plane * gz * sizeof(T));
  };
  auto ld_e = [&](unsigned ec, unsigned gz) {
    if constexpr (sizeof(T) != 8) return 0.0f;
    if (!fold) return 0.0f;
    const unsigned eo = (ec >> 1) * 4u + (((ec ^ gz) & 1u) ? halfB : 0u);
    return bload<float>(rfe, eo, (unsigned)hxn * (unsigned)ny * gz * 4u);
  };
  auto pe_add = [&](T v, float e) { return fold ? v + (T)e : v; };
  auto ld_rhs = [&](unsigned o, unsigned gz) {
    return bload<TR>(rrh, o, (unsigned)plane * gz * sizeof(TR));
  };
  
  // Prefetch registers: step k of the march publishes old planes
  // s_k+1 .. s_k+PPS and rhs planes s_k .. s_k+PPS-1, s_k = z0-1 + k*PPS.
  T po[PD][PPS][2];
  float pe[PD][PPS][2];
  TR pr[PD][PPS][2];
  {
    const unsigned g2 = zq_lo(z0 - 2), g1 = zq_lo(z0 - 1);
    oldr[slot<NO>(-2)][tid] = pe_add(ld_old(o2a, g2), ld_e(eca, g2));
    oldr[slot<NO>(-1)][tid] = pe_add(ld_old(o2a, g1), ld_e(eca, g1));
    if (v2b) {
      oldr[slot<NO>(-2)][c2b] = pe_add(ld_old(o2b, g2), ld_e(ecb, g2));
      oldr[slot<NO>(-1)][c2b] = pe_add(ld_old(o2b, g1), ld_e(ecb, g1));
    }
#pragma unroll
    for (int k = 0; k < PD; ++k)
#pragma unroll
      for (int q = 0; q < PPS; ++q) {
        const int sk = z0 - 1 + k * PPS + q;
        const unsigned go = zq_lo(sk + 1), gr = zq_lo(sk);
        po[k][q][0] = ld_old(o2a, go);
        po[k][q][1] = ld_old(o2b, go);
        pe[k][q][0] = ld_e(eca, go);
        pe[k][q][1] = ld_e(ecb, go);
        pr[k][q][0] = ld_rhs(o1a, gr);
        pr[k][q][1] = ld_rhs(o1b, gr);
      }
  }

  // Largest un-relaxed update, in the sweep's own precision (exact: the
  // maximum of |pn - pc| values that are themselves values of type T).
  T dmax = T(0);
  auto upd = [&](T d) { dmax = d > dmax ? d : dmax; };
  // The owned nodes' new values, summed (FP64 judge sweep only).
  const bool want_sum = sizeof(T) == 8 && a.psum != nullptr;
  double vsum = 0.0;
  auto step = [&](int s, auto K) {
    constexpr int k = decltype(K)::value;
    // Slots of planes s + j: s = z0 - 1 + k*PPS + (a multiple of every
    // depth) and z0 is a multiple of every depth (sor_tiling), so the slots
    // are compile-time constants.
    auto so = [&](int j) { return slot<NO>(-1 + k * PPS + j); };
    auto sd = [&](int j) { return slot<NRD>(-1 + k * PPS + j); };
    auto sh = [&](int j) { return slot<NRH>(-1 + k * PPS + j); };
    // Phase 1: publish old(s+1..s+PPS) and rhs(s..s+PPS-1); prefetch the
    // planes PD steps ahead into the same registers.
#pragma unroll
    for (int q = 0; q < PPS; ++q) {
      T *od = oldr[so(1 + q)];
      TR *rd = rhsr[sh(q)];
      od[tid] = pe_add(po[k][q][0], pe[k][q][0]);
      rd[tid] = pr[k][q][0];
      if (PAD || v2b) od[c2b] = pe_add(po[k][q][1], pe[k][q][1]);
      if (PAD || v1b) rd[c2b] = pr[k][q][1];
      const unsigned go = zq(s + 1 + q + PD * PPS), gr = zq(s + q + PD * PPS);
      po[k][q][0] = ld_old(o2a, go);
      po[k][q][1] = ld_old(o2b, go);
      pe[k][q][0] = ld_e(eca, go);
      pe[k][q][1] = ld_e(ecb, go);
      pr[k][q][0] = ld_rhs(o1a, gr);
      pr[k][q][1] = ld_rhs(o1b, gr);
    }
    __syncthreads();
    // Phase 2: red (colour 0) on planes s..s+PPS-1 over R1 from old values
    // only; keep this thread's old value and rhs of planes s-1..s+PPS-2 for
    // the black pass.
    T bold[PPS];
    TR brhs[PPS];
#pragma unroll
    for (int q = 0; q < PPS; ++q) {
      const int sr = s + q;
      if (red_lane) {
        const int a1 = 2 * rk + ((rpar + sr) & 1);
        const int X = x0 - 1 + a1, Y = y0 - 1 + rb1;
        const int i2 = (rb1 + 1) * R2W + a1 + 1;
        const int i1 = rb1 * R1W + a1;
        const T *o0 = oldr[so(q)];
        const T *om_ = oldr[so(q - 1)];
        const T *op = oldr[so(q + 1)];
        const T pc = o0[i2];
        const T pn = sor_gs<T, TR>(o0[i2 - 1], o0[i2 + 1], o0[i2 - R2W], o0[i2 + R2W], om_[i2],
                                   op[i2], rhsr[sh(q)][i1], ox, oy, oz, A, invA);
        redr[sd(q)][i1] = w1 * pc + w * pn;
        const bool own = a1 >= 1 && a1 <= SX && rb1 >= 1 && rb1 <= SY && X < nx && Y < ny;
        if (sr >= z0 && sr < z1) upd(own ? T(fabs(pn - pc)) : T(0));
      }
      bold[q] = oldr[so(q - 1)][b2i];
      brhs[q] = rhsr[sh(q - 1)][b1i];
    }
    __syncthreads();
    // Phase 3: black (colour 1) on planes s-1..s+PPS-2 over the tile.
#pragma unroll
    for (int q = 0; q < PPS; ++q) {
      const int sb = s - 1 + q;
      if (sb >= z0 && sb < z1) {
        // Every lane evaluates the black update (branch-free); red lanes
        // keep their new red value.
        const T *rm = redr[sd(q - 2)];
        const T *r0 = redr[sd(q - 1)];
        const T *rp = redr[sd(q)];
        const T c = r0[b1i];
        const T pn = sor_gs<T, TR>(r0[b1i - 1], r0[b1i + 1], r0[b1i - R1W], r0[b1i + R1W],
                                   rm[b1i], rp[b1i], brhs[q], ox, oy, oz, A, invA);
        const bool blk = ((bpar + sb) & 1) != 0;
        const T v = blk ? w1 * bold[q] + w * pn : c;
        upd(blk && bown ? T(fabs(pn - bold[q])) : T(0));
        if (bown) {
          (a.pout + (size_t)plane * sb)[(unsigned)(bi + nx * bj)] = v;
          if (want_sum) vsum += (double)v;
        }
      }
    }
  };
  // Whole groups of PD steps in the loop (every back edge comes from the
  // last step of a group, so the loads of a group stay in flight across the
  // next one), then the remaining steps.
  int s = z0 - 1;
  for (; s + (PD - 1) * PPS <= z1; s += PD * PPS)
    unroll_steps(std::make_integer_sequence<int, PD>{},
                 [&](auto K) { step(s + decltype(K)::value * PPS, K); });
  unroll_steps(std::make_integer_sequence<int, PD>{}, [&](auto K) {
    const int sk = s + decltype(K)::value * PPS;
    if (sk <= z1) step(sk, K);
  });
  // Block max, one atomic per block.
  const double dm = block_max<SNW>((double)dmax, scr);
  if (threadIdx.x == 0 && dm != 0.0)
    atomicMax(&st->dmax_bits[it & 7], (unsigned long long)__double_as_longlong(dm));
  if (want_sum) {
    const double bs = block_sum<SNW>(vsum, scr);
    if (threadIdx.x == 0) a.psum[t] = bs;
  }
}

// ------------------------------------------------ colour-split half sweeps
// The FP32 correction sweeps of the mixed form keep the iterate and the rhs
// in colour-split storage: every row (j, k) stores its red nodes
// (i + j + k even) in the red half-array and its black nodes in the black
// one, node i at entry i >> 1 of its colour's row (nx even).  An iteration
// is two in-place launches, red then black; each updates one colour from the
// other with 16-byte loads and stores, no LDS and no halo recomputation
// (16 B/cell per iteration instead of 12, at close to streaming rates).
// Red node (c, j, k) of row parity q = (j + k) & 1 sits at i = 2c + q; its
// x neighbours are black entries c - 1 + q and c + q; black node c sits at
// i = 2c + 1 - q with red x neighbours c - q and c + 1 - q; the y and z
// neighbours of either colour are entry c of the other colour's row.
// Each launch applies the same point update (sor_gs, then the relaxation) in
// the same operation order as sor_kernel, so the iterates are identical.
// dmax is reduced per block and max-accumulated into one of kSplitSlots
// words (on separate cache lines) so that thousands of blocks do not contend
// for one address; the next red launch reduces the slots with one wave.
// Entries per 16-byte vector: 4 in FP32, 2 in FP64.
template <typename T>
constexpr int split_v() { return 16 / (int)sizeof(T); }
constexpr int kSplitZC = 2, kSplitBX = 32, kSplitBY = 8;
constexpr int kSplitNT = kSplitBX * kSplitBY;
constexpr int kSplitSlots = 64, kSplitStride = 16;

template <typename T>
bool sor_split_supported(int nx, int ny, int nz) {
  (void)ny; (void)nz;
  return nx % (2 * split_v<T>()) == 0;
}
size_t sor_split_slot_words() { return (size_t)8 * kSplitSlots * kSplitStride; }

__device__ __forceinline__ int sor_ghost1(int q, int n, int periodic) {
  if (q < 0) return periodic ? q + n : -q;
  if (q >= n) return periodic ? q - n : 2 * (n - 1) - q;
  return q;
}

template <typename T, int COLOUR>
__global__ void __launch_bounds__(kSplitNT) sor_split_kernel(const SorSplitArgs<T> a) {
  constexpr int kSplitV = split_v<T>();
  typedef T vec __attribute__((ext_vector_type(kSplitV)));
  __shared__ double scr[kSplitNT / O3D_WAVE];
  SorState *st = a.st;
  if (st->done) return;
  const int it = a.it + st->base;
  const int tl = threadIdx.x + kSplitBX * threadIdx.y;
  const bool first = blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0;
  T *__restrict__ X = COLOUR == 0 ? a.red : a.black;
  const T *__restrict__ Y = COLOUR == 0 ? a.black : a.red;
  const T *__restrict__ RX = COLOUR == 0 ? a.rhs_red : a.rhs_black;
  const int hx = a.nx / 2, ny = a.ny, nz = a.nz;
  const int px = a.per[0], py = a.per[1], pz = a.per[2];
  const int nv = hx / kSplitV;
  const int cv = blockIdx.x * kSplitBX + threadIdx.x;
  const int j = blockIdx.y * kSplitBY + threadIdx.y;
  const int k0 = blockIdx.z * kSplitZC;
  const bool active = cv < nv && j < ny;
  // The dmax slots of the two previous iterations are loaded first and
  // every load of the sweep right after them, so the slot reduction and the
  // stop decision overlap the sweep's loads (the load counter completes in
  // order); the decision only gates the stores.
  double sl1 = 0.0, sl2 = 0.0;
  if (COLOUR == 0) {
    const size_t lane = __lane_id();
    sl1 = bits_to_double(a.dslot[((size_t)((it - 1) & 7) * kSplitSlots + lane) * kSplitStride]);
    sl2 = bits_to_double(a.dslot[((size_t)((it - 2) & 7) * kSplitSlots + lane) * kSplitStride]);
  }
  const int c0 = active ? cv * kSplitV : 0;
  const int jj = active ? j : 0;
  const size_t pl = (size_t)hx * ny;
  const size_t rj = (size_t)hx * jj + c0;
  const size_t rjm = (size_t)hx * sor_ghost1(jj - 1, ny, py) + c0;
  const size_t rjp = (size_t)hx * sor_ghost1(jj + 1, ny, py) + c0;
  // Entry beside the vector: left for s = 0, right for s = 1 (periodic
  // wrap, or the free-slip mirror, which keeps the colour).
  const int cl = c0 - 1 < 0 ? (px ? hx - 1 : 0) : c0 - 1;
  const int cr = c0 + kSplitV >= hx ? (px ? 0 : hx - 1) : c0 + kSplitV;
  vec yz[kSplitZC + 2], ym[kSplitZC], yp[kSplitZC], pc[kSplitZC], rr[kSplitZC];
  T side[kSplitZC];
  yz[0] = *(const vec *)(Y + rj + pl * sor_ghost1(k0 - 1, nz, pz));
#pragma unroll
  for (int kk = 0; kk <= kSplitZC; ++kk)
    yz[kk + 1] = *(const vec *)(Y + rj + pl * sor_ghost1(min(k0 + kk, nz), nz, pz));
#pragma unroll
  for (int kk = 0; kk < kSplitZC; ++kk) {
    const int k = min(k0 + kk, nz - 1);  // planes past the end are not stored
    const int sd = COLOUR == 0 ? ((jj + k) & 1) : 1 - ((jj + k) & 1);
    const size_t pk = pl * k;
    side[kk] = Y[pk + (rj - c0) + (sd == 0 ? cl : cr)];
    ym[kk] = *(const vec *)(Y + rjm + pk);
    yp[kk] = *(const vec *)(Y + rjp + pk);
    pc[kk] = *(const vec *)(X + rj + pk);
    rr[kk] = *(const vec *)(RX + rj + pk);
  }
  double om;
  bool done = false;
  if (COLOUR == 0) {
    double dlast;
    int reason;
    done = sor_decide_with(st, it, a.eps, a.factor, a.kmax, a.idyn, om, reason, dlast,
                           [&](int q) { return wave_max(q == it - 1 ? sl1 : sl2); });
    // No eps or stagnation exit inside the Gauss-Seidel opening: its small
    // updates say little about the smooth error that remains.
    if (done && reason != 2 && it - 1 <= a.gs) done = false;
    if (first && done && tl == 0) {
      st->done = 1;
      st->reason = reason;
      st->exit_iter = reason == 2 ? a.kmax + 1 : it - 1;
      st->sweeps = it - 1;
      st->dmax_final = dlast;
      st->omega_final = om;
      if (a.log && it - 2 < a.log_len) a.log[it - 2] = dlast;
    }
    if (first && !done) {
      if (tl < kSplitSlots)
        a.dslot[((size_t)((it + 1) & 7) * kSplitSlots + tl) * kSplitStride] = 0ull;
      if (tl == 0) {
        st->omega[it & 7] = om;
        if (a.log && it >= 2 && it - 2 < a.log_len) a.log[it - 2] = dlast;
      }
    }
  } else {
    om = st->omega[it & 7];
  }
  if (done) return;
  const T ox = T(a.ox), oy = T(a.oy), oz = T(a.oz), A = T(a.A);
  const T invA = T(1.0 / a.A);
  // The first a.gs iterations are Gauss-Seidel sweeps (omega = 1); the
  // omega ring keeps the solve's omega.
  const double omu = it <= a.gs ? 1.0 : om;
  const T w = T(omu), w1 = T(1.0 - omu);
  T dmax = T(0);
  if (active) {
#pragma unroll
    for (int kk = 0; kk < kSplitZC; ++kk) {
      const int k = k0 + kk;
      if (k < nz) {
        const int sd = COLOUR == 0 ? ((j + k) & 1) : 1 - ((j + k) & 1);
        const vec zm = yz[kk], zc = yz[kk + 1], zp = yz[kk + 2];
        vec out;
#pragma unroll
        for (int v = 0; v < kSplitV; ++v) {
          const T xm = sd == 0 ? (v == 0 ? side[kk] : zc[v - 1]) : zc[v];
          const T xp = sd == 0 ? zc[v] : (v == kSplitV - 1 ? side[kk] : zc[v + 1]);
          const T pn = sor_gs<T, T>(xm, xp, ym[kk][v], yp[kk][v], zm[v], zp[v], rr[kk][v], ox,
                                    oy, oz, A, invA);
          out[v] = w1 * pc[kk][v] + w * pn;
          const T d = T(fabs(pn - pc[kk][v]));
          dmax = d > dmax ? d : dmax;
        }
        *(vec *)(X + rj + pl * k) = out;
      }
    }
  }
  double dm = wave_max((double)dmax);
  if (tl % O3D_WAVE == 0) scr[tl / O3D_WAVE] = dm;
  __syncthreads();
  dm = scr[0];
#pragma unroll
  for (int q = 1; q < kSplitNT / O3D_WAVE; ++q) dm = fmax(dm, scr[q]);
  if (tl == 0 && dm != 0.0) {
    const int b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    atomicMax(&a.dslot[((size_t)(it & 7) * kSplitSlots + (b % kSplitSlots)) * kSplitStride],
              (unsigned long long)__double_as_longlong(dm));
  }
}

__global__ void sor_split_init_kernel(unsigned long long *dslot, int words) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < words) dslot[i] = 0ull;
}

hipError_t launch_sor_split_init(unsigned long long *dslot, hipStream_t s) {
  const int words = (int)sor_split_slot_words();
  sor_split_init_kernel<<<(words + 255) / 256, 256, 0, s>>>(dslot, words);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_sor_split_iteration(const SorSplitArgs<T> &a, hipStream_t s) {
  if (!sor_split_supported<T>(a.nx, a.ny, a.nz))
    return hipErrorInvalidValue;
  const int nv = a.nx / 2 / split_v<T>();
  const dim3 grd((nv + kSplitBX - 1) / kSplitBX, (a.ny + kSplitBY - 1) / kSplitBY,
                 (a.nz + kSplitZC - 1) / kSplitZC);
  const dim3 blk(kSplitBX, kSplitBY);
  sor_split_kernel<T, 0><<<grd, blk, 0, s>>>(a);
  sor_split_kernel<T, 1><<<grd, blk, 0, s>>>(a);
  return hipGetLastError();
}

// Natural <-> colour-split storage (nx even).
__device__ __forceinline__ size_t split_entry(int idx, int nx, int ny, bool &black) {
  const int i = idx % nx, r = idx / nx;  // r = j + ny * k
  const int j = r % ny, k = r / ny;
  black = ((i + j + k) & 1) != 0;
  return (size_t)(nx / 2) * r + (i >> 1);
}

template <typename T>
__global__ void __launch_bounds__(256) split_kernel(const T *__restrict__ f, T *__restrict__ red,
                                                    T *__restrict__ black, int nx, int ny,
                                                    int n) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  bool b;
  const size_t o = split_entry(idx, nx, ny, b);
  (b ? black : red)[o] = f[idx];
}

template <typename T>
__global__ void __launch_bounds__(256) unsplit_kernel(T *__restrict__ f,
                                                      const T *__restrict__ red,
                                                      const T *__restrict__ black, int nx,
                                                      int ny, int n) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  bool b;
  const size_t o = split_entry(idx, nx, ny, b);
  f[idx] = (b ? black : red)[o];
}

template <typename T>
hipError_t launch_split(const T *f, T *red, T *black, int nx, int ny, int nz, hipStream_t s) {
  const int n = nx * ny * nz;
  split_kernel<T><<<(n + 255) / 256, 256, 0, s>>>(f, red, black, nx, ny, n);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_unsplit(T *f, const T *red, const T *black, int nx, int ny, int nz,
                          hipStream_t s) {
  const int n = nx * ny * nz;
  unsplit_kernel<T><<<(n + 255) / 256, 256, 0, s>>>(f, red, black, nx, ny, n);
  return hipGetLastError();
}

template bool sor_split_supported<float>(int, int, int);
template bool sor_split_supported<double>(int, int, int);
template hipError_t launch_sor_split_iteration<float>(const SorSplitArgs<float> &, hipStream_t);
template hipError_t launch_sor_split_iteration<double>(const SorSplitArgs<double> &,
                                                       hipStream_t);
template hipError_t launch_split<float>(const float *, float *, float *, int, int, int,
                                        hipStream_t);
template hipError_t launch_split<double>(const double *, double *, double *, int, int, int,
                                         hipStream_t);
template hipError_t launch_unsplit<float>(float *, const float *, const float *, int, int, int,
                                          hipStream_t);
template hipError_t launch_unsplit<double>(double *, const double *, const double *, int, int,
                                           int, hipStream_t);

template <typename T, typename TR>
hipError_t launch_sor_iteration(const SorArgs<T, TR> &a, hipStream_t s) {
  // The kernel's ring slots assume z chunks that are a multiple of the ring
  // depth, and its buffer descriptors address at most INT_MAX bytes.
  if (a.tl.zc % SorShape<T>::NO != 0 || !sor_buffer_fits<T, TR>(a.nx, a.ny, a.nz))
    return hipErrorInvalidValue;
  sor_kernel<T, TR><<<dim3(a.tl.blocks()), dim3(SNT), 0, s>>>(a);
  return hipGetLastError();
}

__global__ void sor_init_kernel(SorState *st, double omega, int base) {
  if (threadIdx.x == 0) {
    for (int q = 0; q < 8; ++q) { st->dmax_bits[q] = 0ull; st->omega[q] = omega; }
    st->omega_final = omega;
    st->dmax_final = 0.0;
    st->done = 0;
    st->exit_iter = 0;
    st->reason = -1;
    st->sweeps = 0;
    st->base = base;
  }
}

hipError_t launch_sor_init(SorState *st, double omega, int base, hipStream_t s) {
  sor_init_kernel<<<1, 64, 0, s>>>(st, omega, base);
  return hipGetLastError();
}

// ------------------------------------------------------ mixed precision
// The residual kernels march z over RZC planes of a 64 x 4 (i, j) tile per
// block, so a launch has many blocks and no index divisions: the plain
// residuals with the plane's tile and halo in LDS (tile_march), the guess
// kernels (two and three fields per node) with one column per thread and
// every plane's loads issued before the first use, which keeps more loads in
// flight than a plane-by-plane LDS march at their register count (measured:
// the adaptive guess took 210 us per 256^3 launch as an LDS march, 163 us
// as columns).  The residual keeps the operation order of the red-black
// sweep's Laplacian.
constexpr int RBX = 64, RBY = 4, RZC = 8, RNT = RBX * RBY;

// The tile of this block: the linear block index remapped XCD-aware (tiles
// x fastest, then y, then z; see xcd_tile), so that tiles sharing halo rows
// run on one XCD and find them in its L2.  b is the tile's linear index, the
// slot of its per-block partial sums.
struct RTile {
  int bx, by, bz, b;
};
__device__ __forceinline__ RTile residual_tile() {
  const int gx = (int)gridDim.x, gy = (int)gridDim.y;
  const int lin = (int)blockIdx.x + gx * ((int)blockIdx.y + gy * (int)blockIdx.z);
  const int t = xcd_tile(lin, gx * gy * (int)gridDim.z);
  RTile r;
  r.bx = t % gx;
  r.by = (t / gx) % gy;
  r.bz = t / (gx * gy);
  r.b = t;
  return r;
}

dim3 residual_grid(int nx, int ny, int nz) {
  return dim3((nx + RBX - 1) / RBX, (ny + RBY - 1) / RBY, (nz + RZC - 1) / RZC);
}
int residual_blocks(int nx, int ny, int nz) {
  const dim3 g = residual_grid(nx, ny, nz);
  return (int)(g.x * g.y * g.z);
}

// LDS-tiled z march of the residual kernels over the block's RBX x RBY tile
// (XCD-aware order, RZC planes).  Per plane, the NF node values of the tile
// and of its one-node x and y halo (no corners) are staged in LDS, double
// buffered (one barrier per plane); each thread keeps its own column's
// planes k-1, k, k+1 (and k+2 in flight) and the rhs in registers.  A node
// value is then loaded once per plane, plus the halo, instead of once per
// neighbour reading it.  ld(o, v) fills v[0..NF) for node offset o;
// f(o, i, j, k, nb, rh) receives, per field, the neighbour values
// nb[f][0..6] = (x-1, x+1, y-1, y+1, z-1, z+1, centre) and the rhs, in the
// operation order every residual kernel uses.  Every thread of the block
// runs the march (it holds barriers); threads outside the grid load nothing
// and call f for nothing.
template <int NF, typename TR, typename LD, typename F>
__device__ __forceinline__ void tile_march(LD &&ld, const TR *__restrict__ rhs, int nx, int ny,
                                           int nz, int px, int py, int pz, F &&f) {
  __shared__ double tl[2][NF][RBY + 2][RBX + 2];
  const RTile T = residual_tile();
  const int tx = threadIdx.x % RBX, ty = threadIdx.x / RBX;
  const int i = T.bx * RBX + tx, j = T.by * RBY + ty;
  const bool act = i < nx && j < ny;
  const int k0 = T.bz * RZC;
  const int kend = min(k0 + RZC, nz);
  const size_t plane = (size_t)nx * ny;
  // Halo duties: the first and last valid column and row of the tile.
  const int ie = min(T.bx * RBX + RBX, nx) - 1, je = min(T.by * RBY + RBY, ny) - 1;
  const bool hxl = act && tx == 0, hxr = act && i == ie;
  const bool hyl = act && ty == 0, hyr = act && j == je;
  const int c = act ? i + nx * j : 0;
  const int cxl = act ? sor_ghost(i - 1, nx, px) + nx * j : 0;
  const int cxr = act ? sor_ghost(i + 1, nx, px) + nx * j : 0;
  const int cyl = act ? i + nx * sor_ghost(j - 1, ny, py) : 0;
  const int cyr = act ? i + nx * sor_ghost(j + 1, ny, py) : 0;
  auto zo = [&](int k) { return plane * (size_t)sor_ghost(k, nz, pz); };
  double cm[NF], c0[NF], cp[NF], cn[NF], h[4][NF];
  double rh = 0.0, rn = 0.0;
  __syncthreads();  // a previous march of this block may still read the tile
  if (act) {
    ld(c + zo(k0 - 1), cm);
    ld(c + zo(k0), c0);
    ld(c + zo(k0 + 1), cp);
    rh = (double)rhs[c + plane * k0];
    if (hxl) ld(cxl + zo(k0), h[0]);
    if (hxr) ld(cxr + zo(k0), h[1]);
    if (hyl) ld(cyl + zo(k0), h[2]);
    if (hyr) ld(cyr + zo(k0), h[3]);
#pragma unroll
    for (int q = 0; q < NF; ++q) {
      tl[0][q][ty + 1][tx + 1] = c0[q];
      if (hxl) tl[0][q][ty + 1][tx] = h[0][q];
      if (hxr) tl[0][q][ty + 1][tx + 2] = h[1][q];
      if (hyl) tl[0][q][ty][tx + 1] = h[2][q];
      if (hyr) tl[0][q][ty + 2][tx + 1] = h[3][q];
    }
  }
  __syncthreads();
  for (int k = k0; k < kend; ++k) {
    const int b = (k - k0) & 1;
    const bool more = k + 1 < kend;
    if (act) {
      // Loads for later planes first: the column's plane k+2, the rhs and
      // the halo of plane k+1.
      ld(c + zo(k + 2), cn);
      if (more) {
        rn = (double)rhs[c + plane * (k + 1)];
        if (hxl) ld(cxl + zo(k + 1), h[0]);
        if (hxr) ld(cxr + zo(k + 1), h[1]);
        if (hyl) ld(cyl + zo(k + 1), h[2]);
        if (hyr) ld(cyr + zo(k + 1), h[3]);
      }
      double nb[NF][7];
#pragma unroll
      for (int q = 0; q < NF; ++q) {
        nb[q][0] = tl[b][q][ty + 1][tx];
        nb[q][1] = tl[b][q][ty + 1][tx + 2];
        nb[q][2] = tl[b][q][ty][tx + 1];
        nb[q][3] = tl[b][q][ty + 2][tx + 1];
        nb[q][4] = cm[q];
        nb[q][5] = cp[q];
        nb[q][6] = c0[q];
      }
      f(c + plane * k, i, j, k, nb, rh);
      if (more) {
#pragma unroll
        for (int q = 0; q < NF; ++q) {
          tl[1 - b][q][ty + 1][tx + 1] = cp[q];
          if (hxl) tl[1 - b][q][ty + 1][tx] = h[0][q];
          if (hxr) tl[1 - b][q][ty + 1][tx + 2] = h[1][q];
          if (hyl) tl[1 - b][q][ty][tx + 1] = h[2][q];
          if (hyr) tl[1 - b][q][ty + 2][tx + 1] = h[3][q];
        }
      }
#pragma unroll
      for (int q = 0; q < NF; ++q) {
        cm[q] = c0[q];
        c0[q] = cp[q];
        cp[q] = cn[q];
      }
      rh = rn;
    }
    if (more) __syncthreads();
  }
}

// f(offset of (i, j, k), i, j, k, p(i, j, k), residual) for every node of
// the block's tile, the iterate's value at offset o being ld(o).
template <typename TR, typename LD, typename F>
__device__ __forceinline__ void residual_column_ld(LD &&ld, const TR *__restrict__ rhs, int nx,
                                                   int ny, int nz, int px, int py, int pz,
                                                   double ox, double oy, double oz, double A,
                                                   F &&f) {
  tile_march<1>([&](size_t o, double *v) { v[0] = ld(o); }, rhs, nx, ny, nz, px, py, pz,
                [&](size_t o, int i, int j, int k, const double (*nb)[7], double rh) {
                  const double lap = ox * (nb[0][0] + nb[0][1]) + oy * (nb[0][2] + nb[0][3]) +
                                     oz * (nb[0][4] + nb[0][5]) + A * nb[0][6];
                  f(o, i, j, k, nb[0][6], rh - lap);
                });
}

template <typename TR, typename F>
__device__ __forceinline__ void residual_column(const double *__restrict__ p,
                                                const TR *__restrict__ rhs, int nx, int ny,
                                                int nz, int px, int py, int pz, double ox,
                                                double oy, double oz, double A, F &&f) {
  residual_column_ld([&](size_t o) { return p[o]; }, rhs, nx, ny, nz, px, py, pz, ox, oy, oz,
                     A, [&](size_t o, int i, int j, int k, double, double r) {
                       f(o, i, j, k, r);
                     });
}

// q = p + e (the extrapolated guess) and r = rhs - L q stored in FP32, in one
// pass: the residual is formed from p + e as loaded, with the operation
// order of residual_f32_kernel, so both equal the two-kernel result.  With
// part0, the block's (max |r0|, sum r0^2) of r0 = rhs - L p as well (from the
// same loads), with the arithmetic of residual_norm_pair_kernel's p0 part.
template <typename TR>
__global__ void __launch_bounds__(RNT) guess_residual_kernel(const double *__restrict__ p,
                                                             const float *__restrict__ e,
                                                             const TR *__restrict__ rhs,
                                                             double *__restrict__ q,
                                                             float *__restrict__ r,
                                                             double *__restrict__ part0, int nx,
                                                             int ny, int nz, int px, int py,
                                                             int pz, double ox, double oy,
                                                             double oz, double A) {
  __shared__ double sc[RNT / O3D_WAVE];
  double m0 = 0.0, s0 = 0.0;
  const RTile T = residual_tile();
  const int i = T.bx * RBX + (threadIdx.x % RBX);
  const int j = T.by * RBY + (threadIdx.x / RBX);
  if (i < nx && j < ny) {
    const int k0 = T.bz * RZC;
    const size_t plane = (size_t)nx * ny;
    const int im = sor_ghost(i - 1, nx, px), ip = sor_ghost(i + 1, nx, px);
    const int jm = sor_ghost(j - 1, ny, py), jp = sor_ghost(j + 1, ny, py);
    const int c = i + nx * j;
    const int cl[4] = {im + nx * j, ip + nx * j, i + nx * jm, i + nx * jp};
    double zp[RZC + 2], lp[4][RZC], rh[RZC];
    float ze[RZC + 2], le[4][RZC];
#pragma unroll
    for (int kk = 0; kk < RZC + 2; ++kk) {
      const size_t o = c + plane * sor_ghost(min(k0 + kk - 1, nz), nz, pz);
      zp[kk] = p[o];
      ze[kk] = e[o];
    }
#pragma unroll
    for (int kk = 0; kk < RZC; ++kk) {
      const size_t o = plane * min(k0 + kk, nz - 1);
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        lp[d][kk] = p[cl[d] + o];
        le[d][kk] = e[cl[d] + o];
      }
      rh[kk] = (double)rhs[c + o];
    }
    auto qv = [](double a, float b) { return a + 1.0 * (double)b; };
#pragma unroll
    for (int kk = 0; kk < RZC; ++kk) {
      const int k = k0 + kk;
      if (k < nz) {
        const size_t o = c + plane * k;
        const double lapq = ox * (qv(lp[0][kk], le[0][kk]) + qv(lp[1][kk], le[1][kk])) +
                            oy * (qv(lp[2][kk], le[2][kk]) + qv(lp[3][kk], le[3][kk])) +
                            oz * (qv(zp[kk], ze[kk]) + qv(zp[kk + 2], ze[kk + 2])) +
                            A * qv(zp[kk + 1], ze[kk + 1]);
        q[o] = qv(zp[kk + 1], ze[kk + 1]);
        r[o] = (float)(rh[kk] - lapq);
        if (part0) {
          const double lapp = ox * (lp[0][kk] + lp[1][kk]) + oy * (lp[2][kk] + lp[3][kk]) +
                              oz * (zp[kk] + zp[kk + 2]) + A * zp[kk + 1];
          const double r0 = rh[kk] - lapp;
          m0 = fmax(m0, fabs(r0));
          s0 += r0 * r0;
        }
      }
    }
  }
  if (part0) {
    m0 = block_max<RNT / O3D_WAVE>(m0, sc);
    s0 = block_sum<RNT / O3D_WAVE>(s0, sc);
    if (threadIdx.x == 0) {
      const size_t b = (size_t)T.b;
      part0[2 * b] = m0;
      part0[2 * b + 1] = s0;
    }
  }
}

template <typename TR>
hipError_t launch_guess_residual(const double *p, const float *e, const TR *rhs, double *q,
                                 float *r, int nx, int ny, int nz, const int per[3], double ox,
                                 double oy, double oz, double A, hipStream_t s, double *part0) {
  guess_residual_kernel<TR><<<residual_grid(nx, ny, nz), RNT, 0, s>>>(
      p, e, rhs, q, r, part0, nx, ny, nz, per[0], per[1], per[2], ox, oy, oz, A);
  return hipGetLastError();
}

// The adaptive extrapolation q = p + e1 + beta (e1 - e2) (e1, e2 the last
// two pressure increments: beta = 0 is the linear, beta = 1 the quadratic
// extrapolation) and r = rhs - L q stored in FP32, in one pass; with part0
// the block's (max |r0|, sum r0^2) of r0 = rhs - L p as in
// guess_residual_kernel; and per block the two sums that give the
// least-squares beta of this rhs, (sum (r + beta Ld) Ld, sum Ld Ld) with
// d = e1 - e2: the weight that minimises the L2 residual of p + e1 + beta d,
// which the next step uses.
template <typename TR>
__global__ void __launch_bounds__(RNT) guess_adapt_kernel(
    const double *__restrict__ p, const float *__restrict__ e1, const float *__restrict__ e2,
    const TR *__restrict__ rhs, double beta, double *__restrict__ q, float *__restrict__ r,
    double *__restrict__ part0, double *__restrict__ partb, int nx, int ny, int nz, int px,
    int py, int pz, double ox, double oy, double oz, double A) {
  __shared__ double sc[RNT / O3D_WAVE];
  double m0 = 0.0, s0 = 0.0, sb = 0.0, sd = 0.0;
  const RTile T = residual_tile();
  const int i = T.bx * RBX + (threadIdx.x % RBX);
  const int j = T.by * RBY + (threadIdx.x / RBX);
  if (i < nx && j < ny) {
    const int k0 = T.bz * RZC;
    const size_t plane = (size_t)nx * ny;
    const int im = sor_ghost(i - 1, nx, px), ip = sor_ghost(i + 1, nx, px);
    const int jm = sor_ghost(j - 1, ny, py), jp = sor_ghost(j + 1, ny, py);
    const int c = i + nx * j;
    const int cl[4] = {im + nx * j, ip + nx * j, i + nx * jm, i + nx * jp};
    // The column (planes k0-1 .. k0+RZC) of p, e1 and d = e1 - e2, then the
    // lateral neighbours plane by plane.
    double zp[RZC + 2];
    float ze[RZC + 2], zd[RZC + 2];
#pragma unroll
    for (int kk = 0; kk < RZC + 2; ++kk) {
      const size_t o = c + plane * sor_ghost(min(k0 + kk - 1, nz), nz, pz);
      zp[kk] = p[o];
      const float a = e1[o], b = e2[o];
      ze[kk] = a;
      zd[kk] = a - b;
    }
#pragma unroll
    for (int kk = 0; kk < RZC; ++kk) {
      const int k = k0 + kk;
      const size_t ok = plane * min(k, nz - 1);
      double lp[4];
      float le[4], ld[4];
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        lp[d] = p[cl[d] + ok];
        const float a = e1[cl[d] + ok], b = e2[cl[d] + ok];
        le[d] = a;
        ld[d] = a - b;
      }
      const double rh = (double)rhs[c + ok];
      if (k < nz) {
        auto qv = [&](double pv, float ev, float dv) {
          return pv + (double)ev + beta * (double)dv;
        };
        const double lapq =
            ox * (qv(lp[0], le[0], ld[0]) + qv(lp[1], le[1], ld[1])) +
            oy * (qv(lp[2], le[2], ld[2]) + qv(lp[3], le[3], ld[3])) +
            oz * (qv(zp[kk], ze[kk], zd[kk]) + qv(zp[kk + 2], ze[kk + 2], zd[kk + 2])) +
            A * qv(zp[kk + 1], ze[kk + 1], zd[kk + 1]);
        const double rq = rh - lapq;
        q[c + plane * k] = qv(zp[kk + 1], ze[kk + 1], zd[kk + 1]);
        r[c + plane * k] = (float)rq;
        const double lapd = ox * ((double)ld[0] + (double)ld[1]) +
                            oy * ((double)ld[2] + (double)ld[3]) +
                            oz * ((double)zd[kk] + (double)zd[kk + 2]) + A * (double)zd[kk + 1];
        sb += (rq + beta * lapd) * lapd;
        sd += lapd * lapd;
        if (part0) {
          const double lapp = ox * (lp[0] + lp[1]) + oy * (lp[2] + lp[3]) +
                              oz * (zp[kk] + zp[kk + 2]) + A * zp[kk + 1];
          const double r0 = rh - lapp;
          m0 = fmax(m0, fabs(r0));
          s0 += r0 * r0;
        }
      }
    }
  }
  const size_t b = (size_t)T.b;
  if (part0) {
    m0 = block_max<RNT / O3D_WAVE>(m0, sc);
    s0 = block_sum<RNT / O3D_WAVE>(s0, sc);
    if (threadIdx.x == 0) {
      part0[2 * b] = m0;
      part0[2 * b + 1] = s0;
    }
  }
  sb = block_sum<RNT / O3D_WAVE>(sb, sc);
  sd = block_sum<RNT / O3D_WAVE>(sd, sc);
  if (threadIdx.x == 0) {
    partb[2 * b] = sb;
    partb[2 * b + 1] = sd;
  }
}

// The two beta sums over the blocks (fixed order), out[0..1].
__global__ void __launch_bounds__(1024) beta_final_kernel(const double *__restrict__ partb,
                                                          int nb, double *__restrict__ out) {
  __shared__ double sc[1024 / O3D_WAVE];
  double a = 0.0, d = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    a += partb[2 * b];
    d += partb[2 * b + 1];
  }
  a = block_sum<1024 / O3D_WAVE>(a, sc);
  d = block_sum<1024 / O3D_WAVE>(d, sc);
  if (threadIdx.x == 0) {
    out[0] = a;
    out[1] = d;
  }
}

template <typename TR>
hipError_t launch_guess_adapt(const double *p, const float *e1, const float *e2, const TR *rhs,
                              double beta, double *q, float *r, int nx, int ny, int nz,
                              const int per[3], double ox, double oy, double oz, double A,
                              hipStream_t s, double *part0, double *partb, double *sums) {
  guess_adapt_kernel<TR><<<residual_grid(nx, ny, nz), RNT, 0, s>>>(
      p, e1, e2, rhs, beta, q, r, part0, partb, nx, ny, nz, per[0], per[1], per[2], ox, oy, oz,
      A);
  beta_final_kernel<<<1, 1024, 0, s>>>(partb, residual_blocks(nx, ny, nz), sums);
  return hipGetLastError();
}

template <typename TR>
__global__ void __launch_bounds__(RNT) residual_f32_kernel(const double *__restrict__ p,
                                                           const TR *__restrict__ rhs,
                                                           float *__restrict__ r,
                                                           float *__restrict__ e, int nx, int ny,
                                                           int nz, int px, int py, int pz,
                                                           double ox, double oy, double oz,
                                                           double A) {
  residual_column(p, rhs, nx, ny, nz, px, py, pz, ox, oy, oz, A,
                  [&](size_t o, int, int, int, double res) {
                    r[o] = (float)res;
                    e[o] = 0.0f;
                  });
}

// The same residual stored colour-split (r and e hold the red half-array
// followed by the black one).
template <typename TR>
__global__ void __launch_bounds__(RNT) residual_split_kernel(const double *__restrict__ p,
                                                             const TR *__restrict__ rhs,
                                                             float *__restrict__ r,
                                                             float *__restrict__ e, int nx,
                                                             int ny, int nz, int px, int py,
                                                             int pz, double ox, double oy,
                                                             double oz, double A) {
  const size_t half = (size_t)nx * ny * nz / 2;
  residual_column(p, rhs, nx, ny, nz, px, py, pz, ox, oy, oz, A,
                  [&](size_t, int i, int j, int k, double res) {
                    const size_t o = (size_t)(nx / 2) * (j + (size_t)ny * k) + (i >> 1) +
                                     (((i + j + k) & 1) ? half : 0);
                    r[o] = (float)res;
                    e[o] = 0.0f;
                  });
}

template <typename TR>
hipError_t launch_residual_split(const double *p, const TR *rhs, float *r, float *e, int nx,
                                 int ny, int nz, const int per[3], double ox, double oy,
                                 double oz, double A, hipStream_t s) {
  residual_split_kernel<TR><<<residual_grid(nx, ny, nz), RNT, 0, s>>>(
      p, rhs, r, e, nx, ny, nz, per[0], per[1], per[2], ox, oy, oz, A);
  return hipGetLastError();
}

// p += e with e colour-split (red half-array, then black).
__global__ void __launch_bounds__(256) fold_split_kernel(double *__restrict__ p,
                                                         const float *__restrict__ e, int nx,
                                                         int ny, int nz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= nx) return;
  const size_t half = (size_t)nx * ny * nz / 2;
  const size_t row = j + (size_t)ny * k;
  const size_t o = (size_t)(nx / 2) * row + (i >> 1) + (((i + j + k) & 1) ? half : 0);
  const size_t idx = (size_t)nx * row + i;
  p[idx] = p[idx] + (double)e[o];
}

hipError_t launch_fold_split(double *p, const float *e, int nx, int ny, int nz, hipStream_t s) {
  fold_split_kernel<<<dim3((nx + 255) / 256, ny, nz), 256, 0, s>>>(p, e, nx, ny, nz);
  return hipGetLastError();
}

template <typename TR>
hipError_t launch_residual_f32(const double *p, const TR *rhs, float *r, float *e, int nx, int ny,
                               int nz, const int per[3], double ox, double oy, double oz,
                               double A, hipStream_t s) {
  residual_f32_kernel<TR><<<residual_grid(nx, ny, nz), RNT, 0, s>>>(
      p, rhs, r, e, nx, ny, nz, per[0], per[1], per[2], ox, oy, oz, A);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) fold_kernel(double *__restrict__ p,
                                                   const float *__restrict__ e, int n) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) p[idx] = p[idx] + (double)e[idx];
}

hipError_t launch_fold(double *p, const float *e, int n, hipStream_t s) {
  fold_kernel<<<(n + 255) / 256, 256, 0, s>>>(p, e, n);
  return hipGetLastError();
}

// Residual norms of two iterates against one rhs in one launch: per block
// (max |r0|, sum r0^2, max |r1|, sum r1^2).  With r1s, the residual of p1 is
// also stored colour-split in FP32 (r1s, the first correction solve's rhs
// when p1 becomes the iterate) and e1s is zeroed.
template <typename TR>
__global__ void __launch_bounds__(RNT) residual_norm_pair_kernel(
    const double *__restrict__ p0, const double *__restrict__ p1, const TR *__restrict__ rhs,
    double *__restrict__ part, float *__restrict__ r1s, float *__restrict__ e1s, int nx, int ny,
    int nz, int px, int py, int pz, double ox, double oy, double oz, double A) {
  __shared__ double sc[RNT / O3D_WAVE];
  double m0 = 0.0, s0 = 0.0, m1 = 0.0, s1 = 0.0;
  const size_t half = (size_t)nx * ny * nz / 2;
  if (p0)
    residual_column(p0, rhs, nx, ny, nz, px, py, pz, ox, oy, oz, A,
                    [&](size_t, int, int, int, double r) {
                      m0 = fmax(m0, fabs(r));
                      s0 += r * r;
                    });
  residual_column(p1, rhs, nx, ny, nz, px, py, pz, ox, oy, oz, A,
                  [&](size_t, int i, int j, int k, double r) {
                    m1 = fmax(m1, fabs(r));
                    s1 += r * r;
                    if (r1s) {
                      const size_t o = (size_t)(nx / 2) * (j + (size_t)ny * k) + (i >> 1) +
                                       (((i + j + k) & 1) ? half : 0);
                      r1s[o] = (float)r;
                      e1s[o] = 0.0f;
                    }
                  });
  m0 = block_max<RNT / O3D_WAVE>(m0, sc);
  s0 = block_sum<RNT / O3D_WAVE>(s0, sc);
  m1 = block_max<RNT / O3D_WAVE>(m1, sc);
  s1 = block_sum<RNT / O3D_WAVE>(s1, sc);
  if (threadIdx.x == 0) {
    const size_t b = (size_t)residual_tile().b;
    part[4 * b] = m0;
    part[4 * b + 1] = s0;
    part[4 * b + 2] = m1;
    part[4 * b + 3] = s1;
  }
}

// out[0..3] = (max, sum, max, sum) over the block records, in a fixed order;
// with part0 the first pair comes from its (max, sum) records instead.
__global__ void __launch_bounds__(1024) residual_norm_final_kernel(const double *__restrict__ part,
                                                                  const double *__restrict__ part0,
                                                                  int nb, double *out) {
  __shared__ double sc[1024 / O3D_WAVE];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    v[0] = fmax(v[0], part0 ? part0[2 * b] : part[4 * b]);
    v[1] += part0 ? part0[2 * b + 1] : part[4 * b + 1];
    v[2] = fmax(v[2], part[4 * b + 2]);
    v[3] += part[4 * b + 3];
  }
  v[0] = block_max<1024 / O3D_WAVE>(v[0], sc);
  v[1] = block_sum<1024 / O3D_WAVE>(v[1], sc);
  v[2] = block_max<1024 / O3D_WAVE>(v[2], sc);
  v[3] = block_sum<1024 / O3D_WAVE>(v[3], sc);
  if (threadIdx.x == 0)
    for (int q = 0; q < 4; ++q) out[q] = v[q];
}

template <typename TR>
hipError_t launch_residual_norm_pair(const double *p0, const double *p1, const TR *rhs,
                                     double *part, double *out, int nx, int ny, int nz,
                                     const int per[3], double ox, double oy, double oz, double A,
                                     hipStream_t s, float *r1s, float *e1s,
                                     const double *part0) {
  if (r1s && (nx % 2 != 0 || !e1s)) return hipErrorInvalidValue;
  if (!p0 && !part0) return hipErrorInvalidValue;
  residual_norm_pair_kernel<TR><<<residual_grid(nx, ny, nz), RNT, 0, s>>>(
      p0, p1, rhs, part, r1s, e1s, nx, ny, nz, per[0], per[1], per[2], ox, oy, oz, A);
  residual_norm_final_kernel<<<1, 1024, 0, s>>>(part, p0 ? nullptr : part0,
                                                residual_blocks(nx, ny, nz), out);
  return hipGetLastError();
}

template <typename T>
__global__ void __launch_bounds__(256) copy_kernel(T *__restrict__ d, const T *__restrict__ s,
                                                   int n) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) d[idx] = s[idx];
}
template <typename T>
hipError_t launch_copy(T *dst, const T *src, int n, hipStream_t s) {
  copy_kernel<T><<<(n + 255) / 256, 256, 0, s>>>(dst, src, n);
  return hipGetLastError();
}

template <typename TD, typename TS>
__global__ void __launch_bounds__(256) convert_kernel(TD *__restrict__ d,
                                                      const TS *__restrict__ s, int n) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) d[idx] = (TD)s[idx];
}
template <typename TD, typename TS>
hipError_t launch_convert(TD *dst, const TS *src, int n, hipStream_t s) {
  convert_kernel<TD, TS><<<(n + 255) / 256, 256, 0, s>>>(dst, src, n);
  return hipGetLastError();
}

template <typename T>
__global__ void __launch_bounds__(256) fill_kernel(T *__restrict__ d, T v, int n) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) d[idx] = v;
}
hipError_t launch_fill_f64(double *dst, double v, int n, hipStream_t s) {
  fill_kernel<double><<<(n + 255) / 256, 256, 0, s>>>(dst, v, n);
  return hipGetLastError();
}
hipError_t launch_fill_f32(float *dst, float v, int n, hipStream_t s) {
  fill_kernel<float><<<(n + 255) / 256, 256, 0, s>>>(dst, v, n);
  return hipGetLastError();
}

template int sor_slots<double, double>();
template int sor_slots<double, float>();
template int sor_slots<float, float>();
template SorTiling sor_tiling<double, double>(int, int, int);
template SorTiling sor_tiling<double, float>(int, int, int);
template SorTiling sor_tiling<float, float>(int, int, int);
template hipError_t launch_sor_iteration<double, double>(const SorArgs<double, double> &,
                                                         hipStream_t);
template hipError_t launch_sor_iteration<double, float>(const SorArgs<double, float> &,
                                                        hipStream_t);
template hipError_t launch_sor_iteration<float, float>(const SorArgs<float, float> &,
                                                       hipStream_t);
template hipError_t launch_residual_f32<double>(const double *, const double *, float *, float *,
                                                int, int, int, const int *, double, double,
                                                double, double, hipStream_t);
template hipError_t launch_residual_f32<float>(const double *, const float *, float *, float *,
                                               int, int, int, const int *, double, double, double,
                                               double, hipStream_t);
template hipError_t launch_guess_adapt<float>(const double *, const float *, const float *,
                                              const float *, double, double *, float *, int,
                                              int, int, const int *, double, double, double,
                                              double, hipStream_t, double *, double *,
                                              double *);
template hipError_t launch_guess_adapt<double>(const double *, const float *, const float *,
                                               const double *, double, double *, float *, int,
                                               int, int, const int *, double, double, double,
                                               double, hipStream_t, double *, double *,
                                               double *);
template hipError_t launch_guess_residual<float>(const double *, const float *, const float *,
                                                 double *, float *, int, int, int, const int *,
                                                 double, double, double, double, hipStream_t,
                                                 double *);
template hipError_t launch_guess_residual<double>(const double *, const float *, const double *,
                                                  double *, float *, int, int, int, const int *,
                                                  double, double, double, double, hipStream_t,
                                                  double *);
template hipError_t launch_residual_split<float>(const double *, const float *, float *, float *,
                                                 int, int, int, const int *, double, double,
                                                 double, double, hipStream_t);
template hipError_t launch_residual_norm_pair<double>(const double *, const double *,
                                                      const double *, double *, double *, int,
                                                      int, int, const int *, double, double,
                                                      double, double, hipStream_t, float *,
                                                      float *, const double *);
template hipError_t launch_residual_norm_pair<float>(const double *, const double *,
                                                     const float *, double *, double *, int, int,
                                                     int, const int *, double, double, double,
                                                     double, hipStream_t, float *, float *,
                                                     const double *);
template hipError_t launch_copy<double>(double *, const double *, int, hipStream_t);
template hipError_t launch_copy<float>(float *, const float *, int, hipStream_t);
template hipError_t launch_convert<float, double>(float *, const double *, int, hipStream_t);
template hipError_t launch_convert<double, float>(double *, const float *, int, hipStream_t);
template hipError_t launch_convert<float, float>(float *, const float *, int, hipStream_t);

}  // namespace o3d
