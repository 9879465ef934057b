// C ABI of libo3d.so (declarations and conventions in include/o3d.h).
//
// Two layers:
//  * operator entry points with the reference routines' argument lists; host
//    arrays are staged through device buffers, device arrays used in place;
//  * a device-resident solver that runs whole time steps with all fields in
//    HBM (history ring rotated by index, not copied; statistics reduced on
//    the device; one small read-back per SOR batch and per step).
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/o3d.h"
#include "../../include/o3d_nn.h"
#include "o3d_nn.h"
#include "o3d_mg.h"
#include "o3d_sor.h"
#include "o3d_stencil.h"

using namespace o3d;

namespace {

hipStream_t g_stream = nullptr;

// The coarse-grid correction works on 2^3 block averages (its level 1 is
// the network's grid of O3D_WARM_COARSEN^3 blocks).
constexpr int kMgCoarsen = 2;
// SOR sweeps after the Gauss-Seidel opening before a correction solve that
// starts from a coarse-grid correction ends (O3D_MIXED, inner_max = 0).
constexpr int kMgInnerSor = 8;
std::string g_err;
bool g_init = false;

int fail(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return -1;
}

#define HIPCHK(...)                                                                       \
  do {                                                                                    \
    hipError_t e_ = (__VA_ARGS__);                                                        \
    if (e_ != hipSuccess)                                                                 \
      return fail("%s failed: %s (%s:%d)", #__VA_ARGS__, hipGetErrorString(e_), __FILE__,   \
                  __LINE__);                                                              \
  } while (0)
#define CHK(x)                \
  do {                        \
    int r_ = (x);             \
    if (r_ != 0) return r_;   \
  } while (0)

int ensure_init() {
  if (g_init) return 0;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    return fail("no HIP device available (%s)", hipGetErrorString(e));
  HIPCHK(hipSetDevice(0));
  g_init = true;
  return 0;
}

bool is_device_ptr(const void *p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.isManaged;
}

int check_grid(int nx, int ny, int nz) {
  if (nx < 6 || ny < 6 || nz < 6)
    return fail("grid %dx%dx%d: every axis needs at least 6 nodes (6th-order stencil)", nx, ny,
                nz);
  if ((long long)nx * ny * nz >= (1ll << 31)) return fail("grid too large for 32-bit indexing");
  return 0;
}

// xdir: the caller supports x-Dirichlet (DIRICHLET on axis 0 only).
int check_bc(const int *bc, bool xdir = false) {
  if (!bc) return fail("bc is NULL");
  for (int a = 0; a < 3; ++a) {
    if (bc[a] == DIRICHLET && a == 0 && xdir) continue;
    if (bc[a] != PERIODIC && bc[a] != FREE_SLIP)
      return fail("Unrecognized boundary type %d on axis %d%s", bc[a], a,
                  bc[a] == DIRICHLET ? (a == 0 ? " (x-Dirichlet is not supported by this entry "
                                                 "point)" : " (Dirichlet only on x)") : "");
  }
  return 0;
}

// Device-field registry of the drop-in (o3d_register): host arrays the
// caller keeps for the run (the driver's velocity, pressure, predicted
// velocity and Adams-Bashforth history) get a device mirror that outlives
// the call.  A registered input whose device copy is valid is not copied
// again; a registered output stays on the device, and is copied back at the
// end of the call only when the entry mirrors its host copy (host code reads
// it).  Contract: the host writes a registered array only before the
// library first uses it or after o3d_mark_host_dirty.  Host arrays that are
// not registered are staged through temporary device buffers per call.
struct RegEntry {
  char *host;
  size_t bytes;
  void *dev;
  bool dev_valid, host_valid, mirror;
};
std::vector<RegEntry> g_reg;
double g_h2d = 0.0, g_d2h = 0.0;  // bytes moved between host and device

RegEntry *reg_exact(const void *p) {
  for (RegEntry &e : g_reg)
    if (e.host == (const char *)p) return &e;
  return nullptr;
}

// Before host memory [p, p + bytes) is read or written outside its own entry:
// entries overlapping it whose host copy is stale are copied back, and when
// the range is about to be written, their device copies are invalidated.
int reg_touch(const void *p, size_t bytes, bool write) {
  const char *lo = (const char *)p, *hi = lo + bytes;
  for (RegEntry &e : g_reg) {
    if (e.host == lo || e.host + e.bytes <= lo || hi <= e.host) continue;
    if (!e.host_valid) {
      HIPCHK(hipMemcpyAsync(e.host, e.dev, e.bytes, hipMemcpyDeviceToHost, g_stream));
      HIPCHK(hipStreamSynchronize(g_stream));
      g_d2h += (double)e.bytes;
      e.host_valid = true;
    }
    if (write) e.dev_valid = false;
  }
  return 0;
}

// Staging of host arrays through device buffers for one call.
struct Staged {
  struct Out { void *dev; void *host; size_t bytes; };
  std::vector<void *> allocs;
  std::vector<Out> outs;
  std::vector<RegEntry *> written;
  ~Staged() {
    for (void *p : allocs) (void)hipFree(p);
  }
  template <typename T>
  int in(const T *p, size_t n, const T **dev) {
    if (!p) { *dev = nullptr; return 0; }
    if (is_device_ptr(p)) { *dev = p; return 0; }
    const size_t bytes = n * sizeof(T);
    RegEntry *e = reg_exact(p);
    if (e && e->bytes >= bytes) {
      if (!e->dev_valid) {
        HIPCHK(hipMemcpyAsync(e->dev, e->host, e->bytes, hipMemcpyHostToDevice, g_stream));
        g_h2d += (double)e->bytes;
        e->dev_valid = true;
      }
      *dev = (const T *)e->dev;
      return 0;
    }
    CHK(reg_touch(p, bytes, false));
    void *d = nullptr;
    HIPCHK(hipMalloc(&d, bytes));
    allocs.push_back(d);
    HIPCHK(hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, g_stream));
    g_h2d += (double)bytes;
    *dev = (const T *)d;
    return 0;
  }
  // Output (copy_in: the array is also read, e.g. inout).
  template <typename T>
  int out(T *p, size_t n, T **dev, bool copy_in) {
    if (!p) { *dev = nullptr; return 0; }
    if (is_device_ptr(p)) { *dev = p; return 0; }
    const size_t bytes = n * sizeof(T);
    RegEntry *e = reg_exact(p);
    if (e && e->bytes >= bytes) {
      if (copy_in && !e->dev_valid) {
        HIPCHK(hipMemcpyAsync(e->dev, e->host, e->bytes, hipMemcpyHostToDevice, g_stream));
        g_h2d += (double)e->bytes;
      }
      e->dev_valid = true;
      e->host_valid = false;
      written.push_back(e);
      *dev = (T *)e->dev;
      return 0;
    }
    CHK(reg_touch(p, bytes, true));
    void *d = nullptr;
    HIPCHK(hipMalloc(&d, bytes));
    allocs.push_back(d);
    if (copy_in) {
      HIPCHK(hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, g_stream));
      g_h2d += (double)bytes;
    }
    outs.push_back({d, (void *)p, bytes});
    *dev = (T *)d;
    return 0;
  }
  template <typename T>
  int scratch(size_t n, T **dev) {
    void *d = nullptr;
    HIPCHK(hipMalloc(&d, n * sizeof(T)));
    allocs.push_back(d);
    *dev = (T *)d;
    return 0;
  }
  int finish() {
    for (const Out &o : outs) {
      HIPCHK(hipMemcpyAsync(o.host, o.dev, o.bytes, hipMemcpyDeviceToHost, g_stream));
      g_d2h += (double)o.bytes;
    }
    for (RegEntry *e : written)
      if (e->mirror && !e->host_valid) {
        HIPCHK(hipMemcpyAsync(e->host, e->dev, e->bytes, hipMemcpyDeviceToHost, g_stream));
        g_d2h += (double)e->bytes;
        e->host_valid = true;
      }
    HIPCHK(hipStreamSynchronize(g_stream));
    return 0;
  }
};

template <typename T>
void fill_coef(T dst[3][3], const double h[3], bool second) {
  for (int a = 0; a < 3; ++a) {
    if (second) {
      const D2Coef c = d2coef(h[a]);
      dst[a][0] = T(c.a); dst[a][1] = T(c.b); dst[a][2] = T(c.c);
    } else {
      const D1Coef c = d1coef(h[a]);
      dst[a][0] = T(c.a); dst[a][1] = T(c.b); dst[a][2] = T(c.c);
    }
  }
}

// function_stats layout from a finalized record (src/functions.f90:57-62).
void stats6_from(const FieldPartial &r, int nx, int ny, int nz, double *st) {
  st[0] = r.vmin;
  st[1] = r.vmax;
  st[2] = r.sum / (double)((long long)nx * ny * nz);
  const int ix = r.imax == 0x7fffffff ? 0 : r.imax;
  st[3] = ix % nx + 1;
  st[4] = (ix / nx) % ny + 1;
  st[5] = ix / (nx * ny) + 1;
}

// Solver kind that schemes() binds from the x and y boundary types (the
// x-Dirichlet kinds from y and z).
int solver_from_bc(const int *bc) {
  if (bc[0] == DIRICHLET)
    return SOR_XDIR + (bc[1] != PERIODIC ? 1 : 0) + (bc[2] != PERIODIC ? 2 : 0);
  if (bc[0] == PERIODIC && bc[1] == PERIODIC) return SOR_0000;
  if (bc[0] == PERIODIC && bc[1] == FREE_SLIP) return SOR_0011;
  if (bc[0] == FREE_SLIP && bc[1] == FREE_SLIP) return SOR_111111;
  return -1;
}

int check_red_black(int kind, int nx, int ny, int nz) {
  int per[3];
  solver_periodic(kind, per);
  const int n[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a)
    if (per[a] && (n[a] & 1))
      return fail("red-black SOR needs an even node count along periodic axis %d (n=%d)", a,
                  n[a]);
  return 0;
}

struct SorConst { double ox, oy, oz, A; };
SorConst sor_const(double dx, double dy, double dz) {
  // src/poisson.f90:42-51
  SorConst s;
  const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
  s.ox = 1.0 / dx2; s.oy = 1.0 / dy2; s.oz = 1.0 / dz2;
  const double tx = 2.0 * s.ox, ty = 2.0 * s.oy, tz = 2.0 * s.oz;
  s.A = -(tx + ty + tz);
  return s;
}

// Batched launcher of red-black iterations with device-side stop control.
// Buffers: buf[0] holds the start iterate; launch L reads buf[L & 1] and
// writes buf[(L + 1) & 1].  With `split` (T == TR) each launch is one
// iteration of the colour-split in-place sweeps on `sa` (buf unused).
template <typename T, typename TR>
struct SorRun {
  SorArgs<T, TR> a;
  SorState *st = nullptr;    // device
  SorState *hst = nullptr;   // pinned host copy
  T *buf[2] = {nullptr, nullptr};
  bool split = false;
  SorSplitArgs<T> sa{};
  // Timing of complete batches (all launches did their iterations).  The
  // start event is recorded after a batch's first iteration, so it completes
  // while the queue is busy and the interval holds back-to-back iterations
  // only (recorded before it, on the queue left idle by the host's read of
  // the state, it would also hold the host's submission of the batch).
  double ms_full = 0.0;
  int sweeps_full = 0;
  int launches_full = 0;
  std::vector<hipEvent_t> *events = nullptr;
  // Optional: recorded after iteration a.gs of the first batch.  Those
  // iterations always run in full (they end the solve only on kmax), so a
  // solve that ends inside its first batch still times iterations 2..a.gs.
  hipEvent_t mid = nullptr;

  hipError_t launch() {
    if constexpr (std::is_same<T, TR>::value) {
      if (split) {
        sa.it = a.it;
        return launch_sor_split_iteration<T>(sa, g_stream);
      }
    }
    return launch_sor_iteration<T, TR>(a, g_stream);
  }

  // Buffer holding the result after `sweeps` iterations.
  T *result(int sweeps) const { return buf[sweeps & 1]; }

  int run(double omega, int first_batch, int *sweeps_out) {
    HIPCHK(launch_sor_init(st, omega, 0, g_stream));
    if (split) HIPCHK(launch_sor_split_init(sa.dslot, g_stream));
    int it = 1, L = 0;
    int batch = std::max(1, first_batch);
    bool first = true, mid_set = false;
    while (true) {
      hipEvent_t e0 = (*events)[0], e1 = (*events)[1];
      bool e0_set = false;
      int it0 = 0, L0 = 0;  // iterations and launches before e0
      for (int b = 0; b < batch; ++b) {
        a.it = it;
        a.pin = buf[L & 1];
        a.pout = buf[(L + 1) & 1];
        HIPCHK(launch());
        ++L;
        ++it;
        if (b == 0 && batch > 1) {
          HIPCHK(hipEventRecord(e0, g_stream));
          e0_set = true;
          it0 = it - 1;
          L0 = L;
        }
        if (mid && first && L == a.gs && a.gs >= 2 && a.gs <= a.kmax && e0_set) {
          HIPCHK(hipEventRecord(mid, g_stream));
          mid_set = true;
        }
        if (it - 1 > a.kmax) break;  // that launch closes the state
      }
      HIPCHK(hipEventRecord(e1, g_stream));
      HIPCHK(hipMemcpyAsync(hst, st, sizeof(SorState), hipMemcpyDeviceToHost, g_stream));
      HIPCHK(hipStreamSynchronize(g_stream));
      const int launched = it - 1;  // iterations numbered 1..launched were launched
      if (!hst->done) {
        if (e0_set && launched > it0) {
          float ms = 0.f;
          HIPCHK(hipEventElapsedTime(&ms, e0, e1));
          ms_full += ms;
          sweeps_full += launched - it0;
          launches_full += L - L0;
        }
        if (launched > a.kmax) return fail("SOR state did not close after kmax");
        batch = 8;
        first = false;
        continue;
      }
      if (mid_set && first && hst->sweeps >= a.gs) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, mid));
        ms_full += ms;
        sweeps_full += a.gs - it0;
        launches_full += a.gs - L0;
      }
      break;
    }
    *sweeps_out = hst->sweeps;
    return 0;
  }
};

}  // namespace

// ================================================================ runtime
extern "C" int o3d_init(int device) {
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail("device %d out of range (%d devices)", device, n);
  HIPCHK(hipSetDevice(device));
  g_init = true;
  return 0;
}

extern "C" int o3d_set_stream(void *stream) {
  g_stream = (hipStream_t)stream;
  return 0;
}

extern "C" const char *o3d_last_error(void) { return g_err.c_str(); }

extern "C" int o3d_sync(void) {
  HIPCHK(hipStreamSynchronize(g_stream));
  return 0;
}

extern "C" const char *o3d_version(void) { return "osinco3d-amd 0.1 gfx950"; }

// ============================================================ registry
extern "C" int o3d_register(const void *host, size_t bytes, int mirror) {
  CHK(ensure_init());
  if (!host || bytes == 0) return fail("o3d_register: NULL or empty array");
  if (is_device_ptr(host)) return 0;  // device arrays are used in place anyway
  if (RegEntry *e = reg_exact(host)) {
    if (e->bytes == bytes) {
      e->mirror = e->mirror || mirror != 0;
      return 0;
    }
    return fail("o3d_register: %p is registered with %zu bytes, not %zu", host, e->bytes, bytes);
  }
  const char *lo = (const char *)host;
  for (const RegEntry &e : g_reg)
    if (!(e.host + e.bytes <= lo || lo + bytes <= e.host))
      return fail("o3d_register: %p overlaps a registered array", host);
  void *d = nullptr;
  HIPCHK(hipMalloc(&d, bytes));
  g_reg.push_back({(char *)host, bytes, d, false, true, mirror != 0});
  return 0;
}

extern "C" int o3d_unregister(const void *host) {
  for (size_t q = 0; q < g_reg.size(); ++q)
    if (g_reg[q].host == (const char *)host) {
      RegEntry &e = g_reg[q];
      if (!e.host_valid) {
        HIPCHK(hipMemcpyAsync(e.host, e.dev, e.bytes, hipMemcpyDeviceToHost, g_stream));
        g_d2h += (double)e.bytes;
      }
      HIPCHK(hipStreamSynchronize(g_stream));
      HIPCHK(hipFree(e.dev));
      g_reg.erase(g_reg.begin() + q);
      return 0;
    }
  return fail("o3d_unregister: %p is not registered", host);
}

extern "C" int o3d_mark_host_dirty(const void *host) {
  RegEntry *e = reg_exact(host);
  if (!e) return fail("o3d_mark_host_dirty: %p is not registered", host);
  e->dev_valid = false;
  e->host_valid = true;
  return 0;
}

extern "C" int o3d_sync_host(const void *host) {
  RegEntry *e = reg_exact(host);
  if (!e) return fail("o3d_sync_host: %p is not registered", host);
  if (!e->host_valid) {
    HIPCHK(hipMemcpyAsync(e->host, e->dev, e->bytes, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    g_d2h += (double)e->bytes;
    e->host_valid = true;
  }
  return 0;
}

extern "C" int o3d_transfer_stats(double *h2d_bytes, double *d2h_bytes, int reset) {
  if (h2d_bytes) *h2d_bytes = g_h2d;
  if (d2h_bytes) *d2h_bytes = g_d2h;
  if (reset) g_h2d = g_d2h = 0.0;
  return 0;
}

// ============================================================ measurement
namespace {
// Streaming triad c = a + s b over float4 vectors: two reads and one write
// per element, no reuse, the HBM ceiling the stencil and SOR kernels are
// compared with.  Each thread handles 4 vectors per round (their loads in
// flight together) with non-temporal loads and stores (no cache pollution).
typedef float v4f __attribute__((ext_vector_type(4)));
// Streaming triad c = a + s b: each thread moves TRIAD_VEC 16-byte vectors
// spaced a block apart (every load instruction of a wave covers 1 KiB
// contiguously), one pass over the grid, no loop.  NT selects nontemporal
// loads and stores.
constexpr int TRIAD_VEC = 4;
template <bool NT>
__global__ void __launch_bounds__(256) stream_triad_kernel(const float4 *__restrict__ a4,
                                                           const float4 *__restrict__ b4,
                                                           float4 *__restrict__ c4, size_t n4,
                                                           float s) {
  const v4f *a = reinterpret_cast<const v4f *>(a4);
  const v4f *b = reinterpret_cast<const v4f *>(b4);
  v4f *c = reinterpret_cast<v4f *>(c4);
  const size_t base = (size_t)blockIdx.x * (blockDim.x * TRIAD_VEC) + threadIdx.x;
  v4f x[TRIAD_VEC], y[TRIAD_VEC];
#pragma unroll
  for (int u = 0; u < TRIAD_VEC; ++u) {
    const size_t i = base + (size_t)u * blockDim.x;
    if (i < n4) {
      x[u] = NT ? __builtin_nontemporal_load(a + i) : a[i];
      y[u] = NT ? __builtin_nontemporal_load(b + i) : b[i];
    }
  }
#pragma unroll
  for (int u = 0; u < TRIAD_VEC; ++u) {
    const size_t i = base + (size_t)u * blockDim.x;
    if (i < n4) {
      if (NT)
        __builtin_nontemporal_store(x[u] + s * y[u], c + i);
      else
        c[i] = x[u] + s * y[u];
    }
  }
}
}  // namespace

extern "C" int o3d_stream_probe(size_t bytes_per_buffer, int reps, double *gbs, double *ms) {
  CHK(ensure_init());
  if (reps < 1 || bytes_per_buffer < (1u << 20)) return fail("stream probe: bad arguments");
  const size_t n4 = bytes_per_buffer / 16;
  float4 *buf[3] = {nullptr, nullptr, nullptr};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  for (int q = 0; q < 3 && rc == 0; ++q)
    if (hipMalloc((void **)&buf[q], n4 * 16) != hipSuccess) rc = fail("stream probe: hipMalloc");
  if (rc == 0 && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail("stream probe: events");
  if (rc == 0) {
    for (int q = 0; q < 3; ++q) (void)hipMemsetAsync(buf[q], 0, n4 * 16, g_stream);
    const unsigned blocks = (unsigned)((n4 + 256 * TRIAD_VEC - 1) / (256 * TRIAD_VEC));
    double best = 0.0;
    for (int nt = 0; nt < 2 && rc == 0; ++nt) {
      auto launch = [&](int r) {
        if (nt)
          stream_triad_kernel<true><<<blocks, 256, 0, g_stream>>>(buf[r & 1], buf[1 - (r & 1)],
                                                                  buf[2], n4, 0.5f);
        else
          stream_triad_kernel<false><<<blocks, 256, 0, g_stream>>>(buf[r & 1], buf[1 - (r & 1)],
                                                                   buf[2], n4, 0.5f);
      };
      for (int w = 0; w < 3; ++w) launch(w);
      (void)hipEventRecord(e0, g_stream);
      for (int r = 0; r < reps; ++r) launch(r);
      (void)hipEventRecord(e1, g_stream);
      float t = 0.f;
      if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
        rc = fail("stream probe: timing failed: %s", hipGetErrorString(hipGetLastError()));
        break;
      }
      const double per = (double)t / reps;
      const double g = 3.0 * (double)n4 * 16.0 / (per * 1e-3) / 1e9;
      if (g > best) {
        best = g;
        if (ms) *ms = per;
        if (gbs) *gbs = g;
      }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (int q = 0; q < 3; ++q)
    if (buf[q]) (void)hipFree(buf[q]);
  return rc;
}

// ================================================================ operators
extern "C" int o3d_der(int which, double *df, const double *f, int nx, int ny, int nz,
                       double d) {
  CHK(ensure_init());
  CHK(check_grid(nx, ny, nz));
  if (which < 0 || which > 21) return fail("derivative selector %d out of range", which);
  const size_t n = (size_t)nx * ny * nz;
  Staged s;
  const double *fd;
  double *dfd;
  CHK(s.in(f, n, &fd));
  CHK(s.out(df, n, &dfd, false));
  HIPCHK(launch_der<double>(which, dfd, fd, nx, ny, nz, d, g_stream));
  return s.finish();
}

extern "C" int o3d_der_f32(int which, float *df, const float *f, int nx, int ny, int nz,
                           double d) {
  CHK(ensure_init());
  CHK(check_grid(nx, ny, nz));
  if (which < 0 || which > 21) return fail("derivative selector %d out of range", which);
  const size_t n = (size_t)nx * ny * nz;
  Staged s;
  const float *fd;
  float *dfd;
  CHK(s.in(f, n, &fd));
  CHK(s.out(df, n, &dfd, false));
  HIPCHK(launch_der<float>(which, dfd, fd, nx, ny, nz, d, g_stream));
  return s.finish();
}

namespace {
int finalize_to_host(const FieldPartial *part, int nparts, int stride, int first,
                     FieldPartial *dev_out, FieldPartial *host_out) {
  HIPCHK(launch_finalize_partials(part, nparts, stride, first, dev_out, g_stream));
  HIPCHK(hipMemcpyAsync(host_out, dev_out, sizeof(FieldPartial), hipMemcpyDeviceToHost,
                        g_stream));
  return 0;
}
}  // namespace

extern "C" int o3d_divergence(double *divf, const double *fx, const double *fy, const double *fz,
                              double dx, double dy, double dz, int nx, int ny, int nz,
                              const int *bc, int odd, double divisor, double *stats6) {
  CHK(ensure_init());
  CHK(check_grid(nx, ny, nz));
  CHK(check_bc(bc, true));
  const size_t n = (size_t)nx * ny * nz;
  Staged s;
  DivArgs<double> a;
  CHK(s.in(fx, n, &a.f[0]));
  CHK(s.in(fy, n, &a.f[1]));
  CHK(s.in(fz, n, &a.f[2]));
  CHK(s.out(divf, n, &a.out, false));
  a.nx = nx; a.ny = ny; a.nz = nz;
  for (int q = 0; q < 3; ++q) a.bc[q] = bc[q];
  a.odd = odd != 0;
  a.tl = make_tiling(nx, ny, nz);
  const double h[3] = {dx, dy, dz};
  fill_coef<double>(a.d1, h, false);
  a.divisor = divisor;
  FieldPartial *fin;
  CHK(s.scratch(a.tl.blocks(), &a.part));
  CHK(s.scratch(1, &fin));
  HIPCHK(launch_divergence<double>(a, g_stream));
  FieldPartial r;
  CHK(finalize_to_host(a.part, a.tl.blocks(), 1, 0, fin, &r));
  CHK(s.finish());
  if (stats6) stats6_from(r, nx, ny, nz, stats6);
  return 0;
}

namespace {
// Shared by o3d_nu_t and o3d_predict_velocity: the fused predict kernel with
// the reference's slot semantics.
int predict_common(double *up[3], const double *u[3], double *f1[3], const double *f2[3],
                   const double *f3[3], double *nut, int nh, bool les, double re, double adu,
                   double bdu, double cdu, double cs, double delta, const double h[3], int nx,
                   int ny, int nz, const int *bc, FieldPartial *nut_part) {
  PredictArgs<double> a;
  for (int c = 0; c < 3; ++c) {
    a.u[c] = u[c]; a.up[c] = up[c]; a.f1[c] = f1[c]; a.f2[c] = f2[c]; a.f3[c] = f3[c];
  }
  a.nut = nut;
  a.nut_part = nut_part;
  a.nx = nx; a.ny = ny; a.nz = nz;
  for (int q = 0; q < 3; ++q) a.bc[q] = bc[q];
  a.tl = predict_tiling<double>(nx, ny, nz, les);
  fill_coef<double>(a.d1, h, false);
  fill_coef<double>(a.d2, h, true);
  a.adu = adu; a.bdu = bdu; a.cdu = cdu;
  a.onere = 1.0 / re;
  const double csd = cs * delta;
  a.csd2 = csd * csd;
  HIPCHK(launch_predict<double>(a, nh, les, g_stream));
  return 0;
}
}  // namespace

extern "C" int o3d_nu_t(double *nu_t, const double *ux, const double *uy, const double *uz,
                        double dx, double dy, double dz, double cs, double delta, int nx, int ny,
                        int nz, const int *bc, double *stats6) {
  CHK(ensure_init());
  CHK(check_grid(nx, ny, nz));
  CHK(check_bc(bc, true));
  const size_t n = (size_t)nx * ny * nz;
  Staged s;
  const double *u[3];
  double *nut;
  CHK(s.in(ux, n, &u[0]));
  CHK(s.in(uy, n, &u[1]));
  CHK(s.in(uz, n, &u[2]));
  CHK(s.out(nu_t, n, &nut, false));
  // nu_t shares the fused kernel: the velocity outputs go to scratch.
  double *up[3], *f1[3];
  for (int c = 0; c < 3; ++c) {
    CHK(s.scratch(n, &up[c]));
    CHK(s.scratch(n, &f1[c]));
  }
  const double *f2[3] = {nullptr, nullptr, nullptr};
  const Tiling tl = make_tiling(nx, ny, nz);
  FieldPartial *part, *fin;
  CHK(s.scratch(tl.blocks(), &part));
  CHK(s.scratch(1, &fin));
  const double h[3] = {dx, dy, dz};
  CHK(predict_common(up, u, f1, f2, f2, nut, 1, true, 1.0, 0.0, 0.0, 0.0, cs, delta, h, nx, ny,
                     nz, bc, part));
  FieldPartial r;
  CHK(finalize_to_host(part, tl.blocks(), 1, 0, fin, &r));
  CHK(s.finish());
  if (stats6) stats6_from(r, nx, ny, nz, stats6);
  return 0;
}

extern "C" int o3d_predict_velocity(double *ux_pred, double *uy_pred, double *uz_pred,
                                    const double *ux, const double *uy, const double *uz,
                                    double *fux, double *fuy, double *fuz, double re,
                                    const double *adt, const double *bdt, const double *cdt,
                                    int itime, int itscheme, double dx, double dy, double dz,
                                    int nx, int ny, int nz, int iles, double cs, double delta,
                                    double *nu_t, const int *bc) {
  CHK(ensure_init());
  CHK(check_grid(nx, ny, nz));
  CHK(check_bc(bc, true));
  // Coefficient choice, src/integration.f90:84-105.
  int ci;
  if (itscheme == 1 || itime == 1) ci = 0;
  else if (itscheme == 2 || itime == 2) ci = 1;
  else if (itscheme == 3) ci = 2;
  else { fail("itscheme: %d unrecognized", itscheme); return -2; }
  const size_t n = (size_t)nx * ny * nz;
  Staged s;
  const double *u[3];
  double *up[3], *F[3], *nut;
  CHK(s.in(ux, n, &u[0]));
  CHK(s.in(uy, n, &u[1]));
  CHK(s.in(uz, n, &u[2]));
  CHK(s.out(ux_pred, n, &up[0], false));
  CHK(s.out(uy_pred, n, &up[1], false));
  CHK(s.out(uz_pred, n, &up[2], false));
  CHK(s.out(fux, 3 * n, &F[0], true));
  CHK(s.out(fuy, 3 * n, &F[1], true));
  CHK(s.out(fuz, 3 * n, &F[2], true));
  const bool les = iles == 1;
  nut = nullptr;
  if (les) {
    CHK(s.out(nu_t, n, &nut, false));
  } else if (nu_t) {
    // Without LES the reference sets nu_t = 0 (src/integration.f90:106-113):
    // on the device for a device array, on both copies of a registered one,
    // on the host otherwise (no transfer either way).
    RegEntry *e = is_device_ptr(nu_t) ? nullptr : reg_exact(nu_t);
    if (is_device_ptr(nu_t) || e) {
      HIPCHK(launch_fill_f64(e ? (double *)e->dev : nu_t, 0.0, (int)n, g_stream));
      if (e) { e->dev_valid = true; e->host_valid = true; }
    }
    if (!is_device_ptr(nu_t)) {
      CHK(reg_touch(nu_t, n * sizeof(double), true));
      memset(nu_t, 0, n * sizeof(double));
    }
  }
  double *f1[3];
  const double *f2[3], *f3[3];
  for (int c = 0; c < 3; ++c) { f1[c] = F[c]; f2[c] = F[c] + n; f3[c] = F[c] + 2 * n; }
  const Tiling tl = make_tiling(nx, ny, nz);
  FieldPartial *part;
  CHK(s.scratch(tl.blocks(), &part));
  const double h[3] = {dx, dy, dz};
  CHK(predict_common(up, u, f1, f2, f3, les ? nut : nullptr, ci + 1, les, re, adt[ci], bdt[ci],
                     cdt[ci], cs, delta, h, nx, ny, nz, bc, part));
  // History shift by copy, src/integration.f90:176-188.
  for (int c = 0; c < 3; ++c) {
    if (itscheme == 2) {
      HIPCHK(launch_copy<double>(F[c] + n, F[c], (int)n, g_stream));
    } else if (itscheme == 3) {
      HIPCHK(launch_copy<double>(F[c] + 2 * n, F[c] + n, (int)n, g_stream));
      HIPCHK(launch_copy<double>(F[c] + n, F[c], (int)n, g_stream));
    }
  }
  return s.finish();
}

// The colour-split sweep arguments matching `a` (red, black and the rhs
// halves already in split storage).
template <typename T>
SorSplitArgs<T> split_args(const SorArgs<T, T> &a, T *red, T *black, const T *rhs_red,
                           const T *rhs_black, unsigned long long *dslot) {
  SorSplitArgs<T> sa{};
  sa.red = red;
  sa.black = black;
  sa.rhs_red = rhs_red;
  sa.rhs_black = rhs_black;
  sa.st = a.st;
  sa.dslot = dslot;
  sa.log = a.log;
  sa.log_len = a.log_len;
  sa.nx = a.nx; sa.ny = a.ny; sa.nz = a.nz;
  for (int q = 0; q < 3; ++q) sa.per[q] = a.per[q];
  sa.ox = a.ox; sa.oy = a.oy; sa.oz = a.oz; sa.A = a.A;
  sa.eps = a.eps;
  sa.factor = a.factor;
  sa.kmax = a.kmax;
  sa.idyn = a.idyn;
  sa.it = a.it;
  sa.gs = a.gs;
  return sa;
}

namespace {
// variant: 0 one fused iteration per launch, 2 colour-split half sweeps,
// -1 the colour-split sweeps where the grid allows them, else the fused
// kernel (identical iterates in every case).
template <typename T>
int poisson_dev(T *pp, const T *rhs, double dx, double dy, double dz, int nx, int ny, int nz,
                double *omega, double eps, int kmax, int idyn, int solver, int *exit_iter,
                double *dmax, int *reason, Staged &s, int variant = -1) {
  if (solver < 0 || solver >= SOR_KINDS)
    return fail("no Poisson solver bound for these boundaries");
  CHK(check_red_black(solver, nx, ny, nz));
  if (kmax < 1) return fail("kmax must be >= 1");
  if (!sor_buffer_fits<T, T>(nx, ny, nz))
    return fail("SOR: a %dx%dx%d field exceeds the 2 GiB a buffer descriptor addresses", nx, ny,
                nz);
  const int n = nx * ny * nz;
  SorRun<T, T> r;
  CHK(s.scratch(1, &r.st));
  HIPCHK(hipHostMalloc((void **)&r.hst, sizeof(SorState)));
  std::vector<hipEvent_t> ev(2);
  for (auto &e : ev) HIPCHK(hipEventCreate(&e));
  r.events = &ev;
  T *other;
  CHK(s.scratch(n, &other));
  r.buf[0] = pp;
  r.buf[1] = other;
  SorArgs<T, T> &a = r.a;
  a.rhs = rhs;
  a.st = r.st;
  a.log = nullptr;
  a.log_len = 0;
  a.nx = nx; a.ny = ny; a.nz = nz;
  solver_periodic(solver, a.per);
  a.tl = sor_tiling<T, T>(nx, ny, nz);
  const SorConst c = sor_const(dx, dy, dz);
  a.ox = c.ox; a.oy = c.oy; a.oz = c.oz; a.A = c.A;
  a.eps = eps;
  a.factor = solver == SOR_0011 ? 1.01 : 1.05;  // src/poisson.f90:38,162,285
  a.kmax = kmax;
  a.idyn = idyn;
  a.it = 0;
  a.gs = 0;
  int sweeps = 0;
  int rc = 0;
  if (variant == 2 || (variant < 0 && sor_split_supported<T>(nx, ny, nz))) {
    // Colour-split copies of the iterate (in `other`) and of the rhs.
    T *rs;
    unsigned long long *dslot;
    CHK(s.scratch(n, &rs));
    CHK(s.scratch(sor_split_slot_words(), &dslot));
    const int h = n / 2;
    HIPCHK(launch_split<T>(pp, other, other + h, nx, ny, nz, g_stream));
    HIPCHK(launch_split<T>(rhs, rs, rs + h, nx, ny, nz, g_stream));
    r.split = true;
    r.sa = split_args<T>(a, other, other + h, rs, rs + h, dslot);
    rc = r.run(*omega, 16, &sweeps);
    if (rc == 0) HIPCHK(launch_unsplit<T>(pp, other, other + h, nx, ny, nz, g_stream));
  }
  if (!r.split) {
    rc = r.run(*omega, 16, &sweeps);
    if (rc == 0 && r.result(sweeps) != pp)
      rc = launch_copy<T>(pp, other, n, g_stream) == hipSuccess ? 0 : fail("copy failed");
  }
  if (rc == 0) {
    *omega = r.hst->omega_final;
    if (exit_iter) *exit_iter = r.hst->exit_iter;
    if (dmax) *dmax = r.hst->dmax_final;
    if (reason) *reason = r.hst->reason;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  (void)hipHostFree(r.hst);
  return rc;
}
}  // namespace

extern "C" int o3d_poisson(double *pp, const double *rhs, double dx, double dy, double dz, int nx,
                           int ny, int nz, double *omega, double eps, int kmax, int idyn,
                           int solver, int *exit_iter, double *dmax, int *reason) {
  CHK(ensure_init());
  CHK(check_grid(nx, ny, nz));
  if (!omega) return fail("omega is NULL");
  const size_t n = (size_t)nx * ny * nz;
  Staged s;
  double *p;
  const double *r;
  CHK(s.out(pp, n, &p, true));
  CHK(s.in(rhs, n, &r));
  CHK(poisson_dev(p, r, dx, dy, dz, nx, ny, nz, omega, eps, kmax, idyn, solver, exit_iter, dmax,
                  reason, s));
  return s.finish();
}

extern "C" int o3d_multigrid(float *u, const float *f, int mx, int my, int mz, double hx,
                             double hy, double hz, const int *per, int cycles) {
  CHK(ensure_init());
  if (mx < 1 || my < 1 || mz < 1 || (long long)mx * my * mz >= (1ll << 31))
    return fail("multigrid grid %dx%dx%d out of range", mx, my, mz);
  if (!per || !u || !f) return fail("multigrid: NULL argument");
  if (cycles < 1) return fail("multigrid: cycles must be >= 1");
  const size_t n = (size_t)mx * my * mz;
  Staged s;
  float *du;
  const float *df;
  CHK(s.out(u, n, &du, false));
  CHK(s.in(f, n, &df));
  MgHierarchy H;
  const int m0[3] = {mx, my, mz};
  const double h[3] = {hx, hy, hz};
  const int p01[3] = {per[0] != 0, per[1] != 0, per[2] != 0};
  float *buf;
  CHK(s.scratch(mg_plan(H, m0, h, p01), &buf));
  mg_bind(H, buf);
  HIPCHK(hipMemcpyAsync(H.f[0], df, n * sizeof(float), hipMemcpyDeviceToDevice, g_stream));
  HIPCHK(mg_solve(H, cycles, 2, 2, g_stream));
  HIPCHK(hipMemcpyAsync(du, H.u[0], n * sizeof(float), hipMemcpyDeviceToDevice, g_stream));
  return s.finish();
}

namespace {
// ---- x-Dirichlet inflow / outflow (no counterpart in the reference) --------
// Flux weight of node q of an axis of n nodes: 1/2 on the end nodes of a
// mirrored axis (the trapezoid weights that make the mirrored 7-point
// operator's rows sum to zero), else 1.
__host__ __device__ __forceinline__ double edge_weight(int q, int n, int mirrored) {
  return (mirrored && (q == 0 || q == n - 1)) ? 0.5 : 1.0;
}

// SplitMix64 (Steele, Lea, Flood 2014) as a counter-based generator: output
// number counter + 1 of the stream seeded with `seed`, as a uniform double
// in [-1, 1) from the top 53 bits.
__host__ __device__ __forceinline__ double noise_uniform(unsigned long long seed,
                                                         long long counter) {
  unsigned long long z = seed + (unsigned long long)(counter + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * 0x1.0p-53 * 2.0 - 1.0;
}

// After predict, one thread per (j, k): the inflow plane i = 0 gets the
// profile plus the noise (amp w)(x1 C + x2 S) with x1, x2 drawn per node and
// component, C, S the cosine and sine of the inflow phase; the outflow plane
// i = nx - 1 gets first-order upwind convection at the convective velocity
// (cu = Uc dt / dx) from the velocity before the step.  The x components of
// both planes go to pin / pout for the flux balance.
template <typename T>
__global__ void __launch_bounds__(256) jet_edges_kernel(
    T *__restrict__ up0, T *__restrict__ up1, T *__restrict__ up2, const T *__restrict__ u0,
    const T *__restrict__ u1, const T *__restrict__ u2, const double *__restrict__ base,
    const double *__restrict__ w, double amp, double cph, double sph, unsigned long long seed,
    double cu, int nx, int ny, int nz, double *__restrict__ pin, double *__restrict__ pout) {
  const int P = ny * nz;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= P) return;
  const int j = q % ny, k = q / ny;
  const int row = nx * (j + ny * k);
  T *up[3] = {up0, up1, up2};
  const T *u[3] = {u0, u1, u2};
  const double aw = amp * w[q];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const long long cnt = 2 * (3 * (long long)q + c);
    const double x1 = noise_uniform(seed, cnt), x2 = noise_uniform(seed, cnt + 1);
    const double vin = base[c * P + q] + aw * (x1 * cph + x2 * sph);
    const double ub = (double)u[c][row + nx - 1], ub1 = (double)u[c][row + nx - 2];
    const double vout = ub - cu * (ub - ub1);
    up[c][row] = T(vin);
    up[c][row + nx - 1] = T(vout);
    if (c == 0) {
      pin[q] = (double)T(vin);
      pout[q] = (double)T(vout);
    }
  }
}

// The plane sums in a fixed order (per z row the sequential sum over j of
// wy v, then one thread's sequential sum over k of wz row), and the shift
// (Q_in - Q_out) / W that makes the outflow flux equal the inflow flux.
__global__ void __launch_bounds__(256) jet_flux_kernel(const double *__restrict__ pin,
                                                       const double *__restrict__ pout,
                                                       double *__restrict__ rows, int ny,
                                                       int nz, int my, int mz, double W,
                                                       double *__restrict__ flux) {
  for (int k = threadIdx.x; k < nz; k += blockDim.x) {
    double a = 0.0, b = 0.0;
    for (int j = 0; j < ny; ++j) {
      const double wy = edge_weight(j, ny, my);
      a = a + wy * pin[j + ny * k];
      b = b + wy * pout[j + ny * k];
    }
    rows[k] = a;
    rows[nz + k] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double qi = 0.0, qo = 0.0;
    for (int k = 0; k < nz; ++k) {
      const double wz = edge_weight(k, nz, mz);
      qi = qi + wz * rows[k];
      qo = qo + wz * rows[nz + k];
    }
    flux[0] = qi;
    flux[1] = qo;
    flux[2] = (qi - qo) / W;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) jet_shift_kernel(T *__restrict__ up0,
                                                        const double *__restrict__ pout,
                                                        const double *__restrict__ flux, int nx,
                                                        int ny, int nz) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ny * nz) return;
  const int j = q % ny, k = q / ny;
  up0[nx * (j + ny * k) + nx - 1] = T(pout[q] + flux[2]);
}

// Weighted sum of a field (f null: of the weights alone) with the trapezoid
// weights of the mirrored axes: kWsumBlocks chunks of ceil(n / kWsumBlocks)
// cells, thread t of a chunk summing cells lo + t, lo + t + 256, ... in
// order, the chunk total by block_sum<4>.
constexpr int kWsumBlocks = 1024;
template <typename T>
__global__ void __launch_bounds__(256) wsum_kernel(const T *__restrict__ f, int nx, int ny,
                                                   int nz, int mx, int my, int mz,
                                                   double *__restrict__ part) {
  __shared__ double sv[4];
  const long long n = (long long)nx * ny * nz;
  const long long per = (n + kWsumBlocks - 1) / kWsumBlocks;
  const long long lo = (long long)blockIdx.x * per;
  const long long hi = lo + per < n ? lo + per : n;
  double acc = 0.0;
  for (long long q = lo + threadIdx.x; q < hi; q += 256) {
    const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((long long)nx * ny));
    const double wq = (edge_weight(i, nx, mx) * edge_weight(j, ny, my)) * edge_weight(k, nz, mz);
    acc = acc + (f ? wq * (double)f[q] : wq);
  }
  acc = block_sum<4>(acc, sv);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// The chunk totals summed by 1024 threads (one each) with block_sum<16>;
// out = total, or total / *den when den is given.
__global__ void __launch_bounds__(1024) wsum_final_kernel(const double *__restrict__ part,
                                                          const double *__restrict__ den,
                                                          double *__restrict__ out) {
  __shared__ double sv[16];
  double v = 0.0;
  if (threadIdx.x < kWsumBlocks) v = v + part[threadIdx.x];
  v = block_sum<16>(v, sv);
  if (threadIdx.x == 0) *out = den ? v / *den : v;
}

template <typename T>
__global__ void __launch_bounds__(256) sub_mean_kernel(T *__restrict__ f, long long n,
                                                       const double *__restrict__ mean) {
  const double m = *mean;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x)
    f[q] = T((double)f[q] - m);
}

}  // namespace

// Multigrid solve of the pressure Poisson problem of the SOR solvers (same
// 7-point operator and neighbour rules).  A cycle: the FP64 residual of pp
// (stored colour-split in FP32), stop when its maximum is below tol, the
// coarse-grid correction (residual averaged over 2^3 blocks, one V(npre,
// npost) cycle on the block grid and its coarser levels, trilinear
// interpolation), npre + npost red-black Gauss-Seidel sweeps of the
// correction equation on the full grid from that correction, and pp += the
// correction in FP64.  The residual is formed in FP64 every cycle, so the
// FP32 corrections do not limit the attainable accuracy.
extern "C" int o3d_poisson_multigrid(double *pp, const double *rhs, double dx, double dy,
                                     double dz, int nx, int ny, int nz, int solver, int nlevels,
                                     int npre, int npost, double tol, int max_cycles,
                                     int *cycles, double *resmax, int *reason) {
  CHK(ensure_init());
  CHK(check_grid(nx, ny, nz));
  if (solver < 0 || solver >= SOR_KINDS)
    return fail("no Poisson solver bound for these boundaries");
  CHK(check_red_black(solver, nx, ny, nz));
  if (!sor_split_supported<float>(nx, ny, nz) || ny % 2 || nz % 2)
    return fail("multigrid needs nx a multiple of 8 and ny, nz even (grid %dx%dx%d)", nx, ny,
                nz);
  if (npre < 0 || npost < 0 || npre + npost < 1)
    return fail("multigrid: npre, npost must be >= 0 with npre + npost >= 1");
  if (max_cycles < 1) return fail("multigrid: max_cycles must be >= 1");
  if (!(tol >= 0.0)) return fail("multigrid: tol must be >= 0");
  const size_t n = (size_t)nx * ny * nz;
  Staged s;
  double *p;
  const double *rhs_d;
  CHK(s.out(pp, n, &p, true));
  CHK(s.in(rhs, n, &rhs_d));
  int per[3];
  solver_periodic(solver, per);
  const SorConst c = sor_const(dx, dy, dz);
  if (!per[0] || !per[1] || !per[2]) {
    // A mirrored axis: the operator's columns sum to zero under the
    // trapezoid weights (1/2 on the end nodes), so only the rhs with zero
    // weighted mean can be met.  The solve works on that compatible part
    // (the rhs minus its weighted mean), and max |r| is judged on it: an
    // incompatible rhs would otherwise leave a part no cycle can reduce.
    double *rc, *wp;
    CHK(s.scratch(n, &rc));
    CHK(s.scratch((size_t)kWsumBlocks + 2, &wp));
    wsum_kernel<double><<<kWsumBlocks, 256, 0, g_stream>>>(nullptr, nx, ny, nz, !per[0],
                                                           !per[1], !per[2], wp);
    wsum_final_kernel<<<1, 1024, 0, g_stream>>>(wp, nullptr, wp + kWsumBlocks);
    wsum_kernel<double><<<kWsumBlocks, 256, 0, g_stream>>>(rhs_d, nx, ny, nz, !per[0],
                                                           !per[1], !per[2], wp);
    wsum_final_kernel<<<1, 1024, 0, g_stream>>>(wp, wp + kWsumBlocks, wp + kWsumBlocks + 1);
    HIPCHK(hipMemcpyAsync(rc, rhs_d, n * sizeof(double), hipMemcpyDeviceToDevice, g_stream));
    sub_mean_kernel<double><<<2048, 256, 0, g_stream>>>(rc, (long long)n, wp + kWsumBlocks + 1);
    HIPCHK(hipGetLastError());
    rhs_d = rc;
  }
  // The coarse levels: the grid of 2^3 blocks and the levels below it, at
  // most nlevels - 1 of them (nlevels counts the full grid).
  MgHierarchy H;
  const int m0[3] = {nx / 2, ny / 2, nz / 2};
  const double h[3] = {2.0 * dx, 2.0 * dy, 2.0 * dz};
  float *mgbuf;
  CHK(s.scratch(mg_plan(H, m0, h, per), &mgbuf));
  mg_bind(H, mgbuf);
  H.levels = std::min(H.levels, std::max(nlevels - 1, 0));
  float *R, *E;
  CHK(s.scratch(n, &R));
  CHK(s.scratch(n, &E));
  const int nb = residual_blocks(nx, ny, nz);
  double *part;
  CHK(s.scratch((size_t)4 * nb + 4, &part));
  double *fin = part + 4 * (size_t)nb;
  // Gauss-Seidel sweeps of A e = r (omega = 1 for all of them, ending on
  // the sweep count only).
  SorRun<float, float> run;
  CHK(s.scratch(1, &run.st));
  unsigned long long *dslot;
  CHK(s.scratch(sor_split_slot_words(), &dslot));
  HIPCHK(hipHostMalloc((void **)&run.hst, sizeof(SorState)));
  double *hfin = nullptr;
  if (hipHostMalloc((void **)&hfin, 4 * sizeof(double)) != hipSuccess) {
    (void)hipHostFree(run.hst);
    return fail("multigrid: pinned allocation failed");
  }
  std::vector<hipEvent_t> ev(2);
  for (auto &e : ev) (void)hipEventCreate(&e);
  run.events = &ev;
  SorArgs<float, float> &a = run.a;
  a.rhs = nullptr;
  a.st = run.st;
  a.log = nullptr;
  a.log_len = 0;
  a.nx = nx; a.ny = ny; a.nz = nz;
  for (int q = 0; q < 3; ++q) a.per[q] = per[q];
  a.ox = c.ox; a.oy = c.oy; a.oz = c.oz; a.A = c.A;
  a.eps = 0.0;
  a.factor = 1.0;
  a.kmax = npre + npost;
  a.idyn = 0;
  a.it = 0;
  a.gs = npre + npost;
  const size_t half = n / 2;
  run.split = true;
  run.sa = split_args<float>(a, E, E + half, R, R + half, dslot);
  int rc = 0, done = 0, why = 2;
  double rmax = 0.0, rmax_old = 1609.0;
  while (rc == 0) {
    // r = rhs - L pp (FP64 arithmetic), colour-split in FP32, and e = 0.
    if (launch_residual_norm_pair<double>(p, p, rhs_d, part, fin, nx, ny, nz, per, c.ox, c.oy,
                                          c.oz, c.A, g_stream, R, E) != hipSuccess ||
        hipMemcpyAsync(hfin, fin, 4 * sizeof(double), hipMemcpyDeviceToHost, g_stream) !=
            hipSuccess ||
        hipStreamSynchronize(g_stream) != hipSuccess) {
      rc = fail("multigrid: residual failed: %s", hipGetErrorString(hipGetLastError()));
      break;
    }
    rmax = hfin[2];
    // The stop rules of the SOR solvers (src/poisson.f90:95-122) on max |r|:
    // below tol (or NaN), stagnation (a change below tol / 1000 between
    // cycles: on a mirrored grid the part of an incompatible rhs that the
    // operator cannot reach keeps max |r| from falling below tol), or out
    // of cycles.
    if (!(rmax >= tol)) { why = 0; break; }
    if (done > 0 && fabs(rmax - rmax_old) < tol / 1000.0) { why = 1; break; }
    if (done >= max_cycles) { why = 2; break; }
    rmax_old = rmax;
    if (H.levels > 0) {
      if (mg_pool_split(R, H.f[0], nx, ny, nz, g_stream) != hipSuccess ||
          mg_solve(H, 1, std::max(npre, 1), std::max(npost, 1), g_stream) != hipSuccess ||
          mg_prolong_split(E, H.u[0], nx, ny, nz, per, g_stream) != hipSuccess) {
        rc = fail("multigrid: coarse-grid correction failed");
        break;
      }
    }
    int sweeps = 0;
    if ((rc = run.run(1.0, 16, &sweeps)) != 0) break;
    if (launch_fold_split(p, E, nx, ny, nz, g_stream) != hipSuccess) {
      rc = fail("multigrid: fold failed");
      break;
    }
    ++done;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  (void)hipHostFree(run.hst);
  (void)hipHostFree(hfin);
  if (rc != 0) return rc;
  if (cycles) *cycles = done;
  if (resmax) *resmax = rmax;
  if (reason) *reason = why;
  return s.finish();
}

extern "C" int o3d_poisson_f32(float *pp, const float *rhs, double dx, double dy, double dz,
                               int nx, int ny, int nz, double *omega, double eps, int kmax,
                               int idyn, int solver, int variant, int *exit_iter, double *dmax,
                               int *reason) {
  CHK(ensure_init());
  CHK(check_grid(nx, ny, nz));
  if (!omega) return fail("omega is NULL");
  if (variant != 0 && variant != 2) return fail("SOR kernel variant %d unknown", variant);
  if (variant == 2 && !sor_split_supported<float>(nx, ny, nz))
    return fail("colour-split sweeps need nx a multiple of 8 (nx = %d)", nx);
  const size_t n = (size_t)nx * ny * nz;
  Staged s;
  float *p;
  const float *r;
  CHK(s.out(pp, n, &p, true));
  CHK(s.in(rhs, n, &r));
  CHK(poisson_dev<float>(p, r, dx, dy, dz, nx, ny, nz, omega, eps, kmax, idyn, solver, exit_iter,
                         dmax, reason, s, variant));
  return s.finish();
}

extern "C" int o3d_correct_pression(double *pp, const double *ux_pred, const double *uy_pred,
                                    const double *uz_pred, double dx, double dy, double dz,
                                    int nx, int ny, int nz, double dt, double *omega, double eps,
                                    int kmax, int idyn, const int *bc, int *exit_iter,
                                    double *dmax, int *reason) {
  CHK(ensure_init());
  CHK(check_grid(nx, ny, nz));
  CHK(check_bc(bc));
  if (!omega) return fail("omega is NULL");
  const size_t n = (size_t)nx * ny * nz;
  Staged s;
  double *p, *rhs;
  DivArgs<double> a;
  CHK(s.out(pp, n, &p, true));
  CHK(s.in(ux_pred, n, &a.f[0]));
  CHK(s.in(uy_pred, n, &a.f[1]));
  CHK(s.in(uz_pred, n, &a.f[2]));
  CHK(s.scratch(n, &rhs));
  a.out = rhs;
  a.nx = nx; a.ny = ny; a.nz = nz;
  for (int q = 0; q < 3; ++q) a.bc[q] = bc[q];
  a.odd = 1;  // src/integration.f90:235-236
  a.tl = div_tiling<double>(nx, ny, nz);
  const double h[3] = {dx, dy, dz};
  fill_coef<double>(a.d1, h, false);
  a.divisor = dt;
  CHK(s.scratch(a.tl.blocks(), &a.part));
  HIPCHK(launch_divergence<double>(a, g_stream));
  CHK(poisson_dev(p, rhs, dx, dy, dz, nx, ny, nz, omega, eps, kmax, idyn, solver_from_bc(bc),
                  exit_iter, dmax, reason, s));
  return s.finish();
}

extern "C" int o3d_correct_velocity(double *ux, double *uy, double *uz, const double *ux_pred,
                                    const double *uy_pred, const double *uz_pred,
                                    const double *pp, double dt, double dx, double dy, double dz,
                                    int nx, int ny, int nz, const int *bc, double *vstats) {
  CHK(ensure_init());
  CHK(check_grid(nx, ny, nz));
  CHK(check_bc(bc, true));
  const size_t n = (size_t)nx * ny * nz;
  Staged s;
  CorrectArgs<double, double> a;
  CHK(s.out(ux, n, &a.u[0], false));
  CHK(s.out(uy, n, &a.u[1], false));
  CHK(s.out(uz, n, &a.u[2], false));
  CHK(s.in(ux_pred, n, &a.up[0]));
  CHK(s.in(uy_pred, n, &a.up[1]));
  CHK(s.in(uz_pred, n, &a.up[2]));
  CHK(s.in(pp, n, &a.p));
  a.nx = nx; a.ny = ny; a.nz = nz;
  for (int q = 0; q < 3; ++q) a.bc[q] = bc[q];
  a.tl = correct_tiling<double, double>(nx, ny, nz);
  const double h[3] = {dx, dy, dz};
  fill_coef<double>(a.d1, h, false);
  a.dt = dt;
  FieldPartial *fin;
  CHK(s.scratch(3 * a.tl.blocks(), &a.part));
  CHK(s.scratch(3, &fin));
  HIPCHK(launch_correct<double, double>(a, g_stream));
  FieldPartial r[3];
  for (int c = 0; c < 3; ++c) CHK(finalize_to_host(a.part, a.tl.blocks(), 3, c, fin + c, &r[c]));
  CHK(s.finish());
  int diverged = 0;
  for (int c = 0; c < 3; ++c) {
    if (vstats) {
      vstats[4 * c + 0] = r[c].vmin;
      vstats[4 * c + 1] = r[c].vmax;
      vstats[4 * c + 2] = r[c].amax;
      vstats[4 * c + 3] = r[c].nan;
    }
    if (r[c].nan || r[c].vmax > 1000.0) diverged = 1;
  }
  return diverged;
}

extern "C" int o3d_function_stats(const double *f, int nx, int ny, int nz, double *stats6) {
  CHK(ensure_init());
  if (!stats6) return fail("stats6 is NULL");
  const size_t n = (size_t)nx * ny * nz;
  Staged s;
  const double *fd;
  CHK(s.in(f, n, &fd));
  const int nb = std::max(1, std::min(2048, (int)((n + 4095) / 4096)));
  FieldPartial *part, *fin;
  CHK(s.scratch(nb, &part));
  CHK(s.scratch(1, &fin));
  HIPCHK(launch_field_stats<double>(fd, (int)n, part, nb, g_stream));
  FieldPartial r;
  CHK(finalize_to_host(part, nb, 1, 0, fin, &r));
  CHK(s.finish());
  stats6_from(r, nx, ny, nz, stats6);
  return 0;
}

namespace {
// statistics_calc on device fields with the per-block partial sums in
// `part` (16 doubles per block of make_tiling) and the 16 finalized sums in
// `fin`; out17 is host memory.
template <typename T>
int statistics_run(const T *u[3], int nx, int ny, int nz, double dx, double dy, double dz,
                   double re, double t, const int *bc, double *out17, double *part,
                   double *fin) {
  StatsArgs<T> a;
  for (int c = 0; c < 3; ++c) a.u[c] = u[c];
  a.nx = nx; a.ny = ny; a.nz = nz;
  for (int q = 0; q < 3; ++q) a.bc[q] = bc[q];
  a.tl = make_tiling(nx, ny, nz);
  const double h[3] = {dx, dy, dz};
  fill_coef<T>(a.d1, h, false);
  fill_coef<T>(a.d2, h, true);
  a.xnu = 1.0 / re;
  a.part = part;
  HIPCHK(launch_statistics<T>(a, g_stream));
  HIPCHK(launch_finalize_sums(a.part, a.tl.blocks(), 16, fin, g_stream));
  double h16[16];
  HIPCHK(hipMemcpyAsync(h16, fin, sizeof(h16), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  const double N = (double)((long long)nx * ny * nz);
  out17[0] = t;
  out17[1] = h16[0] / N;  // E_k
  out17[2] = h16[1] / N;  // eps
  out17[3] = h16[2] / N;  // eps2
  out17[4] = h16[3] / N;  // enstrophy
  for (int m = 0; m < 3; ++m) out17[5 + m] = h16[4 + m] / N;
  for (int m = 0; m < 9; ++m) out17[8 + m] = h16[7 + m] / N;
  return 0;
}

template <typename T>
int statistics_dev(const T *u[3], int nx, int ny, int nz, double dx, double dy, double dz,
                   double re, double t, const int *bc, double *out17, Staged &s) {
  double *part, *fin;
  CHK(s.scratch((size_t)16 * make_tiling(nx, ny, nz).blocks(), &part));
  CHK(s.scratch(16, &fin));
  CHK(statistics_run<T>(u, nx, ny, nz, dx, dy, dz, re, t, bc, out17, part, fin));
  return s.finish();
}
}  // namespace

extern "C" int o3d_statistics(const double *ux, const double *uy, const double *uz, int nx,
                              int ny, int nz, double dx, double dy, double dz, double re,
                              double t, const int *bc, double *out17) {
  CHK(ensure_init());
  CHK(check_grid(nx, ny, nz));
  CHK(check_bc(bc, true));
  if (!out17) return fail("out17 is NULL");
  const size_t n = (size_t)nx * ny * nz;
  Staged s;
  const double *u[3];
  CHK(s.in(ux, n, &u[0]));
  CHK(s.in(uy, n, &u[1]));
  CHK(s.in(uz, n, &u[2]));
  return statistics_dev<double>(u, nx, ny, nz, dx, dy, dz, re, t, bc, out17, s);
}

// ---------------------------------------------- statistics without allocation
// Device memory of o3d_statistics_noalloc: the per-block partial sums (16
// doubles for each of at most kNoallocStatBlocks blocks) and the 16
// finalized sums, 2 MiB + 128 B per device for the life of the process.
// Host code never passes `scratch` itself to a kernel: its device address
// on the current device comes from hipGetSymbolAddress (noalloc_scratch).
constexpr int kNoallocStatBlocks = 16384;
constexpr size_t kNoallocScratchDoubles = (size_t)16 * kNoallocStatBlocks + 16;
static __device__ __align__(256) double scratch[kNoallocScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
double *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (with that device current).
int noalloc_scratch(double **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail("device %d beyond the scratch table", dev);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

extern "C" int o3d_statistics_noalloc(const double *ux, const double *uy, const double *uz,
                                      int nx, int ny, int nz, double dx, double dy, double dz,
                                      double re, double t, const int *bc, double *out17) {
  CHK(ensure_init());
  CHK(check_grid(nx, ny, nz));
  CHK(check_bc(bc, true));
  if (!out17) return fail("out17 is NULL");
  if (!is_device_ptr(ux) || !is_device_ptr(uy) || !is_device_ptr(uz))
    return fail("o3d_statistics_noalloc: ux, uy and uz must be device arrays");
  const int blocks = make_tiling(nx, ny, nz).blocks();
  if (blocks > kNoallocStatBlocks)
    return fail("o3d_statistics_noalloc: grid %dx%dx%d needs %d partial records, the "
                "scratch holds %d", nx, ny, nz, blocks, kNoallocStatBlocks);
  double *sp;
  CHK(noalloc_scratch(&sp));
  const double *u[3] = {ux, uy, uz};
  // Every block writes its own record and the finalizing launch writes the
  // 16 sums, so nothing of a previous call is read.
  return statistics_run<double>(u, nx, ny, nz, dx, dy, dz, re, t, bc, out17, sp,
                                sp + (size_t)16 * kNoallocStatBlocks);
}

namespace {
int diag_dev(int mode, double *out0, double *out1, double *out2, const double *ux,
             const double *uy, const double *uz, double dx, double dy, double dz, int nx, int ny,
             int nz, const int *bc) {
  CHK(ensure_init());
  CHK(check_grid(nx, ny, nz));
  CHK(check_bc(bc));
  const size_t n = (size_t)nx * ny * nz;
  Staged s;
  DiagArgs<double> a;
  CHK(s.in(ux, n, &a.u[0]));
  CHK(s.in(uy, n, &a.u[1]));
  CHK(s.in(uz, n, &a.u[2]));
  double *o[3] = {out0, out1, out2};
  for (int c = 0; c < (mode == 0 ? 3 : 1); ++c) {
    if (!o[c]) return fail("NULL output");
    CHK(s.out(o[c], n, &a.out[c], false));
  }
  a.mode = mode;
  a.nx = nx; a.ny = ny; a.nz = nz;
  for (int q = 0; q < 3; ++q) a.bc[q] = bc[q];
  a.tl = make_tiling(nx, ny, nz);
  const double h[3] = {dx, dy, dz};
  fill_coef<double>(a.d1, h, false);
  HIPCHK(launch_diag<double>(a, g_stream));
  return s.finish();
}
}  // namespace

extern "C" int o3d_transeq(double *phi, const double *ux, const double *uy, const double *uz,
                           const double *src, double *fphi, double re, double sc,
                           const double *adt, const double *bdt, const double *cdt, int itime,
                           int itscheme, double dx, double dy, double dz, int nx, int ny, int nz,
                           int iles, const double *nu_t, const int *bc) {
  CHK(ensure_init());
  CHK(check_grid(nx, ny, nz));
  CHK(check_bc(bc));
  if (!adt || !bdt || !cdt) return fail("NULL coefficient array");
  int ci;
  if (itscheme == 1 || itime == 1) ci = 0;
  else if (itscheme == 2 || itime == 2) ci = 1;
  else if (itscheme == 3) ci = 2;
  else { fail("itscheme: %d unrecognized", itscheme); return -2; }
  const size_t n = (size_t)nx * ny * nz;
  Staged s;
  TranseqArgs<double> a;
  double *ph, *F;
  CHK(s.out(phi, n, &ph, true));
  CHK(s.out(fphi, 3 * n, &F, true));
  CHK(s.in(ux, n, &a.u[0]));
  CHK(s.in(uy, n, &a.u[1]));
  CHK(s.in(uz, n, &a.u[2]));
  CHK(s.in(src, n, &a.src));
  a.les = iles == 1;
  if (a.les) {
    if (!nu_t) return fail("nu_t is NULL with iles = 1");
    CHK(s.in(nu_t, n, &a.nu_t));
  } else {
    a.nu_t = nullptr;
  }
  a.phi = ph;
  a.phi_out = ph;
  CHK(s.scratch(n, &a.phi_new));
  a.f1 = F;
  a.f2 = F + n;
  a.f3 = F + 2 * n;
  a.nh = ci + 1;
  a.adu = adt[ci]; a.bdu = bdt[ci]; a.cdu = cdt[ci];
  a.alpha0 = 1.0 / (re * sc);  // src/integration.f90:407,410
  a.sc = sc;
  a.nx = nx; a.ny = ny; a.nz = nz;
  for (int q = 0; q < 3; ++q) a.bc[q] = bc[q];
  a.tl = make_tiling(nx, ny, nz);
  const double h[3] = {dx, dy, dz};
  fill_coef<double>(a.d1, h, false);
  fill_coef<double>(a.d2, h, true);
  CHK(s.scratch(3 * a.tl.blocks(), &a.part));
  double *sums;
  CHK(s.scratch(3, &sums));
  HIPCHK(launch_transeq<double>(a, sums, g_stream));
  // History shift by copy, as the reference (src/integration.f90:453-460).
  if (itscheme == 3) HIPCHK(launch_copy<double>(F + 2 * n, F + n, (int)n, g_stream));
  if (itscheme >= 2) HIPCHK(launch_copy<double>(F + n, F, (int)n, g_stream));
  return s.finish();
}

extern "C" int o3d_rotational(double *rotx, double *roty, double *rotz, const double *ux,
                              const double *uy, const double *uz, double dx, double dy,
                              double dz, int nx, int ny, int nz, const int *bc) {
  return diag_dev(0, rotx, roty, rotz, ux, uy, uz, dx, dy, dz, nx, ny, nz, bc);
}

extern "C" int o3d_q_criterion(double *q, const double *ux, const double *uy, const double *uz,
                               double dx, double dy, double dz, int nx, int ny, int nz,
                               const int *bc) {
  return diag_dev(1, q, nullptr, nullptr, ux, uy, uz, dx, dy, dz, nx, ny, nz, bc);
}

// ====================================================== device-resident solver
namespace {

struct Solver {
  o3d_params p;
  int n = 0;
  bool mixed = false;
  Tiling tl;      // predict (its own occupancy)
  Tiling tl_div;  // divergence (its own occupancy)
  Tiling tl_cor;  // correction
  int kind = 0;
  SorConst sc;
  double omega = 0.0;
  // Fields (element size 8 for FP64, 4 for the FP32 form; p always FP64).
  void *u[3] = {}, *up[3] = {}, *F[3][3] = {}, *rhs = nullptr, *nut = nullptr;
  double *P[2] = {nullptr, nullptr};
  int cur = 0;                 // P[cur] holds the pressure
  float *E[2] = {nullptr, nullptr}, *R = nullptr;  // mixed: correction ping-pong, residual
  unsigned long long *dslot = nullptr;  // colour-split sweeps: dmax slots
  bool r_ready = false;  // R, E[0] hold the split residual of P[cur] and e = 0
  double *rhs_split = nullptr;          // FP64 form: the rhs in colour-split storage
  int ring = 0;                // predict calls so far (history ring position)
  // Reductions.
  FieldPartial *part_div = nullptr, *part_div2 = nullptr, *part_nut = nullptr,
               *part_cor = nullptr, *fin = nullptr;
  FieldPartial *hfin = nullptr;  // pinned, 6 records
  SorState *st = nullptr, *st64 = nullptr, *hst = nullptr;
  std::vector<hipEvent_t> ev;
  int last_sweeps[2] = {0, 0};
  int last_first = 0;          // sweeps of the previous step's first FP32 correction solve
  double sor_ms[2] = {0, 0};
  int sor_sweeps[2] = {0, 0};
  int sor_launches[2] = {0, 0};
  void *nn = nullptr;          // learned warm start
  float *cap_r = nullptr, *cap_eprev = nullptr, *cap_e = nullptr;
  // Learned warm start / capture state.
  double *Pstart = nullptr;    // pressure at the start of the solve (p_prev)
  float *Eprev = nullptr;      // previous step's increment
  float *Eprev2 = nullptr;     // the one before
  float *Rl = nullptr;         // residual of p_prev (float32)
  float *Feat = nullptr;       // network input, 2 coarse planes
  float *Out = nullptr;        // fine-grid scratch
  float *Outc = nullptr;       // network output (coarse)
  MgHierarchy mg;              // coarse-grid correction (coarse_cycles > 0)
  float *mgbuf = nullptr;
  double *npart = nullptr;     // reduction partials (residual norms, sums of squares)
  double *p0part = nullptr;    // per-block norms of the residual of p0_of (guess pass)
  const double *p0_of = nullptr;  // the pressure those norms belong to (null: none)
  double *hnpart = nullptr;    // pinned copy
  double *epart = nullptr;     // sum-of-squares partials of e_prev (one per correction block)
  double *hepart = nullptr;    // pinned copy, valid after the step's final sync
  int ne = 0;                  // entries of epart / hepart
  double *mpart = nullptr;     // mean-pressure partials + the mean (mixed, every axis periodic)
  double *jpart = nullptr;     // per-tile sums of p from the FP64 judge sweep (with mpart)
  int jblocks = 0;             // tiles of the judge sweep
  // Adaptive extrapolation (o3d_params.adaptive_extrapolation): the weight of
  // e_prev - e_prev2 and the device sums behind the next one.
  double beta = 0.0;
  double *bpart = nullptr, *bsum = nullptr, *hbsum = nullptr;
  double beta_ls = 0.0;        // the unclamped least-squares weight beta came from
  bool beta_pending = false;
  bool adapt_now = false;      // this step's guess is the adaptive extrapolation
  bool mean_pending = false;   // the mean is formed, its subtraction not yet done
  // The rhs loses its weighted mean before the solve (x-Dirichlet, or the
  // mixed form on a grid with a mirrored axis unless rhs_as_given).
  bool project = false;
  // x-Dirichlet inflow / outflow (bc[0] == DIRICHLET), set by o3d_solver_set_inflow.
  bool xdir = false, in_set = false;
  double *in_base = nullptr;   // 3 planes of ny nz values: the inflow profile
  double *in_w = nullptr;      // plane: noise weight
  double *in_pl = nullptr;     // 2 planes (inflow u, outflow u before the shift) + 2 nz rows
  double *in_flux = nullptr;   // Q_in, Q_out, shift
  double *hflux = nullptr;     // pinned copy, valid after the step's final sync
  double *wpart = nullptr;     // weighted-sum partials (1024) + the sum of the weights + mean
  double in_amp = 0.0, in_period = 1.0, in_cu = 0.0, in_W = 0.0;
  unsigned long long in_seed = 0;
  int hist = 0;                // steps with an increment available
  std::vector<void *> allocs;

  size_t esz() const { return mixed ? 4 : 8; }
  template <typename T>
  int alloc(T **p_, size_t count) {
    void *d = nullptr;
    HIPCHK(hipMalloc(&d, count * sizeof(T)));
    HIPCHK(hipMemset(d, 0, count * sizeof(T)));
    allocs.push_back(d);
    *p_ = (T *)d;
    return 0;
  }
  int allocv(void **p_, size_t bytes) {
    void *d = nullptr;
    HIPCHK(hipMalloc(&d, bytes));
    HIPCHK(hipMemset(d, 0, bytes));
    allocs.push_back(d);
    *p_ = d;
    return 0;
  }
  ~Solver() {
    for (void *a : allocs) (void)hipFree(a);
    for (auto e : ev) (void)hipEventDestroy(e);
    if (hfin) (void)hipHostFree(hfin);
    if (hst) (void)hipHostFree(hst);
    if (hnpart) (void)hipHostFree(hnpart);
    if (hepart) (void)hipHostFree(hepart);
    if (hflux) (void)hipHostFree(hflux);
    if (hbsum) (void)hipHostFree(hbsum);
  }
};

// The inflow and outflow planes of u* after predict, with the outflow flux
// set equal to the inflow flux (x-Dirichlet).
template <typename T>
int solver_jet_edges(Solver &S, int itime) {
  const o3d_params &p = S.p;
  if (!S.in_set) return fail("x-Dirichlet: no inflow set (o3d_solver_set_inflow)");
  const int P = p.ny * p.nz;
  const double ph = 2.0 * M_PI * ((double)itime * p.dt) / S.in_period;
  const double cph = cos(ph), sph = sin(ph);
  double *pin = S.in_pl, *pout = S.in_pl + P, *rows = S.in_pl + 2 * (size_t)P;
  const int g = (P + 255) / 256;
  jet_edges_kernel<T><<<g, 256, 0, g_stream>>>(
      (T *)S.up[0], (T *)S.up[1], (T *)S.up[2], (const T *)S.u[0], (const T *)S.u[1],
      (const T *)S.u[2], S.in_base, S.in_w, S.in_amp, cph, sph, S.in_seed, S.in_cu, p.nx, p.ny,
      p.nz, pin, pout);
  HIPCHK(hipGetLastError());
  jet_flux_kernel<<<1, 256, 0, g_stream>>>(pin, pout, rows, p.ny, p.nz, p.bc[1] != PERIODIC,
                                           p.bc[2] != PERIODIC, S.in_W, S.in_flux);
  HIPCHK(hipGetLastError());
  jet_shift_kernel<T><<<g, 256, 0, g_stream>>>((T *)S.up[0], pout, S.in_flux, p.nx, p.ny, p.nz);
  HIPCHK(hipGetLastError());
  return 0;
}

// The compatibility projection of the Poisson rhs on grids with a mirrored
// axis (x-Dirichlet; the mixed form with free-slip axes): minus its weighted
// mean (trapezoid weights on the mirrored axes, whose weighted column sums
// of the mirrored operator vanish), the part of the rhs the operator cannot
// reach, so SOR converges to eps.
template <typename T>
int solver_rhs_project(Solver &S) {
  const o3d_params &p = S.p;
  const int mx = p.bc[0] != PERIODIC, my = p.bc[1] != PERIODIC, mz = p.bc[2] != PERIODIC;
  wsum_kernel<T><<<kWsumBlocks, 256, 0, g_stream>>>((const T *)S.rhs, p.nx, p.ny, p.nz, mx, my,
                                                     mz, S.wpart);
  HIPCHK(hipGetLastError());
  wsum_final_kernel<<<1, 1024, 0, g_stream>>>(S.wpart, S.wpart + kWsumBlocks,
                                              S.wpart + kWsumBlocks + 1);
  HIPCHK(hipGetLastError());
  sub_mean_kernel<T><<<2048, 256, 0, g_stream>>>((T *)S.rhs, (long long)S.n,
                                                  S.wpart + kWsumBlocks + 1);
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename T>
int solver_predict(Solver &S, int itime) {
  const o3d_params &p = S.p;
  int ci;
  if (p.itscheme == 1 || itime == 1) ci = 0;
  else if (p.itscheme == 2 || itime == 2) ci = 1;
  else if (p.itscheme == 3) ci = 2;
  else { fail("itscheme: %d unrecognized", p.itscheme); return -2; }
  // History ring: F1 of this step goes to slot ring % 3; slots of the two
  // previous steps hold F2 and F3 (the reference shifts by copying,
  // src/integration.f90:176-188; unused slots start at zero).
  const int s1 = S.ring % 3, s2 = (S.ring + 2) % 3, s3 = (S.ring + 1) % 3;
  PredictArgs<T> a;
  for (int c = 0; c < 3; ++c) {
    a.u[c] = (const T *)S.u[c];
    a.up[c] = (T *)S.up[c];
    a.f1[c] = (T *)S.F[s1][c];
    a.f2[c] = (const T *)S.F[s2][c];
    a.f3[c] = (const T *)S.F[s3][c];
  }
  const bool les = p.iles == 1;
  a.nut = les ? (T *)S.nut : nullptr;
  a.nut_part = S.part_nut;
  a.nx = p.nx; a.ny = p.ny; a.nz = p.nz;
  for (int q = 0; q < 3; ++q) a.bc[q] = p.bc[q];
  a.tl = S.tl;
  const double h[3] = {p.dx, p.dy, p.dz};
  fill_coef<T>(a.d1, h, false);
  fill_coef<T>(a.d2, h, true);
  a.adu = T(p.adt[ci]); a.bdu = T(p.bdt[ci]); a.cdu = T(p.cdt[ci]);
  a.onere = T(1.0 / p.re);
  const double csd = p.cs * p.delta;
  a.csd2 = T(csd * csd);
  HIPCHK(launch_predict<T>(a, ci + 1, les, g_stream));
  if (p.itscheme >= 2) S.ring++;
  return 0;
}

template <typename T>
int solver_divergence(Solver &S, const void *const f[3], T *out, double divisor,
                      FieldPartial *part) {
  DivArgs<T> a;
  for (int c = 0; c < 3; ++c) a.f[c] = (const T *)f[c];
  a.out = out;
  a.part = part;
  a.nx = S.p.nx; a.ny = S.p.ny; a.nz = S.p.nz;
  for (int q = 0; q < 3; ++q) a.bc[q] = S.p.bc[q];
  a.odd = 1;
  a.tl = S.tl_div;
  const double h[3] = {S.p.dx, S.p.dy, S.p.dz};
  fill_coef<T>(a.d1, h, false);
  a.divisor = T(divisor);
  HIPCHK(launch_divergence<T>(a, g_stream));
  return 0;
}

// Partials of the FP64 pressure sum (p_sum_kernel); the mean follows them.
constexpr int kMeanBlocks = 1024;

// The velocity correction; with a pending mean, the pressure's shift
// (p - mean into the spare pressure buffer, which becomes the current one);
// with incr, the increment e_prev = p - p_prev and its sum-of-squares
// partials, all in the one pass over p.
template <typename T>
int solver_correct(Solver &S, bool incr = false) {
  CorrectArgs<T, double> a;
  for (int c = 0; c < 3; ++c) { a.u[c] = (T *)S.u[c]; a.up[c] = (const T *)S.up[c]; }
  a.p = S.P[S.cur];
  a.part = S.part_cor;
  a.nx = S.p.nx; a.ny = S.p.ny; a.nz = S.p.nz;
  for (int q = 0; q < 3; ++q) a.bc[q] = S.p.bc[q];
  a.tl = S.tl_cor;
  const double h[3] = {S.p.dx, S.p.dy, S.p.dz};
  fill_coef<double>(a.d1, h, false);
  a.dt = S.p.dt;
  if (S.mean_pending) {
    a.mean = S.mpart + kMeanBlocks;
    a.pout = S.P[1 - S.cur];
  }
  if (incr) {
    a.p0 = S.Pstart;
    a.e = S.Eprev;
    a.epart = S.epart;
  }
  HIPCHK(launch_correct<T, double>(a, g_stream));
  if (a.pout) S.cur = 1 - S.cur;
  S.mean_pending = false;
  return 0;
}

template <typename T, typename TR>
void sor_args_common(Solver &S, SorArgs<T, TR> &a, const TR *rhs, SorState *st, int kmax) {
  a.rhs = rhs;
  a.st = st;
  a.log = nullptr;
  a.log_len = 0;
  a.nx = S.p.nx; a.ny = S.p.ny; a.nz = S.p.nz;
  solver_periodic(S.kind, a.per);
  a.tl = sor_tiling<T, TR>(S.p.nx, S.p.ny, S.p.nz);
  a.ox = S.sc.ox; a.oy = S.sc.oy; a.oz = S.sc.oz; a.A = S.sc.A;
  a.eps = S.p.eps;
  a.factor = S.kind == SOR_0011 ? 1.01 : 1.05;
  a.kmax = kmax;
  a.idyn = S.p.idyn;
  a.it = 0;
  a.gs = 0;
}

// Plain FP64 Poisson solve on P[cur] (src/poisson.f90 semantics, red-black).
int solver_poisson_f64(Solver &S, o3d_step_info *info) {
  SorRun<double, double> r;
  r.st = S.st;
  r.hst = S.hst;
  r.events = &S.ev;
  r.buf[0] = S.P[S.cur];
  r.buf[1] = S.P[1 - S.cur];
  sor_args_common<double, double>(S, r.a, (const double *)S.rhs, S.st, S.p.kmax);
  int sweeps = 0;
  const int first = S.last_sweeps[0] > 0 ? std::max(4, S.last_sweeps[0] - 2) : 16;
  if (S.rhs_split) {
    // Colour-split sweeps (identical iterates): the iterate in split storage
    // in the spare pressure buffer, back into P[cur] after the solve.
    const int n = S.n, h = n / 2;
    const int nx = S.p.nx, ny = S.p.ny, nz = S.p.nz;
    double *q = S.P[1 - S.cur];
    HIPCHK(launch_split<double>(S.P[S.cur], q, q + h, nx, ny, nz, g_stream));
    HIPCHK(launch_split<double>((const double *)S.rhs, S.rhs_split, S.rhs_split + h, nx, ny, nz,
                                g_stream));
    r.split = true;
    r.sa = split_args<double>(r.a, q, q + h, S.rhs_split, S.rhs_split + h, S.dslot);
    CHK(r.run(S.omega, first, &sweeps));
    HIPCHK(launch_unsplit<double>(S.P[S.cur], q, q + h, nx, ny, nz, g_stream));
  } else {
    CHK(r.run(S.omega, first, &sweeps));
    if (sweeps & 1) S.cur = 1 - S.cur;
  }
  S.last_sweeps[0] = sweeps;
  S.omega = S.hst->omega_final;
  S.sor_ms[0] += r.ms_full;
  S.sor_sweeps[0] += r.sweeps_full;
  S.sor_launches[0] += r.launches_full;
  info->sor_exit_iter = S.hst->exit_iter;
  info->sor_reason = S.hst->reason;
  info->sweeps_f64 = sweeps;
  info->sweeps_f32 = 0;
  info->refinements = 0;
  info->sor_dmax = S.hst->dmax_final;
  info->omega = S.omega;
  return 0;
}

// Mixed precision: FP32 red-black sweeps on the correction e of A e = r,
// r = rhs - A p formed in FP64; p += e in FP64; convergence judged by an FP64
// red-black sweep of p with the reference's stop rules.
// Periodic mean-pressure removal (mixed forms, every axis periodic: the
// pressure is defined up to a constant, which the FP32 sweeps let drift).
// The FP64 sum is formed in a fixed order: per block over a contiguous
// chunk (2 doubles per thread and load), then the partials in index order.
__global__ void __launch_bounds__(256) p_sum_kernel(const double *__restrict__ p, int n,
                                                    double *__restrict__ part) {
  __shared__ double sv[4];
  const int n2 = n / 2;
  const int per = (n2 + gridDim.x - 1) / gridDim.x;
  const int lo = blockIdx.x * per, hi = min(n2, lo + per);
  const double2 *p2 = (const double2 *)p;
  double s = 0.0;
  for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const double2 v = p2[i];
    s += v.x;
    s += v.y;
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0 && (n & 1)) s += p[n - 1];
  s = block_sum<4>(s, sv);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// The mean from the partials (1024 threads, one partial each, the same
// order every time): mp[0] = sum / n.
__global__ void __launch_bounds__(1024) p_mean_kernel(const double *__restrict__ part,
                                                      int nparts, int n, double *__restrict__ mp) {
  __shared__ double sv[16];
  double s = 0.0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) s += part[b];
  s = block_sum<16>(s, sv);
  if (threadIdx.x == 0) mp[0] = s / (double)n;
}

// The mean of the pressure into part[kMeanBlocks] (the velocity correction
// pass applies the shift).
int pressure_mean(double *p, int n, double *part) {
  p_sum_kernel<<<kMeanBlocks, 256, 0, g_stream>>>(p, n, part);
  HIPCHK(hipGetLastError());
  p_mean_kernel<<<1, 1024, 0, g_stream>>>(part, kMeanBlocks, n, part + kMeanBlocks);
  HIPCHK(hipGetLastError());
  return 0;
}
// The same mean from the judge sweep's per-tile sums of the p it wrote.
int pressure_mean_from(const double *tiles, int ntiles, int n, double *mp) {
  p_mean_kernel<<<1, 1024, 0, g_stream>>>(tiles, ntiles, n, mp);
  HIPCHK(hipGetLastError());
  return 0;
}

int solver_poisson_mixed(Solver &S, o3d_step_info *info) {
  const o3d_params &p = S.p;
  const int n = S.n;
  // With the coarse-grid correction, every correction solve is one
  // two-grid cycle: the coarse correction of its residual, the Gauss-Seidel
  // opening and a few SOR sweeps; a solve that has not converged by then
  // ends, and the next refinement cycle starts from the new residual.
  const bool mgc = S.mgbuf != nullptr && S.dslot != nullptr;
  const int inner_max = p.inner_max > 0 ? p.inner_max
                                        : (mgc ? p.smooth_sweeps + kMgInnerSor : 400);
  int total = 0, sw32 = 0, sw64 = 0, cycles = 0;
  int reason = 2;
  double judge_prev = 1609.0, dlast = 0.0;
  int capped = 0;  // consecutive correction solves that ended on their sweep cap
  // The judge sweep summed the p it wrote and nothing has changed p since.
  bool sum_ok = false;
  while (total < p.kmax) {
    sum_ok = false;
    int per[3];
    solver_periodic(S.kind, per);
    // Colour-split correction sweeps where the grid allows (nx a multiple
    // of 8), else one fused iteration per launch; both give the same
    // iterates.
    const bool split = S.dslot != nullptr;
    if (split && cycles == 0 && S.r_ready) {
      // r and e = 0 already written by the warm-start acceptance check.
    } else if (split) {
      HIPCHK(launch_residual_split<float>(S.P[S.cur], (const float *)S.rhs, S.R, S.E[0], p.nx,
                                          p.ny, p.nz, per, S.sc.ox, S.sc.oy, S.sc.oz, S.sc.A,
                                          g_stream));
      if (mgc) {
        // The start of this correction solve: the coarse-grid correction of
        // its residual (cycle 0 starts from a guess that already had one).
        HIPCHK(mg_pool_split(S.R, S.mg.f[0], p.nx, p.ny, p.nz, g_stream));
        HIPCHK(mg_solve(S.mg, p.refine_cycles > 0 ? p.refine_cycles : p.coarse_cycles, 1, 1,
                        g_stream));
        HIPCHK(mg_prolong_split(S.E[0], S.mg.u[0], p.nx, p.ny, p.nz, per, g_stream));
      }
    }
    else
      HIPCHK(launch_residual_f32<float>(S.P[S.cur], (const float *)S.rhs, S.R, S.E[0], p.nx,
                                        p.ny, p.nz, per, S.sc.ox, S.sc.oy, S.sc.oz, S.sc.A,
                                        g_stream));
    SorRun<float, float> r;
    r.st = S.st;
    r.hst = S.hst;
    r.events = &S.ev;
    r.mid = S.ev[6];
    r.buf[0] = S.E[0];
    r.buf[1] = S.E[1];
    const int kin = std::min(inner_max, p.kmax - total);
    sor_args_common<float, float>(S, r.a, S.R, S.st, kin);
    // The first correction solve of the step opens with Gauss-Seidel sweeps
    // (smoothing the high-frequency error a warm start leaves), and so does
    // every one that starts from a coarse-grid correction.
    if (cycles == 0 || mgc) r.a.gs = p.smooth_sweeps;
    if (split) {
      const int h = n / 2;
      r.split = true;
      r.sa = split_args<float>(r.a, S.E[0], S.E[0] + h, S.R, S.R + h, S.dslot);
    }
    int sweeps = 0;
    // The first solve of a step usually takes as many sweeps as the
    // previous step's: launch that many plus the launch that sees the stop,
    // so the solve ends without a second batch or unused launches.
    const int first = cycles == 0 && S.last_first > 0 ? std::min(kin + 1, S.last_first + 1)
                                                      : std::min(kin, 16);
    CHK(r.run(S.omega, first, &sweeps));
    S.sor_ms[1] += r.ms_full;
    S.sor_sweeps[1] += r.sweeps_full;
    S.sor_launches[1] += r.launches_full;
    const int rin = S.hst->reason;
    S.omega = S.hst->omega_final;
    total += sweeps;
    sw32 += sweeps;
    if (cycles == 0) S.last_first = sweeps;
    ++cycles;
    // A solve that ended on its sweep cap continues from a refreshed
    // residual; every fourth such solve in a row is followed by the judge
    // sweep anyway, so a problem that cannot converge (an rhs incompatible
    // with the free-slip mirror) still reaches the stagnation rule.
    const bool to_judge = !(rin == 2 && total < p.kmax) || capped + 1 >= 4;
    // p += e: with colour-split sweeps and a judge sweep next, folded into
    // the judge's loads (it reads p + e and writes the other buffer) instead
    // of a pass of its own; the sums are the same, so are the iterates.
    const bool fuse = split && to_judge;
    if (!fuse) {
      if (split) HIPCHK(launch_fold_split(S.P[S.cur], S.E[0], p.nx, p.ny, p.nz, g_stream));
      else HIPCHK(launch_fold(S.P[S.cur], r.result(sweeps), n, g_stream));
    }
    if (!to_judge) {
      ++capped;
      continue;
    }
    capped = 0;
    // FP64 judge sweep (applied, like the reference's last sweep).
    SorRun<double, float> j;
    j.st = S.st64;
    j.hst = S.hst;
    j.events = &S.ev;
    j.buf[0] = S.P[S.cur];
    j.buf[1] = S.P[1 - S.cur];
    sor_args_common<double, float>(S, j.a, (const float *)S.rhs, S.st64, 1);
    if (fuse) j.a.efold = S.E[0];
    j.a.psum = S.jpart;
    int jsw = 0;
    CHK(j.run(S.omega, 2, &jsw));
    if (jsw & 1) S.cur = 1 - S.cur;
    sum_ok = S.jpart != nullptr && jsw == 1;
    sw64 += jsw;
    total += jsw;
    dlast = S.hst->dmax_final;
    if (dlast < p.eps) { reason = 0; break; }
    if (fabs(judge_prev - dlast) < p.eps / 1000.0) { reason = 1; break; }
    judge_prev = dlast;
  }
  S.last_sweeps[1] = sw32;
  S.r_ready = false;
  if (S.mpart) {
    if (sum_ok) CHK(pressure_mean_from(S.jpart, S.jblocks, n, S.mpart + kMeanBlocks));
    else CHK(pressure_mean(S.P[S.cur], n, S.mpart));
    S.mean_pending = true;
  }
  info->sor_exit_iter = reason == 2 ? p.kmax + 1 : total;
  info->sor_reason = reason;
  info->sweeps_f64 = sw64;
  info->sweeps_f32 = sw32;
  info->refinements = cycles;
  info->sor_dmax = dlast;
  info->omega = S.omega;
  return 0;
}

// ---- learned warm start and training capture ------------------------------
// The network works on a grid O3D_WARM_COARSEN times coarser per axis: its
// inputs are block averages of the fine planes, and its correction to the
// extrapolation is interpolated back trilinearly, so the correction is smooth.
// (At the near-optimal omega every SOR error mode decays at about the same
// rate, and the dmax stop rule weighs high-frequency error most, so a guess
// carrying BF16-level noise at the fine scale would cost sweeps.)
__global__ void lw_pool_kernel(const float *__restrict__ in, float *__restrict__ out, int nx,
                               int ny, int nz, int F, float scale) {
  const int mx = nx / F, my = ny / F, mz = nz / F;
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= mx * my * mz) return;
  const int i0 = (I % mx) * F, j0 = ((I / mx) % my) * F, k0 = (I / (mx * my)) * F;
  float s = 0.f;
  for (int k = 0; k < F; ++k)
    for (int j = 0; j < F; ++j) {
      const float *row = in + (size_t)i0 + (size_t)nx * ((j0 + j) + (size_t)ny * (k0 + k));
      for (int i = 0; i < F; ++i) s += row[i];
    }
  out[I] = s * (scale / (float)(F * F * F));
}

// The same for F = 2 and 4: unrolled, each row segment one 8- or 16-byte
// load (i0 is a multiple of F), summed in the same order (bit-identical).
template <int F>
__global__ void __launch_bounds__(256) lw_pool_f_kernel(const float *__restrict__ in,
                                                        float *__restrict__ out, int nx, int ny,
                                                        int nz, float scale) {
  const int mx = nx / F, my = ny / F, mz = nz / F;
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= mx * my * mz) return;
  const int i0 = (I % mx) * F, j0 = ((I / mx) % my) * F, k0 = (I / (mx * my)) * F;
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < F; ++k)
#pragma unroll
    for (int j = 0; j < F; ++j) {
      const float *row = in + (size_t)i0 + (size_t)nx * ((j0 + j) + (size_t)ny * (k0 + k));
      if constexpr (F == 4) {
        const float4 v = *reinterpret_cast<const float4 *>(row);
        s += v.x;
        s += v.y;
        s += v.z;
        s += v.w;
      } else {
        const float2 v = *reinterpret_cast<const float2 *>(row);
        s += v.x;
        s += v.y;
      }
    }
  out[I] = s * (scale / (float)(F * F * F));
}

// out = the F^3 block averages of in, times scale.
hipError_t launch_lw_pool(const float *in, float *out, int nx, int ny, int nz, int F,
                          float scale) {
  const int nc = (nx / F) * (ny / F) * (nz / F);
  const int blocks = (nc + 255) / 256;
  const uintptr_t a = reinterpret_cast<uintptr_t>(in);
  if (F == 4 && nx % 4 == 0 && a % 16 == 0)
    lw_pool_f_kernel<4><<<blocks, 256, 0, g_stream>>>(in, out, nx, ny, nz, scale);
  else if (F == 2 && nx % 2 == 0 && a % 8 == 0)
    lw_pool_f_kernel<2><<<blocks, 256, 0, g_stream>>>(in, out, nx, ny, nz, scale);
  else
    lw_pool_kernel<<<blocks, 256, 0, g_stream>>>(in, out, nx, ny, nz, F, scale);
  return hipGetLastError();
}

// Position of fine node i on the coarse grid of cell averages: lower cell and
// weight of the upper one; periodic axes wrap, others clamp.
__device__ __forceinline__ void lw_axis(int i, int F, int m, int per, int &c0, int &c1, float &w) {
  const float X = ((float)i - 0.5f * (float)(F - 1)) / (float)F;
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

// q = p + e_prev + s * trilinear(out - in1), out and in1 on the coarse grid.
__global__ void lw_guess_coarse_kernel(const double *__restrict__ p, const float *__restrict__ ep,
                                       const float *__restrict__ out,
                                       const float *__restrict__ in1, double *__restrict__ q,
                                       int nx, int ny, int nz, int F, int px, int py, int pz,
                                       double s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= nx) return;
  const int mx = nx / F, my = ny / F, mz = nz / F;
  int x0, x1, y0, y1, z0, z1;
  float wx, wy, wz;
  lw_axis(i, F, mx, px, x0, x1, wx);
  lw_axis(j, F, my, py, y0, y1, wy);
  lw_axis(k, F, mz, pz, z0, z1, wz);
  auto c = [&](int a, int b, int d) {
    const size_t o = (size_t)a + (size_t)mx * (b + (size_t)my * d);
    return out[o] - in1[o];
  };
  const float v =
      (1.f - wz) * ((1.f - wy) * ((1.f - wx) * c(x0, y0, z0) + wx * c(x1, y0, z0)) +
                    wy * ((1.f - wx) * c(x0, y1, z0) + wx * c(x1, y1, z0))) +
      wz * ((1.f - wy) * ((1.f - wx) * c(x0, y0, z1) + wx * c(x1, y0, z1)) +
            wy * ((1.f - wx) * c(x0, y1, z1) + wx * c(x1, y1, z1)));
  const size_t o = (size_t)i + (size_t)nx * (j + (size_t)ny * k);
  q[o] = p[o] + (double)ep[o] + s * (double)v;
}

// u0 = s * (out - in1) on the coarse grid: the network's estimate of the
// coarse correction of the extrapolation, the multigrid's starting iterate.
// u0 = s (out - in1): the network's correction of the guess whose residual
// it was given (out estimates e_prev + A^-1 r on the block grid, in1 is the
// block average of e_prev over s).
__global__ void lw_coarse_init_kernel(float *__restrict__ u0, const float *__restrict__ out,
                                      const float *__restrict__ in1, int nc, float s) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I < nc) u0[I] = s * (out[I] - in1[I]);
}

// q += trilinear interpolation of the coarse correction c (the multigrid
// solution on the grid of F^3 block averages), one thread per pair of cells
// (2t, 2t + 1) of a row (nx even), 16-byte accesses.
__global__ void __launch_bounds__(256) mg_prolong_fine_kernel(double *__restrict__ q,
                                                              const float *__restrict__ c,
                                                              int nx, int ny, int nz, int F,
                                                              int px, int py, int pz) {
  const int t = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (2 * t >= nx || j >= ny) return;
  const int mx = nx / F, my = ny / F, mz = nz / F;
  int y0, y1, z0, z1;
  float wy, wz;
  lw_axis(j, F, my, py, y0, y1, wy);
  lw_axis(k, F, mz, pz, z0, z1, wz);
  auto v = [&](int a, int b, int d) { return c[(size_t)a + (size_t)mx * (b + (size_t)my * d)]; };
  float tv[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    int x0, x1;
    float wx;
    lw_axis(2 * t + h, F, mx, px, x0, x1, wx);
    tv[h] = (1.f - wz) * ((1.f - wy) * ((1.f - wx) * v(x0, y0, z0) + wx * v(x1, y0, z0)) +
                          wy * ((1.f - wx) * v(x0, y1, z0) + wx * v(x1, y1, z0))) +
            wz * ((1.f - wy) * ((1.f - wx) * v(x0, y0, z1) + wx * v(x1, y0, z1)) +
                  wy * ((1.f - wx) * v(x0, y1, z1) + wx * v(x1, y1, z1)));
  }
  double2 *o = reinterpret_cast<double2 *>(q + (size_t)nx * (j + (size_t)ny * k)) + t;
  double2 w = *o;
  w.x = w.x + (double)tv[0];
  w.y = w.y + (double)tv[1];
  *o = w;
}

__global__ void __launch_bounds__(256) lw_sumsq_kernel(const float *a, int n, double *part) {
  __shared__ double sc[4];
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    s += (double)a[i] * a[i];
  s = block_sum<4>(s, sc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

constexpr int kLwBlocks = 512;


bool mg_level1_is(const MgHierarchy &H, int mx, int my, int mz) {
  return H.levels >= 2 && H.m[1][0] == mx && H.m[1][1] == my && H.m[1][2] == mz;
}

// Partial-sum buffers of the warm-start reductions.
struct LwSums {
  double *part = nullptr;   // device, 2 * kLwBlocks
  double *hpart = nullptr;  // pinned host copy
};

int lw_sums_alloc(LwSums &w) {
  if (w.part) return 0;
  HIPCHK(hipMalloc((void **)&w.part, 2 * kLwBlocks * sizeof(double)));
  HIPCHK(hipHostMalloc((void **)&w.hpart, 2 * kLwBlocks * sizeof(double)));
  return 0;
}

int lw_sumsq(LwSums &w, const float *a, int n, double *out) {
  lw_sumsq_kernel<<<kLwBlocks, 256, 0, g_stream>>>(a, n, w.part);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(w.hpart, w.part, kLwBlocks * sizeof(double), hipMemcpyDeviceToHost,
                        g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  double s = 0.0;
  for (int b = 0; b < kLwBlocks; ++b) s += w.hpart[b];
  *out = s;
  return 0;
}

// Network input (2 coarse planes: r dx^2 F^2 / RMS(e_prev) and
// e_prev / RMS(e_prev), F the coarsening factor, so that the first plane is the
// coarse grid's dimensionless residual in the target's units) and, when e is
// given, the target (the coarse e / RMS(e_prev)).  Returns 1 (nothing
// written) when RMS(e_prev) vanishes.
int lw_features(LwSums &w, const float *r, const float *ep, const float *e, int nx, int ny,
                int nz, double h2, float *inp, float *tgt, double *scale,
                double se2_known = -1.0) {
  const int n = nx * ny * nz;
  const int F = O3D_WARM_COARSEN;
  const int nc = n / (F * F * F);
  double se2 = se2_known;
  if (se2 < 0.0) CHK(lw_sumsq(w, ep, n, &se2));
  const double re = sqrt(se2 / n);
  *scale = re;
  if (!(re > 0.0)) return 1;
  HIPCHK(launch_lw_pool(r, inp, nx, ny, nz, F, (float)(h2 * F * F / re)));
  HIPCHK(launch_lw_pool(ep, inp + nc, nx, ny, nz, F, (float)(1.0 / re)));
  if (e && tgt) HIPCHK(launch_lw_pool(e, tgt, nx, ny, nz, F, (float)(1.0 / re)));
  return 0;
}

int lw_alloc(Solver &S) {
  if (S.Pstart) return 0;
  const size_t n = S.n;
  const size_t nc = n / (O3D_WARM_COARSEN * O3D_WARM_COARSEN * O3D_WARM_COARSEN);
  CHK(S.alloc(&S.Pstart, n));
  CHK(S.alloc(&S.Eprev, n));
  CHK(S.alloc(&S.Eprev2, n));
  CHK(S.alloc(&S.Rl, n));
  CHK(S.alloc(&S.Out, n));
  CHK(S.alloc(&S.Feat, 2 * nc));
  CHK(S.alloc(&S.Outc, nc));
  if (S.p.coarse_cycles > 0) {
    const int F = kMgCoarsen;
    const int m0[3] = {S.p.nx / F, S.p.ny / F, S.p.nz / F};
    const double h[3] = {S.p.dx * F, S.p.dy * F, S.p.dz * F};
    int per[3];
    solver_periodic(S.kind, per);
    const size_t floats = mg_plan(S.mg, m0, h, per);
    CHK(S.alloc(&S.mgbuf, floats));
    mg_bind(S.mg, S.mgbuf);
  }
  // Shared by the residual-norm pair (4 per block + 4) and lw_sumsq
  // (kLwBlocks partials, copied whole to the host).
  const int np = std::max(4 * residual_blocks(S.p.nx, S.p.ny, S.p.nz) + 4, 2 * kLwBlocks);
  CHK(S.alloc(&S.npart, np));
  CHK(S.alloc(&S.p0part, 2 * residual_blocks(S.p.nx, S.p.ny, S.p.nz)));
  HIPCHK(hipHostMalloc((void **)&S.hnpart, 2 * kLwBlocks * sizeof(double)));
  S.ne = S.tl_cor.blocks();
  CHK(S.alloc(&S.epart, (size_t)S.ne));
  HIPCHK(hipHostMalloc((void **)&S.hepart, (size_t)S.ne * sizeof(double)));
  return 0;
}

LwSums lw_of(Solver &S) {
  LwSums w;
  w.part = S.npart;
  w.hpart = S.hnpart;
  return w;
}

// L2 norms of the FP64 residuals rhs - L p of two fields, one host sync.
int lw_residual_l2_pair(Solver &S, const double *p0, const double *p1, double *out) {
  int per[3];
  solver_periodic(S.kind, per);
  const int nb = residual_blocks(S.p.nx, S.p.ny, S.p.nz);
  double *fin = S.npart + 4 * nb;
  // In the mixed form with colour-split sweeps, the residual of p1 (the
  // guess) is kept for the first correction solve (S.r_ready).
  S.r_ready = S.mixed && S.dslot != nullptr;
  // The norms of p0's residual, when the guess pass of this step formed them.
  const bool have0 = S.p0_of != nullptr && S.p0_of == p0;
  S.p0_of = nullptr;
  if (S.mixed)
    HIPCHK(launch_residual_norm_pair<float>(have0 ? nullptr : p0, p1, (const float *)S.rhs,
                                            S.npart, fin, S.p.nx, S.p.ny, S.p.nz, per, S.sc.ox,
                                            S.sc.oy, S.sc.oz, S.sc.A, g_stream,
                                            S.r_ready ? S.R : nullptr,
                                            S.r_ready ? S.E[0] : nullptr, S.p0part));
  else
    HIPCHK(launch_residual_norm_pair<double>(have0 ? nullptr : p0, p1, (const double *)S.rhs,
                                             S.npart, fin, S.p.nx, S.p.ny, S.p.nz, per, S.sc.ox,
                                             S.sc.oy, S.sc.oz, S.sc.A, g_stream, nullptr,
                                             nullptr, S.p0part));
  HIPCHK(hipMemcpyAsync(S.hnpart, fin, 4 * sizeof(double), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  out[0] = sqrt(S.hnpart[1]);
  out[1] = sqrt(S.hnpart[3]);
  return 0;
}

// Before the solve: keep p_prev; with an increment available, form the
// extrapolated pressure p_prev + e_prev (in the spare pressure buffer) and its
// residual r = rhs - L (p_prev + e_prev) in float32; on the first step r is
// the residual of p_prev.
int solver_learning_prepare(Solver &S) {
  CHK(lw_alloc(S));
  S.p0_of = nullptr;
  const int n = S.n;
  int per[3];
  solver_periodic(S.kind, per);
  // With a network, the adaptive guess is used only under the coarse-grid
  // correction, whose start takes the network's estimate relative to it.
  S.adapt_now = S.hist >= 2 && S.p.adaptive_extrapolation &&
                (!S.nn || S.p.coarse_cycles > 0);
  if (S.adapt_now) {
    // p_prev + e_prev + beta (e_prev - e_prev2), beta the least-squares
    // weight of the previous step's rhs, and its residual, with the
    // residual norms of p_prev and the sums for the next beta.
    if (S.mixed)
      HIPCHK(launch_guess_adapt<float>(S.P[S.cur], S.Eprev, S.Eprev2, (const float *)S.rhs,
                                       S.beta, S.P[1 - S.cur], S.Rl, S.p.nx, S.p.ny, S.p.nz,
                                       per, S.sc.ox, S.sc.oy, S.sc.oz, S.sc.A, g_stream,
                                       S.p0part, S.bpart, S.bsum));
    else
      HIPCHK(launch_guess_adapt<double>(S.P[S.cur], S.Eprev, S.Eprev2, (const double *)S.rhs,
                                        S.beta, S.P[1 - S.cur], S.Rl, S.p.nx, S.p.ny, S.p.nz,
                                        per, S.sc.ox, S.sc.oy, S.sc.oz, S.sc.A, g_stream,
                                        S.p0part, S.bpart, S.bsum));
    HIPCHK(hipMemcpyAsync(S.hbsum, S.bsum, 2 * sizeof(double), hipMemcpyDeviceToHost,
                          g_stream));
    S.beta_pending = true;
    S.p0_of = S.P[S.cur];
  } else if (S.hist >= 1) {
    // The extrapolation into the spare buffer and its residual, one pass,
    // with the residual norms of p_prev for the fallback check.
    if (S.mixed)
      HIPCHK(launch_guess_residual<float>(S.P[S.cur], S.Eprev, (const float *)S.rhs,
                                          S.P[1 - S.cur], S.Rl, S.p.nx, S.p.ny, S.p.nz, per,
                                          S.sc.ox, S.sc.oy, S.sc.oz, S.sc.A, g_stream,
                                          S.p0part));
    else
      HIPCHK(launch_guess_residual<double>(S.P[S.cur], S.Eprev, (const double *)S.rhs,
                                           S.P[1 - S.cur], S.Rl, S.p.nx, S.p.ny, S.p.nz, per,
                                           S.sc.ox, S.sc.oy, S.sc.oz, S.sc.A, g_stream,
                                           S.p0part));
    S.p0_of = S.P[S.cur];
  } else if (S.mixed) {
    // The residual kernel also zeroes its second output; Out is scratch here.
    HIPCHK(launch_residual_f32<float>(S.P[S.cur], (const float *)S.rhs, S.Rl, S.Out, S.p.nx,
                                      S.p.ny, S.p.nz, per, S.sc.ox, S.sc.oy, S.sc.oz, S.sc.A,
                                      g_stream));
  } else {
    HIPCHK(launch_residual_f32<double>(S.P[S.cur], (const double *)S.rhs, S.Rl, S.Out, S.p.nx,
                                       S.p.ny, S.p.nz, per, S.sc.ox, S.sc.oy, S.sc.oz, S.sc.A,
                                       g_stream));
  }
  if (S.cap_r)
    HIPCHK(hipMemcpyAsync(S.cap_r, S.Rl, (size_t)n * 4, hipMemcpyDeviceToDevice, g_stream));
  if (S.cap_eprev)
    HIPCHK(hipMemcpyAsync(S.cap_eprev, S.Eprev, (size_t)n * 4, hipMemcpyDeviceToDevice,
                          g_stream));
  return 0;
}

// Warm guess, accepted when its L2 residual is below that of p_prev:
//  * a network attached: p_prev + e_prev + RMS(e_prev) * trilinear(NN(features)
//    - features[1]), the network's coarse-grid correction of the linear
//    extrapolation, from the extrapolation's residual and e_prev;
//  * O3D_WARM_EXTRAPOLATE: p_prev + e_prev, or with adaptive_extrapolation
//    p_prev + e_prev + beta (e_prev - e_prev2) (formed before, with its
//    residual, by solver_learning_prepare).
int solver_warm_guess(Solver &S, int *used, double *warm) {
  *used = 0;
  warm[0] = warm[1] = 0.0;
  if (S.hist < 1) return 0;
  const int n = S.n;
  double *q = S.P[1 - S.cur];
  bool nn_coarse = false;  // the network's output is the coarse solve's start
  if (S.nn) {
    LwSums w = lw_of(S);
    double re = 0.0;
    const double h2 = S.p.dx * S.p.dx;
    // Sum of squares of e_prev: reduced when e_prev was formed, read back by
    // the previous step's final synchronisation.
    double se2 = 0.0;
    for (int b = 0; b < S.ne; ++b) se2 += S.hepart[b];
    const int rc = lw_features(w, S.Rl, S.Eprev, nullptr, S.p.nx, S.p.ny, S.p.nz, h2, S.Feat,
                               nullptr, &re, se2);
    if (rc < 0) return rc;
    if (rc == 1) return 0;
    CHK(o3d_nn_forward(S.nn, S.Feat, S.Outc));
    const int F = O3D_WARM_COARSEN;
    const int nc = n / (F * F * F);
    int per[3];
    solver_periodic(S.kind, per);
    if (S.p.coarse_cycles > 0 && mg_level1_is(S.mg, S.p.nx / F, S.p.ny / F, S.p.nz / F)) {
      // The network's coarse correction (on the multigrid's level 1, the
      // grid of F^3 blocks) starts the multigrid solve of the coarse-grid
      // correction below; q keeps the extrapolation.
      lw_coarse_init_kernel<<<(nc + 255) / 256, 256, 0, g_stream>>>(S.mg.u[1], S.Outc,
                                                                    S.Feat + nc, nc, (float)re);
      nn_coarse = true;
    } else if (S.p.coarse_cycles == 0) {
      dim3 grid((S.p.nx + 255) / 256, S.p.ny, S.p.nz);
      lw_guess_coarse_kernel<<<grid, 256, 0, g_stream>>>(S.P[S.cur], S.Eprev, S.Outc,
                                                         S.Feat + nc, q, S.p.nx, S.p.ny,
                                                         S.p.nz, F, per[0], per[1], per[2], re);
    }
  }
  // Otherwise q already holds the extrapolation (solver_learning_prepare).
  HIPCHK(hipGetLastError());
  if (S.p.coarse_cycles > 0) {
    // Coarse-grid correction: the residual of q averaged over F^3 blocks,
    // the coarse Poisson problem solved by multigrid V-cycles, its trilinear
    // interpolation added to q.  S.Rl holds the residual of the extrapolation
    // (linear or adaptive); other guesses need their own.
    const int F = kMgCoarsen;
    int per[3];
    solver_periodic(S.kind, per);
    const bool own = S.nn && !nn_coarse;
    if (own)
      HIPCHK(launch_residual_f32<float>(q, (const float *)S.rhs, S.Rl, S.Out, S.p.nx, S.p.ny,
                                        S.p.nz, per, S.sc.ox, S.sc.oy, S.sc.oz, S.sc.A,
                                        g_stream));
    HIPCHK(launch_lw_pool(S.Rl, S.mg.f[0], S.p.nx, S.p.ny, S.p.nz, F, 1.0f));
    if (nn_coarse) HIPCHK(mg_seed_from_level1(S.mg, g_stream));
    // From the network's coarse estimate the first correction may run fewer
    // V-cycles (coarse_cycles_learned) than from zero.
    const int cyc = nn_coarse && S.p.coarse_cycles_learned > 0 ? S.p.coarse_cycles_learned
                                                                : S.p.coarse_cycles;
    HIPCHK(mg_solve(S.mg, cyc, 1, 1, g_stream, !nn_coarse));
    dim3 grid((S.p.nx / 2 + 63) / 64, (S.p.ny + 3) / 4, S.p.nz);
    mg_prolong_fine_kernel<<<grid, dim3(64, 4), 0, g_stream>>>(q, S.mg.u[0], S.p.nx, S.p.ny,
                                                               S.p.nz, F, per[0], per[1], per[2]);
    HIPCHK(hipGetLastError());
  }
  CHK(lw_residual_l2_pair(S, S.P[S.cur], q, warm));
  if (warm[1] < warm[0]) {
    S.cur = 1 - S.cur;
    *used = 1;
  } else {
    S.r_ready = false;  // the solve starts from p_prev, whose residual was not kept
  }
  return 0;
}

// After the correction pass formed e_prev = p_new - p_prev: its sum-of-squares
// partials to the host (read at the next step), the capture, the history.
int solver_learning_finish(Solver &S) {
  HIPCHK(hipMemcpyAsync(S.hepart, S.epart, (size_t)S.ne * sizeof(double),
                        hipMemcpyDeviceToHost, g_stream));
  if (S.cap_e)
    HIPCHK(hipMemcpyAsync(S.cap_e, S.Eprev, (size_t)S.n * 4, hipMemcpyDeviceToDevice,
                          g_stream));
  S.hist++;
  return 0;
}

template <typename T>
int solver_step_t(Solver &S, int itime, o3d_step_info *info) {
  const o3d_params &p = S.p;
  HIPCHK(hipEventRecord(S.ev[2], g_stream));
  CHK(solver_predict<T>(S, itime));
  if (S.xdir) CHK(solver_jet_edges<T>(S, itime));
  CHK(solver_divergence<T>(S, S.up, (T *)S.rhs, p.dt, S.part_div));
  if (S.project) CHK(solver_rhs_project<T>(S));
  HIPCHK(hipEventRecord(S.ev[4], g_stream));
  const bool guess =
      S.nn || p.warm_start == O3D_WARM_EXTRAPOLATE;
  const bool learning = guess || S.cap_r || S.cap_eprev || S.cap_e;
  int warm = 0;
  if (learning) CHK(solver_learning_prepare(S));
  if (guess) CHK(solver_warm_guess(S, &warm, info->warm_dmax));
  if (learning) {
    // Keep p_prev for the increment: with an accepted guess it sits in the
    // other pressure buffer, which trades places with the spare one (the
    // solve ping-pongs between P[cur] and P[1 - cur]); otherwise the solve
    // works on it in place, so it is copied.
    if (warm) std::swap(S.P[1 - S.cur], S.Pstart);
    else HIPCHK(hipMemcpyAsync(S.Pstart, S.P[S.cur], (size_t)S.n * 8, hipMemcpyDeviceToDevice,
                               g_stream));
  }
  if (S.mixed) CHK(solver_poisson_mixed(S, info));
  else CHK(solver_poisson_f64(S, info));
  info->warm_used = warm;
  HIPCHK(hipEventRecord(S.ev[5], g_stream));
  // The correction pass also applies the pending mean shift and, in the
  // warm forms, forms e_prev = p - p_prev (the next step's increment).
  if (learning) std::swap(S.Eprev, S.Eprev2);
  CHK(solver_correct<T>(S, learning));
  if (learning) CHK(solver_learning_finish(S));
  CHK(solver_divergence<T>(S, S.u, nullptr, 1.0, S.part_div2));
  // The step's statistics (their partials stay untouched until here): the
  // rhs divergence, the diagnostic divergence, the three corrected
  // components and, with LES, nu_t.
  {
    FinalizeSet fs[6];
    int nf = 0;
    fs[nf++] = {S.part_div, S.tl_div.blocks(), 1, 0, S.fin + 0};
    fs[nf++] = {S.part_div2, S.tl_div.blocks(), 1, 0, S.fin + 1};
    for (int c = 0; c < 3; ++c) fs[nf++] = {S.part_cor, S.tl_cor.blocks(), 3, c, S.fin + 2 + c};
    if (p.iles == 1) fs[nf++] = {S.part_nut, S.tl.blocks(), 1, 0, S.fin + 5};
    HIPCHK(launch_finalize_sets(fs, nf, g_stream));
  }
  HIPCHK(hipEventRecord(S.ev[3], g_stream));
  HIPCHK(hipMemcpyAsync(S.hfin, S.fin, 6 * sizeof(FieldPartial), hipMemcpyDeviceToHost,
                        g_stream));
  if (S.xdir)
    HIPCHK(hipMemcpyAsync(S.hflux, S.in_flux, 3 * sizeof(double), hipMemcpyDeviceToHost,
                          g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  if (S.xdir) {
    info->flow_rate[0] = S.hflux[0] / S.in_W;
    info->flow_rate[1] = S.hflux[1] / S.in_W;
  } else {
    info->flow_rate[0] = info->flow_rate[1] = 0.0;
  }
  info->extrap_beta = S.beta;
  info->extrap_beta_ls = S.beta_ls;
  if (S.beta_pending) {
    // The least-squares weight of this step's rhs, for the next step's
    // guess, kept within [0, 1.5] (between the linear and somewhat beyond
    // the quadratic extrapolation).
    if (S.hbsum[1] > 0.0) {
      S.beta_ls = S.hbsum[0] / S.hbsum[1];
      S.beta = std::min(1.5, std::max(0.0, S.beta_ls));
    }
    S.beta_pending = false;
  }
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, S.ev[2], S.ev[3]));
  info->step_ms = ms;
  HIPCHK(hipEventElapsedTime(&ms, S.ev[4], S.ev[5]));
  info->poisson_ms = ms;
  stats6_from(S.hfin[0], p.nx, p.ny, p.nz, info->divu_pred);
  stats6_from(S.hfin[1], p.nx, p.ny, p.nz, info->divu);
  if (p.iles == 1) stats6_from(S.hfin[5], p.nx, p.ny, p.nz, info->nu_t);
  else for (int q = 0; q < 6; ++q) info->nu_t[q] = 0.0;
  const double h[3] = {p.dx, p.dy, p.dz};
  info->diverged = 0;
  info->has_nan = 0;
  for (int c = 0; c < 3; ++c) {
    const FieldPartial &r = S.hfin[2 + c];
    info->vel_min[c] = r.vmin;
    info->vel_max[c] = r.vmax;
    info->vel_amax[c] = r.amax;
    info->cfl[c] = r.amax * p.dt / h[c];
    if (r.nan) info->has_nan = 1;
    if (r.nan || r.vmax > 1000.0) info->diverged = 1;
  }
  return info->diverged ? 1 : 0;
}

}  // namespace

// Accessors used by the learned warm start (o3d_nn.hip).
namespace o3d {
double *solver_pressure(void *s) { return ((Solver *)s)->P[((Solver *)s)->cur]; }
double *solver_pressure_other(void *s) { return ((Solver *)s)->P[1 - ((Solver *)s)->cur]; }
void solver_swap_pressure(void *s) { ((Solver *)s)->cur = 1 - ((Solver *)s)->cur; }
const void *solver_rhs(void *s) { return ((Solver *)s)->rhs; }
bool solver_mixed(void *s) { return ((Solver *)s)->mixed; }
void solver_dims(void *s, int *nx, int *ny, int *nz) {
  *nx = ((Solver *)s)->p.nx; *ny = ((Solver *)s)->p.ny; *nz = ((Solver *)s)->p.nz;
}
int solver_kind(void *s) { return ((Solver *)s)->kind; }
void solver_sor_const(void *s, double *c4) {
  const SorConst &c = ((Solver *)s)->sc;
  c4[0] = c.ox; c4[1] = c.oy; c4[2] = c.oz; c4[3] = c.A;
}
double solver_omega(void *s) { return ((Solver *)s)->omega; }
double solver_eps(void *s) { return ((Solver *)s)->p.eps; }
hipStream_t current_stream() { return g_stream; }
int set_error(const char *msg) { return fail("%s", msg); }
}  // namespace o3d

extern "C" int o3d_solver_create(const o3d_params *p, void **handle) {
  CHK(ensure_init());
  if (!p || !handle) return fail("NULL argument");
  CHK(check_grid(p->nx, p->ny, p->nz));
  CHK(check_bc(p->bc, true));
  if (p->bc[0] == DIRICHLET && p->nx < 8)
    return fail("x-Dirichlet needs at least 8 nodes in x (nx = %d)", p->nx);
  if (p->precision == O3D_FP64 && !sor_buffer_fits<double, double>(p->nx, p->ny, p->nz))
    return fail("grid %dx%dx%d: an FP64 field exceeds the 2 GiB the SOR kernels' buffer "
                "descriptors address", p->nx, p->ny, p->nz);
  if (p->precision == O3D_MIXED && !sor_buffer_fits<float, double>(p->nx, p->ny, p->nz))
    return fail("grid %dx%dx%d: an FP64 field exceeds the 2 GiB the SOR kernels' buffer "
                "descriptors address", p->nx, p->ny, p->nz);
  if (p->itscheme < 1 || p->itscheme > 3) {
    fail("itscheme: %d unrecognized", p->itscheme);
    return -2;
  }
  if (p->precision != O3D_FP64 && p->precision != O3D_MIXED)
    return fail("unknown precision form %d", p->precision);
  if (p->warm_start < O3D_WARM_PREVIOUS || p->warm_start > O3D_WARM_EXTRAPOLATE)
    return fail("unknown warm start %d", p->warm_start);
  const int kind = solver_from_bc(p->bc);
  if (kind < 0) return fail("no Poisson solver is bound for x free-slip with y periodic");
  if (p->coarse_cycles < 0 || p->smooth_sweeps < 0 || p->coarse_cycles_learned < 0 ||
      p->refine_cycles < 0)
    return fail("coarse_cycles, coarse_cycles_learned, refine_cycles and smooth_sweeps must be "
                ">= 0");
  if ((p->coarse_cycles > 0 || p->smooth_sweeps > 0) && p->precision != O3D_MIXED)
    return fail("coarse_cycles and smooth_sweeps apply to the mixed form only");
  if (p->coarse_cycles > 0 && (p->nx % 2 || p->ny % 2 || p->nz % 2))
    return fail("the coarse-grid correction needs every axis even");
  CHK(check_red_black(kind, p->nx, p->ny, p->nz));
  Solver *S = new Solver();
  S->p = *p;
  S->n = p->nx * p->ny * p->nz;
  S->mixed = p->precision == O3D_MIXED;
  S->tl = S->mixed ? predict_tiling<float>(p->nx, p->ny, p->nz, p->iles == 1)
                   : predict_tiling<double>(p->nx, p->ny, p->nz, p->iles == 1);
  S->tl_div = S->mixed ? div_tiling<float>(p->nx, p->ny, p->nz)
                       : div_tiling<double>(p->nx, p->ny, p->nz);
  S->tl_cor = S->mixed ? correct_tiling<float, double>(p->nx, p->ny, p->nz)
                       : correct_tiling<double, double>(p->nx, p->ny, p->nz);
  S->kind = kind;
  S->sc = sor_const(p->dx, p->dy, p->dz);
  S->omega = p->omega;
  const size_t n = S->n, es = S->esz();
  int rc = 0;
  for (int c = 0; c < 3 && !rc; ++c) {
    rc |= S->allocv(&S->u[c], n * es);
    rc |= S->allocv(&S->up[c], n * es);
    for (int s = 0; s < 3 && !rc; ++s) rc |= S->allocv(&S->F[s][c], n * es);
  }
  if (!rc) rc |= S->allocv(&S->rhs, n * es);
  if (!rc && p->iles == 1) rc |= S->allocv(&S->nut, n * es);
  if (!rc) rc |= S->alloc(&S->P[0], n);
  if (!rc) rc |= S->alloc(&S->P[1], n);
  if (!rc && S->mixed) {
    rc |= S->alloc(&S->E[0], n);
    rc |= S->alloc(&S->E[1], n);
    rc |= S->alloc(&S->R, n);
    if (!rc && sor_split_supported<float>(p->nx, p->ny, p->nz))
      rc |= S->alloc(&S->dslot, sor_split_slot_words());
    int per[3];
    solver_periodic(kind, per);
    if (!rc && per[0] && per[1] && per[2]) {
      rc |= S->alloc(&S->mpart, kMeanBlocks + 1);
      S->jblocks = (int)sor_tiling<double, float>(p->nx, p->ny, p->nz).blocks();
      if (!rc) rc |= S->alloc(&S->jpart, (size_t)S->jblocks);
    }
  }
  if (!rc && !S->mixed && sor_split_supported<double>(p->nx, p->ny, p->nz)) {
    rc |= S->alloc(&S->dslot, sor_split_slot_words());
    rc |= S->alloc(&S->rhs_split, n);
  }
  if (!rc && p->adaptive_extrapolation) {
    rc |= S->alloc(&S->bpart, 2 * (size_t)residual_blocks(p->nx, p->ny, p->nz));
    if (!rc) rc |= S->alloc(&S->bsum, 2);
    if (!rc && hipHostMalloc((void **)&S->hbsum, 2 * sizeof(double)) != hipSuccess)
      rc = fail("hipHostMalloc failed");
  }
  const bool mirrored =
      p->bc[0] != PERIODIC || p->bc[1] != PERIODIC || p->bc[2] != PERIODIC;
  S->project = p->bc[0] == DIRICHLET || (S->mixed && mirrored && !p->rhs_as_given);
  if (!rc && S->project) {
    rc |= S->alloc(&S->wpart, kWsumBlocks + 2);
    if (!rc) {
      // The sum of the weights, once, by the kernels that form the weighted sums.
      wsum_kernel<double><<<kWsumBlocks, 256, 0, g_stream>>>(
          nullptr, p->nx, p->ny, p->nz, p->bc[0] != PERIODIC, p->bc[1] != PERIODIC,
          p->bc[2] != PERIODIC, S->wpart);
      wsum_final_kernel<<<1, 1024, 0, g_stream>>>(S->wpart, nullptr, S->wpart + kWsumBlocks);
      if (hipGetLastError() != hipSuccess || hipStreamSynchronize(g_stream) != hipSuccess)
        rc = fail("weight sum failed");
    }
  }
  if (!rc && p->bc[0] == DIRICHLET) {
    S->xdir = true;
    const size_t P = (size_t)p->ny * p->nz;
    rc |= S->alloc(&S->in_base, 3 * P);
    if (!rc) rc |= S->alloc(&S->in_w, P);
    if (!rc) rc |= S->alloc(&S->in_pl, 2 * P + 2 * (size_t)p->nz);
    if (!rc) rc |= S->alloc(&S->in_flux, 3);
    if (!rc && hipHostMalloc((void **)&S->hflux, 3 * sizeof(double)) != hipSuccess)
      rc = fail("hipHostMalloc failed");
  }
  const int nb = std::max(S->tl.blocks(), std::max(S->tl_div.blocks(), S->tl_cor.blocks()));
  if (!rc) rc |= S->alloc(&S->part_div, nb);
  if (!rc) rc |= S->alloc(&S->part_div2, nb);
  if (!rc) rc |= S->alloc(&S->part_nut, nb);
  if (!rc) rc |= S->alloc(&S->part_cor, 3 * nb);
  if (!rc) rc |= S->alloc(&S->fin, 6);
  if (!rc) rc |= S->alloc(&S->st, 1);
  if (!rc) rc |= S->alloc(&S->st64, 1);
  if (!rc && hipHostMalloc((void **)&S->hfin, 6 * sizeof(FieldPartial)) != hipSuccess)
    rc = fail("hipHostMalloc failed");
  if (!rc && hipHostMalloc((void **)&S->hst, sizeof(SorState)) != hipSuccess)
    rc = fail("hipHostMalloc failed");
  S->ev.resize(7);
  for (auto &e : S->ev)
    if (!rc && hipEventCreate(&e) != hipSuccess) rc = fail("hipEventCreate failed");
  if (rc) { delete S; return rc; }
  *handle = S;
  return 0;
}

extern "C" int o3d_solver_destroy(void *handle) {
  if (handle) {
    (void)hipStreamSynchronize(g_stream);
    delete (Solver *)handle;
  }
  return 0;
}

namespace {
int upload_field(Solver &S, void *dst, const double *src) {
  const size_t n = S.n;
  Staged s;
  const double *d;
  CHK(s.in(src, n, &d));
  if (S.mixed) HIPCHK((launch_convert<float, double>((float *)dst, d, (int)n, g_stream)));
  else HIPCHK(hipMemcpyAsync(dst, d, n * 8, hipMemcpyDeviceToDevice, g_stream));
  return s.finish();
}
int download_field(Solver &S, double *dst, const void *src) {
  const size_t n = S.n;
  Staged s;
  double *d;
  CHK(s.out(dst, n, &d, false));
  if (S.mixed) HIPCHK((launch_convert<double, float>(d, (const float *)src, (int)n, g_stream)));
  else HIPCHK(hipMemcpyAsync(d, src, n * 8, hipMemcpyDeviceToDevice, g_stream));
  return s.finish();
}
}  // namespace

namespace {
// A pressure written from outside (upload, restart) has no increment
// history: the extrapolated / learned warm start re-primes from it, so a
// fresh solver and one that has already stepped continue alike.
void reset_warm_state(Solver &S) {
  S.hist = 0;
  S.beta = 0.0;
  S.beta_ls = 0.0;
  S.p0_of = nullptr;
  S.r_ready = false;
}
}  // namespace

extern "C" int o3d_solver_upload(void *handle, const double *ux, const double *uy,
                                 const double *uz, const double *pp) {
  if (!handle) return fail("NULL handle");
  Solver &S = *(Solver *)handle;
  const double *u[3] = {ux, uy, uz};
  for (int c = 0; c < 3; ++c)
    if (u[c]) CHK(upload_field(S, S.u[c], u[c]));
  if (pp) {
    Staged s;
    const double *d;
    CHK(s.in(pp, S.n, &d));
    HIPCHK(hipMemcpyAsync(S.P[S.cur], d, (size_t)S.n * 8, hipMemcpyDeviceToDevice, g_stream));
    CHK(s.finish());
    reset_warm_state(S);
  }
  return 0;
}

extern "C" int o3d_solver_download(void *handle, double *ux, double *uy, double *uz, double *pp) {
  if (!handle) return fail("NULL handle");
  Solver &S = *(Solver *)handle;
  double *u[3] = {ux, uy, uz};
  for (int c = 0; c < 3; ++c)
    if (u[c]) CHK(download_field(S, u[c], S.u[c]));
  if (pp) {
    Staged s;
    double *d;
    CHK(s.out(pp, S.n, &d, false));
    HIPCHK(hipMemcpyAsync(d, S.P[S.cur], (size_t)S.n * 8, hipMemcpyDeviceToDevice, g_stream));
    CHK(s.finish());
  }
  return 0;
}

extern "C" int o3d_solver_step(void *handle, int itime, o3d_step_info *info) {
  if (!handle || !info) return fail("NULL argument");
  Solver &S = *(Solver *)handle;
  memset(info, 0, sizeof(*info));
  S.sor_ms[0] = S.sor_ms[1] = 0.0;
  S.sor_sweeps[0] = S.sor_sweeps[1] = 0;
  S.sor_launches[0] = S.sor_launches[1] = 0;
  return S.mixed ? solver_step_t<float>(S, itime, info) : solver_step_t<double>(S, itime, info);
}

extern "C" int o3d_solver_statistics(void *handle, double t, double *out17) {
  if (!handle || !out17) return fail("NULL argument");
  Solver &S = *(Solver *)handle;
  Staged s;
  if (S.mixed) {
    const float *u[3] = {(const float *)S.u[0], (const float *)S.u[1], (const float *)S.u[2]};
    return statistics_dev<float>(u, S.p.nx, S.p.ny, S.p.nz, S.p.dx, S.p.dy, S.p.dz, S.p.re, t,
                                 S.p.bc, out17, s);
  }
  const double *u[3] = {(const double *)S.u[0], (const double *)S.u[1], (const double *)S.u[2]};
  return statistics_dev<double>(u, S.p.nx, S.p.ny, S.p.nz, S.p.dx, S.p.dy, S.p.dz, S.p.re, t,
                                S.p.bc, out17, s);
}

extern "C" int o3d_copy(void *dst, const void *src, size_t bytes) {
  CHK(ensure_init());
  if (bytes == 0) return 0;
  if (!dst || !src) return fail("o3d_copy: NULL pointer");
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  return 0;
}

extern "C" void *o3d_solver_field(void *handle, int which) {
  if (!handle) return nullptr;
  Solver &S = *(Solver *)handle;
  switch (which) {
    case 0: case 1: case 2: return S.u[which];
    case 3: case 4: case 5: return S.up[which - 3];
    case 6: return S.P[S.cur];
    case 7: return S.rhs;
    case 8: return S.nut;
    default: return nullptr;
  }
}

extern "C" int o3d_solver_history(void *handle, int write, void *hist, double *state2) {
  if (!handle || !hist || !state2) return fail("NULL argument");
  Solver &S = *(Solver *)handle;
  const size_t bytes = (size_t)S.n * S.esz();
  if (write) {
    const double r = state2[0];
    if (!(r >= 0.0) || r > 1e9 || r != (double)(long long)r)
      return fail("history depth %g is not a count", r);
    S.ring = (int)r;
    S.omega = state2[1];
    reset_warm_state(S);
  }
  for (int age = 0; age < 3; ++age) {
    const int slot = ((S.ring - 1 - age) % 3 + 3) % 3;
    for (int c = 0; c < 3; ++c) {
      char *h = (char *)hist + (size_t)(3 * age + c) * bytes;
      if (write)
        HIPCHK(hipMemcpyAsync(S.F[slot][c], h, bytes, hipMemcpyHostToDevice, g_stream));
      else
        HIPCHK(hipMemcpyAsync(h, S.F[slot][c], bytes, hipMemcpyDeviceToHost, g_stream));
    }
  }
  HIPCHK(hipStreamSynchronize(g_stream));
  if (!write) {
    state2[0] = S.ring;
    state2[1] = S.omega;
  }
  return 0;
}

extern "C" int o3d_solver_sor_timing(void *handle, double *ms2, int *sweeps2, int *launches2) {
  if (!handle || !ms2 || !sweeps2) return fail("NULL argument");
  Solver &S = *(Solver *)handle;
  for (int q = 0; q < 2; ++q) {
    ms2[q] = S.sor_ms[q];
    sweeps2[q] = S.sor_sweeps[q];
    if (launches2) launches2[q] = S.sor_launches[q];
  }
  return 0;
}

extern "C" int o3d_solver_sor_period(void *handle, int iterations, double *us) {
  if (!handle || !us) return fail("NULL argument");
  Solver &S = *(Solver *)handle;
  if (!S.mixed || !S.dslot || !S.R || !S.E[0] || !S.E[1])
    return fail("sor period: needs the mixed form with colour-split sweeps");
  if (iterations < 2 || iterations > 100000)
    return fail("sor period: iterations %d outside [2, 100000]", iterations);
  SorRun<float, float> r;
  r.st = S.st;
  r.hst = S.hst;
  r.events = &S.ev;
  r.buf[0] = S.E[0];
  r.buf[1] = S.E[1];
  sor_args_common<float, float>(S, r.a, S.R, S.st, iterations);
  r.a.eps = 0.0;  // never converged, never stagnant: every launch sweeps
  const int h = S.n / 2;
  r.split = true;
  r.sa = split_args<float>(r.a, S.E[0], S.E[0] + h, S.R, S.R + h, S.dslot);
  HIPCHK(launch_sor_init(S.st, S.omega, 0, g_stream));
  HIPCHK(launch_sor_split_init(S.dslot, g_stream));
  hipEvent_t e0 = S.ev[0], e1 = S.ev[1];
  for (int it = 1; it <= iterations; ++it) {
    r.a.it = it;
    HIPCHK(r.launch());
    if (it == 1) HIPCHK(hipEventRecord(e0, g_stream));
  }
  HIPCHK(hipEventRecord(e1, g_stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *us = 1e3 * (double)ms / (iterations - 1);
  return 0;
}

extern "C" int o3d_solver_set_warm_model(void *handle, void *nn) {
  if (!handle) return fail("NULL handle");
  Solver &S = *(Solver *)handle;
  if (nn) {
    const int F = O3D_WARM_COARSEN;
    if (S.p.nx % (8 * F) || S.p.ny % (8 * F) || S.p.nz % (8 * F))
      return fail("learned warm start: every grid axis must be a multiple of %d (got %dx%dx%d)",
                  8 * F, S.p.nx, S.p.ny, S.p.nz);
    CHK(o3d_nn_resize(nn, S.p.nx / F, S.p.ny / F, S.p.nz / F));
  }
  S.nn = nn;
  return 0;
}

namespace {
LwSums g_lw;  // reduction buffers of o3d_warm_features
}

extern "C" int o3d_warm_features(const float *r, const float *eprev, const float *e, int nx,
                                 int ny, int nz, double dx, float *inp, float *tgt,
                                 double *scale) {
  const int F = O3D_WARM_COARSEN;
  if (!r || !eprev || !inp || !scale) return fail("NULL argument");
  if (nx <= 0 || ny <= 0 || nz <= 0 || nx % F || ny % F || nz % F)
    return fail("grid %dx%dx%d: every axis must be a positive multiple of %d", nx, ny, nz, F);
  CHK(lw_sums_alloc(g_lw));
  const int rc = lw_features(g_lw, r, eprev, e, nx, ny, nz, dx * dx, inp, tgt, scale);
  if (rc < 0) return rc;
  HIPCHK(hipStreamSynchronize(g_stream));
  return rc;
}

extern "C" int o3d_solver_set_inflow(void *handle, const double *base3, const double *weight,
                                     double amp, double period, unsigned long long seed,
                                     double uc) {
  if (!handle) return fail("NULL handle");
  Solver &S = *(Solver *)handle;
  if (!S.xdir) return fail("o3d_solver_set_inflow: the solver has no x-Dirichlet axis");
  if (!base3 || !weight) return fail("o3d_solver_set_inflow: NULL plane");
  if (!(period > 0.0)) return fail("o3d_solver_set_inflow: period must be > 0");
  const o3d_params &p = S.p;
  const size_t P = (size_t)p.ny * p.nz;
  Staged s;
  const double *b, *w;
  CHK(s.in(base3, 3 * P, &b));
  CHK(s.in(weight, P, &w));
  HIPCHK(hipMemcpyAsync(S.in_base, b, 3 * P * sizeof(double), hipMemcpyDeviceToDevice,
                        g_stream));
  HIPCHK(hipMemcpyAsync(S.in_w, w, P * sizeof(double), hipMemcpyDeviceToDevice, g_stream));
  CHK(s.finish());
  // W = the weighted plane sum of ones, in the order jet_flux_kernel sums.
  const int my = p.bc[1] != PERIODIC, mz = p.bc[2] != PERIODIC;
  double W = 0.0;
  for (int k = 0; k < p.nz; ++k) {
    double r = 0.0;
    for (int j = 0; j < p.ny; ++j) r = r + edge_weight(j, p.ny, my) * 1.0;
    W = W + edge_weight(k, p.nz, mz) * r;
  }
  S.in_W = W;
  S.in_amp = amp;
  S.in_period = period;
  S.in_seed = seed;
  S.in_cu = uc * p.dt / p.dx;
  S.in_set = true;
  return 0;
}

extern "C" int o3d_solver_capture(void *handle, float *r_out, float *eprev_out, float *e_out) {
  if (!handle) return fail("NULL handle");
  Solver &S = *(Solver *)handle;
  S.cap_r = r_out;
  S.cap_eprev = eprev_out;
  S.cap_e = e_out;
  return 0;
}
