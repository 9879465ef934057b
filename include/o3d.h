/* osinco3d time-step path on MI355X: the C ABI of libo3d.so.
 *
 * This is the drop-in boundary.  The reference is a serial Fortran code whose
 * time step calls module procedures by name and derivative / Poisson routines
 * through procedure pointers (src/initialization.f90:86-110, bound in
 * schemes() at :226-304).  The Fortran shim modules in osinco3d_amd/fortran/
 * keep those module and procedure names and dummy-argument lists and call the
 * functions below through ISO_C_BINDING.
 *
 * Conventions
 *  - Fields are contiguous column-major (nx, ny, nz) arrays, x fastest, as in
 *    the reference.  float64 unless the name ends in _f32.
 *  - Pointers may be host or device memory: host arrays are staged through
 *    device buffers (correct, slower); device arrays are used in place.
 *  - bc[3]: per-axis boundary type, 0 = periodic, 1 = free-slip
 *    (src/initialization.f90:77-78).  Both faces of an axis share one type,
 *    as schemes() requires.  2 = x-Dirichlet inflow / outflow (x axis only;
 *    the reference's README lists it, its schemes() stops on it,
 *    src/initialization.f90:238-241): accepted by o3d_der (which 20, 21),
 *    o3d_divergence, o3d_nu_t, o3d_predict_velocity, o3d_correct_velocity,
 *    o3d_statistics and the device-resident solver.  Velocity derivatives
 *    along x use the cubic-extrapolated ends of o3d_der 20 / 21; the
 *    pressure is mirrored in x (homogeneous Neumann), its x derivative zero
 *    on the end nodes; o3d_correct_velocity leaves the x-end planes at u*.
 *  - Status: 0 = ok; 1 = velocity diverged (the reference prints
 *    "Velocity has diverged" and stops, src/integration.f90:309-325);
 *    < 0 = runtime or argument error, text from o3d_last_error().
 *  - One caller thread, one HIP stream; every call returns with its host
 *    outputs valid.  No C++ exception crosses this interface.
 */
#ifndef O3D_H
#define O3D_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- runtime */
int o3d_init(int device);
/* Launch stream (a hipStream_t); NULL = the default stream. */
int o3d_set_stream(void *stream);
const char *o3d_last_error(void);
int o3d_sync(void);
/* Library version string. */
const char *o3d_version(void);
/* Measurement: a hand-written streaming triad c = a + s b over three device
 * buffers of bytes_per_buffer each (two reads and one write per element; at
 * 1 GiB per buffer the working set is far beyond the 256 MB Infinity Cache),
 * timed with HIP events over `reps` launches, once with plain and once with
 * nontemporal loads and stores; the faster is reported.  gbs = 3 *
 * bytes_per_buffer / time per launch (GB/s), ms = time per launch; either may
 * be NULL.  No
 * counterpart in the reference: the HBM ceiling the kernels are compared
 * with. */
int o3d_stream_probe(size_t bytes_per_buffer, int reps, double *gbs, double *ms);

/* ----------------------------------------------- device-field registry
 * Host arrays the caller keeps for the whole run (the driver's fields,
 * src/initialization.f90:50-59) can be registered: the library then keeps a
 * device copy across calls (the SURVEY's "registry keyed by host address
 * plus size, with host-valid / device-valid flags").  A registered input
 * whose device copy is current is not copied again; a registered output
 * stays on the device and is copied back to the host at the end of the call
 * only when `mirror` is set (host code reads it).  Other host arrays, and
 * any part of a registered array passed on its own, are staged per call as
 * before (a stale host copy is refreshed first).  Contract: host code writes
 * a registered array only before its first use by the library, or calls
 * o3d_mark_host_dirty after writing it.  No counterpart in the reference
 * (its arrays are host memory only). */
int o3d_register(const void *host, size_t bytes, int mirror);
int o3d_unregister(const void *host);      /* copies a stale host copy back */
int o3d_mark_host_dirty(const void *host); /* the host copy changed */
int o3d_sync_host(const void *host);       /* make the host copy current now */
/* Bytes moved host -> device and device -> host by the staging and the
 * registry so far; reset != 0 zeroes the counters. */
int o3d_transfer_stats(double *h2d_bytes, double *d2h_bytes, int reset);

/* ------------------------------------------------------ operator entry points */

/* One routine of src/derivation.f90 (interface der_type(df, f, d),
 * src/initialization.f90:86-91).  which = axis*6 + (order-1)*3 + variant,
 * variant 0 = "_00" periodic, 1 = "p_11" free-slip even, 2 = "i_11" free-slip
 * odd; 18 = derz_2dsim, 19 = derzz_2dsim.  E.g. which 0 = derx_00
 * (src/derivation.f90:6-60), 8 = deryi_11 (:269-321), 17 = derzzi_11
 * (:836-883).  20 / 21: the x first / second derivative with x-Dirichlet
 * ends (no counterpart in the reference): the values 1..3 nodes beyond an
 * end are the cubic through the four nodes nearest it, so the stencil is
 * exact on cubics up to the end nodes. */
int o3d_der(int which, double *df, const double *f, int nx, int ny, int nz, double d);
int o3d_der_f32(int which, float *df, const float *f, int nx, int ny, int nz, double d);

/* divergence(divf, fx, fy, fz, dx, dy, dz, nx, ny, nz, odd),
 * src/differential_operators.f90:7-38.  divf = div / divisor (divisor = 1 for
 * the plain operator, dt for the Poisson rhs of src/integration.f90:239).
 * divf may be NULL (statistics only).  stats6 (host, may be NULL) receives
 * function_stats of the undivided divergence (src/functions.f90:27-63). */
int o3d_divergence(double *divf, const double *fx, const double *fy, const double *fz,
                   double dx, double dy, double dz, int nx, int ny, int nz, const int *bc,
                   int odd, double divisor, double *stats6);

/* calculate_nu_t(nu_t, ux, uy, uz, dx, dy, dz, cs, delta),
 * src/les_turbulence.f90:10-97; stats6 (host, may be NULL) = function_stats
 * of nu_t (the reference prints min, max, mean). */
int o3d_nu_t(double *nu_t, const double *ux, const double *uy, const double *uz, double dx,
             double dy, double dz, double cs, double delta, int nx, int ny, int nz,
             const int *bc, double *stats6);

/* predict_velocity, src/integration.f90:14-197, same argument meaning:
 * fux/fuy/fuz are (nx,ny,nz,3) history arrays, slot 1 = newest, shifted by
 * copy at the end as the reference does; adt/bdt/cdt are the 3-element
 * coefficient arrays of src/initialization.f90:194-202.  Returns -2 for an
 * unrecognized itscheme (the reference stops). */
int o3d_predict_velocity(double *ux_pred, double *uy_pred, double *uz_pred, const double *ux,
                         const double *uy, const double *uz, double *fux, double *fuy,
                         double *fuz, double re, const double *adt, const double *bdt,
                         const double *cdt, int itime, int itscheme, double dx, double dy,
                         double dz, int nx, int ny, int nz, int iles, double cs, double delta,
                         double *nu_t, const int *bc);

/* poisson_solver_0000 / _0011 / _111111 (interface poi_type,
 * src/initialization.f90:93-102; bodies src/poisson.f90:6-381), red-black
 * ordering.  solver: 0 = "0000", 1 = "0011", 2 = "111111"; 3 + (y mirrored)
 * + 2 (z mirrored) = the x-Dirichlet kinds (x mirrored; no counterpart in
 * the reference).  omega is in/out
 * (dynamic omega persists, idyn = 1).  exit_iter = the iteration the
 * reference prints on exit (kmax+1 when kmax ran out); reason 0 = dmax < eps,
 * 1 = stagnation, 2 = kmax.  Any of the three out pointers may be NULL. */
int o3d_poisson(double *pp, const double *rhs, double dx, double dy, double dz, int nx, int ny,
                int nz, double *omega, double eps, int kmax, int idyn, int solver,
                int *exit_iter, double *dmax, int *reason);

/* The red-black SOR of o3d_poisson in FP32 (the sweeps of the mixed form),
 * same rules and outputs.  variant selects the kernels, both with
 * bit-identical iterates: 0 = one fused red+black iteration per launch;
 * 2 = colour-split storage, one in-place launch per colour (nx a multiple
 * of 8; the form the mixed solver uses).  No counterpart in the reference
 * (its solver is FP64). */
int o3d_poisson_f32(float *pp, const float *rhs, double dx, double dy, double dz, int nx, int ny,
                    int nz, double *omega, double eps, int kmax, int idyn, int solver,
                    int variant, int *exit_iter, double *dmax, int *reason);

/* The coarse-grid solver of the warm starts, on its own: `cycles` multigrid
 * V(2,2) cycles on L u = f from u = 0 (the solver's coarse-grid correction
 * runs the same cycles as V(1,1)), L the 7-point Laplacian of a
 * cell-centred mx x my x mz grid with spacings hx, hy, hz; per[a] != 0 makes
 * axis a periodic, else the ghost cell equals the edge cell (homogeneous
 * Neumann).  Levels are halved while every axis is even and >= 4; the
 * coarsest (at most 64 cells, or where halving stops) is solved by
 * Gauss-Seidel sweeps (30, or 60 above 64 cells).  u, f: host or device
 * float32.  No counterpart in the reference (the mixed forms' coarse-grid
 * correction, o3d_params.coarse_cycles). */
int o3d_multigrid(float *u, const float *f, int mx, int my, int mz, double hx, double hy,
                  double hz, const int *per, int cycles);

/* solve_poisson_multigrid (src/poisson_multigrid.f90:10-18, called from
 * correct_pression when multigrid = 1, src/integration.f90:242-244; that
 * solver does not run as written: its arrays are declared with ghost layers
 * the caller does not pass and its level count is kmax).  A working
 * multigrid solve of the problem the SOR solvers solve (same operator and
 * neighbour rules, solver as for o3d_poisson): cycles of FP64 residual, stop
 * when max |rhs - L pp| < tol (the reference's stated criterion), coarse-grid
 * correction on the 2^3 block averages by one V(npre, npost) cycle over at
 * most nlevels - 1 coarser levels, npre + npost red-black Gauss-Seidel sweeps
 * on the full grid, pp += correction.  With a mirrored (free-slip) axis the
 * operator's columns sum to zero under the trapezoid weights, so only an
 * rhs of zero weighted mean can be met: the solve works on the rhs minus its
 * weighted mean (the part any solver can satisfy) and r is judged on it.
 * Stop rules as the SOR solvers' (src/poisson.f90:95-122) on max |r|:
 * reason 0 = max |r| < tol, 1 = stagnation (max |r| changed by less than
 * tol / 1000 since the last cycle), 2 = max_cycles cycles done.  cycles,
 * resmax (the final max |r|) and reason may be NULL.  Needs nx a multiple
 * of 8 and ny, nz even. */
int o3d_poisson_multigrid(double *pp, const double *rhs, double dx, double dy, double dz,
                          int nx, int ny, int nz, int solver, int nlevels, int npre, int npost,
                          double tol, int max_cycles, int *cycles, double *resmax, int *reason);

/* correct_pression, src/integration.f90:199-255 (SOR branch): divergence of
 * the predicted velocity (odd variants), rhs = div/dt, then the Poisson solver
 * that schemes() binds for bc (src/initialization.f90:283-301). */
int o3d_correct_pression(double *pp, const double *ux_pred, const double *uy_pred,
                         const double *uz_pred, double dx, double dy, double dz, int nx, int ny,
                         int nz, double dt, double *omega, double eps, int kmax, int idyn,
                         const int *bc, int *exit_iter, double *dmax, int *reason);

/* correct_velocity, src/integration.f90:257-330.  vstats (host, may be
 * NULL) receives per component c: [4c+0] min, [4c+1] max, [4c+2] max |u|,
 * [4c+3] NaN flag.  Returns 1 when the reference's guard (NaN, or
 * maxval(u) > 1000) fires. */
int o3d_correct_velocity(double *ux, double *uy, double *uz, const double *ux_pred,
                         const double *uy_pred, const double *uz_pred, const double *pp,
                         double dt, double dx, double dy, double dz, int nx, int ny, int nz,
                         const int *bc, double *vstats);

/* function_stats, src/functions.f90:27-63: min, max, mean, then the 1-based
 * (i, j, k) of the first maximum in column-major order. */
int o3d_function_stats(const double *f, int nx, int ny, int nz, double *stats6);

/* statistics_calc, src/utils.f90:243-375: out17 = t, E_k, eps, eps2,
 * enstrophy, <u^2>, <v^2>, <w^2>, then the nine <(du_c/dx_a)^2> in the order
 * of the stats.dat columns (src/IOfunctions.f90:552-555). */
int o3d_statistics(const double *ux, const double *uy, const double *uz, int nx, int ny, int nz,
                   double dx, double dy, double dz, double re, double t, const int *bc,
                   double *out17);

/* o3d_statistics without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments
 * and results (bit for bit) as o3d_statistics, but ux, uy, uz must be device
 * arrays, and the per-block partial sums live in a fixed 2 MiB device array
 * of the library (one per GPU, held until the process ends): grids needing
 * more than 16384 partial records (the blocks of the stencil tiling, e.g.
 * 256^3 needs 1024) are rejected with an error.  Calls on one GPU share
 * that array, so calls that overlap on one GPU (other threads or streams)
 * race. */
int o3d_statistics_noalloc(const double *ux, const double *uy, const double *uz, int nx,
                           int ny, int nz, double dx, double dy, double dz, double re, double t,
                           const int *bc, double *out17);

/* transeq, src/integration.f90:332-468: passive scalar transport.  fphi is
 * the (nx,ny,nz,3) history (slot 1 = newest, shifted by copy as the
 * reference does); the p (even) derivative variants; alpha_eff =
 * 1/(re sc) (+ nu_t/sc with iles = 1, nu_t read only then); the AB update;
 * then the conservative clipping to [0, 1] with FP64 global sums (their
 * summation order differs from the Fortran sum's, so clipped values agree to
 * round-off; without clipping the result is bit-identical).  A zero weight
 * sum (every value at a bound, where the reference divides 0 by 0) leaves the
 * clipped field unchanged.  Returns -2 for an unrecognized itscheme. */
int o3d_transeq(double *phi, const double *ux, const double *uy, const double *uz,
                const double *src, double *fphi, double re, double sc, const double *adt,
                const double *bdt, const double *cdt, int itime, int itscheme, double dx,
                double dy, double dz, int nx, int ny, int nz, int iles, const double *nu_t,
                const int *bc);

/* rotational(rotx, roty, rotz, ux, uy, uz, dx, dy, dz, nx, ny, nz),
 * src/differential_operators.f90:40-76: the curl from the p (even) D1
 * variants.  One fused pass over (ux, uy, uz); bc as for o3d_divergence. */
int o3d_rotational(double *rotx, double *roty, double *rotz, const double *ux, const double *uy,
                   const double *uz, double dx, double dy, double dz, int nx, int ny, int nz,
                   const int *bc);

/* calculate_Q_criterion(Q, ux, uy, uz, dx, dy, dz, nx, ny, nz),
 * src/differential_operators.f90:78-108: Q = -1/2 (ux_x^2 + uy_y^2 + uz_z^2)
 * - ux_y uy_x - ux_z uz_x - uy_z uz_y, diagonal terms from the i (odd)
 * variants, the others from the p variants, in the reference's order. */
int o3d_q_criterion(double *q, const double *ux, const double *uy, const double *uz, double dx,
                    double dy, double dz, int nx, int ny, int nz, const int *bc);

/* ------------------------------------------------- device-resident solver */

/* Precision forms. */
#define O3D_FP64 0  /* plain FP64: the drop-in (parity with the reference) */
#define O3D_MIXED 1 /* FP32 stencils, AB update and SOR sweeps; FP64-refined Poisson */

/* Warm start of the Poisson solve. */
#define O3D_WARM_PREVIOUS 0 /* previous step's pressure (the reference's behaviour) */
#define O3D_WARM_LEARNED 1  /* learned correction, previous pressure as fallback */
#define O3D_WARM_EXTRAPOLATE 2 /* p_prev + (p_prev - p_prev2), previous pressure as fallback
                                  (see adaptive_extrapolation for the weighted second
                                  difference) */

typedef struct o3d_params {
  int nx, ny, nz;
  double dx, dy, dz;
  int bc[3];
  int itscheme;         /* 1 Euler, 2 AB2, 3 AB3 */
  double dt, re;
  double adt[3], bdt[3], cdt[3];
  int iles;
  double cs, delta;
  double omega, eps;
  int kmax, idyn;
  int precision;        /* O3D_FP64 or O3D_MIXED */
  int warm_start;       /* O3D_WARM_PREVIOUS, _LEARNED or _EXTRAPOLATE */
  int inner_max;        /* O3D_MIXED: FP32 sweeps between FP64 refinements (0 = default) */
  int coarse_cycles;    /* O3D_MIXED with a warm start: multigrid V-cycles of the coarse-grid
                           correction added to the guess, on 2^3 block averages (0 = none;
                           every axis even) */
  int smooth_sweeps;    /* O3D_MIXED: Gauss-Seidel sweeps (omega = 1) that open the first FP32
                           correction solve of each step (0 = none) */
  int coarse_cycles_learned; /* with the learned warm start: V-cycles of the first coarse-grid
                                correction, which starts from the network's coarse estimate
                                instead of zero (0 = coarse_cycles) */
  int adaptive_extrapolation; /* O3D_WARM_EXTRAPOLATE / _LEARNED: extrapolate
                                 p_prev + e_prev + beta (e_prev - e_prev2) (beta 0: linear, 1:
                                 quadratic) with beta the least-squares weight of the previous
                                 step's rhs (the L2 residual of the guess minimised over beta),
                                 kept in [0, 1.5]; 0 = linear.  With a network attached it
                                 applies only when coarse_cycles > 0: the network's input is
                                 then the residual of this guess, and the correction it
                                 estimates from it starts the coarse-grid solve */
  int rhs_as_given;           /* mixed form on a grid with a mirrored axis (free-slip y or z,
                                 x-Dirichlet excepted): 0 = remove the rhs's weighted mean
                                 (trapezoid weights on the mirrored axes), the part the
                                 mirrored operator cannot reach, so every solve can end on eps;
                                 1 = solve the rhs as the divergence gives it, as the FP64 form
                                 does (its solves then end on the stagnation rule).  The FP64
                                 form always keeps the reference's rhs; x-Dirichlet grids
                                 always project */
  int refine_cycles;          /* with the coarse-grid correction: V-cycles of the correction
                                 that opens each later refinement cycle of a step (0 =
                                 coarse_cycles) */
} o3d_params;

typedef struct o3d_step_info {
  int sor_exit_iter;      /* as the reference prints it */
  int sor_reason;         /* 0 eps, 1 stagnation, 2 kmax */
  int sweeps_f64;         /* FP64 red-black sweeps this step */
  int sweeps_f32;         /* FP32 red-black sweeps this step (O3D_MIXED) */
  int refinements;        /* FP64 residual refinements (O3D_MIXED) */
  int warm_used;          /* 1 when the learned warm start was accepted */
  int diverged;           /* reference's "Velocity has diverged" guard */
  int has_nan;
  double sor_dmax;        /* dmax of the last (FP64) sweep */
  double omega;           /* omega after the solve */
  double poisson_ms;      /* device time of the Poisson solve */
  double step_ms;         /* device time of the whole step */
  double warm_dmax[2];    /* L2 residual of the initial guess: [0] previous pressure,
                             [1] the warm guess (0 when none was formed) */
  double divu_pred[6];    /* function_stats of div(u*) */
  double divu[6];         /* function_stats of div(u) */
  double vel_min[3], vel_max[3], vel_amax[3];
  double cfl[3];          /* max|u_c| dt / d_c, src/utils.f90:178-208 */
  double nu_t[6];         /* function_stats of nu_t (iles = 1) */
  double extrap_beta;     /* adaptive extrapolation: the weight beta used this step */
  double flow_rate[2];    /* x-Dirichlet: mean inflow and outflow velocity (the weighted
                             plane fluxes / plane weight, outflow before its flux shift), as
                             print_in_outflow_rate prints them (src/IOfunctions.f90:342-350) */
  double extrap_beta_ls;  /* adaptive extrapolation: the unclamped least-squares weight
                             that extrap_beta was clamped from (0 before the first) */
} o3d_step_info;

int o3d_solver_create(const o3d_params *p, void **handle);
int o3d_solver_destroy(void *handle);
/* Host or device float64 arrays in, copied into the solver's own fields. */
int o3d_solver_upload(void *handle, const double *ux, const double *uy, const double *uz,
                      const double *pp);
int o3d_solver_download(void *handle, double *ux, double *uy, double *uz, double *pp);
/* One time step of src/osinco3d_main.f90:97-128: predict_velocity,
 * correct_pression, correct_velocity and the driver's per-step diagnostics
 * (both divergences, velocity min/max, CFL), with itime as in the driver. */
int o3d_solver_step(void *handle, int itime, o3d_step_info *info);
int o3d_solver_statistics(void *handle, double t, double *out17);
/* Device pointer of a solver field: 0..2 u, 3..5 u*, 6 p (float64), 7 rhs,
 * 8 nu_t.  Element type is float64 for O3D_FP64 and float32 otherwise
 * (p is float64 in both forms). */
void *o3d_solver_field(void *handle, int which);
/* Copy `bytes` between host and/or device memory (either side may be a
 * device pointer such as o3d_solver_field's, or host memory), on the launch
 * stream, and wait for it.  No counterpart in the reference: how a C, Fortran
 * or Python caller reads a solver field without a HIP binding of its own. */
int o3d_copy(void *dst, const void *src, size_t bytes);
/* Time of SOR sweeps measured with HIP events on the launch stream over the
 * last step: total ms of complete sweep batches, the number of sweeps
 * (iterations) and of kernel launches in them, per precision (index 0 FP64,
 * 1 FP32; one iteration per launch: a fused red+black launch, or with
 * colour-split storage the red and the black launch counted as one).
 * launches2 may be NULL. */
int o3d_solver_sor_timing(void *handle, double *ms2, int *sweeps2, int *launches2);
/* Mixed form with colour-split sweeps: `iterations` back-to-back FP32
 * red-black iterations (two launches each) of the solver's correction
 * problem, eps 0 so none stops, on its correction and residual buffers
 * (scratch between solves), timed with HIP events on the launch stream from
 * the end of the first iteration to the end of the last; us = the mean time
 * per iteration.  For measurement: the event interval spans many iterations,
 * so it holds the kernels and the gaps between back-to-back launches only.
 * No counterpart in the reference. */
int o3d_solver_sor_period(void *handle, int iterations, double *us);
/* Adams-Bashforth history for checkpoint / restart (the reference has none:
 * it restarts with an Euler step).  hist: host array of 9 n elements, float64
 * in O3D_FP64 and float32 in O3D_MIXED, ordered by age (0 = the newest
 * right-hand side F, as fux(:,:,:,1) after predict_velocity) then component
 * (x, y, z); state2 = {history depth (predict calls so far), omega}.
 * write = 0 reads the solver's history into hist / state2, write = 1 sets it.
 * With the fields (o3d_solver_upload) this restores a run with the previous
 * pressure as warm start (the reference's) exactly.  The extrapolated and
 * learned warm starts keep no history across a restart: writing the history
 * or uploading a pressure re-primes them (the first step after it starts
 * from the previous pressure), so a restart continues exactly as any other
 * restart from the same checkpoint, not as the uninterrupted run. */
int o3d_solver_history(void *handle, int write, void *hist, double *state2);
/* Attach the learned warm start (an o3d_nn network, include/o3d_nn.h), or
 * NULL to detach; the network is re-sized to the grid coarsened by
 * O3D_WARM_COARSEN (every axis must be a multiple of 8 * O3D_WARM_COARSEN).
 * Each step then forms, before the Poisson solve, the linear extrapolation
 * p_prev + e_prev of the pressure (e_prev = p_prev - p_prev2, the previous
 * step's increment) and its residual r = rhs - L (p_prev + e_prev), feeds the
 * network their block averages (o3d_warm_features), and offers
 * p_prev + e_prev + RMS(e_prev) * trilinear(output - input plane 1) as the
 * initial guess; the previous pressure stays the initial guess unless the
 * guess has the smaller L2 residual. */
int o3d_solver_set_warm_model(void *handle, void *nn);
/* x-Dirichlet inflow (bc[0] == 2; no counterpart in the reference, which
 * stops on this boundary, src/initialization.f90:238-241).  Every step,
 * after predict:
 *  - inflow plane i = 1: u*_c(j, k) = base3[c][j, k]
 *      + (amp weight[j, k]) (x1 cos(phi) + x2 sin(phi)),
 *    phi = 2 pi itime dt / period, x1, x2 uniform in [-1, 1) from SplitMix64
 *    of seed + 2 (3 (j + ny k) + c) (+ 1);
 *  - outflow plane i = nx: first-order upwind convection at velocity uc,
 *    u*_c = u_c - (uc dt / dx) (u_c(nx) - u_c(nx - 1)), then u*_x shifted by
 *    one constant so that the outflow flux equals the inflow flux (plane sums
 *    with weight 1/2 on the end nodes of mirrored axes);
 *  - the Poisson rhs minus its weighted mean (the same weights in 3D), the
 *    part the x-mirrored operator cannot reach, so SOR converges to eps;
 *  - o3d_correct_velocity's rule: the two planes keep u*.
 * base3: 3 planes of ny nz values (x, y, z components, j fastest), weight one
 * plane; host or device float64.  Must be called before the first step. */
int o3d_solver_set_inflow(void *handle, const double *base3, const double *weight, double amp,
                          double period, unsigned long long seed, double uc);

/* Training data capture: when set, every step writes float32 planes of n
 * cells each: r_out = rhs - L (p_prev + e_prev) before the solve (rhs - L p_prev
 * while no increment exists), eprev_out = the previous step's pressure
 * increment e_prev, e_out = this step's increment p_new - p_prev.  NULL
 * pointers disable the capture. */
int o3d_solver_capture(void *handle, float *r_out, float *eprev_out, float *e_out);

/* Coarsening factor of the learned warm start: the network runs on block
 * averages over O3D_WARM_COARSEN^3 cells, and its correction is interpolated
 * back trilinearly. */
#define O3D_WARM_COARSEN 4
/* The network's input and target for one captured step, as the solver forms
 * them (device float32 arrays, n = nx*ny*nz fine cells, nc = n / 64 coarse
 * cells, F = O3D_WARM_COARSEN): inp (2 nc) = block averages of
 * r dx^2 F^2 / RMS(e_prev) and of e_prev / RMS(e_prev); tgt (nc, may be NULL,
 * then e may be NULL) = block averages of e / RMS(e_prev); scale =
 * RMS(e_prev).  r, e_prev, e as o3d_solver_capture writes them.  Returns 1,
 * with nothing written, when RMS(e_prev) vanishes. */
int o3d_warm_features(const float *r, const float *eprev, const float *e, int nx, int ny, int nz,
                      double dx, float *inp, float *tgt, double *scale);

#ifdef __cplusplus
}
#endif
#endif
