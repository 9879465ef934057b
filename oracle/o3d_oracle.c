/* Test infrastructure only: CPU restatement of the reference osinco3d
 * time-step path (the oracle).  Nothing in the product links or calls it;
 * the tests, smoke() and bench.py's cpu_baseline leg load it via ctypes.
 *
 * Every expression keeps the reference's operation order, so that without
 * FMA contraction (x86-64 baseline, gcc -O2) results are bit-identical to the
 * reference compiled with amdflang -O3.  Citations are to the reference tree.
 *
 * Pinned by tests/test_oracle.py against: the reference itself (oracle/_ref,
 * every function here except orc_sor_rb and orc_residual, which restate the
 * build's own red-black ordering and residual), the golden fixtures under
 * tests/golden/, the published t = 0 statistics row
 * (examples/tgv_re1600_dns/tgv_stats_re1600_dns.dat:18) and closed forms
 * (modified wavenumbers, discrete Laplacian eigenfunctions).
 */
#include "o3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j, k) ((size_t)(i) + (size_t)nx * ((size_t)(j) + (size_t)ny * (size_t)(k)))

/* x-Dirichlet ends (kind 3; no counterpart in the reference, which stops
 * on this boundary type, src/initialization.f90:238-241): a ghost m nodes
 * beyond an end (m = 1, 2, 3) is the cubic through the four nodes nearest
 * that end, evaluated there: f(-m) = XW[m-1] . (f(0), f(1), f(2), f(3)),
 * summed in that order, and mirrored at the far end.  Every stencil that
 * reaches past the end is then exact on cubics. */
static const double XW[3][4] = {{4.0, -6.0, 4.0, -1.0},
                                {10.0, -20.0, 15.0, -4.0},
                                {20.0, -45.0, 36.0, -10.0}};

/* Neighbour index along an axis of n nodes, with the ghost value's sign.
 * Periodic: index n+1 wraps to 1 (src/derivation.f90:34-56), so the period
 * is n nodes.  Free-slip: mirror about the end node, ghost f(1-m) = +-f(1+m)
 * and f(n+m) = +-f(n-m) (boundary rows of src/derivation.f90:62-163). */
static inline int ghost(int q, int n, int kind, double *sgn) {
    *sgn = 1.0;
    if (kind == 0) {
        q %= n;
        if (q < 0) q += n;
        return q;
    }
    if (q < 0) {
        q = -q;
        if (kind == 2) *sgn = -1.0;
    } else if (q > n - 1) {
        q = 2 * (n - 1) - q;
        if (kind == 2) *sgn = -1.0;
    }
    return q;
}

static inline double at(const double *f, int nx, int ny, int nz, int axis,
                        int i, int j, int k, int s, int kind) {
    double sg;
    (void)nz;
    if (kind == 3) {
        const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
        const int p = (axis == 0 ? i : (axis == 1 ? j : k)) + s;
        const size_t st = axis == 0 ? 1 : (axis == 1 ? (size_t)nx : (size_t)nx * ny);
        const size_t base = IDX(i, j, k) - (size_t)(axis == 0 ? i : (axis == 1 ? j : k)) * st;
        if (p >= 0 && p < n) return f[base + (size_t)p * st];
        const int m = p < 0 ? -p : p - (n - 1);
        const size_t e = p < 0 ? base : base + (size_t)(n - 1) * st;
        const long d = p < 0 ? (long)st : -(long)st;
        const double *w = XW[m - 1];
        return ((w[0] * f[e] + w[1] * f[e + d]) + w[2] * f[e + 2 * d]) + w[3] * f[e + 3 * d];
    }
    if (axis == 0) { int q = ghost(i + s, nx, kind, &sg); return sg == 1.0 ? f[IDX(q, j, k)] : -f[IDX(q, j, k)]; }
    if (axis == 1) { int q = ghost(j + s, ny, kind, &sg); return sg == 1.0 ? f[IDX(i, q, k)] : -f[IDX(i, q, k)]; }
    { int q = ghost(k + s, nz, kind, &sg); return sg == 1.0 ? f[IDX(i, j, q)] : -f[IDX(i, j, q)]; }
}

/* 6th-order first derivative: src/derivation.f90:6-479.  Coefficients as
 * precomputed at :26-30; even free-slip forces df = 0 on the end nodes
 * (:87,105). */
void orc_der1(double *df, const double *f, int nx, int ny, int nz, int axis,
              int kind, double h) {
    const double sixty = 60.0 * h;
    const double a = 1.0 / sixty, b = 9.0 / sixty, c = 45.0 / sixty;
    const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int p = axis == 0 ? i : (axis == 1 ? j : k);
                if (kind == 1 && (p == 0 || p == n - 1)) { df[IDX(i, j, k)] = 0.0; continue; }
                const double f3p = at(f, nx, ny, nz, axis, i, j, k, 3, kind);
                const double f3m = at(f, nx, ny, nz, axis, i, j, k, -3, kind);
                const double f2p = at(f, nx, ny, nz, axis, i, j, k, 2, kind);
                const double f2m = at(f, nx, ny, nz, axis, i, j, k, -2, kind);
                const double f1p = at(f, nx, ny, nz, axis, i, j, k, 1, kind);
                const double f1m = at(f, nx, ny, nz, axis, i, j, k, -1, kind);
                df[IDX(i, j, k)] = a * (f3p - f3m) - b * (f2p - f2m) + c * (f1p - f1m);
            }
}

/* 4th-order second derivative: src/derivation.f90:497-932, coefficients
 * at :517-521. */
void orc_der2(double *df, const double *f, int nx, int ny, int nz, int axis,
              int kind, double h) {
    const double t = 12.0 * h * h;
    const double a = 1.0 / t, b = 16.0 / t, c = 30.0 / t;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double f2m = at(f, nx, ny, nz, axis, i, j, k, -2, kind);
                const double f2p = at(f, nx, ny, nz, axis, i, j, k, 2, kind);
                const double f1m = at(f, nx, ny, nz, axis, i, j, k, -1, kind);
                const double f1p = at(f, nx, ny, nz, axis, i, j, k, 1, kind);
                df[IDX(i, j, k)] = -a * (f2m + f2p) + b * (f1m + f1p) - c * f[IDX(i, j, k)];
            }
}

/* Variant bound for a field that is odd along `axis` (the "i" pointers) or
 * even (the "p" pointers): schemes(), src/initialization.f90:226-282.  A
 * velocity on an x-Dirichlet axis (bc 2) uses the extrapolated ends. */
static inline int kind_of(const int *bc, int axis, int odd) {
    if (bc[axis] == 2) return 3;
    return bc[axis] == 0 ? 0 : (odd ? 2 : 1);
}
/* The pressure's variant: the p pointers (src/integration.f90:298-300); on an
 * x-Dirichlet axis the pressure is mirrored (homogeneous Neumann), its
 * derivative zero on the end nodes. */
static inline int kind_p(const int *bc, int axis) {
    return bc[axis] == 0 ? 0 : 1;
}

/* src/differential_operators.f90:7-38 */
void orc_divergence(double *divf, const double *fx, const double *fy,
                    const double *fz, double dx, double dy, double dz, int nx,
                    int ny, int nz, const int *bc, int odd) {
    const size_t N = (size_t)nx * ny * nz;
    double *a = malloc(N * sizeof(double)), *b = malloc(N * sizeof(double)),
           *c = malloc(N * sizeof(double));
    orc_der1(a, fx, nx, ny, nz, 0, kind_of(bc, 0, odd != 0), dx);
    orc_der1(b, fy, nx, ny, nz, 1, kind_of(bc, 1, odd != 0), dy);
    orc_der1(c, fz, nx, ny, nz, 2, kind_of(bc, 2, odd != 0), dz);
    for (size_t q = 0; q < N; ++q) divf[q] = a[q] + b[q] + c[q];
    free(a); free(b); free(c);
}

/* The nine first derivatives with the parities predict_velocity,
 * calculate_nu_t and statistics_calc use: u_c is odd along axis c.
 * g[3*c + axis]. */
static void grad9(double **g, const double *ux, const double *uy,
                  const double *uz, double dx, double dy, double dz, int nx,
                  int ny, int nz, const int *bc) {
    const double *u[3] = {ux, uy, uz};
    const double h[3] = {dx, dy, dz};
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a)
            orc_der1(g[3 * c + a], u[c], nx, ny, nz, a, kind_of(bc, a, a == c), h[a]);
}

/* rotational, src/differential_operators.f90:40-76: curl from the p (even)
 * variants (the off-diagonal entries of grad9). */
void orc_rotational(double *rx, double *ry, double *rz, const double *ux,
                    const double *uy, const double *uz, double dx, double dy,
                    double dz, int nx, int ny, int nz, const int *bc) {
    const size_t N = (size_t)nx * ny * nz;
    double *g[9];
    for (int q = 0; q < 9; ++q) g[q] = malloc(N * sizeof(double));
    grad9(g, ux, uy, uz, dx, dy, dz, nx, ny, nz, bc);
    for (size_t q = 0; q < N; ++q) {
        rx[q] = g[7][q] - g[5][q]; /* duzdy - duydz (:64-66) */
        ry[q] = g[2][q] - g[6][q]; /* duxdz - duzdx (:68-70) */
        rz[q] = g[3][q] - g[1][q]; /* duydx - duxdy (:72-74) */
    }
    for (int q = 0; q < 9; ++q) free(g[q]);
}

/* calculate_Q_criterion, src/differential_operators.f90:78-108: diagonal
 * terms from the i (odd) variants, the others from the p variants;
 * expression order of :103-104. */
void orc_q_criterion(double *Q, const double *ux, const double *uy,
                     const double *uz, double dx, double dy, double dz, int nx,
                     int ny, int nz, const int *bc) {
    const size_t N = (size_t)nx * ny * nz;
    double *g[9];
    for (int q = 0; q < 9; ++q) g[q] = malloc(N * sizeof(double));
    grad9(g, ux, uy, uz, dx, dy, dz, nx, ny, nz, bc);
    for (size_t q = 0; q < N; ++q) {
        const double a = g[0][q], b = g[4][q], c = g[8][q];
        Q[q] = -(0.5 * (a * a + b * b + c * c)) - g[1][q] * g[3][q] -
               g[2][q] * g[6][q] - g[5][q] * g[7][q];
    }
    for (int q = 0; q < 9; ++q) free(g[q]);
}

static void lap9(double **g, const double *ux, const double *uy,
                 const double *uz, double dx, double dy, double dz, int nx,
                 int ny, int nz, const int *bc) {
    const double *u[3] = {ux, uy, uz};
    const double h[3] = {dx, dy, dz};
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a)
            orc_der2(g[3 * c + a], u[c], nx, ny, nz, a, kind_of(bc, a, a == c), h[a]);
}

static double **alloc9(size_t N) {
    double **g = malloc(9 * sizeof(double *));
    for (int q = 0; q < 9; ++q) g[q] = malloc(N * sizeof(double));
    return g;
}
static void free9(double **g) {
    for (int q = 0; q < 9; ++q) free(g[q]);
    free(g);
}

/* src/les_turbulence.f90:10-97 (without the statistics print). */
void orc_nu_t(double *nu_t, const double *ux, const double *uy,
              const double *uz, double dx, double dy, double dz, double cs,
              double delta, int nx, int ny, int nz, const int *bc) {
    const size_t N = (size_t)nx * ny * nz;
    double **g = alloc9(N);
    grad9(g, ux, uy, uz, dx, dy, dz, nx, ny, nz, bc);
    const double csd = cs * delta;
    const double csd2 = csd * csd;
    for (size_t q = 0; q < N; ++q) {
        const double s11 = g[0][q], s22 = g[4][q], s33 = g[8][q];
        const double s12 = 0.5 * (g[1][q] + g[3][q]);
        const double s13 = 0.5 * (g[2][q] + g[6][q]);
        const double s23 = 0.5 * (g[5][q] + g[7][q]);
        const double smag = sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 +
                                        2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
        nu_t[q] = csd2 * smag;
    }
    free9(g);
}

/* src/integration.f90:14-197.  f?? hold 3 history slots, slot 1 = newest:
 * f[s*N + q]. Returns 0, or 1 for an unrecognized itscheme (reference: stop). */
int orc_predict(double *ux_pred, double *uy_pred, double *uz_pred,
                const double *ux, const double *uy, const double *uz,
                double *fux, double *fuy, double *fuz, double re,
                const double *adt, const double *bdt, const double *cdt,
                int itime, int itscheme, double dx, double dy, double dz,
                int nx, int ny, int nz, int iles, double cs, double delta,
                double *nu_t, const int *bc) {
    const size_t N = (size_t)nx * ny * nz;
    double adu, bdu, cdu;
    if (itscheme == 1 || itime == 1) { adu = adt[0]; bdu = bdt[0]; cdu = cdt[0]; }
    else if (itscheme == 2 || itime == 2) { adu = adt[1]; bdu = bdt[1]; cdu = cdt[1]; }
    else if (itscheme == 3) { adu = adt[2]; bdu = bdt[2]; cdu = cdt[2]; }
    else return 1;
    const double onere = 1.0 / re;
    if (iles == 1) orc_nu_t(nu_t, ux, uy, uz, dx, dy, dz, cs, delta, nx, ny, nz, bc);
    else memset(nu_t, 0, N * sizeof(double));
    double **g = alloc9(N), **l = alloc9(N);
    grad9(g, ux, uy, uz, dx, dy, dz, nx, ny, nz, bc);
    lap9(l, ux, uy, uz, dx, dy, dz, nx, ny, nz, bc);
    double *fu[3] = {fux, fuy, fuz};
    double *up[3] = {ux_pred, uy_pred, uz_pred};
    const double *u[3] = {ux, uy, uz};
    for (int c = 0; c < 3; ++c) {
        double *F = fu[c];
        for (size_t q = 0; q < N; ++q) {
            const double nu_eff = onere + nu_t[q];
            F[q] = nu_eff * (l[3 * c][q] + l[3 * c + 1][q] + l[3 * c + 2][q]) -
                   (ux[q] * g[3 * c][q] + uy[q] * g[3 * c + 1][q] + uz[q] * g[3 * c + 2][q]);
            up[c][q] = u[c][q] + adu * F[q] + bdu * F[N + q] + cdu * F[2 * N + q];
        }
    }
    for (int c = 0; c < 3; ++c) {
        double *F = fu[c];
        if (itscheme == 2) {
            memcpy(F + N, F, N * sizeof(double));
        } else if (itscheme == 3) {
            memcpy(F + 2 * N, F + N, N * sizeof(double));
            memcpy(F + N, F, N * sizeof(double));
        }
    }
    free9(g); free9(l);
    return 0;
}

/* transeq, src/integration.f90:332-468: scalar transport with the p (even)
 * variants (:415-421), alpha_eff (:405-411), F1 and the AB update
 * (:422-426), conservative clipping (:428-450, sums in index order) and the
 * history shift (:453-460).  fphi: 3 planes, slot 1 first.  A zero weight
 * sum leaves the clipped field unchanged (the reference divides 0 by 0). */
int orc_transeq(double *phi, const double *ux, const double *uy,
                const double *uz, const double *src, double *fphi, double re,
                double sc, const double *adt, const double *bdt,
                const double *cdt, int itime, int itscheme, double dx,
                double dy, double dz, int nx, int ny, int nz, int iles,
                const double *nu_t, const int *bc) {
    const size_t N = (size_t)nx * ny * nz;
    double adu, bdu, cdu;
    if (itscheme == 1 || itime == 1) { adu = adt[0]; bdu = bdt[0]; cdu = cdt[0]; }
    else if (itscheme == 2 || itime == 2) { adu = adt[1]; bdu = bdt[1]; cdu = cdt[1]; }
    else if (itscheme == 3) { adu = adt[2]; bdu = bdt[2]; cdu = cdt[2]; }
    else return 1;
    const double h[3] = {dx, dy, dz};
    double *g[3], *l[3];
    for (int a = 0; a < 3; ++a) {
        g[a] = malloc(N * sizeof(double));
        l[a] = malloc(N * sizeof(double));
        orc_der1(g[a], phi, nx, ny, nz, a, kind_of(bc, a, 0), h[a]);
        orc_der2(l[a], phi, nx, ny, nz, a, kind_of(bc, a, 0), h[a]);
    }
    const double alpha0 = 1.0 / (re * sc);
    for (size_t q = 0; q < N; ++q) {
        const double alpha = iles == 1 ? alpha0 + nu_t[q] / sc : alpha0;
        const double F = alpha * (l[0][q] + l[1][q] + l[2][q]) -
                         (ux[q] * g[0][q] + uy[q] * g[1][q] + uz[q] * g[2][q]) + src[q];
        fphi[q] = F;
        phi[q] = phi[q] + adu * F + bdu * fphi[N + q] + cdu * fphi[2 * N + q];
    }
    double s_old = 0.0, s_new = 0.0, s_w = 0.0;
    for (size_t q = 0; q < N; ++q) s_old += phi[q];
    for (size_t q = 0; q < N; ++q) {
        phi[q] = fmax(0.0, fmin(1.0, phi[q]));
        s_new += phi[q];
    }
    const double excess = s_old / (double)N - s_new / (double)N;
    for (size_t q = 0; q < N; ++q) s_w += fmin(phi[q], 1.0 - phi[q]);
    if (s_w != 0.0)
        for (size_t q = 0; q < N; ++q) {
            const double w = fmin(phi[q], 1.0 - phi[q]) / s_w;
            phi[q] = fmax(0.0, fmin(1.0, phi[q] + excess * w));
        }
    if (itscheme == 3) memcpy(fphi + 2 * N, fphi + N, N * sizeof(double));
    if (itscheme >= 2) memcpy(fphi + N, fphi, N * sizeof(double));
    for (int a = 0; a < 3; ++a) { free(g[a]); free(l[a]); }
    return 0;
}

/* Neighbour rules of the three SOR solvers (src/poisson.f90:57-92,
 * :186-218, :310-344): 0000 all periodic, 0011 free-slip in y only,
 * 111111 free-slip in x, y and z.  Free-slip: m1 = p1 = 2 at node 1 and
 * m1 = p1 = n-1 at node n. */
static inline void nbr(int q, int n, int periodic, int *m, int *p) {
    if (q == 0) { *m = periodic ? n - 1 : 1; *p = 1; }
    else if (q == n - 1) { *m = n - 2; *p = periodic ? 0 : n - 2; }
    else { *m = q - 1; *p = q + 1; }
}
static inline void solver_bc(int solver, int *px, int *py, int *pz) {
    if (solver >= 3) {
        /* x-Dirichlet velocity (no reference counterpart): the pressure is
         * mirrored in x; y and z periodic unless bit 0 / bit 1 of
         * solver - 3 says mirrored. */
        *px = 0; *py = !((solver - 3) & 1); *pz = !((solver - 3) & 2);
        return;
    }
    *px = solver != 2; *py = solver == 0; *pz = solver != 2;
}

struct sor_const { double ox, oy, oz, A; };
static struct sor_const sor_setup(double dx, double dy, double dz) {
    /* src/poisson.f90:42-51 */
    struct sor_const s;
    const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
    s.ox = 1.0 / dx2; s.oy = 1.0 / dy2; s.oz = 1.0 / dz2;
    const double tx = 2.0 * s.ox, ty = 2.0 * s.oy, tz = 2.0 * s.oz;
    s.A = -(tx + ty + tz);
    return s;
}

static inline double sor_point(const double *pp, const double *rhs, int nx,
                               int ny, int nz, int i, int j, int k, int px,
                               int py, int pz, const struct sor_const *s) {
    int im, ip, jm, jp, km, kp;
    nbr(i, nx, px, &im, &ip);
    nbr(j, ny, py, &jm, &jp);
    nbr(k, nz, pz, &km, &kp);
    /* src/poisson.f90:95-98 */
    return (-s->ox * (pp[IDX(im, j, k)] + pp[IDX(ip, j, k)]) -
            s->oy * (pp[IDX(i, jm, k)] + pp[IDX(i, jp, k)]) -
            s->oz * (pp[IDX(i, j, km)] + pp[IDX(i, j, kp)]) + rhs[IDX(i, j, k)]) / s->A;
}

/* Stop and omega rules after a sweep: src/poisson.f90:110-122.
 * Returns 1 to stop (reason set). */
static int sor_after(int iter, double dmax, double *dmax_old, double *omega,
                     double eps, int idyn, int solver, int *reason) {
    const double factor = solver == 1 ? 1.01 : 1.05;
    if (dmax < eps) { *reason = 0; return 1; }
    if (fabs(*dmax_old - dmax) < eps / 1000.0) { *reason = 1; return 1; }
    if (iter > 1 && idyn == 1) {
        if (dmax > *dmax_old) *omega = *omega * (2.0 - factor);
        else if (dmax < 0.1 * *dmax_old) *omega = fmin(*omega * factor, 2.0);
    }
    *dmax_old = dmax;
    return 0;
}

/* Lexicographic SOR, src/poisson.f90:6-130 / 132-255 / 257-381. */
int orc_sor_lex(double *pp, const double *rhs, double dx, double dy, double dz,
                int nx, int ny, int nz, int solver, double *omega, double eps,
                int kmax, int idyn, double *dmax_out, int *reason) {
    int px, py, pz;
    solver_bc(solver, &px, &py, &pz);
    const struct sor_const s = sor_setup(dx, dy, dz);
    double dmax_old = 1609.0, dmax = 0.0;
    int iter;
    *reason = 2;
    for (iter = 1; iter <= kmax; ++iter) {
        dmax = 0.0;
        const double om = *omega;
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const double pn = sor_point(pp, rhs, nx, ny, nz, i, j, k, px, py, pz, &s);
                    const double d = fabs(pn - pp[IDX(i, j, k)]);
                    dmax = d > dmax ? d : dmax;
                    pp[IDX(i, j, k)] = (1.0 - om) * pp[IDX(i, j, k)] + om * pn;
                }
        if (sor_after(iter, dmax, &dmax_old, omega, eps, idyn, solver, reason)) break;
    }
    *dmax_out = dmax;
    return iter;
}

/* Red-black ordering of the same point update and stop rules: colour
 * (i+j+k)&1, colour 0 swept first, dmax over both colours. */
int orc_sor_rb(double *pp, const double *rhs, double dx, double dy, double dz,
               int nx, int ny, int nz, int solver, double *omega, double eps,
               int kmax, int idyn, double *dmax_out, int *reason,
               double *dmax_log, int log_len) {
    int px, py, pz;
    solver_bc(solver, &px, &py, &pz);
    const struct sor_const s = sor_setup(dx, dy, dz);
    double dmax_old = 1609.0, dmax = 0.0;
    int iter;
    *reason = 2;
    for (iter = 1; iter <= kmax; ++iter) {
        dmax = 0.0;
        const double om = *omega;
        for (int colour = 0; colour < 2; ++colour)
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j)
                    for (int i = (j + k + colour) & 1; i < nx; i += 2) {
                        const double pn = sor_point(pp, rhs, nx, ny, nz, i, j, k, px, py, pz, &s);
                        const double d = fabs(pn - pp[IDX(i, j, k)]);
                        dmax = d > dmax ? d : dmax;
                        pp[IDX(i, j, k)] = (1.0 - om) * pp[IDX(i, j, k)] + om * pn;
                    }
        if (dmax_log && iter <= log_len) dmax_log[iter - 1] = dmax;
        if (sor_after(iter, dmax, &dmax_old, omega, eps, idyn, solver, reason)) break;
    }
    *dmax_out = dmax;
    return iter;
}

/* r = rhs - L p with the solver's 7-point Laplacian and neighbour rules. */
void orc_residual(double *r, const double *pp, const double *rhs, double dx,
                  double dy, double dz, int nx, int ny, int nz, int solver) {
    int px, py, pz;
    solver_bc(solver, &px, &py, &pz);
    const struct sor_const s = sor_setup(dx, dy, dz);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int im, ip, jm, jp, km, kp;
                nbr(i, nx, px, &im, &ip);
                nbr(j, ny, py, &jm, &jp);
                nbr(k, nz, pz, &km, &kp);
                const double lap = s.ox * (pp[IDX(im, j, k)] + pp[IDX(ip, j, k)]) +
                                   s.oy * (pp[IDX(i, jm, k)] + pp[IDX(i, jp, k)]) +
                                   s.oz * (pp[IDX(i, j, km)] + pp[IDX(i, j, kp)]) +
                                   s.A * pp[IDX(i, j, k)];
                r[IDX(i, j, k)] = rhs[IDX(i, j, k)] - lap;
            }
}

/* src/functions.f90:157-192 */
static int has_nan(const double *f, size_t N) {
    for (size_t q = 0; q < N; ++q)
        if (isnan(f[q])) return 1;
    return 0;
}

/* src/integration.f90:257-330.  Returns 1 when the reference would print
 * "Velocity has diverged" and stop. */
int orc_correct_velocity(double *ux, double *uy, double *uz,
                         const double *ux_pred, const double *uy_pred,
                         const double *uz_pred, const double *pp, double dt,
                         double dx, double dy, double dz, int nx, int ny,
                         int nz, const int *bc) {
    const size_t N = (size_t)nx * ny * nz;
    double *g = malloc(N * sizeof(double));
    double *u[3] = {ux, uy, uz};
    const double *up[3] = {ux_pred, uy_pred, uz_pred};
    const double h[3] = {dx, dy, dz};
    for (int a = 0; a < 3; ++a) {
        orc_der1(g, pp, nx, ny, nz, a, kind_p(bc, a), h[a]);
        for (size_t q = 0; q < N; ++q) u[a][q] = up[a][q] - dt * g[q];
    }
    free(g);
    if (bc[0] == 2)
        /* x-Dirichlet: the inflow and outflow planes keep u* (set by
         * orc_jet_edges), no pressure correction there. */
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j) {
                    u[a][IDX(0, j, k)] = up[a][IDX(0, j, k)];
                    u[a][IDX(nx - 1, j, k)] = up[a][IDX(nx - 1, j, k)];
                }
    for (int a = 0; a < 3; ++a) {
        double mx = -INFINITY;
        for (size_t q = 0; q < N; ++q) mx = u[a][q] > mx ? u[a][q] : mx;
        if (has_nan(u[a], N) || mx > 1000.0) return 1;
    }
    return 0;
}

/* src/functions.f90:27-63: min, max, mean, first argmax (1-based). */
void orc_function_stats(const double *f, int nx, int ny, int nz, double *st) {
    double sum = 0.0, fmn = 1.79769313486231570815e+308, fmx = -1.79769313486231570815e+308;
    int im = 1, jm = 1, km = 1;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double v = f[IDX(i, j, k)];
                sum = sum + v;
                if (v < fmn) fmn = v;
                if (v > fmx) { fmx = v; im = i + 1; jm = j + 1; km = k + 1; }
            }
    st[0] = fmn; st[1] = fmx;
    st[2] = sum / (double)(nx * ny * nz);
    st[3] = im; st[4] = jm; st[5] = km;
}

/* average_3d_array, src/functions.f90:123-150: i outermost, k innermost. */
static double avg3(const double *t, int nx, int ny, int nz, int square) {
    double sum = 0.0;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const double v = t[IDX(i, j, k)];
                sum = sum + (square ? v * v : v);
            }
    return sum / (double)(nx * ny * nz);
}

/* src/utils.f90:243-375 */
void orc_statistics(const double *ux, const double *uy, const double *uz,
                    int nx, int ny, int nz, double dx, double dy, double dz,
                    double re, double t, const int *bc, double *out) {
    const size_t N = (size_t)nx * ny * nz;
    double **g = alloc9(N);
    out[0] = t;
    out[5] = avg3(ux, nx, ny, nz, 1);
    out[6] = avg3(uy, nx, ny, nz, 1);
    out[7] = avg3(uz, nx, ny, nz, 1);
    grad9(g, ux, uy, uz, dx, dy, dz, nx, ny, nz, bc);
    for (int q = 0; q < 9; ++q) out[8 + q] = avg3(g[q], nx, ny, nz, 1);
    const double xnu = 1.0 / re;
    double dzeta = 0.0, eps = 0.0, e_k = 0.0;
    for (size_t q = 0; q < N; ++q) {
        e_k = e_k + 0.5 * (ux[q] * ux[q] + uy[q] * uy[q] + uz[q] * uz[q]);
        const double a = 2.0 * g[0][q], b = 2.0 * g[4][q], c = 2.0 * g[8][q];
        const double s12 = g[1][q] + g[3][q], s13 = g[2][q] + g[6][q], s23 = g[5][q] + g[7][q];
        eps = eps + 0.5 * xnu * (a * a + b * b + c * c + 2.0 * (s12 * s12) +
                                 2.0 * (s13 * s13) + 2.0 * (s23 * s23));
        const double w1 = g[7][q] - g[5][q], w2 = g[2][q] - g[6][q], w3 = g[3][q] - g[1][q];
        dzeta = dzeta + 0.5 * (w1 * w1 + w2 * w2 + w3 * w3);
    }
    const double nn = (double)(nx * ny * nz);
    out[1] = e_k / nn;
    out[2] = eps / nn;
    out[4] = dzeta / nn;
    lap9(g, ux, uy, uz, dx, dy, dz, nx, ny, nz, bc);
    double eps2 = 0.0;
    for (size_t q = 0; q < N; ++q) {
        const double t1 = (-xnu) * (ux[q] * (g[0][q] + g[1][q] + g[2][q]) +
                                    uy[q] * (g[3][q] + g[4][q] + g[5][q]) +
                                    uz[q] * (g[6][q] + g[7][q] + g[8][q]));
        eps2 = eps2 + t1;
    }
    out[3] = eps2 / nn;
    free9(g);
}

/* ------------------------------------------------------------------------
 * x-Dirichlet inflow / outflow (no counterpart in the reference).
 * ---------------------------------------------------------------------- */

/* SplitMix64 (Steele, Lea, Flood 2014), output number counter + 1 of the
 * stream seeded with `seed` (state seed + (counter + 1) * golden gamma, then
 * the finalizer): the counter-based generator of the inflow noise, as a
 * uniform double in [-1, 1) from the top 53 bits. */
unsigned long long orc_splitmix64(unsigned long long seed, long long counter) {
    unsigned long long z = seed + (unsigned long long)(counter + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double orc_noise_uniform(unsigned long long seed, long long counter) {
    const unsigned long long h = orc_splitmix64(seed, counter);
    return (double)(h >> 11) * 0x1.0p-53 * 2.0 - 1.0;
}

/* Weight of a plane node in the flux sums: 1/2 on the end nodes of a
 * mirrored (non-periodic) axis, else 1 (the trapezoid rule the mirrored
 * 7-point operator's compatibility condition uses). */
static inline double wgt(int q, int n, int mirrored) {
    return (mirrored && (q == 0 || q == n - 1)) ? 0.5 : 1.0;
}

/* Weighted sum over the (y, z) plane of the values v[j + ny k], in the order
 * the device uses: per z row the sequential sum over j of wy * v, then the
 * sequential sum over k of wz * row. */
static double plane_sum(const double *v, int ny, int nz, int my, int mz) {
    double q = 0.0;
    for (int k = 0; k < nz; ++k) {
        double r = 0.0;
        for (int j = 0; j < ny; ++j) r = r + wgt(j, ny, my) * v[j + (size_t)ny * k];
        q = q + wgt(k, nz, mz) * r;
    }
    return q;
}

/* After predict (which leaves every node's u* from the interior formula):
 *  - inflow plane i = 0: u*_c = base_c(j, k) + (amp w(j, k)) (x1 C + x2 S),
 *    x1, x2 = orc_noise_uniform(seed, 2 (3 (j + ny k) + c) + 0 / 1),
 *    C = cos, S = sin of the inflow phase 2 pi t / T (given);
 *  - outflow plane i = nx-1: first-order upwind convection at Uc = the mean
 *    inflow velocity of the base profile: u*_c = u_c - cu (u_c - u_c(i-1)),
 *    cu = Uc dt / dx, from the velocity before the step;
 *  - then u*_x on the outflow plane is shifted by (Q_in - Q_out) / W so that
 *    the outflow flux equals the inflow flux (trapezoid-weighted plane sums,
 *    plane_sum order).
 * base: 3 planes of ny nz values; w: one plane.  flux[0..2] = Q_in, Q_out
 * before the shift, W. */
void orc_jet_edges(double *uxp, double *uyp, double *uzp, const double *ux,
                   const double *uy, const double *uz, const double *base,
                   const double *w, double amp, double cph, double sph,
                   unsigned long long seed, double cu, int nx, int ny, int nz,
                   const int *bc, double *flux) {
    double *up[3] = {uxp, uyp, uzp};
    const double *u[3] = {ux, uy, uz};
    const size_t P = (size_t)ny * nz;
    double *pin = malloc(P * sizeof(double)), *pout = malloc(P * sizeof(double));
    const int my = bc[1] != 0, mz = bc[2] != 0;
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) {
                const size_t q = j + (size_t)ny * k;
                const long long cnt = 2 * (3 * (long long)q + c);
                const double x1 = orc_noise_uniform(seed, cnt);
                const double x2 = orc_noise_uniform(seed, cnt + 1);
                const double vin = base[c * P + q] + (amp * w[q]) * (x1 * cph + x2 * sph);
                const double ub = u[c][IDX(nx - 1, j, k)], ub1 = u[c][IDX(nx - 2, j, k)];
                const double vout = ub - cu * (ub - ub1);
                up[c][IDX(0, j, k)] = vin;
                up[c][IDX(nx - 1, j, k)] = vout;
                if (c == 0) { pin[q] = vin; pout[q] = vout; }
            }
    for (size_t q = 0; q < P; ++q) pin[q] = 1.0;
    const double W = plane_sum(pin, ny, nz, my, mz);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j) pin[j + (size_t)ny * k] = up[0][IDX(0, j, k)];
    const double qin = plane_sum(pin, ny, nz, my, mz);
    const double qout = plane_sum(pout, ny, nz, my, mz);
    const double shift = (qin - qout) / W;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j) up[0][IDX(nx - 1, j, k)] = pout[j + (size_t)ny * k] + shift;
    flux[0] = qin; flux[1] = qout; flux[2] = W;
    free(pin); free(pout);
}

/* Weighted mean of a field with the trapezoid weights of the mirrored axes
 * (mirror[a] != 0), in the device's order: 1024 chunks of ceil(n / 1024)
 * cells; within a chunk 256 threads sum cells lo + t, lo + t + 256, ...;
 * each wave of 64 combines its lanes by the xor butterfly (offsets 32, 16,
 * 8, 4, 2, 1) and the chunk sum is ((0 + w0) + w1) + w2 + w3 over its four
 * waves; the 1024 chunk sums are combined the same way by 1024 threads (one
 * each, 16 waves).  Returns sum(w f) / sum(w). */
static double wave_butterfly(double *s) {
    double t[64];
    for (int o = 32; o > 0; o >>= 1) {
        for (int l = 0; l < 64; ++l) t[l] = s[l] + s[l ^ o];
        memcpy(s, t, sizeof(t));
    }
    return s[0];
}
static double block_tree(double *lanes, int nwaves) {
    double r = 0.0;
    for (int w = 0; w < nwaves; ++w) r = r + wave_butterfly(lanes + 64 * w);
    return r;
}
static double weighted_sum_dev(const double *f, int nx, int ny, int nz, const int *mirror,
                               int ones) {
    const long long n = (long long)nx * ny * nz;
    const int B = 1024, T = 256;
    const long long per = (n + B - 1) / B;
    double *parts = malloc(B * sizeof(double));
    double lanes[256];
    for (int b = 0; b < B; ++b) {
        const long long lo = (long long)b * per, hi = lo + per < n ? lo + per : n;
        for (int t = 0; t < T; ++t) {
            double acc = 0.0;
            for (long long q = lo + t; q < hi; q += T) {
                const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((long long)nx * ny));
                const double wq = (wgt(i, nx, mirror[0]) * wgt(j, ny, mirror[1])) * wgt(k, nz, mirror[2]);
                acc = acc + (ones ? wq : wq * f[q]);
            }
            lanes[t] = acc;
        }
        parts[b] = block_tree(lanes, 4);
    }
    double big[1024];
    memcpy(big, parts, sizeof(big));
    const double r = block_tree(big, 16);
    free(parts);
    return r;
}
double orc_weighted_mean(const double *f, int nx, int ny, int nz, const int *mirror) {
    return weighted_sum_dev(f, nx, ny, nz, mirror, 0) / weighted_sum_dev(f, nx, ny, nz, mirror, 1);
}
