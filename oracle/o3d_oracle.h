/* Test infrastructure only: a CPU restatement of the reference osinco3d
 * time-step path (the oracle), used by the tests, smoke() and the CPU
 * baseline.  It is never linked into, called by or shipped with the product.
 *
 * Arrays are float64, column-major (x fastest), 0-based: f[i + nx*(j + ny*k)].
 * bc[axis]: 0 = periodic, 1 = free-slip (src/initialization.f90:76-78).
 * kind (derivative variant): 0 = periodic "_00", 1 = free-slip even "p_11",
 * 2 = free-slip odd "i_11" (src/derivation.f90).
 */
#ifndef O3D_ORACLE_H
#define O3D_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

void orc_der1(double *df, const double *f, int nx, int ny, int nz, int axis,
              int kind, double h);
void orc_der2(double *df, const double *f, int nx, int ny, int nz, int axis,
              int kind, double h);
void orc_divergence(double *divf, const double *fx, const double *fy,
                    const double *fz, double dx, double dy, double dz, int nx,
                    int ny, int nz, const int *bc, int odd);
void orc_nu_t(double *nu_t, const double *ux, const double *uy,
              const double *uz, double dx, double dy, double dz, double cs,
              double delta, int nx, int ny, int nz, const int *bc);
int orc_predict(double *ux_pred, double *uy_pred, double *uz_pred,
                const double *ux, const double *uy, const double *uz,
                double *fux, double *fuy, double *fuz, double re,
                const double *adt, const double *bdt, const double *cdt,
                int itime, int itscheme, double dx, double dy, double dz,
                int nx, int ny, int nz, int iles, double cs, double delta,
                double *nu_t, const int *bc);
/* solver: 0 -> "0000", 1 -> "0011", 2 -> "111111".  Returns the reference's
 * printed exit iteration (kmax+1 when the loop ran out); *dmax_out = last dmax;
 * *reason = 0 eps, 1 stagnation, 2 kmax. */
int orc_sor_lex(double *pp, const double *rhs, double dx, double dy, double dz,
                int nx, int ny, int nz, int solver, double *omega, double eps,
                int kmax, int idyn, double *dmax_out, int *reason);
/* Red-black ordering of the same update (colour (i+j+k)&1, colour 0 first),
 * with the reference's stop rules; restates the GPU algorithm for checking. */
int orc_sor_rb(double *pp, const double *rhs, double dx, double dy, double dz,
               int nx, int ny, int nz, int solver, double *omega, double eps,
               int kmax, int idyn, double *dmax_out, int *reason,
               double *dmax_log, int log_len);
int orc_correct_velocity(double *ux, double *uy, double *uz,
                         const double *ux_pred, const double *uy_pred,
                         const double *uz_pred, const double *pp, double dt,
                         double dx, double dy, double dz, int nx, int ny,
                         int nz, const int *bc);
void orc_function_stats(const double *f, int nx, int ny, int nz, double *st);
/* out[0..16] = t, e_k, eps, eps2, dzeta, <u^2>,<v^2>,<w^2>, then the nine
 * mean-square gradients, as statistics_calc writes them. */
void orc_statistics(const double *ux, const double *uy, const double *uz,
                    int nx, int ny, int nz, double dx, double dy, double dz,
                    double re, double t, const int *bc, double *out);
void orc_residual(double *r, const double *pp, const double *rhs, double dx,
                  double dy, double dz, int nx, int ny, int nz, int solver);
/* x-Dirichlet inflow / outflow (no reference counterpart).  kind 3 of
 * orc_der1 / orc_der2 = ends by cubic extrapolation; solver 3..6 of the SOR
 * routines = x mirrored, y / z periodic unless bit 0 / 1 of solver - 3. */
unsigned long long orc_splitmix64(unsigned long long seed, long long counter);
double orc_noise_uniform(unsigned long long seed, long long counter);
void orc_jet_edges(double *uxp, double *uyp, double *uzp, const double *ux,
                   const double *uy, const double *uz, const double *base,
                   const double *w, double amp, double cph, double sph,
                   unsigned long long seed, double cu, int nx, int ny, int nz,
                   const int *bc, double *flux);
double orc_weighted_mean(const double *f, int nx, int ny, int nz, const int *mirror);

#ifdef __cplusplus
}
#endif
void orc_rotational(double *rx, double *ry, double *rz, const double *ux,
                    const double *uy, const double *uz, double dx, double dy,
                    double dz, int nx, int ny, int nz, const int *bc);
void orc_q_criterion(double *Q, const double *ux, const double *uy,
                     const double *uz, double dx, double dy, double dz, int nx,
                     int ny, int nz, const int *bc);
int orc_transeq(double *phi, const double *ux, const double *uy,
                const double *uz, const double *src, double *fphi, double re,
                double sc, const double *adt, const double *bdt,
                const double *cdt, int itime, int itscheme, double dx,
                double dy, double dz, int nx, int ny, int nz, int iles,
                const double *nu_t, const int *bc);
#endif
