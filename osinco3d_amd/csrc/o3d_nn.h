// Learned warm start of the Poisson solve (o3d_nn.hip) and the solver
// accessors it uses (defined in o3d_api.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace o3d {
double *solver_pressure(void *s);
double *solver_pressure_other(void *s);
void solver_swap_pressure(void *s);
const void *solver_rhs(void *s);
bool solver_mixed(void *s);
void solver_dims(void *s, int *nx, int *ny, int *nz);
int solver_kind(void *s);
void solver_sor_const(void *s, double *c4);
double solver_omega(void *s);
double solver_eps(void *s);
hipStream_t current_stream();
int set_error(const char *msg);
}  // namespace o3d

// Replace the solver's initial pressure by the learned guess when its first
// red-black sweep moves the field less than the previous pressure's does.
// used = 1 when accepted; warm_dmax[0] / [1] = first-sweep dmax of the
// previous pressure / of the learned guess.
extern "C" int o3d_nn_warm_start(void *nn, void *solver, int itime, int *used, double *warm_dmax);
