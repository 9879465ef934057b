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

