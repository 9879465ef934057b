// Learned warm start: placeholder until the network kernels land.
#include "o3d_nn.h"

extern "C" int o3d_nn_warm_start(void *nn, void *solver, int itime, int *used,
                                 double *warm_dmax) {
  (void)nn; (void)solver; (void)itime; (void)warm_dmax;
  *used = 0;
  return o3d::set_error("learned warm start is not built");
}
