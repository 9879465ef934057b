/* Learned warm start of the Poisson solve: C ABI of the network in libo3d.so.
 *
 * No counterpart in the reference (it always starts SOR from the previous
 * step's pressure, src/integration.f90:247 with pp in/out).  The network is a
 * 3D U-Net-lite: 4 levels with 16/32/64/128 channels, two 3x3x3 convolutions
 * per level and side, average pooling, nearest up-sampling, additive skips,
 * leaky-ReLU, a 1x1x1 FP32 head whose output is added to the second input
 * plane (the network corrects that plane).  Weights and activations are BF16, products
 * accumulate in FP32 on MFMA (implicit-GEMM convolutions); FP32 master
 * weights are trained with Adam.  Convolutions pad periodically.
 *
 * Fields are float32 planes in the solver's layout (nx fastest); every grid
 * axis must be a multiple of 8.  Pointers are device memory.
 */
#ifndef O3D_NN_H
#define O3D_NN_H
#ifdef __cplusplus
extern "C" {
#endif

int o3d_nn_create(int nx, int ny, int nz, unsigned long long seed, void **nn);
int o3d_nn_destroy(void *nn);
/* Re-size the activation buffers for another grid (weights are kept). */
int o3d_nn_resize(void *nn, int nx, int ny, int nz);
int o3d_nn_num_params(void *nn, long long *n);
/* FP32 master weights / gradients, flat (host or device float array). */
int o3d_nn_get_params(void *nn, float *out);
int o3d_nn_set_params(void *nn, const float *in);
int o3d_nn_get_grads(void *nn, float *out);
int o3d_nn_zero_grad(void *nn);
/* in: 2 float32 planes (channel 0 then channel 1), out: 1 float32 plane. */
int o3d_nn_forward(void *nn, const float *in, float *out);
/* Back-propagates d(loss)/d(out) through the last forward; accumulates the
 * weight gradients.  The input of that forward must still hold its values. */
int o3d_nn_backward(void *nn, const float *dout);
/* Mean squared error of out against target (FP64 sum), writing
 * dout = scale * 2 (out - target) / n.  loss (host) receives the mean. */
int o3d_nn_mse(void *nn, const float *out, const float *target, float scale, float *dout,
               double *loss);
int o3d_nn_adam_step(void *nn, float lr, float beta1, float beta2, float eps, int t);
/* Time of the last forward's convolutions (ms, HIP events) and their FLOPs. */
int o3d_nn_last_timing(void *nn, double *ms, double *flops);

#ifdef __cplusplus
}
#endif
#endif
