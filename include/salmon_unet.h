/*
 * salmon_unet.h — learned Hartree+XC potential predictor for the predictor-corrector
 * step (no counterpart in the reference, whose predictor is a full Taylor step with
 * V(t), src/rt/time_evolution_step.f90:354-423).
 *
 * The network is a 3-level U-Net on the periodic real-space grid: 3x3x3 periodic
 * convolutions with ReLU, 16/32/64 channels, 2x2x2 average-pool down, nearest
 * upsampling, skip concatenation, a 1x1x1 head.  Weights and activations are
 * BF16, products accumulate in FP32 on MFMA (gfx950 v_mfma_f32_16x16x32_bf16), master
 * weights and the Adam state are FP32.  Grid extents must be multiples of 4.
 *
 * Tensors handed over are DEVICE arrays, channels-last: in[p][c] (c = 0..3) and
 * out[p], p = ix + nx*(iy + ny*iz), float32.  All work is enqueued on the
 * handle's stream.  Returns 0 or a SALMON_HIP_E* code (salmon_hip_last_error()).
 */
#ifndef SALMON_UNET_H
#define SALMON_UNET_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct salmon_unet *salmon_unet_t;

/* He-normal initial weights from `seed` (zero biases). */
int salmon_unet_create(int nx, int ny, int nz, unsigned long long seed, salmon_unet_t *out);
int salmon_unet_destroy(salmon_unet_t u);
int salmon_unet_set_stream(salmon_unet_t u, void *stream);

/* FP32 master parameters in a flat host array (layer by layer: weights [27][co][ci], bias [co];
   head weights [16], head bias). */
int salmon_unet_num_params(salmon_unet_t u, long *n);
int salmon_unet_get_params(salmon_unet_t u, float *host);
int salmon_unet_set_params(salmon_unet_t u, const float *host);

/* out[p] = net(in)[p]. */
int salmon_unet_forward(salmon_unet_t u, const float *in_dev, float *out_dev);

/*
 * One training sample: forward, mean-squared-error loss against target[p], backward;
 * parameter gradients accumulate over calls.  loss_dev (may be null) receives the
 * sample's loss.  salmon_unet_adam_step applies one Adam update (beta1 0.9,
 * beta2 0.999, eps 1e-8) with the accumulated gradients divided by `nsamples`,
 * then clears them.
 */
int salmon_unet_train_sample(salmon_unet_t u, const float *in_dev, const float *target_dev, float *loss_dev);
int salmon_unet_adam_step(salmon_unet_t u, float lr, int nsamples);

/* Parameter gradients accumulated so far, flat as in get_params (for checks). */
int salmon_unet_get_grads(salmon_unet_t u, float *host);

/* Activation `index` of the last forward pass as float (for checks): -1 = packed input x0
   (n0 x 16); 0..11 = e1a, e1b, p1, e2a, e2b, p2, e3a, e3b, d2a, d2b, d1a, d1b (channels-last).
   n must equal the tensor's size. */
int salmon_unet_get_activation(salmon_unet_t u, int index, float *host, long n);

/*
 * Glue between FP64 potentials V_local(ix,iy,iz) (DEVICE) and the network, for the
 * history v0 = V(t), v1 = V(t-dt), v2, v3:
 *   features:  in[p] = ((v0-v1)/s_diff, (v1-v2)/s_diff, (v2-v3)/s_diff, (v0-vpsl)/s_abs)
 *   predict:   vpred = v0 + (v0-v1)/2 + s_out * nn     (midpoint potential V(t + dt/2))
 *   target:    t = (vmid - v0 - (v0-v1)/2) / s_out    (training target for nn)
 *   residual:  *out_dev = max_p |vpred - (v0 + vnew)/2|, the self-consistency residual of a
 *              step propagated with vpred whose new potential is vnew
 *              (the corrector average of src/rt/time_evolution_step.f90:402-405).
 */
int salmon_unet_features(salmon_unet_t u, const double *v0, const double *v1, const double *v2, const double *v3,
                         const double *vpsl, double s_diff, double s_abs, float *in_dev);
int salmon_unet_predict(salmon_unet_t u, const double *v0, const double *v1, const float *nn_dev, double s_out,
                        double *vpred_dev);
int salmon_unet_target(salmon_unet_t u, const double *vmid, const double *v0, const double *v1, double s_out,
                       float *t_dev);
int salmon_unet_residual(salmon_unet_t u, const double *vpred, const double *v0, const double *vnew, double *out_dev);

/*
 * Acceptance decided on the device (no host synchronisation): *flag_dev = (*resid_dev <= tol);
 * stats_dev[0] += the flag, stats_dev[1] += 1 (accepted / attempted steps); while
 * stats_dev[1] < log_cap, log_dev[stats_dev[1]] = the residual (log_dev may be null with
 * log_cap = 0).  The flag is meant for salmon_hip_set_skip / salmon_hip_copy_if.
 */
int salmon_unet_gate(salmon_unet_t u, const double *resid_dev, double tol, int *flag_dev, int *stats_dev,
                     double *log_dev, int log_cap);

#ifdef __cplusplus
}
#endif
#endif
