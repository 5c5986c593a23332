/*
 * cable_mlp.h -- C ABI of the canopy surrogate in libcable_gpu.so.
 *
 * A 24 -> 128 -> 128 -> 128 -> 8 multilayer perceptron with hard-swish
 * activations (z min(max(z + 3, 0), 6) / 6), trained on the GPU by
 * hand-written kernels: BF16 operands on the matrix cores with FP32
 * accumulation, FP32 master weights, Adam
 * (beta1 0.9, beta2 0.999, eps 1e-8), mean-squared error on normalised
 * outputs.  It maps a tile's forcing, state and parameters at the start of a
 * step to the converged state of the reference's leaf fixed point
 * (src/science/canopy/cbl_dryLeaf.F90:10-666); the hybrid step uses it to
 * warm-start that iteration.  No counterpart in the reference.
 *
 * Parameters are one packed FP32 vector of cable_mlp_nparam() = 37,256
 * values: W1 (24x128), b1 (128), W2 (128x128), b2, W3 (128x128), b3,
 * W4 (128x8), b4 (8); each W row-major [in][out].  Normalisation: 64
 * values, mu_x[24], sd_x[24], mu_y[8], sd_y[8]; the network sees
 * (x - mu_x) / sd_x and predicts (y - mu_y) / sd_y.
 *
 * Every function returns 0 on success; cable_mlp_last_error() gives the
 * message of the calling thread's last failure.
 */
#ifndef CABLE_MLP_H
#define CABLE_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cable_mlp cable_mlp;

enum { CABLE_MLP_OK = 0, CABLE_MLP_E_ARG = 1, CABLE_MLP_E_HIP = 2, CABLE_MLP_E_STATE = 3 };
enum { CABLE_MLP_MAX_BATCH = 65536 };

int cable_mlp_nparam(void);
const char *cable_mlp_last_error(void);

/* a network with buffers for batches of up to `batch` rows */
int cable_mlp_create(cable_mlp **m, int batch);
int cable_mlp_destroy(cable_mlp *m);

/* parameters (host FP32, packed as above); set also resets Adam's state */
int cable_mlp_set_params(cable_mlp *m, const float *params);
int cable_mlp_get_params(cable_mlp *m, float *params);
int cable_mlp_set_norm(cable_mlp *m, const float *norm);

/* y[n x 8] = network(x[n x 24]) on normalised values (host arrays, n <= batch) */
int cable_mlp_forward(cable_mlp *m, const float *x, int n, float *y);
/* gradient of the mean-squared error of (x, t) w.r.t. every parameter,
   readable with cable_mlp_get_grads; *loss receives the loss */
int cable_mlp_gradient(cable_mlp *m, const float *x, const float *t, int n, float *loss);
int cable_mlp_get_grads(cable_mlp *m, float *grads);

/* training samples, raw (unnormalised): x[n x 24], y[n x 8]; copied to the
   device, or (use_samples) device arrays the caller keeps alive */
int cable_mlp_load_samples(cable_mlp *m, const float *x, const float *y, int64_t n);
int cable_mlp_use_samples(cable_mlp *m, const float *x_dev, const float *y_dev, int64_t n);

/* nbatches Adam steps; batch b uses samples order[b*batch .. b*batch+batch-1];
   losses[b] (optional) receives each batch's loss */
int cable_mlp_train(cable_mlp *m, const int64_t *order, int64_t nbatches, float lr, float *losses);

/* nbatches Adam steps in an order the device draws itself: batch b takes
   positions (first_batch + b) * batch .. + batch - 1 of an endless sequence
   of epochs, each a permutation of the loaded samples keyed by (seed,
   epoch); losses[b] (optional) receives each batch's loss.  The same steps
   as cable_mlp_train with the order cable_mlp_shuffled_order gives, with no
   order to compute on the host or copy to the device. */
int cable_mlp_train_shuffled(cable_mlp *m, uint64_t seed, int64_t first_batch, int64_t nbatches, float lr,
                             float *losses);
/* the sample indices at positions first .. first + count - 1 of that
   sequence over n samples (host only) */
int cable_mlp_shuffled_order(uint64_t seed, int64_t n, int64_t first, int64_t count, int64_t *out);

#ifdef __cplusplus
}
#endif
#endif /* CABLE_MLP_H */
