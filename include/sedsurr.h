/* sedsurr.h -- C ABI of the spin-up surrogate: a multilayer perceptron
 * n_in -> n_hidden x n_layers (GELU) -> n_out that maps the 12 forcing and
 * parameter scalars of a sediment column to its near-steady log10 profiles,
 * trained and evaluated on the GPU.
 *
 * No reference counterpart: the reference spin-up is a single-column
 * presimulation copied to every column (fabm_sediment_component.F90:557-632);
 * the surrogate supplies a per-column warm start for sedcol_spinup
 * (sedcol.h) instead.
 *
 * Arithmetic: weights and activations in bfloat16 (round to nearest even),
 * matrix products on the MFMA units with FP32 accumulation, biases, GELU
 * (tanh form), loss, gradients and Adam state in FP32.  Inputs and targets
 * cross the ABI as FP32, already normalised by the caller; every function
 * returns a sedcol_status code (sedcol.h) and sets sedsurr_last_error().
 * A handle belongs to the device that is current at sedsurr_create; every
 * call runs on that device and leaves the caller's current device as it
 * found it. */
#ifndef SEDSURR_H
#define SEDSURR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sedsurr_handle sedsurr_handle;

/* Network shape and the largest batch of one training step or one
 * prediction pass (predictions of more rows run in passes). */
typedef struct {
  int32_t n_in;      /* inputs (<= 64) */
  int32_t n_hidden;  /* width of every hidden layer (multiple of 64) */
  int32_t n_layers;  /* hidden layers (>= 1) */
  int32_t n_out;     /* outputs */
  int32_t max_batch; /* rows per pass (multiple of 64) */
  int32_t reserved;  /* must be 0 */
} sedsurr_shape;

/* Adam (Kingma & Ba 2015) settings of sedsurr_train. */
typedef struct {
  float lr, beta1, beta2, eps;
} sedsurr_adam;

int sedsurr_create(const sedsurr_shape *shape, sedsurr_handle **out);
int sedsurr_destroy(sedsurr_handle *h);
const char *sedsurr_last_error(void);

/* Parameters as one FP32 vector: for each layer l (hidden layers, then the
 * output layer) W_l (fan_in x fan_out, row-major) followed by b_l (fan_out).
 * set_params also resets the Adam state. */
int sedsurr_num_params(sedsurr_handle *h, int64_t *n);
int sedsurr_set_params(sedsurr_handle *h, const float *params);
int sedsurr_get_params(sedsurr_handle *h, float *params);

/* Training set, copied to the device and kept there: X (n x n_in),
 * Y (n x n_out), row-major, normalised. */
int sedsurr_set_data(sedsurr_handle *h, const float *X, const float *Y, int64_t n);

/* `steps` Adam steps on minibatches of `batch` rows; rows[steps * batch]
 * are the row indices of every minibatch (into the training set).
 * loss (steps entries, optional) receives each step's mean squared error
 * before its update. */
int sedsurr_train(sedsurr_handle *h, int32_t steps, int32_t batch, const int32_t *rows,
                  const sedsurr_adam *adam, float *loss);

/* Loss (mean over rows and outputs of the squared error) and its gradient
 * with respect to every parameter (sedsurr_num_params entries, the layout
 * of sedsurr_get_params) for one batch, without an update. */
int sedsurr_loss_grad(sedsurr_handle *h, const float *X, const float *Y, int32_t n, float *loss,
                      float *grad);

/* Y (n x n_out) = network(X (n x n_in)), normalised units. */
int sedsurr_predict(sedsurr_handle *h, const float *X, int64_t n, float *Y);
/* The same with Y left on the GPU, for example as the input of
 * sedcol_set_state_profiles, so that a warm start never passes through host
 * memory.  X is host memory, n >= 1; *Y_device receives the address of a
 * device buffer of the handle (n x n_out, row-major) that stays valid until the
 * next sedsurr_predict_device call or sedsurr_destroy.  The buffer grows to
 * the largest n asked for and is freed with the handle. */
int sedsurr_predict_device(sedsurr_handle *h, const float *X, int64_t n, const float **Y_device);

/* Timing of the last sedsurr_train / sedsurr_predict (HIP events on the
 * handle's stream): milliseconds and kernel launches. */
int sedsurr_last_timing(sedsurr_handle *h, double *ms, int32_t *launches);

#ifdef __cplusplus
}
#endif
#endif
