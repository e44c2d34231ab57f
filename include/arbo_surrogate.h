/* arbo_surrogate.h — C ABI of the learned surrogate of the simulator.
 *
 * No counterpart in the reference: its MCMC evaluates every proposal with
 * C_MAXSIMRUN full simulations (main.cpp:262-277).  The surrogate is a
 * multilayer perceptron (GELU hidden layers, linear output) mapping model
 * parameters to the log-mean weekly incidence; it is trained on simulations
 * from the engine and screens proposals as the first stage of
 * delayed-acceptance MCMC, the simulator staying the second stage.
 *
 * Every matrix product runs on the bfloat16 MFMA units with float32
 * accumulation, each float32 operand split into two bfloat16 halves (three
 * products per step: about 16 significant bits per operand); parameters,
 * activations, gradients and the optimiser state are float32.
 * Host buffers are row-major float32: inputs [n * width_0], outputs
 * [n * width_last].  Parameters flatten layer by layer as W [width_{l+1} x
 * width_l] (row-major) then b [width_{l+1}].
 *
 * Conventions as include/arbo_c.h: 0 or a negative ARBO_E_* code, message in
 * arbo_sur_last_error(); one surrogate per thread.
 */
#ifndef ARBO_SURROGATE_H
#define ARBO_SURROGATE_H
#include <stdint.h>

#include "arbo_c.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ARBO_SUR_MAX_LAYERS 8u
#define ARBO_SUR_MAX_WIDTH 4096u

typedef struct ArboSurrogate ArboSurrogate;

/* n_widths = layers + 1 (2..ARBO_SUR_MAX_LAYERS + 1); rows per call up to
 * max_batch.  Weights start He-normal (Philox seed), biases at zero. */
ARBO_API int arbo_sur_create(int32_t device, uint32_t n_widths, const uint32_t* widths, uint32_t max_batch,
                             uint64_t seed, ArboSurrogate** out);
ARBO_API int arbo_sur_n_params(const ArboSurrogate* s, uint64_t* out);
/* Input and output widths. */
ARBO_API int arbo_sur_shape(const ArboSurrogate* s, uint32_t* inputs, uint32_t* outputs);
ARBO_API int arbo_sur_get_params(ArboSurrogate* s, float* out);
ARBO_API int arbo_sur_set_params(ArboSurrogate* s, const float* in);

/* out = f(x) for n rows. */
ARBO_API int arbo_sur_forward(ArboSurrogate* s, uint32_t n, const float* x, float* out);

/* out = f(x) for n rows on device buffers, allocating nothing.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  x and out are
 * device pointers (x [n * width_0] float32, out [n * width_last] float32);
 * the work is queued on `stream` (a hipStream_t; NULL: the surrogate's own
 * stream) and the call returns without waiting.  The activations between
 * layers live in a fixed 8 MiB scratch array of the library on each GPU
 * (two buffers of 2^20 float32), so n x (the widest hidden layer) must not
 * exceed 2^20 (4,096 rows of a 256-wide network), else ARBO_E_CAPACITY;
 * every call on one GPU shares that array, so calls that overlap on one GPU
 * race.  The results are those of arbo_sur_forward, bit for bit. */
ARBO_API int arbo_sur_forward_noalloc(ArboSurrogate* s, uint32_t n, const float* x, float* out, void* stream);

/* out = f(x) and dx = (d f / d x)^T dout, row by row. */
ARBO_API int arbo_sur_input_grad(ArboSurrogate* s, uint32_t n, const float* x, const float* dout, float* out,
                                 float* dx);

/* Poisson loss L = mean over rows and outputs of exp(f) - y f (its minimiser
 * is f = log E[y]) and its gradient with respect to every parameter. */
ARBO_API int arbo_sur_loss_grad(ArboSurrogate* s, uint32_t n, const float* x, const float* y, double* loss,
                                float* grad);

/* Adam (Kingma & Ba 2015; beta1 0.9, beta2 0.999, eps 1e-8) on the Poisson
 * loss over minibatches of a data set of n rows, reshuffled every epoch
 * (Philox seed); epoch_loss[epochs] gets each epoch's mean loss (every
 * minibatch scored before its own update).  The
 * optimiser state persists across calls. */
ARBO_API int arbo_sur_train(ArboSurrogate* s, uint32_t n, const float* x, const float* y, uint32_t epochs,
                            uint32_t batch, float lr, uint64_t seed, double* epoch_loss);

ARBO_API void arbo_sur_destroy(ArboSurrogate* s);
ARBO_API const char* arbo_sur_last_error(const ArboSurrogate* s);

#ifdef __cplusplus
}
#endif
#endif /* ARBO_SURROGATE_H */
