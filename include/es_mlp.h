/* es_mlp.h -- C ABI of the wheel-planner surrogate network (part of libes_sieve.so).
 *
 * No reference counterpart: the reference plans wheels with the analytic knapsack of
 * sieve/gpu_sieve_v9.py:146-227 only.  This network (64 inputs, two ReLU layers of 128, 2
 * outputs, FP32) predicts per candidate wheel its log survivor density and its measured
 * ns per residue position; the planner uses it to shortlist wheels for an exact cost and a
 * short GPU probe.  Forward, backward and Adam run on the GPU (es_mlp.hip).
 *
 * Parameters are one packed float vector of ES_MLP_PARAMS entries, weights row-major as
 * (input, output): W1 [64 x 128], b1 [128], W2 [128 x 128], b2 [128], W3 [128 x 16],
 * b3 [16]; only the first 2 of the 16 output columns are used, the others are held at zero.
 * Loss: (1/n) sum_rows sum_{o<2} w_o (y_o - t_o)^2.  All calls synchronise. */
#ifndef ES_MLP_H
#define ES_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ES_MLP_IN 64
#define ES_MLP_HIDDEN 128
#define ES_MLP_OUT 2
#define ES_MLP_PARAMS 26896

typedef struct es_mlp es_mlp;

/* Network with device buffers for batches of up to batch_cap rows; parameters start at 0. */
int es_mlp_create(uint32_t batch_cap, es_mlp** out);
int es_mlp_destroy(es_mlp* n);
/* Replace the parameters (count == ES_MLP_PARAMS) and reset the Adam state and step. */
int es_mlp_set_params(es_mlp* n, const float* params, uint32_t count);
int es_mlp_get_params(es_mlp* n, float* params, uint32_t count);
/* Y [count x 2] = network(X [count x 64]); any count (processed in batch_cap chunks). */
int es_mlp_forward(es_mlp* n, const float* X, uint64_t count, float* Y);
/* Loss and its gradient (packed like the parameters, may be NULL) over one batch of
 * count <= batch_cap rows (X [count x 64], T [count x 2]); no update. */
int es_mlp_gradients(es_mlp* n, const float* X, const float* T, uint32_t count,
                     const float w[2], float* grads, float* loss);
/* es_mlp_gradients, then one Adam update (bias-corrected, step counter advanced). */
int es_mlp_adam_step(es_mlp* n, const float* X, const float* T, uint32_t count,
                     const float w[2], float lr, float b1, float b2, float eps, float* loss);
/* Keep a training set resident on the device (X [count x 64], T [count x 2]). */
int es_mlp_upload_dataset(es_mlp* n, const float* X, const float* T, uint64_t count);
/* Epochs of minibatch Adam over the resident set, rows shuffled each epoch (xorshift64* from
 * seed); *last_loss = loss of the final minibatch before its update, *ms = device time. */
int es_mlp_train(es_mlp* n, uint32_t epochs, uint32_t batch, uint64_t seed, const float w[2],
                 float lr, float b1, float b2, float eps, float* last_loss, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* ES_MLP_H */
