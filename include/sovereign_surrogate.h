/* sovereign_surrogate.h - the learned branching-score surrogate (C ABI).
 *
 * A multilayer perceptron  x (in) -> 256 -> 256 -> 1  with ReLU hidden
 * layers scores strong-branching candidates from per-candidate features, so
 * that only the top-ranked few get real (batched) strong branching. Matrix
 * products run on BF16 MFMA (v_mfma_f32_16x16x16bf16_1k): weights and
 * activations are rounded to BF16 (round to nearest even) as they enter an
 * instruction, accumulation is FP32, master weights are FP32 and the
 * optimiser is Adam. Training is on the GPU with these kernels (forward,
 * backward, update).
 *
 * No counterpart in the reference, which ranks strong-branching candidates
 * by fractionality and pseudocosts only (solver/milp/src/branch_and_bound.cpp
 * pick_strong_branch / pick_reliability_branch, :726-899). Status codes as in
 * sovereign_hip.h. Arrays are row-major: x is batch x in, targets and outputs
 * batch long. */
#ifndef SOVEREIGN_SURROGATE_H
#define SOVEREIGN_SURROGATE_H

#include "sovereign_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sov_mlp sov_mlp;

/* Network shape: in features, two hidden layers, one output. */
typedef struct {
  int32_t in;      /* input features (32 for the branching features) */
  int32_t hidden1; /* 256 */
  int32_t hidden2; /* 256 */
  int32_t pad;
} sov_mlp_shape;

/* Creates a network with Glorot-uniform weights and biases from `seed`. */
int sov_mlp_create(const sov_mlp_shape* shape, uint64_t seed, sov_mlp** out);
/* Number of FP32 parameters: packed as W1 (in x h1), b1 (h1), W2 (h1 x h2),
 * b2 (h2), W3 (h2), b3 (1), each row-major. */
int64_t sov_mlp_param_count(const sov_mlp* net);
int sov_mlp_get_params(const sov_mlp* net, float* params);
int sov_mlp_set_params(sov_mlp* net, const float* params);
/* y[b] = net(x[b, :]) for b < batch (host arrays). */
int sov_mlp_forward(sov_mlp* net, const float* x, int32_t batch, float* y);
/* Loss 0.5 mean_b (net(x_b) - t_b)^2 and its gradient with respect to the
 * packed parameters (no update). loss may be NULL. */
int sov_mlp_gradients(sov_mlp* net, const float* x, const float* t, int32_t batch, float* grad, float* loss);
/* Adam training on a device-resident copy of (x, t): `epochs` passes in
 * minibatches of `batch`, shuffled by `seed`. loss_per_epoch (epochs
 * entries, may be NULL) receives the mean training loss of each pass. */
int sov_mlp_train(sov_mlp* net, const float* x, const float* t, int64_t n, int32_t batch, int32_t epochs,
                  float learning_rate, uint64_t seed, float* loss_per_epoch);
int sov_mlp_free(sov_mlp* net);

#ifdef __cplusplus
}
#endif

#endif /* SOVEREIGN_SURROGATE_H */
