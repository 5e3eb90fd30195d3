/*
 * mcb_mlp.h - C ABI of the learned FVS vertex scorer in libmcb_hip.so.
 *
 * No counterpart in the reference: the reference orders its reverse-delete FVS pass by degree
 * (include/FVS.cpp:166-227).  This scorer is a 4 -> 64 -> 64 -> 1 MLP (ReLU hidden layers, logistic
 * output) over the per-vertex features of mcb_vertex_features (include/mcb_host.h).  Its score orders
 * the pass through mcb_build_options.fvs_order (lowest score tried first); the pass itself keeps the
 * result a minimal feedback vertex set for any order.  Weights and activations are BF16 on the device
 * with FP32 accumulation (the 64 x 64 layer and its gradients on MFMA); the parameters are kept and
 * trained as FP32 masters.
 *
 * Parameters, MCB_MLP_PARAMS floats in this order: W1 [4][64] (row = input feature), b1 [64],
 * W2 [64][64] (row = input unit), b2 [64], W3 [64], b3.
 *
 * Status-returning functions return MCB_OK or a negative MCB_E_* code (mcb_host.h);
 * mcb_mlp_last_error() gives the calling thread's last message.
 */
#ifndef MCB_MLP_H
#define MCB_MLP_H

#include <stdint.h>

#include "mcb_host.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MCB_MLP_PARAMS 4545

typedef struct mcb_mlp mcb_mlp;

/* Device parameters (all zero), Adam state and a stream.  MCB_E_DEVICE without a GPU. */
int mcb_mlp_create(mcb_mlp **out);

/* Replaces the parameters (and resets the Adam state); reads them back. */
int mcb_mlp_set_params(mcb_mlp *m, const float *params);
int mcb_mlp_get_params(mcb_mlp *m, float *params);

/* Logits of n rows of 4 features (host arrays). */
int mcb_mlp_forward(mcb_mlp *m, const float *x, int64_t n, float *logits);

/* Gradient of the mean binary cross-entropy over n rows with labels y in {0, 1} (host arrays), and
 * that loss.  Leaves the parameters unchanged. */
int mcb_mlp_gradient(mcb_mlp *m, const float *x, const float *y, int64_t n, float *grad, float *loss);

/* Uploads a training set (n rows of 4 features, labels) that stays on the device. */
int mcb_mlp_set_data(mcb_mlp *m, const float *x, const float *y, int64_t n);

/* One epoch of Adam (beta 0.9 / 0.999, eps 1e-8) over the training set in the given row order
 * (a permutation of 0..n-1), `batch` rows per step; *mean_loss receives the epoch's mean loss
 * (each step's loss taken before its update). */
int mcb_mlp_train_epoch(mcb_mlp *m, const int32_t *order, int64_t batch, float lr, float *mean_loss);

void mcb_mlp_destroy(mcb_mlp *m);

const char *mcb_mlp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MCB_MLP_H */
