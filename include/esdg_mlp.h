/*
 * esdg_mlp.h -- learned troubled-cell indicator for the artificial-viscosity path (N = 7).
 *
 * No counterpart in the reference, which detects troubled elements with the modal sensor of
 * okl/ShockCapturing/ShockCapturing.okl only.  A small MLP classifies each element as smooth or
 * troubled from 72 features; the modal sensor stays the judge (it labels the training data and
 * measures the false negatives) and the fallback (it decides where the network is unsure, and it
 * always sets the viscosity magnitude of a troubled element).
 *
 * Features of element e (ngl = 8, node n = j*8 + i):
 *   x[n], n < 64      (h_n - mean h) / max_n |h_n - mean h|   (0 for a constant element)
 *   x[64 + k], k < 8  max(log10(E_k / E), -16) / 8, E_k the energy of modal shell k (modes (l, m)
 *                     with max(l, m) = k) of m = VdmInv (h + b) VdmInv^T, E the total energy
 *                     (-2 for E = 0)
 * Network: 72 -> 32 (ReLU) -> 32 (ReLU) -> 1 (sigmoid).  Weights are BF16, activations are rounded
 * to BF16 before each product, accumulation is FP32 (MFMA 16x16x32 BF16 on gfx950), biases FP32.
 * Training keeps FP32 master weights and runs the same rounded forward (straight-through).
 *
 * Parameter vector (ESDG_MLP_NPARAM floats, this order): W1[72][32], b1[32], W2[32][32], b2[32],
 * w3[32], b3.
 */
#ifndef ESDG_MLP_H
#define ESDG_MLP_H

#include "esdg.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ESDG_MLP_NIN 72
#define ESDG_MLP_NHID 32
#define ESDG_MLP_NPARAM (72 * 32 + 32 + 32 * 32 + 32 + 32 + 1)

typedef struct esdg_mlp esdg_mlp;

/* A network on the context's device, initialised with seeded He-uniform weights. */
int esdg_mlp_create(esdg_ctx* ctx, uint64_t seed, esdg_mlp** out);
int esdg_mlp_destroy(esdg_mlp* m);
/* Host FP32 parameter vectors (layout above).  set rounds the inference copy to BF16. */
int esdg_mlp_set_params(esdg_mlp* m, const float* host_params);
int esdg_mlp_get_params(esdg_mlp* m, float* host_params);

/* Features of Nelem elements (device arrays: q, B [Nelem][64], VdmInv [8][8]) into feat
 * (device, [Nelem][72] FP32).  Requires ngl = 8. */
int esdg_mlp_features(esdg_ctx* ctx, int Nelem, const double* q, const double* B, const double* VdmInv,
                      float* feat);
/* p[i] = network(feat[i]) for n feature rows (device arrays); BF16 MFMA inference. */
int esdg_mlp_infer(esdg_mlp* m, int n, const float* feat, float* p);
/* Features and inference fused: p[e] for every element of the state q (device arrays). */
int esdg_mlp_indicator(esdg_mlp* m, int Nelem, const double* q, const double* B, const double* VdmInv, float* p);

/* Mean weighted binary cross-entropy over the rows idx[0..n) of (feat, label) and its gradient
 * with respect to the parameters (device arrays; grad has ESDG_MLP_NPARAM floats, loss 1 float).
 * label is 0 or 1; positives weigh pos_weight.  idx may be null (rows 0..n-1). */
int esdg_mlp_loss_grad(esdg_mlp* m, int n, const float* feat, const float* label, const int* idx,
                       float pos_weight, float* grad, float* loss);
/* One Adam step (lr, beta1 0.9, beta2 0.999, eps 1e-8) on the rows idx[0..n); *loss (host, may be
 * null) receives the minibatch loss before the step. */
int esdg_mlp_train_step(esdg_mlp* m, int n, const float* feat, const float* label, const int* idx,
                        float pos_weight, float lr, float* loss);

/* Learned indicator in a solver's artificial-viscosity path (fused schedules 1 and 2, N = 7): at every
 * stage the network classifies the elements; p >= hi: troubled, p <= lo: smooth, in between
 * unsure.  Smooth elements get no viscosity; troubled and unsure ones the modal sensor's value,
 * which the indicator kernel evaluates from the modal coefficients its features already formed.
 * Elements with no viscosity among themselves and their face neighbours skip the BR1 and viscous
 * work (as they do without the indicator).  m = null restores the modal sensor everywhere. */
int esdg_solver_set_indicator(esdg_solver* s, esdg_mlp* m, float lo, float hi);
/* audit = 1 counts the sensor's verdict on the elements the network declared smooth too (the
 * solution is unchanged), for the false-negative counts below. */
int esdg_solver_indicator_audit(esdg_solver* s, int audit);
/* Counts since the last call, in element-stages: with viscosity; the sensor's positives among
 * those the network declared smooth (false negatives, audit only); the sensor's positives (audit
 * only); not declared smooth by the network (troubled or unsure); all. */
int esdg_solver_indicator_counts(esdg_solver* s, int64_t* viscous, int64_t* false_negatives,
                                 int64_t* sensor_positives, int64_t* not_smooth, int64_t* element_stages);

#ifdef __cplusplus
}
#endif
#endif
