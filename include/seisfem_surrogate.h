/*
 * seisfem_surrogate.h -- learned surrogate of the FEM loop: an MLP mapping
 * layered-model parameters + source position to receiver traces, trained and
 * served on the GPU by the library's own kernels (bf16 MFMA GEMMs with FP32
 * accumulation, GELU, MSE, Adam with FP32 master weights).
 *
 * No counterpart in the reference (it has no learned component); the FEM
 * loop generates the training shots and measures the surrogate's error.
 * Host pointers throughout; returns SFG_OK or a negative SFG_E_* code
 * (message in sfg_mlp_last_error()).
 */
#ifndef SEISFEM_SURROGATE_H
#define SEISFEM_SURROGATE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sfg_mlp sfg_mlp;

const char *sfg_mlp_last_error(void);
/* nin -> hidden (x nhidden_layers, GELU) -> nout; fixed batch size (multiple of 32). */
int sfg_mlp_create(sfg_mlp **m, int nin, int hidden, int nhidden_layers, int nout, int batch,
                   unsigned long long seed);
void sfg_mlp_destroy(sfg_mlp *m);
long long sfg_mlp_param_count(const sfg_mlp *m);
/* Copies the training set X [n][nin], Y [n][nout] to the device. */
int sfg_mlp_set_data(sfg_mlp *m, const float *X, const float *Y, int n);
/* One Adam step on samples idx[batch] of the resident set; *loss = batch MSE. */
int sfg_mlp_train_step(sfg_mlp *m, const int *idx, float lr, double *loss);
/* epochs over the resident set (shuffled), lr = lr0 * cosine decay over
 * total_steps Adam steps (0 = the steps of this call). */
int sfg_mlp_train(sfg_mlp *m, int epochs, float lr0, long long total_steps, unsigned long long seed,
                  double *last_loss, double *ms_per_step);
/* bf16 inference Y [n][nout] = net(X [n][nin]); *ms_total = device time of the forward passes. */
int sfg_mlp_predict(sfg_mlp *m, const float *X, int n, float *Y, double *ms_total);
/* Loss and the gradients of one layer for the batch Xb/Yb (no update). */
int sfg_mlp_gradients(sfg_mlp *m, const float *Xb, const float *Yb, double *loss, int layer, float *dW,
                      float *db);
/* FP32 master weights of one layer: W [dout][din], b [dout]. */
int sfg_mlp_get_layer(sfg_mlp *m, int layer, float *W, float *b);
/* Sets the FP32 master weights of one layer (and their bf16 copy): W [dout][din], b [dout]. */
int sfg_mlp_set_layer(sfg_mlp *m, int layer, const float *W, const float *b);
/* C [M][N] = bf16(A) [M][K] . bf16(B)^T [K][N] with FP32 accumulation (kernel test entry). */
int sfg_gemm_bf16_nt(int M, int N, int K, const float *A, const float *B, float *C);
/* Times `reps` launches of the same GEMM on device-resident bf16 operands with
   pseudo-random entries (K a multiple of 32): *ms = average device time of one
   launch (HIP events around the launches on their stream). */
int sfg_gemm_bf16_time(int M, int N, int K, int reps, double *ms);

#ifdef __cplusplus
}
#endif
#endif
