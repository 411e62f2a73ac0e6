/* Coordinate MLP of the learned warm start, on gfx950 (BF16 MFMA with FP32
 * accumulation).  No counterpart in the reference: the reference always
 * starts from the free stream (Euler2D::initializeData, Euler2D.cpp:153-171);
 * this library predicts a starting field that usg_set_prim then loads.
 *
 * Network: nlayer linear layers; layer 0 maps nin inputs to `width`,
 * layers 1..nlayer-2 map width to width, the last maps width to nout; SiLU
 * after every layer but the last.  Arithmetic (the CPU oracle restates it):
 *   H_0 = bf16(x)
 *   Z_l = fp32( sum_k H_{l}[k] * bf16(W_l)[k][n] ) + b_l[n]
 *   H_{l+1} = bf16(silu(Z_l))            (l < nlayer-1),  out = Z_{nlayer-1}
 *   loss = mean over points and outputs of (out - y)^2
 * backward with dZ_l = bf16(dH_l * silu'(Z_l)) in both products,
 * dW_l = sum_p H_l[p][k] dZ_l[p][n], db_l = sum_p dZ_l[p][n] (fp32 sums in a
 * fixed order: results are reproducible), Adam on FP32 master weights.
 * Weights are W_l[k][n] row-major (fan-in major), all layers concatenated;
 * biases likewise.  All calls are synchronous on the library's stream. */
#ifndef USG_MLP_H
#define USG_MLP_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct usg_mlp usg_mlp;

typedef struct usg_mlp_desc {
  int nin;      /* inputs, <= 128                        */
  int width;    /* hidden width, multiple of 64, <= 256  */
  int nlayer;   /* linear layers, >= 2                   */
  int nout;     /* outputs, <= 16                        */
} usg_mlp_desc;

/* Number of weights (sum of fan_in*fan_out) and biases of a network. */
int usg_mlp_sizes(const usg_mlp_desc *d, long *nweight, long *nbias);
int usg_mlp_create(const usg_mlp_desc *d, usg_mlp **out);
int usg_mlp_destroy(usg_mlp *m);
/* FP32 master parameters in/out (the layouts above); set resets Adam. */
int usg_mlp_set_params(usg_mlp *m, const float *w, const float *b);
int usg_mlp_get_params(usg_mlp *m, float *w, float *b);
/* Outputs of n points: x[n*nin] -> y[n*nout] (host arrays). */
int usg_mlp_predict(usg_mlp *m, const float *x, long n, float *y);
/* Loss and gradients of one batch without updating: dw, db as the parameter
 * layouts (host arrays). */
int usg_mlp_gradients(usg_mlp *m, const float *x, const float *y, int n, float *loss, float *dw, float *db);
/* Training set held on the GPU: x[n*nin], y[n*nout] (host arrays, copied). */
int usg_mlp_set_data(usg_mlp *m, const float *x, const float *y, long n);
/* One pass over the training set in the order perm[n] (host array), in
 * batches of `batch` points, one Adam step (lr, beta1 0.9, beta2 0.999,
 * eps 1e-8) per batch; loss_out (may be NULL) receives the mean batch loss. */
int usg_mlp_train_epoch(usg_mlp *m, const int *perm, long n, int batch, float lr, float *loss_out);
/* BF16 MFMA operand-layout check: d[16*16] = a[16*32] * b[32*16] through one
 * v_mfma_f32_16x16x32_bf16 (row-major host arrays of floats, rounded to bf16). */
int usg_mlp_mfma_check(const float *a, const float *b, float *d);
/* Last error text of the calling thread. */
const char *usg_mlp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
