/* ifea_mlp.h — the learned warm start's per-node network on MI355X: forward, backward and Adam in hand-written HIP.
 *
 * No counterpart in the reference: it feeds the warm-start slot LoadVelocityGuess
 * (src/ImpressionFEA.Core/ThreeD/Forging3DSolver.cs:854-869), whose own guesses are the previous increment's velocity
 * and the ExtrapolateStart trend (:44, :856-859).
 *
 * Network: n_in -> width (GELU) -> ... (depth hidden layers) -> n_out, row-major activations [nodes x features].
 * Every product runs on BF16 MFMA (v_mfma_f32_16x16x16_bf16) with both operands rounded to BF16 and FP32 accumulation;
 * activations and gradients stay FP32 between the products (the bias gradient is an FP32 sum), and FP32 master weights
 * and Adam moments are kept on the device.
 * Parameter vector layout (FP32): for each layer l: W_l [fan_in x fan_out] row-major, then b_l [fan_out].
 * Reductions over nodes (weight gradients, loss) use fixed-order partial sums: results are bit-reproducible.
 */
#ifndef IFEA_MLP_H
#define IFEA_MLP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ifea_mlp* ifea_mlp_t;

int ifea_mlp_create(ifea_mlp_t* h, int n_in, int width, int depth, int n_out);
/* Number of FP32 parameters. */
long long ifea_mlp_num_params(ifea_mlp_t h);
int ifea_mlp_set_params(ifea_mlp_t h, const float* params);
int ifea_mlp_get_params(ifea_mlp_t h, float* params);
/* y[n x n_out] = net(x[n x n_in]) for host arrays. */
int ifea_mlp_forward(ifea_mlp_t h, int n, const float* x, float* y);
/* Mean-squared-error loss over n rows and its gradient with respect to every parameter (FP32, grad[num_params]). */
int ifea_mlp_loss_grad(ifea_mlp_t h, int n, const float* x, const float* y, float* loss, float* grad);
/* Upload a training set once (x[n x n_in], y[n x n_out]). */
int ifea_mlp_set_dataset(ifea_mlp_t h, int n, const float* x, const float* y);
/* One epoch over the dataset in the order perm[n] with minibatches of `batch` rows and Adam (beta1 0.9, beta2 0.999,
 * eps 1e-8); returns the mean minibatch loss. */
int ifea_mlp_train_epoch(ifea_mlp_t h, const int* perm, int batch, float lr, float* mean_loss);
void ifea_mlp_destroy(ifea_mlp_t h);

#ifdef __cplusplus
}
#endif
#endif
