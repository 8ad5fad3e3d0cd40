/* couplb_mlp.h -- C ABI of the warm-start network of the multi-direct-forcing
 * IBM (couplb_ibm_mdf_device).  No counterpart in the reference, whose IBM is
 * a single penalty pass (fix_couplb.cpp:779-785).
 *
 * Network: 67 inputs -> 64 -> 64 -> 3 outputs, GELU(z) = z/2 (1 + erf(z/sqrt 2))
 * after both hidden layers.  FP32 master weights; products run on the matrix
 * cores with BF16 operands (inputs, weights, activations and back-propagated
 * errors rounded to BF16) and FP32 accumulation.  Loss: mean over samples of
 * the squared error summed over the 3 outputs.  Training: Adam.
 *
 * Parameter vector (FP32), row-major with the input index first:
 *   W1[67][64], b1[64], W2[64][64], b2[64], W3[64][3], b3[3]   (8707 values)
 *
 * All array arguments named *_dev are device memory; the calls are
 * synchronous.  Status codes are those of couplb_hip.h. */
#ifndef COUPLB_MLP_H
#define COUPLB_MLP_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct couplb_mlp couplb_mlp;

#define COUPLB_MLP_INPUTS 67
#define COUPLB_MLP_OUTPUTS 3
#define COUPLB_MLP_NPARAM 8707

/* New network on `device`, weights drawn from N(0, 1/fan_in) with `seed`
 * (biases zero), Adam state cleared. */
int couplb_mlp_create(int device, unsigned long long seed, couplb_mlp** out);
int couplb_mlp_destroy(couplb_mlp* net);
const char* couplb_mlp_last_error(const couplb_mlp* net);

/* host copies of the FP32 parameter vector; set clears the Adam state */
int couplb_mlp_get_params(couplb_mlp* net, float* params);
int couplb_mlp_set_params(couplb_mlp* net, const float* params);

/* Y_dev[n][3] = net(X_dev[n][67]) */
int couplb_mlp_forward_device(couplb_mlp* net, int n, const float* X_dev, float* Y_dev);

/* gradient of the loss over the n samples (no update): grad[8707] on the
 * host, *loss the loss value */
int couplb_mlp_gradient_device(couplb_mlp* net, int n, const float* X_dev, const float* T_dev,
                               float* grad, double* loss);

/* Adam training on n samples (X_dev [n][67], targets T_dev [n][3]): `epochs`
 * passes over a seeded shuffle in mini-batches of `batch`; *loss receives
 * the mean loss of the last epoch's mini-batches. */
int couplb_mlp_train_device(couplb_mlp* net, int n, const float* X_dev, const float* T_dev,
                            int batch, int epochs, float lr, unsigned long long seed,
                            double* loss);

#ifdef __cplusplus
}
#endif
#endif
