/* libtonto_gpu: the learned initial guess's network (training and inference kernels).
 *
 * No counterpart in the reference: Tonto starts an SCF from a core or promolecule guess
 * (foofiles/molecule.scf.foo:4723-4759, 4827-4844).  The network predicts density-matrix
 * blocks of atom pairs; the SCF that follows is unchanged, so only the iteration count
 * depends on it.
 *
 * One network = three dense layers in FP32, n_in -> nh -> nh -> n_out, SiLU after the
 * first two.  Parameters are one contiguous float array:
 *   W1 [n_in][nh], b1 [nh], W2 [nh][nh], b2 [nh], W3 [nh][n_out], b3 [n_out]
 * (row-major; y = x W + b for a row vector x).  Inputs X [M][n_in], outputs Y [M][n_out].
 * All pointers are device pointers; stream is a hipStream_t (NULL = default stream).
 * Functions return 0 on success, TMLP_E_* otherwise.
 */
#ifndef TONTO_MLP_H
#define TONTO_MLP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TMLP_OK 0
#define TMLP_E_ARG 1
#define TMLP_E_HIP 2

/* number of floats in the parameter (and gradient) array */
long tmlp_n_params(int n_in, int nh, int n_out);

/* floats of workspace tmlp_forward needs for M rows (pre-activations and activations of
 * the two hidden layers, kept for tmlp_backward) */
long tmlp_forward_workspace(long M, int nh);

/* floats of extra workspace tmlp_backward needs for M rows */
long tmlp_backward_workspace(long M, int nh);

/* Y = f(X): ws (tmlp_forward_workspace floats) receives Z1, A1, Z2, A2 */
int tmlp_forward(const float *params, int n_in, int nh, int n_out, const float *X, long M,
                 float *ws, float *Y, void *stream);

/* grads = d loss / d params for the batch whose forward pass left ws, given dY = d loss / dY
 * (M x n_out).  grads is overwritten.  ws2: tmlp_backward_workspace floats. */
int tmlp_backward(const float *params, int n_in, int nh, int n_out, const float *X, long M,
                  const float *ws, const float *dY, float *grads, float *ws2, void *stream);

/* loss = scale * sum (Y - T)^2 over n entries, written to *loss (device float);
 * dY = 2 scale (Y - T) */
int tmlp_mse(const float *Y, const float *T, long n, float scale, float *dY, float *loss,
             void *stream);

/* one Adam step (Kingma and Ba 2015, bias-corrected; step counts from 1) */
int tmlp_adam(float *params, const float *grads, float *m, float *v, long n, float lr,
              float beta1, float beta2, float eps, int step, void *stream);

#ifdef __cplusplus
}
#endif
#endif
