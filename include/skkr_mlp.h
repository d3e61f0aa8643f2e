/*
 * skkr_mlp.h -- C-ABI of the error-prediction MLP of the adaptive BZ mesh.
 *
 * No reference counterpart: the reference integrates over fixed k-sets
 * (kmesh2d.f).  A 3-layer perceptron (nin -> nhid tanh -> nhid tanh -> 1,
 * FP32) predicts the log10 local integration error of a BZ triangle from
 * features of its vertices; forward, backward and Adam kernels are HIP
 * (gfx950), training runs on the GPU.
 *
 * Parameter vector layout (float): W1[nhid][nin], b1[nhid], W2[nhid][nhid],
 * b2[nhid], w3[nhid], b3.  Inputs are row-major n x nin.
 */
#ifndef SKKR_MLP_H
#define SKKR_MLP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct skkr_mlp skkr_mlp;

int skkr_mlp_create(skkr_mlp **m, int nin, int nhid, int device, unsigned long long seed);
void skkr_mlp_destroy(skkr_mlp *m);
int skkr_mlp_nparams(const skkr_mlp *m);
int skkr_mlp_get_params(skkr_mlp *m, float *p);
int skkr_mlp_set_params(skkr_mlp *m, const float *p);
/* y[n] = f(x[n][nin]) */
int skkr_mlp_predict(skkr_mlp *m, int n, const float *x, float *y);
/* mean squared error over n samples and its gradient (nparams) */
int skkr_mlp_loss_grad(skkr_mlp *m, int n, const float *x, const float *t, float *loss,
                       float *grad);
/* full-batch Adam for `steps` steps; loss_history (steps, optional) */
int skkr_mlp_train(skkr_mlp *m, int n, const float *x, const float *t, int steps, float lr,
                   float *loss_history);

#ifdef __cplusplus
}
#endif
#endif /* SKKR_MLP_H */
