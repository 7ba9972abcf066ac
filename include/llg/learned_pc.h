/*
 * learned_pc.h -- a learned correction to the block-Jacobi preconditioner
 * (three 3x3 convolution layers, 32 channels, BF16 on MFMA with FP32
 * accumulation), with its training kernels.  No reference counterpart.
 *
 * For the Newton system (I - gamma J) x = r the preconditioner applies
 *   u = P_bj^-1 r,   z = u + active * C3 C2 C1 (gamma [u, m x u, (m.u) m])
 * which is linear in r (no activations; m is fixed during a solve), so it
 * is a valid left preconditioner for GMRES.  With zero weights (the initial
 * state) z = u: plain block-Jacobi.
 *
 * Training pairs (u = P_bj^-1 b, x) come from the solver's own FP64
 * operator: x is sampled, b = (I - gamma J) x (LLG_Atimes), so each pair is
 * an exact solution of the Newton system at state m and gamma.
 */
#ifndef LLG_LEARNED_PC_H
#define LLG_LEARNED_PC_H

#include "sundials_compat.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct llg_ctx llg_ctx;
typedef struct llg_lpc llg_lpc;

/* network bound to a context's grid (ng % 16 == 0) and mask; periodic
 * selects periodic padding of the convolutions (else zero padding) */
llg_lpc* LLG_LpcCreate(llg_ctx* ctx, int periodic);
void     LLG_LpcDestroy(llg_lpc* p);
const char* LLG_LpcLastError(void);
/* FP32 master weights: W1 [9][16][32], W2 [9][32][32], W3 [9][32][16] */
int      LLG_LpcNumWeights(llg_lpc* p);
int      LLG_LpcSetWeights(llg_lpc* p, const float* w_host);   /* resets Adam */
int      LLG_LpcGetWeights(llg_lpc* p, float* w_host);
int      LLG_LpcGetGrad(llg_lpc* p, float* g_host);
/* network output (FP32 [cell][4], 3 used) for input u at state y */
int      LLG_LpcForward(llg_lpc* p, const double* u_dev, const double* y_dev, double gamma,
                        double scale, float* out_host);
/* loss = sum_active |scale*(u + N(u)) - scale*x|^2 / ncell and its weight
 * gradient (kept on the device, LLG_LpcGetGrad) */
int      LLG_LpcLossGrad(llg_lpc* p, const double* u_dev, const double* y_dev,
                         const double* x_dev, double gamma, double scale, double* loss);
int      LLG_LpcAdamStep(llg_lpc* p, float lr);
int      LLG_LpcTrainStep(llg_lpc* p, const double* u_dev, const double* y_dev,
                          const double* x_dev, double gamma, double scale, float lr,
                          double* loss);
/* z (holding u = P_bj^-1 r) += active * N(u) at state y */
int      LLG_LpcCorrect(llg_lpc* p, const double* y_dev, double gamma, double* z_dev);
/* z = u + active * N(u) at state y, u and z distinct (3n doubles each) */
int      LLG_LpcApply(llg_lpc* p, const double* y_dev, double gamma, const double* u_dev,
                      double* z_dev);
/* a 3n-double device buffer owned by the network, where PrecondSolve puts
 * the block-Jacobi output before LLG_LpcApply */
double*  LLG_LpcInputBuffer(llg_lpc* p);
/* inference form: 1 (default) one fused kernel per application -- features,
 * the three convolutions on v_mfma_f32_16x16x32_bf16 with the hidden layers
 * in LDS, and the update; 0 the layer-by-layer kernels the training uses
 * (hidden layers through HBM) */
int      LLG_LpcSetFused(llg_lpc* p, int on);

/* PrecondSolve applies the correction after block-Jacobi while set (NULL:
 * plain block-Jacobi, the fallback) */
int      LLG_SetLearnedPrecond(llg_ctx* ctx, llg_lpc* p);

#ifdef __cplusplus
}
#endif
#endif /* LLG_LEARNED_PC_H */
