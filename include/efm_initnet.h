/*
 * efm_initnet.h — C ABI of the learned moment-tensor / location initializer
 * kernels in libefm.so (gfx950).  No counterpart in the reference: its
 * inversions start from a fixed moment tensor (mainPropElastic3D_MT.py
 * M:2777, Mmod = StressTensor(0.5, 0, 0.5, 0, 0.5, 0)) at the catalogue
 * location; this network proposes both from the recorded gathers.
 *
 * Building blocks of a 1-D convolutional network trained and run on the GPU:
 * a BF16 MFMA GEMM with FP32 accumulation and fused bias/ReLU epilogue,
 * im2col / col2im for strided 1-D convolutions, ReLU and bias backward, MSE
 * loss, Adam, and the packing of propagator gathers into network inputs.
 *
 * BF16 values are stored as uint16 (the upper half of an IEEE float32, rounded
 * to nearest even).  Activations are laid out (C, B, L): channel-major, then
 * example, then sample, so a convolution's GEMM output (Cout, B*Lout) is the
 * next layer's input without a transpose.  All pointers are device pointers;
 * every call is asynchronous on `stream` and returns 0 or a negative EFM_E_*.
 */
#ifndef EFM_INITNET_H
#define EFM_INITNET_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* y[i] = bf16(x[i]). */
int efm_net_cast_bf16(const float *x, uint16_t *y, long long n, void *stream);

/* C[m, n] = sum_k A[m, k] * B[k, n] for m < M, n < N, k < K, where
 * A[m, k] = A[m*sam + k*sak] and B[k, n] = B[k*sbk + n*sbn] (BF16), FP32
 * accumulation, C[m*ldc + n] FP32.  epilogue 0: C = AB; 1: C = max(AB + bias[m], 0);
 * 2: C = AB + bias[m].  Cb (nullable) receives bf16(C) at the same offsets. */
int efm_net_gemm(int M, int N, int K, const uint16_t *A, long long sam, long long sak,
                 const uint16_t *B, long long sbk, long long sbn, float *C, long long ldc,
                 const float *bias, uint16_t *Cb, int epilogue, void *stream);

/* Lout = (L + 2*pad - ks)/stride + 1.
 * col[(c*ks + k), b*Lout + l] = x[c, b, l*stride + k - pad] (0 outside [0, L)). */
int efm_net_im2col(const uint16_t *x, uint16_t *col, int C, int B, int L, int ks, int stride,
                   int pad, void *stream);
/* Transpose of im2col: dx[c, b, j] = sum over (k, l) with l*stride + k - pad = j
 * of dcol[(c*ks + k), b*Lout + l]. */
int efm_net_col2im(const float *dcol, float *dx, int C, int B, int L, int ks, int stride,
                   int pad, void *stream);

/* dpre = dy * (y > 0); dpre_b (nullable) = bf16(dpre). */
int efm_net_relu_backward(const float *dy, const float *y, float *dpre, uint16_t *dpre_b,
                          long long n, void *stream);
/* db[m] = sum_n dpre[m*N + n]. */
int efm_net_bias_grad(const float *dpre, float *db, int M, long long N, void *stream);

/* Mean-squared error over n values: *loss += sum (out - t)^2 * inv_n (a device
 * double the caller zeroes); dout = 2 (out - t) inv_n; dout_b (nullable) = bf16(dout). */
int efm_net_mse(const float *out, const float *target, float *dout, uint16_t *dout_b,
                double *loss, long long n, float inv_n, void *stream);

/* Adam step on n parameters: m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
 * w -= lr * (m*c1) / (sqrt(v*c2) + eps) with c1 = 1/(1-b1^t), c2 = 1/(1-b2^t);
 * wb (nullable) = bf16(w). */
int efm_net_adam(float *w, const float *g, float *m, float *v, uint16_t *wb, long long n,
                 float lr, float b1, float b2, float eps, float c1, float c2, void *stream);

/* X[(c*L + l), b] = act[c, b, l] (flatten for the dense head) and its transpose
 * dact[c, b, l] = dX[(c*L + l), b]. */
int efm_net_flatten(const uint16_t *act, uint16_t *X, int C, int B, int L, void *stream);
int efm_net_unflatten(const float *dX, float *dact, int C, int B, int L, void *stream);

/* Network inputs from propagator gathers (S, ncomp, ntmax, nrec) FP32:
 * data[s0 + s, comp*nrec + r, l] = bf16(g[s, comp, l, r] / max|g[s]|) for
 * l < L (L <= ntmax), written into a (Ntotal, ncomp*nrec, L) dataset. */
int efm_net_pack_gathers(const float *g, int S, int ncomp, int ntmax, int nrec, int L,
                         uint16_t *data, long long s0, void *stream);
/* Mini-batch assembly from a dataset of ntotal examples: x[c, b, l] =
 * data[idx[b], c, l] (C channels, L samples) and t[j, b] = targets[idx[b]*nout + j];
 * an idx[b] outside [0, ntotal) yields zeros for that example. */
int efm_net_gather_batch(const uint16_t *data, const float *targets, long long ntotal,
                         const int *idx, int B, int C, int L, int nout, uint16_t *x, float *t,
                         void *stream);

#ifdef __cplusplus
}
#endif
#endif
