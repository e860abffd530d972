// Kernels of the learned moment-tensor / location initializer (a 1-D
// convolutional network trained and run on the GPU).  No counterpart in the
// reference, whose inversions start from a fixed moment tensor.
//
// BF16 operands are uint16 upper halves of IEEE float32 (round to nearest
// even); products accumulate in FP32 on the matrix cores with gfx950's
// v_mfma_f32_16x16x32_bf16 (one 32-deep k step per instruction).
// Activations are (C, B, L).
#include "common.hpp"
#include "efm_initnet.h"

namespace efm {
namespace net {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);                      // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((unsigned)h << 16); }

inline int blocks_for(long long n, int bs) {
    long long b = (n + bs - 1) / bs;
    return (int)(b > 65535LL * 8 ? 65535LL * 8 : b);
}

// ---- GEMM ---------------------------------------------------------------------
// Block tile 64 (m) x 64 (n), k-slab 32, 4 waves; wave w owns the 32 x 32
// quadrant (w/2, w%2) as 2 x 2 MFMA 16x16 tiles, one 16x16x32 instruction per
// tile and k-slab.  Operands are staged through LDS as [row][k] (A) and
// [col][k] (B) so that each lane reads its 8 consecutive k values as one
// 16-byte word.  MFMA 16x16x32 lane layout: lane l supplies A[l%16][8(l/16)+j]
// and B[8(l/16)+j][l%16], j < 8, and receives D[4(l/16)+j][l%16], j < 4.
constexpr int GBM = 64, GBN = 64, GBK = 32, GLD = GBK + 8;

// The 2 x 2 MFMA tiles of one k-slab from the LDS stages.
__device__ __forceinline__ void mfma_slab(const uint16_t *As, const uint16_t *Bs, int wm, int wn,
                                          int lane, f32x4 (&acc)[2][2]) {
    bf16x8 a[2], b[2];
    const int ko = 8 * (lane >> 4);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        a[i] = *reinterpret_cast<const bf16x8 *>(&As[(wm + 16 * i + (lane & 15)) * GLD + ko]);
        b[i] = *reinterpret_cast<const bf16x8 *>(&Bs[(wn + 16 * i + (lane & 15)) * GLD + ko]);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
}

__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const uint16_t *__restrict__ A,
                                              long long sam, long long sak,
                                              const uint16_t *__restrict__ B, long long sbk,
                                              long long sbn, float *__restrict__ C, long long ldc,
                                              const float *__restrict__ bias,
                                              uint16_t *__restrict__ Cb, int epi) {
    __shared__ __attribute__((aligned(16))) uint16_t As[GBM * GLD];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[GBN * GLD];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.y * GBM, n0 = blockIdx.x * GBN;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
    const bool a_kfast = (sak == 1);  // choose the staging order that coalesces
    const bool b_nfast = (sbn == 1);
    for (int k0 = 0; k0 < K; k0 += GBK) {
#pragma unroll
        for (int i = 0; i < (GBM * GBK) / 256; ++i) {
            const int idx = t + 256 * i;
            int r, k;
            if (a_kfast) { r = idx >> 5; k = idx & 31; } else { r = idx & 63; k = idx >> 6; }
            const int gm = m0 + r, gk = k0 + k;
            uint16_t v = 0;
            if (gm < M && gk < K) v = A[(long long)gm * sam + (long long)gk * sak];
            As[r * GLD + k] = v;
        }
#pragma unroll
        for (int i = 0; i < (GBN * GBK) / 256; ++i) {
            const int idx = t + 256 * i;
            int c, k;
            if (b_nfast) { c = idx & 63; k = idx >> 6; } else { c = idx >> 5; k = idx & 31; }
            const int gn = n0 + c, gk = k0 + k;
            uint16_t v = 0;
            if (gn < N && gk < K) v = B[(long long)gk * sbk + (long long)gn * sbn];
            Bs[c * GLD + k] = v;
        }
        __syncthreads();
        mfma_slab(As, Bs, wm, wn, lane, acc);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
                const int col = n0 + wn + 16 * j + (lane & 15);
                if (row < M && col < N) {
                    float v = acc[i][j][r];
                    if (epi >= 1) v += bias[row];
                    if (epi == 1) v = fmaxf(v, 0.f);
                    const long long o = (long long)row * ldc + col;
                    C[o] = v;
                    if (Cb) Cb[o] = f2bf(v);
                }
            }
}

// Vectorised variant for operands whose contiguous dimension has unit stride
// (A: k-fast when AK, else m-fast; B: n-fast when BN, else k-fast) and whose
// other stride is a multiple of 8 elements: each thread moves one 16-byte
// vector (8 values) of each operand per k-slab, and the next slab is fetched
// into registers while the matrix cores work on the current one (two LDS
// stages).  Vectors that cross the M/N/K edge fall back to element loads.
template <bool AK, bool BN>
__device__ __forceinline__ void gemm_fetch(uint4 &va, uint4 &vb, int t, int m0, int n0, int k0,
                                           int M, int N, int K, const uint16_t *__restrict__ A,
                                           long long lda, const uint16_t *__restrict__ B,
                                           long long ldb) {
    // A: AK -> row m = t/4, k = 8(t%4); else k = t/8, m = 8(t%8)
    {
        const int r = AK ? (t >> 2) : ((t & 7) << 3), k = AK ? ((t & 3) << 3) : (t >> 3);
        const int gm = m0 + r, gk = k0 + k;
        if (AK ? (gm < M && gk + 7 < K) : (gk < K && gm + 7 < M)) {
            va = *reinterpret_cast<const uint4 *>(AK ? A + (long long)gm * lda + gk
                                                     : A + (long long)gk * lda + gm);
        } else {
            uint16_t e[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int mm = AK ? gm : gm + i, kk = AK ? gk + i : gk;
                e[i] = (mm < M && kk < K) ? (AK ? A[(long long)mm * lda + kk] : A[(long long)kk * lda + mm])
                                          : (uint16_t)0;
            }
            va = make_uint4(e[0] | (unsigned)e[1] << 16, e[2] | (unsigned)e[3] << 16,
                            e[4] | (unsigned)e[5] << 16, e[6] | (unsigned)e[7] << 16);
        }
    }
    // B: BN -> k = t/8, n = 8(t%8); else n = t/4, k = 8(t%4)
    {
        const int c = BN ? ((t & 7) << 3) : (t >> 2), k = BN ? (t >> 3) : ((t & 3) << 3);
        const int gn = n0 + c, gk = k0 + k;
        if (BN ? (gk < K && gn + 7 < N) : (gn < N && gk + 7 < K)) {
            vb = *reinterpret_cast<const uint4 *>(BN ? B + (long long)gk * ldb + gn
                                                     : B + (long long)gn * ldb + gk);
        } else {
            uint16_t e[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int nn = BN ? gn + i : gn, kk = BN ? gk : gk + i;
                e[i] = (nn < N && kk < K) ? (BN ? B[(long long)kk * ldb + nn] : B[(long long)nn * ldb + kk])
                                          : (uint16_t)0;
            }
            vb = make_uint4(e[0] | (unsigned)e[1] << 16, e[2] | (unsigned)e[3] << 16,
                            e[4] | (unsigned)e[5] << 16, e[6] | (unsigned)e[7] << 16);
        }
    }
}

template <bool AK, bool BN>
__device__ __forceinline__ void gemm_stage(uint16_t *As, uint16_t *Bs, const uint4 &va,
                                           const uint4 &vb, int t) {
    if (AK) {
        *reinterpret_cast<uint4 *>(&As[(t >> 2) * GLD + ((t & 3) << 3)]) = va;
    } else {
        const uint16_t *e = reinterpret_cast<const uint16_t *>(&va);
        const int m = (t & 7) << 3, k = t >> 3;
#pragma unroll
        for (int i = 0; i < 8; ++i) As[(m + i) * GLD + k] = e[i];
    }
    if (BN) {
        const uint16_t *e = reinterpret_cast<const uint16_t *>(&vb);
        const int n = (t & 7) << 3, k = t >> 3;
#pragma unroll
        for (int i = 0; i < 8; ++i) Bs[(n + i) * GLD + k] = e[i];
    } else {
        *reinterpret_cast<uint4 *>(&Bs[(t >> 2) * GLD + ((t & 3) << 3)]) = vb;
    }
}

template <bool AK, bool BN>
__global__ __launch_bounds__(256) void k_gemm_vec(int M, int N, int K,
                                                  const uint16_t *__restrict__ A, long long lda,
                                                  const uint16_t *__restrict__ B, long long ldb,
                                                  float *__restrict__ C, long long ldc,
                                                  const float *__restrict__ bias,
                                                  uint16_t *__restrict__ Cb, int epi) {
    __shared__ __attribute__((aligned(16))) uint16_t As[2][GBM * GLD];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][GBN * GLD];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.y * GBM, n0 = blockIdx.x * GBN;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
    uint4 va, vb;
    gemm_fetch<AK, BN>(va, vb, t, m0, n0, 0, M, N, K, A, lda, B, ldb);
    gemm_stage<AK, BN>(As[0], Bs[0], va, vb, t);
    __syncthreads();
    int cur = 0;
    for (int k0 = 0; k0 < K; k0 += GBK) {
        const bool more = k0 + GBK < K;
        if (more) gemm_fetch<AK, BN>(va, vb, t, m0, n0, k0 + GBK, M, N, K, A, lda, B, ldb);
        mfma_slab(As[cur], Bs[cur], wm, wn, lane, acc);
        if (more) gemm_stage<AK, BN>(As[cur ^ 1], Bs[cur ^ 1], va, vb, t);
        __syncthreads();
        cur ^= 1;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
                const int col = n0 + wn + 16 * j + (lane & 15);
                if (row < M && col < N) {
                    float v = acc[i][j][r];
                    if (epi >= 1) v += bias[row];
                    if (epi == 1) v = fmaxf(v, 0.f);
                    const long long o = (long long)row * ldc + col;
                    C[o] = v;
                    if (Cb) Cb[o] = f2bf(v);
                }
            }
}

// ---- elementwise / layout kernels -----------------------------------------------
__global__ void k_cast(const float *__restrict__ x, uint16_t *__restrict__ y, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        y[i] = f2bf(x[i]);
}

__global__ void k_im2col(const uint16_t *__restrict__ x, uint16_t *__restrict__ col, int C, int B,
                         int L, int ks, int stride, int pad, int Lout) {
    const long long ncol = (long long)B * Lout;
    const long long n = (long long)C * ks * ncol;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const long long row = i / ncol, cc = i - row * ncol;
        const int c = (int)(row / ks), k = (int)(row - (long long)c * ks);
        const int b = (int)(cc / Lout), l = (int)(cc - (long long)b * Lout);
        const int p = l * stride + k - pad;
        col[i] = (p >= 0 && p < L) ? x[((long long)c * B + b) * L + p] : (uint16_t)0;
    }
}

__global__ void k_col2im(const float *__restrict__ dcol, float *__restrict__ dx, int C, int B, int L,
                         int ks, int stride, int pad, int Lout) {
    const long long n = (long long)C * B * L;
    const long long ncol = (long long)B * Lout;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(i % L);
        const long long cb = i / L;
        const int b = (int)(cb % B), c = (int)(cb / B);
        float s = 0.f;
        for (int k = 0; k < ks; ++k) {
            const int q = j + pad - k;
            if (q < 0 || q % stride) continue;
            const int l = q / stride;
            if (l >= Lout) continue;
            s += dcol[((long long)c * ks + k) * ncol + (long long)b * Lout + l];
        }
        dx[i] = s;
    }
}

__global__ void k_relu_bwd(const float *__restrict__ dy, const float *__restrict__ y,
                           float *__restrict__ dp, uint16_t *__restrict__ dpb, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const float v = y[i] > 0.f ? dy[i] : 0.f;
        dp[i] = v;
        if (dpb) dpb[i] = f2bf(v);
    }
}

__global__ __launch_bounds__(256) void k_bias_grad(const float *__restrict__ dp,
                                                   float *__restrict__ db, long long N) {
    __shared__ float red[256];
    const float *row = dp + (long long)blockIdx.x * N;
    float s = 0.f;
    for (long long i = threadIdx.x; i < N; i += 256) s += row[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ out,
                                             const float *__restrict__ tg, float *__restrict__ d,
                                             uint16_t *__restrict__ db, double *loss, long long n,
                                             float inv_n) {
    __shared__ double red[256];
    double s = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const float r = out[i] - tg[i];
        s += (double)r * (double)r;
        const float g = 2.f * r * inv_n;
        d[i] = g;
        if (db) db[i] = f2bf(g);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] * (double)inv_n);
}

__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, uint16_t *__restrict__ wb, long long n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
        w[i] = wi;
        if (wb) wb[i] = f2bf(wi);
    }
}

__global__ void k_flatten(const uint16_t *__restrict__ a, uint16_t *__restrict__ X, int C, int B,
                          int L) {
    const long long n = (long long)C * B * L;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        // i indexes X[(c*L + l), b]
        const int b = (int)(i % B);
        const long long f = i / B;
        const int c = (int)(f / L), l = (int)(f - (long long)c * L);
        X[i] = a[((long long)c * B + b) * L + l];
    }
}

__global__ void k_unflatten(const float *__restrict__ dX, float *__restrict__ da, int C, int B,
                            int L) {
    const long long n = (long long)C * B * L;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        // i indexes da[c, b, l]
        const int l = (int)(i % L);
        const long long cb = i / L;
        const int b = (int)(cb % B), c = (int)(cb / B);
        da[i] = dX[((long long)c * L + l) * B + b];
    }
}

// One block per shot: max |g| over the shot's gathers, then the normalised,
// channel-major copy.
__global__ __launch_bounds__(256) void k_pack(const float *__restrict__ g, int ncomp, int ntmax,
                                              int nrec, int L, uint16_t *__restrict__ data,
                                              long long s0) {
    __shared__ float red[256];
    const int s = blockIdx.x;
    const long long per = (long long)ncomp * ntmax * nrec;
    const float *gs = g + (long long)s * per;
    float mx = 0.f;
    for (long long i = threadIdx.x; i < per; i += 256) mx = fmaxf(mx, fabsf(gs[i]));
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    const float inv = red[0] > 0.f ? 1.f / red[0] : 0.f;
    const int C = ncomp * nrec;
    uint16_t *ds = data + (s0 + s) * (long long)C * L;
    const long long n = (long long)C * L;
    for (long long i = threadIdx.x; i < n; i += 256) {
        const int l = (int)(i % L);
        const int c = (int)(i / L);
        const int comp = c / nrec, r = c - comp * nrec;
        ds[i] = f2bf(gs[((long long)comp * ntmax + l) * nrec + r] * inv);
    }
}

__global__ void k_gather_batch(const uint16_t *__restrict__ data, const float *__restrict__ tg,
                               long long ntotal, const int *__restrict__ idx, int B, int C, int L,
                               int nout, uint16_t *__restrict__ x, float *__restrict__ t) {
    const long long n = (long long)C * B * L;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int l = (int)(i % L);
        const long long cb = i / L;
        const int b = (int)(cb % B), c = (int)(cb / B);
        const long long e = idx[b];
        x[i] = (e >= 0 && e < ntotal) ? data[(e * C + c) * L + l] : (uint16_t)0;
        if (i < (long long)nout * B) {
            const int j = (int)(i / B), bb = (int)(i % B);
            const long long eb = idx[bb];
            t[i] = (eb >= 0 && eb < ntotal) ? tg[eb * nout + j] : 0.f;
        }
    }
}

}  // namespace net
}  // namespace efm

using namespace efm;
using namespace efm::net;

#define NET_CHECK(cond, msg) \
    if (!(cond)) return fail(EFM_E_ARG, msg)

extern "C" int efm_net_cast_bf16(const float *x, uint16_t *y, long long n, void *stream) {
    NET_CHECK(x && y && n >= 0, "efm_net_cast_bf16: bad arguments");
    if (n == 0) return EFM_OK;
    k_cast<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(x, y, n);
    return check_launch("efm_net_cast_bf16");
}

extern "C" int efm_net_gemm(int M, int N, int K, const uint16_t *A, long long sam, long long sak,
                            const uint16_t *B, long long sbk, long long sbn, float *C,
                            long long ldc, const float *bias, uint16_t *Cb, int epilogue,
                            void *stream) {
    NET_CHECK(A && B && C && M > 0 && N > 0 && K > 0, "efm_net_gemm: null pointer or empty shape");
    NET_CHECK(sam >= 0 && sak >= 0 && sbk >= 0 && sbn >= 0 && ldc >= N,
              "efm_net_gemm: bad strides");
    NET_CHECK(epilogue >= 0 && epilogue <= 2, "efm_net_gemm: epilogue must be 0, 1 or 2");
    NET_CHECK(epilogue == 0 || bias, "efm_net_gemm: bias epilogue needs a bias");
    NET_CHECK((M + GBM - 1) / GBM <= 65535, "efm_net_gemm: M too large");
    dim3 grid((N + GBN - 1) / GBN, (M + GBM - 1) / GBM);
    hipStream_t st = (hipStream_t)stream;
    // vectorised path: unit stride along one dimension of each operand, the
    // other stride a multiple of 8, 16-byte aligned bases
    const bool a_ok = ((sak == 1 && sam % 8 == 0) || (sam == 1 && sak % 8 == 0)) &&
                      ((uintptr_t)A % 16 == 0) && !(sak == 1 && sam == 1 && M > 1 && K > 1);
    const bool b_ok = ((sbn == 1 && sbk % 8 == 0) || (sbk == 1 && sbn % 8 == 0)) &&
                      ((uintptr_t)B % 16 == 0) && !(sbn == 1 && sbk == 1 && N > 1 && K > 1);
    if (a_ok && b_ok) {
        const bool ak = (sak == 1), bn = (sbn == 1);
        const long long lda = ak ? sam : sak, ldb = bn ? sbk : sbn;
        if (ak && bn)
            k_gemm_vec<true, true><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, bias, Cb, epilogue);
        else if (ak)
            k_gemm_vec<true, false><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, bias, Cb, epilogue);
        else if (bn)
            k_gemm_vec<false, true><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, bias, Cb, epilogue);
        else
            k_gemm_vec<false, false><<<grid, 256, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, bias, Cb, epilogue);
        return check_launch("efm_net_gemm");
    }
    k_gemm<<<grid, 256, 0, st>>>(M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, bias, Cb, epilogue);
    return check_launch("efm_net_gemm");
}

static bool conv_ok(int C, int B, int L, int ks, int stride, int pad) {
    return C > 0 && B > 0 && L > 0 && ks > 0 && stride > 0 && pad >= 0 && L + 2 * pad >= ks;
}

extern "C" int efm_net_im2col(const uint16_t *x, uint16_t *col, int C, int B, int L, int ks,
                              int stride, int pad, void *stream) {
    NET_CHECK(x && col && conv_ok(C, B, L, ks, stride, pad), "efm_net_im2col: bad arguments");
    const int Lout = (L + 2 * pad - ks) / stride + 1;
    const long long n = (long long)C * ks * B * Lout;
    k_im2col<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(x, col, C, B, L, ks, stride, pad,
                                                                    Lout);
    return check_launch("efm_net_im2col");
}

extern "C" int efm_net_col2im(const float *dcol, float *dx, int C, int B, int L, int ks,
                              int stride, int pad, void *stream) {
    NET_CHECK(dcol && dx && conv_ok(C, B, L, ks, stride, pad), "efm_net_col2im: bad arguments");
    const int Lout = (L + 2 * pad - ks) / stride + 1;
    const long long n = (long long)C * B * L;
    k_col2im<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(dcol, dx, C, B, L, ks, stride,
                                                                    pad, Lout);
    return check_launch("efm_net_col2im");
}

extern "C" int efm_net_relu_backward(const float *dy, const float *y, float *dpre,
                                     uint16_t *dpre_b, long long n, void *stream) {
    NET_CHECK(dy && y && dpre && n >= 0, "efm_net_relu_backward: bad arguments");
    if (n == 0) return EFM_OK;
    k_relu_bwd<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(dy, y, dpre, dpre_b, n);
    return check_launch("efm_net_relu_backward");
}

extern "C" int efm_net_bias_grad(const float *dpre, float *db, int M, long long N, void *stream) {
    NET_CHECK(dpre && db && M > 0 && M <= 65535 * 8 && N > 0, "efm_net_bias_grad: bad arguments");
    k_bias_grad<<<M, 256, 0, (hipStream_t)stream>>>(dpre, db, N);
    return check_launch("efm_net_bias_grad");
}

extern "C" int efm_net_mse(const float *out, const float *target, float *dout, uint16_t *dout_b,
                           double *loss, long long n, float inv_n, void *stream) {
    NET_CHECK(out && target && dout && loss && n > 0, "efm_net_mse: bad arguments");
    k_mse<<<blocks_for(n, 256) < 1024 ? blocks_for(n, 256) : 1024, 256, 0, (hipStream_t)stream>>>(
        out, target, dout, dout_b, loss, n, inv_n);
    return check_launch("efm_net_mse");
}

extern "C" int efm_net_adam(float *w, const float *g, float *m, float *v, uint16_t *wb,
                            long long n, float lr, float b1, float b2, float eps, float c1,
                            float c2, void *stream) {
    NET_CHECK(w && g && m && v && n > 0, "efm_net_adam: bad arguments");
    k_adam<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(w, g, m, v, wb, n, lr, b1, b2, eps,
                                                                  c1, c2);
    return check_launch("efm_net_adam");
}

extern "C" int efm_net_flatten(const uint16_t *act, uint16_t *X, int C, int B, int L,
                               void *stream) {
    NET_CHECK(act && X && C > 0 && B > 0 && L > 0, "efm_net_flatten: bad arguments");
    const long long n = (long long)C * B * L;
    k_flatten<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(act, X, C, B, L);
    return check_launch("efm_net_flatten");
}

extern "C" int efm_net_unflatten(const float *dX, float *dact, int C, int B, int L, void *stream) {
    NET_CHECK(dX && dact && C > 0 && B > 0 && L > 0, "efm_net_unflatten: bad arguments");
    const long long n = (long long)C * B * L;
    k_unflatten<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(dX, dact, C, B, L);
    return check_launch("efm_net_unflatten");
}

extern "C" int efm_net_pack_gathers(const float *g, int S, int ncomp, int ntmax, int nrec, int L,
                                    uint16_t *data, long long s0, void *stream) {
    NET_CHECK(g && data && S > 0 && S <= 65535 * 8 && ncomp > 0 && nrec > 0 && L > 0 &&
                  L <= ntmax && s0 >= 0,
              "efm_net_pack_gathers: bad arguments");
    k_pack<<<S, 256, 0, (hipStream_t)stream>>>(g, ncomp, ntmax, nrec, L, data, s0);
    return check_launch("efm_net_pack_gathers");
}

extern "C" int efm_net_gather_batch(const uint16_t *data, const float *targets, long long ntotal,
                                    const int *idx, int B, int C, int L, int nout, uint16_t *x,
                                    float *t, void *stream) {
    NET_CHECK(data && targets && idx && x && t && ntotal > 0 && B > 0 && C > 0 && L > 0 &&
                  nout > 0 && (long long)nout * B <= (long long)C * B * L,
              "efm_net_gather_batch: bad arguments");
    const long long n = (long long)C * B * L;
    k_gather_batch<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(data, targets, ntotal, idx,
                                                                          B, C, L, nout, x, t);
    return check_launch("efm_net_gather_batch");
}
