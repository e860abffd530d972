// Gradient conditioning and model maps of the FWI driver, on the device.
//
// Reference: fwi_Vp_Vs_Rho (mainPropElastic3D_MT.py M:2349-2719) smooths the
// summed λ/μ/ρ images with scipy.ndimage.gaussian_filter(σ = 1; defaults
// mode='reflect', truncate=4.0) and zeroes the top 10 cells and the CPML+2
// zones (M:2479-2504), then updates λ and μ, maps them back to Vp and Vs
// (Model.compute_Vp_Vs, M:1200-1202), ρ from Vp by the Nafe-Drake polynomial
// (Model.compute_Rho, M:1204-1206) and restores ρ in the top 10 cells
// (M:2580-2583, 2656-2661, 2688-2694).  The reference does all of this with
// numpy/scipy on the host; here the volumes stay in HBM.
//
// Smoothing: one pass per axis in scipy's order (axis 0 = y, 1 = z, 2 = x of
// the (ny, nz, nx) volume), each output cell summed in double over its
// 2r+1 neighbours in scipy's order (centre first, then the symmetric pairs
// from the outermost inwards) and rounded to float32 per pass, as
// scipy.ndimage.correlate1d does for float32 arrays; 'reflect' extends a line
// as (d c b a | a b c d | d c b a).  The masks are applied while the last pass
// writes.  All passes are HBM-bound streaming kernels (8 B per cell and pass;
// the neighbours come from the caches).
#include <cmath>

#include "common.hpp"

namespace efm {

constexpr int kMaxRadius = 64;

struct GaussTaps {
    double w[kMaxRadius + 1];   // w[j] for j = 0 .. r (symmetric kernel)
    int r;
};

// scipy.ndimage 'reflect' index for a line of length n (any j).
__device__ __forceinline__ int reflect_index(int j, int n) {
    const int p = 2 * n;
    int m = j % p;
    if (m < 0) m += p;
    return m >= n ? p - 1 - m : m;
}

struct Margins {
    int xl, xh, yl, yh, zl, zh;
};

__device__ __forceinline__ bool masked(const Margins &mg, int x, int y, int z, int nx, int ny, int nz) {
    return x < mg.xl || x >= nx - mg.xh || y < mg.yl || y >= ny - mg.yh || z < mg.zl || z >= nz - mg.zh;
}

// One separable pass along `axis` (0 = y, 1 = z, 2 = x); mask applied when
// `mask` is set (the last pass).
__global__ void k_gauss_pass(const float *__restrict__ in, float *__restrict__ out, int nx, int ny,
                             int nz, int axis, GaussTaps taps, int mask, Margins mg) {
#pragma clang fp contract(off)
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long N = (long long)nx * ny * nz;
    if (i >= N) return;
    const int x = (int)(i % nx);
    const int z = (int)((i / nx) % nz);
    const int y = (int)(i / ((long long)nx * nz));
    if (mask && masked(mg, x, y, z, nx, ny, nz)) {
        out[i] = 0.f;
        return;
    }
    int n, c;
    long long stride, base;
    if (axis == 0) { n = ny; c = y; stride = (long long)nx * nz; base = i - (long long)y * stride; }
    else if (axis == 1) { n = nz; c = z; stride = nx; base = i - (long long)z * nx; }
    else { n = nx; c = x; stride = 1; base = i - x; }
    double acc = (double)in[base + (long long)c * stride] * taps.w[0];
    for (int j = taps.r; j >= 1; --j) {
        const double a = (double)in[base + (long long)reflect_index(c - j, n) * stride];
        const double b = (double)in[base + (long long)reflect_index(c + j, n) * stride];
        acc += (a + b) * taps.w[j];
    }
    out[i] = (float)acc;
}

__global__ void k_mask_copy(const float *__restrict__ in, float *__restrict__ out, int nx, int ny,
                            int nz, Margins mg) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long N = (long long)nx * ny * nz;
    if (i >= N) return;
    const int x = (int)(i % nx);
    const int z = (int)((i / nx) % nz);
    const int y = (int)(i / ((long long)nx * nz));
    out[i] = masked(mg, x, y, z, nx, ny, nz) ? 0.f : in[i];
}

// Model.compute_Rho (M:1204-1206) in the reference's float32 numpy arithmetic:
// v = vp/1e3, the five terms each a float32 product with the powers of v
// (v² as v·v, v³..v⁵ rounded once from exact double products, as a correctly
// rounded powf gives), summed left to right in float32, times 1e3.
__device__ __forceinline__ float nafe_drake_f32(float vp) {
#pragma clang fp contract(off)
    const float v = vp / 1000.0f;
    const double vd = (double)v;
    const float v2 = v * v;
    const float v3 = (float)(vd * vd * vd);
    const float v4 = (float)((vd * vd) * (vd * vd));
    const float v5 = (float)((vd * vd) * (vd * vd) * vd);
    float r = 1.6612f * v - 0.4712f * v2;
    r = r + 0.0671f * v3;
    r = r - 0.0043f * v4;
    r = r + 0.000106f * v5;
    return r * 1000.0f;
}

// lam = f32(lam0 + α_λ·dlam), mu likewise (double, one rounding: the
// reference's float32 model plus a float64 step); vp, vs from (lam, mu, rho_in) in
// float32 (compute_Vp_Vs); rho = Nafe-Drake(vp) (compute_Rho) except at
// z < top, where rho_keep is kept.
__global__ void k_model_update(const float *lam0, const float *dlam, double alam, const float *mu0,
                               const float *dmu, double amu, const float *rho_in,
                               const float *rho_keep, int top, int nx, int nz, long long N,
                               float *lam, float *mu, float *vp, float *vs, float *rho) {
#pragma clang fp contract(off)
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float L = dlam ? (float)((double)lam0[i] + alam * (double)dlam[i]) : lam0[i];
    const float M = dmu ? (float)((double)mu0[i] + amu * (double)dmu[i]) : mu0[i];
    const float R = rho_in[i];
    const float p = sqrtf((L + 2.0f * M) / R);
    const float s = sqrtf(M / R);
    lam[i] = L;
    mu[i] = M;
    vp[i] = p;
    vs[i] = s;
    if (rho) {
        const int z = (int)((i / nx) % nz);
        if (z < top && rho_keep) {
            rho[i] = rho_keep[i];
        } else {
            rho[i] = nafe_drake_f32(p);
        }
    }
}

__global__ void k_rho_nafe_drake(const float *vp, float *rho, long long n) {
#pragma clang fp contract(off)
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    rho[i] = nafe_drake_f32(vp[i]);
}

__global__ void k_axpy(float a, const float *x, float *y, long long n) {
#pragma clang fp contract(off)
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = y[i] + a * x[i];
}

// Cross-correlation residual of one trace per thread (computeResidualCrossCorrelation,
// M:2145-2179): lanes hold consecutive receivers, so each time sample is one
// coalesced row read; sums in double.  J += the trace's 1 - a/(|m||o|).
__global__ void k_xcorr(const float *__restrict__ mod, const float *__restrict__ obs,
                        float *__restrict__ res, double *J, long long ntraces, int ntmax, int nrec) {
    const long long tr = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tr >= ntraces) return;
    const long long sc = tr / nrec;
    const int r = (int)(tr - sc * nrec);
    const long long base = sc * ntmax * nrec + r;
    double mm = 0.0, oo = 0.0, mo = 0.0;
    for (int t = 0; t < ntmax; ++t) {
        const double m = mod[base + (long long)t * nrec], o = obs[base + (long long)t * nrec];
        mm += m * m;
        oo += o * o;
        mo += m * o;
    }
    double nm = sqrt(mm), no = sqrt(oo);
    if (nm == 0.0) nm = 1.0;
    if (no == 0.0) no = 1.0;
    const double ag = mo == 0.0 ? 1.0 : mo;
    const double k1 = ag / (no * nm * nm * nm), k2 = 1.0 / (nm * no);
    for (int t = 0; t < ntmax; ++t) {
        const long long i = base + (long long)t * nrec;
        res[i] = (float)((double)mod[i] * k1 - (double)obs[i] * k2);
    }
    atomicAdd(J, 1.0 - mo / (nm * no));
}

static inline unsigned nblocks(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace efm

using namespace efm;

extern "C" int efm_condition_gradient(const EfmSmooth *p, int nx, int ny, int nz, const float *in,
                                      float *out, float *work, void *stream) {
    if (!p || !in || !out) return fail(EFM_E_ARG, "efm_condition_gradient: null argument");
    if (nx < 1 || ny < 1 || nz < 1) return fail(EFM_E_ARG, "efm_condition_gradient: empty volume");
    if (in == out) return fail(EFM_E_ARG, "efm_condition_gradient: in and out must differ");
    if (!(p->sigma >= 0.f) || !(p->truncate >= 0.f))
        return fail(EFM_E_ARG, "efm_condition_gradient: sigma and truncate must be >= 0");
    for (int k = 0; k < 6; ++k)
        if (p->margin[k] < 0) return fail(EFM_E_ARG, "efm_condition_gradient: negative margin");
    const long long N = (long long)nx * ny * nz;
    const Margins mg{p->margin[0], p->margin[1], p->margin[2], p->margin[3], p->margin[4],
                     p->margin[5]};
    hipStream_t st = (hipStream_t)stream;
    // scipy: radius = int(truncate·σ + 0.5); no filtering when σ is (nearly) zero
    const int r = p->sigma > 1e-15f ? (int)((double)p->truncate * (double)p->sigma + 0.5) : 0;
    if (r > kMaxRadius) return fail(EFM_E_UNSUPPORTED, "efm_condition_gradient: kernel radius above 64");
    if (p->sigma <= 1e-15f || r == 0) {
        // a radius-0 Gaussian is the identity (its single weight normalises to 1)
        hipLaunchKernelGGL(k_mask_copy, dim3(nblocks(N)), dim3(256), 0, st, in, out, nx, ny, nz, mg);
        return check_launch("gradient mask");
    }
    if (!work || work == in || work == out)
        return fail(EFM_E_ARG, "efm_condition_gradient: a separate work volume is required");
    GaussTaps taps{};
    taps.r = r;
    // scipy.ndimage._gaussian_kernel1d: exp(-x²/(2σ²)) over x = -r..r, normalised
    const double s2 = (double)p->sigma * (double)p->sigma;
    double phi[2 * kMaxRadius + 1], sum = 0.0;
    for (int j = -r; j <= r; ++j) sum += (phi[j + r] = std::exp(-0.5 / s2 * (double)j * (double)j));
    for (int j = 0; j <= r; ++j) taps.w[j] = phi[j + r] / sum;
    const Margins none{0, 0, 0, 0, 0, 0};
    hipLaunchKernelGGL(k_gauss_pass, dim3(nblocks(N)), dim3(256), 0, st, in, out, nx, ny, nz, 0, taps, 0, none);
    if (int rc = check_launch("gaussian pass y")) return rc;
    hipLaunchKernelGGL(k_gauss_pass, dim3(nblocks(N)), dim3(256), 0, st, out, work, nx, ny, nz, 1, taps, 0, none);
    if (int rc = check_launch("gaussian pass z")) return rc;
    hipLaunchKernelGGL(k_gauss_pass, dim3(nblocks(N)), dim3(256), 0, st, work, out, nx, ny, nz, 2, taps, 1, mg);
    return check_launch("gaussian pass x");
}

extern "C" int efm_model_update(const float *lam0, const float *dlam, double alpha_lam,
                                const float *mu0, const float *dmu, double alpha_mu,
                                const float *rho_in,
                                const float *rho_keep, int top, int nx, int ny, int nz, float *lam,
                                float *mu, float *vp, float *vs, float *rho, void *stream) {
    if (!lam0 || !mu0 || !rho_in || !lam || !mu || !vp || !vs)
        return fail(EFM_E_ARG, "efm_model_update: null argument");
    if (nx < 1 || ny < 1 || nz < 1 || top < 0) return fail(EFM_E_ARG, "efm_model_update: bad size");
    if (rho && (rho == rho_in || (rho_keep && rho == rho_keep)))
        return fail(EFM_E_ARG, "efm_model_update: rho must not alias rho_in or rho_keep");
    const long long N = (long long)nx * ny * nz;
    hipLaunchKernelGGL(k_model_update, dim3(nblocks(N)), dim3(256), 0, (hipStream_t)stream, lam0,
                       dlam, alpha_lam, mu0, dmu, alpha_mu, rho_in, rho_keep, top, nx, nz, N, lam, mu, vp, vs, rho);
    return check_launch("model update");
}

extern "C" int efm_model_rho_nafe_drake(const float *vp, float *rho, long long n, void *stream) {
    if (!vp || !rho || n < 0) return fail(EFM_E_ARG, "efm_model_rho_nafe_drake: bad argument");
    if (n == 0) return EFM_OK;
    hipLaunchKernelGGL(k_rho_nafe_drake, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, vp, rho, n);
    return check_launch("rho nafe-drake");
}

extern "C" int efm_volume_axpy(float a, const float *x, float *y, long long n, void *stream) {
    if (!x || !y || n < 0) return fail(EFM_E_ARG, "efm_volume_axpy: bad argument");
    if (n == 0) return EFM_OK;
    hipLaunchKernelGGL(k_axpy, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, a, x, y, n);
    return check_launch("axpy");
}

extern "C" int efm_xcorr_residual(const float *mod, const float *obs, float *res, double *J,
                                  int nshot, int ntmax, int nrec, void *stream) {
    if (nshot < 0 || ntmax < 0 || nrec < 0) return fail(EFM_E_ARG, "efm_xcorr_residual: bad size");
    const long long ntr = 3LL * nshot * nrec;
    if (ntr == 0 || ntmax == 0) return EFM_OK;
    if (!mod || !obs || !res || !J) return fail(EFM_E_ARG, "efm_xcorr_residual: null argument");
    hipLaunchKernelGGL(k_xcorr, dim3(nblocks(ntr)), dim3(256), 0, (hipStream_t)stream, mod, obs, res, J,
                       ntr, ntmax, nrec);
    return check_launch("cross-correlation residual");
}
