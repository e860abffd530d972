// Reference-style adjoint and imaging, for drop-in parity with
// Elastic3D.backwardFields (M:2056-2082) and
// Elastic3D.backwardFields_FowardFieldsRemake (M:1929-2054).
//
// The reference runs the forward operator (+dt) on a zeroed adjoint state with
// the residual injected time-reversed into v (kernel_add_adjoint_source,
// M:310-322), rebuilds the forward wavefield in reverse time from the last
// state with the saved CPML boundaries (kernel_set_border, M:422-462), zeroes
// the x/y/bottom CPML zones (kernel_setZero_CPMLzone, M:466-499), and forms
// continuous-adjoint imaging conditions (kernel_Gradients_Rho/_Lambda/_Mu,
// M:191-294; kernel_Gradients_Mij, M:337-372).  It keeps an extra copy of v
// (kernel_DTD_copy, M:501-511); here the reverse velocity sweep writes
// out of place and the two v buffers swap roles, so no copy is made.
#include <string>
#include <utility>

#include "common.hpp"

namespace efm {

int launch_stress_src(const Geo &g, const EfmMedium &m, const ShotPtrs &sh, const StatePtrs &in,
                      const StatePtrs &out, int it, int src_mode, hipStream_t st);

__global__ void k_inject_ref(Geo g, ShotPtrs sh, StatePtrs a, const float *res, int it) {
    const int s = blockIdx.x;
    const int nts = sh.nt[s];
    const int n = nts - it - 1;
    if (n < 0 || n >= nts) return;
    for (int r = threadIdx.x; r < g.nrec; r += blockDim.x) {
        const int cell = sh.rec[(long long)s * g.nrec + r];
        if (cell < 0) continue;
        const long long t = (long long)s * g.N + cell;
        const long long base = ((long long)s * 3 * g.ntmax + n) * g.nrec + r;
        const long long cs = (long long)g.ntmax * g.nrec;
        atomicAdd(&a.v[0][t], res[base]);
        atomicAdd(&a.v[1][t], res[base + cs]);
        atomicAdd(&a.v[2][t], res[base + 2 * cs]);
    }
}

// Single thread per shot, like the reference (which uses thread 0 of one block);
// the reference evaluates this with double literals, so it is done in double.
__global__ void k_image_mt_ref(Geo g, const float *lam, const float *mu, ShotPtrs sh, StatePtrs a,
                               double *gm6, int it) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.S) return;
    const int nts = sh.nt[s];
    if (it >= nts) return;
    const int *sp = sh.src + 3 * s;
    const long long nx = g.nx, sxz = g.sxz;
    const long long lc = sp[0] + sp[2] * nx + sp[1] * sxz;
    const long long t = (long long)s * g.N + lc;
    const double Mu = mu[lc], L = lam[lc];
    const double Mi = mu[lc + 1], Mk = mu[lc + nx], Mj = mu[lc + sxz];
    const double Mik = mu[lc + 1 + nx], Mjk = mu[lc + nx + sxz], Mij = mu[lc + 1 + sxz];
    const double mxz = 1.0 / (0.25 / Mu + 0.25 / Mi + 0.25 / Mk + 0.25 / Mik);
    const double mxy = 1.0 / (0.25 / Mu + 0.25 / Mi + 0.25 / Mj + 0.25 / Mij);
    const double myz = 1.0 / (0.25 / Mu + 0.25 / Mj + 0.25 / Mk + 0.25 / Mjk);
    const double w = sh.wav[(long long)s * g.ntmax + (nts - it - 1)];
    const double txx = a.t[0][t], tyy = a.t[1][t], tzz = a.t[2][t];
    const double txy = a.t[3][t], txz = a.t[4][t], tyz = a.t[5][t];
    const double k = 1.0 / (2.0 * Mu * (3.0 * L + 2.0 * Mu));
    double *o = gm6 + 6 * s;
    // each term is stored to float in the reference before the host sum
    o[0] += (float)(w * k * (2.0 * (L + Mu) * txx - L * tyy - L * tzz));
    o[1] += (float)(w * k * (2.0 * (L + Mu) * tyy - L * txx - L * tzz));
    o[2] += (float)(w * k * (2.0 * (L + Mu) * tzz - L * txx - L * tyy));
    o[3] += (float)(w * (1.0 / mxy) * txy);
    o[4] += (float)(w * (1.0 / mxz) * txz);
    o[5] += (float)(w * (1.0 / myz) * tyz);
}

// λ/μ/ρ imaging conditions, accumulated in float32 (the reference adds the
// per-step device volumes on the host in float32).  Divisions by the medium
// use the hardware reciprocal (1 ulp) and those by dt, dx, dy, dz the
// reciprocals of the steps: the images move by a few ulp, far inside the
// 1e-4 the gradients are held to.  Index arithmetic is 32-bit (N < 2^31).
__global__ void k_image_ref(Geo g, const float *lam, const float *mu, StatePtrs a, StatePtrs f,
                            StatePtrs fp, float dt, float dx, float dy, float dz, float *glam,
                            float *gmu, float *grho) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (i >= g.N) return;
    const int row = i / g.nx;
    const int x = i - row * g.nx;
    const int y = row / g.nz;
    const int z = row - y * g.nz;
    if (!interior(x, y, z, g)) return;
    const int nx = g.nx, sxz = g.sxz32;
    const long long so = (long long)s * g.N;
    const float *__restrict__ av0 = a.v[0] + so, *__restrict__ av1 = a.v[1] + so, *__restrict__ av2 = a.v[2] + so;
    const float *__restrict__ vx = fp.v[0] + so, *__restrict__ vy = fp.v[1] + so, *__restrict__ vz = fp.v[2] + so;
    const float *__restrict__ fx = f.v[0] + so, *__restrict__ fy = f.v[1] + so, *__restrict__ fz = f.v[2] + so;
    const float idt = 1.f / dt, idx = 1.f / dx, idy = 1.f / dy, idz = 1.f / dz;
    // M:206
    const float r = -((av0[i] * (vx[i] - fx[i]) * idt) + (av1[i] * (vy[i] - fy[i]) * idt) +
                      (av2[i] * (vz[i] - fz[i]) * idt));
    const float L = lam[i], M = mu[i];
    // M:234-238
    const float dvx = 0.5f * (vx[i] - vx[i - 1]) * idx + 0.5f * (vx[i - 1] - vx[i - 2]) * idx;
    const float dvy = 0.5f * (vy[i] - vy[i - sxz]) * idy + 0.5f * (vy[i - sxz] - vy[i - 2 * sxz]) * idy;
    const float dvz = 0.5f * (vz[i] - vz[i - nx]) * idz + 0.5f * (vz[i - nx] - vz[i - 2 * nx]) * idz;
    const float txx = a.t[0][so + i], tyy = a.t[1][so + i], tzz = a.t[2][so + i];
    const float k3 = frcp(3.f * L + 2.f * M);
    const float gl = (txx + tyy + tzz) * (dvx + dvy + dvz) * k3;
    // M:262-291: 1/μ̄ for the three shear planes (the reference inverts the
    // harmonic mean and then inverts it back; here the sum of quarter
    // reciprocals is used directly)
    const float q0 = 0.25f * frcp(M), qi = 0.25f * frcp(mu[i + 1]), qk = 0.25f * frcp(mu[i + nx]);
    const float qj = 0.25f * frcp(mu[i + sxz]);
    const float imxz = q0 + qi + qk + 0.25f * frcp(mu[i + 1 + nx]);
    const float imxy = q0 + qi + qj + 0.25f * frcp(mu[i + 1 + sxz]);
    const float imyz = q0 + qj + qk + 0.25f * frcp(mu[i + nx + sxz]);
    const float ex = vx[i] - vx[i - 1], ey = vy[i] - vy[i - sxz], ez = vz[i] - vz[i - nx];
    const float dvx_dz = vx[i + nx] - vx[i], dvx_dy = vx[i + sxz] - vx[i];
    const float dvy_dx = vy[i + 1] - vy[i], dvy_dz = vy[i + nx] - vy[i];
    const float dvz_dx = vz[i + 1] - vz[i], dvz_dy = vz[i + sxz] - vz[i];
    const float k = frcp(M) * k3;
    const float gm = k * (2.f * (L + M) * txx - L * tyy - L * tzz) * ex
                   + k * (2.f * (L + M) * tyy - L * txx - L * tzz) * ey
                   + k * (2.f * (L + M) * tzz - L * txx - L * tyy) * ez
                   + imxy * a.t[3][so + i] * (dvy_dx + dvx_dy)
                   + imxz * a.t[4][so + i] * (dvx_dz + dvz_dx)
                   + imyz * a.t[5][so + i] * (dvz_dy + dvy_dz);
    glam[so + i] += gl;
    gmu[so + i] += gm;
    grho[so + i] += r;
}

__global__ void k_zero_cpml(Geo g, StatePtrs f, int W) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (i >= g.N) return;
    const int row = i / g.nx;
    const int x = i - row * g.nx;
    const int y = row / g.nz;
    const int z = row - y * g.nz;
    if (!(x < W || x >= g.nx - W || z >= g.nz - W || y < W || y >= g.ny - W)) return;
    const long long t = (long long)s * g.N + i;
    for (int k = 0; k < 3; ++k) f.v[k][t] = 0.f;
    for (int k = 0; k < 6; ++k) f.t[k][t] = 0.f;
}

}  // namespace efm

using namespace efm;

static inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

static int check_common(const EfmDims *d, const EfmMedium *m, const EfmShots *sh) {
    return validate(d, m, sh);
}

extern "C" int efm_inject_residual(const EfmDims *d, const EfmShots *sh, const EfmState *st,
                                   const float *res, int it, void *stream) {
    if (!d || !sh || !st || (!res && d->nrec > 0))
        return fail(EFM_E_ARG, "efm_inject_residual: null argument");
    if (int rc = check_dims(d)) return rc;
    if (int rc = validate_shots(d, sh)) return rc;
    if (int rc = validate_state(d, st)) return rc;
    const Geo g = make_geo(*d);
    if (g.nrec <= 0) return EFM_OK;
    hipLaunchKernelGGL(k_inject_ref, dim3(g.S), dim3(64), 0, (hipStream_t)stream, g,
                       make_shots(*sh), make_state(*st), res, it);
    return check_launch("inject residual");
}

extern "C" int efm_image_mt_ref(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                                const EfmState *adj, double *gm6, int it, void *stream) {
    int rc = check_common(d, m, sh);
    if (rc) return rc;
    if (!adj || !gm6) return fail(EFM_E_ARG, "efm_image_mt_ref: null argument");
    if ((rc = validate_state(d, adj))) return rc;
    const Geo g = make_geo(*d);
    hipLaunchKernelGGL(k_image_mt_ref, dim3(nblk(g.S, 64)), dim3(64), 0, (hipStream_t)stream, g,
                       m->lam, m->mu, make_shots(*sh), make_state(*adj), gm6, it);
    return check_launch("image mt");
}

extern "C" int efm_backward_mt_ref(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                                   const EfmState *adj, const float *res, double *gm6, int nsteps,
                                   void *stream) {
    int rc = check_common(d, m, sh);
    if (rc) return rc;
    if (!adj || (!res && d->nrec > 0) || !gm6)
        return fail(EFM_E_ARG, "efm_backward_mt_ref: null argument");
    if ((rc = validate_state(d, adj))) return rc;
    if (nsteps < 0 || nsteps > d->ntmax) return fail(EFM_E_ARG, "efm_backward_mt_ref: nsteps out of range");
    Geo g = make_geo(*d);
    if (g.c < 0) g.c = -g.c;  // the adjoint runs the forward operator with +dt (M:2066-2068)
    hipStream_t st = (hipStream_t)stream;
    const ShotPtrs shp = make_shots(*sh);
    const StatePtrs a = make_state(*adj);
    for (int it = 0; it < nsteps; ++it) {
        if ((rc = launch_stress_src(g, *m, shp, a, a, it, 0, st))) return rc;
        if ((rc = launch_velocity(g, *m, shp, a, a, it, st))) return rc;
        if (g.nrec > 0) {
            hipLaunchKernelGGL(k_inject_ref, dim3(g.S), dim3(64), 0, st, g, shp, a, res, it);
            if ((rc = check_launch("inject residual"))) return rc;
        }
        hipLaunchKernelGGL(k_image_mt_ref, dim3(nblk(g.S, 64)), dim3(64), 0, st, g, m->lam, m->mu,
                           shp, a, gm6, it);
        if ((rc = check_launch("image mt"))) return rc;
    }
    return EFM_OK;
}

extern "C" long long efm_border_floats(const EfmDims *d, int W) {
    if (!d || check_dims(d)) return -1;
    return border_floats(make_geo(*d), W);
}

// Iterations it0 .. it0+nsteps-1 of the reverse loop.  The present velocity
// alternates between the forward state's v and vpres (swapped every
// iteration), so iteration `it` finds it in fwd.v when it is even and in vpres
// when it is odd: a run split into ranges gives the same launches in the same
// order as one run over all nt-1 iterations.
static int backward_remake_ref(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                               const EfmState *fwd, float *vpres, const float *borders, int W,
                               const EfmState *adj, const float *res, float dt, float dx, float dy,
                               float dz, float *glam, float *gmu, float *grho, int it0, int nsteps,
                               void *stream, const char *who) {
    int rc = check_common(d, m, sh);
    if (rc) return rc;
    if (!fwd || !vpres || !borders || !adj || (!res && d->nrec > 0) || !glam || !gmu || !grho)
        return fail(EFM_E_ARG, std::string(who) + ": null argument");
    if ((rc = validate_state(d, fwd)) || (rc = validate_state(d, adj))) return rc;
    if (W + H + 1 > d->nx / 2 || W + H + 1 > d->ny / 2 || W + H + 1 > d->nz / 2)
        return fail(EFM_E_ARG, std::string(who) + ": CPML too wide for the grid");
    const int nt = d->ntmax;
    if (it0 < 0 || nsteps < 0 || it0 + nsteps > nt - 1)
        return fail(EFM_E_ARG, std::string(who) + ": iterations outside [0, nt-1)");
    Geo ga = make_geo(*d);
    if (ga.c < 0) ga.c = -ga.c;
    Geo gf = ga;
    gf.c = -ga.c;  // flagMode = 1 reverse stepping (M:1951, 1964)
    hipStream_t st = (hipStream_t)stream;
    const ShotPtrs shp = make_shots(*sh);
    const StatePtrs a = make_state(*adj);
    StatePtrs cur = make_state(*fwd);           // v buffers: cur = present
    StatePtrs nxt = cur;                         // nxt.v = the other buffer
    const long long vol = (long long)ga.S * ga.N;
    nxt.v[0] = vpres; nxt.v[1] = vpres + vol; nxt.v[2] = vpres + 2 * vol;
    if (it0 & 1) {
        std::swap(cur.v[0], nxt.v[0]); std::swap(cur.v[1], nxt.v[1]); std::swap(cur.v[2], nxt.v[2]);
    }
    for (int it = it0; it < it0 + nsteps; ++it) {
        // adjoint step with the forward operator
        if ((rc = launch_stress_src(ga, *m, shp, a, a, it, 0, st))) return rc;
        if ((rc = launch_velocity(ga, *m, shp, a, a, it, st))) return rc;
        if (ga.nrec > 0) {
            hipLaunchKernelGGL(k_inject_ref, dim3(ga.S), dim3(64), 0, st, ga, shp, a, res, it);
            if ((rc = check_launch("inject residual"))) return rc;
        }
        // forward remake: v (out of place into nxt), borders, CPML zeroing, τ, source
        if ((rc = launch_velocity(gf, *m, shp, cur, nxt, it, st))) return rc;
        if ((rc = launch_border(gf, shp, nxt, const_cast<float *>(borders), W, nt - it - 2, 0, -1, st)))
            return rc;
        hipLaunchKernelGGL(k_zero_cpml, dim3(nblk(ga.N, 256), ga.S), dim3(256), 0, st, ga, nxt, W);
        if ((rc = check_launch("zero cpml"))) return rc;
        if ((rc = launch_stress_src(gf, *m, shp, nxt, nxt, it, -1, st))) return rc;
        // imaging with v (new) and v_present (old)
        hipLaunchKernelGGL(k_image_ref, dim3(nblk(ga.N, 256), ga.S), dim3(256), 0, st, ga, m->lam,
                           m->mu, a, nxt, cur, dt, dx, dy, dz, glam, gmu, grho);
        if ((rc = check_launch("image"))) return rc;
        std::swap(cur.v[0], nxt.v[0]); std::swap(cur.v[1], nxt.v[1]); std::swap(cur.v[2], nxt.v[2]);
    }
    return EFM_OK;
}

extern "C" int efm_backward_remake_ref(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                                       const EfmState *fwd, float *vpres, const float *borders,
                                       int W, const EfmState *adj, const float *res, float dt,
                                       float dx, float dy, float dz, float *glam, float *gmu,
                                       float *grho, void *stream) {
    if (!d) return fail(EFM_E_ARG, "efm_backward_remake_ref: null dims");
    return backward_remake_ref(d, m, sh, fwd, vpres, borders, W, adj, res, dt, dx, dy, dz, glam, gmu,
                               grho, 0, d->ntmax - 1, stream, "efm_backward_remake_ref");
}

extern "C" int efm_backward_remake_ref_steps(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                                             const EfmState *fwd, float *vpres, const float *borders,
                                             int W, const EfmState *adj, const float *res, float dt,
                                             float dx, float dy, float dz, float *glam, float *gmu,
                                             float *grho, int it0, int nsteps, void *stream) {
    return backward_remake_ref(d, m, sh, fwd, vpres, borders, W, adj, res, dt, dx, dy, dz, glam, gmu,
                               grho, it0, nsteps, stream, "efm_backward_remake_ref_steps");
}
