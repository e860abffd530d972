// Forward sweeps of the 3-D elastic velocity-stress staggered-grid FDTD
// (O(2,2), CPML), written for gfx950.
//
// Reference: kernel_Tau (mainPropElastic3D_MT.py M:59-128), addSource
// (M:1728-1734), kernel_Vx_Vy_Vz (M:133-188), kernel_get_Station (M:324-335),
// kernel_get_border / kernel_set_border (M:378-462).
//
// Mapping: a block is BX (x, one 64-lane wave per row) × BZ (z) threads; each
// thread owns one (x, z) column and marches NYC cells along y.  x is the
// contiguous axis, so every wave issues fully coalesced 256-byte row loads; the
// ±1 neighbours along x and z are the same rows shifted and are served from the
// L1/LDS-backed vector cache of the CU, the ±y neighbour from the previous
// march step.  ψ memory variables live only in CPML slabs (identically zero
// elsewhere, where a = b = 0).  The moment-tensor source is added by the thread
// that owns the source cell right after it writes τ (the reference adds it with
// six separate gpuarray launches).
#include "common.hpp"

namespace efm {

struct Med {
    const float *lam, *mu, *rho;
};

// ---------------------------------------------------------------------------
// stress sweep
// ---------------------------------------------------------------------------
// Addressing: per-shot base pointers are wave-uniform (scalar registers) and
// every per-lane offset is a 32-bit int, so loads use the base+offset form.
template <int BX, int BZ, int NYC, int MAP>
__global__ EFM_KERNEL_ATTRS(BX, BZ, MAP) void k_stress(Geo g, Med m, ShotPtrs sh, StatePtrs in,
                                                       StatePtrs out, int it, int src_mode) {
    if (g.itp) it += *g.itp;
    int x, z, s, yb;
    if (!tile_coords<BX, BZ, NYC, MAP>(g, x, z, s, yb)) return;
    if (x < H || x >= g.nx - H || z < H || z >= g.nz - H) return;
    if (it >= sh.nt[s]) return;
    const int y0 = H + yb * NYC;
    const int y1 = min(y0 + NYC, g.ny - H);
    const int nx = g.nx, sxz = g.sxz32;
    const long long so = (long long)s * g.N;
    const float c = g.c;

    const int sxs = slab_coord(x, g.nx, g.wxl, g.wxh);
    const int szs = slab_coord(z, g.nz, g.wzl, g.wzh);
    float ax = 0.f, bx = 0.f, az = 0.f, bz = 0.f;
    if (sxs >= 0) { ax = sh.ax[s * g.nx + x]; bx = sh.bx[s * g.nx + x]; }
    if (szs >= 0) { az = sh.az[s * g.nz + z]; bz = sh.bz[s * g.nz + z]; }
    // src_mode: 1 adds M·w[it] (forward), -1 subtracts M·w[nt-1-it] (reverse-time
    // reconstruction, flagMode = 1 in M:1964), 0 injects nothing (adjoint runs)
    const int *sp = sh.src + 3 * s;
    const bool src_col = (src_mode != 0) && (x == sp[0]) && (z == sp[2]);
    const int srcy = sp[1];

    const float *__restrict__ vx = in.v[0] + so;
    const float *__restrict__ vy = in.v[1] + so;
    const float *__restrict__ vz = in.v[2] + so;
    const float *__restrict__ mu = m.mu;
    const float *__restrict__ lam = m.lam;
    const float *pxi = in.px + px_shot(g, s), *pyi = in.py + py_shot(g, s), *pzi = in.pz + pz_shot(g, s);
    float *pxo = out.px + px_shot(g, s), *pyo = out.py + py_shot(g, s), *pzo = out.pz + pz_shot(g, s);
    // τ in/out may alias (in-place update): no restrict on them
    const float *ti0 = in.t[0] + so, *ti1 = in.t[1] + so, *ti2 = in.t[2] + so;
    const float *ti3 = in.t[3] + so, *ti4 = in.t[4] + so, *ti5 = in.t[5] + so;
    float *to0 = out.t[0] + so, *to1 = out.t[1] + so, *to2 = out.t[2] + so;
    float *to3 = out.t[3] + so, *to4 = out.t[4] + so, *to5 = out.t[5] + so;

    for (int y = y0; y < y1; ++y) {
        const int t = x + z * nx + y * sxz;
        const float vxc = vx[t], vyc = vy[t], vzc = vz[t];
        const float dvx_dx = vxc - vx[t - 1];
        const float dvy_dy = vyc - vy[t - sxz];
        const float dvz_dz = vzc - vz[t - nx];
        const float dvx_dz = vx[t + nx] - vxc;
        const float dvx_dy = vx[t + sxz] - vxc;
        const float dvy_dx = vy[t + 1] - vyc;
        const float dvy_dz = vy[t + nx] - vyc;
        const float dvz_dx = vz[t + 1] - vzc;
        const float dvz_dy = vz[t + sxz] - vzc;
        // strain-like terms e = Δ + ψ(new)   (ψ = b·ψ + a·Δ, updated then used)
        float exx = dvx_dx, eyx = dvy_dx, ezx = dvz_dx;
        float exy = dvx_dy, eyy = dvy_dy, ezy = dvz_dy;
        float exz = dvx_dz, eyz = dvy_dz, ezz = dvz_dz;
        if (sxs >= 0) {
            const int pi = px_off(g, y, z, sxs), cs = g.pxc;
            const float p0 = bx * pxi[pi] + ax * dvx_dx;
            const float p1 = bx * pxi[pi + cs] + ax * dvy_dx;
            const float p2 = bx * pxi[pi + 2 * cs] + ax * dvz_dx;
            pxo[pi] = p0; pxo[pi + cs] = p1; pxo[pi + 2 * cs] = p2;
            exx += p0; eyx += p1; ezx += p2;
        }
        const int sys = slab_coord(y, g.ny, g.wyl, g.wyh);
        if (sys >= 0) {
            const float ay = sh.ay[s * g.ny + y], by = sh.by[s * g.ny + y];
            const int pi = py_off(g, sys, z, x), cs = g.pyc;
            const float p0 = by * pyi[pi] + ay * dvx_dy;
            const float p1 = by * pyi[pi + cs] + ay * dvy_dy;
            const float p2 = by * pyi[pi + 2 * cs] + ay * dvz_dy;
            pyo[pi] = p0; pyo[pi + cs] = p1; pyo[pi + 2 * cs] = p2;
            exy += p0; eyy += p1; ezy += p2;
        }
        if (szs >= 0) {
            const int pi = pz_off(g, y, szs, x), cs = g.pzc;
            const float p0 = bz * pzi[pi] + az * dvx_dz;
            const float p1 = bz * pzi[pi + cs] + az * dvy_dz;
            const float p2 = bz * pzi[pi + 2 * cs] + az * dvz_dz;
            pzo[pi] = p0; pzo[pi + cs] = p1; pzo[pi + 2 * cs] = p2;
            exz += p0; eyz += p1; ezz += p2;
        }
        // moduli: λ, μ and the three 4-point harmonic means of μ (M:80-89)
        const float L = lam[t];
        const float Mu = mu[t];
        const float r0 = frcp(Mu);
        const float ri = frcp(mu[t + 1]);
        const float rk = frcp(mu[t + nx]);
        const float rj = frcp(mu[t + sxz]);
        const float rik = frcp(mu[t + 1 + nx]);
        const float rjk = frcp(mu[t + nx + sxz]);
        const float rij = frcp(mu[t + 1 + sxz]);
        const float mxz = 4.f * frcp(r0 + ri + rk + rik);
        const float mxy = 4.f * frcp(r0 + ri + rj + rij);
        const float myz = 4.f * frcp(r0 + rj + rk + rjk);
        const float l2m = L + 2.f * Mu;
        float txx = ti0[t] + c * (l2m * exx + L * (eyy + ezz));
        float tyy = ti1[t] + c * (l2m * eyy + L * (exx + ezz));
        float tzz = ti2[t] + c * (l2m * ezz + L * (exx + eyy));
        float txy = ti3[t] + c * mxy * (exy + eyx);
        float txz = ti4[t] + c * mxz * (exz + ezx);
        float tyz = ti5[t] + c * myz * (eyz + ezy);
        if (src_col && y == srcy) {
            const int iw = src_mode > 0 ? it : sh.nt[s] - 1 - it;
            const float w0 = sh.wav[(long long)s * g.ntmax + iw];
            const float w = src_mode > 0 ? w0 : -w0;
            const float *m6 = sh.m6 + 6 * s;
            txx += m6[0] * w; tyy += m6[1] * w; tzz += m6[2] * w;
            txy += m6[3] * w; txz += m6[4] * w; tyz += m6[5] * w;
        }
        to0[t] = txx; to1[t] = tyy; to2[t] = tzz;
        to3[t] = txy; to4[t] = txz; to5[t] = tyz;
    }
}

// ---------------------------------------------------------------------------
// velocity sweep
// ---------------------------------------------------------------------------
template <int BX, int BZ, int NYC, int MAP>
__global__ EFM_KERNEL_ATTRS(BX, BZ, MAP) void k_velocity(Geo g, Med m, ShotPtrs sh, StatePtrs in,
                                                         StatePtrs out, int it, float *gathers) {
    if (g.itp) it += *g.itp;
    int x, z, s, yb;
    if (!tile_coords<BX, BZ, NYC, MAP>(g, x, z, s, yb)) return;
    if (x < H || x >= g.nx - H || z < H || z >= g.nz - H) return;
    if (it >= sh.nt[s]) return;
    const int y0 = H + yb * NYC;
    const int y1 = min(y0 + NYC, g.ny - H);
    const int nx = g.nx, sxz = g.sxz32;
    const long long so = (long long)s * g.N;
    const float c = g.c;

    const int sxs = slab_coord(x, g.nx, g.wxl, g.wxh);
    const int szs = slab_coord(z, g.nz, g.wzl, g.wzh);
    float ax = 0.f, bx = 0.f, az = 0.f, bz = 0.f;
    if (sxs >= 0) { ax = sh.ax[s * g.nx + x]; bx = sh.bx[s * g.nx + x]; }
    if (szs >= 0) { az = sh.az[s * g.nz + z]; bz = sh.bz[s * g.nz + z]; }

    // τ is read from `out` (the stress sweep of this step wrote it there)
    const float *__restrict__ txx = out.t[0] + so;
    const float *__restrict__ tyy = out.t[1] + so;
    const float *__restrict__ tzz = out.t[2] + so;
    const float *__restrict__ txy = out.t[3] + so;
    const float *__restrict__ txz = out.t[4] + so;
    const float *__restrict__ tyz = out.t[5] + so;
    const float *__restrict__ rho = m.rho;
    const float *pxi = in.px + px_shot(g, s), *pyi = in.py + py_shot(g, s), *pzi = in.pz + pz_shot(g, s);
    float *pxo = out.px + px_shot(g, s), *pyo = out.py + py_shot(g, s), *pzo = out.pz + pz_shot(g, s);
    // v in/out may alias (in-place update)
    const float *vi0 = in.v[0] + so, *vi1 = in.v[1] + so, *vi2 = in.v[2] + so;
    float *vo0 = out.v[0] + so, *vo1 = out.v[1] + so, *vo2 = out.v[2] + so;

    for (int y = y0; y < y1; ++y) {
        const int t = x + z * nx + y * sxz;
        const float dtxx_dx = txx[t + 1] - txx[t];
        const float dtzz_dz = tzz[t + nx] - tzz[t];
        const float dtyy_dy = tyy[t + sxz] - tyy[t];
        const float txyc = txy[t], txzc = txz[t], tyzc = tyz[t];
        const float dtxy_dy = txyc - txy[t - sxz];
        const float dtxy_dx = txyc - txy[t - 1];
        const float dtxz_dz = txzc - txz[t - nx];
        const float dtxz_dx = txzc - txz[t - 1];
        const float dtyz_dz = tyzc - tyz[t - nx];
        const float dtyz_dy = tyzc - tyz[t - sxz];
        float fxx = dtxx_dx, fyx = dtxy_dx, fzx = dtxz_dx;
        float fxy = dtxy_dy, fyy = dtyy_dy, fzy = dtyz_dy;
        float fxz = dtxz_dz, fyz = dtyz_dz, fzz = dtzz_dz;
        if (sxs >= 0) {
            const int cs = g.pxc, pi = px_off(g, y, z, sxs) + 3 * cs;
            const float p0 = bx * pxi[pi] + ax * dtxx_dx;
            const float p1 = bx * pxi[pi + cs] + ax * dtxy_dx;
            const float p2 = bx * pxi[pi + 2 * cs] + ax * dtxz_dx;
            pxo[pi] = p0; pxo[pi + cs] = p1; pxo[pi + 2 * cs] = p2;
            fxx += p0; fyx += p1; fzx += p2;
        }
        const int sys = slab_coord(y, g.ny, g.wyl, g.wyh);
        if (sys >= 0) {
            const float ay = sh.ay[s * g.ny + y], by = sh.by[s * g.ny + y];
            const int cs = g.pyc, pi = py_off(g, sys, z, x) + 3 * cs;
            const float p0 = by * pyi[pi] + ay * dtxy_dy;
            const float p1 = by * pyi[pi + cs] + ay * dtyy_dy;
            const float p2 = by * pyi[pi + 2 * cs] + ay * dtyz_dy;
            pyo[pi] = p0; pyo[pi + cs] = p1; pyo[pi + 2 * cs] = p2;
            fxy += p0; fyy += p1; fzy += p2;
        }
        if (szs >= 0) {
            const int cs = g.pzc, pi = pz_off(g, y, szs, x) + 3 * cs;
            const float p0 = bz * pzi[pi] + az * dtxz_dz;
            const float p1 = bz * pzi[pi + cs] + az * dtyz_dz;
            const float p2 = bz * pzi[pi + 2 * cs] + az * dtzz_dz;
            pzo[pi] = p0; pzo[pi + cs] = p1; pzo[pi + 2 * cs] = p2;
            fxz += p0; fyz += p1; fzz += p2;
        }
        // averaged buoyancy b̄ = 2bb'/(b+b') = 2/(ρ+ρ')  (M:153-159)
        const float r = rho[t];
        const float bbx = 2.f * frcp(r + rho[t + 1]);
        const float bby = 2.f * frcp(r + rho[t + sxz]);
        const float bbz = 2.f * frcp(r + rho[t + nx]);
        const float n0 = vi0[t] + c * bbx * (fxx + fxy + fxz);
        const float n1 = vi1[t] + c * bby * (fyx + fyy + fyz);
        const float n2 = vi2[t] + c * bbz * (fzx + fzy + fzz);
        vo0[t] = n0;
        vo1[t] = n1;
        vo2[t] = n2;
        // receiver sampling by the cell's own thread (kernel_get_Station, M:324-335)
        if (gathers && is_receiver(g, sh, s, t)) {
            const int *rs = sh.rec + (long long)s * g.nrec;
            const long long cs = (long long)g.ntmax * g.nrec;
            const long long base = ((long long)s * 3 * g.ntmax + it) * g.nrec;
            for (int r = 0; r < g.nrec; ++r)
                if (rs[r] == t) {
                    gathers[base + r] = n0;
                    gathers[base + cs + r] = n1;
                    gathers[base + 2 * cs + r] = n2;
                }
        }
    }
}

// ---------------------------------------------------------------------------
// receivers
// ---------------------------------------------------------------------------
__global__ void k_record(Geo g, ShotPtrs sh, StatePtrs st, float *gathers, int it) {
    if (g.itp) it += *g.itp;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (r >= g.nrec || it >= sh.nt[s]) return;
    const int cell = sh.rec[(long long)s * g.nrec + r];
    if (cell < 0) return;
    const long long t = (long long)s * g.N + cell;
    const long long base = ((long long)s * 3 * g.ntmax + it) * g.nrec + r;
    const long long cstride = (long long)g.ntmax * g.nrec;
    gathers[base] = st.v[0][t];
    gathers[base + cstride] = st.v[1][t];
    gathers[base + 2 * cstride] = st.v[2][t];
}

// ---------------------------------------------------------------------------
// reference boundary save / restore of v (2 layers inside the CPML on x, y,
// bottom; z in [2,4) at the top)
// ---------------------------------------------------------------------------
struct BorderLayout {
    long long left, right, top, bottom, front, back, comp;  // offsets, comp = floats per component
};
__host__ __device__ inline BorderLayout border_layout(int nx, int ny, int nz) {
    BorderLayout b;
    const long long lr = (long long)ny * nz * H, tb = (long long)ny * H * nx, fb = (long long)H * nz * nx;
    b.left = 0; b.right = lr; b.top = 2 * lr; b.bottom = 2 * lr + tb; b.front = 2 * lr + 2 * tb;
    b.back = 2 * lr + 2 * tb + fb; b.comp = 2 * lr + 2 * tb + 2 * fb;
    return b;
}

long long border_floats(const Geo &g, int W) {
    (void)W;
    return 3 * border_layout(g.nx, g.ny, g.nz).comp;
}

// One thread per (component, face element).  borders: (S, nslot, 3, comp)
__global__ void k_border(Geo g, ShotPtrs sh, StatePtrs st, float *borders, int W, int slot,
                         int nslot, int save, int it_gate) {
    const BorderLayout L = border_layout(g.nx, g.ny, g.nz);
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (i >= 3 * L.comp) return;
    if (it_gate >= 0 && it_gate >= sh.nt[s]) return;
    const int comp = (int)(i / L.comp);
    long long e = i - comp * L.comp;
    int x, y, z;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (e < L.top) {                     // left / right: (ny, nz, H)
        const int side = e >= L.right;
        long long k = e - (side ? L.right : 0);
        const int h = (int)(k % H); k /= H;
        z = (int)(k % nz); y = (int)(k / nz);
        x = side ? nx - W - H + h : W + h;
    } else if (e < L.front) {            // top / bottom: (ny, H, nx)
        const int side = e >= L.bottom;
        long long k = e - (side ? L.bottom : L.top);
        x = (int)(k % nx); k /= nx;
        const int h = (int)(k % H); y = (int)(k / H);
        z = side ? nz - W - H + h : H + h;
    } else {                             // front / back: (H, nz, nx)
        const int side = e >= L.back;
        long long k = e - (side ? L.back : L.front);
        x = (int)(k % nx); k /= nx;
        z = (int)(k % nz); const int h = (int)(k / nz);
        y = side ? ny - W - H + h : W + h;
    }
    const long long t = (long long)s * g.N + x + (long long)z * nx + (long long)y * g.sxz;
    float *buf = borders + (((long long)s * nslot + slot) * 3 + comp) * L.comp + e;
    if (save) *buf = st.v[comp][t];
    else st.v[comp][t] = *buf;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int launch_stress_src(const Geo &g, const EfmMedium &m, const ShotPtrs &sh, const StatePtrs &in,
                      const StatePtrs &out, int it, int src_mode, hipStream_t st) {
    const Tile t = get_tile(0);
    Med md{m.lam, m.mu, m.rho};
    bool launched = false;
#define X(a, b, c, o) EFM_LAUNCH_IF(a, b, c, o, k_stress, g, md, sh, in, out, it, src_mode)
    EFM_TILE_LIST(X)
#undef X
    if (!launched) return fail(EFM_E_UNSUPPORTED, "unsupported stress tile");
    return check_launch("stress sweep");
}

int launch_stress(const Geo &g, const EfmMedium &m, const ShotPtrs &sh, const StatePtrs &in,
                  const StatePtrs &out, int it, hipStream_t st) {
    return launch_stress_src(g, m, sh, in, out, it, 1, st);
}

int launch_velocity(const Geo &g, const EfmMedium &m, const ShotPtrs &sh, const StatePtrs &in,
                    const StatePtrs &out, int it, hipStream_t st, float *gathers) {
    const Tile t = get_tile(1);
    Med md{m.lam, m.mu, m.rho};
    bool launched = false;
    // the sweep samples the receivers only when the shots carry a receiver bitmask
    float *gat = (sh.recmask && g.nrec > 0) ? gathers : nullptr;
#define X(a, b, c, o) EFM_LAUNCH_IF(a, b, c, o, k_velocity, g, md, sh, in, out, it, gat)
    EFM_TILE_LIST(X)
#undef X
    if (!launched) return fail(EFM_E_UNSUPPORTED, "unsupported velocity tile");
    return check_launch("velocity sweep");
}

int launch_record(const Geo &g, const ShotPtrs &sh, const StatePtrs &s, float *gathers, int it,
                  hipStream_t st) {
    if (g.nrec <= 0) return EFM_OK;
    dim3 grid((g.nrec + 63) / 64, g.S);
    hipLaunchKernelGGL(k_record, grid, dim3(64), 0, st, g, sh, s, gathers, it);
    return check_launch("record");
}

int launch_border(const Geo &g, const ShotPtrs &sh, const StatePtrs &s, float *borders, int W,
                  int slot, int save, int it_gate, hipStream_t st) {
    const long long n = 3 * border_layout(g.nx, g.ny, g.nz).comp;
    dim3 grid((unsigned)((n + 255) / 256), g.S);
    hipLaunchKernelGGL(k_border, grid, dim3(256), 0, st, g, sh, s, borders, W, slot, g.ntmax,
                       save, it_gate);
    return check_launch("border");
}

}  // namespace efm
