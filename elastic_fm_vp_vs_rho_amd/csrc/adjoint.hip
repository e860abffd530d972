// Exact discrete adjoint of the forward recurrence in forward.hip, plus the
// model maps.  The reference has no exact adjoint (its backwardFields re-uses
// the forward operator, M:2056-2082); these kernels are the transpose of the
// two sweeps, derived by hand:
//
//   forward step n     τ' = τ + c·S(λ,μ)·e(v,ψv),  τ'' = τ' + M·w[n]δ_src,
//                      v' = v + c·B̄(ρ)·F(τ'',ψτ),  d[n] = R v'
//   ψ recurrences      ψ' = bψ + aΔ,  e = Δ + ψ'   (likewise for ψτ)
//
//   reverse step n     R1 v̄[rec] += r[n]
//                      R2 u = c·b̄·v̄;  h̄ = u + a(C+u);  C ← b(C+u);  τ̄ += D_τᵀ h̄
//                         ρ-accumulators  += v̄·(v^{n+1} - v^n)   (= v̄·c·b̄·F)
//                      R3 M̄ += w[n]·τ̄(src)
//                      R4 ē = c·Sᵀτ̄;  ḡ = ē + a(C+ē);  C ← b(C+ē);  v̄ += D_vᵀ ḡ
//                         λ/μ accumulators += τ̄·(∂S/∂λ,μ)e
//
// D_ᵀ are applied in gather form (each thread writes only its own cell), which
// needs h̄/ḡ at the ±1 neighbours; C is read at neighbours while its owner
// rewrites it, so C is ping-ponged between two buffers.  The derived medium
// (μ̄, b̄) is zero outside the interior and τ̄, v̄, C are only ever written on
// the interior, so neighbour terms vanish there without branches.
#include <mutex>
#include <string>
#include <utility>

#include "common.hpp"

namespace efm {

// Loads of the stored forward state carry the streaming (nontemporal) hint:
// each value is read by one step of the reverse sweep only, and the hint keeps
// the reused adjoint state, medium and accumulators cached.
#define LDF(p) __builtin_nontemporal_load(p)

// CPML coefficient of cell i (0 outside the slab, slab coordinate < 0).  The
// load is unconditional (every interior cell's ±1 neighbour indexes the
// profile), so it issues with the others instead of behind a branch.
static __device__ __forceinline__ float coef(const float *p, int i, int slab) {
    const float v = p[i];
    return slab >= 0 ? v : 0.f;
}

struct AdjArgs {
    const float *lam, *mu;
    const float *mbar, *bbar;      // (3, N) each
    StatePtrs fn, fn1;             // forward S^n, S^{n+1}
    StatePtrs a, a2;               // adjoint: a = live (v̄, τ̄, C in), a2 = C out
    float *acc;                    // per shot 8 padded components: λ, μ, μ_xy, μ_xz sums
                                   // interleaved (one float4 per cell) in components 0-3,
                                   // the ρ sum in 4, μ_yz in 7 (5 and 6 unused: 8 components
                                   // keep every shot's float4 base 16-byte aligned for any N)
    ShotPtrs sh;
    const float *resid;            // ∂J/∂gathers, used by R4 when the shots carry a receiver mask
    double *gm6;                   // ∂J/∂M, likewise
};

// ---------------------------------------------------------------------------
// R2: transpose of the velocity sweep
// ---------------------------------------------------------------------------
template <int BX, int BZ, int NYC, int MAP>
__global__ EFM_KERNEL_ATTRS(BX, BZ, MAP) void k_adj_velocity(Geo g, AdjArgs A, int n) {
    int x, z, s, yb;
    if (!tile_coords<BX, BZ, NYC, MAP>(g, x, z, s, yb)) return;
    if (x < H || x >= g.nx - H || z < H || z >= g.nz - H) return;
    if (n >= A.sh.nt[s]) return;
    const int y0 = H + yb * NYC;
    const int y1 = min(y0 + NYC, g.ny - H);
    const int nx = g.nx, sxz = g.sxz32;
    const long long so = (long long)s * g.N;
    const float c = g.c;
    const ShotPtrs &sh = A.sh;

    // slab coordinates and coefficients of this column and its x/z neighbours
    const int sx0 = slab_coord(x, g.nx, g.wxl, g.wxh);
    const int sxm = slab_coord(x - 1, g.nx, g.wxl, g.wxh);
    const int sxp = slab_coord(x + 1, g.nx, g.wxl, g.wxh);
    const int sz0 = slab_coord(z, g.nz, g.wzl, g.wzh);
    const int szm = slab_coord(z - 1, g.nz, g.wzl, g.wzh);
    const int szp = slab_coord(z + 1, g.nz, g.wzl, g.wzh);
    const float *axs = sh.ax + (long long)s * g.nx, *bxs = sh.bx + (long long)s * g.nx;
    const float *ays = sh.ay + (long long)s * g.ny, *bys = sh.by + (long long)s * g.ny;
    const float *azs = sh.az + (long long)s * g.nz, *bzs = sh.bz + (long long)s * g.nz;
    const float ax0 = coef(axs, x, sx0), bx0 = coef(bxs, x, sx0);
    const float axm = coef(axs, x - 1, sxm), axp = coef(axs, x + 1, sxp);
    const float az0 = coef(azs, z, sz0), bz0 = coef(bzs, z, sz0);
    const float azm = coef(azs, z - 1, szm), azp = coef(azs, z + 1, szp);
    const int csx = g.pxc, csy = g.pyc, csz = g.pzc;

    const float *__restrict__ bbx = A.bbar;
    const float *__restrict__ bby = A.bbar + g.ps;
    const float *__restrict__ bbz = A.bbar + 2 * g.ps;
    const float *__restrict__ vbx = A.a.v[0] + so;
    const float *__restrict__ vby = A.a.v[1] + so;
    const float *__restrict__ vbz = A.a.v[2] + so;
    float *__restrict__ tb0 = A.a.t[0] + so;
    float *__restrict__ tb1 = A.a.t[1] + so;
    float *__restrict__ tb2 = A.a.t[2] + so;
    float *__restrict__ tb3 = A.a.t[3] + so;
    float *__restrict__ tb4 = A.a.t[4] + so;
    float *__restrict__ tb5 = A.a.t[5] + so;
    // forward velocities after and before step n: their difference is c·b̄·F, the
    // velocity update the ρ products need (read at the cell and at p-x, p-y, p-z, so
    // without the streaming hint)
    const float *__restrict__ Wx1 = A.fn1.v[0] + so;
    const float *__restrict__ Wy1 = A.fn1.v[1] + so;
    const float *__restrict__ Wz1 = A.fn1.v[2] + so;
    const float *__restrict__ Wx0 = A.fn.v[0] + so;
    const float *__restrict__ Wy0 = A.fn.v[1] + so;
    const float *__restrict__ Wz0 = A.fn.v[2] + so;
    const float *__restrict__ Cx = A.a.px + px_shot(g, s);
    const float *__restrict__ Cy = A.a.py + py_shot(g, s);
    const float *__restrict__ Cz = A.a.pz + pz_shot(g, s);
    float *__restrict__ Cxo = A.a2.px + px_shot(g, s);
    float *__restrict__ Cyo = A.a2.py + py_shot(g, s);
    float *__restrict__ Czo = A.a2.pz + pz_shot(g, s);
    // ρ sums: component 4 of the shot
    float *__restrict__ accr = A.acc + ((long long)s * 8 + 4) * g.ps;

    for (int y = y0; y < y1; ++y) {
        const int t = x + z * nx + y * sxz;
        const int sy0 = slab_coord(y, g.ny, g.wyl, g.wyh);
        const int sym = slab_coord(y - 1, g.ny, g.wyl, g.wyh);
        const int syp = slab_coord(y + 1, g.ny, g.wyl, g.wyh);
        const float ay0 = coef(ays, y, sy0), by0 = coef(bys, y, sy0);
        const float aym = coef(ays, y - 1, sym), ayp = coef(ays, y + 1, syp);

        // Every load of the cell that does not depend on the CPML slabs is issued
        // here, before the first store: the compiler cannot prove that the stores
        // below miss these arrays, so a load placed after a store would wait for it.
        const float tbo0 = tb0[t], tbo1 = tb1[t], tbo2 = tb2[t];
        const float tbo3 = tb3[t], tbo4 = tb4[t], tbo5 = tb5[t];
        const float acco = accr[t];
        const float dwx = Wx1[t] - Wx0[t], dwy = Wy1[t] - Wy0[t], dwz = Wz1[t] - Wz0[t];
        const float dwx_m = Wx1[t - 1] - Wx0[t - 1];
        const float dwy_m = Wy1[t - sxz] - Wy0[t - sxz];
        const float dwz_m = Wz1[t - nx] - Wz0[t - nx];
        // u = c·b̄·v̄ at p and at the neighbours the transposes need
        const float vx0 = vbx[t], vy0 = vby[t], vz0 = vbz[t];
        const float ux = c * bbx[t] * vx0;
        const float uy = c * bby[t] * vy0;
        const float uz = c * bbz[t] * vz0;
        const float ux_mx = c * bbx[t - 1] * vbx[t - 1];        // txx_x at p-x
        const float uy_my = c * bby[t - sxz] * vby[t - sxz];    // tyy_y at p-y
        const float uz_mz = c * bbz[t - nx] * vbz[t - nx];      // tzz_z at p-z
        const float ux_py = c * bbx[t + sxz] * vbx[t + sxz];    // txy_y at p+y
        const float uy_px = c * bby[t + 1] * vby[t + 1];        // txy_x at p+x
        const float ux_pz = c * bbx[t + nx] * vbx[t + nx];      // txz_z at p+z
        const float uz_px = c * bbz[t + 1] * vbz[t + 1];        // txz_x at p+x
        const float uy_pz = c * bby[t + nx] * vby[t + nx];      // tyz_z at p+z
        const float uz_py = c * bbz[t + sxz] * vbz[t + sxz];    // tyz_y at p+y
        // ρ: this cell enters the averaged buoyancy of its own three faces and of the
        // faces at p-x, p-y, p-z; each face contributes u·Δv (= c·b̄·v̄ times this step's
        // forward velocity change).  u vanishes off the interior (b̄ = 0 there).
        const float px = ux * dwx, py = uy * dwy, pz = uz * dwz;
        const float accn = acco + ((px + ux_mx * dwx_m) + (py + uy_my * dwy_m) + (pz + uz_mz * dwz_m));

        // h̄ = u + a(C + u) with C read from the "in" buffer
        float hb_xx_m = ux_mx, hb_xx_0 = ux;           // txx_x
        float hb_yy_m = uy_my, hb_yy_0 = uy;           // tyy_y
        float hb_zz_m = uz_mz, hb_zz_0 = uz;           // tzz_z
        float hb_xyy_0 = ux, hb_xyy_p = ux_py;         // txy_y
        float hb_xyx_0 = uy, hb_xyx_p = uy_px;         // txy_x
        float hb_xzz_0 = ux, hb_xzz_p = ux_pz;         // txz_z
        float hb_xzx_0 = uz, hb_xzx_p = uz_px;         // txz_x
        float hb_yzz_0 = uy, hb_yzz_p = uy_pz;         // tyz_z
        float hb_yzy_0 = uz, hb_yzy_p = uz_py;         // tyz_y
        // x-slab components 3 txx_x, 4 txy_x, 5 txz_x
        if (sxm >= 0) hb_xx_m += axm * (Cx[px_off(g, y, z, sxm) + 3 * csx] + ux_mx);
        if (sxp >= 0) {
            const int pi = px_off(g, y, z, sxp) + 4 * csx;
            hb_xyx_p += axp * (Cx[pi] + uy_px);
            hb_xzx_p += axp * (Cx[pi + csx] + uz_px);
        }
        if (sx0 >= 0) {
            const int pi = px_off(g, y, z, sx0) + 3 * csx;
            const float c0 = Cx[pi] + ux, c1 = Cx[pi + csx] + uy, c2 = Cx[pi + 2 * csx] + uz;
            hb_xx_0 += ax0 * c0; hb_xyx_0 += ax0 * c1; hb_xzx_0 += ax0 * c2;
            Cxo[pi] = bx0 * c0; Cxo[pi + csx] = bx0 * c1; Cxo[pi + 2 * csx] = bx0 * c2;
        }
        // y-slab components 3 txy_y, 4 tyy_y, 5 tyz_y
        if (sym >= 0) hb_yy_m += aym * (Cy[py_off(g, sym, z, x) + 4 * csy] + uy_my);
        if (syp >= 0) {
            const int pi = py_off(g, syp, z, x) + 3 * csy;
            hb_xyy_p += ayp * (Cy[pi] + ux_py);
            hb_yzy_p += ayp * (Cy[pi + 2 * csy] + uz_py);
        }
        if (sy0 >= 0) {
            const int pi = py_off(g, sy0, z, x) + 3 * csy;
            const float c0 = Cy[pi] + ux, c1 = Cy[pi + csy] + uy, c2 = Cy[pi + 2 * csy] + uz;
            hb_xyy_0 += ay0 * c0; hb_yy_0 += ay0 * c1; hb_yzy_0 += ay0 * c2;
            Cyo[pi] = by0 * c0; Cyo[pi + csy] = by0 * c1; Cyo[pi + 2 * csy] = by0 * c2;
        }
        // z-slab components 3 txz_z, 4 tyz_z, 5 tzz_z
        if (szm >= 0) hb_zz_m += azm * (Cz[pz_off(g, y, szm, x) + 5 * csz] + uz_mz);
        if (szp >= 0) {
            const int pi = pz_off(g, y, szp, x) + 3 * csz;
            hb_xzz_p += azp * (Cz[pi] + ux_pz);
            hb_yzz_p += azp * (Cz[pi + csz] + uy_pz);
        }
        if (sz0 >= 0) {
            const int pi = pz_off(g, y, sz0, x) + 3 * csz;
            const float c0 = Cz[pi] + ux, c1 = Cz[pi + csz] + uy, c2 = Cz[pi + 2 * csz] + uz;
            hb_xzz_0 += az0 * c0; hb_yzz_0 += az0 * c1; hb_zz_0 += az0 * c2;
            Czo[pi] = bz0 * c0; Czo[pi + csz] = bz0 * c1; Czo[pi + 2 * csz] = bz0 * c2;
        }
        // τ̄ += D_τᵀ h̄
        tb0[t] = tbo0 + (hb_xx_m - hb_xx_0);
        tb1[t] = tbo1 + (hb_yy_m - hb_yy_0);
        tb2[t] = tbo2 + (hb_zz_m - hb_zz_0);
        tb3[t] = tbo3 + ((hb_xyy_0 - hb_xyy_p) + (hb_xyx_0 - hb_xyx_p));
        tb4[t] = tbo4 + ((hb_xzz_0 - hb_xzz_p) + (hb_xzx_0 - hb_xzx_p));
        tb5[t] = tbo5 + ((hb_yzz_0 - hb_yzz_p) + (hb_yzy_0 - hb_yzy_p));

        // A face on the last interior plane also feeds the cell beyond it, which no
        // thread of this sweep owns; that cell has no other face in the interior, so
        // the face's thread adds it there (one writer per cell).
        accr[t] = accn;
        if (x == g.nx - H - 1) accr[t + 1] += px;
        if (y == g.ny - H - 1) accr[t + sxz] += py;
        if (z == g.nz - H - 1) accr[t + nx] += pz;
    }
}

// ---------------------------------------------------------------------------
// R4: transpose of the stress sweep
// ---------------------------------------------------------------------------
// At least 7 waves per SIMD: the compiler then fits the sweep in 71 VGPRs (73 by
// itself, which allows only 6) by moving some SGPRs to VGPR lanes; the C2 adjoint
// step gets 2-3 % faster.
template <int BX, int BZ, int NYC, int MAP>
__global__ EFM_KERNEL_ATTRS(BX, BZ, MAP) __attribute__((amdgpu_waves_per_eu(7, 8))) void k_adj_stress(Geo g, AdjArgs A, int n) {
    int x, z, s, yb;
    if (!tile_coords<BX, BZ, NYC, MAP>(g, x, z, s, yb)) return;
    if (x < H || x >= g.nx - H || z < H || z >= g.nz - H) return;
    const int nts = A.sh.nt[s];
    const int y0 = H + yb * NYC;
    const int y1 = min(y0 + NYC, g.ny - H);
    const int nx = g.nx, sxz = g.sxz32;
    const long long so = (long long)s * g.N;
    const float c = g.c;
    const ShotPtrs &sh = A.sh;
    // with a receiver mask this sweep also injects the residual of step n-1 (R1
    // of the next reverse step) and sums ∂J/∂M of step n (R3)
    const bool fused = sh.recmask != nullptr;
    const int *rs = sh.rec + (long long)s * g.nrec;
    const long long rcs = (long long)g.ntmax * g.nrec;
    const long long rbase = ((long long)s * 3 * g.ntmax + (n - 1)) * g.nrec;
    if (n >= nts) {
        // the shot's record ends at step nts - 1: its first residual is injected here
        if (fused && n == nts && n >= 1 && g.nrec > 0)
            for (int y = y0; y < y1; ++y) {
                const int t = x + z * nx + y * sxz;
                if (!is_receiver(g, sh, s, t)) continue;
                for (int r = 0; r < g.nrec; ++r)
                    if (rs[r] == t) {
                        A.a.v[0][so + t] += A.resid[rbase + r];
                        A.a.v[1][so + t] += A.resid[rbase + rcs + r];
                        A.a.v[2][so + t] += A.resid[rbase + 2 * rcs + r];
                    }
            }
        return;
    }
    const int *sp = sh.src + 3 * s;
    const int tsrc = sp[0] + sp[2] * nx + sp[1] * sxz;
    const bool inject = fused && n >= 1 && g.nrec > 0;

    const int sx0 = slab_coord(x, g.nx, g.wxl, g.wxh);
    const int sxm = slab_coord(x - 1, g.nx, g.wxl, g.wxh);
    const int sxp = slab_coord(x + 1, g.nx, g.wxl, g.wxh);
    const int sz0 = slab_coord(z, g.nz, g.wzl, g.wzh);
    const int szm = slab_coord(z - 1, g.nz, g.wzl, g.wzh);
    const int szp = slab_coord(z + 1, g.nz, g.wzl, g.wzh);
    const float *axs = sh.ax + (long long)s * g.nx, *bxs = sh.bx + (long long)s * g.nx;
    const float *ays = sh.ay + (long long)s * g.ny, *bys = sh.by + (long long)s * g.ny;
    const float *azs = sh.az + (long long)s * g.nz, *bzs = sh.bz + (long long)s * g.nz;
    const float ax0 = sx0 >= 0 ? axs[x] : 0.f, bx0 = sx0 >= 0 ? bxs[x] : 0.f;
    const float axm = sxm >= 0 ? axs[x - 1] : 0.f, axp = sxp >= 0 ? axs[x + 1] : 0.f;
    const float az0 = sz0 >= 0 ? azs[z] : 0.f, bz0 = sz0 >= 0 ? bzs[z] : 0.f;
    const float azm = szm >= 0 ? azs[z - 1] : 0.f, azp = szp >= 0 ? azs[z + 1] : 0.f;
    const int csx = g.pxc, csy = g.pyc, csz = g.pzc;

    const float *__restrict__ mxy = A.mbar;
    const float *__restrict__ mxz = A.mbar + g.ps;
    const float *__restrict__ myz = A.mbar + 2 * g.ps;
    const float *__restrict__ lam = A.lam;
    const float *__restrict__ mu = A.mu;
    const float *__restrict__ Txx = A.a.t[0] + so;
    const float *__restrict__ Tyy = A.a.t[1] + so;
    const float *__restrict__ Tzz = A.a.t[2] + so;
    const float *__restrict__ Txy = A.a.t[3] + so;
    const float *__restrict__ Txz = A.a.t[4] + so;
    const float *__restrict__ Tyz = A.a.t[5] + so;
    float *__restrict__ vb0 = A.a.v[0] + so;
    float *__restrict__ vb1 = A.a.v[1] + so;
    float *__restrict__ vb2 = A.a.v[2] + so;
    const float *__restrict__ V0 = A.fn.v[0] + so;
    const float *__restrict__ V1 = A.fn.v[1] + so;
    const float *__restrict__ V2 = A.fn.v[2] + so;
    const float *__restrict__ Cx = A.a.px + px_shot(g, s);
    const float *__restrict__ Cy = A.a.py + py_shot(g, s);
    const float *__restrict__ Cz = A.a.pz + pz_shot(g, s);
    float *__restrict__ Cxo = A.a2.px + px_shot(g, s);
    float *__restrict__ Cyo = A.a2.py + py_shot(g, s);
    float *__restrict__ Czo = A.a2.pz + pz_shot(g, s);
    const float *__restrict__ Fx = A.fn.px + px_shot(g, s);
    const float *__restrict__ Fy = A.fn.py + py_shot(g, s);
    const float *__restrict__ Fz = A.fn.pz + pz_shot(g, s);
    // λ, μ-normal, μ_xy, μ_xz sums: one float4 per cell at the shot's base;
    // μ_yz sum: a scalar array at component offset 7
    float4 *__restrict__ acc4v = reinterpret_cast<float4 *>(A.acc + (long long)s * 8 * g.ps);
    float *__restrict__ accyz = A.acc + ((long long)s * 8 + 7) * g.ps;

    for (int y = y0; y < y1; ++y) {
        const int t = x + z * nx + y * sxz;
        const int sy0 = slab_coord(y, g.ny, g.wyl, g.wyh);
        const int sym = slab_coord(y - 1, g.ny, g.wyl, g.wyh);
        const int syp = slab_coord(y + 1, g.ny, g.wyl, g.wyh);
        const float ay0 = sy0 >= 0 ? ays[y] : 0.f, by0 = sy0 >= 0 ? bys[y] : 0.f;
        const float aym = sym >= 0 ? ays[y - 1] : 0.f, ayp = syp >= 0 ? ays[y + 1] : 0.f;

        const float txx = Txx[t], tyy = Tyy[t], tzz = Tzz[t];
        const float txy = Txy[t], txz = Txz[t], tyz = Tyz[t];

        // ē at p
        const float L0 = lam[t], M0 = mu[t];
        const float e_xx = c * ((L0 + 2.f * M0) * txx + L0 * (tyy + tzz));
        const float e_yy = c * ((L0 + 2.f * M0) * tyy + L0 * (txx + tzz));
        const float e_zz = c * ((L0 + 2.f * M0) * tzz + L0 * (txx + tyy));
        const float e_xy = c * mxy[t] * txy;
        const float e_xz = c * mxz[t] * txz;
        const float e_yz = c * myz[t] * tyz;
        // ē at the neighbours
        float exx_px, eyy_py, ezz_pz;
        {
            const int o = t + 1;
            const float L = lam[o], M = mu[o];
            exx_px = c * ((L + 2.f * M) * Txx[o] + L * (Tyy[o] + Tzz[o]));   // vx_x at p+x
        }
        {
            const int o = t + sxz;
            const float L = lam[o], M = mu[o];
            eyy_py = c * ((L + 2.f * M) * Tyy[o] + L * (Txx[o] + Tzz[o]));   // vy_y at p+y
        }
        {
            const int o = t + nx;
            const float L = lam[o], M = mu[o];
            ezz_pz = c * ((L + 2.f * M) * Tzz[o] + L * (Txx[o] + Tyy[o]));   // vz_z at p+z
        }
        const float exz_mz = c * mxz[t - nx] * Txz[t - nx];     // vx_z at p-z
        const float exy_my = c * mxy[t - sxz] * Txy[t - sxz];   // vx_y at p-y
        const float exy_mx = c * mxy[t - 1] * Txy[t - 1];       // vy_x at p-x
        const float eyz_mz = c * myz[t - nx] * Tyz[t - nx];     // vy_z at p-z
        const float exz_mx = c * mxz[t - 1] * Txz[t - 1];       // vz_x at p-x
        const float eyz_my = c * myz[t - sxz] * Tyz[t - sxz];   // vz_y at p-y

        float g_vxx_0 = e_xx, g_vxx_p = exx_px;     // vx_x
        float g_vyy_0 = e_yy, g_vyy_p = eyy_py;     // vy_y
        float g_vzz_0 = e_zz, g_vzz_p = ezz_pz;     // vz_z
        float g_vxz_m = exz_mz, g_vxz_0 = e_xz;     // vx_z
        float g_vxy_m = exy_my, g_vxy_0 = e_xy;     // vx_y
        float g_vyx_m = exy_mx, g_vyx_0 = e_xy;     // vy_x
        float g_vyz_m = eyz_mz, g_vyz_0 = e_yz;     // vy_z
        float g_vzx_m = exz_mx, g_vzx_0 = e_xz;     // vz_x
        float g_vzy_m = eyz_my, g_vzy_0 = e_yz;     // vz_y
        // x-slab components 0 vx_x, 1 vy_x, 2 vz_x
        if (sxp >= 0) g_vxx_p += axp * (Cx[px_off(g, y, z, sxp)] + exx_px);
        if (sxm >= 0) {
            const int pi = px_off(g, y, z, sxm) + csx;
            g_vyx_m += axm * (Cx[pi] + exy_mx);
            g_vzx_m += axm * (Cx[pi + csx] + exz_mx);
        }
        if (sx0 >= 0) {
            const int pi = px_off(g, y, z, sx0);
            const float c0 = Cx[pi] + e_xx, c1 = Cx[pi + csx] + e_xy, c2 = Cx[pi + 2 * csx] + e_xz;
            g_vxx_0 += ax0 * c0; g_vyx_0 += ax0 * c1; g_vzx_0 += ax0 * c2;
            Cxo[pi] = bx0 * c0; Cxo[pi + csx] = bx0 * c1; Cxo[pi + 2 * csx] = bx0 * c2;
        }
        // y-slab components 0 vx_y, 1 vy_y, 2 vz_y
        if (syp >= 0) g_vyy_p += ayp * (Cy[py_off(g, syp, z, x) + csy] + eyy_py);
        if (sym >= 0) {
            const int pi = py_off(g, sym, z, x);
            g_vxy_m += aym * (Cy[pi] + exy_my);
            g_vzy_m += aym * (Cy[pi + 2 * csy] + eyz_my);
        }
        if (sy0 >= 0) {
            const int pi = py_off(g, sy0, z, x);
            const float c0 = Cy[pi] + e_xy, c1 = Cy[pi + csy] + e_yy, c2 = Cy[pi + 2 * csy] + e_yz;
            g_vxy_0 += ay0 * c0; g_vyy_0 += ay0 * c1; g_vzy_0 += ay0 * c2;
            Cyo[pi] = by0 * c0; Cyo[pi + csy] = by0 * c1; Cyo[pi + 2 * csy] = by0 * c2;
        }
        // z-slab components 0 vx_z, 1 vy_z, 2 vz_z
        if (szp >= 0) g_vzz_p += azp * (Cz[pz_off(g, y, szp, x) + 2 * csz] + ezz_pz);
        if (szm >= 0) {
            const int pi = pz_off(g, y, szm, x);
            g_vxz_m += azm * (Cz[pi] + exz_mz);
            g_vyz_m += azm * (Cz[pi + csz] + eyz_mz);
        }
        if (sz0 >= 0) {
            const int pi = pz_off(g, y, sz0, x);
            const float c0 = Cz[pi] + e_xz, c1 = Cz[pi + csz] + e_yz, c2 = Cz[pi + 2 * csz] + e_zz;
            g_vxz_0 += az0 * c0; g_vyz_0 += az0 * c1; g_vzz_0 += az0 * c2;
            Czo[pi] = bz0 * c0; Czo[pi + csz] = bz0 * c1; Czo[pi + 2 * csz] = bz0 * c2;
        }
        // v̄ += D_vᵀ ḡ
        float nv0 = vb0[t] + ((g_vxx_0 - g_vxx_p) + (g_vxz_m - g_vxz_0) + (g_vxy_m - g_vxy_0));
        float nv1 = vb1[t] + ((g_vyy_0 - g_vyy_p) + (g_vyx_m - g_vyx_0) + (g_vyz_m - g_vyz_0));
        float nv2 = vb2[t] + ((g_vzz_0 - g_vzz_p) + (g_vzx_m - g_vzx_0) + (g_vzy_m - g_vzy_0));
        if (inject && is_receiver(g, sh, s, t)) {      // R1 of step n-1 at this cell's receivers
            for (int r = 0; r < g.nrec; ++r)
                if (rs[r] == t) {
                    nv0 += A.resid[rbase + r];
                    nv1 += A.resid[rbase + rcs + r];
                    nv2 += A.resid[rbase + 2 * rcs + r];
                }
        }
        vb0[t] = nv0;
        vb1[t] = nv1;
        vb2[t] = nv2;
        if (fused && t == tsrc) {                       // R3 of step n: ∂J/∂M += w[n]·τ̄(src)
            const double w = sh.wav[(long long)s * g.ntmax + n];
            double *gm = A.gm6 + 6 * s;
            gm[0] += w * (double)txx; gm[1] += w * (double)tyy; gm[2] += w * (double)tzz;
            gm[3] += w * (double)txy; gm[4] += w * (double)txz; gm[5] += w * (double)tyz;
        }

        // λ, μ accumulators: forward e at p from v^n and ψv^n
        const float v0 = LDF(&V0[t]), v1 = LDF(&V1[t]), v2 = LDF(&V2[t]);
        const float d_xx = v0 - LDF(&V0[t - 1]), d_yy = v1 - LDF(&V1[t - sxz]), d_zz = v2 - LDF(&V2[t - nx]);
        const float d_xz = LDF(&V0[t + nx]) - v0, d_xy = LDF(&V0[t + sxz]) - v0;
        const float d_yx = LDF(&V1[t + 1]) - v1, d_yz = LDF(&V1[t + nx]) - v1;
        const float d_zx = LDF(&V2[t + 1]) - v2, d_zy = LDF(&V2[t + sxz]) - v2;
        float fxx = d_xx, fyy = d_yy, fzz = d_zz, fxz = d_xz, fxy = d_xy;
        float fyx = d_yx, fyz = d_yz, fzx = d_zx, fzy = d_zy;
        if (sx0 >= 0) {
            const int pi = px_off(g, y, z, sx0);
            fxx += bx0 * LDF(&Fx[pi]) + ax0 * d_xx;
            fyx += bx0 * LDF(&Fx[pi + csx]) + ax0 * d_yx;
            fzx += bx0 * LDF(&Fx[pi + 2 * csx]) + ax0 * d_zx;
        }
        if (sy0 >= 0) {
            const int pi = py_off(g, sy0, z, x);
            fxy += by0 * LDF(&Fy[pi]) + ay0 * d_xy;
            fyy += by0 * LDF(&Fy[pi + csy]) + ay0 * d_yy;
            fzy += by0 * LDF(&Fy[pi + 2 * csy]) + ay0 * d_zy;
        }
        if (sz0 >= 0) {
            const int pi = pz_off(g, y, sz0, x);
            fxz += bz0 * LDF(&Fz[pi]) + az0 * d_xz;
            fyz += bz0 * LDF(&Fz[pi + csz]) + az0 * d_yz;
            fzz += bz0 * LDF(&Fz[pi + 2 * csz]) + az0 * d_zz;
        }
        float4 a = acc4v[t];
        a.x += (txx + tyy + tzz) * (fxx + fyy + fzz);
        a.y += 2.f * (txx * fxx + tyy * fyy + tzz * fzz);
        a.z += txy * (fxy + fyx);
        a.w += txz * (fxz + fzx);
        acc4v[t] = a;
        accyz[t] += tyz * (fyz + fzy);
    }
}

// ---------------------------------------------------------------------------
// R1 + R3 (one block per shot): moment-tensor sum of step n+1, residual of step n
// ---------------------------------------------------------------------------
__global__ void k_adj_points(Geo g, ShotPtrs sh, StatePtrs a, const float *resid, double *gm6,
                             int n, int n_m) {
    const int s = blockIdx.x;
    const int nts = sh.nt[s];
    if (threadIdx.x == 0 && n_m >= 0 && n_m < nts) {
        const int *sp = sh.src + 3 * s;
        const long long t = (long long)s * g.N + sp[0] + (long long)sp[2] * g.nx + (long long)sp[1] * g.sxz;
        const double w = sh.wav[(long long)s * g.ntmax + n_m];
        for (int k = 0; k < 6; ++k) gm6[6 * s + k] += w * (double)a.t[k][t];
    }
    if (n < 0 || n >= nts) return;
    for (int r = threadIdx.x; r < g.nrec; r += blockDim.x) {
        const int cell = sh.rec[(long long)s * g.nrec + r];
        if (cell < 0) continue;
        const long long t = (long long)s * g.N + cell;
        const long long base = ((long long)s * 3 * g.ntmax + n) * g.nrec + r;
        const long long cs = (long long)g.ntmax * g.nrec;
        atomicAdd(&a.v[0][t], resid[base]);
        atomicAdd(&a.v[1][t], resid[base + cs]);
        atomicAdd(&a.v[2][t], resid[base + 2 * cs]);
    }
}

// ---------------------------------------------------------------------------
// derived medium, gradient combine, model maps
// ---------------------------------------------------------------------------
__global__ void k_derived(Geo g, const float *mu, const float *rho, float *mbar, float *bbar) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    const int x = (int)(i % g.nx);
    const int z = (int)((i / g.nx) % g.nz);
    const int y = (int)(i / g.sxz);
    float mxy = 0.f, mxz = 0.f, myz = 0.f, bx = 0.f, by = 0.f, bz = 0.f;
    if (interior(x, y, z, g)) {
        const long long nx = g.nx, sxz = g.sxz;
        // the forward sweeps' own operations (forward.hip), so that the adjoint
        // is the transpose of the forward operator as implemented
        const float q0 = frcp(mu[i]), qi = frcp(mu[i + 1]), qk = frcp(mu[i + nx]);
        const float qj = frcp(mu[i + sxz]), qik = frcp(mu[i + 1 + nx]);
        const float qjk = frcp(mu[i + nx + sxz]), qij = frcp(mu[i + 1 + sxz]);
        mxz = 4.f * frcp(q0 + qi + qk + qik);
        mxy = 4.f * frcp(q0 + qi + qj + qij);
        myz = 4.f * frcp(q0 + qj + qk + qjk);
        const float r = rho[i];
        bx = 2.f * frcp(r + rho[i + 1]);
        by = 2.f * frcp(r + rho[i + sxz]);
        bz = 2.f * frcp(r + rho[i + nx]);
    }
    mbar[i] = mxy; mbar[i + g.ps] = mxz; mbar[i + 2 * g.ps] = myz;
    bbar[i] = bx; bbar[i + g.ps] = by; bbar[i + 2 * g.ps] = bz;
}

__global__ void k_combine(Geo g, const float *mu, const float *mbar, const float *bbar,
                          const float *acc, float *glam, float *gmu, float *grho) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    const int x = (int)(i % g.nx);
    const int z = (int)((i / g.nx) % g.nz);
    const int y = (int)(i / g.sxz);
    const long long N = g.ps, nx = g.nx, sxz = g.sxz;   // N: padded component stride
    const float c = g.c;
    const float *mxy = mbar, *mxz = mbar + N, *myz = mbar + 2 * N;
    (void)bbar;                   // the ρ sums already carry b̄
    float sl = 0.f, sm = 0.f, shear = 0.f, sr = 0.f;
    for (int s = 0; s < g.S; ++s) {
        const float *A = acc + (long long)s * 8 * N;
        // component k of cell j: 0-3 interleaved at the base, 4 (μ_yz) at 7·N
        auto AC = [&](int k, long long j) -> float { return k < 4 ? A[4 * j + k] : A[7 * N + j]; };
        sl += AC(0, i);
        sm += AC(1, i);
        auto sh4 = [&](int comp, const float *mb, long long o1, bool v1, long long o2, bool v2) {
            float r = AC(comp, i) * mb[i] * mb[i];
            if (v1) r += AC(comp, i - o1) * mb[i - o1] * mb[i - o1];
            if (v2) r += AC(comp, i - o2) * mb[i - o2] * mb[i - o2];
            if (v1 && v2) r += AC(comp, i - o1 - o2) * mb[i - o1 - o2] * mb[i - o1 - o2];
            return r;
        };
        shear += sh4(2, mxy, 1, x > 0, sxz, y > 0);
        shear += sh4(3, mxz, 1, x > 0, nx, z > 0);
        shear += sh4(4, myz, sxz, y > 0, nx, z > 0);
        // the ρ sum holds Σ c·b̄·v̄·Δv over the cell's six faces; per face
        // ∂J/∂ρ = -(c/2)·b̄²·v̄·F = -(1/(2c))·(c·b̄·v̄·Δv)
        sr += A[4 * N + i];
    }
    const float m = mu[i];
    glam[i] = c * sl;
    gmu[i] = c * (sm + (m != 0.f ? 0.25f * shear / (m * m) : 0.f));
    grho[i] = -0.5f / c * sr;
}

// r = mod - obs and J += ½Σr² (block partial sums in double, one atomic per block)
__global__ void k_residual(const float *mod, const float *obs, float *res, double *J, long long n) {
    __shared__ double part[256];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const float r = mod[i] - obs[i];
        res[i] = r;
        v = 0.5 * (double)r * (double)r;
    }
    part[threadIdx.x] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(J, part[0]);
}

__global__ void k_lame(const float *vp, const float *vs, const float *rho, float *lam, float *mu,
                       long long n) {
#pragma clang fp contract(off)
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float m = vs[i] * vs[i] * rho[i];
    mu[i] = m;
    lam[i] = vp[i] * vp[i] * rho[i] - 2.0f * m;
}

__global__ void k_lame_adjoint(const float *vp, const float *vs, const float *rho,
                               const float *glam, const float *gmu, const float *grho_in,
                               float *gvp, float *gvs, float *grho_out, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float p = vp[i], s = vs[i], r = rho[i], gl = glam[i], gm = gmu[i];
    gvp[i] = 2.f * r * p * gl;
    gvs[i] = 2.f * r * s * (gm - 2.f * gl);
    grho_out[i] = (p * p - 2.f * s * s) * gl + s * s * gm + (grho_in ? grho_in[i] : 0.f);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int launch_adj_velocity(const Geo &g, const AdjArgs &A, int n, hipStream_t st) {
    const Tile t = get_tile(2);
    bool launched = false;
#define X(a, b, c, o) EFM_LAUNCH_IF(a, b, c, o, k_adj_velocity, g, A, n)
    EFM_TILE_LIST(X)
#undef X
    if (!launched) return fail(EFM_E_UNSUPPORTED, "unsupported adjoint-velocity tile");
    return check_launch("adjoint velocity sweep");
}

int launch_adj_stress(const Geo &g, const AdjArgs &A, int n, hipStream_t st) {
    const Tile t = get_tile(3);
    bool launched = false;
#define X(a, b, c, o) EFM_LAUNCH_IF(a, b, c, o, k_adj_stress, g, A, n)
    EFM_TILE_LIST(X)
#undef X
    if (!launched) return fail(EFM_E_UNSUPPORTED, "unsupported adjoint-stress tile");
    return check_launch("adjoint stress sweep");
}

}  // namespace efm

using namespace efm;

static inline unsigned blocks_for(long long n, int b) { return (unsigned)((n + b - 1) / b); }

extern "C" int efm_derived_medium(const EfmDims *d, const EfmMedium *m, float *mbar, float *bbar,
                                  void *stream) {
    if (!d || !m || !mbar || !bbar || !m->mu || !m->rho)
        return fail(EFM_E_ARG, "efm_derived_medium: null argument");
    if (int rc = check_dims(d)) return rc;
    const Geo g = make_geo(*d);
    hipLaunchKernelGGL(k_derived, dim3(blocks_for(g.N, 256)), dim3(256), 0, (hipStream_t)stream, g,
                       m->mu, m->rho, mbar, bbar);
    return check_launch("derived medium");
}

namespace efm {

// The reverse sweep of steps n1-1 .. n0 (the body of efm_adjoint, shared with
// efm_gradient_static).  `copy_back` leaves the live ψ adjoint in adj's
// buffers when the ping-pong ended in adj2's.
static int run_adjoint(const EfmDims *d, const Geo &g, AdjArgs A, const EfmState *const *fwd,
                       const EfmState *adj, const EfmState *adj2, int n0, int n1, int last,
                       bool copy_back, hipStream_t st) {
    StatePtrs live = make_state(*adj), other = make_state(*adj2);
    const float *resid = A.resid;
    double *gm6 = A.gm6;
    const bool fused = A.sh.recmask != nullptr;
    if (fused && n1 == d->ntmax) {
        // first residual injection of the run; the sweeps do the rest
        hipLaunchKernelGGL(k_adj_points, dim3(g.S), dim3(64), 0, st, g, A.sh, live, resid, gm6,
                           n1 - 1, -1);
        int rc = check_launch("adjoint points");
        if (rc) return rc;
    }
    for (int n = n1 - 1; n >= n0; --n) {
        int rc = EFM_OK;
        if (!fused) {
            hipLaunchKernelGGL(k_adj_points, dim3(g.S), dim3(64), 0, st, g, A.sh, live, resid, gm6,
                               n, n + 1 < d->ntmax ? n + 1 : -1);
            if ((rc = check_launch("adjoint points"))) return rc;
        }
        A.fn = make_state(*fwd[n - n0]);
        A.fn1 = make_state(*fwd[n - n0 + 1]);
        A.a = live;
        A.a2 = other;
        A.a2.v[0] = live.v[0]; A.a2.v[1] = live.v[1]; A.a2.v[2] = live.v[2];
        for (int k = 0; k < 6; ++k) A.a2.t[k] = live.t[k];
        if ((rc = launch_adj_velocity(g, A, n, st))) return rc;
        // R4 reads the τ̄-side C from `other` (just written) and its own v̄-side C from `live`:
        // the two sweeps touch disjoint ψ components, so swap only after both.
        if ((rc = launch_adj_stress(g, A, n, st))) return rc;
        std::swap(live.px, other.px); std::swap(live.py, other.py); std::swap(live.pz, other.pz);
    }
    if (!fused && last && n0 == 0) {
        hipLaunchKernelGGL(k_adj_points, dim3(g.S), dim3(64), 0, st, g, A.sh, live, resid, gm6, -1, 0);
        int rc = check_launch("adjoint points");
        if (rc) return rc;
    }
    // leave the live ψ adjoint in adj's buffers (copy back when they were swapped)
    if (copy_back && live.px != adj->psi_x) {
        const long long fx = (long long)g.S * 6 * g.ny * g.nz * g.Lx;
        const long long fy = (long long)g.S * 6 * g.Ly * g.nz * g.nx;
        const long long fz = (long long)g.S * 6 * g.ny * g.Lz * g.nx;
        if (fx && hipMemcpyAsync(adj->psi_x, live.px, fx * 4, hipMemcpyDeviceToDevice, st))
            return fail(EFM_E_HIP, "efm_adjoint: psi copy");
        if (fy && hipMemcpyAsync(adj->psi_y, live.py, fy * 4, hipMemcpyDeviceToDevice, st))
            return fail(EFM_E_HIP, "efm_adjoint: psi copy");
        if (fz && hipMemcpyAsync(adj->psi_z, live.pz, fz * 4, hipMemcpyDeviceToDevice, st))
            return fail(EFM_E_HIP, "efm_adjoint: psi copy");
    }
    return EFM_OK;
}

}  // namespace efm

extern "C" int efm_adjoint(const EfmDims *d, const EfmMedium *m, const float *mbar,
                           const float *bbar, const EfmShots *sh, const EfmState *const *fwd,
                           const EfmState *adj, const EfmState *adj2, const float *resid,
                           float *acc, double *gm6, int n0, int n1, int last, void *stream) {
    if (!d || !m || !mbar || !bbar || !sh || !fwd || !adj || !adj2 || (!resid && d->nrec > 0) ||
        !acc || !gm6)
        return fail(EFM_E_ARG, "efm_adjoint: null argument");
    if (int rc = validate(d, m, sh)) return rc;
    if (n0 < 0 || n1 <= n0 || n1 > d->ntmax) return fail(EFM_E_ARG, "efm_adjoint: bad step range");
    if (int rc = validate_state(d, adj)) return rc;
    // adj2 carries only the second ψ buffer
    if ((d->wx_lo + d->wx_hi > 0 && !adj2->psi_x) || (d->wy_lo + d->wy_hi > 0 && !adj2->psi_y) ||
        (d->wz_lo + d->wz_hi > 0 && !adj2->psi_z))
        return fail(EFM_E_ARG, "efm_adjoint: adj2 psi slab pointer is null");
    for (int k = 0; k <= n1 - n0; ++k)
        if (int rc = validate_state(d, fwd[k])) return rc;
    const Geo g = make_geo(*d);
    AdjArgs A;
    A.lam = m->lam; A.mu = m->mu; A.mbar = mbar; A.bbar = bbar; A.acc = acc;
    A.sh = make_shots(*sh);
    A.resid = resid;
    A.gm6 = gm6;
    return run_adjoint(d, g, A, fwd, adj, adj2, n0, n1, last, true, (hipStream_t)stream);
}

extern "C" int efm_gradient_combine(const EfmDims *d, const EfmMedium *m, const float *mbar,
                                    const float *bbar, const float *acc, float *glam, float *gmu,
                                    float *grho, void *stream) {
    if (!d || !m || !m->mu || !mbar || !bbar || !acc || !glam || !gmu || !grho)
        return fail(EFM_E_ARG, "efm_gradient_combine: null argument");
    if (int rc = check_dims(d)) return rc;
    const Geo g = make_geo(*d);
    hipLaunchKernelGGL(k_combine, dim3(blocks_for(g.N, 256)), dim3(256), 0, (hipStream_t)stream, g,
                       m->mu, mbar, bbar, acc, glam, gmu, grho);
    return check_launch("gradient combine");
}

extern "C" int efm_residual(const float *mod, const float *obs, float *res, double *J, long long n,
                            void *stream) {
    if (n == 0) return EFM_OK;                   // no receivers: nothing to compare
    if (!mod || !obs || !res || !J || n < 0) return fail(EFM_E_ARG, "efm_residual: bad argument");
    hipLaunchKernelGGL(k_residual, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, mod,
                       obs, res, J, n);
    return check_launch("residual");
}

extern "C" int efm_model_lame(const float *vp, const float *vs, const float *rho, float *lam,
                              float *mu, long long n, void *stream) {
    if (!vp || !vs || !rho || !lam || !mu || n < 0) return fail(EFM_E_ARG, "efm_model_lame: bad argument");
    if (n == 0) return EFM_OK;
    hipLaunchKernelGGL(k_lame, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, vp, vs,
                       rho, lam, mu, n);
    return check_launch("model lame");
}

extern "C" int efm_model_lame_adjoint(const float *vp, const float *vs, const float *rho,
                                      const float *glam, const float *gmu, const float *grho_in,
                                      float *gvp, float *gvs, float *grho_out, long long n,
                                      void *stream) {
    if (!vp || !vs || !rho || !glam || !gmu || !gvp || !gvs || !grho_out || n < 0)
        return fail(EFM_E_ARG, "efm_model_lame_adjoint: bad argument");
    if (n == 0) return EFM_OK;
    hipLaunchKernelGGL(k_lame_adjoint, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       vp, vs, rho, glam, gmu, grho_in, gvp, gvs, grho_out, n);
    return check_launch("model lame adjoint");
}

// ---------------------------------------------------------------------------
// Allocation-free whole gradient (no reference counterpart): the derived medium,
// the adjoint accumulators and the second ψ-adjoint buffer live in a
// file-scope device array instead of caller workspaces.
// ---------------------------------------------------------------------------
namespace efm {

constexpr long long kScratchFloats = 16LL << 20;   // 64 MiB per device
static __device__ __align__(256) float scratch[kScratchFloats];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static float *g_scratch_dev[kMaxDevices];   // device address of `scratch`, per device

// The calling thread's current device's copy of `scratch` (looked up once per
// device, with that device current).
static int scratch_address(float **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return fail(EFM_E_HIP, "efm_gradient_static: no current device");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(EFM_E_HIP, std::string("efm_gradient_static: hipGetSymbolAddress: ") +
                                       hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<float *>(p);
    }
    *out = g_scratch_dev[dev];
    return EFM_OK;
}

template <typename T>
__global__ void k_fill_zero(T *p, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = T(0);
}

template <typename T>
static int zero_fill(T *p, long long n, hipStream_t st) {
    if (n <= 0 || !p) return EFM_OK;
    const long long b = (n + 255) / 256;
    hipLaunchKernelGGL(k_fill_zero<T>, dim3((unsigned)(b < 65536 ? b : 65536)), dim3(256), 0, st, p, n);
    return check_launch("zero fill");
}

}  // namespace efm

extern "C" int efm_gradient_static(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                                   const EfmState *const *fwd, const EfmState *adj,
                                   const float *resid, float *glam, float *gmu, float *grho,
                                   double *gm6, void *stream) {
    if (!d || !m || !sh || !fwd || !adj || (!resid && d->nrec > 0) || !glam || !gmu || !grho || !gm6)
        return fail(EFM_E_ARG, "efm_gradient_static: null argument");
    if (int rc = validate(d, m, sh)) return rc;
    if (int rc = validate_state(d, adj)) return rc;
    for (int k = 0; k <= d->ntmax; ++k)
        if (int rc = validate_state(d, fwd[k])) return rc;
    const Geo g = make_geo(*d);
    const long long psi = (long long)g.S * 6 * ((long long)g.pxc + g.pyc + g.pzc);
    const long long need = (6 + 8LL * g.S) * g.ps + psi;
    if (need > kScratchFloats)
        return fail(EFM_E_ARG, "efm_gradient_static: the batch needs " + std::to_string(need * 4) +
                                   " bytes of scratch, above the library's static " +
                                   std::to_string(kScratchFloats * 4) + " (use efm_adjoint with "
                                   "caller workspaces)");
    float *base = nullptr;
    if (int rc = scratch_address(&base)) return rc;
    hipStream_t st = (hipStream_t)stream;
    float *mbar = base, *bbar = base + 3 * g.ps, *acc = base + 6 * g.ps;
    float *psi2 = acc + 8LL * g.S * g.ps;
    // what must start at zero: accumulators, the second ψ buffer, the adjoint
    // state and ∂J/∂M (fill kernels, no memset: callers may capture this call)
    int rc = zero_fill(acc, 8LL * g.S * g.ps + psi, st);
    const long long vol = (long long)g.S * g.N;
    float *fields[9] = {adj->vx, adj->vy, adj->vz, adj->txx, adj->tyy, adj->tzz, adj->txy, adj->txz, adj->tyz};
    for (int k = 0; k < 9 && !rc; ++k) rc = zero_fill(fields[k], vol, st);
    if (!rc) rc = zero_fill(adj->psi_x, (long long)g.S * 6 * g.pxc, st);
    if (!rc) rc = zero_fill(adj->psi_y, (long long)g.S * 6 * g.pyc, st);
    if (!rc) rc = zero_fill(adj->psi_z, (long long)g.S * 6 * g.pzc, st);
    if (!rc) rc = zero_fill(gm6, 6LL * g.S, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_derived, dim3(blocks_for(g.N, 256)), dim3(256), 0, st, g, m->mu, m->rho, mbar, bbar);
    if ((rc = check_launch("derived medium"))) return rc;
    EfmState second = *adj;
    second.psi_x = psi2;
    second.psi_y = psi2 + (long long)g.S * 6 * g.pxc;
    second.psi_z = second.psi_y + (long long)g.S * 6 * g.pyc;
    AdjArgs A;
    A.lam = m->lam; A.mu = m->mu; A.mbar = mbar; A.bbar = bbar; A.acc = acc;
    A.sh = make_shots(*sh);
    A.resid = resid;
    A.gm6 = gm6;
    if ((rc = run_adjoint(d, g, A, fwd, adj, &second, 0, d->ntmax, 1, false, st))) return rc;
    hipLaunchKernelGGL(k_combine, dim3(blocks_for(g.N, 256)), dim3(256), 0, st, g, m->mu, mbar, bbar,
                       acc, glam, gmu, grho);
    return check_launch("gradient combine");
}
