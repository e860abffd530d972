/* Scalar-loop FP64 restatement of one forward time step of the reference
 * propagator (kernel_Tau M:59-128, addSource M:1728-1734, kernel_Vx_Vy_Vz
 * M:133-188 of mainPropElastic3D_MT.py).
 *
 * TEST INFRASTRUCTURE ONLY: it exists to pin the vectorised NumPy oracle
 * (oracle/elastic_fp64.py) with an independent transcription that indexes
 * cells one at a time exactly like the reference (tid = x + z*nx + y*nz*nx).
 *
 * Arrays are double, (ny, nz, nx) x-fastest.  f[0..8] = vx vy vz txx tyy tzz
 * txy txz tyz; pv[0..8] = ψ of vx_x vy_x vz_x vx_y vy_y vz_y vx_z vy_z vz_z;
 * pt[0..8] = ψ of txx_x txy_x txz_x txy_y tyy_y tyz_y txz_z tyz_z tzz_z.
 */
#include <stddef.h>

#define H 2

void twin_step(int nx, int ny, int nz, double c,
               double **f, double **pv, double **pt,
               const double *lam, const double *mu, const double *rho,
               const double *ax, const double *bx, const double *ay,
               const double *by, const double *az, const double *bz,
               int sx, int sy, int sz, const double *m6, double amp)
{
    const size_t sxz = (size_t)nx * nz;
    double *vx = f[0], *vy = f[1], *vz = f[2];
    double *txx = f[3], *tyy = f[4], *tzz = f[5], *txy = f[6], *txz = f[7], *tyz = f[8];

    /* stress sweep */
    for (int y = H; y < ny - H; ++y)
        for (int z = H; z < nz - H; ++z)
            for (int x = H; x < nx - H; ++x) {
                size_t t = (size_t)x + (size_t)z * nx + (size_t)y * sxz;
                double Mu = mu[t], L = lam[t];
                double Mi = mu[t + 1], Mk = mu[t + nx], Mj = mu[t + sxz];
                double Mik = mu[t + 1 + nx], Mjk = mu[t + nx + sxz], Mij = mu[t + 1 + sxz];
                double mxz = 1.0 / (0.25 / Mu + 0.25 / Mi + 0.25 / Mk + 0.25 / Mik);
                double mxy = 1.0 / (0.25 / Mu + 0.25 / Mi + 0.25 / Mj + 0.25 / Mij);
                double myz = 1.0 / (0.25 / Mu + 0.25 / Mj + 0.25 / Mk + 0.25 / Mjk);
                double dvx_dx = vx[t] - vx[t - 1];
                double dvy_dy = vy[t] - vy[t - sxz];
                double dvz_dz = vz[t] - vz[t - nx];
                double dvx_dz = vx[t + nx] - vx[t];
                double dvx_dy = vx[t + sxz] - vx[t];
                double dvy_dx = vy[t + 1] - vy[t];
                double dvy_dz = vy[t + nx] - vy[t];
                double dvz_dx = vz[t + 1] - vz[t];
                double dvz_dy = vz[t + sxz] - vz[t];
                pv[0][t] = bx[x] * pv[0][t] + ax[x] * dvx_dx;
                pv[1][t] = bx[x] * pv[1][t] + ax[x] * dvy_dx;
                pv[2][t] = bx[x] * pv[2][t] + ax[x] * dvz_dx;
                pv[3][t] = by[y] * pv[3][t] + ay[y] * dvx_dy;
                pv[4][t] = by[y] * pv[4][t] + ay[y] * dvy_dy;
                pv[5][t] = by[y] * pv[5][t] + ay[y] * dvz_dy;
                pv[6][t] = bz[z] * pv[6][t] + az[z] * dvx_dz;
                pv[7][t] = bz[z] * pv[7][t] + az[z] * dvy_dz;
                pv[8][t] = bz[z] * pv[8][t] + az[z] * dvz_dz;
                double l2m = L + 2.0 * Mu;
                txx[t] += c * (l2m * (dvx_dx + pv[0][t]) + L * (dvy_dy + dvz_dz + pv[4][t] + pv[8][t]));
                tyy[t] += c * (l2m * (dvy_dy + pv[4][t]) + L * (dvx_dx + dvz_dz + pv[0][t] + pv[8][t]));
                tzz[t] += c * (l2m * (dvz_dz + pv[8][t]) + L * (dvx_dx + dvy_dy + pv[0][t] + pv[4][t]));
                txy[t] += c * mxy * ((dvx_dy + pv[3][t]) + (dvy_dx + pv[1][t]));
                txz[t] += c * mxz * ((dvx_dz + pv[6][t]) + (dvz_dx + pv[2][t]));
                tyz[t] += c * myz * ((dvy_dz + pv[7][t]) + (dvz_dy + pv[5][t]));
            }

    /* source */
    {
        size_t t = (size_t)sx + (size_t)sz * nx + (size_t)sy * sxz;
        for (int k = 0; k < 6; ++k) f[3 + k][t] += m6[k] * amp;
    }

    /* velocity sweep */
    for (int y = H; y < ny - H; ++y)
        for (int z = H; z < nz - H; ++z)
            for (int x = H; x < nx - H; ++x) {
                size_t t = (size_t)x + (size_t)z * nx + (size_t)y * sxz;
                double b = 1.0 / rho[t], bi = 1.0 / rho[t + 1];
                double bk = 1.0 / rho[t + nx], bj = 1.0 / rho[t + sxz];
                double b_x = 2.0 * b * bi / (b + bi);
                double b_y = 2.0 * b * bj / (b + bj);
                double b_z = 2.0 * b * bk / (b + bk);
                double dtxx_dx = txx[t + 1] - txx[t];
                double dtzz_dz = tzz[t + nx] - tzz[t];
                double dtyy_dy = tyy[t + sxz] - tyy[t];
                double dtxy_dy = txy[t] - txy[t - sxz];
                double dtxy_dx = txy[t] - txy[t - 1];
                double dtxz_dz = txz[t] - txz[t - nx];
                double dtxz_dx = txz[t] - txz[t - 1];
                double dtyz_dz = tyz[t] - tyz[t - nx];
                double dtyz_dy = tyz[t] - tyz[t - sxz];
                pt[0][t] = bx[x] * pt[0][t] + ax[x] * dtxx_dx;
                pt[1][t] = bx[x] * pt[1][t] + ax[x] * dtxy_dx;
                pt[2][t] = bx[x] * pt[2][t] + ax[x] * dtxz_dx;
                pt[3][t] = by[y] * pt[3][t] + ay[y] * dtxy_dy;
                pt[4][t] = by[y] * pt[4][t] + ay[y] * dtyy_dy;
                pt[5][t] = by[y] * pt[5][t] + ay[y] * dtyz_dy;
                pt[6][t] = bz[z] * pt[6][t] + az[z] * dtxz_dz;
                pt[7][t] = bz[z] * pt[7][t] + az[z] * dtyz_dz;
                pt[8][t] = bz[z] * pt[8][t] + az[z] * dtzz_dz;
                vx[t] += c * b_x * ((dtxx_dx + pt[0][t]) + (dtxy_dy + pt[3][t]) + (dtxz_dz + pt[6][t]));
                vy[t] += c * b_y * ((dtxy_dx + pt[1][t]) + (dtyy_dy + pt[4][t]) + (dtyz_dz + pt[7][t]));
                vz[t] += c * b_z * ((dtxz_dx + pt[2][t]) + (dtyz_dy + pt[5][t]) + (dtzz_dz + pt[8][t]));
            }
}
