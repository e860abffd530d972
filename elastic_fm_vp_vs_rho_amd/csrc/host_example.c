/* A C host of libefm.so without Python or torch: one shot's forward run with
 * every pre-step state kept, the L2 residual against observed gathers, and the
 * exact gradient through the allocation-free entry point.  It links ROCm's HIP
 * runtime itself, which libefm.so's HIP symbols then bind to.
 *
 *   efm_host_example IN OUT
 *
 * IN (little-endian): int32 nx, ny, nz, wx_lo, wx_hi, wy_lo, wy_hi, wz_lo, wz_hi,
 * nt, nrec; float32 dtdx; float32 lam, mu, rho (ny·nz·nx each, x fastest);
 * float32 ax, bx (nx), ay, by (ny), az, bz (nz); int32 src (x, y, z); float32
 * m6 (6); float32 wavelet (nt); int32 rec (nrec linear cell indices
 * x + nx·z + nx·nz·y); float32 obs (3·nt·nrec).
 * OUT: float32 gathers (3·nt·nrec); float64 J; float32 glam, gmu, grho
 * (ny·nz·nx each); float64 gm6 (6).
 * Exit status 0 on success; a failing call prints efm_last_error(). */
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "efm.h"

#define HIPCHECK(x)                                                          \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess) {                                              \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));          \
            exit(2);                                                         \
        }                                                                    \
    } while (0)

static void efm_check(int rc, const char *what) {
    if (rc != EFM_OK) {
        char msg[512];
        efm_last_error(msg, sizeof msg);
        fprintf(stderr, "%s failed (%d): %s\n", what, rc, msg);
        exit(3);
    }
}

static void read_exact(FILE *f, void *p, size_t bytes) {
    if (fread(p, 1, bytes, f) != bytes) {
        fprintf(stderr, "short input\n");
        exit(4);
    }
}

/* Host array of n items of size sz read from f, copied to a new device buffer. */
static void *to_device(FILE *f, size_t n, size_t sz) {
    void *h = malloc(n * sz), *d = NULL;
    if (!h) exit(5);
    read_exact(f, h, n * sz);
    HIPCHECK(hipMalloc(&d, n * sz));
    HIPCHECK(hipMemcpy(d, h, n * sz, hipMemcpyHostToDevice));
    free(h);
    return d;
}

static float *zeros(size_t n) {
    float *d = NULL;
    HIPCHECK(hipMalloc((void **)&d, n * sizeof(float)));
    HIPCHECK(hipMemset(d, 0, n * sizeof(float)));
    return d;
}

static void free_state(EfmState *s) {
    float *v[12] = {s->vx, s->vy, s->vz, s->txx, s->tyy, s->tzz, s->txy, s->txz, s->tyz,
                    s->psi_x, s->psi_y, s->psi_z};
    for (int k = 0; k < 12; ++k) HIPCHECK(hipFree(v[k]));
}

static EfmState new_state(size_t N, size_t npx, size_t npy, size_t npz) {
    EfmState s;
    float **v[9] = {&s.vx, &s.vy, &s.vz, &s.txx, &s.tyy, &s.tzz, &s.txy, &s.txz, &s.tyz};
    for (int k = 0; k < 9; ++k) *v[k] = zeros(N);
    s.psi_x = zeros(npx);
    s.psi_y = zeros(npy);
    s.psi_z = zeros(npz);
    return s;
}

int main(int argc, char **argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
        return 1;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 1;
    int hdr[11];
    float dtdx;
    read_exact(f, hdr, sizeof hdr);
    read_exact(f, &dtdx, sizeof dtdx);
    EfmDims d;
    d.nx = hdr[0]; d.ny = hdr[1]; d.nz = hdr[2]; d.nshot = 1;
    d.wx_lo = hdr[3]; d.wx_hi = hdr[4]; d.wy_lo = hdr[5]; d.wy_hi = hdr[6];
    d.wz_lo = hdr[7]; d.wz_hi = hdr[8];
    d.ntmax = hdr[9]; d.nrec = hdr[10]; d.dtdx = dtdx;
    const size_t N = (size_t)d.nx * d.ny * d.nz;
    const int nt = d.ntmax, nrec = d.nrec;

    EfmMedium m;
    m.lam = to_device(f, N, 4);
    m.mu = to_device(f, N, 4);
    m.rho = to_device(f, N, 4);
    EfmShots sh;
    sh.ax = to_device(f, d.nx, 4); sh.bx = to_device(f, d.nx, 4);
    sh.ay = to_device(f, d.ny, 4); sh.by = to_device(f, d.ny, 4);
    sh.az = to_device(f, d.nz, 4); sh.bz = to_device(f, d.nz, 4);
    sh.src = to_device(f, 3, 4);
    sh.m6 = to_device(f, 6, 4);
    sh.wavelet = to_device(f, nt, 4);
    sh.rec = to_device(f, nrec, 4);
    sh.recmask = NULL;           /* receivers by the separate per-step kernels */
    int *ntd = NULL;
    HIPCHECK(hipMalloc((void **)&ntd, sizeof(int)));
    HIPCHECK(hipMemcpy(ntd, &nt, sizeof(int), hipMemcpyHostToDevice));
    sh.nt = ntd;
    const size_t ng = (size_t)3 * nt * nrec;
    float *obs = to_device(f, ng, 4);
    fclose(f);

    /* ψ slabs: psi_x (6, ny, nz, Lx), psi_y (6, Ly, nz, nx), psi_z (6, ny, Lz, nx) */
    const size_t Lx = d.wx_lo + d.wx_hi, Ly = d.wy_lo + d.wy_hi, Lz = d.wz_lo + d.wz_hi;
    const size_t npx = 6 * (size_t)d.ny * d.nz * Lx;
    const size_t npy = 6 * Ly * d.nz * d.nx;
    const size_t npz = 6 * (size_t)d.ny * Lz * d.nx;
    EfmState *states = malloc((size_t)(nt + 1) * sizeof(EfmState));
    const EfmState **fwd = malloc((size_t)(nt + 1) * sizeof(EfmState *));
    if (!states || !fwd) return 5;
    for (int k = 0; k <= nt; ++k) {
        states[k] = new_state(N, npx, npy, npz);
        fwd[k] = &states[k];
    }
    EfmState adj = new_state(N, npx, npy, npz);
    float *gathers = zeros(ng), *resid = zeros(ng);
    float *glam = zeros(N), *gmu = zeros(N), *grho = zeros(N);
    double *J = NULL, *gm6 = NULL;
    HIPCHECK(hipMalloc((void **)&J, sizeof(double)));
    HIPCHECK(hipMemset(J, 0, sizeof(double)));
    HIPCHECK(hipMalloc((void **)&gm6, 6 * sizeof(double)));

    hipStream_t st;
    HIPCHECK(hipStreamCreate(&st));
    efm_check(efm_forward_states(&d, &m, &sh, fwd, gathers, 0, nt, st), "efm_forward_states");
    efm_check(efm_residual(gathers, obs, resid, J, (long long)ng, st), "efm_residual");
    efm_check(efm_gradient_static(&d, &m, &sh, fwd, &adj, resid, glam, gmu, grho, gm6, st),
              "efm_gradient_static");
    HIPCHECK(hipStreamSynchronize(st));

    FILE *o = fopen(argv[2], "wb");
    if (!o) return 1;
    float *hb = malloc((ng > N ? ng : N) * sizeof(float));
    double hd[6];
    if (!hb) return 5;
    HIPCHECK(hipMemcpy(hb, gathers, ng * 4, hipMemcpyDeviceToHost));
    fwrite(hb, 4, ng, o);
    HIPCHECK(hipMemcpy(hd, J, 8, hipMemcpyDeviceToHost));
    fwrite(hd, 8, 1, o);
    const float *vols[3] = {glam, gmu, grho};
    for (int k = 0; k < 3; ++k) {
        HIPCHECK(hipMemcpy(hb, vols[k], N * 4, hipMemcpyDeviceToHost));
        fwrite(hb, 4, N, o);
    }
    HIPCHECK(hipMemcpy(hd, gm6, 48, hipMemcpyDeviceToHost));
    fwrite(hd, 8, 6, o);
    fclose(o);
    free(hb);

    for (int k = 0; k <= nt; ++k) free_state(&states[k]);
    free_state(&adj);
    free(states);
    free(fwd);
    const void *bufs[] = {m.lam, m.mu, m.rho, sh.ax, sh.bx, sh.ay, sh.by, sh.az, sh.bz, sh.src,
                          sh.m6, sh.wavelet, sh.rec, ntd, obs, gathers, resid, glam, gmu, grho, J, gm6};
    for (size_t k = 0; k < sizeof bufs / sizeof bufs[0]; ++k) HIPCHECK(hipFree((void *)bufs[k]));
    HIPCHECK(hipStreamDestroy(st));
    printf("efm_host_example: %dx%dx%d, %d steps, %d receivers: done\n", d.nx, d.ny, d.nz, nt, nrec);
    return 0;
}
