// extern "C" entry points of libefm.so that are not defined next to their
// kernels, plus error plumbing and the tile registry used by the autotuner.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>

#include "common.hpp"

namespace efm {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(EFM_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return EFM_OK;
}

// Tiles: 0 stress, 1 velocity, 2 adjoint velocity, 3 adjoint stress.  The
// default, blocks of 128 consecutive interior cells in the XCD-aware order
// (map 16), keeps every wave full whatever nx is: measured within 0-2 % of the
// best row tile at 128^3, 4-8 % behind it where rows fill whole waves
// (256 wide), and 15-25 % ahead where they do not (200 and 132 wide).
static std::mutex g_tile_mu;
static Tile g_tiles[4] = {{128, 1, 1, 16}, {128, 1, 1, 16}, {128, 1, 1, 16}, {128, 1, 1, 16}};
#define EFM_TILE_ENTRY(a, b, c, o) {a, b, c, o},
static const Tile kTiles[] = {EFM_TILE_LIST(EFM_TILE_ENTRY)};
#undef EFM_TILE_ENTRY

Tile get_tile(int kernel) {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    return g_tiles[kernel];
}

int set_tile(int kernel, Tile t) {
    if (kernel < 0 || kernel > 3) return fail(EFM_E_ARG, "tile kernel id out of range");
    for (const Tile &k : kTiles)
        if (k.bx == t.bx && k.bz == t.bz && k.ny == t.ny && k.map == t.map) {
            std::lock_guard<std::mutex> lk(g_tile_mu);
            g_tiles[kernel] = t;
            return EFM_OK;
        }
    return fail(EFM_E_UNSUPPORTED, "tile not compiled into this library");
}

// Work-items of one sweep dispatch with tile t (sweep_grid's launch shape,
// counted in 64 bits).
static long long sweep_items(const Geo &g, const Tile &t) {
    const long long per = (long long)t.bx * t.bz;
    if (t.map > 0) return xcd_blocks(g, t) * per;
    const long long nbx = (g.nx + t.bx - 1) / t.bx, nbz = (g.nz + t.bz - 1) / t.bz;
    const long long nyc = (g.ny - 2 * H + t.ny - 1) / t.ny;
    return nbx * nbz * nyc * g.S * per;
}

int check_dims(const EfmDims *d) {
    if (!d) return fail(EFM_E_ARG, "null dimensions");
    if (d->nx < 5 || d->ny < 5 || d->nz < 5 || d->nshot < 1)
        return fail(EFM_E_ARG, "grid must be at least 5 cells per axis and nshot >= 1");
    if (d->wx_lo < 0 || d->wx_hi < 0 || d->wy_lo < 0 || d->wy_hi < 0 || d->wz_lo < 0 ||
        d->wz_hi < 0 || d->wx_lo + d->wx_hi > d->nx || d->wy_lo + d->wy_hi > d->ny ||
        d->wz_lo + d->wz_hi > d->nz)
        return fail(EFM_E_ARG, "CPML widths out of range");
    if (d->ntmax < 1 || d->nrec < 0) return fail(EFM_E_ARG, "ntmax must be >= 1 and nrec >= 0");
    // the sweeps address a shot's cells and CPML slab components with 32-bit offsets
    const long long N = (long long)d->nx * d->ny * d->nz;
    const long long slab = 6LL * std::max({(long long)d->ny * d->nz * (d->wx_lo + d->wx_hi),
                                           (long long)(d->wy_lo + d->wy_hi) * d->nz * d->nx,
                                           (long long)d->ny * (d->wz_lo + d->wz_hi) * d->nx});
    if (N > INT_MAX || slab > INT_MAX)
        return fail(EFM_E_ARG, "grid too large: a shot's cells and its CPML slab values must each "
                               "number below 2^31");
    // every shot of the batch is in one dispatch per sweep
    const Geo g = make_geo(*d);
    for (int k = 0; k < 4; ++k)
        if (sweep_items(g, get_tile(k)) > (long long)UINT32_MAX)
            return fail(EFM_E_ARG, "shot batch too large for one sweep dispatch (2^32 work-items): "
                                   "split it into smaller batches");
    return EFM_OK;
}

int launch_stress_src(const Geo &g, const EfmMedium &m, const ShotPtrs &sh, const StatePtrs &in,
                      const StatePtrs &out, int it, int src_mode, hipStream_t st);

int validate_shots(const EfmDims *d, const EfmShots *sh) {
    if (!sh) return fail(EFM_E_ARG, "null shot description");
    if (!sh->ax || !sh->bx || !sh->ay || !sh->by || !sh->az || !sh->bz || !sh->src || !sh->m6 ||
        !sh->wavelet || !sh->nt || (d->nrec > 0 && !sh->rec))
        return fail(EFM_E_ARG, "shot pointer is null");
    return EFM_OK;
}

int validate(const EfmDims *d, const EfmMedium *m, const EfmShots *sh) {
    if (!d || !m || !sh) return fail(EFM_E_ARG, "null argument");
    int rc = check_dims(d);
    if (rc) return rc;
    if (!m->lam || !m->mu || !m->rho) return fail(EFM_E_ARG, "medium pointer is null");
    return validate_shots(d, sh);
}

int validate_state(const EfmDims *d, const EfmState *s) {
    if (!s || !s->vx || !s->vy || !s->vz || !s->txx || !s->tyy || !s->tzz || !s->txy ||
        !s->txz || !s->tyz)
        return fail(EFM_E_ARG, "state pointer is null");
    // the ψ slabs of an axis are touched whenever that axis has a CPML layer
    if ((d->wx_lo + d->wx_hi > 0 && !s->psi_x) || (d->wy_lo + d->wy_hi > 0 && !s->psi_y) ||
        (d->wz_lo + d->wz_hi > 0 && !s->psi_z))
        return fail(EFM_E_ARG, "state psi slab pointer is null");
    return EFM_OK;
}

}  // namespace efm

using namespace efm;

extern "C" int efm_version(void) { return 10000; }  // 1.0.0

extern "C" int efm_last_error(char *buf, int len) {
    if (!buf || len <= 0) return EFM_E_ARG;
    std::strncpy(buf, g_err.c_str(), (size_t)len - 1);
    buf[len - 1] = '\0';
    return (int)g_err.size();
}

extern "C" int efm_component_pad(void) { return EFM_PAD; }

extern "C" int efm_set_tile(int kernel, int bx, int bz, int ny, int map) {
    return set_tile(kernel, Tile{bx, bz, ny, map});
}

extern "C" int efm_get_tile(int kernel, int *bx, int *bz, int *ny, int *map) {
    if (kernel < 0 || kernel > 3 || !bx || !bz || !ny || !map)
        return fail(EFM_E_ARG, "efm_get_tile: bad argument");
    const Tile t = get_tile(kernel);
    *bx = t.bx; *bz = t.bz; *ny = t.ny; *map = t.map;
    return EFM_OK;
}

extern "C" int efm_tile_variants(int *out, int max) {
    const int n = (int)(sizeof(kTiles) / sizeof(kTiles[0]));
    if (out)
        for (int i = 0; i < n && i < max; ++i) {
            out[4 * i] = kTiles[i].bx; out[4 * i + 1] = kTiles[i].bz;
            out[4 * i + 2] = kTiles[i].ny; out[4 * i + 3] = kTiles[i].map;
        }
    return n;
}

extern "C" int efm_stress_step(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                               const EfmState *in, const EfmState *out, int it, void *stream) {
    int rc = validate(d, m, sh);
    if (rc || (rc = validate_state(d, in)) || (rc = validate_state(d, out))) return rc;
    return launch_stress_src(make_geo(*d), *m, make_shots(*sh), make_state(*in), make_state(*out),
                             it, 1, (hipStream_t)stream);
}

extern "C" int efm_velocity_step(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                                 const EfmState *in, const EfmState *out, int it, void *stream) {
    int rc = validate(d, m, sh);
    if (rc || (rc = validate_state(d, in)) || (rc = validate_state(d, out))) return rc;
    return launch_velocity(make_geo(*d), *m, make_shots(*sh), make_state(*in), make_state(*out), it,
                           (hipStream_t)stream);
}

extern "C" int efm_record(const EfmDims *d, const EfmShots *sh, const EfmState *st, float *gathers,
                          int it, void *stream) {
    if (!d || !sh || !gathers) return fail(EFM_E_ARG, "efm_record: null argument");
    int rc = check_dims(d);
    if (rc) return rc;
    if ((rc = validate_shots(d, sh)) || (rc = validate_state(d, st))) return rc;
    if (it < 0 || it >= d->ntmax) return fail(EFM_E_ARG, "efm_record: it out of range");
    return launch_record(make_geo(*d), make_shots(*sh), make_state(*st), gathers, it,
                         (hipStream_t)stream);
}

extern "C" int efm_forward(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                           const EfmState *st, float *gathers, float *borders, int W, int it0,
                           int nsteps, void *stream) {
    int rc = validate(d, m, sh);
    if (rc || (rc = validate_state(d, st))) return rc;
    if (it0 < 0 || nsteps < 0 || it0 + nsteps > d->ntmax)
        return fail(EFM_E_ARG, "efm_forward: step range outside [0, ntmax)");
    if (borders && (W < 0 || W + H + 1 > d->nx / 2 || W + H + 1 > d->ny / 2 || W + H + 1 > d->nz / 2))
        return fail(EFM_E_ARG, "efm_forward: CPML too wide for the border save");
    const Geo g = make_geo(*d);
    const ShotPtrs shp = make_shots(*sh);
    const StatePtrs s = make_state(*st);
    hipStream_t hs = (hipStream_t)stream;
    for (int it = it0; it < it0 + nsteps; ++it) {
        if ((rc = launch_stress_src(g, *m, shp, s, s, it, 1, hs))) return rc;
        if ((rc = launch_velocity(g, *m, shp, s, s, it, hs, gathers))) return rc;
        if (gathers && !shp.recmask && (rc = launch_record(g, shp, s, gathers, it, hs))) return rc;
        if (borders && (rc = launch_border(g, shp, s, borders, W, it, 1, it, hs))) return rc;
    }
    return EFM_OK;
}

namespace efm {
__global__ void k_counter_set(int *c, int v) {
    if (threadIdx.x == 0) *c = v;
}
__global__ void k_counter_add(int *c, int v) {
    if (threadIdx.x == 0) *c += v;
}
}  // namespace efm

extern "C" int efm_forward_graph(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                                 const EfmState *st, float *gathers, int it0, int nsteps, int chunk,
                                 int *counter, void *stream) {
    int rc = validate(d, m, sh);
    if (rc || (rc = validate_state(d, st))) return rc;
    if (it0 < 0 || nsteps < 0 || it0 + nsteps > d->ntmax)
        return fail(EFM_E_ARG, "efm_forward_graph: step range outside [0, ntmax)");
    if (chunk < 1 || !counter) return fail(EFM_E_ARG, "efm_forward_graph: chunk >= 1 and a counter are required");
    hipStream_t hs = (hipStream_t)stream;
    if (!hs) return fail(EFM_E_ARG, "efm_forward_graph: graph capture needs a non-default stream");
    const ShotPtrs shp = make_shots(*sh);
    const StatePtrs s = make_state(*st);
    const int nchunks = nsteps / chunk;
    if (nchunks > 0) {
        hipLaunchKernelGGL(k_counter_set, dim3(1), dim3(64), 0, hs, counter, it0);
        if ((rc = check_launch("counter set"))) return rc;
        Geo g = make_geo(*d);
        g.itp = counter;
        if (hipStreamBeginCapture(hs, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return fail(EFM_E_HIP, "efm_forward_graph: hipStreamBeginCapture failed");
        int crc = EFM_OK;
        for (int k = 0; k < chunk && crc == EFM_OK; ++k) {
            if ((crc = launch_stress_src(g, *m, shp, s, s, k, 1, hs))) break;
            if ((crc = launch_velocity(g, *m, shp, s, s, k, hs, gathers))) break;
            if (gathers && !shp.recmask) crc = launch_record(g, shp, s, gathers, k, hs);
        }
        if (crc == EFM_OK) {
            hipLaunchKernelGGL(k_counter_add, dim3(1), dim3(64), 0, hs, counter, chunk);
            crc = check_launch("counter add");
        }
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(hs, &graph);
        if (crc) { if (graph) (void)hipGraphDestroy(graph); return crc; }
        if (ec != hipSuccess || !graph) return fail(EFM_E_HIP, "efm_forward_graph: hipStreamEndCapture failed");
        hipGraphExec_t exec = nullptr;
        if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return fail(EFM_E_HIP, "efm_forward_graph: hipGraphInstantiate failed");
        }
        hipError_t el = hipSuccess;
        for (int c = 0; c < nchunks && el == hipSuccess; ++c) el = hipGraphLaunch(exec, hs);
        const hipError_t es = hipStreamSynchronize(hs);   // the exec is released below
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        if (el != hipSuccess || es != hipSuccess) return fail(EFM_E_HIP, "efm_forward_graph: graph launch failed");
    }
    const Geo g = make_geo(*d);
    for (int it = it0 + nchunks * chunk; it < it0 + nsteps; ++it) {
        if ((rc = launch_stress_src(g, *m, shp, s, s, it, 1, hs))) return rc;
        if ((rc = launch_velocity(g, *m, shp, s, s, it, hs, gathers))) return rc;
        if (gathers && !shp.recmask && (rc = launch_record(g, shp, s, gathers, it, hs))) return rc;
    }
    return EFM_OK;
}

extern "C" int efm_forward_states(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                                  const EfmState *const *states, float *gathers, int it0,
                                  int nsteps, void *stream) {
    int rc = validate(d, m, sh);
    if (rc) return rc;
    if (!states) return fail(EFM_E_ARG, "efm_forward_states: null state list");
    if (it0 < 0 || nsteps < 0 || it0 + nsteps > d->ntmax)
        return fail(EFM_E_ARG, "efm_forward_states: step range outside [0, ntmax)");
    for (int k = 0; k <= nsteps; ++k)
        if ((rc = validate_state(d, states[k]))) return rc;
    const Geo g = make_geo(*d);
    const ShotPtrs shp = make_shots(*sh);
    hipStream_t hs = (hipStream_t)stream;
    for (int k = 0; k < nsteps; ++k) {
        const StatePtrs in = make_state(*states[k]), out = make_state(*states[k + 1]);
        if ((rc = launch_stress_src(g, *m, shp, in, out, it0 + k, 1, hs))) return rc;
        if ((rc = launch_velocity(g, *m, shp, in, out, it0 + k, hs, gathers))) return rc;
        if (gathers && !shp.recmask && (rc = launch_record(g, shp, out, gathers, it0 + k, hs)))
            return rc;
    }
    return EFM_OK;
}
