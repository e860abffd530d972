// Internal header shared by the gfx950 kernels of libefm.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "efm.h"

namespace efm {

constexpr int H = 2;  // interior guard: 2 <= x,y,z < n-2 (reference M:54, 77, 152)

// Everything a sweep kernel needs, passed by value (kernel-argument memory).
struct Geo {
    int nx, ny, nz, S;
    int wxl, wxh, wyl, wyh, wzl, wzh;  // CPML widths per face
    int Lx, Ly, Lz;                    // slab extents wxl+wxh, ...
    int ntmax, nrec;
    float c;                           // signed dt/dx
    long long N;                       // nx*ny*nz
    long long sxz;                     // nx*nz (y stride)
    long long ps;                      // padded component stride N + EFM_PAD (acc, mbar, bbar)
    int sxz32;                         // nx*nz as int (per-shot offsets are 32-bit)
    int pxc, pyc, pzc;                 // per-component strides of the ψ slabs
    const int *itp;                    // device time-step counter added to `it` (graph replay), or null
    int recw;                          // words per shot of the receiver bitmask, ceil(N/32)
};

struct StatePtrs {
    float *v[3];    // vx vy vz
    float *t[6];    // txx tyy tzz txy txz tyz
    float *px, *py, *pz;
};

struct ShotPtrs {
    const float *ax, *bx, *ay, *by, *az, *bz;
    const int *src;
    const float *m6;
    const float *wav;
    const int *nt;
    const int *rec;
    const unsigned *recmask;   // (S, recw) receiver bitmask, or null
};

inline Geo make_geo(const EfmDims &d) {
    Geo g;
    g.nx = d.nx; g.ny = d.ny; g.nz = d.nz; g.S = d.nshot;
    g.wxl = d.wx_lo; g.wxh = d.wx_hi; g.wyl = d.wy_lo; g.wyh = d.wy_hi;
    g.wzl = d.wz_lo; g.wzh = d.wz_hi;
    g.Lx = d.wx_lo + d.wx_hi; g.Ly = d.wy_lo + d.wy_hi; g.Lz = d.wz_lo + d.wz_hi;
    g.ntmax = d.ntmax; g.nrec = d.nrec; g.c = d.dtdx;
    g.sxz = (long long)d.nx * d.nz;
    g.N = g.sxz * d.ny;
    g.ps = g.N + EFM_PAD;
    g.sxz32 = d.nx * d.nz;
    g.pxc = d.ny * d.nz * g.Lx;
    g.pyc = g.Ly * d.nz * d.nx;
    g.pzc = d.ny * g.Lz * d.nx;
    g.itp = nullptr;
    g.recw = (int)((g.N + 31) / 32);
    return g;
}

inline StatePtrs make_state(const EfmState &s) {
    StatePtrs p;
    p.v[0] = s.vx; p.v[1] = s.vy; p.v[2] = s.vz;
    p.t[0] = s.txx; p.t[1] = s.tyy; p.t[2] = s.tzz;
    p.t[3] = s.txy; p.t[4] = s.txz; p.t[5] = s.tyz;
    p.px = s.psi_x; p.py = s.psi_y; p.pz = s.psi_z;
    return p;
}

inline ShotPtrs make_shots(const EfmShots &s) {
    ShotPtrs p;
    p.ax = s.ax; p.bx = s.bx; p.ay = s.ay; p.by = s.by; p.az = s.az; p.bz = s.bz;
    p.src = s.src; p.m6 = s.m6; p.wav = s.wavelet; p.nt = s.nt; p.rec = s.rec;
    p.recmask = s.recmask;
    return p;
}

// Per-shot ψ slab bases: comp c of shot s starts at base + c*pxc (int offsets below).
__device__ __forceinline__ long long px_shot(const Geo &g, int s) { return (long long)s * 6 * g.pxc; }
__device__ __forceinline__ long long py_shot(const Geo &g, int s) { return (long long)s * 6 * g.pyc; }
__device__ __forceinline__ long long pz_shot(const Geo &g, int s) { return (long long)s * 6 * g.pzc; }
__device__ __forceinline__ int px_off(const Geo &g, int y, int z, int sx) { return (y * g.nz + z) * g.Lx + sx; }
__device__ __forceinline__ int py_off(const Geo &g, int sy, int z, int x) { return (sy * g.nz + z) * g.nx + x; }
__device__ __forceinline__ int pz_off(const Geo &g, int y, int sz, int x) { return (y * g.Lz + sz) * g.nx + x; }

// ---- slab addressing (ψ lives only in the CPML slabs) ----------------------
// Returns the slab coordinate along an axis of length n with widths lo/hi, or -1.
__device__ __forceinline__ int slab_coord(int i, int n, int lo, int hi) {
    if (i < lo) return i;
    if (i >= n - hi) return lo + (i - (n - hi));
    return -1;
}
// psi_x (S, 6, ny, nz, Lx)
__device__ __forceinline__ long long px_index(const Geo &g, int s, int comp, int y, int z, int sx) {
    return ((((long long)s * 6 + comp) * g.ny + y) * g.nz + z) * g.Lx + sx;
}
__device__ __forceinline__ long long px_cstride(const Geo &g) { return (long long)g.ny * g.nz * g.Lx; }
// psi_y (S, 6, Ly, nz, nx)
__device__ __forceinline__ long long py_index(const Geo &g, int s, int comp, int sy, int z, int x) {
    return ((((long long)s * 6 + comp) * g.Ly + sy) * g.nz + z) * g.nx + x;
}
__device__ __forceinline__ long long py_cstride(const Geo &g) { return (long long)g.Ly * g.nz * g.nx; }
// psi_z (S, 6, ny, Lz, nx)
__device__ __forceinline__ long long pz_index(const Geo &g, int s, int comp, int y, int sz, int x) {
    return ((((long long)s * 6 + comp) * g.ny + y) * g.Lz + sz) * g.nx + x;
}
__device__ __forceinline__ long long pz_cstride(const Geo &g) { return (long long)g.ny * g.Lz * g.nx; }

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Whether cell t (per-shot linear index) of shot s holds a receiver.
__device__ __forceinline__ bool is_receiver(const Geo &g, const ShotPtrs &sh, int s, int t) {
    return (sh.recmask[(long long)s * g.recw + (t >> 5)] >> (t & 31)) & 1u;
}

__device__ __forceinline__ bool interior(int x, int y, int z, const Geo &g) {
    return x >= H && x < g.nx - H && y >= H && y < g.ny - H && z >= H && z < g.nz - H;
}

// ---- error plumbing ---------------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int check_launch(const char *what);
// Grid, CPML widths, record length and the size limits of the sweep kernels
// (32-bit offsets inside a shot, one dispatch of < 2^32 work-items per sweep).
int check_dims(const EfmDims *d);
// Null checks of the C-ABI structs (EFM_E_ARG with a message): the shot
// description's arrays, dims + medium + shots, and one state (ψ slab pointers
// are required for every axis that has a CPML layer).
int validate_shots(const EfmDims *d, const EfmShots *sh);
int validate(const EfmDims *d, const EfmMedium *m, const EfmShots *sh);
int validate_state(const EfmDims *d, const EfmState *s);

// ---- tile configuration chosen by the autotuner ------------------------------
struct Tile {
    int bx, bz, ny;  // threads in x, threads in z, cells per thread along y
    int map;         // block mapping: 0 = linear 3-D grid; 8 = XCD-aware over 8 XCDs;
                     // 16 = XCD-aware over the compact interior plane (see tile_coords)
};

// Compiled tile variants (bx, bz, ny, map).  Whole 64-lane waves along x.
#define EFM_TILE_LIST(X) \
    X(64, 1, 1, 0)       \
    X(64, 2, 1, 0)       \
    X(64, 4, 1, 0)       \
    X(128, 1, 1, 0)      \
    X(128, 2, 1, 0)      \
    X(64, 4, 2, 0)       \
    X(64, 1, 1, 8)       \
    X(64, 2, 1, 8)       \
    X(64, 4, 1, 8)       \
    X(64, 8, 1, 8)       \
    X(128, 1, 1, 8)      \
    X(128, 2, 1, 8)      \
    X(256, 1, 1, 8)      \
    X(64, 4, 2, 8)       \
    X(64, 1, 1, 16)      \
    X(128, 1, 1, 16)     \
    X(256, 1, 1, 16)     \
    X(128, 1, 2, 16)

// XCDs the blocks of an XCD-aware mapping are dealt over.
__host__ __device__ constexpr int map_xcds(int map) { return map == 16 ? 8 : map; }

// Number of tiles (blocks × y-chunks × shots) of a sweep: x-blocks × z-blocks,
// or, for the compact mapping, blocks of bx·bz consecutive interior (x, z) cells.
inline long long sweep_tiles(const Geo &g, int bx, int bz, int ny, int map) {
    const long long nyc = (g.ny - 2 * H + ny - 1) / ny;
    if (map == 16) {
        const long long q = (long long)(g.nx - 2 * H) * (g.nz - 2 * H);
        return (q + bx * bz - 1) / (bx * bz) * nyc * g.S;
    }
    const long long nbx = (g.nx + bx - 1) / bx, nbz = (g.nz + bz - 1) / bz;
    return nbx * nbz * nyc * g.S;
}

// Blocks of an XCD-aware sweep: each shot's tiles are dealt over the k XCDs in
// runs of ceil(tiles/k) (see tile_coords), so a shot takes k·ceil(tiles/k).
inline long long xcd_blocks(const Geo &g, const Tile &t) {
    const long long per_shot = sweep_tiles(g, t.bx, t.bz, t.ny, t.map) / g.S;
    const int k = map_xcds(t.map);
    return k * ((per_shot + k - 1) / k) * g.S;
}

inline dim3 sweep_grid(const Geo &g, const Tile &t) {
    if (t.map > 0) return dim3((unsigned)xcd_blocks(g, t));
    const int nyc = (g.ny - 2 * H + t.ny - 1) / t.ny;
    return dim3((g.nx + t.bx - 1) / t.bx, (g.nz + t.bz - 1) / t.bz, g.S * nyc);
}

// Tile coordinates of the calling block.  MAP = 0: the 3-D grid as launched.
// MAP = k: workgroups are dealt to the k XCDs round-robin (w mod k), each with
// its own L2, so block w is given a tile of its XCD's contiguous run: every XCD
// walks a run of a shot's tiles (x fastest, then z, then y), and the ±z rows and
// ±y planes a tile reads were fetched into the same L2 by its neighbours.  The
// runs are per shot (XCD k takes the k-th run of shot 0, then of shot 1, ...), so
// all XCDs work on the same shot at a time and a batch's sweep keeps one shot's
// working set in the Infinity Cache, as a one-shot sweep does.
// MAP = 16: the same XCD order over blocks of BX·BZ consecutive cells of the
// interior (x, z) plane numbered x fastest with rows wrapped, so that every
// 64-lane wave is full whatever nx is (with nx - 4 not a multiple of 64 the
// other mappings leave part-empty waves, which cost the HBM-bound sweeps their
// share of bytes in flight); a wave then reads at most two row segments.
// Lanes past the last interior cell get x = -1 (the kernels return on it).
// Returns false for the padding blocks past the last tile.
template <int BX, int BZ, int NYC, int MAP>
__device__ __forceinline__ bool tile_coords(const Geo &g, int &x, int &z, int &s, int &yb) {
    const int nyc = (g.ny - 2 * H + NYC - 1) / NYC;
    if (MAP == 16) {
        const int nxi = g.nx - 2 * H;
        const int Q = nxi * (g.nz - 2 * H);
        const int nqb = (Q + BX * BZ - 1) / (BX * BZ);
        const int per_shot = nqb * nyc;
        const int run = (per_shot + 7) / 8;
        const int w = blockIdx.x;
        const int j = w / 8;
        s = j / run;
        const int gi = (w % 8) * run + (j - s * run);
        if (s >= g.S || gi >= per_shot) return false;
        const int qb = gi % nqb;
        yb = gi / nqb;
        const int q = qb * (BX * BZ) + threadIdx.y * BX + threadIdx.x;
        if (q >= Q) {
            x = z = -1;
            return true;
        }
        const int zi = q / nxi;
        z = H + zi;
        x = H + (q - zi * nxi);
        return true;
    }
    if (MAP == 0) {
        x = blockIdx.x * BX + threadIdx.x;
        z = blockIdx.y * BZ + threadIdx.y;
        s = blockIdx.z / nyc;
        yb = blockIdx.z - s * nyc;
        return true;
    }
    const int nbx = (g.nx + BX - 1) / BX, nbz = (g.nz + BZ - 1) / BZ;
    constexpr int K = MAP > 0 ? MAP : 1;
    const int per_shot = nbx * nbz * nyc;
    const int run = (per_shot + K - 1) / K;
    const int w = blockIdx.x;
    const int j = w / K;
    s = j / run;
    const int gi = (w % K) * run + (j - s * run);
    if (s >= g.S || gi >= per_shot) return false;
    const int xb = gi % nbx;
    const int r = gi / nbx;
    const int zb = r % nbz;
    yb = r / nbz;
    x = xb * BX + threadIdx.x;
    z = zb * BZ + threadIdx.y;
    return true;
}

// Launch KER<bx,bz,ny,map> for the selected tile; sets `launched`.
#define EFM_LAUNCH_IF(BX_, BZ_, NY_, MAP_, KER, ...)                                           \
    if (!launched && t.bx == BX_ && t.bz == BZ_ && t.ny == NY_ && t.map == MAP_) {              \
        hipLaunchKernelGGL((KER<BX_, BZ_, NY_, MAP_>), sweep_grid(g, t), dim3(BX_, BZ_), 0, st, \
                           __VA_ARGS__);                                                        \
        launched = true;                                                                       \
    }
#define EFM_KERNEL_ATTRS(BX, BZ, MAP) __launch_bounds__(BX *BZ)

Tile get_tile(int kernel);  // 0 stress, 1 velocity, 2 adjoint-velocity, 3 adjoint-stress
int set_tile(int kernel, Tile t);

// ---- launchers (defined in the .hip files) ------------------------------------
int launch_stress_src(const Geo &g, const EfmMedium &m, const ShotPtrs &sh, const StatePtrs &in,
                      const StatePtrs &out, int it, int src_mode, hipStream_t st);
int launch_stress(const Geo &g, const EfmMedium &m, const ShotPtrs &sh, const StatePtrs &in,
                  const StatePtrs &out, int it, hipStream_t st);
int launch_velocity(const Geo &g, const EfmMedium &m, const ShotPtrs &sh, const StatePtrs &in,
                    const StatePtrs &out, int it, hipStream_t st, float *gathers = nullptr);
int launch_record(const Geo &g, const ShotPtrs &sh, const StatePtrs &s, float *gathers, int it,
                  hipStream_t st);
int launch_border(const Geo &g, const ShotPtrs &sh, const StatePtrs &s, float *borders, int W,
                  int slot, int save, int it_gate, hipStream_t st);
long long border_floats(const Geo &g, int W);

}  // namespace efm
