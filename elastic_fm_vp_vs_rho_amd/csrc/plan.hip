// Persistent plans: HIP graphs of the forward and adjoint time loops, captured
// once per set of buffers and replayed (no reference counterpart: the
// reference issues 9-29 launches per time step from Python, M:1812-2082).
//
// A plan belongs to one shot batch (dims, medium, shots).  Each entry point
// looks its arguments up among the graphs the plan already holds (the buffers,
// the step range and the tiles installed at capture are the key); on a miss it
// records the eager launch sequence on the plan's own capture stream (nothing
// runs during capture), instantiates it and keeps it.  Replays go to the
// caller's stream with no host synchronisation.  The in-place forward graph
// covers `chunk` steps and reads its step index from a device counter the
// plan owns, so one graph serves any step range; the adjoint and
// segment-forward graphs bake in their step range, which the gradient's
// checkpoint segments repeat from one gradient to the next.
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace efm {

__global__ void k_plan_counter(int *c, int v) {
    if (threadIdx.x == 0) *c = v;
}
__global__ void k_plan_counter_add(int *c, int v) {
    if (threadIdx.x == 0) *c += v;
}

struct Graph {
    std::vector<unsigned char> key;
    hipGraphExec_t exec = nullptr;
};

static void key_add(std::vector<unsigned char> &k, const void *p, size_t n) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    k.insert(k.end(), b, b + n);
}
template <typename T>
static void key_add(std::vector<unsigned char> &k, const T &v) {
    key_add(k, &v, sizeof(T));
}

}  // namespace efm

struct EfmPlan {
    EfmDims d;
    EfmMedium m;
    EfmShots sh;
    int chunk;
    int *counter = nullptr;          // device step counter of the forward graph
    hipStream_t cap = nullptr;       // capture stream (records, never runs work)
    std::mutex mu;                   // one call at a time per plan
    std::vector<efm::Graph> graphs;
    efm::Tile tiles[4];
};

using namespace efm;

static void plan_tiles(Tile *t) {
    for (int k = 0; k < 4; ++k) t[k] = get_tile(k);
}

static bool tiles_changed(EfmPlan *p) {
    Tile now[4];
    plan_tiles(now);
    for (int k = 0; k < 4; ++k)
        if (now[k].bx != p->tiles[k].bx || now[k].bz != p->tiles[k].bz || now[k].ny != p->tiles[k].ny ||
            now[k].map != p->tiles[k].map)
            return true;
    return false;
}

static void drop_graphs(EfmPlan *p) {
    for (Graph &g : p->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    p->graphs.clear();
}

// Replay the graph for `key`, capturing it with `record` (which launches on the
// stream it is given) when the plan does not hold it yet.
template <typename F>
static int replay(EfmPlan *p, const std::vector<unsigned char> &key, hipStream_t st, F record) {
    if (tiles_changed(p)) {            // graphs hold the kernels of the tiles they were captured with
        drop_graphs(p);
        plan_tiles(p->tiles);
    }
    for (Graph &g : p->graphs)
        if (g.key == key) {
            if (hipGraphLaunch(g.exec, st) != hipSuccess) return fail(EFM_E_HIP, "efm_plan: graph launch failed");
            return EFM_OK;
        }
    if (hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return fail(EFM_E_HIP, "efm_plan: hipStreamBeginCapture failed");
    const int rc = record(p->cap);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(p->cap, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess || !graph) return fail(EFM_E_HIP, "efm_plan: hipStreamEndCapture failed");
    Graph g;
    g.key = key;
    const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(EFM_E_HIP, "efm_plan: hipGraphInstantiate failed");
    p->graphs.push_back(std::move(g));
    if (hipGraphLaunch(p->graphs.back().exec, st) != hipSuccess)
        return fail(EFM_E_HIP, "efm_plan: graph launch failed");
    return EFM_OK;
}

extern "C" int efm_plan_create(const EfmDims *d, const EfmMedium *m, const EfmShots *sh, int chunk,
                               EfmPlan **out) {
    if (!out) return fail(EFM_E_ARG, "efm_plan_create: null output");
    *out = nullptr;
    if (int rc = validate(d, m, sh)) return rc;
    if (chunk < 1) return fail(EFM_E_ARG, "efm_plan_create: chunk must be >= 1");
    std::unique_ptr<EfmPlan> p(new EfmPlan());
    p->d = *d;
    p->m = *m;
    p->sh = *sh;
    p->chunk = chunk;
    plan_tiles(p->tiles);
    if (hipMalloc(&p->counter, sizeof(int)) != hipSuccess)
        return fail(EFM_E_HIP, "efm_plan_create: counter allocation failed");
    if (hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking) != hipSuccess) {
        (void)hipFree(p->counter);
        return fail(EFM_E_HIP, "efm_plan_create: stream creation failed");
    }
    *out = p.release();
    return EFM_OK;
}

extern "C" int efm_plan_destroy(EfmPlan *p) {
    if (!p) return EFM_OK;
    {
        std::lock_guard<std::mutex> lk(p->mu);
        drop_graphs(p);
    }
    if (p->cap) (void)hipStreamDestroy(p->cap);
    if (p->counter) (void)hipFree(p->counter);
    delete p;
    return EFM_OK;
}

extern "C" int efm_plan_graphs(const EfmPlan *p) { return p ? (int)p->graphs.size() : -1; }

extern "C" int efm_plan_forward(EfmPlan *p, const EfmState *state, float *gathers, int it0, int nsteps,
                                void *stream) {
    if (!p) return fail(EFM_E_ARG, "efm_plan_forward: null plan");
    if (int rc = validate_state(&p->d, state)) return rc;
    if (it0 < 0 || nsteps < 0 || it0 + nsteps > p->d.ntmax)
        return fail(EFM_E_ARG, "efm_plan_forward: step range outside [0, ntmax)");
    std::lock_guard<std::mutex> lk(p->mu);
    hipStream_t st = (hipStream_t)stream;
    const ShotPtrs shp = make_shots(p->sh);
    const StatePtrs s = make_state(*state);
    const int K = p->chunk, nchunks = nsteps / K;
    if (nchunks > 0) {
        hipLaunchKernelGGL(k_plan_counter, dim3(1), dim3(64), 0, st, p->counter, it0);
        if (int rc = check_launch("plan counter")) return rc;
        std::vector<unsigned char> key;
        key_add(key, 'F');
        key_add(key, *state);
        key_add(key, gathers);
        for (int c = 0; c < nchunks; ++c) {
            const int rc = replay(p, key, st, [&](hipStream_t cs) {
                Geo g = make_geo(p->d);
                g.itp = p->counter;
                for (int k = 0; k < K; ++k) {
                    if (int r = launch_stress_src(g, p->m, shp, s, s, k, 1, cs)) return r;
                    if (int r = launch_velocity(g, p->m, shp, s, s, k, cs, gathers)) return r;
                    if (gathers && !shp.recmask)
                        if (int r = launch_record(g, shp, s, gathers, k, cs)) return r;
                }
                hipLaunchKernelGGL(k_plan_counter_add, dim3(1), dim3(64), 0, cs, p->counter, K);
                return check_launch("plan counter add");
            });
            if (rc) return rc;
        }
    }
    const Geo g = make_geo(p->d);
    for (int it = it0 + nchunks * K; it < it0 + nsteps; ++it) {
        if (int rc = launch_stress_src(g, p->m, shp, s, s, it, 1, st)) return rc;
        if (int rc = launch_velocity(g, p->m, shp, s, s, it, st, gathers)) return rc;
        if (gathers && !shp.recmask)
            if (int rc = launch_record(g, shp, s, gathers, it, st)) return rc;
    }
    return EFM_OK;
}

extern "C" int efm_plan_forward_states(EfmPlan *p, const EfmState *const *states, float *gathers,
                                       int it0, int nsteps, void *stream) {
    if (!p || !states) return fail(EFM_E_ARG, "efm_plan_forward_states: null argument");
    if (it0 < 0 || nsteps < 0 || it0 + nsteps > p->d.ntmax)
        return fail(EFM_E_ARG, "efm_plan_forward_states: step range outside [0, ntmax)");
    for (int k = 0; k <= nsteps; ++k)
        if (int rc = validate_state(&p->d, states[k])) return rc;
    if (nsteps == 0) return EFM_OK;
    std::lock_guard<std::mutex> lk(p->mu);
    std::vector<unsigned char> key;
    key_add(key, 'S');
    key_add(key, gathers);
    key_add(key, it0);
    key_add(key, nsteps);
    for (int k = 0; k <= nsteps; ++k) key_add(key, *states[k]);
    return replay(p, key, (hipStream_t)stream, [&](hipStream_t cs) {
        const Geo g = make_geo(p->d);
        const ShotPtrs shp = make_shots(p->sh);
        for (int k = 0; k < nsteps; ++k) {
            const StatePtrs in = make_state(*states[k]), out = make_state(*states[k + 1]);
            if (int r = launch_stress_src(g, p->m, shp, in, out, it0 + k, 1, cs)) return r;
            if (int r = launch_velocity(g, p->m, shp, in, out, it0 + k, cs, gathers)) return r;
            if (gathers && !shp.recmask)
                if (int r = launch_record(g, shp, out, gathers, it0 + k, cs)) return r;
        }
        return EFM_OK;
    });
}

extern "C" int efm_plan_adjoint(EfmPlan *p, const float *mbar, const float *bbar,
                                const EfmState *const *fwd, const EfmState *adj, const EfmState *adj2,
                                const float *resid, float *acc, double *gm6, int n0, int n1, int last,
                                void *stream) {
    if (!p || !mbar || !bbar || !fwd || !adj || !adj2 || (!resid && p->d.nrec > 0) || !acc || !gm6)
        return fail(EFM_E_ARG, "efm_plan_adjoint: null argument");
    if (n0 < 0 || n1 <= n0 || n1 > p->d.ntmax) return fail(EFM_E_ARG, "efm_plan_adjoint: bad step range");
    for (int k = 0; k <= n1 - n0; ++k)
        if (int rc = validate_state(&p->d, fwd[k])) return rc;
    if (int rc = validate_state(&p->d, adj)) return rc;
    std::lock_guard<std::mutex> lk(p->mu);
    std::vector<unsigned char> key;
    key_add(key, 'A');
    key_add(key, mbar); key_add(key, bbar); key_add(key, resid); key_add(key, acc); key_add(key, gm6);
    key_add(key, *adj); key_add(key, *adj2);
    key_add(key, n0); key_add(key, n1); key_add(key, last);
    for (int k = 0; k <= n1 - n0; ++k) key_add(key, *fwd[k]);
    return replay(p, key, (hipStream_t)stream, [&](hipStream_t cs) {
        return efm_adjoint(&p->d, &p->m, mbar, bbar, &p->sh, fwd, adj, adj2, resid, acc, gm6, n0, n1,
                           last, cs);
    });
}
