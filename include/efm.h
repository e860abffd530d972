/*
 * efm.h — C ABI of libefm.so, the gfx950 (MI355X) elastic FDTD propagator.
 *
 * The library replaces the device side of mainPropElastic3D_MT.py (the PyCUDA
 * SourceModule kernels, M:42-530, and the per-step launch sequences of
 * Elastic3D, M:1491-2082).  Every entry point is asynchronous on the HIP
 * stream it is given (efm_forward_graph returns after its work is done),
 * allocates no device buffers (efm_forward_graph builds, replays and releases
 * one HIP graph per call), and returns 0 or a negative EFM_E_* code;
 * efm_last_error() returns the message of the last failure on the calling
 * thread.
 *
 * Memory conventions (all pointers are device pointers, float32 unless noted):
 *   volume           (ny, nz, nx), x fastest — the reference layout (M:72, 758)
 *   batched volume   (S, ny, nz, nx), shot-major
 *   ψ slabs          psi_x (S, 6, ny, nz, Lx), psi_y (S, 6, Ly, nz, nx),
 *                    psi_z (S, 6, ny, Lz, nx) with L = lo + hi CPML widths of
 *                    the axis.  Components: x: vx_x vy_x vz_x txx_x txy_x txz_x;
 *                    y: vx_y vy_y vz_y txy_y tyy_y tyz_y; z: vx_z vy_z vz_z
 *                    txz_z tyz_z tzz_z.  (The reference keeps 18 full volumes,
 *                    M:811-862; ψ is identically zero outside the CPML.)
 *   gathers          (S, 3, ntmax, nrec): vx, vy, vz, time-major per shot like
 *                    the reference Gather (M:1132-1140)
 *   receivers        (S, nrec) int32 linear cell index x + nx*z + nx*nz*y,
 *                    -1 = unused slot
 *   sources          (S, 3) int32 (x, y, z); must be interior (2 <= . < n-2)
 *   moment tensors   (S, 6) float32 (xx, yy, zz, xy, xz, yz) (M:1374-1389)
 *
 * Size limits (checked by every entry point that takes EfmDims, EFM_E_ARG):
 * one shot's grid below 2^31 cells and each CPML slab array below 2^31 values
 * per shot (the kernels use 32-bit offsets within a shot); each sweep runs all
 * shots of the batch in one dispatch of fewer than 2^32 work-items (about
 * 4e9 cells over all shots with the default tile), so larger batches are
 * refused and must be split.
 */
#ifndef EFM_H
#define EFM_H

#ifdef __cplusplus
extern "C" {
#endif

/* Component stride padding (floats) of the multi-component scratch arrays
 * (adjoint accumulators, derived medium): component k of such an array starts
 * at k*(N + EFM_PAD), N = nx*ny*nz, so that equal-index elements of different
 * components do not fall on power-of-two address strides. */
#define EFM_PAD 1088

#define EFM_OK 0
#define EFM_E_ARG (-1)
#define EFM_E_HIP (-2)
#define EFM_E_UNSUPPORTED (-3)

typedef struct EfmDims {
    int nx, ny, nz;   /* grid */
    int nshot;        /* S */
    int wx_lo, wx_hi, wy_lo, wy_hi, wz_lo, wz_hi;  /* CPML widths per face (ψ slab extents) */
    int ntmax;        /* time samples per shot buffer (gathers, wavelets) */
    int nrec;         /* receiver slots per shot */
    float dtdx;       /* signed dt/dx; the reference's flagMode sign (M:1502) */
} EfmDims;

typedef struct EfmMedium {     /* (ny, nz, nx), shared by all shots */
    const float *lam, *mu, *rho;
} EfmMedium;

typedef struct EfmState {      /* batched wavefield (S, ny, nz, nx) + ψ slabs */
    float *vx, *vy, *vz, *txx, *tyy, *tzz, *txy, *txz, *tyz;
    float *psi_x, *psi_y, *psi_z;
} EfmState;

typedef struct EfmShots {      /* per-shot description */
    const float *ax, *bx;      /* (S, nx) CPML profiles (CPML.CPML_i, M:1340-1372) */
    const float *ay, *by;      /* (S, ny) */
    const float *az, *bz;      /* (S, nz) */
    const int   *src;          /* (S, 3) */
    const float *m6;           /* (S, 6) */
    const float *wavelet;      /* (S, ntmax) */
    const int   *nt;           /* (S) record length of each shot */
    const int   *rec;          /* (S, nrec) */
    /* (S, ceil(N/32)) bit set for every cell that holds a receiver of the shot
     * (bit c%32 of word c/32), or null.  When given, the velocity sweep samples
     * the receivers itself and the adjoint stress sweep injects the residual
     * and sums ∂J/∂M, instead of separate small kernels each step. */
    const unsigned *recmask;
} EfmShots;

/* ---- library ---------------------------------------------------------- */
int efm_version(void);
/* Copies the last error message of the calling thread into buf (NUL-terminated). */
int efm_last_error(char *buf, int len);

/* ---- tile selection (no reference counterpart: the reference hard-codes
 * block=(32,2,4), M:2764-2766).  kernel: 0 stress, 1 velocity, 2 adjoint
 * velocity, 3 adjoint stress.  A tile is bx × bz threads, ny cells per thread
 * along y; map 0 launches a linear 3-D grid, map 8 deals contiguous runs of
 * tiles to each of the 8 XCDs (workgroups are dispatched to XCDs round-robin),
 * map 16 does the same over blocks of bx·bz consecutive interior (x, z) cells
 * with rows wrapped (full 64-lane waves whatever nx is).  efm_tile_variants
 * writes the compiled tiles as quadruples and returns their count.  The
 * autotuner (autotune.py) drives these. */
int efm_set_tile(int kernel, int bx, int bz, int ny, int map);
int efm_get_tile(int kernel, int *bx, int *bz, int *ny, int *map);
int efm_tile_variants(int *out, int max);
/* EFM_PAD as compiled into the library. */
int efm_component_pad(void);

/* ---- model maps -------------------------------------------------------- */
/* Model.compute_Lambda_Mu (M:1196-1198): mu = vs*vs*rho, lam = vp*vp*rho - 2*mu,
 * float32 with no contraction (bit-identical to the reference's numpy). */
int efm_model_lame(const float *vp, const float *vs, const float *rho,
                   float *lam, float *mu, long long n, void *stream);
/* Chain rule back to velocities: gvp = 2ρVp·gλ, gvs = 2ρVs·(gμ - 2gλ),
 * grho_out = (Vp²-2Vs²)·gλ + Vs²·gμ + grho_in. */
int efm_model_lame_adjoint(const float *vp, const float *vs, const float *rho,
                           const float *glam, const float *gmu, const float *grho_in,
                           float *gvp, float *gvs, float *grho_out, long long n, void *stream);

/* computeResidual / computeCostFunction (M:2128-2142): res = mod - obs
 * elementwise and J += ½Σres² (J is a device double the caller zeroes). */
int efm_residual(const float *mod, const float *obs, float *res, double *J, long long n,
                 void *stream);

/* ---- forward sweeps ---------------------------------------------------- */
/* kernel_Tau (M:59-128) + addSource (M:1728-1734): τ/ψ_v of `out` from v, τ,
 * ψ_v of `in` (in may alias out); the source term M·w[it] is added to all six
 * τ at each shot's source cell.  Shots with it >= nt[s] are left untouched. */
int efm_stress_step(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                    const EfmState *in, const EfmState *out, int it, void *stream);
/* kernel_Vx_Vy_Vz (M:133-188): v/ψ_τ of `out` from τ of `out` and v, ψ_τ of `in`. */
int efm_velocity_step(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                      const EfmState *in, const EfmState *out, int it, void *stream);
/* kernel_get_Station (M:324-335): gathers[s, c, it, r] = v_c[rec[s, r]]. */
int efm_record(const EfmDims *d, const EfmShots *sh, const EfmState *st,
               float *gathers, int it, void *stream);
/* forwardFields' loop (M:1812-1848) for steps it0 .. it0+nsteps-1, in place:
 * stress+source, velocity, record.  `borders` (nullable) receives the
 * reference boundary save of v every step (getBorder_Vx_Vy_Vz, M:1524-1562);
 * W is the CPML width used for it. */
int efm_forward(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                const EfmState *st, float *gathers, float *borders, int W,
                int it0, int nsteps, void *stream);
/* efm_forward without borders, with the time loop captured once as a hipGraph
 * of `chunk` steps and replayed (no reference counterpart: the reference
 * issues 9+ launches from Python per step).  The kernels read the step index
 * from `counter`, a caller-provided device int, which the graph advances by
 * `chunk` per replay; steps beyond the last whole chunk run eagerly.  Needs a
 * non-default stream; returns after the work has completed. */
int efm_forward_graph(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                      const EfmState *st, float *gathers, int it0, int nsteps, int chunk,
                      int *counter, void *stream);
/* Out-of-place forward for steps it0 .. it0+nsteps-1 (no reference
 * counterpart; used to rebuild a checkpoint segment): states[k] holds the
 * pre-step state S^(it0+k), k = 0 .. nsteps; step it0+k reads states[k] and
 * writes states[k+1].  Only interior cells are written, so the state buffers
 * must start zeroed outside the interior.  gathers is nullable. */
int efm_forward_states(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                       const EfmState *const *states, float *gathers, int it0, int nsteps,
                       void *stream);

/* ---- exact discrete adjoint -------------------------------------------- */
/* Derived medium for the adjoint: mbar (3 components) = harmonic-mean μ_xy,
 * μ_xz, μ_yz; bbar (3) = 2/(ρ+ρ') along x, y, z; component stride N + EFM_PAD.
 * Interior only (zero elsewhere). */
int efm_derived_medium(const EfmDims *d, const EfmMedium *m, float *mbar, float *bbar,
                       void *stream);
/* Reverse sweep of steps n = n1-1 down to n0 (n0 < n1).  fwd[k] is the
 * forward pre-step state S^(n0+k), k = 0 .. n1-n0 (so fwd[n1-n0] is S^n1); only
 * its velocities and ψ slabs are read (its stress pointers must be valid but may
 * all share one buffer).
 * adj/adj2 are the adjoint state and a second ψ buffer (ping-pong; on return
 * the live ψ adjoint is in adj).  resid is ∂J/∂gathers (S, 3, ntmax, nrec).
 * acc (S·8 components of stride N + EFM_PAD, zeroed by the caller) accumulates
 * the λ, μ_n, μ_xy, μ_xz sums (interleaved per cell in components 0-3), the ρ
 * sum (component 4; it also covers the cells one past the interior on the high
 * x, y, z faces) and the μ_yz sum (component 7) of each shot; components 5
 * and 6 are not used.  Only efm_gradient_combine reads it.  gm6 (S, 6) double
 * accumulates ∂J/∂M.  The run is
 * covered by calls over consecutive descending ranges, the first with
 * n1 = ntmax.  Without a receiver mask `last` = 1 marks the call with n0 = 0
 * (it closes the moment-tensor sum of step 0); with one, the adjoint stress
 * sweep injects each next step's residual and sums ∂J/∂M itself and `last`
 * is not used. */
int efm_adjoint(const EfmDims *d, const EfmMedium *m, const float *mbar, const float *bbar,
                const EfmShots *sh, const EfmState *const *fwd, const EfmState *adj,
                const EfmState *adj2, const float *resid, float *acc, double *gm6,
                int n0, int n1, int last, void *stream);
/* The whole exact gradient of a stored forward history without any device
 * allocation and without caller workspaces.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops).  fwd[k] = S^k for k = 0 .. ntmax (every pre-step state);
 * adj is overwritten (zeroed, then the adjoint state); resid as for
 * efm_adjoint; glam, gmu, grho and gm6 are written (not accumulated).  The
 * derived medium, the accumulators and the second ψ-adjoint buffer live in one
 * static device array of 64 MiB per GPU, held until the process ends:
 * (6 + 8·S)·(N + EFM_PAD) + 6·S·(ny·nz·Lx + Ly·nz·nx + ny·Lz·nx) floats must
 * fit in 16 Mi floats (one 96³ shot with 10-cell layers does), else EFM_E_ARG.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (two streams or threads) race and must be serialised by the caller.  Uses
 * fill kernels, not memsets, so that it can be captured. */
int efm_gradient_static(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                        const EfmState *const *fwd, const EfmState *adj, const float *resid,
                        float *glam, float *gmu, float *grho, double *gm6, void *stream);
/* ∂J/∂λ, ∂J/∂μ, ∂J/∂ρ (ny, nz, nx) summed over shots from acc. */
int efm_gradient_combine(const EfmDims *d, const EfmMedium *m, const float *mbar,
                         const float *bbar, const float *acc, float *glam, float *gmu,
                         float *grho, void *stream);

/* ---- reference-style adjoint and imaging (drop-in parity) -------------- */
/* kernel_add_adjoint_source (M:310-322): v[rec] += res[nt-it-1]. */
int efm_inject_residual(const EfmDims *d, const EfmShots *sh, const EfmState *st,
                        const float *res, int it, void *stream);
/* kernel_Gradients_Mij (M:337-372), accumulated: gm6[s] += w[nt-it-1]·C(τ†(src)). */
int efm_image_mt_ref(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                     const EfmState *adj, double *gm6, int it, void *stream);
/* backwardFields (M:2056-2082, mode 0): nsteps adjoint steps (nt there; nt-1
 * inside backwardFields_FowardFieldsRemake) with the forward operator (+dt),
 * residual injection and MT imaging; adj must be zero on entry. */
int efm_backward_mt_ref(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                        const EfmState *adj, const float *res, double *gm6, int nsteps,
                        void *stream);
/* backwardFields_FowardFieldsRemake (M:1929-2054, mode 1), nt-1 iterations.
 * fwd is the final forward state (consumed), vpres is a 3-volume scratch
 * (S, 3, ny, nz, nx), borders from efm_forward, adj zero on entry.  Imaging
 * kernel_Gradients_Rho/_Lambda/_Mu (M:191-294) accumulated into glam, gmu,
 * grho (S, ny, nz, nx) in float32 like the reference's host sums. */
int efm_backward_remake_ref(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                            const EfmState *fwd, float *vpres, const float *borders, int W,
                            const EfmState *adj, const float *res, float dt, float dx,
                            float dy, float dz, float *glam, float *gmu, float *grho,
                            void *stream);
/* Iterations it0 .. it0+nsteps-1 of efm_backward_remake_ref (the body of the
 * reference's per-step loop, M:1939-2054, which it drives from Python one
 * iteration at a time).  Ranges that cover 0 .. nt-2 in order give the
 * launches of one efm_backward_remake_ref call. */
int efm_backward_remake_ref_steps(const EfmDims *d, const EfmMedium *m, const EfmShots *sh,
                                  const EfmState *fwd, float *vpres, const float *borders, int W,
                                  const EfmState *adj, const float *res, float dt, float dx,
                                  float dy, float dz, float *glam, float *gmu, float *grho,
                                  int it0, int nsteps, void *stream);
/* Size in floats of the boundary buffer efm_forward writes for one shot. */
long long efm_border_floats(const EfmDims *d, int W);

/* ---- persistent plans: the time loops as replayed HIP graphs ----------------
 * No reference counterpart (the reference issues 9-29 launches per step from
 * Python, M:1812-2082).  A plan belongs to one shot batch: efm_plan_create
 * copies d, m and sh (their arrays must outlive the plan), allocates a device
 * step counter and a capture stream, and keeps graphs keyed by the buffers,
 * step range and installed tiles of each call; a call whose key it holds
 * replays that graph on `stream`, otherwise the same launches as the eager
 * entry point are captured (not run), instantiated, kept and replayed.  No
 * call synchronises the host; results are bit-identical to the eager entry
 * points.  Changing a tile drops the plan's graphs.  One call at a time per
 * plan (calls are serialised internally). */
typedef struct EfmPlan EfmPlan;
int efm_plan_create(const EfmDims *d, const EfmMedium *m, const EfmShots *sh, int chunk,
                    EfmPlan **out);
int efm_plan_destroy(EfmPlan *p);
/* Number of graphs the plan holds. */
int efm_plan_graphs(const EfmPlan *p);
/* efm_forward without borders: whole chunks of `chunk` steps replay one graph
 * per (state, gathers) that reads the step index from the plan's counter; the
 * remaining steps launch eagerly. */
int efm_plan_forward(EfmPlan *p, const EfmState *state, float *gathers, int it0, int nsteps,
                     void *stream);
/* efm_forward_states as one graph per (states, gathers, it0, nsteps). */
int efm_plan_forward_states(EfmPlan *p, const EfmState *const *states, float *gathers, int it0,
                            int nsteps, void *stream);
/* efm_adjoint as one graph per (buffers, n0, n1, last); the ψ copy-back that
 * efm_adjoint may end with is a copy node of the graph. */
int efm_plan_adjoint(EfmPlan *p, const float *mbar, const float *bbar, const EfmState *const *fwd,
                     const EfmState *adj, const EfmState *adj2, const float *resid, float *acc,
                     double *gm6, int n0, int n1, int last, void *stream);

/* ---- FWI gradient conditioning and model maps (fwi_Vp_Vs_Rho, M:2479-2504,
 * 2577-2583, 2656-2661, 2688-2694; Model M:1200-1206) -------------------- */
typedef struct EfmSmooth {
    float sigma;       /* Gaussian σ in cells; 0 = no smoothing (masks only) */
    float truncate;    /* kernel radius = int(truncate·σ + 0.5) cells, at most 64 (scipy's rule) */
    int margin[6];     /* cells zeroed from the x_lo, x_hi, y_lo, y_hi, z_lo, z_hi faces */
} EfmSmooth;
/* out = mask(gaussian_filter(in, σ, mode='reflect', truncate)) of one
 * (ny, nz, nx) volume: scipy.ndimage.gaussian_filter's separable passes in its
 * axis order, each summed in double and rounded to float32, then the face
 * margins zeroed (the reference zeroes z < 10 and the CPML+2 zone of the other
 * five faces, M:2479-2504).  work is a caller volume of the same size, needed
 * when σ > 0; in, out and work must not overlap. */
int efm_condition_gradient(const EfmSmooth *p, int nx, int ny, int nz, const float *in,
                           float *out, float *work, void *stream);
/* The FWI model update of one (ny, nz, nx) model (M:2577-2583, 2688-2694):
 * lam = lam0 + alpha_lam·dlam and mu = mu0 + alpha_mu·dmu (in double, rounded once;
 * a null direction keeps lam0 / mu0), then Model.compute_Vp_Vs with rho_in
 * (vp = sqrt((lam + 2mu)/rho_in), vs = sqrt(mu/rho_in), float32, M:1200-1202)
 * and, when rho is not null, Model.compute_Rho (Nafe-Drake of vp, M:1204-1206)
 * except at z < top, where rho_keep (if not null) is copied.  Outputs may
 * alias the lam0/mu0 inputs; rho must not alias rho_in or rho_keep. */
int efm_model_update(const float *lam0, const float *dlam, double alpha_lam, const float *mu0,
                     const float *dmu, double alpha_mu, const float *rho_in, const float *rho_keep, int top,
                     int nx, int ny, int nz, float *lam, float *mu, float *vp, float *vs,
                     float *rho, void *stream);
/* computeResidualCrossCorrelation (M:2145-2179) for gathers (S, 3, ntmax, nrec):
 * per trace, with a = <mod, obs>, res = mod·a/(|obs|·|mod|³) - obs/(|mod|·|obs|)
 * (zero |mod|, |obs| or a replaced by 1, as the reference does), sums in
 * double; J (a device double the caller zeroes) += Σ over traces of
 * 1 - a/(|mod|·|obs|), the misfit this residual is the derivative of. */
int efm_xcorr_residual(const float *mod, const float *obs, float *res, double *J,
                       int nshot, int ntmax, int nrec, void *stream);
/* Model.compute_Rho (M:1204-1206): rho = Nafe-Drake(vp), n values. */
int efm_model_rho_nafe_drake(const float *vp, float *rho, long long n, void *stream);
/* y += a·x over n floats (per-shot gradient sums, as the reference's
 * float32 tempLambda += gk.Lambda, M:2466-2468). */
int efm_volume_axpy(float a, const float *x, float *y, long long n, void *stream);

#ifdef __cplusplus
}
#endif
#endif
