/*
 * lmhip.h -- C ABI of the MI355X (gfx950) backend for LangevInMedium.jl's per-particle
 * Langevin step.
 *
 * The library (liblmhip.so) owns the particle state in HBM as FP64 structure-of-arrays, the
 * medium and the host-built transport tables, and one HIP stream per context.  The reference
 * reaches its CPU and CUDA kernels through `simulate_ensemble_bulk(::Backend, ...)`; a new
 * backend type whose method calls these entry points replaces the CUDA driver
 * (src/simulate_gpu.jl:26-462) and its kernels (src/kernels_gpu.jl, src/kernels_gpu_GC.jl),
 * and follows the CPU kernels (src/kernels_cpu.jl, src/kernels_cpu_GC.jl) as the numerical
 * contract.
 *
 * Conventions
 *   - Host arrays are borrowed for the duration of a call only.
 *   - Matrices use Julia's column-major layout: a (rows, N) particle matrix holds element
 *     (d, i) at [d + rows*i]; an (n_r, n_t) medium matrix holds (i, j) at [i + n_r*j]; a
 *     2+1D grid field (nx, ny, nt) holds (ix, iy, it) at [ix + nx*(iy + ny*it)].
 *   - Every entry point returns LMH_OK (0) or a negative status; lmh_last_error() gives the
 *     message.  No C++ exception crosses the ABI.
 *   - A context is not thread-safe; use one per host thread (the reference runs one Julia task).
 */
#ifndef LMHIP_H
#define LMHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------- */
#define LMH_OK            0
#define LMH_E_INVALID    -1  /* an argument is out of range or inconsistent                   */
#define LMH_E_STATE      -2  /* call order: no medium, particles or table set yet             */
#define LMH_E_HIP        -3  /* HIP runtime error (message carries hipGetErrorString)         */
#define LMH_E_NONFINITE  -4  /* NaN/Inf in the state after a step (simulate_cpu.jl:293-303)   */
#define LMH_E_NOMEM      -5  /* device allocation failed                                      */

/* ---- enumerations ---------------------------------------------------------------------- */
/* coordinates: Cartesian step (simulate_cpu.jl:215-316) or the general-coordinate (tau, r)
 * step (simulate_cpu_general_coords.jl:83-126), or MILNE (extension): position rows
 * (x, y, eta_s) and momentum rows (p_x, p_y, p_z) with p_z the longitudinal momentum in the
 * frame comoving with the Bjorken flow at eta_s, on the radial medium.  A MILNE step is the
 * Cartesian step with bjorken_redshift (kernels_cpu.jl:427-435) on (x, y, p) -- the medium at
 * the transverse radius -- followed by the space-time rapidity of the free path over the step
 * tau_a = t0 + (step - 1) dt -> tau_a + dt (p_z tau conserved, y - eta_s = asinh(p_z/m_T), the
 * note at kernels_cpu.jl:419-426): eta_s += asinh(k b) - asinh(b), b = p_z/m_T at the end of the
 * step, k = 1 + dt/tau_a.  Needs dims == pdim == 3, t0 > 0, no position diffusion, reflecting
 * wall or freeze-out capture.                                                               */
#define LMH_CARTESIAN          0
#define LMH_GENERAL_COORDS     1
#define LMH_MILNE              2

/* momentum update branch (simulate_cpu.jl:229-278):
 *   LANGEVIN      forces + update + boost back                     (default :langevin)
 *   RTA           BGK redraw with probability dt/tau_n + boost back   (collision_mode=:rta)
 *   FLUID         glue to the flow, no boost back  (momentum_langevin=false or DsT == 0)
 *   LANGEVIN_RTA  the Langevin kick followed by the BGK redraw, then the boost back; an
 *                 extension with no reference counterpart (the reference's modes are
 *                 exclusive)                                                              */
#define LMH_LANGEVIN           0
#define LMH_RTA                1
#define LMH_FLUID              2
#define LMH_LANGEVIN_RTA       3

/* noise source: FIXED reads a host- or device-supplied record per step; PHILOX generates the
 * record in registers.  Block b (0 kick and collision, 2 position diffusion) of particle i at
 * step s is four 64-bit words from two Philox4x32-10 evaluations (half h gives words 2h, 2h+1,
 * each lo | hi << 32 of consecutive outputs) of counter (i, s, stream mod 2^32,
 * (stream >> 32) mod 2^29 | (2b + h) << 29) under key (seed mod 2^32, seed >> 32); i and s must
 * stay below 2^32 and stream below 2^61.  With lo(w), hi(w) a word's 32-bit halves and
 * d = 1.m (m a word's top 52 bits):
 *   block 0 (kick and collision): the kick normals are the first three components of a 4-D
 *     standard normal vector R (sqrt(v) cos(beta), sqrt(v) sin(beta), sqrt(1 - v) cos(gamma), .)
 *     with R^2 = -2 ln(u1 u2) (u1 = 2 - d of word 0, u2 = (lo(w1) + 1)/2^32), beta = pi (q/2 + r)
 *     (q the top 2 bits of hi(w1), r = lo30(hi(w1))/2^31 - 1/4), v = lo(w2)/2^32, and cos(gamma)
 *     = +-cos(pi x) with x = lo31(hi(w2))/2^32 and the sign bit 31 of hi(w2).  The collision
 *     uniform is lo31(w3)/2^31, the redraw's quantile u_samp = hi(w3)/2^32, and the redraw's
 *     Gaussian g is the three kick normals (drawn for the redraw when the step has no kick): a
 *     colliding particle's kick is replaced by the redraw, so its normals serve it, and the
 *     normalised Gaussian is the reference's isotropic direction (kernels_cpu.jl:497-503).  The
 *     general-coordinate step's one normal is the first of the Box-Muller pair below of words
 *     0 and 1.
 *   block 2 (position diffusion): Box-Muller pairs of words (0, 1) and (2, 3): the radius
 *     sqrt(-2 ln(2 - d)) of the first word, the angle pi (q/2 + r) of the second (q its top 2
 *     bits, r = d/2 - 3/4 in [-1/4, 1/4) with d from its next 52 bits).                     */
#define LMH_NOISE_FIXED        0
#define LMH_NOISE_PHILOX       1

/* integrator of the Ornstein-Uhlenbeck kick:
 *   PREPOINT   exact OU with the pre-point drag (kernels_cpu.jl:329-346)
 *   MIDPOINT   drift-midpoint drag (kernels_gpu.jl:368-381, integrator_mode=1)
 *   POSTPOINT  drag evaluated at the updated momentum by three fixed-point sweeps
 *              (extension; reduces to PREPOINT when the drag does not depend on p)         */
#define LMH_PREPOINT           0
#define LMH_MIDPOINT           1
#define LMH_POSTPOINT          2

/* drag model:
 *   TABLE    eta_D = 1/tau_drag(T) from the host-built table, times m/E when relativistic
 *            (kernels_cpu.jl:281-290,339)
 *   GAMMA_P  eta(p,T) = gamma0*T^2/(1+p^2/Lambda^2) with kappa = 2*m*T*gamma0*T^2
 *            (extension)                                                                  */
#define LMH_DRAG_TABLE         0
#define LMH_DRAG_GAMMA_P       1

/* time tables built on the host by the reference's transport.jl builders */
#define LMH_TABLE_DRAG         0   /* build_tau_drag_spline (transport.jl:208-209)          */
#define LMH_TABLE_RTA          1   /* build_taun_current_spline (transport.jl:221-222), or
                                      any relaxation-time law on the same uniform T grid    */

#define LMH_MAX_SPECIES        4

typedef struct lmh_params {
    double   dt;                  /* Δt [fm/c]                                               */
    double   t0;                  /* initial_time; the kick of step s samples s*dt + t0      */
    double   DsT;                 /* D_s·T: gates the drag (kernels_cpu.jl:270,281) and the
                                     RTA time (kernels_cpu.jl:488); position diffusion uses it */
    double   psi2;                /* Ψ₂ of the optional v₂ flow modulation (kernels_cpu.jl:86) */
    double   gamma0;              /* LMH_DRAG_GAMMA_P coefficient [fm^-1 GeV^-2]             */
    double   Lambda;              /* LMH_DRAG_GAMMA_P momentum scale [GeV]                   */
    uint64_t seed;                /* Philox key word 0                                       */
    uint64_t stream;              /* Philox key word 1                                       */
    int32_t  coords;              /* LMH_CARTESIAN | LMH_GENERAL_COORDS | LMH_MILNE          */
    int32_t  collision;           /* LMH_LANGEVIN | LMH_RTA | LMH_FLUID | LMH_LANGEVIN_RTA    */
    int32_t  noise;               /* LMH_NOISE_FIXED | LMH_NOISE_PHILOX                      */
    int32_t  integrator;          /* LMH_PREPOINT | LMH_MIDPOINT | LMH_POSTPOINT             */
    int32_t  drag_model;          /* LMH_DRAG_TABLE | LMH_DRAG_GAMMA_P                       */
    int32_t  relativistic;        /* Lorentz vs Galilean kinematics (simulate.jl:54-61)      */
    int32_t  bjorken_redshift;    /* p_z <- p_z*tau_a/(tau_a+dt) (kernels_cpu.jl:427-435)   */
    int32_t  position_diffusion;  /* overdamped position noise (kernels_cpu.jl:553-565,591-607) */
    int32_t  reflecting_boundary; /* wall at r_max = last(xgrid) (kernels_cpu.jl:574-581,609-634) */
    int32_t  check_finite;        /* 1: flag the first step whose state is non-finite        */
} lmh_params;

/* Fixed-noise record of one step: a (C, N) row-major block of doubles, channel c of
 * particle i at [c*N + i]; a run of nsteps reads (nsteps, C, N).  Channels, in order:
 *   xi[pdim]            momentum kick normals            (LANGEVIN, LANGEVIN_RTA; the GC
 *                                                         step uses one channel, for p^r)
 *   u_col, u_samp       collision decision and |p*| quantile, U[0,1)   (RTA, LANGEVIN_RTA)
 *   g[pdim]             raw Gaussian direction of the redraw           (RTA, LANGEVIN_RTA)
 *   xpos[dims]          position-diffusion normals                     (position_diffusion)
 * lmh_noise_channels() returns C for a configuration.                                      */
int32_t lmh_noise_channels(const lmh_params* p, int32_t dims, int32_t pdim);

/* observables over the current state (no reference counterpart: the reference returns raw
 * histories, simulate_cpu.jl:318-322).  out receives, in order:
 *   [0] N_sel  [1] sum p_T  [2] sum p_T^2  [3] sum cos(2 phi)  [4] sum p_T*cos(2 phi)
 *   [5 .. 5+nbins)  p_T histogram on [pt_min, pt_max), equal bins, or its Gaussian-smoothed
 *                   counterpart when sigma > 0 (each particle adds
 *                   Phi((hi-pT)/sigma) - Phi((lo-pT)/sigma) to every bin)
 * cos(2 phi) = (px^2 - py^2)/p_T^2; particles with p_T == 0 add 0 to it.  species < 0
 * selects all species.                                                                      */
typedef struct lmh_obs_spec {
    int32_t nbins;
    int32_t species;
    double  pt_min, pt_max;
    double  sigma;
} lmh_obs_spec;

/* launch shape, chosen per context (by the host autotuner or by hand) */
typedef struct lmh_tuning {
    int32_t block;             /* threads per block: 64, 128 or 256 (default 256)               */
    int32_t steps_per_launch;  /* steps a particle stays in registers when lmh_run gets 0 (1)  */
    int32_t slice;             /* 2+1D grid, one-step launches: 1 interpolates the medium in tau
                                  once per step into a slice the particles read (default 1;
                                  lmh_adjoint_forward always runs this way)                    */
    int32_t defer;             /* 2+1D grid, one-step launches with RTA: 1 queues the colliding
                                  particles and finishes them in a second kernel, so the
                                  redraw runs in colliding lanes only; 0 redraws them in the
                                  step kernel (default 0: the faster on MI355X at a few % of
                                  collisions per step; identical results)                     */
    int32_t specialize;        /* 1: the common configuration (relativistic, pre-point, drag
                                  table, no Bjorken redshift, no position diffusion) runs
                                  kernels compiled for it, with identical results (default 1) */
    int32_t sort_every;        /* 2+1D grid, Philox noise, no freeze-out capture: lmh_run keeps
                                  the particles ordered by grid cell (species segments kept),
                                  sorting them on the device before its first step and again
                                  every sort_every steps (the specialised sliced step applies
                                  the sort's permutation by reading through it), so that
                                  neighbouring threads read
                                  neighbouring medium cells.  Each particle keeps its identity
                                  (its noise counters), so results do not depend on it; every
                                  other entry point first restores the caller's order.  0: never
                                  (default 64)                                                  */
} lmh_tuning;

/* device time per launch class, recorded with HIP events on the context's stream between
 * lmh_profile_start and lmh_profile_stop                                                     */
#define LMH_PROF_STEP      0   /* the step kernels                                         */
#define LMH_PROF_REDRAW    1   /* the deferred-collision kernels                           */
#define LMH_PROF_SLICE     2   /* the medium-slice kernels                                 */
#define LMH_PROF_SORT      3   /* the cell sort (keys, radix sort, permutation)            */
typedef struct lmh_profile {
    int64_t launches[4];
    double  ms[4];
} lmh_profile;

/* pathwise adjoint (extension): the forward pass keeps the state every checkpoint_every steps.
 * dtau_dDsT is (d tau_drag / d DsT) / tau_drag for the drag table in use: 1/DsT when the table
 * is proportional to DsT (effective_DsT without modifiers or with the quadratic law,
 * transport.jl:139-156), 0 for the linear law, which does not depend on DsT.                */
typedef struct lmh_adjoint_spec {
    int32_t checkpoint_every;  /* segment length K: a checkpoint every K steps                  */
    int32_t keep_tape;         /* 1: keep every step's tape row and medium slice in device
                                  memory (nsteps x (2 + pdim) x N doubles: the state, no noise; plus
                                  nsteps x 96 B per grid cell), so that the pullback only runs
                                  the reverse sweep; 0: checkpoints only, the pullback re-runs
                                  each segment                                                 */
    double  dtau_dDsT;
} lmh_adjoint_spec;

/* one run of a batch (extension): its particle count, mass and D_s T, the index of its medium
 * (lmh_batch_add_medium_grid2p1) and of its tables (lmh_batch_add_time_table,
 * lmh_batch_add_juttner_table; -1 for none)                                                  */
typedef struct lmh_run_desc {
    int64_t count;
    double  mass;
    double  DsT;
    int32_t medium;
    int32_t drag_table;
    int32_t rta_table;
    int32_t juttner_table;
} lmh_run_desc;

/* launch shape of a batch, chosen by lmh_batch_autotune or by hand                          */
typedef struct lmh_batch_tuning {
    int32_t block;             /* threads per block (= particles per tile): 64, 128 or 256     */
    int32_t steps_per_launch;  /* steps a particle stays in registers per launch               */
    int32_t persistent;        /* 1: a fixed grid of blocks walks the tiles; 0: a block a tile */
    int32_t graph_steps;       /* > 0: lmh_batch_run replays a captured HIP graph of this many
                                  steps (the step index advances on the device); 0: direct    */
    int32_t sort_every;        /* > 0: lmh_batch_run keeps each run's quarks ordered by grid
                                  cell, sorting them on the device before its first step and
                                  then every sort_every steps (rounded up to whole graph
                                  replays); the results do not depend on it; 0: never        */
    int32_t slice_min;         /* > 0: the runs of at least slice_min quarks (on grids of at
                                  most 2^27 nodes) read their medium from per-step node slices
                                  (each grid node interpolated in tau once per step and run,
                                  the cell's coefficients then formed as lmh_run's one-step
                                  launches form them, with the same results as such a run
                                  alone; in the specialised configuration they step in a kernel
                                  of their own, one step per launch); the other runs
                                  interpolate each corner in tau per quark; 0: no slices       */
} lmh_batch_tuning;

typedef struct lmh_ctx lmh_ctx;
typedef struct lmh_batch lmh_batch;

/* ---- lifetime ---------------------------------------------------------------------------- */
int         lmh_create(lmh_ctx** out, int32_t device);
void        lmh_destroy(lmh_ctx* ctx);
const char* lmh_last_error(const lmh_ctx* ctx);
/* the context's HIP stream (a hipStream_t), for events and host synchronisation */
void*       lmh_stream(lmh_ctx* ctx);
int         lmh_synchronize(lmh_ctx* ctx);

int         lmh_set_tuning(lmh_ctx* ctx, const lmh_tuning* t);
int         lmh_get_tuning(lmh_ctx* ctx, lmh_tuning* t);

/* ---- medium ------------------------------------------------------------------------------ */
/* Radial (r, t) medium of the reference: T, v and the optional v2 are (nr, nt) matrices on
 * sorted grids (kernels_cpu.jl:22-51, simulate.jl:21-27).  v2 may be NULL.                  */
int lmh_set_medium_radial(lmh_ctx* ctx,
                          const double* rgrid, int32_t nr,
                          const double* tgrid, int32_t nt,
                          const double* T, const double* v, const double* v2);

/* 2+1D (x, y, tau) medium (extension): uniform x and y nodes x0 + k*dx, sorted tgrid, fields
 * T, vx, vy as (nx, ny, nt) column-major.  Queries are clamped to the grid.                */
int lmh_set_medium_grid2p1(lmh_ctx* ctx,
                           double x0, double dx, int32_t nx,
                           double y0, double dy, int32_t ny,
                           const double* tgrid, int32_t nt,
                           const double* T, const double* vx, const double* vy);

/* ---- host-built transport tables (transport.jl:165-196,277-365) -------------------------- */
/* uniform-grid linear time table: node k holds the time at Tmin + k/invdT                   */
int lmh_set_time_table(lmh_ctx* ctx, int32_t which, int32_t species,
                       double Tmin, double invdT, const double* vals, int32_t n);
/* inverse-CDF Jüttner table, invcdf[jT*nU + iu] = p*(u_iu; T_jT) (transport.jl:268-270)   */
int lmh_set_juttner_table(lmh_ctx* ctx, int32_t species,
                          const double* invcdf, int32_t nU, int32_t nT,
                          double Tmin, double invdT);

/* ---- initial-state samplers on the device (utils.jl:33-172, 499-558) --------------------- */
/* The inverse-CDF tables of sample_particles_from_FONLL, built on the host as utils.jl:52-71
 * (p|r) and 130-146 (polar r marginal) or 80-90 (Cartesian spatial density) build them.       */
typedef struct lmh_fonll_tables {
    const double* r_grid;      /* nr, ascending                                               */
    const double* p_grid;      /* np, ascending                                               */
    const double* cdf_r;       /* nr: CDF of the r marginal (polar sampling; else unused)     */
    const double* n_r;         /* nr: spatial density on r_grid (Cartesian sampling)          */
    const double* cdf_p;       /* (np, nr) column-major: the p|r CDF of every r node          */
    const double* p_ok;        /* nr: 1 where the p|r density is positive (else |p| = 0)      */
    int32_t nr, np;
    int32_t cartesian;         /* 0: polar (r, phi) sampling, 1: rejection in the (x, y) square */
    int32_t pad_;
    double  rmax, fmax;        /* Cartesian: the square's half side and the envelope          */
} lmh_fonll_tables;
/* sample_particles_from_FONLL (utils.jl:33-172) for N quarks, with the antithetic pairing of
 * simulate_cpu.jl:66-89 when antithetic != 0 (quarks 2k, 2k+1 share a position and have
 * opposite momenta; an odd last quark is drawn alone).  Quark k's draws are Philox4x32-10 words
 * of noise block 3 under key (seed, stream); x_out and p_out receive (2, N) column-major
 * matrices.  Needs no particles or medium; N < 2^32.  Synchronises.                          */
int lmh_sample_fonll(lmh_ctx* ctx, const lmh_fonll_tables* tables, int64_t N, int32_t antithetic,
                     uint64_t seed, uint64_t stream, double* x_out, double* p_out);
/* append_thermal_pz (utils.jl:533-558): out (3, N) receives the (2, N) transverse momenta mom
 * and a p_z row drawn by sample_pz_conditional_juttner (utils.jl:499-530) at m_T =
 * sqrt(m^2 + p_T^2) and T = max(T(|x|, t0), 0) of the context's radial medium (pos is (dims,
 * N)); antithetic != 0 gives quark 2k+1 minus the p_z of quark 2k.  Philox words of block 3
 * under key (seed, stream).  Synchronises.                                                   */
int lmh_append_thermal_pz(lmh_ctx* ctx, const double* mom, const double* pos, int32_t dims, int64_t N,
                          double m, double t0, int32_t antithetic, uint64_t seed, uint64_t stream,
                          double* out);

/* ---- particles --------------------------------------------------------------------------- */
/* pos (dims, N) and mom (pdim, N), Julia layout.  Species are contiguous segments:
 * counts[k] particles of mass masses[k], in order; sum(counts) == N.
 * Cartesian: (dims, pdim) in {(1,1),(2,2),(3,3),(2,3)}, the check_momentum_dims contract
 * (utils.jl:561-573).  General coordinates: dims == pdim == 2, rows (tau, r), (p^tau, p^r). */
int lmh_set_particles(lmh_ctx* ctx,
                      const double* pos, int32_t dims,
                      const double* mom, int32_t pdim,
                      int64_t N, int32_t nspecies,
                      const int64_t* counts, const double* masses);
int lmh_get_particles(lmh_ctx* ctx, double* pos, double* mom);

/* ---- stepping ---------------------------------------------------------------------------- */
/* The driver's initial boost of the LRF momenta to the lab frame at `step`
 * (simulate_cpu.jl:160-163; step 0).  Cartesian only.                                       */
int lmh_boost_to_lab(lmh_ctx* ctx, const lmh_params* p, int64_t step);

/* Steps step_begin .. step_begin+nsteps-1 of the driver loop (simulate_cpu.jl:215-303 or
 * simulate_cpu_general_coords.jl:83-115).  noise: NULL for LMH_NOISE_PHILOX, else a host
 * array (nsteps, C, N).  steps_per_launch: 0 lets the library choose; k >= 1 keeps the state
 * in registers for k steps per kernel launch (the results do not depend on k).             */
int lmh_run(lmh_ctx* ctx, const lmh_params* p, int64_t step_begin, int64_t nsteps,
            const double* noise, int32_t steps_per_launch);
/* as lmh_run, with the fixed-noise record already in device memory                         */
int lmh_run_device_noise(lmh_ctx* ctx, const lmh_params* p, int64_t step_begin,
                         int64_t nsteps, const double* d_noise, int32_t steps_per_launch);
/* step of the first non-finite state seen since the last lmh_set_particles, or -1           */
int64_t lmh_nonfinite_step(lmh_ctx* ctx);

/* ---- timing of the launches (adds two events per launch while active) -------------------- */
int lmh_profile_start(lmh_ctx* ctx);
int lmh_profile_stop(lmh_ctx* ctx, lmh_profile* out);     /* synchronises */

/* ---- observables ------------------------------------------------------------------------- */
/* out has 5 + spec->nbins doubles (see lmh_obs_spec); synchronises                          */
int lmh_observables(lmh_ctx* ctx, const lmh_obs_spec* spec, double* out);

/* The same observables into device memory with no allocation -- no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  d_out (5 + nbins doubles, device memory) receives
 * the sums of the particles in their current slot order (the cell sort of lmh_run may have
 * reordered them: the sums agree with lmh_observables to rounding); the kernels are enqueued on
 * the context's stream and the call returns without synchronising.  The per-block partial sums
 * live in one static device array per GPU (1024 x 125 doubles, 1 MB, held until the process
 * ends): nbins is at most 120, and calls on one GPU that overlap in time (several contexts or
 * streams) race on that array.  Calls no hipMalloc, hipFree or allocating library routine.   */
int lmh_observables_noalloc(lmh_ctx* ctx, const lmh_obs_spec* spec, double* d_out);

/* ---- snapshots (the saved states of simulate_cpu.jl:305-316, simulate_gpu.jl:385-410) ------ */
/* Takes the current state in the order of the context's stream (after every step enqueued so
 * far) and copies it to pinned host memory on a second stream while later steps run; returns
 * immediately with the snapshot's index.  lmh_snapshot_get waits for that copy, writes pos
 * (dims, N) and mom (pdim, N) and frees the index for reuse.                                  */
int lmh_snapshot(lmh_ctx* ctx, int32_t* index);
int lmh_snapshot_get(lmh_ctx* ctx, int32_t index, double* pos, double* mom);

/* ---- on-the-fly freeze-out capture (kernels_gpu.jl:583-635, simulate_gpu.jl:53-75,256-265,
 * 375-402; the GPU driver's freezeout_capture/freezeout_interp) ---------------------------- */
/* Tfo > 0 enables the capture for the particles set (and for every later lmh_set_particles),
 * with nothing latched; Tfo <= 0 disables it.  After each step of lmh_run, a particle not yet
 * latched whose medium temperature at its new position and the step's time t is below Tfo is
 * latched: its state interpolated linearly in T between the step's input and output state (the
 * last state above Tfo; interp = 0 or no earlier T above Tfo books the output), the time
 * interpolated the same way between t - dt and t, flag 1.  T is the step's medium
 * interpolation (interpolate_2d_cpu semantics on the radial medium, the 2+1D grid's otherwise).
 * Cartesian coordinates; the capture runs in the generic (not the specialised or deferred)
 * kernels.                                                                                   */
int lmh_freezeout_enable(lmh_ctx* ctx, double Tfo, int32_t interp);
/* pos (dims, N), mom (pdim, N), tau (N), flag (N: 1 latched, 0 not yet)                       */
int lmh_freezeout_get(lmh_ctx* ctx, double* pos, double* mom, double* tau, double* flag);

/* ---- pathwise adjoint (extension; the reference has no differentiation, Project.toml:6-27) -- */
/* The forward half of a ChainRulesCore.rrule / torch.autograd.Function: runs steps
 * step_begin .. step_begin+nsteps-1 exactly as lmh_run does with one step per launch and the
 * sliced medium (the default tuning: same final state, same noise) and keeps checkpoints; the
 * pullback re-runs segments from them through the same per-step medium slices.  With
 * spec->keep_tape the forward pass writes the whole tape instead (no freeze-out capture).  The 2+1D grid in Cartesian coordinates, Philox noise, the relativistic
 * pre-point Langevin kick with or without RTA collisions, drag TABLE or GAMMA_P.            */
int lmh_adjoint_forward(lmh_ctx* ctx, const lmh_params* p, int64_t step_begin, int64_t nsteps,
                        const lmh_adjoint_spec* spec);
/* The pullback: for the loss L = sum_k cotangent[k] * O_k over the observables O of
 * lmh_observables(spec) of the final state, grad receives dL/d(DsT, gamma0, Lambda) along the
 * forward pass's noise path (collision decisions held fixed); dpos (dims, N) and dmom
 * (pdim, N), both or neither, receive dL/d(state before the forward pass).  Re-runs each
 * checkpoint segment (not with keep_tape); may be called again with another cotangent.  Fails with LMH_E_STATE
 * when the particles, medium or tables changed since lmh_adjoint_forward.  Synchronises.   */
int lmh_adjoint_backward(lmh_ctx* ctx, const lmh_obs_spec* spec, const double* cotangent,
                         double* grad, double* dpos, double* dmom);

/* ---- batches of independent runs (extension; the reference runs one ensemble per call with
 * scalar m and DsT, simulate_cpu.jl:20-21) -------------------------------------------------- */
/* Runs r = 0..nruns-1 on the 2+1D grid, each with its own medium, tables, mass, D_s T and
 * particle count, advanced together by one launch per step (or per steps_per_launch steps).
 * Run r draws Philox noise with key (seed, stream + r) and counter (local index, step, block):
 * it follows the trajectories lmh_run gives for the same particles alone with stream + r and
 * tuning slice = 0 (the batch interpolates each medium corner in tau).                       */
int         lmh_batch_create(lmh_batch** out, int32_t device);
void        lmh_batch_destroy(lmh_batch* b);
const char* lmh_batch_last_error(const lmh_batch* b);
void*       lmh_batch_stream(lmh_batch* b);
int lmh_batch_add_medium_grid2p1(lmh_batch* b, double x0, double dx, int32_t nx,
                                 double y0, double dy, int32_t ny,
                                 const double* tgrid, int32_t nt,
                                 const double* T, const double* vx, const double* vy,
                                 int32_t* index);
int lmh_batch_add_time_table(lmh_batch* b, double Tmin, double invdT, const double* vals,
                             int32_t n, int32_t* index);
int lmh_batch_add_juttner_table(lmh_batch* b, const double* invcdf, int32_t nU, int32_t nT,
                                double Tmin, double invdT, int32_t* index);
/* pos (2, N) and mom (pdim, N) of all runs, run after run; N = sum of the counts            */
int lmh_batch_set_runs(lmh_batch* b, const lmh_run_desc* runs, int32_t nruns,
                       const double* pos, const double* mom, int32_t pdim);
int lmh_batch_get_particles(lmh_batch* b, double* pos, double* mom);
int lmh_batch_set_tuning(lmh_batch* b, const lmh_batch_tuning* t);
int lmh_batch_get_tuning(lmh_batch* b, lmh_batch_tuning* t);
/* the initial boost of every run to the lab at `step` (simulate_cpu.jl:160-163)            */
int lmh_batch_boost_to_lab(lmh_batch* b, const lmh_params* p, int64_t step);
/* steps step_begin .. step_begin+nsteps-1 of every run; p->DsT is ignored (each run's own)  */
int lmh_batch_run(lmh_batch* b, const lmh_params* p, int64_t step_begin, int64_t nsteps);
/* times trial_steps steps of every launch shape (block 64/128/256 x persistent 0/1 x
 * steps_per_launch 1/2/4 x slice_min 0/2^14/2^16/2^18) on this batch, then 200 steps at the fastest shape
 * for each sort cadence (sort_every 0, 50, 100, 200), keeps the fastest (graph_steps unchanged) and
 * restores the state; the results do not depend on the block, persistence, steps per launch or
 * cadence (slice_min selects each run's form of the medium interpolation, whose results agree
 * to rounding).  Synchronises.                                                              */
int lmh_batch_autotune(lmh_batch* b, const lmh_params* p, int32_t trial_steps,
                       lmh_batch_tuning* best, double* best_ms_per_step);
/* out (nruns, 5 + nbins): lmh_observables of every run (spec->species is ignored)           */
int lmh_batch_observables(lmh_batch* b, const lmh_obs_spec* spec, double* out);
int64_t lmh_batch_nonfinite_step(lmh_batch* b);

#ifdef __cplusplus
}
#endif
#endif /* LMHIP_H */
