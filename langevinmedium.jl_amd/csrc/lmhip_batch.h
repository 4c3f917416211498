// lmhip_batch.h -- many independent 2+1D runs in one launch (included by lmhip.hip inside its
// anonymous namespace, after the forward kernels it reuses).
//
// A batch holds runs r = 0..R-1, each with its own particle segment (contiguous in one SoA
// state), medium, tables, mass and D_s T.  The launch is cut into tiles of blockDim particles,
// every tile inside one run; tile_run[t] names the run.  Run r's noise is Philox with key
// (seed, stream + r) and counter (local particle index, step, block), so a run inside the batch
// follows the trajectories the same run gives alone with stream + r.  The first step of a launch
// is read from device memory (*step_dev + k_off), so a captured graph of launches advances by
// itself when a final one-thread kernel moves *step_dev on.

struct BatchRun {
    int64_t begin, count, tile0;
    GridMedium g;               // g.slice: the run's ring of per-step node slices (node_sample;
                                // one per step, slice_stride nodes apart), or null
    SpeciesArgs sp;
    double DsT;
    int64_t slice_stride;
};

struct BatchArgs {
    StepArgs base;              // state rows over all runs, the shared parameters
    const BatchRun* runs;
    const int32_t* tile_run;
    int64_t ntiles;             // tiles the launch walks: all, or the entries of tile_list
    const int32_t* tile_list;   // null: tile t is tile t; else tile t is tile_list[t]
    const int64_t* step_dev;
    // the tau cell and fraction of run r at the launch's step s, tcell[r * tcell_k + s] = (j, td),
    // written by k_batch_tcell before the launch (null: each thread finds them itself)
    const double2* tcell;
    int32_t tcell_k;
};

// A block-uniform struct read from device memory into scalar registers.  The compiler issues
// vector loads for it (it cannot tell that the kernel never writes the run table), so every
// 32-bit word is made wave-uniform with readfirstlane: what derives from the run (medium and
// table pointers, sizes, mass) then stays in scalar registers instead of occupying VGPRs.
template <typename T>
__device__ __forceinline__ T uniform_read(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "uniform_read: size must be a multiple of 4 bytes");
    uint32_t w[sizeof(T) / 4];
    __builtin_memcpy(w, p, sizeof(T));
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); ++k) w[k] = __builtin_amdgcn_readfirstlane(w[k]);
    T out;
    __builtin_memcpy(&out, w, sizeof(T));
    return out;
}

// this tile's run as a single-run view: state rows (read and written) and Philox stream of the run
__device__ __forceinline__ void batch_view(const BatchArgs& b, const BatchRun& run, int r, StepArgs* a) {
    *a = b.base;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (a->pos[d]) a->pos[d] += run.begin;
        if (a->mom[d]) a->mom[d] += run.begin;
        if (a->wpos[d]) a->wpos[d] += run.begin;
        if (a->wmom[d]) a->wmom[d] += run.begin;
    }
    a->N = run.count;
    a->p.DsT = run.DsT;
    a->p.stream = b.base.p.stream + (uint64_t)r;
}

// the fused step of k_cart_grid without deferral (Philox noise, the medium interpolated in tau
// per corner), the same device functions in the same order
// SLICED: the run reads the step's slice g.slice (the sliced-run kernel); otherwise the form is
// chosen per run at run time
// tc: the step's precomputed tau cell and fraction (k_batch_tcell), or null
template <int PDIM, int COLL, bool SLICED = false>
__device__ __forceinline__ void fused_grid_step(const StepArgs& a, const lmh_params& p, const GridMedium& g,
                                                const SpeciesArgs& sp, int64_t i, int64_t step,
                                                double* x0, double* x1, double* q,
                                                const double2* tc = nullptr) {
    constexpr bool HAS_KICK = COLL == LMH_LANGEVIN || COLL == LMH_LANGEVIN_RTA;
    constexpr bool HAS_RTA = COLL == LMH_RTA || COLL == LMH_LANGEVIN_RTA;
    const double m = sp.m;
    const double t_now = (double)step * p.dt + p.t0;
    double td = 0.0;
    int j = 0;
    if (!SLICED) {
        if (tc) {
            const double2 v = *tc;
            j = (int)v.x;
            td = v.y;
        } else {
            j = axis_cell(g.tg, g.nt, t_now, &td);
        }
    }
    double Tm, vx, vy;
    // the medium's reads are issued at wave priority 1, ahead of the other waves' arithmetic (the
    // batch kernel is issue bound too: 1.8 % per step on 256 runs): the cell's four nodes of the
    // step's node slice where the run has slices (g.slice, wave-uniform), else the corners
    // interpolated in tau
    __builtin_amdgcn_s_setprio(1);
    if (SLICED || g.slice) node_sample(g, *x0, *x1, &Tm, &vx, &vy);
    else grid_sample<false>(g, *x0, *x1, j, td, &Tm, &vx, &vy);
    __builtin_amdgcn_s_setprio(0);
    const FlowBoost fb = flow_boost(vx, vy);
    apply_boost<PDIM>(q, m, fb, -1, p.relativistic);
    if (PDIM == 3 && p.bjorken_redshift) bjorken(p, step, &q[2 % PDIM]);
    if (COLL == LMH_FLUID) {
        glue_to_flow<PDIM>(q, m, vx, vy, p.relativistic);
    } else {
        const double T = jl_max(Tm, 0.0);
        const bool read_drag = HAS_KICK && drag_table_read(p);
        double tn = 0.0, tau_d = 0.0;
        if (HAS_RTA && read_drag) {
            eval_time_table_pair(sp.rta, sp.drag, sp.shared_axis != 0, T, &tn, &tau_d);
        } else {
            if (HAS_RTA) tn = rta_lookup(sp.rta, T);
            if (read_drag) tau_d = eval_time_table(sp.drag, T);
        }
        double etaD = 0.0, kT = 0.0;
        if (HAS_KICK) drag_coeffs_tau<false>(p, tau_d, m, T, &etaD, &kT);
        double xi[PDIM];
        double ucol = 1.0, usamp_p = 0.0;
        noise_kick<PDIM, LMH_NOISE_PHILOX, HAS_KICK ? 1 : 0, HAS_RTA ? 1 : 0>(a, i, step, 0, xi, &ucol, &usamp_p);
        if (HAS_KICK) ou_kick<PDIM, false>(p, etaD, kT, T, m, a.sqrt_dt, q, xi);
        if (HAS_RTA && rta_decide<true>(p, tn, T, ucol)) {
            double usamp, gg[PDIM];
            noise_redraw<PDIM, LMH_NOISE_PHILOX>(a, i, 0, xi, usamp_p, &usamp, gg);
            rta_redraw<PDIM, false, true, true>(sp.jt, T, q, usamp, gg);
        }
        apply_boost<PDIM>(q, m, fb, +1, p.relativistic);
    }
    if (SLICED || g.slice) grid_push<PDIM, LMH_NOISE_PHILOX, 2>(a, p, g, i, step, 0, j, td, m, x0, x1, q);
    else grid_push<PDIM, LMH_NOISE_PHILOX, 0>(a, p, g, i, step, 0, j, td, m, x0, x1, q);
}

// PERSIST: a fixed grid of blocks walks the tiles (tile = blockIdx.x + k * gridDim.x);
// otherwise one block per tile.  SPEC as in k_cart_grid.  At least four waves per SIMD (128
// VGPRs, a few spilled bytes; left to itself the compiler takes 145 and three waves, 6 % slower
// on the 256-run workload; five waves spill more and are 9 % slower).
constexpr int kBatchWaves = 4;

template <int PDIM, int COLL, bool PERSIST, bool SPEC>
__global__ __launch_bounds__(kBlock, kBatchWaves) void k_batch_grid(BatchArgs b, int32_t k_off, int32_t nsteps) {
    const int64_t step0 = *b.step_dev + k_off;
    const int64_t stride = PERSIST ? (int64_t)gridDim.x : b.ntiles;
    for (int64_t t = blockIdx.x; t < b.ntiles; t += stride) {
        const int64_t tile = b.tile_list ? (int64_t)b.tile_list[t] : t;
        const int r = __builtin_amdgcn_readfirstlane(b.tile_run[tile]);
        const BatchRun run = uniform_read(&b.runs[r]);
        const int64_t i = (tile - run.tile0) * (int64_t)blockDim.x + threadIdx.x;
        if (i >= run.count) continue;
        StepArgs a;
        batch_view(b, run, r, &a);
        const lmh_params p = step_params<SPEC>(a.p);
        const SpeciesArgs sp = run.sp;
        // the launch after a sort (b.base.perm set, wave-uniform) reads slot i's quark from slot
        // src of the run's previous order and writes it, with its identity, to slot i of the
        // other buffers; every other launch reads and writes slot i in place
        const int64_t src = b.base.perm ? (int64_t)b.base.perm[run.begin + i] - run.begin : i;
        double x0 = a.pos[0][src], x1 = a.pos[1][src];
        double q[PDIM];
#pragma unroll
        for (int d = 0; d < PDIM; ++d) q[d] = a.mom[d][src];
        // the quark's identity keys its noise: its original slot in the run (the slot itself
        // unless the run is ordered by grid cell)
        const int64_t pid = b.base.id ? (int64_t)b.base.id[run.begin + src] - run.begin : src;
        int64_t bad = -1;
        GridMedium gs = run.g;
        const double2* tc = b.tcell ? b.tcell + (int64_t)r * b.tcell_k : nullptr;
        for (int s = 0; s < nsteps; ++s) {
            fused_grid_step<PDIM, COLL>(a, p, gs, sp, pid, step0 + s, &x0, &x1, q, tc ? tc + s : nullptr);
            if (gs.slice) gs.slice += run.slice_stride;
            if (bad < 0 && !(isfinite(x0) && isfinite(x1) && finite_all(q, PDIM))) bad = step0 + s;
        }
        a.wpos[0][i] = x0;
        a.wpos[1][i] = x1;
#pragma unroll
        for (int d = 0; d < PDIM; ++d) a.wmom[d][i] = q[d];
        if (b.base.perm) b.base.wid[run.begin + i] = (uint32_t)(pid + run.begin);
        if (bad >= 0) atomicMin(b.base.bad, (unsigned long long)bad);
    }
}

// the tau cell and fraction (axis_cell, as every step kernel finds them) of every run r < R at
// the launch's steps s = 0 .. nk - 1: out[r * nk + s]
__global__ __launch_bounds__(kBlock) void k_batch_tcell(BatchArgs b, const BatchRun* __restrict__ runs, int32_t R,
                                                        int32_t k_off, int32_t nk, double2* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= (int64_t)R * nk) return;
    const int r = (int)(k / nk), s = (int)(k % nk);
    const GridMedium& g = runs[r].g;
    const double t_now = (double)(*b.step_dev + k_off + s) * b.base.p.dt + b.base.p.t0;
    double td;
    const int j = axis_cell(g.tg, g.nt, t_now, &td);
    out[k] = make_double2((double)j, td);
}

// The node slices of the sliced runs for nk steps from the launch's first: blockIdx.y picks the
// sliced run srun[y], thread k of the x dimension the grid node k.  For each step s the node's
// (T, vx, vy) at the step's time, interpolated in tau between the two time layers of the step's
// tau cell (b.tcell) as the record slices interpolate them, into slot s of the run's ring.  A
// thread reloads the layers only where the tau cell changes from one step to the next, so a
// launch of nk steps reads each layer about once and writes 32 bytes per node and step.
__global__ __launch_bounds__(kBlock) void k_batch_nodes(BatchArgs b, const BatchRun* __restrict__ runs,
                                                        const int32_t* __restrict__ srun, int32_t nk) {
    const int r = srun[blockIdx.y];
    const BatchRun& run = runs[r];
    const GridMedium& g = run.g;
    const int64_t nxy = (int64_t)g.nx * g.ny;
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= nxy) return;   // a smaller grid than the launch's largest
    const double2* tc = b.tcell + (int64_t)r * b.tcell_k;
    double4* dst = const_cast<double4*>(g.slice) + k;
    int jl = -1;
    double4 n0 = make_double4(0.0, 0.0, 0.0, 0.0), n1 = n0;
    for (int s = 0; s < nk; ++s) {
        const double2 jt = tc[s];
        const int j = (int)jt.x;
        const double td = jt.y;
        if (j != jl) {
            n0 = ld_global4(g.nodes + nxy * j + k);
            n1 = ld_global4(g.nodes + nxy * (j + 1) + k);
            jl = j;
        }
        const double u = 1.0 - td;
        dst[s * run.slice_stride] = make_double4(n0.x * u + n1.x * td, n0.y * u + n1.y * td,
                                                 n0.z * u + n1.z * td, 0.0);
    }
}

// The sliced runs' step in the specialised configuration, one step per launch: block b takes the
// tile tiles[b] (a tile of a run of at least slice_min quarks), whose run reads the step's node
// slice, slot `slot` of its ring.  fused_grid_step on the slice, at the single-run sliced
// kernel's occupancy; the other runs take k_batch_grid.
// The run table comes in as a read-only argument of its own, so that the block-uniform record is
// read with scalar loads.
template <int PDIM, int COLL>
__global__ __launch_bounds__(kBlock, kGridWaves) void k_batch_grid_sl(BatchArgs b, const BatchRun* __restrict__ runs,
                                                                      const int32_t* __restrict__ tiles,
                                                                      const int32_t* __restrict__ tile_run,
                                                                      int32_t k_off, int32_t slot) {
    const int64_t step = *b.step_dev + k_off;
    const int64_t tile = tiles[blockIdx.x];
    const int r = tile_run[tile];
    const BatchRun& run = runs[r];
    const int64_t i = (tile - run.tile0) * (int64_t)blockDim.x + threadIdx.x;
    if (i >= run.count) return;
    StepArgs a;
    batch_view(b, run, r, &a);
    const lmh_params p = step_params<true>(a.p);
    // read through the sort's permutation after a sort, as in k_batch_grid
    const int64_t src = b.base.perm ? (int64_t)b.base.perm[run.begin + i] - run.begin : i;
    double x0 = a.pos[0][src], x1 = a.pos[1][src];
    double q[PDIM];
#pragma unroll
    for (int d = 0; d < PDIM; ++d) q[d] = a.mom[d][src];
    const int64_t pid = b.base.id ? (int64_t)b.base.id[run.begin + src] - run.begin : src;
    GridMedium g = run.g;
    g.slice += slot * run.slice_stride;
    fused_grid_step<PDIM, COLL, true>(a, p, g, run.sp, pid, step, &x0, &x1, q);
    a.wpos[0][i] = x0;
    a.wpos[1][i] = x1;
#pragma unroll
    for (int d = 0; d < PDIM; ++d) a.wmom[d][i] = q[d];
    if (b.base.perm) b.base.wid[run.begin + i] = (uint32_t)(pid + run.begin);
    if (!(isfinite(x0) && isfinite(x1) && finite_all(q, PDIM))) atomicMin(b.base.bad, (unsigned long long)step);
}

// the cell-order keys of every run: key = run << cell_bits | cell (the cell grid_sample reads on
// the run's grid), value = slot; a stable sort of them orders each run's quarks by cell and keeps
// the runs' segments in place
__global__ __launch_bounds__(kBlock) void k_batch_sort_keys(BatchArgs b, int32_t cell_bits, uint32_t* keys,
                                                            uint32_t* vals) {
    for (int64_t tile = blockIdx.x; tile < b.ntiles; tile += gridDim.x) {
        const int r = __builtin_amdgcn_readfirstlane(b.tile_run[tile]);
        const BatchRun run = uniform_read(&b.runs[r]);
        const int64_t i = (tile - run.tile0) * (int64_t)blockDim.x + threadIdx.x;
        if (i >= run.count) continue;
        const int64_t s = run.begin + i;
        const GridMedium& g = run.g;
        const double xc = fmin(fmax(b.base.pos[0][s], g.x0), g.xlast);
        const double yc = fmin(fmax(b.base.pos[1][s], g.y0), g.ylast);
        const int ix = grid_cell((xc - g.x0) * g.inv_dx, g.nx);
        const int iy = grid_cell((yc - g.y0) * g.inv_dy, g.ny);
        keys[s] = ((uint32_t)r << cell_bits) | (uint32_t)(ix + (g.nx - 1) * iy);
        vals[s] = (uint32_t)s;
    }
}

__global__ void k_batch_advance(int64_t* step_dev, int64_t n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step_dev += n;
}

template <int PDIM>
__global__ __launch_bounds__(kBlock) void k_batch_boost_lab(BatchArgs b, int64_t step) {
    for (int64_t tile = blockIdx.x; tile < b.ntiles; tile += gridDim.x) {
        const int r = __builtin_amdgcn_readfirstlane(b.tile_run[tile]);
        const BatchRun run = uniform_read(&b.runs[r]);
        const int64_t i = (tile - run.tile0) * (int64_t)blockDim.x + threadIdx.x;
        if (i >= run.count) continue;
        StepArgs a;
        batch_view(b, run, r, &a);
        double q[PDIM];
#pragma unroll
        for (int d = 0; d < PDIM; ++d) q[d] = a.mom[d][i];
        const double t_now = (double)step * a.p.dt + a.p.t0;
        double td;
        const int j = axis_cell(run.g.tg, run.g.nt, t_now, &td);
        double Tm, vx, vy;
        grid_sample<false>(run.g, a.pos[0][i], a.pos[1][i], j, td, &Tm, &vx, &vy);
        apply_boost<PDIM>(q, run.sp.m, flow_boost(vx, vy), +1, a.p.relativistic);
#pragma unroll
        for (int d = 0; d < PDIM; ++d) a.mom[d][i] = q[d];
    }
}

// per-run observables: blockIdx.y is the run, k_obs_partial's sums over its segment
__global__ __launch_bounds__(kObsBlock) void k_batch_obs_partial(const double* __restrict__ px,
                                                                  const double* __restrict__ py,
                                                                  const int64_t* __restrict__ begin,
                                                                  lmh_obs_spec s,
                                                                  double* __restrict__ partial) {
    const int r = blockIdx.y;
    obs_block(px, py, begin[r], begin[r + 1], s, blockIdx.x, gridDim.x,
              partial + ((int64_t)r * gridDim.x + blockIdx.x) * (5 + s.nbins));
}

// one block per run: out[r*K + k] = sum over the run's partials, in block order
__global__ __launch_bounds__(kObsBlock) void k_batch_obs_final(const double* __restrict__ partial, int nblk,
                                                                int K, double* __restrict__ out) {
    const int r = blockIdx.x;
    for (int k = threadIdx.x; k < K; k += kObsBlock) {
        double s = 0.0;
        for (int b = 0; b < nblk; ++b) s += partial[((int64_t)r * nblk + b) * K + k];
        out[(int64_t)r * K + k] = s;
    }
}
