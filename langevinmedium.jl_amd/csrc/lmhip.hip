// lmhip.hip -- fused FP64 Langevin step kernels for gfx950 and the C ABI of liblmhip.so.
//
// One thread owns one particle.  The state lives in HBM as structure-of-arrays rows (x, y[, z],
// px, py[, pz]), read once and written once per launch; a launch may keep a particle in
// registers for several steps.  The reference runs each step as 4-6 separate passes over all
// particles (src/simulate_cpu.jl:215-303, src/simulate_gpu.jl:283-392); here the whole step is
// one pass:
//     medium sample -> boost to the LRF -> [Bjorken redshift] -> Langevin kick | BGK redraw |
//     glue to the flow -> boost back -> position push -> non-finite check.
// Noise is either a fixed record read from HBM (for per-particle parity with the CPU kernels)
// or Philox4x32-10 generated in registers.
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <optional>
#include <new>
#include <string>
#include <vector>

#include "lmhip_device.h"

using namespace lmh;

// Scratch of lmh_observables_noalloc: the per-block partial sums of one observables launch
// (1024 blocks x at most 5 + 120 values), a file-scope device array instead of an allocation.
// Each GPU holds its own copy until the process ends; host code reaches it only through the
// address hipGetSymbolAddress returns on that device (the array's name is a host-side stand-in).
constexpr size_t kScratchDoubles = 1024 * (5 + 120);
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kBlock = 256;

struct SpeciesArgs {
    double m;
    TimeTable drag;
    TimeTable rta;
    JuttnerTable jt;
    int32_t shared_axis;  // 1: drag and rta lie on one temperature axis (eval_time_table_pair)
};

// whether a species' drag and RTA tables share their temperature axis
inline int32_t tables_share_axis(const TimeTable& a, const TimeTable& b) {
    return a.vals && b.vals && a.Tmin == b.Tmin && a.invdT == b.invdT && a.n == b.n ? 1 : 0;
}

// Species are contiguous particle segments; one launch covers all of them: block b belongs to
// the segment whose first block is the largest block0[k] <= b.
struct StepArgs {
    double* pos[3];
    double* mom[3];
    double* wpos[3];            // the rows k_cart_grid writes: pos and mom themselves (in place),
    double* wmom[3];            // or the next row of an adjoint segment's tape
    int64_t N;                  // particles (the noise row stride)
    int64_t step_begin;
    int32_t nsteps;
    int32_t C, oxi, orta, opos; // fixed-noise record layout
    int32_t nseg;
    int64_t begin[LMH_MAX_SPECIES + 1];
    int64_t block0[LMH_MAX_SPECIES + 1];
    const double* noise;
    unsigned long long* bad;    // first non-finite step (atomicMin), ULLONG_MAX if none
    SpeciesArgs species[LMH_MAX_SPECIES];  // per-species mass and tables, read with scalar
                                           // loads from the kernel arguments (block-uniform index)
    const SpeciesArgs* species_dev;        // the same in device memory, for per-lane indices
    const uint32_t* id;         // each slot's particle identity (its noise counters): the slot
                                // itself unless the particles are ordered by grid cell
    const uint32_t* perm;       // GATHER launches: slot i's particle is read from slot perm[i] of
    uint32_t* wid;              // pos/mom/id and written to slot i of wpos/wmom/wid
    double4* qrec;              // deferred collisions: block b's records (two double4 each:
                                // x, y, T, vx | vy, identity, slot, 0) at [b*qblock, ...)
    int32_t* qcount;            // deferred collisions per block of the step kernel
    int32_t qblock;             // the step kernel's block size
    lmh_params p;
    double sqrt_dt;
    // on-the-fly freeze-out capture (kernels_gpu.jl:583-635): the latched state of each particle
    // (rows as the state), its time, a flag, and T after the particle's previous step
    struct {
        int32_t on, interp;
        double Tfo;
        double* pos;
        double* mom;
        double* tau;
        double* flag;
        double* lastT;
    } fo;
};

// |v| <= sqrt(1 - 1e-12), kernels_cpu.jl:102
__device__ __forceinline__ double vmax_() { return sqrt(1.0 - 1e-12); }

// The freeze-out latch after a step (kernels_gpu.jl:583-635): T is the medium at the pushed
// position and the step's time t.  Above Tfo, T is remembered; the first step below it books the
// state interpolated linearly in T between the step's input state (the last state above Tfo)
// and its output (frac = 1, the raw output, without interpolation or without a previous T above
// Tfo).  tprev is the previous step's time.
template <int DIMS, int PDIM>
__device__ __forceinline__ void freeze_latch(const StepArgs& a, int64_t i, double T, double t,
                                             double tprev, const double* xin, const double* qin,
                                             const double* x, const double* q, bool* frozen,
                                             double* lastT) {
    const double Tfo = a.fo.Tfo;
    if (!(T < Tfo)) {
        *lastT = T;
        return;
    }
    const double Tl = *lastT;
    const double frac = (a.fo.interp && Tl > Tfo && Tl > T) ? (Tl - Tfo) / (Tl - T) : 1.0;
    const int64_t N = a.N;
#pragma unroll
    for (int d = 0; d < DIMS; ++d) a.fo.pos[(int64_t)d * N + i] = xin[d] + frac * (x[d] - xin[d]);
#pragma unroll
    for (int d = 0; d < PDIM; ++d) a.fo.mom[(int64_t)d * N + i] = qin[d] + frac * (q[d] - qin[d]);
    a.fo.tau[i] = tprev + frac * (t - tprev);
    a.fo.flag[i] = 1.0;
    *frozen = true;
}

// The common configuration -- relativistic, pre-point kick, drag table, no Bjorken redshift, no
// position diffusion -- as compile-time constants: with SPEC the option branches it does not
// take drop out of a kernel, which then needs far fewer registers.  The values are the same.
template <bool SPEC>
__device__ __forceinline__ lmh_params step_params(const lmh_params& in) {
    lmh_params p = in;
    if (SPEC) {
        p.relativistic = 1;
        p.integrator = LMH_PREPOINT;
        p.drag_model = LMH_DRAG_TABLE;
        p.bjorken_redshift = 0;
        p.position_diffusion = 0;
    }
    return p;
}

// this block's species segment and this thread's particle (-1 past the segment's end); the
// segment and its first particle are block-uniform (32-bit block numbers: scalar compares)
__device__ __forceinline__ int64_t particle_of_thread(const StepArgs& a, int* seg) {
    const uint32_t b = blockIdx.x;
    int k = 0;
#pragma unroll
    for (int j = 1; j < LMH_MAX_SPECIES; ++j)
        if (j < a.nseg && b >= (uint32_t)a.block0[j]) k = j;
    *seg = k;
    const int64_t first = a.begin[k] + (int64_t)(b - (uint32_t)a.block0[k]) * blockDim.x;
    const int64_t i = first + threadIdx.x;
    return i < a.begin[k + 1] ? i : -1;
}

struct RadialMedium {
    const double* xg;
    const double* tg;
    const double* T;
    const double* V;
    const double* V2;
    int32_t nx, nt;
    double r_axis_eps, r_max;   // kernels_cpu.jl:528-534
    double xinv, tinv;          // (n - 1)/(g[n-1] - g[0]) of an equally spaced grid, else 0
};

struct GridMedium {
    const double4* nodes;       // (T, vx, vy, 0) at ix + nx*(iy + ny*it)
    const double4* slice;       // the medium at one time: per cell ix + (nx-1)*iy, the bilinear
                                // coefficients (a, b, c, d) of T, vx, vy as three records
    const double* tg;
    int32_t nx, ny, nt;
    double x0, inv_dx, xlast;
    double y0, inv_dy, ylast;
};

// ---------------------------------------------------------------------------------------------
// Noise of one particle-step.  FIXED reads the record layout of include/lmhip.h.  PHILOX draws
// block 0 once: the kick normals, the collision uniform and the redraw's quantile; the redraw's
// direction is the kick normals' direction (a colliding particle's kick is replaced by the
// redraw, so its normals serve it); block 2 holds the position-diffusion normals.
// KICK, RTA: whether the step draws the kick normals and the collision uniform, when the kernel
// knows it at compile time (-1: from the record layout at run time).  usamp receives the
// redraw's quantile in PHILOX mode (FIXED reads it with the direction in noise_redraw).
template <int PDIM, int NOISE, int KICK = -1, int RTA = -1>
__device__ __forceinline__ void noise_kick(const StepArgs& a, int64_t i, int64_t step, int s_off,
                                           double* xi, double* ucol, double* usamp) {
    const bool kick = KICK < 0 ? a.oxi >= 0 : KICK != 0;
    const bool rta = RTA < 0 ? a.orta >= 0 : RTA != 0;
    if (NOISE == LMH_NOISE_FIXED) {
        const double* rec = a.noise + (int64_t)s_off * a.C * a.N + i;
        if (kick) {
#pragma unroll
            for (int d = 0; d < PDIM; ++d) xi[d] = rec[(int64_t)(a.oxi + d) * a.N];
        }
        if (rta) *ucol = rec[(int64_t)a.orta * a.N];
    } else {
        if (!kick && !rta) return;
        uint64_t w[4];
        noise_block(a.p.seed, a.p.stream, i, step, 0, w);
        kick_normals<PDIM>(w, xi);
        if (rta) {
            *ucol = u_low31(w[3]);
            *usamp = u_hi32(w[3]);
        }
    }
}

// The redraw's quantile and direction: FIXED reads them from the record; PHILOX takes the
// quantile noise_kick drew and the kick normals xi
template <int PDIM, int NOISE>
__device__ __forceinline__ void noise_redraw(const StepArgs& a, int64_t i, int s_off, const double* xi,
                                             double usamp_philox, double* usamp, double* g) {
    if (NOISE == LMH_NOISE_FIXED) {
        const double* rec = a.noise + (int64_t)s_off * a.C * a.N + i;
        *usamp = rec[(int64_t)(a.orta + 1) * a.N];
#pragma unroll
        for (int d = 0; d < PDIM; ++d) g[d] = rec[(int64_t)(a.orta + 2 + d) * a.N];
    } else {
        *usamp = usamp_philox;
#pragma unroll
        for (int d = 0; d < PDIM; ++d) g[d] = xi[d];
    }
}

template <int DIMS, int NOISE>
__device__ __forceinline__ void noise_diffusion(const StepArgs& a, int64_t i, int64_t step, int s_off,
                                                double* xp) {
    if (NOISE == LMH_NOISE_FIXED) {
        const double* rec = a.noise + (int64_t)s_off * a.C * a.N + i;
#pragma unroll
        for (int d = 0; d < DIMS; ++d) xp[d] = rec[(int64_t)(a.opos + d) * a.N];
    } else {
        uint64_t w[4];
        noise_block(a.p.seed, a.p.stream, i, step, 2, w);
        double n0, n1;
        box_muller<true>(w[0], w[1], &n0, &n1);
        xp[0] = n0;
        if (DIMS >= 2) xp[1 % DIMS] = n1;
        if (DIMS >= 3) {
            double n2, n3;
            box_muller<true>(w[2], w[3], &n2, &n3);
            xp[2 % DIMS] = n2;
        }
    }
}

// kernel_boost_to_rest_frame_cpu! (sign -1, kernels_cpu.jl:89-145) and
// kernel_boost_to_lab_frame_cpu! (sign +1, :188-241) for one particle with r >= eps
template <int DIMS, int PDIM, bool RADIAL>
__device__ __forceinline__ void radial_boost(double* q, const double* x, double r, double v,
                                             double m, int sign, int relativistic, double vmax) {
    if (!relativistic) {
        if (RADIAL) {
            q[0] = sign > 0 ? q[0] + m * v : q[0] - m * v;
        } else {
            double s0 = m * v * (x[0] / r);
            double s1 = m * v * (x[1 % DIMS] / r);
            q[0] = sign > 0 ? q[0] + s0 : q[0] - s0;
            q[1 % PDIM] = sign > 0 ? q[1 % PDIM] + s1 : q[1 % PDIM] - s1;
        }
        return;
    }
    v = jl_clamp(v, -vmax, vmax);
    double v2 = v * v;
    double gam = 1.0 / sqrt(1.0 - v2 + 1e-10);
    double p2 = 0.0;
#pragma unroll
    for (int d = 0; d < PDIM; ++d) p2 += q[d] * q[d];
    double E = sqrt(p2 + m * m);
    if (RADIAL) {
        q[0] = sign > 0 ? gam * (q[0] + v * E) : gam * (q[0] - v * E);
    } else {
        double rhatx = x[0] / r, rhaty = x[1 % DIMS] / r;
        double px = q[0], py = q[1 % PDIM];
        double p_par = px * rhatx + py * rhaty;
        double p_perp_x = px - p_par * rhatx;
        double p_perp_y = py - p_par * rhaty;
        double p_par_new = sign > 0 ? gam * (p_par + v * E) : gam * (p_par - v * E);
        q[0] = p_par_new * rhatx + p_perp_x;
        q[1 % PDIM] = p_par_new * rhaty + p_perp_y;
    }
}

__device__ __forceinline__ void bjorken(const lmh_params& p, int64_t step, double* q2) {
    // kernel_bjorken_redshift_cpu!, kernels_cpu.jl:427-435
    double tau_a = p.t0 + (double)(step - 1) * p.dt;
    if (tau_a > 0.0) {
        double f = tau_a / (tau_a + p.dt);
        *q2 *= f;
    }
}

// The space-time rapidity over the step tau_a -> tau_a + dt of a quark whose comoving
// longitudinal momentum at the end of the step is p_z (Milne coordinates): p_z tau is conserved
// along a free path and y - eta_s = asinh(p_z/m_T) (the note at kernels_cpu.jl:419-426), so
// eta_s advances by asinh(k b) - asinh(b) with b = p_z/m_T, k = 1 + dt/tau_a, written without
// cancellation as asinh(b (k - 1)(k + 1) / (k sqrt(1 + b^2) + sqrt(1 + (k b)^2))).
__device__ __forceinline__ double milne_deta(double px, double py, double pz, double m, double tau_a,
                                             double dt) {
    if (!(tau_a > 0.0)) return 0.0;
    const double mT = sqrt(m * m + px * px + py * py);
    const double b = pz / mT;
    const double km1 = dt / tau_a;
    const double k = 1.0 + km1;
    const double kb = k * b;
    const double num = b * (km1 * (k + 1.0));
    const double den = k * sqrt(1.0 + b * b) + sqrt(1.0 + kb * kb);
    return asinh(num / den);
}

// milne_deta with refined reciprocal and square-root estimates (an ulp or two from the divisions
// and square roots) and asinh by its Taylor series for |z| <= 1/16 (through z^13: truncation
// below 2e-19 relative; z is the step's rapidity-gap change, a few hundredths at most for the
// steps of the workloads), libm asinh above
__device__ __forceinline__ double milne_deta_fast(double px, double py, double pz, double m, double tau_a,
                                                  double dt) {
    if (!(tau_a > 0.0)) return 0.0;
    const double mT2 = m * m + px * px + py * py;
    const double b = pz * rsqrt_pos(mT2);
    const double km1 = dt / tau_a;
    const double k = 1.0 + km1;
    const double kb = k * b;
    const double num = b * (km1 * (k + 1.0));
    const double den = k * sqrt_pos<true>(1.0 + b * b) + sqrt_pos<true>(1.0 + kb * kb);
    const double z = num * rcp_pos(den);
    if (!(fabs(z) <= 0.0625)) return asinh(z);
    const double z2 = z * z;
    double s = -231.0 / 13312.0;
    s = fma(s, z2, 63.0 / 2816.0);
    s = fma(s, z2, -35.0 / 1152.0);
    s = fma(s, z2, 5.0 / 112.0);
    s = fma(s, z2, -3.0 / 40.0);
    s = fma(s, z2, 1.0 / 6.0);
    return fma(-(z * z2), s, z);
}

// radial_boost's relativistic 2+1D form with the flow factors of the step given (gam = 1/sqrt(1 -
// v^2 + 1e-10) and the unit vector (rhx, rhy) = x/r), E from the refined square root
__device__ __forceinline__ void radial_boost_fast(double* q, double rhx, double rhy, double v, double gam,
                                                  double m, int sign) {
    double p2 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) p2 += q[d] * q[d];
    const double E = sqrt_pos<true>(p2 + m * m);
    const double px = q[0], py = q[1];
    const double p_par = px * rhx + py * rhy;
    const double p_perp_x = px - p_par * rhx;
    const double p_perp_y = py - p_par * rhy;
    const double p_par_new = sign > 0 ? gam * (p_par + v * E) : gam * (p_par - v * E);
    q[0] = p_par_new * rhx + p_perp_x;
    q[1] = p_par_new * rhy + p_perp_y;
}

__device__ __forceinline__ bool finite_all(const double* a, int n) {
    bool ok = true;
    for (int d = 0; d < n; ++d) ok = ok && isfinite(a[d]);
    return ok;
}

// ---------------------------------------------------------------------------------------------
// Cartesian step on the radial (r, t) medium: the reference's own medium (simulate_cpu.jl).
template <int DIMS, int PDIM, int NOISE, int COLL, bool SPEC>
__global__ __launch_bounds__(kBlock) void k_cart_radial(StepArgs a, RadialMedium med) {
    int seg;
    const int64_t i = particle_of_thread(a, &seg);
    if (i < 0) return;
    const SpeciesArgs sp = a.species[seg];
    constexpr bool RADIAL = DIMS == 1;
    const lmh_params pp = step_params<SPEC>(a.p);
    const lmh_params& p = pp;
    const double m = sp.m;
    // Milne coordinates: rows (x, y, eta_s); the medium at the transverse radius, the Bjorken
    // redshift of p_z, then the push of eta_s
    const bool milne = DIMS == 3 && PDIM == 3 && !SPEC && p.coords == LMH_MILNE;
    double x[DIMS], q[PDIM];
#pragma unroll
    for (int d = 0; d < DIMS; ++d) x[d] = a.pos[d][i];
#pragma unroll
    for (int d = 0; d < PDIM; ++d) q[d] = a.mom[d][i];
    // freeze-out capture runs in the generic kernels only
    const bool fo = !SPEC && a.fo.on;
    bool frozen = false;
    double lastT = 0.0;
    if (fo) {
        frozen = a.fo.flag[i] != 0.0;
        lastT = a.fo.lastT[i];
    }
    int64_t bad = -1;
    for (int s_off = 0; s_off < a.nsteps; ++s_off) {
        const int64_t step = a.step_begin + s_off;
        const double t_now = (double)step * p.dt + p.t0;
        double yd;
        const int j = axis_cell_guess(med.tg, med.nt, med.tinv, t_now, &yd);
        double xin[DIMS], qin[PDIM];
        if (fo && !frozen) {
#pragma unroll
            for (int d = 0; d < DIMS; ++d) xin[d] = x[d];
#pragma unroll
            for (int d = 0; d < PDIM; ++d) qin[d] = q[d];
        }
        double xi[PDIM];
        double ucol = 1.0, usamp_p = 0.0;
        noise_kick<PDIM, NOISE>(a, i, step, s_off, xi, &ucol, &usamp_p);

        double r2 = 0.0;
#pragma unroll
        for (int d = 0; d < DIMS; ++d)
            if (!(milne && d == 2)) r2 += x[d] * x[d];
        const double r = sqrt(r2);
        double xd;
        const int ir = axis_cell_guess(med.xg, med.nx, med.xinv, r, &xd);
        const bool skip = r < kEps;
        double v = 0.0;
        if (!skip) {
            v = bilinear(med.V, med.nx, ir, j, xd, yd);
            if (DIMS >= 2 && med.V2 != nullptr) {
                double v2m = bilinear(med.V2, med.nx, ir, j, xd, yd);
                double phi = atan2(x[1 % DIMS], x[0]);
                v *= (1.0 + 2.0 * v2m * cos(2.0 * (phi - p.psi2)));
            }
            radial_boost<DIMS, PDIM, RADIAL>(q, x, r, v, m, -1, p.relativistic, vmax_());
        }
        if (PDIM == 3 && (p.bjorken_redshift || milne)) bjorken(p, step, &q[2 % PDIM]);

        if (COLL == LMH_FLUID) {
            // kernel_set_to_fluid_velocity_cpu!, kernels_cpu.jl:683-733: r from rows 1-2
            double rf = RADIAL ? x[0] : sqrt(x[0] * x[0] + x[1 % DIMS] * x[1 % DIMS]);
            if (!(rf < kEps)) {
                double xdf;
                int irf = axis_cell_guess(med.xg, med.nx, med.xinv, rf, &xdf);
                double vf = bilinear(med.V, med.nx, irf, j, xdf, yd);
                if (!p.relativistic) {
                    if (RADIAL) {
                        q[0] = m * vf;
                    } else {
                        q[0] = m * vf * (x[0] / rf);
                        q[1 % PDIM] = m * vf * (x[1 % DIMS] / rf);
                    }
                } else {
                    vf = jl_clamp(vf, -vmax_(), vmax_());
                    double gam = 1.0 / sqrt(1.0 - vf * vf + 1e-10);
                    if (RADIAL) {
                        q[0] = m * gam * vf;
                    } else {
                        q[0] = m * gam * vf * (x[0] / rf);
                        q[1 % PDIM] = m * gam * vf * (x[1 % DIMS] / rf);
                    }
                }
            }
        } else {
            const double T = jl_max(bilinear(med.T, med.nx, ir, j, xd, yd), 0.0);
            if (COLL == LMH_LANGEVIN || COLL == LMH_LANGEVIN_RTA) {
                double etaD, kT;
                drag_coeffs<RADIAL>(p, sp.drag, m, T, &etaD, &kT);
                ou_kick<PDIM, RADIAL>(p, etaD, kT, T, m, a.sqrt_dt, q, xi);
            }
            if ((COLL == LMH_RTA || COLL == LMH_LANGEVIN_RTA) && rta_collides(p, sp.rta, T, ucol)) {
                double usamp, g[PDIM];
                noise_redraw<PDIM, NOISE>(a, i, s_off, xi, usamp_p, &usamp, g);
                rta_redraw<PDIM, RADIAL>(sp.jt, T, q, usamp, g);
            }
            if (!skip) radial_boost<DIMS, PDIM, RADIAL>(q, x, r, v, m, +1, p.relativistic, vmax_());
        }

        // kernel_update_positions_cpu!, kernels_cpu.jl:535-636
        double E2 = m * m;
#pragma unroll
        for (int d = 0; d < PDIM; ++d) E2 += q[d] * q[d];
        const double E = p.relativistic ? sqrt(E2) : m;
        if (RADIAL) {
            double r0 = x[0];
            double pr = q[0];
            double r_abs = fabs(r0);
            double r_safe = (r_abs < med.r_axis_eps) ? med.r_axis_eps : r_abs;
            double dr = (pr / E) * p.dt;
            if (p.position_diffusion) {
                double xdd;
                int ird = axis_cell_guess(med.xg, med.nx, med.xinv, r_safe, &xdd);
                double Td = bilinear(med.T, med.nx, ird, j, xdd, yd);
                double D = (p.DsT / jl_max(Td, kEps)) / kFmGeV;
                if (D > 0.0) {
                    double xp[DIMS];
                    noise_diffusion<DIMS, NOISE>(a, i, step, s_off, xp);
                    dr += (D / r_safe) * p.dt + sqrt(2.0 * D * p.dt) * xp[0];
                }
            }
            double r_new = r0 + dr;
            if (r_new < 0.0) {
                r_new = -r_new;
                q[0] = -q[0];
            }
            if (p.reflecting_boundary && med.r_max > 0.0 && r_new > med.r_max) {
                r_new = 2.0 * med.r_max - r_new;
                q[0] = -q[0];
                if (r_new < 0.0) {
                    r_new = -r_new;
                    q[0] = -q[0];
                }
            }
            x[0] = r_new;
        } else {
#pragma unroll
            for (int d = 0; d < DIMS; ++d)
                if (!(milne && d == 2)) x[d] = x[d] + p.dt * q[d % PDIM] / E;
            if (milne) x[2 % DIMS] += milne_deta(q[0], q[1 % PDIM], q[2 % PDIM], m,
                                                 p.t0 + (double)(step - 1) * p.dt, p.dt);
            if (p.position_diffusion) {
                double rr = 0.0;
#pragma unroll
                for (int d = 0; d < DIMS; ++d) rr += x[d] * x[d];
                double r_now = sqrt(rr);
                double xdd;
                int ird = axis_cell_guess(med.xg, med.nx, med.xinv, r_now, &xdd);
                double Td = bilinear(med.T, med.nx, ird, j, xdd, yd);
                double D = (p.DsT / jl_max(Td, kEps)) / kFmGeV;
                if (D > 0.0) {
                    double xp[DIMS];
                    noise_diffusion<DIMS, NOISE>(a, i, step, s_off, xp);
                    double sigma = sqrt(2.0 * D * p.dt);
#pragma unroll
                    for (int d = 0; d < DIMS; ++d) x[d] = x[d] + sigma * xp[d];
                }
            }
            if (p.reflecting_boundary && med.r_max > 0.0) {
                double rr = 0.0;
#pragma unroll
                for (int d = 0; d < DIMS; ++d) rr += x[d] * x[d];
                double r_now = sqrt(rr);
                if (r_now > med.r_max) {
                    double invr = 1.0 / jl_max(r_now, med.r_axis_eps);
                    double ppar = 0.0;
#pragma unroll
                    for (int d = 0; d < DIMS; ++d) ppar += q[d % PDIM] * (x[d] * invr);
#pragma unroll
                    for (int d = 0; d < DIMS; ++d) q[d % PDIM] = q[d % PDIM] - 2.0 * ppar * (x[d] * invr);
                    double scale = (2.0 * med.r_max - r_now) * invr;
#pragma unroll
                    for (int d = 0; d < DIMS; ++d) x[d] *= scale;
                }
            }
        }
        if (fo && !frozen) {
            // T at the pushed position: r from all position rows (|x| in radial mode)
            double rr = 0.0;
#pragma unroll
            for (int d = 0; d < DIMS; ++d) rr += x[d] * x[d];
            const double rf = RADIAL ? fabs(x[0]) : sqrt(rr);
            double xdf;
            const int irf = axis_cell_guess(med.xg, med.nx, med.xinv, rf, &xdf);
            const double Tf = bilinear(med.T, med.nx, irf, j, xdf, yd);
            freeze_latch<DIMS, PDIM>(a, i, Tf, t_now, (double)(step - 1) * p.dt + p.t0, xin, qin, x, q,
                                     &frozen, &lastT);
        }
        if (bad < 0 && !(finite_all(x, DIMS) && finite_all(q, PDIM))) bad = step;
    }
#pragma unroll
    for (int d = 0; d < DIMS; ++d) a.pos[d][i] = x[d];
#pragma unroll
    for (int d = 0; d < PDIM; ++d) a.mom[d][i] = q[d];
    if (fo) a.fo.lastT[i] = lastT;
    if (bad >= 0) atomicMin(a.bad, (unsigned long long)bad);
}

// The Milne step (rows x, y, eta_s; p_x, p_y, p_z) in its common configurations, with the options
// it does not take compiled out: the relativistic kick with the integrator INTEG and the drag
// model DRAG, no position diffusion, no reflecting wall, no freeze-out capture, no elliptic flow
// term.  The operations of k_cart_radial's Milne path in the same order, with refined reciprocal
// and square-root estimates in place of the divisions and square roots of the boosts' energy and
// direction, the kick's GAMMA_P denominator, the push and the eta_s step (within a few ulps of the
// generic kernel's values); the medium lookup keeps its divisions, so cells and fractions are the
// reference's, and so does the boosts' gamma.  At least four waves per SIMD (128 VGPRs, no
// spills; left to itself the compiler takes 135 and three waves, 8 % slower).
template <int NOISE, int COLL, int INTEG, int DRAG>
__global__ __launch_bounds__(kBlock, 4) void k_milne(StepArgs a, RadialMedium med) {
    int seg;
    const int64_t i = particle_of_thread(a, &seg);
    if (i < 0) return;
    const SpeciesArgs sp = a.species[seg];
    lmh_params pp = a.p;
    pp.relativistic = 1;
    pp.integrator = INTEG;
    pp.drag_model = DRAG;
    pp.position_diffusion = 0;
    pp.reflecting_boundary = 0;
    const lmh_params& p = pp;
    const double m = sp.m;
    double x[3], q[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = a.pos[d][i];
#pragma unroll
    for (int d = 0; d < 3; ++d) q[d] = a.mom[d][i];
    int64_t bad = -1;
    for (int s_off = 0; s_off < a.nsteps; ++s_off) {
        const int64_t step = a.step_begin + s_off;
        const double t_now = (double)step * p.dt + p.t0;
        double yd;
        const int j = axis_cell_guess(med.tg, med.nt, med.tinv, t_now, &yd);
        double xi[3];
        double ucol = 1.0, usamp_p = 0.0;
        noise_kick<3, NOISE>(a, i, step, s_off, xi, &ucol, &usamp_p);
        // the medium at the transverse radius
        double r2 = 0.0;
        r2 += x[0] * x[0];
        r2 += x[1] * x[1];
        const double r = sqrt(r2);
        double xd;
        const int ir = axis_cell_guess(med.xg, med.nx, med.xinv, r, &xd);
        const bool skip = r < kEps;
        double v = 0.0, gam = 1.0, rhx = 0.0, rhy = 0.0;
        if (!skip) {
            v = jl_clamp(bilinear(med.V, med.nx, ir, j, xd, yd), -vmax_(), vmax_());
            // gam as the reference rounds it: with the fluid at rest it is the same factor every
            // step, so an ulp of a refined estimate would accumulate over the run
            gam = 1.0 / sqrt(1.0 - v * v + 1e-10);
            const double ir_ = rcp_pos(r);
            rhx = x[0] * ir_;
            rhy = x[1] * ir_;
            radial_boost_fast(q, rhx, rhy, v, gam, m, -1);
        }
        bjorken(p, step, &q[2]);
        if (COLL == LMH_FLUID) {
            double rf = sqrt(x[0] * x[0] + x[1] * x[1]);
            if (!(rf < kEps)) {
                double xdf;
                int irf = axis_cell_guess(med.xg, med.nx, med.xinv, rf, &xdf);
                double vf = bilinear(med.V, med.nx, irf, j, xdf, yd);
                vf = jl_clamp(vf, -vmax_(), vmax_());
                double gam = 1.0 / sqrt(1.0 - vf * vf + 1e-10);
                q[0] = m * gam * vf * (x[0] / rf);
                q[1] = m * gam * vf * (x[1] / rf);
            }
        } else {
            const double T = jl_max(bilinear(med.T, med.nx, ir, j, xd, yd), 0.0);
            if (COLL == LMH_LANGEVIN || COLL == LMH_LANGEVIN_RTA) {
                double etaD, kT;
                drag_coeffs<false>(p, sp.drag, m, T, &etaD, &kT);
                ou_kick<3, false, true>(p, etaD, kT, T, m, a.sqrt_dt, q, xi);
            }
            if ((COLL == LMH_RTA || COLL == LMH_LANGEVIN_RTA) &&
                rta_decide<NOISE == LMH_NOISE_PHILOX>(p, rta_lookup(sp.rta, T), T, ucol)) {
                double usamp, g[3];
                noise_redraw<3, NOISE>(a, i, s_off, xi, usamp_p, &usamp, g);
                rta_redraw<3, false, true, NOISE == LMH_NOISE_PHILOX>(sp.jt, T, q, usamp, g);
            }
            if (!skip) radial_boost_fast(q, rhx, rhy, v, gam, m, +1);
        }
        // kernel_update_positions_cpu! (kernels_cpu.jl:535-636) on x, y as x + p (dt/E), dt/E from
        // one reciprocal square root; the eta_s push
        double E2 = m * m;
#pragma unroll
        for (int d = 0; d < 3; ++d) E2 += q[d] * q[d];
        const double dtE = p.dt * rsqrt_pos(E2);
        x[0] = fma(q[0], dtE, x[0]);
        x[1] = fma(q[1], dtE, x[1]);
        x[2] += milne_deta_fast(q[0], q[1], q[2], m, p.t0 + (double)(step - 1) * p.dt, p.dt);
        if (bad < 0 && !(finite_all(x, 3) && finite_all(q, 3))) bad = step;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) a.pos[d][i] = x[d];
#pragma unroll
    for (int d = 0; d < 3; ++d) a.mom[d][i] = q[d];
    if (bad >= 0) atomicMin(a.bad, (unsigned long long)bad);
}

// The driver's initial boost to the lab at `step` (simulate_cpu.jl:160-163), radial medium
template <int DIMS, int PDIM>
__global__ __launch_bounds__(kBlock) void k_boost_lab_radial(StepArgs a, RadialMedium med, int64_t step) {
    int seg;
    const int64_t i = particle_of_thread(a, &seg);
    if (i < 0) return;
    const SpeciesArgs sp = a.species[seg];
    constexpr bool RADIAL = DIMS == 1;
    const lmh_params& p = a.p;
    double x[DIMS], q[PDIM];
#pragma unroll
    for (int d = 0; d < DIMS; ++d) x[d] = a.pos[d][i];
#pragma unroll
    for (int d = 0; d < PDIM; ++d) q[d] = a.mom[d][i];
    double r2 = 0.0;
#pragma unroll
    for (int d = 0; d < DIMS; ++d)
        if (!(DIMS == 3 && p.coords == LMH_MILNE && d == 2)) r2 += x[d] * x[d];
    const double r = sqrt(r2);
    if (r < kEps) return;
    const double t_now = (double)step * p.dt + p.t0;
    double yd, xd;
    const int j = axis_cell_guess(med.tg, med.nt, med.tinv, t_now, &yd);
    const int ir = axis_cell_guess(med.xg, med.nx, med.xinv, r, &xd);
    double v = bilinear(med.V, med.nx, ir, j, xd, yd);
    if (DIMS >= 2 && med.V2 != nullptr) {
        double v2m = bilinear(med.V2, med.nx, ir, j, xd, yd);
        double phi = atan2(x[1 % DIMS], x[0]);
        v *= (1.0 + 2.0 * v2m * cos(2.0 * (phi - p.psi2)));
    }
    radial_boost<DIMS, PDIM, RADIAL>(q, x, r, v, sp.m, +1, p.relativistic, vmax_());
#pragma unroll
    for (int d = 0; d < PDIM; ++d) a.mom[d][i] = q[d];
}

// ---------------------------------------------------------------------------------------------
// 2+1D grid medium (extension): sample at (x, y) with the time cell (j, td) of the step.  Each
// corner is interpolated in tau first, then the corners in x, then in y.  SLICE reads the cell's
// bilinear coefficients of the tau-interpolated corners, prepared once per step by k_grid_slices:
// one 96-byte record per particle instead of four scattered corners; the same bilinear function,
// rounded within a few ulps of the corner form.
__device__ __forceinline__ int grid_cell(double f, int n) {
    // clamp(floor(f), 0, n - 2) with min/max (a NaN f takes cell 0)
    return (int)fmin(fmax(floor(f), 0.0), (double)(n - 2));
}

__device__ __forceinline__ double4 lerp4(const double4& a, const double4& b, double td) {
    double4 c;
    c.x = a.x * (1.0 - td) + b.x * td;
    c.y = a.y * (1.0 - td) + b.y * td;
    c.z = a.z * (1.0 - td) + b.z * td;
    c.w = 0.0;
    return c;
}

// the slice path in two parts, so that a kernel can place independent work between the gather
// and its use: the cell's record (and the fractions), then the bilinear combination.  The query
// clamp of interpolate_2d_cpu (kernels_cpu.jl:27-28) in cell units: the cell index clamped to
// [0, n - 2] and the fraction to [0, 1] (a NaN coordinate takes cell 0, fraction 0)
__device__ __forceinline__ void cell_frac(double x, double x0, double inv_d, int n, int* i, double* frac) {
    const double f = (x - x0) * inv_d;
    const int k = (int)floor(f);
    *i = k < 0 ? 0 : (k > n - 2 ? n - 2 : k);
    *frac = fmin(fmax(f - (double)*i, 0.0), 1.0);
}

__device__ __forceinline__ const double4* slice_cell(const GridMedium& g, double x, double y, double* xd,
                                                     double* yd) {
    int ix, iy;
    cell_frac(x, g.x0, g.inv_dx, g.nx, &ix, xd);
    cell_frac(y, g.y0, g.inv_dy, g.ny, &iy, yd);
    // a 32-bit byte offset into the slice (3 records of 32 bytes per cell)
    const uint32_t off = (uint32_t)(ix + (g.nx - 1) * iy) * 96u;
    return (const double4*)((const char*)g.slice + off);
}

__device__ __forceinline__ void slice_combine(const double4& cT, const double4& cx, const double4& cy,
                                              double xd, double yd, double* T, double* vx, double* vy) {
    *T = fma(fma(cT.w, xd, cT.z), yd, fma(cT.y, xd, cT.x));
    *vx = fma(fma(cx.w, xd, cx.z), yd, fma(cx.y, xd, cx.x));
    *vy = fma(fma(cy.w, xd, cy.z), yd, fma(cy.y, xd, cy.x));
}

// The medium at (x, y) from a node slice (the batch's form): g.slice holds (T, vx, vy) at the
// step's time for every grid node, one double4 per node in the nodes' row-major order.  The
// cell's coefficients are formed from its four nodes in the order the record slices are written
// (k_grid_slices), so the result equals grid_sample<true> on the same step's records, bit for
// bit; the slice is a third of the records' size for a few more subtractions per quark.
__device__ __forceinline__ void node_sample(const GridMedium& g, double x, double y, double* T, double* vx,
                                            double* vy) {
    int ix, iy;
    double xd, yd;
    cell_frac(x, g.x0, g.inv_dx, g.nx, &ix, &xd);
    cell_frac(y, g.y0, g.inv_dy, g.ny, &iy, &yd);
    // 32-bit byte offsets (the batch slices only grids of at most 2^27 nodes)
    const uint32_t off = (uint32_t)(ix + g.nx * iy) * 32u;
    const double4* r0 = (const double4*)((const char*)g.slice + off);
    const double4* r1 = (const double4*)((const char*)g.slice + off + (uint32_t)g.nx * 32u);
    const double4 c00 = r0[0], c10 = r0[1], c01 = r1[0], c11 = r1[1];
    const double4 cT = make_double4(c00.x, c10.x - c00.x, c01.x - c00.x, (c11.x - c10.x) - (c01.x - c00.x));
    const double4 cx = make_double4(c00.y, c10.y - c00.y, c01.y - c00.y, (c11.y - c10.y) - (c01.y - c00.y));
    const double4 cy = make_double4(c00.z, c10.z - c00.z, c01.z - c00.z, (c11.z - c10.z) - (c01.z - c00.z));
    slice_combine(cT, cx, cy, xd, yd, T, vx, vy);
}

template <bool SLICE>
__device__ __forceinline__ void grid_sample(const GridMedium& g, double x, double y, int j,
                                            double td, double* T, double* vx, double* vy) {
    // the query clamped to the grid as interpolate_2d_cpu does (kernels_cpu.jl:27-28); min/max
    // take a NaN coordinate to the lower edge (its state is non-finite already)
    double xc = fmin(fmax(x, g.x0), g.xlast);
    double yc = fmin(fmax(y, g.y0), g.ylast);
    double fx = (xc - g.x0) * g.inv_dx;
    double fy = (yc - g.y0) * g.inv_dy;
    int ix = grid_cell(fx, g.nx), iy = grid_cell(fy, g.ny);
    double xd = fx - (double)ix, yd = fy - (double)iy;
    int64_t n00 = ix + (int64_t)g.nx * iy;
    double4 c00, c10, c01, c11;
    if (SLICE) {
        // the cell's bilinear coefficients (a, b, c, d) of T, vx, vy: f = a + b xd + (c + d xd) yd,
        // the cell and fractions of slice_cell (the specialised kernels' form)
        double sxd, syd;
        const double4* cell = slice_cell(g, x, y, &sxd, &syd);
        const double4 cT = cell[0], cx = cell[1], cy = cell[2];
        slice_combine(cT, cx, cy, sxd, syd, T, vx, vy);
        return;
    } else {
        const int64_t nxy = (int64_t)g.nx * g.ny;
        const double4* s0 = g.nodes + nxy * j + n00;
        const double4* s1 = s0 + nxy;
        c00 = lerp4(ld_global4(s0), ld_global4(s1), td);
        c10 = lerp4(ld_global4(s0 + 1), ld_global4(s1 + 1), td);
        c01 = lerp4(ld_global4(s0 + g.nx), ld_global4(s1 + g.nx), td);
        c11 = lerp4(ld_global4(s0 + g.nx + 1), ld_global4(s1 + g.nx + 1), td);
    }
    double a0 = c00.x * (1.0 - xd) + c10.x * xd;
    double a1 = c01.x * (1.0 - xd) + c11.x * xd;
    *T = a0 * (1.0 - yd) + a1 * yd;
    double b0 = c00.y * (1.0 - xd) + c10.y * xd;
    double b1 = c01.y * (1.0 - xd) + c11.y * xd;
    *vx = b0 * (1.0 - yd) + b1 * yd;
    double e0 = c00.z * (1.0 - xd) + c10.z * xd;
    double e1 = c01.z * (1.0 - xd) + c11.z * xd;
    *vy = e0 * (1.0 - yd) + e1 * yd;
}

// The frame change on the grid: the reference's boost along r_hat with the radial speed
// (kernels_cpu.jl:101-144, 199-241) with r_hat -> v_hat = (vx, vy)/|v| and v -> |v|, skipped when
// |v| < eps as the reference skips r < eps.  The flow-dependent factors are formed once per step
// and applied for both directions.
struct FlowBoost {
    bool on;
    double nx, ny, vmag, v, gam;
};

__device__ __forceinline__ FlowBoost flow_boost(double vx, double vy) {
    FlowBoost b;
    double vm2 = fma(vy, vy, vx * vx);
    // |v| and 1/|v| from one reciprocal square root (an ulp from sqrt and vx/|v|); v = 0 gives
    // |v| = 0 and the boost is off
    const double inv = rsqrt_pos(fmax(vm2, 1e-300));
    b.vmag = vm2 * inv;
    b.on = !(b.vmag < kEps);
    b.nx = vx * inv;
    b.ny = vy * inv;
    double vmax = vmax_();
    b.v = b.vmag > vmax ? vmax : b.vmag;
    b.gam = rsqrt_pos(1.0 - b.v * b.v + 1e-10);
    return b;
}

template <int PDIM>
__device__ __forceinline__ void apply_boost(double* q, double m, const FlowBoost& b, int sign,
                                            int relativistic) {
    if (!b.on) return;
    if (!relativistic) {
        double s = sign > 0 ? 1.0 : -1.0;
        q[0] += s * (m * b.vmag * b.nx);
        q[1] += s * (m * b.vmag * b.ny);
        return;
    }
    // the reference's expressions with each product-sum as one fused multiply-add
    double p2 = q[0] * q[0];
#pragma unroll
    for (int d = 1; d < PDIM; ++d) p2 = fma(q[d], q[d], p2);
    const double E = sqrt_pos<true>(fma(m, m, p2));
    const double px = q[0], py = q[1];
    const double p_par = fma(py, b.ny, px * b.nx);
    const double p_perp_x = fma(-p_par, b.nx, px);
    const double p_perp_y = fma(-p_par, b.ny, py);
    const double p_par_new = b.gam * (sign > 0 ? fma(b.v, E, p_par) : fma(-b.v, E, p_par));
    q[0] = fma(p_par_new, b.nx, p_perp_x);
    q[1] = fma(p_par_new, b.ny, p_perp_y);
}

// kernel_set_to_fluid_velocity_cpu! (kernels_cpu.jl:713-731) with r_hat -> v_hat; a fluid at
// rest sets the transverse momentum to 0
template <int PDIM>
__device__ __forceinline__ void glue_to_flow(double* q, double m, double vx, double vy, int relativistic) {
    double vmag = sqrt(vx * vx + vy * vy);
    if (vmag < kEps) {
        q[0] = 0.0;
        q[1] = 0.0;
        return;
    }
    double nx = vx / vmag, ny = vy / vmag;
    if (!relativistic) {
        q[0] = m * vmag * nx;
        q[1] = m * vmag * ny;
        return;
    }
    double vmax = vmax_();
    double v = vmag > vmax ? vmax : vmag;
    double gam = 1.0 / sqrt(1.0 - v * v + 1e-10);
    q[0] = m * gam * v * nx;
    q[1] = m * gam * v * ny;
}

// position push (kernels_cpu.jl:537-541, 586-607) with the optional position diffusion; i is
// the particle's identity (its noise counters, the record column in fixed-noise mode).  SLICE:
// the medium of the diffusion from the corners (0), the step's record slice (1) or node slice (2)
template <int PDIM, int NOISE, int SLICE>
__device__ __forceinline__ void grid_push(const StepArgs& a, const lmh_params& p, const GridMedium& g,
                                          int64_t i, int64_t step, int s_off, int j, double td,
                                          double m, double* x0, double* x1, const double* q) {
    if (!p.position_diffusion) {
        double E2 = m * m;
#pragma unroll
        for (int d = 0; d < PDIM; ++d) E2 = fma(q[d], q[d], E2);
        // dt*p/E as p*(dt/E) with dt/E = dt/sqrt(E^2) from one reciprocal square root, within an
        // ulp or two of the reference form
        const double dtE = p.relativistic ? p.dt * rsqrt_pos(E2) : p.dt / m;
        *x0 = fma(q[0], dtE, *x0);
        *x1 = fma(q[1], dtE, *x1);
    } else {
        // with position diffusion the reference form x + dt*p/E (kernels_cpu.jl:586-589): the
        // jumps that follow, large where T is small, would carry the ulps of the shorter form
        double E2 = m * m;
#pragma unroll
        for (int d = 0; d < PDIM; ++d) E2 += q[d] * q[d];
        const double E = p.relativistic ? sqrt(E2) : m;
        *x0 = *x0 + p.dt * q[0] / E;
        *x1 = *x1 + p.dt * q[1] / E;

        double Td, ux, uy;
        if (SLICE == 2) node_sample(g, *x0, *x1, &Td, &ux, &uy);
        else grid_sample<SLICE == 1>(g, *x0, *x1, j, td, &Td, &ux, &uy);
        double D = (p.DsT / jl_max(Td, kEps)) / kFmGeV;
        if (D > 0.0) {
            double xp[2];
            noise_diffusion<2, NOISE>(a, i, step, s_off, xp);
            double sigma = sqrt(2.0 * D * p.dt);
            *x0 = *x0 + sigma * xp[0];
            *x1 = *x1 + sigma * xp[1];
        }
    }
}

// Rank of this lane among the block's queued lanes, in lane order (no atomics): the waves'
// counts are combined through LDS, and thread 0 records the block's count.  Every thread of the
// block must call it.
__device__ __forceinline__ int block_rank(bool want, int32_t* block_count) {
    __shared__ int wave_count[kBlock / 64];
    const unsigned long long mask = __ballot(want);
    const int lane = (int)__lane_id();
    const int wave = (int)(threadIdx.x >> 6);
    if (lane == 0) wave_count[wave] = __popcll(mask);
    __syncthreads();
    int before = 0, total = 0;
    const int nw = (int)((blockDim.x + 63) >> 6);
    for (int w = 0; w < nw; ++w) {
        before += w < wave ? wave_count[w] : 0;
        total += wave_count[w];
    }
    if (threadIdx.x == 0) *block_count = total;
    const unsigned long long below = mask & ((1ull << lane) - 1ull);
    return want ? before + __popcll(below) : -1;
}

// the medium at one time (tau cell j, fraction td), per cell: f = a + b xd + (c + d xd) yd with
// a = f00, b = f10 - f00, c = f01 - f00, d = f11 - f10 - f01 + f00 of the tau-interpolated corners.
// One thread per (cell, field) record: field fl (0 T, 1 vx, 2 vy) of cell k is slice[3 k + fl], so
// consecutive lanes write consecutive 32-byte records.
__device__ __forceinline__ double4 slice_coeffs(const GridMedium& g, int j, double td, int64_t k, int fl) {
    const int64_t ncx = g.nx - 1;
    const int64_t ix = k % ncx, iy = k / ncx;
    const int64_t nxy = (int64_t)g.nx * g.ny;
    const double* s0 = reinterpret_cast<const double*>(g.nodes + nxy * j + ix + (int64_t)g.nx * iy) + fl;
    const double* s1 = s0 + 4 * nxy;
    const int64_t up = 4 * (int64_t)g.nx;
    const double c00 = s0[0] * (1.0 - td) + s1[0] * td;
    const double c10 = s0[4] * (1.0 - td) + s1[4] * td;
    const double c01 = s0[up] * (1.0 - td) + s1[up] * td;
    const double c11 = s0[up + 4] * (1.0 - td) + s1[up + 4] * td;
    return make_double4(c00, c10 - c00, c01 - c00, (c11 - c10) - (c01 - c00));
}

// the slices of steps step0 .. step0 + nk - 1 (time cell of each step as the step kernels find it)
__global__ __launch_bounds__(kBlock) void k_grid_slices(GridMedium g, double4* slices, int64_t step0,
                                                        int32_t nk, double dt, double t0) {
    const int64_t nrec = 3 * (int64_t)(g.nx - 1) * (g.ny - 1);
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= nrec * nk) return;
    const int64_t s = k / nrec, r = k - s * nrec;
    const double t_now = (double)(step0 + s) * dt + t0;
    double td;
    const int j = axis_cell(g.tg, g.nt, t_now, &td);
    slices[k] = slice_coeffs(g, j, td, r / 3, (int)(r % 3));
}

// ---------------------------------------------------------------------------------------------
// Cell sorting (extension).  The quarks are kept ordered by the grid cell of their position
// within each species segment, so that the lanes of a wave read the same few medium records and
// the slice gathers coalesce.  A sort is a stable radix sort of (segment, cell) keys with the
// slot as value, then one gather of every state row and of the identity array through the
// permutation.  The identity array carries each quark's original slot, which keys its Philox
// noise, so a sorted run draws the same noise for every quark as the unsorted run.  Restoring
// the caller's order is the same gather after a radix sort by identity.
struct SortKeyArgs {
    const double* x;
    const double* y;
    int64_t begin[LMH_MAX_SPECIES + 1];
    int32_t nseg;
    int32_t cell_bits;
    int64_t N;
};

// key = segment << cell_bits | cell (the cell grid_sample reads), value = slot
__global__ __launch_bounds__(kBlock) void k_sort_keys(SortKeyArgs a, GridMedium g, uint32_t* keys, uint32_t* vals) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= a.N) return;
    int seg = 0;
#pragma unroll
    for (int k = 1; k < LMH_MAX_SPECIES; ++k)
        if (k < a.nseg && s >= a.begin[k]) seg = k;
    const double xc = fmin(fmax(a.x[s], g.x0), g.xlast);
    const double yc = fmin(fmax(a.y[s], g.y0), g.ylast);
    const int ix = grid_cell((xc - g.x0) * g.inv_dx, g.nx);
    const int iy = grid_cell((yc - g.y0) * g.inv_dy, g.ny);
    keys[s] = ((uint32_t)seg << a.cell_bits) | (uint32_t)(ix + (g.nx - 1) * iy);
    vals[s] = (uint32_t)s;
}

__global__ __launch_bounds__(kBlock) void k_iota(uint32_t* id, int64_t N) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s < N) id[s] = (uint32_t)s;
}

// key = identity, value = slot (the sort that restores the caller's order)
__global__ __launch_bounds__(kBlock) void k_unsort_keys(const uint32_t* id, int64_t N, uint32_t* keys, uint32_t* vals) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= N) return;
    keys[s] = id[s];
    vals[s] = (uint32_t)s;
}

// the state in the caller's order: dst row r, slot id[s] = src row r, slot s (an adjoint
// checkpoint taken while the particles are cell-sorted)
__global__ __launch_bounds__(kBlock) void k_scatter_rows(const double* __restrict__ src, double* __restrict__ dst,
                                                         int32_t rows, int64_t N, const uint32_t* __restrict__ id) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= N) return;
    const int64_t to = id[s];
    for (int r = 0; r < rows; ++r) dst[(int64_t)r * N + to] = src[(int64_t)r * N + s];
}

// the inverse of k_scatter_rows: dst row r, slot s = src row r, slot id[s] (rows in the caller's
// order into the order whose identities are id)
__global__ __launch_bounds__(kBlock) void k_gather_rows(const double* __restrict__ src, double* __restrict__ dst,
                                                        int32_t rows, int64_t N, const uint32_t* __restrict__ id) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= N) return;
    const int64_t from = id[s];
    for (int r = 0; r < rows; ++r) dst[(int64_t)r * N + s] = src[(int64_t)r * N + from];
}

// dst row r, slot s = src row r, slot perm[s]; the identity likewise.  Every row is read before
// any is written (the buffers never alias), so a thread's loads are in flight together.
constexpr int kMaxRows = 6;
__global__ __launch_bounds__(kBlock) void k_permute(const double* __restrict__ src, double* __restrict__ dst,
                                                    int32_t rows, int64_t N, const uint32_t* __restrict__ perm,
                                                    const uint32_t* __restrict__ src_id,
                                                    uint32_t* __restrict__ dst_id) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= N) return;
    const int64_t from = perm[s];
    double v[kMaxRows];
#pragma unroll
    for (int r = 0; r < kMaxRows; ++r)
        if (r < rows) v[r] = src[(int64_t)r * N + from];
    const uint32_t id = src_id[from];
#pragma unroll
    for (int r = 0; r < kMaxRows; ++r)
        if (r < rows) dst[(int64_t)r * N + s] = v[r];
    dst_id[s] = id;
}

// the one-step launches of lmh_run read the medium slices of up to kSliceRing steps computed by
// one k_grid_slices launch
constexpr int kSliceRing = 16;

// the adjoint's tape kernel and the batch kernel: at least three waves per SIMD
constexpr int kStepWaves = 3;

// the specialised sliced 2+1D step kernels (the common configuration): at least seven waves per
// SIMD, which they reach without spills; left to itself the compiler gives the one with the
// collision redraw 85 VGPRs and five waves, 4 % slower.  The generic kernels keep three.
constexpr int kGridWaves = 7;

template <int PDIM, int NOISE, bool SLICE, int COLL, bool DEFER, bool SPEC, bool GATHER = false>
__global__ __launch_bounds__(kBlock, (SLICE && SPEC) ? kGridWaves : kStepWaves) void k_cart_grid(StepArgs a, GridMedium g) {
    static_assert(!GATHER || (SPEC && !DEFER), "the gathering step is the specialised in-kernel-collision one");
    // The specialised kernels are VALU-issue bound with seven waves on a SIMD: a wave issues its
    // memory requests (the state loads here, the slice gather below) at wave priority 1, ahead
    // of the other waves' arithmetic, so that the loads' latency starts as early as it can, and
    // drops back to 0 for its own arithmetic (1.8 % per launch at 10^7 quarks)
    if (SPEC) __builtin_amdgcn_s_setprio(1);
    int seg;
    const int64_t i_thread = particle_of_thread(a, &seg);
    // with DEFER every thread reaches the block-wide ranking; threads past the segment's end
    // compute on the segment's first particle and neither store nor queue
    const bool valid = i_thread >= 0;
    if (!DEFER && !valid) return;
    const int64_t i = valid ? i_thread : a.begin[seg];
    const SpeciesArgs sp = a.species[seg];
    constexpr bool HAS_KICK = COLL == LMH_LANGEVIN || COLL == LMH_LANGEVIN_RTA;
    constexpr bool HAS_RTA = COLL == LMH_RTA || COLL == LMH_LANGEVIN_RTA;
    const lmh_params pp = step_params<SPEC>(a.p);
    const lmh_params& p = pp;
    const double m = sp.m;
    // the specialised kernels run for N < 2^29 only, so a state row is addressed by a 32-bit byte
    // offset from its base (no 64-bit address arithmetic per row)
    const uint32_t o8 = (uint32_t)i * 8u;
    // GATHER: the particle comes from slot perm[i] of the order the last cell sort found (the
    // sort's permutation applied by this step's reads instead of a pass of its own)
    const uint32_t src = GATHER ? *(const __attribute__((address_space(1))) uint32_t*)(
                                      (const char*)a.perm + (uint32_t)i * 4u)
                                : (uint32_t)i;
    const uint32_t r8 = GATHER ? src * 8u : o8;
    double x0 = SPEC ? ld_off(a.pos[0], r8) : a.pos[0][i];
    double x1 = SPEC ? ld_off(a.pos[1], r8) : a.pos[1][i];
    double q[PDIM];
#pragma unroll
    for (int d = 0; d < PDIM; ++d) q[d] = SPEC ? ld_off(a.mom[d], r8) : a.mom[d][i];
    // the particle's identity keys its noise (the slot unless the particles are cell-sorted)
    const int64_t pid = SPEC ? (int64_t)*(const __attribute__((address_space(1))) uint32_t*)(
                                   (const char*)a.id + (r8 >> 1))
                             : (int64_t)a.id[i];
    // every state load is issued before the medium sample's address arithmetic waits for x, y
    __builtin_amdgcn_sched_barrier(0);
    if (SPEC) __builtin_amdgcn_s_setprio(0);
    // freeze-out capture runs in the generic, non-deferred kernels only
    const bool fo = !SPEC && !DEFER && a.fo.on;
    bool frozen = false;
    double lastT = 0.0;
    if (fo) {
        frozen = a.fo.flag[i] != 0.0;
        lastT = a.fo.lastT[i];
    }
    int64_t bad = -1;
    // the sliced and the deferred launches are one step each (the slice holds one time); a
    // loop-free body keeps the per-step constants out of the registers of a loop
    const int nsteps = (SLICE || DEFER) ? 1 : a.nsteps;
    for (int s_off = 0; s_off < nsteps; ++s_off) {
        const int64_t step = a.step_begin + s_off;
        const double t_now = (double)step * p.dt + p.t0;
        double td;
        const int j = axis_cell(g.tg, g.nt, t_now, &td);
        double xin[2], qin[PDIM];
        if (fo && !frozen) {
            xin[0] = x0;
            xin[1] = x1;
#pragma unroll
            for (int d = 0; d < PDIM; ++d) qin[d] = q[d];
        }
        double Tm, vx, vy;
        double4 cT, cx, cy;
        double xd, yd;
        if (SLICE) {
            // the slice record is requested first ...
            if (SPEC) __builtin_amdgcn_s_setprio(1);
            const double4* cell = slice_cell(g, x0, x1, &xd, &yd);
            cT = cell[0];
            cx = cell[1];
            cy = cell[2];
            __builtin_amdgcn_sched_barrier(0);
            if (SPEC) __builtin_amdgcn_s_setprio(0);
        } else {
            grid_sample<false>(g, x0, x1, j, td, &Tm, &vx, &vy);
        }
        // ... and the kick noise, which needs no memory, is computed around it (the compiler
        // starts the Philox rounds, which need the identity only, ahead of the gather's address
        // arithmetic; an empty asm that reads the noise keeps it from sinking the noise to its uses)
        double xi[PDIM];
        double ucol = 1.0, usamp_p = 0.0;
        noise_kick<PDIM, NOISE, HAS_KICK ? 1 : 0, HAS_RTA ? 1 : 0>(a, pid, step, s_off, xi, &ucol, &usamp_p);
        if (NOISE == LMH_NOISE_PHILOX) {
#pragma unroll
            for (int d = 0; d < PDIM; ++d)
                if (HAS_KICK) asm volatile("" ::"v"(xi[d]));
            if (HAS_RTA) asm volatile("" ::"v"(ucol));
        }
        if (SLICE) slice_combine(cT, cx, cy, xd, yd, &Tm, &vx, &vy);
        const FlowBoost fb = flow_boost(vx, vy);
        apply_boost<PDIM>(q, m, fb, -1, p.relativistic);
        if (PDIM == 3 && p.bjorken_redshift) bjorken(p, step, &q[2 % PDIM]);
        if (COLL == LMH_FLUID) {
            glue_to_flow<PDIM>(q, m, vx, vy, p.relativistic);
        } else {
            const double T = jl_max(Tm, 0.0);
            // both time-table reads are issued together, before the kick needs the first
            const bool read_drag = HAS_KICK && drag_table_read(p);
            double tn = 0.0, tau_d = 0.0;
            if (HAS_RTA && read_drag) {
                eval_time_table_pair<SPEC>(sp.rta, sp.drag, sp.shared_axis != 0, T, &tn, &tau_d);
            } else {
                if (HAS_RTA) tn = rta_lookup(sp.rta, T);
                if (read_drag) tau_d = eval_time_table(sp.drag, T);
            }
            double etaD = 0.0, kT = 0.0;
            if (HAS_KICK) drag_coeffs_tau<false>(p, tau_d, m, T, &etaD, &kT);
            if (HAS_KICK) ou_kick<PDIM, false>(p, etaD, kT, T, m, a.sqrt_dt, q, xi);
            if (HAS_RTA) {
                const bool col = rta_decide<NOISE == LMH_NOISE_PHILOX>(p, tn, T, ucol);
                if (DEFER) {
                    const int rank = block_rank(valid && col, &a.qcount[blockIdx.x]);
                    if (!valid) return;
                    if (rank >= 0) {
                        // everything the redraw needs: the position, T and the flow at it, the
                        // identity and the slot (the LRF momentum is replaced by the redraw)
                        double4* r = a.qrec + 2 * ((int64_t)blockIdx.x * a.qblock + rank);
                        r[0] = make_double4(x0, x1, T, vx);
                        r[1] = make_double4(vy, __longlong_as_double((long long)pid),
                                            __longlong_as_double((long long)i), 0.0);
                        return;
                    }
                } else if (col) {
                    double usamp, gg[PDIM];
                    noise_redraw<PDIM, NOISE>(a, pid, s_off, xi, usamp_p, &usamp, gg);
                    rta_redraw<PDIM, false, true, NOISE == LMH_NOISE_PHILOX>(sp.jt, T, q, usamp, gg);
                }
            }
            apply_boost<PDIM>(q, m, fb, +1, p.relativistic);
        }
        grid_push<PDIM, NOISE, SLICE>(a, p, g, pid, step, s_off, j, td, m, &x0, &x1, q);
        if (fo && !frozen) {
            double Tf, ux, uy;
            grid_sample<SLICE>(g, x0, x1, j, td, &Tf, &ux, &uy);
            const double xo[2] = {x0, x1};
            freeze_latch<2, PDIM>(a, i, Tf, t_now, (double)(step - 1) * p.dt + p.t0, xin, qin, xo, q,
                                  &frozen, &lastT);
        }
        if (bad < 0 && !(isfinite(x0) && isfinite(x1) && finite_all(q, PDIM))) bad = step;
    }
    if (SPEC) {
        st_off(a.wpos[0], o8, x0);
        st_off(a.wpos[1], o8, x1);
#pragma unroll
        for (int d = 0; d < PDIM; ++d) st_off(a.wmom[d], o8, q[d]);
        if (GATHER) *(__attribute__((address_space(1))) uint32_t*)((char*)a.wid + (o8 >> 1)) = (uint32_t)pid;
    } else {
        a.wpos[0][i] = x0;
        a.wpos[1][i] = x1;
#pragma unroll
        for (int d = 0; d < PDIM; ++d) a.wmom[d][i] = q[d];
    }
    if (fo) a.fo.lastT[i] = lastT;
    if (bad >= 0) atomicMin(a.bad, (unsigned long long)bad);
}

// The deferred collisions of one step: the redraw of kernel_rta_collision_gpu!
// (kernels_gpu.jl:673-701), the boost back and the push, for every queued particle.  The
// particle's medium sample and flow factors are recomputed from the same inputs, so the result is
// the one the fused path gives.  Each wave takes kRedrawSpan consecutive step-kernel blocks and
// packs their queued particles into its lanes.
constexpr int kRedrawSpan = 8;

template <int PDIM, int NOISE, bool SLICE, bool SPEC>
__global__ __launch_bounds__(kBlock) void k_redraw_grid(StepArgs a, GridMedium g, int64_t nblocks) {
    const int64_t step = a.step_begin;
    const lmh_params pp = step_params<SPEC>(a.p);
    const lmh_params& p = pp;
    const double t_now = (double)step * p.dt + p.t0;
    double td;
    const int j = axis_cell(g.tg, g.nt, t_now, &td);
    const int lane = (int)__lane_id();
    const int64_t nwaves = (nblocks + kRedrawSpan - 1) / kRedrawSpan;
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < nwaves; w += wstride) {
        const int64_t b0 = w * kRedrawSpan;
        // inclusive prefix of the span's counts, held by lanes 0..kRedrawSpan-1
        const int cnt = (lane < kRedrawSpan && b0 + lane < nblocks) ? a.qcount[b0 + lane] : 0;
        int incl = cnt;
#pragma unroll
        for (int o = 1; o < kRedrawSpan; o <<= 1) {
            const int up = __shfl_up(incl, o);
            if (lane >= o) incl += up;
        }
        // broadcast the span's intervals while every lane of the wave is active (wave-uniform
        // values, read into scalar registers)
        int end_of[kRedrawSpan], cnt_of[kRedrawSpan];
#pragma unroll
        for (int t = 0; t < kRedrawSpan; ++t) {
            end_of[t] = __builtin_amdgcn_readlane(incl, t);
            cnt_of[t] = __builtin_amdgcn_readlane(cnt, t);
        }
        const int total = end_of[kRedrawSpan - 1];
        for (int k = lane; k < total; k += 64) {
            int bsel = 0, start = 0;
#pragma unroll
            for (int t = 0; t < kRedrawSpan; ++t) {
                if (k >= end_of[t] - cnt_of[t] && k < end_of[t]) {
                    bsel = t;
                    start = end_of[t] - cnt_of[t];
                }
            }
            const double4* r = a.qrec + 2 * ((b0 + bsel) * a.qblock + (k - start));
            const double4 r0 = r[0], r1 = r[1];
            double x0 = r0.x, x1 = r0.y;
            const double T = r0.z, vx = r0.w, vy = r1.x;
            const int64_t pid = (int64_t)__double_as_longlong(r1.y);
            const int64_t i = (int64_t)__double_as_longlong(r1.z);
            int seg = 0;
#pragma unroll
            for (int s2 = 1; s2 < LMH_MAX_SPECIES; ++s2)
                if (s2 < a.nseg && i >= a.begin[s2]) seg = s2;
            const SpeciesArgs sp = a.species_dev[seg];
            const double m = sp.m;
            double q[PDIM];
            const FlowBoost fb = flow_boost(vx, vy);
            // PHILOX: the step's block-0 noise again (the redraw's quantile and the kick normals);
            // FIXED: noise_redraw reads the record
            constexpr int DRAW = NOISE == LMH_NOISE_PHILOX ? 1 : 0;
            double xi[PDIM], ucol = 1.0, usamp_p = 0.0, usamp, gg[PDIM];
            noise_kick<PDIM, NOISE, DRAW, DRAW>(a, pid, step, 0, xi, &ucol, &usamp_p);
            noise_redraw<PDIM, NOISE>(a, pid, 0, xi, usamp_p, &usamp, gg);
            rta_redraw<PDIM, false, true>(sp.jt, T, q, usamp, gg);
            apply_boost<PDIM>(q, m, fb, +1, p.relativistic);
            grid_push<PDIM, NOISE, SLICE>(a, p, g, pid, step, 0, j, td, m, &x0, &x1, q);
            a.pos[0][i] = x0;
            a.pos[1][i] = x1;
#pragma unroll
            for (int d = 0; d < PDIM; ++d) a.mom[d][i] = q[d];
            if (!(isfinite(x0) && isfinite(x1) && finite_all(q, PDIM))) atomicMin(a.bad, (unsigned long long)step);
        }
    }
}

template <int PDIM>
__global__ __launch_bounds__(kBlock) void k_boost_lab_grid(StepArgs a, GridMedium g, int64_t step) {
    int seg;
    const int64_t i = particle_of_thread(a, &seg);
    if (i < 0) return;
    const SpeciesArgs sp = a.species[seg];
    const lmh_params& p = a.p;
    double q[PDIM];
#pragma unroll
    for (int d = 0; d < PDIM; ++d) q[d] = a.mom[d][i];
    const double t_now = (double)step * p.dt + p.t0;
    double td;
    const int j = axis_cell(g.tg, g.nt, t_now, &td);
    double Tm, vx, vy;
    grid_sample<false>(g, a.pos[0][i], a.pos[1][i], j, td, &Tm, &vx, &vy);
    apply_boost<PDIM>(q, sp.m, flow_boost(vx, vy), +1, p.relativistic);
#pragma unroll
    for (int d = 0; d < PDIM; ++d) a.mom[d][i] = q[d];
}

// ---------------------------------------------------------------------------------------------
// general-coordinate (tau, r) step: kernels_cpu_GC.jl:64-102 (boosts, regulators 1e-10 and
// 1e-8), :110-174 (forces: Gamma == 0, no m/E, T unclamped), :244-256 (mass shell), :259-274
// (push); order of simulate_cpu_general_coords.jl:83-115.
template <int NOISE>
__global__ __launch_bounds__(kBlock) void k_gc(StepArgs a, RadialMedium med) {
    int seg;
    const int64_t i = particle_of_thread(a, &seg);
    if (i < 0) return;
    const SpeciesArgs sp = a.species[seg];
    const lmh_params& p = a.p;
    const double m = sp.m;
    double tau = a.pos[0][i], rr = a.pos[1][i];
    double pt = a.mom[0][i], pr = a.mom[1][i];
    int64_t bad = -1;
    for (int s_off = 0; s_off < a.nsteps; ++s_off) {
        const int64_t step = a.step_begin + s_off;
        const double t_now = (double)step * p.dt + p.t0;
        double xi;
        if (NOISE == LMH_NOISE_FIXED) {
            xi = a.noise[(int64_t)s_off * a.N + i];
        } else {
            uint64_t w[4];
            noise_block<1>(p.seed, p.stream, i, step, 0, w);
            double n1;
            box_muller(w[0], w[1], &xi, &n1);
        }
        double yd, xd;
        const int j = axis_cell_gc(med.tg, med.nt, t_now, &yd);
        const int ir = axis_cell_gc(med.xg, med.nx, fabs(rr), &xd);
        const double v = bilinear(med.V, med.nx, ir, j, xd, yd);
        double gam = 1.0 / sqrt(1.0 - v * v + 1e-10);
        double npt = gam * (pt - v * pr);
        double npr = gam * (pr - v * pt);
        pt = npt;
        pr = npr;
        const double T = bilinear(med.T, med.nx, ir, j, xd, yd);
        double DsT_safe = p.DsT <= 0.0 ? 0.0 : p.DsT;
        double tau_n = DsT_safe == 0.0 ? 0.0 : eval_time_table(sp.drag, T);
        bool ok = tau_n > 0.0 && isfinite(tau_n);
        double etaD = ok ? (1.0 / tau_n) : 0.0;
        double kappa = ok ? ((2.0 * m * T) / tau_n) : 0.0;
        double kT = sqrt(kappa);
        double det, sto;
        if (etaD > 0.0 && p.dt > 0.0) {
            double aa = exp(-etaD * p.dt);
            double noise_pref = kT * sqrt((1.0 - aa * aa) / (2.0 * etaD * p.dt));
            det = (aa - 1.0) * pr + 0.0;
            sto = noise_pref * xi;
        } else {
            det = 0.0;
            sto = kT * xi;
        }
        pr = pr + (det + a.sqrt_dt * sto);
        pt = sqrt(m * m + pr * pr);
        double gam2 = 1.0 / sqrt(1.0 - v * v + 1e-8);
        npt = gam2 * (pt + v * pr);
        npr = gam2 * (pr + v * pt);
        pt = npt;
        pr = npr;
        const double E = pt;
        tau = tau + p.dt * pt / E;
        rr = rr + p.dt * pr / E;
        if (bad < 0 && !(isfinite(tau) && isfinite(rr) && isfinite(pt) && isfinite(pr))) bad = step;
    }
    a.pos[0][i] = tau;
    a.pos[1][i] = rr;
    a.mom[0][i] = pt;
    a.mom[1][i] = pr;
    if (bad >= 0) atomicMin(a.bad, (unsigned long long)bad);
}

// ---------------------------------------------------------------------------------------------
// observables: per-block partial sums, then one block per observable adds the partials in block
// order.  A thread accumulates the five moments of its particles in registers, in a fixed order;
// each wave reduces them with a butterfly of shuffles (every lane ends with the same bits) and the
// block adds its waves in order.  The plain histogram counts in LDS with integer atomics (exact,
// so the order does not matter); the smoothed histogram reduces each bin over the wave with the
// same butterfly and lane 0 adds it to its wave's LDS row.  The result does not depend on
// scheduling.
constexpr int kObsBlock = 256;
constexpr int kObsWaves = kObsBlock / 64;
constexpr int kObsMaxBins = 120;
constexpr int kObsMaxK = 5 + kObsMaxBins;
constexpr int kObsBlocks = 1024;           // blocks of a partial-sum launch (grid-stride)
static_assert((size_t)kObsBlocks * kObsMaxK <= kScratchDoubles, "observables partials exceed the scratch array");

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

struct ObsMoments {
    double n = 0.0, pt = 0.0, pt2 = 0.0, c2 = 0.0, ptc2 = 0.0;
    __device__ __forceinline__ void add(double a, double b, double* pt_out) {
        const double p2 = a * a + b * b;
        const double p = sqrt(p2);
        const double c = p2 > 0.0 ? (a * a - b * b) / p2 : 0.0;
        n += 1.0;
        pt += p;
        pt2 += p2;
        c2 += c;
        ptc2 += p * c;
        *pt_out = p;
    }
};

// The sums of particles [i0, i1) that block blk of an nblk-block launch takes (particle
// i0 + blk*kObsBlock + t + k*nblk*kObsBlock for thread t), written to partial[0, 5 + nbins):
// N, sum p_T, sum p_T^2, sum cos 2phi, sum p_T cos 2phi, then the bins.  Every thread of the
// block must call it.
__device__ __forceinline__ void obs_block(const double* __restrict__ px, const double* __restrict__ py,
                                          int64_t i0, int64_t i1, const lmh_obs_spec& s, int64_t blk,
                                          int64_t nblk, double* __restrict__ partial) {
    __shared__ double wmom[kObsWaves][5];
    __shared__ double whist[kObsWaves][kObsMaxBins];
    __shared__ unsigned int hist[kObsMaxBins];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int nb = s.nbins;
    const bool smooth = nb > 0 && s.sigma > 0.0;
    for (int k = t; k < nb; k += kObsBlock) {
        hist[k] = 0u;
#pragma unroll
        for (int w = 0; w < kObsWaves; ++w) whist[w][k] = 0.0;
    }
    __syncthreads();
    const double width = (s.pt_max - s.pt_min) / (double)(nb > 0 ? nb : 1);
    const double inv_w = (double)nb / (s.pt_max - s.pt_min);
    ObsMoments m;
    const int64_t stride = nblk * kObsBlock;
    if (smooth) {
        // every bin b adds Phi((hi - p_T)/sigma) - Phi((lo - p_T)/sigma) between the edges
        // pt_min + b width and pt_min + (b + 1) width, each edge's erfc serving the two bins it
        // bounds; all lanes of the block take the same rounds (past the end they add 0), so that
        // the shuffles see every lane
        const double inv_s = 1.0 / (s.sigma * sqrt(2.0));
        for (int64_t base = i0 + blk * kObsBlock; base < i1; base += stride) {
            const int64_t i = base + t;
            const bool ok = i < i1;
            double p = 0.0;
            if (ok) m.add(px[i], py[i], &p);
            double e_lo = erfc((p - s.pt_min) * inv_s);
            for (int k = 0; k < nb; ++k) {
                const double hi = s.pt_min + (double)(k + 1) * width;
                const double e_hi = erfc((p - hi) * inv_s);
                double c = ok ? 0.5 * (e_hi - e_lo) : 0.0;
                e_lo = e_hi;
                c = wave_sum(c);
                if (lane == 0) whist[wave][k] += c;
            }
        }
    } else {
        int64_t i = i0 + blk * kObsBlock + t;
        // four particles in flight per thread (eight was no faster)
        constexpr int U = 4;
        for (; i + (U - 1) * stride < i1; i += U * stride) {
            double a[U], b[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                a[u] = px[i + u * stride];
                b[u] = py[i + u * stride];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                double p;
                m.add(a[u], b[u], &p);
                if (nb > 0) {
                    const double fl = floor((p - s.pt_min) * inv_w);
                    if (fl >= 0.0 && fl < (double)nb) atomicAdd(&hist[(int)fl], 1u);
                }
            }
        }
        for (; i < i1; i += stride) {
            double p;
            m.add(px[i], py[i], &p);
            if (nb > 0) {
                const double fl = floor((p - s.pt_min) * inv_w);
                if (fl >= 0.0 && fl < (double)nb) atomicAdd(&hist[(int)fl], 1u);
            }
        }
    }
    const double v[5] = {wave_sum(m.n), wave_sum(m.pt), wave_sum(m.pt2), wave_sum(m.c2), wave_sum(m.ptc2)};
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) wmom[wave][k] = v[k];
    }
    __syncthreads();
    if (t < 5) {
        double a = wmom[0][t];
#pragma unroll
        for (int w = 1; w < kObsWaves; ++w) a += wmom[w][t];
        partial[t] = a;
    }
    for (int k = t; k < nb; k += kObsBlock) {
        double a = whist[0][k];
#pragma unroll
        for (int w = 1; w < kObsWaves; ++w) a += whist[w][k];
        partial[5 + k] = smooth ? a : (double)hist[k];
    }
}

__global__ __launch_bounds__(kObsBlock) void k_obs_partial(const double* __restrict__ px,
                                                            const double* __restrict__ py,
                                                            int64_t i0, int64_t i1, lmh_obs_spec s,
                                                            double* __restrict__ partial) {
    obs_block(px, py, i0, i1, s, blockIdx.x, gridDim.x, partial + (int64_t)blockIdx.x * (5 + s.nbins));
}

// one block per observable: the block's threads add strided partials, then a fixed tree
constexpr int kObsFinalBlock = 256;

__global__ __launch_bounds__(kObsFinalBlock) void k_obs_final(const double* __restrict__ partial, int nblk,
                                                               int K, double* __restrict__ out) {
    __shared__ double acc[kObsFinalBlock];
    const int k = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kObsFinalBlock) s += partial[(int64_t)b * K + k];
    acc[threadIdx.x] = s;
    __syncthreads();
    for (int st = kObsFinalBlock / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) acc[threadIdx.x] += acc[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = acc[0];
}

#include "lmhip_adjoint.h"
#include "lmhip_batch.h"
#include "lmhip_sample.h"

}  // namespace

// =============================================================================================
// host side
struct lmh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // particles
    int dims = 0, pdim = 0;
    int64_t N = 0;
    int nspecies = 0;
    int64_t counts[LMH_MAX_SPECIES] = {0, 0, 0, 0};
    double mass[LMH_MAX_SPECIES] = {0, 0, 0, 0};
    double* d_state = nullptr;      // (dims + pdim) rows of N doubles
    unsigned long long* d_bad = nullptr;
    double* d_scratch = nullptr;    // this device's copy of the static scratch array

    // medium
    int medium = -1;                // 0 radial, 1 grid
    double* d_rg = nullptr;
    double* d_tg = nullptr;
    double* d_T = nullptr;
    double* d_V = nullptr;
    double* d_V2 = nullptr;
    int nr = 0, nt = 0;
    double r_axis_eps = 1e-6, r_max = 0.0;
    double r_inv = 0.0, t_inv = 0.0;  // the radial grids' equal_spacing_inv
    double4* d_nodes = nullptr;
    double4* d_slice = nullptr;
    int gx = 0, gy = 0;
    double gx0 = 0, gdx = 1, gy0 = 0, gdy = 1;

    lmh_tuning tune = {256, 1, 1, 0, 1, 64};

    // cell sorting: the alternate state and identity buffers of the permutation, the radix
    // sort's keys, values and workspace; is_sorted says the state is in cell order with d_id
    // holding each slot's identity
    bool is_sorted = false;
    int64_t steps_since_sort = 0;
    bool perm_pending = false;        // the last sort's permutation (d_keys + 3N) not yet applied:
                                      // the next step reads through it (or k_permute applies it)
    uint64_t order_gen = 0;           // bumped by every permutation of the state
    double* d_state_alt = nullptr;
    uint32_t* d_id = nullptr;
    uint32_t* d_id_alt = nullptr;
    uint32_t* d_keys = nullptr;       // keys in, keys out, values in, values out: 4 N words
    void* d_sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int64_t sort_n = 0;               // N the sort buffers were allocated for

    // tables
    struct Tab { double* d = nullptr; double Tmin = 0, invdT = 1; int n = 0; };
    struct JTab { double* d = nullptr; double Tmin = 0, invdT = 1; int nU = 0, nT = 0; };
    Tab drag[LMH_MAX_SPECIES], rta[LMH_MAX_SPECIES];
    JTab jt[LMH_MAX_SPECIES];

    // per-species mass and tables as the kernels read them
    SpeciesArgs h_species[LMH_MAX_SPECIES];
    SpeciesArgs* d_species = nullptr;
    // the deferred collisions of the queued mode (tuning defer 1): records and per-block counts
    double4* d_qrec = nullptr;        // two double4 per queued particle
    int32_t* d_qcount = nullptr;      // one count per step-kernel block

    // launch timing
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    struct Mark { int cls; hipEvent_t a, b; };
    std::vector<Mark> marks;

    // snapshots: a device copy of the state taken on the step stream, then copied to pinned
    // host memory on the copy stream while the next steps run
    struct Snap {
        double* d = nullptr;          // device copy of the state
        double* h = nullptr;          // pinned host copy
        size_t n = 0;                 // doubles
        hipEvent_t taken = nullptr;   // the device copy is complete (step stream)
        hipEvent_t done = nullptr;    // the host copy is complete (copy stream)
        bool busy = false;
    };
    std::vector<Snap> snaps;
    hipStream_t copy_stream = nullptr;

    // on-the-fly freeze-out capture: rows fo_pos (dims), fo_mom (pdim), tau, flag, last T
    bool fo_on = false;
    int32_t fo_interp = 1;
    double fo_Tfo = 0.0;
    double* d_fo = nullptr;

    // bumped by every entry point that changes the particles, the medium or the tables; an
    // adjoint backward pass needs the state its forward pass left
    uint64_t gen = 0;
    // pathwise adjoint: the last adjoint forward pass and its checkpoints, the backward workspace
    struct AdjRecord {
        bool valid = false;
        lmh_params p;
        int64_t step_begin = 0, nsteps = 0;
        int32_t K = 1;
        bool keep = false;            // the forward pass kept the whole tape and every slice
        double dtau_dDsT = 0.0;
        uint64_t gen = 0;
        // the order of each checkpoint's rows: -1 the caller's, else the index of the identity
        // array (d_ckpt_id + index * N) of the cell order the segment ran in
        std::vector<int32_t> ord;
    } adj;
    double* d_ckpt = nullptr;
    size_t ckpt_cap = 0;
    uint32_t* d_ckpt_id = nullptr;
    size_t ckpt_id_cap = 0;
    double* d_tape = nullptr;
    size_t tape_cap = 0;
    double* d_adjslice = nullptr;     // the medium slices of a segment's steps, two buffers
    size_t adjslice_cap = 0;
    double* d_adjw = nullptr;         // lam, per-particle gradients, partial sums, cotangent
    size_t adjw_cap = 0;
    // the backward pass re-runs segment c-1 on a second stream while segment c is reversed;
    // the two tape buffers alternate, ordered by events
    hipStream_t tape_stream = nullptr;
    hipEvent_t ev_tape[2] = {nullptr, nullptr};
    hipEvent_t ev_rev[2] = {nullptr, nullptr};

    // scratch
    double* d_noise = nullptr;
    size_t noise_cap = 0;
    double* d_obs = nullptr;
    size_t obs_cap = 0;

    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
};

namespace {

#define LMH_HIP(ctx, expr)                                                                  \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return (ctx)->fail(LMH_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

int upload(lmh_ctx* c, double** dst, const double* src, size_t n) {
    dfree(*dst);
    if (hipMalloc((void**)dst, n * sizeof(double)) != hipSuccess) {
        *dst = nullptr;
        return c->fail(LMH_E_NOMEM, "device allocation failed");
    }
    LMH_HIP(c, hipMemcpy(*dst, src, n * sizeof(double), hipMemcpyHostToDevice));
    return LMH_OK;
}

bool sorted(const double* v, int n) {
    for (int k = 1; k < n; ++k)
        if (!(v[k] >= v[k - 1])) return false;
    return true;
}

int noise_layout(const lmh_params* p, int dims, int pdim, int* oxi, int* orta, int* opos) {
    int c = 0;
    *oxi = *orta = *opos = -1;
    if (p->coords == LMH_GENERAL_COORDS) {
        *oxi = 0;
        return 1;
    }
    if (p->collision == LMH_LANGEVIN || p->collision == LMH_LANGEVIN_RTA) { *oxi = c; c += pdim; }
    if (p->collision == LMH_RTA || p->collision == LMH_LANGEVIN_RTA) { *orta = c; c += 2 + pdim; }
    if (p->position_diffusion) { *opos = c; c += dims; }
    return c;
}

int64_t species_begin(const lmh_ctx* c, int k) {
    int64_t b = 0;
    for (int j = 0; j < k; ++j) b += c->counts[j];
    return b;
}

// the kernels' view of the species (masses and table pointers), copied to the device whenever a
// table or the particles change
int upload_species(lmh_ctx* c) {
    for (int k = 0; k < LMH_MAX_SPECIES; ++k) {
        SpeciesArgs& s = c->h_species[k];
        s.m = c->mass[k];
        s.drag = TimeTable{c->drag[k].d, c->drag[k].Tmin, c->drag[k].invdT, c->drag[k].n};
        s.rta = TimeTable{c->rta[k].d, c->rta[k].Tmin, c->rta[k].invdT, c->rta[k].n};
        s.jt = JuttnerTable{c->jt[k].d, c->jt[k].Tmin, c->jt[k].invdT, c->jt[k].nU, c->jt[k].nT};
        s.shared_axis = tables_share_axis(s.drag, s.rta);
    }
    if (!c->d_species && hipMalloc((void**)&c->d_species, sizeof(c->h_species)) != hipSuccess) {
        c->d_species = nullptr;
        return c->fail(LMH_E_NOMEM, "species table allocation failed");
    }
    LMH_HIP(c, hipMemcpy(c->d_species, c->h_species, sizeof(c->h_species), hipMemcpyHostToDevice));
    return LMH_OK;
}

// kernel arguments for all species segments; *nblocks receives the grid size
StepArgs make_args(lmh_ctx* c, const lmh_params* p, int block, unsigned* nblocks) {
    StepArgs a;
    memset(&a, 0, sizeof(a));
    for (int d = 0; d < 3; ++d) {
        a.pos[d] = d < c->dims ? c->d_state + (int64_t)d * c->N : nullptr;
        a.mom[d] = d < c->pdim ? c->d_state + (int64_t)(c->dims + d) * c->N : nullptr;
        a.wpos[d] = a.pos[d];
        a.wmom[d] = a.mom[d];
    }
    a.N = c->N;
    a.p = *p;
    a.bad = c->d_bad;
    a.nseg = c->nspecies;
    int64_t b = 0, blk = 0;
    for (int k = 0; k < LMH_MAX_SPECIES; ++k) {
        a.begin[k] = b;
        a.block0[k] = blk;
        if (k < c->nspecies) {
            b += c->counts[k];
            blk += (c->counts[k] + block - 1) / block;
        }
    }
    a.begin[LMH_MAX_SPECIES] = b;
    a.block0[LMH_MAX_SPECIES] = blk;
    // segments past nseg are empty and never selected
    memcpy(a.species, c->h_species, sizeof(a.species));
    a.species_dev = c->d_species;
    a.id = c->d_id;
    a.qrec = c->d_qrec;
    a.qcount = c->d_qcount;
    a.qblock = block;
    a.sqrt_dt = std::sqrt(p->dt);
    a.C = noise_layout(p, c->dims, c->pdim, &a.oxi, &a.orta, &a.opos);
    if (c->fo_on && c->d_fo) {
        const int64_t N = c->N;
        a.fo.on = 1;
        a.fo.interp = c->fo_interp;
        a.fo.Tfo = c->fo_Tfo;
        a.fo.pos = c->d_fo;
        a.fo.mom = c->d_fo + (int64_t)c->dims * N;
        a.fo.tau = c->d_fo + (int64_t)(c->dims + c->pdim) * N;
        a.fo.flag = a.fo.tau + N;
        a.fo.lastT = a.fo.flag + N;
    }
    *nblocks = (unsigned)blk;
    return a;
}

RadialMedium make_radial(const lmh_ctx* c) {
    RadialMedium m;
    m.xg = c->d_rg;
    m.tg = c->d_tg;
    m.T = c->d_T;
    m.V = c->d_V;
    m.V2 = c->d_V2;
    m.nx = c->nr;
    m.nt = c->nt;
    m.r_axis_eps = c->r_axis_eps;
    m.r_max = c->r_max;
    m.xinv = c->r_inv;
    m.tinv = c->t_inv;
    return m;
}

GridMedium make_grid(const lmh_ctx* c) {
    GridMedium g;
    g.nodes = c->d_nodes;
    g.slice = c->d_slice;
    g.tg = c->d_tg;
    g.nx = c->gx;
    g.ny = c->gy;
    g.nt = c->nt;
    g.x0 = c->gx0;
    g.inv_dx = 1.0 / c->gdx;
    g.xlast = c->gx0 + (double)(c->gx - 1) * c->gdx;
    g.y0 = c->gy0;
    g.inv_dy = 1.0 / c->gdy;
    g.ylast = c->gy0 + (double)(c->gy - 1) * c->gdy;
    return g;
}

unsigned grid_blocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// freeze-out arrays for the current particles, all zero (nothing latched, no previous T)
int fo_reset(lmh_ctx* c) {
    LMH_HIP(c, hipStreamSynchronize(c->stream));
    dfree(c->d_fo);
    const size_t n = (size_t)(c->dims + c->pdim + 3) * (size_t)c->N;
    if (hipMalloc((void**)&c->d_fo, n * sizeof(double)) != hipSuccess) {
        c->d_fo = nullptr;
        c->fo_on = false;
        return c->fail(LMH_E_NOMEM, "freeze-out: device allocation failed");
    }
    LMH_HIP(c, hipMemsetAsync(c->d_fo, 0, n * sizeof(double), c->stream));
    LMH_HIP(c, hipStreamSynchronize(c->stream));
    return LMH_OK;
}

hipEvent_t next_event(lmh_ctx* c) {
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

// brackets one launch with events while profiling
struct LaunchMark {
    lmh_ctx* c;
    int cls;
    hipEvent_t a = nullptr;
    LaunchMark(lmh_ctx* ctx, int k) : c(ctx), cls(k) {
        if (c->profiling && (a = next_event(c))) (void)hipEventRecord(a, c->stream);
    }
    ~LaunchMark() {
        if (!a) return;
        hipEvent_t b = next_event(c);
        if (!b) return;
        (void)hipEventRecord(b, c->stream);
        c->marks.push_back({cls, a, b});
    }
};

// ---- cell sorting (host side) ----
int bits_for(uint64_t n) {   // bits that hold the values 0 .. n - 1
    int b = 0;
    while (b < 32 && ((uint64_t)1 << b) < n) ++b;
    return b;
}

int sort_alloc(lmh_ctx* c) {
    if (c->sort_n == c->N && c->d_state_alt) return LMH_OK;
    const size_t N = (size_t)c->N, rows = (size_t)(c->dims + c->pdim);
    dfree(c->d_state_alt);
    dfree(c->d_id_alt);
    dfree(c->d_keys);
    dfree(c->d_sort_tmp);
    c->sort_tmp_bytes = 0;
    c->sort_n = 0;
    if (hipMalloc((void**)&c->d_state_alt, rows * N * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&c->d_id_alt, N * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void**)&c->d_keys, 4 * N * sizeof(uint32_t)) != hipSuccess)
        return c->fail(LMH_E_NOMEM, "cell sort: buffer allocation failed");
    uint32_t* k = c->d_keys;
    LMH_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, c->sort_tmp_bytes, k, k + N, k + 2 * N, k + 3 * N,
                                                   (int)N, 0, 32, c->stream));
    if (hipMalloc(&c->d_sort_tmp, std::max<size_t>(c->sort_tmp_bytes, 1)) != hipSuccess)
        return c->fail(LMH_E_NOMEM, "cell sort: workspace allocation failed");
    c->sort_n = c->N;
    return LMH_OK;
}

// the state and identities now in the alternate buffers become the current ones
void swap_order(lmh_ctx* c) {
    std::swap(c->d_state, c->d_state_alt);
    std::swap(c->d_id, c->d_id_alt);
    ++c->order_gen;
}

// apply a pending permutation (d_keys + 3N) to the state and the identities
int finish_permutation(lmh_ctx* c) {
    if (!c->perm_pending) return LMH_OK;
    const int64_t N = c->N;
    hipLaunchKernelGGL(k_permute, dim3(grid_blocks(N)), dim3(kBlock), 0, c->stream, c->d_state, c->d_state_alt,
                       (int32_t)(c->dims + c->pdim), N, c->d_keys + 3 * N, c->d_id, c->d_id_alt);
    LMH_HIP(c, hipGetLastError());
    swap_order(c);
    c->perm_pending = false;
    return LMH_OK;
}

// radix-sort the keys d_keys[0, N) with the values d_keys[2N, 3N) over bits [0, end_bit): the
// sorted values d_keys[3N, 4N) are the permutation (slot s takes the particle of slot perm[s]),
// applied to the state and the identities at once
int sort_apply(lmh_ctx* c, int end_bit) {
    const int64_t N = c->N;
    uint32_t* k = c->d_keys;
    size_t bytes = c->sort_tmp_bytes;
    LMH_HIP(c, hipcub::DeviceRadixSort::SortPairs(c->d_sort_tmp, bytes, k, k + N, k + 2 * N, k + 3 * N,
                                                   (int)N, 0, end_bit, c->stream));
    c->perm_pending = true;
    return finish_permutation(c);
}

// order the particles by (species segment, grid cell); the 2+1D grid medium must be set.
// defer_permute: the permutation is left for the next step launch to apply (perm_pending).
// mark: time the sort's launches as one LMH_PROF_SORT entry (the caller times them otherwise).
int sort_by_cell(lmh_ctx* c, bool defer_permute = false, bool mark = true) {
    int rc = finish_permutation(c);
    if (rc) return rc;
    if ((rc = sort_alloc(c))) return rc;
    std::optional<LaunchMark> mk;
    if (mark) mk.emplace(c, LMH_PROF_SORT);
    const int64_t N = c->N;
    SortKeyArgs ka;
    memset(&ka, 0, sizeof(ka));
    ka.x = c->d_state;
    ka.y = c->d_state + N;
    int64_t b = 0;
    for (int k = 0; k <= LMH_MAX_SPECIES; ++k) {
        ka.begin[k] = b;
        if (k < c->nspecies) b += c->counts[k];
    }
    ka.nseg = c->nspecies;
    ka.cell_bits = bits_for((uint64_t)(c->gx - 1) * (uint64_t)(c->gy - 1));
    ka.N = N;
    const int end_bit = ka.cell_bits + bits_for((uint64_t)c->nspecies);
    if (end_bit > 32) return c->fail(LMH_E_INVALID, "cell sort: too many grid cells");
    hipLaunchKernelGGL(k_sort_keys, dim3(grid_blocks(N)), dim3(kBlock), 0, c->stream, ka, make_grid(c),
                       c->d_keys, c->d_keys + 2 * N);
    LMH_HIP(c, hipGetLastError());
    // each species segment is radix-sorted on its own over the cell bits (the segment bits are
    // the same within it): a pass fewer than one sort over (segment, cell) when there are two
    // species.  The sorted values d_keys[3N, 4N) are the permutation.
    uint32_t* k = c->d_keys;
    for (int sg = 0; sg < c->nspecies; ++sg) {
        const int64_t b0 = ka.begin[sg], n = ka.begin[sg + 1] - b0;
        if (n <= 0) continue;
        size_t bytes = c->sort_tmp_bytes;
        LMH_HIP(c, hipcub::DeviceRadixSort::SortPairs(c->d_sort_tmp, bytes, k + b0, k + N + b0, k + 2 * N + b0,
                                                       k + 3 * N + b0, (int)n, 0, std::max(1, ka.cell_bits),
                                                       c->stream));
    }
    c->perm_pending = true;
    if (!defer_permute && (rc = finish_permutation(c))) return rc;
    c->is_sorted = true;
    c->steps_since_sort = 0;
    return LMH_OK;
}

// restore the caller's order (a no-op unless the particles are cell-sorted)
int ensure_unsorted(lmh_ctx* c) {
    if (!c->is_sorted) return LMH_OK;
    LMH_HIP(c, hipSetDevice(c->device));
    int rc0 = finish_permutation(c);
    if (rc0) return rc0;
    LaunchMark mk(c, LMH_PROF_SORT);
    const int64_t N = c->N;
    hipLaunchKernelGGL(k_unsort_keys, dim3(grid_blocks(N)), dim3(kBlock), 0, c->stream, c->d_id, N, c->d_keys,
                       c->d_keys + 2 * N);
    LMH_HIP(c, hipGetLastError());
    int rc = sort_apply(c, bits_for((uint64_t)N));
    if (rc) return rc;
    c->is_sorted = false;
    return LMH_OK;
}

}  // namespace

// The address of the static scratch array on `device` (which must be current), looked up once
// per device and kept for every later context on it.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

hipError_t scratch_address(int device, double** out) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = static_cast<double*>(p);
    }
    *out = g_scratch[device];
    return hipSuccess;
}

extern "C" {

int32_t lmh_noise_channels(const lmh_params* p, int32_t dims, int32_t pdim) {
    int a, b, c;
    if (!p) return 0;
    return noise_layout(p, dims, pdim, &a, &b, &c);
}

int lmh_create(lmh_ctx** out, int32_t device) {
    if (!out) return LMH_E_INVALID;
    *out = nullptr;
    lmh_ctx* c = new (std::nothrow) lmh_ctx();
    if (!c) return LMH_E_NOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return LMH_E_HIP;
    }
    if (hipMalloc((void**)&c->d_bad, sizeof(unsigned long long)) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return LMH_E_NOMEM;
    }
    unsigned long long none = ULLONG_MAX;
    if (hipMemcpy(c->d_bad, &none, sizeof(none), hipMemcpyHostToDevice) != hipSuccess ||
        scratch_address(device, &c->d_scratch) != hipSuccess) {
        (void)hipFree(c->d_bad);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return LMH_E_HIP;
    }
    *out = c;
    return LMH_OK;
}

void lmh_destroy(lmh_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    dfree(c->d_state);
    dfree(c->d_bad);
    dfree(c->d_rg);
    dfree(c->d_tg);
    dfree(c->d_T);
    dfree(c->d_V);
    dfree(c->d_V2);
    dfree(c->d_nodes);
    dfree(c->d_slice);
    dfree(c->d_ckpt);
    dfree(c->d_ckpt_id);
    dfree(c->d_tape);
    dfree(c->d_adjslice);
    dfree(c->d_adjw);
    dfree(c->d_fo);
    for (int k = 0; k < LMH_MAX_SPECIES; ++k) {
        dfree(c->drag[k].d);
        dfree(c->rta[k].d);
        dfree(c->jt[k].d);
    }
    dfree(c->d_noise);
    dfree(c->d_obs);
    dfree(c->d_species);
    dfree(c->d_qrec);
    dfree(c->d_qcount);
    dfree(c->d_state_alt);
    dfree(c->d_id);
    dfree(c->d_id_alt);
    dfree(c->d_keys);
    dfree(c->d_sort_tmp);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->tape_stream) {
        (void)hipStreamSynchronize(c->tape_stream);
        (void)hipStreamDestroy(c->tape_stream);
    }
    if (c->copy_stream) {
        (void)hipStreamSynchronize(c->copy_stream);
        (void)hipStreamDestroy(c->copy_stream);
    }
    for (auto& sn : c->snaps) {
        dfree(sn.d);
        if (sn.h) (void)hipHostFree(sn.h);
        if (sn.taken) (void)hipEventDestroy(sn.taken);
        if (sn.done) (void)hipEventDestroy(sn.done);
    }
    for (int k = 0; k < 2; ++k) {
        if (c->ev_tape[k]) (void)hipEventDestroy(c->ev_tape[k]);
        if (c->ev_rev[k]) (void)hipEventDestroy(c->ev_rev[k]);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* lmh_last_error(const lmh_ctx* c) { return c ? c->err.c_str() : "null context"; }

void* lmh_stream(lmh_ctx* c) { return c ? (void*)c->stream : nullptr; }

int lmh_synchronize(lmh_ctx* c) {
    if (!c) return LMH_E_INVALID;
    LMH_HIP(c, hipSetDevice(c->device));
    LMH_HIP(c, hipStreamSynchronize(c->stream));
    return LMH_OK;
}

// (n - 1)/(g[n-1] - g[0]) when every node lies within 1e-9 of a spacing of its place on an
// equally spaced grid (the kernels' first guess of a cell is then off by at most one), else 0
static double equal_spacing_inv(const double* g, int32_t n) {
    const double span = g[n - 1] - g[0];
    if (!(span > 0.0) || !std::isfinite(span)) return 0.0;
    const double h = span / (double)(n - 1);
    for (int32_t k = 0; k < n; ++k)
        if (!(std::fabs(g[k] - (g[0] + (double)k * h)) <= 1e-9 * h)) return 0.0;
    return (double)(n - 1) / span;
}

int lmh_set_medium_radial(lmh_ctx* c, const double* rgrid, int32_t nr, const double* tgrid,
                          int32_t nt, const double* T, const double* v, const double* v2) {
    if (!c) return LMH_E_INVALID;
    ++c->gen;
    if (!rgrid || !tgrid || !T || !v || nr < 2 || nt < 2)
        return c->fail(LMH_E_INVALID, "set_medium_radial: grids need >= 2 nodes and T, v must be given");
    if (!sorted(rgrid, nr) || !sorted(tgrid, nt))
        return c->fail(LMH_E_INVALID, "set_medium_radial: grids must be sorted");
    LMH_HIP(c, hipSetDevice(c->device));
    LMH_HIP(c, hipStreamSynchronize(c->stream));
    size_t n = (size_t)nr * (size_t)nt;
    int rc;
    if ((rc = upload(c, &c->d_rg, rgrid, nr)) || (rc = upload(c, &c->d_tg, tgrid, nt)) ||
        (rc = upload(c, &c->d_T, T, n)) || (rc = upload(c, &c->d_V, v, n)))
        return rc;
    if (v2) {
        if ((rc = upload(c, &c->d_V2, v2, n))) return rc;
    } else {
        dfree(c->d_V2);
    }
    dfree(c->d_nodes);
    dfree(c->d_slice);
    c->nr = nr;
    c->nt = nt;
    c->r_axis_eps = std::max(1e-12, 0.5 * std::fabs(rgrid[1] - rgrid[0]));
    c->r_max = rgrid[nr - 1];
    c->r_inv = equal_spacing_inv(rgrid, nr);
    c->t_inv = equal_spacing_inv(tgrid, nt);
    c->medium = 0;
    return LMH_OK;
}

int lmh_set_medium_grid2p1(lmh_ctx* c, double x0, double dx, int32_t nx, double y0, double dy,
                           int32_t ny, const double* tgrid, int32_t nt, const double* T,
                           const double* vx, const double* vy) {
    if (!c) return LMH_E_INVALID;
    ++c->gen;
    if (!tgrid || !T || !vx || !vy || nx < 2 || ny < 2 || nt < 2 || !(dx > 0.0) || !(dy > 0.0) ||
        !std::isfinite(x0) || !std::isfinite(y0))
        return c->fail(LMH_E_INVALID, "set_medium_grid2p1: need nx, ny, nt >= 2, dx, dy > 0 and all fields");
    if (!sorted(tgrid, nt)) return c->fail(LMH_E_INVALID, "set_medium_grid2p1: tgrid must be sorted");
    LMH_HIP(c, hipSetDevice(c->device));
    LMH_HIP(c, hipStreamSynchronize(c->stream));
    size_t nxy = (size_t)nx * (size_t)ny, n = nxy * (size_t)nt;
    std::vector<double4> nodes;
    try {
        nodes.resize(n);
    } catch (...) {
        return c->fail(LMH_E_NOMEM, "set_medium_grid2p1: host allocation failed");
    }
    for (size_t k = 0; k < n; ++k) nodes[k] = make_double4(T[k], vx[k], vy[k], 0.0);
    dfree(c->d_nodes);
    dfree(c->d_slice);
    if (hipMalloc((void**)&c->d_nodes, n * sizeof(double4)) != hipSuccess ||
        hipMalloc((void**)&c->d_slice, kSliceRing * 3 * (size_t)(nx - 1) * (size_t)(ny - 1) * sizeof(double4)) != hipSuccess)
        return c->fail(LMH_E_NOMEM, "set_medium_grid2p1: device allocation failed");
    LMH_HIP(c, hipMemcpy(c->d_nodes, nodes.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    int rc = upload(c, &c->d_tg, tgrid, nt);
    if (rc) return rc;
    dfree(c->d_rg);
    dfree(c->d_T);
    dfree(c->d_V);
    dfree(c->d_V2);
    c->gx = nx;
    c->gy = ny;
    c->nt = nt;
    c->gx0 = x0;
    c->gdx = dx;
    c->gy0 = y0;
    c->gdy = dy;
    c->medium = 1;
    return LMH_OK;
}

int lmh_set_time_table(lmh_ctx* c, int32_t which, int32_t species, double Tmin, double invdT,
                       const double* vals, int32_t n) {
    if (!c) return LMH_E_INVALID;
    ++c->gen;
    if (species < 0 || species >= LMH_MAX_SPECIES || (which != LMH_TABLE_DRAG && which != LMH_TABLE_RTA) ||
        !vals || n < 1)
        return c->fail(LMH_E_INVALID, "set_time_table: bad species, table kind or size");
    LMH_HIP(c, hipSetDevice(c->device));
    LMH_HIP(c, hipStreamSynchronize(c->stream));
    lmh_ctx::Tab& t = which == LMH_TABLE_DRAG ? c->drag[species] : c->rta[species];
    int rc = upload(c, &t.d, vals, (size_t)n);
    if (rc) return rc;
    t.Tmin = Tmin;
    t.invdT = invdT;
    t.n = n;
    return upload_species(c);
}

int lmh_set_juttner_table(lmh_ctx* c, int32_t species, const double* invcdf, int32_t nU,
                          int32_t nT, double Tmin, double invdT) {
    if (!c) return LMH_E_INVALID;
    ++c->gen;
    if (species < 0 || species >= LMH_MAX_SPECIES || !invcdf || nU < 2 || nT < 2)
        return c->fail(LMH_E_INVALID, "set_juttner_table: bad species or size");
    LMH_HIP(c, hipSetDevice(c->device));
    LMH_HIP(c, hipStreamSynchronize(c->stream));
    lmh_ctx::JTab& t = c->jt[species];
    int rc = upload(c, &t.d, invcdf, (size_t)nU * (size_t)nT);
    if (rc) return rc;
    t.Tmin = Tmin;
    t.invdT = invdT;
    t.nU = nU;
    t.nT = nT;
    return upload_species(c);
}

int lmh_set_particles(lmh_ctx* c, const double* pos, int32_t dims, const double* mom, int32_t pdim,
                      int64_t N, int32_t nspecies, const int64_t* counts, const double* masses) {
    if (!c) return LMH_E_INVALID;
    ++c->gen;
    if (!pos || !mom || N < 1 || dims < 1 || dims > 3 || pdim < 1 || pdim > 3)
        return c->fail(LMH_E_INVALID, "set_particles: need N >= 1 and 1 <= dims, pdim <= 3");
    if (N > INT32_MAX) return c->fail(LMH_E_INVALID, "set_particles: at most 2^31 - 1 particles per context");
    if (nspecies < 1 || nspecies > LMH_MAX_SPECIES || !counts || !masses)
        return c->fail(LMH_E_INVALID, "set_particles: need 1..4 species with counts and masses");
    int64_t sum = 0;
    for (int k = 0; k < nspecies; ++k) {
        if (counts[k] < 0 || !std::isfinite(masses[k]))
            return c->fail(LMH_E_INVALID, "set_particles: species counts must be >= 0 and masses finite");
        sum += counts[k];
    }
    if (sum != N) return c->fail(LMH_E_INVALID, "set_particles: species counts must sum to N");
    if (N > (int64_t)INT32_MAX) return c->fail(LMH_E_INVALID, "set_particles: at most 2^31-1 particles per context");
    LMH_HIP(c, hipSetDevice(c->device));
    LMH_HIP(c, hipStreamSynchronize(c->stream));
    size_t rows = (size_t)(dims + pdim);
    std::vector<double> soa;
    try {
        soa.resize(rows * (size_t)N);
    } catch (...) {
        return c->fail(LMH_E_NOMEM, "set_particles: host allocation failed");
    }
    for (int64_t i = 0; i < N; ++i) {
        for (int d = 0; d < dims; ++d) soa[(size_t)d * N + i] = pos[d + (int64_t)dims * i];
        for (int d = 0; d < pdim; ++d) soa[(size_t)(dims + d) * N + i] = mom[d + (int64_t)pdim * i];
    }
    dfree(c->d_state);
    c->is_sorted = false;
    c->perm_pending = false;
    c->steps_since_sort = 0;
    if (c->sort_n != N || (c->d_state_alt && c->dims + c->pdim != dims + pdim)) {
        dfree(c->d_state_alt);
        dfree(c->d_id_alt);
        dfree(c->d_keys);
        dfree(c->d_sort_tmp);
        c->sort_tmp_bytes = 0;
        c->sort_n = 0;
    }
    if (hipMalloc((void**)&c->d_state, rows * (size_t)N * sizeof(double)) != hipSuccess) {
        c->d_state = nullptr;
        c->N = 0;
        return c->fail(LMH_E_NOMEM, "set_particles: device allocation failed");
    }
    LMH_HIP(c, hipMemcpy(c->d_state, soa.data(), rows * (size_t)N * sizeof(double), hipMemcpyHostToDevice));
    // identities: each slot its own until the particles are sorted by cell
    dfree(c->d_id);
    if (hipMalloc((void**)&c->d_id, (size_t)N * sizeof(uint32_t)) != hipSuccess)
        return c->fail(LMH_E_NOMEM, "set_particles: identity allocation failed");
    hipLaunchKernelGGL(k_iota, dim3(grid_blocks(N)), dim3(kBlock), 0, c->stream, c->d_id, N);
    LMH_HIP(c, hipGetLastError());
    LMH_HIP(c, hipStreamSynchronize(c->stream));
    unsigned long long none = ULLONG_MAX;
    LMH_HIP(c, hipMemcpy(c->d_bad, &none, sizeof(none), hipMemcpyHostToDevice));
    c->dims = dims;
    c->pdim = pdim;
    c->N = N;
    c->nspecies = nspecies;
    for (int k = 0; k < LMH_MAX_SPECIES; ++k) {
        c->counts[k] = k < nspecies ? counts[k] : 0;
        c->mass[k] = k < nspecies ? masses[k] : 0.0;
    }
    // a full block region per step-kernel block: the step kernel's blocks cover each species'
    // count rounded up to the block size, so the regions end below N + nspecies * block; at
    // most N/64 + one partial block per species counts
    dfree(c->d_qrec);
    dfree(c->d_qcount);
    const size_t max_blocks = (size_t)(N / 64) + LMH_MAX_SPECIES + 1;
    const size_t max_rec = (size_t)N + (size_t)(LMH_MAX_SPECIES + 1) * kBlock;
    if (hipMalloc((void**)&c->d_qrec, max_rec * 2 * sizeof(double4)) != hipSuccess ||
        hipMalloc((void**)&c->d_qcount, max_blocks * sizeof(int32_t)) != hipSuccess) {
        return c->fail(LMH_E_NOMEM, "set_particles: queue allocation failed");
    }
    if (c->fo_on) {
        int rc2 = fo_reset(c);
        if (rc2) return rc2;
    }
    return upload_species(c);
}

int lmh_get_particles(lmh_ctx* c, double* pos, double* mom) {
    if (!c) return LMH_E_INVALID;
    if (!c->d_state) return c->fail(LMH_E_STATE, "get_particles: no particles set");
    if (!pos || !mom) return c->fail(LMH_E_INVALID, "get_particles: null output");
    LMH_HIP(c, hipSetDevice(c->device));
    if (int rc = ensure_unsorted(c)) return rc;
    size_t rows = (size_t)(c->dims + c->pdim);
    std::vector<double> soa;
    try {
        soa.resize(rows * (size_t)c->N);
    } catch (...) {
        return c->fail(LMH_E_NOMEM, "get_particles: host allocation failed");
    }
    LMH_HIP(c, hipMemcpyAsync(soa.data(), c->d_state, rows * (size_t)c->N * sizeof(double),
                              hipMemcpyDeviceToHost, c->stream));
    LMH_HIP(c, hipStreamSynchronize(c->stream));
    const int64_t N = c->N;
    for (int64_t i = 0; i < N; ++i) {
        for (int d = 0; d < c->dims; ++d) pos[d + (int64_t)c->dims * i] = soa[(size_t)d * N + i];
        for (int d = 0; d < c->pdim; ++d) mom[d + (int64_t)c->pdim * i] = soa[(size_t)(c->dims + d) * N + i];
    }
    return LMH_OK;
}

}  // extern "C"

namespace {

int obs_check(lmh_ctx* c, const lmh_obs_spec* s) {
    if (!s) return c->fail(LMH_E_INVALID, "observables: null spec");
    if (!c->d_state) return c->fail(LMH_E_STATE, "observables: no particles set");
    if (c->pdim < 2) return c->fail(LMH_E_INVALID, "observables: need >= 2 momentum rows");
    if (s->nbins < 0 || s->nbins > kObsMaxBins || (s->nbins > 0 && !(s->pt_max > s->pt_min)))
        return c->fail(LMH_E_INVALID, "observables: 0 <= nbins <= 120 and pt_max > pt_min");
    if (s->species >= c->nspecies) return c->fail(LMH_E_INVALID, "observables: species out of range");
    return LMH_OK;
}

// the partial sums of the selected particles into `partial` (kObsBlocks x (5 + nbins)) and their
// total into `out` (5 + nbins, device memory), on the context's stream, in the current slot order
int obs_enqueue(lmh_ctx* c, const lmh_obs_spec* s, double* partial, double* out) {
    int64_t i0 = 0, i1 = c->N;
    if (s->species >= 0) {
        i0 = species_begin(c, s->species);
        i1 = i0 + c->counts[s->species];
    }
    const int K = 5 + s->nbins;
    const double* px = c->d_state + (int64_t)c->dims * c->N;
    const double* py = px + c->N;
    hipLaunchKernelGGL(k_obs_partial, dim3(kObsBlocks), dim3(kObsBlock), 0, c->stream, px, py, i0, i1, *s, partial);
    hipLaunchKernelGGL(k_obs_final, dim3(K), dim3(kObsFinalBlock), 0, c->stream, partial, kObsBlocks, K, out);
    LMH_HIP(c, hipGetLastError());
    return LMH_OK;
}

int check_run(lmh_ctx* c, const lmh_params* p) {
    if (!p) return c->fail(LMH_E_INVALID, "null params");
    if (!c->d_state) return c->fail(LMH_E_STATE, "no particles set");
    if (c->medium < 0) return c->fail(LMH_E_STATE, "no medium set");
    if (!(p->dt == p->dt)) return c->fail(LMH_E_INVALID, "dt is NaN");
    if (p->noise != LMH_NOISE_FIXED && p->noise != LMH_NOISE_PHILOX)
        return c->fail(LMH_E_INVALID, "unknown noise mode");
    if (p->collision < LMH_LANGEVIN || p->collision > LMH_LANGEVIN_RTA)
        return c->fail(LMH_E_INVALID, "unknown collision mode");
    if (p->integrator < LMH_PREPOINT || p->integrator > LMH_POSTPOINT)
        return c->fail(LMH_E_INVALID, "unknown integrator");
    if (p->drag_model != LMH_DRAG_TABLE && p->drag_model != LMH_DRAG_GAMMA_P)
        return c->fail(LMH_E_INVALID, "unknown drag model");
    if (p->coords == LMH_GENERAL_COORDS) {
        if (c->medium != 0) return c->fail(LMH_E_STATE, "general coordinates need the radial medium");
        if (c->fo_on) return c->fail(LMH_E_INVALID, "freeze-out capture runs in Cartesian coordinates");
        if (c->dims != 2 || c->pdim != 2)
            return c->fail(LMH_E_INVALID, "general coordinates need dims == pdim == 2 (tau, r)");
        if (p->collision != LMH_LANGEVIN || p->drag_model != LMH_DRAG_TABLE ||
            p->integrator != LMH_PREPOINT || p->position_diffusion || p->reflecting_boundary ||
            p->bjorken_redshift)
            return c->fail(LMH_E_INVALID, "general coordinates support the pre-point Langevin step only");
    } else if (p->coords == LMH_MILNE) {
        if (c->medium != 0) return c->fail(LMH_E_STATE, "Milne coordinates need the radial medium");
        if (c->dims != 3 || c->pdim != 3)
            return c->fail(LMH_E_INVALID, "Milne coordinates need dims == pdim == 3 (x, y, eta_s)");
        if (!(p->t0 > 0.0)) return c->fail(LMH_E_INVALID, "Milne coordinates need t0 > 0 (tau)");
        if (p->position_diffusion || p->reflecting_boundary || c->fo_on)
            return c->fail(LMH_E_INVALID, "Milne coordinates: no position diffusion, reflecting wall or freeze-out capture");
    } else if (p->coords == LMH_CARTESIAN) {
        bool ok = (c->dims == c->pdim) || (c->dims == 2 && c->pdim == 3);
        if (!ok) return c->fail(LMH_E_INVALID, "momentum_dimensions with dimensions is not supported (utils.jl:561-573)");
        if (p->bjorken_redshift && c->pdim != 3)
            return c->fail(LMH_E_INVALID, "bjorken_redshift requires momentum row 3");
        if (c->medium == 1 && (c->dims != 2 || p->reflecting_boundary))
            return c->fail(LMH_E_INVALID, "the 2+1D grid needs dims == 2 and no reflecting boundary");
    } else {
        return c->fail(LMH_E_INVALID, "unknown coordinates");
    }
    bool uses_drag = p->collision == LMH_LANGEVIN || p->collision == LMH_LANGEVIN_RTA;
    bool uses_rta = p->collision == LMH_RTA || p->collision == LMH_LANGEVIN_RTA;
    for (int k = 0; k < c->nspecies; ++k) {
        if (c->counts[k] == 0) continue;
        if (uses_drag && p->drag_model == LMH_DRAG_TABLE && p->DsT > 0.0 && !c->drag[k].d)
            return c->fail(LMH_E_STATE, "no drag table for species " + std::to_string(k));
        if (uses_rta && (!c->rta[k].d || !c->jt[k].d))
            return c->fail(LMH_E_STATE, "RTA needs a relaxation-time table and a Juttner table for species " + std::to_string(k));
    }
    return LMH_OK;
}

// the Philox counter holds the step and the particle index in 32 bits and the stream in 61
int check_steps(lmh_ctx* c, const lmh_params* p, int64_t step_begin, int64_t nsteps) {
    if (step_begin < 0 || nsteps < 0 || step_begin + nsteps > ((int64_t)1 << 32))
        return c->fail(LMH_E_INVALID, "steps must lie in [0, 2^32)");
    if (p->noise == LMH_NOISE_PHILOX && p->stream >= ((uint64_t)1 << 61))
        return c->fail(LMH_E_INVALID, "the Philox stream must be below 2^61");
    return LMH_OK;
}

// the parameters are those the SPEC kernels fix at compile time, and every species' drag and
// relaxation-time tables (where set) have two nodes or more
bool spec_config(const lmh_ctx* c, const lmh_params* p) {
    for (int k = 0; k < c->nspecies; ++k) {
        const SpeciesArgs& s = c->h_species[k];
        if ((s.drag.vals && s.drag.n < 2) || (s.rta.vals && s.rta.n < 2)) return false;
    }
    return c->tune.specialize && !c->fo_on && c->N < ((int64_t)1 << 29) && p->coords == LMH_CARTESIAN &&
           p->relativistic &&
           p->integrator == LMH_PREPOINT &&
           p->drag_model == LMH_DRAG_TABLE && !p->bjorken_redshift && !p->position_diffusion;
}

template <int DIMS, int PDIM, int NOISE, bool SPEC>
void launch_radial_n(lmh_ctx* c, const StepArgs& a, unsigned nb, const RadialMedium& med, int coll) {
    const dim3 G(nb), B(c->tune.block);
    switch (coll) {
        case LMH_LANGEVIN: hipLaunchKernelGGL((k_cart_radial<DIMS, PDIM, NOISE, LMH_LANGEVIN, SPEC>), G, B, 0, c->stream, a, med); break;
        case LMH_RTA: hipLaunchKernelGGL((k_cart_radial<DIMS, PDIM, NOISE, LMH_RTA, SPEC>), G, B, 0, c->stream, a, med); break;
        case LMH_FLUID: hipLaunchKernelGGL((k_cart_radial<DIMS, PDIM, NOISE, LMH_FLUID, SPEC>), G, B, 0, c->stream, a, med); break;
        default: hipLaunchKernelGGL((k_cart_radial<DIMS, PDIM, NOISE, LMH_LANGEVIN_RTA, SPEC>), G, B, 0, c->stream, a, med); break;
    }
}

template <int DIMS, int PDIM>
void launch_radial(lmh_ctx* c, const StepArgs& a, unsigned nb, const RadialMedium& med, int noise, int coll, bool spec) {
    if (noise == LMH_NOISE_FIXED) {
        if (spec) launch_radial_n<DIMS, PDIM, LMH_NOISE_FIXED, true>(c, a, nb, med, coll);
        else launch_radial_n<DIMS, PDIM, LMH_NOISE_FIXED, false>(c, a, nb, med, coll);
    } else {
        if (spec) launch_radial_n<DIMS, PDIM, LMH_NOISE_PHILOX, true>(c, a, nb, med, coll);
        else launch_radial_n<DIMS, PDIM, LMH_NOISE_PHILOX, false>(c, a, nb, med, coll);
    }
}

// the specialised Milne kernel's configurations: (pre-point, drag table) and (post-point,
// gamma(p, T)), relativistic, no position diffusion, wall, freeze-out capture or elliptic term
bool milne_spec(const lmh_ctx* c, const lmh_params* p) {
    const bool combo = (p->integrator == LMH_PREPOINT && p->drag_model == LMH_DRAG_TABLE) ||
                       (p->integrator == LMH_POSTPOINT && p->drag_model == LMH_DRAG_GAMMA_P);
    return c->tune.specialize && c->medium == 0 && c->dims == 3 && c->pdim == 3 && p->coords == LMH_MILNE &&
           p->relativistic && !p->position_diffusion && !p->reflecting_boundary && !c->fo_on &&
           c->d_V2 == nullptr && combo;
}

template <int NOISE, int INTEG, int DRAG>
void launch_milne_c(lmh_ctx* c, const StepArgs& a, unsigned nb, const RadialMedium& med, int coll) {
    const dim3 G(nb), B(c->tune.block);
    switch (coll) {
        case LMH_LANGEVIN: hipLaunchKernelGGL((k_milne<NOISE, LMH_LANGEVIN, INTEG, DRAG>), G, B, 0, c->stream, a, med); break;
        case LMH_RTA: hipLaunchKernelGGL((k_milne<NOISE, LMH_RTA, INTEG, DRAG>), G, B, 0, c->stream, a, med); break;
        case LMH_FLUID: hipLaunchKernelGGL((k_milne<NOISE, LMH_FLUID, INTEG, DRAG>), G, B, 0, c->stream, a, med); break;
        default: hipLaunchKernelGGL((k_milne<NOISE, LMH_LANGEVIN_RTA, INTEG, DRAG>), G, B, 0, c->stream, a, med); break;
    }
}

template <int NOISE>
void launch_milne_n(lmh_ctx* c, const StepArgs& a, unsigned nb, const RadialMedium& med, const lmh_params* p) {
    if (p->integrator == LMH_POSTPOINT)
        launch_milne_c<NOISE, LMH_POSTPOINT, LMH_DRAG_GAMMA_P>(c, a, nb, med, p->collision);
    else
        launch_milne_c<NOISE, LMH_PREPOINT, LMH_DRAG_TABLE>(c, a, nb, med, p->collision);
}

void launch_milne(lmh_ctx* c, const StepArgs& a, unsigned nb, const RadialMedium& med, const lmh_params* p) {
    if (p->noise == LMH_NOISE_FIXED) launch_milne_n<LMH_NOISE_FIXED>(c, a, nb, med, p);
    else launch_milne_n<LMH_NOISE_PHILOX>(c, a, nb, med, p);
}

template <int PDIM, int NOISE, bool SLICE, bool SPEC>
void launch_grid_s(lmh_ctx* c, const StepArgs& a, unsigned nb, const GridMedium& g, int coll, bool defer) {
    const dim3 G(nb), B(c->tune.block);
    {
    LaunchMark mk(c, LMH_PROF_STEP);
    if (coll == LMH_LANGEVIN) {
        hipLaunchKernelGGL((k_cart_grid<PDIM, NOISE, SLICE, LMH_LANGEVIN, false, SPEC>), G, B, 0, c->stream, a, g);
    } else if (coll == LMH_FLUID) {
        hipLaunchKernelGGL((k_cart_grid<PDIM, NOISE, SLICE, LMH_FLUID, false, SPEC>), G, B, 0, c->stream, a, g);
    } else if (coll == LMH_RTA) {
        if (defer) hipLaunchKernelGGL((k_cart_grid<PDIM, NOISE, SLICE, LMH_RTA, true, SPEC>), G, B, 0, c->stream, a, g);
        else hipLaunchKernelGGL((k_cart_grid<PDIM, NOISE, SLICE, LMH_RTA, false, SPEC>), G, B, 0, c->stream, a, g);
    } else {
        if (defer) hipLaunchKernelGGL((k_cart_grid<PDIM, NOISE, SLICE, LMH_LANGEVIN_RTA, true, SPEC>), G, B, 0, c->stream, a, g);
        else hipLaunchKernelGGL((k_cart_grid<PDIM, NOISE, SLICE, LMH_LANGEVIN_RTA, false, SPEC>), G, B, 0, c->stream, a, g);
    }
    }
    if (defer) {
        LaunchMark mk(c, LMH_PROF_REDRAW);
        // one wave per kRedrawSpan step-kernel blocks
        const int64_t nwaves = ((int64_t)nb + kRedrawSpan - 1) / kRedrawSpan;
        const unsigned nr = (unsigned)std::max<int64_t>(1, (nwaves + (kBlock / 64) - 1) / (kBlock / 64));
        hipLaunchKernelGGL((k_redraw_grid<PDIM, NOISE, SLICE, SPEC>), dim3(nr), dim3(kBlock), 0, c->stream, a, g, (int64_t)nb);
    }
}

// the specialised sliced step with in-kernel collisions, reading through a sort's permutation
// (timed by the caller with the sort)
template <int PDIM>
void launch_grid_gather(lmh_ctx* c, const StepArgs& a, unsigned nb, const GridMedium& g, int coll) {
    const dim3 G(nb), B(c->tune.block);
    constexpr int P = LMH_NOISE_PHILOX;
    switch (coll) {
        case LMH_LANGEVIN: hipLaunchKernelGGL((k_cart_grid<PDIM, P, true, LMH_LANGEVIN, false, true, true>), G, B, 0, c->stream, a, g); break;
        case LMH_RTA: hipLaunchKernelGGL((k_cart_grid<PDIM, P, true, LMH_RTA, false, true, true>), G, B, 0, c->stream, a, g); break;
        case LMH_FLUID: hipLaunchKernelGGL((k_cart_grid<PDIM, P, true, LMH_FLUID, false, true, true>), G, B, 0, c->stream, a, g); break;
        default: hipLaunchKernelGGL((k_cart_grid<PDIM, P, true, LMH_LANGEVIN_RTA, false, true, true>), G, B, 0, c->stream, a, g); break;
    }
}

template <int PDIM, bool SPEC>
void launch_grid_n(lmh_ctx* c, const StepArgs& a, unsigned nb, const GridMedium& g, int noise, bool slice,
                   int coll, bool defer) {
    if (noise == LMH_NOISE_FIXED) {
        if (slice) launch_grid_s<PDIM, LMH_NOISE_FIXED, true, SPEC>(c, a, nb, g, coll, defer);
        else launch_grid_s<PDIM, LMH_NOISE_FIXED, false, SPEC>(c, a, nb, g, coll, defer);
    } else {
        if (slice) launch_grid_s<PDIM, LMH_NOISE_PHILOX, true, SPEC>(c, a, nb, g, coll, defer);
        else launch_grid_s<PDIM, LMH_NOISE_PHILOX, false, SPEC>(c, a, nb, g, coll, defer);
    }
}

template <int PDIM>
void launch_grid(lmh_ctx* c, const StepArgs& a, unsigned nb, const GridMedium& g, int noise, bool slice,
                 int coll, bool defer, bool spec) {
    if (spec) launch_grid_n<PDIM, true>(c, a, nb, g, noise, slice, coll, defer);
    else launch_grid_n<PDIM, false>(c, a, nb, g, noise, slice, coll, defer);
}

// cell sorting applies to the 2+1D grid with Philox noise (the noise follows the identity) and
// without freeze-out capture (its records are per slot); anything else runs in the caller's order
bool run_sortable(const lmh_ctx* c, const lmh_params* p) {
    return c->medium == 1 && p->coords == LMH_CARTESIAN && p->noise == LMH_NOISE_PHILOX && !c->fo_on &&
           c->tune.sort_every > 0;
}

int run_impl(lmh_ctx* c, const lmh_params* p, int64_t step_begin, int64_t nsteps,
             const double* d_noise, int32_t spl) {
    const bool grid = c->medium == 1 && p->coords == LMH_CARTESIAN;
    const int32_t k = spl > 0 ? spl : std::max(1, c->tune.steps_per_launch);
    RadialMedium rmed = make_radial(c);
    GridMedium gmed = make_grid(c);
    const bool sortable = run_sortable(c, p);
    if (!sortable) {
        const int rc = ensure_unsorted(c);
        if (rc) return rc;
    }
    unsigned nb = 0;
    StepArgs a = make_args(c, p, c->tune.block, &nb);
    if (nb == 0) return LMH_OK;
    const bool has_rta = p->collision == LMH_RTA || p->collision == LMH_LANGEVIN_RTA;
    const bool defer = grid && k == 1 && has_rta && c->tune.defer && !c->fo_on;
    const bool spec = spec_config(c, p);
    int64_t ring0 = -1;       // first step of the slices in the ring
    for (int64_t s0 = 0; s0 < nsteps; s0 += k) {
        const int32_t n = (int32_t)std::min<int64_t>(k, nsteps - s0);
        const bool slice = grid && n == 1 && c->tune.slice;
        // the specialised sliced step with in-kernel collisions applies a sort's permutation by
        // reading through it (GATHER) instead of a permutation pass of its own
        const bool gather_ok = grid && slice && spec && !defer && p->noise == LMH_NOISE_PHILOX;
        GridMedium gstep = gmed;
        if (slice) {
            // the slices of the next kSliceRing steps in one launch, then one per step from the ring
            const int64_t rec = 3 * (int64_t)(c->gx - 1) * (c->gy - 1);
            if (ring0 < 0 || s0 - ring0 >= kSliceRing) {
                ring0 = s0;
                const int32_t nk = (int32_t)std::min<int64_t>(kSliceRing, nsteps - s0);
                LaunchMark mk(c, LMH_PROF_SLICE);
                hipLaunchKernelGGL(k_grid_slices, dim3(grid_blocks(rec * nk)), dim3(kBlock), 0, c->stream, gmed,
                                   c->d_slice, step_begin + s0, nk, p->dt, p->t0);
            }
            gstep.slice = c->d_slice + rec * (s0 - ring0);
        }
        // a sort that the step applies is timed with that step, as one sort entry
        std::optional<LaunchMark> sort_mark;
        if (sortable && (!c->is_sorted || c->steps_since_sort >= c->tune.sort_every)) {
            if (gather_ok) sort_mark.emplace(c, LMH_PROF_SORT);
            const int rc = sort_by_cell(c, gather_ok, !gather_ok);
            if (rc) return rc;
            a = make_args(c, p, c->tune.block, &nb);
        }
        a.step_begin = step_begin + s0;
        a.nsteps = n;
        a.noise = d_noise ? d_noise + s0 * (int64_t)a.C * c->N : nullptr;
        if (p->coords == LMH_GENERAL_COORDS) {
            LaunchMark mk(c, LMH_PROF_STEP);
            if (p->noise == LMH_NOISE_FIXED)
                hipLaunchKernelGGL((k_gc<LMH_NOISE_FIXED>), dim3(nb), dim3(c->tune.block), 0, c->stream, a, rmed);
            else
                hipLaunchKernelGGL((k_gc<LMH_NOISE_PHILOX>), dim3(nb), dim3(c->tune.block), 0, c->stream, a, rmed);
        } else if (grid && c->perm_pending && gather_ok) {
            // read the old order through the permutation, write the new order into the alternate
            // buffers, then make them current
            StepArgs ag = a;
            const int64_t N = c->N;
            for (int d = 0; d < 3; ++d) {
                ag.wpos[d] = d < c->dims ? c->d_state_alt + (int64_t)d * N : nullptr;
                ag.wmom[d] = d < c->pdim ? c->d_state_alt + (int64_t)(c->dims + d) * N : nullptr;
            }
            ag.perm = c->d_keys + 3 * N;
            ag.wid = c->d_id_alt;
            if (c->pdim == 2) launch_grid_gather<2>(c, ag, nb, gstep, p->collision);
            else launch_grid_gather<3>(c, ag, nb, gstep, p->collision);
            sort_mark.reset();
            swap_order(c);
            c->perm_pending = false;
            a = make_args(c, p, c->tune.block, &nb);
        } else if (grid) {
            if (c->perm_pending) {
                const int rc = finish_permutation(c);
                if (rc) return rc;
                a = make_args(c, p, c->tune.block, &nb);
            }
            if (c->pdim == 2) launch_grid<2>(c, a, nb, gstep, p->noise, slice, p->collision, defer && n == 1, spec);
            else launch_grid<3>(c, a, nb, gstep, p->noise, slice, p->collision, defer && n == 1, spec);
        } else if (milne_spec(c, p)) {
            LaunchMark mk(c, LMH_PROF_STEP);
            launch_milne(c, a, nb, rmed, p);
        } else {
            LaunchMark mk(c, LMH_PROF_STEP);
            if (c->dims == 1) launch_radial<1, 1>(c, a, nb, rmed, p->noise, p->collision, spec);
            else if (c->dims == 2 && c->pdim == 2) launch_radial<2, 2>(c, a, nb, rmed, p->noise, p->collision, spec);
            else if (c->dims == 2 && c->pdim == 3) launch_radial<2, 3>(c, a, nb, rmed, p->noise, p->collision, spec);
            else launch_radial<3, 3>(c, a, nb, rmed, p->noise, p->collision, spec);
        }
        LMH_HIP(c, hipGetLastError());
        c->steps_since_sort += n;
    }
    {
        const int rc = finish_permutation(c);
        if (rc) return rc;
    }
    if (p->check_finite) {
        unsigned long long bad = ULLONG_MAX;
        LMH_HIP(c, hipMemcpyAsync(&bad, c->d_bad, sizeof(bad), hipMemcpyDeviceToHost, c->stream));
        LMH_HIP(c, hipStreamSynchronize(c->stream));
        if (bad != ULLONG_MAX)
            return c->fail(LMH_E_NONFINITE, "Breaking simulation due to NaN/Inf at step " + std::to_string((long long)bad));
    }
    return LMH_OK;
}

}  // namespace

extern "C" {

int lmh_set_tuning(lmh_ctx* c, const lmh_tuning* t) {
    if (!c) return LMH_E_INVALID;
    if (!t || (t->block != 64 && t->block != 128 && t->block != 256) || t->steps_per_launch < 1 ||
        (t->slice != 0 && t->slice != 1) || (t->defer != 0 && t->defer != 1) ||
        (t->specialize != 0 && t->specialize != 1) || t->sort_every < 0)
        return c->fail(LMH_E_INVALID, "set_tuning: block in {64, 128, 256}, steps_per_launch >= 1, slice, defer "
                                      "and specialize in {0, 1}, sort_every >= 0");
    c->tune = *t;
    return LMH_OK;
}

int lmh_get_tuning(lmh_ctx* c, lmh_tuning* t) {
    if (!c || !t) return LMH_E_INVALID;
    *t = c->tune;
    return LMH_OK;
}

int lmh_profile_start(lmh_ctx* c) {
    if (!c) return LMH_E_INVALID;
    c->profiling = true;
    c->ev_used = 0;
    c->marks.clear();
    return LMH_OK;
}

int lmh_profile_stop(lmh_ctx* c, lmh_profile* out) {
    if (!c) return LMH_E_INVALID;
    if (!out) return c->fail(LMH_E_INVALID, "profile_stop: null output");
    LMH_HIP(c, hipSetDevice(c->device));
    LMH_HIP(c, hipStreamSynchronize(c->stream));
    memset(out, 0, sizeof(*out));
    for (const auto& m : c->marks) {
        float ms = 0.0f;
        LMH_HIP(c, hipEventElapsedTime(&ms, m.a, m.b));
        out->launches[m.cls] += 1;
        out->ms[m.cls] += ms;
    }
    c->profiling = false;
    c->marks.clear();
    c->ev_used = 0;
    return LMH_OK;
}

int lmh_boost_to_lab(lmh_ctx* c, const lmh_params* p, int64_t step) {
    if (!c) return LMH_E_INVALID;
    ++c->gen;
    int rc = check_run(c, p);
    if (rc) return rc;
    if (p->coords == LMH_GENERAL_COORDS) return c->fail(LMH_E_INVALID, "boost_to_lab: Cartesian or Milne only");
    LMH_HIP(c, hipSetDevice(c->device));
    if ((rc = ensure_unsorted(c))) return rc;
    RadialMedium rmed = make_radial(c);
    GridMedium gmed = make_grid(c);
    unsigned nb = 0;
    StepArgs a = make_args(c, p, c->tune.block, &nb);
    if (nb == 0) return LMH_OK;
    const int B = c->tune.block;
    if (c->medium == 1) {
        if (c->pdim == 2) hipLaunchKernelGGL((k_boost_lab_grid<2>), dim3(nb), dim3(B), 0, c->stream, a, gmed, step);
        else hipLaunchKernelGGL((k_boost_lab_grid<3>), dim3(nb), dim3(B), 0, c->stream, a, gmed, step);
    } else {
        if (c->dims == 1) hipLaunchKernelGGL((k_boost_lab_radial<1, 1>), dim3(nb), dim3(B), 0, c->stream, a, rmed, step);
        else if (c->pdim == 2) hipLaunchKernelGGL((k_boost_lab_radial<2, 2>), dim3(nb), dim3(B), 0, c->stream, a, rmed, step);
        else if (c->dims == 2) hipLaunchKernelGGL((k_boost_lab_radial<2, 3>), dim3(nb), dim3(B), 0, c->stream, a, rmed, step);
        else hipLaunchKernelGGL((k_boost_lab_radial<3, 3>), dim3(nb), dim3(B), 0, c->stream, a, rmed, step);
    }
    LMH_HIP(c, hipGetLastError());
    return LMH_OK;
}

int lmh_run(lmh_ctx* c, const lmh_params* p, int64_t step_begin, int64_t nsteps,
            const double* noise, int32_t steps_per_launch) {
    if (!c) return LMH_E_INVALID;
    ++c->gen;
    int rc = check_run(c, p);
    if (rc) return rc;
    if (nsteps < 0 || steps_per_launch < 0) return c->fail(LMH_E_INVALID, "nsteps and steps_per_launch must be >= 0");
    if ((rc = check_steps(c, p, step_begin, nsteps))) return rc;
    if (nsteps == 0) return LMH_OK;
    LMH_HIP(c, hipSetDevice(c->device));
    const double* d_noise = nullptr;
    if (p->noise == LMH_NOISE_FIXED) {
        if (!noise) return c->fail(LMH_E_INVALID, "fixed-noise mode needs a noise record");
        int a, b, d;
        size_t C = (size_t)noise_layout(p, c->dims, c->pdim, &a, &b, &d);
        size_t n = C * (size_t)nsteps * (size_t)c->N;
        if (n > c->noise_cap) {
            LMH_HIP(c, hipStreamSynchronize(c->stream));
            dfree(c->d_noise);
            c->noise_cap = 0;
            if (hipMalloc((void**)&c->d_noise, n * sizeof(double)) != hipSuccess) {
                c->d_noise = nullptr;
                return c->fail(LMH_E_NOMEM, "run: noise allocation failed");
            }
            c->noise_cap = n;
        }
        if (n) {
            LMH_HIP(c, hipMemcpyAsync(c->d_noise, noise, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
            LMH_HIP(c, hipStreamSynchronize(c->stream));
        }
        d_noise = c->d_noise;
    }
    return run_impl(c, p, step_begin, nsteps, d_noise, steps_per_launch);
}

int lmh_run_device_noise(lmh_ctx* c, const lmh_params* p, int64_t step_begin, int64_t nsteps,
                         const double* d_noise, int32_t steps_per_launch) {
    if (!c) return LMH_E_INVALID;
    ++c->gen;
    int rc = check_run(c, p);
    if (rc) return rc;
    if (nsteps < 0 || steps_per_launch < 0) return c->fail(LMH_E_INVALID, "nsteps and steps_per_launch must be >= 0");
    if (p->noise == LMH_NOISE_FIXED && !d_noise) return c->fail(LMH_E_INVALID, "fixed-noise mode needs a noise record");
    if ((rc = check_steps(c, p, step_begin, nsteps))) return rc;
    if (nsteps == 0) return LMH_OK;
    LMH_HIP(c, hipSetDevice(c->device));
    return run_impl(c, p, step_begin, nsteps, d_noise, steps_per_launch);
}

int64_t lmh_nonfinite_step(lmh_ctx* c) {
    if (!c || !c->d_bad) return -1;
    unsigned long long bad = ULLONG_MAX;
    if (hipSetDevice(c->device) != hipSuccess) return -1;
    if (hipMemcpyAsync(&bad, c->d_bad, sizeof(bad), hipMemcpyDeviceToHost, c->stream) != hipSuccess) return -1;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return -1;
    return bad == ULLONG_MAX ? -1 : (int64_t)bad;
}

int lmh_observables(lmh_ctx* c, const lmh_obs_spec* s, double* out) {
    if (!c) return LMH_E_INVALID;
    if (!out) return c->fail(LMH_E_INVALID, "observables: null output");
    int rc = obs_check(c, s);
    if (rc) return rc;
    LMH_HIP(c, hipSetDevice(c->device));
    if ((rc = ensure_unsorted(c))) return rc;
    const int K = 5 + s->nbins;
    const size_t need = (size_t)K * (kObsBlocks + 1);
    if (need > c->obs_cap) {
        LMH_HIP(c, hipStreamSynchronize(c->stream));
        dfree(c->d_obs);
        c->obs_cap = 0;
        if (hipMalloc((void**)&c->d_obs, need * sizeof(double)) != hipSuccess) {
            c->d_obs = nullptr;
            return c->fail(LMH_E_NOMEM, "observables: allocation failed");
        }
        c->obs_cap = need;
    }
    double* fin = c->d_obs + (size_t)K * kObsBlocks;
    if ((rc = obs_enqueue(c, s, c->d_obs, fin))) return rc;
    LMH_HIP(c, hipMemcpyAsync(out, fin, K * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    LMH_HIP(c, hipStreamSynchronize(c->stream));
    return LMH_OK;
}

int lmh_observables_noalloc(lmh_ctx* c, const lmh_obs_spec* s, double* d_out) {
    if (!c) return LMH_E_INVALID;
    if (!d_out) return c->fail(LMH_E_INVALID, "observables_noalloc: null output");
    int rc = obs_check(c, s);
    if (rc) return rc;
    if ((size_t)(5 + s->nbins) * kObsBlocks > kScratchDoubles)
        return c->fail(LMH_E_INVALID, "observables_noalloc: the partial sums exceed the scratch array");
    if (!c->d_scratch) return c->fail(LMH_E_STATE, "observables_noalloc: no scratch address for this device");
    LMH_HIP(c, hipSetDevice(c->device));
    return obs_enqueue(c, s, c->d_scratch, d_out);
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// initial-state samplers (utils.jl:33-172, 499-558): tables and inputs copied to one device
// allocation for the call, one thread per quark, results copied back
extern "C" {

int lmh_sample_fonll(lmh_ctx* c, const lmh_fonll_tables* tb, int64_t N, int32_t antithetic,
                     uint64_t seed, uint64_t stream, double* x_out, double* p_out) {
    if (!c) return LMH_E_INVALID;
    if (!tb || !x_out || !p_out || !tb->r_grid || !tb->p_grid || !tb->cdf_p || !tb->p_ok)
        return c->fail(LMH_E_INVALID, "sample_fonll: null tables or output");
    if (tb->nr < 2 || tb->np < 2 || N < 0 || N >= ((int64_t)1 << 32))
        return c->fail(LMH_E_INVALID, "sample_fonll: need nr, np >= 2 and 0 <= N < 2^32");
    if (tb->cartesian ? (!tb->n_r || !(tb->rmax > 0.0) || !(tb->fmax > 0.0)) : !tb->cdf_r)
        return c->fail(LMH_E_INVALID, "sample_fonll: the polar sampler needs cdf_r, the Cartesian one n_r, "
                                      "rmax > 0 and fmax > 0");
    if (!sorted(tb->r_grid, tb->nr) || !sorted(tb->p_grid, tb->np))
        return c->fail(LMH_E_INVALID, "sample_fonll: the grids must be ascending");
    if (tb->cartesian) {
        // the rejection loop accepts nothing from a density that vanishes everywhere
        double nmax = 0.0;
        for (int32_t k = 0; k < tb->nr; ++k) nmax = std::isfinite(tb->n_r[k]) ? std::max(nmax, tb->n_r[k]) : nmax;
        if (!(nmax > 0.0)) return c->fail(LMH_E_INVALID, "sample_fonll: the spatial density n(r) vanishes everywhere");
    }
    if (N == 0) return LMH_OK;
    LMH_HIP(c, hipSetDevice(c->device));
    const size_t nr = (size_t)tb->nr, np = (size_t)tb->np, n = (size_t)N;
    const size_t ntab = 5 * nr + np + np * nr;
    double* d = nullptr;
    int32_t* d_fail = nullptr;
    if (hipMalloc((void**)&d, (ntab + 4 * n) * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&d_fail, sizeof(int32_t)) != hipSuccess) {
        dfree(d);
        return c->fail(LMH_E_NOMEM, "sample_fonll: device allocation failed");
    }
    std::vector<double> h;
    try {
        h.assign(ntab, 0.0);
    } catch (...) {
        dfree(d);
        dfree(d_fail);
        return c->fail(LMH_E_NOMEM, "sample_fonll: host allocation failed");
    }
    double* rg = h.data();
    double* cdf_r = rg + nr;
    double* n_r = cdf_r + nr;
    double* p_ok = n_r + nr;
    double* pg = p_ok + nr;
    double* cdf_p = pg + np;
    memcpy(rg, tb->r_grid, nr * sizeof(double));
    if (tb->cdf_r) memcpy(cdf_r, tb->cdf_r, nr * sizeof(double));
    if (tb->n_r) memcpy(n_r, tb->n_r, nr * sizeof(double));
    memcpy(p_ok, tb->p_ok, nr * sizeof(double));
    memcpy(pg, tb->p_grid, np * sizeof(double));
    memcpy(cdf_p, tb->cdf_p, np * nr * sizeof(double));
    int rc = LMH_OK;
    const hipError_t e1 = hipMemcpyAsync(d, h.data(), ntab * sizeof(double), hipMemcpyHostToDevice, c->stream);
    const hipError_t e2 = hipMemsetAsync(d_fail, 0, sizeof(int32_t), c->stream);
    if (e1 != hipSuccess || e2 != hipSuccess) rc = c->fail(LMH_E_HIP, "sample_fonll: table upload failed");
    if (!rc) {
        FonllTables t;
        t.rg = d;
        t.cdf_r = d + nr;
        t.fsp = d + 2 * nr;
        t.p_ok = d + 3 * nr;
        t.pg = d + 4 * nr;
        t.cdf_p = d + 4 * nr + np;
        t.nr = tb->nr;
        t.np = tb->np;
        t.cartesian = tb->cartesian ? 1 : 0;
        t.antithetic = antithetic ? 1 : 0;
        t.rmax = tb->rmax;
        t.fmax = tb->fmax;
        t.N = N;
        t.nhalf = antithetic ? N / 2 : 0;
        t.seed = seed;
        t.stream = stream;
        t.max_attempts = 100000;
        double* dx = d + ntab;
        double* dp = dx + 2 * n;
        hipLaunchKernelGGL(k_sample_fonll, dim3(grid_blocks(N)), dim3(kBlock), 0, c->stream, t, dx, dp, d_fail);
        int32_t fail = 0;
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(x_out, dx, 2 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipMemcpyAsync(p_out, dp, 2 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipMemcpyAsync(&fail, d_fail, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = c->fail(LMH_E_HIP, "sample_fonll: launch or download failed");
        else if (fail)
            rc = c->fail(LMH_E_STATE, "sample_fonll: the Cartesian rejection sampler accepted no point within "
                                      "100000 attempts for some quark");
    }
    (void)hipStreamSynchronize(c->stream);
    dfree(d);
    dfree(d_fail);
    return rc;
}

int lmh_append_thermal_pz(lmh_ctx* c, const double* mom, const double* pos, int32_t dims, int64_t N,
                          double m, double t0, int32_t antithetic, uint64_t seed, uint64_t stream,
                          double* out) {
    if (!c) return LMH_E_INVALID;
    if (!mom || !pos || !out || dims < 1 || dims > 3 || N < 0 || N >= ((int64_t)1 << 32) || !std::isfinite(m))
        return c->fail(LMH_E_INVALID, "append_thermal_pz: need momenta, positions, 1 <= dims <= 3, 0 <= N < 2^32 "
                                      "and a finite mass");
    if (c->medium != 0) return c->fail(LMH_E_STATE, "append_thermal_pz: needs the radial medium (T(r, t0))");
    if (N == 0) return LMH_OK;
    LMH_HIP(c, hipSetDevice(c->device));
    const size_t n = (size_t)N;
    double* d = nullptr;
    if (hipMalloc((void**)&d, (2 + (size_t)dims + 3) * n * sizeof(double)) != hipSuccess)
        return c->fail(LMH_E_NOMEM, "append_thermal_pz: device allocation failed");
    double* dm = d;
    double* dx = dm + 2 * n;
    double* dout = dx + (size_t)dims * n;
    int rc = LMH_OK;
    if (hipMemcpyAsync(dm, mom, 2 * n * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(dx, pos, (size_t)dims * n * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess) {
        rc = c->fail(LMH_E_HIP, "append_thermal_pz: upload failed");
    } else {
        hipLaunchKernelGGL(k_thermal_pz, dim3(grid_blocks(N)), dim3(kBlock), 0, c->stream, dm, dx, dims, N, m, t0,
                           antithetic ? 1 : 0, seed, stream, make_radial(c), dout);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out, dout, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = c->fail(LMH_E_HIP, "append_thermal_pz: launch or download failed");
    }
    (void)hipStreamSynchronize(c->stream);
    dfree(d);
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// snapshots (simulate_cpu.jl:305-316, simulate_gpu.jl:385-410: the saved states of the history)
extern "C" {

int lmh_snapshot(lmh_ctx* c, int32_t* index) {
    if (!c) return LMH_E_INVALID;
    if (!index) return c->fail(LMH_E_INVALID, "snapshot: null index");
    if (!c->d_state) return c->fail(LMH_E_STATE, "snapshot: no particles set");
    LMH_HIP(c, hipSetDevice(c->device));
    if (int rc = ensure_unsorted(c)) return rc;
    if (!c->copy_stream) LMH_HIP(c, hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    const size_t n = (size_t)(c->dims + c->pdim) * (size_t)c->N;
    int k = -1;
    for (size_t j = 0; j < c->snaps.size(); ++j)
        if (!c->snaps[j].busy && c->snaps[j].n == n) { k = (int)j; break; }
    if (k < 0) {
        lmh_ctx::Snap sn;
        sn.n = n;
        if (hipMalloc((void**)&sn.d, n * sizeof(double)) != hipSuccess) return c->fail(LMH_E_NOMEM, "snapshot: device allocation failed");
        if (hipHostMalloc((void**)&sn.h, n * sizeof(double), hipHostMallocDefault) != hipSuccess) {
            dfree(sn.d);
            return c->fail(LMH_E_NOMEM, "snapshot: pinned host allocation failed");
        }
        if (hipEventCreateWithFlags(&sn.taken, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&sn.done, hipEventDisableTiming) != hipSuccess) {
            dfree(sn.d);
            (void)hipHostFree(sn.h);
            return c->fail(LMH_E_HIP, "snapshot: event creation failed");
        }
        c->snaps.push_back(sn);
        k = (int)c->snaps.size() - 1;
    }
    lmh_ctx::Snap& sn = c->snaps[k];
    LMH_HIP(c, hipMemcpyAsync(sn.d, c->d_state, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    LMH_HIP(c, hipEventRecord(sn.taken, c->stream));
    LMH_HIP(c, hipStreamWaitEvent(c->copy_stream, sn.taken, 0));
    LMH_HIP(c, hipMemcpyAsync(sn.h, sn.d, n * sizeof(double), hipMemcpyDeviceToHost, c->copy_stream));
    LMH_HIP(c, hipEventRecord(sn.done, c->copy_stream));
    sn.busy = true;
    *index = k;
    return LMH_OK;
}

int lmh_snapshot_get(lmh_ctx* c, int32_t index, double* pos, double* mom) {
    if (!c) return LMH_E_INVALID;
    if (index < 0 || index >= (int32_t)c->snaps.size() || !c->snaps[index].busy)
        return c->fail(LMH_E_INVALID, "snapshot_get: no such snapshot");
    if (!pos || !mom) return c->fail(LMH_E_INVALID, "snapshot_get: null output");
    lmh_ctx::Snap& sn = c->snaps[index];
    const size_t rows = (size_t)(c->dims + c->pdim);
    if (sn.n != rows * (size_t)c->N) return c->fail(LMH_E_STATE, "snapshot_get: the particles changed since the snapshot");
    LMH_HIP(c, hipSetDevice(c->device));
    LMH_HIP(c, hipEventSynchronize(sn.done));
    const int64_t N = c->N;
    const double* h = sn.h;
    for (int64_t i = 0; i < N; ++i) {
        for (int d = 0; d < c->dims; ++d) pos[d + (int64_t)c->dims * i] = h[(size_t)d * N + i];
        for (int d = 0; d < c->pdim; ++d) mom[d + (int64_t)c->pdim * i] = h[(size_t)(c->dims + d) * N + i];
    }
    sn.busy = false;
    return LMH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// on-the-fly freeze-out capture (kernels_gpu.jl:583-635, simulate_gpu.jl:53-75,375-402)
extern "C" {

int lmh_freezeout_enable(lmh_ctx* c, double Tfo, int32_t interp) {
    if (!c) return LMH_E_INVALID;
    if (!c->d_state) return c->fail(LMH_E_STATE, "freezeout_enable: set the particles first");
    if (!std::isfinite(Tfo)) return c->fail(LMH_E_INVALID, "freezeout_enable: Tfo must be finite");
    LMH_HIP(c, hipSetDevice(c->device));
    if (int rc = ensure_unsorted(c)) return rc;
    if (!(Tfo > 0.0)) {
        c->fo_on = false;
        LMH_HIP(c, hipStreamSynchronize(c->stream));
        dfree(c->d_fo);
        return LMH_OK;
    }
    c->fo_on = true;
    c->fo_Tfo = Tfo;
    c->fo_interp = interp ? 1 : 0;
    return fo_reset(c);
}

int lmh_freezeout_get(lmh_ctx* c, double* pos, double* mom, double* tau, double* flag) {
    if (!c) return LMH_E_INVALID;
    if (!c->fo_on || !c->d_fo) return c->fail(LMH_E_STATE, "freezeout_get: freeze-out capture is not enabled");
    if (!pos || !mom || !tau || !flag) return c->fail(LMH_E_INVALID, "freezeout_get: null output");
    LMH_HIP(c, hipSetDevice(c->device));
    const int64_t N = c->N;
    const size_t rows = (size_t)(c->dims + c->pdim + 2);
    std::vector<double> h;
    try {
        h.resize(rows * (size_t)N);
    } catch (...) {
        return c->fail(LMH_E_NOMEM, "freezeout_get: host allocation failed");
    }
    LMH_HIP(c, hipMemcpyAsync(h.data(), c->d_fo, rows * (size_t)N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    LMH_HIP(c, hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < N; ++i) {
        for (int d = 0; d < c->dims; ++d) pos[d + (int64_t)c->dims * i] = h[(size_t)d * N + i];
        for (int d = 0; d < c->pdim; ++d) mom[d + (int64_t)c->pdim * i] = h[(size_t)(c->dims + d) * N + i];
        tau[i] = h[(size_t)(c->dims + c->pdim) * N + i];
        flag[i] = h[(size_t)(c->dims + c->pdim + 1) * N + i];
    }
    return LMH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// pathwise adjoint (no reference counterpart)
namespace {

int ensure_dev(lmh_ctx* c, double** ptr, size_t* cap, size_t n, const char* what) {
    if (n <= *cap) return LMH_OK;
    LMH_HIP(c, hipStreamSynchronize(c->stream));
    dfree(*ptr);
    *cap = 0;
    if (hipMalloc((void**)ptr, n * sizeof(double)) != hipSuccess) {
        *ptr = nullptr;
        return c->fail(LMH_E_NOMEM, std::string(what) + ": device allocation failed");
    }
    *cap = n;
    return LMH_OK;
}

int check_adjoint(lmh_ctx* c, const lmh_params* p) {
    if (c->medium != 1 || p->coords != LMH_CARTESIAN || c->dims != 2 || (c->pdim != 2 && c->pdim != 3))
        return c->fail(LMH_E_INVALID, "adjoint: the 2+1D grid medium in Cartesian coordinates (dims 2, pdim 2 or 3)");
    if (p->noise != LMH_NOISE_PHILOX)
        return c->fail(LMH_E_INVALID, "adjoint: Philox noise (common random numbers are regenerated from counters)");
    if (p->collision != LMH_LANGEVIN && p->collision != LMH_LANGEVIN_RTA)
        return c->fail(LMH_E_INVALID, "adjoint: the Langevin kick, with or without RTA collisions");
    if (p->integrator != LMH_PREPOINT || !p->relativistic || p->position_diffusion)
        return c->fail(LMH_E_INVALID, "adjoint: the relativistic pre-point step without position diffusion");
    if (!(p->dt > 0.0)) return c->fail(LMH_E_INVALID, "adjoint: dt > 0");
    if (p->drag_model == LMH_DRAG_GAMMA_P && !(p->Lambda > 0.0))
        return c->fail(LMH_E_INVALID, "adjoint: Lambda > 0");
    return LMH_OK;
}

// one adjoint segment launch: the tape kernel (tape = true; on stream st, with the final state
// to state_out when it is the forward pass) or the reverse kernel (on the context's stream)
struct AdjLaunch {
    const StepArgs* a;
    unsigned nb;
    const GridMedium* g;
    const double4* sl;
    int64_t step0;
    int32_t nk;
    double* tp;
    const double* ckpt = nullptr;
    double* state_out = nullptr;
    hipStream_t st = nullptr;
    double* lam = nullptr;
    double* gacc = nullptr;
    double dtau = 0.0;
    const double* row0 = nullptr;   // the reverse kernel: step 0's rows (the checkpoint) if set
};

template <int PDIM, int COLL, int DRAG>
void launch_adj_t(lmh_ctx* c, bool tape, const AdjLaunch& l) {
    const dim3 G(l.nb), B(c->tune.block);
    if (tape)
        hipLaunchKernelGGL((k_adj_tape<PDIM, COLL, DRAG>), G, B, 0, l.st, *l.a, *l.g, l.sl, l.ckpt, l.step0, l.nk,
                           l.tp, l.state_out);
    else
        hipLaunchKernelGGL((k_adj_reverse<PDIM, COLL, DRAG>), G, B, 0, c->stream, *l.a, *l.g, l.sl, l.step0, l.nk,
                           l.tp, l.row0, l.lam, l.gacc, l.dtau);
}

template <int PDIM>
void launch_adj_p(lmh_ctx* c, bool tape, int coll, int drag, const AdjLaunch& l) {
    if (coll == LMH_LANGEVIN_RTA) {
        if (drag == LMH_DRAG_GAMMA_P) launch_adj_t<PDIM, LMH_LANGEVIN_RTA, LMH_DRAG_GAMMA_P>(c, tape, l);
        else launch_adj_t<PDIM, LMH_LANGEVIN_RTA, LMH_DRAG_TABLE>(c, tape, l);
    } else {
        if (drag == LMH_DRAG_GAMMA_P) launch_adj_t<PDIM, LMH_LANGEVIN, LMH_DRAG_GAMMA_P>(c, tape, l);
        else launch_adj_t<PDIM, LMH_LANGEVIN, LMH_DRAG_TABLE>(c, tape, l);
    }
}

void launch_adj(lmh_ctx* c, bool tape, const lmh_params& p, const AdjLaunch& l) {
    if (c->pdim == 2) launch_adj_p<2>(c, tape, p.collision, p.drag_model, l);
    else launch_adj_p<3>(c, tape, p.collision, p.drag_model, l);
}

// The forward steps of an adjoint segment again, into its tape, on stream st: step k of the
// segment reads tape row k (the checkpoint for k = 0) and writes row k + 1, one launch of
// lmh_run's step kernel per step with the segment's medium slices -- the forward pass's kernel
// and medium, so the tape is its trajectory bit for bit.  The last step's output (the next
// checkpoint) is not needed.
template <int PDIM, int COLL, bool SPEC>
void adj_tape_steps_t(const StepArgs& a, unsigned nb, const GridMedium& g, int block, hipStream_t st) {
    hipLaunchKernelGGL((k_cart_grid<PDIM, LMH_NOISE_PHILOX, true, COLL, false, SPEC>), dim3(nb), dim3(block), 0, st, a, g);
}

template <int PDIM>
void adj_tape_steps_p(const StepArgs& a, unsigned nb, const GridMedium& g, int block, hipStream_t st, int coll,
                      bool spec) {
    if (coll == LMH_LANGEVIN_RTA) {
        if (spec) adj_tape_steps_t<PDIM, LMH_LANGEVIN_RTA, true>(a, nb, g, block, st);
        else adj_tape_steps_t<PDIM, LMH_LANGEVIN_RTA, false>(a, nb, g, block, st);
    } else {
        if (spec) adj_tape_steps_t<PDIM, LMH_LANGEVIN, true>(a, nb, g, block, st);
        else adj_tape_steps_t<PDIM, LMH_LANGEVIN, false>(a, nb, g, block, st);
    }
}

void adj_tape_steps(lmh_ctx* c, const lmh_params& p, StepArgs a, unsigned nb, const GridMedium& g,
                    const double4* sl, int64_t step0, int32_t nk, const double* ckpt, double* tape,
                    hipStream_t st) {
    const int64_t N = c->N;
    const int64_t R = c->dims + c->pdim;
    const int64_t ncell = (int64_t)(c->gx - 1) * (c->gy - 1);
    const bool spec = spec_config(c, &p);
    a.fo.on = 0;
    a.nsteps = 1;
    for (int32_t k = 0; k + 1 < nk; ++k) {
        double* in = const_cast<double*>(k == 0 ? ckpt : tape + (size_t)k * R * N);
        double* out = tape + (size_t)(k + 1) * R * N;
        for (int d = 0; d < 3; ++d) {
            a.pos[d] = d < c->dims ? in + d * N : nullptr;
            a.mom[d] = d < c->pdim ? in + (c->dims + d) * N : nullptr;
            a.wpos[d] = d < c->dims ? out + d * N : nullptr;
            a.wmom[d] = d < c->pdim ? out + (c->dims + d) * N : nullptr;
        }
        a.step_begin = step0 + k;
        GridMedium gk = g;
        gk.slice = sl + (size_t)k * 3 * (size_t)ncell;
        if (c->pdim == 2) adj_tape_steps_p<2>(a, nb, gk, c->tune.block, st, p.collision, spec);
        else adj_tape_steps_p<3>(a, nb, gk, c->tune.block, st, p.collision, spec);
    }
}

// the forward pass of the adjoint steps through the per-step medium slices, one step per launch,
// which the backward pass rebuilds for its recomputation: both see the same medium, bit for bit
struct SliceTuning {
    lmh_ctx* c;
    lmh_tuning saved;
    explicit SliceTuning(lmh_ctx* c_) : c(c_), saved(c_->tune) {
        c->tune.slice = 1;
        c->tune.steps_per_launch = 1;
    }
    ~SliceTuning() { c->tune = saved; }
};

// Adjoint checkpoint cp: the state in its current order, and that order -- -1 for the caller's,
// else the index of a copy of the identity array, shared by the consecutive checkpoints no
// permutation separates (*last_gen: the order generation of the last copy).  The backward pass
// re-runs and reverses the segment in this order, so its medium reads stay cell-coherent.
int take_checkpoint(lmh_ctx* c, int64_t cp, double* dst, int32_t* nid, uint64_t* last_gen) {
    const int64_t N = c->N;
    const int rows = c->dims + c->pdim;
    LMH_HIP(c, hipMemcpyAsync(dst, c->d_state, (size_t)rows * N * sizeof(double), hipMemcpyDeviceToDevice,
                              c->stream));
    int32_t ord = -1;
    if (c->is_sorted) {
        if (*nid > 0 && *last_gen == c->order_gen) {
            ord = *nid - 1;
        } else {
            ord = (*nid)++;
            *last_gen = c->order_gen;
            LMH_HIP(c, hipMemcpyAsync(c->d_ckpt_id + (size_t)ord * N, c->d_id, (size_t)N * sizeof(uint32_t),
                                      hipMemcpyDeviceToDevice, c->stream));
        }
    }
    c->adj.ord[cp] = ord;
    return LMH_OK;
}

}  // namespace

extern "C" {

int lmh_adjoint_forward(lmh_ctx* c, const lmh_params* p, int64_t step_begin, int64_t nsteps,
                        const lmh_adjoint_spec* s) {
    if (!c) return LMH_E_INVALID;
    c->adj.valid = false;
    int rc = check_run(c, p);
    if (rc) return rc;
    if (!s || s->checkpoint_every < 1) return c->fail(LMH_E_INVALID, "adjoint_forward: checkpoint_every >= 1");
    if (s->keep_tape != 0 && s->keep_tape != 1) return c->fail(LMH_E_INVALID, "adjoint_forward: keep_tape is 0 or 1");
    if (nsteps < 1) return c->fail(LMH_E_INVALID, "adjoint_forward: nsteps >= 1");
    if ((rc = check_adjoint(c, p))) return rc;
    if ((rc = check_steps(c, p, step_begin, nsteps))) return rc;
    if (s->keep_tape && c->fo_on)
        return c->fail(LMH_E_INVALID, "adjoint_forward: keep_tape runs without freeze-out capture");
    LMH_HIP(c, hipSetDevice(c->device));
    if ((rc = ensure_unsorted(c))) return rc;
    const int64_t K = s->checkpoint_every;
    const int64_t ncp = (nsteps + K - 1) / K;
    const size_t R = (size_t)(c->dims + c->pdim);
    const size_t per = R * (size_t)c->N;
    ++c->gen;
    try {
        c->adj.ord.assign((size_t)ncp, -1);
    } catch (...) {
        return c->fail(LMH_E_NOMEM, "adjoint_forward: host allocation failed");
    }
    if (s->keep_tape) {
        // the whole tape and every step's medium slice stay in device memory: the pullback only
        // runs the reverse sweep.  The forward steps are the tape kernel's, segment by segment.
        const size_t TR = R;                 // the tape kernels' rows: the state (no noise)
        const int64_t ncell = (int64_t)(c->gx - 1) * (c->gy - 1);
        if ((rc = ensure_dev(c, &c->d_tape, &c->tape_cap, (size_t)nsteps * TR * (size_t)c->N,
                             "adjoint_forward whole tape (keep_tape)")))
            return rc;
        if ((rc = ensure_dev(c, &c->d_adjslice, &c->adjslice_cap, (size_t)nsteps * 12 * (size_t)ncell,
                             "adjoint_forward medium slices (keep_tape)")))
            return rc;
        unsigned nb = 0;
        StepArgs a = make_args(c, p, c->tune.block, &nb);
        GridMedium g = make_grid(c);
        for (int64_t cp = 0; cp < ncp; ++cp) {
            const int32_t nk = (int32_t)std::min<int64_t>(K, nsteps - cp * K);
            const int64_t step0 = step_begin + cp * K;
            double4* sl = reinterpret_cast<double4*>(c->d_adjslice) + (size_t)cp * K * 3 * (size_t)ncell;
            hipLaunchKernelGGL(k_grid_slices, dim3(grid_blocks(3 * ncell * nk)), dim3(kBlock), 0, c->stream, g, sl,
                               step0, nk, p->dt, p->t0);
            if (nb > 0) {
                AdjLaunch l{&a, nb, &g, sl, step0, nk, c->d_tape + (size_t)cp * K * TR * (size_t)c->N};
                l.ckpt = c->d_state;
                l.state_out = c->d_state;
                l.st = c->stream;
                LaunchMark mk(c, LMH_PROF_STEP);
                launch_adj(c, true, *p, l);
            }
            LMH_HIP(c, hipGetLastError());
        }
        if (p->check_finite) {
            unsigned long long bad = ULLONG_MAX;
            LMH_HIP(c, hipMemcpyAsync(&bad, c->d_bad, sizeof(bad), hipMemcpyDeviceToHost, c->stream));
            LMH_HIP(c, hipStreamSynchronize(c->stream));
            if (bad != ULLONG_MAX)
                return c->fail(LMH_E_NONFINITE, "Breaking simulation due to NaN/Inf at step " + std::to_string((long long)bad));
        }
    } else {
        // lmh_run's steps (cell-sorted as it sorts, one step per launch through the medium slices),
        // a checkpoint before every segment (after the segment's cell sort when one is due, so
        // that the checkpoint's order is the one the segment starts in), one finiteness check at
        // the end
        if ((rc = ensure_dev(c, &c->d_ckpt, &c->ckpt_cap, per * (size_t)ncp, "adjoint_forward checkpoints"))) return rc;
        const bool sortable = run_sortable(c, p);
        if (sortable && c->ckpt_id_cap < (size_t)ncp * (size_t)c->N) {
            LMH_HIP(c, hipStreamSynchronize(c->stream));
            dfree(c->d_ckpt_id);
            c->ckpt_id_cap = 0;
            if (hipMalloc((void**)&c->d_ckpt_id, (size_t)ncp * (size_t)c->N * sizeof(uint32_t)) != hipSuccess) {
                c->d_ckpt_id = nullptr;
                return c->fail(LMH_E_NOMEM, "adjoint_forward checkpoint orders: device allocation failed");
            }
            c->ckpt_id_cap = (size_t)ncp * (size_t)c->N;
        }
        SliceTuning sliced(c);
        lmh_params pseg = *p;
        pseg.check_finite = 0;
        int32_t nid = 0;
        uint64_t last_gen = 0;
        for (int64_t cp = 0; cp < ncp; ++cp) {
            if (sortable && (!c->is_sorted || c->steps_since_sort >= c->tune.sort_every))
                if ((rc = sort_by_cell(c))) return rc;
            if ((rc = take_checkpoint(c, cp, c->d_ckpt + per * (size_t)cp, &nid, &last_gen))) return rc;
            const int64_t n = std::min<int64_t>(K, nsteps - cp * K);
            if ((rc = run_impl(c, &pseg, step_begin + cp * K, n, nullptr, 1))) return rc;
        }
        if (p->check_finite) {
            unsigned long long bad = ULLONG_MAX;
            LMH_HIP(c, hipMemcpyAsync(&bad, c->d_bad, sizeof(bad), hipMemcpyDeviceToHost, c->stream));
            LMH_HIP(c, hipStreamSynchronize(c->stream));
            if (bad != ULLONG_MAX)
                return c->fail(LMH_E_NONFINITE, "Breaking simulation due to NaN/Inf at step " + std::to_string((long long)bad));
        }
    }
    c->adj.p = *p;
    c->adj.step_begin = step_begin;
    c->adj.nsteps = nsteps;
    c->adj.K = (int32_t)K;
    c->adj.keep = s->keep_tape != 0;
    c->adj.dtau_dDsT = s->dtau_dDsT;
    c->adj.gen = c->gen;
    c->adj.valid = true;
    return LMH_OK;
}

int lmh_adjoint_backward(lmh_ctx* c, const lmh_obs_spec* os, const double* cotangent, double* grad,
                         double* dpos, double* dmom) {
    if (!c) return LMH_E_INVALID;
    if (!c->adj.valid || c->adj.gen != c->gen)
        return c->fail(LMH_E_STATE, "adjoint_backward: no adjoint forward pass for the current state "
                                    "(the particles, medium or tables changed since)");
    if (!os || !cotangent || !grad) return c->fail(LMH_E_INVALID, "adjoint_backward: null spec, cotangent or gradient");
    if ((dpos == nullptr) != (dmom == nullptr))
        return c->fail(LMH_E_INVALID, "adjoint_backward: give both state-gradient outputs or neither");
    if (os->nbins < 0 || os->nbins > kObsMaxBins || (os->nbins > 0 && !(os->pt_max > os->pt_min)))
        return c->fail(LMH_E_INVALID, "adjoint_backward: 0 <= nbins <= 120 and pt_max > pt_min");
    if (os->species >= c->nspecies) return c->fail(LMH_E_INVALID, "adjoint_backward: species out of range");
    LMH_HIP(c, hipSetDevice(c->device));
    if (int rc = ensure_unsorted(c)) return rc;
    const lmh_params& p = c->adj.p;
    const int64_t N = c->N;
    const int R = c->dims + c->pdim;
    const int TR = R;                    // the tape kernels' rows: the state (no noise)
    const int64_t K = c->adj.K;
    const int64_t ncp = (c->adj.nsteps + K - 1) / K;
    const int nw = 5 + os->nbins;
    int rc;
    const bool keep = c->adj.keep;
    const size_t tape_n = (size_t)K * TR * (size_t)N;
    const int64_t ncell = (int64_t)(c->gx - 1) * (c->gy - 1);
    const size_t slices_n = (size_t)K * 12 * (size_t)ncell;     // K slices of ncell x 3 double4
    if (!keep) {
        if ((rc = ensure_dev(c, &c->d_tape, &c->tape_cap, 2 * tape_n, "adjoint_backward tape"))) return rc;
        if ((rc = ensure_dev(c, &c->d_adjslice, &c->adjslice_cap, 2 * slices_n, "adjoint_backward medium slices")))
            return rc;
    }
    if (!keep && !c->tape_stream) {
        LMH_HIP(c, hipStreamCreateWithFlags(&c->tape_stream, hipStreamNonBlocking));
        for (int k = 0; k < 2; ++k) {
            LMH_HIP(c, hipEventCreateWithFlags(&c->ev_tape[k], hipEventDisableTiming));
            LMH_HIP(c, hipEventCreateWithFlags(&c->ev_rev[k], hipEventDisableTiming));
        }
    }
    // lam and the per-particle gradients are adjacent rows, which move together between orders
    // through the reorder buffer
    const int nmov = R + kAdjThetas;
    const size_t o_lam = 0, o_g = (size_t)R * N, o_part = o_g + (size_t)kAdjThetas * N;
    const size_t o_out = o_part + (size_t)kAdjThetas * kSumBlocks, o_w = o_out + kAdjThetas;
    const size_t o_mov = o_w + nw;
    bool reorder = false;
    for (const int32_t o : c->adj.ord) reorder = reorder || o >= 0;
    if ((rc = ensure_dev(c, &c->d_adjw, &c->adjw_cap, o_mov + (reorder ? (size_t)nmov * N : 0),
                         "adjoint_backward workspace")))
        return rc;
    double* lam = c->d_adjw + o_lam;
    double* gacc = c->d_adjw + o_g;
    LMH_HIP(c, hipMemcpyAsync(c->d_adjw + o_w, cotangent, nw * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int64_t i0 = 0, i1 = N;
    if (os->species >= 0) {
        i0 = species_begin(c, os->species);
        i1 = i0 + c->counts[os->species];
    }
    const double* px = c->d_state + (int64_t)c->dims * N;
    hipLaunchKernelGGL(k_adj_seed, dim3(grid_blocks(N)), dim3(kBlock), 0, c->stream, px, px + N, N, i0, i1,
                       c->pdim, *os, c->d_adjw + o_w, lam);
    LMH_HIP(c, hipMemsetAsync(gacc, 0, (size_t)kAdjThetas * N * sizeof(double), c->stream));
    unsigned nb = 0;
    StepArgs a = make_args(c, &p, c->tune.block, &nb);
    GridMedium g = make_grid(c);
    const size_t per = (size_t)R * N;
    if (keep) {
        // the forward pass kept every step's tape row and medium slice: reverse sweeps only
        for (int64_t cp = ncp - 1; cp >= 0; --cp) {
            const int32_t nk = (int32_t)std::min<int64_t>(K, c->adj.nsteps - cp * K);
            const int64_t step0 = c->adj.step_begin + cp * K;
            AdjLaunch l{&a, nb, &g, reinterpret_cast<const double4*>(c->d_adjslice) + (size_t)cp * K * 3 * (size_t)ncell,
                        step0, nk, c->d_tape + (size_t)cp * tape_n};
            l.lam = lam;
            l.gacc = gacc;
            l.dtau = c->adj.dtau_dDsT;
            LaunchMark mk(c, LMH_PROF_STEP);
            launch_adj(c, false, p, l);
            LMH_HIP(c, hipGetLastError());
        }
    } else {
        // the tape stream starts after the checkpoints are final (the forward pass and the seed
        // were enqueued on the main stream)
        LMH_HIP(c, hipEventRecord(c->ev_rev[0], c->stream));
        LMH_HIP(c, hipStreamWaitEvent(c->tape_stream, c->ev_rev[0], 0));
        // lam and the gradients start in the caller's order; each segment is re-run and reversed
        // in its checkpoint's order (the identities key the noise), and lam and the gradients
        // move to that order first when it differs (through the caller's order)
        const uint32_t* caller_id = c->d_id;
        double* mov = lam;
        double* buf_mov = c->d_adjw + o_mov;
        int32_t lam_ord = -1;
        auto move_to = [&](int32_t ord) -> int {
            if (ord == lam_ord) return LMH_OK;
            const unsigned gb = grid_blocks(N);
            if (lam_ord >= 0) {
                hipLaunchKernelGGL(k_scatter_rows, dim3(gb), dim3(kBlock), 0, c->stream, mov, buf_mov,
                                   (int32_t)nmov, N, c->d_ckpt_id + (size_t)lam_ord * N);
            } else {
                LMH_HIP(c, hipMemcpyAsync(buf_mov, mov, (size_t)nmov * N * sizeof(double), hipMemcpyDeviceToDevice,
                                          c->stream));
            }
            if (ord >= 0) {
                hipLaunchKernelGGL(k_gather_rows, dim3(gb), dim3(kBlock), 0, c->stream, buf_mov, mov, (int32_t)nmov,
                                   N, c->d_ckpt_id + (size_t)ord * N);
            } else {
                LMH_HIP(c, hipMemcpyAsync(mov, buf_mov, (size_t)nmov * N * sizeof(double), hipMemcpyDeviceToDevice,
                                          c->stream));
            }
            LMH_HIP(c, hipGetLastError());
            lam_ord = ord;
            return LMH_OK;
        };
        for (int64_t cp = ncp - 1; cp >= 0; --cp) {
            const int32_t nk = (int32_t)std::min<int64_t>(K, c->adj.nsteps - cp * K);
            const int64_t step0 = c->adj.step_begin + cp * K;
            const int buf = (int)(cp & 1);
            const int32_t ord = c->adj.ord[(size_t)cp];
            a.id = ord >= 0 ? c->d_ckpt_id + (size_t)ord * N : caller_id;
            double4* sl = reinterpret_cast<double4*>(c->d_adjslice + slices_n * buf);
            AdjLaunch l{&a, nb, &g, sl, step0, nk, c->d_tape + tape_n * buf};
            l.ckpt = c->d_ckpt + per * cp;
            l.st = c->tape_stream;
            l.lam = lam;
            l.gacc = gacc;
            l.dtau = c->adj.dtau_dDsT;
            // re-run segment cp into tape and slice buffers buf, once the reverse of segment
            // cp + 2 (their previous reader) is done
            if (cp + 2 <= ncp - 1) LMH_HIP(c, hipStreamWaitEvent(c->tape_stream, c->ev_rev[buf], 0));
            hipLaunchKernelGGL(k_grid_slices, dim3(grid_blocks(3 * ncell * nk)), dim3(kBlock), 0, c->tape_stream,
                               g, sl, step0, nk, p.dt, p.t0);
            adj_tape_steps(c, p, a, nb, g, sl, step0, nk, l.ckpt, l.tp, c->tape_stream);
            l.row0 = l.ckpt;
            LMH_HIP(c, hipEventRecord(c->ev_tape[buf], c->tape_stream));
            // reverse segment cp on the main stream
            if ((rc = move_to(ord))) return rc;
            LMH_HIP(c, hipStreamWaitEvent(c->stream, c->ev_tape[buf], 0));
            {
                LaunchMark mk(c, LMH_PROF_STEP);
                launch_adj(c, false, p, l);
            }
            LMH_HIP(c, hipEventRecord(c->ev_rev[buf], c->stream));
            LMH_HIP(c, hipGetLastError());
        }
        if ((rc = move_to(-1))) return rc;
    }
    hipLaunchKernelGGL(k_rows_partial, dim3(kSumBlocks), dim3(kSumBlock), 0, c->stream, gacc, kAdjThetas, N, c->d_adjw + o_part);
    hipLaunchKernelGGL(k_rows_final, dim3(kAdjThetas), dim3(kSumBlock), 0, c->stream, c->d_adjw + o_part, kSumBlocks, c->d_adjw + o_out);
    LMH_HIP(c, hipGetLastError());
    LMH_HIP(c, hipMemcpyAsync(grad, c->d_adjw + o_out, kAdjThetas * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    LMH_HIP(c, hipStreamSynchronize(c->stream));
    if (dpos) {
        std::vector<double> h;
        try {
            h.resize(per);
        } catch (...) {
            return c->fail(LMH_E_NOMEM, "adjoint_backward: host allocation failed");
        }
        LMH_HIP(c, hipMemcpy(h.data(), lam, per * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < N; ++i) {
            for (int d = 0; d < c->dims; ++d) dpos[d + (int64_t)c->dims * i] = h[(size_t)d * N + i];
            for (int d = 0; d < c->pdim; ++d) dmom[d + (int64_t)c->pdim * i] = h[(size_t)(c->dims + d) * N + i];
        }
    }
    return LMH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// batches of independent runs (no reference counterpart: the reference runs one ensemble per
// call with scalar m and DsT, simulate_cpu.jl:20-21)
struct lmh_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    struct Med { double4* nodes = nullptr; double* tg = nullptr; int nx = 0, ny = 0, nt = 0; double x0 = 0, dx = 1, y0 = 0, dy = 1; };
    struct Tab { double* d = nullptr; double Tmin = 0, invdT = 1; int n = 0; };
    struct JTab { double* d = nullptr; double Tmin = 0, invdT = 1; int nU = 0, nT = 0; };
    std::vector<Med> media;
    std::vector<Tab> tabs;
    std::vector<JTab> jtabs;

    std::vector<lmh_run_desc> runs;
    std::vector<int64_t> begin;       // nruns + 1
    int pdim = 0;
    int64_t N = 0;
    double* d_state = nullptr;        // x, y, px, py[, pz] rows of N doubles
    unsigned long long* d_bad = nullptr;
    int64_t* d_step = nullptr;
    int64_t* d_begin = nullptr;
    BatchRun* d_runs = nullptr;
    int32_t* d_tile_run = nullptr;
    int64_t ntiles = 0;
    int tiles_block = 0;              // block size the tile table was built for
    int layout_slice_min = 0;         // ... and slice_min and steps per launch (the slice rings)
    int layout_spl = 0;
    // per-step node slices of the runs of at least slice_min quarks: run srun[k]'s ring of
    // batch_ring(steps_per_launch) slices in d_bslice, filled by k_batch_nodes ahead of the step
    // launches that read them
    double4* d_bslice = nullptr;
    size_t bslice_cap = 0;            // double4 nodes
    int32_t* d_srun = nullptr;
    int32_t nsr = 0;
    int64_t max_node_blocks = 0;      // k_batch_nodes blocks of the largest sliced grid
    // the tiles of the sliced runs and of the others, for the launches that split them
    int32_t* d_tile_sl = nullptr;
    int32_t* d_tile_co = nullptr;
    int64_t ntiles_sl = 0, ntiles_co = 0;
    double2* d_tcell = nullptr;       // runs x steps_per_launch tau cells (k_batch_tcell)
    size_t tcell_cap = 0;

    lmh_batch_tuning tune = {256, 1, 0, 0, 0, 0};

    // cell order (tune.sort_every > 0): each run's quarks ordered by grid cell within the run's
    // segment; d_id[s] is the original slot of the quark in slot s (its Philox counter is d_id[s]
    // minus the run's first slot).  The step after a sort reads through the sort's permutation
    // and writes into the alternate buffers, which then become the current ones, so the state
    // alternates between two addresses (a captured graph for each).
    uint32_t* d_id = nullptr;
    uint32_t* d_id_alt = nullptr;
    double* d_state_alt = nullptr;
    uint32_t* d_keys = nullptr;       // 4 N words: keys, sorted keys, values, sorted values
    void* d_sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    bool is_sorted = false;
    int64_t steps_since_sort = 0;
    // a sort's permutation (the sorted values d_keys[3N, 4N)) not yet applied: the next step
    // launch reads through it (batch_gather_step), or k_permute applies it
    bool perm_pending = false;

    // captured step loops, one for each of the state's two addresses: slot k valid for gkey[k]
    // (parameters, tuning, the state and identity addresses); glast the slot replayed last
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    std::vector<unsigned char> gkey[2];
    int glast = 0;

    double* d_obs = nullptr;
    size_t obs_cap = 0;
    double* d_save = nullptr;
    size_t save_cap = 0;
    int cus = 0;

    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
};

namespace {

#define LMB_HIP(b, expr)                                                                    \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return (b)->fail(LMH_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

void batch_drop_graph_slot(lmh_batch* b, int k) {
    if (b->gexec[k]) (void)hipGraphExecDestroy(b->gexec[k]);
    if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
    b->gexec[k] = nullptr;
    b->graph[k] = nullptr;
    b->gkey[k].clear();
}

void batch_drop_graph(lmh_batch* b) {
    batch_drop_graph_slot(b, 0);
    batch_drop_graph_slot(b, 1);
}

void batch_free_order(lmh_batch* b) {
    dfree(b->d_id);
    dfree(b->d_id_alt);
    dfree(b->d_state_alt);
    dfree(b->d_keys);
    dfree(b->d_sort_tmp);
    b->d_id = b->d_id_alt = nullptr;
    b->d_state_alt = nullptr;
    b->d_keys = nullptr;
    b->d_sort_tmp = nullptr;
    b->sort_tmp_bytes = 0;
    b->is_sorted = false;
    b->perm_pending = false;
    b->steps_since_sort = 0;
}

// the buffers of the cell order (alternate state and identities, keys and values, the radix
// sort's workspace), allocated at the first sort
int batch_sort_alloc(lmh_batch* b) {
    if (b->d_keys) return LMH_OK;
    const int64_t N = b->N;
    const size_t rows = (size_t)(2 + b->pdim);
    if (hipMalloc((void**)&b->d_id_alt, (size_t)N * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void**)&b->d_state_alt, rows * (size_t)N * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&b->d_keys, 4 * (size_t)N * sizeof(uint32_t)) != hipSuccess)
        return b->fail(LMH_E_NOMEM, "batch cell sort: device allocation failed");
    size_t bytes = 0;
    uint32_t* k = b->d_keys;
    LMB_HIP(b, hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k, k + N, k + 2 * N, k + 3 * N, (int)N, 0, 32,
                                                   b->stream));
    if (hipMalloc(&b->d_sort_tmp, std::max<size_t>(bytes, 1)) != hipSuccess)
        return b->fail(LMH_E_NOMEM, "batch cell sort: device allocation failed");
    b->sort_tmp_bytes = bytes;
    return LMH_OK;
}

// permute the state and the identities through a pending sort's values into the alternate
// buffers, which then become the current ones
int batch_finish_permutation(lmh_batch* b) {
    if (!b->perm_pending) return LMH_OK;
    const int64_t N = b->N;
    const int rows = 2 + b->pdim;
    hipLaunchKernelGGL(k_permute, dim3(grid_blocks(N)), dim3(kBlock), 0, b->stream, b->d_state, b->d_state_alt,
                       (int32_t)rows, N, b->d_keys + 3 * N, b->d_id, b->d_id_alt);
    LMB_HIP(b, hipGetLastError());
    std::swap(b->d_state, b->d_state_alt);
    std::swap(b->d_id, b->d_id_alt);
    b->perm_pending = false;
    return LMH_OK;
}

// sort the keys d_keys[0, N) with the values d_keys[2N, 3N) over bits [0, end_bit); the sorted
// values are the permutation, applied by the next step launch (defer) or here
int batch_sort_apply(lmh_batch* b, int end_bit, bool defer) {
    const int64_t N = b->N;
    uint32_t* k = b->d_keys;
    size_t bytes = b->sort_tmp_bytes;
    LMB_HIP(b, hipcub::DeviceRadixSort::SortPairs(b->d_sort_tmp, bytes, k, k + N, k + 2 * N, k + 3 * N, (int)N, 0,
                                                   end_bit, b->stream));
    b->perm_pending = true;
    return defer ? LMH_OK : batch_finish_permutation(b);
}

BatchArgs batch_args(lmh_batch* b, const lmh_params* p);

// order every run's quarks by grid cell (within the run's segment)
int batch_sort(lmh_batch* b, const lmh_params* p) {
    int rc = batch_sort_alloc(b);
    if (rc) return rc;
    int cell_bits = 1;
    for (const auto& m : b->media) cell_bits = std::max(cell_bits, bits_for((uint64_t)(m.nx - 1) * (uint64_t)(m.ny - 1)));
    const int end_bit = cell_bits + bits_for((uint64_t)b->runs.size());
    if (end_bit > 32) return b->fail(LMH_E_INVALID, "batch cell sort: too many runs or grid cells");
    const BatchArgs a = batch_args(b, p);
    const unsigned G = (unsigned)std::min<int64_t>(b->ntiles, 65535);
    hipLaunchKernelGGL(k_batch_sort_keys, dim3(G), dim3(b->tune.block), 0, b->stream, a, (int32_t)cell_bits,
                       b->d_keys, b->d_keys + 2 * b->N);
    LMB_HIP(b, hipGetLastError());
    if ((rc = batch_sort_apply(b, end_bit, true))) return rc;
    b->is_sorted = true;
    b->steps_since_sort = 0;
    return LMH_OK;
}

// restore the caller's order (a no-op unless the runs are cell-sorted)
int batch_ensure_unsorted(lmh_batch* b) {
    if (int rc = batch_finish_permutation(b)) return rc;
    if (!b->is_sorted) return LMH_OK;
    const int64_t N = b->N;
    hipLaunchKernelGGL(k_unsort_keys, dim3(grid_blocks(N)), dim3(kBlock), 0, b->stream, b->d_id, N, b->d_keys,
                       b->d_keys + 2 * N);
    LMB_HIP(b, hipGetLastError());
    int rc = batch_sort_apply(b, bits_for((uint64_t)N), false);
    if (rc) return rc;
    b->is_sorted = false;
    return LMH_OK;
}

int batch_dev(lmh_batch* b, double** ptr, size_t* cap, size_t n, const char* what) {
    if (n <= *cap) return LMH_OK;
    LMB_HIP(b, hipStreamSynchronize(b->stream));
    dfree(*ptr);
    *cap = 0;
    if (hipMalloc((void**)ptr, n * sizeof(double)) != hipSuccess) {
        *ptr = nullptr;
        return b->fail(LMH_E_NOMEM, std::string(what) + ": device allocation failed");
    }
    *cap = n;
    return LMH_OK;
}

// the slots of a sliced run's ring: the steps of one corner-path launch, or of one k_batch_nodes
// launch of the split path (kSliceRing steps)
int batch_ring(int steps_per_launch) { return std::max(std::max(1, steps_per_launch), kSliceRing); }

// a run reads node slices when it holds at least slice_min quarks and its grid fits node_sample's
// 32-bit byte offsets
bool batch_run_sliced(const lmh_run_desc& d, const lmh_batch::Med& md, int smin) {
    return smin > 0 && d.count >= smin && (int64_t)md.nx * md.ny <= ((int64_t)1 << 27);
}

// tile table and per-run records for the current block size
int batch_layout(lmh_batch* b, int block) {
    const int K = std::max(1, b->tune.steps_per_launch);
    const int smin = b->tune.slice_min;
    if (b->tiles_block == block && b->layout_slice_min == smin && b->layout_spl == K && b->d_runs) return LMH_OK;
    const int R = (int)b->runs.size();
    const int ring = batch_ring(K);
    std::vector<BatchRun> h(R);
    std::vector<int32_t> tr, srun, tsl, tco;
    int64_t t0 = 0;
    // the slice rings: `ring` steps of one double4 per grid node for every sliced run
    int64_t nnode = 0, max_blocks = 0;
    for (int r = 0; r < R; ++r) {
        const lmh_run_desc& d = b->runs[r];
        const lmh_batch::Med& md = b->media[d.medium];
        if (!batch_run_sliced(d, md, smin)) continue;
        const int64_t nxy = (int64_t)md.nx * md.ny;
        nnode += ring * nxy;
        max_blocks = std::max<int64_t>(max_blocks, (nxy + kBlock - 1) / kBlock);
        srun.push_back(r);
    }
    if ((size_t)R * ring > b->tcell_cap) {
        dfree(b->d_tcell);
        b->tcell_cap = 0;
        if (hipMalloc((void**)&b->d_tcell, (size_t)R * ring * sizeof(double2)) != hipSuccess)
            return b->fail(LMH_E_NOMEM, "batch: allocation failed");
        b->tcell_cap = (size_t)R * ring;
    }
    if ((size_t)nnode > b->bslice_cap) {
        dfree(b->d_bslice);
        b->bslice_cap = 0;
        if (hipMalloc((void**)&b->d_bslice, (size_t)nnode * sizeof(double4)) != hipSuccess)
            return b->fail(LMH_E_NOMEM, "batch: medium slice allocation failed");
        b->bslice_cap = (size_t)nnode;
    }
    int64_t soff = 0;
    for (int r = 0; r < R; ++r) {
        const lmh_run_desc& d = b->runs[r];
        const lmh_batch::Med& md = b->media[d.medium];
        BatchRun& o = h[r];
        memset(&o, 0, sizeof(o));
        o.begin = b->begin[r];
        o.count = d.count;
        o.tile0 = t0;
        o.g.nodes = md.nodes;
        o.g.slice = nullptr;
        o.g.tg = md.tg;
        o.g.nx = md.nx;
        o.g.ny = md.ny;
        o.g.nt = md.nt;
        o.g.x0 = md.x0;
        o.g.inv_dx = 1.0 / md.dx;
        o.g.xlast = md.x0 + (double)(md.nx - 1) * md.dx;
        o.g.y0 = md.y0;
        o.g.inv_dy = 1.0 / md.dy;
        o.g.ylast = md.y0 + (double)(md.ny - 1) * md.dy;
        o.sp.m = d.mass;
        if (d.drag_table >= 0) {
            const lmh_batch::Tab& t = b->tabs[d.drag_table];
            o.sp.drag = TimeTable{t.d, t.Tmin, t.invdT, t.n};
        }
        if (d.rta_table >= 0) {
            const lmh_batch::Tab& t = b->tabs[d.rta_table];
            o.sp.rta = TimeTable{t.d, t.Tmin, t.invdT, t.n};
        }
        if (d.juttner_table >= 0) {
            const lmh_batch::JTab& t = b->jtabs[d.juttner_table];
            o.sp.jt = JuttnerTable{t.d, t.Tmin, t.invdT, t.nU, t.nT};
        }
        o.sp.shared_axis = tables_share_axis(o.sp.drag, o.sp.rta);
        o.DsT = d.DsT;
        if (batch_run_sliced(d, md, smin)) {
            o.slice_stride = (int64_t)md.nx * md.ny;
            o.g.slice = b->d_bslice + soff;
            soff += ring * o.slice_stride;
        }
        const int64_t nt = (d.count + block - 1) / block;
        for (int64_t k = 0; k < nt; ++k) {
            tr.push_back(r);
            (o.g.slice ? tsl : tco).push_back((int32_t)(t0 + k));
        }
        t0 += nt;
    }
    batch_drop_graph(b);
    dfree(b->d_runs);
    dfree(b->d_tile_run);
    dfree(b->d_srun);
    dfree(b->d_tile_sl);
    dfree(b->d_tile_co);
    b->nsr = 0;
    b->ntiles_sl = b->ntiles_co = 0;
    if (hipMalloc((void**)&b->d_runs, std::max<size_t>(1, R) * sizeof(BatchRun)) != hipSuccess ||
        hipMalloc((void**)&b->d_tile_run, std::max<size_t>(1, tr.size()) * sizeof(int32_t)) != hipSuccess ||
        hipMalloc((void**)&b->d_srun, std::max<size_t>(1, srun.size()) * sizeof(int32_t)) != hipSuccess ||
        hipMalloc((void**)&b->d_tile_sl, std::max<size_t>(1, tsl.size()) * sizeof(int32_t)) != hipSuccess ||
        hipMalloc((void**)&b->d_tile_co, std::max<size_t>(1, tco.size()) * sizeof(int32_t)) != hipSuccess)
        return b->fail(LMH_E_NOMEM, "batch: device allocation failed");
    if (!tsl.empty()) LMB_HIP(b, hipMemcpy(b->d_tile_sl, tsl.data(), tsl.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!tco.empty()) LMB_HIP(b, hipMemcpy(b->d_tile_co, tco.data(), tco.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    b->ntiles_sl = (int64_t)tsl.size();
    b->ntiles_co = (int64_t)tco.size();
    if (R) LMB_HIP(b, hipMemcpy(b->d_runs, h.data(), R * sizeof(BatchRun), hipMemcpyHostToDevice));
    if (!tr.empty()) LMB_HIP(b, hipMemcpy(b->d_tile_run, tr.data(), tr.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!srun.empty()) LMB_HIP(b, hipMemcpy(b->d_srun, srun.data(), srun.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    b->ntiles = t0;
    b->tiles_block = block;
    b->layout_slice_min = smin;
    b->layout_spl = K;
    b->nsr = (int32_t)srun.size();
    b->max_node_blocks = max_blocks;
    return LMH_OK;
}

BatchArgs batch_args(lmh_batch* b, const lmh_params* p) {
    BatchArgs a;
    memset(&a, 0, sizeof(a));
    for (int d = 0; d < 3; ++d) {
        a.base.pos[d] = d < 2 ? b->d_state + (int64_t)d * b->N : nullptr;
        a.base.mom[d] = d < b->pdim ? b->d_state + (int64_t)(2 + d) * b->N : nullptr;
        a.base.wpos[d] = a.base.pos[d];
        a.base.wmom[d] = a.base.mom[d];
    }
    a.base.N = b->N;
    a.base.p = *p;
    a.base.bad = b->d_bad;
    a.base.id = b->d_id;
    a.base.nseg = 1;
    a.base.sqrt_dt = std::sqrt(p->dt);
    a.base.C = noise_layout(p, 2, b->pdim, &a.base.oxi, &a.base.orta, &a.base.opos);
    a.runs = b->d_runs;
    a.tile_run = b->d_tile_run;
    a.ntiles = b->ntiles;
    a.step_dev = b->d_step;
    return a;
}

int batch_check(lmh_batch* b, const lmh_params* p) {
    if (!p) return b->fail(LMH_E_INVALID, "null params");
    if (b->runs.empty() || !b->d_state) return b->fail(LMH_E_STATE, "batch: no runs set");
    if (p->coords != LMH_CARTESIAN || p->noise != LMH_NOISE_PHILOX)
        return b->fail(LMH_E_INVALID, "batch: Cartesian coordinates and Philox noise");
    if (p->collision < LMH_LANGEVIN || p->collision > LMH_LANGEVIN_RTA)
        return b->fail(LMH_E_INVALID, "unknown collision mode");
    if (p->integrator < LMH_PREPOINT || p->integrator > LMH_POSTPOINT)
        return b->fail(LMH_E_INVALID, "unknown integrator");
    if (p->drag_model != LMH_DRAG_TABLE && p->drag_model != LMH_DRAG_GAMMA_P)
        return b->fail(LMH_E_INVALID, "unknown drag model");
    if (p->reflecting_boundary) return b->fail(LMH_E_INVALID, "the 2+1D grid has no reflecting boundary");
    if (p->bjorken_redshift && b->pdim != 3) return b->fail(LMH_E_INVALID, "bjorken_redshift requires momentum row 3");
    const bool drag = p->collision == LMH_LANGEVIN || p->collision == LMH_LANGEVIN_RTA;
    const bool rta = p->collision == LMH_RTA || p->collision == LMH_LANGEVIN_RTA;
    for (size_t r = 0; r < b->runs.size(); ++r) {
        const lmh_run_desc& d = b->runs[r];
        if (drag && p->drag_model == LMH_DRAG_TABLE && d.DsT > 0.0 && d.drag_table < 0)
            return b->fail(LMH_E_STATE, "batch: run " + std::to_string(r) + " has no drag table");
        if (rta && (d.rta_table < 0 || d.juttner_table < 0))
            return b->fail(LMH_E_STATE, "batch: RTA needs a relaxation-time and a Juttner table for run " + std::to_string(r));
    }
    return LMH_OK;
}

template <int PDIM, bool PERSIST, bool SPEC>
void batch_launch_s(lmh_batch* b, const BatchArgs& a, int coll, int32_t k_off, int32_t n, unsigned grid) {
    const dim3 G(grid), B(b->tune.block);
    switch (coll) {
        case LMH_LANGEVIN: hipLaunchKernelGGL((k_batch_grid<PDIM, LMH_LANGEVIN, PERSIST, SPEC>), G, B, 0, b->stream, a, k_off, n); break;
        case LMH_RTA: hipLaunchKernelGGL((k_batch_grid<PDIM, LMH_RTA, PERSIST, SPEC>), G, B, 0, b->stream, a, k_off, n); break;
        case LMH_FLUID: hipLaunchKernelGGL((k_batch_grid<PDIM, LMH_FLUID, PERSIST, SPEC>), G, B, 0, b->stream, a, k_off, n); break;
        default: hipLaunchKernelGGL((k_batch_grid<PDIM, LMH_LANGEVIN_RTA, PERSIST, SPEC>), G, B, 0, b->stream, a, k_off, n); break;
    }
}

template <int PDIM, bool PERSIST>
void batch_launch_p(lmh_batch* b, const BatchArgs& a, int coll, int32_t k_off, int32_t n, unsigned grid) {
    const lmh_params& p = a.base.p;
    const bool spec = p.relativistic && p.integrator == LMH_PREPOINT && p.drag_model == LMH_DRAG_TABLE &&
                      !p.bjorken_redshift && !p.position_diffusion;
    if (spec) batch_launch_s<PDIM, PERSIST, true>(b, a, coll, k_off, n, grid);
    else batch_launch_s<PDIM, PERSIST, false>(b, a, coll, k_off, n, grid);
}

bool batch_spec(const lmh_params& p) {
    return p.relativistic && p.integrator == LMH_PREPOINT && p.drag_model == LMH_DRAG_TABLE &&
           !p.bjorken_redshift && !p.position_diffusion;
}

template <int PDIM>
void batch_launch_sl(lmh_batch* b, const BatchArgs& a, int coll, int32_t k_off, int32_t slot) {
    const dim3 G((unsigned)a.ntiles), B(b->tune.block);
    switch (coll) {
        case LMH_LANGEVIN: hipLaunchKernelGGL((k_batch_grid_sl<PDIM, LMH_LANGEVIN>), G, B, 0, b->stream, a, a.runs, a.tile_list, a.tile_run, k_off, slot); break;
        case LMH_RTA: hipLaunchKernelGGL((k_batch_grid_sl<PDIM, LMH_RTA>), G, B, 0, b->stream, a, a.runs, a.tile_list, a.tile_run, k_off, slot); break;
        case LMH_FLUID: hipLaunchKernelGGL((k_batch_grid_sl<PDIM, LMH_FLUID>), G, B, 0, b->stream, a, a.runs, a.tile_list, a.tile_run, k_off, slot); break;
        default: hipLaunchKernelGGL((k_batch_grid_sl<PDIM, LMH_LANGEVIN_RTA>), G, B, 0, b->stream, a, a.runs, a.tile_list, a.tile_run, k_off, slot); break;
    }
}

// the tau cells of every run for the launch's n steps from k_off, into a->tcell
void batch_tcell(lmh_batch* b, BatchArgs* a, int32_t k_off, int32_t n) {
    const int32_t R = (int32_t)b->runs.size();
    hipLaunchKernelGGL(k_batch_tcell, dim3((unsigned)(((int64_t)R * n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       b->stream, *a, a->runs, R, k_off, n, b->d_tcell);
    a->tcell = b->d_tcell;
    a->tcell_k = n;
}

// steps *d_step + 0 .. nsteps-1 as launches of steps_per_launch steps, then *d_step += nsteps.
// In the specialised configuration the runs with medium slices take k_batch_grid_sl, one step per
// launch, reading the node slices that one k_batch_nodes launch wrote for up to kSliceRing steps;
// the others k_batch_grid over their own tiles.
void batch_enqueue(lmh_batch* b, const BatchArgs& a0, int coll, int64_t nsteps) {
    const int32_t k = std::max(1, b->tune.steps_per_launch);
    const bool split = b->nsr > 0 && batch_spec(a0.base.p);
    BatchArgs a = a0;
    if (split) {
        a.tile_list = b->d_tile_co;
        a.ntiles = b->ntiles_co;
    }
    unsigned grid = (unsigned)a.ntiles;
    if (b->tune.persistent) {
        // enough blocks to fill every CU at the kernel's occupancy, walking the tiles
        const int per_cu = std::max(1, 2048 / std::max(64, b->tune.block));
        grid = (unsigned)std::min<int64_t>(a.ntiles, (int64_t)b->cus * per_cu);
    }
    if (split) {
        BatchArgs as = a0;
        as.tile_list = b->d_tile_sl;
        as.ntiles = b->ntiles_sl;
        const dim3 G((unsigned)b->max_node_blocks, (unsigned)b->nsr, 1u);
        for (int64_t s0 = 0; s0 < nsteps; s0 += kSliceRing) {
            const int32_t n = (int32_t)std::min<int64_t>(kSliceRing, nsteps - s0);
            BatchArgs at = a0;
            batch_tcell(b, &at, (int32_t)s0, n);
            hipLaunchKernelGGL(k_batch_nodes, G, dim3(kBlock), 0, b->stream, at, at.runs, b->d_srun, n);
            for (int32_t s = 0; s < n; ++s) {
                if (b->pdim == 2) batch_launch_sl<2>(b, as, coll, (int32_t)(s0 + s), s);
                else batch_launch_sl<3>(b, as, coll, (int32_t)(s0 + s), s);
            }
        }
    }
    for (int64_t s0 = 0; s0 < nsteps && a.ntiles > 0; s0 += k) {
        const int32_t n = (int32_t)std::min<int64_t>(k, nsteps - s0);
        batch_tcell(b, &a, (int32_t)s0, n);
        if (b->nsr > 0 && !split) {
            // the sliced runs' medium for the launch's n steps
            const dim3 G((unsigned)b->max_node_blocks, (unsigned)b->nsr, 1u);
            hipLaunchKernelGGL(k_batch_nodes, G, dim3(kBlock), 0, b->stream, a, a.runs, b->d_srun, n);
        }
        if (b->pdim == 2) {
            if (b->tune.persistent) batch_launch_p<2, true>(b, a, coll, (int32_t)s0, n, grid);
            else batch_launch_p<2, false>(b, a, coll, (int32_t)s0, n, grid);
        } else {
            if (b->tune.persistent) batch_launch_p<3, true>(b, a, coll, (int32_t)s0, n, grid);
            else batch_launch_p<3, false>(b, a, coll, (int32_t)s0, n, grid);
        }
    }
    hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(64), 0, b->stream, b->d_step, nsteps);
}

// the steps through captured graph replays of graph_steps steps (the graph built for the
// current parameters and tuning), the rest as direct launches
int batch_steps(lmh_batch* b, const lmh_params* p, const BatchArgs& a, int64_t nsteps) {
    const int64_t S = b->tune.graph_steps;
    int64_t done = 0;
    if (S > 0 && nsteps >= S) {
        // the key: parameters, tuning (steps per replay included) and the state's and identities'
        // addresses; the graph holds the kernel arguments
        std::vector<unsigned char> key(sizeof(lmh_params) + sizeof(lmh_batch_tuning) + 2 * sizeof(void*));
        unsigned char* kp = key.data();
        memcpy(kp, p, sizeof(lmh_params));
        memcpy(kp += sizeof(lmh_params), &b->tune, sizeof(lmh_batch_tuning));
        memcpy(kp += sizeof(lmh_batch_tuning), &b->d_state, sizeof(void*));
        memcpy(kp + sizeof(void*), &b->d_id, sizeof(void*));
        int k = -1;
        for (int c = 0; c < 2; ++c)
            if (b->gexec[c] && b->gkey[c] == key) k = c;
        if (k < 0) {
            // an empty slot, else the one not replayed last
            k = !b->gexec[0] ? 0 : (!b->gexec[1] ? 1 : b->glast ^ 1);
            batch_drop_graph_slot(b, k);
            LMB_HIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            batch_enqueue(b, a, p->collision, S);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (e != hipSuccess) return b->fail(LMH_E_HIP, std::string("batch: graph capture: ") + hipGetErrorString(e));
            b->graph[k] = g;
            e = hipGraphInstantiate(&b->gexec[k], g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                batch_drop_graph_slot(b, k);
                return b->fail(LMH_E_HIP, std::string("batch: graph instantiation: ") + hipGetErrorString(e));
            }
            b->gkey[k] = key;
        }
        b->glast = k;
        for (; done + S <= nsteps; done += S) LMB_HIP(b, hipGraphLaunch(b->gexec[k], b->stream));
    }
    if (done < nsteps) batch_enqueue(b, a, p->collision, nsteps - done);
    LMB_HIP(b, hipGetLastError());
    return LMH_OK;
}

int batch_run_impl(lmh_batch* b, const lmh_params* p, int64_t step_begin, int64_t nsteps) {
    int rc = batch_layout(b, b->tune.block);
    if (rc) return rc;
    if (b->ntiles == 0 || nsteps == 0) return LMH_OK;
    LMB_HIP(b, hipMemcpyAsync(b->d_step, &step_begin, sizeof(int64_t), hipMemcpyHostToDevice, b->stream));
    LMB_HIP(b, hipStreamSynchronize(b->stream));
    // cell order: a sort before the first step and whenever sort_every steps have passed, checked
    // between chunks of whole graph replays
    const int64_t E = b->tune.sort_every;
    const bool sorting = E > 0 && b->d_id != nullptr;
    if (!sorting && (rc = batch_ensure_unsorted(b))) return rc;
    const int64_t S = std::max<int64_t>(1, b->tune.graph_steps);
    for (int64_t done = 0; done < nsteps;) {
        int64_t chunk = nsteps - done;
        if (sorting) {
            if (!b->is_sorted || b->steps_since_sort >= E) {
                if ((rc = batch_sort(b, p))) return rc;
            }
            chunk = std::min<int64_t>(chunk, ((E - b->steps_since_sort + S - 1) / S) * S);
        }
        if (b->perm_pending) {
            // the first step after the sort, launched directly, reads every quark through the
            // permutation and writes the new order into the alternate buffers, which then become
            // the current ones
            BatchArgs g = batch_args(b, p);
            for (int d = 0; d < 3; ++d) {
                if (g.base.wpos[d]) g.base.wpos[d] = b->d_state_alt + (g.base.pos[d] - b->d_state);
                if (g.base.wmom[d]) g.base.wmom[d] = b->d_state_alt + (g.base.mom[d] - b->d_state);
            }
            g.base.perm = b->d_keys + 3 * b->N;
            g.base.wid = b->d_id_alt;
            batch_enqueue(b, g, p->collision, 1);
            LMB_HIP(b, hipGetLastError());
            std::swap(b->d_state, b->d_state_alt);
            std::swap(b->d_id, b->d_id_alt);
            b->perm_pending = false;
            done += 1;
            b->steps_since_sort += 1;
            if (--chunk == 0) continue;
        }
        // the arguments after the sort, which moves the state to the other buffer
        const BatchArgs a = batch_args(b, p);
        if ((rc = batch_steps(b, p, a, chunk))) return rc;
        done += chunk;
        b->steps_since_sort += chunk;
    }
    if (p->check_finite) {
        unsigned long long bad = ULLONG_MAX;
        LMB_HIP(b, hipMemcpyAsync(&bad, b->d_bad, sizeof(bad), hipMemcpyDeviceToHost, b->stream));
        LMB_HIP(b, hipStreamSynchronize(b->stream));
        if (bad != ULLONG_MAX)
            return b->fail(LMH_E_NONFINITE, "Breaking simulation due to NaN/Inf at step " + std::to_string((long long)bad));
    }
    return LMH_OK;
}

}  // namespace

extern "C" {

int lmh_batch_create(lmh_batch** out, int32_t device) {
    if (!out) return LMH_E_INVALID;
    *out = nullptr;
    lmh_batch* b = new (std::nothrow) lmh_batch();
    if (!b) return LMH_E_NOMEM;
    b->device = device;
    hipDeviceProp_t prop;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return LMH_E_HIP;
    }
    b->cus = prop.multiProcessorCount;
    unsigned long long none = ULLONG_MAX;
    if (hipMalloc((void**)&b->d_bad, sizeof(none)) != hipSuccess || hipMalloc((void**)&b->d_step, sizeof(int64_t)) != hipSuccess ||
        hipMemcpy(b->d_bad, &none, sizeof(none), hipMemcpyHostToDevice) != hipSuccess) {
        dfree(b->d_bad);
        dfree(b->d_step);
        (void)hipStreamDestroy(b->stream);
        delete b;
        return LMH_E_HIP;
    }
    *out = b;
    return LMH_OK;
}

void lmh_batch_destroy(lmh_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_drop_graph(b);
    for (auto& m : b->media) {
        dfree(m.nodes);
        dfree(m.tg);
    }
    for (auto& t : b->tabs) dfree(t.d);
    for (auto& t : b->jtabs) dfree(t.d);
    dfree(b->d_state);
    dfree(b->d_bad);
    dfree(b->d_step);
    dfree(b->d_begin);
    dfree(b->d_runs);
    dfree(b->d_tile_run);
    dfree(b->d_srun);
    dfree(b->d_tile_sl);
    dfree(b->d_tile_co);
    dfree(b->d_bslice);
    dfree(b->d_tcell);
    dfree(b->d_obs);
    dfree(b->d_save);
    batch_free_order(b);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char* lmh_batch_last_error(const lmh_batch* b) { return b ? b->err.c_str() : "null batch"; }

void* lmh_batch_stream(lmh_batch* b) { return b ? (void*)b->stream : nullptr; }

int lmh_batch_add_medium_grid2p1(lmh_batch* b, double x0, double dx, int32_t nx, double y0, double dy,
                                 int32_t ny, const double* tgrid, int32_t nt, const double* T,
                                 const double* vx, const double* vy, int32_t* index) {
    if (!b) return LMH_E_INVALID;
    if (!index || !tgrid || !T || !vx || !vy || nx < 2 || ny < 2 || nt < 2 || !(dx > 0.0) || !(dy > 0.0) ||
        !std::isfinite(x0) || !std::isfinite(y0))
        return b->fail(LMH_E_INVALID, "batch_add_medium_grid2p1: need nx, ny, nt >= 2, dx, dy > 0 and all fields");
    if (!sorted(tgrid, nt)) return b->fail(LMH_E_INVALID, "batch_add_medium_grid2p1: tgrid must be sorted");
    LMB_HIP(b, hipSetDevice(b->device));
    const size_t n = (size_t)nx * (size_t)ny * (size_t)nt;
    std::vector<double4> nodes;
    try {
        nodes.resize(n);
    } catch (...) {
        return b->fail(LMH_E_NOMEM, "batch_add_medium_grid2p1: host allocation failed");
    }
    for (size_t k = 0; k < n; ++k) nodes[k] = make_double4(T[k], vx[k], vy[k], 0.0);
    lmh_batch::Med m;
    if (hipMalloc((void**)&m.nodes, n * sizeof(double4)) != hipSuccess ||
        hipMalloc((void**)&m.tg, (size_t)nt * sizeof(double)) != hipSuccess) {
        dfree(m.nodes);
        dfree(m.tg);
        return b->fail(LMH_E_NOMEM, "batch_add_medium_grid2p1: device allocation failed");
    }
    b->media.push_back(m);     // owned from here on, freed by lmh_batch_destroy
    LMB_HIP(b, hipMemcpy(m.nodes, nodes.data(), n * sizeof(double4), hipMemcpyHostToDevice));
    LMB_HIP(b, hipMemcpy(m.tg, tgrid, (size_t)nt * sizeof(double), hipMemcpyHostToDevice));
    lmh_batch::Med& mm = b->media.back();
    mm.nx = nx;
    mm.ny = ny;
    mm.nt = nt;
    mm.x0 = x0;
    mm.dx = dx;
    mm.y0 = y0;
    mm.dy = dy;
    *index = (int32_t)b->media.size() - 1;
    return LMH_OK;
}

int lmh_batch_add_time_table(lmh_batch* b, double Tmin, double invdT, const double* vals, int32_t n,
                             int32_t* index) {
    if (!b) return LMH_E_INVALID;
    if (!index || !vals || n < 1) return b->fail(LMH_E_INVALID, "batch_add_time_table: need values");
    LMB_HIP(b, hipSetDevice(b->device));
    lmh_batch::Tab t;
    if (hipMalloc((void**)&t.d, (size_t)n * sizeof(double)) != hipSuccess)
        return b->fail(LMH_E_NOMEM, "batch_add_time_table: device allocation failed");
    b->tabs.push_back(t);
    LMB_HIP(b, hipMemcpy(t.d, vals, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    b->tabs.back().Tmin = Tmin;
    b->tabs.back().invdT = invdT;
    b->tabs.back().n = n;
    *index = (int32_t)b->tabs.size() - 1;
    return LMH_OK;
}

int lmh_batch_add_juttner_table(lmh_batch* b, const double* invcdf, int32_t nU, int32_t nT, double Tmin,
                                double invdT, int32_t* index) {
    if (!b) return LMH_E_INVALID;
    if (!index || !invcdf || nU < 2 || nT < 2) return b->fail(LMH_E_INVALID, "batch_add_juttner_table: bad size");
    LMB_HIP(b, hipSetDevice(b->device));
    lmh_batch::JTab t;
    const size_t n = (size_t)nU * (size_t)nT;
    if (hipMalloc((void**)&t.d, n * sizeof(double)) != hipSuccess)
        return b->fail(LMH_E_NOMEM, "batch_add_juttner_table: device allocation failed");
    b->jtabs.push_back(t);
    LMB_HIP(b, hipMemcpy(t.d, invcdf, n * sizeof(double), hipMemcpyHostToDevice));
    lmh_batch::JTab& tt = b->jtabs.back();
    tt.Tmin = Tmin;
    tt.invdT = invdT;
    tt.nU = nU;
    tt.nT = nT;
    *index = (int32_t)b->jtabs.size() - 1;
    return LMH_OK;
}

int lmh_batch_set_runs(lmh_batch* b, const lmh_run_desc* runs, int32_t nruns, const double* pos,
                       const double* mom, int32_t pdim) {
    if (!b) return LMH_E_INVALID;
    if (!runs || nruns < 1 || !pos || !mom || (pdim != 2 && pdim != 3))
        return b->fail(LMH_E_INVALID, "batch_set_runs: need runs, positions and momenta (pdim 2 or 3)");
    int64_t N = 0;
    for (int32_t r = 0; r < nruns; ++r) {
        const lmh_run_desc& d = runs[r];
        if (d.count < 0 || d.count > INT32_MAX || !(d.mass > 0.0) || d.medium < 0 || d.medium >= (int32_t)b->media.size() ||
            d.drag_table >= (int32_t)b->tabs.size() || d.rta_table >= (int32_t)b->tabs.size() ||
            d.juttner_table >= (int32_t)b->jtabs.size())
            return b->fail(LMH_E_INVALID, "batch_set_runs: run " + std::to_string(r) +
                                              " has a negative count, a non-positive mass or an unknown medium or table");
        N += d.count;
    }
    if (N < 1 || N > ((int64_t)1 << 31) * 64) return b->fail(LMH_E_INVALID, "batch_set_runs: bad total particle count");
    LMB_HIP(b, hipSetDevice(b->device));
    LMB_HIP(b, hipStreamSynchronize(b->stream));
    const size_t rows = (size_t)(2 + pdim);
    std::vector<double> soa;
    try {
        soa.resize(rows * (size_t)N);
    } catch (...) {
        return b->fail(LMH_E_NOMEM, "batch_set_runs: host allocation failed");
    }
    for (int64_t i = 0; i < N; ++i) {
        for (int d = 0; d < 2; ++d) soa[(size_t)d * N + i] = pos[d + 2 * i];
        for (int d = 0; d < pdim; ++d) soa[(size_t)(2 + d) * N + i] = mom[d + (int64_t)pdim * i];
    }
    batch_drop_graph(b);
    dfree(b->d_state);
    dfree(b->d_begin);
    dfree(b->d_runs);
    dfree(b->d_tile_run);
    dfree(b->d_srun);
    dfree(b->d_tile_sl);
    dfree(b->d_tile_co);
    b->nsr = 0;
    b->ntiles_sl = b->ntiles_co = 0;
    batch_free_order(b);
    b->tiles_block = 0;
    b->runs.assign(runs, runs + nruns);
    b->begin.assign(nruns + 1, 0);
    for (int32_t r = 0; r < nruns; ++r) b->begin[r + 1] = b->begin[r] + runs[r].count;
    b->N = N;
    b->pdim = pdim;
    if (hipMalloc((void**)&b->d_state, rows * (size_t)N * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&b->d_begin, (size_t)(nruns + 1) * sizeof(int64_t)) != hipSuccess)
        return b->fail(LMH_E_NOMEM, "batch_set_runs: device allocation failed");
    LMB_HIP(b, hipMemcpy(b->d_state, soa.data(), rows * (size_t)N * sizeof(double), hipMemcpyHostToDevice));
    LMB_HIP(b, hipMemcpy(b->d_begin, b->begin.data(), (size_t)(nruns + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    // identities: every quark in its own slot
    if (N < ((int64_t)1 << 32)) {
        if (hipMalloc((void**)&b->d_id, (size_t)N * sizeof(uint32_t)) != hipSuccess)
            return b->fail(LMH_E_NOMEM, "batch_set_runs: device allocation failed");
        hipLaunchKernelGGL(k_iota, dim3(grid_blocks(N)), dim3(kBlock), 0, b->stream, b->d_id, N);
        LMB_HIP(b, hipGetLastError());
    }
    unsigned long long none = ULLONG_MAX;
    LMB_HIP(b, hipMemcpy(b->d_bad, &none, sizeof(none), hipMemcpyHostToDevice));
    return batch_layout(b, b->tune.block);
}

int lmh_batch_get_particles(lmh_batch* b, double* pos, double* mom) {
    if (!b) return LMH_E_INVALID;
    if (!b->d_state) return b->fail(LMH_E_STATE, "batch_get_particles: no runs set");
    if (!pos || !mom) return b->fail(LMH_E_INVALID, "batch_get_particles: null output");
    LMB_HIP(b, hipSetDevice(b->device));
    if (int rc = batch_ensure_unsorted(b)) return rc;
    const size_t rows = (size_t)(2 + b->pdim);
    const int64_t N = b->N;
    std::vector<double> soa;
    try {
        soa.resize(rows * (size_t)N);
    } catch (...) {
        return b->fail(LMH_E_NOMEM, "batch_get_particles: host allocation failed");
    }
    LMB_HIP(b, hipMemcpyAsync(soa.data(), b->d_state, rows * (size_t)N * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    LMB_HIP(b, hipStreamSynchronize(b->stream));
    for (int64_t i = 0; i < N; ++i) {
        for (int d = 0; d < 2; ++d) pos[d + 2 * i] = soa[(size_t)d * N + i];
        for (int d = 0; d < b->pdim; ++d) mom[d + (int64_t)b->pdim * i] = soa[(size_t)(2 + d) * N + i];
    }
    return LMH_OK;
}

int lmh_batch_set_tuning(lmh_batch* b, const lmh_batch_tuning* t) {
    if (!b) return LMH_E_INVALID;
    if (!t || (t->block != 64 && t->block != 128 && t->block != 256) || t->steps_per_launch < 1 ||
        (t->persistent != 0 && t->persistent != 1) || t->graph_steps < 0 || t->sort_every < 0 ||
        t->slice_min < 0)
        return b->fail(LMH_E_INVALID, "batch_set_tuning: block in {64, 128, 256}, steps_per_launch >= 1, "
                                      "persistent in {0, 1}, graph_steps >= 0, sort_every >= 0, "
                                      "slice_min >= 0");
    b->tune = *t;
    return LMH_OK;
}

int lmh_batch_get_tuning(lmh_batch* b, lmh_batch_tuning* t) {
    if (!b || !t) return LMH_E_INVALID;
    *t = b->tune;
    return LMH_OK;
}

int lmh_batch_boost_to_lab(lmh_batch* b, const lmh_params* p, int64_t step) {
    if (!b) return LMH_E_INVALID;
    int rc = batch_check(b, p);
    if (rc) return rc;
    LMB_HIP(b, hipSetDevice(b->device));
    if ((rc = batch_layout(b, b->tune.block))) return rc;
    if (b->ntiles == 0) return LMH_OK;
    const BatchArgs a = batch_args(b, p);
    const dim3 G((unsigned)b->ntiles), B(b->tune.block);
    if (b->pdim == 2) hipLaunchKernelGGL((k_batch_boost_lab<2>), G, B, 0, b->stream, a, step);
    else hipLaunchKernelGGL((k_batch_boost_lab<3>), G, B, 0, b->stream, a, step);
    LMB_HIP(b, hipGetLastError());
    return LMH_OK;
}

int lmh_batch_run(lmh_batch* b, const lmh_params* p, int64_t step_begin, int64_t nsteps) {
    if (!b) return LMH_E_INVALID;
    int rc = batch_check(b, p);
    if (rc) return rc;
    if (nsteps < 0 || step_begin < 0 || step_begin + nsteps > ((int64_t)1 << 32))
        return b->fail(LMH_E_INVALID, "batch_run: steps must lie in [0, 2^32)");
    if (p->stream >= ((uint64_t)1 << 61) - b->runs.size())
        return b->fail(LMH_E_INVALID, "batch_run: the Philox stream plus the run count must stay below 2^61");
    LMB_HIP(b, hipSetDevice(b->device));
    return batch_run_impl(b, p, step_begin, nsteps);
}

int lmh_batch_autotune(lmh_batch* b, const lmh_params* p, int32_t trial_steps, lmh_batch_tuning* best,
                       double* best_ms_per_step) {
    if (!b) return LMH_E_INVALID;
    int rc = batch_check(b, p);
    if (rc) return rc;
    if (trial_steps < 1) return b->fail(LMH_E_INVALID, "batch_autotune: trial_steps >= 1");
    LMB_HIP(b, hipSetDevice(b->device));
    if ((rc = batch_ensure_unsorted(b))) return rc;
    const size_t n = (size_t)(2 + b->pdim) * (size_t)b->N;
    if ((rc = batch_dev(b, &b->d_save, &b->save_cap, n, "batch_autotune"))) return rc;
    LMB_HIP(b, hipMemcpyAsync(b->d_save, b->d_state, n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
    lmh_params q = *p;
    q.check_finite = 0;
    const lmh_batch_tuning keep = b->tune;
    lmh_batch_tuning pick = keep;
    double pick_ms = 0.0;
    hipEvent_t e0, e1;
    LMB_HIP(b, hipEventCreate(&e0));
    LMB_HIP(b, hipEventCreate(&e1));
    const int blocks[3] = {64, 128, 256};
    const int spls[3] = {1, 2, 4};
    // medium slices for no run, or for the runs of at least 2^14, 2^16 or 2^18 quarks
    const int slice_mins[4] = {0, 1 << 14, 1 << 16, 1 << 18};
    for (int bi = 0; bi < 3 && rc == LMH_OK; ++bi)
        for (int pe = 0; pe < 2 && rc == LMH_OK; ++pe)
            for (int si = 0; si < 3 * 4 && rc == LMH_OK; ++si) {
                b->tune = lmh_batch_tuning{blocks[bi], spls[si % 3], pe, 0, 0, slice_mins[si / 3]};
                // one untimed launch for the code object and caches, then the timed steps
                if ((rc = batch_run_impl(b, &q, 1, spls[si % 3]))) break;
                if (hipEventRecord(e0, b->stream) != hipSuccess) { rc = b->fail(LMH_E_HIP, "batch_autotune: event"); break; }
                if ((rc = batch_run_impl(b, &q, 1 + spls[si % 3], trial_steps))) break;
                if (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                    rc = b->fail(LMH_E_HIP, "batch_autotune: event");
                    break;
                }
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                const double per = ms / trial_steps;
                if (pick_ms == 0.0 || per < pick_ms) {
                    pick_ms = per;
                    pick = b->tune;
                }
                if (hipMemcpyAsync(b->d_state, b->d_save, n * sizeof(double), hipMemcpyDeviceToDevice, b->stream) != hipSuccess) {
                    rc = b->fail(LMH_E_HIP, "batch_autotune: restore");
                    break;
                }
            }
    // the sort cadence at the fastest shape: kTuneSortSteps steps from the same state for each
    // cadence, the sorts included (the first before the first step)
    constexpr int64_t kTuneSortSteps = 200;
    const int cadences[4] = {0, 50, 100, 200};
    double pick_sort_ms = 0.0;
    int pick_sort = 0;
    for (int ci = 0; ci < 4 && rc == LMH_OK && b->d_id; ++ci) {
        b->tune = pick;
        b->tune.graph_steps = 0;
        b->tune.sort_every = cadences[ci];
        if (hipEventRecord(e0, b->stream) != hipSuccess) { rc = b->fail(LMH_E_HIP, "batch_autotune: event"); break; }
        if ((rc = batch_run_impl(b, &q, 1, kTuneSortSteps))) break;
        if (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            rc = b->fail(LMH_E_HIP, "batch_autotune: event");
            break;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double per = ms / kTuneSortSteps;
        if (ci == 0 || per < pick_sort_ms) {
            pick_sort_ms = per;
            pick_sort = cadences[ci];
        }
        // the state as before, every quark in its own slot again
        if (hipMemcpyAsync(b->d_state, b->d_save, n * sizeof(double), hipMemcpyDeviceToDevice, b->stream) != hipSuccess) {
            rc = b->fail(LMH_E_HIP, "batch_autotune: restore");
            break;
        }
        hipLaunchKernelGGL(k_iota, dim3(grid_blocks(b->N)), dim3(kBlock), 0, b->stream, b->d_id, b->N);
        b->is_sorted = false;
        b->steps_since_sort = 0;
    }
    pick.sort_every = pick_sort;
    if (pick_sort_ms > 0.0) pick_ms = pick_sort_ms;
    if (b->d_id && b->is_sorted) {
        hipLaunchKernelGGL(k_iota, dim3(grid_blocks(b->N)), dim3(kBlock), 0, b->stream, b->d_id, b->N);
        b->is_sorted = false;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // the state as before the trials, the first non-finite step cleared
    LMB_HIP(b, hipMemcpyAsync(b->d_state, b->d_save, n * sizeof(double), hipMemcpyDeviceToDevice, b->stream));
    unsigned long long none = ULLONG_MAX;
    LMB_HIP(b, hipMemcpyAsync(b->d_bad, &none, sizeof(none), hipMemcpyHostToDevice, b->stream));
    LMB_HIP(b, hipStreamSynchronize(b->stream));
    if (rc) {
        b->tune = keep;
        return rc;
    }
    pick.graph_steps = keep.graph_steps;
    b->tune = pick;
    if (best) *best = pick;
    if (best_ms_per_step) *best_ms_per_step = pick_ms;
    return LMH_OK;
}

int lmh_batch_observables(lmh_batch* b, const lmh_obs_spec* s, double* out) {
    if (!b) return LMH_E_INVALID;
    if (!s || !out) return b->fail(LMH_E_INVALID, "batch_observables: null spec or output");
    if (!b->d_state) return b->fail(LMH_E_STATE, "batch_observables: no runs set");
    if (s->nbins < 0 || s->nbins > kObsMaxBins || (s->nbins > 0 && !(s->pt_max > s->pt_min)))
        return b->fail(LMH_E_INVALID, "batch_observables: 0 <= nbins <= 120 and pt_max > pt_min");
    LMB_HIP(b, hipSetDevice(b->device));
    // the sums in the caller's order (the same bits as the run alone)
    if (int rc = batch_ensure_unsorted(b)) return rc;
    const int K = 5 + s->nbins;
    const int R = (int)b->runs.size();
    const int nblk = 16;
    int rc = batch_dev(b, &b->d_obs, &b->obs_cap, (size_t)K * R * (nblk + 1), "batch_observables");
    if (rc) return rc;
    const double* px = b->d_state + 2 * b->N;
    hipLaunchKernelGGL(k_batch_obs_partial, dim3(nblk, R), dim3(kObsBlock), 0, b->stream, px, px + b->N,
                       b->d_begin, *s, b->d_obs);
    double* fin = b->d_obs + (size_t)K * R * nblk;
    hipLaunchKernelGGL(k_batch_obs_final, dim3(R), dim3(kObsBlock), 0, b->stream, b->d_obs, nblk, K, fin);
    LMB_HIP(b, hipGetLastError());
    LMB_HIP(b, hipMemcpyAsync(out, fin, (size_t)K * R * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    LMB_HIP(b, hipStreamSynchronize(b->stream));
    return LMH_OK;
}

int64_t lmh_batch_nonfinite_step(lmh_batch* b) {
    if (!b || !b->d_bad) return -1;
    unsigned long long bad = ULLONG_MAX;
    if (hipSetDevice(b->device) != hipSuccess) return -1;
    if (hipMemcpyAsync(&bad, b->d_bad, sizeof(bad), hipMemcpyDeviceToHost, b->stream) != hipSuccess) return -1;
    if (hipStreamSynchronize(b->stream) != hipSuccess) return -1;
    return bad == ULLONG_MAX ? -1 : (int64_t)bad;
}

}  // extern "C"
