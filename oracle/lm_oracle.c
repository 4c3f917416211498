/*
 * lm_oracle.c -- CPU restatement of LangevInMedium.jl's per-particle Langevin step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker that the HIP backend is compared against.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it.
 *
 * The reference is Julia and cannot run in this environment, so every function below restates
 * one reference function, in the reference's own order of operations, and cites its file:line
 * (paths relative to the LangevInMedium.jl repository root).  The step keeps the reference's
 * structure: one pass over all particles per kernel, with the noise the driver would draw
 * (`randn`, `rand`) supplied instead as a fixed-noise record.  Built with
 * -ffp-contract=off so that no multiply-add is fused that Julia would round twice.
 *
 * Functions with no reference counterpart (the 2+1D grid medium, species segments, the
 * LANGEVIN_RTA composite, the GAMMA_P drag, the post-point integrator, Philox noise and the
 * observables) are build extensions; their definitions are stated where they appear.
 */
#include "lm_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define JL_EPS DBL_EPSILON            /* Julia eps() */
static const double FMGEV = 1.0 / 0.197327;   /* constants.jl:13 (GevInvTofm), :23 (fmGeV) */

double lmo_fmGeV(void) { return FMGEV; }

/* Julia's clamp(x, lo, hi) = x > hi ? hi : (x < lo ? lo : x) */
static inline double jl_clamp(double x, double lo, double hi) {
    return x > hi ? hi : (x < lo ? lo : x);
}
/* Julia's max/min for the finite, unsigned-zero cases used here */
static inline double jl_max(double a, double b) { return (b > a) ? b : a; }

/* Base.searchsortedlast on a sorted vector: the last 1-based index k with v[k] <= x, 0 if none. */
int64_t lmo_searchsortedlast(const double* v, int64_t n, double x) {
    int64_t lo = 0, hi = n + 1;
    while (lo < hi - 1) {
        int64_t m = lo + ((hi - lo) >> 1);
        if (x < v[m - 1]) hi = m; else lo = m;
    }
    return lo;
}

static inline int64_t clamp_i64(int64_t i, int64_t lo, int64_t hi) {
    return i > hi ? hi : (i < lo ? lo : i);
}

/* interpolate_2d_cpu, kernels_cpu.jl:22-51 (query clamp, searchsortedlast, eps guard) */
double lmo_interpolate_2d(const double* x, int32_t nx, const double* y, int32_t ny,
                          const double* values, double xi, double yi) {
    xi = jl_clamp(xi, x[0], x[nx - 1]);
    yi = jl_clamp(yi, y[0], y[ny - 1]);
    int64_t i = lmo_searchsortedlast(x, nx, xi);
    int64_t j = lmo_searchsortedlast(y, ny, yi);
    i = clamp_i64(i, 1, nx - 1);
    j = clamp_i64(j, 1, ny - 1);
    double x0 = x[i - 1], x1 = x[i];
    double y0 = y[j - 1], y1 = y[j];
    double v00 = values[(i - 1) + (int64_t)nx * (j - 1)], v10 = values[i + (int64_t)nx * (j - 1)];
    double v01 = values[(i - 1) + (int64_t)nx * j], v11 = values[i + (int64_t)nx * j];
    double dx = x1 - x0;
    double dy = y1 - y0;
    double xd = fabs(dx) < JL_EPS ? 0.0 : (xi - x0) / dx;
    double yd = fabs(dy) < JL_EPS ? 0.0 : (yi - y0) / dy;
    double c0 = v00 * (1.0 - xd) + v10 * xd;
    double c1 = v01 * (1.0 - xd) + v11 * xd;
    return c0 * (1.0 - yd) + c1 * yd;
}

/* interpolate_2d_cpu of the GC module, kernels_cpu_GC.jl:37-57: no query clamp, no eps guard,
 * so queries off the grid extrapolate the edge cell linearly. */
double lmo_interpolate_2d_gc(const double* x, int32_t nx, const double* y, int32_t ny,
                             const double* values, double xi, double yi) {
    int64_t i = lmo_searchsortedlast(x, nx, xi);
    int64_t j = lmo_searchsortedlast(y, ny, yi);
    i = clamp_i64(i, 1, nx - 1);
    j = clamp_i64(j, 1, ny - 1);
    double x0 = x[i - 1], x1 = x[i];
    double y0 = y[j - 1], y1 = y[j];
    double v00 = values[(i - 1) + (int64_t)nx * (j - 1)], v10 = values[i + (int64_t)nx * (j - 1)];
    double v01 = values[(i - 1) + (int64_t)nx * j], v11 = values[i + (int64_t)nx * j];
    double xd = (xi - x0) / (x1 - x0);
    double yd = (yi - y0) / (y1 - y0);
    double c0 = v00 * (1.0 - xd) + v10 * xd;
    double c1 = v01 * (1.0 - xd) + v11 * xd;
    return c0 * (1.0 - yd) + c1 * yd;
}

/* eval_tau_n_spline, transport.jl:231-253: linear, index clamped, fraction NOT clamped. */
double lmo_eval_time_table(double T, double Tmin, double invdT, const double* vals, int32_t n) {
    if (n < 2) return vals[0];
    if (!isfinite(T)) return 0.0;
    double u = (T - Tmin) * invdT;
    double fu = floor(u);
    int64_t i;                                   /* 1-based, clamp(Int(floor(u)) + 1, 1, n-1) */
    if (!(fu >= 0.0)) i = 1;
    else if (fu > (double)(n - 2)) i = n - 1;
    else i = (int64_t)fu + 1;
    double t = u - (double)(i - 1);
    double y0 = vals[i - 1];
    double y1 = vals[i];
    return (1.0 - t) * y0 + t * y1;
}

/* tau_drag, transport.jl:132-137: m*DsT/T^2 in GeV^-1, converted to fm. */
double lmo_tau_drag(double T, double m, double DsT) {
    if (!isfinite(T) || !isfinite(m) || !isfinite(DsT) || T <= 0.0 || m <= 0.0 || DsT <= 0.0)
        return 0.0;
    return m * DsT / (T * T) / FMGEV;
}

/* effective_DsT, transport.jl:139-156 */
double lmo_effective_DsT(double T, double DsT, int32_t linear, double slope, double offset,
                         double Tfo, int32_t quad, double Tref, int32_t* err) {
    if (err) *err = 0;
    if (quad) {
        if (linear || !(Tref > 0.0)) { if (err) *err = 1; return 0.0; }
        double x = T / Tref;
        return DsT * x * x;
    }
    if (!linear) return DsT;
    double Tf = jl_max(T, Tfo);
    return slope * Tf + offset;
}

/*
 * Exponentially scaled modified Bessel function K_nu(z)*exp(z), the quantity Bessels.besselkx
 * returns (transport.jl:77-78,87-88).  Own implementation from the integral representation
 *     K_nu(z) e^z = int_0^inf exp(-z (cosh t - 1)) cosh(nu t) dt ,   z > 0,
 * by the trapezoid rule, which converges geometrically for this analytic, doubly-exponentially
 * decaying integrand; cosh t - 1 = 2 sinh^2(t/2) avoids cancellation near t = 0.  The step is
 * kept below a quarter of the integrand's width 1/sqrt(z) for large z.
 */
double lmo_besselkx(double nu, double z) {
    if (!(z > 0.0)) return NAN;
    double h = 1.0 / 32.0;
    double hz = 0.25 / sqrt(z);
    if (hz < h) h = hz;
    double sum = 0.5;                               /* t = 0, half weight */
    for (int64_t k = 1; k < 2000000; ++k) {
        double t = (double)k * h;
        double s = sinh(0.5 * t);
        double e = -2.0 * z * s * s;
        double term = 0.5 * (exp(e + nu * t) + exp(e - nu * t));
        sum += term;
        if (term < 1e-20 * sum && e < -50.0) break;
    }
    return sum * h;
}

static inline double jl_sign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : x); }

/* tau_n_main3, transport.jl:58-104 (z > 50 fast path at :81-85) */
double lmo_tau_n_main3(double T, double m, double DsT) {
    if (!isfinite(T) || !isfinite(m) || !isfinite(DsT) || T <= 0.0 || m <= 0.0 || DsT <= 0.0)
        return 0.0;
    double z = m / T;
    if (z <= 0.0) return 0.0;
    double K2x = lmo_besselkx(2.0, z);
    double K3x = lmo_besselkx(3.0, z);
    double denominator = jl_max(fabs(K2x), 1e-300);
    if (z > 50.0) {
        double ratio_K3K2 = K3x / denominator * jl_sign(K2x == 0.0 ? 1.0 : K2x);
        double tau_GeVinv = (DsT / T) * z * ratio_K3K2;
        double r = tau_GeVinv / FMGEV;
        return r < 1e20 ? r : 1e20;
    }
    double K1x = lmo_besselkx(1.0, z);
    double K5x = lmo_besselkx(5.0, z);
    double numerator = 2.0 * K1x - 3.0 * K3x + K5x;
    double ratio = numerator / denominator * jl_sign(K2x == 0.0 ? 1.0 : K2x);
    double z3_over_Tm = z * z * z / T;
    if (!isfinite(z3_over_Tm) || z3_over_Tm > 1e50) z3_over_Tm = 1e50;
    double tau_GeVinv = (DsT / 48.0) * z3_over_Tm * ratio;
    if (!isfinite(tau_GeVinv) || fabs(tau_GeVinv) > 1e50) return 1e50 * jl_sign(tau_GeVinv);
    return tau_GeVinv / FMGEV;
}

/* _build_time_spline, transport.jl:165-196; which = LMH_TABLE_DRAG uses tau_drag
 * (build_tau_drag_spline, :208-209), LMH_TABLE_RTA uses tau_n_main3
 * (build_taun_current_spline, :221-222).  scale is LV_TAUN_SCALE (:9-13,193). */
int lmo_build_time_spline(int32_t which, double m, double DsT, double Tmin, double Tmax,
                          int32_t nT, int32_t linear, double slope, double offset, double Tfo,
                          int32_t quad, double Tref, double scale,
                          double* Tmin_out, double* invdT_out, double* vals) {
    if (nT < 2) return LMO_E_INVALID;
    if (!(Tmin > 0.0)) return LMO_E_INVALID;
    if (!(Tmax > Tmin)) return LMO_E_INVALID;
    double dT = (Tmax - Tmin) / (double)(nT - 1);
    double invdT = 1.0 / dT;
    for (int32_t i = 0; i < nT; ++i) {
        double Ti = Tmin + (double)i * dT;
        int32_t err = 0;
        double DsT_eff = lmo_effective_DsT(Ti, DsT, linear, slope, offset, Tfo, quad, Tref, &err);
        if (err) return LMO_E_INVALID;
        double t = which == LMH_TABLE_DRAG ? lmo_tau_drag(Ti, m, DsT_eff)
                                           : lmo_tau_n_main3(Ti, m, DsT_eff);
        vals[i] = t * scale;
    }
    *Tmin_out = Tmin;
    *invdT_out = invdT;
    return LMO_OK;
}

/* p^a for the small non-negative integer exponents used here; Julia's Float64^Int is
 * correctly rounded for a <= 2, which plain products reproduce. */
static inline double ipow(double p, int a) {
    if (a == 0) return 1.0;
    if (a == 1) return p;
    if (a == 2) return p * p;
    return p * p * p;
}

/* build_juttner_invcdf, transport.jl:277-365 */
int lmo_build_juttner_invcdf(double m, int32_t d, double Tmin, double Tmax, int32_t nT,
                             int32_t nU, int32_t np, double quantile_cap,
                             double* invcdf, double* Tmin_out, double* invdT_out) {
    double M = m;
    int a = d - 1;
    if (nT < 2 || nU < 2 || np < 2 || a < 0 || a > 3) return LMO_E_INVALID;
    double Tmin_f = jl_max(Tmin, 1e-4);
    double Tmax_f = jl_max(Tmax, Tmin_f * (1.0 + 1e-6));
    double span = Tmax_f - Tmin_f;
    Tmin_f = jl_max(Tmin_f - 0.02 * span, 1e-4);
    Tmax_f = Tmax_f + 0.02 * span;
    double dT = (Tmax_f - Tmin_f) / (double)(nT - 1);
    double invdT = 1.0 / dT;
    double* pgrid = (double*)malloc(sizeof(double) * (size_t)np);
    double* cdf = (double*)malloc(sizeof(double) * (size_t)np);
    if (!pgrid || !cdf) { free(pgrid); free(cdf); return LMO_E_INVALID; }
    for (int32_t jT = 0; jT < nT; ++jT) {
        double T = Tmin_f + (double)jT * dT;
        double pmax = 16.0 * sqrt(M * T) + 16.0 * T;
        double dp = pmax / (double)(np - 1);
        pgrid[0] = 0.0;
        double wprev = (a == 0) ? 1.0 : 0.0;
        cdf[0] = 0.0;
        for (int32_t k = 1; k < np; ++k) {
            double p = (double)k * dp;
            pgrid[k] = p;
            double E = sqrt(M * M + p * p);
            double w = (a == 0) ? exp(-(E - M) / T) : ipow(p, a) * exp(-(E - M) / T);
            cdf[k] = cdf[k - 1] + 0.5 * (wprev + w) * dp;
            wprev = w;
        }
        double norm = cdf[np - 1];
        double invnorm = norm > 0.0 ? 1.0 / norm : 0.0;
        for (int32_t k = 0; k < np; ++k) cdf[k] *= invnorm;
        cdf[np - 1] = 1.0;
        double p_hi = pgrid[np - 1];
        double thr = 1.0 - quantile_cap;
        for (int32_t k = 0; k < np; ++k) {
            if (cdf[k] >= thr) { p_hi = pgrid[k]; break; }
        }
        int64_t base = (int64_t)jT * nU;
        int32_t k = 0;
        for (int32_t iu = 0; iu < nU; ++iu) {
            double u = (double)iu / (double)(nU - 1);
            while (k < np - 1 && cdf[k + 1] < u) ++k;
            double c0 = cdf[k], c1 = cdf[k + 1];
            double p0 = pgrid[k], p1 = pgrid[k + 1];
            double frac = (c1 > c0) ? (u - c0) / (c1 - c0) : 0.0;
            frac = frac < 0.0 ? 0.0 : (frac > 1.0 ? 1.0 : frac);
            double pval = p0 + frac * (p1 - p0);
            invcdf[base + iu] = pval < p_hi ? pval : p_hi;
        }
    }
    free(pgrid);
    free(cdf);
    *Tmin_out = Tmin_f;
    *invdT_out = invdT;
    return LMO_OK;
}

/* _juttner_invcdf_lookup, kernels_gpu.jl:88-117 (bilinear in (u, T), both indices clamped) */
double lmo_juttner_invcdf_lookup(const double* invcdf, int32_t nU, int32_t nT,
                                 double Ttab_min, double Ttab_invdT, double u, double T) {
    double uu = u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u);
    double fu = uu * (double)(nU - 1);
    int64_t iu = (int64_t)floor(fu);
    iu = iu < 0 ? 0 : (iu > nU - 2 ? nU - 2 : iu);
    double tu = fu - (double)iu;
    double fT = (T - Ttab_min) * Ttab_invdT;
    double fl = floor(fT);
    int64_t jT;
    if (!(fl >= 0.0)) jT = 0;
    else if (fl > (double)(nT - 2)) jT = nT - 2;
    else jT = (int64_t)fl;
    double tT = fT - (double)jT;
    tT = tT < 0.0 ? 0.0 : (tT > 1.0 ? 1.0 : tT);
    int64_t b0 = jT * nU;
    int64_t b1 = (jT + 1) * nU;
    double v00 = invcdf[b0 + iu];
    double v10 = invcdf[b0 + iu + 1];
    double v01 = invcdf[b1 + iu];
    double v11 = invcdf[b1 + iu + 1];
    double c0 = v00 * (1.0 - tu) + v10 * tu;
    double c1 = v01 * (1.0 - tu) + v11 * tu;
    return c0 * (1.0 - tT) + c1 * tT;
}

/* splitmix64: the private uniform stream of the rejection sampler below */
static inline uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline double u01(uint64_t* s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

/* _draw_juttner_pstar_lib, kernels_cpu.jl:442-458: rejection with the exact-mode envelope.
 * Julia's rand() is not reproducible here; a splitmix64 stream replaces it, so this sampler
 * is compared statistically only. */
int lmo_draw_juttner_rejection(double m, double T, int32_t d, uint64_t seed, int64_t n,
                               double* out) {
    uint64_t s = seed;
    for (int64_t j = 0; j < n; ++j) {
        if (T <= 0.0) { out[j] = 0.0; continue; }
        int a = d - 1;
        double E_peak = (a * T + sqrt((double)(a * a) * T * T + 4.0 * m * m)) / 2.0;
        double p_peak = sqrt(jl_max(E_peak * E_peak - m * m, 1e-12));
        double gpeak = ipow(p_peak, a) * exp(-(E_peak - m) / T);
        double wmax = 1.3 * jl_max(gpeak, 1e-300);
        double pmax = 16.0 * sqrt(m * T) + 16.0 * T;
        double res = p_peak;
        for (int it = 0; it < 100000; ++it) {
            double p = pmax * u01(&s);
            double E = sqrt(m * m + p * p);
            double w = ipow(p, a) * exp(-(E - m) / T);
            if (u01(&s) * wmax <= w) { res = p; break; }
        }
        out[j] = res;
    }
    return LMO_OK;
}

/* threads of the per-particle loops (OpenMP): the passes over particles are independent per
 * particle, so the results do not depend on the thread count */
void lmo_set_threads(int32_t n) { omp_set_num_threads(n > 0 ? n : 1); }
int32_t lmo_get_threads(void) { return (int32_t)omp_get_max_threads(); }

/* ============================== Philox4x32-10 and the noise record ==========================
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11, the Random123 reference generator):
 * multipliers 0xD2511F53 and 0xCD9E8D57, Weyl key increments 0x9E3779B9 and 0xBB67AE85, ten
 * rounds of the round function of the paper's Table 2.                                        */
void lmo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Block b (0 kick and collision, 2 position diffusion, 3 initial-state samplers) of particle i
 * at step s: four 64-bit words from two Philox4x32-10 evaluations (half h = 0 gives w0, w1;
 * h = 1 gives w2, w3; each word is lo | hi << 32 of consecutive outputs) of counter
 *     (i mod 2^32, s mod 2^32, stream mod 2^32, (stream >> 32) mod 2^29 | (2 b + h) << 29)
 * under key (seed mod 2^32, seed >> 32).  The host keeps i, s < 2^32 and stream < 2^61.       */
void lmo_noise_block(uint64_t seed, uint64_t stream, int64_t i, int64_t step, int32_t block,
                     uint64_t w[4]) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int h = 0; h < 2; ++h) {
        const uint32_t ctr[4] = {(uint32_t)(uint64_t)i, (uint32_t)(uint64_t)step, (uint32_t)stream,
                                 ((uint32_t)(stream >> 32) & 0x1FFFFFFFu) | ((uint32_t)(2 * block + h) << 29)};
        uint32_t x[4];
        lmo_philox4x32_10(ctr, key, x);
        w[2 * h] = (uint64_t)x[0] | ((uint64_t)x[1] << 32);
        w[2 * h + 1] = (uint64_t)x[2] | ((uint64_t)x[3] << 32);
    }
}

/* uniforms from the top 52 bits of a word: d = 1.m in [1, 2) (the bits as the mantissa of
 * exponent 0), then d - 1 in [0, 1) or 2 - d in (0, 1], both exact */
static inline double unit_mantissa(uint64_t w) {
    uint64_t b = (w >> 12) | 0x3FF0000000000000ULL;
    double d;
    memcpy(&d, &b, 8);
    return d;
}
static inline double u_co(uint64_t w) { return unit_mantissa(w) - 1.0; }
static inline double u_oc(uint64_t w) { return 2.0 - unit_mantissa(w); }

/* The two transcendental functions of the Box-Muller transform, written with correctly
 * rounded operations only (+, -, *, /, sqrt, fma, floor and exact bit manipulation) so that the
 * same definition evaluates to the same bits on any IEEE-754 machine.
 *
 * noise_log(u), u in [2^-52, 1]: u = 2^e m with m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s),
 * s = (m - 1)/(m + 1), |s| <= 0.1716, by the Taylor series of atanh through s^19 (truncation
 * < 3e-17 relative); ln 2 split into a head whose products with e are exact and a tail.       */
static const double LN2_HI = 6.93147180369123816490e-01;
static const double LN2_LO = 1.90821492927058770002e-10;

static inline double noise_log(double u) {
    uint64_t b;
    memcpy(&b, &u, 8);
    int64_t e = (int64_t)((b >> 52) & 0x7FF) - 1023;
    uint64_t mb = (b & 0x000FFFFFFFFFFFFFULL) | 0x3FF0000000000000ULL;
    double m;
    memcpy(&m, &mb, 8);
    if (m > 1.4142135623730951) { m = 0.5 * m; e += 1; }
    double s = (m - 1.0) / (m + 1.0);
    double z = s * s;
    double p = 1.0 / 19.0;
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    double lm = 2.0 * s + (2.0 * s * z) * p;
    double de = (double)e;
    return fma(de, LN2_HI, fma(de, LN2_LO, lm));
}

/* noise_sincospi(x): sin(pi x) and cos(pi x) for x in [0, 2): x = k/2 + r with k the nearest
 * half-integer count and |r| <= 1/4 (exact), the Taylor polynomials of sin(pi r) through r^15
 * and cos(pi r) through r^16 (truncation below 7e-17 and 2e-18 relative), then the quadrant
 * k mod 4.                                                                                     */
static const double SP[8] = {  /* pi^(2j+1) (-1)^j / (2j+1)!, rounded to double */
    3.141592653589793, -5.16771278004997, 2.5501640398773455, -0.5992645293207921,
    0.08214588661112823, -0.0073704309457143504, 0.00046630280576761255,
    -2.1915353447830217e-05};
static const double CP[9] = {  /* pi^(2j) (-1)^j / (2j)!, rounded to double */
    1.0, -4.934802200544679, 4.0587121264167685, -1.3352627688545895, 0.2353306303588932,
    -0.02580689139001406, 0.0019295743094039231, -0.0001046381049248457, 4.303069587032947e-06};

static inline void noise_sincospi(double x, double* sn, double* cs) {
    double k = floor(2.0 * x + 0.5);
    double r = x - 0.5 * k;
    double z = r * r;
    double ps = SP[7];
    for (int j = 6; j >= 0; --j) ps = fma(ps, z, SP[j]);
    double pc = CP[8];
    for (int j = 7; j >= 0; --j) pc = fma(pc, z, CP[j]);
    double sr = r * ps, cr = pc;
    int q = (int)k & 3;
    double s1 = (q & 1) ? cr : sr;
    double c1 = (q & 1) ? sr : cr;
    *sn = (q & 2) ? -s1 : s1;
    *cs = ((q + 1) & 2) ? -c1 : c1;
}

/* (sin, cos) of a uniform angle from one word: the quadrant q = the top 2 bits, the angle
 * pi (q/2 + r) with r = d/2 - 3/4 in [-1/4, 1/4), d = 1.m from the next 52 bits (exact); the
 * polynomials and quadrant logic of noise_sincospi */
static inline void sincos_word(uint64_t w, double* sn, double* cs) {
    double r = fma(0.5, unit_mantissa(w << 2), -0.75);
    double z = r * r;
    double ps = SP[7];
    for (int j = 6; j >= 0; --j) ps = fma(ps, z, SP[j]);
    double pc = CP[8];
    for (int j = 7; j >= 0; --j) pc = fma(pc, z, CP[j]);
    double sr = r * ps, cr = pc;
    int q = (int)(uint32_t)(w >> 62);
    double s1 = (q & 1) ? cr : sr;
    double c1 = (q & 1) ? sr : cr;
    *sn = (q & 2) ? -s1 : s1;
    *cs = ((q + 1) & 2) ? -c1 : c1;
}

/* Box-Muller: radius from the (0,1] uniform of wa, the uniform angle of wb (sincos_word) */
static inline void box_muller(uint64_t wa, uint64_t wb, double* n0, double* n1) {
    double r = sqrt(-2.0 * noise_log(u_oc(wa)));
    double s, c;
    sincos_word(wb, &s, &c);
    *n0 = r * c;
    *n1 = r * s;
}

/* noise_cospi_half(x) = cos(pi x) for x in [0, 1/2]: Taylor in x^2 through x^20, Horner */
static const double CH[11] = {  /* pi^(2j) (-1)^j / (2j)!, rounded to double */
    1.0, -4.934802200544679, 4.0587121264167685, -1.3352627688545895, 0.2353306303588932,
    -0.02580689139001406, 0.0019295743094039231, -0.0001046381049248457, 4.303069587032947e-06,
    -1.3878952462213771e-07, 3.604730797462501e-09};

static inline double noise_cospi_half(double x) {
    double z = x * x;
    double p = CH[10];
    for (int j = 9; j >= 0; --j) p = fma(p, z, CH[j]);
    return p;
}

/* (sin, cos) of a uniform angle from a 32-bit word x: the quadrant q = the top 2 bits and the
 * angle pi (q/2 + r), r = lo30(x)/2^31 - 1/4 in [-1/4, 1/4) (exact); the polynomials and quadrant
 * logic of noise_sincospi */
static inline void sincos_u32(uint32_t x, double* sn, double* cs) {
    double r = fma((double)(x & 0x3FFFFFFFu), 0x1.0p-31, -0.25);
    double z = r * r;
    double ps = SP[7];
    for (int j = 6; j >= 0; --j) ps = fma(ps, z, SP[j]);
    double pc = CP[8];
    for (int j = 7; j >= 0; --j) pc = fma(pc, z, CP[j]);
    double sr = r * ps, cr = pc;
    int q = (int)(x >> 30);
    double s1 = (q & 1) ? cr : sr;
    double c1 = (q & 1) ? sr : cr;
    *sn = (q & 2) ? -s1 : s1;
    *cs = ((q + 1) & 2) ? -c1 : c1;
}

/* block 0's kick normals: the first three components of a 4-D standard normal vector
 * R (sqrt(v) cos(beta), sqrt(v) sin(beta), sqrt(1 - v) cos(gamma), sqrt(1 - v) sin(gamma)), with
 * R^2 = -2 ln(u1 u2) ~ chi^2(4) (u1 the (0,1] uniform of w0's top 52 bits, u2 = (lo(w1) + 1)/2^32),
 * beta the uniform angle of hi(w1), v = lo(w2)/2^32 uniform on [0, 1) (the squared modulus of the
 * first complex coordinate of a uniform point on the 3-sphere), and cos(gamma) a random sign
 * (bit 31 of hi(w2)) times cos(pi x), x = lo31(hi(w2))/2^32 uniform on [0, 1/2) */
static inline void kick_normals(const uint64_t w[4], double n[3]) {
    double P = u_oc(w[0]) * (((double)(uint32_t)w[1] + 1.0) * 0x1.0p-32);
    double L = -2.0 * noise_log(P);
    double v = (double)(uint32_t)w[2] * 0x1.0p-32;
    double a = sqrt(v * L), b = sqrt((1.0 - v) * L);
    double s, c;
    sincos_u32((uint32_t)(w[1] >> 32), &s, &c);
    n[0] = a * c;
    n[1] = a * s;
    uint32_t hg = (uint32_t)(w[2] >> 32);
    double nb = b * noise_cospi_half((double)(hg & 0x7FFFFFFFu) * 0x1.0p-32);
    n[2] = (hg >> 31) ? -nb : nb;
}

void lmo_noise_transforms(double u, double x, double out[4]) {
    out[0] = noise_log(u);
    noise_sincospi(x, &out[1], &out[2]);
    out[3] = noise_cospi_half(0.25 * x);
}

static int noise_layout(const lmh_params* p, int32_t dims, int32_t pdim,
                        int32_t* off_xi, int32_t* off_rta, int32_t* off_pos) {
    int32_t c = 0;
    *off_xi = *off_rta = *off_pos = -1;
    if (p->coords == LMH_GENERAL_COORDS) {
        *off_xi = 0;
        return 1;
    }
    if (p->collision == LMH_LANGEVIN || p->collision == LMH_LANGEVIN_RTA) { *off_xi = c; c += pdim; }
    if (p->collision == LMH_RTA || p->collision == LMH_LANGEVIN_RTA) { *off_rta = c; c += 2 + pdim; }
    if (p->position_diffusion) { *off_pos = c; c += dims; }
    return c;
}

int32_t lmo_noise_channels(const lmh_params* p, int32_t dims, int32_t pdim) {
    int32_t a, b, c;
    return noise_layout(p, dims, pdim, &a, &b, &c);
}

/* the collision uniform: the low 31 bits of a word over 2^31 */
static inline double u_low31(uint64_t w) { return (double)(uint32_t)(w & 0x7FFFFFFFu) * 0x1.0p-31; }

/* the redraw's quantile: the 32-bit uniform hi(w)/2^32 in [0, 1) */
static inline double u_hi32(uint64_t w) { return (double)(uint32_t)(w >> 32) * 0x1.0p-32; }

/* The PHILOX noise record of one step (include/lmhip.h): block b of particle i at step s is
 * lmo_noise_block(seed, stream, i, s, b).
 *   block 0: the normals n0, n1, n2 of kick_normals -> xi[0..pdim) (GC: xi = the first normal of
 *            the Box-Muller pair of (w0, w1)); u_col = the low 31 bits of w3; u_samp =
 *            hi(w3)/2^32; g = n0..n(pdim-1), the kick normals (rta_collision normalises them)
 *   block 2: (xpos0, xpos1) = BoxMuller(w0, w1), xpos2 = first of BoxMuller(w2, w3)         */
void lmo_noise_record(const lmh_params* p, int32_t dims, int32_t pdim, int64_t N, int64_t step,
                      double* rec) {
    int32_t oxi, orta, opos;
    noise_layout(p, dims, pdim, &oxi, &orta, &opos);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        uint64_t w[4];
        double n[4];
        if (oxi >= 0 || orta >= 0) {
            lmo_noise_block(p->seed, p->stream, i, step, 0, w);
            if (p->coords == LMH_GENERAL_COORDS) box_muller(w[0], w[1], &n[0], &n[1]);
            else kick_normals(w, n);
            if (oxi >= 0) {
                int32_t k = p->coords == LMH_GENERAL_COORDS ? 1 : pdim;
                for (int32_t d = 0; d < k; ++d) rec[(int64_t)(oxi + d) * N + i] = n[d];
            }
            if (orta >= 0) {
                rec[(int64_t)orta * N + i] = u_low31(w[3]);
                rec[(int64_t)(orta + 1) * N + i] = u_hi32(w[3]);
                for (int32_t d = 0; d < pdim; ++d) rec[(int64_t)(orta + 2 + d) * N + i] = n[d];
            }
        }
        if (opos >= 0) {
            lmo_noise_block(p->seed, p->stream, i, step, 2, w);
            box_muller(w[0], w[1], &n[0], &n[1]);
            box_muller(w[2], w[3], &n[2], &n[3]);
            for (int32_t d = 0; d < dims; ++d) rec[(int64_t)(opos + d) * N + i] = n[d];
        }
    }
}

/* ===================== initial-state samplers (utils.jl:33-172, 499-558) ====================== */
/* Quark k's draws are the words of noise block 3 under key (seed, stream), counter word 1 selecting
 * the draw: position attempts 0, 1, ..., the momentum 0xFFFFFFFF, thermal p_z attempts
 * 0x80000000 + a.  The inverse CDFs follow Interpolations.LinearInterpolation(cdf, values) with
 * flat extrapolation, as utils.jl:68,141 build them.                                             */
static int64_t sortedlast(const double* v, int64_t n, double x) {   /* Base.searchsortedlast */
    int64_t lo = 0, hi = n + 1;
    while (lo < hi - 1) {
        int64_t m = lo + ((hi - lo) >> 1);
        if (x < v[m - 1]) hi = m; else lo = m;
    }
    return lo;
}

static double inverse_cdf(const double* cdf, const double* vals, int64_t n, double u) {
    if (!(u > cdf[0])) return vals[0];
    if (!(u < cdf[n - 1])) return vals[n - 1];
    int64_t k = sortedlast(cdf, n, u);
    double t = (u - cdf[k - 1]) / (cdf[k] - cdf[k - 1]);
    return (1.0 - t) * vals[k - 1] + t * vals[k];
}

static double linear_or_zero(const double* x, const double* f, int64_t n, double q) {
    if (!(q >= x[0] && q <= x[n - 1])) return 0.0;
    int64_t k = sortedlast(x, n, q);
    if (k > n - 1) k = n - 1;
    double t = (q - x[k - 1]) / (x[k] - x[k - 1]);
    return (1.0 - t) * f[k - 1] + t * f[k];
}

static inline double clampd(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }

int lmo_sample_fonll(const double* rg, const double* pg, const double* cdf_r, const double* n_r,
                     const double* cdf_p, const double* p_ok, int32_t nr, int32_t np, int32_t cartesian,
                     double rmax, double fmax, int64_t N, int32_t antithetic, uint64_t seed,
                     uint64_t stream, double* x_out, double* p_out) {
    const int64_t nhalf = antithetic ? N / 2 : 0;
    int failed = 0;
#pragma omp parallel for schedule(static) reduction(| : failed)
    for (int64_t i = 0; i < N; ++i) {
        int64_t k = i;
        double sgn = 1.0, x = 0.0, y = 0.0, r = 0.0, s, c;
        uint64_t w[4];
        if (antithetic && i < 2 * nhalf) {
            k = i / 2;
            sgn = (i % 2) ? -1.0 : 1.0;
        } else if (antithetic) {
            k = nhalf;
        }
        if (!cartesian) {
            /* utils.jl:160-166 */
            lmo_noise_block(seed, stream, k, 0, 3, w);
            r = clampd(inverse_cdf(cdf_r, rg, nr, u_co(w[0])), rg[0], rg[nr - 1]);
            noise_sincospi(2.0 * u_co(w[1]), &s, &c);
            x = r * c;
            y = r * s;
        } else {
            /* utils.jl:98-127 */
            int done = 0;
            for (int32_t a = 0; a < 100000 && !done; ++a) {
                lmo_noise_block(seed, stream, k, a, 3, w);
                double xa = rmax * (2.0 * u_co(w[0]) - 1.0);
                double ya = rmax * (2.0 * u_co(w[1]) - 1.0);
                double ra = sqrt(xa * xa + ya * ya);
                if (ra > rmax) continue;
                if (u_co(w[2]) * fmax > linear_or_zero(rg, n_r, nr, ra)) continue;
                x = xa;
                y = ya;
                r = ra;
                done = 1;
            }
            if (!done) failed |= 1;
        }
        /* utils.jl:119-125, 168-172: |p| from the p|r CDF of the r node, azimuth uniform */
        lmo_noise_block(seed, stream, k, 0xFFFFFFFFLL, 3, w);
        double rc = clampd(r, rg[0], rg[nr - 1]);
        int64_t j = sortedlast(rg, nr, rc);
        j = j < 1 ? 1 : (j > nr ? nr : j);
        double pm = p_ok[j - 1] != 0.0 ? inverse_cdf(cdf_p + (j - 1) * (int64_t)np, pg, np, u_co(w[0])) : 0.0;
        noise_sincospi(2.0 * u_co(w[1]), &s, &c);
        x_out[2 * i] = x;
        x_out[2 * i + 1] = y;
        p_out[2 * i] = sgn * (pm * c);
        p_out[2 * i + 1] = sgn * (pm * s);
    }
    return failed ? LMO_E_INVALID : LMO_OK;
}

/* sample_pz_conditional_juttner, utils.jl:499-530 */
static double thermal_pz(uint64_t seed, uint64_t stream, int64_t k, double mT, double T) {
    uint64_t w[4];
    if (!(T > 0.0 && mT > 0.0)) return 0.0;
    double a = mT / T;
    if (a > 1.05) {
        double c = fmax(2.0, a / (a - 1.0));
        double sig = sqrt(c / a);
        for (int it = 0; it < 100000; ++it) {
            double n0, n1;
            lmo_noise_block(seed, stream, k, (int64_t)(0x80000000u + (uint32_t)it), 3, w);
            box_muller(w[0], w[1], &n0, &n1);
            double y = sig * n0;
            double ch = cosh(y);
            double acc = ch * exp(-a * (ch - 1.0) + a * y * y / (2.0 * c));
            if (u_co(w[2]) <= acc) return mT * sinh(y);
        }
        return 0.0;
    }
    double b = sqrt(2.0) * T;
    for (int it = 0; it < 1000000; ++it) {
        lmo_noise_block(seed, stream, k, (int64_t)(0x80000000u + (uint32_t)it), 3, w);
        double pz = -b * noise_log(u_oc(w[0])) * (u_co(w[1]) < 0.5 ? -1.0 : 1.0);
        double acc = exp(-(sqrt(mT * mT + pz * pz) - (mT + fabs(pz)) / sqrt(2.0)) / T);
        if (u_co(w[2]) <= acc) return pz;
    }
    return 0.0;
}

/* append_thermal_pz, utils.jl:533-558, with T_of_r the radial medium at t0 */
void lmo_append_thermal_pz(const double* mom, const double* pos, int32_t dims, int64_t N, double m,
                           double t0, int32_t antithetic, uint64_t seed, uint64_t stream,
                           const lmo_radial_medium* med, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        int mirror = antithetic && (i % 2);
        int64_t k = mirror ? i - 1 : i;
        double r2 = 0.0;
        for (int32_t d = 0; d < dims; ++d) r2 += pos[(int64_t)dims * k + d] * pos[(int64_t)dims * k + d];
        double T = lmo_interpolate_2d(med->xgrid, med->nx, med->tgrid, med->nt, med->T, sqrt(r2), t0);
        T = 0.0 > T ? 0.0 : T;                       /* Julia max(T, 0.0) */
        double px = mom[2 * k], py = mom[2 * k + 1];
        double mT = sqrt(m * m + px * px + py * py);
        double pz = thermal_pz(seed, stream, k, mT, T);
        out[3 * i] = mom[2 * i];
        out[3 * i + 1] = mom[2 * i + 1];
        out[3 * i + 2] = mirror ? -pz : pz;
    }
}

/* ===================== Cartesian step on the radial (r, t) medium =========================== */

static inline double interp_med(const lmo_radial_medium* med, const double* f, double r, double t) {
    return lmo_interpolate_2d(med->xgrid, med->nx, med->tgrid, med->nt, f, r, t);
}

/* kernel_boost_to_rest_frame_cpu!, kernels_cpu.jl:53-148 */
void lmo_kernel_boost_to_rest(double* mom, const double* pos, const lmo_radial_medium* med,
                              double m, int64_t N, int64_t step, double dt, double t0,
                              int32_t dims, int32_t pdim, int32_t radial_mode, double psi2,
                              int32_t relativistic) {
    for (int64_t i = 0; i < N; ++i) {
        double t_now = (double)step * dt + t0;
        double r2 = 0.0;
        for (int32_t d = 0; d < dims; ++d) r2 += pos[d + dims * i] * pos[d + dims * i];
        double r = sqrt(r2);
        if (r < JL_EPS) continue;
        double v = interp_med(med, med->v, r, t_now);
        if (med->v2 && dims >= 2) {
            double v2m = interp_med(med, med->v2, r, t_now);
            double phi = atan2(pos[1 + dims * i], pos[0 + dims * i]);
            v *= (1.0 + 2.0 * v2m * cos(2.0 * (phi - psi2)));
        }
        if (!relativistic) {
            if (radial_mode) {
                mom[0 + pdim * i] -= m * v;
            } else {
                double x = pos[0 + dims * i], y = pos[1 + dims * i];
                mom[0 + pdim * i] -= m * v * (x / r);
                mom[1 + pdim * i] -= m * v * (y / r);
            }
            continue;
        }
        double vmax = sqrt(1.0 - 1e-12);
        v = jl_clamp(v, -vmax, vmax);
        double v2 = v * v;
        double gam = 1.0 / sqrt(1.0 - v2 + 1e-10);
        double p2 = 0.0;
        for (int32_t d = 0; d < pdim; ++d) p2 += mom[d + pdim * i] * mom[d + pdim * i];
        double E = sqrt(p2 + m * m);
        if (radial_mode) {
            mom[0 + pdim * i] = gam * (mom[0 + pdim * i] - v * E);
        } else {
            double x = pos[0 + dims * i], y = pos[1 + dims * i];
            double rhatx = x / r, rhaty = y / r;
            double px = mom[0 + pdim * i], py = mom[1 + pdim * i];
            double p_par = px * rhatx + py * rhaty;
            double p_perp_x = px - p_par * rhatx;
            double p_perp_y = py - p_par * rhaty;
            double p_par_new = gam * (p_par - v * E);
            mom[0 + pdim * i] = p_par_new * rhatx + p_perp_x;
            mom[1 + pdim * i] = p_par_new * rhaty + p_perp_y;
        }
    }
}

/* kernel_boost_to_lab_frame_cpu!, kernels_cpu.jl:159-244 */
void lmo_kernel_boost_to_lab(double* mom, const double* pos, const lmo_radial_medium* med,
                             double m, int64_t N, int64_t step, double dt, double t0,
                             int32_t dims, int32_t pdim, int32_t radial_mode, double psi2,
                             int32_t relativistic) {
    for (int64_t i = 0; i < N; ++i) {
        double r2 = 0.0;
        for (int32_t d = 0; d < dims; ++d) r2 += pos[d + dims * i] * pos[d + dims * i];
        double r = sqrt(r2);
        if (r < JL_EPS) continue;
        double t_now = (double)step * dt + t0;
        double v = interp_med(med, med->v, r, t_now);
        if (med->v2 && dims >= 2) {
            double v2m = interp_med(med, med->v2, r, t_now);
            double phi = atan2(pos[1 + dims * i], pos[0 + dims * i]);
            v *= (1.0 + 2.0 * v2m * cos(2.0 * (phi - psi2)));
        }
        if (!relativistic) {
            if (radial_mode) {
                mom[0 + pdim * i] += m * v;
            } else {
                double x = pos[0 + dims * i], y = pos[1 + dims * i];
                mom[0 + pdim * i] += m * v * (x / r);
                mom[1 + pdim * i] += m * v * (y / r);
            }
            continue;
        }
        double vmax = sqrt(1.0 - 1e-12);
        v = jl_clamp(v, -vmax, vmax);
        double v2 = v * v;
        double gam = 1.0 / sqrt(1.0 - v2 + 1e-10);
        double p2 = 0.0;
        for (int32_t d = 0; d < pdim; ++d) p2 += mom[d + pdim * i] * mom[d + pdim * i];
        double E = sqrt(p2 + m * m);
        if (radial_mode) {
            mom[0 + pdim * i] = gam * (mom[0 + pdim * i] + v * E);
        } else {
            double x = pos[0 + dims * i], y = pos[1 + dims * i];
            double rhatx = x / r, rhaty = y / r;
            double px = mom[0 + pdim * i], py = mom[1 + pdim * i];
            double p_par = px * rhatx + py * rhaty;
            double p_perp_x = px - p_par * rhatx;
            double p_perp_y = py - p_par * rhaty;
            double p_par_new = gam * (p_par + v * E);
            mom[0 + pdim * i] = p_par_new * rhatx + p_perp_x;
            mom[1 + pdim * i] = p_par_new * rhaty + p_perp_y;
        }
    }
}

/* kernel_bjorken_redshift_cpu!, kernels_cpu.jl:427-435 (p_z is row 3) */
static void kernel_bjorken(double* mom, int32_t pdim, int64_t step, double dt, double t0,
                           int64_t N) {
    double tau_a = t0 + (double)(step - 1) * dt;
    if (!(tau_a > 0.0)) return;
    double f = tau_a / (tau_a + dt);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) mom[2 + pdim * i] *= f;
}

/*
 * The drag/noise coefficients of one particle in the LRF and its momentum update.
 *   Reference (PREPOINT, TABLE): kernel_compute_all_forces_cpu!, kernels_cpu.jl:326-352
 *   (exact OU with eta_eff = eta_D*m/E when relativistic), then
 *   kernel_update_momenta_LRF_cpu!, kernels_cpu.jl:406-417: p += det + sqrt(dt)*sto.
 *   MIDPOINT: kernels_gpu.jl:366-381 (noise-free half step, eta_eff re-evaluated there).
 *   POSTPOINT (extension): eta_eff evaluated at the updated momentum by three fixed-point
 *   sweeps q_k = p + ((a_k - 1) p + sqrt(dt) noise_pref_k xi), q_0 = p, eta_k = eta(q_{k-1}).
 *   GAMMA_P (extension): eta(q) = eta_D / (1 + |q|^2/Lambda^2) with eta_D = gamma0*T^2 and
 *   kappa = 2 m T eta_D; the m/E factor is not applied.
 * ηD and kT are the per-particle coefficients of kernels_cpu.jl:281-290.                     */
static inline double eta_of(const lmh_params* p, double etaD, double m, const double* q,
                            int32_t pdim) {
    double q2 = 0.0;
    for (int32_t d = 0; d < pdim; ++d) q2 += q[d] * q[d];
    if (p->drag_model == LMH_DRAG_GAMMA_P) {
        return etaD / (1.0 + q2 / (p->Lambda * p->Lambda));
    }
    double E = sqrt(q2 + m * m);
    return p->relativistic ? etaD * m / E : etaD;
}

static void ou_forces(const lmh_params* p, double etaD, double kT, double m, const double* q,
                      const double* xi, int32_t pdim, double* det, double* sto) {
    double dt = p->dt;
    if (etaD > 0.0) {
        double eta_eff = eta_of(p, etaD, m, q, pdim);
        if (p->integrator == LMH_MIDPOINT && dt > 0.0) {
            double a_half = exp(-eta_eff * (0.5 * dt));
            double qm[3];
            for (int32_t d = 0; d < pdim; ++d) qm[d] = a_half * q[d];
            eta_eff = eta_of(p, etaD, m, qm, pdim);
        }
        double a = exp(-eta_eff * dt);
        double noise_pref = (dt > 0.0 && eta_eff > 0.0)
                                ? (kT * sqrt((1.0 - a * a) / (2.0 * eta_eff * dt))) : 0.0;
        for (int32_t d = 0; d < pdim; ++d) {
            det[d] = (a - 1.0) * q[d];
            sto[d] = noise_pref * xi[d];
        }
    } else {
        for (int32_t d = 0; d < pdim; ++d) {
            det[d] = 0.0;
            sto[d] = kT * xi[d];
        }
    }
}

/* coefficients eta_D, kT from the local T (kernels_cpu.jl:263-290) or the GAMMA_P model */
static void drag_coeffs(const lmh_params* p, const lmo_table* tab, double m, double T,
                        double* etaD, double* kT) {
    if (p->drag_model == LMH_DRAG_GAMMA_P) {
        double e = p->gamma0 * T * T;
        double kappa = 2.0 * m * T * e;
        *etaD = e;
        *kT = sqrt(kappa);
        return;
    }
    double DsT_safe = p->DsT <= 0.0 ? 0.0 : p->DsT;
    double tau = DsT_safe == 0.0 ? 0.0 : lmo_eval_time_table(T, tab->Tmin, tab->invdT, tab->vals, tab->n);
    double etaD_ = (tau > 0.0 && isfinite(tau)) ? (1.0 / tau) : 0.0;
    double kappa = (tau > 0.0 && isfinite(tau)) ? ((2.0 * m * T) / tau) : 0.0;
    *etaD = etaD_;
    *kT = sqrt(kappa);
}

/* kernel_compute_all_forces_cpu! + kernel_update_momenta_LRF_cpu! over all particles, with
 * T[i] the clamped local temperature of each particle.  Radial mode (kernels_cpu.jl:293-323):
 * Euler drag det = -eta_eff p dt and sto = kT xi (the (kL-kT) projector term is exactly zero
 * because kL == kT, :288-289). */
static void forces_and_update(const lmh_params* p, const lmo_table* tab, double m,
                              const double* Tloc, double* mom, const double* xi /* (pdim,N) */,
                              int32_t pdim, int32_t radial_mode, int64_t N,
                              double* det /* (pdim,N) */, double* sto) {
    double dt = p->dt;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        double T = Tloc[i];
        double etaD, kT;
        drag_coeffs(p, tab, m, T, &etaD, &kT);
        double* q = mom + (int64_t)pdim * i;
        const double* x = xi + (int64_t)pdim * i;
        double* de = det + (int64_t)pdim * i;
        double* st = sto + (int64_t)pdim * i;
        if (radial_mode) {
            double s = 0.0;
            for (int32_t d = 0; d < pdim; ++d) s += q[d] * q[d];
            double pm = sqrt(s);
            double E_r = sqrt(pm * pm + m * m);
            double eta_r = p->relativistic ? etaD * m / E_r : etaD;
            for (int32_t d = 0; d < pdim; ++d) {
                de[d] = -eta_r * q[d] * dt;
                st[d] = kT * 1.0 * x[d];
            }
        } else if (p->integrator == LMH_POSTPOINT && etaD > 0.0 && dt > 0.0) {
            double qk[3], dk[3], sk[3];
            for (int32_t d = 0; d < pdim; ++d) qk[d] = q[d];
            double sqrt_dt = sqrt(dt);
            for (int it = 0; it < 3; ++it) {
                double eta_eff = eta_of(p, etaD, m, qk, pdim);
                double a = exp(-eta_eff * dt);
                double noise_pref = eta_eff > 0.0
                                        ? (kT * sqrt((1.0 - a * a) / (2.0 * eta_eff * dt))) : 0.0;
                for (int32_t d = 0; d < pdim; ++d) {
                    dk[d] = (a - 1.0) * q[d];
                    sk[d] = noise_pref * x[d];
                }
                for (int32_t d = 0; d < pdim; ++d) qk[d] = q[d] + (dk[d] + sqrt_dt * sk[d]);
            }
            for (int32_t d = 0; d < pdim; ++d) { de[d] = dk[d]; st[d] = sk[d]; }
        } else {
            ou_forces(p, etaD, kT, m, q, x, pdim, de, st);
        }
    }
    double sqrt_dt = sqrt(dt);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i)
        for (int32_t d = 0; d < pdim; ++d) {
            int64_t k = d + (int64_t)pdim * i;
            mom[k] += det[k] + sqrt_dt * sto[k];
        }
}

/* kernel_rta_collision_cpu!, kernels_cpu.jl:474-508, with the redraw of
 * kernel_rta_collision_gpu!, kernels_gpu.jl:653-704: |p*| from the inverse-CDF table at the
 * supplied quantile, direction from the supplied Gaussians (radial mode: the sign of g0). */
static void rta_collision(const lmh_params* p, const lmo_table* taun, const lmo_jtable* jt,
                          const double* Tloc, double* mom, int32_t pdim, int32_t radial_mode,
                          int64_t N, const double* ucol, const double* usamp,
                          const double* g /* (pdim,N) */) {
    double dt = p->dt;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        double T = Tloc[i];
        double tau_n = (p->DsT > 0.0 && T > 0.0)
                           ? lmo_eval_time_table(T, taun->Tmin, taun->invdT, taun->vals, taun->n)
                           : 0.0;
        double Pcol = (tau_n > 0.0 && isfinite(tau_n)) ? jl_clamp(dt / tau_n, 0.0, 1.0) : 1.0;
        if (!(ucol[i] < Pcol)) continue;
        double pstar = lmo_juttner_invcdf_lookup(jt->vals, jt->nU, jt->nT, jt->Tmin, jt->invdT,
                                                 usamp[i], T);
        if (radial_mode) {
            mom[0 + pdim * i] = g[0 + pdim * i] < 0.0 ? -pstar : pstar;
        } else {
            double nrm2 = 0.0;
            for (int32_t d = 0; d < pdim; ++d) nrm2 += g[d + pdim * i] * g[d + pdim * i];
            double inv = pstar / sqrt(jl_max(nrm2, 1e-300));
            for (int32_t d = 0; d < pdim; ++d) mom[d + pdim * i] = g[d + pdim * i] * inv;
        }
    }
}

/* kernel_set_to_fluid_velocity_cpu!, kernels_cpu.jl:671-735 */
static void set_to_fluid(double* mom, const double* pos, const lmo_radial_medium* med, double m,
                         int64_t N, int64_t step, double dt, double t0, int32_t dims,
                         int32_t pdim, int32_t radial_mode, int32_t relativistic) {
    for (int64_t i = 0; i < N; ++i) {
        double r, x = 0.0, y = 0.0;
        if (radial_mode) {
            r = pos[0 + dims * i];
        } else {
            x = pos[0 + dims * i];
            y = pos[1 + dims * i];
            r = sqrt(x * x + y * y);
        }
        if (r < JL_EPS) continue;
        double tau_now = (double)step * dt + t0;
        double v = interp_med(med, med->v, r, tau_now);
        if (!relativistic) {
            if (radial_mode) {
                mom[0 + pdim * i] = m * v;
            } else {
                mom[0 + pdim * i] = m * v * (x / r);
                mom[1 + pdim * i] = m * v * (y / r);
            }
            continue;
        }
        double vmax = sqrt(1.0 - 1e-12);
        v = jl_clamp(v, -vmax, vmax);
        double gam = 1.0 / sqrt(1.0 - v * v + 1e-10);
        if (radial_mode) {
            mom[0 + pdim * i] = m * gam * v;
        } else {
            mom[0 + pdim * i] = m * gam * v * (x / r);
            mom[1 + pdim * i] = m * gam * v * (y / r);
        }
    }
}

/* kernel_update_positions_cpu!, kernels_cpu.jl:510-637, with the position-diffusion randn()
 * draws (:562, :604) supplied as xpos (dims, N). */
static void update_positions(const lmh_params* p, const lmo_radial_medium* med, double* pos,
                             double* mom, double m, int64_t N, int64_t step, int32_t dims,
                             int32_t pdim, int32_t radial_mode, const double* xpos) {
    double dt = p->dt, t0 = p->t0;
    const double* xg = med->xgrid;
    double r_axis_eps = med->nx >= 2 ? jl_max(1e-12, 0.5 * fabs(xg[1] - xg[0])) : 1e-6;
    double r_max = med->nx >= 1 ? xg[med->nx - 1] : 0.0;
    for (int64_t i = 0; i < N; ++i) {
        double E2 = m * m;
        for (int32_t d = 0; d < pdim; ++d) E2 += mom[d + pdim * i] * mom[d + pdim * i];
        double E = p->relativistic ? sqrt(E2) : m;
        if (radial_mode) {
            double r = pos[0 + dims * i];
            double pr = mom[0 + pdim * i];
            double r_abs = fabs(r);
            double r_safe = (r_abs < r_axis_eps) ? r_axis_eps : r_abs;
            double dr = (pr / E) * dt;
            if (p->position_diffusion) {
                double T = interp_med(med, med->T, r_safe, (double)step * dt + t0);
                double D = (p->DsT / jl_max(T, JL_EPS)) / FMGEV;
                if (D > 0.0) {
                    double xi = xpos[0 + dims * i];
                    dr += (D / r_safe) * dt + sqrt(2.0 * D * dt) * xi;
                }
            }
            double r_new = r + dr;
            if (r_new < 0.0) {
                r_new = -r_new;
                mom[0 + pdim * i] = -mom[0 + pdim * i];
            }
            if (p->reflecting_boundary && r_max > 0.0 && r_new > r_max) {
                r_new = 2.0 * r_max - r_new;
                mom[0 + pdim * i] = -mom[0 + pdim * i];
                if (r_new < 0.0) {
                    r_new = -r_new;
                    mom[0 + pdim * i] = -mom[0 + pdim * i];
                }
            }
            pos[0 + dims * i] = r_new;
        } else {
            for (int32_t d = 0; d < dims; ++d)
                pos[d + dims * i] += dt * mom[d + pdim * i] / E;
            if (p->position_diffusion) {
                double r2 = 0.0;
                for (int32_t d = 0; d < dims; ++d) r2 += pos[d + dims * i] * pos[d + dims * i];
                double r_now = sqrt(r2);
                double T = interp_med(med, med->T, r_now, (double)step * dt + t0);
                double D = (p->DsT / jl_max(T, JL_EPS)) / FMGEV;
                if (D > 0.0) {
                    double sigma = sqrt(2.0 * D * dt);
                    for (int32_t d = 0; d < dims; ++d)
                        pos[d + dims * i] += sigma * xpos[d + dims * i];
                }
            }
            if (p->reflecting_boundary && r_max > 0.0) {
                double r2 = 0.0;
                for (int32_t d = 0; d < dims; ++d) r2 += pos[d + dims * i] * pos[d + dims * i];
                double r_now = sqrt(r2);
                if (r_now > r_max) {
                    double invr = 1.0 / jl_max(r_now, r_axis_eps);
                    double ppar = 0.0;
                    for (int32_t d = 0; d < dims; ++d)
                        ppar += mom[d + pdim * i] * (pos[d + dims * i] * invr);
                    for (int32_t d = 0; d < dims; ++d)
                        mom[d + pdim * i] -= 2.0 * ppar * (pos[d + dims * i] * invr);
                    double scale = (2.0 * r_max - r_now) * invr;
                    for (int32_t d = 0; d < dims; ++d) pos[d + dims * i] *= scale;
                }
            }
        }
    }
}

/* scratch of one segment */
typedef struct {
    double *xi, *det, *sto, *T, *ucol, *usamp, *g, *xpos, *aux;
} work_t;

static int work_alloc(work_t* w, int64_t N, int32_t pdim, int32_t dims) {
    size_t n = (size_t)N;
    memset(w, 0, sizeof(*w));
    w->xi = (double*)calloc(n * (size_t)pdim, sizeof(double));
    w->det = (double*)calloc(n * (size_t)pdim, sizeof(double));
    w->sto = (double*)calloc(n * (size_t)pdim, sizeof(double));
    w->T = (double*)calloc(n, sizeof(double));
    w->ucol = (double*)calloc(n, sizeof(double));
    w->usamp = (double*)calloc(n, sizeof(double));
    w->g = (double*)calloc(n * (size_t)pdim, sizeof(double));
    w->xpos = (double*)calloc(n * (size_t)dims, sizeof(double));
    w->aux = (double*)calloc(n * 3, sizeof(double));
    return (w->xi && w->det && w->sto && w->T && w->ucol && w->usamp && w->g && w->xpos && w->aux)
               ? LMO_OK : LMO_E_INVALID;
}
static void work_free(work_t* w) {
    free(w->xi); free(w->det); free(w->sto); free(w->T); free(w->ucol); free(w->usamp);
    free(w->g); free(w->xpos); free(w->aux);
}

/* channels of the record (row-major (C, Ntot), segment starting at i0) into Julia-layout
 * (rows, n) scratch: the driver's `ξ .= randn(pdim, N)` (simulate_cpu.jl:216) and the draws of
 * the RTA and diffusion kernels, read from the fixed-noise record instead. */
static void gather_channels(const double* rec, int64_t Ntot, int64_t i0, int64_t n, int32_t c0,
                            int32_t rows, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        for (int32_t d = 0; d < rows; ++d)
            out[d + (int64_t)rows * i] = rec[(int64_t)(c0 + d) * Ntot + i0 + i];
}

static int all_finite(const double* a, int64_t n) {
    int ok = 1;
#pragma omp parallel for schedule(static) reduction(&& : ok)
    for (int64_t k = 0; k < n; ++k) ok = ok && isfinite(a[k]);
    return ok;
}

static int64_t species_begin(const lmo_species* sp, int32_t k) {
    int64_t b = 0;
    for (int32_t j = 0; j < k; ++j) b += sp->counts[j];
    return b;
}

/* One step of simulate_cpu.jl:215-290 on one species segment (pos/mom point at its start). */
static void step_cartesian_segment(const lmh_params* p, const lmo_radial_medium* med,
                                   const lmo_species* sp, int32_t k, int32_t dims,
                                   int32_t pdim, double* pos, double* mom, int64_t n,
                                   int64_t step, const double* rec, int64_t Ntot, int64_t i0,
                                   work_t* w) {
    int32_t oxi, orta, opos;
    noise_layout(p, dims, pdim, &oxi, &orta, &opos);
    double m = sp->mass[k];
    int32_t radial_mode = dims == 1;
    double dt = p->dt, t0 = p->t0;
    if (oxi >= 0) gather_channels(rec, Ntot, i0, n, oxi, pdim, w->xi);
    if (orta >= 0) {
        gather_channels(rec, Ntot, i0, n, orta, 1, w->ucol);
        gather_channels(rec, Ntot, i0, n, orta + 1, 1, w->usamp);
        gather_channels(rec, Ntot, i0, n, orta + 2, pdim, w->g);
    }
    if (opos >= 0) gather_channels(rec, Ntot, i0, n, opos, dims, w->xpos);

    /* 1. boost to the LRF (simulate_cpu.jl:219-222) */
    lmo_kernel_boost_to_rest(mom, pos, med, m, n, step, dt, t0, dims, pdim, radial_mode,
                             p->psi2, p->relativistic);
    /* 1b. Bjorken redshift (simulate_cpu.jl:225-227) */
    if (p->bjorken_redshift) kernel_bjorken(mom, pdim, step, dt, t0, n);

    if (p->collision == LMH_FLUID) {
        /* simulate_cpu.jl:229-233 */
        set_to_fluid(mom, pos, med, m, n, step, dt, t0, dims, pdim, radial_mode, p->relativistic);
    } else {
        /* local temperature, clamped at 0 (kernels_cpu.jl:264-267, :486-487) */
        for (int64_t i = 0; i < n; ++i) {
            double r2 = 0.0;
            for (int32_t d = 0; d < dims; ++d) r2 += pos[d + dims * i] * pos[d + dims * i];
            double r = sqrt(r2);
            double T = interp_med(med, med->T, r, (double)step * dt + t0);
            w->T[i] = jl_max(T, 0.0);
        }
        if (p->collision == LMH_LANGEVIN || p->collision == LMH_LANGEVIN_RTA) {
            /* 2.-3. forces and update (simulate_cpu.jl:255-271) */
            forces_and_update(p, &sp->drag[k], m, w->T, mom, w->xi, pdim, radial_mode, n,
                              w->det, w->sto);
        }
        if (p->collision == LMH_RTA || p->collision == LMH_LANGEVIN_RTA) {
            /* simulate_cpu.jl:240-245 */
            rta_collision(p, &sp->rta[k], &sp->jt[k], w->T, mom, pdim, radial_mode, n, w->ucol,
                          w->usamp, w->g);
        }
        /* 4. boost back to the lab (simulate_cpu.jl:248-251, :274-277) */
        lmo_kernel_boost_to_lab(mom, pos, med, m, n, step, dt, t0, dims, pdim, radial_mode,
                                p->psi2, p->relativistic);
    }
    /* 5. positions (simulate_cpu.jl:281-290) */
    update_positions(p, med, pos, mom, m, n, step, dims, pdim, radial_mode, w->xpos);
}

static int check_cartesian_dims(int32_t dims, int32_t pdim) {
    if (dims == pdim && dims >= 1 && dims <= 3) return 1;
    return dims == 2 && pdim == 3;
}

/* ---- Milne coordinates (extension): position rows (x, y, eta_s), momentum rows (p_x, p_y, p_z)
 * with p_z the longitudinal momentum in the frame comoving with the Bjorken flow at eta_s.  The
 * step is the reference's Bjorken-mode Cartesian step on (x, y, p) (dimensions 2, momentum rows
 * 3, bjorken_redshift: simulate_cpu.jl:215-290, kernels_cpu.jl:427-435), the medium sampled at
 * the transverse radius, followed by the push of eta_s. */

/* copies of rows 0 and 1 of a (3, N) Milne position matrix into (2, N) and back */
static void milne_split(const double* pos, double* xy, int64_t N) {
    for (int64_t i = 0; i < N; ++i) { xy[2 * i] = pos[3 * i]; xy[2 * i + 1] = pos[3 * i + 1]; }
}
static void milne_join(const double* xy, double* pos, int64_t N) {
    for (int64_t i = 0; i < N; ++i) { pos[3 * i] = xy[2 * i]; pos[3 * i + 1] = xy[2 * i + 1]; }
}

/* The space-time rapidity over the step tau_a -> tau_b = tau_a + dt (tau_a = t0 + (step - 1) dt,
 * the interval of the Bjorken redshift) of a quark whose comoving longitudinal momentum at tau_b
 * is p_z: along a free path p_z tau is conserved and y - eta_s = asinh(p_z/m_T) (the note at
 * kernels_cpu.jl:419-426: dη/dτ = tanh(y−η)/τ), so eta_s advances by asinh(k b) - asinh(b) with
 * b = p_z/m_T and k = tau_b/tau_a, written without cancellation as
 *     asinh(b (k - 1)(k + 1) / (k sqrt(1 + b^2) + sqrt(1 + (k b)^2))),   k - 1 = dt/tau_a.    */
double lmo_milne_deta(double px, double py, double pz, double m, double tau_a, double dt) {
    if (!(tau_a > 0.0)) return 0.0;
    double mT = sqrt(m * m + px * px + py * py);
    double b = pz / mT;
    double km1 = dt / tau_a;
    double k = 1.0 + km1;
    double kb = k * b;
    double num = b * (km1 * (k + 1.0));
    double den = k * sqrt(1.0 + b * b) + sqrt(1.0 + kb * kb);
    return asinh(num / den);
}

/* The driver's initial boost at step 0, simulate_cpu.jl:160-163 */
int lmo_boost_to_lab_all(const lmh_params* p, const lmo_radial_medium* med,
                         const lmo_species* sp, int32_t dims, int32_t pdim,
                         double* pos, double* mom, int64_t N, int64_t step) {
    if (p->coords == LMH_MILNE) {
        if (dims != 3 || pdim != 3) return LMO_E_INVALID;
        double* xy = (double*)malloc((size_t)(2 * N) * sizeof(double));
        if (!xy) return LMO_E_INVALID;
        milne_split(pos, xy, N);
        lmh_params q = *p;
        q.coords = LMH_CARTESIAN;
        int rc = lmo_boost_to_lab_all(&q, med, sp, 2, 3, xy, mom, N, step);
        free(xy);
        return rc;
    }
    if (!check_cartesian_dims(dims, pdim)) return LMO_E_INVALID;
    if (species_begin(sp, sp->nspecies) != N) return LMO_E_INVALID;
    for (int32_t k = 0; k < sp->nspecies; ++k) {
        int64_t b = species_begin(sp, k);
        lmo_kernel_boost_to_lab(mom + (int64_t)pdim * b, pos + (int64_t)dims * b, med,
                                sp->mass[k], sp->counts[k], step, p->dt, p->t0, dims, pdim,
                                dims == 1, p->psi2, p->relativistic);
    }
    return LMO_OK;
}

/* Steps step_begin .. step_begin+nsteps-1 of simulate_cpu.jl:215-303; noise is the fixed-noise
 * record (nsteps, C, N).  The NaN/Inf scan of :293-303 stops the run at the first bad step. */
static int run_milne(const lmh_params* p, const lmo_radial_medium* med, const lmo_species* sp,
                     double* pos, double* mom, int64_t N, int64_t step_begin, int64_t nsteps,
                     const double* noise, int64_t* bad_step);

int lmo_run_cartesian(const lmh_params* p, const lmo_radial_medium* med, const lmo_species* sp,
                      int32_t dims, int32_t pdim, double* pos, double* mom, int64_t N,
                      int64_t step_begin, int64_t nsteps, const double* noise,
                      int64_t* bad_step) {
    if (p->coords == LMH_MILNE) {
        if (dims != 3 || pdim != 3 || p->position_diffusion || p->reflecting_boundary || !(p->t0 > 0.0))
            return LMO_E_INVALID;
        return run_milne(p, med, sp, pos, mom, N, step_begin, nsteps, noise, bad_step);
    }
    if (!check_cartesian_dims(dims, pdim)) return LMO_E_INVALID;
    if (bad_step) *bad_step = -1;
    int32_t C = lmo_noise_channels(p, dims, pdim);
    work_t w;
    if (work_alloc(&w, N, pdim, dims) != LMO_OK) { work_free(&w); return LMO_E_INVALID; }
    int rc = LMO_OK;
    for (int64_t s = 0; s < nsteps; ++s) {
        int64_t step = step_begin + s;
        const double* rec = noise ? noise + s * (int64_t)C * N : NULL;
        for (int32_t k = 0; k < sp->nspecies; ++k) {
            int64_t b = species_begin(sp, k);
            step_cartesian_segment(p, med, sp, k, dims, pdim, pos + (int64_t)dims * b,
                                   mom + (int64_t)pdim * b, sp->counts[k], step, rec, N, b, &w);
        }
        if (!all_finite(mom, (int64_t)pdim * N) || !all_finite(pos, (int64_t)dims * N)) {
            if (bad_step) *bad_step = step;
            rc = LMO_E_NONFINITE;
            break;
        }
    }
    work_free(&w);
    return rc;
}

/* Milne steps: the Bjorken-mode Cartesian step on (x, y, p) (the noise record of dims 2), then
 * eta_s += lmo_milne_deta(end-of-step momentum) for every quark */
static int run_milne(const lmh_params* p, const lmo_radial_medium* med, const lmo_species* sp,
                     double* pos, double* mom, int64_t N, int64_t step_begin, int64_t nsteps,
                     const double* noise, int64_t* bad_step) {
    if (bad_step) *bad_step = -1;
    lmh_params q = *p;
    q.coords = LMH_CARTESIAN;
    q.bjorken_redshift = 1;
    const int32_t C = lmo_noise_channels(&q, 2, 3);
    double* xy = (double*)malloc((size_t)(2 * N) * sizeof(double));
    work_t w;
    if (!xy || work_alloc(&w, N, 3, 2) != LMO_OK) { free(xy); work_free(&w); return LMO_E_INVALID; }
    milne_split(pos, xy, N);
    int rc = LMO_OK;
    for (int64_t s = 0; s < nsteps; ++s) {
        int64_t step = step_begin + s;
        const double* rec = noise ? noise + s * (int64_t)C * N : NULL;
        for (int32_t k = 0; k < sp->nspecies; ++k) {
            int64_t b = species_begin(sp, k);
            step_cartesian_segment(&q, med, sp, k, 2, 3, xy + 2 * b, mom + 3 * b, sp->counts[k], step,
                                   rec, N, b, &w);
            const double tau_a = p->t0 + (double)(step - 1) * p->dt;
            const double m = sp->mass[k];
#pragma omp parallel for schedule(static)
            for (int64_t i = b; i < b + sp->counts[k]; ++i)
                pos[2 + 3 * i] += lmo_milne_deta(mom[3 * i], mom[3 * i + 1], mom[3 * i + 2], m, tau_a, p->dt);
        }
        milne_join(xy, pos, N);
        if (!all_finite(mom, 3 * N) || !all_finite(pos, 3 * N)) {
            if (bad_step) *bad_step = step;
            rc = LMO_E_NONFINITE;
            break;
        }
    }
    work_free(&w);
    free(xy);
    return rc;
}

/* ======================= general-coordinate (tau, r) step ==================================== */

/* One step of simulate_cpu_general_coords.jl:83-115 on one segment, dims == 2:
 *   kernel_boost_to_rest_frame_general_coords_cpu!   kernels_cpu_GC.jl:64-82   (regulator 1e-10)
 *   kernel_compute_all_forces_general_coords_cpu!     kernels_cpu_GC.jl:110-174 (Gamma == 0,
 *        rows 2..dims, eta_D without the m/E factor, T not clamped)
 *   kernel_update_momenta_LRF_general_coords_cpu!     kernels_cpu_GC.jl:244-256 (mass shell)
 *   kernel_boost_to_lab_frame_general_coords_cpu!     kernels_cpu_GC.jl:85-102  (regulator 1e-8)
 *   kernel_update_positions_general_coords_cpu!       kernels_cpu_GC.jl:259-274
 * The per-particle p_mags/p_units/eta outputs of the force kernel feed nothing later and are
 * not formed.  Returns LMO_E_DOMAIN where Julia's sqrt would throw (kappa < 0 from T < 0). */
static int step_gc_segment(const lmh_params* p, const lmo_radial_medium* med, double m,
                           const lmo_table* tab, double* pos, double* mom, int64_t n,
                           int64_t step, const double* rec, int64_t Ntot, int64_t i0, work_t* w) {
    double dt = p->dt, t0 = p->t0;
    const int32_t dims = 2;
    /* 1. boost to the LRF */
    for (int64_t i = 0; i < n; ++i) {
        double r = fabs(pos[1 + dims * i]);
        double v = lmo_interpolate_2d_gc(med->xgrid, med->nx, med->tgrid, med->nt, med->v, r,
                                         (double)step * dt + t0);
        double gam = 1.0 / sqrt(1.0 - v * v + 1e-10);
        double ptau = mom[0 + dims * i], pr = mom[1 + dims * i];
        mom[0 + dims * i] = gam * (ptau - v * pr);
        mom[1 + dims * i] = gam * (pr - v * ptau);
    }
    /* 2. forces (row 2 only) */
    for (int64_t i = 0; i < n; ++i) {
        double T = lmo_interpolate_2d_gc(med->xgrid, med->nx, med->tgrid, med->nt, med->T,
                                         fabs(pos[1 + dims * i]), (double)step * dt + t0);
        double DsT_safe = p->DsT <= 0.0 ? 0.0 : p->DsT;
        double tau_n = DsT_safe == 0.0 ? 0.0 : lmo_eval_time_table(T, tab->Tmin, tab->invdT,
                                                                    tab->vals, tab->n);
        double etaD = (tau_n > 0.0 && isfinite(tau_n)) ? (1.0 / tau_n) : 0.0;
        double kappa = (tau_n > 0.0 && isfinite(tau_n)) ? ((2.0 * m * T) / tau_n) : 0.0;
        if (kappa < 0.0) return LMO_E_DOMAIN;
        double kT = sqrt(kappa);
        double xi = rec[(int64_t)0 * Ntot + i0 + i];
        double geo_term = 0.0;          /* Christoffel drift with Gamma == 0 (:105-108, :146-154) */
        double det, sto;
        if (etaD > 0.0 && dt > 0.0) {
            double a = exp(-etaD * dt);
            double noise_pref = kT * sqrt((1.0 - a * a) / (2.0 * etaD * dt));
            det = (a - 1.0) * mom[1 + dims * i] + geo_term;
            sto = noise_pref * xi;
        } else {
            det = geo_term;
            sto = kT * xi;
        }
        w->det[i] = det;
        w->sto[i] = sto;
    }
    /* 3. update + mass shell */
    for (int64_t i = 0; i < n; ++i) {
        mom[1 + dims * i] += w->det[i] + sqrt(dt) * w->sto[i];
        double psq = mom[1 + dims * i] * mom[1 + dims * i];
        mom[0 + dims * i] = sqrt(m * m + psq);
    }
    /* 4. boost back */
    for (int64_t i = 0; i < n; ++i) {
        double r = fabs(pos[1 + dims * i]);
        double v = lmo_interpolate_2d_gc(med->xgrid, med->nx, med->tgrid, med->nt, med->v, r,
                                         (double)step * dt + t0);
        double gam = 1.0 / sqrt(1.0 - v * v + 1e-8);
        double ptau = mom[0 + dims * i], pr = mom[1 + dims * i];
        mom[0 + dims * i] = gam * (ptau + v * pr);
        mom[1 + dims * i] = gam * (pr + v * ptau);
    }
    /* 5. push x^mu += dt p^mu / p^tau */
    for (int64_t i = 0; i < n; ++i) {
        double E = mom[0 + dims * i];
        for (int32_t mu = 0; mu < dims; ++mu) pos[mu + dims * i] += dt * mom[mu + dims * i] / E;
    }
    return LMO_OK;
}

int lmo_run_gc(const lmh_params* p, const lmo_radial_medium* med, const lmo_species* sp,
               double* pos, double* mom, int64_t N, int64_t step_begin, int64_t nsteps,
               const double* noise, int64_t* bad_step) {
    if (bad_step) *bad_step = -1;
    if (!noise) return LMO_E_INVALID;
    work_t w;
    if (work_alloc(&w, N, 2, 2) != LMO_OK) { work_free(&w); return LMO_E_INVALID; }
    int rc = LMO_OK;
    for (int64_t s = 0; s < nsteps && rc == LMO_OK; ++s) {
        int64_t step = step_begin + s;
        const double* rec = noise + s * N;
        for (int32_t k = 0; k < sp->nspecies && rc == LMO_OK; ++k) {
            int64_t b = species_begin(sp, k);
            rc = step_gc_segment(p, med, sp->mass[k], &sp->drag[k], pos + 2 * b, mom + 2 * b,
                                 sp->counts[k], step, rec, N, b, &w);
        }
        if (rc == LMO_OK && (!all_finite(mom, 2 * N) || !all_finite(pos, 2 * N))) {
            if (bad_step) *bad_step = step;
            rc = LMO_E_NONFINITE;
        }
    }
    work_free(&w);
    return rc;
}

/* =========================== 2+1D grid medium (extension) ===================================
 * Medium: uniform x and y nodes x0 + k*dx, a sorted tau grid, fields (nx, ny, nt).
 * Sample at (x, y, t): the query is clamped to the grid; the cell index is
 * clamp(floor(f), 0, n-2) with f = (x_c - x0)*(1/dx), and the fraction is f - index; the tau
 * cell and fraction follow interpolate_2d_cpu (kernels_cpu.jl:27-45).  Each of the four corner
 * values is first interpolated in tau, then in x, then in y, in interpolate_2d_cpu's form.   */
static inline int64_t grid_cell(double f, int32_t n) {
    double fl = floor(f);
    if (!(fl >= 0.0)) return 0;
    if (fl > (double)(n - 2)) return n - 2;
    return (int64_t)fl;
}

void lmo_grid_sample(const lmo_grid_medium* g, double x, double y, double t, double out[3]) {
    double xlast = g->x0 + (double)(g->nx - 1) * g->dx;
    double ylast = g->y0 + (double)(g->ny - 1) * g->dy;
    double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy;
    double xc = jl_clamp(x, g->x0, xlast);
    double yc = jl_clamp(y, g->y0, ylast);
    double fx = (xc - g->x0) * inv_dx;
    double fy = (yc - g->y0) * inv_dy;
    int64_t ix = grid_cell(fx, g->nx), iy = grid_cell(fy, g->ny);
    double xd = fx - (double)ix, yd = fy - (double)iy;
    const double* tg = g->tgrid;
    double tc = jl_clamp(t, tg[0], tg[g->nt - 1]);
    int64_t j = clamp_i64(lmo_searchsortedlast(tg, g->nt, tc), 1, g->nt - 1);
    double ta = tg[j - 1], tb = tg[j];
    double dtt = tb - ta;
    double td = fabs(dtt) < JL_EPS ? 0.0 : (tc - ta) / dtt;
    int64_t nx = g->nx, nxy = (int64_t)g->nx * g->ny;
    int64_t s0 = nxy * (j - 1), s1 = nxy * j;
    const double* F[3] = {g->T, g->vx, g->vy};
    for (int f = 0; f < 3; ++f) {
        const double* A = F[f];
        int64_t n00 = ix + nx * iy, n10 = n00 + 1, n01 = n00 + nx, n11 = n01 + 1;
        double c00 = A[s0 + n00] * (1.0 - td) + A[s1 + n00] * td;
        double c10 = A[s0 + n10] * (1.0 - td) + A[s1 + n10] * td;
        double c01 = A[s0 + n01] * (1.0 - td) + A[s1 + n01] * td;
        double c11 = A[s0 + n11] * (1.0 - td) + A[s1 + n11] * td;
        double cx0 = c00 * (1.0 - xd) + c10 * xd;
        double cx1 = c01 * (1.0 - xd) + c11 * xd;
        out[f] = cx0 * (1.0 - yd) + cx1 * yd;
    }
}

/*
 * The frame change on the grid: the reference boosts along r_hat with the radial flow speed
 * (kernels_cpu.jl:101-144, 199-241); the grid medium carries the flow vector (vx, vy), so the
 * boost runs along v_hat = (vx, vy)/|v| with speed |v|, in the same operation order (r_hat ->
 * v_hat, v -> |v|), skipped when |v| < eps as the reference skips r < eps.
 */
static void grid_boost(double* q, int32_t pdim, double m, double vx, double vy, int32_t sign,
                       int32_t relativistic) {
    double vm2 = vx * vx + vy * vy;
    double vmag = sqrt(vm2);
    if (vmag < JL_EPS) return;
    double nx = vx / vmag, ny = vy / vmag;
    if (!relativistic) {
        double s = sign > 0 ? 1.0 : -1.0;
        q[0] += s * (m * vmag * nx);
        q[1] += s * (m * vmag * ny);
        return;
    }
    double vmax = sqrt(1.0 - 1e-12);
    double v = vmag > vmax ? vmax : vmag;
    double gam = 1.0 / sqrt(1.0 - v * v + 1e-10);
    double p2 = 0.0;
    for (int32_t d = 0; d < pdim; ++d) p2 += q[d] * q[d];
    double E = sqrt(p2 + m * m);
    double px = q[0], py = q[1];
    double p_par = px * nx + py * ny;
    double p_perp_x = px - p_par * nx;
    double p_perp_y = py - p_par * ny;
    double p_par_new = sign > 0 ? gam * (p_par + v * E) : gam * (p_par - v * E);
    q[0] = p_par_new * nx + p_perp_x;
    q[1] = p_par_new * ny + p_perp_y;
}

static void step_grid_segment(const lmh_params* p, const lmo_grid_medium* g,
                              const lmo_species* sp, int32_t k, int32_t pdim, double* pos,
                              double* mom, int64_t n, int64_t step, const double* rec,
                              int64_t Ntot, int64_t i0, work_t* w) {
    const int32_t dims = 2;
    int32_t oxi, orta, opos;
    noise_layout(p, dims, pdim, &oxi, &orta, &opos);
    double m = sp->mass[k];
    double dt = p->dt, t0 = p->t0;
    double t_now = (double)step * dt + t0;
    if (oxi >= 0) gather_channels(rec, Ntot, i0, n, oxi, pdim, w->xi);
    if (orta >= 0) {
        gather_channels(rec, Ntot, i0, n, orta, 1, w->ucol);
        gather_channels(rec, Ntot, i0, n, orta + 1, 1, w->usamp);
        gather_channels(rec, Ntot, i0, n, orta + 2, pdim, w->g);
    }
    if (opos >= 0) gather_channels(rec, Ntot, i0, n, opos, dims, w->xpos);
    /* medium at each particle: aux = (T, vx, vy) */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double s3[3];
        lmo_grid_sample(g, pos[0 + dims * i], pos[1 + dims * i], t_now, s3);
        w->aux[3 * i] = s3[0]; w->aux[3 * i + 1] = s3[1]; w->aux[3 * i + 2] = s3[2];
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        grid_boost(mom + (int64_t)pdim * i, pdim, m, w->aux[3 * i + 1], w->aux[3 * i + 2], -1,
                   p->relativistic);
    if (p->bjorken_redshift) kernel_bjorken(mom, pdim, step, dt, t0, n);
    if (p->collision == LMH_FLUID) {
        /* kernels_cpu.jl:713-731 with r_hat -> v_hat; a fluid at rest sets p_perp to 0 */
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            double vx = w->aux[3 * i + 1], vy = w->aux[3 * i + 2];
            double vmag = sqrt(vx * vx + vy * vy);
            double* q = mom + (int64_t)pdim * i;
            if (vmag < JL_EPS) { q[0] = 0.0; q[1] = 0.0; continue; }
            double nx = vx / vmag, ny = vy / vmag;
            if (!p->relativistic) { q[0] = m * vmag * nx; q[1] = m * vmag * ny; continue; }
            double vmax = sqrt(1.0 - 1e-12);
            double v = vmag > vmax ? vmax : vmag;
            double gam = 1.0 / sqrt(1.0 - v * v + 1e-10);
            q[0] = m * gam * v * nx;
            q[1] = m * gam * v * ny;
        }
    } else {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) w->T[i] = jl_max(w->aux[3 * i], 0.0);
        if (p->collision == LMH_LANGEVIN || p->collision == LMH_LANGEVIN_RTA)
            forces_and_update(p, &sp->drag[k], m, w->T, mom, w->xi, pdim, 0, n, w->det, w->sto);
        if (p->collision == LMH_RTA || p->collision == LMH_LANGEVIN_RTA)
            rta_collision(p, &sp->rta[k], &sp->jt[k], w->T, mom, pdim, 0, n, w->ucol, w->usamp,
                          w->g);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i)
            grid_boost(mom + (int64_t)pdim * i, pdim, m, w->aux[3 * i + 1], w->aux[3 * i + 2], +1,
                       p->relativistic);
    }
    /* push (kernels_cpu.jl:537-541, 586-607) */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double E2 = m * m;
        for (int32_t d = 0; d < pdim; ++d) E2 += mom[d + pdim * i] * mom[d + pdim * i];
        double E = p->relativistic ? sqrt(E2) : m;
        for (int32_t d = 0; d < dims; ++d) pos[d + dims * i] += dt * mom[d + pdim * i] / E;
        if (p->position_diffusion) {
            double s3[3];
            lmo_grid_sample(g, pos[0 + dims * i], pos[1 + dims * i], t_now, s3);
            double D = (p->DsT / jl_max(s3[0], JL_EPS)) / FMGEV;
            if (D > 0.0) {
                double sigma = sqrt(2.0 * D * dt);
                for (int32_t d = 0; d < dims; ++d) pos[d + dims * i] += sigma * w->xpos[d + dims * i];
            }
        }
    }
}

int lmo_boost_to_lab_grid(const lmh_params* p, const lmo_grid_medium* g, const lmo_species* sp,
                          int32_t pdim, const double* pos, double* mom, int64_t N, int64_t step) {
    if (pdim != 2 && pdim != 3) return LMO_E_INVALID;
    if (species_begin(sp, sp->nspecies) != N) return LMO_E_INVALID;
    double t = (double)step * p->dt + p->t0;
    for (int32_t k = 0; k < sp->nspecies; ++k) {
        int64_t b = species_begin(sp, k), n = sp->counts[k];
        double m = sp->mass[k];
#pragma omp parallel for schedule(static)
        for (int64_t i = b; i < b + n; ++i) {
            double med[3];
            lmo_grid_sample(g, pos[2 * i], pos[2 * i + 1], t, med);
            grid_boost(mom + (int64_t)pdim * i, pdim, m, med[1], med[2], +1, p->relativistic);
        }
    }
    return LMO_OK;
}

int lmo_run_grid(const lmh_params* p, const lmo_grid_medium* g, const lmo_species* sp,
                 int32_t pdim, double* pos, double* mom, int64_t N, int64_t step_begin,
                 int64_t nsteps, const double* noise, int64_t* bad_step) {
    if (bad_step) *bad_step = -1;
    if (pdim != 2 && pdim != 3) return LMO_E_INVALID;
    if (p->reflecting_boundary) return LMO_E_INVALID;
    int32_t C = lmo_noise_channels(p, 2, pdim);
    work_t w;
    if (work_alloc(&w, N, pdim, 2) != LMO_OK) { work_free(&w); return LMO_E_INVALID; }
    int rc = LMO_OK;
    for (int64_t s = 0; s < nsteps; ++s) {
        int64_t step = step_begin + s;
        const double* rec = noise ? noise + s * (int64_t)C * N : NULL;
        for (int32_t k = 0; k < sp->nspecies; ++k) {
            int64_t b = species_begin(sp, k);
            step_grid_segment(p, g, sp, k, pdim, pos + 2 * b, mom + (int64_t)pdim * b,
                              sp->counts[k], step, rec, N, b, &w);
        }
        if (!all_finite(mom, (int64_t)pdim * N) || !all_finite(pos, 2 * N)) {
            if (bad_step) *bad_step = step;
            rc = LMO_E_NONFINITE;
            break;
        }
    }
    work_free(&w);
    return rc;
}

/* ======================== on-the-fly freeze-out capture ====================================
 * kernel_capture_freezeout_gpu! (kernels_gpu.jl:583-635), applied after a step: for each
 * particle not yet latched, T at its position and tau = step*dt + t0; above Tfo the state, tau
 * and T are cached in last_*; the first time below, the state and tau are interpolated linearly
 * in T between the cache and the current state (frac = 1 without interp or without a cached T
 * above Tfo) and the flag is set.  T comes from the step's interpolation: interpolate_2d_cpu on
 * the radial medium (r = |x| in radial mode, else from all position rows), the grid sample on
 * the 2+1D grid (exactly one of rmed, gmed is given).                                          */
void lmo_freezeout_capture(const lmo_radial_medium* rmed, const lmo_grid_medium* gmed,
                           const double* pos, const double* mom, int32_t dims, int32_t pdim,
                           int64_t N, int64_t step, double dt, double t0, double Tfo,
                           int32_t interp, double* fo_pos, double* fo_mom, double* fo_tau,
                           double* fo_flag, double* last_pos, double* last_mom,
                           double* last_tau, double* last_T) {
    const double tau = (double)step * dt + t0;
    for (int64_t i = 0; i < N; ++i) {
        if (fo_flag[i] != 0.0) continue;
        double T;
        if (gmed) {
            double s3[3];
            lmo_grid_sample(gmed, pos[0 + dims * i], pos[1 + dims * i], tau, s3);
            T = s3[0];
        } else {
            double r;
            if (dims == 1) {
                r = fabs(pos[i]);
            } else {
                double r2 = 0.0;
                for (int32_t d = 0; d < dims; ++d) r2 += pos[d + dims * i] * pos[d + dims * i];
                r = sqrt(r2);
            }
            T = lmo_interpolate_2d(rmed->xgrid, rmed->nx, rmed->tgrid, rmed->nt, rmed->T, r, tau);
        }
        if (T < Tfo) {
            double Tl = last_T[i];
            double frac = (interp && Tl > Tfo && Tl > T) ? (Tl - Tfo) / (Tl - T) : 1.0;
            for (int32_t d = 0; d < dims; ++d) {
                int64_t k = d + (int64_t)dims * i;
                fo_pos[k] = last_pos[k] + frac * (pos[k] - last_pos[k]);
            }
            for (int32_t d = 0; d < pdim; ++d) {
                int64_t k = d + (int64_t)pdim * i;
                fo_mom[k] = last_mom[k] + frac * (mom[k] - last_mom[k]);
            }
            fo_tau[i] = last_tau[i] + frac * (tau - last_tau[i]);
            fo_flag[i] = 1.0;
        } else {
            for (int32_t d = 0; d < dims; ++d) last_pos[d + (int64_t)dims * i] = pos[d + (int64_t)dims * i];
            for (int32_t d = 0; d < pdim; ++d) last_mom[d + (int64_t)pdim * i] = mom[d + (int64_t)pdim * i];
            last_tau[i] = tau;
            last_T[i] = T;
        }
    }
}

/* ================================ observables (extension) ===================================
 * over particles [i0, i1) of a (pdim, N) momentum matrix, see lmh_obs_spec in lmhip.h      */
int lmo_observables(const double* mom, int32_t pdim, int64_t i0, int64_t i1,
                    const lmh_obs_spec* spec, double* out) {
    if (pdim < 2 || spec->nbins < 0) return LMO_E_INVALID;
    int32_t nb = spec->nbins;
    for (int32_t b = 0; b < 5 + nb; ++b) out[b] = 0.0;
    double width = (spec->pt_max - spec->pt_min) / (double)(nb > 0 ? nb : 1);
    double inv_w = (double)nb / (spec->pt_max - spec->pt_min);
    for (int64_t i = i0; i < i1; ++i) {
        double px = mom[0 + (int64_t)pdim * i], py = mom[1 + (int64_t)pdim * i];
        double pt2 = px * px + py * py;
        double pt = sqrt(pt2);
        double c2 = pt2 > 0.0 ? (px * px - py * py) / pt2 : 0.0;
        out[0] += 1.0;
        out[1] += pt;
        out[2] += pt2;
        out[3] += c2;
        out[4] += pt * c2;
        if (nb == 0) continue;
        if (spec->sigma > 0.0) {
            /* bin b spans the edges pt_min + b width and pt_min + (b + 1) width: each edge's erfc
               serves the two bins it bounds */
            double inv_s = 1.0 / (spec->sigma * sqrt(2.0));
            double e_lo = erfc((pt - spec->pt_min) * inv_s);
            for (int32_t b = 0; b < nb; ++b) {
                double hi = spec->pt_min + (double)(b + 1) * width;
                double e_hi = erfc((pt - hi) * inv_s);
                out[5 + b] += 0.5 * (e_hi - e_lo);
                e_lo = e_hi;
            }
        } else {
            double f = (pt - spec->pt_min) * inv_w;
            double fl = floor(f);
            if (fl >= 0.0 && fl < (double)nb) out[5 + (int64_t)fl] += 1.0;
        }
    }
    return LMO_OK;
}
