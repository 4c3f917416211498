// lmhip_device.h -- device building blocks of the fused Langevin step (gfx950, FP64).
//
// Every function restates one piece of the reference's CPU kernels (LangevInMedium.jl,
// src/kernels_cpu.jl and src/kernels_cpu_GC.jl) in the reference's order of operations, so that
// the fused kernels reproduce the CPU path to rounding.  The library is compiled with
// -ffp-contract=off: no multiply-add is fused that the reference rounds twice.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lmhip.h"

namespace lmh {

constexpr double kEps = 2.220446049250313e-16;          // Julia eps()
constexpr double kFmGeV = 1.0 / 0.197327;                // constants.jl:13,23

__device__ __forceinline__ double jl_clamp(double x, double lo, double hi) {
    return x > hi ? hi : (x < lo ? lo : x);
}
__device__ __forceinline__ double jl_max(double a, double b) { return b > a ? b : a; }

// Base.searchsortedlast: last 1-based index k with v[k] <= x (0 if none)
__device__ __forceinline__ int searchsortedlast(const double* __restrict__ v, int n, double x) {
    int lo = 0, hi = n + 1;
    while (lo < hi - 1) {
        int m = lo + ((hi - lo) >> 1);
        if (x < v[m - 1]) hi = m; else lo = m;
    }
    return lo;
}

// Cell and fraction of interpolate_2d_cpu along one axis (kernels_cpu.jl:27-45): clamp the
// query, searchsortedlast, clamp the cell to [1, n-1], eps-guarded fraction.  Returns the
// 0-based left node.
__device__ __forceinline__ int axis_cell(const double* __restrict__ g, int n, double q, double* frac) {
    double qc = jl_clamp(q, g[0], g[n - 1]);
    int i = searchsortedlast(g, n, qc);
    i = i < 1 ? 1 : (i > n - 1 ? n - 1 : i);
    double a = g[i - 1], b = g[i];
    double d = b - a;
    *frac = fabs(d) < kEps ? 0.0 : (qc - a) / d;
    return i - 1;
}

// axis_cell with a first guess for a grid of equal spacing (inv = (n - 1)/(g[n-1] - g[0]), the
// host's check; inv = 0 for any other grid): the guessed cell l is taken where its nodes show it
// to be the cell searchsortedlast finds -- g[l] <= q < g[l+1] (or l the last cell and g[l] <= q) --
// and the search runs otherwise, so the result is axis_cell's for every q.  Two independent loads
// instead of a chain of log2(n) dependent ones.
__device__ __forceinline__ int axis_cell_guess(const double* __restrict__ g, int n, double inv, double q,
                                               double* frac) {
    const double qc = jl_clamp(q, g[0], g[n - 1]);
    int i = 0;   // searchsortedlast clamped to [1, n - 1] (1-based)
    bool found = false;
    if (inv > 0.0) {
        const double f = floor((qc - g[0]) * inv);
        const int l = (int)fmin(fmax(f, 0.0), (double)(n - 2));   // a NaN query takes cell 0 here
        const double a = g[l], b = g[l + 1];
        if (a <= qc && (qc < b || l == n - 2)) {
            i = l + 1;
            found = true;
        }
    }
    if (!found) {
        i = searchsortedlast(g, n, qc);
        i = i < 1 ? 1 : (i > n - 1 ? n - 1 : i);
    }
    const double a = g[i - 1], b = g[i];
    const double d = b - a;
    *frac = fabs(d) < kEps ? 0.0 : (qc - a) / d;
    return i - 1;
}

// The same without the query clamp and the eps guard (kernels_cpu_GC.jl:38-51)
__device__ __forceinline__ int axis_cell_gc(const double* __restrict__ g, int n, double q, double* frac) {
    int i = searchsortedlast(g, n, q);
    i = i < 1 ? 1 : (i > n - 1 ? n - 1 : i);
    double a = g[i - 1], b = g[i];
    *frac = (q - a) / (b - a);
    return i - 1;
}

// bilinear combination of interpolate_2d_cpu (kernels_cpu.jl:39-50), column-major (nx, nt)
__device__ __forceinline__ double bilinear(const double* __restrict__ f, int nx, int i, int j,
                                           double xd, double yd) {
    const double* c = f + i + (int64_t)nx * j;
    double v00 = c[0], v10 = c[1];
    double v01 = c[nx], v11 = c[nx + 1];
    double c0 = v00 * (1.0 - xd) + v10 * xd;
    double c1 = v01 * (1.0 - xd) + v11 * xd;
    return c0 * (1.0 - yd) + c1 * yd;
}

// A load of read-only table data through the global address space: a global rather than a
// flat access (the tables' pointers come from memory, where the compiler cannot tell their
// address space), so that waiting for it does not also wait for LDS and scalar loads.
__device__ __forceinline__ double ld_global(const double* p) {
    return *(const __attribute__((address_space(1))) double*)p;
}

// A read-only load at a 32-bit byte offset from a base the whole launch shares: the global
// access with the base in scalar registers and the offset in one vector register, with no 64-bit
// address arithmetic per load (tables and slices are far below 4 GB)
__device__ __forceinline__ double ld_off(const double* base, uint32_t byte_off) {
    return *(const __attribute__((address_space(1))) double*)((const char*)base + byte_off);
}

// the store counterpart of ld_off
__device__ __forceinline__ void st_off(double* base, uint32_t byte_off, double v) {
    *(__attribute__((address_space(1))) double*)((char*)base + byte_off) = v;
}

// (component by component: a copy of the vector type through an address-space pointer would
// go back through a generic pointer)
__device__ __forceinline__ double4 ld_global4(const double4* p) {
    const __attribute__((address_space(1))) double* q = (const __attribute__((address_space(1))) double*)p;
    return make_double4(q[0], q[1], q[2], q[3]);
}

// eval_tau_n_spline, transport.jl:231-253 (index clamped, fraction not clamped)
struct TimeTable {
    const double* vals;
    double Tmin, invdT;
    int n;
};

__device__ __forceinline__ double eval_time_table(const TimeTable& t, double T) {
    double u = (T - t.Tmin) * t.invdT;
    // the 1-based segment i = clamp(floor(u) + 1, 1, n - 1) as a clamp of floor(u) in double
    // (min/max instead of branches; a NaN u takes the first segment), node 0 for a one-node
    // table.  The loads are in bounds for any T, so they are issued unconditionally and the
    // special cases select afterwards.
    const double fc = fmin(fmax(floor(u), 0.0), (double)(t.n - 2));
    const bool one = t.n < 2;
    const int i = one ? 0 : (int)fc + 1;
    double w = u - fc;
    double y0 = ld_off(t.vals, (uint32_t)(one ? 0 : i - 1) * 8u);
    double y1 = ld_off(t.vals, (uint32_t)i * 8u);
    const double r = (1.0 - w) * y0 + w * y1;
    return one ? y0 : (isfinite(T) ? r : 0.0);
}

// Two time tables read at one T.  Tables built over one medium's temperature range share their
// axis (the same Tmin, 1/dT and node count; the drag and RTA tables of a species are, which the
// host records in `shared`): the segment and the weight are then found once, and each value is
// the one eval_time_table gives.  MULTI: the host has checked that the tables have two nodes or
// more (the one-node case compiled out).
template <bool MULTI = false>
__device__ __forceinline__ void eval_time_table_pair(const TimeTable& a, const TimeTable& b, bool shared,
                                                     double T, double* ya, double* yb) {
    if (!shared) {
        *ya = eval_time_table(a, T);
        *yb = eval_time_table(b, T);
        return;
    }
    double u = (T - a.Tmin) * a.invdT;
    const double fc = fmin(fmax(floor(u), 0.0), (double)(a.n - 2));
    const bool one = !MULTI && a.n < 2;
    const int i = one ? 0 : (int)fc + 1;
    double w = u - fc;
    const uint32_t o0 = (uint32_t)(one ? 0 : i - 1) * 8u, o1 = (uint32_t)i * 8u;
    const double a0 = ld_off(a.vals, o0), a1 = ld_off(a.vals, o1);
    const double b0 = ld_off(b.vals, o0), b1 = ld_off(b.vals, o1);
    const double ra = (1.0 - w) * a0 + w * a1;
    const double rb = (1.0 - w) * b0 + w * b1;
    const bool fin = isfinite(T);
    *ya = one ? a0 : (fin ? ra : 0.0);
    *yb = one ? b0 : (fin ? rb : 0.0);
}

// inverse-CDF Jüttner table and its bilinear lookup, kernels_gpu.jl:88-117
struct JuttnerTable {
    const double* vals;
    double Tmin, invdT;
    int nU, nT;
};

// UNIT: u is known to lie in [0, 1) (a Philox quantile), where the clamps of u and of its cell
// change nothing
template <bool UNIT = false>
__device__ __forceinline__ double juttner_lookup(const JuttnerTable& jt, double u, double T) {
    double uu = UNIT ? u : (u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u));
    double fu = uu * (double)(jt.nU - 1);
    int iu = (int)floor(fu);
    if (!UNIT) iu = iu < 0 ? 0 : (iu > jt.nU - 2 ? jt.nU - 2 : iu);
    double tu = fu - (double)iu;
    double fT = (T - jt.Tmin) * jt.invdT;
    double fl = floor(fT);
    // the cell clamped to [0, nT - 2] (a NaN fraction takes cell 0)
    const int jT = (int)fmin(fmax(fl, 0.0), (double)(jt.nT - 2));
    double tT = fT - (double)jT;
    tT = tT < 0.0 ? 0.0 : (tT > 1.0 ? 1.0 : tT);
    const uint32_t o0 = (uint32_t)(jT * jt.nU + iu) * 8u;
    const uint32_t o1 = o0 + (uint32_t)jt.nU * 8u;
    double v00 = ld_off(jt.vals, o0), v10 = ld_off(jt.vals, o0 + 8u);
    double v01 = ld_off(jt.vals, o1), v11 = ld_off(jt.vals, o1 + 8u);
    double c0 = v00 * (1.0 - tu) + v10 * tu;
    double c1 = v01 * (1.0 - tu) + v11 * tu;
    return c0 * (1.0 - tT) + c1 * tT;
}

// ---------------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Block b of particle i at step s: four 64-bit words from
// two evaluations (half h = 0 -> w0, w1; h = 1 -> w2, w3; a word is lo | hi << 32 of consecutive
// outputs) of counter (i, s, stream_lo, stream_hi mod 2^29 | (2b + h) << 29) under key
// (seed_lo, seed_hi).  The host keeps i, s < 2^32 and stream < 2^61.
__device__ __forceinline__ void philox4x32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                           uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        // one 32x32 -> 64-bit multiply per product, one three-input XOR (bitop3 0x96) per word
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32(hi1, c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32(hi0, c3, k1, 0x96);
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// HALVES = 1 fills w0, w1 only
template <int HALVES = 2>
__device__ __forceinline__ void noise_block(uint64_t seed, uint64_t stream, int64_t i, int64_t step,
                                            int block, uint64_t w[4]) {
#pragma unroll
    for (int h = 0; h < HALVES; ++h) {
        uint32_t x[4];
        philox4x32((uint32_t)(uint64_t)i, (uint32_t)(uint64_t)step, (uint32_t)stream,
                   ((uint32_t)(stream >> 32) & 0x1FFFFFFFu) | ((uint32_t)(2 * block + h) << 29),
                   (uint32_t)seed, (uint32_t)(seed >> 32), x);
        w[2 * h] = (uint64_t)x[0] | ((uint64_t)x[1] << 32);
        w[2 * h + 1] = (uint64_t)x[2] | ((uint64_t)x[3] << 32);
    }
}

// uniforms from the top 52 bits of a word: d = 1.m in [1, 2) by setting the exponent, then
// d - 1 in [0, 1) or 2 - d in (0, 1], both exact (three instructions, no integer conversion)
__device__ __forceinline__ double unit_mantissa(uint64_t w) {
    return __longlong_as_double((long long)((w >> 12) | 0x3FF0000000000000ULL));
}
__device__ __forceinline__ double u_co(uint64_t w) { return unit_mantissa(w) - 1.0; }
__device__ __forceinline__ double u_oc(uint64_t w) { return 2.0 - unit_mantissa(w); }

// the collision uniform: a 31-bit uniform in [0,1) from the low 31 bits of a word
__device__ __forceinline__ double u_low31(uint64_t w) { return (double)(uint32_t)(w & 0x7FFFFFFFu) * 0x1.0p-31; }


// 1/sqrt(x) for x > 0 (normal): the hardware estimate y (relative error e ~ 2^-24) refined by
// one third-order step y (1 + d/2 + 3 d^2/8), d = 1 - x y^2 (truncation ~ 5 e^3/16 < 2^-70;
// d from one fused multiply-add, so its rounding stays ~2^-54 absolute): within an ulp of the
// correctly rounded 1/sqrt(x) in five instructions (two Newton steps take seven)
__device__ __forceinline__ double rsqrt_pos(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double d = fma(-(x * y), y, 1.0);
    return fma(y, d * fma(d, 0.375, 0.5), y);
}

// 1/x for finite x > 0 (normal): the hardware estimate y refined by one third-order step
// y (1 + e + e^2), e = 1 - x y: within an ulp of the correctly rounded 1/x in three instructions
__device__ __forceinline__ double rcp_pos(double x) {
    const double y = __builtin_amdgcn_rcp(x);
    const double e = fma(-x, y, 1.0);
    return fma(fma(e, e, e), y, y);
}

// sqrt(x) for x >= 0: g = x/sqrt(x) from the once-refined estimate, then one residual step
// g + (x - g^2)/(2 sqrt(x)); 0 and NaN pass through.  No rescaling of tiny or huge arguments
// (the callers' arguments are energies, rates and variances far from the exponent limits).
// POS: x > 0 is known (an energy squared, at least m^2), so 0 needs no pass-through.
template <bool POS = false>
__device__ __forceinline__ double sqrt_pos(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    y = fma(y, fma(-(h * y), y, 0.5), y);
    const double g = x * y;
    const double r = fma(fma(-g, g, x), 0.5 * y, g);
    return (POS || x > 0.0) ? r : x;
}

// sqrt(x) for x = 0 or x >= 1e-300 (noise radii and kick widths: a product of factors each far
// from the underflow range, or exactly 0): the estimate taken at max(x, 1e-300), whose residual
// step returns exactly 0 for x = 0 (g = 0), so no select is needed; otherwise sqrt_pos.
__device__ __forceinline__ double sqrt_nn(double x) {
    double y = __builtin_amdgcn_rsq(fmax(x, 1e-300));
    const double h = 0.5 * x;
    y = fma(y, fma(-(h * y), y, 0.5), y);
    const double g = x * y;
    return fma(fma(-g, g, x), 0.5 * y, g);
}

// fma(a, b, c) for a wave-uniform constant c, read from a scalar register pair by the v_fma_f64
// itself.  The polynomial coefficients below are such constants; left to itself the compiler
// copies some of them into vector register pairs first (two v_mov_b32 per term, and two more
// VGPRs).  The same IEEE fused multiply-add as fma().
__device__ __forceinline__ double fma_sc(double a, double b, double c) {
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
    return r;
}

// The two transcendental functions of the Box-Muller transform, the same definition as the host
// restatement of the noise record (series, splitting and quadrant logic, fused multiply-adds in
// the same places), from the refined reciprocal and square-root estimates (within an ulp of the
// correctly rounded division and square root the host uses), so the two records agree within
// an ulp or two.
//
// noise_log(u), u in [2^-52, 1]: u = 2^e m, m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s) with
// s = (m - 1)/(m + 1), |s| <= 0.1716, by the atanh series through s^19 (truncation < 3e-17).
constexpr double kLn2Hi = 6.93147180369123816490e-01;   // e * kLn2Hi is exact for |e| < 2^11
constexpr double kLn2Lo = 1.90821492927058770002e-10;

// EXACT: the correctly rounded division the host uses (bit-identical to the host record)
template <bool EXACT = false>
__device__ __forceinline__ double noise_log(double u) {
    uint64_t b = (uint64_t)__double_as_longlong(u);
    int e = (int)((b >> 52) & 0x7FF) - 1023;
    double m = __longlong_as_double((long long)((b & 0x000FFFFFFFFFFFFFULL) | 0x3FF0000000000000ULL));
    if (m > 1.4142135623730951) {
        m = 0.5 * m;
        e += 1;
    }
    // the quotient from the refined reciprocal y and one residual step q + (a - b q) y, which
    // rounds it as the correctly rounded division does (Markstein), short of the scaling and
    // special cases the division's operands here (|a| < 1/2, b in [1.7, 2.5]) never need
    double s;
    if (EXACT) {
        s = (m - 1.0) / (m + 1.0);
    } else {
        const double a = m - 1.0, b = m + 1.0;
        // the estimate refined once (within ~20 ulp); the residual step corrects the quotient
        double y = __builtin_amdgcn_rcp(b);
        y = fma(y, fma(-b, y, 1.0), y);
        const double q = a * y;
        s = fma(fma(-b, q, a), y, q);
    }
    double z = s * s;
    double p = fma_sc(1.0 / 19.0, z, 1.0 / 17.0);
    p = fma_sc(p, z, 1.0 / 15.0);
    p = fma_sc(p, z, 1.0 / 13.0);
    p = fma_sc(p, z, 1.0 / 11.0);
    p = fma_sc(p, z, 1.0 / 9.0);
    p = fma_sc(p, z, 1.0 / 7.0);
    p = fma_sc(p, z, 1.0 / 5.0);
    p = fma_sc(p, z, 1.0 / 3.0);
    double lm = 2.0 * s + (2.0 * s * z) * p;
    double de = (double)e;
    return fma(de, kLn2Hi, fma(de, kLn2Lo, lm));
}

// noise_sincospi(x), x in [0, 2): x = k/2 + r with |r| <= 1/4 (exact), Taylor polynomials of
// sin(pi r) through r^15 and cos(pi r) through r^16 (truncation below 7e-17 and 2e-18 relative),
// then the quadrant.
__device__ __forceinline__ void noise_sincospi(double x, double* sn, double* cs) {
    constexpr double SP[8] = {3.141592653589793, -5.16771278004997, 2.5501640398773455,
                              -0.5992645293207921, 0.08214588661112823, -0.0073704309457143504,
                              0.00046630280576761255, -2.1915353447830217e-05};
    constexpr double CP[9] = {1.0, -4.934802200544679, 4.0587121264167685, -1.3352627688545895,
                              0.2353306303588932, -0.02580689139001406, 0.0019295743094039231,
                              -0.0001046381049248457, 4.303069587032947e-06};
    double k = floor(2.0 * x + 0.5);
    double r = x - 0.5 * k;
    double z = r * r;
    double ps = SP[7];
#pragma unroll
    for (int j = 6; j >= 0; --j) ps = fma_sc(ps, z, SP[j]);
    double pc = CP[8];
#pragma unroll
    for (int j = 7; j >= 0; --j) pc = fma_sc(pc, z, CP[j]);
    double sr = r * ps, cr = pc;
    int q = (int)k & 3;
    double s1 = (q & 1) ? cr : sr;
    double c1 = (q & 1) ? sr : cr;
    *sn = (q & 2) ? -s1 : s1;
    *cs = ((q + 1) & 2) ? -c1 : c1;
}

// (sin, cos) of the angle pi (q/2 + r), q in 0..3, r in [-1/4, 1/4): the polynomials and the
// quadrant logic of noise_sincospi
__device__ __forceinline__ void sincos_qr(int q, double r, double* sn, double* cs) {
    constexpr double SP[8] = {3.141592653589793, -5.16771278004997, 2.5501640398773455,
                              -0.5992645293207921, 0.08214588661112823, -0.0073704309457143504,
                              0.00046630280576761255, -2.1915353447830217e-05};
    constexpr double CP[9] = {1.0, -4.934802200544679, 4.0587121264167685, -1.3352627688545895,
                              0.2353306303588932, -0.02580689139001406, 0.0019295743094039231,
                              -0.0001046381049248457, 4.303069587032947e-06};
    const double z = r * r;
    double ps = SP[7];
#pragma unroll
    for (int j = 6; j >= 0; --j) ps = fma_sc(ps, z, SP[j]);
    double pc = CP[8];
#pragma unroll
    for (int j = 7; j >= 0; --j) pc = fma_sc(pc, z, CP[j]);
    const double sr = r * ps, cr = pc;
    const double s1 = (q & 1) ? cr : sr;
    const double c1 = (q & 1) ? sr : cr;
    *sn = (q & 2) ? -s1 : s1;
    *cs = ((q + 1) & 2) ? -c1 : c1;
}

// (sin, cos) of a uniform angle from a 32-bit word x: the quadrant q = the top 2 bits and
// r = lo30(x)/2^31 - 1/4 in [-1/4, 1/4) (exact)
__device__ __forceinline__ void sincos_u32(uint32_t x, double* sn, double* cs) {
    sincos_qr((int)(x >> 30), fma((double)(x & 0x3FFFFFFFu), 0x1.0p-31, -0.25), sn, cs);
}

// (sin, cos) of a uniform angle from one word: the quadrant q is the word's top 2 bits and the
// angle pi (q/2 + r), r = d/2 - 3/4 in [-1/4, 1/4) with d = 1.m from the next 52 bits (exact);
// the polynomials and the quadrant of noise_sincospi
__device__ __forceinline__ void sincos_word(uint64_t w, double* sn, double* cs) {
    constexpr double SP[8] = {3.141592653589793, -5.16771278004997, 2.5501640398773455,
                              -0.5992645293207921, 0.08214588661112823, -0.0073704309457143504,
                              0.00046630280576761255, -2.1915353447830217e-05};
    constexpr double CP[9] = {1.0, -4.934802200544679, 4.0587121264167685, -1.3352627688545895,
                              0.2353306303588932, -0.02580689139001406, 0.0019295743094039231,
                              -0.0001046381049248457, 4.303069587032947e-06};
    const double r = fma(0.5, unit_mantissa(w << 2), -0.75);
    const double z = r * r;
    double ps = SP[7];
#pragma unroll
    for (int j = 6; j >= 0; --j) ps = fma_sc(ps, z, SP[j]);
    double pc = CP[8];
#pragma unroll
    for (int j = 7; j >= 0; --j) pc = fma_sc(pc, z, CP[j]);
    const double sr = r * ps, cr = pc;
    const int q = (int)(uint32_t)(w >> 62);
    const double s1 = (q & 1) ? cr : sr;
    const double c1 = (q & 1) ? sr : cr;
    *sn = (q & 2) ? -s1 : s1;
    *cs = ((q + 1) & 2) ? -c1 : c1;
}

// noise_cospi_half(x) = cos(pi x) for x in [0, 1/2]: the Taylor polynomial in x^2 through x^20
// (truncation below 2e-17), Horner's rule with fused multiply-adds; the same definition on the host
__device__ __forceinline__ double noise_cospi_half(double x) {
    constexpr double C[11] = {1.0, -4.934802200544679, 4.0587121264167685, -1.3352627688545895, 0.2353306303588932,
                              -0.02580689139001406, 0.0019295743094039231, -0.0001046381049248457,
                              4.303069587032947e-06, -1.3878952462213771e-07, 3.604730797462501e-09};
    const double z = x * x;
    double p = C[10];
#pragma unroll
    for (int j = 9; j >= 0; --j) p = fma_sc(p, z, C[j]);
    return p;
}

// Block 0's kick normals: the first NN (<= 3) components of a 4-D standard normal vector
// R (sqrt(v) cos(beta), sqrt(v) sin(beta), sqrt(1 - v) cos(gamma), sqrt(1 - v) sin(gamma)).
// R^2 = -2 ln(u1 u2) is chi^2(4) (the sum of two independent chi^2(2) = -2 ln u variables); v,
// the squared modulus of the first complex coordinate of a uniform point on the 3-sphere, is
// uniform on [0, 1); beta and gamma are uniform angles.  u1 is the (0, 1] uniform of w0's top
// 52 bits, u2 = (lo(w1) + 1)/2^32, beta the angle of hi(w1) (sincos_u32), v = lo(w2)/2^32, and
// cos(gamma) a random sign (bit 31 of hi(w2)) times cos(pi x), x = lo31(hi(w2))/2^32 in
// [0, 1/2).  One logarithm for three normals (Box-Muller takes one per pair); w3 stays for the
// collision uniform and the redraw's quantile.
template <int NN>
__device__ __forceinline__ void kick_normals(const uint64_t w[4], double* n) {
    const double P = u_oc(w[0]) * (((double)(uint32_t)w[1] + 1.0) * 0x1.0p-32);
    const double L = -2.0 * noise_log(P);
    const double v = (double)(uint32_t)w[2] * 0x1.0p-32;
    const double a = sqrt_nn(v * L);
    double s, c;
    sincos_u32((uint32_t)(w[1] >> 32), &s, &c);
    n[0] = a * c;
    if (NN >= 2) n[1 % NN] = a * s;
    if (NN >= 3) {
        const double b = sqrt_nn((1.0 - v) * L);
        const uint32_t hg = (uint32_t)(w[2] >> 32);
        const double nb = b * noise_cospi_half((double)(hg & 0x7FFFFFFFu) * 0x1.0p-32);
        n[2 % NN] = (hg >> 31) ? -nb : nb;
    }
}

// the redraw's quantile: the 32-bit uniform hi(w3)/2^32 in [0, 1)
__device__ __forceinline__ double u_hi32(uint64_t w) { return (double)(uint32_t)(w >> 32) * 0x1.0p-32; }

// Box-Muller: radius from the (0,1] uniform of wa, the uniform angle of wb (sincos_word).
// EXACT takes the correctly rounded division and square root of the host restatement, so the
// pair is the host's to the bit (the position-diffusion block, whose normals are multiplied by
// jumps that D ~ 1/T makes large in a cold medium)
template <bool EXACT = false>
__device__ __forceinline__ void box_muller(uint64_t wa, uint64_t wb, double* n0, double* n1) {
    const double l = -2.0 * noise_log<EXACT>(u_oc(wa));
    double r = EXACT ? sqrt(l) : sqrt_pos(l);
    double s, c;
    sincos_word(wb, &s, &c);
    *n0 = r * c;
    *n1 = r * s;
}

// ---------------------------------------------------------------------------------------------
// exp(-x) for x >= 0.  Up to x = 1/16 (every drag attenuation of a step of a few hundredths of
// tau) the Taylor polynomial through x^9, whose remainder (< 3e-19) is far below the rounding
// of the result; larger or non-finite arguments take exp.
__device__ __forceinline__ double exp_neg(double x) {
    if (!(x <= 0.0625)) return exp(-x);
    // Horner's rule in y = -x, r_k = r_(k+1) y + 1/k!, written in x: s_k = (-1)^k r_k =
    // s_(k+1) x + (-1)^k/k! (each step the same fused multiply-add up to sign, so the same bits)
    double r = fma_sc(-1.0 / 362880.0, x, 1.0 / 40320.0);
    r = fma_sc(r, x, -1.0 / 5040.0);
    r = fma_sc(r, x, 1.0 / 720.0);
    r = fma_sc(r, x, -1.0 / 120.0);
    r = fma_sc(r, x, 1.0 / 24.0);
    r = fma_sc(r, x, -1.0 / 6.0);
    r = fma(r, x, 0.5);
    r = fma(r, x, -1.0);
    return fma(r, x, 1.0);
}

// drag_table_read: whether the drag time table is read (DsT > 0 and the TABLE model; a NaN DsT
// reads it); drag_coeffs_tau takes the table value tau read at T (0 where it is not read).
__device__ __forceinline__ bool drag_table_read(const lmh_params& p) {
    return p.drag_model != LMH_DRAG_GAMMA_P && !(p.DsT <= 0.0);
}

// drag/noise coefficients from the local (clamped) T: kernels_cpu.jl:270-290, eta_D = 1/tau and
// kT = sqrt(kappa), kappa = 2 m T / tau; or the GAMMA_P model eta_D = gamma0*T^2,
// kappa = 2 m T eta_D.  eta_D = 0 gives kappa = 0 in both models.  EXACT keeps the reference's
// divisions (the radial-mode Euler kick uses kT); otherwise 1/tau is the refined reciprocal and
// kappa = 2 m T eta_D (the exact-OU kick below needs kT only in its radial form).
template <bool EXACT>
__device__ __forceinline__ void drag_coeffs_tau(const lmh_params& p, double tau, double m, double T,
                                                double* etaD, double* kT) {
    if (p.drag_model == LMH_DRAG_GAMMA_P) {
        double e = p.gamma0 * T * T;
        *etaD = e;
        *kT = sqrt(2.0 * m * T * e);
        return;
    }
    bool ok = tau > 0.0 && isfinite(tau);
    if (EXACT) {
        *etaD = ok ? (1.0 / tau) : 0.0;
        double kappa = ok ? ((2.0 * m * T) / tau) : 0.0;
        *kT = sqrt_pos(kappa);
        return;
    }
    const double e = ok ? rcp_pos(tau) : 0.0;
    *etaD = e;
    *kT = sqrt_pos((2.0 * m * T) * e);
}

template <bool EXACT>
__device__ __forceinline__ void drag_coeffs(const lmh_params& p, const TimeTable& tab, double m,
                                            double T, double* etaD, double* kT) {
    const double tau = drag_table_read(p) ? eval_time_table(tab, T) : 0.0;
    drag_coeffs_tau<EXACT>(p, tau, m, T, etaD, kT);
}

// effective drag at momentum q -- eta_D*m/E (relativistic) or eta_D (kernels_cpu.jl:334-339),
// or eta_D/(1+|q|^2/Lambda^2) for GAMMA_P -- and c = kappa/(2 eta_eff) = m T eta_D/eta_eff:
// T E (relativistic), m T, or m T (1+|q|^2/Lambda^2) (GAMMA_P)
// FAST: the GAMMA_P denominator's two divisions as products with refined reciprocals (an ulp or
// two from the divisions)
template <int PDIM, bool FAST = false>
__device__ __forceinline__ double eta_of(const lmh_params& p, double etaD, double m, double T,
                                         const double* q, double* c) {
    double q2 = q[0] * q[0];
#pragma unroll
    for (int d = 1; d < PDIM; ++d) q2 = fma(q[d], q[d], q2);
    if (p.drag_model == LMH_DRAG_GAMMA_P) {
        const double den = FAST ? fma(q2, rcp_pos(p.Lambda * p.Lambda), 1.0) : 1.0 + q2 / (p.Lambda * p.Lambda);
        *c = (m * T) * den;
        return FAST ? etaD * rcp_pos(den) : etaD / den;
    }
    if (p.relativistic) {
        const double E2 = fma(m, m, q2);
        const double iE = rsqrt_pos(E2);
        *c = T * (E2 * iE);
        return (etaD * m) * iE;
    }
    *c = m * T;
    return etaD;
}

// The LRF momentum kick: kernel_compute_all_forces_cpu! (kernels_cpu.jl:293-352) followed by
// kernel_update_momenta_LRF_cpu! (:406-417), p += det + sqrt(dt)*sto.  RADIAL selects the Euler
// form of radial mode (:293-323, sto = kT*xi since kL == kT).  MIDPOINT follows
// kernels_gpu.jl:368-381; POSTPOINT re-evaluates the drag at the updated momentum by three
// fixed-point sweeps.
//
// The exact-OU noise sqrt(dt)*sto = sqrt(dt) kT sqrt((1 - a^2)/(2 eta_eff dt)) xi (:340-349) is
// written as sqrt((1 - a^2) kappa/(2 eta_eff)) xi = sqrt((1 - a^2) c) xi with c from eta_of: the
// same quantity without the divisions (within a few ulps of the reference's rounding).
template <int PDIM, bool RADIAL, bool FAST = false>
__device__ __forceinline__ void ou_kick(const lmh_params& p, double etaD, double kT, double T,
                                        double m, double sqrt_dt, double* q, const double* xi) {
    const double dt = p.dt;
    if (RADIAL) {
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < PDIM; ++d) s += q[d] * q[d];
        double pm = sqrt(s);
        double E_r = sqrt(pm * pm + m * m);
        double eta_r = p.relativistic ? etaD * m / E_r : etaD;
#pragma unroll
        for (int d = 0; d < PDIM; ++d) {
            double det = -eta_r * q[d] * dt;
            double sto = kT * 1.0 * xi[d];
            q[d] = q[d] + (det + sqrt_dt * sto);
        }
        return;
    }
    if (!(etaD > 0.0)) {
        // eta_D = 0: kappa = 0, the kick is sqrt(dt) * (0 * xi)
#pragma unroll
        for (int d = 0; d < PDIM; ++d) q[d] = q[d] + (0.0 + sqrt_dt * (0.0 * xi[d]));
        return;
    }
    double c;
    if (p.integrator == LMH_POSTPOINT && dt > 0.0) {
        double qk[PDIM];
#pragma unroll
        for (int d = 0; d < PDIM; ++d) qk[d] = q[d];
        for (int it = 0; it < 3; ++it) {
            const double eta_eff = eta_of<PDIM, FAST>(p, etaD, m, T, qk, &c);
            const double a = exp_neg(eta_eff * dt);
            const double s = sqrt_pos(fma(-a, a, 1.0) * c);
#pragma unroll
            for (int d = 0; d < PDIM; ++d) qk[d] = q[d] + fma(s, xi[d], (a - 1.0) * q[d]);
        }
#pragma unroll
        for (int d = 0; d < PDIM; ++d) q[d] = qk[d];
        return;
    }
    double eta_eff = eta_of<PDIM, FAST>(p, etaD, m, T, q, &c);
    if (p.integrator == LMH_MIDPOINT && dt > 0.0) {
        const double a_half = exp_neg(eta_eff * (0.5 * dt));
        double qm[PDIM];
#pragma unroll
        for (int d = 0; d < PDIM; ++d) qm[d] = a_half * q[d];
        eta_eff = eta_of<PDIM, FAST>(p, etaD, m, T, qm, &c);
    }
    const double a = exp_neg(eta_eff * dt);
    const double s = dt > 0.0 ? sqrt_pos(fma(-a, a, 1.0) * c) : 0.0;
#pragma unroll
    for (int d = 0; d < PDIM; ++d) q[d] = q[d] + fma(s, xi[d], (a - 1.0) * q[d]);
}

// BGK collision in the LRF: kernel_rta_collision_cpu! (kernels_cpu.jl:485-505) with the
// inverse-CDF draw of kernel_rta_collision_gpu! (kernels_gpu.jl:673-701).  The decision and the
// redraw are split so that the redraw's noise is generated only for colliding particles.
// The collision test, split so that the table read can be issued early and its value used late:
// rta_lookup reads tau_n(T) (the table must exist), rta_decide applies the reference's cases
// (no DsT or T <= 0: tau_n = 0, a certain collision).
__device__ __forceinline__ double rta_lookup(const TimeTable& taun, double T) {
    return eval_time_table(taun, T);
}

// UNIT: ucol lies in [0, 1) (a Philox uniform), so ucol < 1 holds: a collision is certain unless
// D_sT > 0, T > 0 and tau_n is a finite positive number (one class test), and otherwise decided
// by ucol tau_n < dt -- the same decision in fewer instructions.
template <bool UNIT = false>
__device__ __forceinline__ bool rta_decide(const lmh_params& p, double tn, double T, double ucol) {
    if (UNIT) {
        const bool pos = __builtin_amdgcn_class(tn, 0x180);   // +subnormal or +normal
        return !(p.DsT > 0.0 && T > 0.0 && pos) || ucol * tn < p.dt;
    }
    const double tau_n = (p.DsT > 0.0 && T > 0.0) ? tn : 0.0;
    // ucol < clamp(dt/tau_n, 0, 1) with ucol in [0, 1) is ucol * tau_n < dt for tau_n > 0 (no
    // division; the two forms differ only when ucol * tau_n is within an ulp of dt)
    if (!(tau_n > 0.0 && isfinite(tau_n))) return ucol < 1.0;
    return ucol * tau_n < p.dt;
}

__device__ __forceinline__ bool rta_collides(const lmh_params& p, const TimeTable& taun, double T,
                                             double ucol) {
    return rta_decide(p, rta_lookup(taun, T), T, ucol);
}

// The redraw of kernel_rta_collision_gpu! (kernels_gpu.jl:673-701): |p*| from the table, the
// direction g normalised (the normalised Gaussians of kernels_cpu.jl:497-503).  FAST takes
// p* / |g| from the refined reciprocal square root (within an ulp or two of the division by the
// square root); UNIT as in juttner_lookup.
template <int PDIM, bool RADIAL, bool FAST = false, bool UNIT = false>
__device__ __forceinline__ void rta_redraw(const JuttnerTable& jt, double T, double* q,
                                           double usamp, const double* g) {
    double pstar = juttner_lookup<UNIT>(jt, usamp, T);
    if (RADIAL) {
        q[0] = g[0] < 0.0 ? -pstar : pstar;
    } else if (FAST) {
        double nrm2 = g[0] * g[0];
#pragma unroll
        for (int d = 1; d < PDIM; ++d) nrm2 = fma(g[d], g[d], nrm2);
        // fmax, not jl_max: a NaN norm (a NaN g) gives NaN components either way
        const double inv = pstar * rsqrt_pos(fmax(nrm2, 1e-300));
#pragma unroll
        for (int d = 0; d < PDIM; ++d) q[d] = g[d] * inv;
    } else {
        double nrm2 = 0.0;
#pragma unroll
        for (int d = 0; d < PDIM; ++d) nrm2 += g[d] * g[d];
        double inv = pstar / sqrt(jl_max(nrm2, 1e-300));
#pragma unroll
        for (int d = 0; d < PDIM; ++d) q[d] = g[d] * inv;
    }
}

}  // namespace lmh
