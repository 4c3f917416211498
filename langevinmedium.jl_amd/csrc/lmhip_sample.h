// lmhip_sample.h -- initial-state samplers on the device (included by lmhip.hip inside its
// anonymous namespace, after the step kernels).
//
// The reference samples the initial quarks on the host before every run: positions and transverse
// momenta from the FONLL density (src/utils.jl:33-172, sample_particles_from_FONLL, called by
// src/simulate_cpu.jl:66-94) and, for three momentum rows, a thermal p_z (src/utils.jl:499-558,
// append_thermal_pz / sample_pz_conditional_juttner).  Here one thread draws one quark.  The
// inverse-CDF tables are built on the host exactly as the reference builds them; the draws are
// Philox4x32-10 words of noise block 3 under the run's key (seed, stream) and the draw's index,
// with counter word 1 selecting the draw: position attempts 0, 1, 2, ..., the momentum
// 0xFFFFFFFF, the thermal p_z attempts 0x80000000 + k.  Uniforms are the 52-bit uniforms of the
// words.  The steps use blocks 0 and 2, so no sampler word repeats a step's.

constexpr int kSampleBlock = 3;
constexpr uint32_t kSampleMomentum = 0xFFFFFFFFu;
constexpr uint32_t kSamplePz = 0x80000000u;

// the inverse CDF of utils.jl: Interpolations.LinearInterpolation(cdf, vals) with flat
// extrapolation, cdf strictly increasing from 0 to 1 (the reference's monotone fix-up)
__device__ __forceinline__ double inverse_cdf(const double* __restrict__ cdf, const double* __restrict__ vals,
                                              int n, double u) {
    if (!(u > cdf[0])) return vals[0];
    if (!(u < cdf[n - 1])) return vals[n - 1];
    const int k = searchsortedlast(cdf, n, u);     // 1-based, 1 <= k <= n - 1 here
    const double t = (u - cdf[k - 1]) / (cdf[k] - cdf[k - 1]);
    return (1.0 - t) * vals[k - 1] + t * vals[k];
}

// Interpolations.LinearInterpolation(r_grid, f) with extrapolation_bc = 0 (the Cartesian
// sampler's spatial density)
__device__ __forceinline__ double linear_or_zero(const double* __restrict__ x, const double* __restrict__ f,
                                                 int n, double q) {
    if (!(q >= x[0] && q <= x[n - 1])) return 0.0;
    int k = searchsortedlast(x, n, q);
    k = k > n - 1 ? n - 1 : k;
    const double t = (q - x[k - 1]) / (x[k] - x[k - 1]);
    return (1.0 - t) * f[k - 1] + t * f[k];
}

struct FonllTables {
    const double* rg;       // r grid (nr)
    const double* pg;       // p grid (np)
    const double* cdf_r;    // polar: CDF of the r marginal (nr)
    const double* fsp;      // cartesian: spatial density n(r) on the r grid (nr)
    const double* cdf_p;    // the p|r CDFs, column j for r-grid node j (np x nr)
    const double* p_ok;     // 1 where column j has a positive density, else 0 (p = 0)
    int32_t nr, np;
    int32_t cartesian;
    int32_t antithetic;
    double rmax, fmax;      // cartesian: the square's half side and the rejection envelope
    int64_t N;              // quarks to draw
    int64_t nhalf;          // antithetic: pairs (2k, 2k+1) share draw k; an odd last quark draws nhalf
    uint64_t seed, stream;
    int32_t max_attempts;   // cartesian rejection cap (the reference loops until accepted)
};

// |p| from the p|r CDF of the r-grid node at r (utils.jl:119-125: clamp, searchsortedlast)
__device__ __forceinline__ double fonll_pmag(const FonllTables& t, double r, double u) {
    const double rc = jl_clamp(r, t.rg[0], t.rg[t.nr - 1]);
    int j = searchsortedlast(t.rg, t.nr, rc);
    j = j < 1 ? 1 : (j > t.nr ? t.nr : j);
    const int c = j - 1;
    if (!(t.p_ok[c] != 0.0)) return 0.0;
    return inverse_cdf(t.cdf_p + (int64_t)c * t.np, t.pg, t.np, u);
}

// one quark per thread; x_out, p_out are (2, N) column-major
__global__ __launch_bounds__(kBlock) void k_sample_fonll(FonllTables t, double* __restrict__ x_out,
                                                         double* __restrict__ p_out, int32_t* __restrict__ fail) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= t.N) return;
    int64_t k = i;
    double sgn = 1.0;
    if (t.antithetic && i < 2 * t.nhalf) {
        k = i >> 1;
        sgn = (i & 1) ? -1.0 : 1.0;
    } else if (t.antithetic) {
        k = t.nhalf;
    }
    double x = 0.0, y = 0.0, r = 0.0;
    if (!t.cartesian) {
        // utils.jl:160-166: r from the inverse CDF of the marginal, clamped to the grid; phi uniform
        uint64_t w[4];
        noise_block<1>(t.seed, t.stream, k, 0, kSampleBlock, w);
        r = jl_clamp(inverse_cdf(t.cdf_r, t.rg, t.nr, u_co(w[0])), t.rg[0], t.rg[t.nr - 1]);
        double s, c;
        noise_sincospi(2.0 * u_co(w[1]), &s, &c);
        x = r * c;
        y = r * s;
    } else {
        // utils.jl:98-127: uniform (x, y) in the square, kept inside the disk with probability
        // n(r)/f_max
        bool done = false;
        for (int32_t a = 0; a < t.max_attempts && !done; ++a) {
            uint64_t w[4];
            noise_block<2>(t.seed, t.stream, k, (int64_t)a, kSampleBlock, w);
            const double xa = t.rmax * (2.0 * u_co(w[0]) - 1.0);
            const double ya = t.rmax * (2.0 * u_co(w[1]) - 1.0);
            const double ra = sqrt(xa * xa + ya * ya);
            if (ra > t.rmax) continue;
            if (u_co(w[2]) * t.fmax > linear_or_zero(t.rg, t.fsp, t.nr, ra)) continue;
            x = xa;
            y = ya;
            r = ra;
            done = true;
        }
        if (!done) atomicOr(fail, 1);
    }
    // utils.jl:119-125 / 168-172: |p| from the p|r CDF, its azimuth uniform
    uint64_t w[4];
    noise_block<1>(t.seed, t.stream, k, (int64_t)kSampleMomentum, kSampleBlock, w);
    const double pm = fonll_pmag(t, r, u_co(w[0]));
    double s, c;
    noise_sincospi(2.0 * u_co(w[1]), &s, &c);
    x_out[2 * i] = x;
    x_out[2 * i + 1] = y;
    p_out[2 * i] = sgn * (pm * c);
    p_out[2 * i + 1] = sgn * (pm * s);
}

// sample_pz_conditional_juttner (utils.jl:499-530): one draw of p_z ~ exp(-sqrt(mT^2+pz^2)/T) by
// the reference's rejection samplers, attempt a from block 3's counter word 0x80000000 + a
__device__ __forceinline__ double thermal_pz(uint64_t seed, uint64_t stream, int64_t k, double mT, double T) {
    if (!(T > 0.0 && mT > 0.0)) return 0.0;
    const double a = mT / T;
    if (a > 1.05) {
        // Gaussian envelope in the rapidity y, p_z = mT sinh(y)
        const double c = fmax(2.0, a / (a - 1.0));
        const double sig = sqrt(c / a);
        for (int it = 0; it < 100000; ++it) {
            uint64_t w[4];
            noise_block<2>(seed, stream, k, (int64_t)(kSamplePz + (uint32_t)it), kSampleBlock, w);
            double n0, n1;
            box_muller(w[0], w[1], &n0, &n1);
            const double y = sig * n0;
            const double ch = cosh(y);
            const double acc = ch * exp(-a * (ch - 1.0) + a * y * y / (2.0 * c));
            if (u_co(w[2]) <= acc) return mT * sinh(y);
        }
        return 0.0;
    }
    // Laplace envelope in p_z
    const double b = sqrt(2.0) * T;
    for (int it = 0; it < 1000000; ++it) {
        uint64_t w[4];
        noise_block<2>(seed, stream, k, (int64_t)(kSamplePz + (uint32_t)it), kSampleBlock, w);
        const double pz = -b * noise_log(u_oc(w[0])) * (u_co(w[1]) < 0.5 ? -1.0 : 1.0);
        const double acc = exp(-(sqrt(mT * mT + pz * pz) - (mT + fabs(pz)) / sqrt(2.0)) / T);
        if (u_co(w[2]) <= acc) return pz;
    }
    return 0.0;
}

// append_thermal_pz (utils.jl:533-558): out (3, N) = the (2, N) transverse momenta and p_z drawn
// at T = max(T(r, t0), 0) of the radial medium, r = |x|; antithetic mirrors p_z across the pairs
// (2k, 2k+1) (0-based: an odd quark takes minus its predecessor's draw)
__global__ __launch_bounds__(kBlock) void k_thermal_pz(const double* __restrict__ mom, const double* __restrict__ pos,
                                                       int32_t dims, int64_t N, double m, double t0, int32_t antithetic,
                                                       uint64_t seed, uint64_t stream, RadialMedium med,
                                                       double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    const bool mirror = antithetic && (i & 1);
    const int64_t k = mirror ? i - 1 : i;
    double r2 = 0.0;
    for (int d = 0; d < dims; ++d) r2 += pos[(int64_t)dims * k + d] * pos[(int64_t)dims * k + d];
    double xd, yd;
    const int ir = axis_cell_guess(med.xg, med.nx, med.xinv, sqrt(r2), &xd);
    const int jt = axis_cell_guess(med.tg, med.nt, med.tinv, t0, &yd);
    const double T = jl_max(bilinear(med.T, med.nx, ir, jt, xd, yd), 0.0);
    const double px = mom[2 * k], py = mom[2 * k + 1];
    const double mT = sqrt(m * m + px * px + py * py);
    const double pz = thermal_pz(seed, stream, k, mT, T);
    out[3 * i] = mom[2 * i];
    out[3 * i + 1] = mom[2 * i + 1];
    out[3 * i + 2] = mirror ? -pz : pz;
}
