// lmhip_adjoint.h -- pathwise adjoint of the fused 2+1D Langevin step (included by lmhip.hip
// inside its anonymous namespace, after the forward kernels it reuses).
//
// The loss is L = sum_k w_k O_k over the observables O of the final lab momenta (N, sum p_T,
// sum p_T^2, sum cos 2phi, sum p_T cos 2phi, the [smoothed] p_T histogram); w is the cotangent a
// ChainRulesCore.rrule / torch.autograd.Function pullback receives.  The parameters are
// theta = (DsT, gamma0, Lambda): DsT scales the drag-time table, (gamma0, Lambda) are the
// GAMMA_P drag eta = gamma0 T^2 / (1 + p^2/Lambda^2).
//
// Common random numbers: the noise of the particle with identity i at step s is Philox4x32-10 of
// a counter made of (i, s, stream, block), so each segment's recomputation draws the forward
// pass's noise again and no noise is kept between the passes.  The forward pass keeps the state
// every K steps (checkpoints, in the cell order the segment runs in, with its identities); the
// backward pass walks the segments from the last to the first: it prepares the medium slices of
// the segment's steps (the per-cell coefficients the forward step kernels read), re-runs the
// segment forward from its checkpoint with the step kernel, writing each step's input state to
// a segment tape -- the forward pass's trajectory, bit for bit -- then runs the reverse steps over
// the tape, drawing each step's noise again.
// One thread owns one particle throughout a segment, so no adjoint is exchanged between threads;
// between segments of different orders the adjoint moves with its particle.
//
// The RTA collision decision is a discontinuous function of the state; the pathwise derivative
// keeps each step's decision fixed (a colliding step forgets the incoming momentum and depends
// on the state only through T in the inverse-CDF lookup).

constexpr int kAdjThetas = 3;   // d/dDsT, d/dgamma0, d/dLambda

// d/dT of eval_time_table: the slope of the segment the evaluation uses (transport.jl:231-253
// extrapolates linearly, so the end segments keep their slope outside the table)
__device__ __forceinline__ double eval_time_table_slope(const TimeTable& t, double T) {
    if (t.n < 2 || !isfinite(T)) return 0.0;
    double u = (T - t.Tmin) * t.invdT;
    double fu = floor(u);
    int i;
    if (!(fu >= 0.0)) i = 1;
    else if (fu > (double)(t.n - 2)) i = t.n - 1;
    else i = (int)fu + 1;
    return (t.vals[i] - t.vals[i - 1]) * t.invdT;
}

// eval_time_table and eval_time_table_slope from one read of the segment's two nodes
__device__ __forceinline__ double eval_time_table_and_slope(const TimeTable& t, double T, double* slope) {
    double u = (T - t.Tmin) * t.invdT;
    const double fc = fmin(fmax(floor(u), 0.0), (double)(t.n - 2));
    const bool one = t.n < 2;
    const int i = one ? 0 : (int)fc + 1;
    double w = u - fc;
    const double y0 = ld_off(t.vals, (uint32_t)(one ? 0 : i - 1) * 8u);
    const double y1 = ld_off(t.vals, (uint32_t)i * 8u);
    const bool fin = isfinite(T);
    *slope = (one || !fin) ? 0.0 : (y1 - y0) * t.invdT;
    const double r = (1.0 - w) * y0 + w * y1;
    return one ? y0 : (fin ? r : 0.0);
}

// d/dT of juttner_lookup at fixed quantile u (0 where the T fraction is clamped)
__device__ __forceinline__ double juttner_lookup_dT(const JuttnerTable& jt, double u, double T) {
    double uu = u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u);
    double fu = uu * (double)(jt.nU - 1);
    int iu = (int)floor(fu);
    iu = iu < 0 ? 0 : (iu > jt.nU - 2 ? jt.nU - 2 : iu);
    double tu = fu - (double)iu;
    double fT = (T - jt.Tmin) * jt.invdT;
    double fl = floor(fT);
    int jT;
    if (!(fl >= 0.0)) jT = 0;
    else if (fl > (double)(jt.nT - 2)) jT = jt.nT - 2;
    else jT = (int)fl;
    double tT = fT - (double)jT;
    if (!(tT >= 0.0 && tT <= 1.0)) return 0.0;
    const double* b0 = jt.vals + (int64_t)jT * jt.nU + iu;
    const double* b1 = b0 + jt.nU;
    double c0 = b0[0] * (1.0 - tu) + b0[1] * tu;
    double c1 = b1[0] * (1.0 - tu) + b1[1] * tu;
    return (c1 - c0) * jt.invdT;
}

// The medium at (x, y) from the step's slice (g.slice), as grid_sample<true> computes it (same
// operations, same rounding), and its derivatives with respect to x and y (0 along an axis where
// the position is clamped): d/dxd of a + b xd + (c + d xd) yd is b + d yd, d/dyd is c + d xd.
struct SampleGrad {
    double T, vx, vy;
    double dTx, dTy, dvxx, dvxy, dvyx, dvyy;
};

// in two parts, so that independent work can be placed between the gather and its use
struct SampleCell {
    double4 cT, cx, cy;
    double xd, yd;
};

__device__ __forceinline__ SampleCell grid_sample_cell(const GridMedium& g, double x, double y) {
    SampleCell c;
    double xc = fmin(fmax(x, g.x0), g.xlast);
    double yc = fmin(fmax(y, g.y0), g.ylast);
    double fx = (xc - g.x0) * g.inv_dx;
    double fy = (yc - g.y0) * g.inv_dy;
    int ix = grid_cell(fx, g.nx), iy = grid_cell(fy, g.ny);
    c.xd = fx - (double)ix;
    c.yd = fy - (double)iy;
    const double4* cell = g.slice + 3 * (ix + (int64_t)(g.nx - 1) * iy);
    c.cT = cell[0];
    c.cx = cell[1];
    c.cy = cell[2];
    return c;
}

__device__ __forceinline__ void grid_sample_combine(const GridMedium& g, double x, double y, const SampleCell& c,
                                                    SampleGrad* s) {
    const double xd = c.xd, yd = c.yd;
    const double4 cT = c.cT, cx = c.cx, cy = c.cy;
    s->T = fma(fma(cT.w, xd, cT.z), yd, fma(cT.y, xd, cT.x));
    s->vx = fma(fma(cx.w, xd, cx.z), yd, fma(cx.y, xd, cx.x));
    s->vy = fma(fma(cy.w, xd, cy.z), yd, fma(cy.y, xd, cy.x));
    const double gx = (x >= g.x0 && x <= g.xlast) ? g.inv_dx : 0.0;
    const double gy = (y >= g.y0 && y <= g.ylast) ? g.inv_dy : 0.0;
    s->dTx = fma(cT.w, yd, cT.y) * gx;
    s->dvxx = fma(cx.w, yd, cx.y) * gx;
    s->dvyx = fma(cy.w, yd, cy.y) * gx;
    s->dTy = fma(cT.w, xd, cT.z) * gy;
    s->dvxy = fma(cx.w, xd, cx.z) * gy;
    s->dvyy = fma(cy.w, xd, cy.z) * gy;
}

__device__ __forceinline__ void grid_sample_grad(const GridMedium& g, double x, double y, SampleGrad* s) {
    grid_sample_combine(g, x, y, grid_sample_cell(g, x, y), s);
}

// One forward step, exactly the fused k_cart_grid step (non-deferred, Philox, no position
// diffusion, the medium from the step's slice g.slice), for the tape recomputation.  xi, ucol and
// usamp are the step's block-0 noise (noise_kick).
template <int PDIM, int COLL>
__device__ __forceinline__ void adj_forward_step(const StepArgs& a, const lmh_params& p,
                                                 const GridMedium& g, const SpeciesArgs& sp,
                                                 int64_t i, int64_t step, double* x0, double* x1,
                                                 double* q, const double* xi, double ucol, double usamp_p) {
    constexpr bool HAS_RTA = COLL == LMH_LANGEVIN_RTA;
    const double m = sp.m;
    // the slice holds the step's time; no time cell is needed (no position diffusion here)
    const int j = 0;
    const double td = 0.0;
    double Tm, vx, vy;
    grid_sample<true>(g, *x0, *x1, j, td, &Tm, &vx, &vy);
    const FlowBoost fb = flow_boost(vx, vy);
    apply_boost<PDIM>(q, m, fb, -1, p.relativistic);
    if (PDIM == 3 && p.bjorken_redshift) bjorken(p, step, &q[2 % PDIM]);
    const double T = jl_max(Tm, 0.0);
    double etaD = 0.0, kT = 0.0;
    drag_coeffs<false>(p, sp.drag, m, T, &etaD, &kT);
    ou_kick<PDIM, false>(p, etaD, kT, T, m, a.sqrt_dt, q, xi);
    if (HAS_RTA && rta_decide<true>(p, rta_lookup(sp.rta, T), T, ucol)) {
        double usamp, gg[PDIM];
        noise_redraw<PDIM, LMH_NOISE_PHILOX>(a, i, 0, xi, usamp_p, &usamp, gg);
        rta_redraw<PDIM, false, true, true>(sp.jt, T, q, usamp, gg);
    }
    apply_boost<PDIM>(q, m, fb, +1, p.relativistic);
    grid_push<PDIM, LMH_NOISE_PHILOX, false>(a, p, g, i, step, 0, j, td, m, x0, x1, q);
}

// Reciprocal and reciprocal square root for values that only scale derivatives, where an ulp or
// two from 1/x and 1/sqrt(x) does not matter: the hardware estimates refined by one Newton step
// (three instructions instead of a correctly rounded division's ten).
__device__ __forceinline__ double drv_rcp(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

__device__ __forceinline__ double drv_rsq(double x) {
    const double r = __builtin_amdgcn_rsq(x);
    return fma(r, fma(-(0.5 * x * r), r, 0.5), r);
}

// One reverse step of the particle with identity i.  (x0, x1, qin) is the step's input state; its
// block-0 noise is drawn again from the Philox counters of (i, step).
// lam holds the adjoint of the step's output state (x0, x1, q) on entry and of its input state
// on return; th accumulates dL/dtheta.  The pre-point exact-OU kick, relativistic boosts.  The
// forward values are recomputed from the input state; the collision decision sees the forward
// pass's temperature (the same grid interpolation), everything else here only scales the
// derivatives.
template <int PDIM, int COLL>
__device__ __forceinline__ void adj_reverse_step(const StepArgs& a, const lmh_params& p,
                                                 const GridMedium& g, const SpeciesArgs& sp,
                                                 int64_t i, int64_t step, double x0, double x1,
                                                 const double* qin, double dtau_dDsT, double* lam, double* th) {
    constexpr bool HAS_RTA = COLL == LMH_LANGEVIN_RTA;
    const double m = sp.m, m2 = sp.m * sp.m;
    const double dt = p.dt;
    const double idt = drv_rcp(dt);

    // ---- forward values: the medium cell is requested first; the step's noise, drawn again from
    // its Philox counters (common random numbers), is computed while the cell is in flight (an
    // empty asm that reads the noise keeps the compiler from sinking it to its uses)
    const SampleCell cell = grid_sample_cell(g, x0, x1);
    double xi[PDIM];
    double ucol = 1.0, usamp_p = 0.0;
    noise_kick<PDIM, LMH_NOISE_PHILOX, 1, COLL == LMH_LANGEVIN_RTA ? 1 : 0>(a, i, step, 0, xi, &ucol, &usamp_p);
#pragma unroll
    for (int d = 0; d < PDIM; ++d) asm volatile("" ::"v"(xi[d]));
    SampleGrad s;
    grid_sample_combine(g, x0, x1, cell, &s);
    const double vm2 = s.vx * s.vx + s.vy * s.vy;
    const double ivm0 = rsqrt_pos(fmax(vm2, 1e-300));
    const double vmag = vm2 * ivm0;
    const bool on = !(vmag < kEps);
    const double ivm = on ? ivm0 : 0.0;
    const double nx = s.vx * ivm, ny = s.vy * ivm;
    const double vmax = vmax_();
    const double v = vmag > vmax ? vmax : vmag;
    const double gam = drv_rsq(1.0 - v * v + 1e-10);
    // boost to the LRF: rows 0, 1 move by d1 along n
    double p2in = 0.0;
#pragma unroll
    for (int d = 0; d < PDIM; ++d) p2in += qin[d] * qin[d];
    const double iE1 = drv_rsq(p2in + m2);
    const double E1 = (p2in + m2) * iE1;
    const double ppar1 = qin[0] * nx + qin[1] * ny;
    const double in1 = ppar1 - v * E1;
    const double d1 = on ? gam * in1 - ppar1 : 0.0;
    double qB[PDIM];
#pragma unroll
    for (int d = 0; d < PDIM; ++d) qB[d] = qin[d];
    qB[0] += d1 * nx;
    qB[1] += d1 * ny;
    double fbj = 1.0;
    if (PDIM == 3 && p.bjorken_redshift) {
        const double tau_a = p.t0 + (double)(step - 1) * dt;
        if (tau_a > 0.0) fbj = tau_a / (tau_a + dt);
        qB[2 % PDIM] *= fbj;
    }
    const double T = jl_max(s.T, 0.0);
    double etaD = 0.0, kT = 0.0, kappa = 0.0, tau = 0.0, itau = 0.0, dtau_dT = 0.0;
    bool tab_ok = false;
    if (p.drag_model == LMH_DRAG_GAMMA_P) {
        etaD = p.gamma0 * T * T;
        kappa = 2.0 * m * T * etaD;
        kT = sqrt_pos(kappa);
    } else {
        const double DsT_safe = p.DsT <= 0.0 ? 0.0 : p.DsT;
        double slope = 0.0;
        tau = DsT_safe == 0.0 ? 0.0 : eval_time_table_and_slope(sp.drag, T, &slope);
        tab_ok = tau > 0.0 && isfinite(tau);
        if (tab_ok) {
            itau = drv_rcp(tau);
            etaD = itau;
            kappa = 2.0 * m * T * itau;
            kT = sqrt_pos(kappa);
            dtau_dT = slope;
        }
    }
    const double ikT = kT > 0.0 ? drv_rcp(kT) : 0.0;
    // kick (ou_kick, pre-point branch)
    double q2B = 0.0;
#pragma unroll
    for (int d = 0; d < PDIM; ++d) q2B += qB[d] * qB[d];
    const bool kick = etaD > 0.0;
    double eta = 0.0, aa = 1.0, rr = 0.0, srr = 0.0, isrr = 0.0, ieta = 0.0, iE2 = 0.0, iden = 1.0, iL = 0.0;
    double qK[PDIM];
    if (kick) {
        if (p.drag_model == LMH_DRAG_GAMMA_P) {
            iL = drv_rcp(p.Lambda);
            iden = drv_rcp(1.0 + q2B * iL * iL);
            eta = etaD * iden;
        } else {
            iE2 = drv_rsq(q2B + m2);
            eta = etaD * m * iE2;
        }
        aa = exp_neg(eta * dt);
        ieta = drv_rcp(eta);
        rr = (1.0 - aa * aa) * (0.5 * ieta * idt);
        isrr = rr > 0.0 ? drv_rsq(rr) : 0.0;
        srr = rr * isrr;
#pragma unroll
        for (int d = 0; d < PDIM; ++d) qK[d] = aa * qB[d] + a.sqrt_dt * (kT * srr) * xi[d];
    } else {
#pragma unroll
        for (int d = 0; d < PDIM; ++d) qK[d] = qB[d] + a.sqrt_dt * kT * xi[d];
    }
    // collision (decision fixed)
    bool col = false;
    double qC[PDIM], ghat[PDIM], dpstar = 0.0;
#pragma unroll
    for (int d = 0; d < PDIM; ++d) qC[d] = qK[d];
    if (HAS_RTA && rta_decide<true>(p, rta_lookup(sp.rta, T), T, ucol)) {
        col = true;
        double usamp, gg[PDIM];
        noise_redraw<PDIM, LMH_NOISE_PHILOX>(a, i, 0, xi, usamp_p, &usamp, gg);
        double nrm2 = 0.0;
#pragma unroll
        for (int d = 0; d < PDIM; ++d) nrm2 += gg[d] * gg[d];
        const double inv = drv_rsq(jl_max(nrm2, 1e-300));
        const double pstar = juttner_lookup<true>(sp.jt, usamp, T);
#pragma unroll
        for (int d = 0; d < PDIM; ++d) {
            ghat[d] = gg[d] * inv;
            qC[d] = pstar * ghat[d];
        }
        dpstar = juttner_lookup_dT(sp.jt, usamp, T);
    }
    // boost back to the lab: rows 0, 1 move by d3 along n
    double p2C = 0.0;
#pragma unroll
    for (int d = 0; d < PDIM; ++d) p2C += qC[d] * qC[d];
    const double iE3 = drv_rsq(p2C + m2);
    const double E3 = (p2C + m2) * iE3;
    const double ppar3 = qC[0] * nx + qC[1] * ny;
    const double in3 = ppar3 + v * E3;
    const double d3 = on ? gam * in3 - ppar3 : 0.0;
    double qF[PDIM];
#pragma unroll
    for (int d = 0; d < PDIM; ++d) qF[d] = qC[d];
    qF[0] += d3 * nx;
    qF[1] += d3 * ny;
    double p2F = 0.0;
#pragma unroll
    for (int d = 0; d < PDIM; ++d) p2F += qF[d] * qF[d];
    const double iE4 = drv_rsq(p2F + m2);

    // ---- reverse sweep
    double qb[PDIM];
#pragma unroll
    for (int d = 0; d < PDIM; ++d) qb[d] = lam[2 + d];
    double x0b = lam[0], x1b = lam[1];
    // push: x += dt * q / E
    {
        const double w = dt * iE4;
        qb[0] += lam[0] * w;
        qb[1] += lam[1] * w;
        const double Eb = -(lam[0] * qF[0] + lam[1] * qF[1]) * w * iE4;      // d/dE of dt*q/E
#pragma unroll
        for (int d = 0; d < PDIM; ++d) qb[d] += Eb * qF[d] * iE4;
    }
    double nxb = 0.0, nyb = 0.0, vb = 0.0, gamb = 0.0, Tb = 0.0;
    // boost back (q' = q + (gam (ppar + v E) - ppar) n), input qC
    if (on) {
        const double deltab = qb[0] * nx + qb[1] * ny;
        nxb += qb[0] * d3;
        nyb += qb[1] * d3;
        gamb += deltab * in3;
        vb += deltab * gam * E3;
        const double Eb = deltab * gam * v;
        const double pparb = deltab * (gam - 1.0);
        nxb += pparb * qC[0];
        nyb += pparb * qC[1];
        qb[0] += pparb * nx;
        qb[1] += pparb * ny;
#pragma unroll
        for (int d = 0; d < PDIM; ++d) qb[d] += Eb * iE3 * qC[d];
    }
    if (col) {
        double s_ = 0.0;
#pragma unroll
        for (int d = 0; d < PDIM; ++d) s_ += qb[d] * ghat[d];
        Tb += s_ * dpstar;
#pragma unroll
        for (int d = 0; d < PDIM; ++d) qb[d] = 0.0;
    }
    // kick: qK = a qB + sqrt(dt) kT sqrt(r) xi, a = exp(-eta dt), r = (1 - a^2)/(2 eta dt)
    double kTb = 0.0, etaDb = 0.0;
    {
        double xq = 0.0;
#pragma unroll
        for (int d = 0; d < PDIM; ++d) xq += qb[d] * xi[d];
        if (kick) {
            double ab = 0.0;
#pragma unroll
            for (int d = 0; d < PDIM; ++d) ab += qb[d] * qB[d];
            const double npb = a.sqrt_dt * xq;
            double etab = 0.0;
            if (rr > 0.0) {
                kTb += npb * srr;
                const double rrb = npb * kT * 0.5 * isrr;
                ab += -rrb * aa * ieta * idt;
                etab += -rrb * rr * ieta;
            }
            etab += -ab * dt * aa;
#pragma unroll
            for (int d = 0; d < PDIM; ++d) qb[d] *= aa;
            if (p.drag_model == LMH_DRAG_GAMMA_P) {
                etaDb += etab * iden;
                const double denb = -etab * eta * iden;          // eta = etaD / den
                const double iL2 = iL * iL;
#pragma unroll
                for (int d = 0; d < PDIM; ++d) qb[d] += denb * 2.0 * qB[d] * iL2;
                th[2] += denb * (-2.0 * q2B * iL2 * iL);
            } else {
                etaDb += etab * m * iE2;                        // eta = etaD m / E
                const double E2b = -etab * eta * iE2;
#pragma unroll
                for (int d = 0; d < PDIM; ++d) qb[d] += E2b * qB[d] * iE2;
            }
        } else {
            kTb += a.sqrt_dt * xq;
        }
    }
    // drag coefficients
    if (p.drag_model == LMH_DRAG_GAMMA_P) {
        const double kappab = kTb * 0.5 * ikT;                 // kappa = 2 m T etaD
        Tb += kappab * 2.0 * m * etaD;
        etaDb += kappab * 2.0 * m * T;
        Tb += etaDb * 2.0 * p.gamma0 * T;
        th[1] += etaDb * T * T;
    } else if (tab_ok) {
        const double kappab = kTb * 0.5 * ikT;                 // kappa = 2 m T / tau
        Tb += kappab * 2.0 * m * itau;
        const double taub = -etaDb * itau * itau - kappab * kappa * itau;
        Tb += taub * dtau_dT;
        th[0] += taub * tau * dtau_dDsT;
    }
    const double Tmb = s.T > 0.0 ? Tb : 0.0;
    // Bjorken redshift, then the boost to the LRF (ppar - v E), input qin
    if (PDIM == 3) qb[2 % PDIM] *= fbj;
    if (on) {
        const double deltab = qb[0] * nx + qb[1] * ny;
        nxb += qb[0] * d1;
        nyb += qb[1] * d1;
        gamb += deltab * in1;
        vb -= deltab * gam * E1;
        const double Eb = -deltab * gam * v;
        const double pparb = deltab * (gam - 1.0);
        nxb += pparb * qin[0];
        nyb += pparb * qin[1];
        qb[0] += pparb * nx;
        qb[1] += pparb * ny;
#pragma unroll
        for (int d = 0; d < PDIM; ++d) qb[d] += Eb * iE1 * qin[d];
        // flow factors: gam(v), v = min(|v|, vmax), n = v_vec / |v|
        vb += gamb * gam * gam * gam * v;
        double vmagb = vmag > vmax ? 0.0 : vb;
        vmagb -= (nxb * s.vx + nyb * s.vy) * ivm * ivm;
        const double vxb = nxb * ivm + vmagb * s.vx * ivm;
        const double vyb = nyb * ivm + vmagb * s.vy * ivm;
        x0b += Tmb * s.dTx + vxb * s.dvxx + vyb * s.dvyx;
        x1b += Tmb * s.dTy + vxb * s.dvxy + vyb * s.dvyy;
    } else {
        x0b += Tmb * s.dTx;
        x1b += Tmb * s.dTy;
    }
    lam[0] = x0b;
    lam[1] = x1b;
#pragma unroll
    for (int d = 0; d < PDIM; ++d) lam[2 + d] = qb[d];
}

// dL/dq of the final lab momenta for the cotangent w of the observables (k_obs_partial's
// definitions); positions and p_z do not enter.
__global__ __launch_bounds__(kBlock) void k_adj_seed(const double* __restrict__ px,
                                                      const double* __restrict__ py, int64_t N,
                                                      int64_t i0, int64_t i1, int pdim,
                                                      lmh_obs_spec s, const double* __restrict__ w,
                                                      double* __restrict__ lam) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    const int R = 2 + pdim;
    double gx = 0.0, gy = 0.0;
    if (i >= i0 && i < i1) {
        const double a = px[i], b = py[i];
        const double pt2 = a * a + b * b;
        if (pt2 > 0.0) {
            const double pt = sqrt(pt2);
            const double c2 = (a * a - b * b) / pt2;
            // d pt, d pt^2, d cos2phi
            const double ptx = a / pt, pty = b / pt;
            const double c2x = 4.0 * a * b * b / (pt2 * pt2), c2y = -4.0 * b * a * a / (pt2 * pt2);
            double gpt = w[1] + w[4] * c2;
            if (s.nbins > 0 && s.sigma > 0.0) {
                const double width = (s.pt_max - s.pt_min) / (double)s.nbins;
                const double inv_s = 1.0 / (s.sigma * sqrt(2.0));
                const double k = inv_s / sqrt(M_PI);
                for (int bin = 0; bin < s.nbins; ++bin) {
                    const double lo = s.pt_min + (double)bin * width;
                    const double hi = lo + width;
                    const double zl = (pt - lo) * inv_s, zh = (pt - hi) * inv_s;
                    gpt += w[5 + bin] * k * (exp(-zl * zl) - exp(-zh * zh));
                }
            }
            gx = gpt * ptx + w[2] * 2.0 * a + (w[3] + w[4] * pt) * c2x;
            gy = gpt * pty + w[2] * 2.0 * b + (w[3] + w[4] * pt) * c2y;
        }
    }
    lam[0 * N + i] = 0.0;
    lam[1 * N + i] = 0.0;
    lam[2 * N + i] = gx;
    lam[3 * N + i] = gy;
    for (int r = 4; r < R; ++r) lam[(int64_t)r * N + i] = 0.0;
}

// The options the adjoint supports are fixed at compile time (relativistic pre-point kick, no
// position diffusion, the drag model DRAG); the Bjorken redshift stays a run-time choice.
template <int DRAG>
__device__ __forceinline__ lmh_params adjoint_params(const lmh_params& in) {
    lmh_params p = in;
    p.relativistic = 1;
    p.integrator = LMH_PREPOINT;
    p.position_diffusion = 0;
    p.drag_model = DRAG;
    return p;
}

// A backward segment, first half: re-run the segment's nk steps from its checkpoint, writing
// each step's input state to the tape (rows x0, x1, q...; no noise: the reverse kernel draws the
// step's noise again from its Philox counters).  With state_out (the forward pass that keeps the whole tape) the kernel is the
// forward pass itself: it also takes the last step, writes the final state there (in place over
// ckpt is fine: each thread reads its particle before writing it) and flags non-finite states.
template <int PDIM, int COLL, int DRAG>
__global__ __launch_bounds__(kBlock, kStepWaves) void k_adj_tape(StepArgs a, GridMedium g,
                                                                  const double4* __restrict__ slices,
                                                                  const double* ckpt,
                                                                  int64_t seg_step0, int32_t nk,
                                                                  double* __restrict__ tape,
                                                                  double* state_out) {
    int seg;
    const int64_t i = particle_of_thread(a, &seg);
    if (i < 0) return;
    const SpeciesArgs sp = a.species[seg];
    const lmh_params p = adjoint_params<DRAG>(a.p);
    const int64_t N = a.N;
    constexpr int R = 2 + PDIM;
    constexpr int TR = R;
    double x0 = ckpt[i], x1 = ckpt[N + i];
    double q[PDIM];
#pragma unroll
    for (int d = 0; d < PDIM; ++d) q[d] = ckpt[(int64_t)(2 + d) * N + i];
    int64_t bad = -1;
    for (int k = 0; k < nk; ++k) {
        const int64_t step = seg_step0 + k;
        double* tk = tape + (int64_t)k * TR * N + i;
        tk[0] = x0;
        tk[N] = x1;
#pragma unroll
        for (int d = 0; d < PDIM; ++d) tk[(int64_t)(2 + d) * N] = q[d];
        // the last step's output is the next checkpoint: only its inputs go to the tape
        if (k + 1 < nk || state_out) {
            double xi[PDIM];
            double ucol = 1.0, usamp = 0.0;
            noise_kick<PDIM, LMH_NOISE_PHILOX, 1, COLL == LMH_LANGEVIN_RTA ? 1 : 0>(a, i, step, 0, xi, &ucol, &usamp);
            GridMedium gk = g;
            gk.slice = slices + (int64_t)k * 3 * ((int64_t)(g.nx - 1) * (g.ny - 1));
            adj_forward_step<PDIM, COLL>(a, p, gk, sp, i, step, &x0, &x1, q, xi, ucol, usamp);
            if (state_out && bad < 0 && !(isfinite(x0) && isfinite(x1) && finite_all(q, PDIM))) bad = step;
        }
    }
    if (state_out) {
        state_out[i] = x0;
        state_out[N + i] = x1;
#pragma unroll
        for (int d = 0; d < PDIM; ++d) state_out[(int64_t)(2 + d) * N + i] = q[d];
        if (bad >= 0) atomicMin(a.bad, (unsigned long long)bad);
    }
}

// Second half: the reverse steps over the tape from the last to the first.  lam carries the
// state adjoint between segments (rows as the state), gacc the per-particle dL/dtheta (3 rows),
// both in the tape's order; a.id holds each slot's identity, which keys its noise.
template <int PDIM, int COLL, int DRAG>
__global__ __launch_bounds__(kBlock) void k_adj_reverse(StepArgs a, GridMedium g,
                                                        const double4* __restrict__ slices,
                                                        int64_t seg_step0,
                                                        int32_t nk, const double* __restrict__ tape,
                                                        const double* __restrict__ row0,
                                                        double* __restrict__ lam,
                                                        double* __restrict__ gacc, double dtau_dDsT) {
    int seg;
    const int64_t i = particle_of_thread(a, &seg);
    if (i < 0) return;
    const SpeciesArgs sp = a.species[seg];
    const lmh_params p = adjoint_params<DRAG>(a.p);
    const int64_t N = a.N;
    const int64_t pid = a.id[i];
    constexpr int R = 2 + PDIM;
    constexpr int TR = R;
    double lm[R];
#pragma unroll
    for (int r = 0; r < R; ++r) lm[r] = lam[(int64_t)r * N + i];
    double th[kAdjThetas] = {0.0, 0.0, 0.0};
    // the rows of step k: the tape's, or for k = 0 the checkpoint's when row0 is given (the tape
    // recomputation then starts from it and leaves tape row 0 unwritten); the row of step k - 1 is
    // loaded while step k is reversed
    auto row_of = [&](int k) -> const double* {
        return (k == 0 && row0) ? row0 + i : tape + (int64_t)k * TR * N + i;
    };
    double row[TR];
    {
        const double* tk = row_of(nk - 1);
#pragma unroll
        for (int r = 0; r < TR; ++r) row[r] = tk[(int64_t)r * N];
    }
    for (int k = nk - 1; k >= 0; --k) {
        const int64_t step = seg_step0 + k;
        double cur[TR];
#pragma unroll
        for (int r = 0; r < TR; ++r) cur[r] = row[r];
        if (k > 0) {
            const double* tn = row_of(k - 1);
#pragma unroll
            for (int r = 0; r < TR; ++r) row[r] = tn[(int64_t)r * N];
        }
        GridMedium gk = g;
        gk.slice = slices + (int64_t)k * 3 * ((int64_t)(g.nx - 1) * (g.ny - 1));
        adj_reverse_step<PDIM, COLL>(a, p, gk, sp, pid, step, cur[0], cur[1], cur + 2, dtau_dDsT, lm, th);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) lam[(int64_t)r * N + i] = lm[r];
#pragma unroll
    for (int t = 0; t < kAdjThetas; ++t) gacc[(int64_t)t * N + i] += th[t];
}

// Deterministic sums of `rows` rows of n doubles: per-block partials in a fixed tree, then one
// block per row sums the partials in a fixed tree.
constexpr int kSumBlock = 256;
constexpr int kSumBlocks = 1024;

__global__ __launch_bounds__(kSumBlock) void k_rows_partial(const double* __restrict__ x, int rows,
                                                            int64_t n, double* __restrict__ partial) {
    __shared__ double acc[kSumBlock];
    for (int r = 0; r < rows; ++r) {
        double s = 0.0;
        for (int64_t i = (int64_t)blockIdx.x * kSumBlock + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * kSumBlock)
            s += x[(int64_t)r * n + i];
        acc[threadIdx.x] = s;
        __syncthreads();
        for (int st = kSumBlock / 2; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st) acc[threadIdx.x] += acc[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[(int64_t)r * gridDim.x + blockIdx.x] = acc[0];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kSumBlock) void k_rows_final(const double* __restrict__ partial, int nblk,
                                                          double* __restrict__ out) {
    __shared__ double acc[kSumBlock];
    const int r = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kSumBlock) s += partial[(int64_t)r * nblk + b];
    acc[threadIdx.x] = s;
    __syncthreads();
    for (int st = kSumBlock / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) acc[threadIdx.x] += acc[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[r] = acc[0];
}
