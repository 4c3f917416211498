/*
 * lm_oracle.h -- CPU restatement of LangevInMedium.jl's per-particle Langevin step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP backend is compared against.
 * Only the tests, the smoke check and the CPU-baseline leg of the benchmark may load it; the
 * product (liblmhip.so and the Python host package) never does.
 *
 * Arrays use the reference's Julia layout: a (rows, N) particle matrix holds (d, i) at
 * [d + rows*i]; an (n_r, n_t) medium matrix holds (i, j) at [i + n_r*j].  The fixed-noise
 * record of one step is the (C, N) row-major block described in include/lmhip.h.
 */
#ifndef LM_ORACLE_H
#define LM_ORACLE_H

#include <stdint.h>
#include "../include/lmhip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LMO_OK          0
#define LMO_E_INVALID  -1
#define LMO_E_NONFINITE -4
#define LMO_E_DOMAIN   -6   /* a Julia DomainError (sqrt of a negative number) */

typedef struct lmo_radial_medium {
    const double* xgrid;
    const double* tgrid;
    const double* T;       /* (nx, nt) */
    const double* v;       /* (nx, nt) */
    const double* v2;      /* (nx, nt) or NULL */
    int32_t nx, nt;
} lmo_radial_medium;

typedef struct lmo_grid_medium {
    double x0, dx, y0, dy;
    const double* tgrid;
    const double* T;       /* (nx, ny, nt) */
    const double* vx;
    const double* vy;
    int32_t nx, ny, nt, _pad;
} lmo_grid_medium;

typedef struct lmo_table {     /* uniform-grid linear time table (transport.jl:231-253) */
    const double* vals;
    double Tmin, invdT;
    int32_t n, _pad;
} lmo_table;

typedef struct lmo_jtable {    /* inverse-CDF Jüttner table (transport.jl:268-270) */
    const double* vals;
    double Tmin, invdT;
    int32_t nU, nT;
} lmo_jtable;

typedef struct lmo_species {
    int32_t nspecies, _pad;
    int64_t counts[LMH_MAX_SPECIES];
    double  mass[LMH_MAX_SPECIES];
    lmo_table  drag[LMH_MAX_SPECIES];
    lmo_table  rta[LMH_MAX_SPECIES];
    lmo_jtable jt[LMH_MAX_SPECIES];
} lmo_species;

/* ---- threads of the per-particle loops ------------------------------------------------------ */
void lmo_set_threads(int32_t n);
int32_t lmo_get_threads(void);

/* ---- scalar building blocks ---------------------------------------------------------------- */
double lmo_fmGeV(void);
int64_t lmo_searchsortedlast(const double* v, int64_t n, double x);
double lmo_interpolate_2d(const double* x, int32_t nx, const double* y, int32_t ny,
                          const double* values, double xi, double yi);
double lmo_interpolate_2d_gc(const double* x, int32_t nx, const double* y, int32_t ny,
                             const double* values, double xi, double yi);
double lmo_eval_time_table(double T, double Tmin, double invdT, const double* vals, int32_t n);
double lmo_tau_drag(double T, double m, double DsT);
double lmo_effective_DsT(double T, double DsT, int32_t linear, double slope, double offset,
                         double Tfo, int32_t quad, double Tref, int32_t* err);
double lmo_besselkx(double nu, double z);
double lmo_tau_n_main3(double T, double m, double DsT);
int lmo_build_time_spline(int32_t which, double m, double DsT, double Tmin, double Tmax,
                          int32_t nT, int32_t linear, double slope, double offset, double Tfo,
                          int32_t quad, double Tref, double scale,
                          double* Tmin_out, double* invdT_out, double* vals);
int lmo_build_juttner_invcdf(double m, int32_t d, double Tmin, double Tmax, int32_t nT,
                             int32_t nU, int32_t np, double quantile_cap,
                             double* invcdf, double* Tmin_out, double* invdT_out);
double lmo_juttner_invcdf_lookup(const double* invcdf, int32_t nU, int32_t nT,
                                 double Ttab_min, double Ttab_invdT, double u, double T);
int lmo_draw_juttner_rejection(double m, double T, int32_t d, uint64_t seed, int64_t n,
                               double* out);

/* ---- Philox4x32-10 and the noise record ------------------------------------------------------ */
void lmo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void lmo_noise_block(uint64_t seed, uint64_t stream, int64_t i, int64_t step, int32_t block,
                     uint64_t w[4]);
void lmo_noise_transforms(double u, double x, double out[4]);
int32_t lmo_noise_channels(const lmh_params* p, int32_t dims, int32_t pdim);
void lmo_noise_record(const lmh_params* p, int32_t dims, int32_t pdim, int64_t N, int64_t step,
                      double* rec);

/* ---- initial-state samplers (utils.jl:33-172, 499-558) with the draws of block 3 -------------- */
int lmo_sample_fonll(const double* rg, const double* pg, const double* cdf_r, const double* n_r,
                     const double* cdf_p, const double* p_ok, int32_t nr, int32_t np, int32_t cartesian,
                     double rmax, double fmax, int64_t N, int32_t antithetic, uint64_t seed,
                     uint64_t stream, double* x_out, double* p_out);
void lmo_append_thermal_pz(const double* mom, const double* pos, int32_t dims, int64_t N, double m,
                           double t0, int32_t antithetic, uint64_t seed, uint64_t stream,
                           const lmo_radial_medium* med, double* out);

/* ---- Cartesian step on the radial (r, t) medium: kernels_cpu.jl + simulate_cpu.jl ------------ */
void lmo_kernel_boost_to_rest(double* mom, const double* pos, const lmo_radial_medium* med,
                              double m, int64_t N, int64_t step, double dt, double t0,
                              int32_t dims, int32_t pdim, int32_t radial_mode, double psi2,
                              int32_t relativistic);
void lmo_kernel_boost_to_lab(double* mom, const double* pos, const lmo_radial_medium* med,
                             double m, int64_t N, int64_t step, double dt, double t0,
                             int32_t dims, int32_t pdim, int32_t radial_mode, double psi2,
                             int32_t relativistic);
int lmo_boost_to_lab_all(const lmh_params* p, const lmo_radial_medium* med,
                         const lmo_species* sp, int32_t dims, int32_t pdim,
                         double* pos, double* mom, int64_t N, int64_t step);
int lmo_run_cartesian(const lmh_params* p, const lmo_radial_medium* med, const lmo_species* sp,
                      int32_t dims, int32_t pdim, double* pos, double* mom, int64_t N,
                      int64_t step_begin, int64_t nsteps, const double* noise,
                      int64_t* bad_step);

/* Milne coordinates (extension, coords LMH_MILNE): lmo_run_cartesian and lmo_boost_to_lab_all
 * with dims == pdim == 3, position rows (x, y, eta_s); the step is the Bjorken-mode Cartesian
 * step on (x, y, p) followed by eta_s += lmo_milne_deta(p_x, p_y, p_z, m, tau_a, dt) */
double lmo_milne_deta(double px, double py, double pz, double m, double tau_a, double dt);

/* ---- general-coordinate (tau, r) step: kernels_cpu_GC.jl + simulate_cpu_general_coords.jl ---- */
int lmo_run_gc(const lmh_params* p, const lmo_radial_medium* med, const lmo_species* sp,
               double* pos, double* mom, int64_t N, int64_t step_begin, int64_t nsteps,
               const double* noise, int64_t* bad_step);

/* ---- 2+1D grid medium (extension) ------------------------------------------------------------ */
void lmo_grid_sample(const lmo_grid_medium* g, double x, double y, double t, double out[3]);
int lmo_run_grid(const lmh_params* p, const lmo_grid_medium* g, const lmo_species* sp,
                 int32_t pdim, double* pos, double* mom, int64_t N, int64_t step_begin,
                 int64_t nsteps, const double* noise, int64_t* bad_step);
/* the driver's initial boost to the lab (simulate_cpu.jl:160-163) on the grid: every quark's
 * momentum boosted along the flow sampled at its (x, y) and t = step*dt + t0 */
int lmo_boost_to_lab_grid(const lmh_params* p, const lmo_grid_medium* g, const lmo_species* sp,
                          int32_t pdim, const double* pos, double* mom, int64_t N, int64_t step);

/* ---- on-the-fly freeze-out capture ------------------------------------------------------------ */
void lmo_freezeout_capture(const lmo_radial_medium* rmed, const lmo_grid_medium* gmed,
                           const double* pos, const double* mom, int32_t dims, int32_t pdim,
                           int64_t N, int64_t step, double dt, double t0, double Tfo,
                           int32_t interp, double* fo_pos, double* fo_mom, double* fo_tau,
                           double* fo_flag, double* last_pos, double* last_mom,
                           double* last_tau, double* last_T);

/* ---- observables ----------------------------------------------------------------------------- */
int lmo_observables(const double* mom, int32_t pdim, int64_t i0, int64_t i1,
                    const lmh_obs_spec* spec, double* out);

#ifdef __cplusplus
}
#endif
#endif
