/*
 * tm_oracle.h -- CPU oracle for the Dornic / Rietkerk stochastic step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, never the product: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (trophic_maelstorm_amd/libtm.so) neither links nor calls it.
 *
 * A plain-C restatement of the reference's single-species hot path
 * (paths relative to the reference's simulations/ directory):
 *   Rietkerk  Rietkerk_FPE/rietkerk_bjork_FKE.cpp:3842-3865 (constants),
 *             :4154-4211 (Dornic P step), :2363-2444 (f_2Dor + RK4),
 *             :4352-4465 (checks, flush, copy), :4017-4031 + :1238-1280 (rho(t)).
 *   DP        Percolation_FD/multiSPDP.cpp:2480-2504 (constants),
 *             :2901-2958 (Dornic rho step), :1647-1664 + :1821-1898 (RK4 of -b rho^2),
 *             :3231-3262 (flush, copy); exact update stochasticSPDP.cpp:1064-1068.
 * fed by counter-based Philox4x32-10 uniforms instead of the reference's
 * std::mt19937_64 stream (FKE:3896-3899), with samplers written from the
 * published methods named at each function.  Every integer decision uses only
 * IEEE-exact operations (+ - * / sqrt floor) and this file's own exp/log, so a
 * GPU build that performs the same operations in the same order reproduces the
 * draws bit-for-bit.  Build with -ffp-contract=off.
 *
 * Parity pins (tests/test_oracle_*.py): Philox against the Random123
 * known-answer vectors and rocRAND's host Philox; exp/log against numpy; the
 * normal quantile against scipy.special.ndtri; Poisson/Gamma samplers by
 * exact pmf chi-square and moment tests; the Dornic step by the closed-form
 * CIR transition moments; the Laplacian by its Fourier eigenvalues; RK4 by the
 * Rietkerk mean-field fixed points and the P=0 closed form; the DP nonlinear
 * sub-steps by the closed form rho/(1+b rho t); the whole DP step by ensemble
 * comparison with the reference's own percolationDornic_2D built from source;
 * the replicate statistics against the reference's own
 * var_mean_incremental_surv_runs (stochasticSPDP.cpp:173-207) built from
 * source, and against numpy means and unbiased variances; the consumer RK4
 * (tmo_fke_rk4) by the single-species RK4 bit for bit with no grazers, the
 * two-species form bit for bit with no predators, and the ODE (solve_ivp) on
 * uniform fields; the vegetation Dornic step (tmo_fke_veg_dornic) by the
 * single-species step with no plant growth; the separable movement by the
 * scatter restatement (summation order apart), mass conservation and its
 * independence of S; the movement statistics by numpy's mean and variance.
 */
#ifndef TM_ORACLE_H
#define TM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- parameter packs (mirror the product's C-ABI structs field for field) ---- */
typedef struct tmo_dp_params {
    double a, b, D, sigma, dt, dx;
    double a_end;          /* quench target; used when ramp_steps > 0 */
    int64_t ramp_steps;    /* a(n) = a + (a_end - a) * min(n, ramp_steps) / ramp_steps */
    double flush_eps;      /* multiSPDP.cpp:2426  1e-12 */
    double flush_neg;      /* multiSPDP.cpp:3257  -1e-80 */
    int32_t scheme;        /* 0 Euler, 1 RK4 (multiSPDP.cpp:1821), 2 exact (stochasticSPDP.cpp:1066) */
    int32_t sampler;       /* 0 exact Poisson/Gamma, 1 reference Gaussian shortcuts above 100 */
} tmo_dp_params;

typedef struct tmo_rk_params {
    double R, c, gmax, alpha, rW, W0, d;
    double D[3];           /* Dp, Dw, Do */
    double K[3];           /* k0 (unused by the 1-species RHS), k1, k2 */
    double sigma;          /* sigma_P */
    double dt, dx;
    double flush_eps;      /* rietkerk_bjork_basic.h:119  1e-10 */
    int32_t sampler;
    int32_t pad_;
} tmo_rk_params;

/* Vegetation + consumers, the multi-species DP step (multiSPDP.cpp with SPB = 2 or 3,
 * MultiSPDP.h:344-346 argument list).  Species 0 is the vegetation, 1 the grazer,
 * 2 the predator.  b is the vegetation's self-limitation in the trophic RK4; the
 * reference's call (multiSPDP.cpp:3222) passes a there, so pass b = a to reproduce it. */
typedef struct tmo_dpc_params {
    double a, b;
    double D[3], sigma[3];
    double v[3];           /* consumer speeds; v[0] unused */
    double M[3];           /* consumer mortality; M[0] unused */
    double E[3];           /* consumer efficiency; E[0] unused */
    double A[3][3];        /* attack rates; A[0][1] grazer on vegetation, A[1][2] predator on grazer */
    double H[3][3];        /* handling times, same indices */
    double pR[3];          /* perception radius, length units; pR[0] = 0 */
    double dt, dx;
    double a_end;
    int64_t ramp_steps;
    double flush_eps, flush_neg;
    int32_t sampler;
    int32_t pad_;
} tmo_dpc_params;

/* ---- primitives exposed for the oracle's own pins ---- */
void tmo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double tmo_exp(double x);
double tmo_log(double x);
double tmo_norm_icdf(double p);
double tmo_logfact(int64_t k);

/* Draw n samples from one stream (key = (seed, inst), counter = (cell, step, draw)).
 * kind: 0 uniform, 1 normal, 2 poisson(param, sampler mode = mode), 3 gamma(param). */
void tmo_sample(int kind, double param, int mode, uint32_t seed, uint32_t inst,
                uint32_t cell, int64_t step, int64_t n, double* out);

/* Dornic constants (FKE:3847-3849, MSP:2485-2486): out = {beta, exp(beta dt), lambda} */
void tmo_dornic_constants(double linear, double Dcoef, double sigma, double dt, double dx, double out[3]);

/* One Dornic transition for a single cell with given alpha_i and rho (moment pins). */
double tmo_dornic_cell(double alpha_i, double rho, double lambda, double lambda_exp, double sigma,
                       int sampler, uint32_t seed, uint32_t inst, uint32_t cell, int64_t step,
                       int64_t* n_out);

/* n independent transitions (cells 0..n-1 of step `step`) of one cell state. */
void tmo_dornic_batch(double alpha_i, double rho, double lambda, double lambda_exp, double sigma,
                      int sampler, uint32_t seed, uint32_t inst, int64_t step, int64_t n,
                      double* out, int64_t* n_out);

/* i.i.d. Uniform(lo[f], hi[f]) initial fields, species-major [nf][L*L]. */
void tmo_init_uniform(int L, int nf, uint32_t seed, uint32_t inst,
                      const double* lo, const double* hi, double* out);

/* Run nsteps steps in place on host fields (species-major SoA, FKE:3902).
 * rows (nullable): [nsteps][2*nf] = per field (#cells > 0, spatial mean) after each step.
 * draws (nullable): [L*L] Poisson integer of the LAST step, -1 where the cell was skipped.
 * Returns 0, or 1 if a negative / non-finite value was seen (first one in *fault_cell). */
int tmo_dp_run(const tmo_dp_params* p, int L, uint32_t seed, uint32_t inst, int64_t step0,
               int64_t nsteps, double* rho, double* rows, int64_t* draws, int64_t* fault_cell);
int tmo_rk_run(const tmo_rk_params* p, int L, uint32_t seed, uint32_t inst, int64_t step0,
               int64_t nsteps, double* fields, double* rows, int64_t* draws, int64_t* fault_cell);

/* Deterministic pieces exposed for closed-form pins. */
void tmo_laplacian(int L, const double* in, double* out);            /* 5-point periodic, FKE:1634-1653 */
void tmo_rk_rhs(const tmo_rk_params* p, int L, const double* y, double* f); /* f_2Dor, FKE:2363-2390 */
void tmo_rk4_deterministic(const tmo_rk_params* p, int L, int64_t nsteps, double* y); /* FKE:2393-2429 */
double tmo_dp_nonlinear(int scheme, double b, double dt, double rho);  /* one nonlinear sub-step */
void tmo_reduce(int L, int nf, const double* fields, double* row);    /* (#>0, mean) per field */
double tmo_flush(double x, double eps, double neg);                   /* post-step flush rule */

/* Speckle frame (FKE:409-431 rule): u from Philox(seed, inst; cell, 0xFFFFFFFE,
 * 0xFFFFFFFF, 0) word 0; field f = high[f] where u <= perc, else low[f]. */
void tmo_init_speckle(int L, int nf, uint32_t seed, uint32_t inst, double perc,
                      const double* high, const double* low, double* out);
/* Gaussian frame (FKE:1089-1199): fields f < n_gauss Gaussian bumps at the
 * corner centres with sd = L / sd_div[min(f, n_sd-1)], the rest amp[f]. */
void tmo_init_gauss(int L, int nf, int n_gauss, const double* sd_div, int n_sd,
                    const double* amp, double* out);

/* Across-replicate statistics of a sweep (FKE:4578-4602 first replicate,
 * FKE:1319-1369 / MSP:1080-1125 later ones).  series: [n_inst][n_meas][2*nf]
 * rows (#>0, mean per field); params_index[n_inst] groups the instances, in
 * instance order = replicate order.  flags: 1 = integer-division variance
 * term (FKE:1351), 2 = rows after a replicate's first extinct measurement read
 * as zero (FKE:3086,3135).  out: [n_params][n_meas][nf+3] = means of fields
 * 1..nf-1 over all replicates, then mean, variance, surviving count and mean
 * occupied count of field 0 over the surviving replicates. */
void tmo_replicate_stats(const double* series, int n_inst, int n_meas, int nf,
                         const int32_t* params_index, int n_params, int flags, double* out);

/* ---- consumer species (multiSPDP.cpp, SPB = 2, 3) ---- */
/* sin and cos of 2 pi u (quarter-turn reduction + Taylor polynomials); the
 * direction of a consumer's advection draw (MSP:3100). */
void tmo_sincos2pi(double u, double* s, double* c);
/* The fixed-order sum that stands for mean_of_vector's sequential sum where the
 * means steer the step (chunks of 4096, 256 lanes, pairwise lane fold). */
double tmo_fixed_sum(const double* x, int64_t n);
/* RK4 of the trophic terms at one cell, y[ns] in place (MSP:1670-1697, :1789-1819, :1821-1898). */
void tmo_dpc_rk4(const tmo_dpc_params* p, int ns, double* y);
/* Offsets (row, column) of the sites within radius R of a site, in the order of
 * computeNeighboringSitesCentral (MSP:1442-1461): x-major over [-R, R]^2, then
 * std::sort by squared distance.  out[2 k], out[2 k + 1]; returns the count, or
 * -1 if cap is too small (perception.cpp). */
int tmo_perception_offsets(int R, int32_t* out, int cap);
/* nsteps of the ns-species step on host fields [2 ns - 1][L*L] = rho_0 .. rho_{ns-1},
 * gamma_1 .. gamma_{ns-1}.  disk: perception offsets of R = (int)(max pR / dx),
 * disk_len of them.  Consumer s draws from the stream of counter block s << 24:
 * word 0 is its advection direction, the rest its Dornic sampler.  rows as
 * tmo_dp_run, over all 2 ns - 1 fields.  Returns 1 on a negative or non-finite
 * density (the smallest species-major index s L^2 + cell in *fault_index). */
int tmo_dpc_run(const tmo_dpc_params* p, int ns, int L, const int32_t* disk, int disk_len,
                uint32_t seed, uint32_t inst, int64_t step0, int64_t nsteps, double* fields,
                double* rows, int64_t* fault_index);

/* Multi-species replicate statistics (multiSPDP.cpp:3376-3391, :1080-1123):
 * out [n_params][n_meas][4 nsp] = per species (mean, variance, surviving count,
 * mean occupied count) over the replicates it still occupies; flags as above,
 * flag 2 ending a replicate at the vegetation's (field 0's) first empty row. */
void tmo_species_stats(const double* series, int n_inst, int n_meas, int nf, int nsp,
                       const int32_t* params_index, int n_params, int flags, double* out);

/* Rietkerk consumers' movement (advdiff_FKE_MultiSp, FKE:3614-3695): the
 * per-species constants of set_global_FKE_params (FKE:195-250) with the stencils
 * of precompute_gaussian (FKE:1735-1749; returns -1 if a stencil exceeds its cap),
 * and the scatter of one species' field.  out is overwritten. */
int tmo_fke_move_constants(double D, double v, double sigma, double dt_analytic, double dt_adv,
                           int mv_invariance, double dx, int CI, int32_t* rD, int32_t* rV,
                           double* mu_pref, double* stencil_D, int cap_D, double* stencil_V, int cap_V);
void tmo_fke_movement(int L, const double* rho, const double* gamma, const double* vx, const double* vy,
                      int rD, int rV, double mu_pref, const double* SD, const double* SV, double* out);

/* Perception gamma and velocities of the Rietkerk consumers (FKE:2534-2606,
 * :3371-3482); gamma[ns-1][L*L] updated in place, veff[ns][L*L][2] drawn with
 * three species. */
void tmo_fke_perception(int ns, int L, const double* pR, double dx, const double* v, const int32_t* disk,
                        int disk_len, uint32_t seed, uint32_t inst, int64_t step, const double* rho,
                        double* gamma, double* veff);

/* The Rietkerk consumer model's RK4 (RK4_Integrate_Stochastic_2Sp, FKE:2476-2533,
 * and _MultiSp, :2652-2720, with f_DorFKE_2Sp, :2449-2474, and f_DorFKE_3Sp,
 * :2611-2650): one step of y[ns + 3][L*L] = P, G, (Pr,) W, O in place; returns the
 * number of values left negative or non-finite.  dt2 = dt_rk / 2 and
 * dt6 = dt_rk / 6 as set_global_FKE_params sets them (FKE:206-210); dt is the
 * RK4's own step (its third stage). */
typedef struct tmo_fke_rk_params {
    double a, c, gmax, alpha, rW, W0;
    double K[3];
    double A[3][3], H[3][3], E[3], M[3];
    double D_W, D_O;
    double dt, dt_rk, dx;
} tmo_fke_rk_params;
int64_t tmo_fke_rk4(const tmo_fke_rk_params* p, int ns, int L, double* y);

/* The Rietkerk consumer model's vegetation Dornic step (FKE:4150-4211): Pstar
 * from the old vegetation P, with the single-species step's constants and
 * transition (p->d is the vegetation's mortality M[0]).  No flush. */
void tmo_fke_veg_dornic(const tmo_rk_params* p, int L, uint32_t seed, uint32_t inst, int64_t step,
                        const double* P, double* Pstar);

/* One species' movement step in separable form.  With each source's terms
 * (c.x, c.y, c.z) = (gamma^2 rho, (gamma - gamma^2) rho, (1 - gamma)^2 rho) and its
 * advected mean (mx, my), as tmo_fke_movement forms them:
 *   cx(p, q) = c.x of source (p, q) (0 where rho < 1e-10);
 *   Y1(p, q) = sum of c.y over the sources of row p with my = q;
 *   Y2(p, q) = sum of c.y over the sources of column q with mx = p;
 *   Z(p, q)  = sum of c.z over the sources with (mx, my) = (p, q);
 * each over the (2S + 1)^2 sources around (p, q), rows then columns ascending
 * (S >= every source's shift |m - c|).  Then along each row
 *   P = sum_b SD|b| cx(p, q - b) + sum_b SV|b| Y1(p, q - b),
 *   Q = sum_b SD|b| Y2(p, q - b) + sum_b SV|b| Z(p, q - b),
 * and along each column out(x, y) = sum_a SD|a| P(x - a, y) + sum_a SV|a| Q(x - a, y),
 * taps from -r to r, indices mod L.  The same four terms as tmo_fke_movement (D x D,
 * D x V, V x D, V x V), factored by rows and columns and summed in this fixed
 * order.  Needs 2 max(rD, rV + S) + 1 <= L. */
void tmo_fke_movement_sep(int L, const double* rho, const double* gamma, const double* vx, const double* vy,
                          int rD, int rV, int S, double mu_pref, const double* SD, const double* SV, double* out);

/* Movement statistics of one consumer (Rho_Mov, FKE:4034-4060): out[0] mean and
 * out[1] population variance of gamma, out[2], out[3] means of vx and vy, each
 * sum by tmo_fixed_sum and divided by L^2; the variance sums (g - mean)^2. */
void tmo_fke_mov_stats(int L, const double* gamma, const double* vx, const double* vy, double out[4]);

#ifdef __cplusplus
}
#endif
#endif
