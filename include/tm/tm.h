/*
 * tm.h -- C ABI of libtm, the MI355X (gfx950) Dornic / Rietkerk step library.
 *
 * This is the drop-in boundary for the reference's single-species hot loop:
 * the body of the time loop of
 *   rietkerk_DorFPE_2D_MultiSp   Rietkerk_FPE/rietkerk_bjork_FKE.cpp:3697 (loop body :4005-4468)
 *   dP_Dornic_2D_MultiSp         Percolation_FD/multiSPDP.cpp:2421        (loop body :2640-3265)
 * (paths relative to the reference's simulations/ directory).  The reference
 * has no FFI for this path -- the step is inline C++ -- so each entry point
 * below names the reference statements it replaces.  A caller keeps the
 * reference's drivers (argv contract Rietkerk_FPE/order_1stoc_test_rietkerk.cpp:99-110,
 * sweep rietkerk_bjork_FKE.cpp:4665) and replaces the loop body with
 * tm_run() blocks between measurement times.
 *
 * Conventions: plain C types only; host arrays are FP64, species-major SoA
 * [field][L*L] row-major (i = p*L + q), as the reference's
 * std::vector<std::vector<double>> fields (rietkerk_bjork_FKE.cpp:3902-3904).
 * Every function returns tm_status; nothing throws across the ABI.  A context
 * owns its device memory and one HIP stream and is used from one host thread
 * at a time (the reference's one-OpenMP-thread-per-parameter usage maps to one
 * context per thread, or better, to one batched context).
 */
#ifndef TM_TM_H
#define TM_TM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum tm_status {
    TM_OK = 0,
    TM_ERR_ARG = -1,       /* bad argument (sizes, null pointers, unknown enum) */
    TM_ERR_HIP = -2,       /* a HIP runtime call failed (see tm_last_error)     */
    TM_ERR_NOMEM = -3,     /* device or pinned allocation failed                */
    TM_ERR_STATE = -4,     /* call out of order (e.g. tm_run before configure)  */
    TM_ERR_NODEV = -5,     /* no usable gfx950 device                           */
    TM_ERR_TOO_LARGE = -6, /* beyond an allocation-free entry point's fixed scratch */
    TM_FAULT_NUMERIC = 1   /* a negative or non-finite value was produced; see  *
                            * tm_fault_info.  Replaces the reference's error-   *
                            * frame dump + exit(3) (rietkerk_bjork_FKE.cpp:4489-4560) */
} tm_status;

typedef enum tm_model {
    TM_MODEL_DP = 1,        /* one field rho                         (multiSPDP.cpp, Sp = 1) */
    TM_MODEL_DP2 = 2,       /* vegetation + grazer: fields rho_0, rho_1, gamma_1
                             * (multiSPDP.cpp built with SPB = 2, MultiSPDP_constants_2Sp.h:5) */
    TM_MODEL_RIETKERK = 3,  /* three fields P, W, O (rietkerk_bjork_constants_1Sp.h:5, Sp = 3) */
    TM_MODEL_DP3 = 4,       /* vegetation + grazer + predator: fields rho_0, rho_1, rho_2,
                             * gamma_1, gamma_2 (SPB = 3, MultiSPDP_constants_3Sp.h:5) */
    TM_MODEL_FKE2 = 5,      /* Rietkerk vegetation + grazer: fields P, G, W, O, gamma_1
                             * (rietkerk_bjork_FKE.cpp built with SPB = 2) */
    TM_MODEL_FKE3 = 6       /* Rietkerk vegetation + grazer + predator: fields P, G, Pr, W, O,
                             * gamma_1, gamma_2 (SPB = 3) */
} tm_model;

typedef enum tm_precision {
    TM_F64 = 0,   /* FP64 storage and arithmetic: bit-exact with the CPU restatement */
    TM_F32 = 1    /* FP32 storage and stepping; FP64 Poisson mean, ramp and rho(t) sums */
} tm_precision;

typedef enum tm_sampler {
    TM_SAMPLER_EXACT = 0,   /* exact Poisson (inversion / PTRS) and Gamma (Marsaglia-Tsang) */
    TM_SAMPLER_COMPAT = 1   /* the reference's Gaussian shortcuts above 100
                             * (rietkerk_bjork_FKE.cpp:4175-4191) */
} tm_sampler;

typedef enum tm_dp_scheme {
    TM_DP_EULER = 0,   /* rho <- rho_star - b rho_star^2 dt                           */
    TM_DP_RK4 = 1,     /* RK4 of -b rho^2 (multiSPDP.cpp:1821-1898, f at :1662).  The
                        * reference's call site multiSPDP.cpp:3222 passes a in b's slot;
                        * pass b = a to reproduce that run exactly.                  */
    TM_DP_EXACT = 2    /* rho_star / (1 + b rho_star dt) (stochasticSPDP.cpp:1064-1068) */
} tm_dp_scheme;

/* DP-Langevin parameters: dP_Dornic_2D_MultiSp(..., a, b, c, D[Sp], ..., sigma[], ..., dt, dx, ...)
 * (Percolation_FD/MultiSPDP.h:344-346).  The quench ramp is new: a(n) = a + (a_end - a) *
 * min(n, ramp_steps) / ramp_steps at instance step n (constant a when ramp_steps == 0). */
typedef struct tm_dp_params {
    double a, b, D, sigma, dt, dx;
    double a_end;
    int64_t ramp_steps;
    double flush_eps;   /* 1e-12 (multiSPDP.cpp:2426)  */
    double flush_neg;   /* -1e-80 (multiSPDP.cpp:3257) */
    int32_t scheme;     /* tm_dp_scheme */
    int32_t sampler;    /* tm_sampler   */
} tm_dp_params;

/* Vegetation + consumer parameters: dP_Dornic_2D_MultiSp(..., a, b, c, D[Sp], v[SpB], sigma[], ...,
 * A[SpB][SpB], H[SpB][SpB], E[SpB], M[SpB], pR[], dtV[SpB], ..., dt, dx, ...)
 * (Percolation_FD/MultiSPDP.h:344-346).  Species 0 is the vegetation, 1 the grazer, 2 the
 * predator (TM_MODEL_DP3 only).  a drives the vegetation's Dornic step (ramped as in
 * tm_dp_params); b is the vegetation's self-limitation in the trophic RK4 -- the reference's
 * call (multiSPDP.cpp:3222) passes a there, so pass b = a to reproduce its runs.  The
 * perception disk has radius R = (int)(max pR / dx) sites (1 <= R <= TM_DPC_MAX_RADIUS).
 * c and dtV have no effect on the reference's step and are not taken. */
#define TM_DPC_MAX_RADIUS 12
typedef struct tm_dpc_params {
    double a, b;
    double D[3], sigma[3];
    double v[3];        /* consumer speeds, length/time; v[0] unused    */
    double M[3];        /* consumer mortality; M[0] unused              */
    double E[3];        /* consumer efficiency; E[0] unused             */
    double A[3][3];     /* attack rates: A[0][1] grazer on vegetation, A[1][2] predator on grazer */
    double H[3][3];     /* handling times, same indices                 */
    double pR[3];       /* perception radius, length units; pR[0] = 0  */
    double dt, dx;
    double a_end;
    int64_t ramp_steps;
    double flush_eps;   /* 1e-12 (multiSPDP.cpp:2426)  */
    double flush_neg;   /* -1e-80 (multiSPDP.cpp:3257) */
    int32_t sampler;    /* tm_sampler */
    int32_t reserved;
} tm_dpc_params;

/* Rietkerk parameters: rietkerk_DorFPE_2D_MultiSp(..., a(=R), c, gmax, alpha, rW, W0, D[Sp], ...,
 * K[3], sigma[], ..., M[](d = M[0]), ..., dt, dx, ...) (Rietkerk_FPE/rietkerk_bjork_basic.h:450-452).
 * Units are whatever the caller uses (the model is unit-agnostic). */
typedef struct tm_rietkerk_params {
    double R, c, gmax, alpha, rW, W0, d;
    double D[3];        /* D_P, D_W, D_O                              */
    double K[3];        /* k0 (unused by the 1-species RHS), k1, k2   */
    double sigma;       /* sigma_P                                     */
    double dt, dx;
    double flush_eps;   /* 1e-10 (rietkerk_bjork_basic.h:119)          */
    int32_t sampler;    /* tm_sampler */
    int32_t reserved;
} tm_rietkerk_params;

/* Rietkerk vegetation + consumer parameters (TM_MODEL_FKE2 / TM_MODEL_FKE3):
 * rietkerk_DorFPE_2D_MultiSp(..., a, c, gmax, alpha, rW, W0, D[Sp], v[SpB], K[3], sigma[], ...,
 * A, H, E, M, pR[], ..., dt, dx, ...) (Rietkerk_FPE/rietkerk_bjork_FKE.cpp:3697-3699) with
 * set_global_FKE_params(D, v, dtAdv, Dt_Analytic, dx, L, CI) (:195-250).  Species 0 is the
 * vegetation P, 1 the grazer G, 2 the predator Pr (TM_MODEL_FKE3 only).  One step is the
 * loop body :4100-4462: perception gamma (and, with the predator, velocities) from the old
 * densities, the vegetation's Dornic step, the RK4 of the trophic and water terms, the
 * consumers' Fokker-Planck movement, checks and flush.  The reference's D[Sp] is split
 * into D_P (vegetation, its Dornic step), D_mov[s] (consumer s, its movement) and
 * D_W / D_O. */
typedef struct tm_fke_params {
    double a;                     /* rainfall (R)                                               */
    double c, gmax, alpha, rW, W0;
    double K[3];                  /* k0 (unused), k1, k2                                        */
    double A[3][3], H[3][3];      /* attack rates and handling times: [0][1] grazer on P,
                                   * [1][2] predator on G                                       */
    double E[3], M[3];            /* efficiencies (E[0] unused) and mortalities; M[0] is the
                                   * vegetation's d in its Dornic step (FKE:3847)               */
    double D_P, D_W, D_O;
    double D_mov[3];              /* consumers' movement diffusion; [0] unused                  */
    double v[3];                  /* speeds: movement and, with the predator, the velocity
                                   * draws of every species                                     */
    double sigma[3];              /* sigma[0]: vegetation noise; sigma[1], sigma[2]: width of
                                   * the consumers' movement stencils (FKE:3764-3772)           */
    double pR[3];                 /* perception radii, length units; pR[0] = 0 as the
                                   * reference's                                                */
    double dt;                    /* time step: the Dornic step and the RK4's third stage       */
    double dt_analytic;           /* Dt_Analytic: RK4 half and sixth steps (FKE:206-210) and the
                                   * movement kernels                                          */
    double dt_adv[3];             /* per consumer, with mv_invariance = 1                       */
    double dx;
    double flush_eps;             /* 1e-10 (rietkerk_bjork_basic.h:119)                         */
    int32_t mv_invariance;        /* 0: beta_vel = 3 / dt_analytic; 1: 3 / dt_adv[s]            */
    int32_t CI;                   /* stencil half-widths in sigmas, 4 in the reference          */
    int32_t sampler;              /* tm_sampler, of the vegetation's Dornic step                */
    int32_t reserved;
} tm_fke_params;

/* One lattice instance (a replicate at one parameter point).  Instances of
 * different L share one launch.  The Philox key is (seed, instance_id); the
 * counter is (cell, step, block), so a run is reproducible from (seed,
 * instance_id, step) alone -- the reference's random_device seeding
 * (rietkerk_bjork_FKE.cpp:3896) becomes a caller choice. */
typedef struct tm_instance_desc {
    int32_t L;             /* lattice side g                                  */
    int32_t params_index;  /* index into the params array of tm_configure     */
    uint32_t seed;
    uint32_t instance_id;
    int64_t step0;         /* instance step number of the next step to take   */
} tm_instance_desc;

typedef struct tm_ctx tm_ctx;

/* Context lifetime.  device: HIP ordinal.  Replaces the per-call cudaMalloc /
 * cudaFree of the reference's CUDA seam (rietkerk_bjork_FKE.cpp:3709-3712, :4636-4639). */
tm_status tm_create(tm_ctx** out, int device);
void tm_destroy(tm_ctx* ctx);
const char* tm_last_error(const tm_ctx* ctx);

/* Declare the batch: model, precision, params[n_params] (tm_dp_params,
 * tm_rietkerk_params, or tm_dpc_params for TM_MODEL_DP2 / TM_MODEL_DP3, which are
 * TM_F64 only), instances[n_inst].  Allocates the double-buffered fields (zeroed)
 * and the per-instance descriptor and tile tables.  Replaces the per-replicate
 * field allocation of rietkerk_bjork_FKE.cpp:3902-3913 (multiSPDP.cpp:2576-2585
 * with the perception gamma fields, :2581). */
tm_status tm_configure(tm_ctx* ctx, int model, int precision, const void* params, int n_params,
                       const tm_instance_desc* instances, int n_inst);

/* Host <-> device fields of one instance, FP64 [nfields][L*L] (converted for TM_F32).
 * Replace the init_* frame builders' output hand-off (rietkerk_bjork_FKE.cpp:3964-3971)
 * and give save_frame (rietkerk_bjork_FKE.cpp:2829-2876) its data. */
tm_status tm_upload_fields(tm_ctx* ctx, int inst, const double* host);
tm_status tm_download_fields(tm_ctx* ctx, int inst, double* host);

/* i.i.d. Uniform(lo[f], hi[f]) initial fields for every instance, drawn on the
 * device from Philox(seed, instance_id; cell, 0xFFFFFFFF, 0xFFFFFFFF, f / 4), word f % 4. */
tm_status tm_init_uniform(tm_ctx* ctx, const double* lo, const double* hi);

/* Speckle initial frames: every cell draws u ~ Uniform(0,1) from
 * Philox(seed, instance_id; cell, 0xFFFFFFFE, 0xFFFFFFFF, 0), word 0, and field f
 * takes high[p][f] where u <= perc[p], low[p][f] elsewhere, p = the instance's
 * params_index (perc[n_params], high/low [n_params][nfields]).  Replaces
 * init_randconstframe (rietkerk_bjork_FKE.cpp:409-431) and, with values from
 * tm_mft_frame_values or the caller, init_randbistableframe (:434-451),
 * init_exprtk_homogenousMFTframe (:453-531) and init_exprtk_randbiMFTframe
 * (:533-642). */
tm_status tm_init_speckle(tm_ctx* ctx, const double* perc, const double* high, const double* low);

/* Gaussian initial frames (init_gaussframe + init_gaussiantear,
 * rietkerk_bjork_FKE.cpp:1089-1199): fields f < n_gauss hold
 * amp[p][f] exp(-((i-cx)^2 + (j-cy)^2) / (2 sd sd)) at cell i*L + j, centred at
 * the corners (L/8 or 7L/8, integer division) counter-clockwise from the top
 * left, with sd = L / sd_div[f] (the reference's default is sd_div = {8, 16};
 * missing entries repeat the last); fields f >= n_gauss hold amp[p][f]. */
tm_status tm_init_gauss(tm_ctx* ctx, int n_gauss, const double* sd_div, int n_sd, const double* amp);

/* Host only.  Per-field (high, low) values of the reference's mean-field
 * frames at one parameter point: mft[2*nf] are the caller's mean-field values
 * at R (the reference's exprtk MFT_Vec_CoexExpr, evaluated), c_spread[2*nf] its
 * scale factors; the first nf entries apply below R_c, the last nf at or above.
 * homogeneous != 0: init_exprtk_homogenousMFTframe (rietkerk_bjork_FKE.cpp:453-531),
 * high == low.  Otherwise init_exprtk_randbiMFTframe (:533-642): field 0 is
 * max(dP, c_spread E_0) in the speckles.  Returns 0, or -1 on a bad argument. */
int tm_mft_frame_values(int homogeneous, double R, double R_c, double dP, int nf, const double* c_spread,
                        const double* mft, double* high, double* low);

/* Advance every instance n_steps steps: Dornic Poisson-Gamma on P (rho), RK4
 * (Rietkerk) or the DP nonlinear sub-step, checks, flush and copy -- the loop
 * body rietkerk_bjork_FKE.cpp:4154-4465 / multiSPDP.cpp:2901-3262.
 * rho_rows (nullable, host): [n_inst][n_steps][2*nfields] FP64 with, per field,
 * (#cells > 0, spatial mean) AFTER each step -- occupied_sites_of_vector and
 * mean_of_vector (rietkerk_bjork_FKE.cpp:1238-1280) as used at :4017-4031, every
 * step, with Kahan-compensated FP64 sums.  Returns TM_FAULT_NUMERIC if a
 * negative or non-finite value appeared (the step still completes). */
tm_status tm_run(tm_ctx* ctx, int64_t n_steps, double* rho_rows);

/* rho(t) row of the current state: [n_inst][2*nfields]. */
tm_status tm_measure(tm_ctx* ctx, double* rows);

/* TM_MODEL_FKE2 / TM_MODEL_FKE3: the movement statistics the reference logs at each
 * measurement time into Rho_Mov (rietkerk_bjork_FKE.cpp:4034-4060), of the current state:
 * out [n_inst][4 SpB] (SpB = 2 or 3 species); for each consumer s = 1 .. SpB-1, columns
 * 4s .. 4s+3 hold the spatial mean and variance of gamma_s and the spatial means of its
 * velocity components vx, vy in the last step (columns 0-3 are zero, as there).  The
 * velocities are those the last step drew (zero for TM_MODEL_FKE2, which draws none, and
 * before the first step).  The variance is the population variance,
 * sum (gamma - mean)^2 / L^2 (variance_of_vector, :1248-1257), in FP64 with the
 * consumer models' fixed-order sums; the reference accumulates that sum in float. */
tm_status tm_mov_stats(tm_ctx* ctx, double* out);

/* Debug aid for parity tests: when enabled, tm_run records the Poisson integer
 * of every cell of the last step (-1 where the absorbing-state skip applied,
 * rietkerk_bjork_FKE.cpp:4161-4164); tm_download_draws returns them [L*L]. */
tm_status tm_set_record_draws(tm_ctx* ctx, int enable);
tm_status tm_download_draws(tm_ctx* ctx, int inst, int64_t* host);

/* First numeric fault since configure: instance, cell, step; all -1 if none. */
tm_status tm_fault_info(tm_ctx* ctx, int32_t* inst, int64_t* cell, int64_t* step);

/* Current instance step numbers (step0 + steps taken). */
tm_status tm_get_step(tm_ctx* ctx, int inst, int64_t* step);

/* Rewind every instance to its step0 and clear the fault record; fields are
 * left as they are (call tm_init_uniform / tm_upload_fields next).  Launch
 * shape and tuning are kept.  Starts a fresh replicate of the batch,
 * the role of the replicate loop head rietkerk_bjork_FKE.cpp:3887-3899. */
tm_status tm_reset(tm_ctx* ctx);

/* Launch shape.  variant: index into the kernel table for the configured
 * model/precision (-1 = default); graph_steps: steps per launch block (the
 * per-step constants, the step launches, the rho(t) fold and the counter bump,
 * enqueued together; 0 = blocks of at most 64 steps), rounded up to a whole
 * number of launches for a variant that advances several steps per launch.
 * (The name is kept from when blocks were replayed as hipGraphs.)
 * tm_variant_count / tm_variant_name describe the table. */
tm_status tm_set_launch(tm_ctx* ctx, int variant, int graph_steps);
/* The launch shape in use: variant index and block length (either nullable). */
tm_status tm_get_launch(const tm_ctx* ctx, int* variant, int* graph_steps);
int tm_variant_count(int model, int precision);
const char* tm_variant_name(int model, int precision, int variant);

/* Autotuner: time each kernel variant (and pick a block length) on the configured
 * batch and its current fields without advancing it, and pick the fastest; the
 * variant in place when tuning starts is kept unless another is at least 5%
 * faster.  The choice persists in cache_path (NULL = $TM_TUNE_CACHE or
 * ~/.cache/trophic_maelstorm_amd/tune.txt), keyed by (device arch, model,
 * precision, batch shape); a cached choice is reused without timing.  Tune on a
 * representative state (e.g. the initial fields of the run).
 * out_variant / out_graph_steps (nullable) get the choice. */
tm_status tm_autotune(tm_ctx* ctx, const char* cache_path, int* out_variant, int* out_graph_steps);

/* Timing aid: runs n_steps steps (as tm_run, rows discarded) and reports the
 * device time between HIP events recorded on the context's stream around the
 * step kernels only (ms_total), and the number of step-kernel launches. */
tm_status tm_time_steps(tm_ctx* ctx, int64_t n_steps, double* ms_total, int64_t* launches);

/* Device time of the last tm_run: elapsed time between HIP events recorded on
 * the context's stream before its first launch and after its last (step
 * kernels, per-block constants, rho(t) fold and counter bump; the rows copy to
 * the host is outside).  n_steps (nullable) gets that call's step count. */
tm_status tm_last_run_time(const tm_ctx* ctx, double* ms, int64_t* n_steps);

/* Steps that tm_run / tm_run_series advanced since tm_configure by single-step
 * launches and by multi-step (temporally blocked) launches of the DP variants
 * named _xK / _xKa.  A timing aid (e.g. to weigh per-launch traffic). */
tm_status tm_mode_steps(const tm_ctx* ctx, int64_t* single_steps, int64_t* blocked_steps);

/* The context's HIP stream (hipStream_t) for callers that order their own work. */
void* tm_stream(tm_ctx* ctx);

/* ---- Output layer: CSV text, snapshots, burn-in input ----------------------
 * Numbers are printed as the reference's std::ostream prints them: a double
 * with precision p as printf("%.*g", p, x) -- p = 6 is the reference's default
 * (its FRAME and PRELIM files), 16 its sweep file -- and integers in decimal.
 * Host-only functions return 0 on success and a negative value on failure. */

/* FRAME file of one instance (save_frame, rietkerk_bjork_FKE.cpp:2829-2876):
 * header (may be NULL; the reference's is "a_c,  x,  P(x; t), W(x; t), O(x; t) \n",
 * rietkerk_bjork_constants_1Sp.h:6), then per cell i "a,i,f0,...\n" with `digits`
 * significant digits (6 = the reference's). */
tm_status tm_write_frame_csv(tm_ctx* ctx, int inst, const char* path, double a, const char* header, int digits);

/* Host only: the same file from a host frame [nf][L*L]. */
int tm_format_frame_csv(const char* path, const double* fields, int nf, int L, double a, const char* header,
                        int digits);

/* Host only: PRELIM file (save_prelimframe, rietkerk_bjork_FKE.cpp:2723-2765):
 * header, then per row i "a,r+1,g,t[i]," and each of stats[i][0..n_cols-1]
 * followed by a comma -- e.g. the rows of tm_replicate_stats. */
int tm_write_prelim_csv(const char* path, const char* header, double a, int r, int g, const double* t,
                        const double* stats, int n_rows, int n_cols);

/* Host only: the sweep's final file (rietkerk_bjork_FKE.cpp:4852-4886): the
 * header line, then rows[n_rows][n_cols] comma-separated with precision 16. */
int tm_write_sweep_csv(const char* path, const char* header, const double* rows, int n_rows, int n_cols);

/* Binary snapshot of one instance: a 64-byte header ("TMSNAP1", model,
 * precision, nf, L, seed, instance_id, step, params_index, data bytes) and the
 * fields in their stored precision, species-major.  A fast stand-in for the
 * FRAME text at large L; tm_snapshot_to_frame_csv converts it afterwards. */
tm_status tm_save_snapshot(tm_ctx* ctx, int inst, const char* path);

/* Load a snapshot into instance inst (same model and L; the precision is
 * converted as tm_upload_fields does).  step (nullable) gets the recorded
 * instance step, for a resumed run's step0. */
tm_status tm_load_snapshot(tm_ctx* ctx, int inst, const char* path, int64_t* step);

/* Host only: a snapshot as the FRAME file tm_write_frame_csv would write. */
int tm_snapshot_to_frame_csv(const char* snap, const char* csv, double a, const char* header, int digits);

/* Host only: burn-in input (init_exprtk_readCSVcolumns_frame,
 * rietkerk_bjork_FKE.cpp:644-935).  Reads a FRAME file: the header names with
 * white space removed select the columns cols[n_cols]; "W(x;t)" and "O(x;t)"
 * go to fields nf-2 and nf-1, any other column to field (its position - 2).
 * Cells that do not parse as numbers read as 0; fields not named are left as
 * they are in out[nf][L*L].  Returns the number of rows read (<= L*L), -2 if a
 * column is missing, -1 on other errors. */
int tm_read_frame_csv(const char* path, int L, int nf, const char* const* cols, int n_cols, double* out);

/* ---- Sweep measurement and replicate statistics ----------------------------
 * The reference's sweep (first_order_critical_exp_delta_stochastic_MultiSp,
 * rietkerk_bjork_FKE.cpp:4665-4887 / multiSPDP.cpp:3456-3702) runs one OpenMP
 * thread per parameter value and, inside it, r replicates one after another,
 * logging (#occupied, mean) per field at the times t_meas and folding each
 * replicate into running statistics over the replicates still alive.  Here the
 * whole sweep is one batch: instance k is replicate j of parameter point
 * params_index, where j counts the earlier instances with the same
 * params_index (instance order is replicate order). */

/* Host only (no device needed).  The measurement times of the reference:
 * logarithm10_time_bins(t_max, dt) (rietkerk_bjork_FKE.cpp:1497-1536,
 * multiSPDP.cpp:1251-1287): 10^(0.04 k) rounded down to the dt grid, kept above
 * 1, duplicates and points closer than dt/2 removed.  Writes at most cap values
 * to out (nullable) and returns how many there are (0 when t_max <= 1). */
int tm_log10_time_bins(double t_max, double dt, double* out, int cap);

/* ---- Rietkerk consumers' movement (the reference's CUDA seam) ----------------
 * advdiff_FKE_MultiSp (Rietkerk_FPE/rietkerk_bjork_FKE.cpp:3614-3695) and its CUDA
 * twin updateMovement (kernel_advdiff_FKE2D.cu:92-180) move each consumer species
 * one analytic Fokker-Planck step: every cell with rho >= 1e-10 spreads rho through
 * the product of a row and a column kernel, each gamma x (Gaussian centred on the
 * cell) + (1 - gamma) x (Gaussian centred on the advected mean cell + v mu_prefactor,
 * rounded), the two Gaussians being the stencils below. */
typedef struct tm_fke_move_species {
    int32_t rD;            /* sig_D_ScaledBounds[s]: diffusion stencil half-width, sites   */
    int32_t rV;            /* sig_vD_ScaledBounds[s]: advection stencil half-width, sites  */
    double mu_prefactor;   /* mu_vel_prefactor[s] (FKE:248)                                 */
} tm_fke_move_species;

/* Host only.  The constants set_global_FKE_params derives for one consumer species
 * (FKE:222-250: beta = 3 / dt_analytic, or 3 / dt_adv with mv_invariance == 1;
 * sigma_D = sqrt(2 D dt_analytic); the velocity OU spread; CI = 4 by default) and
 * its two stencils, precompute_gaussian(sigma, rD) and (sigma, rV) (FKE:1735-1749,
 * :3764-3772 -- the reference passes the species' noise sigma).  stencil_D gets
 * rD + 1 values, stencil_V rV + 1.  Returns 0, or -1 if a cap is too small. */
int tm_fke_move_constants(double D, double v, double sigma, double dt_analytic, double dt_adv, int mv_invariance,
                          double dx, int CI, tm_fke_move_species* out, double* stencil_D, int cap_D,
                          double* stencil_V, int cap_V);

/* The movement step of ns consumer species on one L x L lattice, on `device`:
 * host arrays rho, gamma [ns][L*L], v_eff [ns][L*L][2] (vx, vy as the reference's
 * std::pair), stencils concatenated by species (rD_s + 1 and rV_s + 1 values);
 * out [ns][L*L] receives the moved fields (it may alias rho).  Replaces the
 * reference's host copies, cudaMemset, launchAdvDiffKernel_MultiSp and copy back
 * (FKE:4276-4301, rietkerk_bjork_basic.h:461-467) and the CPU path (FKE:4303-4309).
 * The kernels factor the four terms by rows and columns and every cell pulls its
 * own sums, so no atomics are needed; stencils and advection shifts must fit the
 * lattice (2 max(rD, rV + max |v| mu_prefactor + 1) + 1 <= L).  Returns 0 or a
 * negative tm_status. */
int tm_fke_movement_host(int device, int ns, int L, const tm_fke_move_species* sp, const double* stencil_D,
                         const double* stencil_V, const double* rho, const double* gamma, const double* v_eff,
                         double* out);

/* The same step on device arrays (hipMalloc'd by the caller, same layouts), on
 * `stream` (a hipStream_t, or null): the drop-in for launchAdvDiffKernel_MultiSp
 * (kernel_advdiff_FKE2D.cu:224-236) with the constants passed in place of the
 * reference's __constant__ copies (copyToDeviceConstantMemory_AdvTerms, :20-86).
 * out is overwritten, so the reference's cudaMemset of Rho_tsar is not needed;
 * out must not alias rho.  A first small kernel finds the largest speed, which
 * the host reads back (one stream synchronisation) to size the shift reach.
 * Returns 0 or a negative tm_status. */
int tm_fke_movement(int ns, int L, const tm_fke_move_species* sp, const double* stencil_D, const double* stencil_V,
                    const double* rho, const double* gamma, const double* v_eff, double* out, void* stream);

/* Perception gamma and velocities of the Rietkerk model's consumers, one step:
 * calc_gamma_vel_2Sp_NonRefugia (rietkerk_bjork_FKE.cpp:2534-2606) for ns = 2 and
 * calc_gamma_vel_3Sp_NonRefugia (:3371-3482) for ns = 3, on `device`, host arrays.
 * rho [ns][L*L]: the old densities (vegetation, grazer, predator); gamma [ns-1][L*L]:
 * the consumers' previous gamma, updated in place (both forms add onto it);
 * v_eff [ns][L*L][2]: with ns = 3 every species draws v[s] (cos, sin)(2 pi u), u =
 * word 0 of Philox(seed, inst; cell, step, block (s << 24) | 0x800000); with ns = 2 it
 * is left as it is (the reference draws none).  pR [ns] and dx set the perception
 * radius as in tm_dpc_params; an extinct consumer gets gamma 1.  The spatial means
 * use the consumer models' fixed-order sum.  Returns 0 or a negative tm_status. */
int tm_fke_perception_host(int device, int ns, int L, const double* pR, double dx, const double* v, uint32_t seed,
                           uint32_t inst, int64_t step, const double* rho, double* gamma, double* v_eff);

/* Parameters of the Rietkerk consumer model's RK4 (f_DorFKE_2Sp / f_DorFKE_3Sp,
 * rietkerk_bjork_FKE.cpp:2449-2474, :2611-2650).  Names follow the reference's
 * arguments: a is the rainfall; K[1] and K[2] enter (K[0] does not); A, H, E, M
 * are indexed [prey][consumer] as there.  D_W and D_O are the diffusion constants
 * of W and O (the reference's Dxd2 = D / dx^2, FKE:3849-3852).  dt is the step the
 * RK4 receives, used by its third stage; dt_rk is Dt_Analytic, which sets the
 * half and sixth steps dt2 = dt_rk / 2 and dt6 = dt_rk / 6 (set_global_FKE_params,
 * FKE:206-210).  dt_rk = dt gives the textbook RK4. */
typedef struct tm_fke_rk_params {
    double a, c, gmax, alpha, rW, W0;
    double K[3];
    double A[3][3], H[3][3], E[3], M[3];
    double D_W, D_O;
    double dt, dt_rk, dx;
} tm_fke_rk_params;

/* One RK4 step of the Rietkerk consumer model's reaction and water terms, in
 * place: RK4_Integrate_Stochastic_2Sp (rietkerk_bjork_FKE.cpp:2476-2533) for
 * ns = 1 consumer and RK4_Integrate_Stochastic_MultiSp (:2652-2720) for ns = 2.
 * fields [ns + 3][L*L] in the reference's order: P, G, (Pr,) W, O; periodic
 * lattice.  Device arrays (hipMalloc'd), on `stream` (a hipStream_t, or null); the
 * library allocates its scratch (3 (ns + 3) L^2 doubles) and waits for the stream
 * before returning.  n_bad (may be null) gets the number of values that ended
 * negative or non-finite, which the reference logs and carries on from
 * (FKE:2511-2529).  Returns 0 or a negative tm_status. */
int tm_fke_rk4(int ns, int L, const tm_fke_rk_params* p, double* fields, int64_t* n_bad, void* stream);

/* The same step on host arrays, on `device`. */
int tm_fke_rk4_host(int device, int ns, int L, const tm_fke_rk_params* p, double* fields, int64_t* n_bad);

/* tm_fke_rk4 without allocating: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  The same kernels and result as
 * tm_fke_rk4; the stage buffers and the negative-value counter live in a fixed
 * library array on each GPU (3 x 5 x 256^2 doubles, 7.5 MiB, held until the
 * process ends), so L <= 256 (ns <= 2), else TM_ERR_TOO_LARGE.  Runs on the
 * current device, whose stream `stream` must be; waits for it before returning.
 * Calls that overlap on one GPU share the array and race. */
int tm_fke_rk4_noalloc(int ns, int L, const tm_fke_rk_params* p, double* fields, int64_t* n_bad, void* stream);

/* The Rietkerk consumer model's vegetation Dornic sub-step (rietkerk_bjork_FKE.cpp:
 * 4150-4211): Pstar [L*L] from the old vegetation P [L*L], periodic lattice, with
 * the constants of FKE:3847-3849 -- pass the vegetation's mortality M[0] as p->d,
 * and p->D[0], p->sigma, p->dt, p->dx, p->sampler; the other fields are unused.
 * The same transition, Philox stream (seed, inst; cell, step) and absorbing skip as
 * the single-species step's P Dornic.  No flush (the reference flushes after the
 * movement step, FKE:4453-4462).  Device arrays on `stream` (Pstar must not alias
 * P), or host arrays on `device`.  Returns 0 or a negative tm_status. */
int tm_fke_veg_dornic(int L, const tm_rietkerk_params* p, uint32_t seed, uint32_t inst, int64_t step,
                      const double* P, double* Pstar, void* stream);
int tm_fke_veg_dornic_host(int device, int L, const tm_rietkerk_params* p, uint32_t seed, uint32_t inst,
                           int64_t step, const double* P, double* Pstar);

/* Host only.  The perception neighbourhood of the consumer models: offsets
 * (row, column) of the sites within Euclidean radius R of a site, in the order of
 * computeNeighboringSitesCentral (multiSPDP.cpp:1442-1461): generated row-major
 * over [-R, R]^2 and std::sort-ed by squared distance alone, ties as that sort
 * leaves them.  out[2 k], out[2 k + 1]; returns the count, or -1 if it exceeds cap. */
int tm_perception_offsets(int R, int32_t* out, int cap);

/* Host only.  The step numbers at which the reference's time loop logs each
 * t_meas[i]: t accumulates t += dt from 0 and the state after n steps is logged
 * for the current index when t == 0 or t_meas[index] - dt/2 <= t < t_meas[index]
 * + dt/2 (rietkerk_bjork_FKE.cpp:4005-4017,4467 / multiSPDP.cpp:2640-2650,3264),
 * while t < t_meas[n_meas-1] + dt.  Returns how many points were reached (a
 * point off the dt grid stalls the rest, as in the reference). */
int tm_measure_steps(const double* t_meas, int n_meas, double dt, int64_t* steps);

/* Advance every instance to step steps[n_meas-1] (counted from now) as tm_run
 * would, and record each instance's rho(t) row (tm_run's row format) at each
 * steps[i] into a device series [n_inst][n_meas][2*nfields] kept by the context
 * for tm_replicate_stats; steps[i] == 0 records the current state.  steps must
 * be non-decreasing and >= 0.  series (nullable, host) gets a copy.  Replaces
 * the per-replicate Rho_M logging of rietkerk_bjork_FKE.cpp:4017-4031. */
tm_status tm_run_series(tm_ctx* ctx, const int64_t* steps, int n_meas, double* series);

/* Flags of tm_replicate_stats. */
#define TM_STATS_INTDIV_VARIANCE 1     /* the reference's (1/(r-1)) in integer arithmetic
                                        * (rietkerk_bjork_FKE.cpp:1351, multiSPDP.cpp:1112):
                                        * the old-variance term drops for r >= 3          */
#define TM_STATS_STOP_AT_EXTINCTION 2  /* a replicate's rows after its first measurement
                                        * with no occupied cell of field 0 read as zero: the
                                        * reference ends the replicate there
                                        * (rietkerk_bjork_FKE.cpp:3086,3135,4068-4072)    */
#define TM_STATS_REFERENCE 3           /* both: the reference's numbers                   */

/* Across-replicate statistics of the last tm_run_series, on the device, one
 * row per (parameter point, measurement): stats [n_params][n_meas][nfields+3] =
 *   means over all replicates of fields 1..nfields-1 (W, O),
 *   then for field 0: mean over surviving replicates of the spatial mean, its
 *   variance over them, the number of surviving replicates, and their mean
 *   number of occupied cells
 * -- rho_rep_avg_var columns 1..(nfields+2) of rietkerk_bjork_FKE.cpp:4578-4602
 * (first replicate) and var_mean_incremental_surv_runs (rietkerk_bjork_FKE.cpp:
 * 1319-1369, multiSPDP.cpp:1080-1125; later replicates, in order).  A replicate
 * survives at a measurement when field 0 is occupied there and at every earlier
 * one.  Without TM_STATS_INTDIV_VARIANCE the variance is the unbiased sample
 * variance.  Parameter points with no instance get zero rows.
 * TM_MODEL_DP2 / TM_MODEL_DP3: stats [n_params][n_meas][4 * species] = per
 * species (vegetation, grazer, predator) those four statistics of that species,
 * each over the replicates it still occupies (multiSPDP.cpp:3376-3391 first
 * replicate, :1080-1123 later ones); with TM_STATS_STOP_AT_EXTINCTION a
 * replicate's rows after the vegetation's first extinct measurement read as zero. */
tm_status tm_replicate_stats(tm_ctx* ctx, int flags, double* stats);

#ifdef __cplusplus
}
#endif
#endif /* TM_TM_H */
