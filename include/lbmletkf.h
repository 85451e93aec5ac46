/*
 * lbmletkf.h -- C-ABI of the MI355X-native LBM-LETKF assimilation cycle.
 *
 * One context holds a whole ensemble of M members as one structure-of-arrays
 * batch in a single GPU's HBM:  f[member][q][j][i]  (FP32, double-buffered),
 * with the D2Q9 population order of the reference, q = (ex+1) + 3*(ey+1)
 * (src/config.h:256-258) and node order ij = i + nx*j (src/config.h:210).
 *
 * Every entry point returns an int status (LBMLETKF_OK == 0, negative on error);
 * lbmletkf_last_error() returns a message.  No C++ exception crosses this ABI.
 * Calls on one context are ordered on that context's HIP stream; entry points
 * that return data to the host synchronise that stream.  A context is not
 * thread-safe; distinct contexts are independent.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   update(data&, int)                         src/lbm.h:16-29
 *   data(int) / stdata::init_lbm / init_force  src/data.h:45-64, src/data.cu:25-237
 *   DataAssimilator(const util::mpi&)          src/assimilate.h:31-40
 *   DataAssimilator::assimilate(data&, mpi, t) src/assimilate.h:54-68 -> src/letkf.cu:120-360
 *   util::letkf_solver<T>                      src/util/letkf_solver.h:25-400
 *   ensemble_stat (RMSE / spread)              ensemble_stat/main.cpp:108-177
 */
#ifndef LBMLETKF_H_
#define LBMLETKF_H_

#ifdef __cplusplus
extern "C" {
#endif

#define LBMLETKF_OK              0
#define LBMLETKF_ERR_ARG        -1  /* invalid argument / unsupported size        */
#define LBMLETKF_ERR_HIP        -2  /* HIP runtime error (allocation, launch ...)  */
#define LBMLETKF_ERR_NUMERIC    -3  /* NaN / non-finite value detected on device   */
#define LBMLETKF_ERR_STATE      -4  /* call out of order (e.g. no obs network set) */

/* forcing kinds */
#define LBMLETKF_FORCE_NONE          0
#define LBMLETKF_FORCE_WATANABE97    1   /* src/data.cu:161-350 */
/* initial-condition kinds */
#define LBMLETKF_IC_REFERENCE        0   /* src/data.cu:25-141: mt19937 phases, quadrant modes, max-speed scaling */
#define LBMLETKF_IC_K4_SPECTRUM      1   /* E(k) ~ k^4 exp(-2 (k/k0)^2), half-plane modes, rms-speed scaling */
/* observed variables (bit mask) -- src/letkf.cu:41,183-185, DA_OBS_RUV src/config.h:44 */
#define LBMLETKF_OBS_U   1
#define LBMLETKF_OBS_V   2
#define LBMLETKF_OBS_RHO 4
/* localization kinds */
#define LBMLETKF_LOC_GC_WINDOW       0   /* reference window + Gaspari-Cohn R-localization, src/letkf.cu:44-56,97-113 */
#define LBMLETKF_LOC_GAUSS           1   /* exp(-d^2 / 2L^2) on periodic distance, hard cutoff */
/* eddy-viscosity closures: Smagorinsky (src/lbm.cu:24-65) or the learned per-node MLP */
#define LBMLETKF_CLOSURE_SMAGORINSKY 0
#define LBMLETKF_CLOSURE_MLP         1

/* analysis variants */
#define LBMLETKF_VARIANT_PLAIN       0   /* per-member LBM launches, FP32 Gram on VALU, FP64 Jacobi */
#define LBMLETKF_VARIANT_MIXED       1   /* batched LBM + fused LETKF, mixed precision */

/* All physical / numerical parameters of a run.  The reference fixes these at
 * compile time in src/config.h and the makefile macros
 * (config/makefile.macro.in:1-69); here they are a runtime struct. */
typedef struct lbmletkf_config {
    int    nx, ny;              /* grid (config.h:133-134)                                  */
    double c;                   /* lattice speed c = u_ref/cfl (config.h:248)               */
    double dx, dt;              /* config.h:250-251                                         */
    double nu;                  /* molecular viscosity (config.h:157)                       */
    double rho_ref;             /* config.h:156                                             */
    double cs_smag;             /* Smagorinsky constant (literal 0.2 at lbm.cu:59)          */
    double u_ref;               /* velocity scale multiplying e_q in the force term (lbm.h:47) */
    double friction_rate;       /* linear friction per step (config.h:180, lbm.h:52)        */
    /* stochastic forcing, Watanabe (1997) -- data.cu:161-350 */
    int    force_kind;
    double force_kf;            /* injected wavenumber (config.h:161)                       */
    double force_amp_total;     /* fkf * u_ref / t0 (config.h:171, data.cu:218-223)         */
    double force_sigma_k;       /* Gaussian width in |k| (data.cu:170)                      */
    double force_dk;            /* clipping half-width in |k| (data.cu:171)                 */
    double force_rand_stdev;    /* N(0, s^2) draws, s = 2^-1/4 (data.cu:243)                */
    unsigned long long force_seed;
    /* initial condition -- data.cu:25-141 */
    int    ic_kind;
    double ic_amp;              /* max speed (REFERENCE) or rms speed (K4_SPECTRUM)         */
    double ic_k0, ic_dk, ic_sigma;
    /* observation network and LETKF -- letkf.cu:29-56, letkf_solver.h:235-349 */
    int    obs_stride;          /* DA_XYPRUNE (config.h:59-61)                              */
    int    obs_offset;          /* lattice offset of the observed sites                     */
    int    obs_vars;            /* LBMLETKF_OBS_* mask                                      */
    double obs_sigma_u;         /* OBS_ERROR_U  (config.h:190-202)                          */
    double obs_sigma_rho;       /* OBS_ERROR_RHO                                            */
    int    loc_kind;            /* LBMLETKF_LOC_*                                           */
    int    loc_rloc_len;        /* LETKF_RLOC_LEN (0 = not defined)  letkf.cu:29-39         */
    double loc_scale;           /* Gaussian L (cells)                                       */
    double loc_cutoff;          /* Gaussian hard cutoff (cells)                             */
    double inflation;           /* LETKF_COVINF beta (config.h:48-50)                       */
    int    da_interval;         /* LBM steps per DA cycle (iiter*DAPRUNE, config.h:136,142) */
    int    variant;             /* LBMLETKF_VARIANT_*                                       */
    double jacobi_tol;          /* relative off-diagonal tolerance of the eigensolver       */
    int    jacobi_max_sweeps;   /* cusolver syevj uses tol 1e-6, 100 sweeps (cu_matrix_eigensolver_wrapper.hpp:192-193) */
    int    gram_passes;         /* MIXED: 1 = bf16 operands, 3 = split bf16 (hi*hi + hi*lo + lo*hi) */
    int    jacobi_fp64_sweeps;  /* MIXED: at most this many FP64 sweeps after the FP32 phase (presets: 0) */
    /* eddy-viscosity closure of the fused LBM kernel (MIXED variant) */
    int    closure_kind;         /* LBMLETKF_CLOSURE_*                                       */
    double closure_strain_scale; /* MLP inputs are strain components / this scale            */
    double closure_nu_scale;     /* MLP output nu_t = this scale * softplus(z)               */
    double closure_nu_max;       /* nu_t above this (or not finite) falls back to Smagorinsky */
    /* MIXED: FP32 stop tolerance (0: 3e-6).  Without FP64 sweeps the phase stops
     * once the columns' largest pair cosine, measured after a sweep, is at most a
     * third of this, or one sweep after that measurement, below 1e-4, stops halving;
     * with FP64 sweeps it stops after the sweep whose largest pair cosine met is at
     * most this, or once that cosine, below 1e-4, no longer halves per sweep */
    double jacobi_tol32;
} lbmletkf_config;

typedef struct lbmletkf_ctx lbmletkf_ctx;

/* Library version string; never fails, needs no GPU. */
const char* lbmletkf_version(void);

/* Fill *cfg with a named preset: "ref" (src/config.h constants at SCALER=4,
 * Cs=0.2, Watanabe97 forcing, GC window) or "baseline" (lattice units, tau=0.503,
 * Cs=0.1, decaying turbulence, (u,v) every 8th node, Gaussian L=12).  Needs no GPU. */
int lbmletkf_config_preset(const char* name, lbmletkf_config* cfg);

/* Validate a configuration without creating a context.  Needs no GPU. */
int lbmletkf_config_check(const lbmletkf_config* cfg, int n_members);

/* Create a context for n_members ensemble members on the current HIP device.
 * Replaces data(int) (src/data.h:55) for all members at once and
 * DataAssimilator(mpi) (src/assimilate.h:31).  All state starts at rest (rho=rho_ref). */
int lbmletkf_create(const lbmletkf_config* cfg, int n_members, lbmletkf_ctx** out);
int lbmletkf_destroy(lbmletkf_ctx* ctx);

/* Last error message of ctx (or of the calling thread when ctx is NULL). */
const char* lbmletkf_last_error(const lbmletkf_ctx* ctx);

/* Initial condition of one member from cfg.ic_* with the given seed; replaces
 * stdata::init_lbm (src/data.cu:25-141; the reference passes ensemble_id[+334]). */
int lbmletkf_init_member(lbmletkf_ctx* ctx, int member, unsigned long long seed);

/* Host copies of one member's populations, layout [q][j][i] (src/config.h:212). */
int lbmletkf_set_state(lbmletkf_ctx* ctx, int member, const float* f_host);
int lbmletkf_get_state(lbmletkf_ctx* ctx, int member, float* f_host);

/* f := feq(rho,u,v) for one member (src/lbm.h:31-39); host arrays [j][i]. */
int lbmletkf_set_macro(lbmletkf_ctx* ctx, int member, const float* rho, const float* u, const float* v);

/* rho,u,v of one member from its populations (src/lbm.h:72-109).  Host arrays [j][i];
 * any pointer may be NULL.  lbmletkf_get_macro_device writes device arrays, stream-ordered. */
int lbmletkf_get_macro(lbmletkf_ctx* ctx, int member, float* rho, float* u, float* v);
int lbmletkf_get_macro_device(lbmletkf_ctx* ctx, int member, float* rho, float* u, float* v);

/* Advance all members by n_steps LBM steps; replaces n calls of
 * update(data&, t) (src/lbm.h:16-29): forcing -> pull-stream + moments ->
 * Smagorinsky nu_t -> BGK collide + force -> swap.  Stream-ordered, non-blocking. */
int lbmletkf_forecast(lbmletkf_ctx* ctx, int n_steps);

/* Step counter that indexes the forcing random stream (the reference consumes
 * its cuRAND buffer once per update(), src/data.cu:240-253). */
int lbmletkf_set_step(lbmletkf_ctx* ctx, long long step);
int lbmletkf_get_step(lbmletkf_ctx* ctx, long long* step);

/* Observation network and localization from cfg.obs_* / cfg.loc_*; builds the
 * per-gridpoint local observation sets on the device.  Replaces init_letkf /
 * set_rR (src/letkf.cu:59-118).  Called implicitly by the first analysis. */
int lbmletkf_setup_obs(lbmletkf_ctx* ctx);

/* Explicit observation network: sites at the given nodes (i + nx*j, distinct),
 * site s = position in the list, replacing the lattice of cfg.obs_stride (uneven
 * site densities; a gridpoint may have no local observation, in which case its
 * analysis is xbar + sqrt(beta) dX).  Needs the Gaussian localization; the local
 * set of a gridpoint lists the sites within the cutoff scanning the window by
 * rows, then columns, ascending.  No counterpart in the reference, whose network
 * is the DA_XYPRUNE lattice (src/letkf.cu:29-56). */
int lbmletkf_set_obs_sites(lbmletkf_ctx* ctx, int n_sites, const int* nodes);

/* Number of observed sites and of scalar observations (sites x observed vars).
 * Observation vectors are ordered site-major: o = site*n_vars + var, with site
 * s = sx + sy*(nx/stride) and vars in the order u, v, rho (those in the mask). */
int lbmletkf_obs_count(lbmletkf_ctx* ctx, int* n_sites, int* n_obs);

/* Synthetic observations of one member (normally the nature run, member 0 of
 * a one-member context): y = H(x) + sigma*N(0,1), noise from Philox4x32-10 keyed
 * by (seed, counter); noise=0 gives H(x).  Output has n_obs floats; device or host.
 * Replaces the OBSERVE branch of output() (src/io.cu:159-246). */
int lbmletkf_observe(lbmletkf_ctx* ctx, int member, unsigned long long seed,
                     unsigned long long counter, int noise, float* yo, int yo_on_device);

/* Gather observations from full fields laid out like the reference's
 * io/observed files (float [ny][nx], NaN where unobserved; src/letkf.cu:197-221,
 * 237-269).  Any field may be NULL if its variable is not observed. Host arrays. */
int lbmletkf_obs_from_fields(lbmletkf_ctx* ctx, const float* rho, const float* u,
                             const float* v, float* yo_host);

/* One LETKF analysis of all members at all gridpoints with observation vector yo
 * (n_obs floats, device or host pointer).  Replaces DataAssimilator::assimilate
 * (src/assimilate.h:54-68 -> src/letkf.cu:120-360 and letkf_solver::solve_evd /
 * solve_xa, src/util/letkf_solver.h:235-349).  Stream-ordered; a NaN in the
 * analysis is reported by the next synchronising call (LBMLETKF_ERR_NUMERIC). */
int lbmletkf_analysis(lbmletkf_ctx* ctx, const float* yo, int yo_on_device);

/* One full DA cycle: analysis(yo) then forecast(cfg.da_interval) -- the order of
 * src/main.cu:95-113.  yo is a device pointer.  In the MIXED variant, with phase
 * timing off, the kernels of the cycle are captured once into a hipGraph (per parity
 * of the population double buffer) and replayed; set_dump drops the graphs. */
int lbmletkf_cycle(lbmletkf_ctx* ctx, const float* yo_device);

/* Ensemble statistics against a truth (ensemble_stat/main.cpp:108-177, with the
 * actual grid size rather than its hard-coded 128*128).  truth arrays [j][i], device
 * or host; out = {rmse_uv, spread_uv, rmse_rho, spread_rho}.  Spread uses the biased
 * (1/M) standard deviation as the reference does. */
int lbmletkf_ensemble_stats(lbmletkf_ctx* ctx, const float* truth_rho, const float* truth_u,
                            const float* truth_v, int truth_on_device, double out[4]);

/* Per-node fields over the members ([ny][nx] each; a NULL output is skipped): the
 * ensemble mean of u and v, the mean and standard deviation of the vorticity
 * (central differences, src/io.cu:61-86), and the standard deviations of u and v;
 * deviations biased (1/M).  The ENSEMBLE_STAT outputs of src/io.cu:90-145
 * (ens_mean_u, ens_mean_v, ens_vor_mean, ens_vor_std) plus the spread fields,
 * without gathering the members on the host (one reduction over the members on
 * the device; the buffers are allocated at the first call). */
int lbmletkf_ensemble_fields(lbmletkf_ctx* ctx, float* mean_u, float* mean_v, float* vor_mean, float* vor_std,
                             float* sd_u, float* sd_v);

/* lbmletkf_ensemble_stats with the truth on the device and no allocation: no
 * counterpart in the reference; for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops).  The per-block partial sums
 * go to a static device array of the library (2 MiB per device, grids of up to
 * 2^24 nodes; larger grids return LBMLETKF_ERR_ARG).  Calls that overlap on one
 * device share that array and race.  Same result as lbmletkf_ensemble_stats. */
int lbmletkf_ensemble_stats_noalloc(lbmletkf_ctx* ctx, const float* truth_rho_device,
                                    const float* truth_u_device, const float* truth_v_device, double out[4]);

/* Local observation set of one gridpoint (parity diagnostics): site indices and
 * localization weights in gather order.  *n receives the count; at most cap are written. */
int lbmletkf_local_obs(lbmletkf_ctx* ctx, int gridpoint, int* n, int* sites, float* weights, int cap);

/* Record the ensemble transform of gridpoints [g0, g0+count) at every following
 * analysis (count=0 disables); lbmletkf_get_dump returns, per gridpoint, W (MxM
 * row-major), w (M) and the ascending eigenvalues of C (M), in FP64.
 * Diagnostics counterpart of the reference's HDF5 dumps (src/util/letkf_solver.h:244-345). */
int lbmletkf_set_dump(lbmletkf_ctx* ctx, int g0, int count);
int lbmletkf_get_dump(lbmletkf_ctx* ctx, double* W, double* w, double* evals);
/* The eigensolve side of the same dump (src/util/letkf_solver.h:235-272, the _Q and
 * _V matrices of solve_evd): per gridpoint C = (M-1)/beta I + dY^T R~^-1 dY as the
 * kernels formed it (M x M row-major), the eigenvectors V (row-major, column j the
 * eigenvector of evals[j]) and the eigenvalues in the solver's order (unsorted, as
 * syevjBatched returns them with sorting off), in FP64. */
int lbmletkf_get_dump_evd(lbmletkf_ctx* ctx, double* C, double* V, double* evals);

/* The observation-space ensemble of the last analysis (src/util/letkf_solver.h:
 * 238-243: dY = Y - ybar, dyo = yo - ybar): dy [n_obs][M] and dyo [n_obs] in the
 * observation order of lbmletkf_obs_count (site-major, variables u, v, rho as
 * configured).  With lbmletkf_local_obs they give each gridpoint's _dYf and _dyo. */
int lbmletkf_get_obs_ensemble(lbmletkf_ctx* ctx, float* dy, float* dyo);

/* Phase timers on the context's stream (HIP events around each phase), the
 * counterpart of the reference's util::timer tags (src/util/timer.hpp:17-85,
 * src/main.cu:95-115, src/letkf.cu:120-360).  enable(1) clears and starts them.
 * Phases: "forecast" (all LBM kernels incl. forcing), "obs" (H(x), innovations),
 * "gram" (PLAIN: C and b), "solve" (PLAIN: eigensolve, transform, state update;
 * MIXED: gather, MFMA Gram and the FP32 Jacobi sweeps), "finish" (MIXED: FP64
 * sweeps, transform, state update).
 * read() returns the accumulated seconds, the number of kernel launches and a
 * work count: node updates (forecast), observations (obs), local scalar
 * observations summed over gridpoints (gram), Jacobi sweeps summed over
 * gridpoints (solve; FP32 sweeps in MIXED).  The pseudo-phase "solve_fp64" reads
 * the MIXED variant's FP64 sweeps summed over gridpoints as its work count.
 * read() synchronises the stream.  While timing is on, lbmletkf_cycle launches
 * eagerly instead of replaying its graph. */
int lbmletkf_profile_enable(lbmletkf_ctx* ctx, int on);
int lbmletkf_profile_read(lbmletkf_ctx* ctx, const char* phase, double* seconds,
                          long long* launches, long long* work);

/* Newtonian nudging of every member towards observed fields, the reference's
 * DA_NUDGING analysis (src/nudging.cu:40-142, da_w_itp): the observation lattice of
 * stride xyprune is interpolated bilinearly (periodic) to every node and
 * f_q := alpha feq_q(rho_o, u_o, v_o) + (1 - alpha) f_q.  rho, u, v: [ny][nx] fields
 * (device or host) of which only the lattice nodes are read (xyprune = 1: every
 * node); rho may be NULL (observations of u, v only: rho_o is the member's own
 * density).  quadra (DA_QUADRA): odd leaves i >= nx/2 alone, >= 2 leaves j >= ny/2
 * alone. */
int lbmletkf_nudge(lbmletkf_ctx* ctx, const float* rho, const float* u, const float* v, int on_device,
                   double alpha, int xyprune, int quadra);

/* Perturbation of the Lyapunov runs (src/data.cu:426-456, src/main.cu:73-76):
 * random velocities u_ref N(0,1) per node from std::mt19937(ensemble_id + 132333)
 * and std::normal_distribution<float> (u then v, rows outer), and
 * f_q := epsilon feq_q(rho, u_r, v_r) + (1 - epsilon) f_q. */
int lbmletkf_perturb(lbmletkf_ctx* ctx, int member, unsigned long long ensemble_id, double epsilon);

/* ---- learned eddy-viscosity closure ------------------------------------------
 * No counterpart in the reference (its LES model is Smagorinsky only,
 * src/lbm.cu:24-65): a per-node network 18 -> 64 -> 64 -> 1 (tanh-form GELU,
 * softplus output) maps the 3x3 stencil of the streamed velocity's strain
 * (a = (ux - vy)/2, b = (uy + vx)/2, divided by closure_strain_scale; stencil point
 * s = (dj+1)*3 + (di+1) gives inputs 2s, 2s+1) to nu_t = closure_nu_scale *
 * softplus(z).  Parameters are one FP32 vector of LBMLETKF_CLOSURE_NPARAM values:
 * W1[64][18], b1[64], W2[64][64], b2[64], w3[64], b3. */
#define LBMLETKF_CLOSURE_NPARAM 5441
/* training sample: 18 inputs, a, b (unscaled), t1 = (tau11 - tau22)/2, t2 = tau12 */
#define LBMLETKF_CLOSURE_SAMPLE 22

/* Initial parameters from a seed (Glorot-uniform weights, zero biases, output
 * layer scaled by 0.1 and b3 = softplus^-1(1): nu_t ~ closure_nu_scale). */
int lbmletkf_closure_init_params(unsigned long long seed, float* params);

/* Load the network into a context created with closure_kind = MLP; every later
 * LBM step of that context (forecast, cycle) runs the fused kernel with it
 * (weights in bf16, FP32 accumulation).  Nodes where the network's nu_t is not
 * finite or outside [0, closure_nu_max] use the Smagorinsky nu_t.  Requires nx a
 * multiple of 64 and ny of 16. */
int lbmletkf_set_closure(lbmletkf_ctx* ctx, const float* params);

/* nu_t (after the fallback) that the next step would use for one member, [j][i]
 * host array: the same kernel as the forecast, its populations discarded. */
int lbmletkf_closure_nut(lbmletkf_ctx* ctx, int member, float* nut);

/* Box filter of one member by an integer factor: rho_bar, u_bar, v_bar on the
 * (nx/factor) x (ny/factor) grid, host arrays [J][I]. */
int lbmletkf_filtered_macro(lbmletkf_ctx* fine, int member, int factor, float* rho, float* u, float* v);

/* Trainer: samples in device memory, parameters and Adam state, the forward and
 * backward kernels.  coarse_cfg describes the coarse grid (dx, closure scales). */
typedef struct lbmletkf_trainer lbmletkf_trainer;
int lbmletkf_trainer_create(const lbmletkf_config* coarse_cfg, long long capacity, lbmletkf_trainer** out);
int lbmletkf_trainer_destroy(lbmletkf_trainer* tr);
/* Append the (nx/factor)*(ny/factor) samples of one member of a fine-grid context:
 * box-filtered velocity and the subgrid stress tau_ij = <u_i u_j> - <u_i><u_j>. */
int lbmletkf_trainer_add_samples(lbmletkf_trainer* tr, lbmletkf_ctx* fine, int member, int factor);
int lbmletkf_trainer_set_samples(lbmletkf_trainer* tr, const float* samples, long long n);
int lbmletkf_trainer_get_samples(lbmletkf_trainer* tr, long long first, long long n, float* samples);
int lbmletkf_trainer_count(lbmletkf_trainer* tr, long long* n);
/* set_params also clears the Adam moments */
int lbmletkf_trainer_set_params(lbmletkf_trainer* tr, const float* params);
int lbmletkf_trainer_get_params(lbmletkf_trainer* tr, float* params);
/* Mean loss w_norm ((2 nu_t a + t1)^2 + (2 nu_t b + t2)^2) over samples
 * [first, first+n) and its gradient (NPARAM floats, may be NULL); the loss is the
 * model error of tau_dev = -2 nu_t S_dev.  w_norm <= 0 takes 1 / mean(t1^2 + t2^2)
 * over all samples. */
int lbmletkf_trainer_loss_grad(lbmletkf_trainer* tr, long long first, long long n, double w_norm, double* loss,
                               float* grad);
/* steps Adam updates (beta 0.9, 0.999, eps 1e-8) on minibatches walking a
 * permutation of the samples; loss_history (steps floats, may be NULL) receives
 * each minibatch's loss before its update. */
int lbmletkf_trainer_train(lbmletkf_trainer* tr, int batch, int steps, double lr, double w_norm,
                           float* loss_history);

/* ---- tuning ----------------------------------------------------------------------
 * Launch shapes of the MIXED variant: the shape of the fused LBM step --
 * lbm_tiles_per_block = k > 0: k tiles of 64 x 16 per workgroup; = -h: full-row
 * strips of h rows (h = 8, 16, 32; nx a multiple of 64 up to 1024, one thread per
 * column); = -(100 + h): full-row strips of h rows advanced two steps per launch
 * (h = 16, 32; nx a multiple of 64 up to 1024); = -(200 + h): three steps per
 * launch (h = 16, 32; nx a multiple of 64 up to 512).  The multi-step strips serve
 * flows without body force or friction (others take one step per launch) and are
 * bit-identical to single steps -- and the chain length of the analysis (gridpoints
 * of one column per wavefront, each warm-started from the previous one's
 * eigenvectors; the result changes by FP32 rounding).
 * No counterpart in the reference. */
int lbmletkf_set_tuning(lbmletkf_ctx* ctx, int lbm_tiles_per_block, int mix_chain);
int lbmletkf_get_tuning(lbmletkf_ctx* ctx, int* lbm_tiles_per_block, int* mix_chain);
/* Autotuner: times each candidate on this device with the context's own state
 * and observation network (the LBM into the idle half of the double buffer, the
 * solve kernel alone on observations of member 0: the ensemble is not changed)
 * and keeps the fastest.  cache_path (may be NULL) is a text file of
 * "<device/problem key> <tiles per block> <chain>" lines: a hit is applied without
 * timing, a miss is appended.  report (may be NULL) receives the timings. */
int lbmletkf_autotune(lbmletkf_ctx* ctx, const char* cache_path, char* report, int report_len);

/* ---- groups: several OSSEs in shared launches ----------------------------------
 * No counterpart in the reference (one OSSE per run).  A group of 1-8 contexts of
 * different grids and ensemble sizes (MIXED variant, grids multiples of 64 x 16,
 * one gram_passes / FP64-sweep setting) is analysed by ONE launch of each
 * analysis kernel over every gridpoint of every context, and stepped by each
 * context's own tuned LBM launches (a context with the learned closure by its
 * closure kernel, which needs lbmletkf_set_closure first; the group calls check
 * that before any launch): the first context on the group's stream, the others
 * on side streams of the group forked from it by an event and joined back to it
 * (also when a launch fails), so that work ordered after a group call on the
 * group's stream sees every context's steps; the group cycle (analysis, then
 * da_interval steps, equal in all contexts) is captured as one hipGraph.  The
 * group works on its own streams after draining the contexts' streams; do not use
 * a member context concurrently with a group call.  yo_device[k] is the observation vector (device) of the k-th context as
 * passed to create.  Errors of the group calls read from lbmletkf_last_error(NULL). */
typedef struct lbmletkf_group lbmletkf_group;
int lbmletkf_group_create(lbmletkf_ctx* const* ctxs, int n, lbmletkf_group** out);
int lbmletkf_group_destroy(lbmletkf_group* g);
int lbmletkf_group_forecast(lbmletkf_group* g, int n_steps);
int lbmletkf_group_analysis(lbmletkf_group* g, const float* const* yo_device);
int lbmletkf_group_cycle(lbmletkf_group* g, const float* const* yo_device);
/* waits for the group's streams and reports a NaN flagged by any member analysis */
int lbmletkf_group_synchronize(lbmletkf_group* g);
void* lbmletkf_group_get_stream(lbmletkf_group* g);

/* Stream access for callers that time or order work around the context. */
void* lbmletkf_get_stream(lbmletkf_ctx* ctx);
/* Run the context's work on the caller's hipStream_t from now on (NULL: back to a
 * stream the context owns).  Work queued before the call completes first; the
 * caller's stream must outlive its use by the context (destroy the context or set
 * another stream first).  Replaces the reference's implicit per-thread default
 * stream (config/makefile.compiler.in:43, --default-stream per-thread), on which
 * update() and assimilate() run. */
int lbmletkf_set_stream(lbmletkf_ctx* ctx, void* stream);
int lbmletkf_synchronize(lbmletkf_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* LBMLETKF_H_ */
