/*
 * FP64 CPU oracle of the LBM-LETKF cycle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++/OpenMP restatement of the reference's algorithm, used by the
 * tests as the parity checker and by bench.py as the CPU baseline.  Product
 * code never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository).  How each function is pinned (golden vector,
 * closed form, library cross-check, invariant) is listed in tests/ and in the
 * design notes; the cuRAND MTGP32 forcing stream of the reference is replaced by
 * Philox4x32-10 and is therefore not pinned to the reference's draws.
 *
 * The configuration struct has the field order of lbmletkf_config.
 */
#ifndef LBMLETKF_ORACLE_H_
#define LBMLETKF_ORACLE_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_config {
    int    nx, ny;
    double c, dx, dt, nu, rho_ref, cs_smag, u_ref, friction_rate;
    int    force_kind;
    double force_kf, force_amp_total, force_sigma_k, force_dk, force_rand_stdev;
    unsigned long long force_seed;
    int    ic_kind;
    double ic_amp, ic_k0, ic_dk, ic_sigma;
    int    obs_stride, obs_offset, obs_vars;
    double obs_sigma_u, obs_sigma_rho;
    int    loc_kind, loc_rloc_len;
    double loc_scale, loc_cutoff, inflation;
    int    da_interval, variant;
    double jacobi_tol;
    int    jacobi_max_sweeps;
    int    gram_passes, jacobi_fp64_sweeps;  /* GPU precision knobs; unused by the oracle */
    int    closure_kind;
    double closure_strain_scale, closure_nu_scale, closure_nu_max;
    double jacobi_tol32;
} orc_config;

int  orc_sizeof_config(void);
void orc_set_threads(int n);
/* Use LAPACK dsyev (a function pointer with the Fortran LP64 signature) for the
   local eigensolves of orc_letkf_transform; NULL: the cyclic Jacobi below. */
void orc_set_dsyev(void* dsyev);

/* Philox4x32-10 block (Salmon et al. SC'11) and the Box-Muller normals built on it */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_philox_normal4(uint64_t key, const uint32_t ctr[4], double out[4]);

/* forcing: src/data.cu:161-223 (modes) and :239-350 (field) */
int  orc_force_modes(const orc_config* cfg, double* kx, double* ky, double* amp, int cap);
void orc_force_field(const orc_config* cfg, long long step, double* fx, double* fy);

/* D2Q9 pieces: src/lbm.h:31-109, src/lbm.cu:24-160, src/lbm_collision_srt.h:8-16 */
void orc_feq(const orc_config* cfg, const double* rho, const double* u, const double* v, double* f);
void orc_macro(const orc_config* cfg, const double* f, double* rho, double* u, double* v);
void orc_stream(const orc_config* cfg, const double* f, double* fn);
void orc_les(const orc_config* cfg, const double* u, const double* v, double* nus);
/* one full step of M members, f[M][9][ny][nx]; step indexes the forcing stream */
void orc_lbm_step(const orc_config* cfg, int M, const double* fin, double* fout, long long step);
/* the same step with a given body-force field Fx, Fy [ny][nx] (shared by the members) */
void orc_lbm_step_forced(const orc_config* cfg, int M, const double* fin, double* fout, const double* Fx,
                         const double* Fy);

/* initial velocity of one member, src/data.cu:25-141 (direct mode sums) */
void orc_initial_velocity(const orc_config* cfg, unsigned long long seed, double* u, double* v);

/* observation network and local sets: src/letkf.cu:29-56, 97-113, 150-187 */
void orc_obs_count(const orc_config* cfg, int* n_sites, int* n_obs);
int  orc_local_obs(const orc_config* cfg, int gridpoint, int* sites, double* weights, int cap);
double orc_gaspari_cohn(double d);
void orc_obs_moments(const orc_config* cfg, int M, const double* f, double* hx);
void orc_observe(const orc_config* cfg, const double* hx, unsigned long long seed,
                 unsigned long long counter, int noise, double* yo);

/* symmetric eigen-decomposition by row-cyclic Jacobi; returns sweeps used.
 * A (n x n, row-major) is destroyed; V receives eigenvectors as columns. */
int  orc_jacobi_eig(int n, double* A, double* V, double* d, double tol, int max_sweeps);

/* LETKF transform of one local problem, src/util/letkf_solver.h:235-349:
 * dY [p][M] (rows = observations), rinv [p], dyo [p] -> W [M][M], w [M], evals [M] (ascending) */
int  orc_letkf_transform(int M, int p, const double* dY, const double* rinv, const double* dyo,
                         double beta, double tol, int max_sweeps, double* W, double* w, double* evals);

/* full analysis of f[M][9][ny][nx] in place; optional dump of W, w, evals for
 * gridpoints [dump_g0, dump_g0+dump_n).  g_stride > 1 analyses only every
 * g_stride-th gridpoint (used to time a bounded sample).  Returns 0, or -3 on NaN. */
int  orc_letkf_analysis(const orc_config* cfg, int M, double* f, const double* yo,
                        int dump_g0, int dump_n, double* W, double* w, double* evals, int g_stride);

/* the analysis of gridpoints [g0, g0+n) only (f updated there), with W, w, evals
 * of each; the observation moments still come from the whole ensemble.  Returns
 * 0, -1 on a range outside the grid, -3 on NaN. */
int  orc_letkf_analysis_strip(const orc_config* cfg, int M, double* f, const double* yo, int g0, int n,
                              double* W, double* w, double* evals);

/* ensemble_stat/main.cpp:108-177 with N = nx*ny: out = {rmse_uv, spread_uv, rmse_rho, spread_rho} */
void orc_ensemble_stats(const orc_config* cfg, int M, const double* f, const double* trho,
                        const double* tu, const double* tv, double out[4]);

/* ---- learned eddy-viscosity closure (an extension beyond the reference; no reference code) ----
 * features at node n: for the 3x3 stencil s = (dj+1)*3 + (di+1), the deviatoric
 * strain a = (ux - vy)/2 and shear b = (uy + vx)/2 at node (i+di, j+dj) from the
 * central differences of src/lbm.cu:44-47, divided by strain_scale:
 * feat[n][2s] = a, feat[n][2s+1] = b.  ab[n] = (a, b) at the node itself, unscaled.
 * MLP 18 -> 64 -> 64 -> 1: h1 = gelu(W1 x + b1), h2 = gelu(W2 h1 + b2),
 * z = w3 . h2 + b3, nu_t = nu_scale * softplus(z); gelu(x) = x/2 (1 + tanh(k (x +
 * 0.044715 x^3))), k = sqrt(2/pi).  params: W1[64][18], b1[64], W2[64][64], b2[64],
 * w3[64], b3 (5441 values).  bf16 != 0 rounds x, W1, W2 and h1 to bf16 (the fused
 * kernel's MFMA operands).
 * A sample is 22 values: 18 features, a, b (unscaled), t1 = (tau11 - tau22)/2,
 * t2 = tau12 (the filtered SGS stress); loss = mean w_norm ((2 nu a + t1)^2 +
 * (2 nu b + t2)^2), the squared misfit of tau_dev = -2 nu_t S_dev. */
#define ORC_CLOSURE_NPARAM 5441
#define ORC_CLOSURE_SAMPLE 22
void   orc_closure_features(int nx, int ny, double dx, double strain_scale, const double* u, const double* v,
                            double* feat, double* ab);
void   orc_closure_nut(const double* params, const double* feat, long long n, double nu_scale, int bf16,
                       double* nut);
double orc_closure_loss_grad(const double* params, const double* samples, long long n, double nu_scale,
                             double w_norm, double* grad);
/* box filter by F of fine fields [nyf][nxf] -> coarse [3][ny][nx] (rho, u, v) and
 * samples [ny*nx][22]; coarse dx = dx_coarse */
void   orc_coarse_grain(int nxf, int nyf, int F, const double* rho, const double* u, const double* v,
                        double dx_coarse, double strain_scale, double* samples, double* coarse);
/* one LBM step (as orc_lbm_step) with nu_t from the MLP where it is finite and
 * within [0, closure_nu_max], Smagorinsky elsewhere */
void   orc_lbm_step_closure(const orc_config* cfg, int M, const double* fin, double* fout, long long step,
                            const double* params, int bf16);

/* Explicit observation sites (node indices) replacing the lattice of cfg in every
 * observation function that follows (n < 0 or nodes NULL: back to the lattice).
 * Local sets then scan the cutoff window by rows, then columns, ascending. */
void   orc_set_obs_sites(int n, const int* nodes);

/* ---- nudging and the Lyapunov perturbation ---------------------------------- */
/* src/nudging.cu:40-142 (da_w_itp): M members f [M][9][ny][nx] updated in place;
 * rho may be NULL (rho_o = the member's own density) */
void   orc_nudge(const orc_config* cfg, int M, double* f, const double* rho, const double* u, const double* v,
                 double alpha, int xyprune, int quadra);
/* src/data.cu:432-447: the random velocities of ensemble id (std::mt19937(id + 132333),
 * std::normal_distribution<float>, u then v per node, rows outer, times u_ref) */
void   orc_perturb_draws(const orc_config* cfg, unsigned long long ensemble_id, double* ur, double* vr);
/* src/data.cu:443-454: f [9][ny][nx] := eps feq(rho, ur, vr) + (1 - eps) f */
void   orc_perturb(const orc_config* cfg, double* f, const double* ur, const double* vr, double eps);

#ifdef __cplusplus
}
#endif
#endif
