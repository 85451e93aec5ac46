// Internal state of a context and the launchers shared between translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/lbmletkf.h"

namespace lbmletkf {

constexpr int Q = 9;
constexpr int MAX_MEMBERS = 64;      // one wavefront of lanes per ensemble matrix
constexpr int MAX_FORCE_MODES = 100; // src/data.cu:191-212 prunes to 100 modes
constexpr int MIXED_GROUP_MAX = 8;   // contexts in one group launch
// doubles per gridpoint of the transform dump (layout at Ctx::d_dump)
__host__ __device__ constexpr long long dump_record(int M) { return 3LL * M * M + 2LL * M; }
constexpr int LBM_TWO_STEP = 100;    // shape code offset of the two-step strips
constexpr int LBM_THREE_STEP = 200;  // shape code offset of the three-step strips

// Per-step LBM constants, in the arithmetic type of the kernels (FP32).
struct LbmParams {
    int nx, ny, n_members;
    long long nodes;          // nx*ny
    long long member_stride;  // Q*nx*ny floats
    float c, inv_c, inv_c2;   // c, 1/c, 1/c^2
    float k_tau;              // tau = 0.5 + k_tau*(nu + nu_t), k_tau = 3/(c^2 dt)   (lbm.cu:133-138)
    float nu;
    float les_coef;           // Cs^2 dx^2                                           (lbm.cu:59-60)
    float inv_2dx;            // central difference 1/(2 dx)                         (lbm.cu:44-47)
    float force_scale;        // u_ref * 3/c^2: e_q*u_ref/cs2 of force_acc            (lbm.h:47-54)
    float fric_over_dt;       // friction_rate/dt                                   (lbm.h:52-53)
    float dt;
    int has_force;
};

// Forcing modes of Watanabe (1997) as built by init_forces (src/data.cu:161-223).
struct ForceTable {
    int n = 0;
    float kx[MAX_FORCE_MODES], ky[MAX_FORCE_MODES], amp[MAX_FORCE_MODES];
};

// Observed lattice and per-variable error weights.
struct ObsNet {
    int stride = 0, offset = 0;
    int n_sx = 0, n_sy = 0, n_sites = 0;
    int n_vars = 0;
    int var_id[3] = {0, 0, 0};          // 0=u 1=v 2=rho, in observation order
    float inv_sigma2[3] = {0, 0, 0};    // R^-1 diagonal per variable (letkf.cu:89-94)
    // sites: a lattice (stride, offset) or, when `custom`, an explicit node list
    bool custom = false;
    std::vector<int> site_nodes;        // node of each site (filled by setup for both kinds)
    int max_local = 0;                  // largest local site count over gridpoints
    long long n_entries = 0;            // total local (site, weight) pairs
};

// MIXED: PH_SOLVE = gather + Gram + FP32 sweeps, PH_FINISH = FP64 sweeps + assembly + update
enum Phase { PH_FORECAST = 0, PH_OBS, PH_GRAM, PH_SOLVE, PH_FINISH, PH_COUNT };

struct ClosureWeights;

struct ProfRec {
    int phase;
    hipEvent_t a, b;
    long long launches, work;
};

struct Ctx {
    lbmletkf_config cfg{};
    int M = 0;
    int nx = 0, ny = 0;
    long long nodes = 0;
    LbmParams lp{};
    ForceTable force{};
    hipStream_t stream = nullptr;
    bool own_stream = true;  // false while the caller's stream (lbmletkf_set_stream) is in use
    int device = 0;
    std::string last_error;

    // state, double-buffered SoA f[member][q][j][i]
    float* f[2] = {nullptr, nullptr};
    int cur = 0;
    // plain-variant intermediates, one member at a time: rho,u,v after streaming, nu_t
    float* plain_ruv = nullptr;   // [3][nodes]
    float* plain_nus = nullptr;   // [nodes]
    // forcing field of the current step, shared by all members: [2][nodes]
    float* force_xy = nullptr;
    ForceTable* d_force = nullptr;
    long long* d_step = nullptr;  // device copy of the step counter (graph-replay safe)
    long long step = 0;

    // device status word: bit 0 = NaN in analysis
    int* d_status = nullptr;
    int* h_status = nullptr;      // pinned

    // observation network and local sets (CSR over gridpoints)
    bool obs_ready = false;
    ObsNet obs{};
    int* d_site_node = nullptr;   // [n_sites] node index of each site
    int* d_loc_ptr = nullptr;     // [nodes+1]
    int* d_loc_site = nullptr;    // [n_entries]
    float* d_loc_w = nullptr;     // [n_entries] localization weight (without R^-1)
    float* d_obs_coef = nullptr;  // [3] R^-1 per variable, device copy
    int* d_site_map = nullptr;    // custom networks: [nodes] site index of each node or -1

    // analysis buffers
    float* d_yo = nullptr;        // [n_obs]
    double* d_hx = nullptr;       // [n_obs][M] H(x) of every member (FP64: rho ~ 1 + O(Ma^2))
    double* d_hx1 = nullptr;      // [n_obs] H(x) of one member (observe)
    float* d_dy = nullptr;        // [n_obs][M] perturbations dY = Y - ybar
    float* d_dyo = nullptr;       // [n_obs] innovation yo - ybar
    float* d_gram = nullptr;      // plain: [nodes][M][M] C matrices
    float* d_bvec = nullptr;      // plain: [nodes][M] dY^T R^-1 dyo
    // MIXED workspace of the two-kernel form (FP64 sweeps requested; N = mix_width,
    // M padded to 16/32/64); the fused form of the presets needs none
    float* d_mix_u = nullptr;     // [nodes][N][N] FP32 columns of U = C V after the FP32 sweeps
    float* d_mix_b = nullptr;     // [nodes][N] b = dY^T R^-1 dyo
    int mix_width = 0;
    int mix_chain = 8;            // consecutive gridpoints per wavefront in k_mixed_solve
    // shape of the fused LBM step: k > 0 tiles of 64x16 per workgroup (k_lbm_fused),
    // -h full-row strips of h rows (k_lbm_rows), -(LBM_TWO_STEP + h) full-row strips
    // of h rows advanced two steps per launch (k_lbm_rows2), -(LBM_THREE_STEP + h)
    // three steps per launch (k_lbm_rows3)
    int lbm_tiles_per_block = 1;

    // transform dump: per gridpoint a record of dump_record(M) doubles,
    //   [0, M^2)          W (PLAIN) or V (MIXED), row-major
    //   [M^2, M^2+M)      w (PLAIN) or the eigenvalues d (MIXED)
    //   [M^2+M, M^2+2M)   d (PLAIN) or b = G^T sqrt(R^-1) dyo (MIXED)
    //   [M^2+2M, 2M^2+2M) C, the matrix the eigensolve took, row-major
    //   [2M^2+2M, 3M^2+2M) V (PLAIN), row-major
    int dump_g0 = 0, dump_n = 0;
    double* d_dump = nullptr;     // [dump_n][dump_record(M)]
    // lbmletkf_ensemble_fields: members' (u, v) and the six output fields (allocated on first use)
    float* d_ens_uv = nullptr;
    float* d_ens_out = nullptr;

    // scratch for reductions
    double* d_red = nullptr;      // [red_cap]
    int red_cap = 0;

    // macro scratch for stats
    float* d_truth = nullptr;     // [3][nodes]
    double* stats_scratch = nullptr;  // device address of the static partial sums (noalloc stats)

    // phase timers
    bool prof = false;
    std::vector<ProfRec> prof_pending;
    std::vector<hipEvent_t> event_pool;
    double prof_sec[PH_COUNT] = {};
    long long prof_launch[PH_COUNT] = {};
    long long prof_work[PH_COUNT] = {};
    unsigned long long* d_sweeps = nullptr;    // Jacobi sweeps summed over gridpoints
    unsigned long long* d_sweeps64 = nullptr;  // MIXED: FP64 sweeps summed over gridpoints
    int dump_kind = 0;  // 0: W, w, evals (plain); 1: V, d, b (mixed), assembled on the host

    // MIXED cycle (analysis + da_interval steps) as an instantiated hipGraph, one per
    // parity of `cur` (the captured kernels name the current population buffer)
    hipGraphExec_t cycle_exec[2] = {nullptr, nullptr};
    int cycle_cur_after[2] = {0, 1};  // `cur` after a replay of cycle_exec[k]
    void drop_graphs();
    unsigned graph_gen = 0;       // bumped by drop_graphs: graphs of groups holding this context are stale

    // learned closure (closure_kind = MLP): the network's weight tiles
    ClosureWeights* d_closure = nullptr;
    bool closure_ready = false;

    ~Ctx();
};

// status code + message for the last-error strings of the C-ABI (ctx may be null)
int report_error(Ctx* c, int code, const std::string& msg);

// ---- launchers (lbm.hip) ------------------------------------------------------
hipError_t launch_force_field(Ctx& c, int step_offset);
// one launch of the fused step in the tuned shape, advancing `steps` steps (up to
// lbm_launch_steps(c))
hipError_t launch_lbm_fused(Ctx& c, const float* fin, float* fout, int steps = 1);
// steps one launch of the tuned shape can take without forcing: 3 for the
// three-step strips (lbm_tiles_per_block <= -LBM_THREE_STEP), 2 for the two-step
// strips (<= -LBM_TWO_STEP), else 1; 1 for any forced flow
int lbm_launch_steps(const Ctx& c);
// the MIXED analysis of several contexts (obs already prepared) in one launch of
// each of its two kernels, on `stream`
hipError_t launch_letkf_mixed_group(Ctx* const* cs, int n, hipStream_t stream);
// one step with the learned closure for tiles [tile0, tile0 + n_tiles) of the
// 64x16 tiling (member-major); nut_out (may be NULL) receives nu_t of those
// tiles' nodes, indexed by node within the member
hipError_t launch_lbm_mlp(Ctx& c, const float* fin, float* fout, int tile0, int n_tiles, float* nut_out);
hipError_t launch_lbm_plain_member(Ctx& c, const float* fin, float* fout, int member);
hipError_t launch_nudge(Ctx& c, const float* ro, const float* uo, const float* vo, float alpha, int xyprune,
                        int skip_east, int skip_north, int use_obs_rho);
hipError_t launch_perturb(Ctx& c, int member, const float* ur, const float* vr, float eps);
hipError_t launch_set_feq(Ctx& c, int member, const float* d_rho, const float* d_u, const float* d_v);
hipError_t launch_macro(Ctx& c, const float* f_member, float* rho, float* u, float* v);
// uv: [M][2][nodes] scratch; out: [6][nodes] (mean u, mean v, mean and sd of the
// vorticity, sd of u and of v)
hipError_t launch_ensemble_fields(Ctx& c, float* uv, float* out);
hipError_t launch_ensemble_stats(Ctx& c, const float* truth_rho, const float* truth_u,
                                 const float* truth_v, double* partials, int n_blocks);
int ensemble_stats_blocks(const Ctx& c);
// the static partial-sum array of lbmletkf_ensemble_stats_noalloc: blocks it holds
// (grids of up to 2^24 nodes) and its device address on the current device
constexpr int STATS_SCRATCH_BLOCKS = 65536;
hipError_t stats_scratch_address(double** out);
hipError_t launch_step_advance(Ctx& c, int n);
bool lbm_fused_supported(int nx, int ny);
// full-row strips of h rows (the alternative shape of the fused step)
bool lbm_rows_supported(int nx, int ny, int h);
// full-row strips of h rows advanced two steps per launch (k_lbm_rows2)
bool lbm_rows2_supported(int nx, int ny, int h);
bool lbm_rows3_supported(int nx, int ny, int h);

// ---- launchers (letkf.hip) ----------------------------------------------------
hipError_t launch_local_count(Ctx& c, int* d_count);
hipError_t launch_local_fill(Ctx& c);
hipError_t launch_obs_moments(Ctx& c, const float* f, double* hx, int member0, int n_members);
hipError_t launch_obs_prepare(Ctx& c);
hipError_t launch_observe(Ctx& c, const double* hx_member, float* yo, uint64_t seed,
                          uint64_t counter, int noise);
hipError_t launch_gram_fp32(Ctx& c);
hipError_t launch_solve_fp64(Ctx& c, float* f, unsigned long long* sweeps);

// ---- launchers (letkf_mixed.hip) ----------------------------------------------
int mixed_width(int M);
// the presets' form: one fused kernel (no FP64 sweeps requested)
bool mixed_fused(const Ctx& c);
long long mixed_fused_blocks(int nx, int ny, int chain);
// Fused form: one kernel; the events are not used; solve_only skips the state update.
// Two-kernel form: ev_mid_end / ev_mid_begin (may be null) are recorded between the
// two kernels; solve_only launches the first kernel alone (the state untouched).
hipError_t launch_letkf_mixed(Ctx& c, float* f, unsigned long long* sweeps32,
                              unsigned long long* sweeps64, int passes, hipEvent_t ev_mid_end,
                              hipEvent_t ev_mid_begin, bool solve_only = false);

}  // namespace lbmletkf
