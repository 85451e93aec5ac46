// C-ABI of the LBM-LETKF cycle: configuration presets, context lifetime, the
// initial conditions (host, once per member) and the orchestration of the
// forecast and analysis kernels on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <new>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "context.hpp"
#include "philox.hpp"

using namespace lbmletkf;

namespace {

thread_local std::string g_thread_error;
constexpr double PI = 3.14159265358979323846;

int fail(Ctx* c, int code, const std::string& msg) {
    if (c) c->last_error = msg;
    g_thread_error = msg;
    return code;
}

}  // namespace

namespace lbmletkf {
int report_error(Ctx* c, int code, const std::string& msg) { return fail(c, code, msg); }
}  // namespace lbmletkf

namespace {

#define HIP_TRY(ctx, expr)                                                              \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail((ctx), LBMLETKF_ERR_HIP,                                        \
                        std::string(#expr) + ": " + hipGetErrorString(e_));             \
    } while (0)

template <class T>
hipError_t dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
}


// ---------------------------------------------------------------------------
// forcing modes, src/data.cu:161-223 (kx, ky in [0, kf+dk), kf-dk <= |k| <= kf+dk,
// amp ~ exp(-(|k|-kf)^2/sigma) normalised so that sum(amp) = fkf*u_ref/t0)
// ---------------------------------------------------------------------------
void build_force_table(const lbmletkf_config& cfg, ForceTable& t) {
    t.n = 0;
    if (cfg.force_kind != LBMLETKF_FORCE_WATANABE97) return;
    const int kmax_search = int(cfg.force_kf + cfg.force_dk);
    long double asum = 0;
    std::vector<float> kx, ky, amp;
    for (int y = 0; y < kmax_search; ++y) {
        for (int x = 0; x < kmax_search; ++x) {
            if (x == 0 && y == 0) continue;
            const double k = std::sqrt(double(y * y + x * x));
            if (cfg.force_kf - cfg.force_dk <= k && k <= cfg.force_kf + cfg.force_dk) {
                const float a = float(std::exp(-(k - cfg.force_kf) * (k - cfg.force_kf) / cfg.force_sigma_k));
                asum += a;
                kx.push_back(float(x));
                ky.push_back(float(y));
                amp.push_back(a);
            }
        }
    }
    const int n = std::min<int>(int(kx.size()), MAX_FORCE_MODES);
    t.n = n;
    for (int m = 0; m < n; ++m) {
        t.kx[m] = kx[m];
        t.ky[m] = ky[m];
        t.amp[m] = float((long double)amp[m] * (long double)cfg.force_amp_total / asum);
    }
}

// ---------------------------------------------------------------------------
// initial velocity fields.  Both kinds are the velocity of a random-phase
// streamfunction psi = sum_k p(k) sin(kx i + ky j + theta_k):
//   u = sum p ky cos(kx i + ky j + theta),  v = -sum p kx cos(...)
// (src/data.cu:76-110), evaluated here by two separable passes in FP64.
// ---------------------------------------------------------------------------
struct Mode { int ki, kj; double p, theta; };

void modes_reference(const lbmletkf_config& cfg, unsigned long long seed, int nx,
                     std::vector<Mode>& modes) {
    // src/data.cu:41-73: std::mt19937(seed), uniform_real_distribution<float>(-1,1),
    // p(k) = exp(-(k-k0)^2/sigma) on k0-dk <= k <= k0+dk, quadrant 0 <= ki,kj < nx/2
    const int nk = nx / 2;
    std::mt19937 eng(static_cast<std::mt19937::result_type>(seed));
    std::uniform_real_distribution<float> dist(-1, 1);
    const float k0 = float(cfg.ic_k0), dk = float(cfg.ic_dk), sigma = float(cfg.ic_sigma);
    for (int kj = 0; kj < nk; ++kj) {
        for (int ki = 0; ki < nk; ++ki) {
            const float k = std::sqrt(float(ki * ki + kj * kj));
            const float pk = (k0 - dk <= k && k <= k0 + dk) ? std::exp(-(k - k0) * (k - k0) / sigma) : 0.f;
            const float theta = float(dist(eng) * PI);
            if (pk != 0.f) modes.push_back({ki, kj, double(pk), double(theta)});
        }
    }
}

void modes_k4(const lbmletkf_config& cfg, unsigned long long seed, int nx, int ny,
              std::vector<Mode>& modes) {
    // E(k) ~ k^4 exp(-2 (k/k0)^2) over the half plane kj >= 0 (one of each +-k
    // pair); streamfunction amplitude p = sqrt(E(k)/k^3) = sqrt(k) exp(-(k/k0)^2).
    const int nk = std::min(nx, ny) / 2;
    for (int kj = 0; kj < nk; ++kj) {
        for (int ki = -nk + 1; ki < nk; ++ki) {
            if (kj == 0 && ki <= 0) continue;
            const double k = std::sqrt(double(ki * ki + kj * kj));
            if (k >= nk) continue;
            u32x4 ctr{{uint32_t(ki + nk), uint32_t(kj), 0u, STREAM_IC}};
            const u32x4 r = philox4x32_10(ctr, uint32_t(seed), uint32_t(seed >> 32));
            const double theta = PI * (2.0 * u32_to_open01(r.v[0]) - 1.0);
            const double p = std::sqrt(k) * std::exp(-(k / cfg.ic_k0) * (k / cfg.ic_k0));
            modes.push_back({ki, kj, p, theta});
        }
    }
}

template <class F>
void parallel_for(int n, F fn) {
    int nt = int(std::thread::hardware_concurrency());
    nt = std::max(1, std::min(nt, 16));
    nt = std::min(nt, n);
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([=, &fn] {
            for (int a = t; a < n; a += nt) fn(a);
        });
    for (auto& x : th) x.join();
}

void velocity_from_modes(const std::vector<Mode>& modes, int nx, int ny, std::vector<double>& u,
                         std::vector<double>& v) {
    using cd = std::complex<double>;
    int kj_max = 0;
    for (const auto& m : modes) kj_max = std::max(kj_max, m.kj);
    const int nkj = kj_max + 1;
    std::vector<cd> ex(nx), ey(ny);
    for (int a = 0; a < nx; ++a) ex[a] = std::polar(1.0, 2 * PI * a / nx);
    for (int a = 0; a < ny; ++a) ey[a] = std::polar(1.0, 2 * PI * a / ny);
    // group modes by kj
    std::vector<std::vector<int>> by_kj(nkj);
    for (int n = 0; n < int(modes.size()); ++n) by_kj[modes[n].kj].push_back(n);
    std::vector<cd> Au(size_t(nkj) * nx), Av(size_t(nkj) * nx);
    parallel_for(nkj, [&](int kj) {
        for (int i = 0; i < nx; ++i) {
            cd su(0, 0), sv(0, 0);
            for (int n : by_kj[kj]) {
                const Mode& m = modes[n];
                const int idx = ((m.ki * i) % nx + nx) % nx;
                const cd e = ex[idx] * std::polar(1.0, m.theta);
                const double kx = 2 * PI * m.ki / nx;
                su += m.p * e;
                sv += -m.p * kx * e;
            }
            Au[size_t(kj) * nx + i] = su;
            Av[size_t(kj) * nx + i] = sv;
        }
    });
    u.assign(size_t(nx) * ny, 0.0);
    v.assign(size_t(nx) * ny, 0.0);
    parallel_for(ny, [&](int j) {
        for (int i = 0; i < nx; ++i) {
            double su = 0, sv = 0;
            for (int kj = 0; kj < nkj; ++kj) {
                if (by_kj[kj].empty()) continue;
                const cd e = ey[(size_t(kj) * j) % ny];
                const double ky = 2 * PI * kj / ny;
                su += ky * (Au[size_t(kj) * nx + i] * e).real();
                sv += (Av[size_t(kj) * nx + i] * e).real();
            }
            u[size_t(j) * nx + i] = su;
            v[size_t(j) * nx + i] = sv;
        }
    });
}

int check_config(const lbmletkf_config* cfg, int n_members, std::string& why) {
    auto bad = [&](const std::string& m) { why = m; return LBMLETKF_ERR_ARG; };
    if (!cfg) return bad("config is NULL");
    if (n_members < 1 || n_members > MAX_MEMBERS) return bad("n_members must be in [1, 64]");
    if (cfg->nx < 4 || cfg->ny < 4) return bad("nx and ny must be >= 4");
    if ((long long)cfg->nx * cfg->ny > (1LL << 30)) return bad("grid too large");
    if (!(cfg->c > 0 && cfg->dx > 0 && cfg->dt > 0)) return bad("c, dx and dt must be positive");
    if (!(cfg->nu >= 0) || !(cfg->cs_smag >= 0)) return bad("nu and cs_smag must be >= 0");
    if (!(cfg->rho_ref > 0)) return bad("rho_ref must be positive");
    if (cfg->force_kind != LBMLETKF_FORCE_NONE && cfg->force_kind != LBMLETKF_FORCE_WATANABE97)
        return bad("unknown force_kind");
    if (cfg->ic_kind != LBMLETKF_IC_REFERENCE && cfg->ic_kind != LBMLETKF_IC_K4_SPECTRUM)
        return bad("unknown ic_kind");
    if (cfg->obs_stride < 1 || cfg->nx % cfg->obs_stride || cfg->ny % cfg->obs_stride)
        return bad("obs_stride must divide nx and ny (src/letkf.cu:30)");
    if (cfg->obs_offset < 0 || cfg->obs_offset >= cfg->obs_stride) return bad("obs_offset must be in [0, obs_stride)");
    if (cfg->obs_vars <= 0 || cfg->obs_vars > 7) return bad("obs_vars must be a non-empty mask of u|v|rho");
    if (!(cfg->obs_sigma_u > 0) || !(cfg->obs_sigma_rho > 0)) return bad("observation errors must be positive");
    if (cfg->loc_kind == LBMLETKF_LOC_GC_WINDOW) {
        if (cfg->obs_offset != 0) return bad("the GC window localization needs obs_offset == 0");
        if (cfg->loc_rloc_len < 0) return bad("loc_rloc_len must be >= 0");
    } else if (cfg->loc_kind == LBMLETKF_LOC_GAUSS) {
        if (!(cfg->loc_scale > 0) || !(cfg->loc_cutoff >= 0)) return bad("Gaussian localization needs scale > 0, cutoff >= 0");
        const int R = int(std::floor(cfg->loc_cutoff));
        if (2 * R + 1 > std::min(cfg->nx, cfg->ny)) return bad("localization cutoff must be below half the domain");
    } else {
        return bad("unknown loc_kind");
    }
    if (!(cfg->inflation > 0)) return bad("inflation must be positive");
    if (cfg->da_interval < 1) return bad("da_interval must be >= 1");
    if (cfg->variant != LBMLETKF_VARIANT_PLAIN && cfg->variant != LBMLETKF_VARIANT_MIXED)
        return bad("unknown variant");
    if (!(cfg->jacobi_tol > 0) || cfg->jacobi_max_sweeps < 1) return bad("jacobi_tol > 0 and jacobi_max_sweeps >= 1 required");
    if (cfg->gram_passes != 1 && cfg->gram_passes != 3) return bad("gram_passes must be 1 or 3");
    if (cfg->jacobi_fp64_sweeps < 0 || cfg->jacobi_fp64_sweeps > 20) return bad("jacobi_fp64_sweeps must be in [0, 20]");
    if (!(cfg->jacobi_tol32 >= 0 && cfg->jacobi_tol32 < 1)) return bad("jacobi_tol32 must be in [0, 1)");
    if (cfg->closure_kind != LBMLETKF_CLOSURE_SMAGORINSKY && cfg->closure_kind != LBMLETKF_CLOSURE_MLP)
        return bad("unknown closure_kind");
    if (cfg->closure_kind == LBMLETKF_CLOSURE_MLP &&
        !(cfg->closure_strain_scale > 0 && cfg->closure_nu_scale > 0 && cfg->closure_nu_max > 0))
        return bad("the MLP closure needs positive strain, nu and nu_max scales");
    if (cfg->closure_kind == LBMLETKF_CLOSURE_MLP && (cfg->nx % 64 || cfg->ny % 16))
        return bad("the MLP closure needs nx a multiple of 64 and ny a multiple of 16");
    return LBMLETKF_OK;
}

LbmParams make_lbm_params(const lbmletkf_config& cfg, int M) {
    LbmParams p{};
    p.nx = cfg.nx;
    p.ny = cfg.ny;
    p.n_members = M;
    p.nodes = (long long)cfg.nx * cfg.ny;
    p.member_stride = Q * p.nodes;
    p.c = float(cfg.c);
    p.inv_c = float(1.0 / cfg.c);
    p.inv_c2 = float(1.0 / (cfg.c * cfg.c));
    p.k_tau = float(3.0 / (cfg.c * cfg.c * cfg.dt));
    p.nu = float(cfg.nu);
    p.les_coef = float(cfg.cs_smag * cfg.cs_smag * cfg.dx * cfg.dx);
    p.inv_2dx = float(1.0 / (2.0 * cfg.dx));
    p.force_scale = float(cfg.u_ref * 3.0 / (cfg.c * cfg.c));
    p.fric_over_dt = float(cfg.friction_rate / cfg.dt);
    p.dt = float(cfg.dt);
    p.has_force = cfg.force_kind != LBMLETKF_FORCE_NONE;
    return p;
}

int sync_and_check(Ctx* c) {
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    int st = 0;
    HIP_TRY(c, hipMemcpy(&st, c->d_status, sizeof(int), hipMemcpyDeviceToHost));
    if (st) {
        HIP_TRY(c, hipMemset(c->d_status, 0, sizeof(int)));
        std::string m;
        if (st & 1) m += "non-finite analysis state (the reference traps the same condition, src/util/letkf_solver.h:352-364); ";
        if (st & 2) m += "non-finite observation at a used site; ";
        return fail(c, LBMLETKF_ERR_NUMERIC, m);
    }
    return LBMLETKF_OK;
}

int setup_obs_impl(Ctx* c) {
    if (c->obs_ready) return LBMLETKF_OK;
    const auto& cfg = c->cfg;
    ObsNet& o = c->obs;
    o.n_vars = 0;
    const double isu = 1.0 / (cfg.obs_sigma_u * cfg.obs_sigma_u);
    const double isr = 1.0 / (cfg.obs_sigma_rho * cfg.obs_sigma_rho);
    if (cfg.obs_vars & LBMLETKF_OBS_U) { o.var_id[o.n_vars] = 0; o.inv_sigma2[o.n_vars++] = float(isu); }
    if (cfg.obs_vars & LBMLETKF_OBS_V) { o.var_id[o.n_vars] = 1; o.inv_sigma2[o.n_vars++] = float(isu); }
    if (cfg.obs_vars & LBMLETKF_OBS_RHO) { o.var_id[o.n_vars] = 2; o.inv_sigma2[o.n_vars++] = float(isr); }
    if (o.custom) {
        // explicit site list: the site map lets the localization scan find them
        o.stride = 1;
        o.offset = 0;
        o.n_sites = int(o.site_nodes.size());
        o.n_sx = o.n_sites;
        o.n_sy = 1;
        std::vector<int> map(c->nodes, -1);
        for (int s = 0; s < o.n_sites; ++s) map[o.site_nodes[s]] = s;
        HIP_TRY(c, dalloc(&c->d_site_map, c->nodes));
        HIP_TRY(c, hipMemcpy(c->d_site_map, map.data(), sizeof(int) * c->nodes, hipMemcpyHostToDevice));
    } else {
        o.stride = cfg.obs_stride;
        o.offset = cfg.obs_offset;
        o.n_sx = c->nx / o.stride;
        o.n_sy = c->ny / o.stride;
        o.n_sites = o.n_sx * o.n_sy;
        o.site_nodes.resize(o.n_sites);
        for (int sy = 0; sy < o.n_sy; ++sy)
            for (int sx = 0; sx < o.n_sx; ++sx)
                o.site_nodes[sx + sy * o.n_sx] = (sx * o.stride + o.offset) + (sy * o.stride + o.offset) * c->nx;
    }
    const int n_obs = o.n_sites * o.n_vars;
    HIP_TRY(c, dalloc(&c->d_site_node, std::max(o.n_sites, 1)));
    if (o.n_sites)
        HIP_TRY(c, hipMemcpy(c->d_site_node, o.site_nodes.data(), sizeof(int) * o.n_sites, hipMemcpyHostToDevice));

    // local sets: count on device, scan on host, fill on device
    int* d_count = nullptr;
    HIP_TRY(c, dalloc(&d_count, c->nodes));
    HIP_TRY(c, launch_local_count(*c, d_count));
    std::vector<int> count(c->nodes);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(count.data(), d_count, sizeof(int) * c->nodes, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipFree(d_count));
    std::vector<int> ptr(c->nodes + 1);
    long long acc = 0;
    int mx = 0;
    for (long long g = 0; g < c->nodes; ++g) {
        ptr[g] = int(acc);
        acc += count[g];
        mx = std::max(mx, count[g]);
        if (acc > (1LL << 31) - 1) return fail(c, LBMLETKF_ERR_ARG, "local observation sets exceed 2^31 entries");
    }
    ptr[c->nodes] = int(acc);
    o.max_local = mx;
    o.n_entries = acc;
    HIP_TRY(c, dalloc(&c->d_loc_ptr, c->nodes + 1));
    HIP_TRY(c, hipMemcpy(c->d_loc_ptr, ptr.data(), sizeof(int) * (c->nodes + 1), hipMemcpyHostToDevice));
    HIP_TRY(c, dalloc(&c->d_loc_site, std::max(acc, 1LL)));
    HIP_TRY(c, dalloc(&c->d_loc_w, std::max(acc, 1LL)));
    HIP_TRY(c, launch_local_fill(*c));

    float coef[3] = {o.inv_sigma2[0], o.inv_sigma2[1], o.inv_sigma2[2]};
    HIP_TRY(c, dalloc(&c->d_obs_coef, 3));
    HIP_TRY(c, hipMemcpy(c->d_obs_coef, coef, sizeof(coef), hipMemcpyHostToDevice));
    const int n_obs1 = std::max(n_obs, 1);
    HIP_TRY(c, dalloc(&c->d_yo, n_obs1));
    HIP_TRY(c, dalloc(&c->d_hx, (size_t)n_obs1 * c->M));
    HIP_TRY(c, dalloc(&c->d_hx1, n_obs1));
    HIP_TRY(c, dalloc(&c->d_dy, (size_t)n_obs1 * c->M));
    HIP_TRY(c, dalloc(&c->d_dyo, n_obs1));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->obs_ready = true;
    return LBMLETKF_OK;
}

int ensure_plain_buffers(Ctx* c) {
    if (c->d_gram) return LBMLETKF_OK;
    HIP_TRY(c, dalloc(&c->d_gram, (size_t)c->nodes * c->M * c->M));
    HIP_TRY(c, dalloc(&c->d_bvec, (size_t)c->nodes * c->M));
    return LBMLETKF_OK;
}

// the MIXED variant's workspace (allocated outside any capture): U and b between
// the two kernels of the form with FP64 sweeps; the fused kernel needs none
int ensure_mixed_buffers(Ctx* c) {
    const int N = mixed_width(c->M);
    if (!mixed_fused(*c) && !(c->d_mix_u && c->mix_width == N)) {
        HIP_TRY(c, dalloc(&c->d_mix_u, (size_t)c->nodes * N * N));
        HIP_TRY(c, dalloc(&c->d_mix_b, (size_t)c->nodes * N));
    }
    c->mix_width = N;
    return LBMLETKF_OK;
}

int ensure_scratch(Ctx* c) {
    if (c->d_truth) return LBMLETKF_OK;
    HIP_TRY(c, dalloc(&c->d_truth, 3 * c->nodes));
    c->red_cap = ensemble_stats_blocks(*c) * 4;
    HIP_TRY(c, dalloc(&c->d_red, c->red_cap));
    return LBMLETKF_OK;
}

// ---- phase timers: an event pair per phase occurrence, summed lazily -----------
hipEvent_t prof_event(Ctx* c) {
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

hipEvent_t prof_begin(Ctx* c) {
    if (!c->prof) return nullptr;
    hipEvent_t a = prof_event(c);
    if (a) (void)hipEventRecord(a, c->stream);
    return a;
}

void prof_end(Ctx* c, int phase, hipEvent_t a, long long launches, long long work) {
    if (!c->prof || !a) return;
    hipEvent_t b = prof_event(c);
    if (!b) return;
    (void)hipEventRecord(b, c->stream);
    c->prof_pending.push_back({phase, a, b, launches, work});
}

int prof_collect(Ctx* c) {
    for (const ProfRec& r : c->prof_pending) {
        HIP_TRY(c, hipEventSynchronize(r.b));
        float ms = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&ms, r.a, r.b));
        c->prof_sec[r.phase] += 1e-3 * ms;
        c->prof_launch[r.phase] += r.launches;
        c->prof_work[r.phase] += r.work;
        c->event_pool.push_back(r.a);
        c->event_pool.push_back(r.b);
    }
    c->prof_pending.clear();
    return LBMLETKF_OK;
}

// n LBM steps of all members (src/lbm.h:16-29 applied n times): stream-ordered
// launches only, so a cycle can be captured into a graph.  Advances cur and step.
int forecast_launches(Ctx* c, int n_steps) {
    const bool mlp = c->cfg.closure_kind == LBMLETKF_CLOSURE_MLP;
    if (mlp && !c->closure_ready) return fail(c, LBMLETKF_ERR_STATE, "closure_kind is MLP: call lbmletkf_set_closure first");
    const bool fused = mlp || (c->cfg.variant == LBMLETKF_VARIANT_MIXED && lbm_fused_supported(c->nx, c->ny));
    hipEvent_t ev = prof_begin(c);
    // the two- and three-step strips take several steps per launch (no forcing: no
    // per-step field); the last steps of a forecast take what is left
    const int per_launch = (fused && !mlp) ? lbm_launch_steps(*c) : 1;
    long long launches = 0;
    for (int s = 0; s < n_steps;) {
        const int k = std::min(per_launch, n_steps - s);
        if (c->lp.has_force) HIP_TRY(c, launch_force_field(*c, s));
        const float* fin = c->f[c->cur];
        float* fout = c->f[1 - c->cur];
        if (mlp) {
            HIP_TRY(c, launch_lbm_mlp(*c, fin, fout, 0, (c->nx / 64) * (c->ny / 16) * c->M, nullptr));
            ++launches;
        } else if (fused) {
            HIP_TRY(c, launch_lbm_fused(*c, fin, fout, k));
            ++launches;
        } else {
            for (int m = 0; m < c->M; ++m) HIP_TRY(c, launch_lbm_plain_member(*c, fin, fout, m));
            launches += 3LL * c->M;
        }
        if (c->lp.has_force) ++launches;
        c->cur = 1 - c->cur;
        s += k;
    }
    if (n_steps > 0) HIP_TRY(c, launch_step_advance(*c, n_steps));
    prof_end(c, PH_FORECAST, ev, launches + (n_steps > 0), (long long)n_steps * c->M * c->nodes);
    c->step += n_steps;
    return LBMLETKF_OK;
}

// The analysis kernels on the observation vector already in d_yo (the observation
// network must be set up): stream-ordered launches only, apart from the one-time
// allocation of the PLAIN variant's per-gridpoint matrices (that variant is never
// captured into a graph).
int analysis_launches(Ctx* c) {
    if (c->cfg.variant == LBMLETKF_VARIANT_PLAIN)
        if (int st = ensure_plain_buffers(c)) return st;
    const int n_obs = c->obs.n_sites * c->obs.n_vars;
    hipEvent_t e0 = prof_begin(c);
    HIP_TRY(c, launch_obs_moments(*c, c->f[c->cur], c->d_hx, 0, c->M));
    HIP_TRY(c, launch_obs_prepare(*c));
    prof_end(c, PH_OBS, e0, 2, n_obs);
    if (c->cfg.variant == LBMLETKF_VARIANT_MIXED && mixed_fused(*c)) {
        // k_mixed_analysis: gather, MFMA Gram, FP32 sweeps, assembly and update in
        // one kernel, without workspace
        hipEvent_t e2 = prof_begin(c);
        HIP_TRY(c, launch_letkf_mixed(*c, c->f[c->cur], c->prof ? c->d_sweeps : nullptr, nullptr,
                                      c->cfg.gram_passes, nullptr, nullptr));
        prof_end(c, PH_SOLVE, e2, 1, 0);
        c->dump_kind = 1;
        return LBMLETKF_OK;
    }
    if (c->cfg.variant == LBMLETKF_VARIANT_MIXED) {
        // k_mixed_solve (gather + MFMA Gram + FP32 sweeps), then k_mixed_finish
        // (FP64 sweeps + assembly + update); the workspace is allocated by the caller
        hipEvent_t e2 = prof_begin(c);
        hipEvent_t mid_end = c->prof ? prof_event(c) : nullptr;
        hipEvent_t mid_begin = c->prof ? prof_event(c) : nullptr;
        HIP_TRY(c, launch_letkf_mixed(*c, c->f[c->cur], c->prof ? c->d_sweeps : nullptr,
                                      c->prof ? c->d_sweeps64 : nullptr, c->cfg.gram_passes, mid_end,
                                      mid_begin));
        if (c->prof && e2 && mid_end && mid_begin) {
            c->prof_pending.push_back({PH_SOLVE, e2, mid_end, 1, 0});
            prof_end(c, PH_FINISH, mid_begin, 1, 0);
        }
        c->dump_kind = 1;
        return LBMLETKF_OK;
    }
    hipEvent_t e1 = prof_begin(c);
    HIP_TRY(c, launch_gram_fp32(*c));
    prof_end(c, PH_GRAM, e1, 1, c->obs.n_entries * c->obs.n_vars);
    hipEvent_t e2 = prof_begin(c);
    HIP_TRY(c, launch_solve_fp64(*c, c->f[c->cur], c->prof ? c->d_sweeps : nullptr));
    prof_end(c, PH_SOLVE, e2, 1, 0);
    c->dump_kind = 0;
    return LBMLETKF_OK;
}

// One MIXED cycle (analysis on d_yo, then da_interval steps) replayed from a graph.
// The graph is captured on first use for the current parity of `cur` and rebuilt
// after anything that changes a captured argument (dump range, observation buffers).
// The captured work contains kernels only: no memset or memcpy nodes.
int cycle_graph(Ctx* c) {
    const int k = c->cur;
    if (!c->cycle_exec[k]) {
        const int cur0 = c->cur;
        const long long step0 = c->step;
        HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int st = analysis_launches(c);
        if (st == LBMLETKF_OK) st = forecast_launches(c, c->cfg.da_interval);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(c->stream, &g);
        c->cycle_cur_after[k] = c->cur;
        c->cur = cur0;
        c->step = step0;
        if (st != LBMLETKF_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        HIP_TRY(c, ec);
        hipGraphExec_t x = nullptr;
        const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIP_TRY(c, ei);
        c->cycle_exec[k] = x;
    }
    HIP_TRY(c, hipGraphLaunch(c->cycle_exec[k], c->stream));
    c->cur = c->cycle_cur_after[k];
    c->step += c->cfg.da_interval;
    c->dump_kind = 1;
    return LBMLETKF_OK;
}

#define CHECK_CTX(c) \
    do { if (!(c)) return fail(nullptr, LBMLETKF_ERR_ARG, "context is NULL"); } while (0)
#define CHECK_MEMBER(c, m) \
    do { if ((m) < 0 || (m) >= (c)->M) return fail((c), LBMLETKF_ERR_ARG, "member index out of range"); } while (0)

}  // namespace

namespace lbmletkf {
void Ctx::drop_graphs() {
    ++graph_gen;
    if ((cycle_exec[0] || cycle_exec[1]) && stream) (void)hipStreamSynchronize(stream);
    for (hipGraphExec_t& x : cycle_exec) {
        if (x) (void)hipGraphExecDestroy(x);
        x = nullptr;
    }
}

Ctx::~Ctx() {
    drop_graphs();
    void* ptrs[] = {f[0], f[1], plain_ruv, plain_nus, force_xy, d_force, d_step, d_status,
                    d_site_node, d_loc_ptr, d_loc_site, d_loc_w, d_obs_coef, d_yo, d_hx, d_hx1, d_dy,
                    d_dyo, d_gram, d_bvec, d_mix_u, d_mix_b, d_dump, d_red, d_truth, d_sweeps,
                    d_sweeps64, d_closure, d_site_map, d_ens_uv, d_ens_out};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (const ProfRec& r : prof_pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
    if (stream && own_stream) (void)hipStreamDestroy(stream);
}
}  // namespace lbmletkf

extern "C" {

const char* lbmletkf_version(void) { return "lbmletkf 0.1 (gfx950)"; }

int lbmletkf_config_preset(const char* name, lbmletkf_config* cfg) {
    if (!name || !cfg) return fail(nullptr, LBMLETKF_ERR_ARG, "NULL argument");
    lbmletkf_config c{};
    const std::string n(name);
    if (n == "ref") {
        // src/config.h at SCALER=4, src/lbm.cu:59 (Cs=0.2), src/data.cu:161-350
        const double u_ref = 1.0, cfl = 0.05, h_ref = 2 * PI;
        c.nx = c.ny = 64 * 4;
        c.c = u_ref / cfl;
        c.dx = h_ref / (c.ny - 1);
        c.dt = c.dx / c.c;
        c.nu = 1e-4;
        c.rho_ref = 1.0;
        c.cs_smag = 0.2;
        c.u_ref = u_ref;
        c.friction_rate = 5e-4;
        c.force_kind = LBMLETKF_FORCE_WATANABE97;
        c.force_kf = 4;
        c.force_amp_total = 5.6 * u_ref / 1.0;
        c.force_sigma_k = 1;
        c.force_dk = 2;
        c.force_rand_stdev = 0.84089641525;
        c.force_seed = 1991;
        c.ic_kind = LBMLETKF_IC_REFERENCE;
        c.ic_amp = 0.01 * u_ref;
        c.ic_k0 = 4;
        c.ic_dk = 10;
        c.ic_sigma = 5;
        c.obs_stride = 1;
        c.obs_offset = 0;
        c.obs_vars = LBMLETKF_OBS_U | LBMLETKF_OBS_V | LBMLETKF_OBS_RHO;
        c.obs_sigma_u = 0.1;     // run/jserr:146-147
        c.obs_sigma_rho = 0.01;
        c.loc_kind = LBMLETKF_LOC_GC_WINDOW;
        c.loc_rloc_len = 0;
        c.loc_scale = 0;
        c.loc_cutoff = 0;
        c.inflation = 1.0;
        c.da_interval = 5 * 4 * 10;  // iiter * DAPRUNE
        c.variant = LBMLETKF_VARIANT_PLAIN;
        c.jacobi_tol = 1e-10;
        c.jacobi_max_sweeps = 100;
        c.gram_passes = 3;
        c.jacobi_fp64_sweeps = 0;  // the FP32 phase reaches the FP32 floor; FP64 sweeps optional
        c.closure_kind = LBMLETKF_CLOSURE_SMAGORINSKY;
        c.closure_strain_scale = 1e-2 * c.c / c.dx;  // ~ u_rms k0 2pi/n at 256^2, k0 = 8
        c.closure_nu_scale = 1e-4 * c.c * c.dx;      // Smagorinsky nu_t at that strain, Cs dx = 0.1 dx
        c.closure_nu_max = 0.05 * c.c * c.dx;
    } else if (n == "baseline") {
        c.nx = c.ny = 256;
        c.c = 1.0;
        c.dx = 1.0;
        c.dt = 1.0;
        c.nu = (0.503 - 0.5) / 3.0;
        c.rho_ref = 1.0;
        c.cs_smag = 0.1;
        c.u_ref = 1.0;
        c.friction_rate = 0.0;
        c.force_kind = LBMLETKF_FORCE_NONE;
        c.force_kf = 4;
        c.force_amp_total = 0;
        c.force_sigma_k = 1;
        c.force_dk = 2;
        c.force_rand_stdev = 0.84089641525;
        c.force_seed = 1991;
        c.ic_kind = LBMLETKF_IC_K4_SPECTRUM;
        c.ic_amp = 0.05;
        c.ic_k0 = 8;
        c.ic_dk = 0;
        c.ic_sigma = 0;
        c.obs_stride = 8;
        c.obs_offset = 0;
        c.obs_vars = LBMLETKF_OBS_U | LBMLETKF_OBS_V;
        c.obs_sigma_u = 0.2 * 0.05;
        c.obs_sigma_rho = 0.01;
        c.loc_kind = LBMLETKF_LOC_GAUSS;
        c.loc_rloc_len = 0;
        c.loc_scale = 12.0;
        c.loc_cutoff = 43.8;
        c.inflation = 1.1;
        c.da_interval = 100;
        c.variant = LBMLETKF_VARIANT_MIXED;
        c.jacobi_tol = 1e-10;
        c.jacobi_max_sweeps = 100;
        c.gram_passes = 3;
        c.jacobi_fp64_sweeps = 0;  // the FP32 phase reaches the FP32 floor; FP64 sweeps optional
        c.closure_kind = LBMLETKF_CLOSURE_SMAGORINSKY;
        c.closure_strain_scale = 1e-2 * c.c / c.dx;  // ~ u_rms k0 2pi/n at 256^2, k0 = 8
        c.closure_nu_scale = 1e-4 * c.c * c.dx;      // Smagorinsky nu_t at that strain, Cs dx = 0.1 dx
        c.closure_nu_max = 0.05 * c.c * c.dx;
    } else {
        return fail(nullptr, LBMLETKF_ERR_ARG, "unknown preset: " + n);
    }
    *cfg = c;
    return LBMLETKF_OK;
}

int lbmletkf_config_check(const lbmletkf_config* cfg, int n_members) {
    std::string why;
    const int st = check_config(cfg, n_members, why);
    if (st) return fail(nullptr, st, why);
    return LBMLETKF_OK;
}

const char* lbmletkf_last_error(const lbmletkf_ctx* ctx) {
    if (ctx) return reinterpret_cast<const Ctx*>(ctx)->last_error.c_str();
    return g_thread_error.c_str();
}

int lbmletkf_create(const lbmletkf_config* cfg, int n_members, lbmletkf_ctx** out) {
    if (!out) return fail(nullptr, LBMLETKF_ERR_ARG, "out is NULL");
    *out = nullptr;
    std::string why;
    if (int st = check_config(cfg, n_members, why)) return fail(nullptr, st, why);
    Ctx* c = new (std::nothrow) Ctx();
    if (!c) return fail(nullptr, LBMLETKF_ERR_ARG, "out of host memory");
    c->cfg = *cfg;
    c->M = n_members;
    c->nx = cfg->nx;
    c->ny = cfg->ny;
    c->nodes = (long long)cfg->nx * cfg->ny;
    c->lp = make_lbm_params(*cfg, n_members);
    build_force_table(*cfg, c->force);
    auto bail = [&](int st) { g_thread_error = c->last_error; delete c; return st; };
#define CREATE_TRY(expr)                                                                      \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            fail(c, LBMLETKF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
            return bail(LBMLETKF_ERR_HIP);                                                    \
        }                                                                                     \
    } while (0)
    CREATE_TRY(hipGetDevice(&c->device));
    CREATE_TRY(stats_scratch_address(&c->stats_scratch));
    CREATE_TRY(hipStreamCreate(&c->stream));
    const size_t fsz = size_t(c->M) * Q * c->nodes;
    CREATE_TRY(dalloc(&c->f[0], fsz));
    CREATE_TRY(dalloc(&c->f[1], fsz));
    CREATE_TRY(dalloc(&c->plain_ruv, 3 * c->nodes));
    CREATE_TRY(dalloc(&c->plain_nus, c->nodes));
    CREATE_TRY(dalloc(&c->force_xy, 2 * c->nodes));
    CREATE_TRY(hipMemset(c->force_xy, 0, sizeof(float) * 2 * c->nodes));
    CREATE_TRY(dalloc(&c->d_force, 1));
    CREATE_TRY(hipMemcpy(c->d_force, &c->force, sizeof(ForceTable), hipMemcpyHostToDevice));
    CREATE_TRY(dalloc(&c->d_step, 1));
    CREATE_TRY(hipMemset(c->d_step, 0, sizeof(long long)));
    CREATE_TRY(dalloc(&c->d_status, 1));
    CREATE_TRY(hipMemset(c->d_status, 0, sizeof(int)));
    // every member at rest: f = feq(rho_ref, 0, 0)
    {
        float *r = nullptr, *z = nullptr;
        CREATE_TRY(dalloc(&r, c->nodes));
        CREATE_TRY(dalloc(&z, c->nodes));
        std::vector<float> ones(c->nodes, float(cfg->rho_ref));
        CREATE_TRY(hipMemcpy(r, ones.data(), sizeof(float) * c->nodes, hipMemcpyHostToDevice));
        CREATE_TRY(hipMemset(z, 0, sizeof(float) * c->nodes));
        for (int m = 0; m < c->M; ++m) CREATE_TRY(launch_set_feq(*c, m, r, z, z));
        CREATE_TRY(hipStreamSynchronize(c->stream));
        CREATE_TRY(hipFree(r));
        CREATE_TRY(hipFree(z));
    }
#undef CREATE_TRY
    *out = reinterpret_cast<lbmletkf_ctx*>(c);
    return LBMLETKF_OK;
}

int lbmletkf_destroy(lbmletkf_ctx* ctx) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c) return LBMLETKF_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete c;
    return LBMLETKF_OK;
}

int lbmletkf_set_macro(lbmletkf_ctx* ctx, int member, const float* rho, const float* u,
                       const float* v) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    CHECK_MEMBER(c, member);
    if (!rho || !u || !v) return fail(c, LBMLETKF_ERR_ARG, "rho, u and v are required");
    if (int st = ensure_scratch(c)) return st;
    const size_t n = size_t(c->nodes);
    HIP_TRY(c, hipMemcpyAsync(c->d_truth, rho, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->d_truth + n, u, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->d_truth + 2 * n, v, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, launch_set_feq(*c, member, c->d_truth, c->d_truth + n, c->d_truth + 2 * n));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LBMLETKF_OK;
}

int lbmletkf_nudge(lbmletkf_ctx* ctx, const float* rho, const float* u, const float* v, int on_device,
                   double alpha, int xyprune, int quadra) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    if (!u || !v) return fail(c, LBMLETKF_ERR_ARG, "u and v are required");
    if (!std::isfinite(alpha)) return fail(c, LBMLETKF_ERR_ARG, "alpha must be finite");
    if (xyprune < 1 || xyprune > c->nx || xyprune > c->ny) return fail(c, LBMLETKF_ERR_ARG, "xyprune out of range");
    if (quadra < 0 || quadra > 3) return fail(c, LBMLETKF_ERR_ARG, "quadra must be 0..3");
    const float *dr = rho, *du = u, *dv = v;
    if (!on_device) {
        if (int st = ensure_scratch(c)) return st;
        const size_t n = size_t(c->nodes);
        if (rho) HIP_TRY(c, hipMemcpyAsync(c->d_truth, rho, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(c->d_truth + n, u, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(c->d_truth + 2 * n, v, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
        dr = rho ? c->d_truth : nullptr;
        du = c->d_truth + n;
        dv = c->d_truth + 2 * n;
    }
    HIP_TRY(c, launch_nudge(*c, dr, du, dv, float(alpha), xyprune, quadra % 2 == 1, quadra >= 2, rho != nullptr));
    if (!on_device) HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LBMLETKF_OK;
}

int lbmletkf_perturb(lbmletkf_ctx* ctx, int member, unsigned long long ensemble_id, double epsilon) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    CHECK_MEMBER(c, member);
    if (!std::isfinite(epsilon)) return fail(c, LBMLETKF_ERR_ARG, "epsilon must be finite");
    if (int st = ensure_scratch(c)) return st;
    // src/data.cu:432-447: std::mt19937(id + 132333), normal_distribution<float>(0, 1),
    // u then v per node, rows j outer, scaled by u_ref
    std::mt19937 eng(uint32_t(ensemble_id + 132333ull));
    std::normal_distribution<float> nd(0.f, 1.f);
    const size_t n = size_t(c->nodes);
    std::vector<float> uv(2 * n);
    const float u_ref = float(c->cfg.u_ref);
    for (size_t ij = 0; ij < n; ++ij) {
        uv[ij] = u_ref * nd(eng);
        uv[n + ij] = u_ref * nd(eng);
    }
    HIP_TRY(c, hipMemcpyAsync(c->d_truth, uv.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, launch_perturb(*c, member, c->d_truth, c->d_truth + n, float(epsilon)));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return LBMLETKF_OK;
}

int lbmletkf_init_member(lbmletkf_ctx* ctx, int member, unsigned long long seed) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    CHECK_MEMBER(c, member);
    const int nx = c->nx, ny = c->ny;
    std::vector<Mode> modes;
    if (c->cfg.ic_kind == LBMLETKF_IC_REFERENCE) modes_reference(c->cfg, seed, nx, modes);
    else modes_k4(c->cfg, seed, nx, ny, modes);
    std::vector<double> u, v;
    velocity_from_modes(modes, nx, ny, u, v);
    const size_t n = size_t(nx) * ny;
    double scale = 0.0;
    if (c->cfg.ic_kind == LBMLETKF_IC_REFERENCE) {
        // src/data.cu:112-120: scale to max speed p_amp*u_ref
        double vmax = 1e-30 * c->cfg.u_ref;
        for (size_t a = 0; a < n; ++a) vmax = std::max(vmax, std::sqrt(u[a] * u[a] + v[a] * v[a]));
        scale = c->cfg.ic_amp / vmax;
    } else {
        double e = 0.0;
        for (size_t a = 0; a < n; ++a) e += u[a] * u[a] + v[a] * v[a];
        const double rms = std::sqrt(e / double(n));
        scale = rms > 0 ? c->cfg.ic_amp / rms : 0.0;
    }
    std::vector<float> rr(n, float(c->cfg.rho_ref)), uf(n), vf(n);
    for (size_t a = 0; a < n; ++a) {
        uf[a] = float(u[a] * scale);
        vf[a] = float(v[a] * scale);
    }
    return lbmletkf_set_macro(ctx, member, rr.data(), uf.data(), vf.data());
}

int lbmletkf_set_state(lbmletkf_ctx* ctx, int member, const float* f_host) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    CHECK_MEMBER(c, member);
    if (!f_host) return fail(c, LBMLETKF_ERR_ARG, "f is NULL");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(c->f[c->cur] + (size_t)member * c->lp.member_stride, f_host,
                         sizeof(float) * c->lp.member_stride, hipMemcpyHostToDevice));
    return LBMLETKF_OK;
}

int lbmletkf_get_state(lbmletkf_ctx* ctx, int member, float* f_host) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    CHECK_MEMBER(c, member);
    if (!f_host) return fail(c, LBMLETKF_ERR_ARG, "f is NULL");
    if (int st = sync_and_check(c)) return st;
    HIP_TRY(c, hipMemcpy(f_host, c->f[c->cur] + (size_t)member * c->lp.member_stride,
                         sizeof(float) * c->lp.member_stride, hipMemcpyDeviceToHost));
    return LBMLETKF_OK;
}

int lbmletkf_get_macro_device(lbmletkf_ctx* ctx, int member, float* rho, float* u, float* v) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    CHECK_MEMBER(c, member);
    HIP_TRY(c, launch_macro(*c, c->f[c->cur] + (size_t)member * c->lp.member_stride, rho, u, v));
    return LBMLETKF_OK;
}

int lbmletkf_get_macro(lbmletkf_ctx* ctx, int member, float* rho, float* u, float* v) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    CHECK_MEMBER(c, member);
    if (int st = ensure_scratch(c)) return st;
    const size_t n = size_t(c->nodes);
    HIP_TRY(c, launch_macro(*c, c->f[c->cur] + (size_t)member * c->lp.member_stride, c->d_truth,
                            c->d_truth + n, c->d_truth + 2 * n));
    if (int st = sync_and_check(c)) return st;
    if (rho) HIP_TRY(c, hipMemcpy(rho, c->d_truth, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (u) HIP_TRY(c, hipMemcpy(u, c->d_truth + n, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (v) HIP_TRY(c, hipMemcpy(v, c->d_truth + 2 * n, sizeof(float) * n, hipMemcpyDeviceToHost));
    return LBMLETKF_OK;
}

int lbmletkf_forecast(lbmletkf_ctx* ctx, int n_steps) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    if (n_steps < 0) return fail(c, LBMLETKF_ERR_ARG, "n_steps must be >= 0");
    return forecast_launches(c, n_steps);
}

int lbmletkf_set_step(lbmletkf_ctx* ctx, long long step) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(c->d_step, &step, sizeof(long long), hipMemcpyHostToDevice));
    c->step = step;
    return LBMLETKF_OK;
}

int lbmletkf_get_step(lbmletkf_ctx* ctx, long long* step) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    if (!step) return fail(c, LBMLETKF_ERR_ARG, "step is NULL");
    *step = c->step;
    return LBMLETKF_OK;
}

// drop the observation network and everything sized by it
void release_obs(Ctx* c) {
    c->drop_graphs();
    for (void** p : {(void**)&c->d_site_node, (void**)&c->d_loc_ptr, (void**)&c->d_loc_site, (void**)&c->d_loc_w,
                     (void**)&c->d_obs_coef, (void**)&c->d_yo, (void**)&c->d_hx, (void**)&c->d_hx1,
                     (void**)&c->d_dy, (void**)&c->d_dyo, (void**)&c->d_site_map}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    c->obs_ready = false;
}

int lbmletkf_set_obs_sites(lbmletkf_ctx* ctx, int n_sites, const int* nodes) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    if (n_sites < 0 || (n_sites > 0 && !nodes)) return fail(c, LBMLETKF_ERR_ARG, "bad site list");
    if (c->cfg.loc_kind != LBMLETKF_LOC_GAUSS)
        return fail(c, LBMLETKF_ERR_ARG, "an explicit site list needs the Gaussian localization");
    std::vector<char> seen(c->nodes, 0);
    for (int s = 0; s < n_sites; ++s) {
        if (nodes[s] < 0 || nodes[s] >= c->nodes) return fail(c, LBMLETKF_ERR_ARG, "site node out of the grid");
        if (seen[nodes[s]]++) return fail(c, LBMLETKF_ERR_ARG, "duplicate site node");
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    release_obs(c);
    c->obs.custom = true;
    c->obs.site_nodes.assign(nodes, nodes + n_sites);
    return setup_obs_impl(c);
}

int lbmletkf_setup_obs(lbmletkf_ctx* ctx) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    return setup_obs_impl(c);
}

int lbmletkf_obs_count(lbmletkf_ctx* ctx, int* n_sites, int* n_obs) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    if (int st = setup_obs_impl(c)) return st;
    if (n_sites) *n_sites = c->obs.n_sites;
    if (n_obs) *n_obs = c->obs.n_sites * c->obs.n_vars;
    return LBMLETKF_OK;
}

int lbmletkf_observe(lbmletkf_ctx* ctx, int member, unsigned long long seed,
                     unsigned long long counter, int noise, float* yo, int yo_on_device) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    CHECK_MEMBER(c, member);
    if (!yo) return fail(c, LBMLETKF_ERR_ARG, "yo is NULL");
    if (int st = setup_obs_impl(c)) return st;
    const int n_obs = c->obs.n_sites * c->obs.n_vars;
    // H(x) of the member into d_dyo (free outside an analysis), then noise into the target
    HIP_TRY(c, launch_obs_moments(*c, c->f[c->cur], c->d_hx1, member, 1));
    if (yo_on_device) {
        HIP_TRY(c, launch_observe(*c, c->d_hx1, yo, seed, counter, noise));
        return LBMLETKF_OK;
    }
    HIP_TRY(c, launch_observe(*c, c->d_hx1, c->d_yo, seed, counter, noise));
    if (int st = sync_and_check(c)) return st;
    HIP_TRY(c, hipMemcpy(yo, c->d_yo, sizeof(float) * n_obs, hipMemcpyDeviceToHost));
    return LBMLETKF_OK;
}

int lbmletkf_obs_from_fields(lbmletkf_ctx* ctx, const float* rho, const float* u, const float* v,
                             float* yo_host) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    if (!yo_host) return fail(c, LBMLETKF_ERR_ARG, "yo is NULL");
    if (int st = setup_obs_impl(c)) return st;
    const float* fields[3] = {u, v, rho};
    for (int a = 0; a < c->obs.n_vars; ++a)
        if (!fields[c->obs.var_id[a]]) return fail(c, LBMLETKF_ERR_ARG, "a field of an observed variable is NULL");
    const ObsNet& o = c->obs;
    for (int s = 0; s < o.n_sites; ++s) {
        const long long node = o.site_nodes[s];
        for (int a = 0; a < o.n_vars; ++a) yo_host[s * o.n_vars + a] = fields[o.var_id[a]][node];
    }
    return LBMLETKF_OK;
}

int lbmletkf_analysis(lbmletkf_ctx* ctx, const float* yo, int yo_on_device) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    if (!yo) return fail(c, LBMLETKF_ERR_ARG, "yo is NULL");
    if (c->M < 2) return fail(c, LBMLETKF_ERR_ARG, "the LETKF needs at least two members ((M-1) I is singular for M=1)");
    if (int st = setup_obs_impl(c)) return st;
    if (c->cfg.variant == LBMLETKF_VARIANT_MIXED)
        if (int st = ensure_mixed_buffers(c)) return st;
    const int n_obs = c->obs.n_sites * c->obs.n_vars;
    HIP_TRY(c, hipMemcpyAsync(c->d_yo, yo, sizeof(float) * n_obs,
                              yo_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                              c->stream));
    return analysis_launches(c);
}

int lbmletkf_cycle(lbmletkf_ctx* ctx, const float* yo_device) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    if (!yo_device) return fail(c, LBMLETKF_ERR_ARG, "yo is NULL");
    if (c->M < 2) return fail(c, LBMLETKF_ERR_ARG, "the LETKF needs at least two members ((M-1) I is singular for M=1)");
    if (int st = setup_obs_impl(c)) return st;
    if (c->cfg.variant == LBMLETKF_VARIANT_MIXED)
        if (int st = ensure_mixed_buffers(c)) return st;
    const int n_obs = c->obs.n_sites * c->obs.n_vars;
    HIP_TRY(c, hipMemcpyAsync(c->d_yo, yo_device, sizeof(float) * n_obs, hipMemcpyDeviceToDevice, c->stream));
    // timed phases need their events between the kernels: run eagerly then
    if (c->cfg.variant == LBMLETKF_VARIANT_MIXED && !c->prof) return cycle_graph(c);
    if (int st = analysis_launches(c)) return st;
    return forecast_launches(c, c->cfg.da_interval);
}

int lbmletkf_ensemble_stats(lbmletkf_ctx* ctx, const float* truth_rho, const float* truth_u,
                            const float* truth_v, int truth_on_device, double out[4]) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    if (!truth_rho || !truth_u || !truth_v || !out) return fail(c, LBMLETKF_ERR_ARG, "NULL argument");
    if (int st = ensure_scratch(c)) return st;
    const size_t n = size_t(c->nodes);
    const float* tr = truth_rho;
    const float* tu = truth_u;
    const float* tv = truth_v;
    if (!truth_on_device) {
        HIP_TRY(c, hipMemcpyAsync(c->d_truth, truth_rho, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(c->d_truth + n, truth_u, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(c->d_truth + 2 * n, truth_v, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
        tr = c->d_truth;
        tu = c->d_truth + n;
        tv = c->d_truth + 2 * n;
    }
    const int nb = ensemble_stats_blocks(*c);
    HIP_TRY(c, launch_ensemble_stats(*c, tr, tu, tv, c->d_red, nb));
    if (int st = sync_and_check(c)) return st;
    std::vector<double> part(size_t(nb) * 4);
    HIP_TRY(c, hipMemcpy(part.data(), c->d_red, sizeof(double) * part.size(), hipMemcpyDeviceToHost));
    long double s[4] = {0, 0, 0, 0};
    for (int b = 0; b < nb; ++b)
        for (int a = 0; a < 4; ++a) s[a] += part[size_t(b) * 4 + a];
    out[0] = double(std::sqrt(s[0] / (2.0L * n)));
    out[1] = double(std::sqrt(s[1] / (2.0L * n)));
    out[2] = double(std::sqrt(s[2] / (long double)n));
    out[3] = double(std::sqrt(s[3] / (long double)n));
    return LBMLETKF_OK;
}

int lbmletkf_ensemble_fields(lbmletkf_ctx* ctx, float* mean_u, float* mean_v, float* vor_mean, float* vor_std,
                             float* sd_u, float* sd_v) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    const size_t n = size_t(c->nodes);
    if (!c->d_ens_uv) HIP_TRY(c, dalloc(&c->d_ens_uv, size_t(2) * c->M * n));
    if (!c->d_ens_out) HIP_TRY(c, dalloc(&c->d_ens_out, 6 * n));
    HIP_TRY(c, launch_ensemble_fields(*c, c->d_ens_uv, c->d_ens_out));
    if (int st = sync_and_check(c)) return st;
    float* outs[6] = {mean_u, mean_v, vor_mean, vor_std, sd_u, sd_v};
    for (int k = 0; k < 6; ++k)
        if (outs[k]) HIP_TRY(c, hipMemcpy(outs[k], c->d_ens_out + k * n, sizeof(float) * n, hipMemcpyDeviceToHost));
    return LBMLETKF_OK;
}

int lbmletkf_ensemble_stats_noalloc(lbmletkf_ctx* ctx, const float* truth_rho, const float* truth_u,
                                    const float* truth_v, double out[4]) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    if (!truth_rho || !truth_u || !truth_v || !out) return fail(c, LBMLETKF_ERR_ARG, "NULL argument");
    const int nb = ensemble_stats_blocks(*c);
    if (nb > STATS_SCRATCH_BLOCKS)
        return fail(c, LBMLETKF_ERR_ARG, "grid too large for lbmletkf_ensemble_stats_noalloc: at most " +
                                              std::to_string((long long)STATS_SCRATCH_BLOCKS * 256) + " nodes");
    if (!c->stats_scratch) return fail(c, LBMLETKF_ERR_STATE, "static scratch address not resolved");
    // every block writes its four partial sums: nothing to clear
    HIP_TRY(c, launch_ensemble_stats(*c, truth_rho, truth_u, truth_v, c->stats_scratch, nb));
    if (int st = sync_and_check(c)) return st;
    // the partials summed in block order (as lbmletkf_ensemble_stats does), read back
    // through a fixed host buffer
    long double s[4] = {0, 0, 0, 0};
    double part[4 * 512];
    for (int b0 = 0; b0 < nb; b0 += 512) {
        const int k = std::min(512, nb - b0);
        HIP_TRY(c, hipMemcpy(part, c->stats_scratch + 4 * size_t(b0), sizeof(double) * 4 * size_t(k),
                             hipMemcpyDeviceToHost));
        for (int b = 0; b < k; ++b)
            for (int a = 0; a < 4; ++a) s[a] += part[size_t(b) * 4 + a];
    }
    const size_t n = size_t(c->nodes);
    out[0] = double(std::sqrt(s[0] / (2.0L * n)));
    out[1] = double(std::sqrt(s[1] / (2.0L * n)));
    out[2] = double(std::sqrt(s[2] / (long double)n));
    out[3] = double(std::sqrt(s[3] / (long double)n));
    return LBMLETKF_OK;
}

int lbmletkf_local_obs(lbmletkf_ctx* ctx, int gridpoint, int* n, int* sites, float* weights, int cap) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    if (!n) return fail(c, LBMLETKF_ERR_ARG, "n is NULL");
    if (gridpoint < 0 || gridpoint >= c->nodes) return fail(c, LBMLETKF_ERR_ARG, "gridpoint out of range");
    if (int st = setup_obs_impl(c)) return st;
    int pr[2];
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(pr, c->d_loc_ptr + gridpoint, sizeof(int) * 2, hipMemcpyDeviceToHost));
    *n = pr[1] - pr[0];
    const int k = std::min(*n, std::max(cap, 0));
    if (k > 0 && sites) HIP_TRY(c, hipMemcpy(sites, c->d_loc_site + pr[0], sizeof(int) * k, hipMemcpyDeviceToHost));
    if (k > 0 && weights) HIP_TRY(c, hipMemcpy(weights, c->d_loc_w + pr[0], sizeof(float) * k, hipMemcpyDeviceToHost));
    return LBMLETKF_OK;
}

int lbmletkf_set_dump(lbmletkf_ctx* ctx, int g0, int count) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    if (count < 0 || g0 < 0 || (long long)g0 + count > c->nodes) return fail(c, LBMLETKF_ERR_ARG, "dump range out of bounds");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->drop_graphs();  // the captured analysis names the dump buffer and range
    if (c->d_dump) {
        HIP_TRY(c, hipFree(c->d_dump));
        c->d_dump = nullptr;
    }
    c->dump_g0 = g0;
    c->dump_n = count;
    if (count > 0) {
        const size_t per = size_t(dump_record(c->M));
        HIP_TRY(c, dalloc(&c->d_dump, per * count));
        HIP_TRY(c, hipMemset(c->d_dump, 0, sizeof(double) * per * count));
    }
    return LBMLETKF_OK;
}

int lbmletkf_get_dump(lbmletkf_ctx* ctx, double* W, double* w, double* evals) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    if (c->dump_n == 0) return fail(c, LBMLETKF_ERR_STATE, "no dump range set");
    if (int st = sync_and_check(c)) return st;
    const int M = c->M;
    const size_t per = size_t(dump_record(M));
    std::vector<double> h(per * c->dump_n);
    HIP_TRY(c, hipMemcpy(h.data(), c->d_dump, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (int g = 0; g < c->dump_n; ++g) {
        const double* src = h.data() + per * g;
        if (c->dump_kind == 1) {
            // the fused kernel records V (row-major), d and b; the transform it applies
            // is W = sqrt(M-1) V D^-1/2 V^T and w = V D^-1 V^T b (letkf_solver.h:288-306)
            const double* V = src;
            const double* d = src + size_t(M) * M;
            const double* b = d + M;
            std::vector<double> hv(M);
            for (int i = 0; i < M; ++i) {
                double s = 0;
                for (int k = 0; k < M; ++k) s += V[size_t(k) * M + i] * b[k];
                hv[i] = s / d[i];
            }
            const double sq = std::sqrt(double(M - 1));
            for (int k = 0; k < M; ++k) {
                if (w) {
                    double s = 0;
                    for (int i = 0; i < M; ++i) s += V[size_t(k) * M + i] * hv[i];
                    w[size_t(M) * g + k] = s;
                }
                if (W)
                    for (int l = 0; l < M; ++l) {
                        double s = 0;
                        for (int i = 0; i < M; ++i) s += V[size_t(k) * M + i] * V[size_t(l) * M + i] / std::sqrt(d[i]);
                        W[size_t(M) * M * g + size_t(k) * M + l] = sq * s;
                    }
            }
            if (evals) {
                std::vector<double> e(d, d + M);
                std::sort(e.begin(), e.end());
                std::copy(e.begin(), e.end(), evals + size_t(M) * g);
            }
            continue;
        }
        if (W) std::copy(src, src + size_t(M) * M, W + size_t(M) * M * g);
        if (w) std::copy(src + size_t(M) * M, src + size_t(M) * M + M, w + size_t(M) * g);
        if (evals) {
            std::vector<double> d(src + size_t(M) * M + M, src + size_t(M) * M + 2 * M);
            std::sort(d.begin(), d.end());
            std::copy(d.begin(), d.end(), evals + size_t(M) * g);
        }
    }
    return LBMLETKF_OK;
}

int lbmletkf_get_dump_evd(lbmletkf_ctx* ctx, double* C, double* V, double* evals) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    if (c->dump_n == 0) return fail(c, LBMLETKF_ERR_STATE, "no dump range set");
    if (int st = sync_and_check(c)) return st;
    const int M = c->M;
    const size_t per = size_t(dump_record(M)), MM = size_t(M) * M;
    std::vector<double> h(per * c->dump_n);
    HIP_TRY(c, hipMemcpy(h.data(), c->d_dump, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (int g = 0; g < c->dump_n; ++g) {
        const double* src = h.data() + per * g;
        // MIXED records V and d first; PLAIN records W, w, d, and V after C
        const double* Vs = c->dump_kind == 1 ? src : src + MM + 2 * M + MM;
        const double* ds = c->dump_kind == 1 ? src + MM : src + MM + M;
        if (C) std::copy(src + MM + 2 * M, src + MM + 2 * M + MM, C + MM * g);
        if (V) std::copy(Vs, Vs + MM, V + MM * g);
        if (evals) std::copy(ds, ds + M, evals + size_t(M) * g);
    }
    return LBMLETKF_OK;
}

int lbmletkf_get_obs_ensemble(lbmletkf_ctx* ctx, float* dy, float* dyo) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    if (!c->obs_ready || !c->d_dy) return fail(c, LBMLETKF_ERR_STATE, "no analysis has run");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const size_t n_obs = size_t(c->obs.n_sites) * c->obs.n_vars;
    if (dy) HIP_TRY(c, hipMemcpy(dy, c->d_dy, sizeof(float) * n_obs * c->M, hipMemcpyDeviceToHost));
    if (dyo) HIP_TRY(c, hipMemcpy(dyo, c->d_dyo, sizeof(float) * n_obs, hipMemcpyDeviceToHost));
    return LBMLETKF_OK;
}

int lbmletkf_profile_enable(lbmletkf_ctx* ctx, int on) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (int st = prof_collect(c)) return st;
    for (int p = 0; p < PH_COUNT; ++p) {
        c->prof_sec[p] = 0;
        c->prof_launch[p] = 0;
        c->prof_work[p] = 0;
    }
    if (!c->d_sweeps) HIP_TRY(c, dalloc(&c->d_sweeps, 1));
    if (!c->d_sweeps64) HIP_TRY(c, dalloc(&c->d_sweeps64, 1));
    HIP_TRY(c, hipMemset(c->d_sweeps, 0, sizeof(unsigned long long)));
    HIP_TRY(c, hipMemset(c->d_sweeps64, 0, sizeof(unsigned long long)));
    c->prof = on != 0;
    return LBMLETKF_OK;
}

int lbmletkf_profile_read(lbmletkf_ctx* ctx, const char* phase, double* seconds,
                          long long* launches, long long* work) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    if (!phase) return fail(c, LBMLETKF_ERR_ARG, "phase is NULL");
    static const char* names[PH_COUNT] = {"forecast", "obs", "gram", "solve", "finish"};
    int ph = -1;
    for (int p = 0; p < PH_COUNT; ++p)
        if (std::strcmp(phase, names[p]) == 0) ph = p;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (int st = prof_collect(c)) return st;
    if (std::strcmp(phase, "solve_fp64") == 0) {
        // FP64 sweeps of the MIXED variant (its time is part of "solve")
        unsigned long long sw = 0;
        if (c->d_sweeps64) HIP_TRY(c, hipMemcpy(&sw, c->d_sweeps64, sizeof(sw), hipMemcpyDeviceToHost));
        if (seconds) *seconds = 0.0;
        if (launches) *launches = 0;
        if (work) *work = (long long)sw;
        return LBMLETKF_OK;
    }
    if (ph < 0) return fail(c, LBMLETKF_ERR_ARG, std::string("unknown phase: ") + phase);
    long long wk = c->prof_work[ph];
    if (ph == PH_SOLVE && c->d_sweeps) {
        unsigned long long sw = 0;
        HIP_TRY(c, hipMemcpy(&sw, c->d_sweeps, sizeof(sw), hipMemcpyDeviceToHost));
        wk = (long long)sw;
    }
    if (seconds) *seconds = c->prof_sec[ph];
    if (launches) *launches = c->prof_launch[ph];
    if (work) *work = wk;
    return LBMLETKF_OK;
}

void* lbmletkf_get_stream(lbmletkf_ctx* ctx) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    return c ? reinterpret_cast<void*>(c->stream) : nullptr;
}

int lbmletkf_set_stream(lbmletkf_ctx* ctx, void* stream) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (s == c->stream || (s == nullptr && c->own_stream)) return LBMLETKF_OK;
    // the work queued so far completes on the old stream; captured cycles are
    // dropped and re-captured on the new one at the next cycle call
    c->drop_graphs();
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (s == nullptr) {
        hipStream_t own = nullptr;
        HIP_TRY(c, hipStreamCreate(&own));
        s = own;
    }
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    c->own_stream = (stream == nullptr);
    c->stream = s;
    return LBMLETKF_OK;
}

int lbmletkf_synchronize(lbmletkf_ctx* ctx) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    return sync_and_check(c);
}

// ---- groups: several contexts stepped and analysed by shared launches -------------
}  // extern "C"

namespace {

struct Group {
    std::vector<Ctx*> cs;       // launch order: widest analyses first
    std::vector<int> user_pos;  // cs[k] is the caller's context user_pos[k]
    hipStream_t stream = nullptr;
    struct Exec {
        std::vector<int> cur;        // parities of the contexts' double buffers at capture
        std::vector<unsigned> gen;
        hipGraphExec_t x;
        std::vector<int> cur_after;  // their parities after a replay
    };
    std::vector<Exec> execs;
    std::string last_error;
    // the forecast of cs[k], k >= 1, runs on side[k-1], forked from `stream` by
    // ev[0] and joined back by ev[k] (default priority, like `stream`)
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> ev;
    void drop() {
        if (!execs.empty() && stream) (void)hipStreamSynchronize(stream);
        for (Exec& e : execs) (void)hipGraphExecDestroy(e.x);
        execs.clear();
    }
    ~Group() {
        drop();
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipStream_t s : side) (void)hipStreamSynchronize(s);
        for (hipStream_t s : side) (void)hipStreamDestroy(s);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// Points every context's stream at the group's for the duration of a group call
// (their launchers use Ctx::stream); the contexts' own streams were drained first.
struct StreamLend {
    Group* g;
    std::vector<hipStream_t> saved;
    explicit StreamLend(Group* g_) : g(g_) {
        for (Ctx* c : g->cs) {
            saved.push_back(c->stream);
            c->stream = g->stream;
        }
    }
    ~StreamLend() {
        for (size_t k = 0; k < g->cs.size(); ++k) g->cs[k]->stream = saved[k];
    }
};

int gfail(Group* g, int code, const std::string& msg) {
    if (g) g->last_error = msg;
    return fail(nullptr, code, msg);
}

#define G_HIP(g, expr)                                                                                 \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) return gfail((g), LBMLETKF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// every context with the learned closure has its network (checked before any launch)
int group_closures_ready(Group* g) {
    for (Ctx* c : g->cs)
        if (c->cfg.closure_kind == LBMLETKF_CLOSURE_MLP && !c->closure_ready)
            return gfail(g, LBMLETKF_ERR_STATE, "a context with closure_kind = MLP has no network: call lbmletkf_set_closure first");
    return LBMLETKF_OK;
}

int group_drain(Group* g) {
    for (Ctx* c : g->cs) G_HIP(g, hipStreamSynchronize(c->stream));
    return LBMLETKF_OK;
}

// per-context observation work (sizes fixed) before any capture
int group_prepare(Group* g, const float* const* yo) {
    for (size_t k = 0; k < g->cs.size(); ++k) {
        Ctx* c = g->cs[k];
        if (!yo[g->user_pos[k]]) return gfail(g, LBMLETKF_ERR_ARG, "yo of a context is NULL");
        if (int st = setup_obs_impl(c)) return gfail(g, st, c->last_error);
        if (int st = ensure_mixed_buffers(c)) return gfail(g, st, c->last_error);
    }
    return LBMLETKF_OK;
}

int group_analysis_launches(Group* g) {
    for (Ctx* c : g->cs) {
        G_HIP(g, launch_obs_moments(*c, c->f[c->cur], c->d_hx, 0, c->M));
        G_HIP(g, launch_obs_prepare(*c));
    }
    G_HIP(g, launch_letkf_mixed_group(g->cs.data(), int(g->cs.size()), g->stream));
    for (Ctx* c : g->cs) c->dump_kind = 1;
    return LBMLETKF_OK;
}

// Each context steps in its own tuned shape: the tuned shapes (one- to three-step row
// strips) beat one tile launch over all contexts, which measured the same as the
// sum of the contexts' one-step launches.  A context with the learned closure steps
// with the closure kernel, one step per launch, as it does on its own.
int group_context_steps(Group* g, Ctx* c, int n_steps) {
    const bool mlp = c->cfg.closure_kind == LBMLETKF_CLOSURE_MLP;
    const int per_launch = mlp ? 1 : lbm_launch_steps(*c);
    for (int s = 0; s < n_steps;) {
        const int k = std::min(per_launch, n_steps - s);
        if (c->lp.has_force) G_HIP(g, launch_force_field(*c, s));
        if (mlp)
            G_HIP(g, launch_lbm_mlp(*c, c->f[c->cur], c->f[1 - c->cur], 0, (c->nx / 64) * (c->ny / 16) * c->M,
                                    nullptr));
        else
            G_HIP(g, launch_lbm_fused(*c, c->f[c->cur], c->f[1 - c->cur], k));
        c->cur = 1 - c->cur;
        s += k;
    }
    if (n_steps > 0) G_HIP(g, launch_step_advance(*c, n_steps));
    c->step += n_steps;
    return LBMLETKF_OK;
}

// The first context (the widest analysis, launched first) steps on the group's
// stream; the others step on side streams forked from it and joined back, so the
// launches of small grids, which do not fill the device, overlap with it.  The
// same fork and join are captured into the cycle graph as parallel branches.
int group_forecast_launches(Group* g, int n_steps) {
    const size_t n = g->cs.size();
    const bool fork = n > 1 && n_steps > 0;
    size_t forked = 0;  // cs[1..forked] step on side streams that wait on ev[0]
    int st = LBMLETKF_OK;
    if (fork) {
        hipError_t e = hipEventRecord(g->ev[0], g->stream);
        for (size_t k = 1; e == hipSuccess && k < n; ++k) {
            e = hipStreamWaitEvent(g->side[k - 1], g->ev[0], 0);
            if (e == hipSuccess) {
                g->cs[k]->stream = g->side[k - 1];  // StreamLend restores the context's own stream
                forked = k;
            }
        }
        if (e != hipSuccess)
            st = gfail(g, LBMLETKF_ERR_HIP, std::string("forking the group's forecast: ") + hipGetErrorString(e));
    }
    if (st == LBMLETKF_OK)
        for (Ctx* c : g->cs)
            if ((st = group_context_steps(g, c, n_steps)) != LBMLETKF_OK) break;
    // Every forked side stream is joined back, also after a failed launch: a capture
    // ends only with all its streams joined, and eager work left on a side stream
    // would not be covered by a wait on the group's stream.
    for (size_t k = 1; k <= forked; ++k) {
        g->cs[k]->stream = g->stream;
        hipError_t e = hipEventRecord(g->ev[k], g->side[k - 1]);
        if (e == hipSuccess) e = hipStreamWaitEvent(g->stream, g->ev[k], 0);
        if (e != hipSuccess && st == LBMLETKF_OK)
            st = gfail(g, LBMLETKF_ERR_HIP, std::string("joining the group's forecast: ") + hipGetErrorString(e));
    }
    return st;
}

}  // namespace

extern "C" {

int lbmletkf_group_create(lbmletkf_ctx* const* ctxs, int n, lbmletkf_group** out) {
    if (!ctxs || !out) return fail(nullptr, LBMLETKF_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (n < 1 || n > MIXED_GROUP_MAX) return fail(nullptr, LBMLETKF_ERR_ARG, "a group holds 1 to 8 contexts");
    std::vector<int> order(n);
    for (int k = 0; k < n; ++k) {
        Ctx* c = reinterpret_cast<Ctx*>(ctxs[k]);
        if (!c) return fail(nullptr, LBMLETKF_ERR_ARG, "context is NULL");
        for (int j = 0; j < k; ++j)
            if (ctxs[j] == ctxs[k]) return fail(nullptr, LBMLETKF_ERR_ARG, "a context appears twice");
        if (c->cfg.variant != LBMLETKF_VARIANT_MIXED) return fail(nullptr, LBMLETKF_ERR_ARG, "group contexts must use the MIXED variant");
        if (!lbm_fused_supported(c->nx, c->ny)) return fail(nullptr, LBMLETKF_ERR_ARG, "group grids must be multiples of 64 x 16");
        if (c->M < 2) return fail(nullptr, LBMLETKF_ERR_ARG, "group contexts need at least two members");
        Ctx* c0 = reinterpret_cast<Ctx*>(ctxs[0]);
        if (c->cfg.gram_passes != c0->cfg.gram_passes || (c->cfg.jacobi_fp64_sweeps > 0) != (c0->cfg.jacobi_fp64_sweeps > 0))
            return fail(nullptr, LBMLETKF_ERR_ARG, "group contexts must share gram_passes and the FP64 sweep switch");
        if (c->device != c0->device) return fail(nullptr, LBMLETKF_ERR_ARG, "group contexts must be on one device");
        order[k] = k;
    }
    // widest analyses (then largest grids) first: the heaviest workgroups start first
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const Ctx* ca = reinterpret_cast<Ctx*>(ctxs[a]);
        const Ctx* cb = reinterpret_cast<Ctx*>(ctxs[b]);
        const int na = mixed_width(ca->M), nb = mixed_width(cb->M);
        return na != nb ? na > nb : ca->nodes > cb->nodes;
    });
    Group* g = new Group();
    for (int k : order) {
        g->cs.push_back(reinterpret_cast<Ctx*>(ctxs[k]));
        g->user_pos.push_back(k);
    }
    hipError_t e = hipStreamCreate(&g->stream);
    for (int k = 1; k < n && e == hipSuccess; ++k) {
        hipStream_t s = nullptr;
        e = hipStreamCreate(&s);
        if (e == hipSuccess) g->side.push_back(s);
    }
    for (int k = 0; k < n && e == hipSuccess; ++k) {
        hipEvent_t ev = nullptr;
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) g->ev.push_back(ev);
    }
    if (e != hipSuccess) {
        delete g;
        return fail(nullptr, LBMLETKF_ERR_HIP, std::string("group streams and events: ") + hipGetErrorString(e));
    }
    *out = reinterpret_cast<lbmletkf_group*>(g);
    return LBMLETKF_OK;
}

int lbmletkf_group_destroy(lbmletkf_group* grp) {
    delete reinterpret_cast<Group*>(grp);
    return LBMLETKF_OK;
}

int lbmletkf_group_forecast(lbmletkf_group* grp, int n_steps) {
    Group* g = reinterpret_cast<Group*>(grp);
    if (!g) return fail(nullptr, LBMLETKF_ERR_ARG, "group is NULL");
    if (n_steps < 0) return gfail(g, LBMLETKF_ERR_ARG, "n_steps must be >= 0");
    if (int st = group_closures_ready(g)) return st;
    if (int st = group_drain(g)) return st;
    StreamLend lend(g);
    return group_forecast_launches(g, n_steps);
}

int lbmletkf_group_analysis(lbmletkf_group* grp, const float* const* yo_device) {
    Group* g = reinterpret_cast<Group*>(grp);
    if (!g || !yo_device) return fail(nullptr, LBMLETKF_ERR_ARG, "NULL argument");
    if (int st = group_drain(g)) return st;
    if (int st = group_prepare(g, yo_device)) return st;
    StreamLend lend(g);
    for (size_t k = 0; k < g->cs.size(); ++k) {
        Ctx* c = g->cs[k];
        const int n_obs = c->obs.n_sites * c->obs.n_vars;
        G_HIP(g, hipMemcpyAsync(c->d_yo, yo_device[g->user_pos[k]], sizeof(float) * n_obs, hipMemcpyDeviceToDevice,
                                g->stream));
    }
    return group_analysis_launches(g);
}

int lbmletkf_group_cycle(lbmletkf_group* grp, const float* const* yo_device) {
    Group* g = reinterpret_cast<Group*>(grp);
    if (!g || !yo_device) return fail(nullptr, LBMLETKF_ERR_ARG, "NULL argument");
    const int steps = g->cs[0]->cfg.da_interval;
    for (Ctx* c : g->cs)
        if (c->cfg.da_interval != steps) return gfail(g, LBMLETKF_ERR_ARG, "group cycles need one da_interval");
    if (int st = group_closures_ready(g)) return st;
    if (int st = group_drain(g)) return st;
    if (int st = group_prepare(g, yo_device)) return st;
    StreamLend lend(g);
    for (size_t k = 0; k < g->cs.size(); ++k) {
        Ctx* c = g->cs[k];
        const int n_obs = c->obs.n_sites * c->obs.n_vars;
        G_HIP(g, hipMemcpyAsync(c->d_yo, yo_device[g->user_pos[k]], sizeof(float) * n_obs, hipMemcpyDeviceToDevice,
                                g->stream));
    }
    // one graph per (parities of the double buffers, buffer generations)
    std::vector<int> cur;
    std::vector<unsigned> gen;
    for (Ctx* c : g->cs) {
        cur.push_back(c->cur);
        gen.push_back(c->graph_gen);
    }
    hipGraphExec_t x = nullptr;
    std::vector<int> cur_after;
    for (auto it = g->execs.begin(); it != g->execs.end();) {
        if (it->gen != gen) {  // a context changed its buffers or shapes since: stale
            (void)hipGraphExecDestroy(it->x);
            it = g->execs.erase(it);
            continue;
        }
        if (it->cur == cur) {
            x = it->x;
            cur_after = it->cur_after;
        }
        ++it;
    }
    if (!x) {
        std::vector<long long> step0;
        for (Ctx* c : g->cs) step0.push_back(c->step);
        G_HIP(g, hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
        int st = group_analysis_launches(g);
        if (st == LBMLETKF_OK) st = group_forecast_launches(g, steps);
        hipGraph_t gr = nullptr;
        const hipError_t ec = hipStreamEndCapture(g->stream, &gr);
        cur_after.clear();
        for (size_t k = 0; k < g->cs.size(); ++k) {
            cur_after.push_back(g->cs[k]->cur);
            g->cs[k]->cur = cur[k];
            g->cs[k]->step = step0[k];
        }
        if (st != LBMLETKF_OK) {
            if (gr) (void)hipGraphDestroy(gr);
            return st;
        }
        G_HIP(g, ec);
        const hipError_t ei = hipGraphInstantiate(&x, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        G_HIP(g, ei);
        g->execs.push_back({cur, gen, x, cur_after});
    }
    G_HIP(g, hipGraphLaunch(x, g->stream));
    for (size_t k = 0; k < g->cs.size(); ++k) {
        Ctx* c = g->cs[k];
        c->cur = cur_after[k];
        c->step += steps;
        c->dump_kind = 1;
    }
    return LBMLETKF_OK;
}

int lbmletkf_group_synchronize(lbmletkf_group* grp) {
    Group* g = reinterpret_cast<Group*>(grp);
    if (!g) return fail(nullptr, LBMLETKF_ERR_ARG, "group is NULL");
    G_HIP(g, hipStreamSynchronize(g->stream));
    for (hipStream_t s : g->side) G_HIP(g, hipStreamSynchronize(s));
    for (Ctx* c : g->cs)
        if (int st = sync_and_check(c)) return gfail(g, st, c->last_error);
    return LBMLETKF_OK;
}

void* lbmletkf_group_get_stream(lbmletkf_group* grp) {
    Group* g = reinterpret_cast<Group*>(grp);
    return g ? reinterpret_cast<void*>(g->stream) : nullptr;
}

// ---- tuning ------------------------------------------------------------------------

int lbmletkf_set_tuning(lbmletkf_ctx* ctx, int lbm_tiles_per_block, int mix_chain) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    const bool three = lbm_tiles_per_block <= -LBM_THREE_STEP;
    const bool two = lbm_tiles_per_block <= -LBM_TWO_STEP && !three;
    if (lbm_tiles_per_block == 0 || lbm_tiles_per_block > 16 || mix_chain < 1 || mix_chain > 64 ||
        (lbm_tiles_per_block < -32 && !two && !three))
        return fail(c, LBMLETKF_ERR_ARG,
                    "the LBM shape must be 1..16 tiles, -8/-16/-32 rows, -116/-132 two-step or -216/-232 three-step "
                    "rows, the chain in [1, 64]");
    if (lbm_tiles_per_block > 0 && lbm_fused_supported(c->nx, c->ny) &&
        ((long long)(c->nx / 64) * (c->ny / 16) * c->M) % lbm_tiles_per_block)
        return fail(c, LBMLETKF_ERR_ARG, "tiles per block must divide the number of tiles");
    if (lbm_tiles_per_block < 0 && !two && !three && !lbm_rows_supported(c->nx, c->ny, -lbm_tiles_per_block))
        return fail(c, LBMLETKF_ERR_ARG, "row strips need nx a multiple of 64 up to 1024 and ny a multiple of the height");
    if (two && !lbm_rows2_supported(c->nx, c->ny, -lbm_tiles_per_block - LBM_TWO_STEP))
        return fail(c, LBMLETKF_ERR_ARG,
                    "two-step row strips need nx a multiple of 64 up to 1024 and ny a multiple of the height (16 or 32)");
    if (three && !lbm_rows3_supported(c->nx, c->ny, -lbm_tiles_per_block - LBM_THREE_STEP))
        return fail(c, LBMLETKF_ERR_ARG,
                    "three-step row strips need nx a multiple of 64 up to 512 and ny a multiple of the height (16 or 32)");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->drop_graphs();
    c->lbm_tiles_per_block = lbm_tiles_per_block;
    c->mix_chain = mix_chain;
    return LBMLETKF_OK;
}

int lbmletkf_get_tuning(lbmletkf_ctx* ctx, int* lbm_tiles_per_block, int* mix_chain) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    if (lbm_tiles_per_block) *lbm_tiles_per_block = c->lbm_tiles_per_block;
    if (mix_chain) *mix_chain = c->mix_chain;
    return LBMLETKF_OK;
}

}  // extern "C"

namespace {

// mean time of `reps` launches of fn on the context's stream after `warm` untimed ones
template <class Fn>
int time_launches(Ctx* c, int warm, int reps, Fn fn, double& us) {
    for (int k = 0; k < warm; ++k) HIP_TRY(c, fn());
    hipEvent_t a = prof_event(c), b = prof_event(c);
    if (!a || !b) return fail(c, LBMLETKF_ERR_HIP, "hipEventCreate failed");
    struct Back {
        Ctx* c;
        hipEvent_t a, b;
        ~Back() {
            c->event_pool.push_back(a);
            c->event_pool.push_back(b);
        }
    } back{c, a, b};
    HIP_TRY(c, hipEventRecord(a, c->stream));
    for (int k = 0; k < reps; ++k) HIP_TRY(c, fn());
    HIP_TRY(c, hipEventRecord(b, c->stream));
    HIP_TRY(c, hipEventSynchronize(b));
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, a, b));
    us = 1e3 * double(ms) / reps;
    return LBMLETKF_OK;
}

std::string tune_key(const Ctx* c) {
    hipDeviceProp_t prop{};
    std::string dev = "unknown";
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) dev = prop.gcnArchName;
    for (char& ch : dev)
        if (ch == ' ' || ch == ':') ch = '_';
    const auto& g = c->cfg;
    return dev + "/" + std::to_string(c->nx) + "x" + std::to_string(c->ny) + "/M" + std::to_string(c->M) + "/v" +
           std::to_string(g.variant) + "/p" + std::to_string(g.gram_passes) + "/f" +
           std::to_string(g.jacobi_fp64_sweeps > 0) + "/s" + std::to_string(c->obs.n_sites) + "/e" +
           std::to_string(c->obs.n_entries) + "/t4";  // t4: three-step strips, fused analysis with column chains
}

}  // namespace

extern "C" {

int lbmletkf_autotune(lbmletkf_ctx* ctx, const char* cache_path, char* report, int report_len) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    CHECK_CTX(c);
    if (c->prof) return fail(c, LBMLETKF_ERR_STATE, "switch phase timing off before tuning");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const bool mixed = c->cfg.variant == LBMLETKF_VARIANT_MIXED;
    if (mixed) {
        if (c->M < 2) return fail(c, LBMLETKF_ERR_ARG, "the analysis needs at least two members");
        if (int st = setup_obs_impl(c)) return st;
        if (int st = ensure_mixed_buffers(c)) return st;
    }
    const std::string key = tune_key(c);
    std::string rep;
    auto put_report = [&]() {
        if (report && report_len > 0) {
            std::strncpy(report, rep.c_str(), size_t(report_len) - 1);
            report[report_len - 1] = '\0';
        }
    };
    // a cached choice for this device and problem
    if (cache_path) {
        if (std::FILE* fp = std::fopen(cache_path, "r")) {
            char line[512];
            while (std::fgets(line, sizeof line, fp)) {
                char k[400];
                int tpb = 0, ch = 0;
                if (std::sscanf(line, "%399s %d %d", k, &tpb, &ch) == 3 && key == k) {
                    std::fclose(fp);
                    if (int st = lbmletkf_set_tuning(ctx, tpb, ch)) return st;
                    rep = "cached " + key + ": lbm_tiles_per_block=" + std::to_string(tpb) +
                          " mix_chain=" + std::to_string(ch);
                    put_report();
                    return LBMLETKF_OK;
                }
            }
            std::fclose(fp);
        }
    }
    c->drop_graphs();
    char buf[128];
    // 1. row-band height of the fused LBM kernel: 1, 2 or 4 tiles of 64x16 per workgroup
    //    (the step writes the idle half of the double buffer only)
    int best_tpb = c->lbm_tiles_per_block;
    if (mixed && lbm_fused_supported(c->nx, c->ny)) {
        const long long tiles = (long long)(c->nx / 64) * (c->ny / 16) * c->M;
        double best = 1e300;
        rep += "lbm us/step:";
        for (int tpb : {1, 2, 4, -8, -16, -LBM_TWO_STEP - 16, -LBM_TWO_STEP - 32, -LBM_THREE_STEP - 16,
                        -LBM_THREE_STEP - 32}) {
            if (tpb > 0 && tiles % tpb) continue;
            if (tpb < 0 && tpb > -LBM_TWO_STEP && !lbm_rows_supported(c->nx, c->ny, -tpb)) continue;
            if (tpb <= -LBM_TWO_STEP && tpb > -LBM_THREE_STEP && !lbm_rows2_supported(c->nx, c->ny, -tpb - LBM_TWO_STEP))
                continue;
            if (tpb <= -LBM_THREE_STEP && !lbm_rows3_supported(c->nx, c->ny, -tpb - LBM_THREE_STEP)) continue;
            c->lbm_tiles_per_block = tpb;
            const int k = lbm_launch_steps(*c);  // time per step (a launch may take several)
            double us = 0;
            if (int st = time_launches(c, 2, 10, [&] { return launch_lbm_fused(*c, c->f[c->cur], c->f[1 - c->cur], k); }, us))
                return st;
            us /= k;
            std::snprintf(buf, sizeof buf, " %d:%.1f", tpb, us);
            rep += buf;
            if (us < best) {
                best = us;
                best_tpb = tpb;
            }
        }
        c->lbm_tiles_per_block = best_tpb;
    }
    // 2. chain length of the FP32 solve kernel (consecutive gridpoints per wavefront,
    //    warm-started from each other), timed on the ensemble's own local problems with
    //    observations of member 0
    int best_chain = c->mix_chain;
    if (mixed) {
        HIP_TRY(c, launch_obs_moments(*c, c->f[c->cur], c->d_hx1, 0, 1));
        HIP_TRY(c, launch_observe(*c, c->d_hx1, c->d_yo, 0x7475ull, 1, 1));
        HIP_TRY(c, launch_obs_moments(*c, c->f[c->cur], c->d_hx, 0, c->M));
        HIP_TRY(c, launch_obs_prepare(*c));
        double best = 1e300;
        rep += std::string(rep.empty() ? "" : "; ") + "solve us:";
        for (int ch : {1, 2, 4, 8, 16, 32}) {
            c->mix_chain = ch;
            double us = 0;
            if (int st = time_launches(c, 1, 3, [&] {
                    return launch_letkf_mixed(*c, c->f[c->cur], nullptr, nullptr, c->cfg.gram_passes, nullptr, nullptr, true);
                }, us))
                return st;
            std::snprintf(buf, sizeof buf, " %d:%.1f", ch, us);
            rep += buf;
            if (us < best) {
                best = us;
                best_chain = ch;
            }
        }
        c->mix_chain = best_chain;
        if (int st = sync_and_check(c)) return st;
    }
    rep = key + ": lbm_tiles_per_block=" + std::to_string(best_tpb) + " mix_chain=" + std::to_string(best_chain) +
          " (" + rep + ")";
    if (cache_path) {
        if (std::FILE* fp = std::fopen(cache_path, "a")) {
            std::fprintf(fp, "%s %d %d\n", key.c_str(), best_tpb, best_chain);
            std::fclose(fp);
        }
    }
    put_report();
    return LBMLETKF_OK;
}

}  // extern "C"
