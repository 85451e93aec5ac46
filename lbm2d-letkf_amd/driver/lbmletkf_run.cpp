// OSSE front-end on the C-ABI: the reference's run modes (src/main.cu:22-150,
// src/config.h:21-150) with its raw output layout (src/io.cu:22-258).
//
//   lbmletkf_run nature   [options]   nature run, writes io/nature/ens0000/*
//   lbmletkf_run observe  [options]   nature run + noisy observations io/observed/ens0000/*
//   lbmletkf_run da_letkf [options]   ensemble with LETKF analyses read from io/observed
//   lbmletkf_run free     [options]   the same ensemble without analyses (DA_DUMMY)
//   lbmletkf_run da_nudging [options] ensemble nudged towards io/observed (DA_NUDGING)
//   lbmletkf_run lyapunov [options]   ensemble perturbed after the spin-up (LYAPNOV)
//
// Main loop, as the reference's: for tt = 0 .. outputs, t = tt*iiter:
//   analysis if t > 0 and (t/iiter) % daprune == 0      (src/letkf.cu:120-123)
//   output   if t % (iiter*ioprune) == 0                (src/io.cu:22-23)
//   iiter LBM steps of every member
//
// Options:
//   --preset ref|baseline   configuration preset (default ref)
//   --members M             ensemble size for da_letkf/free (default 8)
//   --ens-offset K          ensemble id of the first member (default 1; the nature run is id 0)
//   --prefix DIR            output root (default io)
//   --outputs N             number of output intervals in the main loop (default 200)
//   --iiter n               LBM steps per output interval (default da_interval / daprune)
//   --ioprune n             write every n-th interval (default 10)
//   --daprune n             analyse every n-th interval (default 10)
//   --spinup n              LBM steps before the main loop (default 0)
//   --nx n                  grid size (square), overriding the preset
//   --variant plain|mixed   LETKF variant (default: the preset's)
//   --obs-stride p          observation lattice stride (DA_XYPRUNE)
//   --loc-scale L           Gaussian localization scale (cells)
//   --loc-cutoff d          Gaussian localization cutoff (cells)
//   --inflation b           multiplicative covariance inflation (LETKF_COVINF)
//   --obs-seed S            seed of the observation noise (default 334, src/io.cu:166)
//   --no-nan                keep values at unobserved nodes (default: NaN, OBS_XYPRUNE_NAN)
//   --stats                 also write the ensemble mean/std files (ENSEMBLE_STAT)
//   --timing                print the phase timers at the end
//   --nud-rate a            nudging rate alpha (DA_NUD_RATE, default 0.1)
//   --quadra q              nudged quadrants (DA_QUADRA, default 0 = all)
//   --ly-epsilon e          perturbation amplitude of lyapunov runs (LY_EPSILON, default 1e-6)
//   --savemat [g0:count]    at each analysis write the LETKF matrices of gridpoints
//                           [g0, g0+count) (default all) to <prefix>/letkf/mat/ens0
//                           (src/util/letkf_solver.h:244-345, src/letkf.cu:277-306)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lbmletkf.h"
#include "datio.hpp"
#include "matio.hpp"

namespace {

using lbmletkf_io::ens_dir;
using lbmletkf_io::make_dirs;
using lbmletkf_io::read_field;
using lbmletkf_io::step_file;
using lbmletkf_io::write_field;

struct Options {
    std::string mode;
    std::string preset = "ref";
    int members = 8;
    int ens_offset = 1;
    std::string prefix = "io";
    int outputs = 200;
    int iiter = 0;
    int ioprune = 10;
    int daprune = 10;
    long long spinup = 0;
    int nx = 0;
    int variant = -1;
    int obs_stride = 0;
    double loc_scale = 0, loc_cutoff = 0, inflation = 0;
    unsigned obs_seed = 334;
    bool nan_unobserved = true;
    bool stats = false;
    bool timing = false;
    double nud_rate = 0.1;
    int quadra = 0;
    double ly_epsilon = 1e-6;
    bool savemat = false;
    long long savemat_g0 = 0, savemat_n = -1;
};

[[noreturn]] void usage(const char* msg) {
    std::fprintf(stderr, "lbmletkf_run: %s\n", msg);
    std::fprintf(stderr, "usage: lbmletkf_run nature|observe|da_letkf|free|da_nudging|lyapunov [--preset ref|baseline]\n"
                         "       [--members M] [--nud-rate a] [--quadra q] [--ly-epsilon e]\n"
                         "       [--ens-offset K] [--prefix DIR] [--outputs N] [--iiter n] [--ioprune n]\n"
                         "       [--daprune n] [--spinup n] [--nx n] [--variant plain|mixed] [--obs-stride p]\n"
                         "       [--loc-scale L] [--loc-cutoff d] [--inflation b] [--obs-seed S] [--no-nan]\n"
                         "       [--stats] [--timing] [--savemat [g0:count]]\n");
    std::exit(2);
}

Options parse(int argc, char** argv) {
    if (argc < 2) usage("missing mode");
    Options o;
    o.mode = argv[1];
    if (o.mode != "nature" && o.mode != "observe" && o.mode != "da_letkf" && o.mode != "free" &&
        o.mode != "da_nudging" && o.mode != "lyapunov")
        usage(("unknown mode " + o.mode).c_str());
    for (int a = 2; a < argc; ++a) {
        const std::string k = argv[a];
        auto val = [&]() -> std::string {
            if (a + 1 >= argc) usage(("missing value for " + k).c_str());
            return argv[++a];
        };
        if (k == "--preset") o.preset = val();
        else if (k == "--members") o.members = std::stoi(val());
        else if (k == "--ens-offset") o.ens_offset = std::stoi(val());
        else if (k == "--prefix") o.prefix = val();
        else if (k == "--outputs") o.outputs = std::stoi(val());
        else if (k == "--iiter") o.iiter = std::stoi(val());
        else if (k == "--ioprune") o.ioprune = std::stoi(val());
        else if (k == "--daprune") o.daprune = std::stoi(val());
        else if (k == "--spinup") o.spinup = std::stoll(val());
        else if (k == "--nx") o.nx = std::stoi(val());
        else if (k == "--variant") {
            const std::string v = val();
            if (v == "plain") o.variant = LBMLETKF_VARIANT_PLAIN;
            else if (v == "mixed") o.variant = LBMLETKF_VARIANT_MIXED;
            else usage("variant must be plain or mixed");
        } else if (k == "--obs-stride") o.obs_stride = std::stoi(val());
        else if (k == "--loc-scale") o.loc_scale = std::stod(val());
        else if (k == "--loc-cutoff") o.loc_cutoff = std::stod(val());
        else if (k == "--inflation") o.inflation = std::stod(val());
        else if (k == "--obs-seed") o.obs_seed = unsigned(std::stoul(val()));
        else if (k == "--no-nan") o.nan_unobserved = false;
        else if (k == "--stats") o.stats = true;
        else if (k == "--timing") o.timing = true;
        else if (k == "--nud-rate") o.nud_rate = std::stod(val());
        else if (k == "--quadra") o.quadra = std::stoi(val());
        else if (k == "--ly-epsilon") o.ly_epsilon = std::stod(val());
        else if (k == "--savemat") {
            o.savemat = true;
            if (a + 1 < argc && argv[a + 1][0] != '-') {
                const std::string v = argv[++a];
                const size_t c = v.find(':');
                if (c == std::string::npos) usage("--savemat range must be g0:count");
                o.savemat_g0 = std::stoll(v.substr(0, c));
                o.savemat_n = std::stoll(v.substr(c + 1));
            }
        }
        else usage(("unknown option " + k).c_str());
    }
    if (o.members < 1 || o.outputs < 0 || o.ioprune < 1 || o.daprune < 1 || o.spinup < 0)
        usage("counts must be positive");
    return o;
}

// a context with the error text of the library on failure
struct Ctx {
    lbmletkf_ctx* h = nullptr;
    Ctx(const lbmletkf_config& cfg, int m) {
        if (int st = lbmletkf_create(&cfg, m, &h))
            throw std::runtime_error("lbmletkf_create failed (" + std::to_string(st) + "): " +
                                     lbmletkf_last_error(nullptr));
    }
    ~Ctx() { lbmletkf_destroy(h); }
    void ck(int st, const char* what) const {
        if (st) throw std::runtime_error(std::string(what) + " failed (" + std::to_string(st) + "): " +
                                         lbmletkf_last_error(h));
    }
};

struct Fields {
    std::vector<float> r, u, v, vor;
    explicit Fields(size_t n) : r(n), u(n), v(n), vor(n) {}
};

// macroscopic fields and vorticity of one member
void member_fields(const Ctx& c, int m, int nx, int ny, float dx, Fields& f) {
    c.ck(lbmletkf_get_macro(c.h, m, f.r.data(), f.u.data(), f.v.data()), "get_macro");
    lbmletkf_io::vorticity(f.u.data(), f.v.data(), nx, ny, dx, f.vor.data());
}

void write_fields(const std::string& dir, long long step, const Fields& f) {
    make_dirs(dir);
    const size_t n = f.r.size();
    write_field(step_file(dir, "rho", step), f.r.data(), n);
    write_field(step_file(dir, "u", step), f.u.data(), n);
    write_field(step_file(dir, "v", step), f.v.data(), n);
    write_field(step_file(dir, "vor", step), f.vor.data(), n);
}

// Observed fields of the reference's OBSERVE mode with OBSERVE_ERROR_RUV and
// OBS_XYPRUNE_NAN (src/io.cu:159-246): per node in order rho, u, v, each truth plus
// sigma times a draw of std::normal_distribution<float>(0, 1) on one
// std::mt19937(seed) engine that lives for the whole run; vorticity from the noisy
// velocities; NaN off the observation lattice (i or j not a multiple of the stride).
void observe_fields(std::mt19937& eng, std::normal_distribution<float>& nd, const Fields& truth,
                    const lbmletkf_config& cfg, bool nan_off_lattice, Fields& obs) {
    const int nx = cfg.nx, ny = cfg.ny;
    lbmletkf_io::add_obs_noise(eng, nd, truth.r.data(), truth.u.data(), truth.v.data(), truth.r.size(),
                               float(cfg.obs_sigma_rho), float(cfg.obs_sigma_u), obs.r.data(), obs.u.data(),
                               obs.v.data());
    lbmletkf_io::vorticity(obs.u.data(), obs.v.data(), nx, ny, float(cfg.dx), obs.vor.data());
    if (!nan_off_lattice) return;
    const int p = cfg.obs_stride, off = cfg.obs_offset;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            if ((i - off) % p != 0 || (j - off) % p != 0 || i < off || j < off) {
                const size_t ij = size_t(i) + size_t(nx) * j;
                obs.r[ij] = obs.u[ij] = obs.v[ij] = obs.vor[ij] = NAN;
            }
}

// ensemble mean of u, v and mean / standard deviation of the vorticity (src/io.cu:90-145)
struct EnsembleStat {
    std::vector<double> su, sv, svor, svor2;
    int count = 0;
    explicit EnsembleStat(size_t n) : su(n), sv(n), svor(n), svor2(n) {}
    void add(const Fields& f) {
        for (size_t i = 0; i < su.size(); ++i) {
            su[i] += f.u[i];
            sv[i] += f.v[i];
            svor[i] += f.vor[i];
            svor2[i] += double(f.vor[i]) * f.vor[i];
        }
        ++count;
    }
    void write(const std::string& prefix, long long step) const {
        const size_t n = su.size();
        std::vector<float> mu(n), mv(n), mvor(n), sdvor(n);
        for (size_t i = 0; i < n; ++i) {
            mu[i] = float(su[i] / count);
            mv[i] = float(sv[i] / count);
            const double m = svor[i] / count;
            mvor[i] = float(m);
            const double var = svor2[i] / count - m * m;
            sdvor[i] = float(var > 0 ? std::sqrt(var) : 0.0);
        }
        const std::string d = prefix + "/calc";
        make_dirs(d);
        write_field(step_file(d, "ens_mean_u", step), mu.data(), n);
        write_field(step_file(d, "ens_mean_v", step), mv.data(), n);
        write_field(step_file(d, "ens_vor_mean", step), mvor.data(), n);
        write_field(step_file(d, "ens_vor_std", step), sdvor.data(), n);
    }
};

// Member states of gridpoints [g0, g0+n) as the reference's X matrices: per
// gridpoint 9 x M column-major (row = population q, column = member).
std::vector<float> state_matrices(const Ctx& c, int M, size_t nodes, long long g0, long long n) {
    std::vector<float> X(size_t(9) * M * size_t(n)), f(size_t(9) * nodes);
    for (int m = 0; m < M; ++m) {
        c.ck(lbmletkf_get_state(c.h, m, f.data()), "get_state");
        for (long long b = 0; b < n; ++b)
            for (int q = 0; q < 9; ++q) X[q + 9 * (m + size_t(M) * b)] = f[size_t(q) * nodes + size_t(g0 + b)];
    }
    return X;
}

// The reference's solve_xa diagnostics from the transform dump and the states
// before and after the analysis: _Xf, _dXf, _Pa = V D^-1 V^T = W W / (M-1),
// _dWa = W, _Xa, _Wa = W + w 1^T, _AI = Xa - Xf, _Z = Wa - I, _Lambda = diag(d),
// _Wma = w 1^T.  W, w and d are the FP64 results of the analysis kernels; the
// eigenvalues are in the solver's order (that of the columns of _V).
void save_matrices(const std::string& dir, long long t, int M, long long n, const std::vector<float>& Xf,
                   const std::vector<float>& Xa, const std::vector<double>& W, const std::vector<double>& w,
                   const std::vector<double>& d) {
    using lbmletkf_io::save_matrix;
    const size_t MM = size_t(M) * M, N = size_t(n);
    std::vector<float> dX(Xf.size()), AI(Xf.size());
    for (size_t b = 0; b < N; ++b)
        for (int q = 0; q < 9; ++q) {
            double mean = 0;
            for (int m = 0; m < M; ++m) mean += Xf[q + 9 * (m + M * b)];
            mean /= M;
            for (int m = 0; m < M; ++m) {
                const size_t k = q + 9 * (m + M * b);
                dX[k] = float(Xf[k] - mean);
                AI[k] = Xa[k] - Xf[k];
            }
        }
    std::vector<double> dWa(MM * N), Wa(MM * N), Z(MM * N), Lam(MM * N, 0.0), Wma(MM * N), Pa(MM * N);
    for (size_t b = 0; b < N; ++b) {
        const double* Wb = W.data() + MM * b;  // row-major W[i*M + j]
        for (int c = 0; c < M; ++c)
            for (int r = 0; r < M; ++r) {
                const size_t k = r + size_t(M) * c + MM * b;
                dWa[k] = Wb[size_t(r) * M + c];
                Wa[k] = dWa[k] + w[size_t(M) * b + r];
                Z[k] = Wa[k] - (r == c ? 1.0 : 0.0);
                Wma[k] = w[size_t(M) * b + r];
                double s = 0;
                for (int k2 = 0; k2 < M; ++k2) s += Wb[size_t(r) * M + k2] * Wb[size_t(k2) * M + c];
                Pa[k] = s / (M - 1);
            }
        for (int r = 0; r < M; ++r) Lam[r + size_t(M) * r + MM * b] = d[size_t(M) * b + r];
    }
    save_matrix(dir, "_Xf", 9, M, N, t, Xf);
    save_matrix(dir, "_dXf", 9, M, N, t, dX);
    save_matrix(dir, "_Pa", M, M, N, t, Pa);
    save_matrix(dir, "_dWa", M, M, N, t, dWa);
    save_matrix(dir, "_Xa", 9, M, N, t, Xa);
    save_matrix(dir, "_Wa", M, M, N, t, Wa);
    save_matrix(dir, "_AI", 9, M, N, t, AI);
    save_matrix(dir, "_Z", M, M, N, t, Z);
    save_matrix(dir, "_Lambda", M, M, N, t, Lam);
    save_matrix(dir, "_Wma", M, M, N, t, Wma);
}

// The reference's solve_evd diagnostics (src/util/letkf_solver.h:235-272): per
// gridpoint _dYf (local observations x members: the rows of dY = Y - ybar the
// gridpoint's analysis gathers, in its gather order), _dyo (local observations x 1),
// _Q (C = (M-1)/beta I + dY^T R~^-1 dY as the kernels formed it) and _V (its
// eigenvectors, column j for eigenvalue j of _Lambda).  The reference's windows have
// a fixed row count; here the local sets vary per gridpoint (Gaussian localization)
// and the row count is the largest of the range, shorter sets padded with zeros.
void save_evd_matrices(Ctx& c, const std::string& dir, long long t, int M, long long g0, long long n,
                       int n_sites, int n_obs) {
    using lbmletkf_io::save_matrix;
    const size_t MM = size_t(M) * M, N = size_t(n);
    const int nv = n_sites > 0 ? n_obs / n_sites : 1;
    std::vector<double> C(MM * N), V(MM * N), ev(size_t(M) * N);
    c.ck(lbmletkf_get_dump_evd(c.h, C.data(), V.data(), ev.data()), "get_dump_evd");
    std::vector<float> dy(size_t(std::max(n_obs, 1)) * M), dyo(size_t(std::max(n_obs, 1)));
    c.ck(lbmletkf_get_obs_ensemble(c.h, dy.data(), dyo.data()), "get_obs_ensemble");
    std::vector<std::vector<int>> sets(N);
    size_t rows = 1;
    for (size_t b = 0; b < N; ++b) {
        int k = 0;
        std::vector<int> sites(size_t(std::max(n_sites, 1)));
        c.ck(lbmletkf_local_obs(c.h, int(g0 + (long long)b), &k, sites.data(), nullptr, int(sites.size())), "local_obs");
        sites.resize(size_t(std::min(k, int(sites.size()))));
        sets[b] = sites;
        rows = std::max(rows, sites.size() * size_t(nv));
    }
    std::vector<float> dYf(rows * M * N, 0.f), dyol(rows * N, 0.f);
    for (size_t b = 0; b < N; ++b)
        for (size_t e = 0; e < sets[b].size(); ++e)
            for (int a = 0; a < nv; ++a) {
                const size_t r = e * nv + a, o = size_t(sets[b][e]) * nv + a;
                for (int m = 0; m < M; ++m) dYf[r + rows * (m + size_t(M) * b)] = dy[o * M + m];
                dyol[r + rows * b] = dyo[o];
            }
    std::vector<double> Q(MM * N), Vc(MM * N);
    for (size_t b = 0; b < N; ++b)
        for (int col = 0; col < M; ++col)
            for (int r = 0; r < M; ++r) {
                Q[r + size_t(M) * col + MM * b] = C[MM * b + size_t(r) * M + col];
                Vc[r + size_t(M) * col + MM * b] = V[MM * b + size_t(r) * M + col];
            }
    save_matrix(dir, "_dYf", rows, M, N, t, dYf);
    save_matrix(dir, "_dyo", rows, 1, N, t, dyol);
    save_matrix(dir, "_Q", M, M, N, t, Q);
    save_matrix(dir, "_V", M, M, N, t, Vc);
}

int run(const Options& o) {
    lbmletkf_config cfg;
    if (lbmletkf_config_preset(o.preset.c_str(), &cfg)) throw std::runtime_error(lbmletkf_last_error(nullptr));
    if (o.nx > 0) cfg.nx = cfg.ny = o.nx;
    if (o.variant >= 0) cfg.variant = o.variant;
    if (o.obs_stride > 0) cfg.obs_stride = o.obs_stride;
    if (o.loc_scale > 0) cfg.loc_scale = o.loc_scale;
    if (o.loc_cutoff > 0) cfg.loc_cutoff = o.loc_cutoff;
    if (o.inflation > 0) cfg.inflation = o.inflation;
    const int iiter = o.iiter > 0 ? o.iiter : std::max(1, cfg.da_interval / o.daprune);
    cfg.da_interval = iiter * o.daprune;
    const bool ensemble = o.mode == "da_letkf" || o.mode == "free" || o.mode == "da_nudging" || o.mode == "lyapunov";
    const bool assimilate = o.mode == "da_letkf" || o.mode == "da_nudging";
    const int M = ensemble ? o.members : 1;
    const int first_id = ensemble ? o.ens_offset : 0;
    if (int st = lbmletkf_config_check(&cfg, M)) {
        throw std::runtime_error("configuration rejected (" + std::to_string(st) + "): " + lbmletkf_last_error(nullptr));
    }
    const int nx = cfg.nx, ny = cfg.ny;
    const size_t n = size_t(nx) * ny;
    std::printf("lbmletkf_run %s: preset %s, %dx%d, %d member(s) from id %d, iiter %d, outputs %d, ioprune %d, "
                "daprune %d, %s\n",
                o.mode.c_str(), o.preset.c_str(), nx, ny, M, first_id, iiter, o.outputs, o.ioprune, o.daprune,
                lbmletkf_version());

    Ctx c(cfg, M);
    for (int m = 0; m < M; ++m) c.ck(lbmletkf_init_member(c.h, m, (unsigned long long)(first_id + m)), "init_member");
    if (o.spinup > 0) {
        std::printf("spin-up: %lld steps\n", o.spinup);
        c.ck(lbmletkf_forecast(c.h, int(o.spinup)), "forecast");
        c.ck(lbmletkf_set_step(c.h, 0), "set_step");
    }
    if (o.mode == "lyapunov")  // src/main.cu:73-76
        for (int m = 0; m < M; ++m) c.ck(lbmletkf_perturb(c.h, m, (unsigned long long)(first_id + m), o.ly_epsilon), "perturb");
    if (o.timing) c.ck(lbmletkf_profile_enable(c.h, 1), "profile_enable");

    std::mt19937 eng(o.obs_seed);
    std::normal_distribution<float> nd(0.f, 1.f);
    Fields f(n), obs(n);
    int n_obs = 0, n_sites = 0;
    if (o.mode == "da_letkf") c.ck(lbmletkf_obs_count(c.h, &n_sites, &n_obs), "obs_count");
    std::vector<float> yo(size_t(std::max(n_obs, 1)));
    const long long mat_g0 = o.savemat_g0, mat_n = o.savemat_n < 0 ? (long long)n - o.savemat_g0 : o.savemat_n;
    const std::string mat_dir = o.prefix + "/letkf/mat/ens0";
    if (o.savemat && o.mode == "da_letkf") {
        if (mat_g0 < 0 || mat_n < 1 || mat_g0 + mat_n > (long long)n) usage("--savemat range out of the grid");
        c.ck(lbmletkf_set_dump(c.h, int(mat_g0), int(mat_n)), "set_dump");
        make_dirs(mat_dir);
    }

    for (int tt = 0; tt <= o.outputs; ++tt) {
        const long long t = (long long)tt * iiter;
        const long long step = t / iiter;
        if (assimilate && t > 0 && step % o.daprune == 0) {
            const std::string od = ens_dir(o.prefix, "observed", 0);
            read_field(step_file(od, "rho", step), obs.r.data(), n);
            read_field(step_file(od, "u", step), obs.u.data(), n);
            read_field(step_file(od, "v", step), obs.v.data(), n);
        }
        if (o.mode == "da_nudging" && t > 0 && step % o.daprune == 0) {
            // src/nudging.cu:40-142 with DA_OBS_RUV: rho, u, v from the observation lattice
            const bool obs_rho = (cfg.obs_vars & LBMLETKF_OBS_RHO) != 0;
            c.ck(lbmletkf_nudge(c.h, obs_rho ? obs.r.data() : nullptr, obs.u.data(), obs.v.data(), 0, o.nud_rate,
                                cfg.obs_stride, o.quadra), "nudge");
            std::printf("t=%lld: nudging, rate %g\n", t, o.nud_rate);
        }
        if (o.mode == "da_letkf" && t > 0 && step % o.daprune == 0) {
            c.ck(lbmletkf_obs_from_fields(c.h, obs.r.data(), obs.u.data(), obs.v.data(), yo.data()), "obs_from_fields");
            std::vector<float> Xf;
            if (o.savemat) Xf = state_matrices(c, M, n, mat_g0, mat_n);
            c.ck(lbmletkf_analysis(c.h, yo.data(), 0), "analysis");
            std::printf("t=%lld: analysis with %d observations\n", t, n_obs);
            if (o.savemat) {
                const size_t nm = size_t(mat_n);
                std::vector<double> W(nm * M * M), w(nm * M), d(nm * M), ev(nm * M);
                c.ck(lbmletkf_get_dump(c.h, W.data(), w.data(), d.data()), "get_dump");
                // _Lambda in the solver's order, the order of the columns of _V
                c.ck(lbmletkf_get_dump_evd(c.h, nullptr, nullptr, ev.data()), "get_dump_evd");
                save_matrices(mat_dir, t, M, mat_n, Xf, state_matrices(c, M, n, mat_g0, mat_n), W, w, ev);
                save_evd_matrices(c, mat_dir, t, M, mat_g0, mat_n, n_sites, n_obs);
                std::printf("t=%lld: LETKF matrices of %lld gridpoints -> %s\n", t, mat_n, mat_dir.c_str());
            }
        }
        if (t % ((long long)iiter * o.ioprune) == 0) {
            EnsembleStat es(o.stats ? n : 0);
            for (int m = 0; m < M; ++m) {
                member_fields(c, m, nx, ny, float(cfg.dx), f);
                if (ensemble) {
                    write_fields(ens_dir(o.prefix, "calc", first_id + m), step, f);
                    if (o.stats) es.add(f);
                } else {
                    write_fields(ens_dir(o.prefix, "nature", 0), step, f);
                    if (o.mode == "observe") {
                        observe_fields(eng, nd, f, cfg, o.nan_unobserved, obs);
                        write_fields(ens_dir(o.prefix, "observed", 0), step, obs);
                    }
                }
            }
            if (ensemble && o.stats) es.write(o.prefix, step);
            std::printf("t=%lld: output step %lld\n", t, step);
        }
        if (tt < o.outputs) c.ck(lbmletkf_forecast(c.h, iiter), "forecast");
    }
    c.ck(lbmletkf_synchronize(c.h), "synchronize");
    if (o.timing) {
        for (const char* ph : {"forecast", "obs", "gram", "solve", "finish"}) {
            double sec = 0;
            long long launches = 0, work = 0;
            c.ck(lbmletkf_profile_read(c.h, ph, &sec, &launches, &work), "profile_read");
            std::printf("timer %-8s %10.3f ms  launches %lld\n", ph, sec * 1e3, launches);
        }
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return run(parse(argc, argv));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "lbmletkf_run: %s\n", e.what());
        return 1;
    }
}
