// FP64 CPU oracle of the LBM-LETKF cycle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
// Straight loops in double precision, parallelised with OpenMP over nodes,
// members or gridpoints.  Nothing here is shipped or called by the product.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace {

constexpr double PI = 3.14159265358979323846;
constexpr int Q = 9;

inline int ex_of(int q) { return q % 3 - 1; }  // src/config.h:257
inline int ey_of(int q) { return q / 3 - 1; }  // src/config.h:258
inline double wq_of(int q) {                    // src/config.h:260-263
    const int ee = ex_of(q) * ex_of(q) + ey_of(q) * ey_of(q);
    return ee == 0 ? 4.0 / 9.0 : (ee == 1 ? 1.0 / 9.0 : 1.0 / 36.0);
}
inline int pmod(int a, int n) { int r = a % n; return r < 0 ? r + n : r; }
// pmod for -n <= a < 2n (neighbour indices), without a division
inline int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// src/lbm.h:31-39 (non-ICLBM branch)
inline double feq_q(const orc_config& c, double rho, double u, double v, int q) {
    const double uc = (u * ex_of(q) + v * ey_of(q)) / c.c;
    const double uu = (u * u + v * v) / (c.c * c.c);
    return rho * wq_of(q) * (1.0 + 3.0 * uc + 4.5 * uc * uc - 1.5 * uu);
}

// src/lbm.h:41-70 (first-order term only; FORCE_SECOND_ORDER and
// VELOCITY_FORCE_CORRECTION are not defined in src/config.h:16-17)
inline double force_q(const orc_config& c, double fx_, double fy_, int q, double rho, double u, double v) {
    const double cx = c.u_ref * ex_of(q);
    const double cy = c.u_ref * ey_of(q);
    const double cs2 = c.c * c.c / 3.0;
    const double fx = (fx_ - c.friction_rate / c.dt * u) * c.dt;
    const double fy = (fy_ - c.friction_rate / c.dt * v) * c.dt;
    return (cx * fx + cy * fy) / cs2 * wq_of(q) * rho;
}

void philox_block(const uint32_t in[4], uint32_t k0, uint32_t k1, uint32_t out[4]) {
    uint32_t x[4] = {in[0], in[1], in[2], in[3]};
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * x[0];
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * x[2];
        const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
        const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
        const uint32_t y0 = hi1 ^ x[1] ^ k0, y2 = hi0 ^ x[3] ^ k1;
        x[0] = y0;
        x[1] = lo1;
        x[2] = y2;
        x[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    for (int a = 0; a < 4; ++a) out[a] = x[a];
}

void normal4(uint64_t key, const uint32_t ctr[4], double out[4]) {
    uint32_t r[4];
    philox_block(ctr, uint32_t(key), uint32_t(key >> 32), r);
    for (int h = 0; h < 2; ++h) {
        const double u1 = (double(r[2 * h]) + 0.5) / 4294967296.0;
        const double u2 = (double(r[2 * h + 1]) + 0.5) / 4294967296.0;
        const double rad = std::sqrt(-2.0 * std::log(u1));
        out[2 * h] = rad * std::cos(2 * PI * u2);
        out[2 * h + 1] = rad * std::sin(2 * PI * u2);
    }
}

// LAPACK dsyev (Fortran calling convention, 32-bit integers), when the caller
// provides one through orc_set_dsyev; else the eigensolve is orc_jacobi_eig
typedef void (*dsyev_fn)(const char*, const char*, const int*, double*, const int*, double*, double*,
                         const int*, int*);
dsyev_fn g_dsyev = nullptr;

constexpr uint32_t STREAM_FORCE = 0x464f5243u, STREAM_IC = 0x49434b34u, STREAM_OBS = 0x4f425345u;

struct ObsInfo {
    int n_sx, n_sy, n_sites, n_vars;
    int var_id[3];
    double rinv[3];
};

// explicit observation sites (orc_set_obs_sites); empty: the lattice of cfg
std::vector<int> g_sites;
bool g_custom = false;
std::vector<int> g_site_map;  // site index of each node, built for g_map_nodes nodes
long long g_map_nodes = -1;

const std::vector<int>& site_map(long long nodes) {
#pragma omp critical(orc_site_map)
    if (g_map_nodes != nodes) {
        g_site_map.assign(size_t(nodes), -1);
        for (int s = 0; s < int(g_sites.size()); ++s) g_site_map[g_sites[s]] = s;
        g_map_nodes = nodes;
    }
    return g_site_map;
}

ObsInfo obs_info(const orc_config& c) {
    ObsInfo o{};
    o.n_sx = c.nx / c.obs_stride;
    o.n_sy = c.ny / c.obs_stride;
    o.n_sites = g_custom ? int(g_sites.size()) : o.n_sx * o.n_sy;
    const double ru = 1.0 / (c.obs_sigma_u * c.obs_sigma_u), rr = 1.0 / (c.obs_sigma_rho * c.obs_sigma_rho);
    if (c.obs_vars & 1) { o.var_id[o.n_vars] = 0; o.rinv[o.n_vars++] = ru; }
    if (c.obs_vars & 2) { o.var_id[o.n_vars] = 1; o.rinv[o.n_vars++] = ru; }
    if (c.obs_vars & 4) { o.var_id[o.n_vars] = 2; o.rinv[o.n_vars++] = rr; }
    return o;
}

void moments(const orc_config& c, const double* f, long long N, long long node, double& r, double& u, double& v) {
    // src/lbm.h:72-109
    double rr = 0, jx = 0, jy = 0;
    for (int q = 0; q < Q; ++q) {
        const double fq = f[q * N + node];
        rr += fq;
        jx += fq * c.c * ex_of(q);
        jy += fq * c.c * ey_of(q);
    }
    r = rr;
    u = jx / rr;
    v = jy / rr;
}

// Smagorinsky eddy viscosity of node (i, j) from the central differences of the
// streamed velocities, src/lbm.cu:36-60
double les_node(const orc_config& c, const double* u, const double* v, int i, int j) {
    const int nx = c.nx, ny = c.ny;
    const double dx = c.dx, cs = c.cs_smag;
    const long long e = wrap(i + 1, nx) + (long long)nx * j, w = wrap(i - 1, nx) + (long long)nx * j;
    const long long n = i + (long long)nx * wrap(j + 1, ny), s = i + (long long)nx * wrap(j - 1, ny);
    const double ux = (u[e] - u[w]) / (2 * dx), uy = (u[n] - u[s]) / (2 * dx);
    const double vx = (v[e] - v[w]) / (2 * dx), vy = (v[n] - v[s]) / (2 * dx);
    const double s00 = ux, s01 = 0.5 * (uy + vx), s10 = s01, s11 = vy;
    const double ss = std::sqrt(2 * (s00 * s00 + s01 * s10 + s10 * s01 + s11 * s11));
    return cs * cs * dx * dx * ss;
}

}  // namespace

extern "C" {

int orc_sizeof_config(void) { return int(sizeof(orc_config)); }
void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
void orc_set_dsyev(void* fn) { g_dsyev = reinterpret_cast<dsyev_fn>(fn); }

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox_block(ctr, key[0], key[1], out);
}
void orc_philox_normal4(uint64_t key, const uint32_t ctr[4], double out[4]) { normal4(key, ctr, out); }

// src/data.cu:161-223 (mode search :172-189, amplitudes normalised to fkf*u_ref :218-223)
int orc_force_modes(const orc_config* cfg, double* kx, double* ky, double* amp, int cap) {
    const orc_config& c = *cfg;
    if (c.force_kind != 1) return 0;
    const int kmax_search = int(c.force_kf + c.force_dk);
    std::vector<double> vkx, vky, va;
    long double asum = 0;
    for (int y = 0; y < kmax_search; y++) {
        for (int x = 0; x < kmax_search; x++) {
            if (x == 0 && y == 0) continue;
            const double k = std::sqrt(double(y * y + x * x));
            if (c.force_kf - c.force_dk <= k && k <= c.force_kf + c.force_dk) {
                const double a = std::exp(-(k - c.force_kf) * (k - c.force_kf) / c.force_sigma_k);
                asum += a;
                vkx.push_back(x);
                vky.push_back(y);
                va.push_back(a);
            }
        }
    }
    const int n = int(std::min<size_t>(vkx.size(), 100));
    for (int m = 0; m < n && m < cap; ++m) {
        kx[m] = vkx[m];
        ky[m] = vky[m];
        amp[m] = double((long double)va[m] * c.force_amp_total / asum);
    }
    return n;
}

// src/data.cu:239-350 (field :305-349), randoms from Philox instead of cuRAND MTGP32
void orc_force_field(const orc_config* cfg, long long step, double* fx, double* fy) {
    const orc_config& c = *cfg;
    double kx[100], ky[100], amp[100];
    const int n = orc_force_modes(cfg, kx, ky, amp, 100);
    std::vector<double> rnd(size_t(n) * 4);
    for (int m = 0; m < n; ++m) {
        const uint32_t ctr[4] = {uint32_t(step), uint32_t(uint64_t(step) >> 32), uint32_t(m), STREAM_FORCE};
        double z[4];
        normal4(c.force_seed, ctr, z);
        for (int a = 0; a < 4; ++a) rnd[m * 4 + a] = z[a] * c.force_rand_stdev;
    }
    const int nx = c.nx, ny = c.ny;
#pragma omp parallel for
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
            const double x = double(i - nx / 2) / nx, y = double(j - ny / 2) / ny;
            double sx = 0, sy = 0;
            for (int m = 0; m < n; ++m) {
                const double th = 2 * PI * (kx[m] * x + ky[m] * y);
                const double* r = &rnd[m * 4];
                const double fek = (r[0] * kx[m] + r[1] * ky[m]) * std::sin(th) + (r[2] * kx[m] + r[3] * ky[m]) * std::cos(th);
                sx += amp[m] * 2 * PI * ky[m] * fek;
                sy += amp[m] * 2 * PI * (-kx[m]) * fek;
            }
            fx[i + nx * j] = sx;
            fy[i + nx * j] = sy;
        }
    }
}

void orc_feq(const orc_config* cfg, const double* rho, const double* u, const double* v, double* f) {
    const long long N = (long long)cfg->nx * cfg->ny;
#pragma omp parallel for
    for (long long n = 0; n < N; ++n)
        for (int q = 0; q < Q; ++q) f[q * N + n] = feq_q(*cfg, rho[n], u[n], v[n], q);
}

void orc_macro(const orc_config* cfg, const double* f, double* rho, double* u, double* v) {
    const long long N = (long long)cfg->nx * cfg->ny;
#pragma omp parallel for
    for (long long n = 0; n < N; ++n) {
        double r, uu, vv;
        moments(*cfg, f, N, n, r, uu, vv);
        if (rho) rho[n] = r;
        if (u) u[n] = uu;
        if (v) v[n] = vv;
    }
}

// src/lbm.cu:67-103 (pull streaming :81-89; the moments :92-100 are orc_macro)
void orc_stream(const orc_config* cfg, const double* f, double* fn) {
    const int nx = cfg->nx, ny = cfg->ny;
    const long long N = (long long)nx * ny;
#pragma omp parallel for
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            for (int q = 0; q < Q; ++q) {
                const int iu = pmod(i - ex_of(q), nx), ju = pmod(j - ey_of(q), ny);
                fn[q * N + i + (long long)nx * j] = f[q * N + iu + (long long)nx * ju];
            }
}

// src/lbm.cu:24-65 (Smagorinsky :36-60, LES_CSM not defined)
void orc_les(const orc_config* cfg, const double* u, const double* v, double* nus) {
    const int nx = cfg->nx, ny = cfg->ny;
#pragma omp parallel for
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) nus[i + (long long)nx * j] = les_node(*cfg, u, v, i, j);
}

// src/lbm.h:16-29: stream + moments, LES, collide + force, for every member.
// Two passes over (member, row): the pull streaming and the moments of every
// node, then the Smagorinsky viscosity (it needs the streamed velocities of the
// four neighbours) and the collision.
void orc_lbm_step_forced(const orc_config* cfg, int M, const double* fin, double* fout, const double* Fx,
                         const double* Fy) {
    const orc_config& c = *cfg;
    const int nx = c.nx, ny = c.ny;
    const long long N = (long long)nx * ny;
    std::vector<double> r(size_t(M) * N), u(size_t(M) * N), v(size_t(M) * N);
#pragma omp parallel for collapse(2)
    for (int k = 0; k < M; ++k)
        for (int j = 0; j < ny; ++j) {
            const double* f = fin + (long long)k * Q * N;
            double* fn = fout + (long long)k * Q * N;
            for (int i = 0; i < nx; ++i) {
                const long long n = i + (long long)nx * j;
                for (int q = 0; q < Q; ++q) {                       // src/lbm.cu:81-89
                    const int iu = wrap(i - ex_of(q), nx), ju = wrap(j - ey_of(q), ny);
                    fn[q * N + n] = f[q * N + iu + (long long)nx * ju];
                }
                moments(c, fn, N, n, r[k * N + n], u[k * N + n], v[k * N + n]);   // src/lbm.cu:92-100
            }
        }
#pragma omp parallel for collapse(2)
    for (int k = 0; k < M; ++k)
        for (int j = 0; j < ny; ++j) {
            double* fn = fout + (long long)k * Q * N;
            const double *rk = r.data() + k * N, *uk = u.data() + k * N, *vk = v.data() + k * N;
            for (int i = 0; i < nx; ++i) {                           // src/lbm.cu:105-160
                const long long n = i + (long long)nx * j;
                const double nu = c.nu + les_node(c, uk, vk, i, j);         // src/lbm.cu:24-65
                const double tau = 0.5 + 3.0 * nu / c.c / c.c / c.dt;   // src/lbm.cu:134-138
                const double omega = 1.0 / tau;
                for (int q = 0; q < Q; ++q) {
                    double fc = fn[q * N + n];
                    fc -= omega * (fc - feq_q(c, rk[n], uk[n], vk[n], q));   // lbm_collision_srt.h:8-16
                    fc += force_q(c, Fx[n], Fy[n], q, rk[n], uk[n], vk[n]);  // src/lbm.cu:147-152
                    fn[q * N + n] = fc;
                }
            }
        }
}

void orc_lbm_step(const orc_config* cfg, int M, const double* fin, double* fout, long long step) {
    const long long N = (long long)cfg->nx * cfg->ny;
    std::vector<double> Fx(N, 0.0), Fy(N, 0.0);
    if (cfg->force_kind == 1) orc_force_field(cfg, step, Fx.data(), Fy.data());   // src/lbm.cu:20-22
    orc_lbm_step_forced(cfg, M, fin, fout, Fx.data(), Fy.data());
}

// src/data.cu:25-141 (REFERENCE kind); the K4 kind uses the same streamfunction
// form over the half plane with Philox phases.  Direct sums over all modes.
void orc_initial_velocity(const orc_config* cfg, unsigned long long seed, double* u, double* v) {
    const orc_config& c = *cfg;
    const int nx = c.nx, ny = c.ny;
    struct M4 { double kx, ky, p, th; };
    std::vector<M4> modes;
    if (c.ic_kind == 0) {
        const int nk = nx / 2;
        std::mt19937 eng(static_cast<std::mt19937::result_type>(seed));
        std::uniform_real_distribution<float> dist(-1, 1);
        const float k0 = float(c.ic_k0), dk = float(c.ic_dk), sg = float(c.ic_sigma);
        for (int kj = 0; kj < nk; kj++)
            for (int ki = 0; ki < nk; ki++) {
                const float k = std::sqrt(float(ki * ki + kj * kj));
                const float pk = (k0 - dk <= k && k <= k0 + dk) ? std::exp(-(k - k0) * (k - k0) / sg) : 0.f;
                const float th = float(dist(eng) * PI);
                if (pk != 0.f) modes.push_back({2 * PI * ki / nx, 2 * PI * kj / ny, double(pk), double(th)});
            }
    } else {
        const int nk = std::min(nx, ny) / 2;
        for (int kj = 0; kj < nk; ++kj)
            for (int ki = -nk + 1; ki < nk; ++ki) {
                if (kj == 0 && ki <= 0) continue;
                const double k = std::sqrt(double(ki * ki + kj * kj));
                if (k >= nk) continue;
                const uint32_t ctr[4] = {uint32_t(ki + nk), uint32_t(kj), 0u, STREAM_IC};
                const uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32)};
                uint32_t r[4];
                orc_philox4x32_10(ctr, key, r);
                const double th = PI * (2.0 * (double(r[0]) + 0.5) / 4294967296.0 - 1.0);
                const double p = std::sqrt(k) * std::exp(-(k / c.ic_k0) * (k / c.ic_k0));
                modes.push_back({2 * PI * ki / nx, 2 * PI * kj / ny, p, th});
            }
    }
    const long long N = (long long)nx * ny;
#pragma omp parallel for
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double su = 0, sv = 0;
            for (const M4& m : modes) {
                const double cs = std::cos(m.kx * i + m.ky * j + m.th);
                su += m.p * m.ky * cs;
                sv += m.p * -m.kx * cs;
            }
            u[i + (long long)nx * j] = su;
            v[i + (long long)nx * j] = sv;
        }
    double scale;
    if (c.ic_kind == 0) {
        double vmax = 1e-30 * c.u_ref;
        for (long long n = 0; n < N; ++n) vmax = std::max(vmax, std::sqrt(u[n] * u[n] + v[n] * v[n]));
        scale = c.ic_amp / vmax;
    } else {
        double e = 0;
        for (long long n = 0; n < N; ++n) e += u[n] * u[n] + v[n] * v[n];
        const double rms = std::sqrt(e / double(N));
        scale = rms > 0 ? c.ic_amp / rms : 0;
    }
    for (long long n = 0; n < N; ++n) {
        u[n] *= scale;
        v[n] *= scale;
    }
}

void orc_set_obs_sites(int n, const int* nodes) {
    g_custom = nodes != nullptr && n >= 0;
    g_sites.assign(nodes, nodes + (g_custom ? n : 0));
    g_map_nodes = -1;
}

void orc_obs_count(const orc_config* cfg, int* n_sites, int* n_obs) {
    const ObsInfo o = obs_info(*cfg);
    if (n_sites) *n_sites = o.n_sites;
    if (n_obs) *n_obs = o.n_sites * o.n_vars;
}

// src/letkf.cu:105-109
double orc_gaspari_cohn(double d) {
    return (d <= 0) ? 1
         : (d <= 1) ? -d * d * d * d * d / 4. + d * d * d * d / 2. + d * d * d * 5. / 8. - d * d * 5. / 3. + 1
         : (d <= 2) ? d * d * d * d * d / 12. - d * d * d * d / 2. + d * d * d * 5. / 8. + d * d * 5. / 3. - d * 5 + 4 - 2. / 3. / d
         : 0;
}

int orc_local_obs(const orc_config* cfg, int gridpoint, int* sites, double* weights, int cap) {
    const orc_config& c = *cfg;
    const ObsInfo o = obs_info(c);
    const int i = gridpoint % c.nx, j = gridpoint / c.nx;
    const int p = c.obs_stride;
    int n = 0;
    auto put = [&](int s, double w) {
        if (n < cap) {
            if (sites) sites[n] = s;
            if (weights) weights[n] = w;
        }
        ++n;
    };
    if (c.loc_kind == 0) {
        // window geometry, src/letkf.cu:45-55
        int n_obs_local, obs_offset, n_obs_x;
        double d_local;
        if (c.loc_rloc_len <= 0) {
            d_local = p - 1;
            n_obs_local = p == 1 ? 0 : 2;
            obs_offset = p == 1 ? 0 : 1;
            n_obs_x = p == 1 ? 1 : 4;
        } else {
            d_local = double(c.loc_rloc_len) * p;
            obs_offset = 0;
            n_obs_local = c.loc_rloc_len * 2;
            n_obs_x = n_obs_local * 2 + 1;
        }
        for (int jr = 0; jr < n_obs_x; ++jr)
            for (int ir = 0; ir < n_obs_x; ++ir) {
                // src/letkf.cu:164-165 and 248-254
                const int io = (int(i / p) + ir - (n_obs_local - obs_offset)) * p;
                const int jo = (int(j / p) + jr - (n_obs_local - obs_offset)) * p;
                const int di = std::abs(i - io), dj = std::abs(j - jo);
                const int iop = (io + c.nx) % c.nx, jop = (jo + c.ny) % c.ny;
                double w = 1.0;
                if (p > 1) {  // src/letkf.cu:97-113 (R-localization)
                    const double cutoff = d_local + 1e-6;
                    w = orc_gaspari_cohn(std::sqrt(double(di) * di + double(dj) * dj) / cutoff);
                }
                if (di > n_obs_local * p || dj > n_obs_local * p) w = 0.0;  // src/letkf.cu:181,264
                put(iop / p + (jop / p) * o.n_sx, w);
            }
    } else if (g_custom) {
        // explicit sites inside the cutoff, window rows then columns ascending
        const int R = int(std::floor(c.loc_cutoff));
        const std::vector<int>& map = site_map((long long)c.nx * c.ny);
        for (int dy = -R; dy <= R; ++dy)
            for (int dx = -R; dx <= R; ++dx) {
                const double d2 = double(dx * dx + dy * dy);
                if (d2 > c.loc_cutoff * c.loc_cutoff) continue;
                const int s = map[pmod(i + dx, c.nx) + (size_t)c.nx * pmod(j + dy, c.ny)];
                if (s >= 0) put(s, std::exp(-d2 / (2 * c.loc_scale * c.loc_scale)));
            }
    } else {
        const int R = int(std::floor(c.loc_cutoff));
        for (int sy = (int)std::ceil(double(j - c.obs_offset - R) / p); sy * p + c.obs_offset <= j + R; ++sy)
            for (int sx = (int)std::ceil(double(i - c.obs_offset - R) / p); sx * p + c.obs_offset <= i + R; ++sx) {
                const int dx = i - (sx * p + c.obs_offset), dy = j - (sy * p + c.obs_offset);
                const double d2 = double(dx * dx + dy * dy);
                if (d2 <= c.loc_cutoff * c.loc_cutoff)
                    put(pmod(sx, o.n_sx) + pmod(sy, o.n_sy) * o.n_sx, std::exp(-d2 / (2 * c.loc_scale * c.loc_scale)));
            }
    }
    return n;
}

// H(x) at every site for every member: hx[(s*n_vars+a)*M + k]  (src/letkf.cu:150-187)
void orc_obs_moments(const orc_config* cfg, int M, const double* f, double* hx) {
    const orc_config& c = *cfg;
    const ObsInfo o = obs_info(c);
    const long long N = (long long)c.nx * c.ny;
#pragma omp parallel for
    for (int s = 0; s < o.n_sites; ++s) {
        const int sx = s % o.n_sx, sy = s / o.n_sx;
        const long long node = g_custom ? (long long)g_sites[s]
                                        : (sx * c.obs_stride + c.obs_offset) + (long long)c.nx * (sy * c.obs_stride + c.obs_offset);
        for (int k = 0; k < M; ++k) {
            double r, u, v;
            moments(c, f + (long long)k * Q * N, N, node, r, u, v);
            const double val[3] = {u, v, r};
            for (int a = 0; a < o.n_vars; ++a) hx[((long long)s * o.n_vars + a) * M + k] = val[o.var_id[a]];
        }
    }
}

void orc_observe(const orc_config* cfg, const double* hx, unsigned long long seed,
                 unsigned long long counter, int noise, double* yo) {
    const ObsInfo o = obs_info(*cfg);
    const int n_obs = o.n_sites * o.n_vars;
    for (int a = 0; a < n_obs; ++a) {
        double y = hx[a];
        if (noise) {
            const uint32_t ctr[4] = {uint32_t(counter), uint32_t(counter >> 32), uint32_t(a / 4), STREAM_OBS};
            double z[4];
            normal4(seed, ctr, z);
            const double sig = o.var_id[a % o.n_vars] == 2 ? cfg->obs_sigma_rho : cfg->obs_sigma_u;
            y += sig * z[a % 4];
        }
        yo[a] = y;
    }
}

// Row-cyclic Jacobi (Golub & Van Loan, Matrix Computations, Alg. 8.5.3).
int orc_jacobi_eig(int n, double* A, double* V, double* d, double tol, int max_sweeps) {
    for (int a = 0; a < n * n; ++a) V[a] = 0.0;
    for (int a = 0; a < n; ++a) V[a * n + a] = 1.0;
    int sweep = 0;
    for (; sweep < max_sweeps; ++sweep) {
        double off = 0, tot = 0;
        for (int a = 0; a < n; ++a)
            for (int b = 0; b < n; ++b) {
                const double x = A[a * n + b];
                tot += x * x;
                if (a != b) off += x * x;
            }
        if (off <= tol * tol * tot) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[p * n + q];
                if (apq == 0.0) continue;
                const double theta = (A[q * n + q] - A[p * n + p]) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                const double cs = 1 / std::sqrt(t * t + 1), sn = t * cs;
                for (int k = 0; k < n; ++k) {  // columns p, q
                    const double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = cs * akp - sn * akq;
                    A[k * n + q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < n; ++k) {  // rows p, q
                    const double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = cs * apk - sn * aqk;
                    A[q * n + k] = sn * apk + cs * aqk;
                }
                A[p * n + q] = 0.0;
                A[q * n + p] = 0.0;
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = cs * vkp - sn * vkq;
                    V[k * n + q] = sn * vkp + cs * vkq;
                }
            }
    }
    for (int a = 0; a < n; ++a) d[a] = A[a * n + a];
    return sweep;
}

int orc_letkf_transform(int M, int p, const double* dY, const double* rinv, const double* dyo,
                        double beta, double tol, int max_sweeps, double* W, double* w, double* evals) {
    // src/util/letkf_solver.h:251-254: C = (M-1)/beta I + dY^T rR dY
    std::vector<double> C(size_t(M) * M, 0.0), V(size_t(M) * M), d(M), b(M, 0.0);
    for (int a = 0; a < M; ++a) C[a * M + a] = (M - 1) / beta;
    for (int o = 0; o < p; ++o) {
        const double* y = dY + (size_t)o * M;
        const double r = rinv[o];
        if (r == 0.0) continue;
        for (int a = 0; a < M; ++a) {
            const double ra = r * y[a];
            b[a] += ra * dyo[o];
            for (int bb = 0; bb < M; ++bb) C[a * M + bb] += ra * y[bb];
        }
    }
    int sweeps = 0;
    if (g_dsyev) {
        // LAPACK dsyev on the lower triangle (column-major; C is symmetric).  The
        // eigenvectors come back as the columns of A: V[k][i] = A[k + M i].
        std::vector<double> A(C), work(size_t(66) * M);
        const int lwork = int(work.size());
        int info = 0;
        g_dsyev("V", "L", &M, A.data(), &M, d.data(), work.data(), &lwork, &info);
        if (info != 0) return -1;
        for (int k = 0; k < M; ++k)
            for (int i = 0; i < M; ++i) V[k * M + i] = A[k + (size_t)M * i];
    } else {
        sweeps = orc_jacobi_eig(M, C.data(), V.data(), d.data(), tol, max_sweeps);
    }
    // letkf_solver.h:288-306: Pa = V D^-1 V^T, w = Pa dY^T rR dyo, W = sqrt(M-1) V D^-1/2 V^T
    std::vector<double> g(M, 0.0);
    for (int i = 0; i < M; ++i) {
        for (int k = 0; k < M; ++k) g[i] += V[k * M + i] * b[k];
        g[i] /= d[i];
    }
    for (int k = 0; k < M; ++k) {
        double s = 0;
        for (int i = 0; i < M; ++i) s += V[k * M + i] * g[i];
        w[k] = s;
    }
    const double sq = std::sqrt(double(M - 1));
    std::vector<double> dis(M);
    for (int i = 0; i < M; ++i) dis[i] = 1.0 / std::sqrt(d[i]);
    for (int k = 0; k < M; ++k)
        for (int l = 0; l < M; ++l) {
            double s = 0;
            for (int i = 0; i < M; ++i) s += V[k * M + i] * V[l * M + i] * dis[i];
            W[k * M + l] = sq * s;
        }
    if (evals) {
        std::vector<double> e(d);
        std::sort(e.begin(), e.end());
        std::copy(e.begin(), e.end(), evals);
    }
    return sweeps;
}

// the analysis of gridpoints g_begin, g_begin + g_stride, ... < g_end
static int letkf_analysis_range(const orc_config* cfg, int M, double* f, const double* yo, long long g_begin,
                                long long g_end, int g_stride, int dump_g0, int dump_n, double* W, double* w,
                                double* evals) {
    const orc_config& c = *cfg;
    const ObsInfo o = obs_info(c);
    const long long N = (long long)c.nx * c.ny;
    const int n_obs = o.n_sites * o.n_vars;
    if (g_stride < 1) g_stride = 1;
    std::vector<double> hx(size_t(n_obs) * M), ybar(n_obs);
    orc_obs_moments(cfg, M, f, hx.data());
    for (int a = 0; a < n_obs; ++a) {  // letkf_solver.h:239 column mean
        double s = 0;
        for (int k = 0; k < M; ++k) s += hx[(size_t)a * M + k];
        ybar[a] = s / M;
    }
    int bad = 0;
    const int cap = 4096;
#pragma omp parallel
    {
        std::vector<int> sites(cap);
        std::vector<double> wl(cap), dY, rinv, dyo, Wl(size_t(M) * M), wv(M), ev(M);
#pragma omp for schedule(dynamic, 64)
        for (long long g = g_begin; g < g_end; g += g_stride) {
            const int n = std::min(orc_local_obs(cfg, int(g), sites.data(), wl.data(), cap), cap);
            dY.assign(size_t(n) * o.n_vars * M, 0.0);
            rinv.assign(size_t(n) * o.n_vars, 0.0);
            dyo.assign(size_t(n) * o.n_vars, 0.0);
            int p = 0;
            for (int e = 0; e < n; ++e)
                for (int a = 0; a < o.n_vars; ++a, ++p) {
                    const long long ob = (long long)sites[e] * o.n_vars + a;
                    for (int k = 0; k < M; ++k) dY[(size_t)p * M + k] = hx[ob * M + k] - ybar[ob];
                    rinv[p] = wl[e] * o.rinv[a];
                    dyo[p] = yo[ob] - ybar[ob];
                }
            orc_letkf_transform(M, p, dY.data(), rinv.data(), dyo.data(), c.inflation, c.jacobi_tol,
                                c.jacobi_max_sweeps, Wl.data(), wv.data(), ev.data());
            if (g >= dump_g0 && g < (long long)dump_g0 + dump_n) {
                const long long t = g - dump_g0;
                if (W) std::copy(Wl.begin(), Wl.end(), W + t * M * M);
                if (w) std::copy(wv.begin(), wv.end(), w + t * M);
                if (evals) std::copy(ev.begin(), ev.end(), evals + t * M);
            }
            // Xa = xbar + dX (W + w 1^T)   (letkf_solver.h:313-316)
            for (int q = 0; q < Q; ++q) {
                std::vector<double> x(M);
                double xbar = 0;
                for (int k = 0; k < M; ++k) {
                    x[k] = f[(long long)k * Q * N + q * N + g];
                    xbar += x[k];
                }
                xbar /= M;
                for (int l = 0; l < M; ++l) {
                    double s = xbar;
                    for (int k = 0; k < M; ++k) s += (x[k] - xbar) * (Wl[k * M + l] + wv[k]);
                    if (!std::isfinite(s)) bad = 1;
                    f[(long long)l * Q * N + q * N + g] = s;
                }
            }
        }
    }
    return bad ? -3 : 0;
}

int orc_letkf_analysis(const orc_config* cfg, int M, double* f, const double* yo, int dump_g0,
                       int dump_n, double* W, double* w, double* evals, int g_stride) {
    return letkf_analysis_range(cfg, M, f, yo, 0, (long long)cfg->nx * cfg->ny, g_stride, dump_g0, dump_n, W, w,
                                evals);
}

int orc_letkf_analysis_strip(const orc_config* cfg, int M, double* f, const double* yo, int g0, int n,
                             double* W, double* w, double* evals) {
    const long long N = (long long)cfg->nx * cfg->ny;
    if (g0 < 0 || n < 0 || g0 + (long long)n > N) return -1;
    return letkf_analysis_range(cfg, M, f, yo, g0, (long long)g0 + n, 1, g0, n, W, w, evals);
}

void orc_ensemble_stats(const orc_config* cfg, int M, const double* f, const double* trho,
                        const double* tu, const double* tv, double out[4]) {
    const orc_config& c = *cfg;
    const long long N = (long long)c.nx * c.ny;
    long double serr_v = 0, ssp_v = 0, serr_r = 0, ssp_r = 0;
    for (long long n = 0; n < N; ++n) {
        long double su = 0, ssu = 0, sv = 0, ssv = 0, sr = 0, ssr = 0;
        for (int k = 0; k < M; ++k) {
            double r, u, v;
            moments(c, f + (long long)k * Q * N, N, n, r, u, v);
            su += u; ssu += (long double)u * u;
            sv += v; ssv += (long double)v * v;
            sr += r; ssr += (long double)r * r;
        }
        auto sd = [&](long double s, long double ss) {
            const long double m = s / M, var = ss / M - m * m;
            return var <= 0 ? 0.0L : std::sqrt(var);
        };
        const long double mu = su / M, mv = sv / M, mr = sr / M;
        const long double du = sd(su, ssu), dv = sd(sv, ssv), dr = sd(sr, ssr);
        serr_v += (mu - tu[n]) * (mu - tu[n]) + (mv - tv[n]) * (mv - tv[n]);
        ssp_v += du * du + dv * dv;
        serr_r += (mr - trho[n]) * (mr - trho[n]);
        ssp_r += dr * dr;
    }
    out[0] = double(std::sqrt(serr_v / (2 * N)));
    out[1] = double(std::sqrt(ssp_v / (2 * N)));
    out[2] = double(std::sqrt(serr_r / N));
    out[3] = double(std::sqrt(ssp_r / N));
}

}  // extern "C"

// ---------------------------------------------------------------------------
// learned eddy-viscosity closure (an extension beyond the reference: no reference code)
// ---------------------------------------------------------------------------
namespace {

constexpr int CL_IN = 18, CL_H = 64;
constexpr int CL_W1 = 0, CL_B1 = CL_W1 + CL_H * CL_IN, CL_W2 = CL_B1 + CL_H, CL_B2 = CL_W2 + CL_H * CL_H,
              CL_W3 = CL_B2 + CL_H, CL_B3 = CL_W3 + CL_H, CL_N = CL_B3 + 1;
static_assert(CL_N == ORC_CLOSURE_NPARAM, "parameter count");

// round to the nearest bf16 (ties to even), as the GPU's conversion does
double bf16r(double x) {
    float f = float(x);
    uint32_t b;
    std::memcpy(&b, &f, 4);
    if ((b & 0x7f800000u) == 0x7f800000u) return x;  // inf / nan
    const uint32_t lsb = (b >> 16) & 1u;
    b += 0x7fffu + lsb;
    b &= 0xffff0000u;
    std::memcpy(&f, &b, 4);
    return double(f);
}

const double GELU_K = std::sqrt(2.0 / PI);
inline double gelu(double x) { return 0.5 * x * (1.0 + std::tanh(GELU_K * (x + 0.044715 * x * x * x))); }
inline double gelu_d(double x) {
    const double t = std::tanh(GELU_K * (x + 0.044715 * x * x * x));
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * GELU_K * (1.0 + 3.0 * 0.044715 * x * x);
}
inline double softplus(double z) { return std::max(z, 0.0) + std::log1p(std::exp(-std::fabs(z))); }
inline double sigmoid(double z) { return 1.0 / (1.0 + std::exp(-z)); }

struct Fwd {
    double p1[CL_H], h1[CL_H], p2[CL_H], h2[CL_H], z;
};

void mlp_forward(const double* P, const double* x_in, int bf16, Fwd& F) {
    double x[CL_IN];
    for (int i = 0; i < CL_IN; ++i) x[i] = bf16 ? bf16r(x_in[i]) : x_in[i];
    for (int o = 0; o < CL_H; ++o) {
        double s = 0;
        for (int i = 0; i < CL_IN; ++i) s += (bf16 ? bf16r(P[CL_W1 + o * CL_IN + i]) : P[CL_W1 + o * CL_IN + i]) * x[i];
        F.p1[o] = s + P[CL_B1 + o];
        F.h1[o] = gelu(F.p1[o]);
    }
    for (int o = 0; o < CL_H; ++o) {
        double s = 0;
        for (int i = 0; i < CL_H; ++i) {
            const double w = bf16 ? bf16r(P[CL_W2 + o * CL_H + i]) : P[CL_W2 + o * CL_H + i];
            s += w * (bf16 ? bf16r(F.h1[i]) : F.h1[i]);
        }
        F.p2[o] = s + P[CL_B2 + o];
        F.h2[o] = gelu(F.p2[o]);
    }
    double z = P[CL_B3];
    for (int i = 0; i < CL_H; ++i) z += P[CL_W3 + i] * F.h2[i];
    F.z = z;
}

// strain (a, b) of node (i, j) by the central differences of src/lbm.cu:44-47
inline void strain_ab(int nx, int ny, double dx, const double* u, const double* v, int i, int j, double& a,
                      double& b) {
    const long long e = pmod(i + 1, nx) + (long long)nx * j, w = pmod(i - 1, nx) + (long long)nx * j;
    const long long n = i + (long long)nx * pmod(j + 1, ny), s = i + (long long)nx * pmod(j - 1, ny);
    const double ux = (u[e] - u[w]) / (2 * dx), uy = (u[n] - u[s]) / (2 * dx);
    const double vx = (v[e] - v[w]) / (2 * dx), vy = (v[n] - v[s]) / (2 * dx);
    a = 0.5 * (ux - vy);
    b = 0.5 * (uy + vx);
}

}  // namespace

void orc_closure_features(int nx, int ny, double dx, double strain_scale, const double* u, const double* v,
                          double* feat, double* ab) {
#pragma omp parallel for
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const long long n = i + (long long)nx * j;
            for (int dj = -1; dj <= 1; ++dj)
                for (int di = -1; di <= 1; ++di) {
                    double a, b;
                    strain_ab(nx, ny, dx, u, v, pmod(i + di, nx), pmod(j + dj, ny), a, b);
                    const int s = (dj + 1) * 3 + (di + 1);
                    feat[n * CL_IN + 2 * s] = a / strain_scale;
                    feat[n * CL_IN + 2 * s + 1] = b / strain_scale;
                    if (di == 0 && dj == 0 && ab) {
                        ab[2 * n] = a;
                        ab[2 * n + 1] = b;
                    }
                }
        }
}

void orc_closure_nut(const double* params, const double* feat, long long n, double nu_scale, int bf16, double* nut) {
#pragma omp parallel for
    for (long long k = 0; k < n; ++k) {
        Fwd F;
        mlp_forward(params, feat + k * CL_IN, bf16, F);
        nut[k] = nu_scale * softplus(F.z);
    }
}

double orc_closure_loss_grad(const double* P, const double* samples, long long n, double nu_scale, double w_norm,
                             double* grad) {
    std::vector<double> g(CL_N, 0.0);
    long double loss = 0;
    for (long long k = 0; k < n; ++k) {
        const double* smp = samples + k * ORC_CLOSURE_SAMPLE;
        Fwd F;
        mlp_forward(P, smp, 0, F);
        const double a = smp[18], b = smp[19], t1 = smp[20], t2 = smp[21];
        const double nu = nu_scale * softplus(F.z);
        const double r1 = 2 * nu * a + t1, r2 = 2 * nu * b + t2;
        loss += w_norm * (r1 * r1 + r2 * r2);
        const double dnu = w_norm * (4 * r1 * a + 4 * r2 * b);
        const double dz = dnu * nu_scale * sigmoid(F.z);
        double dp2[CL_H], dh1[CL_H];
        for (int i = 0; i < CL_H; ++i) {
            g[CL_W3 + i] += dz * F.h2[i];
            dp2[i] = dz * P[CL_W3 + i] * gelu_d(F.p2[i]);
        }
        g[CL_B3] += dz;
        for (int i = 0; i < CL_H; ++i) dh1[i] = 0;
        for (int o = 0; o < CL_H; ++o) {
            g[CL_B2 + o] += dp2[o];
            for (int i = 0; i < CL_H; ++i) {
                g[CL_W2 + o * CL_H + i] += dp2[o] * F.h1[i];
                dh1[i] += P[CL_W2 + o * CL_H + i] * dp2[o];
            }
        }
        for (int o = 0; o < CL_H; ++o) {
            const double dp1 = dh1[o] * gelu_d(F.p1[o]);
            g[CL_B1 + o] += dp1;
            for (int i = 0; i < CL_IN; ++i) g[CL_W1 + o * CL_IN + i] += dp1 * smp[i];
        }
    }
    for (int i = 0; i < CL_N; ++i) grad[i] = g[i] / double(n);
    return double(loss / n);
}

void orc_coarse_grain(int nxf, int nyf, int F, const double* rho, const double* u, const double* v,
                      double dx_coarse, double strain_scale, double* samples, double* coarse) {
    const int nx = nxf / F, ny = nyf / F;
    const long long N = (long long)nx * ny;
    std::vector<double> uu(N), uv(N), vv(N);
    double* rc = coarse;
    double* uc = coarse + N;
    double* vc = coarse + 2 * N;
    for (int J = 0; J < ny; ++J)
        for (int I = 0; I < nx; ++I) {
            double sr = 0, su = 0, sv = 0, suu = 0, suv = 0, svv = 0;
            for (int b = 0; b < F; ++b)
                for (int a = 0; a < F; ++a) {
                    const long long k = (long long)(F * I + a) + (long long)nxf * (F * J + b);
                    sr += rho[k];
                    su += u[k];
                    sv += v[k];
                    suu += u[k] * u[k];
                    suv += u[k] * v[k];
                    svv += v[k] * v[k];
                }
            const double inv = 1.0 / (F * F);
            const long long n = I + (long long)nx * J;
            rc[n] = sr * inv;
            uc[n] = su * inv;
            vc[n] = sv * inv;
            uu[n] = suu * inv;
            uv[n] = suv * inv;
            vv[n] = svv * inv;
        }
    std::vector<double> feat(N * CL_IN), ab(2 * N);
    orc_closure_features(nx, ny, dx_coarse, strain_scale, uc, vc, feat.data(), ab.data());
    for (long long n = 0; n < N; ++n) {
        double* s = samples + n * ORC_CLOSURE_SAMPLE;
        for (int i = 0; i < CL_IN; ++i) s[i] = feat[n * CL_IN + i];
        s[18] = ab[2 * n];
        s[19] = ab[2 * n + 1];
        const double t11 = uu[n] - uc[n] * uc[n], t12 = uv[n] - uc[n] * vc[n], t22 = vv[n] - vc[n] * vc[n];
        s[20] = 0.5 * (t11 - t22);
        s[21] = t12;
    }
}

void orc_lbm_step_closure(const orc_config* cfg, int M, const double* fin, double* fout, long long step,
                          const double* params, int bf16) {
    const orc_config& c = *cfg;
    const long long N = (long long)c.nx * c.ny;
    std::vector<double> Fx(N, 0.0), Fy(N, 0.0);
    if (c.force_kind == 1) orc_force_field(cfg, step, Fx.data(), Fy.data());
    std::vector<double> r(N), u(N), v(N), nus(N), nut(N), feat(N * CL_IN);
    for (int k = 0; k < M; ++k) {
        const double* f = fin + (long long)k * Q * N;
        double* fn = fout + (long long)k * Q * N;
        orc_stream(cfg, f, fn);
        orc_macro(cfg, fn, r.data(), u.data(), v.data());
        orc_les(cfg, u.data(), v.data(), nus.data());
        orc_closure_features(c.nx, c.ny, c.dx, c.closure_strain_scale, u.data(), v.data(), feat.data(), nullptr);
        orc_closure_nut(params, feat.data(), N, c.closure_nu_scale, bf16, nut.data());
#pragma omp parallel for
        for (long long n = 0; n < N; ++n) {
            const double nt = (std::isfinite(nut[n]) && nut[n] >= 0 && nut[n] <= c.closure_nu_max) ? nut[n] : nus[n];
            const double tau = 0.5 + 3.0 * (c.nu + nt) / c.c / c.c / c.dt;
            const double omega = 1.0 / tau;
            for (int q = 0; q < Q; ++q) {
                double fc = fn[q * N + n];
                fc -= omega * (fc - feq_q(c, r[n], u[n], v[n], q));
                fc += force_q(c, Fx[n], Fy[n], q, r[n], u[n], v[n]);
                fn[q * N + n] = fc;
            }
        }
    }
}

// ---- nudging and the Lyapunov perturbation ---------------------------------------
void orc_nudge(const orc_config* cfg, int M, double* f, const double* rho, const double* u, const double* v,
               double alpha, int xyprune, int quadra) {
    const orc_config& c = *cfg;
    const int nx = c.nx, ny = c.ny;
    const long long N = (long long)nx * ny;
    const bool skip_east = quadra % 2 == 1, skip_north = quadra >= 2;  // src/nudging.cu:26-27
#pragma omp parallel for
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            if ((skip_east && i >= nx / 2) || (skip_north && j >= ny / 2)) continue;
            const long long ij = i + (long long)nx * j;
            double ro = 0, uo, vo;
            if (xyprune == 1) {
                uo = u[ij];
                vo = v[ij];
                if (rho) ro = rho[ij];
            } else {  // src/nudging.cu:100-126: bilinear from the lattice corners
                const int iw = (i / xyprune) * xyprune, js = (j / xyprune) * xyprune;
                const int ie = pmod(iw + xyprune, nx), jn = pmod(js + xyprune, ny);
                const double ir = double(i - iw) / xyprune, jr = double(j - js) / xyprune;
                auto itp = [&](const double* a) {
                    return (1 - ir) * (1 - jr) * a[iw + (long long)nx * js] + ir * (1 - jr) * a[ie + (long long)nx * js] +
                           (1 - ir) * jr * a[iw + (long long)nx * jn] + ir * jr * a[ie + (long long)nx * jn];
                };
                uo = itp(u);
                vo = itp(v);
                if (rho) ro = itp(rho);
            }
            for (int m = 0; m < M; ++m) {
                double* fm = f + (long long)m * Q * N;
                double r = ro;
                if (!rho) {  // src/nudging.cu:128-131
                    r = 0;
                    for (int q = 0; q < Q; ++q) r += fm[q * N + ij];
                }
                for (int q = 0; q < Q; ++q)  // src/nudging.cu:133-138
                    fm[q * N + ij] = alpha * feq_q(c, r, uo, vo, q) + (1 - alpha) * fm[q * N + ij];
            }
        }
}

void orc_perturb_draws(const orc_config* cfg, unsigned long long ensemble_id, double* ur, double* vr) {
    std::mt19937 eng(uint32_t(ensemble_id + 132333ull));
    std::normal_distribution<float> nd(0.f, 1.f);
    const long long N = (long long)cfg->nx * cfg->ny;
    const float u_ref = float(cfg->u_ref);
    for (long long ij = 0; ij < N; ++ij) {
        ur[ij] = u_ref * nd(eng);
        vr[ij] = u_ref * nd(eng);
    }
}

void orc_perturb(const orc_config* cfg, double* f, const double* ur, const double* vr, double eps) {
    const orc_config& c = *cfg;
    const long long N = (long long)c.nx * c.ny;
#pragma omp parallel for
    for (long long ij = 0; ij < N; ++ij) {
        double r = 0;
        for (int q = 0; q < Q; ++q) r += f[q * N + ij];
        for (int q = 0; q < Q; ++q) f[q * N + ij] = eps * feq_q(c, r, ur[ij], vr[ij], q) + (1 - eps) * f[q * N + ij];
    }
}
