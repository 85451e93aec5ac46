// Raw field files of the reference's OSSE layout (src/io.cu:22-258):
//   <prefix>/{nature,observed,calc}/ens%04d/{rho,u,v,vor}_step%010d.dat
// float32, row-major [ny][nx] (index i + nx*j), no header; the step number is the
// output index t / iiter.  Ensemble statistics go to <prefix>/calc/ens_<name>_step....
#pragma once

#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace lbmletkf_io {

// zero-padded decimal (src/util/stringutils.hpp:10-20): 0 <= t < 10^digits
inline std::string aligned(long long t, int digits) {
    long long lim = 1;
    for (int k = 0; k < digits; ++k) lim *= 10;
    if (t < 0 || t >= lim) throw std::runtime_error("value out of range for " + std::to_string(digits) + " digits");
    std::string s = std::to_string(t);
    return std::string(size_t(digits) - s.size(), '0') + s;
}

inline std::string ens_dir(const std::string& prefix, const std::string& kind, int ens) {
    return prefix + "/" + kind + "/ens" + aligned(ens, 4);
}

inline std::string step_file(const std::string& dir, const std::string& var, long long step) {
    return dir + "/" + var + "_step" + aligned(step, 10) + ".dat";
}

// mkdir -p
inline void make_dirs(const std::string& path) {
    std::string cur;
    size_t pos = 0;
    while (pos != std::string::npos) {
        pos = path.find('/', pos + 1);
        cur = path.substr(0, pos);
        if (cur.empty()) continue;
        if (::mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
            throw std::runtime_error("could not create directory " + cur + ": " + std::strerror(errno));
    }
}

inline void write_field(const std::string& fname, const float* data, size_t n) {
    std::FILE* fp = std::fopen(fname.c_str(), "wb");
    if (!fp) throw std::runtime_error("could not open file: " + fname);
    const size_t w = std::fwrite(data, sizeof(float), n, fp);
    std::fclose(fp);
    if (w != n) throw std::runtime_error("short write: " + fname);
}

inline void read_field(const std::string& fname, float* data, size_t n) {
    std::FILE* fp = std::fopen(fname.c_str(), "rb");
    if (!fp) throw std::runtime_error("could not open file: " + fname);
    const size_t r = std::fread(data, sizeof(float), n, fp);
    std::fclose(fp);
    if (r != n) throw std::runtime_error("short read (" + std::to_string(r) + " of " + std::to_string(n) + "): " + fname);
}

// Observation noise of the reference's OBSERVE mode with OBSERVE_ERROR_RUV
// (src/io.cu:165-179): per node in order rho, u, v, the truth plus sigma times a draw
// of std::normal_distribution<float>(0, 1) on one std::mt19937 engine (seeded
// 334 + ensemble id, alive for the whole run).  The draws are taken at every node,
// observed or not.
template <class Engine, class Normal>
void add_obs_noise(Engine& eng, Normal& nd, const float* r, const float* u, const float* v, size_t n,
                   float sigma_rho, float sigma_u, float* ro, float* uo, float* vo) {
    for (size_t ij = 0; ij < n; ++ij) {
        ro[ij] = r[ij] + sigma_rho * nd(eng);
        uo[ij] = u[ij] + sigma_u * nd(eng);
        vo[ij] = v[ij] + sigma_u * nd(eng);
    }
}

// vorticity vx - uy by periodic central differences (src/io.cu:61-86), FP32
inline void vorticity(const float* u, const float* v, int nx, int ny, float dx, float* vor) {
    for (int j = 0; j < ny; ++j) {
        const int js = (j - 1 + ny) % ny, jn = (j + 1) % ny;
        for (int i = 0; i < nx; ++i) {
            const int iw = (i - 1 + nx) % nx, ie = (i + 1) % nx;
            const float uy = (u[i + nx * jn] - u[i + nx * js]) / (2 * dx);
            const float vx = (v[ie + nx * j] - v[iw + nx * j]) / (2 * dx);
            vor[i + nx * j] = vx - uy;
        }
    }
}

}  // namespace lbmletkf_io
