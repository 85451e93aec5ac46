// Offline validation tools on the raw field files (host only, no GPU):
//
//   lbmletkf_stat ensemble NATURE_DIR CALC_DIR T_ITER --members N --nx NX [--ny NY]
//                 [--skip S] [--first-member K] [--grid-quirk]
//       ensemble_stat/main.cpp:35-177: for every output step t <= T_ITER with
//       t % S == 0, the ensemble mean and biased (1/N) standard deviation of u, v,
//       rho over the members CALC_DIR/ens%04d (K..K+N-1), then
//         spread = sqrt(sum_i (sd_u^2 + sd_v^2) / 2n),
//         rmse   = sqrt(sum_i ((mean_u - u_nat)^2 + (mean_v - v_nat)^2) / 2n)
//       and the same for rho, against NATURE_DIR/ens0000.  The reference hard-codes
//       n = 128*128 whatever the grid (ensemble_stat/main.cpp:22); the actual grid
//       size is used unless --grid-quirk asks for the reference's count.
//       Prints "t_step, spread, rmse, rho_spread, rho_rmse, max(debug) = ..." per step
//       on stdout and the CSV row "t, spread, rmse, rho_spread, rho_rmse" on stderr.
//
//   lbmletkf_stat bindiff DIR0 DIR1 T_ITER [--skip S] [--vars u,v]
//       bindiff/main.cpp:30-66: the RMS difference of the u and v files of two
//       directories at every output step; prints "t,rmse" on stdout and
//       "t: rmse (count = n)" on stderr.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "datio.hpp"

namespace {

using lbmletkf_io::aligned;
using lbmletkf_io::read_field;
using lbmletkf_io::step_file;

[[noreturn]] void usage(const char* msg) {
    std::fprintf(stderr, "lbmletkf_stat: %s\n", msg);
    std::fprintf(stderr,
                 "usage: lbmletkf_stat ensemble NATURE_DIR CALC_DIR T_ITER --members N --nx NX [--ny NY]\n"
                 "                     [--skip S] [--first-member K] [--grid-quirk]\n"
                 "       lbmletkf_stat bindiff DIR0 DIR1 T_ITER [--skip S] [--vars u,v]\n");
    std::exit(2);
}

int ensemble(int argc, char** argv) {
    if (argc < 5) usage("ensemble needs NATURE_DIR CALC_DIR T_ITER");
    const std::string nat = argv[2], cal = argv[3];
    const int t_iter = std::stoi(argv[4]);
    int members = 0, nx = 0, ny = 0, skip = 10, first = 0;
    bool quirk = false;
    for (int a = 5; a < argc; ++a) {
        const std::string k = argv[a];
        auto val = [&]() -> int {
            if (a + 1 >= argc) usage(("missing value for " + k).c_str());
            return std::stoi(argv[++a]);
        };
        if (k == "--members") members = val();
        else if (k == "--nx") nx = val();
        else if (k == "--ny") ny = val();
        else if (k == "--skip") skip = val();
        else if (k == "--first-member") first = val();
        else if (k == "--grid-quirk") quirk = true;
        else usage(("unknown option " + k).c_str());
    }
    if (ny == 0) ny = nx;
    if (members < 1 || nx < 1 || skip < 1) usage("--members and --nx are required");
    const size_t n = size_t(nx) * ny;
    // the reference reads and averages only the first 128*128 values of each file
    const size_t count = quirk ? std::min<size_t>(n, 128 * 128) : n;
    std::printf("dirinfo: nature=%s -- ensemble=%s\n", nat.c_str(), cal.c_str());
    std::vector<float> un(n), vn(n), rn(n);
    std::vector<std::vector<float>> ue(members, std::vector<float>(n)), ve(members, std::vector<float>(n)),
        re(members, std::vector<float>(n));
    for (int t = 0; t <= t_iter; ++t) {
        if (t % skip != 0) continue;
        const std::string nd = nat + "/ens0000";
        read_field(step_file(nd, "u", t), un.data(), n);
        read_field(step_file(nd, "v", t), vn.data(), n);
        read_field(step_file(nd, "rho", t), rn.data(), n);
        for (int k = 0; k < members; ++k) {
            const std::string ed = cal + "/ens" + aligned(first + k, 4);
            read_field(step_file(ed, "u", t), ue[k].data(), n);
            read_field(step_file(ed, "v", t), ve[k].data(), n);
            read_field(step_file(ed, "rho", t), re[k].data(), n);
        }
        long double sp = 0, se = 0, rsp = 0, rse = 0, mx = 0;
        for (size_t i = 0; i < count; ++i) {
            // member sums in long double, as the reference's sum_t (ensemble_stat/main.cpp:118-131)
            long double s[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
            for (int k = 0; k < members; ++k) {
                const long double x[3] = {ue[k][i], ve[k][i], re[k][i]};
                for (int a = 0; a < 3; ++a) {
                    s[a] += x[a];
                    s2[a] += x[a] * x[a];
                }
            }
            float mean[3], sd[3];
            for (int a = 0; a < 3; ++a) {
                const long double m = s[a] / members, ms = s2[a] / members;
                const long double var = ms - m * m;
                mean[a] = float(m);  // stored as float, as the reference's buffers
                sd[a] = float(var <= 0 ? 0.0L : std::sqrt(var));
            }
            const long double eu = (long double)mean[0] - un[i], ev = (long double)mean[1] - vn[i];
            se += eu * eu + ev * ev;
            sp += (long double)sd[0] * sd[0] + (long double)sd[1] * sd[1];
            const long double er = (long double)mean[2] - rn[i];
            rse += er * er;
            rsp += (long double)sd[2] * sd[2];
            mx = std::max<long double>(mx, std::max(mean[0], mean[1]));  // starts at 0, as the reference
        }
        const long double cvel = 2.0L * count, crho = count;
        const double spread = double(std::sqrt(sp / cvel)), rmse = double(std::sqrt(se / cvel));
        const double rspread = double(std::sqrt(rsp / crho)), rrmse = double(std::sqrt(rse / crho));
        std::printf("t_step, spread, rmse, rho_spread, rho_rmse, max(debug) = %d, %.9g, %.9g, %.9g, %.9g, %.9g\n", t,
                    spread, rmse, rspread, rrmse, double(mx));
        std::fprintf(stderr, "%d, %.9g, %.9g, %.9g, %.9g\n", t, spread, rmse, rspread, rrmse);
    }
    return 0;
}

int bindiff(int argc, char** argv) {
    if (argc < 5) usage("bindiff needs DIR0 DIR1 T_ITER");
    const std::string d0 = argv[2], d1 = argv[3];
    const int t_iter = std::stoi(argv[4]);
    int skip = 10;
    std::vector<std::string> vars = {"u", "v"};
    for (int a = 5; a < argc; ++a) {
        const std::string k = argv[a];
        if (a + 1 >= argc) usage(("missing value for " + k).c_str());
        const std::string v = argv[++a];
        if (k == "--skip") skip = std::stoi(v);
        else if (k == "--vars") {
            vars.clear();
            size_t p = 0;
            while (p <= v.size()) {
                const size_t q = v.find(',', p);
                vars.push_back(v.substr(p, q == std::string::npos ? std::string::npos : q - p));
                if (q == std::string::npos) break;
                p = q + 1;
            }
        } else usage(("unknown option " + k).c_str());
    }
    if (skip < 1) usage("--skip must be positive");
    for (int t = 0; t <= t_iter; ++t) {
        if (t % skip != 0) continue;
        long long count = 0;
        long double sse = 0;
        for (const std::string& var : vars) {
            const std::string f0 = step_file(d0, var, t), f1 = step_file(d1, var, t);
            std::FILE* a = std::fopen(f0.c_str(), "rb");
            if (!a) throw std::runtime_error("could not read file: " + f0);
            std::FILE* b = std::fopen(f1.c_str(), "rb");
            if (!b) {
                std::fclose(a);
                throw std::runtime_error("could not read file: " + f1);
            }
            // element by element up to the shorter file, as the reference's read loop
            float x0 = 0, x1 = 0;
            while (std::fread(&x0, sizeof(float), 1, a) == 1 && std::fread(&x1, sizeof(float), 1, b) == 1) {
                const float e = x0 - x1;
                sse += (long double)e * e;
                ++count;
            }
            std::fclose(a);
            std::fclose(b);
        }
        const double rmse = count ? double(std::sqrt(sse / count)) : NAN;
        std::fprintf(stderr, "%d: %.9g (count = %lld)\n", t, rmse, count);
        std::printf("%d,%.9g\n", t, rmse);
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc < 2) usage("missing command");
        const std::string cmd = argv[1];
        if (cmd == "ensemble") return ensemble(argc, argv);
        if (cmd == "bindiff") return bindiff(argc, argv);
        usage(("unknown command " + cmd).c_str());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "lbmletkf_stat: %s\n", e.what());
        return 1;
    }
}
