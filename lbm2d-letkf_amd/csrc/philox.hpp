// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11), used for every random draw of the
// library: forcing phases, initial spectra of the K4 kind and observation noise.
// It replaces cuRAND MTGP32 (src/data.cu:227-247 of the reference), whose stream
// cannot be reproduced without cuRAND.  Being counter based, the same draw is
// obtained on the host and on the device from (key, counter) alone.
#pragma once
#include <cstdint>
#include <cmath>

#if defined(__HIPCC__)
#define LBMLETKF_HD __host__ __device__ inline
#else
#define LBMLETKF_HD inline
#endif

namespace lbmletkf {

struct u32x4 { uint32_t v[4]; };

LBMLETKF_HD void philox_mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
    const uint64_t p = uint64_t(a) * uint64_t(b);
    hi = uint32_t(p >> 32);
    lo = uint32_t(p);
}

LBMLETKF_HD u32x4 philox4x32_10(u32x4 ctr, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        philox_mulhilo(M0, ctr.v[0], hi0, lo0);
        philox_mulhilo(M1, ctr.v[2], hi1, lo1);
        u32x4 n;
        n.v[0] = hi1 ^ ctr.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ ctr.v[3] ^ k1;
        n.v[3] = lo0;
        ctr = n;
        k0 += W0;
        k1 += W1;
    }
    return ctr;
}

// uniform in the open interval (0,1)
LBMLETKF_HD double u32_to_open01(uint32_t x) { return (double(x) + 0.5) * (1.0 / 4294967296.0); }

// Four N(0,1) deviates from one Philox block by the Box-Muller transform, in FP64.
LBMLETKF_HD void philox_normal4(uint64_t key, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                double out[4]) {
    u32x4 ctr{{c0, c1, c2, c3}};
    const u32x4 r = philox4x32_10(ctr, uint32_t(key), uint32_t(key >> 32));
    const double two_pi = 6.283185307179586476925286766559;
    for (int h = 0; h < 2; ++h) {
        const double u1 = u32_to_open01(r.v[2 * h]);
        const double u2 = u32_to_open01(r.v[2 * h + 1]);
        const double rad = sqrt(-2.0 * log(u1));
        out[2 * h] = rad * cos(two_pi * u2);
        out[2 * h + 1] = rad * sin(two_pi * u2);
    }
}

// Counter words that separate the streams of the library.
enum : uint32_t {
    STREAM_FORCE = 0x464f5243u,  // 'FORC'
    STREAM_IC    = 0x49434b34u,  // 'ICK4'
    STREAM_OBS   = 0x4f425345u,  // 'OBSE'
};

}  // namespace lbmletkf
