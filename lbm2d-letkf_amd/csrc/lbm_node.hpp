// Per-node pieces of the D2Q9 BGK step shared by the forecast kernels: pull
// streaming, moments, the Smagorinsky viscosity, equilibrium and collision.
#pragma once
#include "context.hpp"

namespace lbmletkf {
namespace lbmnode {

constexpr float W0 = 4.f / 9.f, W1 = 1.f / 9.f, W2 = 1.f / 36.f;

// Multiply-adds are fused only inside one statement in the per-node functions, so
// their rounding does not depend on the kernel they are inlined into: every shape
// of the step (tiles, row strips, two-step strips, the learned-closure step's
// fallback) gives the same bits for the same node.
#define LBMNODE_CONTRACT_PER_STATEMENT _Pragma("clang fp contract(on)")

__device__ __forceinline__ int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// Pull streaming of the 9 populations into node (i,j): fc[q] = f[q](i-ex_q, j-ey_q)
// with ex_q = q%3-1, ey_q = q/3-1 (src/config.h:256-258, src/lbm.cu:81-89).
// im/ip are the wrapped columns i-1/i+1, rm/r0/rp the row offsets of j-1/j/j+1.
__device__ __forceinline__ void pull9(const float* __restrict__ f, long long nodes, int im, int i,
                                      int ip, long long rm, long long r0, long long rp,
                                      float fc[Q]) {
    fc[0] = f[0 * nodes + rp + ip];
    fc[1] = f[1 * nodes + rp + i];
    fc[2] = f[2 * nodes + rp + im];
    fc[3] = f[3 * nodes + r0 + ip];
    fc[4] = f[4 * nodes + r0 + i];
    fc[5] = f[5 * nodes + r0 + im];
    fc[6] = f[6 * nodes + rm + ip];
    fc[7] = f[7 * nodes + rm + i];
    fc[8] = f[8 * nodes + rm + im];
}

// rho, u = c sum f ex / rho, v = c sum f ey / rho   (src/lbm.h:72-109), the
// quotients through the hardware reciprocal of rho (1 ulp; rho ~ 1).  The sums
// are grouped symmetrically under the reflection y -> -y (q = 0, 1, 2 <-> 6, 7, 8):
// on the reflected populations they give the same rho and u and exactly -v, so a
// kernel may step the reflected lattice and reflect back bit for bit.
__device__ __forceinline__ void moments(const float fc[Q], float c, float& rho, float& u, float& v) {
    LBMNODE_CONTRACT_PER_STATEMENT
    const float r = (((fc[0] + fc[6]) + (fc[1] + fc[7])) + ((fc[2] + fc[8]) + (fc[3] + fc[5]))) + fc[4];
    const float jx = ((fc[2] + fc[8]) + fc[5]) - ((fc[0] + fc[6]) + fc[3]);
    const float jy = (fc[6] + fc[7] + fc[8]) - (fc[0] + fc[1] + fc[2]);
    const float ir = __builtin_amdgcn_rcpf(r);
    rho = r;
    u = (c * jx) * ir;
    v = (c * jy) * ir;
}

// nu_t = Cs^2 dx^2 sqrt(2 S:S) from central differences (src/lbm.cu:36-60).
__device__ __forceinline__ float smagorinsky(float uE, float uW, float uN, float uS, float vE,
                                             float vW, float vN, float vS, const LbmParams& p) {
    LBMNODE_CONTRACT_PER_STATEMENT
    const float ux = (uE - uW) * p.inv_2dx;
    const float uy = (uN - uS) * p.inv_2dx;
    const float vx = (vE - vW) * p.inv_2dx;
    const float vy = (vN - vS) * p.inv_2dx;
    const float s01 = 0.5f * (uy + vx);
    const float ss = __builtin_amdgcn_sqrtf(2.f * (ux * ux + 2.f * s01 * s01 + vy * vy));  // 1 ulp
    return p.les_coef * ss;
}

// Equilibrium populations feq_q(rho,u,v) (src/lbm.h:31-39).  The rest population
// is taken as rho - sum of the others: in FP32 the weights 4/9, 1/9, 1/36 sum to
// 1 + 8e-9, and with omega ~ 2 that bias would add ~1.6e-8 rho of mass per node
// and step; the subtraction keeps sum_q feq = rho up to unbiased rounding.
__device__ __forceinline__ void equilibrium9(float rho, float u, float v, const LbmParams& p,
                                             float feq[Q]) {
    LBMNODE_CONTRACT_PER_STATEMENT
    // opposite directions share A = 1 + 4.5 uc^2 - 1.5 uu and differ in the sign of
    // 3 uc: feq_q = rho w_q (A + 3 uc), feq_q' = rho w_q (A - 3 uc)
    const float uu = (u * u + v * v) * p.inv_c2;
    const float base = 1.f - 1.5f * uu;
    const float rw1 = rho * W1, rw2 = rho * W2;
    const float ux = u * p.inv_c, uy = v * p.inv_c;
    const float upv = ux + uy, umv = ux - uy;
    const float ax = base + 4.5f * ux * ux, ay = base + 4.5f * uy * uy;
    const float ap = base + 4.5f * upv * upv, am = base + 4.5f * umv * umv;
    feq[5] = rw1 * (ax + 3.f * ux);    // (+1,  0)
    feq[3] = rw1 * (ax - 3.f * ux);    // (-1,  0)
    feq[7] = rw1 * (ay + 3.f * uy);    // ( 0, +1)
    feq[1] = rw1 * (ay - 3.f * uy);    // ( 0, -1)
    feq[8] = rw2 * (ap + 3.f * upv);   // (+1, +1)
    feq[0] = rw2 * (ap - 3.f * upv);   // (-1, -1)
    feq[2] = rw2 * (am + 3.f * umv);   // (+1, -1)
    feq[6] = rw2 * (am - 3.f * umv);   // (-1, +1)
    // grouped symmetrically under y -> -y, as in moments()
    const float moving = ((feq[0] + feq[6]) + (feq[1] + feq[7])) + ((feq[2] + feq[8]) + (feq[3] + feq[5]));
    feq[4] = rho - moving;
}

// BGK relaxation towards feq plus the first-order force term with linear
// friction (src/lbm_collision_srt.h:8-16, src/lbm.h:41-70).  FORCE = false is for
// runs with neither forcing nor friction, where the force term is exactly zero.
template <bool FORCE>
__device__ __forceinline__ void collide9(float fc[Q], float rho, float u, float v, float nut,
                                         float Fx, float Fy, const LbmParams& p) {
    LBMNODE_CONTRACT_PER_STATEMENT
    const float tau = 0.5f + p.k_tau * (p.nu + nut);
    const float omega = __builtin_amdgcn_rcpf(tau);  // 1 ulp
    const float fxp = (Fx - p.fric_over_dt * u) * p.dt;
    const float fyp = (Fy - p.fric_over_dt * v) * p.dt;
    float feq[Q];
    equilibrium9(rho, u, v, p, feq);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int ex = q % 3 - 1, ey = q / 3 - 1;
        const float wq = (ex == 0 && ey == 0) ? W0 : ((ex == 0 || ey == 0) ? W1 : W2);
        const float fq = fc[q] - omega * (fc[q] - feq[q]);
        if constexpr (FORCE)
            fc[q] = fq + wq * rho * p.force_scale * (float(ex) * fxp + float(ey) * fyp);
        else
            fc[q] = fq;
    }
}

}  // namespace lbmnode
}  // namespace lbmletkf
