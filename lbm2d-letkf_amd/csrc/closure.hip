// Learned eddy-viscosity closure: training data from coarse-grained runs, the
// training kernels (forward, backward, Adam) and the upload of the network to a
// context for the fused LBM kernel.  The network is defined in closure.hpp.
//
// Training data (one coarse node = one sample): a fine run of the same LBM is
// box-filtered by an integer factor F, u_bar = <u>, and the filtered subgrid
// stress tau_ij = <u_i u_j> - u_bar_i u_bar_j is the target of the eddy-viscosity
// model tau_dev = -2 nu_t S_dev(u_bar): loss = w ((2 nu a + t1)^2 + (2 nu b + t2)^2)
// with (a, b) the deviatoric strain of u_bar and t1 = (tau11 - tau22)/2, t2 = tau12.
//
// Training kernels run in FP32 on the vector ALU: one sample per lane, the
// parameters in LDS (broadcast reads), per-workgroup partial gradients summed in a
// fixed order (deterministic), then Adam.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "closure.hpp"
#include "context.hpp"
#include "lbm_node.hpp"
#include "philox.hpp"

namespace lbmletkf {

namespace {

__device__ __forceinline__ float gelu_f(float x) {
    const float k = 0.7978845608028654f;  // sqrt(2/pi)
    return 0.5f * x * (1.f + tanhf(k * (x + 0.044715f * x * x * x)));
}
__device__ __forceinline__ float gelu_df(float x) {
    const float k = 0.7978845608028654f;
    const float t = tanhf(k * (x + 0.044715f * x * x * x));
    return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * k * (1.f + 3.f * 0.044715f * x * x);
}
__device__ __forceinline__ float softplus_f(float z) { return fmaxf(z, 0.f) + log1pf(expf(-fabsf(z))); }
__device__ __forceinline__ float sigmoid_f(float z) { return 1.f / (1.f + expf(-z)); }

// Box filter by F of one fine member: per coarse node rho_bar, u_bar, v_bar and the
// subgrid stresses t11, t12, t22, summed in FP64.  The fine velocity is the
// moment c sum(f e) / rho of the stored populations (src/lbm.h:72-109).
__global__ void k_coarse_fields(const float* __restrict__ f, long long nodes_f, int nxf, int F, int nx, int ny,
                                float c_lat, float* __restrict__ out) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nc = (long long)nx * ny;
    if (n >= nc) return;
    const int I = int(n % nx), J = int(n / nx);
    double sr = 0, su = 0, sv = 0, suu = 0, suv = 0, svv = 0;
    for (int b = 0; b < F; ++b)
        for (int a = 0; a < F; ++a) {
            const long long k = (long long)(F * I + a) + (long long)nxf * (F * J + b);
            float fq[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) fq[q] = f[q * nodes_f + k];
            float r = 0.f;
#pragma unroll
            for (int q = 0; q < Q; ++q) r += fq[q];
            const float jx = (fq[2] + fq[5] + fq[8]) - (fq[0] + fq[3] + fq[6]);
            const float jy = (fq[6] + fq[7] + fq[8]) - (fq[0] + fq[1] + fq[2]);
            const double u = double(c_lat * jx / r), v = double(c_lat * jy / r);
            sr += r;
            su += u;
            sv += v;
            suu += u * u;
            suv += u * v;
            svv += v * v;
        }
    const double inv = 1.0 / double(F * F);
    const double ub = su * inv, vb = sv * inv;
    out[0 * nc + n] = float(sr * inv);
    out[1 * nc + n] = float(ub);
    out[2 * nc + n] = float(vb);
    out[3 * nc + n] = float(suu * inv - ub * ub);
    out[4 * nc + n] = float(suv * inv - ub * vb);
    out[5 * nc + n] = float(svv * inv - vb * vb);
}

// strain (a, b) at node (i, j) of the coarse velocity by central differences
__device__ __forceinline__ void strain_ab(const float* u, const float* v, int nx, int ny, int i, int j, float inv_2dx,
                                          float& a, float& b) {
    const int ie = i + 1 == nx ? 0 : i + 1, iw = i == 0 ? nx - 1 : i - 1;
    const int jn = j + 1 == ny ? 0 : j + 1, js = j == 0 ? ny - 1 : j - 1;
    const float ux = (u[ie + nx * j] - u[iw + nx * j]) * inv_2dx;
    const float uy = (u[i + nx * jn] - u[i + nx * js]) * inv_2dx;
    const float vx = (v[ie + nx * j] - v[iw + nx * j]) * inv_2dx;
    const float vy = (v[i + nx * jn] - v[i + nx * js]) * inv_2dx;
    a = 0.5f * (ux - vy);
    b = 0.5f * (uy + vx);
}

// samples [n][22] from the coarse fields (rho, u, v, t11, t12, t22)
__global__ void k_closure_samples(const float* __restrict__ fields, int nx, int ny, float inv_2dx, float inv_s0,
                                  float* __restrict__ samples) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nc = (long long)nx * ny;
    if (n >= nc) return;
    const int i = int(n % nx), j = int(n / nx);
    const float* u = fields + nc;
    const float* v = fields + 2 * nc;
    float* s = samples + n * CL_SAMPLE;
    for (int dj = -1; dj <= 1; ++dj)
        for (int di = -1; di <= 1; ++di) {
            const int ii = (i + di + nx) % nx, jj = (j + dj + ny) % ny;
            float a, b;
            strain_ab(u, v, nx, ny, ii, jj, inv_2dx, a, b);
            const int st = (dj + 1) * 3 + (di + 1);
            s[2 * st] = a * inv_s0;
            s[2 * st + 1] = b * inv_s0;
            if (di == 0 && dj == 0) {
                s[18] = a;
                s[19] = b;
            }
        }
    s[20] = 0.5f * (fields[3 * nc + n] - fields[5 * nc + n]);
    s[21] = fields[4 * nc + n];
}

// ---- training ------------------------------------------------------------------
constexpr int TB = 64;  // samples (lanes) per workgroup
constexpr int PST = CL_NPARAM + 1;  // partial record: gradient + loss

// Loss and gradient of `count` samples starting at `first` (through the
// permutation idx = (pa * k + pc) mod n_total when permute != 0); each workgroup
// writes its partial sums to partial[blockIdx.x][0..CL_NPARAM].
__global__ __launch_bounds__(TB) void k_closure_grad(const float* __restrict__ params,
                                                     const float* __restrict__ samples, long long n_total,
                                                     long long first, long long count, unsigned long long pa,
                                                     unsigned long long pc, int permute, float nu_scale, float w_norm,
                                                     float* __restrict__ partial) {
    __shared__ float P[CL_NPARAM];
    __shared__ float sx[TB][CL_IN + 1];
    __shared__ float sh1[TB][CL_H + 1];
    __shared__ float sdp1[TB][CL_H + 1];
    __shared__ float sdp2[TB][CL_H + 1];
    __shared__ float sh2[TB][CL_H + 1];
    __shared__ float sdz[TB], sloss[TB];
    const int t = threadIdx.x;
    for (int k = t; k < CL_NPARAM; k += TB) P[k] = params[k];
    __syncthreads();

    const long long k = first + (long long)blockIdx.x * TB + t;
    const bool valid = k < first + count;
    long long idx = k;
    if (permute) idx = (long long)((pa * (unsigned long long)k + pc) % (unsigned long long)n_total);
    if (!valid) idx = 0;
    const float* smp = samples + idx * CL_SAMPLE;
    float x[CL_IN];
#pragma unroll
    for (int i = 0; i < CL_IN; ++i) x[i] = smp[i];
    const float a = smp[18], b = smp[19], t1 = smp[20], t2 = smp[21];

    // forward; sdp1/sdp2 first hold gelu'(p1), gelu'(p2)
    for (int o = 0; o < CL_H; ++o) {
        float s = P[CL_B1 + o];
#pragma unroll
        for (int i = 0; i < CL_IN; ++i) s = fmaf(P[CL_W1 + o * CL_IN + i], x[i], s);
        sh1[t][o] = gelu_f(s);
        sdp1[t][o] = gelu_df(s);
    }
    float z = P[CL_B3];
    for (int o = 0; o < CL_H; ++o) {
        float s = P[CL_B2 + o];
        for (int i = 0; i < CL_H; ++i) s = fmaf(P[CL_W2 + o * CL_H + i], sh1[t][i], s);
        const float h = gelu_f(s);
        sh2[t][o] = h;
        sdp2[t][o] = gelu_df(s);
        z = fmaf(P[CL_W3 + o], h, z);
    }
    const float nu = nu_scale * softplus_f(z);
    const float r1 = 2.f * nu * a + t1, r2 = 2.f * nu * b + t2;
    sloss[t] = valid ? w_norm * (r1 * r1 + r2 * r2) : 0.f;
    const float dz = valid ? w_norm * 4.f * (r1 * a + r2 * b) * nu_scale * sigmoid_f(z) : 0.f;
    sdz[t] = dz;

    // backward to the hidden layers
    float dh1[CL_H];
#pragma unroll
    for (int i = 0; i < CL_H; ++i) dh1[i] = 0.f;
    for (int o = 0; o < CL_H; ++o) {
        const float dp2 = dz * P[CL_W3 + o] * sdp2[t][o];
        sdp2[t][o] = dp2;
#pragma unroll
        for (int i = 0; i < CL_H; ++i) dh1[i] = fmaf(P[CL_W2 + o * CL_H + i], dp2, dh1[i]);
    }
#pragma unroll
    for (int i = 0; i < CL_H; ++i) sdp1[t][i] *= dh1[i];
#pragma unroll
    for (int i = 0; i < CL_IN; ++i) sx[t][i] = x[i];
    __syncthreads();

    // parameter gradients: lane o owns row o of W1 and W2, b1[o], b2[o], w3[o]
    const int o = t;
    float gw2[CL_H], gw1[CL_IN];
#pragma unroll
    for (int i = 0; i < CL_H; ++i) gw2[i] = 0.f;
#pragma unroll
    for (int i = 0; i < CL_IN; ++i) gw1[i] = 0.f;
    float gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f, loss = 0.f;
    for (int s = 0; s < TB; ++s) {
        const float d2 = sdp2[s][o], d1 = sdp1[s][o];
#pragma unroll
        for (int i = 0; i < CL_H; ++i) gw2[i] = fmaf(d2, sh1[s][i], gw2[i]);
#pragma unroll
        for (int i = 0; i < CL_IN; ++i) gw1[i] = fmaf(d1, sx[s][i], gw1[i]);
        gb1 += d1;
        gb2 += d2;
        gw3 = fmaf(sdz[s], sh2[s][o], gw3);
        gb3 += sdz[s];
        loss += sloss[s];
    }
    float* pr = partial + (long long)blockIdx.x * PST;
#pragma unroll
    for (int i = 0; i < CL_H; ++i) pr[CL_W2 + o * CL_H + i] = gw2[i];
#pragma unroll
    for (int i = 0; i < CL_IN; ++i) pr[CL_W1 + o * CL_IN + i] = gw1[i];
    pr[CL_B1 + o] = gb1;
    pr[CL_B2 + o] = gb2;
    pr[CL_W3 + o] = gw3;
    if (o == 0) {
        pr[CL_B3] = gb3;
        pr[CL_NPARAM] = loss;
    }
}

// grad[p] = sum over workgroups (in order) / count; out[CL_NPARAM] = mean loss
__global__ void k_closure_reduce(const float* __restrict__ partial, int nb, float inv_count, float* __restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > CL_NPARAM) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += partial[(long long)b * PST + p];
    out[p] = float(s * inv_count);
}

__global__ void k_adam(float* __restrict__ params, const float* __restrict__ grad, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= CL_NPARAM) return;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
}

// mean of t1^2 + t2^2 over the samples (per-workgroup partial sums in FP64)
__global__ void k_target_power(const float* __restrict__ samples, long long n, double* __restrict__ part) {
    __shared__ double red[256];
    double s = 0.0;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
        const double t1 = samples[k * CL_SAMPLE + 20], t2 = samples[k * CL_SAMPLE + 21];
        s += t1 * t1 + t2 * t2;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// ---- the forecast step with the learned closure -------------------------------
using namespace lbmnode;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MX = 64, MY = 16, MNT = 256;   // tile of one member, threads per workgroup
constexpr int MROWS = MNT / MX;              // 4 rows per pass
constexpr int MNODES = MY / MROWS;           // 4 nodes per thread
constexpr int VX = MX + 4, VY = MY + 4;      // streamed velocity with a 2-node halo
constexpr int SX = MX + 2, SY = MY + 2;      // strain with a 1-node halo
constexpr int RING = VX * VY - MX * MY;      // 336 halo nodes

// gelu(x) = x/2 (1 + tanh(k (x + 0.044715 x^3))) = x / (1 + exp(-2k (x + 0.044715 x^3))),
// one exp2 and one reciprocal; the limits come out as 0 (x -> -inf) and x (x -> +inf)
__device__ __forceinline__ float gelu_mlp(float x) {
    const float c1 = -2.f * 0.7978845608028654f * 1.4426950408889634f;
    const float c3 = c1 * 0.044715f;
    const float e = __builtin_amdgcn_exp2f(x * fmaf(c3, x * x, c1));
    return x * __builtin_amdgcn_rcpf(1.f + e);
}

// softplus(z) = max(z, 0) + log1p(e), e = exp(-|z|), on the hardware exp2 and log2:
// log1p(e) = log(u) e / (u - 1) with u = 1 + e rounded (the rounding of u cancels),
// and e (1 - e/2) below e = 2^-10 (relative error under e^2/3).  A NaN z stays NaN.
__device__ __forceinline__ float softplus_hw(float z) {
    const float e = __builtin_amdgcn_exp2f(-fabsf(z) * 1.4426950408889634f);
    const float u = 1.f + e;
    const float big = __builtin_amdgcn_logf(u) * 0.6931471805599453f * (e * __builtin_amdgcn_rcpf(u - 1.f));
    const float lp = e < 0x1p-10f ? e * fmaf(-0.5f, e, 1.f) : big;
    return fmaxf(z, 0.f) + lp;
}

// nu_t of 16 nodes (lane % 16) from their bf16 features (this lane's k-block
// lane / 16), returned on every lane of the node; sb: this lane's bias and
// output-weight vectors in LDS, b1v at sb[64 I], b2v at sb[64 (4 + I)], w3v at
// sb[64 (8 + I)]
__device__ __forceinline__ float mlp_nu(const bf16x8 (&A1)[4], const bf16x8 (&A2)[4][2], const f32x4* sb, bf16x8 x,
                                        float b3, float nu_scale) {
    bf16x8 h[2];
#pragma unroll
    for (int I = 0; I < 4; ++I) {
        const f32x4 a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1[I], x, sb[64 * I], 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) h[I / 2][(I % 2) * 4 + r] = (__bf16)gelu_mlp(a[r]);
    }
    float zp = 0.f;
#pragma unroll
    for (int I = 0; I < 4; ++I) {
        f32x4 a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A2[I][0], h[0], sb[64 * (4 + I)], 0, 0, 0);
        a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A2[I][1], h[1], a, 0, 0, 0);
        const f32x4 w3 = sb[64 * (8 + I)];
#pragma unroll
        for (int r = 0; r < 4; ++r) zp = fmaf(w3[r], gelu_mlp(a[r]), zp);
    }
    zp += __shfl_xor(zp, 16);
    zp += __shfl_xor(zp, 32);
    return nu_scale * softplus_hw(zp + b3);
}

// One LBM step of tiles [tile0, tile0 + gridDim.x) with nu_t from the network:
// pull streaming of the tile and a 2-node halo ring, the strain (a, b) of the
// streamed velocity on the tile plus a 1-node ring (LDS), the network on MFMA for
// 16 nodes at a time (a wave takes 4 rows of 64 nodes), then the BGK collision with
// that nu_t.  Where the network gives a non-finite value or one outside
// [0, nu_max] the node falls back to the Smagorinsky nu_t of src/lbm.cu:24-65.
template <bool FORCE>
__global__ __launch_bounds__(MNT) __attribute__((amdgpu_waves_per_eu(4))) void k_lbm_mlp(const float* __restrict__ fin, float* __restrict__ fout,
                                                  const float* __restrict__ force, LbmParams p,
                                                  const ClosureWeights* __restrict__ cw, int tiles_x,
                                                  int tiles_per_member, int tile0, float* __restrict__ nut_out) {
    const int nb = gridDim.x;
    const int b = blockIdx.x;
    const int bb = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
    const int t = tile0 + bb;
    const int nx = p.nx, ny = p.ny;
    const long long nodes = p.nodes;
    const int member = t / tiles_per_member;
    const int tt = t - member * tiles_per_member;
    const int x0 = (tt % tiles_x) * MX;
    const int y0 = (tt / tiles_x) * MY;
    const float* f = fin + (long long)member * p.member_stride;
    float* fo = fout + (long long)member * p.member_stride;

    __shared__ float su[VY][VX];
    __shared__ float sv[VY][VX];
    __shared__ float2 sab[SY][SX];
    __shared__ float snu[MY][MX];
    __shared__ f32x4 sbw[12 * 64];

    const int lx = threadIdx.x % MX;
    const int ly0 = threadIdx.x / MX;
    const int i = x0 + lx;
    const int im = wrap(i - 1, nx), ip = wrap(i + 1, nx);

    // the streamed velocity of the tile; the populations are pulled again for the
    // collision (from L2), so that none is held in registers through the network
#pragma unroll
    for (int r = 0; r < MNODES; ++r) {
        const int ly = ly0 + r * MROWS;
        const int j = y0 + ly;
        const long long r0 = (long long)j * nx;
        const long long rm = (long long)wrap(j - 1, ny) * nx;
        const long long rp = (long long)wrap(j + 1, ny) * nx;
        float fc[Q], rr, uu, vv;
        pull9(f, nodes, im, i, ip, rm, r0, rp, fc);
        moments(fc, p.c, rr, uu, vv);
        su[ly + 2][lx + 2] = uu;
        sv[ly + 2][lx + 2] = vv;
    }
    // halo ring: rows -2, -1 and MY, MY+1 over VX columns, then columns -2, -1, MX,
    // MX+1 over the MY tile rows (frame coordinates hx, hy from the tile corner)
    for (int h = threadIdx.x; h < RING; h += MNT) {
        int hx, hy;
        if (h < 4 * VX) {
            const int row = h / VX;
            hy = row < 2 ? row - 2 : MY + row - 2;
            hx = h % VX - 2;
        } else {
            const int k = h - 4 * VX;
            const int col = k / MY;
            hx = col < 2 ? col - 2 : MX + col - 2;
            hy = k % MY;
        }
        const int gi = wrap(x0 + hx, nx), gj = wrap(y0 + hy, ny);
        const long long r0 = (long long)gj * nx;
        const long long rm = (long long)wrap(gj - 1, ny) * nx;
        const long long rp = (long long)wrap(gj + 1, ny) * nx;
        float fh[Q], rh, uh, vh;
        pull9(f, nodes, wrap(gi - 1, nx), gi, wrap(gi + 1, nx), rm, r0, rp, fh);
        moments(fh, p.c, rh, uh, vh);
        su[hy + 2][hx + 2] = uh;
        sv[hy + 2][hx + 2] = vh;
    }
    __syncthreads();
    // scaled strain on the tile plus a 1-node ring
    const float s_inv = cw->inv_strain_scale * p.inv_2dx;
    for (int k = threadIdx.x; k < SX * SY; k += MNT) {
        const int sy = k / SX, sx = k % SX;
        const int cy = sy + 1, cx = sx + 1;  // centre in the velocity frame
        const float ux = su[cy][cx + 1] - su[cy][cx - 1], uy = su[cy + 1][cx] - su[cy - 1][cx];
        const float vx = sv[cy][cx + 1] - sv[cy][cx - 1], vy = sv[cy + 1][cx] - sv[cy - 1][cx];
        sab[sy][sx] = make_float2(0.5f * (ux - vy) * s_inv, 0.5f * (uy + vx) * s_inv);
    }

    // network weights of this lane (layout in closure.hpp): the bf16 matrices in
    // registers, the bias and output-weight vectors (b1v, b2v, w3v, contiguous) in LDS
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    bf16x8 A1[4], A2[4][2];
#pragma unroll
    for (int I = 0; I < 4; ++I) {
        A1[I] = *reinterpret_cast<const bf16x8*>(cw->a1[I][lane]);
        A2[I][0] = *reinterpret_cast<const bf16x8*>(cw->a2[I][0][lane]);
        A2[I][1] = *reinterpret_cast<const bf16x8*>(cw->a2[I][1][lane]);
    }
    for (int k = threadIdx.x; k < 12 * 64; k += MNT) sbw[k] = reinterpret_cast<const f32x4*>(cw->b1v)[k];
    const float b3 = cw->b3, nu_scale = cw->nu_scale;
    __syncthreads();

    // features of lane: stencil points s = 4 kq .. 4 kq + 3 (s < 9) of node lane % 16
    const int col = lane % 16, kq = lane / 16;
#pragma unroll 1
    for (int g = 0; g < 16; ++g) {
        const int ly = 4 * wave + g / 4, gx = 16 * (g % 4) + col;
        bf16x8 x;
#pragma unroll
        for (int tq = 0; tq < 4; ++tq) {
            const int s = 4 * kq + tq;
            float2 ab = make_float2(0.f, 0.f);
            if (s < 9) ab = sab[ly + s / 3][gx + s % 3];
            x[2 * tq] = (__bf16)ab.x;
            x[2 * tq + 1] = (__bf16)ab.y;
        }
        // an opaque lane index keeps the LDS vectors from being hoisted into registers
        int ln = lane;
        asm volatile("" : "+v"(ln));
        const float nu = mlp_nu(A1, A2, sbw + ln, x, b3, nu_scale);
        if (kq == 0) snu[ly][gx] = nu;
    }
    __syncthreads();

    const float nu_max = cw->nu_max;
    // a pointer the compiler cannot see through: the second pull is not merged
    // with the first (whose values would otherwise stay live through the network)
    const float* f2 = f;
    asm volatile("" : "+s"(f2));
#pragma unroll
    for (int r = 0; r < MNODES; ++r) {
        const int ly = ly0 + r * MROWS;
        const int j = y0 + ly;
        const int cy = ly + 2, cx = lx + 2;
        float fc[Q], rr, uu, vv;
        {
            const long long r0 = (long long)j * nx;
            const long long rm = (long long)wrap(j - 1, ny) * nx;
            const long long rp = (long long)wrap(j + 1, ny) * nx;
            pull9(f2, nodes, im, i, ip, rm, r0, rp, fc);
            moments(fc, p.c, rr, uu, vv);
        }
        float nut = snu[ly][lx];
        if (!(nut >= 0.f && nut <= nu_max))  // also catches NaN
            nut = smagorinsky(su[cy][cx + 1], su[cy][cx - 1], su[cy + 1][cx], su[cy - 1][cx], sv[cy][cx + 1],
                              sv[cy][cx - 1], sv[cy + 1][cx], sv[cy - 1][cx], p);
        const long long node = (long long)j * nx + i;
        if (nut_out) nut_out[node] = nut;
        float Fx = 0.f, Fy = 0.f;
        if (FORCE && p.has_force) {
            Fx = force[node];
            Fy = force[nodes + node];
        }
        collide9<FORCE>(fc, rr, uu, vv, nut, Fx, Fy, p);
#pragma unroll
        for (int q = 0; q < Q; ++q) fo[q * nodes + node] = fc[q];
    }
}

unsigned short bf16_rne(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    if ((b & 0x7f800000u) == 0x7f800000u) return (unsigned short)(b >> 16);
    b += 0x7fffu + ((b >> 16) & 1u);
    return (unsigned short)(b >> 16);
}

}  // namespace

// the weight tiles of the fused kernel (layout in closure.hpp)
void closure_weights_host(const float* P, const lbmletkf_config& cfg, ClosureWeights& w) {
    std::memset(&w, 0, sizeof(w));
    for (int I = 0; I < 4; ++I)
        for (int l = 0; l < 64; ++l)
            for (int t = 0; t < 8; ++t) {
                const int row = 16 * I + l % 16, k = 8 * (l / 16) + t;
                w.a1[I][l][t] = k < CL_IN ? bf16_rne(P[CL_W1 + row * CL_IN + k]) : 0;
            }
    for (int I = 0; I < 4; ++I)
        for (int K = 0; K < 2; ++K)
            for (int l = 0; l < 64; ++l)
                for (int t = 0; t < 8; ++t) {
                    const int row = 16 * I + l % 16, kq = l / 16;
                    const int hidden = 16 * (2 * K + t / 4) + 4 * kq + t % 4;
                    w.a2[I][K][l][t] = bf16_rne(P[CL_W2 + row * CL_H + hidden]);
                }
    for (int J = 0; J < 4; ++J)
        for (int l = 0; l < 64; ++l)
            for (int r = 0; r < 4; ++r) {
                const int h = 16 * J + 4 * (l / 16) + r;
                w.b1v[J][l][r] = P[CL_B1 + h];
                w.b2v[J][l][r] = P[CL_B2 + h];
                w.w3v[J][l][r] = P[CL_W3 + h];
            }
    w.b3 = P[CL_B3];
    w.inv_strain_scale = float(1.0 / cfg.closure_strain_scale);
    w.nu_scale = float(cfg.closure_nu_scale);
    w.nu_max = float(cfg.closure_nu_max);
}

hipError_t launch_lbm_mlp(Ctx& c, const float* fin, float* fout, int tile0, int n_tiles, float* nut_out) {
    const int tiles_x = c.nx / MX, tiles = tiles_x * (c.ny / MY);
    if (!c.d_closure || c.nx % MX || c.ny % MY || tile0 < 0 || n_tiles < 1 || tile0 + n_tiles > tiles * c.M)
        return hipErrorInvalidValue;
    const bool force = c.lp.has_force || c.lp.fric_over_dt != 0.f;
    if (force)
        k_lbm_mlp<true><<<n_tiles, MNT, 0, c.stream>>>(fin, fout, c.force_xy, c.lp, c.d_closure, tiles_x, tiles, tile0,
                                                      nut_out);
    else
        k_lbm_mlp<false><<<n_tiles, MNT, 0, c.stream>>>(fin, fout, c.force_xy, c.lp, c.d_closure, tiles_x, tiles,
                                                       tile0, nut_out);
    return hipGetLastError();
}

// ---- trainer ---------------------------------------------------------------------
struct Trainer {
    lbmletkf_config cfg{};
    hipStream_t stream = nullptr;
    long long cap = 0, n = 0;
    float* d_samples = nullptr;
    float* d_params = nullptr;
    float* d_m = nullptr;
    float* d_v = nullptr;
    float* d_grad = nullptr;     // [CL_NPARAM + 1]: gradient, loss
    float* d_partial = nullptr;  // [max_blocks][PST]
    int max_blocks = 0;
    double* d_pow = nullptr;     // [256]
    float* d_fields = nullptr;   // coarse-graining scratch [6][nodes_c]
    long long fields_cap = 0;
    long long adam_t = 0;
    ~Trainer() {
        for (void* p : {(void*)d_samples, (void*)d_params, (void*)d_m, (void*)d_v, (void*)d_grad, (void*)d_partial,
                        (void*)d_pow, (void*)d_fields})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

#define TR_TRY(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return report_error(nullptr, LBMLETKF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int ensure_fields(Trainer* t, long long nc) {
    if (t->fields_cap >= nc) return LBMLETKF_OK;
    if (t->d_fields) TR_TRY(hipFree(t->d_fields));
    t->d_fields = nullptr;
    TR_TRY(hipMalloc(&t->d_fields, sizeof(float) * 6 * nc));
    t->fields_cap = nc;
    return LBMLETKF_OK;
}

// coarse fields of one fine member into t->d_fields, on the fine context's stream
int coarse_fields(Trainer* t, Ctx* fine, int member, int F, long long& nc, int& nx, int& ny) {
    if (F < 1 || fine->nx % F || fine->ny % F) return report_error(fine, LBMLETKF_ERR_ARG, "the grid is not divisible by the filter factor");
    if (member < 0 || member >= fine->M) return report_error(fine, LBMLETKF_ERR_ARG, "member index out of range");
    nx = fine->nx / F;
    ny = fine->ny / F;
    nc = (long long)nx * ny;
    if (int st = ensure_fields(t, nc)) return st;
    const float* f = fine->f[fine->cur] + (long long)member * fine->lp.member_stride;
    k_coarse_fields<<<unsigned((nc + 255) / 256), 256, 0, fine->stream>>>(f, fine->nodes, fine->nx, F, nx, ny,
                                                                       float(fine->cfg.c), t->d_fields);
    TR_TRY(hipGetLastError());
    return LBMLETKF_OK;
}

int grad_launch(Trainer* t, long long first, long long count, bool permute, unsigned long long pa,
                unsigned long long pc, float w_norm) {
    const int nb = int((count + TB - 1) / TB);
    if (nb > t->max_blocks) {
        if (t->d_partial) TR_TRY(hipFree(t->d_partial));
        t->d_partial = nullptr;
        TR_TRY(hipMalloc(&t->d_partial, sizeof(float) * (size_t)nb * PST));
        t->max_blocks = nb;
    }
    k_closure_grad<<<nb, TB, 0, t->stream>>>(t->d_params, t->d_samples, t->n, first, count, pa, pc, permute ? 1 : 0,
                                             float(t->cfg.closure_nu_scale), w_norm, t->d_partial);
    TR_TRY(hipGetLastError());
    k_closure_reduce<<<(CL_NPARAM + 1 + 255) / 256, 256, 0, t->stream>>>(t->d_partial, nb, float(1.0 / double(count)),
                                                                        t->d_grad);
    TR_TRY(hipGetLastError());
    return LBMLETKF_OK;
}

int auto_w_norm(Trainer* t, double w_norm, float& out) {
    if (w_norm > 0) {
        out = float(w_norm);
        return LBMLETKF_OK;
    }
    k_target_power<<<256, 256, 0, t->stream>>>(t->d_samples, t->n, t->d_pow);
    TR_TRY(hipGetLastError());
    std::vector<double> part(256);
    TR_TRY(hipMemcpyAsync(part.data(), t->d_pow, sizeof(double) * 256, hipMemcpyDeviceToHost, t->stream));
    TR_TRY(hipStreamSynchronize(t->stream));
    double s = 0;
    for (double x : part) s += x;
    const double mean = s / double(t->n);
    if (!(mean > 0)) return report_error(nullptr, LBMLETKF_ERR_NUMERIC, "the training targets are all zero");
    out = float(1.0 / mean);
    return LBMLETKF_OK;
}

}  // namespace
}  // namespace lbmletkf

using namespace lbmletkf;

extern "C" {

int lbmletkf_closure_init_params(unsigned long long seed, float* params) {
    if (!params) return report_error(nullptr, LBMLETKF_ERR_ARG, "params is NULL");
    // uniform +-sqrt(6/(fan_in+fan_out)) from Philox, biases 0; the output layer is
    // scaled by 0.1 and b3 = softplus^-1(1), so the untrained network gives nu_t ~ nu_scale
    auto fill = [&](int off, int n, double lim, uint32_t tag) {
        for (int k = 0; k < n; k += 4) {
            u32x4 ctr{{uint32_t(k), tag, 0u, 0x434c4f53u}};
            const u32x4 r = philox4x32_10(ctr, uint32_t(seed), uint32_t(seed >> 32));
            for (int q = 0; q < 4 && k + q < n; ++q)
                params[off + k + q] = float(lim * (2.0 * u32_to_open01(r.v[q]) - 1.0));
        }
    };
    for (int k = 0; k < CL_NPARAM; ++k) params[k] = 0.f;
    fill(CL_W1, CL_H * CL_IN, std::sqrt(6.0 / (CL_IN + CL_H)), 1);
    fill(CL_W2, CL_H * CL_H, std::sqrt(6.0 / (CL_H + CL_H)), 2);
    fill(CL_W3, CL_H, 0.1 * std::sqrt(6.0 / (CL_H + 1)), 3);
    params[CL_B3] = float(std::log(std::expm1(1.0)));
    return LBMLETKF_OK;
}

int lbmletkf_trainer_create(const lbmletkf_config* coarse_cfg, long long capacity, lbmletkf_trainer** out) {
    if (!coarse_cfg || !out) return report_error(nullptr, LBMLETKF_ERR_ARG, "NULL argument");
    *out = nullptr;
    if (capacity < 1) return report_error(nullptr, LBMLETKF_ERR_ARG, "capacity must be positive");
    if (!(coarse_cfg->closure_strain_scale > 0 && coarse_cfg->closure_nu_scale > 0))
        return report_error(nullptr, LBMLETKF_ERR_ARG, "closure_strain_scale and closure_nu_scale must be positive");
    Trainer* t = new Trainer();
    t->cfg = *coarse_cfg;
    t->cap = capacity;
    auto bail = [&](hipError_t e, const char* what) {
        delete t;
        return report_error(nullptr, LBMLETKF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipStreamCreate(&t->stream)) != hipSuccess) return bail(e, "hipStreamCreate");
    if ((e = hipMalloc(&t->d_samples, sizeof(float) * CL_SAMPLE * (size_t)capacity)) != hipSuccess) return bail(e, "hipMalloc samples");
    if ((e = hipMalloc(&t->d_params, sizeof(float) * CL_NPARAM)) != hipSuccess) return bail(e, "hipMalloc params");
    if ((e = hipMalloc(&t->d_m, sizeof(float) * CL_NPARAM)) != hipSuccess) return bail(e, "hipMalloc m");
    if ((e = hipMalloc(&t->d_v, sizeof(float) * CL_NPARAM)) != hipSuccess) return bail(e, "hipMalloc v");
    if ((e = hipMalloc(&t->d_grad, sizeof(float) * (CL_NPARAM + 1))) != hipSuccess) return bail(e, "hipMalloc grad");
    if ((e = hipMalloc(&t->d_pow, sizeof(double) * 256)) != hipSuccess) return bail(e, "hipMalloc pow");
    std::vector<float> p(CL_NPARAM);
    lbmletkf_closure_init_params(1, p.data());
    if ((e = hipMemcpy(t->d_params, p.data(), sizeof(float) * CL_NPARAM, hipMemcpyHostToDevice)) != hipSuccess)
        return bail(e, "hipMemcpy params");
    if ((e = hipMemset(t->d_m, 0, sizeof(float) * CL_NPARAM)) != hipSuccess) return bail(e, "hipMemset m");
    if ((e = hipMemset(t->d_v, 0, sizeof(float) * CL_NPARAM)) != hipSuccess) return bail(e, "hipMemset v");
    *out = reinterpret_cast<lbmletkf_trainer*>(t);
    return LBMLETKF_OK;
}

int lbmletkf_trainer_destroy(lbmletkf_trainer* tr) {
    Trainer* t = reinterpret_cast<Trainer*>(tr);
    if (!t) return LBMLETKF_OK;
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    delete t;
    return LBMLETKF_OK;
}

int lbmletkf_trainer_add_samples(lbmletkf_trainer* tr, lbmletkf_ctx* fine_ctx, int member, int factor) {
    Trainer* t = reinterpret_cast<Trainer*>(tr);
    Ctx* fine = reinterpret_cast<Ctx*>(fine_ctx);
    if (!t || !fine) return report_error(nullptr, LBMLETKF_ERR_ARG, "NULL argument");
    long long nc;
    int nx, ny;
    if (int st = coarse_fields(t, fine, member, factor, nc, nx, ny)) return st;
    if (t->n + nc > t->cap) return report_error(nullptr, LBMLETKF_ERR_ARG, "trainer capacity exceeded");
    const float inv_2dx = float(1.0 / (2.0 * t->cfg.dx));
    k_closure_samples<<<unsigned((nc + 255) / 256), 256, 0, fine->stream>>>(
        t->d_fields, nx, ny, inv_2dx, float(1.0 / t->cfg.closure_strain_scale), t->d_samples + t->n * CL_SAMPLE);
    TR_TRY(hipGetLastError());
    TR_TRY(hipStreamSynchronize(fine->stream));
    t->n += nc;
    return LBMLETKF_OK;
}

int lbmletkf_trainer_set_samples(lbmletkf_trainer* tr, const float* host, long long n) {
    Trainer* t = reinterpret_cast<Trainer*>(tr);
    if (!t || (!host && n > 0)) return report_error(nullptr, LBMLETKF_ERR_ARG, "NULL argument");
    if (n < 0 || n > t->cap) return report_error(nullptr, LBMLETKF_ERR_ARG, "sample count out of range");
    if (n) TR_TRY(hipMemcpy(t->d_samples, host, sizeof(float) * CL_SAMPLE * (size_t)n, hipMemcpyHostToDevice));
    t->n = n;
    return LBMLETKF_OK;
}

int lbmletkf_trainer_count(lbmletkf_trainer* tr, long long* n) {
    Trainer* t = reinterpret_cast<Trainer*>(tr);
    if (!t || !n) return report_error(nullptr, LBMLETKF_ERR_ARG, "NULL argument");
    *n = t->n;
    return LBMLETKF_OK;
}

int lbmletkf_trainer_get_samples(lbmletkf_trainer* tr, long long first, long long n, float* host) {
    Trainer* t = reinterpret_cast<Trainer*>(tr);
    if (!t || !host) return report_error(nullptr, LBMLETKF_ERR_ARG, "NULL argument");
    if (first < 0 || n < 0 || first + n > t->n) return report_error(nullptr, LBMLETKF_ERR_ARG, "sample range out of bounds");
    TR_TRY(hipStreamSynchronize(t->stream));
    if (n) TR_TRY(hipMemcpy(host, t->d_samples + first * CL_SAMPLE, sizeof(float) * CL_SAMPLE * (size_t)n, hipMemcpyDeviceToHost));
    return LBMLETKF_OK;
}

int lbmletkf_trainer_set_params(lbmletkf_trainer* tr, const float* params) {
    Trainer* t = reinterpret_cast<Trainer*>(tr);
    if (!t || !params) return report_error(nullptr, LBMLETKF_ERR_ARG, "NULL argument");
    TR_TRY(hipStreamSynchronize(t->stream));
    TR_TRY(hipMemcpy(t->d_params, params, sizeof(float) * CL_NPARAM, hipMemcpyHostToDevice));
    TR_TRY(hipMemset(t->d_m, 0, sizeof(float) * CL_NPARAM));
    TR_TRY(hipMemset(t->d_v, 0, sizeof(float) * CL_NPARAM));
    t->adam_t = 0;
    return LBMLETKF_OK;
}

int lbmletkf_trainer_get_params(lbmletkf_trainer* tr, float* params) {
    Trainer* t = reinterpret_cast<Trainer*>(tr);
    if (!t || !params) return report_error(nullptr, LBMLETKF_ERR_ARG, "NULL argument");
    TR_TRY(hipStreamSynchronize(t->stream));
    TR_TRY(hipMemcpy(params, t->d_params, sizeof(float) * CL_NPARAM, hipMemcpyDeviceToHost));
    return LBMLETKF_OK;
}

int lbmletkf_trainer_loss_grad(lbmletkf_trainer* tr, long long first, long long n, double w_norm, double* loss,
                               float* grad) {
    Trainer* t = reinterpret_cast<Trainer*>(tr);
    if (!t) return report_error(nullptr, LBMLETKF_ERR_ARG, "NULL argument");
    if (first < 0 || n < 1 || first + n > t->n) return report_error(nullptr, LBMLETKF_ERR_ARG, "sample range out of bounds");
    float wn;
    if (int st = auto_w_norm(t, w_norm, wn)) return st;
    if (int st = grad_launch(t, first, n, false, 0, 0, wn)) return st;
    std::vector<float> g(CL_NPARAM + 1);
    TR_TRY(hipMemcpyAsync(g.data(), t->d_grad, sizeof(float) * (CL_NPARAM + 1), hipMemcpyDeviceToHost, t->stream));
    TR_TRY(hipStreamSynchronize(t->stream));
    if (grad) std::memcpy(grad, g.data(), sizeof(float) * CL_NPARAM);
    if (loss) *loss = g[CL_NPARAM];
    return LBMLETKF_OK;
}

int lbmletkf_trainer_train(lbmletkf_trainer* tr, int batch, int steps, double lr, double w_norm,
                           float* loss_history) {
    Trainer* t = reinterpret_cast<Trainer*>(tr);
    if (!t) return report_error(nullptr, LBMLETKF_ERR_ARG, "NULL argument");
    if (batch < 1 || steps < 0 || !(lr > 0) || t->n < 1) return report_error(nullptr, LBMLETKF_ERR_ARG, "bad training arguments");
    if (batch > t->n) batch = int(t->n);
    float wn;
    if (int st = auto_w_norm(t, w_norm, wn)) return st;
    // minibatches walk a permutation k -> (pa k + pc) mod n, a new multiplier each epoch
    static const unsigned long long primes[] = {2654435761ull, 40503ull, 2246822519ull, 3266489917ull,
                                                668265263ull, 374761393ull, 1000003ull, 998244353ull};
    const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
    const long long per_epoch = std::max<long long>(1, t->n / batch);
    float* d_hist = nullptr;
    if (loss_history && steps > 0) TR_TRY(hipMalloc(&d_hist, sizeof(float) * steps));
    int st = LBMLETKF_OK;
    for (int s = 0; s < steps && st == LBMLETKF_OK; ++s) {
        const long long epoch = s / per_epoch, pos = (s % per_epoch) * batch;
        unsigned long long pa = primes[epoch % 8];
        while (std::gcd(pa, (unsigned long long)t->n) != 1) pa += 2;
        const unsigned long long pc = (unsigned long long)(epoch * 7919 + 17) % (unsigned long long)t->n;
        st = grad_launch(t, pos, batch, true, pa % (unsigned long long)t->n, pc, wn);
        if (st != LBMLETKF_OK) break;
        ++t->adam_t;
        const float c1 = float(1.0 / (1.0 - std::pow(b1, double(t->adam_t))));
        const float c2 = float(1.0 / (1.0 - std::pow(b2, double(t->adam_t))));
        k_adam<<<(CL_NPARAM + 255) / 256, 256, 0, t->stream>>>(t->d_params, t->d_grad, t->d_m, t->d_v, float(lr),
                                                               float(b1), float(b2), float(eps), c1, c2);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess && d_hist)
            e = hipMemcpyAsync(d_hist + s, t->d_grad + CL_NPARAM, sizeof(float), hipMemcpyDeviceToDevice, t->stream);
        if (e != hipSuccess) st = report_error(nullptr, LBMLETKF_ERR_HIP, std::string("train: ") + hipGetErrorString(e));
    }
    hipError_t e = hipStreamSynchronize(t->stream);
    if (e == hipSuccess && d_hist && st == LBMLETKF_OK)
        e = hipMemcpy(loss_history, d_hist, sizeof(float) * steps, hipMemcpyDeviceToHost);
    if (d_hist) (void)hipFree(d_hist);
    if (st != LBMLETKF_OK) return st;
    if (e != hipSuccess) return report_error(nullptr, LBMLETKF_ERR_HIP, std::string("train: ") + hipGetErrorString(e));
    return LBMLETKF_OK;
}

int lbmletkf_filtered_macro(lbmletkf_ctx* fine_ctx, int member, int factor, float* rho, float* u, float* v) {
    Ctx* fine = reinterpret_cast<Ctx*>(fine_ctx);
    if (!fine || !rho || !u || !v) return report_error(nullptr, LBMLETKF_ERR_ARG, "NULL argument");
    Trainer t;  // scratch only
    long long nc;
    int nx, ny;
    if (int st = coarse_fields(&t, fine, member, factor, nc, nx, ny)) return st;
    TR_TRY(hipStreamSynchronize(fine->stream));
    TR_TRY(hipMemcpy(rho, t.d_fields, sizeof(float) * nc, hipMemcpyDeviceToHost));
    TR_TRY(hipMemcpy(u, t.d_fields + nc, sizeof(float) * nc, hipMemcpyDeviceToHost));
    TR_TRY(hipMemcpy(v, t.d_fields + 2 * nc, sizeof(float) * nc, hipMemcpyDeviceToHost));
    return LBMLETKF_OK;
}

int lbmletkf_set_closure(lbmletkf_ctx* ctx, const float* params) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !params) return report_error(c, LBMLETKF_ERR_ARG, "NULL argument");
    if (c->cfg.closure_kind != LBMLETKF_CLOSURE_MLP)
        return report_error(c, LBMLETKF_ERR_STATE, "the context was not created with closure_kind = MLP");
    ClosureWeights w;
    closure_weights_host(params, c->cfg, w);
    if (!c->d_closure) {
        hipError_t e = hipMalloc(&c->d_closure, sizeof(ClosureWeights));
        if (e != hipSuccess) return report_error(c, LBMLETKF_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    hipError_t e = hipMemcpyAsync(c->d_closure, &w, sizeof(w), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return report_error(c, LBMLETKF_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    c->closure_ready = true;
    return LBMLETKF_OK;
}

int lbmletkf_closure_nut(lbmletkf_ctx* ctx, int member, float* nut) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!c || !nut) return report_error(c, LBMLETKF_ERR_ARG, "NULL argument");
    if (member < 0 || member >= c->M) return report_error(c, LBMLETKF_ERR_ARG, "member index out of range");
    if (!c->closure_ready) return report_error(c, LBMLETKF_ERR_STATE, "no closure loaded: call lbmletkf_set_closure first");
    float* d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(float) * c->nodes);
    if (e != hipSuccess) return report_error(c, LBMLETKF_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    // the populations go to the idle half of the double buffer, which the next step overwrites
    const int tiles = (c->nx / 64) * (c->ny / 16);
    e = launch_lbm_mlp(*c, c->f[c->cur], c->f[1 - c->cur], member * tiles, tiles, d);
    if (e == hipSuccess) e = hipMemcpyAsync(nut, d, sizeof(float) * c->nodes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return report_error(c, LBMLETKF_ERR_HIP, std::string("closure_nut: ") + hipGetErrorString(e));
    return LBMLETKF_OK;
}

}  // extern "C"
