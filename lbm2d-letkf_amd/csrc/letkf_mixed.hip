// Mixed-precision fused LETKF analysis for gfx950: one wavefront per gridpoint
// does the whole local analysis of src/util/letkf_solver.h:235-349 without a
// round trip through HBM:
//
//  1. gather: the rows of dY (and the innovation) of the gridpoint's local
//     observations, scaled by sqrt(R^-1 o localization), straight from L2;
//  2. Gram on the matrix cores: C = (M-1)/beta I + G^T G and b = G^T (sqrt(Rl^-1) dyo)
//     with v_mfma_f32_16x16x32_bf16 (FP32 accumulation); PASSES = 3 splits each
//     operand into bf16 hi + lo parts (hi*hi + hi*lo + lo*hi);
//  3. cyclic Jacobi in FP32: A in LDS (16.6 KB at N = 64), the eigenvector
//     matrix V in registers with one column per lane, so a rotation of the
//     column pair (p, p^m) is a lane-pair exchange (__shfl_xor by m);
//  4. Jacobi sweeps in FP64 on the converged matrix (packed upper triangle in
//     the same LDS, V widened to FP64 in registers);
//  5. inverse-square-root assembly in FP64 without forming W:
//       Xa = xbar + dX (W + w 1^T) = xbar + sqrt(M-1) (dX V) D^-1/2 V^T + (dX w) 1^T,
//       w = V D^-1 V^T b,
//     with V transposed through LDS a quarter of its columns at a time.
#include "context.hpp"

namespace lbmletkf {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct MixedParams {
    int M, nodes, n_vars;
    long long member_stride;
    float diag;             // (M-1)/beta
    float coef[3];          // R^-1 per observed variable
    double sqm1;            // sqrt(M-1)
    float tol32_2;          // FP32 phase: stop when off(A)^2 <= tol32^2 ||A||^2
    double tol64_2;         // FP64 phase
    int max32, max64;
    int dump_g0, dump_n;
    unsigned long long* sweeps32;
    unsigned long long* sweeps64;
};

__device__ inline double wave_sum_d(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    return x;
}

// pair k of round m (XOR schedule): p has bit hb(m) clear, q = p ^ m
__device__ inline void xpair(int k, int m, int& p, int& q) {
    const int hb = 31 - __clz(m);
    p = ((k >> hb) << (hb + 1)) | (k & ((1 << hb) - 1));
    q = p ^ m;
}
// index of the pair whose low member is p (bit hb(m) of p clear) in round m
__device__ inline int pair_of(int p, int m) {
    const int c = p;
    const int hb = 31 - __clz(m);
    return ((c >> (hb + 1)) << hb) | (c & ((1 << hb) - 1));
}

// FP32 symmetric matrix stored full (row stride N+1), writes mirrored
template <int N>
struct FullF32 {
    float* a;
    __device__ float ld(int i, int j) const { return a[i * (N + 1) + j]; }
    __device__ void st_pair(int i, int j, float x) const {
        a[i * (N + 1) + j] = x;
        a[j * (N + 1) + i] = x;
    }
    __device__ void st(int i, int j, float x) const { a[i * (N + 1) + j] = x; }
};

// FP64 symmetric matrix stored as packed upper triangle
template <int N>
struct TriF64 {
    double* a;
    __device__ static int idx(int i, int j) {
        const int lo = min(i, j), hi = max(i, j);
        return lo * N - (lo * (lo - 1)) / 2 + (hi - lo);
    }
    __device__ double ld(int i, int j) const { return a[idx(i, j)]; }
    __device__ void st_pair(int i, int j, double x) const { a[idx(i, j)] = x; }
    __device__ void st(int i, int j, double x) const { a[idx(i, j)] = x; }
};

template <int N>
struct Blocks {
    static constexpr int NP = N / 2;
    static constexpr int NBLK = NP * (NP + 1) / 2;
    static constexpr int PER_LANE = (NBLK + 63) / 64;
};

// One sweep of N-1 XOR rounds.  v[] is the lane's column of V.
template <int N, class T, class Store>
__device__ void jacobi_sweep(const Store& A, T* rc, T* rs, T* rt, T (&v)[N], const int (&bI)[Blocks<N>::PER_LANE],
                             const int (&bJ)[Blocks<N>::PER_LANE], int lane) {
    constexpr int NP = N / 2;
    for (int m = 1; m < N; ++m) {
        if (lane < NP) {
            int p, q;
            xpair(lane, m, p, q);
            const T app = A.ld(p, p), aqq = A.ld(q, q), apq = A.ld(p, q);
            T c = T(1), s = T(0), t = T(0);
            if (apq != T(0)) {
                const T theta = (aqq - app) / (T(2) * apq);
                t = (theta >= T(0) ? T(1) : T(-1)) / (fabs(theta) + sqrt(theta * theta + T(1)));
                c = T(1) / sqrt(t * t + T(1));
                s = t * c;
            }
            rc[lane] = c;
            rs[lane] = s;
            rt[lane] = t;
        }
        __syncthreads();
        // two-sided update of the 2x2 blocks (I <= J) owned by this lane
#pragma unroll
        for (int u = 0; u < Blocks<N>::PER_LANE; ++u) {
            const int I = bI[u], J = bJ[u];
            if (I < 0) continue;
            int p, q;
            xpair(I, m, p, q);
            if (I == J) {
                const T apq = A.ld(p, q), t = rt[I];
                A.st(p, p, A.ld(p, p) - t * apq);
                A.st(q, q, A.ld(q, q) + t * apq);
                A.st_pair(p, q, T(0));
            } else {
                int r, s2;
                xpair(J, m, r, s2);
                const T cI = rc[I], sI = rs[I], cJ = rc[J], sJ = rs[J];
                const T b11 = A.ld(p, r), b12 = A.ld(p, s2), b21 = A.ld(q, r), b22 = A.ld(q, s2);
                const T x11 = cI * b11 - sI * b21, x12 = cI * b12 - sI * b22;
                const T x21 = sI * b11 + cI * b21, x22 = sI * b12 + cI * b22;
                A.st_pair(p, r, cJ * x11 - sJ * x12);
                A.st_pair(p, s2, sJ * x11 + cJ * x12);
                A.st_pair(q, r, cJ * x21 - sJ * x22);
                A.st_pair(q, s2, sJ * x21 + cJ * x22);
            }
        }
        // V <- V J: column `lane` mixes with column lane^m (both below N, as m < N)
        if (lane < N) {
            const bool low = (lane & (1 << (31 - __clz(m)))) == 0;
            // the pair is indexed by its low member p; the high member is p ^ m,
            // which also differs from p in the bits of m below its top bit
            const int k = pair_of(low ? lane : (lane ^ m), m);
            const T c = rc[k], s = low ? -rs[k] : rs[k];
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const T other = __shfl_xor(v[i], m);
                v[i] = c * v[i] + s * other;
            }
        }
        __syncthreads();
    }
}

// off-diagonal and total squared Frobenius norms
template <int N, class T, class Store>
__device__ void frob(const Store& A, int lane, double& off, double& tot) {
    off = 0.0;
    tot = 0.0;
    for (int idx = lane; idx < N * N; idx += 64) {
        const int i = idx / N, j = idx % N;
        const double x = double(A.ld(i, j));
        tot += x * x;
        if (i != j) off += x * x;
    }
    off = wave_sum_d(off);
    tot = wave_sum_d(tot);
}

template <int N>
struct MixedLds {
    static constexpr int A32_BYTES = N * (N + 1) * 4;
    static constexpr int TRI_BYTES = N * (N + 1) / 2 * 8;
    static constexpr int TQ = 16;  // columns of V transposed at a time
    static constexpr int T_STRIDE = TQ + 1;
    static constexpr int ASM_BYTES = N * T_STRIDE * 8 + Q * N * 8 + Q * N * 4 + N * 8;
    static constexpr int MAIN_BYTES =
        (A32_BYTES > TRI_BYTES ? (A32_BYTES > ASM_BYTES ? A32_BYTES : ASM_BYTES)
                               : (TRI_BYTES > ASM_BYTES ? TRI_BYTES : ASM_BYTES));
};

template <int N, int PASSES>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_letkf_mixed(
    const int* __restrict__ loc_ptr, const int* __restrict__ loc_site, const float* __restrict__ loc_w,
    const float* __restrict__ dy, const float* __restrict__ dyo, float* __restrict__ f, MixedParams mp,
    double* __restrict__ dump, int* __restrict__ status) {
    constexpr int NT = N / 16;              // member tiles of 16
    constexpr int NTILE = NT * (NT + 1) / 2; // upper-triangular tiles of C
    constexpr int NP = N / 2;
    __shared__ __align__(16) unsigned char smem[MixedLds<N>::MAIN_BYTES];
    __shared__ float rc32[NP], rs32[NP], rt32[NP];
    __shared__ double rc64[NP], rs64[NP], rt64[NP];
    __shared__ double bvec[N];
    __shared__ double xbar[Q];

    const int g = blockIdx.x;
    const int lane = threadIdx.x;
    const int M = mp.M;
    const int nv = mp.n_vars;

    // ---- 1+2. gather and Gram on MFMA -------------------------------------------
    f32x4 acc[NTILE + NT];
#pragma unroll
    for (int a = 0; a < NTILE + NT; ++a) acc[a] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int e0 = loc_ptr[g], e1 = loc_ptr[g + 1];
    const int P = (e1 - e0) * nv;
    const int col = lane & 15;   // member within a tile (A rows / B columns)
    const int kq = lane >> 4;    // which 8 of the 32 k-values this lane holds
    for (int base = 0; base < P; base += 32) {
        float x[NT][8], inn[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int jj = base + 8 * kq + t;
#pragma unroll
            for (int I = 0; I < NT; ++I) x[I][t] = 0.f;
            inn[t] = 0.f;
            if (jj < P) {
                const int e = e0 + jj / nv, va = jj % nv;
                const long long o = (long long)loc_site[e] * nv + va;
                const float sw = sqrtf(loc_w[e] * mp.coef[va]);
#pragma unroll
                for (int I = 0; I < NT; ++I) {
                    const int mm = 16 * I + col;
                    x[I][t] = (mm < M) ? dy[o * M + mm] * sw : 0.f;
                }
                inn[t] = (col == 0) ? dyo[o] * sw : 0.f;
            }
        }
        bf16x8 hi[NT + 1], lo[NT + 1];
#pragma unroll
        for (int I = 0; I <= NT; ++I) {
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const float xv = (I < NT) ? x[I][t] : inn[t];
                const __bf16 h = (__bf16)xv;
                hi[I][t] = h;
                lo[I][t] = (__bf16)(xv - (float)h);
            }
        }
        int tix = 0;
#pragma unroll
        for (int I = 0; I < NT; ++I) {
#pragma unroll
            for (int J = I; J < NT; ++J, ++tix) {
                acc[tix] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi[I], hi[J], acc[tix], 0, 0, 0);
                if (PASSES == 3) {
                    acc[tix] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi[I], lo[J], acc[tix], 0, 0, 0);
                    acc[tix] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lo[I], hi[J], acc[tix], 0, 0, 0);
                }
            }
        }
#pragma unroll
        for (int I = 0; I < NT; ++I) {
            acc[NTILE + I] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi[I], hi[NT], acc[NTILE + I], 0, 0, 0);
            if (PASSES == 3) {
                acc[NTILE + I] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi[I], lo[NT], acc[NTILE + I], 0, 0, 0);
                acc[NTILE + I] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lo[I], hi[NT], acc[NTILE + I], 0, 0, 0);
            }
        }
    }
    // result tile (I,J): lane holds rows 16I + 4*kq + r, column 16J + col
    FullF32<N> A32{reinterpret_cast<float*>(smem)};
    {
        int tix = 0;
#pragma unroll
        for (int I = 0; I < NT; ++I) {
#pragma unroll
            for (int J = I; J < NT; ++J, ++tix) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * I + 4 * kq + r, cc = 16 * J + col;
                    float val = acc[tix][r];
                    if (row == cc) val += (row < M) ? mp.diag : 1.f;
                    A32.st(row, cc, val);
                    if (I != J) A32.st(cc, row, val);
                }
            }
        }
#pragma unroll
        for (int I = 0; I < NT; ++I)
            if (col == 0) {
#pragma unroll
                for (int r = 0; r < 4; ++r) bvec[16 * I + 4 * kq + r] = double(acc[NTILE + I][r]);
            }
    }
    __syncthreads();

    // ---- 3. FP32 Jacobi -----------------------------------------------------------
    int bI[Blocks<N>::PER_LANE], bJ[Blocks<N>::PER_LANE];
#pragma unroll
    for (int u = 0; u < Blocks<N>::PER_LANE; ++u) {
        const int b = lane + 64 * u;
        bI[u] = -1;
        bJ[u] = -1;
        if (b < Blocks<N>::NBLK) {
            int I = 0, rem = b;
            while (rem >= NP - I) {
                rem -= NP - I;
                ++I;
            }
            bI[u] = I;
            bJ[u] = I + rem;
        }
    }
    float v[N];
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = (i == lane) ? 1.f : 0.f;
    int sw32 = 0;
    double prev_off = 0.0;
    for (; sw32 < mp.max32; ++sw32) {
        double off, tot;
        frob<N, float>(A32, lane, off, tot);
        // stop at the tolerance, or once FP32 rounding stalls the (by then
        // quadratic) decrease below a relative off-norm of 1e-4
        if (off <= double(mp.tol32_2) * tot || (sw32 > 0 && off <= 1e-8 * tot && off > 0.25 * prev_off)) break;
        prev_off = off;
        jacobi_sweep<N, float>(A32, rc32, rs32, rt32, v, bI, bJ, lane);
    }

    // ---- 4. FP64 sweeps on the packed triangle ------------------------------------
    constexpr int TRI = N * (N + 1) / 2;
    constexpr int TRI_PER_LANE = (TRI + 63) / 64;
    float tmp[TRI_PER_LANE];
#pragma unroll
    for (int u = 0; u < TRI_PER_LANE; ++u) {
        const int t = lane + 64 * u;
        tmp[u] = 0.f;
        if (t < TRI) {
            // row i of packed index t: start(i) = i N - i (i-1)/2 <= t < start(i+1)
            const float b2 = float(2 * N + 1);
            int i = int((b2 - sqrtf(b2 * b2 - 8.f * float(t))) * 0.5f);
            i = max(0, min(i, N - 1));
            while (i > 0 && i * N - (i * (i - 1)) / 2 > t) --i;
            while (i < N - 1 && (i + 1) * N - ((i + 1) * i) / 2 <= t) ++i;
            const int j = i + (t - (i * N - (i * (i - 1)) / 2));
            tmp[u] = A32.ld(i, j);
        }
    }
    __syncthreads();
    TriF64<N> A64{reinterpret_cast<double*>(smem)};
#pragma unroll
    for (int u = 0; u < TRI_PER_LANE; ++u) {
        const int t = lane + 64 * u;
        if (t < TRI) A64.a[t] = double(tmp[u]);
    }
    double vd[N];
#pragma unroll
    for (int i = 0; i < N; ++i) vd[i] = double(v[i]);
    __syncthreads();
    int sw64 = 0;
    for (; sw64 < mp.max64; ++sw64) {
        double off, tot;
        frob<N, double>(A64, lane, off, tot);
        if (off <= mp.tol64_2 * tot) break;
        jacobi_sweep<N, double>(A64, rc64, rs64, rt64, vd, bI, bJ, lane);
    }
    if (lane == 0) {
        if (mp.sweeps32) atomicAdd(mp.sweeps32, (unsigned long long)sw32);
        if (mp.sweeps64) atomicAdd(mp.sweeps64, (unsigned long long)sw64);
    }

    // ---- 5. assembly in FP64 --------------------------------------------------------
    // column layout (lane = eigen-index i): d_i, h_i = (V^T b)_i / d_i, s_i = sqrt(M-1)/sqrt(d_i)
    double d_i = 1.0, h_i = 0.0, s_i = 0.0;
    if (lane < N) {
        d_i = A64.ld(lane, lane);
        double gs = 0.0;
#pragma unroll
        for (int r = 0; r < N; ++r) gs += vd[r] * bvec[r];
        h_i = gs / d_i;
        s_i = mp.sqm1 / sqrt(d_i);
    }
    const bool dumping = (g >= mp.dump_g0) && (g < mp.dump_g0 + mp.dump_n);
    if (dumping && lane < M) {
        double* dg = dump + (long long)(g - mp.dump_g0) * (M * M + 2 * M);
#pragma unroll
        for (int k = 0; k < N; ++k)  // static indices keep vd[] in registers
            if (k < M) dg[k * M + lane] = vd[k];  // V[k][lane]
        dg[M * M + lane] = d_i;
        dg[M * M + M + lane] = bvec[lane];
    }
    __syncthreads();  // A64 is dead from here on; its LDS is reused
    constexpr int TQ = MixedLds<N>::TQ, TS = MixedLds<N>::T_STRIDE;
    double* Tq = reinterpret_cast<double*>(smem);       // [N][TS] chunk of V columns
    double* Zs = Tq + N * TS;                            // [Q][N] s_i (dX V)[q][i]
    float* dXs = reinterpret_cast<float*>(Zs + Q * N);   // [Q][N]
    double* hv = reinterpret_cast<double*>(dXs + Q * N); // [N]
    // members: row layout, lane = member l
    float xk[Q];
    const int member = lane < M ? lane : 0;
    const long long base = (long long)member * mp.member_stride + g;
#pragma unroll
    for (int q = 0; q < Q; ++q) xk[q] = 0.f;
    if (lane < M) {
#pragma unroll
        for (int q = 0; q < Q; ++q) xk[q] = f[base + (long long)q * mp.nodes];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const double s = wave_sum_d(double(xk[q]));
        if (lane == 0) xbar[q] = s / double(M);
    }
    if (lane < N) hv[lane] = h_i;
    __syncthreads();
    float dxk[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        dxk[q] = (lane < M) ? float(double(xk[q]) - xbar[q]) : 0.f;
        if (lane < N) dXs[q * N + lane] = dxk[q];
    }
    __syncthreads();
    // Z'[q][i] = s_i sum_r dX[q][r] V[r][i]   (column layout)
    if (lane < N) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            double z = 0.0;
#pragma unroll
            for (int r = 0; r < N; ++r) z += double(dXs[q * N + r]) * vd[r];
            Zs[q * N + lane] = s_i * z;
        }
    }
    // dXa[q][l] = sum_i Z'[q][i] V[l][i],  w_l = sum_i V[l][i] h_i   (row layout)
    double dxa[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) dxa[q] = 0.0;
    double wl = 0.0;
#pragma unroll
    for (int c0 = 0; c0 < N; c0 += TQ) {
        __syncthreads();
        if (lane >= c0 && lane < c0 + TQ) {
#pragma unroll
            for (int r = 0; r < N; ++r) Tq[r * TS + (lane - c0)] = vd[r];
        }
        __syncthreads();
        if (lane < N) {
#pragma unroll
            for (int ii = 0; ii < TQ; ++ii) {
                const double vli = Tq[lane * TS + ii];  // V[lane][c0+ii]
                wl += vli * hv[c0 + ii];
#pragma unroll
                for (int q = 0; q < Q; ++q) dxa[q] += Zs[q * N + c0 + ii] * vli;
            }
        }
    }
    // Xa = xbar + dXa + (dX w) 1^T
    bool bad = false;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const double dxw = wave_sum_d((lane < M) ? double(dxk[q]) * wl : 0.0);
        if (lane < M) {
            const float xa = float(xbar[q] + dxa[q] + dxw);
            bad |= !isfinite(xa);
            f[base + (long long)q * mp.nodes] = xa;
        }
    }
    if (bad) atomicOr(status, 1);
}

}  // namespace

hipError_t launch_letkf_mixed(Ctx& c, float* f, unsigned long long* sweeps32, unsigned long long* sweeps64,
                              int passes) {
    MixedParams mp;
    mp.M = c.M;
    mp.nodes = int(c.nodes);
    mp.n_vars = c.obs.n_vars;
    mp.member_stride = c.lp.member_stride;
    mp.diag = float(double(c.M - 1) / c.cfg.inflation);
    for (int a = 0; a < 3; ++a) mp.coef[a] = c.obs.inv_sigma2[a];
    mp.sqm1 = sqrt(double(c.M - 1));
    mp.tol32_2 = 1e-6f * 1e-6f;
    mp.tol64_2 = c.cfg.jacobi_tol * c.cfg.jacobi_tol;
    mp.max32 = 40;
    mp.max64 = 4;
    mp.dump_g0 = c.dump_g0;
    mp.dump_n = c.dump_n;
    mp.sweeps32 = sweeps32;
    mp.sweeps64 = sweeps64;
    const dim3 grid(unsigned(c.nodes)), block(64);
    int N = 16;
    while (N < c.M) N *= 2;
#define LAUNCH_MIXED(NN, PP)                                                                          \
    k_letkf_mixed<NN, PP><<<grid, block, 0, c.stream>>>(c.d_loc_ptr, c.d_loc_site, c.d_loc_w, c.d_dy, \
                                                         c.d_dyo, f, mp, c.d_dump, c.d_status)
    if (passes == 3) {
        switch (N) {
            case 16: LAUNCH_MIXED(16, 3); break;
            case 32: LAUNCH_MIXED(32, 3); break;
            case 64: LAUNCH_MIXED(64, 3); break;
            default: return hipErrorInvalidValue;
        }
    } else {
        switch (N) {
            case 16: LAUNCH_MIXED(16, 1); break;
            case 32: LAUNCH_MIXED(32, 1); break;
            case 64: LAUNCH_MIXED(64, 1); break;
            default: return hipErrorInvalidValue;
        }
    }
#undef LAUNCH_MIXED
    return hipGetLastError();
}

}  // namespace lbmletkf
