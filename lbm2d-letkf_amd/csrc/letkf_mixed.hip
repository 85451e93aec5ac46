// Mixed-precision LETKF analysis for gfx950: the local analysis of
// src/util/letkf_solver.h:235-349 at every gridpoint, one wavefront per gridpoint,
// in two kernels:
//
// k_mixed_solve (FP32, ~200 VGPRs, C in 16 KB of LDS; a chain of neighbouring
// gridpoints per wavefront, each warm-started from its predecessor's eigenvectors)
//  1. gather: the rows of dY (and the innovation) of the gridpoint's local
//     observations, scaled by sqrt(R^-1 o localization), straight from L2;
//  2. Gram on the matrix cores: C = (M-1)/beta I + G^T G and b = G^T (sqrt(Rl^-1) dyo)
//     with v_mfma_f32_16x16x32_bf16 (FP32 accumulation); PASSES = 3 splits each
//     operand into bf16 hi + lo parts (hi*hi + hi*lo + lo*hi);
//  3. cyclic one-sided (Hestenes) Jacobi in FP32 on U = C, register resident:
//     lane j owns column j; round m rotates the column pairs (p, p^m).  For most
//     rounds the two lanes of a pair each hold half the rows of both columns, so
//     a round is a half inner product per lane, one exchange of the partial sum
//     and one FMA per held element; between rounds the partner's rows move by DPP
//     lane permutations, with few crossbar exchanges.  No LDS memory and no
//     barrier inside a sweep.  As C is symmetric positive definite,
//     U -> C V = V D, so the eigenpairs are d_j = ||u_j||, v_j = u_j / d_j and V
//     is never accumulated.  U and b go to a workspace in HBM.
//
// k_mixed_finish (FP64, two wavefronts per gridpoint, each with half the rows)
//  4. the same sweeps in FP64 on the widened columns (partner half-column held in
//     registers, pair inner products summed across the two waves through LDS)
//     until the pairwise cosines fall below the eigensolver tolerance or the sweep
//     cap is reached;
//  5. inverse-square-root assembly in FP64 without forming W:
//       Xa = xbar + dX (W + w 1^T) = xbar + sqrt(M-1) (dX V) D^-1/2 V^T + (dX w) 1^T,
//       w = V D^-1 V^T b,
//     with V transposed through LDS a quarter of its columns at a time.
//
// The workspace round trip (4 N^2 bytes per gridpoint each way) costs ~0.3 ms at
// 256^2 x 64 and lets each kernel have the register budget of its own phase.
#include "context.hpp"

namespace lbmletkf {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct MixedParams {
    int M, nodes, n_vars;
    int n_obs;              // scalar observations (rows of dY)
    int nx, ny;             // grid (k_mixed_analysis walks its chains along columns)
    long long member_stride;
    float diag;             // (M-1)/beta
    float coef[3];          // R^-1 per observed variable
    double sqm1;            // sqrt(M-1)
    float check_tol;        // fused FP32 phase: stop when the columns' Gram (matrix cores) has
                            //   no pair cosine above this
    float tol32;            // FP32 phase: stop after a sweep whose largest pair cosine is below
    double tol64;           // FP64 phase: the same test, then at most max64 sweeps
    int max32, max64;
    int dump_g0, dump_n;
    int chain;              // k_mixed_solve: consecutive gridpoints per wavefront
    int write_state;        // k_mixed_analysis: 0 skips the state update (timing only)
    unsigned long long* sweeps32;
    unsigned long long* sweeps64;
};

// one analysis of a group launch
struct MixedTask {
    const int* loc_ptr;
    const int* loc_site;
    const float* loc_w;
    const float* dy;
    const float* dyo;
    float* ws_u;
    float* ws_b;
    float* f;
    double* dump;
    int* status;
    MixedParams mp;
    int N;
    int b0_solve, nb_solve;    // workgroup range of k_mixed_solve_group
    int b0_finish, nb_finish;  // workgroup range of k_mixed_finish_group
};
struct MixedGroup {
    int n;
    MixedTask t[MIXED_GROUP_MAX];
};

__device__ inline double wave_sum_d(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    return x;
}

template <class T>
__device__ inline T wave_max(T x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off));
    return x;
}

// Workgroups are dealt round-robin to the 8 XCDs; this order gives each XCD a
// contiguous range of gridpoints, so the strided member reads and writes of
// neighbouring gridpoints share cache lines in that XCD's L2.
__device__ inline int xcd_order(int b, int nb) { return (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b; }

// Rotation parameters: the FP32 sweeps use the hardware reciprocal, square root and
// reciprocal square root (about 1 ulp); the FP64 sweeps round correctly.  Any
// (c, t) accurate to a few ulp gives a rotation orthogonal to working precision.
[[maybe_unused]] __device__ inline float rot_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
[[maybe_unused]] __device__ inline double rot_rcp(double x) { return 1.0 / x; }
[[maybe_unused]] __device__ inline float rot_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
[[maybe_unused]] __device__ inline double rot_sqrt(double x) { return sqrt(x); }
[[maybe_unused]] __device__ inline float rot_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
[[maybe_unused]] __device__ inline double rot_rsq(double x) { return 1.0 / sqrt(x); }
// a + b rounded once, never fused with the product that formed a: the two lanes of
// a pair add the same two values in opposite order and must agree bit for bit
__device__ inline float add_uncontracted(float a, float b) {
#pragma clang fp contract(off)
    return a + b;
}


// Cross-lane reads of the sweeps.  A lane reads the value of lane ^ m: through the
// LDS crossbar (ds_bpermute, any m) or, for the masks the data-parallel primitives
// express within a row of 16 lanes, as a VALU move: quad_perm [1,0,3,2] (^1),
// [2,3,0,1] (^2), [3,2,1,0] (^3), row_ror:8 (^8).
struct XorBperm {
    int addr;  // (lane ^ m) * 4
    __device__ float operator()(float v) const {
        return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
    }
};
template <int CTRL>
struct XorDpp {
    __device__ float operator()(float v) const {
        return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
    }
};
constexpr int DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_XOR3 = 0x1B, DPP_XOR8 = 0x128;

// In-place FMAs whose first factor comes from another lane of the same quad, as
// the DPP source operand of v_fmac_f32: lane l reads lane l ^ D (D = 1, 2, 3), one
// instruction per element and no register for the moved value (the compiler folds
// DPP moves into untied VALU operations only, not into v_fmac).  Each block opens
// with the two wait states the hardware requires between a VALU write of a
// register and a DPP read of it; inside a block no instruction writes a register
// that a later one reads through DPP.  All lanes of the wave must be active.  An
// instruction reads all its operands before it writes, so d[i] += d[i](lane ^ D) b
// takes the partner's value from before the update.
#define LML_QP1 "quad_perm:[1,0,3,2]"
#define LML_QP2 "quad_perm:[2,3,0,1]"
#define LML_QP3 "quad_perm:[3,2,1,0]"
#define LML_F(QP, d, s0, s1) \
    "v_fmac_f32_dpp %[" #d "], %[" #s0 "], %[" #s1 "] " QP " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
#define LML_V8(n, v, r, K)                                                                                  \
    [n##0] r(v[K]), [n##1] r(v[K + 1]), [n##2] r(v[K + 2]), [n##3] r(v[K + 3]), [n##4] r(v[K + 4]),       \
        [n##5] r(v[K + 5]), [n##6] r(v[K + 6]), [n##7] r(v[K + 7])
#define LML_V16(n, v, r, K)                                                                                 \
    LML_V8(n, v, r, K), [n##8] r(v[K + 8]), [n##9] r(v[K + 9]), [n##10] r(v[K + 10]), [n##11] r(v[K + 11]), \
        [n##12] r(v[K + 12]), [n##13] r(v[K + 13]), [n##14] r(v[K + 14]), [n##15] r(v[K + 15])
#define LML_S4 [s0] "+v"(s[0]), [s1] "+v"(s[1]), [s2] "+v"(s[2]), [s3] "+v"(s[3])
#define LML_S4W [s0] "=v"(s[0]), [s1] "=v"(s[1]), [s2] "=v"(s[2]), [s3] "=v"(s[3])
// the lines of a block: d[i] += a[i](lane ^ D) * b, d[i] += d[i](lane ^ D) * b,
// s[i & 3] += a[i](lane ^ D) * c[i], s[i & 3] += a[i](lane ^ D) * a[i] (the sums s are
// set by the first block, not accumulated onto)
#define LML_AX(QP, i) LML_F(QP, d##i, a##i, b)
#define LML_SAX(QP, i) LML_F(QP, d##i, d##i, b)
#define LML_DOT(QP, i, j) LML_F(QP, s##j, a##i, c##i)
#define LML_SDOT(QP, i, j) LML_F(QP, s##j, a##i, a##i)
#define LML_M(QP, d, s0, s1) \
    "v_mul_f32_dpp %[" #d "], %[" #s0 "], %[" #s1 "] " QP " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
#define LML_DOT0(QP, i, j) LML_M(QP, s##j, a##i, c##i)
#define LML_SDOT0(QP, i, j) LML_M(QP, s##j, a##i, a##i)
#define LML_8(M, QP)                                                                                        \
    M(QP, 0) M(QP, 1) M(QP, 2) M(QP, 3) M(QP, 4) M(QP, 5) M(QP, 6) M(QP, 7)
#define LML_16(M, QP) LML_8(M, QP) M(QP, 8) M(QP, 9) M(QP, 10) M(QP, 11) M(QP, 12) M(QP, 13) M(QP, 14) M(QP, 15)
#define LML_D8(M, M0, QP)                                                                                   \
    M0(QP, 0, 0) M0(QP, 1, 1) M0(QP, 2, 2) M0(QP, 3, 3) M(QP, 4, 0) M(QP, 5, 1) M(QP, 6, 2) M(QP, 7, 3)
#define LML_D16(M, M0, QP)                                                                                  \
    LML_D8(M, M0, QP) M(QP, 8, 0) M(QP, 9, 1) M(QP, 10, 2) M(QP, 11, 3) M(QP, 12, 0) M(QP, 13, 1)           \
        M(QP, 14, 2) M(QP, 15, 3)
// blocks of W = 16 elements (8 for a last part of 8), each opening with the wait states
#define LML_AXPY(W, QP) \
    asm("s_nop 1\n\t" LML_##W(LML_AX, QP) : LML_V##W(d, d, "+v", K) : LML_V##W(a, a, "v", K), [b] "v"(b))
#define LML_SAXPY(W, QP) asm("s_nop 1\n\t" LML_##W(LML_SAX, QP) : LML_V##W(d, d, "+v", K) : [b] "v"(b))
// (the first block of a sum starts its four partial sums with products)
#define LML_DOTB(W, QP)                                                                                     \
    if constexpr (K == 0)                                                                                   \
        asm("s_nop 1\n\t" LML_D##W(LML_DOT, LML_DOT0, QP)                                                    \
            : LML_S4W                                                                                       \
            : LML_V##W(a, a, "v", K), LML_V##W(c, c, "v", K));                                              \
    else                                                                                                    \
        asm("s_nop 1\n\t" LML_D##W(LML_DOT, LML_DOT, QP) : LML_S4 : LML_V##W(a, a, "v", K), LML_V##W(c, c, "v", K))
#define LML_SDOTB(W, QP)                                                                                    \
    if constexpr (K == 0)                                                                                   \
        asm("s_nop 1\n\t" LML_D##W(LML_SDOT, LML_SDOT0, QP) : LML_S4W : LML_V##W(a, a, "v", K));            \
    else                                                                                                    \
        asm("s_nop 1\n\t" LML_D##W(LML_SDOT, LML_SDOT, QP) : LML_S4 : LML_V##W(a, a, "v", K))
#define LML_BY_D(OP, W) \
    if constexpr (D == 1) OP(W, LML_QP1); \
    else if constexpr (D == 2) OP(W, LML_QP2); \
    else OP(W, LML_QP3)
#define LML_BY_W(OP)                            \
    if constexpr (K + 16 <= H) {                \
        LML_BY_D(OP, 16);                       \
    } else {                                    \
        LML_BY_D(OP, 8);                        \
    }
#define LML_NEXT (K + 16 <= H ? K + 16 : K + 8)

template <int D, int H, int K = 0>
__device__ __forceinline__ void dpp_axpy(float (&d)[H], const float (&a)[H], float b) {
    static_assert(D >= 1 && D <= 3 && H % 8 == 0, "quad lane offset, blocks of 8 or 16");
    LML_BY_W(LML_AXPY)
    if constexpr (LML_NEXT < H) dpp_axpy<D, H, LML_NEXT>(d, a, b);
}
template <int D, int H, int K = 0>
__device__ __forceinline__ void dpp_self_axpy(float (&d)[H], float b) {
    static_assert(D >= 1 && D <= 3 && H % 8 == 0, "quad lane offset, blocks of 8 or 16");
    LML_BY_W(LML_SAXPY)
    if constexpr (LML_NEXT < H) dpp_self_axpy<D, H, LML_NEXT>(d, b);
}
template <int D, int H, int K = 0>
__device__ __forceinline__ void dpp_dot(float (&s)[4], const float (&a)[H], const float (&c)[H]) {
    static_assert(D >= 1 && D <= 3 && H % 8 == 0, "quad lane offset, blocks of 8 or 16");
    LML_BY_W(LML_DOTB)
    if constexpr (LML_NEXT < H) dpp_dot<D, H, LML_NEXT>(s, a, c);
}
template <int D, int H, int K = 0>
__device__ __forceinline__ void dpp_self_dot(float (&s)[4], const float (&a)[H]) {
    static_assert(D >= 1 && D <= 3 && H % 8 == 0, "quad lane offset, blocks of 8 or 16");
    LML_BY_W(LML_SDOTB)
    if constexpr (LML_NEXT < H) dpp_self_dot<D, H, LML_NEXT>(s, a);
}
#undef LML_NEXT
#undef LML_BY_W
#undef LML_BY_D
#undef LML_SDOTB
#undef LML_DOTB
#undef LML_SAXPY
#undef LML_AXPY
#undef LML_D16
#undef LML_D8
#undef LML_16
#undef LML_8
#undef LML_SDOT0
#undef LML_DOT0
#undef LML_M
#undef LML_SDOT
#undef LML_DOT
#undef LML_SAX
#undef LML_AX
#undef LML_S4W
#undef LML_S4
#undef LML_V16
#undef LML_V8
#undef LML_F
#undef LML_QP3
#undef LML_QP2
#undef LML_QP1

// the DPP control of the read from lane ^ D inside a quad
template <int D>
constexpr int dpp_xor_ctrl() {
    return D == 1 ? DPP_XOR1 : D == 2 ? DPP_XOR2 : DPP_XOR3;
}

// Jacobi rotation of the 2x2 Gram matrix [[a_p, g], [g, a_q]] of a column pair,
// u_p' = c (u_p - t u_q), u_q' = c (u_q + t u_p), with t = tan(theta) of the smaller
// angle: for d = a_q - a_p and e = 2g (scaled by the column scales),
//   t = sgn(d) e / (|d| + sqrt(d^2 + e^2)),   c = 1 / sqrt(1 + t^2),
// which is the textbook t = sgn(z) / (|z| + sqrt(1 + z^2)), z = d / e, multiplied
// through by e: one square root and one reciprocal, no quotient d / e, and t = 0
// for e = 0.  d^2 + e^2 stays finite while the squared column norms stay below
// ~1e19, which the caller ensures by scaling U by a power of two.
__device__ inline void jacobi_tc(float ap, float aq, float gam, float& t, float& c) {
    const float d = aq - ap, e = 2.f * gam;
    const float den = fabsf(d) + rot_sqrt(fmaf(d, d, e * e));
    t = den > 0.f ? (d >= 0.f ? e : -e) * rot_rcp(den) : 0.f;
    c = rot_rsq(fmaf(t, t, 1.f));
}

// One round of the one-sided (Hestenes) Jacobi method with whole columns: lane j
// holds column j as sc * u[], its reciprocal scale isc and al = ||column||^2; its
// pair is lane j ^ D.  The rotation's factor c goes into the column scale (and
// 1/c = (1 + t^2) c into isc), so the update is one FMA per element, reading the
// partner's element through DPP.  Both lanes of a pair form the rotation from the
// same operands in the same order, so they agree on it bit for bit.  Returns
// |g| / sqrt(a_p a_q) before the rotation (COS), or 0.
template <int N, int D, bool COS>
__device__ inline float hestenes_round(float (&u)[N], float& sc, float& isc, float& al, bool low) {
    float g4[4];
    dpp_self_dot<D, N>(g4, u);
    const XorDpp<dpp_xor_ctrl<D>()> xch;
    const float sc_o = xch(sc), al_o = xch(al);
    const float gam = (sc * sc_o) * ((g4[0] + g4[1]) + (g4[2] + g4[3]));
    const float ap = low ? al : al_o, aq = low ? al_o : al;
    float t, c;
    jacobi_tc(ap, aq, gam, t, c);
    const float coef = (low ? -t : t) * (sc_o * isc);
    dpp_self_axpy<D, N>(u, coef);
    sc *= c;
    isc *= fmaf(t, t, 1.f) * c;
    al = low ? ap - t * gam : aq + t * gam;
    if constexpr (!COS) return 0.f;
    const float den = ap * aq;
    return den > 0.f ? fabsf(gam) * rot_rsq(den) : 0.f;
}

// a value of lane ^ D (D = 0: the lane's own)
template <int D>
__device__ __forceinline__ float from_quad(float v) {
    if constexpr (D == 0)
        return v;
    else
        return XorDpp<dpp_xor_ctrl<D>()>{}(v);
}

// One round with split columns.  For the pair (P, Q) = (p, p ^ m), p with bit b
// of m clear, the lane of P holds rows [0, N/2) of column P in x[] and the lane of
// Q rows [N/2, N) of column Q; the other half of each column (the lane's "guest"
// rows: rows [0, N/2) of Q for the lane of P) lies in the y[] of lane ^ D, where
// D = 0..3 is how far the walk over the masks has gone since the guest rows last
// moved.  The lane keeps the scale, reciprocal scale and squared norm of its own
// half (sx, ix, ax), and those of the rows it holds in y[] (sy, iy, ay: the guest
// of lane ^ D), which it reads of lane ^ D through DPP operands and updates with
// the rotation of lane ^ D.  A column's norm, kept by the lane of its first half
// and by the holder of its other half, is updated alike in both (a_own - t' g and
// a_held + t g with the same rounded product), so the norms are bit-identical in
// both lanes of a pair; the two halves of one column carry scales of their own.
// The pair inner product is the sum of the two lanes' scaled half sums (one
// exchange): both lanes must form it bit for bit alike, since they each derive
// the rotation from it, so the sum is rounded once and not fused with the product
// that scaled either half (a fused form rounds the partner's half only, and the
// halves of a column then turn by angles that differ in the last bits: with nearly
// equal eigenvalues the sweeps stop converging at cosines ~1e-5).
// The rotation is written from the lane's own view, d' = a_guest - a_own: the
// lane of Q sees d' = -d and so t' = -t exactly (sign flips are exact, and a tie
// d = 0 takes t > 0 from P's view).  It is applied as two shears, so that both
// updates are in place and each may read the other lane through DPP:
//   x' = x - t' (sy/sx) y,  then  y'' = y + t' c^2 (sx/sy) x',  sx' = sx c,  sy' = sy q c,
// with q = 1 + t'^2 (c^2 = 1/q), which gives sx' x' = c (X - t' Y) and
// sy' y'' = c (Y + t' X) for X = sx x, Y = sy y (the holder of the guest rows
// applies the second shear with t' c^2 sx of lane ^ D, whose guest they are, and
// its own 1/sy).  The two halves of a column take
// the shears in opposite order, so they agree to the rounding of c^2 q = 1 and of
// the tracked reciprocal scales, a few ulp per rotation: measured on nearly scalar
// and rank-deficient problems, the sweeps converge as far as with whole-column
// rotations (and as with 1/q and the reciprocals refined by a Newton step).
template <int H, int D, bool COS>
__device__ inline float shear_round(float (&x)[H], float (&y)[H], float& sx, float& ix, float& ax, float& sy,
                                    float& iy, float& ay, bool hi, XorBperm xch) {
    float s4[4];
    if constexpr (D == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) s4[i] = x[i] * y[i];
#pragma unroll
        for (int i = 4; i < H; ++i) s4[i & 3] = fmaf(x[i], y[i], s4[i & 3]);
    } else {
        dpp_dot<D, H>(s4, y, x);
    }
    const float gsy = from_quad<D>(sy), gay = from_quad<D>(ay);  // the guest's, held by lane ^ D
    const float ps = ((s4[0] + s4[1]) + (s4[2] + s4[3])) * (sx * gsy);
    const float g = add_uncontracted(ps, xch(ps));
    const float dp = gay - ax, e = 2.f * g;
    const float den = fabsf(dp) + rot_sqrt(fmaf(dp, dp, e * e));
    // t > 0 for d' > 0, and for a tie d' = 0 in the lane of P: one integer compare
    // of the bits of d' with 0 in the lane of P and 1 in the lane of Q (+0 and the
    // positive floats have bits >= 0, the positive floats >= 1, the negative < 0)
    const bool pos = __float_as_int(dp) >= int(hi);
    // den = 0 only where e = 0, and then any finite 1/den gives t = 0
    const float t = (pos ? e : -e) * rot_rcp(fmaxf(den, 0x1p-126f));
    const float q = fmaf(t, t, 1.f);
    const float c = rot_rsq(q);
    const float co = -t * (gsy * ix);
    // the second shear's coefficient without the held rows' reciprocal scale
    const float w = (t * sx) * (c * c);
    const float ic = q * c, tg = t * g;
    if constexpr (D == 0) {
#pragma unroll
        for (int i = 0; i < H; ++i) x[i] = fmaf(co, y[i], x[i]);
        const float ka = w * iy;
#pragma unroll
        for (int i = 0; i < H; ++i) y[i] = fmaf(ka, x[i], y[i]);
    } else {
        dpp_axpy<D, H>(x, y, co);
        dpp_axpy<D, H>(y, x, from_quad<D>(w) * iy);
    }
    const float den2 = ax * gay;
    sx *= c;
    ix *= ic;
    ax = add_uncontracted(ax, -tg);
    // the held rows take the rotation of lane ^ D, whose guest they are (in the
    // block's last round as DPP source operands written out, after the two wait
    // states a DPP read needs: the compiler leaves the moves there unfolded)
    if constexpr (D == 2) {
        asm("s_nop 1\n\t"
            "v_mul_f32_dpp %0, %3, %0 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
            "v_mul_f32_dpp %1, %4, %1 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
            "v_add_f32_dpp %2, %5, %2 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf bound_ctrl:1"
            : "+v"(sy), "+v"(iy), "+v"(ay)
            : "v"(ic), "v"(c), "v"(tg));
    } else {
        sy *= from_quad<D>(ic);
        iy *= from_quad<D>(c);
        ay = add_uncontracted(ay, from_quad<D>(tg));
    }
    if constexpr (!COS) return 0.f;
    return den2 > 0.f ? fabsf(g) * rot_rsq(den2) : 0.f;
}

// One cyclic sweep (every column pair once) of the FP32 one-sided Jacobi method
// on columns held whole (lane j = column j, N columns), returning the largest pair
// cosine met.  The rounds m = 4..N-1 go in groups of equal top bit b of m: for
// group b, lanes with bit b clear keep the first N/2 rows of their column and
// receive the same rows of their partner's, the others the last N/2 rows, so each
// round costs half the inner-product FMAs and no column exchange.  Inside a group
// the masks follow a Gray code, in blocks of four rounds m0, m0 ^ 1, m0 ^ 3, m0 ^ 2:
// within a block the guest rows stay where they are and are read through DPP
// (D = 0, 1, 3, 2), and between blocks (a flip of bit j >= 2) they move once through
// the crossbar.  A group ends with the other half of each lane's column back in the
// lane; the next group keeps whichever half its split gives the lane and sends the
// other half to its new partner (one crossbar exchange, the halves keeping scales of
// their own), and only the last group rebuilds whole columns.  The rounds m = 1, 2, 3
// pair lanes inside a quad and read whole columns through DPP.  With COS the sweep
// returns the largest pair cosine it met (before each rotation), else 0.
template <int N, bool COS>
__device__ inline float hestenes_sweep(float (&u)[N], float& sc, float& isc, float& al, int lane) {
    constexpr int H = N / 2;
    constexpr int LG = (N == 64) ? 6 : (N == 32) ? 5 : 4;
    static_assert(N == 16 || N == 32 || N == 64, "column count");
    float mc = 0.f;
    int b = LG - 1;
    bool hi = ((lane >> b) & 1) != 0;
    int gsize = 1 << b;
    float x[H], y[H];
    {
        const XorBperm to{(lane ^ gsize) << 2};
#pragma unroll
        for (int i = 0; i < H; ++i) {
            x[i] = hi ? u[H + i] : u[i];
            y[i] = to(hi ? u[i] : u[H + i]);
        }
    }
    float sx = sc, ix = isc, ax = al;
    float sy = XorBperm{(lane ^ gsize) << 2}(sc), iy = XorBperm{(lane ^ gsize) << 2}(isc),
          ay = XorBperm{(lane ^ gsize) << 2}(al);
    float so, io;  // scale and reciprocal scale of the other half of the lane's column
    for (;;) {
        int m = gsize;
        for (int k = 0;; k += 4) {
            mc = fmaxf(mc, shear_round<H, 0, COS>(x, y, sx, ix, ax, sy, iy, ay, hi, XorBperm{(lane ^ m) << 2}));
            m ^= 1;
            mc = fmaxf(mc, shear_round<H, 1, COS>(x, y, sx, ix, ax, sy, iy, ay, hi, XorBperm{(lane ^ m) << 2}));
            m ^= 2;
            mc = fmaxf(mc, shear_round<H, 3, COS>(x, y, sx, ix, ax, sy, iy, ay, hi, XorBperm{(lane ^ m) << 2}));
            m ^= 1;
            mc = fmaxf(mc, shear_round<H, 2, COS>(x, y, sx, ix, ax, sy, iy, ay, hi, XorBperm{(lane ^ m) << 2}));
            // the guest rows move once, on either path: after the group's last round
            // the other half of the lane's column (the guest of lane ^ m, in lane
            // ^ m ^ 2) comes home; otherwise, with the Gray-code bit j >= 2 flipping,
            // the lane's new guest is the guest of lane ^ (1 << j), in lane ^ (1 << j) ^ 2
            const bool last = k + 4 == gsize;
            const int j = __builtin_ctz(k + 4);
            const XorBperm mv{(lane ^ (last ? m : 1 << j) ^ 2) << 2};
#pragma unroll
            for (int i = 0; i < H; ++i) y[i] = mv(y[i]);
            // their scales and norm go with them
            sy = mv(sy);
            iy = mv(iy);
            if (last) break;
            ay = mv(ay);
            m ^= 1 << j;
        }
        so = sy;
        io = iy;
        if (b == 2) break;
        // into group b - 1: keep the half of the column that its split gives the lane,
        // send the other half (and its scales) to lane ^ 2^(b-1), receive that lane's
        const bool nhi = ((lane >> (b - 1)) & 1) != 0;
        const bool same = hi == nhi;
        gsize = 1 << (b - 1);
        const XorBperm to{(lane ^ gsize) << 2};
        if (!same) {
            // the lanes whose half changes swap their two halves in place (v_swap_b32
            // under the lane mask: one instruction per element, no third register)
#pragma unroll
            for (int i = 0; i < H; ++i) asm volatile("v_swap_b32 %0, %1" : "+v"(x[i]), "+v"(y[i]));
        }
#pragma unroll
        for (int i = 0; i < H; ++i) y[i] = to(y[i]);
        sy = to(same ? so : sx);
        iy = to(same ? io : ix);
        ay = to(ax);
        sx = same ? sx : so;
        ix = same ? ix : io;
        hi = nhi;
        --b;
    }
    {
        // whole columns again, the other half brought to this half's scale sx
        const float r = so / sx;
#pragma unroll
        for (int i = 0; i < H; ++i) {
            const float tv = y[i] * r;
            u[i] = hi ? tv : x[i];
            u[H + i] = hi ? x[i] : tv;
        }
    }
    sc = sx;
    isc = 1.f / sx;
    al = ax;
    mc = fmaxf(mc, hestenes_round<N, 1, COS>(u, sc, isc, al, (lane & 1) == 0));
    mc = fmaxf(mc, hestenes_round<N, 2, COS>(u, sc, isc, al, (lane & 2) == 0));
    mc = fmaxf(mc, hestenes_round<N, 3, COS>(u, sc, isc, al, (lane & 2) == 0));
    return mc;
}

// fold the column scale into u and refresh the squared norm
template <int N, class T>
__device__ inline void hestenes_renorm(T (&u)[N], T& sc, T& al) {
    T a4[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int i = 0; i < N; ++i) {
        u[i] *= sc;
        a4[i & 3] = fma(u[i], u[i], a4[i & 3]);
    }
    sc = T(1);
    al = (a4[0] + a4[1]) + (a4[2] + a4[3]);
}

// ---------------------------------------------------------------------------------
// gather + MFMA Gram of gridpoint g into accumulators: G^T G in the tiles (I, J) of C,
// all NT x NT of them (FULL, tile I NT + J) or those with I <= J (the upper
// triangle, in row order), then b = G^T sqrt(R^-1) dyo in the NT tiles after them.
// Tile (I, J): lane (col, kq) holds rows 16 I + 4 kq + r, column 16 J + col.
// ---------------------------------------------------------------------------------
template <int N, bool FULL>
struct GramTiles {
    static constexpr int NT = N / 16;
    static constexpr int C = FULL ? NT * NT : NT * (NT + 1) / 2;
    static constexpr int ALL = C + NT;
};
// accumulator count of the FULL Gram (a name without a comma, for array bounds in
// parameter declarations)
template <int N>
constexpr int GRAM_FULL = GramTiles<N, true>::ALL;

template <int N, int PASSES, bool FULL>
__device__ __forceinline__ void gram_acc(int g, int lane, const int* __restrict__ loc_ptr,
                                         const int* __restrict__ loc_site, const float* __restrict__ loc_w,
                                         const float* __restrict__ dy, const float* __restrict__ dyo,
                                         const MixedParams& mp, f32x4 (&acc)[GramTiles<N, FULL>::ALL]) {
    constexpr int NT = N / 16;                       // member tiles of 16
    constexpr int NTILE = GramTiles<N, FULL>::C;     // tiles of C
    const int M = mp.M;
    const int nv = mp.n_vars;
#pragma unroll
    for (int a = 0; a < NTILE + NT; ++a) acc[a] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int e0 = loc_ptr[g], e1 = loc_ptr[g + 1];
    const int P = (e1 - e0) * nv;
    const int col = lane & 15;   // member within a tile (A rows / B columns)
    const int kq = lane >> 4;    // which 8 of the 32 k-values this lane holds
    // Observation index j of the local set = entry e0 + j / nv, variable j % nv.  A
    // lane's 8 consecutive k-values step (e, va) without dividing; values beyond P
    // read a valid element with weight zero, so the loop has no branches and all its
    // loads are in flight together.
    // the R^-1 weights as uniform values (a select between them, not an indexed load)
    const float cf0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(mp.coef[0])));
    const float cf1 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(mp.coef[1])));
    const float cf2 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(mp.coef[2])));
    // dY and the innovations through buffer loads: a 32-bit offset per observation,
    // the member tiles 64 bytes apart as immediate offsets, and reads past the end of
    // dY return 0.  Members 16 I + col >= M read the next row's elements (finite);
    // their rows and columns of C and rows of b are cleared after the loop.
    const __amdgpu_buffer_rsrc_t rdy =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(dy), 0, mp.n_obs * M * 4, 0x00020000);
    const __amdgpu_buffer_rsrc_t rdyo =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(dyo), 0, mp.n_obs * 4, 0x00020000);
    const float cmask = col == 0 ? 1.f : 0.f;
    for (int base = 0; base < P; base += 32) {
        const int jj0 = base + 8 * kq;
        int e = e0 + jj0 / nv, va = jj0 % nv;
        int oo[8];
        float sw[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const bool ok = jj0 + t < P;
            const int ec = ok ? e : e0, vc = ok ? va : 0;
            const float cf = vc == 0 ? cf0 : (vc == 1 ? cf1 : cf2);
            oo[t] = loc_site[ec] * nv + vc;
            const float w = loc_w[ec];
            sw[t] = __builtin_amdgcn_sqrtf(w * cf) * (ok ? 1.f : 0.f);  // 1 ulp; 0 beyond P
            const bool wrapv = va + 1 == nv;
            va = wrapv ? 0 : va + 1;
            e += wrapv ? 1 : 0;
        }
        float x[NT + 1][8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int vo = (oo[t] * M + col) * 4;
#pragma unroll
            for (int I = 0; I < NT; ++I)
                x[I][t] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rdy, vo + 64 * I, 0, 0)) *
                          sw[t];
            x[NT][t] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rdyo, oo[t] * 4, 0, 0)) *
                       (sw[t] * cmask);
        }
        // split x = hi + lo, both bf16 (hi rounded to nearest), two values per conversion
        bf16x8 hi[NT + 1], lo[NT + 1];
#pragma unroll
        for (int I = 0; I <= NT; ++I) {
            bf16x2 h[4], l[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float a = x[I][2 * t], b = x[I][2 * t + 1];
                h[t] = __builtin_convertvector(f32x2{a, b}, bf16x2);
                const unsigned bits = __builtin_bit_cast(unsigned, h[t]);
                const float ha = __builtin_bit_cast(float, bits << 16);
                const float hb = __builtin_bit_cast(float, bits & 0xffff0000u);
                l[t] = __builtin_convertvector(f32x2{a - ha, b - hb}, bf16x2);
            }
            hi[I] = __builtin_shufflevector(__builtin_shufflevector(h[0], h[1], 0, 1, 2, 3),
                                            __builtin_shufflevector(h[2], h[3], 0, 1, 2, 3), 0, 1, 2, 3, 4, 5, 6, 7);
            lo[I] = __builtin_shufflevector(__builtin_shufflevector(l[0], l[1], 0, 1, 2, 3),
                                            __builtin_shufflevector(l[2], l[3], 0, 1, 2, 3), 0, 1, 2, 3, 4, 5, 6, 7);
        }
        int tix = 0;
#pragma unroll
        for (int I = 0; I < NT; ++I) {
#pragma unroll
            for (int J = FULL ? 0 : I; J < NT; ++J, ++tix) {
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
    if (M < N) {
        // members >= M: rows and columns of C, rows of b (tile element r of lane (col,
        // kq) is row 16 I + 4 kq + r, column 16 J + col), by factors formed here from an
        // opaque copy of the lane index (hoisted out of the chain loop, the per-element
        // conditions would be held across it)
        int ln = lane;
        asm volatile("" : "+v"(ln));
        const int cl = ln & 15, kl = ln >> 4;
        float rowm[NT][4], colm[NT];
#pragma unroll
        for (int I = 0; I < NT; ++I) {
            colm[I] = 16 * I + cl < M ? 1.f : 0.f;
#pragma unroll
            for (int r = 0; r < 4; ++r) rowm[I][r] = 16 * I + 4 * kl + r < M ? 1.f : 0.f;
        }
        int tz = 0;
#pragma unroll
        for (int I = 0; I < NT; ++I) {
#pragma unroll
            for (int J = FULL ? 0 : I; J < NT; ++J, ++tz) {
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[tz][r] *= rowm[I][r] * colm[J];
            }
        }
#pragma unroll
        for (int I = 0; I < NT; ++I)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[NTILE + I][r] *= rowm[I][r];
    }
}

// gather + MFMA Gram of gridpoint g: C (full, row stride N) into LDS, b into bg[0, N)
template <int N, int PASSES>
__device__ inline void gram_to_lds(int g, int lane, const int* __restrict__ loc_ptr,
                                   const int* __restrict__ loc_site, const float* __restrict__ loc_w,
                                   const float* __restrict__ dy, const float* __restrict__ dyo,
                                   float* __restrict__ bg, const MixedParams& mp, float* cfull) {
    constexpr int NT = N / 16;
    constexpr int NTILE = GramTiles<N, false>::C;
    const int M = mp.M;
    const int col = lane & 15, kq = lane >> 4;
    f32x4 acc[GramTiles<N, false>::ALL];
    gram_acc<N, PASSES, false>(g, lane, loc_ptr, loc_site, loc_w, dy, dyo, mp, acc);
    // result tile (I,J), I <= J: lane holds rows 16I + 4*kq + r, column 16J + col;
    // the tiles below the diagonal are the mirrors of those above
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
                cfull[row * N + cc] = val;
                if (I != J) cfull[cc * N + row] = val;
            }
        }
    }
#pragma unroll
    for (int I = 0; I < NT; ++I)
        if (col == 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) bg[16 * I + 4 * kq + r] = acc[NTILE + I][r];
        }
}

// Warm start: U0 = C V_prev on the FP32 matrix cores (v_mfma_f32_16x16x4f32), where
// V_prev is the previous gridpoint's U (row-major in the workspace, just written by
// this wave) with unit columns.  The product takes U_prev as it is and scales column
// n of the result by 1 / ||u_n|| afterwards.  MFMA lane layout: A[m][k] from lane
// (m = lane % 16, k = lane / 16), B[k][n] from lane (k = lane / 16, n = lane % 16),
// D[4 (lane / 16) + r][lane % 16].  Each lane's B values are a quarter of the rows
// of columns 16 tj + lane % 16, so the column norms come from its squares summed
// over the four lanes of a column (the same sum in all four).  U0 replaces C in
// cfull (row-major) after every read of C is done.
// V_prev row-major (up[k N + n] = V[k][n]); the loads of row k give the lane
// columns 16 tj + col
template <int N>
struct RowMajorRows {
    const float* __restrict__ up;
    int col;
    __device__ void operator()(int k, float (&b)[N / 16]) const {
#pragma unroll
        for (int tj = 0; tj < N / 16; ++tj) b[tj] = up[k * N + 16 * tj + col];
    }
};
template <int N, class Rows>
__device__ inline void warm_start(Rows rows, float* cfull, int lane) {
    constexpr int NT = N / 16;
    const int col = lane & 15, kq = lane >> 4;
    f32x4 acc[NT * NT];
#pragma unroll
    for (int a = 0; a < NT * NT; ++a) acc[a] = f32x4{0.f, 0.f, 0.f, 0.f};
    float nrm[NT];
#pragma unroll
    for (int tj = 0; tj < NT; ++tj) nrm[tj] = 0.f;
#pragma unroll
    for (int ks = 0; ks < N / 4; ++ks) {
        const int k = 4 * ks + kq;
        float a[NT], b[NT];
#pragma unroll
        for (int ti = 0; ti < NT; ++ti) a[ti] = cfull[k * N + 16 * ti + col];  // C symmetric: no bank conflicts
        rows(k, b);
#pragma unroll
        for (int tj = 0; tj < NT; ++tj) nrm[tj] = fmaf(b[tj], b[tj], nrm[tj]);
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
            for (int tj = 0; tj < NT; ++tj)
                acc[ti * NT + tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ti], b[tj], acc[ti * NT + tj], 0, 0, 0);
    }
    float inv[NT];
#pragma unroll
    for (int tj = 0; tj < NT; ++tj) {
        float nn = nrm[tj] + __shfl_xor(nrm[tj], 16);
        nn = nn + __shfl_xor(nn, 32);
        inv[tj] = nn > 0.f ? rot_rsq(nn) : 0.f;
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
        for (int tj = 0; tj < NT; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                cfull[(16 * ti + 4 * kq + r) * N + 16 * tj + col] = acc[ti * NT + tj][r] * inv[tj];
}

// U0 (C, or the warm start) from LDS into the registers (lane j = column j; C is
// symmetric, so rows and columns read alike), then the FP32 one-sided sweeps.  On
// return the columns are sc * u[] and the matrix U = uscale * sc * u (uscale a power
// of two that keeps the squared column norms in range).
template <int N>
__device__ float gram_max_cos(const float (&u)[N], float* cv, int lane);

// With CHECK, the columns' largest pair cosine is measured on the matrix cores
// (check_lds: scratch for gram_max_cos) after each sweep from sweep `first` on, and
// that measurement alone decides the stop; `sweeps` returns the sweep count.
template <int N, bool CHECK>
__device__ __forceinline__ float jacobi_fp32(float (&u)[N], float& sc, const float* cfull, int lane,
                                             const MixedParams& mp, float* check_lds, int first, int& sweeps) {
    const int l = lane & (N - 1);
    float a0 = 0.f;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const float x = cfull[i * N + l];
        u[i] = (lane < N) ? x : 0.f;
        a0 = fmaf(u[i], u[i], a0);
    }
    // keep the squared column norms below 1e18 (the range of the rotation
    // formula): scale U by 2^-32 when needed, exactly, and undo it on output
    float uscale = 1.f;
    if (wave_max(a0) > 1e18f) {
        uscale = 0x1p32f;
#pragma unroll
        for (int i = 0; i < N; ++i) u[i] *= 0x1p-32f;
    }

    // ---- 3. FP32 one-sided Jacobi ------------------------------------------------
    // lanes >= N (N < 64) hold zero columns: their pair Gram is zero, they never rotate
    // Stop at the tolerance, or when the largest cosine below 1e-4 no longer
    // halves from one sweep to the next: that is the FP32 rounding floor of
    // this matrix (convergence, even slowed by clustered eigenvalues, keeps
    // dividing it by more than 2).
    float al = 0.f;
    sc = 1.f;
    int sw32 = 0;
    if constexpr (CHECK) {
        // The sweeps form no cosine of their own (a product, a reciprocal square root
        // and a maximum per round).  The cosine measured after sweep k is, to first
        // order, the largest one that sweep k + 1 meets before its rotations, so when
        // the rate test on it fires, one more sweep runs, as the same test on the
        // cosines met by the sweeps would have it: on clustered eigenvalues that sweep
        // still halves the columns' departure from orthogonality.
        float gm_prev = 1.f;
        bool last = false;
        while (sw32 < mp.max32) {
            hestenes_renorm<N, float>(u, sc, al);
            float isc = 1.f;
            hestenes_sweep<N, false>(u, sc, isc, al, lane);
            ++sw32;
            if (last) break;
            if (sw32 < first) continue;
            const float gm = gram_max_cos<N>(u, check_lds, lane);
            if (gm <= mp.check_tol) break;
            last = gm <= 1e-4f && gm > 0.5f * gm_prev;
            gm_prev = gm;
        }
    } else {
        float mc_prev = 1.f;
        while (sw32 < mp.max32) {
            hestenes_renorm<N, float>(u, sc, al);
            float isc = 1.f;
            float mc = hestenes_sweep<N, true>(u, sc, isc, al, lane);
            mc = wave_max(mc);
            ++sw32;
            if (mc <= mp.tol32 || (mc <= 1e-4f && mc > 0.5f * mc_prev)) break;
            mc_prev = mc;
        }
    }
    if (lane == 0 && mp.sweeps32) atomicAdd(mp.sweeps32, (unsigned long long)sw32);
    sweeps = sw32;
    return uscale;
}

// ---------------------------------------------------------------------------------
// kernel 1: gather, MFMA Gram, FP32 sweeps for a chain of mp.chain consecutive
// gridpoints.  The first starts from U = C; each later one from U = C V_prev, the
// eigenvectors of its neighbour (V_prev = the neighbour's U with unit columns),
// which leaves far fewer sweeps: the local observation sets of neighbouring
// gridpoints share almost all sites and weights.  U = C V with V orthogonal either
// way, so the eigenpairs still come out as column norms and directions.
// ws_u[g][i][j] = U[i][j], ws_b[g][i] = b_i.
// ---------------------------------------------------------------------------------
// Workgroup bx of nbx (nbx a multiple of 8, or the whole grid) of one analysis.
template <int N, int PASSES>
__device__ __forceinline__ void solve_body(int bx, int nbx, const int* __restrict__ loc_ptr,
                                           const int* __restrict__ loc_site, const float* __restrict__ loc_w,
                                           const float* __restrict__ dy, const float* __restrict__ dyo,
                                           float* __restrict__ ws_u, float* __restrict__ ws_b,
                                           const MixedParams& mp, float* cfull, double* __restrict__ dump) {
    // cfull: N*N floats of LDS, 16-byte aligned
    const int lane = threadIdx.x;
    const int g_begin = xcd_order(bx, nbx) * mp.chain;
    const int g_end = min(g_begin + mp.chain, mp.nodes);
    for (int g = g_begin; g < g_end; ++g) {
        __syncthreads();  // the previous gridpoint's reads of cfull are done
        gram_to_lds<N, PASSES>(g, lane, loc_ptr, loc_site, loc_w, dy, dyo, ws_b + (long long)g * N, mp, cfull);
        __syncthreads();
        if (dump && g >= mp.dump_g0 && g < mp.dump_g0 + mp.dump_n) {
            double* dc = dump + (long long)(g - mp.dump_g0) * dump_record(mp.M) + mp.M * mp.M + 2 * mp.M;
#pragma unroll 1
            for (int idx = lane; idx < mp.M * mp.M; idx += 64) dc[idx] = double(cfull[(idx / mp.M) * N + idx % mp.M]);
        }
        float u[N];
        if (g > g_begin) {
            warm_start<N>(RowMajorRows<N>{ws_u + (long long)(g - 1) * N * N, lane & 15}, cfull, lane);
            __syncthreads();
        }
        float sc;
        int sweeps;
        const float uscale = jacobi_fp32<N, false>(u, sc, cfull, lane, mp, nullptr, 0, sweeps);
        if (lane < N) {
            float* ug = ws_u + (long long)g * N * N;
            const float so = sc * uscale;
#pragma unroll
            for (int i = 0; i < N; ++i) ug[i * N + lane] = u[i] * so;
        }
    }
}

template <int N, int PASSES>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_mixed_solve(
    const int* __restrict__ loc_ptr, const int* __restrict__ loc_site, const float* __restrict__ loc_w,
    const float* __restrict__ dy, const float* __restrict__ dyo, float* __restrict__ ws_u,
    float* __restrict__ ws_b, MixedParams mp, double* __restrict__ dump) {
    __shared__ __align__(16) float cfull[N * N];
    solve_body<N, PASSES>(blockIdx.x, gridDim.x, loc_ptr, loc_site, loc_w, dy, dyo, ws_u, ws_b, mp, cfull, dump);
}

// ---------------------------------------------------------------------------------
// The fused analysis (the presets: no FP64 sweeps).  The gather, Gram and FP32
// sweeps of k_mixed_solve, then in the same wavefront the eigenpairs and the update
// of the gridpoint's members:
//   lambda_i = ||u_i||, v_i = u_i / lambda_i          (column norms in FP64)
//   s_i = sqrt(M-1) / sqrt(lambda_i), h_i = (V^T b)_i / lambda_i   (FP64)
//   Y = [dX; b^T] V                                   (FP32 matrix cores)
//   Xa[q][l] = xbar[q] + sum_i Y[q][i] (s_i V[l][i] + h_i)   (FP32 matrix cores)
// which is Xa = xbar 1^T + dX (W + w 1^T) with W = sqrt(M-1) V D^-1/2 V^T and
// w = V D^-1 V^T b (letkf_solver.h:288-316), since dX w = dX V D^-1 V^T b = Y h.
// The next gridpoint of the chain (the next row of the same column) starts from
// C V with this V, which stays in LDS; C stays in the Gram's accumulators.  No
// intermediate leaves the CU: the kernel reads the local observation lists, dY and
// the state, and writes the analysed state.
// ---------------------------------------------------------------------------------
constexpr int ASM_ROWS = Q + 1;  // dX rows and b^T: the A operand of the first product

template <int N>
struct FusedShared {
    float cv[N * N];             // C, then U0, then V with swizzled rows (vswz)
    float stage[N * ASM_ROWS];   // [k][ASM_ROWS]: dX^T and b (first product), then Y^T (second)
    float bvec[N];               // b = G^T sqrt(R^-1) dyo from the Gram phase
    float coef[2][N];            // s_i, h_i
};

// Element (k, i) of V sits at cv[k N + (i ^ vswz(k))].  The MFMA operand reads of
// the warm start (lane (col, kq) reading rows k = 16 I + 4 kq + r, columns 16 tj +
// col), of the first product (rows 4 s + grp, columns 16 T + c) and of the second
// (rows 16 T + c, columns 4 s + grp), and the writes of a row by lane = column,
// each meet 64 distinct banks: vswz puts bits 2-3 of k into bits 2-3 and
// (bits 0-1 XOR bits 2-3) of k into bits 4-5.
template <int N>
__device__ inline int vswz(int k) {
    return (4 * ((k >> 2) & 3) + 16 * ((k & 3) ^ ((k >> 2) & 3))) & (N - 1);
}

// Largest pair cosine |u_p . u_q| / (|u_p| |u_q|) of the columns (lane l holds
// column l in u[0..N), up to a per-column scale, which cancels): the columns go to
// cv (row k at cv[k N + (col ^ vswz(k))]) and the tiles (I, J), I <= J, of U^T U are
// formed on the FP32 matrix cores (v_mfma_f32_16x16x4f32, exact products, FP32
// sums: the cosine is resolved to ~1e-6).  Uses all of cv; wave-uniform result.
template <int N>
__device__ float gram_max_cos(const float (&u)[N], float* cv, int lane) {
    constexpr int NT = N / 16;
    constexpr int NTU = NT * (NT + 1) / 2;
    int ln = lane;
    asm volatile("" : "+v"(ln));  // addresses formed here, not hoisted into the sweep loop
    if (ln < N) {
#pragma unroll
        for (int k = 0; k < N; ++k) cv[k * N + (ln ^ vswz<N>(k))] = u[k];
    }
    __syncthreads();
    const int c = ln & 15, g = ln >> 4;
    f32x4 acc[NTU];
#pragma unroll
    for (int t = 0; t < NTU; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < N / 4; ++s) {
        const int k = 4 * s + g;
        float a[NT];
#pragma unroll
        for (int I = 0; I < NT; ++I) a[I] = cv[k * N + ((16 * I + c) ^ vswz<N>(k))];
        int t = 0;
#pragma unroll
        for (int I = 0; I < NT; ++I)
#pragma unroll
            for (int J = I; J < NT; ++J, ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[I], a[J], acc[t], 0, 0, 0);
    }
    // tile (I, J): lane (c, g) holds G[16 I + 4 g + r][16 J + c]; the diagonal from the
    // tiles (I, I) where 4 g + r == c
    __syncthreads();  // every read of the columns is done
    {
        int t = 0;
#pragma unroll
        for (int I = 0; I < NT; ++I) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (4 * g + r == c) cv[16 * I + c] = acc[t][r];
            t += NT - I;
        }
    }
    __syncthreads();
    float mx = 0.f;
    int t = 0;
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
        for (int J = I; J < NT; ++J, ++t) {
            const float dq = cv[16 * J + c];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int p = 16 * I + 4 * g + r, q = 16 * J + c;
                if (p < q) {
                    const float d = cv[p] * dq;
                    const float cs = d > 0.f ? fabsf(acc[t][r]) * __builtin_amdgcn_rsqf(d) : 0.f;
                    mx = fmaxf(mx, cs);
                }
            }
        }
    __syncthreads();  // cv is free again
    return wave_max(mx);
}

// C with its diagonal from the FULL Gram accumulators: (M-1)/beta on the members'
// diagonal, 1 on the padding columns (a decoupled identity block)
template <int N>
__device__ inline void gram_add_diag(f32x4 (&acc)[GRAM_FULL<N>], int lane, int M, float diag) {
    constexpr int NT = N / 16;
    const int col = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (4 * kq + r == col) acc[I * NT + I][r] += (16 * I + col < M) ? diag : 1.f;
}

// C from the FULL accumulators into a dump record (row-major M x M, FP64); the
// element offsets come from an opaque copy of the lane index (the dump is rare and
// its addresses must not be hoisted out of the chain loop)
template <int N>
__device__ inline void dump_c_acc(const f32x4 (&acc)[GRAM_FULL<N>], double* dc, int M, int lane) {
    constexpr int NT = N / 16;
    int ln = lane;
    asm volatile("" : "+v"(ln));
    const int col = ln & 15, kq = ln >> 4;
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
        for (int J = 0; J < NT; ++J)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * I + 4 * kq + r, cc = 16 * J + col;
                if (row < M && cc < M) dc[row * M + cc] = double(acc[I * NT + J][r]);
            }
}

// C from the FULL accumulators into cv, row-major (the first gridpoint of a chain)
template <int N>
__device__ inline void gram_to_cv(const f32x4 (&acc)[GRAM_FULL<N>], float* cv, int lane) {
    constexpr int NT = N / 16;
    const int col = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int I = 0; I < NT; ++I)
#pragma unroll
        for (int J = 0; J < NT; ++J)
#pragma unroll
            for (int r = 0; r < 4; ++r) cv[(16 * I + 4 * kq + r) * N + 16 * J + col] = acc[I * NT + J][r];
}

// Warm start of a chain's later gridpoints without leaving the CU: U0 = C V_prev on
// the FP32 matrix cores (v_mfma_f32_16x16x4f32), with C read from the Gram's FULL
// accumulators as the A operand and V_prev, the previous gridpoint's eigenvectors
// (unit columns), from cv where its update left them.  The sum over k runs in the
// order k = 16 (s >> 2) + 4 kq + (s & 3) for K-step s and lane group kq, so that
// the A value C[16 ti + col][k] = C[k][16 ti + col] (C is symmetric) is register
// s & 3 of the lane's own accumulator tile (s >> 2, ti): no data moves.  MFMA lane
// layout: A[m][kk] from lane (m = lane % 16, kk = lane / 16), B[kk][n] from lane
// (kk = lane / 16, n = lane % 16), D[4 (lane / 16) + r][lane % 16].  U0 then
// replaces V_prev in cv (row-major).
template <int N>
__device__ inline void warm_start_acc(const f32x4 (&acc)[GRAM_FULL<N>], float* cv, int lane) {
    constexpr int NT = N / 16;
    // the swizzled read addresses are formed from an opaque copy of the lane index:
    // hoisted out of the chain loop, they would stay live across the sweeps and spill
    int ln = lane;
    asm volatile("" : "+v"(ln));
    const int col = ln & 15, kq = ln >> 4;
    f32x4 d[NT * NT];
#pragma unroll
    for (int a = 0; a < NT * NT; ++a) d[a] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < N / 4; ++s) {
        const int k = 16 * (s >> 2) + 4 * kq + (s & 3);
        float b[NT];
#pragma unroll
        for (int tj = 0; tj < NT; ++tj) b[tj] = cv[k * N + ((16 * tj + col) ^ vswz<N>(k))];
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
            for (int tj = 0; tj < NT; ++tj)
                d[ti * NT + tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(acc[(s >> 2) * NT + ti][s & 3], b[tj],
                                                                       d[ti * NT + tj], 0, 0, 0);
    }
    __syncthreads();  // every read of V_prev is done
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
        for (int tj = 0; tj < NT; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r) cv[(16 * ti + 4 * kq + r) * N + 16 * tj + col] = d[ti * NT + tj][r];
}

template <int N, int PASSES>
__device__ __forceinline__ void analysis_body(int bx, int nbx, const int* __restrict__ loc_ptr,
                                              const int* __restrict__ loc_site, const float* __restrict__ loc_w,
                                              const float* __restrict__ dy, const float* __restrict__ dyo,
                                              float* __restrict__ f, const MixedParams& mp,
                                              double* __restrict__ dump, int* __restrict__ status,
                                              FusedShared<N>& sh) {
    constexpr int NT = N / 16;
    const int lane = threadIdx.x;
    const int M = mp.M;
    // Chain blk covers rows [gy0, gy1) of column gx: the chains that run at the same
    // time on one XCD (consecutive blk) are neighbouring columns of the same rows,
    // so they share the cache lines of the member states (a line holds 32 columns
    // of one member and population) instead of each holding its own
    const int blk = xcd_order(bx, nbx);
    const int gx = blk % mp.nx;
    const int gy0 = (blk / mp.nx) * mp.chain;
    const int gy1 = min(gy0 + mp.chain, mp.ny);
    bool bad = false;
    int sweeps = 2;
    for (int gy = gy0; gy < gy1; ++gy) {
        const int g = gy * mp.nx + gx;
        f32x4 acc[GRAM_FULL<N>];
        gram_acc<N, PASSES, true>(g, lane, loc_ptr, loc_site, loc_w, dy, dyo, mp, acc);
        gram_add_diag<N>(acc, lane, M, mp.diag);
        if (mp.write_state && g >= mp.dump_g0 && g < mp.dump_g0 + mp.dump_n)
            dump_c_acc<N>(acc, dump + (long long)(g - mp.dump_g0) * dump_record(M) + M * M + 2 * M, M, lane);
        __syncthreads();  // the previous gridpoint's reads of stage and coef are done
        if ((lane & 15) == 0) {
#pragma unroll
            for (int I = 0; I < NT; ++I)
#pragma unroll
                for (int r = 0; r < 4; ++r) sh.bvec[16 * I + 4 * (lane >> 4) + r] = acc[NT * NT + I][r];
        }
        // U0: C itself for the first gridpoint of the chain, else C V_prev
        if (gy == gy0)
            gram_to_cv<N>(acc, sh.cv, lane);
        else
            warm_start_acc<N>(acc, sh.cv, lane);
        __syncthreads();
        float u[N];
        float sc;
        // The first measurement follows the second sweep, or for a warm start the
        // sweep count of the previous gridpoint up to 3: the cosines of a warm start
        // fall ~0.1 -> 1e-3 -> floor, and a measurement that fails costs about a
        // seventh of a sweep (0.22 ms per analysis at 256^2 x 64 when made after
        // every warm start's second sweep).
        const int first = gy == gy0 ? 2 : min(max(sweeps, 2), 3);
        const float uscale = jacobi_fp32<N, true>(u, sc, sh.cv, lane, mp, sh.cv, first, sweeps);

        // ---- eigenpairs: lambda_i = ||u_i|| in FP64, v_i = u_i / ||u_i|| ------------------
        double n4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int i = 0; i < N; ++i) {
            u[i] *= sc;
            n4[i & 3] = fma(double(u[i]), double(u[i]), n4[i & 3]);
        }
        const bool col = lane < N;
        const double nrm = sqrt((n4[0] + n4[1]) + (n4[2] + n4[3]));
        const double lam = (col && nrm > 0.0) ? nrm * double(uscale) : 1.0;
        {
            const float inv = (col && nrm > 0.0) ? float(1.0 / nrm) : 0.f;
#pragma unroll
            for (int i = 0; i < N; ++i) u[i] *= inv;
        }
        // lane-dependent LDS addresses of the assembly are formed from an opaque copy of
        // the lane index, so that they are not hoisted out of the chain loop (held
        // across the sweeps they would spill)
        int ln = lane;
        asm volatile("" : "+v"(ln));
        const int c = ln & 15, grp = ln >> 4;  // MFMA layout
        __syncthreads();  // every read of U0 in cv is done
        if (col) {
#pragma unroll
            for (int k = 0; k < N; ++k) sh.cv[k * N + (ln ^ vswz<N>(k))] = u[k];
        }
        const bool dumping = (g >= mp.dump_g0) && (g < mp.dump_g0 + mp.dump_n) && mp.write_state;
        if (dumping && lane < M) {
            // V[k][lane]: the lane's own column read back from cv and normalised again in
            // FP64, so that the dumped columns have unit norm to FP64 precision (rolled
            // loops: the dump is rare and its addresses must not be hoisted)
            double* dg = dump + (long long)(g - mp.dump_g0) * dump_record(M);
            double vv = 0.0;
#pragma unroll 1
            for (int k = 0; k < N; ++k) {
                const double x = double(sh.cv[k * N + (ln ^ vswz<N>(k))]);
                vv = fma(x, x, vv);
            }
            const double iv = vv > 0.0 ? 1.0 / sqrt(vv) : 0.0;
#pragma unroll 1
            for (int k = 0; k < M; ++k) dg[k * M + lane] = double(sh.cv[k * N + (ln ^ vswz<N>(k))]) * iv;
            dg[M * M + lane] = lam;
            dg[M * M + M + lane] = double(sh.bvec[lane]);
        }
        // ---- member perturbations: lane = member --------------------------------------
        const int member = lane < M ? lane : 0;
        const long long base = (long long)member * mp.member_stride + g;
        double xbar[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const float x = f[base + (long long)q * mp.nodes];
            const double xd = lane < M ? double(x) : 0.0;
            xbar[q] = wave_sum_d(xd) / double(M);
            if (col) sh.stage[lane * ASM_ROWS + q] = lane < M ? float(xd - xbar[q]) : 0.f;
        }
        if (col) sh.stage[lane * ASM_ROWS + Q] = sh.bvec[lane];
        __syncthreads();
        // ---- first product: Y[r][i] = sum_k A[r][k] V[k][i], A = [dX; b^T] ------------------
        // lane (c, grp) holds Y[4 grp + r][16 T + c]
        f32x4 y[NT];
#pragma unroll
        for (int T = 0; T < NT; ++T) y[T] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < N / 4; ++s) {
            const int k = 4 * s + grp;
            const float a = c < ASM_ROWS ? sh.stage[k * ASM_ROWS + c] : 0.f;
#pragma unroll
            for (int T = 0; T < NT; ++T) {
                const float b = sh.cv[k * N + ((16 * T + c) ^ vswz<N>(k))];
                y[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, y[T], 0, 0, 0);
            }
        }
        // ---- coefficients s_i, h_i (FP64), lane i -------------------------------------------
        // (V^T b)_i = Y[Q][i] sits in lane (i % 16, grp = Q / 4), element Q % 4 of tile i / 16
        {
            float vb = 0.f;
#pragma unroll
            for (int T = 0; T < NT; ++T) {
                const float t = __shfl(y[T][Q % 4], (Q / 4) * 16 + (lane & 15));
                vb = (lane >> 4) == T ? t : vb;
            }
            if (col) {
                sh.coef[0][lane] = float(mp.sqm1 / sqrt(lam));
                sh.coef[1][lane] = float(double(vb) / lam);
            }
        }
        __syncthreads();  // the first product's reads of stage are done; coef written
        // Y^T into stage (rows q < Q only)
#pragma unroll
        for (int T = 0; T < NT; ++T)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int q = 4 * grp + r;
                if (q < Q) sh.stage[(16 * T + c) * ASM_ROWS + q] = y[T][r];
            }
        __syncthreads();
        // ---- second product: P[q][l] = sum_i Y[q][i] (s_i V[l][i] + h_i) ---------------------
        f32x4 p[NT];
#pragma unroll
        for (int T = 0; T < NT; ++T) p[T] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < N / 4; ++s) {
            const int i = 4 * s + grp;
            const float a = c < Q ? sh.stage[i * ASM_ROWS + c] : 0.f;
            const float si = sh.coef[0][i], hi = sh.coef[1][i];
#pragma unroll
            for (int T = 0; T < NT; ++T) {
                const int l = 16 * T + c;
                const float b = fmaf(si, sh.cv[l * N + (i ^ vswz<N>(l))], hi);
                p[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, p[T], 0, 0, 0);
            }
        }
        // ---- Xa = xbar + P: lane (c, grp) holds members 16 T + c, variables 4 grp + r --------
        if (mp.write_state) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int q = 4 * grp + r;
                const double xb = grp == 0 ? xbar[r] : (grp == 1 ? xbar[4 + r] : (r == 0 ? xbar[8] : 0.0));
#pragma unroll
                for (int T = 0; T < NT; ++T) {
                    const int l = 16 * T + c;
                    if (q < Q && l < M) {
                        const float xa = float(xb + double(p[T][r]));
                        bad |= !isfinite(xa);
                        f[(long long)l * mp.member_stride + (long long)q * mp.nodes + g] = xa;
                    }
                }
            }
        }
    }
    if (bad) atomicOr(status, 1);
}

template <int N, int PASSES>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_mixed_analysis(
    const int* __restrict__ loc_ptr, const int* __restrict__ loc_site, const float* __restrict__ loc_w,
    const float* __restrict__ dy, const float* __restrict__ dyo, float* __restrict__ f, MixedParams mp,
    double* __restrict__ dump, int* __restrict__ status) {
    __shared__ __align__(16) FusedShared<N> sh;
    analysis_body<N, PASSES>(blockIdx.x, gridDim.x, loc_ptr, loc_site, loc_w, dy, dyo, f, mp, dump, status, sh);
}

// ---------------------------------------------------------------------------------
// kernel 2: two wavefronts per gridpoint.  Wave w holds rows [w N/2, (w+1) N/2) of
// every column (lane j = column j), so a pair inner product is the sum of the two
// waves' partial sums, exchanged through LDS with one barrier per round (the
// exchange buffer alternates between rounds).  Both waves then form the same
// rotation from the same operands in the same order.
// ---------------------------------------------------------------------------------
template <int N>
struct FinishLds {
    static constexpr int H = N / 2;    // rows per wave
    static constexpr int TQ = 16;      // columns of V transposed at a time
    static constexpr int T_STRIDE = TQ + 1;
};

// Sum of the two waves' values of x at this lane (same order in both waves).
__device__ inline double pair_sum(double x, double* buf, int w, int lane) {
    buf[w * 64 + lane] = x;
    __syncthreads();
    return buf[lane] + buf[64 + lane];
}

// one rotation round of the split columns; `part` is this round's exchange buffer
template <int H>
__device__ inline double hestenes_round_split(double (&u)[H], double& sc, double& al, int m, int lane, int w,
                                              double* part) {
    const bool low = (lane & (1 << (31 - __clz(m)))) == 0;
    double o[H];
    double g4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < H; ++i) {
        o[i] = __shfl_xor(u[i], m);
        g4[i & 3] = fma(u[i], o[i], g4[i & 3]);
    }
    const double gp = pair_sum((g4[0] + g4[1]) + (g4[2] + g4[3]), part, w, lane);
    const double sc_o = __shfl_xor(sc, m), al_o = __shfl_xor(al, m);
    const double gam = (sc * sc_o) * gp;
    const double ap = low ? al : al_o, aq = low ? al_o : al;
    double t = 0.0, c = 1.0;
    if (gam != 0.0) {
        const double z = (aq - ap) / (2.0 * gam);
        t = (z >= 0.0 ? 1.0 : -1.0) / (fabs(z) + sqrt(z * z + 1.0));
        c = 1.0 / sqrt(t * t + 1.0);
    }
    const double coef = (low ? -t : t) * (sc_o / sc);
#pragma unroll
    for (int i = 0; i < H; ++i) u[i] = fma(coef, o[i], u[i]);
    sc *= c;
    al = low ? ap - t * gam : aq + t * gam;
    const double den = ap * aq;
    return den > 0.0 ? fabs(gam) / sqrt(den) : 0.0;
}

// fold the column scale into the wave's rows and refresh the squared column norm
template <int H>
__device__ inline void renorm_split(double (&u)[H], double& sc, double& al, int lane, int w, double* buf) {
    double a4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < H; ++i) {
        u[i] *= sc;
        a4[i & 3] = fma(u[i], u[i], a4[i & 3]);
    }
    sc = 1.0;
    al = pair_sum((a4[0] + a4[1]) + (a4[2] + a4[3]), buf, w, lane);
}

template <int N>
struct FinishShared {
    static constexpr int TS = FinishLds<N>::T_STRIDE;
    double xpart[3][2 * 64];  // pair-sum buffers: two alternating rounds + norms/h
    double bvec[N];
    double xbar[Q];
    double Tq[N * TS];        // a TQ-column chunk of V, row-major
    double Zp[2][Q * N];      // each wave's part of s_i (dX V)[q][i]
    float dXs[Q * N];
    double hv[N];
};

template <int N, bool SWEEP64>
__device__ __forceinline__ void finish_body(int bx, int nbx, const float* __restrict__ ws_u,
                                            const float* __restrict__ ws_b, float* __restrict__ f,
                                            const MixedParams& mp, double* __restrict__ dump,
                                            int* __restrict__ status, FinishShared<N>& sh) {
    constexpr int H = FinishLds<N>::H, TQ = FinishLds<N>::TQ, TS = FinishLds<N>::T_STRIDE;
    auto& xpart = sh.xpart;
    auto& bvec = sh.bvec;
    auto& xbar = sh.xbar;
    auto& Tq = sh.Tq;
    auto& Zp = sh.Zp;
    auto& dXs = sh.dXs;
    auto& hv = sh.hv;

    const int g = xcd_order(bx, nbx);
    if (g >= mp.nodes) return;  // padding workgroup of a group launch (uniform per workgroup)
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int M = mp.M;
    const int l = lane & (N - 1);
    const int r0 = w * H;

    double ud[H];
    {
        const float* ug = ws_u + (long long)g * N * N;
#pragma unroll
        for (int i = 0; i < H; ++i) {
            const float x = ug[(r0 + i) * N + l];
            ud[i] = (lane < N) ? double(x) : 0.0;
        }
        if (w == 0 && lane < N) bvec[lane] = double(ws_b[(long long)g * N + lane]);
    }
    double sc = 1.0, al = 0.0;

    // ---- 4. FP64 sweeps on the widened columns ------------------------------------------
    if constexpr (SWEEP64) {
        int sw64 = 0;
        while (sw64 < mp.max64) {
            renorm_split<H>(ud, sc, al, lane, w, xpart[2]);
            double mc = 0.0;
            for (int m = 1; m < N; ++m)
                mc = fmax(mc, hestenes_round_split<H>(ud, sc, al, m, lane, w, xpart[m & 1]));
            mc = wave_max(mc);  // identical in both waves, so both leave the loop together
            ++sw64;
            if (mc <= mp.tol64) break;
        }
        if (threadIdx.x == 0 && mp.sweeps64) atomicAdd(mp.sweeps64, (unsigned long long)sw64);
    }
    // eigenpairs: d = ||u||, v = u / d  (U = C V with orthogonal columns)
    __syncthreads();
    renorm_split<H>(ud, sc, al, lane, w, xpart[2]);
    const double lam = (lane < N) ? sqrt(al) : 1.0;
    {
        const double inv = 1.0 / lam;
#pragma unroll
        for (int i = 0; i < H; ++i) ud[i] *= inv;
    }

    // ---- 5. assembly in FP64 --------------------------------------------------------
    // column layout (lane = eigen-index i): d_i, h_i = (V^T b)_i / d_i, s_i = sqrt(M-1)/sqrt(d_i)
    double hpart = 0.0;
#pragma unroll
    for (int r = 0; r < H; ++r) hpart += ud[r] * bvec[r0 + r];
    __syncthreads();
    const double hs = pair_sum(hpart, xpart[2], w, lane);
    const double d_i = lam;
    const double h_i = (lane < N) ? hs / d_i : 0.0;
    const double s_i = (lane < N) ? mp.sqm1 / sqrt(d_i) : 0.0;
    const bool dumping = (g >= mp.dump_g0) && (g < mp.dump_g0 + mp.dump_n);
    if (dumping && lane < M) {
        double* dg = dump + (long long)(g - mp.dump_g0) * dump_record(M);
#pragma unroll
        for (int k = 0; k < H; ++k)  // static indices keep ud[] in registers
            if (r0 + k < M) dg[(r0 + k) * M + lane] = ud[k];  // V[r0+k][lane]
        if (w == 0) {
            dg[M * M + lane] = d_i;
            dg[M * M + M + lane] = bvec[lane];
        }
    }
    // members: row layout in wave 0, lane = member
    float xk[Q];
    const int member = lane < M ? lane : 0;
    const long long base = (long long)member * mp.member_stride + g;
#pragma unroll
    for (int q = 0; q < Q; ++q) xk[q] = 0.f;
    if (w == 0 && lane < M) {
#pragma unroll
        for (int q = 0; q < Q; ++q) xk[q] = f[base + (long long)q * mp.nodes];
    }
    if (w == 0) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const double s = wave_sum_d(double(xk[q]));
            if (lane == 0) xbar[q] = s / double(M);
        }
        if (lane < N) hv[lane] = h_i;
    }
    __syncthreads();
    if (w == 0 && lane < N) {
#pragma unroll
        for (int q = 0; q < Q; ++q) dXs[q * N + lane] = (lane < M) ? float(double(xk[q]) - xbar[q]) : 0.f;
    }
    __syncthreads();
    // Z'[q][i] = s_i sum_r dX[q][r] V[r][i], each wave over its rows  (column layout)
    if (lane < N) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            double z = 0.0;
#pragma unroll
            for (int r = 0; r < H; ++r) z += double(dXs[q * N + r0 + r]) * ud[r];
            Zp[w][q * N + lane] = s_i * z;
        }
    }
    __syncthreads();
    if (w == 0 && lane < N) {
#pragma unroll
        for (int q = 0; q < Q; ++q) Zp[0][q * N + lane] += Zp[1][q * N + lane];
    }
    // dXa[q][l] = sum_i Z'[q][i] V[l][i],  w_l = sum_i V[l][i] h_i   (row layout, lane =
    // member l): wave 0 forms q = 0..4 and w, wave 1 q = 5..8
    constexpr int QW0 = 5;
    double dxa[QW0];
#pragma unroll
    for (int k = 0; k < QW0; ++k) dxa[k] = 0.0;
    double wl = 0.0;
#pragma unroll
    for (int c0 = 0; c0 < N; c0 += TQ) {
        __syncthreads();
        if (lane >= c0 && lane < c0 + TQ) {
#pragma unroll
            for (int r = 0; r < H; ++r) Tq[(r0 + r) * TS + (lane - c0)] = ud[r];
        }
        __syncthreads();
        if (lane < N) {
            if (w == 0) {
#pragma unroll 2
                for (int ii = 0; ii < TQ; ++ii) {
                    const double vli = Tq[lane * TS + ii];  // V[lane][c0+ii]
                    wl += vli * hv[c0 + ii];
#pragma unroll
                    for (int k = 0; k < QW0; ++k) dxa[k] += Zp[0][k * N + c0 + ii] * vli;
                }
            } else {
#pragma unroll 2
                for (int ii = 0; ii < TQ; ++ii) {
                    const double vli = Tq[lane * TS + ii];
#pragma unroll
                    for (int k = 0; k < Q - QW0; ++k) dxa[k] += Zp[0][(QW0 + k) * N + c0 + ii] * vli;
                }
            }
        }
    }
    // w to both waves; Xa = xbar + dXa + (dX w) 1^T, each wave for its q
    __syncthreads();
    if (w == 0) xpart[0][lane] = wl;
    __syncthreads();
    const double wlv = xpart[0][lane];
    const int qb = w == 0 ? 0 : QW0, nq = w == 0 ? QW0 : Q - QW0;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < QW0; ++k) {
        if (k < nq) {  // uniform per wave
            const int q = qb + k;
            const double dxw = wave_sum_d((lane < M) ? double(dXs[q * N + lane]) * wlv : 0.0);
            if (lane < M) {
                const float xa = float(xbar[q] + dxa[k] + dxw);
                bad |= !isfinite(xa);
                f[base + (long long)q * mp.nodes] = xa;
            }
        }
    }
    if (bad) atomicOr(status, 1);
}

template <int N, bool SWEEP64>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(3))) void k_mixed_finish(
    const float* __restrict__ ws_u, const float* __restrict__ ws_b, float* __restrict__ f, MixedParams mp,
    double* __restrict__ dump, int* __restrict__ status) {
    __shared__ FinishShared<N> sh;
    finish_body<N, SWEEP64>(blockIdx.x, gridDim.x, ws_u, ws_b, f, mp, dump, status, sh);
}

// ---------------------------------------------------------------------------------
// group launches: the analyses of several contexts (different grids and ensemble
// sizes) in one launch of each kernel.  Each task owns a range of workgroups,
// padded to a multiple of 8 so that its XCD-contiguous order holds; a workgroup
// finds its task by scanning the (at most MIXED_GROUP_MAX) range starts.
// ---------------------------------------------------------------------------------
template <int PASSES>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_mixed_solve_group(MixedGroup G) {
    int k = 0;
    while (k + 1 < G.n && int(blockIdx.x) >= G.t[k + 1].b0_solve) ++k;
    const MixedTask& T = G.t[k];
    const int bx = int(blockIdx.x) - T.b0_solve;
    __shared__ __align__(16) float cfull[64 * 64];  // one buffer for every width
    switch (T.N) {
        case 16: solve_body<16, PASSES>(bx, T.nb_solve, T.loc_ptr, T.loc_site, T.loc_w, T.dy, T.dyo, T.ws_u, T.ws_b, T.mp, cfull, T.dump); break;
        case 32: solve_body<32, PASSES>(bx, T.nb_solve, T.loc_ptr, T.loc_site, T.loc_w, T.dy, T.dyo, T.ws_u, T.ws_b, T.mp, cfull, T.dump); break;
        default: solve_body<64, PASSES>(bx, T.nb_solve, T.loc_ptr, T.loc_site, T.loc_w, T.dy, T.dyo, T.ws_u, T.ws_b, T.mp, cfull, T.dump); break;
    }
}

template <int PASSES>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_mixed_analysis_group(MixedGroup G) {
    int k = 0;
    while (k + 1 < G.n && int(blockIdx.x) >= G.t[k + 1].b0_solve) ++k;
    const MixedTask& T = G.t[k];
    const int bx = int(blockIdx.x) - T.b0_solve;
    // one buffer for every width (the layouts of smaller widths fit inside N = 64's)
    __shared__ __align__(16) FusedShared<64> sh;
    const int* lp = T.loc_ptr;
    const int* ls = T.loc_site;
    const float* lw = T.loc_w;
    const float* ty = T.dy;
    const float* to = T.dyo;
    float* tf = T.f;
    double* td = T.dump;
    int* ts = T.status;
    switch (T.N) {
        case 16:
            analysis_body<16, PASSES>(bx, T.nb_solve, lp, ls, lw, ty, to, tf, T.mp, td, ts,
                                      *reinterpret_cast<FusedShared<16>*>(&sh));
            break;
        case 32:
            analysis_body<32, PASSES>(bx, T.nb_solve, lp, ls, lw, ty, to, tf, T.mp, td, ts,
                                      *reinterpret_cast<FusedShared<32>*>(&sh));
            break;
        default: analysis_body<64, PASSES>(bx, T.nb_solve, lp, ls, lw, ty, to, tf, T.mp, td, ts, sh); break;
    }
}

template <bool SWEEP64>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(3))) void k_mixed_finish_group(MixedGroup G) {
    int k = 0;
    while (k + 1 < G.n && int(blockIdx.x) >= G.t[k + 1].b0_finish) ++k;
    const MixedTask& T = G.t[k];
    const int bx = int(blockIdx.x) - T.b0_finish;
    // one buffer for every width (the layouts of smaller widths fit inside N = 64's)
    __shared__ FinishShared<64> sh;
    switch (T.N) {
        case 16: finish_body<16, SWEEP64>(bx, T.nb_finish, T.ws_u, T.ws_b, T.f, T.mp, T.dump, T.status,
                                          *reinterpret_cast<FinishShared<16>*>(&sh)); break;
        case 32: finish_body<32, SWEEP64>(bx, T.nb_finish, T.ws_u, T.ws_b, T.f, T.mp, T.dump, T.status,
                                          *reinterpret_cast<FinishShared<32>*>(&sh)); break;
        default: finish_body<64, SWEEP64>(bx, T.nb_finish, T.ws_u, T.ws_b, T.f, T.mp, T.dump, T.status, sh); break;
    }
}

MixedParams mixed_params(const Ctx& c, unsigned long long* sweeps32, unsigned long long* sweeps64) {
    MixedParams mp{};
    mp.M = c.M;
    mp.nodes = int(c.nodes);
    mp.nx = c.nx;
    mp.ny = c.ny;
    mp.n_vars = c.obs.n_vars;
    mp.n_obs = c.obs.n_sites * c.obs.n_vars;
    mp.member_stride = c.lp.member_stride;
    mp.diag = float(double(c.M - 1) / c.cfg.inflation);
    for (int a = 0; a < 3; ++a) mp.coef[a] = c.obs.inv_sigma2[a];
    mp.sqm1 = sqrt(double(c.M - 1));
    // The FP32 phase runs to the FP32 floor (largest pair cosine ~1e-6): the chained
    // warm start takes the normalised columns of the previous gridpoint's U as its
    // basis, and any non-orthogonality left there becomes an error of the next
    // gridpoint's eigenpairs that no later sweep removes.
    mp.tol32 = c.cfg.jacobi_tol32 > 0 ? float(c.cfg.jacobi_tol32) : 3e-6f;
#ifndef LBMLETKF_CHECK_DIV
#define LBMLETKF_CHECK_DIV 3.f
#endif
    mp.check_tol = mp.tol32 / LBMLETKF_CHECK_DIV;
    mp.tol64 = c.cfg.jacobi_tol;
    mp.max32 = c.cfg.jacobi_max_sweeps;
    mp.max64 = c.cfg.jacobi_fp64_sweeps;
    mp.dump_g0 = c.dump_g0;
    mp.dump_n = c.dump_n;
    mp.sweeps32 = sweeps32;
    mp.sweeps64 = sweeps64;
    mp.chain = c.mix_chain;
    mp.write_state = 1;
    return mp;
}

}  // namespace

int mixed_width(int M) {
    int N = 16;
    while (N < M) N *= 2;
    return N;
}

bool mixed_fused(const Ctx& c) { return c.cfg.jacobi_fp64_sweeps <= 0; }

// workgroups of k_mixed_analysis: one per chain of `chain` rows of one column
long long mixed_fused_blocks(int nx, int ny, int chain) { return (long long)nx * ((ny + chain - 1) / chain); }

hipError_t launch_letkf_mixed(Ctx& c, float* f, unsigned long long* sweeps32, unsigned long long* sweeps64,
                              int passes, hipEvent_t ev_mid_end, hipEvent_t ev_mid_begin, bool solve_only) {
    const int N = mixed_width(c.M);
    if (!c.d_loc_ptr || !c.d_dy || !c.d_dyo || !f || (c.dump_n > 0 && !c.d_dump)) return hipErrorInvalidValue;
    if (c.mix_width != N || c.mix_chain < 1) return hipErrorInvalidValue;
    MixedParams mp = mixed_params(c, sweeps32, sweeps64);
    if (mixed_fused(c)) {
        // one kernel, no workspace
        const long long nb = mixed_fused_blocks(c.nx, c.ny, mp.chain);
        if (solve_only) {
            mp.write_state = 0;
            mp.dump_n = 0;
        }
        const dim3 grid_a{unsigned(nb), 1u, 1u};
#define LAUNCH_FUSED(NN, PP)                                                                                   \
    k_mixed_analysis<NN, PP><<<grid_a, dim3(64), 0, c.stream>>>(c.d_loc_ptr, c.d_loc_site, c.d_loc_w, c.d_dy,  \
                                                                c.d_dyo, f, mp, c.d_dump, c.d_status)
        switch (N * 10 + (passes == 3 ? 3 : 1)) {
            case 163: LAUNCH_FUSED(16, 3); break;
            case 161: LAUNCH_FUSED(16, 1); break;
            case 323: LAUNCH_FUSED(32, 3); break;
            case 321: LAUNCH_FUSED(32, 1); break;
            case 643: LAUNCH_FUSED(64, 3); break;
            case 641: LAUNCH_FUSED(64, 1); break;
            default: return hipErrorInvalidValue;
        }
#undef LAUNCH_FUSED
        return hipGetLastError();
    }
    if (!c.d_mix_u || !c.d_mix_b) return hipErrorInvalidValue;
    const dim3 grid(unsigned(c.nodes)), block(64), block2(128);
    const dim3 grid_solve(unsigned((c.nodes + mp.chain - 1) / mp.chain));
    const bool sweep64 = mp.max64 > 0;
#define LAUNCH_SOLVE(NN, PP)                                                                           \
    k_mixed_solve<NN, PP><<<grid_solve, block, 0, c.stream>>>(c.d_loc_ptr, c.d_loc_site, c.d_loc_w, c.d_dy, \
                                                         c.d_dyo, c.d_mix_u, c.d_mix_b, mp, \
                                                         solve_only ? nullptr : c.d_dump)
#define LAUNCH_FINISH(NN)                                                                                     \
    (sweep64 ? k_mixed_finish<NN, true><<<grid, block2, 0, c.stream>>>(c.d_mix_u, c.d_mix_b, f, mp, c.d_dump,  \
                                                                        c.d_status)                             \
             : k_mixed_finish<NN, false><<<grid, block2, 0, c.stream>>>(c.d_mix_u, c.d_mix_b, f, mp, c.d_dump, \
                                                                         c.d_status))
    switch (N * 10 + (passes == 3 ? 3 : 1)) {
        case 163: LAUNCH_SOLVE(16, 3); break;
        case 161: LAUNCH_SOLVE(16, 1); break;
        case 323: LAUNCH_SOLVE(32, 3); break;
        case 321: LAUNCH_SOLVE(32, 1); break;
        case 643: LAUNCH_SOLVE(64, 3); break;
        case 641: LAUNCH_SOLVE(64, 1); break;
        default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || solve_only) return e;
    for (hipEvent_t ev : {ev_mid_end, ev_mid_begin}) {
        if (!ev) continue;
        e = hipEventRecord(ev, c.stream);
        if (e != hipSuccess) return e;
    }
    switch (N) {
        case 16: LAUNCH_FINISH(16); break;
        case 32: LAUNCH_FINISH(32); break;
        case 64: LAUNCH_FINISH(64); break;
        default: return hipErrorInvalidValue;
    }
#undef LAUNCH_SOLVE
#undef LAUNCH_FINISH
    return hipGetLastError();
}

hipError_t launch_letkf_mixed_group(Ctx* const* cs, int n, hipStream_t stream) {
    if (n < 1 || n > MIXED_GROUP_MAX) return hipErrorInvalidValue;
    MixedGroup G{};
    G.n = n;
    const int passes = cs[0]->cfg.gram_passes;
    const bool sweep64 = cs[0]->cfg.jacobi_fp64_sweeps > 0;
    long long bs = 0, bf = 0;
    for (int k = 0; k < n; ++k) {
        Ctx& c = *cs[k];
        const int N = mixed_width(c.M);
        if (!c.d_loc_ptr || !c.d_dy || !c.d_dyo || c.mix_width != N || (c.dump_n > 0 && !c.d_dump))
            return hipErrorInvalidValue;
        if (sweep64 && (!c.d_mix_u || !c.d_mix_b)) return hipErrorInvalidValue;
        if (c.cfg.gram_passes != passes || (c.cfg.jacobi_fp64_sweeps > 0) != sweep64) return hipErrorInvalidValue;
        MixedTask& T = G.t[k];
        T.loc_ptr = c.d_loc_ptr;
        T.loc_site = c.d_loc_site;
        T.loc_w = c.d_loc_w;
        T.dy = c.d_dy;
        T.dyo = c.d_dyo;
        T.ws_u = c.d_mix_u;
        T.ws_b = c.d_mix_b;
        T.f = c.f[c.cur];
        T.dump = c.d_dump;
        T.status = c.d_status;
        T.mp = mixed_params(c, nullptr, nullptr);
        T.N = N;
        const long long chains = sweep64 ? (c.nodes + T.mp.chain - 1) / T.mp.chain
                                         : mixed_fused_blocks(c.nx, c.ny, T.mp.chain);
        T.nb_solve = int((chains + 7) / 8 * 8);
        T.nb_finish = int((c.nodes + 7) / 8 * 8);
        T.b0_solve = int(bs);
        T.b0_finish = int(bf);
        bs += T.nb_solve;
        bf += T.nb_finish;
    }
    if (bs > (1LL << 31) - 1 || bf > (1LL << 31) - 1) return hipErrorInvalidValue;
    if (!sweep64) {
        if (passes == 3)
            k_mixed_analysis_group<3><<<unsigned(bs), 64, 0, stream>>>(G);
        else
            k_mixed_analysis_group<1><<<unsigned(bs), 64, 0, stream>>>(G);
        return hipGetLastError();
    }
    if (passes == 3)
        k_mixed_solve_group<3><<<unsigned(bs), 64, 0, stream>>>(G);
    else
        k_mixed_solve_group<1><<<unsigned(bs), 64, 0, stream>>>(G);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (sweep64)
        k_mixed_finish_group<true><<<unsigned(bf), 128, 0, stream>>>(G);
    else
        k_mixed_finish_group<false><<<unsigned(bf), 128, 0, stream>>>(G);
    return hipGetLastError();
}

}  // namespace lbmletkf
