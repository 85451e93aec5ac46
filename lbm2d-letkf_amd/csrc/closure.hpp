// Learned eddy-viscosity closure: the network's parameter layout and the
// device-side weight tiles the fused LBM kernel reads.
//
// Network (18 -> 64 -> 64 -> 1):
//   h1 = gelu(W1 x + b1), h2 = gelu(W2 h1 + b2), z = w3 . h2 + b3,
//   nu_t = nu_scale * softplus(z),  gelu(x) = x/2 (1 + tanh(k (x + 0.044715 x^3))).
// x: at the 3x3 stencil s = (dj+1)*3 + (di+1) around the node, the deviatoric
// strain a = (ux - vy)/2 and the shear b = (uy + vx)/2 of the streamed velocity
// (central differences as the Smagorinsky model's), divided by strain_scale:
// x[2s] = a, x[2s+1] = b.
// Flat FP32 parameter vector: W1[64][18], b1[64], W2[64][64], b2[64], w3[64], b3.
#pragma once
#include <cstddef>

namespace lbmletkf {

constexpr int CL_IN = 18, CL_H = 64;
constexpr int CL_W1 = 0, CL_B1 = CL_W1 + CL_H * CL_IN, CL_W2 = CL_B1 + CL_H, CL_B2 = CL_W2 + CL_H * CL_H,
              CL_W3 = CL_B2 + CL_H, CL_B3 = CL_W3 + CL_H, CL_NPARAM = CL_B3 + 1;
constexpr int CL_SAMPLE = 22;  // 18 features, a, b (unscaled), t1 = (tau11 - tau22)/2, t2 = tau12

// Weights of the fused kernel, as MFMA operand tiles of v_mfma_f32_16x16x32_bf16
// computing H^T = W X^T (rows = hidden units, columns = 16 nodes):
//   a1[I][lane]    : W1 rows 16I + lane%16, inputs 8(lane/16)..+7 (inputs >= 18 are 0)
//   a2[I][K][lane] : W2' rows 16I + lane%16, k' = 32K + 8(lane/16)..+7, where W2' is W2
//                    with its columns permuted so that the layer-1 accumulators of a
//                    lane (hidden 16J + 4(lane/16) + r, J = 2K, 2K+1, r = 0..3) are its
//                    B operand directly: k' = 32K + 8 kq + t  <->  hidden
//                    16 (2K + t/4) + 4 kq + t%4
//   b1v/b2v/w3v[J][lane][r] : b1, b2, w3 at hidden 16J + 4(lane/16) + r (FP32)
struct ClosureWeights {
    unsigned short a1[4][64][8];
    unsigned short a2[4][2][64][8];
    float b1v[4][64][4];
    float b2v[4][64][4];
    float w3v[4][64][4];
    float b3;
    float inv_strain_scale, nu_scale, nu_max;
};
// the collide kernel copies b1v, b2v, w3v to LDS as one block of 16-byte vectors
static_assert(offsetof(ClosureWeights, b1v) % 16 == 0, "b1v alignment");
static_assert(offsetof(ClosureWeights, b2v) == offsetof(ClosureWeights, b1v) + sizeof(float) * 4 * 64 * 4, "b2v");
static_assert(offsetof(ClosureWeights, w3v) == offsetof(ClosureWeights, b2v) + sizeof(float) * 4 * 64 * 4, "w3v");

}  // namespace lbmletkf
