// Device-side pieces of the warm-start MLP (12 -> 64 -> 64 -> 1, exact-erf
// GELU) shared by the standalone network kernels (mlp.hip) and the fused
// warm-start Jacobi launch (particles.hip), so both evaluate a row with the
// same operations in the same order.
//
// Parameters: one float32 vector of 5057 values, W1[12][64] b1[64] W2[64][64]
// b2[64] W3[64] b3[1] (row-major, input x output).  The two 64-wide products
// run on v_mfma_f32_16x16x16_bf16: inputs, hidden activations and W1/W2 are
// rounded to bfloat16 (nearest even), products accumulate in float32; the
// output layer and every elementwise op stay in float32.
//
// A wavefront evaluates 16 rows at a time.  MFMA operand layout (16x16x16):
//   A (M x K):  lane l holds A[l % 16][4 (l / 16) + j], j = 0..3
//   B (K x N):  lane l holds B[4 (l / 16) + j][l % 16]
//   D (M x N):  lane l holds D[4 (l / 16) + j][l % 16]
#pragma once

#include <hip/hip_runtime.h>

namespace xpbd_mlp {

constexpr int NIN = 12, NH = 64;
constexpr int P_W1 = 0, P_B1 = P_W1 + NIN * NH, P_W2 = P_B1 + NH, P_B2 = P_W2 + NH * NH,
              P_W3 = P_B2 + NH, P_B3 = P_W3 + NH, NPARAM = P_B3 + 1;
static_assert(NPARAM == 5057, "parameter count");

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned short to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float from_bf16(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// The network in LDS: W1 padded to 16 input rows (12..15 zero), bfloat16.
struct Weights {
  unsigned short w1[16][NH];
  unsigned short w2[NH][NH];
  float b1[NH], b2[NH], w3[NH];
};

// Fills `s` from theta with threads [tid, ...) stepping by nthreads; the
// caller synchronises before use.
__device__ inline void load_weights(Weights& s, const float* __restrict__ theta, int tid,
                                    int nthreads) {
  for (int i = tid; i < 16 * NH; i += nthreads) {
    const int k = i / NH, n = i % NH;
    s.w1[k][n] = k < NIN ? to_bf16(theta[P_W1 + k * NH + n]) : (unsigned short)0;
  }
  for (int i = tid; i < NH * NH; i += nthreads) s.w2[i / NH][i % NH] = to_bf16(theta[P_W2 + i]);
  for (int i = tid; i < NH; i += nthreads) {
    s.b1[i] = theta[P_B1 + i];
    s.b2[i] = theta[P_B2 + i];
    s.w3[i] = theta[P_W3 + i];
  }
}

// Forward of one 16-row tile held by a wave.  feat(r, k) returns the
// bfloat16 bits of feature k (< 16; zero for k >= 12) of tile row r.
// `tile` is this wave's 16 x 64 LDS transposition tile.  On return z1/z2
// hold the pre-activations (D layout, 4 column tiles x 4 rows per lane), y
// the outputs of rows 4g+j (identical across the 16 lanes of group g) and
// xa the layer-1 A operand.
template <class Feat>
__device__ inline void forward16(const Weights& s, unsigned short (*tile)[NH], Feat feat, float b3,
                                 f32x4 z1[4], f32x4 z2[4], float y[4], bf16x4& xa) {
  const int l = threadIdx.x & 63, r16 = l & 15, g = l >> 4;
  for (int j = 0; j < 4; ++j) xa[j] = (short)feat(r16, 4 * g + j);
  for (int n = 0; n < 4; ++n) {
    bf16x4 b;
    for (int j = 0; j < 4; ++j) b[j] = (short)s.w1[4 * g + j][16 * n + r16];
    z1[n] = mfma(xa, b, f32x4{0.f, 0.f, 0.f, 0.f});
    for (int j = 0; j < 4; ++j) {
      z1[n][j] += s.b1[16 * n + r16];
      tile[4 * g + j][16 * n + r16] = to_bf16(gelu(z1[n][j]));
    }
  }
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  // layer 2: A = H1 tile read back in A layout
  bf16x4 a[4];
  for (int sK = 0; sK < 4; ++sK)
    for (int j = 0; j < 4; ++j) a[sK][j] = (short)tile[r16][16 * sK + 4 * g + j];
  for (int n = 0; n < 4; ++n) {
    f32x4 acc{0.f, 0.f, 0.f, 0.f};
    for (int sK = 0; sK < 4; ++sK) {
      bf16x4 b;
      for (int j = 0; j < 4; ++j) b[j] = (short)s.w2[16 * sK + 4 * g + j][16 * n + r16];
      acc = mfma(a[sK], b, acc);
    }
    for (int j = 0; j < 4; ++j) acc[j] += s.b2[16 * n + r16];
    z2[n] = acc;
  }
  // layer 3 (float32): y[row] = sum_c gelu(z2[row][c]) w3[c] + b3
  for (int j = 0; j < 4; ++j) {
    float p = 0.0f;
    for (int n = 0; n < 4; ++n) p += gelu(z2[n][j]) * s.w3[16 * n + r16];
    for (int o = 1; o < 16; o <<= 1) p += __shfl_xor(p, o, 64);
    y[j] = p + b3;
  }
  __builtin_amdgcn_wave_barrier();
}

}  // namespace xpbd_mlp
