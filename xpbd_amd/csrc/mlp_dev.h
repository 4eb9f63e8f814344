// Device-side pieces of the warm-start MLP (12 -> 64 -> 64 -> 1, GELU in
// its tanh form) shared by the standalone network kernels (mlp.hip) and the
// warm-start launch (particles.hip), so both evaluate a row with the same
// operations in the same order.
//
// Parameters: one float32 vector of 5057 values, W1[12][64] b1[64] W2[64][64]
// b2[64] W3[64] b3[1] (row-major, input x output).  Inputs, hidden
// activations and W1/W2 are rounded to bfloat16 (nearest even), products
// accumulate in float32 on the matrix cores; the output layer and every
// elementwise op stay in float32.  Contractions over the 64 hidden units
// (layer 2 and its backward) run on gfx950's v_mfma_f32_16x16x32_bf16, two
// per 16 x 16 tile; those over 16 (the 12 features padded, or the 16 rows of
// a tile in the weight gradients) on v_mfma_f32_16x16x16_bf16.
//
// A wavefront evaluates 16 rows at a time.  MFMA operand layouts (lane l,
// r = l % 16, g = l / 16; checked against a CPU product on the device):
//   16x16x16:  A[r][4g + j], B[4g + j][r], j = 0..3
//   16x16x32:  A[r][8g + j], B[8g + j][r], j = 0..7
//   D (both):  D[4g + j][r], j = 0..3
// In a 32-wide contraction, K slice m (m = 0, 1) puts hidden unit
// 32m + 4g + j at position 8g + j and 32m + 16 + 4g + j at 8g + 4 + j, so
// the two 16x16 D tiles 2m and 2m + 1 of the previous layer are, lane for
// lane, the B operand (h[2m], h[2m + 1]) with no data movement; the weight
// operands are read in the same order.
//
// Two evaluation orders share these weights: forward16 (rows as M, used by
// the training kernels, whose backward pass works in that layout) and
// forward16_t (the transposed products Z^T = W^T X^T, used for inference:
// a layer's output tile lands in the lanes exactly as the next layer's B
// operand, so no LDS transposition; the weight operands stay in registers).
#pragma once

#include <hip/hip_runtime.h>

namespace xpbd_mlp {

constexpr int NIN = 12, NH = 64;
constexpr int P_W1 = 0, P_B1 = P_W1 + NIN * NH, P_W2 = P_B1 + NH, P_B2 = P_W2 + NH * NH,
              P_W3 = P_B2 + NH, P_B3 = P_W3 + NH, NPARAM = P_B3 + 1;
static_assert(NPARAM == 5057, "parameter count");

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16, round to nearest even (v_cvt_pk_bf16_f32 on gfx950;
// for finite values the same bits as adding 0x7FFF + lsb to the pattern)
__device__ inline unsigned short to_bf16(float f) {
  return __builtin_bit_cast(unsigned short, (__bf16)f);
}
__device__ inline float from_bf16(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

// GELU, tanh form (Hendrycks & Gimpel 2016): 0.5 z (1 + tanh(sqrt(2/pi) (z +
// 0.044715 z^3))) = z sigmoid(u), u = 2 sqrt(2/pi) z (1 + 0.044715 z^2);
// sigmoid from the hardware exp2 and reciprocal (v_exp_f32, v_rcp_f32) of
// -u log2(e) = z (k0 + k1 z^2) with the constants folded: five VALU
// operations and two transcendentals instead of the ~40 of an erf.
constexpr float kGeluC0 = 1.5957691216057308f;  // 2 sqrt(2 / pi)
constexpr float kGeluC1 = 0.044715f;
constexpr float kLog2e = 1.4426950408889634f;
constexpr float kGeluK0 = (float)(-1.4426950408889634 * 1.5957691216057308);
constexpr float kGeluK1 = (float)(-1.4426950408889634 * 1.5957691216057308 * 0.044715);
__device__ inline float gelu_sigmoid(float z) {
  const float v = z * fmaf(kGeluK1, z * z, kGeluK0);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(v));
}
__device__ inline float gelu(float z) { return z * gelu_sigmoid(z); }
__device__ inline float gelu_grad(float z) {
  const float s = gelu_sigmoid(z);
  return s + z * s * (1.0f - s) * (kGeluC0 * fmaf(3.0f * kGeluC1, z * z, 1.0f));
}

__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Two bfloat16 quadruples as one 8-element operand (positions 8g .. 8g+3,
// then 8g+4 .. 8g+7 of a 32-wide K slice): one shuffle of the 16-bit
// patterns and one bit cast of the whole vector.  (Inserting bit-cast
// __bf16 values element by element gave wrong operands with this compiler;
// checked against a CPU product on the device.)
typedef short s16x8 __attribute__((ext_vector_type(8)));
__device__ inline bf16x8 cat8(bf16x4 lo, bf16x4 hi) {
  const s16x8 t = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_bit_cast(bf16x8, t);
}

__device__ inline f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// The network in LDS, bfloat16 weights stored transposed (output-major:
// w1t[n][k] = W1[k][n] with k padded to 16, w2t[n][k] = W2[k][n]) so that a
// lane's four B-operand values are contiguous (one 8-byte LDS read).
struct alignas(16) Weights {
  unsigned short w1t[NH][16];
  unsigned short w2t[NH][NH];
  float b1[NH], b2[NH], w3[NH];
};
static_assert(sizeof(Weights) % 16 == 0, "weights copied in 16-byte pieces");

// The network as one packed image (Weights followed by b3), built once from
// theta by xpbd_mlp_pack and copied into LDS with 16-byte loads.
struct alignas(16) Image {
  Weights w;
  float b3, pad[3];
};
static_assert(sizeof(Image) % 16 == 0, "image copied in 16-byte pieces");

// Fills `s` from theta with threads [tid, ...) stepping by nthreads; the
// caller synchronises before use.
__device__ inline void load_weights(Weights& s, const float* __restrict__ theta, int tid,
                                    int nthreads) {
  for (int i = tid; i < 16 * NH; i += nthreads) {
    const int k = i / NH, n = i % NH;
    s.w1t[n][k] = k < NIN ? to_bf16(theta[P_W1 + k * NH + n]) : (unsigned short)0;
  }
  for (int i = tid; i < NH * NH; i += nthreads) s.w2t[i % NH][i / NH] = to_bf16(theta[P_W2 + i]);
  for (int i = tid; i < NH; i += nthreads) {
    s.b1[i] = theta[P_B1 + i];
    s.b2[i] = theta[P_B2 + i];
    s.w3[i] = theta[P_W3 + i];
  }
}

// Copies a packed image's weights into `s` (16 bytes per load).
__device__ inline void load_image(Weights& s, const Image* __restrict__ img, int tid,
                                  int nthreads) {
  const uint4* src = (const uint4*)img;
  uint4* dst = (uint4*)&s;
  for (int i = tid; i < (int)(sizeof(Weights) / 16); i += nthreads) dst[i] = src[i];
}

// Forward of one 16-row tile held by a wave.  feat(r, k) returns the
// bfloat16 bits of feature k (< 16; zero for k >= 12) of tile row r.
// `tile` is this wave's 16 x 64 LDS transposition tile (8-byte aligned).  On return z1/z2
// hold the pre-activations (D layout, 4 column tiles x 4 rows per lane), y
// the outputs of rows 4g+j (identical across the 16 lanes of group g) and
// xa the layer-1 A operand.
template <class Feat>
__device__ inline void forward16(const Weights& s, unsigned short (*tile)[NH], Feat feat, float b3,
                                 f32x4 z1[4], f32x4 z2[4], float y[4], bf16x4& xa) {
  const int l = threadIdx.x & 63, r16 = l & 15, g = l >> 4;
  for (int j = 0; j < 4; ++j) xa[j] = (short)feat(r16, 4 * g + j);
  for (int n = 0; n < 4; ++n) {
    const bf16x4 b = *(const bf16x4*)&s.w1t[16 * n + r16][4 * g];
    z1[n] = mfma(xa, b, f32x4{0.f, 0.f, 0.f, 0.f});
    for (int j = 0; j < 4; ++j) {
      z1[n][j] += s.b1[16 * n + r16];
      tile[4 * g + j][16 * n + r16] = to_bf16(gelu(z1[n][j]));
    }
  }
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  // layer 2: A = H1 tile read back in A layout, two 32-wide K slices
  bf16x8 a[2];
  for (int m = 0; m < 2; ++m)
    a[m] = cat8(*(const bf16x4*)&tile[r16][32 * m + 4 * g],
                *(const bf16x4*)&tile[r16][32 * m + 16 + 4 * g]);
  for (int n = 0; n < 4; ++n) {
    f32x4 acc{0.f, 0.f, 0.f, 0.f};
    for (int m = 0; m < 2; ++m) {
      const bf16x8 b = cat8(*(const bf16x4*)&s.w2t[16 * n + r16][32 * m + 4 * g],
                            *(const bf16x4*)&s.w2t[16 * n + r16][32 * m + 16 + 4 * g]);
      acc = mfma32(a[m], b, acc);
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

// Inference weights held in registers for a wave's whole life: the W1^T and
// W2^T tiles as MFMA A operands (40 VGPRs).  Lane (r, g) = (l % 16, l / 16)
// holds a1[n][j] = W1[4g + j][16n + r] and a2[n][m] = (W2[32m + 4g + j][16n + r],
// W2[32m + 16 + 4g + j][16n + r]), j = 0..3.
struct RegWeights {
  bf16x4 a1[4];
  bf16x8 a2[4][2];
};

__device__ inline void load_reg_weights(RegWeights& w, const Weights& s) {
  const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
  for (int n = 0; n < 4; ++n) {
    w.a1[n] = *(const bf16x4*)&s.w1t[16 * n + r][4 * g];
    for (int m = 0; m < 2; ++m)
      w.a2[n][m] = cat8(*(const bf16x4*)&s.w2t[16 * n + r][32 * m + 4 * g],
                        *(const bf16x4*)&s.w2t[16 * n + r][32 * m + 16 + 4 * g]);
  }
}

// Forward of one 16-row tile in the transposed order.  xb is the layer-1 B
// operand: lane (r, g) holds the bfloat16 features 4g .. 4g+3 of tile row r
// (zero beyond feature 11).  b1, b2, w3 point at the float vectors (LDS).
// Returns y of row r in the four lanes (r, 0..3): layer 3 accumulates, per
// lane, gelu(z2) w3 over its 16 hidden units with fused multiply-adds (n
// outer, j inner), then sums the four lanes of the row ((g0 + g1) + (g2 +
// g3)) and adds b3.
__device__ inline float forward16_t(const RegWeights& w, const float* b1, const float* b2,
                                   const float* w3, bf16x4 xb, float b3) {
  const int l = threadIdx.x & 63, g = l >> 4;
  bf16x4 h[4];
  for (int n = 0; n < 4; ++n) {  // Z1^T tile n: lane holds Z1[r][16n + 4g + j]
    // the bias enters as the accumulator's start value
    const float4 b = *(const float4*)&b1[16 * n + 4 * g];
    const f32x4 z = mfma(w.a1[n], xb, f32x4{b.x, b.y, b.z, b.w});
    h[n][0] = (short)to_bf16(gelu(z[0]));
    h[n][1] = (short)to_bf16(gelu(z[1]));
    h[n][2] = (short)to_bf16(gelu(z[2]));
    h[n][3] = (short)to_bf16(gelu(z[3]));
  }
  float p = 0.0f;
  const bf16x8 hb[2] = {cat8(h[0], h[1]), cat8(h[2], h[3])};  // B slices of H1^T
  for (int n = 0; n < 4; ++n) {  // Z2^T tile n = b2 + sum_m W2^T(n, m) H1^T(m)
    const float4 b = *(const float4*)&b2[16 * n + 4 * g];
    f32x4 acc{b.x, b.y, b.z, b.w};
    for (int m = 0; m < 2; ++m) acc = mfma32(w.a2[n][m], hb[m], acc);
    const float4 c = *(const float4*)&w3[16 * n + 4 * g];
    p = fmaf(gelu(acc[0]), c.x, p);
    p = fmaf(gelu(acc[1]), c.y, p);
    p = fmaf(gelu(acc[2]), c.z, p);
    p = fmaf(gelu(acc[3]), c.w, p);
  }
  p += __shfl_xor(p, 16, 64);
  p += __shfl_xor(p, 32, 64);
  return p + b3;
}

}  // namespace xpbd_mlp
