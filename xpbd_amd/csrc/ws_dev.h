// Row features and the bounded start of the learned lambda warm start,
// shared by the standalone feature kernel (warmstart.hip) and the fused
// warm-start Jacobi launch (particles.hip).  No reference counterpart.
//
// Features of a row (12, in this order): one-hot type (4), slog(C/l),
// log10(w)/4, alpha~/(w + alpha~), 0 (slot of a previous-substep lambda,
// unused), slog(v_n h / l), slog(|v_t| h / l), slog(pen / l),
// log10(1 + alpha~)/8, with w = w_a + w_b (|grad C|^2 weighted by inverse
// mass), alpha~ = alpha / h^2, l = 1 mm, slog(x) = sign(x) ln(1 + |x|) and
// the relative velocity v = (x - x_prev) / h of the predicted positions.
//
// Start: lambda0 = 4 sigmoid(y) * (-C / (w + alpha~)), a multiple (0..4) of
// the row's first plain step (one-sided rows only while violated).
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"

namespace xpbd_ws {

constexpr float kLen = 1.0e-3f;
constexpr float kOmegaMax = 4.0f;

// Logarithms from the hardware log2 (v_log_f32, inputs >= 1e-30 here, all
// normal): features are rounded to bfloat16 before use, far coarser than
// the ~1e-7 relative error of log2(1 + |x|) ln 2.
constexpr float kLn2 = 0.69314718055994531f, kLog10of2 = 0.30102999566398120f;
__device__ inline float slog(float x) {
  return copysignf(__builtin_amdgcn_logf(1.0f + fabsf(x)) * kLn2, x);
}

__device__ inline void row_features(int t, float C, float w, float alpha, float vn, float vt,
                                    float pen, float h, float f[12]) {
  const float s = w + alpha;
  for (int k = 0; k < 4; ++k) f[k] = (t == k) ? 1.0f : 0.0f;
  f[4] = slog(C / kLen);
  f[5] = __builtin_amdgcn_logf(fmaxf(w, 1e-30f)) * (kLog10of2 * 0.25f);
  f[6] = s > 0.0f ? alpha / s : 0.0f;
  f[7] = 0.0f;
  f[8] = slog(vn * h / kLen);
  f[9] = slog(vt * h / kLen);
  f[10] = slog(pen / kLen);
  f[11] = __builtin_amdgcn_logf(1.0f + alpha) * (kLog10of2 * 0.125f);
}

// Whether the learned start is in use for a launch: ws_on (device flag of
// the frame plan's judge; NULL = always) and `invert` (the judge runs the
// start the solver is not using).
// The network's output per row is kept as bfloat16 bits (round to nearest
// even): the start is a learned estimate, and the per-substep launches read
// it for every row.
__device__ inline unsigned short y_bits(float y) {
  return __builtin_bit_cast(unsigned short, (__bf16)y);
}
__device__ inline float y_of(const unsigned short* Y, int j) {
  return __uint_as_float((unsigned)Y[j] << 16);
}

__device__ inline bool start_active(const int* ws_on, int invert) {
  const bool on = ws_on ? (*ws_on != 0) : true;
  return on != (invert != 0);
}

// lambda0 of a row from the network output y; `step` is the row's first
// plain step -C / (w + alpha~) (zero when the row is inactive).
__device__ inline float start_lambda(float y, float step) {
  constexpr float kLog2e = 1.4426950408889634f;
  const float s = __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(y * -kLog2e));
  const float lam0 = kOmegaMax * s * step;
  return isfinite(lam0) ? lam0 : 0.0f;
}

}  // namespace xpbd_ws
