// Learned lambda warm start for the 3-DOF solver (no reference
// counterpart; it replaces the zero reset of solver.py:431 / Alg. 1 line 4).
//
// At the start of a substep (after integrate), for every row j:
//   features F[j][12]  one-hot type (4), slog(C/l), log10(w)/4,
//                      alpha~/(w + alpha~), slog(lambda_prev (w + alpha~)/l)
//                      (zero when LAM is NULL),
//                      slog(v_n h / l), slog(|v_t| h / l), slog(pen / l),
//                      log10(1 + alpha~)/8
//   with w = w_a + w_b (|grad C|^2 weighted by inverse mass), alpha~ = alpha/h^2,
//   l = 1 mm and slog(x) = sign(x) ln(1 + |x|).
// The MLP predicts y = logit(omega / 8), omega = beta lambda* / lambda1
// (lambda* = the multiplier after many plain iterations, beta = 0.5, lambda1
// = -C / (w + alpha~) the first plain step from zero); the warm start sets
// lambda0 = 8 sigmoid(y) lambda1 (zero for separated one-sided rows) and
// moves every particle by the average over its rows of w grad C lambda0
// (times the relaxation), the position the Jacobi sweep itself reaches when
// its multipliers add up to lambda0: Jacobi applies the mean, not the sum,
// of a particle's corrections (kernels.py:326-339), so the summed offset
// x~ + M^-1 grad C^T lambda0 of Eq. 18 would overshoot by the row count.
#include <cmath>

#include "common.h"
#include "xpbd.h"

namespace {

constexpr int kB = 256;
constexpr float kLen = 1.0e-3f;
constexpr float kOmegaMax = 8.0f;  // largest multiple of the first plain step

__device__ inline float slog(float x) { return copysignf(log1pf(fabsf(x)), x); }

struct RowGeo {
  int t, a, b;
  float C, w, alpha, vn, vt, pen;
  f3 n;
  bool ok;
};

// Constraint value, gradient direction and relative velocity of one row at
// the current positions (kernels.py:131-208 conventions).
__device__ inline RowGeo row_geo(const float4* __restrict__ X, const float4* __restrict__ V,
                                 const float4* __restrict__ ANC, int4 row, int j, float inv_dt2) {
  RowGeo g;
  g.t = row_type(row);
  g.a = row_a(row);
  g.b = row.y;
  const float p0 = __int_as_float(row.z), p1 = __int_as_float(row.w);
  const float4 xa = X[g.a];
  const f3 va = xyz(V[g.a]);
  g.alpha = 0.0f;
  g.pen = 0.0f;
  g.ok = true;
  f3 vrel = va;
  if (g.t == T_FLOOR) {
    g.n = mk3(0.0f, 1.0f, 0.0f);
    g.C = xa.y - p0;
    g.w = xa.w;
    g.pen = fmaxf(0.0f, -g.C);
  } else {
    f3 pb;
    float wb = 0.0f;
    if (g.t == T_ATTACH) {
      pb = xyz(ANC[j]);
    } else {
      const float4 xb = X[g.b];
      pb = xyz(xb);
      wb = xb.w;
      vrel = va - xyz(V[g.b]);
    }
    const f3 d = xyz(xa) - pb;
    const float L = length(d);
    g.ok = L >= XPBD_EPS;
    g.n = g.ok ? d / L : mk3(0.0f, 1.0f, 0.0f);
    g.C = L - p0;
    g.w = xa.w + wb;
    if (g.t == T_CONTACT) g.pen = fmaxf(0.0f, -g.C);
    else g.alpha = p1 * inv_dt2;
  }
  g.vn = dot(vrel, g.n);
  g.vt = length(vrel - g.n * g.vn);
  return g;
}

__global__ __launch_bounds__(kB) void k_ws_features(const float4* __restrict__ X,
                                                    const float4* __restrict__ V,
                                                    const int4* __restrict__ R,
                                                    const float4* __restrict__ ANC,
                                                    const float* __restrict__ LAM, float h,
                                                    float* __restrict__ F, int m) {
  const int j = blockIdx.x * kB + threadIdx.x;
  if (j >= m) return;
  const RowGeo g = row_geo(X, V, ANC, R[j], j, 1.0f / (h * h));
  const float s = g.w + g.alpha;
  float f[12];
  for (int k = 0; k < 4; ++k) f[k] = (g.t == k) ? 1.0f : 0.0f;
  f[4] = slog(g.C / kLen);
  f[5] = log10f(fmaxf(g.w, 1e-30f)) * 0.25f;
  f[6] = s > 0.0f ? g.alpha / s : 0.0f;
  f[7] = LAM ? slog(LAM[j] * s / kLen) : 0.0f;
  f[8] = slog(g.vn * h / kLen);
  f[9] = slog(g.vt * h / kLen);
  f[10] = slog(g.pen / kLen);
  f[11] = log10f(1.0f + g.alpha) * 0.125f;
  float4* o = (float4*)(F + (size_t)j * 12);
  o[0] = make_float4(f[0], f[1], f[2], f[3]);
  o[1] = make_float4(f[4], f[5], f[6], f[7]);
  o[2] = make_float4(f[8], f[9], f[10], f[11]);
}

// y -> lambda0 per row, written to LAM; w grad C lambda0 and a count of one
// per endpoint summed into ACC (averaged by the particle pass's apply).
__global__ __launch_bounds__(kB) void k_ws_apply(const float4* __restrict__ X,
                                                 const float4* __restrict__ V,
                                                 const int4* __restrict__ R,
                                                 const float4* __restrict__ ANC,
                                                 const float* __restrict__ Y, float h,
                                                 float* __restrict__ LAM,
                                                 float4* __restrict__ ACC, int m) {
  const int j = blockIdx.x * kB + threadIdx.x;
  if (j >= m) return;
  const int4 row = R[j];
  const RowGeo g = row_geo(X, V, ANC, row, j, 1.0f / (h * h));
  const float s = g.w + g.alpha;
  // lambda0 = omega * (the first plain step from zero, -C / (w + alpha~)),
  // omega = kOmegaMax * sigmoid(y): bounded by the row's own violation now,
  // so nothing from earlier substeps is fed back.  One-sided rows push only
  // while violated (the projection leaves a separated row alone).
  float step = (s > 0.0f && g.ok) ? -g.C / s : 0.0f;
  if (g.t == T_FLOOR || g.t == T_CONTACT) step = g.C < 0.0f ? step : 0.0f;
  const float omega = kOmegaMax / (1.0f + expf(-Y[j]));
  float lam0 = omega * step;
  if (!isfinite(lam0)) lam0 = 0.0f;
  LAM[j] = lam0;
  if (lam0 == 0.0f) return;
  const float wa = X[g.a].w;
  const f3 ca = g.n * (wa * lam0);
  atomicAdd(&ACC[g.a].x, ca.x);
  atomicAdd(&ACC[g.a].y, ca.y);
  atomicAdd(&ACC[g.a].z, ca.z);
  atomicAdd(&ACC[g.a].w, 1.0f);
  if (g.t == T_DISTANCE || g.t == T_CONTACT) {
    const float wb = X[g.b].w;
    const f3 cb = g.n * (-wb * lam0);
    atomicAdd(&ACC[g.b].x, cb.x);
    atomicAdd(&ACC[g.b].y, cb.y);
    atomicAdd(&ACC[g.b].z, cb.z);
    atomicAdd(&ACC[g.b].w, 1.0f);
  }
}

// Residual: sum over rows of r^2, r = C + alpha~ lambda (two-sided rows),
// max(0, -C) (one-sided rows).
__global__ __launch_bounds__(kB) void k_ws_residual(const float4* __restrict__ X,
                                                    const int4* __restrict__ R,
                                                    const float4* __restrict__ ANC,
                                                    const float* __restrict__ LAM, float h,
                                                    double* __restrict__ out, int m) {
  __shared__ double sh[kB];
  const int j = blockIdx.x * kB + threadIdx.x;
  double r2 = 0.0;
  if (j < m) {
    const int4 row = R[j];
    const RowGeo g = row_geo(X, X, ANC, row, j, 1.0f / (h * h));
    float r;
    if (g.t == T_FLOOR || g.t == T_CONTACT) r = fmaxf(0.0f, -g.C);
    else r = g.C + g.alpha * LAM[j];
    r2 = (double)r * (double)r;
  }
  sh[threadIdx.x] = r2;
  __syncthreads();
  for (int s = kB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, sh[0]);
}

inline int nblk(int n) { return (n + kB - 1) / kB; }

}  // namespace

extern "C" {

int xpbd_ws_features(void* stream, const float* X, const float* V, const int* R, const float* ANC,
                     const float* LAM, float h, float* F, int n_rows) {
  if (n_rows <= 0) return 0;
  hipLaunchKernelGGL(k_ws_features, dim3(nblk(n_rows)), dim3(kB), 0, (hipStream_t)stream,
                     (const float4*)X, (const float4*)V, (const int4*)R, (const float4*)ANC, LAM,
                     h, F, n_rows);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd_ws_apply(void* stream, const float* X, const float* V, const int* R, const float* ANC,
                  const float* Y, float h, float* LAM, float* ACC, int n_rows) {
  if (n_rows <= 0) return 0;
  hipLaunchKernelGGL(k_ws_apply, dim3(nblk(n_rows)), dim3(kB), 0, (hipStream_t)stream,
                     (const float4*)X, (const float4*)V, (const int4*)R, (const float4*)ANC, Y, h,
                     LAM, (float4*)ACC, n_rows);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd_ws_residual(void* stream, const float* X, const int* R, const float* ANC,
                     const float* LAM, float h, double* out_sum, int n_rows) {
  if (n_rows <= 0) return 0;
  hipLaunchKernelGGL(k_ws_residual, dim3(nblk(n_rows)), dim3(kB), 0, (hipStream_t)stream,
                     (const float4*)X, (const int4*)R, (const float4*)ANC, LAM, h, out_sum,
                     n_rows);
  XPBD_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
