// Learned lambda warm start for the 3-DOF solver (no reference
// counterpart; it replaces the zero reset of solver.py:431 / Alg. 1 line 4):
// the row features of the training data and the residual the warm start is
// judged by.  The start itself runs inside the first Jacobi launch of a
// substep (k_jacobi_tile_warm, particles.hip): lambda0 from the network, then
// that iteration's update continues from lambda0 and moves the particles by
// w grad C (lambda0 + dlambda).  Features and lambda0: ws_dev.h.
#include <cmath>

#include "common.h"
#include "xpbd.h"
#include "ws_dev.h"

namespace {

using namespace xpbd_ws;

constexpr int kB = 256;

struct RowGeo {
  int t, a, b;
  float C, w, alpha, vn, vt, pen;
  f3 n;
  bool ok;
};

// Constraint value, gradient direction and relative velocity (x - x_prev) / h
// of one row at the current positions (kernels.py:131-208 conventions).
__device__ inline RowGeo row_geo(const float4* __restrict__ X, const float4* __restrict__ XP,
                                 const float4* __restrict__ ANC, int4 row, int j, float h) {
  RowGeo g;
  g.t = row_type(row);
  g.a = row_a(row);
  g.b = row.y;
  const float p0 = __int_as_float(row.z), p1 = __int_as_float(row.w);
  const float4 xa = X[g.a];
  const float inv_h = 1.0f / h;
  auto vel = [&](int i, float4 x4) { return (xyz(x4) - xyz(XP[i])) * inv_h; };
  g.alpha = 0.0f;
  g.pen = 0.0f;
  g.ok = true;
  f3 vrel = XP ? vel(g.a, xa) : mk3(0.0f, 0.0f, 0.0f);
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
      if (XP) vrel = vrel - vel(g.b, xb);
    }
    const f3 d = xyz(xa) - pb;
    const float L = length(d);
    g.ok = L >= XPBD_EPS;
    g.n = g.ok ? d / L : mk3(0.0f, 1.0f, 0.0f);
    g.C = L - p0;
    g.w = xa.w + wb;
    if (g.t == T_CONTACT) g.pen = fmaxf(0.0f, -g.C);
    else g.alpha = p1 * (1.0f / (h * h));  // as the projection computes alpha~
  }
  g.vn = dot(vrel, g.n);
  g.vt = length(vrel - g.n * g.vn);
  return g;
}

__global__ __launch_bounds__(kB) void k_ws_features(const float4* __restrict__ X,
                                                    const float4* __restrict__ XP,
                                                    const int4* __restrict__ R,
                                                    const float4* __restrict__ ANC, float h,
                                                    float* __restrict__ F, int m) {
  const int j = blockIdx.x * kB + threadIdx.x;
  if (j >= m) return;
  const RowGeo g = row_geo(X, XP, ANC, R[j], j, h);
  float f[12];
  row_features(g.t, g.C, g.w, g.alpha, g.vn, g.vt, g.pen, h, f);
  float4* o = (float4*)(F + (size_t)j * 12);
  o[0] = make_float4(f[0], f[1], f[2], f[3]);
  o[1] = make_float4(f[4], f[5], f[6], f[7]);
  o[2] = make_float4(f[8], f[9], f[10], f[11]);
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
    const RowGeo g = row_geo(X, nullptr, ANC, row, j, h);
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

int xpbd_ws_features(void* stream, const float* X, const float* XP, const int* R,
                     const float* ANC, float h, float* F, int n_rows) {
  if (n_rows <= 0) return 0;
  hipLaunchKernelGGL(k_ws_features, dim3(nblk(n_rows)), dim3(kB), 0, (hipStream_t)stream,
                     (const float4*)X, (const float4*)XP, (const int4*)R, (const float4*)ANC, h, F,
                     n_rows);
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
