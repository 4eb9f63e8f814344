// Learned lambda warm start for the 3-DOF solver (no reference
// counterpart; it replaces the zero reset of solver.py:431 / Alg. 1 line 4):
// the row features, the network pass of a substep (features + MLP -> y per
// row) and the residual the warm start is measured by.  The start itself is
// applied by the warm-start tile launch (k_jacobi_tile_warm, particles.hip)
// from y.  Features and lambda0: ws_dev.h; the network: mlp_dev.h.
#include <cmath>

#include "common.h"
#include "mlp_dev.h"
#include "ws_dev.h"
#include "xpbd.h"

namespace {

using namespace xpbd_ws;

constexpr int kB = 256;

struct RowGeo {
  int t, a, b;
  float C, w, alpha, vn, vt, pen;
  f3 n;
  bool ok;
};

// Constraint value, gradient direction and relative velocity
// (x - x_prev) / h of one row (kernels.py:131-208 conventions); pos(i) and
// prev(i) give a particle's position (x, y, z, inverse mass) and x_prev.
template <bool VEL, class Pos, class Prev>
__device__ inline RowGeo row_geo(Pos pos, Prev prev, const float4* __restrict__ ANC, int4 row,
                                 int j, float h) {
  RowGeo g;
  g.t = row_type(row);
  g.a = row_a(row);
  g.b = row.y;
  const float p0 = __int_as_float(row.z), p1 = __int_as_float(row.w);
  const float4 xa = pos(g.a);
  const float inv_h = 1.0f / h;
  g.alpha = 0.0f;
  g.pen = 0.0f;
  g.ok = true;
  f3 vrel = VEL ? (xyz(xa) - prev(g.a)) * inv_h : mk3(0.0f, 0.0f, 0.0f);
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
      const float4 xb = pos(g.b);
      pb = xyz(xb);
      wb = xb.w;
      if (VEL) vrel = vrel - (xyz(xb) - prev(g.b)) * inv_h;
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
  const RowGeo g = row_geo<true>([&](int i) { return X[i]; }, [&](int i) { return xyz(XP[i]); },
                                 ANC, R[j], j, h);
  float f[12];
  row_features(g.t, g.C, g.w, g.alpha, g.vn, g.vt, g.pen, h, f);
  float4* o = (float4*)(F + (size_t)j * 12);
  o[0] = make_float4(f[0], f[1], f[2], f[3]);
  o[1] = make_float4(f[4], f[5], f[6], f[7]);
  o[2] = make_float4(f[8], f[9], f[10], f[11]);
}

// Network pass of a substep: every wave takes ROWS consecutive rows; lanes
// 0 .. ROWS-1 compute their features (as k_ws_features, at the predicted
// positions when PREDICT: x~ = x + (v + g h) h from Xin and V, x_prev = Xin,
// the arithmetic of integrate) into LDS as bfloat16, then the wave runs
// forward16_t on each 16-row tile and writes y per row.  The weight operands
// live in registers (loaded once per wave from the packed image), the
// biases in LDS; waves loop over row groups.  ROWS = 64 (one feature pass
// of all 64 lanes per four tiles) unless the row set is too small to give
// every CU a wave, then 16.
constexpr int kNetWaves = 4, kNetBlock = 64 * kNetWaves, kNetMaxBlocks = 2048;

template <int ROWS>
struct alignas(16) NetLds {
  float b1[xpbd_mlp::NH], b2[xpbd_mlp::NH], w3[xpbd_mlp::NH];
  unsigned short ftile[kNetWaves][ROWS][16];
};

template <bool PREDICT, int ROWS>
__global__ __launch_bounds__(kNetBlock) void k_ws_net(
    const float4* __restrict__ Xin, const float4* __restrict__ XP, const float4* __restrict__ V,
    const int4* __restrict__ R, const float4* __restrict__ ANC,
    const xpbd_mlp::Image* __restrict__ img, unsigned short* __restrict__ Y, float h, float gx, float gy,
    float gz, const float* __restrict__ gscale, int m, const int* __restrict__ ws_on,
    int invert) {
  using namespace xpbd_mlp;
  // the learned start switched off for this launch: nothing to predict
  if (!xpbd_ws::start_active(ws_on, invert)) return;
  __shared__ NetLds<ROWS> s;
  for (int i = threadIdx.x; i < NH; i += kNetBlock) {
    s.b1[i] = img->w.b1[i];
    s.b2[i] = img->w.b2[i];
    s.w3[i] = img->w.w3[i];
  }
  RegWeights rw;
  load_reg_weights(rw, img->w);
  const float b3 = img->b3;
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, r16 = lane & 15, g = lane >> 4;
  auto pos = [&](int i) -> float4 {
    const float4 x4 = Xin[i];
    if (!PREDICT || x4.w == 0.0f) return x4;
    f3 gr = mk3(gx, gy, gz);
    if (gscale) gr = gr * gscale[i];
    const f3 v = xyz(V[i]) + gr * h;
    const f3 x = xyz(x4) + v * h;
    return make_float4(x.x, x.y, x.z, x4.w);
  };
  auto prev = [&](int i) { return PREDICT ? xyz(Xin[i]) : xyz(XP[i]); };
  const int groups = (m + ROWS - 1) / ROWS;
  for (int c = blockIdx.x * kNetWaves + wv; c < groups; c += gridDim.x * kNetWaves) {
    if (lane < ROWS) {
      const int j = c * ROWS + lane;
      float f[12];
      for (int q = 0; q < 12; ++q) f[q] = 0.0f;
      if (j < m) {
        const RowGeo rg = row_geo<true>(pos, prev, ANC, R[j], j, h);
        row_features(rg.t, rg.C, rg.w, rg.alpha, rg.vn, rg.vt, rg.pen, h, f);
      }
      unsigned pk[8];
      for (int q = 0; q < 6; ++q)
        pk[q] = (unsigned)to_bf16(f[2 * q]) | ((unsigned)to_bf16(f[2 * q + 1]) << 16);
      pk[6] = pk[7] = 0u;
      uint4* dst = (uint4*)&s.ftile[wv][lane][0];
      dst[0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
      dst[1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    for (int t = 0; t < ROWS / 16; ++t) {
      const int j = c * ROWS + 16 * t + r16;
      if (c * ROWS + 16 * t >= m) break;
      const bf16x4 xb = *(const bf16x4*)&s.ftile[wv][16 * t + r16][4 * g];
      const float y = forward16_t(rw, s.b1, s.b2, s.w3, xb, b3);
      if (g == 0 && j < m) Y[j] = xpbd_ws::y_bits(y);
    }
    __builtin_amdgcn_wave_barrier();
  }
}

template <int ROWS>
void launch_net(hipStream_t s, const float* Xin, const float* XP, const float* V, const int* R,
                const float* ANC, const xpbd_mlp::Image* img, unsigned short* Y, float h, float gx,
                float gy, float gz, const float* gscale, int n_rows, int predict,
                const int* ws_on, int invert) {
  const int groups = (n_rows + ROWS - 1) / ROWS;
  int nb = (groups + kNetWaves - 1) / kNetWaves;
  if (nb > kNetMaxBlocks) nb = kNetMaxBlocks;
  if (predict)
    hipLaunchKernelGGL((k_ws_net<true, ROWS>), dim3(nb), dim3(kNetBlock), 0, s,
                       (const float4*)Xin, (const float4*)XP, (const float4*)V, (const int4*)R,
                       (const float4*)ANC, img, Y, h, gx, gy, gz, gscale, n_rows, ws_on, invert);
  else
    hipLaunchKernelGGL((k_ws_net<false, ROWS>), dim3(nb), dim3(kNetBlock), 0, s,
                       (const float4*)Xin, (const float4*)XP, (const float4*)V, (const int4*)R,
                       (const float4*)ANC, img, Y, h, gx, gy, gz, gscale, n_rows, ws_on, invert);
}

// The judge's decision (one thread): sums[0] = squared residual of the
// substep the solver ran, sums[1] = that of the same substep run with the
// other start; the learned start stays (or becomes) active when its residual
// is no larger than the zero start's.  stats: decisions for the learned
// start, decisions for the zero start, switches.  The sums are cleared.
__global__ void k_ws_decide(double* sums, int* ws_on, int* stats) {
  const int active = *ws_on != 0;
  const double learned = active ? sums[0] : sums[1];
  const double zero = active ? sums[1] : sums[0];
  const int next = learned <= zero ? 1 : 0;
  if (stats) {
    stats[next ? 0 : 1] += 1;
    stats[2] += next != active;
  }
  *ws_on = next;
  sums[0] = 0.0;
  sums[1] = 0.0;
}

// Residual: sum over rows of r^2, r = C + alpha~ lambda (two-sided rows),
// max(0, -C) (one-sided rows).
// Grid-stride over the rows with at most kResidualBlocks blocks, so that one
// double atomic per block lands on the sum (one per 256 rows serialised on
// the single address and dominated the judge's cost on 11 M rows).
constexpr int kResidualBlocks = 1024;

__global__ __launch_bounds__(kB) void k_ws_residual(const float4* __restrict__ X,
                                                    const int4* __restrict__ R,
                                                    const float4* __restrict__ ANC,
                                                    const float* __restrict__ LAM, float h,
                                                    double* __restrict__ out, int m) {
  __shared__ double sh[kB];
  double r2 = 0.0;
  for (int j = blockIdx.x * kB + threadIdx.x; j < m; j += gridDim.x * kB) {
    const int4 row = R[j];
    const RowGeo g = row_geo<false>([&](int i) { return X[i]; },
                                    [&](int i) { return mk3(0.0f, 0.0f, 0.0f); }, ANC, row, j, h);
    float r;
    if (g.t == T_FLOOR || g.t == T_CONTACT) r = fmaxf(0.0f, -g.C);
    else r = g.C + g.alpha * LAM[j];
    r2 += (double)r * (double)r;
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

int xpbd_ws_net(void* stream, const float* Xin, const float* XP, const float* V, const int* R,
                const float* ANC, const void* image, unsigned short* Y, float h, float gx, float gy,
                float gz, const float* gscale, int n_rows, int predict, const int* ws_on,
                int invert) {
  if (n_rows <= 0) return 0;
  if (!image || !Y) return (int)hipErrorInvalidValue;
  const xpbd_mlp::Image* img = (const xpbd_mlp::Image*)image;
  hipStream_t s = (hipStream_t)stream;
  // 64-row groups from 256 groups up (measured on cloths: 113 k rows 14.5
  // against 21.2 us with 16-row groups, 1.8 M rows 110 against 147 us)
  if (n_rows >= 64 * 256)
    launch_net<64>(s, Xin, XP, V, R, ANC, img, Y, h, gx, gy, gz, gscale, n_rows, predict, ws_on,
                   invert);
  else
    launch_net<16>(s, Xin, XP, V, R, ANC, img, Y, h, gx, gy, gz, gscale, n_rows, predict, ws_on,
                   invert);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd_ws_decide(void* stream, double* sums, int* ws_on, int* stats) {
  if (!sums || !ws_on) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_ws_decide, dim3(1), dim3(1), 0, (hipStream_t)stream, sums, ws_on, stats);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd_ws_residual(void* stream, const float* X, const int* R, const float* ANC,
                     const float* LAM, float h, double* out_sum, int n_rows) {
  if (n_rows <= 0) return 0;
  const int nb = nblk(n_rows) < kResidualBlocks ? nblk(n_rows) : kResidualBlocks;
  hipLaunchKernelGGL(k_ws_residual, dim3(nb), dim3(kB), 0, (hipStream_t)stream,
                     (const float4*)X, (const int4*)R, (const float4*)ANC, LAM, h, out_sum,
                     n_rows);
  XPBD_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
