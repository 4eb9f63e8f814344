// Warm-start MLP (12 -> 64 -> 64 -> 1, tanh-form GELU) on gfx950 MFMA.
//
// No reference counterpart: the learned lambda warm start is new; it
// replaces the zero reset of the multipliers at the start of a substep
// (src/xpbd3d/solver.py:431, solver_6dof.py:374-377).
//
// The network layout, bfloat16 rounding and the per-wave forward live in
// mlp_dev.h (shared with the fused warm-start Jacobi launch).  A D tile of
// one layer is already the A^T / B operand of the weight gradient, and only
// the forward H1 and backward dZ2 tiles take a trip through LDS to become A
// operands.
//
// Training: each block accumulates its gradient in LDS and writes one
// partial vector; a second kernel sums the partials in block order
// (deterministic), and an Adam kernel updates the parameters.
#include <cmath>

#include "common.h"
#include "xpbd.h"

#include "mlp_dev.h"

namespace {

using namespace xpbd_mlp;

constexpr int kWaves = 4, kBlk = 64 * kWaves;

struct Lds {
  Weights w;
  unsigned short tile[kWaves][16][NH];  // per-wave transposition tile
  float grad[NPARAM];                    // block gradient (training only)
};

__device__ void load_params(Lds& s, const float* __restrict__ theta, bool zero_grad) {
  load_weights(s.w, theta, threadIdx.x, kBlk);
  if (zero_grad)
    for (int i = threadIdx.x; i < NPARAM; i += kBlk) s.grad[i] = 0.0f;
  __syncthreads();
}

// Forward of rows [row0, row0 + 16) of F (rows >= m read as zero).
__device__ inline void forward_tile(Lds& s, const float* __restrict__ F, int row0, int m,
                                    float b3, f32x4 z1[4], f32x4 z2[4], float y[4],
                                    bf16x4& xa) {
  const int wv = threadIdx.x >> 6;
  auto feat = [&](int r, int k) -> unsigned short {
    const int row = row0 + r;
    return (row < m && k < NIN) ? to_bf16(F[(size_t)row * NIN + k]) : (unsigned short)0;
  };
  forward16(s.w, s.tile[wv], feat, b3, z1, z2, y, xa);
}

__global__ __launch_bounds__(kBlk) void k_mlp_pack(const float* __restrict__ theta,
                                                   Image* __restrict__ img) {
  load_weights(img->w, theta, threadIdx.x, kBlk);
  if (threadIdx.x == 0) {
    img->b3 = theta[P_B3];
    img->pad[0] = img->pad[1] = img->pad[2] = 0.0f;
  }
}

// Inference over the rows of F (forward16_t, the order of the network pass
// of a substep, so both give the same bits for the same features).
__global__ __launch_bounds__(kBlk) void k_mlp_forward(const float* __restrict__ theta,
                                                      const float* __restrict__ F,
                                                      float* __restrict__ Y, int m) {
  __shared__ Weights w;
  load_weights(w, theta, threadIdx.x, kBlk);
  __syncthreads();
  RegWeights rw;
  load_reg_weights(rw, w);
  const int l = threadIdx.x & 63, r16 = l & 15, g = l >> 4, wv = threadIdx.x >> 6;
  const float b3 = theta[P_B3];
  const int ntiles = (m + 15) / 16;
  for (int t = blockIdx.x * kWaves + wv; t < ntiles; t += gridDim.x * kWaves) {
    const int row = t * 16 + r16;
    bf16x4 xb;
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * g + j;
      xb[j] = (short)((row < m && k < NIN) ? to_bf16(F[(size_t)row * NIN + k])
                                           : (unsigned short)0);
    }
    const float y = forward16_t(rw, w.b1, w.b2, w.w3, xb, b3);
    if (g == 0 && row < m) Y[row] = y;
  }
}

// Forward + backward of the mean-squared error over rows [0, m); writes the
// block's gradient (not yet divided by anything but m) and squared-error sum.
__global__ __launch_bounds__(kBlk) void k_mlp_train_partial(const float* __restrict__ theta,
                                                            const float* __restrict__ F,
                                                            const float* __restrict__ T, int m,
                                                            float* __restrict__ partial,
                                                            float* __restrict__ loss_partial) {
  __shared__ Lds s;
  load_params(s, theta, true);
  const int l = threadIdx.x & 63, r16 = l & 15, g = l >> 4, wv = threadIdx.x >> 6;
  const float b3 = theta[P_B3];
  const float inv_m2 = 2.0f / (float)m;
  f32x4 gW2[4][4], gW1[4];
  for (int a = 0; a < 4; ++a) {
    gW1[a] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int b = 0; b < 4; ++b) gW2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gw3[4] = {0, 0, 0, 0}, gb3 = 0.0f;
  float sq = 0.0f;
  const int ntiles = (m + 15) / 16;
  for (int t = blockIdx.x * kWaves + wv; t < ntiles; t += gridDim.x * kWaves) {
    f32x4 z1[4], z2[4];
    float y[4];
    bf16x4 xa;
    forward_tile(s, F, t * 16, m, b3, z1, z2, y, xa);
    float dy[4];
    for (int j = 0; j < 4; ++j) {
      const int row = t * 16 + 4 * g + j;
      const float r = row < m ? y[j] - T[row] : 0.0f;
      dy[j] = r * inv_m2;
      if (r16 == 0) {
        sq += r * r;
        gb3 += dy[j];
      }
    }
    // output layer and dZ2
    f32x4 dz2[4];
    for (int n = 0; n < 4; ++n) {
      const float w3 = s.w.w3[16 * n + r16];
      for (int j = 0; j < 4; ++j) {
        gw3[n] += gelu(z2[n][j]) * dy[j];
        dz2[n][j] = dy[j] * w3 * gelu_grad(z2[n][j]);
        gb2[n] += dz2[n][j];
      }
    }
    // dW2 = H1^T dZ2: the H1 and dZ2 D tiles are the A and B operands
    bf16x4 h1a[4], dzb[4];
    for (int n = 0; n < 4; ++n)
      for (int j = 0; j < 4; ++j) {
        h1a[n][j] = (short)to_bf16(gelu(z1[n][j]));
        dzb[n][j] = (short)to_bf16(dz2[n][j]);
      }
    for (int mt = 0; mt < 4; ++mt)
      for (int nt = 0; nt < 4; ++nt) gW2[mt][nt] = mfma(h1a[mt], dzb[nt], gW2[mt][nt]);
    // dH1 = dZ2 W2^T: dZ2 goes through LDS into the A layout
    for (int n = 0; n < 4; ++n)
      for (int j = 0; j < 4; ++j) s.tile[wv][4 * g + j][16 * n + r16] = (unsigned short)dzb[n][j];
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    bf16x8 a[2];  // two 32-wide K slices over the layer-2 units (mlp_dev.h)
    for (int m = 0; m < 2; ++m) {
      bf16x4 lo, hi;
      for (int j = 0; j < 4; ++j) {
        lo[j] = (short)s.tile[wv][r16][32 * m + 4 * g + j];
        hi[j] = (short)s.tile[wv][r16][32 * m + 16 + 4 * g + j];
      }
      a[m] = cat8(lo, hi);
    }
    __builtin_amdgcn_wave_barrier();
    bf16x4 dz1b[4];
    for (int nt = 0; nt < 4; ++nt) {
      f32x4 acc{0.f, 0.f, 0.f, 0.f};
      for (int m = 0; m < 2; ++m) {
        bf16x4 lo, hi;  // B[k][n] = W2[n][k]
        for (int j = 0; j < 4; ++j) {
          lo[j] = (short)s.w.w2t[32 * m + 4 * g + j][16 * nt + r16];
          hi[j] = (short)s.w.w2t[32 * m + 16 + 4 * g + j][16 * nt + r16];
        }
        acc = mfma32(a[m], cat8(lo, hi), acc);
      }
      for (int j = 0; j < 4; ++j) {
        const float dz = acc[j] * gelu_grad(z1[nt][j]);
        gb1[nt] += dz;
        dz1b[nt][j] = (short)to_bf16(dz);
      }
    }
    // dW1 = X^T dZ1: A[m][k] = X[row k][feature m]
    bf16x4 xt;
    {
      const int row0 = t * 16;
      for (int j = 0; j < 4; ++j) {
        const int row = row0 + 4 * g + j;
        const float v = (row < m && r16 < NIN) ? F[(size_t)row * NIN + r16] : 0.0f;
        xt[j] = (short)to_bf16(v);
      }
    }
    for (int nt = 0; nt < 4; ++nt) gW1[nt] = mfma(xt, dz1b[nt], gW1[nt]);
  }
  // block reduction into LDS (each wave adds its register partials)
  for (int mt = 0; mt < 4; ++mt)
    for (int nt = 0; nt < 4; ++nt)
      for (int j = 0; j < 4; ++j)
        atomicAdd(&s.grad[P_W2 + (16 * mt + 4 * g + j) * NH + 16 * nt + r16], gW2[mt][nt][j]);
  for (int nt = 0; nt < 4; ++nt)
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * g + j;
      if (k < NIN) atomicAdd(&s.grad[P_W1 + k * NH + 16 * nt + r16], gW1[nt][j]);
    }
  for (int n = 0; n < 4; ++n) {
    atomicAdd(&s.grad[P_B1 + 16 * n + r16], gb1[n]);
    atomicAdd(&s.grad[P_B2 + 16 * n + r16], gb2[n]);
    atomicAdd(&s.grad[P_W3 + 16 * n + r16], gw3[n]);
  }
  if (r16 == 0) atomicAdd(&s.grad[P_B3], gb3);
  __shared__ float sq_sh[kBlk];
  sq_sh[threadIdx.x] = sq;
  __syncthreads();
  for (int i = threadIdx.x; i < NPARAM; i += kBlk) partial[(size_t)blockIdx.x * NPARAM + i] = s.grad[i];
  if (threadIdx.x == 0) {
    float tot = 0.0f;
    for (int i = 0; i < kBlk; ++i) tot += sq_sh[i];
    loss_partial[blockIdx.x] = tot;
  }
}

__global__ __launch_bounds__(256) void k_reduce_partials(const float* __restrict__ partial,
                                                         int nblocks, float* __restrict__ grad,
                                                         const float* __restrict__ loss_partial,
                                                         float* __restrict__ loss, int m) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < NPARAM) {
    float acc = 0.0f;
    for (int b = 0; b < nblocks; ++b) acc += partial[(size_t)b * NPARAM + i];
    grad[i] = acc;
  }
  if (i == 0 && loss) {
    float tot = 0.0f;
    for (int b = 0; b < nblocks; ++b) tot += loss_partial[b];
    *loss = tot / (float)m;
  }
}

// Adam (Kingma & Ba 2015); bias corrections passed in as 1 - beta^t.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ theta, float* __restrict__ mom,
                                              float* __restrict__ vel,
                                              const float* __restrict__ grad, float lr, float b1,
                                              float b2, float eps, float c1, float c2, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = grad[i];
  const float m = b1 * mom[i] + (1.0f - b1) * gi;
  const float v = b2 * vel[i] + (1.0f - b2) * gi * gi;
  mom[i] = m;
  vel[i] = v;
  theta[i] = theta[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
}

}  // namespace

extern "C" {

int xpbd_mlp_param_count(void) { return NPARAM; }

int xpbd_mlp_forward(void* stream, const float* theta, const float* F, float* Y, int m) {
  if (m <= 0) return 0;
  int nb = (m + 16 * kWaves - 1) / (16 * kWaves);
  if (nb > 2048) nb = 2048;
  hipLaunchKernelGGL(k_mlp_forward, dim3(nb), dim3(kBlk), 0, (hipStream_t)stream, theta, F, Y, m);
  XPBD_LAUNCH_CHECK();
  return 0;
}

long long xpbd_mlp_image_bytes(void) { return (long long)sizeof(Image); }

int xpbd_mlp_pack(void* stream, const float* theta, void* image) {
  if (!theta || !image) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(kBlk), 0, (hipStream_t)stream, theta,
                     (Image*)image);
  XPBD_LAUNCH_CHECK();
  return 0;
}

long long xpbd_mlp_train_workspace_bytes(int nblocks) {
  return (long long)sizeof(float) * ((long long)nblocks * NPARAM + nblocks);
}

int xpbd_mlp_grad(void* stream, const float* theta, const float* F, const float* T, int m,
                  int nblocks, void* ws, float* grad, float* loss) {
  if (m <= 0 || nblocks <= 0) return (int)hipErrorInvalidValue;
  float* partial = (float*)ws;
  float* lossp = partial + (size_t)nblocks * NPARAM;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_mlp_train_partial, dim3(nblocks), dim3(kBlk), 0, s, theta, F, T, m,
                     partial, lossp);
  hipLaunchKernelGGL(k_reduce_partials, dim3((NPARAM + 255) / 256), dim3(256), 0, s, partial,
                     nblocks, grad, lossp, loss, m);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd_mlp_adam(void* stream, float* theta, float* mom, float* vel, const float* grad,
                  float lr, float beta1, float beta2, float eps, int step) {
  if (step < 1) return (int)hipErrorInvalidValue;
  const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, (hipStream_t)stream, theta,
                     mom, vel, grad, lr, beta1, beta2, eps, c1, c2, NPARAM);
  XPBD_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
