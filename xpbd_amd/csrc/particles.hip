// 3-DOF particle XPBD kernels for gfx950 and their C-ABI entry points.
//
// Device layout (all buffers owned by the caller):
//   X   float4[N]  (x, y, z, inv_mass)      position + inverse mass
//   V   float4[N]  (vx, vy, vz, 0)          velocity
//   XP  float4[N]  (x, y, z, 0)             substep-start position (x_prev)
//   R   int4[M]    packed constraint rows   (layout in common.h)
//   ANC float4[M]  ATTACH world anchors     (only read for ATTACH rows)
//   LAM float[M]   accumulated multipliers
//   ACC float4[N]  Jacobi accumulators      (sum dx.xyz, count in .w)
//
// Reference counterparts: src/xpbd3d/kernels.py (integrate 49-73,
// friction_delta 80-98, solve_constraints_color 105-208,
// solve_constraints_jacobi 219-322, apply_jacobi 326-339, finalize_velocity
// 346-358, cap_velocity 362-371) and the frame loop of Solver.step
// (src/xpbd3d/solver.py:412-467).
#include <cmath>

#include "common.h"
#include "ws_dev.h"
#include "xpbd.h"

namespace {

// kernels.py:80-98
__device__ inline f3 friction_delta(f3 dp_rel, f3 n_hat, float penetration, float mu) {
  f3 dp_t = dp_rel - n_hat * dot(dp_rel, n_hat);
  float lt = length(dp_t);
  if (lt < XPBD_EPS) return mk3(0.0f, 0.0f, 0.0f);
  float scale = 1.0f;
  float cone = mu * penetration;
  if (lt >= cone) scale = cone / lt;
  return dp_t * (-scale);
}

// Result of projecting one row against positions that are constant for the
// whole Jacobi launch: up to two body corrections (kernels.py:219-322).
struct Proj {
  int na;       // number of corrections (0, 1 or 2)
  int a, b;
  f3 ca, cb;
  bool upd;     // lambda was updated
  float lam;
};

// `pos(i)` returns the (x, y, z, inverse mass) the launch sees for particle
// i, `prev(i)` its x_prev and `anchor()` the row's ATTACH anchor; the loads
// happen only on the branches that use them.
// FAST_NORMAL: n = d * (1 / L) instead of d / L (one ulp apart at most).
template <bool FAST_NORMAL = false, class Pos, class Prev, class Anchor>
__device__ inline Proj project_row(Pos pos, Prev prev, Anchor anchor, int4 row, float lam,
                                   float inv_dt2) {
  Proj p;
  p.na = 0;
  p.upd = false;
  p.lam = lam;
  const int t = row_type(row);
  const int a = row_a(row);
  const int bb = row.y;
  const float p0 = __int_as_float(row.z);
  const float p1 = __int_as_float(row.w);
  p.a = a;
  p.b = bb;
  if (t == T_DISTANCE || t == T_ATTACH) {
    float4 xa4 = pos(a);
    float wa = xa4.w, wb;
    f3 pb;
    if (t == T_ATTACH) {
      pb = xyz(anchor());
      wb = 0.0f;
    } else {
      float4 xb4 = pos(bb);
      pb = xyz(xb4);
      wb = xb4.w;
    }
    float wsum = wa + wb;
    if (wsum == 0.0f) return p;
    f3 d = xyz(xa4) - pb;
    float L = length(d);
    if (L < XPBD_EPS) return p;
    f3 n_hat = FAST_NORMAL ? d * (1.0f / L) : d / L;
    float C = L - p0;
    float alpha = p1 * inv_dt2;
    float dlam = (-C - alpha * lam) / (wsum + alpha);
    p.lam = lam + dlam;
    p.upd = true;
    p.ca = n_hat * (wa * dlam);
    p.na = 1;
    if (t == T_DISTANCE) {
      p.cb = (-n_hat) * (wb * dlam);
      p.na = 2;
    }
    return p;
  }
  if (t == T_FLOOR) {
    float4 xa4 = pos(a);
    float wa = xa4.w;
    if (wa == 0.0f) return p;
    float C = xa4.y - p0;
    if (C >= 0.0f) return p;
    float dlam = -C / wa;
    p.lam = lam + dlam;
    p.upd = true;
    f3 corr = mk3(0.0f, wa * dlam, 0.0f);
    float mu = p1;
    if (mu > 0.0f) {
      f3 dp_rel = (xyz(xa4) + corr) - prev(a);
      corr = corr + friction_delta(dp_rel, mk3(0.0f, 1.0f, 0.0f), -C, mu);
    }
    p.ca = corr;
    p.na = 1;
    return p;
  }
  if (t == T_CONTACT) {
    float4 xa4 = pos(a), xb4 = pos(bb);
    float wa = xa4.w, wb = xb4.w;
    float wsum = wa + wb;
    if (wsum == 0.0f) return p;
    f3 d = xyz(xa4) - xyz(xb4);
    float L = length(d);
    if (L < XPBD_EPS) return p;
    f3 n_hat = d / L;
    float C = L - p0;
    if (C >= 0.0f) return p;
    float dlam = -C / wsum;
    p.lam = lam + dlam;
    p.upd = true;
    f3 ca = n_hat * (wa * dlam);
    f3 cb = (-n_hat) * (wb * dlam);
    float mu = p1;
    if (mu > 0.0f) {
      f3 dp_rel = ((xyz(xa4) + ca) - prev(a)) - ((xyz(xb4) + cb) - prev(bb));
      f3 fcorr = friction_delta(dp_rel, n_hat, -C, mu);
      ca = ca + fcorr * (wa / wsum);
      cb = cb - fcorr * (wb / wsum);
    }
    p.ca = ca;
    p.cb = cb;
    p.na = 2;
    return p;
  }
  return p;
}

__device__ inline Proj project_jacobi(const float4* __restrict__ X, const float4* __restrict__ XP,
                                      int4 row, float4 anchor_or_zero, float lam, float inv_dt2) {
  return project_row([&](int i) { return X[i]; }, [&](int i) { return xyz(XP[i]); },
                     [&]() { return anchor_or_zero; }, row, lam, inv_dt2);
}

// ---------------------------------------------------------------------------
// Jacobi projection with LDS-staged accumulation.
//
// Rows arrive sorted by their smallest endpoint, so the first row of a tile
// holds the tile's smallest particle index `base`.  Each block keeps a window
// of `window` accumulators [base, base+window) in LDS (SoA: x, y, z, count),
// adds every endpoint that falls inside it with LDS float atomics, and sends
// the rest straight to global atomics.  After the tile, the non-empty window
// entries are flushed to ACC with one global atomic per component.
// ---------------------------------------------------------------------------
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_jacobi_project(
    const float4* __restrict__ X, const float4* __restrict__ XP, const int4* __restrict__ R,
    const float4* __restrict__ ANC, float* __restrict__ LAM, float4* __restrict__ ACC, float h,
    int first, int n_rows, int rows_per_thread, int window) {
  extern __shared__ float lds[];
  float* sx = lds;
  float* sy = lds + window;
  float* sz = lds + 2 * window;
  float* sn = lds + 3 * window;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tile_rows = BLOCK * rows_per_thread;
  const int row0 = tile * tile_rows;
  if (row0 >= n_rows) return;
  const int4 r0 = R[row0];
  const int base = (r0.y >= 0 && r0.y < row_a(r0)) ? r0.y : row_a(r0);
  for (int k = threadIdx.x; k < 4 * window; k += BLOCK) lds[k] = 0.0f;
  __syncthreads();

  const float inv_dt2 = 1.0f / (h * h);
  const int row_end = min(row0 + tile_rows, n_rows);
  for (int j = row0 + threadIdx.x; j < row_end; j += BLOCK) {
    int4 row = R[j];
    float lam = first ? 0.0f : LAM[j];
    float4 anc = make_float4(0.f, 0.f, 0.f, 0.f);
    if (row_type(row) == T_ATTACH) anc = ANC[j];
    Proj p = project_jacobi(X, XP, row, anc, lam, inv_dt2);
    if (p.upd || first) LAM[j] = p.lam;
    if (p.na >= 1) {
      unsigned off = (unsigned)(p.a - base);
      if (off < (unsigned)window) {
        atomicAdd(&sx[off], p.ca.x);
        atomicAdd(&sy[off], p.ca.y);
        atomicAdd(&sz[off], p.ca.z);
        atomicAdd(&sn[off], 1.0f);
      } else {
        atomicAdd(&ACC[p.a].x, p.ca.x);
        atomicAdd(&ACC[p.a].y, p.ca.y);
        atomicAdd(&ACC[p.a].z, p.ca.z);
        atomicAdd(&ACC[p.a].w, 1.0f);
      }
    }
    if (p.na == 2) {
      unsigned off = (unsigned)(p.b - base);
      if (off < (unsigned)window) {
        atomicAdd(&sx[off], p.cb.x);
        atomicAdd(&sy[off], p.cb.y);
        atomicAdd(&sz[off], p.cb.z);
        atomicAdd(&sn[off], 1.0f);
      } else {
        atomicAdd(&ACC[p.b].x, p.cb.x);
        atomicAdd(&ACC[p.b].y, p.cb.y);
        atomicAdd(&ACC[p.b].z, p.cb.z);
        atomicAdd(&ACC[p.b].w, 1.0f);
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < window; k += BLOCK) {
    float c = sn[k];
    if (c != 0.0f) {
      float4* dst = &ACC[base + k];
      atomicAdd(&dst->x, sx[k]);
      atomicAdd(&dst->y, sy[k]);
      atomicAdd(&dst->z, sz[k]);
      atomicAdd(&dst->w, c);
    }
  }
}

// Plain per-row global-atomic Jacobi projection (no LDS window); kept as the
// reference point the LDS-staged variant is measured against.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_jacobi_project_global(
    const float4* __restrict__ X, const float4* __restrict__ XP, const int4* __restrict__ R,
    const float4* __restrict__ ANC, float* __restrict__ LAM, float4* __restrict__ ACC, float h,
    int first, int n_rows) {
  const int j = blockIdx.x * BLOCK + threadIdx.x;
  if (j >= n_rows) return;
  const float inv_dt2 = 1.0f / (h * h);
  int4 row = R[j];
  float lam = first ? 0.0f : LAM[j];
  float4 anc = make_float4(0.f, 0.f, 0.f, 0.f);
  if (row_type(row) == T_ATTACH) anc = ANC[j];
  Proj p = project_jacobi(X, XP, row, anc, lam, inv_dt2);
  if (p.upd || first) LAM[j] = p.lam;
  if (p.na >= 1) {
    atomicAdd(&ACC[p.a].x, p.ca.x);
    atomicAdd(&ACC[p.a].y, p.ca.y);
    atomicAdd(&ACC[p.a].z, p.ca.z);
    atomicAdd(&ACC[p.a].w, 1.0f);
  }
  if (p.na == 2) {
    atomicAdd(&ACC[p.b].x, p.cb.x);
    atomicAdd(&ACC[p.b].y, p.cb.y);
    atomicAdd(&ACC[p.b].z, p.cb.z);
    atomicAdd(&ACC[p.b].w, 1.0f);
  }
}

// Projection of a row segment whose live length is on the device (rows
// generated on the GPU, e.g. CONTACT rows): plain global atomics into ACC.
// PREDICT: the first iteration of a substep with integrate folded in, as in
// the tile kernel: positions are predicted from X and V on the fly (the
// arithmetic of integrate, gscale optional), x_prev is X itself.
template <int BLOCK, bool PREDICT>
__global__ __launch_bounds__(BLOCK) void k_jacobi_project_dyn(
    const float4* __restrict__ X, const float4* __restrict__ XP, const float4* __restrict__ V,
    const int4* __restrict__ R, float* __restrict__ LAM, float4* __restrict__ ACC, float h,
    int first, float gx, float gy, float gz, const float* __restrict__ gscale,
    const int* __restrict__ n_dev, int cap) {
  const int j = blockIdx.x * BLOCK + threadIdx.x;
  if (j >= min(*n_dev, cap)) return;
  const float inv_dt2 = 1.0f / (h * h);
  const int4 row = R[j];
  const float lam = first ? 0.0f : LAM[j];
  auto pos = [&](int i) -> float4 {
    const float4 x4 = X[i];
    if (!PREDICT || x4.w == 0.0f) return x4;
    f3 g = mk3(gx, gy, gz);
    if (gscale) g = g * gscale[i];
    const f3 v = xyz(V[i]) + g * h;
    const f3 x = xyz(x4) + v * h;
    return make_float4(x.x, x.y, x.z, x4.w);
  };
  auto prev = [&](int i) { return PREDICT ? xyz(X[i]) : xyz(XP[i]); };
  const Proj p = project_row(pos, prev, [&]() { return make_float4(0.f, 0.f, 0.f, 0.f); }, row,
                             lam, inv_dt2);
  if (p.upd || first) LAM[j] = p.lam;
  if (p.na >= 1) {
    atomicAdd(&ACC[p.a].x, p.ca.x);
    atomicAdd(&ACC[p.a].y, p.ca.y);
    atomicAdd(&ACC[p.a].z, p.ca.z);
    atomicAdd(&ACC[p.a].w, 1.0f);
  }
  if (p.na == 2) {
    atomicAdd(&ACC[p.b].x, p.cb.x);
    atomicAdd(&ACC[p.b].y, p.cb.y);
    atomicAdd(&ACC[p.b].z, p.cb.z);
    atomicAdd(&ACC[p.b].w, 1.0f);
  }
}

// ---------------------------------------------------------------------------
// Fused Jacobi iteration over particle tiles (projection + apply in one
// launch, no global accumulators or atomics).
//
// Particles are split into tiles of P consecutive indices; a row is owned by
// the tile of its smallest endpoint (rows are grouped by that tile, so tile t
// owns the contiguous rows [tile_row_start[t], tile_row_start[t+1]); inside a
// tile they are ordered by type and by rank among the rows of the same
// smallest endpoint, so neighbouring lanes take the same branch and add into
// different particles).  A row whose other endpoint lies in another tile is
// also listed as a halo row of that tile and projected again there: both
// tiles read the same X_in / LAM_in, so they compute the same correction;
// each adds only its own endpoints and only the owner writes LAM_out.
//
// Accumulation: corrections are summed per particle in LDS as 2^-40 m fixed
// point with 64-bit integer atomics (plus a 32-bit count).  Integer sums are
// exact and independent of the order in which lanes arrive, so an iteration
// is bitwise reproducible from run to run; 64-bit integer LDS atomics also
// run several times faster than 32-bit float LDS atomics on CDNA.
// Conversion in two instructions: (double)c + 6144.0 lies in [4096, 8192),
// a binade whose ulp is 2^-40, so the add rounds c to the nearest multiple
// of 2^-40 (ties to even) and the bit pattern of the sum is
// bits(6144.0) + round(c * 2^40).  Those bit patterns are added (modulo
// 2^64) and the count x bits(6144.0) is subtracted at apply.  Range:
// |correction| < 2^11 m per row; a larger or non-finite correction marks the
// particle, whose position then becomes NaN (as a float sum would).
//
// The unit normal of the tile kernel's rows is d * (1 / L) rather than the
// reference's three divides d / L (one ulp apart at most); C, the lambda
// update and the correction magnitude keep the reference's arithmetic.
//
// After a barrier every thread applies the averaged correction to its tile's
// particles and writes X_out (ping-pong buffers), so one launch is one
// complete Jacobi iteration (kernels.py:219-339).  On the last iteration of a
// substep the same pass also runs finalize_velocity and cap_velocity
// (kernels.py:346-371).
// ---------------------------------------------------------------------------
constexpr double kFixBias = 6144.0;                    // 1.5 * 2^12, ulp 2^-40
constexpr unsigned long long kFixBiasBits = 0x40B8000000000000ull;  // bits of 6144.0
constexpr double kFixInv = 1.0 / 1099511627776.0;      // 2^-40
constexpr float kFixLimit = 2047.0f;                   // |correction| per row, metres
constexpr unsigned kFixPoison = 1u << 24;              // count unit of a bad correction

__device__ inline unsigned long long fix_bits(float c) {
  return (unsigned long long)__double_as_longlong((double)c + kFixBias);
}
// Sum of `adds` fix_bits values -> metres.
__device__ inline float unfix(unsigned long long s, unsigned adds) {
  return (float)((double)(long long)(s - (unsigned long long)adds * kFixBiasBits) * kFixInv);
}

// LDS of a tile: x, y, z sums (int64, P each), count (u32, P), then the
// tile's positions (float4, P) at a 16-byte aligned offset.
__host__ __device__ inline size_t tile_lds_pos_offset(int P) {
  return ((size_t)P * (3 * sizeof(long long) + sizeof(unsigned)) + 15) & ~(size_t)15;
}
__host__ __device__ inline size_t tile_lds_bytes(int P) { return tile_lds_pos_offset(P) + (size_t)P * sizeof(float4); }

//
// PREDICT: the first iteration of a substep with `integrate` (kernels.py:49-73)
// folded in.  Xin then holds the positions at the start of the substep; every
// position the launch reads is predicted on the fly, x~ = x + (v + g h) h,
// from Xin and V (the same arithmetic as k_integrate_to, so the result is
// identical to an integrate launch followed by the plain iteration), x_prev
// is written for the tile's own particles, and lambda starts from zero.  V
// is left as it was: finalize, in a later launch of the same substep,
// rewrites it, so a predict launch must never also finalize.
//
// NEXT (with FINALIZE): the last iteration of a substep also runs the next
// substep's integrate on the finalized velocity held in registers: x_prev <-
// x, x <- x + (v + g h) h (k_integrate_to's arithmetic on the value finalize
// would have written to V, so the result is identical to finalize, then
// integrate), and V is not written: the next substep's finalize derives the
// velocity from x and x_prev again, and the frame's last substep finalizes
// without NEXT.  Other tiles must not read x_prev during such a launch, so it
// is used only when no row of the launch reads the x_prev of a particle
// outside its own tile (no CONTACT rows: FLOOR rows read their own
// particle's).
//
// The launch mode is a template argument (bits of TileMode) so that a
// launch pays only for what it uses: every optional branch left in the
// kernel costs registers and scheduling freedom in the row loop (a runtime
// finalize / ACC branch alone measured ~10 % on 8 M-particle cloths).
// FIRST: the first iteration of a substep whose positions are already
// integrated (lambda starts from zero: no multiplier loads).
enum TileMode : int {
  TM_PREDICT = 1, TM_FINALIZE = 2, TM_ACC = 4, TM_ANCHOR = 8, TM_NEXT = 16, TM_FIRST = 32
};

// From this many particles the new positions are written with non-temporal
// stores: 64 MB of positions plus the rows stream through the caches anyway,
// and not allocating the written lines measured 2 % faster on 8 x 1024^2
// cloths (0.366 -> 0.359 ms); small scenes re-read them from L2.
constexpr int kStreamOutParticles = 1 << 22;

template <int BLOCK, int MODE>
__global__ __launch_bounds__(BLOCK) void k_jacobi_tile(
    const float4* __restrict__ Xin, float4* __restrict__ Xout, float4* __restrict__ XP,
    float4* __restrict__ V, const int4* __restrict__ R, const float4* __restrict__ ANC,
    const float* __restrict__ LAMin, float* __restrict__ LAMout,
    const int* __restrict__ tile_row_start, const int* __restrict__ halo_start,
    const int* __restrict__ halo_rows, float h, int first, int P, int n, float relax,
    int finalize, float damping, float max_speed, int do_cap, float4* __restrict__ ACC, float gx,
    float gy, float gz, const float* __restrict__ gscale) {
  constexpr bool PREDICT = MODE & TM_PREDICT, FINALIZE = MODE & TM_FINALIZE;
  constexpr bool HAS_ACC = MODE & TM_ACC, HAS_ANCHOR = MODE & TM_ANCHOR;
  constexpr bool NEXT = MODE & TM_NEXT;
  static_assert(!NEXT || (FINALIZE && !PREDICT), "NEXT extends a finalizing launch");
  constexpr bool ZERO_LAM = PREDICT || (MODE & TM_FIRST);
  extern __shared__ unsigned long long lds64[];
  unsigned long long* sx = lds64;
  unsigned long long* sy = lds64 + P;
  unsigned long long* sz = lds64 + 2 * P;
  unsigned* cnt = (unsigned*)(lds64 + 3 * P);
  float4* spos = (float4*)(lds64 + tile_lds_pos_offset(P) / sizeof(unsigned long long));
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int p0 = tile * P;
  if (p0 >= n) return;
  const int p1 = min(p0 + P, n);
  const float inv_dt2 = 1.0f / (h * h);
  auto predict = [&](float4 x4, int i) -> float4 {
    if (!PREDICT || x4.w == 0.0f) return x4;
    f3 g = mk3(gx, gy, gz);
    if (gscale) g = g * gscale[i];
    const f3 v = xyz(V[i]) + g * h;
    const f3 x = xyz(x4) + v * h;
    return make_float4(x.x, x.y, x.z, x4.w);
  };
  for (int k = threadIdx.x; k < 3 * P; k += BLOCK) lds64[k] = 0ull;
  for (int k = threadIdx.x; k < P; k += BLOCK) cnt[k] = 0u;
  const int rb = tile_row_start[tile], re = tile_row_start[tile + 1];
  const int hb = halo_start[tile], he = halo_start[tile + 1];
  const int owned = re - rb;
  const int total = owned + (he - hb);
  constexpr int U = 2;
  int4 row[U];
  float lam[U];
  int jj[U];
  auto load_pass = [&](int k0) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u * BLOCK;
      jj[u] = -1;
      if (k < total) {
        jj[u] = k < owned ? rb + k : halo_rows[hb + (k - owned)];
        row[u] = R[jj[u]];
        lam[u] = (ZERO_LAM || first) ? 0.0f : LAMin[jj[u]];
      }
    }
  };
  load_pass(threadIdx.x);
  // the tile's own positions (predicted in a PREDICT launch) go to LDS once;
  // rows read endpoints of this tile from there and others from HBM / L2
  for (int i = p0 + threadIdx.x; i < p1; i += BLOCK) {
    const float4 x0 = Xin[i];
    spos[i - p0] = predict(x0, i);
    if (PREDICT) XP[i] = make_float4(x0.x, x0.y, x0.z, 0.0f);
  }
  __syncthreads();
  auto pos = [&](int i) -> float4 {
    const unsigned off = (unsigned)(i - p0);
    if (off < (unsigned)(p1 - p0)) return spos[off];
    return predict(Xin[i], i);
  };
  auto prev = [&](int i) { return PREDICT ? xyz(Xin[i]) : xyz(XP[i]); };
  auto add = [&](int i, f3 c) {
    const unsigned off = (unsigned)(i - p0);
    if (off >= (unsigned)P) return;
    // branch-free: out-of-range or NaN components are clamped for the
    // conversion and flag the particle through the count
    const bool ok = fabsf(c.x) <= kFixLimit && fabsf(c.y) <= kFixLimit && fabsf(c.z) <= kFixLimit;
    atomicAdd(&sx[off], fix_bits(fminf(fmaxf(c.x, -kFixLimit), kFixLimit)));
    atomicAdd(&sy[off], fix_bits(fminf(fmaxf(c.y, -kFixLimit), kFixLimit)));
    atomicAdd(&sz[off], fix_bits(fminf(fmaxf(c.z, -kFixLimit), kFixLimit)));
    atomicAdd(&cnt[off], ok ? 1u : (kFixPoison | 1u));
  };
  // two rows per thread per pass, their loads issued together: those of the
  // first pass before the staging barrier (they overlap the position
  // staging), those of each next pass before the current pass is projected
  // (61 VGPRs: still 8 waves per SIMD)
  for (int k0 = threadIdx.x; k0 < total; k0 += BLOCK * U) {
    int4 nrow[U];
    float nlam[U];
    int njj[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + (U + u) * BLOCK;
      njj[u] = -1;
      if (k < total) {
        njj[u] = k < owned ? rb + k : halo_rows[hb + (k - owned)];
        nrow[u] = R[njj[u]];
        nlam[u] = (ZERO_LAM || first) ? 0.0f : LAMin[njj[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (jj[u] < 0) continue;
      const int j = jj[u];
      const Proj p = project_row<true>(
          pos, prev, [&]() { return HAS_ANCHOR ? ANC[j] : make_float4(0.f, 0.f, 0.f, 0.f); },
          row[u], lam[u], inv_dt2);
      if (k0 + u * BLOCK < owned) LAMout[jj[u]] = p.lam;
      if (p.na >= 1) add(p.a, p.ca);
      if (p.na == 2) add(p.b, p.cb);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      row[u] = nrow[u];
      lam[u] = nlam[u];
      jj[u] = njj[u];
    }
  }
  __syncthreads();
  const bool stream_out = n >= kStreamOutParticles;
  for (int i = p0 + threadIdx.x; i < p1; i += BLOCK) {
    const int o = i - p0;
    const float4 x4 = spos[o];
    const unsigned c = cnt[o];
    const unsigned adds = c & (kFixPoison - 1u);
    float4 a = make_float4(0.f, 0.f, 0.f, (float)adds);
    if (c) a = make_float4(unfix(sx[o], adds), unfix(sy[o], adds), unfix(sz[o], adds), a.w);
    if (c >= kFixPoison) a.x = __int_as_float(0x7fc00000);
    if (HAS_ACC) {  // corrections of the device-counted row segment
      const float4 g = ACC[i];
      if (g.w != 0.0f) {
        a = make_float4(a.x + g.x, a.y + g.y, a.z + g.z, a.w + g.w);
        ACC[i] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
    f3 x = xyz(x4);
    if (a.w > 0.0f && x4.w > 0.0f) x = x + xyz(a) * (relax / a.w);
    if (FINALIZE) {
      f3 v;
      if (x4.w == 0.0f) {
        v = mk3(0.0f, 0.0f, 0.0f);
      } else {
        v = ((x - xyz(XP[i])) / h) * (1.0f - damping);
      }
      if (do_cap) {
        const float speed = length(v);
        if (speed > max_speed) v = v * (max_speed / speed);
      }
      if (NEXT) {  // the next substep's integrate (k_integrate_to)
        XP[i] = make_float4(x.x, x.y, x.z, 0.0f);
        if (x4.w != 0.0f) {
          f3 g = mk3(gx, gy, gz);
          if (gscale) g = g * gscale[i];
          const f3 vn = v + g * h;
          x = x + vn * h;
        }
      } else {
        V[i] = make_float4(v.x, v.y, v.z, 0.0f);
      }
    }
    if (stream_out) {  // a frame's positions far exceed the caches: do not keep them
      typedef float f4v __attribute__((ext_vector_type(4)));
      const f4v o = {x.x, x.y, x.z, x4.w};
      __builtin_nontemporal_store(o, (f4v*)&Xout[i]);
    } else {
      Xout[i] = make_float4(x.x, x.y, x.z, x4.w);
    }
  }
}

// ---------------------------------------------------------------------------
// The fused iteration over the compact tables (xpbd_tile_compact).
//
// A tile's own DISTANCE rows come as 12 bytes (a tile-relative endpoint pair
// + rest in an int2, the compliance in a float: 4 bytes fewer than the int4
// row), its other rows from R (GEN), and, with PF, the FLOOR row of each
// particle that has exactly one is projected by the apply pass from 8 bytes
// per particle: its correction joins the exact fixed-point sum like any
// other row's, so the iteration is the int4 kernel's bit for bit; its
// multiplier is updated in place, read and written only where the row is
// active (or reset in a zero-start launch).  On 8 x 1024^2 hanging cloths
// these cut the bytes a launch moves by ~17 %.
//
// The compact rows' loads are unconditional, their index clamped into the
// segment (a clamped load re-reads one cached row), and they are issued one
// pass ahead: with no branch around them the compiler can count the loads in
// flight and wait only for the pass it projects (with the loads under a
// branch it waits for all of them, the next pass's included).  Every option
// is a template argument (the zero start and the streaming stores
// included): the same kernel with a runtime flag tested in the apply pass
// measured 7 % slower on the 8 x 1024^2 cloths (0.323 against 0.305 ms per
// launch).
// ---------------------------------------------------------------------------
template <int MODE, bool GEN, bool PF, bool STREAM>
__global__ __launch_bounds__(512) void k_jacobi_tile_cx(
    const float4* __restrict__ Xin, float4* __restrict__ Xout, float4* __restrict__ XP,
    float4* __restrict__ V, const int4* __restrict__ R, const float4* __restrict__ ANC,
    const float* __restrict__ LAMin, float* __restrict__ LAMout,
    const int* __restrict__ tile_row_start, const int* __restrict__ halo_start,
    const int* __restrict__ dist_end, const int* __restrict__ gen_end,
    const int2* __restrict__ CI, const float* __restrict__ CC, const int2* __restrict__ H2,
    const float2* __restrict__ PFT, float* __restrict__ PFLAM, float h, int P, int n,
    float relax, float damping, float max_speed, int do_cap, float4* __restrict__ ACC, float gx,
    float gy, float gz, const float* __restrict__ gscale, const int* __restrict__ acc_tiles) {
  constexpr int BLOCK = 512;
  constexpr int U = 2;
  constexpr bool PREDICT = MODE & TM_PREDICT, FINALIZE = MODE & TM_FINALIZE;
  constexpr bool HAS_ACC = MODE & TM_ACC, HAS_ANCHOR = MODE & TM_ANCHOR;
  constexpr bool NEXT = MODE & TM_NEXT;
  constexpr bool ZERO_LAM = PREDICT || (MODE & TM_FIRST);
  static_assert(!NEXT || (FINALIZE && !PREDICT), "NEXT extends a finalizing launch");
  static_assert(!HAS_ANCHOR || GEN, "ATTACH rows are read from R");
  extern __shared__ unsigned long long lds64[];
  unsigned long long* sx = lds64;
  unsigned long long* sy = lds64 + P;
  unsigned long long* sz = lds64 + 2 * P;
  unsigned* cnt = (unsigned*)(lds64 + 3 * P);
  float4* spos = (float4*)(lds64 + tile_lds_pos_offset(P) / sizeof(unsigned long long));
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int p0 = tile * P;
  if (p0 >= n) return;
  const int p1 = min(p0 + P, n);
  const int tid = threadIdx.x;
  const float inv_dt2 = 1.0f / (h * h);
  // tiles without a particle that device rows can reach skip the ACC read
  const bool tile_acc = HAS_ACC && (!acc_tiles || acc_tiles[tile] != 0);
  const int rb = tile_row_start[tile];
  const int rd = dist_end[tile];
  const int hb = halo_start[tile], he = halo_start[tile + 1];
  const int n_cmp = rd - rb;
  const int last = max(rd - 1, 0);  // a valid row index even for an empty segment
  int2 q[U];
  float cm[U], lm[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int j = min(rb + tid + u * BLOCK, last);
    q[u] = CI[j];
    cm[u] = CC[j];
    lm[u] = ZERO_LAM ? 0.0f : LAMin[j];
  }
  auto predict = [&](float4 x4, int i) -> float4 {
    if (!PREDICT || x4.w == 0.0f) return x4;
    f3 g = mk3(gx, gy, gz);
    if (gscale) g = g * gscale[i];
    const f3 v = xyz(V[i]) + g * h;
    const f3 x = xyz(x4) + v * h;
    return make_float4(x.x, x.y, x.z, x4.w);
  };
  for (int k = tid; k < 3 * P; k += BLOCK) lds64[k] = 0ull;
  for (int k = tid; k < P; k += BLOCK) cnt[k] = 0u;
  for (int i = p0 + tid; i < p1; i += BLOCK) {
    const float4 x0 = Xin[i];
    spos[i - p0] = predict(x0, i);
    if (PREDICT) XP[i] = make_float4(x0.x, x0.y, x0.z, 0.0f);
  }
  __syncthreads();
  auto pos = [&](int i) -> float4 {
    const unsigned off = (unsigned)(i - p0);
    if (off < (unsigned)(p1 - p0)) return spos[off];
    return predict(Xin[i], i);
  };
  auto prev = [&](int i) { return PREDICT ? xyz(Xin[i]) : xyz(XP[i]); };
  auto add = [&](int i, f3 c) {
    const unsigned off = (unsigned)(i - p0);
    if (off >= (unsigned)P) return;
    const bool ok = fabsf(c.x) <= kFixLimit && fabsf(c.y) <= kFixLimit && fabsf(c.z) <= kFixLimit;
    atomicAdd(&sx[off], fix_bits(fminf(fmaxf(c.x, -kFixLimit), kFixLimit)));
    atomicAdd(&sy[off], fix_bits(fminf(fmaxf(c.y, -kFixLimit), kFixLimit)));
    atomicAdd(&sz[off], fix_bits(fminf(fmaxf(c.z, -kFixLimit), kFixLimit)));
    atomicAdd(&cnt[off], ok ? 1u : (kFixPoison | 1u));
  };
  auto no_anchor = [] { return make_float4(0.f, 0.f, 0.f, 0.f); };
  // a compact row of the tile based at `base`: DISTANCE, no type bits
  auto compact = [&](int2 qq, float c, float l, int base) {
    const unsigned v = (unsigned)qq.x;
    const int4 row = make_int4(base + (int)(v & 1023u), base + (int)(v >> 10), qq.y,
                               __float_as_int(c));
    return project_row<true>(pos, prev, no_anchor, row, l, inv_dt2);
  };
  for (int k0 = tid; k0 < n_cmp; k0 += BLOCK * U) {
    int2 nq[U];
    float ncm[U], nlm[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = min(rb + k0 + (U + u) * BLOCK, last);
      nq[u] = CI[j];
      ncm[u] = CC[j];
      nlm[u] = ZERO_LAM ? 0.0f : LAMin[j];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u * BLOCK;
      if (k < n_cmp) {
        const Proj p = compact(q[u], cm[u], lm[u], p0);
        LAMout[rb + k] = p.lam;
        if (p.na >= 1) add(p.a, p.ca);
        if (p.na == 2) add(p.b, p.cb);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      q[u] = nq[u];
      cm[u] = ncm[u];
      lm[u] = nlm[u];
    }
  }
  if (GEN) {  // rows read whole: ATTACH, CONTACT, far DISTANCE, FLOOR without PF
    const int rg = PF ? gen_end[tile] : tile_row_start[tile + 1];
    for (int j = rd + tid; j < rg; j += BLOCK) {
      const float l = ZERO_LAM ? 0.0f : LAMin[j];
      const Proj p = project_row<true>(
          pos, prev, [&]() { return HAS_ANCHOR ? ANC[j] : make_float4(0.f, 0.f, 0.f, 0.f); },
          R[j], l, inv_dt2);
      LAMout[j] = p.lam;
      if (p.na >= 1) add(p.a, p.ca);
      if (p.na == 2) add(p.b, p.cb);
    }
  }
  // halo rows: owned by an earlier tile, projected again for this tile's
  // endpoint (the owner writes their multiplier)
  for (int k = hb + tid; k < he; k += BLOCK) {
    const int2 e = H2[k];
    const int j = e.x;
    const float l = ZERO_LAM ? 0.0f : LAMin[j];
    Proj p;
    if (GEN && e.y < 0) {
      p = project_row<true>(
          pos, prev, [&]() { return HAS_ANCHOR ? ANC[j] : make_float4(0.f, 0.f, 0.f, 0.f); },
          R[j], l, inv_dt2);
    } else {
      p = compact(CI[j], CC[j], l, e.y);
    }
    if (p.na >= 1) add(p.a, p.ca);
    if (p.na == 2) add(p.b, p.cb);
  }
  __syncthreads();
  for (int i = p0 + tid; i < p1; i += BLOCK) {
    const int o = i - p0;
    const float4 x4 = spos[o];
    unsigned c = cnt[o];
    unsigned long long s0 = sx[o], s1 = sy[o], s2 = sz[o];
    if (PF) {  // the particle's FLOOR row (kernels.py:272-291)
      const float2 pf = PFT[i];
      const float C = x4.y - pf.x;
      if (x4.w != 0.0f && C < 0.0f) {
        const float wa = x4.w;
        const float dlam = -C / wa;
        PFLAM[i] = (ZERO_LAM ? 0.0f : PFLAM[i]) + dlam;
        f3 corr = mk3(0.0f, wa * dlam, 0.0f);
        if (pf.y > 0.0f) {  // x_prev: the substep's start (PREDICT wrote it above)
          const f3 dp_rel = (xyz(x4) + corr) - xyz(XP[i]);
          corr = corr + friction_delta(dp_rel, mk3(0.0f, 1.0f, 0.0f), -C, pf.y);
        }
        const bool ok = fabsf(corr.x) <= kFixLimit && fabsf(corr.y) <= kFixLimit &&
                        fabsf(corr.z) <= kFixLimit;
        s0 += fix_bits(fminf(fmaxf(corr.x, -kFixLimit), kFixLimit));
        s1 += fix_bits(fminf(fmaxf(corr.y, -kFixLimit), kFixLimit));
        s2 += fix_bits(fminf(fmaxf(corr.z, -kFixLimit), kFixLimit));
        c += ok ? 1u : (kFixPoison | 1u);
      } else if (ZERO_LAM && pf.x != -INFINITY) {
        PFLAM[i] = 0.0f;
      }
    }
    const unsigned adds = c & (kFixPoison - 1u);
    float4 a = make_float4(0.f, 0.f, 0.f, (float)adds);
    if (c) a = make_float4(unfix(s0, adds), unfix(s1, adds), unfix(s2, adds), a.w);
    if (c >= kFixPoison) a.x = __int_as_float(0x7fc00000);
    if (tile_acc) {  // corrections of the device-counted row segment
      const float4 g = ACC[i];
      if (g.w != 0.0f) {
        a = make_float4(a.x + g.x, a.y + g.y, a.z + g.z, a.w + g.w);
        ACC[i] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
    f3 x = xyz(x4);
    if (a.w > 0.0f && x4.w > 0.0f) x = x + xyz(a) * (relax / a.w);
    if (FINALIZE) {
      f3 v;
      if (x4.w == 0.0f) {
        v = mk3(0.0f, 0.0f, 0.0f);
      } else {
        v = ((x - xyz(XP[i])) / h) * (1.0f - damping);
      }
      if (do_cap) {
        const float speed = length(v);
        if (speed > max_speed) v = v * (max_speed / speed);
      }
      if (NEXT) {  // the next substep's integrate (k_integrate_to)
        XP[i] = make_float4(x.x, x.y, x.z, 0.0f);
        if (x4.w != 0.0f) {
          f3 g = mk3(gx, gy, gz);
          if (gscale) g = g * gscale[i];
          const f3 vn = v + g * h;
          x = x + vn * h;
        }
      } else {
        V[i] = make_float4(v.x, v.y, v.z, 0.0f);
      }
    }
    if (STREAM) {  // a frame's positions far exceed the caches: do not keep them
      typedef float f4v __attribute__((ext_vector_type(4)));
      const f4v ov = {x.x, x.y, x.z, x4.w};
      __builtin_nontemporal_store(ov, (f4v*)&Xout[i]);
    } else {
      Xout[i] = make_float4(x.x, x.y, x.z, x4.w);
    }
  }
}

// The learned start of one row (no reference counterpart): lambda0 from
// the network output yv and the row's first plain step -C / (w + alpha~)
// (one-sided rows only while violated), and the direction its endpoints
// move along: a by nrm * wa * lam0, b (two) by -nrm * wb * lam0.
struct WarmRow {
  float lam0, wa, wb;
  f3 nrm;
  int a, b;
  bool two;
};
template <class Pos, class Anchor>
__device__ inline WarmRow warm_row(Pos pos, Anchor anchor, int4 row, float yv, float inv_dt2) {
  WarmRow r;
  const int t = row_type(row);
  r.a = row_a(row);
  r.b = row.y;
  const float p0r = __int_as_float(row.z);
  const float4 xa4 = pos(r.a);
  r.wa = xa4.w;
  r.wb = 0.0f;
  float C, w, alpha = 0.0f;
  r.nrm = mk3(0.0f, 1.0f, 0.0f);
  bool ok = true;
  if (t == T_FLOOR) {
    C = xa4.y - p0r;
    w = r.wa;
  } else {
    f3 pb;
    if (t == T_ATTACH) {
      pb = xyz(anchor());
    } else {
      const float4 xb4 = pos(r.b);
      pb = xyz(xb4);
      r.wb = xb4.w;
    }
    const f3 d = xyz(xa4) - pb;
    const float L = length(d);
    ok = L >= XPBD_EPS;
    if (ok) r.nrm = d * __builtin_amdgcn_rcpf(L);
    C = L - p0r;
    w = r.wa + r.wb;
    if (t != T_CONTACT) alpha = __int_as_float(row.w) * inv_dt2;
  }
  const float sden = w + alpha;
  float step = (sden > 0.0f && ok) ? -C * __builtin_amdgcn_rcpf(sden) : 0.0f;
  if (t == T_FLOOR || t == T_CONTACT) step = C < 0.0f ? step : 0.0f;
  r.lam0 = xpbd_ws::start_lambda(yv, step);
  r.two = t == T_DISTANCE || t == T_CONTACT;
  return r;
}

// ---------------------------------------------------------------------------
// A whole two-iteration substep in one launch (small scenes, where every
// launch costs its latency rather than its bytes): the first iteration
// (xpbd_jacobi_tile with first = 1) followed by the second with finalize
// and, unless LAST, the next substep's integrate (xpbd_jacobi_tile_next), or
// finalize + cap (xpbd_jacobi_tile with finalize) when LAST.
// WARM: the first iteration is replaced by the learned start of
// xpbd_jacobi_tile_warm (network output Y per row, start flag ws_on), the
// second is the substep's one plain iteration from its multipliers.
// PREDICT: the substep's integrate too (the frame's first substep,
// xpbd_jacobi_tile_first): every local particle is predicted from X and V
// at staging, its x_prev is X (XPin is not read); not with LAST, whose
// velocity stores other blocks would read.
//
// A tile's second iteration reads the first iteration's positions of the
// particles one row away (its ring 1), so the block also runs the first
// iteration for them: every row incident to the tile or its ring 1
// (phase 1, endpoints in the tile, ring 1 or ring 2), projected from the
// launch's input positions with lambda = 0, summed with the same exact
// fixed-point accumulation and applied the same way: the ring's first
// iteration positions are bit for bit those of the separate launch.  Phase
// 2 then projects every row incident to the tile from those positions and
// the phase-1 multipliers, and applies, finalizes and integrates the tile.
// Rows and multipliers in LDS between the phases; the owner tile writes a
// row's multiplier (as the separate launches do).  x_prev is read from XPin
// (rows of ring particles read their x_prev) and written to XPout, a second
// buffer: the next substep's start positions, or (LAST) a copy of XPin, so
// that the last substep's x_prev ends in the buffer the plan names.
//
// Host tables per tile t (xpbd_amd/solver.py: tile_pair_tables):
//   ring_start[t] .. ring_start[t+1]: global indices of ring 1, then ring 2
//     (ring1_n[t] of them in ring 1); local index = P + position.
//   p1_start[t] .. p1_start[t+1]: phase-1 entries int4 (row j, local a,
//     local b or -1, 0).
//   p2_start[t] .. p2_start[t+1]: phase-2 entries, the phase-1 position of
//     the row | (1 << 31) when the tile owns the row.
// Dynamic LDS (tile_pair_lds_bytes): positions, x_prev and global indices
// of the local particles (lmax), fixed-point sums of the tile and ring 1
// (amax), phase-1 rows, row indices and multipliers (k1max).
//
// The launch is latency-bound (a 128^2 cloth has 256 tiles of 64): every
// global load whose address the tables give (the thread's first phase-1
// rows through their row index, its first phase-2 entries, the local
// particles' positions and x_prev through ring_idx) is issued before the
// staging barrier, so the block waits for two load latencies instead of one
// per phase and pass.  Blocks of 1024 threads (4 waves per SIMD with one
// block per CU) hide more of the phases' LDS and ALU latency than 512: 7.05
// against 7.64 us per launch on the 128^2 cloth, same results.
// ---------------------------------------------------------------------------
__host__ __device__ inline size_t tile_pair_lds_bytes(int lmax, int amax, int k1max) {
  return (size_t)amax * 3 * sizeof(unsigned long long) + (size_t)lmax * 2 * sizeof(float4) +
         (size_t)k1max * sizeof(int4) + (size_t)amax * sizeof(unsigned) +
         (size_t)lmax * sizeof(int) + (size_t)k1max * (sizeof(int) + sizeof(float));
}

template <int BLOCK, bool LAST, bool PREDICT, bool WARM>
__global__ __launch_bounds__(BLOCK) void k_jacobi_tile_pair(
    const float4* __restrict__ Xin, float4* __restrict__ Xout, const float4* XPin,
    float4* XPout, float4* __restrict__ V, const int4* __restrict__ R,
    const float4* __restrict__ ANC, float* __restrict__ LAMout,
    const int* __restrict__ ring_start, const int* __restrict__ ring1_n,
    const int* __restrict__ ring_idx, const int* __restrict__ p1_start,
    const int4* __restrict__ p1, const int* __restrict__ p2_start, const int* __restrict__ p2,
    float h, int P, int n, float relax, float damping, float max_speed, int do_cap, float gx,
    float gy, float gz, const float* __restrict__ gscale, int lmax, int amax, int k1max,
    const unsigned short* __restrict__ Y, const int* __restrict__ ws_on, int invert) {
  constexpr int U1 = 3;  // phase-1 rows per thread loaded ahead (3 x BLOCK per tile)
  constexpr int U2 = 2;  // phase-2 entries per thread loaded ahead
  extern __shared__ unsigned long long lds64[];
  unsigned long long* sx = lds64;
  unsigned long long* sy = sx + amax;
  unsigned long long* sz = sy + amax;
  float4* spos = (float4*)(sz + amax);
  float4* sprev = spos + lmax;
  int4* srow = (int4*)(sprev + lmax);
  unsigned* cnt = (unsigned*)(srow + k1max);
  int* gidx = (int*)(cnt + amax);
  int* sj = gidx + lmax;
  float* slam = (float*)(sj + k1max);
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int p0 = tile * P;
  if (p0 >= n) return;
  const int np = min(p0 + P, n) - p0;
  const int rb = ring_start[tile], re = ring_start[tile + 1];
  const int nr1 = ring1_n[tile];
  const int nloc = P + (re - rb);
  const int nacc = P + nr1;  // tile and ring 1 get sums
  const int k1b = p1_start[tile], n1 = p1_start[tile + 1] - k1b;
  const int k2b = p2_start[tile], n2 = p2_start[tile + 1] - k2b;
  const float inv_dt2 = 1.0f / (h * h);
  // rows of phase 1 in local form: body index -> local index of endpoint a,
  // b -> local index of endpoint b (or -1)
  auto local_row = [&](int k, int& j) {
    const int4 e = p1[k1b + k];
    const int4 row = R[e.x];
    j = e.x;
    return make_int4((row.x & ~ROW_BODY_MASK) | e.y, e.z, row.z, row.w);
  };
  int4 prow[U1];
  int pj[U1];
  float py[U1];  // WARM: the rows' network outputs
#pragma unroll
  for (int u = 0; u < U1; ++u) {
    const int k = threadIdx.x + u * BLOCK;
    pj[u] = -1;
    py[u] = 0.0f;
    if (k < n1) {
      prow[u] = local_row(k, pj[u]);
      if (WARM) py[u] = xpbd_ws::y_of(Y, pj[u]);
    }
  }
  // WARM with the start switched off: lambda0 = 0 and no move (the zero start)
  const bool active = !WARM || xpbd_ws::start_active(ws_on, invert);
  int pe[U2];
#pragma unroll
  for (int u = 0; u < U2; ++u) {
    const int k = threadIdx.x + u * BLOCK;
    pe[u] = k < n2 ? p2[k2b + k] : 0;
  }
  for (int k = threadIdx.x; k < nacc; k += BLOCK) {
    sx[k] = sy[k] = sz[k] = 0ull;
    cnt[k] = 0u;
  }
  for (int l = threadIdx.x; l < nloc; l += BLOCK) {
    int g = -1;
    if (l < P) {
      if (l < np) g = p0 + l;
    } else {
      g = ring_idx[rb + (l - P)];
    }
    gidx[l] = g;
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 x4 = g >= 0 ? Xin[g] : z;
    if (PREDICT) {  // integrate (kernels.py:49-73): x_prev = x, x += (v + g h) h
      const float4 v4 = g >= 0 ? V[g] : z;
      sprev[l] = make_float4(x4.x, x4.y, x4.z, 0.0f);
      if (g >= 0 && x4.w != 0.0f) {
        f3 gv = mk3(gx, gy, gz);
        if (gscale) gv = gv * gscale[g];
        const f3 vn = xyz(v4) + gv * h;
        const f3 x = xyz(x4) + vn * h;
        x4 = make_float4(x.x, x.y, x.z, x4.w);
      }
    } else {
      sprev[l] = g >= 0 ? XPin[g] : z;
    }
    spos[l] = x4;
  }
  __syncthreads();
  auto pos = [&](int l) { return spos[l]; };
  auto prev = [&](int l) { return xyz(sprev[l]); };
  auto add = [&](int l, f3 c, int lim) {
    if ((unsigned)l >= (unsigned)lim) return;
    const bool ok = fabsf(c.x) <= kFixLimit && fabsf(c.y) <= kFixLimit && fabsf(c.z) <= kFixLimit;
    atomicAdd(&sx[l], fix_bits(fminf(fmaxf(c.x, -kFixLimit), kFixLimit)));
    atomicAdd(&sy[l], fix_bits(fminf(fmaxf(c.y, -kFixLimit), kFixLimit)));
    atomicAdd(&sz[l], fix_bits(fminf(fmaxf(c.z, -kFixLimit), kFixLimit)));
    atomicAdd(&cnt[l], ok ? 1u : (kFixPoison | 1u));
  };
  // phase 1: the first iteration of the tile and ring 1 (lambda from zero),
  // or (WARM) the learned start of k_jacobi_tile_warm on the same rows
  auto phase1 = [&](int k, int4 lrow, int j, float yv) {
    srow[k] = lrow;
    sj[k] = j;
    if (WARM) {
      float lam0 = 0.0f;
      if (active) {
        const WarmRow wr = warm_row(
            pos, [&]() { return ANC ? ANC[j] : make_float4(0.f, 0.f, 0.f, 0.f); }, lrow, yv,
            inv_dt2);
        lam0 = wr.lam0;
        if (lam0 != 0.0f) {
          add(wr.a, wr.nrm * (wr.wa * lam0), nacc);
          if (wr.two) add(wr.b, wr.nrm * (-wr.wb * lam0), nacc);
        }
      }
      slam[k] = lam0;
      return;
    }
    const Proj p = project_row<true>(
        pos, prev, [&]() { return ANC ? ANC[j] : make_float4(0.f, 0.f, 0.f, 0.f); }, lrow,
        0.0f, inv_dt2);
    slam[k] = p.lam;
    if (p.na >= 1) add(p.a, p.ca, nacc);
    if (p.na == 2) add(p.b, p.cb, nacc);
  };
#pragma unroll
  for (int u = 0; u < U1; ++u)
    if (pj[u] >= 0) phase1(threadIdx.x + u * BLOCK, prow[u], pj[u], py[u]);
  for (int k = threadIdx.x + U1 * BLOCK; k < n1; k += BLOCK) {
    int j;
    const int4 lrow = local_row(k, j);
    phase1(k, lrow, j, WARM ? xpbd_ws::y_of(Y, j) : 0.0f);
  }
  __syncthreads();
  // the first iteration's apply (k_jacobi_tile without finalize), in LDS
  auto sums = [&](int o) {
    const unsigned c = cnt[o];
    const unsigned adds = c & (kFixPoison - 1u);
    float4 a = make_float4(0.f, 0.f, 0.f, (float)adds);
    if (c) a = make_float4(unfix(sx[o], adds), unfix(sy[o], adds), unfix(sz[o], adds), a.w);
    if (c >= kFixPoison) a.x = __int_as_float(0x7fc00000);
    return a;
  };
  for (int o = threadIdx.x; o < nacc; o += BLOCK) {
    const float4 a = sums(o);
    if (o < P) {  // cleared as read: phase 2 sums the tile again
      sx[o] = sy[o] = sz[o] = 0ull;
      cnt[o] = 0u;
    }
    if (gidx[o] < 0) continue;
    const float4 x4 = spos[o];
    f3 x = xyz(x4);
    if (a.w > 0.0f && x4.w > 0.0f) x = x + xyz(a) * (relax / a.w);
    spos[o] = make_float4(x.x, x.y, x.z, x4.w);
  }
  __syncthreads();
  // phase 2: the second iteration of the tile's rows
  auto phase2 = [&](int e) {
    const int k1 = e & 0x7fffffff;
    const int j = sj[k1];
    const Proj p = project_row<true>(
        pos, prev, [&]() { return ANC ? ANC[j] : make_float4(0.f, 0.f, 0.f, 0.f); }, srow[k1],
        slam[k1], inv_dt2);
    if (e < 0) LAMout[j] = p.lam;
    if (p.na >= 1) add(p.a, p.ca, P);
    if (p.na == 2) add(p.b, p.cb, P);
  };
#pragma unroll
  for (int u = 0; u < U2; ++u)
    if (threadIdx.x + u * BLOCK < n2) phase2(pe[u]);
  for (int k = threadIdx.x + U2 * BLOCK; k < n2; k += BLOCK) phase2(p2[k2b + k]);
  __syncthreads();
  // apply + finalize + cap, then the next substep's integrate
  // (k_jacobi_tile with FINALIZE and NEXT) or the velocities (LAST)
  for (int o = threadIdx.x; o < np; o += BLOCK) {
    const int i = p0 + o;
    const float4 x4 = spos[o];
    const float4 a = sums(o);
    f3 x = xyz(x4);
    if (a.w > 0.0f && x4.w > 0.0f) x = x + xyz(a) * (relax / a.w);
    const float4 xp4 = sprev[o];
    f3 v;
    if (x4.w == 0.0f) {
      v = mk3(0.0f, 0.0f, 0.0f);
    } else {
      v = ((x - xyz(xp4)) / h) * (1.0f - damping);
    }
    if (do_cap) {
      const float speed = length(v);
      if (speed > max_speed) v = v * (max_speed / speed);
    }
    if (!LAST) {
      XPout[i] = make_float4(x.x, x.y, x.z, 0.0f);
      if (x4.w != 0.0f) {
        f3 g = mk3(gx, gy, gz);
        if (gscale) g = g * gscale[i];
        const f3 vn = v + g * h;
        x = x + vn * h;
      }
    } else {
      if (XPout != XPin) XPout[i] = xp4;
      V[i] = make_float4(v.x, v.y, v.z, 0.0f);
    }
    Xout[i] = make_float4(x.x, x.y, x.z, x4.w);
  }
}

// ---------------------------------------------------------------------------
// Learned warm start launch (no reference counterpart; it replaces the zero
// start of solver.py:431), run once per substep after the network pass
// (xpbd_ws_net, which writes each row's output y) and before the Jacobi
// iterations, which then continue from its multipliers (first = 0).
//
// Same tiles, staging (with integrate folded in when PREDICT) and exact
// accumulation as k_jacobi_tile.  Every row (owned and halo) starts from
// lambda0 = 4 sigmoid(y) lambda1, lambda1 = -C / (w + alpha~) its first
// plain step (one-sided rows only while violated), writes lambda0 (owner)
// and adds w_i grad C lambda0 to its endpoints; each particle moves by
// relax x the mean of those offsets, the position a Jacobi sweep reaches
// when its multipliers add up to lambda0 (Jacobi applies the mean of a
// particle's corrections, kernels.py:326-339).  The start is a learned
// estimate, so the normal, lambda1 and the sigmoid use the hardware
// reciprocal and exp2 (v_rcp_f32, v_exp_f32; a few ulp): 4.29 against
// 4.65 us per launch on the 128^2 cloth with IEEE divides and expf.
// ---------------------------------------------------------------------------
template <int BLOCK, bool PREDICT>
__global__ __launch_bounds__(BLOCK) void k_jacobi_tile_warm(
    const float4* __restrict__ Xin, float4* __restrict__ Xout, float4* __restrict__ XP,
    const float4* __restrict__ V, const int4* __restrict__ R, const float4* __restrict__ ANC,
    float* __restrict__ LAMout, const int* __restrict__ tile_row_start,
    const int* __restrict__ halo_start, const int* __restrict__ halo_rows,
    const unsigned short* __restrict__ Y, float h, int P, int n, float relax, float gx, float gy,
    float gz, const float* __restrict__ gscale, const int* __restrict__ ws_on, int invert) {
  extern __shared__ unsigned long long lds64[];
  // switched off: lambda0 = 0 for every row, positions as predicted (the
  // following iterations then run exactly as from the zero start)
  const bool active = xpbd_ws::start_active(ws_on, invert);
  unsigned long long* sx = lds64;
  unsigned long long* sy = lds64 + P;
  unsigned long long* sz = lds64 + 2 * P;
  unsigned* cnt = (unsigned*)(lds64 + 3 * P);
  float4* spos = (float4*)(lds64 + tile_lds_pos_offset(P) / sizeof(unsigned long long));
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int p0 = tile * P;
  if (p0 >= n) return;
  const int p1 = min(p0 + P, n);
  const float inv_dt2 = 1.0f / (h * h);
  auto predict = [&](float4 x4, int i) -> float4 {
    if (!PREDICT || x4.w == 0.0f) return x4;
    f3 g = mk3(gx, gy, gz);
    if (gscale) g = g * gscale[i];
    const f3 v = xyz(V[i]) + g * h;
    const f3 x = xyz(x4) + v * h;
    return make_float4(x.x, x.y, x.z, x4.w);
  };
  if (!active) {  // the zero start: positions as predicted, multipliers 0
    for (int i = p0 + threadIdx.x; i < p1; i += BLOCK) {
      const float4 x0 = Xin[i];
      Xout[i] = predict(x0, i);
      if (PREDICT) XP[i] = make_float4(x0.x, x0.y, x0.z, 0.0f);
    }
    for (int j = tile_row_start[tile] + threadIdx.x; j < tile_row_start[tile + 1]; j += BLOCK)
      LAMout[j] = 0.0f;
    return;
  }
  for (int k = threadIdx.x; k < 3 * P; k += BLOCK) lds64[k] = 0ull;
  for (int k = threadIdx.x; k < P; k += BLOCK) cnt[k] = 0u;
  for (int i = p0 + threadIdx.x; i < p1; i += BLOCK) {
    const float4 x0 = Xin[i];
    spos[i - p0] = predict(x0, i);
    if (PREDICT) XP[i] = make_float4(x0.x, x0.y, x0.z, 0.0f);
  }
  __syncthreads();
  const int rb = tile_row_start[tile], re = tile_row_start[tile + 1];
  const int hb = halo_start[tile], he = halo_start[tile + 1];
  const int owned = re - rb;
  const int total = owned + (he - hb);
  auto pos = [&](int i) -> float4 {
    const unsigned off = (unsigned)(i - p0);
    if (off < (unsigned)(p1 - p0)) return spos[off];
    return predict(Xin[i], i);
  };
  auto add = [&](int i, f3 c) {
    const unsigned off = (unsigned)(i - p0);
    if (off >= (unsigned)P) return;
    const bool ok = fabsf(c.x) <= kFixLimit && fabsf(c.y) <= kFixLimit && fabsf(c.z) <= kFixLimit;
    atomicAdd(&sx[off], fix_bits(fminf(fmaxf(c.x, -kFixLimit), kFixLimit)));
    atomicAdd(&sy[off], fix_bits(fminf(fmaxf(c.y, -kFixLimit), kFixLimit)));
    atomicAdd(&sz[off], fix_bits(fminf(fmaxf(c.z, -kFixLimit), kFixLimit)));
    atomicAdd(&cnt[off], ok ? 1u : (kFixPoison | 1u));
  };
  for (int k = threadIdx.x; k < total; k += BLOCK) {
    const int j = k < owned ? rb + k : halo_rows[hb + (k - owned)];
    const WarmRow wr = warm_row(pos, [&]() { return ANC[j]; }, R[j], xpbd_ws::y_of(Y, j), inv_dt2);
    if (k < owned) LAMout[j] = wr.lam0;
    if (wr.lam0 == 0.0f) continue;
    add(wr.a, wr.nrm * (wr.wa * wr.lam0));
    if (wr.two) add(wr.b, wr.nrm * (-wr.wb * wr.lam0));
  }
  __syncthreads();
  for (int i = p0 + threadIdx.x; i < p1; i += BLOCK) {
    const int o = i - p0;
    const float4 x4 = spos[o];
    const unsigned c = cnt[o];
    const unsigned adds = c & (kFixPoison - 1u);
    float4 a = make_float4(0.f, 0.f, 0.f, (float)adds);
    if (c) a = make_float4(unfix(sx[o], adds), unfix(sy[o], adds), unfix(sz[o], adds), a.w);
    if (c >= kFixPoison) a.x = __int_as_float(0x7fc00000);
    f3 x = xyz(x4);
    if (a.w > 0.0f && x4.w > 0.0f) x = x + xyz(a) * (relax / a.w);
    Xout[i] = make_float4(x.x, x.y, x.z, x4.w);
  }
}

// The learned start over the compact tables (xpbd_tile_compact): the same
// rows, lambda0 and moves as k_jacobi_tile_warm, the compact rows' loads
// unconditional and issued one pass ahead as in k_jacobi_tile_cx.  GEN:
// some rows of [dist_end, gen_end) (PF) or [dist_end, end) are read from R.
// PF: the single FLOOR rows [gen_end, end) are read from R too (the launch
// needs their network output), and their lambda0 goes to PFLAM, where the
// following k_jacobi_tile_cx launches keep it.
template <bool PREDICT, bool GEN, bool PF>
__global__ __launch_bounds__(512) void k_jacobi_tile_warm_cx(
    const float4* __restrict__ Xin, float4* __restrict__ Xout, float4* __restrict__ XP,
    const float4* __restrict__ V, const int4* __restrict__ R, const float4* __restrict__ ANC,
    float* __restrict__ LAMout, const int* __restrict__ tile_row_start,
    const int* __restrict__ halo_start, const int* __restrict__ dist_end,
    const int* __restrict__ gen_end, const int2* __restrict__ CI, const float* __restrict__ CC,
    const int2* __restrict__ H2, float* __restrict__ PFLAM, const unsigned short* __restrict__ Y,
    float h, int P, int n, float relax, float gx, float gy, float gz,
    const float* __restrict__ gscale, const int* __restrict__ ws_on, int invert) {
  constexpr int BLOCK = 512;
  constexpr int U = 2;
  extern __shared__ unsigned long long lds64[];
  const bool active = xpbd_ws::start_active(ws_on, invert);
  unsigned long long* sx = lds64;
  unsigned long long* sy = lds64 + P;
  unsigned long long* sz = lds64 + 2 * P;
  unsigned* cnt = (unsigned*)(lds64 + 3 * P);
  float4* spos = (float4*)(lds64 + tile_lds_pos_offset(P) / sizeof(unsigned long long));
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int p0 = tile * P;
  if (p0 >= n) return;
  const int p1 = min(p0 + P, n);
  const int tid = threadIdx.x;
  const float inv_dt2 = 1.0f / (h * h);
  auto predict = [&](float4 x4, int i) -> float4 {
    if (!PREDICT || x4.w == 0.0f) return x4;
    f3 g = mk3(gx, gy, gz);
    if (gscale) g = g * gscale[i];
    const f3 v = xyz(V[i]) + g * h;
    const f3 x = xyz(x4) + v * h;
    return make_float4(x.x, x.y, x.z, x4.w);
  };
  const int rb = tile_row_start[tile], re = tile_row_start[tile + 1];
  if (!active) {  // the zero start: positions as predicted, multipliers 0
    for (int i = p0 + tid; i < p1; i += BLOCK) {
      const float4 x0 = Xin[i];
      Xout[i] = predict(x0, i);
      if (PREDICT) XP[i] = make_float4(x0.x, x0.y, x0.z, 0.0f);
    }
    for (int j = rb + tid; j < re; j += BLOCK) LAMout[j] = 0.0f;
    if (PF) {  // the following launches start the single FLOOR rows from PFLAM
      for (int j = gen_end[tile] + tid; j < re; j += BLOCK) PFLAM[row_a(R[j])] = 0.0f;
    }
    return;
  }
  const int rd = dist_end[tile];
  const int hb = halo_start[tile], he = halo_start[tile + 1];
  const int n_cmp = rd - rb;
  const int last = max(rd - 1, 0);
  int2 q[U];
  float cm[U], yv[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int j = min(rb + tid + u * BLOCK, last);
    q[u] = CI[j];
    cm[u] = CC[j];
    yv[u] = xpbd_ws::y_of(Y, j);
  }
  for (int k = tid; k < 3 * P; k += BLOCK) lds64[k] = 0ull;
  for (int k = tid; k < P; k += BLOCK) cnt[k] = 0u;
  for (int i = p0 + tid; i < p1; i += BLOCK) {
    const float4 x0 = Xin[i];
    spos[i - p0] = predict(x0, i);
    if (PREDICT) XP[i] = make_float4(x0.x, x0.y, x0.z, 0.0f);
  }
  __syncthreads();
  auto pos = [&](int i) -> float4 {
    const unsigned off = (unsigned)(i - p0);
    if (off < (unsigned)(p1 - p0)) return spos[off];
    return predict(Xin[i], i);
  };
  auto add = [&](int i, f3 c) {
    const unsigned off = (unsigned)(i - p0);
    if (off >= (unsigned)P) return;
    const bool ok = fabsf(c.x) <= kFixLimit && fabsf(c.y) <= kFixLimit && fabsf(c.z) <= kFixLimit;
    atomicAdd(&sx[off], fix_bits(fminf(fmaxf(c.x, -kFixLimit), kFixLimit)));
    atomicAdd(&sy[off], fix_bits(fminf(fmaxf(c.y, -kFixLimit), kFixLimit)));
    atomicAdd(&sz[off], fix_bits(fminf(fmaxf(c.z, -kFixLimit), kFixLimit)));
    atomicAdd(&cnt[off], ok ? 1u : (kFixPoison | 1u));
  };
  auto no_anchor = [] { return make_float4(0.f, 0.f, 0.f, 0.f); };
  auto compact = [](int2 qq, float c, int base) {  // DISTANCE: no type bits
    const unsigned v = (unsigned)qq.x;
    return make_int4(base + (int)(v & 1023u), base + (int)(v >> 10), qq.y, __float_as_int(c));
  };
  auto move = [&](const WarmRow& wr) {
    if (wr.lam0 == 0.0f) return;
    add(wr.a, wr.nrm * (wr.wa * wr.lam0));
    if (wr.two) add(wr.b, wr.nrm * (-wr.wb * wr.lam0));
  };
  for (int k0 = tid; k0 < n_cmp; k0 += BLOCK * U) {
    int2 nq[U];
    float ncm[U], nyv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = min(rb + k0 + (U + u) * BLOCK, last);
      nq[u] = CI[j];
      ncm[u] = CC[j];
      nyv[u] = xpbd_ws::y_of(Y, j);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u * BLOCK;
      if (k < n_cmp) {
        const WarmRow wr = warm_row(pos, no_anchor, compact(q[u], cm[u], p0), yv[u], inv_dt2);
        LAMout[rb + k] = wr.lam0;
        move(wr);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      q[u] = nq[u];
      cm[u] = ncm[u];
      yv[u] = nyv[u];
    }
  }
  const int rg = PF ? gen_end[tile] : re;
  if (GEN) {  // every other owned row from R
    for (int j = rd + tid; j < rg; j += BLOCK) {
      const WarmRow wr = warm_row(pos, [&]() { return ANC[j]; }, R[j], xpbd_ws::y_of(Y, j), inv_dt2);
      LAMout[j] = wr.lam0;
      move(wr);
    }
  }
  if (PF) {  // the single FLOOR rows: lambda0 per particle
    for (int j = rg + tid; j < re; j += BLOCK) {
      const WarmRow wr = warm_row(pos, no_anchor, R[j], xpbd_ws::y_of(Y, j), inv_dt2);
      PFLAM[wr.a] = wr.lam0;
      move(wr);
    }
  }
  for (int k = hb + tid; k < he; k += BLOCK) {
    const int2 e = H2[k];
    const int j = e.x;
    const int4 row = (GEN && e.y < 0) ? R[j] : compact(CI[j], CC[j], e.y);
    move(warm_row(pos, [&]() { return ANC[j]; }, row, xpbd_ws::y_of(Y, j), inv_dt2));
  }
  __syncthreads();
  for (int i = p0 + tid; i < p1; i += BLOCK) {
    const int o = i - p0;
    const float4 x4 = spos[o];
    const unsigned c = cnt[o];
    const unsigned adds = c & (kFixPoison - 1u);
    float4 a = make_float4(0.f, 0.f, 0.f, (float)adds);
    if (c) a = make_float4(unfix(sx[o], adds), unfix(sy[o], adds), unfix(sz[o], adds), a.w);
    if (c >= kFixPoison) a.x = __int_as_float(0x7fc00000);
    f3 x = xyz(x4);
    if (a.w > 0.0f && x4.w > 0.0f) x = x + xyz(a) * (relax / a.w);
    Xout[i] = make_float4(x.x, x.y, x.z, x4.w);
  }
}

// integrate (kernels.py:49-73) from Xin into Xout (which may alias Xin).
// gscale (optional): per-particle gravity factor, g_i = g * gscale[i].
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_integrate_to(const float4* Xin, float4* Xout,
                                                        float4* __restrict__ V,
                                                        float4* __restrict__ XP, float gx,
                                                        float gy, float gz, float h,
                                                        const float* __restrict__ gscale, int n) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const float4 x4 = Xin[i];
  f3 x = xyz(x4);
  XP[i] = make_float4(x.x, x.y, x.z, 0.0f);
  if (x4.w != 0.0f) {
    f3 g = mk3(gx, gy, gz);
    if (gscale) g = g * gscale[i];
    const f3 v = xyz(V[i]) + g * h;
    x = x + v * h;
    V[i] = make_float4(v.x, v.y, v.z, 0.0f);
  }
  Xout[i] = make_float4(x.x, x.y, x.z, x4.w);
}

// ---------------------------------------------------------------------------
// Colored Gauss-Seidel: one launch per color over a contiguous slice of the
// color-sorted rows; rows of one color touch disjoint particles, so the
// in-place updates are race-free (kernels.py:105-208).
// ---------------------------------------------------------------------------
// One Gauss-Seidel row, in place on X (global memory, or the LDS copy of
// the single-workgroup frame kernel): kernels.py:105-208.
__device__ inline void gs_row(float4* __restrict__ X, const float4* __restrict__ XP,
                              const int4* __restrict__ R, const float4* __restrict__ ANC,
                              float* __restrict__ LAM, float inv_dt2, int first, int j) {
  const int4 row = R[j];
  const int t = row_type(row);
  const int a = row_a(row);
  const int bb = row.y;
  const float p0 = __int_as_float(row.z);
  const float p1 = __int_as_float(row.w);
  float lam = first ? 0.0f : LAM[j];
  if (first) LAM[j] = 0.0f;

  if (t == T_DISTANCE || t == T_ATTACH) {
    float4 xa4 = X[a];
    float wa = xa4.w, wb;
    f3 pb;
    float4 xb4;
    if (t == T_ATTACH) {
      pb = xyz(ANC[j]);
      wb = 0.0f;
    } else {
      xb4 = X[bb];
      pb = xyz(xb4);
      wb = xb4.w;
    }
    float wsum = wa + wb;
    if (wsum == 0.0f) return;
    f3 d = xyz(xa4) - pb;
    float L = length(d);
    if (L < XPBD_EPS) return;
    f3 n_hat = d / L;
    float C = L - p0;
    float alpha = p1 * inv_dt2;
    float dlam = (-C - alpha * lam) / (wsum + alpha);
    LAM[j] = lam + dlam;
    X[a] = with_xyz(xa4, xyz(xa4) + n_hat * (wa * dlam));
    if (t == T_DISTANCE) X[bb] = with_xyz(xb4, xyz(xb4) - n_hat * (wb * dlam));
    return;
  }
  if (t == T_FLOOR) {
    float4 xa4 = X[a];
    float wa = xa4.w;
    if (wa == 0.0f) return;
    float C = xa4.y - p0;
    if (C >= 0.0f) return;
    float dlam = -C / wa;
    LAM[j] = lam + dlam;
    f3 xa = xyz(xa4) + mk3(0.0f, wa * dlam, 0.0f);
    float mu = p1;
    if (mu > 0.0f) {
      f3 dp_rel = xa - xyz(XP[a]);
      f3 corr = friction_delta(dp_rel, mk3(0.0f, 1.0f, 0.0f), -C, mu);
      xa = xa + corr;
    }
    X[a] = with_xyz(xa4, xa);
    return;
  }
  if (t == T_CONTACT) {
    float4 xa4 = X[a], xb4 = X[bb];
    float wa = xa4.w, wb = xb4.w;
    float wsum = wa + wb;
    if (wsum == 0.0f) return;
    f3 d = xyz(xa4) - xyz(xb4);
    float L = length(d);
    if (L < XPBD_EPS) return;
    f3 n_hat = d / L;
    float C = L - p0;
    if (C >= 0.0f) return;
    float dlam = -C / wsum;
    LAM[j] = lam + dlam;
    f3 xa = xyz(xa4) + n_hat * (wa * dlam);
    f3 xb = xyz(xb4) - n_hat * (wb * dlam);
    float mu = p1;
    if (mu > 0.0f) {
      f3 dp_rel = (xa - xyz(XP[a])) - (xb - xyz(XP[bb]));
      f3 corr = friction_delta(dp_rel, n_hat, -C, mu);
      xa = xa + corr * (wa / wsum);
      xb = xb - corr * (wb / wsum);
    }
    X[a] = with_xyz(xa4, xa);
    X[bb] = with_xyz(xb4, xb);
    return;
  }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_gs_color(float4* __restrict__ X,
                                                    const float4* __restrict__ XP,
                                                    const int4* __restrict__ R,
                                                    const float4* __restrict__ ANC,
                                                    float* __restrict__ LAM, float h, int first,
                                                    int start, int count) {
  const int tid = xcd_tile(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
  if (tid >= count) return;
  gs_row(X, XP, R, ANC, LAM, 1.0f / (h * h), first, start + tid);
}

// ---------------------------------------------------------------------------
// Whole frames of a small scene in one workgroup (solver.py:412-467).
//
// A scene of a few thousand particles leaves a one-launch-per-phase frame
// bound by launch latency (a 64-link chain in GS mode is 90 launches a
// frame).  Here one workgroup keeps the positions in LDS for the whole run
// and separates the phases with workgroup barriers: integrate, then per
// iteration either the color slices (GS, in place, gs_row) or a Jacobi pass
// (the fused tile kernel's row arithmetic and exact fixed-point sums, then
// the averaged apply), then finalize + cap.  Velocities, x_prev and lambda
// stay in global memory (each is touched by one thread per phase).  The
// arithmetic is the launch path's, so GS results are identical bit for bit
// and Jacobi results equal the tile kernel's.
// ---------------------------------------------------------------------------
template <bool GS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_frame_block(
    float4* __restrict__ X, float4* __restrict__ V, float4* __restrict__ XP,
    const int4* __restrict__ R, const float4* __restrict__ ANC, float* __restrict__ LAM,
    const int* __restrict__ color_starts, int num_colors, int n, int m, int substeps,
    int iterations, float h, float gx, float gy, float gz, float damping, float relax,
    float max_speed, int do_cap, int frames, const int* __restrict__ order,
    const int* __restrict__ num_colors_dev) {
  extern __shared__ unsigned long long lds64[];
  unsigned long long* sx = lds64;
  unsigned long long* sy = lds64 + n;
  unsigned long long* sz = lds64 + 2 * n;
  unsigned* cnt = (unsigned*)(lds64 + 3 * n);
  float4* spos = GS ? (float4*)lds64
                    : (float4*)(lds64 + tile_lds_pos_offset(n) / sizeof(unsigned long long));
  const int tid = threadIdx.x;
  const float inv_dt2 = 1.0f / (h * h);
  // GS colours: a count on the device (rows coloured on the device; < 0 =
  // colouring failed, no row is projected) and the rows of colour c at
  // order[color_starts[c] ..] (NULL: the rows themselves, sorted by colour)
  const int ncol = num_colors_dev ? *num_colors_dev : num_colors;
  for (int i = tid; i < n; i += BLOCK) spos[i] = X[i];
  __syncthreads();
  auto pos = [&](int i) { return spos[i]; };
  auto prev = [&](int i) { return xyz(XP[i]); };
  for (int f = 0; f < frames; ++f) {
    for (int sub = 0; sub < substeps; ++sub) {
      // integrate (kernels.py:49-73)
      for (int i = tid; i < n; i += BLOCK) {
        const float4 x4 = spos[i];
        f3 x = xyz(x4);
        XP[i] = make_float4(x.x, x.y, x.z, 0.0f);
        if (x4.w != 0.0f) {
          const f3 v = xyz(V[i]) + mk3(gx, gy, gz) * h;
          x = x + v * h;
          V[i] = make_float4(v.x, v.y, v.z, 0.0f);
          spos[i] = make_float4(x.x, x.y, x.z, x4.w);
        }
      }
      __syncthreads();
      for (int it = 0; it < iterations && m > 0; ++it) {
        const int first = it == 0;
        if (GS) {
          for (int c = 0; c < ncol; ++c) {
            const int c0 = color_starts[c], c1 = color_starts[c + 1];
            for (int k = c0 + tid; k < c1; k += BLOCK)
              gs_row(spos, XP, R, ANC, LAM, inv_dt2, first, order ? order[k] : k);
            __syncthreads();
          }
        } else {
          for (int i = tid; i < n; i += BLOCK) {
            sx[i] = 0ull;
            sy[i] = 0ull;
            sz[i] = 0ull;
            cnt[i] = 0u;
          }
          __syncthreads();
          for (int j = tid; j < m; j += BLOCK) {
            const float lam = first ? 0.0f : LAM[j];
            const Proj p = project_row<true>(pos, prev, [&]() { return ANC[j]; }, R[j], lam,
                                             inv_dt2);
            LAM[j] = p.lam;
            auto add = [&](int i, f3 c) {
              const bool ok =
                  fabsf(c.x) <= kFixLimit && fabsf(c.y) <= kFixLimit && fabsf(c.z) <= kFixLimit;
              atomicAdd(&sx[i], fix_bits(fminf(fmaxf(c.x, -kFixLimit), kFixLimit)));
              atomicAdd(&sy[i], fix_bits(fminf(fmaxf(c.y, -kFixLimit), kFixLimit)));
              atomicAdd(&sz[i], fix_bits(fminf(fmaxf(c.z, -kFixLimit), kFixLimit)));
              atomicAdd(&cnt[i], ok ? 1u : (kFixPoison | 1u));
            };
            if (p.na >= 1) add(p.a, p.ca);
            if (p.na == 2) add(p.b, p.cb);
          }
          __syncthreads();
          for (int i = tid; i < n; i += BLOCK) {
            const float4 x4 = spos[i];
            const unsigned c = cnt[i];
            const unsigned adds = c & (kFixPoison - 1u);
            float4 a = make_float4(0.f, 0.f, 0.f, (float)adds);
            if (c) a = make_float4(unfix(sx[i], adds), unfix(sy[i], adds), unfix(sz[i], adds), a.w);
            if (c >= kFixPoison) a.x = __int_as_float(0x7fc00000);
            if (a.w > 0.0f && x4.w > 0.0f) {
              const f3 x = xyz(x4) + xyz(a) * (relax / a.w);
              spos[i] = make_float4(x.x, x.y, x.z, x4.w);
            }
          }
          __syncthreads();
        }
      }
      // finalize_velocity + cap_velocity (kernels.py:346-371), per particle
      for (int i = tid; i < n; i += BLOCK) {
        const float4 x4 = spos[i];
        f3 v;
        if (x4.w == 0.0f) {
          v = mk3(0.0f, 0.0f, 0.0f);
        } else {
          v = ((xyz(x4) - xyz(XP[i])) / h) * (1.0f - damping);
        }
        if (do_cap) {
          const float speed = length(v);
          if (speed > max_speed) v = v * (max_speed / speed);
        }
        V[i] = make_float4(v.x, v.y, v.z, 0.0f);
      }
      __syncthreads();
    }
  }
  for (int i = tid; i < n; i += BLOCK) X[i] = spos[i];
}

// ---------------------------------------------------------------------------
// Per-particle streaming pass.  One kernel covers, in this order and each
// behind a flag: apply_jacobi (kernels.py:326-339), finalize_velocity
// (346-358), cap_velocity (362-371) and the next substep's integrate (49-73).
// Fusing them is exact: each is a pure per-particle map.
// ---------------------------------------------------------------------------
enum : int { F_APPLY = 1, F_FINALIZE = 2, F_CAP = 4, F_INTEGRATE = 8 };

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_particles(float4* __restrict__ X, float4* __restrict__ V,
                                                     float4* __restrict__ XP,
                                                     float4* __restrict__ ACC, int flags,
                                                     float relax, float h, float damping,
                                                     float max_speed, float gx, float gy, float gz,
                                                     int n) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  float4 x4 = X[i];
  const float w = x4.w;
  f3 x = xyz(x4);
  bool x_dirty = false;
  if (flags & F_APPLY) {
    float4 acc = ACC[i];
    float c = acc.w;
    if (c > 0.0f && w > 0.0f) {
      x = x + xyz(acc) * (relax / c);
      x_dirty = true;
    }
    ACC[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  f3 v;
  bool v_loaded = false;
  f3 xp;
  if (flags & F_FINALIZE) {
    xp = xyz(XP[i]);
    if (w == 0.0f) {
      v = mk3(0.0f, 0.0f, 0.0f);
    } else {
      v = ((x - xp) / h) * (1.0f - damping);
    }
    v_loaded = true;
    if (flags & F_CAP) {
      float speed = length(v);
      if (speed > max_speed) v = v * (max_speed / speed);
    }
  }
  if (flags & F_INTEGRATE) {
    XP[i] = make_float4(x.x, x.y, x.z, 0.0f);
    if (w != 0.0f) {
      if (!v_loaded) {
        v = xyz(V[i]);
        v_loaded = true;
      }
      v = v + mk3(gx, gy, gz) * h;
      x = x + v * h;
      x_dirty = true;
    }
  }
  if (v_loaded) V[i] = make_float4(v.x, v.y, v.z, 0.0f);
  if (x_dirty) X[i] = make_float4(x.x, x.y, x.z, w);
}

constexpr int kBlock = 256;
constexpr int kTileBlock = 512;

inline int blocks_for(long long n, int per_block) { return (int)((n + per_block - 1) / per_block); }

template <int BLOCK, int MODE>
void launch_tile_mode(hipStream_t s, int nb, size_t shmem, const float* Xin, float* Xout,
                      float* XP, float* V, const int* R, const float* ANC, const float* LAMin,
                      float* LAMout, const int* trs, const int* hs, const int* hr, float h,
                      int first, int tile_size, int n, float relax, float damping,
                      float max_speed, int do_cap, float* ACC, float gx, float gy, float gz,
                      const float* gscale) {
  hipLaunchKernelGGL((k_jacobi_tile<BLOCK, MODE>), dim3(nb), dim3(BLOCK), shmem, s,
                     (const float4*)Xin, (float4*)Xout, (float4*)XP, (float4*)V, (const int4*)R,
                     (const float4*)ANC, LAMin, LAMout, trs, hs, hr, h, first, tile_size, n, relax,
                     (MODE & TM_FINALIZE) ? 1 : 0, damping, max_speed, do_cap, (float4*)ACC, gx,
                     gy, gz, gscale);
}

template <int BLOCK>
int launch_tile_block(int mode, hipStream_t s, int nb, size_t shmem, const float* Xin,
                      float* Xout, float* XP, float* V, const int* R, const float* ANC,
                      const float* LAMin, float* LAMout, const int* trs, const int* hs,
                      const int* hr, float h, int first, int tile_size, int n, float relax,
                      float damping, float max_speed, int do_cap, float* ACC, float gx, float gy,
                      float gz, const float* gscale) {
#define XPBD_TILE_MODE(M)                                                                      \
  case M:                                                                                      \
    launch_tile_mode<BLOCK, M>(s, nb, shmem, Xin, Xout, XP, V, R, ANC, LAMin, LAMout, trs, hs, \
                               hr, h, first, tile_size, n, relax, damping, max_speed, do_cap,  \
                               ACC, gx, gy, gz, gscale);                                      \
    return 0;
  switch (mode) {
    XPBD_TILE_MODE(0)
    XPBD_TILE_MODE(TM_PREDICT)
    XPBD_TILE_MODE(TM_FINALIZE)
    XPBD_TILE_MODE(TM_ACC)
    XPBD_TILE_MODE(TM_FINALIZE | TM_ACC)
    XPBD_TILE_MODE(TM_ANCHOR)
    XPBD_TILE_MODE(TM_PREDICT | TM_ANCHOR)
    XPBD_TILE_MODE(TM_PREDICT | TM_ACC)
    XPBD_TILE_MODE(TM_PREDICT | TM_ACC | TM_ANCHOR)
    XPBD_TILE_MODE(TM_FINALIZE | TM_ANCHOR)
    XPBD_TILE_MODE(TM_ACC | TM_ANCHOR)
    XPBD_TILE_MODE(TM_FINALIZE | TM_ACC | TM_ANCHOR)
    XPBD_TILE_MODE(TM_FINALIZE | TM_NEXT)
    XPBD_TILE_MODE(TM_FINALIZE | TM_NEXT | TM_ACC)
    XPBD_TILE_MODE(TM_FINALIZE | TM_NEXT | TM_ANCHOR)
    XPBD_TILE_MODE(TM_FINALIZE | TM_NEXT | TM_ACC | TM_ANCHOR)
    XPBD_TILE_MODE(TM_FIRST)
    XPBD_TILE_MODE(TM_FIRST | TM_ACC)
    XPBD_TILE_MODE(TM_FIRST | TM_ANCHOR)
    XPBD_TILE_MODE(TM_FIRST | TM_ACC | TM_ANCHOR)
  }
#undef XPBD_TILE_MODE
  return (int)hipErrorInvalidValue;  // predict never combines with finalize
}

// ANC may be NULL when no row is ATTACH, ACC when no device-counted rows
// join the iteration.
// finalize: 0 no, 1 finalize + cap, 2 also the next substep's integrate.
int launch_tile(bool predict, hipStream_t s, const float* Xin, float* Xout, float* XP, float* V,
                const int* R, const float* ANC, const float* LAMin, float* LAMout,
                const int* tile_row_start, const int* halo_start, const int* halo_rows, float h,
                int first, int tile_size, int n, float relax, int finalize, float damping,
                float max_speed, float* ACC, float gx, float gy, float gz, const float* gscale) {
  if (n <= 0) return 0;
  if (tile_size < 1 || tile_size > XPBD_MAX_TILE) return (int)hipErrorInvalidValue;
  const int do_cap = (max_speed > 0.0f && std::isfinite(max_speed)) ? 1 : 0;
  const int nb = blocks_for(n, tile_size);
  const size_t shmem = tile_lds_bytes(tile_size);
  // a first, non-finalizing launch on integrated positions: lambda = 0 as
  // a template argument (the finalizing one keeps the runtime flag)
  const bool zero_lam = first && !predict && !finalize;
  const int mode = (predict ? TM_PREDICT : 0) | (finalize ? TM_FINALIZE : 0) |
                   (finalize == 2 ? TM_NEXT : 0) | (ACC ? TM_ACC : 0) | (ANC ? TM_ANCHOR : 0) |
                   (zero_lam ? TM_FIRST : 0);
  int rc;
  // 512 threads from 64-particle tiles up: a 64-particle tile of the 128^2
  // cloth has ~700 rows (owned + halo), one pass of 512 threads x 2 rows
  // (measured: 0.177 against 0.203 ms per frame with 256 threads)
  if (tile_size < 64)
    rc = launch_tile_block<kBlock>(mode, s, nb, shmem, Xin, Xout, XP, V, R, ANC, LAMin, LAMout,
                                   tile_row_start, halo_start, halo_rows, h, first, tile_size, n,
                                   relax, damping, max_speed, do_cap, ACC, gx, gy, gz, gscale);
  else
    rc = launch_tile_block<kTileBlock>(mode, s, nb, shmem, Xin, Xout, XP, V, R, ANC, LAMin,
                                       LAMout, tile_row_start, halo_start, halo_rows, h, first,
                                       tile_size, n, relax, damping, max_speed, do_cap, ACC, gx,
                                       gy, gz, gscale);
  if (rc) return rc;
  XPBD_LAUNCH_CHECK();
  return 0;
}



// Launches of the compact-table kernel: every mode, with and without rows
// read whole (GEN) and per-particle floors (PF).
template <int MODE, bool GEN, bool PF>
void launch_cx_mode(hipStream_t s, int nb, size_t shmem, const float* Xin, float* Xout,
                    float* XP, float* V, const int* R, const float* ANC, const float* LAMin,
                    float* LAMout, const int* trs, const int* hs, const xpbd_tile_compact* cx,
                    float h, int P, int n, float relax, float damping, float max_speed,
                    int do_cap, int stream_out, float* ACC, float gx, float gy, float gz,
                    const float* gscale) {
  auto k = stream_out ? k_jacobi_tile_cx<MODE, GEN, PF, true>
                      : k_jacobi_tile_cx<MODE, GEN, PF, false>;
  hipLaunchKernelGGL(k, dim3(nb), dim3(512), shmem, s,
                     (const float4*)Xin, (float4*)Xout, (float4*)XP, (float4*)V, (const int4*)R,
                     (const float4*)ANC, LAMin, LAMout, trs, hs, cx->dist_end, cx->gen_end,
                     (const int2*)cx->ci, cx->cc, (const int2*)cx->halo2,
                     (const float2*)cx->pf, cx->pflam, h, P, n, relax, damping, max_speed,
                     do_cap, (float4*)ACC, gx, gy, gz, gscale, cx->acc_tiles);
}

template <bool GEN, bool PF>
int launch_cx_gp(int mode, hipStream_t s, int nb, size_t shmem, const float* Xin, float* Xout,
                 float* XP, float* V, const int* R, const float* ANC, const float* LAMin,
                 float* LAMout, const int* trs, const int* hs, const xpbd_tile_compact* cx,
                 float h, int P, int n, float relax, float damping, float max_speed, int do_cap,
                 int stream_out, float* ACC, float gx, float gy, float gz, const float* gscale) {
#define XPBD_CX_MODE(M)                                                                        \
  case M:                                                                                      \
    launch_cx_mode<M, GEN, PF>(s, nb, shmem, Xin, Xout, XP, V, R, ANC, LAMin, LAMout, trs, hs, \
                               cx, h, P, n, relax, damping, max_speed, do_cap, stream_out,     \
                               ACC, gx, gy, gz, gscale);                                      \
    return 0;
#define XPBD_CX_BASES(X)                                                                       \
  XPBD_CX_MODE(0 | (X))                                                                        \
  XPBD_CX_MODE(TM_FIRST | (X))                                                                 \
  XPBD_CX_MODE(TM_FINALIZE | (X))                                                              \
  XPBD_CX_MODE(TM_FINALIZE | TM_FIRST | (X))                                                   \
  XPBD_CX_MODE(TM_FINALIZE | TM_NEXT | (X))                                                    \
  XPBD_CX_MODE(TM_FINALIZE | TM_NEXT | TM_FIRST | (X))                                         \
  XPBD_CX_MODE(TM_PREDICT | (X))
  switch (mode) {
    XPBD_CX_BASES(0)
    XPBD_CX_BASES(TM_ACC)
  }
  if constexpr (GEN) {
    switch (mode) {
      XPBD_CX_BASES(TM_ANCHOR)
      XPBD_CX_BASES(TM_ANCHOR | TM_ACC)
    }
  }
#undef XPBD_CX_BASES
#undef XPBD_CX_MODE
  return (int)hipErrorInvalidValue;
}

// The compact-table form of launch_tile (tile_size >= 64: 512-thread blocks).
int launch_tile_cx(bool predict, hipStream_t s, const float* Xin, float* Xout, float* XP,
                   float* V, const int* R, const float* ANC, const float* LAMin, float* LAMout,
                   const int* tile_row_start, const int* halo_start, const xpbd_tile_compact* cx,
                   float h, int first, int tile_size, int n, float relax, int finalize,
                   float damping, float max_speed, float* ACC, float gx, float gy, float gz,
                   const float* gscale) {
  if (n <= 0) return 0;
  if (!cx || tile_size < 64 || tile_size > XPBD_MAX_TILE || !cx->dist_end || !cx->ci ||
      !cx->cc || !cx->halo2 || (cx->pf && (!cx->gen_end || !cx->pflam)))
    return (int)hipErrorInvalidValue;
  const bool gen = cx->general != 0;
  const bool pf = cx->pf != nullptr;
  const int do_cap = (max_speed > 0.0f && std::isfinite(max_speed)) ? 1 : 0;
  const int nb = blocks_for(n, tile_size);
  const size_t shmem = tile_lds_bytes(tile_size);
  // ATTACH rows are read whole: without such rows no anchor is read
  const int mode = (predict ? TM_PREDICT : 0) | (finalize ? TM_FINALIZE : 0) |
                   (finalize == 2 ? TM_NEXT : 0) | (ACC ? TM_ACC : 0) |
                   (ANC && gen ? TM_ANCHOR : 0) | (first && !predict ? TM_FIRST : 0);
  const int so = n >= kStreamOutParticles ? 1 : 0;
  int rc;
#define XPBD_CX_ARGS                                                                           \
  mode, s, nb, shmem, Xin, Xout, XP, V, R, ANC, LAMin, LAMout, tile_row_start, halo_start, cx, \
      h, tile_size, n, relax, damping, max_speed, do_cap, so, ACC, gx, gy, gz, gscale
  if (gen)
    rc = pf ? launch_cx_gp<true, true>(XPBD_CX_ARGS) : launch_cx_gp<true, false>(XPBD_CX_ARGS);
  else
    rc = pf ? launch_cx_gp<false, true>(XPBD_CX_ARGS) : launch_cx_gp<false, false>(XPBD_CX_ARGS);
#undef XPBD_CX_ARGS
  if (rc) return rc;
  XPBD_LAUNCH_CHECK();
  return 0;
}

}  // namespace

// ===========================================================================
// C-ABI entry points
// ===========================================================================
extern "C" {

constexpr size_t kPairLdsMax = 64 * 1024;  // dynamic LDS of one pair launch, at most
constexpr int kPairBlock = 1024;           // threads per pair-launch block

size_t xpbd_jacobi_tile_pair_lds_bytes(int lmax, int amax, int k1max) {
  return tile_pair_lds_bytes(lmax, amax, k1max);
}

int xpbd_jacobi_tile_pair(void* stream, const float* Xin, float* Xout, const float* XPin,
                          float* XPout, float* V, const int* R, const float* ANC, float* LAMout,
                          const int* ring_start, const int* ring1_n, const int* ring_idx,
                          const int* p1_start, const int* p1, const int* p2_start, const int* p2,
                          float h, int tile_size, int n, float relax, float damping,
                          float max_speed, float gx, float gy, float gz, const float* gscale,
                          int lmax, int amax, int k1max, int last, int predict) {
  return xpbd_jacobi_tile_pair_warm(stream, Xin, Xout, XPin, XPout, V, R, ANC, LAMout,
                                    ring_start, ring1_n, ring_idx, p1_start, p1, p2_start, p2, h,
                                    tile_size, n, relax, damping, max_speed, gx, gy, gz, gscale,
                                    lmax, amax, k1max, last, predict, nullptr, nullptr, 0);
}

int xpbd_jacobi_tile_pair_warm(void* stream, const float* Xin, float* Xout, const float* XPin,
                               float* XPout, float* V, const int* R, const float* ANC,
                               float* LAMout, const int* ring_start, const int* ring1_n,
                               const int* ring_idx, const int* p1_start, const int* p1,
                               const int* p2_start, const int* p2, float h, int tile_size, int n,
                               float relax, float damping, float max_speed, float gx, float gy,
                               float gz, const float* gscale, int lmax, int amax, int k1max,
                               int last, int predict, const unsigned short* Y, const int* ws_on,
                               int invert) {
  if (n <= 0) return 0;
  if (tile_size < 1 || tile_size > XPBD_MAX_TILE || amax < tile_size || lmax < amax ||
      k1max < 0 || !Xin || !Xout || !XPin || !XPout || !R || !LAMout || !ring_start ||
      !ring1_n || !ring_idx || !p1_start || !p1 || !p2_start || !p2 || ((last || predict) && !V) ||
      (last && predict) || Xin == Xout || (XPin == XPout && !last))
    return (int)hipErrorInvalidValue;
  const size_t shmem = tile_pair_lds_bytes(lmax, amax, k1max);
  if (shmem > kPairLdsMax) return (int)hipErrorInvalidValue;
  const int do_cap = (max_speed > 0.0f && std::isfinite(max_speed)) ? 1 : 0;
  const int nb = blocks_for(n, tile_size);
  hipStream_t s = (hipStream_t)stream;
#define XPBD_PAIR_LAUNCH(L, PR, W)                                                            \
  hipLaunchKernelGGL((k_jacobi_tile_pair<kPairBlock, L, PR, W>), dim3(nb), dim3(kPairBlock),   \
                     shmem, s, (const float4*)Xin, (float4*)Xout, (const float4*)XPin,         \
                     (float4*)XPout, (float4*)V, (const int4*)R, (const float4*)ANC, LAMout,   \
                     ring_start, ring1_n, ring_idx, p1_start, (const int4*)p1, p2_start, p2, h, \
                     tile_size, n, relax, damping, max_speed, do_cap, gx, gy, gz, gscale, lmax, \
                     amax, k1max, Y, ws_on, invert)
  if (Y) {
    if (last)
      XPBD_PAIR_LAUNCH(true, false, true);
    else if (predict)
      XPBD_PAIR_LAUNCH(false, true, true);
    else
      XPBD_PAIR_LAUNCH(false, false, true);
  } else {
    if (last)
      XPBD_PAIR_LAUNCH(true, false, false);
    else if (predict)
      XPBD_PAIR_LAUNCH(false, true, false);
    else
      XPBD_PAIR_LAUNCH(false, false, false);
  }
#undef XPBD_PAIR_LAUNCH
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd_particles_pass(void* stream, float* X, float* V, float* XP, float* ACC, int flags,
                        float relax, float h, float damping, float max_speed, float gx, float gy,
                        float gz, int n) {
  if (n <= 0) return 0;
  if ((flags & F_APPLY) && ACC == nullptr) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_particles<kBlock>, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s,
                     (float4*)X, (float4*)V, (float4*)XP, (float4*)ACC, flags, relax, h, damping,
                     max_speed, gx, gy, gz, n);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd_integrate(void* stream, float* X, float* V, float* XP, float gx, float gy, float gz,
                   float h, int n) {
  return xpbd_particles_pass(stream, X, V, XP, nullptr, F_INTEGRATE, 1.0f, h, 0.0f, 0.0f, gx, gy,
                             gz, n);
}

int xpbd_apply_jacobi(void* stream, float* X, float* ACC, float relax, int n) {
  return xpbd_particles_pass(stream, X, nullptr, nullptr, ACC, F_APPLY, relax, 1.0f, 0.0f, 0.0f,
                             0.0f, 0.0f, 0.0f, n);
}

int xpbd_finalize_velocity(void* stream, const float* X, const float* XP, float* V, float h,
                           float damping, float max_speed, int n) {
  int flags = F_FINALIZE;
  if (max_speed > 0.0f && std::isfinite(max_speed)) flags |= F_CAP;
  return xpbd_particles_pass(stream, (float*)X, V, (float*)XP, nullptr, flags, 1.0f, h, damping,
                             max_speed, 0.0f, 0.0f, 0.0f, n);
}

int xpbd_solve_jacobi(void* stream, const float* X, const float* XP, const int* R,
                      const float* ANC, float* LAM, float* ACC, float h, int first, int n_rows,
                      int rows_per_thread, int window) {
  if (n_rows <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  if (window <= 0) {
    hipLaunchKernelGGL(k_jacobi_project_global<kBlock>, dim3(blocks_for(n_rows, kBlock)),
                       dim3(kBlock), 0, s, (const float4*)X, (const float4*)XP, (const int4*)R,
                       (const float4*)ANC, LAM, (float4*)ACC, h, first, n_rows);
  } else {
    if (rows_per_thread < 1) rows_per_thread = 1;
    if (window > XPBD_MAX_WINDOW) return (int)hipErrorInvalidValue;
    int nb = blocks_for(n_rows, kBlock * rows_per_thread);
    size_t shmem = sizeof(float) * 4 * (size_t)window;
    hipLaunchKernelGGL(k_jacobi_project<kBlock>, dim3(nb), dim3(kBlock), shmem, s,
                       (const float4*)X, (const float4*)XP, (const int4*)R, (const float4*)ANC,
                       LAM, (float4*)ACC, h, first, n_rows, rows_per_thread, window);
  }
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd_jacobi_tile(void* stream, const float* Xin, float* Xout, const float* XP, float* V,
                     const int* R, const float* ANC, const float* LAMin, float* LAMout,
                     const int* tile_row_start, const int* halo_start, const int* halo_rows,
                     float h, int first, int tile_size, int n_particles, float relax,
                     int finalize, float damping, float max_speed, float* ACC) {
  // the plain launch only reads XP (friction, finalize)
  return launch_tile(false, (hipStream_t)stream, Xin, Xout, const_cast<float*>(XP), V, R, ANC,
                     LAMin, LAMout, tile_row_start, halo_start, halo_rows, h, first, tile_size,
                     n_particles, relax, finalize ? 1 : 0, damping, max_speed, ACC, 0.0f, 0.0f,
                     0.0f, nullptr);
}

int xpbd_jacobi_tile_next(void* stream, const float* Xin, float* Xout, float* XP, const int* R,
                          const float* ANC, const float* LAMin, float* LAMout,
                          const int* tile_row_start, const int* halo_start, const int* halo_rows,
                          float h, int first, float gx, float gy, float gz, const float* gscale,
                          int tile_size, int n_particles, float relax, float damping,
                          float max_speed, float* ACC) {
  if (Xin == Xout) return (int)hipErrorInvalidValue;  // neighbours read Xin while tiles write
  return launch_tile(false, (hipStream_t)stream, Xin, Xout, XP, nullptr, R, ANC, LAMin, LAMout,
                     tile_row_start, halo_start, halo_rows, h, first, tile_size, n_particles,
                     relax, 2, damping, max_speed, ACC, gx, gy, gz, gscale);
}

int xpbd_jacobi_tile_first(void* stream, const float* X, float* Xout, float* XP, const float* V,
                           const int* R, const float* ANC, float* LAMout,
                           const int* tile_row_start, const int* halo_start, const int* halo_rows,
                           float h, float gx, float gy, float gz, const float* gscale,
                           int tile_size, int n_particles, float relax, float* ACC) {
  if (X == Xout) return (int)hipErrorInvalidValue;  // neighbours read X while tiles write
  return launch_tile(true, (hipStream_t)stream, X, Xout, XP, const_cast<float*>(V), R, ANC,
                     nullptr, LAMout, tile_row_start, halo_start, halo_rows, h, 1, tile_size,
                     n_particles, relax, 0, 0.0f, 0.0f, ACC, gx, gy, gz, gscale);
}

int xpbd_jacobi_tile_cx(void* stream, const float* Xin, float* Xout, const float* XP, float* V,
                        const int* R, const float* ANC, const float* LAMin, float* LAMout,
                        const int* tile_row_start, const int* halo_start,
                        const xpbd_tile_compact* cx, float h, int first, int tile_size,
                        int n_particles, float relax, int finalize, float damping,
                        float max_speed, float* ACC) {
  // the compact loop's multiplier loads are unconditional in a launch with
  // lambda from LAMin: a zero start never reads it
  return launch_tile_cx(false, (hipStream_t)stream, Xin, Xout, const_cast<float*>(XP), V, R,
                        ANC, LAMin, LAMout, tile_row_start, halo_start, cx, h, first, tile_size,
                        n_particles, relax, finalize ? 1 : 0, damping, max_speed, ACC, 0.0f,
                        0.0f, 0.0f, nullptr);
}

int xpbd_jacobi_tile_next_cx(void* stream, const float* Xin, float* Xout, float* XP,
                             const int* R, const float* ANC, const float* LAMin, float* LAMout,
                             const int* tile_row_start, const int* halo_start,
                             const xpbd_tile_compact* cx, float h, int first, float gx,
                             float gy, float gz, const float* gscale, int tile_size,
                             int n_particles, float relax, float damping, float max_speed,
                             float* ACC) {
  if (Xin == Xout) return (int)hipErrorInvalidValue;  // neighbours read Xin while tiles write
  return launch_tile_cx(false, (hipStream_t)stream, Xin, Xout, XP, nullptr, R, ANC, LAMin,
                        LAMout, tile_row_start, halo_start, cx, h, first, tile_size,
                        n_particles, relax, 2, damping, max_speed, ACC, gx, gy, gz, gscale);
}

int xpbd_jacobi_tile_first_cx(void* stream, const float* X, float* Xout, float* XP,
                              const float* V, const int* R, const float* ANC, float* LAMout,
                              const int* tile_row_start, const int* halo_start,
                              const xpbd_tile_compact* cx, float h, float gx, float gy, float gz,
                              const float* gscale, int tile_size, int n_particles, float relax,
                              float* ACC) {
  if (X == Xout) return (int)hipErrorInvalidValue;  // neighbours read X while tiles write
  return launch_tile_cx(true, (hipStream_t)stream, X, Xout, XP, const_cast<float*>(V), R, ANC,
                        nullptr, LAMout, tile_row_start, halo_start, cx, h, 1, tile_size,
                        n_particles, relax, 0, 0.0f, 0.0f, ACC, gx, gy, gz, gscale);
}

int xpbd_jacobi_tile_warm(void* stream, const float* Xin, float* Xout, float* XP, const float* V,
                          const int* R, const float* ANC, float* LAMout,
                          const int* tile_row_start, const int* halo_start, const int* halo_rows,
                          const unsigned short* Y, float h, float gx, float gy, float gz,
                          const float* gscale, int tile_size, int n_particles, float relax,
                          int predict, const int* ws_on, int invert) {
  if (n_particles <= 0) return 0;
  if (tile_size < 1 || tile_size > XPBD_MAX_TILE || !Y || Xin == Xout)
    return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  const int nb = blocks_for(n_particles, tile_size);
  const size_t shmem = tile_lds_bytes(tile_size);
#define XPBD_WARM_LAUNCH(B, PR)                                                                \
  hipLaunchKernelGGL((k_jacobi_tile_warm<B, PR>), dim3(nb), dim3(B), shmem, s,                 \
                     (const float4*)Xin, (float4*)Xout, (float4*)XP, (const float4*)V,         \
                     (const int4*)R, (const float4*)ANC, LAMout, tile_row_start, halo_start,   \
                     halo_rows, Y, h, tile_size, n_particles, relax, gx, gy, gz, gscale, ws_on,  \
                     invert)
  if (tile_size < 64) {
    if (predict) XPBD_WARM_LAUNCH(kBlock, true);
    else XPBD_WARM_LAUNCH(kBlock, false);
  } else {
    if (predict) XPBD_WARM_LAUNCH(kTileBlock, true);
    else XPBD_WARM_LAUNCH(kTileBlock, false);
  }
#undef XPBD_WARM_LAUNCH
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd_jacobi_tile_warm_cx(void* stream, const float* Xin, float* Xout, float* XP,
                             const float* V, const int* R, const float* ANC, float* LAMout,
                             const int* tile_row_start, const int* halo_start,
                             const xpbd_tile_compact* cx, const unsigned short* Y, float h, float gx,
                             float gy, float gz, const float* gscale, int tile_size,
                             int n_particles, float relax, int predict, const int* ws_on,
                             int invert) {
  if (n_particles <= 0) return 0;
  if (!cx || !cx->dist_end || !cx->ci || !cx->cc || !cx->halo2 || tile_size < 64 ||
      tile_size > XPBD_MAX_TILE || !Y || Xin == Xout ||
      (cx->pf && (!cx->gen_end || !cx->pflam)))
    return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  const int nb = blocks_for(n_particles, tile_size);
  const size_t shmem = tile_lds_bytes(tile_size);
#define XPBD_WARM_CX(PR, GN, F)                                                               \
  hipLaunchKernelGGL((k_jacobi_tile_warm_cx<PR, GN, F>), dim3(nb), dim3(512), shmem, s,       \
                     (const float4*)Xin, (float4*)Xout, (float4*)XP, (const float4*)V,        \
                     (const int4*)R, (const float4*)ANC, LAMout, tile_row_start, halo_start,  \
                     cx->dist_end, cx->gen_end, (const int2*)cx->ci, cx->cc,                  \
                     (const int2*)cx->halo2, cx->pflam, Y, h, tile_size, n_particles, relax,  \
                     gx, gy, gz, gscale, ws_on, invert)
#define XPBD_WARM_CX_PF(PR, GN)                                                               \
  if (cx->pf) XPBD_WARM_CX(PR, GN, true);                                                     \
  else XPBD_WARM_CX(PR, GN, false);
  const bool gen = cx->general != 0;
  if (predict) {
    if (gen) { XPBD_WARM_CX_PF(true, true) }
    else { XPBD_WARM_CX_PF(true, false) }
  } else {
    if (gen) { XPBD_WARM_CX_PF(false, true) }
    else { XPBD_WARM_CX_PF(false, false) }
  }
#undef XPBD_WARM_CX_PF
#undef XPBD_WARM_CX
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd_integrate_to(void* stream, const float* Xin, float* Xout, float* V, float* XP, float gx,
                      float gy, float gz, float h, const float* gscale, int n) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_integrate_to<kBlock>, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0,
                     (hipStream_t)stream, (const float4*)Xin, (float4*)Xout, (float4*)V,
                     (float4*)XP, gx, gy, gz, h, gscale, n);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd_solve_jacobi_dyn(void* stream, const float* X, const float* XP, const int* R,
                          float* LAM, float* ACC, float h, int first, const int* n_rows_dev,
                          int cap) {
  if (cap <= 0) return 0;
  hipLaunchKernelGGL((k_jacobi_project_dyn<kBlock, false>), dim3(blocks_for(cap, kBlock)),
                     dim3(kBlock), 0, (hipStream_t)stream, (const float4*)X, (const float4*)XP,
                     nullptr, (const int4*)R, LAM, (float4*)ACC, h, first, 0.0f, 0.0f, 0.0f,
                     nullptr, n_rows_dev, cap);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd_solve_jacobi_dyn_first(void* stream, const float* X, const float* V, const int* R,
                                float* LAM, float* ACC, float h, float gx, float gy, float gz,
                                const float* gscale, const int* n_rows_dev, int cap) {
  if (cap <= 0) return 0;
  hipLaunchKernelGGL((k_jacobi_project_dyn<kBlock, true>), dim3(blocks_for(cap, kBlock)),
                     dim3(kBlock), 0, (hipStream_t)stream, (const float4*)X, nullptr,
                     (const float4*)V, (const int4*)R, LAM, (float4*)ACC, h, 1, gx, gy, gz,
                     gscale, n_rows_dev, cap);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd_frame_block_capacity(int gs) {
  // dynamic LDS within the default 64 KB: 16 B (GS) or 44 B (Jacobi) per particle
  return gs ? 4096 : (int)((65536 - 16) / (3 * sizeof(long long) + sizeof(unsigned) +
                                          sizeof(float4)));
}

int xpbd_frame_block(void* stream, float* X, float* V, float* XP, const int* R, const float* ANC,
                     float* LAM, const int* color_starts, int num_colors, int n_particles,
                     int n_rows, int substeps, int iterations, float h, float gx, float gy,
                     float gz, float damping, float relax, float max_speed, int gs, int frames) {
  if (n_particles <= 0 || frames <= 0) return 0;
  if (n_particles > xpbd_frame_block_capacity(gs) || substeps < 0 || iterations < 0 ||
      (gs && n_rows > 0 && (!color_starts || num_colors < 1)))
    return (int)hipErrorInvalidValue;
  constexpr int kWG = 1024;
  const int do_cap = (max_speed > 0.0f && std::isfinite(max_speed)) ? 1 : 0;
  hipStream_t s = (hipStream_t)stream;
  if (gs)
    hipLaunchKernelGGL((k_frame_block<true, kWG>), dim3(1), dim3(kWG),
                       sizeof(float4) * (size_t)n_particles, s, (float4*)X, (float4*)V,
                       (float4*)XP, (const int4*)R, (const float4*)ANC, LAM, color_starts,
                       num_colors, n_particles, n_rows, substeps, iterations, h, gx, gy, gz,
                       damping, relax, max_speed, do_cap, frames, nullptr, nullptr);
  else
    hipLaunchKernelGGL((k_frame_block<false, kWG>), dim3(1), dim3(kWG),
                       tile_lds_bytes(n_particles), s, (float4*)X, (float4*)V, (float4*)XP,
                       (const int4*)R, (const float4*)ANC, LAM, color_starts, num_colors,
                       n_particles, n_rows, substeps, iterations, h, gx, gy, gz, damping, relax,
                       max_speed, do_cap, frames, nullptr, nullptr);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd_frame_block_ordered(void* stream, float* X, float* V, float* XP, const int* R,
                             const float* ANC, float* LAM, const int* color_starts,
                             const int* order, const int* num_colors_dev, int n_particles,
                             int substeps, int iterations, float h, float gx, float gy, float gz,
                             float damping, float relax, float max_speed, int frames) {
  if (n_particles <= 0 || frames <= 0) return 0;
  if (n_particles > xpbd_frame_block_capacity(1) || substeps < 0 || iterations < 0 ||
      !color_starts || !order || !num_colors_dev)
    return (int)hipErrorInvalidValue;
  constexpr int kWG = 1024;
  const int do_cap = (max_speed > 0.0f && std::isfinite(max_speed)) ? 1 : 0;
  hipLaunchKernelGGL((k_frame_block<true, kWG>), dim3(1), dim3(kWG),
                     sizeof(float4) * (size_t)n_particles, (hipStream_t)stream, (float4*)X,
                     (float4*)V, (float4*)XP, (const int4*)R, (const float4*)ANC, LAM,
                     color_starts, 0, n_particles, 1, substeps, iterations, h, gx, gy, gz,
                     damping, relax, max_speed, do_cap, frames, order, num_colors_dev);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd_solve_color(void* stream, float* X, const float* XP, const int* R, const float* ANC,
                     float* LAM, float h, int first, int start, int count) {
  if (count <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_gs_color<kBlock>, dim3(blocks_for(count, kBlock)), dim3(kBlock), 0, s,
                     (float4*)X, (const float4*)XP, (const int4*)R, (const float4*)ANC, LAM, h,
                     first, start, count);
  XPBD_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
