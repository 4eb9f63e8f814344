// GPU sphere-sphere CONTACT generation for the 3-DOF solver — the device
// counterpart of Solver._rebuild_contacts (src/xpbd3d/solver.py:275-342)
// with the row order of its grid broad phase (_grid_candidate_pairs,
// solver.py:520-562):
//   margin = min(contact_margin + 3 dt max|v|, r_max)   (solver.py:309-316)
//   cell   = 2 r_max + margin                           (solver.py:324)
//   candidates: for each of the 14 cell offsets (the cell itself, then the
//   13 forward neighbours in the reference's order), for each body k in
//   ascending index, the bodies of cell(k) + offset in ascending index (the
//   cell itself: only those after k)
//   keep |x_k - x_p|^2 <= (r_k + r_p + margin)^2, skip pairs already joined
//   by a constraint, row = CONTACT(k, p, rest = r_k + r_p, mu = sqrt(mu_k mu_p))
// "Index" is the API particle index (`api`, when the device layout permutes
// the particles), so the rows come out in exactly the reference's order.
// Count per (offset, body), exclusive scan in offset-major order, emit; the
// live row count stays on the device and gates the Jacobi projection.  The
// bodies are the colliding ones only (clist: their device indices in
// ascending API index), so the counts, the scans and the cell table scale
// with the colliders, not with every particle of the scene.
#include <algorithm>
#include <cmath>

#include "common.h"
#include "scan.h"
#include "xpbd.h"

namespace {

// |v| max over all particles and components, as float bits (non-negative
// floats order like their int bits), plus the derived margin and cell.
__global__ __launch_bounds__(kB) void k_vmax(const float4* __restrict__ V, int n, int* vmax_bits) {
  __shared__ float sh[kB];
  float m = 0.0f;
  for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
    float4 v = V[i];
    m = fmaxf(m, fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z))));
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = kB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(vmax_bits, __float_as_int(sh[0]));
}

// params[0] = margin, params[1] = cell (float32, computed as the reference
// does in float64 host arithmetic, then rounded).
__global__ void k_contact_params(const int* vmax_bits, double contact_margin, double dt,
                                 float r_max, float* params) {
  double vmax = (double)__int_as_float(*vmax_bits);
  double margin = fmin(contact_margin + 3.0 * dt * vmax, (double)r_max);
  params[0] = (float)margin;
  params[1] = (float)((double)r_max * 2.0 + margin);
}

__host__ __device__ inline int chash(int cx, int cy, int cz, int T) {
  unsigned h = (unsigned)cx * 73856093u ^ (unsigned)cy * 19349663u ^ (unsigned)cz * 83492791u;
  return (int)(h & (unsigned)(T - 1));
}

// Body k of the candidate list: its device index (clist NULL: k itself).
__device__ inline int body_of(const int* clist, int k) { return clist ? clist[k] : k; }

__global__ __launch_bounds__(kB) void k_sphere_cells(const float4* __restrict__ X,
                                                     const float* __restrict__ radius,
                                                     const float* __restrict__ params, int T,
                                                     int4* cells, int* slot_count,
                                                     const int* __restrict__ clist, int nb) {
  const int k = blockIdx.x * kB + threadIdx.x;
  if (k >= nb) return;
  const int i = body_of(clist, k);
  if (!(radius[i] > 0.0f)) {
    cells[i] = make_int4(0, 0, 0, -1);
    return;
  }
  const float cell = params[1];
  const float4 x = X[i];
  const int cx = (int)floorf(x.x / cell), cy = (int)floorf(x.y / cell), cz = (int)floorf(x.z / cell);
  const int s = chash(cx, cy, cz, T);
  cells[i] = make_int4(cx, cy, cz, s);
  atomicAdd(&slot_count[s], 1);
}

// Bucket records in slot order, so that a candidate is one load of each:
// bkey = (cell x, y, z, API index), bpos = (x, y, z, radius), bref = (device
// index, friction bits).
__global__ __launch_bounds__(kB) void k_sphere_scatter(
    const int4* __restrict__ cells, const int* __restrict__ slot_start, int* cursor,
    const float4* __restrict__ X, const float* __restrict__ radius, const float* __restrict__ mu,
    const int* __restrict__ api, int4* __restrict__ bkey, float4* __restrict__ bpos,
    int2* __restrict__ bref, const int* __restrict__ clist, int nb) {
  const int k = blockIdx.x * kB + threadIdx.x;
  if (k >= nb) return;
  const int i = body_of(clist, k);
  const int4 c = cells[i];
  if (c.w < 0) return;
  const int p = slot_start[c.w] + atomicAdd(&cursor[c.w], 1);
  const float4 x = X[i];
  bkey[p] = make_int4(c.x, c.y, c.z, api ? api[i] : i);
  bpos[p] = make_float4(x.x, x.y, x.z, radius[i]);
  bref[p] = make_int2(i, mu ? __float_as_int(mu[i]) : 0);
}

constexpr int kOffsets = 14;

// Cell offset o of the reference's enumeration: the cell itself, then
// (1,-1,-1) .. (1,1,1) in x-major order, (0,1,-1), (0,1,0), (0,1,1), (0,0,1).
__device__ inline int3 cell_offset(int o) {
  if (o == 0) return make_int3(0, 0, 0);
  if (o <= 9) return make_int3(1, (o - 1) / 3 - 1, (o - 1) % 3 - 1);
  if (o <= 12) return make_int3(0, 1, o - 11);
  return make_int3(0, 0, 1);
}

// Pair (lo, hi) of API indices in the sorted key list (lo << 32 | hi)?
__device__ inline bool excluded(const long long* __restrict__ excl, int n_excl, int a, int b) {
  const int lo = min(a, b), hi = max(a, b);
  const long long key = ((long long)lo << 32) | (long long)(unsigned)hi;
  int l = 0, r = n_excl;
  while (l < r) {
    const int m = (l + r) >> 1;
    if (excl[m] < key) l = m + 1;
    else r = m;
  }
  return l < n_excl && excl[l] == key;
}

// One thread per (offset o, body k), offset-major.  Bodies in API order:
// with clist, body k (device index clist[k]) has rank k; without, device
// body i has rank api(i).  EMIT = false: cnt[o nb + rank] = number of rows
// the body contributes at offset o.  EMIT = true: those rows at
// off[o nb + rank], partners in ascending API index.
template <bool EMIT>
__global__ __launch_bounds__(kB) void k_sphere_query(
    const float4* __restrict__ X, const float* __restrict__ radius,
    const float* __restrict__ mu, const float* __restrict__ params,
    const int4* __restrict__ cells, const int* __restrict__ slot_start,
    const int4* __restrict__ bkey, const float4* __restrict__ bpos,
    const int2* __restrict__ bref, const int* __restrict__ api,
    const long long* __restrict__ excl, int n_excl, int T, const int* __restrict__ clist,
    int n, int* __restrict__ cnt, const int* __restrict__ off, int4* __restrict__ rows,
    int cap) {
  const long long t = (long long)blockIdx.x * kB + threadIdx.x;
  if (t >= (long long)kOffsets * n) return;
  const int o = (int)(t / n), k = (int)(t - (long long)o * n);
  const int i = body_of(clist, k);
  const int ai = api ? api[i] : i;
  const int rk = clist ? k : ai;
  const int4 ci = cells[i];
  if (ci.w < 0) {
    if (!EMIT) cnt[o * n + rk] = 0;
    return;
  }
  const float margin = params[0];
  const float4 xi = X[i];
  const float ri = radius[i];
  const int3 d = cell_offset(o);
  const int cx = ci.x + d.x, cy = ci.y + d.y, cz = ci.z + d.z;
  const int s = chash(cx, cy, cz, T);
  const int base = EMIT ? off[o * n + rk] : 0;
  int found = 0;
  const int pe = slot_start[s + 1];
  for (int p = slot_start[s]; p < pe; ++p) {
    const int4 key = bkey[p];
    const float4 xj = bpos[p];
    if (key.x != cx || key.y != cy || key.z != cz) continue;
    const int aj = key.w;
    if (aj == ai || (o == 0 && aj <= ai)) continue;
    const float d0 = xi.x - xj.x, d1 = xi.y - xj.y, d2 = xi.z - xj.z;
    const float dd = d0 * d0 + d1 * d1 + d2 * d2;
    const float rs = ri + xj.w + margin;
    if (!(dd <= rs * rs)) continue;
    if (n_excl > 0 && excluded(excl, n_excl, ai, aj)) continue;
    if (EMIT && base + found < cap) {
      const int2 ref = bref[p];
      const float rest = ri + xj.w;
      const float m = sqrtf(mu[i] * __int_as_float(ref.y));
      rows[base + found] =
          make_int4(i | (T_CONTACT << 28), ref.x, __float_as_int(rest), __float_as_int(m));
    }
    ++found;
  }
  if (EMIT) {  // partners in ascending API index (insertion sort of the segment)
    int4* seg = rows + base;
    const int m = min(found, max(cap - base, 0));
    for (int q = 1; q < m; ++q) {
      const int4 v = seg[q];
      const int kv = api ? api[v.y] : v.y;
      int r = q - 1;
      while (r >= 0 && (api ? api[seg[r].y] : seg[r].y) > kv) {
        seg[r + 1] = seg[r];
        --r;
      }
      seg[r + 1] = v;
    }
  } else {
    cnt[o * n + rk] = found;
  }
}

__global__ __launch_bounds__(kB) void k_fill(int* p, int v, int len) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i < len) p[i] = v;
}

int slots_for(int n) {
  int t = 1024;
  while (t < 2 * n) t <<= 1;
  return t;
}

struct Ws {
  int* vmax;
  float* params;
  int4* cells;
  int *slot_count, *slot_start, *cursor, *cnt, *off;
  int4* bkey;
  float4* bpos;
  int2* bref;
  ScanWs scan;
};

size_t carve(char* base, Ws& w, int n) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    off = (off + 255) & ~(size_t)255;
    char* p = base ? base + off : nullptr;
    off += bytes ? bytes : 1;
    return p;
  };
  const int T = slots_for(n);
  w.vmax = (int*)take(sizeof(int));
  w.params = (float*)take(2 * sizeof(float));
  w.cells = (int4*)take(sizeof(int4) * (size_t)n);
  w.slot_count = (int*)take(sizeof(int) * ((size_t)T + 1));
  w.slot_start = (int*)take(sizeof(int) * ((size_t)T + 1));
  w.cursor = (int*)take(sizeof(int) * (size_t)T);
  w.bkey = (int4*)take(sizeof(int4) * (size_t)n);
  w.bpos = (float4*)take(sizeof(float4) * (size_t)n);
  w.bref = (int2*)take(sizeof(int2) * (size_t)n);
  const long long m = (long long)kOffsets * n;
  w.cnt = (int*)take(sizeof(int) * ((size_t)m + 1));
  w.off = (int*)take(sizeof(int) * ((size_t)m + 1));
  w.scan.block_sums = (int*)take(sizeof(int) * scan_block_sums(std::max<long long>(T + 1, m + 1)));
  return off;
}

}  // namespace

extern "C" {

long long xpbd_contacts_workspace_bytes(int n) {
  Ws w;
  return (long long)carve(nullptr, w, n < 1 ? 1 : n) + 256;
}

int xpbd_sphere_contacts(void* stream, void* ws, const float* X, const float* V,
                         const float* radius, const float* mu, const int* api,
                         const int* clist, int nc, const long long* excl, int n_excl,
                         double contact_margin, double dt, float r_max, int n, int cap,
                         int* ROWS, int* count) {
  if (n_excl < 0 || (n_excl > 0 && !excl) || (long long)kOffsets * n >= (1ll << 31) ||
      (clist && (nc < 0 || nc > n)))
    return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  const int nbod = clist ? nc : n;  // bodies the candidate search visits
  if (nbod <= 1 || !(r_max > 0.0f)) {
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(kB), 0, s, count, 0, 1);
    XPBD_LAUNCH_CHECK();
    return 0;
  }
  Ws w;
  carve((char*)ws, w, n);  // sized for every particle: nbod <= n fits
  const int T = slots_for(nbod);
  const int nb = nblk(nbod, kB);
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(kB), 0, s, w.vmax, 0, 1);
  hipLaunchKernelGGL(k_fill, dim3(nblk(T + 1, kB)), dim3(kB), 0, s, w.slot_count, 0, T + 1);
  hipLaunchKernelGGL(k_fill, dim3(nblk(T, kB)), dim3(kB), 0, s, w.cursor, 0, T);
  // the margin uses the largest velocity of every particle (solver.py:309)
  hipLaunchKernelGGL(k_vmax, dim3(min(nblk(n, kB), 1024)), dim3(kB), 0, s, (const float4*)V, n,
                     w.vmax);
  hipLaunchKernelGGL(k_contact_params, dim3(1), dim3(1), 0, s, w.vmax, contact_margin, dt, r_max,
                     w.params);
  hipLaunchKernelGGL(k_sphere_cells, dim3(nb), dim3(kB), 0, s, (const float4*)X, radius,
                     w.params, T, w.cells, w.slot_count, clist, nbod);
  int rc = exclusive_scan(s, w.slot_count, w.slot_start, T + 1, w.scan, nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL(k_sphere_scatter, dim3(nb), dim3(kB), 0, s, w.cells, w.slot_start, w.cursor,
                     (const float4*)X, radius, mu, api, w.bkey, w.bpos, w.bref, clist, nbod);
  const int nq = nblk((long long)kOffsets * nbod, kB);
  hipLaunchKernelGGL(k_sphere_query<false>, dim3(nq), dim3(kB), 0, s, (const float4*)X, radius,
                     mu, w.params, w.cells, w.slot_start, w.bkey, w.bpos, w.bref, api, excl,
                     n_excl, T, clist, nbod, w.cnt, w.off, (int4*)ROWS, cap);
  rc = exclusive_scan(s, w.cnt, w.off, kOffsets * nbod, w.scan, count);
  if (rc) return rc;
  hipLaunchKernelGGL(k_sphere_query<true>, dim3(nq), dim3(kB), 0, s, (const float4*)X, radius,
                     mu, w.params, w.cells, w.slot_start, w.bkey, w.bpos, w.bref, api, excl,
                     n_excl, T, clist, nbod, w.cnt, w.off, (int4*)ROWS, cap);
  XPBD_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
