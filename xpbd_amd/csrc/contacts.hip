// GPU sphere-sphere CONTACT generation for the 3-DOF solver — the device
// counterpart of Solver._rebuild_contacts (src/xpbd3d/solver.py:275-342):
//   margin = min(contact_margin + 3 dt max|v|, r_max)   (solver.py:309-316)
//   cell   = 2 r_max + margin                           (solver.py:324)
//   keep pairs in the same / adjacent cells with |x_a - x_b|^2 <= (r_a + r_b + margin)^2
//   row    = CONTACT(a, b, rest = r_a + r_b, mu = sqrt(mu_a mu_b))  (solver.py:334-341)
// Pairs come out in (a, b) order with a < b (count, scan, emit, segment
// sort), written as packed rows into a caller buffer; the live row count
// stays on the device and gates the Jacobi projection of those rows.
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
__global__ void k_contact_params(const int* vmax_bits, float contact_margin, float dt,
                                 float r_max, float* params) {
  double vmax = (double)__int_as_float(*vmax_bits);
  double margin = fmin((double)contact_margin + 3.0 * (double)dt * vmax, (double)r_max);
  params[0] = (float)margin;
  params[1] = (float)((double)r_max * 2.0 + margin);
}

__host__ __device__ inline int chash(int cx, int cy, int cz, int T) {
  unsigned h = (unsigned)cx * 73856093u ^ (unsigned)cy * 19349663u ^ (unsigned)cz * 83492791u;
  return (int)(h & (unsigned)(T - 1));
}

__global__ __launch_bounds__(kB) void k_sphere_cells(const float4* __restrict__ X,
                                                     const float* __restrict__ radius,
                                                     const float* __restrict__ params, int T,
                                                     int4* cells, int* slot_count, int n) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
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

__global__ __launch_bounds__(kB) void k_sphere_scatter(const int4* __restrict__ cells,
                                                       const int* __restrict__ slot_start,
                                                       int* cursor, int* bucket, int n) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const int s = cells[i].w;
  if (s < 0) return;
  bucket[slot_start[s] + atomicAdd(&cursor[s], 1)] = i;
}

__device__ inline void sort_rows_by_b(int4* seg, int m) {
  for (int a = 1; a < m; ++a) {
    int4 v = seg[a];
    int b = a - 1;
    while (b >= 0 && seg[b].y > v.y) {
      seg[b + 1] = seg[b];
      --b;
    }
    seg[b + 1] = v;
  }
}

template <bool EMIT>
__global__ __launch_bounds__(kB) void k_sphere_query(
    const float4* __restrict__ X, const float* __restrict__ radius,
    const float* __restrict__ mu, const float* __restrict__ params,
    const int4* __restrict__ cells, const int* __restrict__ slot_start,
    const int* __restrict__ bucket, int T, int n, int* __restrict__ cnt,
    const int* __restrict__ off, int4* __restrict__ rows, int cap) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const int4 ci = cells[i];
  int found = 0;
  if (ci.w >= 0) {
    const float margin = params[0];
    const float4 xi = X[i];
    const float ri = radius[i];
    const int base = EMIT ? off[i] : 0;
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          const int cx = ci.x + dx, cy = ci.y + dy, cz = ci.z + dz;
          const int s = chash(cx, cy, cz, T);
          for (int p = slot_start[s]; p < slot_start[s + 1]; ++p) {
            const int j = bucket[p];
            if (j <= i) continue;
            const int4 cj = cells[j];
            if (cj.x != cx || cj.y != cy || cj.z != cz) continue;
            const float4 xj = X[j];
            const float d0 = xi.x - xj.x, d1 = xi.y - xj.y, d2 = xi.z - xj.z;
            const float dd = d0 * d0 + d1 * d1 + d2 * d2;
            const float rs = ri + radius[j] + margin;
            if (!(dd <= rs * rs)) continue;
            if (EMIT && base + found < cap) {
              const float rest = ri + radius[j];
              const float m = sqrtf(mu[i] * mu[j]);
              rows[base + found] =
                  make_int4(i | (T_CONTACT << 28), j, __float_as_int(rest), __float_as_int(m));
            }
            ++found;
          }
        }
    if (EMIT && found > 0) {
      const int m = min(found, max(cap - base, 0));
      if (m > 0) sort_rows_by_b(rows + base, m);
    }
  }
  if (!EMIT) cnt[i] = found;
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
  int *slot_count, *slot_start, *cursor, *bucket, *cnt, *off;
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
  w.bucket = (int*)take(sizeof(int) * (size_t)n);
  w.cnt = (int*)take(sizeof(int) * ((size_t)n + 1));
  w.off = (int*)take(sizeof(int) * ((size_t)n + 1));
  w.scan.block_sums = (int*)take(sizeof(int) * scan_block_sums(std::max<long long>(T + 1, n + 1)));
  return off;
}

}  // namespace

extern "C" {

long long xpbd_contacts_workspace_bytes(int n) {
  Ws w;
  return (long long)carve(nullptr, w, n < 1 ? 1 : n) + 256;
}

int xpbd_sphere_contacts(void* stream, void* ws, const float* X, const float* V,
                         const float* radius, const float* mu, float contact_margin, float dt,
                         float r_max, int n, int cap, int* ROWS, int* count) {
  hipStream_t s = (hipStream_t)stream;
  if (n <= 1 || !(r_max > 0.0f)) {
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(kB), 0, s, count, 0, 1);
    XPBD_LAUNCH_CHECK();
    return 0;
  }
  Ws w;
  carve((char*)ws, w, n);
  const int T = slots_for(n);
  const int nb = nblk(n, kB);
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(kB), 0, s, w.vmax, 0, 1);
  hipLaunchKernelGGL(k_fill, dim3(nblk(T + 1, kB)), dim3(kB), 0, s, w.slot_count, 0, T + 1);
  hipLaunchKernelGGL(k_fill, dim3(nblk(T, kB)), dim3(kB), 0, s, w.cursor, 0, T);
  hipLaunchKernelGGL(k_vmax, dim3(min(nb, 1024)), dim3(kB), 0, s, (const float4*)V, n, w.vmax);
  hipLaunchKernelGGL(k_contact_params, dim3(1), dim3(1), 0, s, w.vmax, contact_margin, dt, r_max,
                     w.params);
  hipLaunchKernelGGL(k_sphere_cells, dim3(nb), dim3(kB), 0, s, (const float4*)X, radius,
                     w.params, T, w.cells, w.slot_count, n);
  int rc = exclusive_scan(s, w.slot_count, w.slot_start, T + 1, w.scan, nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL(k_sphere_scatter, dim3(nb), dim3(kB), 0, s, w.cells, w.slot_start, w.cursor,
                     w.bucket, n);
  hipLaunchKernelGGL(k_sphere_query<false>, dim3(nb), dim3(kB), 0, s, (const float4*)X, radius,
                     mu, w.params, w.cells, w.slot_start, w.bucket, T, n, w.cnt, w.off,
                     (int4*)ROWS, cap);
  rc = exclusive_scan(s, w.cnt, w.off, n, w.scan, count);
  if (rc) return rc;
  hipLaunchKernelGGL(k_sphere_query<true>, dim3(nb), dim3(kB), 0, s, (const float4*)X, radius,
                     mu, w.params, w.cells, w.slot_start, w.bucket, T, n, w.cnt, w.off,
                     (int4*)ROWS, cap);
  XPBD_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
