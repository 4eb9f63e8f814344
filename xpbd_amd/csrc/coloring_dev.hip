// Device colouring of the 3-DOF rows (src/xpbd3d/coloring.py:26-89) for
// Gauss-Seidel frames whose CONTACT rows are generated on the device every
// frame, so that the rows never travel to the host.
//
// The reference colours greedily: rows in the order of a stable sort on
// -(deg a + deg b) (deg = rows incident to a body), each taking the smallest
// colour no earlier row sharing one of its bodies took.  The same colours
// come out of Jones-Plassmann rounds with that order as the priority: a row
// takes the smallest colour free among its earlier neighbours as soon as all
// of them are coloured, and a row's colour depends only on those.  On the
// contact graphs this serves (sphere stacks and piles) that is 4-10 rounds.
// One workgroup does everything (degrees and the incidence lists in LDS and
// scratch, the rounds separated by barriers, then the colour ranges), so it
// runs inside a frame without a host round trip.
#include <cstdint>

#include "common.h"
#include "xpbd.h"

namespace {

constexpr int kCB = 1024;           // threads of the colouring workgroup
constexpr int kColorWords = 4;      // up to 256 colours
constexpr int kMaxBodies = 4096;    // bodies whose degrees fit the LDS

__device__ inline int row_b_of(int4 r) { return r.y; }

// LDS: deg[n], start[n + 1], cursor[n] (ints); colour counters [256].
__global__ __launch_bounds__(kCB) void k_color_rows(const int4* __restrict__ R, int m_static,
                                                    const int* __restrict__ dyn_count,
                                                    int dyn_cap, int n, int* __restrict__ inc,
                                                    int* __restrict__ color,
                                                    int* __restrict__ order,
                                                    int* __restrict__ starts,
                                                    int* __restrict__ ncolors) {
  __shared__ int deg[kMaxBodies];
  __shared__ int start[kMaxBodies + 1];
  __shared__ int cursor[kMaxBodies];
  __shared__ int ccount[kColorWords * 64 + 1];
  __shared__ int s_remaining, s_maxc, s_bad;
  const int tid = threadIdx.x;
  const int m = m_static + (dyn_count ? min(max(*dyn_count, 0), dyn_cap) : 0);
  for (int i = tid; i < n; i += kCB) deg[i] = 0;
  for (int c = tid; c <= kColorWords * 64; c += kCB) ccount[c] = 0;
  if (tid == 0) {
    s_maxc = -1;
    s_bad = 0;
  }
  __syncthreads();
  // degrees (coloring.py:45-52)
  for (int j = tid; j < m; j += kCB) {
    const int4 r = R[j];
    atomicAdd(&deg[row_a(r)], 1);
    if (row_b_of(r) >= 0) atomicAdd(&deg[row_b_of(r)], 1);
  }
  __syncthreads();
  // incidence lists: exclusive scan of the degrees (one thread; n <= 4096)
  if (tid == 0) {
    int acc = 0;
    for (int i = 0; i < n; ++i) {
      start[i] = acc;
      acc += deg[i];
    }
    start[n] = acc;
  }
  __syncthreads();
  for (int i = tid; i < n; i += kCB) cursor[i] = start[i];
  __syncthreads();
  for (int j = tid; j < m; j += kCB) {
    const int4 r = R[j];
    inc[atomicAdd(&cursor[row_a(r)], 1)] = j;
    if (row_b_of(r) >= 0) inc[atomicAdd(&cursor[row_b_of(r)], 1)] = j;
    color[j] = -1;
  }
  __syncthreads();
  auto dsum = [&](int4 r) { return deg[row_a(r)] + (row_b_of(r) >= 0 ? deg[row_b_of(r)] : 0); };
  // Jones-Plassmann rounds; k is earlier than j when its degree sum is
  // larger, or equal with a smaller index (the stable sort's order)
  for (int round = 0; round <= m; ++round) {
    if (tid == 0) s_remaining = 0;
    __syncthreads();
    for (int j = tid; j < m; j += kCB) {
      if (__atomic_load_n(&color[j], __ATOMIC_RELAXED) >= 0) continue;
      const int4 rj = R[j];
      const int dj = dsum(rj);
      unsigned long long used[kColorWords] = {0ull, 0ull, 0ull, 0ull};
      bool ready = true;
      const int bodies[2] = {row_a(rj), row_b_of(rj)};
      for (int q = 0; q < 2 && ready; ++q) {
        const int body = bodies[q];
        if (body < 0) continue;
        for (int p = start[body]; p < start[body + 1]; ++p) {
          const int k = inc[p];
          if (k == j) continue;
          const int dk = dsum(R[k]);
          if (dk > dj || (dk == dj && k < j)) {
            const int c = __atomic_load_n(&color[k], __ATOMIC_RELAXED);
            if (c < 0) {
              ready = false;
              break;
            }
            used[c >> 6] |= 1ull << (c & 63);
          }
        }
      }
      if (!ready) {
        atomicAdd(&s_remaining, 1);
        continue;
      }
      int c = -1;
      for (int w = 0; w < kColorWords; ++w) {
        if (~used[w]) {
          c = w * 64 + __builtin_ctzll(~used[w]);
          break;
        }
      }
      if (c < 0) {  // more colours than the ranges hold: no GS frame runs
        s_bad = 1;
        c = 0;
      }
      __atomic_store_n(&color[j], c, __ATOMIC_RELAXED);
      atomicMax(&s_maxc, c);
    }
    __syncthreads();
    const int remaining = s_remaining;
    __syncthreads();  // every thread has read it before thread 0 clears it
    if (remaining == 0) break;
  }
  __syncthreads();
  // colour ranges (coloring.py:74-89); inside a colour the rows share no
  // body, so their order does not change what Gauss-Seidel computes
  const int nc = s_maxc + 1;
  for (int j = tid; j < m; j += kCB)
    if (color[j] >= 0) atomicAdd(&ccount[color[j]], 1);
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int c = 0; c < nc; ++c) {
      const int k = ccount[c];
      ccount[c] = acc;
      starts[c] = acc;
      acc += k;
    }
    starts[nc] = acc;
    *ncolors = (s_bad || s_remaining != 0) ? -1 : nc;
  }
  __syncthreads();
  for (int j = tid; j < m; j += kCB)
    if (color[j] >= 0) order[atomicAdd(&ccount[color[j]], 1)] = j;
}

}  // namespace

extern "C" {

int xpbd_color_rows_max_bodies(void) { return kMaxBodies; }

int xpbd_color_rows_device(void* stream, const int* R, int n_rows_static, const int* dyn_count,
                           int dyn_cap, int n_bodies, int* inc, int* color, int* order,
                           int* color_starts, int* num_colors) {
  if (n_bodies < 0 || n_bodies > kMaxBodies || n_rows_static < 0 || dyn_cap < 0 || !R ||
      !inc || !color || !order || !color_starts || !num_colors)
    return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_color_rows, dim3(1), dim3(kCB), 0, (hipStream_t)stream, (const int4*)R,
                     n_rows_static, dyn_count, dyn_cap, n_bodies, inc, color, order, color_starts,
                     num_colors);
  XPBD_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
