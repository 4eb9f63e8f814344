// GPU broad phase of the 6-DOF solver: per-body inflated world AABBs
// (kernels_bvh.py:21-48), a Morton-coded LBVH (Karras 2012) built with an
// LDS-ranked LSD radix sort, and a counting-sort uniform hash grid.  Both
// query paths apply the reference's pair filters (kernels_bvh.py:52-85,
// kernels_hashgrid.py:37-88) and emit pairs deterministically: a count pass,
// an exclusive scan, then an emit pass where every body writes its own
// partners j > i into its segment sorted ascending — the list is ordered by
// (i, j) no matter how the tree is traversed.
//
// All scratch lives in a caller-owned workspace (xpbd6_*_workspace_bytes).
#include <cmath>
#include <cstdint>
#include <mutex>

#include "common.h"
#include "rigid_math.h"
#include "scan.h"
#include "xpbd.h"

namespace {

constexpr int kSortItems = 8;                   // keys per thread in a sort tile
constexpr int kSortTile = kB * kSortItems;      // 2048
constexpr int kStack = 64;                      // traversal stack depth

// ---------------------------------------------------------------------------
// workspace carving
// ---------------------------------------------------------------------------
struct Arena {
  char* base;
  size_t off = 0;
  template <class T>
  T* take(size_t count) {
    off = (off + 255) & ~(size_t)255;
    T* p = (T*)(base ? base + off : nullptr);
    off += sizeof(T) * (count ? count : 1);
    return p;
  }
};

struct LbvhWs {
  float4 *lo, *hi;                  // [n] body AABBs
  float4* bounds;                   // [2] centroid bounds
  float4* part;                     // [2 * nb] partial bounds
  unsigned *key_a, *key_b;          // [n]
  int *val_a, *val_b;               // [n]
  int* hist;                        // [256 * ntiles]
  int *left, *right, *parent, *flag;  // [n-1]
  int* leaf_parent;                 // [n]
  float4 *nlo, *nhi;                // [n-1] internal node boxes
  int *cnt, *off;                   // [n+1]
  ScanWs scan;
};

struct GridWs {
  int4* cell;        // [n] integer cell coords, .w = hash slot
  int* slot_count;   // [T+1]
  int* slot_start;   // [T+1]
  int* cursor;       // [T]
  int* bucket;       // [n] body ids grouped by slot
  int *cnt, *off;    // [n+1]
  ScanWs scan;
};

size_t scan_ws(Arena& a, ScanWs& s, long long items) {
  s.block_sums = a.take<int>(scan_block_sums(items));
  return a.off;
}

int hist_len(int n) { return 256 * nblk(n, kSortTile); }

size_t carve_lbvh(Arena& a, LbvhWs& w, int n) {
  int ni = n > 1 ? n - 1 : 1;
  w.lo = a.take<float4>(n);
  w.hi = a.take<float4>(n);
  w.bounds = a.take<float4>(2);
  w.part = a.take<float4>(2 * (size_t)nblk(n, kB));
  w.key_a = a.take<unsigned>(n);
  w.key_b = a.take<unsigned>(n);
  w.val_a = a.take<int>(n);
  w.val_b = a.take<int>(n);
  w.hist = a.take<int>((size_t)hist_len(n) + 1);
  w.left = a.take<int>(ni);
  w.right = a.take<int>(ni);
  w.parent = a.take<int>(ni);
  w.flag = a.take<int>(ni);
  w.leaf_parent = a.take<int>(n);
  w.nlo = a.take<float4>(ni);
  w.nhi = a.take<float4>(ni);
  w.cnt = a.take<int>((size_t)n + 1);
  w.off = a.take<int>((size_t)n + 1);
  long long scan_items = std::max<long long>((long long)n + 1, hist_len(n) + 1);
  scan_ws(a, w.scan, scan_items);
  return a.off;
}

int grid_slots(int n) {
  int t = 1024;
  while (t < 2 * n) t <<= 1;
  return t;
}

size_t carve_grid(Arena& a, GridWs& w, int n) {
  int T = grid_slots(n);
  w.cell = a.take<int4>(n);
  w.slot_count = a.take<int>((size_t)T + 1);
  w.slot_start = a.take<int>((size_t)T + 1);
  w.cursor = a.take<int>(T);
  w.bucket = a.take<int>(n);
  w.cnt = a.take<int>((size_t)n + 1);
  w.off = a.take<int>((size_t)n + 1);
  scan_ws(a, w.scan, std::max<long long>((long long)n + 1, (long long)T + 1));
  return a.off;
}

// ---------------------------------------------------------------------------
// AABBs (kernels_bvh.py:21-48) and centroid bounds
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kB) void k_body_aabb(const float4* __restrict__ X,
                                                  const float4* __restrict__ Q,
                                                  const float4* __restrict__ V,
                                                  const float4* __restrict__ W,
                                                  const float4* __restrict__ HE, float man_margin,
                                                  float dt, float4* LO, float4* HI, int n) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  body_aabb(X[i], Q[i], V[i], W[i], HE[i], man_margin, dt, LO[i], HI[i]);
}

__global__ __launch_bounds__(kB) void k_bounds_partial(const float4* __restrict__ LO,
                                                       const float4* __restrict__ HI,
                                                       float4* part, int n) {
  __shared__ float smin[3][kB], smax[3][kB];
  const int i = blockIdx.x * kB + threadIdx.x;
  f3 c = mk3(INFINITY, INFINITY, INFINITY), d = mk3(-INFINITY, -INFINITY, -INFINITY);
  if (i < n) {
    f3 cc = (xyz(LO[i]) + xyz(HI[i])) * 0.5f;
    c = cc;
    d = cc;
  }
  smin[0][threadIdx.x] = c.x; smin[1][threadIdx.x] = c.y; smin[2][threadIdx.x] = c.z;
  smax[0][threadIdx.x] = d.x; smax[1][threadIdx.x] = d.y; smax[2][threadIdx.x] = d.z;
  __syncthreads();
  for (int st = kB / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      for (int k = 0; k < 3; ++k) {
        smin[k][threadIdx.x] = fminf(smin[k][threadIdx.x], smin[k][threadIdx.x + st]);
        smax[k][threadIdx.x] = fmaxf(smax[k][threadIdx.x], smax[k][threadIdx.x + st]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = make_float4(smin[0][0], smin[1][0], smin[2][0], 0.f);
    part[2 * blockIdx.x + 1] = make_float4(smax[0][0], smax[1][0], smax[2][0], 0.f);
  }
}

__global__ __launch_bounds__(kB) void k_bounds_final(const float4* __restrict__ part, int np,
                                                     float4* bounds) {
  __shared__ float smin[3][kB], smax[3][kB];
  f3 c = mk3(INFINITY, INFINITY, INFINITY), d = mk3(-INFINITY, -INFINITY, -INFINITY);
  for (int k = threadIdx.x; k < np; k += kB) {
    float4 a = part[2 * k], b = part[2 * k + 1];
    c = mk3(fminf(c.x, a.x), fminf(c.y, a.y), fminf(c.z, a.z));
    d = mk3(fmaxf(d.x, b.x), fmaxf(d.y, b.y), fmaxf(d.z, b.z));
  }
  smin[0][threadIdx.x] = c.x; smin[1][threadIdx.x] = c.y; smin[2][threadIdx.x] = c.z;
  smax[0][threadIdx.x] = d.x; smax[1][threadIdx.x] = d.y; smax[2][threadIdx.x] = d.z;
  __syncthreads();
  for (int st = kB / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      for (int k = 0; k < 3; ++k) {
        smin[k][threadIdx.x] = fminf(smin[k][threadIdx.x], smin[k][threadIdx.x + st]);
        smax[k][threadIdx.x] = fmaxf(smax[k][threadIdx.x], smax[k][threadIdx.x + st]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    bounds[0] = make_float4(smin[0][0], smin[1][0], smin[2][0], 0.f);
    bounds[1] = make_float4(smax[0][0], smax[1][0], smax[2][0], 0.f);
  }
}

// 30-bit Morton code: 10 bits per axis of the centroid, normalised to the
// centroid bounds, bits interleaved x2 y1 z0 from the top (Karras 2012).
__host__ __device__ inline unsigned expand_bits(unsigned v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}
__host__ __device__ inline unsigned morton3(float x, float y, float z) {
  x = fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
  y = fminf(fmaxf(y * 1024.0f, 0.0f), 1023.0f);
  z = fminf(fmaxf(z * 1024.0f, 0.0f), 1023.0f);
  return (expand_bits((unsigned)x) << 2) | (expand_bits((unsigned)y) << 1) | expand_bits((unsigned)z);
}

__global__ __launch_bounds__(kB) void k_morton(const float4* __restrict__ LO,
                                               const float4* __restrict__ HI,
                                               const float4* __restrict__ bounds, unsigned* key,
                                               int* val, int n) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const f3 lo = xyz(bounds[0]), hi = xyz(bounds[1]);
  const f3 ext = hi - lo;
  const f3 c = (xyz(LO[i]) + xyz(HI[i])) * 0.5f;
  const float nx = ext.x > 0.0f ? (c.x - lo.x) / ext.x : 0.0f;
  const float ny = ext.y > 0.0f ? (c.y - lo.y) / ext.y : 0.0f;
  const float nz = ext.z > 0.0f ? (c.z - lo.z) / ext.z : 0.0f;
  key[i] = morton3(nx, ny, nz);
  val[i] = i;
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits.  Scatter ranks each 256-key chunk of
// a tile with a wave multisplit (8 ballots give every lane the mask of lanes
// holding its digit) and per-wave digit counts in LDS.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kB) void k_radix_hist(const unsigned* __restrict__ key, int n,
                                                   int shift, int* hist, int ntiles) {
  __shared__ int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int base = blockIdx.x * kSortTile;
  for (int k = 0; k < kSortItems; ++k) {
    int idx = base + k * kB + threadIdx.x;
    if (idx < n) atomicAdd(&h[(key[idx] >> shift) & 255u], 1);
  }
  __syncthreads();
  hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kB) void k_radix_scatter(const unsigned* __restrict__ key_in,
                                                      const int* __restrict__ val_in,
                                                      unsigned* __restrict__ key_out,
                                                      int* __restrict__ val_out, int n, int shift,
                                                      const int* __restrict__ offs, int ntiles) {
  __shared__ int run[256];
  __shared__ int start[256];
  __shared__ int wcnt[kB / 64][256];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const unsigned long long lt = (1ull << lane) - 1ull;
  start[t] = offs[t * ntiles + blockIdx.x];
  run[t] = 0;
  const int base = blockIdx.x * kSortTile;
  for (int k = 0; k < kSortItems; ++k) {
    for (int w = 0; w < kB / 64; ++w) wcnt[w][t] = 0;
    __syncthreads();
    const int idx = base + k * kB + t;
    const bool active = idx < n;
    const unsigned kk = active ? key_in[idx] : 0u;
    const int vv = active ? val_in[idx] : 0;
    const unsigned digit = (kk >> shift) & 255u;
    unsigned long long peers = __ballot(active);
    for (int b = 0; b < 8; ++b) {
      const bool bit = (digit >> b) & 1u;
      const unsigned long long bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    if (!active) peers = 0;
    const int rank_w = __popcll(peers & lt);
    if (active && rank_w == 0) wcnt[wave][digit] = __popcll(peers);
    __syncthreads();
    if (active) {
      int r = rank_w;
      for (int w = 0; w < wave; ++w) r += wcnt[w][digit];
      const int pos = start[digit] + run[digit] + r;
      key_out[pos] = kk;
      val_out[pos] = vv;
    }
    __syncthreads();
    int add = 0;
    for (int w = 0; w < kB / 64; ++w) add += wcnt[w][t];
    run[t] += add;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Karras 2012 hierarchy.  Children: >= 0 internal node, < 0 leaf ~k.
// ---------------------------------------------------------------------------
// Leaves hold L consecutive bodies of the Morton order (L = the leaf size;
// leaf i's key is that of its first body); the hierarchy is Karras's over
// the n = ceil(bodies / L) leaves.
__device__ inline int delta(const unsigned* key, int n, int L, int i, int j) {
  if (j < 0 || j >= n) return -1;
  unsigned a = key[i * L], b = key[j * L];
  if (a == b) return 32 + __clz((unsigned)(i ^ j));
  return __clz(a ^ b);
}

__global__ __launch_bounds__(kB) void k_karras(const unsigned* __restrict__ key, int n, int L,
                                               int* left, int* right, int* parent,
                                               int* leaf_parent, int* flag) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n - 1) return;
  const int d = (delta(key, n, L, i, i + 1) - delta(key, n, L, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(key, n, L, i, i - d);
  int lmax = 2;
  while (delta(key, n, L, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(key, n, L, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = delta(key, n, L, i, j);
  int s = 0;
  int t = l;
  do {
    t = (t + 1) >> 1;
    if (delta(key, n, L, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int gamma = i + s * d + min(d, 0);
  const int lo = min(i, j), hi = max(i, j);
  const int lc = (lo == gamma) ? ~gamma : gamma;
  const int rc = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
  left[i] = lc;
  right[i] = rc;
  if (lc < 0) leaf_parent[~lc] = i; else parent[lc] = i;
  if (rc < 0) leaf_parent[~rc] = i; else parent[rc] = i;
  if (i == 0) parent[0] = -1;
  flag[i] = 0;
}

__device__ inline float4 ld_coherent(const float4* p) {
  const float* f = (const float*)p;
  return make_float4(__hip_atomic_load(f + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                     __hip_atomic_load(f + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                     __hip_atomic_load(f + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), 0.0f);
}

// Box of leaf k: the union of its bodies' boxes (bodies val[k L .. k L + L)
// clipped to nb bodies).
__device__ inline void leaf_box(const int* __restrict__ val, const float4* __restrict__ LO,
                                const float4* __restrict__ HI, int k, int L, int nb, float4& lo,
                                float4& hi) {
  const int b0 = k * L, b1 = min(b0 + L, nb);
  lo = LO[val[b0]];
  hi = HI[val[b0]];
  for (int t = b0 + 1; t < b1; ++t) {
    const float4 l = LO[val[t]], h = HI[val[t]];
    lo = make_float4(fminf(lo.x, l.x), fminf(lo.y, l.y), fminf(lo.z, l.z), 0.f);
    hi = make_float4(fmaxf(hi.x, h.x), fmaxf(hi.y, h.y), fmaxf(hi.z, h.z), 0.f);
  }
}

// Bottom-up refit over n leaves of L bodies (nb bodies): the second thread
// to reach a node unions its children.
__global__ __launch_bounds__(kB) void k_refit(const int* __restrict__ val,
                                              const float4* __restrict__ LO,
                                              const float4* __restrict__ HI,
                                              const int* __restrict__ left,
                                              const int* __restrict__ right,
                                              const int* __restrict__ parent,
                                              const int* __restrict__ leaf_parent, int* flag,
                                              float4* nlo, float4* nhi, int n, int L, int nb) {
  const int k = blockIdx.x * kB + threadIdx.x;
  if (k >= n) return;
  int node = leaf_parent[k];
  while (node >= 0) {
    __threadfence();
    if (atomicAdd(&flag[node], 1) == 0) return;
    __threadfence();
    const int c0 = left[node], c1 = right[node];
    float4 l0, h0, l1, h1;
    if (c0 < 0) leaf_box(val, LO, HI, ~c0, L, nb, l0, h0);
    else { l0 = ld_coherent(&nlo[c0]); h0 = ld_coherent(&nhi[c0]); }
    if (c1 < 0) leaf_box(val, LO, HI, ~c1, L, nb, l1, h1);
    else { l1 = ld_coherent(&nlo[c1]); h1 = ld_coherent(&nhi[c1]); }
    nlo[node] = make_float4(fminf(l0.x, l1.x), fminf(l0.y, l1.y), fminf(l0.z, l1.z), 0.f);
    nhi[node] = make_float4(fmaxf(h0.x, h1.x), fmaxf(h0.y, h1.y), fmaxf(h0.z, h1.z), 0.f);
    node = parent[node];
  }
}

// Insertion sort of one body's partner segment (ascending j).
__device__ inline void sort_segment(int2* seg, int m) {
  for (int a = 1; a < m; ++a) {
    int2 v = seg[a];
    int b = a - 1;
    while (b >= 0 && seg[b].y > v.y) {
      seg[b + 1] = seg[b];
      --b;
    }
    seg[b + 1] = v;
  }
}

// All-pairs test for small scenes: one wavefront per body i sweeps the
// bodies j > i in 64-wide chunks against the same AABBs and filter as the
// LBVH query.  COUNT writes the partner count of i; EMIT writes the pairs
// at off[i] in ascending j (each lane's slot = the hits of lower lanes in
// its chunk, from the ballot), i.e. exactly the sorted segments of the LBVH
// path, in 4 launches (AABBs, count, scan, emit) instead of ~25.
constexpr int kAllPairsMax = 4096;
constexpr int kMaxLeaf = 64;  // bodies per LBVH leaf

template <bool EMIT>
__global__ __launch_bounds__(kB) void k_allpairs(const float4* __restrict__ X,
                                                 const float4* __restrict__ HE,
                                                 const float4* __restrict__ LO,
                                                 const float4* __restrict__ HI, int n,
                                                 int* __restrict__ cnt, const int* __restrict__ off,
                                                 int2* __restrict__ pairs, int cap) {
  const int i = blockIdx.x * (kB / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;  // whole wave
  const float4 lo = LO[i], hi = HI[i];
  const float imi = X[i].w;
  const int gi = __float_as_int(HE[i].w);
  int k = EMIT ? off[i] : 0;
  for (int j0 = i + 1; j0 < n; j0 += 64) {
    const int j = j0 + lane;
    bool hit = false;
    if (j < n)
      hit = overlap(lo, hi, LO[j], HI[j]) &&
            keep_pair(i, j, imi, X[j].w, gi, __float_as_int(HE[j].w));
    const unsigned long long m = __ballot(hit);
    if (EMIT && hit) {
      const int slot = k + __popcll(m & ((1ull << lane) - 1ull));
      if (slot < cap) pairs[slot] = make_int2(i, j);
    }
    k += __popcll(m);
  }
  if (!EMIT && lane == 0) cnt[i] = k;
}

template <bool EMIT>
__global__ __launch_bounds__(kB) void k_bvh_query(
    const float4* __restrict__ X, const float4* __restrict__ HE, const float4* __restrict__ LO,
    const float4* __restrict__ HI, const int* __restrict__ val, const int* __restrict__ left,
    const int* __restrict__ right, const float4* __restrict__ nlo,
    const float4* __restrict__ nhi, int n, int L, int nl, int* __restrict__ cnt,
    const int* __restrict__ off, int2* __restrict__ pairs, int cap) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const float4 lo = LO[i], hi = HI[i];
  const float imi = X[i].w;
  const int gi = __float_as_int(HE[i].w);
  int found = 0;
  const int base = EMIT ? off[i] : 0;
  int stack[kStack];
  int sp = 0;
  stack[sp++] = (nl == 1) ? ~0 : 0;
  while (sp > 0) {
    const int node = stack[--sp];
    if (node < 0) {  // the leaf's bodies
      const int b0 = (~node) * L, b1 = min(b0 + L, n);
      for (int t = b0; t < b1; ++t) {
        const int j = val[t];
        if (!overlap(lo, hi, LO[j], HI[j])) continue;
        if (!keep_pair(i, j, imi, X[j].w, gi, __float_as_int(HE[j].w))) continue;
        if (EMIT && base + found < cap) pairs[base + found] = make_int2(i, j);
        ++found;
      }
      continue;
    }
    if (!overlap(lo, hi, nlo[node], nhi[node])) continue;
    if (sp + 2 <= kStack) {
      stack[sp++] = right[node];
      stack[sp++] = left[node];
    }
  }
  if (!EMIT) {
    cnt[i] = found;
  } else if (found > 0) {
    const int m = min(found, max(cap - base, 0));
    if (m > 0) sort_segment(pairs + base, m);
  }
}

// ---------------------------------------------------------------------------
// Uniform hash grid (kernels_hashgrid.py:37-88)
// ---------------------------------------------------------------------------
__host__ __device__ inline int grid_hash(int cx, int cy, int cz, int T) {
  unsigned h = (unsigned)cx * 73856093u ^ (unsigned)cy * 19349663u ^ (unsigned)cz * 83492791u;
  return (int)(h & (unsigned)(T - 1));
}
__device__ inline int cell_of(float p, float cell) { return (int)floorf(p / cell); }

__global__ __launch_bounds__(kB) void k_grid_cells(const float4* __restrict__ X, float cell,
                                                   int T, int4* cells, int* slot_count, int n) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const f3 x = xyz(X[i]);
  const int cx = cell_of(x.x, cell), cy = cell_of(x.y, cell), cz = cell_of(x.z, cell);
  const int s = grid_hash(cx, cy, cz, T);
  cells[i] = make_int4(cx, cy, cz, s);
  atomicAdd(&slot_count[s], 1);
}

__global__ __launch_bounds__(kB) void k_fill_int(int* p, int v, int len) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i < len) p[i] = v;
}

__global__ __launch_bounds__(kB) void k_grid_scatter(const int4* __restrict__ cells,
                                                     const int* __restrict__ slot_start,
                                                     int* cursor, int* bucket, int n) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const int s = cells[i].w;
  bucket[slot_start[s] + atomicAdd(&cursor[s], 1)] = i;
}

template <bool EMIT>
__global__ __launch_bounds__(kB) void k_grid_query(
    const float4* __restrict__ X, const float4* __restrict__ Q, const float4* __restrict__ V,
    const float4* __restrict__ W, const float4* __restrict__ HE, const int4* __restrict__ cells,
    const int* __restrict__ slot_start, const int* __restrict__ bucket, float man_margin,
    float dt, float rs_max, float cell, int T, int n, int* __restrict__ cnt,
    const int* __restrict__ off, int2* __restrict__ pairs, int cap) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const float4 hei = HE[i];
  const f3 ei = xyz(hei);
  const float rsi = length(ei);
  const float mi = fminf(dt * (length(xyz(V[i])) + length(xyz(W[i])) * rsi), rsi);
  const float qr = fminf(rsi + rs_max + man_margin + mi, cell);
  const f3 hi = aabb_half(ldq(Q[i]), ei);
  const float4 xi4 = X[i];
  const f3 xi = xyz(xi4);
  const int gi = __float_as_int(hei.w);
  const int lx = cell_of(xi.x - qr, cell), ux = cell_of(xi.x + qr, cell);
  const int ly = cell_of(xi.y - qr, cell), uy = cell_of(xi.y + qr, cell);
  const int lz = cell_of(xi.z - qr, cell), uz = cell_of(xi.z + qr, cell);
  int found = 0;
  const int base = EMIT ? off[i] : 0;
  for (int cx = lx; cx <= ux; ++cx)
    for (int cy = ly; cy <= uy; ++cy)
      for (int cz = lz; cz <= uz; ++cz) {
        const int s = grid_hash(cx, cy, cz, T);
        for (int p = slot_start[s]; p < slot_start[s + 1]; ++p) {
          const int j = bucket[p];
          const int4 cj = cells[j];
          if (cj.x != cx || cj.y != cy || cj.z != cz) continue;
          const float4 hej = HE[j];
          if (!keep_pair(i, j, xi4.w, X[j].w, gi, __float_as_int(hej.w))) continue;
          const f3 ej = xyz(hej);
          const float rsj = length(ej);
          const float mj = fminf(dt * (length(xyz(V[j])) + length(xyz(W[j])) * rsj), rsj);
          const f3 hj = aabb_half(ldq(Q[j]), ej);
          const f3 dpos = xi - xyz(X[j]);
          const float pad = man_margin + mi + mj;
          const float ox = fabsf(dpos.x) - (hi.x + hj.x + pad);
          const float oy = fabsf(dpos.y) - (hi.y + hj.y + pad);
          const float oz = fabsf(dpos.z) - (hi.z + hj.z + pad);
          if (ox <= 0.0f && oy <= 0.0f && oz <= 0.0f) {
            if (EMIT && base + found < cap) pairs[base + found] = make_int2(i, j);
            ++found;
          }
        }
      }
  if (!EMIT) {
    cnt[i] = found;
  } else if (found > 0) {
    const int m = min(found, max(cap - base, 0));
    if (m > 0) sort_segment(pairs + base, m);
  }
}

// ---------------------------------------------------------------------------
// Fixed scratch of the allocation-free LBVH broad phase
// (xpbd6_broadphase_lbvh_noalloc): one copy per GPU, held until the process
// ends, shared by every call on that GPU.  7 MiB holds the LBVH workspace of
// XPBD_NOALLOC_MAX_BODIES bodies (6.8 MiB at 65536).  Host code never passes
// the symbol itself to a kernel: it uses the device address that
// hipGetSymbolAddress returns for the current device, kept per device.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = (size_t)7 << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per
// device, with that device current).
int scratch_address(void** out) {
  int dev = 0;
  XPBD_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    XPBD_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return (int)hipErrorInvalidSymbol;
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

long long xpbd6_lbvh_workspace_bytes(int n) {
  if (n < 1) n = 1;
  Arena a{nullptr};
  LbvhWs w;
  return (long long)carve_lbvh(a, w, n) + 256;
}

long long xpbd6_hashgrid_workspace_bytes(int n) {
  if (n < 1) n = 1;
  Arena a{nullptr};
  GridWs w;
  return (long long)carve_grid(a, w, n) + 256;
}

int xpbd6_body_aabb(void* stream, const float* X, const float* Q, const float* V, const float* W,
                    const float* HE, float man_margin, float dt, float* LO, float* HI, int n) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_body_aabb, dim3(nblk(n, kB)), dim3(kB), 0, (hipStream_t)stream,
                     (const float4*)X, (const float4*)Q, (const float4*)V, (const float4*)W,
                     (const float4*)HE, man_margin, dt, (float4*)LO, (float4*)HI, n);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd6_broadphase_lbvh(void* stream, void* ws, const float* X, const float* Q, const float* V,
                          const float* W, const float* HE, float man_margin, float dt, int n,
                          int cap, int* PAIRS, int* PC) {
  return xpbd6_broadphase_lbvh_leaf(stream, ws, X, Q, V, W, HE, man_margin, dt, n, cap, PAIRS,
                                    PC, n <= kAllPairsMax ? 0 : 1);
}

int xpbd6_broadphase_lbvh_leaf(void* stream, void* ws, const float* X, const float* Q,
                               const float* V, const float* W, const float* HE, float man_margin,
                               float dt, int n, int cap, int* PAIRS, int* PC, int leaf_size) {
  hipStream_t s = (hipStream_t)stream;
  if (leaf_size < 0 || leaf_size > kMaxLeaf) return (int)hipErrorInvalidValue;
  if (n <= 1) {
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(kB), 0, s, PC, 0, 1);
    XPBD_LAUNCH_CHECK();
    return 0;
  }
  Arena a{(char*)ws};
  LbvhWs w;
  carve_lbvh(a, w, n);
  const int nb = nblk(n, kB);
  hipLaunchKernelGGL(k_body_aabb, dim3(nb), dim3(kB), 0, s, (const float4*)X, (const float4*)Q,
                     (const float4*)V, (const float4*)W, (const float4*)HE, man_margin, dt, w.lo,
                     w.hi, n);
  if (leaf_size == 0) {  // all pairs: the same pair list, fewer launches
    const int nw = nblk(n, kB / 64);
    hipLaunchKernelGGL(k_allpairs<false>, dim3(nw), dim3(kB), 0, s, (const float4*)X,
                       (const float4*)HE, w.lo, w.hi, n, w.cnt, w.off, (int2*)PAIRS, cap);
    XPBD_LAUNCH_CHECK();
    int rc = exclusive_scan(s, w.cnt, w.off, n, w.scan, PC);
    if (rc) return rc;
    hipLaunchKernelGGL(k_allpairs<true>, dim3(nw), dim3(kB), 0, s, (const float4*)X,
                       (const float4*)HE, w.lo, w.hi, n, w.cnt, w.off, (int2*)PAIRS, cap);
    XPBD_LAUNCH_CHECK();
    return 0;
  }
  hipLaunchKernelGGL(k_bounds_partial, dim3(nb), dim3(kB), 0, s, w.lo, w.hi, w.part, n);
  hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(kB), 0, s, w.part, nb, w.bounds);
  hipLaunchKernelGGL(k_morton, dim3(nb), dim3(kB), 0, s, w.lo, w.hi, w.bounds, w.key_a, w.val_a,
                     n);
  XPBD_LAUNCH_CHECK();
  const int ntiles = nblk(n, kSortTile);
  unsigned *kin = w.key_a, *kout = w.key_b;
  int *vin = w.val_a, *vout = w.val_b;
  for (int shift = 0; shift < 32; shift += 8) {
    hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(kB), 0, s, kin, n, shift, w.hist, ntiles);
    int rc = exclusive_scan(s, w.hist, w.hist, 256 * ntiles, w.scan, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(k_radix_scatter, dim3(ntiles), dim3(kB), 0, s, kin, vin, kout, vout, n,
                       shift, w.hist, ntiles);
    XPBD_LAUNCH_CHECK();
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  // after 4 passes the sorted keys are back in key_a / val_a
  const int L = leaf_size;
  const int nl = (n + L - 1) / L;  // leaves
  if (nl > 1) {
    hipLaunchKernelGGL(k_karras, dim3(nblk(nl - 1, kB)), dim3(kB), 0, s, kin, nl, L, w.left,
                       w.right, w.parent, w.leaf_parent, w.flag);
    hipLaunchKernelGGL(k_refit, dim3(nblk(nl, kB)), dim3(kB), 0, s, vin, w.lo, w.hi, w.left,
                       w.right, w.parent, w.leaf_parent, w.flag, w.nlo, w.nhi, nl, L, n);
  }
  hipLaunchKernelGGL(k_bvh_query<false>, dim3(nb), dim3(kB), 0, s, (const float4*)X,
                     (const float4*)HE, w.lo, w.hi, vin, w.left, w.right, w.nlo, w.nhi, n, L, nl,
                     w.cnt, w.off, (int2*)PAIRS, cap);
  XPBD_LAUNCH_CHECK();
  int rc = exclusive_scan(s, w.cnt, w.off, n, w.scan, PC);
  if (rc) return rc;
  hipLaunchKernelGGL(k_bvh_query<true>, dim3(nb), dim3(kB), 0, s, (const float4*)X,
                     (const float4*)HE, w.lo, w.hi, vin, w.left, w.right, w.nlo, w.nhi, n, L, nl,
                     w.cnt, w.off, (int2*)PAIRS, cap);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd6_broadphase_lbvh_noalloc(void* stream, const float* X, const float* Q, const float* V,
                                  const float* W, const float* HE, float man_margin, float dt,
                                  int n, int cap, int* PAIRS, int* PC) {
  if (n > XPBD_NOALLOC_MAX_BODIES) return XPBD_ERR_NOALLOC_TOO_LARGE;
  if ((size_t)xpbd6_lbvh_workspace_bytes(n) > kNoAllocScratchBytes)
    return XPBD_ERR_NOALLOC_TOO_LARGE;
  void* ws = nullptr;
  int rc = scratch_address(&ws);
  if (rc) return rc;
  return xpbd6_broadphase_lbvh(stream, ws, X, Q, V, W, HE, man_margin, dt, n, cap, PAIRS, PC);
}

int xpbd6_broadphase_hashgrid(void* stream, void* ws, const float* X, const float* Q,
                              const float* V, const float* W, const float* HE, float man_margin,
                              float dt, float rs_max, float cell, int n, int cap, int* PAIRS,
                              int* PC) {
  hipStream_t s = (hipStream_t)stream;
  if (n <= 1 || !(cell > 0.0f)) {
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(kB), 0, s, PC, 0, 1);
    XPBD_LAUNCH_CHECK();
    return 0;
  }
  Arena a{(char*)ws};
  GridWs w;
  carve_grid(a, w, n);
  const int T = grid_slots(n);
  const int nb = nblk(n, kB);
  hipLaunchKernelGGL(k_fill_int, dim3(nblk(T + 1, kB)), dim3(kB), 0, s, w.slot_count, 0, T + 1);
  hipLaunchKernelGGL(k_fill_int, dim3(nblk(T, kB)), dim3(kB), 0, s, w.cursor, 0, T);
  hipLaunchKernelGGL(k_grid_cells, dim3(nb), dim3(kB), 0, s, (const float4*)X, cell, T, w.cell,
                     w.slot_count, n);
  XPBD_LAUNCH_CHECK();
  int rc = exclusive_scan(s, w.slot_count, w.slot_start, T + 1, w.scan, nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL(k_grid_scatter, dim3(nb), dim3(kB), 0, s, w.cell, w.slot_start, w.cursor,
                     w.bucket, n);
  hipLaunchKernelGGL(k_grid_query<false>, dim3(nb), dim3(kB), 0, s, (const float4*)X,
                     (const float4*)Q, (const float4*)V, (const float4*)W, (const float4*)HE,
                     w.cell, w.slot_start, w.bucket, man_margin, dt, rs_max, cell, T, n, w.cnt,
                     w.off, (int2*)PAIRS, cap);
  XPBD_LAUNCH_CHECK();
  rc = exclusive_scan(s, w.cnt, w.off, n, w.scan, PC);
  if (rc) return rc;
  hipLaunchKernelGGL(k_grid_query<true>, dim3(nb), dim3(kB), 0, s, (const float4*)X,
                     (const float4*)Q, (const float4*)V, (const float4*)W, (const float4*)HE,
                     w.cell, w.slot_start, w.bucket, man_margin, dt, rs_max, cell, T, n, w.cnt,
                     w.off, (int2*)PAIRS, cap);
  XPBD_LAUNCH_CHECK();
  return 0;
}

// Morton code of a normalised point (exported for the encode/decode tests).
unsigned xpbd_morton3(float x, float y, float z) { return morton3(x, y, z); }

}  // extern "C"
