// Device-wide exclusive scan of int32 arrays (three launches: per-tile scan,
// scan of the tile sums in one block, add-back), shared by the broad-phase
// and contact-generation kernels.
#pragma once
#include "common.h"

namespace {

constexpr int kB = 256;
constexpr int kScanItems = 4;               // items per thread in a scan block
constexpr int kScanTile = kB * kScanItems;  // 1024

inline int nblk(long long n, int per) { return (int)((n + per - 1) / per); }

struct ScanWs {
  int* block_sums;  // scan_block_sums(len) ints
};

inline size_t scan_block_sums(long long len) { return (size_t)nblk(len, kScanTile) + 1; }

__device__ inline int block_exclusive_scan(int v, int* sh) {
  // sh: kB ints.  Returns exclusive prefix; sh[kB-1] ends as inclusive of the last
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int d = 1; d < kB; d <<= 1) {
    int add = t >= d ? sh[t - d] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  return sh[t] - v;
}

__global__ __launch_bounds__(kB) void k_scan_tiles(const int* __restrict__ in, int* __restrict__ out,
                                                   int* __restrict__ sums, int len) {
  __shared__ int sh[kB];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  int local = 0;
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < len) ? in[base + k] : 0;
    local += v[k];
  }
  int ex = block_exclusive_scan(local, sh);
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < len) out[base + k] = ex;
    ex += v[k];
  }
  if (threadIdx.x == kB - 1) sums[blockIdx.x] = sh[kB - 1];
}

__global__ __launch_bounds__(kB) void k_scan_sums(int* sums, int nsums, int* total_out) {
  __shared__ int sh[kB];
  int carry = 0;
  for (int start = 0; start < nsums; start += kB) {
    int i = start + threadIdx.x;
    int v = i < nsums ? sums[i] : 0;
    int ex = block_exclusive_scan(v, sh);
    int tot = sh[kB - 1];
    if (i < nsums) sums[i] = ex + carry;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}

__global__ __launch_bounds__(kB) void k_scan_add(int* out, const int* sums, int len) {
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  const int add = sums[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < len) out[base + k] += add;
}

// Up to kScanOneMax items: one workgroup of kScanOneT threads, kScanItems
// items each, scans everything in one launch (in may equal out: each thread
// reads its items before the first barrier and writes only them).
constexpr int kScanOneT = 1024;
constexpr int kScanOneMax = kScanOneT * kScanItems;  // 4096

__global__ __launch_bounds__(kScanOneT) void k_scan_one(const int* in, int* out, int len,
                                                         int* total_out) {
  __shared__ int sh[kScanOneT];
  const int t = threadIdx.x;
  const int base = t * kScanItems;
  int v[kScanItems];
  int local = 0;
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < len) ? in[base + k] : 0;
    local += v[k];
  }
  sh[t] = local;
  __syncthreads();
  for (int d = 1; d < kScanOneT; d <<= 1) {
    const int add = t >= d ? sh[t - d] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  int ex = sh[t] - local;
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < len) out[base + k] = ex;
    ex += v[k];
  }
  if (t == kScanOneT - 1 && total_out) *total_out = sh[kScanOneT - 1];
}

int exclusive_scan(hipStream_t s, const int* in, int* out, int len, ScanWs& ws, int* total) {
  if (len <= kScanOneMax) {
    hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(kScanOneT), 0, s, in, out, len, total);
    XPBD_LAUNCH_CHECK();
    return 0;
  }
  int nb = nblk(len, kScanTile);
  hipLaunchKernelGGL(k_scan_tiles, dim3(nb), dim3(kB), 0, s, in, out, ws.block_sums, len);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kB), 0, s, ws.block_sums, nb, total);
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kB), 0, s, out, ws.block_sums, len);
  XPBD_LAUNCH_CHECK();
  return 0;
}

}  // namespace
