// Shared helpers for the XPBD HIP kernels (gfx950 / CDNA4).
//
// Every arithmetic helper evaluates in the order the reference's Warp builtins
// define (dot = (a0*b0 + a1*b1) + a2*b2, length = sqrt(dot), vec/scalar divides
// each component) and the library is compiled with -ffp-contract=off, so a
// kernel that follows the reference statement order reproduces its FP32
// results bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define XPBD_EPS 1.0e-9f

// Constraint type codes (src/xpbd3d/kernels.py:37-40).
enum : int { T_DISTANCE = 0, T_ATTACH = 1, T_FLOOR = 2, T_CONTACT = 3 };

// A packed 3-DOF constraint row is one int4:
//   .x = body_a | (type << 28)    .y = body_b (-1 for one-body rows)
//   .z = bits of p0               .w = bits of p1
// DISTANCE: p0 = rest, p1 = compliance     ATTACH: p0 = rest, p1 = compliance
// FLOOR:    p0 = floor_y, p1 = mu          CONTACT: p0 = rest, p1 = mu
// ATTACH rows read their world anchor from a separate float4 array.
#define ROW_BODY_MASK 0x0FFFFFFF
__host__ __device__ inline int row_a(int4 r) { return r.x & ROW_BODY_MASK; }
__host__ __device__ inline int row_type(int4 r) { return (int)((unsigned)r.x >> 28); }

struct f3 {
  float x, y, z;
};
__host__ __device__ inline f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ inline f3 xyz(float4 a) { return f3{a.x, a.y, a.z}; }
__host__ __device__ inline f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__host__ __device__ inline f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__host__ __device__ inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline float length(f3 a) { return sqrtf(dot(a, a)); }
__host__ __device__ inline f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline float4 with_xyz(float4 a, f3 p) { return make_float4(p.x, p.y, p.z, a.w); }

// Launch geometry: consecutive blocks of a 1-D grid are dealt round-robin to
// the 8 XCDs of an MI355X.  For kernels whose neighbouring tiles share data
// (Jacobi windows, colour slices of a mesh) remap so that each XCD owns one
// contiguous run of tiles and its private L2 sees the shared lines.
// The map is a bijection on [0, nblocks): the largest multiple-of-8 prefix is
// transposed (block b -> (b % 8) * per + b / 8), the ragged tail maps to itself.
__host__ __device__ inline int xcd_tile(int bid, int nblocks) {
  const int nx = 8;
  int per = nblocks / nx;
  if (bid >= per * nx) return bid;
  return (bid % nx) * per + bid / nx;
}

#define XPBD_CHECK(expr)                       \
  do {                                         \
    hipError_t _e = (expr);                    \
    if (_e != hipSuccess) return (int)_e;      \
  } while (0)

#define XPBD_LAUNCH_CHECK() XPBD_CHECK(hipGetLastError())
