// Quaternion / box helpers of the 6-DOF kernels (quaternions are xyzw).
//
// Each helper evaluates in a fixed, documented order (the mathematical
// definitions of the reference's Warp builtins):
//   rotate(q,v)     = (v*(2w*w-1) + qv*(2 dot(qv,v))) + cross(qv,v)*(2w)
//   rotate_inv(q,v) = (v*(2w*w-1) + qv*(2 dot(qv,v))) - cross(qv,v)*(2w)
//   to_matrix(q)    = columns rotate(q, e_k)
//   a*b             = Hamilton product, normalize(q) = q / |q|
#pragma once
#include "common.h"

struct q4 {
  float x, y, z, w;
};

__host__ __device__ inline q4 ldq(float4 a) { return q4{a.x, a.y, a.z, a.w}; }
__host__ __device__ inline float4 stq(q4 q) { return make_float4(q.x, q.y, q.z, q.w); }

__host__ __device__ inline f3 qrot(q4 q, f3 v) {
  f3 qv = mk3(q.x, q.y, q.z);
  float c = 2.0f * q.w * q.w - 1.0f;
  float d = 2.0f * dot(qv, v);
  return (v * c + qv * d) + cross(qv, v) * (q.w * 2.0f);
}
__host__ __device__ inline f3 qrot_inv(q4 q, f3 v) {
  f3 qv = mk3(q.x, q.y, q.z);
  float c = 2.0f * q.w * q.w - 1.0f;
  float d = 2.0f * dot(qv, v);
  return (v * c + qv * d) - cross(qv, v) * (q.w * 2.0f);
}
__host__ __device__ inline q4 qmul(q4 a, q4 b) {
  return q4{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z,
            a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x,
            a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
__host__ __device__ inline q4 qadd(q4 a, q4 b) { return q4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
__host__ __device__ inline q4 qscale(q4 a, float s) { return q4{a.x * s, a.y * s, a.z * s, a.w * s}; }
__host__ __device__ inline q4 qnormalize(q4 q) {
  float l = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  return q4{q.x / l, q.y / l, q.z / l, q.w / l};
}
__host__ __device__ inline q4 qconj(q4 q) { return q4{-q.x, -q.y, -q.z, q.w}; }

struct m3 {
  f3 c[3];  // columns
};
__host__ __device__ inline m3 qmat(q4 q) {
  m3 R;
  R.c[0] = qrot(q, mk3(1.0f, 0.0f, 0.0f));
  R.c[1] = qrot(q, mk3(0.0f, 1.0f, 0.0f));
  R.c[2] = qrot(q, mk3(0.0f, 0.0f, 1.0f));
  return R;
}
// transpose(R) * v
__host__ __device__ inline f3 mtv(const m3& R, f3 v) {
  return mk3(dot(R.c[0], v), dot(R.c[1], v), dot(R.c[2], v));
}
__host__ __device__ inline float comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
// Column k of R for a k the compiler can resolve (unrolled loops); a
// runtime axis uses qrot(q, unit_axis(k)) instead, which computes the same
// column from the quaternion without indexing a private array
__host__ __device__ inline f3 col(const m3& R, int k) { return R.c[k]; }
__host__ __device__ inline f3 unit_axis(int k) {
  return mk3(k == 0 ? 1.0f : 0.0f, k == 1 ? 1.0f : 0.0f, k == 2 ? 1.0f : 0.0f);
}
__host__ __device__ inline f3 sel3(bool c, f3 a, f3 b) {
  return mk3(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z);
}
__host__ __device__ inline q4 selq(bool c, q4 a, q4 b) {
  return q4{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z, c ? a.w : b.w};
}
__host__ __device__ inline f3 mulc(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }

// kernels_6dof.py:29-35
__host__ __device__ inline f3 box_corner(f3 he, int i) {
  float sx = (i & 1) ? 1.0f : -1.0f, sy = (i & 2) ? 1.0f : -1.0f, sz = (i & 4) ? 1.0f : -1.0f;
  return mk3(sx * he.x, sy * he.y, sz * he.z);
}
// kernels_6dof.py:38-44: R (I^-1 (R^T u))
__host__ __device__ inline f3 world_invI_mul(q4 q, f3 iI, f3 u) {
  f3 ub = qrot_inv(q, u);
  ub = mk3(ub.x * iI.x, ub.y * iI.y, ub.z * iI.z);
  return qrot(q, ub);
}
// kernels_6dof.py:47-53: 1/m + (r x n)^T I^-1 (r x n)
__host__ __device__ inline float gen_inv_mass(float im, q4 q, f3 iI, f3 r, f3 n) {
  f3 rn = cross(r, n);
  return im + dot(rn, world_invI_mul(q, iI, rn));
}
// kernels_6dof.py:56-61
__host__ __device__ inline q4 quat_integrate(q4 q, f3 w, float h) {
  q4 wq{w.x, w.y, w.z, 0.0f};
  q = qadd(q, qscale(qmul(wq, q), 0.5f * h));
  return qnormalize(q);
}
// kernels_bvh.py:37-43 / kernels_hashgrid.py:27-34: |R| he
__host__ __device__ inline f3 aabb_half(q4 q, f3 e) {
  m3 R = qmat(q);
  return mk3(fabsf(R.c[0].x) * e.x + fabsf(R.c[1].x) * e.y + fabsf(R.c[2].x) * e.z,
             fabsf(R.c[0].y) * e.x + fabsf(R.c[1].y) * e.y + fabsf(R.c[2].y) * e.z,
             fabsf(R.c[0].z) * e.x + fabsf(R.c[1].z) * e.y + fabsf(R.c[2].z) * e.z);
}

// kernels_bvh.py:21-48 `compute_body_aabb`: |R| he plus a velocity margin
__device__ inline void body_aabb(float4 x4, float4 q4v, float4 v4, float4 w4, float4 he4,
                                 float man_margin, float dt, float4& lo4, float4& hi4) {
  const f3 e = xyz(he4);
  const f3 half = aabb_half(ldq(q4v), e);
  const float rr = length(e);
  const float m = man_margin + dt * (length(xyz(v4)) + length(xyz(w4)) * rr);
  const f3 pad = mk3(m, m, m);
  const f3 x = xyz(x4);
  const f3 lo = (x - half) - pad, hi = (x + half) + pad;
  lo4 = make_float4(lo.x, lo.y, lo.z, 0.0f);
  hi4 = make_float4(hi.x, hi.y, hi.z, 0.0f);
}

// kernels_bvh.py:52-85 filter: j > i, not static-static, not same group > 0
__device__ inline bool keep_pair(int i, int j, float imi, float imj, int gi, int gj) {
  if (j <= i) return false;
  if (imi == 0.0f && imj == 0.0f) return false;
  if (gi == gj && gi > 0) return false;
  return true;
}

__device__ inline bool overlap(float4 alo, float4 ahi, float4 blo, float4 bhi) {
  return !(alo.x > bhi.x || blo.x > ahi.x || alo.y > bhi.y || blo.y > ahi.y || alo.z > bhi.z ||
           blo.z > ahi.z);
}
