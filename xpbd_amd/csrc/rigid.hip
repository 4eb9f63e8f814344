// 6-DOF rigid-body XPBD kernels for gfx950 and their C-ABI entry points
// (reference: src/xpbd3d/kernels_6dof.py; frame sequence of
// Solver6DOF._record, src/xpbd3d/solver_6dof.py:357-426).
//
// Device layout (float4 per body so each gather is one 16-byte load):
//   X  (x, y, z, inv_mass)   Q  (qx, qy, qz, qw)   V, W  (v / omega, 0)
//   XP, QP  substep-start position / orientation
//   IB (inverse body inertia diagonal, 0)   HE (half extents, group as int bits)
//   DX (sum dx, count)  DR (sum d-rotation, 0)    position Jacobi accumulators
//   DV (sum dv, count)  DW (sum d-omega, 0)       velocity-pass accumulators
//   LAMF float[8n]  floor-corner multipliers
// Contact manifold (generated once per substep, compacted):
//   MC int[cap]      points per candidate pair (the reference's m_count)
//   PTI int4[8cap]   (pair, incident body, reference body, slot)
//   PTN/PTOI/PTOR float4[8cap]  normal, incident / reference body-frame offsets
//   LAMP float[8cap] point multipliers;  PTC int[1] live point count
// Joints: JI int2 (a, b), JA float4 (anchor_a, rest), JB float4 (anchor_b,
// compliance), LAMJ float.
#include <cmath>

#include "common.h"
#include "rigid_math.h"
#include "rigid_sub.h"
#include "xpbd.h"

#define F4(p) ((float4*)(p))
#define CF4(p) ((const float4*)(p))

namespace {

constexpr int kB = 256;

__device__ inline void atomic_add3(float4* p, f3 v) {
  atomicAdd(&p->x, v.x);
  atomicAdd(&p->y, v.y);
  atomicAdd(&p->z, v.z);
}

// A counter written by an earlier launch (or atomically in this one).
__device__ inline int ld_count(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// kernels_6dof.py:108-130 with atomics (B may be -1 / static)
__device__ inline void accumulate_correction(float4* DX, float4* DR, int a, int b, float ima,
                                             float imb, q4 qa, q4 qb, f3 iIa, f3 iIb, f3 ra,
                                             f3 rb, f3 p) {
  if (ima > 0.0f) {
    atomic_add3(&DX[a], p * ima);
    atomic_add3(&DR[a], world_invI_mul(qa, iIa, cross(ra, p)));
    atomicAdd(&DX[a].w, 1.0f);
  }
  if (b >= 0 && imb > 0.0f) {
    atomic_add3(&DX[b], (-p) * imb);
    atomic_add3(&DR[b], -world_invI_mul(qb, iIb, cross(rb, p)));
    atomicAdd(&DX[b].w, 1.0f);
  }
}

// kernels_6dof.py:67-102 on values: one substep's prediction of a dynamic
// body from its start pose (x0, q) and velocity (v, w); returns the
// predicted position and updates q, v, w in place (w only with gyro).
__device__ inline f3 integrate_vals(f3 x0, q4& q, f3& v, f3& w, f3 ii, f3 g, float h,
                                    int gyro) {
  v = v + g * h;
  const f3 x = x0 + v * h;
  if (gyro != 0) {
    f3 wb = qrot_inv(q, w);
    float ix = ii.x > 0.0f ? 1.0f / ii.x : 0.0f;
    float iy = ii.y > 0.0f ? 1.0f / ii.y : 0.0f;
    float iz = ii.z > 0.0f ? 1.0f / ii.z : 0.0f;
    f3 Iw = mk3(ix * wb.x, iy * wb.y, iz * wb.z);
    f3 tau = -cross(wb, Iw);
    f3 dwb = mk3(ii.x * tau.x, ii.y * tau.y, ii.z * tau.z) * h;
    w = w + qrot(q, dwb);
  }
  q = quat_integrate(q, w, h);
  return x;
}

// gravity of body i (gscale: optional per-body factor, batched scenes)
__device__ inline f3 body_gravity(int i, float gx, float gy, float gz,
                                  const float* __restrict__ gscale) {
  f3 g = mk3(gx, gy, gz);
  if (gscale) g = g * gscale[i];
  return g;
}

// kernels_6dof.py:67-102 (+ per-substep lambda_floor reset and contact
// point counter reset, both exact rewrites of solver_6dof.py:374-375)
// gscale (optional): per-body gravity factor (batched independent scenes)
__device__ inline void integrate_elem(int i, float4* X, float4* Q, float4* V, float4* W,
                                                   float4* XP, float4* QP, const float4* IB,
                                                   float* LAMF, float gx, float gy,
                                                   float gz, float h, int gyro,
                                                   const float* __restrict__ gscale, int n) {
  float4 x4 = X[i];
  q4 q = ldq(Q[i]);
  XP[i] = make_float4(x4.x, x4.y, x4.z, 0.0f);
  QP[i] = stq(q);
  float4* lf = (float4*)(LAMF + 8 * (size_t)i);
  lf[0] = make_float4(0.f, 0.f, 0.f, 0.f);
  lf[1] = make_float4(0.f, 0.f, 0.f, 0.f);
  if (x4.w == 0.0f) return;
  f3 v = xyz(V[i]), w = xyz(W[i]);
  const f3 x = integrate_vals(xyz(x4), q, v, w, xyz(IB[i]), body_gravity(i, gx, gy, gz, gscale),
                              h, gyro);
  V[i] = make_float4(v.x, v.y, v.z, 0.0f);
  if (gyro != 0) W[i] = make_float4(w.x, w.y, w.z, 0.0f);
  X[i] = make_float4(x.x, x.y, x.z, x4.w);
  Q[i] = stq(q);
}

__global__ __launch_bounds__(kB) void k6_integrate(float4* X, float4* Q, float4* V, float4* W,
                                                   float4* XP, float4* QP, const float4* IB,
                                                   float* LAMF, int* PTC, float gx, float gy,
                                                   float gz, float h, int gyro,
                                                   const float* __restrict__ gscale, int n) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i == 0 && PTC) *PTC = 0;
  if (i >= n) return;
  integrate_elem(i, X, Q, V, W, XP, QP, IB, LAMF, gx, gy, gz, h, gyro, gscale, n);
}

// kernels_6dof.py:226-230
__device__ inline float proj_radius(f3 L, const m3& R, f3 e) {
  return fabsf(e.x * dot(L, R.c[0])) + fabsf(e.y * dot(L, R.c[1])) + fabsf(e.z * dot(L, R.c[2]));
}

// kernels_6dof.py:233-382: face-axis SAT + Sutherland-Hodgman clip of the
// incident face against the reference face, <= 8 points per pair.  The
// points of a wavefront are compacted into the point arrays with one atomic
// per wave (inclusive scan of the per-lane counts over the 64 lanes); all 64
// lanes of a wave call it together (pid >= live: no pair).
// The clipping polygons of one pair (two buffers of up to 8 points).
// LocalPoly: private arrays (scratch memory; the whole-frame workgroup
// kernel, whose LDS holds the body state).  LdsPoly: this lane's slice of a
// block-wide LDS array, lane-minor so the lanes of a wave hit consecutive
// banks (the fused launch: a dependent scratch access costs several times an
// LDS access, and the clip loop is a chain of them).
struct LocalPoly {
  f3 p[2][8];
  __device__ f3 get(int b, int k) const { return p[b][k]; }
  __device__ void set(int b, int k, f3 v) { p[b][k] = v; }
};
template <int STRIDE>
struct LdsPoly {
  float* base;  // &array[0][lane]
  __device__ f3 get(int b, int k) const {
    const int o = 3 * (8 * b + k) * STRIDE;
    return mk3(base[o], base[o + STRIDE], base[o + 2 * STRIDE]);
  }
  __device__ void set(int b, int k, f3 v) {
    const int o = 3 * (8 * b + k) * STRIDE;
    base[o] = v.x;
    base[o + STRIDE] = v.y;
    base[o + 2 * STRIDE] = v.z;
  }
};
constexpr int kPolyFloats = 2 * 8 * 3;

// The two boxes of a candidate pair as the manifold sees them (pose and
// half extents; inverse mass and inertia for the solve of its points).
struct PairBodies {
  int a, b;
  f3 cA, cB, eA, eB, iIA, iIB;
  q4 qa, qb;
  float imA, imB;
};

// wbase / wtotal: the wavefront's points are [wbase, wbase + wtotal) (the
// slots from 8 cap on are not written).
template <class Poly>
__device__ inline void box_manifold_pair(
    int pid, bool valid, const PairBodies& P, float margin, int cap, int* __restrict__ MC,
    int4* __restrict__ PTI, float4* __restrict__ PTN, float4* __restrict__ PTOI,
    float4* __restrict__ PTOR, float* __restrict__ LAMP, int* __restrict__ PTC, Poly& poly,
    int& wbase, int& wtotal) {
  const int lane = threadIdx.x & 63;
  int nc = 0, inc_b = 0, ref_b = 0;
  f3 ref_face_n = mk3(0.f, 0.f, 0.f);
  // the surviving polygon and the frames its points are expressed in
  int cur = 0, plen = 0;
  f3 ref_face_c = mk3(0.f, 0.f, 0.f), c_inc = ref_face_c, c_ref = ref_face_c;
  f3 ri0 = c_inc, ri1 = c_inc, ri2 = c_inc, rr0 = c_inc, rr1 = c_inc, rr2 = c_inc;  // frames
  if (valid) {
    const int a = P.a, b = P.b;
    const f3 cA = P.cA, cB = P.cB;
    const f3 eA = P.eA, eB = P.eB;
    const q4 qa = P.qa, qb = P.qb;
    const m3 RA = qmat(qa), RB = qmat(qb);
    const f3 cBA = cA - cB;
    float best_overlap = 1.0e30f;
    f3 best_axis = mk3(0.0f, 1.0f, 0.0f);
    int ref_is_a = 1, best_k = 0;
    bool sep = false;
#pragma unroll
    for (int k = 0; k < 3 && !sep; ++k) {
      f3 L = col(RA, k);
      float ov = comp(eA, k) + proj_radius(L, RB, eB) - fabsf(dot(L, cBA));
      if (ov < 0.0f) sep = true;
      else if (ov < best_overlap) { best_overlap = ov; best_axis = L; ref_is_a = 1; best_k = k; }
    }
#pragma unroll
    for (int k = 0; k < 3 && !sep; ++k) {
      f3 L = col(RB, k);
      float ov = comp(eB, k) + proj_radius(L, RA, eA) - fabsf(dot(L, cBA));
      if (ov < 0.0f) sep = true;
      else if (ov < best_overlap) { best_overlap = ov; best_axis = L; ref_is_a = 0; best_k = k; }
    }
    if (!sep) {
      f3 nBA = best_axis;
      if (dot(nBA, cBA) < 0.0f) nBA = -nBA;
      // reference and incident box by value selects; their frames are
      // recomputed from the quaternions (qmat / qrot of a unit axis: the
      // same operations on the same values as RA / RB), never picked out of
      // a matrix by a runtime index, so everything stays in registers
      const bool ra = ref_is_a == 1;
      ref_b = ra ? a : b;
      inc_b = ra ? b : a;
      c_ref = sel3(ra, cA, cB);
      c_inc = sel3(ra, cB, cA);
      const f3 e_ref = sel3(ra, eA, eB), e_inc = sel3(ra, eB, eA);
      const q4 q_ref = selq(ra, qa, qb), q_inc = selq(ra, qb, qa);
      ref_face_n = ra ? -nBA : nBA;
      const m3 R_ref = qmat(q_ref), R_inc = qmat(q_inc);
      const int ref_axis = best_k;
      ref_face_c = c_ref + ref_face_n * comp(e_ref, ref_axis);
      const int ax1 = (ref_axis + 1) % 3, ax2 = (ref_axis + 2) % 3;
      const f3 u_ref = qrot(q_ref, unit_axis(ax1)), v_ref = qrot(q_ref, unit_axis(ax2));
      const float eu = comp(e_ref, ax1), ev = comp(e_ref, ax2);
      int inc_axis = 0;
      float inc_sign = 1.0f, best_dot = -1.0e30f;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        float dp = -dot(col(R_inc, k), ref_face_n);
        if (dp > best_dot) { best_dot = dp; inc_axis = k; inc_sign = 1.0f; }
        if (-dp > best_dot) { best_dot = -dp; inc_axis = k; inc_sign = -1.0f; }
      }
      const f3 inc_n = qrot(q_inc, unit_axis(inc_axis)) * inc_sign;
      const f3 inc_face_c = c_inc + inc_n * comp(e_inc, inc_axis);
      const int iax1 = (inc_axis + 1) % 3, iax2 = (inc_axis + 2) % 3;
      const f3 iu = qrot(q_inc, unit_axis(iax1)), iv = qrot(q_inc, unit_axis(iax2));
      const float ieu = comp(e_inc, iax1), iev = comp(e_inc, iax2);
      poly.set(0, 0, (inc_face_c + iu * ieu) + iv * iev);
      poly.set(0, 1, (inc_face_c - iu * ieu) + iv * iev);
      poly.set(0, 2, (inc_face_c - iu * ieu) - iv * iev);
      poly.set(0, 3, (inc_face_c + iu * ieu) - iv * iev);
      plen = 4;
      bool dead = false;
      for (int pi = 0; pi < 4; ++pi) {
        f3 pp, pn;
        if (pi == 0) { pp = ref_face_c + u_ref * eu; pn = u_ref; }
        else if (pi == 1) { pp = ref_face_c - u_ref * eu; pn = -u_ref; }
        else if (pi == 2) { pp = ref_face_c + v_ref * ev; pn = v_ref; }
        else { pp = ref_face_c - v_ref * ev; pn = -v_ref; }
        if (plen == 0) { dead = true; break; }
        int olen = 0;
        const int nx = cur ^ 1;
        for (int k = 0; k < plen; ++k) {
          f3 ap = poly.get(cur, k);
          f3 bp = poly.get(cur, (k + 1) % plen);
          float da = dot(ap - pp, pn), db = dot(bp - pp, pn);
          if (da <= 0.0f) {
            if (olen < 8) poly.set(nx, olen++, ap);
            if (db > 0.0f) {
              float t = da / (da - db);
              if (olen < 8) poly.set(nx, olen++, ap + (bp - ap) * t);
            }
          } else if (db <= 0.0f) {
            float t = da / (da - db);
            if (olen < 8) poly.set(nx, olen++, ap + (bp - ap) * t);
          }
        }
        cur = nx;
        plen = olen;
      }
      if (dead) plen = 0;
      ri0 = R_inc.c[0], ri1 = R_inc.c[1], ri2 = R_inc.c[2];
      rr0 = R_ref.c[0], rr1 = R_ref.c[1], rr2 = R_ref.c[2];
      // count the points below the margin (the first 8); emitted below
      for (int k = 0; k < plen && nc < 8; ++k) {
        const f3 p_inc = poly.get(cur, k);
        if (dot(p_inc - ref_face_c, ref_face_n) < margin) nc += 1;
      }
    }
  }
  // wave-level compaction of the points
  int incl = nc;
  for (int d = 1; d < 64; d <<= 1) {
    int t = __shfl_up(incl, d, 64);
    if (lane >= d) incl += t;
  }
  const int total = __shfl(incl, 63, 64);
  int base = 0;
  if (lane == 63 && total > 0) base = atomicAdd(PTC, total);
  base = __shfl(base, 63, 64);
  const int first = base + incl - nc;
  for (int k = 0, c = 0; k < plen && c < nc; ++k) {
    const f3 p_inc = poly.get(cur, k);
    const float d_signed = dot(p_inc - ref_face_c, ref_face_n);
    if (!(d_signed < margin)) continue;
    const f3 p_ref = p_inc - ref_face_n * d_signed;
    const f3 vi = p_inc - c_inc, vr = p_ref - c_ref;
    const f3 oi = mk3(dot(ri0, vi), dot(ri1, vi), dot(ri2, vi));   // mtv(R_inc, vi)
    const f3 orf = mk3(dot(rr0, vr), dot(rr1, vr), dot(rr2, vr));  // mtv(R_ref, vr)
    const int q = first + c;
    if (q >= 8 * cap) break;  // a counter left non-zero: drop, never write past
    PTI[q] = make_int4(pid, inc_b, ref_b, c);
    PTN[q] = make_float4(ref_face_n.x, ref_face_n.y, ref_face_n.z, 0.0f);
    PTOI[q] = make_float4(oi.x, oi.y, oi.z, 0.0f);
    PTOR[q] = make_float4(orf.x, orf.y, orf.z, 0.0f);
    LAMP[q] = 0.0f;
    ++c;
  }
  if (pid < cap) MC[pid] = nc;
  wbase = base;
  wtotal = total;
}

template <class Poly>
__device__ inline void box_manifold_elem(
    int pid, int live, const float4* __restrict__ X, const float4* __restrict__ Q,
    const float4* __restrict__ HE, const int2* __restrict__ PAIRS, float margin, int cap,
    int* __restrict__ MC, int4* __restrict__ PTI, float4* __restrict__ PTN,
    float4* __restrict__ PTOI, float4* __restrict__ PTOR, float* __restrict__ LAMP,
    int* __restrict__ PTC, Poly& poly) {
  PairBodies P;
  const bool valid = pid < live;
  if (valid) {
    const int2 pr = PAIRS[pid];
    P.a = pr.x;
    P.b = pr.y;
    P.cA = xyz(X[P.a]);
    P.cB = xyz(X[P.b]);
    P.eA = xyz(HE[P.a]);
    P.eB = xyz(HE[P.b]);
    P.qa = ldq(Q[P.a]);
    P.qb = ldq(Q[P.b]);
  }
  int wbase, wtotal;
  box_manifold_pair(pid, valid, P, margin, cap, MC, PTI, PTN, PTOI, PTOR, LAMP, PTC, poly, wbase,
                    wtotal);
}

__global__ __launch_bounds__(kB) void k6_box_manifold(
    const float4* __restrict__ X, const float4* __restrict__ Q, const float4* __restrict__ HE,
    const int2* __restrict__ PAIRS, const int* __restrict__ PC, float margin, int cap,
    int* __restrict__ MC, int4* __restrict__ PTI, float4* __restrict__ PTN,
    float4* __restrict__ PTOI, float4* __restrict__ PTOR, float* __restrict__ LAMP,
    int* __restrict__ PTC) {
  __shared__ float psh[kPolyFloats * kB];
  LdsPoly<kB> poly{psh + threadIdx.x};
  box_manifold_elem(blockIdx.x * kB + threadIdx.x, min(*PC, cap), X, Q, HE, PAIRS, margin, cap,
                    MC, PTI, PTN, PTOI, PTOR, LAMP, PTC, poly);
}

// kernels_6dof.py:440-483 (lambda reset folded into the first iteration) on
// the poses of the joint's bodies: xa4 / xb4 = (position, inverse mass).
__device__ inline void solve_joint_vals(int k, int a, int b, float4 xa4, q4 qa, f3 iIa,
                                        float4 xb4, q4 qb, f3 iIb, float4 ja, float4 jb,
                                        float* __restrict__ LAMJ, float h, int first,
                                        float4* DX, float4* DR) {
  float lam = first ? 0.0f : LAMJ[k];
  if (first) LAMJ[k] = 0.0f;
  f3 pa = xyz(xa4) + qrot(qa, xyz(ja));
  f3 pb = xyz(xb4) + qrot(qb, xyz(jb));
  f3 d = pa - pb;
  float dist = length(d);
  if (dist < XPBD_EPS) return;
  f3 nn = d / dist;
  float C = dist - ja.w;
  f3 ra = pa - xyz(xa4), rb = pb - xyz(xb4);
  float w_a = gen_inv_mass(xa4.w, qa, iIa, ra, nn);
  float w_b = gen_inv_mass(xb4.w, qb, iIb, rb, nn);
  float wsum = w_a + w_b;
  if (wsum <= 0.0f) return;
  float alpha = jb.w / (h * h);
  float dlam = (-C - alpha * lam) / (wsum + alpha);
  LAMJ[k] = lam + dlam;
  f3 p = nn * dlam;
  accumulate_correction(DX, DR, a, b, xa4.w, xb4.w, qa, qb, iIa, iIb, ra, rb, p);
}

__device__ inline void solve_joints_elem(int k,
    const float4* __restrict__ X, const float4* __restrict__ Q, const float4* __restrict__ IB,
    const int2* __restrict__ JI, const float4* __restrict__ JA, const float4* __restrict__ JB,
    float* __restrict__ LAMJ, float h, int first, float4* DX, float4* DR, int nj) {
  const int2 ab = JI[k];
  const int a = ab.x, b = ab.y;
  solve_joint_vals(k, a, b, X[a], ldq(Q[a]), xyz(IB[a]), X[b], ldq(Q[b]), xyz(IB[b]), JA[k],
                   JB[k], LAMJ, h, first, DX, DR);
}

__global__ __launch_bounds__(kB) void k6_solve_joints(
    const float4* __restrict__ X, const float4* __restrict__ Q, const float4* __restrict__ IB,
    const int2* __restrict__ JI, const float4* __restrict__ JA, const float4* __restrict__ JB,
    float* __restrict__ LAMJ, float h, int first, float4* DX, float4* DR, int nj) {
  const int k = blockIdx.x * kB + threadIdx.x;
  if (k >= nj) return;
  solve_joints_elem(k, X, Q, IB, JI, JA, JB, LAMJ, h, first, DX, DR, nj);
}

// kernels_6dof.py:158-211.  One thread owns one body, so its corners add to
// DX/DR in corner order with plain loads and stores.
template <bool ATOMIC>
__device__ inline void solve_floor_elem(int i, 
    const float4* __restrict__ X, const float4* __restrict__ Q, const float4* __restrict__ XP,
    const float4* __restrict__ QP, const float4* __restrict__ IB, const float4* __restrict__ HE,
    float floor_y, float mu_s, float* __restrict__ LAMF, float4* DX, float4* DR, int n) {
  const float4 x4 = X[i];
  const float im = x4.w;
  if (im == 0.0f) return;
  const f3 x = xyz(x4);
  const q4 q = ldq(Q[i]);
  const f3 iI = xyz(IB[i]);
  const f3 he = xyz(HE[i]);
  const f3 s = mk3(0.0f, 1.0f, 0.0f);
  // ATOMIC (launches shared with other row kinds): sum this body's corners
  // from zero and add the sums atomically; otherwise one thread owns DX[i]
  float4 dx = ATOMIC ? make_float4(0.f, 0.f, 0.f, 0.f) : DX[i];
  float4 dr = ATOMIC ? make_float4(0.f, 0.f, 0.f, 0.f) : DR[i];
  bool touched = false;
  float lam[8];
  for (int c = 0; c < 8; ++c) {
    lam[c] = 0.0f;
    const f3 corner = box_corner(he, c);
    const f3 p1 = x + qrot(q, corner);
    const float d = floor_y - p1.y;
    if (d <= 0.0f) continue;
    const f3 r = p1 - x;
    const float w = gen_inv_mass(im, q, iI, r, s);
    if (w <= 0.0f) continue;
    const float dlam = d / w;
    lam[c] = dlam;
    const f3 p = s * dlam;
    const f3 lin = p * im, ang = world_invI_mul(q, iI, cross(r, p));
    dx.x += lin.x; dx.y += lin.y; dx.z += lin.z;
    dr.x += ang.x; dr.y += ang.y; dr.z += ang.z;
    dx.w += 1.0f;
    touched = true;
    if (mu_s > 0.0f) {
      const f3 p1b = xyz(XP[i]) + qrot(ldq(QP[i]), corner);
      const f3 dp = p1 - p1b;
      const f3 dpt = dp - s * dot(dp, s);
      const float lt = length(dpt);
      if (lt > XPBD_EPS && lt < mu_s * d) {
        const float w2 = gen_inv_mass(im, q, iI, r, dpt / lt);
        if (w2 > 0.0f) {
          const f3 pf = (-dpt) / w2;
          const f3 lf = pf * im, af = world_invI_mul(q, iI, cross(r, pf));
          dx.x += lf.x; dx.y += lf.y; dx.z += lf.z;
          dr.x += af.x; dr.y += af.y; dr.z += af.z;
        }
      }
    }
  }
  float4* lf = (float4*)(LAMF + 8 * (size_t)i);
  lf[0] = make_float4(lam[0], lam[1], lam[2], lam[3]);
  lf[1] = make_float4(lam[4], lam[5], lam[6], lam[7]);
  if (touched) {
    if (ATOMIC) {
      atomic_add3(&DX[i], xyz(dx));
      atomicAdd(&DX[i].w, dx.w);
      atomic_add3(&DR[i], xyz(dr));
    } else {
      DX[i] = dx;
      DR[i] = dr;
    }
  }
}

__global__ __launch_bounds__(kB) void k6_solve_floor(
    const float4* __restrict__ X, const float4* __restrict__ Q, const float4* __restrict__ XP,
    const float4* __restrict__ QP, const float4* __restrict__ IB, const float4* __restrict__ HE,
    float floor_y, float mu_s, float* __restrict__ LAMF, float4* DX, float4* DR, int n) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  solve_floor_elem<false>(i, X, Q, XP, QP, IB, HE, floor_y, mu_s, LAMF, DX, DR, n);
}

// kernels_6dof.py:385-430, one thread per compacted manifold point
__device__ inline void solve_box_elem(int k, 
    const float4* __restrict__ X, const float4* __restrict__ Q, const float4* __restrict__ IB,
    const int4* __restrict__ PTI, const float4* __restrict__ PTN,
    const float4* __restrict__ PTOI, const float4* __restrict__ PTOR, float* __restrict__ LAMP,
    const int* __restrict__ PTC, float4* DX, float4* DR, int max_points) {
  const int4 id = PTI[k];
  const int inc = id.y, ref = id.z;
  const f3 s = xyz(PTN[k]);
  const float4 xi4 = X[inc], xr4 = X[ref];
  const q4 qi = ldq(Q[inc]), qr = ldq(Q[ref]);
  const f3 p_inc = xyz(xi4) + qrot(qi, xyz(PTOI[k]));
  const f3 p_ref = xyz(xr4) + qrot(qr, xyz(PTOR[k]));
  const float C = dot(p_inc - p_ref, s);
  if (C >= 0.0f) {
    LAMP[k] = 0.0f;
    return;
  }
  const float d = -C;
  const f3 r_inc = p_inc - xyz(xi4), r_ref = p_ref - xyz(xr4);
  const f3 iIi = xyz(IB[inc]), iIr = xyz(IB[ref]);
  const float w_inc = gen_inv_mass(xi4.w, qi, iIi, r_inc, s);
  const float w_ref = gen_inv_mass(xr4.w, qr, iIr, r_ref, s);
  const float wsum = w_inc + w_ref;
  if (wsum <= 0.0f) return;
  const float dlam = d / wsum;
  LAMP[k] = LAMP[k] + dlam;
  const f3 p = s * dlam;
  accumulate_correction(DX, DR, inc, ref, xi4.w, xr4.w, qi, qr, iIi, iIr, r_inc, r_ref, p);
}

__global__ __launch_bounds__(kB) void k6_solve_box(
    const float4* __restrict__ X, const float4* __restrict__ Q, const float4* __restrict__ IB,
    const int4* __restrict__ PTI, const float4* __restrict__ PTN,
    const float4* __restrict__ PTOI, const float4* __restrict__ PTOR, float* __restrict__ LAMP,
    const int* __restrict__ PTC, float4* DX, float4* DR, int max_points) {
  const int k = blockIdx.x * kB + threadIdx.x;
  if (k >= min(*PTC, max_points)) return;
  solve_box_elem(k, X, Q, IB, PTI, PTN, PTOI, PTOR, LAMP, PTC, DX, DR, max_points);
}

// kernels_6dof.py:486-506 on values: the Jacobi sums (acc = sum dx, count;
// drs = sum d-rotation) applied to the pose (x4 = position, inverse mass).
__device__ inline void apply_jacobi_vals(float4 acc, f3 drs, float relax, float4& x4, q4& q) {
  if (acc.w > 0.0f && x4.w > 0.0f) {
    const float inv_n = relax / acc.w;
    const f3 x = xyz(x4) + xyz(acc) * inv_n;
    const f3 dw = drs * inv_n;
    const q4 wq{dw.x, dw.y, dw.z, 0.0f};
    x4 = make_float4(x.x, x.y, x.z, x4.w);
    q = qnormalize(qadd(q, qscale(qmul(wq, q), 0.5f)));
  }
}

__device__ inline void apply_jacobi_elem(int i, float4* X, float4* Q, float4* DX,
                                                      float4* DR, float relax, int n) {
  const float4 acc = DX[i];
  float4 x4 = X[i];
  if (acc.w > 0.0f && x4.w > 0.0f) {
    q4 q = ldq(Q[i]);
    apply_jacobi_vals(acc, xyz(DR[i]), relax, x4, q);
    X[i] = x4;
    Q[i] = stq(q);
  }
  if (acc.w != 0.0f || acc.x != 0.0f || acc.y != 0.0f || acc.z != 0.0f) {
    DX[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    DR[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

__global__ __launch_bounds__(kB) void k6_apply_jacobi(float4* X, float4* Q, float4* DX,
                                                      float4* DR, float relax, int n) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  apply_jacobi_elem(i, X, Q, DX, DR, relax, n);
}

// kernels_6dof.py:512-535 on values: velocities of a body from its pose
// (x4 = position, inverse mass) and substep-start pose; zero when static.
__device__ inline void finalize_vals(float4 x4, q4 q, float4 xp4, q4 qp, float h, float lin_damp,
                                     float ang_damp, f3& v, f3& w) {
  if (x4.w == 0.0f) {
    v = mk3(0.f, 0.f, 0.f);
    w = v;
    return;
  }
  v = ((xyz(x4) - xyz(xp4)) / h) * (1.0f - lin_damp);
  const q4 dq = qmul(q, qconj(qp));
  w = mk3(dq.x, dq.y, dq.z) * (2.0f / h);
  if (dq.w < 0.0f) w = -w;
  w = w * (1.0f - ang_damp);
}

__device__ inline void finalize_elem(int i, const float4* __restrict__ X,
                                                  const float4* __restrict__ Q,
                                                  const float4* __restrict__ XP,
                                                  const float4* __restrict__ QP, float4* V,
                                                  float4* W, float h, float lin_damp,
                                                  float ang_damp, int n) {
  const float4 x4 = X[i];
  f3 v, w;
  finalize_vals(x4, ldq(Q[i]), XP[i], ldq(QP[i]), h, lin_damp, ang_damp, v, w);
  V[i] = make_float4(v.x, v.y, v.z, 0.0f);
  W[i] = make_float4(w.x, w.y, w.z, 0.0f);
}

__global__ __launch_bounds__(kB) void k6_finalize(const float4* __restrict__ X,
                                                  const float4* __restrict__ Q,
                                                  const float4* __restrict__ XP,
                                                  const float4* __restrict__ QP, float4* V,
                                                  float4* W, float h, float lin_damp,
                                                  float ang_damp, int n) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  finalize_elem(i, X, Q, XP, QP, V, W, h, lin_damp, ang_damp, n);
}

// Shared tail of velocity_floor / velocity_box (kernels_6dof.py:575-583):
// dynamic friction Eq. 30 and restitution.
__device__ inline f3 velocity_impulse(f3 vrel, f3 s, float lam_n, float mu_d, float restitution,
                                      float h) {
  const float vn = dot(vrel, s);
  const f3 vt = vrel - s * vn;
  const float ltan = length(vt);
  f3 impulse = mk3(0.0f, 0.0f, 0.0f);
  if (mu_d > 0.0f && ltan > XPBD_EPS) {
    const float fn = lam_n / (h * h);
    const float dvm = fminf(h * mu_d * fabsf(fn), ltan);
    impulse = impulse - (vt / ltan) * dvm;
  }
  if (vn < 0.0f) impulse = impulse + s * (-vn + fmaxf(-restitution * vn, 0.0f));
  return impulse;
}

// kernels_6dof.py:541-590; one thread owns one body (plain RMW of DV/DW)
template <bool ATOMIC>
__device__ inline void velocity_floor_elem(int i, 
    const float4* __restrict__ X, const float4* __restrict__ Q, const float4* __restrict__ IB,
    const float4* __restrict__ HE, float mu_d, float restitution,
    const float* __restrict__ LAMF, float h, const float4* __restrict__ V,
    const float4* __restrict__ W, float4* DV, float4* DW, int n) {
  const float im = X[i].w;
  if (im == 0.0f) return;
  const float4* lf4 = (const float4*)(LAMF + 8 * (size_t)i);
  const float4 l0 = lf4[0], l1 = lf4[1];
  const float lam[8] = {l0.x, l0.y, l0.z, l0.w, l1.x, l1.y, l1.z, l1.w};
  const q4 q = ldq(Q[i]);
  const f3 iI = xyz(IB[i]), he = xyz(HE[i]);
  const f3 v = xyz(V[i]), w = xyz(W[i]);
  const f3 s = mk3(0.0f, 1.0f, 0.0f);
  float4 dv = ATOMIC ? make_float4(0.f, 0.f, 0.f, 0.f) : DV[i];
  float4 dw = ATOMIC ? make_float4(0.f, 0.f, 0.f, 0.f) : DW[i];
  bool touched = false;
  for (int c = 0; c < 8; ++c) {
    const float lam_n = lam[c];
    if (lam_n == 0.0f) continue;
    const f3 r = qrot(q, box_corner(he, c));
    const f3 vc = v + cross(w, r);
    const f3 impulse = velocity_impulse(vc, s, lam_n, mu_d, restitution, h);
    if (length(impulse) > XPBD_EPS) {
      const float wg = gen_inv_mass(im, q, iI, r, impulse / length(impulse));
      if (wg > 0.0f) {
        const f3 p = impulse / wg;
        const f3 lin = p * im, ang = world_invI_mul(q, iI, cross(r, p));
        dv.x += lin.x; dv.y += lin.y; dv.z += lin.z; dv.w += 1.0f;
        dw.x += ang.x; dw.y += ang.y; dw.z += ang.z;
        touched = true;
      }
    }
  }
  if (touched) {
    if (ATOMIC) {
      atomic_add3(&DV[i], xyz(dv));
      atomicAdd(&DV[i].w, dv.w);
      atomic_add3(&DW[i], xyz(dw));
    } else {
      DV[i] = dv;
      DW[i] = dw;
    }
  }
}

__global__ __launch_bounds__(kB) void k6_velocity_floor(
    const float4* __restrict__ X, const float4* __restrict__ Q, const float4* __restrict__ IB,
    const float4* __restrict__ HE, float mu_d, float restitution,
    const float* __restrict__ LAMF, float h, const float4* __restrict__ V,
    const float4* __restrict__ W, float4* DV, float4* DW, int n) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  velocity_floor_elem<false>(i, X, Q, IB, HE, mu_d, restitution, LAMF, h, V, W, DV, DW, n);
}

// kernels_6dof.py:593-659 on values: the impulse p of one manifold point
// (multiplier lam_n, normal s, body-frame offsets oi / orf) between bodies
// i and j, given their poses (xi4 / xj4 = position, inverse mass),
// velocities and inverse inertia; false when it applies none.  ri / rj:
// the point's arms.
__device__ inline bool velocity_box_impulse(float lam_n, f3 s, f3 oi, f3 orf, float4 xi4, q4 qi,
                                            f3 vi, f3 wi, f3 iIi, float4 xj4, q4 qj, f3 vj,
                                            f3 wj, f3 iIj, float mu_d, float restitution,
                                            float h, f3& p, f3& ri, f3& rj) {
  const f3 p_inc = xyz(xi4) + qrot(qi, oi);
  const f3 p_ref = xyz(xj4) + qrot(qj, orf);
  ri = p_inc - xyz(xi4);
  rj = p_ref - xyz(xj4);
  const f3 vci = vi + cross(wi, ri);
  const f3 vcj = vj + cross(wj, rj);
  const f3 impulse = velocity_impulse(vci - vcj, s, lam_n, mu_d, restitution, h);
  const float li = length(impulse);
  if (!(li > XPBD_EPS)) return false;
  const f3 ndir = impulse / li;
  const float wi_ = gen_inv_mass(xi4.w, qi, iIi, ri, ndir);
  const float wj_ = gen_inv_mass(xj4.w, qj, iIj, rj, ndir);
  const float wsum = wi_ + wj_;
  if (!(wsum > 0.0f)) return false;
  p = impulse / wsum;
  return true;
}

__device__ inline void velocity_box_vals(float lam_n, int ci, int cj, f3 s, f3 oi, f3 orf,
                                         float4 xi4, q4 qi, f3 vi, f3 wi, f3 iIi, float4 xj4,
                                         q4 qj, f3 vj, f3 wj, f3 iIj, float mu_d,
                                         float restitution, float h, float4* DV, float4* DW) {
  f3 p, ri, rj;
  if (!velocity_box_impulse(lam_n, s, oi, orf, xi4, qi, vi, wi, iIi, xj4, qj, vj, wj, iIj, mu_d,
                            restitution, h, p, ri, rj))
    return;
  if (xi4.w > 0.0f) {
    atomic_add3(&DV[ci], p * xi4.w);
    atomic_add3(&DW[ci], world_invI_mul(qi, iIi, cross(ri, p)));
    atomicAdd(&DV[ci].w, 1.0f);
  }
  if (xj4.w > 0.0f) {
    atomic_add3(&DV[cj], (-p) * xj4.w);
    atomic_add3(&DW[cj], -world_invI_mul(qj, iIj, cross(rj, p)));
    atomicAdd(&DV[cj].w, 1.0f);
  }
}

// kernels_6dof.py:593-659, one thread per compacted manifold point
__device__ inline void velocity_box_elem(int k,
    const float4* __restrict__ X, const float4* __restrict__ Q, const float4* __restrict__ IB,
    const int4* __restrict__ PTI, const float4* __restrict__ PTN,
    const float4* __restrict__ PTOI, const float4* __restrict__ PTOR,
    const float* __restrict__ LAMP, const int* __restrict__ PTC, float mu_d, float restitution,
    float h, const float4* __restrict__ V, const float4* __restrict__ W, float4* DV, float4* DW,
    int max_points) {
  const float lam_n = LAMP[k];
  if (lam_n == 0.0f) return;
  const int4 id = PTI[k];
  const int ci = id.y, cj = id.z;
  velocity_box_vals(lam_n, ci, cj, xyz(PTN[k]), xyz(PTOI[k]), xyz(PTOR[k]), X[ci], ldq(Q[ci]),
                    xyz(V[ci]), xyz(W[ci]), xyz(IB[ci]), X[cj], ldq(Q[cj]), xyz(V[cj]),
                    xyz(W[cj]), xyz(IB[cj]), mu_d, restitution, h, DV, DW);
}

__global__ __launch_bounds__(kB) void k6_velocity_box(
    const float4* __restrict__ X, const float4* __restrict__ Q, const float4* __restrict__ IB,
    const int4* __restrict__ PTI, const float4* __restrict__ PTN,
    const float4* __restrict__ PTOI, const float4* __restrict__ PTOR,
    const float* __restrict__ LAMP, const int* __restrict__ PTC, float mu_d, float restitution,
    float h, const float4* __restrict__ V, const float4* __restrict__ W, float4* DV, float4* DW,
    int max_points) {
  const int k = blockIdx.x * kB + threadIdx.x;
  if (k >= min(*PTC, max_points)) return;
  velocity_box_elem(k, X, Q, IB, PTI, PTN, PTOI, PTOR, LAMP, PTC, mu_d, restitution, h, V, W, DV, DW, max_points);
}

// kernels_6dof.py:662-691 on values: apply_velocity_6dof (acc = sum dv,
// count; dws = sum d-omega) then cap_velocity_6dof; true when v or w changed.
__device__ inline bool apply_velocity_vals(float4 acc, f3 dws, float im, float max_lin,
                                           float max_ang, int do_cap, f3& v, f3& w) {
  bool dirty = false;
  if (acc.w > 0.0f && im > 0.0f) {
    const float inv_n = 1.0f / acc.w;
    v = v + xyz(acc) * inv_n;
    w = w + dws * inv_n;
    dirty = true;
  }
  if (do_cap) {
    const float sp = length(v);
    if (max_lin > 0.0f && sp > max_lin) { v = v * (max_lin / sp); dirty = true; }
    const float sa = length(w);
    if (max_ang > 0.0f && sa > max_ang) { w = w * (max_ang / sa); dirty = true; }
  }
  return dirty;
}

// kernels_6dof.py:662-691: apply_velocity_6dof then cap_velocity_6dof
__device__ inline void apply_velocity_elem(int i, const float4* __restrict__ X, float4* V,
                                                        float4* W, float4* DV, float4* DW,
                                                        float max_lin, float max_ang, int do_cap,
                                                        int n) {
  const float4 acc = DV[i];
  f3 v = xyz(V[i]), w = xyz(W[i]);
  const bool dirty =
      apply_velocity_vals(acc, xyz(DW[i]), X[i].w, max_lin, max_ang, do_cap, v, w);
  if (acc.w != 0.0f || acc.x != 0.0f || acc.y != 0.0f || acc.z != 0.0f) {
    DV[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    DW[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  if (dirty) {
    V[i] = make_float4(v.x, v.y, v.z, 0.0f);
    W[i] = make_float4(w.x, w.y, w.z, 0.0f);
  }
}

__global__ __launch_bounds__(kB) void k6_apply_velocity(const float4* __restrict__ X, float4* V,
                                                        float4* W, float4* DV, float4* DW,
                                                        float max_lin, float max_ang, int do_cap,
                                                        int n) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  apply_velocity_elem(i, X, V, W, DV, DW, max_lin, max_ang, do_cap, n);
}

// ---------------------------------------------------------------------------
// Fused launches of the substep (frame plan): consecutive per-body kernels
// run as one pass, and the row kinds of one Jacobi sweep (joints, floor
// corners, box points) as one launch over their concatenated index ranges.
// Rows of different kinds then add into the same bodies concurrently, so the
// floor rows add their per-body sums atomically too; the accumulation order
// of a Jacobi sweep is arbitrary in the reference as well (atomic_add).
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Floor rows of the fused launches, one box corner per lane: 8 consecutive
// lanes (a group starting at a lane index divisible by 8) take body i's
// corners 0..7; each lane stages its corner's contributions in LDS and the
// group's first lane adds them in corner order, exactly as the per-body loop
// of solve_floor_elem / velocity_floor_elem does, then adds the body's sums
// atomically.  Same results, a quarter of the dependent work per lane.
// `slot` is this block's LDS staging area, 4 float4 per lane.
// ---------------------------------------------------------------------------
__device__ inline void stage_wait() {
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
}

// Body i's pose (x4 = position, inverse mass; q), substep-start pose (xs,
// qs), inverse inertia and half extents come in as values.
__device__ inline void solve_floor_corner_vals(int i, int c, float4 x4, q4 q, f3 xs, q4 qs,
    f3 iI, f3 he, float floor_y, float mu_s, float* __restrict__ LAMF, float4* DX, float4* DR,
    float4* slot) {
  const float im = x4.w;
  f3 lin = mk3(0.f, 0.f, 0.f), ang = lin, lf = lin, af = lin;
  float touched = 0.0f, fric = 0.0f, lam = 0.0f;
  if (im != 0.0f) {
    const f3 x = xyz(x4);
    const f3 s = mk3(0.0f, 1.0f, 0.0f);
    const f3 corner = box_corner(he, c);
    const f3 p1 = x + qrot(q, corner);
    const float d = floor_y - p1.y;
    if (d > 0.0f) {
      const f3 r = p1 - x;
      const float w = gen_inv_mass(im, q, iI, r, s);
      if (w > 0.0f) {
        const float dlam = d / w;
        lam = dlam;
        const f3 p = s * dlam;
        lin = p * im;
        ang = world_invI_mul(q, iI, cross(r, p));
        touched = 1.0f;
        if (mu_s > 0.0f) {
          const f3 p1b = xs + qrot(qs, corner);
          const f3 dp = p1 - p1b;
          const f3 dpt = dp - s * dot(dp, s);
          const float lt = length(dpt);
          if (lt > XPBD_EPS && lt < mu_s * d) {
            const float w2 = gen_inv_mass(im, q, iI, r, dpt / lt);
            if (w2 > 0.0f) {
              const f3 pf = (-dpt) / w2;
              lf = pf * im;
              af = world_invI_mul(q, iI, cross(r, pf));
              fric = 1.0f;
            }
          }
        }
      }
    }
    LAMF[8 * (size_t)i + c] = lam;
  }
  const int lane = threadIdx.x;
  slot[4 * lane + 0] = make_float4(lin.x, lin.y, lin.z, touched);
  slot[4 * lane + 1] = make_float4(ang.x, ang.y, ang.z, fric);
  slot[4 * lane + 2] = make_float4(lf.x, lf.y, lf.z, 0.0f);
  slot[4 * lane + 3] = make_float4(af.x, af.y, af.z, 0.0f);
  stage_wait();
  if (c != 0 || im == 0.0f) return;
  float4 dx = make_float4(0.f, 0.f, 0.f, 0.f), dr = dx;
  bool any = false;
  for (int k = 0; k < 8; ++k) {
    const float4 a = slot[4 * (lane + k) + 0];
    if (a.w == 0.0f) continue;
    const float4 b = slot[4 * (lane + k) + 1];
    dx.x += a.x; dx.y += a.y; dx.z += a.z;
    dr.x += b.x; dr.y += b.y; dr.z += b.z;
    dx.w += 1.0f;
    any = true;
    if (b.w != 0.0f) {
      const float4 e = slot[4 * (lane + k) + 2], g = slot[4 * (lane + k) + 3];
      dx.x += e.x; dx.y += e.y; dx.z += e.z;
      dr.x += g.x; dr.y += g.y; dr.z += g.z;
    }
  }
  if (any) {
    atomic_add3(&DX[i], xyz(dx));
    atomicAdd(&DX[i].w, dx.w);
    atomic_add3(&DR[i], xyz(dr));
  }
}

__device__ inline void solve_floor_corner(int i, int c,
    const float4* __restrict__ X, const float4* __restrict__ Q, const float4* __restrict__ XP,
    const float4* __restrict__ QP, const float4* __restrict__ IB, const float4* __restrict__ HE,
    float floor_y, float mu_s, float* __restrict__ LAMF, float4* DX, float4* DR,
    float4* slot) {
  solve_floor_corner_vals(i, c, X[i], ldq(Q[i]), xyz(XP[i]), ldq(QP[i]), xyz(IB[i]),
                          xyz(HE[i]), floor_y, mu_s, LAMF, DX, DR, slot);
}

// Body i's inverse mass, orientation, inverse inertia, half extents,
// velocities and this corner's multiplier (lam, read only when im != 0)
// come in as values.
__device__ inline void velocity_floor_corner_vals(int i, int c, float im, q4 q, f3 iI, f3 he,
    f3 v, f3 w, float lam, float mu_d, float restitution, float h, float4* DV, float4* DW,
    float4* slot) {
  f3 lin = mk3(0.f, 0.f, 0.f), ang = lin;
  float touched = 0.0f;
  const float lam_n = im != 0.0f ? lam : 0.0f;
  if (lam_n != 0.0f) {
    const f3 r = qrot(q, box_corner(he, c));
    const f3 vc = v + cross(w, r);
    const f3 s = mk3(0.0f, 1.0f, 0.0f);
    const f3 impulse = velocity_impulse(vc, s, lam_n, mu_d, restitution, h);
    if (length(impulse) > XPBD_EPS) {
      const float wg = gen_inv_mass(im, q, iI, r, impulse / length(impulse));
      if (wg > 0.0f) {
        const f3 p = impulse / wg;
        lin = p * im;
        ang = world_invI_mul(q, iI, cross(r, p));
        touched = 1.0f;
      }
    }
  }
  const int lane = threadIdx.x;
  slot[4 * lane + 0] = make_float4(lin.x, lin.y, lin.z, touched);
  slot[4 * lane + 1] = make_float4(ang.x, ang.y, ang.z, 0.0f);
  stage_wait();
  if (c != 0 || im == 0.0f) return;
  float4 dv = make_float4(0.f, 0.f, 0.f, 0.f), dw = dv;
  bool any = false;
  for (int k = 0; k < 8; ++k) {
    const float4 a = slot[4 * (lane + k) + 0];
    if (a.w == 0.0f) continue;
    const float4 b = slot[4 * (lane + k) + 1];
    dv.x += a.x; dv.y += a.y; dv.z += a.z; dv.w += 1.0f;
    dw.x += b.x; dw.y += b.y; dw.z += b.z;
    any = true;
  }
  if (any) {
    atomic_add3(&DV[i], xyz(dv));
    atomicAdd(&DV[i].w, dv.w);
    atomic_add3(&DW[i], xyz(dw));
  }
}

__device__ inline void velocity_floor_corner(int i, int c,
    const float4* __restrict__ X, const float4* __restrict__ Q, const float4* __restrict__ IB,
    const float4* __restrict__ HE, float mu_d, float restitution,
    const float* __restrict__ LAMF, float h, const float4* __restrict__ V,
    const float4* __restrict__ W, float4* DV, float4* DW, float4* slot) {
  velocity_floor_corner_vals(i, c, X[i].w, ldq(Q[i]), xyz(IB[i]), xyz(HE[i]), xyz(V[i]),
                             xyz(W[i]), LAMF[8 * (size_t)i + c], mu_d, restitution, h, DV, DW,
                             slot);
}

// First lane of the floor corners in k6_solve_all: joints end at nj, the
// corners start at the next multiple of 64, so every 8-lane body group lies
// inside one wavefront.
__host__ __device__ inline int floor_lane0(int nj) { return (nj + 63) & ~63; }

// apply_velocity + cap of the previous substep (apply != 0), then integrate
__global__ __launch_bounds__(kB) void k6_begin_substep(
    float4* X, float4* Q, float4* V, float4* W, float4* XP, float4* QP, const float4* IB,
    float* LAMF, int* PTC, float4* DV, float4* DW, float max_lin, float max_ang, int do_cap,
    int apply, float gx, float gy, float gz, float h, int gyro, const float* __restrict__ gscale,
    int n) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i == 0 && PTC) *PTC = 0;
  if (i >= n) return;
  if (apply) apply_velocity_elem(i, X, V, W, DV, DW, max_lin, max_ang, do_cap, n);
  integrate_elem(i, X, Q, V, W, XP, QP, IB, LAMF, gx, gy, gz, h, gyro, gscale, n);
}

// joints [0, nj), floor corners [f0, f0 + 8 n) with f0 = floor_lane0(nj), box
// points [f0 + 8 n, f0 + 8 n + max_points)
__global__ __launch_bounds__(kB) void k6_solve_all(
    const float4* __restrict__ X, const float4* __restrict__ Q, const float4* __restrict__ XP,
    const float4* __restrict__ QP, const float4* __restrict__ IB, const float4* __restrict__ HE,
    const int2* __restrict__ JI, const float4* __restrict__ JA, const float4* __restrict__ JB,
    float* __restrict__ LAMJ, float h, int first, int nj, float floor_y, float mu_s,
    float* __restrict__ LAMF, int n, const int4* __restrict__ PTI,
    const float4* __restrict__ PTN, const float4* __restrict__ PTOI,
    const float4* __restrict__ PTOR, float* __restrict__ LAMP, const int* __restrict__ PTC,
    int max_points, float4* DX, float4* DR) {
  __shared__ float4 slot[4 * kB];
  const int k = blockIdx.x * kB + threadIdx.x;
  const int f0 = floor_lane0(nj);
  if (k < nj) {
    solve_joints_elem(k, X, Q, IB, JI, JA, JB, LAMJ, h, first, DX, DR, nj);
  } else if (k < f0) {
    return;
  } else if (k < f0 + 8 * n) {
    const int t = k - f0;
    solve_floor_corner(t >> 3, t & 7, X, Q, XP, QP, IB, HE, floor_y, mu_s, LAMF, DX, DR, slot);
  } else {
    const int p = k - f0 - 8 * n;
    if (p < min(*PTC, max_points))
      solve_box_elem(p, X, Q, IB, PTI, PTN, PTOI, PTOR, LAMP, PTC, DX, DR, max_points);
  }
}

// apply_jacobi, then finalize_velocity on the last iteration
__global__ __launch_bounds__(kB) void k6_apply_finalize(float4* X, float4* Q, float4* DX,
                                                        float4* DR, float relax, int finalize,
                                                        const float4* __restrict__ XP,
                                                        const float4* __restrict__ QP, float4* V,
                                                        float4* W, float h, float lin_damp,
                                                        float ang_damp, int n) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  apply_jacobi_elem(i, X, Q, DX, DR, relax, n);
  if (finalize) finalize_elem(i, X, Q, XP, QP, V, W, h, lin_damp, ang_damp, n);
}

// velocity pass: floor corners [0, 8 n), box points [8 n, 8 n + max_points)
__global__ __launch_bounds__(kB) void k6_velocity_all(
    const float4* __restrict__ X, const float4* __restrict__ Q, const float4* __restrict__ IB,
    const float4* __restrict__ HE, float mu_d, float restitution, const float* __restrict__ LAMF,
    float h, const float4* __restrict__ V, const float4* __restrict__ W, int n,
    const int4* __restrict__ PTI, const float4* __restrict__ PTN,
    const float4* __restrict__ PTOI, const float4* __restrict__ PTOR,
    const float* __restrict__ LAMP, const int* __restrict__ PTC, int max_points, float4* DV,
    float4* DW) {
  __shared__ float4 slot[4 * kB];
  const int k = blockIdx.x * kB + threadIdx.x;
  if (k < 8 * n) {
    velocity_floor_corner(k >> 3, k & 7, X, Q, IB, HE, mu_d, restitution, LAMF, h, V, W, DV, DW,
                          slot);
  } else {
    const int p = k - 8 * n;
    if (p < min(*PTC, max_points))
      velocity_box_elem(p, X, Q, IB, PTI, PTN, PTOI, PTOR, LAMP, PTC, mu_d, restitution, h, V, W,
                        DV, DW, max_points);
  }
}

// ---------------------------------------------------------------------------
// Two launches per substep (frame plan, iterations >= 1).  Each per-body
// phase is folded into the launch that needs its result: a lane recomputes
// the per-body operations it depends on from the values the previous launch
// left (the same operations on the same values as the separate launches, so
// the same results), and the corner-0 lane of each body stores them once.
//   contacts (substep s): apply_velocity + cap of substep s-1 (s > 0) and
//     integrate, then joints, floor corners and the box manifold of every
//     candidate pair with each point solved at once: the first Jacobi
//     iteration.  Stores the predicted pose and the substep-start pose.
//   velocity (substep s): apply_jacobi of the last iteration and
//     finalize_velocity, then floor corners and box points of the velocity
//     pass.  Stores x, q, v, w.
//   end of frame: apply_velocity + cap of the last substep.
// Buffer parities: rigid_sub.h.  Lanes: contacts [0, nj) joints, [f0, f0 +
// 8 n) floor corners, [p0, p0 + cap) pairs; velocity [0, 8 n) floor
// corners, [pv0, pv0 + 8 cap) points; p0 and pv0 start a workgroup, so the
// pair and point workgroups hold nothing else (they meet at block barriers).
// A pair workgroup runs the manifold of its pairs, one per lane, then the
// first Jacobi iteration of the points they emitted, one per lane.
// ---------------------------------------------------------------------------

// Body i at the end of the contacts launch's per-body phase: predicted pose
// (x4 = position, inverse mass; q), substep-start pose, inverse inertia.
struct Pred {
  float4 x4;
  q4 q;
  f3 xs;
  q4 qs;
  f3 iI;
};

__device__ inline Pred predict_body(const xpbd_rigid_desc& d, const SubBufs& b, int i) {
  const float4 x4 = CF4(d.x)[i];
  const q4 q0 = ldq(CF4(d.q)[i]);
  f3 v = xyz(CF4(d.v)[i]), w = xyz(CF4(d.w)[i]);
  Pred p;
  p.iI = xyz(CF4(d.inv_I)[i]);
  p.xs = xyz(x4);
  p.qs = q0;
  p.x4 = x4;
  p.q = q0;
  if (b.apply)
    (void)apply_velocity_vals(b.DVin[i], xyz(b.DWin[i]), x4.w, d.max_speed, d.max_omega,
                              d.max_speed > 0.0f, v, w);
  if (x4.w != 0.0f) {
    q4 q = q0;
    const f3 x = integrate_vals(p.xs, q, v, w, p.iI, body_gravity(i, d.gx, d.gy, d.gz, d.gscale),
                                d.h, d.gyroscopic);
    p.x4 = make_float4(x.x, x.y, x.z, x4.w);
    p.q = q;
  }
  return p;
}

// First lane of the box points in the velocity launch (and of the pairs in
// the contacts launch): a workgroup boundary, so the point workgroups hold
// nothing else and every lane of them reaches their barrier.
__host__ __device__ inline int block_lane0(int lanes) { return (lanes + kB - 1) / kB * kB; }

// One point's contributions to the two bodies of its pair (the terms
// accumulate_correction / velocity_box_elem add atomically).
struct PointSums {
  f3 li, ai, lj, aj;  // linear / angular terms on body i and body j
  float ni, nj;       // 1 when the point adds to that body
};

// Block-wide: the contributions of consecutive lanes with the same key (the
// points of one pair, at most 8, in point order) are summed in the first
// lane of each run and added once to its bodies bi / bj (A: sum of linear
// terms and count, B: sum of angular terms).  One chain of atomics per body
// and pair instead of one per point.  key < 0: no contribution.  Every lane
// of the block calls it.
__device__ inline void add_pair_runs(int key, const PointSums& c, int bi, int bj, float4* A,
                                     float4* B, float4* sh, int* skey) {
  const int t = threadIdx.x;
  sh[4 * t + 0] = make_float4(c.li.x, c.li.y, c.li.z, c.ni);
  sh[4 * t + 1] = make_float4(c.ai.x, c.ai.y, c.ai.z, 0.0f);
  sh[4 * t + 2] = make_float4(c.lj.x, c.lj.y, c.lj.z, c.nj);
  sh[4 * t + 3] = make_float4(c.aj.x, c.aj.y, c.aj.z, 0.0f);
  skey[t] = key;
  __syncthreads();
  if (key < 0 || (t > 0 && skey[t - 1] == key)) return;
  float4 si = make_float4(0.f, 0.f, 0.f, 0.f), ri = si, sj = si, rj = si;
  for (int u = t; u < kB && u < t + 8 && skey[u] == key; ++u) {
    const float4 e0 = sh[4 * u + 0], e1 = sh[4 * u + 1], e2 = sh[4 * u + 2], e3 = sh[4 * u + 3];
    if (e0.w != 0.0f) {
      si.x += e0.x; si.y += e0.y; si.z += e0.z; si.w += e0.w;
      ri.x += e1.x; ri.y += e1.y; ri.z += e1.z;
    }
    if (e2.w != 0.0f) {
      sj.x += e2.x; sj.y += e2.y; sj.z += e2.z; sj.w += e2.w;
      rj.x += e3.x; rj.y += e3.y; rj.z += e3.z;
    }
  }
  if (si.w > 0.0f) {
    atomic_add3(&A[bi], xyz(si));
    atomic_add3(&B[bi], xyz(ri));
    atomicAdd(&A[bi].w, si.w);
  }
  if (sj.w > 0.0f) {
    atomic_add3(&A[bj], xyz(sj));
    atomic_add3(&B[bj], xyz(rj));
    atomicAdd(&A[bj].w, sj.w);
  }
}

// Workgroup-local: the first point of the solve phase's point r (of the
// points the block's four wavefronts emitted) and its global slot.
__device__ inline int block_point_slot(int r, const int* wbase, const int* wtot) {
  for (int w = 0; w < kB / 64; ++w) {
    if (r < wtot[w]) return wbase[w] + r;
    r -= wtot[w];
  }
  return -1;
}

__global__ __launch_bounds__(kB) void k6_sub_contacts(xpbd_rigid_desc d, SubBufs b) {
  __shared__ float psh[kPolyFloats * kB];
  __shared__ float4 slot[4 * kB];
  __shared__ float4 pose[6 * kB];  // per pair: (c, im), q, iI of its bodies a and b
  __shared__ int skey[kB];
  __shared__ int wbase[kB / 64], wtot[kB / 64];
  const int k = blockIdx.x * kB + threadIdx.x;
  const int n = d.n, nj = d.nj;
  const int f0 = floor_lane0(nj), p0 = block_lane0(f0 + 8 * n);
  const float4* HE = CF4(d.he);
  const int pc = ld_count(d.pair_count);
  const bool skip = pc > d.cap;  // broad-phase overflow: the frame writes nothing
  if (k < nj) {
    const int2 ab = ((const int2*)d.j_idx)[k];
    const Pred pa = predict_body(d, b, ab.x), pb = predict_body(d, b, ab.y);
    if (skip) return;
    solve_joint_vals(k, ab.x, ab.y, pa.x4, pa.q, pa.iI, pb.x4, pb.q, pb.iI, CF4(d.j_a)[k],
                     CF4(d.j_b)[k], d.lam_j, d.h, 1, b.DX, b.DR);
    return;
  }
  if (k < p0) {
    if (k < f0 || k >= f0 + 8 * n) return;
    const int t = k - f0, i = t >> 3, c = t & 7;
    const Pred p = predict_body(d, b, i);
    if (skip) return;
    if (c == 0) {
      b.XPR[i] = p.x4;
      b.QPR[i] = stq(p.q);
      F4(d.x_prev)[i] = make_float4(p.xs.x, p.xs.y, p.xs.z, 0.0f);
      F4(d.q_prev)[i] = stq(p.qs);
      b.DV[i] = make_float4(0.f, 0.f, 0.f, 0.f);
      b.DW[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    solve_floor_corner_vals(i, c, p.x4, p.q, p.xs, p.qs, p.iI, xyz(HE[i]), d.floor_y, d.friction,
                            d.lam_floor, b.DX, b.DR, slot);
    return;
  }
  // candidate pairs: whole workgroups from here on (block barriers below);
  // skip and the live count are uniform over the grid / the workgroup
  const int t = threadIdx.x;
  const int pid0 = blockIdx.x * kB - p0;  // this workgroup's first pair
  const int live = min(pc, d.cap);
  if (skip || pid0 >= live) return;
  const int pid = pid0 + t;
  PairBodies P;
  const bool valid = pid < live;
  if (valid) {
    const int2 pr = ((const int2*)d.pairs)[pid];
    const Pred pa = predict_body(d, b, pr.x), pb = predict_body(d, b, pr.y);
    P.a = pr.x;
    P.b = pr.y;
    P.cA = xyz(pa.x4);
    P.cB = xyz(pb.x4);
    P.qa = pa.q;
    P.qb = pb.q;
    P.imA = pa.x4.w;
    P.imB = pb.x4.w;
    P.iIA = pa.iI;
    P.iIB = pb.iI;
    P.eA = xyz(HE[pr.x]);
    P.eB = xyz(HE[pr.y]);
    pose[6 * t + 0] = pa.x4;
    pose[6 * t + 1] = stq(pa.q);
    pose[6 * t + 2] = make_float4(pa.iI.x, pa.iI.y, pa.iI.z, __int_as_float(pr.x));
    pose[6 * t + 3] = pb.x4;
    pose[6 * t + 4] = stq(pb.q);
    pose[6 * t + 5] = make_float4(pb.iI.x, pb.iI.y, pb.iI.z, __int_as_float(pr.y));
  }
  // the box manifold (solver_6dof.py:376-384), one pair per lane
  LdsPoly<kB> poly{psh + t};
  int wb, wt;
  box_manifold_pair(pid, valid, P, d.man_margin, d.cap, d.m_count, (int4*)d.pt_idx, F4(d.pt_n),
                    F4(d.pt_oi), F4(d.pt_or), d.lam_pt, b.PTC, poly, wb, wt);
  const int maxp = 8 * d.cap;
  if ((t & 63) == 0) {
    wbase[t >> 6] = wb;
    wtot[t >> 6] = max(0, min(wt, maxp - wb));  // slots past 8 cap were not written
  }
  __syncthreads();  // the block's points and their pairs' poses are visible
  int total = 0;
  for (int w = 0; w < kB / 64; ++w) total += wtot[w];
  // the first Jacobi iteration over the block's points, one per lane
  // (solve_box_elem on the stored values: the body-frame offsets taken back
  // to the world through the same predicted poses); each pair's corrections
  // summed in LDS and added once (solver_6dof.py:385-401)
  for (int r0 = 0; r0 < total; r0 += kB) {
    const int q = r0 + t < total ? block_point_slot(r0 + t, wbase, wtot) : -1;
    PointSums c;
    c.li = c.ai = c.lj = c.aj = mk3(0.f, 0.f, 0.f);
    c.ni = c.nj = 0.0f;
    int key = -1, bi = 0, bj = 0;
    if (q >= 0) {
      const int4 id = ((const int4*)d.pt_idx)[q];
      const f3 s = xyz(CF4(d.pt_n)[q]);
      const f3 oi = xyz(CF4(d.pt_oi)[q]), orf = xyz(CF4(d.pt_or)[q]);
      const int lp = id.x - pid0;  // the pair's lane in this block
      const bool inc_a = __float_as_int(pose[6 * lp + 2].w) == id.y;
      const int ii = 6 * lp + (inc_a ? 0 : 3), ir = 6 * lp + (inc_a ? 3 : 0);
      const float4 xi4 = pose[ii], xr4 = pose[ir];
      const q4 qi = ldq(pose[ii + 1]), qr = ldq(pose[ir + 1]);
      const f3 iIi = xyz(pose[ii + 2]), iIr = xyz(pose[ir + 2]);
      key = id.x;
      bi = id.y;
      bj = id.z;
      const f3 pi = xyz(xi4) + qrot(qi, oi);
      const f3 pr = xyz(xr4) + qrot(qr, orf);
      const float C = dot(pi - pr, s);
      if (!(C >= 0.0f)) {
        const f3 r_inc = pi - xyz(xi4), r_ref = pr - xyz(xr4);
        const float wsum =
            gen_inv_mass(xi4.w, qi, iIi, r_inc, s) + gen_inv_mass(xr4.w, qr, iIr, r_ref, s);
        if (!(wsum <= 0.0f)) {
          const float dlam = -C / wsum;
          d.lam_pt[q] = 0.0f + dlam;
          const f3 pp = s * dlam;  // accumulate_correction's terms
          if (xi4.w > 0.0f) {
            c.li = pp * xi4.w;
            c.ai = world_invI_mul(qi, iIi, cross(r_inc, pp));
            c.ni = 1.0f;
          }
          if (xr4.w > 0.0f) {
            c.lj = (-pp) * xr4.w;
            c.aj = -world_invI_mul(qr, iIr, cross(r_ref, pp));
            c.nj = 1.0f;
          }
        }
      }
    }
    add_pair_runs(key, c, bi, bj, b.DX, b.DR, slot, skey);
    __syncthreads();  // slot / skey are reused by the next round
  }
}

// Body i after the last iteration's apply_jacobi and finalize_velocity.
struct Fin {
  float4 x4;
  q4 q;
  f3 v, w, iI;
};

__device__ inline Fin finalize_body(const xpbd_rigid_desc& d, const SubBufs& b, int i) {
  Fin f;
  f.x4 = b.XPR[i];
  f.q = ldq(b.QPR[i]);
  const float4 acc = b.DX[i];
  const f3 drs = xyz(b.DR[i]);
  const float4 xp4 = CF4(d.x_prev)[i];
  const q4 qp = ldq(CF4(d.q_prev)[i]);
  f.iI = xyz(CF4(d.inv_I)[i]);
  apply_jacobi_vals(acc, drs, d.relax, f.x4, f.q);
  finalize_vals(f.x4, f.q, xp4, qp, d.h, d.lin_damp, d.ang_damp, f.v, f.w);
  return f;
}

__global__ __launch_bounds__(kB) void k6_sub_velocity(xpbd_rigid_desc d, SubBufs b) {
  __shared__ float4 slot[4 * kB];
  __shared__ int skey[kB];
  const int k = blockIdx.x * kB + threadIdx.x;
  const int n = d.n;
  const bool skip = ld_count(d.pair_count) > d.cap;  // as the contacts launch
  const int pv0 = block_lane0(8 * n);
  if (k < pv0) {
    if (k >= 8 * n) return;
    const int i = k >> 3, c = k & 7;
    const Fin f = finalize_body(d, b, i);
    const float lam = d.lam_floor[8 * (size_t)i + c];
    if (skip) return;
    if (c == 0) {
      F4(d.x)[i] = f.x4;
      F4(d.q)[i] = stq(f.q);
      F4(d.v)[i] = make_float4(f.v.x, f.v.y, f.v.z, 0.0f);
      F4(d.w)[i] = make_float4(f.w.x, f.w.y, f.w.z, 0.0f);
      b.DXn[i] = make_float4(0.f, 0.f, 0.f, 0.f);
      b.DRn[i] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (i == 0) *b.PTCn = 0;
    }
    velocity_floor_corner_vals(i, c, f.x4.w, f.q, f.iI, xyz(CF4(d.he)[i]), f.v, f.w, lam,
                               d.friction, d.restitution, d.h, b.DV, b.DW, slot);
    return;
  }
  // box points, one per lane; whole workgroups from here on (the block
  // barrier of add_pair_runs): a workgroup leaves together or not at all
  const int p = k - pv0, maxp = 8 * d.cap;
  const int cnt = min(ld_count(b.PTC), maxp);
  if (skip || (p - (int)threadIdx.x) >= cnt) return;  // uniform per workgroup
  float lam_n = 0.0f;
  int4 id = make_int4(-1, 0, 0, 0);
  float4 sn = make_float4(0.f, 0.f, 0.f, 0.f), oi = sn, orf = sn;
  if (p < cnt) {
    lam_n = d.lam_pt[p];
    id = ((const int4*)d.pt_idx)[p];
    sn = CF4(d.pt_n)[p];
    oi = CF4(d.pt_oi)[p];
    orf = CF4(d.pt_or)[p];
  }
  PointSums c;
  c.li = c.ai = c.lj = c.aj = mk3(0.f, 0.f, 0.f);
  c.ni = c.nj = 0.0f;
  if (p < cnt && lam_n != 0.0f) {  // velocity_box_elem
    const Fin fi = finalize_body(d, b, id.y), fj = finalize_body(d, b, id.z);
    f3 pp, ri, rj;
    if (velocity_box_impulse(lam_n, xyz(sn), xyz(oi), xyz(orf), fi.x4, fi.q, fi.v, fi.w, fi.iI,
                             fj.x4, fj.q, fj.v, fj.w, fj.iI, d.friction, d.restitution, d.h, pp,
                             ri, rj)) {
      if (fi.x4.w > 0.0f) {
        c.li = pp * fi.x4.w;
        c.ai = world_invI_mul(fi.q, fi.iI, cross(ri, pp));
        c.ni = 1.0f;
      }
      if (fj.x4.w > 0.0f) {
        c.lj = (-pp) * fj.x4.w;
        c.aj = -world_invI_mul(fj.q, fj.iI, cross(rj, pp));
        c.nj = 1.0f;
      }
    }
  }
  add_pair_runs(p < cnt ? id.x : -1, c, id.y, id.z, b.DV, b.DW, slot, skey);
}

// apply_velocity + cap of the frame's last substep (its sums in last.DV /
// last.DW) unless the frame was skipped (broad-phase overflow); every sum
// buffer and both point counters back to zero either way.  With a status
// ring (host-mapped int[2 * ring_len]) thread 0 also reports the frame:
// (pair count, frame number + 1) in slot frame % ring_len, the number
// stored last and with release order, so a host that reads it sees the count.
__global__ __launch_bounds__(kB) void k6_sub_end(xpbd_rigid_desc d, SubBufs last, int* ring,
                                                 int ring_len, int* frames) {
  const int i = blockIdx.x * kB + threadIdx.x;
  const int pc = ld_count(d.pair_count);
  if (i == 0) {
    d.pt_count[0] = 0;
    d.pt_count[1] = 0;
    if (ring) {
      const int f = frames[0];
      frames[0] = f + 1;
      int* slot = ring + 2 * (f % ring_len);
      __hip_atomic_store(slot, pc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(slot + 1, (f + 1) & 0x7fffffff, __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
  if (i >= d.n) return;
  const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
  if (pc <= d.cap) {
    apply_velocity_elem(i, CF4(d.x), F4(d.v), F4(d.w), last.DV, last.DW, d.max_speed,
                        d.max_omega, d.max_speed > 0.0f, d.n);
  } else {
    last.DV[i] = z;
    last.DW[i] = z;
  }
  float4* DVo = (float4*)last.DVin;  // the other parity
  float4* DWo = (float4*)last.DWin;
  DVo[i] = z;
  DWo[i] = z;
  last.DX[i] = z;
  last.DR[i] = z;
  last.DXn[i] = z;
  last.DRn[i] = z;
}

// ---------------------------------------------------------------------------
// Whole frames of a small rigid scene in one workgroup (no reference
// counterpart as a kernel: it runs solver_6dof.py:428-454 with the LBVH
// filter and the substep loop of solver_6dof.py:357-426 for `frames`
// frames).  The broad phase tests all pairs (i, j > i) against the same
// AABBs and filter as the LBVH path and emits them in (i, j) order through a
// workgroup scan of the per-body counts, so the pair list equals the LBVH
// path's; the substep phases are the elem functions of the fused launches,
// separated by workgroup barriers (__syncthreads is a workgroup-scope
// release/acquire: all waves of the workgroup share one CU's L1, so stores,
// and the L2 atomics that sum the accumulators, are visible to every wave
// after it without a device-scope fence).  Counters written inside the
// launch (pair and point counts) are read back with atomic loads, never
// through the scalar cache.  The body state (12 float4 arrays and the floor
// multipliers, 240 B per body) lives in LDS for the whole launch: up to
// kFBMax bodies.  cap >= n (n - 1) / 2, so the pair list cannot overflow.
// ---------------------------------------------------------------------------
constexpr int kFB = 1024;
constexpr int kFBMax = 256;

__device__ inline void fb_sync() { __syncthreads(); }


__global__ __launch_bounds__(kFB) void k6_frame_block(xpbd_rigid_desc d, float4* __restrict__ LO,
                                                     float4* __restrict__ HI, float dt,
                                                     int frames) {
  __shared__ int scnt[kFBMax + 1];
  __shared__ float4 sb[12][kFBMax];
  __shared__ float slf[8 * kFBMax];
  __shared__ float4 fslot[4 * kFB];  // floor-corner staging (4 float4 per lane)
  const int tid = threadIdx.x;
  const int n = d.n;
  float4* const G[12] = {F4(d.x),      F4(d.q),       F4(d.v),     F4(d.w),
                         F4(d.x_prev), F4(d.q_prev),  F4(d.inv_I), F4(d.he),
                         F4(d.dx),     F4(d.drot),    F4(d.dv),    F4(d.dw)};
  for (int k = 0; k < 12; ++k)
    for (int i = tid; i < n; i += kFB) sb[k][i] = G[k][i];
  for (int i = tid; i < 8 * n; i += kFB) slf[i] = d.lam_floor[i];
  __syncthreads();
  float4* X = sb[0];
  float4* Q = sb[1];
  float4* V = sb[2];
  float4* W = sb[3];
  float4* XP = sb[4];
  float4* QP = sb[5];
  const float4* IB = sb[6];
  const float4* HE = sb[7];
  float4* DX = sb[8];
  float4* DR = sb[9];
  float4* DV = sb[10];
  float4* DW = sb[11];
  float* LAMF = slf;
  int2* PAIRS = (int2*)d.pairs;
  int4* PTI = (int4*)d.pt_idx;
  float4* PTN = F4(d.pt_n);
  float4* PTOI = F4(d.pt_oi);
  float4* PTOR = F4(d.pt_or);
  const int2* JI = (const int2*)d.j_idx;
  const float4* JA = CF4(d.j_a);
  const float4* JB = CF4(d.j_b);
  const int maxp = 8 * d.cap;
  const int cap_v = d.max_speed > 0.0f;
  for (int fr = 0; fr < frames; ++fr) {
    // ---- broad phase: AABBs, per-body partner counts, scan, ordered emit
    for (int i = tid; i < n; i += kFB)
      body_aabb(X[i], Q[i], V[i], W[i], HE[i], d.man_margin, dt, LO[i], HI[i]);
    fb_sync();
    for (int i = tid; i < kFBMax; i += kFB) {
      int c = 0;
      if (i < n) {
        const float4 lo = LO[i], hi = HI[i];
        const float imi = X[i].w;
        const int gi = __float_as_int(HE[i].w);
        for (int j = i + 1; j < n; ++j)
          if (overlap(lo, hi, LO[j], HI[j]) && keep_pair(i, j, imi, X[j].w, gi, __float_as_int(HE[j].w)))
            ++c;
      }
      scnt[i] = c;
    }
    __syncthreads();
    if (tid == 0) {  // exclusive scan (n <= 1024: serial in LDS)
      int acc = 0;
      for (int i = 0; i < kFBMax; ++i) {
        const int c = scnt[i];
        scnt[i] = acc;
        acc += c;
      }
      scnt[kFBMax] = acc;
      ((int*)d.pair_count)[0] = acc;  // this launch is the broad phase too
    }
    __syncthreads();
    for (int i = tid; i < n; i += kFB) {
      int k = scnt[i];
      const float4 lo = LO[i], hi = HI[i];
      const float imi = X[i].w;
      const int gi = __float_as_int(HE[i].w);
      for (int j = i + 1; j < n; ++j)
        if (overlap(lo, hi, LO[j], HI[j]) && keep_pair(i, j, imi, X[j].w, gi, __float_as_int(HE[j].w))) {
          if (k < d.cap) PAIRS[k] = make_int2(i, j);
          ++k;
        }
    }
    const int live = min(scnt[kFBMax], d.cap);
    fb_sync();
    // ---- substeps (the fused launches' phases)
    for (int sub = 0; sub < d.substeps; ++sub) {
      if (tid == 0) d.pt_count[0] = 0;
      for (int i = tid; i < n; i += kFB) {
        if (sub > 0) apply_velocity_elem(i, X, V, W, DV, DW, d.max_speed, d.max_omega, cap_v, n);
        integrate_elem(i, X, Q, V, W, XP, QP, IB, LAMF, d.gx, d.gy, d.gz, d.h,
                       d.gyroscopic, d.gscale, n);
      }
      fb_sync();
      for (int base = 0; base < live; base += kFB) {  // whole waves call the manifold
        LocalPoly poly;
        box_manifold_elem(base + tid, live, X, Q, HE, PAIRS, d.man_margin, d.cap, d.m_count, PTI,
                          PTN, PTOI, PTOR, d.lam_pt, d.pt_count, poly);
      }
      fb_sync();
      const int npts = min(ld_count(d.pt_count), maxp);
      for (int it = 0; it < d.iterations; ++it) {
        // joints, floor corners (one per lane, as k6_solve_all), box points;
        // the stride is a multiple of 64, so corner groups stay in a wave.
        // Corners per lane only while everything fits one pass of the
        // workgroup (a second pass costs more than the per-body loop saves)
        const int f0 = floor_lane0(d.nj);
        if (f0 + 8 * n + npts > kFB) {
          for (int k = tid; k < d.nj + n + npts; k += kFB) {
            if (k < d.nj)
              solve_joints_elem(k, X, Q, IB, JI, JA, JB, d.lam_j, d.h, it == 0, DX, DR, d.nj);
            else if (k < d.nj + n)
              solve_floor_elem<true>(k - d.nj, X, Q, XP, QP, IB, HE, d.floor_y, d.friction,
                                     LAMF, DX, DR, n);
            else
              solve_box_elem(k - d.nj - n, X, Q, IB, PTI, PTN, PTOI, PTOR, d.lam_pt,
                             d.pt_count, DX, DR, maxp);
          }
        } else {
        for (int k = tid; k < f0 + 8 * n + npts; k += kFB) {
          if (k < d.nj) {
            solve_joints_elem(k, X, Q, IB, JI, JA, JB, d.lam_j, d.h, it == 0, DX, DR, d.nj);
          } else if (k < f0) {
            continue;
          } else if (k < f0 + 8 * n) {
            const int t = k - f0;
            solve_floor_corner(t >> 3, t & 7, X, Q, XP, QP, IB, HE, d.floor_y, d.friction, LAMF,
                               DX, DR, fslot);
          } else {
            solve_box_elem(k - f0 - 8 * n, X, Q, IB, PTI, PTN, PTOI, PTOR, d.lam_pt,
                           d.pt_count, DX, DR, maxp);
          }
        }
        }
        fb_sync();
        for (int i = tid; i < n; i += kFB) {
          apply_jacobi_elem(i, X, Q, DX, DR, d.relax, n);
          if (it + 1 == d.iterations)
            finalize_elem(i, X, Q, XP, QP, V, W, d.h, d.lin_damp, d.ang_damp, n);
        }
        fb_sync();
      }
      if (d.iterations == 0) {
        for (int i = tid; i < n; i += kFB)
          finalize_elem(i, X, Q, XP, QP, V, W, d.h, d.lin_damp, d.ang_damp, n);
        fb_sync();
      }
      if (8 * n + npts > kFB) {  // as the solve phase: corners only in one pass
        for (int k = tid; k < n + npts; k += kFB) {
          if (k < n)
            velocity_floor_elem<true>(k, X, Q, IB, HE, d.friction, d.restitution, LAMF, d.h,
                                      V, W, DV, DW, n);
          else
            velocity_box_elem(k - n, X, Q, IB, PTI, PTN, PTOI, PTOR, d.lam_pt, d.pt_count,
                              d.friction, d.restitution, d.h, V, W, DV, DW, maxp);
        }
      } else {
        for (int k = tid; k < 8 * n + npts; k += kFB) {
          if (k < 8 * n)
            velocity_floor_corner(k >> 3, k & 7, X, Q, IB, HE, d.friction, d.restitution,
                                  LAMF, d.h, V, W, DV, DW, fslot);
          else
            velocity_box_elem(k - 8 * n, X, Q, IB, PTI, PTN, PTOI, PTOR, d.lam_pt,
                              d.pt_count, d.friction, d.restitution, d.h, V, W, DV, DW, maxp);
        }
      }
      fb_sync();
    }
    for (int i = tid; i < n; i += kFB)
      apply_velocity_elem(i, X, V, W, DV, DW, d.max_speed, d.max_omega, cap_v, n);
    fb_sync();
  }
  for (int k = 0; k < 12; ++k)
    if (k != 6 && k != 7)  // inverse inertia and half extents are read-only
      for (int i = tid; i < n; i += kFB) G[k][i] = sb[k][i];
  for (int i = tid; i < 8 * n; i += kFB) d.lam_floor[i] = slf[i];
}

inline int nblk(long long n) { return (int)((n + kB - 1) / kB); }

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

int xpbd6_integrate(void* stream, float* X, float* Q, float* V, float* W, float* XP, float* QP,
                    const float* IB, float* LAMF, int* PTC, float gx, float gy, float gz,
                    float h, int gyro, const float* gscale, int n) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k6_integrate, dim3(nblk(n)), dim3(kB), 0, (hipStream_t)stream, F4(X), F4(Q),
                     F4(V), F4(W), F4(XP), F4(QP), CF4(IB), LAMF, PTC, gx, gy, gz, h, gyro,
                     gscale, n);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd6_box_manifold(void* stream, const float* X, const float* Q, const float* HE,
                       const int* PAIRS, const int* PC, float margin, int cap, int* MC, int* PTI,
                       float* PTN, float* PTOI, float* PTOR, float* LAMP, int* PTC) {
  if (cap <= 0) return 0;
  hipLaunchKernelGGL(k6_box_manifold, dim3(nblk(cap)), dim3(kB), 0, (hipStream_t)stream, CF4(X),
                     CF4(Q), CF4(HE), (const int2*)PAIRS, PC, margin, cap, MC, (int4*)PTI,
                     F4(PTN), F4(PTOI), F4(PTOR), LAMP, PTC);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd6_solve_joints(void* stream, const float* X, const float* Q, const float* IB,
                       const int* JI, const float* JA, const float* JB, float* LAMJ, float h,
                       int first, float* DX, float* DR, int nj) {
  if (nj <= 0) return 0;
  hipLaunchKernelGGL(k6_solve_joints, dim3(nblk(nj)), dim3(kB), 0, (hipStream_t)stream, CF4(X),
                     CF4(Q), CF4(IB), (const int2*)JI, CF4(JA), CF4(JB), LAMJ, h, first, F4(DX),
                     F4(DR), nj);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd6_solve_floor(void* stream, const float* X, const float* Q, const float* XP,
                      const float* QP, const float* IB, const float* HE, float floor_y,
                      float mu_s, float* LAMF, float* DX, float* DR, int n) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k6_solve_floor, dim3(nblk(n)), dim3(kB), 0, (hipStream_t)stream, CF4(X),
                     CF4(Q), CF4(XP), CF4(QP), CF4(IB), CF4(HE), floor_y, mu_s, LAMF, F4(DX),
                     F4(DR), n);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd6_solve_box(void* stream, const float* X, const float* Q, const float* IB,
                    const int* PTI, const float* PTN, const float* PTOI, const float* PTOR,
                    float* LAMP, const int* PTC, float* DX, float* DR, int max_points) {
  if (max_points <= 0) return 0;
  hipLaunchKernelGGL(k6_solve_box, dim3(nblk(max_points)), dim3(kB), 0, (hipStream_t)stream,
                     CF4(X), CF4(Q), CF4(IB), (const int4*)PTI, CF4(PTN), CF4(PTOI), CF4(PTOR),
                     LAMP, PTC, F4(DX), F4(DR), max_points);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd6_apply_jacobi(void* stream, float* X, float* Q, float* DX, float* DR, float relax,
                       int n) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k6_apply_jacobi, dim3(nblk(n)), dim3(kB), 0, (hipStream_t)stream, F4(X),
                     F4(Q), F4(DX), F4(DR), relax, n);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd6_finalize_velocity(void* stream, const float* X, const float* Q, const float* XP,
                            const float* QP, float* V, float* W, float h, float lin_damp,
                            float ang_damp, int n) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k6_finalize, dim3(nblk(n)), dim3(kB), 0, (hipStream_t)stream, CF4(X),
                     CF4(Q), CF4(XP), CF4(QP), F4(V), F4(W), h, lin_damp, ang_damp, n);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd6_velocity_floor(void* stream, const float* X, const float* Q, const float* IB,
                         const float* HE, float mu_d, float restitution, const float* LAMF,
                         float h, const float* V, const float* W, float* DV, float* DW, int n) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k6_velocity_floor, dim3(nblk(n)), dim3(kB), 0, (hipStream_t)stream, CF4(X),
                     CF4(Q), CF4(IB), CF4(HE), mu_d, restitution, LAMF, h, CF4(V), CF4(W),
                     F4(DV), F4(DW), n);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd6_velocity_box(void* stream, const float* X, const float* Q, const float* IB,
                       const int* PTI, const float* PTN, const float* PTOI, const float* PTOR,
                       const float* LAMP, const int* PTC, float mu_d, float restitution, float h,
                       const float* V, const float* W, float* DV, float* DW, int max_points) {
  if (max_points <= 0) return 0;
  hipLaunchKernelGGL(k6_velocity_box, dim3(nblk(max_points)), dim3(kB), 0, (hipStream_t)stream,
                     CF4(X), CF4(Q), CF4(IB), (const int4*)PTI, CF4(PTN), CF4(PTOI), CF4(PTOR),
                     LAMP, PTC, mu_d, restitution, h, CF4(V), CF4(W), F4(DV), F4(DW),
                     max_points);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd6_begin_substep(void* stream, float* X, float* Q, float* V, float* W, float* XP,
                        float* QP, const float* IB, float* LAMF, int* PTC, float* DV, float* DW,
                        float max_lin, float max_ang, int do_cap, int apply, float gx, float gy,
                        float gz, float h, int gyro, const float* gscale, int n) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k6_begin_substep, dim3(nblk(n)), dim3(kB), 0, (hipStream_t)stream, F4(X),
                     F4(Q), F4(V), F4(W), F4(XP), F4(QP), CF4(IB), LAMF, PTC, F4(DV), F4(DW),
                     max_lin, max_ang, do_cap, apply, gx, gy, gz, h, gyro, gscale, n);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd6_solve_all(void* stream, const float* X, const float* Q, const float* XP,
                    const float* QP, const float* IB, const float* HE, const int* JI,
                    const float* JA, const float* JB, float* LAMJ, float h, int first, int nj,
                    float floor_y, float mu_s, float* LAMF, int n, const int* PTI,
                    const float* PTN, const float* PTOI, const float* PTOR, float* LAMP,
                    const int* PTC, int max_points, float* DX, float* DR) {
  if (nj < 0 || n < 0 || max_points < 0) return (int)hipErrorInvalidValue;
  const long long total = (long long)floor_lane0(nj) + 8ll * n + max_points;
  if (total <= 0) return 0;
  if (total >= (1ll << 31)) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k6_solve_all, dim3(nblk(total)), dim3(kB), 0, (hipStream_t)stream, CF4(X),
                     CF4(Q), CF4(XP), CF4(QP), CF4(IB), CF4(HE), (const int2*)JI, CF4(JA),
                     CF4(JB), LAMJ, h, first, nj, floor_y, mu_s, LAMF, n, (const int4*)PTI,
                     CF4(PTN), CF4(PTOI), CF4(PTOR), LAMP, PTC, max_points, F4(DX), F4(DR));
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd6_apply_finalize(void* stream, float* X, float* Q, float* DX, float* DR, float relax,
                         int finalize, const float* XP, const float* QP, float* V, float* W,
                         float h, float lin_damp, float ang_damp, int n) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k6_apply_finalize, dim3(nblk(n)), dim3(kB), 0, (hipStream_t)stream, F4(X),
                     F4(Q), F4(DX), F4(DR), relax, finalize, CF4(XP), CF4(QP), F4(V), F4(W), h,
                     lin_damp, ang_damp, n);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd6_velocity_all(void* stream, const float* X, const float* Q, const float* IB,
                       const float* HE, float mu_d, float restitution, const float* LAMF,
                       float h, const float* V, const float* W, int n, const int* PTI,
                       const float* PTN, const float* PTOI, const float* PTOR, const float* LAMP,
                       const int* PTC, int max_points, float* DV, float* DW) {
  if (n < 0 || max_points < 0) return (int)hipErrorInvalidValue;
  const long long total = 8ll * n + max_points;
  if (total <= 0) return 0;
  if (total >= (1ll << 31)) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k6_velocity_all, dim3(nblk(total)), dim3(kB), 0, (hipStream_t)stream,
                     CF4(X), CF4(Q), CF4(IB), CF4(HE), mu_d, restitution, LAMF, h, CF4(V), CF4(W),
                     n, (const int4*)PTI, CF4(PTN), CF4(PTOI), CF4(PTOR), LAMP, PTC, max_points,
                     F4(DV), F4(DW));
  XPBD_LAUNCH_CHECK();
  return 0;
}

static int check_sub_desc(const xpbd_rigid_desc* desc, int sub) {
  if (!desc || !desc->aux || !desc->pt_count || desc->n < 0 || desc->cap < 0 || desc->nj < 0 ||
      sub < 0)
    return (int)hipErrorInvalidValue;
  const long long lanes = (long long)floor_lane0(desc->nj) + 8ll * desc->n + kB + desc->cap;
  if (lanes >= (1ll << 31) || 8ll * desc->n + kB + 8ll * desc->cap >= (1ll << 31))
    return (int)hipErrorInvalidValue;
  return 0;
}

int xpbd6_substep_contacts(void* stream, const xpbd_rigid_desc* desc, int sub) {
  int rc = check_sub_desc(desc, sub);
  if (rc) return rc;
  const xpbd_rigid_desc& d = *desc;
  if (d.n == 0) return 0;
  const long long lanes = (long long)block_lane0(floor_lane0(d.nj) + 8 * d.n) + d.cap;
  hipLaunchKernelGGL(k6_sub_contacts, dim3(nblk(lanes)), dim3(kB), 0, (hipStream_t)stream, d,
                     sub_bufs(d, sub));
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd6_substep_velocity(void* stream, const xpbd_rigid_desc* desc, int sub) {
  int rc = check_sub_desc(desc, sub);
  if (rc) return rc;
  const xpbd_rigid_desc& d = *desc;
  if (d.n == 0) return 0;
  hipLaunchKernelGGL(k6_sub_velocity, dim3(nblk((long long)block_lane0(8 * d.n) + 8ll * d.cap)),
                     dim3(kB), 0,
                     (hipStream_t)stream, d, sub_bufs(d, sub));
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd6_substep_end_report(void* stream, const xpbd_rigid_desc* desc, int* ring,
                             int ring_len, int* frames) {
  int rc = check_sub_desc(desc, 0);
  if (rc) return rc;
  const xpbd_rigid_desc& d = *desc;
  if (d.substeps < 1 || (ring && (ring_len < 1 || !frames))) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k6_sub_end, dim3(nblk(d.n > 0 ? d.n : 1)), dim3(kB), 0,
                     (hipStream_t)stream, d, sub_bufs(d, d.substeps - 1), ring, ring_len,
                     frames);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd6_substep_end(void* stream, const xpbd_rigid_desc* desc) {
  if (desc && desc->n == 0) return 0;
  return xpbd6_substep_end_report(stream, desc, nullptr, 0, nullptr);
}

int xpbd6_frame_block_capacity(void) { return kFBMax; }

int xpbd6_frame_block(void* stream, const xpbd_rigid_desc* desc, float* LO, float* HI, float dt,
                      int frames) {
  if (!desc || frames < 0) return (int)hipErrorInvalidValue;
  const xpbd_rigid_desc& d = *desc;
  if (d.n <= 0 || frames == 0) return 0;
  if (d.n > kFBMax || (long long)d.cap < (long long)d.n * (d.n - 1) / 2 || !LO || !HI)
    return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k6_frame_block, dim3(1), dim3(kFB), 0, (hipStream_t)stream, d, F4(LO),
                     F4(HI), dt, frames);
  XPBD_LAUNCH_CHECK();
  return 0;
}

int xpbd6_apply_velocity(void* stream, const float* X, float* V, float* W, float* DV, float* DW,
                         float max_lin, float max_ang, int do_cap, int n) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k6_apply_velocity, dim3(nblk(n)), dim3(kB), 0, (hipStream_t)stream, CF4(X),
                     F4(V), F4(W), F4(DV), F4(DW), max_lin, max_ang, do_cap, n);
  XPBD_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
