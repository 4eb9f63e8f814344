/*
 * ORACLE — test infrastructure only.  Serial CPU restatement of the reference
 * 6-DOF rigid-body kernels (src/xpbd3d/kernels_6dof.py), the broad-phase
 * AABB kernel (kernels_bvh.py:22-48) and the pair filters of the LBVH and
 * hash-grid queries (kernels_bvh.py:52-85, kernels_hashgrid.py:38-88),
 * driven like Solver6DOF._record / step (src/xpbd3d/solver_6dof.py:357-529).
 *
 * Semantics: Warp CPU launch model (serial tid loop, atomics in tid order),
 * FP32, -ffp-contract=off.  Warp's quaternion/matrix builtins are restated
 * from their mathematical definitions (quat xyzw):
 *   rotate(q,v)     = v*(2w^2-1) + qv*(2 dot(qv,v)) + cross(qv,v)*(2w)
 *   rotate_inv(q,v) = v*(2w^2-1) + qv*(2 dot(qv,v)) - cross(qv,v)*(2w)
 *   to_matrix(q)    = columns rotate(q, e_k)
 *   a*b             = Hamilton product, normalize(q) = q / |q|
 *   inverse(q)      = conjugate
 * The exact operation order inside Warp's builtins cannot be checked here
 * (warp-lang is not installed), so those are pinned to ulp level only by the
 * textbook identities in tests/test_oracle6.py.
 *
 * Broad phase: the candidate set is defined by brute force over all pairs
 * (the LBVH query returns every body whose AABB overlaps, inclusive bounds),
 * emitted in (i ascending, j ascending) order.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EPS6 1.0e-9f
#define MAXC 8

typedef struct { float x, y, z; } v3;
typedef struct { float x, y, z, w; } q4;
typedef struct { float m[3][3]; } m33;

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 dvs(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float len(v3 a) { return sqrtf(dot(a, a)); }
static inline v3 cross(v3 a, v3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline v3 ld(const float *p, int i) { return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static inline void st(float *p, int i, v3 a) { p[3 * i] = a.x; p[3 * i + 1] = a.y; p[3 * i + 2] = a.z; }
static inline q4 ldq(const float *p, int i) { q4 q = {p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]}; return q; }
static inline void stq(float *p, int i, q4 q) { p[4 * i] = q.x; p[4 * i + 1] = q.y; p[4 * i + 2] = q.z; p[4 * i + 3] = q.w; }
static inline void aadd(float *p, int i, v3 a) { p[3 * i] += a.x; p[3 * i + 1] += a.y; p[3 * i + 2] += a.z; }
static inline float fabsf_(float a) { return a < 0.0f ? -a : a; }
static inline float fminf_(float a, float b) { return a < b ? a : b; }
static inline float fmaxf_(float a, float b) { return a > b ? a : b; }

static inline v3 qrot(q4 q, v3 v) {
  v3 qv = mk(q.x, q.y, q.z);
  float c = 2.0f * q.w * q.w - 1.0f;
  float d = 2.0f * dot(qv, v);
  return add(add(mul(v, c), mul(qv, d)), mul(cross(qv, v), q.w * 2.0f));
}
static inline v3 qrot_inv(q4 q, v3 v) {
  v3 qv = mk(q.x, q.y, q.z);
  float c = 2.0f * q.w * q.w - 1.0f;
  float d = 2.0f * dot(qv, v);
  return sub(add(mul(v, c), mul(qv, d)), mul(cross(qv, v), q.w * 2.0f));
}
static inline q4 qmul(q4 a, q4 b) {
  q4 r;
  r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
  r.y = a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z;
  r.z = a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x;
  r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
  return r;
}
static inline q4 qadd(q4 a, q4 b) { q4 r = {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; return r; }
static inline q4 qscale(q4 a, float s) { q4 r = {a.x * s, a.y * s, a.z * s, a.w * s}; return r; }
static inline q4 qnormalize(q4 q) {
  float l = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  q4 r = {q.x / l, q.y / l, q.z / l, q.w / l};
  return r;
}
static inline q4 qconj(q4 q) { q4 r = {-q.x, -q.y, -q.z, q.w}; return r; }
static inline m33 qmat(q4 q) {
  m33 R;
  v3 c0 = qrot(q, mk(1.0f, 0.0f, 0.0f)), c1 = qrot(q, mk(0.0f, 1.0f, 0.0f)),
     c2 = qrot(q, mk(0.0f, 0.0f, 1.0f));
  R.m[0][0] = c0.x; R.m[1][0] = c0.y; R.m[2][0] = c0.z;
  R.m[0][1] = c1.x; R.m[1][1] = c1.y; R.m[2][1] = c1.z;
  R.m[0][2] = c2.x; R.m[1][2] = c2.y; R.m[2][2] = c2.z;
  return R;
}
static inline v3 col(m33 R, int k) { return mk(R.m[0][k], R.m[1][k], R.m[2][k]); }
static inline v3 mtv(m33 R, v3 v) { /* transpose(R) * v */
  return mk(R.m[0][0] * v.x + R.m[1][0] * v.y + R.m[2][0] * v.z,
            R.m[0][1] * v.x + R.m[1][1] * v.y + R.m[2][1] * v.z,
            R.m[0][2] * v.x + R.m[1][2] * v.y + R.m[2][2] * v.z);
}
static inline float comp(v3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

/* kernels_6dof.py:29-35 */
static inline v3 box_corner(v3 he, int i) {
  float sx = (i & 1) ? 1.0f : -1.0f, sy = (i & 2) ? 1.0f : -1.0f, sz = (i & 4) ? 1.0f : -1.0f;
  return mk(sx * he.x, sy * he.y, sz * he.z);
}
/* kernels_6dof.py:38-44 */
static inline v3 world_invI_mul(q4 q, v3 iI, v3 u) {
  v3 ub = qrot_inv(q, u);
  ub = mk(ub.x * iI.x, ub.y * iI.y, ub.z * iI.z);
  return qrot(q, ub);
}
/* kernels_6dof.py:47-53 */
static inline float gen_inv_mass(float im, q4 q, v3 iI, v3 r, v3 n) {
  v3 rn = cross(r, n);
  return im + dot(rn, world_invI_mul(q, iI, rn));
}
/* kernels_6dof.py:56-61 */
static inline q4 quat_integrate(q4 q, v3 w, float h) {
  q4 wq = {w.x, w.y, w.z, 0.0f};
  q = qadd(q, qscale(qmul(wq, q), 0.5f * h));
  return qnormalize(q);
}

typedef struct {
  int n;
  float *x, *q, *v, *w, *x_prev, *q_prev;
  const float *inv_mass, *inv_I, *he;
  const int *group;
  float *dx, *drot, *dv, *dw;
  int *dcount, *dvc;
  float *lam_floor; /* 8n */
} bodies_t;

/* kernels_6dof.py:67-102 */
void orc6_integrate(float *x, float *q, float *v, float *w, const float *inv_mass,
                    const float *inv_I, float gx, float gy, float gz, float h, int gyro,
                    float *x_prev, float *q_prev, int n) {
  v3 g = mk(gx, gy, gz);
  for (int i = 0; i < n; ++i) {
    st(x_prev, i, ld(x, i));
    stq(q_prev, i, ldq(q, i));
    if (inv_mass[i] == 0.0f) continue;
    st(v, i, add(ld(v, i), mul(g, h)));
    st(x, i, add(ld(x, i), mul(ld(v, i), h)));
    if (gyro != 0) {
      v3 wb = qrot_inv(ldq(q, i), ld(w, i));
      v3 ii = ld(inv_I, i);
      float ix = ii.x > 0.0f ? 1.0f / ii.x : 0.0f;
      float iy = ii.y > 0.0f ? 1.0f / ii.y : 0.0f;
      float iz = ii.z > 0.0f ? 1.0f / ii.z : 0.0f;
      v3 Iw = mk(ix * wb.x, iy * wb.y, iz * wb.z);
      v3 tau = neg(cross(wb, Iw));
      v3 dwb = mul(mk(ii.x * tau.x, ii.y * tau.y, ii.z * tau.z), h);
      st(w, i, add(ld(w, i), qrot(ldq(q, i), dwb)));
    }
    stq(q, i, quat_integrate(ldq(q, i), ld(w, i), h));
  }
}

/* kernels_6dof.py:108-130 */
static void accumulate_correction(float *dx, float *drot, int *dcount, int a, int b, float ima,
                                  float imb, q4 qa, q4 qb, v3 iIa, v3 iIb, v3 ra, v3 rb, v3 p) {
  if (ima > 0.0f) {
    aadd(dx, a, mul(p, ima));
    aadd(drot, a, world_invI_mul(qa, iIa, cross(ra, p)));
    dcount[a] += 1;
  }
  if (b >= 0 && imb > 0.0f) {
    aadd(dx, b, mul(neg(p), imb));
    aadd(drot, b, neg(world_invI_mul(qb, iIb, cross(rb, p))));
    dcount[b] += 1;
  }
}

/* kernels_6dof.py:158-211 */
void orc6_solve_floor(const float *x, const float *q, const float *x_prev, const float *q_prev,
                      const float *inv_mass, const float *inv_I, const float *he, float floor_y,
                      float mu_s, float *lam_floor, float h, float *dx, float *drot, int *dcount,
                      int n) {
  (void)h;
  v3 s = mk(0.0f, 1.0f, 0.0f);
  for (int i = 0; i < n; ++i) {
    if (inv_mass[i] == 0.0f) continue;
    for (int c = 0; c < 8; ++c) {
      v3 corner = box_corner(ld(he, i), c);
      v3 p1 = add(ld(x, i), qrot(ldq(q, i), corner));
      float d = floor_y - p1.y;
      int slot = i * 8 + c;
      if (d <= 0.0f) {
        lam_floor[slot] = 0.0f;
        continue;
      }
      v3 r = sub(p1, ld(x, i));
      float wgt = gen_inv_mass(inv_mass[i], ldq(q, i), ld(inv_I, i), r, s);
      if (wgt <= 0.0f) continue;
      float dlam = d / wgt;
      lam_floor[slot] = dlam;
      v3 p = mul(s, dlam);
      aadd(dx, i, mul(p, inv_mass[i]));
      aadd(drot, i, world_invI_mul(ldq(q, i), ld(inv_I, i), cross(r, p)));
      dcount[i] += 1;
      if (mu_s > 0.0f) {
        v3 p1b = add(ld(x_prev, i), qrot(ldq(q_prev, i), corner));
        v3 dp = sub(p1, p1b);
        v3 dpt = sub(dp, mul(s, dot(dp, s)));
        float lt = len(dpt);
        if (lt > EPS6 && lt < mu_s * d) {
          float w2 = gen_inv_mass(inv_mass[i], ldq(q, i), ld(inv_I, i), r, dvs(dpt, lt));
          if (w2 > 0.0f) {
            v3 pf = dvs(neg(dpt), w2);
            aadd(dx, i, mul(pf, inv_mass[i]));
            aadd(drot, i, world_invI_mul(ldq(q, i), ld(inv_I, i), cross(r, pf)));
          }
        }
      }
    }
  }
}

/* kernels_6dof.py:226-230 */
static inline float proj_radius(v3 L, m33 R, v3 e) {
  return fabsf_(e.x * dot(L, col(R, 0))) + fabsf_(e.y * dot(L, col(R, 1))) +
         fabsf_(e.z * dot(L, col(R, 2)));
}

/* kernels_6dof.py:233-382.  poly scratch: 16 vec3 per pair. */
void orc6_box_manifold(const float *x, const float *q, const float *he, const int *pair_a,
                       const int *pair_b, int pair_count, float margin, float *poly,
                       int *m_count, int *m_inc, int *m_ref, float *m_normal, float *m_off_inc,
                       float *m_off_ref, int cap) {
  for (int pid = 0; pid < cap; ++pid) {
    m_count[pid] = 0;
    if (pid >= pair_count) continue;
    int a = pair_a[pid], b = pair_b[pid];
    v3 cA = ld(x, a), cB = ld(x, b);
    v3 eA = ld(he, a), eB = ld(he, b);
    m33 RA = qmat(ldq(q, a)), RB = qmat(ldq(q, b));
    v3 cBA = sub(cA, cB);
    float best_overlap = 1.0e30f;
    v3 best_axis = mk(0.0f, 1.0f, 0.0f);
    int ref_is_a = 1, best_k = 0, sep = 0;
    for (int k = 0; k < 3 && !sep; ++k) {
      v3 L = col(RA, k);
      float ov = comp(eA, k) + proj_radius(L, RB, eB) - fabsf_(dot(L, cBA));
      if (ov < 0.0f) { sep = 1; break; }
      if (ov < best_overlap) { best_overlap = ov; best_axis = L; ref_is_a = 1; best_k = k; }
    }
    for (int k = 0; k < 3 && !sep; ++k) {
      v3 L = col(RB, k);
      float ov = comp(eB, k) + proj_radius(L, RA, eA) - fabsf_(dot(L, cBA));
      if (ov < 0.0f) { sep = 1; break; }
      if (ov < best_overlap) { best_overlap = ov; best_axis = L; ref_is_a = 0; best_k = k; }
    }
    if (sep) continue;
    v3 nBA = best_axis;
    if (dot(nBA, cBA) < 0.0f) nBA = neg(nBA);
    int ref_b, inc_b;
    v3 c_ref, e_ref, c_inc, e_inc, ref_face_n;
    m33 R_ref, R_inc;
    if (ref_is_a == 1) {
      ref_b = a; inc_b = b; c_ref = cA; R_ref = RA; e_ref = eA;
      c_inc = cB; R_inc = RB; e_inc = eB; ref_face_n = neg(nBA);
    } else {
      ref_b = b; inc_b = a; c_ref = cB; R_ref = RB; e_ref = eB;
      c_inc = cA; R_inc = RA; e_inc = eA; ref_face_n = nBA;
    }
    int ref_axis = best_k;
    v3 ref_face_c = add(c_ref, mul(ref_face_n, comp(e_ref, ref_axis)));
    int ax1 = (ref_axis + 1) % 3, ax2 = (ref_axis + 2) % 3;
    v3 u_ref = col(R_ref, ax1), v_ref = col(R_ref, ax2);
    float eu = comp(e_ref, ax1), ev = comp(e_ref, ax2);
    int inc_axis = 0;
    float inc_sign = 1.0f, best_dot = -1.0e30f;
    for (int k = 0; k < 3; ++k) {
      v3 c = col(R_inc, k);
      float dp = -dot(c, ref_face_n);
      if (dp > best_dot) { best_dot = dp; inc_axis = k; inc_sign = 1.0f; }
      if (-dp > best_dot) { best_dot = -dp; inc_axis = k; inc_sign = -1.0f; }
    }
    v3 inc_n = mul(col(R_inc, inc_axis), inc_sign);
    v3 inc_face_c = add(c_inc, mul(inc_n, comp(e_inc, inc_axis)));
    int iax1 = (inc_axis + 1) % 3, iax2 = (inc_axis + 2) % 3;
    v3 iu = col(R_inc, iax1), iv = col(R_inc, iax2);
    float ieu = comp(e_inc, iax1), iev = comp(e_inc, iax2);
    int cur = pid * 16, nxt = pid * 16 + 8;
    st(poly, cur + 0, add(add(inc_face_c, mul(iu, ieu)), mul(iv, iev)));
    st(poly, cur + 1, add(sub(inc_face_c, mul(iu, ieu)), mul(iv, iev)));
    st(poly, cur + 2, sub(sub(inc_face_c, mul(iu, ieu)), mul(iv, iev)));
    st(poly, cur + 3, sub(add(inc_face_c, mul(iu, ieu)), mul(iv, iev)));
    int plen = 4, dead = 0;
    for (int pi = 0; pi < 4; ++pi) {
      v3 pp, pn;
      if (pi == 0) { pp = add(ref_face_c, mul(u_ref, eu)); pn = u_ref; }
      else if (pi == 1) { pp = sub(ref_face_c, mul(u_ref, eu)); pn = neg(u_ref); }
      else if (pi == 2) { pp = add(ref_face_c, mul(v_ref, ev)); pn = v_ref; }
      else { pp = sub(ref_face_c, mul(v_ref, ev)); pn = neg(v_ref); }
      if (plen == 0) { dead = 1; break; }
      int olen = 0;
      for (int k = 0; k < plen; ++k) {
        v3 ap = ld(poly, cur + k), bp = ld(poly, cur + (k + 1) % plen);
        float da = dot(sub(ap, pp), pn), db = dot(sub(bp, pp), pn);
        if (da <= 0.0f) {
          if (olen < 8) { st(poly, nxt + olen, ap); olen += 1; }
          if (db > 0.0f) {
            float t = da / (da - db);
            if (olen < 8) { st(poly, nxt + olen, add(ap, mul(sub(bp, ap), t))); olen += 1; }
          }
        } else if (db <= 0.0f) {
          float t = da / (da - db);
          if (olen < 8) { st(poly, nxt + olen, add(ap, mul(sub(bp, ap), t))); olen += 1; }
        }
      }
      int tmp = cur; cur = nxt; nxt = tmp;
      plen = olen;
    }
    if (dead || plen == 0) continue;
    int nc = 0;
    for (int k = 0; k < plen; ++k) {
      if (nc >= 8) break;
      v3 p_inc = ld(poly, cur + k);
      float d_signed = dot(sub(p_inc, ref_face_c), ref_face_n);
      if (d_signed < margin) {
        v3 p_ref = sub(p_inc, mul(ref_face_n, d_signed));
        st(m_off_inc, pid * 8 + nc, mtv(R_inc, sub(p_inc, c_inc)));
        st(m_off_ref, pid * 8 + nc, mtv(R_ref, sub(p_ref, c_ref)));
        nc += 1;
      }
    }
    m_count[pid] = nc;
    m_inc[pid] = inc_b;
    m_ref[pid] = ref_b;
    st(m_normal, pid, ref_face_n);
  }
}

/* kernels_6dof.py:385-430 */
void orc6_solve_box(const float *x, const float *q, const float *inv_mass, const float *inv_I,
                    const int *m_count, const int *m_inc, const int *m_ref,
                    const float *m_normal, const float *m_off_inc, const float *m_off_ref,
                    float *lam_pair, float *dx, float *drot, int *dcount, int cap) {
  for (int pid = 0; pid < cap; ++pid) {
    int nc = m_count[pid];
    if (nc == 0) continue;
    int inc = m_inc[pid], ref = m_ref[pid];
    v3 s = ld(m_normal, pid);
    for (int c = 0; c < nc; ++c) {
      int slot = pid * 8 + c;
      v3 p_inc = add(ld(x, inc), qrot(ldq(q, inc), ld(m_off_inc, slot)));
      v3 p_ref = add(ld(x, ref), qrot(ldq(q, ref), ld(m_off_ref, slot)));
      float C = dot(sub(p_inc, p_ref), s);
      if (C >= 0.0f) { lam_pair[slot] = 0.0f; continue; }
      float d = -C;
      v3 r_inc = sub(p_inc, ld(x, inc)), r_ref = sub(p_ref, ld(x, ref));
      float w_inc = gen_inv_mass(inv_mass[inc], ldq(q, inc), ld(inv_I, inc), r_inc, s);
      float w_ref = gen_inv_mass(inv_mass[ref], ldq(q, ref), ld(inv_I, ref), r_ref, s);
      float wsum = w_inc + w_ref;
      if (wsum <= 0.0f) continue;
      float dlam = d / wsum;
      lam_pair[slot] = lam_pair[slot] + dlam;
      v3 p = mul(s, dlam);
      accumulate_correction(dx, drot, dcount, inc, ref, inv_mass[inc], inv_mass[ref],
                            ldq(q, inc), ldq(q, ref), ld(inv_I, inc), ld(inv_I, ref), r_inc,
                            r_ref, p);
    }
  }
}

/* kernels_6dof.py:440-483 */
void orc6_solve_joints(const float *x, const float *q, const float *inv_mass, const float *inv_I,
                       const int *j_a, const int *j_b, const float *j_anchor_a,
                       const float *j_anchor_b, const float *j_rest, const float *j_compliance,
                       float *lam_joint, float h, float *dx, float *drot, int *dcount, int nj) {
  for (int k = 0; k < nj; ++k) {
    int a = j_a[k], b = j_b[k];
    v3 pa = add(ld(x, a), qrot(ldq(q, a), ld(j_anchor_a, k)));
    v3 pb = add(ld(x, b), qrot(ldq(q, b), ld(j_anchor_b, k)));
    v3 d = sub(pa, pb);
    float dist = len(d);
    if (dist < EPS6) continue;
    v3 nn = dvs(d, dist);
    float C = dist - j_rest[k];
    v3 ra = sub(pa, ld(x, a)), rb = sub(pb, ld(x, b));
    float wa = gen_inv_mass(inv_mass[a], ldq(q, a), ld(inv_I, a), ra, nn);
    float wb = gen_inv_mass(inv_mass[b], ldq(q, b), ld(inv_I, b), rb, nn);
    float wsum = wa + wb;
    if (wsum <= 0.0f) continue;
    float alpha = j_compliance[k] / (h * h);
    float dlam = (-C - alpha * lam_joint[k]) / (wsum + alpha);
    lam_joint[k] = lam_joint[k] + dlam;
    v3 p = mul(nn, dlam);
    accumulate_correction(dx, drot, dcount, a, b, inv_mass[a], inv_mass[b], ldq(q, a),
                          ldq(q, b), ld(inv_I, a), ld(inv_I, b), ra, rb, p);
  }
}

/* kernels_6dof.py:486-506 */
void orc6_apply_jacobi(float *x, float *q, const float *inv_mass, float relax, float *dx,
                       float *drot, int *dcount, int n) {
  for (int i = 0; i < n; ++i) {
    int c = dcount[i];
    if (c > 0 && inv_mass[i] > 0.0f) {
      float inv_n = relax / (float)c;
      st(x, i, add(ld(x, i), mul(ld(dx, i), inv_n)));
      v3 dwv = mul(ld(drot, i), inv_n);
      q4 wq = {dwv.x, dwv.y, dwv.z, 0.0f};
      q4 qi = ldq(q, i);
      stq(q, i, qnormalize(qadd(qi, qscale(qmul(wq, qi), 0.5f))));
    }
    st(dx, i, mk(0.0f, 0.0f, 0.0f));
    st(drot, i, mk(0.0f, 0.0f, 0.0f));
    dcount[i] = 0;
  }
}

/* kernels_6dof.py:512-535 */
void orc6_finalize_velocity(const float *x, const float *q, const float *x_prev,
                            const float *q_prev, const float *inv_mass, float h, float lin_damp,
                            float ang_damp, float *v, float *w, int n) {
  for (int i = 0; i < n; ++i) {
    if (inv_mass[i] == 0.0f) {
      st(v, i, mk(0.0f, 0.0f, 0.0f));
      st(w, i, mk(0.0f, 0.0f, 0.0f));
      continue;
    }
    st(v, i, mul(dvs(sub(ld(x, i), ld(x_prev, i)), h), 1.0f - lin_damp));
    q4 dq = qmul(ldq(q, i), qconj(ldq(q_prev, i)));
    v3 om = mul(mk(dq.x, dq.y, dq.z), 2.0f / h);
    if (dq.w < 0.0f) om = neg(om);
    st(w, i, mul(om, 1.0f - ang_damp));
  }
}

/* kernels_6dof.py:541-590 */
void orc6_velocity_floor(const float *x, const float *q, const float *inv_mass,
                         const float *inv_I, const float *he, float floor_y, float mu_d,
                         float restitution, const float *lam_floor, float h, const float *v,
                         const float *w, float *dv, float *dw, int *dvc, int n) {
  (void)x; (void)floor_y;
  v3 s = mk(0.0f, 1.0f, 0.0f);
  for (int i = 0; i < n; ++i) {
    if (inv_mass[i] == 0.0f) continue;
    for (int c = 0; c < 8; ++c) {
      int slot = i * 8 + c;
      float lam_n = lam_floor[slot];
      if (lam_n == 0.0f) continue;
      v3 corner = box_corner(ld(he, i), c);
      v3 r = qrot(ldq(q, i), corner);
      v3 vc = add(ld(v, i), cross(ld(w, i), r));
      float vn = dot(vc, s);
      v3 vt = sub(vc, mul(s, vn));
      float ltan = len(vt);
      v3 impulse = mk(0.0f, 0.0f, 0.0f);
      if (mu_d > 0.0f && ltan > EPS6) {
        float fn = lam_n / (h * h);
        float dvm = fminf_(h * mu_d * fabsf_(fn), ltan);
        impulse = sub(impulse, mul(dvs(vt, ltan), dvm));
      }
      if (vn < 0.0f) impulse = add(impulse, mul(s, -vn + fmaxf_(-restitution * vn, 0.0f)));
      if (len(impulse) > EPS6) {
        float wg = gen_inv_mass(inv_mass[i], ldq(q, i), ld(inv_I, i), r, dvs(impulse, len(impulse)));
        if (wg > 0.0f) {
          v3 p = dvs(impulse, wg);
          aadd(dv, i, mul(p, inv_mass[i]));
          aadd(dw, i, world_invI_mul(ldq(q, i), ld(inv_I, i), cross(r, p)));
          dvc[i] += 1;
        }
      }
    }
  }
}

/* kernels_6dof.py:593-659 */
void orc6_velocity_box(const float *x, const float *q, const float *inv_mass, const float *inv_I,
                       const int *m_count, const int *m_inc, const int *m_ref,
                       const float *m_normal, const float *m_off_inc, const float *m_off_ref,
                       float mu_d, float restitution, const float *lam_pair, float h,
                       const float *v, const float *w, float *dv, float *dw, int *dvc, int cap) {
  for (int pid = 0; pid < cap; ++pid) {
    int nc = m_count[pid];
    if (nc == 0) continue;
    int ci = m_inc[pid], cj = m_ref[pid];
    v3 s = ld(m_normal, pid);
    for (int c = 0; c < nc; ++c) {
      int slot = pid * 8 + c;
      float lam_n = lam_pair[slot];
      if (lam_n == 0.0f) continue;
      v3 p_inc = add(ld(x, ci), qrot(ldq(q, ci), ld(m_off_inc, slot)));
      v3 p_ref = add(ld(x, cj), qrot(ldq(q, cj), ld(m_off_ref, slot)));
      v3 ri = sub(p_inc, ld(x, ci)), rj = sub(p_ref, ld(x, cj));
      v3 vci = add(ld(v, ci), cross(ld(w, ci), ri));
      v3 vcj = add(ld(v, cj), cross(ld(w, cj), rj));
      v3 vrel = sub(vci, vcj);
      float vn = dot(vrel, s);
      v3 vt = sub(vrel, mul(s, vn));
      float ltan = len(vt);
      v3 impulse = mk(0.0f, 0.0f, 0.0f);
      if (mu_d > 0.0f && ltan > EPS6) {
        float fn = lam_n / (h * h);
        float dvm = fminf_(h * mu_d * fabsf_(fn), ltan);
        impulse = sub(impulse, mul(dvs(vt, ltan), dvm));
      }
      if (vn < 0.0f) impulse = add(impulse, mul(s, -vn + fmaxf_(-restitution * vn, 0.0f)));
      float li = len(impulse);
      if (li > EPS6) {
        v3 ndir = dvs(impulse, li);
        float wi = gen_inv_mass(inv_mass[ci], ldq(q, ci), ld(inv_I, ci), ri, ndir);
        float wj = gen_inv_mass(inv_mass[cj], ldq(q, cj), ld(inv_I, cj), rj, ndir);
        float wsum = wi + wj;
        if (wsum > 0.0f) {
          v3 p = dvs(impulse, wsum);
          if (inv_mass[ci] > 0.0f) {
            aadd(dv, ci, mul(p, inv_mass[ci]));
            aadd(dw, ci, world_invI_mul(ldq(q, ci), ld(inv_I, ci), cross(ri, p)));
            dvc[ci] += 1;
          }
          if (inv_mass[cj] > 0.0f) {
            aadd(dv, cj, mul(neg(p), inv_mass[cj]));
            aadd(dw, cj, neg(world_invI_mul(ldq(q, cj), ld(inv_I, cj), cross(rj, p))));
            dvc[cj] += 1;
          }
        }
      }
    }
  }
}

/* kernels_6dof.py:662-679 */
void orc6_apply_velocity(const float *inv_mass, float *dv, float *dw, int *dvc, float *v,
                         float *w, int n) {
  for (int i = 0; i < n; ++i) {
    int c = dvc[i];
    if (c > 0 && inv_mass[i] > 0.0f) {
      float inv_n = 1.0f / (float)c;
      st(v, i, add(ld(v, i), mul(ld(dv, i), inv_n)));
      st(w, i, add(ld(w, i), mul(ld(dw, i), inv_n)));
    }
    st(dv, i, mk(0.0f, 0.0f, 0.0f));
    st(dw, i, mk(0.0f, 0.0f, 0.0f));
    dvc[i] = 0;
  }
}

/* kernels_6dof.py:682-691 */
void orc6_cap_velocity(float *v, float *w, float max_lin, float max_ang, int n) {
  for (int i = 0; i < n; ++i) {
    float sp = len(ld(v, i));
    if (max_lin > 0.0f && sp > max_lin) st(v, i, mul(ld(v, i), max_lin / sp));
    float sa = len(ld(w, i));
    if (max_ang > 0.0f && sa > max_ang) st(w, i, mul(ld(w, i), max_ang / sa));
  }
}

/* kernels_bvh.py:21-48 */
void orc6_body_aabb(const float *x, const float *q, const float *v, const float *w,
                    const float *he, float man_margin, float dt, float *lowers, float *uppers,
                    int n) {
  for (int i = 0; i < n; ++i) {
    m33 R = qmat(ldq(q, i));
    v3 e = ld(he, i);
    float ax = fabsf_(R.m[0][0]) * e.x + fabsf_(R.m[0][1]) * e.y + fabsf_(R.m[0][2]) * e.z;
    float ay = fabsf_(R.m[1][0]) * e.x + fabsf_(R.m[1][1]) * e.y + fabsf_(R.m[1][2]) * e.z;
    float az = fabsf_(R.m[2][0]) * e.x + fabsf_(R.m[2][1]) * e.y + fabsf_(R.m[2][2]) * e.z;
    v3 half = mk(ax, ay, az);
    float rr = len(e);
    float m = man_margin + dt * (len(ld(v, i)) + len(ld(w, i)) * rr);
    v3 pad = mk(m, m, m);
    st(lowers, i, sub(sub(ld(x, i), half), pad));
    st(uppers, i, add(add(ld(x, i), half), pad));
  }
}

/* kernels_bvh.py:52-85 — the LBVH query reports every body whose box
 * overlaps (closed intervals); the pair filter keeps j > i, drops
 * static-static and same-group (> 0) pairs.  Emitted in (i, j) order,
 * clamped to cap; returns the full count. */
int orc6_pairs_aabb(const float *lowers, const float *uppers, const float *inv_mass,
                    const int *group, int n, int cap, int *pair_a, int *pair_b) {
  int count = 0;
  for (int i = 0; i < n; ++i) {
    v3 lo = ld(lowers, i), hi = ld(uppers, i);
    for (int j = i + 1; j < n; ++j) {
      v3 lj = ld(lowers, j), hj = ld(uppers, j);
      if (lo.x > hj.x || lj.x > hi.x || lo.y > hj.y || lj.y > hi.y || lo.z > hj.z || lj.z > hi.z)
        continue;
      if (inv_mass[i] == 0.0f && inv_mass[j] == 0.0f) continue;
      if (group[i] == group[j] && group[i] > 0) continue;
      if (count < cap) { pair_a[count] = i; pair_b[count] = j; }
      count += 1;
    }
  }
  return count;
}

/* kernels_hashgrid.py:27-34 */
static inline v3 aabb_half(q4 q, v3 e) {
  m33 R = qmat(q);
  return mk(fabsf_(R.m[0][0]) * e.x + fabsf_(R.m[0][1]) * e.y + fabsf_(R.m[0][2]) * e.z,
            fabsf_(R.m[1][0]) * e.x + fabsf_(R.m[1][1]) * e.y + fabsf_(R.m[1][2]) * e.z,
            fabsf_(R.m[2][0]) * e.x + fabsf_(R.m[2][1]) * e.y + fabsf_(R.m[2][2]) * e.z);
}

/* kernels_hashgrid.py:37-88 — the hash-grid candidate filter.  Body i
 * queries a sphere of radius qr = min(rs_i + rs_max + margin + m_i, cell)
 * around x_i; the grid walk visits the cells floor((x_i -/+ qr)/cell) per
 * axis, so a candidate j must have its cell floor(x_j/cell) inside that
 * range; then the exact per-axis AABB test with pad = margin + m_i + m_j. */
int orc6_pairs_hashgrid(const float *x, const float *q, const float *v, const float *w,
                        const float *he, const float *inv_mass, const int *group,
                        float man_margin, float dt, float rs_max, float cell, int n, int cap,
                        int *pair_a, int *pair_b) {
  int count = 0;
  for (int i = 0; i < n; ++i) {
    v3 ei = ld(he, i);
    float rsi = len(ei);
    float mi = fminf_(dt * (len(ld(v, i)) + len(ld(w, i)) * rsi), rsi);
    float qr = fminf_(rsi + rs_max + man_margin + mi, cell);
    v3 hi = aabb_half(ldq(q, i), ei);
    v3 xi = ld(x, i);
    int lo_c[3], hi_c[3];
    float xc[3] = {xi.x, xi.y, xi.z};
    for (int k = 0; k < 3; ++k) {
      lo_c[k] = (int)floorf((xc[k] - qr) / cell);
      hi_c[k] = (int)floorf((xc[k] + qr) / cell);
    }
    for (int j = i + 1; j < n; ++j) {
      if (inv_mass[i] == 0.0f && inv_mass[j] == 0.0f) continue;
      if (group[i] == group[j] && group[i] > 0) continue;
      v3 xj = ld(x, j);
      float xjc[3] = {xj.x, xj.y, xj.z};
      int inside = 1;
      for (int k = 0; k < 3; ++k) {
        int c = (int)floorf(xjc[k] / cell);
        if (c < lo_c[k] || c > hi_c[k]) inside = 0;
      }
      if (!inside) continue;
      v3 ej = ld(he, j);
      float rsj = len(ej);
      float mj = fminf_(dt * (len(ld(v, j)) + len(ld(w, j)) * rsj), rsj);
      v3 hj = aabb_half(ldq(q, j), ej);
      v3 dpos = sub(xi, xj);
      float pad = man_margin + mi + mj;
      float ox = fabsf_(dpos.x) - (hi.x + hj.x + pad);
      float oy = fabsf_(dpos.y) - (hi.y + hj.y + pad);
      float oz = fabsf_(dpos.z) - (hi.z + hj.z + pad);
      if (ox <= 0.0f && oy <= 0.0f && oz <= 0.0f) {
        if (count < cap) { pair_a[count] = i; pair_b[count] = j; }
        count += 1;
      }
    }
  }
  return count;
}
