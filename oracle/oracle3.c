/*
 * ORACLE — test infrastructure only.  Serial CPU restatement of the reference
 * 3-DOF XPBD particle kernels (src/xpbd3d/kernels.py) and of the greedy
 * constraint coloring (src/xpbd3d/coloring.py).  Nothing in the product links
 * or calls this file; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg load it, as the checker.
 *
 * Semantics followed:
 *   - Warp CPU launch model: every kernel is a serial loop over tid = 0..dim-1,
 *     so the Jacobi atomic adds land in tid order.
 *   - Reference data layout: AoS vec3 (12 B) / float / int32 arrays.
 *   - FP32 throughout, operation order as written in kernels.py; built with
 *     -ffp-contract=off so no FMA contraction happens.
 *   - vec3 builtins restated from their definitions: dot(a,b) is
 *     (a0*b0 + a1*b1) + a2*b2, length(a) = sqrtf(dot(a,a)), vec/scalar divides
 *     each component, vec*scalar multiplies each component.
 *
 * Parity pins (see tests/test_oracle3.py):
 *   integrate/finalize/cap : closed-form discrete free fall, static bodies,
 *                            speed clamp direction/magnitude
 *   DISTANCE/ATTACH        : single-row closed form dlam = -C/(wa+wb+alpha~);
 *                            steady-state stretch m*g*alpha (test_solver.py:50)
 *   FLOOR/CONTACT+friction : Coulomb cone cases (inside -> full cancel,
 *                            outside -> mu*d), push-out depth
 *   Jacobi vs GS           : identical on a single color class with relax=1
 *   coloring               : validity + brute-force greedy certificate
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EPS 1.0e-9f

enum { DISTANCE = 0, ATTACH = 1, FLOOR_ = 2, CONTACT = 3 };

typedef struct { float x, y, z; } v3;

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 dvs(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float len(v3 a) { return sqrtf(dot(a, a)); }
static inline v3 ld(const float *p, int i) { return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static inline void st(float *p, int i, v3 a) { p[3 * i] = a.x; p[3 * i + 1] = a.y; p[3 * i + 2] = a.z; }

/* kernels.py:49-73 */
void orc_integrate(float *x, float *v, const float *inv_mass, float gx, float gy,
                   float gz, float dt, float *x_prev, int n) {
  v3 g = mk(gx, gy, gz);
  for (int i = 0; i < n; ++i) {
    st(x_prev, i, ld(x, i));
    if (inv_mass[i] == 0.0f) continue;
    st(v, i, add(ld(v, i), mul(g, dt)));
    st(x, i, add(ld(x, i), mul(ld(v, i), dt)));
  }
}

/* kernels.py:80-98 */
static v3 friction_delta(v3 dp_rel, v3 n_hat, float penetration, float mu) {
  v3 dp_t = sub(dp_rel, mul(n_hat, dot(dp_rel, n_hat)));
  float lt = len(dp_t);
  if (lt < EPS) return mk(0.0f, 0.0f, 0.0f);
  float scale = 1.0f;
  float cone = mu * penetration;
  if (lt >= cone) scale = cone / lt;
  return mul(dp_t, -scale);
}

/* kernels.py:105-208 — one color slice, in place. */
void orc_solve_color(float *x, const float *x_prev, const float *inv_mass,
                     const int *c_type, const int *c_a, const int *c_b,
                     const float *c_rest, const float *c_compliance,
                     const float *c_anchor, const float *c_mu, float *c_lambda,
                     const int *c_active, const int *color_order, int color_start,
                     int count, float dt) {
  for (int tid = 0; tid < count; ++tid) {
    int j = color_order[color_start + tid];
    if (c_active[j] == 0) continue;
    int t = c_type[j];
    float inv_dt2 = 1.0f / (dt * dt);
    if (t == DISTANCE || t == ATTACH) {
      int a = c_a[j];
      float wa = inv_mass[a], wb;
      v3 pb;
      if (t == ATTACH) {
        pb = ld(c_anchor, j);
        wb = 0.0f;
      } else {
        int b = c_b[j];
        pb = ld(x, b);
        wb = inv_mass[b];
      }
      float wsum = wa + wb;
      if (wsum == 0.0f) continue;
      v3 d = sub(ld(x, a), pb);
      float L = len(d);
      if (L < EPS) continue;
      v3 n_hat = dvs(d, L);
      float C = L - c_rest[j];
      float alpha = c_compliance[j] * inv_dt2;
      float dlam = (-C - alpha * c_lambda[j]) / (wsum + alpha);
      c_lambda[j] = c_lambda[j] + dlam;
      st(x, a, add(ld(x, a), mul(n_hat, wa * dlam)));
      if (t == DISTANCE) {
        int b = c_b[j];
        st(x, b, sub(ld(x, b), mul(n_hat, wb * dlam)));
      }
      continue;
    }
    if (t == FLOOR_) {
      int a = c_a[j];
      float wa = inv_mass[a];
      if (wa == 0.0f) continue;
      float floor_y = c_anchor[3 * j + 1];
      float C = ld(x, a).y - floor_y;
      if (C >= 0.0f) continue;
      float dlam = -C / wa;
      c_lambda[j] = c_lambda[j] + dlam;
      st(x, a, add(ld(x, a), mk(0.0f, wa * dlam, 0.0f)));
      float mu = c_mu[j];
      if (mu > 0.0f) {
        v3 dp_rel = sub(ld(x, a), ld(x_prev, a));
        v3 corr = friction_delta(dp_rel, mk(0.0f, 1.0f, 0.0f), -C, mu);
        st(x, a, add(ld(x, a), corr));
      }
      continue;
    }
    if (t == CONTACT) {
      int a = c_a[j], b = c_b[j];
      float wa = inv_mass[a], wb = inv_mass[b];
      float wsum = wa + wb;
      if (wsum == 0.0f) continue;
      v3 d = sub(ld(x, a), ld(x, b));
      float L = len(d);
      if (L < EPS) continue;
      v3 n_hat = dvs(d, L);
      float C = L - c_rest[j];
      if (C >= 0.0f) continue;
      float dlam = -C / wsum;
      c_lambda[j] = c_lambda[j] + dlam;
      st(x, a, add(ld(x, a), mul(n_hat, wa * dlam)));
      st(x, b, sub(ld(x, b), mul(n_hat, wb * dlam)));
      float mu = c_mu[j];
      if (mu > 0.0f) {
        v3 dp_rel = sub(sub(ld(x, a), ld(x_prev, a)), sub(ld(x, b), ld(x_prev, b)));
        v3 corr = friction_delta(dp_rel, n_hat, -C, mu);
        st(x, a, add(ld(x, a), mul(corr, wa / wsum)));
        st(x, b, sub(ld(x, b), mul(corr, wb / wsum)));
      }
      continue;
    }
  }
}

static inline void atomic_add3(float *dx, int i, v3 c) {
  dx[3 * i] += c.x; dx[3 * i + 1] += c.y; dx[3 * i + 2] += c.z;
}

/* kernels.py:219-322 — serial tid order stands in for the atomics. */
void orc_solve_jacobi(const float *x, const float *x_prev, const float *inv_mass,
                      const int *c_type, const int *c_a, const int *c_b,
                      const float *c_rest, const float *c_compliance,
                      const float *c_anchor, const float *c_mu, float *c_lambda,
                      const int *c_active, float dt, float *dx, int *dn, int n_c) {
  for (int j = 0; j < n_c; ++j) {
    if (c_active[j] == 0) continue;
    int t = c_type[j];
    float inv_dt2 = 1.0f / (dt * dt);
    if (t == DISTANCE || t == ATTACH) {
      int a = c_a[j];
      float wa = inv_mass[a], wb;
      v3 pb;
      if (t == ATTACH) {
        pb = ld(c_anchor, j);
        wb = 0.0f;
      } else {
        int b = c_b[j];
        pb = ld(x, b);
        wb = inv_mass[b];
      }
      float wsum = wa + wb;
      if (wsum == 0.0f) continue;
      v3 d = sub(ld(x, a), pb);
      float L = len(d);
      if (L < EPS) continue;
      v3 n_hat = dvs(d, L);
      float C = L - c_rest[j];
      float alpha = c_compliance[j] * inv_dt2;
      float dlam = (-C - alpha * c_lambda[j]) / (wsum + alpha);
      c_lambda[j] = c_lambda[j] + dlam;
      atomic_add3(dx, a, mul(n_hat, wa * dlam));
      dn[a] += 1;
      if (t == DISTANCE) {
        atomic_add3(dx, c_b[j], mul(neg(n_hat), wb * dlam));
        dn[c_b[j]] += 1;
      }
      continue;
    }
    if (t == FLOOR_) {
      int a = c_a[j];
      float wa = inv_mass[a];
      if (wa == 0.0f) continue;
      float floor_y = c_anchor[3 * j + 1];
      float C = ld(x, a).y - floor_y;
      if (C >= 0.0f) continue;
      float dlam = -C / wa;
      c_lambda[j] = c_lambda[j] + dlam;
      v3 corr = mk(0.0f, wa * dlam, 0.0f);
      float mu = c_mu[j];
      if (mu > 0.0f) {
        v3 dp_rel = sub(add(ld(x, a), corr), ld(x_prev, a));
        corr = add(corr, friction_delta(dp_rel, mk(0.0f, 1.0f, 0.0f), -C, mu));
      }
      atomic_add3(dx, a, corr);
      dn[a] += 1;
      continue;
    }
    if (t == CONTACT) {
      int a = c_a[j], b = c_b[j];
      float wa = inv_mass[a], wb = inv_mass[b];
      float wsum = wa + wb;
      if (wsum == 0.0f) continue;
      v3 d = sub(ld(x, a), ld(x, b));
      float L = len(d);
      if (L < EPS) continue;
      v3 n_hat = dvs(d, L);
      float C = L - c_rest[j];
      if (C >= 0.0f) continue;
      float dlam = -C / wsum;
      c_lambda[j] = c_lambda[j] + dlam;
      v3 ca = mul(n_hat, wa * dlam);
      v3 cb = mul(neg(n_hat), wb * dlam);
      float mu = c_mu[j];
      if (mu > 0.0f) {
        v3 dp_rel = sub(sub(add(ld(x, a), ca), ld(x_prev, a)),
                        sub(add(ld(x, b), cb), ld(x_prev, b)));
        v3 fcorr = friction_delta(dp_rel, n_hat, -C, mu);
        ca = add(ca, mul(fcorr, wa / wsum));
        cb = sub(cb, mul(fcorr, wb / wsum));
      }
      atomic_add3(dx, a, ca);
      dn[a] += 1;
      atomic_add3(dx, b, cb);
      dn[b] += 1;
      continue;
    }
  }
}

/* kernels.py:326-339 */
void orc_apply_jacobi(float *x, float *dx, int *dn, const float *inv_mass,
                      float relax, int n) {
  for (int i = 0; i < n; ++i) {
    int c = dn[i];
    if (c > 0 && inv_mass[i] > 0.0f) st(x, i, add(ld(x, i), mul(ld(dx, i), relax / (float)c)));
    st(dx, i, mk(0.0f, 0.0f, 0.0f));
    dn[i] = 0;
  }
}

/* kernels.py:346-358 */
void orc_finalize_velocity(const float *x, const float *x_prev,
                           const float *inv_mass, float dt, float damping,
                           float *v, int n) {
  for (int i = 0; i < n; ++i) {
    if (inv_mass[i] == 0.0f) {
      st(v, i, mk(0.0f, 0.0f, 0.0f));
      continue;
    }
    st(v, i, mul(dvs(sub(ld(x, i), ld(x_prev, i)), dt), 1.0f - damping));
  }
}

/* kernels.py:362-371 */
void orc_cap_velocity(float *v, float max_speed, int n) {
  for (int i = 0; i < n; ++i) {
    float speed = len(ld(v, i));
    if (speed > max_speed) st(v, i, mul(ld(v, i), max_speed / speed));
  }
}

/*
 * One frame of Solver.step (solver.py:412-442) without self-collision:
 * substeps x [integrate, lambda<-0, iterations x sweep, finalize, cap].
 * mode 0 = gs (needs color_order/color_starts), 1 = jacobi.
 * `h` is the substep length already rounded to float32 (solver.py:423).
 * Scratch dx (3n floats) and dn (n ints) must be zero on entry; they are zero
 * again on exit.
 */
void orc_step3(float *x, float *v, float *x_prev, const float *inv_mass, int n_b,
               const int *c_type, const int *c_a, const int *c_b,
               const float *c_rest, const float *c_compliance,
               const float *c_anchor, const float *c_mu, float *c_lambda,
               const int *c_active, int n_c, const int *color_order,
               const int *color_starts, int num_colors, float gx, float gy,
               float gz, float h, int substeps, int iterations, int mode,
               float damping, float jacobi_relax, float max_speed,
               int frames, float *dx, int *dn) {
  for (int f = 0; f < frames; ++f) {
    for (int s = 0; s < substeps; ++s) {
      orc_integrate(x, v, inv_mass, gx, gy, gz, h, x_prev, n_b);
      if (n_c > 0) {
        memset(c_lambda, 0, sizeof(float) * (size_t)n_c);
        for (int it = 0; it < iterations; ++it) {
          if (mode == 0) {
            for (int c = 0; c < num_colors; ++c) {
              int start = color_starts[c];
              int count = color_starts[c + 1] - start;
              if (count <= 0) continue;
              orc_solve_color(x, x_prev, inv_mass, c_type, c_a, c_b, c_rest,
                              c_compliance, c_anchor, c_mu, c_lambda, c_active,
                              color_order, start, count, h);
            }
          } else {
            orc_solve_jacobi(x, x_prev, inv_mass, c_type, c_a, c_b, c_rest,
                             c_compliance, c_anchor, c_mu, c_lambda, c_active, h,
                             dx, dn, n_c);
            orc_apply_jacobi(x, dx, dn, inv_mass, jacobi_relax, n_b);
          }
        }
      }
      orc_finalize_velocity(x, x_prev, inv_mass, h, damping, v, n_b);
      if (max_speed > 0.0f && isfinite(max_speed)) orc_cap_velocity(v, max_speed, n_b);
    }
  }
}

/* ---------------------------------------------------------------------------
 * coloring.py:26-71 — greedy Welsh-Powell coloring.
 * Order: stable sort of constraint ids by key -(deg[a] + (deg[b] if b>=0)).
 * The stable sort is a merge sort on (key, index) so ties keep index order.
 * Each constraint takes the smallest color absent from used[a] U used[b].
 * Per-body used sets are dynamic bitsets of 64-bit words.
 * ------------------------------------------------------------------------- */
typedef struct { int64_t key; int idx; } kv_t;

static void merge_sort_kv(kv_t *a, kv_t *tmp, int n) {
  for (int width = 1; width < n; width *= 2) {
    for (int lo = 0; lo < n; lo += 2 * width) {
      int mid = lo + width < n ? lo + width : n;
      int hi = lo + 2 * width < n ? lo + 2 * width : n;
      int i = lo, j = mid, k = lo;
      while (i < mid && j < hi) tmp[k++] = (a[j].key < a[i].key) ? a[j++] : a[i++];
      while (i < mid) tmp[k++] = a[i++];
      while (j < hi) tmp[k++] = a[j++];
    }
    memcpy(a, tmp, sizeof(kv_t) * (size_t)n);
  }
}

int orc_color_constraints(int n_bodies, const int *c_a, const int *c_b, int n_c,
                          int *color) {
  if (n_c == 0) return 0;
  int64_t *deg = (int64_t *)calloc((size_t)n_bodies + 1, sizeof(int64_t));
  for (int j = 0; j < n_c; ++j) {
    if (c_a[j] >= 0) deg[c_a[j]] += 1;
    if (c_b[j] >= 0) deg[c_b[j]] += 1;
  }
  kv_t *kv = (kv_t *)malloc(sizeof(kv_t) * (size_t)n_c);
  kv_t *tmp = (kv_t *)malloc(sizeof(kv_t) * (size_t)n_c);
  for (int j = 0; j < n_c; ++j) {
    int64_t s = deg[c_a[j]] + (c_b[j] >= 0 ? deg[c_b[j]] : 0);
    kv[j].key = -s;
    kv[j].idx = j;
  }
  merge_sort_kv(kv, tmp, n_c);
  int words = 1;
  uint64_t *used = (uint64_t *)calloc((size_t)n_bodies * (size_t)words, sizeof(uint64_t));
  int max_color = -1;
  for (int k = 0; k < n_c; ++k) {
    int j = kv[k].idx;
    int a = c_a[j], b = c_b[j];
    int c = 0;
    for (;; ++c) {
      int w = c >> 6;
      uint64_t bit = (uint64_t)1 << (c & 63);
      int taken = 0;
      if (w < words) {
        if (a >= 0 && (used[(size_t)a * words + w] & bit)) taken = 1;
        if (b >= 0 && (used[(size_t)b * words + w] & bit)) taken = 1;
      }
      if (!taken) break;
    }
    if ((c >> 6) >= words) { /* grow every body's bitset */
      int nw = (c >> 6) + 1;
      uint64_t *nu = (uint64_t *)calloc((size_t)n_bodies * (size_t)nw, sizeof(uint64_t));
      for (int i = 0; i < n_bodies; ++i)
        memcpy(nu + (size_t)i * nw, used + (size_t)i * words, sizeof(uint64_t) * (size_t)words);
      free(used);
      used = nu;
      words = nw;
    }
    color[j] = c;
    if (c > max_color) max_color = c;
    uint64_t bit = (uint64_t)1 << (c & 63);
    if (a >= 0) used[(size_t)a * words + (c >> 6)] |= bit;
    if (b >= 0) used[(size_t)b * words + (c >> 6)] |= bit;
  }
  free(used); free(kv); free(tmp); free(deg);
  return max_color + 1;
}
