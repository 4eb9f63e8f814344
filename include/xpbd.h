/*
 * xpbd.h — C-ABI of libxpbd_hip.so, the MI355X (gfx950) XPBD solver library.
 *
 * Every entry point takes plain device pointers, sizes and scalars, launches
 * asynchronously on the given HIP stream (NULL = the default stream) and
 * returns 0 on success or a hipError_t code.  The library never allocates or
 * frees caller buffers; plans and graphs live behind opaque handles with an
 * explicit create/destroy.  Host-only entry points (coloring) return -1 on
 * invalid input.
 *
 * Each declaration names the reference interface it replaces, as a
 * `wp.launch(kernel, ...)` site or a host routine of nanli9/XPBD
 * (paths relative to the reference repository).
 *
 * Device layouts (3-DOF particle solver):
 *   X   float4[N]  (x, y, z, inv_mass)     V  float4[N] (vx, vy, vz, 0)
 *   XP  float4[N]  x_prev (x, y, z, 0)     ACC float4[N] Jacobi sums (dx, count)
 *   R   int4[M]    rows: .x = body_a | type << 28, .y = body_b (-1 if none),
 *                  .z/.w = float bits of (rest, compliance) for DISTANCE/ATTACH,
 *                  (floor_y, mu) for FLOOR, (rest, mu) for CONTACT
 *   ANC float4[M]  ATTACH world anchor     LAM float[M] multipliers
 */
#ifndef XPBD_H_
#define XPBD_H_

#ifdef __cplusplus
extern "C" {
#endif

#define XPBD_TYPE_DISTANCE 0
#define XPBD_TYPE_ATTACH 1
#define XPBD_TYPE_FLOOR 2
#define XPBD_TYPE_CONTACT 3

#define XPBD_MODE_GS 0
#define XPBD_MODE_JACOBI 1
/* Jacobi with one fused launch per iteration; after a frame the positions
 * are in x and the multipliers in lam (iterations even) or lam_alt (odd). */
#define XPBD_MODE_JACOBI_TILE 2

/* Largest LDS accumulator window (particles) of the Jacobi projection. */
#define XPBD_MAX_WINDOW 8192
/* Largest particle tile of the fused Jacobi iteration (44 B of LDS per particle:
 * three 64-bit fixed-point sums, a count and the staged position). */
#define XPBD_MAX_TILE 1024

/* ---- 3-DOF particle kernels ------------------------------------------- */

/* Fused per-particle pass; flags: 1 apply_jacobi, 2 finalize_velocity,
 * 4 cap_velocity, 8 integrate (applied in that order).
 * Replaces wp.launch(apply_jacobi / finalize_velocity / cap_velocity /
 * integrate) at src/xpbd3d/solver.py:465, 437, 441, 427. */
int xpbd_particles_pass(void* stream, float* X, float* V, float* XP, float* ACC, int flags,
                        float relax, float h, float damping, float max_speed, float gx, float gy,
                        float gz, int n);

/* kernels.py:49-73 `integrate`, launched at solver.py:427-429. */
int xpbd_integrate(void* stream, float* X, float* V, float* XP, float gx, float gy, float gz,
                   float h, int n);

/* kernels.py:326-339 `apply_jacobi`, launched at solver.py:465-467. */
int xpbd_apply_jacobi(void* stream, float* X, float* ACC, float relax, int n);

/* kernels.py:346-371 `finalize_velocity` + `cap_velocity` (cap skipped when
 * max_speed <= 0 or non-finite), launched at solver.py:437-442. */
int xpbd_finalize_velocity(void* stream, const float* X, const float* XP, float* V, float h,
                           float damping, float max_speed, int n);

/* kernels.py:219-322 `solve_constraints_jacobi`, launched at solver.py:459-464.
 * Rows must be sorted by their smallest endpoint when window > 0.
 * first != 0 treats lambda as 0 (the per-substep reset, solver.py:431) and
 * writes every row's lambda.  window = LDS accumulator window in particles
 * (0 = plain global atomics), rows_per_thread = rows per thread per tile. */
int xpbd_solve_jacobi(void* stream, const float* X, const float* XP, const int* R,
                      const float* ANC, float* LAM, float* ACC, float h, int first, int n_rows,
                      int rows_per_thread, int window);

/* One full Jacobi iteration, kernels.py:219-339 `solve_constraints_jacobi`
 * + `apply_jacobi` fused (solver.py:458-467), over particle tiles of
 * tile_size particles: reads Xin / LAMin, writes Xout / LAMout (distinct
 * buffers).  Rows grouped by the tile of their smallest endpoint (any order
 * inside a tile); tile_row_start[t] is the first row of tile t (n_tiles + 1
 * entries); halo_rows[halo_start[t] .. halo_start[t+1]) lists the rows whose
 * larger endpoint lies in tile t and smaller endpoint in an earlier tile.
 * Per-particle sums are exact (2^-40 m fixed point), so results do not depend
 * on the order of the atomic adds.  ANC may be NULL when no row is ATTACH.
 * finalize != 0 also applies finalize_velocity + cap_velocity
 * (kernels.py:346-371) and writes V.  ACC (optional, float4[n]) holds sums
 * from xpbd_solve_jacobi_dyn of this iteration; they join the average and
 * are cleared. */
int xpbd_jacobi_tile(void* stream, const float* Xin, float* Xout, const float* XP, float* V,
                     const int* R, const float* ANC, const float* LAMin, float* LAMout,
                     const int* tile_row_start, const int* halo_start, const int* halo_rows,
                     float h, int first, int tile_size, int n_particles, float relax,
                     int finalize, float damping, float max_speed, float* ACC);

/* The first iteration of a substep with kernels.py:49-73 `integrate` fused
 * in (solver.py:431-458: integrate, lambda reset, first projection + apply):
 * X holds the positions at the start of the substep and is only read; every
 * position the launch needs is predicted from X and V exactly as
 * xpbd_integrate_to computes it; x_prev (XP) of every particle is written,
 * lambda starts from zero (LAMout written for every row), Xout receives the
 * iterated positions (Xout != X).  V is not written: the caller finishes the
 * substep with at least one xpbd_jacobi_tile launch with finalize = 1.
 * Same tile tables as xpbd_jacobi_tile; gscale as in xpbd_integrate_to; ACC
 * (optional) as in xpbd_jacobi_tile, filled by xpbd_solve_jacobi_dyn_first. */
int xpbd_jacobi_tile_first(void* stream, const float* X, float* Xout, float* XP, const float* V,
                           const int* R, const float* ANC, float* LAMout,
                           const int* tile_row_start, const int* halo_start, const int* halo_rows,
                           float h, float gx, float gy, float gz, const float* gscale,
                           int tile_size, int n_particles, float relax, float* ACC);

/* The last iteration of a substep followed by the next substep's
 * kernels.py:49-73 `integrate` (solver.py:431-458: apply, finalize_velocity,
 * cap_velocity of substep k, then integrate of substep k + 1) in one launch:
 * as xpbd_jacobi_tile with finalize = 1, except that the finalized velocity
 * stays in registers: XP receives the iterated positions, Xout the positions
 * predicted from them (v + g h, gscale optional, exactly as
 * xpbd_integrate_to computes it) and V is not written, so the next substep
 * continues with a plain first launch (first = 1) and the frame's last
 * substep ends with xpbd_jacobi_tile (finalize = 1), which writes V.  Other
 * tiles read XP while a tile writes it: no row may be CONTACT (their
 * friction reads the other endpoint's x_prev); Xin != Xout. */
int xpbd_jacobi_tile_next(void* stream, const float* Xin, float* Xout, float* XP, const int* R,
                          const float* ANC, const float* LAMin, float* LAMout,
                          const int* tile_row_start, const int* halo_start, const int* halo_rows,
                          float h, int first, float gx, float gy, float gz, const float* gscale,
                          int tile_size, int n_particles, float relax, float damping,
                          float max_speed, float* ACC);

/* Compact description of the rows of the fused tile launches (built by
 * xpbd_amd/solver.py compact_tile_tables; no reference counterpart: it is
 * the device layout of the same rows).  The owned rows of tile t,
 * [tile_row_start[t], tile_row_start[t+1]), are ordered in three segments:
 *   [tile_row_start[t], dist_end[t])  DISTANCE rows stored compactly: ci
 *       int2 {(lo - t*tile_size) | (hi - t*tile_size) << 10, bits of rest}
 *       and cc float compliance, lo / hi the smaller / larger endpoint
 *       (hi - t*tile_size < 2^22; tile_size <= 1024).  Projecting (lo, hi)
 *       gives bit for bit the corrections of (a, b) in either order.
 *   [dist_end[t], gen_end[t])  every other row, read from R (int4).
 *   [gen_end[t], tile_row_start[t+1])  FLOOR rows of particles with exactly
 *       one FLOOR row, in particle order.  With pf set they are projected per
 *       particle by the apply pass from pf (float2 {floor_y, mu}, floor_y =
 *       -inf for a particle without one), their multipliers kept in place in
 *       pflam (float[n_particles]; the LAM entries of these rows are then
 *       neither read nor written); with pf NULL they are read from R like
 *       the general rows.
 * halo2 (int2, one per halo row, same order as halo_rows) holds {row, base}
 * with base = owner tile * tile_size for a compact row and -1 for a row read
 * from R.  ci and cc hold entries for every row index (only the compact
 * ones are read).  general must be nonzero when any row is read from R: a
 * nonempty [dist_end[t], gen_end[t]) segment, a halo entry with base -1, or,
 * with pf NULL, a nonempty FLOOR segment (with general = 0 such rows are
 * skipped).  tile_size >= 64. */
typedef struct xpbd_tile_compact {
  const int* dist_end;
  const int* gen_end;
  const int* ci;
  const float* cc;
  const int* halo2;
  const float* pf;
  float* pflam;
  int general;
  /* optional, launches with device-row sums (ACC): acc_tiles[t] = 0 when no
   * particle of tile t can be an endpoint of a device row (no colliding
   * particle), so the tile skips its ACC reads; NULL: every tile reads */
  const int* acc_tiles;
} xpbd_tile_compact;

/* xpbd_jacobi_tile, xpbd_jacobi_tile_first and xpbd_jacobi_tile_next over
 * the compact tables (cx replaces halo_rows): same results bit for bit, the
 * FLOOR multipliers in cx->pflam when cx->pf is set.  The launch of the
 * plain zero-start Jacobi plan (solver.py:427-467). */
int xpbd_jacobi_tile_cx(void* stream, const float* Xin, float* Xout, const float* XP, float* V,
                        const int* R, const float* ANC, const float* LAMin, float* LAMout,
                        const int* tile_row_start, const int* halo_start,
                        const xpbd_tile_compact* cx, float h, int first, int tile_size,
                        int n_particles, float relax, int finalize, float damping,
                        float max_speed, float* ACC);
int xpbd_jacobi_tile_first_cx(void* stream, const float* X, float* Xout, float* XP,
                              const float* V, const int* R, const float* ANC, float* LAMout,
                              const int* tile_row_start, const int* halo_start,
                              const xpbd_tile_compact* cx, float h, float gx, float gy, float gz,
                              const float* gscale, int tile_size, int n_particles, float relax,
                              float* ACC);
int xpbd_jacobi_tile_next_cx(void* stream, const float* Xin, float* Xout, float* XP,
                             const int* R, const float* ANC, const float* LAMin, float* LAMout,
                             const int* tile_row_start, const int* halo_start,
                             const xpbd_tile_compact* cx, float h, int first, float gx,
                             float gy, float gz, const float* gscale, int tile_size,
                             int n_particles, float relax, float damping, float max_speed,
                             float* ACC);

/* kernels.py:49-73 `integrate` reading Xin and writing Xout (may alias);
 * gscale (optional, float[n]) scales gravity per particle (batched scenes). */
int xpbd_integrate_to(void* stream, const float* Xin, float* Xout, float* V, float* XP, float gx,
                      float gy, float gz, float h, const float* gscale, int n);

/* solve_constraints_jacobi (kernels.py:219-322) over rows [0, min(*n_rows_dev,
 * cap)) of a segment generated on the device (CONTACT rows), accumulating
 * into ACC with global atomics. */
int xpbd_solve_jacobi_dyn(void* stream, const float* X, const float* XP, const int* R,
                          float* LAM, float* ACC, float h, int first, const int* n_rows_dev,
                          int cap);
/* The same for the first iteration of a substep with integrate folded in
 * (as xpbd_jacobi_tile_first: X = positions at the start of the substep,
 * predicted from X and V on the fly, x_prev = X; lambda from zero). */
int xpbd_solve_jacobi_dyn_first(void* stream, const float* X, const float* V, const int* R,
                                float* LAM, float* ACC, float h, float gx, float gy, float gz,
                                const float* gscale, const int* n_rows_dev, int cap);

/* Device CONTACT generation, Solver._rebuild_contacts (solver.py:275-342)
 * with the row order of _grid_candidate_pairs (solver.py:520-562):
 * margin = min(contact_margin + 3 dt max|v|, r_max) (float64, as the host
 * computes it), cell = 2 r_max + margin; for each of the 14 cell offsets
 * (self, then the 13 forward neighbours in the reference's order), for each
 * body k in ascending API index, the bodies p of the offset cell in
 * ascending API index (same cell: p after k) with |xk - xp|^2 <= (rk + rp +
 * margin)^2 and (min, max) of their API indices not in `excl` give rows
 * CONTACT(k, p, rest rk + rp, mu sqrt(mu_k mu_p)), written with device
 * indices into ROWS (int4, clamped to cap); the full count goes to
 * count[0].  api[i] = API index of device particle i (NULL: identity);
 * excl = sorted keys (lo << 32 | hi) of the API index pairs joined by a
 * constraint (NULL with n_excl = 0: none).  radius[i] = 0 excludes
 * particle i.  clist (optional): the device indices of the colliding
 * particles (radius > 0) in ascending API index, nc of them; the search,
 * its counts and scans then cover those only (the rows are the same).
 * NULL: every particle is visited. */
long long xpbd_contacts_workspace_bytes(int n);
int xpbd_sphere_contacts(void* stream, void* ws, const float* X, const float* V,
                         const float* radius, const float* mu, const int* api,
                         const int* clist, int nc, const long long* excl, int n_excl,
                         double contact_margin, double dt, float r_max, int n, int cap,
                         int* ROWS, int* count);

/* Whole frames of a small scene in one workgroup (solver.py:412-467): the
 * positions stay in LDS, phases are separated by workgroup barriers.
 * gs = 1: Gauss-Seidel over the color slices color_starts[0..num_colors]
 * (device array; rows sorted by color), bit-identical to xpbd_solve_color
 * launches; gs = 0: Jacobi with the fused tile kernel's arithmetic and exact
 * sums (identical to xpbd_jacobi_tile).  Runs `frames` frames of `substeps`
 * substeps x `iterations` iterations; n_particles <=
 * xpbd_frame_block_capacity(gs). */
int xpbd_frame_block(void* stream, float* X, float* V, float* XP, const int* R, const float* ANC,
                     float* LAM, const int* color_starts, int num_colors, int n_particles,
                     int n_rows, int substeps, int iterations, float h, float gx, float gy,
                     float gz, float damping, float relax, float max_speed, int gs, int frames);
/* Learned warm start of a substep (no reference counterpart; it replaces the
 * zero start of solver.py:431), launched after xpbd_ws_net and before the
 * substep's Jacobi iterations, which then run with first = 0 from LAMout.
 * Every row starts from lambda0 = 4 sigmoid(Y[row]) (-C / (w + alpha~))
 * (Y: the network pass's bfloat16 outputs)
 * (one-sided rows only while violated), written to LAMout; each particle
 * moves by relax x the mean over its rows of w_i grad C lambda0 (Xin ->
 * Xout, Xin != Xout).  predict = 1 folds integrate in as
 * xpbd_jacobi_tile_first does (Xin = positions at the start of the substep;
 * XP written; V read), predict = 0 reads integrated positions.  Tile tables
 * as xpbd_jacobi_tile.  ws_on (optional device int, NULL = on) switches the
 * start off when *ws_on == 0 (invert != 0: when *ws_on != 0): lambda0 = 0
 * for every row and Xout = the (predicted) positions, after which the
 * iterations run exactly as from the zero start. */
int xpbd_jacobi_tile_warm(void* stream, const float* Xin, float* Xout, float* XP, const float* V,
                          const int* R, const float* ANC, float* LAMout,
                          const int* tile_row_start, const int* halo_start, const int* halo_rows,
                          const unsigned short* Y, float h, float gx, float gy, float gz,
                          const float* gscale, int tile_size, int n_particles, float relax,
                          int predict, const int* ws_on, int invert);
/* xpbd_jacobi_tile_warm over the compact tables (cx replaces halo_rows):
 * the same lambda0 and positions; with cx->pf the single FLOOR rows' lambda0
 * goes to cx->pflam (zero when the start is off), where the following
 * xpbd_jacobi_tile_cx launches read it. */
int xpbd_jacobi_tile_warm_cx(void* stream, const float* Xin, float* Xout, float* XP,
                             const float* V, const int* R, const float* ANC, float* LAMout,
                             const int* tile_row_start, const int* halo_start,
                             const xpbd_tile_compact* cx, const unsigned short* Y, float h, float gx,
                             float gy, float gz, const float* gscale, int tile_size,
                             int n_particles, float relax, int predict, const int* ws_on,
                             int invert);
/* Largest particle count of xpbd_frame_block (LDS-bound). */
int xpbd_frame_block_capacity(int gs);

/* kernels.py:105-208 `solve_constraints_color` over rows [start, start+count)
 * of a color-sorted row array, launched per color at solver.py:444-456. */
int xpbd_solve_color(void* stream, float* X, const float* XP, const int* R, const float* ANC,
                     float* LAM, float h, int first, int start, int count);

/* ---- 3-DOF frame driver (Solver.step, solver.py:412-467) -------------- */

typedef struct xpbd_particles_desc {
  float* x;              /* float4[n_particles] */
  float* v;              /* float4[n_particles] */
  float* x_prev;         /* float4[n_particles] */
  float* acc;            /* float4[n_particles], zero on entry (Jacobi) */
  const int* rows;       /* int4[n_rows]; color-sorted (GS) or endpoint-sorted (Jacobi) */
  const float* anchors;  /* float4[n_rows] */
  float* lam;            /* float[n_rows] */
  int n_particles;
  int n_rows;
  const int* color_starts; /* host int[num_colors + 1] (GS only; copied) */
  int num_colors;
  int mode;              /* XPBD_MODE_GS or XPBD_MODE_JACOBI */
  int substeps;
  int iterations;
  float h;               /* substep length dt / substeps, rounded to float */
  float gx, gy, gz;
  float damping;
  float relax;
  float max_speed;       /* <= 0 or non-finite disables the cap */
  int rows_per_thread;   /* Jacobi tile shape */
  int window;
  /* fused-tile Jacobi (mode XPBD_MODE_JACOBI_TILE) */
  float* x_alt;          /* float4[n_particles], second position buffer */
  float* lam_alt;        /* float[n_rows], second multiplier buffer */
  const int* tile_row_start; /* int[n_tiles + 1] */
  const int* halo_start;     /* int[n_tiles + 1] */
  const int* halo_rows;      /* int[halo_start[n_tiles]] */
  int tile_size;
  /* device-generated row segment (tile mode): projected each iteration with
   * global atomics into acc, merged by the tile kernel's apply */
  const int* dyn_rows;   /* int4[dyn_cap] */
  float* dyn_lam;        /* float[dyn_cap] */
  const int* dyn_count;  /* int[1], device */
  int dyn_cap;
  const float* gscale;   /* float[n_particles] gravity factor, or NULL (tile mode) */
  /* learned warm start (XPBD_MODE_JACOBI_TILE): the network packed by
   * xpbd_mlp_pack, or NULL for the reference's zero start; every substep
   * then starts with xpbd_ws_net (outputs into ws_y) and
   * xpbd_jacobi_tile_warm */
  const void* ws_image;
  unsigned short* ws_y;  /* bfloat16 bits [n_rows] */
  /* Tile mode: 1 folds integrate into the first iteration of every substep
   * (xpbd_jacobi_tile_first; >= 2 iterations or the learned warm start), 2
   * folds it into the first substep's first iteration and into the previous
   * substep's last iteration for the others (xpbd_jacobi_tile_next; no
   * CONTACT rows among `rows`; any iteration count), 0 keeps a separate
   * xpbd_integrate_to launch per substep.  Results are identical. */
  int fuse_first;
  /* Small scenes (n_particles <= xpbd_frame_block_capacity): 1 runs whole
   * frames in one workgroup (xpbd_frame_block, GS or window-Jacobi mode; one
   * launch per xpbd_frame3_run call); color_starts_dev is the device copy of
   * color_starts (GS). */
  int block;
  const int* color_starts_dev;
  /* learned warm start: the network pass runs on substeps sub % ws_every ==
   * 0 of each frame (<= 1: every substep); the other substeps reuse ws_y
   * with their own first plain steps */
  int ws_every;
  /* The plain zero start as judge and fallback of the learned start
   * (tile mode with ws_image, fuse_first 2, substeps >= 2): ws_on (device
   * int, 1 = learned start) is read by every warm-start launch.  With
   * ws_judge = k >= 1, the first substep of every k-th frame (counted by
   * the plan from its first run) is first run with the other start from the
   * same state into the judge buffers (ws_jx, ws_jx_alt, ws_jxp float4[n];
   * ws_jlam, ws_jlam_alt float[n_rows], ws_jy bfloat16 bits[n_rows]; ws_jacc float4[n] and
   * ws_jdlam float[dyn_cap] with device rows), then with the solver's own;
   * both squared residuals go to ws_rsum (double[2], zero at the start) and
   * xpbd_ws_decide sets ws_on for the rest of the frame and the following
   * ones (stats in ws_stats, int[3], optional).  Judged and other frames
   * are two captured graphs. */
  int* ws_on;
  int ws_judge;
  float* ws_jx;
  float* ws_jx_alt;
  float* ws_jxp;
  float* ws_jlam;
  float* ws_jlam_alt;
  unsigned short* ws_jy;
  float* ws_jacc;
  float* ws_jdlam;
  double* ws_rsum;
  int* ws_stats;
  /* Two-iteration substeps in one launch (tile mode, iterations = 2,
   * fuse_first = 2, no learned start, no device rows; see
   * xpbd_jacobi_tile_pair): when pair_p1 is set, every substep after a
   * frame's first runs as one xpbd_jacobi_tile_pair launch with these
   * tables, its x_prev ping-ponging between x_prev and x_prev_alt
   * (float4[n_particles]); the last substep leaves it in x_prev. */
  float* x_prev_alt;
  const int* pair_ring_start;
  const int* pair_ring1_n;
  const int* pair_ring_idx;
  const int* pair_p1_start;
  const int* pair_p1;
  const int* pair_p2_start;
  const int* pair_p2;
  int pair_lmax, pair_amax, pair_k1max;
  /* Compact tile tables (tile mode, tile_size >= 64, optional; copied at
   * create): the plain and warm-start tile launches become their _cx forms.
   * pf may be set only without the pair launch, which reads FLOOR
   * multipliers from lam. */
  const xpbd_tile_compact* compact;
  /* optional: a second judge x_prev buffer (float4[n]), so that judged
   * substeps can also run as one launch (xpbd_jacobi_tile_pair_warm) */
  float* ws_jxp_alt;
} xpbd_particles_desc;

/* A whole two-iteration substep of the fused tile plan in one launch, for
 * small scenes where launches cost latency: the first iteration
 * (xpbd_jacobi_tile, first = 1, on predicted positions) and the second with
 * finalize + cap and, unless `last`, the next substep's integrate
 * (xpbd_jacobi_tile_next; with `last`, xpbd_jacobi_tile with finalize),
 * results bit for bit those of the two launches.  Each tile also runs the
 * first iteration of the particles one row away (ring 1), from every row
 * incident to them: tables of xpbd_amd/solver.py tile_pair_tables (rings,
 * phase-1 rows with local endpoints, phase-2 rows).  x_prev is read from
 * XPin and written to XPout (!= XPin unless last): the next substep's start
 * positions, or with `last` a copy of XPin.  Multipliers of the second
 * iteration go to LAMout (owner tiles).  predict = 1 (not with last): the
 * substep's integrate too (solver.py:427, kernels.py:49-73; the frame's first
 * substep, as xpbd_jacobi_tile_first): Xin and V are the substep's start,
 * x_prev is Xin (XPin is not read).  Dynamic LDS:
 * xpbd_jacobi_tile_pair_lds_bytes(lmax, amax, k1max) <= 64 KiB. */
int xpbd_jacobi_tile_pair(void* stream, const float* Xin, float* Xout, const float* XPin,
                          float* XPout, float* V, const int* R, const float* ANC, float* LAMout,
                          const int* ring_start, const int* ring1_n, const int* ring_idx,
                          const int* p1_start, const int* p1, const int* p2_start, const int* p2,
                          float h, int tile_size, int n, float relax, float damping,
                          float max_speed, float gx, float gy, float gz, const float* gscale,
                          int lmax, int amax, int k1max, int last, int predict);
size_t xpbd_jacobi_tile_pair_lds_bytes(int lmax, int amax, int k1max);
/* The learned-start form of xpbd_jacobi_tile_pair (no reference counterpart;
 * it replaces the zero start of solver.py:431): a one-iteration substep
 * with the learned start, i.e. xpbd_jacobi_tile_warm (Y, ws_on, invert as
 * there) followed by one xpbd_jacobi_tile / xpbd_jacobi_tile_next launch,
 * bit for bit, with the same tables; Y NULL gives xpbd_jacobi_tile_pair. */
int xpbd_jacobi_tile_pair_warm(void* stream, const float* Xin, float* Xout, const float* XPin,
                               float* XPout, float* V, const int* R, const float* ANC,
                               float* LAMout, const int* ring_start, const int* ring1_n,
                               const int* ring_idx, const int* p1_start, const int* p1,
                               const int* p2_start, const int* p2, float h, int tile_size, int n,
                               float relax, float damping, float max_speed, float gx, float gy,
                               float gz, const float* gscale, int lmax, int amax, int k1max,
                               int last, int predict, const unsigned short* Y, const int* ws_on,
                               int invert);

int xpbd_frame3_create(const xpbd_particles_desc* desc, void** handle);
/* Runs `frames` frames on `stream`; use_graph != 0 captures the frame once
 * into a hipGraph (replacing the per-launch loop of solver.py:426-442). */
int xpbd_frame3_run(void* handle, void* stream, int frames, int use_graph);
int xpbd_frame3_destroy(void* handle);

/* ---- 6-DOF rigid bodies (src/xpbd3d/kernels_6dof.py) --------------------
 * Body buffers are float4[n]: X (x, inv_mass), Q (quat xyzw), V, W (v, omega),
 * XP, QP (substep start), IB (inverse body inertia diagonal), HE (half
 * extents, .w = collision group as int bits); accumulators DX (dx, count),
 * DR (d-rotation), DV (dv, count), DW (d-omega); LAMF float[8n].
 * Pairs int2[cap] with the live count in PC[0]; manifold points are
 * compacted: PTI int4 (pair, incident, reference, slot), PTN/PTOI/PTOR
 * float4, LAMP float, live count PTC[0], capacity 8*cap; MC int[cap] holds
 * the points per pair.  Joints: JI int2 (a, b), JA float4 (anchor_a, rest),
 * JB float4 (anchor_b, compliance), LAMJ float. */

/* kernels_6dof.py:67-102 `integrate_bodies` (solver_6dof.py:370-373); also
 * zeroes LAMF and PTC[0] (the lambda resets of solver_6dof.py:374-375).
 * gscale (optional, float[n]) scales gravity per body (batched scenes). */
int xpbd6_integrate(void* stream, float* X, float* Q, float* V, float* W, float* XP, float* QP,
                    const float* IB, float* LAMF, int* PTC, float gx, float gy, float gz,
                    float h, int gyro, const float* gscale, int n);
/* kernels_6dof.py:233-382 `generate_box_manifold` (solver_6dof.py:378-382). */
int xpbd6_box_manifold(void* stream, const float* X, const float* Q, const float* HE,
                       const int* PAIRS, const int* PC, float margin, int cap, int* MC, int* PTI,
                       float* PTN, float* PTOI, float* PTOR, float* LAMP, int* PTC);
/* kernels_6dof.py:440-483 `solve_joints` (solver_6dof.py:385-390); first != 0
 * resets lambda (solver_6dof.py:376-377). */
int xpbd6_solve_joints(void* stream, const float* X, const float* Q, const float* IB,
                       const int* JI, const float* JA, const float* JB, float* LAMJ, float h,
                       int first, float* DX, float* DR, int nj);
/* kernels_6dof.py:158-211 `solve_floor_contacts` (solver_6dof.py:391-395). */
int xpbd6_solve_floor(void* stream, const float* X, const float* Q, const float* XP,
                      const float* QP, const float* IB, const float* HE, float floor_y,
                      float mu_s, float* LAMF, float* DX, float* DR, int n);
/* kernels_6dof.py:385-430 `solve_box_manifold` (solver_6dof.py:396-401). */
int xpbd6_solve_box(void* stream, const float* X, const float* Q, const float* IB,
                    const int* PTI, const float* PTN, const float* PTOI, const float* PTOR,
                    float* LAMP, const int* PTC, float* DX, float* DR, int max_points);
/* kernels_6dof.py:486-506 `apply_jacobi_6dof` (solver_6dof.py:402-404). */
int xpbd6_apply_jacobi(void* stream, float* X, float* Q, float* DX, float* DR, float relax,
                       int n);
/* kernels_6dof.py:512-535 `finalize_velocity_6dof` (solver_6dof.py:405-408). */
int xpbd6_finalize_velocity(void* stream, const float* X, const float* Q, const float* XP,
                            const float* QP, float* V, float* W, float h, float lin_damp,
                            float ang_damp, int n);
/* kernels_6dof.py:541-590 `velocity_floor` (solver_6dof.py:409-413). */
int xpbd6_velocity_floor(void* stream, const float* X, const float* Q, const float* IB,
                         const float* HE, float mu_d, float restitution, const float* LAMF,
                         float h, const float* V, const float* W, float* DV, float* DW, int n);
/* kernels_6dof.py:593-659 `velocity_box` (solver_6dof.py:414-419). */
int xpbd6_velocity_box(void* stream, const float* X, const float* Q, const float* IB,
                       const int* PTI, const float* PTN, const float* PTOI, const float* PTOR,
                       const float* LAMP, const int* PTC, float mu_d, float restitution, float h,
                       const float* V, const float* W, float* DV, float* DW, int max_points);
/* kernels_6dof.py:662-691 `apply_velocity_6dof` + `cap_velocity_6dof`
 * (solver_6dof.py:420-426). */
int xpbd6_apply_velocity(void* stream, const float* X, float* V, float* W, float* DV, float* DW,
                         float max_lin, float max_ang, int do_cap, int n);

/* Fused launches of the same kernels, one launch per phase of a substep
 * (solver_6dof.py:357-426 records the phases as separate launches).
 * begin_substep: apply_velocity of the previous substep when apply != 0
 * (solver_6dof.py:420-426), then integrate_bodies and the lambda resets
 * (solver_6dof.py:370-375).  solve_all: one thread per joint, per box
 * corner on the floor (8 lanes per body, summed in corner order) and per
 * live manifold point (solver_6dof.py:385-401); every body's correction is
 * added atomically into DX/DR.  apply_finalize: apply_jacobi and, when
 * finalize != 0, finalize_velocity (solver_6dof.py:402-408).  velocity_all:
 * velocity_floor (per corner, as solve_all) and velocity_box
 * (solver_6dof.py:409-419), added atomically into DV/DW. */
int xpbd6_begin_substep(void* stream, float* X, float* Q, float* V, float* W, float* XP,
                        float* QP, const float* IB, float* LAMF, int* PTC, float* DV, float* DW,
                        float max_lin, float max_ang, int do_cap, int apply, float gx, float gy,
                        float gz, float h, int gyro, const float* gscale, int n);
int xpbd6_solve_all(void* stream, const float* X, const float* Q, const float* XP,
                    const float* QP, const float* IB, const float* HE, const int* JI,
                    const float* JA, const float* JB, float* LAMJ, float h, int first, int nj,
                    float floor_y, float mu_s, float* LAMF, int n, const int* PTI,
                    const float* PTN, const float* PTOI, const float* PTOR, float* LAMP,
                    const int* PTC, int max_points, float* DX, float* DR);
int xpbd6_apply_finalize(void* stream, float* X, float* Q, float* DX, float* DR, float relax,
                         int finalize, const float* XP, const float* QP, float* V, float* W,
                         float h, float lin_damp, float ang_damp, int n);
int xpbd6_velocity_all(void* stream, const float* X, const float* Q, const float* IB,
                       const float* HE, float mu_d, float restitution, const float* LAMF,
                       float h, const float* V, const float* W, int n, const int* PTI,
                       const float* PTN, const float* PTOI, const float* PTOR, const float* LAMP,
                       const int* PTC, int max_points, float* DV, float* DW);

typedef struct xpbd_rigid_desc {
  float *x, *q, *v, *w, *x_prev, *q_prev; /* float4[n] */
  const float *inv_I, *he;                 /* float4[n] */
  float *dx, *drot, *dv, *dw;              /* float4[n], zero on entry */
  float* lam_floor;                        /* float[8n] */
  int n;
  const int* pairs;                        /* int2[cap] */
  const int* pair_count;                   /* int[1], device */
  int cap;
  int* m_count;                            /* int[cap] */
  int* pt_idx;                             /* int4[8cap] */
  float *pt_n, *pt_oi, *pt_or;             /* float4[8cap] */
  float* lam_pt;                           /* float[8cap] */
  int* pt_count;                           /* int[1]; int[2] with aux */
  const int* j_idx;                        /* int2[nj] */
  const float *j_a, *j_b;                  /* float4[nj] */
  float* lam_j;                            /* float[nj] */
  int nj;
  int substeps, iterations, gyroscopic;
  float h, gx, gy, gz, floor_y, friction, restitution, lin_damp, ang_damp, relax;
  float max_speed, max_omega, man_margin;
  const float* gscale;                     /* float[n] gravity factor, or NULL */
  float* aux;  /* float4[6n] for the two-launch substep (zero on entry), or NULL */
} xpbd_rigid_desc;

/* Solver6DOF._record (solver_6dof.py:357-426) as a plan; run() replays one
 * hipGraph per frame (solver_6dof.py:519-527). */
/* Whole frames of a small rigid scene in one workgroup (n <=
 * xpbd6_frame_block_capacity() = 256 bodies, state held in LDS):
 * per frame the LBVH path's broad phase (compute_body_aabb + the emit_pairs
 * filter of kernels_bvh.py:21-85, every pair (i, j > i) tested, pairs
 * written in (i, j) order with the count in pair_count[0]) followed by the
 * substep loop of xpbd_frame6_run (solver_6dof.py:357-426), `frames` times.
 * desc->cap must be >= n (n - 1) / 2 (no overflow); LO/HI float4[n] scratch;
 * dt = the frame length (AABB velocity margin).  pairs and pair_count are
 * written despite their const in the descriptor. */
int xpbd6_frame_block(void* stream, const xpbd_rigid_desc* desc, float* LO, float* HI, float dt,
                      int frames);
int xpbd6_frame_block_capacity(void);

/* The substep of solver_6dof.py:357-426 in two launches (plus one per
 * further iteration), iterations >= 1; the frame plan's sequence when
 * desc->aux is set (float4[6n]: predicted pose, then the parity-1 copies of
 * dx, drot, dv, dw; pt_count int[2]; all zero on the first call).
 * substep_contacts(sub): apply_velocity + cap of substep sub-1 (sub > 0;
 * solver_6dof.py:420-426) and integrate_bodies (solver_6dof.py:370-373),
 * recomputed on values by every lane that needs a body, then the first
 * Jacobi iteration: joints, floor corners and, in workgroups of their own,
 * the box manifold (solver_6dof.py:376-384) one pair per lane followed by
 * the points those pairs emitted, one per lane (solver_6dof.py:385-401);
 * stores the predicted pose in aux and the start pose in x_prev / q_prev.  Further iterations: xpbd6_apply_finalize
 * (finalize 0) and xpbd6_solve_all (first 0) on the predicted pose.
 * substep_velocity(sub): apply_jacobi of the last iteration and
 * finalize_velocity on values (solver_6dof.py:402-408), then the velocity
 * pass (solver_6dof.py:409-419), one point per lane, each pair's terms
 * summed before they are added; stores x, q, v, w.  substep_end: the
 * frame's last apply_velocity + cap; every sum buffer and both point
 * counters back to zero.  Sums of substep sub use parity sub & 1 of the
 * sum buffers, so no lane reads what another lane of the same launch
 * writes; results equal those of the separate launches up to the order of
 * the atomic sums. */
int xpbd6_substep_contacts(void* stream, const xpbd_rigid_desc* desc, int sub);
int xpbd6_substep_velocity(void* stream, const xpbd_rigid_desc* desc, int sub);
int xpbd6_substep_end(void* stream, const xpbd_rigid_desc* desc);
/* substep_end that also reports the frame: when ring (device address of
 * host-mapped int[2 ring_len]) is set, slot frames[0] % ring_len gets
 * (pair_count[0], frames[0] + 1) and frames[0] (device int) advances.
 * All three substep launches write no state when pair_count[0] > cap (the
 * broad phase overflowed): the frame is skipped, and a caller that grows
 * the pair buffers runs it again. */
int xpbd6_substep_end_report(void* stream, const xpbd_rigid_desc* desc, int* ring, int ring_len,
                             int* frames);

int xpbd_frame6_create(const xpbd_rigid_desc* desc, void** handle);
int xpbd_frame6_run(void* handle, void* stream, int frames, int use_graph);
int xpbd_frame6_destroy(void* handle);
/* The frame's broad phase inside the plan (two-launch substep plans only;
 * solver_6dof.py:428-454 runs it on the host's schedule and reads the count
 * back every frame): kind 1 LBVH (xpbd6_broadphase_lbvh, or _leaf with
 * leaf >= 0), 2 hash grid (rsmax, cell); ws = the broad-phase workspace,
 * dt = the frame length.  The frame then needs no host round trip: its
 * status goes to a host-mapped ring that xpbd_frame6_status reads.
 * kind 0: the caller supplies the pairs again. */
int xpbd_frame6_set_broadphase(void* handle, int kind, void* ws, int leaf, float rsmax,
                               float cell, float dt);
/* Status of the frames run so far (wait != 0: after synchronizing the device):
 * out[0] frames not yet reported, out[1] the first frame whose pair count
 * exceeded cap (-1: none), out[2] that count, out[3] frames run from it on,
 * all of which were skipped (solver_6dof.py:443-453 grows the pair buffers
 * and emits again; the caller grows them, makes a new plan and runs those
 * frames again). */
int xpbd_frame6_status(void* handle, int wait, long long* out);

/* ---- 6-DOF broad phase (kernels_bvh.py, kernels_hashgrid.py) ------------
 * Both write the candidate pairs (i < j, (i, j) ascending) into PAIRS
 * (clamped to cap) and the full count into PC[0]; the caller reads PC[0],
 * grows cap and calls again on overflow (solver_6dof.py:443-453). */

/* kernels_bvh.py:21-48 `compute_body_aabb`: LO/HI float4[n]. */
int xpbd6_body_aabb(void* stream, const float* X, const float* Q, const float* V, const float* W,
                    const float* HE, float man_margin, float dt, float* LO, float* HI, int n);
/* Bytes of the caller-owned workspace for n bodies. */
long long xpbd6_lbvh_workspace_bytes(int n);
long long xpbd6_hashgrid_workspace_bytes(int n);
/* compute_body_aabb + wp.Bvh(lbvh).rebuild + emit_pairs
 * (solver_6dof.py:428-454; kernels_bvh.py:52-85). */
int xpbd6_broadphase_lbvh(void* stream, void* ws, const float* X, const float* Q, const float* V,
                          const float* W, const float* HE, float man_margin, float dt, int n,
                          int cap, int* PAIRS, int* PC);
/* The same with the hierarchy's leaf size chosen by the caller: leaf_size
 * bodies (consecutive in Morton order) per LBVH leaf, 1..64, or 0 for the
 * all-pairs test; xpbd6_broadphase_lbvh uses 0 up to 4096 bodies and 1
 * above.  The pair list is the same for every leaf size (the reference's
 * leaf size is internal to wp.Bvh and unobservable); an autotuner picks it
 * by timing. */
int xpbd6_broadphase_lbvh_leaf(void* stream, void* ws, const float* X, const float* Q,
                               const float* V, const float* W, const float* HE, float man_margin,
                               float dt, int n, int cap, int* PAIRS, int* PC, int leaf_size);
/* Allocation-free variant of xpbd6_broadphase_lbvh (no counterpart in the
 * reference): for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same inputs, outputs and pair list as xpbd6_broadphase_lbvh, but its
 * scratch is a fixed 7 MiB array inside the library, one per GPU, held until
 * the process ends; it calls no allocating routine.  n above
 * XPBD_NOALLOC_MAX_BODIES returns XPBD_ERR_NOALLOC_TOO_LARGE before any
 * launch.  Every call on a GPU shares that GPU's array: calls that overlap
 * on one GPU (two streams at once) race; calls ordered on one stream do not.
 * Nothing in the array needs to start at zero (each pass writes what the
 * next reads). */
#define XPBD_NOALLOC_MAX_BODIES 65536
#define XPBD_ERR_NOALLOC_TOO_LARGE (-2)
int xpbd6_broadphase_lbvh_noalloc(void* stream, const float* X, const float* Q, const float* V,
                                  const float* W, const float* HE, float man_margin, float dt,
                                  int n, int cap, int* PAIRS, int* PC);
/* wp.HashGrid.build + emit_pairs_hashgrid (solver_6dof.py:456-479;
 * kernels_hashgrid.py:37-88); cell = 2 rs_max + margin (solver_6dof.py:335-337). */
int xpbd6_broadphase_hashgrid(void* stream, void* ws, const float* X, const float* Q,
                              const float* V, const float* W, const float* HE, float man_margin,
                              float dt, float rs_max, float cell, int n, int cap, int* PAIRS,
                              int* PC);
/* 30-bit Morton code of a point in [0,1)^3 (10 bits per axis, x highest);
 * no reference counterpart (Warp builds its LBVH internally). */
unsigned xpbd_morton3(float x, float y, float z);

/* ---- learned lambda warm start (no reference counterpart) -----------------
 * MLP 12 -> 64 -> 64 -> 1 (GELU, tanh form), 5057 float32 parameters
 * W1[12][64] b1[64] W2[64][64] b2[64] W3[64] b3; BF16 MFMA products with
 * FP32 accumulation.  It predicts the multipliers that replace the zero reset
 * of solver.py:431.  F is float[m][12] (features), Y/T float[m]. */
int xpbd_mlp_param_count(void);
int xpbd_mlp_forward(void* stream, const float* theta, const float* F, float* Y, int m);
/* The network packed for the warm-start network pass (bfloat16 weights in
 * the LDS layout, then b3): image of xpbd_mlp_image_bytes() bytes, 16-byte
 * aligned. */
long long xpbd_mlp_image_bytes(void);
int xpbd_mlp_pack(void* stream, const float* theta, void* image);
long long xpbd_mlp_train_workspace_bytes(int nblocks);
/* Gradient of the mean squared error over m rows (grad float[5057]) and the
 * loss (loss float[1]); nblocks partial sums reduced in block order. */
int xpbd_mlp_grad(void* stream, const float* theta, const float* F, const float* T, int m,
                  int nblocks, void* ws, float* grad, float* loss);
/* One Adam step (step >= 1) on theta with moment buffers mom/vel. */
int xpbd_mlp_adam(void* stream, float* theta, float* mom, float* vel, const float* grad,
                  float lr, float beta1, float beta2, float eps, int step);

/* Row features of the MLP (float[m][12]) at the predicted positions X with
 * x_prev XP (velocities (X - XP) / h), as the warm-start launch computes
 * them; and the squared-residual sum (r = C + alpha~ lambda for two-sided
 * rows, max(0, -C) for one-sided rows) added to *out_sum. */
int xpbd_ws_features(void* stream, const float* X, const float* XP, const int* R,
                     const float* ANC, float h, float* F, int n_rows);
/* The network pass of a warm-started substep: per row the features (as
 * xpbd_ws_features; at the positions integrate would predict from Xin and V
 * when predict = 1, x_prev = Xin; otherwise at Xin with x_prev XP) and the
 * network's output Y[row] (as xpbd_mlp_forward on those features, stored
 * as bfloat16 bits, round to nearest even); image =
 * the network packed by xpbd_mlp_pack.  ws_on / invert as in
 * xpbd_jacobi_tile_warm: with the start off the launch does nothing. */
int xpbd_ws_net(void* stream, const float* Xin, const float* XP, const float* V, const int* R,
                const float* ANC, const void* image, unsigned short* Y, float h, float gx, float gy,
                float gz, const float* gscale, int n_rows, int predict, const int* ws_on,
                int invert);
int xpbd_ws_residual(void* stream, const float* X, const int* R, const float* ANC,
                     const float* LAM, float h, double* out_sum, int n_rows);
/* The plain zero start as judge and fallback of the learned start (one
 * thread): sums[0] is the squared residual of the substep the solver ran,
 * sums[1] that of the same substep run with the other start; *ws_on becomes
 * 1 (learned start) when the learned start's sum is no larger, else 0 (zero
 * start); stats (optional, int[3]) counts the decisions for the learned
 * start, those for the zero start, and the switches; the sums are cleared. */
int xpbd_ws_decide(void* stream, double* sums, int* ws_on, int* stats);

/* ---- device coloring (src/xpbd3d/coloring.py:26-89) ----------------------
 * The reference's greedy colouring (stable order on -(deg a + deg b), each
 * row the smallest colour its earlier neighbours left free) computed on the
 * device as Jones-Plassmann rounds with that order as the priority (the same
 * colours), in one workgroup: rows R[0 .. n_rows_static + min(*dyn_count,
 * dyn_cap)) (int4, API order: static rows, then the device CONTACT rows;
 * dyn_count NULL = none) over n_bodies <= xpbd_color_rows_max_bodies()
 * bodies.  Outputs: color[row]; order = the rows grouped by colour (any
 * order inside a colour), color_starts[c] its offsets (<= 257 ints),
 * *num_colors (-1 when more than 256 colours were needed).  inc: scratch of
 * 2 (n_rows_static + dyn_cap) ints.  Replaces the host colouring of
 * solver.py:394-401 for frames with device CONTACT rows. */
int xpbd_color_rows_max_bodies(void);
int xpbd_color_rows_device(void* stream, const int* R, int n_rows_static, const int* dyn_count,
                           int dyn_cap, int n_bodies, int* inc, int* color, int* order,
                           int* color_starts, int* num_colors);
/* Whole Gauss-Seidel frames in one workgroup (as xpbd_frame_block with
 * gs = 1) over rows coloured on the device: R, ANC and LAM in API order,
 * colour c's rows at order[color_starts[c] .. color_starts[c + 1]), the
 * colour count read from *num_colors_dev (a negative count projects no
 * row).  solver.py:412-467 for self-collision frames without the host. */
int xpbd_frame_block_ordered(void* stream, float* X, float* V, float* XP, const int* R,
                             const float* ANC, float* LAM, const int* color_starts,
                             const int* order, const int* num_colors_dev, int n_particles,
                             int substeps, int iterations, float h, float gx, float gy, float gz,
                             float damping, float relax, float max_speed, int frames);

/* ---- host coloring (src/xpbd3d/coloring.py) ----------------------------- */

/* coloring.py:26-71 `color_constraints`; returns the number of colors. */
int xpbd_color_constraints(int n_bodies, const int* body_a, const int* body_b, int n_rows,
                           int* color_out);
/* coloring.py:74-89 `color_ranges`: order_out[n_rows], starts_out[num_colors+1]. */
int xpbd_color_ranges(const int* color, int n_rows, int num_colors, int* order_out,
                      int* starts_out);

/* ---- library info -------------------------------------------------------- */
/* Returns a static string naming the build (architecture, flags). */
const char* xpbd_build_info(void);
/* Message for a return code of any entry point. */
const char* xpbd_error_string(int code);

#ifdef __cplusplus
}
#endif
#endif /* XPBD_H_ */
