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

/* Largest LDS accumulator window (particles) of the Jacobi projection. */
#define XPBD_MAX_WINDOW 8192

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
} xpbd_particles_desc;

int xpbd_frame3_create(const xpbd_particles_desc* desc, void** handle);
/* Runs `frames` frames on `stream`; use_graph != 0 captures the frame once
 * into a hipGraph (replacing the per-launch loop of solver.py:426-442). */
int xpbd_frame3_run(void* handle, void* stream, int frames, int use_graph);
int xpbd_frame3_destroy(void* handle);

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
