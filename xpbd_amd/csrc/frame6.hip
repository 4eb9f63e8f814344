// Frame driver of the 6-DOF solver: the fixed per-substep kernel sequence of
// Solver6DOF._record (src/xpbd3d/solver_6dof.py:357-426), issued eagerly or
// replayed from one hipGraph captured per plan (the reference captures a
// CUDA graph, solver_6dof.py:519-529).  Box kernels launch at fixed capacity
// and read the live pair / point counts from device memory, so the graph
// topology does not depend on the frame.
#include <new>

#include "common.h"
#include "rigid_sub.h"
#include "xpbd.h"

namespace {

constexpr int kRing = 64;  // frames whose status the host has not read yet, at most

struct Frame6 {
  xpbd_rigid_desc d;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t capture_stream = nullptr;
  // broad phase inside the frame (two-launch substep only): 0 none (the
  // caller supplies the pairs), 1 LBVH (leaf < 0: its default), 2 hash grid
  int bp_kind = 0;
  void* bp_ws = nullptr;
  int bp_leaf = -1;
  float bp_rsmax = 0.0f, bp_cell = 0.0f, bp_dt = 0.0f;
  // per-frame status: host-mapped ring of (pair count, frame + 1), the
  // device's frame counter, and what the host has read of it
  int* ring = nullptr;
  int* ring_dev = nullptr;
  int* frames_dev = nullptr;
  long long launched = 0, checked = 0;
  long long bad_frame = -1;
  int bad_count = 0;
};

// Per substep five launches: apply_velocity of the previous substep fused
// with integrate; the manifold; per iteration one launch over every row kind
// (joints, floor corners, box points) and one apply (with finalize on the
// last iteration); one velocity pass over floor bodies and box points.  The
// frame's last apply_velocity is its own launch.  Order of operations as
// Solver6DOF._record (solver_6dof.py:357-426).
int issue(Frame6* f, hipStream_t s) {
  const xpbd_rigid_desc& d = f->d;
  const int n = d.n;
  if (n <= 0) return 0;
  const int maxp = 8 * d.cap;
  const int cap_v = d.max_speed > 0.0f;
  int rc;
  for (int sub = 0; sub < d.substeps; ++sub) {
    rc = xpbd6_begin_substep(s, d.x, d.q, d.v, d.w, d.x_prev, d.q_prev, d.inv_I, d.lam_floor,
                             d.pt_count, d.dv, d.dw, d.max_speed, d.max_omega, cap_v, sub > 0,
                             d.gx, d.gy, d.gz, d.h, d.gyroscopic, d.gscale, n);
    if (rc) return rc;
    rc = xpbd6_box_manifold(s, d.x, d.q, d.he, d.pairs, d.pair_count, d.man_margin, d.cap,
                            d.m_count, d.pt_idx, d.pt_n, d.pt_oi, d.pt_or, d.lam_pt, d.pt_count);
    if (rc) return rc;
    for (int it = 0; it < d.iterations; ++it) {
      rc = xpbd6_solve_all(s, d.x, d.q, d.x_prev, d.q_prev, d.inv_I, d.he, d.j_idx, d.j_a, d.j_b,
                           d.lam_j, d.h, it == 0, d.nj, d.floor_y, d.friction, d.lam_floor, n,
                           d.pt_idx, d.pt_n, d.pt_oi, d.pt_or, d.lam_pt, d.pt_count, maxp, d.dx,
                           d.drot);
      if (rc) return rc;
      rc = xpbd6_apply_finalize(s, d.x, d.q, d.dx, d.drot, d.relax, it + 1 == d.iterations,
                                d.x_prev, d.q_prev, d.v, d.w, d.h, d.lin_damp, d.ang_damp, n);
      if (rc) return rc;
    }
    if (d.iterations == 0) {
      rc = xpbd6_finalize_velocity(s, d.x, d.q, d.x_prev, d.q_prev, d.v, d.w, d.h, d.lin_damp,
                                   d.ang_damp, n);
      if (rc) return rc;
    }
    rc = xpbd6_velocity_all(s, d.x, d.q, d.inv_I, d.he, d.friction, d.restitution, d.lam_floor,
                            d.h, d.v, d.w, n, d.pt_idx, d.pt_n, d.pt_oi, d.pt_or, d.lam_pt,
                            d.pt_count, maxp, d.dv, d.dw);
    if (rc) return rc;
  }
  return xpbd6_apply_velocity(s, d.x, d.v, d.w, d.dv, d.dw, d.max_speed, d.max_omega, cap_v, n);
}

// Two launches per substep (desc.aux set, iterations >= 1): the contacts
// launch (apply_velocity + integrate folded in, first iteration solved with
// the manifold), for each further iteration apply_jacobi and one solve
// launch over every row kind on the predicted pose, then the velocity launch
// (last apply_jacobi and finalize_velocity folded in); one launch ends the
// frame.  Same operations in the same order as issue().
int issue_two_launch(Frame6* f, hipStream_t s) {
  const xpbd_rigid_desc& d = f->d;
  const int n = d.n;
  if (n <= 0) return 0;
  const int maxp = 8 * d.cap;
  int rc;
  for (int sub = 0; sub < d.substeps; ++sub) {
    rc = xpbd6_substep_contacts(s, &d, sub);
    if (rc) return rc;
    const SubBufs b = sub_bufs(d, sub);
    float* XPR = (float*)b.XPR;
    float* QPR = (float*)b.QPR;
    for (int it = 1; it < d.iterations; ++it) {
      rc = xpbd6_apply_finalize(s, XPR, QPR, (float*)b.DX, (float*)b.DR, d.relax, 0, d.x_prev,
                                d.q_prev, d.v, d.w, d.h, d.lin_damp, d.ang_damp, n);
      if (rc) return rc;
      rc = xpbd6_solve_all(s, XPR, QPR, d.x_prev, d.q_prev, d.inv_I, d.he, d.j_idx, d.j_a, d.j_b,
                           d.lam_j, d.h, 0, d.nj, d.floor_y, d.friction, d.lam_floor, n, d.pt_idx,
                           d.pt_n, d.pt_oi, d.pt_or, d.lam_pt, b.PTC, maxp, (float*)b.DX,
                           (float*)b.DR);
      if (rc) return rc;
    }
    rc = xpbd6_substep_velocity(s, &d, sub);
    if (rc) return rc;
  }
  return xpbd6_substep_end_report(s, &d, f->ring_dev, kRing, f->frames_dev);
}

bool two_launch(const Frame6* f) { return f->d.aux && f->d.iterations >= 1; }

// The frame's broad phase (kernels only: it records into the graph).
int issue_broadphase(Frame6* f, hipStream_t s) {
  const xpbd_rigid_desc& d = f->d;
  float* X = d.x;
  int* PAIRS = (int*)d.pairs;
  int* PC = (int*)d.pair_count;
  if (f->bp_kind == 1 && f->bp_leaf < 0)
    return xpbd6_broadphase_lbvh(s, f->bp_ws, X, d.q, d.v, d.w, d.he, d.man_margin, f->bp_dt,
                                 d.n, d.cap, PAIRS, PC);
  if (f->bp_kind == 1)
    return xpbd6_broadphase_lbvh_leaf(s, f->bp_ws, X, d.q, d.v, d.w, d.he, d.man_margin,
                                      f->bp_dt, d.n, d.cap, PAIRS, PC, f->bp_leaf);
  return xpbd6_broadphase_hashgrid(s, f->bp_ws, X, d.q, d.v, d.w, d.he, d.man_margin, f->bp_dt,
                                   f->bp_rsmax, f->bp_cell, d.n, d.cap, PAIRS, PC);
}

int issue_frame(Frame6* f, hipStream_t s) {
  if (!two_launch(f)) return issue(f, s);
  if (f->bp_kind) {
    int rc = issue_broadphase(f, s);
    if (rc) return rc;
  }
  return issue_two_launch(f, s);
}

// Reads the status of the frames the device has reported, in order; the
// first frame whose pair count exceeded cap is kept (it and every later
// frame wrote nothing: their broad phase saw the same state).
void scan_status(Frame6* f) {
  if (!f->ring) {
    f->checked = f->launched;
    return;
  }
  while (f->checked < f->launched) {
    const int* slot = f->ring + 2 * (f->checked % kRing);
    const int mark = __atomic_load_n(slot + 1, __ATOMIC_ACQUIRE);
    if (mark != (int)((f->checked + 1) & 0x7fffffff)) break;
    const int count = __atomic_load_n(slot, __ATOMIC_RELAXED);
    if (count > f->d.cap && f->bad_frame < 0) {
      f->bad_frame = f->checked;
      f->bad_count = count;
    }
    ++f->checked;
  }
}

void free_status(Frame6* f) {
  if (f->ring) {
    (void)hipDeviceSynchronize();  // no frame in flight still reports into the ring
    (void)hipHostFree(f->ring);
  }
  if (f->frames_dev) (void)hipFree(f->frames_dev);
  f->ring = f->ring_dev = f->frames_dev = nullptr;
}

void drop_graph(Frame6* f) {
  if (f->exec) (void)hipGraphExecDestroy(f->exec);
  if (f->graph) (void)hipGraphDestroy(f->graph);
  f->exec = nullptr;
  f->graph = nullptr;
}

}  // namespace

extern "C" {

int xpbd_frame6_create(const xpbd_rigid_desc* desc, void** handle) {
  if (!desc || !handle || desc->n < 0 || desc->cap < 0) return (int)hipErrorInvalidValue;
  Frame6* f = new (std::nothrow) Frame6();
  if (!f) return (int)hipErrorOutOfMemory;
  f->d = *desc;
  *handle = f;
  return 0;
}

int xpbd_frame6_set_broadphase(void* handle, int kind, void* ws, int leaf, float rsmax,
                               float cell, float dt) {
  Frame6* f = (Frame6*)handle;
  if (!f || kind < 0 || kind > 2 || (kind && !ws) || !two_launch(f))
    return (int)hipErrorInvalidValue;
  if (f->launched != f->checked) return (int)hipErrorNotReady;
  drop_graph(f);
  f->bp_kind = kind;
  f->bp_ws = ws;
  f->bp_leaf = leaf;
  f->bp_rsmax = rsmax;
  f->bp_cell = cell;
  f->bp_dt = dt;
  if (kind && !f->ring) {
    void* ring = nullptr;
    XPBD_CHECK(hipHostMalloc(&ring, 2 * kRing * sizeof(int),
                             hipHostMallocMapped | hipHostMallocCoherent));
    f->ring = (int*)ring;
    for (int k = 0; k < 2 * kRing; ++k) f->ring[k] = 0;
    void* dev = nullptr;
    hipError_t e = hipHostGetDevicePointer(&dev, ring, 0);
    if (e == hipSuccess) e = hipMalloc((void**)&f->frames_dev, sizeof(int));
    if (e == hipSuccess) e = hipMemset(f->frames_dev, 0, sizeof(int));
    if (e != hipSuccess) {
      free_status(f);
      return (int)e;
    }
    f->ring_dev = (int*)dev;
  }
  return 0;
}

int xpbd_frame6_status(void* handle, int wait, long long* out) {
  Frame6* f = (Frame6*)handle;
  if (!f || !out) return (int)hipErrorInvalidValue;
  // the frames may have run on any stream (a caller's side stream, say)
  if (wait) XPBD_CHECK(hipDeviceSynchronize());
  scan_status(f);
  if (wait && f->checked != f->launched) return (int)hipErrorUnknown;  // a frame never reported
  out[0] = f->launched - f->checked;
  out[1] = f->bad_frame;
  out[2] = f->bad_count;
  out[3] = f->bad_frame < 0 ? 0 : f->launched - f->bad_frame;
  return 0;
}

int xpbd_frame6_run(void* handle, void* stream, int frames, int use_graph) {
  Frame6* f = (Frame6*)handle;
  if (!f || frames < 0) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  if (f->ring && frames > kRing - 1) return (int)hipErrorInvalidValue;
  if (f->ring && f->launched + frames - f->checked > kRing - 1) {
    // the ring would overwrite a status the host has not read: wait for it
    XPBD_CHECK(hipStreamSynchronize(s));
    scan_status(f);
  }
  if (!use_graph) {
    for (int k = 0; k < frames; ++k) {
      int rc = issue_frame(f, s);
      if (rc) return rc;
      ++f->launched;
    }
    return 0;
  }
  if (!f->exec) {
    if (!f->capture_stream) XPBD_CHECK(hipStreamCreateWithFlags(&f->capture_stream, hipStreamNonBlocking));
    XPBD_CHECK(hipStreamBeginCapture(f->capture_stream, hipStreamCaptureModeThreadLocal));
    int rc = issue_frame(f, f->capture_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(f->capture_stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    XPBD_CHECK(e);
    f->graph = g;
    e = hipGraphInstantiate(&f->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      drop_graph(f);
      return (int)e;
    }
  }
  for (int k = 0; k < frames; ++k) {
    XPBD_CHECK(hipGraphLaunch(f->exec, s));
    ++f->launched;
  }
  return 0;
}

int xpbd_frame6_destroy(void* handle) {
  Frame6* f = (Frame6*)handle;
  if (!f) return 0;
  drop_graph(f);
  if (f->capture_stream) (void)hipStreamDestroy(f->capture_stream);
  free_status(f);
  delete f;
  return 0;
}

}  // extern "C"
