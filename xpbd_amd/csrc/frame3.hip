// Frame driver for the 3-DOF solver: issues one frame of the substep loop
// (src/xpbd3d/solver.py:412-467) on a stream, either eagerly or as one
// replay of a hipGraph captured once per plan.
//
// Launch sequence per frame (S substeps, I iterations):
//   integrate(s=0)
//   for s: for it: jacobi: project(first=it==0) [+ apply if it < I-1]
//                  gs:     for color: solve_color(first=it==0)
//          particles_pass(apply? + finalize + cap? + integrate(s+1)?)
// The lambda reset of Alg. 1 line 4 is folded into the first iteration's
// projection (it reads lambda = 0 and writes every row), and apply of the
// last iteration, finalize, cap and the next integrate share one streaming
// pass: both are exact rewrites of the reference sequence.
#include <cmath>
#include <cstring>
#include <new>
#include <utility>
#include <vector>

#include "common.h"
#include "xpbd.h"

namespace {

enum : int { F_APPLY = 1, F_FINALIZE = 2, F_CAP = 4, F_INTEGRATE = 8 };

struct Frame3 {
  xpbd_particles_desc d;
  xpbd_tile_compact cx;  // copy of *desc.compact (d.compact points here)
  std::vector<int> color_starts;
  // captured frames: [0] plain, [1] with the judge of the learned start
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  hipStream_t capture_stream = nullptr;
  long long frame_no = 0;  // frames run since the plan was created
};

// Fused-tile Jacobi frame.
//
// With fuse_first, each substep is `iterations` k_jacobi_tile launches only
// (plus, with device-counted rows, their projection before each): the first
// has integrate folded in (xpbd_jacobi_tile_first, and
// xpbd_solve_jacobi_dyn_first for the device rows), the last runs finalize +
// cap.  With fuse_first = 2 only the frame's first substep does so (or, with
// one iteration, runs an integrate pass in place); the last launch of every
// other substep also integrates the next one (xpbd_jacobi_tile_next), whose
// first launch is then a plain one with first = 1.  Launches
// ping-pong between X and X_alt; `loc` tracks which one holds the state, and
// a frame that ends in X_alt copies it back.
// With the two-iteration tables (pair_p1; no device rows, no learned start)
// each substep is one xpbd_jacobi_tile_pair launch, the frame's first with
// its integrate folded in (predict) unless it is also the last.
//
// Otherwise (contact rows projected by a separate launch before each
// iteration, or fuse_first off) a substep is one integrate pass (X -> X or
// X_alt, chosen so that after the substep's launches the state is back in X)
// and the plain launches.
//
// With the learned warm start (ws_image) every substep first runs
// xpbd_jacobi_tile_warm (integrate folded in on the fused path), preceded
// on every ws_every-th substep by the network pass xpbd_ws_net (the other
// substeps reuse its outputs), and then all `iterations` plain launches with
// first = 0.  Multipliers ping-pong as the
// positions do: a substep's first launch writes lam_alt.
// Buffers of one run of the fused substep: the position and multiplier
// ping-pong pairs, x_prev, the network outputs and the device-row sums and
// multipliers.  `src0` (optional) is read by the first launch instead of
// x[loc] (the judge starts from the solver's state without writing it).
struct TileRun {
  float* x[2];
  float* lam[2];
  float* xp;      // x_prev the substep reads (and, without pair launches, writes)
  float* xp_alt;  // the other x_prev buffer of the pair launches
  unsigned short* y;
  float* acc;
  float* dlam;
  const float* src0;
  int loc = 0;   // x[loc] holds the state
  int li = 0;    // lam[li] holds the multipliers after the last launch
};

// The plain tile launches of a plan, over the compact tables when the plan
// has them (identical results).
int tile_launch(const xpbd_particles_desc& d, hipStream_t s, const float* xin, float* xout,
                float* xp, const float* lin, float* lout, int first, int finalize, float* acc) {
  if (d.compact)
    return xpbd_jacobi_tile_cx(s, xin, xout, xp, d.v, d.rows, d.anchors, lin, lout,
                               d.tile_row_start, d.halo_start, d.compact, d.h, first, d.tile_size,
                               d.n_particles, d.relax, finalize, d.damping, d.max_speed, acc);
  return xpbd_jacobi_tile(s, xin, xout, xp, d.v, d.rows, d.anchors, lin, lout, d.tile_row_start,
                          d.halo_start, d.halo_rows, d.h, first, d.tile_size, d.n_particles,
                          d.relax, finalize, d.damping, d.max_speed, acc);
}

int tile_launch_first(const xpbd_particles_desc& d, hipStream_t s, const float* xin, float* xout,
                      float* xp, float* lout, float* acc) {
  if (d.compact)
    return xpbd_jacobi_tile_first_cx(s, xin, xout, xp, d.v, d.rows, d.anchors, lout,
                                     d.tile_row_start, d.halo_start, d.compact, d.h, d.gx, d.gy,
                                     d.gz, d.gscale, d.tile_size, d.n_particles, d.relax, acc);
  return xpbd_jacobi_tile_first(s, xin, xout, xp, d.v, d.rows, d.anchors, lout, d.tile_row_start,
                                d.halo_start, d.halo_rows, d.h, d.gx, d.gy, d.gz, d.gscale,
                                d.tile_size, d.n_particles, d.relax, acc);
}

int tile_launch_next(const xpbd_particles_desc& d, hipStream_t s, const float* xin, float* xout,
                     float* xp, const float* lin, float* lout, int first, float* acc) {
  if (d.compact)
    return xpbd_jacobi_tile_next_cx(s, xin, xout, xp, d.rows, d.anchors, lin, lout,
                                    d.tile_row_start, d.halo_start, d.compact, d.h, first, d.gx,
                                    d.gy, d.gz, d.gscale, d.tile_size, d.n_particles, d.relax,
                                    d.damping, d.max_speed, acc);
  return xpbd_jacobi_tile_next(s, xin, xout, xp, d.rows, d.anchors, lin, lout, d.tile_row_start,
                               d.halo_start, d.halo_rows, d.h, first, d.gx, d.gy, d.gz, d.gscale,
                               d.tile_size, d.n_particles, d.relax, d.damping, d.max_speed, acc);
}

// One substep of the fused tile plan (see issue_frame_tile) on stream s;
// invert != 0 runs the start the solver is not using (the judge).
int issue_fused_substep(const xpbd_particles_desc& d, hipStream_t s, TileRun& r, int sub,
                        int invert) {
  const int n = d.n_particles;
  const bool have_rows = (d.n_rows > 0 || d.dyn_cap > 0) && d.iterations > 0;
  const int it_n = have_rows ? d.iterations : 0;
  const bool dyn = d.dyn_cap > 0 && d.dyn_rows && d.dyn_count;
  const bool warm = d.ws_image != nullptr && d.n_rows > 0 && it_n > 0;
  const int ws_every = d.ws_every > 1 ? d.ws_every : 1;
  const bool next = d.fuse_first >= 2 && it_n >= 1;
  int rc;
  // predicted: the previous substep's last launch already integrated
  const bool predicted = next && sub > 0;
  const bool integrate_next = next && sub + 1 < d.substeps;
  float* const* bufs = r.x;
  float* const* lams = r.lam;
  auto in = [&]() -> const float* { return r.src0 ? r.src0 : bufs[r.loc]; };
  r.li = 0;
  int it = 0;
  if (warm) {
    if (sub % ws_every == 0) {
      rc = xpbd_ws_net(s, in(), r.xp, d.v, d.rows, d.anchors, d.ws_image, r.y, d.h, d.gx, d.gy,
                       d.gz, d.gscale, d.n_rows, predicted ? 0 : 1, d.ws_on, invert);
      if (rc) return rc;
    }
    const bool last = !integrate_next;
    if (it_n == 1 && d.pair_p1 && r.xp_alt && !dyn && next && !(last && !predicted)) {
      // the learned start and the substep's iteration in one launch
      float* xp_out = last ? d.x_prev : r.xp_alt;
      rc = xpbd_jacobi_tile_pair_warm(
          s, in(), bufs[r.loc ^ 1], r.xp, xp_out, d.v, d.rows, d.anchors, lams[0],
          d.pair_ring_start, d.pair_ring1_n, d.pair_ring_idx, d.pair_p1_start, d.pair_p1,
          d.pair_p2_start, d.pair_p2, d.h, d.tile_size, n, d.relax, d.damping, d.max_speed, d.gx,
          d.gy, d.gz, d.gscale, d.pair_lmax, d.pair_amax, d.pair_k1max, last ? 1 : 0,
          predicted ? 0 : 1, r.y, d.ws_on, invert);
      if (rc) return rc;
      r.src0 = nullptr;
      r.loc ^= 1;
      r.li = 0;
      if (!last) std::swap(r.xp, r.xp_alt);
      return 0;
    }
    if (d.compact)
      rc = xpbd_jacobi_tile_warm_cx(s, in(), bufs[r.loc ^ 1], r.xp, d.v, d.rows, d.anchors,
                                    lams[1], d.tile_row_start, d.halo_start, d.compact, r.y, d.h,
                                    d.gx, d.gy, d.gz, d.gscale, d.tile_size, n, d.relax,
                                    predicted ? 0 : 1, d.ws_on, invert);
    else
      rc = xpbd_jacobi_tile_warm(s, in(), bufs[r.loc ^ 1], r.xp, d.v, d.rows, d.anchors,
                                 lams[1], d.tile_row_start, d.halo_start, d.halo_rows, r.y, d.h,
                                 d.gx, d.gy, d.gz, d.gscale, d.tile_size, n, d.relax,
                                 predicted ? 0 : 1, d.ws_on, invert);
    if (rc) return rc;
    r.src0 = nullptr;
    r.loc ^= 1;
    r.li = 1;
  } else if (!predicted && it_n == 2 && d.pair_p1 && r.xp_alt && !dyn && integrate_next) {
    // the frame's first substep as one pair launch with the integrate folded
    // in (x_prev = the start positions, written to the other x_prev buffer)
    rc = xpbd_jacobi_tile_pair(s, in(), bufs[r.loc ^ 1], r.xp, r.xp_alt, d.v, d.rows, d.anchors,
                               lams[0], d.pair_ring_start, d.pair_ring1_n, d.pair_ring_idx,
                               d.pair_p1_start, d.pair_p1, d.pair_p2_start, d.pair_p2, d.h,
                               d.tile_size, n, d.relax, d.damping, d.max_speed, d.gx, d.gy, d.gz,
                               d.gscale, d.pair_lmax, d.pair_amax, d.pair_k1max, 0, 1);
    if (rc) return rc;
    r.src0 = nullptr;
    r.loc ^= 1;
    r.li = 0;
    std::swap(r.xp, r.xp_alt);
    return 0;
  } else if (!predicted && it_n >= 2) {
    if (dyn) {  // device-counted rows, first iteration on the predicted positions
      rc = xpbd_solve_jacobi_dyn_first(s, in(), d.v, d.dyn_rows, r.dlam, r.acc, d.h, d.gx, d.gy,
                                       d.gz, d.gscale, d.dyn_count, d.dyn_cap);
      if (rc) return rc;
    }
    rc = tile_launch_first(d, s, in(), bufs[r.loc ^ 1], r.xp, lams[1], r.acc);
    if (rc) return rc;
    r.src0 = nullptr;
    r.loc ^= 1;
    r.li = 1;
    it = 1;
  } else if (!predicted) {  // one iteration: integrate in place, then the chain
    rc = xpbd_integrate_to(s, bufs[r.loc], bufs[r.loc], d.v, r.xp, d.gx, d.gy, d.gz, d.h,
                           d.gscale, n);
    if (rc) return rc;
  } else if (it_n == 2 && d.pair_p1 && r.xp_alt && !dyn) {
    // both iterations (and the finalize / next integrate) in one launch;
    // x_prev goes to the other buffer (ring rows read this one), and the
    // last substep copies it back into d.x_prev
    const bool last = !integrate_next;
    float* xp_out = last ? d.x_prev : r.xp_alt;
    rc = xpbd_jacobi_tile_pair(s, bufs[r.loc], bufs[r.loc ^ 1], r.xp, xp_out, d.v, d.rows,
                               d.anchors, lams[0], d.pair_ring_start, d.pair_ring1_n,
                               d.pair_ring_idx, d.pair_p1_start, d.pair_p1, d.pair_p2_start,
                               d.pair_p2, d.h, d.tile_size, n, d.relax, d.damping, d.max_speed,
                               d.gx, d.gy, d.gz, d.gscale, d.pair_lmax, d.pair_amax,
                               d.pair_k1max, last ? 1 : 0, 0);
    if (rc) return rc;
    r.loc ^= 1;
    r.li = 0;
    if (!last) std::swap(r.xp, r.xp_alt);
    return 0;
  }
  for (; it < it_n; ++it) {
    const int first = (!warm && it == 0) ? 1 : 0;
    if (dyn) {  // device rows start from zero (the warm start covers `rows` only)
      rc = xpbd_solve_jacobi_dyn(s, bufs[r.loc], r.xp, d.dyn_rows, r.dlam, r.acc, d.h,
                                 it == 0 ? 1 : 0, d.dyn_count, d.dyn_cap);
      if (rc) return rc;
    }
    if (it + 1 == it_n && integrate_next) {
      rc = tile_launch_next(d, s, bufs[r.loc], bufs[r.loc ^ 1], r.xp, lams[r.li],
                            lams[r.li ^ 1], first, r.acc);
    } else {
      rc = tile_launch(d, s, bufs[r.loc], bufs[r.loc ^ 1], r.xp, lams[r.li], lams[r.li ^ 1],
                       first, it + 1 == it_n ? 1 : 0, r.acc);
    }
    if (rc) return rc;
    r.loc ^= 1;
    r.li ^= 1;
  }
  return 0;
}

// Fused-tile Jacobi frame.
//
// With fuse_first, each substep is `iterations` k_jacobi_tile launches only
// (plus, with device-counted rows, their projection before each): the first
// has integrate folded in (xpbd_jacobi_tile_first, and
// xpbd_solve_jacobi_dyn_first for the device rows), the last runs finalize +
// cap.  With fuse_first = 2 only the frame's first substep does so (or, with
// one iteration, runs an integrate pass in place); the last launch of every
// other substep also integrates the next one (xpbd_jacobi_tile_next), whose
// first launch is then a plain one with first = 1.  Launches
// ping-pong between X and X_alt; `loc` tracks which one holds the state, and
// a frame that ends in X_alt copies it back.
// With the two-iteration tables (pair_p1; no device rows, no learned start)
// each substep is one xpbd_jacobi_tile_pair launch, the frame's first with
// its integrate folded in (predict) unless it is also the last.
//
// Otherwise (contact rows projected by a separate launch before each
// iteration, or fuse_first off) a substep is one integrate pass (X -> X or
// X_alt, chosen so that after the substep's launches the state is back in X)
// and the plain launches.
//
// With the learned warm start (ws_image) every substep first runs
// xpbd_jacobi_tile_warm (integrate folded in on the fused path), preceded
// on every ws_every-th substep by the network pass xpbd_ws_net (the other
// substeps reuse its outputs), and then all `iterations` plain launches with
// first = 0.  Multipliers ping-pong as the positions do: a substep's first
// launch writes lam_alt.
//
// Judged frames (ws_judge, fuse_first 2, >= 2 substeps): the frame's first
// substep is first run with the other start (the zero start while the
// learned one is on, and the reverse) from the same state into the judge
// buffers, then as usual; both substeps' squared residuals (positions after
// the iterations: x_prev, which the substep's last launch sets to them) are
// summed, and xpbd_ws_decide keeps the learned start for the rest of the
// frame and the following ones only if its residual was no larger.  The
// judge runs on the same stream: on small scenes a second stream's
// synchronisation cost more than the judge's launches, on large ones the
// launches fill the GPU anyway.
int issue_frame_tile(Frame3* f, hipStream_t s, bool judge) {
  const xpbd_particles_desc& d = f->d;
  const int n = d.n_particles;
  if (n <= 0) return 0;
  const bool have_rows = (d.n_rows > 0 || d.dyn_cap > 0) && d.iterations > 0;
  const int it_n = have_rows ? d.iterations : 0;
  const bool dyn = d.dyn_cap > 0 && d.dyn_rows && d.dyn_count;
  const bool warm = d.ws_image != nullptr && d.n_rows > 0 && it_n > 0;
  float* bufs[2] = {d.x, d.x_alt};
  float* lams[2] = {d.lam, d.lam_alt};
  int rc;
  const bool next = d.fuse_first >= 2 && it_n >= 1;
  if (d.fuse_first && d.n_rows > 0 && (it_n >= 2 || (warm && it_n >= 1) || next)) {
    TileRun r{{d.x, d.x_alt}, {d.lam, d.lam_alt}, d.x_prev,
              d.x_prev_alt, d.ws_y, dyn ? d.acc : nullptr, d.dyn_lam, nullptr};
    judge = judge && warm && d.ws_judge > 0;
    for (int sub = 0; sub < d.substeps; ++sub) {
      if (judge && sub == 0) {
        TileRun jr{{d.ws_jx, d.ws_jx_alt}, {d.ws_jlam, d.ws_jlam_alt}, d.ws_jxp, d.ws_jxp_alt,
                   d.ws_jy, dyn ? d.ws_jacc : nullptr, d.ws_jdlam, d.x};
        rc = issue_fused_substep(d, s, jr, 0, 1);
        if (rc) return rc;
        rc = xpbd_ws_residual(s, jr.xp, d.rows, d.anchors, jr.lam[jr.li], d.h, d.ws_rsum + 1,
                              d.n_rows);
        if (rc) return rc;
        rc = issue_fused_substep(d, s, r, 0, 0);
        if (rc) return rc;
        rc = xpbd_ws_residual(s, r.xp, d.rows, d.anchors, r.lam[r.li], d.h, d.ws_rsum, d.n_rows);
        if (rc) return rc;
        rc = xpbd_ws_decide(s, d.ws_rsum, d.ws_on, d.ws_stats);
        if (rc) return rc;
        continue;
      }
      rc = issue_fused_substep(d, s, r, sub, 0);
      if (rc) return rc;
    }
    if (r.loc) {
      XPBD_CHECK(hipMemcpyAsync(d.x, d.x_alt, sizeof(float) * 4 * (size_t)n,
                                hipMemcpyDeviceToDevice, s));
    }
    return 0;
  }
  const int ws_every = d.ws_every > 1 ? d.ws_every : 1;
  const int launches = it_n + (warm ? 1 : 0);
  for (int sub = 0; sub < d.substeps; ++sub) {
    int cur = launches & 1;
    rc = xpbd_integrate_to(s, d.x, bufs[cur], d.v, d.x_prev, d.gx, d.gy, d.gz, d.h, d.gscale, n);
    if (rc) return rc;
    int li = 0;
    if (warm) {
      if (sub % ws_every == 0) {
        rc = xpbd_ws_net(s, bufs[cur], d.x_prev, d.v, d.rows, d.anchors, d.ws_image, d.ws_y,
                         d.h, 0.0f, 0.0f, 0.0f, nullptr, d.n_rows, 0, d.ws_on, 0);
        if (rc) return rc;
      }
      if (d.compact)
        rc = xpbd_jacobi_tile_warm_cx(s, bufs[cur], bufs[cur ^ 1], d.x_prev, d.v, d.rows,
                                      d.anchors, lams[1], d.tile_row_start, d.halo_start,
                                      d.compact, d.ws_y, d.h, 0.0f, 0.0f, 0.0f, nullptr,
                                      d.tile_size, n, d.relax, 0, d.ws_on, 0);
      else
        rc = xpbd_jacobi_tile_warm(s, bufs[cur], bufs[cur ^ 1], d.x_prev, d.v, d.rows,
                                   d.anchors, lams[1], d.tile_row_start, d.halo_start,
                                   d.halo_rows, d.ws_y, d.h, 0.0f, 0.0f, 0.0f, nullptr,
                                   d.tile_size, n, d.relax, 0, d.ws_on, 0);
      if (rc) return rc;
      cur ^= 1;
      li = 1;
    }
    for (int it = 0; it < it_n; ++it) {
      const int nxt = cur ^ 1;
      if (dyn) {
        rc = xpbd_solve_jacobi_dyn(s, bufs[cur], d.x_prev, d.dyn_rows, d.dyn_lam, d.acc, d.h,
                                   it == 0, d.dyn_count, d.dyn_cap);
        if (rc) return rc;
      }
      rc = tile_launch(d, s, bufs[cur], bufs[nxt], d.x_prev, lams[li], lams[li ^ 1],
                       it == 0 && !warm ? 1 : 0, it + 1 == it_n ? 1 : 0, dyn ? d.acc : nullptr);
      if (rc) return rc;
      cur = nxt;
      li ^= 1;
    }
    if (it_n == 0) {
      int flags = F_FINALIZE;
      if (d.max_speed > 0.0f && std::isfinite(d.max_speed)) flags |= F_CAP;
      rc = xpbd_particles_pass(s, d.x, d.v, d.x_prev, d.acc, flags, d.relax, d.h, d.damping,
                               d.max_speed, d.gx, d.gy, d.gz, n);
      if (rc) return rc;
    }
  }
  return 0;
}

int issue_frame(Frame3* f, hipStream_t s, bool judge) {
  const xpbd_particles_desc& d = f->d;
  if (d.mode == XPBD_MODE_JACOBI_TILE) return issue_frame_tile(f, s, judge);
  const int n = d.n_particles;
  if (n <= 0) return 0;
  const bool have_rows = d.n_rows > 0 && d.iterations > 0;
  const bool jacobi = d.mode == XPBD_MODE_JACOBI;
  int cap = (d.max_speed > 0.0f && std::isfinite(d.max_speed)) ? F_CAP : 0;
  int rc = xpbd_particles_pass(s, d.x, d.v, d.x_prev, d.acc, F_INTEGRATE, d.relax, d.h, d.damping,
                               d.max_speed, d.gx, d.gy, d.gz, n);
  if (rc) return rc;
  for (int sub = 0; sub < d.substeps; ++sub) {
    if (have_rows) {
      for (int it = 0; it < d.iterations; ++it) {
        const int first = it == 0;
        if (jacobi) {
          rc = xpbd_solve_jacobi(s, d.x, d.x_prev, d.rows, d.anchors, d.lam, d.acc, d.h, first,
                                 d.n_rows, d.rows_per_thread, d.window);
          if (rc) return rc;
          if (it + 1 < d.iterations) {
            rc = xpbd_apply_jacobi(s, d.x, d.acc, d.relax, n);
            if (rc) return rc;
          }
        } else {
          for (int c = 0; c + 1 < (int)f->color_starts.size(); ++c) {
            int start = f->color_starts[c];
            int count = f->color_starts[c + 1] - start;
            rc = xpbd_solve_color(s, d.x, d.x_prev, d.rows, d.anchors, d.lam, d.h, first, start,
                                  count);
            if (rc) return rc;
          }
        }
      }
    }
    int flags = F_FINALIZE | cap;
    if (have_rows && jacobi) flags |= F_APPLY;
    if (sub + 1 < d.substeps) flags |= F_INTEGRATE;
    rc = xpbd_particles_pass(s, d.x, d.v, d.x_prev, d.acc, flags, d.relax, d.h, d.damping,
                             d.max_speed, d.gx, d.gy, d.gz, n);
    if (rc) return rc;
  }
  return 0;
}

void drop_graph(Frame3* f) {
  for (int k = 0; k < 2; ++k) {
    if (f->exec[k]) (void)hipGraphExecDestroy(f->exec[k]);
    if (f->graph[k]) (void)hipGraphDestroy(f->graph[k]);
    f->exec[k] = nullptr;
    f->graph[k] = nullptr;
  }
}

// Whether the plan's next frame is a judged one (every ws_judge-th frame).
bool judged_frame(const Frame3* f) {
  return f->d.ws_judge > 0 && f->frame_no % f->d.ws_judge == 0;
}

int capture(Frame3* f, int k) {
  if (!f->capture_stream) XPBD_CHECK(hipStreamCreateWithFlags(&f->capture_stream, hipStreamNonBlocking));
  XPBD_CHECK(hipStreamBeginCapture(f->capture_stream, hipStreamCaptureModeThreadLocal));
  int rc = issue_frame(f, f->capture_stream, k == 1);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(f->capture_stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  XPBD_CHECK(e);
  f->graph[k] = g;
  e = hipGraphInstantiate(&f->exec[k], g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    f->graph[k] = nullptr;
    f->exec[k] = nullptr;
    return (int)e;
  }
  return 0;
}

}  // namespace

extern "C" {

int xpbd_frame3_create(const xpbd_particles_desc* desc, void** handle) {
  if (!desc || !handle) return (int)hipErrorInvalidValue;
  if (desc->mode != XPBD_MODE_GS && desc->mode != XPBD_MODE_JACOBI &&
      desc->mode != XPBD_MODE_JACOBI_TILE)
    return (int)hipErrorInvalidValue;
  if (desc->mode == XPBD_MODE_JACOBI_TILE && (desc->n_rows > 0 || desc->dyn_cap > 0) &&
      (!desc->x_alt || !desc->lam_alt || !desc->tile_row_start || !desc->halo_start ||
       desc->tile_size < 1 || desc->tile_size > XPBD_MAX_TILE))
    return (int)hipErrorInvalidValue;
  if (desc->ws_image &&
      (desc->mode != XPBD_MODE_JACOBI_TILE || !desc->ws_y))
    return (int)hipErrorInvalidValue;  // learned warm start: tile mode
  if (desc->mode != XPBD_MODE_JACOBI_TILE && (desc->dyn_cap > 0 || desc->gscale))
    return (int)hipErrorInvalidValue;  // device-counted rows / gravity scale: tile mode only
  if (desc->mode == XPBD_MODE_JACOBI_TILE && desc->dyn_cap > 0 &&
      (!desc->acc || !desc->dyn_rows || !desc->dyn_lam || !desc->dyn_count))
    return (int)hipErrorInvalidValue;
  if (desc->mode == XPBD_MODE_GS && desc->n_rows > 0 && (!desc->color_starts || desc->num_colors < 1))
    return (int)hipErrorInvalidValue;
  if (desc->mode == XPBD_MODE_JACOBI && desc->n_rows > 0 && !desc->acc) return (int)hipErrorInvalidValue;
  if (desc->block &&
      (desc->mode == XPBD_MODE_JACOBI_TILE || desc->ws_image ||
       desc->n_particles > xpbd_frame_block_capacity(desc->mode == XPBD_MODE_GS) ||
       (desc->mode == XPBD_MODE_GS && desc->n_rows > 0 && !desc->color_starts_dev)))
    return (int)hipErrorInvalidValue;
  if (desc->ws_judge < 0) return (int)hipErrorInvalidValue;
  const bool judge = desc->ws_judge > 0;
  if (judge &&
      (!desc->ws_image || desc->mode != XPBD_MODE_JACOBI_TILE || desc->block ||
       desc->fuse_first < 2 || desc->substeps < 2 || desc->iterations < 1 || desc->n_rows <= 0 ||
       !desc->ws_on || !desc->ws_rsum || !desc->ws_jx || !desc->ws_jx_alt || !desc->ws_jxp ||
       !desc->ws_jlam || !desc->ws_jlam_alt || !desc->ws_jy ||
       (desc->dyn_cap > 0 && (!desc->ws_jacc || !desc->ws_jdlam))))
    return (int)hipErrorInvalidValue;  // the judge runs in the fused plan's shape only
  if (desc->compact &&
      (desc->mode != XPBD_MODE_JACOBI_TILE || desc->tile_size < 64 || !desc->compact->dist_end ||
       !desc->compact->ci ||
       !desc->compact->cc || !desc->compact->halo2 ||
       (desc->compact->pf &&
        (!desc->compact->gen_end || !desc->compact->pflam || desc->pair_p1))))
    return (int)hipErrorInvalidValue;  // per-particle floors: no kernel reading lam's FLOOR rows
  Frame3* f = new (std::nothrow) Frame3();
  if (!f) return (int)hipErrorOutOfMemory;
  f->d = *desc;
  if (desc->compact) {
    f->cx = *desc->compact;
    f->d.compact = &f->cx;
  }
  if (desc->mode == XPBD_MODE_GS && desc->n_rows > 0)
    f->color_starts.assign(desc->color_starts, desc->color_starts + desc->num_colors + 1);
  f->d.color_starts = nullptr;
  *handle = f;
  return 0;
}

int xpbd_frame3_run(void* handle, void* stream, int frames, int use_graph) {
  Frame3* f = (Frame3*)handle;
  if (!f || frames < 0) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  if (f->d.block) {  // whole frames in one workgroup: one launch, no graph needed
    const xpbd_particles_desc& d = f->d;
    return xpbd_frame_block(s, d.x, d.v, d.x_prev, d.rows, d.anchors, d.lam, d.color_starts_dev,
                            d.num_colors, d.n_particles, d.n_rows, d.substeps, d.iterations, d.h,
                            d.gx, d.gy, d.gz, d.damping, d.relax, d.max_speed,
                            d.mode == XPBD_MODE_GS ? 1 : 0, frames);
  }
  for (int k = 0; k < frames; ++k) {
    const int j = judged_frame(f) ? 1 : 0;
    if (!use_graph) {
      int rc = issue_frame(f, s, j == 1);
      if (rc) return rc;
    } else {
      if (!f->exec[j]) {
        int rc = capture(f, j);
        if (rc) return rc;
      }
      XPBD_CHECK(hipGraphLaunch(f->exec[j], s));
    }
    ++f->frame_no;
  }
  return 0;
}

int xpbd_frame3_destroy(void* handle) {
  Frame3* f = (Frame3*)handle;
  if (!f) return 0;
  drop_graph(f);
  if (f->capture_stream) (void)hipStreamDestroy(f->capture_stream);
  delete f;
  return 0;
}

}  // extern "C"
