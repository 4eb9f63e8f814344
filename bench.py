#!/usr/bin/env python3
"""Benchmark of the XPBD substep loop on one MI355X.

Workload (one "step" = one frame of 1/60 s = 15 substeps x 2 iterations):
the 128 x 128 cloth of the benchmark configuration — 16 384 particles,
97 026 DISTANCE rows (structural + both shears + bend, compliance
log-uniform in [1e-8, 1e-3]) and 16 384 FLOOR rows (mu 0.4), Jacobi mode,
run through the public ``Solver`` API (hipGraph replay per frame).

Metric: constraint projections per second = rows x substeps x iterations /
frame time (the reference's definition, examples/benchmark.py:123-130).

Also reported on the same JSON line:
* ``roofline``       — the kernel that dominates the headline frame (the
  two-iteration k_jacobi_tile_pair launch when the plan uses it, else one
  fused k_jacobi_tile iteration): algorithmic bytes per launch / average
  launch time, timed with HIP events on the launch stream, against 8 TB/s.
* ``roofline_scene`` — the same kernel and a full frame on 8 x 1024^2 cloths
  (58.6M rows, a working set far above the caches), built on the device.
* ``cpu_baseline``   — the serial CPU restatement of the reference kernels
  (oracle/) on a bounded number of frames of the same workload, 1 core.

``python bench.py [--gpus 1] [--steps K] [--warmup W]``
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "constraint proj/s (1 GPU); substeps/s at fixed residual, learned vs plain λ"
HBM_PEAK_GBS = 8000.0


def iter_bytes(n_distance: int, n_floor: int, n_particles: int, x_prev: bool) -> int:
    """Algorithmic HBM bytes of one Jacobi iteration (the unit of one
    k_jacobi_tile launch): 24 B per DISTANCE row (a, b, rest, alpha, lambda
    read + write), 20 B per FLOOR row (a, floor_y, mu, lambda read + write),
    and per particle 28 B (x read + write, inverse mass read), + 12 B (x_prev
    read) when the launch reads x_prev: some FLOOR row with friction is
    active (the particle lies below its floor) in the state it runs on."""
    return 24 * n_distance + 20 * n_floor + (40 if x_prev else 28) * n_particles


def time_kernel(fn, reps: int, rounds: int = 3, warmup: int = 3) -> float:
    """Average duration in ms of ``fn`` (which launches on torch's current
    stream), bracketed by HIP events on that stream: the median over
    ``rounds`` windows of ``reps`` back-to-back launches, after ``warmup``
    launches."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    out = []
    for _ in range(rounds):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        out.append(e0.elapsed_time(e1) / reps)
    return sorted(out)[len(out) // 2]


def time_kernel_graph(fn, reps: int, rounds: int = 3) -> float:
    """Average duration in ms of one ``fn`` launch with the host out of the
    loop: ``reps`` back-to-back launches captured once into a hipGraph
    (torch.cuda.CUDAGraph records the library's launches on the capture
    stream) and replayed on torch's current stream between HIP events; the
    median over ``rounds`` replays.  Small kernels launched one by one from
    Python are host-bound (~8 us per ctypes launch)."""
    import torch

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fn()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(reps):
            fn()
    g.replay()
    torch.cuda.synchronize()
    out = []
    for _ in range(rounds):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        g.replay()
        e1.record()
        torch.cuda.synchronize()
        out.append(e0.elapsed_time(e1) / reps)
    del g
    return sorted(out)[len(out) // 2]


def compact_struct(N, cx: dict, pf: bool = True):
    """xpbd_tile_compact over the tensors of ``cx`` (per-particle floors
    unless ``pf`` is False; ``cx["general"]``: some row is read whole)."""
    cs = N.TileCompact()
    for k in ("dist_end", "gen_end", "ci", "cc", "halo2"):
        setattr(cs, k, N.ptr(cx[k]))
    if pf:
        cs.pf, cs.pflam = N.ptr(cx["pf"]), N.ptr(cx["pflam"])
    cs.general = int(cx.get("general", 1) or not pf)
    return cs


def tile_launcher(N, X, X_B, XP, V, R, ANC, LAM, LAM_B, TRS, HS, HR, h, tile, n, cx=None):
    """One k_jacobi_tile launch (a full Jacobi iteration) on fixed buffers
    (ANC None: the scene has no ATTACH rows), over the compact tables when
    ``cx`` (an xpbd_tile_compact) is given, else over the int4 rows."""
    lib = N.lib()
    args = tuple(None if t is None else N.ptr(t)
                 for t in (X, X_B, XP, V, R, ANC, LAM, LAM_B, TRS, HS))
    hr = N.ptr(HR)

    def fn():
        if cx is not None:
            N.check(lib.xpbd_jacobi_tile_cx(N.current_stream(), *args, N.ctypes.addressof(cx), h,
                                            0, tile, n, 1.0, 0, 0.0, 40.0, None),
                    "xpbd_jacobi_tile_cx")
        else:
            N.check(lib.xpbd_jacobi_tile(N.current_stream(), *args, hr, h, 0, tile, n, 1.0, 0,
                                         0.0, 40.0, None), "xpbd_jacobi_tile")
    return fn


def headline_launch(N, s):
    """The launch that dominates the solver's frame plan, for timing alone:
    the two-iteration pair launch when the plan uses it (every substep after
    a frame's first), else one plain iteration (over the plan's compact
    tables when it has them).  Returns (fn, kernel name, iterations per
    launch); the launch rewrites the solver's second buffers only."""
    import numpy as np

    lib = N.lib()
    h = float(np.float32(s.dt / s.substeps))
    ts, hs, hr = s._tile_dev
    if s._plan_pair:
        _, tabs, (lmax, amax, k1max), xp_alt = s._pair_dev
        ptrs = [N.ptr(t) for t in (s.X, s.X_B, s.XP, xp_alt, s.V, s.R, s.ANC, s.LAM)]
        ptrs += [N.ptr(tabs[k]) for k in ("ring_start", "ring1_n", "ring_idx", "p1_start", "p1",
                                          "p2_start", "p2")]

        def fn():
            N.check(lib.xpbd_jacobi_tile_pair(
                N.current_stream(), *ptrs, h, s._tile, s.num_bodies, s.jacobi_relax, s.damping,
                s.max_speed, *s.gravity, None, lmax, amax, k1max, 0, 0), "xpbd_jacobi_tile_pair")
        return fn, "k_jacobi_tile_pair", 2
    fn = tile_launcher(N, s.X, s.X_B, s.XP, s.V, s.R, s.ANC if s._has_attach else None, s.LAM,
                       s.LAM_B, ts, hs, hr, h, s._tile, s.num_bodies, cx=s._cx_struct)
    return fn, "k_jacobi_tile", 1


def floor_rows_active(X, R) -> int:
    """FLOOR rows whose particle lies below its floor (they read x_prev for
    friction), counted on the device."""
    import torch

    t = (R[:, 0] >> 28) & 15
    fl = t == 2
    a = (R[fl, 0] & 0x0FFFFFFF).long()
    y = X[a, 1]
    return int(torch.count_nonzero(y < R[fl, 2].view(torch.float32)))


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), else the platform's."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(h, s, budget_s: float) -> dict:
    from oracle import oracle as orc

    if not orc.LIB.exists():
        orc.build()
    sc = orc.Scene3(h, dt=s.dt, substeps=s.substeps, iterations=s.iterations, gravity=s.gravity,
                    mode=s.solve_mode, damping=s.damping, relax=s.jacobi_relax,
                    max_speed=s.max_speed)
    t = time.perf_counter()
    sc.step(1)
    one = time.perf_counter() - t
    frames = max(1, min(200, int(budget_s / max(one, 1e-6))))
    t = time.perf_counter()
    sc.step(frames)
    dt = time.perf_counter() - t
    proj = s.num_constraints * s.substeps * s.iterations * frames
    return {"value": proj / dt, "unit": "proj/s", "cores": 1, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{frames} frames of the same workload after 1 warm-up frame "
                      f"({dt:.1f} s), serial C restatement (oracle/), "
                      f"host CPU {cpu_model()}, "
                      f"os.cpu_count()={os.cpu_count()}"}


def roofline_scene(N, reps: int, frames: int) -> dict:
    """k_jacobi_tile on 8 x 1024^2 cloths (58.6 M rows, far beyond the
    caches): the sheets hanging (no FLOOR row active, x_prev never read) and
    lying on the floor (every FLOOR row active with friction).  The launch
    is the product's, over the compact tables; the same launch over the
    int4 rows and halo list is timed beside it."""
    import torch

    from xpbd_amd.scenes import packed_cloth_device
    from xpbd_amd.solver import MODE_JACOBI_TILE

    res, sheets = 1024, 8
    h = float(torch.tensor(1.0 / 60.0 / 15.0, dtype=torch.float32))

    def measure(sc, compact=True):
        cs = compact_struct(N, sc["CX"]) if compact else None
        launch = tile_launcher(N, sc["X"], sc["X_B"], sc["XP"], sc["V"], sc["R"], None,
                               sc["LAM"], sc["LAM_B"], sc["TRS"], sc["HS"], sc["HR"], h,
                               sc["tile"], sc["n_particles"], cx=cs)
        return time_kernel_graph(launch, reps)

    resting = packed_cloth_device(res, sheets, "cuda:0", height=-0.002)
    torch.cuda.synchronize()
    n, m, nd = resting["n_particles"], resting["n_rows"], resting["n_distance"]
    rest_active = floor_rows_active(resting["X"], resting["R"])
    rest_ms = measure(resting)
    rest_int4_ms = measure(resting, compact=False)
    rest_bytes = iter_bytes(nd, m - nd, n, rest_active > 0)
    del resting
    torch.cuda.empty_cache()
    sc = packed_cloth_device(res, sheets, "cuda:0")
    torch.cuda.synchronize()
    hang_active = floor_rows_active(sc["X"], sc["R"])
    ms = measure(sc)
    int4_ms = measure(sc, compact=False)
    nbytes = iter_bytes(nd, m - nd, n, hang_active > 0)
    d = N.ParticlesDesc()
    d.x, d.v, d.x_prev, d.acc = (N.ptr(sc["X"]), N.ptr(sc["V"]), N.ptr(sc["XP"]),
                                 N.ptr(sc["ACC"]))
    d.rows, d.anchors, d.lam = N.ptr(sc["R"]), None, N.ptr(sc["LAM"])
    d.n_particles, d.n_rows = n, m
    d.color_starts, d.num_colors = None, 0
    d.mode, d.substeps, d.iterations = MODE_JACOBI_TILE, 15, 2
    d.h, d.gx, d.gy, d.gz = h, 0.0, -9.81, 0.0
    d.damping, d.relax, d.max_speed = 0.0, 1.0, 40.0
    d.x_alt, d.lam_alt = N.ptr(sc["X_B"]), N.ptr(sc["LAM_B"])
    d.tile_row_start, d.halo_start, d.halo_rows = N.ptr(sc["TRS"]), N.ptr(sc["HS"]), N.ptr(sc["HR"])
    d.tile_size = sc["tile"]
    cs = compact_struct(N, sc["CX"])
    d.compact = N.ctypes.addressof(cs)
    # fuse_first 2: integrate folded into the previous substep's last launch
    # (product path), 1: into each substep's first launch, 0: separate pass
    frame = {2: None, 1: None, 0: None}
    for fuse in ((2, 1, 0) if frames > 0 else ()):
        d.fuse_first = fuse
        plan = N.c_void_p()
        N.check(N.lib().xpbd_frame3_create(N.ctypes.byref(d), N.ctypes.byref(plan)), "create")
        run = lambda: N.check(N.lib().xpbd_frame3_run(plan, N.current_stream(), 1, 1), "run")  # noqa
        frame[fuse] = time_kernel(run, frames)
        N.lib().xpbd_frame3_destroy(plan)
    frame_ms = frame[2]
    achieved = nbytes / (ms * 1e-3) / 1e9
    rest_ach = rest_bytes / (rest_ms * 1e-3) / 1e9
    return {
        "workload": f"{sheets} x {res}^2 cloth: {n} particles, {nd} DISTANCE + {m - nd} FLOOR "
                    f"rows ({sc['n_halo']} halo rows re-projected), jacobi 15x2, hanging at "
                    "y = 0.5 m (no FLOOR row active)",
        "kernel": "k_jacobi_tile_cx",
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_launch": nbytes,
        "floor_rows_active": hang_active,
        "kernel_ms": round(ms, 4),
        "layout": "compact tables (12-byte DISTANCE rows, per-particle FLOOR rows)",
        "int4_rows": {"kernel_ms": round(int4_ms, 4),
                      "frac": round(nbytes / (int4_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "resting": {"workload": "the same sheets lying on the floor (y = -2 mm, x_prev = x): "
                                "every FLOOR row active with friction",
                    "floor_rows_active": rest_active, "bytes_per_launch": rest_bytes,
                    "kernel_ms": round(rest_ms, 4), "achieved": round(rest_ach, 1),
                    "frac": round(rest_ach / HBM_PEAK_GBS, 4),
                    "int4_rows": {"kernel_ms": round(rest_int4_ms, 4),
                                  "frac": round(rest_bytes / (rest_int4_ms * 1e-3) / 1e9
                                                / HBM_PEAK_GBS, 4)}},
        "frame_ms": None if frame_ms is None else round(frame_ms, 3),
        "frame_ms_integrate_in_first_launch": None if frame[1] is None else round(frame[1], 3),
        "frame_ms_separate_integrate": None if frame[0] is None else round(frame[0], 3),
        "tile": sc["tile"],
        "proj_per_s": m * 30 / (frame_ms * 1e-3) if frames > 0 else None,
    }


def time_frames(step, frames: int, warmup: int) -> float:
    """ms per frame of ``step()`` over ``frames`` frames after ``warmup``.
    A solver with ``settle()`` (frames whose overflow check runs late) is
    settled inside the timed region, so a frame run again is timed too."""
    import torch

    settle = getattr(getattr(step, "__self__", None), "settle", None) or (lambda: None)
    for _ in range(warmup):
        step()
    settle()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(frames):
        step()
    settle()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / frames * 1e3


def other_workloads(frames: int) -> dict:
    """Configurations 1, 3 and 4 of the benchmark set and the cloth in GS
    mode, timed the same way (secondary lines; the headline is the cloth in
    Jacobi mode)."""
    import torch

    from xpbd_amd import Solver, Solver6DOF, autotune
    from xpbd_amd.batch import BatchedWorld, draw
    from xpbd_amd.scenes import cloth_sheet, hanging_chain, unified_scene

    out = {}
    g = cloth_sheet(Solver, res=128, solve_mode="gs", device="cuda:0")[0]
    ms = time_frames(g.step, max(frames // 4, 10), 5)
    out["cloth128_gs"] = {"ms_per_frame": round(ms, 4), "rows": g.num_constraints,
                          "colors": int(len(g._color_starts) - 1),
                          "proj_per_s": g.num_constraints * 30 / (ms * 1e-3),
                          "substeps_per_s": 15 / (ms * 1e-3)}
    c = hanging_chain(Solver, device="cuda:0")
    ms = time_frames(c.step, frames, 10)
    out["chain64_gs"] = {"ms_per_frame": round(ms, 4), "rows": c.num_constraints,
                         "proj_per_s": c.num_constraints * 30 / (ms * 1e-3),
                         "substeps_per_s": 15 / (ms * 1e-3)}
    u = unified_scene(Solver6DOF, device="cuda:0")
    u.step()
    leaf, leaf_times = autotune.tune6(u, frames=3)   # LBVH leaf size (untimed setup)
    # 30 untimed frames: the boxes land and the pair and manifold buffers
    # reach their steady size (a growth re-captures the frame graph)
    ms = time_frames(u.step, max(frames // 4, 10), 30)
    rows = u.num_joints + int(u.contact_counts().sum()) + 8 * int(
        int((u.inv_masses() > 0).sum()))
    out["unified_1440_boxes"] = {"ms_per_frame": round(ms, 4), "bodies": u.num_bodies,
                                 "joints": u.num_joints, "pairs": u.n_pairs,
                                 "manifold_points": int(u.contact_counts().sum()),
                                 "rows": rows, "proj_per_s": rows * 30 / (ms * 1e-3),
                                 "lbvh_leaf": leaf,
                                 "leaf_trials_ms_per_frame": {str(k): round(v, 4)
                                                              for k, v in leaf_times.items()}}
    spec = draw(1024, 0)
    del g, c, u
    # release the allocator's cached blocks first: inside this process (after
    # the headline section) the world otherwise ran ~25 % slower
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    w = BatchedWorld(spec, device="cuda:0")
    ms = time_frames(w.step, max(frames // 4, 10), 30)
    rows = w.rows()
    out["batched_1024_scenes"] = {"ms_per_frame": round(ms, 4), **spec.totals(), "rows": rows,
                                  "proj_per_s": rows * 30 / (ms * 1e-3),
                                  "substeps_per_s": 15 / (ms * 1e-3), "mix": spec.mix}
    return out


# The reference's own published rows (README.md:108-115, 186-191; Warp on an
# RTX 3060 Laptop GPU, ms per frame), for the same scenes built here.
PUBLISHED_MS = {
    ("cloth16", "gs"): 6.6, ("cloth16", "jacobi"): 2.4, ("cloth32", "gs"): 6.5,
    ("cloth32", "jacobi"): 2.3, ("cloth48", "gs"): 6.5, ("cloth48", "jacobi"): 2.3,
    ("stack8x8x5", "gs"): 7.8, ("stack8x8x5", "jacobi"): 5.2,
    ("boxes27", "lbvh"): 1.33, ("boxes100", "lbvh"): 1.38, ("boxes384", "lbvh"): 1.41,
    ("boxes1440", "lbvh"): 1.84,
}


def reference_rows(frames: int = 120, warmup: int = 20) -> dict:
    """The reference benchmark's scenes (examples/benchmark.py:71-107 and the
    6-DOF box piles of README.md:186-191), timed like its time_steps
    (warm-up frames, then frames between device syncs), with the published
    RTX 3060 Laptop figures beside them."""
    from xpbd_amd import Solver, Solver6DOF
    from xpbd_amd.scenes import box_pile, cloth_bench, sphere_stack

    rows = {}

    def put(name, mode, s, ms, key=None):
        n_rows = s.num_constraints + int(getattr(s, "n_contacts", 0) or 0)  # + device rows
        proj = n_rows * s.substeps * s.iterations
        pub = PUBLISHED_MS.get((name, mode))
        rows[key or f"{name}_{mode}"] = {
            "ms_per_frame": round(ms, 4), "bodies": s.num_bodies, "rows": n_rows,
            "proj_per_s": proj / (ms * 1e-3), "published_rtx3060_ms": pub,
            "speedup_vs_published": None if pub is None else round(pub / ms, 1)}

    for mode in ("gs", "jacobi"):
        for res in (16, 32, 48):
            s = cloth_bench(Solver, res, mode, device="cuda:0")
            put(f"cloth{res}", mode, s, time_frames(s.step, frames, warmup))
        s = sphere_stack(Solver, 8, 5, mode, device="cuda:0")
        put("stack8x8x5", mode, s, time_frames(s.step, frames, warmup))
    # the same stacks with their CONTACT rows generated on the device (the
    # host path's rows, in its order) instead of the per-frame host broad
    # phase and re-upload; in GS mode also recoloured on the device and the
    # frame run in one workgroup
    for mode in ("jacobi", "gs"):
        s = sphere_stack(Solver, 8, 5, mode, device="cuda:0")
        s.enable_self_collision(False)
        s.enable_device_contacts(True, margin=0.0)
        put("stack8x8x5", mode, s, time_frames(s.step, frames, warmup),
            key=f"stack8x8x5_{mode}_device_contacts")
    for dims in ((3, 3, 3), (5, 4, 5), (8, 6, 8), (12, 10, 12)):
        s = box_pile(Solver6DOF, *dims, device="cuda:0", broadphase="lbvh")
        n = dims[0] * dims[1] * dims[2]
        ms = time_frames(s.step, frames, warmup)
        pub = PUBLISHED_MS.get((f"boxes{n}", "lbvh"))
        rows[f"boxes{n}_lbvh"] = {"ms_per_frame": round(ms, 4), "bodies": s.num_bodies,
                                  "published_rtx3060_ms": pub,
                                  "speedup_vs_published": round(pub / ms, 1)}
    return rows


def judge_interval(name: str):
    """Frames between the judge's checks of the learned start for a timed
    configuration: the solver's default (every 8th frame) unless the name
    says otherwise (None: leave the default)."""
    if name.endswith("no_judge"):
        return 0
    if name.endswith("judge1"):
        return 1
    return None


def warm_start_section(res: int, frames: int) -> dict:
    """Learned lambda warm start on the cloth: fit (two rounds, on the GPU),
    RMS residual of each start on its own trajectory, and whole-frame
    substeps/s of the captured frame plans (tile kernel; each substep starts
    with the warm-start launch, after the network pass on every substep or,
    "net_per_frame", on the first substep of each frame only, with a network
    trained for that use), each plan autotuned."""
    import torch

    from xpbd_amd import Solver
    from xpbd_amd import autotune
    from xpbd_amd import warmstart as ws
    from xpbd_amd.scenes import cloth_sheet

    def cloth(iters=2):
        return cloth_sheet(Solver, res=res, solve_mode="jacobi", device="cuda:0",
                           iterations=iters)[0]

    t0 = time.perf_counter()
    log = []
    theta = ws.fit(cloth, substeps=120, settle_frames=2, steps=800, log=log)
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t0
    theta_frame = ws.fit(cloth, substeps=120, settle_frames=2, steps=800, log=log, net_every=15,
                         pair_with_net=True)
    s = cloth()
    s.step(20)
    resid, fallback = ws.Evaluator(s, theta).run(substeps=60, plain_iters=(1, 2, 4),
                                                 learned_iters=(1, 2))
    amort, _ = ws.Evaluator(s, theta_frame).run(substeps=60, plain_iters=(),
                                                learned_iters=(1, 2), net_every=15)
    resid.update({f"{k}_net_per_frame": v for k, v in amort.items()})
    sps, shapes, judged = {}, {}, {}
    for name, iters, learned, every in [("plain_4", 4, False, 1), ("plain_2", 2, False, 1),
                                        ("plain_1", 1, False, 1), ("learned_2", 2, True, 1),
                                        ("learned_1", 1, True, 1),
                                        ("learned_2_net_per_frame", 2, True, 15),
                                        ("learned_1_net_per_frame", 1, True, 15),
                                        ("learned_1_net_per_frame_judge1", 1, True, 15),
                                        ("learned_1_net_per_frame_no_judge", 1, True, 15)]:
        c = cloth(iters)
        c.block_frames = False
        if learned:
            c.warm_start_every = every
            if judge_interval(name) is not None:
                c.warm_start_judge = judge_interval(name)
            c.set_warm_start(theta if every == 1 else theta_frame)
        best, _ = autotune.tune(c)
        shapes[name] = list(best) if best else None
        ms = time_frames(c.step, frames, 5)
        sps[name] = round(c.substeps / (ms * 1e-3), 1)
        if learned:
            judged[name] = c.warm_start_stats()
    return {
        "scene": f"cloth {res}x{res}, 15 substeps, Jacobi tile kernel",
        "fit_seconds": round(fit_s, 2), "fit": log[0] if log else None,
        "fit_net_per_frame": log[1] if len(log) > 1 else None,
        "rms_residual": {k: float(f"{v:.4e}") for k, v in resid.items()},
        "fallback_substeps_of_60": fallback,
        "substeps_per_s": sps,
        "judge": judged,
        "launch_shapes": shapes,
        "learned1_vs_plain2": {"residual_ratio": round(resid["learned_1"] / resid["plain_2"], 3),
                               "speedup": round(sps["learned_1"] / sps["plain_2"], 3)},
        "learned2_vs_plain4": {"residual_ratio": round(resid["learned_2"] / resid["plain_4"], 3),
                               "speedup": round(sps["learned_2"] / sps["plain_4"], 3)},
        "learned1_vs_plain4": {"residual_ratio": round(resid["learned_1"] / resid["plain_4"], 3),
                               "speedup": round(sps["learned_1"] / sps["plain_4"], 3)},
        "learned1_net_per_frame_vs_plain4": {
            "residual_ratio": round(resid["learned_1_net_per_frame"] / resid["plain_4"], 3),
            "speedup": round(sps["learned_1_net_per_frame"] / sps["plain_4"], 3)},
        "learned2_net_per_frame_vs_plain4": {
            "residual_ratio": round(resid["learned_2_net_per_frame"] / resid["plain_4"], 3),
            "speedup": round(sps["learned_2_net_per_frame"] / sps["plain_4"], 3)},
        "learned1_net_per_frame_judge1_vs_plain4": {
            "speedup": round(sps["learned_1_net_per_frame_judge1"] / sps["plain_4"], 3)},
        "learned1_net_per_frame_no_judge_vs_plain4": {
            "speedup": round(sps["learned_1_net_per_frame_no_judge"] / sps["plain_4"], 3)},
    }


def batched_particles(fit_log=None):
    """The batched world's particle solver after two frames, with a
    ``restore()`` that puts it back to that state (two plain iterations, no
    learned start), and the network for one pass per frame fitted on the
    world (two rounds, on the GPU) when ``fit_log`` is a list (it receives
    the fit's log); returns (world, solver, restore, theta or None, fit
    seconds)."""
    import torch

    from xpbd_amd import warmstart as ws
    from xpbd_amd.batch import BatchedWorld, draw

    w = BatchedWorld(draw(1024, 0), device="cuda:0")
    s = w.particles
    s.block_frames = False
    s.step(2)
    torch.cuda.synchronize()
    x0, v0 = s.positions(), s.velocities()     # API order: survives a re-layout

    def restore():
        if s._ws_theta is not None:
            s.set_warm_start(None)
        s.iterations = 2
        s.set_state(x0, v0)
        return s

    theta, fit_s = None, 0.0
    if fit_log is not None:
        t0 = time.perf_counter()
        theta = ws.fit(restore, substeps=60, settle_frames=1, steps=800, log=fit_log,
                       net_every=15, pair_with_net=True, every=5)
        torch.cuda.synchronize()
        fit_s = time.perf_counter() - t0
        restore()
    return w, s, restore, theta, fit_s


def warm_start_batched_section(frames: int) -> dict:
    """Learned lambda warm start on the particle solver of the batched world
    (1024 uneven scenes: chains, cloths, sphere piles with device contact
    rows; per-scene gravity, compliance and friction): a network for one
    pass per frame is fitted on the world itself (two rounds, on the GPU),
    then each start runs on its own trajectory from the same state (RMS
    residual over the static rows, 60 substeps) and whole frames are timed
    through the captured frame plan."""
    import torch

    from xpbd_amd import warmstart as ws

    log = []
    w, s, restore, theta, fit_s = batched_particles(fit_log=log)
    s.step(4)
    resid, fallback = ws.Evaluator(s, theta).run(substeps=60, plain_iters=(1, 2, 4),
                                                 learned_iters=(1, 2), net_every=15)
    torch.cuda.empty_cache()
    sps, judged = {}, {}
    for name, iters, learned in [("plain_4", 4, False), ("plain_2", 2, False),
                                 ("plain_1", 1, False), ("learned_2_net_per_frame", 2, True),
                                 ("learned_1_net_per_frame", 1, True),
                                 ("learned_1_net_per_frame_judge1", 1, True),
                                 ("learned_1_net_per_frame_no_judge", 1, True)]:
        restore()
        s.iterations = iters
        if learned:
            s.warm_start_every = 15
            if judge_interval(name) is not None:
                s.warm_start_judge = judge_interval(name)
            s.set_warm_start(theta)
        ms = time_frames(s.step, frames, 3)
        sps[name] = round(s.substeps / (ms * 1e-3), 1)
        if learned:
            judged[name] = s.warm_start_stats()
    restore()
    resid = {(f"{k}_net_per_frame" if k.startswith("learned") else k): v for k, v in resid.items()}
    out = {
        "scene": f"batched world, particle solver: {s.num_bodies} particles, "
                 f"{s.num_constraints} static rows + device CONTACT rows, 1024 scenes, "
                 "15 substeps, Jacobi tile kernel",
        "fit_seconds": round(fit_s, 2), "fit": log[0] if log else None,
        "rms_residual_static_rows": {k: float(f"{v:.4e}") for k, v in resid.items()},
        "fallback_substeps_of_60": fallback,
        "substeps_per_s": sps,
        "judge": judged,
    }
    for k in (1, 2):
        out[f"learned{k}_net_per_frame_vs_plain4"] = {
            "residual_ratio": round(resid[f"learned_{k}_net_per_frame"] / resid["plain_4"], 3),
            "speedup": round(sps[f"learned_{k}_net_per_frame"] / sps["plain_4"], 3)}
    for v in ("judge1", "no_judge"):
        out[f"learned1_net_per_frame_{v}_vs_plain4"] = {
            "speedup": round(sps[f"learned_1_net_per_frame_{v}"] / sps["plain_4"], 3)}
    del w, s
    return out


def load_traffic(kernel: str):
    """HBM bytes per launch of ``kernel`` measured by rocprofv3 --pmc runs
    (FETCH_SIZE and WRITE_SIZE scaled by a plain copy of known size),
    committed under profiles/; None when no measurement exists."""
    p = ROOT / "profiles" / "traffic.json"
    if not p.exists():
        return None
    try:
        e = json.loads(p.read_text()).get(kernel)
    except (ValueError, OSError):
        return None
    return None if e is None else e.get("bytes_per_launch")


def traffic_frac(traffic, kernel_ms):
    """Counter-measured HBM bytes per launch (profiles/traffic.json) over the
    launch time, as a fraction of the HBM peak (None without counters)."""
    if not traffic or not kernel_ms:
        return None
    return round(traffic / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)


def scene_only(name: str, frames: int, warmup: int) -> dict:
    """Time one scene alone (for kernel traces of its frames): ``boxesN``
    (the reference's N-box LBVH pile, N in 27, 100, 384, 1440), ``cloth128``
    (the headline cloth, Jacobi), ``stack-gs-device`` (the 8 x 8 x 5 sphere
    stack in GS mode with device CONTACT rows), ``tile-kernel`` (the
    headline cloth's dominant launch replayed from a graph),
    ``batched-plain4`` / ``batched-learned1`` (the batched world's particle
    solver with four plain iterations, or one after the learned start with
    its network fitted on the world first, judge on every 8th frame)."""
    import torch

    from xpbd_amd import Solver, Solver6DOF
    from xpbd_amd import _native as N
    from xpbd_amd.scenes import box_pile, cloth_sheet, sphere_stack

    dims = {"boxes27": (3, 3, 3), "boxes100": (5, 4, 5), "boxes384": (8, 6, 8),
            "boxes1440": (12, 10, 12)}
    if name in dims:
        s = box_pile(Solver6DOF, *dims[name], device="cuda:0", broadphase="lbvh")
    elif name in ("cloth128", "tile-kernel"):
        s = cloth_sheet(Solver, res=128, solve_mode="jacobi", device="cuda:0")[0]
        s.tile_size = 64
    elif name == "stack-gs-device":
        s = sphere_stack(Solver, 8, 5, "gs", device="cuda:0")
        s.enable_self_collision(False)
        s.enable_device_contacts(True, margin=0.0)
    elif name in ("batched-plain4", "batched-learned1"):
        learned = name == "batched-learned1"
        w, s, _, theta, _ = batched_particles(fit_log=[] if learned else None)
        if learned:
            s.warm_start_every = 15
            s.set_warm_start(theta)
        s.iterations = 1 if learned else 4
    else:
        raise SystemExit(f"unknown scene {name!r}")
    if name == "tile-kernel":
        s.step(warmup)
        launch, kname, iters = headline_launch(N, s)
        return {"scene": name, "kernel": kname, "iterations_per_launch": iters,
                "kernel_ms": time_kernel_graph(launch, frames)}
    return {"scene": name, "ms_per_frame": time_frames(s.step, frames, warmup)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--scene", default=None,
                    help="time only this scene and print its line (see scene_only)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--mode", default="jacobi", choices=["jacobi", "gs"])
    ap.add_argument("--res", type=int, default=128)
    ap.add_argument("--kernel-reps", type=int, default=200)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline-scene", action="store_true")
    ap.add_argument("--roofline-frames", type=int, default=5,
                    help="frames timed on the roofline scene (0: kernel only)")
    ap.add_argument("--no-others", action="store_true",
                    help="skip the chain / unified / batched secondary lines")
    ap.add_argument("--no-reference-rows", action="store_true",
                    help="skip the reference benchmark's own scenes")
    ap.add_argument("--no-warm-start", action="store_true",
                    help="skip the learned lambda warm-start section")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shape instead of the autotuner's pick")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run plain 512 MiB device copies (PMC calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} requested; this benchmark measures one GPU only "
              "(independent scenes would shard across GPUs without communication; the "
              "multi-GPU path is not built)", file=sys.stderr)
        return 2

    import torch

    from xpbd_amd import Solver
    from xpbd_amd import _native as N
    from xpbd_amd.scenes import cloth_sheet

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.scene:
        print(json.dumps(scene_only(args.scene, args.steps, args.warmup)), flush=True)
        return 0
    s, _ = cloth_sheet(Solver, res=args.res, solve_mode=args.mode, device="cuda:0")
    host = s.host_arrays()
    n_dist = int((host["type"] == 0).sum())
    n_floor = int((host["type"] == 2).sum())
    tuned = None
    if args.mode == "jacobi" and not args.no_autotune:
        from xpbd_amd import autotune

        best, trials = autotune.tune(s, frames=4)
        tuned = {"kernel": best[0], "tile": best[1], "fuse_first": best[3],
                 "trials_ms_per_frame": {f"{c[0]}:{c[1]}:{'fused' if c[3] else 'separate'}":
                                         round(v, 4) for c, v in trials.items()}}
        print(f"[bench] autotune picked {best}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        s.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = elapsed / args.steps * 1e3
    proj = s.num_constraints * s.substeps * s.iterations
    value = proj / (ms * 1e-3)
    print(f"[bench] {args.mode} cloth{args.res}: {ms:.4f} ms/frame, {value / 1e9:.3f} G proj/s",
          file=sys.stderr, flush=True)

    roof = None
    if args.mode == "jacobi" and s._tile_dev is not None and s.jacobi_kernel == "tile":
        launch, kname, iters = headline_launch(N, s)
        x_keep, v_keep = s.positions(), s.velocities()
        active = floor_rows_active(s.X, s.R)
        kms = time_kernel_graph(launch, args.kernel_reps)
        s.set_state(x_keep, v_keep)
        # a pair launch runs two iterations (its fused finalize and next
        # integrate are not counted: a lower bound on its bytes)
        nbytes = iters * iter_bytes(n_dist, n_floor, s.num_bodies, active > 0)
        ach = nbytes / (kms * 1e-3) / 1e9
        traffic = load_traffic(kname)
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_frac": traffic_frac(traffic, kms),
                "kernel": kname, "iterations_per_launch": iters, "kernel_ms": round(kms, 5),
                "bytes_per_launch": nbytes, "floor_rows_active": active, "tile": s._tile}
        print(f"[bench] {kname} {kms * 1e3:.2f} us, {ach:.1f} GB/s", file=sys.stderr,
              flush=True)

    if args.calib_copy:
        a = torch.empty(128 << 20, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    scene = None
    if not args.no_roofline_scene:
        scene = roofline_scene(N, reps=20, frames=args.roofline_frames)
        print(f"[bench] roofline scene: kernel {scene['kernel_ms']:.3f} ms "
              f"{scene['achieved']} GB/s, frame {scene['frame_ms']} ms", file=sys.stderr,
              flush=True)
        scene["traffic"] = load_traffic("k_jacobi_tile_cx@8x1024^2")
        scene["traffic_frac"] = traffic_frac(scene["traffic"], scene["kernel_ms"])
        rest = scene["resting"]
        rest["traffic"] = load_traffic("k_jacobi_tile_cx@8x1024^2 resting")
        rest["traffic_frac"] = traffic_frac(rest["traffic"], rest["kernel_ms"])
        for key, part in (("k_jacobi_tile@8x1024^2", scene["int4_rows"]),
                          ("k_jacobi_tile@8x1024^2 resting", rest["int4_rows"])):
            part["traffic"] = load_traffic(key)
            part["traffic_frac"] = traffic_frac(part["traffic"], part["kernel_ms"])
        torch.cuda.empty_cache()

    others = None
    if not args.no_others:
        others = other_workloads(max(args.steps // 3, 30))
        print(f"[bench] others: {json.dumps(others)}", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()

    ref_rows = None
    if not args.no_reference_rows:
        ref_rows = reference_rows()
        print(f"[bench] reference rows: {json.dumps(ref_rows)}", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()

    warm = None
    if not args.no_warm_start and args.mode == "jacobi":
        warm = warm_start_section(args.res, max(args.steps // 3, 30))
        print(f"[bench] warm start: {json.dumps(warm)}", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()

    warm_b = None
    if not args.no_warm_start and args.mode == "jacobi":
        warm_b = warm_start_batched_section(max(args.steps // 10, 10))
        print(f"[bench] warm start batched: {json.dumps(warm_b)}", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(host, s, args.cpu_budget)
        print(f"[bench] cpu baseline {cpu['value'] / 1e6:.2f} M proj/s", file=sys.stderr,
              flush=True)

    line = {
        "metric": METRIC, "value": value, "unit": "proj/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"cloth {args.res}x{args.res}: {s.num_bodies} particles, "
                               f"{n_dist} DISTANCE + {n_floor} FLOOR rows, {args.mode}, "
                               f"15 substeps x 2 iterations, dt 1/60, hipGraph per frame",
                   "rows": s.num_constraints, "substeps": s.substeps,
                   "iterations": s.iterations, "mode": args.mode, "launch": tuned},
        "substeps_per_s": s.substeps / (ms * 1e-3),
        "roofline": roof,
        "roofline_scene": scene,
        "cpu_baseline": cpu,
        "other_workloads": others,
        "learned_warm_start": warm,
        "learned_warm_start_batched": warm_b,
        "reference_benchmark_rows": ref_rows,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
