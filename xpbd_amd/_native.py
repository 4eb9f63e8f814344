"""ctypes binding of ``libxpbd_hip.so`` (the C-ABI declared in include/xpbd.h).

The library is loaded after ``torch`` so that both share one HIP runtime
(torch ships ``libamdhip64.so.7`` under the same soname, and the loader
reuses the copy already mapped).  A missing library is an error: there is
no fallback path.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "lib" / "libxpbd_hip.so"
_lib = None

c_int = ctypes.c_int
c_float = ctypes.c_float
c_void_p = ctypes.c_void_p
P_int = ctypes.POINTER(ctypes.c_int)


class ParticlesDesc(ctypes.Structure):
    """Mirror of ``xpbd_particles_desc`` (include/xpbd.h)."""

    _fields_ = [
        ("x", c_void_p), ("v", c_void_p), ("x_prev", c_void_p), ("acc", c_void_p),
        ("rows", c_void_p), ("anchors", c_void_p), ("lam", c_void_p),
        ("n_particles", c_int), ("n_rows", c_int),
        ("color_starts", c_void_p), ("num_colors", c_int),
        ("mode", c_int), ("substeps", c_int), ("iterations", c_int),
        ("h", c_float), ("gx", c_float), ("gy", c_float), ("gz", c_float),
        ("damping", c_float), ("relax", c_float), ("max_speed", c_float),
        ("rows_per_thread", c_int), ("window", c_int),
        ("x_alt", c_void_p), ("lam_alt", c_void_p), ("tile_row_start", c_void_p),
        ("halo_start", c_void_p), ("halo_rows", c_void_p), ("tile_size", c_int),
        ("dyn_rows", c_void_p), ("dyn_lam", c_void_p), ("dyn_count", c_void_p),
        ("dyn_cap", c_int), ("gscale", c_void_p),
        ("ws_image", c_void_p), ("ws_y", c_void_p),
        ("fuse_first", c_int), ("block", c_int), ("color_starts_dev", c_void_p),
        ("ws_every", c_int),
        ("ws_on", c_void_p), ("ws_judge", c_int), ("ws_jx", c_void_p), ("ws_jx_alt", c_void_p),
        ("ws_jxp", c_void_p), ("ws_jlam", c_void_p), ("ws_jlam_alt", c_void_p),
        ("ws_jy", c_void_p), ("ws_jacc", c_void_p), ("ws_jdlam", c_void_p),
        ("ws_rsum", c_void_p), ("ws_stats", c_void_p),
        ("x_prev_alt", c_void_p), ("pair_ring_start", c_void_p), ("pair_ring1_n", c_void_p),
        ("pair_ring_idx", c_void_p), ("pair_p1_start", c_void_p), ("pair_p1", c_void_p),
        ("pair_p2_start", c_void_p), ("pair_p2", c_void_p),
        ("pair_lmax", c_int), ("pair_amax", c_int), ("pair_k1max", c_int),
        ("compact", c_void_p), ("ws_jxp_alt", c_void_p),
    ]


class TileCompact(ctypes.Structure):
    """Mirror of ``xpbd_tile_compact`` (include/xpbd.h)."""

    _fields_ = ([(k, c_void_p) for k in ("dist_end", "gen_end", "ci", "cc", "halo2", "pf",
                                         "pflam")] + [("general", c_int), ("acc_tiles", c_void_p)])


class RigidDesc(ctypes.Structure):
    """Mirror of ``xpbd_rigid_desc`` (include/xpbd.h)."""

    _fields_ = (
        [(k, c_void_p) for k in ("x", "q", "v", "w", "x_prev", "q_prev", "inv_I", "he", "dx",
                                 "drot", "dv", "dw", "lam_floor")]
        + [("n", c_int), ("pairs", c_void_p), ("pair_count", c_void_p), ("cap", c_int)]
        + [(k, c_void_p) for k in ("m_count", "pt_idx", "pt_n", "pt_oi", "pt_or", "lam_pt",
                                   "pt_count", "j_idx", "j_a", "j_b", "lam_j")]
        + [("nj", c_int), ("substeps", c_int), ("iterations", c_int), ("gyroscopic", c_int)]
        + [(k, c_float) for k in ("h", "gx", "gy", "gz", "floor_y", "friction", "restitution",
                                  "lin_damp", "ang_damp", "relax", "max_speed", "max_omega",
                                  "man_margin")]
        + [("gscale", c_void_p), ("aux", c_void_p)]
    )


_P = c_void_p
_F = c_float
_I = c_int

# name -> (restype, argtypes)
_SIGS = {
    "xpbd_particles_pass": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                    c_float, c_float, c_float, c_float, c_float, c_float,
                                    c_float, c_int]),
    "xpbd_integrate": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float, c_float,
                               c_float, c_int]),
    "xpbd_apply_jacobi": (c_int, [c_void_p, c_void_p, c_void_p, c_float, c_int]),
    "xpbd_finalize_velocity": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float,
                                       c_float, c_int]),
    "xpbd_solve_jacobi": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_float, c_int, c_int, c_int, c_int]),
    "xpbd_jacobi_tile": (c_int, [c_void_p] * 12 + [c_float, c_int, c_int, c_int, c_float, c_int,
                                                    c_float, c_float, c_void_p]),
    "xpbd_jacobi_tile_first": (c_int, [c_void_p] * 11 + [c_float, c_float, c_float, c_float,
                                                         c_void_p, c_int, c_int, c_float,
                                                         c_void_p]),
    "xpbd_jacobi_tile_next": (c_int, [c_void_p] * 11 + [c_float, c_int, c_float, c_float, c_float,
                                                        c_void_p, c_int, c_int, c_float, c_float,
                                                        c_float, c_void_p]),
    # the same launches over compact tables (a TileCompact address in place
    # of halo_rows)
    "xpbd_jacobi_tile_cx": (c_int, [c_void_p] * 12 + [c_float, c_int, c_int, c_int, c_float,
                                                       c_int, c_float, c_float, c_void_p]),
    "xpbd_jacobi_tile_first_cx": (c_int, [c_void_p] * 11 + [c_float, c_float, c_float, c_float,
                                                            c_void_p, c_int, c_int, c_float,
                                                            c_void_p]),
    "xpbd_jacobi_tile_next_cx": (c_int, [c_void_p] * 11 + [c_float, c_int, c_float, c_float,
                                                           c_float, c_void_p, c_int, c_int,
                                                           c_float, c_float, c_float, c_void_p]),
    "xpbd_integrate_to": (c_int, [c_void_p] * 5 + [c_float, c_float, c_float, c_float, c_void_p,
                                                   c_int]),
    "xpbd_solve_jacobi_dyn": (c_int, [c_void_p] * 6 + [c_float, c_int, c_void_p, c_int]),
    "xpbd_solve_jacobi_dyn_first": (c_int, [c_void_p] * 6 + [c_float, c_float, c_float, c_float,
                                                             c_void_p, c_void_p, c_int]),
    "xpbd_contacts_workspace_bytes": (ctypes.c_longlong, [c_int]),
    "xpbd_sphere_contacts": (c_int, [c_void_p] * 8 + [c_int, c_void_p, c_int, ctypes.c_double,
                                                      ctypes.c_double, c_float, c_int, c_int,
                                                      c_void_p, c_void_p]),
    "xpbd_frame_block": (c_int, [c_void_p] * 8 + [c_int] * 5 + [c_float] * 7 + [c_int, c_int]),
    "xpbd_frame_block_capacity": (c_int, [c_int]),
    "xpbd_solve_color": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_float, c_int, c_int, c_int]),
    "xpbd_frame3_create": (c_int, [ctypes.POINTER(ParticlesDesc), ctypes.POINTER(c_void_p)]),
    "xpbd_frame3_run": (c_int, [c_void_p, c_void_p, c_int, c_int]),
    "xpbd_frame3_destroy": (c_int, [c_void_p]),
    "xpbd_color_constraints": (c_int, [c_int, c_void_p, c_void_p, c_int, c_void_p]),
    "xpbd_color_rows_max_bodies": (c_int, []),
    "xpbd_color_rows_device": (c_int, [_P, _P, _I, _P, _I, _I, _P, _P, _P, _P, _P]),
    "xpbd_frame_block_ordered": (c_int, [_P] * 10 + [_I, _I, _I] + [_F] * 7 + [_I]),
    "xpbd_color_ranges": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "xpbd6_integrate": (_I, [_P] * 10 + [_F, _F, _F, _F, _I, _P, _I]),
    "xpbd6_box_manifold": (_I, [_P] * 6 + [_F, _I] + [_P] * 7),
    "xpbd6_solve_joints": (_I, [_P] * 8 + [_F, _I, _P, _P, _I]),
    "xpbd6_solve_floor": (_I, [_P] * 7 + [_F, _F, _P, _P, _P, _I]),
    "xpbd6_solve_box": (_I, [_P] * 12 + [_I]),
    "xpbd6_apply_jacobi": (_I, [_P] * 5 + [_F, _I]),
    "xpbd6_finalize_velocity": (_I, [_P] * 7 + [_F, _F, _F, _I]),
    "xpbd6_velocity_floor": (_I, [_P] * 5 + [_F, _F, _P, _F, _P, _P, _P, _P, _I]),
    "xpbd6_velocity_box": (_I, [_P] * 10 + [_F, _F, _F, _P, _P, _P, _P, _I]),
    "xpbd6_frame_block": (_I, [_P, _P, _P, _P, _F, _I]),
    "xpbd6_frame_block_capacity": (_I, []),
    "xpbd6_apply_velocity": (_I, [_P] * 6 + [_F, _F, _I, _I]),
    "xpbd6_begin_substep": (_I, [_P] * 12 + [_F, _F, _I, _I, _F, _F, _F, _F, _I, _P, _I]),
    "xpbd6_solve_all": (_I, [_P] * 11 + [_F, _I, _I, _F, _F, _P, _I] + [_P] * 6 + [_I, _P, _P]),
    "xpbd6_apply_finalize": (_I, [_P] * 5 + [_F, _I] + [_P] * 4 + [_F, _F, _F, _I]),
    "xpbd6_velocity_all": (_I, [_P] * 5 + [_F, _F, _P, _F, _P, _P, _I] + [_P] * 6 + [_I, _P, _P]),
    "xpbd_frame6_create": (_I, [ctypes.POINTER(RigidDesc), ctypes.POINTER(c_void_p)]),
    "xpbd_frame6_run": (_I, [_P, _P, _I, _I]),
    "xpbd_jacobi_tile_pair": (_I, [_P] * 16 + [_F, _I, _I, _F, _F, _F, _F, _F, _F, _P, _I, _I, _I, _I,
                                              _I]),
    "xpbd_jacobi_tile_pair_lds_bytes": (ctypes.c_size_t, [_I, _I, _I]),
    "xpbd_jacobi_tile_pair_warm": (_I, [_P] * 16 + [_F, _I, _I, _F, _F, _F, _F, _F, _F, _P, _I,
                                                   _I, _I, _I, _I, _P, _P, _I]),
    "xpbd6_substep_contacts": (_I, [_P, _P, _I]),
    "xpbd6_substep_velocity": (_I, [_P, _P, _I]),
    "xpbd6_substep_end": (_I, [_P, _P]),
    "xpbd6_substep_end_report": (_I, [_P, _P, _P, _I, _P]),
    "xpbd_frame6_set_broadphase": (_I, [_P, _I, _P, _I, _F, _F, _F]),
    "xpbd_frame6_status": (_I, [_P, _I, _P]),
    "xpbd_frame6_destroy": (_I, [_P]),
    "xpbd6_body_aabb": (_I, [_P] * 6 + [_F, _F, _P, _P, _I]),
    "xpbd6_lbvh_workspace_bytes": (ctypes.c_longlong, [_I]),
    "xpbd6_hashgrid_workspace_bytes": (ctypes.c_longlong, [_I]),
    "xpbd6_broadphase_lbvh": (_I, [_P] * 7 + [_F, _F, _I, _I, _P, _P]),
    "xpbd6_broadphase_lbvh_leaf": (_I, [_P] * 7 + [_F, _F, _I, _I, _P, _P, _I]),
    "xpbd6_broadphase_lbvh_noalloc": (_I, [_P] * 6 + [_F, _F, _I, _I, _P, _P]),
    "xpbd6_broadphase_hashgrid": (_I, [_P] * 7 + [_F, _F, _F, _F, _I, _I, _P, _P]),
    "xpbd_morton3": (ctypes.c_uint, [_F, _F, _F]),
    "xpbd_mlp_param_count": (c_int, []),
    "xpbd_mlp_forward": (c_int, [_P, _P, _P, _P, _I]),
    "xpbd_mlp_image_bytes": (ctypes.c_longlong, []),
    "xpbd_mlp_pack": (c_int, [_P, _P, _P]),
    "xpbd_mlp_train_workspace_bytes": (ctypes.c_longlong, [_I]),
    "xpbd_mlp_grad": (c_int, [_P, _P, _P, _P, _I, _I, _P, _P, _P]),
    "xpbd_mlp_adam": (c_int, [_P, _P, _P, _P, _P, _F, _F, _F, _F, _I]),
    "xpbd_ws_features": (c_int, [_P] * 5 + [_F, _P, _I]),
    "xpbd_ws_net": (c_int, [_P] * 8 + [_F, _F, _F, _F, _P, _I, _I, _P, _I]),
    "xpbd_jacobi_tile_warm": (c_int, [_P] * 12 + [_F, _F, _F, _F, _P, _I, _I, _F, _I, _P, _I]),
    "xpbd_jacobi_tile_warm_cx": (c_int, [_P] * 12 + [_F, _F, _F, _F, _P, _I, _I, _F, _I, _P,
                                                     _I]),
    "xpbd_ws_decide": (c_int, [_P, _P, _P, _P]),
    "xpbd_ws_residual": (c_int, [_P] * 5 + [_F, _P, _I]),
    "xpbd_build_info": (ctypes.c_char_p, []),
    "xpbd_error_string": (ctypes.c_char_p, [c_int]),
}

_EXTRA_SIGS: dict = {}


def register(sigs: dict) -> None:
    """Add signatures of entry points defined by other modules of the package."""
    _EXTRA_SIGS.update(sigs)
    if _lib is not None:
        _bind(_lib, sigs)


def _bind(lib, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib_path() -> Path:
    return _LIB_PATH


def lib():
    """Load (once) and return the bound library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  (shares its HIP runtime with the library)

    if not _LIB_PATH.exists():
        raise RuntimeError(
            f"{_LIB_PATH} is missing: build it with `python -m xpbd_amd.build` "
            "(the solver has no non-HIP path)")
    mode = os.RTLD_NOW | os.RTLD_GLOBAL
    lb = ctypes.CDLL(str(_LIB_PATH), mode=mode)
    _bind(lb, _SIGS)
    _bind(lb, _EXTRA_SIGS)
    _lib = lb
    return lb


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().xpbd_error_string(int(rc)).decode()
        raise RuntimeError(f"{what or 'xpbd call'} failed: {msg} (code {rc})")


def ptr(t) -> int:
    """Device (or host) address of a contiguous torch tensor or numpy array."""
    if hasattr(t, "data_ptr"):
        assert t.is_contiguous()
        return t.data_ptr()
    return t.ctypes.data


def current_stream():
    import torch

    return c_void_p(torch.cuda.current_stream().cuda_stream)
