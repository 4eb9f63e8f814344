"""The allocation-free LBVH broad phase (xpbd6_broadphase_lbvh_noalloc, a
fixed per-GPU scratch array in place of the caller's workspace) against the
existing entry point (xpbd6_broadphase_lbvh with a caller workspace), called
through the C-ABI: the same pair list and count, bit for bit, on two calls
in a row, the second with a smaller scene (6000 bodies: the LBVH tree path;
then 300 bodies: the all-pairs path), and a scene above the fixed maximum
rejected before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _boxes(n, seed, spread):
    from xpbd_amd import Solver6DOF
    from xpbd_amd.scenes import _random_quats

    rng = np.random.default_rng(seed)
    s = Solver6DOF(dt=1 / 60, substeps=1, iterations=1, device=DEV)
    q = _random_quats(rng, n)
    for i in range(n):
        s.add_box(tuple(rng.uniform(-spread, spread, 3) + [0, spread + 1, 0]),
                  tuple(rng.uniform(0.05, 0.2, 3)), mass=0.0 if i % 17 == 0 else 1.0,
                  quaternion=tuple(q[i]), velocity=tuple(rng.normal(size=3)),
                  omega=tuple(rng.normal(size=3)))
    s._flush()
    return s


def _pairs(fn, s, with_ws):
    import torch

    from xpbd_amd import _native as N

    b, n = s._buf, s.num_bodies
    cap = max(64, min(n * (n - 1) // 2, 200 * n))
    pairs = torch.full((cap, 2), -7, dtype=torch.int32, device=DEV)
    pc = torch.full((1,), -7, dtype=torch.int32, device=DEV)
    args = [N.ptr(b[k]) for k in ("X", "Q", "V", "W", "HE")]
    args += [s._man_margin, s.dt, n, cap, N.ptr(pairs), N.ptr(pc)]
    if with_ws:
        args = [N.ptr(b["WS"])] + args
    N.check(fn(N.current_stream(), *args), fn.__name__)
    torch.cuda.synchronize()
    cnt = int(pc.item())
    assert 0 < cnt <= cap
    return cnt, pairs[:cnt].cpu().numpy()


def test_noalloc_broadphase_matches_workspace_entry_point():
    from xpbd_amd import _native as N

    lib = N.lib()
    for n, seed, spread in ((6000, 2, 3.0), (300, 1, 0.2)):
        s = _boxes(n, seed, spread)
        c_ref, p_ref = _pairs(lib.xpbd6_broadphase_lbvh, s, True)
        c_new, p_new = _pairs(lib.xpbd6_broadphase_lbvh_noalloc, s, False)
        assert c_new == c_ref, n
        assert np.array_equal(p_new, p_ref), n
        print(f"noalloc broad phase, {n} bodies: {c_new} pairs, identical")


def test_noalloc_broadphase_rejects_more_bodies_than_its_scratch():
    """65537 bodies: refused on the host with the library's own error code,
    before any launch (the pointers are never touched)."""
    from xpbd_amd import _native as N

    lib = N.lib()
    rc = lib.xpbd6_broadphase_lbvh_noalloc(N.current_stream(), None, None, None, None, None,
                                           0.02, 1 / 60, 65537, 64, None, None)
    assert rc == -2
    assert "65536" in lib.xpbd_error_string(rc).decode()
    with pytest.raises(RuntimeError, match="allocation-free"):
        N.check(rc, "xpbd6_broadphase_lbvh_noalloc")
