"""GPU parity of the prebuilt group kernels (fbm_cell, layered3) against the
CPU oracle on the same seeded inputs.

Tolerance (stated per the brief): float outputs within 4 ulp or 1e-5
relative, whichever is larger; the kernels evaluate the oracle's operations
in the same order, so the observed difference is expected to be zero and
the exact-match fraction is asserted too.
"""
import numpy as np
import pytest

import _oracle as O
from openshadinglanguage_amd import workloads as W

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def ulp_ok(got, ref, ulps=4, rel=1e-5):
    got = np.asarray(got, np.float32)
    ref = np.asarray(ref, np.float32)
    gi = got.view(np.int32).astype(np.int64)
    ri = ref.view(np.int32).astype(np.int64)
    gi = np.where(gi < 0, -(gi & 0x7fffffff), gi)
    ri = np.where(ri < 0, -(ri & 0x7fffffff), ri)
    relerr = np.abs(got.astype(np.float64) - ref) / np.maximum(np.abs(ref.astype(np.float64)), 1e-30)
    return (np.abs(gi - ri) <= ulps) | (relerr <= rel) | (got == ref)


@pytest.fixture(scope="module")
def dev():
    from openshadinglanguage_amd import check, lib
    check(lib().oslmi_init())
    return torch.device("cuda:0")


def _run_fbm_cell(dev, P):
    import ctypes
    from openshadinglanguage_amd import check, lib
    d = torch.from_numpy(P).to(dev)
    out = torch.empty_like(d)
    check(lib().oslmi_shade_fbm_cell(P.shape[1], ctypes.c_void_p(d.data_ptr()),
                                     ctypes.c_void_p(out.data_ptr()), None))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _run_layered3(dev, w, aov=True):
    import ctypes
    from openshadinglanguage_amd import check, lib
    n = w["P"].shape[1]
    t = {k: torch.from_numpy(w[k]).to(dev) for k in ("P", "N", "wo", "wi")}
    out = torch.empty((4, n), device=dev)
    a = torch.empty((20, n), device=dev) if aov else None
    p = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
    check(lib().oslmi_shade_layered3(
        n, p(t["P"]), p(t["N"]), p(t["wo"]), p(t["wi"]),
        w["dPdx"].ctypes.data_as(ctypes.c_void_p), w["dPdy"].ctypes.data_as(ctypes.c_void_p),
        p(out), p(a) if aov else None, None))
    torch.cuda.synchronize()
    return out.cpu().numpy(), (a.cpu().numpy() if aov else None)


def test_fbm_cell_full_workload(dev):
    w = W.fbm_cell_inputs(256, seed=1)
    ref = O.shade_c1(w["P"], nthreads=8)
    got = _run_fbm_cell(dev, w["P"])
    assert ulp_ok(got, ref).all()
    assert np.mean(got == ref) == 1.0


@pytest.mark.parametrize("n", [1, 255, 4099])
def test_fbm_cell_ragged(dev, n):
    rng = np.random.default_rng(n)
    P = rng.uniform(-300, 300, size=(3, n)).astype(np.float32)
    np.testing.assert_array_equal(_run_fbm_cell(dev, P), O.shade_c1(P))


def test_layered3_parity(dev):
    w = W.layered3_inputs(res=1024, seed=2, n=20000)
    ref, ref_aov = O.shade_c2(w["P"], w["N"], w["wo"], w["wi"], w["dPdx"], w["dPdy"],
                              nthreads=8, debug=True)
    got, aov = _run_layered3(dev, w)
    names = ["fbm", "fbm.dx", "fbm.dy", "vn.x", "vn.y", "vn.z", "vn.dx.x", "vn.dx.y", "vn.dx.z",
             "vn.dy.x", "vn.dy.y", "vn.dy.z", "gab", "gab.dx", "gab.dy", "Cd.r", "Cd.g", "Cd.b",
             "alpha"]
    for k, nm in enumerate(names):
        ok = ulp_ok(aov[k], ref_aov[k])
        assert ok.all(), "%s: %d mismatches, first %s" % (nm, (~ok).sum(), np.argwhere(~ok)[:3].ravel())
    ok = ulp_ok(got, ref)
    assert ok.all(), "eval: %d mismatches" % (~ok).sum()
    assert np.mean(got == ref) > 0.999


def test_layered3_no_aov_matches_aov(dev):
    w = W.layered3_inputs(res=1024, seed=2, n=3001)
    a, _ = _run_layered3(dev, w, aov=True)
    b, _ = _run_layered3(dev, w, aov=False)
    np.testing.assert_array_equal(a, b)
