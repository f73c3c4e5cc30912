"""GPU parity of the wide noise shadeops against the CPU oracle.

Bar: integer outputs (hash, and cellnoise/hashnoise whose float is a pure
function of the integer hash) bit-exact; float noise within 4 ulp or 1e-5
relative, whichever is larger (both paths evaluate the same IEEE operations
in the same order, so in practice they agree exactly).
"""
import zlib

import numpy as np
import pytest

import _oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _ulp_close(got, ref, ulps=4, rel=1e-5):
    got = got.astype(np.float32)
    ref = ref.astype(np.float32)
    gi = got.view(np.int32).astype(np.int64)
    ri = ref.view(np.int32).astype(np.int64)
    gi = np.where(gi < 0, -(gi & 0x7fffffff), gi)
    ri = np.where(ri < 0, -(ri & 0x7fffffff), ri)
    ulp = np.abs(gi - ri)
    relerr = np.abs(got.astype(np.float64) - ref) / np.maximum(np.abs(ref.astype(np.float64)), 1e-30)
    ok = (ulp <= ulps) | (relerr <= rel) | (got == ref)
    return ok, ulp


def _inputs(indim, n, derivs, seed):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-50, 50, size=(indim, n)).astype(np.float32)
    # include lattice points, cell boundaries and negative zero
    p[:, :8] = np.array([0.0, -0.0, 1.0, -1.0, 0.9999, -0.0001, 3.5, -7.25], np.float32)[None, :]
    if derivs:
        d = rng.uniform(-1, 1, size=(2 * indim, n)).astype(np.float32)
        p = np.concatenate([p, d])
    return p


def _seed(*key):
    return zlib.crc32(repr(key).encode()) & 0xffff


@pytest.fixture(scope="module")
def dev():
    from openshadinglanguage_amd import check, lib
    check(lib().oslmi_init())
    return torch.device("cuda:0")


CASES = [(nm, od, idim, dv) for nm in ("perlin", "uperlin", "cell", "hash", "simplex", "usimplex")
         for od in (1, 3) for idim in (1, 2, 3, 4) for dv in (False, True)]


@pytest.mark.parametrize("name,outdim,indim,derivs", CASES)
def test_noise_parity(dev, name, outdim, indim, derivs):
    from openshadinglanguage_amd import shadeops
    n = 4099  # ragged: not a multiple of the 256-thread block
    inp = _inputs(indim, n, derivs, seed=_seed(name, outdim, indim, derivs))
    ref = O.noise(name, outdim, inp, derivs=derivs)
    got = shadeops.noise(name, outdim, torch.from_numpy(inp).to(dev), derivs=derivs).cpu().numpy()
    assert got.shape == ref.shape
    if name in ("cell", "hash"):
        np.testing.assert_array_equal(got.view(np.int32), ref.view(np.int32))
    else:
        ok, ulp = _ulp_close(got, ref)
        assert ok.all(), "max ulp %d at %s" % (ulp.max(), np.argwhere(~ok)[:4])


def test_hash_op_parity(dev):
    from openshadinglanguage_amd import shadeops
    for indim in (1, 2, 3, 4):
        inp = _inputs(indim, 1000, False, seed=indim)
        ref = O.hash_op(inp)
        got = shadeops.hash_op(torch.from_numpy(inp).to(dev)).cpu().numpy()
        np.testing.assert_array_equal(got, ref)


def test_empty_batch(dev):
    from openshadinglanguage_amd import shadeops
    out = shadeops.noise("perlin", 1, torch.zeros((3, 0), device=dev))
    assert out.shape == (1, 0)


PCASES = [(nm, od, idim, dv) for nm in ("perlin", "uperlin", "cell", "hash")
          for od in (1, 3) for idim in (1, 2, 3, 4) for dv in (False, True)]


@pytest.mark.parametrize("name,outdim,indim,derivs", PCASES)
def test_pnoise_parity(dev, name, outdim, indim, derivs):
    """Periodic noise (oslmi_pnoise) against the oracle. Periods per point
    and axis: whole and fractional values, values below 1 (taken as 1) and,
    for perlin, NaN (lattice period 1); coordinates of both signs so the
    negative remainders of imod are exercised."""
    from openshadinglanguage_amd import shadeops
    n = 4099
    seed = _seed("p", name, outdim, indim, derivs)
    inp = _inputs(indim, n, derivs, seed=seed)
    rng = np.random.default_rng(seed + 1)
    per = rng.choice(np.array([0.25, 1.0, 2.0, 3.0, 3.7, 5.0, 8.0, 13.5, -2.0], np.float32),
                     size=(indim, n))
    if name in ("perlin", "uperlin"):
        per[:, 100:140] = np.nan
    ref = O.pnoise(name, outdim, inp, per, derivs=derivs)
    got = shadeops.pnoise(name, outdim, torch.from_numpy(inp).to(dev),
                          torch.from_numpy(per).to(dev), derivs=derivs).cpu().numpy()
    assert got.shape == ref.shape
    if name in ("cell", "hash"):
        np.testing.assert_array_equal(got.view(np.int32), ref.view(np.int32))
    else:
        ok, ulp = _ulp_close(got, ref)
        assert ok.all(), "max ulp %d at %s" % (ulp.max(), np.argwhere(~ok)[:4])


def test_pnoise_rejects_simplex(dev):
    from openshadinglanguage_amd import lib, shadeops
    x = torch.zeros((1, 4), device=dev)
    with pytest.raises(Exception):
        shadeops.pnoise("simplex", 1, x, torch.ones((1, 4), device=dev))
    assert b"unknown" in lib().oslmi_last_error()
