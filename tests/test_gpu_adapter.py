"""The C++ renderer adapter (openshadinglanguage_amd/adapter/
oslmi_batched.hpp), compiled into tests/adapter/adapter_demo.cpp's program
by the build, run as a renderer would run it: the layered3 group built
through ShaderGroupBegin / Shader / ConnectShaders / ShaderGroupEnd,
outputs placed into the program's own 32-byte pixel records by add_symlocs
(output arena, offset + shadeindex * stride), BatchedShaderGlobals<16>
blocks in the reference's byte layout, shaded batch by batch through the
reference-shaped execute() and all at once through execute_frame().

Checked: both paths write the same bytes; every record holds the values
the oracle computes for the point whose shade index names it (the symbols
pattern.fbm, color.Cd and pattern.gab, within 4 ulp or 1e-5 relative; the
emitted kernel runs the oracle's operations, so they are expected to be
bit-identical); the bytes the shader does not own (a tag field) survive.
"""
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
from openshadinglanguage_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "openshadinglanguage_amd")
DEMO = os.path.join(PKG, "_lib", "oslmi_adapter_demo")
PIXEL = np.dtype([("fbm", "<f4"), ("Cd", "<f4", 3), ("gab", "<f4"), ("tag", "<i4"),
                  ("pad", "<f4", 2)])


def ulp_close(got, ref, ulps=4, rel=1e-5):
    got = np.asarray(got, np.float32)
    ref = np.asarray(ref, np.float32)
    gi = got.view(np.int32).astype(np.int64)
    ri = ref.view(np.int32).astype(np.int64)
    gi = np.where(gi < 0, -(gi & 0x7fffffff), gi)
    ri = np.where(ri < 0, -(ri & 0x7fffffff), ri)
    d = np.abs(got.astype(np.float64) - ref)
    return (np.abs(gi - ri) <= ulps) | (d <= rel * np.abs(ref.astype(np.float64))) | (got == ref)


@pytest.mark.parametrize("n", [1, 16, 4099])
def test_adapter_program(tmp_path, n):
    assert os.path.exists(DEMO), "adapter program not built (run __graft_entry__.build())"
    w = W.layered3_inputs(res=1024, seed=12, n=n)
    perm = np.random.default_rng(n).permutation(n).astype(np.int32)
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(np.int64(n).tobytes())
        f.write(np.ascontiguousarray(w["P"], np.float32).tobytes())
        f.write(np.ascontiguousarray(w["N"], np.float32).tobytes())
        f.write(perm.tobytes())
    out = tmp_path / "out.bin"
    print("adapter program: n=%d" % n, flush=True)
    r = subprocess.run([DEMO, os.path.join(PKG, "csrc"), os.path.join(PKG, "_lib", "jitcache"),
                        os.path.join(PKG, "shaders"), str(inp), str(out)],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr + r.stdout
    px = np.fromfile(out, dtype=PIXEL)
    assert px.size == 2 * n
    per_batch, frame = px[:n], px[n:]
    assert per_batch.tobytes() == frame.tobytes()
    np.testing.assert_array_equal(frame["tag"], 7 * np.arange(n) + 3)
    _, aov = O.shade_c2(w["P"], w["N"], w["wo"], w["wi"], w["dPdx"], w["dPdy"], nthreads=8,
                        debug=True)
    rec = frame[perm]  # record of point i is at its shade index
    assert ulp_close(rec["fbm"], aov[0]).all()
    assert ulp_close(rec["gab"], aov[12]).all()
    assert ulp_close(rec["Cd"].T, aov[15:18]).all()
    print("exact: fbm %.4f gab %.4f Cd %.4f" % (np.mean(rec["fbm"] == aov[0]),
                                               np.mean(rec["gab"] == aov[12]),
                                               np.mean(rec["Cd"].T == aov[15:18])))
