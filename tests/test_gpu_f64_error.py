"""Error of the GPU's binary32 shading results against a binary64
evaluation of the same path (the oracle's sources built in binary64 with
exact transcendentals, oracle/f64), reported beside the binary32 oracle's
own error against the same binary64 values.

Cases: fbm_cell 256x256, layered3 (65536 points of the 1024x1024 workload,
derivatives on, emitted group kernel) and the layered closure (65536
points, 1 eval + 4 samples). The GPU's error must be no larger than the
binary32 reference evaluation's; both summaries are printed and, when the
environment names a directory in OSLMI_REPORT_DIR, written there as
fp64_error.json.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import _oracle as O
from openshadinglanguage_amd import workloads as W

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

REPORT = {}


@pytest.fixture(scope="module")
def dev():
    from openshadinglanguage_amd import check, lib
    check(lib().oslmi_init())
    yield torch.device("cuda:0")
    d = os.environ.get("OSLMI_REPORT_DIR")
    if d and REPORT:
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, "fp64_error.json"), "w") as f:
            json.dump(REPORT, f, indent=1, sort_keys=True)


def _record(case, gpu, ref32):
    REPORT[case] = {"gpu": gpu, "oracle_f32": ref32}
    print("\n%s\n  gpu        %s\n  oracle f32 %s" % (case, json.dumps(gpu), json.dumps(ref32)),
          flush=True)


def _no_worse(gpu, ref32):
    for k in ("rel_p50", "rel_p99", "rel_p99_9", "ulp_p50", "ulp_p99"):
        assert gpu[k] <= ref32[k] * 1.05 + 1e-12, (k, gpu[k], ref32[k])
    assert gpu["frac_rel_le1e5"] >= ref32["frac_rel_le1e5"] - 1e-3
    if "frac_flipped" in gpu:
        assert gpu["frac_flipped"] <= ref32["frac_flipped"] + 1e-3


def test_fp64_error_fbm_cell(dev):
    from openshadinglanguage_amd import shadingsys as S
    ss, g = S.build_fbm_cell()
    ss.jit_group(g)
    w = W.fbm_cell_inputs(256, seed=1)
    n = w["P"].shape[1]
    print("fbm_cell: n=%d" % n, flush=True)
    out = torch.empty((3, n), device=dev)
    ss.execute(g, n, {"P": torch.from_numpy(w["P"]).to(dev)}, outputs=out)
    torch.cuda.synchronize()
    r64 = O.shade_c1_f64(w["P"], nthreads=8)
    gpu = O.error_summary(out.cpu().numpy(), r64)
    ref = O.error_summary(O.shade_c1(w["P"], nthreads=8), r64)
    _record("fbm_cell_256x256", gpu, ref)
    _no_worse(gpu, ref)


def test_fp64_error_layered3(dev):
    from openshadinglanguage_amd import shadingsys as S
    ss, g = S.build_layered3()
    ss.jit_group(g)
    n = 1 << 16
    w = W.layered3_inputs(res=1024, seed=2, n=n)
    print("layered3: n=%d" % n, flush=True)
    t = {k: torch.from_numpy(w[k]).to(dev) for k in ("P", "N", "wo", "wi")}
    ev = torch.empty((4, n), device=dev)
    ss.execute(g, n, {"P": t["P"], "N": t["N"], "dPdx": tuple(w["dPdx"]),
                      "dPdy": tuple(w["dPdy"])}, wo=t["wo"], wi=t["wi"], eval_out=ev)
    torch.cuda.synchronize()
    a = (w["P"], w["N"], w["wo"], w["wi"], w["dPdx"], w["dPdy"])
    r64 = O.shade_c2_f64(*a, nthreads=8)
    gpu = O.error_summary(ev.cpu().numpy(), r64)
    ref = O.error_summary(O.shade_c2(*a, nthreads=8), r64)
    _record("layered3_65536", gpu, ref)
    _no_worse(gpu, ref)


def test_fp64_error_layered_closure(dev):
    from openshadinglanguage_amd import check, lib
    n = 1 << 16
    w = W.layered_closure_inputs(n, seed=4)
    mat = W.layered_material()
    print("layered closure: n=%d" % n, flush=True)
    t = {k: torch.from_numpy(np.ascontiguousarray(w[k])).to(dev)
         for k in ("N", "V", "L", "params", "rnd")}
    ev = torch.empty((4, n), device=dev)
    sp = torch.empty((28, n), device=dev)
    p = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
    check(lib().oslmi_layered_closure(ctypes.byref(mat), n, p(t["N"]), p(t["V"]), p(t["L"]),
                                      p(t["params"]), p(t["rnd"]), p(ev), p(sp), None))
    torch.cuda.synchronize()
    a = (mat.as_list(), w["N"], w["V"], w["L"], w["params"], w["rnd"])
    ev64, sp64 = O.mtx_layered_f64(*a, nthreads=8)
    ev32, sp32 = O.mtx_layered(*a, nthreads=8)
    gpu = O.error_summary(ev.cpu().numpy(), ev64)
    ref = O.error_summary(ev32, ev64)
    _record("layered_closure_eval_65536", gpu, ref)
    _no_worse(gpu, ref)
    gpu = O.sample_error_summary(sp.cpu().numpy(), sp64)
    ref = O.sample_error_summary(sp32, sp64)
    _record("layered_closure_samples_65536x4", gpu, ref)
    _no_worse(gpu, ref)
