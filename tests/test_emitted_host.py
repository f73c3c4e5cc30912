"""CPU check of the emitter's closure lowering: each MaterialX closure
program's emitted kernel, compiled for the host (tests/hostkernel/
harness.cpp defines the device qualifiers away and runs the grid-stride
loop on one thread), evaluates its closure exactly as the oracle's
interpreter does (closure trees flattened by testrender's walk). The
kernel code and the interpreter differ in structure (lobe lists built
eagerly at each closure op against a tree walked afterwards), so a slip
in either (a null rule, an opacity, a walk that should stop) shows here
without a GPU. The GPU runs the same sources (test_gpu_mx_closure.py)."""
import os
import subprocess

import numpy as np
import pytest

import _mxprogs as X
import _oracle as O
import _pnprogs as P
from openshadinglanguage_amd import shadingsys as S

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "openshadinglanguage_amd", "csrc")
LUTS = os.path.join(HERE, "..", "oracle", "build", "bsdl_luts.bin")


def _host_run(tmp_path, name, text, inp, export, outputs=(), nplanes=0):
    ss = S.ShadingSystem()
    if export:
        ss.attribute("closure_export", 1)
    g = S.build_generated_group(ss, ([(name, "surface", text)], []), name, outputs=outputs)
    src = tmp_path / (name + ".hip")
    src.write_text(ss.emit(g))
    exe = tmp_path / "harness"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-I", CSRC,
                           "-I", str(tmp_path), '-DKERNEL="%s.hip"' % name, "-o", str(exe),
                           os.path.join(HERE, "hostkernel", "harness.cpp")])
    n = inp["P"].shape[1]
    fin = tmp_path / "in.bin"
    with open(fin, "wb") as f:
        f.write(np.int64(n).tobytes())
        for key in ("P", "N", "dPdu", "dPdv", "wo", "wi"):
            a = inp.get(key)
            f.write(np.ascontiguousarray(np.zeros((3, n)) if a is None else a, np.float32).tobytes())
    fout = tmp_path / "out.bin"
    O.mtx_tables()  # the oracle's table file exists once the oracle has loaded it
    mode = ["export"] if export else ["out%d" % nplanes] if nplanes else []
    subprocess.check_call([str(exe), str(fin), str(fout), LUTS] + mode)
    return np.fromfile(fout, np.float32).reshape(-1, n)


@pytest.mark.parametrize("k", range(len(X.programs())), ids=[p[0] for p in X.programs()])
def test_host_built_kernel_matches_interpreter(tmp_path, k):
    name, o = X.programs()[k]
    inp = X.inputs(1500, 60 + k)
    ev = _host_run(tmp_path, name, o.text(), inp, export=False)
    gr = O.Group([(name, "surface", o.text())], [])
    _, rev, _ = O.group_run_ex(gr, inp["P"], inp["N"], inp["wo"], dPdu=inp["dPdu"],
                               dPdv=inp["dPdv"], wi=inp["wi"])
    assert np.array_equal(ev, rev), (name, np.mean(ev == rev))


@pytest.mark.parametrize("k", range(len(X.programs())), ids=[p[0] for p in X.programs()])
def test_host_built_closure_export_matches_interpreter(tmp_path, k):
    """The exported lobe lists (testrender ids, weights, parameters) equal
    the interpreter's walk of the closure tree."""
    name, o = X.programs()[k]
    inp = X.inputs(800, 70 + k)
    got = _host_run(tmp_path, name, o.text(), inp, export=True)
    gr = O.Group([(name, "surface", o.text())], [])
    ref = O.group_closures(gr, inp["P"], inp["N"], inp["wo"], dPdu=inp["dPdu"], dPdv=inp["dPdv"])
    assert np.array_equal(got, ref), (name, np.argwhere(got != ref)[:4].tolist())
    assert (ref[0] >= 1).all()


@pytest.mark.parametrize("k", range(len(P.SPECS)), ids=[p[0] for p in P.SPECS])
def test_host_built_periodic_noise_matches_interpreter(tmp_path, k):
    """pnoise / psnoise with and without a name, periodic gabor and the
    vector gabor, 1-4 dimensions, float and color results: the emitted
    kernel's outputs equal the interpreter's bit for bit. Periods vary per
    point (N, dPdv), so the lattice wrap runs with runtime periods."""
    name, text = P.programs()[k]
    inp = P.inputs(600 if "gabor" in name else 3000, 80 + k)
    got = _host_run(tmp_path, name, text, inp, export=False, outputs=P.OUTPUTS, nplanes=16)
    gr = O.Group([(name, "surface", text)], [])
    n = inp["P"].shape[1]
    for d in range(1, 5):
        gr.outputs("F%d" % d, "C%d" % d)
        ref, _, _ = O.group_run_ex(gr, inp["P"], inp["N"], np.zeros((3, n), np.float32),
                                   dPdu=inp["dPdu"], dPdv=inp["dPdv"])
        np.testing.assert_array_equal(got[d - 1], ref[0], err_msg="%s F%d" % (name, d))
        c0 = 4 + 3 * (d - 1)
        np.testing.assert_array_equal(got[c0:c0 + 3], ref[1:4], err_msg="%s C%d" % (name, d))
