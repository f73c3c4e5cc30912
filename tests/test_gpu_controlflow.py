"""GPU parity of emitted control flow and closure dataflow: the programs of
tests/_cfprogs.py (if/else chains, for / while / do-while with trip counts
0..8, break, continue, nested loops, return from an inlined function, exit;
closures reused in an add, grown in a loop, and Ci shared between layers)
run by emitted group kernels, against the oracle's interpreter (the same
binary32 operations: bit-identical expected; stated tolerance 4 ulp or 1e-5
relative) and against each program's closed form."""
import numpy as np
import pytest

import _cfprogs as C
import _oracle as O
from openshadinglanguage_amd import shadingsys as S

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

N_PTS = 5000


@pytest.fixture(scope="module")
def dev():
    from openshadinglanguage_amd import check, lib
    check(lib().oslmi_init())
    return torch.device("cuda:0")


@pytest.fixture(scope="module")
def pts():
    return C.inputs(N_PTS, 9)


def run_gpu(dev, osos, conns, pts, name, eval_=False):
    ss = S.ShadingSystem()
    layers = [(o.name, "l%d" % k, o.text()) for k, o in enumerate(osos)]
    g = S.build_generated_group(ss, (layers, conns), name)
    ss.jit_group(g)
    n = pts["P"].shape[1]
    print("control-flow group %s: n=%d eval=%d" % (name, n, eval_), flush=True)
    t = {k: torch.from_numpy(np.ascontiguousarray(pts[k])).to(dev) for k in ("P", "N", "wo", "wi")}
    out = torch.full((4, n), -7.0, device=dev)
    ev = torch.full((4, n), -7.0, device=dev) if eval_ else None
    ss.execute(g, n, {"P": t["P"], "N": t["N"], "dPdx": (0.0, 0.0, 0.0), "dPdy": (0.0, 0.0, 0.0)},
               outputs=out, wo=t["wo"] if eval_ else None, wi=t["wi"] if eval_ else None,
               eval_out=ev)
    torch.cuda.synchronize()
    return out.cpu().numpy(), (ev.cpu().numpy() if eval_ else None)


@pytest.mark.parametrize("name", sorted(C.PROGRAMS))
def test_program_matches_interpreter_and_closed_form(dev, pts, name):
    o, want = C.PROGRAMS[name]()
    got, _ = run_gpu(dev, [o], [], pts, "cf_" + name)
    ref, _ = O.Group([(o.name, "l0", o.text())], []).outputs("fout", "cout").run(pts["P"], pts["N"])
    np.testing.assert_array_equal(got, ref)
    n, a, b = pts["nab"]
    np.testing.assert_array_equal(got[0], np.asarray(want(n, a, b), np.float32))


@pytest.mark.parametrize("name", sorted(C.CLOSURE_PROGRAMS))
def test_closure_program_matches_interpreter(dev, pts, name):
    o, want = C.CLOSURE_PROGRAMS[name]()
    _, ev = run_gpu(dev, [o], [], pts, "cl_" + name, eval_=True)
    _, ref = O.Group([(o.name, "l0", o.text())], []).outputs("fout", "cout").run(
        pts["P"], pts["N"], pts["wo"], pts["wi"])
    np.testing.assert_array_equal(ev, ref)
    n, a, b = pts["nab"]
    for c in range(3):
        np.testing.assert_allclose(ev[c], want(n, a, b), rtol=2e-6)


def test_ci_written_by_upstream_layer(dev, pts):
    osos, conns, want = C.closure_upstream_layers()
    out, ev = run_gpu(dev, osos, conns, pts, "cl_upstream", eval_=True)
    layers = [(o.name, "l%d" % k, o.text()) for k, o in enumerate(osos)]
    rout, rev = O.Group(layers, conns).outputs("fout", "cout").run(
        pts["P"], pts["N"], pts["wo"], pts["wi"])
    np.testing.assert_array_equal(out, rout)
    np.testing.assert_array_equal(ev, rev)
    n, a, b = pts["nab"]
    for c in range(3):
        np.testing.assert_allclose(ev[c], want(n, a, b), rtol=2e-6)
