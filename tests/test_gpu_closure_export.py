"""GPU parity of the closure export (oslmi_execute_batch_closures): Ci
handed out as testrender's list of components (closure id, weight rgb,
parameters in testrender's registration order, src/testrender/
shading.cpp:185-296), against the oracle's interpreter walking the same
closure trees (process_bsdf_closure). Counts and ids exact; weights and
parameters within 4 ulp or 1e-5 relative (the emitted code forms weights
in another multiplication order)."""
import numpy as np
import pytest

import _mxprogs as X
import _oracle as O
from openshadinglanguage_amd import OSLMIError
from openshadinglanguage_amd import shadingsys as S
from test_gpu_batch import close

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

CASES = [(name, o.text()) for name, o in X.programs()] + [
    ("layered3_surface", open(S.os.path.join(S.SHADER_DIR, "layered3_surface.oso")).read())]


@pytest.fixture(scope="module")
def dev():
    from openshadinglanguage_amd import check, lib
    check(lib().oslmi_init())
    return torch.device("cuda:0")


@pytest.mark.parametrize("k", range(len(CASES)), ids=[c[0] for c in CASES])
def test_exported_closures_match_interpreter(dev, k):
    name, text = CASES[k]
    inp = X.inputs(12000, 80 + k)
    n = inp["P"].shape[1]
    ss = S.ShadingSystem()
    ss.attribute("closure_export", 1)
    g = S.build_generated_group(ss, ([(name, "surface", text)], []), name, outputs=())
    t = {key: torch.from_numpy(np.ascontiguousarray(inp[key])).to(dev)
         for key in ("P", "N", "dPdu", "dPdv", "wo")}
    glob = {"P": t["P"], "N": t["N"], "dPdu": t["dPdu"], "dPdv": t["dPdv"],
            "dPdx": (0.0, 0.0, 0.0), "dPdy": (0.0, 0.0, 0.0)}
    out = torch.full((1 + 8 * S.CLOSURE_RECORD, n), -7.0, device=dev)
    print("closure export %s: n=%d" % (name, n), flush=True)
    ss.closures(g, n, glob, out, max_lobes=8, wo=t["wo"])
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ref = O.group_closures(O.Group([(name, "surface", text)], []), inp["P"], inp["N"], inp["wo"],
                           dPdu=inp["dPdu"], dPdv=inp["dPdv"])
    assert np.array_equal(got[0], ref[0]), name
    cnt = ref[0].astype(int)
    for j in range(int(cnt.max())):
        rows = slice(1 + j * S.CLOSURE_RECORD, 1 + (j + 1) * S.CLOSURE_RECORD)
        live = cnt > j
        assert np.array_equal(got[rows][0][live], ref[rows][0][live]), (name, j)  # ids
        ok = close(got[rows][:, live], ref[rows][:, live])
        assert ok.all(), (name, j, np.argwhere(~ok)[:3].tolist())
    print("%s: exact fraction %.4f" % (name, np.mean(got[:, cnt > 0] == ref[:, cnt > 0])))


def test_export_needs_the_attribute(dev):
    name, text = CASES[0]
    ss = S.ShadingSystem()
    g = S.build_generated_group(ss, ([(name, "surface", text)], []), name, outputs=())
    inp = X.inputs(64, 5)
    t = {key: torch.from_numpy(np.ascontiguousarray(inp[key])).to(dev) for key in ("P", "N")}
    out = torch.zeros((1 + S.CLOSURE_RECORD, 64), device=dev)
    with pytest.raises(OSLMIError, match="closure_export"):
        ss.closures(g, 64, {"P": t["P"], "N": t["N"]}, out, max_lobes=1)
