"""GPU parity of periodic and vector noise in emitted kernels: pnoise /
psnoise with and without a name (periodic perlin, cell, hash), the periodic
gabor (pgabor / pgabor3) and the vector gabor (gabor3), in 1-4 dimensions
with float and color results, against the oracle's interpreter. Periods vary
per point. Bar: bit for bit (the emitted code evaluates the interpreter's
IEEE operations in the same order)."""
import numpy as np
import pytest

import _oracle as O
import _pnprogs as PN
from openshadinglanguage_amd import shadingsys as S

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from openshadinglanguage_amd import check, lib
    check(lib().oslmi_init())
    return torch.device("cuda:0")


@pytest.mark.parametrize("k", range(len(PN.SPECS)), ids=[p[0] for p in PN.SPECS])
def test_periodic_noise_programs_match_interpreter(dev, k):
    name, text = PN.programs()[k]
    n = 2000 if "gabor" in name else 20000
    inp = PN.inputs(n, 90 + k)
    ss = S.ShadingSystem()
    g = S.build_generated_group(ss, ([(name, "surface", text)], []), name, outputs=PN.OUTPUTS)
    ss.jit_group(g)
    assert ss.output_planes(g) == 16
    glob = {key: torch.from_numpy(np.ascontiguousarray(inp[key])).to(dev)
            for key in ("P", "N", "dPdu", "dPdv")}
    glob.update({"dPdx": (0.0, 0.0, 0.0), "dPdy": (0.0, 0.0, 0.0)})
    out = torch.full((16, n), -7.0, device=dev)
    print("periodic noise %s: %d points" % (name, n), flush=True)
    ss.execute(g, n, glob, outputs=out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    gr = O.Group([(name, "surface", text)], [])
    for d in range(1, 5):
        gr.outputs("F%d" % d, "C%d" % d)
        ref, _, _ = O.group_run_ex(gr, inp["P"], inp["N"], np.zeros((3, n), np.float32),
                                   dPdu=inp["dPdu"], dPdv=inp["dPdv"], nthreads=8)
        np.testing.assert_array_equal(got[d - 1], ref[0], err_msg="%s F%d" % (name, d))
        c0 = 4 + 3 * (d - 1)
        np.testing.assert_array_equal(got[c0:c0 + 3], ref[1:4], err_msg="%s C%d" % (name, d))
