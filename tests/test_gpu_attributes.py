"""GPU parity of getattribute in an emitted kernel against the oracle's
interpreter (bit-identical expected) and the closed form."""
import numpy as np
import pytest

import _oracle as O
from openshadinglanguage_amd import shadingsys as S
from test_attributes import ATTRS, attr_shader, expected

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_getattribute_matches_interpreter():
    from openshadinglanguage_amd import check, lib
    check(lib().oslmi_init())
    dev = torch.device("cuda:0")
    P = np.random.default_rng(2).uniform(-2, 2, (3, 5000)).astype(np.float32)
    ss = S.ShadingSystem()
    for name, typ, v in ATTRS:
        ss.set_attribute(name, typ, v)
    o = attr_shader()
    g = S.build_generated_group(ss, ([(o.name, "l0", o.text())], []), "attrs")
    ss.jit_group(g)
    print("getattribute group: n=%d" % P.shape[1], flush=True)
    Pt = torch.from_numpy(P).to(dev)
    out = torch.empty((4, P.shape[1]), device=dev)
    ss.execute(g, P.shape[1], {"P": Pt, "N": Pt}, outputs=out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ig = O.Group([(o.name, "l0", o.text())], [])
    for name, typ, v in ATTRS:
        ig.set_attribute(name, typ, v)
    ref, _ = ig.outputs("fout", "cout").run(P, P)
    np.testing.assert_array_equal(got, ref)
    fout, cout = expected(P.astype(np.float64))
    np.testing.assert_allclose(got[0], fout, rtol=1e-6, atol=1e-5)
    np.testing.assert_allclose(got[1:4], cout, rtol=1e-6, atol=1e-5)
