"""GPU parity of named-space transforms in emitted group kernels: the
transform table the shading system builds from the renderer's named spaces,
transform / transformv / transformn by names and by matrices, matrix(from,
to) and getmatrix, against the oracle's interpreter on the same points
(bit-identical expected: the same binary32 operations in the same order;
stated tolerance 4 ulp / 1e-5 rel), the reference's golden output, and the
derivatives of transformed points against their closed form."""
import numpy as np
import pytest

import _oracle as O
from openshadinglanguage_amd import groupgen as G
from openshadinglanguage_amd import shadingsys as S
from test_transforms import GOLDEN, INPUTS, MOBJ, MSHAD, transform_shader

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from openshadinglanguage_amd import check, lib
    check(lib().oslmi_init())
    return torch.device("cuda:0")


def spaces():
    rng = np.random.default_rng(11)
    proj = np.eye(4, dtype=np.float32)
    proj[:3, :3] += rng.normal(0, 0.3, (3, 3)).astype(np.float32)
    proj[:3, 3] = rng.uniform(-0.05, 0.05, 3).astype(np.float32)
    proj[3] = (0.5, -0.25, 1.0, 1.2)
    return {"shader": MSHAD, "object": MOBJ,
            "myspace": np.diag([1.0, 2.0, 1.0, 1.0]).astype(np.float32), "proj": proj}


def run_gpu(dev, oso, P, sp, dPdx=(0.0, 0.0, 0.0), dPdy=(0.0, 0.0, 0.0)):
    ss = S.ShadingSystem()
    for name, m in sp.items():
        ss.name_transform(name, m)
    g = S.build_generated_group(ss, ([(oso.name, "l0", oso.text())], []), oso.name)
    ss.jit_group(g)
    n = P.shape[1]
    print("transform group %s: n=%d" % (oso.name, n), flush=True)
    Pt = torch.from_numpy(np.ascontiguousarray(P)).to(dev)
    out = torch.empty((4, n), device=dev)
    ss.execute(g, n, {"P": Pt, "N": Pt, "dPdx": dPdx, "dPdy": dPdy}, outputs=out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


CASES = [(k, f, t, m) for k in ("point", "vector", "normal")
         for f, t, m in ((None, "shader", False), ("shader", "common", False),
                         ("shader", "object", False), ("object", "proj", False),
                         ("proj", "world", False), ("common", "proj", False),
                         ("myspace", "nowhere", False), ("world", "common", False),
                         ("proj", "myspace", True), ("shader", "object", True))]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s-%s-%s" % c)
def test_transforms_match_interpreter(dev, case):
    kind, frm, to, mi = case
    rng = np.random.default_rng(7)
    P = rng.uniform(-4, 4, (3, 1000)).astype(np.float32)
    sp = spaces()
    oso = transform_shader(kind, frm, to, matrix_input=mi)
    got = run_gpu(dev, oso, P, sp)[1:4]
    g = O.Group([(oso.name, "l0", oso.text())], [])
    for name, m in sp.items():
        g.name_transform(name, m)
    ref = g.outputs("fout", "cout").run(P, P)[0][1:4]
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("case", GOLDEN, ids=lambda c: "%s-%s-%s-%s" % c[:4])
def test_transforms_match_reference_golden(dev, case):
    kind, frm, to, inp, want = case
    P = np.array(INPUTS[inp], np.float32).reshape(3, 1)
    got = run_gpu(dev, transform_shader(kind, frm, to), P, {"shader": MSHAD, "object": MOBJ})
    np.testing.assert_allclose(got[1:4, 0], want, rtol=1e-4, atol=2e-5)


def deriv_shader(to, which):
    o = G.Oso("shader", "xfd_%s_%s" % (to, which))
    o.sym("global", "point", "P")
    o.sym("oparam", "float", "fout", "0")
    o.sym("oparam", "color", "cout", "0 0 0")
    r = o.tmp("point")
    o.op("transform", r, o.const("string", to), "P")
    d = o.tmp("vector")
    o.op(which, d, r)
    o.op("assign", "cout", d)
    return o


@pytest.mark.parametrize("to", ["myspace", "proj"])
def test_transform_derivatives_closed_form(dev, to):
    rng = np.random.default_rng(9)
    P = rng.uniform(-2, 2, (3, 512)).astype(np.float32)
    dPdx, dPdy = (0.01, 0.002, -0.003), (-0.004, 0.02, 0.001)
    sp = spaces()
    Minv = np.linalg.inv(sp[to].astype(np.float64))
    x = np.vstack([P.astype(np.float64), np.ones((1, 512))]).T
    h = x @ Minv
    for which, d in (("Dx", dPdx), ("Dy", dPdy)):
        got = run_gpu(dev, deriv_shader(to, which), P, sp, dPdx=dPdx, dPdy=dPdy)[1:4]
        dh = np.asarray(d, np.float64) @ Minv[:3, :]  # derivative of (a, b, c, w)
        want = ((dh[None, :3] * h[:, 3:4] - h[:, :3] * dh[None, 3:4]) / h[:, 3:4] ** 2).T
        np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-6, err_msg=which)


def test_transforms_and_attributes_in_a_batch(dev):
    """The same groups in one persistent batch launch (oslmi_batch_*) read
    the shading system's table through each item's GroupArgs: outputs are
    bit-identical to the groups' own kernels."""
    from test_attributes import ATTRS, attr_shader
    sp = spaces()
    ss = S.ShadingSystem()
    for name, m in sp.items():
        ss.name_transform(name, m)
    for name, typ, v in ATTRS:
        ss.set_attribute(name, typ, v)
    osos = [transform_shader(*c) for c in CASES[::3]] + [attr_shader()]
    groups = [S.build_generated_group(ss, ([(o.name, "l0", o.text())], []), "b_" + o.name)
              for o in osos]
    rng = np.random.default_rng(4)
    items, outs, alone = [], [], []
    for k, g in enumerate(groups):
        n = 300 + 50 * k
        P = torch.from_numpy(rng.uniform(-3, 3, (3, n)).astype(np.float32)).to(dev)
        glob = {"P": P, "N": P, "dPdx": (0.0, 0.0, 0.0), "dPdy": (0.0, 0.0, 0.0)}
        o = torch.empty((4, n), device=dev)
        items.append((n, glob, o, None, None, None))
        outs.append(o)
        ss.jit_group(g)
        a = torch.empty((4, n), device=dev)
        ss.execute(g, n, glob, outputs=a)
        alone.append(a)
    print("transform batch: %d groups" % len(groups), flush=True)
    b = S.GroupBatch(ss, groups, groups_per_kernel=4)
    b.bind(items)
    b.launch()
    torch.cuda.synchronize()
    for k in range(len(groups)):
        np.testing.assert_array_equal(outs[k].cpu().numpy(), alone[k].cpu().numpy(),
                                      err_msg=osos[k].name)
