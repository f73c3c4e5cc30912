"""Named coordinate systems (renderer services get_matrix /
get_inverse_matrix) and the transform / transformv / transformn, matrix and
getmatrix ops: the oracle's interpreter against the reference's golden
output (testsuite/transform/ref/out.txt, testshade's shader and object
spaces) and against numpy's float64 linear algebra, plus emission and the
hiprtc compile of groups that use them (no GPU needed)."""
import numpy as np
import pytest

import _oracle as O
from openshadinglanguage_amd import OSLMIError
from openshadinglanguage_amd import groupgen as G
from openshadinglanguage_amd import shadingsys as S


def imath_rotate_z_translate(angle, t):
    """testshade's setup_transformations (src/testshade/testshade.cpp:
    925-950): makeIdentity, translate(t), rotate((0, 0, angle)) in float."""
    c, s = np.cos(np.float32(angle)), np.sin(np.float32(angle))
    m = np.eye(4, dtype=np.float32)
    m[0, :2] = (c, s)
    m[1, :2] = (-s, c)
    m[3, :3] = t
    return m


MSHAD = imath_rotate_z_translate(np.pi / 4, (1.0, 0.0, 0.0))
MOBJ = imath_rotate_z_translate(np.pi / 2, (0.0, 1.0, 0.0))


def transform_shader(kind, frm, to, matrix_input=False):
    """One layer: cout = transform{,v,n}([frm,] to, P) (or by the matrix
    'matrix(frm, to)' when matrix_input), fout = 0."""
    o = G.Oso("shader", "xf_%s_%s_%s" % (kind, frm or "c", to))
    o.sym("global", "point", "P")
    o.sym("oparam", "float", "fout", "0")
    o.sym("oparam", "color", "cout", "0 0 0")
    op = {"point": "transform", "vector": "transformv", "normal": "transformn"}[kind]
    src = o.tmp(kind)
    o.op("assign", src, "P")
    r = o.tmp(kind)
    if matrix_input:
        M = o.tmp("matrix")
        o.op("matrix", M, o.const("string", frm), o.const("string", to))
        o.op(op, r, M, src)
    elif frm is None:
        o.op(op, r, o.const("string", to), src)
    else:
        o.op(op, r, o.const("string", frm), o.const("string", to), src)
    o.op("assign", "cout", r)
    return o


def run_oracle(oso, P, spaces):
    g = O.Group([(oso.name, "l0", oso.text())], [])
    for name, m in spaces.items():
        g.name_transform(name, m)
    out, _ = g.outputs("fout", "cout").run(P, np.zeros_like(P))
    return out[1:4]


# testsuite/transform/ref/out.txt (printed with %.5g): p = (1, 1.41421, 0),
# z = (0, 0, 0); (kind, from, to, input) -> result
GOLDEN = [
    ("point", None, "shader", "p", (1, 1, 0)),
    ("vector", None, "shader", "p", (1.7071, 0.29289, 0)),
    ("normal", None, "shader", "p", (1.7071, 0.29289, 0)),
    ("point", None, "shader", "z", (-0.70711, 0.70711, 0)),
    ("vector", None, "shader", "z", (0, 0, 0)),
    ("point", "common", "shader", "p", (1, 1, 0)),
    ("vector", "common", "shader", "p", (1.7071, 0.29289, 0)),
    ("normal", "common", "shader", "p", (1.7071, 0.29289, 0)),
    ("point", "shader", "common", "p", (0.70711, 1.7071, 0)),
    ("vector", "shader", "common", "p", (-0.29289, 1.7071, 0)),
    ("normal", "shader", "common", "p", (-0.29289, 1.7071, 0)),
    ("point", "shader", "common", "z", (1, 0, 0)),
    ("point", "shader", "object", "p", (0.7071, -0.70711, 0)),
    ("vector", "shader", "object", "p", (1.7071, 0.29289, 0)),
    ("normal", "shader", "object", "p", (1.7071, 0.29289, 0)),
    ("point", "shader", "object", "z", (-1, -1, 0)),
]
INPUTS = {"p": (1.0, 1.41421, 0.0), "z": (0.0, 0.0, 0.0)}


@pytest.mark.parametrize("case", GOLDEN, ids=lambda c: "%s-%s-%s-%s" % c[:4])
def test_oracle_matches_reference_golden(case):
    kind, frm, to, inp, want = case
    P = np.array(INPUTS[inp], np.float32).reshape(3, 1)
    got = run_oracle(transform_shader(kind, frm, to), P, {"shader": MSHAD, "object": MOBJ})
    # the golden text has 5 significant digits
    np.testing.assert_allclose(got[:, 0], want, rtol=1e-4, atol=2e-5)


def _random_spaces(rng):
    a = rng.normal(size=(3, 3)) + 2 * np.eye(3)
    aff = np.eye(4)
    aff[:3, :3] = a
    aff[3, :3] = rng.uniform(-2, 2, 3)
    proj = aff.copy()
    proj[:3, 3] = rng.uniform(-0.05, 0.05, 3)
    proj[3, 3] = 1.2
    return aff.astype(np.float32), proj.astype(np.float32)


@pytest.mark.parametrize("kind", ["point", "vector", "normal"])
def test_oracle_inverse_transforms_match_float64(kind):
    rng = np.random.default_rng(3)
    aff, proj = _random_spaces(rng)
    P = rng.uniform(-3, 3, (3, 64)).astype(np.float32)
    x = np.vstack([P.astype(np.float64), np.ones((1, 64))]).T  # rows (x, y, z, 1)
    spaces = {"aff": aff, "proj": proj}
    for name, m in spaces.items():
        m64 = m.astype(np.float64)
        inv = np.linalg.inv(m64)
        for frm, to, M in ((name, "common", m64), ("common", name, inv), ("world", name, inv)):
            got = run_oracle(transform_shader(kind, frm, to), P, spaces)
            if kind == "point":
                h = x @ M
                want = (h[:, :3] / h[:, 3:4]).T
            elif kind == "vector":
                want = (x[:, :3] @ M[:3, :3]).T
            else:
                want = (x[:, :3] @ np.linalg.inv(M).T[:3, :3]).T
            np.testing.assert_allclose(got, want, rtol=2e-4, atol=2e-5,
                                       err_msg="%s %s->%s" % (kind, frm, to))


def test_oracle_from_to_is_product_and_unknown_space_is_identity():
    rng = np.random.default_rng(5)
    aff, proj = _random_spaces(rng)
    P = rng.uniform(-3, 3, (3, 16)).astype(np.float32)
    spaces = {"a": aff, "b": proj}
    got = run_oracle(transform_shader("point", "a", "b"), P, spaces)
    M = aff.astype(np.float64) @ np.linalg.inv(proj.astype(np.float64))
    h = np.vstack([P, np.ones((1, 16))]).T @ M
    np.testing.assert_allclose(got, (h[:, :3] / h[:, 3:4]).T, rtol=2e-4, atol=2e-5)
    # by the matrix op: the same result
    got_m = run_oracle(transform_shader("point", "a", "b", matrix_input=True), P, spaces)
    np.testing.assert_allclose(got_m, got, rtol=1e-6, atol=1e-6)
    # an unknown space leaves the value unchanged
    for frm, to in (("nowhere", "common"), ("a", "nowhere"), ("common", "nowhere")):
        np.testing.assert_array_equal(run_oracle(transform_shader("point", frm, to), P, spaces), P)


def test_emit_and_compile_named_transforms(tmp_path):
    ss = S.ShadingSystem(jit_cache=str(tmp_path))
    for k, (kind, frm, to, mi) in enumerate((("point", None, "shader", False),
                                             ("normal", "object", "myspace", False),
                                             ("vector", "shader", "world", True))):
        o = transform_shader(kind, frm, to, matrix_input=mi)
        g = S.build_generated_group(ss, ([(o.name, "l0", o.text())], []), "xf%d" % k)
        src = ss.emit(g)
        assert "A.xf + " in src
        ss.compile_group(g)


def test_name_transform_rejects_common_space():
    ss = S.ShadingSystem()
    for bad in ("common", "world", ""):
        with pytest.raises(OSLMIError):
            ss.name_transform(bad, np.eye(4))
    ss.name_transform("myspace", np.diag([1.0, 2.0, 1.0, 1.0]))
