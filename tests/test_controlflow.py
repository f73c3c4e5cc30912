"""The oracle's OSO interpreter (the C3 batch's oracle) against closed
forms: if/else chains, for / while / do-while loops with data-dependent
trip counts 0..8, break and continue (also in nested loops), return from an
inlined function inside a loop, exit, and closure values built by add and
mul, reused, grown in a loop, and shared through Ci between layers.

Every closed form is written from the meaning of the OSL source in the
program's docstring (tests/_cfprogs.py), so a mistake in the interpreter's
jump handling (a continue that skips the loop's step, a do-while that tests
first, a return that leaves the loop, a break that leaves both loops, Ci
reset per layer, lobes dropped on reuse) fails one of them.
"""
import numpy as np
import pytest

import _cfprogs as C
import _oracle as O

N_PTS = 3000


@pytest.fixture(scope="module")
def pts():
    return C.inputs(N_PTS, 5)


@pytest.mark.parametrize("name", sorted(C.PROGRAMS))
def test_program_closed_form(pts, name):
    o, want = C.PROGRAMS[name]()
    g = O.Group([(o.name, "l0", o.text())], []).outputs("fout", "cout")
    out, _ = g.run(pts["P"], pts["N"])
    n, a, b = pts["nab"]
    np.testing.assert_array_equal(out[0], np.asarray(want(n, a, b), np.float32))


def test_trip_counts_cover_0_to_8(pts):
    n, a, b = pts["nab"]
    for v in (n, a, b):
        assert set(np.unique(v)) == set(range(9))


def _closure_eval(layers, conns, pts):
    g = O.Group(layers, conns).outputs("fout", "cout")
    return g.run(pts["P"], pts["N"], pts["wo"], pts["wi"])


@pytest.mark.parametrize("name", sorted(C.CLOSURE_PROGRAMS))
def test_closure_closed_form(pts, name):
    o, want = C.CLOSURE_PROGRAMS[name]()
    _, ev = _closure_eval([(o.name, "l0", o.text())], [], pts)
    n, a, b = pts["nab"]
    w = want(n, a, b)
    for c in range(3):
        np.testing.assert_allclose(ev[c], w, rtol=2e-6)
    np.testing.assert_allclose(ev[3], pts["cos_wi"] / np.pi, rtol=2e-5, atol=1e-7)


def test_ci_shared_between_layers(pts):
    (o0, o1), conns, want = C.closure_upstream_layers()
    layers = [(o0.name, "l0", o0.text()), (o1.name, "l1", o1.text())]
    out, ev = _closure_eval(layers, conns, pts)
    n, a, b = pts["nab"]
    np.testing.assert_array_equal(out[0], 2.0)  # fout = fin + 1, fin = 1
    for c in range(3):
        np.testing.assert_allclose(ev[c], want(n, a, b), rtol=2e-6)


def test_planted_mistakes_fail():
    """The closed forms catch the plausible mistakes they are meant to."""
    n = np.arange(9)
    _, want_for = C.for_continue_break()
    # a continue that also skipped the step would loop forever at i == a;
    # treating continue as break ends the sum at a
    a, b = np.full(9, 2), np.full(9, 9)
    as_break = np.array([sum(i + 1 for i in range(min(k, 2))) for k in n], float)
    assert not np.array_equal(as_break, want_for(n, a, b))
    _, want_dw = C.dowhile_continue()
    tested_first = np.array([sum(i for i in range(1, k + 1)) for k in n], float)
    assert not np.array_equal(tested_first, want_dw(n, np.zeros(9, int), b))
    _, want_fn = C.function_return()
    # a return that left the loop would stop after the first call
    assert want_fn(np.array([5]), None, np.array([9]))[0] == 2 + 1 + 2 + 1 + 2
