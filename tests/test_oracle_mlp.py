"""The numpy oracle of the surrogate network (oracle/mlp.py): its analytic
gradients against central finite differences (float64, no bf16 rounding),
and its bf16 rounding against hand-computed cases."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import mlp as M  # noqa: E402


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8,
                  1.0 + 2.0 ** -7 + 2.0 ** -8 + 2.0 ** -10, -2.5, 3.0e38, 1e-40], np.float32)
    got = M.bf16(x)
    # ulp(1) = 2^-7. Ties to even: 1 + 2^-8 -> 1 and 1 + 3 * 2^-8 -> 1 + 2^-6;
    # above the half-way point -> up
    assert got[0] == 1.0 and got[1] == 1.0
    assert got[2] == 1.0 + 2.0 ** -6
    assert got[3] == 1.0 + 2.0 ** -6
    assert got[4] == -2.5
    assert np.isfinite(got[5]) and abs(got[5] / 3.0e38 - 1) < 2 ** -8
    assert abs(got[6] - 1e-40) < 1e-41


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    p = M.init_params(3)
    for k in p:
        if k.startswith("b"):
            p[k] = rng.normal(0, 0.1, p[k].shape)
    p["b5"][M.NOUT:] = 0
    x = rng.uniform(-1, 1, (16, M.NIN))
    y = rng.normal(0, 1, (16, M.NOUT))
    loss, g = M.loss_and_grads(p, x, y)
    eps = 1e-6
    for name in ("W1", "b1", "W3", "b4", "W5", "b5"):
        idx = [tuple(rng.integers(0, s) for s in p[name].shape) for _ in range(6)]
        if name == "W1":
            idx = [(i[0], i[1] % M.NIN) for i in idx]
        if name in ("W5", "b5"):
            idx = [(i[0] % M.NOUT,) + tuple(i[1:]) for i in idx]
        for i in idx:
            q = {k: v.copy() for k, v in p.items()}
            q[name][i] += eps
            lp, _ = M.loss_and_grads(q, x, y)
            q[name][i] -= 2 * eps
            lm, _ = M.loss_and_grads(q, x, y)
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[name][i]) <= 1e-6 * max(1.0, abs(fd)), (name, i, fd, g[name][i])


def test_pack_roundtrip_and_count():
    p = M.init_params(0)
    flat = M.pack(p)
    assert flat.size == 16672 == M.param_count()
    q = M.unpack(flat)
    assert all(np.array_equal(p[k], q[k]) for k in p)
