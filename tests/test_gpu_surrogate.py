"""GPU tests of the closure surrogate's BF16-MFMA kernels.

* gradients of one batch against the numpy oracle (oracle/mlp.py), which
  rounds to bf16 at the same points and sums in float64: the kernels differ
  only by float32 accumulation order;
* the fused inference kernel against the oracle's forward pass;
* a short training run lowers the loss; routing every cell to the exact
  closure reproduces the exact kernel bit for bit.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

import _oracle as O
from openshadinglanguage_amd import surrogate as S
from openshadinglanguage_amd import workloads as W

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import mlp as M  # noqa: E402


@pytest.fixture(scope="module")
def dev():
    from openshadinglanguage_amd import check, lib
    check(lib().oslmi_init())
    return torch.device("cuda:0")


def _norm_inputs(q):
    x = (q.T.astype(np.float32) - np.array(S.IN_OFFSET, np.float32)) * np.array(S.IN_SCALE,
                                                                                   np.float32)
    return x.astype(np.float64)


def _norm_targets(t, mean, std):
    y = np.empty((t.shape[1], 4))
    y[:, :3] = np.log(np.maximum(t[:3].T.astype(np.float32), np.float32(1e-6)))
    y[:, 3] = np.log(np.maximum(t[3].astype(np.float32), np.float32(1e-12)))
    return ((y.astype(np.float32) - np.array(mean, np.float32)) / np.array(std, np.float32)
            ).astype(np.float64)


def _random_params(seed):
    rng = np.random.default_rng(seed)
    p = M.init_params(seed)
    for k in p:
        if k.startswith("b"):
            p[k] = rng.normal(0, 0.2, p[k].shape)
    p["b5"][4:] = 0
    return p


def test_gradients_match_oracle(dev):
    print("surrogate grads vs oracle", flush=True)
    B = 1024
    q = S.upper_hemisphere_queries(B, seed=21)
    t = O.mtx_eval_local(W.layered_material().as_list(), q)
    s = S.Surrogate(max_batch=2048)
    mean, std = [-0.5, -0.6, -0.7, -1.0], [1.3, 1.2, 1.1, 1.5]
    s.set_output_normalization(mean, std)
    p = _random_params(4)
    s.set_params(M.pack(p).astype(np.float32))
    pf = M.unpack(M.pack(p).astype(np.float32).astype(np.float64))
    g_gpu, loss_gpu = s.grads(torch.from_numpy(q).to(dev), torch.from_numpy(t).to(dev))
    loss, g = M.loss_and_grads(pf, _norm_inputs(q), _norm_targets(t, mean, std), bf=True)
    ref = M.pack(g)
    assert abs(loss_gpu - loss) <= 1e-4 * abs(loss)
    for name, shape in M.SHAPES:
        a = M.unpack(g_gpu.astype(np.float64))[name]
        b = M.unpack(ref)[name]
        scale = np.abs(b).max() + 1e-12
        assert np.abs(a - b).max() <= 2e-3 * scale, (name, np.abs(a - b).max(), scale)


def test_inference_matches_oracle_forward(dev):
    print("surrogate inference vs oracle forward", flush=True)
    n = 4096 + 5  # ragged: not a multiple of the 16-query tile
    q = S.upper_hemisphere_queries(n, seed=22)
    s = S.Surrogate(max_batch=2048)
    mean, std = [-0.5, -0.6, -0.7, -1.0], [1.3, 1.2, 1.1, 1.5]
    s.set_output_normalization(mean, std)
    p = _random_params(5)
    s.set_params(M.pack(p).astype(np.float32))
    pf = M.unpack(M.pack(p).astype(np.float32).astype(np.float64))
    tq = torch.from_numpy(q).to(dev)
    out = torch.full((4, n), -1.0, device=dev)
    s.infer(tq, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    pred, _ = M.forward(pf, _norm_inputs(q), bf=True)
    ref = np.exp(pred * np.array(std) + np.array(mean)).T
    rel = np.abs(got - ref) / np.abs(ref)
    assert np.median(rel) < 1e-3 and rel.max() < 3e-2, (np.median(rel), rel.max())


def test_out_of_domain_and_gate(dev):
    s = S.Surrogate(max_batch=2048)
    q = S.upper_hemisphere_queries(64, seed=23)
    q[5, :8] = 2.5       # coat ior outside [1.3, 1.8]
    q[2, 8:16] = -0.5    # L below the surface: exact zero
    out = torch.full((4, 64), -1.0, device=dev)
    s.infer(torch.from_numpy(q).to(dev), out)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    assert np.isnan(o[:, :8]).all()
    assert (o[:, 8:16] == 0).all()
    assert np.isfinite(o[:, 16:]).all() and (o[:, 16:] > 0).all()


def test_training_lowers_loss(dev):
    print("surrogate training run", flush=True)
    n = 1 << 18
    q = torch.from_numpy(S.upper_hemisphere_queries(n, seed=24)).to(dev)
    t = torch.empty((4, n), device=dev)
    from openshadinglanguage_amd import check, lib
    m = W.layered_material()
    check(lib().oslmi_layered_eval_local(ctypes.byref(m), n, S._p(q), S._p(t), None))
    s = S.Surrogate(max_batch=8192)
    s.fit_output_normalization(t)
    qb, tb = q[:, :8192].contiguous(), t[:, :8192].contiguous()
    _, loss0 = s.grads(qb, tb)
    s.train(q, t, steps=400, batch=8192, lr=3e-3)
    _, loss1 = s.grads(qb, tb)
    assert loss1 < 0.25 * loss0, (loss0, loss1)
    # deterministic: the same run gives the same parameters
    p1 = s.get_params()
    s2 = S.Surrogate(max_batch=8192)
    s2.fit_output_normalization(t)
    s2.train(q, t, steps=400, batch=8192, lr=3e-3)
    assert np.array_equal(p1, s2.get_params())


def test_full_routing_reproduces_exact(dev):
    from openshadinglanguage_amd import check, lib
    n = 50000
    q = torch.from_numpy(W.surrogate_queries(n, seed=25)).to(dev)
    m = W.layered_material()
    exact = torch.empty((4, n), device=dev)
    check(lib().oslmi_layered_eval_local(ctypes.byref(m), n, S._p(q), S._p(exact), None))
    s = S.Surrogate(max_batch=2048)
    s.set_routing(np.ones(S.NCELL, np.uint8))
    out = torch.full((4, n), -1.0, device=dev)
    idx = torch.empty(n, dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    s.infer(q, out, material=m, fb_index=idx, fb_count=cnt)
    torch.cuda.synchronize()
    assert torch.equal(out, exact)
    upper = (q[4] > 0).logical_and(q[2] > 0)
    assert int(cnt.item()) == int(upper.sum())
    # the list holds every routed query once (in no particular order)
    listed = np.sort(idx[:int(cnt.item())].cpu().numpy())
    assert np.array_equal(listed, np.flatnonzero(upper.cpu().numpy()))
    s.set_routing(np.zeros(S.NCELL, np.uint8))
    s.infer(q, out, material=m, fb_index=idx, fb_count=cnt)
    torch.cuda.synchronize()
    assert int(cnt.item()) == 0


def test_spot_check_sends_low_predictions_to_exact(dev):
    """With the spot-check threshold above every prediction, all network
    queries join the fallback list and the answer equals the exact closure;
    with a threshold between, exactly the queries predicted below it do."""
    from openshadinglanguage_amd import check, lib
    n = 50000
    q = torch.from_numpy(W.surrogate_queries(n, seed=26)).to(dev)
    m = W.layered_material()
    exact = torch.empty((4, n), device=dev)
    check(lib().oslmi_layered_eval_local(ctypes.byref(m), n, S._p(q), S._p(exact), None))
    s = S.Surrogate(max_batch=2048)
    s.set_routing(np.zeros(S.NCELL, np.uint8))
    idx = torch.empty(n, dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    net = torch.full((4, n), -1.0, device=dev)
    s.infer(q, net)  # network answers for every upper-hemisphere query
    s.set_spot_check(1e30)
    out = torch.full((4, n), -1.0, device=dev)
    s.infer(q, out, material=m, fb_index=idx, fb_count=cnt)
    torch.cuda.synchronize()
    assert torch.equal(out, exact)
    upper = (q[4] > 0).logical_and(q[2] > 0)
    assert int(cnt.item()) == int(upper.sum())
    thr = float(net[:3, upper].min(0).values.median())
    s.set_spot_check(thr)
    s.infer(q, out, material=m, fb_index=idx, fb_count=cnt)
    torch.cuda.synchronize()
    low = upper.logical_and(net[:3].min(0).values < thr)
    assert int(cnt.item()) == int(low.sum())
    assert torch.equal(out[:, low], exact[:, low])
    assert torch.equal(out[:, upper & ~low], net[:, upper & ~low])
    with pytest.raises(Exception):
        s.set_spot_check(-1.0)
