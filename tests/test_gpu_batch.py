"""Multi-group batches on the GPU: the 64 generated groups (seed 3) shaded
by the persistent batch kernel, checked against

  * the oracle's OSO interpreter on the same seeded points (outputs fout,
    cout and the closure evaluation): within 4 ulp or 1e-5 relative, the
    stated bar with no absolute floor; the fraction of bit-identical values
    is printed per group and asserted overall;
  * each group's own standalone kernel (oslmi_execute_batch) at the full
    workload counts: bit-identical, since both run the same emitted code;
  * itself across launch shapes, graph replays and the autotuner's choice:
    bit-identical.
"""
import ctypes

import numpy as np
import pytest

import _oracle as O
from openshadinglanguage_amd import groupgen as G

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NG = 64


def close(got, ref, ulps=4, rel=1e-5):
    got = np.asarray(got, np.float32)
    ref = np.asarray(ref, np.float32)
    gi = got.view(np.int32).astype(np.int64)
    ri = ref.view(np.int32).astype(np.int64)
    gi = np.where(gi < 0, -(gi & 0x7fffffff), gi)
    ri = np.where(ri < 0, -(ri & 0x7fffffff), ri)
    with np.errstate(invalid="ignore"):  # inf - inf where both pdfs are a delta lobe's
        d = np.abs(got.astype(np.float64) - ref)
    return (np.abs(gi - ri) <= ulps) | (d <= rel * np.abs(ref.astype(np.float64))) \
        | (got == ref) | (np.isnan(got) & np.isnan(ref))


@pytest.fixture(scope="module")
def setup():
    from openshadinglanguage_amd import check, lib
    from openshadinglanguage_amd import shadingsys as S
    check(lib().oslmi_init())
    specs, counts = G.make_groups(NG, 3)
    ss = S.ShadingSystem()
    groups = [S.build_generated_group(ss, specs[k], "gen%d" % k) for k in range(NG)]
    batch = S.GroupBatch(ss, groups)
    return {"ss": ss, "specs": specs, "counts": counts, "groups": groups, "batch": batch}


def _items(counts, seed0, dev):
    """Per-group device inputs/outputs; returns (items, host inputs, outs)."""
    items, host, outs = [], [], []
    for k, n in enumerate(counts):
        n = int(n)
        w = G.group_inputs(n, seed0 + k)
        t = {x: torch.from_numpy(w[x]).to(dev) for x in ("P", "N", "wo", "wi")}
        out = torch.full((4, n), -7.0, device=dev)
        ev = torch.full((4, n), -7.0, device=dev)
        glob = {"P": t["P"], "N": t["N"], "dPdx": (0.0, 0.0, 0.0), "dPdy": (0.0, 0.0, 0.0)}
        items.append((n, glob, out, t["wo"], t["wi"], ev))
        host.append(w)
        outs.append((out, ev))
    return items, host, outs


def _cpu(outs):
    return [(o.cpu().numpy(), e.cpu().numpy()) for o, e in outs]


def test_batch_matches_oracle_interpreter(setup):
    dev = torch.device("cuda:0")
    counts = [min(int(c), 700) for c in setup["counts"]]
    counts[5] = 1  # single-point group
    items, host, outs = _items(counts, 100, dev)
    b = setup["batch"]
    b.bind(items)
    b.launch()
    torch.cuda.synchronize()
    got = _cpu(outs)
    bad = []
    same = total = 0
    for k in range(NG):
        layers, conns = setup["specs"][k][0], setup["specs"][k][1]
        ref_out, ref_ev = O.Group(layers, conns).run(host[k]["P"], host[k]["N"], host[k]["wo"],
                                                    host[k]["wi"], nthreads=8)
        ok_o = close(got[k][0], ref_out)
        ok_e = close(got[k][1], ref_ev)
        eq = np.concatenate([(got[k][0] == ref_out) | (np.isnan(got[k][0]) & np.isnan(ref_out)),
                             (got[k][1] == ref_ev) | (np.isnan(got[k][1]) & np.isnan(ref_ev))])
        same += int(eq.sum())
        total += eq.size
        print("group %2d: %5d points, exact %.5f" % (k, counts[k], eq.mean()), flush=True)
        if not (ok_o.all() and ok_e.all()):
            i = np.argwhere(~ok_o)[:3].tolist() + np.argwhere(~ok_e)[:3].tolist()
            bad.append((k, int((~ok_o).sum()), int((~ok_e).sum()), i))
    print("exact-match fraction over all groups: %.6f (%d values)" % (same / total, total))
    assert not bad, "groups with mismatches (group, outputs, evals, first): %s" % bad
    assert same / total > 0.999


def test_batch_matches_standalone_kernels_full_counts(setup):
    dev = torch.device("cuda:0")
    counts = [int(c) for c in setup["counts"]]
    items, host, outs = _items(counts, 1000, dev)
    b = setup["batch"]
    b.bind(items)
    b.launch()
    torch.cuda.synchronize()
    got = _cpu(outs)
    ss = setup["ss"]
    for k, (n, glob, _, wo, wi, _) in enumerate(items):
        out = torch.empty((4, n), device=dev)
        ev = torch.empty((4, n), device=dev)
        ss.execute(setup["groups"][k], n, glob, out, wo, wi, ev)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got[k][0], out.cpu().numpy(), err_msg="group %d outputs" % k)
        np.testing.assert_array_equal(got[k][1], ev.cpu().numpy(), err_msg="group %d eval" % k)


@pytest.mark.parametrize("cfg", [dict(block=64, ppt=1, blocks_per_cu=1, bb_lds=0, order=0,
                                      streams=1),
                                 dict(block=128, ppt=4, blocks_per_cu=2, bb_lds=1, order=0,
                                      streams=4),
                                 dict(block=256, ppt=2, blocks_per_cu=32, bb_lds=1, order=1,
                                      streams=2)])
def test_launch_shapes_agree(setup, cfg):
    dev = torch.device("cuda:0")
    counts = [min(int(c), 5000) for c in setup["counts"]]
    items, _, outs = _items(counts, 2000, dev)
    b = setup["batch"]
    b.bind(items)
    b.set_config(block=256, ppt=1, blocks_per_cu=8, bb_lds=1, order=0, streams=1)
    b.launch()
    torch.cuda.synchronize()
    ref = _cpu(outs)
    for o, e in outs:
        o.fill_(-7.0)
        e.fill_(-7.0)
    b.set_config(**cfg)
    b.launch()
    torch.cuda.synchronize()
    for k, (r, g) in enumerate(zip(ref, _cpu(outs))):
        np.testing.assert_array_equal(r[0], g[0], err_msg="group %d" % k)
        np.testing.assert_array_equal(r[1], g[1], err_msg="group %d" % k)
    # one launch per group gives the same values too
    for o, e in outs:
        o.fill_(-7.0)
        e.fill_(-7.0)
    b.launch_per_group()
    torch.cuda.synchronize()
    for k, (r, g) in enumerate(zip(ref, _cpu(outs))):
        np.testing.assert_array_equal(r[0], g[0], err_msg="group %d" % k)


def test_graph_replays_match_eager(setup):
    dev = torch.device("cuda:0")
    counts = [min(int(c), 20000) for c in setup["counts"]]
    items, _, outs = _items(counts, 3000, dev)
    b = setup["batch"]
    b.bind(items)
    b.launch()
    torch.cuda.synchronize()
    ref = _cpu(outs)
    s = torch.cuda.Stream()
    b.capture(s)
    for rep in range(3):
        for o, e in outs:
            o.fill_(-7.0)
            e.fill_(-7.0)
        torch.cuda.synchronize()
        b.launch(stream=s)
        s.synchronize()
        got = _cpu(outs)
        for k in range(NG):
            np.testing.assert_array_equal(ref[k][0], got[k][0], err_msg="replay %d group %d" % (rep, k))
            np.testing.assert_array_equal(ref[k][1], got[k][1], err_msg="replay %d group %d" % (rep, k))


def test_autotune_keeps_results(setup):
    dev = torch.device("cuda:0")
    counts = [min(int(c), 3000) for c in setup["counts"]]
    items, _, outs = _items(counts, 4000, dev)
    b = setup["batch"]
    b.bind(items)
    b.set_config(block=256, ppt=1, blocks_per_cu=8, bb_lds=1, order=0, streams=1)
    b.launch()
    torch.cuda.synchronize()
    ref = _cpu(outs)
    best, ms = b.autotune(reps=1)
    assert best["block"] in (64, 128, 256) and 1 <= best["ppt"] <= 8 and ms > 0
    assert b.config().as_dict() == best
    for o, e in outs:
        o.fill_(-7.0)
        e.fill_(-7.0)
    b.launch()
    torch.cuda.synchronize()
    for k, (r, g) in enumerate(zip(ref, _cpu(outs))):
        np.testing.assert_array_equal(r[0], g[0], err_msg="group %d" % k)
        np.testing.assert_array_equal(r[1], g[1], err_msg="group %d" % k)
    # tuned as captured graphs: the winner is replayed on the tuning stream
    s = torch.cuda.Stream()
    best, ms = b.autotune(stream=s, reps=1, graph=True)
    assert ms > 0 and b.config().as_dict() == best
    for rep in range(2):
        for o, e in outs:
            o.fill_(-7.0)
            e.fill_(-7.0)
        torch.cuda.synchronize()
        b.launch(stream=s)
        s.synchronize()
        for k, (r, g) in enumerate(zip(ref, _cpu(outs))):
            np.testing.assert_array_equal(r[0], g[0], err_msg="graph group %d" % k)
            np.testing.assert_array_equal(r[1], g[1], err_msg="graph group %d" % k)


def test_empty_items_are_skipped(setup):
    dev = torch.device("cuda:0")
    counts = [0 if k % 3 else 17 for k in range(NG)]
    items, _, outs = _items(counts, 5000, dev)
    b = setup["batch"]
    b.bind(items)
    b.launch()
    b.launch()
    torch.cuda.synchronize()
    for k in range(0, NG, 3):
        assert (outs[k][0] != -7.0).all()


def test_second_group_set_matches_interpreter():
    """16 more generated groups (seed 11), each through its own emitted
    kernel (oslmi_execute_batch), against the oracle's interpreter."""
    from openshadinglanguage_amd import check, lib
    from openshadinglanguage_amd import shadingsys as S
    check(lib().oslmi_init())
    dev = torch.device("cuda:0")
    specs, _ = G.make_groups(16, 11)
    ss = S.ShadingSystem()
    bad = []
    for k, spec in enumerate(specs):
        g = S.build_generated_group(ss, spec, "set11_%d" % k)
        n = 300 + 37 * k
        w = G.group_inputs(n, 500 + k)
        t = {x: torch.from_numpy(w[x]).to(dev) for x in ("P", "N", "wo", "wi")}
        out = torch.full((4, n), -7.0, device=dev)
        ev = torch.full((4, n), -7.0, device=dev)
        glob = {"P": t["P"], "N": t["N"], "dPdx": (0.0, 0.0, 0.0), "dPdy": (0.0, 0.0, 0.0)}
        ss.execute(g, n, glob, out, t["wo"], t["wi"], ev)
        torch.cuda.synchronize()
        ref_out, ref_ev = O.Group(spec[0], spec[1]).run(w["P"], w["N"], w["wo"], w["wi"],
                                                        nthreads=8)
        ok_o = close(out.cpu().numpy(), ref_out)
        ok_e = close(ev.cpu().numpy(), ref_ev)
        if not (ok_o.all() and ok_e.all()):
            bad.append((k, int((~ok_o).sum()), int((~ok_e).sum())))
    assert not bad, "groups with mismatches (group, outputs, evals): %s" % bad
