"""The allocation-free batch entry point (oslmi_batch_launch_noalloc),
called through the C ABI, against the allocating path it stands in for
(oslmi_batch_bind + oslmi_batch_launch) on the 64 generated groups: the
outputs must be bit-identical (the same kernels over the same tiles), on
two calls in a row, the second with fewer points per group, and requests
beyond the fixed workspace must fail cleanly."""
import ctypes

import numpy as np
import pytest

from openshadinglanguage_amd import groupgen as G

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NG = 64


@pytest.fixture(scope="module")
def setup():
    from openshadinglanguage_amd import check, lib
    from openshadinglanguage_amd import shadingsys as S
    check(lib().oslmi_init())
    specs, counts = G.make_groups(NG, 3)
    ss = S.ShadingSystem()
    groups = [S.build_generated_group(ss, specs[k], "gen%d" % k) for k in range(NG)]
    batch = S.GroupBatch(ss, groups)
    return {"counts": counts, "batch": batch}


def _items(counts, seed0, dev):
    items = []
    for k, n in enumerate(counts):
        w = G.group_inputs(int(n), seed0 + k)
        t = {x: torch.from_numpy(w[x]).to(dev) for x in ("P", "N", "wo", "wi")}
        glob = {"P": t["P"], "N": t["N"], "dPdx": (0.0, 0.0, 0.0), "dPdy": (0.0, 0.0, 0.0)}
        items.append((int(n), glob, t["wo"], t["wi"]))
    return items


def _outs(items, dev):
    return [(torch.full((4, n), -7.0, device=dev), torch.full((4, n), -7.0, device=dev))
            for n, _, _, _ in items]


def _c_items(items, outs):
    from openshadinglanguage_amd import shadingsys as S
    arr = (S.BatchItem * len(items))()
    for k, ((n, glob, wo, wi), (o, e)) in enumerate(zip(items, outs)):
        arr[k].globals = S.batched_globals(n, glob)
        arr[k].d_outputs = o.data_ptr()
        arr[k].d_wo = wo.data_ptr()
        arr[k].d_wi = wi.data_ptr()
        arr[k].d_eval = e.data_ptr()
    return arr


def _noalloc(batch, arr, n):
    from openshadinglanguage_amd import lib
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return lib().oslmi_batch_launch_noalloc(ctypes.c_void_p(batch.h), n, arr, st)


def _check_same(batch, items, dev, tag):
    from openshadinglanguage_amd import check
    ref = _outs(items, dev)
    batch.bind([(n, glob, o, wo, wi, e) for (n, glob, wo, wi), (o, e) in zip(items, ref)])
    batch.launch()
    got = _outs(items, dev)
    arr = _c_items(items, got)
    print("no-alloc %s: %d groups, %d points" % (tag, len(items), sum(i[0] for i in items)),
          flush=True)
    check(_noalloc(batch, arr, len(items)))
    torch.cuda.synchronize()
    for k, ((ro, re), (go, ge)) in enumerate(zip(ref, got)):
        np.testing.assert_array_equal(ro.cpu().numpy(), go.cpu().numpy(), err_msg="%s group %d" % (tag, k))
        np.testing.assert_array_equal(re.cpu().numpy(), ge.cpu().numpy(), err_msg="%s group %d" % (tag, k))


def test_noalloc_matches_allocating_path_twice(setup):
    dev = torch.device("cuda:0")
    b = setup["batch"]
    first = _items([min(int(c), 1500) for c in setup["counts"]], 4000, dev)
    _check_same(b, first, dev, "call 1")
    second = _items([min(int(c), 300) for c in setup["counts"]], 5000, dev)
    _check_same(b, second, dev, "call 2 (smaller)")


def test_noalloc_rejects_requests_beyond_the_workspace(setup):
    from openshadinglanguage_amd import lib
    dev = torch.device("cuda:0")
    b = setup["batch"]
    items = _items([1] * NG, 6000, dev)
    outs = _outs(items, dev)
    arr = _c_items(items, outs)
    assert _noalloc(b, arr, NG - 1) != 0  # not one item per group
    assert b"nitems" in lib().oslmi_last_error()
    # 32769 tiles of 64 points in one group exceed the tile capacity
    b.set_config(block=64, ppt=1, blocks_per_cu=8, bb_lds=1, order=0, streams=1)
    big = 64 * 32769
    items[0] = (big,) + _items([big], 6100, dev)[0][1:]
    outs[0] = (torch.empty((4, big), device=dev), torch.empty((4, big), device=dev))
    arr = _c_items(items, outs)
    assert _noalloc(b, arr, NG) != 0
    assert b"tiles exceed" in lib().oslmi_last_error()
    b.set_config(block=256, ppt=1, blocks_per_cu=8, bb_lds=1, order=0, streams=1)
