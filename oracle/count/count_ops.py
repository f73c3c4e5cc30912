#!/usr/bin/env python3
"""Algorithmic 32-bit lane-operations per shading point of the benchmark
groups, counted by the op-counting build of the oracle
(oracle/build/liboslref_count.so, see counted.h), on seeded samples of the
benchmark workloads. Also checks that the counting build computes the same
values as the oracle. Writes openshadinglanguage_amd/roofline_ops.json.

    make -C oracle count && python oracle/count/count_ops.py
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import _oracle as O  # noqa: E402
from openshadinglanguage_amd import workloads as W  # noqa: E402


def main():
    L = ctypes.CDLL(os.path.join(ROOT, "oracle", "build", "liboslref_count.so"))
    P = ctypes.c_void_p
    L.ocount_get.restype = ctypes.c_ulonglong
    L.oref_shade_c1.argtypes = [P, P, ctypes.c_int64, ctypes.c_int]
    L.oref_shade_c2.argtypes = [P, P, P, P, P, P, P, P, ctypes.c_int64, ctypes.c_int]
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    res = {}

    w = W.fbm_cell_inputs(256, 1)
    out = np.zeros_like(w["P"])
    L.ocount_reset()
    L.oref_shade_c1(p(w["P"]), p(out), w["P"].shape[1], 1)
    n = w["P"].shape[1]
    res["fbm_cell"] = {"ops_per_point": L.ocount_get() / n, "points_counted": n}
    assert np.array_equal(out, O.shade_c1(w["P"])), "counting build differs from oracle"

    full = W.layered3_inputs(1024, 2)
    rng = np.random.default_rng(0)
    idx = np.sort(rng.choice(full["P"].shape[1], 32768, replace=False))
    sub = {k: np.ascontiguousarray(v[:, idx]) if v.ndim == 2 else v for k, v in full.items()}
    n = idx.size
    out = np.zeros((4, n), np.float32)
    L.ocount_reset()
    L.oref_shade_c2(p(sub["P"]), p(sub["N"]), p(sub["wo"]), p(sub["wi"]), p(sub["dPdx"]),
                    p(sub["dPdy"]), p(out), None, n, 1)
    res["layered3"] = {"ops_per_point": L.ocount_get() / n, "points_counted": int(n),
                       "sample": "32768 random points of the 1024x1024 workload (seed 0)"}
    ref = O.shade_c2(sub["P"], sub["N"], sub["wo"], sub["wi"], sub["dPdx"], sub["dPdy"])
    assert np.array_equal(out, ref), "counting build differs from oracle"
    # layered closure: 1 eval + 4 samples per point, and eval-only queries
    L.oref_mtx_layered.argtypes = [P] * 8 + [ctypes.c_int64, ctypes.c_int]
    L.oref_mtx_eval_local.argtypes = [P, P, P, ctypes.c_int64, ctypes.c_int]
    mat = np.array(W.layered_material().as_list(), np.float32)
    w = W.layered_closure_inputs(32768, seed=4)
    n = w["N"].shape[1]
    ev = np.zeros((4, n), np.float32)
    smp = np.zeros((28, n), np.float32)
    L.ocount_reset()
    L.oref_mtx_layered(p(mat), p(w["N"]), p(w["V"]), p(w["L"]), p(w["params"]), p(w["rnd"]),
                       p(ev), p(smp), n, 1)
    res["closure"] = {"ops_per_point": L.ocount_get() / n, "points_counted": n,
                      "sample": "first 32768 points of the closure workload (seed 4); "
                                "1 eval + 4 samples per point"}
    rev, rsmp = O.mtx_layered(mat, w["N"], w["V"], w["L"], w["params"], w["rnd"])
    assert np.array_equal(ev, rev) and np.array_equal(smp, rsmp), "counting build differs"
    q = W.surrogate_queries(32768, seed=6)
    out = np.zeros((4, q.shape[1]), np.float32)
    L.ocount_reset()
    L.oref_mtx_eval_local(p(mat), p(q), p(out), q.shape[1], 1)
    res["closure_eval"] = {"ops_per_point": L.ocount_get() / q.shape[1],
                           "points_counted": q.shape[1],
                           "sample": "32768 shading-frame queries (seed 6); 1 eval per query"}
    assert np.array_equal(out, O.mtx_eval_local(mat, q)), "counting build differs"
    # the 64 generated groups (seed 3), fout/cout outputs + 1 closure eval
    from openshadinglanguage_amd import groupgen as G
    specs, counts = G.make_groups(64, 3)
    L.oref_group_new.restype = P
    L.oref_group_layer.argtypes = [P, ctypes.c_char_p, ctypes.c_char_p]
    L.oref_group_connect.argtypes = [P] + [ctypes.c_char_p] * 4
    L.oref_group_run.argtypes = [P, P, P, P, P, ctypes.c_int64, P, P, ctypes.c_int]
    L.oref_group_free.argtypes = [P]
    per_group = []
    for k, spec in enumerate(specs):
        h = L.oref_group_new()
        for _, ln, text in spec[0]:
            assert L.oref_group_layer(h, text.encode(), ln.encode()) == 0
        for c in spec[1]:
            assert L.oref_group_connect(h, *[x.encode() for x in c]) == 0
        w = G.group_inputs(512, 100 + k)
        out = np.zeros((4, 512), np.float32)
        ev = np.zeros((4, 512), np.float32)
        L.ocount_reset()
        L.oref_group_run(h, p(w["P"]), p(w["N"]), p(w["wo"]), p(w["wi"]), 512, p(out), p(ev), 1)
        per_group.append(L.ocount_get() / 512)
        L.oref_group_free(h)
        ref_out, ref_ev = O.Group(spec[0], spec[1]).run(w["P"], w["N"], w["wo"], w["wi"])
        assert np.array_equal(out, ref_out) and np.array_equal(ev, ref_ev), "counting build differs"
    total = float(np.dot(per_group, counts.astype(np.float64)))
    res["batch64"] = {"ops_per_point": total / float(counts.sum()), "ops_total": total,
                      "points": int(counts.sum()), "ops_per_point_by_group": per_group,
                      "sample": "512 points per group (group_inputs seed 100 + group); "
                                "fout, cout and 1 closure eval per point"}
    res["_doc"] =("algorithmic 32-bit lane-ops per point from the op-counting oracle "
                   "(oracle/count/count_ops.py)")
    dst = os.path.join(ROOT, "openshadinglanguage_amd", "roofline_ops.json")
    json.dump(res, open(dst, "w"), indent=1)
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
