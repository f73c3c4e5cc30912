#!/usr/bin/env python3
"""Benchmark: Mpts/s shaded on one MI355X vs the batched CPU path.

Default workload (N=1): the 3-layer "layered3" group over 1024x1024 points,
derivatives on, closure evaluated at one (wo, wi) pair per point (one step =
one batched execution of the group over all points). --workload fbm_cell
runs the single-layer 256x256 group instead.

Prints one JSON line (see the README of the benchmark contract): metric,
value (Mpts/s, whole job), ms_per_step, roofline (VALU-bound kernel, live
HIP-event timing) and cpu_baseline (the CPU oracle timed on the host cores
on the same workload).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X (gfx950) 32-bit VALU peak: 256 CUs x 128 lanes/clk x 2.4 GHz =
# 78.6 T lane-ops/s (the published 157.3 TFLOP/s FP32 vector figure counts an
# FMA as two flops). Measured on MI355X: the emitted layered3 kernel issues
# 1.35 VALU wave-instructions per CU-clock (SQ_INSTS_VALU / GRBM_GUI_ACTIVE),
# which only a 2-per-clock issue rate allows.
VALU_PEAK_OPS = 256 * 128 * 2.4e9
HBM_PEAK_BPS = 8.0e12

# Algorithmic 32-bit lane-ops per point (oracle/count/count_ops.py ->
# openshadinglanguage_amd/roofline_ops.json) and algorithmic HBM bytes per
# point (SoA planes read + written).
WORKLOADS = {
    "layered3": {"res": 1024, "bytes_per_point": (12 * 4) + 16},
    "fbm_cell": {"res": 256, "bytes_per_point": 12 + 12},
}

# HBM traffic per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (profiles/*_pmc.txt), in bytes; None when not measured.
# FETCH_SIZE reads half the bytes of a plain 256 MiB copy on this GPU (x2
# scale), WRITE_SIZE reads them exactly (x1); profiles/layered3_pmc.txt. The
# layered3 figure includes the kernel's register-spill (scratch) traffic.
MEASURED_TRAFFIC = {"layered3": 2 * 41121.9 * 1024 + 85283.5 * 1024, "fbm_cell": None,
                    # the layered closure, 2,097,152 points per launch
                    # (profiles/layered_closure_pmc.txt)
                    "closure": 2 * 107329.7 * 1024 + 262144.0 * 1024}


def _ops_per_point(name):
    path = os.path.join(ROOT, "openshadinglanguage_amd", "roofline_ops.json")
    return json.load(open(path))[name]["ops_per_point"]


def host_threads():
    """Threads for the CPU legs: the host CPUs this process may run on,
    within the box's CPU share (OMP_NUM_THREADS when it is set)."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    share = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(avail or 1, int(share))) if share and share.isdigit() else max(1, avail or 1)


def host_info():
    """nproc, the CPU model and the widest x86-64 vector level it has."""
    model, flags = "unknown", ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and model == "unknown":
                model = line.split(":", 1)[1].strip()
            if line.startswith("flags") and not flags:
                flags = line.split(":", 1)[1]
    except OSError:
        pass
    level = "v4" if " avx512f " in flags + " " and " avx512bw " in flags + " " else "v3"
    return {"nproc": os.cpu_count(), "model": model, "level": level}


def cpu_baseline(name, inputs, nthreads, passes):
    """The reference-algorithm CPU restatement built for throughput
    (oracle/cpubase: one translation unit, -O3 for the host's vector level,
    FMA contraction, 16-lane SoA batches with `omp simd` lane loops, one
    thread per host CPU of the box's share) on the same inputs."""
    info = host_info()
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "build", "liboslref_w16_%s.so" % info["level"]))
    P = ctypes.c_void_p
    n = inputs["P"].shape[1]
    a = {k: np.ascontiguousarray(inputs[k], np.float32) for k in inputs if k in
         ("P", "N", "wo", "wi", "dPdx", "dPdy")}
    if name == "layered3":
        out = np.zeros((4, n), np.float32)
        lib.oref_w16_shade_c2.argtypes = [P] * 7 + [ctypes.c_int64, ctypes.c_int]
        args = [a[k].ctypes.data for k in ("P", "N", "wo", "wi", "dPdx", "dPdy")] + [
            out.ctypes.data, n, nthreads]
        fn = lambda: lib.oref_w16_shade_c2(*args)  # noqa: E731
    else:
        out = np.zeros((3, n), np.float32)
        lib.oref_w16_shade_c1.argtypes = [P, P, ctypes.c_int64, ctypes.c_int]
        fn = lambda: lib.oref_w16_shade_c1(a["P"].ctypes.data, out.ctypes.data, n, nthreads)  # noqa
    fn()  # warm (thread creation, page faults)
    t0 = time.perf_counter()
    for _ in range(passes):
        fn()
    dt = time.perf_counter() - t0
    return {"value": round(n * passes / dt / 1e6, 3), "unit": "Mpts/s", "cores": nthreads,
            "kind": "port", "nproc": info["nproc"], "cpu": info["model"],
            "sample": "%d pass(es) over the full %d-point workload: the reference-algorithm "
                      "restatement in 16-lane SoA batches (omp simd lane loops), -O3 "
                      "-march=x86-64-%s, FMA contraction on, %d threads of %s host CPUs, "
                      "%.1f s wall" % (passes, n, info["level"], nthreads, info["nproc"], dt)}


def _timed(step, stream, steps, warmup):
    """(ms per launch from HIP events on `stream`, wall ms per step)."""
    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return ev0.elapsed_time(ev1) / steps, wall * 1e3 / steps


# closure workload: 2M points, per point 1 eval + 4 samples; bytes per point:
# N, V, L 36 + params 36 + rnd 32 in; eval 16 + samples 4 x 28 out
CLOSURE_POINTS = 1 << 21
CLOSURE_BYTES_PER_POINT = 36 + 36 + 32 + 16 + 112


def bench_closure(steps, warmup, cpu, cpu_passes):
    """The layered MaterialX closure over the 2M-point closure workload."""
    import torch
    from openshadinglanguage_amd import check, lib
    from openshadinglanguage_amd import workloads as W
    dev = torch.device("cuda:0")
    w = W.layered_closure_inputs(CLOSURE_POINTS, seed=4)
    n = w["N"].shape[1]
    t = {k: torch.from_numpy(w[k]).to(dev) for k in ("N", "V", "L", "params", "rnd")}
    ev = torch.empty((4, n), device=dev)
    smp = torch.empty((28, n), device=dev)
    m = W.layered_material()
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    p = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731

    def step():
        check(lib().oslmi_layered_closure(ctypes.byref(m), n, p(t["N"]), p(t["V"]), p(t["L"]),
                                          p(t["params"]), p(t["rnd"]), p(ev), p(smp), sp))
    kernel_ms, ms_per_step = _timed(step, stream, steps, warmup)
    ops = _ops_per_point("closure")
    achieved = ops * n / (kernel_ms / 1e3)
    res = {"workload": "layer(sheen, layer(dielectric coat, generalized-Schlick base)): "
                       "1 eval + 4 samples per point", "points": n,
           "Mpts_per_s": round(n / (ms_per_step / 1e3) / 1e6, 2),
           "closure_evals_per_s": round(5 * n / (ms_per_step / 1e3), 1),
           "ms_per_step": round(ms_per_step, 4),
           "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 3),
                        "peak": round(VALU_PEAK_OPS / 1e12, 3), "unit": "Tops/s",
                        "frac": round(achieved / VALU_PEAK_OPS, 4),
                        "traffic": MEASURED_TRAFFIC["closure"],
                        "ops_per_point": round(ops, 1), "kernel_ms": round(kernel_ms, 4),
                        "hbm_algorithmic_GBps": round(CLOSURE_BYTES_PER_POINT * n
                                                      / (kernel_ms / 1e3) / 1e9, 2)}}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as O
        nthreads = host_threads()
        mat = m.as_list()
        fn = lambda: O.mtx_layered(mat, w["N"], w["V"], w["L"], w["params"], w["rnd"],  # noqa
                                   nthreads=nthreads)
        fn()
        t0 = time.perf_counter()
        for _ in range(cpu_passes):
            fn()
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(n * cpu_passes / dt / 1e6, 3), "unit": "Mpts/s",
                               "cores": nthreads, "kind": "port",
                               "sample": "%d pass(es) over the full %d-point workload, %d "
                                         "threads, %.1f s wall" % (cpu_passes, n, nthreads, dt)}
        res["gpu_over_cpu"] = round(res["Mpts_per_s"] / res["cpu_baseline"]["value"], 1)
    return res


def bench_small(steps, warmup):
    """The single-layer fbm_cell group over 256x256 points: 65,536 points
    are a few microseconds of work, so the launch cost dominates a single
    launch; 100 launches captured in one hipGraph and replayed amortise it."""
    import torch
    from openshadinglanguage_amd import shadingsys as S
    from openshadinglanguage_amd import workloads as W
    dev = torch.device("cuda:0")
    inp = W.fbm_cell_inputs(256, seed=1)
    n = inp["P"].shape[1]
    P = torch.from_numpy(inp["P"]).to(dev)
    out = torch.empty_like(P)
    ss, grp = S.build_fbm_cell()
    ss.jit_group(grp)
    side = torch.cuda.Stream()
    step = ss.bind(grp, n, {"P": P}, outputs=out, stream=side)
    eager_ms, _ = _timed(step, side, steps, warmup)
    torch.cuda.synchronize()
    ref = out.clone()
    reps = 100
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        for _ in range(reps):
            step()
    torch.cuda.synchronize()
    out.fill_(-1.0)
    # replay() launches the graph on the current stream
    graph_ms, _ = _timed(g.replay, torch.cuda.current_stream(), max(3, steps // 10), 1)
    torch.cuda.synchronize()
    same = bool(torch.equal(out, ref))
    per_launch = graph_ms / reps
    ops = _ops_per_point("fbm_cell")
    achieved = ops * n / (per_launch / 1e3)
    return {"workload": "fbm_cell: 4-octave perlin fBm + 0.25 cellnoise(3P) -> smoothstep/mix, "
                        "256x256 points, no derivs",
            "points": n, "eager_ms_per_launch": round(eager_ms, 5),
            "graph_ms_per_launch": round(per_launch, 5), "launches_per_graph": reps,
            "Mpts_per_s": round(n / (per_launch / 1e3) / 1e6, 1),
            "graph_matches_eager": same,
            "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 3),
                         "peak": round(VALU_PEAK_OPS / 1e12, 3), "unit": "Tops/s",
                         "frac": round(achieved / VALU_PEAK_OPS, 4), "traffic": None,
                         "ops_per_point": round(ops, 1), "kernel_ms": round(per_launch, 5)}}


def bench_batch(steps, warmup, cpu, autotune_reps=2):
    """The 64 generated groups (seed 3, log-uniform point counts): one
    persistent batch launch (autotuned, replayed from a hipGraph) against
    the same kernels launched group after group, as the reference's
    one-execute-per-group loop runs them."""
    import torch
    from openshadinglanguage_amd import groupgen as G
    from openshadinglanguage_amd import shadingsys as S
    dev = torch.device("cuda:0")
    specs, counts = G.make_groups(64, 3)
    ss = S.ShadingSystem()
    groups = [S.build_generated_group(ss, specs[k], "gen%d" % k) for k in range(64)]
    b = S.GroupBatch(ss, groups)
    items, host = [], []
    for k, n in enumerate(counts):
        n = int(n)
        w = G.group_inputs(n, 1000 + k)
        t = {x: torch.from_numpy(w[x]).to(dev) for x in ("P", "N", "wo", "wi")}
        glob = {"P": t["P"], "N": t["N"], "dPdx": (0.0, 0.0, 0.0), "dPdy": (0.0, 0.0, 0.0)}
        items.append((n, glob, torch.empty((4, n), device=dev), t["wo"], t["wi"],
                      torch.empty((4, n), device=dev)))
        host.append(w)
    total = int(counts.sum())
    b.bind(items)
    stream = torch.cuda.current_stream()
    seq_ms, _ = _timed(lambda: b.launch_per_group(stream), stream, max(3, steps // 5), 1)
    b.set_config(block=256, ppt=1, blocks_per_cu=8, bb_lds=1, order=0, streams=1)
    untuned_ms, _ = _timed(lambda: b.launch(stream), stream, max(3, steps // 5), 1)
    # eager: 8 groups per kernel, kernels overlapped on side streams, tuned
    side = torch.cuda.Stream()
    t0 = time.perf_counter()
    best_eager, tuner_eager = b.autotune(side, reps=autotune_reps)
    tune_s = time.perf_counter() - t0
    eager_ms, eager_step = _timed(lambda: b.launch(side), side, steps, warmup)
    # hipGraph: the tuner times every candidate as a captured graph and keeps
    # the winner captured; a captured graph runs its kernels one after
    # another, so fewer, larger kernels win there (32 groups per kernel)
    graphs = {}
    for per, bb in ((8, b), (32, None)):
        if bb is None:
            bb = S.GroupBatch(ss, groups, groups_per_kernel=per)
            bb.bind(items)
        best_g, _ = bb.autotune(side, reps=autotune_reps, graph=True)
        g_ms, g_step = _timed(lambda: bb.launch(side), side, steps, warmup)
        graphs[per] = {"graph_ms": round(g_ms, 4), "wall_ms_per_step": round(g_step, 4),
                       "kernels": bb.kernels, "best": best_g}
    gbest = min(graphs, key=lambda k: graphs[k]["graph_ms"])
    kernel_ms = min(eager_ms, graphs[gbest]["graph_ms"])
    ms_per_step = eager_step if eager_ms <= graphs[gbest]["graph_ms"] else \
        graphs[gbest]["wall_ms_per_step"]
    info = json.load(open(os.path.join(ROOT, "openshadinglanguage_amd",
                                       "roofline_ops.json")))["batch64"]
    achieved = info["ops_total"] / (kernel_ms / 1e3)
    res = {"workload": "64 generated groups (seed 3): 2-12 layers of perlin/simplex/cell/gabor/"
                       "hash noise, matrix transforms, color ops, data-dependent if/loops, "
                       "closure sums; point counts log-uniform in [32, 262144]; outputs "
                       "fout, cout + 1 closure eval per point",
           "groups": 64, "points": total,
           "Mpts_per_s": round(total / (kernel_ms / 1e3) / 1e6, 2),
           "best_mode": "eager" if eager_ms <= graphs[gbest]["graph_ms"] else
                        "graph (%d groups per kernel)" % gbest,
           "ms_per_step": round(ms_per_step, 4),
           "eager": {"kernels": b.kernels, "ms": round(eager_ms, 4), "best": best_eager,
                     "tuner_ms": round(tuner_eager, 4), "tune_seconds": round(tune_s, 2)},
           "graph": graphs,
           "untuned_ms": round(untuned_ms, 4), "per_group_launches_ms": round(seq_ms, 4),
           "speedup_vs_per_group": round(seq_ms / kernel_ms, 2),
           "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 3),
                        "peak": round(VALU_PEAK_OPS / 1e12, 3), "unit": "Tops/s",
                        "frac": round(achieved / VALU_PEAK_OPS, 4), "traffic": None,
                        "ops_per_point": round(info["ops_per_point"], 1),
                        "kernel_ms": round(kernel_ms, 4)}}
    if cpu:
        # the oracle's OSO interpreter, group after group, on a bounded
        # sample of every group; the full-workload time is extrapolated
        # per group from its sample
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as O
        nthreads = host_threads()
        est, sampled, wall = 0.0, 0, 0.0
        for k, spec in enumerate(specs):
            m = min(int(counts[k]), 1024)
            w = host[k]
            g = O.Group(spec[0], spec[1])
            P, N = w["P"][:, :m].copy(), w["N"][:, :m].copy()
            wo, wi = w["wo"][:, :m].copy(), w["wi"][:, :m].copy()
            t0 = time.perf_counter()
            g.run(P, N, wo, wi, nthreads=nthreads)
            dt = time.perf_counter() - t0
            wall += dt
            sampled += m
            est += dt * int(counts[k]) / m
        res["cpu_baseline"] = {"value": round(total / est / 1e6, 4), "unit": "Mpts/s",
                               "cores": nthreads, "kind": "port",
                               "sample": "oracle OSO interpreter (not a compiled path), groups "
                                         "one by one, first min(n, 1024) points of each "
                                         "(%d points, %.1f s), extrapolated per group"
                                         % (sampled, wall)}
        res["gpu_over_cpu"] = round(res["Mpts_per_s"] / res["cpu_baseline"]["value"], 1)
    return res


SURROGATE_TRAIN = 1 << 24   # training samples (seed 5)
SURROGATE_CALIB = 1 << 20   # calibration queries (seed 7)
SURROGATE_QUERIES = 1 << 23  # benchmark queries (seed 6)
MLP_FLOP_PER_QUERY = 2 * (14 * 64 + 3 * 64 * 64 + 64 * 4)
# predicted weights below this go to the exact closure (spot check)
SPOT_CHECK_MIN_WEIGHT = 1e-3
MFMA_BF16_PEAK = 2.5e15


def bench_surrogate(steps, warmup, train_steps, batch=16384, budget=0.01):
    """Train the closure surrogate on GPU-labelled data, calibrate its
    routing, then time exact and surrogate evaluation of 8M queries."""
    import torch
    from openshadinglanguage_amd import check, lib
    from openshadinglanguage_amd import surrogate as S
    from openshadinglanguage_amd import workloads as W
    dev = torch.device("cuda:0")
    m = W.layered_material()
    p = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)

    def exact(q):
        out = torch.empty((4, q.shape[1]), device=dev)
        check(lib().oslmi_layered_eval_local(ctypes.byref(m), q.shape[1], p(q), p(out), sp))
        return out

    t0 = time.perf_counter()
    qtr = torch.from_numpy(S.upper_hemisphere_queries(SURROGATE_TRAIN, seed=5)).to(dev)
    ttr = exact(qtr)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    s = S.Surrogate(max_batch=batch)
    s.fit_output_normalization(ttr)
    print("surrogate: %d labelled samples in %.1f s; training" % (SURROGATE_TRAIN, t_data),
          file=sys.stderr, flush=True)
    tr = s.train(qtr, ttr, steps=train_steps, batch=batch, lr=2e-3, log_every=0,
                 progress=lambda k: print("surrogate: step %d" % k, file=sys.stderr, flush=True))
    _, final_loss = s.grads(qtr[:, :batch].contiguous(), ttr[:, :batch].contiguous())
    del qtr, ttr

    # calibration on held-out queries: network error per routing cell
    qc = torch.from_numpy(S.upper_hemisphere_queries(SURROGATE_CALIB, seed=7)).to(dev)
    tc = exact(qc)
    pc = torch.empty_like(tc)
    s.infer(qc, pc)
    torch.cuda.synchronize()
    ew_c, ep_c = S.relative_errors(pc.cpu().numpy(), tc.cpu().numpy())
    table, _, predicted, pfb = S.calibrate(pc.cpu().numpy(), tc.cpu().numpy(), qc.cpu().numpy(),
                                           budget, min_weight=SPOT_CHECK_MIN_WEIGHT)
    calib = {"queries": SURROGATE_CALIB,
             "network_mean_err": [float(ew_c.mean()), float(ep_c.mean())],
             "predicted_mean_err_after_routing": list(predicted),
             "predicted_fallback_fraction": pfb, "margin": 0.8,
             "spot_check_min_weight": SPOT_CHECK_MIN_WEIGHT}
    s.set_routing(table)
    s.set_spot_check(SPOT_CHECK_MIN_WEIGHT)
    del qc, tc, pc

    # benchmark queries from the closure workload's distribution
    qn = W.surrogate_queries(SURROGATE_QUERIES, seed=6)
    q = torch.from_numpy(qn).to(dev)
    n = q.shape[1]
    ex = torch.empty((4, n), device=dev)
    out = torch.empty((4, n), device=dev)
    fbi = torch.empty(n, dtype=torch.int32, device=dev)
    fbn = torch.zeros(1, dtype=torch.int32, device=dev)

    def step_exact():
        check(lib().oslmi_layered_eval_local(ctypes.byref(m), n, p(q), p(ex), sp))

    def step_sur():
        s.infer(q, out, material=m, fb_index=fbi, fb_count=fbn, stream=stream)

    ms_exact, wall_exact = _timed(step_exact, stream, steps, warmup)
    ms_sur, wall_sur = _timed(step_sur, stream, steps, warmup)
    # the network alone (no routing, no fallback) for the MFMA roofline
    net = torch.empty_like(out)
    ms_net, _ = _timed(lambda: s.infer(q, net, stream=stream), stream, steps, warmup)
    torch.cuda.synchronize()
    ew, ep = S.relative_errors(out.cpu().numpy(), ex.cpu().numpy())
    answered = ~((qn[4] > 0) & (qn[2] <= 0))
    nfb = int(fbn.item())
    # the network evaluates only the queries it answers in that run: not the
    # exact zeros (L below the surface), not those outside the training
    # domain (NaN without a routing table)
    n_net = int((answered & np.isfinite(net.cpu().numpy()).all(axis=0)).sum())
    flop = MLP_FLOP_PER_QUERY * n_net / (ms_net / 1e3)
    return {
        "workload": "8M closure queries (seed 6): exact layered closure vs 14-4x64-4 SiLU MLP "
                    "(BF16 MFMA) with calibrated exact fallback",
        "queries": n,
        "exact_evals_per_s": round(n / (ms_exact / 1e3), 1),
        "surrogate_evals_per_s": round(n / (ms_sur / 1e3), 1),
        "speedup": round(ms_exact / ms_sur, 3),
        "exact_kernel_ms": round(ms_exact, 4), "surrogate_ms": round(ms_sur, 4),
        "network_only_ms": round(ms_net, 4),
        "fallback_fraction": round(nfb / n, 4),
        "routed_cells": int(table.sum()),
        "calibration": calib,
        "rel_err_weight": {"mean": float(ew[answered].mean()),
                           "p99": float(np.quantile(ew[answered], 0.99))},
        "rel_err_pdf": {"mean": float(ep[answered].mean()),
                        "p99": float(np.quantile(ep[answered], 0.99))},
        "error_over": "queries with L above the surface (the rest are exact zeros)",
        "training": {"samples": SURROGATE_TRAIN, "steps": tr["steps"], "batch": batch,
                     "seconds": round(tr["seconds"], 2), "final_loss": round(final_loss, 6),
                     "data_seconds": round(t_data, 2)},
        "roofline": {"bound": "mfma", "achieved": round(flop / 1e12, 2),
                     "peak": MFMA_BF16_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": round(flop / MFMA_BF16_PEAK, 4), "traffic": None,
                     "flop_per_query": MLP_FLOP_PER_QUERY, "network_queries": n_net,
                     "kernel_ms": round(ms_net, 4),
                     "note": "one fused kernel reads and routes all queries and runs the "
                             "network on network_queries of them; flop counts those only"},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="layered3", choices=sorted(WORKLOADS))
    ap.add_argument("--path", default="jit", choices=["jit", "prebuilt"],
                    help="jit: group built from OSO and lowered by the emitter (default); "
                         "prebuilt: the ahead-of-time lowering")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-closure", action="store_true",
                    help="skip the layered-closure measurement")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the surrogate training + measurement")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 64-group batch measurement")
    ap.add_argument("--no-small", action="store_true",
                    help="skip the 256x256 single-layer group (graph-replayed) measurement")
    ap.add_argument("--train-steps", type=int, default=120000)
    # 6 passes over the full workload: about 1 s wall, 15-20 s of CPU work
    # on 16 host threads
    ap.add_argument("--cpu-passes", type=int, default=6)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain 256 MiB device copy (calibrates rocprof FETCH/WRITE_SIZE)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the shading path has no "
              "cross-GPU exchange; --gpus %d is not supported)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_calibration:
        nbytes = 256 << 20
        src = torch.ones(nbytes // 4, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": nbytes, "bytes_written": nbytes,
                          "launches": args.steps}))
        return

    from openshadinglanguage_amd import check, lib
    from openshadinglanguage_amd import workloads as W

    check(lib().oslmi_init())
    dev = torch.device("cuda:0")
    name = args.workload
    res = WORKLOADS[name]["res"]
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731

    if name == "layered3":
        inp = W.layered3_inputs(res, seed=2)
        n = inp["P"].shape[1]
        t = {k: torch.from_numpy(inp[k]).to(dev) for k in ("P", "N", "wo", "wi")}
        out = torch.empty((4, n), device=dev)
        dx = inp["dPdx"].ctypes.data_as(ctypes.c_void_p)
        dy = inp["dPdy"].ctypes.data_as(ctypes.c_void_p)

        if args.path == "jit":
            from openshadinglanguage_amd import shadingsys as S
            ss, grp = S.build_layered3()
            ss.jit_group(grp)
            gl = {"P": t["P"], "N": t["N"], "dPdx": tuple(inp["dPdx"]), "dPdy": tuple(inp["dPdy"])}
            step = ss.bind(grp, n, gl, wo=t["wo"], wi=t["wi"], eval_out=out, stream=stream)
        else:
            def step():
                check(lib().oslmi_shade_layered3(n, p(t["P"]), p(t["N"]), p(t["wo"]),
                                                 p(t["wi"]), dx, dy, p(out), None, sp))
        config = {"workload": "layered3: 3-layer group (simplex fBm 6 oct + vector perlin + "
                              "anisotropic 2D gabor -> mix/smoothstep/blackbody/transformc -> "
                              "diffuse*Cd + ggx microfacet, 1 closure eval/pt), derivs on",
                  "points": n, "grid": "%dx%d" % (res, res), "precision": "fp32"}
    else:
        inp = W.fbm_cell_inputs(res, seed=1)
        n = inp["P"].shape[1]
        Pt = torch.from_numpy(inp["P"]).to(dev)
        out = torch.empty_like(Pt)

        if args.path == "jit":
            from openshadinglanguage_amd import shadingsys as S
            ss, grp = S.build_fbm_cell()
            ss.jit_group(grp)

            step = ss.bind(grp, n, {"P": Pt}, outputs=out, stream=stream)
        else:
            def step():
                check(lib().oslmi_shade_fbm_cell(n, p(Pt), p(out), sp))
        config = {"workload": "fbm_cell: 4-octave perlin fBm + 0.25 cellnoise(3P) -> "
                              "smoothstep/mix, no derivs",
                  "points": n, "grid": "%dx%d" % (res, res), "precision": "fp32"}

    config["path"] = ("OSO group -> emitter -> hiprtc kernel" if args.path == "jit"
                      else "prebuilt lowering")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    value = n / (ms_per_step / 1e3) / 1e6

    ops = _ops_per_point(name)
    achieved = ops * n / (kernel_ms / 1e3)
    roof = {"bound": "valu", "achieved": round(achieved / 1e12, 3),
            "peak": round(VALU_PEAK_OPS / 1e12, 3), "unit": "Tops/s",
            "frac": round(achieved / VALU_PEAK_OPS, 4),
            "traffic": MEASURED_TRAFFIC[name],
            "ops_per_point": round(ops, 1),
            "kernel_ms": round(kernel_ms, 4),
            "hbm_algorithmic_GBps": round(WORKLOADS[name]["bytes_per_point"] * n
                                          / (kernel_ms / 1e3) / 1e9, 2),
            "hbm_peak_GBps": HBM_PEAK_BPS / 1e9}

    line = {"metric": "Mpts/s shaded (GPU vs host batched CPU)", "value": round(value, 2),
            "unit": "Mpts/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded)",
            "config": config, "roofline": roof}
    if not args.no_cpu:
        nthreads = host_threads()
        cpu_inp = inp if name == "fbm_cell" else inp
        line["cpu_baseline"] = cpu_baseline(name, cpu_inp, nthreads, args.cpu_passes)
        line["gpu_over_cpu"] = round(value / line["cpu_baseline"]["value"], 1)
    if not args.no_small and name != "fbm_cell":
        line["small"] = bench_small(max(20, args.steps), args.warmup)
    if not args.no_batch:
        line["batch"] = bench_batch(max(5, args.steps // 2), args.warmup, not args.no_cpu)
    if not args.no_closure:
        line["closure"] = bench_closure(max(5, args.steps // 2), args.warmup, not args.no_cpu,
                                        args.cpu_passes)
    if not args.no_surrogate:
        line["surrogate"] = bench_surrogate(max(5, args.steps // 2), args.warmup,
                                            args.train_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
