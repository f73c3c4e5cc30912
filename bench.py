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
# FMA as two flops). Measured on the box: k_layered3 issues 1.24 VALU
# wave-instructions per CU-clock (SQ_INSTS_VALU / GRBM_GUI_ACTIVE), which
# only a 2-per-clock issue rate allows.
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
# scale), WRITE_SIZE reads them exactly (x1); profiles/layered3_pmc.txt.
MEASURED_TRAFFIC = {"layered3": 2 * 24816.98 * 1024 + 16384.0 * 1024, "fbm_cell": None}


def _ops_per_point(name):
    path = os.path.join(ROOT, "openshadinglanguage_amd", "roofline_ops.json")
    return json.load(open(path))[name]["ops_per_point"]


def cpu_baseline(name, inputs, nthreads, passes):
    """The CPU oracle (reference-algorithm restatement) on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    O.lib()
    if name == "layered3":
        n = inputs["P"].shape[1]
        fn = lambda: O.shade_c2(inputs["P"], inputs["N"], inputs["wo"], inputs["wi"],  # noqa
                                inputs["dPdx"], inputs["dPdy"], nthreads=nthreads)
    else:
        n = inputs["P"].shape[1]
        fn = lambda: O.shade_c1(inputs["P"], nthreads=nthreads)  # noqa
    fn()  # warm (thread creation, page faults)
    t0 = time.perf_counter()
    for _ in range(passes):
        fn()
    dt = time.perf_counter() - t0
    return {"value": round(n * passes / dt / 1e6, 3), "unit": "Mpts/s", "cores": nthreads,
            "kind": "port",
            "sample": "%d pass(es) over the full %d-point workload, %d threads, %.1f s wall"
                      % (passes, n, nthreads, dt)}


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
                        "frac": round(achieved / VALU_PEAK_OPS, 4), "traffic": None,
                        "ops_per_point": round(ops, 1), "kernel_ms": round(kernel_ms, 4),
                        "hbm_algorithmic_GBps": round(CLOSURE_BYTES_PER_POINT * n
                                                      / (kernel_ms / 1e3) / 1e9, 2)}}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as O
        nthreads = min(16, os.cpu_count() or 1)
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
    ap.add_argument("--cpu-passes", type=int, default=2)
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
        nthreads = min(16, os.cpu_count() or 1)
        cpu_inp = inp if name == "fbm_cell" else inp
        line["cpu_baseline"] = cpu_baseline(name, cpu_inp, nthreads, args.cpu_passes)
        line["gpu_over_cpu"] = round(value / line["cpu_baseline"]["value"], 1)
    if not args.no_closure:
        line["closure"] = bench_closure(max(5, args.steps // 2), args.warmup, not args.no_cpu,
                                        args.cpu_passes)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
