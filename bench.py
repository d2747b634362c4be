"""Benchmark of the NMF multiplicative-update hot path on one MI355X.

A step is one multiplicative-update iteration (H update then W update) over the
whole matrix.  Default workload: Euclidean NMF, V 16384 x 8192 = W0 H0 + U[0,0.01)
noise (W0, H0 ~ U[0,1)), r = 64, synthetic data from the seeded generator.
Prints ONE JSON line (rank 0).
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

# fp32x: FP32 products as six BF16 piece products each, so its matrix-core ceiling in FP32 flops
# is the BF16 peak / 6
PEAK = {"fp32_mfma_tflops": 157.3, "bf16_mfma_tflops": 2500.0, "fp8_mfma_tflops": 5000.0,
        "fp32x_mfma_tflops": 2500.0 / 6, "hbm_gbs": 8000.0}

WORKLOADS = {
    # name: (rule, n, m, r, r0, seed)
    "euclid-256x128-r16": ("euclidean", 256, 128, 16, 16, 0x5EED + 1),
    "euclid-16384x8192-r64": ("euclidean", 16384, 8192, 64, 64, 0x5EED + 4),
    "euclid-8192x4096-r64": ("euclidean", 8192, 4096, 64, 64, 0x5EED + 2),
    "kl-4096x4096-r32": ("kl", 4096, 4096, 32, 32, 0x5EED + 3),
}


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def load_traffic(workload):
    """HBM bytes per launch of the pass kernels from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by what the counter reads for a copy of known size).
    Counters cannot be read live; an absent summary gives traffic null."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{workload}.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        d = json.load(f)
    out = {k: v["bytes_per_launch"] for k, v in d.get("kernels", {}).items()}
    out["_source"] = f"profiles/pmc_traffic_{workload}.json"
    return out


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="euclid-16384x8192-r64", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="bf16", choices=["fp32", "bf16", "fp8", "fp32x"],
                    help="precision of the headline value (the FP32 path is always measured beside it)")
    ap.add_argument("--extra-legs", default="fp8,fp32x",
                    help="further precisions measured and reported beside the headline (comma list, '' = none)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-single-seconds", type=float, default=0.0,
                    help="also time the reference-faithful single-thread restatement (naive loops, the reference's "
                         "operation order) for about this long, at least one iteration (0 = off; a C4 iteration "
                         "takes tens of seconds)")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--serve-jobs", type=int, default=4096,
                    help="jobs of the batched-serving mix run through the server (0 = skip serving)")
    ap.add_argument("--serve-seq-jobs", type=int, default=4096,
                    help="how many of those jobs are also run one at a time for the comparison")
    ap.add_argument("--serve-rate", type=float, default=2000.0,
                    help="Poisson arrival rate of the jobs (per s; the default saturates the server)")
    ap.add_argument("--serve-low-rate", type=float, default=100.0,
                    help="a second, sub-saturating arrival rate for the batched server (0 = skip)")
    ap.add_argument("--serve-autotune", type=int, default=2,
                    help="timed iterations per candidate of the server's batch autotuner (0 = autotuner off)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 512 MiB device to device three times (the known-size kernel that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters of a rocprofv3 --pmc run) and exit")
    ap.add_argument("--no-reduction", action="store_true",
                    help="skip the reduction-framework leg (sum / min / max over the FP32 V)")
    ap.add_argument("--no-gemm", action="store_true",
                    help="skip the container-GEMM leg (W H of the workload, square 4096^3 GEMMs per operand format)")
    ap.add_argument("--no-class", action="store_true",
                    help="skip the drop-in class legs (DoIteration through the object handle, FP32)")
    ap.add_argument("--no-additive", action="store_true",
                    help="skip the additive-update legs (NMF_AdditiveEuclidian / NMF_AdditiveDivergence, FP32)")
    ap.add_argument("--gap-iterations", default="100,500,1000",
                    help="iteration counts at which the objective gap to the FP32 path is reported ('' = none)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X, and the NMF path "
              "has no multi-GPU mode (the row-sharded design is documented, not built). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    from gpumlib_amd import capi, hip
    from gpumlib_amd import nmf as G

    lib = capi.load()
    if args.copy_calibration:
        nbytes = 512 << 20
        src, dst = hip.Buffer(nbytes), hip.Buffer(nbytes)
        src.fill_zero()
        for _ in range(3):
            dst.copy_from(src, nbytes)
        hip.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}), flush=True)
        return
    rule_name, n, m, r, r0, seed = WORKLOADS[args.workload]
    rule = G.EUCLIDEAN if rule_name == "euclidean" else G.KL
    if rule == G.KL and args.precision not in ("fp32", "bf16"):
        args.precision = "fp32"  # the KL rule runs with FP32 or BF16 storage
    eps = G.REFERENCE_EPS if rule == G.EUCLIDEAN else 1e-9

    t_gen = time.time()
    V = G.gen_euclidean(n, m, r0, seed) if rule == G.EUCLIDEAN else G.gen_poisson(n, m, r0, seed)
    W0, H0 = G.gen_init(n, m, r, seed)
    t_gen = time.time() - t_gen

    obj_kind = G.OBJ_FROBENIUS if rule == G.EUCLIDEAN else G.OBJ_KL
    if rule == G.EUCLIDEAN:
        flops_h = 2.0 * n * m * r + 2.0 * n * r * r
        flops_w = 2.0 * n * m * r + 2.0 * m * r * r
        flops_iter = 4.0 * n * m * r + 4.0 * r * r * (n + m)
    else:
        flops_h = flops_w = 4.0 * n * m * r
        flops_iter = 8.0 * n * m * r

    def leg(prec_name):
        prec = {"fp32": G.FP32, "bf16": G.BF16, "fp8": G.FP8, "fp32x": G.FP32X}[prec_name]
        t_setup = time.perf_counter()
        d = G.DeviceNMF(rule, V, W0, H0, n, m, r, eps=eps, precision=prec)
        t_setup = time.perf_counter() - t_setup
        vbytes_el = {G.FP32: 4, G.BF16: 2, G.FP8: 1, G.FP32X: 4}[prec]
        obj0 = d.objective(obj_kind)
        # Objective (against the FP32 V, FP64 reduction on the GPU) after fixed iteration
        # counts from the initial W, H: the precision gap is stated for every result.  These
        # untimed iterations run first, so the timed ones below start on a device already in
        # its steady state; W and H are then reset to the initial factors.
        gaps = {}
        if gap_iters:
            done = 0
            for t in gap_iters:
                d.launch(t - done)
                done = t
                gaps[t] = d.objective_fp32v(obj_kind)
            d.set_wh(W0, H0)
        d.launch(args.warmup)
        d.stream.synchronize()
        t0 = time.perf_counter()
        d.launch(args.steps)
        d.stream.synchronize()
        t1 = time.perf_counter()
        ms = (t1 - t0) * 1e3 / args.steps
        t_d2h = time.perf_counter()
        d.get_w()
        d.get_h()
        t_d2h = time.perf_counter() - t_d2h
        obj1 = d.objective(obj_kind)
        # The same K iterations with the host hand-over around them: V, W, H from pageable host
        # memory, the V store built, W and H copied back (never the headline value).
        e2e = t_setup + (t1 - t0) + t_d2h
        end_to_end = {"setup_s": t_setup, "upload_s": d.setup_s["upload"], "pack_s": d.setup_s["pack"],
                      "upload_gbs": 4.0 * (n * m + (n + m) * r) / d.setup_s["upload"] / 1e9,
                      "iterations": args.steps, "iterate_s": t1 - t0, "download_s": t_d2h,
                      "iters_per_s_incl_transfers": args.steps / e2e}

        # Per-pass durations: HIP events on the stream the kernels are launched on, with
        # the H and W passes alternating as they do in an iteration (each streams its own
        # copy of V, so neither finds the other's data in the last-level cache).
        def time_passes():
            for _ in range(2):
                d.run_pass(0)
                d.run_pass(1)
            evs = [hip.Event() for _ in range(2 * args.kernel_reps + 1)]
            evs[0].record(d.stream)
            for k in range(args.kernel_reps):
                d.run_pass(0)
                evs[2 * k + 1].record(d.stream)
                d.run_pass(1)
                evs[2 * k + 2].record(d.stream)
            evs[-1].synchronize()
            th = sum(evs[2 * k].elapsed_ms(evs[2 * k + 1]) for k in range(args.kernel_reps)) / args.kernel_reps
            tw = sum(evs[2 * k + 1].elapsed_ms(evs[2 * k + 2]) for k in range(args.kernel_reps)) / args.kernel_reps
            return th, tw

        ms_h, ms_w = time_passes()

        dom = ("pass_h", ms_h, flops_h) if ms_h >= ms_w else ("pass_w", ms_w, flops_w)
        out = {
            "iters_per_s": 1e3 / ms, "ms_per_step": ms,
            "mfma_peak_frac": flops_iter / (ms * 1e-3) / 1e12 / PEAK[f"{prec_name}_mfma_tflops"],
            "v_stream_gbs": 2.0 * vbytes_el * n * m / (ms * 1e-3) / 1e9,
            "objective_entry": obj0, "objective_after": obj1, "objective_at": gaps, "end_to_end": end_to_end,
            "kernels": {"pass_h_ms": ms_h, "pass_w_ms": ms_w,
                        "pass_h_tflops": flops_h / (ms_h * 1e-3) / 1e12,
                        "pass_w_tflops": flops_w / (ms_w * 1e-3) / 1e12},
        }
        tr = traffic.get(f"{prec_name}.{dom[0]}")
        if prec == G.FP32:
            ach = dom[2] / (dom[1] * 1e-3) / 1e12
            out["roofline"] = {"kernel": dom[0], "bound": "mfma", "achieved": ach, "peak": PEAK["fp32_mfma_tflops"],
                               "unit": "TFLOP/s", "frac": ach / PEAK["fp32_mfma_tflops"], "traffic": tr}
        else:
            vb = float(vbytes_el) * n * m  # algorithmic bytes of one pass: V streamed once
            ach = vb / (dom[1] * 1e-3) / 1e9
            out["roofline"] = {"kernel": dom[0], "bound": "hbm", "achieved": ach, "peak": PEAK["hbm_gbs"],
                               "unit": "GB/s", "frac": ach / PEAK["hbm_gbs"], "traffic": tr}
        if tr is not None:
            out["roofline"]["traffic_source"] = traffic["_source"]
        return out

    gap_iters = sorted(int(x) for x in args.gap_iterations.split(",") if x.strip())
    traffic = load_traffic(args.workload)
    legs = {"fp32": leg("fp32")}
    if args.precision != "fp32":
        legs[args.precision] = leg(args.precision)
    # storage precisions each rule runs: the KL rule has FP32 and BF16 storage
    runs = ("fp32", "bf16", "fp8", "fp32x") if rule == G.EUCLIDEAN else ("fp32", "bf16")
    for extra in (x.strip() for x in args.extra_legs.split(",")):
        if extra and extra not in legs and extra in runs:
            legs[extra] = leg(extra)
            legs[extra]["speedup_vs_fp32_path"] = legs[extra]["iters_per_s"] / legs["fp32"]["iters_per_s"]
            if gap_iters:
                legs[extra]["objective_rel_gap_vs_fp32_path"] = {
                    str(t): abs(legs[extra]["objective_at"][t] - legs["fp32"]["objective_at"][t])
                    / legs["fp32"]["objective_at"][t] for t in gap_iters}
    # Additive-update variants on the same workload (FP32, the same pass kernels with the
    # additive update): iterations per second and objectives, for the record.
    def additive_leg(arule):
        d = G.DeviceNMF(arule, V, W0, H0, n, m, r, eps=eps)
        obj0 = d.objective(obj_kind)
        d.launch(args.warmup)
        d.stream.synchronize()
        t0 = time.perf_counter()
        d.launch(args.steps)
        d.stream.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        flops = flops_w if arule == G.ADDITIVE_EUCLIDEAN else flops_iter  # that rule runs the W half only
        return {"iters_per_s": 1e3 / ms, "ms_per_step": ms, "tflops": flops / (ms * 1e-3) / 1e12,
                "objective_entry": obj0, "objective_after": d.objective(obj_kind),
                "iterations_timed": args.steps, "precision": "fp32"}

    # The drop-in C++ classes (NMF_MultiplicativeEuclidianDistance / NMF_MultiplicativeDivergence through
    # the object handle, FP32): what a caller of the reference API sees per DoIteration(), with the
    # default quality cadence (the reference quality, nmf_kernels_quality.cu:35, fused into every
    # iteration's pass over V, as nmf_base.h computes it whenever it can) and with it off.
    def class_leg(quality_interval, read_each=False):
        h = ctypes.c_void_p()
        ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        capi.check(lib.gpumlib_nmf_object_create(rule, G.FP32, n, m, r, ptr(V), ptr(W0), ptr(H0), ctypes.byref(h)),
                   "gpumlib_nmf_object_create")
        try:
            capi.check(lib.gpumlib_nmf_object_set_epsilon(h, eps))
            capi.check(lib.gpumlib_nmf_object_set_quality_interval(h, quality_interval))
            wout = np.empty(n * r, np.float32)
            capi.check(lib.gpumlib_nmf_object_do_iterations(h, args.warmup, 1))
            capi.check(lib.gpumlib_nmf_object_get_w(h, ptr(wout)))  # blocking read-back: the stream is idle
            t0 = time.perf_counter()
            capi.check(lib.gpumlib_nmf_object_get_w(h, ptr(wout)))
            t_get = time.perf_counter() - t0
            t0 = time.perf_counter()
            if read_each:  # the usual caller's loop: DoIteration(), then QualityImprovement()
                for _ in range(args.steps):
                    capi.check(lib.gpumlib_nmf_object_do_iteration(h, 1))
                    lib.gpumlib_nmf_object_quality_improvement(h)
            else:
                capi.check(lib.gpumlib_nmf_object_do_iterations(h, args.steps, 1))
            capi.check(lib.gpumlib_nmf_object_get_w(h, ptr(wout)))
            t = time.perf_counter() - t0 - t_get
            return {"iters_per_s": args.steps / t, "ms_per_step": t * 1e3 / args.steps,
                    "quality_interval": quality_interval, "quality_read_each_iteration": read_each,
                    "quality_improvement": float(lib.gpumlib_nmf_object_quality_improvement(h))}
        finally:
            lib.gpumlib_nmf_object_destroy(h)

    class_api = {}
    if not args.no_class:
        print("class-API legs", file=sys.stderr, flush=True)
        class_api = {"default_quality_cadence": class_leg(1),
                     "default_quality_cadence_read_each_iteration": class_leg(1, read_each=True),
                     "quality_off": class_leg(0),
                     "note": "DoIteration() of the drop-in class through gpumlib_nmf_object_do_iterations (or one "
                             "DoIteration() and one QualityImprovement() call at a time), FP32; timed from the "
                             "launch of the iterations to a blocking GetW(), minus that read-back's own time"}

    additive = {}
    if not args.no_additive:
        names = ({"additive_euclidean": G.ADDITIVE_EUCLIDEAN, "additive_euclidean_with_h": G.ADDITIVE_EUCLIDEAN_WITH_H}
                 if rule == G.EUCLIDEAN else {"additive_kl": G.ADDITIVE_KL})
        for name, arule in names.items():
            print(f"additive leg {name}", file=sys.stderr, flush=True)
            additive[name] = additive_leg(arule)

    # Reduction framework (Reduction::Sum / MinIndex / MaxIndex) over the FP32 V of the
    # workload: HBM-bound, 4 bytes read per element; HIP events on the launching stream.
    def reduction_leg():
        x = hip.Buffer.from_array(np.asarray(V, dtype=np.float32))
        count = n * m
        wsb = int(lib.gpumlib_reduce_workspace_bytes(count))
        ws, out = hip.Buffer(wsb), hip.Buffer(16)
        st = hip.Stream()
        calls = {
            "sum": lambda: lib.gpumlib_reduce_sum(x.ptr, count, 1.0, out.ptr, ws.ptr, wsb, st.handle),
            "min_index": lambda: lib.gpumlib_reduce_min(x.ptr, count, out.ptr, out.ptr + 4, ws.ptr, wsb, st.handle),
            "max_index": lambda: lib.gpumlib_reduce_max(x.ptr, count, out.ptr, out.ptr + 4, ws.ptr, wsb, st.handle),
        }
        res = {}
        reps = max(5, args.kernel_reps)
        for name, fn in calls.items():
            for _ in range(3):
                capi.check(fn(), name)
            e0, e1 = hip.Event(), hip.Event()
            e0.record(st)
            for _ in range(reps):
                capi.check(fn(), name)
            e1.record(st)
            e1.synchronize()
            us = e0.elapsed_ms(e1) * 1e3 / reps
            gbs = 4.0 * count / (us * 1e-6) / 1e9
            res[name] = {"us_per_call": us, "gbs": gbs, "frac_of_hbm_peak": gbs / PEAK["hbm_gbs"]}
        x.close()
        return {"elements": count, "bytes": 4 * count, "calls": res,
                "note": "two launches per call (partials, then one workgroup combining them); "
                        "GB/s = 4 bytes per element / time per call"}

    reduction = None if args.no_reduction else reduction_leg()

    # Container GEMM (DeviceMatrix::Multiply / GetWH): W H of the workload (n x m from n x r times
    # r x m, FP32: bound by writing the n x m product), and a square 4096^3 GEMM in each operand
    # format against that format's matrix-core peak.  HIP events on the launching stream.
    def gemm_leg():
        st = hip.Stream()

        def timed(fn, reps):
            for _ in range(2):
                capi.check(fn(), "gemm")
            e0, e1 = hip.Event(), hip.Event()
            e0.record(st)
            for _ in range(reps):
                capi.check(fn(), "gemm")
            e1.record(st)
            e1.synchronize()
            return e0.elapsed_ms(e1) * 1e3 / reps

        Wd = hip.Buffer.from_array(np.asarray(W0, dtype=np.float32))
        Hd = hip.Buffer.from_array(np.asarray(H0, dtype=np.float32))
        WH = hip.Buffer(4 * n * m)
        us = timed(lambda: lib.gpumlib_nmf_sgemm(b"N", b"N", n, m, r, 1.0, Wd.ptr, n, Hd.ptr, r, 0.0, WH.ptr, n,
                                                 st.handle), max(5, args.kernel_reps))
        out = {"get_wh": {"shape": f"{n}x{m}x{r}", "us_per_call": us,
                          "tflops": 2.0 * n * m * r / (us * 1e-6) / 1e12,
                          "write_gbs": 4.0 * n * m / (us * 1e-6) / 1e9}}
        for x in (Wd, Hd, WH):
            x.close()
        S = 4096
        A, B, Cm = hip.Buffer(4 * S * S), hip.Buffer(4 * S * S), hip.Buffer(4 * S * S)
        A.fill_zero()
        B.fill_zero()
        calls = {
            "fp32": (lambda: lib.gpumlib_nmf_sgemm(b"N", b"N", S, S, S, 1.0, A.ptr, S, B.ptr, S, 0.0, Cm.ptr, S,
                                                   st.handle), PEAK["fp32_mfma_tflops"]),
            "bf16": (lambda: lib.gpumlib_nmf_gemm_bf16(b"N", b"N", S, S, S, 1.0, A.ptr, S, B.ptr, S, 0.0, Cm.ptr, S,
                                                       st.handle), PEAK["bf16_mfma_tflops"]),
            "fp8": (lambda: lib.gpumlib_nmf_gemm_fp8(b"N", b"N", S, S, S, 1.0, A.ptr, S, 127, B.ptr, S, 127, 0.0,
                                                     Cm.ptr, S, st.handle), PEAK["fp8_mfma_tflops"]),
        }
        for name, (fn, peak) in calls.items():
            us = timed(fn, 5)
            tf = 2.0 * S ** 3 / (us * 1e-6) / 1e12
            out[f"square_{name}"] = {"shape": f"{S}x{S}x{S}", "us_per_call": us, "tflops": tf,
                                     "frac_of_mfma_peak": tf / peak}
        for x in (A, B, Cm):
            x.close()
        out["note"] = ("gpumlib_nmf_sgemm / _gemm_bf16 / _gemm_fp8, column-major 'N','N'; 128 x 128 tiles on the "
                       "operand format's matrix cores")
        return out

    gemm = None if args.no_gemm else gemm_leg()

    head = legs[args.precision]
    line = {
        "metric": "NMF iter/s (16384x8192, r=64)" if args.workload.startswith("euclid-16384") else f"NMF iter/s ({args.workload})",
        "value": head["iters_per_s"],
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic (seeded SplitMix64 generator: V = W0 H0 + U[0,0.01) noise)" if rule == G.EUCLIDEAN
        else "synthetic (seeded SplitMix64 generator: V ~ Poisson(W0 H0), W0, H0 ~ Exp(1))",
        "config": {"workload": args.workload, "rule": rule_name, "n": n, "m": m, "r": r,
                   "precision": args.precision, "eps": eps, "parallelism": "single GPU"},
        "mfma_peak_frac": head["mfma_peak_frac"],
        "roofline": head["roofline"],
        "legs": legs,
        "speedup_vs_fp32_path": head["iters_per_s"] / legs["fp32"]["iters_per_s"],
        "objective_rel_gap_vs_fp32_path": {
            str(t): abs(head["objective_at"][t] - legs["fp32"]["objective_at"][t]) / legs["fp32"]["objective_at"][t]
            for t in gap_iters},
        "objective_note": "objectives of both paths' W, H against the FP32 V (FP64 reduction on the GPU) after "
                          "the stated iteration counts from the same initial W, H; the FP32 path itself is held "
                          "to the FP64 host oracle within 1e-4 by the parity tests",
        "gen_seconds": t_gen,
    }
    if class_api:
        line["class_api"] = class_api
    if additive:
        line["additive_legs"] = additive
    if reduction:
        line["reduction"] = reduction
    if gemm:
        line["gemm"] = gemm

    if args.serve_jobs > 0:
        from gpumlib_amd import serving

        print(f"serving: generating {args.serve_jobs} jobs", file=sys.stderr, flush=True)
        jobs = serving.make_jobs(args.serve_jobs)
        data = serving.make_data(jobs)

        def public(res):
            return {k: v for k, v in res.items() if not k.startswith("_")}

        print(f"serving: batched run at {args.serve_rate:g}/s", file=sys.stderr, flush=True)
        srv = serving.run_server(lib, jobs, data, rate=args.serve_rate, autotune=args.serve_autotune)
        line["serving"] = {
            "workload": f"{len(jobs)} jobs, Euclidean/KL 50/50, V n x m in GPUMLib's naming (n rows log-uniform "
                        f"[64,2048], m columns log-uniform [64,4096]), r in {{8,16,32,64}}, to relative objective "
                        f"change < 1e-4 or 500 iterations (checked every 10-iteration round), Poisson arrivals",
            "batched": public(srv),
        }
        if args.serve_low_rate > 0:
            print(f"serving: batched run at {args.serve_low_rate:g}/s", file=sys.stderr, flush=True)
            line["serving"]["batched_low_rate"] = public(
                serving.run_server(lib, jobs, data, rate=args.serve_low_rate, autotune=args.serve_autotune))
        nseq = min(args.serve_seq_jobs, len(jobs))
        if nseq > 0:
            # the unbatched comparison point: the same jobs through the same server, one at a
            # time, each through the single-problem fused iteration (pooled device memory, no
            # host work per round beyond reading the two objectives)
            print(f"serving: one-at-a-time run of {nseq} jobs", file=sys.stderr, flush=True)
            seq = serving.run_server(lib, jobs[:nseq], data[:nseq], rate=args.serve_rate, sequential=1)
            one = public(seq)
            if args.serve_low_rate > 0:
                one["at_low_rate"] = {"arrival_rate_per_s": args.serve_low_rate,
                                      **serving.queued_latency(seq["_service_s"], args.serve_low_rate),
                                      "note": "one FIFO server fed the low-rate Poisson stream, from the service "
                                              "times measured in this run"}
            line["serving"]["one_at_a_time"] = one
            line["serving"]["jobs_per_s_speedup"] = srv["jobs_per_s"] / seq["jobs_per_s"]
            if nseq < len(jobs):
                line["serving"]["note"] = f"one-at-a-time over the first {nseq} of the {len(jobs)} jobs"

    if not args.no_cpu_baseline:
        from oracle import oracle as orc

        Vh = np.ascontiguousarray(V)
        Wh, Hh = W0.copy(), H0.copy()
        fn = orc.euc_iter if rule == G.EUCLIDEAN else orc.kl_iter
        fn(n, m, r, Vh, Wh, Hh, eps, iters=1, omp=True)  # warm-up
        k, t, each = 0, 0.0, []
        while t < args.cpu_seconds or k < 5:
            s0 = time.perf_counter()
            fn(n, m, r, Vh, Wh, Hh, eps, iters=1, omp=True)
            each.append(time.perf_counter() - s0)
            t += each[-1]
            k += 1
        line["cpu_baseline"] = {"value": k / t, "unit": "iter/s", "median_iter_per_s": 1.0 / float(np.median(each)),
                                "cores": orc.omp_threads(), "kind": "port",
                                "cpu": host_cpu_model(), "sample": f"{k} iterations of the same workload ({args.workload}, FP32 update rules, "
                                          f"OpenMP host-cores restatement) after 1 warm-up; value over "
                                          f"their total time, median_iter_per_s from the median one"}
    if args.cpu_single_seconds > 0:
        from oracle import oracle as orc

        Vh = np.ascontiguousarray(V)
        Wh, Hh = W0.copy(), H0.copy()
        fn = orc.euc_iter if rule == G.EUCLIDEAN else orc.kl_iter
        k, t = 0, 0.0
        while t < args.cpu_single_seconds or k < 1:
            s0 = time.perf_counter()
            fn(n, m, r, Vh, Wh, Hh, eps, iters=1, omp=False)
            dt = time.perf_counter() - s0
            t += dt
            k += 1
            if dt > 1.0 or k % 100 == 0:
                print(f"cpu single thread: {k} iterations, {t:.1f} s", file=sys.stderr, flush=True)
        line["cpu_reference_faithful"] = {
            "value": k / t, "unit": "iter/s", "cores": 1, "kind": "port", "cpu": host_cpu_model(),
            "sample": f"{k} iterations of the same workload ({args.workload}), single thread, naive loops in the "
                      f"reference's operation order (no warm-up: each iteration is long)"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
