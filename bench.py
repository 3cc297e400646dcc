"""Benchmark of the MadSDP interior-point hot path on one MI355X.

Default workload: the serving stream.  Independent SDP instances arrive as a
Poisson process (the seed-4 stream: half max-cut with n ~ U{20..600}, half
dense SDPs with n ~ U{16..200}, m ~ U{50..3000}), each solved to the 1e-7
relative gap (NT, EigenStep(0.95), Mehrotra(3), DenseCholesky).  A step is a
window of 16 consecutive requests; `--warmup` windows are served untimed,
then `--steps` windows as one continuous stream by the mixed-precision
continuous-batching loop (value = its solves/s), and the same arrivals by the
plain FP64 path (one solve at a time) for the speed-up, plus the batching
loop in FP64.  p50/p99 are request latencies (completion - arrival).

Other workloads (`--workload maxcut3000|maxcut100|dense250|blocks96`) time
one instance solved repeatedly, its data resident in HBM: a step is one
complete solve.

Output: one JSON line with solves/s (whole job), p50/p99 latency, the live
roofline of the dominant kernel (HIP events stamped on its dispatches over the
timed region) and the CPU baseline (the FP64 oracle on a bounded sample of the
same work, timed on this host).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "madsdp.jl_amd"))

METRIC = "solves/s + p50/p99 latency @1e-7 gap: mixed/batched vs FP64 plain, 1 MI355X"
PEAKS = {"mfma": 78.6, "mfma32": 157.3, "hbm": 8000.0}  # FP64/FP32 MFMA TFLOP/s, HBM GB/s


def build_model(workload):
    from madsdp_amd import instances
    if workload == "maxcut3000":
        return instances.maxcut_3000(), dict(
            workload="max-cut G(3000,0.01) +-1 weights: one 3000x3000 PSD block, m=3000, "
                     "plain FP64, NT + EigenStep + Mehrotra, DenseCholesky, tol 1e-7",
            n=3000, m=3000, seed=1)
    if workload == "maxcut100":
        return instances.maxcut_100(), dict(
            workload="max-cut G(100,0.1) unit weights: one 100x100 PSD block, m=100, plain FP64",
            n=100, m=100, seed=0)
    if workload == "dense250":
        return instances.dense_250_m8000(), dict(
            workload="random dense SDP: one 250x250 PSD block + 400 nonnegative scalars, m=8000 "
                     "dense constraints, NT + EigenStep + Mehrotra, DenseCholesky, tol 1e-7",
            n=250, m=8000, n_scalar=400, seed=2)
    if workload == "blocks96":
        return instances.blocks_96_m6000(), dict(
            workload="block-diagonal SOS-style SDP: 96 PSD blocks of 8..120, 200 scalars, "
                     "m=6000, each constraint dense on 1-3 blocks, tol 1e-7",
            m=6000, n_scalar=200, seed=3)
    if workload.startswith("maxcut"):
        n = int(workload[6:])
        return instances.maxcut(n, 10.0 / n, 1, "pm1"), dict(
            workload=f"max-cut G({n},{10.0 / n:.4f}) +-1 weights, plain FP64", n=n, m=n, seed=1)
    raise SystemExit(f"unknown workload {workload}")


WINDOW = 16  # requests of the serving stream per bench step


def _blas_threads():
    try:
        from threadpoolctl import threadpool_info
        return max((d.get("num_threads", 0) for d in threadpool_info()
                    if d.get("user_api") == "blas"), default=None)
    except Exception:
        return None


def stream_cpu_baseline(models, budget_s):
    """The FP64 oracle on the stream's own instances, in arrival order, as many
    whole solves as fit in about `budget_s` seconds with all BLAS threads, then
    the same instances on one thread (the reference benchmark's convention,
    benchmark/maxcut.jl:16) for at most half that budget."""
    from oracle import madsdp_oracle as O
    threads = _blas_threads() or os.cpu_count()
    done, t0 = 0, time.perf_counter()
    while done < len(models) and (done == 0 or time.perf_counter() - t0 < budget_s):
        O.Solver(models[done]).solve()
        done += 1
    t_all = time.perf_counter() - t0
    one = None
    try:
        from threadpoolctl import threadpool_limits
        with threadpool_limits(limits=1, user_api="blas"):
            k, t1 = 0, time.perf_counter()
            while k < done and (k == 0 or time.perf_counter() - t1 < budget_s / 2):
                O.Solver(models[k]).solve()
                k += 1
            one = {"value": k / (time.perf_counter() - t1), "cores": 1, "solves": k}
    except Exception:
        pass
    multi = {"value": done / t_all, "cores": threads, "solves": done}
    # the faster configuration is the baseline (one thread wins on these
    # small instances: BLAS threading costs more than it brings), the other
    # is reported beside it
    best, other = (one, multi) if one and one["value"] > multi["value"] else (multi, one)
    return {"value": best["value"], "unit": "solves/s", "cores": best["cores"], "kind": "port",
            "sample": f"FP64 oracle (numpy/scipy OpenBLAS, {best['cores']} thread(s)) solving the "
                      f"first {best['solves']} instances of the timed stream back to back "
                      f"(the first {done} took {t_all:.1f} s with {threads} threads)",
            "other_threads": other}


def run_stream(args):
    """Serving stream: the seed-4 Poisson stream of independent instances,
    half max-cut with n ~ U{20..600}, half dense SDPs with n ~ U{16..200},
    m ~ U{50..3000}, n_s = n.  A step is a window of WINDOW consecutive
    requests: the first `warmup` windows warm the library up untimed, the
    next `steps` windows are served as one continuous stream (arrival times
    kept) by the mixed-precision continuous-batching loop (`slots` concurrent
    solves), then over the same arrivals by the plain FP64 path (one solve at
    a time) and by the batching loop in FP64.  Returns the JSON record."""
    import numpy as np
    from madsdp_amd import MadSDPSolver, instances, serve
    import madsdp_amd.lib as L

    nwarm, ntimed = args.warmup * WINDOW, args.steps * WINDOW
    stream = instances.serving_stream(count=nwarm + ntimed, seed=4, rate=args.rate)
    t = time.perf_counter()
    models = [instances.build_stream_instance(kind, params) for _, kind, params in stream]
    print(f"generated {len(models)} instances in {time.perf_counter() - t:.1f} s", file=sys.stderr,
          flush=True)
    warm_models, timed_models = models[:nwarm], models[nwarm:]
    t_first = stream[nwarm][0] if ntimed else 0.0
    arrivals = [a - t_first for a, _, _ in stream[nwarm:]]

    def one(ms, arr, slots, precision, profile=False):
        t = time.perf_counter()
        # concurrent solves share the device: one launch per product
        solvers = [MadSDPSolver(model, device=0, precision=precision, shared_device=int(slots > 1))
                   for model in ms]
        prep = time.perf_counter() - t
        lib = L.load()
        if profile:
            L.check(lib.msdp_profile_begin(3), "msdp_profile_begin")
        res, summary = serve(solvers, arr, slots=slots, mem_budget=args.mem_budget << 30)
        prof = None
        if profile:
            import ctypes as C
            ms_, n_, w_ = C.c_double(), C.c_int64(), C.c_double()
            L.check(lib.msdp_profile_end(C.byref(ms_), C.byref(n_), C.byref(w_)), "msdp_profile_end")
            prof = (ms_.value, n_.value, w_.value)
        st = [int(s.status) for s in solvers]
        its = [int(s.iter) for s in solvers]
        objs = [s.primal_objective() for s in solvers]
        info = [s.info() for s in solvers]
        summary.update(prepare_s=prep, succeeded=int(sum(v == 1 for v in st)),
                       iterations_mean=float(np.mean(its)),
                       lowp_solves=int(sum(i.lowp_factorizations > 0 for i in info)),
                       fp64_fallbacks=int(sum(i.fp64_fallbacks for i in info)),
                       not_succeeded=[(k, stream[nwarm + k][1] if len(ms) == ntimed else "",
                                       st[k], its[k]) for k in range(len(st)) if st[k] != 1])
        print(f"{precision} slots={slots}: {summary}", file=sys.stderr, flush=True)
        return summary, st, its, objs, prof

    # warm-up (untimed): kernels, stream-ordered memory pools, host threads
    if nwarm:
        warm_arr = [a - stream[0][0] for a, _, _ in stream[:nwarm]]
        one(warm_models, warm_arr, args.slots, args.batched_precision)
        one(warm_models[:2], [0.0, 0.0], 1, "fp64")

    batched, st_b, it_b, obj_b, prof = one(timed_models, arrivals, args.slots,
                                           args.batched_precision, profile=True)
    plain, st_p, it_p, obj_p, _ = one(timed_models, arrivals, 1, "fp64")
    other = None
    if args.compare_precision:
        alt = "fp64" if args.batched_precision == "mixed" else "mixed"
        other = one(timed_models, arrivals, args.slots, alt)[0]
        other["precision"] = alt
    dit = [b - a for a, b in zip(it_p, it_b)]
    dobj = max(abs(a - b) / max(1.0, abs(a)) for a, b in zip(obj_p, obj_b))
    roof = None
    if prof and prof[1] > 0 and prof[0] > 0:
        ms_, n_, w_ = prof
        achieved = w_ / (ms_ * 1e-3) / 1e12
        # HBM bytes per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE passes of
        # this kernel (a plain one-stream serving of the stream's first requests:
        # counter passes need one stream)
        traffic, tsample = None, None
        tf = os.path.join(ROOT, "profiles", "trd1_traffic_stream.json")
        if os.path.exists(tf):
            with open(tf) as fh:
                tj = json.load(fh)
            traffic, tsample = tj["traffic_bytes_per_launch"], tj["sample"]
        roof = {"bound": "mfma", "achieved": achieved, "peak": PEAKS["mfma"], "unit": "TFLOP/s",
                "frac": achieved / PEAKS["mfma"], "traffic": traffic, "traffic_sample": tsample,
                "kernel": "trd1_kernel (one-workgroup Householder tridiagonalisation, n <= 640: "
                          "the NT eigenproblem and the step-length eigenvalues of the stream's "
                          "blocks); work = (4/3) n^3 flops per matrix, FP64 vector peak",
                "launches": n_, "avg_launch_us": 1e3 * ms_ / n_,
                "algorithmic_per_launch": w_ / n_,
                "kernel_share_of_timed": ms_ * 1e-3 / batched["span_s"]}
    rec = {
        "metric": METRIC,
        "value": batched["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * batched["span_s"] / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # the arithmetic the served solves actually ran: FP32 assembly only
        # where a solve was eligible for it (lowp_min_gflop, full-rank H)
        "dtype": "f64+f32" if batched.get("lowp_solves", 0) > 0 else "f64",
        "data": "synthetic (seeded numpy PCG64 instances, Poisson arrivals)",
        "config": {"workload": f"serving stream: requests {nwarm}..{nwarm + ntimed - 1} of the "
                               f"seed-4 stream (prefix of the 4096-instance stream; {WINDOW} requests "
                               f"per step, the first {nwarm} served untimed as warm-up): half max-cut "
                               f"n~U{{20..600}}, half dense n~U{{16..200}} m~U{{50..3000}} n_s=n; "
                               f"Poisson rate {args.rate}/s; tol 1e-7; {args.batched_precision} "
                               f"continuous batching with {args.slots} slots vs plain FP64 one at a "
                               f"time over the same arrivals",
                   "requests": ntimed, "window": WINDOW, "slots": args.slots, "rate": args.rate},
        "p50_latency_ms": 1e3 * batched["p50_latency_s"],
        "p99_latency_ms": 1e3 * batched["p99_latency_s"],
        "plain_fp64": plain,
        "batched": dict(batched, precision=args.batched_precision),
        "batched_other_precision": other,
        "speedup_vs_plain": batched["solves_per_s"] / plain["solves_per_s"],
        "same_status": st_b == st_p,
        "max_extra_iterations": int(max(dit)) if dit else 0,
        "max_rel_objective_diff": dobj,
        "roofline": roof,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = stream_cpu_baseline(timed_models, args.cpu_seconds)
    return rec


def cpu_baseline(model, iters_gpu, budget_s):
    """The FP64 oracle on the same instance: initialization plus as many IPM
    iterations as fit in about `budget_s` seconds (at least one); solves/s is
    extrapolated to the GPU solve's iteration count."""
    from oracle import madsdp_oracle as O
    threads = None
    try:
        from threadpoolctl import threadpool_info
        threads = max((d.get("num_threads", 0) for d in threadpool_info()
                       if d.get("user_api") == "blas"), default=None)
    except Exception:
        pass
    s = O.Solver(model)
    t0 = time.perf_counter()
    s.initialize()
    s.update_termination()
    t_init = time.perf_counter() - t0
    done, t_it = 0, 0.0
    while s.status == O.Status.REGULAR and (done == 0 or t_it < budget_s):
        t1 = time.perf_counter()
        s.iterate()
        if s.status == O.Status.REGULAR:
            s.update_termination()
        t_it += time.perf_counter() - t1
        done += 1
    per_iter = t_it / max(done, 1)
    est = t_init + per_iter * iters_gpu
    return {"value": 1.0 / est, "unit": "solves/s", "cores": threads or os.cpu_count(),
            "kind": "port",
            "sample": f"FP64 oracle (numpy/scipy OpenBLAS): initialization + {done} IPM iteration(s) "
                      f"of the same instance, {t_init + t_it:.1f} s; extrapolated to "
                      f"{iters_gpu} iterations ({per_iter:.2f} s/iteration)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="stream",
                    help="stream (default: the serving stream), maxcut3000, maxcut100, "
                         "dense250, blocks96 or maxcut<N> (one instance solved repeatedly)")
    ap.add_argument("--profile", default="trd", choices=["gemm", "trd", "gemm32", "none"],
                    help="kernel timed live for the roofline: trd = trd_panel_product_kernel "
                         "(the dominant kernel of the max-cut solve), gemm = the FP64 GEMM class, "
                         "gemm32 = the FP32 GEMM class")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed", "bf16x3"])
    ap.add_argument("--schur", default="cholesky", choices=["cholesky", "qr", "ldlt"],
                    help="single-instance workloads: Schur solver (DenseCholesky, QRSchur "
                         "or MadNLPSchur)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--rate", type=float, default=50.0, help="stream: Poisson arrival rate (1/s)")
    ap.add_argument("--slots", type=int, default=16, help="stream: concurrent solves")
    ap.add_argument("--mem-budget", type=int, default=128, help="stream: device GiB for admission")
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="stream: GPU_MAX_HW_QUEUES for the run (0: one per slot, 4..32)")
    ap.add_argument("--batched-precision", default="mixed", choices=["mixed", "fp64"],
                    help="stream: precision of the batched path")
    ap.add_argument("--compare-precision", type=int, default=1,
                    help="stream: also serve the timed requests batched in the other precision")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run an elementwise kernel over MIB MiB three times (the "
                         "known-size reference for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the benchmark runs on one GPU; multi-GPU runs are not supported "
              "(independent solves shard across GPUs with no exchange).", file=sys.stderr)
        sys.exit(2)

    if args.workload == "stream":
        # one hardware queue per concurrent solve (HIP's default is 4); set
        # before the HIP runtime starts
        prev = os.environ.get("GPU_MAX_HW_QUEUES")
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, max(4, args.hw_queues or args.slots)))
        print(f"GPU_MAX_HW_QUEUES {prev} -> {os.environ['GPU_MAX_HW_QUEUES']}; "
              f"{len(os.sched_getaffinity(0))} host cores", file=sys.stderr)
        rec = run_stream(args)
        print(json.dumps(rec), flush=True)
        # every request must end with the same status on both paths
        return 0 if rec["same_status"] else 1

    if args.workload == "blocks96":
        # the per-block stages run on 8 auxiliary streams: give them their own
        # hardware queues (HIP's default is 4), before the HIP runtime starts
        os.environ["GPU_MAX_HW_QUEUES"] = "16"

    import ctypes as C
    import numpy as np
    from madsdp_amd import DenseCholesky, MadNLPSchur, MadSDPSolver, QRSchur, Status

    if args.copy_calibration > 0:
        import torch
        a = torch.ones(args.copy_calibration << 17, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 2.0, out=b)  # an elementwise kernel: reads and writes MIB MiB
        torch.cuda.synchronize()
        del a, b
    model, config = build_model(args.workload)
    schur = {"cholesky": DenseCholesky, "qr": QRSchur, "ldlt": MadNLPSchur}[args.schur]()
    s = MadSDPSolver(model, device=0, precision=args.precision, schur=schur)
    config = dict(config, schur=type(schur).__name__)
    for _ in range(args.warmup):
        s.solve()
        print(f"warmup solve: status {s.status.name}, {s.iter} iterations", file=sys.stderr, flush=True)
    cls = {"gemm": 0, "trd": 1, "gemm32": 2, "none": -1}[args.profile]
    if cls >= 0:
        s.lib.msdp_solver_profile_begin(s._h, cls)
    lat = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        t1 = time.perf_counter()
        s.solve()
        lat.append(time.perf_counter() - t1)
        print(f"solve {k}: {lat[-1]:.3f} s, status {s.status.name}, {s.iter} iterations, "
              f"obj {s.primal_objective():.10e}", file=sys.stderr, flush=True)
    total = time.perf_counter() - t0
    roof = None
    if cls >= 0:
        ms, launches, work = C.c_double(), C.c_int64(), C.c_double()
        s.lib.msdp_solver_profile_end(s._h, C.byref(ms), C.byref(launches), C.byref(work))
        if launches.value > 0 and ms.value > 0:
            if cls in (0, 2):
                achieved = work.value / (ms.value * 1e-3) / 1e12
                traffic = None
                tf = os.path.join(ROOT, "profiles", f"gemm_traffic_{args.workload}.json")
                if cls == 0 and args.precision == "fp64" and os.path.exists(tf):
                    with open(tf) as fh:
                        traffic = json.load(fh)["traffic_bytes_per_launch"]
                peak = PEAKS["mfma"] if cls == 0 else PEAKS["mfma32"]
                roof = {"bound": "mfma", "achieved": achieved, "peak": peak,
                        "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
                        "kernel": "FP64 GEMM class (gemm128_f64_kernel, gemm_sk_f64_kernel, "
                                  "gemm_f64_kernel, gemm_splitk_kernel)"
                        if cls == 0
                        else "gemm128_f32_kernel (FP32 MFMA 16x16x4)",
                        "launches": launches.value, "avg_launch_us": 1e3 * ms.value / launches.value,
                        "algorithmic_per_launch": work.value / launches.value,
                        "kernel_share_of_solve": ms.value * 1e-3 / total}
            else:
                achieved = work.value / (ms.value * 1e-3) / 1e9
                traffic = None
                # HBM traffic per launch needs a rocprofv3 --pmc pass of this
                # kernel; none has been made of the current one (null)
                roof = {"bound": "hbm", "achieved": achieved, "peak": PEAKS["hbm"],
                        "unit": "GB/s", "frac": achieved / PEAKS["hbm"], "traffic": traffic,
                        "kernel": "trd_panel_product_kernel (p = A_trailing v of the blocked "
                                  "tridiagonalisation, one launch per column)",
                        "launches": launches.value, "avg_launch_us": 1e3 * ms.value / launches.value,
                        "algorithmic_per_launch": work.value / launches.value,
                        "kernel_share_of_solve": ms.value * 1e-3 / total}
    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": METRIC,
        "value": args.steps / total,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded numpy PCG64 graph)",
        "config": config,
        "p50_latency_ms": float(np.percentile(lat_ms, 50)),
        "p99_latency_ms": float(np.percentile(lat_ms, 99)),
        "precision": args.precision,
        "iterations": int(s.iter),
        "lowp_factorizations": int(s.info().lowp_factorizations),
        "refine_steps": int(s.info().refine_steps),
        "fp64_fallbacks": int(s.info().fp64_fallbacks),
        "status": s.status.name,
        "objective": s.primal_objective(),
        "roofline": roof,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(model, int(s.iter), args.cpu_seconds)
    ok = s.status == Status.SOLVE_SUCCEEDED
    print(json.dumps(out), flush=True)
    s.close()
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
