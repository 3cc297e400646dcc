"""Benchmark of the MadSDP interior-point hot path on one MI355X.

A step is one complete solve (initialize!, then iterate! until
update_termination! stops at the 1e-7 relative gap) of the plain FP64 GPU
path on the max-cut SDP over a seeded G(3000, 0.01) graph with +-1 weights:
one 3000 x 3000 PSD block, m = 3000 one-hot constraints, NT direction,
EigenStep(0.95), Mehrotra(3), DenseCholesky.  The problem data are uploaded
once and stay resident in HBM; the timed region holds only solves.

Output: one JSON line with solves/s (whole job), p50/p99 latency, the live
roofline of the dominant kernel (HIP events on the library's stream over the
timed region) and the CPU baseline (the FP64 oracle, a bounded sample of the
same solve timed on this host).
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


def run_stream(args):
    """Serving stream (c4): Poisson arrivals of independent instances, half
    max-cut with n ~ U{20..600}, half dense SDPs with n ~ U{16..200},
    m ~ U{50..3000}, n_s = n.  The same arrivals are served by the plain FP64
    path (one solve at a time) and by the continuous-batching loop (`--slots`
    concurrent solves) in `--batched-precision` and in the other precision.
    Returns the JSON record."""
    import numpy as np
    from madsdp_amd import MadSDPSolver, instances, serve

    stream = instances.serving_stream(count=args.count, seed=4, rate=args.rate)
    arrivals = [a - stream[0][0] for a, _, _ in stream]

    t = time.perf_counter()
    models = [instances.build_stream_instance(kind, params) for _, kind, params in stream]
    print(f"generated {len(models)} instances in {time.perf_counter() - t:.1f} s", file=sys.stderr,
          flush=True)

    def one(slots, precision):
        t = time.perf_counter()
        solvers = [MadSDPSolver(model, device=0, precision=precision) for model in models]
        prep = time.perf_counter() - t
        res, summary = serve(solvers, arrivals, slots=slots, mem_budget=args.mem_budget << 30)
        st = [int(s.status) for s in solvers]
        its = [int(s.iter) for s in solvers]
        objs = [s.primal_objective() for s in solvers]
        summary.update(prepare_s=prep, succeeded=int(sum(v == 1 for v in st)),
                       iterations_mean=float(np.mean(its)),
                       not_succeeded=[(k, stream[k][1], stream[k][2].get("n"), stream[k][2].get("m"),
                                       st[k], its[k]) for k in range(len(st)) if st[k] != 1])
        print(f"{precision} slots={slots}: {summary}", file=sys.stderr, flush=True)
        return summary, its, objs

    if args.stream_paths == "plain":
        plain, it_p, obj_p = one(1, "fp64")
        batched, it_b, obj_b = plain, it_p, obj_p
        other = None
    elif args.stream_paths == "batched":
        batched, it_b, obj_b = one(args.slots, args.batched_precision)
        plain, it_p, obj_p = batched, it_b, obj_b
        other = None
    else:
        plain, it_p, obj_p = one(1, "fp64")
        batched, it_b, obj_b = one(args.slots, args.batched_precision)
        alt = "mixed" if args.batched_precision == "fp64" else "fp64"
        other = one(args.slots, alt)[0]
        other["precision"] = alt
    dit = [b - a for a, b in zip(it_p, it_b)]
    dobj = max(abs(a - b) / max(1.0, abs(a)) for a, b in zip(obj_p, obj_b))
    return {
        "metric": METRIC,
        "value": batched["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.count,
        "warmup": 0,
        "ms_per_step": 1e3 / batched["solves_per_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded numpy PCG64 instances, Poisson arrivals)",
        "config": {"workload": f"serving stream: first {args.count} of the 4096-instance seed-4 stream "
                               f"(half max-cut n~U{{20..600}}, half dense n~U{{16..200}}, m~U{{50..3000}}), "
                               f"Poisson rate {args.rate}/s, tol 1e-7",
                   "slots": args.slots, "rate": args.rate, "count": args.count},
        "p50_latency_ms": 1e3 * batched["p50_latency_s"],
        "p99_latency_ms": 1e3 * batched["p99_latency_s"],
        "plain_fp64": plain,
        "batched": dict(batched, precision=args.batched_precision),
        "batched_other_precision": other,
        "speedup_vs_plain": batched["solves_per_s"] / plain["solves_per_s"],
        "max_extra_iterations": int(max(dit)),
        "max_rel_objective_diff": dobj,
        "roofline": None,
        "cpu_baseline": None,
    }


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
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="maxcut3000")
    ap.add_argument("--profile", default="trd", choices=["gemm", "trd", "gemm32", "none"],
                    help="kernel timed live for the roofline: trd = trd_panel_product_kernel "
                         "(the dominant kernel of the max-cut solve), gemm = the FP64 GEMM class, "
                         "gemm32 = the FP32 GEMM class")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--count", type=int, default=128, help="stream: number of requests")
    ap.add_argument("--rate", type=float, default=50.0, help="stream: Poisson arrival rate (1/s)")
    ap.add_argument("--slots", type=int, default=16, help="stream: concurrent solves")
    ap.add_argument("--mem-budget", type=int, default=128, help="stream: device GiB for admission")
    ap.add_argument("--stream-paths", default="both", choices=["both", "plain", "batched"])
    ap.add_argument("--batched-precision", default="fp64", choices=["mixed", "fp64"],
                    help="stream: precision of the batched path")
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
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, max(4, args.slots)))
        print(f"GPU_MAX_HW_QUEUES {prev} -> {os.environ['GPU_MAX_HW_QUEUES']}; "
              f"{len(os.sched_getaffinity(0))} host cores", file=sys.stderr)
        rec = run_stream(args)
        print(json.dumps(rec), flush=True)
        # every request must end with the same status on both paths
        same = rec["plain_fp64"]["not_succeeded"] == rec["batched"]["not_succeeded"]
        return 0 if same else 1

    if args.workload == "blocks96":
        # the per-block stages run on 8 auxiliary streams: give them their own
        # hardware queues (HIP's default is 4), before the HIP runtime starts
        os.environ["GPU_MAX_HW_QUEUES"] = "16"

    import ctypes as C
    import numpy as np
    from madsdp_amd import MadSDPSolver, Status

    if args.copy_calibration > 0:
        import torch
        a = torch.ones(args.copy_calibration << 17, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 2.0, out=b)  # an elementwise kernel: reads and writes MIB MiB
        torch.cuda.synchronize()
        del a, b
    model, config = build_model(args.workload)
    s = MadSDPSolver(model, device=0, precision=args.precision)
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
                        "kernel": "FP64 GEMM class (gemm128_f64_kernel, gemm_f64_kernel, "
                                  "gemm_splitk_kernel, vendor DGEMM for large plain products)"
                        if cls == 0
                        else "gemm128_f32_kernel (FP32 MFMA 16x16x4)",
                        "launches": launches.value, "avg_launch_us": 1e3 * ms.value / launches.value,
                        "algorithmic_per_launch": work.value / launches.value,
                        "kernel_share_of_solve": ms.value * 1e-3 / total}
            else:
                achieved = work.value / (ms.value * 1e-3) / 1e9
                traffic = None
                tf = os.path.join(ROOT, "profiles", f"trd_traffic_{args.workload}.json")
                if os.path.exists(tf):
                    with open(tf) as fh:
                        traffic = json.load(fh)["traffic_bytes_per_launch"]
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
