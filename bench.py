#!/usr/bin/env python3
"""Benchmark of the gfx950 sync-free SpTRSV on one MI355X.

Default workload: FP64 forward solve of the lower triangle (incl. diagonal) of
the 3D 7-point Laplacian on a 128^3 grid (n = 2,097,152, nnz = 8,339,456,
382 DAG levels), diag 6, off-diagonal -1, x ~ U(-1, 1) from
numpy default_rng(0), b = L x, one right-hand side.  A step is one solve of
that system on inputs already resident in HBM; W untimed solves, then K timed
solves between two stream synchronisations.

Prints one JSON line.  `roofline.achieved` is the algorithmic bytes of one
solve (CSC values + row indices + column pointers read once, b read once,
x written once) divided by the mean executor-kernel duration measured with HIP
events on the stream the kernel runs on; `roofline.traffic` is the HBM traffic
per launch measured with rocprofv3 counters, read from
profiles/traffic_<workload>.json when that file exists.  `cpu_baseline` times
the reference's own serial executor (compiled from the reference sources into
oracle/_ref) or, without it, the oracle restatement, single-threaded, on the
same matrix.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "benchmark-sptrsv_amd"))

import numpy as np  # noqa: E402

METRIC = "solves/sec + achieved HBM GB/s vs FP64 sync-free at rel err ≤1e-10"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name):
    from sptrsv_amd import matrices as mx
    if name == "lap3d":
        f = mx.laplacian3d_lower(128, 128, 128)
        return f, mx.x_uniform(f.n, seed=0), 382
    if name == "lap2d":
        f = mx.laplacian2d_lower(256, 256)
        return f, np.ones(f.n), 511
    if name == "powerlaw":
        f = mx.power_law_lower(1_000_000, seed=0)
        return f, mx.x_uniform(f.n, seed=1), None
    if name == "bidiag":
        f = mx.bidiagonal(1_000_000)
        return f, mx.x_uniform(f.n, seed=0), 1_000_000
    if name == "chain_far":
        f = mx.chain_far(1_000_000)
        return f, mx.x_uniform(f.n, seed=0), 1_000_000
    if name == "batch":
        f = mx.laplacian2d_batch(512, seed=0)
        return f, mx.x_uniform(f.n, seed=0), None
    if name == "stencil27":
        # level(x, y, z) = x + 2y + 4z
        f = mx.stencil27_lower(100, 100, 60)
        return f, mx.x_uniform(f.n, seed=0), 99 + 2 * 99 + 4 * 59 + 1
    if name == "lap3d_share":
        # one GPU's share of the 3D Laplacian under an 8-way contiguous column
        # split: the diagonal block of 16 planes.  Its right-hand side is
        # L_block x_block, i.e. the contributions of the earlier blocks are
        # already subtracted, as if the remote producers had finished.
        f = mx.laplacian3d_lower(128, 128, 16)
        return f, mx.x_uniform(f.n, seed=0), 127 + 127 + 15 + 1
    raise SystemExit(f"unknown workload {name}")


WORKLOAD_DESC = {
    "lap3d": "lower triangle of 3D 7-point Laplacian 128^3 (n=2,097,152), diag 6, off -1",
    "lap2d": "lower triangle of 2D 5-point Laplacian 256^2 (n=65,536), diag 4, off -1",
    "powerlaw": "random power-law lower triangular n=1,000,000, Zipf(2) cap 1000, window 50,000",
    "bidiag": "bidiagonal chain n=1,000,000, diag 2, sub -1",
    "batch": "512 uneven 2D 5-point Laplacian lower factors (sides log-U[32,256]) in one launch",
    "chain_far": "chain n=1,000,000 (diag 2, sub -1) with a second input 9000 columns back "
                 "(-0.5): no pure links, so it is passed one 64-column batch at a time; "
                 "time / 15,625 = the cost of one batch-to-batch step",
    "stencil27": "lower triangle of 3D 27-point stencil 100x100x60 (n=600,000), diag 26, off -1",
    "lap3d_share": "one GPU's share of the 128^3 3D Laplacian factor under an 8-way "
                   "contiguous column split (diagonal block of 16 planes, n=262,144; "
                   "cross-block contributions pre-accumulated into b)",
}


def per_system_launches(f, b, repeats=3):
    """The batch's systems solved the plain way: one plan and one launch per
    system, back to back on one stream (default executor configuration, since
    tuning 512 plans would dominate the run).  Returns the mean wall time of
    the 512 launches in ms, the solution arrays and each system's span."""
    import sptrsv_amd as S
    sp = f.sys_ptr.astype(np.int64)
    plans = []
    for s in range(len(sp) - 1):
        lo, hi = int(sp[s]), int(sp[s + 1])
        e0, e1 = int(f.colPtr[lo]), int(f.colPtr[hi])
        cp = (f.colPtr[lo:hi + 1] - e0).astype(np.int32)
        ri = (f.rowIdx[e0:e1] - lo).astype(np.int32)
        plan = S.Plan.from_host(cp, ri, f.val[e0:e1])
        db = S.DeviceArray.from_host(np.ascontiguousarray(b[lo:hi]))
        dx = S.DeviceArray((hi - lo,), np.float64)
        plans.append((plan, db, dx, lo, hi))
    stream = S.Stream()
    for plan, db, dx, _, _ in plans:  # warm-up: analysis caches, slots
        plan.solve(db, dx, stream=stream)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(repeats):
        for plan, db, dx, _, _ in plans:
            plan.solve(db, dx, stream=stream)
    stream.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / repeats
    for plan, _, _, _, _ in plans:
        assert plan.check(stream) == 0, "per-system solve timed out"
    return ms, [p[2] for p in plans], [(p[3], p[4]) for p in plans]


def host_cpu():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(f, b, budget_s, substitution=0, rhs=1):
    """Single-threaded serial sync-free executor on the same matrix."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import sptrsv_oracle as orc
    kind = "reference" if orc.ref(np.float64) is not None else "port"
    fn = orc.ref_solve if kind == "reference" else orc.solve
    indeg = orc.analyser(f.rowIdx, f.n)  # analysis is outside the timed executor
    colPtr, rowIdx = f.colPtr, f.rowIdx
    reps, t0 = 0, time.perf_counter()
    times = []
    while time.perf_counter() - t0 < budget_s or reps < 2:
        t = time.perf_counter()
        x_cpu = fn(colPtr, rowIdx, f.val, b, substitution=substitution, rhs=rhs,
                   in_degree=indeg)
        times.append(time.perf_counter() - t)
        reps += 1
    med = float(np.median(times))
    return {"value": round(rhs / med, 3), "unit": "solves/s", "cores": 1, "kind": kind,
            "sample": f"{reps} serial FP64 executor runs on the same matrix and b "
                      f"({rhs} right-hand side{'s' if rhs > 1 else ''}; median "
                      f"{med * 1e3:.2f} ms each, analysis excluded) on 1 core of "
                      f"{host_cpu()}, {os.cpu_count()} host cores visible"}, x_cpu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="lap3d", choices=sorted(WORKLOAD_DESC))
    ap.add_argument("--substitution", default="forward", choices=["forward", "backward"],
                    help="backward solves U = L^T of the workload (diagonal last)")
    ap.add_argument("--mtx", default=None,
                    help="Matrix Market file: its unit-diagonal triangle (main.cu:152-187)")
    ap.add_argument("--mtx-values", default="file", choices=["file", "random"],
                    help="off-diagonal values of an --mtx factor: the file's, or the "
                         "reference's random 1..10")
    ap.add_argument("--mtx-diag", default="file", choices=["file", "unit"],
                    help="diagonal of an --mtx factor: the file's, or the reference's 1")
    ap.add_argument("--csc", default=None, help="CSC text factor (main.cu:247-286)")
    ap.add_argument("--rhs", type=int, default=1)
    ap.add_argument("--form", default="plain", choices=["plain", "ir"])
    ap.add_argument("--ir-iters", type=int, default=1,
                    help="FP64 refinement steps after the first FP32 solve (one reaches "
                         "<= 1e-10 on every benchmark factor)")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--config", default=None,
                    help='executor configuration as JSON (e.g. \'{"engine": 2, '
                         '"lanes_per_column": 64, "rhs_lanes": 64, "rhs_per_lane": 2}\'); '
                         "skips the autotuner")
    ap.add_argument("--graph", default=None, action=argparse.BooleanOptionalAction,
                    help="time replays of a hipGraph of the solve (default: on for "
                         "the batch workload, whose configuration names one graph launch)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-per-system", action="store_true",
                    help="batch workload: skip the 512 separate per-system launches")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run 3 device-to-device copies of this many MiB "
                         "(the known-size kernel that scales HBM counters)")
    args = ap.parse_args()

    if args.calibrate_copy_mib:
        import sptrsv_amd as S
        nbytes = args.calibrate_copy_mib << 20
        src = S.DeviceArray((nbytes,), np.uint8)
        dst = S.DeviceArray((nbytes,), np.uint8)
        src.zero()
        for _ in range(3):
            S.check(S.lib().sptrsv_memcpy_dtod(dst.ptr, src.ptr, nbytes), "dtod")
        S.lib().sptrsv_synchronize()
        print(json.dumps({"copy_bytes": nbytes, "copies": 3}), flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the solver has no multi-GPU path "
              "(a single triangular system is one GPU's work; independent "
              "systems or right-hand-side tiles would shard with no exchange)",
              file=sys.stderr)
        sys.exit(2)

    import sptrsv_amd as S

    if S.device_count() < 1:
        raise SystemExit("no HIP device visible")
    log(f"device: {S.device_name()}")
    t = time.perf_counter()
    bwd = args.substitution == "backward"
    if args.mtx or args.csc:
        from sptrsv_amd import matrices as mx, mtx
        f = (mtx.triangle_from_mtx(args.mtx, substitution=int(bwd), values=args.mtx_values,
                                   diag=args.mtx_diag) if args.mtx
             else mtx.read_csc_text(args.csc))
        x_true, levels = mx.x_uniform(f.n, seed=0), None
        args.workload = "file"
        WORKLOAD_DESC["file"] = f"{'Matrix Market' if args.mtx else 'CSC text'} factor " \
                                f"{os.path.basename(args.mtx or args.csc)} (n={f.n}, nnz={f.nnz})"
    else:
        f, x_true, levels = make_workload(args.workload)
        if bwd:
            from sptrsv_amd import matrices as mx
            f = mx.upper_from_lower(f)
    if args.graph is None:
        args.graph = args.workload == "batch"
    rhs = args.rhs
    if rhs > 1:
        from sptrsv_amd import matrices as mx
        x_true = mx.x_normal(f.n, rhs, seed=1)
    b = f.spmv(x_true)
    log(f"workload {args.workload}: n={f.n} nnz={f.nnz} rhs={rhs} "
        f"(generated in {time.perf_counter() - t:.1f} s)")

    flags = S.PLAN_VALIDATE | (S.PLAN_MIXED if args.form == "ir" else 0)
    mode = 1 if args.form == "ir" else 0
    dP = S.DeviceArray.from_host(f.colPtr)
    dI = S.DeviceArray.from_host(f.rowIdx)
    dV = S.DeviceArray.from_host(f.val)

    def new_plan():
        return S.Plan(dP, dI, dV, f.n, f.nnz, substitution=int(bwd), flags=flags,
                      sys_ptr=f.sys_ptr)

    db = S.DeviceArray.from_host(b)
    dx = S.DeviceArray(b.shape, np.float64)
    stream = S.Stream()
    tuned = None
    if args.config:
        tuned = json.loads(args.config)
    elif not args.no_tune:
        # the tuner's result reaches the timed plan through the tune cache
        # (a cache hit stores it for the solve type of `mode`)
        if not os.environ.get("SPTRSV_TUNE_CACHE"):
            import tempfile
            os.environ["SPTRSV_TUNE_CACHE"] = os.path.join(
                tempfile.mkdtemp(prefix="sptrsv_bench_"), "tune.cache")
        tplan = new_plan()
        tuned, tms = tplan.autotune(rhs=rhs, mode=mode)
        log(f"autotuned config {tuned} ({tms:.3f} ms)")
        tplan.destroy()
    # the timed plan: created afresh, so its analysis times are those of the
    # configuration that is timed (plan creation + per-configuration analysis)
    plan = new_plan()
    # DAG levels first (the tune cache's key holds the depth; the row engine
    # claims rows in level order), timed apart from the executor analysis
    t_lv = time.perf_counter()
    dev_levels = plan.levels(stream)[0]
    levels_ms = (time.perf_counter() - t_lv) * 1e3
    if args.config:
        plan.set_config(rhs, **tuned)
    elif tuned:
        plan.autotune(rhs=rhs, mode=mode)  # cache hit: no timing runs
    plan.prepare(rhs, mode, stream=stream)
    graph = None
    if args.graph:
        graph = plan.graph(db, dx, rhs=rhs, mode=mode, iters=args.ir_iters, stream=stream)

    def step():
        if graph is not None:
            graph.launch(stream)
        elif args.form == "ir":
            plan.solve_ir(db, dx, rhs=rhs, iters=args.ir_iters, stream=stream)
        else:
            plan.solve(db, dx, rhs=rhs, stream=stream)

    for _ in range(args.warmup):
        step()
    stream.synchronize()
    assert plan.check(stream) == 0, "solve timed out"

    # timed region: K steps bracketed by stream synchronisation, HIP events
    # around each launch on the solve stream
    evs = [(S.Event(), S.Event()) for _ in range(args.steps)]
    stream.synchronize()
    t0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record(stream)
        step()
        e1.record(stream)
    stream.synchronize()
    wall = time.perf_counter() - t0
    st = plan.check(stream)
    assert st == 0, f"solve status {st}"
    per = np.array([e0.elapsed_ms(e1) for e0, e1 in evs])
    ms_step = wall * 1e3 / args.steps
    x = dx.to_host()

    # accuracy against the known solution (b was built as L x_true)
    d = np.abs(x - x_true.reshape(x.shape))
    err_true = {"l1": float(d.sum() / np.abs(x_true).sum()),
                "linf": float(d.max() / np.abs(x_true).max())}

    solve_bytes = f.algorithmic_bytes(rhs=rhs, value_bytes=8)
    if args.form == "ir":
        it = args.ir_iters
        s32 = f.algorithmic_bytes(rhs=rhs, value_bytes=4)
        resid = f.nnz * 12 + (f.n + 1) * 4 + 3 * f.n * rhs * 8 + f.n * rhs * 4
        solve_bytes = (it + 1) * s32 + it * resid + f.n * rhs * (8 + 4) + f.n * rhs * 12
    kern_ms = float(np.mean(per))
    achieved = solve_bytes / (kern_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_{args.form}_rhs{rhs}"
                         + ("_bwd" if bwd else "") + ".json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    if levels is None:
        # DAG depth of the factor (findlevel.h's definition), device analysis
        levels = dev_levels
    elif levels != dev_levels:
        raise SystemExit(f"device level analysis gives depth {dev_levels}, expected {levels}")
    pinfo = plan.info()
    out = {
        "metric": METRIC,
        "value": round(rhs * args.steps / wall, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.form == "plain" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": WORKLOAD_DESC[args.workload]
                        + (", upper triangle U = L^T, backward substitution" if bwd else "")
                        + f", nrhs={rhs}, "
                        + ("FP64 sync-free" if args.form == "plain"
                           else f"FP32 solve + {args.ir_iters} FP64 refinement steps"),
            "n": f.n, "nnz": f.nnz, "nrhs": rhs, "dag_levels": levels,
            "executor": tuned or plan.get_config(rhs),
            "timed_as": "hipGraph replays" if graph is not None else "eager launches",
            "analysis_ms": round(pinfo["analysis_ms"] + pinfo["prepare_ms"], 3),
            "analysis_split_ms": {"plan_create": round(pinfo["analysis_ms"], 3),
                                  "dag_levels": round(levels_ms, 3),
                                  "per_configuration": round(max(0.0, pinfo["prepare_ms"]
                                                                 - levels_ms), 3)},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "algorithmic_bytes": solve_bytes,
            "kernel_ms_mean": round(kern_ms, 5),
            "kernel_ms_min": round(float(per.min()), 5),
        },
        "rel_err_vs_x_true": err_true,
    }
    if (args.workload == "batch" and args.form == "plain" and rhs == 1 and not bwd
            and not args.no_per_system):
        # the same 512 systems as 512 separate plain launches
        ps_ms, dxs, spans = per_system_launches(f, b.reshape(-1))
        worst = 0.0
        for dx, (lo, hi) in zip(dxs, spans):
            xs, xt = dx.to_host(), x_true.reshape(-1)[lo:hi]
            worst = max(worst, float(np.abs(xs - xt).max() / np.abs(xt).max()))
        out["per_system_launches"] = {
            "ms": round(ps_ms, 4), "launches": len(dxs),
            "batched_speedup": round(ps_ms / ms_step, 2),
            "rel_err_vs_x_true_linf": worst,
            "note": "same systems, one plan and one launch each, default executor"}
    if not args.no_cpu:
        cpu, x_cpu = cpu_baseline(f, b.reshape(-1), args.cpu_seconds, int(bwd), rhs)
        out["cpu_baseline"] = cpu
        # the CPU run is also the checker: same matrix, same b
        d = np.abs(x.reshape(-1) - x_cpu)
        l1 = float(d.sum() / np.abs(x_cpu).sum())
        linf = float(d.max() / np.abs(x_cpu).max())
        tol = 1e-12 if args.form == "plain" else 1e-10
        out["rel_err_vs_cpu_serial"] = {"l1": l1, "linf": linf, "tol": tol,
                                        "pass": bool(l1 <= tol and linf <= tol)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
