"""Benchmark: orbit-steps/s of the batched DOP853 integrator in the triaxial Dehnen Multipole
potential (lmax=8) on one MI355X - the MLP surrogate path against the FP64 Multipole path,
both held to a per-orbit energy error of at most 1e-6.

Workload ("batching at scale"): 1M orbits with initial conditions drawn from the triaxial
Dehnen density (seed 42), integration times log-uniform between 1 and 300 circular periods,
accuracy 1e-8, final state plus max relative energy error per orbit, FP64 integrator.
One bench step = one integration of the whole batch; inputs and outputs stay resident in HBM.

Default mode runs both forms on the same initial conditions:
  FP64 Multipole: one persistent-kernel launch per step (its time and HIP-event kernel time);
  MLP surrogate: the network is first trained on the GPU (not timed), then each step is one
    surrogate-orbit call - a pilot of every 64th orbit on the surrogate (200 steps), then either
    the surrogate kernel with the FP64 Multipole as energy judge and the FP64 kernel for the
    orbits that exceed dE/E = 1e-6 or leave the training radii, or, when the pilot shows that
    the surrogate cannot hold the batch to the tolerance, the FP64 kernel for the whole batch.
`value` = accepted orbit-steps / s of the surrogate path (the steps of the final results over
the whole call); the FP64 path's figure and the ratio are reported next to it.  `--mode
fp64` runs the FP64 path alone (its figure is then `value`).

Launch counts: every leg runs one warm-up launch (min(--warmup, 1)); the headline leg times
exactly --steps launches, the FP64 leg of the default mode at most 2 (it is reported beside the
headline, and one launch of 1M orbits takes seconds).  The counts are recorded in `config`.

Prints one JSON line (rank 0).  `--gpus N` with N > 1 is not supported (single-GPU build).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "orbit-steps/s at 1M orbits: MLP surrogate vs FP64 Multipole at ΔE/E≤1e-6"

# Algorithmic FP64 operations per unit of work: the reference algorithm's own counts, from the
# oracle (its C++ restatement) compiled with an operation-counting double (+ - * / sqrt and each
# libm call count 1; oracle/flopcount.py on 10^4 orbits of the 1-300-period workload and 2x10^4
# density points of the triaxial Dehnen lmax=8 expansion):
#   per DOP853 step attempt, everything included (12 evaluations, stage sums, error norm and
#     step control, energy bookkeeping, the orbits' starts amortised)           27,795
#   per DPRKN8 step attempt (9 evaluations)                                      20,508
#   per potential + gradient evaluation at density points                        2,194
#   per potential + gradient + Hessian evaluation                                3,044
FLOPS_PER_ATTEMPT = {"dop853": 27795, "dprkn8": 20508}
# The GPU kernels execute less arithmetic for the same work (per-cell polynomial tables instead of
# per-point Hermite weights): FP64 operations per attempt from the hardware counters
# SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F64 on the 1M-orbit launch (an FMA counts 2; every lane of each
# instruction, idle lanes included); reported beside the roofline figure as executed_frac.
GPU_FLOPS_PER_ATTEMPT = {"dop853": 16290, "dprkn8": 12342}
METHOD_FLAG = {"dop853": 0, "dprkn8": 0x200}
# record fetch modes of the Multipole evaluation (agx_set_fetch)
FETCH_MODES = {"scalar": 0, "vec2": 1, "coop": 2, "cell": 3}
# launches timed for a leg that is reported beside the headline figure (not the headline)
COMPARISON_STEPS = 4
PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 (vector and matrix), AMD published figure
# HBM bytes per launch of the orbit path on this workload (the probe and the main persistent
# kernel), from rocprofv3 FETCH_SIZE and WRITE_SIZE runs scaled by a 1 GiB copy
# (profiles/orbitKernel_1M_uneven_triaxial_dehnen_lmax8.md): FETCH 467,243 KiB x 2.000,
# WRITE 354,277 KiB x 1.000
ORBIT_TRAFFIC_BYTES = 467243 * 1024 * 2.0 + 354277 * 1024


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--orbits", type=int, default=1_000_000)
    ap.add_argument("--precision", choices=["fp64", "fp32tables"], default="fp64")
    ap.add_argument("--cpu-orbits", type=int, default=20_000,
                    help="size of the CPU-baseline sample (0 disables)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--mode", choices=["orbit", "fp64", "eval", "hess", "mlp", "copy", "orbitobj", "construct",
                                       "target", "var"],
                    default="orbit",
                    help="orbit: surrogate and FP64 paths (default); fp64: the FP64 path alone; "
                         "eval: potential + force at 1M points drawn from the density; hess: the same "
                         "with the Hessian; mlp: the "
                         "surrogate's fused MFMA evaluation at 1M points; copy: one 1 GiB "
                         "device-to-device copy, the reference for HBM byte counters; orbitobj: "
                         "orbit(dtype=object) - every step recorded, splines built and evaluated; "
                         "construct: Multipole from a density (Poisson solver) on the device vs the host; "
                         "target: orbits feeding a DensitySphHarm target; var: orbits with the six "
                         "deviation vectors (der=True) and with Lyapunov exponents")
    ap.add_argument("--method", choices=["dop853", "dprkn8"], default="dop853",
                    help="integrator of the FP64 Multipole path (the surrogate path uses DOP853)")
    ap.add_argument("--groups", default=None,
                    help="group mode of the FP64 DOP853 kernel, STRIDE[,ORBITS] (STRIDE 0: off): one "
                         "wavefront in STRIDE integrates the first ORBITS orbits of the order with 8 lanes "
                         "per orbit; default: the library's automatic choice (off at 1M orbits)")
    ap.add_argument("--fetch", choices=list(FETCH_MODES), default=None,
                    help="record gather of the FP64 orbit kernel (default: the library's, vec2)")
    ap.add_argument("--periods", type=float, default=0,
                    help="FP64 path only: integrate every orbit for this many circular periods instead "
                         "of the log-uniform 1-300 (a uniform batch, for profiling the bulk)")
    ap.add_argument("--graph", action="store_true",
                    help="fp64 mode: time replays of the batch captured as a HIP graph (agx_orbit_graph_*; "
                         "the work order and probe are computed once, when the graph is created)")
    ap.add_argument("--train-steps", type=int, default=6000)
    ap.add_argument("--train-samples", type=float, default=20e6)
    ap.add_argument("--max-rel-de", type=float, default=1e-6)
    return ap.parse_args()


# algorithmic FP64 operations of one potential+gradient evaluation (eval mode) and of one
# potential+gradient+Hessian evaluation (hess mode): the oracle's counts (see FLOPS_PER_ATTEMPT)
FLOPS_PER_POINT = 2194
FLOPS_PER_POINT_HESS = 3044


def bench_eval(args, hess=False):
    import numpy as np
    import torch
    import agama_amd
    from agama_amd import lib, workloads
    dev = torch.device("cuda:0")
    coef = os.path.join(ROOT, "tests", "golden", "dehnen_triax_lmax8.coef")
    pot = agama_amd.Potential(file=coef)
    n = 1_000_000
    pts_np = workloads.density_points(n, seed=7)
    pts = torch.from_numpy(pts_np).to(dev)
    phi = torch.empty(n, dtype=torch.float64, device=dev)
    grad = torch.empty((n, 3), dtype=torch.float64, device=dev)
    hes = torch.empty((n, 6), dtype=torch.float64, device=dev) if hess else None
    prec = agama_amd.PRECISION_FP64 if args.precision == "fp64" else agama_amd.PRECISION_FP32
    if hess and prec != agama_amd.PRECISION_FP64:
        raise SystemExit("bench.py: the Hessian path is FP64 only")
    stream = torch.cuda.Stream()
    reps = max(args.steps, 1) * 20
    if args.fetch:
        assert lib.agx_set_fetch(FETCH_MODES[args.fetch], 1) == 0

    def launch():
        if hess:
            rc = lib.agx_eval_hess_many_device(pot.handle, n, pts.data_ptr(), 0.0, phi.data_ptr(),
                                               grad.data_ptr(), hes.data_ptr(), prec, stream.cuda_stream)
        else:
            rc = lib.agx_eval_many_device(pot.handle, n, pts.data_ptr(), 0.0, phi.data_ptr(), grad.data_ptr(),
                                          prec, stream.cuda_stream)
        if rc:
            raise RuntimeError(agama_amd._native.last_error())
    with torch.cuda.stream(stream):
        for _ in range(max(args.warmup, 1) * 5):
            launch()
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(reps):
            launch()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    per = ev0.elapsed_time(ev1) / 1e3 / reps
    achieved = n * (FLOPS_PER_POINT_HESS if hess else FLOPS_PER_POINT) / per / 1e12
    cpu = None
    if args.cpu_orbits > 0:
        from oracle import pyoracle
        threads = args.cpu_threads or os.cpu_count()
        fast = pyoracle.OraclePotential(coef, library=pyoracle.fast())
        t0 = time.perf_counter()
        for _ in range(3):
            (fast.eval_hess if hess else fast.eval)(pts_np, nthreads=threads)
        dt = (time.perf_counter() - t0) / 3
        cpu = {"value": n / dt, "unit": "evaluations/s", "cores": threads, "kind": "port",
               "sample": f"the same 1M points, C++ OpenMP restatement (-O2), {threads} threads"}
    what = "potential+force+Hessian" if hess else "potential+force"
    out = {"metric": f"{what} evaluations/s at 1M points (triaxial Dehnen Multipole lmax=8)",
           "value": n / per, "unit": "evaluations/s", "n_gpus": 1, "steps": reps, "warmup": args.warmup,
           "ms_per_step": wall / reps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f64 (fp32 spline tables)",
           "data": "synthetic",
           "config": {"workload": ("Phi, grad Phi and the Hessian" if hess else "Phi and grad Phi") +
                                  " at 1M points drawn from the triaxial Dehnen density"},
           "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                        "kernel": "evalHessKernel" if hess else "evalKernel<double>", "kernel_ms": per * 1e3},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build targets one GPU; --gpus N with N > 1 is not supported "
              "(orbits are independent and would shard by orbit index)", file=sys.stderr)
        sys.exit(2)
    if args.mode in ("eval", "hess"):
        return bench_eval(args, hess=args.mode == "hess")
    if args.mode == "copy":
        import torch
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"mode": "copy", "bytes_read": a.numel() * 8, "bytes_written": b.numel() * 8}))
        return
    if args.mode == "mlp":
        return bench_mlp(args)
    if args.mode == "orbitobj":
        return bench_orbitobj(args)
    if args.mode == "construct":
        return bench_construct(args)
    if args.mode == "target":
        return bench_target(args)
    if args.mode == "var":
        return bench_var(args)
    import numpy as np
    import torch
    import agama_amd
    from agama_amd import lib, workloads

    dev = torch.device("cuda:0")
    coef = os.path.join(ROOT, "tests", "golden", "dehnen_triax_lmax8.coef")
    pot = agama_amd.Potential(file=coef)
    n = args.orbits
    if args.fetch:
        assert lib.agx_set_fetch(1, FETCH_MODES[args.fetch]) == 0
    if args.groups:
        g = [int(v) for v in args.groups.split(",")]
        assert lib.agx_set_orbit_groups(g[0], g[1] if len(g) > 1 else -1, 0) == 0
    if args.periods > 0 and args.mode != "fp64":
        raise SystemExit("bench.py: --periods applies to --mode fp64 only")
    ic, T = workloads.triaxial_dehnen_orbits(n, lambda p: pot.potential(p), pot.coefs(), seed=42,
                                             uneven=args.periods <= 0, periods=args.periods)
    prec = agama_amd.PRECISION_FP64 if args.precision == "fp64" else agama_amd.PRECISION_FP32

    d_ic = torch.from_numpy(ic).to(dev)
    d_T = torch.from_numpy(T).to(dev)
    d_sz = torch.ones(n, dtype=torch.int32, device=dev)
    d_off = torch.arange(n, dtype=torch.int64, device=dev)
    d_time = torch.empty(n, dtype=torch.float64, device=dev)
    d_fin = torch.empty((n, 6), dtype=torch.float64, device=dev)
    d_ns = torch.empty(n, dtype=torch.int64, device=dev)
    d_na = torch.empty(n, dtype=torch.int64, device=dev)
    d_ne = torch.empty(n, dtype=torch.int64, device=dev)
    d_de = torch.empty(n, dtype=torch.float64, device=dev)
    d_st = torch.empty(n, dtype=torch.int32, device=dev)
    stream = torch.cuda.Stream()

    def launch_fp64():
        rc = lib.agx_orbit_batch_device(pot.handle, n, d_ic.data_ptr(), None, d_T.data_ptr(), d_sz.data_ptr(),
                                        d_off.data_ptr(), None, 1e-8, 0.0, int(1e8), d_time.data_ptr(), None,
                                        d_fin.data_ptr(), d_ns.data_ptr(), d_na.data_ptr(), d_ne.data_ptr(),
                                        d_de.data_ptr(), d_st.data_ptr(), prec | METHOD_FLAG[args.method],
                                        stream.cuda_stream)
        if rc:
            raise RuntimeError(agama_amd._native.last_error())

    graph = None
    if args.graph:
        if args.mode != "fp64":
            raise SystemExit("--graph applies to --mode fp64")
        graph = lib.agx_orbit_graph_create(pot.handle, n, d_ic.data_ptr(), None, d_T.data_ptr(), d_sz.data_ptr(),
                                           d_off.data_ptr(), 1e-8, 0.0, int(1e8), d_time.data_ptr(), None,
                                           d_fin.data_ptr(), d_ns.data_ptr(), d_na.data_ptr(), d_ne.data_ptr(),
                                           d_de.data_ptr(), d_st.data_ptr(), prec | METHOD_FLAG[args.method])
        if not graph:
            raise RuntimeError(agama_amd._native.last_error())

        def launch_fp64():   # noqa: F811 (the graph replaces the eager launch)
            if lib.agx_orbit_graph_launch(graph, stream.cuda_stream):
                raise RuntimeError(agama_amd._native.last_error())

    # One bench step is one integration of the whole batch: a persistent launch of several
    # seconds.  After the first launch of a kernel nothing is left to warm up (code object
    # loaded, tables in L2 after the first evaluations), so each leg runs at most one warm-up
    # launch whatever --warmup asks.  The headline leg times all --steps launches; a leg that
    # is only reported beside it (the FP64 path in the default mode) times at most
    # COMPARISON_STEPS.  Each launch is synchronised and reported as it ends (the cost of the
    # synchronisation is microseconds against seconds of integration).
    warmup_used = min(args.warmup, 1)
    headline_fp64 = args.mode == "fp64" or args.precision != "fp64" or args.method != "dop853"
    fp64_steps = args.steps if headline_fp64 else min(args.steps, COMPARISON_STEPS)

    def timed(launch, label, steps):
        with torch.cuda.stream(stream):
            for _ in range(warmup_used):
                t0 = time.perf_counter()
                launch()
                torch.cuda.synchronize()
                print(f"{label}: warm-up launch done in {time.perf_counter() - t0:.2f} s", file=sys.stderr,
                      flush=True)
            torch.cuda.synchronize()
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            t0 = time.perf_counter()
            ev0.record(stream)
            for i in range(steps):
                launch()
                torch.cuda.synchronize()
                print(f"{label}: step {i + 1}/{steps} done at {time.perf_counter() - t0:.2f} s", file=sys.stderr,
                      flush=True)
            ev1.record(stream)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
        return wall, ev0.elapsed_time(ev1) / 1e3

    wall, kernel_s = timed(launch_fp64, f"FP64 Multipole ({args.method})", fp64_steps)
    ns = d_ns.cpu().numpy()
    na = d_na.cpu().numpy()
    de = d_de.cpu().numpy()
    st = d_st.cpu().numpy()
    if graph:
        lib.agx_orbit_graph_destroy(graph)
    steps_per_launch = int(ns.sum())
    attempts = int(na.sum())
    flops = attempts * FLOPS_PER_ATTEMPT[args.method]
    gpu_flops = attempts * GPU_FLOPS_PER_ATTEMPT[args.method]
    per_launch = kernel_s / fp64_steps
    fp64_value = steps_per_launch * fp64_steps / wall
    achieved = flops / per_launch / 1e12
    print(f"steps per orbit: median {np.median(ns):.0f}, 99.9% {np.quantile(ns, 0.999):.0f}, max {ns.max()}",
          file=sys.stderr, flush=True)
    print(f"orbits {n}: {steps_per_launch} accepted steps, {attempts} attempts per launch; "
          f"kernel {per_launch*1e3:.1f} ms/launch, wall {wall:.2f} s; status counts {np.bincount(st)}; "
          f"median/max dE/E {np.median(de):.2e}/{np.max(de):.2e}", file=sys.stderr, flush=True)

    surrogate = None
    value = fp64_value
    if args.mode == "orbit" and args.precision == "fp64" and args.method == "dop853":
        t0 = time.perf_counter()
        sur = agama_amd.Surrogate.train(pot, samples=args.train_samples, steps=args.train_steps)
        train_s = time.perf_counter() - t0
        hist = sur.loss_history()
        print(f"surrogate trained in {train_s:.1f} s: loss {hist[:20].mean():.3e} -> {hist[-50:].mean():.3e}",
              file=sys.stderr, flush=True)
        nfb = np.zeros(1, dtype=np.int64)

        def launch_sur():
            rc = lib.agx_surrogate_orbit_batch_device(
                sur._h, n, d_ic.data_ptr(), None, d_T.data_ptr(), d_sz.data_ptr(), d_off.data_ptr(), 1e-8, 0.0,
                int(1e8), float(args.max_rel_de), d_time.data_ptr(), None, d_fin.data_ptr(), d_ns.data_ptr(),
                d_na.data_ptr(), d_ne.data_ptr(), d_de.data_ptr(), d_st.data_ptr(), nfb.ctypes.data,
                stream.cuda_stream)
            if rc:
                raise RuntimeError(agama_amd._native.last_error())

        swall, _ = timed(launch_sur, "MLP surrogate", args.steps)
        pilot = np.zeros(2, dtype=np.int64)
        lib.agx_surrogate_last_pilot(pilot[:1].ctypes.data, pilot[1:].ctypes.data)
        sns = d_ns.cpu().numpy()
        sde = d_de.cpu().numpy()
        sst = d_st.cpu().numpy()
        value = float(sns.sum()) * args.steps / swall
        nf = int(nfb[0])
        # surrogate accuracy at points drawn from the density (judged by the FP64 Multipole)
        pts = workloads.density_points(200_000, seed=9)
        r = np.linalg.norm(pts, axis=1)
        pts = pts[(r > 0.01) & (r < 100)]
        ophi, og = pot._eval(pts, True, True)
        sphi, sg = sur._eval(pts, True, True)
        ferr = np.linalg.norm(sg - og, axis=1) / np.linalg.norm(og, axis=1)
        perr = np.abs(sphi / ophi - 1)
        surrogate = {
            "value": value, "fp64_multipole_value": fp64_value, "surrogate_over_fp64": value / fp64_value,
            "max_rel_de": args.max_rel_de, "fallback_orbits": nf, "fallback_fraction": nf / n,
            "pilot": {"orbits": int(pilot[0]), "fell_back": int(pilot[1]),
                      "batch_routed_to_fp64": bool(2 * pilot[1] >= pilot[0] > 0)},
            "orbits_kept_on_surrogate": n - nf,
            "max_dE_over_E": float(np.max(sde)), "status_counts": np.bincount(sst & 7).tolist(),
            "force_err_median": float(np.median(ferr)), "force_err_p99": float(np.quantile(ferr, 0.99)),
            "potential_err_median": float(np.median(perr)),
            "training": {"samples": int(args.train_samples), "steps": args.train_steps, "batch": 16384,
                         "seconds": train_s, "final_loss": float(hist[-50:].mean())},
            "wall_s_per_step": swall / args.steps,
        }
        print(f"surrogate path: {value/1e6:.1f} Msteps/s vs FP64 {fp64_value/1e6:.1f} Msteps/s; "
              f"{nf} of {n} orbits fell back (dE/E > {args.max_rel_de:g} or outside the training radii); "
              f"force err median {np.median(ferr):.2e}, potential err median {np.median(perr):.2e}",
              file=sys.stderr, flush=True)

    cpu = None
    if args.cpu_orbits > 0:
        from oracle import pyoracle
        threads = args.cpu_threads or os.cpu_count()
        fast = pyoracle.OraclePotential(coef, library=pyoracle.fast())
        m = min(args.cpu_orbits, n)
        t0 = time.perf_counter()
        r = fast.orbits(ic[:m], T[:m], trajsize=1, accuracy=1e-8, nthreads=threads, method=args.method)
        dt = time.perf_counter() - t0
        cpu = {"value": float(r["nsteps"].sum() / dt), "unit": "orbit-steps/s", "cores": threads,
               "kind": "port",
               "sample": f"first {m} of the {n} orbits (same seeded ICs and times), FP64 Multipole, C++ "
                         f"OpenMP restatement built with -O2 (oracle/build/liboracle_fast.so), "
                         f"{dt:.1f} s on {threads} threads"}
        print(f"cpu baseline: {cpu['value']/1e6:.2f} Msteps/s on {threads} threads ({dt:.1f} s)",
              file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "orbit-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": (surrogate["wall_s_per_step"] if surrogate else wall / fp64_steps) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64 (fp32 spline tables)",
        "data": "synthetic",
        "config": {"workload": f"triaxial Dehnen Multipole lmax=8, {n} orbits, " +
                               (f"{args.periods:g}" if args.periods > 0 else "log-uniform 1-300") +
                               f" circular periods, {args.method.upper()} accuracy 1e-8, final state + max "
                               "dE/E <= 1e-6",
                   "orbits": n, "accepted_steps_per_step": steps_per_launch,
                   "attempts_per_step": attempts, "precision": args.precision,
                   "path": "MLP surrogate with FP64 fallback" if surrogate else "FP64 Multipole",
                   "group_mode": args.groups or "automatic",
                   "hip_graph": bool(args.graph),
                   "warmup_launches_per_leg": warmup_used,
                   "headline_timed_launches": args.steps,
                   "fp64_leg_timed_launches": fp64_steps},
        "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": ORBIT_TRAFFIC_BYTES if (n == 1_000_000 and args.precision == "fp64"
                                                        and args.method == "dop853") else None,
                     "kernel": "orbitKernel (FP64)" if args.method == "dop853" else "orbitKernelRKN (FP64)",
                     "kernel_ms": per_launch * 1e3,
                     "flops_per_attempt": FLOPS_PER_ATTEMPT[args.method],
                     "executed_frac": gpu_flops / per_launch / 1e12 / PEAK_FP64_TFLOPS},
        "cpu_baseline": cpu,
    }
    if surrogate:
        out["surrogate"] = surrogate
        out["fp64_multipole"] = {"value": fp64_value, "ms_per_step": wall / fp64_steps * 1e3,
                                 "timed_launches": fp64_steps}
    print(json.dumps(out), flush=True)


def bench_orbitobj(args):
    """orbit(dtype=object) on the device: 100k orbits x 10 circular periods with every step
    recorded (agx_orbit_record_batch), the quintic splines of the Orbit objects built, and 100
    evaluations per orbit at random times; host arrays in and out as the Python call has them.
    value = recorded rows (steps + 1 per orbit) per second of recording + spline build."""
    import numpy as np
    import agama_amd
    from agama_amd import lib, workloads
    from agama_amd._native import ptr
    coef = os.path.join(ROOT, "tests", "golden", "dehnen_triax_lmax8.coef")
    pot = agama_amd.Potential(file=coef)
    n = 100_000 if args.orbits == 1_000_000 else args.orbits
    ic, T = workloads.triaxial_dehnen_orbits(n, lambda p: pot.potential(p), pot.coefs(), seed=42, periods=10)
    sz = np.zeros(n, dtype=np.int32)
    rows = np.empty(n, dtype=np.int64)
    ns = np.empty(n, dtype=np.int64)
    na = np.empty(n, dtype=np.int64)
    q = 100
    rng = np.random.default_rng(3)
    which = np.repeat(np.arange(n, dtype=np.int64), q)
    tq = (rng.random(n * q) * np.repeat(T, q)).astype(np.float64)
    out = np.empty((n * q, 6))
    nodes = np.empty(n, dtype=np.int64)

    def step():
        t0 = time.perf_counter()
        h = lib.agx_orbit_record_batch(pot.handle, n, ptr(ic), None, ptr(T), ptr(sz), 1e-8, 0.0, int(1e8), 0,
                                       ptr(rows), None, ptr(ns), ptr(na), None, None, None, 0)
        if not h:
            raise RuntimeError(agama_amd._native.last_error())
        t1 = time.perf_counter()
        if lib.agx_orbit_record_series_info(h, ptr(nodes), None):
            raise RuntimeError(agama_amd._native.last_error())
        t2 = time.perf_counter()
        if lib.agx_orbit_record_interp(h, n * q, ptr(which), ptr(tq), 0, ptr(out)):
            raise RuntimeError(agama_amd._native.last_error())
        t3 = time.perf_counter()
        lib.agx_orbit_record_free(h)
        return t1 - t0, t2 - t1, t3 - t2

    for _ in range(args.warmup):
        step()
        print("orbitobj: warmup step done", file=sys.stderr, flush=True)
    times = [step() for _ in range(max(args.steps, 1))]
    rec = sum(t[0] for t in times) / len(times)
    build = sum(t[1] for t in times) / len(times)
    interp = sum(t[2] for t in times) / len(times)
    total_rows = int(rows.sum())
    attempts = int(na.sum())
    flops = attempts * FLOPS_PER_ATTEMPT["dop853"]
    print(f"orbitobj: {n} orbits, {total_rows} rows (median {np.median(rows):.0f} per orbit); record "
          f"{rec*1e3:.1f} ms, spline build {build*1e3:.1f} ms, {n*q} evaluations {interp*1e3:.1f} ms",
          file=sys.stderr, flush=True)
    cpu = None
    if args.cpu_orbits > 0:
        from oracle import pyoracle
        threads = args.cpu_threads or os.cpu_count()
        fast = pyoracle.OraclePotential(coef, library=pyoracle.fast())
        m = min(40000, n)
        t0 = time.perf_counter()
        r = fast.orbits_steps(ic[:m], T[:m], nthreads=threads, capacity=1024)
        for i, (ti, xi) in enumerate(r["series"]):
            pyoracle.orbit_interp(ti, xi, 0.0, tq[i * q:(i + 1) * q])
        dt = time.perf_counter() - t0
        cpu = {"value": float(r["rows"].sum() / dt), "unit": "recorded rows/s", "cores": threads, "kind": "port",
               "sample": f"first {m} of the {n} orbits: every-step integration (one pass, 1024-row blocks) "
                         f"with the C++ OpenMP restatement (-O2, {threads} threads) plus one Orbit built and "
                         f"evaluated {q} times per orbit (single thread), {dt:.1f} s"}
        print(f"cpu baseline: {cpu['value']/1e6:.2f} M rows/s ({dt:.1f} s)", file=sys.stderr, flush=True)
    achieved = flops / rec / 1e12
    res = {"metric": "recorded orbit rows/s of orbit(dtype=object): every step + Orbit splines (triaxial "
                     "Dehnen Multipole lmax=8)",
           "value": total_rows / (rec + build), "unit": "rows/s", "n_gpus": 1, "steps": len(times),
           "warmup": args.warmup, "ms_per_step": (rec + build + interp) * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{n} orbits x 10 circular periods, DOP853 accuracy 1e-8, every step "
                                  f"recorded, quintic splines built, {q} evaluations per orbit",
                      "rows": total_rows, "record_ms": rec * 1e3, "spline_build_ms": build * 1e3,
                      "evaluations": n * q, "evaluate_ms": interp * 1e3,
                      "evaluations_per_s": n * q / interp},
           "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                        "kernel": "orbitKernel (FP64) with every-step recording; wall time of the record call",
                        "kernel_ms": rec * 1e3},
           "cpu_baseline": cpu}
    print(json.dumps(res), flush=True)


def bench_target(args):
    """orbit(targets=[Target(type='DensitySphHarm', gridr=16 nodes, lmax=6, mmax=4)]): 100k orbits
    x 10 circular periods, DOP853 accuracy 1e-8, ten dense-output samples per step added into a
    145-value datacube per orbit; value = accepted steps/s of the whole call (host arrays in,
    float32 matrix out), beside the same orbits without the target"""
    import numpy as np
    import agama_amd
    from agama_amd import workloads
    coef = os.path.join(ROOT, "tests", "golden", "dehnen_triax_lmax8.coef")
    pot = agama_amd.Potential(file=coef)
    n = 100_000 if args.orbits == 1_000_000 else args.orbits
    ic, T = workloads.triaxial_dehnen_orbits(n, lambda p: pot.potential(p), pot.coefs(), seed=42, periods=10)
    gridr = np.concatenate([[0.0], np.logspace(-2, 3, 16)])
    tgt = agama_amd.Target(type="DensitySphHarm", gridr=gridr, lmax=6, mmax=4)
    info = {}
    for _ in range(args.warmup):
        agama_amd.orbit(potential=pot, ic=ic, time=T, targets=tgt)
        print("target: warmup step done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(max(args.steps, 1)):
        mat = agama_amd.orbit(potential=pot, ic=ic, time=T, targets=tgt, info=info)
    wall = (time.perf_counter() - t0) / max(args.steps, 1)
    steps = int(info["nsteps"].sum())
    t0 = time.perf_counter()
    agama_amd.orbit(potential=pot, ic=ic, time=T, info=info)
    plain = time.perf_counter() - t0
    print(f"target: {n} orbits, {steps} steps; with target {wall*1e3:.1f} ms, without {plain*1e3:.1f} ms; "
          f"datacube {mat.shape}", file=sys.stderr, flush=True)
    cpu = None
    if args.cpu_orbits > 0:
        from oracle import pyoracle
        threads = args.cpu_threads or os.cpu_count()
        fast = pyoracle.OraclePotential(coef, library=pyoracle.fast())
        m = min(20000, n)
        t0 = time.perf_counter()
        fast.orbits_target_sphharm(ic[:m], T[:m], gridr, 6, 4, nthreads=threads)
        dt = time.perf_counter() - t0
        frac = m / n
        cpu = {"value": steps * frac / dt, "unit": "orbit-steps/s", "cores": threads, "kind": "port",
               "sample": f"first {m} of the {n} orbits with the same target, C++ OpenMP restatement (-O2, "
                         f"{threads} threads), {dt:.1f} s (steps scaled by the sample's share of orbits)"}
        print(f"cpu baseline: {cpu['value']/1e6:.2f} Msteps/s ({dt:.1f} s)", file=sys.stderr, flush=True)
    out = {"metric": "orbit-steps/s with a DensitySphHarm target (triaxial Dehnen Multipole lmax=8)",
           "value": steps / wall, "unit": "orbit-steps/s", "n_gpus": 1, "steps": max(args.steps, 1),
           "warmup": args.warmup, "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{n} orbits x 10 circular periods, DOP853 accuracy 1e-8, DensitySphHarm target "
                                  "(16 radial nodes, lmax=6, mmax=4: 145 values per orbit)",
                      "accepted_steps": steps, "without_target_ms": plain * 1e3},
           "roofline": None, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def bench_var(args):
    """orbit(der=True) and orbit(lyapunov=True): 100k orbits x 10 circular periods, DOP853
    accuracy 1e-8; der=True integrates the six deviation vectors with the GL4 variational stepper
    (four Hessian evaluations per accepted step) and returns them at the end (trajsize=2);
    value = accepted steps/s of the der=True call (host arrays in, results out), beside the
    lyapunov=True call and the plain orbits"""
    import numpy as np
    import agama_amd
    from agama_amd import workloads
    coef = os.path.join(ROOT, "tests", "golden", "dehnen_triax_lmax8.coef")
    pot = agama_amd.Potential(file=coef)
    n = 100_000 if args.orbits == 1_000_000 else args.orbits
    ic, T = workloads.triaxial_dehnen_orbits(n, lambda p: pot.potential(p), pot.coefs(), seed=42, periods=10)
    info = {}
    for _ in range(args.warmup):
        agama_amd.orbit(potential=pot, ic=ic, time=T, trajsize=2, der=True)
        print("var: warmup step done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(max(args.steps, 1)):
        agama_amd.orbit(potential=pot, ic=ic, time=T, trajsize=2, der=True, info=info)
    wall = (time.perf_counter() - t0) / max(args.steps, 1)
    steps = int(info["nsteps"].sum())
    t0 = time.perf_counter()
    lyap = agama_amd.orbit(potential=pot, ic=ic, time=T, lyapunov=True)
    lyap_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    agama_amd.orbit(potential=pot, ic=ic, time=T, info={})
    plain = time.perf_counter() - t0
    chaotic = float(np.mean(np.asarray(lyap)[:, 0] > 0))
    print(f"var: {n} orbits, {steps} steps; der=True {wall*1e3:.1f} ms, lyapunov=True {lyap_s*1e3:.1f} ms, "
          f"plain {plain*1e3:.1f} ms; chaotic fraction {chaotic:.3f}", file=sys.stderr, flush=True)
    cpu = None
    if args.cpu_orbits > 0:
        from oracle import pyoracle
        threads = args.cpu_threads or os.cpu_count()
        fast = pyoracle.OraclePotential(coef, library=pyoracle.fast())
        m = min(40000, n)
        t0 = time.perf_counter()
        ref = fast.orbits_var(ic[:m], T[:m], trajsize=2, nthreads=threads)
        dt = time.perf_counter() - t0
        cpu = {"value": int(np.sum(ref["nsteps"])) / dt, "unit": "orbit-steps/s", "cores": threads, "kind": "port",
               "sample": f"first {m} of the {n} orbits with der=True, C++ OpenMP restatement (-O2, {threads} "
                         f"threads), {dt:.1f} s"}
        print(f"cpu baseline: {cpu['value']/1e6:.3f} Msteps/s ({dt:.1f} s)", file=sys.stderr, flush=True)
    out = {"metric": "orbit-steps/s with the six deviation vectors (der=True, triaxial Dehnen Multipole lmax=8)",
           "value": steps / wall, "unit": "orbit-steps/s", "n_gpus": 1, "steps": max(args.steps, 1),
           "warmup": args.warmup, "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{n} orbits x 10 circular periods, DOP853 accuracy 1e-8, der=True with "
                                  "trajsize=2; lyapunov=True on the same orbits",
                      "accepted_steps": steps, "lyapunov_ms": lyap_s * 1e3, "plain_ms": plain * 1e3,
                      "chaotic_fraction": chaotic},
           "roofline": None, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def bench_construct(args):
    """Multipole(density=...) construction: the Poisson solver (density at the quadrature
    points, forward spherical-harmonic transform, radial integrals) on the device against the
    host in the reference's order, for the triaxial Dehnen lmax=8 (50 radii) and lmax=16 (80
    radii) expansions; the timed unit is one agx_build_coef_file call (coefficients + file) and
    one agx_potential_create (coefficients + spline tables + upload)."""
    import agama_amd
    from agama_amd import lib
    configs = [("lmax=8, gridSizeR=50", 8, 50), ("lmax=16, gridSizeR=80", 16, 80)]
    reps = max(args.steps, 1) * 10
    rows = []
    for label, lmax, nr in configs:
        params = (f"type=Multipole, density=Dehnen, gamma=1, axisRatioY=0.8, axisRatioZ=0.6, lmax={lmax}, "
                  f"gridSizeR={nr}, rmin=1e-3, rmax=1e3").encode()
        res = {}
        for mode, name in ((0, "device"), (1, "host")):
            assert lib.agx_set_construction(mode) == 0
            path = f"/tmp/agx_bench_{os.getpid()}.coef".encode()
            for _ in range(max(args.warmup, 1)):
                if lib.agx_build_coef_file(params, path):
                    raise RuntimeError(agama_amd._native.last_error())
            t0 = time.perf_counter()
            for _ in range(reps):
                lib.agx_build_coef_file(params, path)
            coef_s = (time.perf_counter() - t0) / reps
            # a new grid every time: the device path also recomputes its grids and transform matrix
            t0 = time.perf_counter()
            for i in range(reps):
                lib.agx_build_coef_file(params.replace(b"rmax=1e3", f"rmax={1000 + i + 1}".encode()), path)
            cold_s = (time.perf_counter() - t0) / reps
            t0 = time.perf_counter()
            for _ in range(reps):
                h = lib.agx_potential_create(params)
                if not h:
                    raise RuntimeError(agama_amd._native.last_error())
                lib.agx_potential_destroy(h)
            create_s = (time.perf_counter() - t0) / reps
            res[name] = (coef_s, create_s, cold_s)
            os.remove(path)
        lib.agx_set_construction(-1)
        rows.append({"config": label, "device_coef_ms": res["device"][0] * 1e3, "host_coef_ms": res["host"][0] * 1e3,
                     "device_coef_new_grid_ms": res["device"][2] * 1e3, "host_coef_new_grid_ms": res["host"][2] * 1e3,
                     "device_create_ms": res["device"][1] * 1e3, "host_create_ms": res["host"][1] * 1e3})
        print(f"construct {label}: coefficients device {res['device'][0]*1e3:.2f} ms (new grid "
              f"{res['device'][2]*1e3:.2f}) / host {res['host'][0]*1e3:.2f} ms (new grid {res['host'][2]*1e3:.2f}); "
              f"potential create device {res['device'][1]*1e3:.2f} ms / host {res['host'][1]*1e3:.2f} ms",
              file=sys.stderr, flush=True)
    big = rows[-1]
    out = {"metric": "Multipole(density=Dehnen triaxial) constructions/s (Poisson solver + tables + upload)",
           "value": 1e3 / big["device_create_ms"], "unit": "constructions/s", "n_gpus": 1, "steps": reps,
           "warmup": args.warmup, "ms_per_step": big["device_create_ms"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "triaxial Dehnen gamma=1 (0.8, 0.6) expanded at lmax=16, 80 radii; lmax=8 / 50 "
                                  "radii alongside", "runs": rows},
           "roofline": None,
           "cpu_baseline": {"value": 1e3 / big["host_create_ms"], "unit": "constructions/s", "cores": 1,
                            "kind": "port", "sample": "the same construction with the Poisson solver on the "
                                                      "host in the reference's order (single thread)"}}
    print(json.dumps(out), flush=True)


# algorithmic bf16 MFMA operations of one surrogate evaluation: 3 hidden 256x256 products in
# the forward and 3 in the input-gradient backward (2 flops per multiply-add)
MLP_MFMA_FLOPS_PER_EVAL = 6 * 2 * 256 * 256
PEAK_BF16_TFLOPS = 2500.0    # MI355X dense bf16 MFMA, AMD published figure


def bench_mlp(args):
    """the surrogate's fused evaluation kernel on 1M points: evaluations/s and MFMA rate"""
    import numpy as np
    import torch
    import agama_amd
    from agama_amd import lib, workloads
    dev = torch.device("cuda:0")
    coef = os.path.join(ROOT, "tests", "golden", "dehnen_triax_lmax8.coef")
    pot = agama_amd.Potential(file=coef)
    sur = agama_amd.Surrogate.train(pot, samples=min(args.train_samples, 4e6), steps=min(args.train_steps, 1000))
    n = 1_000_000
    pts = torch.from_numpy(workloads.density_points(n, seed=7)).to(dev)
    phi = torch.empty(n, dtype=torch.float64, device=dev)
    grad = torch.empty((n, 3), dtype=torch.float64, device=dev)
    stream = torch.cuda.Stream()
    reps = max(args.steps, 1) * 5

    def launch():
        rc = lib.agx_surrogate_eval_many_device(sur._h, n, pts.data_ptr(), phi.data_ptr(), grad.data_ptr(),
                                                stream.cuda_stream)
        if rc:
            raise RuntimeError(agama_amd._native.last_error())
    with torch.cuda.stream(stream):
        for _ in range(max(args.warmup, 1)):
            launch()
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(reps):
            launch()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    per = ev0.elapsed_time(ev1) / 1e3 / reps
    achieved = n * MLP_MFMA_FLOPS_PER_EVAL / per / 1e12
    out = {"metric": "MLP surrogate potential+force evaluations/s at 1M points (bf16 MFMA, fp32 accumulation)",
           "value": n / per, "unit": "evaluations/s", "n_gpus": 1, "steps": reps, "warmup": args.warmup,
           "ms_per_step": wall / reps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "bf16 (fp32 accumulation; f64 base and assembly)", "data": "synthetic",
           "config": {"workload": "surrogate Phi and grad Phi at 1M points drawn from the triaxial Dehnen density"},
           "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_BF16_TFLOPS, "traffic": None,
                        "kernel": "surrogateEvalKernel", "kernel_ms": per * 1e3},
           "cpu_baseline": None}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
