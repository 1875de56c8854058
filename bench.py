"""Benchmark of the DQMC sweep engine on one MI355X.

A step is one single-direction sweep of every chain of the batch (one
k_sweep launch).  The default workload is the batched configuration: 256
independent chains of the 12x12 repulsive Hubbard model (t=1, U=4, mu=0,
beta=6, 60 slices, stabilization pace 10, checkerboard break-up, FP64),
seeds 0..255.  Beside it the same chain shape runs alone (the plain,
one-chain-per-launch path) so the batched/plain ratio is reported, and the
oracle (one chain per process, FP64 CPU restatement of the reference) is
timed on the host cores for the CPU baseline.

Prints ONE JSON line.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dqmc-framework_amd"))

import dqmc_amd as D  # noqa: E402
import numpy as np  # noqa: E402

METRIC = "chain-sweeps/s batched vs plain FP64 GPU at max|dG|<=1e-6; p50/p99 job latency"

WORKLOADS = {
    "batched-12x12": dict(chains=256, side=12, time_size=60, beta=6.0, pace=10, model=D.REPULSIVE,
                          checkerboard=True, U=4.0, mu=0.0, delay=32,
                          desc="256 chains x 12x12 repulsive Hubbard t=1 U=4 mu=0 beta=6 L=60 p=10 checkerboard FP64"),
    "plain-8x8": dict(chains=1, side=8, time_size=160, beta=8.0, pace=10, model=D.ATTRACTIVE,
                      checkerboard=True, U=4.0, mu=-0.5, delay=16,
                      desc="1 chain 8x8 attractive Hubbard t=1 |U|=4 mu=-0.5 beta=8 L=160 p=10 checkerboard k=16 FP64"),
    "plain-4x4": dict(chains=1, side=4, time_size=80, beta=4.0, pace=10, model=D.REPULSIVE,
                      checkerboard=False, U=4.0, mu=0.0, delay=16,
                      desc="1 chain 4x4 repulsive Hubbard t=1 U=4 mu=0 beta=4 L=80 p=10 dense FP64"),
    "mixed-16x16": dict(chains=64, side=16, time_size=200, beta=10.0, pace=8, model=D.REPULSIVE,
                        checkerboard=False, U=6.0, mu=0.0, delay=32, precision=D.PREC_MIXED,
                        mus=[0.0, -0.2, -0.4, -0.6],
                        desc="64 chains x 16x16 repulsive Hubbard t=1 U=6 beta=10 L=200 p=8 dense, FP32 MFMA "
                             "products/wraps + FP64 stacks"),
}

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# rocprofv3 summary holding the measured HBM bytes per k_sweep launch of a workload
# (FETCH_SIZE and WRITE_SIZE passes scaled by a copy of known size)
TRAFFIC_PROFILES = {"batched-12x12": "profiles/k_sweep_256x12x12_L60_checkerboard_fp64.md"}


def profiled_traffic(name):
    """HBM bytes per launch recorded by the workload's profile summary, or None."""
    path = TRAFFIC_PROFILES.get(name)
    if path is None or not os.path.exists(os.path.join(ROOT, path)):
        return None, None
    with open(os.path.join(ROOT, path)) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1]), path
    return None, None


def chain_params(w, chains=None, seed0=0):
    n = w["chains"] if chains is None else chains
    mus = w.get("mus", [w["mu"]])
    return [D.make_params(w["side"], w["time_size"], w["beta"], pace=w["pace"], model=w["model"],
                          checkerboard=w["checkerboard"], delay=w["delay"], U=w["U"], mu=mus[c * len(mus) // n],
                          seed=seed0 + c, precision=w.get("precision", D.PREC_FP64))
            for c in range(n)]


def algorithmic_flops_per_chain_sweep(w, accepted_per_chain_sweep):
    """Per chain-sweep FLOPs (both spins; attractive counts one spin):
    wrap 8LN^3 dense / 64LN^2 checkerboard, slice products 4LN^3 / 32LN^2,
    updates 4 A N^2, stabilization 29.3 n_s N^3 (pivoted-QR UDV, both spins)."""
    N = w["side"] ** 2
    L = w["time_size"]
    ns = -(-L // w["pace"])
    spins = 1 if w["model"] == D.ATTRACTIVE else 2
    if w["checkerboard"]:
        wrap, prod = 64.0 * L * N * N, 32.0 * L * N * N
    else:
        wrap, prod = 8.0 * L * N ** 3, 4.0 * L * N ** 3
    upd = 4.0 * accepted_per_chain_sweep * N * N
    stab = 29.3 * ns * N ** 3
    return (wrap + prod + upd + stab) * spins / 2.0


def algorithmic_flops_per_dynamic_sweep(w):
    """Per chain FLOPs of a time-displaced sweep (both spins; attractive counts one spin):
    the wrap of G(t,t), three slice products per slice (G(t,0), G(0,t) and the stack's
    product), and per stabilization the push plus a recompute with 2N right-hand sides and two
    more N^3 products than the equal-time one (41.3 n_s N^3 for both spins)."""
    N = w["side"] ** 2
    L = w["time_size"]
    ns = -(-L // w["pace"])
    spins = 1 if w["model"] == D.ATTRACTIVE else 2
    if w["checkerboard"]:
        wrap, prod = 64.0 * L * N * N, 32.0 * L * N * N
    else:
        wrap, prod = 8.0 * L * N ** 3, 4.0 * L * N ** 3
    return (wrap + 3.0 * prod + 41.3 * ns * N ** 3) * spins / 2.0


def autotune(w, candidates=(8, 16, 32)):
    """Delayed-update depth and chain -> workgroup mapping for the workload's chain shape, chosen
    by the engine's autotuners (each candidate timed on a full batch of the workload's size)."""
    ctx = D.Context(0)
    best, ms = ctx.autotune_delay(chain_params(w, 1)[0], w["chains"], candidates=candidates, sweeps=2)
    p = chain_params(w, 1)[0]
    p.delay = best
    mode, mode_ms = ctx.autotune_launch(p, w["chains"], sweeps=2)
    ctx.close()
    names = {D.LAUNCH_SINGLE: "single", D.LAUNCH_PAIR: "pair"}
    return best, ms, names[mode], {names[m]: t for m, t in mode_ms.items()}


LAUNCH_MODES = {"auto": D.LAUNCH_AUTO, "single": D.LAUNCH_SINGLE, "pair": D.LAUNCH_PAIR}


def run_gpu(w, steps, warmup, chains=None, graph=True, launch="auto"):
    ctx = D.Context(0)
    b = D.Batch(ctx, chain_params(w, chains))
    b.use_graph(graph)   # the one-workgroup-per-chain launch replayed from a captured HIP graph
    b.set_launch_mode(LAUNCH_MODES[launch])   # AUTO: spin pairs when the batch leaves CUs idle
    b.fields_random()
    b.init()
    for _ in range(warmup):
        b.sweep(1)
    b.synchronize()
    n = b.n
    acc0 = 0
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        b.sweep(1)
        kernel_ms.append(b.last_sweep_ms())
        acc0 += sum(b.info(c).accepted for c in range(min(n, 8)))
    b.synchronize()
    wall = time.perf_counter() - t0
    acc_per_chain_sweep = acc0 / (min(n, 8) * steps)
    wrap_err = max(b.info(c).wrap_error for c in range(n))
    statuses = [b.info(c).status for c in range(n)]
    used = b.last_launch_mode()
    b.close()
    ctx.close()
    mode = ("spin pairs (two workgroups per chain)" if used == D.LAUNCH_PAIR else
            "one workgroup per chain, " + ("hipGraph replay" if graph else "direct"))
    return dict(wall=wall, kernel_ms=kernel_ms, chains=n, accepted=acc_per_chain_sweep, wrap_error=wrap_err,
                status=max(statuses), launch=mode)


def phase_breakdown(w, chains, sweeps=2, launch="auto"):
    """Device time per sweep phase (ms per chain-sweep), from the engine's phase clock (in a
    spin-pair launch, the clock of spin 0's workgroup)."""
    ctx = D.Context(0)
    b = D.Batch(ctx, chain_params(w, chains))
    b.set_launch_mode(LAUNCH_MODES[launch])
    b.fields_random()
    b.init()
    b.sweep(1)
    b.sweep(sweeps, profile=True)
    b.synchronize()
    n = min(b.n, 8)
    nph = len(D.PHASES)
    tot = [0.0] * nph
    for c in range(n):
        ph = b.info(c).phase_seconds
        for q in range(nph):
            tot[q] += ph[q]
    out = {name: 1000.0 * tot[q] / (n * sweeps) for q, name in enumerate(D.PHASES)}
    out["launch_ms_per_sweep"] = b.last_sweep_ms() / sweeps
    b.close()
    ctx.close()
    return out


def phase_rooflines(w, phases, chains, accepted):
    """Chip-level rate of each sweep phase: its algorithmic work per chain-sweep x chains
    over the phase's device time per chain-sweep (the chains run concurrently, one per CU).
    HBM-bound phases in GB/s of algorithmic bytes (G read and written per pass), MFMA-bound ones
    in TFLOP/s.  The Metropolis phase is priced at the reference's unblocked rank-1 updates
    (2 N^2 FLOP and 16 N^2 B per spin and accepted flip), which the delayed block updates
    replace; the same phase's executed MFMA work (2 N^3 + 2 k N^2 per slice and spin) is given
    beside it."""
    N = w["side"] ** 2
    L = w["time_size"]
    ns = -(-L // w["pace"])
    spins = 1 if w["model"] == D.ATTRACTIVE else 2
    k = min(w["delay"], 32)
    cb = w["checkerboard"]
    out = {}

    def rate(name, amount, unit, peak):
        ms = phases.get(name, 0.0)
        if ms <= 0:
            return
        a = amount * chains / (ms * 1e-3) / (1e9 if unit == "GB/s" else 1e12)
        out[name] = {"achieved": a, "unit": unit, "peak": peak, "frac": a / peak, "per_chain_sweep": amount}

    if cb:
        rate("wrap", 4 * 8.0 * N * N * L * spins, "GB/s", PEAK_HBM_GBS)
        rate("product", 2 * 8.0 * N * N * L * spins, "GB/s", PEAK_HBM_GBS)
    else:
        rate("wrap", 8.0 * L * N ** 3 * spins / 2, "TFLOP/s", PEAK_FP64_TFLOPS)
        rate("product", 4.0 * L * N ** 3 * spins / 2, "TFLOP/s", PEAK_FP64_TFLOPS)
    rate("metropolis", 16.0 * accepted * N * N * spins, "GB/s", PEAK_HBM_GBS)
    if "metropolis" in out:
        blocks = L * spins * (2.0 * N ** 3 + 2.0 * k * N * N)
        out["metropolis"]["executed_mfma_tflops"] = blocks * chains / (phases["metropolis"] * 1e-3) / 1e12
    stab_ms = phases.get("push", 0.0) + phases.get("fresh", 0.0)
    if stab_ms > 0:
        amount = 29.3 * ns * N ** 3 * spins / 2
        a = amount * chains / (stab_ms * 1e-3) / 1e12
        out["push+fresh"] = {"achieved": a, "unit": "TFLOP/s", "peak": PEAK_FP64_TFLOPS, "frac": a / PEAK_FP64_TFLOPS,
                             "per_chain_sweep": amount}
    return out


def run_dynamic(w, steps, chains=None, profile=False):
    """The measuring schedule with time-displaced observables (src/dqmc.cpp:48-70): each
    forth-and-back is a dynamic sweep (fields fixed, G(t,0) and G(0,t) with G(t,t), dynamic
    observables accumulated) and an updating backward sweep.  Times the dynamic sweeps."""
    ctx = D.Context(0)
    b = D.Batch(ctx, chain_params(w, chains))
    b.fields_random()
    b.init()
    b.sweep(2)
    dyn_ms = []
    for _ in range(steps):
        b.sweep(1, measure=True, dynamic=True, profile=profile)
        dyn_ms.append(b.last_sweep_ms())
        b.sweep(1, measure=True, dynamic=True)
    b.synchronize()
    n = b.n
    kms = sum(dyn_ms) / len(dyn_ms)
    achieved = algorithmic_flops_per_dynamic_sweep(w) * n / (kms * 1e-3) / 1e12
    out = {"chains": n, "ms_per_dynamic_sweep": kms, "chain_dynamic_sweeps_per_s": n / (kms * 1e-3),
           "achieved_tflops": achieved, "frac_fp64_peak": achieved / PEAK_FP64_TFLOPS,
           "max_wrap_error": max(b.info(c).wrap_error for c in range(n)),
           "measurements_per_chain": b.dynamic_accum(0)["count"]}
    if profile:
        m = min(n, 8)
        nph = len(D.PHASES)
        tot = [0.0] * nph
        for c in range(m):
            ph = b.info(c).phase_seconds
            for q in range(nph):
                tot[q] += ph[q]
        out["phases_ms_per_chain_sweep"] = {name: 1000.0 * tot[q] / (m * steps) for q, name in enumerate(D.PHASES)}
    b.close()
    ctx.close()
    return out


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return "host CPU"


def oracle_run(args, cpus, seconds):
    """Oracle chains, one process per chain, each pinned to its own CPU; chain-sweeps/s of each."""
    exe = os.path.join(ROOT, "oracle", "_build", "refcpu")
    if not os.path.exists(exe):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    base = [exe] + args + ["--warm", "100000", "--bins", "0", "--max-seconds", str(seconds)]
    ps = [subprocess.Popen(base + ["--seed", str(c)], stdout=subprocess.PIPE,
                           preexec_fn=(lambda cpu=cpu: os.sched_setaffinity(0, {cpu})))
          for c, cpu in enumerate(cpus)]
    rates = []
    for p in ps:
        out, _ = p.communicate()
        rates.append(json.loads(out.decode().strip().splitlines()[-1])["sweeps_per_s"])
    return rates


def cpu_baseline(w, seconds=12.0, max_procs=16):
    """The reference's CPU path (one chain per MPI rank, src/dqmc_main.cpp:117) timed as the
    oracle, one chain per process pinned one per core, on every CPU this process may run on up to
    the 16 a one-GPU box gives it.  Calibrated on the reference's example configuration (8x8
    attractive, U 4, beta 8, 160 slices, dense, pace 10: its warm-up ran 512 sweeps in 53.95 s on
    one rank, 9.49 sweeps/s, example/log.out), timed here on one core."""
    cpus = sorted(os.sched_getaffinity(0))[:max_procs]
    args = ["--side", str(w["side"]), "--slices", str(w["time_size"]), "--beta", str(w["beta"]),
            "--pace", str(w["pace"]), "--model", str(w["model"]), "--cb", str(int(w["checkerboard"])),
            "--U", str(w["U"]), "--mu", str(w["mu"])]
    rates = oracle_run(args, cpus, seconds)
    example = ["--side", "8", "--slices", "160", "--beta", "8.0", "--pace", "10", "--model", "1", "--cb", "0",
               "--U", "4.0", "--mu", "0.0"]
    cal = oracle_run(example, cpus[:1], 8.0)[0]
    model = cpu_model()
    return {"value": sum(rates), "unit": "chain-sweeps/s", "cores": len(cpus), "kind": "port",
            "per_core": sum(rates) / len(cpus), "cpu": model,
            "calibration": {"config": "example/config.toml: 8x8 attractive U=4 beta=8 L=160 dense p=10",
                            "oracle_sweeps_per_s_per_core": cal, "reference_sweeps_per_s_per_core": 9.49,
                            "oracle_over_reference": cal / 9.49},
            "sample": f"{len(cpus)} oracle processes pinned one per core (one chain each, {w['side']}x{w['side']}, "
                      f"L={w['time_size']}), ~{seconds:.0f} s of sweeps each, on {model}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--workload", default="batched-12x12", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the single-chain comparison run")
    ap.add_argument("--phases", action="store_true", help="also report device time per sweep phase")
    ap.add_argument("--launch", default="auto", choices=sorted(LAUNCH_MODES),
                    help="chain -> workgroup mapping: auto (spin pairs when the batch leaves CUs idle), single, pair")
    ap.add_argument("--precision", choices=["workload", "fp64", "mixed", "fp32_wrap", "fp32_product"],
                    default="workload", help="override the workload's precision mode")
    ap.add_argument("--dynamic-steps", type=int, default=3,
                    help="time-displaced sweeps timed on the same batch (0 skips them)")
    ap.add_argument("--no-autotune", action="store_true", help="use the workload's fixed delay depth")
    ap.add_argument("--serve-jobs", type=int, default=2000, help="jobs of the serving stream (0 skips it)")
    ap.add_argument("--serve-scale", type=float, default=0.1, help="sweep-count scale of the serving stream")
    ap.add_argument("--serve-rate", type=float, default=40.0, help="Poisson arrivals per second")
    ap.add_argument("--no-graph", action="store_true", help="launch the sweeps directly instead of replaying a graph")
    ap.add_argument("--gemm-timing", default="", metavar="SIZES",
                    help="only run the workgroup GEMM of the N x N x N products in its 64x64-tile form (variant 1) "
                         "and its wide-tile forms (5: 128x128, 6: 160x160) three times each per size (comma-separated "
                         "N), for a kernel trace of k_gemm, and exit")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only fill and copy a device buffer of this size and exit (a known-traffic reference for "
                         "profiler byte counters)")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"fill_bytes_written": n * 8, "copy_bytes_read": n * 8, "copy_bytes_written": n * 8}))
        return
    if args.gemm_timing:
        ctx = D.Context(0)
        rng = np.random.default_rng(0)
        runs = []
        for n in [int(x) for x in args.gemm_timing.split(",")]:
            A = rng.standard_normal((n, n))
            B = rng.standard_normal((n, n))
            ref = ctx.gemm(1, A, B)
            for v in (1, 5, 6):
                for _ in range(3):
                    assert np.array_equal(ctx.gemm(v, A, B), ref)   # the forms give the same bits
                runs.append({"N": n, "variant": v, "launches": 3})
        ctx.close()
        print(json.dumps({"gemm_timing": runs, "order": "k_gemm dispatches: per N, one reference launch (variant 1), "
                                                        "then 3 launches of each variant in the listed order"}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: chains are independent and would be sharded by chain across GPUs, "
              f"but only the single-GPU figure is measured here (asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)
    w = dict(WORKLOADS[args.workload])
    if args.precision != "workload":
        w["precision"] = {"fp64": D.PREC_FP64, "mixed": D.PREC_MIXED, "fp32_wrap": D.PREC_FP32_WRAP,
                          "fp32_product": D.PREC_FP32_PRODUCT}[args.precision]
        w["desc"] += f" (precision override: {args.precision})"
    tune = launch_tune = None
    if not args.no_autotune and w["chains"] > 1:
        best, tune, mapping, launch_tune = autotune(w)
        w["delay"] = best
        if args.launch == "auto":
            args.launch = mapping
        print(f"[bench] autotuned: delay depth {best} (ms per 2 sweeps {tune}), mapping {mapping} ({launch_tune})",
              file=sys.stderr, flush=True)
    print(f"[bench] {w['desc']}: warmup {args.warmup}, steps {args.steps}", file=sys.stderr, flush=True)
    r = run_gpu(w, args.steps, args.warmup, graph=not args.no_graph, launch=args.launch)
    chain_sweeps = r["chains"] * args.steps
    value = chain_sweeps / r["wall"]
    ms_per_step = 1000.0 * r["wall"] / args.steps
    kms = sum(r["kernel_ms"]) / len(r["kernel_ms"])
    flops_launch = algorithmic_flops_per_chain_sweep(w, r["accepted"]) * r["chains"]
    achieved = flops_launch / (kms * 1e-3) / 1e12
    traffic, traffic_src = profiled_traffic(args.workload)
    out = {
        "metric": METRIC, "value": value, "unit": "chain-sweeps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if w.get("precision", D.PREC_FP64) == D.PREC_FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": w["desc"], "chains": r["chains"], "side": w["side"], "time_size": w["time_size"],
                   "beta": w["beta"], "pace": w["pace"], "checkerboard": w["checkerboard"], "delay": w["delay"]},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "kernel": "k_sweep", "kernel_ms": kms, "algorithmic_gflop_per_launch": flops_launch / 1e9},
        "accepted_per_chain_sweep": r["accepted"], "max_wrap_error": r["wrap_error"],
        "launch": r["launch"],
    }
    if tune is not None:
        out["config"]["delay_autotune_ms"] = tune
        out["config"]["launch_autotune_ms"] = launch_tune
    if traffic is not None:
        out["roofline"]["traffic_source"] = f"{traffic_src} (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated)"
        out["roofline"]["traffic_tbs_at_kernel_ms"] = traffic / (kms * 1e-3) / 1e12
    if not args.no_plain and r["chains"] > 1:
        print("[bench] plain path: the same chain alone, one chain per launch", file=sys.stderr, flush=True)
        p = run_gpu(w, max(2, args.steps // 4), 1, chains=1, launch=args.launch)
        plain = p["chains"] * max(2, args.steps // 4) / p["wall"]
        out["plain_chain_sweeps_per_s"] = plain
        out["batched_over_plain"] = value / plain
    if args.phases:
        print("[bench] phase breakdown", file=sys.stderr, flush=True)
        out["phases_ms_per_chain_sweep"] = {"batched": phase_breakdown(w, r["chains"], launch=args.launch),
                                            "single": phase_breakdown(w, 1, launch=args.launch)}
        out["phase_rooflines"] = phase_rooflines(w, out["phases_ms_per_chain_sweep"]["batched"], r["chains"],
                                                 r["accepted"])
    if w.get("precision", D.PREC_FP64) != D.PREC_FP64:
        import precision
        print("[bench] precision check: recomputed G at every stabilization vs FP64", file=sys.stderr, flush=True)
        ctx = D.Context(0)
        picks = [0, r["chains"] // 4] if r["chains"] > 1 else [0]
        dg = precision.stabilization_errors(ctx, [chain_params(w)[c] for c in picks])
        ctx.close()
        out["precision_check"] = {
            "chains": picks, "stabilizations_x_spins": int(dg.size), "dG_recomputed_median": float(np.median(dg)),
            "dG_recomputed_max": float(dg.max()), "wrap_error_max_over_batch": r["wrap_error"],
            "target_1e-6_met": bool(dg.max() <= 1e-6 and r["wrap_error"] <= 1e-6)}
    if args.dynamic_steps > 0:
        print(f"[bench] time-displaced sweeps: {args.dynamic_steps}", file=sys.stderr, flush=True)
        out["dynamic"] = run_dynamic(w, args.dynamic_steps, profile=args.phases)
    if args.serve_jobs > 0:
        import serve
        print(f"[bench] serving sample: {args.serve_jobs} jobs at {args.serve_rate}/s", file=sys.stderr, flush=True)
        jobs = serve.generate_jobs(args.serve_jobs, scale=args.serve_scale)
        ctx = D.Context(0)
        summary, _ = serve.run(ctx, jobs, args.serve_rate, 160 << 30, log=sys.stderr)
        ctx.close()
        summary["job_mix"] = (f"side in {{4..20}}, beta U[2,12], dtau in {{0.05,0.1}}, U U[-8,8], mu U[-1,1], "
                              f"sweeps U[50,500] x {args.serve_scale}, seed = job id, FP64, checkerboard")
        out["serving"] = summary
    if not args.no_cpu:
        print("[bench] CPU baseline (oracle, one chain per process)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
