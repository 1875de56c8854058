"""Serving driver: a Poisson stream of DQMC jobs through the continuous-batching
server (dqmc_server_* of the C-ABI), with p50/p99 job latency and throughput.

Job mix (the serving configuration): side L uniform in {4,6,8,10,12,16,20};
beta uniform in [2,12]; dtau in {0.05, 0.1}, time_size = round(beta/dtau) and
dtau = beta/time_size; U uniform in [-8,8] (U < 0 runs the attractive model
with |U|); mu uniform in [-1,1]; sweeps uniform in [50,500]; seed = job id;
stabilization pace 10.  A fifth of the sweeps (rounded) warm up, the rest are
measured in bins of 10 sweeps.  `scale` shrinks the sweep counts for short
runs; the shape distribution is unchanged.  Jobs run in FP64; the FP32
product modes are an explicit opt-in (--mixed), since on live chains they do
not hold G to the FP64 chain's (the job results carry the wrap error and the
chain's status bits).

Usage: python serve.py --jobs 200 --rate 20 --scale 0.1
       python serve.py --configs a.toml b.toml ...   (jobs from the reference's config files)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import dqmc_amd as D  # noqa: E402

SIDES = [4, 6, 8, 10, 12, 16, 20]


def generate_jobs(n, seed=2024, scale=1.0, precision=D.PREC_FP64, checkerboard=True, bin_size=10):
    rng = np.random.default_rng(seed)
    jobs = []
    for jid in range(n):
        side = int(rng.choice(SIDES))
        beta = float(rng.uniform(2.0, 12.0))
        dtau = float(rng.choice([0.05, 0.1]))
        L = max(1, int(round(beta / dtau)))
        U = float(rng.uniform(-8.0, 8.0))
        mu = float(rng.uniform(-1.0, 1.0))
        sweeps = int(rng.integers(50, 501))
        sweeps = max(2 * bin_size, int(round(sweeps * scale)))
        warm = max(2, int(round(sweeps / 5)))
        meas = max(bin_size, ((sweeps - warm) // bin_size) * bin_size)
        model = D.ATTRACTIVE if U < 0 else D.REPULSIVE
        p = D.make_params(side, L, beta, pace=10, model=model, checkerboard=checkerboard, delay=16,
                          U=abs(U), mu=mu, seed=jid, precision=precision)
        jobs.append(dict(id=jid, params=p, warmup=warm, measure=meas, bin_size=bin_size, side=side, L=L,
                         beta=beta, U=U, mu=mu))
    return jobs


def jobs_from_configs(paths, precision=D.PREC_FP64):
    """Serving jobs from TOML configuration files in the reference's format
    (src/dqmc_initializer.cpp:27-338); job k uses seed k."""
    import jobspec
    jobs = []
    for k, path in enumerate(paths):
        spec = jobspec.load(path)
        j = jobspec.job(spec, seed=k, precision=precision)
        j.update(id=k, side=spec.side, L=spec.time_size, beta=spec.beta, U=spec.onsite_u, mu=spec.chemical_potential,
                 config=path)
        jobs.append(j)
    return jobs


def run(ctx, jobs, rate, budget_bytes, lanes=4, max_active=4096, arrival_seed=7, timeout_s=3000.0, log=None):
    """Submit jobs at Poisson arrival times (rate per second) and wait for all."""
    rng = np.random.default_rng(arrival_seed)
    gaps = rng.exponential(1.0 / rate, size=len(jobs)) if rate > 0 else np.zeros(len(jobs))
    srv = D.Server(ctx, budget_bytes, max_active=max_active, lanes=lanes)
    t0 = time.perf_counter()
    results = {}
    idmap = {}
    t_next = 0.0
    k = 0
    last_log = t0
    while len(results) < len(jobs):
        now = time.perf_counter() - t0
        while k < len(jobs) and now >= t_next:
            j = jobs[k]
            sid = srv.submit(j["params"], j["warmup"], j["measure"], j["bin_size"], dynamic=j.get("dynamic", False))
            idmap[sid] = j["id"]
            k += 1
            if k < len(jobs):
                t_next += gaps[k]
        for r in srv.poll():
            results[idmap[r["id"]]] = r
        if log and time.perf_counter() - last_log > 20:
            st = srv.stats()
            print(f"[serve] t={now:.0f}s submitted {k} done {len(results)} active {st['active']} "
                  f"queued {st['queued']} chain-sweeps {st['chain_sweeps']}", file=log, flush=True)
            last_log = time.perf_counter()
        if now > timeout_s:
            raise RuntimeError(f"serving did not finish in {timeout_s} s ({len(results)}/{len(jobs)} done)")
        time.sleep(0.002)
    wall = time.perf_counter() - t0
    stats = srv.stats()
    srv.close()
    lat = np.array([results[j["id"]]["latency"] for j in jobs])
    sweeps = sum(results[j["id"]]["sweeps"] for j in jobs)
    return {
        "jobs": len(jobs), "rate_per_s": rate, "wall_s": wall, "chain_sweeps": int(sweeps),
        "chain_sweeps_per_s": sweeps / wall, "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)), "mean_latency_s": float(lat.mean()),
        "launches": stats["launches"], "peak_bytes": stats["bytes_peak"], "budget_bytes": stats["budget"],
        "failed": sum(1 for r in results.values() if r["status"] != 0),
    }, results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=200)
    ap.add_argument("--rate", type=float, default=20.0, help="Poisson arrivals per second")
    ap.add_argument("--scale", type=float, default=0.1, help="sweep-count scale of the job mix")
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--lanes", type=int, default=4, help="size-class lanes (N <= 64, <= 144, <= 256, > 256)")
    ap.add_argument("--mixed", action="store_true",
                    help="FP32 matrix-core products and wraps (opt-in; FP64 by default)")
    ap.add_argument("--configs", nargs="+", default=None, help="TOML configuration files to serve as jobs")
    args = ap.parse_args()
    if args.configs:
        jobs = jobs_from_configs(args.configs, precision=D.PREC_MIXED if args.mixed else D.PREC_FP64)
    else:
        jobs = generate_jobs(args.jobs, scale=args.scale, precision=D.PREC_MIXED if args.mixed else D.PREC_FP64)
    ctx = D.Context(0)
    summary, _ = run(ctx, jobs, args.rate, int(args.budget_gb * 2 ** 30), lanes=args.lanes, log=sys.stderr)
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
