"""Continuous-batching server: admission by footprint, jobs joining while
others run, cancellation, and that every job's result is the same alone and
in the mix (the per-chain computation does not depend on co-resident chains)."""
import time

import numpy as np
import pytest

import dqmc_amd as D
import serve

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return D.Context(0)


def batch_reference(ctx, job):
    """The same schedule through the batch API: warm-up, then bins of measured sweeps."""
    b = D.Batch(ctx, [job["params"]])
    b.fields_random()
    b.init()
    b.sweep(job["warmup"])
    bins = []
    for _ in range(job["measure"] // job["bin_size"]):
        b.clear_measure()
        b.sweep(job["bin_size"], measure=True)
        s, c = b.measure_accum(0)
        bins.append(D.normalize_bin(s, c))
    mean, err = D.bin_stats(bins) if len(bins) > 1 else (bins[0], np.zeros(9))
    return mean, err


def small_jobs(n, seed=1):
    jobs = serve.generate_jobs(n, seed=seed, scale=0.06)
    # keep the test quick: lattices up to 8x8 and at most 80 slices
    out = []
    for j in jobs:
        if j["side"] <= 8 and j["L"] <= 80:
            out.append(j)
    return out


def test_server_matches_batch_api_and_is_mix_independent(ctx):
    jobs = small_jobs(60)[:12]
    assert len(jobs) >= 6
    print(f"[gpu-serve] {len(jobs)} jobs", flush=True)
    # alone: each job through its own server
    alone = {}
    for j in jobs[:3]:
        summary, res = serve.run(ctx, [j], rate=0, budget_bytes=1 << 30)
        alone[j["id"]] = res[j["id"]]
        mean, err = batch_reference(ctx, j)
        assert np.array_equal(np.array(res[j["id"]]["mean"]), mean), j["id"]
    # in the mix: all jobs, Poisson arrivals, jobs join while others run
    summary, res = serve.run(ctx, jobs, rate=40.0, budget_bytes=1 << 30)
    assert summary["failed"] == 0
    for jid, r in alone.items():
        assert res[jid]["mean"] == r["mean"]
        assert res[jid]["error"] == r["error"]
        assert res[jid]["sweeps"] == r["sweeps"]
    assert summary["p99_latency_s"] >= summary["p50_latency_s"] > 0


def test_admission_by_footprint_and_rejection(ctx):
    jobs = small_jobs(40, seed=3)[:6]
    sizes = [D.footprint([j["params"]]) for j in jobs]
    budget = max(sizes) * 2 + 4096   # room for about two resident chains
    summary, res = serve.run(ctx, jobs, rate=0, budget_bytes=budget)
    assert summary["failed"] == 0 and len(res) == len(jobs)
    assert summary["peak_bytes"] <= budget
    srv = D.Server(ctx, 1 << 20)
    big = D.make_params(20, 200, 10.0, pace=10)
    with pytest.raises(D.DqmcError):
        srv.submit(big, 10, 10, 10)
    srv.close()


def test_cancel_frees_slot_and_leaves_others_unchanged(ctx):
    jobs = small_jobs(40, seed=5)[:5]
    ref = {}
    for j in jobs:
        ref[j["id"]] = batch_reference(ctx, j)[0]
    srv = D.Server(ctx, 1 << 30)
    ids = {}
    for j in jobs:
        ids[srv.submit(j["params"], j["warmup"], j["measure"], j["bin_size"])] = j["id"]
    victim = list(ids)[2]
    assert srv.cancel(victim)
    assert srv.wait_idle(600)
    got = {}
    for r in srv.poll():
        got[ids[r["id"]]] = r
    assert got[ids[victim]]["status"] in (0, 1)
    for sid, jid in ids.items():
        if sid == victim:
            continue
        assert got[jid]["status"] == 0
        assert np.array_equal(np.array(got[jid]["mean"]), ref[jid])
    st = srv.stats()
    assert st["bytes_in_use"] == 0
    srv.close()


def test_dynamic_job_in_the_mix_matches_batch_api(ctx):
    # a job with the dynamic option (time-displaced sweeps opening each measuring iteration)
    # served among equal-time jobs gives the batch API's dynamic and equal-time results
    print("[gpu-serving] dynamic job in a mix", flush=True)
    p = D.make_params(6, 24, 2.0, pace=6, checkerboard=True, U=3.0, mu=-0.1, seed=77)
    warm, meas, bs = 6, 24, 8
    srv = D.Server(ctx, 1 << 30)
    others = small_jobs(60, seed=5)[:6]
    assert len(others) >= 4
    ids = [srv.submit(j["params"], j["warmup"], j["measure"], j["bin_size"]) for j in others[:3]]
    jid = srv.submit(p, warm, meas, bs, dynamic=True)
    ids += [srv.submit(j["params"], j["warmup"], j["measure"], j["bin_size"]) for j in others[3:]]
    results = {}
    t0 = time.time()
    while len(results) < len(ids) + 1 and time.time() - t0 < 120:
        for r in srv.poll():
            results[r["id"]] = r
        time.sleep(0.01)
    assert len(results) == len(ids) + 1 and results[jid]["status"] == 0
    with pytest.raises(D.DqmcError):
        srv.submit(p, 5, 24, 8, dynamic=True)   # dynamic jobs measure whole forth-and-back iterations
    dm, de = srv.dynamic_result(jid)
    with pytest.raises(D.DqmcError):
        srv.dynamic_result(ids[0])            # an equal-time job has no dynamic results
    srv.close()
    b = D.Batch(ctx, [p])
    b.fields_random(); b.init(); b.sweep(warm)
    eq_bins, dyn_bins = [], []
    for _ in range(meas // bs):
        b.clear_measure()
        b.sweep(bs, measure=True, dynamic=True)
        s, c = b.measure_accum(0)
        eq_bins.append(D.normalize_bin(s, c))
        a = b.dynamic_accum(0)
        dyn_bins.append(np.concatenate([[abs(a["sign"] / a["count"]), a["rho"] / a["sign"]],
                                        (a["gk"] / a["sign"]).T.ravel(), a["dos"] / a["sign"], a["dss"] / a["sign"]]))
    em, _ = D.bin_stats(eq_bins)
    assert np.allclose(results[jid]["mean"], em, rtol=1e-12, atol=1e-12)
    m, e = D.bin_stats(dyn_bins)
    ours = np.concatenate([[dm["sign"], dm["rho"]], dm["gk"].T.ravel(), dm["dos"], dm["dss"]])
    assert np.allclose(ours, m, rtol=1e-12, atol=1e-12)


def test_large_lattices_in_the_mix_match_the_batch_api(ctx):
    """A 16x16 and a 20x20 job (the two largest lattices of the serving mix, in lanes of their
    own) served among small jobs give the batch API's results exactly, in FP64."""
    print("[gpu-serving] 16x16 and 20x20 jobs in a mix", flush=True)
    big = [dict(id=1000, params=D.make_params(16, 40, 2.0, pace=10, checkerboard=True, delay=16, U=4.0, mu=-0.3,
                                              seed=1000), warmup=2, measure=4, bin_size=2),
           dict(id=1001, params=D.make_params(20, 24, 2.4, pace=10, model=D.ATTRACTIVE, checkerboard=True, delay=16,
                                              U=5.0, mu=0.2, seed=1001), warmup=2, measure=4, bin_size=2),
           dict(id=1002, params=D.make_params(20, 30, 3.0, pace=10, checkerboard=True, delay=16, U=6.0, mu=-0.5,
                                              seed=1002), warmup=2, measure=4, bin_size=2)]
    jobs = small_jobs(60, seed=9)[:8] + big
    summary, res = serve.run(ctx, jobs, rate=30.0, budget_bytes=8 << 30)
    assert summary["failed"] == 0
    for j in big:
        assert j["params"].precision == D.PREC_FP64
        mean, err = batch_reference(ctx, j)
        assert np.array_equal(np.array(res[j["id"]]["mean"]), mean), j["id"]
        assert np.array_equal(np.array(res[j["id"]]["error"]), err), j["id"]
        assert res[j["id"]]["chain_status"] == 0 and res[j["id"]]["sweeps"] == 6
    print(f"[gpu-serving] p50 {summary['p50_latency_s']:.2f} s p99 {summary['p99_latency_s']:.2f} s", flush=True)


def test_generated_jobs_are_fp64_by_default():
    assert all(j["params"].precision == D.PREC_FP64 for j in serve.generate_jobs(20))
