"""The reference's example run (example/config.toml: 8 x 8 attractive Hubbard, U = 4, mu = 0,
beta = 8, 160 slices, dense, pace 10, equal-time and dynamic observables) repeated on the GPU
with independent chains, compared with the reference's published results
(example/output/{dos,dss,greens}.out and log.out, packed in
golden/example_observables_L160_N64.npz) within the combined statistical errors.

The GPU run: 64 chains (seeds 0..63), each thermalized with 300 sweeps and then measured over
one bin of 20 forth-and-back iterations (a time-displaced sweep with dynamic measurements,
then an updating backward sweep with equal-time measurements, src/dqmc.cpp:48-70); the chains
are the bins.  A Monte Carlo comparison, so the bar is statistical: |z| < 5 for the scalars
and, for D(tau), DSS(tau) and G(k,tau), at most 5% of the points beyond |z| = 3 and none
beyond |z| = 6, z = (ours - reference) / sqrt(err_ours^2 + err_reference^2).
"""
import os

import numpy as np
import pytest

import dqmc_amd as D

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EQ_NAMES = ["filling_number", "double_occupancy", "kinetic_energy", "momentum_distribution", "local_spin_corr",
            "spin_density_structure_factor", "charge_density_structure_factor", "s_wave_pairing_corr"]


def test_example_run_within_error_bars():
    ref = np.load(os.path.join(HERE, "golden", "example_observables_L160_N64.npz"))
    chains, warm, fb = 64, 300, 20
    ctx = D.Context(0)
    ps = [D.make_params(8, 160, 8.0, pace=10, model=D.ATTRACTIVE, U=4.0, mu=0.0, seed=k) for k in range(chains)]
    b = D.Batch(ctx, ps)
    b.fields_random()
    b.init()
    print(f"[gpu-example] thermalize {warm} sweeps x {chains} chains", flush=True)
    b.sweep(warm)
    b.clear_measure()
    print(f"[gpu-example] measure {fb} forth-and-back iterations", flush=True)
    b.sweep(2 * fb, measure=True, dynamic=True)
    eq_bins, dyn = [], {"dos": [], "dss": [], "gk": [], "rho": []}
    for c in range(chains):
        s, n = b.measure_accum(c)
        eq_bins.append(D.normalize_bin(s, n))
        a = b.dynamic_accum(c)
        assert a["count"] == fb
        for key in dyn:
            dyn[key].append(a[key] / a["sign"])     # sum / (count <sign>)
    print(f"[gpu-example] max wrap error {max(b.info(c).wrap_error for c in range(chains)):.2e}", flush=True)
    eq_m, eq_e = D.bin_stats(eq_bins)
    names = list(ref["scalar_names"])
    worst = []
    for o, name in enumerate(EQ_NAMES, start=1):
        rm, re_ = ref["scalars"][names.index(name)]
        z = (eq_m[o] - rm) / np.hypot(eq_e[o], re_)
        worst.append((abs(z), name))
        print(f"[gpu-example] {name}: {eq_m[o]:.6f} +- {eq_e[o]:.6f} vs {rm:.6f} +- {re_:.6f} (z {z:+.2f})", flush=True)
    rm, re_ = ref["scalars"][names.index("superfluid_stiffness")]
    m, e = D.bin_stats(np.array(dyn["rho"])[:, None])
    z = (m[0] - rm) / np.hypot(e[0], re_)
    worst.append((abs(z), "superfluid_stiffness"))
    print(f"[gpu-example] superfluid_stiffness: {m[0]:.6f} +- {e[0]:.6f} vs {rm:.6f} +- {re_:.6f} (z {z:+.2f})", flush=True)
    assert max(worst)[0] < 5.0, max(worst)
    for key, rk in (("dos", "dos"), ("dss", "dss"), ("gk", "greens")):
        m, e = D.bin_stats(np.array(dyn[key]))
        z = (m - ref[rk][0]) / np.hypot(e, ref[rk][1])
        frac3 = float(np.mean(np.abs(z) > 3))
        print(f"[gpu-example] {key}: max|z| {np.max(np.abs(z)):.2f}, beyond 3: {100 * frac3:.1f}%, "
              f"tau=0 {m.ravel()[0]:.6f} vs {ref[rk][0].ravel()[0]:.6f}", flush=True)
        assert np.max(np.abs(z)) < 6.0, key
        assert frac3 <= 0.05, key
    b.close()
    ctx.close()
