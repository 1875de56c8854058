"""Warm restart from the reference's field file (DqmcIO::output_bosonic_fields /
read_bosonic_fields_from_file, the `--fields` input of src/dqmc_main.cpp:148-160) plus the
random engine's state: a chain stopped at slice 0, written out and resumed in a fresh batch
continues the same Markov chain (fields and accept bits identical, G within the fresh-G bar)."""
import numpy as np
import pytest

import dqmc_amd as D

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cb,model", [(False, D.REPULSIVE), (True, D.ATTRACTIVE)])
def test_resume_from_field_file_continues_the_chain(tmp_path, cb, model):
    print(f"[gpu-checkpoint] cb={cb} model={model}", flush=True)
    ctx = D.Context(0)
    p = D.make_params(6, 40, 3.0, pace=10, model=model, checkerboard=cb, U=4.0, mu=-0.1, seed=11)
    run = D.Batch(ctx, [p], record=True)
    run.fields_random(); run.init()
    run.sweep(6)                                   # three forth-and-back iterations: back at slice 0
    path = str(tmp_path / "fields.out")
    D.write_fields(path, run.fields(0))
    rng = run.info(0).rng
    run.sweep(4, record=True)
    resumed = D.Batch(ctx, [p], record=True)
    resumed.set_fields(0, D.read_fields(path, 40, 36))
    resumed.set_rng_state(0, rng)
    resumed.init()
    resumed.sweep(4, record=True)
    assert np.array_equal(run.fields(0), resumed.fields(0))
    assert np.array_equal(run.accepts(0), resumed.accepts(0))
    assert run.info(0).rng == resumed.info(0).rng and run.info(0).sign == resumed.info(0).sign
    for spin in (1, -1):
        assert np.max(np.abs(run.green(0, spin) - resumed.green(0, spin))) < 1e-10
    ctx.close()
