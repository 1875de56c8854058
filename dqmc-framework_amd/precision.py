"""Error of the FP32 product modes against FP64 on live chains.

At every stabilization of a chain's upward sweeps the recomputed G of the chain (from its own
stacks, whose products may be FP32) is compared with an FP64 recompute of the exact field
configuration the chain had at that moment: rows up to the slice from after the sweep (already
updated), the rest from before.  The FP64 recomputes run as one batch of frozen sweeps, one chain
per configuration.  Returns max|dG| of every (stabilization, spin).
"""
import numpy as np

import dqmc_amd as D


def stabilization_errors(ctx, params, sweeps=4):
    """params: the chains (any precision) to follow; sweeps: single-direction sweeps run from a
    random start, of which the upward ones (0, 2, ...) are checked."""
    r = D.Batch(ctx, params, record=True)
    r.fields_random()
    r.init()
    out = []
    for s in range(sweeps):
        before = [r.fields(k).copy() for k in range(r.n)]
        r.sweep(1, record=True)
        after = [r.fields(k).copy() for k in range(r.n)]
        if s % 2 == 1:
            continue   # a backward sweep keeps the post-update G of each slot, not the recomputed one
        cfgs = []
        for k, p in enumerate(params):
            L, pace = p.time_size, p.stabilization_pace
            for t in (t for t in range(1, L + 1) if t % pace == 0 or t == L):
                h = before[k].copy()
                h[:t] = after[k][:t]
                cfgs.append((k, t, h))
        ref_params = []
        for k, _, _ in cfgs:
            q = D.ChainParams.from_buffer_copy(params[k])
            q.precision = D.PREC_FP64
            ref_params.append(q)
        ref = D.Batch(ctx, ref_params, record=True)
        for q, (k, t, h) in enumerate(cfgs):
            ref.set_fields(q, h)
        ref.init()
        ref.sweep(1, record=True, frozen=True)
        for q, (k, t, h) in enumerate(cfgs):
            for spin in (1, -1):
                out.append(float(np.max(np.abs(r.slice_green(k, t - 1, spin) - ref.slice_green(q, t - 1, spin)))))
        ref.close()
    r.close()
    return np.array(out)
