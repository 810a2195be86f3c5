"""Autotuner for the engine's kernel variants.

Each option of `emc_set_option` selects a kernel variant or launch shape (rows in flight per
gather wave, orientation tiles per wave, frame ranges of the partitioned tomogram, ...); the
results agree to the parity tolerance whatever the choice, only the speed changes, and the
best choice depends on the workload (photons per frame, frames, pixels, orientations). The
tuner times the candidates on the live workload: one `maximize` per candidate (scale factors
not updated), reading the engine's HIP-event stage times for the stages the option affects,
and keeps the fastest value before moving to the next option (greedy, in the order given).
"""

# (option, candidate values (the first is the engine default), stages whose time decides).
# lik_split is left out by default: at the 100k-frame / 50,100-orientation workload it lost
# (8 ranges 561.6 ms, 4 ranges 450.4 ms, off 401.3 ms) and each value re-chunks the
# orientations; add ('lik_split', (1, 8, 4), ('likelihood',)) to try it elsewhere.
PLAIN_SPACE = (
    ('lik_xcd', (1, 2, 0), ('likelihood',)),
    ('lik_unroll', (8, 16, 4), ('likelihood',)),
    ('lik_tiles', (1, 2), ('likelihood',)),
    ('tomo_unroll', (16, 8), ('tomogram',)),
    ('tomo_parts', (8, 16, 4), ('tomogram', 'reduce')),
    ('tomo_width32', (0, 1), ('tomogram', 'reduce')),
)


def tune(engine, iteration=1, need_scaling=0, known_scale=0, space=PLAIN_SPACE, log=None):
    """Tune `engine` in place. Returns {option: chosen value} and the timings
    {option: {value: ms}}. Every candidate runs one full maximize (model2, inter_weight and
    the per-frame outputs are overwritten; scale factors are left alone)."""
    chosen, timings = {}, {}
    for name, values, stages in space:
        times = {}
        for v in values:
            engine.set_option(name, v)
            st = engine.maximize(iteration, need_scaling, 0, known_scale)
            times[v] = sum(st['stage_ms'][s] for s in stages)
        best = min(times, key=times.get)
        engine.set_option(name, best)
        chosen[name] = best
        timings[name] = times
        if log:
            log('autotune %s: %s -> %s' % (name, ', '.join('%s: %.1f ms' % kv for kv in times.items()), best))
    return chosen, timings
