"""Autotuning of one system's launch geometry.

The engine's own brick chooser is a cost model (`choose_brick`: the expected own particles of a brick fill
one round of a workgroup's lanes, among those shapes the fewest staged cells).  `autotune_system` checks it
on the system itself: it times short graph-replayed runs of candidate configurations (cell size as
cutoff / cell_radius, brick shape in cells) and returns the fastest.  Every candidate gives the same
trajectory bit for bit (fixed-point force sums, draws keyed by handle and step), so a choice changes the
speed only; `test_autotune_candidates_give_one_trajectory` checks that.
"""
import time

import numpy as np


def brick_candidates(density, cutoff, lanes=256):
    """Cubes around the one-round brick (about `lanes` expected own particles) and the two-round cube, at
    cells of the cutoff; the engine's own choice ((0, 0, 0)); and cells of half the cutoff with the engine's
    brick (the 125-cell stencil of cell_radius 2)."""
    lam = density * cutoff ** 3  # expected particles per cell
    e1 = max(2, int(round((lanes / lam) ** (1.0 / 3.0))))
    e2 = max(2, int(round((2 * lanes / lam) ** (1.0 / 3.0))))
    shapes = {(0, 0, 0)}
    for e in (e1 - 1, e1, e1 + 1, e2):
        if e >= 2:
            shapes.add((e, e, e))
    cands = [{"cell_radius": 1, "brick": list(b)} for b in sorted(shapes)]
    cands.append({"cell_radius": 2, "brick": [0, 0, 0]})
    return cands


def autotune_system(make_engine, handler, dt, candidates, steps=30, warmup=5, keep_states=False):
    """Times `warmup` + `steps` graph-replayed steps of each candidate.  make_engine(candidate) returns an
    (engine, replica) pair holding the initial state; the engine is closed afterwards.  Returns the fastest
    candidate (with its rate) and every result; with keep_states, each result also carries the final
    (positions, types, handles) of the candidate's run."""
    results = []
    for cand in candidates:
        eng, r = make_engine(cand)
        try:
            eng.run(handler, dt, warmup, graph=True)
            eng.synchronize()
            n0 = eng.n_particles(r)
            t0 = time.perf_counter()
            eng.run(handler, dt, steps, graph=True)  # returns synchronised
            elapsed = time.perf_counter() - t0
            res = dict(cand)
            res["particle_steps_per_s"] = n0 * steps / elapsed
            res["ms_per_step"] = 1e3 * elapsed / steps
            if keep_states:
                p, t, h = eng.get_particles(r)
                order = np.argsort(h)
                res["state"] = (p[order], t[order], h[order])
            results.append(res)
        finally:
            eng.close()
    best = max(results, key=lambda x: x["particle_steps_per_s"])
    return best, results
