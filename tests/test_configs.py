"""The batched-ensemble and serving-stream configurations at their own sizes.

Batched ensemble: all 4096 replicas of the A+B<->C system, particle counts log-uniform in [500, 50000] at
density 0.25 (about 4.3e7 particles in one engine), 20 graph-replayed steps.  Every replica keeps its A + C
and B + C counts (fusion A+B->C and fission C->A+B conserve both), every sampled particle lies in its box,
energies are finite, and 32 replicas spread through the sample end bit for bit where they end when each
is run alone in an engine of its own (handle-keyed Philox draws, fixed-point sums).

Serving stream: the first four requests of the Poisson(200/s) trace with particle counts log-uniform in
[1e3, 2e5] and step counts log-uniform in [1e3, 1e5] (not scaled), handlers 50/50, served with two slots so
that requests wait for, and then share, the launches of others; each request's final state and reaction
totals equal those of the same request run alone.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = pytest.importorskip("readdy_amd._lib")
from readdy_amd import Engine  # noqa: E402
from readdy_amd import ensemble as E  # noqa: E402
from readdy_amd import serving as S  # noqa: E402

STEPS = 20


def test_batched_ensemble_at_full_size():
    n, box = E.sample(4096)
    print(f"building 4096 replicas, {int(n.sum())} particles", flush=True)
    b = E.Batched(n, box, nl_cutoff=2.0)
    eng = b.engine
    c0 = eng.count_types(len(b.replicas)).astype(np.int64)
    assert np.array_equal(c0.sum(axis=1), n)
    print(f"running {STEPS} graph-replayed steps", flush=True)
    b.run(STEPS)
    c1 = eng.count_types(len(b.replicas)).astype(np.int64)
    assert np.array_equal(c1[:, 0] + c1[:, 2], c0[:, 0] + c0[:, 2])
    assert np.array_equal(c1[:, 1] + c1[:, 2], c0[:, 1] + c0[:, 2])
    assert c1[:, 2].sum() > 1000  # the reactions ran
    sample = list(range(0, 4096, 128))
    for r in sample:
        p, t, h = eng.get_particles(r)
        assert len(t) == c1[r].sum() and np.all(np.abs(p) <= box[r] / 2) and np.isfinite(eng.energy(r))
    print("standalone runs of 32 replicas", flush=True)
    pots, rxs = E.chemistry()
    for r in sample:
        a = Engine(3, E.DIFFUSION, potentials=pots, reactions=rxs, nl_cutoff=2.0)
        ra = a.add_replica([box[r]] * 3, (1, 1, 1), seed=(0x5EED << 32) | r, max_particles=2 * int(n[r]))
        a.set_particles(ra, *E.initial_state(n[r], box[r], r))
        a.run(L.HANDLER_UA, E.DT, STEPS, graph=True)
        for x, y in zip(a.get_particles(ra), eng.get_particles(r)):
            assert np.array_equal(x, y), r
        assert np.array_equal(a.reaction_totals(ra), eng.reaction_totals(r)), r
        a.close()
    b.close()


def test_served_stream_at_its_sizes_equals_alone():
    trace = S.poisson_trace(4, rate_per_s=200.0, seed=8)  # n in [1e3, 2e5], steps in [1e3, 1e5], unscaled
    print("requests (particles, steps, handler):", [(r.n, r.steps, r.handler) for r in trace], flush=True)
    assert max(r.n for r in trace) > 50_000 and max(r.steps for r in trace) > 5_000
    assert {r.handler for r in trace} == {"UncontrolledApproximation", "Gillespie"}
    srv = S.Server(hbm_budget_bytes=8 << 30, max_slots=2, interval_steps=500)
    results, summary = srv.run(trace, keep_final=True)
    srv.close()
    assert summary["requests"] == len(trace)
    for req in trace:
        print(f"request {req.rid} alone", flush=True)
        (xa, ta, ha), totals = S.run_alone(req)
        xm, tm, hm = results[req.rid].final
        assert np.array_equal(ha, hm) and np.array_equal(ta, tm) and np.array_equal(xa, xm), req
        assert np.array_equal(totals, results[req.rid].observed_reactions[-1]), req
