"""Ensemble statistics: the HIP engine against the FP64 oracle with independent random streams.

The dynamics are stochastic, so parity here is statistical: an estimate from an ensemble of GPU
replicas and one from an ensemble of oracle runs with other seeds must agree within 3 standard
errors (the combined standard error of the two ensemble means; each replica's time or particle
average is one independent sample).  Covered: the diffusion constant from the mean-squared
displacement (free and with harmonic repulsion), the radial distribution function of a repulsive
fluid, and the equilibrium complex count of A+B<->C under UncontrolledApproximation and Gillespie
(the GPU's keyed Gillespie against the oracle's literal performEvents algorithm).
"""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

L = pytest.importorskip("readdy_amd._lib")
from readdy_amd import Engine  # noqa: E402


def _within(a, b, k=3.0):
    """ensemble samples a, b: |mean a - mean b| < k combined standard errors"""
    a, b = np.asarray(a, float), np.asarray(b, float)
    se = np.sqrt(a.var(ddof=1) / len(a) + b.var(ddof=1) / len(b))
    return abs(a.mean() - b.mean()), se


def _hr(n_types):
    return [(L.POT_HARMONIC_REPULSION, a, b, [10.0, 1.0]) for a in range(n_types) for b in range(a, n_types)]


def _positions(n, box, seed):
    return (np.random.default_rng(seed).random((n, 3)) - 0.5) * box


# ---- mean-squared displacement -----------------------------------------------------------------
def _msd_gpu(n, box, steps, dt, pots, n_rep):
    e = Engine(1, [1.0], potentials=pots, max_replicas=n_rep)
    reps = [e.add_replica([box] * 3, seed=7000 + r, max_particles=2 * n) for r in range(n_rep)]
    for r in reps:
        e.set_particles(r, _positions(n, box, 100 + r), np.zeros(n, np.uint32))
    prev = [e.get_particles(r)[0] for r in reps]
    disp = [np.zeros((n, 3)) for _ in reps]
    for t in range(steps):
        e.run(L.HANDLER_NONE, dt, 1)
        for i, r in enumerate(reps):
            x = e.get_particles(r)[0]
            d = x - prev[i]
            disp[i] += d - box * np.round(d / box)
            prev[i] = x
    e.close()
    return [float((d ** 2).sum(1).mean()) / (6 * steps * dt) for d in disp]


def _msd_oracle(n, box, steps, dt, pots, n_rep):
    out = []
    for r in range(n_rep):
        o = po.System(1, [box] * 3, [1, 1, 1], diffusion=[1.0])
        for kind, a, b, p in pots:
            o.add_pair_potential(kind, a, b, p)
        o.set_seed(90000 + r)
        o.add_particles(_positions(n, box, 100 + r), np.zeros(n, np.int32), ids=np.arange(n))
        o.nl_setup(max(o.max_cutoff(), 1.0))
        o.calculate_forces()
        prev = o.particles()[0][np.argsort(o.particles()[2])]
        disp = np.zeros((n, 3))
        for t in range(steps):
            o.loop_step(dt, po.UA_KEYED, t)
            p, _, ids, _ = o.particles()
            x = p[np.argsort(ids)]
            d = x - prev
            disp += d - box * np.round(d / box)
            prev = x
        out.append(float((disp ** 2).sum(1).mean()) / (6 * steps * dt))
    return out


def test_free_diffusion_constant_from_msd():
    # closed form: <|x(t) - x(0)|^2> = 6 D t with D = 1
    g = np.asarray(_msd_gpu(1000, 12.0, 40, 0.01, [], 8))
    se_g = g.std(ddof=1) / np.sqrt(len(g))
    assert abs(g.mean() - 1.0) < 3 * se_g, (g.mean(), se_g)


def test_diffusion_constant_with_repulsion_matches_oracle():
    pots = _hr(1)
    g = _msd_gpu(500, 12.6, 60, 0.01, pots, 8)  # density 0.25
    o = _msd_oracle(500, 12.6, 60, 0.01, pots, 8)
    delta, se = _within(g, o)
    assert delta < 3 * se, (np.mean(g), np.mean(o), se)
    assert np.mean(g) < 0.99  # repulsion slows diffusion measurably below D = 1


# ---- radial distribution function ----------------------------------------------------------------
def _rdf(x, box, edges):
    d = x[:, None, :] - x[None, :, :]
    d -= box * np.round(d / box)
    r = np.sqrt((d ** 2).sum(-1))[np.triu_indices(len(x), 1)]
    h, _ = np.histogram(r, edges)
    shell = 4.0 / 3.0 * np.pi * (edges[1:] ** 3 - edges[:-1] ** 3)
    rho = len(x) / box ** 3
    return h / (0.5 * len(x) * rho * shell)


def test_rdf_of_repulsive_fluid_matches_oracle():
    n, box, dt, burn, samples, every, n_rep = 400, 9.3, 0.01, 200, 8, 25, 6  # density 0.5
    edges = np.linspace(0.0, 2.5, 26)
    pots = _hr(1)
    e = Engine(1, [1.0], potentials=pots, max_replicas=n_rep)
    reps = [e.add_replica([box] * 3, seed=500 + r, max_particles=2 * n) for r in range(n_rep)]
    for r in reps:
        e.set_particles(r, _positions(n, box, 40 + r), np.zeros(n, np.uint32))
    e.run(L.HANDLER_NONE, dt, burn)
    g_gpu = np.zeros((n_rep, len(edges) - 1))
    for s in range(samples):
        e.run(L.HANDLER_NONE, dt, every)
        for i, r in enumerate(reps):
            g_gpu[i] += _rdf(e.get_particles(r)[0], box, edges) / samples
    e.close()
    g_or = np.zeros((n_rep, len(edges) - 1))
    for i in range(n_rep):
        o = po.System(1, [box] * 3, [1, 1, 1], diffusion=[1.0])
        for kind, a, b, p in pots:
            o.add_pair_potential(kind, a, b, p)
        o.set_seed(31000 + i)
        o.add_particles(_positions(n, box, 40 + i), np.zeros(n, np.int32), ids=np.arange(n))
        o.nl_setup(o.max_cutoff())
        o.calculate_forces()
        t = 0
        for _ in range(burn):
            o.loop_step(dt, po.UA_KEYED, t)
            t += 1
        for s in range(samples):
            for _ in range(every):
                o.loop_step(dt, po.UA_KEYED, t)
                t += 1
            g_or[i] += _rdf(o.particles()[0], box, edges) / samples
    se = np.sqrt(g_gpu.var(0, ddof=1) / n_rep + g_or.var(0, ddof=1) / n_rep) + 1e-9
    z = np.abs(g_gpu.mean(0) - g_or.mean(0)) / se
    informative = (g_gpu.mean(0) + g_or.mean(0)) > 0.05  # bins inside the core are empty on both sides
    assert np.all(z[informative] < 4.0) and np.sum(z[informative] > 3.0) <= 2, (z, g_gpu.mean(0), g_or.mean(0))
    # the structure is real: depletion inside d = 1, contact peak above 1
    gm = g_gpu.mean(0)
    assert gm[:5].mean() < 0.5 and gm[10:14].max() > 1.0


# ---- equilibrium A+B <-> C -----------------------------------------------------------------------
def _eq_system():
    pots = _hr(3)
    rxs = [dict(kind=L.RX_FUSION, educts=[0, 1], products=[2], rate=1.0, educt_distance=1.0),
           dict(kind=L.RX_FISSION, educts=[2], products=[0, 1], rate=0.5, product_distance=1.0)]
    return pots, rxs


@pytest.mark.parametrize("handler,mode", [(0, po.UA_KEYED), (1, po.GILLESPIE_LITERAL)])
def test_equilibrium_complex_count_matches_oracle(handler, mode):
    # one replica's time average has a standard deviation of about 1.6 complexes at <C> ~ 30
    n_half, box, dt, burn, steps, every, n_rep, n_or = 100, 10.0, 0.01, 400, 1200, 10, 64, 32
    pots, rxs = _eq_system()
    init = [(_positions(2 * n_half, box, 300 + r), np.repeat(np.arange(2, dtype=np.uint32), n_half))
            for r in range(n_rep)]
    e = Engine(3, [1.0, 1.0, 0.5], potentials=pots, reactions=rxs, max_replicas=n_rep)
    reps = [e.add_replica([box] * 3, seed=800 + r, max_particles=4 * n_half) for r in range(n_rep)]
    for r, (x, t) in zip(reps, init):
        e.set_particles(r, x, t)
    e.run(handler, dt, burn)
    c_gpu = np.zeros(n_rep)
    for s in range(steps // every):
        e.run(handler, dt, every)
        c_gpu += e.count_types(n_rep)[:, 2] / (steps // every)
    e.close()
    c_or = np.zeros(n_or)
    for i, (x, t) in enumerate(init[:n_or]):
        o = po.System(3, [box] * 3, [1, 1, 1], diffusion=[1.0, 1.0, 0.5])
        for kind, a, b, p in pots:
            o.add_pair_potential(kind, a, b, p)
        for r in rxs:
            o.add_reaction(r["kind"], r["educts"], r["products"], r["rate"], r.get("educt_distance", 0.0),
                           r.get("product_distance", 0.0))
        o.set_seed(55000 + i)
        o.add_particles(x, t.astype(np.int32), ids=np.arange(len(t)))
        o.nl_setup(o.max_cutoff())
        o.calculate_forces()
        for step in range(burn + steps):
            o.loop_step(dt, mode, step)
            if step >= burn and (step - burn + 1) % every == 0:
                c_or[i] += (o.particles()[1] == 2).sum() / (steps // every)
    delta, se = _within(c_gpu, c_or)
    assert delta < 3 * se, (c_gpu.mean(), c_or.mean(), se)
    assert c_gpu.mean() > 5  # the reaction equilibrium is populated
