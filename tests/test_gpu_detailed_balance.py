"""DetailedBalance on the device (rdy_react with RDY_HANDLER_DETAILED_BALANCE) against the oracle:
the reference's integration scenes (readdy/test/TestDetailedBalance.cpp:364-746) event for event, a
mixed chemistry with every kind of reversible pair over several integrated steps, and the Metropolis
ratio of isolated conversions.  Ids and types are compared exactly; positions to 1e-6 after integration
(FP32 pair forces), exactly where no integration step ran.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from readdy_amd import _lib as L
from readdy_amd.engine import Engine

from test_detailed_balance import (A, B, C, D, HR, NO_INTEGRATION, SCENE_OUTCOME, _isolated_pairs, conversion,
                                   enzymatic, fission, fusion, oracle_system, scene)

pytestmark = pytest.mark.gpu


def _compare(e, r, s, box, tol):
    gp, gt, gh = e.get_particles(r)
    op, ot, oid, _ = s.particles()
    order = np.argsort(oid)
    assert np.array_equal(gh, oid[order].astype(np.uint32))
    assert np.array_equal(gt, ot[order].astype(np.uint32))
    d = gp - op[order]
    d -= box * np.round(d / box)
    assert np.abs(d).max() <= tol


def _compare_records(e, r, s, tol=1e-12):
    g, o = e.reaction_records(r), s.reaction_records()
    assert len(g) == len(o)
    for f in ("reaction", "type", "educts", "products", "types_from"):
        assert np.array_equal(np.asarray(g[f]).astype(np.int64), np.asarray(o[f]).astype(np.int64)), f
    assert np.allclose(g["position"], o["position"], rtol=0, atol=tol)


@pytest.mark.parametrize("kind,accept", list(SCENE_OUTCOME))
def test_reference_scenes_on_gpu(kind, accept):
    print(f"scene {kind} accept={accept}", flush=True)
    pots, rxs, pos, ty = scene(kind, accept)
    e = Engine(4, [1., 1., 1., 0.], potentials=pots, reactions=rxs)
    r = e.add_replica([12.] * 3, (1, 1, 1), seed=3, max_particles=2048)
    e.set_particles(r, pos, ty)
    e.record_reactions(True)
    e.react(L.HANDLER_DETAILED_BALANCE, 0.1)
    s = oracle_system(4, [12.] * 3, pots, rxs, diffusion=[1, 1, 1, 0])
    s.add_particles(pos, ty, ids=np.arange(len(ty)))
    s.set_id_counter(len(ty))
    s.set_seed(3)
    s.nl_setup(s.max_cutoff())
    s.set_record_reactions(True)
    n = s.detailed_balance(0.1, NO_INTEGRATION)
    want_n, want_counts, want_max = SCENE_OUTCOME[(kind, accept)]
    assert n == want_n
    _, gt, gh = e.get_particles(r)
    assert np.bincount(gt, minlength=4)[:3].tolist() == want_counts and int(gh.max()) == want_max
    _compare(e, r, s, 12.0, 1e-12)
    assert np.array_equal(e.reaction_counts(r), s.reaction_counts()[:len(rxs)])
    _compare_records(e, r, s)


def _mixed_chemistry():
    pots = [(HR, a, b, [10., 1.]) for a in range(4) for b in range(a, 4)]
    rxs = [fusion(A, B, C, 5., 1.), fission(C, A, B, 3., 1.),          # FusionFission
           conversion(A, D, 1.), conversion(D, A, 2.),                 # ConversionConversion
           enzymatic(B, C, D, 4., 0.8), enzymatic(D, C, B, 2., 0.8),   # EnzymaticEnzymatic (C the catalyst)
           dict(kind=L.RX_DECAY, educts=[D], rate=0.1),                # not reversible
           fission(D, A, A, 0.2, 1.)]                                  # not reversible
    return pots, rxs


@pytest.mark.parametrize("lds", [0, 128])
def test_mixed_chemistry_steps_match_oracle_event_for_event(lds):
    """integrate, neighbour list, DetailedBalance, forces; five steps, the same Philox counters on both
    sides; lds 128: every brick searched by the bin-reading kernel"""
    pots, rxs = _mixed_chemistry()
    n, Lb, dt = 1500, 12.0, 0.01
    rng = np.random.default_rng(12)
    pos = rng.uniform(-Lb / 2, Lb / 2, (n, 3))
    ty = rng.integers(0, 4, n).astype(np.uint32)
    e = Engine(4, [1.] * 4, potentials=pots, reactions=rxs, lds_states=lds)
    r = e.add_replica([Lb] * 3, (1, 1, 1), seed=41, max_particles=4 * n)
    e.set_particles(r, pos, ty)
    e.record_reactions(True)
    s = oracle_system(4, [Lb] * 3, pots, rxs, diffusion=[1.] * 4)
    s.add_particles(pos, ty.astype(np.int32), ids=np.arange(n))
    s.set_id_counter(n)
    s.set_seed(41)
    s.nl_setup(s.max_cutoff())
    s.set_record_reactions(True)
    e.calculate_forces()
    s.calculate_forces()
    total = np.zeros(len(rxs), dtype=np.int64)
    for t in range(5):
        print(f"step {t}", flush=True)
        noise = np.random.default_rng(300 + t).normal(size=(6 * n, 3))
        e.integrate(dt, noise)
        e.update_neighbor_list()
        s.euler_bd(dt, noise.ravel(), step=t)
        s.nl_fill()
        e.react(L.HANDLER_DETAILED_BALANCE, dt)
        s.detailed_balance(dt, t)
        s.nl_fill()
        _compare(e, r, s, Lb, 1e-6)
        c = e.reaction_counts(r)
        assert np.array_equal(c, s.reaction_counts()[:len(rxs)])
        _compare_records(e, r, s, 1e-6)
        total += c.astype(np.int64)
        e.calculate_forces()
        s.calculate_forces()
    # every kind of pair moved in both directions, the others fired too
    assert (total[:6] > 0).all() and total[6:].sum() > 0, total


def test_metropolis_ratio_of_isolated_conversions_on_gpu():
    """A -> B next to a D at distance 1 with HR(B, D) k=2 d=2: acceptance e^-1, 1000 independent moves (one
    replica step resolves at most 1024 accepted events)"""
    n = 1000
    box, a, d = _isolated_pairs(n, 6.0, 1.0)
    e = Engine(4, [1.] * 4, potentials=[(HR, B, D, [2., 2.])], reactions=[conversion(A, B, 1e9), conversion(B, A, 1e-9)],
               events_per_replica=1024)
    r = e.add_replica([box] * 3, (1, 1, 1), seed=21, max_particles=3 * n)
    e.set_particles(r, np.vstack([a, d]), np.array([A] * n + [D] * n, dtype=np.uint32))
    e.react(L.HANDLER_DETAILED_BALANCE, 1.0)
    got = int(e.reaction_counts(r)[0])
    p = math.exp(-1.0)
    assert abs(got - n * p) < 5 * math.sqrt(n * p * (1 - p)), got
    s = oracle_system(4, [box] * 3, [(HR, B, D, [2., 2.])], [conversion(A, B, 1e9), conversion(B, A, 1e-9)])
    s.add_particles(np.vstack([a, d]), np.array([A] * n + [D] * n, dtype=np.int32), ids=np.arange(2 * n))
    s.set_id_counter(2 * n)
    s.set_seed(21)
    s.nl_setup(s.max_cutoff())
    assert s.detailed_balance(1.0, NO_INTEGRATION) == got
    _compare(e, r, s, box, 0.0)


def test_detailed_balance_is_not_a_step_handler():
    """the fused loop and the per-replica handler take UncontrolledApproximation and Gillespie only"""
    e = Engine(2, [1., 1.], reactions=[conversion(0, 1, 1.), conversion(1, 0, 1.)])
    r = e.add_replica([5.] * 3, (1, 1, 1), seed=1, max_particles=64)
    e.set_particles(r, np.zeros((1, 3)), np.zeros(1, dtype=np.uint32))
    with pytest.raises(L.EngineError, match="handler"):
        e.run(L.HANDLER_DETAILED_BALANCE, 0.1, 2)
    with pytest.raises(L.EngineError, match="handler"):
        e.set_handler(r, L.HANDLER_DETAILED_BALANCE)
    e.react(L.HANDLER_DETAILED_BALANCE, 0.1)
    assert e.n_particles(r) == 1


# the pair potentials of every kind in the Metropolis energies (LJ: m, n, cutoff, shift, epsilon, sigma; SE:
# C, kappa, D, sigma, exponent, cutoff; WIPH: k, d0, depth, cutoff), strengths depending on the type pair so
# that every move changes the energy, on a jittered lattice so no pair overlaps
POTENTIAL_CASES = {
    "weak_interaction": (L.POT_WEAK_INTERACTION, lambda s: [10., 1.0, 0.4 + 0.2 * s, 1.6]),
    "lennard_jones": (L.POT_LENNARD_JONES, lambda s: [12., 6., 1.6, 1., 0.3 + 0.1 * s, 0.8]),
    "lennard_jones_unshifted": (L.POT_LENNARD_JONES, lambda s: [12., 6., 1.6, 0., 0.3 + 0.1 * s, 0.8]),
    "screened_electrostatics": (L.POT_SCREENED_ELECTROSTATICS, lambda s: [-0.5 - 0.25 * s, 1.0, 0.5, 0.8, 6., 1.6]),
}


@pytest.mark.parametrize("case", list(POTENTIAL_CASES))
def test_every_potential_kind_in_the_metropolis_energy(case):
    kind, params = POTENTIAL_CASES[case]
    pots = [(kind, a, b, params(a + b)) for a in range(4) for b in range(a, 4)]
    rxs = [fusion(A, B, C, 50., 1.4), fission(C, A, B, 30., 1.4), conversion(A, D, 20.), conversion(D, A, 20.),
           enzymatic(B, C, D, 40., 1.4), enzymatic(D, C, B, 40., 1.4)]
    m, a0, Lb = 8, 1.25, 10.0
    g = np.array([(i, j, k) for i in range(m) for j in range(m) for k in range(m)], float) * a0 - Lb / 2 + 0.1
    rng = np.random.default_rng(3)
    pos = g + rng.uniform(-0.15, 0.15, g.shape)
    ty = rng.integers(0, 4, len(pos)).astype(np.uint32)
    n = len(pos)
    e = Engine(4, [1.] * 4, potentials=pots, reactions=rxs)
    r = e.add_replica([Lb] * 3, (1, 1, 1), seed=13, max_particles=4 * n)
    e.set_particles(r, pos, ty)
    e.record_reactions(True)
    e.react(L.HANDLER_DETAILED_BALANCE, 0.05)
    s = oracle_system(4, [Lb] * 3, pots, rxs)
    s.add_particles(pos, ty.astype(np.int32), ids=np.arange(n))
    s.set_id_counter(n)
    s.set_seed(13)
    s.nl_setup(s.max_cutoff())
    s.set_record_reactions(True)
    done = s.detailed_balance(0.05, NO_INTEGRATION)
    _compare(e, r, s, Lb, 1e-12)
    assert np.array_equal(e.reaction_counts(r), s.reaction_counts()[:len(rxs)])
    _compare_records(e, r, s)
    # some moves were rejected (ids of rejected moves' educts advanced) and some accepted
    _, _, ids, _ = s.particles()
    assert done > 10 and int(ids.max()) > n
