"""Bead chains + crowders with enzymatic binding under Gillespie (the heavy topology configuration).

Reduced size (12 chains of 10 beads, 300 crowders, box 12): five steps of integrate / NL / react / NL
/ forces with injected noise match the FP64 oracle event for event (same events, same products,
positions within the FP32-pair-force bound), and the fused run path reproduces the separate actions
bit for bit (the fix-up after bindings touches beads, whose moves carry bonded forces).  Full size
(10^4 chains of 50 beads, 5 10^5 crowders, box 160): the forces of the initial state against the
oracle, then invariants of a run (species counts conserved where the chemistry conserves them,
positions in the box, energy and forces finite).

The configuration's own time step, dt = 1e-3, is numerically unstable for these chains: where three
beads pass through a line the cosine-dihedral force grows as 1/sin(theta) up to the reference's clamps
(TorsionPotential.cpp:53-159), a bead jumps by F D dt, bonds and crowder overlaps follow, and the FP64
oracle itself reaches forces of 1e18 within two steps at full size (1e7 at 5e-4; stable at 2e-4 for 20
steps).  The device's force sums are 64-bit fixed point, so it stops such a run with RDY_ERR_NUMERIC
(a force beyond 1.68e7) instead of carrying inf or wrapped sums on; the runs below use dt = 2e-4, and
one test shows that at dt = 1e-3 the device reports the blow-up that the oracle shows from the same state.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

from readdy_amd import workloads as C2

pytestmark = pytest.mark.gpu

L = pytest.importorskip("readdy_amd._lib")
from readdy_amd import Engine  # noqa: E402

DIFF = [0.5, 1.0, 1.0]


def _gpu(pos, types, bonds, angles, dihedrals, box, rate, seed=11, **kw):
    e = Engine(3, DIFF, potentials=C2.potentials(L), reactions=C2.reactions(L, rate), topology_flavor=[1, 0, 0],
               **kw)
    r = e.add_replica([box] * 3, (1, 1, 1), seed=seed, max_particles=2 * len(types))
    e.set_particles(r, pos, types)
    e.set_bonds(r, bonds, 100.0, 1.0)
    e.set_angles(r, angles, 10.0, np.pi)
    e.set_dihedrals(r, dihedrals, 1.0, 3, 0.0)
    return e, r


def _oracle(pos, types, bonds, angles, dihedrals, box, rate, seed):
    o = po.System(3, [box] * 3, [1, 1, 1], diffusion=DIFF)
    for kind, a, b, p in C2.potentials(L):
        o.add_pair_potential(kind, a, b, p)
    for x in C2.reactions(L, rate):
        o.add_reaction(x["kind"], x["educts"], x["products"], x["rate"], x["educt_distance"], 0.0)
    o.set_seed(seed)
    o.add_particles(pos, types.astype(np.int32), ids=np.arange(len(types)))
    for b in bonds:
        o.add_bond(int(b[0]), int(b[1]), 100.0, 1.0)
    for a in angles:
        o.add_angle(int(a[0]), int(a[1]), int(a[2]), 10.0, np.pi)
    for d in dihedrals:
        o.add_dihedral(int(d[0]), int(d[1]), int(d[2]), int(d[3]), 1.0, 3, 0.0)
    o.nl_setup(o.max_cutoff())
    return o


def test_reduced_chains_match_oracle_event_for_event():
    # dt 2e-4: near contact the Lennard-Jones curvature makes D dt U'' ~ 0.2, so the FP32-vs-FP64 force
    # difference is amplified by ~1.2 per step (at dt 1e-3 it would double every step)
    box, dt, rate = 12.0, 2e-4, 600.0
    pos, types, bonds, angles, dihedrals = C2.system(12, 10, 300, box, seed=1)
    n = len(types)
    e, r = _gpu(pos, types, bonds, angles, dihedrals, box, rate, seed=99)
    o = _oracle(pos, types, bonds, angles, dihedrals, box, rate, 99)
    e.calculate_forces()
    o.calculate_forces()
    f_gpu, _ = e.get_forces(r)
    f_ref = o.particles()[3][np.argsort(o.particles()[2])]
    assert np.linalg.norm(f_gpu - f_ref) / np.linalg.norm(f_ref) < 1e-5
    total = 0
    for t in range(5):
        noise = np.random.default_rng(200 + t).normal(size=(2 * n, 3))
        e.integrate(dt, noise)
        e.update_neighbor_list()
        o.euler_bd(dt, noise.ravel(), step=t)
        o.nl_fill()
        e.react(1, dt)
        o.reactions(po.GILLESPIE_KEYED, dt, step=t)
        o.nl_fill()
        gp, gt, gh = e.get_particles(r)
        op, ot, oid, _ = o.particles()
        order = np.argsort(oid)
        assert np.array_equal(gh, oid[order].astype(np.uint32))
        assert np.array_equal(gt, ot[order].astype(np.uint32))
        d = gp - op[order]
        d -= box * np.round(d / box)
        # FP32 Lennard-Jones (12th powers: rel. error ~1e-6) on forces up to ~1e3 moves positions by
        # D dt dF ~ 1e-7 per step
        assert np.abs(d).max() < 1e-5
        assert np.array_equal(e.reaction_counts(r), o.reaction_counts())
        total += int(e.reaction_counts(r).sum())
        e.calculate_forces()
        o.calculate_forces()
    assert total > 10


def test_reduced_chains_fused_run_equals_separate_actions():
    box, dt, rate, steps = 12.0, 2e-4, 250.0, 20
    sysm = C2.system(12, 10, 300, box, seed=3)
    e1, r1 = _gpu(*sysm, box, rate)
    e1.run(1, dt, steps)
    e2, r2 = _gpu(*sysm, box, rate)
    e2.calculate_forces()
    for _ in range(steps):
        e2.integrate(dt)
        e2.update_neighbor_list()
        e2.react(1, dt)
        e2.update_neighbor_list()
        e2.calculate_forces()
    for x, y in zip(e1.get_particles(r1), e2.get_particles(r2)):
        assert np.array_equal(x, y)
    assert e1.energy(r1) == e2.energy(r2)
    assert np.array_equal(e1.reaction_totals(r1), e2.reaction_totals(r2))
    assert e1.reaction_totals(r1).sum() > 10


def test_full_size_forces_and_invariants():
    box, dt = 160.0, 2e-4
    print("building the 10^6-particle chain system", flush=True)
    pos, types, bonds, angles, dihedrals = C2.system(10000, 50, 500000, box, seed=2)
    n_beads = int((types == 0).sum())
    e, r = _gpu(pos, types, bonds, angles, dihedrals, box, 0.01, seed=2)
    # forces of the initial state against the FP64 oracle, at full size
    e.calculate_forces()
    f_gpu, _ = e.get_forces(r)
    print("oracle forces", flush=True)
    o = _oracle(pos, types, bonds, angles, dihedrals, box, 0.01, 2)
    o.calculate_forces()
    f_ref = o.particles()[3][np.argsort(o.particles()[2])]
    del o
    assert np.linalg.norm(f_gpu - f_ref) / np.linalg.norm(f_ref) < 1e-5
    print("running", flush=True)
    e.run(1, dt, 20, graph=True)
    p, t, h = e.get_particles(r)
    assert int((t == 0).sum()) == n_beads  # the catalyst is never consumed
    assert int((t == 1).sum() + (t == 2).sum()) == len(types) - n_beads
    assert np.all(np.abs(p) <= box / 2)
    assert np.isfinite(e.energy(r))
    f, _ = e.get_forces(r)
    assert np.all(np.isfinite(f))


def test_full_size_at_the_configured_time_step_reports_the_blow_up():
    """dt = 1e-3: the FP64 oracle's forces pass 1.68e7 within five steps from the initial state, and the
    device's run of those steps stops with RDY_ERR_NUMERIC (every kernel ends normally)."""
    box, dt, steps = 160.0, 1e-3, 5
    pos, types, bonds, angles, dihedrals = C2.system(10000, 50, 500000, box, seed=2)
    o = _oracle(pos, types, bonds, angles, dihedrals, box, 0.01, 2)
    o.set_threads(max(1, min(16, len(__import__("os").sched_getaffinity(0)))))
    o.calculate_forces()
    peak = 0.0
    for t in range(steps):
        o.loop_step(dt, po.GILLESPIE_KEYED, t)
        peak = max(peak, float(np.abs(o.particles()[3]).max()))
    print(f"oracle: largest force within {steps} steps {peak:.2e}", flush=True)
    assert peak > 1.68e7
    del o
    e, r = _gpu(pos, types, bonds, angles, dihedrals, box, 0.01, seed=2)
    with pytest.raises(L.EngineError, match="error -5"):
        e.run(1, dt, steps, graph=True)
