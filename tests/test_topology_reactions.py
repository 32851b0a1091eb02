"""Spatial topology reactions of the topology-particle enzymatic kind, "T1(p1) + (p2) -> T2(p3) + (p4)"
(SpatialTopologyReaction.h:64-236, CPUEvaluateTopologyReactions.cpp:88-475), on the device against the
oracle: the deterministic scenes the oracle is pinned by (rate 1e10: one reaction per topology, one per
free particle, the strict radius gate, the topology's type), event for event on random chain systems
(same Philox counters on both sides), bit for bit between the fused loop and the separate actions, and
the bead-crowder binding of the bead-chain configuration as ReaDDy models it (a topology-particle
reaction instead of a plain enzymatic one), reduced, against the oracle.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

L = pytest.importorskip("readdy_amd._lib")
from readdy_amd import Engine  # noqa: E402
from readdy_amd import workloads as W  # noqa: E402

# species 0 bead (topology), 1 X_free, 2 X_bound, 3 bead' (topology); T0(0) + (1) -> T1(to1) + (2)
FLAVOR = [1, 0, 0, 1]


def _engine(box, rate=1e10, radius=1.2, to1=0, pots=(), n=64, seed=5, nl_cutoff=None):
    e = Engine(4, [1.0, 1.0, 1.0, 1.0], potentials=list(pots), topology_flavor=FLAVOR,
               nl_cutoff=nl_cutoff or max(radius, 1.0))
    r = e.add_replica([box] * 3, (1, 1, 1), seed=seed, max_particles=4 * n + 64)
    return e, r


def _sr(rate=1e10, radius=1.2, to1=0):
    return [dict(type1=0, top_type1=0, type2=1, type_to1=to1, top_type_to1=1, type_to2=2, rate=rate, radius=radius)]


def test_scene_matches_the_oracle():
    pos = np.array([[0, 0, 0], [1, 0, 0], [6, 0, 0], [0.5, 0.5, 0], [6.5, 0, 0], [0, 3, 0], [6, 0.5, 0]], float)
    types = np.array([0, 0, 0, 1, 1, 1, 2], np.uint32)
    e, r = _engine(20.0)
    e.set_particles(r, pos, types)
    e.set_spatial_reactions(_sr())
    e.set_topologies(r, [0, 0], [0, 0, 1])
    e.evaluate_topology_reactions(1.0)
    _, t, h = e.get_particles(r)
    assert list(t) == [0, 0, 0, 2, 2, 1, 2]
    assert list(e.topology_types(r)) == [1, 1] and list(e.topology_reaction_counts(r)) == [2]
    e.evaluate_topology_reactions(1.0)
    assert list(e.topology_reaction_counts(r)) == [0] and list(e.topology_reaction_counts(r, totals=True)) == [2]


def test_reacted_topologies_and_topology_particle_species():
    """the topologies a spatial evaluation changed (the structural reactions of the same evaluation skip them),
    only for that evaluation; topology particles take new species by handle, their topology flag from the
    species' flavour (a normal species takes a particle out of the topology reactions)"""
    pos = np.array([[0, 0, 0], [1, 0, 0], [6, 0, 0], [0.5, 0.5, 0], [0, 3, 0]], float)
    e, r = _engine(20.0)
    e.set_particles(r, pos, np.array([0, 0, 0, 1, 1], np.uint32))
    e.set_spatial_reactions(_sr())
    e.set_topologies(r, [0, 0, 0], [0, 1, 2])
    e.evaluate_topology_reactions(1.0)
    reacted = e.topology_reacted(r)
    assert reacted.sum() == 1 and list(e.topology_types(r)[reacted]) == [1]
    e.evaluate_topology_reactions(1.0)  # nothing left to react
    assert not e.topology_reacted(r).any()
    e.set_particle_types(r, [2], [3])  # a topology species
    e.set_particle_types(r, [1], [1])  # a normal species: now a free partner for topology 0 or 2
    _, t, _ = e.get_particles(r)
    assert t[2] == 3 and t[1] == 1
    with pytest.raises(L.EngineError, match="topology particles"):
        e.set_particle_types(r, [3], [0])  # a free particle
    with pytest.raises(L.EngineError, match="type out of range"):
        e.set_particle_types(r, [0], [9])


def test_shared_free_particle_and_strict_gate():
    e, r = _engine(20.0, radius=1.25, nl_cutoff=1.5)
    # topology particles first (handles 0-2), then the free ones (3: shared by topologies 0 and 1, 4: exactly
    # 1.25 from the bead of topology 2)
    e.set_particles(r, [[0, 0, 0], [1.5, 0, 0], [5, 5, 5], [0.75, 0, 0], [6.25, 5, 5]], [0, 0, 0, 1, 1])
    e.set_spatial_reactions(_sr(radius=1.25, to1=3))
    e.set_topologies(r, [0, 0, 0], [0, 1, 2])
    e.evaluate_topology_reactions(1.0)
    _, t, _ = e.get_particles(r)
    assert t[3] == 2 and sorted(t[:2].tolist()) == [0, 3]  # one of the two topologies bound the shared particle
    assert t[2] == 0 and t[4] == 1                          # exactly at the radius: no reaction
    assert sorted(e.topology_types(r).tolist()) == [0, 0, 1]


def _chains_system(n_chains, beads, n_free, box, seed):
    pos, types, bonds, angles, dihedrals = W.system(n_chains, beads, n_free, box, seed=seed)
    topology_of = np.repeat(np.arange(n_chains), beads)
    return pos, types, bonds, topology_of


def _oracle(pos, types, bonds, topology_of, n_chains, box, rate, radius, seed, pots):
    o = po.System(4, [box] * 3, [1, 1, 1], diffusion=[1.0, 1.0, 1.0, 1.0])
    for kind, a, b, p in pots:
        o.add_pair_potential(kind, a, b, p)
    o.add_spatial_reaction(0, 0, 1, 0, 1, 2, rate, radius)
    o.set_seed(seed)
    o.add_particles(pos, types.astype(np.int32), ids=np.arange(len(types)))
    tof = np.full(len(types), -1, np.int64)
    tof[:len(topology_of)] = topology_of
    o.set_topologies(tof, np.zeros(n_chains, np.int64))
    for b in bonds:
        o.add_bond(int(b[0]), int(b[1]), 100.0, 1.0)
    o.nl_setup(max(o.max_cutoff(), radius))
    return o


@pytest.mark.parametrize("rate", [5.0, 1e10])
def test_random_chains_event_for_event(rate):
    """Chains of beads (one topology each) among free particles: same Philox counters, same positions:
    the same particles react, into the same species, and the topology types agree, step after step (with
    separate integrate / topology-reaction actions and injected noise between them)."""
    box, n_chains, beads, radius, dt = 14.0, 30, 8, 1.2, 0.01
    pos, types, bonds, topology_of = _chains_system(n_chains, beads, 600, box, seed=4)
    pots = [(L.POT_HARMONIC_REPULSION, a, b, [5.0, 0.8]) for a in range(3) for b in range(a, 3) if (a, b) != (0, 0)]
    e, r = _engine(box, n=len(types), pots=pots, seed=9)
    e.set_particles(r, pos, types)
    e.set_bonds(r, bonds, 100.0, 1.0)
    e.set_spatial_reactions(_sr(rate=rate, radius=radius))
    e.set_topologies(r, np.zeros(n_chains), topology_of)
    o = _oracle(pos, types, bonds, topology_of, n_chains, box, rate, radius, 9, pots)
    rng = np.random.default_rng(77)
    e.calculate_forces()
    o.calculate_forces()
    total = 0
    for t in range(4):
        noise = rng.normal(size=(len(types), 3))
        e.integrate(dt, noise)
        e.update_neighbor_list()
        o.euler_bd(dt, noise.ravel(), step=t)
        o.nl_fill()
        e.evaluate_topology_reactions(dt)
        o.topology_reactions(dt, step=t)
        gp, gt, gh = e.get_particles(r)
        op, ot, oid, _ = o.particles()
        order = np.argsort(oid)
        assert np.array_equal(gh, oid[order].astype(np.uint32))
        assert np.array_equal(gt, ot[order].astype(np.uint32)), t
        assert np.array_equal(e.topology_types(r), o.topology_types().astype(np.uint32))
        assert np.array_equal(e.topology_reaction_counts(r), o.topology_reaction_counts())
        total += int(o.topology_reaction_counts().sum())
        e.calculate_forces()
        o.calculate_forces()
    print(f"rate {rate}: {total} topology reactions in 4 steps", flush=True)
    assert total > (20 if rate > 1e9 else 3)


def test_fused_loop_equals_separate_actions_with_topology_reactions():
    """The fused step performs the topology reactions after the step's reactions on the post-reaction state
    and redoes the moves they change; the same run by separate actions gives the same state bit for bit."""
    box, n_chains, beads, dt, steps = 14.0, 30, 8, 0.005, 12
    pos, types, bonds, topology_of = _chains_system(n_chains, beads, 600, box, seed=6)
    pots = [(L.POT_HARMONIC_REPULSION, a, b, [5.0, 0.8]) for a in range(3) for b in range(a, 3) if (a, b) != (0, 0)]
    rxs = [dict(kind=L.RX_CONVERSION, educts=[2], products=[1], rate=20.0)]  # X_bound releases: both kinds mix

    def build():
        e = Engine(4, [1.0] * 4, potentials=pots, reactions=rxs, topology_flavor=FLAVOR, nl_cutoff=1.2)
        r = e.add_replica([box] * 3, (1, 1, 1), seed=21, max_particles=4 * len(types))
        e.set_particles(r, pos, types)
        e.set_bonds(r, bonds, 100.0, 1.0)
        e.set_spatial_reactions([dict(type1=0, top_type1=0, type2=1, type_to1=0, top_type_to1=0, type_to2=2,
                                      rate=40.0, radius=1.2)])
        e.set_topologies(r, np.zeros(n_chains), topology_of)
        return e, r

    e1, r1 = build()
    e1.run(L.HANDLER_GILLESPIE, dt, steps)
    e2, r2 = build()
    e2.calculate_forces()
    for _ in range(steps):
        e2.integrate(dt)
        e2.update_neighbor_list()
        e2.react(L.HANDLER_GILLESPIE, dt)
        e2.evaluate_topology_reactions(dt)
        e2.update_neighbor_list()
        e2.calculate_forces()
    for x, y in zip(e1.get_particles(r1), e2.get_particles(r2)):
        assert np.array_equal(x, y)
    assert np.array_equal(e1.topology_reaction_counts(r1, totals=True), e2.topology_reaction_counts(r2, totals=True))
    assert e1.topology_reaction_counts(r1, totals=True).sum() > 10
    assert np.array_equal(e1.reaction_totals(r1), e2.reaction_totals(r2)) and e1.reaction_totals(r1).sum() > 10
    e3, r3 = build()
    e3.run(L.HANDLER_GILLESPIE, dt, steps, graph=True)
    for x, y in zip(e1.get_particles(r1), e3.get_particles(r3)):
        assert np.array_equal(x, y)


def test_argument_errors():
    e, r = _engine(20.0)
    e.set_particles(r, [[0, 0, 0], [1, 0, 0]], [0, 1])
    with pytest.raises(L.EngineError):
        e.set_spatial_reactions([dict(_sr()[0], type1=1)])  # the topology side must be a topology species
    with pytest.raises(L.EngineError):
        e.set_spatial_reactions([dict(_sr()[0], radius=5.0)])  # beyond the neighbour-list cutoff
    with pytest.raises(L.EngineError):
        e.set_spatial_reactions([dict(_sr()[0], rate=0.0)])
    e.set_spatial_reactions(_sr())
    with pytest.raises(L.EngineError):
        e.set_topologies(r, [0], [3])  # topology 3 does not exist


def test_move_and_append_topology_particles():
    """Recipe steps on the device: a moved topology particle lands in its new cell (its repulsion with a free
    particle it now overlaps shows in the forces), appended particles take the reserve's handles after the
    uploaded topology particles, and the reserve bounds them"""
    print("move and append", flush=True)
    pots = [(L.POT_HARMONIC_REPULSION, 0, 1, [10.0, 1.0])]
    e = Engine(4, [1.0] * 4, potentials=pots, topology_flavor=FLAVOR, nl_cutoff=1.2)
    r = e.add_replica([20.0] * 3, (1, 1, 1), seed=2, max_particles=256)
    e.reserve_topology_particles(r, 2)
    e.set_particles(r, [[0, 0, 0], [1, 0, 0], [5, 5, 5]], [0, 0, 1])  # handles 0, 1 (beads), 4 (free X)
    e.set_topologies(r, [0], [0, 0])
    e.calculate_forces()
    f, h = e.get_forces(r)
    assert h.tolist() == [0, 1, 4] and abs(f).max() == 0
    e.set_particle_positions(r, [1], [[5.5, 5, 5]])  # 0.5 from X: overlapping
    e.calculate_forces()
    f, h = e.get_forces(r)
    # harmonic repulsion k (1 - 0.5) = 5 pushes the bead (x 5.5) away from X (x 5.0)
    assert f[1, 0] == -f[2, 0] and abs(f[1, 0] - 5.0) < 1e-6 and f[0].tolist() == [0, 0, 0]
    assert np.allclose(e.get_particles(r)[0][1], [5.5, 5, 5], rtol=0, atol=0)
    new = e.append_topology_particles(r, [3, 0], [[-1, 0, 0], [9, 9, 9]])
    assert new.tolist() == [2, 3]
    pos, types, handles = e.get_particles(r)
    assert handles.tolist() == [0, 1, 2, 3, 4] and types.tolist() == [0, 0, 3, 0, 1]
    assert np.array_equal(pos[2], [-1, 0, 0])
    e.set_topologies(r, [0], [0, 0, 0, 0])  # four topology particles now
    with pytest.raises(L.EngineError, match="reserve"):
        e.append_topology_particles(r, [0], [[1, 1, 1]])
    with pytest.raises(L.EngineError, match="topology species"):
        e.append_topology_particles(r, [1], [[1, 1, 1]])
    with pytest.raises(L.EngineError, match="topology particles"):
        e.set_particle_positions(r, [4], [[0, 0, 0]])
    e.update_neighbor_list()
    e.calculate_forces()
    assert e.get_forces(r)[0].shape == (5, 3)
