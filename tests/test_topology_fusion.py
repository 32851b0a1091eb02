"""Topology-particle fusion, "T1(p1) + (p2) -> T2(p3--p4)" (SpatialTopologyReaction.h:64-236, STRMode::TP_FUSION;
handleTopologyParticleReaction with appendParticle, CPUEvaluateTopologyReactions.cpp:369-404).

The oracle is pinned on a hand-built polymer-growth scene whose outcome follows from the rules alone (rate
1e10 accepts with probability 1 in FP64): every chain end with a free monomer inside the radius grows by that
monomer, which joins under the next reserve id and is the new end, so that the next evaluation grows the chain
from it (a wrong id or topology for the joining particle, or a dropped type change, fails this); a monomer
shared by two ends joins one of them; a fusion beyond the reserve is an error.  On the device (gpu): the same
scenes, random chains growing among monomers event for event against the oracle (ids, species, fusion rows,
the bonds of the grown edges), and the fused loop against the separate actions bit for bit.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

MID, END, MONO = 0, 1, 2   # chain species (topology flavour) and the free monomer
FLAVOR = [1, 1, 0]
GROW = dict(type1=END, top_type1=0, type2=MONO, type_to1=MID, top_type_to1=0, type_to2=END, rate=1e10, radius=1.0,
            fusion=1)

# two chains (ids 0-1 and 2-3, each "mid end" along x), a reserve of three ids (4, 5, 6), monomers 7-10:
# 7 is 0.7 from chain 0's end, 8 is 0.8 from chain 1's end, 9 is 1.5 from chain 0's end but 0.8 from 7, 10 far
SCENE_POS = np.array([[-5, 0, 0], [-4, 0, 0], [5, 0, 0], [6, 0, 0],
                      [-3.3, 0, 0], [6.8, 0, 0], [-2.5, 0, 0], [9, 0, 0]], float)
SCENE_TYPES = np.array([MID, END, MID, END, MONO, MONO, MONO, MONO])
SCENE_IDS = np.array([0, 1, 2, 3, 7, 8, 9, 10])


def _oracle(pos, types, ids, topology_of_id, top_types, reserve, box=20.0, reaction=GROW, pots=(), seed=3):
    o = po.System(3, [box] * 3, [1, 1, 1], diffusion=[1.0, 1.0, 1.0])
    for kind, a, b, p in pots:
        o.add_pair_potential(kind, a, b, p)
    x = reaction
    o.add_spatial_reaction(x["type1"], x["top_type1"], x["type2"], x["type_to1"], x["top_type_to1"], x["type_to2"],
                           x["rate"], x["radius"], fusion=bool(x.get("fusion")))
    o.set_seed(seed)
    o.add_particles(pos, np.asarray(types, np.int32), ids=np.asarray(ids))
    o.set_topologies(np.asarray(topology_of_id), np.asarray(top_types))
    o.set_topology_reserve(*reserve)
    o.nl_setup(max(o.max_cutoff(), x["radius"]))
    return o


def _by_id(o):
    _, t, ids, _ = o.particles()
    return {int(i): int(ty) for i, ty in zip(ids, t)}


def test_oracle_polymer_growth_scene():
    o = _oracle(SCENE_POS, SCENE_TYPES, SCENE_IDS, [0, 0, 1, 1], [0, 0], (4, 7))
    assert o.topology_reactions(1.0, step=0) == 2
    rows = o.topology_fusions()
    assert sorted(map(tuple, rows[:, :3].tolist())) == [(0, 1, 7), (1, 3, 8)]
    assert sorted(rows[:, 3].tolist()) == [4, 5]  # reserve ids, in the order performed
    new7 = int(rows[rows[:, 2] == 7, 3][0])
    t = _by_id(o)
    assert t[1] == MID and t[3] == MID and t[new7] == END and 7 not in t and 8 not in t
    # the joined monomer is chain 0's end now: 9 joins it (0.8 away), chain 1 finds nothing
    assert o.topology_reactions(1.0, step=1) == 1
    assert o.topology_fusions().tolist() == [[0, new7, 9, 6]]
    t = _by_id(o)
    assert t[new7] == MID and t[6] == END and t[10] == MONO
    assert o.topology_reactions(1.0, step=2) == 0 and len(o.topology_fusions()) == 0
    assert o.topology_types().tolist() == [0, 0]


def test_oracle_fusion_reserve_and_shared_monomer():
    # a monomer midway between the two chain ends (0.5 from each): exactly one chain takes it
    pos = np.array([[-1.5, 0, 0], [-0.5, 0, 0], [1.5, 0, 0], [0.5, 0, 0], [0, 0, 0]], float)
    o = _oracle(pos, [MID, END, MID, END, MONO], [0, 1, 2, 3, 5], [0, 0, 1, 1], [0, 0], (4, 5))
    assert o.topology_reactions(1.0, step=0) == 1
    (top, end, mono, new), = o.topology_fusions().tolist()
    assert (end, mono, new) == ((1 if top == 0 else 3), 5, 4)
    # two fusions wanted, one reserve id: an error
    o = _oracle(SCENE_POS, SCENE_TYPES, SCENE_IDS, [0, 0, 1, 1], [0, 0], (4, 5))
    with pytest.raises(RuntimeError, match="reserve"):
        o.topology_reactions(1.0, step=0)


def test_oracle_enzymatic_and_fusion_differ_only_in_the_join():
    """the same scene with the enzymatic kind: the same events, no join, ids kept"""
    enz = dict(GROW, fusion=0, type_to2=MONO)
    o = _oracle(SCENE_POS, SCENE_TYPES, SCENE_IDS, [0, 0, 1, 1], [0, 0], (4, 7), reaction=enz)
    assert o.topology_reactions(1.0, step=0) == 2 and len(o.topology_fusions()) == 0
    t = _by_id(o)
    assert t[1] == MID and t[3] == MID and t[7] == MONO and t[8] == MONO
    assert o.topology_reactions(1.0, step=1) == 0  # no chain end left


# ------------------------------------------------------------------------------------------------ device
def _engine(box, reserve, n, pots=(), rxs=(), seed=3, nl_cutoff=1.0):
    from readdy_amd import Engine
    e = Engine(3, [1.0, 1.0, 1.0], potentials=list(pots), reactions=list(rxs), topology_flavor=FLAVOR,
               nl_cutoff=nl_cutoff)
    r = e.add_replica([box] * 3, (1, 1, 1), seed=seed, max_particles=4 * n + 64)
    e.reserve_topology_particles(r, reserve)
    return e, r


@pytest.mark.gpu
def test_polymer_growth_scene_on_the_device():
    print("fusion scene", flush=True)
    e, r = _engine(20.0, 3, len(SCENE_TYPES))
    e.set_particles(r, SCENE_POS, SCENE_TYPES.astype(np.uint32))
    _, _, h = e.get_particles(r)
    assert h.tolist() == SCENE_IDS.tolist()  # free particles numbered after the reserve
    e.set_spatial_reactions([GROW])
    e.set_topologies(r, [0, 0], [0, 0, 1, 1])
    o = _oracle(SCENE_POS, SCENE_TYPES, SCENE_IDS, [0, 0, 1, 1], [0, 0], (4, 7))
    zero = np.zeros((16, 3))
    for step in range(3):
        e.integrate(1.0, zero)  # no force, no noise: the positions stay, the step counter advances
        e.evaluate_topology_reactions(1.0)
        o.topology_reactions(1.0, step=step)
        _, gt, gh = e.get_particles(r)
        assert dict(zip(gh.tolist(), gt.tolist())) == _by_id(o), step
        assert e.topology_fusions(r).tolist() == o.topology_fusions().tolist(), step
    assert e.topology_reaction_counts(r, totals=True).tolist() == [3]
    # the topology particles are handles 0 .. 6 now: the assignment takes seven entries
    e.set_topologies(r, [0, 0], [0, 0, 1, 1, 0, 1, 0])
    with pytest.raises(Exception):
        e.set_topologies(r, [0, 0], [0, 0, 1, 1])


@pytest.mark.gpu
def test_fusion_beyond_the_reserve_is_a_capacity_error():
    from readdy_amd import _lib as L
    print("fusion reserve", flush=True)
    e, r = _engine(20.0, 1, len(SCENE_TYPES))
    pos, types = SCENE_POS, SCENE_TYPES.astype(np.uint32)
    e.set_particles(r, pos, types)
    e.set_spatial_reactions([GROW])
    e.set_topologies(r, [0, 0], [0, 0, 1, 1])
    with pytest.raises(L.EngineError, match="reserve"):
        e.evaluate_topology_reactions(1.0)
    with pytest.raises(L.EngineError, match="joining particle"):
        e.set_spatial_reactions([dict(GROW, type_to2=MONO)])  # a fusion's joining particle must be a topology species


def _growth_system(n_chains, n_mono, box, seed):
    rng = np.random.default_rng(seed)
    pos, types, bonds, top = [], [], [], []
    for c in range(n_chains):
        p0 = rng.uniform(-box / 2 + 3.5, box / 2 - 3.5, 3)  # the whole chain inside the box
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        for k in range(4):
            pos.append(p0 + 0.9 * k * u)
            types.append(END if k == 3 else MID)
            top.append(c)
            if k:
                bonds.append((4 * c + k - 1, 4 * c + k))
    pos += list(rng.uniform(-box / 2, box / 2, (n_mono, 3)))
    types += [MONO] * n_mono
    return np.array(pos), np.array(types, np.uint32), bonds, top


@pytest.mark.gpu
@pytest.mark.parametrize("rate", [5.0, 1e10])
def test_random_growth_event_for_event(rate):
    """40 chains grow among 1500 monomers over five integrated steps: the same monomers join the same chains
    under the same reserve handles, the species and the fusion rows agree, and the grown edges get their
    bonds on both sides before the next step"""
    from readdy_amd import _lib as L
    print(f"fusion growth rate {rate}", flush=True)
    box, n_chains, n_mono, reserve, dt = 16.0, 40, 1500, 400, 0.01
    pos, types, bonds, top = _growth_system(n_chains, n_mono, box, seed=2)
    n_topo = 4 * n_chains
    pots = [(L.POT_HARMONIC_REPULSION, a, MONO, [5.0, 0.8]) for a in range(3)]
    e, r = _engine(box, reserve, len(types), pots=pots, seed=11)
    e.set_particles(r, pos, types)
    e.set_spatial_reactions([dict(GROW, rate=rate)])
    e.set_topologies(r, np.zeros(n_chains), top)
    ids = np.concatenate([np.arange(n_topo), np.arange(n_topo, len(types)) + reserve])
    tof = np.full(int(ids.max()) + 1, -1)
    tof[:n_topo] = top
    o = _oracle(pos, types, ids, tof, np.zeros(n_chains, int), (n_topo, n_topo + reserve), box=box,
                reaction=dict(GROW, rate=rate), pots=pots, seed=11)
    slot = {int(i): k for k, i in enumerate(ids)}  # the oracle's entries stay in their slots
    edges = list(bonds)
    e.set_bonds(r, edges, 100.0, 0.9)
    for a, b in edges:
        o.add_bond(slot[a], slot[b], 100.0, 0.9)
    rng = np.random.default_rng(5)
    e.calculate_forces()
    o.calculate_forces()
    grown = 0
    for t in range(5):
        noise = rng.normal(size=(n_topo + reserve + n_mono, 3))
        e.integrate(dt, noise)
        e.update_neighbor_list()
        o.euler_bd(dt, noise.ravel(), step=t)
        o.nl_fill()
        e.evaluate_topology_reactions(dt)
        o.topology_reactions(dt, step=t)
        rows = e.topology_fusions(r)
        assert rows.tolist() == o.topology_fusions().tolist(), t
        _, gt, gh = e.get_particles(r)
        assert dict(zip(gh.tolist(), gt.tolist())) == _by_id(o), t
        for tp, a, old, new in rows.tolist():
            slot[new] = slot.pop(old)
            edges.append((a, new))
            o.add_bond(slot[a], slot[new], 100.0, 0.9)
        grown += len(rows)
        e.set_bonds(r, edges, 100.0, 0.9)
        e.calculate_forces()
        o.calculate_forces()
        _, fh = e.get_forces(r)
        assert np.array_equal(fh, np.sort(o.particles()[2]).astype(np.uint32))
    print(f"{grown} monomers joined", flush=True)
    assert grown > (40 if rate > 1e9 else 5)
    # every chain is a path: each joined monomer is bonded to the previous end
    assert len(edges) == 3 * n_chains + grown


@pytest.mark.gpu
def test_fused_loop_equals_separate_actions_with_fusion():
    from readdy_amd import _lib as L
    print("fusion fused loop", flush=True)
    box, n_chains, n_mono, reserve, dt, steps = 16.0, 40, 1500, 600, 0.005, 12
    pos, types, bonds, top = _growth_system(n_chains, n_mono, box, seed=7)
    pots = [(L.POT_HARMONIC_REPULSION, a, MONO, [5.0, 0.8]) for a in range(3)]

    def build():
        e, r = _engine(box, reserve, len(types), pots=pots, seed=23)
        e.set_particles(r, pos, types)
        e.set_bonds(r, bonds, 100.0, 0.9)
        e.set_spatial_reactions([dict(GROW, rate=40.0)])
        e.set_topologies(r, np.zeros(n_chains), top)
        return e, r

    e1, r1 = build()
    e1.run(L.HANDLER_UA, dt, steps)
    e2, r2 = build()
    e2.calculate_forces()
    for _ in range(steps):
        e2.integrate(dt)
        e2.update_neighbor_list()
        e2.react(L.HANDLER_UA, dt)
        e2.evaluate_topology_reactions(dt)
        e2.update_neighbor_list()
        e2.calculate_forces()
    for x, y in zip(e1.get_particles(r1), e2.get_particles(r2)):
        assert np.array_equal(x, y)
    tot = e1.topology_reaction_counts(r1, totals=True)
    assert np.array_equal(tot, e2.topology_reaction_counts(r2, totals=True)) and tot.sum() > 10
    _, t1, h1 = e1.get_particles(r1)
    n_topo = 4 * n_chains
    joined = (h1 >= n_topo) & (h1 < n_topo + reserve)
    assert joined.sum() == tot.sum() and set(t1[joined].tolist()) <= {MID, END}


# ---------------------------------------------------------------------------------- topology-topology enzymatic
# species 0 A, 1 B, 2 A', 3 B' (all topology species); R1 "T0(A) + T1(B) -> T2(A') + T3(B')", R2 "T4(A) + T4(B) ->
# T5(A') + T5(B')".  Topologies (one or two particles each, all 0.5 apart in pairs): 0 {A} + 1 {B} react by R1;
# 2 {B, type 1} + 3 {A, type 0} react by R1 with its sides swapped (the lower id is the B); 4 {A, B} of type 4 is one
# topology (no reaction within a topology); 5 {A} + 6 {B} of type 4 react by R2; 7 {A} with two B topologies 8, 9
# on either side reacts once.
TT_POS = np.array([[0, 0, 0], [0.5, 0, 0], [5, 0, 0], [5.5, 0, 0], [10, 0, 0], [10.5, 0, 0], [12, 5, 0],
                   [12.5, 5, 0], [-5, 5, 0], [-4.5, 5, 0], [-5.5, 5, 0]], float)
TT_TYPES = np.array([0, 1, 1, 0, 0, 1, 0, 1, 0, 1, 1])
TT_TOP_OF = [0, 1, 2, 3, 4, 4, 5, 6, 7, 8, 9]
TT_TOP_TYPES = [0, 1, 1, 0, 4, 4, 4, 0, 1, 1]
TT_R1 = dict(type1=0, top_type1=0, type2=1, top_type2=1, type_to1=2, top_type_to1=2, type_to2=3, top_type_to2=3,
             rate=1e10, radius=1.0, topology_topology=1)
TT_R2 = dict(TT_R1, top_type1=4, top_type2=4, top_type_to1=5, top_type_to2=5)


def _tt_oracle(pos, types, top_of, top_types, reactions, box=30.0, seed=3, pots=()):
    o = po.System(4, [box] * 3, [1, 1, 1], diffusion=[1.0] * 4)
    for kind, a, b, p in pots:
        o.add_pair_potential(kind, a, b, p)
    for x in reactions:
        o.add_spatial_tt(x["type1"], x["top_type1"], x["type2"], x["top_type2"], x["type_to1"], x["top_type_to1"],
                         x["type_to2"], x["top_type_to2"], x["rate"], x["radius"])
    o.set_seed(seed)
    o.add_particles(pos, np.asarray(types, np.int32), ids=np.arange(len(types)))
    o.set_topologies(np.asarray(top_of), np.asarray(top_types))
    o.nl_setup(max(o.max_cutoff(), max(x["radius"] for x in reactions)))
    return o


def test_oracle_topology_topology_enzymatic_scene():
    o = _tt_oracle(TT_POS, TT_TYPES, TT_TOP_OF, TT_TOP_TYPES, [TT_R1, TT_R2])
    assert o.topology_reactions(1.0, step=0) == 4
    t = _by_id(o)
    assert [t[i] for i in range(8)] == [2, 3, 3, 2, 0, 1, 2, 3]
    tt = o.topology_types().tolist()
    assert tt[:7] == [2, 3, 3, 2, 4, 5, 5]
    # topology 7 reacted with exactly one of 8 and 9
    assert t[8] == 2 and sorted([t[9], t[10]]) == [1, 3] and tt[7] == 2 and sorted(tt[8:]) == [1, 3]
    assert o.topology_reaction_counts().tolist() == [3, 1]
    assert o.topology_reactions(1.0, step=1) == 0  # no A in a topology of type 0 next to a B in one of type 1


@pytest.mark.gpu
def test_topology_topology_enzymatic_scene_on_the_device():
    print("tt enzymatic scene", flush=True)
    from readdy_amd import Engine
    e = Engine(4, [1.0] * 4, topology_flavor=[1, 1, 1, 1], nl_cutoff=1.0)
    r = e.add_replica([30.0] * 3, (1, 1, 1), seed=3, max_particles=128)
    e.set_particles(r, TT_POS, TT_TYPES.astype(np.uint32))
    e.set_spatial_reactions([TT_R1, TT_R2])
    e.set_topologies(r, TT_TOP_TYPES, TT_TOP_OF)
    o = _tt_oracle(TT_POS, TT_TYPES, TT_TOP_OF, TT_TOP_TYPES, [TT_R1, TT_R2])
    e.integrate(1.0, np.zeros((11, 3)))  # no force, no noise: the step counter advances
    e.evaluate_topology_reactions(1.0)
    o.topology_reactions(1.0, step=0)
    _, gt, gh = e.get_particles(r)
    assert dict(zip(gh.tolist(), gt.tolist())) == _by_id(o)
    assert e.topology_types(r).tolist() == o.topology_types().tolist()
    assert e.topology_reaction_counts(r).tolist() == [3, 1]
    reacted = e.topology_reacted(r).tolist()
    assert reacted == [True, True, True, True, False, True, True, True] + [t == 3 for t in e.topology_types(r)[8:]]


@pytest.mark.gpu
@pytest.mark.parametrize("rate", [5.0, 1e10])
def test_topology_topology_random_chains_event_for_event(rate):
    """40 chains of four beads (species 0, topology type 0) among each other: bead pairs of different chains
    within 1.2 react "T0(0) + T0(0) -> T1(2) + T1(2)" (a symmetric reaction); the same pairs react on both
    sides over four integrated steps, one reaction per topology and step"""
    from readdy_amd import Engine
    from readdy_amd import _lib as L
    print(f"tt chains rate {rate}", flush=True)
    box, n_chains, dt = 12.0, 40, 0.01
    pos, types, bonds, top = _growth_system(n_chains, 0, box, seed=8)
    types = np.zeros(len(types), np.uint32)
    rx = dict(type1=0, top_type1=0, type2=0, top_type2=0, type_to1=2, top_type_to1=1, type_to2=2, top_type_to2=1,
              rate=rate, radius=1.2, topology_topology=1)
    pots = [(L.POT_HARMONIC_REPULSION, 0, 0, [5.0, 0.8])]
    e = Engine(4, [1.0] * 4, potentials=pots, topology_flavor=[1, 1, 1, 1], nl_cutoff=1.2)
    r = e.add_replica([box] * 3, (1, 1, 1), seed=17, max_particles=4 * len(types))
    e.set_particles(r, pos, types)
    e.set_bonds(r, bonds, 100.0, 0.9)
    e.set_spatial_reactions([rx])
    e.set_topologies(r, np.zeros(n_chains), top)
    o = _tt_oracle(pos, types, top, np.zeros(n_chains, int), [rx], box=box, seed=17, pots=pots)
    for a, b in bonds:
        o.add_bond(a, b, 100.0, 0.9)
    rng = np.random.default_rng(9)
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
        _, gt, gh = e.get_particles(r)
        assert dict(zip(gh.tolist(), gt.tolist())) == _by_id(o), t
        assert e.topology_types(r).tolist() == o.topology_types().tolist(), t
        total += int(o.topology_reaction_counts().sum())
        e.calculate_forces()
        o.calculate_forces()
    print(f"{total} topology-topology reactions", flush=True)
    assert total > (5 if rate > 1e9 else 1)


# ----------------------------------------------------------------------------------- topology-topology fusion
# species 0 A, 1 B, 2 A', 3 B'; M "T0(A) + T1(B) -> T2(A'--B')" merges the B side's topology into the A side's
# (the side of the lower id survives), of type 2; C "T2(A') + T1(B') -> T2(A') + T1(B)" can only fire between
# different topologies, so after correct merges it finds nothing (an absorbed particle left in its old topology,
# of type 1, would react).  Topologies: 0 {A} + 1 {B}; 2 {B} + 3 {A} (reverse order); 4 {A} with 5 {B}, 6 {B}
# on either side (merges with one).
TM_POS = np.array([[0, 0, 0], [0.5, 0, 0], [5, 0, 0], [5.5, 0, 0], [-5, 5, 0], [-4.5, 5, 0], [-5.5, 5, 0]], float)
TM_TYPES = np.array([0, 1, 1, 0, 0, 1, 1])
TM_TOP_OF = [0, 1, 2, 3, 4, 5, 6]
TM_TOP_TYPES = [0, 1, 1, 0, 0, 1, 1]
TM_M = dict(type1=0, top_type1=0, type2=1, top_type2=1, type_to1=2, top_type_to1=2, type_to2=3, top_type_to2=2,
            rate=1e10, radius=1.0, topology_topology=1, fusion=1)
TM_C = dict(type1=2, top_type1=2, type2=3, top_type2=1, type_to1=2, top_type_to1=2, type_to2=1, top_type_to2=1,
            rate=1e10, radius=1.0, topology_topology=1)


def _tm_oracle(pos, types, top_of, top_types, reactions, box=30.0, seed=3, pots=()):
    o = po.System(4, [box] * 3, [1, 1, 1], diffusion=[1.0] * 4)
    for kind, a, b, p in pots:
        o.add_pair_potential(kind, a, b, p)
    for x in reactions:
        if x.get("fusion"):
            o.add_spatial_tt_fusion(x["type1"], x["top_type1"], x["type2"], x["top_type2"], x["type_to1"],
                                    x["type_to2"], x["top_type_to1"], x["rate"], x["radius"])
        else:
            o.add_spatial_tt(x["type1"], x["top_type1"], x["type2"], x["top_type2"], x["type_to1"], x["top_type_to1"],
                             x["type_to2"], x["top_type_to2"], x["rate"], x["radius"])
    o.set_seed(seed)
    o.add_particles(pos, np.asarray(types, np.int32), ids=np.arange(len(types)))
    o.set_topologies(np.asarray(top_of), np.asarray(top_types))
    o.nl_setup(max(o.max_cutoff(), max(x["radius"] for x in reactions)))
    return o


def test_oracle_topology_topology_fusion_scene():
    o = _tm_oracle(TM_POS, TM_TYPES, TM_TOP_OF, TM_TOP_TYPES, [TM_M, TM_C])
    assert o.topology_reactions(1.0, step=0) == 3
    rows = sorted(map(tuple, o.topology_merges().tolist()))
    assert rows[:2] == [(0, 1, 0, 1), (2, 3, 2, 3)] and rows[2][:3] in [(4, 5, 4), (4, 6, 4)]
    t = _by_id(o)
    assert [t[i] for i in range(5)] == [2, 3, 3, 2, 2]  # the reverse pair: the B (lower id) takes B'
    tt = o.topology_types().tolist()
    assert tt[0] == 2 and tt[2] == 2 and tt[4] == 2 and tt[1] == 1 and tt[3] == 0
    # the check reaction finds no A' and B' in different topologies
    assert o.topology_reactions(1.0, step=1) == 0


@pytest.mark.gpu
def test_topology_topology_fusion_scene_on_the_device():
    print("tt fusion scene", flush=True)
    from readdy_amd import Engine
    e = Engine(4, [1.0] * 4, topology_flavor=[1, 1, 1, 1], nl_cutoff=1.0)
    r = e.add_replica([30.0] * 3, (1, 1, 1), seed=3, max_particles=128)
    e.set_particles(r, TM_POS, TM_TYPES.astype(np.uint32))
    e.set_spatial_reactions([TM_M, TM_C])
    e.set_topologies(r, TM_TOP_TYPES, TM_TOP_OF)
    o = _tm_oracle(TM_POS, TM_TYPES, TM_TOP_OF, TM_TOP_TYPES, [TM_M, TM_C])
    for step in range(2):
        e.integrate(1.0, np.zeros((7, 3)))
        e.evaluate_topology_reactions(1.0)
        o.topology_reactions(1.0, step=step)
        _, gt, gh = e.get_particles(r)
        assert dict(zip(gh.tolist(), gt.tolist())) == _by_id(o), step
        assert e.topology_types(r).tolist() == o.topology_types().tolist(), step
        assert e.topology_merges(r).tolist() == o.topology_merges().tolist(), step
    assert e.topology_reaction_counts(r, totals=True).tolist() == [3, 0]


@pytest.mark.gpu
@pytest.mark.parametrize("rate", [50.0, 1e10])
def test_topology_topology_fusion_random_chains_event_for_event(rate):
    """40 two-bead chains (ends of species 0 in topologies of type 0, the other bead species 1) merge pairwise
    by "T0(0) + T0(0) -> T1(2--2)", then merged chains of type 1 merge on by "T1(0) + T0(0) -> T1(2--2)" and
    "T1(0) + T1(0) -> T1(2--2)" while free ends meet: the same merges on both sides over four integrated steps,
    the merged topologies never reacting within themselves"""
    from readdy_amd import Engine
    from readdy_amd import _lib as L
    print(f"tt fusion chains rate {rate}", flush=True)
    box, n, dt = 8.0, 40, 0.01
    rng = np.random.default_rng(21)
    pos, types, top = [], [], []
    for c in range(n):
        p0 = rng.uniform(-box / 2 + 1, box / 2 - 1, 3)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        pos += [p0, p0 + 0.9 * u]
        types += [1, 0]
        top += [c, c]
    pos, types = np.array(pos), np.array(types, np.uint32)
    bonds = [(2 * c, 2 * c + 1) for c in range(n)]
    base = dict(type1=0, type2=0, type_to1=2, type_to2=2, top_type_to1=1, top_type_to2=1, rate=rate, radius=1.2,
                topology_topology=1, fusion=1)
    rxs = [dict(base, top_type1=0, top_type2=0), dict(base, top_type1=1, top_type2=0),
           dict(base, top_type1=1, top_type2=1)]
    pots = [(L.POT_HARMONIC_REPULSION, 1, 1, [5.0, 0.8])]
    e = Engine(3, [1.0] * 3, potentials=pots, topology_flavor=[1, 1, 1], nl_cutoff=1.2)
    r = e.add_replica([box] * 3, (1, 1, 1), seed=29, max_particles=4 * len(types))
    e.set_particles(r, pos, types)
    e.set_bonds(r, bonds, 100.0, 0.9)
    e.set_spatial_reactions(rxs)
    e.set_topologies(r, np.zeros(n), top)
    o = po.System(3, [box] * 3, [1, 1, 1], diffusion=[1.0] * 3)
    o.add_pair_potential(*pots[0])
    for x in rxs:
        o.add_spatial_tt_fusion(x["type1"], x["top_type1"], x["type2"], x["top_type2"], 2, 2, 1, rate, 1.2)
    o.set_seed(29)
    o.add_particles(pos, types.astype(np.int32), ids=np.arange(len(types)))
    o.set_topologies(np.asarray(top), np.zeros(n, int))
    for a, b in bonds:
        o.add_bond(a, b, 100.0, 0.9)
    o.nl_setup(max(o.max_cutoff(), 1.2))
    e.calculate_forces()
    o.calculate_forces()
    merged = 0
    for t in range(4):
        noise = rng.normal(size=(len(types), 3))
        e.integrate(dt, noise)
        e.update_neighbor_list()
        o.euler_bd(dt, noise.ravel(), step=t)
        o.nl_fill()
        e.evaluate_topology_reactions(dt)
        o.topology_reactions(dt, step=t)
        _, gt, gh = e.get_particles(r)
        assert dict(zip(gh.tolist(), gt.tolist())) == _by_id(o), t
        assert e.topology_types(r).tolist() == o.topology_types().tolist(), t
        assert e.topology_merges(r).tolist() == o.topology_merges().tolist(), t
        merged += len(o.topology_merges())
        e.calculate_forces()
        o.calculate_forces()
    print(f"{merged} merges", flush=True)
    assert merged > 3


@pytest.mark.gpu
def test_fused_loop_equals_separate_actions_with_topology_topology_reactions():
    """two-bead chains with topology-topology enzymatic and fusion reactions in the fused step (after the step's
    reactions, merges rewriting topology indices on the device) and by separate actions: the same state bit for
    bit, the same reaction counts"""
    from readdy_amd import Engine
    from readdy_amd import _lib as L
    print("tt fused loop", flush=True)
    box, n, dt, steps = 8.0, 60, 0.005, 12
    rng = np.random.default_rng(33)
    pos, types, top = [], [], []
    for c in range(n):
        p0 = rng.uniform(-box / 2 + 1, box / 2 - 1, 3)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        pos += [p0, p0 + 0.9 * u]
        types += [1, 0]
        top += [c, c]
    pos, types = np.array(pos), np.array(types, np.uint32)
    bonds = [(2 * c, 2 * c + 1) for c in range(n)]
    rxs = [dict(type1=0, top_type1=0, type2=0, top_type2=0, type_to1=2, top_type_to1=1, type_to2=2, top_type_to2=1,
                rate=30.0, radius=1.2, topology_topology=1, fusion=1),
           dict(type1=1, top_type1=0, type2=1, top_type2=0, type_to1=1, top_type_to1=2, type_to2=1, top_type_to2=2,
                rate=30.0, radius=1.2, topology_topology=1)]
    pots = [(L.POT_HARMONIC_REPULSION, 1, 1, [5.0, 0.8])]

    def build():
        e = Engine(3, [1.0] * 3, potentials=pots, topology_flavor=[1, 1, 1], nl_cutoff=1.2)
        r = e.add_replica([box] * 3, (1, 1, 1), seed=41, max_particles=4 * len(types))
        e.set_particles(r, pos, types)
        e.set_bonds(r, bonds, 100.0, 0.9)
        e.set_spatial_reactions(rxs)
        e.set_topologies(r, np.zeros(n), top)
        return e, r

    e1, r1 = build()
    e1.run(L.HANDLER_UA, dt, steps)
    e2, r2 = build()
    e2.calculate_forces()
    for _ in range(steps):
        e2.integrate(dt)
        e2.update_neighbor_list()
        e2.react(L.HANDLER_UA, dt)
        e2.evaluate_topology_reactions(dt)
        e2.update_neighbor_list()
        e2.calculate_forces()
    for x, y in zip(e1.get_particles(r1), e2.get_particles(r2)):
        assert np.array_equal(x, y)
    tot = e1.topology_reaction_counts(r1, totals=True)
    assert np.array_equal(tot, e2.topology_reaction_counts(r2, totals=True)) and tot[0] > 0 and tot[1] > 0
    assert np.array_equal(e1.topology_types(r1), e2.topology_types(r2))


# self-fusion: a U-shaped chain of four (handles 0-3, ends 0.9 apart), a bonded pair (4, 5) and two single
# particles (6, 7) in topologies of their own; every topology has type 0 and species A=0 / B=1 / C=2
TS_POS = np.array([[0, 0, 0], [0.9, 0, 0], [0.9, 0.9, 0], [0, 0.9, 0], [5, 0, 0], [5.5, 0, 0], [10, 0, 0],
                   [10.5, 0, 0]], float)
TS_TYPES = np.array([0, 2, 2, 1, 0, 1, 0, 1])
TS_TOP_OF = [0, 0, 0, 0, 1, 1, 2, 3]
TS_BONDS = [(0, 1), (1, 2), (2, 3), (4, 5)]
TS_R = dict(type1=0, top_type1=0, type2=1, top_type2=0, type_to1=2, top_type_to1=1, type_to2=2, top_type_to2=1,
            rate=1e10, radius=1.0, topology_topology=1, fusion=1, self_fusion=1)


def _ts_oracle(reaction):
    o = po.System(3, [30.0] * 3, [1, 1, 1], diffusion=[1.0] * 3)
    o.add_spatial_tt_fusion(0, 0, 1, 0, 2, 2, 1, reaction["rate"], reaction["radius"])
    if reaction["self_fusion"]:
        o.allow_self_fusion()
    o.set_seed(3)
    o.add_particles(TS_POS, TS_TYPES.astype(np.int32), ids=np.arange(8))
    o.set_topologies(np.asarray(TS_TOP_OF), np.zeros(4, int))
    for a, b in TS_BONDS:
        o.add_bond(a, b, 100.0, 0.9)
    o.nl_setup(1.0)
    return o


@pytest.mark.parametrize("self_fusion", [1, 0])
def test_oracle_self_fusion_closes_rings_of_unbonded_ends_only(self_fusion):
    o = _ts_oracle(dict(TS_R, self_fusion=self_fusion))
    n = o.topology_reactions(1.0, step=0)
    rows = sorted(map(tuple, o.topology_merges().tolist()))
    t = _by_id(o)
    if self_fusion:  # the chain's ends join (a merge row naming topology 0 twice); the bonded pair does not
        assert n == 2 and rows == [(0, 0, 0, 3), (2, 3, 6, 7)]
        assert [t[i] for i in range(8)] == [2, 2, 2, 2, 0, 1, 2, 2]
        assert o.topology_types().tolist() == [1, 0, 1, 0]
    else:
        assert n == 1 and rows == [(2, 3, 6, 7)]
        assert [t[i] for i in range(8)] == [0, 2, 2, 1, 0, 1, 2, 2]
        assert o.topology_types().tolist() == [0, 0, 1, 0]
    with pytest.raises(Exception):
        po.System(3, [30.0] * 3, [1, 1, 1], diffusion=[1.0] * 3).allow_self_fusion()


@pytest.mark.gpu
@pytest.mark.parametrize("self_fusion", [1, 0])
def test_self_fusion_scene_on_the_device(self_fusion):
    print("self fusion scene", self_fusion, flush=True)
    from readdy_amd import Engine
    x = dict(TS_R, self_fusion=self_fusion)
    e = Engine(3, [1.0] * 3, topology_flavor=[1, 1, 1], nl_cutoff=1.0)
    r = e.add_replica([30.0] * 3, (1, 1, 1), seed=3, max_particles=64)
    e.set_particles(r, TS_POS, TS_TYPES.astype(np.uint32))
    e.set_spatial_reactions([x])
    e.set_topologies(r, [0, 0, 0, 0], TS_TOP_OF)
    e.set_bonds(r, TS_BONDS, 100.0, 0.9)
    o = _ts_oracle(x)
    e.evaluate_topology_reactions(1.0)
    o.topology_reactions(1.0, step=0)
    _, gt, gh = e.get_particles(r)
    assert dict(zip(gh.tolist(), gt.tolist())) == _by_id(o)
    assert e.topology_types(r).tolist() == o.topology_types().tolist()
    assert sorted(map(tuple, e.topology_merges(r).tolist())) == sorted(map(tuple, o.topology_merges().tolist()))
    assert e.topology_reaction_counts(r, totals=True).tolist() == [2 if self_fusion else 1]


@pytest.mark.gpu
def test_self_fusion_needs_a_topology_topology_fusion():
    print("self fusion validation", flush=True)
    from readdy_amd import Engine
    e = Engine(3, [1.0] * 3, topology_flavor=[1, 1, 1], nl_cutoff=1.0)
    with pytest.raises(Exception):
        e.set_spatial_reactions([dict(TS_R, fusion=0)])
