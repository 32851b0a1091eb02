"""The Python API (readdy_amd.api): the system's registries, the reaction descriptor grammar of
ReactionRegistry::add (readdy/main/model/reactions/ReactionRegistry.cpp:154-298) with the reference's error
cases, units and kBT, and a simulation through the API against the engine driven directly (GPU)."""
import math

import numpy as np
import pytest

from readdy_amd import _lib as L
from readdy_amd.api import BOLTZMANN_KJ_PER_MOL_K, ReactionDiffusionSystem


def _system(**kw):
    s = ReactionDiffusionSystem([10., 10., 10.], unit_system=None, **kw)
    for n in ("A", "B", "C"):
        s.add_species(n, 1.)
    return s


def test_descriptors_of_every_reaction_kind():
    """the descriptor examples of ReactionRegistry's tests (readdy/test/TestReactions.cpp) map to the same
    reactions as the explicit registry calls"""
    s = _system()
    s.reactions.add("decay: A ->", 1.)
    s.reactions.add("conv: A -> B", 2.)
    s.reactions.add("fus: A +(2.0) B -> C [0.3, 0.7]", 3.)
    s.reactions.add("fis: C -> A +(1.5) B", 4.)
    s.reactions.add("enz: A +(0.5) C -> B + C", 5.)
    rx = {r["name"]: r for r in s.reactions._rx}
    assert rx["decay"] == dict(name="decay", kind=L.RX_DECAY, educts=[0], products=[], rate=1.)
    assert rx["conv"] == dict(name="conv", kind=L.RX_CONVERSION, educts=[0], products=[1], rate=2.)
    assert rx["fus"] == dict(name="fus", kind=L.RX_FUSION, educts=[0, 1], products=[2], rate=3., educt_distance=2.,
                             weight1=0.3, weight2=0.7)
    assert rx["fis"] == dict(name="fis", kind=L.RX_FISSION, educts=[2], products=[0, 1], rate=4.,
                             product_distance=1.5, weight1=0.5, weight2=0.5)
    # "A +(r) C -> B + C": A the educt converted to B, C the catalyst (educts {from, catalyst})
    assert rx["enz"] == dict(name="enz", kind=L.RX_ENZYMATIC, educts=[0, 2], products=[1, 2], rate=5.,
                             educt_distance=0.5)
    assert s.calculate_max_cutoff() == 2.0


@pytest.mark.parametrize("descriptor,message", [
    ("A B", "must contain an arrow"),
    ("x: A -> B -> C", "more than one arrow"),
    ("A -> B", "colon"),
    ("x: A + B -> C", "parentheses"),
    ("x: C -> A + B", "parentheses"),
    ("x: A +(1) B ->", "did not contain one educt"),
    ("x: A +(1) C -> B + A", "have to be equal"),
    ("x: A +(1) B -> C [0.5 0.5]", "comma"),
    ("x: A +(r) B -> C", "not a number"),
    ("x: A -> D", "Could not find type"),
])
def test_descriptor_errors(descriptor, message):
    s = _system()
    with pytest.raises(ValueError, match=message):
        s.reactions.add(descriptor, 1.)


def test_registry_errors_and_type_names():
    s = _system()
    s.reactions.add_conversion("c", "A", "B", 1.)
    with pytest.raises(ValueError, match="exists already"):
        s.reactions.add_decay("c", "A", 1.)
    with pytest.raises(ValueError, match="negative"):
        s.reactions.add_decay("d", "A", -1.)
    with pytest.raises(ValueError, match="already exists"):
        s.add_species("A", 2.)
    for bad in ("A+B", "A(1)", "A->B", "x:y", "[A]", "A--B", " A", ""):
        with pytest.raises(ValueError):
            s.add_species(bad)
    s.add_species("A B", 1.)  # inner spaces are allowed
    with pytest.raises(NotImplementedError):
        s.potentials.add_box("A", 1., [0, 0, 0], [1, 1, 1])


def test_units_and_kbt():
    s = ReactionDiffusionSystem([10., 10., 10.])  # default units: 293 K
    assert math.isclose(s.kbt, 293. * 0.0083144626181532, rel_tol=1e-15)
    assert math.isclose(s.temperature, 293.)
    s.temperature = 300.
    assert math.isclose(s.kbt, 300. * BOLTZMANN_KJ_PER_MOL_K)
    with pytest.raises(ValueError):
        s.kbt = 2.
    u = ReactionDiffusionSystem([10., 10., 10.], unit_system=None)
    assert u.kbt == 1.
    u.kbt = 2.5
    assert u.kbt == 2.5
    with pytest.raises(ValueError):
        ReactionDiffusionSystem([10., 10., 10.], temperature=300., unit_system=None)
    with pytest.raises(ValueError):
        u.temperature
    assert u.periodic_boundary_conditions == [True, True, True] and u.box_volume == 1000.


def test_potentials_and_cutoff():
    s = _system()
    s.potentials.add_harmonic_repulsion("A", "B", 10., 1.2)
    s.potentials.add_lennard_jones("A", "A", 1., 1.)            # default cutoff 2.5 sigma
    s.potentials.add_weak_interaction_piecewise_harmonic("B", "C", 10., 1., 0.5, 1.8)
    s.potentials.add_screened_electrostatics("C", "C", -1., 1., 1., 1., 6., 3.1)
    assert s.calculate_max_cutoff() == 3.1
    kinds = [p[0] for p in s.potentials._pots]
    assert kinds == [L.POT_HARMONIC_REPULSION, L.POT_LENNARD_JONES, L.POT_WEAK_INTERACTION,
                     L.POT_SCREENED_ELECTROSTATICS]
    assert s.potentials._pots[1][3] == [12., 6., 2.5, 1., 1., 1.]
    with pytest.raises(AssertionError):
        s.potentials.add_lennard_jones("A", "A", 1., 1., shift=1)


def test_simulation_arguments():
    s = _system()
    with pytest.raises(ValueError):
        s.simulation(kernel="CPU")
    sim = s.simulation()
    assert sim.reaction_handler == "Gillespie"
    with pytest.raises(ValueError):
        sim.reaction_handler = "Fancy"
    sim.reaction_handler = "DetailedBalance"
    sim.add_particles("A", np.zeros((3, 3)))
    sim.add_particle("C", [1., 2., 3.])
    got = sim.current_particles()
    assert [g[0] for g in got] == ["A", "A", "A", "C"] and np.array_equal(got[3][1], [1., 2., 3.])
    with pytest.raises(ValueError):
        sim.run(10, 0.)


@pytest.mark.gpu
@pytest.mark.parametrize("handler", ["UncontrolledApproximation", "Gillespie"])
def test_simulation_through_the_api_equals_the_engine(tmp_path, handler):
    """A +(1) B <-> C with repulsion: the API's run equals the engine driven directly on the same seed, the
    callbacks see every evaluation, and the file holds the observables and the trajectory"""
    from readdy_amd import h5io
    from readdy_amd.engine import Engine
    print(f"api {handler}", flush=True)
    s = ReactionDiffusionSystem([12., 12., 12.], unit_system=None)
    s.add_species("A", 1.)
    s.add_species("B", 0.5)
    s.add_species("C", 0.2)
    s.potentials.add_harmonic_repulsion("A", "B", 10., 1.)
    s.reactions.add("fus: A +(1) B -> C", 2.)
    s.reactions.add("fis: C -> A +(1) B", 0.5)
    rng = np.random.default_rng(4)
    pa, pb = rng.uniform(-6, 6, (400, 3)), rng.uniform(-6, 6, (400, 3))
    sim = s.simulation(output_file=str(tmp_path / "api.h5"), reaction_handler=handler)
    sim.seed = 9
    sim.add_particles("A", pa)
    sim.add_particles("B", pb)
    seen = []
    sim.observe.number_of_particles(5, ["A", "C"], callback=lambda v: seen.append(v.tolist()))
    sim.observe.reaction_counts(5)
    energies, virials = [], []
    sim.observe.energy(5, callback=energies.append)
    sim.observe.virial(5, callback=virials.append)
    sim.record_trajectory(10)
    sim.run(20, 0.01)

    e = Engine(3, [1., 0.5, 0.2], potentials=s.potentials._pots, reactions=s.reactions._rx)
    r = e.add_replica([12.] * 3, (1, 1, 1), seed=9, max_particles=3200)
    e.set_particles(r, np.vstack([pa, pb]), np.array([0] * 400 + [1] * 400, dtype=np.uint32))
    e.record_virial(True)
    e.calculate_forces()
    h = L.HANDLER_UA if handler == "UncontrolledApproximation" else L.HANDLER_GILLESPIE
    counts = [e.count_types(1)[0][[0, 2]].tolist()]
    want_e, want_v = [e.energy(r)], [e.virial(r)]
    for _ in range(4):
        e.run(h, 0.01, 5)
        counts.append(e.count_types(1)[0][[0, 2]].tolist())
        want_e.append(e.energy(r))
        want_v.append(e.virial(r))
    assert energies == want_e and all(np.array_equal(a, b) for a, b in zip(virials, want_v))
    assert want_e[0] > 0 and abs(want_v[0]).max() > 0  # overlapping A-B pairs repel
    pos, types, _ = e.get_particles(r)
    names = ["A", "B", "C"]
    got = sim.current_particles()
    assert [g[0] for g in got] == [names[t] for t in types]
    assert np.array_equal(np.array([g[1] for g in got]), pos)
    assert seen == counts
    assert counts[-1][1] > 0
    e.close()
    path = str(tmp_path / "api.h5")
    assert h5io.read_u64(path, "/readdy/observables/n_particles/time").tolist() == [0, 5, 10, 15, 20]
    assert h5io.read_u64(path, "/readdy/observables/n_particles/data").reshape(5, 2).tolist() == counts
    assert h5io.read_f64(path, "/readdy/observables/energy/data").tolist() == want_e
    assert np.array_equal(h5io.read_f64(path, "/readdy/observables/virial/data").reshape(5, 3, 3), np.array(want_v))
    t, frames = h5io.read_trajectory(path, "")
    assert t.tolist() == [0, 10, 20] and np.array_equal(frames[-1][3], pos)


@pytest.mark.gpu
def test_detailed_balance_through_the_api():
    """the DetailedBalance handler runs the loop as separate actions; A +(1) B <-> C reaches a mix"""
    print("api DetailedBalance", flush=True)
    s = ReactionDiffusionSystem([10., 10., 10.], unit_system=None)
    for n in ("A", "B", "C"):
        s.add_species(n, 1.)
    s.reactions.add("fus: A +(1) B -> C", 5.)
    s.reactions.add("fis: C -> A +(1) B", 1.)
    sim = s.simulation(reaction_handler="DetailedBalance")
    rng = np.random.default_rng(5)
    sim.add_particles("A", rng.uniform(-5, 5, (200, 3)))
    sim.add_particles("B", rng.uniform(-5, 5, (200, 3)))
    last, total = [], []
    sim.observe.number_of_particles(1, ["A", "B", "C"], callback=lambda v: last.append(v.tolist()))
    sim.observe.number_of_particles(1, callback=lambda v: total.append(v.tolist()), save=dict(name="all"))
    sim.observe.reaction_counts(1)
    sim.run(20, 0.01)
    assert len(last) == 21
    nA, nB, nC = last[-1]
    assert nA == nB and nA + nC == 200 and nC > 0
    assert total == [[sum(x)] for x in last]  # without types: the total of every type
