"""Pins the CPU oracle against the reference's known answers, closed forms and brute force.

Every check here is independent of the oracle's own formulas: reference golden values,
finite differences of energies, brute-force pair sets, a half-shell (Newton's third law)
traversal, and exact geometric facts.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))
KAT = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "philox_kat.json")))


def test_philox_known_answers():
    for v in KAT["vectors"]:
        out = po.philox4x32_10([int(x, 16) for x in v["ctr"]], [int(x, 16) for x in v["key"]])
        assert ["%08x" % x for x in out] == v["out"]


def test_u53_range_and_resolution():
    assert po.u53(0, 0) == 0.0
    assert po.u53(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53
    assert po.u53(1 << 5, 0) == 2.0 ** -27
    assert po.u53(0, 1 << 6) == 2.0 ** -53


def test_shortest_difference_boundary_rule():
    box, pbc = [10.0, 10.0, 10.0], [1, 1, 1]
    # dv = +L/2 stays (strict >), dv = -L/2 flips (<=): boundary_condition_operations.h:70-80
    d = po.shortest_difference([0, 0, 0], [5.0, -5.0, 4.0], box, pbc)
    assert list(d) == [5.0, 5.0, 4.0]
    d = po.shortest_difference([0, 0, 0], [5.0, 0, 0], box, [0, 0, 0])
    assert list(d) == [5.0, 0.0, 0.0]
    v = po.fix_position([5.0, -5.0, 27.0], box, pbc)
    assert list(v) == [-5.0, -5.0, -3.0]


def _build(scene, types_extra=()):
    names = scene["types"]
    tid = {n: i for i, n in enumerate(names)}
    s = po.System(len(names), scene["box"], scene["pbc"])
    for p in scene.get("potentials", []):
        t1, t2 = tid[p["types"][0]], tid[p["types"][1]]
        if p["kind"] == "HarmonicRepulsion":
            s.add_pair_potential(po.POT_HARMONIC_REPULSION, t1, t2, [p["force_constant"], p["interaction_distance"]])
        elif p["kind"] == "LennardJones":
            s.add_pair_potential(po.POT_LENNARD_JONES, t1, t2,
                                 [p["m"], p["n"], p["cutoff"], float(p["shift"]), p["epsilon"], p["sigma"]])
        elif p["kind"] == "ScreenedElectrostatics":
            s.add_pair_potential(po.POT_SCREENED_ELECTROSTATICS, t1, t2,
                                 [p["electrostatic_strength"], p["inverse_screening_depth"], p["repulsion_strength"],
                                  p["repulsion_distance"], p["exponent"], p["cutoff"]])
    pos = [p[1] for p in scene["particles"]]
    types = [tid[p[0]] for p in scene["particles"]]
    s.add_particles(pos, types, ids=np.arange(len(types)))
    s.nl_setup(s.max_cutoff())
    return s


def _forces_by_id(s):
    pos, types, ids, f = s.particles()
    out = np.zeros((len(ids), 3))
    out[ids.astype(int)] = f
    return out


def test_harmonic_repulsion_two_particles():
    g = GOLDEN["harmonic_repulsion_two_particles"]
    s = _build(g)
    e1 = s.calculate_forces()
    e2 = s.calculate_forces()  # calculating twice must not accumulate (TestStateModel.cpp:86)
    assert e1 == e2
    assert e1 == pytest.approx(g["energy"], rel=1e-12)
    np.testing.assert_allclose(_forces_by_id(s), g["forces"], atol=1e-12)


def test_harmonic_repulsion_six_particles():
    g = GOLDEN["harmonic_repulsion_six_particles"]
    s = _build(g)
    e = s.calculate_forces()
    pf = {k: np.array(v) for k, v in g["pair_forces"].items()}
    expected = np.zeros((6, 3))
    expected[0] = pf["03"] + pf["05"]
    expected[1] = pf["13"] + pf["15"]
    expected[2] = pf["23"] + pf["25"]
    expected[3] = -pf["03"] - pf["13"] - pf["23"]
    expected[5] = -pf["05"] - pf["15"] - pf["25"]
    np.testing.assert_allclose(_forces_by_id(s), expected, atol=1e-12)
    assert e == pytest.approx(sum(g["pair_energies"].values()), rel=1e-12)


def test_lennard_jones_known_answer():
    g = GOLDEN["lennard_jones"]
    s = _build(g)
    e = s.calculate_forces()
    assert e == pytest.approx(g["energy"], rel=g["energy_rel_tolerance"])
    np.testing.assert_allclose(_forces_by_id(s), g["forces"], atol=1e-5)


def test_screened_electrostatics_known_answer():
    g = GOLDEN["screened_electrostatics"]
    s = _build(g)
    e = s.calculate_forces()
    assert e == pytest.approx(g["energy"], rel=1e-9)
    np.testing.assert_allclose(_forces_by_id(s), g["forces"], atol=g["force_tolerance"])


def _fd_force(energy, x, h=1e-6):
    g = np.zeros(3)
    for d in range(3):
        xp, xm = np.array(x, float), np.array(x, float)
        xp[d] += h
        xm[d] -= h
        g[d] = (energy(xp) - energy(xm)) / (2 * h)
    return -g


@pytest.mark.parametrize("kind,params,xs", [
    (po.POT_HARMONIC_REPULSION, [10.0, 1.0], [[0.3, -0.2, 0.5], [0.1, 0.0, 0.0], [0.6, 0.5, 0.4]]),
    (po.POT_WEAK_INTERACTION, [10.0, 1.0, 0.5, 2.0], [[0.3, -0.2, 0.5], [0.8, 0.6, 0.3], [1.2, 0.9, 0.7],
                                                     [0.2, 1.5, 0.1]]),
    (po.POT_LENNARD_JONES, [12, 6, 2.5, 1.0, 1.0, 1.0], [[0.9, 0.2, 0.1], [1.3, -0.4, 0.6], [2.0, 0.1, 0.5]]),
    (po.POT_SCREENED_ELECTROSTATICS, [-1.0, 1.0, 1.0, 1.0, 6, 8.0], [[1.2, 1.5, -1.7], [0.9, 0.2, 0.1]]),
])
def test_pair_force_is_minus_energy_gradient(kind, params, xs):
    # x_ij = neighbour - self, so moving "self" by +h changes x_ij by -h: F_self = +dE/dx_ij
    for x in xs:
        f, _ = po.pair_force_energy(kind, params, x)
        g = -_fd_force(lambda y: po.pair_force_energy(kind, params, y)[1], x)
        np.testing.assert_allclose(f, g, rtol=1e-6, atol=1e-8)


def test_weak_interaction_continuity_values():
    # closed-form energies at the piece boundaries: -h at d0, -h/2 at d0 + l/2, 0 at rc
    k, d0, h, rc = 10.0, 1.0, 0.5, 2.0
    for r, want in [(d0, -h), (d0 + 0.5 * (rc - d0), -h / 2), (rc - 1e-12, 0.0)]:
        _, e = po.pair_force_energy(po.POT_WEAK_INTERACTION, [k, d0, h, rc], [r, 0.0, 0.0])
        assert e == pytest.approx(want, abs=1e-9)
    _, e = po.pair_force_energy(po.POT_WEAK_INTERACTION, [k, d0, h, rc], [0.0, 0.0, 0.0])
    assert e == pytest.approx(0.5 * k * d0 * d0 - h)


def test_harmonic_bond_known_answer():
    g = GOLDEN["harmonic_bond"]
    (fi, fj), e = po.bond_force_energy(g["force_constant"], g["length"], *g["positions"], g["box"], g["pbc"])
    assert list(fi) == g["forces"][0] and list(fj) == g["forces"][1]  # asserted with == in the reference
    assert e == g["energy"]


@pytest.mark.parametrize("name", ["harmonic_angle", "harmonic_angle_general"])
def test_harmonic_angle_known_answers(name):
    g = GOLDEN[name]
    f, e = po.angle_force_energy(g["force_constant"], g["theta0"], *g["positions"], g["box"], g["pbc"])
    assert e == pytest.approx(g["energy"], rel=1e-12)
    np.testing.assert_allclose(np.array(f), g["forces"], atol=g["force_tolerance"])


@pytest.mark.parametrize("name", ["cosine_dihedral", "cosine_dihedral_steep"])
def test_cosine_dihedral_known_answers(name):
    g = GOLDEN[name]
    f, e = po.dihedral_force_energy(g["force_constant"], g["multiplicity"], g["phi0"], *g["positions"], g["box"],
                                    g["pbc"])
    assert e == pytest.approx(g["energy"], rel=1e-12)
    np.testing.assert_allclose(np.array(f), g["forces"], atol=g["force_tolerance"])


def test_bonded_forces_are_energy_gradients():
    rng = np.random.default_rng(7)
    for _ in range(20):
        x = rng.normal(size=(4, 3)) * 0.8 + np.array([[0, 0, 0], [1, 0, 0], [1.5, 0.9, 0], [2.4, 1.0, 0.7]])
        k, mult, phi0 = 1.3, 3, 0.4
        f, _ = po.dihedral_force_energy(k, mult, phi0, *x)
        for a in range(4):
            def en(y, a=a):
                xx = x.copy()
                xx[a] = y
                return po.dihedral_force_energy(k, mult, phi0, *xx)[1]
            np.testing.assert_allclose(f[a], _fd_force(en, x[a]), rtol=1e-5, atol=1e-6)
        f, _ = po.angle_force_energy(2.0, 2.0, *x[:3])
        for a in range(3):
            def ea(y, a=a):
                xx = x[:3].copy()
                xx[a] = y
                return po.angle_force_energy(2.0, 2.0, *xx)[1]
            np.testing.assert_allclose(f[a], _fd_force(ea, x[a]), rtol=1e-5, atol=1e-6)


def test_dihedral_zero_force_branch():
    # |cos phi| < 1e-4 returns zero force but keeps the energy (TorsionPotential.cpp:91-93)
    x = [[-1, 0, 0], [0, 0, 0], [0, 0, 1], [0, 1, 1]]  # phi = 90 degrees
    f, e = po.dihedral_force_energy(1.0, 3, 0.0, *x)
    assert all(np.all(fi == 0) for fi in f)
    # phi = -pi/2 here, so E = k (1 + cos(-3 pi / 2)) = 1
    assert e == pytest.approx(1.0, abs=1e-12)


def _random_system(n, box, pbc, seed, n_types=2):
    rng = np.random.default_rng(seed)
    pos = (rng.random((n, 3)) - 0.5) * np.array(box)
    types = rng.integers(0, n_types, n)
    s = po.System(n_types, box, pbc)
    s.add_particles(pos, types, ids=np.arange(n))
    return s, pos, types


def _pairset(a):
    return set(map(tuple, a.astype(np.int64).tolist()))


@pytest.mark.parametrize("box,pbc,radius,cutoff", [
    ([10, 10, 10], [1, 1, 1], 1, 1.0), ([10, 10, 10], [1, 1, 1], 2, 1.0), ([10, 10, 10], [1, 1, 1], 3, 1.0),
    ([7.3, 5.1, 12.2], [1, 0, 1], 1, 1.3), ([3.0, 3.0, 3.0], [1, 1, 1], 1, 1.2), ([2.1, 9.0, 4.0], [0, 0, 0], 1, 1.0),
    ([1.5, 4.0, 1.5], [0, 0, 0], 1, 1.2), ([4, 4, 4], [1, 1, 1], 1, 1.9),
])
def test_cell_list_pairs_equal_brute_force(box, pbc, radius, cutoff):
    s, _, _ = _random_system(600, box, pbc, seed=11)
    s.nl_setup(cutoff, radius)
    nl = s.nl_pairs(cutoff)
    bf = s.brute_pairs(cutoff)
    assert len(nl) == len(_pairset(nl)), "a pair was visited twice"
    assert _pairset(nl) == _pairset(bf)


def test_neighbor_list_reference_scenes():
    g = GOLDEN["neighbor_list_scenes"]
    for sc in g["scenes"]:
        s = po.System(1, sc["box"], sc["pbc"])
        s.add_particles(sc["positions"], [0] * len(sc["positions"]), ids=np.arange(len(sc["positions"])))
        s.nl_setup(g["cutoff"], sc["radius"])
        pairs = _pairset(s.nl_pairs(g["cutoff"]))
        if sc.get("all_pairs"):
            n = len(sc["positions"])
            assert pairs == {(i, j) for i in range(n) for j in range(n) if i != j}
        for p in sc.get("must_contain", []):
            assert tuple(p) in pairs


def test_forces_match_half_shell_traversal():
    """Full-neighbour cell-list traversal with E/2 equals an i<j brute-force sum with Newton's third law."""
    box, pbc = [12.0, 12.0, 12.0], [1, 1, 1]
    s, pos, types = _random_system(700, box, pbc, seed=3)
    hr = [10.0, 1.0]
    wi = [10.0, 1.0, 0.5, 2.0]
    for t1 in range(2):
        for t2 in range(t1, 2):
            s.add_pair_potential(po.POT_HARMONIC_REPULSION, t1, t2, hr)
    s.add_pair_potential(po.POT_WEAK_INTERACTION, 0, 0, wi)
    s.nl_setup(s.max_cutoff())
    e = s.calculate_forces()
    f = _forces_by_id(s)
    n = len(types)
    F = np.zeros((n, 3))
    E = 0.0
    for i in range(n):
        for j in range(i + 1, n):
            x = po.shortest_difference(pos[i], pos[j], box, pbc)
            d2 = x @ x
            pots = [(po.POT_HARMONIC_REPULSION, hr, 1.0)]
            if types[i] == 0 and types[j] == 0:
                pots.append((po.POT_WEAK_INTERACTION, wi, 2.0))
            for kind, prm, rc in pots:
                if d2 < rc * rc:
                    fi, ei = po.pair_force_energy(kind, prm, x)
                    F[i] += fi
                    F[j] -= fi
                    E += ei
    np.testing.assert_allclose(f, F, rtol=1e-10, atol=1e-10)
    assert e == pytest.approx(E, rel=1e-10)


def test_euler_bd_closed_form_with_injected_noise():
    box, pbc = [10.0, 10.0, 10.0], [1, 1, 1]
    s = po.System(2, box, pbc, kbt=2.0, diffusion=[1.0, 0.5])
    pos = np.array([[4.9, 0.0, -4.9], [0.0, 1.0, 2.0]])
    s.add_particles(pos, [0, 1], ids=[0, 1])
    s.add_pair_potential(po.POT_HARMONIC_REPULSION, 0, 1, [10.0, 1.0])
    s.nl_setup(1.0)
    s.calculate_forces()  # particles too far apart: zero force
    noise = np.array([[1.0, -2.0, 0.5], [0.3, 0.2, -0.1]])
    dt = 0.01
    s.euler_bd(dt, noise.ravel())
    got, _, ids, _ = s.particles()
    for i, D in enumerate([1.0, 0.5]):
        want = pos[i] + noise[i] * math.sqrt(2 * D * dt)
        want = po.fix_position(want, box, pbc)
        np.testing.assert_array_equal(got[list(ids).index(i)], want)


def test_euler_bd_drift_term_at_kbt_not_one():
    """x += sqrt(2 D dt) xi + F D dt / kBT (CPUEulerBDIntegrator.cpp:63-73) with a nonzero force at kBT = 1.7.

    The force is the harmonic repulsion's closed form k (r - d) / r * x_ij (PotentialsOrder2.h:83-91),
    written out here, so a dropped 1/kBT, a flipped drift sign or a missing D in the drift fails."""
    box, pbc = [10.0, 10.0, 10.0], [1, 1, 1]
    kbt, D, dt, k, d = 1.7, [1.0, 0.4], 0.01, 10.0, 1.0
    s = po.System(2, box, pbc, kbt=kbt, diffusion=D)
    pos = np.array([[0.1, 0.2, 0.3], [0.5, -0.1, 0.55]])  # 0.56 apart: inside d = 1
    s.add_particles(pos, [0, 1], ids=[0, 1])
    s.add_pair_potential(po.POT_HARMONIC_REPULSION, 0, 1, [k, d])
    s.nl_setup(1.0)
    s.calculate_forces()
    x01 = pos[1] - pos[0]
    r = math.sqrt(x01 @ x01)
    f0 = k * (r - d) / r * x01  # force on particle 0 (pushed away from 1)
    forces = np.array([f0, -f0])
    assert f0[0] < 0  # repulsive: particle 0 moves away from particle 1 along x
    noise = np.array([[0.3, -1.1, 0.7], [-0.4, 0.9, 0.2]])
    s.euler_bd(dt, noise.ravel())
    got, _, ids, _ = s.particles()
    for i in range(2):
        want = pos[i] + noise[i] * math.sqrt(2 * D[i] * dt) + forces[i] * D[i] * dt / kbt
        np.testing.assert_allclose(got[list(ids).index(i)], want, rtol=0, atol=1e-14)
        wrong = pos[i] + noise[i] * math.sqrt(2 * D[i] * dt) + forces[i] * D[i] * dt  # without 1/kBT
        assert np.abs(got[list(ids).index(i)] - wrong).max() > 1e-3


def test_virial_known_answer_sign():
    """Virial -x_ij (x) f_i per pair (CPUCalculateForces.cpp:247-249): two particles 1.8 apart along z under
    harmonic repulsion k = 1, d = 2 (the scene of TestStateModel.cpp:57-100: F = -0.2 z on the first)
    give W_zz = -(1.8)(-0.2) = +0.36 and zero elsewhere, counted once for the pair."""
    s = po.System(1, [10.0] * 3, [1, 1, 1])
    s.add_pair_potential(po.POT_HARMONIC_REPULSION, 0, 0, [1.0, 2.0])
    s.add_particles([[0.0, 0.0, 0.0], [0.0, 0.0, 1.8]], [0, 0], ids=[0, 1])
    s.nl_setup(s.max_cutoff())
    e, w = s.calculate_forces(virial=True)
    want = np.zeros((3, 3))
    want[2, 2] = 0.36
    np.testing.assert_allclose(w, want, rtol=1e-12, atol=1e-15)
    assert e == pytest.approx(0.02, rel=1e-12)


def test_virial_equals_pair_sum_brute_force():
    """The oracle's virial over a periodic fluid equals sum over unordered pairs of -x_ij (x) f_i, formed here
    from all O(N^2) minimum-image pairs with the pinned pair forces: a double count, a dropped off-diagonal
    term or a transposed product (the WIPH pieces make x and f non-parallel only through summation, so the
    check also compares the full matrix) fails."""
    rng = np.random.default_rng(11)
    box, pbc = [6.0, 7.0, 8.0], [1, 1, 1]
    n = 120
    pos = (rng.random((n, 3)) - 0.5) * np.array(box)
    types = rng.integers(0, 2, n)
    hr, wi = [10.0, 1.0], [10.0, 1.0, 0.5, 2.0]
    s = po.System(2, box, pbc)
    for a in range(2):
        for b in range(a, 2):
            s.add_pair_potential(po.POT_HARMONIC_REPULSION, a, b, hr)
    s.add_pair_potential(po.POT_WEAK_INTERACTION, 0, 0, wi)
    s.add_particles(pos, types, ids=np.arange(n))
    s.nl_setup(s.max_cutoff())
    _, w = s.calculate_forces(virial=True)
    W = np.zeros((3, 3))
    pairs = 0
    for i in range(n):
        for j in range(i + 1, n):
            x = po.shortest_difference(pos[i], pos[j], box, pbc)
            d2 = x @ x
            pots = [(po.POT_HARMONIC_REPULSION, hr, 1.0)]
            if types[i] == 0 and types[j] == 0:
                pots.append((po.POT_WEAK_INTERACTION, wi, 2.0))
            for kind, prm, rc in pots:
                if d2 < rc * rc:
                    fi, _ = po.pair_force_energy(kind, prm, x)
                    W += np.outer(-x, fi)
                    pairs += 1
    assert pairs > 100
    np.testing.assert_allclose(w, W, rtol=1e-10, atol=1e-10)
    assert np.abs(W).max() > 1.0


def test_fusion_and_fission_placement():
    g = GOLDEN["fusion_placement"]
    s = po.System(3, g["box"], g["pbc"])
    s.add_reaction(po.RX_FUSION, [0, 1], [2], 1e16, g["educt_distance"] * 10, 0, g["weight1"], g["weight2"])
    s.add_particles([g["A"], g["B"]], [0, 1], ids=[0, 1])
    s.nl_setup(s.max_cutoff())
    s.reactions(po.UA_KEYED, 1.0)
    pos, types, _, _ = s.particles()
    assert list(types) == [2]
    np.testing.assert_allclose(pos[0], g["product"], atol=1e-15)

    g = GOLDEN["fission_distance"]
    s = po.System(3, g["box"], g["pbc"])
    s.add_reaction(po.RX_FISSION, [2], [0, 1], 1e16, 0, g["product_distance"], g["weight1"], g["weight2"])
    s.add_particles([g["C"]], [2], ids=[0])
    s.reactions(po.UA_KEYED, 1.0)
    pos, types, ids, _ = s.particles()
    assert sorted(types) == [0, 1]
    assert np.linalg.norm(pos[0] - pos[1]) == pytest.approx(g["product_distance"], rel=1e-12)
    # product ids: the appended product is drawn first, entry1's new id second (ReactionUtils.h:194-201)
    assert ids[list(types).index(1)] == 1 and ids[list(types).index(0)] == 2


def test_fusion_through_periodic_boundary():
    g = GOLDEN["fusion_through_periodic_boundary"]
    for mode in (po.UA_KEYED, po.GILLESPIE_LITERAL, po.GILLESPIE_KEYED):
        s = po.System(1, g["box"], g["pbc"])
        s.add_reaction(po.RX_FUSION, [0, 0], [0], g["rate"], g["radius"])
        s.add_particles(g["particles"], [0, 0], ids=[0, 1])
        s.nl_setup(s.max_cutoff())
        s.reactions(mode, 1.0)
        pos, _, _, _ = s.particles()
        assert len(pos) == 1
        np.testing.assert_allclose(pos[0], g["product"], rtol=g["product_rel_tolerance"])


def test_fission_near_boundary():
    g = GOLDEN["fission_near_boundary"]
    for seed in range(5):
        s = po.System(1, g["box"], g["pbc"])
        s.set_seed(seed)
        s.add_reaction(po.RX_FISSION, [0], [0, 0], g["rate"], 0, g["product_distance"])
        s.add_particles([g["particle"]], [0], ids=[0])
        s.reactions(po.UA_KEYED, 1.0)
        pos, _, _, _ = s.particles()
        assert len(pos) == 2
        assert (pos[0][0] <= 0) != (pos[1][0] <= 0)


def _gillespie_scene():
    g = GOLDEN["parallel_gillespie_scene"]
    names = g["types"]
    tid = {n: i for i, n in enumerate(names)}
    s = po.System(3, g["box"], g["pbc"], diffusion=[0.25] * 3)
    for r in g["reactions"]:
        s.add_reaction(po.RX_FUSION, [tid[r["educts"][0]], tid[r["educts"][1]]], [tid[r["product"]]], r["rate"],
                       r["radius"])
    s.add_particles([p[1] for p in g["particles"]], [tid[p[0]] for p in g["particles"]],
                    ids=np.arange(len(g["particles"])))
    s.nl_setup(s.max_cutoff())
    return s, g, tid


@pytest.mark.parametrize("mode", [po.UA_KEYED, po.GILLESPIE_LITERAL, po.GILLESPIE_KEYED])
def test_parallel_gillespie_scene(mode):
    s, g, tid = _gillespie_scene()
    s.reactions(mode, g["timestep"])
    pos, types, _, _ = s.particles()
    assert len(types) == g["expected_count"]
    have = [(int(t), tuple(p)) for t, p in zip(types, pos)]
    for name, p in g["expected_exact"]:
        assert (tid[name], tuple(float(x) for x in p)) in have
    for name, p in g["expected_close"]:
        assert any(t == tid[name] and np.allclose(q, p, rtol=1e-14) for t, q in have)


@pytest.mark.parametrize("mode", [po.UA_KEYED, po.GILLESPIE_LITERAL, po.GILLESPIE_KEYED])
def test_decay_scene_empties_box(mode):
    g = GOLDEN["decay_scene"]
    s = po.System(1, g["box"], g["pbc"], diffusion=[g["diffusion"]])
    s.add_reaction(po.RX_DECAY, [0], [], g["decay_rate"])
    s.add_reaction(po.RX_FISSION, [0], [0, 0], g["fission_rate"], 0, g["fission_distance"])
    s.add_particles(np.zeros((g["n_particles"], 3)), [0] * g["n_particles"], ids=np.arange(g["n_particles"]))
    for t in range(g["steps"]):
        s.calculate_forces()
        s.euler_bd(g["timestep"], step=t)
        s.reactions(mode, g["timestep"], step=t)
    assert s.n_active == g["expected_remaining"]


def test_gillespie_keyed_selection_matches_literal_in_distribution():
    """Efraimidis-Spirakis keys + greedy over accepted events reproduce performEvents' law.

    Small conflict graph (a chain plus an order-1 event) with unequal rates and acceptance
    probabilities; the frequency of each outcome pattern is compared by a chi-square test.
    """
    p1 = [0, 1, 2, 0, 3]
    p2 = [1, 2, 3, -1, -1]
    rate = [1.0, 3.0, 0.5, 2.0, 1.5]
    acc = [0.6, 0.3, 0.9, 0.5, 0.7]
    trials = 20000
    counts = {}
    for mode in (po.GILLESPIE_LITERAL, po.GILLESPIE_KEYED):
        c = {}
        for t in range(trials):
            out = po.perform_events_abstract(p1, p2, rate, acc, mode, seed=99 + mode, trial=t)
            key = tuple(out.tolist())
            c[key] = c.get(key, 0) + 1
        counts[mode] = c
    keys = sorted(set(counts[1]) | set(counts[2]))
    a = np.array([counts[1].get(k, 0) for k in keys], float)
    b = np.array([counts[2].get(k, 0) for k in keys], float)
    mask = (a + b) >= 20
    chi2 = np.sum((a[mask] - b[mask]) ** 2 / (a[mask] + b[mask]))
    dof = mask.sum() - 1
    from scipy.stats import chi2 as chi2dist
    assert chi2dist.sf(chi2, dof) > 1e-3, (chi2, dof, keys, a, b)


def test_ua_conserves_complex_counts():
    # readdy/test/TestReactions.cpp:75-118: #A + #AB and #B + #AB stay constant
    rng = np.random.default_rng(5)
    s = po.System(3, [5, 5, 5], [1, 1, 1], diffusion=[1.0, 1.0, 0.0])
    s.add_reaction(po.RX_FUSION, [0, 1], [2], 0.5, 1.0)
    s.add_reaction(po.RX_FISSION, [2], [0, 1], 0.5, 0, 1.0)
    pos = rng.uniform(-2.5, 2.5, (100, 3))
    s.add_particles(pos, [0, 1] * 50, ids=np.arange(100))
    s.nl_setup(s.max_cutoff())
    s.calculate_forces()
    for t in range(10):
        for mode in (po.UA_KEYED,):
            s.loop_step(1.0, mode, t)
        _, types, _, _ = s.particles()
        nA, nB, nAB = [(types == k).sum() for k in range(3)]
        assert nA + nAB == 50 and nB + nAB == 50


@pytest.mark.parametrize("mode", [po.UA_KEYED, po.GILLESPIE_KEYED])
def test_threaded_timing_baseline_equals_one_thread(mode):
    # the multithreaded oracle (timing baseline only) follows the one-thread run exactly: each particle
    # sums its own pairs in the same order, and the keyed handlers sort candidates by their keys
    rng = np.random.default_rng(8)
    pos = rng.uniform(-6, 6, (2000, 3))
    types = rng.integers(0, 3, 2000).astype(np.int32)
    runs = []
    for threads in (1, 4):
        s = po.System(3, [12, 12, 12], [1, 1, 1], diffusion=[1.0, 1.0, 0.5])
        s.set_seed(3)
        for a in range(3):
            for b in range(a, 3):
                s.add_pair_potential(po.POT_HARMONIC_REPULSION, a, b, [10.0, 1.0])
        s.add_pair_potential(po.POT_WEAK_INTERACTION, 0, 0, [10.0, 1.0, 0.5, 2.0])
        s.add_reaction(po.RX_FUSION, [0, 1], [2], 5.0, 1.0)
        s.add_reaction(po.RX_FISSION, [2], [0, 1], 2.0, 0.0, 1.0)
        s.add_particles(pos, types, ids=np.arange(2000))
        s.set_threads(threads)
        s.nl_setup(s.max_cutoff())
        s.calculate_forces()
        energies = [s.loop_step(0.01, mode, t) for t in range(5)]
        runs.append((s.particles(), energies, s.reaction_counts()))
    (p1, e1, c1), (p4, e4, c4) = runs
    for a, b in zip(p1, p4):
        assert np.array_equal(a, b)
    assert np.array_equal(c1, c4) and c1.sum() > 0
    np.testing.assert_allclose(e1, e4, rtol=1e-12)


def _rdf_numpy(pos, types, ids, box, pbc, borders, frm, to):
    """RadialDistribution counts by numpy: minimum-image distances of all ordered pairs, np.searchsorted
    (side='right' is std::upper_bound)"""
    d = pos[None, :, :] - pos[:, None, :]
    for a in range(3):
        if pbc[a]:
            d[..., a] = np.where(d[..., a] > box[a] / 2, d[..., a] - box[a], d[..., a])
            d[..., a] = np.where(d[..., a] <= -box[a] / 2, d[..., a] + box[a], d[..., a])
    dist = np.sqrt((d[..., 0] ** 2 + d[..., 1] ** 2) + d[..., 2] ** 2)
    use = np.isin(types, frm)[:, None] & np.isin(types, to)[None, :] & (ids[:, None] != ids[None, :])
    idx = np.searchsorted(borders, dist[use], side="right")
    idx = idx[(idx > 0) & (idx < len(borders))]
    return np.bincount(idx - 1, minlength=len(borders) - 1).astype(np.uint64), int(np.isin(types, frm).sum())


def test_rdf_two_particles_known_answer():
    # the reference's observable test scene (test_observables_io.py:201-238): A at (-2.5, 0, 0), B at the
    # origin, borders arange(0, 5, 0.01), counted from A to B: one pair at distance 2.5
    s = po.System(2, [10.0] * 3, [1, 1, 1])
    s.add_particles(np.array([[-2.5, 0, 0], [0, 0, 0]]), [0, 1], ids=np.arange(2))
    borders = np.arange(0, 5, 0.01)
    counts, nf = s.rdf_counts(borders, [0], [1])
    assert nf == 1 and counts.sum() == 1
    k = int(np.nonzero(counts)[0][0])
    assert borders[k] <= 2.5 < borders[k + 1]
    counts, nf = s.rdf_counts(borders, [1], [1])  # no B-B pair
    assert nf == 1 and counts.sum() == 0


@pytest.mark.parametrize("pbc", [(1, 1, 1), (1, 0, 1)])
def test_rdf_counts_equal_numpy_brute_force(pbc):
    rng = np.random.default_rng(12)
    box = np.array([8.0, 9.0, 10.0])
    pos = (rng.random((300, 3)) - 0.5) * box
    types = rng.integers(0, 3, 300).astype(np.int32)
    ids = np.arange(300)
    s = po.System(3, box, list(pbc))
    s.add_particles(pos, types, ids=ids)
    borders = np.concatenate([[0.3], np.linspace(0.5, 4.0, 36)])  # first border above 0: closer pairs uncounted
    for frm, to in [([0, 1, 2], [0, 1, 2]), ([0], [1, 2]), ([2], [2])]:
        counts, nf = s.rdf_counts(borders, frm, to)
        want, wnf = _rdf_numpy(pos, types, ids, box, pbc, borders, frm, to)
        assert nf == wnf
        assert np.array_equal(counts, want)


def test_histogram_along_axis_equals_numpy():
    # CPUObservables.cpp:85-136: upper_bound over the borders; bins 1 .. n_bins - 1 of the upper bound
    # (the CPU kernel tests the index against the bin count, so the last bin never fills)
    rng = np.random.default_rng(3)
    box = [10.0, 12.0, 8.0]
    pos = (rng.random((2000, 3)) - 0.5) * np.array(box)
    types = rng.integers(0, 3, 2000).astype(np.int32)
    s = po.System(3, box, [1, 1, 1])
    s.add_particles(pos, types, ids=np.arange(2000))
    borders = np.linspace(-5.0, 5.0, 41)
    for axis in range(3):
        for sel in ([0, 1, 2], [1], [0, 2]):
            got = s.histogram_along_axis(borders, sel, axis)
            x = pos[np.isin(types, sel), axis]
            u = np.searchsorted(borders, x, side="right")
            u = u[(u >= 1) & (u < len(borders) - 1)]
            want = np.bincount(u - 1, minlength=len(borders) - 1).astype(np.uint64)
            assert np.array_equal(got, want)
            assert got[-1] == 0


def test_cas_filled_cell_list_has_the_serial_pair_set():
    """The timing baseline's cell list, filled from all threads by atomic head exchange as
    CellLinkedList.cpp:171-228 fills it, holds the same pairs as the serial fill (only the order within
    a cell may differ), and its forces agree to rounding."""
    rng = np.random.default_rng(21)
    n, box = 20000, [30.0, 30.0, 30.0]
    pos = (rng.random((n, 3)) - 0.5) * 30.0
    out = []
    for cas in (False, True):
        s = po.System(1, box, [1, 1, 1])
        s.add_pair_potential(po.POT_HARMONIC_REPULSION, 0, 0, [10.0, 1.5])
        s.add_particles(pos, np.zeros(n, dtype=np.int32), ids=np.arange(n))
        s.set_threads(8)
        s.set_cas_fill(cas)
        s.nl_setup(1.5)
        pairs = s.nl_pairs(1.5)
        e = s.calculate_forces()
        out.append((set(map(tuple, pairs.tolist())), e, s.particles()[3]))
    assert out[0][0] == out[1][0] and len(out[0][0]) > 10000
    assert out[1][1] == pytest.approx(out[0][1], rel=1e-12)
    np.testing.assert_allclose(out[1][2], out[0][2], rtol=1e-12, atol=1e-12)


def _tp_system(pos, types, topology_of, top_types, rate=1e10, radius=1.2, to1=0, box=20.0):
    """Species 0 bead (topology), 1 X_free, 2 X_bound, 3 bead'; reaction T0(0) + (1) -> T1(to1) + (2)."""
    s = po.System(4, [box] * 3, [1, 1, 1])
    s.add_spatial_reaction(0, 0, 1, to1, 1, 2, rate, radius)
    s.add_particles(np.asarray(pos, float), np.asarray(types, np.int32), ids=np.arange(len(types)))
    s.set_topologies(topology_of, top_types)
    s.nl_setup(s.max_cutoff())
    return s


def _types_by_id(s):
    _, t, ids, _ = s.particles()
    return t[np.argsort(ids)]


def test_topology_particle_reactions_scene():
    """rate 1e10 (acceptance 1 in FP64): topology 0 (beads 0, 1) has one free particle near both beads, so
    exactly one of its two candidate events fires (events of one topology depend on each other,
    CPUEvaluateTopologyReactions.cpp:455-464); topology 1 (bead 2) binds its neighbour; a free particle
    beyond the radius, and one of a species with no reaction, stay; both topologies change type, so a
    second evaluation finds nothing (the reaction needs topology type 0)."""
    pos = [[0, 0, 0], [1, 0, 0], [6, 0, 0], [0.5, 0.5, 0], [6.5, 0, 0], [0, 3, 0], [6, 0.5, 0]]
    types = [0, 0, 0, 1, 1, 1, 2]
    s = _tp_system(pos, types, [0, 0, 1, -1, -1, -1, -1], [0, 0])
    assert s.max_cutoff() == 1.2
    assert s.topology_reactions(1.0, step=0) == 2
    assert list(_types_by_id(s)) == [0, 0, 0, 2, 2, 1, 2]
    assert list(s.topology_types()) == [1, 1] and list(s.topology_reaction_counts()) == [2]
    assert s.topology_reactions(1.0, step=1) == 0 and list(s.topology_reaction_counts()) == [0]


def test_topology_particle_reactions_share_the_free_particle():
    """One free particle within reach of two topologies reacts once: topology-particle events that share
    the free particle depend on each other; the topology particle takes type_to1, the free one type_to2."""
    pos = [[0, 0, 0], [1.5, 0, 0], [0.75, 0, 0]]
    s = _tp_system(pos, [0, 0, 1], [0, 1, -1], [0, 0], to1=3)
    assert s.topology_reactions(1.0, step=0) == 1
    t = _types_by_id(s)
    assert t[2] == 2 and sorted(t[:2].tolist()) == [0, 3]
    tt = s.topology_types()
    assert sorted(tt.tolist()) == [0, 1] and tt[int(np.argmax(t[:2] == 3))] == 1


def test_topology_particle_reaction_gate_is_strict():
    """distSquared < radius^2: a free particle at exactly the radius (1.25, exact in FP64) does not react,
    one just inside does."""
    for d, want in [(1.25, 0), (1.25 - 1e-12, 1)]:
        s = _tp_system([[0, 0, 0], [d, 0, 0]], [0, 1], [0, -1], [0], radius=1.25)
        assert s.topology_reactions(1.0) == want


def test_topology_keyed_selection_matches_literal_in_distribution():
    """The keyed selection with the topology dependencies (all events of one topology conflict; events
    sharing a free particle conflict) follows performEvents' law: events as (topology token, free particle
    token) pairs in the abstract performEvents of the oracle, chi-square against the literal algorithm.
    Topology 10 has three candidate partners, free particle 1 is shared by topologies 10 and 11."""
    p1 = [10, 10, 10, 11, 12]
    p2 = [1, 2, 3, 1, 3]
    rate = [1.0, 2.5, 0.7, 1.8, 3.0]
    acc = [0.5, 0.4, 0.8, 0.6, 0.3]
    trials = 20000
    counts = {}
    for mode in (po.GILLESPIE_LITERAL, po.GILLESPIE_KEYED):
        c = {}
        for t in range(trials):
            key = tuple(po.perform_events_abstract(p1, p2, rate, acc, mode, seed=7 + mode, trial=t).tolist())
            c[key] = c.get(key, 0) + 1
        counts[mode] = c
    keys = sorted(set(counts[po.GILLESPIE_LITERAL]) | set(counts[po.GILLESPIE_KEYED]))
    a = np.array([counts[po.GILLESPIE_LITERAL].get(k, 0) for k in keys], float)
    b = np.array([counts[po.GILLESPIE_KEYED].get(k, 0) for k in keys], float)
    mask = (a + b) >= 20
    chi2 = np.sum((a[mask] - b[mask]) ** 2 / (a[mask] + b[mask]))
    from scipy.stats import chi2 as chi2dist
    assert chi2dist.sf(chi2, mask.sum() - 1) > 1e-3, (chi2, keys, a, b)
