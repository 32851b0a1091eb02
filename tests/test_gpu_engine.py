"""GPU parity tests: the HIP engine (through its C ABI) against the CPU oracle.

Bars: neighbour pair sets and reaction candidate sets bit-exact; forces and energies within
1e-5 relative of the FP64 oracle (FP32 pair arithmetic on FP64 min-image displacements);
free diffusion with injected noise bit-exact; reaction steps identical event-for-event when both
sides draw from the same Philox counters.
"""
import json
import os

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

L = pytest.importorskip("readdy_amd._lib")
from readdy_amd import Engine  # noqa: E402

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))

FORCE_TOL = 1e-5   # relative, ||dF|| / ||F_ref|| over all particles
ENERGY_TOL = 1e-5  # relative


def _pairset(a):
    return set(map(tuple, np.asarray(a, dtype=np.int64).tolist()))


def _oracle(n_types, box, pbc, pots, rxs=(), diffusion=None, kbt=1.0):
    o = po.System(n_types, box, pbc, kbt=kbt, diffusion=diffusion)
    for kind, a, b, p in pots:
        o.add_pair_potential(kind, a, b, p)
    for r in rxs:
        o.add_reaction(r["kind"], r.get("educts", []), r.get("products", []), r["rate"], r.get("educt_distance", 0.0),
                       r.get("product_distance", 0.0), r.get("weight1", 0.5), r.get("weight2", 0.5))
    return o


def _engine(n_types, box, pbc, pots, rxs=(), diffusion=None, seed=0, n=1000, kbt=1.0, **kw):
    diffusion = [1.0] * n_types if diffusion is None else diffusion
    e = Engine(n_types, diffusion, kbt=kbt, potentials=pots, reactions=list(rxs), **kw)
    r = e.add_replica(box, pbc, seed=seed, max_particles=4 * n + 64)
    return e, r


def _forces_sorted(o):
    _, _, ids, f = o.particles()
    return f[np.argsort(ids)]


@pytest.mark.parametrize("box,pbc,radius,cutoff,n", [
    ([10, 10, 10], [1, 1, 1], 1, 1.0, 2000), ([10, 10, 10], [1, 1, 1], 2, 1.0, 2000),
    ([7.3, 5.1, 12.2], [1, 0, 1], 1, 1.3, 1500), ([3.0, 3.0, 3.0], [1, 1, 1], 1, 1.2, 300),
    ([2.1, 9.0, 4.0], [0, 0, 0], 1, 1.0, 600), ([4, 4, 4], [1, 1, 1], 1, 1.9, 400),
    ([20, 20, 20], [1, 1, 1], 1, 2.0, 4000), ([1.5, 4.0, 1.5], [0, 0, 0], 1, 1.2, 50),
])
def test_pair_set_bit_exact(box, pbc, radius, cutoff, n):
    rng = np.random.default_rng(17)
    pos = (rng.random((n, 3)) - 0.5) * np.array(box)
    pos[: n // 10] = np.round(pos[: n // 10] * 4) / 4  # lattice points: exact half-box and cell-edge ties
    types = np.zeros(n, dtype=np.uint32)
    pots = [(L.POT_HARMONIC_REPULSION, 0, 0, [1.0, cutoff])]
    e, r = _engine(1, box, pbc, pots, n=n, cell_radius=radius)
    e.set_particles(r, pos, types)
    gpu = e.neighbor_pairs(r, cutoff)
    o = _oracle(1, box, pbc, [])
    o.add_particles(pos, types.astype(np.int32), ids=np.arange(n))
    ref = o.brute_pairs(cutoff)
    assert len(gpu) == len(_pairset(gpu)), "pair visited twice"
    assert _pairset(gpu) == _pairset(ref)


def test_neighbor_list_reference_scenes():
    g = GOLDEN["neighbor_list_scenes"]
    for sc in g["scenes"]:
        n = len(sc["positions"])
        pots = [(L.POT_HARMONIC_REPULSION, 0, 0, [1.0, g["cutoff"]])]
        e, r = _engine(1, sc["box"], sc["pbc"], pots, n=n, cell_radius=min(sc["radius"], 3))
        e.set_particles(r, sc["positions"], np.zeros(n, dtype=np.uint32))
        pairs = _pairset(e.neighbor_pairs(r, g["cutoff"]))
        if sc.get("all_pairs"):
            assert pairs == {(i, j) for i in range(n) for j in range(n) if i != j}
        for p in sc.get("must_contain", []):
            assert tuple(p) in pairs


def _system_c0like(n=2000, L_box=20.0, seed=0):
    rng = np.random.default_rng(seed)
    pos = (rng.random((n, 3)) - 0.5) * L_box
    types = np.repeat(np.arange(2, dtype=np.uint32), n // 2)
    return pos, types


def _force_case(pots, n_types, pos, types, box, pbc=(1, 1, 1), **kw):
    e, r = _engine(n_types, box, pbc, pots, n=len(types), **kw)
    e.set_particles(r, pos, types)
    e.calculate_forces()
    f, hs = e.get_forces(r)
    assert np.array_equal(hs, np.arange(len(types)))
    o = _oracle(n_types, box, pbc, pots)
    o.add_particles(pos, types.astype(np.int32), ids=np.arange(len(types)))
    o.nl_setup(o.max_cutoff())
    e_ref = o.calculate_forces()
    f_ref = _forces_sorted(o)
    rel = np.linalg.norm(f - f_ref) / np.linalg.norm(f_ref)
    pair_scale = np.abs(f_ref).sum()
    e_rel = abs(e.energy(r) - e_ref) / abs(e_ref)
    return rel, e_rel, np.linalg.norm(f - f_ref) / pair_scale


def test_forces_formed_on_demand_before_any_neighbour_list():
    """Exports and the separate integrator form the forces on demand when the exported ones are stale; straight
    after an upload nothing has built the brick tables (nor the bin-reading kernel's brick list) yet, so the
    on-demand evaluation builds them first.  Switching the virial on right after the upload makes the exported
    forces stale without a walk; every first call then gives what an explicit force evaluation gives."""
    print("on-demand forces after upload", flush=True)
    pots = [(L.POT_HARMONIC_REPULSION, 0, 0, [10.0, 1.0])]
    rng = np.random.default_rng(31)
    pos = rng.uniform(-10, 10, (8, 3))
    pos[1] = pos[0] + [0.6, 0, 0]  # one overlapping pair: a nonzero force
    types = np.zeros(8, np.uint32)
    noise = rng.normal(size=(8, 3))

    def fresh(explicit):
        e, r = _engine(1, [20.0] * 3, (1, 1, 1), pots, n=8)
        e.set_particles(r, pos, types)
        e.record_virial(True)
        if explicit:
            e.calculate_forces()
        return e, r

    e, r = fresh(True)
    f_ref, e_ref, v_ref = e.get_forces(r)[0], e.energy(r), e.virial(r)
    e.integrate(0.01, noise)
    x_ref = e.get_particles(r)[0]
    e.close()
    assert abs(f_ref).max() > 0 and e_ref > 0
    for first in ("forces", "energy", "virial", "integrate"):
        e, r = fresh(False)
        if first == "forces":
            assert np.array_equal(e.get_forces(r)[0], f_ref)
        elif first == "energy":
            assert e.energy(r) == e_ref
        elif first == "virial":
            assert np.array_equal(e.virial(r), v_ref)
        else:
            e.integrate(0.01, noise)
            assert np.array_equal(e.get_particles(r)[0], x_ref)
        e.close()


def test_forces_harmonic_repulsion_c0():
    pos, types = _system_c0like()
    pots = [(L.POT_HARMONIC_REPULSION, a, b, [10.0, 1.0]) for a in range(3) for b in range(a, 3)]
    rel, e_rel, _ = _force_case(pots, 3, pos, types, [20.0] * 3)
    assert rel < FORCE_TOL and e_rel < ENERGY_TOL, (rel, e_rel)


# lds_states: 0 sizes the LDS staging automatically (every brick staged); 600 leaves about half of
# the bricks (~600 halo records each) to the bin-reading kernel; 128 leaves all of them to it
@pytest.mark.parametrize("lds", [0, 600, 128])
def test_forces_weak_interaction_c1_chemistry(lds):
    # C1 chemistry at C1 density (0.125) in a smaller box: HR everywhere + WIPH between A and A
    n, Lb = 8000, 40.0
    rng = np.random.default_rng(1)
    pos = (rng.random((n, 3)) - 0.5) * Lb
    types = rng.choice(np.array([0, 1, 2], dtype=np.uint32), n, p=[0.4, 0.4, 0.2])
    pots = [(L.POT_HARMONIC_REPULSION, a, b, [10.0, 1.0]) for a in range(3) for b in range(a, 3)]
    pots.append((L.POT_WEAK_INTERACTION, 0, 0, [10.0, 1.0, 0.5, 2.0]))
    rel, e_rel, _ = _force_case(pots, 3, pos, types, [Lb] * 3, lds_states=lds)
    assert rel < FORCE_TOL and e_rel < ENERGY_TOL, (rel, e_rel)


def test_forces_lennard_jones_and_screened():
    n, Lb = 3000, 16.0
    rng = np.random.default_rng(2)
    # jittered lattice keeps LJ pairs away from the singularity
    g = np.stack(np.meshgrid(*[np.arange(15)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n].astype(float)
    pos = (g + 0.5) * (Lb / 15) - Lb / 2 + rng.normal(scale=0.05, size=(n, 3))
    types = rng.integers(0, 2, n).astype(np.uint32)
    pots = [(L.POT_LENNARD_JONES, 0, 0, [12, 6, 2.5, 1.0, 1.0, 1.0]),
            (L.POT_LENNARD_JONES, 0, 1, [12, 6, 2.5, 1.0, 1.0, 1.0]),
            (L.POT_SCREENED_ELECTROSTATICS, 1, 1, [-1.0, 1.0, 1.0, 1.0, 6, 3.0])]
    rel, e_rel, _ = _force_case(pots, 2, pos, types, [Lb] * 3)
    assert rel < FORCE_TOL and e_rel < ENERGY_TOL, (rel, e_rel)


@pytest.mark.parametrize("name", ["harmonic_repulsion_two_particles", "harmonic_repulsion_six_particles",
                                  "lennard_jones", "screened_electrostatics"])
def test_reference_known_answers_on_gpu(name):
    g = GOLDEN[name]
    tid = {t: i for i, t in enumerate(g["types"])}
    pots = []
    for p in g["potentials"]:
        a, b = tid[p["types"][0]], tid[p["types"][1]]
        if p["kind"] == "HarmonicRepulsion":
            pots.append((L.POT_HARMONIC_REPULSION, a, b, [p["force_constant"], p["interaction_distance"]]))
        elif p["kind"] == "LennardJones":
            pots.append((L.POT_LENNARD_JONES, a, b, [p["m"], p["n"], p["cutoff"], float(p["shift"]), p["epsilon"],
                                                     p["sigma"]]))
        else:
            pots.append((L.POT_SCREENED_ELECTROSTATICS, a, b, [p["electrostatic_strength"],
                                                               p["inverse_screening_depth"], p["repulsion_strength"],
                                                               p["repulsion_distance"], p["exponent"], p["cutoff"]]))
    pos = np.array([q[1] for q in g["particles"]], dtype=float)
    types = np.array([tid[q[0]] for q in g["particles"]], dtype=np.uint32)
    e, r = _engine(len(tid), g["box"], g["pbc"], pots, n=len(types))
    e.set_particles(r, pos, types)
    e.calculate_forces()
    f, _ = e.get_forces(r)
    if "forces" in g:
        np.testing.assert_allclose(f, g["forces"], rtol=1e-5, atol=1e-6)
    en = e.energy(r)
    want = g["energy"] if "energy" in g else sum(g["pair_energies"].values())
    assert en == pytest.approx(want, rel=1e-5)


def _chain(nb, rng):
    steps = rng.normal(size=(nb, 3))
    steps /= np.linalg.norm(steps, axis=1)[:, None]
    return np.cumsum(steps, axis=0) * 1.0


def test_bonded_terms_vs_oracle():
    rng = np.random.default_rng(3)
    Lb, nch, nb = 30.0, 20, 12
    pos = []
    for c in range(nch):
        x = _chain(nb, rng)
        x = x - x.mean(0) + (rng.random(3) - 0.5) * (Lb - 12)
        pos.append(x)
    pos = np.concatenate(pos)
    pos = (pos + Lb / 2) % Lb - Lb / 2
    n = len(pos)
    types = np.zeros(n, dtype=np.uint32)
    bonds = [(c * nb + i, c * nb + i + 1) for c in range(nch) for i in range(nb - 1)]
    angles = [(c * nb + i, c * nb + i + 1, c * nb + i + 2) for c in range(nch) for i in range(nb - 2)]
    dihs = [(c * nb + i, c * nb + i + 1, c * nb + i + 2, c * nb + i + 3) for c in range(nch) for i in range(nb - 3)]
    e = Engine(1, [1.0], topology_flavor=[1], nl_cutoff=2.0)
    r = e.add_replica([Lb] * 3, (1, 1, 1), max_particles=2 * n)
    e.set_particles(r, pos, types)
    e.set_bonds(r, bonds, 100.0, 1.0)
    e.set_angles(r, angles, 10.0, np.pi)
    e.set_dihedrals(r, dihs, 1.0, 3, 0.0)
    e.calculate_forces()
    f, _ = e.get_forces(r)
    o = po.System(1, [Lb] * 3, [1, 1, 1])
    o.add_particles(pos, types.astype(np.int32), ids=np.arange(n))
    for b in bonds:
        o.add_bond(*b, 100.0, 1.0)
    for a in angles:
        o.add_angle(*a, 10.0, np.pi)
    for d in dihs:
        o.add_dihedral(*d, 1.0, 3, 0.0)
    e_ref = o.calculate_forces()
    f_ref = _forces_sorted(o)
    assert np.linalg.norm(f - f_ref) / np.linalg.norm(f_ref) < 1e-9
    assert e.energy(r) == pytest.approx(e_ref, rel=1e-7)


def test_free_diffusion_injected_noise_bit_exact():
    n, Lb = 1000, 10.0
    rng = np.random.default_rng(4)
    pos = (rng.random((n, 3)) - 0.5) * Lb
    types = rng.integers(0, 2, n).astype(np.uint32)
    D = [1.0, 0.37]
    e = Engine(2, D, kbt=1.3)
    r = e.add_replica([Lb, Lb, Lb], (1, 0, 1), max_particles=2 * n)
    e.set_particles(r, pos, types)
    o = po.System(2, [Lb] * 3, [1, 0, 1], kbt=1.3, diffusion=D)
    o.add_particles(pos, types.astype(np.int32), ids=np.arange(n))
    for t in range(10):
        noise = rng.normal(size=(n, 3))
        e.integrate(0.01, noise)
        e.update_neighbor_list()
        o.euler_bd(0.01, noise.ravel())
    p, _, hs = e.get_particles(r)
    op, _, ids, _ = o.particles()
    assert np.array_equal(p, op[np.argsort(ids)])


def test_drift_with_forces_at_kbt_not_one():
    """EulerBD with pair forces at kBT = 1.7 and unequal D, injected noise, five integrate / NL / forces
    steps against the oracle.  A dropped 1/kBT or D, or a flipped drift, moves a particle by about
    |F| D dt (1 - 1/kBT) ~ 1e-3 per step here, far above the 1e-7 the FP32 pair forces allow."""
    n, Lb, kbt, dt = 1500, 12.0, 1.7, 0.01
    D = [1.0, 0.4]
    rng = np.random.default_rng(41)
    pos = (rng.random((n, 3)) - 0.5) * Lb
    types = rng.integers(0, 2, n).astype(np.uint32)
    pots = [(L.POT_HARMONIC_REPULSION, a, b, [10.0, 1.0]) for a in range(2) for b in range(a, 2)]
    pots.append((L.POT_WEAK_INTERACTION, 0, 0, [10.0, 1.0, 0.5, 2.0]))
    e, r = _engine(2, [Lb] * 3, (1, 1, 1), pots, diffusion=D, kbt=kbt, n=n)
    e.set_particles(r, pos, types)
    o = _oracle(2, [Lb] * 3, [1, 1, 1], pots, diffusion=D, kbt=kbt)
    o.add_particles(pos, types.astype(np.int32), ids=np.arange(n))
    o.nl_setup(o.max_cutoff())
    e.calculate_forces()
    o.calculate_forces()
    f_ref = _forces_sorted(o)
    assert np.abs(f_ref).max() > 1.0  # the drift term is exercised
    for t in range(5):
        noise = rng.normal(size=(n, 3))
        e.integrate(dt, noise)
        e.update_neighbor_list()
        e.calculate_forces()
        o.euler_bd(dt, noise.ravel())
        o.nl_fill()
        o.calculate_forces()
    p, _, _ = e.get_particles(r)
    op, _, ids, _ = o.particles()
    d = p - op[np.argsort(ids)]
    d -= Lb * np.round(d / Lb)
    err = np.abs(d).max()
    print(f"drift at kBT={kbt}: max |dx| = {err:.2e}", flush=True)
    assert err < 1e-7


@pytest.mark.parametrize("fused", [False, True])
def test_virial_vs_oracle(fused):
    """Context::recordVirial: the device virial (sum over pairs of -x_ij (x) f_i, CPUCalculateForces.cpp:
    247-249) against the FP64 oracle's within 1e-5 relative (Frobenius), for a separate force evaluation
    and for the state after fused steps (formed on demand)."""
    n, Lb = 4000, 20.0
    rng = np.random.default_rng(42)
    pos = (rng.random((n, 3)) - 0.5) * Lb
    types = rng.integers(0, 3, n).astype(np.uint32)
    pots = [(L.POT_HARMONIC_REPULSION, a, b, [10.0, 1.0]) for a in range(3) for b in range(a, 3)]
    pots.append((L.POT_WEAK_INTERACTION, 0, 0, [10.0, 1.0, 0.5, 2.0]))
    e, r = _engine(3, [Lb] * 3, (1, 1, 1), pots, n=n, seed=9)
    e.set_particles(r, pos, types)
    e.record_virial(True)
    if fused:
        e.run(L.HANDLER_NONE, 0.005, 4, graph=True)
        pos, types, hs = e.get_particles(r)
        order = np.argsort(hs)
        pos, types = pos[order], types[order]
    else:
        e.calculate_forces()
    w = e.virial(r)
    o = _oracle(3, [Lb] * 3, [1, 1, 1], pots)
    o.add_particles(pos, types.astype(np.int32), ids=np.arange(len(types)))
    o.nl_setup(o.max_cutoff())
    _, w_ref = o.calculate_forces(virial=True)
    rel = np.linalg.norm(w - w_ref) / np.linalg.norm(w_ref)
    print(f"virial rel err {rel:.2e}", flush=True)
    assert rel < 1e-5 and np.linalg.norm(w_ref) > 1.0
    e.record_virial(False)
    with pytest.raises(L.EngineError):
        e.virial(r)


def test_force_beyond_fixed_point_range_is_a_numeric_error():
    """A near-coincident Lennard-Jones pair (r = 0.002 sigma: |F| ~ 1e36) gives a force beyond the 2^-32
    fixed-point range of the force sums.  The force action and the fused step report RDY_ERR_NUMERIC
    (the reference would carry inf/NaN on), every kernel ends normally, and the engine keeps working."""
    lj = (L.POT_LENNARD_JONES, 0, 0, [12.0, 6.0, 2.5, 1.0, 1.0, 1.0])
    rng = np.random.default_rng(43)
    n, Lb = 200, 12.0
    pos = (rng.random((n, 3)) - 0.5) * Lb
    pos[1] = pos[0] + np.array([0.002, 0.0, 0.0])
    types = np.zeros(n, dtype=np.uint32)
    e, r = _engine(1, [Lb] * 3, (1, 1, 1), [lj], n=n)
    e.set_particles(r, pos, types)
    with pytest.raises(L.EngineError, match="error -5"):
        e.calculate_forces()
    with pytest.raises(L.EngineError, match="error -5"):
        e.run(L.HANDLER_NONE, 1e-6, 2)
    # a clean state afterwards: the engine is usable, the forces finite and below the range
    good = np.array([[x, y, z] for x in range(-5, 6, 2) for y in range(-5, 6, 2) for z in range(-5, 6, 2)], float)
    e.set_particles(r, good, np.zeros(len(good), dtype=np.uint32))
    e.calculate_forces()
    f, _ = e.get_forces(r)
    assert np.all(np.isfinite(f)) and len(f) == len(good)


def test_more_accepted_events_than_a_step_holds_is_a_capacity_error():
    """A step accepts at most 1024 reaction events per replica (the resolve workgroup keeps them in LDS).
    2000 particles of a species that decays at rate 1e10 all fire in one step: the separate reaction action
    and the fused step report RDY_ERR_CAPACITY naming the event limit, every kernel ends normally, and the
    engine keeps working on a state within the limit."""
    rng = np.random.default_rng(47)
    n, Lb = 2000, 30.0
    pos = (rng.random((n, 3)) - 0.5) * Lb
    types = np.zeros(n, dtype=np.uint32)
    decay = [dict(kind=L.RX_DECAY, educts=[0], products=[], rate=1e10)]
    e, r = _engine(2, [Lb] * 3, (1, 1, 1), [], rxs=decay, n=n)
    e.set_particles(r, pos, types)
    with pytest.raises(L.EngineError, match="error -3.*(event|1024)"):
        e.react(L.HANDLER_UA, 0.01)
    e.set_particles(r, pos, types)
    with pytest.raises(L.EngineError, match="error -3.*(event|1024)"):
        e.run(L.HANDLER_UA, 0.01, 2)
    # within the limit: 500 decaying particles and 1500 of an inert species all behave
    types[:1500] = 1
    e.set_particles(r, pos, types)
    e.run(L.HANDLER_UA, 0.01, 2)
    assert list(e.count_types(1)[0]) == [0, 1500]


def test_philox_noise_matches_oracle_stream():
    """Philox-driven BD: identical integer draws, the same FP32 Box-Muller transform on both sides.  The
    device's FP32 logarithm, sine and cosine and the oracle's correctly rounded ones may differ by an FP32
    unit, so each draw agrees to a few 2^-24 relative: at most 5 steps x sqrt(2 D dt) = 0.14 x 5 x 3e-7
    here, 2e-6 with margin; a different transform or counter would differ by O(0.1)."""
    n, Lb = 500, 10.0
    rng = np.random.default_rng(5)
    pos = (rng.random((n, 3)) - 0.5) * Lb
    types = np.zeros(n, dtype=np.uint32)
    e = Engine(1, [1.0])
    r = e.add_replica([Lb] * 3, (1, 1, 1), seed=1234, max_particles=2 * n)
    e.set_particles(r, pos, types)
    o = po.System(1, [Lb] * 3, [1, 1, 1])
    o.set_seed(1234)
    o.add_particles(pos, types.astype(np.int32), ids=np.arange(n))
    for t in range(5):
        e.integrate(0.01)
        o.euler_bd(0.01, step=t)
    p, _, _ = e.get_particles(r)
    op, _, ids, _ = o.particles()
    d = p - op[np.argsort(ids)]
    d -= Lb * np.round(d / Lb)
    print(f"Philox BD stream: max |dx| = {np.abs(d).max():.2e}", flush=True)
    assert np.abs(d).max() < 2e-6


def _c0_chemistry():
    pots = [(L.POT_HARMONIC_REPULSION, a, b, [10.0, 1.0]) for a in range(3) for b in range(a, 3)]
    rxs = [dict(kind=L.RX_FUSION, educts=[0, 1], products=[2], rate=0.1, educt_distance=1.0),
           dict(kind=L.RX_FISSION, educts=[2], products=[0, 1], rate=0.05, product_distance=1.0)]
    return pots, rxs


def test_reaction_candidates_bit_exact():
    pots, rxs = _c0_chemistry()
    rxs = rxs + [dict(kind=L.RX_FUSION, educts=[0, 0], products=[2], rate=0.3, educt_distance=0.7),
                 dict(kind=L.RX_ENZYMATIC, educts=[1, 2], products=[0, 2], rate=0.2, educt_distance=1.3),
                 dict(kind=L.RX_DECAY, educts=[1], rate=0.01), dict(kind=L.RX_CONVERSION, educts=[0], products=[1],
                                                                    rate=0.0)]
    rng = np.random.default_rng(6)
    n, Lb = 3000, 20.0
    pos = (rng.random((n, 3)) - 0.5) * Lb
    types = rng.integers(0, 3, n).astype(np.uint32)
    e, r = _engine(3, [Lb] * 3, (1, 1, 1), pots, rxs, n=n)
    e.set_particles(r, pos, types)
    gpu = e.reaction_candidates(r)
    o = _oracle(3, [Lb] * 3, [1, 1, 1], pots, rxs)
    o.add_particles(pos, types.astype(np.int32), ids=np.arange(n))
    o.nl_setup(o.max_cutoff())
    ref = o.reaction_candidates()
    g = {(int(a), int(b), int(c), int(d)) for a, b, c, d in gpu}
    f = {(int(a), int(b), int(c), int(d)) for a, b, c, d in ref}
    assert len(g) == len(gpu)
    assert g == f


# lds 256: about half of the bricks (~250 halo records) searched by the bin-reading kernel; 128: all
@pytest.mark.parametrize("handler,mode,lds", [(0, po.UA_KEYED, 0), (1, po.GILLESPIE_KEYED, 0), (0, po.UA_KEYED, 256),
                                              (1, po.GILLESPIE_KEYED, 128)])
def test_reaction_steps_match_oracle_event_for_event(handler, mode, lds):
    """Same positions, same Philox counters: the same events fire and products land identically."""
    pots, rxs = _c0_chemistry()
    rxs[0]["rate"] = 20.0
    rxs[1]["rate"] = 10.0
    n, Lb, dt = 2000, 20.0, 0.01
    pos, types = _system_c0like(n, Lb, seed=7)
    e, r = _engine(3, [Lb] * 3, (1, 1, 1), pots, rxs, diffusion=[1.0, 1.0, 0.5], seed=99, n=n, lds_states=lds)
    e.set_particles(r, pos, types)
    o = _oracle(3, [Lb] * 3, [1, 1, 1], pots, rxs, diffusion=[1.0, 1.0, 0.5])
    o.set_seed(99)
    o.add_particles(pos, types.astype(np.int32), ids=np.arange(n))
    o.nl_setup(o.max_cutoff())
    total = 0
    for t in range(5):
        noise = np.random.default_rng(100 + t).normal(size=(4 * n, 3))
        e.integrate(dt, noise)
        e.update_neighbor_list()
        o.euler_bd(dt, noise.ravel(), step=t)
        o.nl_fill()
        e.react(handler, dt)
        o.reactions(mode, dt, step=t)
        o.nl_fill()
        gp, gt, gh = e.get_particles(r)
        op, ot, oid, _ = o.particles()
        order = np.argsort(oid)
        assert np.array_equal(gh, oid[order].astype(np.uint32))
        assert np.array_equal(gt, ot[order].astype(np.uint32))
        d = gp - op[order]
        d -= Lb * np.round(d / Lb)
        # FP32 pair forces (rel. error ~1e-7) move positions by D dt dF ~ 1e-8 per step
        assert np.abs(d).max() < 1e-6
        assert np.array_equal(e.reaction_counts(r), o.reaction_counts())
        total += int(e.reaction_counts(r).sum())
        # forces for the next step are taken from the same state on both sides
        e.calculate_forces()
        o.calculate_forces()
    assert total > 50


@pytest.mark.parametrize("handler", [0, 1])
def test_parallel_gillespie_scene_on_gpu(handler):
    g = GOLDEN["parallel_gillespie_scene"]
    tid = {t: i for i, t in enumerate(g["types"])}
    rxs = [dict(kind=L.RX_FUSION, educts=[tid[x["educts"][0]], tid[x["educts"][1]]], products=[tid[x["product"]]],
                rate=x["rate"], educt_distance=x["radius"]) for x in g["reactions"]]
    pos = np.array([p[1] for p in g["particles"]], dtype=float)
    types = np.array([tid[p[0]] for p in g["particles"]], dtype=np.uint32)
    e, r = _engine(3, g["box"], g["pbc"], [], rxs, diffusion=[0.25] * 3, n=len(types))
    e.set_particles(r, pos, types)
    e.react(handler, g["timestep"])
    p, t, _ = e.get_particles(r)
    assert len(t) == g["expected_count"]
    have = [(int(a), tuple(b)) for a, b in zip(t, p)]
    for name, q in g["expected_exact"]:
        assert (tid[name], tuple(float(x) for x in q)) in have
    for name, q in g["expected_close"]:
        assert any(a == tid[name] and np.allclose(b, q, rtol=1e-14) for a, b in have)


@pytest.mark.parametrize("handler", [0, 1])
def test_decay_scene_on_gpu(handler):
    g = GOLDEN["decay_scene"]
    rxs = [dict(kind=L.RX_DECAY, educts=[0], rate=g["decay_rate"]),
           dict(kind=L.RX_FISSION, educts=[0], products=[0, 0], rate=g["fission_rate"],
                product_distance=g["fission_distance"])]
    e, r = _engine(1, g["box"], g["pbc"], [], rxs, diffusion=[g["diffusion"]], n=4 * g["n_particles"])
    e.set_particles(r, np.zeros((g["n_particles"], 3)), np.zeros(g["n_particles"], dtype=np.uint32))
    for t in range(g["steps"]):
        e.calculate_forces()
        e.integrate(g["timestep"])
        e.react(handler, g["timestep"])
    assert e.n_particles(r) == g["expected_remaining"]


def test_fusion_through_boundary_and_fission_near_boundary_on_gpu():
    g = GOLDEN["fusion_through_periodic_boundary"]
    for handler in (0, 1):
        rx = [dict(kind=L.RX_FUSION, educts=[0, 0], products=[0], rate=g["rate"], educt_distance=g["radius"])]
        e, r = _engine(1, g["box"], g["pbc"], [], rx, diffusion=[0.0], n=2)
        e.set_particles(r, g["particles"], np.zeros(2, dtype=np.uint32))
        e.react(handler, 1.0)
        p, _, _ = e.get_particles(r)
        assert len(p) == 1
        np.testing.assert_allclose(p[0], g["product"], rtol=g["product_rel_tolerance"])
    g = GOLDEN["fission_near_boundary"]
    for seed in range(4):
        rx = [dict(kind=L.RX_FISSION, educts=[0], products=[0, 0], rate=g["rate"],
                   product_distance=g["product_distance"])]
        e, r = _engine(1, g["box"], g["pbc"], [], rx, diffusion=[0.0], seed=seed, n=2)
        e.set_particles(r, [g["particle"]], np.zeros(1, dtype=np.uint32))
        e.react(0, 1.0)
        p, _, _ = e.get_particles(r)
        assert len(p) == 2 and ((p[0][0] <= 0) != (p[1][0] <= 0))


def _c0_engine(graph=False, n=2000, seed=11, max_replicas=1, **kw):
    pots, rxs = _c0_chemistry()
    pos, types = _system_c0like(n, 20.0, seed=3)
    e = Engine(3, [1.0, 1.0, 0.5], potentials=pots, reactions=rxs, max_replicas=max_replicas, **kw)
    return e, pos, types


@pytest.mark.parametrize("lds,graph", [(128, False), (256, True)])
def test_bin_reading_bricks_equal_staged_bricks(lds, graph):
    """Bricks left to the bin-reading kernels give the staged kernels' trajectory bit for bit
    (fixed-point force sums, the same pair arithmetic, the same candidate rule), also under graph
    replay, where the brick lists are cleared by the kernels themselves."""
    e1, pos, types = _c0_engine()
    r1 = e1.add_replica([20.0] * 3, seed=6, max_particles=8000)
    e1.set_particles(r1, pos, types)
    e2, _, _ = _c0_engine(lds_states=lds)
    r2 = e2.add_replica([20.0] * 3, seed=6, max_particles=8000)
    e2.set_particles(r2, pos, types)
    for chunk in (7, 13):
        e1.run(0, 0.01, chunk)
        e2.run(0, 0.01, chunk, graph=graph)
        for x, y in zip(e1.get_particles(r1), e2.get_particles(r2)):
            assert np.array_equal(x, y)
        assert e1.energy(r1) == e2.energy(r2)
        assert np.array_equal(e1.reaction_totals(r1), e2.reaction_totals(r2))


def test_fused_loop_equals_separate_actions():
    """rdy_run (fused kernels) reproduces integrate / NL / react / NL / forces exactly."""
    e1, pos, types = _c0_engine()
    r1 = e1.add_replica([20.0] * 3, seed=5, max_particles=8000)
    e1.set_particles(r1, pos, types)
    e1.run(0, 0.01, 30)
    e2, _, _ = _c0_engine()
    r2 = e2.add_replica([20.0] * 3, seed=5, max_particles=8000)
    e2.set_particles(r2, pos, types)
    e2.calculate_forces()
    for t in range(30):
        e2.integrate(0.01)
        e2.update_neighbor_list()
        e2.react(0, 0.01)
        e2.update_neighbor_list()
        e2.calculate_forces()
    a = e1.get_particles(r1)
    b = e2.get_particles(r2)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert e1.energy(r1) == e2.energy(r2)
    assert np.array_equal(e1.reaction_totals(r1), e2.reaction_totals(r2))


@pytest.mark.parametrize("handler", [0, 1])
def test_fused_loop_equals_separate_actions_every_reaction_kind(handler):
    """A high-rate chemistry with every reaction kind (decay, conversion, enzymatic, fusion, fission)
    kills, inserts and retypes tens of particles per step, often next to each other and across
    species with different cutoffs.  The fused step stores the forces of the walk before the
    reactions and corrects them around every edit; the result must equal integrate / NL / react /
    NL / forces bit for bit, under UA and under Gillespie."""
    pots = [(L.POT_HARMONIC_REPULSION, a, b, [10.0, 0.8 + 0.1 * (a + b)]) for a in range(4) for b in range(a, 4)]
    pots.append((L.POT_WEAK_INTERACTION, 0, 3, [10.0, 1.0, 0.5, 1.6]))
    rxs = [dict(kind=L.RX_FUSION, educts=[0, 1], products=[2], rate=5.0, educt_distance=1.0),
           dict(kind=L.RX_FISSION, educts=[2], products=[0, 1], rate=3.0, product_distance=0.9),
           dict(kind=L.RX_ENZYMATIC, educts=[1, 2], products=[3, 2], rate=4.0, educt_distance=1.2),
           dict(kind=L.RX_CONVERSION, educts=[3], products=[0], rate=2.0),
           dict(kind=L.RX_DECAY, educts=[3], rate=0.5)]
    diff = [1.0, 0.8, 0.5, 1.2]
    rng = np.random.default_rng(17)
    n, Lb, dt, steps = 2400, 16.0, 0.02, 25
    pos = (rng.random((n, 3)) - 0.5) * Lb
    types = rng.choice(np.arange(4, dtype=np.uint32), n)

    def make():
        e = Engine(4, diff, potentials=pots, reactions=rxs)
        r = e.add_replica([Lb] * 3, seed=31, max_particles=2 * n)
        e.set_particles(r, pos, types)
        return e, r

    e1, r1 = make()
    e1.run(handler, dt, steps)
    e2, r2 = make()
    e2.calculate_forces()
    for _ in range(steps):
        e2.integrate(dt)
        e2.update_neighbor_list()
        e2.react(handler, dt)
        e2.update_neighbor_list()
        e2.calculate_forces()
    for x, y in zip(e1.get_particles(r1), e2.get_particles(r2)):
        assert np.array_equal(x, y)
    assert e1.energy(r1) == e2.energy(r2)
    f1, h1 = e1.get_forces(r1)
    f2, h2 = e2.get_forces(r2)
    o1, o2 = np.argsort(h1), np.argsort(h2)
    assert np.array_equal(h1[o1], h2[o2]) and np.array_equal(f1[o1], f2[o2])
    totals = e1.reaction_totals(r1)
    assert np.array_equal(totals, e2.reaction_totals(r2))
    assert np.all(totals > 0) and totals.sum() > 40 * steps  # every kind fired, tens per step


def test_graph_replay_equals_eager():
    """Graph replays, with host reads in between, track an eager engine step for step."""
    e1, pos, types = _c0_engine()
    r1 = e1.add_replica([20.0] * 3, seed=8, max_particles=8000)
    e1.set_particles(r1, pos, types)
    e2, _, _ = _c0_engine()
    r2 = e2.add_replica([20.0] * 3, seed=8, max_particles=8000)
    e2.set_particles(r2, pos, types)
    for chunk in (10, 20, 11, 40):
        e1.run(1, 0.01, chunk)
        e2.run(1, 0.01, chunk, graph=True)
        a = e1.get_particles(r1)
        b = e2.get_particles(r2)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        assert np.array_equal(e1.count_types(1), e2.count_types(1))


def test_replica_alone_equals_replica_in_batch():
    """Deterministic fixed-point forces and handle-keyed Philox make a replica batch-invariant."""
    pots, rxs = _c0_chemistry()
    e1 = Engine(3, [1.0, 1.0, 0.5], potentials=pots, reactions=rxs)
    pos, types = _system_c0like(1000, 16.0, seed=21)
    r = e1.add_replica([16.0] * 3, seed=77, max_particles=4000)
    e1.set_particles(r, pos, types)
    e1.run(0, 0.01, 25)
    alone = e1.get_particles(r)
    eb = Engine(3, [1.0, 1.0, 0.5], potentials=pots, reactions=rxs, max_replicas=6, max_cells=6 * 16 ** 3,
                max_particles=40000)
    ids = []
    for k in range(6):
        L_k = [12.0, 16.0, 20.0, 16.0, 9.0, 14.0][k]
        rk = eb.add_replica([L_k] * 3, seed=77 if k == 3 else 1000 + k)
        ids.append(rk)
        if k == 3:
            eb.set_particles(rk, pos, types)
        else:
            p, t = _system_c0like(int(0.25 * L_k ** 3) // 2 * 2, L_k, seed=k)
            eb.set_particles(rk, p, t)
    eb.run(0, 0.01, 25)
    mixed = eb.get_particles(ids[3])
    for x, y in zip(alone, mixed):
        assert np.array_equal(x, y)


def test_c1_full_size_invariants():
    """1M particles, C1 chemistry: A+C and B+C conserved, positions wrapped, energy finite."""
    n, Lb = 1_000_000, 200.0
    rng = np.random.default_rng(1)
    pos = (rng.random((n, 3)) - 0.5) * Lb
    types = rng.choice(np.array([0, 1, 2], dtype=np.uint32), n, p=[0.4, 0.4, 0.2])
    pots = [(L.POT_HARMONIC_REPULSION, a, b, [10.0, 1.0]) for a in range(3) for b in range(a, 3)]
    pots.append((L.POT_WEAK_INTERACTION, 0, 0, [10.0, 1.0, 0.5, 2.0]))
    _, rxs = _c0_chemistry()
    e = Engine(3, [1.0, 1.0, 0.5], potentials=pots, reactions=rxs)
    r = e.add_replica([Lb] * 3, seed=1, max_particles=n)
    e.set_particles(r, pos, types)
    c0 = e.count_types(1)[0].astype(np.int64)
    e.run(0, 0.005, 20)
    c1 = e.count_types(1)[0].astype(np.int64)
    assert c0[0] + c0[2] == c1[0] + c1[2] and c0[1] + c0[2] == c1[1] + c1[2]
    assert np.isfinite(e.energy(r))
    assert e.reaction_totals(r).sum() > 100
    p, _, _ = e.get_particles(r)
    assert np.all(p >= -Lb / 2) and np.all(p < Lb / 2)


def test_ensemble_batched_equals_plain():
    """Replicas of uneven size packed into one engine step exactly like one engine per replica
    (same Philox keys; fixed-point forces make the result independent of the packing), with the
    ensemble's own cell size and bin capacity."""
    from readdy_amd import ensemble as E
    n, box = E.sample(12, seed=11, n_min=200, n_max=3000)
    b = E.Batched(n, box, nl_cutoff=2.0)
    p = E.Plain(n, box, nl_cutoff=1.0)
    b.run(15, graph=True)
    for e, _ in p.engines:
        e.run(L.HANDLER_UA, E.DT, 15)
    for rid, (e, r) in zip(b.replicas, p.engines):
        xb, tb, hb = b.engine.get_particles(rid)
        xp, tp, hp = e.get_particles(r)
        assert np.array_equal(hb, hp) and np.array_equal(tb, tp)
        assert np.array_equal(xb, xp)
    b.close()
    p.close()


def test_set_particles_replaces_and_fresh_replicas_keep_their_neighbours():
    """A second upload into a replica drops its previous particles; uploading into a fresh replica (no
    drop pass) leaves the other replicas' particles alone."""
    rng = np.random.default_rng(12)
    e = Engine(2, [1.0, 1.0], potentials=[(L.POT_HARMONIC_REPULSION, 0, 1, [10.0, 1.0])], max_replicas=3)
    r0 = e.add_replica([10.0] * 3, seed=1, max_particles=1000)
    e.set_particles(r0, (rng.random((300, 3)) - 0.5) * 10, rng.integers(0, 2, 300).astype(np.uint32))
    r1 = e.add_replica([12.0] * 3, seed=2, max_particles=1000)
    p1 = (rng.random((200, 3)) - 0.5) * 12
    e.set_particles(r1, p1, np.zeros(200, np.uint32))
    assert e.n_particles(r0) == 300 and e.n_particles(r1) == 200
    p0 = (rng.random((120, 3)) - 0.5) * 10
    e.set_particles(r0, p0, np.ones(120, np.uint32))  # replaces r0's 300
    x0, t0, _ = e.get_particles(r0)
    assert len(t0) == 120 and np.all(t0 == 1) and np.array_equal(x0, p0)
    x1, _, _ = e.get_particles(r1)
    assert np.array_equal(x1, p1)
    assert np.array_equal(e.count_types(2), [[0, 120], [200, 0]])
