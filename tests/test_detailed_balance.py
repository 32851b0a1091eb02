"""DetailedBalance: the reversible-pair configuration (product C ABI and oracle) against the
reference's known answers, and the oracle's handler against the reference's integration scenes and
closed forms.  No GPU: rdy_reversible_reactions is a host computation.

Reference: readdy/test/TestDetailedBalance.cpp (configs :58-212, search :214-341, scenes :364-834),
readdy/main/model/Actions.cpp:67-519, kernels/singlecpu/src/actions/SCPUReactionImpls.cpp:407-813.
"""
import math

import numpy as np
import pytest
from scipy import integrate as si
from scipy import stats

from oracle import pyoracle as po
from readdy_amd import _lib as L
from readdy_amd import engine as E

A, B, C, D = 0, 1, 2, 3
HR = L.POT_HARMONIC_REPULSION


def fusion(e1, e2, p, rate, radius, w1=0.5, w2=0.5):
    return dict(kind=L.RX_FUSION, educts=[e1, e2], products=[p], rate=rate, educt_distance=radius, weight1=w1,
                weight2=w2)


def fission(e, p1, p2, rate, distance, w1=0.5, w2=0.5):
    return dict(kind=L.RX_FISSION, educts=[e], products=[p1, p2], rate=rate, product_distance=distance, weight1=w1,
                weight2=w2)


def conversion(e, p, rate):
    return dict(kind=L.RX_CONVERSION, educts=[e], products=[p], rate=rate)


def enzymatic(frm, cat, to, rate, radius):
    """"frm +(radius) cat -> to + cat" (educts {from, catalyst}, products {to, catalyst})"""
    return dict(kind=L.RX_ENZYMATIC, educts=[frm, cat], products=[to, cat], rate=rate, educt_distance=radius)


def oracle_system(n_types, box, pots, rxs, kbt=1.0, diffusion=None):
    s = po.System(n_types, box, (1, 1, 1), kbt=kbt, diffusion=diffusion)
    for kind, a, b, p in pots:
        s.add_pair_potential(kind, a, b, p)
    for r in rxs:
        s.add_reaction(r["kind"], r["educts"], r.get("products", []), r["rate"], r.get("educt_distance", 0.0),
                       r.get("product_distance", 0.0), r.get("weight1", 0.5), r.get("weight2", 0.5))
    return s


def both(n_types, box, pots, rxs, kbt=1.0):
    """the reversible pairs from the product (C ABI) and from the oracle, as lists of dicts"""
    vol = float(np.prod(box))
    prod = E.reversible_reactions(n_types, pots, rxs, kbt, vol)
    orc = oracle_system(n_types, box, pots, rxs, kbt).reversible_configs()
    return prod, orc


def approx(a, b, eps=1e-4):  # Catch::Approx(b).epsilon(eps)
    return abs(a - b) <= eps * max(abs(a), abs(b))


# ---- ReversibleReactionConfig known answers (TestDetailedBalance.cpp:65-158) ----
def test_config_fusion_fission_known_answers():
    L0 = 21.544346900318832
    pots = [(HR, A, B, [2., 3.])]
    rxs = [fusion(A, B, C, 1., 5.), fission(C, A, B, 2., 5.)]
    prod, orc = both(3, [L0] * 3, pots, rxs)
    assert len(prod) == 1 and len(orc) == 1
    p, o = prod[0], orc[0]
    assert p["kind"] == 0 and o["kind"] == 0
    assert (p["forward"], p["backward"]) == (0, 1) == (o["forward"], o["backward"])
    assert p["n_lhs"] == 2 and p["n_rhs"] == 1
    assert p["lhs"] == [A, B] and p["rhs"][0] == C
    assert approx(p["total_volume"], 10000.) and p["kbt"] == 1.
    assert p["lhs_interaction_radius"] == 3. and p["reaction_radius"] == 5.
    for key, okey, want in [("lhs_interaction_volume", "lhs_volume", 113.09733),
                            ("eff_lhs_interaction_volume", "eff_lhs_interaction", 68.09910),
                            ("eff_lhs_reaction_volume", "eff_lhs_reaction", 478.60057),
                            ("equilibrium_constant", "equilibrium_constant", 41.60045),
                            ("macro_forward_rate", "macro_forward", 480.76392)]:
        assert approx(p[key], want), (key, p[key])
        assert approx(o[okey], want), (okey, o[okey])
        # the two quadratures (adaptive Gauss-Kronrod, adaptive Simpson) agree far below the known answers' digits
        assert approx(p[key], o[okey], 1e-9), (key, p[key], o[okey])
    assert p["macro_backward_rate"] == 2. and p["acceptance_prefactor"] == 1.


def test_config_conversion_known_answers():
    prod, orc = both(3, [10.] * 3, [], [conversion(A, B, 1.), conversion(B, A, 2.)])
    p, o = prod[0], orc[0]
    assert p["kind"] == 1 == o["kind"]
    assert approx(p["total_volume"], 1000.)
    assert p["n_lhs"] == 1 and p["n_rhs"] == 1 and p["lhs"][0] == A and p["rhs"][0] == B
    for k in ("lhs_interaction_radius", "lhs_interaction_volume", "rhs_interaction_radius", "rhs_interaction_volume",
              "eff_lhs_interaction_volume", "eff_lhs_reaction_volume", "eff_rhs_interaction_volume",
              "eff_rhs_reaction_volume"):
        assert p[k] == 0., k
    assert p["acceptance_prefactor"] == 1.
    assert p["macro_forward_rate"] == 1. and p["macro_backward_rate"] == 2. and p["equilibrium_constant"] == 2.
    assert o["equilibrium_constant"] == 2. and o["eff_lhs_reaction"] == 0.


def test_config_enzymatic_known_answers():
    pots = [(HR, A, C, [2., 3.]), (HR, B, C, [2., 3.])]
    rxs = [enzymatic(A, C, B, 1., 5.), enzymatic(B, C, A, 2., 5.)]
    prod, orc = both(3, [10.] * 3, pots, rxs)
    p, o = prod[0], orc[0]
    assert p["kind"] == 2 == o["kind"]
    assert p["lhs"] == [A, C] and p["rhs"] == [B, C]
    assert p["lhs_interaction_radius"] == 3. and p["rhs_interaction_radius"] == 3.
    v = 4. / 3. * math.pi * 27.
    assert approx(p["lhs_interaction_volume"], v, 1e-12) and approx(p["rhs_interaction_volume"], v, 1e-12)
    for k in ("eff_lhs_interaction_volume", "eff_rhs_interaction_volume"):
        assert approx(p[k], 68.09910), k
    for k in ("eff_lhs_reaction_volume", "eff_rhs_reaction_volume"):
        assert approx(p[k], 478.60057), k
    assert approx(p["acceptance_prefactor"], 1.)
    assert approx(p["macro_forward_rate"], 501.15150) and approx(p["macro_backward_rate"], 1002.30300)
    assert p["equilibrium_constant"] == 2.
    assert approx(o["macro_forward"], 501.15150) and approx(o["equilibrium_constant"], 2., 1e-12)


def test_config_closed_forms_without_potentials():
    """no lhs potentials: the effective volumes are the plain sphere volumes 4/3 pi R^3"""
    prod, orc = both(3, [10.] * 3, [], [enzymatic(A, C, B, 1., 4.), enzymatic(B, C, A, 2., 4.)])
    v = 4. / 3. * math.pi * 64.
    assert approx(prod[0]["eff_lhs_reaction_volume"], v, 1e-12) and approx(orc[0]["eff_lhs_reaction"], v, 1e-12)
    assert prod[0]["acceptance_prefactor"] == 1.  # TestDetailedBalance.cpp:202-210


def test_fission_cdf_against_direct_quadrature():
    """the oracle's fission-distance table against scipy's quadrature of 4 pi r^2 exp(-U(r))"""
    k, d = 2., 3.
    s = oracle_system(3, [20.] * 3, [(HR, A, B, [k, d])], [fusion(A, B, C, 1., 5.), fission(C, A, B, 2., 5.)])
    s.reversible_configs()
    radii, cdf = s.fission_table(0)
    assert len(radii) == 1001 and radii[0] == 0. and abs(radii[-1] - 5.) < 1e-12

    def f(r):
        u = 0.5 * k * (r - d) ** 2 if r < d else 0.
        return 4 * math.pi * r * r * math.exp(-u)

    total = si.quad(f, 0., 5., points=[d], epsabs=0, epsrel=1e-13, limit=200)[0]
    for i in (1, 100, 333, 600, 999):
        want = si.quad(f, 0., radii[i], points=[d] if radii[i] > d else None, epsabs=0, epsrel=1e-13, limit=200)[0]
        assert abs(cdf[i] - want / total) < 1e-11, (i, cdf[i], want / total)
    assert cdf[-1] == 1.


# ---- searchReversibleReactions (TestDetailedBalance.cpp:222-341) ----
def _kinds(prod):
    return [(p["kind"], p["forward"], p["backward"]) for p in prod]


@pytest.mark.parametrize("case", ["enzymatic", "fusion_fission", "fission_fusion", "michaelis_menten", "many_mixed",
                                  "wrong_weights"])
def test_search_reversible_reactions(case):
    if case == "enzymatic":
        rxs, want = [enzymatic(A, C, B, 1., 4.), enzymatic(B, C, A, 2., 4.)], [(2, 0, 1)]
    elif case == "fusion_fission":
        rxs, want = [fusion(A, B, C, 1., 2.), fission(C, A, B, 2., 2.)], [(0, 0, 1)]
    elif case == "fission_fusion":  # fission first, the fusion's educts in the other order
        rxs, want = [fission(C, A, B, 2., 2.), fusion(B, A, C, 1., 2.)], [(0, 1, 0)]
    elif case == "michaelis_menten":  # A + B <-> C <-> A + D, and B -> D
        rxs = [fusion(A, B, C, 2., 2.), fission(C, A, B, 1., 2.), fission(C, A, D, 3., 3.), fusion(A, D, C, 4., 3.),
               conversion(B, D, 5.)]
        want = [(0, 0, 1), (0, 3, 2)]
    elif case == "many_mixed":
        rxs = [fusion(A, B, C, 2., 2.), fission(C, A, B, 1., 2.), fusion(A, C, B, 4., 3.),
               enzymatic(A, B, C, 3., 2.), enzymatic(C, B, A, 5., 2.), enzymatic(B, A, C, 6., 4.),
               conversion(A, B, 7.), conversion(B, A, 8.), conversion(C, B, 9.)]
        want = [(0, 0, 1), (1, 6, 7), (2, 3, 4)]
    else:  # wrong weights: no pair
        rxs, want = [fusion(A, B, C, 1., 2., 0.5, 0.5), fission(C, A, B, 1., 2., 0.3, 0.7)], []
    prod, orc = both(4, [10.] * 3, [], rxs)
    assert _kinds(prod) == want
    assert [(o["kind"], o["forward"], o["backward"]) for o in orc] == want


@pytest.mark.parametrize("case", ["duplicate", "species"])
def test_search_rejects_ambiguous_pairs(case):
    if case == "duplicate":  # TestDetailedBalance.cpp:322-327
        rxs = [fusion(A, B, C, 2., 2.), fission(C, A, B, 1., 2.), fusion(A, B, C, 2., 2.)]
        msg = "Duplicate"
    else:  # A + B -> C with C -> A + A: paired by the search's repeated comparison, then not reversible
        rxs = [fusion(A, B, C, 1., 2.), fission(C, A, A, 2., 2.)]
        msg = "particle species"
    with pytest.raises(L.EngineError, match=msg):
        E.reversible_reactions(3, [], rxs, 1.0, 1000.)
    with pytest.raises(RuntimeError, match=msg):
        oracle_system(3, [10.] * 3, [], rxs).reversible_configs()


# ---- the oracle's handler ----
R2 = 1.2599210498948732
SCENE_POTS = [(HR, A, B, [10., 2.]), (HR, B, B, [10., 2.]), (HR, A, A, [10., 2.]), (HR, A, C, [10., 1 + R2]),
              (HR, B, C, [10., 1 + R2]), (HR, C, C, [10., 2 * R2])]
NO_INTEGRATION = 0xFFFFFFFF  # the step counter of a replica that has not been integrated yet


def scene(kind, accept, seed=3):
    """TestDetailedBalance.cpp:364-746: one reaction next to 500 D particles ~ N(0, 1) in a 12^3 box"""
    pots = list(SCENE_POTS)
    if kind == "conversion":
        rxs = [conversion(A, B, 1e15), conversion(B, A, 1e-15)]
        pots.append((HR, A if accept else B, D, [10., 2.]))
        pos, ty = [[0., 0., 0.]], [A]
    elif kind == "fusion":
        rxs = [fusion(A, B, C, 1e15, 2.), fission(C, A, B, 1e-15, 2.)]
        pots += [(HR, A, D, [10., 2.]), (HR, B, D, [10., 2.])] if accept else [(HR, C, D, [10., 2.])]
        pos, ty = [[.1, .1, .1], [-.1, -.1, -.1]], [A, B]
    elif kind == "fission":
        rxs = [fusion(A, B, C, 1e-15, 2.), fission(C, A, B, 1e15, 2.)]
        pots += [(HR, C, D, [10., 2.])] if accept else [(HR, A, D, [10., 2.]), (HR, B, D, [10., 2.])]
        pos, ty = [[-.1, -.1, -.1]], [C]
    else:  # enzymatic, C the catalyst
        rxs = [enzymatic(A, C, B, 1e15, 2.), enzymatic(B, C, A, 1e-15, 2.)]
        pots.append((HR, A if accept else B, D, [10., 2.]))
        pos, ty = [[0., 0., 0.], [0., 0., 1.]], [A, C]
    d = np.random.default_rng(seed).normal(0., 1., (500, 3))
    return pots, rxs, np.vstack([pos, d]), np.array(ty + [D] * 500, dtype=np.int32)


# performed events, final counts of A, B, C, and the largest id after one call (ids: 0 .. n-1 uploaded;
# a reversible event takes ids for its products and then for the educts' backups, accepted or not)
SCENE_OUTCOME = {
    ("conversion", True): (1, [0, 1, 0], 501), ("conversion", False): (0, [1, 0, 0], 502),
    ("fusion", True): (1, [0, 0, 1], 502), ("fusion", False): (0, [1, 1, 0], 504),
    ("fission", True): (1, [1, 1, 0], 502), ("fission", False): (0, [0, 0, 1], 503),
    ("enzymatic", True): (1, [0, 1, 1], 502), ("enzymatic", False): (0, [1, 0, 1], 503),
}


@pytest.mark.parametrize("kind,accept", list(SCENE_OUTCOME))
def test_oracle_reference_scenes(kind, accept):
    pots, rxs, pos, ty = scene(kind, accept)
    s = oracle_system(4, [12.] * 3, pots, rxs, diffusion=[1, 1, 1, 0])
    s.add_particles(pos, ty, ids=np.arange(len(ty)))
    s.set_id_counter(len(ty))
    s.set_seed(3)
    s.nl_setup(s.max_cutoff())
    s.set_record_reactions(True)
    n = s.detailed_balance(0.1, NO_INTEGRATION)
    _, t, ids, _ = s.particles()
    want_n, want_counts, want_max = SCENE_OUTCOME[(kind, accept)]
    assert n == want_n
    assert np.bincount(t, minlength=4)[:3].tolist() == want_counts and np.sum(t == D) == 500
    assert int(ids.max()) == want_max
    assert sorted(s.reaction_counts().tolist()) == [0, want_n]
    recs = s.reaction_records()
    assert len(recs) == want_n
    if want_n:  # the reversible path fills no products (SCPUReactionImpls.cpp:579-742)
        assert recs["products"].tolist() == [[0, 0]]
        if kind == "enzymatic":  # educt first, then the catalyst
            assert recs["educts"].tolist() == [[0, 1]] and recs["types_from"].tolist() == [[A, C]]


def _isolated_pairs(n, spacing, offset):
    """n pairs (first at the lattice point, second at +offset along x) on a cubic lattice"""
    m = int(math.ceil(n ** (1 / 3)))
    g = np.array([(i, j, k) for i in range(m) for j in range(m) for k in range(m)][:n], dtype=float)
    box = m * spacing
    first = g * spacing - box / 2 + 0.5 * spacing
    second = first + np.array([offset, 0., 0.])
    return box, first, second


def test_oracle_metropolis_ratio_of_a_conversion():
    """A -> B next to a D at distance 1, HR(B, D) k=2 d=2: dE = 1, acceptance e^-1 (2000 independent pairs)"""
    n = 2000
    box, a, d = _isolated_pairs(n, 6.0, 1.0)
    s = oracle_system(4, [box] * 3, [(HR, B, D, [2., 2.])], [conversion(A, B, 1e9), conversion(B, A, 1e-9)])
    s.add_particles(np.vstack([a, d]), np.array([A] * n + [D] * n, dtype=np.int32), ids=np.arange(2 * n))
    s.set_id_counter(2 * n)
    s.set_seed(21)
    s.nl_setup(s.max_cutoff())
    got = s.detailed_balance(1.0, 4)
    p = math.exp(-1.0)
    assert abs(got - n * p) < 5 * math.sqrt(n * p * (1 - p)), got


def test_oracle_enzymatic_prefactor_only_forward():
    """A + C -> B + C with HR(A, C) k=2 d=3 and no B-C potential: the forward move's acceptance is
    prefactor * exp(U_AC(r)) with prefactor = V_eff,lhs / V_eff,rhs < 1; the backward B + C -> A + C gets
    prefactor 1 (SCPUReactionImpls.cpp:482-489 tests the forward educt order in both branches), so its
    acceptance is exp(-U_AC(r))"""
    k, dd, R, r, n = 2., 3., 2., 1.9, 2000
    u = 0.5 * k * (r - dd) ** 2
    veff = si.quad(lambda x: 4 * math.pi * x * x * math.exp(-0.5 * k * (x - dd) ** 2), 0, R, epsrel=1e-13)[0]
    pref = veff / (4. / 3. * math.pi * R ** 3)
    box, a, c = _isolated_pairs(n, 8.0, r)
    for forward in (True, False):
        rxs = [enzymatic(A, C, B, 1e9 if forward else 1e-9, R), enzymatic(B, C, A, 1e-9 if forward else 1e9, R)]
        s = oracle_system(3, [box] * 3, [(HR, A, C, [k, dd])], rxs)
        cfg = s.reversible_configs()[0]
        assert abs(cfg["acceptance_prefactor"] - pref) < 1e-10
        first = A if forward else B
        s.add_particles(np.vstack([a, c]), np.array([first] * n + [C] * n, dtype=np.int32), ids=np.arange(2 * n))
        s.set_id_counter(2 * n)
        s.set_seed(5)
        s.nl_setup(s.max_cutoff())
        got = s.detailed_balance(1.0, 2)
        p = min(1., pref * math.exp(u)) if forward else math.exp(-u)
        assert 0.05 < p < 0.95
        assert abs(got - n * p) < 5 * math.sqrt(n * p * (1 - p)), (forward, got, n * p)


def test_oracle_fission_distances_follow_the_boltzmann_weighted_volume():
    """C -> A + B with HR(A, B): the products' mutual energy is removed from the Metropolis ratio, so every
    isolated fission is accepted and the A-B distances follow the CDF of 4 pi r^2 exp(-U(r)) on [0, R]"""
    k, dd, R, n = 2., 3., 2.5, 2000
    box, c, _ = _isolated_pairs(n, 8.0, 0.)
    s = oracle_system(3, [box] * 3, [(HR, A, B, [k, dd])], [fusion(A, B, C, 1e-9, R), fission(C, A, B, 1e9, R)])
    s.add_particles(c, np.array([C] * n, dtype=np.int32), ids=np.arange(n))
    s.set_id_counter(n)
    s.set_seed(8)
    s.nl_setup(s.max_cutoff())
    assert s.detailed_balance(1.0, 0) == n
    pos, t, ids, _ = s.particles()
    assert np.sum(t == A) == n and np.sum(t == B) == n
    # each fission's products got consecutive ids (products[1] = B first, then A), pairs share the id pair
    order = np.argsort(ids)
    p = pos[order]
    dist = np.linalg.norm(p[0::2] - p[1::2], axis=1)
    f = lambda x: 4 * math.pi * x * x * math.exp(-0.5 * k * (x - dd) ** 2)
    tot = si.quad(f, 0, R, epsrel=1e-12)[0]
    cdf = np.vectorize(lambda x: si.quad(f, 0, min(max(x, 0.), R), epsrel=1e-10)[0] / tot)
    # the table is discrete (1000 radii): distances sit on the grid, shift by half a cell for the KS test
    res = stats.kstest(dist - 0.5 * R / 1000, cdf)
    assert res.pvalue > 1e-4, res


def test_oracle_without_potentials_equals_gillespie():
    """no potentials: every reversible move has acceptance 1, so the handler performs exactly the events of
    the keyed Gillespie selection (fission distances and ids differ, the counts and species do not)"""
    rxs = [conversion(A, B, 3.), conversion(B, A, 2.), enzymatic(A, C, B, 4., 1.), enzymatic(B, C, A, 1., 1.),
           fusion(A, B, D, 2., 1.), fission(D, A, B, 1., 1.), conversion(C, A, 0.5)]
    rng = np.random.default_rng(4)
    pos = rng.uniform(-5, 5, (600, 3))
    ty = rng.integers(0, 4, 600).astype(np.int32)
    out = []
    for handler in ("db", "gillespie"):
        s = oracle_system(4, [10.] * 3, [], rxs)
        s.add_particles(pos, ty, ids=np.arange(600))
        s.set_id_counter(600)
        s.set_seed(77)
        s.nl_setup(s.max_cutoff())
        n = s.detailed_balance(0.05, 3) if handler == "db" else s.reactions(po.GILLESPIE_KEYED, 0.05, 3)
        _, t, _, _ = s.particles()
        out.append((n, s.reaction_counts().tolist(), np.bincount(t, minlength=4).tolist()))
    assert out[0] == out[1]
    assert out[0][0] > 30


@pytest.mark.parametrize("kind,params", [
    (L.POT_WEAK_INTERACTION, [10., 1.0, 0.8, 1.6]),
    (L.POT_LENNARD_JONES, [12., 6., 1.6, 1., 0.5, 0.8]),
    (L.POT_LENNARD_JONES, [12., 6., 1.6, 0., 0.5, 0.8]),
    (L.POT_SCREENED_ELECTROSTATICS, [-1.0, 1.0, 0.5, 0.8, 6., 1.6]),
])
def test_configs_of_every_potential_kind(kind, params):
    """the effective volumes of each potential kind: the C ABI (adaptive Gauss-Kronrod) against the oracle
    (adaptive Simpson) and scipy's quadrature of 4 pi r^2 exp(-U(r)) with the oracle's U"""
    pots = [(kind, A, B, params), (kind, B, C, params)]
    rxs = [fusion(A, B, C, 1., 1.4), fission(C, A, B, 2., 1.4), enzymatic(A, C, D, 1., 1.4),
           enzymatic(D, C, A, 2., 1.4)]
    prod, orc = both(4, [10.] * 3, pots, rxs)
    assert [p["kind"] for p in prod] == [0, 2] == [o["kind"] for o in orc]
    u = lambda r: po.pair_force_energy(kind, params, [r, 0., 0.])[1]
    f = lambda r: 4 * math.pi * r * r * math.exp(-u(r)) if r > 0 else 0.
    want = si.quad(f, 1e-9, 1.4, epsabs=0, epsrel=1e-12, limit=500)[0]
    for p, o in zip(prod, orc):
        assert approx(p["eff_lhs_reaction_volume"], o["eff_lhs_reaction"], 1e-9)
        assert approx(p["eff_lhs_interaction_volume"], o["eff_lhs_interaction"], 1e-9)
    assert approx(prod[0]["eff_lhs_reaction_volume"], want, 1e-8)
