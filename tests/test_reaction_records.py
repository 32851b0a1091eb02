"""Reaction records (ReactionRecord.h:54-63) as the CPU kernel's performReaction fills them
(ReactionUtils.h:155-225): reaction id and type, educts' ids and types before the reaction, products'
ids after it, position = (pos1 + pos2) / 2 with fixPosition.

CPU: the oracle's records against the reference's own scene (test_observables_io.py:323-392: a
conversion at the origin, a fusion at (1.05, 1, 1)) and, for every reaction kind, against the states
before and after each reaction call.  GPU: the same scene, the engine's records equal to the oracle's
event for event, and the fused (eager and graph) step's records equal to the separate actions'.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

CONV, FUS, FIS, ENZ, DEC = po.RX_CONVERSION, po.RX_FUSION, po.RX_FISSION, po.RX_ENZYMATIC, po.RX_DECAY

# test_observables_io.py:325-339: types A, B, C with D = 0; "mylabel" A->B at 1e-5, "A->B" at 1e16,
# "B+C->A" fusion at 1e16 within 1.0, weights 0.5
SCENE_RXS = [dict(kind=CONV, educts=[0], products=[1], rate=1e-5),
             dict(kind=CONV, educts=[0], products=[1], rate=1e16),
             dict(kind=FUS, educts=[1, 2], products=[0], rate=1e16, educt_distance=1.0)]
SCENE_POS = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [1.1, 1.0, 1.0]])
SCENE_TYPES = np.array([0, 1, 2], dtype=np.uint32)

# every reaction kind, high rates
ALL_RXS = [dict(kind=FUS, educts=[0, 1], products=[2], rate=5.0, educt_distance=1.0),
           dict(kind=FIS, educts=[2], products=[0, 1], rate=3.0, product_distance=0.9),
           dict(kind=ENZ, educts=[1, 2], products=[3, 2], rate=4.0, educt_distance=1.2),
           dict(kind=CONV, educts=[3], products=[0], rate=2.0),
           dict(kind=DEC, educts=[3], rate=0.5)]


def _oracle(n_types, box, rxs, diffusion):
    o = po.System(n_types, box, [1, 1, 1], diffusion=diffusion)
    for r in rxs:
        o.add_reaction(r["kind"], r.get("educts", []), r.get("products", []), r["rate"], r.get("educt_distance", 0.0),
                       r.get("product_distance", 0.0))
    return o


def _check_scene(rec, n_before=3):
    """test_observables_io.py:381-392: one conversion (label A->B) at the origin, one fusion (B+C->A)
    at (1.05, 1, 1); here also the ids and types around them."""
    assert len(rec) == 2
    conv = rec[rec["type"] == CONV]
    fus = rec[rec["type"] == FUS]
    assert len(conv) == 1 and len(fus) == 1
    c, f = conv[0], fus[0]
    assert c["reaction"] == 1 and np.array_equal(c["position"], [0.0, 0.0, 0.0])
    assert list(c["educts"]) == [0, 0] and list(c["types_from"]) == [0, 0]
    assert f["reaction"] == 2
    np.testing.assert_allclose(f["position"], [1.05, 1.0, 1.0], rtol=0, atol=1e-15)
    assert sorted(f["educts"]) == [1, 2] and sorted(f["types_from"]) == [1, 2]
    assert c["products"][1] == 0 and f["products"][1] == 0
    fresh = {int(c["products"][0]), int(f["products"][0])}
    assert len(fresh) == 2 and min(fresh) >= n_before


def test_oracle_records_reference_scene():
    o = _oracle(3, [10.0] * 3, SCENE_RXS, [0.0] * 3)
    o.set_record_reactions(True)
    o.add_particles(SCENE_POS, SCENE_TYPES.astype(np.int32), ids=np.arange(3))
    o.nl_setup(o.max_cutoff())
    o.reactions(po.UA_KEYED, 0.01, step=0)
    rec = o.reaction_records()
    _check_scene(rec)
    _, types, ids, _ = o.particles()
    after = dict(zip(ids.tolist(), types.tolist()))
    assert after == {int(rec[rec["type"] == CONV][0]["products"][0]): 1,
                     int(rec[rec["type"] == FUS][0]["products"][0]): 0}


@pytest.mark.parametrize("mode", [po.UA_KEYED, po.GILLESPIE_KEYED])
def test_oracle_records_match_the_states_around_them(mode):
    """Each record names educts that existed with those types and positions before the call, a
    position that is fixPosition of their plain mean, and products that exist afterwards with the
    reaction's product types; consumed educts are gone, the enzymatic catalyst stays."""
    rng = np.random.default_rng(5)
    n, box = 1500, 14.0
    o = _oracle(4, [box] * 3, ALL_RXS, [1.0, 0.8, 0.5, 1.2])
    o.set_record_reactions(True)
    o.set_seed(11)
    o.add_particles((rng.random((n, 3)) - 0.5) * box, rng.integers(0, 4, n).astype(np.int32), ids=np.arange(n))
    o.nl_setup(o.max_cutoff())
    seen = set()
    for t in range(4):
        o.euler_bd(0.01, step=t)
        o.nl_fill()
        p0, t0, i0, _ = o.particles()
        before = {int(i): (p0[k], int(t0[k])) for k, i in enumerate(i0)}
        o.reactions(mode, 0.02, step=t)
        o.nl_fill()
        rec = o.reaction_records()
        assert len(rec) == int(o.reaction_counts().sum()) > 10
        p1, t1, i1, _ = o.particles()
        after = {int(i): int(t1[k]) for k, i in enumerate(i1)}
        for r in rec:
            rx = ALL_RXS[r["reaction"]]
            assert r["type"] == rx["kind"]
            e1, e2 = int(r["educts"][0]), int(r["educts"][1])
            (x1, ty1), (x2, ty2) = before[e1], before[e2]
            assert [ty1, ty2] == list(r["types_from"])
            if rx["kind"] in (FUS, ENZ):
                assert e1 != e2 and sorted([ty1, ty2]) == sorted(rx["educts"])
            else:
                assert e1 == e2 and ty1 == rx["educts"][0]
            assert np.array_equal(r["position"], po.fix_position((x1 + x2) / 2., [box] * 3, [1, 1, 1]))
            prods = [int(q) for q in r["products"]]
            if rx["kind"] == DEC:
                assert prods == [0, 0] and e1 not in after
            elif rx["kind"] in (CONV, FUS):
                assert prods[1] == 0 and after[prods[0]] == rx["products"][0]
                assert e1 not in after and e2 not in after
            elif rx["kind"] == FIS:
                assert after[prods[0]] == rx["products"][0] and after[prods[1]] == rx["products"][1]
                assert e1 not in after
            else:  # enzymatic: educts {from, catalyst}; the catalyst keeps its id
                cat, conv = (e1, e2) if ty1 == rx["educts"][1] else (e2, e1)
                assert cat in after and conv not in after
                assert set(prods) == {cat, next(q for q in prods if q != cat)}
                assert after[next(q for q in prods if q != cat)] == rx["products"][0]
            seen.add(int(r["type"]))
    assert seen == {CONV, FUS, FIS, ENZ, DEC}


# ------------------------------------------------------------------------------------------ GPU
L = None


def _gpu():
    global L
    L = pytest.importorskip("readdy_amd._lib")
    from readdy_amd import Engine
    return Engine


def _rxs_gpu(rxs):
    return [dict(r, kind={CONV: L.RX_CONVERSION, FUS: L.RX_FUSION, FIS: L.RX_FISSION, ENZ: L.RX_ENZYMATIC,
                          DEC: L.RX_DECAY}[r["kind"]]) for r in rxs]


@pytest.mark.gpu
@pytest.mark.parametrize("handler", [0, 1])
def test_gpu_records_reference_scene(handler):
    Engine = _gpu()
    e = Engine(3, [0.0] * 3, reactions=_rxs_gpu(SCENE_RXS))
    r = e.add_replica([10.0] * 3, seed=4, max_particles=64)
    e.set_particles(r, SCENE_POS, SCENE_TYPES)
    e.record_reactions(True)
    e.run(handler, 0.01, 1)
    rec = e.reaction_records(r)
    print("scene records", rec, flush=True)
    _check_scene(rec)
    _, t, h = e.get_particles(r)
    assert dict(zip(h.tolist(), t.tolist())) == {int(rec[rec["type"] == CONV][0]["products"][0]): 1,
                                                 int(rec[rec["type"] == FUS][0]["products"][0]): 0}


def _by_educts(rec):
    return rec[np.lexsort((rec["educts"][:, 1], rec["educts"][:, 0]))]


@pytest.mark.gpu
@pytest.mark.parametrize("handler,mode", [(0, po.UA_KEYED), (1, po.GILLESPIE_KEYED)])
def test_gpu_records_equal_oracle_event_for_event(handler, mode):
    """No potentials and injected noise, so both sides hold bit-identical states: the same events
    fire and their records agree in every field."""
    Engine = _gpu()
    rng = np.random.default_rng(8)
    n, box, dt = 1500, 14.0, 0.02
    diff = [1.0, 0.8, 0.5, 1.2]
    pos = (rng.random((n, 3)) - 0.5) * box
    types = rng.integers(0, 4, n).astype(np.uint32)
    e = Engine(4, diff, reactions=_rxs_gpu(ALL_RXS))
    r = e.add_replica([box] * 3, seed=21, max_particles=4 * n)
    e.set_particles(r, pos, types)
    e.record_reactions(True)
    o = _oracle(4, [box] * 3, ALL_RXS, diff)
    o.set_seed(21)
    o.set_record_reactions(True)
    o.add_particles(pos, types.astype(np.int32), ids=np.arange(n))
    o.nl_setup(o.max_cutoff())
    total = 0
    for t in range(5):
        print("records step", t, "handler", handler, flush=True)
        noise = np.random.default_rng(300 + t).normal(size=(4 * n, 3))
        e.integrate(dt, noise)
        e.update_neighbor_list()
        o.euler_bd(dt, noise.ravel(), step=t)
        o.nl_fill()
        e.react(handler, dt)
        o.reactions(mode, dt, step=t)
        o.nl_fill()
        g, w = _by_educts(e.reaction_records(r)), _by_educts(o.reaction_records())
        assert len(g) == len(w) == int(e.reaction_counts(r).sum())
        for f in ("reaction", "type", "educts", "products", "types_from"):
            assert np.array_equal(g[f].astype(np.int64), w[f].astype(np.int64)), f
        assert np.abs(g["position"] - w["position"]).max(initial=0.0) < 1e-12
        total += len(g)
    assert total > 100


@pytest.mark.gpu
@pytest.mark.parametrize("handler,graph", [(0, False), (1, False), (0, True)])
def test_gpu_fused_step_records_equal_separate_actions(handler, graph):
    """The fused step (eager or graph-replayed) keeps the records of its last step, equal to those of
    integrate / NL / react / NL / forces; one replica among two."""
    Engine = _gpu()
    pots = [(L.POT_HARMONIC_REPULSION, a, b, [10.0, 0.8 + 0.1 * (a + b)]) for a in range(4) for b in range(a, 4)]
    rng = np.random.default_rng(17)
    n, box, dt, steps = 2000, 15.0, 0.02, 12
    scenes = [((rng.random((n, 3)) - 0.5) * box, rng.integers(0, 4, n).astype(np.uint32)) for _ in range(2)]

    def make():
        e = Engine(4, [1.0, 0.8, 0.5, 1.2], potentials=pots, reactions=_rxs_gpu(ALL_RXS), max_replicas=2)
        rs = []
        for k, (pos, types) in enumerate(scenes):
            rs.append(e.add_replica([box] * 3, seed=40 + k, max_particles=3 * n))
            e.set_particles(rs[-1], pos, types)
        e.record_reactions(True)
        return e, rs

    e1, r1 = make()
    e1.run(handler, dt, steps, graph=graph)
    e2, r2 = make()
    e2.calculate_forces()
    for _ in range(steps):
        e2.integrate(dt)
        e2.update_neighbor_list()
        e2.react(handler, dt)
        e2.update_neighbor_list()
        e2.calculate_forces()
    for a, b in zip(r1, r2):
        g, w = e1.reaction_records(a), e2.reaction_records(b)
        assert len(g) > 10 and g.tobytes() == w.tobytes()  # same fields, same key order
        assert np.array_equal(np.bincount(g["reaction"], minlength=5).astype(np.uint64), e2.reaction_counts(b))


@pytest.mark.gpu
def test_gpu_reused_slot_starts_without_records():
    """A replica added into the slot of a removed one reports no records before its first reaction step."""
    Engine = _gpu()
    e = Engine(3, [0.0] * 3, reactions=_rxs_gpu(SCENE_RXS), max_replicas=1)
    r = e.add_replica([10.0] * 3, seed=4, max_particles=64)
    e.set_particles(r, SCENE_POS, SCENE_TYPES)
    e.record_reactions(True)
    e.run(0, 0.01, 1)
    assert len(e.reaction_records(r)) == 2
    e.remove_replica(r)
    r2 = e.add_replica([10.0] * 3, seed=5, max_particles=64)
    assert r2 == r and len(e.reaction_records(r2)) == 0
