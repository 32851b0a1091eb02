"""The single-system autotuner (readdy_amd/tune.py): its candidate list, and on the GPU that every candidate
geometry (cell size, brick shape) gives the same trajectory bit for bit, so that the tuner's choice changes
the speed only.  The system is the 10^6-particle workload's chemistry and density in a 62^3 box (about 30k
particles), a few fused UA steps with reactions firing."""
import numpy as np
import pytest

from readdy_amd.tune import autotune_system, brick_candidates


def test_candidates_cover_the_engine_choice_and_cubes_around_one_round():
    c = brick_candidates(0.125, 2.0)
    assert {"cell_radius": 1, "brick": [0, 0, 0]} in c
    assert {"cell_radius": 2, "brick": [0, 0, 0]} in c
    cubes = [x["brick"][0] for x in c if x["cell_radius"] == 1 and x["brick"][0]]
    assert 6 in cubes and all(x["brick"][0] == x["brick"][1] == x["brick"][2] for x in c)
    # denser systems get smaller bricks
    assert max(x["brick"][0] for x in brick_candidates(1.0, 2.0)) < max(cubes)


@pytest.mark.gpu
def test_autotune_candidates_give_one_trajectory():
    from readdy_amd import Engine
    from readdy_amd import _lib as L
    from readdy_amd.workloads import c1_chemistry

    box = 62.0
    n = int(0.125 * box ** 3)
    rng = np.random.default_rng(5)
    pos = (rng.random((n, 3)) - 0.5) * box
    types = rng.choice(np.array([0, 1, 2], dtype=np.uint32), n, p=[0.4, 0.4, 0.2])
    pots, rxs = c1_chemistry(L)
    rxs = [dict(x, rate=20.0 * x["rate"]) for x in rxs]  # reactions fire in every step

    def make_engine(cand):
        e = Engine(3, [1.0, 1.0, 0.5], potentials=pots, reactions=rxs, cell_radius=cand["cell_radius"],
                   brick=tuple(cand["brick"]))
        r = e.add_replica([box] * 3, (1, 1, 1), seed=9, max_particles=2 * n)
        e.set_particles(r, pos, types)
        return e, r

    cands = brick_candidates(0.125, 2.0) + [{"cell_radius": 1, "brick": [4, 6, 9]}]
    print(f"{n} particles, {len(cands)} candidates", flush=True)
    best, res = autotune_system(make_engine, L.HANDLER_UA, 0.005, cands, steps=6, warmup=2, keep_states=True)
    assert best in res and best["particle_steps_per_s"] == max(x["particle_steps_per_s"] for x in res)
    p0, t0, h0 = res[0]["state"]
    assert not np.array_equal(np.sort(t0), np.sort(types))  # reactions happened
    for x in res[1:]:
        p, t, h = x["state"]
        assert np.array_equal(h, h0) and np.array_equal(t, t0) and np.array_equal(p, p0), x


def test_ensemble_brick_candidates_around_the_half_round_cube():
    from readdy_amd.ensemble import brick_candidates, half_round_brick
    for cell in (1.0, 1.5, 2.0):
        e = half_round_brick(cell)[0]
        c = brick_candidates(cell)
        assert c[0] == (0, 0, 0) and (e, e, e) in c and (e, e, e + 1) in c and (e + 1,) * 3 in c
