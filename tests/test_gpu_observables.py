"""Device radial distribution against the oracle's brute-force restatement of
RadialDistribution::evaluate (RadialDistribution.cpp:75-124): the integer counts per bin are bit-exact
(the same FP64 minimum-image distance, sqrt and upper_bound on both sides), for stencils wider than
one cell, boxes smaller than twice the largest border (the de-duplicated wrapped stencil), first
borders above zero, particles outside a non-periodic box, and one replica among several.  At one
million particles the pair count below the neighbour-list cutoff equals the cell list's pair count.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

L = pytest.importorskip("readdy_amd._lib")
from readdy_amd import Engine  # noqa: E402

HR = [10.0, 1.0]


def _pots(nt):
    return [(L.POT_HARMONIC_REPULSION, a, b, HR) for a in range(nt) for b in range(a, nt)]


def _scene(n, box, seed, nt=3, outside_axis=None):
    rng = np.random.default_rng(seed)
    box = np.asarray(box, dtype=np.float64)
    pos = (rng.random((n, 3)) - 0.5) * box
    if outside_axis is not None:  # a few particles beyond a non-periodic wall
        k = n // 50
        pos[:k, outside_axis] = np.where(rng.random(k) < 0.5, -1, 1) * (box[outside_axis] / 2 + rng.random(k) * 2.0)
    types = rng.integers(0, nt, n).astype(np.uint32)
    return pos, types


def _oracle_counts(pos, types, box, pbc, borders, frm, to, nt=3):
    o = po.System(nt, list(box), list(pbc))
    o.add_particles(pos, types.astype(np.int32), ids=np.arange(len(types)))
    return o.rdf_counts(borders, frm, to)


CASES = [
    # box, pbc, n, borders
    ([20.0, 20.0, 20.0], (1, 1, 1), 2000, np.arange(0, 5, 0.01)),
    ([20.0, 18.0, 22.0], (1, 1, 1), 2500, np.concatenate([[0.3], np.linspace(0.5, 3.5, 61)])),
    ([6.0, 6.0, 7.0], (1, 1, 1), 400, np.linspace(0.0, 4.5, 46)),        # borders beyond L/2: wrapped stencil
    ([9.0, 7.0, 11.0], (1, 0, 1), 900, np.linspace(0.0, 3.0, 31)),       # a wall, particles beyond it
    ([8.0, 8.0, 8.0], (0, 0, 0), 600, np.linspace(0.1, 2.5, 25)),
]


@pytest.mark.parametrize("box,pbc,n,borders", CASES)
def test_rdf_counts_bit_exact(box, pbc, n, borders):
    outside = 1 if pbc == (1, 0, 1) else (0 if pbc == (0, 0, 0) else None)
    pos, types = _scene(n, box, seed=n, outside_axis=outside)
    e = Engine(3, [1.0] * 3, potentials=_pots(3))
    r = e.add_replica(box, pbc, seed=1, max_particles=4 * n)
    e.set_particles(r, pos, types)
    for frm, to in [([0, 1, 2], [0, 1, 2]), ([0], [1, 2]), ([2], [2]), ([1], [0])]:
        print("rdf case", box, pbc, n, frm, to, flush=True)
        got, nf = e.rdf_counts(r, borders, frm, to)
        want, wnf = _oracle_counts(pos, types, box, pbc, borders, frm, to)
        assert nf == wnf
        assert np.array_equal(got, want)
        assert want.sum() > 0


def test_rdf_of_one_replica_among_several():
    box = [15.0, 15.0, 15.0]
    e = Engine(3, [1.0] * 3, potentials=_pots(3), max_replicas=3)
    scenes = []
    for k in range(3):
        pos, types = _scene(800 + 100 * k, box, seed=40 + k)
        r = e.add_replica(box, (1, 1, 1), seed=k, max_particles=4000)
        e.set_particles(r, pos, types)
        scenes.append((r, pos, types))
    borders = np.linspace(0.0, 4.0, 81)
    for r, pos, types in scenes:
        got, nf = e.rdf_counts(r, borders, [0, 1], [1, 2])
        want, wnf = _oracle_counts(pos, types, box, (1, 1, 1), borders, [0, 1], [1, 2])
        assert nf == wnf and np.array_equal(got, want)


def test_rdf_after_steps_and_normalisation():
    box = [12.0, 12.0, 12.0]
    pos, types = _scene(1200, box, seed=5)
    e = Engine(3, [1.0, 1.0, 0.5], potentials=_pots(3))
    r = e.add_replica(box, (1, 1, 1), seed=3, max_particles=5000)
    e.set_particles(r, pos, types)
    e.run(2, 0.01, 7)  # no reactions: the state after seven fused steps
    p, t, _ = e.get_particles(r)
    borders = np.arange(0, 5, 0.05)
    got, nf = e.rdf_counts(r, borders, [0], [0, 1, 2])
    want, wnf = _oracle_counts(p, t, box, (1, 1, 1), borders, [0], [0, 1, 2])
    assert nf == wnf and np.array_equal(got, want)
    # RadialDistribution.cpp:105-123: counts / (4/3 pi (r_hi^3 - r_lo^3) n_from density)
    density = len(t) / np.prod(box)
    centers, dist = e.radial_distribution(r, borders, [0], [0, 1, 2], density)
    shell = 4. / 3. * np.pi * (np.power(borders[1:], 3.) - np.power(borders[:-1], 3.))
    assert np.array_equal(centers, (borders[:-1] + borders[1:]) / 2)
    assert np.array_equal(dist, want.astype(np.float64) / (shell * wnf * density))
    # an ideal-gas-like system: g(r) near 1 beyond the repulsion range
    assert abs(dist[(centers > 2.0) & (centers < 4.5)].mean() - 1.0) < 0.1


def test_rdf_at_one_million_particles_counts_the_cell_list_pairs():
    import bench
    pos, types = bench.c1_system()
    pots, rxs = bench.c1_chemistry(L)
    e = Engine(3, [1.0, 1.0, 0.5], potentials=pots, reactions=rxs)
    r = e.add_replica([200.0] * 3, seed=1, max_particles=len(types))
    e.set_particles(r, pos, types)
    counts, nf = e.rdf_counts(r, [0.0, 2.0], [0, 1, 2], [0, 1, 2])
    assert nf == len(types)
    n_pairs = len(e.neighbor_pairs(r, 2.0))
    # sqrt(d2) < 2 <=> d2 < 4 but for d2 within an ulp below 4 (where sqrt rounds to 2): none at these sizes
    assert int(counts[0]) == n_pairs


@pytest.mark.parametrize("pbc", [(1, 1, 1), (1, 0, 1)])
def test_histogram_along_axis_bit_exact(pbc):
    box = [14.0, 11.0, 9.0]
    outside = 1 if pbc == (1, 0, 1) else None
    pos, types = _scene(3000, box, seed=21, outside_axis=outside)
    e = Engine(3, [1.0] * 3, potentials=_pots(3))
    r = e.add_replica(box, pbc, seed=1, max_particles=12000)
    e.set_particles(r, pos, types)
    o = po.System(3, box, list(pbc))
    o.add_particles(pos, types.astype(np.int32), ids=np.arange(len(types)))
    borders = np.linspace(-7.5, 7.5, 151)
    for axis in range(3):
        for sel in ([0, 1, 2], [2], [0, 1]):
            got = e.histogram_along_axis(r, borders, sel, axis)
            want = o.histogram_along_axis(borders, sel, axis)
            assert np.array_equal(got, want) and want.sum() > 0
