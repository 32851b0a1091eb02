"""The allocation-free histogram entry point against the allocating one it mirrors.

rdy_histogram_along_axis_noalloc computes HistogramAlongAxis (CPUObservables.cpp:85-136) in a static
per-GPU workspace.  Its counts must equal rdy_histogram_along_axis' bit for bit: on two calls in a row,
the second with fewer borders, fewer species and another replica and axis (the workspace keeps the
first call's contents, which the second must not see), and after fused steps changed the state.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = pytest.importorskip("readdy_amd._lib")
from readdy_amd import Engine  # noqa: E402


def _engine():
    pots = [(L.POT_HARMONIC_REPULSION, a, b, [10.0, 1.0]) for a in range(3) for b in range(a, 3)]
    rxs = [dict(kind=L.RX_FUSION, educts=[0, 1], products=[2], rate=0.1, educt_distance=1.0),
           dict(kind=L.RX_FISSION, educts=[2], products=[0, 1], rate=0.05, product_distance=1.0)]
    e = Engine(3, [1.0, 1.0, 0.5], potentials=pots, reactions=rxs, max_replicas=2)
    rng = np.random.default_rng(31)
    rs = []
    for q, (n, Lb) in enumerate([(20000, 40.0), (5000, 25.0)]):
        r = e.add_replica([Lb] * 3, (1, 1, 0), seed=q + 1, max_particles=2 * n)
        e.set_particles(r, (rng.random((n, 3)) - 0.5) * Lb, rng.integers(0, 3, n).astype(np.uint32))
        rs.append(r)
    return e, rs


def test_noalloc_histogram_equals_the_allocating_entry_point():
    e, (r0, r1) = _engine()
    print("no-alloc histogram: call 1 (replica 0, 201 borders, types 0 and 2, axis x)", flush=True)
    b1 = np.linspace(-20.0, 20.0, 201)
    want1 = e.histogram_along_axis(r0, b1, [0, 2], 0)
    got1 = e.histogram_along_axis_noalloc(r0, b1, [0, 2], 0)
    assert want1.sum() > 10000 and np.array_equal(got1, want1)
    print("no-alloc histogram: call 2 (replica 1, 33 borders, type 1, axis z)", flush=True)
    b2 = np.linspace(-12.5, 12.5, 33)
    want2 = e.histogram_along_axis(r1, b2, [1], 2)
    got2 = e.histogram_along_axis_noalloc(r1, b2, [1], 2)
    assert want2.sum() > 1000 and np.array_equal(got2, want2)
    print("no-alloc histogram: after fused steps, the maximum border count", flush=True)
    e.run(L.HANDLER_UA, 0.01, 5, graph=True)
    b3 = np.linspace(-20.0, 20.0, 4097)
    assert np.array_equal(e.histogram_along_axis_noalloc(r0, b3, [0, 1, 2], 1), e.histogram_along_axis(r0, b3, [0, 1, 2], 1))


def test_noalloc_histogram_rejects_sizes_beyond_its_workspace():
    e, (r0, _) = _engine()
    with pytest.raises(L.EngineError, match="4097"):
        e.histogram_along_axis_noalloc(r0, np.linspace(-1.0, 1.0, 4098), [0], 0)
    with pytest.raises(L.EngineError):
        e.histogram_along_axis_noalloc(r0, np.array([0.0]), [0], 0)
    with pytest.raises(L.EngineError):
        e.histogram_along_axis_noalloc(r0, np.linspace(-1.0, 1.0, 5), [7], 0)
    # a rejected call leaves the entry point working
    b = np.linspace(-20.0, 20.0, 11)
    assert np.array_equal(e.histogram_along_axis_noalloc(r0, b, [0], 0), e.histogram_along_axis(r0, b, [0], 0))
