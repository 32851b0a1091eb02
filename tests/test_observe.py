"""Recorder: the reference's observable timing (evaluated before the first step and after every step
whose number is a multiple of the stride, stride 0 = every step; SimulationLoop.h:268-300,
Observable.h:144-147) around the engine's fused steps, and the HDF5 file it writes.

CPU: the schedule against a stand-in engine that counts steps.  GPU: every observable on a reacting
system; the series agree with each other (particle counts change by the reactions' stoichiometry,
one record per counted event, the last trajectory frame is the final state) and with direct calls.
"""
import numpy as np
import pytest

from readdy_amd import h5io
from readdy_amd.observe import Recorder


class _Steps:
    """Stands in for an engine: the 'state' is the number of steps run."""

    def __init__(self):
        self.steps, self.runs = 0, []

    def run(self, handler, dt, k, graph=True):
        assert k > 0
        self.steps += k
        self.runs.append(k)

    def count_types(self, n):
        return np.array([[self.steps, 0]], dtype=np.uint64)

    def reaction_counts(self, r):
        return np.array([self.steps % 3], dtype=np.uint64)


def test_schedule_matches_the_reference_loop():
    e = _Steps()
    rec = Recorder(e)
    rec.n_particles(3, name="every_third")
    rec.reaction_counts(0, name="every_step")
    rec.n_particles(4, name="every_fourth")
    rec.run(0, 0.1, 10)
    rec.run(0, 0.1, 5)  # continues at step 10 without evaluating it twice
    t3, v3 = rec.series("every_third")
    assert list(t3) == [0, 3, 6, 9, 12, 15] and [int(x[0]) for x in v3] == list(t3)
    t1, _ = rec.series("every_step")
    assert list(t1) == list(range(16))
    t4, v4 = rec.series("every_fourth")
    assert list(t4) == [0, 4, 8, 12] and [int(x[0]) for x in v4] == [0, 4, 8, 12]
    assert sum(e.runs) == 15


def test_only_strided_observables_chunk_the_run():
    e = _Steps()
    rec = Recorder(e)
    rec.n_particles(25)
    rec.run(0, 0.1, 100)
    assert e.runs == [25, 25, 25, 25]
    with pytest.raises(ValueError):
        rec.n_particles(5)  # the same name twice


def test_write_without_gpu(tmp_path):
    e = _Steps()
    rec = Recorder(e)
    rec.n_particles(2)
    rec.reaction_counts(5)
    rec.run(0, 0.1, 10)
    path = str(tmp_path / "obs.h5")
    rec.write(path, ["A", "B"], [10.0] * 3, diffusion=[1.0, 0.5],
              reactions=[dict(kind=0, educts=[0], products=[1], rate=1.0, name="conv")])
    assert list(h5io.read_u64(path, "/readdy/observables/n_particles/time")) == [0, 2, 4, 6, 8, 10]
    d = h5io.read_u64(path, "/readdy/observables/n_particles/data").reshape(6, 2)
    assert list(d[:, 0]) == [0, 2, 4, 6, 8, 10]
    assert list(h5io.read_u64(path, "/readdy/observables/reaction_counts/counts/0")) == [0, 2, 1]


@pytest.mark.gpu
def test_recorder_on_the_device(tmp_path):
    L = pytest.importorskip("readdy_amd._lib")
    from readdy_amd import Engine
    pots = [(L.POT_HARMONIC_REPULSION, a, b, [10.0, 1.0]) for a in range(3) for b in range(a, 3)]
    rxs = [dict(kind=L.RX_FUSION, educts=[0, 1], products=[2], rate=2.0, educt_distance=1.0, name="fus"),
           dict(kind=L.RX_FISSION, educts=[2], products=[0, 1], rate=1.0, product_distance=1.0, name="fis")]
    rng = np.random.default_rng(9)
    n, box = 4000, 25.0
    pos = (rng.random((n, 3)) - 0.5) * box
    types = rng.integers(0, 3, n).astype(np.uint32)
    e = Engine(3, [1.0, 1.0, 0.5], potentials=pots, reactions=[{k: v for k, v in r.items() if k != "name"}
                                                                for r in rxs])
    r = e.add_replica([box] * 3, seed=5, max_particles=3 * n)
    e.set_particles(r, pos, types)
    rec = Recorder(e, r)
    rec.n_particles(1)
    rec.reaction_counts(1)
    rec.reactions(1)
    rec.trajectory(4)
    rec.positions(4)
    rec.forces(10)
    rec.particles(20)
    borders = np.linspace(0.0, 4.0, 41)
    rec.radial_distribution(5, borders, [0], [0, 1, 2], n / box ** 3)
    hb = np.linspace(-box / 2, box / 2, 26)
    rec.histogram_along_axis(5, hb, [2], 0)
    print("recorder run", flush=True)
    rec.run(0, 0.01, 12)
    rec.run(0, 0.01, 8)
    t, counts = rec.series("n_particles")
    _, rcounts = rec.series("reaction_counts")
    _, recs = rec.series("reactions")
    assert list(t) == list(range(21))
    counts = np.array(counts, dtype=np.int64)
    rcounts = np.array(rcounts, dtype=np.int64)
    # fusion A+B->C: (-1,-1,+1); fission C->A+B: (+1,+1,-1)
    stoich = np.array([[-1, -1, 1], [1, 1, -1]])
    assert np.array_equal(np.diff(counts, axis=0), rcounts[1:] @ stoich)
    assert rcounts[1:].sum() > 20 and not rcounts[0].any()
    assert [len(x) for x in recs] == list(rcounts.sum(axis=1))
    tt, frames = rec.series("", "trajectory")
    assert list(tt) == [0, 4, 8, 12, 16, 20]
    xf, tf, hf = e.get_particles(r)
    assert np.array_equal(frames[-1][0], hf) and np.array_equal(frames[-1][2], xf)
    assert np.array_equal(np.bincount(frames[2][1], minlength=3), counts[8])
    _, rdf = rec.series("radial_distribution")
    assert np.array_equal(rdf[-1], e.radial_distribution(r, borders, [0], [0, 1, 2], n / box ** 3)[1])
    _, hist = rec.series("histogram_along_axis")
    assert np.array_equal(hist[-1], e.histogram_along_axis(r, hb, [2], 0).astype(np.float64))
    _, ps = rec.series("positions")
    assert all(np.array_equal(p, fr[2]) for p, fr in zip(ps, frames))  # positions equal the trajectory's
    tf_, fs = rec.series("forces")
    f_now, h_now = e.get_forces(r)
    assert list(tf_) == [0, 10, 20] and np.array_equal(fs[-1], f_now[np.argsort(h_now)])
    _, parts = rec.series("particles")
    assert np.array_equal(parts[-1][1], hf) and np.array_equal(parts[-1][2], xf)
    path = str(tmp_path / "run.h5")
    rec.write(path, ["A", "B", "C"], [box] * 3, diffusion=[1.0, 1.0, 0.5], reactions=rxs)
    ft, fframes = h5io.read_trajectory(path, "")
    assert list(ft) == list(tt) and np.array_equal(fframes[-1][3], xf)
    rt, rrec = h5io.read_reactions(path, "reactions")
    assert [len(x) for x in rrec] == [len(x) for x in recs]
    assert np.array_equal(h5io.read_f64(path, "/readdy/observables/radial_distribution/distribution").reshape(5, -1),
                          np.array(rdf))
    assert np.array_equal(h5io.read_vlen(path, "/readdy/observables/forces/data", 0)[-1], fs[-1])
    assert np.array_equal(h5io.read_vlen(path, "/readdy/observables/particles/ids", 2)[-1], hf)


def test_streaming_appends_every_flush_stride(tmp_path):
    """open() streams: no observable holds flush_stride frames in memory, and the file after close() equals
    the one write() makes from a recorder that kept everything"""
    def recorder():
        r = Recorder(_Steps())
        r.n_particles(1)
        r.reaction_counts(2)
        return r

    path_s, path_w = str(tmp_path / "stream.h5"), str(tmp_path / "write.h5")
    rx = [dict(kind=0, educts=[0], products=[1], rate=1.0, name="conv")]
    rs = recorder()
    rs.open(path_s, ["A", "B"], [10.0] * 3, reactions=rx, flush_stride=3)
    for _ in range(5):
        rs.run(0, 0.1, 4)
        assert all(len(o["time"]) < 3 for o in rs._obs)
    rs.close()
    rw = recorder()
    for _ in range(5):
        rw.run(0, 0.1, 4)
    rw.write(path_w, ["A", "B"], [10.0] * 3, reactions=rx)
    for ds in ("n_particles/time", "n_particles/data", "reaction_counts/time", "reaction_counts/counts/0"):
        a = h5io.read_u64(path_s, "/readdy/observables/" + ds)
        b = h5io.read_u64(path_w, "/readdy/observables/" + ds)
        assert np.array_equal(a, b), ds
    assert list(h5io.read_u64(path_s, "/readdy/observables/n_particles/time")) == list(range(21))
    with pytest.raises(ValueError):
        recorder().open(str(tmp_path / "x.h5"), ["A", "B"], [10.0] * 3, flush_stride=0)
