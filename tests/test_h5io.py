"""HDF5 output in ReaDDy's layout (CPU): FlatTrajectory records/limits/time (Trajectory.cpp:112-158,
record fields as in Types.cpp:79-88) and RadialDistribution bin_centers/distribution/time
(RadialDistribution.cpp:158-170) written through libreaddy_h5.so and read back; the dataset paths,
the compound's field names and types checked with h5dump where the image has it."""
import os
import shutil
import subprocess

import numpy as np
import pytest

h5io = pytest.importorskip("readdy_amd.h5io")
try:
    h5io.lib()
except OSError:
    pytest.skip("libreaddy_h5.so not built", allow_module_level=True)


def _frames(rng, sizes):
    out = []
    for n in sizes:
        ids = rng.permutation(10 * n + 1)[:n].astype(np.uint64)
        out.append((ids, rng.integers(0, 7, n).astype(np.uint16), rng.integers(0, 2, n).astype(np.uint8),
                    rng.normal(size=(n, 3))))
    return out


def test_flat_trajectory_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    frames = _frames(rng, [5, 0, 17, 1])  # an empty frame keeps its [k, k) limits
    time = [0, 10, 20, 35]
    path = str(tmp_path / "traj.h5")
    with h5io.File(path) as f:
        f.flat_trajectory("traj", time, frames)
        f.attribute("note", "round trip")
    t, got = h5io.read_trajectory(path, "traj")
    assert np.array_equal(t, time)
    lim = h5io.read_u64(path, "/readdy/trajectory/traj/limits").reshape(-1, 2)
    ends = np.cumsum([len(fr[0]) for fr in frames])
    assert np.array_equal(lim[:, 1], ends) and np.array_equal(lim[1:, 0], ends[:-1]) and lim[0, 0] == 0
    for (i1, t1, f1, p1), (i2, t2, f2, p2) in zip(frames, got):
        assert np.array_equal(i1, i2) and np.array_equal(t1, t2) and np.array_equal(f1, f2)
        assert np.array_equal(p1, p2)  # FP64 stored as FP64


def test_radial_distribution_round_trip(tmp_path):
    path = str(tmp_path / "rdf.h5")
    borders = np.linspace(0.0, 3.0, 31)
    centers = 0.5 * (borders[1:] + borders[:-1])
    dist = np.random.default_rng(2).random((3, 30))
    with h5io.File(path) as f:
        f.radial_distribution("rdf", [0, 100, 200], centers, dist)
        f.n_particles("n_particles", [0, 100, 200], [[1, 2], [3, 4], [5, 6]])
    g = "/readdy/observables/rdf"
    assert np.array_equal(h5io.read_f64(path, g + "/bin_centers"), centers)
    assert np.array_equal(h5io.read_f64(path, g + "/distribution").reshape(3, 30), dist)
    assert np.array_equal(h5io.read_u64(path, g + "/time"), [0, 100, 200])
    assert np.array_equal(h5io.read_u64(path, "/readdy/observables/n_particles/data"), [1, 2, 3, 4, 5, 6])


def test_missing_dataset_and_bad_frames_raise(tmp_path):
    path = str(tmp_path / "x.h5")
    with h5io.File(path) as f:
        with pytest.raises(ValueError):
            f.flat_trajectory("t", [0, 1], _frames(np.random.default_rng(3), [2]))
        f.n_particles("n", [0], [[1]])
    with pytest.raises(OSError):
        h5io.read_u64(path, "/readdy/observables/none/time")


@pytest.mark.skipif(shutil.which("h5dump") is None and not os.path.exists("/opt/conda/bin/h5dump"),
                    reason="h5dump not in the image")
def test_layout_seen_by_h5dump(tmp_path):
    path = str(tmp_path / "traj.h5")
    with h5io.File(path) as f:
        f.flat_trajectory("traj", [0], _frames(np.random.default_rng(4), [3]))
    exe = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    out = subprocess.run([exe, "-H", path], capture_output=True, text=True, check=True).stdout
    for needle in ('GROUP "readdy"', 'GROUP "trajectory"', 'GROUP "traj"', 'DATASET "records"',
                   'H5T_STD_U64LE "id"', 'H5T_STD_U16LE "typeId"', 'H5T_STD_U8LE "flavor"',
                   'H5T_ARRAY { [3] H5T_IEEE_F64LE } "pos"', 'DATASET "limits"', 'DATASET "time"'):
        assert needle in out, needle


def _h5dump():
    return shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def test_config_group(tmp_path):
    """/readdy/config as IOUtils.cpp:58-176 writes it: the general JSON (kbt, box_volume, box_size, pbc),
    particle types with ids in order, registered reactions with their educt/product types."""
    import json
    path = str(tmp_path / "cfg.h5")
    rxs = [dict(name="fus", id=0, educts=[0, 1], products=[2], rate=0.1, educt_distance=1.0),
           dict(name="fis", id=1, educts=[2], products=[0, 1], rate=0.05, product_distance=1.5)]
    with h5io.File(path) as f:
        f.config([("A", 1.0), ("B", 1.0), ("C", 0.5, "TOPOLOGY")], [10.0, 12.0, 14.0], (True, False, True), 2.5, rxs)
    g = json.loads(h5io.read_string(path, "/readdy/config/general"))
    assert g == {"kbt": 2.5, "box_volume": 1680.0, "box_size": [10.0, 12.0, 14.0], "pbc": [True, False, True]}
    exe = _h5dump()
    if exe is None:
        return
    out = subprocess.run([exe, "-d", "/readdy/config/particle_types", path], capture_output=True, text=True,
                         check=True).stdout
    flat = " ".join(out.split())
    for rec in ('"A", "NORMAL", 0, 1', '"B", "NORMAL", 1, 1', '"C", "TOPOLOGY", 2, 0.5'):
        assert rec in flat, (rec, flat)
    out = subprocess.run([exe, "-d", "/readdy/config/registered_reactions", path], capture_output=True, text=True,
                         check=True).stdout
    flat = " ".join(out.split())
    assert '"fus", 0, 2, 1, 0.1, 1, 0, [ 0, 1 ], [ 2, 0 ]' in flat, flat
    assert '"fis", 1, 1, 2, 0.05, 0, 1.5, [ 2, 0 ], [ 0, 1 ]' in flat, flat


def test_reactions_observable_round_trip(tmp_path):
    """Reactions.cpp:72-86: one variable-length list of records per time point (empty lists too)."""
    rng = np.random.default_rng(6)
    frames = []
    for n in (0, 4, 1, 0, 7):
        f = np.zeros(n, h5io.REACTION_DTYPE)
        f["reaction"] = rng.integers(0, 5, n)
        f["type"] = rng.integers(0, 5, n)
        f["educts"] = rng.integers(0, 1 << 40, (n, 2))
        f["products"] = rng.integers(0, 1 << 40, (n, 2))
        f["types_from"] = rng.integers(0, 9, (n, 2))
        f["position"] = rng.normal(size=(n, 3))
        frames.append(f)
    path = str(tmp_path / "rx.h5")
    with h5io.File(path) as fh:
        fh.reactions("reactions", [0, 1, 2, 3, 4], frames)
    t, got = h5io.read_reactions(path, "reactions")
    assert list(t) == [0, 1, 2, 3, 4] and [len(g) for g in got] == [len(f) for f in frames]
    for f, g in zip(frames, got):
        assert f.tobytes() == g.tobytes()
    exe = _h5dump()
    if exe:
        out = subprocess.run([exe, "-H", path], capture_output=True, text=True, check=True).stdout
        for needle in ('DATASET "records"', 'H5T_VLEN', 'H5T_STD_U64LE "reaction_id"', 'H5T_STD_I32LE "reaction_type"',
                       '"educts"', '"products"', '"position"', '"types_from"'):
            assert needle in out, needle


def test_positions_forces_particles_round_trip(tmp_path):
    """Positions / Forces (Positions.cpp:71-86, Forces.cpp:80-91): a variable-length list of {x, y, z}
    per time point; Particles (Particles.cpp:62-92): types, ids and positions lists."""
    rng = np.random.default_rng(8)
    sizes = (3, 0, 5)
    pos = [rng.normal(size=(n, 3)) for n in sizes]
    types = [rng.integers(0, 9, n).astype(np.uint16) for n in sizes]
    ids = [rng.integers(0, 1 << 50, n).astype(np.uint64) for n in sizes]
    path = str(tmp_path / "p.h5")
    with h5io.File(path) as f:
        f.positions("positions", [0, 5, 10], pos)
        f.forces("forces", [0, 5, 10], pos[::-1])
        f.particles("particles", [0, 5, 10], list(zip(types, ids, pos)))
    got = h5io.read_vlen(path, "/readdy/observables/positions/data", 0)
    assert [g.shape for g in got] == [(3, 3), (0, 3), (5, 3)] and all(np.array_equal(a, b) for a, b in zip(got, pos))
    assert all(np.array_equal(a, b) for a, b in zip(h5io.read_vlen(path, "/readdy/observables/forces/data", 0),
                                                   pos[::-1]))
    g = "/readdy/observables/particles/"
    assert all(np.array_equal(a, b) for a, b in zip(h5io.read_vlen(path, g + "types", 1), types))
    assert all(np.array_equal(a, b) for a, b in zip(h5io.read_vlen(path, g + "ids", 2), ids))
    assert all(np.array_equal(a, b) for a, b in zip(h5io.read_vlen(path, g + "positions", 0), pos))
    assert list(h5io.read_u64(path, g + "time")) == [0, 5, 10]
    exe = _h5dump()
    if exe:
        out = subprocess.run([exe, "-H", path], capture_output=True, text=True, check=True).stdout
        flat = "".join(out.split())
        assert 'H5T_VLEN{H5T_COMPOUND{H5T_IEEE_F64LE"x";H5T_IEEE_F64LE"y";H5T_IEEE_F64LE"z";}}' in flat
        assert 'H5T_VLEN{H5T_STD_U16LE}' in flat and 'H5T_VLEN{H5T_STD_U64LE}' in flat


def test_topologies_observable_round_trip(tmp_path):
    """Topologies::append (Topologies.cpp:136-170): particles with a count prefix per topology, edge rows
    with an (edge count, 0) prefix, [begin, end) limits per frame, the topology types per frame."""
    frames = [[(0, [0, 1, 2], [(0, 1), (1, 2)])],
              [(0, [0, 1], [(0, 1)]), (0, [2], [])],        # a broken bond: two topologies
              []]                                           # no topologies at all
    path = str(tmp_path / "tops.h5")
    with h5io.File(path) as f:
        f.topologies("topologies", [0, 10, 20], frames)
    t, got = h5io.read_topologies(path, "topologies")
    assert list(t) == [0, 10, 20]
    for want, have in zip(frames, got):
        assert [(ty, list(idx), ed) for ty, idx, ed in want] == [(ty, list(idx), ed) for ty, idx, ed in have]
    g = "/readdy/observables/topologies/"
    assert list(h5io.read_u64(path, g + "particles")) == [3, 0, 1, 2, 2, 0, 1, 1, 2]
    assert h5io.read_u64(path, g + "limitsParticles").reshape(-1, 2).tolist() == [[0, 4], [4, 9], [9, 9]]
    assert h5io.read_u64(path, g + "edges").reshape(-1, 2).tolist() == [[2, 0], [0, 1], [1, 2], [1, 0], [0, 1],
                                                                          [0, 0]]
    assert h5io.read_u64(path, g + "limitsEdges").reshape(-1, 2).tolist() == [[0, 3], [3, 6], [6, 6]]


def test_writers_append_to_chunked_series(tmp_path):
    """A writer called again for one observable appends (the flat trajectory's limits and the topologies'
    limits continue after the earlier frames); every time-indexed dataset is chunked with the file's
    chunk rows and unlimited along time, as the reference creates them with its flushStride."""
    path = str(tmp_path / "app.h5")
    rng = np.random.default_rng(8)
    frames = _frames(rng, [3, 0, 2, 4, 1])
    tops = [[(0, [0, 1], [(0, 1)])], [(1, [2, 3, 4], [(0, 1), (1, 2)]), (0, [5], [])], [],
            [(2, [6, 7], [(0, 1)])]]
    with h5io.File(path, chunk=4) as f:
        f.n_particles("n", [0, 1, 2], np.arange(6).reshape(3, 2))
        f.n_particles("n", [3, 4], np.arange(6, 10).reshape(2, 2))
        f.reaction_counts("rc", [0], [[1, 2]])
        f.reaction_counts("rc", [1, 2], [[3, 4], [5, 6]])
        f.flat_trajectory("traj", [0, 1], frames[:2])
        f.flat_trajectory("traj", [2, 3, 4], frames[2:])
        f.topologies("tops", [0, 1], tops[:2])
        f.topologies("tops", [2, 3], tops[2:])
        f.positions("pos", [0], [frames[0][3]])
        f.positions("pos", [1, 2], [frames[2][3], frames[3][3]])
    assert list(h5io.read_u64(path, "/readdy/observables/n/time")) == [0, 1, 2, 3, 4]
    assert list(h5io.read_u64(path, "/readdy/observables/n/data")) == list(range(10))
    assert list(h5io.read_u64(path, "/readdy/observables/rc/counts/1")) == [2, 4, 6]
    t, got = h5io.read_trajectory(path, "traj")
    assert list(t) == [0, 1, 2, 3, 4]
    for (ids, ty, fl, pos), want in zip(got, frames):
        assert np.array_equal(ids, want[0]) and np.array_equal(pos, np.asarray(want[3]).reshape(-1, 3))
    tt, tf = h5io.read_topologies(path, "tops")
    assert list(tt) == [0, 1, 2, 3]
    for got_f, want_f in zip(tf, tops):
        assert [(ty, list(map(int, idx)), ed) for ty, idx, ed in got_f] == \
               [(ty, idx, [tuple(e) for e in ed]) for ty, idx, ed in want_f]
    pv = h5io.read_vlen(path, "/readdy/observables/pos/data", 0)
    assert len(pv) == 3 and np.array_equal(pv[2], np.asarray(frames[3][3]).reshape(-1, 3))
    exe = _h5dump()
    if exe:
        out = subprocess.run([exe, "-p", "-H", "-d", "/readdy/observables/n/data", path], capture_output=True,
                             text=True, check=True).stdout
        assert "CHUNKED ( 4, 2 )" in out and "H5S_UNLIMITED" in out, out


def test_energy_and_virial_series(tmp_path):
    """Energy's data is one value per frame ([T], Energy.cpp:69-71), Virial's a 3x3 matrix per frame
    ([T][3][3], Virial.cpp:76-78); both append like the other series"""
    path = str(tmp_path / "ev.h5")
    v = np.arange(27, dtype=np.float64).reshape(3, 3, 3) * 0.5
    with h5io.File(path, chunk=2) as f:
        f.energy("energy", [0, 5], [1.5, -2.25])
        f.energy("energy", [10], [3.0])
        f.virial("virial", [0], v[:1])
        f.virial("virial", [5, 10], v[1:])
    assert list(h5io.read_u64(path, "/readdy/observables/energy/time")) == [0, 5, 10]
    assert list(h5io.read_f64(path, "/readdy/observables/energy/data")) == [1.5, -2.25, 3.0]
    assert np.array_equal(h5io.read_f64(path, "/readdy/observables/virial/data"), v.ravel())
    exe = _h5dump()
    if exe:
        out = subprocess.run([exe, "-H", "-d", "/readdy/observables/energy/data", path], capture_output=True,
                             text=True, check=True).stdout
        assert "( 3 ) / ( H5S_UNLIMITED )" in out, out
        out = subprocess.run([exe, "-H", "-d", "/readdy/observables/virial/data", path], capture_output=True,
                             text=True, check=True).stdout
        assert "( 3, 3, 3 ) / ( H5S_UNLIMITED, 3, 3 )" in out, out
