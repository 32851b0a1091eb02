"""ctypes binding of libreaddy_h5.so: observable and trajectory datasets in ReaDDy's HDF5 layout.

Writers (one file per call sequence: create, datasets, attributes, close):
  n_particles          /readdy/observables/<name>/{data [T][n_types], time [T]}
                       (NParticles, readdy/main/model/observables/Observables.cpp)
  reaction_counts      /readdy/observables/<name>/{counts/<reaction id> [T], time [T]}
                       (ReactionCounts, readdy/main/model/observables/Observables.cpp)
  radial_distribution  /readdy/observables/<name>/{bin_centers [B], distribution [T][B], time [T]}
                       (RadialDistribution.cpp:158-170)
  histogram_along_axis /readdy/observables/<name>/{data [T][B] f64, time [T]} (HistogramAlongAxis.cpp:83-96)
  positions, forces     /readdy/observables/<name>/{data [T] (variable-length lists of {x, y, z}), time}
                       (Positions.cpp:71-86, Forces.cpp:80-91)
  particles            /readdy/observables/<name>/{types, ids, positions [T] (variable-length), time}
                       (Particles.cpp:62-92)
  topologies           /readdy/observables/<name>/{particles, limitsParticles [T][2], edges [E][2],
                       limitsEdges [T][2], types [T] (variable-length), time} (Topologies.cpp:60-170)
  reactions            /readdy/observables/<name>/{records [T] (variable-length lists of reaction
                       records), time [T]} (Reactions.cpp:72-86; record fields as in Types.cpp:66-77)
  flat_trajectory      /readdy/trajectory/<name>/{records [R], limits [T][2], time [T]}
                       (FlatTrajectory, Trajectory.cpp:112-158; records are {id u64, typeId u16,
                       flavor u8, pos f64[3]} as in Types.cpp:79-88, limits a frame's [begin, end))
  config               /readdy/config/{particle_types, registered_reactions, general}
                       (IOUtils.cpp:58-176: what readdy.Trajectory reads for type names, reactions,
                       kBT, box and periodicity)
Every time-indexed dataset is chunked and unlimited along time (chunk rows: File.set_chunk, the
reference's flushStride), and a writer called again for the same observable appends its frames, so a
recorder can flush every flushStride frames (Observable.h:157-199).
Readers (tests and tools): read_u64, read_f64, read_records, read_string, read_reactions, read_vlen.
"""
import ctypes
import json
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_lib = None

# a reaction record as read back (the fields of the reference's ReactionRecord type)
REACTION_DTYPE = np.dtype([("reaction", np.uint64), ("type", np.int32), ("educts", np.uint64, 2),
                           ("products", np.uint64, 2), ("types_from", np.uint16, 2), ("position", np.float64, 3)])


def lib():
    """The loaded library (raises OSError when it is not built)."""
    global _lib
    if _lib is None:
        h = ctypes.CDLL(os.path.join(_HERE, "libreaddy_h5.so"))
        h.rdy_h5_create.restype = ctypes.c_int64
        h.rdy_h5_create.argtypes = [ctypes.c_char_p]
        h.rdy_h5_set_chunk.argtypes = [ctypes.c_int64, ctypes.c_uint64]
        for fn in ("rdy_h5_n_particles", "rdy_h5_reaction_counts"):
            getattr(h, fn).argtypes = [ctypes.c_int64, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint32, _u64p,
                                       _u64p]
        h.rdy_h5_radial_distribution.argtypes = [ctypes.c_int64, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint32,
                                                 _u64p, _f64p, _f64p]
        h.rdy_h5_flat_trajectory.argtypes = [ctypes.c_int64, ctypes.c_char_p, ctypes.c_uint64, _u64p, _u64p, _u64p,
                                             _u16p, _u8p, _f64p]
        cpp = ctypes.POINTER(ctypes.c_char_p)
        h.rdy_h5_particle_types.argtypes = [ctypes.c_int64, ctypes.c_uint32, cpp, _f64p, cpp]
        h.rdy_h5_registered_reactions.argtypes = [ctypes.c_int64, ctypes.c_uint32, cpp, _u64p, _u64p, _u64p, _f64p,
                                                  _f64p, _f64p, _u16p, _u16p]
        h.rdy_h5_general.argtypes = [ctypes.c_int64, ctypes.c_char_p]
        h.rdy_h5_read_string.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_char_p]
        h.rdy_h5_reactions.argtypes = [ctypes.c_int64, ctypes.c_char_p, ctypes.c_uint64, _u64p, _u64p, _u64p, _i32p,
                                       _u64p, _u64p, _f64p, _u16p]
        h.rdy_h5_read_reactions.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint64, _u64p, ctypes.c_uint64,
                                            _u64p, _i32p, _u64p, _u64p, _f64p, _u16p, _u64p, _u64p]
        h.rdy_h5_series_f64.argtypes = [ctypes.c_int64, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint32, _u64p,
                                        _f64p]
        h.rdy_h5_series_f64_nd.argtypes = [ctypes.c_int64, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint32,
                                           ctypes.POINTER(ctypes.c_uint32), _u64p, _f64p]
        h.rdy_h5_u64_table.argtypes = [ctypes.c_int64, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint64,
                                       ctypes.c_uint32, _u64p]
        h.rdy_h5_vlen_series.argtypes = [ctypes.c_int64, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                         ctypes.c_uint64, _u64p, ctypes.c_void_p, _u64p]
        h.rdy_h5_read_vlen.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_uint64, _u64p,
                                       ctypes.c_uint64, ctypes.c_void_p, _u64p, _u64p]
        h.rdy_h5_attribute.argtypes = [ctypes.c_int64, ctypes.c_char_p, ctypes.c_char_p]
        h.rdy_h5_close.argtypes = [ctypes.c_int64]
        h.rdy_h5_read_u64.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint64, _u64p, _u64p]
        h.rdy_h5_read_f64.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint64, _f64p, _u64p]
        h.rdy_h5_read_records.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint64, _u64p, _u16p, _u8p,
                                          _f64p, _u64p]
        _lib = h
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


class File:
    """One HDF5 file being written; every write raises on failure."""

    def __init__(self, path, chunk=64):
        self.path = path
        self.f = lib().rdy_h5_create(path.encode())
        if self.f < 0:
            raise OSError(f"cannot create {path}")
        self._top_totals = {}  # topologies observable -> (particle rows, edge rows) written so far
        self.set_chunk(chunk)

    def set_chunk(self, rows):
        """chunk rows (time points) of the datasets created from now on: the reference's flushStride"""
        self._check(lib().rdy_h5_set_chunk(self.f, int(rows)), "chunk size")

    def _check(self, rc, what):
        if rc:
            raise OSError(f"HDF5 write of {what} failed in {self.path}")

    def n_particles(self, name, time, counts):
        t = np.ascontiguousarray(time, dtype=np.uint64)
        c = np.ascontiguousarray(counts, dtype=np.uint64).reshape(len(t), -1)
        self._check(lib().rdy_h5_n_particles(self.f, name.encode(), len(t), c.shape[1], _p(t, _u64p),
                                             _p(c, _u64p)), name)

    def reaction_counts(self, name, time, counts):
        t = np.ascontiguousarray(time, dtype=np.uint64)
        c = np.ascontiguousarray(counts, dtype=np.uint64).reshape(len(t), -1)
        self._check(lib().rdy_h5_reaction_counts(self.f, name.encode(), len(t), c.shape[1], _p(t, _u64p),
                                                 _p(c, _u64p)), name)

    def radial_distribution(self, name, time, bin_centers, distribution):
        t = np.ascontiguousarray(time, dtype=np.uint64)
        c = np.ascontiguousarray(bin_centers, dtype=np.float64)
        d = np.ascontiguousarray(distribution, dtype=np.float64).reshape(len(t), len(c))
        self._check(lib().rdy_h5_radial_distribution(self.f, name.encode(), len(t), len(c), _p(t, _u64p),
                                                     _p(c, _f64p), _p(d, _f64p)), name)

    def histogram_along_axis(self, name, time, data):
        t = np.ascontiguousarray(time, dtype=np.uint64)
        d = np.ascontiguousarray(data, dtype=np.float64).reshape(len(t), -1)
        self._check(lib().rdy_h5_series_f64(self.f, name.encode(), len(t), d.shape[1], _p(t, _u64p), _p(d, _f64p)),
                    name)

    def energy(self, name, time, values):
        """Energy (Energy.cpp:69-71): data [T], one value per frame"""
        self._series_nd(name, time, np.asarray(values, dtype=np.float64).reshape(len(time)))

    def virial(self, name, time, values):
        """Virial (Virial.cpp:76-78): data [T][3][3]"""
        self._series_nd(name, time, np.asarray(values, dtype=np.float64).reshape(len(time), 3, 3))

    def _series_nd(self, name, time, data):
        t = np.ascontiguousarray(time, dtype=np.uint64)
        d = np.ascontiguousarray(data, dtype=np.float64)
        dims = (ctypes.c_uint32 * 3)(*(list(d.shape[1:]) + [0, 0, 0])[:3])
        self._check(lib().rdy_h5_series_f64_nd(self.f, name.encode(), len(t), d.ndim - 1, dims, _p(t, _u64p),
                                               _p(d, _f64p)), name)

    def _vlen(self, name, dataset, elem, time, frames, with_time):
        dt, width = _VLEN[elem]
        t = np.ascontiguousarray(time, dtype=np.uint64)
        if len(frames) != len(t):
            raise ValueError("one frame per time point")
        sizes = np.array([len(f) for f in frames], dtype=np.uint64)
        data = (np.ascontiguousarray(np.concatenate([np.asarray(f, dtype=dt).reshape(-1, width) for f in frames]))
                if int(sizes.sum()) else np.zeros((1, width), dt))
        self._check(lib().rdy_h5_vlen_series(self.f, name.encode(), dataset.encode(), elem, len(t), _p(sizes, _u64p),
                                             data.ctypes.data_as(ctypes.c_void_p), _p(t, _u64p) if with_time else None),
                    f"{name}/{dataset}")

    def positions(self, name, time, frames):
        """frames: one [n][3] array of positions per entry of time."""
        self._vlen(name, "data", 0, time, frames, True)

    def forces(self, name, time, frames):
        self._vlen(name, "data", 0, time, frames, True)

    def particles(self, name, time, frames):
        """frames: one (type ids, ids, positions [n][3]) tuple per entry of time."""
        self._vlen(name, "types", 1, time, [f[0] for f in frames], False)
        self._vlen(name, "ids", 2, time, [f[1] for f in frames], False)
        self._vlen(name, "positions", 0, time, [f[2] for f in frames], True)

    def topologies(self, name, time, frames):
        """frames: per time point a list of topologies (type id, particle indices, edges as vertex pairs).
        The layout of Topologies::append (Topologies.cpp:136-170): per topology its particle count then its
        particles, and an edge row (edge count, 0) then its edges; limits are each frame's [begin, end)."""
        t = np.ascontiguousarray(time, dtype=np.uint64)
        if len(frames) != len(t):
            raise ValueError("one frame per time point")
        parts, edges, lp, le, types = [], [], [], [], []
        done_p, done_e = self._top_totals.get(name, (0, 0))  # rows of earlier appends of this observable
        for tops in frames:
            p0, e0 = len(parts), len(edges)
            for ty, idx, ed in tops:
                parts.append(len(idx))
                parts.extend(int(i) for i in idx)
                edges.append((len(ed), 0))
                edges.extend((int(a), int(b)) for a, b in ed)
            lp.append((done_p + p0, done_p + len(parts)))
            le.append((done_e + e0, done_e + len(edges)))
            types.append(np.array([ty for ty, _, _ in tops], dtype=np.uint16))

        def table(dataset, rows, cols):
            a = np.ascontiguousarray(np.asarray(rows, dtype=np.uint64).reshape(-1, cols) if len(rows) else
                                     np.zeros((0, cols), np.uint64))
            self._check(lib().rdy_h5_u64_table(self.f, name.encode(), dataset.encode(), a.shape[0], cols,
                                               _p(a, _u64p)), f"{name}/{dataset}")
        table("particles", parts, 1)
        table("limitsParticles", lp, 2)
        table("edges", edges, 2)
        table("limitsEdges", le, 2)
        self._vlen(name, "types", 1, time, types, True)
        self._top_totals[name] = (done_p + len(parts), done_e + len(edges))

    def flat_trajectory(self, name, time, frames):
        """frames: one (ids, type ids, flavors, positions [n][3]) tuple per entry of time."""
        t = np.ascontiguousarray(time, dtype=np.uint64)
        if len(frames) != len(t):
            raise ValueError("one frame per time point")
        sizes = np.array([len(f[0]) for f in frames], dtype=np.uint64)
        cat = lambda k, dt: np.ascontiguousarray(np.concatenate([np.asarray(f[k], dtype=dt).ravel() for f in frames])
                                                 if frames else np.zeros(0, dt))
        ids, types, flav, pos = cat(0, np.uint64), cat(1, np.uint16), cat(2, np.uint8), cat(3, np.float64)
        if not (len(ids) == len(types) == len(flav) == len(pos) // 3 == int(sizes.sum())):
            raise ValueError("frame arrays of unequal length")
        self._check(lib().rdy_h5_flat_trajectory(self.f, name.encode(), len(t), _p(t, _u64p), _p(sizes, _u64p),
                                                 _p(ids, _u64p), _p(types, _u16p), _p(flav, _u8p),
                                                 _p(pos, _f64p)), name)

    def reactions(self, name, time, frames):
        """frames: one array of reaction records per entry of time, each with the fields reaction, type,
        educts [2], products [2], types_from [2], position [3] (Engine.reaction_records' layout)."""
        t = np.ascontiguousarray(time, dtype=np.uint64)
        if len(frames) != len(t):
            raise ValueError("one frame per time point")
        sizes = np.array([len(f) for f in frames], dtype=np.uint64)
        cat = lambda k, dt, w: (np.ascontiguousarray(np.concatenate([np.asarray(f[k], dtype=dt).reshape(-1, w)
                                                                     for f in frames]), dtype=dt)
                                if int(sizes.sum()) else np.zeros((0, w), dt))
        rid, typ = cat("reaction", np.uint64, 1), cat("type", np.int32, 1)
        ed, pr = cat("educts", np.uint64, 2), cat("products", np.uint64, 2)
        tf, pos = cat("types_from", np.uint16, 2), cat("position", np.float64, 3)
        self._check(lib().rdy_h5_reactions(self.f, name.encode(), len(t), _p(t, _u64p), _p(sizes, _u64p),
                                           _p(rid, _u64p), _p(typ, _i32p), _p(ed, _u64p), _p(pr, _u64p),
                                           _p(pos, _f64p), _p(tf, _u16p)), name)

    def config(self, types, box, pbc=(True, True, True), kbt=1.0, reactions=()):
        """types: (name, diffusion constant[, flavor]) per type id; reactions: dicts with name, id, educts,
        products (type ids), rate, educt_distance, product_distance."""
        cstr = lambda xs: (ctypes.c_char_p * max(1, len(xs)))(*[x.encode() for x in xs])
        names = [t[0] for t in types]
        flav = [t[2] if len(t) > 2 else "NORMAL" for t in types]
        diff = np.ascontiguousarray([t[1] for t in types], dtype=np.float64)
        self._check(lib().rdy_h5_particle_types(self.f, len(types), cstr(names), _p(diff, _f64p), cstr(flav)),
                    "particle_types")
        if reactions:
            col = lambda k, dt: np.ascontiguousarray([r[k] for r in reactions], dtype=dt)
            pad = lambda xs: (list(xs) + [0, 0])[:2]
            ed = np.ascontiguousarray([pad(r.get("educts", [])) for r in reactions], dtype=np.uint16)
            pr = np.ascontiguousarray([pad(r.get("products", [])) for r in reactions], dtype=np.uint16)
            ne = np.ascontiguousarray([len(r.get("educts", [])) for r in reactions], dtype=np.uint64)
            npr = np.ascontiguousarray([len(r.get("products", [])) for r in reactions], dtype=np.uint64)
            ids, rate = col("id", np.uint64), col("rate", np.float64)
            edist = np.ascontiguousarray([r.get("educt_distance", 0.0) for r in reactions], dtype=np.float64)
            pdist = np.ascontiguousarray([r.get("product_distance", 0.0) for r in reactions], dtype=np.float64)
            self._check(lib().rdy_h5_registered_reactions(
                self.f, len(reactions), cstr([r["name"] for r in reactions]), _p(ids, _u64p), _p(ne, _u64p),
                _p(npr, _u64p), _p(rate, _f64p), _p(edist, _f64p), _p(pdist, _f64p), _p(ed, _u16p),
                _p(pr, _u16p)), "registered_reactions")
        box = [float(b) for b in box]
        general = {"kbt": float(kbt), "box_volume": float(np.prod(box)), "box_size": box,
                   "pbc": [bool(p) for p in pbc]}
        self._check(lib().rdy_h5_general(self.f, json.dumps(general).encode()), "general")

    def attribute(self, key, value):
        self._check(lib().rdy_h5_attribute(self.f, key.encode(), str(value).encode()), key)

    def close(self):
        if self.f >= 0:
            rc = lib().rdy_h5_close(self.f)
            self.f = -1
            self._check(rc, "close")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _size(path, dataset):
    """Element count of a dataset (a read into a zero-capacity buffer reports it)."""
    n = ctypes.c_uint64(1 << 63)
    one = np.zeros(1, np.uint64)
    lib().rdy_h5_read_u64(path.encode(), dataset.encode(), 0, _p(one, _u64p), ctypes.byref(n))
    if n.value == 1 << 63:
        raise OSError(f"no dataset {dataset} in {path}")
    return n.value


def read_u64(path, dataset):
    out = np.zeros(max(1, _size(path, dataset)), dtype=np.uint64)
    n = ctypes.c_uint64()
    if lib().rdy_h5_read_u64(path.encode(), dataset.encode(), len(out), _p(out, _u64p), ctypes.byref(n)):
        raise OSError(f"cannot read {dataset} of {path}")
    return out[:n.value]


def read_f64(path, dataset):
    out = np.zeros(max(1, _size(path, dataset)), dtype=np.float64)
    n = ctypes.c_uint64()
    if lib().rdy_h5_read_f64(path.encode(), dataset.encode(), len(out), _p(out, _f64p), ctypes.byref(n)):
        raise OSError(f"cannot read {dataset} of {path}")
    return out[:n.value]


def read_records(path, dataset):
    """(ids, type ids, flavors, positions [R][3]) of a trajectory's records dataset."""
    cap = max(1, _size(path, dataset))
    ids = np.zeros(cap, np.uint64)
    types = np.zeros(cap, np.uint16)
    flav = np.zeros(cap, np.uint8)
    pos = np.zeros((cap, 3), np.float64)
    n = ctypes.c_uint64()
    if lib().rdy_h5_read_records(path.encode(), dataset.encode(), cap, _p(ids, _u64p), _p(types, _u16p),
                                 _p(flav, _u8p), _p(pos, _f64p), ctypes.byref(n)):
        raise OSError(f"cannot read {dataset} of {path}")
    k = n.value
    return ids[:k], types[:k], flav[:k], pos[:k]


def read_string(path, dataset, cap=1 << 16):
    buf = ctypes.create_string_buffer(cap)
    if lib().rdy_h5_read_string(path.encode(), dataset.encode(), cap, buf):
        raise OSError(f"cannot read {dataset} of {path}")
    return buf.value.decode()


def read_reactions(path, name):
    """(time, [records per frame]) of /readdy/observables/<name> (REACTION_DTYPE)."""
    g = f"/readdy/observables/{name}"
    t = read_u64(path, g + "/time")
    nf, n = ctypes.c_uint64(), ctypes.c_uint64()
    sizes = np.zeros(max(1, len(t)), np.uint64)
    one = np.zeros(1, np.uint64)
    # first pass: the record count (a zero-capacity read reports it)
    lib().rdy_h5_read_reactions(path.encode(), (g + "/records").encode(), len(sizes), _p(sizes, _u64p), 0,
                                _p(one, _u64p), None, None, None, None, None, ctypes.byref(nf), ctypes.byref(n))
    cap = max(1, n.value)
    rid, typ = np.zeros(cap, np.uint64), np.zeros(cap, np.int32)
    ed, pr = np.zeros((cap, 2), np.uint64), np.zeros((cap, 2), np.uint64)
    tf, pos = np.zeros((cap, 2), np.uint16), np.zeros((cap, 3), np.float64)
    if lib().rdy_h5_read_reactions(path.encode(), (g + "/records").encode(), len(sizes), _p(sizes, _u64p), cap,
                                   _p(rid, _u64p), _p(typ, _i32p), _p(ed, _u64p), _p(pr, _u64p), _p(pos, _f64p),
                                   _p(tf, _u16p), ctypes.byref(nf), ctypes.byref(n)):
        raise OSError(f"cannot read {g}/records of {path}")
    out = np.zeros(n.value, REACTION_DTYPE)
    out["reaction"], out["type"], out["educts"] = rid[:n.value], typ[:n.value], ed[:n.value]
    out["products"], out["types_from"], out["position"] = pr[:n.value], tf[:n.value], pos[:n.value]
    ends = np.cumsum(sizes[:nf.value].astype(np.int64))
    return t, [out[b - int(s):b] for b, s in zip(ends, sizes[:nf.value])]


_VLEN = {0: (np.float64, 3), 1: (np.uint16, 1), 2: (np.uint64, 1)}


def read_vlen(path, dataset, elem):
    """The frames of a variable-length series (elem 0: {x, y, z} rows, 1: uint16, 2: uint64)."""
    dt, width = _VLEN[elem]
    nf, n = ctypes.c_uint64(), ctypes.c_uint64()
    one = np.zeros(1, np.uint64)
    lib().rdy_h5_read_vlen(path.encode(), dataset.encode(), elem, 0, _p(one, _u64p), 0, None, ctypes.byref(nf),
                           ctypes.byref(n))
    sizes = np.zeros(max(1, nf.value), np.uint64)
    out = np.zeros((max(1, n.value), width), dt)
    if lib().rdy_h5_read_vlen(path.encode(), dataset.encode(), elem, len(sizes), _p(sizes, _u64p), n.value,
                              out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nf), ctypes.byref(n)):
        raise OSError(f"cannot read {dataset} of {path}")
    ends = np.cumsum(sizes[:nf.value].astype(np.int64))
    frames = [out[b - int(s):b] for b, s in zip(ends, sizes[:nf.value])]
    return frames if width == 3 else [f.ravel() for f in frames]


def read_topologies(path, name):
    """(time, frames) of /readdy/observables/<name>, frames as written by File.topologies."""
    g = f"/readdy/observables/{name}/"
    t = read_u64(path, g + "time")
    parts = read_u64(path, g + "particles")
    lp = read_u64(path, g + "limitsParticles").reshape(-1, 2)
    edges = read_u64(path, g + "edges").reshape(-1, 2)
    le = read_u64(path, g + "limitsEdges").reshape(-1, 2)
    types = read_vlen(path, g + "types", 1)
    frames = []
    for k in range(len(t)):
        tops, p, e = [], int(lp[k, 0]), int(le[k, 0])
        for ty in types[k]:
            n = int(parts[p])
            idx = parts[p + 1:p + 1 + n]
            ne = int(edges[e, 0])
            tops.append((int(ty), idx, [tuple(int(v) for v in row) for row in edges[e + 1:e + 1 + ne]]))
            p += 1 + n
            e += 1 + ne
        assert p == lp[k, 1] and e == le[k, 1]
        frames.append(tops)
    return t, frames


def read_trajectory(path, name):
    """The frames of /readdy/trajectory/<name>: (time, [(ids, types, flavors, positions) per frame])."""
    g = f"/readdy/trajectory/{name}" if name else "/readdy/trajectory"
    t = read_u64(path, g + "/time")
    lim = read_u64(path, g + "/limits").reshape(len(t), 2)
    ids, types, flav, pos = read_records(path, g + "/records")
    return t, [(ids[a:b], types[a:b], flav[a:b], pos[a:b]) for a, b in lim]
