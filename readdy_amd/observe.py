"""Observables of one replica recorded at a stride and written to HDF5 in ReaDDy's layout.

The reference's simulation loop evaluates every registered observable before the first step and after
each step t whose number is a multiple of the observable's stride (SimulationLoop.h:268-300,
Observable.h:144-147: stride 0 means every step).  Recorder does the same around the engine's fused
steps: it runs the replica in chunks that end where the next observable is due, evaluates the due ones
on the device state, and keeps the series on the host; write() puts them into an HDF5 file with the
configuration group that readdy.Trajectory reads.  open() / close() instead stream the series into a
file: every observable's frames are appended each flush_stride evaluations and leave host memory, as
the reference's observables flush their buffered frames every flushStride (Observable.h:157-199).

  n_particles           NParticles: particles per type                       /readdy/observables/<name>
  reaction_counts       ReactionCounts: events per reaction of the last step  (its handler call)
  reactions             Reactions: the last step's reaction records
  radial_distribution   RadialDistribution (RadialDistribution.cpp:75-124), on the device
  histogram_along_axis  HistogramAlongAxis (CPUObservables.cpp:85-136), on the device
  energy                Energy: the potential energy of the state (pair and bonded terms)
  virial                Virial: the 3x3 pair virial (Context::recordVirial is switched on)
  positions, forces     Positions / Forces: every particle's position or force (ordered by id)
  particles             Particles: every particle's type, id and position
  trajectory            FlatTrajectory: every particle's id, type and position /readdy/trajectory/<name>
"""
import math

import numpy as np

from . import h5io


class Recorder:
    def __init__(self, engine, replica=0):
        self.e = engine
        self.r = replica
        self.t = 0                 # steps run through this recorder
        self._obs = []             # dicts: kind, name, stride, args, time, values
        self._evaluated_at = None  # the step whose observables were evaluated last
        self._file = None          # h5io.File while streaming (open() .. close())
        self._flush_stride = 0
        self._flavors = None

    def _add(self, kind, name, stride, callback=None, **args):
        if any(o["name"] == name and o["group"] == _group(kind) for o in self._obs):
            raise ValueError(f"an observable named {name!r} is already recorded")
        if kind == "reactions":
            self.e.record_reactions(True)
        self._obs.append(dict(kind=kind, name=name, group=_group(kind), stride=int(stride), args=args, time=[],
                              values=[], callback=callback))

    def n_particles(self, stride, name="n_particles", types=None, callback=None):
        """particles per type; types: the listed type ids only, in that order, or (an empty list) their total
        (CPUNParticles::evaluate); None: every type.  callback(value) runs at each evaluation."""
        self._add("n_particles", name, stride, callback, types=None if types is None else list(types))

    def reaction_counts(self, stride, name="reaction_counts", callback=None):
        self._add("reaction_counts", name, stride, callback)

    def reactions(self, stride, name="reactions", callback=None):
        self._add("reactions", name, stride, callback)

    def radial_distribution(self, stride, borders, from_types, to_types, density, name="radial_distribution",
                            callback=None):
        self._add("rdf", name, stride, callback, borders=np.asarray(borders, dtype=np.float64), frm=list(from_types),
                  to=list(to_types), density=float(density))

    def histogram_along_axis(self, stride, borders, types, axis, name="histogram_along_axis", callback=None):
        self._add("histogram", name, stride, callback, borders=np.asarray(borders, dtype=np.float64),
                  types=list(types), axis=int(axis))

    def energy(self, stride, name="energy", callback=None):
        self._add("energy", name, stride, callback)

    def virial(self, stride, name="virial", callback=None):
        self.e.record_virial(True)
        self._add("virial", name, stride, callback)

    def positions(self, stride, name="positions", types=None, callback=None):
        """positions ordered by id; types: only particles of these type ids"""
        self._add("positions", name, stride, callback, types=None if types is None else list(types))

    def forces(self, stride, name="forces", types=None, callback=None):
        self._add("forces", name, stride, callback, types=None if types is None else list(types))

    def particles(self, stride, name="particles", callback=None):
        self._add("particles", name, stride, callback)

    def trajectory(self, stride, name=""):
        self._add("trajectory", name, stride)

    # ------------------------------------------------------------------------------------ running
    def _due(self, o, t):
        return o["stride"] == 0 or t % o["stride"] == 0

    def _evaluate(self, t):
        if self._evaluated_at == t:
            return
        self._evaluated_at = t
        e, r = self.e, self.r
        for o in self._obs:
            if not self._due(o, t):
                continue
            k, a = o["kind"], o["args"]
            if k == "n_particles":
                v = e.count_types(r + 1)[r].copy()
                if a.get("types") is not None:  # the listed types, or their total when none is listed
                    v = v[a["types"]] if a["types"] else np.array([v.sum()], dtype=v.dtype)
            elif k == "reaction_counts":
                v = e.reaction_counts(r).copy()
            elif k == "reactions":
                v = e.reaction_records(r)
            elif k == "rdf":
                v = e.radial_distribution(r, a["borders"], a["frm"], a["to"], a["density"])[1]
            elif k == "histogram":
                v = e.histogram_along_axis(r, a["borders"], a["types"], a["axis"]).astype(np.float64)
            elif k == "energy":
                v = e.energy(r)
            elif k == "virial":
                v = e.virial(r)
            elif k == "positions":
                pos, types, _ = e.get_particles(r)
                v = pos if a.get("types") is None else pos[np.isin(types, a["types"])]
            elif k == "forces":
                f, h = e.get_forces(r)
                v = f[np.argsort(h, kind="stable")]
                if a.get("types") is not None:
                    v = v[np.isin(e.get_particles(r)[1], a["types"])]
            elif k == "particles":
                pos, types, handles = e.get_particles(r)
                v = (types, handles, pos)
            else:  # trajectory: (handles, types, positions)
                pos, types, handles = e.get_particles(r)
                v = (handles, types, pos)
            o["time"].append(t)
            o["values"].append(v)
            if o["callback"] is not None:
                o["callback"](v)
            if self._file is not None and len(o["time"]) >= self._flush_stride:
                self._flush(o)

    def _next_due(self, t):
        nxt = [t + 1 if o["stride"] == 0 else (t // o["stride"] + 1) * o["stride"] for o in self._obs]
        return min(nxt) if nxt else math.inf

    def run(self, handler, dt, n_steps, graph=True):
        """n_steps fused steps of the replica's engine, observables evaluated where due (and at the
        recorder's first step 0)."""
        if self._evaluated_at is None and any(o["kind"] in ("forces", "energy", "virial") for o in self._obs):
            self.e.calculate_forces()  # the loop forms the initial forces before step 0's observables
        self._evaluate(self.t)
        end = self.t + int(n_steps)
        while self.t < end:
            k = int(min(self._next_due(self.t), end) - self.t)
            self.e.run(handler, dt, k, graph=graph)
            self.t += k
            self._evaluate(self.t)

    def series(self, name, kind=None):
        """(time, values) of the observable `name`."""
        for o in self._obs:
            if o["name"] == name and (kind is None or o["kind"] == kind):
                return np.asarray(o["time"], dtype=np.uint64), o["values"]
        raise KeyError(name)

    # ------------------------------------------------------------------------------------ output
    def open(self, path, type_names, box, pbc=(True, True, True), kbt=1.0, diffusion=None, reactions=(),
             topology_types=(), flush_stride=64):
        """Streams the series into one HDF5 file: /readdy/config now, then every observable's frames in
        appends of flush_stride evaluations (also the datasets' chunk rows); close() writes the rest.
        reactions: the engine's reaction dicts (kind, educts, products, rate, ...) with a "name"."""
        if self._file is not None:
            raise RuntimeError("the recorder already streams into a file")
        if int(flush_stride) < 1:
            raise ValueError("flush_stride must be at least 1")
        diffusion = [0.0] * len(type_names) if diffusion is None else list(diffusion)
        f = h5io.File(path, chunk=int(flush_stride))
        try:
            f.config([(n, d, "TOPOLOGY" if n in topology_types else "NORMAL") for n, d in zip(type_names, diffusion)],
                     box, pbc, kbt, [dict(rx, id=i, name=rx.get("name", f"reaction_{i}")) for i, rx in
                                     enumerate(reactions)])
        except Exception:
            f.close()
            raise
        self._file, self._flush_stride = f, int(flush_stride)
        # topology particles would carry flavor 1 (TOPOLOGY)
        self._flavors = np.array([1 if n in topology_types else 0 for n in type_names], dtype=np.uint8)
        for o in self._obs:  # frames recorded before open() go first
            if o["time"]:
                self._flush(o)

    def _flush(self, o):
        """appends the buffered frames of observable o to the open file and drops them"""
        f, k, t, v = self._file, o["kind"], o["time"], o["values"]
        if k == "n_particles":
            f.n_particles(o["name"], t, np.asarray(v, dtype=np.uint64).reshape(len(t), -1))
        elif k == "reaction_counts":
            f.reaction_counts(o["name"], t, np.asarray(v, dtype=np.uint64).reshape(len(t), -1))
        elif k == "reactions":
            f.reactions(o["name"], t, v)
        elif k == "rdf":
            b = o["args"]["borders"]
            f.radial_distribution(o["name"], t, (b[:-1] + b[1:]) / 2, np.asarray(v).reshape(len(t), -1))
        elif k == "histogram":
            f.histogram_along_axis(o["name"], t, np.asarray(v).reshape(len(t), -1))
        elif k == "energy":
            f.energy(o["name"], t, v)
        elif k == "virial":
            f.virial(o["name"], t, v)
        elif k == "positions":
            f.positions(o["name"], t, v)
        elif k == "forces":
            f.forces(o["name"], t, v)
        elif k == "particles":
            f.particles(o["name"], t, v)
        else:
            flav = self._flavors
            f.flat_trajectory(o["name"], t, [(h, ty, flav[np.asarray(ty, dtype=np.int64)], x) for h, ty, x in v])
        o["time"], o["values"] = [], []

    def close(self):
        """Appends the frames still buffered and closes the file open() started."""
        if self._file is None:
            return
        try:
            for o in self._obs:
                if o["time"]:
                    self._flush(o)
        finally:
            f, self._file = self._file, None
            f.close()

    def write(self, path, type_names, box, pbc=(True, True, True), kbt=1.0, diffusion=None, reactions=(),
              topology_types=()):
        """One HDF5 file: the recorded series plus /readdy/config (types, reactions, general).  The frames
        written leave the recorder, as after a stream's close()."""
        n = max([len(o["time"]) for o in self._obs] + [1])
        self.open(path, type_names, box, pbc, kbt, diffusion, reactions, topology_types, flush_stride=n)
        self.close()


def _group(kind):
    return "trajectory" if kind == "trajectory" else "observables"
