"""Python handle on the HIP engine (C ABI in include/readdy_amd.h).

An Engine is one chemistry (species, pair potentials, reactions) plus any number of
replicas that step together in shared launches.  All arrays cross the boundary as plain
NumPy buffers; device memory stays inside the library.
"""
import ctypes
import math

import numpy as np

from . import _lib as L

_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def bin_capacity_for(particles_per_cell):
    """Cell-bin capacity for an expected mean occupancy: Poisson mean + 7 sigma + 6, multiple of 4."""
    lam = max(float(particles_per_cell), 0.0)
    c = int(math.ceil(lam + 7.0 * math.sqrt(lam) + 6.0))
    return max(8, (c + 3) // 4 * 4)


def _potential_array(potentials):
    a = (L.PairPotential * max(len(potentials), 1))()
    for i, (kind, t1, t2, params) in enumerate(potentials):
        a[i].kind = kind
        a[i].type1 = t1
        a[i].type2 = t2
        for k, v in enumerate(params):
            a[i].params[k] = float(v)
    return a


def _reaction_array(reactions):
    a = (L.Reaction * max(len(reactions), 1))()
    for i, r in enumerate(reactions):
        x = a[i]
        x.kind = r["kind"]
        ed = list(r.get("educts", [])) + [0, 0]
        pr = list(r.get("products", [])) + [0, 0]
        x.educts[0], x.educts[1] = ed[0], ed[1]
        x.products[0], x.products[1] = pr[0], pr[1]
        x.rate = r["rate"]
        x.educt_distance = r.get("educt_distance", 0.0)
        x.product_distance = r.get("product_distance", 0.0)
        x.weight1 = r.get("weight1", 0.5)
        x.weight2 = r.get("weight2", 0.5)
    return a


REVERSIBLE_KINDS = {0: "FusionFission", 1: "ConversionConversion", 2: "EnzymaticEnzymatic"}


def reversible_reactions(n_types, potentials=(), reactions=(), kbt=1.0, total_volume=1.0):
    """DetailedBalance's reversible pairs of a chemistry (rdy_reversible_reactions; host computation, no
    GPU needed): one dict per pair with the ReversibleReactionConfig values
    (include/readdy/model/actions/DetailedBalance.h:64-139)."""
    lib = L.load()
    pots, rx = _potential_array(potentials), _reaction_array(reactions)
    diff = np.ones(max(int(n_types), 1))
    ctx = L.ContextDesc()
    ctx.n_types = int(n_types)
    ctx.kbt = kbt
    ctx.diffusion = _ptr(diff, _dp)
    ctx.topology_flavor = None
    ctx.n_potentials = len(potentials)
    ctx.potentials = pots
    ctx.n_reactions = len(reactions)
    ctx.reactions = rx
    n = ctypes.c_uint64()
    L.check(lib.rdy_reversible_reactions(ctypes.byref(ctx), float(total_volume), 0, None, ctypes.byref(n)))
    out = (L.ReversibleReaction * max(n.value, 1))()
    L.check(lib.rdy_reversible_reactions(ctypes.byref(ctx), float(total_volume), n.value, out, ctypes.byref(n)))
    res = []
    for c in out[:n.value]:
        d = {f: getattr(c, f) for f, _ in L.ReversibleReaction._fields_}
        d["lhs"], d["rhs"] = list(c.lhs), list(c.rhs)
        d["kind_name"] = REVERSIBLE_KINDS[c.kind]
        res.append(d)
    return res


class Engine:
    def __init__(self, n_types, diffusion, kbt=1.0, potentials=(), reactions=(), topology_flavor=None,
                 nl_cutoff=0.0, cell_radius=1, bin_capacity=0, max_replicas=1, max_cells=0, max_particles=0,
                 events_per_replica=0, block_size=0, brick=(0, 0, 0), lds_states=0):
        self._lib = L.load()
        self.n_types = int(n_types)
        self._diff = np.ascontiguousarray(diffusion, dtype=np.float64)
        assert self._diff.size == self.n_types
        self._pots = _potential_array(potentials)
        self._rx = _reaction_array(reactions)
        self.reaction_kinds = [r["kind"] for r in reactions]
        self.n_reactions = len(reactions)
        self._flavor = None
        if topology_flavor is not None:
            self._flavor = np.ascontiguousarray(topology_flavor, dtype=np.uint8)
        ctx = L.ContextDesc()
        ctx.n_types = self.n_types
        ctx.kbt = kbt
        ctx.diffusion = _ptr(self._diff, _dp)
        ctx.topology_flavor = _ptr(self._flavor, ctypes.POINTER(ctypes.c_uint8)) if self._flavor is not None else None
        ctx.n_potentials = len(potentials)
        ctx.potentials = self._pots
        ctx.n_reactions = len(reactions)
        ctx.reactions = self._rx
        cfg = L.EngineConfig()
        cfg.nl_cutoff = nl_cutoff
        cfg.cell_radius = cell_radius
        cfg.bin_capacity = bin_capacity
        cfg.max_replicas = max_replicas
        cfg.max_cells = max_cells
        cfg.max_particles = max_particles
        cfg.events_per_replica = events_per_replica
        cfg.block_size = block_size
        for k in range(3):
            cfg.brick[k] = int(brick[k])
        cfg.lds_states = lds_states
        h = ctypes.c_void_p()
        L.check(self._lib.rdy_engine_create(ctypes.byref(ctx), ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.replicas = []

    def close(self):
        if getattr(self, "_h", None):
            self._lib.rdy_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- replicas and particles ----
    def add_replica(self, box, pbc=(1, 1, 1), seed=0, max_particles=0):
        d = L.ReplicaDesc()
        for k in range(3):
            d.box[k] = float(box[k])
            d.pbc[k] = int(bool(pbc[k]))
        d.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        d.max_particles = int(max_particles)
        r = ctypes.c_uint32()
        L.check(self._lib.rdy_replica_add(self._h, ctypes.byref(d), ctypes.byref(r)))
        self.replicas.append(r.value)
        return r.value

    def footprint(self, box, max_particles):
        d = L.ReplicaDesc()
        for k in range(3):
            d.box[k] = float(box[k])
            d.pbc[k] = 1
        d.max_particles = int(max_particles)
        b = ctypes.c_uint64()
        L.check(self._lib.rdy_replica_footprint(self._h, ctypes.byref(d), ctypes.byref(b)))
        return b.value

    def remove_replica(self, r):
        L.check(self._lib.rdy_replica_remove(self._h, r))
        if r in self.replicas:
            self.replicas.remove(r)

    def set_particles(self, r, pos, types):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        types = np.ascontiguousarray(types, dtype=np.uint32)
        assert len(pos) == len(types)
        L.check(self._lib.rdy_set_particles(self._h, r, len(types), _ptr(pos, _dp), _ptr(types, _u32p)))

    def n_particles(self, r=0):
        n = ctypes.c_uint64()
        L.check(self._lib.rdy_get_particles(self._h, r, 0, None, None, None, ctypes.byref(n)))
        return n.value

    def get_particles(self, r=0):
        """(positions, types, handles) of live particles, sorted by handle."""
        n = self.n_particles(r)
        pos = np.zeros((n, 3))
        types = np.zeros(n, dtype=np.uint32)
        hs = np.zeros(n, dtype=np.uint32)
        m = ctypes.c_uint64()
        L.check(self._lib.rdy_get_particles(self._h, r, n, _ptr(pos, _dp), _ptr(types, _u32p), _ptr(hs, _u32p),
                                            ctypes.byref(m)))
        o = np.argsort(hs, kind="stable")
        return pos[o], types[o], hs[o]

    def get_forces(self, r=0):
        """(forces, handles) of the last force evaluation, sorted by handle."""
        n = self.n_particles(r)
        f = np.zeros((n, 3))
        hs = np.zeros(n, dtype=np.uint32)
        m = ctypes.c_uint64()
        L.check(self._lib.rdy_get_forces(self._h, r, n, _ptr(f, _dp), _ptr(hs, _u32p), ctypes.byref(m)))
        o = np.argsort(hs, kind="stable")
        return f[o], hs[o]

    def set_bonds(self, r, idx, k, length):
        idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1, 2)
        k = np.ascontiguousarray(np.broadcast_to(k, (len(idx),)), dtype=np.float64)
        length = np.ascontiguousarray(np.broadcast_to(length, (len(idx),)), dtype=np.float64)
        L.check(self._lib.rdy_set_bonds(self._h, r, len(idx), _ptr(idx, _u32p), _ptr(k, _dp), _ptr(length, _dp)))
        self._bond_terms = getattr(self, "_bond_terms", {})
        self._bond_terms[r] = len(idx)

    # rdy_break_pair (include/readdy_amd.h)
    BREAK_PAIR_DTYPE = np.dtype([("type1", np.uint32), ("type2", np.uint32), ("threshold_energy", np.float64),
                                 ("rate", np.float64)], align=True)

    def break_bonds(self, dt, pairs):
        """BreakBonds' device pass: pairs = [(type1, type2, threshold energy, rate)]; returns the broken
        edges as rows (replica, h1, h2), sorted.  The bonded terms stay as set: the caller sets them again
        without the broken edges (and the angles and dihedrals across them)."""
        p = np.zeros(len(pairs), dtype=self.BREAK_PAIR_DTYPE)
        for i, (a, b, thr, rate) in enumerate(pairs):
            p[i] = (a, b, thr, rate)
        cap = max(1, sum(getattr(self, "_bond_terms", {}).values()))  # an edge breaks once; edges <= terms
        out = np.zeros((cap, 3), dtype=np.uint32)
        n = ctypes.c_uint64()
        L.check(self._lib.rdy_break_bonds(self._h, float(dt), len(p), p.ctypes.data_as(ctypes.c_void_p), cap,
                                          _ptr(out, _u32p), ctypes.byref(n)))
        if n.value > cap:
            raise RuntimeError("more broken bonds than the output holds")
        return out[:n.value]

    def set_angles(self, r, idx, k, theta0):
        idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1, 3)
        k = np.ascontiguousarray(np.broadcast_to(k, (len(idx),)), dtype=np.float64)
        t0 = np.ascontiguousarray(np.broadcast_to(theta0, (len(idx),)), dtype=np.float64)
        L.check(self._lib.rdy_set_angles(self._h, r, len(idx), _ptr(idx, _u32p), _ptr(k, _dp), _ptr(t0, _dp)))

    def set_dihedrals(self, r, idx, k, multiplicity, phi0):
        idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1, 4)
        n = len(idx)
        k = np.ascontiguousarray(np.broadcast_to(k, (n,)), dtype=np.float64)
        m = np.ascontiguousarray(np.broadcast_to(multiplicity, (n,)), dtype=np.uint32)
        p0 = np.ascontiguousarray(np.broadcast_to(phi0, (n,)), dtype=np.float64)
        L.check(self._lib.rdy_set_dihedrals(self._h, r, n, _ptr(idx, _u32p), _ptr(k, _dp), _ptr(m, _u32p),
                                            _ptr(p0, _dp)))

    # ---- actions ----
    def update_neighbor_list(self):
        L.check(self._lib.rdy_update_neighbor_list(self._h))

    def calculate_forces(self):
        L.check(self._lib.rdy_calculate_forces(self._h))

    def energy(self, r=0):
        e = ctypes.c_double()
        L.check(self._lib.rdy_get_energy(self._h, r, ctypes.byref(e)))
        return e.value

    def record_virial(self, on=True):
        """Context::recordVirial: later force evaluations also form the pair virial."""
        L.check(self._lib.rdy_record_virial(self._h, 1 if on else 0))

    def virial(self, r=0):
        """3x3 virial of the last force evaluation: sum over unordered pairs of -x_ij (x) f_i."""
        v = np.zeros(9)
        L.check(self._lib.rdy_get_virial(self._h, r, _ptr(v, _dp)))
        return v.reshape(3, 3)

    def integrate(self, dt, noise=None):
        if noise is None:
            L.check(self._lib.rdy_integrate(self._h, dt, None, 0))
        else:
            noise = np.ascontiguousarray(noise, dtype=np.float64).reshape(-1, 3)
            L.check(self._lib.rdy_integrate(self._h, dt, _ptr(noise, _dp), len(noise)))

    def react(self, handler, dt):
        L.check(self._lib.rdy_react(self._h, handler, dt))

    def reaction_counts(self, r=0):
        c = np.zeros(max(self.n_reactions, 1), dtype=np.uint64)
        L.check(self._lib.rdy_get_reaction_counts(self._h, r, _ptr(c, _u64p)))
        return c[:self.n_reactions]

    def reaction_totals(self, r=0):
        c = np.zeros(max(self.n_reactions, 1), dtype=np.uint64)
        L.check(self._lib.rdy_get_reaction_totals(self._h, r, _ptr(c, _u64p)))
        return c[:self.n_reactions]

    # rdy_reaction_record (include/readdy_amd.h), a ReactionRecord (ReactionRecord.h:54-63)
    RECORD_DTYPE = np.dtype([("reaction", np.uint32), ("type", np.int32), ("educts", np.uint32, 2),
                             ("products", np.uint32, 2), ("types_from", np.uint32, 2), ("position", np.float64, 3)])

    def record_reactions(self, on=True):
        """Keep every reaction step's records (Context::recordReactionsWithPositions)."""
        L.check(self._lib.rdy_record_reactions(self._h, 1 if on else 0))

    def reaction_records(self, r=0):
        """The records of the replica's last reaction step, in event-key order (RECORD_DTYPE)."""
        n = ctypes.c_uint64()
        L.check(self._lib.rdy_get_reaction_records(self._h, r, 0, None, ctypes.byref(n)))
        out = np.zeros(max(1, n.value), dtype=self.RECORD_DTYPE)
        L.check(self._lib.rdy_get_reaction_records(self._h, r, len(out), out.ctypes.data_as(ctypes.c_void_p),
                                                   ctypes.byref(n)))
        return out[:n.value]

    def run(self, handler, dt, n_steps, graph=False):
        fn = self._lib.rdy_run_graph if graph else self._lib.rdy_run
        L.check(fn(self._h, handler, dt, int(n_steps)))

    # ---- spatial topology reactions (topology-particle enzymatic and fusion) ----
    def set_spatial_reactions(self, reactions):
        """reactions: dicts with type1, top_type1, type2, type_to1, top_type_to1, type_to2, rate, radius and
        optionally fusion ("T1(type1) + (type2) -> T2(type_to1) + (type_to2)", with fusion=1
        "T1(type1) + (type2) -> T2(type_to1--type_to2)"), or topology_topology=1 with top_type2 and top_type_to2
        ("T1(type1) + T3(type2) -> T2(type_to1) + T4(type_to2)")"""
        arr = (L.SpatialReaction * max(len(reactions), 1))()
        for i, x in enumerate(reactions):
            for f in ("type1", "top_type1", "type2", "type_to1", "top_type_to1", "type_to2", "rate", "radius"):
                setattr(arr[i], f, x[f])
            arr[i].fusion = int(x.get("fusion", 0))
            arr[i].topology_topology = int(x.get("topology_topology", 0))
            arr[i].top_type2 = int(x.get("top_type2", 0))
            arr[i].top_type_to2 = int(x.get("top_type_to2", 0))
            arr[i].self_fusion = int(x.get("self_fusion", 0))
        self._sr, self._n_sr = arr, len(reactions)
        L.check(self._lib.rdy_set_spatial_reactions(self._h, len(reactions), ctypes.cast(arr, ctypes.c_void_p)))

    def set_topologies(self, r, topology_types, topology_of):
        """topology_types[k]: type of topology k; topology_of[h]: topology of topology particle h"""
        tt = np.ascontiguousarray(topology_types, dtype=np.uint32)
        to = np.ascontiguousarray(topology_of, dtype=np.uint32)
        L.check(self._lib.rdy_set_topologies(self._h, r, len(tt), _ptr(tt, _u32p), _ptr(to, _u32p)))

    def topology_types(self, r=0):
        n = ctypes.c_uint32()
        L.check(self._lib.rdy_get_topology_types(self._h, r, 0, None, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.uint32)
        L.check(self._lib.rdy_get_topology_types(self._h, r, n.value, _ptr(out, _u32p), ctypes.byref(n)))
        return out[:n.value]

    def reserve_topology_particles(self, r, n):
        """handles for particles that join topologies by fusion; takes effect with the next set_particles"""
        L.check(self._lib.rdy_reserve_topology_particles(self._h, r, int(n)))

    def topology_fusions(self, r=0):
        """rows (topology, topology particle handle, joining particle's old handle, its new handle) of the last
        evaluation's fusions, in the order they were performed"""
        n = ctypes.c_uint32()
        L.check(self._lib.rdy_get_topology_fusions(self._h, r, 0, None, ctypes.byref(n)))
        out = np.zeros((max(n.value, 1), 4), dtype=np.uint32)
        L.check(self._lib.rdy_get_topology_fusions(self._h, r, n.value, _ptr(out, _u32p), ctypes.byref(n)))
        return out[:n.value]

    def topology_merges(self, r=0):
        """rows (surviving topology, absorbed topology, surviving side's particle, absorbed side's particle) of the
        last evaluation's topology-topology fusions, in key order"""
        n = ctypes.c_uint32()
        L.check(self._lib.rdy_get_topology_merges(self._h, r, 0, None, ctypes.byref(n)))
        out = np.zeros((max(n.value, 1), 4), dtype=np.uint32)
        L.check(self._lib.rdy_get_topology_merges(self._h, r, n.value, _ptr(out, _u32p), ctypes.byref(n)))
        return out[:n.value]

    def evaluate_topology_reactions(self, dt):
        L.check(self._lib.rdy_evaluate_topology_reactions(self._h, dt))

    def topology_reacted(self, r=0):
        """per topology, whether the last evaluate_topology_reactions changed it"""
        n = ctypes.c_uint32()
        L.check(self._lib.rdy_get_topology_reacted(self._h, r, 0, None, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.uint8)
        L.check(self._lib.rdy_get_topology_reacted(self._h, r, n.value, out.ctypes.data_as(ctypes.POINTER(
            ctypes.c_uint8)), ctypes.byref(n)))
        return out[:n.value].astype(bool)

    def set_particle_types(self, r, handles, types):
        """new species of topology particles (handles below the replica's topology count)"""
        h = np.ascontiguousarray(handles, dtype=np.uint32)
        t = np.ascontiguousarray(types, dtype=np.uint32)
        L.check(self._lib.rdy_set_particle_types(self._h, r, len(h), _ptr(h, _u32p), _ptr(t, _u32p)))

    def set_particle_positions(self, r, handles, pos):
        """moves topology particles (by handle); the engine re-bins its particles"""
        h = np.ascontiguousarray(handles, dtype=np.uint32)
        p = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        assert len(h) == len(p)
        L.check(self._lib.rdy_set_particle_positions(self._h, r, len(h), _ptr(h, _u32p), _ptr(p, _dp)))

    def append_topology_particles(self, r, types, pos):
        """new topology particles under the next handles of the topology reserve; returns their handles"""
        t = np.ascontiguousarray(types, dtype=np.uint32)
        p = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        assert len(t) == len(p)
        h = np.zeros(max(len(t), 1), dtype=np.uint32)
        L.check(self._lib.rdy_append_topology_particles(self._h, r, len(t), _ptr(t, _u32p), _ptr(p, _dp),
                                                        _ptr(h, _u32p)))
        return h[:len(t)]

    def topology_reaction_counts(self, r=0, totals=False):
        n = getattr(self, "_n_sr", 0)
        c = np.zeros(max(n, 1), dtype=np.uint64)
        L.check(self._lib.rdy_get_topology_reaction_counts(self._h, r, 1 if totals else 0, _ptr(c, _u64p)))
        return c[:n]

    def run_begin(self, handler, dt, n_steps):
        """Opens a run of n_steps taken a few steps at a time (no host synchronisation in between)."""
        L.check(self._lib.rdy_run_begin(self._h, handler, dt, int(n_steps)))

    def run_enqueue(self, k=1):
        L.check(self._lib.rdy_run_enqueue(self._h, int(k)))

    def run_end(self):
        L.check(self._lib.rdy_run_end(self._h))

    def set_steps(self, r, steps):
        L.check(self._lib.rdy_replica_set_steps(self._h, r, int(steps)))

    def set_handler(self, r, handler):
        """Reaction handler of replica r (L.HANDLER_*; -1: the handler passed to run/react)."""
        L.check(self._lib.rdy_replica_set_handler(self._h, r, int(handler)))

    def steps_done(self, r):
        s = ctypes.c_uint64()
        L.check(self._lib.rdy_replica_steps_done(self._h, r, ctypes.byref(s)))
        return s.value

    def synchronize(self):
        L.check(self._lib.rdy_synchronize(self._h))

    # ---- parity exports ----
    def neighbor_pairs(self, r, cutoff):
        n = ctypes.c_uint64()
        L.check(self._lib.rdy_neighbor_pairs(self._h, r, cutoff, 0, None, ctypes.byref(n)))
        cap = n.value
        out = np.zeros(2 * max(cap, 1), dtype=np.uint32)
        L.check(self._lib.rdy_neighbor_pairs(self._h, r, cutoff, cap, _ptr(out, _u32p), ctypes.byref(n)))
        return out[:2 * cap].reshape(-1, 2)

    def reaction_candidates(self, r=0):
        n = ctypes.c_uint64()
        L.check(self._lib.rdy_reaction_candidates(self._h, r, 0, None, ctypes.byref(n)))
        cap = n.value
        out = np.zeros(4 * max(cap, 1), dtype=np.uint32)
        L.check(self._lib.rdy_reaction_candidates(self._h, r, cap, _ptr(out, _u32p), ctypes.byref(n)))
        return out[:4 * cap].reshape(-1, 4)

    def count_types(self, n_replicas=None):
        n = n_replicas if n_replicas is not None else (max(self.replicas) + 1 if self.replicas else 1)
        c = np.zeros(n * self.n_types, dtype=np.uint64)
        L.check(self._lib.rdy_count_types(self._h, _ptr(c, _u64p), n))
        return c.reshape(n, self.n_types)

    def rdf_counts(self, replica, borders, from_types, to_types):
        """Radial-distribution pair counts per bin and the number of 'from' particles (the pair loop of
        RadialDistribution::evaluate, RadialDistribution.cpp:75-103, over the cell list on the device)"""
        b = np.ascontiguousarray(borders, dtype=np.float64)
        f = np.ascontiguousarray(from_types, dtype=np.uint32)
        t = np.ascontiguousarray(to_types, dtype=np.uint32)
        c = np.zeros(max(len(b) - 1, 0), dtype=np.uint64)
        nf = np.zeros(1, dtype=np.uint64)
        L.check(self._lib.rdy_rdf_counts(self._h, replica, _ptr(b, _dp), len(b), _ptr(f, _u32p), len(f),
                                         _ptr(t, _u32p), len(t), _ptr(c, _u64p), _ptr(nf, _u64p)))
        return c, int(nf[0])

    def histogram_along_axis(self, replica, borders, types, axis):
        """HistogramAlongAxis counts as the reference's CPU kernel forms them (CPUObservables.cpp:85-136)"""
        b = np.ascontiguousarray(borders, dtype=np.float64)
        t = np.ascontiguousarray(types, dtype=np.uint32)
        c = np.zeros(max(len(b) - 1, 0), dtype=np.uint64)
        L.check(self._lib.rdy_histogram_along_axis(self._h, replica, _ptr(b, _dp), len(b), _ptr(t, _u32p), len(t),
                                                   int(axis), _ptr(c, _u64p)))
        return c

    def histogram_along_axis_noalloc(self, replica, borders, types, axis):
        """The same counts from the entry point that allocates nothing (static per-GPU workspace)."""
        b = np.ascontiguousarray(borders, dtype=np.float64)
        t = np.ascontiguousarray(types, dtype=np.uint32)
        c = np.zeros(max(len(b) - 1, 0), dtype=np.uint64)
        L.check(self._lib.rdy_histogram_along_axis_noalloc(self._h, replica, _ptr(b, _dp), len(b), _ptr(t, _u32p),
                                                           len(t), int(axis), _ptr(c, _u64p)))
        return c

    def radial_distribution(self, replica, borders, from_types, to_types, particle_to_density):
        """(bin centers, distribution) as RadialDistribution::evaluate forms them
        (RadialDistribution.cpp:105-123, 136-144): counts / (4/3 pi (r_hi^3 - r_lo^3) n_from density)"""
        b = np.asarray(borders, dtype=np.float64)
        counts, nf = self.rdf_counts(replica, b, from_types, to_types)
        centers = (b[:-1] + b[1:]) / 2
        shell = 4. / 3. * np.pi * (np.power(b[1:], 3.) - np.power(b[:-1], 3.))
        return centers, counts.astype(np.float64) / (shell * nf * particle_to_density)

    def profile_reset(self):
        L.check(self._lib.rdy_profile_reset(self._h))

    def profile_get(self):
        us = ctypes.c_double()
        n = ctypes.c_uint64()
        L.check(self._lib.rdy_profile_get(self._h, ctypes.byref(us), ctypes.byref(n)))
        return us.value, n.value


def device_info():
    lib = L.load()
    name = ctypes.create_string_buffer(256)
    mem = ctypes.c_uint64()
    cus = ctypes.c_int32()
    L.check(lib.rdy_device_info(name, 256, ctypes.byref(mem), ctypes.byref(cus)))
    return {"name": name.value.decode(), "hbm_bytes": mem.value, "compute_units": cus.value}
