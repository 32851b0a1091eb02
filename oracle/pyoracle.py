"""ctypes wrapper around oracle/liboracle.so, the CPU FP64 restatement of ReaDDy's CPU kernel.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, always as the checker and never as the code being measured or shipped.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

POT_HARMONIC_REPULSION = 0
POT_WEAK_INTERACTION = 1
POT_LENNARD_JONES = 2
POT_SCREENED_ELECTROSTATICS = 3

RX_CONVERSION = 0
RX_FUSION = 1
RX_FISSION = 2
RX_ENZYMATIC = 3
RX_DECAY = 4

UA_KEYED = 0
GILLESPIE_LITERAL = 1
GILLESPIE_KEYED = 2

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_create.restype = ctypes.c_void_p
        L.or_create.argtypes = [ctypes.c_int, _dp, _ip, ctypes.c_double, _dp]
        L.or_destroy.argtypes = [ctypes.c_void_p]
        L.or_add_pair_potential.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.or_add_reaction.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double]
        L.or_set_seed.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        L.or_set_id_counter.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        L.or_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_set_cas_fill.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_phase_times.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int]
        L.or_phase_times.restype = None
        L.or_histogram_along_axis.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int,
                                              _u64p]
        L.or_rdf_counts.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _ip, ctypes.c_int, _ip, ctypes.c_int, _u64p,
                                    _u64p]
        L.or_add_particles.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _ip, _u64p]
        L.or_add_bond.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_double,
                                  ctypes.c_double]
        L.or_add_angle.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                   ctypes.c_double, ctypes.c_double]
        L.or_add_dihedral.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                      ctypes.c_size_t, ctypes.c_double, ctypes.c_int, ctypes.c_double]
        L.or_n_entries.restype = ctypes.c_size_t
        L.or_n_entries.argtypes = [ctypes.c_void_p]
        L.or_n_active.restype = ctypes.c_size_t
        L.or_n_active.argtypes = [ctypes.c_void_p]
        L.or_get_particles.restype = ctypes.c_size_t
        L.or_get_particles.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _dp, _ip, _u64p, _dp]
        L.or_nl_setup.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
        L.or_nl_fill.argtypes = [ctypes.c_void_p]
        L.or_nl_dims.argtypes = [ctypes.c_void_p, _ip]
        L.or_nl_pairs.restype = ctypes.c_size_t
        L.or_nl_pairs.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_size_t, _u64p]
        L.or_brute_pairs.restype = ctypes.c_size_t
        L.or_brute_pairs.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_size_t, _u64p]
        L.or_max_cutoff.restype = ctypes.c_double
        L.or_max_cutoff.argtypes = [ctypes.c_void_p]
        L.or_calculate_forces.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.or_euler_bd.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp, ctypes.c_size_t, ctypes.c_uint32]
        L.or_reaction_candidates.restype = ctypes.c_size_t
        L.or_reaction_candidates.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _u64p]
        L.or_reactions.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_uint32, _u64p]
        L.or_set_record_reactions.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_set_record_reactions.restype = None
        L.or_reaction_records.restype = ctypes.c_void_p
        L.or_reaction_records.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
        L.or_reaction_counts.restype = _u64p
        L.or_reaction_counts.argtypes = [ctypes.c_void_p, _ip]
        L.or_loop_step.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_uint32, _dp,
                                   ctypes.c_size_t, _dp]
        L.or_perform_events_abstract.argtypes = [ctypes.c_int, _i64p, _i64p, _dp, _dp, ctypes.c_int,
                                                 ctypes.c_uint64, ctypes.c_uint32, _u8p]
        L.or_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.or_u53.restype = ctypes.c_double
        L.or_u53.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.or_fix_position.argtypes = [_dp, _dp, _ip]
        L.or_shortest_difference.argtypes = [_dp, _dp, _dp, _ip, _dp]
        L.or_dist2.restype = ctypes.c_double
        L.or_dist2.argtypes = [_dp, _dp, _dp, _ip]
        L.or_pair_force_energy.restype = ctypes.c_double
        L.or_pair_force_energy.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.or_bond_force_energy.restype = ctypes.c_double
        L.or_bond_force_energy.argtypes = [ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _ip, _dp, _dp]
        L.or_angle_force_energy.restype = ctypes.c_double
        L.or_angle_force_energy.argtypes = [ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp, _ip, _dp, _dp,
                                            _dp]
        L.or_dihedral_force_energy.restype = ctypes.c_double
        L.or_dihedral_force_energy.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp,
                                               _dp, _ip, _dp, _dp, _dp, _dp]
        L.or_add_spatial_reaction.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 6 + [ctypes.c_double] * 2
        L.or_add_spatial_fusion.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 6 + [ctypes.c_double] * 2
        L.or_set_topology_reserve.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64]
        L.or_add_spatial_tt.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 8 + [ctypes.c_double] * 2
        L.or_add_spatial_tt_fusion.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 7 + [ctypes.c_double] * 2
        L.or_allow_self_fusion.argtypes = [ctypes.c_void_p]
        L.or_topology_merges.restype = _u64p
        L.or_topology_merges.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
        L.or_topology_fusions.restype = _u64p
        L.or_topology_fusions.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
        L.or_set_topologies.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _ip, ctypes.c_size_t, _ip]
        L.or_topology_reactions.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_uint32, _u64p]
        L.or_topology_types.restype = _ip
        L.or_topology_types.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
        L.or_topology_reaction_counts.restype = _u64p
        L.or_topology_reaction_counts.argtypes = [ctypes.c_void_p, _ip]
        L.or_reversible_configs.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _dp, ctypes.POINTER(ctypes.c_size_t)]
        L.or_fission_table.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.or_detailed_balance.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_uint32, _u64p]
        L.or_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().or_last_error().decode())


def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return list(o)


def u53(a, b):
    return lib().or_u53(a, b)


def shortest_difference(lhs, rhs, box, pbc):
    lhs, lp = _d(lhs)
    rhs, rp = _d(rhs)
    box, bp = _d(box)
    pbc, pp = _i(pbc)
    out = np.zeros(3)
    lib().or_shortest_difference(lp, rp, bp, pp, out.ctypes.data_as(_dp))
    return out


def fix_position(v, box, pbc):
    v = np.array(v, dtype=np.float64)
    box, bp = _d(box)
    pbc, pp = _i(pbc)
    lib().or_fix_position(v.ctypes.data_as(_dp), bp, pp)
    return v


def pair_force_energy(kind, params, x_ij):
    p = np.zeros(8)
    p[:len(params)] = params
    p, pp = _d(p)
    x, xp = _d(x_ij)
    f = np.zeros(3)
    e = lib().or_pair_force_energy(kind, pp, xp, f.ctypes.data_as(_dp))
    return f, e


def bond_force_energy(k, l0, xi, xj, box=(1e9,) * 3, pbc=(0, 0, 0)):
    xi, a = _d(xi)
    xj, b = _d(xj)
    box, bp = _d(box)
    pbc, pp = _i(pbc)
    fi, fj = np.zeros(3), np.zeros(3)
    e = lib().or_bond_force_energy(k, l0, a, b, bp, pp, fi.ctypes.data_as(_dp), fj.ctypes.data_as(_dp))
    return (fi, fj), e


def angle_force_energy(k, theta0, xi, xj, xk, box=(1e9,) * 3, pbc=(0, 0, 0)):
    xi, a = _d(xi)
    xj, b = _d(xj)
    xk, c = _d(xk)
    box, bp = _d(box)
    pbc, pp = _i(pbc)
    f = [np.zeros(3) for _ in range(3)]
    e = lib().or_angle_force_energy(k, theta0, a, b, c, bp, pp, *[x.ctypes.data_as(_dp) for x in f])
    return f, e


def dihedral_force_energy(k, mult, phi0, xi, xj, xk, xl, box=(1e9,) * 3, pbc=(0, 0, 0)):
    xi, a = _d(xi)
    xj, b = _d(xj)
    xk, c = _d(xk)
    xl, d = _d(xl)
    box, bp = _d(box)
    pbc, pp = _i(pbc)
    f = [np.zeros(3) for _ in range(4)]
    e = lib().or_dihedral_force_energy(k, mult, phi0, a, b, c, d, bp, pp, *[x.ctypes.data_as(_dp) for x in f])
    return f, e


def perform_events_abstract(p1, p2, rate, accept_p, mode, seed, trial):
    n = len(p1)
    p1 = np.ascontiguousarray(p1, dtype=np.int64)
    p2 = np.ascontiguousarray(p2, dtype=np.int64)
    rate, rp = _d(rate)
    acc, ap = _d(accept_p)
    out = np.zeros(n, dtype=np.uint8)
    _check(lib().or_perform_events_abstract(n, p1.ctypes.data_as(_i64p), p2.ctypes.data_as(_i64p), rp, ap, mode,
                                            seed, trial, out.ctypes.data_as(_u8p)))
    return out


class System:
    """One ReaDDy system (Context + CPU kernel state) held by the oracle."""

    def __init__(self, n_types, box, pbc=(1, 1, 1), kbt=1.0, diffusion=None):
        diffusion = np.ones(n_types) if diffusion is None else np.asarray(diffusion, dtype=np.float64)
        self._box, bp = _d(box)
        self._pbc, pp = _i(pbc)
        d, dp = _d(diffusion)
        self._h = lib().or_create(n_types, bp, pp, kbt, dp)
        if not self._h:
            raise RuntimeError(lib().or_last_error().decode())

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_destroy(self._h)
            self._h = None

    def add_pair_potential(self, kind, t1, t2, params):
        p = np.zeros(8)
        p[:len(params)] = params
        p, pp = _d(p)
        _check(lib().or_add_pair_potential(self._h, kind, t1, t2, pp))

    def add_reaction(self, kind, educts, products, rate, educt_distance=0.0, product_distance=0.0, w1=0.5, w2=0.5):
        e = list(educts) + [0] * (2 - len(educts))
        p = list(products) + [0] * (2 - len(products))
        _check(lib().or_add_reaction(self._h, kind, e[0], e[1], p[0], p[1], rate, educt_distance, product_distance,
                                     w1, w2))

    def set_seed(self, seed):
        lib().or_set_seed(self._h, seed)

    def set_id_counter(self, next_id):
        lib().or_set_id_counter(self._h, next_id)

    def rdf_counts(self, borders, from_types, to_types):
        """RadialDistribution counts (brute force over ordered pairs) and the number of 'from' particles"""
        b, bp = _d(np.asarray(borders, dtype=np.float64))
        f = np.ascontiguousarray(from_types, dtype=np.int32)
        t = np.ascontiguousarray(to_types, dtype=np.int32)
        counts = np.zeros(max(len(b) - 1, 0), dtype=np.uint64)
        nf = np.zeros(1, dtype=np.uint64)
        _check(lib().or_rdf_counts(self._h, bp, len(b), f.ctypes.data_as(_ip), len(f), t.ctypes.data_as(_ip), len(t),
                                   counts.ctypes.data_as(_u64p), nf.ctypes.data_as(_u64p)))
        return counts, int(nf[0])

    def histogram_along_axis(self, borders, types, axis):
        """CPUHistogramAlongAxis counts (the CPU kernel's bins: the last one stays empty)"""
        b, bp = _d(np.asarray(borders, dtype=np.float64))
        t = np.ascontiguousarray(types, dtype=np.int32)
        counts = np.zeros(max(len(b) - 1, 0), dtype=np.uint64)
        _check(lib().or_histogram_along_axis(self._h, bp, len(b), t.ctypes.data_as(_ip), len(t), int(axis),
                                             counts.ctypes.data_as(_u64p)))
        return counts

    def phase_times(self, reset=False):
        """seconds spent in loop_step's phases: integrator, neighborList, reactions, forces"""
        out = np.zeros(4)
        lib().or_phase_times(self._h, out.ctypes.data_as(_dp), int(bool(reset)))
        return dict(zip(("integrator", "neighborList", "reactions", "forces"), out.tolist()))

    def set_threads(self, n):
        """OpenMP threads for forces, integration and the keyed candidate search (default 1)"""
        lib().or_set_threads(self._h, int(n))

    def set_cas_fill(self, on=True):
        """Fill the cell list from all threads by atomic exchange, as the reference does (timing only:
        the order within a cell varies from run to run)"""
        lib().or_set_cas_fill(self._h, 1 if on else 0)

    def add_particles(self, pos, types, ids=None):
        pos, pp = _d(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
        types, tp = _i(types)
        if ids is not None:
            ids = np.ascontiguousarray(ids, dtype=np.uint64)
            ip = ids.ctypes.data_as(_u64p)
        else:
            ip = None
        _check(lib().or_add_particles(self._h, len(types), pp, tp, ip))

    def add_bond(self, i, j, k, l0):
        _check(lib().or_add_bond(self._h, i, j, k, l0))

    def add_angle(self, i, j, k_, k, theta0):
        _check(lib().or_add_angle(self._h, i, j, k_, k, theta0))

    def add_dihedral(self, i, j, k_, l, k, mult, phi0):
        _check(lib().or_add_dihedral(self._h, i, j, k_, l, k, mult, phi0))

    @property
    def n_active(self):
        return lib().or_n_active(self._h)

    def particles(self):
        n = lib().or_get_particles(self._h, 0, None, None, None, None)
        pos = np.zeros((n, 3))
        types = np.zeros(n, dtype=np.int32)
        ids = np.zeros(n, dtype=np.uint64)
        forces = np.zeros((n, 3))
        lib().or_get_particles(self._h, n, pos.ctypes.data_as(_dp), types.ctypes.data_as(_ip),
                               ids.ctypes.data_as(_u64p), forces.ctypes.data_as(_dp))
        return pos, types, ids, forces

    def max_cutoff(self):
        return lib().or_max_cutoff(self._h)

    def nl_setup(self, cutoff, radius=1):
        _check(lib().or_nl_setup(self._h, cutoff, radius))

    def nl_fill(self):
        _check(lib().or_nl_fill(self._h))

    def nl_dims(self):
        d = (ctypes.c_int * 3)()
        lib().or_nl_dims(self._h, d)
        return tuple(d)

    def _pairs(self, fn, cutoff):
        n = fn(self._h, cutoff, 0, None)
        out = np.zeros(2 * max(n, 1), dtype=np.uint64)
        fn(self._h, cutoff, n, out.ctypes.data_as(_u64p))
        return out[:2 * n].reshape(-1, 2)

    def nl_pairs(self, cutoff):
        return self._pairs(lib().or_nl_pairs, cutoff)

    def brute_pairs(self, cutoff):
        return self._pairs(lib().or_brute_pairs, cutoff)

    def calculate_forces(self, virial=False):
        e = ctypes.c_double()
        v = np.zeros(9)
        _check(lib().or_calculate_forces(self._h, ctypes.byref(e), v.ctypes.data_as(_dp)))
        return (e.value, v.reshape(3, 3)) if virial else e.value

    def euler_bd(self, dt, noise_by_id=None, step=0):
        if noise_by_id is not None:
            nb, npp = _d(noise_by_id)
            _check(lib().or_euler_bd(self._h, dt, npp, nb.size, step))
        else:
            _check(lib().or_euler_bd(self._h, dt, None, 0, step))

    def reaction_candidates(self):
        n = lib().or_reaction_candidates(self._h, 0, None)
        out = np.zeros(4 * max(n, 1), dtype=np.uint64)
        lib().or_reaction_candidates(self._h, n, out.ctypes.data_as(_u64p))
        return out[:4 * n].reshape(-1, 4)

    def reactions(self, mode, dt, step=0):
        n = ctypes.c_uint64()
        _check(lib().or_reactions(self._h, mode, dt, step, ctypes.byref(n)))
        return n.value

    REV_FIELDS = ("kind", "forward", "backward", "lhs0", "lhs1", "rhs0", "rhs1", "reaction_radius", "lhs_radius",
                  "rhs_radius", "lhs_volume", "rhs_volume", "eff_lhs_interaction", "eff_rhs_interaction",
                  "eff_lhs_reaction", "eff_rhs_reaction", "acceptance_prefactor", "total_volume",
                  "equilibrium_constant", "macro_forward", "macro_backward")

    def reversible_configs(self):
        """DetailedBalance's reversible pairs and their configuration, one dict per pair"""
        n = ctypes.c_size_t()
        _check(lib().or_reversible_configs(self._h, 0, None, ctypes.byref(n)))
        rows = np.zeros((max(n.value, 1), len(self.REV_FIELDS)))
        _check(lib().or_reversible_configs(self._h, n.value, rows.ctypes.data_as(_dp), ctypes.byref(n)))
        out = []
        for r in rows[:n.value]:
            d = dict(zip(self.REV_FIELDS, r.tolist()))
            for k in ("kind", "forward", "backward", "lhs0", "lhs1", "rhs0", "rhs1"):
                d[k] = int(d[k])
            out.append(d)
        return out

    def fission_table(self, k):
        radii, cdf = np.zeros(1001), np.zeros(1001)
        rc = lib().or_fission_table(self._h, k, radii.ctypes.data_as(_dp), cdf.ctypes.data_as(_dp))
        if rc < 0:
            raise RuntimeError(lib().or_last_error().decode())
        return radii[:rc], cdf[:rc]

    def detailed_balance(self, dt, step=0):
        """one DetailedBalance handler call; returns the number of performed events"""
        n = ctypes.c_uint64()
        _check(lib().or_detailed_balance(self._h, dt, step, ctypes.byref(n)))
        return n.value

    # or_reaction_record (readdy_oracle.h)
    RECORD_DTYPE = np.dtype([("reaction", np.uint32), ("type", np.int32), ("educts", np.uint64, 2),
                             ("products", np.uint64, 2), ("types_from", np.uint32, 2), ("position", np.float64, 3)])

    def set_record_reactions(self, on=True):
        lib().or_set_record_reactions(self._h, 1 if on else 0)

    def reaction_records(self):
        """Records of the last reactions() call, in the order the reactions were performed."""
        n = ctypes.c_size_t()
        p = lib().or_reaction_records(self._h, ctypes.byref(n))
        if n.value == 0:
            return np.zeros(0, dtype=self.RECORD_DTYPE)
        raw = ctypes.string_at(p, n.value * self.RECORD_DTYPE.itemsize)
        return np.frombuffer(raw, dtype=self.RECORD_DTYPE).copy()

    def reaction_counts(self):
        n = ctypes.c_int()
        p = lib().or_reaction_counts(self._h, ctypes.byref(n))
        return np.array([p[i] for i in range(n.value)], dtype=np.uint64)

    # ---- spatial topology reactions of the topology-particle kinds ----
    def add_spatial_reaction(self, type1, top_type1, type2, type_to1, top_type_to1, type_to2, rate, radius,
                             fusion=False):
        """T1(type1) + (type2) -> T2(type_to1) + (type_to2), or with fusion T1(type1) + (type2) ->
        T2(type_to1--type_to2); topology types as integers"""
        f = lib().or_add_spatial_fusion if fusion else lib().or_add_spatial_reaction
        _check(f(self._h, type1, top_type1, type2, type_to1, top_type_to1, type_to2, rate, radius))

    def add_spatial_tt(self, type1, top_type1, type2, top_type2, type_to1, top_type_to1, type_to2, top_type_to2, rate,
                       radius):
        """T1(type1) + T2(type2) -> T3(type_to1) + T4(type_to2), both sides in topologies"""
        _check(lib().or_add_spatial_tt(self._h, type1, top_type1, type2, top_type2, type_to1, top_type_to1, type_to2,
                                       top_type_to2, rate, radius))

    def add_spatial_tt_fusion(self, type1, top_type1, type2, top_type2, type_to1, type_to2, top_type_to, rate, radius):
        """T1(type1) + T2(type2) -> T3(type_to1--type_to2): the second topology merges into the first"""
        _check(lib().or_add_spatial_tt_fusion(self._h, type1, top_type1, type2, top_type2, type_to1, type_to2,
                                              top_type_to, rate, radius))

    def allow_self_fusion(self):
        """the last topology-topology fusion also joins two unbonded particles of one topology"""
        _check(lib().or_allow_self_fusion(self._h))

    def topology_merges(self):
        """rows (surviving topology, absorbed topology, surviving side's id, absorbed side's id)"""
        n = ctypes.c_size_t()
        p = lib().or_topology_merges(self._h, ctypes.byref(n))
        return np.array([p[i] for i in range(4 * n.value)], dtype=np.uint64).reshape(n.value, 4)

    def set_topology_reserve(self, next_id, end_id):
        """ids next_id .. end_id - 1 for particles that join topologies by fusion"""
        _check(lib().or_set_topology_reserve(self._h, next_id, end_id))

    def topology_fusions(self):
        """rows (topology, topology particle id, old id, new id) of the last call's fusions"""
        n = ctypes.c_size_t()
        p = lib().or_topology_fusions(self._h, ctypes.byref(n))
        return np.array([p[i] for i in range(4 * n.value)], dtype=np.uint64).reshape(n.value, 4)

    def set_topologies(self, topology_of_id, top_types):
        """topology_of_id[id] = topology index of the particle with that id (-1: none)"""
        a, ap = _i(np.asarray(topology_of_id, dtype=np.int32))
        b, bp = _i(np.asarray(top_types, dtype=np.int32))
        _check(lib().or_set_topologies(self._h, a.size, ap, b.size, bp))

    def topology_reactions(self, dt, step=0):
        n = ctypes.c_uint64()
        _check(lib().or_topology_reactions(self._h, dt, step, ctypes.byref(n)))
        return n.value

    def topology_types(self):
        n = ctypes.c_size_t()
        p = lib().or_topology_types(self._h, ctypes.byref(n))
        return np.array([p[i] for i in range(n.value)], dtype=np.int64)

    def topology_reaction_counts(self):
        n = ctypes.c_int()
        p = lib().or_topology_reaction_counts(self._h, ctypes.byref(n))
        return np.array([p[i] for i in range(n.value)], dtype=np.uint64)

    def loop_step(self, dt, mode, step, noise_by_id=None):
        e = ctypes.c_double()
        if noise_by_id is not None:
            nb, npp = _d(noise_by_id)
            _check(lib().or_loop_step(self._h, dt, mode, step, npp, nb.size, ctypes.byref(e)))
        else:
            _check(lib().or_loop_step(self._h, dt, mode, step, None, 0, ctypes.byref(e)))
        return e.value
