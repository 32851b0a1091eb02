"""Loads libreaddy_amd.so (the HIP engine) and declares its C ABI (include/readdy_amd.h).

There is no CPU fallback: if the library is missing or cannot be loaded the import fails
with an error that says how to build it.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libreaddy_amd.so")

RDY_OK = 0
POT_HARMONIC_REPULSION = 0
POT_WEAK_INTERACTION = 1
POT_LENNARD_JONES = 2
POT_SCREENED_ELECTROSTATICS = 3
RX_CONVERSION = 0
RX_FUSION = 1
RX_FISSION = 2
RX_ENZYMATIC = 3
RX_DECAY = 4
HANDLER_UA = 0
HANDLER_GILLESPIE = 1
HANDLER_NONE = 2
HANDLER_DETAILED_BALANCE = 3


class PairPotential(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("type1", ctypes.c_uint32), ("type2", ctypes.c_uint32),
                ("params", ctypes.c_double * 8)]


class SpatialReaction(ctypes.Structure):
    _fields_ = [("type1", ctypes.c_uint32), ("top_type1", ctypes.c_uint32), ("type2", ctypes.c_uint32),
                ("type_to1", ctypes.c_uint32), ("top_type_to1", ctypes.c_uint32), ("type_to2", ctypes.c_uint32),
                ("rate", ctypes.c_double), ("radius", ctypes.c_double), ("fusion", ctypes.c_uint32),
                ("topology_topology", ctypes.c_uint32), ("top_type2", ctypes.c_uint32),
                ("top_type_to2", ctypes.c_uint32), ("self_fusion", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class Reaction(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("educts", ctypes.c_uint32 * 2), ("products", ctypes.c_uint32 * 2),
                ("rate", ctypes.c_double), ("educt_distance", ctypes.c_double),
                ("product_distance", ctypes.c_double), ("weight1", ctypes.c_double), ("weight2", ctypes.c_double)]


class ContextDesc(ctypes.Structure):
    _fields_ = [("n_types", ctypes.c_uint32), ("kbt", ctypes.c_double),
                ("diffusion", ctypes.POINTER(ctypes.c_double)), ("topology_flavor", ctypes.POINTER(ctypes.c_uint8)),
                ("n_potentials", ctypes.c_uint32), ("potentials", ctypes.POINTER(PairPotential)),
                ("n_reactions", ctypes.c_uint32), ("reactions", ctypes.POINTER(Reaction))]


class EngineConfig(ctypes.Structure):
    _fields_ = [("nl_cutoff", ctypes.c_double), ("cell_radius", ctypes.c_int32), ("bin_capacity", ctypes.c_uint32),
                ("max_replicas", ctypes.c_uint32), ("max_cells", ctypes.c_uint64),
                ("max_particles", ctypes.c_uint64), ("events_per_replica", ctypes.c_uint32),
                ("block_size", ctypes.c_uint32), ("brick", ctypes.c_uint32 * 3), ("lds_states", ctypes.c_uint32)]


class ReplicaDesc(ctypes.Structure):
    _fields_ = [("box", ctypes.c_double * 3), ("pbc", ctypes.c_int32 * 3), ("seed", ctypes.c_uint64),
                ("max_particles", ctypes.c_uint64)]


class ReversibleReaction(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("forward", ctypes.c_uint32), ("backward", ctypes.c_uint32),
                ("n_lhs", ctypes.c_uint32), ("n_rhs", ctypes.c_uint32), ("lhs", ctypes.c_uint32 * 2),
                ("rhs", ctypes.c_uint32 * 2), ("reaction_radius", ctypes.c_double),
                ("lhs_interaction_radius", ctypes.c_double), ("rhs_interaction_radius", ctypes.c_double),
                ("lhs_interaction_volume", ctypes.c_double), ("rhs_interaction_volume", ctypes.c_double),
                ("eff_lhs_interaction_volume", ctypes.c_double), ("eff_rhs_interaction_volume", ctypes.c_double),
                ("eff_lhs_reaction_volume", ctypes.c_double), ("eff_rhs_reaction_volume", ctypes.c_double),
                ("acceptance_prefactor", ctypes.c_double), ("total_volume", ctypes.c_double),
                ("kbt", ctypes.c_double), ("equilibrium_constant", ctypes.c_double),
                ("macro_forward_rate", ctypes.c_double), ("macro_backward_rate", ctypes.c_double),
                ("micro_forward_rate", ctypes.c_double), ("micro_backward_rate", ctypes.c_double)]


# every symbol the header declares, with (restype, argtypes)
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_int = ctypes.c_int
_d = ctypes.c_double

SIGNATURES = {
    "rdy_engine_create": (_int, [ctypes.POINTER(ContextDesc), ctypes.POINTER(EngineConfig), ctypes.POINTER(_vp)]),
    "rdy_engine_destroy": (_int, [_vp]),
    "rdy_last_error": (ctypes.c_char_p, []),
    "rdy_device_info": (_int, [ctypes.c_char_p, ctypes.c_size_t, _u64p, _i32p]),
    "rdy_replica_add": (_int, [_vp, ctypes.POINTER(ReplicaDesc), _u32p]),
    "rdy_replica_remove": (_int, [_vp, _u32]),
    "rdy_replica_footprint": (_int, [_vp, ctypes.POINTER(ReplicaDesc), _u64p]),
    "rdy_set_particles": (_int, [_vp, _u32, _u64, _dp, _u32p]),
    "rdy_particle_count": (_u64, [_vp, _u32]),
    "rdy_get_particles": (_int, [_vp, _u32, _u64, _dp, _u32p, _u32p, _u64p]),
    "rdy_get_forces": (_int, [_vp, _u32, _u64, _dp, _u32p, _u64p]),
    "rdy_set_bonds": (_int, [_vp, _u32, _u64, _u32p, _dp, _dp]),
    "rdy_set_angles": (_int, [_vp, _u32, _u64, _u32p, _dp, _dp]),
    "rdy_set_dihedrals": (_int, [_vp, _u32, _u64, _u32p, _dp, _u32p, _dp]),
    "rdy_update_neighbor_list": (_int, [_vp]),
    "rdy_calculate_forces": (_int, [_vp]),
    "rdy_get_energy": (_int, [_vp, _u32, _dp]),
    "rdy_record_virial": (_int, [_vp, _int]),
    "rdy_get_virial": (_int, [_vp, _u32, _dp]),
    "rdy_integrate": (_int, [_vp, _d, _dp, _u64]),
    "rdy_react": (_int, [_vp, _int, _d]),
    "rdy_get_topology_reacted": (_int, [_vp, _u32, _u32, ctypes.POINTER(ctypes.c_uint8), _u32p]),
    "rdy_set_particle_types": (_int, [_vp, _u32, _u64, _u32p, _u32p]),
    "rdy_reversible_reactions": (_int, [ctypes.POINTER(ContextDesc), _d, _u64, ctypes.POINTER(ReversibleReaction),
                                        _u64p]),
    "rdy_get_reaction_counts": (_int, [_vp, _u32, _u64p]),
    "rdy_get_reaction_totals": (_int, [_vp, _u32, _u64p]),
    "rdy_record_reactions": (_int, [_vp, _int]),
    "rdy_break_bonds": (_int, [_vp, ctypes.c_double, _u32, _vp, _u64, _u32p, _u64p]),
    "rdy_get_reaction_records": (_int, [_vp, _u32, _u64, _vp, _u64p]),
    "rdy_run": (_int, [_vp, _int, _d, _u64]),
    "rdy_run_graph": (_int, [_vp, _int, _d, _u64]),
    "rdy_set_spatial_reactions": (_int, [_vp, _u32, _vp]),
    "rdy_set_topologies": (_int, [_vp, _u32, _u32, _u32p, _u32p]),
    "rdy_reserve_topology_particles": (_int, [_vp, _u32, _u32]),
    "rdy_get_topology_fusions": (_int, [_vp, _u32, _u32, _u32p, _u32p]),
    "rdy_get_topology_merges": (_int, [_vp, _u32, _u32, _u32p, _u32p]),
    "rdy_set_particle_positions": (_int, [_vp, _u32, _u64, _u32p, _dp]),
    "rdy_append_topology_particles": (_int, [_vp, _u32, _u64, _u32p, _dp, _u32p]),
    "rdy_get_topology_types": (_int, [_vp, _u32, _u32, _u32p, _u32p]),
    "rdy_evaluate_topology_reactions": (_int, [_vp, _d]),
    "rdy_get_topology_reaction_counts": (_int, [_vp, _u32, _int, _u64p]),
    "rdy_run_begin": (_int, [_vp, _int, _d, _u64]),
    "rdy_run_enqueue": (_int, [_vp, _u64]),
    "rdy_run_end": (_int, [_vp]),
    "rdy_replica_set_steps": (_int, [_vp, _u32, _u64]),
    "rdy_replica_steps_done": (_int, [_vp, _u32, _u64p]),
    "rdy_replica_set_handler": (_int, [_vp, _u32, _int]),
    "rdy_synchronize": (_int, [_vp]),
    "rdy_neighbor_pairs": (_int, [_vp, _u32, _d, _u64, _u32p, _u64p]),
    "rdy_reaction_candidates": (_int, [_vp, _u32, _u64, _u32p, _u64p]),
    "rdy_count_types": (_int, [_vp, _u64p, _u32]),
    "rdy_rdf_counts": (_int, [_vp, _u32, _dp, _u32, _u32p, _u32, _u32p, _u32, _u64p, _u64p]),
    "rdy_histogram_along_axis": (_int, [_vp, _u32, _dp, _u32, _u32p, _u32, _u32, _u64p]),
    "rdy_histogram_along_axis_noalloc": (_int, [_vp, _u32, _dp, _u32, _u32p, _u32, _u32, _u64p]),
    "rdy_profile_reset": (_int, [_vp]),
    "rdy_profile_get": (_int, [_vp, _dp, _u64p]),
}

_lib = None


class EngineError(RuntimeError):
    pass


def load():
    """Returns the loaded library; raises with build instructions if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP engine first (python -c 'import __graft_entry__ as g; g.build()' "
            f"or make -C readdy_amd/csrc). There is no CPU fallback.")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != RDY_OK:
        raise EngineError(f"libreaddy_amd error {rc}: {load().rdy_last_error().decode()}")
    return rc
