"""The user-facing Python API of ReaDDy, for the part this kernel runs: a ReactionDiffusionSystem with its
species, pair potentials and reactions, and the Simulation it creates, driving the MI355X engine.

Mirrors readdy.ReactionDiffusionSystem (wrappers/python/src/python/readdy/api/reaction_diffusion_system.py:
62-400) and readdy.Simulation (simulation.py:54-620) with their registries (registry/potentials.py:78-160,
registry/reactions.py:48-230, registry/observables.py:58-240):

    system = ReactionDiffusionSystem([20., 20., 20.], unit_system=None)
    system.add_species("A", 1.)
    system.potentials.add_harmonic_repulsion("A", "A", 10., 1.)
    system.reactions.add("fus: A +(1) A -> A", 0.1)
    sim = system.simulation(kernel="MI355X", output_file="out.h5", reaction_handler="UncontrolledApproximation")
    sim.add_particles("A", positions)
    sim.observe.number_of_particles(10, ["A"])
    sim.record_trajectory(100)
    sim.run(1000, 0.01)

Quantities are plain numbers.  With unit_system=None they are in units of 1 and kBT is 1 unless set; with
the default unit system (nm, ns, kJ/mol, K; the reference converts pint quantities into these) numbers are
read in those units and kBT = k_B T with T = 293 K unless given.  External potentials, compartments,
topologies, checkpoints and custom loops are not offered here (the C++ host core carries the topologies).
"""
import math

import numpy as np

from . import _lib as L
from .engine import Engine
from .observe import Recorder

BOLTZMANN_KJ_PER_MOL_K = 0.0083144626181532  # the default unit system's k_B, kJ/(mol K)

HANDLERS = {"UncontrolledApproximation": L.HANDLER_UA, "Gillespie": L.HANDLER_GILLESPIE,
            "DetailedBalance": L.HANDLER_DETAILED_BALANCE}


class ReactionDiffusionSystem:
    def __init__(self, box_size, temperature=None, periodic_boundary_conditions=None, unit_system="default"):
        self._units = unit_system is not None and unit_system != {} and unit_system is not False
        if temperature is not None and not self._units:
            raise ValueError("Setting the temperature without a unit system is not supported. "
                             "If working without units, set kbt instead.")
        self._species = []  # (name, diffusion constant, topology flavour)
        self._kbt = 1.0
        if self._units:
            self.temperature = 293.0 if temperature is None else temperature
        self.box_size = box_size
        self.periodic_boundary_conditions = [True, True, True] if periodic_boundary_conditions is None else \
            periodic_boundary_conditions
        self.potentials = _PotentialRegistry(self)
        self.reactions = _ReactionRegistry(self)

    # ---- units, box ----
    @property
    def kbt(self):
        return self._kbt

    @kbt.setter
    def kbt(self, value):
        if self._units:
            raise ValueError("Setting kbt directly is only supported without a unit system; set the temperature")
        self._kbt = float(value)

    @property
    def temperature(self):
        if not self._units:
            raise ValueError("No temperature without a unit system")
        return self._kbt / BOLTZMANN_KJ_PER_MOL_K

    @temperature.setter
    def temperature(self, value):
        if not self._units:
            raise ValueError("Setting the temperature without a unit system is not supported.")
        self._kbt = BOLTZMANN_KJ_PER_MOL_K * float(value)

    @property
    def box_size(self):
        return list(self._box)

    @box_size.setter
    def box_size(self, value):
        v = [float(x) for x in value]
        if len(v) != 3 or any(not (x > 0) for x in v):
            raise ValueError("the box size must be three positive numbers")
        self._box = v

    @property
    def box_volume(self):
        return float(np.prod(self._box))

    @property
    def periodic_boundary_conditions(self):
        return list(self._pbc)

    @periodic_boundary_conditions.setter
    def periodic_boundary_conditions(self, value):
        self._pbc = [bool(x) for x in value]
        if len(self._pbc) != 3:
            raise ValueError("three periodicity flags")

    # ---- species ----
    def add_species(self, name, diffusion_constant=1.):
        self._add_species(name, diffusion_constant, False)

    def add_topology_species(self, name, diffusion_constant=1.):
        self._add_species(name, diffusion_constant, True)

    def _add_species(self, name, d, topology):
        _validate_type_name(name)
        if name in self.registered_species():
            raise ValueError(f"A particle type with name {name} already exists.")
        if not (float(d) >= 0):
            raise ValueError("the diffusion constant must not be negative")
        self._species.append((name, float(d), topology))

    def registered_species(self):
        return [s[0] for s in self._species]

    def _type_id(self, name):
        for i, s in enumerate(self._species):
            if s[0] == name:
                return i
        raise ValueError(f"Could not find type \"{name}\"")

    @property
    def diffusion_constants(self):
        return {s[0]: s[1] for s in self._species}

    def calculate_max_cutoff(self):
        """Context::calculateMaxCutoff (Context.cpp:126-144): pair potentials and order-2 educt distances"""
        c = 0.0
        for kind, _, _, p in self.potentials._pots:
            c = max(c, _cutoff(kind, p))
        for r in self.reactions._rx:
            if r["kind"] in (L.RX_FUSION, L.RX_ENZYMATIC):
                c = max(c, r["educt_distance"])
        return c

    def simulation(self, kernel="MI355X", output_file="", integrator="EulerBDIntegrator", reaction_handler="Gillespie",
                   evaluate_topology_reactions=True, evaluate_forces=True, evaluate_observables=True, skin=0):
        if kernel != "MI355X":
            raise ValueError(f"this package runs the MI355X kernel, not {kernel!r}")
        if integrator != "EulerBDIntegrator":
            raise ValueError(f"unknown integrator {integrator!r}")
        if skin:
            raise ValueError("the MI355X kernel rebuilds its cell list every step: no skin")
        return Simulation(self, output_file, reaction_handler, evaluate_forces, evaluate_observables)


def _validate_type_name(name):
    """validateTypeName (readdy/main/model/Utils.cpp:171-190)"""
    for cs in ("[", "]", "(", ")", "->", "--", ":", "+"):
        if cs in name:
            raise ValueError(f"Encountered invalid character sequence \"{cs}\" in type name \"{name}\"")
    if not name or name[0] == " ":
        raise ValueError(f"Type name \"{name}\" is empty or contained leading/trailing whitespaces.")


def _cutoff(kind, p):
    """the parameter that holds a pair potential's cutoff in the engine's layout for its kind"""
    return p[{L.POT_HARMONIC_REPULSION: 1, L.POT_WEAK_INTERACTION: 3, L.POT_LENNARD_JONES: 2,
              L.POT_SCREENED_ELECTROSTATICS: 5}[kind]]


class _PotentialRegistry:  # registry/potentials.py:78-160; the engine's parameter layout per kind
    def __init__(self, system):
        self._s = system
        self._pots = []

    def _add(self, kind, t1, t2, params):
        self._pots.append((kind, self._s._type_id(t1), self._s._type_id(t2), [float(x) for x in params]))

    def add_harmonic_repulsion(self, particle_type1, particle_type2, force_constant, interaction_distance):
        self._add(L.POT_HARMONIC_REPULSION, particle_type1, particle_type2, [force_constant, interaction_distance])

    def add_weak_interaction_piecewise_harmonic(self, particle_type1, particle_type2, force_constant, desired_distance,
                                                depth, cutoff):
        self._add(L.POT_WEAK_INTERACTION, particle_type1, particle_type2, [force_constant, desired_distance, depth,
                                                                           cutoff])

    def add_lennard_jones(self, particle_type1, particle_type2, epsilon, sigma, cutoff=None, m=12, n=6, shift=True):
        if not isinstance(shift, bool):
            raise AssertionError("shift can only be bool")
        cutoff = 2.5 * sigma if cutoff is None else cutoff
        self._add(L.POT_LENNARD_JONES, particle_type1, particle_type2, [m, n, cutoff, 1.0 if shift else 0.0, epsilon,
                                                                        sigma])

    def add_screened_electrostatics(self, particle_type1, particle_type2, electrostatic_strength,
                                    inverse_screening_depth, repulsion_strength, repulsion_distance, exponent, cutoff):
        self._add(L.POT_SCREENED_ELECTROSTATICS, particle_type1, particle_type2,
                  [electrostatic_strength, inverse_screening_depth, repulsion_strength, repulsion_distance, exponent,
                   cutoff])

    def _external(self, *args, **kwargs):
        raise NotImplementedError("external (order-1) potentials are not available on the MI355X kernel")

    add_box = add_sphere = add_capsule = add_cylinder = add_spherical_barrier = add_harmonic_geometry = _external
    add_custom_external = add_custom_pair = _external


class _ReactionRegistry:  # registry/reactions.py:48-230; ReactionRegistry.cpp:154-305 for descriptors
    def __init__(self, system):
        self._s = system
        self._rx = []

    def _append(self, r):
        if any(x["name"] == r["name"] for x in self._rx):
            raise ValueError(f"A reaction with the name {r['name']} exists already")
        if not (r["rate"] >= 0):
            raise ValueError("the rate must not be negative")
        self._rx.append(r)

    def add_decay(self, name, particle_type, rate):
        self._append(dict(name=name, kind=L.RX_DECAY, educts=[self._s._type_id(particle_type)], products=[],
                          rate=float(rate)))

    def add_conversion(self, name, type_from, type_to, rate):
        self._append(dict(name=name, kind=L.RX_CONVERSION, educts=[self._s._type_id(type_from)],
                          products=[self._s._type_id(type_to)], rate=float(rate)))

    def add_fusion(self, name, type_from1, type_from2, type_to, rate, educt_distance, weight1=0.5, weight2=0.5):
        self._append(dict(name=name, kind=L.RX_FUSION, educts=[self._s._type_id(type_from1),
                                                               self._s._type_id(type_from2)],
                          products=[self._s._type_id(type_to)], rate=float(rate), educt_distance=float(educt_distance),
                          weight1=float(weight1), weight2=float(weight2)))

    def add_fission(self, name, type_from, type_to1, type_to2, rate, product_distance, weight1=0.5, weight2=0.5):
        self._append(dict(name=name, kind=L.RX_FISSION, educts=[self._s._type_id(type_from)],
                          products=[self._s._type_id(type_to1), self._s._type_id(type_to2)], rate=float(rate),
                          product_distance=float(product_distance), weight1=float(weight1), weight2=float(weight2)))

    def add_enzymatic(self, name, type_catalyst, type_from, type_to, rate, educt_distance):
        c = self._s._type_id(type_catalyst)
        self._append(dict(name=name, kind=L.RX_ENZYMATIC, educts=[self._s._type_id(type_from), c],
                          products=[self._s._type_id(type_to), c], rate=float(rate),
                          educt_distance=float(educt_distance)))

    def add(self, descriptor, rate):
        """"name: A ->" decay, "name: A -> B" conversion, "name: A +(r) B -> C [w1, w2]" fusion,
        "name: A -> B +(d) C [w1, w2]" fission, "name: A +(r) C -> B + C" enzymatic (C the catalyst)"""
        if "->" not in descriptor:
            raise ValueError("the descriptor must contain an arrow (\"->\") to indicate lhs and rhs.")
        if descriptor.count("->") > 1:
            raise ValueError("the descriptor must not contain more than one arrow (\"->\").")
        lhs, rhs = (x.strip() for x in descriptor.split("->"))
        if ":" not in lhs:
            raise ValueError("The descriptor did not contain a colon ':' to specify the end of the name.")
        name, lhs = (x.strip() for x in lhs.split(":", 1))
        w1 = w2 = 0.5
        if "[" in rhs:
            lb, rb = rhs.find("["), rhs.rfind("]")
            if rb < lb:
                raise ValueError(f"The term \"{rhs}\" did not contain a pair of brackets '[*]'")
            inner = rhs[lb + 1:rb]
            if "," not in inner:
                raise ValueError(f"The term \"{rhs}\" did not contain a comma ','")
            w1, w2 = (_number(x, "weight") for x in inner.split(",", 1))
            rhs = rhs[:lb].strip()
        two = "+" in lhs
        e1, e2, radius = _split_plus(lhs, True) if two else (lhs, None, None)
        if not rhs:
            if two:
                raise ValueError(f"Left hand side (\"{lhs}\") did not contain one educt")
            return self.add_decay(name, e1, rate)
        if "+" not in rhs:
            if two:
                return self.add_fusion(name, e1, e2, rhs, rate, _number(radius, "educt distance"), w1, w2)
            return self.add_conversion(name, e1, rhs, rate)
        if not two:
            p1, p2, dist = _split_plus(rhs, True)
            return self.add_fission(name, e1, p1, p2, rate, _number(dist, "product distance"), w1, w2)
        p1, p2, _ = _split_plus(rhs, False)
        if e2 != p2:
            raise ValueError(f"In enzymatic reaction, educt2 (\"{e2}\") and product2 (\"{p2}\") have to be equal")
        return self.add_enzymatic(name, e2, e1, p1, rate, _number(radius, "educt distance"))


def _split_plus(side, distance):
    plus = side.find("+")
    first = side[:plus].strip()
    if not distance:
        return first, side[plus + 1:].strip(), None
    o, c = side.find("("), side.find(")")
    if o < 0 or c < 0 or c < o:
        raise ValueError(f"The side \"{side}\" did not contain a pair of parentheses '(*)'")
    return first, side[c + 1:].strip(), side[o + 1:c].strip()


def _number(s, what):
    try:
        return float(s)
    except (TypeError, ValueError):
        raise ValueError(f"The {what} \"{s}\" is not a number") from None


class _Observables:  # registry/observables.py:58-240, recorded by observe.Recorder
    def __init__(self, sim):
        self._sim = sim
        self._list = []

    def _ids(self, types):
        return None if types is None else [self._sim._system._type_id(t) for t in types]

    def number_of_particles(self, stride, types=None, callback=None, save="default"):
        # NParticles: the listed types' counts, or the total of all particles without types
        self._list.append(("n_particles", dict(stride=stride, types=self._ids(types) if types else [],
                                               callback=callback, name=_save_name(save, "n_particles"))))

    def reaction_counts(self, stride, callback=None, save="default"):
        self._list.append(("reaction_counts", dict(stride=stride, callback=callback,
                                                   name=_save_name(save, "reaction_counts"))))

    def reactions(self, stride, callback=None, save="default"):
        self._list.append(("reactions", dict(stride=stride, callback=callback, name=_save_name(save, "reactions"))))

    def particle_positions(self, stride, types=None, callback=None, save="default"):
        self._list.append(("positions", dict(stride=stride, types=self._ids(types), callback=callback,
                                             name=_save_name(save, "particle_positions"))))

    def particles(self, stride, callback=None, save="default"):
        self._list.append(("particles", dict(stride=stride, callback=callback, name=_save_name(save, "particles"))))

    def forces(self, stride, types=None, callback=None, save="default"):
        self._list.append(("forces", dict(stride=stride, types=self._ids(types), callback=callback,
                                          name=_save_name(save, "forces"))))

    def rdf(self, stride, bin_borders, types_count_from, types_count_to, particle_to_density, callback=None,
            save="default"):
        self._list.append(("radial_distribution", dict(
            stride=stride, borders=bin_borders, from_types=self._ids(types_count_from),
            to_types=self._ids(types_count_to), density=particle_to_density, callback=callback,
            name=_save_name(save, "rdf"))))

    def histogram_along_axis(self, stride, bin_borders, types, axis, callback=None, save="default"):
        self._list.append(("histogram_along_axis", dict(stride=stride, borders=bin_borders, types=self._ids(types),
                                                        axis=axis, callback=callback,
                                                        name=_save_name(save, "histogram_along_axis"))))

    def energy(self, stride, callback=None, save="default"):
        self._list.append(("energy", dict(stride=stride, callback=callback, name=_save_name(save, "energy"))))

    def virial(self, stride, callback=None, save="default"):
        self._list.append(("virial", dict(stride=stride, callback=callback, name=_save_name(save, "virial"))))

    def _unsupported(self, *args, **kwargs):
        raise NotImplementedError("this observable is not available on the MI355X kernel")

    pressure = topologies = _unsupported


def _save_name(save, default):
    if isinstance(save, dict):
        return save.get("name", default)
    return default


class Simulation:
    """One system's simulation on the MI355X engine: particles, observables, the trajectory and run()."""

    def __init__(self, system, output_file, reaction_handler, evaluate_forces, evaluate_observables):
        self._system = system
        self.output_file = output_file
        self.reaction_handler = reaction_handler
        self.evaluate_forces = bool(evaluate_forces)
        self.evaluate_observables = bool(evaluate_observables)
        self.observe = _Observables(self)
        self._trajectory = None
        self._pos = np.zeros((0, 3))
        self._types = np.zeros(0, np.uint32)
        self.seed = 0
        self._t = 0

    @property
    def reaction_handler(self):
        return self._handler

    @reaction_handler.setter
    def reaction_handler(self, value):
        if value not in HANDLERS:
            raise ValueError(f"Requested reaction scheduler {value} is not available.")
        self._handler = value

    def add_particle(self, type, position):
        self.add_particles(type, np.asarray(position, dtype=float).reshape(1, 3))

    def add_particles(self, type, positions):
        p = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        t = self._system._type_id(type)
        if self._system._species[t][2]:
            raise ValueError("topology particles are added with topologies, which this API does not offer")
        self._pos = np.vstack([self._pos, p])
        self._types = np.concatenate([self._types, np.full(len(p), t, np.uint32)])

    def current_particles(self):
        """(type name, position) of every particle, ordered by id"""
        names = self._system.registered_species()
        return [(names[t], p.copy()) for t, p in zip(self._types, self._pos)]

    def record_trajectory(self, stride=1, name="", chunk_size=1000):
        self._trajectory = (int(stride), name, int(chunk_size))

    def _engine(self):
        s = self._system
        if not s.reactions._rx and not s._species:
            raise ValueError("the system has no species")
        flavor = [1 if sp[2] else 0 for sp in s._species]
        e = Engine(len(s._species), [sp[1] for sp in s._species], kbt=s.kbt, potentials=s.potentials._pots,
                   reactions=s.reactions._rx, topology_flavor=flavor if any(flavor) else None)
        r = e.add_replica(s.box_size, s.periodic_boundary_conditions, seed=self.seed,
                          max_particles=max(4 * len(self._types), 1024))
        e.set_particles(r, self._pos, self._types)
        return e, r

    def run(self, n_steps, timestep, show_summary=False):
        """n_steps steps of the simulation loop (SimulationLoop.h:256-306) from the current particles;
        observables and the trajectory go to output_file when one is set"""
        s = self._system
        if not (timestep > 0):
            raise ValueError("the time step must be positive")
        e, r = self._engine()
        try:
            rec = Recorder(e, r)
            if self.evaluate_observables:
                self._register(rec)
            if self.output_file:
                stride = self._trajectory[2] if self._trajectory else 1000
                rec.open(self.output_file, s.registered_species(), s.box_size, s.periodic_boundary_conditions, s.kbt,
                         [sp[1] for sp in s._species], s.reactions._rx,
                         topology_types=[sp[0] for sp in s._species if sp[2]], flush_stride=max(1, stride))
            handler = HANDLERS[self._handler] if s.reactions._rx else L.HANDLER_NONE
            if not self.evaluate_forces:
                raise NotImplementedError("the MI355X loop always evaluates forces")
            if handler == L.HANDLER_DETAILED_BALANCE:
                self._run_separate(e, rec, handler, float(timestep), int(n_steps))
            else:
                rec.run(handler, float(timestep), int(n_steps))
            rec.close()
            pos, types, _ = e.get_particles(r)
            self._pos, self._types = pos, types
            self._t += int(n_steps)
        finally:
            e.close()
        if show_summary:
            print(f"simulated {n_steps} steps of {len(self._types)} particles on the MI355X kernel")

    def _run_separate(self, e, rec, handler, dt, n_steps):
        """the loop as separate actions (integrate, neighbour list, reactions, neighbour list, forces), for a
        handler the fused device loop does not run"""
        e.update_neighbor_list()
        e.calculate_forces()
        rec._evaluate(rec.t)
        for _ in range(n_steps):
            e.integrate(dt)
            e.update_neighbor_list()
            e.react(handler, dt)
            e.update_neighbor_list()
            e.calculate_forces()
            rec.t += 1
            rec._evaluate(rec.t)

    def _register(self, rec):
        for kind, a in self.observe._list:
            cb, name, stride = a["callback"], a["name"], a["stride"]
            if kind == "n_particles":
                rec.n_particles(stride, name=name, types=a["types"], callback=cb)
            elif kind == "reaction_counts":
                rec.reaction_counts(stride, name=name, callback=cb)
            elif kind == "reactions":
                rec.reactions(stride, name=name, callback=cb)
            elif kind == "positions":
                rec.positions(stride, name=name, types=a["types"], callback=cb)
            elif kind == "forces":
                rec.forces(stride, name=name, types=a["types"], callback=cb)
            elif kind == "particles":
                rec.particles(stride, name=name, callback=cb)
            elif kind == "energy":
                rec.energy(stride, name=name, callback=cb)
            elif kind == "virial":
                rec.virial(stride, name=name, callback=cb)
            elif kind == "radial_distribution":
                rec.radial_distribution(stride, a["borders"], a["from_types"], a["to_types"], a["density"], name=name,
                                        callback=cb)
            else:
                rec.histogram_along_axis(stride, a["borders"], a["types"], a["axis"], name=name, callback=cb)
        if self._trajectory is not None:
            rec.trajectory(self._trajectory[0], name=self._trajectory[1])
