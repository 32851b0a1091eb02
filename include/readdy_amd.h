/*
 * readdy_amd.h -- C ABI of the MI355X ReaDDy compute engine (libreaddy_amd.so).
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  A failing call
 * returns a negative status and leaves a message for rdy_last_error() (thread-local).
 * HIP errors never escape as anything else.
 *
 * One engine holds one chemistry (a ReaDDy Context: species, pair potentials, reactions)
 * and any number of replicas (independent systems with their own box, particles and
 * Philox seed) that step together in shared launches.  A single ReaDDy simulation is an
 * engine with one replica.
 *
 * Each entry point names the reference interface it replaces (paths relative to the
 * readdy/readdy source tree).
 */
#ifndef READDY_AMD_H
#define READDY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RDY_OK 0
#define RDY_ERR_INVALID (-1)  /* bad argument (std::invalid_argument in the reference)           */
#define RDY_ERR_HIP (-2)      /* HIP runtime error                                               */
#define RDY_ERR_CAPACITY (-3) /* a device pool (cells, bins, events, particles) is exhausted     */
#define RDY_ERR_STATE (-4)    /* call out of order (std::logic_error in the reference)           */
#define RDY_ERR_NUMERIC (-5)  /* non-finite position, or a force beyond the fixed-point range of the
                                 force sums (1.68e7 per pair or particle)                            */

/* pair potentials: include/readdy/model/potentials/PotentialsOrder2.h:54-323 */
#define RDY_POT_HARMONIC_REPULSION 0       /* params: force_constant, interaction_distance          */
#define RDY_POT_WEAK_INTERACTION 1         /* params: force_constant, desired_distance, depth, cutoff */
#define RDY_POT_LENNARD_JONES 2            /* params: m, n, cutoff, shift, epsilon, sigma           */
#define RDY_POT_SCREENED_ELECTROSTATICS 3  /* params: C, kappa, D, sigma, exponent, cutoff          */

/* reactions: include/readdy/model/reactions/{Conversion,Fusion,Fission,Enzymatic,Decay}.h */
#define RDY_RX_CONVERSION 0
#define RDY_RX_FUSION 1
#define RDY_RX_FISSION 2
#define RDY_RX_ENZYMATIC 3 /* educts = {from, catalyst}, products = {to, catalyst} (Enzymatic.h:56-57) */
#define RDY_RX_DECAY 4

/* reaction handlers: kernels/cpu/src/actions/reactions/{CPUUncontrolledApproximation,CPUGillespie}.cpp */
#define RDY_HANDLER_UNCONTROLLED_APPROXIMATION 0
#define RDY_HANDLER_GILLESPIE 1
#define RDY_HANDLER_NONE 2
/* DetailedBalance (include/readdy/model/actions/Actions.h:157-176, SCPUReactionImpls.cpp:407-555): the
 * Gillespie selection, then each selected event of a reversible pair a Metropolis move on its energy
 * change.  rdy_react only (not a step handler of rdy_run / rdy_run_graph / rdy_run_begin). */
#define RDY_HANDLER_DETAILED_BALANCE 3

typedef struct rdy_engine rdy_engine;

typedef struct {
    uint32_t kind;
    uint32_t type1, type2;
    double params[8];
} rdy_pair_potential;

typedef struct {
    uint32_t kind;
    uint32_t educts[2];
    uint32_t products[2];
    double rate;
    double educt_distance;
    double product_distance;
    double weight1, weight2;
} rdy_reaction;

/* The system description: readdy::model::Context (include/readdy/model/Context.h:69-261). */
typedef struct {
    uint32_t n_types;
    double kbt;
    const double *diffusion;          /* [n_types]                                            */
    const uint8_t *topology_flavor;   /* [n_types] or NULL: 1 = topology species (bonded)      */
    uint32_t n_potentials;
    const rdy_pair_potential *potentials;
    uint32_t n_reactions;
    const rdy_reaction *reactions;
} rdy_context_desc;

/* Device pools, sized once (288 GB of HBM allows generous defaults). 0 = automatic. */
typedef struct {
    double nl_cutoff;         /* neighbour-list cutoff; 0 = Context::calculateMaxCutoff (Context.cpp:126-144) */
    int32_t cell_radius;      /* cll_radius (KernelConfiguration.h:63): cells of cutoff/radius          */
    uint32_t bin_capacity;    /* particles per cell slot block; 0 = from the expected density           */
    uint32_t max_replicas;
    uint64_t max_cells;       /* total cells over all replicas (0 = max_replicas x the first replica's)  */
    uint64_t max_particles;   /* total live particles over all replicas (0 = max_replicas x the first
                                 replica's 2 max_particles + 1024, at least 4096)                       */
    uint32_t events_per_replica; /* accepted reaction events per replica per step (0 = automatic)       */
    uint32_t block_size;      /* threads per block of the per-particle kernels (0 = 256)               */
    uint32_t brick[3];        /* cells per LDS-staged brick along x, y, z (0 = from the density)        */
    uint32_t lds_states;      /* particle records a brick may stage in LDS (0 = from the density)        */
} rdy_engine_config;

typedef struct {
    double box[3];
    int32_t pbc[3];
    uint64_t seed;            /* Philox key of this replica                                             */
    uint64_t max_particles;   /* expected peak live particles (sizes its cell grid check and events)    */
} rdy_replica_desc;

/* ---- lifecycle: replaces KernelProvider::create + CPUKernel construction/initialize
 *      (readdy/main/plugin/KernelProvider.cpp:183-189, kernels/cpu/src/CPUKernel.cpp:52-83) ---- */
int rdy_engine_create(const rdy_context_desc *ctx, const rdy_engine_config *cfg, rdy_engine **out);
int rdy_engine_destroy(rdy_engine *e);
const char *rdy_last_error(void);
int rdy_device_info(char *name, size_t name_len, uint64_t *hbm_bytes, int32_t *n_cus);

/* add/remove a replica.  Removing frees its cells and events for the next admission. */
int rdy_replica_add(rdy_engine *e, const rdy_replica_desc *desc, uint32_t *replica_id);
int rdy_replica_remove(rdy_engine *e, uint32_t replica);
/* bytes of HBM a replica of this description would occupy (serving admission) */
int rdy_replica_footprint(const rdy_engine *e, const rdy_replica_desc *desc, uint64_t *bytes);

/* ---- particles: replaces StateModel::addParticles / getParticles / removeAllParticles
 *      (include/readdy/model/StateModel.h:66-117, DefaultDataContainer.h:92-118) ----
 * Particle ids on the device are replica-local uint32 handles.  Uploaded particles get
 * handles 0..n-1 in upload order (topology species must come first); reaction products
 * get fresh handles from the replica's counter (Particle::nextId, Particle.h:112-114). */
int rdy_set_particles(rdy_engine *e, uint32_t replica, uint64_t n, const double *pos, const uint32_t *types);
uint64_t rdy_particle_count(rdy_engine *e, uint32_t replica);
/* copies up to cap live particles (order unspecified); returns the count via *n */
int rdy_get_particles(rdy_engine *e, uint32_t replica, uint64_t cap, double *pos, uint32_t *types,
                      uint32_t *handles, uint64_t *n);
/* forces of the last force evaluation, aligned with rdy_get_particles' order of the same state */
int rdy_get_forces(rdy_engine *e, uint32_t replica, uint64_t cap, double *forces, uint32_t *handles, uint64_t *n);

/* ---- topology bonded terms: replaces GraphTopology::configure's flattened lists and
 *      CPUCalculate{HarmonicBond,HarmonicAngle,CosineDihedral}Potential
 *      (readdy/main/model/topologies/GraphTopology.cpp:59-143, CPUTopologyActions.cpp:55-123).
 * Indices are handles of topology particles of that replica. */
int rdy_set_bonds(rdy_engine *e, uint32_t replica, uint64_t n, const uint32_t *idx2, const double *k,
                  const double *length);
int rdy_set_angles(rdy_engine *e, uint32_t replica, uint64_t n, const uint32_t *idx3, const double *k,
                   const double *theta0);
int rdy_set_dihedrals(rdy_engine *e, uint32_t replica, uint64_t n, const uint32_t *idx4, const double *k,
                      const uint32_t *multiplicity, const double *phi0);

/* ---- actions (include/readdy/model/actions/Actions.h:81-454) ---- */
/* CellLinkedList::setUpBins (kernels/cpu/src/nl/CellLinkedList.cpp:230-247) */
int rdy_update_neighbor_list(rdy_engine *e);
/* CPUCalculateForces::perform (kernels/cpu/src/actions/CPUCalculateForces.cpp:46-212) */
int rdy_calculate_forces(rdy_engine *e);
/* total potential energy per replica of the last force evaluation (StateModel::energy) */
int rdy_get_energy(rdy_engine *e, uint32_t replica, double *energy);
/* Context::recordVirial (include/readdy/model/Context.h:168-173): on = 1 makes every force evaluation
 * also form the virial of the pair potentials, as CPUCalculateForces::calculateOrder2 does
 * (kernels/cpu/src/actions/CPUCalculateForces.cpp:214-268): sum over unordered pairs of -x_ij (x) f_i,
 * x_ij = shortestDifference(x_i, x_j), f_i the pair force on i.  Off by default. */
int rdy_record_virial(rdy_engine *e, int on);
/* the virial of the replica's last force evaluation, row-major 3x3 (StateModel::virial, read by the
 * Virial observable, kernels/cpu/src/observables/CPUObservables.cpp:241-245) */
int rdy_get_virial(rdy_engine *e, uint32_t replica, double *virial9);
/* CPUEulerBDIntegrator::perform (kernels/cpu/src/actions/CPUEulerBDIntegrator.cpp:51-103).
 * noise: NULL for the Philox stream, else a device-independent host array [max_handle][3]
 * of N(0,1) draws indexed by handle (single-replica engines only; injected-noise parity). */
int rdy_integrate(rdy_engine *e, double dt, const double *noise, uint64_t noise_handles);
/* CPUUncontrolledApproximation::perform / CPUGillespie::perform / SCPUDetailedBalance::perform
 * (handler RDY_HANDLER_DETAILED_BALANCE; the reversible pairs are those rdy_reversible_reactions lists) */
int rdy_react(rdy_engine *e, int handler, double dt);
/* reaction counts of the last handler call, [n_reactions] (StateModel::reactionCounts) */
int rdy_get_reaction_counts(rdy_engine *e, uint32_t replica, uint64_t *counts);
/* reaction counts accumulated since the replica was added, [n_reactions] */
int rdy_get_reaction_totals(rdy_engine *e, uint32_t replica, uint64_t *counts);
/* one performed reaction: readdy::model::reactions::ReactionRecord (ReactionRecord.h:54-63), filled as
 * the CPU kernel's performReaction fills it (ReactionUtils.h:155-225): educts' handles and types before
 * the reaction (both entries equal for order 1), products' handles after it (0 where none), position =
 * (pos1 + pos2) / 2 wrapped into the box; reaction = index in rdy_context_desc.reactions, type = RDY_RX_* */
typedef struct {
    uint32_t reaction;
    int32_t type;
    uint32_t educts[2];
    uint32_t products[2];
    uint32_t types_from[2];
    double position[3];
} rdy_reaction_record;
/* Context::recordReactionsWithPositions (Context.h:69-261): on = 1 makes every reaction step keep its
 * records (one pool of 64 B x 1024 per replica slot, allocated on the first switch-on); off by default */
int rdy_record_reactions(rdy_engine *e, int on);
/* the records of the replica's last reaction step (StateModel::reactionRecords, as the Reactions
 * observable reads them, Reactions.cpp:60-90), in event-key order; *n = their number (records beyond
 * cap are not written) */
int rdy_get_reaction_records(rdy_engine *e, uint32_t replica, uint64_t cap, rdy_reaction_record *out,
                             uint64_t *n);

/* ReversibleReactionConfig (include/readdy/model/actions/DetailedBalance.h:64-139) as
 * DetailedBalance::searchReversibleReactions pairs them (readdy/main/model/Actions.cpp:432-519):
 * forward / backward index into rdy_context_desc.reactions (for FusionFission the forward reaction is the
 * fusion), lhs / rhs the species of the two sides (second entry 0 for a single species).  Host computation,
 * no device needed; total_volume is the box volume.  Errors as the reference throws them (ambiguous
 * or inconsistent pairs) return RDY_ERR_INVALID with the reference's message.  Writes at most cap rows;
 * *n = the number of pairs. */
#define RDY_REVERSIBLE_FUSION_FISSION 0
#define RDY_REVERSIBLE_CONVERSION_CONVERSION 1
#define RDY_REVERSIBLE_ENZYMATIC_ENZYMATIC 2
typedef struct {
    int32_t kind;
    uint32_t forward, backward;
    uint32_t n_lhs, n_rhs;
    uint32_t lhs[2], rhs[2];
    double reaction_radius;
    double lhs_interaction_radius, rhs_interaction_radius;
    double lhs_interaction_volume, rhs_interaction_volume;
    double eff_lhs_interaction_volume, eff_rhs_interaction_volume;
    double eff_lhs_reaction_volume, eff_rhs_reaction_volume;
    double acceptance_prefactor;
    double total_volume;
    double kbt;
    double equilibrium_constant;
    double macro_forward_rate, macro_backward_rate;
    double micro_forward_rate, micro_backward_rate;
} rdy_reversible_reaction;
int rdy_reversible_reactions(const rdy_context_desc *ctx, double total_volume, uint64_t cap,
                             rdy_reversible_reaction *out, uint64_t *n);

/* BreakBonds (include/readdy/model/actions/Actions.h:218-341) with its BreakConfig (Actions.h:190-211):
 * a bond edge (the consecutive bond terms of one replica on one handle pair, i.e. all bond
 * configurations of a topology edge) whose end types form a listed unordered pair and whose summed
 * harmonic energy exceeds the pair's threshold breaks with probability 1 - exp(-rate dt) (Philox
 * uniform per edge and call).  Writes the broken edges as (replica, h1, h2) rows sorted ascending, at
 * most cap of them; *n_broken = their number.  The bonded terms are left as they are: the caller
 * removes the edges from its graphs and sets the replica's bonds, angles and dihedrals again
 * (GraphTopology::configure), as the reference's structural reaction does.  rate <= 0 is an error
 * (BreakConfig::addBreakablePair). */
typedef struct {
    uint32_t type1, type2;
    double threshold_energy, rate;
} rdy_break_pair;
int rdy_break_bonds(rdy_engine *e, double dt, uint32_t n_pairs, const rdy_break_pair *pairs, uint64_t cap,
                    uint32_t *broken, uint64_t *n_broken);

/* ---- spatial topology reactions of the topology-particle kinds: enzymatic "T1(p1) + (p2) -> T2(p3) + (p4)"
 *      and fusion "T1(p1) + (p2) -> T2(p3--p4)": SpatialTopologyReaction.h:64-236 (STRMode::TP_ENZYMATIC,
 *      TP_FUSION), evaluated as CPUEvaluateTopologyReactions.cpp:88-475 does.  A particle of species type1 in
 *      a topology of type top_type1 and a particle of no topology of species type2 closer than radius react
 *      with probability 1 - exp(-rate dt); at most one reaction per topology and per free particle and step
 *      (performEvents with eventsDependent); the topology particle becomes type_to1, the free one type_to2,
 *      the topology top_type_to1.  Fusion (handleTopologyParticleReaction, :369-404, appendParticle): the free
 *      particle joins the topology, bonded to the topology particle; it takes the next handle of the
 *      replica's topology reserve (rdy_reserve_topology_particles), and rdy_get_topology_fusions lists the
 *      joins.  Topology-topology enzymatic (STRMode::TT_ENZYMATIC, handleTopologyTopologyReaction, :406-475):
 *      two topology particles of different topologies whose (species, topology type) pairs match the reaction
 *      in either order react once per pair; each particle and topology takes the product of its side, and
 *      both topologies count as reacted (eventsDependent).  Topology types are the caller's integers.  The
 *      bonded terms stay as they are: the caller sets them again when a reaction changes the topology's
 *      bonded configuration (every fusion does). */
typedef struct {
    uint32_t type1, top_type1;   /* the topology particle's species and its topology's type */
    uint32_t type2;              /* the free particle's species                               */
    uint32_t type_to1, top_type_to1, type_to2;
    double rate, radius;
    uint32_t fusion;             /* 0: TP_ENZYMATIC, 1: TP_FUSION (type_to2 a topology species) */
    uint32_t topology_topology;  /* 1: TT_ENZYMATIC "T1(p1) + T2(p2) -> T3(p3) + T4(p4)": the second particle is
                                    a topology particle of another topology, of type top_type2, which becomes
                                    top_type_to2; with fusion 1, TT_FUSION "T1(p1) + T2(p2) -> T3(p3--p4)": the
                                    second topology merges into the first, which becomes top_type_to1 */
    uint32_t top_type2, top_type_to2;
    uint32_t self_fusion;        /* topology-topology fusion "[self=true]" (TT_FUSION_ALLOW_SELF): two particles of
                                    one topology that are not bonded to each other fuse as well, adding the edge */
    uint32_t reserved;           /* 0 */
} rdy_spatial_reaction;
/* the engine's spatial topology reactions (TopologyRegistry::addSpatialReaction, TopologyRegistry.cpp:195-223);
 * radius <= the engine's neighbour-list cutoff, at most 64 reactions */
int rdy_set_spatial_reactions(rdy_engine *e, uint32_t n, const rdy_spatial_reaction *reactions);
/* the replica's topologies: their types [n_topologies] and, for each topology particle (handles 0 ..
 * n_topo - 1: the last upload's and those fusions added), the index of its topology [n_topo]
 * (StateModel::addTopology), or UINT32_MAX for one that left its topology (a structural reaction made it a
 * normal particle) */
int rdy_set_topologies(rdy_engine *e, uint32_t replica, uint32_t n_topologies, const uint32_t *topology_types,
                       const uint32_t *topology_of);
/* Topology-particle fusion needs handles for the particles that join topologies.  Called before
 * rdy_set_particles, this reserves `n` handles after the upload's topology particles: the upload then
 * numbers its topology particles 0 .. n_topo - 1 and its free particles from n_topo + n on, and each
 * fusion gives the joining particle the next reserved handle (its id changes; ReaDDy keeps the id).  A
 * fusion beyond the reserve is a capacity error.  The reserve holds for the replica's next uploads. */
int rdy_reserve_topology_particles(rdy_engine *e, uint32_t replica, uint32_t n);
/* the fusions of the last rdy_evaluate_topology_reactions call (or fused step) of the replica, in the order
 * they were performed: rows [n][4] of (topology index, topology particle's handle, the joining particle's
 * handle before, its handle now); *n = their count.  The replica's topology particles are then handles
 * 0 .. n_topo - 1 with n_topo grown by the fusions (rdy_set_topologies takes that many entries). */
int rdy_get_topology_fusions(rdy_engine *e, uint32_t replica, uint32_t cap, uint32_t *rows, uint32_t *n);
/* Recipe steps of structural topology reactions that move or add topology particles (TopologyReactionActions:
 * ChangeParticlePosition, AppendParticle; Recipe.h:40-132).  rdy_set_particle_positions moves topology particles
 * (by handle) and re-bins the engine's particles, as a neighbour-list update does.  rdy_append_topology_particles
 * adds n particles of topology species at the given positions under the next handles of the replica's topology
 * reserve (rdy_reserve_topology_particles), returned in handles[n]; the caller then assigns them to topologies
 * (rdy_set_topologies) and sets the bonded terms.  Beyond the reserve: RDY_ERR_CAPACITY. */
int rdy_set_particle_positions(rdy_engine *e, uint32_t replica, uint64_t n, const uint32_t *handles, const double *pos);
int rdy_append_topology_particles(rdy_engine *e, uint32_t replica, uint64_t n, const uint32_t *types,
                                  const double *pos, uint32_t *handles);
/* the topology-topology fusions of the last evaluation (or fused step) of the replica, in key order: rows [n][4]
 * of (surviving topology, absorbed topology, the surviving side's particle, the absorbed side's particle); the
 * absorbed topology keeps its index with no particles until the caller sets the topologies again
 * (Topology::appendTopology and the erase of the absorbed one, CPUEvaluateTopologyReactions.cpp:433-437) */
int rdy_get_topology_merges(rdy_engine *e, uint32_t replica, uint32_t cap, uint32_t *rows, uint32_t *n);
/* the replica's topology types as the topology reactions left them; *n = the number of topologies */
int rdy_get_topology_types(rdy_engine *e, uint32_t replica, uint32_t cap, uint32_t *types, uint32_t *n);
/* per topology, 1 if a spatial topology reaction changed it in the last rdy_evaluate_topology_reactions call:
 * the structural reactions of the same evaluation skip those topologies (one topology reaction per topology
 * and evaluation, eventsDependent in CPUEvaluateTopologyReactions.cpp:455-464); *n = the number of topologies */
int rdy_get_topology_reacted(rdy_engine *e, uint32_t replica, uint32_t cap, uint8_t *reacted, uint32_t *n);
/* TopologyReactionActionFactory::createChangeParticleType (TopologyReactionActions.cpp:40-60) for the recipes
 * of structural topology reactions: topology particles (handles below the replica's topology count) take
 * the given species; a species of normal flavour takes the particle out of the bonded-term bookkeeping (the
 * caller drops it from its topology).  Forces and energies are formed anew by the next force evaluation. */
int rdy_set_particle_types(rdy_engine *e, uint32_t replica, uint64_t n, const uint32_t *handles,
                           const uint32_t *types);
/* EvaluateTopologyReactions (Actions.h, CPUEvaluateTopologyReactions::perform) on the current state; the
 * fused loop (rdy_run, rdy_run_graph) performs it in every step after the reactions (SimulationLoop.h:286-287) */
int rdy_evaluate_topology_reactions(rdy_engine *e, double dt);
/* per spatial reaction: performed in the last evaluation (totals = 0) or since the replica was added (1) */
int rdy_get_topology_reaction_counts(rdy_engine *e, uint32_t replica, int totals, uint64_t *counts);

/* ---- the simulation loop: replaces api::SimulationLoop::run's body
 *      (include/readdy/api/SimulationLoop.h:256-306): per step integrate, neighbour list,
 *      reactions, neighbour list, forces -- fused into shared launches over all replicas.
 * Runs n_steps for every active replica; rdy_run_graph captures one step in a hipGraph
 * and replays it. */
int rdy_run(rdy_engine *e, int handler, double dt, uint64_t n_steps);
int rdy_run_graph(rdy_engine *e, int handler, double dt, uint64_t n_steps);
/* The same run taken a few steps at a time, with no host synchronisation in between: rdy_run_begin
 * enqueues the run's first walk, each rdy_run_enqueue(k) enqueues the next k fused steps (the run's last
 * step copies instead of moving, as in rdy_run), and rdy_run_end waits for the device and reports errors.
 * Several engines can interleave their steps this way on their own streams.  While a run is open the
 * engine's other calls that change state are refused (RDY_ERR_STATE).  No reference counterpart
 * (SimulationLoop::run performs one whole run, SimulationLoop.h:256-306). */
int rdy_run_begin(rdy_engine *e, int handler, double dt, uint64_t n_steps);
int rdy_run_enqueue(rdy_engine *e, uint64_t k);
int rdy_run_end(rdy_engine *e);
/* per-replica step budget for serving: a replica stops stepping after this many steps */
int rdy_replica_set_steps(rdy_engine *e, uint32_t replica, uint64_t steps);
int rdy_replica_steps_done(rdy_engine *e, uint32_t replica, uint64_t *steps);
/* per-replica reaction handler (RDY_HANDLER_*; -1 = the handler the step is launched with), so that
 * UncontrolledApproximation and Gillespie systems share launches; a launch with RDY_HANDLER_NONE
 * performs no reactions for any replica.  No reference counterpart (one handler per simulation,
 * SimulationLoop.h:104-109). */
int rdy_replica_set_handler(rdy_engine *e, uint32_t replica, int handler);
int rdy_synchronize(rdy_engine *e);

/* ---- test exports (parity): the pair set of the cell list, the candidate list ---- */
/* ordered pairs (handle_i, handle_j) with d2 < cutoff^2 seen by the cell-list traversal */
int rdy_neighbor_pairs(rdy_engine *e, uint32_t replica, double cutoff, uint64_t cap, uint32_t *pairs,
                       uint64_t *n);
/* candidates before the Bernoulli draw: rows {n_educts, handle1, handle2, reaction} */
int rdy_reaction_candidates(rdy_engine *e, uint32_t replica, uint64_t cap, uint32_t *rows, uint64_t *n);

/* observables: particles per type for every replica, [n_replicas][n_types] */
int rdy_count_types(rdy_engine *e, uint64_t *counts, uint32_t n_replicas);
/* radial distribution counts of one replica, replacing the pair loop of RadialDistribution::evaluate
 * (readdy/main/model/observables/RadialDistribution.cpp:75-103): every ordered pair of live particles
 * (from, to), from's type in from_types and to's in to_types, adds one to counts[b] where borders[b] <=
 * |shortestDifference| < borders[b + 1] (upper_bound); *n_from = live particles of the from types.
 * 2 <= n_borders <= 4097, borders ascending.  The distribution itself, counts / (4/3 pi (r_b+1^3 -
 * r_b^3) n_from particleToDensity) (:105-123), is formed by the caller. */
int rdy_rdf_counts(rdy_engine *e, uint32_t replica, const double *borders, uint32_t n_borders,
                   const uint32_t *from_types, uint32_t n_from_types, const uint32_t *to_types,
                   uint32_t n_to_types, uint64_t *counts, uint64_t *n_from);
/* particle histogram along one axis of one replica, replacing CPUHistogramAlongAxis::evaluate
 * (kernels/cpu/src/observables/CPUObservables.cpp:85-136): a live particle of the listed types adds
 * one to counts[u - 1], u = upper_bound(borders, position[axis]), for 1 <= u < n_borders - 1 (the CPU
 * kernel's bounds: its last bin stays empty).  2 <= n_borders <= 4097. */
int rdy_histogram_along_axis(rdy_engine *e, uint32_t replica, const double *borders, uint32_t n_borders,
                             const uint32_t *types, uint32_t n_types, uint32_t axis, uint64_t *counts);

/* The same histogram as rdy_histogram_along_axis, computed without allocating: no hipMalloc, hipFree or
 * other allocating routine is called.  No counterpart in the reference; it is for callers that cannot
 * allocate device memory (HIP graph capture around it, tight serving loops, harnesses that forbid it).
 * Its workspace (borders, species mask, counts: 66,048 bytes) is a static device array, one per GPU, held
 * until the process ends; 2 <= n_borders <= 4097.  Every call on one GPU shares that GPU's workspace and
 * leaves its contents behind (the counts are cleared at the start of each call): calls that overlap on
 * one GPU (several host threads, or several streams) race and must be serialised by the caller. */
int rdy_histogram_along_axis_noalloc(rdy_engine *e, uint32_t replica, const double *borders, uint32_t n_borders,
                                     const uint32_t *types, uint32_t n_types, uint32_t axis, uint64_t *counts);

/* kernel timing of the dominant (fused force/integrate/rebin) kernel, measured with HIP
 * events on the engine's stream: average microseconds over the launches since the last reset */
int rdy_profile_reset(rdy_engine *e);
int rdy_profile_get(rdy_engine *e, double *avg_us_fused, uint64_t *launches);

#ifdef __cplusplus
}
#endif
#endif
