/*
 * readdy_oracle.h -- CPU FP64 restatement of ReaDDy's "CPU" kernel per-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine in
 * readdy_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product path never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the readdy/readdy source tree).  Parity status per function is listed in the
 * header comment of readdy_oracle.c.
 */
#ifndef READDY_ORACLE_H
#define READDY_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* pair potential kinds (include/readdy/model/potentials/PotentialsOrder2.h) */
enum {
    OR_POT_HARMONIC_REPULSION = 0,   /* p = {k, d}                                  */
    OR_POT_WEAK_INTERACTION = 1,     /* p = {k, d0, depth, cutoff}                  */
    OR_POT_LENNARD_JONES = 2,        /* p = {m, n, cutoff, shift, epsilon, sigma}   */
    OR_POT_SCREENED_ELECTROSTATICS = 3 /* p = {C, kappa, D, sigma, exponent, cutoff} */
};

/* reaction kinds (include/readdy/model/reactions/{Conversion,Fusion,Fission,Enzymatic,Decay}.h) */
enum {
    OR_RX_CONVERSION = 0,
    OR_RX_FUSION = 1,
    OR_RX_FISSION = 2,
    OR_RX_ENZYMATIC = 3,
    OR_RX_DECAY = 4
};

/* reaction handler modes */
enum {
    OR_UA_KEYED = 0,          /* UA: shuffle realised as ascending Philox u64 keys          */
    OR_GILLESPIE_LITERAL = 1, /* Gillespie: algorithm.h performEvents with a Philox stream  */
    OR_GILLESPIE_KEYED = 2    /* Gillespie: Efraimidis-Spirakis keys + greedy over accepted */
};

typedef struct or_system or_system;

/* ---- Philox4x32-10 (Salmon et al., SC'11) and derived draws ---- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double or_u53(uint32_t a, uint32_t b);

/* ---- periodic boundary helpers (include/readdy/common/boundary_condition_operations.h) ---- */
void or_fix_position(double *v, const double *box, const int *pbc);
void or_shortest_difference(const double *lhs, const double *rhs, const double *box, const int *pbc, double *out);
double or_dist2(const double *lhs, const double *rhs, const double *box, const int *pbc);

/* ---- single-term potentials (energy returned, force accumulated into f) ---- */
double or_pair_force_energy(int kind, const double *params, const double *x_ij, double *f);
double or_bond_force_energy(double k, double l0, const double *xi, const double *xj, const double *box,
                            const int *pbc, double *fi, double *fj);
double or_angle_force_energy(double k, double theta0, const double *xi, const double *xj, const double *xk,
                             const double *box, const int *pbc, double *fi, double *fj, double *fk);
double or_dihedral_force_energy(double k, int multiplicity, double phi0, const double *xi, const double *xj,
                                const double *xk, const double *xl, const double *box, const int *pbc,
                                double *fi, double *fj, double *fk, double *fl);

/* ---- system ---- */
or_system *or_create(int n_types, const double *box, const int *pbc, double kbt, const double *diffusion);
void or_destroy(or_system *s);
int or_add_pair_potential(or_system *s, int kind, int t1, int t2, const double *params);
int or_add_reaction(or_system *s, int kind, int educt1, int educt2, int product1, int product2, double rate,
                    double educt_distance, double product_distance, double weight1, double weight2);
int or_set_seed(or_system *s, uint64_t seed);
int or_set_id_counter(or_system *s, uint64_t next_id);
int or_set_threads(or_system *s, int n);
/* on = 1: with more than one thread, the cell list is filled from all threads by atomic head exchange
 * (CellLinkedList.cpp:171-228); the order within a cell then varies from run to run (timing only) */
int or_set_cas_fill(or_system *s, int on);
/* accumulated wall seconds of or_loop_step's phases: integrator, neighbour list, reactions, forces */
void or_phase_times(or_system *s, double *out4, int reset);
/* CPUHistogramAlongAxis::evaluate (kernels/cpu/src/observables/CPUObservables.cpp:85-136): each live
 * particle of the listed types adds one to bin upper_bound(borders, pos[axis]) - 1 when that index is
 * at least 0 and below n_borders - 2 (the CPU kernel compares against the bin count, so the last bin
 * stays empty). */
int or_histogram_along_axis(const or_system *s, const double *borders, int n_borders, const int *types, int n_types,
                            int axis, uint64_t *counts);
/* RadialDistribution::evaluate counts (RadialDistribution.cpp:75-103): every ordered pair (from, to) of
 * live particles with distinct ids, type of from in from_types and of to in to_types, adds one to the bin
 * upper_bound(borders, sqrt(distSquared)) - 1 when the distance lies below the last border; *n_from is
 * the number of live particles of the from types.  Brute force over all pairs. */
int or_rdf_counts(const or_system *s, const double *borders, int n_borders, const int *from_types, int n_from,
                  const int *to_types, int n_to, uint64_t *counts, uint64_t *n_from_out);
int or_add_particles(or_system *s, int n, const double *pos, const int *types, const uint64_t *ids);
int or_add_bond(or_system *s, size_t i1, size_t i2, double k, double l0);
int or_add_angle(or_system *s, size_t i1, size_t i2, size_t i3, double k, double theta0);
int or_add_dihedral(or_system *s, size_t i1, size_t i2, size_t i3, size_t i4, double k, int multiplicity,
                    double phi0);
size_t or_n_entries(const or_system *s);
size_t or_n_active(const or_system *s);
/* active particles in slot order */
size_t or_get_particles(const or_system *s, size_t cap, double *pos, int *types, uint64_t *ids, double *forces);

/* ---- neighbour list (kernels/cpu/src/nl/CellLinkedList.cpp) ---- */
int or_nl_setup(or_system *s, double cutoff, int radius);
int or_nl_fill(or_system *s);
void or_nl_dims(const or_system *s, int *dims);
/* ordered pairs (id_i, id_j) visited by the cell list with d2 < cutoff^2; returns count */
size_t or_nl_pairs(const or_system *s, double cutoff, size_t cap, uint64_t *out);
size_t or_brute_pairs(const or_system *s, double cutoff, size_t cap, uint64_t *out);

/* ---- actions ---- */
double or_max_cutoff(const or_system *s);
int or_calculate_forces(or_system *s, double *energy, double *virial9);
int or_euler_bd(or_system *s, double dt, const double *noise_by_id, size_t noise_len, uint32_t step);
/* candidate list before Bernoulli draws: rows {nEducts, id1, id2, reaction_index}; returns count */
size_t or_reaction_candidates(or_system *s, size_t cap, uint64_t *out);
int or_reactions(or_system *s, int mode, double dt, uint32_t step, uint64_t *n_performed);
const uint64_t *or_reaction_counts(const or_system *s, int *n);
/* ReactionRecord (include/readdy/model/reactions/ReactionRecord.h:54-63) as performReaction fills it
 * (kernels/cpu ReactionUtils.h:155-225); kept for the last or_reactions call when switched on
 * (Context::recordReactionsWithPositions), cleared at its start (CPUUncontrolledApproximation.cpp:143-145) */
typedef struct {
    uint32_t reaction;
    int32_t type;
    uint64_t educts[2];
    uint64_t products[2];
    uint32_t types_from[2];
    double where[3];
} or_reaction_record;
void or_set_record_reactions(or_system *s, int on);
const or_reaction_record *or_reaction_records(const or_system *s, size_t *n);
/* one SimulationLoop step (include/readdy/api/SimulationLoop.h:256-306): integrate, NL, reactions, NL, forces */
int or_loop_step(or_system *s, double dt, int mode, uint32_t step, const double *noise_by_id, size_t noise_len,
                 double *energy);

/* ---- generic performEvents on abstract events, for distribution tests ---- */
/* events: n rows of {particle1, particle2 (or -1), rate}; accept_p per event; out: performed flags.
 * mode OR_GILLESPIE_LITERAL or OR_GILLESPIE_KEYED; draws from Philox(key=seed, ctr=(trial, i, stream)). */
int or_perform_events_abstract(int n, const int64_t *p1, const int64_t *p2, const double *rate,
                               const double *accept_p, int mode, uint64_t seed, uint32_t trial,
                               uint8_t *performed);

/* ---- DetailedBalance (kernels/singlecpu/src/actions/SCPUReactionImpls.cpp:407-813) ----
 * or_reversible_configs: the reversible pairs (Actions.cpp:432-519) and their ReversibleReactionConfig
 * (Actions.cpp:67-367), total volume = box volume, one row of OR_REV_ROW doubles each:
 * {kind, forward, backward, lhs0, lhs1, rhs0, rhs1, reaction_radius, lhs_radius, rhs_radius, lhs_volume,
 *  rhs_volume, eff_lhs_interaction, eff_rhs_interaction, eff_lhs_reaction, eff_rhs_reaction,
 *  acceptance_prefactor, total_volume, equilibrium_constant, macro_forward, macro_backward}
 * (order-1 sides list the species and 0).  or_fission_table: radii and cumulative probabilities of pair
 * k (FusionFission), returns their count.  or_detailed_balance: one handler call. */
#define OR_REV_ROW 21
int or_reversible_configs(or_system *s, size_t cap, double *rows, size_t *n);
int or_fission_table(or_system *s, int k, double *radii, double *cdf);
int or_detailed_balance(or_system *s, double dt, uint32_t step, uint64_t *n_performed);

const char *or_last_error(void);

/* ---- spatial topology reactions of the topology-particle kinds (enzymatic, fusion)
 *      (kernels/cpu/src/actions/CPUEvaluateTopologyReactions.cpp:88-475) ---- */
int or_add_spatial_reaction(or_system *s, int type1, int top_type1, int type2, int type_to1, int top_type_to1,
                            int type_to2, double rate, double radius);
/* the fusion kind "T1(p1) + (p2) -> T2(p3--p4)": the free particle joins the topology under the next id of
 * the reserve next_id .. end_id - 1; or_topology_fusions lists the last call's joins as rows (topology,
 * topology particle id, old id, new id) */
int or_add_spatial_fusion(or_system *s, int type1, int top_type1, int type2, int type_to1, int top_type_to1,
                          int type_to2, double rate, double radius);
int or_set_topology_reserve(or_system *s, uint64_t next_id, uint64_t end_id);
/* the topology-topology enzymatic kind "T1(p1) + T2(p2) -> T3(p3) + T4(p4)" (STRMode::TT_ENZYMATIC) */
int or_add_spatial_tt(or_system *s, int type1, int top_type1, int type2, int top_type2, int type_to1, int top_type_to1,
                      int type_to2, int top_type_to2, double rate, double radius);
/* the topology-topology fusion kind "T1(p1) + T2(p2) -> T3(p3--p4)" (STRMode::TT_FUSION): the second topology's
 * particles join the first, which becomes top_type_to; or_topology_merges lists (surviving, absorbed, ids) */
int or_add_spatial_tt_fusion(or_system *s, int type1, int top_type1, int type2, int top_type2, int type_to1,
                             int type_to2, int top_type_to, double rate, double radius);
/* the last topology-topology fusion also joins two unbonded particles of one topology (STRMode::
 * TT_FUSION_ALLOW_SELF with the default minimum graph distance); its merge row names the topology twice */
int or_allow_self_fusion(or_system *s);
const uint64_t *or_topology_merges(const or_system *s, size_t *n);
const uint64_t *or_topology_fusions(const or_system *s, size_t *n);
/* topology_of_id[id] = the topology of the particle with that id (-1: none), for ids below n_ids */
int or_set_topologies(or_system *s, size_t n_ids, const int *topology_of_id, size_t n_tops, const int *top_types);
int or_topology_reactions(or_system *s, double dt, uint32_t step, uint64_t *n_performed);
const int *or_topology_types(const or_system *s, size_t *n);
const uint64_t *or_topology_reaction_counts(const or_system *s, int *n);

#ifdef __cplusplus
}
#endif
#endif
