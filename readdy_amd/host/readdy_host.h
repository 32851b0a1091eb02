// Host-side C++ core of the MI355X ReaDDy kernel: the object contract a ReaDDy simulation talks to
// (Context, Kernel, StateModel, ActionFactory / Action, KernelProvider, SimulationLoop), shaped like
// the reference's readdy::model / readdy::plugin / readdy::api classes so that a caller written
// against those reads the same here.  Every device operation goes through the C ABI of
// libreaddy_amd.so (include/readdy_amd.h); this layer only owns the host bookkeeping (registries,
// particle-id mapping, topology graphs) and turns status codes into exceptions.
//
// Reference interfaces mirrored (paths relative to the readdy/readdy tree):
//   Context                 include/readdy/model/Context.h:69-261
//   ParticleTypeRegistry    include/readdy/model/_internal/ParticleTypeRegistry.h:60-199
//   PotentialRegistry       include/readdy/model/potentials/PotentialRegistry.h:55-380
//   ReactionRegistry        include/readdy/model/reactions/ReactionRegistry.h:59-210
//   TopologyRegistry (potential configuration only)
//                           include/readdy/model/topologies/TopologyRegistry.h, api/PotentialConfiguration.h:136-161
//   Kernel                  include/readdy/model/Kernel.h:71-235
//   StateModel              include/readdy/model/StateModel.h:66-117
//   Action / ActionFactory  include/readdy/model/actions/Action.h:58-79, ActionFactory.h:93-128
//   KernelProvider          include/readdy/plugin/KernelProvider.h, readdy/main/plugin/KernelProvider.cpp:68-209
//   SimulationLoop          include/readdy/api/SimulationLoop.h:71-306
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_map>
#include <utility>
#include <vector>

namespace readdy_amd {

using scalar = double;

struct Vec3 {
    scalar x = 0, y = 0, z = 0;
    Vec3() = default;
    Vec3(scalar a, scalar b, scalar c) : x(a), y(b), z(c) {}
    scalar operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

using ParticleTypeId = uint16_t;  // common.h:68
using ParticleId = uint64_t;      // common.h:69

namespace model {

// ---- registries ---------------------------------------------------------------------------------
struct ParticleTypeInfo {
    std::string name;
    ParticleTypeId typeId;
    scalar diffusionConstant;
    bool topologyFlavor;  // particleflavor::TOPOLOGY (ParticleTypeRegistry.h:52-58)
};

class ParticleTypeRegistry {
public:
    void add(const std::string &name, scalar diffusionConstant, bool topologyFlavor = false);
    void addTopologyType(const std::string &name, scalar diffusionConstant) { add(name, diffusionConstant, true); }
    ParticleTypeId idOf(const std::string &name) const;
    const std::string &nameOf(ParticleTypeId id) const;
    const ParticleTypeInfo &infoOf(ParticleTypeId id) const;
    scalar diffusionConstantOf(const std::string &name) const { return infoOf(idOf(name)).diffusionConstant; }
    std::size_t nTypes() const { return _types.size(); }
    const std::vector<ParticleTypeInfo> &types() const { return _types; }

private:
    std::vector<ParticleTypeInfo> _types;
    std::unordered_map<std::string, ParticleTypeId> _byName;
};

struct PairPotentialDesc {
    uint32_t kind;  // RDY_POT_*
    ParticleTypeId type1, type2;
    std::array<scalar, 8> params{};
    scalar cutoff() const;
};

class PotentialRegistry {
public:
    explicit PotentialRegistry(const ParticleTypeRegistry *types) : _types(types) {}
    void rebind(const ParticleTypeRegistry *types) { _types = types; }  // after a Context copy
    // PotentialRegistry.h:97-160 (order 2)
    void addHarmonicRepulsion(const std::string &t1, const std::string &t2, scalar forceConstant,
                              scalar interactionDistance);
    void addWeakInteractionPiecewiseHarmonic(const std::string &t1, const std::string &t2, scalar forceConstant,
                                             scalar desiredDistance, scalar depth, scalar cutoff);
    void addLennardJones(const std::string &t1, const std::string &t2, unsigned m, unsigned n, scalar cutoff,
                         bool shift, scalar epsilon, scalar sigma);
    void addScreenedElectrostatics(const std::string &t1, const std::string &t2, scalar electrostaticStrength,
                                   scalar inverseScreeningDepth, scalar repulsionStrength, scalar repulsionDistance,
                                   unsigned exponent, scalar cutoff);
    const std::vector<PairPotentialDesc> &potentialsOrder2() const { return _order2; }

private:
    const ParticleTypeRegistry *_types;
    std::vector<PairPotentialDesc> _order2;
};

enum class ReactionType { Conversion, Fusion, Fission, Enzymatic, Decay };  // Reaction.h:60-61

struct ReactionDesc {
    std::string name;
    ReactionType type;
    std::vector<ParticleTypeId> educts, products;
    scalar rate = 0, eductDistance = 0, productDistance = 0, weight1 = .5, weight2 = .5;
};

// one performed reaction (ReactionRecord.h:54-63): type as ReactionType, educts' ids and types before
// the reaction, products' ids after it, where = (pos1 + pos2) / 2 in the box, id = reaction index
struct ReactionRecord {
    int type = 0;
    std::array<ParticleId, 2> educts{{0, 0}};
    std::array<ParticleId, 2> products{{0, 0}};
    std::array<ParticleTypeId, 2> types_from{{0, 0}};
    Vec3 where;
    std::size_t id = 0;
};

class ReactionRegistry {
public:
    explicit ReactionRegistry(const ParticleTypeRegistry *types) : _types(types) {}
    void rebind(const ParticleTypeRegistry *types) { _types = types; }  // after a Context copy
    // ReactionRegistry.h:106-158
    void addConversion(const std::string &name, const std::string &from, const std::string &to, scalar rate);
    void addEnzymatic(const std::string &name, const std::string &catalyst, const std::string &from,
                      const std::string &to, scalar rate, scalar eductDistance);
    void addFission(const std::string &name, const std::string &from, const std::string &to1,
                    const std::string &to2, scalar rate, scalar productDistance, scalar weight1 = .5,
                    scalar weight2 = .5);
    void addFusion(const std::string &name, const std::string &from1, const std::string &from2,
                   const std::string &to, scalar rate, scalar eductDistance, scalar weight1 = .5,
                   scalar weight2 = .5);
    void addDecay(const std::string &name, const std::string &type, scalar rate);
    // the descriptor form, ReactionRegistry::add (ReactionRegistry.cpp:154-305):
    //   "name: A ->"                    decay        "name: A -> B"              conversion
    //   "name: A +(r) B -> C [w1, w2]"  fusion       "name: A -> B +(d) C [w1, w2]"  fission
    //   "name: A +(r) C -> B + C"       enzymatic (C the catalyst); weights optional (0.5, 0.5)
    // returns the reaction's index; malformed descriptors throw std::invalid_argument
    std::size_t add(const std::string &descriptor, scalar rate);
    const std::vector<ReactionDesc> &reactions() const { return _reactions; }
    std::size_t idOf(const std::string &name) const;

private:
    void append(ReactionDesc r);  // rejects a name that exists already (ReactionRegistry.cpp:62-63)
    const ParticleTypeRegistry *_types;
    std::vector<ReactionDesc> _reactions;
};

struct BondConfiguration { scalar forceConstant, length; };                         // PotentialConfiguration.h
struct AngleConfiguration { scalar forceConstant, equilibriumAngle; };
struct TorsionConfiguration { scalar forceConstant; unsigned multiplicity; scalar phi_0; };

enum class STRMode { TT_ENZYMATIC, TT_FUSION, TT_FUSION_ALLOW_SELF, TP_ENZYMATIC, TP_FUSION };  // SpatialTopologyReaction.h:64-67

// a spatial topology reaction as STRParser reads it (SpatialTopologyReaction.cpp:81-220); the second
// topology type is EmptyTopologyId for the topology-particle kinds
struct SpatialReactionDesc {
    std::string name;
    STRMode mode;
    ParticleTypeId type1, type2, typeTo1, typeTo2;
    std::size_t topType1, topType2, topTypeTo1, topTypeTo2;
    scalar rate, radius;
    unsigned minGraphDistance = 0;  // self-fusion: pairs at a graph distance up to this do not react ("distance>N")
    bool isTopologyParticle() const { return mode == STRMode::TP_ENZYMATIC || mode == STRMode::TP_FUSION; }
};

class StructuralTopologyReaction;

class TopologyRegistry {
public:
    static constexpr std::size_t EmptyTopologyId = static_cast<std::size_t>(-1);
    explicit TopologyRegistry(const ParticleTypeRegistry *types) : _types(types) {}
    // topology types (TopologyRegistry::addType); ids in registration order
    std::size_t addType(const std::string &name);
    std::size_t idOf(const std::string &name) const;  // throws std::invalid_argument for an unknown type
    bool hasType(const std::string &name) const;
    const std::string &nameOf(std::size_t id) const;
    // TopologyRegistry::addSpatialReaction(descriptor, rate, radius) (TopologyRegistry.cpp:204-223):
    //   "name: T1(p1) + (p2) -> T2(p3) + (p4)"  topology-particle enzymatic
    //   "name: T1(p1) + (p2) -> T2(p3--p4)"     topology-particle fusion
    //   "name: T1(p1) + T2(p2) -> T3(p3) + T4(p4)" / "... -> T3(p3--p4) [self=true]"  topology-topology kinds
    // Malformed descriptors, unknown types and a non-positive rate throw std::invalid_argument.
    void addSpatialReaction(const std::string &descriptor, scalar rate, scalar radius);
    const std::vector<SpatialReactionDesc> &spatialReactions() const { return _spatial; }
    bool isSpatialReactionType(ParticleTypeId t) const;
    // TopologyRegistry::addStructuralReaction (TopologyRegistry.h:108-122): a reaction of one topology type
    void addStructuralReaction(const std::string &type, StructuralTopologyReaction reaction);
    const std::vector<StructuralTopologyReaction> &structuralReactionsOf(const std::string &type) const;
    bool hasStructuralReactions() const { return !_structural.empty(); }
    void rebind(const ParticleTypeRegistry *types) { _types = types; }  // after a Context copy
    void configureBondPotential(const std::string &t1, const std::string &t2, BondConfiguration cfg);
    void configureAnglePotential(const std::string &t1, const std::string &t2, const std::string &t3,
                                 AngleConfiguration cfg);
    void configureTorsionPotential(const std::string &t1, const std::string &t2, const std::string &t3,
                                   const std::string &t4, TorsionConfiguration cfg);
    // the potential configuration, keyed by ordered type tuples (both directions are registered)
    std::map<std::tuple<ParticleTypeId, ParticleTypeId>, std::vector<BondConfiguration>> bonds;
    std::map<std::tuple<ParticleTypeId, ParticleTypeId, ParticleTypeId>, std::vector<AngleConfiguration>> angles;
    std::map<std::tuple<ParticleTypeId, ParticleTypeId, ParticleTypeId, ParticleTypeId>,
             std::vector<TorsionConfiguration>>
        torsions;

private:
    const ParticleTypeRegistry *_types;
    std::vector<std::string> _topologyTypes;
    std::vector<SpatialReactionDesc> _spatial;
    std::map<std::string, std::vector<StructuralTopologyReaction>> _structural;
};

class Context {
public:
    Context();
    Context(const Context &o);
    Context &operator=(const Context &o);
    ParticleTypeRegistry &particleTypes() { return _types; }
    const ParticleTypeRegistry &particleTypes() const { return _types; }
    PotentialRegistry &potentials() { return _potentials; }
    const PotentialRegistry &potentials() const { return _potentials; }
    ReactionRegistry &reactions() { return _reactions; }
    const ReactionRegistry &reactions() const { return _reactions; }
    TopologyRegistry &topologyRegistry() { return _topologies; }
    const TopologyRegistry &topologyRegistry() const { return _topologies; }
    std::array<scalar, 3> &boxSize() { return _box; }
    const std::array<scalar, 3> &boxSize() const { return _box; }
    std::array<bool, 3> &periodicBoundaryConditions() { return _pbc; }
    const std::array<bool, 3> &periodicBoundaryConditions() const { return _pbc; }
    scalar &kBT() { return _kbt; }
    scalar kBT() const { return _kbt; }
    // Context::recordReactionsWithPositions (Context.h): keep each reaction step's records
    bool &recordReactionsWithPositions() { return _recordReactions; }
    bool recordReactionsWithPositions() const { return _recordReactions; }
    // Context::recordVirial (Context.h:168-173): force evaluations also form the pair virial
    bool &recordVirial() { return _recordVirial; }
    bool recordVirial() const { return _recordVirial; }
    // Context::calculateMaxCutoff (Context.cpp:126-144): pair potentials and order-2 educt distances
    scalar calculateMaxCutoff() const;

private:
    void rebind();
    ParticleTypeRegistry _types;
    PotentialRegistry _potentials{&_types};
    ReactionRegistry _reactions{&_types};
    TopologyRegistry _topologies{&_types};
    std::array<scalar, 3> _box{{1, 1, 1}};
    std::array<bool, 3> _pbc{{true, true, true}};
    scalar _kbt = 1;
    bool _recordReactions = false;
    bool _recordVirial = false;
};

class Particle {  // Particle.h:60-120
public:
    Particle(scalar x, scalar y, scalar z, ParticleTypeId type, ParticleId id = 0) : _pos(x, y, z), _type(type), _id(id) {}
    Particle(Vec3 pos, ParticleTypeId type, ParticleId id = 0) : _pos(pos), _type(type), _id(id) {}
    const Vec3 &pos() const { return _pos; }
    ParticleTypeId type() const { return _type; }
    ParticleId id() const { return _id; }

private:
    Vec3 _pos;
    ParticleTypeId _type;
    ParticleId _id;
};

class Kernel;

// A linear or branched topology: its particles and graph edges (GraphTopology.h; the graph
// library's findNTuples enumerates its bonds, angles and dihedrals, Graph_detail.h:39-78)
class GraphTopology {
public:
    GraphTopology(std::string type, std::vector<std::size_t> particleIndices)
        : _type(std::move(type)), _particles(std::move(particleIndices)), _neighbors(_particles.size()) {}
    void addEdge(std::size_t v1, std::size_t v2);  // vertex indices into particleIndices()
    // GraphTopology::appendParticle (GraphTopology.cpp): a new vertex for particle `particle`, bonded to the
    // vertex of particle `counterpart`
    void appendParticle(std::size_t particle, std::size_t counterpart);
    // a new vertex for particle `particle`, without edges; returns its vertex index
    std::size_t addVertex(std::size_t particle);
    // GraphTopology::appendTopology: the other topology's vertices and edges join this one, with an edge between
    // the vertices of particle `thisParticle` (here) and `otherParticle` (there)
    void appendTopology(const GraphTopology &other, std::size_t otherParticle, std::size_t thisParticle);
    bool removeEdge(std::size_t v1, std::size_t v2);  // false when there is no such edge
    bool hasEdge(std::size_t v1, std::size_t v2) const;
    std::size_t nParticles() const { return _particles.size(); }
    // the connected components as topologies of the same type, vertices in ascending order and each
    // vertex's neighbours in their present order (one component: a copy of this topology)
    std::vector<GraphTopology> connectedComponents() const;
    const std::vector<std::size_t> &particleIndices() const { return _particles; }
    const std::string &type() const { return _type; }
    void setType(std::string type) { _type = std::move(type); }
    const std::vector<std::size_t> &neighborsOf(std::size_t v) const { return _neighbors.at(v); }
    bool isConnected() const;  // Graph::isConnected (a topology of one vertex is connected)
    // the structural reactions' rates for this topology (GraphTopology::updateReactionRates)
    void updateReactionRates(const std::vector<StructuralTopologyReaction> &reactions);
    const std::vector<scalar> &rates() const { return _rates; }
    // one particle of normal flavour: no longer a topology (GraphTopology::isNormalParticle)
    bool isNormalParticle(const Kernel &kernel) const;
    // visits every edge, 3-path and 4-path once, in the reference's enumeration order
    void findNTuples(const std::function<void(std::size_t, std::size_t)> &pair,
                     const std::function<void(std::size_t, std::size_t, std::size_t)> &triple,
                     const std::function<void(std::size_t, std::size_t, std::size_t, std::size_t)> &quad) const;

private:
    std::string _type;
    std::vector<std::size_t> _particles;
    std::vector<std::vector<std::size_t>> _neighbors;
    std::vector<scalar> _rates;
};

// ---- structural topology reactions (StructuralTopologyReaction.h:58-220, Recipe.h:40-132) -------
// thrown when a reaction leaves a topology that was to stay connected disconnected or with particles
// of normal flavour (TopologyReactionException.h)
class TopologyReactionException : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

// the operations a structural reaction applies to its topology, in order; vertices are indices into
// the topology's particleIndices()
class Recipe {
public:
    struct Step {
        enum Kind { ChangeParticleType, AddEdge, RemoveEdge, SeparateVertex, ChangeTopologyType, ChangeParticlePosition,
                    AppendNewParticle };
        Kind kind;
        std::size_t v1 = 0, v2 = 0;
        ParticleTypeId type = 0;
        std::string name;  // a species name to resolve, or the new topology type
        bool byName = false;
        Vec3 pos;
        std::vector<std::size_t> neighbors;
    };
    explicit Recipe(GraphTopology &topology) : _topology(&topology) {}
    Recipe &changeParticleType(std::size_t vertex, ParticleTypeId to);
    Recipe &changeParticleType(std::size_t vertex, const std::string &to);
    Recipe &addEdge(std::size_t v1, std::size_t v2);
    Recipe &removeEdge(std::size_t v1, std::size_t v2);
    Recipe &separateVertex(std::size_t vertex);  // removes every edge of the vertex (Recipe.cpp:56-62)
    Recipe &changeTopologyType(const std::string &type);
    // in the reference's recipe; a kernel that cannot move or create particles in a topology throws when
    // the reaction executes
    Recipe &changeParticlePosition(std::size_t vertex, const Vec3 &pos);
    Recipe &appendNewParticle(const std::vector<std::size_t> &neighbors, const std::string &type, const Vec3 &pos);
    const std::vector<Step> &steps() const { return _steps; }
    GraphTopology &topology() { return *_topology; }
    const GraphTopology &topology() const { return *_topology; }

private:
    GraphTopology *_topology;
    std::vector<Step> _steps;
};

class StructuralTopologyReaction {
public:
    using reaction_function = std::function<Recipe(GraphTopology &)>;
    using rate_function = std::function<scalar(const GraphTopology &)>;
    StructuralTopologyReaction(std::string name, reaction_function reaction, rate_function rate);
    StructuralTopologyReaction(std::string name, reaction_function reaction, scalar rate);
    scalar rate(const GraphTopology &topology) const { return _rate(topology); }
    Recipe operations(GraphTopology &topology) const { return _reaction(topology); }
    // Mode (StructuralTopologyReaction.h:58-70): by default the topology may fall apart into children
    bool expects_connected_after_reaction() const { return _expectConnected; }
    void expect_connected_after_reaction() { _expectConnected = true; }
    bool creates_child_topologies_after_reaction() const { return !_expectConnected; }
    void create_child_topologies_after_reaction() { _expectConnected = false; }
    const std::string &name() const { return _name; }
    std::size_t id() const { return _id; }
    // StructuralTopologyReaction::execute (StructuralTopologyReaction.cpp:61-131): the recipe through the
    // kernel's topology action factory, then the post-reaction checks.  Returns the connected components
    // when the graph fell apart (create-children mode); otherwise empty, the topology reconfigured.
    std::vector<GraphTopology> execute(GraphTopology &topology, Kernel *kernel) const;

private:
    static std::size_t counter;
    std::string _name;
    reaction_function _reaction;
    rate_function _rate;
    bool _expectConnected = false;
    std::size_t _id;
};

// ---- actions -----------------------------------------------------------------------------------
class Action {  // Action.h:58-79
public:
    virtual ~Action() = default;
    virtual void perform() = 0;
};

class TimeStepDependentAction : public Action {
public:
    explicit TimeStepDependentAction(scalar dt) : _dt(dt) {}
    scalar timeStep() const { return _dt; }
    void setTimeStep(scalar dt) { _dt = dt; }

protected:
    scalar _dt;
};

// ---- DetailedBalance (Actions.h:157-176, DetailedBalance.h:49-148) ----------------------------
enum class ReversibleType { FusionFission, ConversionConversion, EnzymaticEnzymatic };

// ReversibleReactionConfig (DetailedBalance.h:64-139): a pair of reactions that undo each other and the
// values derived from it; ids are reaction indices, the FusionFission forward reaction is the fusion
struct ReversibleReactionConfig {
    std::size_t forwardId = 0, backwardId = 0;
    std::string forwardName, backwardName;
    ReversibleType reversibleType = ReversibleType::ConversionConversion;
    std::uint8_t numberLhsTypes = 0, numberRhsTypes = 0;
    std::array<ParticleTypeId, 2> lhsTypes{{0, 0}}, rhsTypes{{0, 0}};
    std::array<std::string, 2> lhsNames, rhsNames;
    scalar microForwardRate = 0, microBackwardRate = 0, reactionRadius = 0;
    scalar lhsInteractionRadius = 0, rhsInteractionRadius = 0, lhsInteractionVolume = 0, rhsInteractionVolume = 0;
    scalar effectiveLhsInteractionVolume = 0, effectiveRhsInteractionVolume = 0;
    scalar effectiveLhsReactionVolume = 0, effectiveRhsReactionVolume = 0;
    scalar totalVolume = 0, kbt = 0, acceptancePrefactor = 1;
    scalar equilibriumConstant = 0, macroForwardRate = 0, macroBackwardRate = 0;
};

// the DetailedBalance reaction handler: the reversible pairs are searched when the action is created
// (DetailedBalance::searchReversibleReactions, Actions.cpp:432-519), which throws std::logic_error for
// ambiguous or inconsistent pairs
class DetailedBalance : public TimeStepDependentAction {
public:
    using TimeStepDependentAction::TimeStepDependentAction;
    const std::vector<ReversibleReactionConfig> &reversibleReactions() const { return _reversibleReactions; }
    std::string describe() const;  // Actions.cpp:369-430: one block per pair

protected:
    std::vector<ReversibleReactionConfig> _reversibleReactions;
};

// BreakConfig (Actions.h:190-211): per unordered pair of particle types, the energy above which a bond
// may break and the rate at which it then breaks
class BreakConfig {
public:
    struct Pair {
        ParticleTypeId type1, type2;
        scalar thresholdEnergy, rate;
    };
    void addBreakablePair(ParticleTypeId type1, ParticleTypeId type2, scalar thresholdEnergy, scalar rate);
    const std::vector<Pair> &pairs() const { return _pairs; }

private:
    std::vector<Pair> _pairs;
};

class ActionFactory {  // ActionFactory.h:93-128
public:
    virtual ~ActionFactory() = default;
    virtual std::unique_ptr<Action> addParticles(const std::vector<Particle> &particles) = 0;
    virtual std::unique_ptr<TimeStepDependentAction> eulerBDIntegrator(scalar dt) = 0;
    virtual std::unique_ptr<Action> calculateForces() = 0;
    virtual std::unique_ptr<Action> createNeighborList(scalar interactionDistance) = 0;
    virtual std::unique_ptr<Action> updateNeighborList() = 0;
    virtual std::unique_ptr<Action> clearNeighborList() = 0;
    virtual std::unique_ptr<Action> evaluateCompartments() = 0;
    virtual std::unique_ptr<TimeStepDependentAction> uncontrolledApproximation(scalar dt) = 0;
    virtual std::unique_ptr<TimeStepDependentAction> gillespie(scalar dt) = 0;
    virtual std::unique_ptr<DetailedBalance> detailedBalance(scalar dt) = 0;
    // ActionFactory::breakBonds (Actions.h:218-341)
    virtual std::unique_ptr<TimeStepDependentAction> breakBonds(scalar dt, BreakConfig config) = 0;
    virtual std::unique_ptr<TimeStepDependentAction> evaluateTopologyReactions(scalar dt) = 0;
    virtual std::unique_ptr<Action> initializeKernel() = 0;
};

class StateModel {  // StateModel.h:66-117
public:
    virtual ~StateModel() = default;
    virtual void addParticle(const Particle &p) = 0;
    virtual void addParticles(const std::vector<Particle> &p) = 0;
    virtual std::vector<Particle> getParticles() const = 0;     // ordered by particle id
    virtual std::vector<Vec3> getParticlePositions() const = 0;  // aligned with getParticles()
    virtual std::vector<Vec3> getParticleForces() const = 0;     // aligned with getParticles()
    virtual void initializeNeighborList(scalar interactionDistance) = 0;
    virtual void updateNeighborList() = 0;
    virtual void clearNeighborList() = 0;
    virtual scalar energy() const = 0;
    virtual scalar time() const = 0;
    virtual void setTime(scalar t) = 0;
    // adds a topology: its particles are placed first on the device (handles 0..), the graph edges
    // and the topology registry's configuration give its bonds, angles and dihedrals
    virtual GraphTopology *addTopology(const std::string &type, const std::vector<Particle> &particles) = 0;
    virtual std::vector<GraphTopology *> getTopologies() = 0;
    virtual void configureTopologies() = 0;  // GraphTopology::configure for all topologies
    virtual std::vector<uint64_t> reactionCounts() const = 0;  // of the last reaction handler call
    // of the last reaction handler call, when the context records them (empty otherwise)
    virtual std::vector<ReactionRecord> reactionRecords() const = 0;
    virtual void clear() = 0;
    // observable inputs (the reference's observables read them from the state model): the pair virial
    // of the last force evaluation (row-major, needs Context::recordVirial), particles per type, and the
    // integer counts of HistogramAlongAxis (CPUObservables.cpp:85-136) and of the radial distribution's
    // pair loop (RadialDistribution.cpp:75-103, with the number of "from" particles).  A kernel that does
    // not provide one throws std::logic_error.
    virtual std::array<scalar, 9> virial() const;
    virtual std::vector<unsigned long> particlesPerType() const;
    virtual std::vector<uint64_t> histogramAlongAxis(const std::vector<scalar> &borders,
                                                     const std::vector<ParticleTypeId> &types, unsigned axis) const;
    virtual std::vector<uint64_t> rdfCounts(const std::vector<scalar> &borders, const std::vector<ParticleTypeId> &from,
                                            const std::vector<ParticleTypeId> &to, uint64_t *nFrom) const;
    // changes the species of the particle with this id (TopologyReactionActionFactory::createChangeParticleType
    // acts through it); a kernel that cannot throws std::logic_error
    virtual void changeParticleType(ParticleId id, ParticleTypeId type);
    // moves the topology particle with this id (createChangeParticlePosition acts through it)
    virtual void changeParticlePosition(ParticleId id, const Vec3 &pos);
    // a new particle of a topology species; returns its id (createAppendParticle acts through it)
    virtual ParticleId appendTopologyParticle(ParticleTypeId type, const Vec3 &pos);
};

class Kernel;

// ---- topology actions (TopologyActionFactory.h; TopologyReactionActionFactory.h:55-101) --------
// The bonded potentials of topologies are evaluated inside the kernel's force action (the device sums
// every topology's bonds, angles and dihedrals with the pair forces), so the potential half of the
// reference factory has no separate actions here; the reaction half edits a topology: its graph, its
// type and its particles' species.  The graph edits take effect with the next configureTopologies().
class TopologyReactionAction : public Action {
public:
    explicit TopologyReactionAction(GraphTopology *topology) : _topology(topology) {
        if (!topology) throw std::invalid_argument("topology reaction action without a topology");
    }
    virtual void undo() = 0;  // TopologyReactionAction::undo (reverts the change when a reaction is rolled back)

protected:
    GraphTopology *_topology;
};

class TopologyActionFactory {
public:
    explicit TopologyActionFactory(StateModel *state) : _state(state) {}
    using ActionPtr = std::unique_ptr<TopologyReactionAction>;
    // vertex indices are positions in GraphTopology::particleIndices()
    ActionPtr createChangeParticleType(GraphTopology *topology, std::size_t vertex, ParticleTypeId typeTo) const;
    ActionPtr createAddEdge(GraphTopology *topology, std::size_t v1, std::size_t v2) const;
    ActionPtr createRemoveEdge(GraphTopology *topology, std::size_t v1, std::size_t v2) const;
    ActionPtr createChangeTopologyType(GraphTopology *topology, const std::string &typeTo) const;
    // TopologyReactionActions ChangeParticlePosition / AppendParticle: the particle of `vertex` moves to pos; a new
    // particle of `type` at pos joins the topology, bonded to the vertices `neighbors`
    ActionPtr createChangeParticlePosition(GraphTopology *topology, std::size_t vertex, const Vec3 &pos) const;
    ActionPtr createAppendParticle(GraphTopology *topology, const std::vector<std::size_t> &neighbors,
                                   ParticleTypeId type, const Vec3 &pos) const;

private:
    StateModel *_state;
};

// ---- observables (ObservableFactory.h:75-178, Observable.h:85-205) -----------------------------
namespace observables {

using Stride = std::size_t;
using TimeStep = std::size_t;
using Matrix33 = std::array<scalar, 9>;  // row-major

// An observable is evaluated at time step t when the stride divides t (stride 0: every step) and it
// was not evaluated at t already (Observable.h:144-147); after each evaluation its callback receives
// the result.
class ObservableBase {
public:
    ObservableBase(Kernel *kernel, Stride stride) : kernel(kernel), _stride(stride) {}
    virtual ~ObservableBase() = default;
    ObservableBase(const ObservableBase &) = delete;
    ObservableBase &operator=(const ObservableBase &) = delete;
    Stride stride() const { return _stride; }
    TimeStep currentTimeStep() const { return t_current; }
    bool shouldEvaluate(TimeStep t) const { return (t_current != t || firstCall) && (_stride == 0 || t % _stride == 0); }
    void call(TimeStep t) {
        if (!shouldEvaluate(t)) return;
        firstCall = false;
        t_current = t;
        evaluate();
        fire();
    }
    virtual void evaluate() = 0;
    virtual std::string type() const = 0;

protected:
    virtual void fire() = 0;
    Kernel *kernel;
    Stride _stride;
    TimeStep t_current = 0;
    bool firstCall = true;
};

template <class Result>
class Observable : public ObservableBase {
public:
    using result_type = Result;
    using callback_type = std::function<void(const Result &)>;
    Observable(Kernel *kernel, Stride stride) : ObservableBase(kernel, stride) {}
    const Result &getResult() const { return result; }
    callback_type &callback() { return _callback; }

protected:
    void fire() override {
        if (_callback) _callback(result);
    }
    Result result{};
    callback_type _callback;
};

// The observables a kernel provides; each evaluate() reads the device state through the C ABI.
struct Energy : Observable<scalar> {
    using Observable::Observable;
    void evaluate() override;
    std::string type() const override { return "Energy"; }
};
struct Virial : Observable<Matrix33> {  // needs Context::recordVirial, which the factory switches on
    using Observable::Observable;
    void evaluate() override;
    std::string type() const override { return "Virial"; }
};
struct NParticles : Observable<std::vector<unsigned long>> {  // total (no types) or per listed type
    NParticles(Kernel *k, Stride s, std::vector<ParticleTypeId> types) : Observable(k, s), typesToCount(std::move(types)) {}
    void evaluate() override;
    std::string type() const override { return "NParticles"; }
    std::vector<ParticleTypeId> typesToCount;
};
struct Forces : Observable<std::vector<Vec3>> {  // of the listed types (all when none), particle order
    Forces(Kernel *k, Stride s, std::vector<ParticleTypeId> types) : Observable(k, s), typesToCount(std::move(types)) {}
    void evaluate() override;
    std::string type() const override { return "Forces"; }
    std::vector<ParticleTypeId> typesToCount;
};
struct Positions : Observable<std::vector<Vec3>> {
    Positions(Kernel *k, Stride s, std::vector<ParticleTypeId> types) : Observable(k, s), typesToCount(std::move(types)) {}
    void evaluate() override;
    std::string type() const override { return "Positions"; }
    std::vector<ParticleTypeId> typesToCount;
};
struct HistogramAlongAxis : Observable<std::vector<scalar>> {
    HistogramAlongAxis(Kernel *k, Stride s, std::vector<scalar> borders, std::vector<ParticleTypeId> types,
                       unsigned axis)
        : Observable(k, s), binBorders(std::move(borders)), typesToCount(std::move(types)), axis(axis) {
        result.assign(binBorders.size() > 1 ? binBorders.size() - 1 : 0, 0.0);
    }
    void evaluate() override;
    std::string type() const override { return "HistogramAlongAxis"; }
    std::vector<scalar> binBorders;
    std::vector<ParticleTypeId> typesToCount;
    unsigned axis;
};
// (bin centers, distribution) as RadialDistribution.cpp:105-123 normalises the pair counts
struct RadialDistribution : Observable<std::pair<std::vector<scalar>, std::vector<scalar>>> {
    RadialDistribution(Kernel *k, Stride s, std::vector<scalar> borders, std::vector<ParticleTypeId> from,
                       std::vector<ParticleTypeId> to, scalar particleToDensity);
    void evaluate() override;
    std::string type() const override { return "RadialDistribution"; }
    std::vector<scalar> binBorders;
    std::vector<ParticleTypeId> typeCountFrom, typeCountTo;
    scalar particleToDensity;
};
// (types, ids, positions) of every particle (Particles.cpp:62-92)
struct Particles : Observable<std::tuple<std::vector<ParticleTypeId>, std::vector<ParticleId>, std::vector<Vec3>>> {
    using Observable::Observable;
    void evaluate() override;
    std::string type() const override { return "Particles"; }
};
struct Reactions : Observable<std::vector<ReactionRecord>> {  // needs recordReactionsWithPositions
    using Observable::Observable;
    void evaluate() override;
    std::string type() const override { return "Reactions"; }
};
struct ReactionCounts : Observable<std::map<std::size_t, std::size_t>> {  // reaction index -> count, last step
    using Observable::Observable;
    void evaluate() override;
    std::string type() const override { return "ReactionCounts"; }
};

class ObservableFactory {
public:
    explicit ObservableFactory(Kernel *kernel) : kernel(kernel) {}
    template <class T>
    using ObsCallback = std::function<void(const typename T::result_type &)>;
    std::unique_ptr<Energy> energy(Stride stride, ObsCallback<Energy> callback = {}) const;
    std::unique_ptr<Virial> virial(Stride stride, ObsCallback<Virial> callback = {}) const;
    std::unique_ptr<HistogramAlongAxis> histogramAlongAxis(Stride stride, std::vector<scalar> binBorders,
                                                           std::vector<std::string> typesToCount, unsigned axis,
                                                           ObsCallback<HistogramAlongAxis> callback = {}) const;
    std::unique_ptr<NParticles> nParticles(Stride stride, std::vector<std::string> typesToCount = {},
                                           ObsCallback<NParticles> callback = {}) const;
    std::unique_ptr<Forces> forces(Stride stride, std::vector<std::string> typesToCount = {},
                                   ObsCallback<Forces> callback = {}) const;
    std::unique_ptr<Positions> positions(Stride stride, std::vector<std::string> typesToCount = {},
                                         ObsCallback<Positions> callback = {}) const;
    std::unique_ptr<RadialDistribution> radialDistribution(Stride stride, std::vector<scalar> binBorders,
                                                           std::vector<std::string> typeCountFrom,
                                                           std::vector<std::string> typeCountTo,
                                                           scalar particleToDensity,
                                                           ObsCallback<RadialDistribution> callback = {}) const;
    std::unique_ptr<Particles> particles(Stride stride, ObsCallback<Particles> callback = {}) const;
    std::unique_ptr<Reactions> reactions(Stride stride, ObsCallback<Reactions> callback = {}) const;
    std::unique_ptr<ReactionCounts> reactionCounts(Stride stride, ObsCallback<ReactionCounts> callback = {}) const;

private:
    std::vector<ParticleTypeId> ids(const std::vector<std::string> &names) const;
    Kernel *kernel;
};

}  // namespace observables

class Kernel {  // Kernel.h:71-235
public:
    explicit Kernel(std::string name) : _name(std::move(name)) {}
    virtual ~Kernel() = default;
    Kernel(const Kernel &) = delete;
    Kernel &operator=(const Kernel &) = delete;
    const std::string &name() const { return _name; }
    Context &context() { return _context; }
    const Context &context() const { return _context; }
    virtual StateModel &stateModel() = 0;
    virtual const StateModel &stateModel() const = 0;
    virtual ActionFactory &actions() = 0;
    virtual bool supportsGillespie() const { return false; }
    virtual void initialize() {}
    // optional: `steps` steps of the SimulationLoop action sequence (integrate, neighbour list,
    // reactions, neighbour list, forces) run on the device without returning to the host in
    // between; false when the kernel has no such path
    virtual bool runFused(const std::string &scheduler, scalar dt, std::size_t steps) {
        (void)scheduler, (void)dt, (void)steps;
        return false;
    }
    // Kernel::observe (Kernel.h:193-199): the factory of this kernel's observables
    observables::ObservableFactory &observe() { return _observe; }
    // Kernel::connectObservable / evaluateObservables: the loop evaluates the connected observables at
    // the start and after every step (each decides by its stride)
    void connectObservable(observables::ObservableBase *o);
    void disconnectObservable(observables::ObservableBase *o);
    void evaluateObservables(observables::TimeStep t);
    const std::vector<observables::ObservableBase *> &connectedObservables() const { return _observables; }
    // Kernel::getTopologyActionFactory (Kernel.h:163-165): nullptr when the kernel has no topologies
    virtual TopologyActionFactory *getTopologyActionFactory() { return nullptr; }

protected:
    std::string _name;
    Context _context;
    observables::ObservableFactory _observe{this};
    std::vector<observables::ObservableBase *> _observables;
};

}  // namespace model

namespace plugin {

// KernelProvider.cpp:68-209: built-in factories (route A) and dlopen'ed plugin libraries exporting
// extern "C" const char* name() and extern "C" readdy_amd::model::Kernel* createKernel() (route B)
class KernelProvider {
public:
    using factory = std::function<model::Kernel *()>;
    struct KernelDeleter {
        std::shared_ptr<void> library;  // keeps the plugin mapped while a kernel lives
        void operator()(model::Kernel *k) const;
    };
    using kernel_ptr = std::unique_ptr<model::Kernel, KernelDeleter>;

    static KernelProvider &getInstance();
    void add(const std::string &name, factory f);
    void loadKernelsFromDirectory(const std::string &directory);
    void loadKernelLibrary(const std::string &path);
    kernel_ptr create(const std::string &name) const;
    std::vector<std::string> availableKernels() const;
    static bool isSharedLibrary(const std::string &path);

private:
    struct Entry {
        factory make;
        std::shared_ptr<void> library;
    };
    std::map<std::string, Entry> _factories;
};

}  // namespace plugin

namespace api {

namespace observables = model::observables;

// SimulationLoop.h:256-306: integrate, neighbour list, reactions, neighbour list, forces per step
class SimulationLoop {
public:
    SimulationLoop(model::Kernel *kernel, scalar timeStep);
    // "UncontrolledApproximation" | "Gillespie" | "DetailedBalance" (ActionFactory.h:80-91)
    void useReactionScheduler(const std::string &name);
    void evaluateForces(bool v) { _forces = v; }
    void evaluateReactions(bool v) { _reactions = v; }
    void evaluateTopologyReactions(bool v) { _topologyReactions = v; }  // SimulationLoop.h:185-190
    // called after every step with the step index (observables); none registered: the loop hands the
    // whole run to the device-fused path (rdy_run), which performs the same actions in the same order
    void addCallback(std::function<void(std::size_t)> cb) { _callbacks.push_back(std::move(cb)); }
    void run(std::size_t steps);
    scalar timeStep() const { return _dt; }
    std::size_t currentStep() const { return _t; }  // time steps taken by this loop so far

private:
    std::size_t _t = 0;
    model::Kernel *_kernel;
    scalar _dt;
    std::string _scheduler = "Gillespie";  // the reference default when supported (SimulationLoop.h:104-109)
    bool _forces = true, _reactions = true, _topologyReactions = true;
    std::vector<std::function<void(std::size_t)>> _callbacks;
};

}  // namespace api

// thrown for every failing device call; carries the engine's message
class DeviceError : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

}  // namespace readdy_amd
