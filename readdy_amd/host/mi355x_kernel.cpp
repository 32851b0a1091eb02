// The "MI355X" kernel plugin (libreaddy_amd_plugin.so): a readdy_amd::model::Kernel whose state
// model and actions drive the HIP engine through its C ABI (include/readdy_amd.h).
//
// Plugin entry points (the reference's route for kernel libraries, KernelProvider.cpp:135-153,
// CPUKernel.cpp:89-95): extern "C" const char* name() and extern "C" Kernel* createKernel().
//
// Particle ids are the engine's replica-local handles: particles uploaded in one batch get handles
// 0..n-1 (topology particles first, in topology order), reaction products fresh handles
// (Particle::nextId, Particle.h:112-114).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/readdy_amd.h"
#include "readdy_host.h"

namespace readdy_amd {
namespace kernel {

using namespace model;

static void ok(int rc, const char *what) {
    if (rc != RDY_OK) {
        const std::string msg = std::string(what) + ": " + rdy_last_error();
        if (rc == RDY_ERR_INVALID) throw std::invalid_argument(msg);
        if (rc == RDY_ERR_STATE) throw std::logic_error(msg);
        throw DeviceError(msg);
    }
}

class MI355XKernel;

class MI355XStateModel : public StateModel {
public:
    explicit MI355XStateModel(MI355XKernel *k) : _k(k) {}
    void addParticle(const Particle &p) override { addParticles({p}); }
    void addParticles(const std::vector<Particle> &ps) override;
    std::vector<Particle> getParticles() const override;
    std::vector<Vec3> getParticlePositions() const override;
    std::vector<Vec3> getParticleForces() const override;
    void initializeNeighborList(scalar) override;
    void updateNeighborList() override;
    void clearNeighborList() override {}
    scalar energy() const override;
    scalar time() const override { return _time; }
    void setTime(scalar t) override { _time = t; }
    GraphTopology *addTopology(const std::string &type, const std::vector<Particle> &particles) override;
    std::vector<GraphTopology *> getTopologies() override {
        syncTopologies();
        std::vector<GraphTopology *> v;
        for (auto &t : _topologies) v.push_back(t.get());
        return v;
    }
    void configureTopologies() override;
    std::vector<uint64_t> reactionCounts() const override;
    std::vector<ReactionRecord> reactionRecords() const override;
    void clear() override;
    std::array<scalar, 9> virial() const override;
    std::vector<unsigned long> particlesPerType() const override;
    std::vector<uint64_t> histogramAlongAxis(const std::vector<scalar> &borders, const std::vector<ParticleTypeId> &types,
                                             unsigned axis) const override;
    std::vector<uint64_t> rdfCounts(const std::vector<scalar> &borders, const std::vector<ParticleTypeId> &from,
                                    const std::vector<ParticleTypeId> &to, uint64_t *nFrom) const override;

    // uploads the pending particles (topology particles first) if the device copy is stale
    void flush();
    // BreakBonds: the device picks the edges that break; their topologies lose the edges, split into
    // their connected components and the bonded terms are configured again
    void breakBonds(scalar dt, const BreakConfig &config);
    // EvaluateTopologyReactions on the device; the topologies' types and their particles' species are read
    // back, and their bonded terms set again when a reaction changed a species they depend on
    void evaluateTopologyReactions(scalar dt);
    // whether a spatial topology reaction changes a species that the bonded configuration names: the
    // device loop keeps the bonded terms as they are, so such a context runs step by step
    bool reactionsChangeBondedTerms() const;
    // after a device loop: the topology types and species are read back when next asked for
    void markTopologiesStale();
    // TopologyReactionActionFactory::createChangeParticleType on the device (topology particles only)
    void changeParticleType(ParticleId id, ParticleTypeId type) override;
    void changeParticlePosition(ParticleId id, const Vec3 &pos) override;
    ParticleId appendTopologyParticle(ParticleTypeId type, const Vec3 &pos) override;

private:
    void configureSpatialReactions();
    void uploadTopologyTypes();
    void syncTopologies();
    void structuralReactions(scalar dt, const std::vector<uint8_t> &reacted);
    void spatialReactions(scalar dt);
    std::vector<uint8_t> _spatialReacted;  // per topology (after merges): changed by the last spatial pass
    MI355XKernel *_k;
    scalar _time = 0;
    std::vector<Particle> _topoParticles, _freeParticles;  // pending host copies (before the first flush)
    std::vector<std::unique_ptr<GraphTopology>> _topologies;
    bool _dirty = false, _uploaded = false, _topoDirty = false;
    bool _spatialSet = false, _topoStale = false;
    uint64_t _structuralCalls = 0;  // keys the host draws of the structural reactions
};

// the context as the engine's C ABI describes it (arrays owned here)
struct ContextDesc {
    std::vector<double> D;
    std::vector<uint8_t> flavor;
    std::vector<rdy_pair_potential> pots;
    std::vector<rdy_reaction> rxs;
    rdy_context_desc cd{};
    explicit ContextDesc(const Context &c) {
        const auto &types = c.particleTypes().types();
        if (types.empty()) throw std::logic_error("the context has no particle types");
        for (const auto &t : types) {
            D.push_back(t.diffusionConstant);
            flavor.push_back(t.topologyFlavor ? 1 : 0);
        }
        for (const auto &p : c.potentials().potentialsOrder2()) {
            rdy_pair_potential q{};
            q.kind = p.kind;
            q.type1 = p.type1;
            q.type2 = p.type2;
            for (int i = 0; i < 8; ++i) q.params[i] = p.params[i];
            pots.push_back(q);
        }
        for (const auto &r : c.reactions().reactions()) {
            rdy_reaction q{};
            switch (r.type) {
            case ReactionType::Conversion: q.kind = RDY_RX_CONVERSION; break;
            case ReactionType::Fusion: q.kind = RDY_RX_FUSION; break;
            case ReactionType::Fission: q.kind = RDY_RX_FISSION; break;
            case ReactionType::Enzymatic: q.kind = RDY_RX_ENZYMATIC; break;
            case ReactionType::Decay: q.kind = RDY_RX_DECAY; break;
            }
            for (std::size_t i = 0; i < r.educts.size() && i < 2; ++i) q.educts[i] = r.educts[i];
            for (std::size_t i = 0; i < r.products.size() && i < 2; ++i) q.products[i] = r.products[i];
            q.rate = r.rate;
            q.educt_distance = r.eductDistance;
            q.product_distance = r.productDistance;
            q.weight1 = r.weight1;
            q.weight2 = r.weight2;
            rxs.push_back(q);
        }
        cd.n_types = (uint32_t)types.size();
        cd.kbt = c.kBT();
        cd.diffusion = D.data();
        cd.topology_flavor = flavor.data();
        cd.n_potentials = (uint32_t)pots.size();
        cd.potentials = pots.data();
        cd.n_reactions = (uint32_t)rxs.size();
        cd.reactions = rxs.data();
    }
};

class MI355XKernel : public Kernel {
public:
    MI355XKernel() : Kernel("MI355X"), _state(this), _actions(this) {
        // the reference passes no configuration to third-party kernels (KernelConfiguration.cpp:91-98),
        // so the Philox key comes from the environment
        if (const char *s = std::getenv("READDY_AMD_SEED")) _seed = std::strtoull(s, nullptr, 0);
    }
    ~MI355XKernel() override {
        if (_engine) rdy_engine_destroy(_engine);
    }
    StateModel &stateModel() override { return _state; }
    const StateModel &stateModel() const override { return _state; }
    ActionFactory &actions() override { return _actions; }
    bool supportsGillespie() const override { return true; }
    void initialize() override { engine(); }
    bool runFused(const std::string &scheduler, scalar dt, std::size_t steps) override {
        engine();
        _state.flush();
        if (_state.reactionsChangeBondedTerms()) return false;
        if (scheduler == "DetailedBalance") return false;  // a separate action (rdy_react only)
        if (context().topologyRegistry().hasStructuralReactions()) return false;  // host recipes
        const int h = scheduler == "Gillespie" ? RDY_HANDLER_GILLESPIE
                      : scheduler == "UncontrolledApproximation" ? RDY_HANDLER_UNCONTROLLED_APPROXIMATION
                                                                 : RDY_HANDLER_NONE;
        syncRecording();
        ok(rdy_run_graph(_engine, h, dt, steps), "run");
        _state.markTopologiesStale();
        return true;
    }

    // switches the engine's reaction records on once the context asks for them (after the first
    // upload: the engine keeps them per replica), and its virial with Context::recordVirial
    void syncRecording() {
        if (_hasReplica && context().recordReactionsWithPositions() != _recording) {
            ok(rdy_record_reactions(_engine, context().recordReactionsWithPositions() ? 1 : 0), "record reactions");
            _recording = context().recordReactionsWithPositions();
        }
        if (_engine && context().recordVirial() != _virial) {
            ok(rdy_record_virial(_engine, context().recordVirial() ? 1 : 0), "record virial");
            _virial = context().recordVirial();
        }
    }
    TopologyActionFactory *getTopologyActionFactory() override { return &_taf; }

    // the engine and its single replica, created from the context on first use
    rdy_engine *engine() {
        if (_engine) return _engine;
        const Context &c = context();
        ContextDesc desc(c);
        const rdy_context_desc &cd = desc.cd;
        rdy_engine_config cfg{};
        cfg.nl_cutoff = _nlCutoff;
        // the engine's own default covers the pair potentials and order-2 reactions; the spatial topology
        // reactions' radii widen it as Context::calculateMaxCutoff does
        if (_nlCutoff == 0 && !c.topologyRegistry().spatialReactions().empty()) cfg.nl_cutoff = c.calculateMaxCutoff();
        cfg.max_replicas = 1;
        ok(rdy_engine_create(&cd, &cfg, &_engine), "engine");
        return _engine;
    }

    // adds the single replica for n particles (box and periodicity from the context)
    void replica(uint64_t n) {
        if (_hasReplica) return;
        rdy_replica_desc d{};
        for (int a = 0; a < 3; ++a) {
            d.box[a] = context().boxSize()[a];
            d.pbc[a] = context().periodicBoundaryConditions()[a] ? 1 : 0;
        }
        d.seed = _seed;
        d.max_particles = std::max<uint64_t>(2 * n, 1024);
        ok(rdy_replica_add(engine(), &d, &_replica), "replica");
        _hasReplica = true;
    }

    uint32_t replicaId() const { return _replica; }
    bool hasReplica() const { return _hasReplica; }
    void setSeed(uint64_t s) { _seed = s; }
    uint64_t seed() const { return _seed; }
    void setNeighborListCutoff(scalar c) {
        if (_engine) throw std::logic_error("the neighbour-list cutoff is fixed once the engine exists");
        _nlCutoff = c;
    }

private:
    MI355XStateModel _state;
    TopologyActionFactory _taf{&_state};

public:
    class Actions : public ActionFactory {
    public:
        explicit Actions(MI355XKernel *k) : _k(k) {}
        std::unique_ptr<Action> addParticles(const std::vector<Particle> &ps) override;
        std::unique_ptr<TimeStepDependentAction> eulerBDIntegrator(scalar dt) override;
        std::unique_ptr<Action> calculateForces() override;
        std::unique_ptr<Action> createNeighborList(scalar d) override;
        std::unique_ptr<Action> updateNeighborList() override;
        std::unique_ptr<Action> clearNeighborList() override;
        std::unique_ptr<Action> evaluateCompartments() override;
        std::unique_ptr<TimeStepDependentAction> uncontrolledApproximation(scalar dt) override;
        std::unique_ptr<TimeStepDependentAction> gillespie(scalar dt) override;
        std::unique_ptr<DetailedBalance> detailedBalance(scalar dt) override;
        std::unique_ptr<TimeStepDependentAction> breakBonds(scalar dt, BreakConfig config) override;
        std::unique_ptr<TimeStepDependentAction> evaluateTopologyReactions(scalar dt) override;
        std::unique_ptr<Action> initializeKernel() override;

    private:
        MI355XKernel *_k;
    };

private:
    Actions _actions;
    rdy_engine *_engine = nullptr;
    uint32_t _replica = 0;
    bool _hasReplica = false;
    bool _recording = false;
    bool _virial = false;
    uint64_t _seed = 0;
    scalar _nlCutoff = 0;  // 0: Context::calculateMaxCutoff

    friend class MI355XStateModel;
    friend class MI355XDetailedBalance;
};

// ---- state model -----------------------------------------------------------------------------------
void MI355XStateModel::addParticles(const std::vector<Particle> &ps) {
    const auto &reg = _k->context().particleTypes();
    for (const auto &p : ps) {
        if (reg.infoOf(p.type()).topologyFlavor)
            throw std::invalid_argument("topology particles are added through addTopology");
        _freeParticles.push_back(p);
    }
    _dirty = true;
}

GraphTopology *MI355XStateModel::addTopology(const std::string &type, const std::vector<Particle> &particles) {
    const auto &reg = _k->context().particleTypes();
    std::vector<std::size_t> idx;
    for (const auto &p : particles) {
        if (!reg.infoOf(p.type()).topologyFlavor)
            throw std::invalid_argument("a topology holds particles of topology species only");
        idx.push_back(_topoParticles.size());
        _topoParticles.push_back(p);
    }
    _topologies.push_back(std::make_unique<GraphTopology>(type, idx));
    _dirty = _topoDirty = true;
    return _topologies.back().get();
}

void MI355XStateModel::flush() {
    if (!_dirty) return;
    // particles already on the device stay; their current state is read back and re-uploaded
    // together with the new ones (the upload assigns handles in order)
    std::vector<Particle> all;
    if (_uploaded) {
        // re-uploading renumbers handles in upload order, which would break the bonded lists
        if (!_topoParticles.empty())
            throw std::logic_error("with topologies present, particles and topologies must be added before the "
                                   "first action runs");
        all = getParticles();
    }
    std::vector<double> pos;
    std::vector<uint32_t> types;
    auto put = [&](const Particle &p) {
        pos.push_back(p.pos().x);
        pos.push_back(p.pos().y);
        pos.push_back(p.pos().z);
        types.push_back(p.type());
    };
    if (!_uploaded)
        for (const auto &p : _topoParticles) put(p);
    for (const auto &p : all) put(p);
    for (const auto &p : _freeParticles) put(p);
    _k->replica(types.size());
    // topology-particle fusion and the recipes that append particles: the new topology particles take handles
    // of a reserve after the uploaded ones (every free particle of this upload joining once, and more)
    bool fusion = _k->context().topologyRegistry().hasStructuralReactions();
    for (const auto &r : _k->context().topologyRegistry().spatialReactions()) fusion = fusion || r.mode == STRMode::TP_FUSION;
    if (fusion && !_uploaded)
        ok(rdy_reserve_topology_particles(_k->engine(), _k->replicaId(),
                                          (uint32_t)std::max<std::size_t>(1024, 2 * _freeParticles.size())),
           "topology reserve");
    ok(rdy_set_particles(_k->engine(), _k->replicaId(), types.size(), pos.data(), types.data()), "set particles");
    _freeParticles.clear();
    _dirty = false;
    _uploaded = true;
    configureSpatialReactions();
    if (_topoDirty) configureTopologies();
}

void MI355XStateModel::markTopologiesStale() {
    _topoStale = !_k->context().topologyRegistry().spatialReactions().empty();
}

// the context's spatial topology reactions, set on the engine once (TopologyRegistry.cpp:195-223)
void MI355XStateModel::configureSpatialReactions() {
    if (_spatialSet) return;
    std::vector<rdy_spatial_reaction> v;
    for (const auto &r : _k->context().topologyRegistry().spatialReactions()) {
        if (r.mode != STRMode::TP_ENZYMATIC && r.mode != STRMode::TP_FUSION && r.mode != STRMode::TT_ENZYMATIC &&
            r.mode != STRMode::TT_FUSION && !(r.mode == STRMode::TT_FUSION_ALLOW_SELF && r.minGraphDistance <= 1))
            throw std::invalid_argument("spatial topology reaction \"" + r.name +
                                        "\": the MI355X kernel performs the topology-particle kinds "
                                        "(T1(p1) + (p2) -> T2(p3) + (p4) and T1(p1) + (p2) -> T2(p3--p4)) and the "
                                        "topology-topology kinds (T1(p1) + T2(p2) -> T3(p3) + T4(p4), "
                                        "T1(p1) + T2(p2) -> T3(p3--p4), with [self=true] at a minimum graph "
                                        "distance of at most 1) only");
        const bool selfFusion = r.mode == STRMode::TT_FUSION_ALLOW_SELF;
        rdy_spatial_reaction q{};
        q.type1 = r.type1;
        q.top_type1 = (uint32_t)r.topType1;
        q.type2 = r.type2;
        q.type_to1 = r.typeTo1;
        q.top_type_to1 = (uint32_t)r.topTypeTo1;
        q.type_to2 = r.typeTo2;
        q.rate = r.rate;
        q.radius = r.radius;
        q.fusion = (r.mode == STRMode::TP_FUSION || r.mode == STRMode::TT_FUSION || selfFusion) ? 1u : 0u;
        q.self_fusion = selfFusion ? 1u : 0u;
        if (r.mode == STRMode::TT_ENZYMATIC || r.mode == STRMode::TT_FUSION || selfFusion) {
            q.topology_topology = 1u;
            q.top_type2 = (uint32_t)r.topType2;
            q.top_type_to2 = (uint32_t)r.topTypeTo2;
        }
        v.push_back(q);
    }
    if (!v.empty()) ok(rdy_set_spatial_reactions(_k->engine(), (uint32_t)v.size(), v.data()), "spatial reactions");
    _spatialSet = true;
}

// the topologies' types and each topology particle's topology, for the device's topology reactions
void MI355XStateModel::uploadTopologyTypes() {
    const auto &reg = _k->context().topologyRegistry();
    if (reg.spatialReactions().empty() || !_uploaded) return;
    std::vector<uint32_t> types, of(_topoParticles.size(), 0xFFFFFFFFu);  // UINT32_MAX: left its topology
    for (std::size_t t = 0; t < _topologies.size(); ++t) {
        types.push_back((uint32_t)reg.idOf(_topologies[t]->type()));
        for (auto i : _topologies[t]->particleIndices()) of.at(i) = (uint32_t)t;
    }
    ok(rdy_set_topologies(_k->engine(), _k->replicaId(), (uint32_t)types.size(), types.data(), of.data()),
       "topologies");
}

void MI355XStateModel::syncTopologies() {
    if (!_topoStale) return;
    _topoStale = false;
    const auto &reg = _k->context().topologyRegistry();
    std::vector<uint32_t> types(_topologies.size() + 1);
    uint32_t n = 0;
    ok(rdy_get_topology_types(_k->engine(), _k->replicaId(), (uint32_t)types.size(), types.data(), &n),
       "topology types");
    if (n != _topologies.size()) throw std::logic_error("the device holds a different number of topologies");
    for (std::size_t t = 0; t < n; ++t) _topologies[t]->setType(reg.nameOf(types[t]));
    // the topology particles are the device handles 0 .. n_topo - 1 (they are uploaded first)
    for (const auto &p : getParticles())
        if (p.id() < _topoParticles.size()) _topoParticles[p.id()] = Particle(p.pos(), p.type(), _topoParticles[p.id()].id());
}

bool MI355XStateModel::reactionsChangeBondedTerms() const {
    const auto &reg = _k->context().topologyRegistry();
    auto named = [&](ParticleTypeId t) {
        for (const auto &kv : reg.bonds)
            if (std::get<0>(kv.first) == t || std::get<1>(kv.first) == t) return true;
        for (const auto &kv : reg.angles)
            if (std::get<0>(kv.first) == t || std::get<1>(kv.first) == t || std::get<2>(kv.first) == t) return true;
        for (const auto &kv : reg.torsions)
            if (std::get<0>(kv.first) == t || std::get<1>(kv.first) == t || std::get<2>(kv.first) == t ||
                std::get<3>(kv.first) == t)
                return true;
        return false;
    };
    for (const auto &r : reg.spatialReactions()) {
        if (r.mode == STRMode::TP_FUSION || r.mode == STRMode::TT_FUSION) return true;  // every fusion adds an edge
        if (r.typeTo1 != r.type1 && (named(r.type1) || named(r.typeTo1))) return true;
        if (r.mode == STRMode::TT_ENZYMATIC && r.typeTo2 != r.type2 && (named(r.type2) || named(r.typeTo2))) return true;
    }
    return false;
}

// CPUEvaluateTopologyReactions::perform (CPUEvaluateTopologyReactions.cpp:88-475): the spatial reactions on
// the device, then the structural reactions of the topologies they left alone
void MI355XStateModel::evaluateTopologyReactions(scalar dt) {
    flush();
    const auto &reg = _k->context().topologyRegistry();
    const auto &sr = reg.spatialReactions();
    if (!_k->hasReplica() || _topologies.empty()) return;
    if (_topoDirty) configureTopologies();
    std::vector<uint8_t> reacted(_topologies.size(), 0);
    if (!sr.empty()) {
        spatialReactions(dt);
        reacted = _spatialReacted;
    }
    if (reg.hasStructuralReactions()) structuralReactions(dt, reacted);
}

// The structural reactions (CPUEvaluateTopologyReactions.cpp:88-220, executeStructuralReaction in
// actions/Utils.h:52-80).  Each topology that no spatial reaction changed in this evaluation draws, for each
// structural reaction of its type with rate r = rate(topology) > 0, the acceptance 1 - exp(-r dt) and an
// Efraimidis-Spirakis key -ln(u) / r; the accepted reaction with the smallest key is performed (one topology
// reaction per topology and evaluation, as performEvents' eventsDependent rule gives), topologies in key
// order.  A topology that fell apart is replaced by its components (children of one normal-flavour particle
// leave the topologies), one left as a single normal-flavour particle leaves them as well.  The reference
// draws structural and spatial events from one performEvents list; here the spatial ones go first, which
// differs only when both kinds are accepted for one topology in one step.
void MI355XStateModel::structuralReactions(scalar dt, const std::vector<uint8_t> &reacted) {
    const auto &reg = _k->context().topologyRegistry();
    std::mt19937_64 gen(_k->seed() ^ (0x9E3779B97F4A7C15ull * ++_structuralCalls));
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    struct Ev {
        double key;
        std::size_t top, rx;
    };
    std::vector<Ev> events;
    for (std::size_t ti = 0; ti < _topologies.size(); ++ti) {
        GraphTopology &t = *_topologies[ti];
        const auto &rs = reg.structuralReactionsOf(t.type());
        if (rs.empty() || (ti < reacted.size() && reacted[ti])) continue;
        t.updateReactionRates(rs);
        Ev best{0.0, ti, SIZE_MAX};
        for (std::size_t j = 0; j < rs.size(); ++j) {
            const double r = t.rates()[j];
            const double u = uni(gen), v = uni(gen);
            if (!(r > 0) || !(u < 1.0 - std::exp(-r * dt))) continue;
            const double key = -std::log(v + 0x1p-54) / r;
            if (best.rx == SIZE_MAX || key < best.key) best = {key, ti, j};
        }
        if (best.rx != SIZE_MAX) events.push_back(best);
    }
    if (events.empty()) return;
    std::sort(events.begin(), events.end(), [](const Ev &a, const Ev &b) { return a.key < b.key; });
    std::vector<char> gone(_topologies.size(), 0);
    std::vector<std::unique_ptr<GraphTopology>> appended;
    for (const Ev &ev : events) {
        GraphTopology &t = *_topologies[ev.top];
        const auto &rx = reg.structuralReactionsOf(t.type()).at(ev.rx);
        auto children = rx.execute(t, _k);
        if (!children.empty()) {
            gone[ev.top] = 1;
            for (auto &c : children)
                if (!c.isNormalParticle(*_k)) appended.push_back(std::make_unique<GraphTopology>(std::move(c)));
        } else if (t.isNormalParticle(*_k)) {
            gone[ev.top] = 1;
        }
    }
    std::vector<std::unique_ptr<GraphTopology>> next;
    for (std::size_t ti = 0; ti < _topologies.size(); ++ti)
        if (!gone[ti]) next.push_back(std::move(_topologies[ti]));
    for (auto &t : appended) next.push_back(std::move(t));
    _topologies = std::move(next);
    _topoDirty = true;
    configureTopologies();
}

void MI355XStateModel::changeParticleType(ParticleId id, ParticleTypeId type) {
    flush();
    if (id >= _topoParticles.size())
        throw std::invalid_argument("the MI355X kernel changes the species of topology particles only");
    const uint32_t h = (uint32_t)id, t = (uint32_t)type;
    ok(rdy_set_particle_types(_k->engine(), _k->replicaId(), 1, &h, &t), "change particle type");
    syncTopologies();
    const Particle &old = _topoParticles[id];
    _topoParticles[id] = Particle(old.pos(), type, old.id());
}

// TopologyReactionActions ChangeParticlePosition: the device record moves and the particles are re-binned
void MI355XStateModel::changeParticlePosition(ParticleId id, const Vec3 &pos) {
    flush();
    if (id >= _topoParticles.size())
        throw std::invalid_argument("the MI355X kernel moves topology particles only");
    const uint32_t h = (uint32_t)id;
    const double p[3] = {pos.x, pos.y, pos.z};
    ok(rdy_set_particle_positions(_k->engine(), _k->replicaId(), 1, &h, p), "change particle position");
    const Particle &old = _topoParticles[id];
    _topoParticles[id] = Particle(pos, old.type(), old.id());
}

// TopologyReactionActions AppendParticle: a new topology particle under the next reserve handle
ParticleId MI355XStateModel::appendTopologyParticle(ParticleTypeId type, const Vec3 &pos) {
    flush();
    if (!_k->context().particleTypes().infoOf(type).topologyFlavor)
        throw std::invalid_argument("an appended particle must be of a topology species");
    const uint32_t t = (uint32_t)type;
    const double p[3] = {pos.x, pos.y, pos.z};
    uint32_t h = 0;
    ok(rdy_append_topology_particles(_k->engine(), _k->replicaId(), 1, &t, p, &h), "append topology particle");
    if (h != _topoParticles.size()) throw std::logic_error("appended particle handle out of order");
    _topoParticles.emplace_back(pos, type, h);
    _topoDirty = true;
    return h;
}

// the spatial pass on the device, its species and topology types read back
void MI355XStateModel::spatialReactions(scalar dt) {
    const auto &sr = _k->context().topologyRegistry().spatialReactions();
    ok(rdy_evaluate_topology_reactions(_k->engine(), dt), "topology reactions");
    // the topologies this pass changed, in the device's present order (merges below remove some)
    _spatialReacted.assign(_topologies.size(), 0);
    uint32_t nr = 0;
    ok(rdy_get_topology_reacted(_k->engine(), _k->replicaId(), (uint32_t)_spatialReacted.size(), _spatialReacted.data(),
                                &nr),
       "topology reactions");
    std::vector<uint64_t> c(sr.size(), 0);
    ok(rdy_get_topology_reaction_counts(_k->engine(), _k->replicaId(), 0, c.data()), "topology reaction counts");
    bool any = false, retyped = false;
    for (std::size_t i = 0; i < sr.size(); ++i)
        if (c[i]) {
            any = true;
            retyped = retyped || sr[i].typeTo1 != sr[i].type1 ||
                      (sr[i].mode == STRMode::TT_ENZYMATIC && sr[i].typeTo2 != sr[i].type2);
        }
    if (!any) return;
    // fusions: the joining particle is the topology particle of the next reserve handle
    // (handleTopologyParticleReaction: appendParticle(idx2, idx1), CPUEvaluateTopologyReactions.cpp:369-404)
    uint32_t nf = 0;
    ok(rdy_get_topology_fusions(_k->engine(), _k->replicaId(), 0, nullptr, &nf), "topology fusions");
    if (nf) {
        std::vector<uint32_t> rows(4ull * nf);
        ok(rdy_get_topology_fusions(_k->engine(), _k->replicaId(), nf, rows.data(), &nf), "topology fusions");
        for (uint32_t i = 0; i < nf; ++i) {
            const uint32_t top = rows[4 * i], h_top = rows[4 * i + 1], h_new = rows[4 * i + 3];
            if (h_new != _topoParticles.size() || top >= _topologies.size())
                throw std::logic_error("topology fusion rows out of order");
            _topoParticles.emplace_back(0., 0., 0., 0, h_new);  // species and position from the device below
            _topologies[top]->appendParticle(h_new, h_top);
        }
        retyped = true;
    }
    _topoStale = true;
    syncTopologies();
    // topology-topology fusions: the absorbed topology's graph joins the surviving one with the new edge, and
    // the absorbed one leaves the list (appendTopology and erase, CPUEvaluateTopologyReactions.cpp:433-437)
    uint32_t nm = 0;
    ok(rdy_get_topology_merges(_k->engine(), _k->replicaId(), 0, nullptr, &nm), "topology merges");
    if (nm) {
        std::vector<uint32_t> rows(4ull * nm);
        ok(rdy_get_topology_merges(_k->engine(), _k->replicaId(), nm, rows.data(), &nm), "topology merges");
        std::vector<char> gone(_topologies.size(), 0);
        for (uint32_t i = 0; i < nm; ++i) {
            const uint32_t a = rows[4 * i], b = rows[4 * i + 1];
            if (a >= _topologies.size() || b >= _topologies.size() || gone[a] || gone[b])
                throw std::logic_error("topology merge rows out of range");
            if (a == b) {  // self-fusion: an edge between two vertices of one graph (the topology stays)
                const auto &pi = _topologies[a]->particleIndices();
                const auto v1 = std::find(pi.begin(), pi.end(), (std::size_t)rows[4 * i + 2]);
                const auto v2 = std::find(pi.begin(), pi.end(), (std::size_t)rows[4 * i + 3]);
                if (v1 == pi.end() || v2 == pi.end()) throw std::logic_error("self-fusion of particles not in the topology");
                _topologies[a]->addEdge((std::size_t)(v1 - pi.begin()), (std::size_t)(v2 - pi.begin()));
                continue;
            }
            _topologies[a]->appendTopology(*_topologies[b], rows[4 * i + 3], rows[4 * i + 2]);
            gone[b] = 1;
        }
        std::vector<std::unique_ptr<GraphTopology>> next;
        std::vector<uint8_t> flags;
        for (std::size_t t = 0; t < _topologies.size(); ++t)
            if (!gone[t]) {
                next.push_back(std::move(_topologies[t]));
                flags.push_back(t < _spatialReacted.size() ? _spatialReacted[t] : 0);
            }
        _topologies = std::move(next);
        _spatialReacted = std::move(flags);
        retyped = true;
    }
    if (retyped) {  // GraphTopology::configure after the reaction (TopologyReactionActions.h)
        _topoDirty = true;
        configureTopologies();
    }
}

// GraphTopology::configure (GraphTopology.cpp:59-143) for every topology, flattened per replica
void MI355XStateModel::configureTopologies() {
    const auto &cfg = _k->context().topologyRegistry();
    std::vector<uint32_t> b2, a3, d4, mult;
    std::vector<double> bk, bl, ak, at, dk, dp;
    for (const auto &top : _topologies) {
        const auto &pi = top->particleIndices();
        auto type = [&](std::size_t v) { return _topoParticles[pi[v]].type(); };
        top->findNTuples(
            [&](std::size_t v1, std::size_t v2) {
                auto it = cfg.bonds.find({type(v1), type(v2)});
                if (it == cfg.bonds.end())
                    throw std::invalid_argument("topology edge " + std::to_string(pi[v1]) + " -- " +
                                                std::to_string(pi[v2]) + " has no bond configured");
                for (const auto &c : it->second) {
                    b2.push_back((uint32_t)pi[v1]);
                    b2.push_back((uint32_t)pi[v2]);
                    bk.push_back(c.forceConstant);
                    bl.push_back(c.length);
                }
            },
            [&](std::size_t v1, std::size_t v2, std::size_t v3) {
                auto it = cfg.angles.find({type(v1), type(v2), type(v3)});
                if (it == cfg.angles.end()) return;
                for (const auto &c : it->second) {
                    a3.insert(a3.end(), {(uint32_t)pi[v1], (uint32_t)pi[v2], (uint32_t)pi[v3]});
                    ak.push_back(c.forceConstant);
                    at.push_back(c.equilibriumAngle);
                }
            },
            [&](std::size_t v1, std::size_t v2, std::size_t v3, std::size_t v4) {
                auto it = cfg.torsions.find({type(v1), type(v2), type(v3), type(v4)});
                if (it == cfg.torsions.end()) return;
                for (const auto &c : it->second) {
                    d4.insert(d4.end(), {(uint32_t)pi[v1], (uint32_t)pi[v2], (uint32_t)pi[v3], (uint32_t)pi[v4]});
                    dk.push_back(c.forceConstant);
                    mult.push_back(c.multiplicity);
                    dp.push_back(c.phi_0);
                }
            });
    }
    if (!_uploaded) {
        flush();  // uploads and comes back here
        return;
    }
    rdy_engine *e = _k->engine();
    const uint32_t r = _k->replicaId();
    ok(rdy_set_bonds(e, r, bk.size(), b2.data(), bk.data(), bl.data()), "bonds");
    ok(rdy_set_angles(e, r, ak.size(), a3.data(), ak.data(), at.data()), "angles");
    ok(rdy_set_dihedrals(e, r, dk.size(), d4.data(), dk.data(), mult.data(), dp.data()), "dihedrals");
    uploadTopologyTypes();
    _topoDirty = false;
}

std::vector<Particle> MI355XStateModel::getParticles() const {
    std::vector<Particle> out;
    if (!_k->hasReplica()) {
        out = _topoParticles;
        out.insert(out.end(), _freeParticles.begin(), _freeParticles.end());
        return out;
    }
    rdy_engine *e = const_cast<MI355XKernel *>(_k)->engine();
    const uint64_t n = rdy_particle_count(e, _k->replicaId());
    std::vector<double> pos(3 * n + 3);
    std::vector<uint32_t> types(n + 1), handles(n + 1);
    uint64_t got = 0;
    ok(rdy_get_particles(e, _k->replicaId(), n, pos.data(), types.data(), handles.data(), &got), "get particles");
    out.reserve(got);
    for (uint64_t i = 0; i < got; ++i)
        out.emplace_back(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], (ParticleTypeId)types[i], handles[i]);
    // the device export lists particles in no particular order; ordering by id makes repeated
    // calls (and getParticleForces) agree
    std::sort(out.begin(), out.end(), [](const Particle &a, const Particle &b) { return a.id() < b.id(); });
    return out;
}

std::vector<Vec3> MI355XStateModel::getParticlePositions() const {
    std::vector<Vec3> v;
    for (const auto &p : getParticles()) v.push_back(p.pos());
    return v;
}

// forces of the last force evaluation, ordered by particle id like getParticles()
std::vector<Vec3> MI355XStateModel::getParticleForces() const {
    std::vector<Vec3> v;
    if (!_k->hasReplica()) return v;
    rdy_engine *e = const_cast<MI355XKernel *>(_k)->engine();
    const uint64_t n = rdy_particle_count(e, _k->replicaId());
    std::vector<double> f(3 * n + 3);
    std::vector<uint32_t> handles(n + 1);
    uint64_t got = 0;
    ok(rdy_get_forces(e, _k->replicaId(), n, f.data(), handles.data(), &got), "get forces");
    std::vector<uint64_t> order(got);
    for (uint64_t i = 0; i < got; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return handles[a] < handles[b]; });
    v.reserve(got);
    for (uint64_t i : order) v.emplace_back(f[3 * i], f[3 * i + 1], f[3 * i + 2]);
    return v;
}

void MI355XStateModel::initializeNeighborList(scalar) {
    flush();
    ok(rdy_update_neighbor_list(_k->engine()), "neighbour list");
}

void MI355XStateModel::updateNeighborList() {
    flush();
    ok(rdy_update_neighbor_list(_k->engine()), "neighbour list");
}

scalar MI355XStateModel::energy() const {
    if (!_k->hasReplica()) return 0;
    double en = 0;
    ok(rdy_get_energy(const_cast<MI355XKernel *>(_k)->engine(), _k->replicaId(), &en), "energy");
    return en;
}

std::array<scalar, 9> MI355XStateModel::virial() const {
    std::array<scalar, 9> v{};
    if (!_k->hasReplica()) return v;
    auto *k = const_cast<MI355XKernel *>(_k);
    k->syncRecording();
    ok(rdy_get_virial(k->engine(), _k->replicaId(), v.data()), "virial");
    return v;
}

std::vector<unsigned long> MI355XStateModel::particlesPerType() const {
    const auto nt = _k->context().particleTypes().nTypes();
    std::vector<unsigned long> out(nt, 0);
    if (!_k->hasReplica()) {
        for (const auto &p : getParticles()) ++out[p.type()];
        return out;
    }
    const uint32_t r = _k->replicaId();
    std::vector<uint64_t> c((std::size_t)(r + 1) * nt, 0);
    ok(rdy_count_types(const_cast<MI355XKernel *>(_k)->engine(), c.data(), r + 1), "particles per type");
    for (std::size_t t = 0; t < nt; ++t) out[t] = (unsigned long)c[(std::size_t)r * nt + t];
    return out;
}

std::vector<uint64_t> MI355XStateModel::histogramAlongAxis(const std::vector<scalar> &borders,
                                                           const std::vector<ParticleTypeId> &types,
                                                           unsigned axis) const {
    if (borders.size() < 2) throw std::invalid_argument("histogram needs at least two bin borders");
    std::vector<uint64_t> c(borders.size() - 1, 0);
    if (!_k->hasReplica()) return c;
    std::vector<uint32_t> t(types.begin(), types.end());
    ok(rdy_histogram_along_axis(const_cast<MI355XKernel *>(_k)->engine(), _k->replicaId(), borders.data(),
                                (uint32_t)borders.size(), t.data(), (uint32_t)t.size(), axis, c.data()),
       "histogram along axis");
    return c;
}

std::vector<uint64_t> MI355XStateModel::rdfCounts(const std::vector<scalar> &borders,
                                                  const std::vector<ParticleTypeId> &from,
                                                  const std::vector<ParticleTypeId> &to, uint64_t *nFrom) const {
    if (borders.size() < 2) throw std::invalid_argument("radial distribution needs at least two bin borders");
    std::vector<uint64_t> c(borders.size() - 1, 0);
    uint64_t nf = 0;
    if (_k->hasReplica()) {
        std::vector<uint32_t> f(from.begin(), from.end()), t(to.begin(), to.end());
        ok(rdy_rdf_counts(const_cast<MI355XKernel *>(_k)->engine(), _k->replicaId(), borders.data(),
                          (uint32_t)borders.size(), f.data(), (uint32_t)f.size(), t.data(), (uint32_t)t.size(),
                          c.data(), &nf),
           "radial distribution");
    }
    if (nFrom) *nFrom = nf;
    return c;
}

std::vector<uint64_t> MI355XStateModel::reactionCounts() const {
    const auto n = _k->context().reactions().reactions().size();
    std::vector<uint64_t> c(n, 0);
    if (_k->hasReplica() && n)
        ok(rdy_get_reaction_counts(const_cast<MI355XKernel *>(_k)->engine(), _k->replicaId(), c.data()),
           "reaction counts");
    return c;
}

std::vector<ReactionRecord> MI355XStateModel::reactionRecords() const {
    std::vector<ReactionRecord> out;
    if (!_k->hasReplica() || !_k->_recording) return out;
    rdy_engine *e = const_cast<MI355XKernel *>(_k)->engine();
    uint64_t n = 0;
    ok(rdy_get_reaction_records(e, _k->replicaId(), 0, nullptr, &n), "reaction records");
    std::vector<rdy_reaction_record> r(n);
    ok(rdy_get_reaction_records(e, _k->replicaId(), n, r.data(), &n), "reaction records");
    for (const auto &x : r) {
        ReactionRecord o;
        o.type = x.type;
        o.id = x.reaction;
        for (int j = 0; j < 2; ++j) {
            o.educts[j] = x.educts[j];
            o.products[j] = x.products[j];
            o.types_from[j] = (ParticleTypeId)x.types_from[j];
        }
        o.where = Vec3(x.position[0], x.position[1], x.position[2]);
        out.push_back(o);
    }
    return out;
}

void MI355XStateModel::breakBonds(scalar dt, const BreakConfig &config) {
    flush();
    if (!_k->hasReplica() || _topologies.empty()) return;
    if (_topoDirty) configureTopologies();
    std::vector<rdy_break_pair> pairs;
    for (const auto &p : config.pairs()) pairs.push_back({p.type1, p.type2, p.thresholdEnergy, p.rate});
    // every edge breaks at most once: room for all of them (a second call would draw anew)
    std::size_t n_edges = 0;
    for (const auto &t : _topologies)
        t->findNTuples([&](std::size_t, std::size_t) { ++n_edges; }, [](std::size_t, std::size_t, std::size_t) {},
                       [](std::size_t, std::size_t, std::size_t, std::size_t) {});
    std::vector<uint32_t> broken(3 * std::max<std::size_t>(n_edges, 1));
    uint64_t nb = 0;
    ok(rdy_break_bonds(_k->engine(), dt, (uint32_t)pairs.size(), pairs.data(), n_edges, broken.data(), &nb),
       "break bonds");
    if (nb == 0) return;
    if (nb > n_edges) throw std::logic_error("break bonds: more broken edges than edges");
    // topology particle index (the device handle) -> (topology, vertex)
    std::vector<std::pair<std::size_t, std::size_t>> where(_topoParticles.size(), {SIZE_MAX, 0});
    for (std::size_t ti = 0; ti < _topologies.size(); ++ti) {
        const auto &pi = _topologies[ti]->particleIndices();
        for (std::size_t v = 0; v < pi.size(); ++v) where.at(pi[v]) = {ti, v};
    }
    std::vector<char> touched(_topologies.size(), 0);
    for (uint64_t k = 0; k < nb; ++k) {
        const auto a = where.at(broken[3 * k + 1]), b = where.at(broken[3 * k + 2]);
        if (a.first == SIZE_MAX || a.first != b.first || !_topologies[a.first]->removeEdge(a.second, b.second))
            throw std::logic_error("break bonds: a broken bond is not an edge of one topology");
        touched[a.first] = 1;
    }
    // a topology that lost edges is replaced by its connected components (in place, then appended), as the
    // structural reaction's recipe does (Actions.h:262-287); every particle here is of topology flavour, so
    // single particles stay one-particle topologies
    std::vector<std::unique_ptr<GraphTopology>> next, appended;
    for (std::size_t ti = 0; ti < _topologies.size(); ++ti) {
        if (!touched[ti]) {
            next.push_back(std::move(_topologies[ti]));
            continue;
        }
        auto parts = _topologies[ti]->connectedComponents();
        for (std::size_t c = 0; c < parts.size(); ++c)
            (c == 0 ? next : appended).push_back(std::make_unique<GraphTopology>(std::move(parts[c])));
    }
    for (auto &t : appended) next.push_back(std::move(t));
    _topologies = std::move(next);
    _topoDirty = true;
    configureTopologies();
}

void MI355XStateModel::clear() {
    if (_k->hasReplica()) ok(rdy_set_particles(_k->engine(), _k->replicaId(), 0, nullptr, nullptr), "clear");
    _topoParticles.clear();
    _freeParticles.clear();
    _topologies.clear();
    _dirty = _topoDirty = _topoStale = false;
    _time = 0;
}

// ---- actions ---------------------------------------------------------------------------------------
namespace {

template <class F>
class FnAction : public Action {
public:
    explicit FnAction(F f) : _f(std::move(f)) {}
    void perform() override { _f(); }

private:
    F _f;
};

template <class F>
class FnStepAction : public TimeStepDependentAction {
public:
    FnStepAction(scalar dt, F f) : TimeStepDependentAction(dt), _f(std::move(f)) {}
    void perform() override { _f(_dt); }

private:
    F _f;
};

template <class F>
std::unique_ptr<Action> action(F f) {
    return std::make_unique<FnAction<F>>(std::move(f));
}

template <class F>
std::unique_ptr<TimeStepDependentAction> step_action(scalar dt, F f) {
    return std::make_unique<FnStepAction<F>>(dt, std::move(f));
}

}  // namespace

std::unique_ptr<Action> MI355XKernel::Actions::addParticles(const std::vector<Particle> &ps) {
    return action([k = _k, ps] { k->stateModel().addParticles(ps); });
}

// CPUEulerBDIntegrator::perform (CPUEulerBDIntegrator.cpp:51-103), Philox normal stream
std::unique_ptr<TimeStepDependentAction> MI355XKernel::Actions::eulerBDIntegrator(scalar dt) {
    return step_action(dt, [k = _k](scalar h) {
        k->_state.flush();
        ok(rdy_integrate(k->engine(), h, nullptr, 0), "integrate");
    });
}

// CPUCalculateForces::perform (CPUCalculateForces.cpp:46-212)
std::unique_ptr<Action> MI355XKernel::Actions::calculateForces() {
    return action([k = _k] {
        k->_state.flush();
        ok(rdy_calculate_forces(k->engine()), "forces");
    });
}

std::unique_ptr<Action> MI355XKernel::Actions::createNeighborList(scalar d) {
    return action([k = _k, d] { k->stateModel().initializeNeighborList(d); });
}

std::unique_ptr<Action> MI355XKernel::Actions::updateNeighborList() {
    return action([k = _k] { k->stateModel().updateNeighborList(); });
}

std::unique_ptr<Action> MI355XKernel::Actions::clearNeighborList() {
    return action([] {});
}

std::unique_ptr<Action> MI355XKernel::Actions::evaluateCompartments() {
    return action([] {});  // no compartments are registered in this context
}

std::unique_ptr<TimeStepDependentAction> MI355XKernel::Actions::uncontrolledApproximation(scalar dt) {
    return step_action(dt, [k = _k](scalar h) {
        k->_state.flush();
        k->syncRecording();
        ok(rdy_react(k->engine(), RDY_HANDLER_UNCONTROLLED_APPROXIMATION, h), "uncontrolled approximation");
    });
}

std::unique_ptr<TimeStepDependentAction> MI355XKernel::Actions::gillespie(scalar dt) {
    return step_action(dt, [k = _k](scalar h) {
        k->_state.flush();
        k->syncRecording();
        ok(rdy_react(k->engine(), RDY_HANDLER_GILLESPIE, h), "gillespie");
    });
}

std::unique_ptr<TimeStepDependentAction> MI355XKernel::Actions::breakBonds(scalar dt, BreakConfig config) {
    return step_action(dt, [k = _k, config](scalar h) { k->_state.breakBonds(h, config); });
}

// SCPUDetailedBalance (kernels/singlecpu/src/actions/SCPUReactionImpls.cpp:407-813) on the device; the
// reversible pairs come from rdy_reversible_reactions when the action is created
class MI355XDetailedBalance : public DetailedBalance {
public:
    MI355XDetailedBalance(MI355XKernel *k, scalar dt) : DetailedBalance(dt), _k(k) {
        const Context &c = k->context();
        ContextDesc desc(c);
        const auto box = c.boxSize();
        const double volume = box[0] * box[1] * box[2];
        uint64_t n = 0;
        if (rdy_reversible_reactions(&desc.cd, volume, 0, nullptr, &n) != RDY_OK)
            throw std::logic_error(rdy_last_error());
        std::vector<rdy_reversible_reaction> rows(n);
        if (n && rdy_reversible_reactions(&desc.cd, volume, n, rows.data(), &n) != RDY_OK)
            throw std::logic_error(rdy_last_error());
        const auto &rx = c.reactions().reactions();
        for (const auto &r : rows) {
            ReversibleReactionConfig q;
            q.forwardId = r.forward;
            q.backwardId = r.backward;
            q.forwardName = rx.at(r.forward).name;
            q.backwardName = rx.at(r.backward).name;
            q.reversibleType = r.kind == RDY_REVERSIBLE_FUSION_FISSION          ? ReversibleType::FusionFission
                               : r.kind == RDY_REVERSIBLE_CONVERSION_CONVERSION ? ReversibleType::ConversionConversion
                                                                                : ReversibleType::EnzymaticEnzymatic;
            q.numberLhsTypes = (std::uint8_t)r.n_lhs;
            q.numberRhsTypes = (std::uint8_t)r.n_rhs;
            for (int i = 0; i < 2; ++i) {
                q.lhsTypes[i] = r.lhs[i];
                q.rhsTypes[i] = r.rhs[i];
                q.lhsNames[i] = i < (int)r.n_lhs ? c.particleTypes().nameOf(r.lhs[i]) : std::string();
                q.rhsNames[i] = i < (int)r.n_rhs ? c.particleTypes().nameOf(r.rhs[i]) : std::string();
            }
            q.microForwardRate = r.micro_forward_rate;
            q.microBackwardRate = r.micro_backward_rate;
            q.reactionRadius = r.reaction_radius;
            q.lhsInteractionRadius = r.lhs_interaction_radius;
            q.rhsInteractionRadius = r.rhs_interaction_radius;
            q.lhsInteractionVolume = r.lhs_interaction_volume;
            q.rhsInteractionVolume = r.rhs_interaction_volume;
            q.effectiveLhsInteractionVolume = r.eff_lhs_interaction_volume;
            q.effectiveRhsInteractionVolume = r.eff_rhs_interaction_volume;
            q.effectiveLhsReactionVolume = r.eff_lhs_reaction_volume;
            q.effectiveRhsReactionVolume = r.eff_rhs_reaction_volume;
            q.totalVolume = r.total_volume;
            q.kbt = r.kbt;
            q.acceptancePrefactor = r.acceptance_prefactor;
            q.equilibriumConstant = r.equilibrium_constant;
            q.macroForwardRate = r.macro_forward_rate;
            q.macroBackwardRate = r.macro_backward_rate;
            _reversibleReactions.push_back(q);
        }
    }
    void perform() override {
        _k->_state.flush();
        _k->syncRecording();
        ok(rdy_react(_k->engine(), RDY_HANDLER_DETAILED_BALANCE, _dt), "detailed balance");
    }

private:
    MI355XKernel *_k;
};

std::unique_ptr<DetailedBalance> MI355XKernel::Actions::detailedBalance(scalar dt) {
    return std::make_unique<MI355XDetailedBalance>(_k, dt);
}

// EvaluateTopologyReactions (Actions.h:343-360): the spatial topology reactions of the topology-particle
// enzymatic kind on the device; the context's structural reactions are not registered on this kernel
std::unique_ptr<TimeStepDependentAction> MI355XKernel::Actions::evaluateTopologyReactions(scalar dt) {
    return step_action(dt, [k = _k](scalar h) { k->_state.evaluateTopologyReactions(h); });
}

std::unique_ptr<Action> MI355XKernel::Actions::initializeKernel() {
    return action([k = _k] { k->initialize(); });
}

}  // namespace kernel
}  // namespace readdy_amd

// ---- plugin entry points -------------------------------------------------------------------------------
extern "C" const char *name() { return "MI355X"; }

extern "C" readdy_amd::model::Kernel *createKernel() { return new readdy_amd::kernel::MI355XKernel(); }
