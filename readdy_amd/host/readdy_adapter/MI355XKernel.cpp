// The MI355X kernel as a plugin of an installed ReaDDy: readdy::model::Kernel, StateModel, ActionFactory and
// ObservableFactory implemented over the C ABI of libreaddy_amd.so (include/readdy_amd.h).  It is compiled
// against ReaDDy's own headers (readdy/model/Kernel.h and what it includes), which need spdlog and the rest of
// ReaDDy's dependencies; see the Makefile beside this file.  It mirrors readdy_amd/host/mi355x_kernel.cpp, the
// plugin of this repository's own host core, which is built and tested here.
//
// Scope: particles, pair potentials, reactions (UncontrolledApproximation, Gillespie, DetailedBalance), the integrator, the
// neighbour list, the device-fused observables and checkpoints.  Topologies are not offered through this
// adapter (getTopologyActionFactory() returns null, so Kernel::supportsTopologies() is false and ReaDDy's API
// refuses topology calls up front); the host core of this repository carries them.
//
// Reference interfaces: Kernel (include/readdy/model/Kernel.h:30-200), StateModel (StateModel.h:66-117),
// ActionFactory (actions/ActionFactory.h:56-128), ObservableFactory (observables/ObservableFactory.h:75-178),
// the plugin entry points name() / createKernel() (kernels/cpu/src/CPUKernel.cpp:89-95,
// KernelProvider.cpp:135-153).

#include <readdy/api/Saver.h>
#include <readdy/model/Kernel.h>
#include <readdy/model/actions/Actions.h>
#include <readdy/model/observables/ObservableFactory.h>
#include <readdy/model/observables/Observables.h>
#include <readdy/model/potentials/PotentialsOrder2.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <functional>
#include <memory>
#include <map>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "readdy_amd.h"

namespace mi355x {

namespace rm = readdy::model;
using readdy::ParticleId;
using readdy::ParticleTypeId;
using readdy::ReactionId;
using readdy::scalar;
using readdy::Stride;
using readdy::TimeStep;
using readdy::Vec3;

// status -> the exception types of ReaDDy's own kernels
void ok(int rc, const char *what) {
    if (rc == RDY_OK) return;
    const std::string msg = std::string(what) + ": " + rdy_last_error();
    if (rc == RDY_ERR_INVALID) throw std::invalid_argument(msg);
    if (rc == RDY_ERR_STATE) throw std::logic_error(msg);
    throw std::runtime_error(msg);
}

// ---- potential parameters -----------------------------------------------------------------------------
// LennardJones and ScreenedElectrostatics keep their parameters protected: a pointer to member formed in a
// derived class reads them from any object of the base class.
struct LJParams : rm::potentials::LennardJones {
    static void read(const rm::potentials::LennardJones &p, double out[8]) {
        out[0] = p.*(&LJParams::m);
        out[1] = p.*(&LJParams::n);
        out[2] = p.*(&LJParams::cutoffDistance);
        out[3] = (p.*(&LJParams::shift)) ? 1.0 : 0.0;
        out[4] = p.*(&LJParams::epsilon);
        out[5] = p.*(&LJParams::sigma);
    }
};

struct SEParams : rm::potentials::ScreenedElectrostatics {
    static void read(const rm::potentials::ScreenedElectrostatics &p, double out[8]) {
        out[0] = p.*(&SEParams::electrostaticStrength);
        out[1] = p.*(&SEParams::inverseScreeningDepth);
        out[2] = p.*(&SEParams::repulsionStrength);
        out[3] = p.*(&SEParams::repulsionDistance);
        out[4] = p.*(&SEParams::exponent);
        out[5] = p.*(&SEParams::cutoff);
    }
};

// WeakInteractionPiecewiseHarmonic keeps its parameters private; describe() prints each with fmt's shortest
// round-trip form (Potentials.cpp:139-144), so reading the numbers back gives the exact doubles.
void read_wiph(const rm::potentials::WeakInteractionPiecewiseHarmonic &p, double out[8]) {
    const std::string d = p.describe();
    const char *keys[4] = {"k=", "desired distance=", "depth=", "cutoff="};
    for (int i = 0; i < 4; ++i) {
        const auto at = d.find(keys[i]);
        if (at == std::string::npos) throw std::logic_error("unexpected description of a weak interaction: " + d);
        out[i] = std::strtod(d.c_str() + at + std::char_traits<char>::length(keys[i]), nullptr);
    }
}

class MI355XKernel;

// ---- state model ----------------------------------------------------------------------------------------
class MI355XStateModel : public rm::StateModel {
public:
    explicit MI355XStateModel(MI355XKernel *k) : _k(k) {}

    std::vector<Vec3> getParticlePositions() const override {
        std::vector<Vec3> v;
        for (const auto &p : getParticles()) v.push_back(p.pos());
        return v;
    }
    std::vector<rm::Particle> getParticles() const override;
    rm::Particle getParticleForIndex(std::size_t index) const override { return getParticles().at(index); }
    ParticleTypeId getParticleType(std::size_t index) const override { return getParticles().at(index).type(); }

    // the engine keeps its cell list in step with every move; these only upload pending particles
    void initializeNeighborList(scalar) override;
    void updateNeighborList() override;
    void clearNeighborList() override {}

    void addParticle(const rm::Particle &p) override { addParticles({p}); }
    void addParticles(const std::vector<rm::Particle> &ps) override {
        _pending.insert(_pending.end(), ps.begin(), ps.end());
    }
    rm::top::GraphTopology *const addTopology(readdy::TopologyTypeId, const std::vector<rm::Particle> &) override {
        throw std::invalid_argument("the MI355X kernel's ReaDDy adapter does not hold topologies");
    }
    std::vector<rm::top::GraphTopology *> getTopologies() override { return {}; }

    void removeParticle(const rm::Particle &p) override {
        auto all = getParticles();
        const auto it = std::find_if(all.begin(), all.end(), [&](const rm::Particle &q) { return q.id() == p.id(); });
        if (it == all.end()) throw std::invalid_argument("no particle with id " + std::to_string(p.id()));
        all.erase(it);
        upload(all);
    }
    void removeAllParticles() override { upload({}); }

    scalar energy() const override;
    scalar &energy() override {
        _energy = static_cast<const MI355XStateModel *>(this)->energy();
        return _energy;
    }
    scalar time() const override { return _time; }
    void setTime(scalar t) override { _time = t; }

    // getParticles() lists live particles only, so the indices it implies are dense already
    void toDenseParticleIndices(std::vector<std::size_t>::iterator, std::vector<std::size_t>::iterator) const override {}

    void clear() override {
        removeAllParticles();
        _time = 0;
    }

    // uploads the pending particles (with the ones already on the device, whose state is read back)
    void flush() {
        if (_pending.empty()) return;
        auto all = deviceParticles();
        all.insert(all.end(), _pending.begin(), _pending.end());
        _pending.clear();
        upload(all);
    }
    // ParticleId of a device handle: uploaded particles keep theirs, reaction products draw the next ids from
    // ReaDDy's process-wide counter when they are first seen (Particle.h:69)
    ParticleId idOf(uint32_t handle) const {
        while (_ids.size() <= handle) _ids.push_back(rm::Particle(Vec3(0, 0, 0), 0).id());
        return _ids[handle];
    }

private:
    std::vector<rm::Particle> deviceParticles() const;  // ordered by handle
    void upload(const std::vector<rm::Particle> &all);

    MI355XKernel *_k;
    scalar _time = 0, _energy = 0;
    std::vector<rm::Particle> _pending;
    mutable std::vector<ParticleId> _ids;  // device handle -> ParticleId
};

// ---- kernel ---------------------------------------------------------------------------------------------
class MI355XActionFactory;
class MI355XObservableFactory;

class MI355XKernel : public rm::Kernel {
public:
    static const std::string name;
    MI355XKernel();
    ~MI355XKernel() override;
    static rm::Kernel *create() { return new MI355XKernel(); }

    void initialize() override {
        rm::Kernel::initialize();
        engine();
    }
    bool supportsGillespie() const override { return true; }
    const rm::StateModel &stateModel() const override { return _state; }
    rm::StateModel &stateModel() override { return _state; }
    const rm::actions::ActionFactory &actions() const override;
    rm::actions::ActionFactory &actions() override;
    const rm::observables::ObservableFactory &observe() const override;
    rm::observables::ObservableFactory &observe() override;
    const rm::top::TopologyActionFactory *const getTopologyActionFactory() const override { return nullptr; }
    rm::top::TopologyActionFactory *const getTopologyActionFactory() override { return nullptr; }

    MI355XStateModel &state() { return _state; }
    const std::vector<ReactionId> &reactionIds() const { return _reactionIds; }

    // the engine, built from the context on first use, and its single replica
    rdy_engine *engine() {
        if (_engine) return _engine;
        const auto &ctx = context();
        const auto &types = ctx.particleTypes();
        std::vector<double> D(types.nTypes());
        std::vector<uint8_t> flavor(types.nTypes(), 0);
        for (const auto &entry : types.typeMapping()) {
            const auto &info = types.infoOf(entry.second);
            D.at(entry.second) = info.diffusionConstant;
            flavor.at(entry.second) = info.flavor == rm::particleflavor::TOPOLOGY ? 1 : 0;
        }
        std::vector<rdy_pair_potential> pots;
        for (const auto &entry : ctx.potentials().potentialsOrder2()) {
            for (const auto *p : entry.second) {
                rdy_pair_potential q{};
                q.type1 = p->particleType1();
                q.type2 = p->particleType2();
                if (const auto *hr = dynamic_cast<const rm::potentials::HarmonicRepulsion *>(p)) {
                    q.kind = RDY_POT_HARMONIC_REPULSION;
                    q.params[0] = hr->getForceConstant();
                    q.params[1] = hr->interactionDistance();
                } else if (const auto *w = dynamic_cast<const rm::potentials::WeakInteractionPiecewiseHarmonic *>(p)) {
                    q.kind = RDY_POT_WEAK_INTERACTION;
                    read_wiph(*w, q.params);
                } else if (const auto *lj = dynamic_cast<const rm::potentials::LennardJones *>(p)) {
                    q.kind = RDY_POT_LENNARD_JONES;
                    LJParams::read(*lj, q.params);
                } else if (const auto *se = dynamic_cast<const rm::potentials::ScreenedElectrostatics *>(p)) {
                    q.kind = RDY_POT_SCREENED_ELECTROSTATICS;
                    SEParams::read(*se, q.params);
                } else {
                    throw std::invalid_argument("pair potential \"" + p->describe() +
                                                "\" is not available on the MI355X kernel");
                }
                pots.push_back(q);
            }
        }
        if (!ctx.potentials().potentialsOrder1().empty())
            throw std::invalid_argument("external (order-1) potentials are not available on the MI355X kernel");
        // reactions in ReactionId order; the engine's table index maps back through _reactionIds
        std::vector<const rm::reactions::Reaction *> rx;
        for (const auto *r : ctx.reactions().order1Flat()) rx.push_back(r);
        for (const auto *r : ctx.reactions().order2Flat()) rx.push_back(r);
        std::sort(rx.begin(), rx.end(), [](const auto *a, const auto *b) { return a->id() < b->id(); });
        std::vector<rdy_reaction> rxs;
        _reactionIds.clear();
        for (const auto *r : rx) {
            rdy_reaction q{};
            switch (r->type()) {
            case rm::reactions::ReactionType::Conversion: q.kind = RDY_RX_CONVERSION; break;
            case rm::reactions::ReactionType::Fusion: q.kind = RDY_RX_FUSION; break;
            case rm::reactions::ReactionType::Fission: q.kind = RDY_RX_FISSION; break;
            case rm::reactions::ReactionType::Enzymatic: q.kind = RDY_RX_ENZYMATIC; break;
            case rm::reactions::ReactionType::Decay: q.kind = RDY_RX_DECAY; break;
            }
            for (std::size_t i = 0; i < r->nEducts(); ++i) q.educts[i] = r->educts()[i];
            for (std::size_t i = 0; i < r->nProducts(); ++i) q.products[i] = r->products()[i];
            q.rate = r->rate();
            q.educt_distance = r->eductDistance();
            q.product_distance = r->productDistance();
            q.weight1 = r->weight1();
            q.weight2 = r->weight2();
            rxs.push_back(q);
            _reactionIds.push_back(r->id());
        }
        rdy_context_desc cd{};
        cd.n_types = (uint32_t)types.nTypes();
        cd.kbt = ctx.kBT();
        cd.diffusion = D.data();
        cd.topology_flavor = flavor.data();
        cd.n_potentials = (uint32_t)pots.size();
        cd.potentials = pots.data();
        cd.n_reactions = (uint32_t)rxs.size();
        cd.reactions = rxs.data();
        rdy_engine_config cfg{};
        cfg.nl_cutoff = ctx.calculateMaxCutoff();
        cfg.max_replicas = 1;
        ok(rdy_engine_create(&cd, &cfg, &_engine), "engine");
        return _engine;
    }

    // the replica for up to n particles (created on the first upload, grown by a new engine when needed)
    uint32_t replica(uint64_t n) {
        if (_hasReplica && n <= _capacity) return _replica;
        if (_hasReplica) throw std::logic_error("the MI355X adapter's particle capacity is fixed by its first upload");
        rdy_replica_desc d{};
        for (int a = 0; a < 3; ++a) {
            d.box[a] = context().boxSize()[a];
            d.pbc[a] = context().periodicBoundaryConditions()[a] ? 1 : 0;
        }
        // ReaDDy hands no configuration to third-party kernels (KernelConfiguration.cpp:91-98)
        const char *s = std::getenv("READDY_AMD_SEED");
        d.seed = s ? std::strtoull(s, nullptr, 0) : 0;
        _capacity = std::max<uint64_t>(4 * n, 4096);
        d.max_particles = _capacity;
        ok(rdy_replica_add(engine(), &d, &_replica), "replica");
        _hasReplica = true;
        syncRecording();
        return _replica;
    }
    bool hasReplica() const { return _hasReplica; }
    uint32_t replicaId() const { return _replica; }

    // the context's recording switches onto the engine
    void syncRecording() {
        if (!_hasReplica) return;
        const bool rec = context().recordReactionsWithPositions();
        if (rec != _recording) ok(rdy_record_reactions(engine(), rec ? 1 : 0), "record reactions");
        _recording = rec;
        const bool vir = context().recordVirial();
        if (vir != _virial) ok(rdy_record_virial(engine(), vir ? 1 : 0), "record virial");
        _virial = vir;
    }

private:
    MI355XStateModel _state;
    std::unique_ptr<MI355XActionFactory> _actions;
    std::unique_ptr<MI355XObservableFactory> _observables;
    rdy_engine *_engine = nullptr;
    uint32_t _replica = 0;
    uint64_t _capacity = 0;
    bool _hasReplica = false, _recording = false, _virial = false;
    std::vector<ReactionId> _reactionIds;
};

const std::string MI355XKernel::name = "MI355X";

// ---- state model bodies -----------------------------------------------------------------------------------
std::vector<rm::Particle> MI355XStateModel::getParticles() const {
    auto out = deviceParticles();
    out.insert(out.end(), _pending.begin(), _pending.end());
    return out;
}

std::vector<rm::Particle> MI355XStateModel::deviceParticles() const {
    std::vector<rm::Particle> out;
    if (!_k->hasReplica()) return out;
    rdy_engine *e = _k->engine();
    const uint64_t n = rdy_particle_count(e, _k->replicaId());
    std::vector<double> pos(3 * n + 3);
    std::vector<uint32_t> types(n + 1), handles(n + 1);
    uint64_t got = 0;
    ok(rdy_get_particles(e, _k->replicaId(), n, pos.data(), types.data(), handles.data(), &got), "get particles");
    std::vector<uint64_t> order(got);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return handles[a] < handles[b]; });
    out.reserve(got);
    for (uint64_t i : order)
        out.emplace_back(Vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]), (ParticleTypeId)types[i], idOf(handles[i]));
    return out;
}

void MI355XStateModel::upload(const std::vector<rm::Particle> &all) {
    std::vector<double> pos;
    std::vector<uint32_t> types;
    pos.reserve(3 * all.size());
    for (const auto &p : all) {
        pos.insert(pos.end(), {p.pos()[0], p.pos()[1], p.pos()[2]});
        types.push_back(p.type());
    }
    const uint32_t r = _k->replica(all.size());
    ok(rdy_set_particles(_k->engine(), r, all.size(), pos.data(), types.data()), "set particles");
    // the upload numbers the particles 0 .. n-1 in this order; each keeps its ParticleId
    _ids.clear();
    for (const auto &p : all) _ids.push_back(p.id());
}

void MI355XStateModel::initializeNeighborList(scalar) {
    flush();
    if (_k->hasReplica()) ok(rdy_update_neighbor_list(_k->engine()), "neighbour list");
}

void MI355XStateModel::updateNeighborList() {
    flush();
    if (_k->hasReplica()) ok(rdy_update_neighbor_list(_k->engine()), "neighbour list");
}

scalar MI355XStateModel::energy() const {
    if (!_k->hasReplica()) return 0;
    double en = 0;
    ok(rdy_get_energy(_k->engine(), _k->replicaId(), &en), "energy");
    return en;
}

// ---- actions ----------------------------------------------------------------------------------------------
class AddParticles : public rm::actions::AddParticles {
public:
    AddParticles(MI355XKernel *k, const std::vector<rm::Particle> &ps) : rm::actions::AddParticles(k, ps) {}
};

class EulerBD : public rm::actions::EulerBDIntegrator {
public:
    EulerBD(MI355XKernel *k, scalar dt) : EulerBDIntegrator(dt), _k(k) {}
    void perform() override {  // CPUEulerBDIntegrator::perform (CPUEulerBDIntegrator.cpp:51-103)
        _k->state().flush();
        ok(rdy_integrate(_k->engine(), timeStep(), nullptr, 0), "integrate");
    }

private:
    MI355XKernel *_k;
};

class CalculateForces : public rm::actions::CalculateForces {
public:
    explicit CalculateForces(MI355XKernel *k) : _k(k) {}
    void perform() override {  // CPUCalculateForces::perform (CPUCalculateForces.cpp:46-212)
        _k->state().flush();
        _k->syncRecording();
        if (_k->hasReplica()) ok(rdy_calculate_forces(_k->engine()), "forces");
    }

private:
    MI355XKernel *_k;
};

class CreateNeighborList : public rm::actions::CreateNeighborList {
public:
    CreateNeighborList(MI355XKernel *k, scalar d) : rm::actions::CreateNeighborList(d), _k(k) {}
    void perform() override { _k->state().initializeNeighborList(cutoffDistance()); }

private:
    MI355XKernel *_k;
};

class UpdateNeighborList : public rm::actions::UpdateNeighborList {
public:
    explicit UpdateNeighborList(MI355XKernel *k) : _k(k) {}
    void perform() override { _k->state().updateNeighborList(); }

private:
    MI355XKernel *_k;
};

class ClearNeighborList : public rm::actions::ClearNeighborList {
public:
    void perform() override {}
};

class EvaluateCompartments : public rm::actions::EvaluateCompartments {
public:
    explicit EvaluateCompartments(MI355XKernel *k) : _k(k) {}
    void perform() override {
        if (!_k->context().compartments().get().empty())
            throw std::invalid_argument("compartments are not available on the MI355X kernel");
    }

private:
    MI355XKernel *_k;
};

template <class Base, int HANDLER>
class React : public Base {
public:
    React(MI355XKernel *k, scalar dt) : Base(dt), _k(k) {}
    void perform() override {  // CPUUncontrolledApproximation / CPUGillespie ::perform
        _k->state().flush();
        _k->syncRecording();
        if (_k->hasReplica()) ok(rdy_react(_k->engine(), HANDLER, this->timeStep()), "reactions");
    }

private:
    MI355XKernel *_k;
};

// SCPUDetailedBalance::perform (kernels/singlecpu/src/actions/SCPUReactionImpls.cpp:407-555) on the device.
// The reversible pairs the base class lists are the reference's own search (Actions.cpp:432-519, which also
// throws for ambiguous pairs); the device pairs the same reactions in registration order, so for an
// EnzymaticEnzymatic pair the reaction that gets the acceptance prefactor is the one registered first,
// where the reference's search walks hash maps keyed by particle type.
class DetailedBalance : public rm::actions::reactions::DetailedBalance {
public:
    DetailedBalance(MI355XKernel *k, scalar dt) : rm::actions::reactions::DetailedBalance(dt), _k(k) {
        searchReversibleReactions(k->context());
    }
    void perform() override {
        _k->state().flush();
        _k->syncRecording();
        if (_k->hasReplica()) ok(rdy_react(_k->engine(), RDY_HANDLER_DETAILED_BALANCE, timeStep()), "detailed balance");
    }

private:
    MI355XKernel *_k;
};

class EvaluateTopologyReactions : public rm::actions::top::EvaluateTopologyReactions {
public:
    using rm::actions::top::EvaluateTopologyReactions::EvaluateTopologyReactions;
    void perform() override {}  // the adapter holds no topologies
};

class BreakBonds : public rm::actions::top::BreakBonds {
public:
    using rm::actions::top::BreakBonds::BreakBonds;
    void perform() override {}  // the adapter holds no topologies
};

class EvaluateObservables : public rm::actions::EvaluateObservables {
public:
    explicit EvaluateObservables(MI355XKernel *k) : _k(k) {}
    void perform(TimeStep t) override { _k->evaluateObservables(t); }

private:
    MI355XKernel *_k;
};

class MakeCheckpoint : public rm::actions::MakeCheckpoint {
public:
    MakeCheckpoint(MI355XKernel *k, const std::string &base, std::size_t maxNSaves) : _k(k), _saver(base, maxNSaves) {}
    void perform(TimeStep t) override { _saver.makeCheckpoint(_k, t); }
    std::string describe() const override { return _saver.describe(); }

private:
    MI355XKernel *_k;
    readdy::api::Saver _saver;
};

class InitializeKernel : public rm::actions::InitializeKernel {
public:
    explicit InitializeKernel(MI355XKernel *k) : _k(k) {}
    void perform() override { _k->initialize(); }

private:
    MI355XKernel *_k;
};

class MI355XActionFactory : public rm::actions::ActionFactory {
public:
    explicit MI355XActionFactory(MI355XKernel *k) : _k(k) {}
    std::unique_ptr<rm::actions::AddParticles> addParticles(const std::vector<rm::Particle> &ps) const override {
        return std::make_unique<AddParticles>(_k, ps);
    }
    std::unique_ptr<rm::actions::EulerBDIntegrator> eulerBDIntegrator(scalar dt) const override {
        return std::make_unique<EulerBD>(_k, dt);
    }
    std::unique_ptr<rm::actions::CalculateForces> calculateForces() const override {
        return std::make_unique<CalculateForces>(_k);
    }
    std::unique_ptr<rm::actions::CreateNeighborList> createNeighborList(scalar d) const override {
        return std::make_unique<CreateNeighborList>(_k, d);
    }
    std::unique_ptr<rm::actions::UpdateNeighborList> updateNeighborList() const override {
        return std::make_unique<UpdateNeighborList>(_k);
    }
    std::unique_ptr<rm::actions::ClearNeighborList> clearNeighborList() const override {
        return std::make_unique<ClearNeighborList>();
    }
    std::unique_ptr<rm::actions::EvaluateCompartments> evaluateCompartments() const override {
        return std::make_unique<EvaluateCompartments>(_k);
    }
    std::unique_ptr<rm::actions::reactions::UncontrolledApproximation> uncontrolledApproximation(scalar dt) const override {
        return std::make_unique<React<rm::actions::reactions::UncontrolledApproximation,
                                      RDY_HANDLER_UNCONTROLLED_APPROXIMATION>>(_k, dt);
    }
    std::unique_ptr<rm::actions::reactions::Gillespie> gillespie(scalar dt) const override {
        return std::make_unique<React<rm::actions::reactions::Gillespie, RDY_HANDLER_GILLESPIE>>(_k, dt);
    }
    std::unique_ptr<rm::actions::reactions::DetailedBalance> detailedBalance(scalar dt) const override {
        return std::make_unique<DetailedBalance>(_k, dt);
    }
    std::unique_ptr<rm::actions::top::EvaluateTopologyReactions> evaluateTopologyReactions(scalar dt) const override {
        return std::make_unique<EvaluateTopologyReactions>(dt);
    }
    std::unique_ptr<rm::actions::top::BreakBonds> breakBonds(scalar dt, rm::actions::top::BreakConfig cfg) const override {
        return std::make_unique<BreakBonds>(dt, std::move(cfg));
    }
    std::unique_ptr<rm::actions::EvaluateObservables> evaluateObservables() const override {
        return std::make_unique<EvaluateObservables>(_k);
    }
    std::unique_ptr<rm::actions::MakeCheckpoint> makeCheckpoint(std::string base, std::size_t maxNSaves) const override {
        return std::make_unique<MakeCheckpoint>(_k, base, maxNSaves);
    }
    std::unique_ptr<rm::actions::InitializeKernel> initializeKernel() const override {
        return std::make_unique<InitializeKernel>(_k);
    }

private:
    MI355XKernel *_k;
};

// ---- observables ------------------------------------------------------------------------------------------
// each evaluate() reads the device state; what the base classes compute themselves (Energy from
// StateModel::energy(), the file layouts) stays theirs
std::vector<ParticleTypeId> type_ids(const rm::Context &ctx, const std::vector<std::string> &names) {
    std::vector<ParticleTypeId> v;
    for (const auto &n : names) v.push_back(ctx.particleTypes().idOf(n));
    return v;
}

class Virial : public rm::observables::Virial {
public:
    Virial(MI355XKernel *k, Stride s) : rm::observables::Virial(k, s), _k(k) {}
    void evaluate() override {
        readdy::Matrix33::data_arr v{};
        if (_k->hasReplica()) {
            _k->syncRecording();
            double w[9];
            ok(rdy_get_virial(_k->engine(), _k->replicaId(), w), "virial");
            std::copy(w, w + 9, v.begin());
        }
        result = readdy::Matrix33(v);
    }

private:
    MI355XKernel *_k;
};

class Positions : public rm::observables::Positions {
public:
    Positions(MI355XKernel *k, Stride s, std::vector<std::string> types)
        : rm::observables::Positions(k, s, std::move(types)), _k(k) {}
    void evaluate() override {
        result.clear();
        for (const auto &p : _k->state().getParticles())
            if (typesToCount.empty() ||
                std::find(typesToCount.begin(), typesToCount.end(), p.type()) != typesToCount.end())
                result.push_back(p.pos());
    }

private:
    MI355XKernel *_k;
};

class Particles : public rm::observables::Particles {
public:
    Particles(MI355XKernel *k, Stride s) : rm::observables::Particles(k, s), _k(k) {}
    void evaluate() override {
        auto &[types, ids, positions] = result;
        types.clear();
        ids.clear();
        positions.clear();
        for (const auto &p : _k->state().getParticles()) {
            types.push_back(p.type());
            ids.push_back(p.id());
            positions.push_back(p.pos());
        }
    }

private:
    MI355XKernel *_k;
};

class HistogramAlongAxis : public rm::observables::HistogramAlongAxis {
public:
    HistogramAlongAxis(MI355XKernel *k, Stride s, std::vector<scalar> borders, std::vector<std::string> types,
                       unsigned axis)
        : rm::observables::HistogramAlongAxis(k, s, std::move(borders), std::move(types), axis), _k(k) {}
    void evaluate() override {  // CPUHistogramAlongAxis::evaluate (CPUObservables.cpp:85-136), on the device
        std::vector<uint64_t> c(binBorders.size() - 1, 0);
        if (_k->hasReplica()) {
            std::vector<uint32_t> t(typesToCount.begin(), typesToCount.end());
            ok(rdy_histogram_along_axis(_k->engine(), _k->replicaId(), binBorders.data(), (uint32_t)binBorders.size(),
                                        t.data(), (uint32_t)t.size(), axis, c.data()),
               "histogram along axis");
        }
        result.assign(c.begin(), c.end());
    }

private:
    MI355XKernel *_k;
};

class NParticles : public rm::observables::NParticles {
public:
    NParticles(MI355XKernel *k, Stride s, std::vector<std::string> types)
        : rm::observables::NParticles(k, s, std::move(types)), _k(k) {}
    void evaluate() override {  // CPUNParticles::evaluate: the total, or one count per listed type
        const auto nt = _k->context().particleTypes().nTypes();
        std::vector<uint64_t> c(nt, 0);
        if (_k->hasReplica()) {
            const uint32_t r = _k->replicaId();
            std::vector<uint64_t> all((std::size_t)(r + 1) * nt, 0);
            ok(rdy_count_types(_k->engine(), all.data(), r + 1), "particles per type");
            std::copy(all.begin() + (std::size_t)r * nt, all.begin() + (std::size_t)(r + 1) * nt, c.begin());
        }
        result.clear();
        if (typesToCount.empty()) {
            result.push_back((unsigned long)std::accumulate(c.begin(), c.end(), uint64_t{0}));
        } else {
            for (auto t : typesToCount) result.push_back((unsigned long)c.at(t));
        }
    }

private:
    MI355XKernel *_k;
};

class Forces : public rm::observables::Forces {
public:
    Forces(MI355XKernel *k, Stride s, std::vector<std::string> types)
        : rm::observables::Forces(k, s, std::move(types)), _k(k) {}
    void evaluate() override {  // the forces of the last force evaluation, in getParticles() order
        result.clear();
        if (!_k->hasReplica()) return;
        rdy_engine *e = _k->engine();
        const uint64_t n = rdy_particle_count(e, _k->replicaId());
        std::vector<double> f(3 * n + 3);
        std::vector<uint32_t> handles(n + 1);
        uint64_t got = 0;
        ok(rdy_get_forces(e, _k->replicaId(), n, f.data(), handles.data(), &got), "forces");
        std::vector<uint64_t> order(got);
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return handles[a] < handles[b]; });
        const auto parts = _k->state().getParticles();
        for (std::size_t i = 0; i < order.size(); ++i) {
            const auto type = parts.at(i).type();
            if (typesToCount.empty() || std::find(typesToCount.begin(), typesToCount.end(), type) != typesToCount.end()) {
                const uint64_t j = order[i];
                result.emplace_back(f[3 * j], f[3 * j + 1], f[3 * j + 2]);
            }
        }
    }

private:
    MI355XKernel *_k;
};

class Reactions : public rm::observables::Reactions {
public:
    Reactions(MI355XKernel *k, Stride s) : rm::observables::Reactions(k, s), _k(k) {}
    void evaluate() override {  // the records of the last reaction step (SCPUObservables.h:248)
        result.clear();
        if (!_k->hasReplica()) return;
        _k->syncRecording();
        uint64_t n = 0;
        ok(rdy_get_reaction_records(_k->engine(), _k->replicaId(), 0, nullptr, &n), "reaction records");
        std::vector<rdy_reaction_record> r(n);
        ok(rdy_get_reaction_records(_k->engine(), _k->replicaId(), n, r.data(), &n), "reaction records");
        auto &st = _k->state();
        for (const auto &x : r) {
            rm::reactions::ReactionRecord rec;
            rec.type = (int)x.type;
            rec.id = _k->reactionIds().at(x.reaction);
            // products per kind (ReactionUtils.h:155-225): decay 0, conversion and fusion 1, fission and
            // enzymatic 2; absent products stay 0 as in the reference, order-1 educts repeat the one educt
            const int np = x.type == RDY_RX_DECAY ? 0 : (x.type == RDY_RX_FISSION || x.type == RDY_RX_ENZYMATIC) ? 2 : 1;
            for (int j = 0; j < 2; ++j) {
                rec.educts[j] = st.idOf(x.educts[j]);
                rec.products[j] = j < np ? st.idOf(x.products[j]) : 0;
                rec.types_from[j] = (ParticleTypeId)x.types_from[j];
            }
            rec.where = Vec3(x.position[0], x.position[1], x.position[2]);
            result.push_back(rec);
        }
    }

private:
    MI355XKernel *_k;
};

class ReactionCounts : public rm::observables::ReactionCounts {
public:
    ReactionCounts(MI355XKernel *k, Stride s) : rm::observables::ReactionCounts(k, s), _k(k) {}
    void evaluate() override {  // the counts of the last reaction step, by ReactionId
        auto &[counts, spatial, structural] = result;
        counts.clear();
        spatial.clear();
        structural.clear();
        const auto &ids = _k->reactionIds();
        std::vector<uint64_t> c(ids.size(), 0);
        if (_k->hasReplica() && !ids.empty())
            ok(rdy_get_reaction_counts(_k->engine(), _k->replicaId(), c.data()), "reaction counts");
        for (std::size_t i = 0; i < ids.size(); ++i) counts[ids[i]] = c[i];
    }

private:
    MI355XKernel *_k;
};

class RadialDistribution : public rm::observables::RadialDistribution {
public:
    RadialDistribution(MI355XKernel *k, Stride s, std::vector<scalar> borders, std::vector<std::string> from,
                       std::vector<std::string> to, scalar density)
        : rm::observables::RadialDistribution(k, s, borders, from, to, density), _k(k) {}
    // the pair loop of RadialDistribution::evaluate (RadialDistribution.cpp:75-103) on the device, the
    // normalisation as :105-123
    void evaluate() override {
        const auto &b = binBorders;
        std::vector<uint64_t> c(b.size() - 1, 0);
        uint64_t nFrom = 0;
        if (_k->hasReplica()) {
            std::vector<uint32_t> f(typeCountFrom.begin(), typeCountFrom.end());
            std::vector<uint32_t> t(typeCountTo.begin(), typeCountTo.end());
            ok(rdy_rdf_counts(_k->engine(), _k->replicaId(), b.data(), (uint32_t)b.size(), f.data(), (uint32_t)f.size(),
                              t.data(), (uint32_t)t.size(), c.data(), &nFrom),
               "radial distribution");
        }
        auto &dist = std::get<1>(result);  // the bin centres were set with the borders (setBinBorders)
        dist.resize(b.size() - 1);
        for (std::size_t i = 0; i + 1 < b.size(); ++i)
            dist[i] = (scalar)c[i] / (4. / 3. * M_PI * (std::pow(b[i + 1], 3.) - std::pow(b[i], 3.)) * (scalar)nFrom *
                                      particleToDensity);
    }

private:
    MI355XKernel *_k;
};

class MI355XObservableFactory : public rm::observables::ObservableFactory {
    template <class T>
    using Cb = typename std::function<void(const typename T::result_type &)>;

public:
    explicit MI355XObservableFactory(MI355XKernel *k) : rm::observables::ObservableFactory(k), _k(k) {}

    std::unique_ptr<rm::observables::Energy> energy(Stride s, Cb<rm::observables::Energy> cb) const override {
        auto o = std::make_unique<rm::observables::Energy>(_k, s);  // StateModel::energy() is the device's
        o->setCallback(cb);
        return o;
    }
    std::unique_ptr<rm::observables::Virial> virial(Stride s, Cb<rm::observables::Virial> cb) const override {
        return make<Virial>(cb, _k, s);
    }
    std::unique_ptr<rm::observables::HistogramAlongAxis>
    histogramAlongAxis(Stride s, std::vector<scalar> borders, std::vector<std::string> types, unsigned int axis,
                       Cb<rm::observables::HistogramAlongAxis> cb) const override {
        return make<HistogramAlongAxis>(cb, _k, s, std::move(borders), std::move(types), axis);
    }
    std::unique_ptr<rm::observables::NParticles>
    nParticles(Stride s, std::vector<std::string> types, Cb<rm::observables::NParticles> cb) const override {
        return make<NParticles>(cb, _k, s, std::move(types));
    }
    std::unique_ptr<rm::observables::Forces>
    forces(Stride s, std::vector<std::string> types, Cb<rm::observables::Forces> cb) const override {
        return make<Forces>(cb, _k, s, std::move(types));
    }
    std::unique_ptr<rm::observables::Positions>
    positions(Stride s, std::vector<std::string> types, Cb<rm::observables::Positions> cb) const override {
        return make<Positions>(cb, _k, s, std::move(types));
    }
    std::unique_ptr<rm::observables::RadialDistribution>
    radialDistribution(Stride s, std::vector<scalar> borders, std::vector<std::string> from, std::vector<std::string> to,
                       scalar density, Cb<rm::observables::RadialDistribution> cb) const override {
        return make<RadialDistribution>(cb, _k, s, std::move(borders), std::move(from), std::move(to), density);
    }
    std::unique_ptr<rm::observables::Particles> particles(Stride s, Cb<rm::observables::Particles> cb) const override {
        return make<Particles>(cb, _k, s);
    }
    std::unique_ptr<rm::observables::Reactions> reactions(Stride s, Cb<rm::observables::Reactions> cb) const override {
        return make<Reactions>(cb, _k, s);
    }
    std::unique_ptr<rm::observables::ReactionCounts>
    reactionCounts(Stride s, Cb<rm::observables::ReactionCounts> cb) const override {
        return make<ReactionCounts>(cb, _k, s);
    }

private:
    template <class T, class C, class... A>
    static std::unique_ptr<T> make(const C &cb, A &&...args) {
        auto o = std::make_unique<T>(std::forward<A>(args)...);
        o->setCallback(cb);
        return o;
    }
    MI355XKernel *_k;
};

// ---- the kernel's remaining bodies --------------------------------------------------------------------------
MI355XKernel::MI355XKernel()
    : rm::Kernel(name), _state(this), _actions(std::make_unique<MI355XActionFactory>(this)),
      _observables(std::make_unique<MI355XObservableFactory>(this)) {}

MI355XKernel::~MI355XKernel() {
    if (_engine) rdy_engine_destroy(_engine);
}

const rm::actions::ActionFactory &MI355XKernel::actions() const { return *_actions; }
rm::actions::ActionFactory &MI355XKernel::actions() { return *_actions; }
const rm::observables::ObservableFactory &MI355XKernel::observe() const { return *_observables; }
rm::observables::ObservableFactory &MI355XKernel::observe() { return *_observables; }

}  // namespace mi355x

// the entry points ReaDDy's KernelProvider looks up in a plugin library (KernelProvider.cpp:135-153)
extern "C" const char *name() { return mi355x::MI355XKernel::name.c_str(); }

extern "C" readdy::model::Kernel *createKernel() { return mi355x::MI355XKernel::create(); }
