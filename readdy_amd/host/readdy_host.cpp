// Registries, topology graphs, the kernel provider and the simulation loop of the host core
// (libreaddy_host.so).  Device work happens in the kernel plugins this library loads.
#include "readdy_host.h"

#include <dirent.h>
#include <dlfcn.h>
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <tuple>

#include "../../include/readdy_amd.h"

namespace readdy_amd {
namespace model {

// ---- ParticleTypeRegistry (ParticleTypeRegistry.h:60-199) --------------------------------------
void ParticleTypeRegistry::add(const std::string &name, scalar diffusionConstant, bool topologyFlavor) {
    if (name.empty() || name.find_first_of(" \t\n") != std::string::npos)
        throw std::invalid_argument("particle type name must be non-empty and contain no whitespace");
    if (_byName.count(name)) throw std::invalid_argument("particle type \"" + name + "\" already exists");
    if (diffusionConstant < 0) throw std::invalid_argument("diffusion constant must be non-negative");
    const auto id = static_cast<ParticleTypeId>(_types.size());
    _types.push_back({name, id, diffusionConstant, topologyFlavor});
    _byName[name] = id;
}

ParticleTypeId ParticleTypeRegistry::idOf(const std::string &name) const {
    auto it = _byName.find(name);
    if (it == _byName.end()) throw std::invalid_argument("unknown particle type \"" + name + "\"");
    return it->second;
}

const std::string &ParticleTypeRegistry::nameOf(ParticleTypeId id) const { return infoOf(id).name; }

const ParticleTypeInfo &ParticleTypeRegistry::infoOf(ParticleTypeId id) const {
    if (id >= _types.size()) throw std::invalid_argument("unknown particle type id " + std::to_string(id));
    return _types[id];
}

// ---- PotentialRegistry (PotentialRegistry.h:97-160; Potentials.cpp:150-202) ----------------------
scalar PairPotentialDesc::cutoff() const {
    switch (kind) {
    case RDY_POT_HARMONIC_REPULSION: return params[1];
    case RDY_POT_WEAK_INTERACTION: return params[3];
    case RDY_POT_LENNARD_JONES: return params[2];
    default: return params[5];
    }
}

void PotentialRegistry::addHarmonicRepulsion(const std::string &t1, const std::string &t2, scalar k, scalar d) {
    if (k <= 0 || d <= 0) throw std::invalid_argument("harmonic repulsion needs positive force constant and distance");
    PairPotentialDesc p{RDY_POT_HARMONIC_REPULSION, _types->idOf(t1), _types->idOf(t2)};
    p.params = {k, d};
    _order2.push_back(p);
}

void PotentialRegistry::addWeakInteractionPiecewiseHarmonic(const std::string &t1, const std::string &t2, scalar k,
                                                            scalar d0, scalar depth, scalar cutoff) {
    if (!(cutoff > d0)) throw std::invalid_argument("weak interaction: cutoff must exceed the desired distance");
    PairPotentialDesc p{RDY_POT_WEAK_INTERACTION, _types->idOf(t1), _types->idOf(t2)};
    p.params = {k, d0, depth, cutoff};
    _order2.push_back(p);
}

void PotentialRegistry::addLennardJones(const std::string &t1, const std::string &t2, unsigned m, unsigned n,
                                        scalar cutoff, bool shift, scalar epsilon, scalar sigma) {
    if (m <= n) throw std::invalid_argument("Lennard-Jones: the repulsive exponent m must exceed n");  // Potentials.cpp:159
    PairPotentialDesc p{RDY_POT_LENNARD_JONES, _types->idOf(t1), _types->idOf(t2)};
    p.params = {(scalar)m, (scalar)n, cutoff, shift ? 1.0 : 0.0, epsilon, sigma};
    _order2.push_back(p);
}

void PotentialRegistry::addScreenedElectrostatics(const std::string &t1, const std::string &t2, scalar C,
                                                  scalar kappa, scalar D, scalar sigma, unsigned exponent,
                                                  scalar cutoff) {
    if (kappa <= 0 || sigma <= 0 || cutoff <= 0)  // Potentials.cpp:187-195
        throw std::invalid_argument("screened electrostatics: screening depth, repulsion distance and cutoff must be positive");
    PairPotentialDesc p{RDY_POT_SCREENED_ELECTROSTATICS, _types->idOf(t1), _types->idOf(t2)};
    p.params = {C, kappa, D, sigma, (scalar)exponent, cutoff};
    _order2.push_back(p);
}

// ---- ReactionRegistry (ReactionRegistry.h:106-158) ----------------------------------------------
static void check_rate(scalar rate) {
    if (!(rate >= 0)) throw std::invalid_argument("reaction rates must be non-negative");
}

void ReactionRegistry::addConversion(const std::string &name, const std::string &from, const std::string &to,
                                     scalar rate) {
    check_rate(rate);
    append({name, ReactionType::Conversion, {_types->idOf(from)}, {_types->idOf(to)}, rate});
}

void ReactionRegistry::addEnzymatic(const std::string &name, const std::string &catalyst, const std::string &from,
                                    const std::string &to, scalar rate, scalar eductDistance) {
    check_rate(rate);
    ReactionDesc r{name, ReactionType::Enzymatic, {_types->idOf(from), _types->idOf(catalyst)},
                   {_types->idOf(to), _types->idOf(catalyst)}, rate};  // Enzymatic.h:56-57
    r.eductDistance = eductDistance;
    append(r);
}

// weights normalised to sum 1 (Reaction.h:137-160)
static std::pair<scalar, scalar> normalised(scalar w1, scalar w2) {
    const scalar s = w1 + w2;
    if (!(s > 0)) throw std::invalid_argument("reaction weights must have a positive sum");
    return {w1 / s, w2 / s};
}

void ReactionRegistry::addFission(const std::string &name, const std::string &from, const std::string &to1,
                                  const std::string &to2, scalar rate, scalar productDistance, scalar w1, scalar w2) {
    check_rate(rate);
    ReactionDesc r{name, ReactionType::Fission, {_types->idOf(from)}, {_types->idOf(to1), _types->idOf(to2)}, rate};
    r.productDistance = productDistance;
    std::tie(r.weight1, r.weight2) = normalised(w1, w2);
    append(r);
}

void ReactionRegistry::addFusion(const std::string &name, const std::string &from1, const std::string &from2,
                                 const std::string &to, scalar rate, scalar eductDistance, scalar w1, scalar w2) {
    check_rate(rate);
    ReactionDesc r{name, ReactionType::Fusion, {_types->idOf(from1), _types->idOf(from2)}, {_types->idOf(to)}, rate};
    r.eductDistance = eductDistance;
    std::tie(r.weight1, r.weight2) = normalised(w1, w2);
    append(r);
}

void ReactionRegistry::addDecay(const std::string &name, const std::string &type, scalar rate) {
    check_rate(rate);
    append({name, ReactionType::Decay, {_types->idOf(type)}, {}, rate});
}

void ReactionRegistry::append(ReactionDesc r) {
    for (const auto &q : _reactions)
        if (q.name == r.name) throw std::invalid_argument("A reaction with the name " + r.name + " exists already");
    _reactions.push_back(std::move(r));
}

namespace {
std::string trimmed(const std::string &s) {
    const auto b = s.find_first_not_of(" \t\n\r"), e = s.find_last_not_of(" \t\n\r");
    return b == std::string::npos ? std::string() : s.substr(b, e - b + 1);
}

// "X +(r) Y" -> (X, Y, r); "X + Y" -> (X, Y, "") when the side carries no distance
std::tuple<std::string, std::string, std::string> split_plus(const std::string &side, std::size_t plus, bool distance) {
    const std::string first = trimmed(side.substr(0, plus));
    if (!distance) return {first, trimmed(side.substr(plus + 1)), std::string()};
    const auto open = side.find('('), close = side.find(')');
    if (open == std::string::npos || close == std::string::npos || close < open)
        throw std::invalid_argument("The side \"" + side + "\" did not contain a pair of parentheses '(*)'");
    return {first, trimmed(side.substr(close + 1)), trimmed(side.substr(open + 1, close - open - 1))};
}

scalar number(const std::string &s, const std::string &what) {
    std::size_t used = 0;
    const std::string t = trimmed(s);
    scalar v = 0;
    try {
        v = std::stod(t, &used);
    } catch (const std::exception &) {
        used = 0;
    }
    if (used == 0) throw std::invalid_argument("The " + what + " \"" + s + "\" is not a number");
    return v;
}
}  // namespace

std::size_t ReactionRegistry::add(const std::string &descriptor, scalar rate) {
    const std::string arrow = "->";
    const auto ap = descriptor.find(arrow);
    if (ap == std::string::npos)
        throw std::invalid_argument("the descriptor must contain an arrow (\"->\") to indicate lhs and rhs.");
    if (descriptor.find(arrow, ap + 1) != std::string::npos)
        throw std::invalid_argument("the descriptor must not contain more than one arrow (\"->\").");
    std::string lhs = trimmed(descriptor.substr(0, ap)), rhs = trimmed(descriptor.substr(ap + arrow.size()));
    const auto colon = lhs.find(':');
    if (colon == std::string::npos)
        throw std::invalid_argument("The descriptor did not contain a colon ':' to specify the end of the name.");
    const std::string name = trimmed(lhs.substr(0, colon));
    lhs = trimmed(lhs.substr(colon + 1));
    scalar w1 = 0.5, w2 = 0.5;
    if (const auto lb = rhs.find('['); lb != std::string::npos) {
        const auto rb = rhs.rfind(']');
        if (rb == std::string::npos || rb < lb)
            throw std::invalid_argument("The term \"" + rhs + "\" did not contain a pair of brackets '[*]'");
        const std::string inner = rhs.substr(lb + 1, rb - lb - 1);
        const auto comma = inner.find(',');
        if (comma == std::string::npos)
            throw std::invalid_argument("The term \"" + rhs + "\" did not contain a comma ','");
        w1 = number(inner.substr(0, comma), "weight");
        w2 = number(inner.substr(comma + 1), "weight");
        rhs = trimmed(rhs.substr(0, lb));
    }
    std::string e1, e2, radius;
    const auto lplus = lhs.find('+');
    const bool two_educts = lplus != std::string::npos;
    if (two_educts) std::tie(e1, e2, radius) = split_plus(lhs, lplus, true);
    else e1 = lhs;
    if (rhs.empty()) {
        if (two_educts) throw std::invalid_argument("Left hand side (\"" + lhs + "\") did not contain one educt");
        addDecay(name, e1, rate);
        return _reactions.size() - 1;
    }
    const auto rplus = rhs.find('+');
    if (rplus == std::string::npos) {
        if (two_educts) addFusion(name, e1, e2, rhs, rate, number(radius, "educt distance"), w1, w2);
        else addConversion(name, e1, rhs, rate);
        return _reactions.size() - 1;
    }
    if (!two_educts) {
        std::string p1, p2, distance;
        std::tie(p1, p2, distance) = split_plus(rhs, rplus, true);
        addFission(name, e1, p1, p2, rate, number(distance, "product distance"), w1, w2);
        return _reactions.size() - 1;
    }
    std::string p1, p2, unused;
    std::tie(p1, p2, unused) = split_plus(rhs, rplus, false);
    if (e2 != p2)
        throw std::invalid_argument("In enzymatic reaction, educt2 (\"" + e2 + "\") and product2 (\"" + p2 +
                                    "\") have to be equal");
    addEnzymatic(name, e2, e1, p1, rate, number(radius, "educt distance"));
    return _reactions.size() - 1;
}

std::size_t ReactionRegistry::idOf(const std::string &name) const {
    for (std::size_t i = 0; i < _reactions.size(); ++i)
        if (_reactions[i].name == name) return i;
    throw std::invalid_argument("unknown reaction \"" + name + "\"");
}

// ---- TopologyRegistry: potential configuration by type tuple, registered in both directions ------
void TopologyRegistry::configureBondPotential(const std::string &t1, const std::string &t2, BondConfiguration cfg) {
    const auto a = _types->idOf(t1), b = _types->idOf(t2);
    bonds[{a, b}].push_back(cfg);
    if (a != b) bonds[{b, a}].push_back(cfg);
}

void TopologyRegistry::configureAnglePotential(const std::string &t1, const std::string &t2, const std::string &t3,
                                               AngleConfiguration cfg) {
    const auto a = _types->idOf(t1), b = _types->idOf(t2), c = _types->idOf(t3);
    angles[{a, b, c}].push_back(cfg);
    if (a != c) angles[{c, b, a}].push_back(cfg);
}

void TopologyRegistry::configureTorsionPotential(const std::string &t1, const std::string &t2, const std::string &t3,
                                                 const std::string &t4, TorsionConfiguration cfg) {
    const auto a = _types->idOf(t1), b = _types->idOf(t2), c = _types->idOf(t3), d = _types->idOf(t4);
    torsions[{a, b, c, d}].push_back(cfg);
    if (std::make_tuple(a, b, c, d) != std::make_tuple(d, c, b, a)) torsions[{d, c, b, a}].push_back(cfg);
}

// ---- Context -------------------------------------------------------------------------------------
Context::Context() = default;

Context::Context(const Context &o)
    : _types(o._types), _potentials(o._potentials), _reactions(o._reactions), _topologies(o._topologies),
      _box(o._box), _pbc(o._pbc), _kbt(o._kbt), _recordReactions(o._recordReactions) {
    rebind();
}

Context &Context::operator=(const Context &o) {
    if (this == &o) return *this;
    _types = o._types;
    _potentials = o._potentials;
    _reactions = o._reactions;
    _topologies = o._topologies;
    _box = o._box;
    _pbc = o._pbc;
    _kbt = o._kbt;
    _recordReactions = o._recordReactions;
    rebind();
    return *this;
}

// the registries refer to this context's type registry, not the source's
void Context::rebind() {
    _potentials.rebind(&_types);
    _reactions.rebind(&_types);
    _topologies.rebind(&_types);
}

scalar Context::calculateMaxCutoff() const {
    scalar c = 0;
    for (const auto &p : _potentials.potentialsOrder2()) c = std::max(c, p.cutoff());
    for (const auto &r : _reactions.reactions())
        if (r.educts.size() == 2) c = std::max(c, r.eductDistance);
    return c;
}

// ---- GraphTopology ----------------------------------------------------------------------------------
void GraphTopology::addEdge(std::size_t v1, std::size_t v2) {
    if (v1 >= _particles.size() || v2 >= _particles.size() || v1 == v2)
        throw std::invalid_argument("topology edge between invalid vertices");
    if (std::find(_neighbors[v1].begin(), _neighbors[v1].end(), v2) != _neighbors[v1].end()) return;
    _neighbors[v1].push_back(v2);
    _neighbors[v2].push_back(v1);
}

bool GraphTopology::hasEdge(std::size_t v1, std::size_t v2) const {
    if (v1 >= _particles.size() || v2 >= _particles.size()) return false;
    return std::find(_neighbors[v1].begin(), _neighbors[v1].end(), v2) != _neighbors[v1].end();
}

bool GraphTopology::removeEdge(std::size_t v1, std::size_t v2) {
    if (!hasEdge(v1, v2)) return false;
    _neighbors[v1].erase(std::find(_neighbors[v1].begin(), _neighbors[v1].end(), v2));
    _neighbors[v2].erase(std::find(_neighbors[v2].begin(), _neighbors[v2].end(), v1));
    return true;
}

std::vector<GraphTopology> GraphTopology::connectedComponents() const {
    const std::size_t n = _particles.size();
    std::vector<std::size_t> comp(n, SIZE_MAX);
    std::size_t nc = 0;
    for (std::size_t s = 0; s < n; ++s) {  // breadth-first labelling, components ordered by lowest vertex
        if (comp[s] != SIZE_MAX) continue;
        std::vector<std::size_t> queue{s};
        comp[s] = nc;
        for (std::size_t q = 0; q < queue.size(); ++q)
            for (std::size_t m : _neighbors[queue[q]])
                if (comp[m] == SIZE_MAX) {
                    comp[m] = nc;
                    queue.push_back(m);
                }
        ++nc;
    }
    std::vector<std::vector<std::size_t>> members(nc);
    std::vector<std::size_t> local(n);
    for (std::size_t v = 0; v < n; ++v) {
        local[v] = members[comp[v]].size();
        members[comp[v]].push_back(v);
    }
    std::vector<GraphTopology> out;
    for (std::size_t c = 0; c < nc; ++c) {
        std::vector<std::size_t> idx;
        for (std::size_t v : members[c]) idx.push_back(_particles[v]);
        GraphTopology t(_type, idx);
        for (std::size_t v : members[c])
            for (std::size_t m : _neighbors[v]) t._neighbors[local[v]].push_back(local[m]);
        out.push_back(std::move(t));
    }
    return out;
}

void BreakConfig::addBreakablePair(ParticleTypeId type1, ParticleTypeId type2, scalar thresholdEnergy, scalar rate) {
    if (!(rate > 0.))  // Actions.h:193-195
        throw std::logic_error("Rate of breaking a bond must be positive, was " + std::to_string(rate));
    _pairs.push_back({type1, type2, thresholdEnergy, rate});
}

// edges once (first seen from the lower-visited vertex), 3-paths (n2, v, n) with n2 < n around each
// vertex v, 4-paths (n', v, n, m) along each edge (v, n) first seen from v (Graph_detail.h:39-78)
void GraphTopology::findNTuples(
    const std::function<void(std::size_t, std::size_t)> &pair,
    const std::function<void(std::size_t, std::size_t, std::size_t)> &triple,
    const std::function<void(std::size_t, std::size_t, std::size_t, std::size_t)> &quad) const {
    std::vector<char> visited(_neighbors.size(), 0);
    for (std::size_t v = 0; v < _neighbors.size(); ++v) {
        visited[v] = 1;
        const auto &nv = _neighbors[v];
        for (std::size_t n : nv) {
            if (!visited[n]) {
                pair(v, n);
                for (std::size_t n2 : nv) {
                    if (n2 == n) continue;
                    for (std::size_t m : _neighbors[n])
                        if (m != v && m != n2) quad(n2, v, n, m);
                }
            }
            for (std::size_t n2 : nv)
                if (n2 != n && n2 < n) triple(n2, v, n);
        }
    }
}

}  // namespace model

// ---- KernelProvider -------------------------------------------------------------------------------
namespace plugin {

void KernelProvider::KernelDeleter::operator()(model::Kernel *k) const {
    delete k;  // before the library reference is dropped
}

KernelProvider &KernelProvider::getInstance() {
    static KernelProvider instance;
    return instance;
}

void KernelProvider::add(const std::string &name, factory f) { _factories[name] = Entry{std::move(f), nullptr}; }

bool KernelProvider::isSharedLibrary(const std::string &path) {  // KernelProvider.cpp:123-133
    struct stat st {};
    if (stat(path.c_str(), &st) != 0 || !S_ISREG(st.st_mode)) return false;
    const auto dot = path.rfind(".so");
    return dot != std::string::npos && (dot + 3 == path.size() || path[dot + 3] == '.');
}

void KernelProvider::loadKernelLibrary(const std::string &path) {  // KernelProvider.cpp:135-175
    void *h = dlopen(path.c_str(), RTLD_LAZY | RTLD_GLOBAL);
    if (!h) throw std::runtime_error(std::string("could not load kernel library ") + path + ": " + dlerror());
    std::shared_ptr<void> lib(h, [](void *p) { dlclose(p); });
    using name_fn = const char *(*)();
    using create_fn = model::Kernel *(*)();
    auto nm = reinterpret_cast<name_fn>(dlsym(h, "name"));
    auto cr = reinterpret_cast<create_fn>(dlsym(h, "createKernel"));
    if (!nm || !cr)
        throw std::runtime_error("kernel library " + path + " does not export name() and createKernel()");
    _factories[nm()] = Entry{cr, lib};
}

void KernelProvider::loadKernelsFromDirectory(const std::string &directory) {
    DIR *d = opendir(directory.c_str());
    if (!d) throw std::runtime_error("no such kernel plugin directory: " + directory);
    std::vector<std::string> files;
    while (dirent *e = readdir(d)) files.push_back(directory + "/" + e->d_name);
    closedir(d);
    std::sort(files.begin(), files.end());
    for (const auto &f : files)
        if (isSharedLibrary(f)) {
            void *probe = dlopen(f.c_str(), RTLD_LAZY | RTLD_LOCAL | RTLD_NOLOAD);
            bool is_plugin = false;
            void *h = probe ? probe : dlopen(f.c_str(), RTLD_LAZY | RTLD_LOCAL);
            if (h) {
                is_plugin = dlsym(h, "name") && dlsym(h, "createKernel");
                dlclose(h);
            }
            if (is_plugin) loadKernelLibrary(f);
        }
}

KernelProvider::kernel_ptr KernelProvider::create(const std::string &name) const {
    auto it = _factories.find(name);
    if (it == _factories.end()) throw std::invalid_argument("no kernel named \"" + name + "\" is available");
    return kernel_ptr(it->second.make(), KernelDeleter{it->second.library});
}

std::vector<std::string> KernelProvider::availableKernels() const {
    std::vector<std::string> v;
    for (const auto &kv : _factories) v.push_back(kv.first);
    return v;
}

}  // namespace plugin

// ---- SimulationLoop (SimulationLoop.h:256-306) -------------------------------------------------------
namespace api {

SimulationLoop::SimulationLoop(model::Kernel *kernel, scalar timeStep) : _kernel(kernel), _dt(timeStep) {
    if (!kernel) throw std::invalid_argument("simulation loop needs a kernel");
    if (!(timeStep > 0)) throw std::invalid_argument("time step must be positive");
    if (!kernel->supportsGillespie()) _scheduler = "UncontrolledApproximation";
}

void SimulationLoop::useReactionScheduler(const std::string &name) {
    if (name != "UncontrolledApproximation" && name != "Gillespie")
        throw std::invalid_argument("unknown reaction scheduler \"" + name + "\"");
    _scheduler = name;
}

void SimulationLoop::run(std::size_t steps) {
    auto &actions = _kernel->actions();
    auto &state = _kernel->stateModel();
    const bool reactions = _reactions && !_kernel->context().reactions().reactions().empty();
    const std::string scheduler = reactions ? _scheduler : std::string("none");
    if (_callbacks.empty() && _forces && _kernel->runFused(scheduler, _dt, steps)) {
        state.setTime(state.time() + _dt * (scalar)steps);
        return;
    }
    auto integrator = actions.eulerBDIntegrator(_dt);
    auto forces = actions.calculateForces();
    auto nl = actions.updateNeighborList();
    std::unique_ptr<model::TimeStepDependentAction> react;
    if (reactions)
        react = scheduler == "Gillespie" ? actions.gillespie(_dt) : actions.uncontrolledApproximation(_dt);
    nl->perform();
    if (_forces) forces->perform();
    for (auto &cb : _callbacks) cb(0);
    for (std::size_t t = 0; t < steps; ++t) {
        integrator->perform();
        nl->perform();
        if (react) {
            react->perform();
            nl->perform();
        }
        if (_forces) forces->perform();
        state.setTime(state.time() + _dt);
        for (auto &cb : _callbacks) cb(t + 1);
    }
}

}  // namespace api
}  // namespace readdy_amd
