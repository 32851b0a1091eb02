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
    // validateTypeName (readdy/main/model/Utils.cpp:171-190): none of the descriptor tokens, no leading space
    // (inner spaces are allowed, "Topology A"); an empty name is refused as well
    for (const char *cs : {"[", "]", "(", ")", "->", "--", ":", "+"})
        if (name.find(cs) != std::string::npos)
            throw std::invalid_argument("Encountered invalid character sequence \"" + std::string(cs) +
                                        "\" in type name \"" + name + "\"");
    if (name.empty() || name[0] == ' ')
        throw std::invalid_argument("Type name \"" + name + "\" is empty or contained leading/trailing whitespaces.");
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

// ---- TopologyRegistry: topology types and spatial reactions ----------------------------------------
std::size_t TopologyRegistry::addType(const std::string &name) {
    if (name.empty()) throw std::invalid_argument("a topology type needs a name");
    for (const auto &t : _topologyTypes)
        if (t == name) throw std::invalid_argument("the topology type \"" + name + "\" exists already");
    _topologyTypes.push_back(name);
    return _topologyTypes.size() - 1;
}

bool TopologyRegistry::hasType(const std::string &name) const {
    return std::find(_topologyTypes.begin(), _topologyTypes.end(), name) != _topologyTypes.end();
}

std::size_t TopologyRegistry::idOf(const std::string &name) const {
    for (std::size_t i = 0; i < _topologyTypes.size(); ++i)
        if (_topologyTypes[i] == name) return i;
    throw std::invalid_argument("unknown topology type \"" + name + "\"");
}

const std::string &TopologyRegistry::nameOf(std::size_t id) const {
    if (id >= _topologyTypes.size()) throw std::invalid_argument("no topology type with id " + std::to_string(id));
    return _topologyTypes[id];
}

bool TopologyRegistry::isSpatialReactionType(ParticleTypeId t) const {
    for (const auto &r : _spatial)
        if (r.type1 == t || r.type2 == t) return true;
    return false;
}

namespace {
std::string trim(const std::string &s) {
    const auto a = s.find_first_not_of(" \t\n\r"), b = s.find_last_not_of(" \t\n\r");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// one term "T(p)" or "(p)": (particle type, topology type or "")
std::pair<std::string, std::string> str_term(const std::string &s) {
    const auto open = s.find('('), close = s.find(')', open == std::string::npos ? 0 : open);
    if (open == std::string::npos || close == std::string::npos)
        throw std::invalid_argument("The term \"" + s + "\" did not contain a particle type.");
    return {trim(s.substr(open + 1, close - open - 1)), trim(s.substr(0, open))};
}
}  // namespace

void TopologyRegistry::addSpatialReaction(const std::string &descriptor, scalar rate, scalar radius) {
    // STRParser::parse_descriptor (SpatialTopologyReaction.cpp:81-236)
    const std::string arrow = "->";
    const auto ap = descriptor.find(arrow);
    if (ap == std::string::npos)
        throw std::invalid_argument("the descriptor must contain an arrow (\"->\") to indicate lhs and rhs.");
    if (descriptor.find(arrow, ap + 1) != std::string::npos)
        throw std::invalid_argument("the descriptor must not contain more than one arrow (\"->\").");
    std::string lhs = trim(descriptor.substr(0, ap));
    const std::string rhs = trim(descriptor.substr(ap + arrow.size()));
    const auto colon = lhs.find(':');
    if (colon == std::string::npos)
        throw std::invalid_argument("The descriptor did not contain a colon ':' to specify the end of the name.");
    SpatialReactionDesc d{};
    d.name = trim(lhs.substr(0, colon));
    lhs = trim(lhs.substr(colon + 1));
    const auto lp = lhs.find('+');
    if (lp == std::string::npos)
        throw std::invalid_argument("The left hand side of the topology reaction did not contain a '+'.");
    const auto e1 = str_term(lhs.substr(0, lp)), e2 = str_term(lhs.substr(lp + 1));
    std::pair<std::string, std::string> p1, p2;
    const auto rp = rhs.find('+');
    const bool fusion = rp == std::string::npos;
    if (fusion) {  // "T(p1--p2)", an option block after it is not part of the term
        const auto f = str_term(rhs);
        const auto bond = f.first.find("--");
        if (bond == std::string::npos)
            throw std::invalid_argument("The right-hand side was of fusion type but there was no bond \"--\" defined.");
        p1 = {trim(f.first.substr(0, bond)), f.second};
        p2 = {trim(f.first.substr(bond + 2)), ""};
    } else {
        const auto r1 = str_term(rhs.substr(0, rp)), r2 = str_term(rhs.substr(rp + 1));
        p1 = r1;
        p2 = r2;
    }
    d.type1 = _types->idOf(e1.first);
    d.type2 = _types->idOf(e2.first);
    d.typeTo1 = _types->idOf(p1.first);
    d.typeTo2 = _types->idOf(p2.first);
    if (e2.second.empty()) {  // the topology-particle case
        d.mode = fusion ? STRMode::TP_FUSION : STRMode::TP_ENZYMATIC;
        d.topType1 = idOf(e1.second);
        d.topTypeTo1 = idOf(p1.second);
        d.topType2 = d.topTypeTo2 = EmptyTopologyId;
    } else {
        d.topType1 = idOf(e1.second);
        d.topType2 = idOf(e2.second);
        d.topTypeTo1 = idOf(p1.second);
        if (fusion) {
            d.mode = STRMode::TT_FUSION;
            d.topTypeTo2 = EmptyTopologyId;
            // "[self=true]" (with an optional distance) allows a topology to fuse with itself
            const auto ob = rhs.find('[');
            const auto cb = ob == std::string::npos ? std::string::npos : rhs.find(']', ob);
            if (cb != std::string::npos) {
                std::string opts;
                for (char ch : rhs.substr(ob + 1, cb - ob - 1))
                    if (ch != ' ' && ch != '\t') opts += ch;
                if (opts.find("self=true") == std::string::npos)
                    throw std::runtime_error("Option \"self=true\" is missing. This option is currently required for "
                                             "all valid options.");
                d.mode = STRMode::TT_FUSION_ALLOW_SELF;
                const auto dp = opts.find("distance>");  // get_distance (SpatialTopologyReaction.cpp:262-269)
                d.minGraphDistance = dp == std::string::npos ? 0u : (unsigned)std::stoi(opts.substr(dp + 9));
            }
        } else {
            d.mode = STRMode::TT_ENZYMATIC;
            d.topTypeTo2 = idOf(p2.second);
        }
    }
    d.rate = rate;
    d.radius = radius;
    // TopologyRegistry::validateSpatialReaction (TopologyRegistry.cpp:238-293)
    if (radius <= 0)
        throw std::invalid_argument("The radius of an structural topology reaction(" + d.name +
                                    ") should always be positive");
    const auto &i1 = _types->infoOf(d.type1), &i2 = _types->infoOf(d.type2);
    const auto &o1 = _types->infoOf(d.typeTo1), &o2 = _types->infoOf(d.typeTo2);
    if (!i1.topologyFlavor && !i2.topologyFlavor)
        throw std::invalid_argument("At least one of the educt types (" + i1.name + " and " + i2.name +
                                    ") of topology reaction " + d.name + " need to be topology flavored!");
    for (const auto &r : _spatial)
        if (r.name == d.name)
            throw std::invalid_argument("An structural topology reaction with the same name (" + d.name +
                                        ") was already registered!");
    const bool isFusion = d.mode == STRMode::TP_FUSION || d.mode == STRMode::TT_FUSION ||
                          d.mode == STRMode::TT_FUSION_ALLOW_SELF;
    if (isFusion) {
        if (!o1.topologyFlavor || !o2.topologyFlavor)
            throw std::invalid_argument("One or both of the target types (" + o1.name + " and " + o2.name +
                                        ") of topology reaction " + d.name +
                                        " did not have the topology flavor and is_fusion was set to true!");
    } else if (i1.topologyFlavor != o1.topologyFlavor || i2.topologyFlavor != o2.topologyFlavor) {
        throw std::invalid_argument("if is_fusion is false, flavors need to remain same (topology reaction " +
                                    d.name + ")");
    }
    _spatial.push_back(d);
}

// ---- Context -------------------------------------------------------------------------------------
Context::Context() = default;

Context::Context(const Context &o)
    : _types(o._types), _potentials(o._potentials), _reactions(o._reactions), _topologies(o._topologies),
      _box(o._box), _pbc(o._pbc), _kbt(o._kbt), _recordReactions(o._recordReactions),
      _recordVirial(o._recordVirial) {
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
    _recordVirial = o._recordVirial;
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
    for (const auto &r : _topologies.spatialReactions()) c = std::max(c, r.radius);  // Context.cpp:126-144
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

std::size_t GraphTopology::addVertex(std::size_t particle) {
    if (std::find(_particles.begin(), _particles.end(), particle) != _particles.end())
        throw std::invalid_argument("addVertex: the particle is in this topology already");
    _particles.push_back(particle);
    _neighbors.emplace_back();
    return _particles.size() - 1;
}

void GraphTopology::appendTopology(const GraphTopology &other, std::size_t otherParticle, std::size_t thisParticle) {
    const auto mine = std::find(_particles.begin(), _particles.end(), thisParticle);
    const auto theirs = std::find(other._particles.begin(), other._particles.end(), otherParticle);
    if (mine == _particles.end() || theirs == other._particles.end())
        throw std::invalid_argument("appendTopology: a particle of the new edge is not in its topology");
    const std::size_t off = _particles.size(), v = (std::size_t)(mine - _particles.begin()),
                      w = off + (std::size_t)(theirs - other._particles.begin());
    for (std::size_t i = 0; i < other._particles.size(); ++i) {
        _particles.push_back(other._particles[i]);
        _neighbors.emplace_back();
        for (std::size_t n : other._neighbors[i]) _neighbors.back().push_back(off + n);
    }
    addEdge(v, w);
}

void GraphTopology::appendParticle(std::size_t particle, std::size_t counterpart) {
    const auto it = std::find(_particles.begin(), _particles.end(), counterpart);
    if (it == _particles.end()) throw std::invalid_argument("appendParticle: the counterpart is not in this topology");
    if (std::find(_particles.begin(), _particles.end(), particle) != _particles.end())
        throw std::invalid_argument("appendParticle: the particle is in this topology already");
    const std::size_t v = (std::size_t)(it - _particles.begin());
    _particles.push_back(particle);
    _neighbors.emplace_back();
    addEdge(v, _particles.size() - 1);
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

bool GraphTopology::isConnected() const {
    if (_particles.size() <= 1) return true;
    return connectedComponents().size() == 1;
}

void GraphTopology::updateReactionRates(const std::vector<StructuralTopologyReaction> &reactions) {
    _rates.clear();
    for (const auto &r : reactions) _rates.push_back(r.rate(*this));
}

bool GraphTopology::isNormalParticle(const Kernel &kernel) const {
    if (_particles.size() != 1) return false;
    const ParticleId id = _particles[0];
    for (const auto &p : kernel.stateModel().getParticles())
        if (p.id() == id) return !kernel.context().particleTypes().infoOf(p.type()).topologyFlavor;
    throw std::logic_error("a topology's particle is not in the state model");
}

// ---- structural topology reactions --------------------------------------------------------------------
std::size_t StructuralTopologyReaction::counter = 0;

StructuralTopologyReaction::StructuralTopologyReaction(std::string name, reaction_function reaction, rate_function rate)
    : _name(std::move(name)), _reaction(std::move(reaction)), _rate(std::move(rate)), _id(counter++) {}

StructuralTopologyReaction::StructuralTopologyReaction(std::string name, reaction_function reaction, scalar rate)
    : StructuralTopologyReaction(std::move(name), std::move(reaction),
                                 [rate](const GraphTopology &) -> scalar { return rate; }) {}

Recipe &Recipe::changeParticleType(std::size_t vertex, ParticleTypeId to) {
    Step st{};
    st.kind = Step::ChangeParticleType;
    st.v1 = vertex;
    st.type = to;
    _steps.push_back(st);
    return *this;
}
Recipe &Recipe::changeParticleType(std::size_t vertex, const std::string &to) {
    Step st{};
    st.kind = Step::ChangeParticleType;
    st.v1 = vertex;
    st.name = to;
    st.byName = true;
    _steps.push_back(st);
    return *this;
}
Recipe &Recipe::addEdge(std::size_t v1, std::size_t v2) {
    Step st{};
    st.kind = Step::AddEdge;
    st.v1 = v1;
    st.v2 = v2;
    _steps.push_back(st);
    return *this;
}
Recipe &Recipe::removeEdge(std::size_t v1, std::size_t v2) {
    Step st{};
    st.kind = Step::RemoveEdge;
    st.v1 = v1;
    st.v2 = v2;
    _steps.push_back(st);
    return *this;
}
Recipe &Recipe::separateVertex(std::size_t vertex) {  // Recipe.cpp:56-62: one edge removal per neighbour
    for (std::size_t n : _topology->neighborsOf(vertex)) removeEdge(vertex, n);
    return *this;
}
Recipe &Recipe::changeTopologyType(const std::string &type) {
    Step st{};
    st.kind = Step::ChangeTopologyType;
    st.name = type;
    _steps.push_back(st);
    return *this;
}
Recipe &Recipe::changeParticlePosition(std::size_t vertex, const Vec3 &pos) {
    Step st{};
    st.kind = Step::ChangeParticlePosition;
    st.v1 = vertex;
    st.pos = pos;
    _steps.push_back(st);
    return *this;
}
Recipe &Recipe::appendNewParticle(const std::vector<std::size_t> &neighbors, const std::string &type, const Vec3 &pos) {
    Step st{};
    st.kind = Step::AppendNewParticle;
    st.neighbors = neighbors;
    st.name = type;
    st.pos = pos;
    _steps.push_back(st);
    return *this;
}

std::vector<GraphTopology> StructuralTopologyReaction::execute(GraphTopology &topology, Kernel *kernel) const {
    const auto &types = kernel->context().particleTypes();
    const auto &registry = kernel->context().topologyRegistry();
    Recipe recipe = operations(topology);
    const auto &steps = recipe.steps();
    if (steps.empty()) return {};
    TopologyActionFactory *taf = kernel->getTopologyActionFactory();
    if (!taf) throw std::logic_error("the kernel has no topology action factory");
    std::vector<TopologyActionFactory::ActionPtr> actions;
    for (const auto &st : steps) {
        switch (st.kind) {
        case Recipe::Step::ChangeParticleType:
            actions.push_back(taf->createChangeParticleType(&topology, st.v1, st.byName ? types.idOf(st.name) : st.type));
            break;
        case Recipe::Step::AddEdge: actions.push_back(taf->createAddEdge(&topology, st.v1, st.v2)); break;
        case Recipe::Step::RemoveEdge:
        case Recipe::Step::SeparateVertex: actions.push_back(taf->createRemoveEdge(&topology, st.v1, st.v2)); break;
        case Recipe::Step::ChangeTopologyType: actions.push_back(taf->createChangeTopologyType(&topology, st.name)); break;
        case Recipe::Step::ChangeParticlePosition:
            actions.push_back(taf->createChangeParticlePosition(&topology, st.v1, st.pos));
            break;
        case Recipe::Step::AppendNewParticle:
            actions.push_back(taf->createAppendParticle(&topology, st.neighbors, types.idOf(st.name), st.pos));
            break;
        }
    }
    for (auto &a : actions) a->perform();
    if (expects_connected_after_reaction()) {
        bool valid = topology.isConnected();
        for (std::size_t id : topology.particleIndices())
            for (const auto &p : kernel->stateModel().getParticles())
                if (p.id() == id && !types.infoOf(p.type()).topologyFlavor) valid = false;
        if (!valid)
            throw TopologyReactionException("The topology was invalid after the reaction, see previous warning messages.");
        kernel->stateModel().configureTopologies();
        topology.updateReactionRates(registry.structuralReactionsOf(topology.type()));
        return {};
    }
    if (!topology.isConnected()) return topology.connectedComponents();
    if (!topology.isNormalParticle(*kernel)) {
        kernel->stateModel().configureTopologies();
        topology.updateReactionRates(registry.structuralReactionsOf(topology.type()));
    }
    return {};
}

void TopologyRegistry::addStructuralReaction(const std::string &type, StructuralTopologyReaction reaction) {
    if (!hasType(type)) throw std::invalid_argument("unknown topology type \"" + type + "\"");
    _structural[type].push_back(std::move(reaction));
}

const std::vector<StructuralTopologyReaction> &TopologyRegistry::structuralReactionsOf(const std::string &type) const {
    static const std::vector<StructuralTopologyReaction> none;
    auto it = _structural.find(type);
    return it == _structural.end() ? none : it->second;
}

std::string DetailedBalance::describe() const {  // Actions.cpp:521-545
    std::string s;
    if (_reversibleReactions.empty()) return s;
    s += "- reversible reactions handled subject to detailed-balance:\n";
    for (const auto &r : _reversibleReactions) {
        s += "     * -> \"" + r.forwardName + "\" and <- \"" + r.backwardName + "\", with types: ";
        switch (r.reversibleType) {
        case ReversibleType::FusionFission:
            s += r.lhsNames[0] + " + " + r.lhsNames[1] + " <--> " + r.rhsNames[0];
            break;
        case ReversibleType::ConversionConversion:
            s += r.lhsNames[0] + " <--> " + r.rhsNames[0];
            break;
        case ReversibleType::EnzymaticEnzymatic:
            s += r.lhsNames[0] + " + " + r.lhsNames[1] + " <--> " + r.rhsNames[0] + " + " + r.rhsNames[1];
            break;
        }
        s += "\n";
    }
    return s;
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

// ---- StateModel observable inputs: kernels without them ------------------------------------------
std::array<scalar, 9> StateModel::virial() const { throw std::logic_error("this kernel provides no virial"); }
std::vector<unsigned long> StateModel::particlesPerType() const {
    throw std::logic_error("this kernel provides no particle counts per type");
}
std::vector<uint64_t> StateModel::histogramAlongAxis(const std::vector<scalar> &, const std::vector<ParticleTypeId> &,
                                                     unsigned) const {
    throw std::logic_error("this kernel provides no histogram along an axis");
}
std::vector<uint64_t> StateModel::rdfCounts(const std::vector<scalar> &, const std::vector<ParticleTypeId> &,
                                            const std::vector<ParticleTypeId> &, uint64_t *) const {
    throw std::logic_error("this kernel provides no radial distribution");
}
void StateModel::changeParticleType(ParticleId, ParticleTypeId) {
    throw std::logic_error("this kernel cannot change a particle's species");
}
void StateModel::changeParticlePosition(ParticleId, const Vec3 &) {
    throw std::logic_error("this kernel cannot move a topology particle");
}
ParticleId StateModel::appendTopologyParticle(ParticleTypeId, const Vec3 &) {
    throw std::logic_error("this kernel cannot add topology particles");
}

// ---- topology reaction actions (TopologyReactionActionFactory.h:55-101) ------------------------------
namespace {

class ChangeParticleType : public TopologyReactionAction {
public:
    ChangeParticleType(GraphTopology *t, StateModel *state, std::size_t vertex, ParticleTypeId to)
        : TopologyReactionAction(t), _state(state), _vertex(vertex), _to(to) {
        if (vertex >= t->nParticles()) throw std::invalid_argument("no such vertex in the topology");
    }
    void perform() override {
        const ParticleId id = _topology->particleIndices()[_vertex];
        bool found = false;
        for (const auto &p : _state->getParticles())
            if (p.id() == id) {
                _from = p.type();
                found = true;
                break;
            }
        if (!found) throw std::logic_error("the topology's particle is not in the state model");
        _state->changeParticleType(id, _to);
        _done = true;
    }
    void undo() override {
        if (_done) _state->changeParticleType(_topology->particleIndices()[_vertex], _from);
        _done = false;
    }

private:
    StateModel *_state;
    std::size_t _vertex;
    ParticleTypeId _to, _from = 0;
    bool _done = false;
};

class EdgeAction : public TopologyReactionAction {
public:
    EdgeAction(GraphTopology *t, std::size_t v1, std::size_t v2, bool add)
        : TopologyReactionAction(t), _v1(v1), _v2(v2), _add(add) {}
    void perform() override {
        if (_add) {
            _changed = !_topology->hasEdge(_v1, _v2);
            _topology->addEdge(_v1, _v2);
        } else {
            _changed = _topology->removeEdge(_v1, _v2);
            if (!_changed) throw std::invalid_argument("the topology has no such edge");
        }
    }
    void undo() override {
        if (!_changed) return;
        if (_add) _topology->removeEdge(_v1, _v2);
        else _topology->addEdge(_v1, _v2);
        _changed = false;
    }

private:
    std::size_t _v1, _v2;
    bool _add, _changed = false;
};

class ChangeParticlePosition : public TopologyReactionAction {
public:
    ChangeParticlePosition(GraphTopology *t, StateModel *state, std::size_t vertex, const Vec3 &pos)
        : TopologyReactionAction(t), _state(state), _vertex(vertex), _to(pos) {
        if (vertex >= t->nParticles()) throw std::invalid_argument("no such vertex in the topology");
    }
    void perform() override {
        const ParticleId id = _topology->particleIndices()[_vertex];
        bool found = false;
        for (const auto &p : _state->getParticles())
            if (p.id() == id) {
                _from = p.pos();
                found = true;
                break;
            }
        if (!found) throw std::logic_error("the topology's particle is not in the state model");
        _state->changeParticlePosition(id, _to);
        _done = true;
    }
    void undo() override {
        if (_done) _state->changeParticlePosition(_topology->particleIndices()[_vertex], _from);
        _done = false;
    }

private:
    StateModel *_state;
    std::size_t _vertex;
    Vec3 _to, _from{};
    bool _done = false;
};

// the new particle stays in the state model on undo (a recipe's step cannot take a particle back)
class AppendParticle : public TopologyReactionAction {
public:
    AppendParticle(GraphTopology *t, StateModel *state, std::vector<std::size_t> neighbors, ParticleTypeId type,
                   const Vec3 &pos)
        : TopologyReactionAction(t), _state(state), _neighbors(std::move(neighbors)), _type(type), _pos(pos) {
        for (std::size_t v : _neighbors)
            if (v >= t->nParticles()) throw std::invalid_argument("no such vertex in the topology");
    }
    void perform() override {
        const ParticleId id = _state->appendTopologyParticle(_type, _pos);
        const std::size_t v = _topology->addVertex((std::size_t)id);
        for (std::size_t n : _neighbors) _topology->addEdge(n, v);
    }
    void undo() override { throw std::logic_error("an appended particle cannot be taken back"); }

private:
    StateModel *_state;
    std::vector<std::size_t> _neighbors;
    ParticleTypeId _type;
    Vec3 _pos;
};

class ChangeTopologyType : public TopologyReactionAction {
public:
    ChangeTopologyType(GraphTopology *t, std::string to) : TopologyReactionAction(t), _to(std::move(to)) {}
    void perform() override {
        _from = _topology->type();
        _topology->setType(_to);
    }
    void undo() override { _topology->setType(_from); }

private:
    std::string _to, _from;
};

}  // namespace

TopologyActionFactory::ActionPtr TopologyActionFactory::createChangeParticleType(GraphTopology *topology,
                                                                                 std::size_t vertex,
                                                                                 ParticleTypeId typeTo) const {
    return std::make_unique<ChangeParticleType>(topology, _state, vertex, typeTo);
}
TopologyActionFactory::ActionPtr TopologyActionFactory::createAddEdge(GraphTopology *topology, std::size_t v1,
                                                                      std::size_t v2) const {
    return std::make_unique<EdgeAction>(topology, v1, v2, true);
}
TopologyActionFactory::ActionPtr TopologyActionFactory::createRemoveEdge(GraphTopology *topology, std::size_t v1,
                                                                         std::size_t v2) const {
    return std::make_unique<EdgeAction>(topology, v1, v2, false);
}
TopologyActionFactory::ActionPtr TopologyActionFactory::createChangeParticlePosition(GraphTopology *topology,
                                                                                     std::size_t vertex,
                                                                                     const Vec3 &pos) const {
    return std::make_unique<ChangeParticlePosition>(topology, _state, vertex, pos);
}
TopologyActionFactory::ActionPtr TopologyActionFactory::createAppendParticle(GraphTopology *topology,
                                                                             const std::vector<std::size_t> &neighbors,
                                                                             ParticleTypeId type, const Vec3 &pos) const {
    return std::make_unique<AppendParticle>(topology, _state, neighbors, type, pos);
}
TopologyActionFactory::ActionPtr TopologyActionFactory::createChangeTopologyType(GraphTopology *topology,
                                                                                 const std::string &typeTo) const {
    return std::make_unique<ChangeTopologyType>(topology, typeTo);
}

// ---- observables (readdy/main/model/observables/*.cpp; kernels/cpu/src/observables/CPUObservables.cpp)
namespace observables {

void Energy::evaluate() { result = kernel->stateModel().energy(); }

void Virial::evaluate() { result = kernel->stateModel().virial(); }

void NParticles::evaluate() {
    const auto per = kernel->stateModel().particlesPerType();
    result.clear();
    if (typesToCount.empty()) {  // CPUNParticles::evaluate: the total when no types are listed
        unsigned long n = 0;
        for (auto c : per) n += c;
        result.push_back(n);
    } else {
        for (auto t : typesToCount) result.push_back(t < per.size() ? per[t] : 0ul);
    }
}

namespace {
bool listed(const std::vector<ParticleTypeId> &types, ParticleTypeId t) {
    return types.empty() || std::find(types.begin(), types.end(), t) != types.end();
}
}  // namespace

void Forces::evaluate() {
    result.clear();
    const auto ps = kernel->stateModel().getParticles();
    const auto fs = kernel->stateModel().getParticleForces();
    for (std::size_t i = 0; i < ps.size(); ++i)
        if (listed(typesToCount, ps[i].type())) result.push_back(fs[i]);
}

void Positions::evaluate() {
    result.clear();
    for (const auto &p : kernel->stateModel().getParticles())
        if (listed(typesToCount, p.type())) result.push_back(p.pos());
}

void HistogramAlongAxis::evaluate() {
    const auto c = kernel->stateModel().histogramAlongAxis(binBorders, typesToCount, axis);
    result.assign(c.begin(), c.end());
}

RadialDistribution::RadialDistribution(Kernel *k, Stride s, std::vector<scalar> borders,
                                       std::vector<ParticleTypeId> from, std::vector<ParticleTypeId> to,
                                       scalar particleToDensity)
    : Observable(k, s), binBorders(std::move(borders)), typeCountFrom(std::move(from)), typeCountTo(std::move(to)),
      particleToDensity(particleToDensity) {
    if (binBorders.size() < 2) throw std::invalid_argument("radial distribution needs at least two bin borders");
    auto &centers = std::get<0>(result);
    for (std::size_t b = 0; b + 1 < binBorders.size(); ++b) centers.push_back(.5 * (binBorders[b] + binBorders[b + 1]));
    std::get<1>(result).assign(centers.size(), 0.0);
}

void RadialDistribution::evaluate() {
    uint64_t nFrom = 0;
    const auto counts = kernel->stateModel().rdfCounts(binBorders, typeCountFrom, typeCountTo, &nFrom);
    auto &dist = std::get<1>(result);
    // RadialDistribution.cpp:105-123
    const scalar pi = 3.14159265358979323846;
    for (std::size_t b = 0; b < counts.size(); ++b) {
        const scalar lo = binBorders[b], hi = binBorders[b + 1];
        dist[b] = (scalar)counts[b] /
                  (4. / 3. * pi * (std::pow(hi, 3.) - std::pow(lo, 3.)) * (scalar)nFrom * particleToDensity);
    }
}

void Particles::evaluate() {
    auto &types = std::get<0>(result);
    auto &ids = std::get<1>(result);
    auto &pos = std::get<2>(result);
    types.clear();
    ids.clear();
    pos.clear();
    for (const auto &p : kernel->stateModel().getParticles()) {
        types.push_back(p.type());
        ids.push_back(p.id());
        pos.push_back(p.pos());
    }
}

void Reactions::evaluate() {
    if (!kernel->context().recordReactionsWithPositions())
        throw std::logic_error("the Reactions observable needs Context::recordReactionsWithPositions");
    result = kernel->stateModel().reactionRecords();
}

void ReactionCounts::evaluate() {
    const auto c = kernel->stateModel().reactionCounts();
    result.clear();
    for (std::size_t i = 0; i < c.size(); ++i) result[i] = (std::size_t)c[i];
}

std::vector<ParticleTypeId> ObservableFactory::ids(const std::vector<std::string> &names) const {
    std::vector<ParticleTypeId> out;
    for (const auto &n : names) out.push_back(kernel->context().particleTypes().idOf(n));
    return out;
}

namespace {
template <class T, class CB>
std::unique_ptr<T> with_callback(std::unique_ptr<T> o, CB cb) {
    o->callback() = std::move(cb);
    return o;
}
}  // namespace

std::unique_ptr<Energy> ObservableFactory::energy(Stride stride, ObsCallback<Energy> cb) const {
    return with_callback(std::make_unique<Energy>(kernel, stride), std::move(cb));
}
std::unique_ptr<Virial> ObservableFactory::virial(Stride stride, ObsCallback<Virial> cb) const {
    kernel->context().recordVirial() = true;  // as CPUObservableFactory::virial does (CPUObservableFactory.cpp:128)
    return with_callback(std::make_unique<Virial>(kernel, stride), std::move(cb));
}
std::unique_ptr<HistogramAlongAxis> ObservableFactory::histogramAlongAxis(Stride stride, std::vector<scalar> borders,
                                                                          std::vector<std::string> types, unsigned axis,
                                                                          ObsCallback<HistogramAlongAxis> cb) const {
    if (axis > 2) throw std::invalid_argument("histogram axis must be 0, 1 or 2");
    return with_callback(std::make_unique<HistogramAlongAxis>(kernel, stride, std::move(borders), ids(types), axis),
                         std::move(cb));
}
std::unique_ptr<NParticles> ObservableFactory::nParticles(Stride stride, std::vector<std::string> types,
                                                          ObsCallback<NParticles> cb) const {
    return with_callback(std::make_unique<NParticles>(kernel, stride, ids(types)), std::move(cb));
}
std::unique_ptr<Forces> ObservableFactory::forces(Stride stride, std::vector<std::string> types,
                                                  ObsCallback<Forces> cb) const {
    return with_callback(std::make_unique<Forces>(kernel, stride, ids(types)), std::move(cb));
}
std::unique_ptr<Positions> ObservableFactory::positions(Stride stride, std::vector<std::string> types,
                                                        ObsCallback<Positions> cb) const {
    return with_callback(std::make_unique<Positions>(kernel, stride, ids(types)), std::move(cb));
}
std::unique_ptr<RadialDistribution> ObservableFactory::radialDistribution(Stride stride, std::vector<scalar> borders,
                                                                          std::vector<std::string> from,
                                                                          std::vector<std::string> to,
                                                                          scalar particleToDensity,
                                                                          ObsCallback<RadialDistribution> cb) const {
    return with_callback(
        std::make_unique<RadialDistribution>(kernel, stride, std::move(borders), ids(from), ids(to), particleToDensity),
        std::move(cb));
}
std::unique_ptr<Particles> ObservableFactory::particles(Stride stride, ObsCallback<Particles> cb) const {
    return with_callback(std::make_unique<Particles>(kernel, stride), std::move(cb));
}
std::unique_ptr<Reactions> ObservableFactory::reactions(Stride stride, ObsCallback<Reactions> cb) const {
    kernel->context().recordReactionsWithPositions() = true;
    return with_callback(std::make_unique<Reactions>(kernel, stride), std::move(cb));
}
std::unique_ptr<ReactionCounts> ObservableFactory::reactionCounts(Stride stride, ObsCallback<ReactionCounts> cb) const {
    return with_callback(std::make_unique<ReactionCounts>(kernel, stride), std::move(cb));
}

}  // namespace observables

// ---- Kernel observable bookkeeping --------------------------------------------------------------
void Kernel::connectObservable(observables::ObservableBase *o) {
    if (!o) throw std::invalid_argument("null observable");
    if (std::find(_observables.begin(), _observables.end(), o) == _observables.end()) _observables.push_back(o);
}

void Kernel::disconnectObservable(observables::ObservableBase *o) {
    _observables.erase(std::remove(_observables.begin(), _observables.end(), o), _observables.end());
}

void Kernel::evaluateObservables(observables::TimeStep t) {
    for (auto *o : _observables) o->call(t);
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
    if (name != "UncontrolledApproximation" && name != "Gillespie" && name != "DetailedBalance")
        throw std::invalid_argument("Requested reaction scheduler " + name + " is not available.");
    if (name == "DetailedBalance") (void)_kernel->actions().detailedBalance(_dt);  // its pairs are checked now
    _scheduler = name;
}

void SimulationLoop::run(std::size_t steps) {
    auto &actions = _kernel->actions();
    auto &state = _kernel->stateModel();
    const bool reactions = _reactions && !_kernel->context().reactions().reactions().empty();
    const std::string scheduler = reactions ? _scheduler : std::string("none");
    const auto &obs = _kernel->connectedObservables();
    const observables::TimeStep t0 = _t;
    // the device loop performs the registered topology reactions in every step
    const auto &treg = _kernel->context().topologyRegistry();
    const bool fusable = (_topologyReactions || treg.spatialReactions().empty()) &&
                         !(_topologyReactions && treg.hasStructuralReactions());
    if (_callbacks.empty() && _forces && fusable) {
        if (obs.empty()) {
            if (_kernel->runFused(scheduler, _dt, steps)) {
                state.setTime(state.time() + _dt * (scalar)steps);
                _t += steps;
                return;
            }
        } else {
            // the loop evaluates the observables before the first step (after the initial forces) and
            // after every step (SimulationLoop.h:270-300); each observable decides by its stride, so the
            // device runs fused chunks that end where the next observable is due
            auto next_due = [&](observables::TimeStep t) {
                observables::TimeStep n = t0 + steps;
                for (auto *o : obs) {
                    const auto s = o->stride();
                    n = std::min<observables::TimeStep>(n, s == 0 ? t + 1 : (t / s + 1) * s);
                }
                return n;
            };
            observables::TimeStep t = t0;
            bool fused = true, started = false;
            while (t < t0 + steps) {
                if (!started) {
                    actions.calculateForces()->perform();  // the loop's runForces() before its first evaluation
                    _kernel->evaluateObservables(t);
                    started = true;
                }
                const observables::TimeStep n = next_due(t);
                if (!_kernel->runFused(scheduler, _dt, n - t)) {
                    fused = false;
                    break;
                }
                state.setTime(state.time() + _dt * (scalar)(n - t));
                t = n;
                _t = t;
                _kernel->evaluateObservables(t);
            }
            if (fused) return;
            steps -= (std::size_t)(t - t0);  // nothing ran fused: the separate actions below take all steps
        }
    }
    auto integrator = actions.eulerBDIntegrator(_dt);
    auto forces = actions.calculateForces();
    auto nl = actions.updateNeighborList();
    std::unique_ptr<model::TimeStepDependentAction> react, topoReact;
    if (reactions)
        react = scheduler == "Gillespie"         ? actions.gillespie(_dt)
                : scheduler == "DetailedBalance" ? std::unique_ptr<model::TimeStepDependentAction>(
                                                       actions.detailedBalance(_dt))
                                                 : actions.uncontrolledApproximation(_dt);
    if (_topologyReactions && (!treg.spatialReactions().empty() || treg.hasStructuralReactions()))
        topoReact = actions.evaluateTopologyReactions(_dt);
    nl->perform();
    if (_forces) forces->perform();
    _kernel->evaluateObservables(_t);
    for (auto &cb : _callbacks) cb(0);
    for (std::size_t t = 0; t < steps; ++t) {
        integrator->perform();
        nl->perform();
        if (react) react->perform();
        if (topoReact) topoReact->perform();  // SimulationLoop.h:286-287
        if (react || topoReact) nl->perform();
        if (_forces) forces->perform();
        state.setTime(state.time() + _dt);
        ++_t;
        _kernel->evaluateObservables(_t);
        for (auto &cb : _callbacks) cb(t + 1);
    }
}

}  // namespace api
}  // namespace readdy_amd
