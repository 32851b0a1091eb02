// C++ boundary test: loads the MI355X kernel plugin the way a ReaDDy host loads kernel libraries
// (KernelProvider::loadKernelsFromDirectory, then create("MI355X")), configures systems through
// the Context registries and drives them through StateModel / ActionFactory / SimulationLoop.
// Results are checked against the FP64 C oracle (oracle/readdy_oracle.c) on identical inputs.
//
// usage: test_plugin <plugin directory> [--provider-only]     exit status 0 = all checks passed
// (--provider-only: plugin loading, kernel creation and error behaviour; no device work)
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "../../readdy_amd/host/readdy_host.h"

extern "C" {
#include "../../oracle/readdy_oracle.h"
}

using namespace readdy_amd;

static int failures = 0;
#define CHECK(cond, ...)                                   \
    do {                                                   \
        if (!(cond)) {                                     \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                      \
            std::printf("\n");                             \
            ++failures;                                    \
        }                                                  \
    } while (0)

struct System {
    std::vector<double> pos;
    std::vector<int> types;
};

static System uniform(int nA, int nB, double L, unsigned seed) {
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> u(-0.5 * L, 0.5 * L);
    System s;
    for (int i = 0; i < nA + nB; ++i) {
        for (int a = 0; a < 3; ++a) s.pos.push_back(u(g));
        s.types.push_back(i < nA ? 0 : 1);
    }
    return s;
}

// C0 chemistry (harmonic repulsion on all pairs, A+B<->C) in the kernel's context
static void c0_context(model::Context &c, double L) {
    c.boxSize() = {L, L, L};
    c.periodicBoundaryConditions() = {true, true, true};
    c.particleTypes().add("A", 1.0);
    c.particleTypes().add("B", 1.0);
    c.particleTypes().add("C", 0.5);
    const char *t[3] = {"A", "B", "C"};
    for (int a = 0; a < 3; ++a)
        for (int b = a; b < 3; ++b) c.potentials().addHarmonicRepulsion(t[a], t[b], 10.0, 1.0);
    c.reactions().addFusion("fus", "A", "B", "C", 20.0, 1.0);
    c.reactions().addFission("fis", "C", "A", "B", 10.0, 1.0);
}

static or_system *c0_oracle(const System &s, double L) {
    const double box[3] = {L, L, L};
    const int pbc[3] = {1, 1, 1};
    const double D[3] = {1.0, 1.0, 0.5};
    or_system *o = or_create(3, box, pbc, 1.0, D);
    const double hr[2] = {10.0, 1.0};
    for (int a = 0; a < 3; ++a)
        for (int b = a; b < 3; ++b) or_add_pair_potential(o, OR_POT_HARMONIC_REPULSION, a, b, hr);
    or_add_reaction(o, OR_RX_FUSION, 0, 1, 2, -1, 20.0, 1.0, 0.0, 0.5, 0.5);
    or_add_reaction(o, OR_RX_FISSION, 2, -1, 0, 1, 10.0, 0.0, 1.0, 0.5, 0.5);
    std::vector<uint64_t> ids(s.types.size());
    for (size_t i = 0; i < ids.size(); ++i) ids[i] = i;
    or_add_particles(o, (int)s.types.size(), s.pos.data(), s.types.data(), ids.data());
    or_nl_setup(o, or_max_cutoff(o), 1);
    return o;
}

struct OracleState {
    std::map<uint64_t, std::array<double, 3>> pos, force;
    std::map<uint64_t, int> type;
};

static OracleState oracle_state(const or_system *o) {
    const size_t n = or_n_active(o);
    std::vector<double> p(3 * n), f(3 * n);
    std::vector<int> t(n);
    std::vector<uint64_t> id(n);
    or_get_particles(o, n, p.data(), t.data(), id.data(), f.data());
    OracleState s;
    for (size_t i = 0; i < n; ++i) {
        s.pos[id[i]] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
        s.force[id[i]] = {f[3 * i], f[3 * i + 1], f[3 * i + 2]};
        s.type[id[i]] = t[i];
    }
    return s;
}

static double wrapd(double d, double L) { return d - L * std::round(d / L); }

static void test_provider_and_errors(const std::string &dir) {
    auto &prov = plugin::KernelProvider::getInstance();
    prov.loadKernelsFromDirectory(dir);
    bool found = false;
    for (const auto &n : prov.availableKernels()) found = found || n == "MI355X";
    CHECK(found, "MI355X not among the available kernels");
    bool threw = false;
    try {
        prov.create("NoSuchKernel");
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    CHECK(threw, "unknown kernel name must throw std::invalid_argument");
    auto k = prov.create("MI355X");
    CHECK(k->name() == "MI355X", "kernel name");
    CHECK(k->supportsGillespie(), "MI355X supports Gillespie");
    std::printf("provider: ok\n");
}

static bool close_rel(double a, double b, double eps) { return std::fabs(a - b) <= eps * std::max(std::fabs(a), std::fabs(b)); }

// DetailedBalance's reversible pairs through the plugin (host computation): the reference's known
// answers (readdy/test/TestDetailedBalance.cpp:65-91, 322-327) and the scheduler names
static void test_detailed_balance_configs() {
    auto k = plugin::KernelProvider::getInstance().create("MI355X");
    auto &c = k->context();
    const double L = 21.544346900318832;
    c.boxSize() = {L, L, L};
    c.particleTypes().add("A", 1.);
    c.particleTypes().add("B", 1.);
    c.particleTypes().add("C", 1.);
    c.potentials().addHarmonicRepulsion("A", "B", 2., 3.);
    c.reactions().add("fusion: A +(5) B -> C", 1.);
    c.reactions().add("fission: C -> A +(5) B", 2.);
    auto db = k->actions().detailedBalance(1.);
    CHECK(db->reversibleReactions().size() == 1, "one reversible pair");
    if (db->reversibleReactions().size() == 1) {
        const auto &r = db->reversibleReactions()[0];
        CHECK(r.reversibleType == model::ReversibleType::FusionFission, "FusionFission");
        CHECK(r.forwardName == "fusion" && r.backwardName == "fission", "names %s %s", r.forwardName.c_str(),
              r.backwardName.c_str());
        CHECK(close_rel(r.totalVolume, 10000., 1e-4), "total volume %g", r.totalVolume);
        CHECK(r.lhsInteractionRadius == 3. && r.reactionRadius == 5., "radii");
        CHECK(close_rel(r.lhsInteractionVolume, 113.09733, 1e-4), "lhs volume %g", r.lhsInteractionVolume);
        CHECK(close_rel(r.effectiveLhsInteractionVolume, 68.09910, 1e-4), "eff lhs int %g", r.effectiveLhsInteractionVolume);
        CHECK(close_rel(r.effectiveLhsReactionVolume, 478.60057, 1e-4), "eff lhs rx %g", r.effectiveLhsReactionVolume);
        CHECK(close_rel(r.equilibriumConstant, 41.60045, 1e-4), "K %g", r.equilibriumConstant);
        CHECK(close_rel(r.macroForwardRate, 480.76392, 1e-4), "kf %g", r.macroForwardRate);
        CHECK(r.macroBackwardRate == 2. && r.acceptancePrefactor == 1., "kb, prefactor");
    }
    CHECK(db->describe().find("A + B <--> C") != std::string::npos, "describe: %s", db->describe().c_str());
    api::SimulationLoop loop(k.get(), 0.1);
    loop.useReactionScheduler("DetailedBalance");
    bool threw = false;
    try {
        loop.useReactionScheduler("NoSuchScheduler");
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    CHECK(threw, "an unknown scheduler throws std::invalid_argument");
    c.reactions().add("anotherFusion: A +(5) B -> C", 1.);
    threw = false;
    try {
        (void)k->actions().detailedBalance(1.);
    } catch (const std::logic_error &) {
        threw = true;
    }
    CHECK(threw, "duplicate reversible pairs throw std::logic_error");
    std::printf("detailed balance configs: ok\n");
}

// Structural topology reactions through the plugin: the reference's scenes (readdy/test/TestTopologyReactions.cpp:
// 82-420) with the recipes executed on the device's topology particles
static void structural_context(model::Context &c) {
    c.boxSize() = {10., 10., 10.};
    c.particleTypes().add("Topology A", 1.0, true);
    c.particleTypes().add("Topology B", 1.0, true);
    c.particleTypes().add("Topology Invalid Type", 1.0, true);
    c.particleTypes().add("A", 1.0);
    c.topologyRegistry().configureBondPotential("Topology A", "Topology A", {10, 10});
    c.topologyRegistry().configureBondPotential("Topology A", "Topology B", {10, 10});
    c.topologyRegistry().configureBondPotential("Topology B", "Topology B", {10, 10});
}

static model::GraphTopology *small_topology(model::Kernel *k, const std::string &type) {
    const auto a = k->context().particleTypes().idOf("Topology A");
    auto *t = k->stateModel().addTopology(type, {model::Particle(0, 0, 0, a), model::Particle(0, 0, 0, a),
                                                 model::Particle(0, 0, 0, a)});
    t->addEdge(0, 1);
    t->addEdge(1, 2);
    return t;
}

static ParticleTypeId type_of(model::Kernel *k, ParticleId id) {
    for (const auto &p : k->stateModel().getParticles())
        if (p.id() == id) return p.type();
    return (ParticleTypeId)-1;
}

static void test_structural_reactions() {
    using model::Recipe;
    using model::StructuralTopologyReaction;
    {  // mode flags (:96-114)
        StructuralTopologyReaction r("r", [](model::GraphTopology &t) { return Recipe(t); }, 0.);
        r.expect_connected_after_reaction();
        CHECK(r.expects_connected_after_reaction() && !r.creates_child_topologies_after_reaction(), "mode flags");
        r.create_child_topologies_after_reaction();
        CHECK(!r.expects_connected_after_reaction() && r.creates_child_topologies_after_reaction(), "mode flags");
    }
    {  // change a particle type, constant and function rates (:116-162)
        auto k = plugin::KernelProvider::getInstance().create("MI355X");
        auto &c = k->context();
        structural_context(c);
        c.topologyRegistry().addType("Reactive Topology Type");
        auto *top = small_topology(k.get(), "Reactive Topology Type");
        const auto B = c.particleTypes().idOf("Topology B"), A = c.particleTypes().idOf("Topology A");
        StructuralTopologyReaction r1("r", [&](model::GraphTopology &t) { return Recipe(t).changeParticleType(0, B); }, 5.);
        r1.expect_connected_after_reaction();
        StructuralTopologyReaction r2("r2", [&](model::GraphTopology &t) { return Recipe(t).changeParticleType(0, A); },
                                      [](const model::GraphTopology &) { return 15.; });
        r2.expect_connected_after_reaction();
        c.topologyRegistry().addStructuralReaction("Reactive Topology Type", r1);
        c.topologyRegistry().addStructuralReaction("Reactive Topology Type", r2);
        k->actions().updateNeighborList()->perform();  // uploads (the particles coincide: no forces)
        const auto &rs = c.topologyRegistry().structuralReactionsOf("Reactive Topology Type");
        top->updateReactionRates(rs);
        CHECK(top->rates().size() == 2 && top->rates()[0] == 5 && top->rates()[1] == 15, "rates 5 and 15");
        CHECK(rs.at(0).execute(*top, k.get()).empty(), "in-place reaction returns no children");
        CHECK(type_of(k.get(), top->particleIndices()[0]) == B, "vertex 0 became Topology B");
        CHECK(rs.at(1).execute(*top, k.get()).empty(), "in-place reaction returns no children");
        CHECK(type_of(k.get(), top->particleIndices()[0]) == A, "vertex 0 is Topology A again");
    }
    {  // add an edge, remove an edge into children (:163-245)
        auto k = plugin::KernelProvider::getInstance().create("MI355X");
        auto &c = k->context();
        structural_context(c);
        c.topologyRegistry().addType("TA");
        auto *top = small_topology(k.get(), "TA");
        StructuralTopologyReaction add("add", [](model::GraphTopology &t) { return Recipe(t).addEdge(0, 2); }, 5.);
        add.expect_connected_after_reaction();
        StructuralTopologyReaction rem("rem", [](model::GraphTopology &t) { return Recipe(t).removeEdge(0, 1); }, 5.);
        k->actions().updateNeighborList()->perform();
        CHECK(add.execute(*top, k.get()).empty() && top->hasEdge(0, 2), "edge 0--2 added");
        CHECK(top->removeEdge(0, 2), "edge removed again");
        const auto parts = top->particleIndices();
        auto children = rem.execute(*top, k.get());
        CHECK(children.size() == 2, "two children after removing the edge");
        if (children.size() == 2) {
            CHECK(children[0].nParticles() == 1 && children[1].nParticles() == 2, "children sizes 1 and 2");
            CHECK(children[1].hasEdge(0, 1) && children[1].neighborsOf(0).size() == 1, "the second child keeps one edge");
            CHECK(children[0].particleIndices()[0] == parts[0] && children[1].particleIndices()[0] == parts[1] &&
                      children[1].particleIndices()[1] == parts[2],
                  "children keep their particles");
        }
    }
    for (int decay = 0; decay < 2; ++decay) {  // a 50-chain split up, with and without decay (:247-420)
        auto k = plugin::KernelProvider::getInstance().create("MI355X");
        auto &c = k->context();
        structural_context(c);
        c.topologyRegistry().addType("TA");
        const auto A = c.particleTypes().idOf("Topology A");
        std::vector<model::Particle> ps;
        for (int i = 0; i < 50; ++i) ps.emplace_back(-5 + i * 10. / 50., 0, 0, A);
        auto *top = k->stateModel().addTopology("TA", ps);
        for (int i = 0; i + 1 < 50; ++i) top->addEdge(i, i + 1);
        auto gen = std::make_shared<std::mt19937>(11 + decay);
        StructuralTopologyReaction split(
            "r",
            [gen](model::GraphTopology &t) {
                Recipe recipe(t);
                const auto n = t.nParticles();
                if (n > 1) {  // a random edge of the chain: vertices are kept in chain order
                    std::vector<std::pair<std::size_t, std::size_t>> edges;
                    for (std::size_t v = 0; v < n; ++v)
                        for (std::size_t w : t.neighborsOf(v))
                            if (v < w) edges.emplace_back(v, w);
                    if (!edges.empty()) {
                        const auto e = edges[std::uniform_int_distribution<std::size_t>(0, edges.size() - 1)(*gen)];
                        recipe.removeEdge(e.first, e.second);
                    }
                }
                return recipe;
            },
            [](const model::GraphTopology &t) { return t.nParticles() > 1 ? t.nParticles() / 50. : 0.; });
        c.topologyRegistry().addStructuralReaction("TA", split);
        if (decay) {
            const auto normal = c.particleTypes().idOf("A");
            StructuralTopologyReaction dec(
                "r2",
                [normal](model::GraphTopology &t) {
                    if (t.nParticles() != 1) throw std::logic_error("decay of a topology with more than one particle");
                    return Recipe(t).changeParticleType(0, normal);
                },
                [](const model::GraphTopology &t) { return t.nParticles() > 1 ? 0. : 1.; });
            c.topologyRegistry().addStructuralReaction("TA", dec);
        }
        auto integrator = k->actions().eulerBDIntegrator(1.0);
        auto forces = k->actions().calculateForces();
        auto topReactions = k->actions().evaluateTopologyReactions(1.0);
        forces->perform();
        for (int t = 1; t < 500; ++t) {
            integrator->perform();
            topReactions->perform();
            forces->perform();
        }
        const auto tops = k->stateModel().getTopologies();
        std::size_t nA = 0;
        for (const auto &p : k->stateModel().getParticles()) nA += p.type() == c.particleTypes().idOf("A");
        if (decay) {
            CHECK(tops.empty(), "every single particle decayed out of the topologies (%zu left)", tops.size());
            CHECK(nA == 50, "50 normal particles (%zu)", nA);
        } else {
            bool single = tops.size() == 50;
            for (auto *t : tops) single = single && t->nParticles() == 1;
            CHECK(single, "the chain split into 50 one-particle topologies (%zu topologies)", tops.size());
        }
    }
    std::printf("structural topology reactions: ok\n");
}

// DetailedBalance on the device through the plugin: the reference's integration scenes
// (TestDetailedBalance.cpp:379-462): a conversion A -> B next to 500 D particles is accepted when A
// repels D and rejected when B does
static void test_detailed_balance_scenes() {
    for (int accept = 0; accept < 2; ++accept) {
        auto k = plugin::KernelProvider::getInstance().create("MI355X");
        auto &c = k->context();
        c.boxSize() = {12., 12., 12.};
        c.particleTypes().add("A", 1.);
        c.particleTypes().add("B", 1.);
        c.particleTypes().add("D", 0.);
        c.reactions().addConversion("convAB", "A", "B", 1e15);
        c.reactions().addConversion("convBA", "B", "A", 1e-15);
        c.potentials().addHarmonicRepulsion(accept ? "A" : "B", "D", 10., 2.);
        std::mt19937 gen(5);
        std::normal_distribution<double> nd(0., 1.);
        std::vector<model::Particle> ps;
        ps.emplace_back(0., 0., 0., c.particleTypes().idOf("A"));
        for (int i = 0; i < 500; ++i) ps.emplace_back(nd(gen), nd(gen), nd(gen), c.particleTypes().idOf("D"));
        k->actions().addParticles(ps)->perform();
        auto db = k->actions().detailedBalance(0.1);
        db->perform();
        std::size_t nA = 0, nB = 0, nD = 0;
        for (const auto &p : k->stateModel().getParticles()) {
            nA += p.type() == c.particleTypes().idOf("A");
            nB += p.type() == c.particleTypes().idOf("B");
            nD += p.type() == c.particleTypes().idOf("D");
        }
        CHECK(nD == 500 && nA + nB == 1, "conservation");
        CHECK(nB == (std::size_t)accept, "conversion %s", accept ? "accepted" : "rejected");
    }
    std::printf("detailed balance scenes: ok\n");
}

// forces and energy of the C0 system against the oracle; then five loop steps event for event
static void test_c0_forces_and_loop(const std::string &) {
    const double L = 20.0;
    auto k = plugin::KernelProvider::getInstance().create("MI355X");
    c0_context(k->context(), L);
    const System s = uniform(1000, 1000, L, 7);
    std::vector<model::Particle> ps;
    for (size_t i = 0; i < s.types.size(); ++i)
        ps.emplace_back(s.pos[3 * i], s.pos[3 * i + 1], s.pos[3 * i + 2], (ParticleTypeId)s.types[i]);
    k->actions().addParticles(ps)->perform();
    k->actions().calculateForces()->perform();
    or_system *o = c0_oracle(s, L);
    double e_ref = 0;
    or_calculate_forces(o, &e_ref, nullptr);
    const auto ref = oracle_state(o);
    const auto parts = k->stateModel().getParticles();
    const auto forces = k->stateModel().getParticleForces();
    CHECK(parts.size() == s.types.size(), "particle count %zu", parts.size());
    double num = 0, den = 0;
    for (size_t i = 0; i < parts.size() && i < forces.size(); ++i) {
        const auto &fr = ref.force.at(parts[i].id());
        for (int a = 0; a < 3; ++a) {
            num += (forces[i][a] - fr[a]) * (forces[i][a] - fr[a]);
            den += fr[a] * fr[a];
        }
    }
    const double rel = std::sqrt(num / den);
    const double e_rel = std::fabs(k->stateModel().energy() - e_ref) / std::fabs(e_ref);
    if (!(rel < 1e-5)) {  // the particles that disagree most
        std::vector<std::pair<double, size_t>> d;
        for (size_t i = 0; i < parts.size() && i < forces.size(); ++i) {
            const auto &fr = ref.force.at(parts[i].id());
            d.emplace_back(std::fabs(forces[i].x - fr[0]) + std::fabs(forces[i].y - fr[1]) + std::fabs(forces[i].z - fr[2]), i);
        }
        std::sort(d.rbegin(), d.rend());
        for (size_t k2 = 0; k2 < 4 && k2 < d.size(); ++k2) {
            const size_t i = d[k2].second;
            const auto &fr = ref.force.at(parts[i].id());
            std::printf("  particle %llu at (%.4f %.4f %.4f): kernel (%.6g %.6g %.6g) oracle (%.6g %.6g %.6g)\n",
                        (unsigned long long)parts[i].id(), parts[i].pos().x, parts[i].pos().y, parts[i].pos().z,
                        forces[i].x, forces[i].y, forces[i].z, fr[0], fr[1], fr[2]);
        }
    }
    CHECK(rel < 1e-5, "force relative error %.3g", rel);
    CHECK(e_rel < 1e-5, "energy relative error %.3g", e_rel);
    std::printf("c0 forces: rel %.3g, energy rel %.3g\n", rel, e_rel);

    // SimulationLoop with a per-step callback (the action-by-action path) against or_loop_step
    // (same Philox counters, UncontrolledApproximation with keyed order)
    const std::uint64_t seed = 0;  // the kernel's Philox key (READDY_AMD_SEED unset)
    or_set_seed(o, seed);
    k->context().recordReactionsWithPositions() = true;  // records compared step by step below
    or_set_record_reactions(o, 1);
    uint64_t records_checked = 0;
    api::SimulationLoop loop(k.get(), 0.01);
    loop.useReactionScheduler("UncontrolledApproximation");
    int step_checked = 0;
    double worst = 0;
    uint64_t events = 0;
    loop.addCallback([&](std::size_t t) {
        if (t == 0) return;
        double e = 0;
        or_loop_step(o, 0.01, OR_UA_KEYED, (uint32_t)(t - 1), nullptr, 0, &e);
        const auto r = oracle_state(o);
        const auto p = k->stateModel().getParticles();
        CHECK(p.size() == r.pos.size(), "step %zu: %zu vs %zu particles", t, p.size(), r.pos.size());
        for (const auto &q : p) {
            auto it = r.pos.find(q.id());
            if (it == r.pos.end()) {
                CHECK(false, "step %zu: handle %llu missing in the oracle", t, (unsigned long long)q.id());
                continue;
            }
            CHECK(r.type.at(q.id()) == q.type(), "step %zu: type of %llu", t, (unsigned long long)q.id());
            for (int a = 0; a < 3; ++a) worst = std::max(worst, std::fabs(wrapd(q.pos()[a] - it->second[a], L)));
        }
        const auto counts = k->stateModel().reactionCounts();
        int nrc = 0;
        const uint64_t *oc = or_reaction_counts(o, &nrc);
        for (int i = 0; i < nrc && i < (int)counts.size(); ++i) {
            CHECK(counts[i] == oc[i], "step %zu: reaction %d count %llu vs %llu", t, i,
                  (unsigned long long)counts[i], (unsigned long long)oc[i]);
            events += counts[i];
        }
        // the step's reaction records: the same events with the same ids, types and positions
        auto recs = k->stateModel().reactionRecords();
        size_t no = 0;
        const or_reaction_record *orr = or_reaction_records(o, &no);
        std::vector<or_reaction_record> want(orr, orr + no);
        auto key = [](uint64_t a, uint64_t b) { return std::make_pair(a, b); };
        std::sort(recs.begin(), recs.end(), [&](const model::ReactionRecord &a, const model::ReactionRecord &b) {
            return key(a.educts[0], a.educts[1]) < key(b.educts[0], b.educts[1]);
        });
        std::sort(want.begin(), want.end(), [&](const or_reaction_record &a, const or_reaction_record &b) {
            return key(a.educts[0], a.educts[1]) < key(b.educts[0], b.educts[1]);
        });
        CHECK(recs.size() == want.size(), "step %zu: %zu records vs %zu", t, recs.size(), want.size());
        for (size_t i = 0; i < recs.size() && i < want.size(); ++i) {
            const auto &g = recs[i];
            const auto &w = want[i];
            bool same = g.type == w.type && g.id == w.reaction;
            for (int j = 0; j < 2; ++j)
                same = same && g.educts[j] == w.educts[j] && g.products[j] == w.products[j] &&
                       g.types_from[j] == w.types_from[j];
            for (int a = 0; a < 3; ++a) same = same && std::fabs(wrapd(g.where[a] - w.where[a], L)) < 1e-6;
            CHECK(same, "step %zu: record %zu differs (educts %llu %llu)", t, i, (unsigned long long)w.educts[0],
                  (unsigned long long)w.educts[1]);
            ++records_checked;
        }
        ++step_checked;
    });
    loop.run(5);
    CHECK(step_checked == 5, "callback steps %d", step_checked);
    CHECK(worst < 1e-6, "largest position difference %.3g", worst);
    CHECK(events > 20, "too few reaction events (%llu) to exercise the handler", (unsigned long long)events);
    CHECK(records_checked == events, "records %llu vs events %llu", (unsigned long long)records_checked,
          (unsigned long long)events);
    std::printf("c0 loop: 5 steps, %llu events, worst position difference %.3g\n", (unsigned long long)events, worst);
    or_destroy(o);

    // device-fused loop: conservation of A + C and B + C
    api::SimulationLoop fused(k.get(), 0.01);
    fused.useReactionScheduler("Gillespie");
    fused.run(50);
    int nA = 0, nB = 0, nC = 0;
    for (const auto &q : k->stateModel().getParticles()) (q.type() == 0 ? nA : q.type() == 1 ? nB : nC)++;
    CHECK(nA + nC == 1000 && nB + nC == 1000, "conservation: A %d B %d C %d", nA, nB, nC);
    CHECK(std::fabs(k->stateModel().time() - (5 + 50) * 0.01) < 1e-12, "simulation time %g", k->stateModel().time());
    std::printf("c0 fused Gillespie: A %d B %d C %d\n", nA, nB, nC);
}

// a bead chain topology: GraphTopology edges -> bonds / angles / dihedrals via the registry
static void test_topology() {
    const double L = 30.0;
    auto k = plugin::KernelProvider::getInstance().create("MI355X");
    auto &c = k->context();
    c.boxSize() = {L, L, L};
    c.particleTypes().addTopologyType("bead", 1.0);
    c.particleTypes().add("X", 1.0);
    c.topologyRegistry().configureBondPotential("bead", "bead", {100.0, 1.0});
    c.topologyRegistry().configureAnglePotential("bead", "bead", "bead", {10.0, M_PI});
    c.topologyRegistry().configureTorsionPotential("bead", "bead", "bead", "bead", {1.0, 3, 0.0});
    c.potentials().addHarmonicRepulsion("X", "X", 10.0, 1.0);
    c.potentials().addHarmonicRepulsion("X", "bead", 10.0, 1.0);
    const int nb = 12;
    std::mt19937_64 g(3);
    std::normal_distribution<double> nd(0, 1);
    std::vector<model::Particle> chain;
    std::vector<double> pos;
    std::vector<int> types;
    double x[3] = {0, 0, 0};
    for (int i = 0; i < nb; ++i) {
        double d[3] = {nd(g), nd(g), nd(g)};
        const double n = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int a = 0; a < 3; ++a) x[a] += i ? d[a] / n : 0.0;
        chain.emplace_back(x[0], x[1], x[2], c.particleTypes().idOf("bead"));
        pos.insert(pos.end(), x, x + 3);
        types.push_back(0);
    }
    auto *top = k->stateModel().addTopology("chain", chain);
    for (int i = 0; i + 1 < nb; ++i) top->addEdge(i, i + 1);
    std::uniform_real_distribution<double> u(-0.5 * L, 0.5 * L);
    std::vector<model::Particle> free;
    for (int i = 0; i < 300; ++i) {
        double p[3] = {u(g), u(g), u(g)};
        free.emplace_back(p[0], p[1], p[2], c.particleTypes().idOf("X"));
        pos.insert(pos.end(), p, p + 3);
        types.push_back(1);
    }
    k->stateModel().addParticles(free);
    k->stateModel().configureTopologies();
    k->actions().calculateForces()->perform();

    const double box[3] = {L, L, L};
    const int pbc[3] = {1, 1, 1};
    const double D[2] = {1.0, 1.0};
    or_system *o = or_create(2, box, pbc, 1.0, D);
    const double hr[2] = {10.0, 1.0};
    or_add_pair_potential(o, OR_POT_HARMONIC_REPULSION, 1, 1, hr);
    or_add_pair_potential(o, OR_POT_HARMONIC_REPULSION, 1, 0, hr);
    std::vector<uint64_t> ids(types.size());
    for (size_t i = 0; i < ids.size(); ++i) ids[i] = i;
    or_add_particles(o, (int)types.size(), pos.data(), types.data(), ids.data());
    for (int i = 0; i + 1 < nb; ++i) or_add_bond(o, i, i + 1, 100.0, 1.0);
    for (int i = 0; i + 2 < nb; ++i) or_add_angle(o, i, i + 1, i + 2, 10.0, M_PI);
    for (int i = 0; i + 3 < nb; ++i) or_add_dihedral(o, i, i + 1, i + 2, i + 3, 1.0, 3, 0.0);
    or_nl_setup(o, or_max_cutoff(o), 1);
    double e_ref = 0;
    or_calculate_forces(o, &e_ref, nullptr);
    const auto ref = oracle_state(o);
    const auto parts = k->stateModel().getParticles();
    const auto forces = k->stateModel().getParticleForces();
    double num = 0, den = 0;
    for (size_t i = 0; i < parts.size(); ++i) {
        const auto &fr = ref.force.at(parts[i].id());
        for (int a = 0; a < 3; ++a) {
            num += (forces[i][a] - fr[a]) * (forces[i][a] - fr[a]);
            den += fr[a] * fr[a];
        }
    }
    const double rel = std::sqrt(num / den), e_rel = std::fabs(k->stateModel().energy() - e_ref) / std::fabs(e_ref);
    CHECK(rel < 1e-5 && e_rel < 1e-5, "topology forces rel %.3g energy rel %.3g", rel, e_rel);
    std::printf("topology: %d beads + 300 crowders, force rel %.3g, energy rel %.3g\n", nb, rel, e_rel);
    or_destroy(o);
}

// BreakBonds on the reference's scenes (readdy/test/TestBreakingBonds.cpp:55-313): bonds of force
// constant 1 and length 1 stretched to 2 carry energy 1; thresholds 0.9 and 1.1 at rate 1e10 and
// dt 1 break them surely or never
static auto bb_kernel() {
    auto k = plugin::KernelProvider::getInstance().create("MI355X");
    auto &c = k->context();
    c.boxSize() = {10., 10., 10.};
    for (const char *t : {"A", "B", "C"}) c.particleTypes().addTopologyType(t, 1.0);
    return k;
}

static std::map<uint64_t, std::pair<int, Vec3>> bb_particles(model::Kernel &k) {
    std::map<uint64_t, std::pair<int, Vec3>> m;
    for (const auto &p : k.stateModel().getParticles()) m[p.id()] = {p.type(), p.pos()};
    return m;
}

static void test_break_bonds() {
    using model::BreakConfig;
    {  // a dimer A-A: breaks under threshold 0.9, not under 1.1; the bond force goes with the bond
        for (const double thr : {0.9, 1.1}) {
            auto k = bb_kernel();
            auto &c = k->context();
            c.topologyRegistry().configureBondPotential("A", "A", {1., 1.});
            const auto A = c.particleTypes().idOf("A");
            auto *t = k->stateModel().addTopology("T", {{0., 0., 0., A}, {0., 0., 2., A}});
            t->addEdge(0, 1);
            k->actions().calculateForces()->perform();
            const auto f0 = k->stateModel().getParticleForces();
            CHECK(f0.size() == 2 && std::fabs(std::fabs(f0[0].z) - 2.0) < 1e-12, "dimer bond force %g", f0[0].z);
            BreakConfig cfg;
            cfg.addBreakablePair(A, A, thr, 1e10);
            k->actions().breakBonds(1., cfg)->perform();
            const auto tops = k->stateModel().getTopologies();
            if (thr < 1.) {
                CHECK(tops.size() == 2 && tops[0]->nParticles() == 1 && tops[1]->nParticles() == 1,
                      "dimer: %zu topologies after breaking", tops.size());
            } else {
                CHECK(tops.size() == 1 && tops[0]->nParticles() == 2 && tops[0]->type() == "T",
                      "dimer: %zu topologies kept", tops.size());
            }
            k->actions().calculateForces()->perform();
            const auto f1 = k->stateModel().getParticleForces();
            const double want = thr < 1. ? 0.0 : 2.0;
            CHECK(f1.size() == 2 && std::fabs(std::fabs(f1[0].z) - want) < 1e-12 &&
                      std::fabs(std::fabs(f1[1].z) - want) < 1e-12,
                  "dimer force after break bonds %g (want %g)", f1.size() ? f1[0].z : -1., want);
            const auto ps = bb_particles(*k);
            CHECK(ps.size() == 2 && ps.at(0).second.z == 0. && ps.at(1).second.z == 2., "dimer particles moved");
        }
        bool threw = false;
        try {
            BreakConfig cfg;
            cfg.addBreakablePair(0, 0, 1., 0.);
        } catch (const std::logic_error &) {
            threw = true;
        }
        CHECK(threw, "a non-positive break rate must throw std::logic_error");
    }
    // a linear trimer A(z=-1) - A(z=1) - B(z=3)
    for (int which = 0; which < 3; ++which) {  // A-B breakable, A-A breakable, both with threshold 1.1
        auto k = bb_kernel();
        auto &c = k->context();
        c.topologyRegistry().configureBondPotential("A", "A", {1., 1.});
        c.topologyRegistry().configureBondPotential("B", "A", {1., 1.});
        const auto A = c.particleTypes().idOf("A"), B = c.particleTypes().idOf("B");
        auto *t = k->stateModel().addTopology("T", {{0., 0., -1., A}, {0., 0., 1., A}, {0., 0., 3., B}});
        t->addEdge(0, 1);
        t->addEdge(1, 2);
        BreakConfig cfg;
        if (which == 0) cfg.addBreakablePair(A, B, 0.9, 1e10);
        if (which == 1) cfg.addBreakablePair(A, A, 0.9, 1e10);
        if (which == 2) {
            cfg.addBreakablePair(A, A, 1.1, 1e10);
            cfg.addBreakablePair(A, B, 1.1, 1e10);
        }
        k->actions().breakBonds(1., cfg)->perform();
        const auto tops = k->stateModel().getTopologies();
        const auto ps = bb_particles(*k);
        if (which == 2) {
            CHECK(tops.size() == 1 && tops[0]->nParticles() == 3, "trimer kept: %zu topologies", tops.size());
        } else {
            CHECK(tops.size() == 2, "trimer: %zu topologies", tops.size());
            const auto *pair = tops.size() == 2 ? (tops[0]->nParticles() == 2 ? tops[0] : tops[1]) : nullptr;
            CHECK(pair && pair->nParticles() == 2, "trimer: no two-particle topology");
            if (pair) {
                std::vector<int> ty;
                for (auto i : pair->particleIndices()) ty.push_back(ps.at(i).first);
                std::sort(ty.begin(), ty.end());
                const std::vector<int> want = which == 0 ? std::vector<int>{A, A} : std::vector<int>{A, B};
                CHECK(ty == want, "trimer case %d: wrong particles in the two-particle topology", which);
            }
        }
        CHECK(ps.size() == 3 && ps.at(0).second.z == -1. && ps.at(1).second.z == 1. && ps.at(2).second.z == 3.,
              "trimer particles moved");
    }
    {  // three dimers A-A, B-B, C-C broken one after the other: six one-particle topologies
        auto k = bb_kernel();
        auto &c = k->context();
        for (const char *n : {"A", "B", "C"}) c.topologyRegistry().configureBondPotential(n, n, {1., 1.});
        for (const char *n : {"A", "B", "C"}) {
            const auto T = c.particleTypes().idOf(n);
            k->stateModel().addTopology("T", {{0., 0., 0., T}, {0., 0., 2., T}})->addEdge(0, 1);
        }
        BreakConfig cfg;
        for (const char *n : {"A", "B", "C"}) {
            const auto T = c.particleTypes().idOf(n);
            cfg.addBreakablePair(T, T, 0.9, 1e10);
            k->actions().breakBonds(1., cfg)->perform();
        }
        const auto tops = k->stateModel().getTopologies();
        CHECK(tops.size() == 6, "three dimers: %zu topologies", tops.size());
        for (const auto *tp : tops) CHECK(tp->nParticles() == 1, "three dimers: a topology of %zu", tp->nParticles());
    }
    std::printf("break bonds: reference scenes passed\n");
}

// host core only (no kernel): edge removal and the component split BreakBonds relies on, on a graph
// with a cycle and a branch, against the same components built edge by edge
using Tuples = std::vector<std::vector<std::size_t>>;
static Tuples enumerate(const model::GraphTopology &t) {
    Tuples out;
    t.findNTuples([&](std::size_t a, std::size_t b) { out.push_back({a, b}); },
                  [&](std::size_t a, std::size_t b, std::size_t c) { out.push_back({a, b, c}); },
                  [&](std::size_t a, std::size_t b, std::size_t c, std::size_t d) { out.push_back({a, b, c, d}); });
    return out;
}

static void test_graph_ops() {
    // particles 10..17; triangle 0-1-2, bridge 2-3, branch 3-4-5 and 3-6-7
    model::GraphTopology g("T", {10, 11, 12, 13, 14, 15, 16, 17});
    for (auto [a, b] : std::vector<std::pair<int, int>>{{0, 1}, {1, 2}, {2, 0}, {2, 3}, {3, 4}, {4, 5}, {3, 6}, {6, 7}})
        g.addEdge(a, b);
    CHECK(!g.removeEdge(0, 5), "removing a missing edge must report false");
    CHECK(g.connectedComponents().size() == 1, "one component before any removal");
    CHECK(g.removeEdge(0, 1) && !g.hasEdge(1, 0), "edge 0-1 removed in both directions");
    CHECK(g.connectedComponents().size() == 1, "a cycle edge does not split");
    CHECK(g.removeEdge(3, 2), "bridge removed (given in the other order)");
    const auto parts = g.connectedComponents();
    CHECK(parts.size() == 2, "the bridge splits into %zu components", parts.size());
    if (parts.size() == 2) {
        CHECK((parts[0].particleIndices() == std::vector<std::size_t>{10, 11, 12}), "first component's particles");
        CHECK((parts[1].particleIndices() == std::vector<std::size_t>{13, 14, 15, 16, 17}), "second component's particles");
        model::GraphTopology a("T", {10, 11, 12}), b("T", {13, 14, 15, 16, 17});
        a.addEdge(1, 2);
        a.addEdge(2, 0);
        b.addEdge(0, 1);
        b.addEdge(1, 2);
        b.addEdge(0, 3);
        b.addEdge(3, 4);
        CHECK(enumerate(parts[0]) == enumerate(a), "first component's bonds / angles / dihedrals");
        CHECK(enumerate(parts[1]) == enumerate(b), "second component's bonds / angles / dihedrals");
        CHECK(enumerate(parts[1]).size() == 4 + 3 + 2, "branch: 4 edges, 3 angles, 2 dihedrals");
        CHECK(parts[0].type() == "T" && parts[1].type() == "T", "components keep the topology type");
    }
    for (const double rate : {0.0, -1.0}) {
        bool threw = false;
        try {
            model::BreakConfig c;
            c.addBreakablePair(0, 1, 1.0, rate);
        } catch (const std::logic_error &) {
            threw = true;
        }
        CHECK(threw, "break rate %g must throw", rate);
    }
    model::BreakConfig c;
    c.addBreakablePair(0, 1, 0.5, 2.0);
    CHECK(c.pairs().size() == 1 && c.pairs()[0].rate == 2.0, "breakable pair kept");
    std::printf("graph operations: passed\n");
}

// the reaction descriptor form on the reference's own cases (readdy/test/TestContext.cpp:222-345)
static void test_descriptors() {
    using RT = model::ReactionType;
    model::Context c;
    for (const char *t : {"A", "B", "C"}) c.particleTypes().add(t, 1.);
    auto &reg = c.reactions();
    const auto A = c.particleTypes().idOf("A"), B = c.particleTypes().idOf("B"), C = c.particleTypes().idOf("C");
    auto get = [&](const char *d, double rate) { return reg.reactions().at(reg.add(d, rate)); };
    auto r = get("mydecay:A->", 2.);
    CHECK(r.type == RT::Decay && r.educts == std::vector<ParticleTypeId>{A} && r.products.empty() && r.rate == 2.,
          "decay descriptor");
    r = get("myconv: A -> B", 3.);
    CHECK(r.type == RT::Conversion && r.educts[0] == A && r.products[0] == B && r.rate == 3., "conversion descriptor");
    r = get("myfus: B +(1.2) B -> C [0.5, 0.5]", 4.);
    CHECK(r.type == RT::Fusion && r.educts == (std::vector<ParticleTypeId>{B, B}) && r.products[0] == C &&
              r.eductDistance == 1.2 && r.weight1 == 0.5 && r.weight2 == 0.5 && r.rate == 4.,
          "fusion descriptor");
    r = get("myfiss: B -> C +(3.0) B [0.1, 0.9]", 5.);
    CHECK(r.type == RT::Fission && r.educts[0] == B && r.products == (std::vector<ParticleTypeId>{C, B}) &&
              r.productDistance == 3.0 && r.weight1 == 0.1 && r.weight2 == 0.9 && r.rate == 5.,
          "fission descriptor");
    r = get("myenz:A +(1.5) C -> B + C", 6.);
    CHECK(r.type == RT::Enzymatic && r.educts == (std::vector<ParticleTypeId>{A, C}) &&
              r.products == (std::vector<ParticleTypeId>{B, C}) && r.eductDistance == 1.5 && r.rate == 6.,
          "enzymatic descriptor");
    r = get("eat:B +(1) A -> A + A", 7.);
    CHECK(r.type == RT::Enzymatic && r.educts == (std::vector<ParticleTypeId>{B, A}) &&
              r.products == (std::vector<ParticleTypeId>{A, A}) && r.eductDistance == 1. && r.rate == 7.,
          "predator-prey descriptor");
    for (const char *bad : {"myinvalid: + A -> B", "noarrow: A B", " : Noname ->", "weights: A + A -> A [0.1, ]",
                            "blub: A (3)+ A -> B", "eat: A +(1) B -> A + A", "two: A -> B -> C", "mydecay: B ->"}) {
        bool threw = false;
        try {
            reg.add(bad, 1.);
        } catch (const std::invalid_argument &) {
            threw = true;
        }
        CHECK(threw, "descriptor \"%s\" must throw std::invalid_argument", bad);
    }
    std::printf("reaction descriptors: passed\n");
}

// spatial topology reaction descriptors on the reference's cases (readdy/test/TestContext.cpp:100-160)
// and the validation of TopologyRegistry::validateSpatialReaction (TopologyRegistry.cpp:238-293)
static void test_spatial_descriptors() {
    using M = model::STRMode;
    model::Context c;
    for (const char *t : {"A", "B", "C"}) c.particleTypes().addTopologyType(t, 1.);
    c.particleTypes().add("P", 1.);
    c.particleTypes().add("Q", 1.);
    auto &reg = c.topologyRegistry();
    CHECK(reg.addType("T") == 0 && reg.addType("TT") == 1 && reg.nameOf(1) == "TT", "topology types in order");
    reg.addSpatialReaction("MyFusionReaction: T(A)+T(B) -> TT(B--A)", 10, 20);
    reg.addSpatialReaction("MySelfFusionReaction: T(A)+T(B) -> TT(B--A) [self=true]", 10, 20);
    reg.addSpatialReaction("MyTPFusion: T(A)+(P) -> TT(B--B)", 10, 20);
    reg.addSpatialReaction("MyEnzymaticTT: T(A)+T(B) -> TT(B) + TT(A)", 10, 20);
    reg.addSpatialReaction("MyEnzymaticTP: T(A) + (P) -> TT(B) + (Q)", 10, 20);
    const auto &sr = reg.spatialReactions();
    CHECK(sr.size() == 5, "five spatial reactions");
    const auto id = [&](const char *n) { return c.particleTypes().idOf(n); };
    CHECK(reg.isSpatialReactionType(id("A")) && !reg.isSpatialReactionType(id("C")), "spatial reaction types");
    const auto by = [&](const std::string &n) -> const model::SpatialReactionDesc & {
        for (const auto &r : sr)
            if (r.name == n) return r;
        throw std::invalid_argument("no reaction " + n);
    };
    const auto &tp = by("MyEnzymaticTP");
    CHECK(tp.mode == M::TP_ENZYMATIC && tp.type1 == id("A") && tp.type2 == id("P") && tp.topType1 == 0 &&
              tp.topTypeTo1 == 1 && tp.typeTo1 == id("B") && tp.typeTo2 == id("Q") && tp.rate == 10 &&
              tp.radius == 20 && tp.isTopologyParticle(),
          "TP_ENZYMATIC descriptor");
    const auto &tt = by("MyEnzymaticTT");
    CHECK(tt.mode == M::TT_ENZYMATIC && tt.type1 == id("A") && tt.type2 == id("B") && tt.topType1 == 0 &&
              tt.topType2 == 0 && tt.topTypeTo1 == 1 && tt.topTypeTo2 == 1 && tt.typeTo1 == id("B") &&
              tt.typeTo2 == id("A") && !tt.isTopologyParticle(),
          "TT_ENZYMATIC descriptor");
    const auto &tf = by("MyFusionReaction");
    CHECK(tf.mode == M::TT_FUSION && tf.typeTo1 == id("B") && tf.typeTo2 == id("A") && tf.topTypeTo1 == 1,
          "TT_FUSION descriptor");
    CHECK(by("MySelfFusionReaction").mode == M::TT_FUSION_ALLOW_SELF, "TT_FUSION_ALLOW_SELF descriptor");
    const auto &pf = by("MyTPFusion");
    CHECK(pf.mode == M::TP_FUSION && pf.typeTo1 == id("B") && pf.typeTo2 == id("B") && pf.topTypeTo1 == 1,
          "TP_FUSION descriptor");
    for (const char *bad : {"noarrow: T(A) + (P)", "two: T(A) + (P) -> T(A) + (P) -> T(A) + (P)",
                            "nocolon T(A) + (P) -> T(B) + (Q)", "noplus: T(A) -> T(B) + (Q)",
                            "nobond: T(A) + T(B) -> TT(A)", "unknowntop: X(A) + (P) -> T(B) + (Q)",
                            "unknowntype: T(Z) + (P) -> T(B) + (Q)", "MyEnzymaticTP: T(A) + (P) -> TT(B) + (Q)",
                            "flavor: T(A) + (P) -> T(P) + (Q)", "notopology: T(P) + (Q) -> T(Q) + (P)",
                            "fusionflavor: T(A) + (P) -> TT(B--P)"}) {
        bool threw = false;
        try {
            reg.addSpatialReaction(bad, 1., 1.);
        } catch (const std::invalid_argument &) {
            threw = true;
        }
        CHECK(threw, "spatial descriptor \"%s\" must throw std::invalid_argument", bad);
    }
    bool threw = false;
    try {
        reg.addSpatialReaction("radius: T(A) + (P) -> T(B) + (Q)", 1., 0.);
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    CHECK(threw, "a non-positive radius throws");
    threw = false;
    try {
        reg.addSpatialReaction("opt: T(A) + T(B) -> TT(A--B) [distance=3]", 1., 1.);
    } catch (const std::runtime_error &) {
        threw = true;
    }
    CHECK(threw, "an option block without self=true throws");
    CHECK(c.calculateMaxCutoff() == 20, "Context::calculateMaxCutoff includes the spatial radii");
    std::printf("spatial topology reaction descriptors: passed\n");
}

// A chain scene for the topology-particle enzymatic reaction through the plugin: 10 chains of 4 beads
// (type "Chain", head first), an X next to the head of chains 0-4 and 5 more X far from every chain.
// bind: Chain(head) + (X) -> Bound(head2) + (Y); each reaction turns one Chain into Bound and one X into Y,
// and a Bound topology matches no reaction, so #Bound == #Y at any time.
static plugin::KernelProvider::kernel_ptr tr_kernel(bool retype) {
    auto k = plugin::KernelProvider::getInstance().create("MI355X");
    auto &c = k->context();
    c.boxSize() = {40., 40., 40.};
    c.particleTypes().addTopologyType("head", 0.1);
    c.particleTypes().addTopologyType("bead", 0.1);
    c.particleTypes().addTopologyType("head2", 0.1);
    c.particleTypes().add("X", 1.);
    c.particleTypes().add("Y", 1.);
    auto &reg = c.topologyRegistry();
    reg.addType("Chain");
    reg.addType("Bound");
    reg.configureBondPotential("head", "bead", {10., 1.});
    reg.configureBondPotential("bead", "bead", {10., 1.});
    reg.configureBondPotential("head2", "bead", {20., 1.5});
    c.potentials().addHarmonicRepulsion("X", "bead", 10., 1.);
    reg.addSpatialReaction(retype ? "bind: Chain(head) + (X) -> Bound(head2) + (Y)"
                                  : "bind: Chain(head) + (X) -> Bound(head) + (Y)",
                           1e10, 1.5);
    auto &sm = k->stateModel();
    const auto head = c.particleTypes().idOf("head"), bead = c.particleTypes().idOf("bead");
    std::vector<model::Particle> xs;
    for (int t = 0; t < 10; ++t) {
        const double y = -15. + 3. * t;
        std::vector<model::Particle> ps;
        for (int i = 0; i < 4; ++i) ps.emplace_back(-5. + i, y, 0., i ? bead : head);
        auto *top = sm.addTopology("Chain", ps);
        for (int i = 0; i < 3; ++i) top->addEdge(i, i + 1);
        xs.emplace_back(t < 5 ? -5.8 : 12., t < 5 ? y : -15. + 6. * (t - 5), t < 5 ? 0.3 : 10., c.particleTypes().idOf("X"));
    }
    sm.addParticles(xs);
    return k;
}

static void tr_counts(model::Kernel &k, int &bound, int &y, int &x, int &h2) {
    bound = y = x = h2 = 0;
    for (auto *t : k.stateModel().getTopologies()) bound += t->type() == "Bound";
    const auto &pt = k.context().particleTypes();
    for (const auto &p : k.stateModel().getParticles()) {
        y += p.type() == pt.idOf("Y");
        x += p.type() == pt.idOf("X");
        h2 += p.type() == pt.idOf("head2");
    }
}

static void test_topology_reactions() {
    for (int retype = 0; retype < 2; ++retype) {
        auto k = tr_kernel(retype != 0);
        k->actions().calculateForces()->perform();
        k->actions().evaluateTopologyReactions(0.01)->perform();
        int bound, y, x, h2;
        tr_counts(*k, bound, y, x, h2);
        // every X next to a head is within the radius, the rate makes the reaction sure, and no other pair
        // is within reach: chains 0-4 react exactly once each
        CHECK(bound == 5 && y == 5 && x == 5 && h2 == (retype ? 5 : 0),
              "one evaluation: Bound %d, Y %d, X %d, head2 %d (retype %d)", bound, y, x, h2, retype);
        const auto tops = k->stateModel().getTopologies();
        for (int t = 0; t < 10; ++t)
            CHECK(tops[t]->type() == (t < 5 ? "Bound" : "Chain"), "topology %d type %s", t, tops[t]->type().c_str());
        if (retype) {
            // the retyped heads carry the head2 bond (k 20, length 1.5): compressed by 0.5 it holds 20 * 0.25
            // (the other bonds are at rest); the energy is formed on the state the reactions left
            k->actions().calculateForces()->perform();
            const double e2 = k->stateModel().energy();
            CHECK(std::fabs(e2 - 5 * 20. * 0.25) < 1e-6 * 5 * 20 * 0.25,
                  "bonded terms set again after the retyping: energy %.9g", e2);
        }
        // a loop of 20 steps: fused on the device when the bonded terms stay, step by step otherwise
        api::SimulationLoop loop(k.get(), 0.01);
        loop.run(20);
        tr_counts(*k, bound, y, x, h2);
        CHECK(bound == y && x + y == 10 && bound >= 5 && h2 == (retype ? bound : 0),
              "after 20 steps: Bound %d, Y %d, X %d, head2 %d (retype %d)", bound, y, x, h2, retype);
        std::printf("topology reactions (retype %d): Bound %d, Y %d after 20 steps\n", retype, bound, y);
    }
}

// Topology-particle fusion through the plugin ("grow: Polymer(end) + (mono) -> Polymer(mid--end)", rate 1e10):
// five chains mid-mid-end, each with a monomer 1.1 from its end and a second one 1.1 beyond it (2.2 from the
// end, outside the radius 1.2).  One evaluation grows every chain by its first monomer (the end becomes mid,
// the monomer the new end, under a topology handle after the 15 chain particles), the next by the second;
// the new edges carry the mid-end bond (k 10, length 1), stretched by 0.1: 10 * 0.01 per join.
static void test_topology_fusion() {
    auto k = plugin::KernelProvider::getInstance().create("MI355X");
    auto &c = k->context();
    c.boxSize() = {40., 40., 40.};
    c.particleTypes().addTopologyType("mid", 0.1);
    c.particleTypes().addTopologyType("end", 0.1);
    c.particleTypes().add("mono", 1.);
    auto &reg = c.topologyRegistry();
    reg.addType("Polymer");
    reg.configureBondPotential("mid", "mid", {10., 1.});
    reg.configureBondPotential("mid", "end", {10., 1.});
    reg.addSpatialReaction("grow: Polymer(end) + (mono) -> Polymer(mid--end)", 1e10, 1.2);
    const auto mid = c.particleTypes().idOf("mid"), end = c.particleTypes().idOf("end"),
               mono = c.particleTypes().idOf("mono");
    auto &sm = k->stateModel();
    std::vector<model::Particle> monos;
    for (int t = 0; t < 5; ++t) {
        const double y = -10. + 4. * t;
        std::vector<model::Particle> ps;
        for (int i = 0; i < 3; ++i) ps.emplace_back((double)i, y, 0., i == 2 ? end : mid);
        auto *top = sm.addTopology("Polymer", ps);
        top->addEdge(0, 1);
        top->addEdge(1, 2);
        monos.emplace_back(3.1, y, 0., mono);
        monos.emplace_back(4.2, y, 0., mono);
    }
    sm.addParticles(monos);
    k->actions().calculateForces()->perform();
    CHECK(std::fabs(k->stateModel().energy()) < 1e-9, "chains at rest: energy %.3g", k->stateModel().energy());
    for (int round = 1; round <= 2; ++round) {
        k->actions().evaluateTopologyReactions(0.01)->perform();
        int n_mid = 0, n_end = 0, n_mono = 0;
        std::size_t max_topo_id = 0;
        for (const auto &p : sm.getParticles()) {
            n_mid += p.type() == mid;
            n_end += p.type() == end;
            n_mono += p.type() == mono;
            if (p.type() != mono) max_topo_id = std::max<std::size_t>(max_topo_id, p.id());
        }
        CHECK(n_mid == 5 * (2 + round) && n_end == 5 && n_mono == 10 - 5 * round,
              "fusion round %d: mid %d end %d mono %d", round, n_mid, n_end, n_mono);
        CHECK(max_topo_id == 15u + 5u * round - 1u, "round %d: joined particles take handles up to %zu", round,
              max_topo_id);
        for (auto *t : sm.getTopologies())
            CHECK(t->nParticles() == (std::size_t)(3 + round) && t->isConnected() && t->type() == "Polymer",
                  "round %d: a chain of %zu particles", round, t->nParticles());
        k->actions().calculateForces()->perform();
        const double e = k->stateModel().energy();
        CHECK(std::fabs(e - 5 * round * 10. * 0.01) < 1e-6, "round %d: bond energy %.9g", round, e);
    }
    api::SimulationLoop loop(k.get(), 0.001);
    loop.run(10);
    CHECK(sm.getParticles().size() == 25u && sm.getTopologies().size() == 5u, "after 10 steps: %zu particles",
          sm.getParticles().size());
    std::printf("topology fusion: 5 chains grew by two monomers each\n");
}

// Topology-topology enzymatic through the plugin ("swap: Left(end) + Right(end) -> Left2(done) + Right2(done)",
// rate 1e10): three pairs of two-bead chains whose ends are 0.6 apart react, a fourth pair 3 apart does not;
// the reacted topologies take the new types, the ends the new species.
static void test_topology_topology_enzymatic() {
    auto k = plugin::KernelProvider::getInstance().create("MI355X");
    auto &c = k->context();
    c.boxSize() = {30., 30., 30.};
    c.particleTypes().addTopologyType("bead", 0.1);
    c.particleTypes().addTopologyType("end", 0.1);
    c.particleTypes().addTopologyType("done", 0.1);
    auto &reg = c.topologyRegistry();
    for (const char *t : {"Left", "Right", "Left2", "Right2"}) reg.addType(t);
    reg.configureBondPotential("bead", "end", {10., 1.});
    reg.configureBondPotential("bead", "done", {10., 1.});
    reg.addSpatialReaction("swap: Left(end) + Right(end) -> Left2(done) + Right2(done)", 1e10, 1.0);
    const auto bead = c.particleTypes().idOf("bead"), end = c.particleTypes().idOf("end"),
               done = c.particleTypes().idOf("done");
    auto &sm = k->stateModel();
    for (int p = 0; p < 4; ++p) {
        const double y = -9. + 6. * p, gap = p < 3 ? 0.6 : 3.0;
        auto *l = sm.addTopology("Left", {model::Particle(-1., y, 0., bead), model::Particle(0., y, 0., end)});
        l->addEdge(0, 1);
        auto *r = sm.addTopology("Right", {model::Particle(gap + 1., y, 0., bead), model::Particle(gap, y, 0., end)});
        r->addEdge(0, 1);
    }
    k->actions().calculateForces()->perform();
    k->actions().evaluateTopologyReactions(0.01)->perform();
    const auto tops = sm.getTopologies();
    int left2 = 0, right2 = 0, n_done = 0;
    for (auto *t : tops) {
        left2 += t->type() == "Left2";
        right2 += t->type() == "Right2";
    }
    for (const auto &p : sm.getParticles()) n_done += p.type() == done;
    CHECK(left2 == 3 && right2 == 3 && n_done == 6, "topology-topology: Left2 %d Right2 %d done %d", left2, right2,
          n_done);
    CHECK(tops[6]->type() == "Left" && tops[7]->type() == "Right", "the distant pair stays: %s %s",
          tops[6]->type().c_str(), tops[7]->type().c_str());
    k->actions().calculateForces()->perform();
    CHECK(std::fabs(sm.energy()) < 1e-9, "bonds configured again at rest: energy %.3g", sm.energy());
    std::printf("topology-topology enzymatic: 3 of 4 pairs reacted\n");
}

// Topology-topology fusion through the plugin ("merge: Left(end) + Right(end) -> Joined(mid--mid)", rate 1e10):
// three pairs of two-bead chains whose ends are 0.6 apart merge into one four-bead topology each (the distant
// fourth pair stays two topologies); the new bond (mid-mid, k 10, length 1) is stretched to 0.6: 10 * 0.4^2 each.
static void test_topology_topology_fusion() {
    auto k = plugin::KernelProvider::getInstance().create("MI355X");
    auto &c = k->context();
    c.boxSize() = {30., 30., 30.};
    c.particleTypes().addTopologyType("bead", 0.1);
    c.particleTypes().addTopologyType("end", 0.1);
    c.particleTypes().addTopologyType("mid", 0.1);
    auto &reg = c.topologyRegistry();
    for (const char *t : {"Left", "Right", "Joined"}) reg.addType(t);
    reg.configureBondPotential("bead", "end", {10., 1.});
    reg.configureBondPotential("bead", "mid", {10., 1.});
    reg.configureBondPotential("mid", "mid", {10., 1.});
    reg.addSpatialReaction("merge: Left(end) + Right(end) -> Joined(mid--mid)", 1e10, 1.0);
    const auto bead = c.particleTypes().idOf("bead"), end = c.particleTypes().idOf("end");
    auto &sm = k->stateModel();
    for (int p = 0; p < 4; ++p) {
        const double y = -9. + 6. * p, gap = p < 3 ? 0.6 : 3.0;
        auto *l = sm.addTopology("Left", {model::Particle(-1., y, 0., bead), model::Particle(0., y, 0., end)});
        l->addEdge(0, 1);
        auto *r = sm.addTopology("Right", {model::Particle(gap + 1., y, 0., bead), model::Particle(gap, y, 0., end)});
        r->addEdge(0, 1);
    }
    k->actions().calculateForces()->perform();
    k->actions().evaluateTopologyReactions(0.01)->perform();
    const auto tops = sm.getTopologies();
    int joined = 0;
    for (auto *t : tops)
        if (t->type() == "Joined") {
            ++joined;
            CHECK(t->nParticles() == 4u && t->isConnected(), "a merged topology of %zu particles", t->nParticles());
        }
    CHECK(tops.size() == 5u && joined == 3, "topology-topology fusion: %zu topologies, %d joined", tops.size(), joined);
    k->actions().calculateForces()->perform();
    const double e = sm.energy(), want = 3 * 10. * 0.16;
    CHECK(std::fabs(e - want) < 1e-6 * want, "merged bonds: energy %.9g, expected %.9g", e, want);
    k->actions().evaluateTopologyReactions(0.01)->perform();
    CHECK(sm.getTopologies().size() == 5u, "nothing left to merge: %zu topologies", sm.getTopologies().size());
    std::printf("topology-topology fusion: 3 of 4 pairs merged\n");
}

// Topology-topology self-fusion ("[self=true]"): a chain end(0,0,0)-bead(1,0,0)-bead(1,0.9,0)-end(0,0.9,0) and a
// bonded pair of ends 0.5 apart, bonds k 10, length 1.  One evaluation closes the chain into a ring of four
// (its ends 0.9 apart are not bonded) and leaves the bonded pair alone: two topologies, the bond energy
// 10 * 0.1^2 twice plus 10 * 0.5^2; "distance>2" is refused.
static void test_topology_self_fusion() {
    auto k = plugin::KernelProvider::getInstance().create("MI355X");
    auto &c = k->context();
    c.boxSize() = {30., 30., 30.};
    c.particleTypes().addTopologyType("bead", 0.1);
    c.particleTypes().addTopologyType("end", 0.1);
    c.particleTypes().addTopologyType("mid", 0.1);
    auto &reg = c.topologyRegistry();
    for (const char *t : {"Chain", "Ring"}) reg.addType(t);
    for (const char *a : {"bead", "end", "mid"})
        for (const char *b : {"bead", "end", "mid"}) reg.configureBondPotential(a, b, {10., 1.});
    reg.addSpatialReaction("ring: Chain(end) + Chain(end) -> Ring(mid--mid) [self=true]", 1e10, 1.0);
    const auto bead = c.particleTypes().idOf("bead"), end = c.particleTypes().idOf("end");
    auto &sm = k->stateModel();
    auto *u = sm.addTopology("Chain", {model::Particle(0., 0., 0., end), model::Particle(1., 0., 0., bead),
                                       model::Particle(1., 0.9, 0., bead), model::Particle(0., 0.9, 0., end)});
    u->addEdge(0, 1);
    u->addEdge(1, 2);
    u->addEdge(2, 3);
    auto *p = sm.addTopology("Chain", {model::Particle(5., 0., 0., end), model::Particle(5.5, 0., 0., end)});
    p->addEdge(0, 1);
    k->actions().calculateForces()->perform();
    k->actions().evaluateTopologyReactions(0.01)->perform();
    const auto tops = sm.getTopologies();
    int rings = 0;
    for (auto *t : tops)
        if (t->type() == "Ring") {
            ++rings;
            CHECK(t->nParticles() == 4u, "the ring holds %zu particles", t->nParticles());
        }
    CHECK(tops.size() == 2u && rings == 1, "self-fusion: %zu topologies, %d rings", tops.size(), rings);
    k->actions().calculateForces()->perform();
    const double e = sm.energy(), want = 2 * 10. * 0.01 + 10. * 0.25;
    CHECK(std::fabs(e - want) < 1e-6 * want, "ring bonds: energy %.9g, expected %.9g", e, want);
    auto k2 = plugin::KernelProvider::getInstance().create("MI355X");
    auto &c2 = k2->context();
    c2.boxSize() = {30., 30., 30.};
    c2.particleTypes().addTopologyType("end", 0.1);
    c2.topologyRegistry().addType("Chain");
    c2.topologyRegistry().configureBondPotential("end", "end", {10., 1.});
    c2.topologyRegistry().addSpatialReaction("far: Chain(end) + Chain(end) -> Chain(end--end) [self=true, distance>2]",
                                             1e10, 1.0);
    auto *q = k2->stateModel().addTopology("Chain", {model::Particle(0., 0., 0., c2.particleTypes().idOf("end"))});
    (void)q;
    bool refused = false;
    try {
        k2->actions().calculateForces()->perform();
        k2->actions().evaluateTopologyReactions(0.01)->perform();
    } catch (const std::exception &) {
        refused = true;
    }
    CHECK(refused, "a minimum graph distance above 1 is refused");
    std::printf("topology self-fusion: one ring closed, the bonded pair left alone\n");
}

// Recipe steps that move and append topology particles (Recipe::changeParticlePosition, appendNewParticle): three
// chains of three beads 1 apart with bonds (k 10, length 1); a structural reaction (rate 1e10 while a chain has
// fewer than four particles) moves bead 0 by 0.3 away from bead 1 and appends a bead 1.2 beyond bead 2, bonded to
// it.  One evaluation: four particles per chain, the new ones under topology handles 9..11 (after the nine
// uploaded ones), the positions as the recipe set them, and the bond energy 10 * 0.3^2 + 10 * 0.2^2 per chain;
// the next evaluation does nothing (rate 0).
static void test_structural_move_and_append() {
    using model::Recipe;
    using model::StructuralTopologyReaction;
    auto k = plugin::KernelProvider::getInstance().create("MI355X");
    auto &c = k->context();
    c.boxSize() = {30., 30., 30.};
    c.particleTypes().addTopologyType("bead", 0.1);
    auto &reg = c.topologyRegistry();
    reg.addType("Chain");
    reg.configureBondPotential("bead", "bead", {10., 1.});
    const auto bead = c.particleTypes().idOf("bead");
    std::vector<double> ys;
    StructuralTopologyReaction grow(
        "grow",
        [&](model::GraphTopology &t) {
            const auto ps = k->stateModel().getParticles();
            Vec3 p0{}, p2{};
            for (const auto &p : ps) {
                if (p.id() == t.particleIndices()[0]) p0 = p.pos();
                if (p.id() == t.particleIndices()[2]) p2 = p.pos();
            }
            return Recipe(t).changeParticlePosition(0, Vec3(p0.x - 0.3, p0.y, p0.z))
                .appendNewParticle({2}, "bead", Vec3(p2.x + 1.2, p2.y, p2.z));
        },
        [](const model::GraphTopology &t) { return t.nParticles() < 4 ? 1e10 : 0.; });
    grow.expect_connected_after_reaction();
    reg.addStructuralReaction("Chain", grow);
    auto &sm = k->stateModel();
    for (int t = 0; t < 3; ++t) {
        const double y = -6. + 6. * t;
        ys.push_back(y);
        std::vector<model::Particle> ps;
        for (int i = 0; i < 3; ++i) ps.emplace_back((double)i, y, 0., bead);
        auto *top = sm.addTopology("Chain", ps);
        top->addEdge(0, 1);
        top->addEdge(1, 2);
    }
    k->actions().calculateForces()->perform();
    CHECK(std::fabs(sm.energy()) < 1e-9, "chains at rest: energy %.3g", sm.energy());
    k->actions().evaluateTopologyReactions(0.01)->perform();
    const auto tops = sm.getTopologies();
    const auto ps = sm.getParticles();
    CHECK(ps.size() == 12u && tops.size() == 3u, "%zu particles, %zu topologies", ps.size(), tops.size());
    for (std::size_t t = 0; t < tops.size(); ++t) {
        const auto &pi = tops[t]->particleIndices();
        CHECK(pi.size() == 4u && pi[3] >= 9u && pi[3] < 12u && tops[t]->hasEdge(2, 3) && tops[t]->isConnected(),
              "chain %zu: %zu particles, new handle %zu", t, pi.size(), pi.size() > 3 ? pi[3] : 0);
        for (const auto &p : ps) {
            if (p.id() == pi[0])
                CHECK(std::fabs(p.pos().x + 0.3) < 1e-12 && std::fabs(p.pos().y - ys[t]) < 1e-12, "bead 0 moved to %.3f",
                      p.pos().x);
            if (pi.size() > 3 && p.id() == pi[3])
                CHECK(std::fabs(p.pos().x - 3.2) < 1e-12 && p.type() == bead, "appended bead at %.3f", p.pos().x);
        }
    }
    k->actions().calculateForces()->perform();
    const double e = sm.energy(), want = 3 * (10. * 0.09 + 10. * 0.04);
    CHECK(std::fabs(e - want) < 1e-6 * want, "bond energy %.9g, expected %.9g", e, want);
    k->actions().evaluateTopologyReactions(0.01)->perform();
    CHECK(sm.getParticles().size() == 12u, "rate 0 at four particles: %zu particles", sm.getParticles().size());
    api::SimulationLoop loop(k.get(), 0.001);
    loop.run(5);
    CHECK(sm.getParticles().size() == 12u && sm.getTopologies().size() == 3u, "after 5 steps");
    std::printf("structural move and append: three chains grew to four beads\n");
}

// Kernel::observe(): observables connected to the kernel and evaluated by the SimulationLoop on the
// device-fused path in chunks that end where one is due (strides 0, 2, 5 and 10 over 10 steps), checked
// against the oracle on the state they saw: energy, virial, particles per type, the radial distribution
// and the histogram along an axis; conservation of A + C and B + C at every step.
static void test_observables() {
    const double L = 12.0;
    auto k = plugin::KernelProvider::getInstance().create("MI355X");
    c0_context(k->context(), L);
    const System s = uniform(400, 400, L, 9);
    std::vector<model::Particle> ps;
    for (size_t i = 0; i < s.types.size(); ++i)
        ps.emplace_back(s.pos[3 * i], s.pos[3 * i + 1], s.pos[3 * i + 2], (ParticleTypeId)s.types[i]);
    k->actions().addParticles(ps)->perform();
    auto &f = k->observe();
    std::vector<std::size_t> e_steps, n_steps;
    std::vector<double> energies;
    std::vector<std::vector<unsigned long>> counts;
    model::observables::Matrix33 last_virial{};
    std::vector<double> last_rdf;
    std::vector<double> last_hist;
    auto en = f.energy(2, [&](const double &e) { energies.push_back(e); });
    auto np = f.nParticles(0, {"A", "B", "C"}, [&](const std::vector<unsigned long> &c) { counts.push_back(c); });
    auto vir = f.virial(5, [&](const model::observables::Matrix33 &v) { last_virial = v; });
    std::vector<double> borders;
    for (int b = 0; b <= 20; ++b) borders.push_back(0.1 * b);
    auto rdf = f.radialDistribution(10, borders, {"A"}, {"B", "C"}, 800.0 / (L * L * L),
                                    [&](const std::pair<std::vector<double>, std::vector<double>> &r) { last_rdf = r.second; });
    std::vector<double> hb;
    for (int b = 0; b <= 24; ++b) hb.push_back(-6.0 + 0.5 * b);
    auto hist = f.histogramAlongAxis(5, hb, {"A", "C"}, 1, [&](const std::vector<double> &h) { last_hist = h; });
    CHECK(k->context().recordVirial(), "the virial observable switches recordVirial on");
    for (model::observables::ObservableBase *o :
         std::initializer_list<model::observables::ObservableBase *>{en.get(), np.get(), vir.get(), rdf.get(), hist.get()})
        k->connectObservable(o);
    api::SimulationLoop loop(k.get(), 0.01);
    loop.useReactionScheduler("UncontrolledApproximation");
    loop.run(10);
    CHECK(loop.currentStep() == 10, "loop step %zu", loop.currentStep());
    CHECK(energies.size() == 6, "energy evaluated %zu times (t = 0, 2, ..., 10)", energies.size());
    CHECK(counts.size() == 11, "particle counts evaluated %zu times (every step)", counts.size());
    for (const auto &c : counts) {
        CHECK(c.size() == 3 && c[0] + c[2] == 400 && c[1] + c[2] == 400, "A + C and B + C conserved (%lu %lu %lu)",
              c[0], c[1], c[2]);
    }
    CHECK(en->currentTimeStep() == 10 && rdf->currentTimeStep() == 10, "last evaluations at t = 10");
    // the oracle on the state the t = 10 evaluations saw
    const auto parts = k->stateModel().getParticles();
    std::vector<double> pos;
    std::vector<int> types;
    std::vector<uint64_t> ids;
    for (const auto &p : parts) {
        pos.insert(pos.end(), {p.pos().x, p.pos().y, p.pos().z});
        types.push_back(p.type());
        ids.push_back(p.id());
    }
    const double box[3] = {L, L, L};
    const int pbc[3] = {1, 1, 1};
    const double D[3] = {1.0, 1.0, 0.5};
    or_system *o = or_create(3, box, pbc, 1.0, D);
    const double hr[2] = {10.0, 1.0};
    for (int a = 0; a < 3; ++a)
        for (int b = a; b < 3; ++b) or_add_pair_potential(o, OR_POT_HARMONIC_REPULSION, a, b, hr);
    or_add_particles(o, (int)types.size(), pos.data(), types.data(), ids.data());
    or_nl_setup(o, 1.0, 1);
    double e_ref = 0, w_ref[9];
    or_calculate_forces(o, &e_ref, w_ref);
    CHECK(std::fabs(energies.back() - e_ref) <= 1e-5 * std::fabs(e_ref), "energy %.10g vs %.10g", energies.back(), e_ref);
    double dn = 0, dd = 0;
    for (int c = 0; c < 9; ++c) {
        dn += (last_virial[c] - w_ref[c]) * (last_virial[c] - w_ref[c]);
        dd += w_ref[c] * w_ref[c];
    }
    CHECK(dd > 0 && std::sqrt(dn / dd) < 1e-5, "virial relative error %.3g", std::sqrt(dn / std::max(dd, 1e-300)));
    const int from[1] = {0}, to[2] = {1, 2};
    std::vector<uint64_t> rc(borders.size() - 1);
    uint64_t nfrom = 0;
    or_rdf_counts(o, borders.data(), (int)borders.size(), from, 1, to, 2, rc.data(), &nfrom);
    CHECK(last_rdf.size() == rc.size(), "rdf bins");
    for (size_t b = 0; b < rc.size() && b < last_rdf.size(); ++b) {
        const double lo = borders[b], hi = borders[b + 1];
        const double want = (double)rc[b] / (4. / 3. * M_PI * (std::pow(hi, 3.) - std::pow(lo, 3.)) * (double)nfrom *
                                             (800.0 / (L * L * L)));
        CHECK(last_rdf[b] == want, "rdf bin %zu: %.10g vs %.10g", b, last_rdf[b], want);
    }
    const int htypes[2] = {0, 2};
    std::vector<uint64_t> hc(hb.size() - 1);
    or_histogram_along_axis(o, hb.data(), (int)hb.size(), htypes, 2, 1, hc.data());
    for (size_t b = 0; b < hc.size() && b < last_hist.size(); ++b)
        CHECK(last_hist[b] == (double)hc[b], "histogram bin %zu: %g vs %llu", b, last_hist[b], (unsigned long long)hc[b]);
    or_destroy(o);
    // the topology action factory: edits of a topology's graph and type, undone in reverse
    CHECK(k->getTopologyActionFactory() != nullptr, "the kernel supports topologies");
    model::GraphTopology g("T", {0, 1, 2});
    auto *taf = k->getTopologyActionFactory();
    auto a1 = taf->createAddEdge(&g, 0, 1);
    auto a2 = taf->createChangeTopologyType(&g, "U");
    a1->perform();
    a2->perform();
    CHECK(g.hasEdge(0, 1) && g.type() == "U", "edge added and type changed");
    a2->undo();
    a1->undo();
    CHECK(!g.hasEdge(0, 1) && g.type() == "T", "both undone");
    bool threw = false;
    try {
        taf->createRemoveEdge(&g, 0, 2)->perform();
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    CHECK(threw, "removing a missing edge throws");
}

int main(int argc, char **argv) {
    if (argc > 1 && std::string(argv[1]) == "--host-only") {
        test_graph_ops();
        test_descriptors();
        test_spatial_descriptors();
        std::printf(failures ? "FAILED (%d)\n" : "ALL PASSED\n", failures);
        return failures ? 1 : 0;
    }
    if (argc < 2) {
        std::printf("usage: %s <plugin directory>\n", argv[0]);
        return 2;
    }
    const bool provider_only = argc > 2 && std::string(argv[2]) == "--provider-only";
    try {
        test_provider_and_errors(argv[1]);
        test_detailed_balance_configs();
        if (!provider_only) {
            test_detailed_balance_scenes();
            test_structural_reactions();
            test_c0_forces_and_loop(argv[1]);
            test_observables();
            test_topology();
            test_break_bonds();
            test_topology_reactions();
            test_topology_fusion();
            test_structural_move_and_append();
            test_topology_topology_enzymatic();
            test_topology_topology_fusion();
            test_topology_self_fusion();
        }
    } catch (const std::exception &e) {
        std::printf("FAIL exception: %s\n", e.what());
        return 1;
    }
    std::printf(failures ? "FAILED (%d)\n" : "ALL PASSED\n", failures);
    return failures ? 1 : 0;
}
