// Reversible reaction pairs for the DetailedBalance handler (host code, FP64).
//
// DetailedBalance::searchReversibleReactions (readdy/main/model/Actions.cpp:432-519) pairs the registered
// reactions into FusionFission (A + B <-> C), ConversionConversion (A <-> B) and EnzymaticEnzymatic
// (A + C <-> B + C), and ReversibleReactionConfig (Actions.cpp:67-367) derives from each pair the interaction
// and reaction volumes, the acceptance prefactor, the inferred macroscopic rates and, for FusionFission, the
// cumulative distribution of the fission distance.  The integrals are the reference's: the radial Boltzmann
// factor 4 pi r^2 exp(-U(r)/kBT) of the lhs (rhs) pair potentials, integrated adaptively to a relative error
// of 1e-12 (the reference bisects the panel with the largest error estimate of a 201-point Gauss-Kronrod
// rule, common/integration.h:388-520; here the same panel strategy runs the 15-point Gauss-Kronrod rule of
// QUADPACK, which reaches the same integrals to the requested accuracy).
//
// Used by the engine: its device tables for the DetailedBalance handler and rdy_reversible_reactions.
// pair_energy_fp64 also runs on the device (the Metropolis energy changes).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace rdy_rev {

enum Kind { FUSION_FISSION = 0, CONVERSION_CONVERSION = 1, ENZYMATIC_ENZYMATIC = 2 };

// reaction kinds as include/readdy_amd.h numbers them
enum RxKind { CONVERSION = 0, FUSION = 1, FISSION = 2, ENZYMATIC = 3, DECAY = 4 };

struct Reaction {  // one registered reaction; entries an order-1 reaction or a single product does not use are 0
    int kind;
    uint32_t educts[2], products[2];
    double rate, educt_distance, product_distance, weight1, weight2;
    int n_educts() const { return (kind == FUSION || kind == ENZYMATIC) ? 2 : 1; }
    int n_products() const { return kind == DECAY ? 0 : ((kind == FISSION || kind == ENZYMATIC) ? 2 : 1); }
};

struct PairPotential {  // include/readdy_amd.h rdy_pair_potential
    int kind;  // 0 harmonic repulsion, 1 weak interaction, 2 Lennard-Jones, 3 screened electrostatics
    uint32_t type1, type2;
    double params[8];
};

// U(r) of one pair potential at distance r (PotentialsOrder2.h calculateEnergy, FP64); q = its
// rdy_pair_potential parameters
__host__ __device__ inline double pair_energy_fp64(int kind, const double *q, double r) {
    switch (kind) {
    case 0: {  // k/2 (r - d)^2 below d (PotentialsOrder2.h:72-79)
        const double k = q[0], d = q[1];
        if (r * r < d * d) return 0.5 * k * (r - d) * (r - d);
        return 0.0;
    }
    case 1: {  // three harmonic pieces (PotentialsOrder2.h:131-152)
        const double k = q[0], d0 = q[1], h = q[2], rc = q[3];
        const double len2 = rc - d0;
        if (r < d0) return 0.5 * k * (r - d0) * (r - d0) - h;
        if (r < d0 + 0.5 * len2) return 0.5 * h * (1.0 / (0.5 * len2)) * (1.0 / (0.5 * len2)) * (r - d0) * (r - d0) - h;
        if (r < rc) return -0.5 * h * (1.0 / (0.5 * len2)) * (1.0 / (0.5 * len2)) * (r - rc) * (r - rc);
        return 0.0;
    }
    case 2: {  // k [(s/r)^m - (s/r)^n], shifted by its value at the cutoff (PotentialsOrder2.h:210-233)
        const double m = q[0], n = q[1], rc = q[2], eps = q[4], s = q[5];
        const bool shift = q[3] != 0.0;
        if (r > rc) return 0.0;
        // k = -eps / ((s / r_min)^m - (s / r_min)^n), r_min = s (n/m)^(1/(n-m))  (Potentials.cpp:157-171)
        const double rmin = s * pow(n / m, 1.0 / (n - m));
        const double k = -eps / (pow(s / rmin, m) - pow(s / rmin, n));
        const double er = k * (pow(s / r, m) - pow(s / r, n));
        return shift ? er - k * (pow(s / rc, m) - pow(s / rc, n)) : er;
    }
    case 3: {  // C e^(-kappa r)/r + D (s/r)^n, with no cutoff test (PotentialsOrder2.h:291-296)
        const double C = q[0], kappa = q[1], D = q[2], s = q[3], n = q[4];
        return C * exp(-kappa * r) / r + D * pow(s / r, n);
    }
    default:
        return 0.0;
    }
}

inline double pair_energy(const PairPotential &p, double r) { return pair_energy_fp64(p.kind, p.params, r); }

inline double cutoff_of(const PairPotential &p) {
    switch (p.kind) {
    case 0: return p.params[1];
    case 1: return p.params[3];
    case 2: return p.params[2];
    case 3: return p.params[5];
    default: return 0.0;
    }
}

// QUADPACK's 15-point Gauss-Kronrod rule on [a, b]: (Kronrod integral, |Kronrod - Gauss|)
template <class F>
std::pair<double, double> gk15(F f, double a, double b) {
    static const double xgk[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                                  0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                                  0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                                  0.207784955007898467600689403773245, 0.0};
    static const double wgk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                                  0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                                  0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                                  0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
    static const double wg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                                 0.381830050505118944950369775488975, 0.417959183673469387755102040816327};
    if (!(a < b)) return {0.0, 0.0};
    const double c = 0.5 * (a + b), h = 0.5 * (b - a);
    const double fc = f(c);
    double k = fc * wgk[7], g = fc * wg[3];
    for (int j = 0; j < 7; ++j) {
        const double s = f(c - h * xgk[j]) + f(c + h * xgk[j]);
        k += wgk[j] * s;
        if (j % 2 == 1) g += wg[j / 2] * s;
    }
    return {k * h, std::fabs((k - g) * h)};
}

// adaptive integration: bisect the panel with the largest error estimate until the relative error of the
// sum is below rel (common/integration.h:466-520)
template <class F>
double integrate(F f, double a, double b, double rel = 1e-12, int maxiter = 1000) {
    struct Panel {
        double a, b, v, err;
    };
    if (!(a < b)) return 0.0;
    std::vector<Panel> panels;
    auto r0 = gk15(f, a, b);
    panels.push_back({a, b, r0.first, r0.second});
    double total = r0.first, err = r0.second;
    if (total == 0.0 && err == 0.0) return 0.0;
    for (int it = 0; it < maxiter && err / std::fabs(total) > rel; ++it) {
        auto worst = std::max_element(panels.begin(), panels.end(),
                                      [](const Panel &x, const Panel &y) { return x.err < y.err; });
        const Panel p = *worst;
        panels.erase(worst);
        const double m = 0.5 * (p.a + p.b);
        const auto lo = gk15(f, p.a, m), hi = gk15(f, m, p.b);
        panels.push_back({p.a, m, lo.first, lo.second});
        panels.push_back({m, p.b, hi.first, hi.second});
        total = 0.0;
        err = 0.0;
        for (const auto &q : panels) {
            total += q.v;
            err += q.err;
        }
    }
    return total;
}

constexpr int FISSION_POINTS = 1000;  // the fission distance table has FISSION_POINTS + 1 radii (Actions.cpp:296)

struct Config {  // ReversibleReactionConfig (include/readdy/model/actions/DetailedBalance.h:64-139)
    Kind kind;
    uint32_t forward, backward;  // indices into the reaction list
    int n_lhs, n_rhs;
    uint32_t lhs[2], rhs[2];
    double micro_forward, micro_backward, reaction_radius = 0;
    double lhs_radius = 0, lhs_volume = 0, rhs_radius = 0, rhs_volume = 0;
    double eff_lhs_interaction = 0, eff_rhs_interaction = 0, eff_lhs_reaction = 0, eff_rhs_reaction = 0;
    double acceptance_prefactor = 1, total_volume = 0, kbt = 1;
    double equilibrium_constant = 0, macro_forward = 0, macro_backward = 0;
    std::vector<double> fission_radii, fission_cdf;  // FusionFission only
};

inline std::vector<const PairPotential *> potentials_of(const std::vector<PairPotential> &pots, uint32_t a, uint32_t b) {
    std::vector<const PairPotential *> v;
    for (const auto &p : pots)
        if ((p.type1 == a && p.type2 == b) || (p.type1 == b && p.type2 == a)) v.push_back(&p);
    return v;
}

inline bool same2(const uint32_t x[2], const uint32_t y[2]) { return x[0] == y[0] && x[1] == y[1]; }

// ReversibleReactionConfig::ReversibleReactionConfig (Actions.cpp:67-367); throws std::logic_error where it does
inline Config make_config(uint32_t fwd, uint32_t bwd, const std::vector<Reaction> &rx,
                          const std::vector<PairPotential> &pots, double kbt, double box_volume) {
    Config c{};
    c.forward = fwd;
    c.backward = bwd;
    const Reaction *F = &rx.at(fwd), *B = &rx.at(bwd);
    int nl = F->n_educts(), nr = B->n_educts();
    if (nl == 2 && nr == 1) {
        c.kind = FUSION_FISSION;
    } else if (nl == 1 && nr == 2) {
        c.kind = FUSION_FISSION;  // the lhs has the two educts: forward and backward swap
        std::swap(F, B);
        std::swap(c.forward, c.backward);
        nl = 2;
        nr = 1;
    } else if (nl == 1 && nr == 1) {
        c.kind = CONVERSION_CONVERSION;
    } else if (nl == 2 && nr == 2) {
        c.kind = ENZYMATIC_ENZYMATIC;
    } else {
        throw std::logic_error("Type of reversible reaction cannot be determined");
    }
    c.n_lhs = nl;
    c.n_rhs = nr;
    // the two-entry educt and product arrays of Reaction.h:176-177 (unused entries 0, as the caller
    // passes them)
    for (int i = 0; i < 2; ++i) {
        c.lhs[i] = F->educts[i];
        c.rhs[i] = B->educts[i];
    }
    // species: lhs == backward products, rhs == forward products (fusion-fission may swap the lhs order)
    if (!same2(c.lhs, B->products) || !same2(c.rhs, F->products)) {
        const bool swapped_ok = c.kind == FUSION_FISSION && c.lhs[0] == B->products[1] && c.lhs[1] == B->products[0] &&
                                F->products[0] == B->educts[0];
        if (!swapped_ok) throw std::logic_error("The given reactions are not reversible with respect to particle species");
    }
    if (c.kind == FUSION_FISSION) {
        if (F->educts[0] == B->products[0] && F->educts[1] == B->products[1]) {
            if (!(F->weight1 == B->weight1 && F->weight2 == B->weight2))
                throw std::logic_error("The given reactions are not reversible with respect to weights");
        } else if (F->educts[0] == B->products[1] && F->educts[1] == B->products[0]) {
            if (!(F->weight1 == B->weight2 && F->weight2 == B->weight1))
                throw std::logic_error("The given reactions are not reversible with respect to weights");
        }
    }
    c.micro_forward = F->rate;
    c.micro_backward = B->rate;
    if (c.kind == FUSION_FISSION) {
        if (F->educt_distance != B->product_distance)
            throw std::logic_error("Reaction radii forward and backward must be equal");
        c.reaction_radius = F->educt_distance;
    } else if (c.kind == ENZYMATIC_ENZYMATIC) {
        if (F->educt_distance != B->educt_distance)
            throw std::logic_error("Reaction radii forward and backward must be equal");
        c.reaction_radius = F->educt_distance;
    }
    std::vector<const PairPotential *> lp, rp;
    if (c.kind == FUSION_FISSION || c.kind == ENZYMATIC_ENZYMATIC) lp = potentials_of(pots, c.lhs[0], c.lhs[1]);
    if (c.kind == ENZYMATIC_ENZYMATIC) rp = potentials_of(pots, c.rhs[0], c.rhs[1]);
    const double pi = 3.14159265358979323846;
    for (const auto *p : lp) c.lhs_radius = std::max(c.lhs_radius, cutoff_of(*p));
    for (const auto *p : rp) c.rhs_radius = std::max(c.rhs_radius, cutoff_of(*p));
    if (!lp.empty()) c.lhs_volume = 4. / 3. * pi * std::pow(c.lhs_radius, 3);
    if (!rp.empty()) c.rhs_volume = 4. / 3. * pi * std::pow(c.rhs_radius, 3);
    auto integrand = [&](const std::vector<const PairPotential *> &ps) {
        return [ps, kbt, pi](double r) {
            double u = 0;
            for (const auto *p : ps) u += pair_energy(*p, r);
            return 4. * pi * r * r * std::exp(-u / kbt);
        };
    };
    const auto fl = integrand(lp), fr = integrand(rp);
    c.kbt = kbt;
    c.eff_lhs_interaction = integrate(fl, 0., c.lhs_radius);
    c.eff_rhs_interaction = integrate(fr, 0., c.rhs_radius);
    c.eff_lhs_reaction = integrate(fl, 0., c.reaction_radius);
    c.eff_rhs_reaction = integrate(fr, 0., c.reaction_radius);
    if (c.kind == FUSION_FISSION) {  // Actions.cpp:292-321
        const double dr = c.reaction_radius / (double)FISSION_POINTS;
        double r = 0.;
        for (int i = 0; i < FISSION_POINTS + 1; ++i) {
            c.fission_radii.push_back(r);
            r += dr;
        }
        for (double ri : c.fission_radii) c.fission_cdf.push_back(integrate(fl, c.fission_radii.front(), ri));
        if (c.fission_cdf.back() >= 1e-16) {
            const double last = c.fission_cdf.back();
            for (auto &p : c.fission_cdf) p /= last;
        } else {
            std::fill(c.fission_cdf.begin(), c.fission_cdf.end() - 1, 0.);
            c.fission_cdf.back() = 1.;
        }
    }
    c.acceptance_prefactor = c.kind == ENZYMATIC_ENZYMATIC ? c.eff_lhs_reaction / c.eff_rhs_reaction : 1.;
    c.total_volume = box_volume;
    switch (c.kind) {  // Actions.cpp:340-366
    case FUSION_FISSION:
        c.equilibrium_constant = (c.micro_backward / c.micro_forward) *
                                 (c.total_volume - c.lhs_volume + c.eff_lhs_interaction) / c.eff_lhs_reaction;
        c.macro_backward = c.micro_backward;
        c.macro_forward = c.macro_backward * c.total_volume / c.equilibrium_constant;
        break;
    case CONVERSION_CONVERSION:
        c.macro_forward = c.micro_forward;
        c.macro_backward = c.micro_backward;
        c.equilibrium_constant = c.macro_backward / c.macro_forward;
        break;
    case ENZYMATIC_ENZYMATIC:
        c.macro_forward = c.micro_forward * c.total_volume * c.eff_lhs_reaction /
                          (c.total_volume - c.lhs_volume + c.eff_lhs_interaction);
        c.macro_backward = c.micro_backward * c.total_volume * c.eff_rhs_reaction /
                           (c.total_volume - c.rhs_volume + c.eff_rhs_interaction);
        c.equilibrium_constant = c.macro_backward / c.macro_forward;
        break;
    }
    return c;
}

// DetailedBalance::searchReversibleReactions (Actions.cpp:432-519): the pairs in the reference's order
// (fusion-fission, conversion-conversion, enzymatic-enzymatic); equivalent pairs throw std::logic_error
inline std::vector<Config> search(const std::vector<Reaction> &rx, const std::vector<PairPotential> &pots, double kbt,
                                  double box_volume) {
    std::vector<uint32_t> o1, o2;
    for (uint32_t i = 0; i < rx.size(); ++i) (rx[i].n_educts() == 1 ? o1 : o2).push_back(i);
    std::vector<Config> out;
    for (uint32_t i2 : o2) {
        const Reaction &r2 = rx[i2];
        for (uint32_t i1 : o1) {
            const Reaction &r1 = rx[i1];
            if (r2.n_products() == 1 && r1.n_products() == 2 && r1.product_distance == r2.educt_distance) {
                // the second alternative repeats its first comparison, as the reference's does (Actions.cpp:453-455)
                const bool plain = r2.educts[0] == r1.products[0] && r2.educts[1] == r1.products[1] &&
                                   r1.weight1 == r2.weight1 && r1.weight2 == r2.weight2;
                const bool swapped = r2.educts[0] == r1.products[1] && r2.educts[0] == r1.products[1] &&
                                     r1.weight1 == r2.weight2 && r1.weight2 == r2.weight1;
                if (r1.educts[0] == r2.products[0] && (plain || swapped))
                    out.push_back(make_config(i2, i1, rx, pots, kbt, box_volume));
            }
        }
    }
    for (std::size_t a = 0; a < o1.size(); ++a)
        for (std::size_t b = a + 1; b < o1.size(); ++b) {
            const Reaction &ri = rx[o1[a]], &rj = rx[o1[b]];
            if (ri.n_products() == 1 && rj.n_products() == 1 && ri.educts[0] == rj.products[0] &&
                ri.products[0] == rj.educts[0])
                out.push_back(make_config(o1[a], o1[b], rx, pots, kbt, box_volume));
        }
    for (std::size_t a = 0; a < o2.size(); ++a)
        for (std::size_t b = a + 1; b < o2.size(); ++b) {
            const Reaction &ri = rx[o2[a]], &rj = rx[o2[b]];
            if (ri.n_products() == 2 && rj.n_products() == 2 && ri.educt_distance == rj.educt_distance &&
                same2(ri.educts, rj.products) && same2(ri.products, rj.educts))
                out.push_back(make_config(o2[a], o2[b], rx, pots, kbt, box_volume));
        }
    for (std::size_t i = 0; i < out.size(); ++i)  // equivalentReversibleReactions (DetailedBalance.h:144-148)
        for (std::size_t j = i + 1; j < out.size(); ++j) {
            const Config &x = out[i], &y = out[j];
            const bool eq = (same2(x.lhs, y.lhs) && same2(x.rhs, y.rhs)) || (same2(x.lhs, y.rhs) && same2(x.rhs, y.lhs));
            if (eq)
                throw std::logic_error("Duplicate found in reversible reactions. Registered reactions might be ambiguous "
                                       "with respect to reversibility.");
        }
    return out;
}

// the fission distance of a uniform u (ReversibleReactionConfig::drawFissionDistance: the first table radius
// whose cumulative probability is not below u)
inline double fission_distance(const Config &c, double u) {
    const auto it = std::lower_bound(c.fission_cdf.begin(), c.fission_cdf.end(), u);
    return c.fission_radii[(std::size_t)std::min<std::ptrdiff_t>(it - c.fission_cdf.begin(),
                                                                 (std::ptrdiff_t)c.fission_radii.size() - 1)];
}

}  // namespace rdy_rev
