/*
 * readdy_oracle.c -- CPU FP64 restatement of ReaDDy's "CPU" kernel per-step path.
 *
 * TEST INFRASTRUCTURE ONLY (see readdy_oracle.h).  Compiled with -ffp-contract=off so
 * every floating-point expression is evaluated in the reference's operation order
 * without fused multiply-add (the reference's x86-64 baseline build has no FMA).
 *
 * Random numbers.  The reference draws from thread-local std::mt19937 seeded from
 * std::random_device (include/readdy/model/RandomProvider.h:55-104), which no two runs
 * reproduce.  The oracle therefore uses a counter-based Philox4x32-10 stream with the
 * counter layout below; the HIP engine implements the same layout independently, so
 * both sides draw identical integers:
 *   key  = (seed & 0xffffffff, seed >> 32)
 *   BD noise of particle h in step t        ctr = (h, t, 0, 0)
 *   order-1 reaction r of particle h        ctr = (h, t, 1 | r << 8, 0)
 *   order-2 reaction r of pair (lo, hi)     ctr = (lo, t, 2 | r << 8, hi)   (lo < hi ids)
 *   fission direction of particle h         ctr = (h, t, 3, 0)
 *   DetailedBalance fission distance of h   ctr = (h, t, 10, 0)        u53(x0, x1)
 *   DetailedBalance Metropolis, reaction r  ctr = (lo, t, 11 | r << 8, hi)   u53(x0, x1), lo <= hi ids
 *   uniform from words (a, b)               u53(a, b) = ((a >> 5) * 2^26 + (b >> 6)) / 2^53
 *   normals (Box-Muller)                    v_k = (x_k + 0.5) / 2^32;
 *        n0 = sqrt(-2 ln v0) cos(2 pi v1), n1 = sqrt(-2 ln v0) sin(2 pi v1), n2 = sqrt(-2 ln v2) cos(2 pi v3)
 *   reaction acceptance                     u53(x0, x1) < 1 - exp(-rate dt)
 *   UA priority (shuffle order)             (x2 << 32) | x3, ascending
 *   Gillespie key (Efraimidis-Spirakis)     -ln(u53(x2, x3) + 2^-54) / rate, ascending
 *
 * Parity status (what pins each function; "parity unpinned" where nothing does):
 *   or_philox4x32_10     Random123 known-answer vectors (tests/golden/philox_kat.json)
 *   bcs helpers          boundary cases + brute force
 *   pair potentials      reference known answers (TestStateModel.cpp, TestPotentials.cpp); WIPH by
 *                        closed-form continuity values and finite differences
 *   bonded terms         reference known answers (TestTopologies.cpp) + finite differences
 *   cell list            brute-force O(N^2) pair sets + reference fixed scenes (TestNeighborList.cpp)
 *   forces/energy        second traversal: brute-force half-shell sum (independent of the cell list)
 *   EulerBD              closed form with injected noise
 *   performReaction      reference known answers (kernels/cpu/test/TestReactions.cpp, readdy/test/TestReactions.cpp)
 *   UA / Gillespie       deterministic reference scenes (kernels/cpu/test/TestReactions.cpp:250-354) and
 *                        the statistical equivalence of literal and keyed Gillespie selection
 *   reversible configs   reference known answers (readdy/test/TestDetailedBalance.cpp: volumes, prefactor,
 *                        rates, pairs found and rejected) and closed forms (free-particle volumes, the
 *                        harmonic-repulsion CDF against direct quadrature)
 *   DetailedBalance      closed forms: acceptance 1 without potentials, the Metropolis ratio of single
 *                        conversions against exp(-dE/kT) of a two-particle scene, the rejected move's id
 *                        bookkeeping; the equilibrium statistics of a conversion pair
 *   virial               two-particle known answer (sign and the once-per-pair count) + O(N^2) brute force
 *   topology reactions   enzymatic kind: hand-built scenes (one event per topology and per free particle,
 *                        the strict radius gate) and a chi-square test of the keyed selection against the
 *                        literal performEvents; fusion kind: polymer-growth scenes whose outcome follows
 *                        from the rules (the joined particle is the next end), the reserve limit;
 *                        topology-topology kinds: scenes with pairs in either order, pairs inside one
 *                        topology, one reaction per topology, and for merges a check reaction that fires
 *                        only on a particle left in its absorbed topology; self-fusion: a ring scene whose
 *                        unbonded chain ends join while a bonded pair does not, and the same scene without
 *                        the flag, where only the two separate topologies merge
 *   break bonds          reference scenes (TestBreakingBonds.cpp) through the energy threshold
 *   RDF / histograms     two-particle known answer, numpy brute-force pair counts and direct binning
 */
#include "readdy_oracle.h"

#include <math.h>
#include <stdio.h>
#include <time.h>
#include <stdlib.h>
#include <string.h>

#define OR_MAX_TYPES 64
#define OR_MAX_POTS 256
#define OR_MAX_RX 256

static __thread char g_err[512];

static int fail(const char *msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return -1;
}

const char *or_last_error(void) { return g_err; }

/* ------------------------------------------------------------------------------------------------
 * Philox4x32-10, written from Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
 * 1, 2, 3", SC'11 (round function and constants as published there).
 * --------------------------------------------------------------------------------------------- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

double or_u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

static void philox_draw(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t out[4]) {
    uint32_t ctr[4] = {c0, c1, c2, c3};
    uint32_t key[2] = {(uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32)};
    or_philox4x32_10(ctr, key, out);
}

static void normals3(const uint32_t x[4], double n[3]) {
    const double two_pi = 6.283185307179586;
    double v0 = ((double)x[0] + 0.5) * (1.0 / 4294967296.0);
    double v1 = ((double)x[1] + 0.5) * (1.0 / 4294967296.0);
    double v2 = ((double)x[2] + 0.5) * (1.0 / 4294967296.0);
    double v3 = ((double)x[3] + 0.5) * (1.0 / 4294967296.0);
    double r0 = sqrt(-2.0 * log(v0));
    double r1 = sqrt(-2.0 * log(v2));
    n[0] = r0 * cos(two_pi * v1);
    n[1] = r0 * sin(two_pi * v1);
    n[2] = r1 * cos(two_pi * v3);
}

/* the Brownian-dynamics noise: three N(0,1) draws by Box-Muller in FP32 from one Philox block, the
 * transform the HIP engine uses: 24-bit uniforms v_k = ((x_k >> 8) + 1/2) 2^-24, r = sqrt(-2 ln v),
 * angle 2 pi v.  The logarithm, sine and cosine are formed in FP64 and rounded to FP32, so each result
 * lies within an FP32 unit of the exact value, as the device's FP32 functions do; the two sides agree to
 * a few FP32 units, not bit for bit. */
static void bd_normals3(const uint32_t x[4], double n[3]) {
    const float s24 = 5.9604644775390625e-08f; /* 2^-24 */
    const float v0 = ((float)(x[0] >> 8) + 0.5f) * s24;
    const float v1 = ((float)(x[1] >> 8) + 0.5f) * s24;
    const float v2 = ((float)(x[2] >> 8) + 0.5f) * s24;
    const float v3 = ((float)(x[3] >> 8) + 0.5f) * s24;
    const float r0 = sqrtf(-2.0f * (float)log((double)v0));
    const float r1 = sqrtf(-2.0f * (float)log((double)v2));
    const double pi = 3.14159265358979323846;
    const float a1 = 2.0f * v1, a3 = 2.0f * v3;
    const float s1 = (float)sin(pi * (double)a1), c1 = (float)cos(pi * (double)a1);
    const float c3 = (float)cos(pi * (double)a3);
    n[0] = (double)(r0 * c1);
    n[1] = (double)(r0 * s1);
    n[2] = (double)(r1 * c3);
}

/* ------------------------------------------------------------------------------------------------
 * Periodic boundary helpers: include/readdy/common/boundary_condition_operations.h:53-91.
 * --------------------------------------------------------------------------------------------- */
void or_fix_position(double *v, const double *box, const int *pbc) {
    for (int d = 0; d < 3; ++d) {
        if (pbc[d]) {
            if (!isfinite(v[d])) continue; /* the reference would spin forever on inf */
            if (!(fabs(v[d]) < 1024.0 * box[d])) v[d] = fmod(v[d], box[d]);
            while (v[d] >= .5 * box[d]) v[d] -= box[d];
            while (v[d] < -.5 * box[d]) v[d] += box[d];
        }
    }
}

void or_shortest_difference(const double *lhs, const double *rhs, const double *box, const int *pbc, double *out) {
    for (int i = 0; i < 3; ++i) {
        double dv = rhs[i] - lhs[i];
        if (pbc[i]) {
            if (dv > box[i] * .5) dv -= box[i];
            else if (dv <= -box[i] * .5) dv += box[i];
        }
        out[i] = dv;
    }
}

/* inner product as std::inner_product from 0: ((x*x + y*y) + z*z) -- ReaDDyVec3.h:332-334 */
static double dot3(const double *a, const double *b) {
    double s = 0.0;
    s = s + a[0] * b[0];
    s = s + a[1] * b[1];
    s = s + a[2] * b[2];
    return s;
}

double or_dist2(const double *lhs, const double *rhs, const double *box, const int *pbc) {
    double dv[3];
    or_shortest_difference(lhs, rhs, box, pbc, dv);
    return dot3(dv, dv);
}

/* ------------------------------------------------------------------------------------------------
 * Pair potentials: include/readdy/model/potentials/PotentialsOrder2.h:54-323,
 * readdy/main/model/potentials/Potentials.cpp:150-202.  x_ij = neighbour - self.
 * Returns the energy; adds the force on "self" into f.
 * --------------------------------------------------------------------------------------------- */
static double lj_k(double m, double n, double eps, double sigma) {
    /* Potentials.cpp:157-171 */
    double r_min = sigma * pow(n / m, 1. / (n - m));
    return -eps / (pow(sigma / r_min, m) - pow(sigma / r_min, n));
}

double or_pair_force_energy(int kind, const double *p, const double *x, double *f) {
    double d2 = dot3(x, x);
    switch (kind) {
    case OR_POT_HARMONIC_REPULSION: { /* PotentialsOrder2.h:72-91 */
        double k = p[0], d = p[1], dd = d * d, e = 0.0;
        if (d2 < dd) {
            double r = sqrt(d2);
            r -= d;
            r *= r;
            e = 0.5 * r * k;
        }
        if (d2 < dd && d2 > 0) {
            double r = sqrt(d2);
            double fac = (k * (r - d)) / r;
            f[0] += fac * x[0];
            f[1] += fac * x[1];
            f[2] += fac * x[2];
        }
        return e;
    }
    case OR_POT_WEAK_INTERACTION: { /* PotentialsOrder2.h:131-183 */
        double k = p[0], d0 = p[1], h = p[2], rc = p[3];
        double dist = sqrt(d2);
        double len2 = rc - d0;
        double e = 0.0;
        if (dist < d0) {
            e = .5 * k * (dist - d0) * (dist - d0) - h;
        } else if (dist < d0 + .5 * len2) {
            e = .5 * h * (1. / (.5 * len2)) * (1. / (.5 * len2)) * (dist - d0) * (dist - d0) - h;
        } else if (dist < rc) {
            e = -.5 * h * (1. / (.5 * len2)) * (1. / (.5 * len2)) * (dist - rc) * (dist - rc);
        }
        double factor = 0;
        if (dist < d0) {
            factor = -1 * k * (d0 - dist);
        } else {
            if (dist < d0 + .5 * len2) {
                factor = -1. * h * (1. / (.5 * len2)) * (1. / (.5 * len2)) * (d0 - dist);
            } else if (dist < rc) {
                factor = h * (1. / (.5 * len2)) * (1. / (.5 * len2)) * (rc - dist);
            }
        }
        if (dist > 0 && factor != 0) {
            f[0] += factor * x[0] / dist;
            f[1] += factor * x[1] / dist;
            f[2] += factor * x[2] / dist;
        }
        return e;
    }
    case OR_POT_LENNARD_JONES: { /* PotentialsOrder2.h:232-262 */
        double m = p[0], n = p[1], rc = p[2], shift = p[3], eps = p[4], sigma = p[5];
        double k = lj_k(m, n, eps, sigma);
        double r = sqrt(d2);
        double e = 0.0;
        if (!(r > rc)) {
            e = k * (pow(sigma / r, m) - pow(sigma / r, n));
            if (shift != 0.0) e -= k * (pow(sigma / rc, m) - pow(sigma / rc, n));
        }
        if (r <= rc) {
            double fac = -1. * k * (1 / (sigma * sigma)) * (m * pow(sigma / r, m + 2) - n * pow(sigma / r, n + 2));
            f[0] += fac * x[0];
            f[1] += fac * x[1];
            f[2] += fac * x[2];
        }
        return e;
    }
    case OR_POT_SCREENED_ELECTROSTATICS: { /* PotentialsOrder2.h:299-315 */
        double C = p[0], kappa = p[1], D = p[2], sigma = p[3], nexp = p[4];
        double r = sqrt(d2);
        double e = C * exp(-kappa * r) / r;
        e += D * pow(sigma / r, nexp);
        double ff = C * exp(-kappa * r);
        ff *= (kappa / r + 1. / pow(r, 2.));
        ff += D * nexp / sigma * pow(sigma / r, nexp + 1.);
        f[0] += ff * (-1. * x[0] / r);
        f[1] += ff * (-1. * x[1] / r);
        f[2] += ff * (-1. * x[2] / r);
        return e;
    }
    default:
        return 0.0;
    }
}

static double pot_cutoff(int kind, const double *p) {
    switch (kind) {
    case OR_POT_HARMONIC_REPULSION: return p[1];
    case OR_POT_WEAK_INTERACTION: return p[3];
    case OR_POT_LENNARD_JONES: return p[2];
    case OR_POT_SCREENED_ELECTROSTATICS: return p[5];
    default: return 0.0;
    }
}

/* ------------------------------------------------------------------------------------------------
 * Bonded terms.
 * Bond:  include/readdy/model/topologies/potentials/BondedPotential.h:105-113 and
 *        kernels/cpu/src/actions/topologies/CPUTopologyActions.cpp:55-74.
 * Angle: readdy/main/model/topologies/potentials/AnglePotential.cpp:58-104, CPUTopologyActions.cpp:82-97.
 * Dihedral energy: TorsionPotential.cpp:53-70.  The dihedral force is the analytic gradient of
 * that energy written from the standard decomposition (Blondel & Karplus, J. Comput. Chem. 17,
 * 1132 (1996)); it keeps the reference's zero-force branch when |cos phi| < 1e-4
 * (TorsionPotential.cpp:91-93) and its 1e-4 floors on squared norms.
 * --------------------------------------------------------------------------------------------- */
double or_bond_force_energy(double k, double l0, const double *xi, const double *xj, const double *box,
                            const int *pbc, double *fi, double *fj) {
    if (k == 0) return 0.0;
    double x[3];
    or_shortest_difference(xi, xj, box, pbc, x);
    double norm = sqrt(dot3(x, x));
    double c = (2. * k * (norm - l0) / norm);
    for (int d = 0; d < 3; ++d) {
        double fu = c * x[d];
        fi[d] += fu;
        fj[d] -= fu;
    }
    return k * (norm - l0) * (norm - l0);
}

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

double or_angle_force_energy(double k, double theta0, const double *xi, const double *xj, const double *xk,
                             const double *box, const int *pbc, double *fi, double *fj, double *fk) {
    const double SMALL = 1e-10;
    double x_ji[3], x_jk[3];
    or_shortest_difference(xj, xi, box, pbc, x_ji);
    or_shortest_difference(xj, xk, box, pbc, x_jk);
    /* energy, AnglePotential.cpp:58-65 */
    double sp = dot3(x_ji, x_jk);
    double nji = sqrt(dot3(x_ji, x_ji)), njk = sqrt(dot3(x_jk, x_jk));
    double theta = acos(sp / (nji * njk));
    double energy = k * (theta - theta0) * (theta - theta0);
    /* force, AnglePotential.cpp:67-104 */
    double nji2 = dot3(x_ji, x_ji);
    if (nji2 < SMALL) nji2 = SMALL;
    double n_ji = sqrt(nji2);
    double njk2 = dot3(x_jk, x_jk);
    if (njk2 < SMALL) njk2 = SMALL;
    double n_jk = sqrt(njk2);
    double nprod = n_ji * n_jk;
    if (nprod < SMALL) nprod = SMALL;
    double inv = 1. / nprod;
    double cos_t = clampd(inv * sp, -1., 1.);
    double sin_inv = sqrt(1. - cos_t * cos_t);
    if (sin_inv < SMALL) sin_inv = SMALL;
    sin_inv = 1. / sin_inv;
    double c = 2. * k * (acos(cos_t) - theta0) * sin_inv;
    for (int d = 0; d < 3; ++d) {
        double force_i = c * cos_t * (1. / nji2) * x_ji[d] - c * inv * x_jk[d];
        double force_k = -c * inv * x_ji[d] + c * cos_t * (1. / njk2) * x_jk[d];
        fi[d] -= force_i;
        fj[d] += force_i + force_k;
        fk[d] -= force_k;
    }
    return energy;
}

static void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

double or_dihedral_force_energy(double k, int mult, double phi0, const double *xi, const double *xj,
                                const double *xk, const double *xl, const double *box, const int *pbc,
                                double *fi, double *fj, double *fk, double *fl) {
    const double SMALL = 1e-4;
    double x_ji[3], x_kj[3], x_kl[3], x_jk[3];
    or_shortest_difference(xj, xi, box, pbc, x_ji);
    or_shortest_difference(xk, xj, box, pbc, x_kj);
    or_shortest_difference(xk, xl, box, pbc, x_kl);
    for (int d = 0; d < 3; ++d) x_jk[d] = -1. * x_kj[d];
    /* energy: TorsionPotential.cpp:53-70 */
    double m[3], n[3], mxn[3];
    cross3(x_ji, x_kj, m);
    cross3(x_kl, x_jk, n);
    double jk_norm = sqrt(dot3(x_jk, x_jk));
    if (jk_norm < SMALL) jk_norm = SMALL;
    double mn_norm = sqrt(dot3(m, m)) * sqrt(dot3(n, n));
    if (mn_norm < SMALL) mn_norm = SMALL;
    cross3(m, n, mxn);
    double sin_e = dot3(mxn, x_jk) / (mn_norm * jk_norm);
    double cos_e = dot3(m, n) / mn_norm;
    double phi_e = -atan2(sin_e, cos_e);
    double energy = k * (1 + cos(mult * phi_e - phi0));

    /* force: gradient of phi.  With b1 = xj - xi, b2 = xk - xj, b3 = xl - xk the reference's
     * m = x_ji x x_kj = b1 x b2 =: A and n = x_kl x x_jk = b3 x b2 = -(b2 x b3) =: -B, so its
     * cos = -A.B/(|A||B|) and sin = -((A x B).b2)/(|A||B||b2|).  With the IUPAC angle
     * psi = atan2(|b2| b1.B, A.B) this gives phi = -atan2(-sin psi, -cos psi) = -psi - pi (mod 2 pi),
     * hence grad phi = -grad psi, and grad psi has the closed form
     *   dpsi/dxi = -|b2| A/|A|^2,  dpsi/dxl = |b2| B/|B|^2,
     *   dpsi/dxj = -(1 + b1.b2/|b2|^2) dpsi/dxi + (b3.b2/|b2|^2) dpsi/dxl,
     *   dpsi/dxk = (b1.b2/|b2|^2) dpsi/dxi - (1 + b3.b2/|b2|^2) dpsi/dxl. */
    double x2 = dot3(x_jk, x_jk);
    if (x2 < SMALL) x2 = SMALL;
    double m2 = dot3(m, m);
    if (m2 < SMALL) m2 = SMALL;
    double n2 = dot3(n, n);
    if (n2 < SMALL) n2 = SMALL;
    double mnn = sqrt(m2) * sqrt(n2);
    if (mnn < SMALL) mnn = SMALL;
    double cos_phi = dot3(n, m) / mnn;
    if (fabs(cos_phi) < SMALL) return energy;
    double sin_phi = dot3(mxn, x_jk) / (mnn * sqrt(x2));
    double phi = -atan2(sin_phi, cos_phi);
    double dVdphi = -k * mult * sin(mult * phi - phi0);

    double b1[3], b2[3], b3[3], A[3], B[3];
    for (int d = 0; d < 3; ++d) {
        b1[d] = -x_ji[d];
        b2[d] = x_jk[d];
        b3[d] = x_kl[d];
    }
    cross3(b1, b2, A);
    cross3(b2, b3, B);
    double A2 = dot3(A, A), B2 = dot3(B, B);
    if (A2 < SMALL) A2 = SMALL;
    if (B2 < SMALL) B2 = SMALL;
    double b2n = sqrt(x2);
    double p12 = dot3(b1, b2) / x2, p32 = dot3(b3, b2) / x2;
    double gi[3], gl[3];
    for (int d = 0; d < 3; ++d) {
        gi[d] = -b2n * A[d] / A2;
        gl[d] = b2n * B[d] / B2;
    }
    /* F = -dV/dphi grad phi = dV/dphi grad psi */
    for (int d = 0; d < 3; ++d) {
        double gj = -(1.0 + p12) * gi[d] + p32 * gl[d];
        double gk = p12 * gi[d] - (1.0 + p32) * gl[d];
        fi[d] += dVdphi * gi[d];
        fj[d] += dVdphi * gj;
        fk[d] += dVdphi * gk;
        fl[d] += dVdphi * gl[d];
    }
    return energy;
}

/* ------------------------------------------------------------------------------------------------
 * System state.  Particle container semantics follow
 * include/readdy/kernel/cpu/data/DefaultDataContainer.h:79-146 (blank reuse on add, update()
 * overwrites removed slots first, appends otherwise, tombstones leftover removals).
 * --------------------------------------------------------------------------------------------- */
typedef struct {
    double pos[3];
    double force[3];
    uint64_t id;
    int type;
    int deactivated;
} entry_t;

typedef struct {
    int kind, t1, t2;
    double p[8];
    double cutoff2;
} pot_t;

typedef struct {
    int kind, educts[2], products[2];
    double rate, educt_distance, educt_distance2, product_distance, w1, w2;
} rx_t;

typedef struct {
    size_t i[4];
    double k, l0;
    int mult;
} bonded_t;

#define OR_MAX_STR 64
/* a topology-particle spatial topology reaction, enzymatic "T1(p1) + (p2) -> T2(p3) + (p4)" or fusion
 * "T1(p1) + (p2) -> T2(p3--p4)" (SpatialTopologyReaction.h:64-236, STRMode::TP_ENZYMATIC / TP_FUSION) */
typedef struct {
    int type1, top_type1, type2, type_to1, top_type_to1, type_to2;
    double rate, radius, radius2;
    int fusion;
    int tt, top_type2, top_type_to2; /* topology-topology enzymatic: the second side in another topology */
    int self_fusion; /* topology-topology fusion that also joins two unbonded particles of one topology */
} str_t;

typedef struct {
    int n_educts;
    size_t idx1, idx2;
    int rx;          /* global reaction index */
    uint64_t key;    /* UA priority or Gillespie key bits */
    double gkey;     /* Gillespie E-S key */
    double rate, cumulative;
    int blocked;
} event_t;

struct or_system {
    int n_types;
    double box[3];
    int pbc[3];
    double kbt;
    double D[OR_MAX_TYPES];
    pot_t pots[OR_MAX_POTS];
    int n_pots;
    rx_t rx[OR_MAX_RX];
    int n_rx;
    entry_t *e;
    size_t n, cap;
    size_t *blanks;
    size_t n_blanks, cap_blanks;
    bonded_t *bonds, *angles, *dihedrals;
    size_t n_bonds, n_angles, n_dihedrals;
    uint64_t seed;
    uint64_t next_id;
    /* cell list */
    int nl_set;
    double nl_cutoff;
    int nl_radius;
    int dims[3];
    double cell_size[3];
    size_t n_cells;
    size_t *adj;       /* [n_cells][1 + max_adj] : count then indices */
    size_t max_adj;
    size_t *head;      /* [n_cells], 1-based particle index, 0 = end */
    size_t *list;      /* [n + 1] */
    size_t list_len;
    uint64_t counts[OR_MAX_RX];
    int record;                 /* keep reaction records (or_set_record_reactions) */
    or_reaction_record *recs;
    size_t n_recs, cap_recs;
    double phase_s[4]; /* or_loop_step wall time: integrator, neighborList (both fills), reactions, forces */
    int threads; /* > 1: forces, integration and the candidate search on that many OpenMP threads */
    int cas_fill; /* fill the cell list from all threads by atomic head exchange (timing baseline) */
    /* topologies for spatial topology reactions: the topology of each particle id (-1: none) and the
     * type of each topology */
    int *topology_of;
    size_t topology_of_len;
    int *top_types;
    size_t n_tops;
    uint64_t fuse_next, fuse_end; /* ids for particles joining topologies by fusion: fuse_next .. fuse_end - 1 */
    uint64_t *fusions;            /* the last call's fusions: (topology, topology particle id, old id, new id) */
    size_t n_fusions;
    uint64_t *merges;             /* the last call's topology merges: (surviving, absorbed, their particles' ids) */
    size_t n_merges;
    str_t str[OR_MAX_STR];
    int n_str;
    uint64_t str_counts[OR_MAX_STR];
    /* DetailedBalance: reversible pairs, built on the first call */
    struct rev_s *rev;
    int n_rev, rev_ready;
    int rev_of_rx[OR_MAX_RX];
};

or_system *or_create(int n_types, const double *box, const int *pbc, double kbt, const double *diffusion) {
    if (n_types <= 0 || n_types > OR_MAX_TYPES) {
        fail("n_types out of range");
        return NULL;
    }
    or_system *s = (or_system *)calloc(1, sizeof(or_system));
    s->n_types = n_types;
    for (int d = 0; d < 3; ++d) {
        s->box[d] = box[d];
        s->pbc[d] = pbc[d];
    }
    s->kbt = kbt;
    for (int t = 0; t < n_types; ++t) s->D[t] = diffusion[t];
    s->seed = 0;
    s->threads = 1;
    return s;
}

void or_destroy(or_system *s) {
    if (!s) return;
    free(s->topology_of);
    free(s->fusions);
    free(s->merges);
    free(s->top_types);
    free(s->e);
    free(s->blanks);
    free(s->bonds);
    free(s->angles);
    free(s->dihedrals);
    free(s->adj);
    free(s->head);
    free(s->list);
    free(s->recs);
    free(s->rev);
    free(s);
}

int or_add_pair_potential(or_system *s, int kind, int t1, int t2, const double *params) {
    if (s->n_pots >= OR_MAX_POTS) return fail("too many potentials");
    pot_t *p = &s->pots[s->n_pots++];
    p->kind = kind;
    p->t1 = t1;
    p->t2 = t2;
    memcpy(p->p, params, sizeof(double) * 8);
    double rc = pot_cutoff(kind, params);
    p->cutoff2 = rc * rc;
    return 0;
}

int or_add_reaction(or_system *s, int kind, int educt1, int educt2, int product1, int product2, double rate,
                    double educt_distance, double product_distance, double weight1, double weight2) {
    if (s->n_rx >= OR_MAX_RX) return fail("too many reactions");
    rx_t *r = &s->rx[s->n_rx++];
    r->kind = kind;
    r->educts[0] = educt1;
    r->educts[1] = educt2;
    r->products[0] = product1;
    r->products[1] = product2;
    r->rate = rate;
    r->educt_distance = educt_distance;
    r->educt_distance2 = educt_distance * educt_distance;
    r->product_distance = product_distance;
    /* Fusion.h:61-68 / Fission.h:63-70: weights normalised to sum 1 */
    double sum = weight1 + weight2;
    if (sum != 1) {
        weight1 /= sum;
        weight2 /= sum;
    }
    r->w1 = weight1;
    r->w2 = weight2;
    return 0;
}

int or_set_seed(or_system *s, uint64_t seed) {
    s->seed = seed;
    return 0;
}

/* Thread count of the timing baseline (the reference's CPU kernel runs a static-partition thread
 * pool, kernels/cpu/include/readdy/kernel/cpu/util/config.h).  Forces and positions are the same
 * for any count; the energy is summed in another order (FP64 rounding may differ in the last bits)
 * and the keyed handlers' candidates are listed in another order (both sort them by their keys
 * before use).  The literal Gillespie path depends on list order and gathers on one thread. */
int or_set_threads(or_system *s, int n) {
    s->threads = n < 1 ? 1 : n;
    return 0;
}

int or_set_cas_fill(or_system *s, int on) {
    s->cas_fill = on != 0;
    return 0;
}

int or_rdf_counts(const or_system *s, const double *borders, int n_borders, const int *from_types, int n_from,
                  const int *to_types, int n_to, uint64_t *counts, uint64_t *n_from_out) {
    if (n_borders < 2) return fail("bin borders need at least two entries");
    for (int b = 0; b + 1 < n_borders; ++b) counts[b] = 0;
    uint64_t nf = 0;
    for (size_t i = 0; i < s->n; ++i) {
        const entry_t *p = &s->e[i];
        if (p->deactivated) continue;
        int is_from = 0;
        for (int k = 0; k < n_from; ++k) is_from |= p->type == from_types[k];
        if (!is_from) continue;
        ++nf;
        for (size_t j = 0; j < s->n; ++j) {
            const entry_t *q = &s->e[j];
            if (q->deactivated || q->id == p->id) continue;
            int is_to = 0;
            for (int k = 0; k < n_to; ++k) is_to |= q->type == to_types[k];
            if (!is_to) continue;
            const double dist = sqrt(or_dist2(p->pos, q->pos, s->box, s->pbc));
            /* std::upper_bound: the first border greater than dist */
            int lo = 0, hi = n_borders;
            while (lo < hi) {
                const int mid = (lo + hi) / 2;
                if (dist < borders[mid]) hi = mid;
                else lo = mid + 1;
            }
            /* lo == 0 (below the first border) would index counts[-1] in the reference; not counted here */
            if (lo > 0 && lo < n_borders) counts[lo - 1]++;
        }
    }
    *n_from_out = nf;
    return 0;
}

int or_histogram_along_axis(const or_system *s, const double *borders, int n_borders, const int *types, int n_types,
                            int axis, uint64_t *counts) {
    if (n_borders < 2) return fail("bin borders need at least two entries");
    if (axis < 0 || axis > 2) return fail("axis must be 0, 1 or 2");
    const int n_bins = n_borders - 1;
    for (int b = 0; b < n_bins; ++b) counts[b] = 0;
    for (size_t i = 0; i < s->n; ++i) {
        const entry_t *p = &s->e[i];
        if (p->deactivated) continue;
        int in = 0;
        for (int k = 0; k < n_types; ++k) in |= p->type == types[k];
        if (!in) continue;
        const double x = p->pos[axis];
        int lo = 0, hi = n_borders;  /* std::upper_bound */
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (x < borders[mid]) hi = mid;
            else lo = mid + 1;
        }
        if (lo < n_borders && lo >= 1 && lo < n_bins) counts[lo - 1]++;  /* CPUObservables.cpp:107-114 */
    }
    return 0;
}

int or_set_id_counter(or_system *s, uint64_t next_id) {
    s->next_id = next_id;
    return 0;
}

static uint64_t next_id(or_system *s) { return s->next_id++; }

static size_t add_entry(or_system *s, const entry_t *en) {
    /* DefaultDataContainer.h:79-90 */
    if (s->n_blanks > 0) {
        size_t idx = s->blanks[--s->n_blanks];
        s->e[idx] = *en;
        return idx;
    }
    if (s->n == s->cap) {
        s->cap = s->cap ? 2 * s->cap : 1024;
        s->e = (entry_t *)realloc(s->e, s->cap * sizeof(entry_t));
    }
    s->e[s->n] = *en;
    return s->n++;
}

static void remove_entry(or_system *s, size_t idx) {
    s->e[idx].deactivated = 1;
    if (s->n_blanks == s->cap_blanks) {
        s->cap_blanks = s->cap_blanks ? 2 * s->cap_blanks : 256;
        s->blanks = (size_t *)realloc(s->blanks, s->cap_blanks * sizeof(size_t));
    }
    s->blanks[s->n_blanks++] = idx;
}

int or_add_particles(or_system *s, int n, const double *pos, const int *types, const uint64_t *ids) {
    for (int i = 0; i < n; ++i) {
        entry_t en;
        memset(&en, 0, sizeof en);
        en.pos[0] = pos[3 * i];
        en.pos[1] = pos[3 * i + 1];
        en.pos[2] = pos[3 * i + 2];
        en.type = types[i];
        en.id = ids ? ids[i] : next_id(s);
        if (ids && ids[i] >= s->next_id) s->next_id = ids[i] + 1;
        add_entry(s, &en);
    }
    return 0;
}

static int push_bonded(bonded_t **arr, size_t *n, const bonded_t *b) {
    *arr = (bonded_t *)realloc(*arr, (*n + 1) * sizeof(bonded_t));
    (*arr)[(*n)++] = *b;
    return 0;
}

int or_add_bond(or_system *s, size_t i1, size_t i2, double k, double l0) {
    bonded_t b = {{i1, i2, 0, 0}, k, l0, 0};
    return push_bonded(&s->bonds, &s->n_bonds, &b);
}

int or_add_angle(or_system *s, size_t i1, size_t i2, size_t i3, double k, double theta0) {
    bonded_t b = {{i1, i2, i3, 0}, k, theta0, 0};
    return push_bonded(&s->angles, &s->n_angles, &b);
}

int or_add_dihedral(or_system *s, size_t i1, size_t i2, size_t i3, size_t i4, double k, int multiplicity,
                    double phi0) {
    bonded_t b = {{i1, i2, i3, i4}, k, phi0, multiplicity};
    return push_bonded(&s->dihedrals, &s->n_dihedrals, &b);
}

size_t or_n_entries(const or_system *s) { return s->n; }

size_t or_n_active(const or_system *s) {
    size_t c = 0;
    for (size_t i = 0; i < s->n; ++i) c += !s->e[i].deactivated;
    return c;
}

size_t or_get_particles(const or_system *s, size_t cap, double *pos, int *types, uint64_t *ids, double *forces) {
    size_t c = 0;
    for (size_t i = 0; i < s->n; ++i) {
        const entry_t *en = &s->e[i];
        if (en->deactivated) continue;
        if (c < cap) {
            if (pos) memcpy(pos + 3 * c, en->pos, 3 * sizeof(double));
            if (types) types[c] = en->type;
            if (ids) ids[c] = en->id;
            if (forces) memcpy(forces + 3 * c, en->force, 3 * sizeof(double));
        }
        ++c;
    }
    return c;
}

/* ------------------------------------------------------------------------------------------------
 * Cell-linked list: kernels/cpu/src/nl/CellLinkedList.cpp:52-142 (setUp), :147-169 (serial
 * fillBins), :230-247 (setUpBins); neighbour visit order CellLinkedList.h:304-313.
 * --------------------------------------------------------------------------------------------- */
static int cmp_size(const void *a, const void *b) {
    size_t x = *(const size_t *)a, y = *(const size_t *)b;
    return x < y ? -1 : (x > y);
}

static size_t cidx(const or_system *s, int i, int j, int k) {
    return ((size_t)i * s->dims[1] + j) * s->dims[2] + k; /* Index.h:137-139 row-major */
}

int or_nl_setup(or_system *s, double cutoff, int radius) {
    if (cutoff <= 0) return fail("The cutoff distance for setting up a neighbor list must be > 0");
    s->nl_cutoff = cutoff;
    s->nl_radius = radius;
    double width = cutoff / (double)radius;
    for (int d = 0; d < 3; ++d) {
        double q = floor(s->box[d] / width);
        s->dims[d] = (int)(q < 1. ? 1. : q);
        s->cell_size[d] = s->box[d] / (double)s->dims[d];
    }
    s->n_cells = (size_t)s->dims[0] * s->dims[1] * s->dims[2];
    size_t nn[3];
    for (int d = 0; d < 3; ++d) nn[d] = (size_t)s->dims[d] < (size_t)(2 * radius + 1) ? (size_t)s->dims[d] : (size_t)(2 * radius + 1);
    s->max_adj = nn[0] * nn[1] * nn[2];
    free(s->adj);
    s->adj = (size_t *)calloc(s->n_cells * (1 + s->max_adj), sizeof(size_t));
    size_t *tmp = (size_t *)malloc(sizeof(size_t) * (2 * radius + 1) * (2 * radius + 1) * (2 * radius + 1));
    for (int i = 0; i < s->dims[0]; ++i)
        for (int j = 0; j < s->dims[1]; ++j)
            for (int k = 0; k < s->dims[2]; ++k) {
                size_t me = cidx(s, i, j, k), cnt = 0;
                for (int ii = i - radius; ii <= i + radius; ++ii)
                    for (int jj = j - radius; jj <= j + radius; ++jj)
                        for (int kk = k - radius; kk <= k + radius; ++kk) {
                            if (ii == i && jj == j && kk == k) continue;
                            int a[3] = {ii, jj, kk};
                            int ok = 1;
                            for (int d = 0; d < 3; ++d) {
                                if (s->pbc[d]) a[d] = (a[d] % s->dims[d] + s->dims[d]) % s->dims[d];
                                if (a[d] < 0 || a[d] >= s->dims[d]) ok = 0;
                            }
                            if (ok) tmp[cnt++] = cidx(s, a[0], a[1], a[2]);
                        }
                qsort(tmp, cnt, sizeof(size_t), cmp_size);
                size_t u = 0;
                for (size_t q = 0; q < cnt; ++q) {
                    if (tmp[q] == me) continue;
                    if (u > 0 && tmp[u - 1] == tmp[q]) continue;
                    tmp[u++] = tmp[q];
                }
                size_t *row = s->adj + me * (1 + s->max_adj);
                row[0] = u;
                memcpy(row + 1, tmp, u * sizeof(size_t));
            }
    free(tmp);
    s->nl_set = 1;
    return or_nl_fill(s);
}

static int in_box(const or_system *s, const double *p) {
    /* CellLinkedList.cpp:188-192 */
    return -.5 * s->box[0] <= p[0] && .5 * s->box[0] > p[0] && -.5 * s->box[1] <= p[1] && .5 * s->box[1] > p[1] &&
           -.5 * s->box[2] <= p[2] && .5 * s->box[2] > p[2];
}

static int cell_coord(const or_system *s, const double *p, int d) {
    int c = (int)floor((p[d] + .5 * s->box[d]) / s->cell_size[d]);
    /* a coordinate a hair below +L/2 may round up to dims; the cell is geometrically the last one */
    if (c >= s->dims[d]) c = s->dims[d] - 1;
    if (c < 0) c = 0;
    return c;
}

int or_nl_fill(or_system *s) {
    if (!s->nl_set) return fail("Attempting to fill neighborlist bins, but cell structure is not set up yet");
    free(s->head);
    free(s->list);
    s->head = (size_t *)calloc(s->n_cells, sizeof(size_t));
    s->list = (size_t *)calloc(s->n + 1, sizeof(size_t));
    s->list_len = s->n + 1;
    if (s->cas_fill && s->threads > 1) {
        /* CellLinkedList.cpp:171-228: every thread pushes its particles onto the cell heads by atomic
         * exchange; the order within a cell then depends on the thread interleaving (the pair set does
         * not), so this fill serves the timing baseline, and the parity checks use the serial fill */
#pragma omp parallel for schedule(static) num_threads(s->threads)
        for (size_t i = 0; i < s->n; ++i) {
            const entry_t *en = &s->e[i];
            if (en->deactivated || !in_box(s, en->pos)) continue;
            size_t c = cidx(s, cell_coord(s, en->pos, 0), cell_coord(s, en->pos, 1), cell_coord(s, en->pos, 2));
            s->list[i + 1] = __atomic_exchange_n(&s->head[c], i + 1, __ATOMIC_RELAXED);
        }
        return 0;
    }
    for (size_t i = 0; i < s->n; ++i) {
        const entry_t *en = &s->e[i];
        if (en->deactivated || !in_box(s, en->pos)) continue;
        size_t c = cidx(s, cell_coord(s, en->pos, 0), cell_coord(s, en->pos, 1), cell_coord(s, en->pos, 2));
        s->list[i + 1] = s->head[c];
        s->head[c] = i + 1;
    }
    return 0;
}

void or_nl_dims(const or_system *s, int *dims) {
    for (int d = 0; d < 3; ++d) dims[d] = s->dims[d];
}

/* visit callback type */
typedef void (*visit_fn)(or_system *s, size_t p, size_t q, void *ud);

static void for_each_neighbor(or_system *s, size_t p, size_t cell, visit_fn fn, void *ud) {
    for (size_t q1 = s->head[cell]; q1 != 0; q1 = s->list[q1])
        if (q1 - 1 != p) fn(s, p, q1 - 1, ud);
    const size_t *row = s->adj + cell * (1 + s->max_adj);
    for (size_t a = 0; a < row[0]; ++a)
        for (size_t q1 = s->head[row[1 + a]]; q1 != 0; q1 = s->list[q1]) fn(s, p, q1 - 1, ud);
}

static void for_each_pair(or_system *s, visit_fn fn, void *ud) {
    for (size_t c = 0; c < s->n_cells; ++c)
        for (size_t p1 = s->head[c]; p1 != 0; p1 = s->list[p1]) for_each_neighbor(s, p1 - 1, c, fn, ud);
}

typedef struct {
    double rc2;
    size_t cap, n;
    uint64_t *out;
} pair_ud;

static void collect_pair(or_system *s, size_t p, size_t q, void *ud) {
    pair_ud *u = (pair_ud *)ud;
    if (s->e[q].deactivated) return;
    if (or_dist2(s->e[p].pos, s->e[q].pos, s->box, s->pbc) < u->rc2) {
        if (u->n < u->cap) {
            u->out[2 * u->n] = s->e[p].id;
            u->out[2 * u->n + 1] = s->e[q].id;
        }
        u->n++;
    }
}

size_t or_nl_pairs(const or_system *s, double cutoff, size_t cap, uint64_t *out) {
    pair_ud u = {cutoff * cutoff, cap, 0, out};
    if (!s->nl_set) return 0;
    for_each_pair((or_system *)s, collect_pair, &u);
    return u.n;
}

size_t or_brute_pairs(const or_system *s, double cutoff, size_t cap, uint64_t *out) {
    double rc2 = cutoff * cutoff;
    size_t c = 0;
    for (size_t i = 0; i < s->n; ++i) {
        if (s->e[i].deactivated || !in_box(s, s->e[i].pos)) continue;
        for (size_t j = 0; j < s->n; ++j) {
            if (i == j || s->e[j].deactivated || !in_box(s, s->e[j].pos)) continue;
            if (or_dist2(s->e[i].pos, s->e[j].pos, s->box, s->pbc) < rc2) {
                if (c < cap) {
                    out[2 * c] = s->e[i].id;
                    out[2 * c + 1] = s->e[j].id;
                }
                ++c;
            }
        }
    }
    return c;
}

/* ------------------------------------------------------------------------------------------------
 * Context::calculateMaxCutoff: readdy/main/model/Context.cpp:126-144.
 * --------------------------------------------------------------------------------------------- */
double or_max_cutoff(const or_system *s) {
    double m = 0.0;
    for (int i = 0; i < s->n_pots; ++i) {
        double c = sqrt(s->pots[i].cutoff2);
        if (c > m) m = c;
    }
    for (int i = 0; i < s->n_rx; ++i) {
        int k = s->rx[i].kind;
        if ((k == OR_RX_FUSION || k == OR_RX_ENZYMATIC) && s->rx[i].educt_distance > m) m = s->rx[i].educt_distance;
    }
    for (int i = 0; i < s->n_str; ++i) /* spatial topology reaction radii (Context.cpp:126-144) */
        if (s->str[i].radius > m) m = s->str[i].radius;
    return m;
}

/* ------------------------------------------------------------------------------------------------
 * CalculateForces: kernels/cpu/src/actions/CPUCalculateForces.cpp:46-212 (zero forces, topology
 * terms, order-2 full traversal with energy x 1/2), :214-268 (calculateOrder2), :270-293 (bonded).
 * --------------------------------------------------------------------------------------------- */
typedef struct {
    double energy;
    double virial[9];
} force_ud;

static void force_visit(or_system *s, size_t p, size_t q, void *ud) {
    force_ud *u = (force_ud *)ud;
    entry_t *ep = &s->e[p];
    const entry_t *eq = &s->e[q];
    if (eq->deactivated) return;
    double x[3];
    int found = 0;
    double my_e = 0.0;
    for (int k = 0; k < s->n_pots; ++k) {
        const pot_t *pt = &s->pots[k];
        if (!((pt->t1 == ep->type && pt->t2 == eq->type) || (pt->t1 == eq->type && pt->t2 == ep->type))) continue;
        if (!found) {
            or_shortest_difference(ep->pos, eq->pos, s->box, s->pbc, x);
            found = 1;
        }
        double d2 = dot3(x, x);
        if (d2 < pt->cutoff2) {
            double fu[3] = {0, 0, 0};
            my_e += or_pair_force_energy(pt->kind, pt->p, x, fu);
            ep->force[0] += fu[0];
            ep->force[1] += fu[1];
            ep->force[2] += fu[2];
            if (p < q) {
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) u->virial[3 * a + b] += (-1. * x[a]) * fu[b];
            }
        }
    }
    u->energy += 0.5 * my_e;
}

int or_calculate_forces(or_system *s, double *energy, double *virial9) {
    force_ud u;
    memset(&u, 0, sizeof u);
    for (size_t i = 0; i < s->n; ++i) s->e[i].force[0] = s->e[i].force[1] = s->e[i].force[2] = 0.0;
    double e_top = 0.0;
    for (size_t b = 0; b < s->n_bonds; ++b) {
        const bonded_t *t = &s->bonds[b];
        e_top += or_bond_force_energy(t->k, t->l0, s->e[t->i[0]].pos, s->e[t->i[1]].pos, s->box, s->pbc,
                                      s->e[t->i[0]].force, s->e[t->i[1]].force);
    }
    for (size_t b = 0; b < s->n_angles; ++b) {
        const bonded_t *t = &s->angles[b];
        e_top += or_angle_force_energy(t->k, t->l0, s->e[t->i[0]].pos, s->e[t->i[1]].pos, s->e[t->i[2]].pos, s->box,
                                       s->pbc, s->e[t->i[0]].force, s->e[t->i[1]].force, s->e[t->i[2]].force);
    }
    for (size_t b = 0; b < s->n_dihedrals; ++b) {
        const bonded_t *t = &s->dihedrals[b];
        e_top += or_dihedral_force_energy(t->k, t->mult, t->l0, s->e[t->i[0]].pos, s->e[t->i[1]].pos,
                                          s->e[t->i[2]].pos, s->e[t->i[3]].pos, s->box, s->pbc,
                                          s->e[t->i[0]].force, s->e[t->i[1]].force, s->e[t->i[2]].force,
                                          s->e[t->i[3]].force);
    }
    if (s->n_pots > 0) {
        if (!s->nl_set) return fail("forces need a neighbour list");
        if (s->threads > 1) {  /* each particle sums its own pairs: static partition over cells */
#pragma omp parallel num_threads(s->threads)
            {
                force_ud local;
                memset(&local, 0, sizeof local);
#pragma omp for schedule(static)
                for (size_t c = 0; c < s->n_cells; ++c)
                    for (size_t p1 = s->head[c]; p1 != 0; p1 = s->list[p1])
                        for_each_neighbor(s, p1 - 1, c, force_visit, &local);
#pragma omp critical
                {
                    u.energy += local.energy;
                    for (int a = 0; a < 9; ++a) u.virial[a] += local.virial[a];
                }
            }
        } else {
            for_each_pair(s, force_visit, &u);
        }
    }
    if (energy) *energy = e_top + u.energy;
    if (virial9) memcpy(virial9, u.virial, sizeof u.virial);
    return 0;
}

/* ------------------------------------------------------------------------------------------------
 * EulerBD: kernels/cpu/src/actions/CPUEulerBDIntegrator.cpp:51-76.  noise_by_id[3*id + d] when given,
 * Philox stream (h=id, t=step, 0, 0) otherwise.
 * --------------------------------------------------------------------------------------------- */
int or_euler_bd(or_system *s, double dt, const double *noise_by_id, size_t noise_len, uint32_t step) {
    double kbt = s->kbt;
    if (noise_by_id)
        for (size_t i = 0; i < s->n; ++i)
            if (!s->e[i].deactivated && 3 * s->e[i].id + 2 >= noise_len)
                return fail("noise buffer too short for particle id");
#pragma omp parallel for schedule(static) num_threads(s->threads) if (s->threads > 1)
    for (size_t i = 0; i < s->n; ++i) {
        entry_t *en = &s->e[i];
        if (en->deactivated) continue;
        double D = s->D[en->type];
        double xi[3];
        if (noise_by_id) {
            xi[0] = noise_by_id[3 * en->id];
            xi[1] = noise_by_id[3 * en->id + 1];
            xi[2] = noise_by_id[3 * en->id + 2];
        } else {
            uint32_t x[4];
            philox_draw(s->seed, (uint32_t)en->id, step, 0, 0, x);
            bd_normals3(x, xi);
        }
        double sc = sqrt(2. * D * dt);
        for (int d = 0; d < 3; ++d) {
            double det = en->force[d] * D * dt / kbt;
            double rnd = xi[d] * sc;
            en->pos[d] += rnd + det;
        }
        or_fix_position(en->pos, s->box, s->pbc);
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------------
 * Reactions.
 *   candidate search: CPUUncontrolledApproximation.cpp:70-133 and ReactionUtils.h:83-144
 *   acceptance: ReactionUtils.h:64-76 with approximateRate = false
 *   UA selection: CPUUncontrolledApproximation.cpp:197-254
 *   Gillespie selection: ReactionUtils.cpp:55-123 + include/readdy/common/algorithm.h:62-136
 *   performReaction: ReactionUtils.h:146-227; data update: DefaultDataContainer.h:121-139
 * Event endpoints are ordered canonically so that both the reference's slot rule and the HIP
 * engine place reaction products identically: entry1 is the particle whose type equals the
 * reaction's first educt; for equal educt types the lower id.
 * --------------------------------------------------------------------------------------------- */
typedef struct {
    event_t *ev;
    size_t n, cap;
    double dt;
    uint32_t step;
    int draw; /* 1: Bernoulli prefilter (UA), 0: all candidates */
} ev_ud;

static void push_event(ev_ud *u, const event_t *e) {
    if (u->n == u->cap) {
        u->cap = u->cap ? 2 * u->cap : 1024;
        u->ev = (event_t *)realloc(u->ev, u->cap * sizeof(event_t));
    }
    u->ev[u->n++] = *e;
}

static double accept_p(double rate, double dt) { return 1 - exp(-rate * dt); }

static void canonical_order(const or_system *s, const rx_t *r, size_t *a, size_t *b) {
    const entry_t *ea = &s->e[*a], *eb = &s->e[*b];
    int swap = 0;
    if (r->educts[0] != r->educts[1]) {
        if (ea->type != r->educts[0]) swap = 1;
    } else if (eb->id < ea->id) {
        swap = 1;
    }
    if (swap) {
        size_t t = *a;
        *a = *b;
        *b = t;
    }
}

static void rx_visit(or_system *s, size_t p, size_t q, void *ud) {
    ev_ud *u = (ev_ud *)ud;
    if (!(p > q)) return; /* CPUUncontrolledApproximation.cpp:101 */
    const entry_t *ep = &s->e[p], *eq = &s->e[q];
    if (eq->deactivated) return;
    double d2 = -1.0;
    for (int r = 0; r < s->n_rx; ++r) {
        const rx_t *rx = &s->rx[r];
        if (!(rx->kind == OR_RX_FUSION || rx->kind == OR_RX_ENZYMATIC)) continue;
        if (!((rx->educts[0] == ep->type && rx->educts[1] == eq->type) ||
              (rx->educts[1] == ep->type && rx->educts[0] == eq->type)))
            continue;
        if (d2 < 0) d2 = or_dist2(eq->pos, ep->pos, s->box, s->pbc);
        if (!(rx->rate > 0 && d2 < rx->educt_distance2)) continue;
        uint64_t lo = ep->id < eq->id ? ep->id : eq->id, hi = ep->id < eq->id ? eq->id : ep->id;
        uint32_t x[4];
        philox_draw(s->seed, (uint32_t)lo, u->step, 2u | ((uint32_t)r << 8), (uint32_t)hi, x);
        if (u->draw && !(or_u53(x[0], x[1]) < accept_p(rx->rate, u->dt))) continue;
        event_t e;
        memset(&e, 0, sizeof e);
        e.n_educts = 2;
        e.idx1 = p;
        e.idx2 = q;
        canonical_order(s, rx, &e.idx1, &e.idx2);
        e.rx = r;
        e.rate = rx->rate;
        e.key = ((uint64_t)x[2] << 32) | x[3];
        e.gkey = -log(or_u53(x[2], x[3]) + 0x1p-54) / rx->rate;
        push_event(u, &e);
    }
}

static void gather(or_system *s, ev_ud *u) {
    for (size_t i = 0; i < s->n; ++i) {
        const entry_t *en = &s->e[i];
        if (en->deactivated) continue;
        for (int r = 0; r < s->n_rx; ++r) {
            const rx_t *rx = &s->rx[r];
            if (rx->kind == OR_RX_FUSION || rx->kind == OR_RX_ENZYMATIC) continue;
            if (rx->educts[0] != en->type || !(rx->rate > 0)) continue;
            uint32_t x[4];
            philox_draw(s->seed, (uint32_t)en->id, u->step, 1u | ((uint32_t)r << 8), 0, x);
            if (u->draw && !(or_u53(x[0], x[1]) < accept_p(rx->rate, u->dt))) continue;
            event_t e;
            memset(&e, 0, sizeof e);
            e.n_educts = 1;
            e.idx1 = e.idx2 = i;
            e.rx = r;
            e.rate = rx->rate;
            e.key = ((uint64_t)x[2] << 32) | x[3];
            e.gkey = -log(or_u53(x[2], x[3]) + 0x1p-54) / rx->rate;
            push_event(u, &e);
        }
    }
    if (!s->nl_set) return;
    if (s->threads > 1 && u->draw) {  /* per-thread candidate lists, appended in thread order */
#pragma omp parallel num_threads(s->threads)
        {
            ev_ud local = *u;
            local.ev = NULL;
            local.n = local.cap = 0;
#pragma omp for schedule(static)
            for (size_t c = 0; c < s->n_cells; ++c)
                for (size_t p1 = s->head[c]; p1 != 0; p1 = s->list[p1]) for_each_neighbor(s, p1 - 1, c, rx_visit, &local);
#pragma omp for ordered schedule(static, 1)
            for (int t = 0; t < s->threads; ++t) {
#pragma omp ordered
                for (size_t k = 0; k < local.n; ++k) push_event(u, &local.ev[k]);
            }
            free(local.ev);
        }
    } else {
        for_each_pair(s, rx_visit, u);
    }
}

size_t or_reaction_candidates(or_system *s, size_t cap, uint64_t *out) {
    ev_ud u;
    memset(&u, 0, sizeof u);
    u.draw = 0;
    gather(s, &u);
    for (size_t i = 0; i < u.n && i < cap; ++i) {
        out[4 * i] = (uint64_t)u.ev[i].n_educts;
        out[4 * i + 1] = s->e[u.ev[i].idx1].id;
        out[4 * i + 2] = s->e[u.ev[i].idx2].id;
        out[4 * i + 3] = (uint64_t)u.ev[i].rx;
    }
    size_t n = u.n;
    free(u.ev);
    return n;
}

typedef struct {
    entry_t *new_entries;
    size_t n_new, cap_new;
    size_t *removed;
    size_t n_removed, cap_removed;
} update_t;

static void upd_new(update_t *u, const entry_t *e) {
    if (u->n_new == u->cap_new) {
        u->cap_new = u->cap_new ? 2 * u->cap_new : 64;
        u->new_entries = (entry_t *)realloc(u->new_entries, u->cap_new * sizeof(entry_t));
    }
    u->new_entries[u->n_new++] = *e;
}

static void upd_removed(update_t *u, size_t idx) {
    if (u->n_removed == u->cap_removed) {
        u->cap_removed = u->cap_removed ? 2 * u->cap_removed : 64;
        u->removed = (size_t *)realloc(u->removed, u->cap_removed * sizeof(size_t));
    }
    u->removed[u->n_removed++] = idx;
}

static void perform_reaction(or_system *s, const event_t *ev, update_t *up, uint32_t step) {
    const rx_t *r = &s->rx[ev->rx];
    entry_t *e1 = &s->e[ev->idx1];
    entry_t *e2 = &s->e[ev->idx2];
    or_reaction_record *rec = NULL;
    if (s->record) {
        /* ReactionUtils.h:155-162: type, where = plain mean of the two positions, educts, types_from;
         * the caller then applies fixPosition (CPUUncontrolledApproximation.cpp:214-217) */
        if (s->n_recs == s->cap_recs) {
            s->cap_recs = s->cap_recs ? 2 * s->cap_recs : 64;
            s->recs = (or_reaction_record *)realloc(s->recs, s->cap_recs * sizeof(or_reaction_record));
        }
        rec = &s->recs[s->n_recs++];
        memset(rec, 0, sizeof *rec);
        rec->reaction = (uint32_t)ev->rx;
        rec->type = r->kind;
        for (int d = 0; d < 3; ++d) rec->where[d] = (e1->pos[d] + e2->pos[d]) / 2.;
        or_fix_position(rec->where, s->box, s->pbc);
        rec->educts[0] = e1->id;
        rec->educts[1] = e2->id;
        rec->types_from[0] = (uint32_t)e1->type;
        rec->types_from[1] = (uint32_t)e2->type;
    }
    switch (r->kind) {
    case OR_RX_DECAY:
        upd_removed(up, ev->idx1);
        break;
    case OR_RX_CONVERSION:
        e1->type = r->products[0];
        e1->id = next_id(s);
        if (rec) rec->products[0] = e1->id;
        break;
    case OR_RX_ENZYMATIC:
        if (e1->type == r->educts[1]) {
            e2->type = r->products[0];
            e2->id = next_id(s);
        } else {
            e1->type = r->products[0];
            e1->id = next_id(s);
        }
        if (rec) {
            rec->products[0] = e1->id;
            rec->products[1] = e2->id;
        }
        break;
    case OR_RX_FISSION: {
        uint32_t x[4];
        double n3[3];
        philox_draw(s->seed, (uint32_t)e1->id, step, 3, 0, x);
        normals3(x, n3);
        double nn = sqrt(dot3(n3, n3));
        for (int d = 0; d < 3; ++d) n3[d] /= nn;
        entry_t ne;
        memset(&ne, 0, sizeof ne);
        for (int d = 0; d < 3; ++d) ne.pos[d] = e1->pos[d] - r->w2 * r->product_distance * n3[d];
        or_fix_position(ne.pos, s->box, s->pbc);
        ne.type = r->products[1];
        ne.id = next_id(s);
        upd_new(up, &ne);
        e1->type = r->products[0];
        e1->id = next_id(s);
        for (int d = 0; d < 3; ++d) e1->pos[d] += r->w1 * r->product_distance * n3[d];
        or_fix_position(e1->pos, s->box, s->pbc);
        if (rec) {
            rec->products[0] = e1->id;
            rec->products[1] = ne.id;
        }
        break;
    }
    case OR_RX_FUSION: {
        double diff[3];
        or_shortest_difference(e1->pos, e2->pos, s->box, s->pbc, diff);
        double w = (r->educts[0] == e1->type) ? r->w1 : r->w2;
        entry_t ne;
        memset(&ne, 0, sizeof ne);
        for (int d = 0; d < 3; ++d) ne.pos[d] = e1->pos[d] + w * diff[d];
        or_fix_position(ne.pos, s->box, s->pbc);
        ne.type = r->products[0];
        ne.id = next_id(s);
        upd_new(up, &ne);
        upd_removed(up, ev->idx1);
        upd_removed(up, ev->idx2);
        if (rec) rec->products[0] = ne.id;
        break;
    }
    }
    s->counts[ev->rx]++;
}

static void apply_update(or_system *s, update_t *up) {
    /* DefaultDataContainer.h:121-139 */
    size_t it_del = 0;
    for (size_t i = 0; i < up->n_new; ++i) {
        if (it_del < up->n_removed) {
            s->e[up->removed[it_del]] = up->new_entries[i];
            ++it_del;
        } else {
            add_entry(s, &up->new_entries[i]);
        }
    }
    while (it_del < up->n_removed) remove_entry(s, up->removed[it_del++]);
    free(up->new_entries);
    free(up->removed);
}

static int shares(const event_t *a, const event_t *b) {
    return a->idx1 == b->idx1 || a->idx1 == b->idx2 || a->idx2 == b->idx1 || a->idx2 == b->idx2;
}

static int cmp_key(const void *a, const void *b) {
    const event_t *x = (const event_t *)a, *y = (const event_t *)b;
    return x->key < y->key ? -1 : (x->key > y->key);
}

static int cmp_gkey(const void *a, const void *b) {
    const event_t *x = (const event_t *)a, *y = (const event_t *)b;
    if (x->gkey < y->gkey) return -1;
    if (x->gkey > y->gkey) return 1;
    return x->key < y->key ? -1 : (x->key > y->key);
}

/* algorithm.h:62-136 literal, with x ~ U(0, cumulative) drawn from Philox (seed, (trial, draw#, 7, 0)) */
static void perform_events_literal(event_t *ev, size_t n, uint64_t seed, uint32_t trial,
                                   int (*eval)(void *, const event_t *), void *eu) {
    if (n == 0) return;
    size_t n_deact = 0;
    double cum = 0.;
    for (size_t i = 0; i < n; ++i) {
        ev[i].cumulative = ev[i].rate + cum;
        cum = ev[i].cumulative;
    }
    uint32_t draw = 0;
    while (n_deact < n) {
        double cr = ev[n - n_deact - 1].cumulative;
        uint32_t x[4];
        philox_draw(seed, trial, draw++, 7, 0, x);
        double xr = or_u53(x[0], x[1]) * cr;
        /* lower_bound on cumulative over [0, n - n_deact) */
        size_t lo = 0, hi = n - n_deact;
        while (lo < hi) {
            size_t mid = (lo + hi) / 2;
            if (ev[mid].cumulative < xr) lo = mid + 1;
            else hi = mid;
        }
        if (lo >= n - n_deact) lo = n - n_deact - 1; /* xr == cr exactly cannot pass the bound */
        event_t *it = &ev[lo];
        /* shouldEvaluate: the acceptance probability of the tried event (carried in gkey) */
        int acc = or_u53(x[2], x[3]) < it->gkey;
        if (acc) {
            eval(eu, it);
            event_t done = *it;
            size_t k = 0;
            double cs = 0.0;
            while (k < n - n_deact) {
                if (shares(&ev[k], &done)) {
                    ++n_deact;
                    event_t t = ev[k];
                    ev[k] = ev[n - n_deact];
                    ev[n - n_deact] = t;
                } else {
                    cs += ev[k].rate;
                    ev[k].cumulative = cs;
                    ++k;
                }
            }
        } else {
            ++n_deact;
            event_t t = *it;
            *it = ev[n - n_deact];
            ev[n - n_deact] = t;
            it->cumulative = it->rate;
            if (it > ev) it->cumulative += (it - 1)->cumulative;
            double cs = it->cumulative;
            for (event_t *j = it + 1; j < ev + (n - n_deact); ++j) {
                cs += j->rate;
                j->cumulative = cs;
            }
        }
    }
}

typedef struct {
    uint8_t *performed;
    const int64_t *tag;
} abs_ud;

static int abs_eval(void *ud, const event_t *e) {
    abs_ud *u = (abs_ud *)ud;
    u->performed[e->rx] = 1;
    return 0;
}

int or_perform_events_abstract(int n, const int64_t *p1, const int64_t *p2, const double *rate,
                               const double *acc, int mode, uint64_t seed, uint32_t trial, uint8_t *performed) {
    event_t *ev = (event_t *)calloc((size_t)n > 0 ? (size_t)n : 1, sizeof(event_t));
    for (int i = 0; i < n; ++i) {
        ev[i].idx1 = (size_t)p1[i];
        ev[i].idx2 = p2[i] < 0 ? (size_t)p1[i] : (size_t)p2[i];
        ev[i].rate = rate[i];
        ev[i].rx = i;
        ev[i].gkey = acc[i]; /* literal path reads the acceptance probability from here */
        performed[i] = 0;
    }
    abs_ud u = {performed, NULL};
    if (mode == OR_GILLESPIE_LITERAL) {
        perform_events_literal(ev, (size_t)n, seed, trial, abs_eval, &u);
    } else {
        /* keyed: every event draws its own acceptance and E-S key; greedy over accepted in key order */
        size_t m = 0;
        for (int i = 0; i < n; ++i) {
            uint32_t x[4];
            philox_draw(seed, trial, (uint32_t)i, 8, 0, x);
            if (!(or_u53(x[0], x[1]) < acc[i])) continue;
            ev[m] = ev[i];
            ev[m].gkey = -log(or_u53(x[2], x[3]) + 0x1p-54) / rate[i];
            ev[m].key = ((uint64_t)x[2] << 32) | x[3];
            ++m;
        }
        qsort(ev, m, sizeof(event_t), cmp_gkey);
        for (size_t i = 0; i < m; ++i) {
            if (ev[i].blocked) continue;
            performed[ev[i].rx] = 1;
            for (size_t j = i + 1; j < m; ++j)
                if (shares(&ev[j], &ev[i])) ev[j].blocked = 1;
        }
    }
    free(ev);
    return 0;
}

int or_reactions(or_system *s, int mode, double dt, uint32_t step, uint64_t *n_performed) {
    /* counts and records are reset at the start of each handler call: CPUUncontrolledApproximation.cpp:143-148 */
    memset(s->counts, 0, sizeof s->counts);
    s->n_recs = 0;
    ev_ud u;
    memset(&u, 0, sizeof u);
    u.dt = dt;
    u.step = step;
    u.draw = 1; /* keyed modes prefilter by the per-event Bernoulli draw */
    if (mode == OR_GILLESPIE_LITERAL) u.draw = 0;
    gather(s, &u);
    update_t up;
    memset(&up, 0, sizeof up);
    uint64_t done = 0;
    if (mode == OR_UA_KEYED || mode == OR_GILLESPIE_KEYED) {
        qsort(u.ev, u.n, sizeof(event_t), mode == OR_UA_KEYED ? cmp_key : cmp_gkey);
        for (size_t i = 0; i < u.n; ++i) {
            event_t *ev = &u.ev[i];
            if (ev->blocked) continue;
            perform_reaction(s, ev, &up, step);
            ++done;
            for (size_t j = i + 1; j < u.n; ++j)
                if (shares(&u.ev[j], ev)) u.ev[j].blocked = 1;
        }
    } else {
        /* literal performEvents: shouldEval draws 1-exp(-rate dt) per tried event */
        for (size_t i = 0; i < u.n; ++i) u.ev[i].gkey = accept_p(u.ev[i].rate, dt);
        event_t *ev = u.ev;
        size_t n = u.n, n_deact = 0;
        double cum = 0.;
        for (size_t i = 0; i < n; ++i) {
            ev[i].cumulative = ev[i].rate + cum;
            cum = ev[i].cumulative;
        }
        uint32_t draw = 0;
        while (n_deact < n) {
            double cr = ev[n - n_deact - 1].cumulative;
            uint32_t x[4];
            philox_draw(s->seed, step, draw++, 9, 0, x);
            double xr = or_u53(x[0], x[1]) * cr;
            size_t lo = 0, hi = n - n_deact;
            while (lo < hi) {
                size_t mid = (lo + hi) / 2;
                if (ev[mid].cumulative < xr) lo = mid + 1;
                else hi = mid;
            }
            if (lo >= n - n_deact) lo = n - n_deact - 1;
            event_t *it = &ev[lo];
            if (or_u53(x[2], x[3]) < it->gkey) {
                perform_reaction(s, it, &up, step);
                ++done;
                event_t d = *it;
                size_t k = 0;
                double cs = 0.0;
                while (k < n - n_deact) {
                    if (shares(&ev[k], &d)) {
                        ++n_deact;
                        event_t t = ev[k];
                        ev[k] = ev[n - n_deact];
                        ev[n - n_deact] = t;
                    } else {
                        cs += ev[k].rate;
                        ev[k].cumulative = cs;
                        ++k;
                    }
                }
            } else {
                ++n_deact;
                event_t t = *it;
                *it = ev[n - n_deact];
                ev[n - n_deact] = t;
                it->cumulative = it->rate;
                if (it > ev) it->cumulative += (it - 1)->cumulative;
                double cs = it->cumulative;
                for (event_t *j = it + 1; j < ev + (n - n_deact); ++j) {
                    cs += j->rate;
                    j->cumulative = cs;
                }
            }
        }
    }
    apply_update(s, &up);
    free(u.ev);
    if (n_performed) *n_performed = done;
    return 0;
}

/* ------------------------------------------------------------------------------------------------
 * DetailedBalance.
 *   reversible pairs: DetailedBalance::searchReversibleReactions (readdy/main/model/Actions.cpp:432-519)
 *   their configuration: ReversibleReactionConfig (Actions.cpp:67-367); the integrals of
 *     4 pi r^2 exp(-U(r)/kBT) by adaptive Simpson quadrature to a relative 1e-13 (the reference uses
 *     adaptive Gauss-Kronrod, common/integration.h; the values agree to the quadrature accuracy)
 *   event loop: SCPUDetailedBalance::perform (kernels/singlecpu/src/actions/SCPUReactionImpls.cpp:407-555)
 *     with the keyed Gillespie selection: the events that pass their Bernoulli draw, in key order;
 *     each one not sharing a particle with an earlier evaluated event is evaluated on the state the
 *     earlier ones left (accepted or not, it blocks the later events that share a particle)
 *   proposal: performReversibleReactionEvent (SCPUReactionImpls.cpp:579-743); rollback with fresh
 *     ids: ParticleBackup / generateBackwardUpdate (SCPUReactionImpls.cpp:362-405, 557-577)
 *   energy: calculateEnergies (SCPUReactionImpls.cpp:745-781), the pair potentials only; here the
 *     change is summed directly over the particles that leave and arrive, brute force over all
 *     particles inside the box, each potential gated by d^2 < cutoff^2 as the force pass gates it
 * --------------------------------------------------------------------------------------------- */
#define OR_FISSION_POINTS 1000

typedef struct rev_s {
    int kind; /* 0 FusionFission, 1 ConversionConversion, 2 EnzymaticEnzymatic */
    int fwd, bwd, n_lhs, n_rhs;
    int lhs[2], rhs[2];
    double reaction_radius, lhs_radius, rhs_radius, lhs_volume, rhs_volume;
    double eff_lhs_int, eff_rhs_int, eff_lhs_rx, eff_rhs_rx;
    double prefactor, total_volume, K, kf, kb;
    double radii[OR_FISSION_POINTS + 1], cdf[OR_FISSION_POINTS + 1];
} rev_t;

/* U(r) of the potentials between types a and b, as calculateEnergy((r, 0, 0)) sums them (no gate) */
static double rev_u(const or_system *s, int a, int b, double r) {
    double x[3] = {r, 0., 0.}, f[3] = {0., 0., 0.}, u = 0.;
    for (int k = 0; k < s->n_pots; ++k) {
        const pot_t *pt = &s->pots[k];
        if ((pt->t1 == a && pt->t2 == b) || (pt->t1 == b && pt->t2 == a)) u += or_pair_force_energy(pt->kind, pt->p, x, f);
    }
    return u;
}

typedef struct {
    const or_system *s;
    int a, b, any;
} rev_f;

static double rev_integrand(const rev_f *f, double r) {
    const double u = f->any ? rev_u(f->s, f->a, f->b, r) : 0.;
    const double v = 4. * 3.14159265358979323846 * r * r * exp(-1. * u / f->s->kbt);
    /* r = 0 under a repulsive core (LJ, SE: inf - inf): the integrand's limit there is 0; the reference's
     * Gauss-Kronrod nodes never reach the end points */
    return isfinite(v) ? v : 0.;
}

static double simpson_rec(const rev_f *f, double a, double b, double fa, double fm, double fb, double whole, double tol,
                          int depth) {
    const double m = .5 * (a + b), lm = .5 * (a + m), rm = .5 * (m + b);
    const double flm = rev_integrand(f, lm), frm = rev_integrand(f, rm);
    const double left = (m - a) / 6. * (fa + 4. * flm + fm), right = (b - m) / 6. * (fm + 4. * frm + fb);
    const double d = left + right - whole;
    /* converged, or the difference is at the rounding level of the panel (a steep wall of the potential
     * would otherwise recurse to the depth limit everywhere around it) */
    if (depth <= 0 || !(fabs(d) > 15. * tol) || fabs(d) <= 64. * 2.220446049250313e-16 * fabs(left + right) ||
        fabs(d) < 1e-290)  /* subnormal values carry no relative precision */
        return left + right + d / 15.;
    return simpson_rec(f, a, m, fa, flm, fm, left, .5 * tol, depth - 1) +
           simpson_rec(f, m, b, fm, frm, fb, right, .5 * tol, depth - 1);
}

/* the integral over [a, b] to an absolute tolerance of 1e-13 times scale (scale <= 0: the integral's own
 * coarse estimate) */
static double rev_integrate_scaled(const rev_f *f, double a, double b, double scale) {
    if (!(a < b)) return 0.;
    /* a first pass on 64 panels fixes the scale of the absolute tolerance */
    double coarse = 0.;
    const int np = 64;
    for (int i = 0; i < np; ++i) {
        const double x0 = a + (b - a) * i / np, x1 = a + (b - a) * (i + 1) / np;
        coarse += (x1 - x0) / 6. * (rev_integrand(f, x0) + 4. * rev_integrand(f, .5 * (x0 + x1)) + rev_integrand(f, x1));
    }
    const double tol = 1e-13 * (scale > 0. ? scale : fabs(coarse)) / np;
    double total = 0.;
    for (int i = 0; i < np; ++i) {
        const double x0 = a + (b - a) * i / np, x1 = a + (b - a) * (i + 1) / np;
        const double f0 = rev_integrand(f, x0), fm = rev_integrand(f, .5 * (x0 + x1)), f1 = rev_integrand(f, x1);
        total += simpson_rec(f, x0, x1, f0, fm, f1, (x1 - x0) / 6. * (f0 + 4. * fm + f1), tol, 40);
    }
    return total;
}

static double rev_integrate(const rev_f *f, double a, double b) { return rev_integrate_scaled(f, a, b, 0.); }

static int n_educts_of(int kind) { return (kind == OR_RX_FUSION || kind == OR_RX_ENZYMATIC) ? 2 : 1; }
static int n_products_of(int kind) {
    return kind == OR_RX_DECAY ? 0 : ((kind == OR_RX_FISSION || kind == OR_RX_ENZYMATIC) ? 2 : 1);
}

static double pot_cutoff_of_pair(const or_system *s, int a, int b, int *any) {
    double m = 0.;
    *any = 0;
    for (int k = 0; k < s->n_pots; ++k) {
        const pot_t *pt = &s->pots[k];
        if ((pt->t1 == a && pt->t2 == b) || (pt->t1 == b && pt->t2 == a)) {
            const double c = sqrt(pt->cutoff2);
            if (!*any || c > m) m = c;
            *any = 1;
        }
    }
    return m;
}

static int rev_configure(const or_system *s, int fwd, int bwd, rev_t *c) {
    memset(c, 0, sizeof *c);
    const rx_t *F = &s->rx[fwd], *B = &s->rx[bwd];
    int nl = n_educts_of(F->kind), nr = n_educts_of(B->kind);
    if (nl == 2 && nr == 1) c->kind = 0;
    else if (nl == 1 && nr == 2) { /* the fusion is the forward reaction */
        const rx_t *t = F;
        F = B;
        B = t;
        int ti = fwd;
        fwd = bwd;
        bwd = ti;
        nl = 2;
        nr = 1;
        c->kind = 0;
    } else if (nl == 1 && nr == 1) c->kind = 1;
    else if (nl == 2 && nr == 2) c->kind = 2;
    else return fail("Type of reversible reaction cannot be determined");
    c->fwd = fwd;
    c->bwd = bwd;
    c->n_lhs = nl;
    c->n_rhs = nr;
    /* Reaction::educts() is a two-entry array; an order-1 reaction's second entry is 0 there, which the
     * species comparison below sees as well */
    c->lhs[0] = F->educts[0];
    c->lhs[1] = nl == 2 ? F->educts[1] : 0;
    c->rhs[0] = B->educts[0];
    c->rhs[1] = nr == 2 ? B->educts[1] : 0;
    int bp[2] = {B->products[0], n_products_of(B->kind) == 2 ? B->products[1] : 0};
    int fp[2] = {F->products[0], n_products_of(F->kind) == 2 ? F->products[1] : 0};
    if (!(c->lhs[0] == bp[0] && c->lhs[1] == bp[1]) || !(c->rhs[0] == fp[0] && c->rhs[1] == fp[1])) {
        if (!(c->kind == 0 && c->lhs[0] == bp[1] && c->lhs[1] == bp[0] && fp[0] == c->rhs[0]))
            return fail("The given reactions are not reversible with respect to particle species");
    }
    if (c->kind == 0) {
        if (F->educts[0] == bp[0] && F->educts[1] == bp[1]) {
            if (!(F->w1 == B->w1 && F->w2 == B->w2)) return fail("The given reactions are not reversible with respect to weights");
        } else if (F->educts[0] == bp[1] && F->educts[1] == bp[0]) {
            if (!(F->w1 == B->w2 && F->w2 == B->w1)) return fail("The given reactions are not reversible with respect to weights");
        }
    }
    if (c->kind == 0) {
        if (F->educt_distance != B->product_distance) return fail("Reaction radii forward and backward must be equal");
        c->reaction_radius = F->educt_distance;
    } else if (c->kind == 2) {
        if (F->educt_distance != B->educt_distance) return fail("Reaction radii forward and backward must be equal");
        c->reaction_radius = F->educt_distance;
    }
    rev_f fl = {s, c->lhs[0], c->lhs[1], 0}, fr = {s, c->rhs[0], c->rhs[1], 0};
    if (c->kind != 1) c->lhs_radius = pot_cutoff_of_pair(s, c->lhs[0], c->lhs[1], &fl.any);
    if (c->kind == 2) c->rhs_radius = pot_cutoff_of_pair(s, c->rhs[0], c->rhs[1], &fr.any);
    const double pi = 3.14159265358979323846;
    if (fl.any) c->lhs_volume = 4. / 3. * pi * pow(c->lhs_radius, 3);
    if (fr.any) c->rhs_volume = 4. / 3. * pi * pow(c->rhs_radius, 3);
    c->eff_lhs_int = rev_integrate(&fl, 0., c->lhs_radius);
    c->eff_rhs_int = rev_integrate(&fr, 0., c->rhs_radius);
    c->eff_lhs_rx = rev_integrate(&fl, 0., c->reaction_radius);
    c->eff_rhs_rx = rev_integrate(&fr, 0., c->reaction_radius);
    if (c->kind == 0) {
        const double dr = c->reaction_radius / (double)OR_FISSION_POINTS;
        double r = 0.;
        for (int i = 0; i <= OR_FISSION_POINTS; ++i) {
            c->radii[i] = r;
            r += dr;
        }
        /* cumulative panels, each to a tolerance relative to the whole range (the reference integrates
         * each cumulative value from 0 to a relative 1e-12) */
        const double whole = c->eff_lhs_rx;
        c->cdf[0] = 0.;
        for (int i = 1; i <= OR_FISSION_POINTS; ++i)
            c->cdf[i] = c->cdf[i - 1] + rev_integrate_scaled(&fl, c->radii[i - 1], c->radii[i], whole);
        const double last = c->cdf[OR_FISSION_POINTS];
        if (last >= 1e-16) {
            for (int i = 0; i <= OR_FISSION_POINTS; ++i) c->cdf[i] /= last;
        } else {
            for (int i = 0; i < OR_FISSION_POINTS; ++i) c->cdf[i] = 0.;
            c->cdf[OR_FISSION_POINTS] = 1.;
        }
    }
    c->prefactor = c->kind == 2 ? c->eff_lhs_rx / c->eff_rhs_rx : 1.;
    c->total_volume = s->box[0] * s->box[1] * s->box[2];
    const double kf = F->rate, kb = B->rate, V = c->total_volume;
    if (c->kind == 0) {
        c->K = (kb / kf) * (V - c->lhs_volume + c->eff_lhs_int) / c->eff_lhs_rx;
        c->kb = kb;
        c->kf = c->kb * V / c->K;
    } else if (c->kind == 1) {
        c->kf = kf;
        c->kb = kb;
        c->K = c->kb / c->kf;
    } else {
        c->kf = kf * V * c->eff_lhs_rx / (V - c->lhs_volume + c->eff_lhs_int);
        c->kb = kb * V * c->eff_rhs_rx / (V - c->rhs_volume + c->eff_rhs_int);
        c->K = c->kb / c->kf;
    }
    return 0;
}

static int rev_build(or_system *s) {
    free(s->rev);
    s->rev = NULL;
    s->n_rev = 0;
    for (int i = 0; i < OR_MAX_RX; ++i) s->rev_of_rx[i] = -1;
    int cap = 0;
    int o1[OR_MAX_RX], o2[OR_MAX_RX], n1 = 0, n2 = 0;
    for (int i = 0; i < s->n_rx; ++i) {
        if (n_educts_of(s->rx[i].kind) == 1) o1[n1++] = i;
        else o2[n2++] = i;
    }
    int pairs[2 * OR_MAX_RX * OR_MAX_RX / 8][2];
    int np = 0;
    const int max_pairs = (int)(sizeof pairs / sizeof pairs[0]);
    /* FusionFission: (fusion, fission) with the fission's product distance = the fusion's educt distance */
    for (int a = 0; a < n2; ++a)
        for (int b = 0; b < n1; ++b) {
            const rx_t *r2 = &s->rx[o2[a]], *r1 = &s->rx[o1[b]];
            if (!(n_products_of(r2->kind) == 1 && n_products_of(r1->kind) == 2 && r1->product_distance == r2->educt_distance))
                continue;
            const int same = r2->educts[0] == r1->products[0] && r2->educts[1] == r1->products[1] && r1->w1 == r2->w1 &&
                             r1->w2 == r2->w2;
            /* Actions.cpp:453-455 compare educts()[0] with products()[1] twice */
            const int flipped = r2->educts[0] == r1->products[1] && r1->w1 == r2->w2 && r1->w2 == r2->w1;
            if (r1->educts[0] == r2->products[0] && (same || flipped)) {
                if (np >= max_pairs) return fail("too many reversible pairs");
                pairs[np][0] = o2[a];
                pairs[np][1] = o1[b];
                ++np;
            }
        }
    for (int i = 0; i < n1; ++i)
        for (int j = i + 1; j < n1; ++j) {
            const rx_t *ri = &s->rx[o1[i]], *rj = &s->rx[o1[j]];
            if (n_products_of(ri->kind) == 1 && n_products_of(rj->kind) == 1 && ri->educts[0] == rj->products[0] &&
                ri->products[0] == rj->educts[0]) {
                if (np >= max_pairs) return fail("too many reversible pairs");
                pairs[np][0] = o1[i];
                pairs[np][1] = o1[j];
                ++np;
            }
        }
    for (int i = 0; i < n2; ++i)
        for (int j = i + 1; j < n2; ++j) {
            const rx_t *ri = &s->rx[o2[i]], *rj = &s->rx[o2[j]];
            if (n_products_of(ri->kind) == 2 && n_products_of(rj->kind) == 2 && ri->educt_distance == rj->educt_distance &&
                ri->educts[0] == rj->products[0] && ri->educts[1] == rj->products[1] && ri->products[0] == rj->educts[0] &&
                ri->products[1] == rj->educts[1]) {
                if (np >= max_pairs) return fail("too many reversible pairs");
                pairs[np][0] = o2[i];
                pairs[np][1] = o2[j];
                ++np;
            }
        }
    cap = np > 0 ? np : 1;
    s->rev = (rev_t *)calloc((size_t)cap, sizeof(rev_t));
    for (int k = 0; k < np; ++k)
        if (rev_configure(s, pairs[k][0], pairs[k][1], &s->rev[k])) return -1;
    s->n_rev = np;
    /* equivalentReversibleReactions (DetailedBalance.h:144-148) */
    for (int i = 0; i < np; ++i)
        for (int j = i + 1; j < np; ++j) {
            const rev_t *x = &s->rev[i], *y = &s->rev[j];
            const int ll = x->lhs[0] == y->lhs[0] && x->lhs[1] == y->lhs[1], rr = x->rhs[0] == y->rhs[0] && x->rhs[1] == y->rhs[1];
            const int lr = x->lhs[0] == y->rhs[0] && x->lhs[1] == y->rhs[1], rl = x->rhs[0] == y->lhs[0] && x->rhs[1] == y->lhs[1];
            if ((ll && rr) || (lr && rl))
                return fail("Duplicate found in reversible reactions. Registered reactions might be ambiguous with respect to reversibility.");
        }
    for (int k = 0; k < np; ++k) {
        s->rev_of_rx[s->rev[k].fwd] = k;
        s->rev_of_rx[s->rev[k].bwd] = k;
    }
    s->rev_ready = 1;
    return 0;
}

int or_reversible_configs(or_system *s, size_t cap, double *rows, size_t *n) {
    if (rev_build(s)) return -1;
    for (int k = 0; k < s->n_rev && (size_t)k < cap; ++k) {
        const rev_t *c = &s->rev[k];
        double *o = rows + (size_t)k * OR_REV_ROW;
        const double v[OR_REV_ROW] = {c->kind, c->fwd, c->bwd, c->lhs[0], c->lhs[1], c->rhs[0], c->rhs[1],
                                      c->reaction_radius, c->lhs_radius, c->rhs_radius, c->lhs_volume, c->rhs_volume,
                                      c->eff_lhs_int, c->eff_rhs_int, c->eff_lhs_rx, c->eff_rhs_rx, c->prefactor,
                                      c->total_volume, c->K, c->kf, c->kb};
        memcpy(o, v, sizeof v);
    }
    *n = (size_t)s->n_rev;
    return 0;
}

int or_fission_table(or_system *s, int k, double *radii, double *cdf) {
    if (!s->rev_ready && rev_build(s)) return -1;
    if (k < 0 || k >= s->n_rev || s->rev[k].kind != 0) return fail("not a FusionFission pair");
    memcpy(radii, s->rev[k].radii, sizeof s->rev[k].radii);
    memcpy(cdf, s->rev[k].cdf, sizeof s->rev[k].cdf);
    return OR_FISSION_POINTS + 1;
}

/* pair energy of a particle of type t at pos against every other particle inside the box (skipping
 * entries x0 and x1), each potential gated by d^2 < cutoff^2 */
static double db_env_energy(const or_system *s, const double *pos, int t, size_t x0, size_t x1) {
    if (!in_box(s, pos)) return 0.;
    double sum = 0.;
    for (size_t i = 0; i < s->n; ++i) {
        const entry_t *o = &s->e[i];
        if (o->deactivated || i == x0 || i == x1 || !in_box(s, o->pos)) continue;
        double x[3], f[3] = {0., 0., 0.};
        or_shortest_difference(pos, o->pos, s->box, s->pbc, x);
        const double d2 = dot3(x, x);
        for (int k = 0; k < s->n_pots; ++k) {
            const pot_t *pt = &s->pots[k];
            if (!((pt->t1 == t && pt->t2 == o->type) || (pt->t1 == o->type && pt->t2 == t))) continue;
            if (d2 < pt->cutoff2) sum += or_pair_force_energy(pt->kind, pt->p, x, f);
        }
    }
    return sum;
}

/* the gated pair energy of two positions of types a and b (both inside the box) */
static double db_pair_energy(const or_system *s, const double *p, int a, const double *q, int b) {
    if (!in_box(s, p) || !in_box(s, q)) return 0.;
    double x[3], f[3] = {0., 0., 0.}, u = 0.;
    or_shortest_difference(p, q, s->box, s->pbc, x);
    const double d2 = dot3(x, x);
    for (int k = 0; k < s->n_pots; ++k) {
        const pot_t *pt = &s->pots[k];
        if (!((pt->t1 == a && pt->t2 == b) || (pt->t1 == b && pt->t2 == a))) continue;
        if (d2 < pt->cutoff2) u += or_pair_force_energy(pt->kind, pt->p, x, f);
    }
    return u;
}

/* gated U at distance r between types a and b (the mutual interaction dU_int of a FusionFission move) */
static double db_u_at(const or_system *s, int a, int b, double r) {
    double x[3] = {r, 0., 0.}, f[3] = {0., 0., 0.}, u = 0.;
    for (int k = 0; k < s->n_pots; ++k) {
        const pot_t *pt = &s->pots[k];
        if (!((pt->t1 == a && pt->t2 == b) || (pt->t1 == b && pt->t2 == a))) continue;
        if (r * r < pt->cutoff2) u += or_pair_force_energy(pt->kind, pt->p, x, f);
    }
    return u;
}

static or_reaction_record *db_record(or_system *s) {
    if (s->n_recs == s->cap_recs) {
        s->cap_recs = s->cap_recs ? 2 * s->cap_recs : 64;
        s->recs = (or_reaction_record *)realloc(s->recs, s->cap_recs * sizeof(or_reaction_record));
    }
    or_reaction_record *rec = &s->recs[s->n_recs++];
    memset(rec, 0, sizeof *rec);
    return rec;
}

/* one evaluated event; returns 1 when it was performed (accepted) */
static int db_event(or_system *s, const event_t *ev, uint32_t step) {
    const int k = s->rev_of_rx[ev->rx];
    if (k < 0) { /* not reversible: performed directly (SCPUReactionImpls.cpp:525-548) */
        update_t up;
        memset(&up, 0, sizeof up);
        perform_reaction(s, ev, &up, step);
        apply_update(s, &up);
        return 1;
    }
    const rx_t *r = &s->rx[ev->rx];
    const rev_t *c = &s->rev[k];
    entry_t *e1 = &s->e[ev->idx1], *e2 = &s->e[ev->idx2];
    size_t educt = ev->idx1; /* the particle that converts (EnzymaticEnzymatic: not the catalyst) */
    double dE = 0., du_int = 0., fdist = 0., n3[3] = {0., 0., 0.};
    double q1[3], q2[3];
    int n_fwd = 1, n_back = 1;
    switch (r->kind) {
    case OR_RX_FISSION: {
        uint32_t x[4];
        philox_draw(s->seed, (uint32_t)e1->id, step, 3, 0, x);
        normals3(x, n3);
        const double nn = sqrt(dot3(n3, n3));
        for (int d = 0; d < 3; ++d) n3[d] /= nn;
        philox_draw(s->seed, (uint32_t)e1->id, step, 10, 0, x);
        const double u = or_u53(x[0], x[1]);
        int lo = 0, hi = OR_FISSION_POINTS + 1; /* lower_bound, DetailedBalance.h:133-138 */
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (c->cdf[mid] < u) lo = mid + 1;
            else hi = mid;
        }
        if (lo > OR_FISSION_POINTS) lo = OR_FISSION_POINTS;
        fdist = c->radii[lo];
        du_int = db_u_at(s, r->products[0], r->products[1], fdist);
        for (int d = 0; d < 3; ++d) {
            q1[d] = e1->pos[d] - r->w2 * fdist * n3[d]; /* products[1] */
            q2[d] = e1->pos[d] + r->w1 * fdist * n3[d]; /* products[0] */
        }
        or_fix_position(q1, s->box, s->pbc);
        or_fix_position(q2, s->box, s->pbc);
        dE = db_env_energy(s, q1, r->products[1], ev->idx1, ev->idx1) + db_env_energy(s, q2, r->products[0], ev->idx1, ev->idx1) +
             db_pair_energy(s, q1, r->products[1], q2, r->products[0]) - db_env_energy(s, e1->pos, e1->type, ev->idx1, ev->idx1);
        n_fwd = 2;
        break;
    }
    case OR_RX_FUSION: {
        double diff[3];
        or_shortest_difference(e1->pos, e2->pos, s->box, s->pbc, diff);
        du_int = -db_u_at(s, r->educts[0], r->educts[1], sqrt(dot3(diff, diff)));
        const double w = (r->educts[0] == e1->type) ? r->w1 : r->w2;
        for (int d = 0; d < 3; ++d) q1[d] = e1->pos[d] + w * diff[d];
        or_fix_position(q1, s->box, s->pbc);
        dE = db_env_energy(s, q1, r->products[0], ev->idx1, ev->idx2) - db_env_energy(s, e1->pos, e1->type, ev->idx1, ev->idx2) -
             db_env_energy(s, e2->pos, e2->type, ev->idx1, ev->idx2) - db_pair_energy(s, e1->pos, e1->type, e2->pos, e2->type);
        n_back = 2;
        break;
    }
    case OR_RX_ENZYMATIC:
        educt = (e1->type == r->educts[1]) ? ev->idx2 : ev->idx1;
        /* fall through */
    default: {
        const entry_t *ed = &s->e[educt];
        dE = db_env_energy(s, ed->pos, r->products[0], educt, educt) - db_env_energy(s, ed->pos, ed->type, educt, educt);
        break;
    }
    }
    const double boltzmann = exp(-1. / s->kbt * (dE - du_int));
    const double pref = (c->kind == 2 && r->educts[0] == c->lhs[0] && r->educts[1] == c->lhs[1]) ? c->prefactor : 1.;
    double acceptance = pref * boltzmann;
    if (!(acceptance < 1.)) acceptance = 1.; /* std::min(1., x) */
    const uint64_t lo = e1->id < e2->id ? e1->id : e2->id, hi = e1->id < e2->id ? e2->id : e1->id;
    uint32_t x[4];
    philox_draw(s->seed, (uint32_t)lo, step, 11u | ((uint32_t)ev->rx << 8), (uint32_t)hi, x);
    const int accept = or_u53(x[0], x[1]) < acceptance;
    const uint64_t base = s->next_id;
    s->next_id += (uint64_t)(n_fwd + n_back); /* products, then the backup particles */
    if (!accept) {
        if (r->kind == OR_RX_FUSION) {
            e1->id = base + 1;
            e2->id = base + 2;
        } else {
            s->e[educt].id = base + (uint64_t)n_fwd;
        }
        return 0;
    }
    or_reaction_record *rec = s->record ? db_record(s) : NULL;
    if (rec) {
        rec->reaction = (uint32_t)ev->rx;
        rec->type = r->kind;
    }
    switch (r->kind) {
    case OR_RX_FISSION: {
        if (rec) {
            for (int d = 0; d < 3; ++d) rec->where[d] = e1->pos[d];
            or_fix_position(rec->where, s->box, s->pbc);
            rec->educts[0] = rec->educts[1] = e1->id;
            rec->types_from[0] = rec->types_from[1] = (uint32_t)e1->type;
        }
        update_t up;
        memset(&up, 0, sizeof up);
        entry_t p1, p2;
        memset(&p1, 0, sizeof p1);
        memset(&p2, 0, sizeof p2);
        memcpy(p1.pos, q1, sizeof q1);
        p1.type = r->products[1];
        p1.id = base;
        memcpy(p2.pos, q2, sizeof q2);
        p2.type = r->products[0];
        p2.id = base + 1;
        upd_new(&up, &p1);
        upd_new(&up, &p2);
        upd_removed(&up, ev->idx1);
        apply_update(s, &up);
        break;
    }
    case OR_RX_FUSION: {
        if (rec) {
            for (int d = 0; d < 3; ++d) rec->where[d] = (e1->pos[d] + e2->pos[d]) / 2.;
            or_fix_position(rec->where, s->box, s->pbc);
            rec->educts[0] = e1->id;
            rec->educts[1] = e2->id;
            rec->types_from[0] = (uint32_t)e1->type;
            rec->types_from[1] = (uint32_t)e2->type;
        }
        update_t up;
        memset(&up, 0, sizeof up);
        entry_t p;
        memset(&p, 0, sizeof p);
        memcpy(p.pos, q1, sizeof q1);
        p.type = r->products[0];
        p.id = base;
        upd_new(&up, &p);
        upd_removed(&up, ev->idx1);
        upd_removed(&up, ev->idx2);
        apply_update(s, &up);
        break;
    }
    default: {
        entry_t *ed = &s->e[educt];
        if (rec) {
            const entry_t *other = (educt == ev->idx1) ? e2 : e1; /* the catalyst, or the educt itself */
            if (r->kind == OR_RX_CONVERSION) other = ed;
            for (int d = 0; d < 3; ++d) rec->where[d] = (ed->pos[d] + other->pos[d]) / 2.;
            or_fix_position(rec->where, s->box, s->pbc);
            rec->educts[0] = ed->id;
            rec->educts[1] = other->id;
            rec->types_from[0] = (uint32_t)ed->type;
            rec->types_from[1] = (uint32_t)other->type;
        }
        ed->type = r->products[0];
        ed->id = base;
        break;
    }
    }
    s->counts[ev->rx]++;
    return 1;
}

int or_detailed_balance(or_system *s, double dt, uint32_t step, uint64_t *n_performed) {
    memset(s->counts, 0, sizeof s->counts);
    s->n_recs = 0;
    if (n_performed) *n_performed = 0;
    if (!s->rev_ready && rev_build(s)) return -1;
    ev_ud u;
    memset(&u, 0, sizeof u);
    u.dt = dt;
    u.step = step;
    u.draw = 1;
    gather(s, &u);
    qsort(u.ev, u.n, sizeof(event_t), cmp_gkey);
    uint64_t done = 0;
    for (size_t i = 0; i < u.n; ++i) {
        event_t *ev = &u.ev[i];
        if (ev->blocked) continue;
        done += (uint64_t)db_event(s, ev, step);
        for (size_t j = i + 1; j < u.n; ++j)
            if (shares(&u.ev[j], ev)) u.ev[j].blocked = 1;
    }
    free(u.ev);
    if (n_performed) *n_performed = done;
    return 0;
}

void or_set_record_reactions(or_system *s, int on) {
    s->record = on != 0;
    s->n_recs = 0;
}

const or_reaction_record *or_reaction_records(const or_system *s, size_t *n) {
    if (n) *n = s->n_recs;
    return s->recs;
}

const uint64_t *or_reaction_counts(const or_system *s, int *n) {
    if (n) *n = s->n_rx;
    return s->counts;
}

/* SimulationLoop::run body: include/readdy/api/SimulationLoop.h:256-306 */
static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int or_loop_step(or_system *s, double dt, int mode, uint32_t step, const double *noise_by_id, size_t noise_len,
                 double *energy) {
    /* phases timed like the reference's readdy::util::Timer scopes (integrator, neighborList, reactions,
     * forces; SimulationLoop.h:280-300) */
    double t0 = now_s();
    if (or_euler_bd(s, dt, noise_by_id, noise_len, step)) return -1;
    double t1 = now_s();
    if (or_nl_fill(s)) return -1;
    double t2 = now_s();
    if (s->n_rx > 0 && or_reactions(s, mode, dt, step, NULL)) return -1;
    double t3 = now_s();
    if (or_nl_fill(s)) return -1;
    double t4 = now_s();
    int rc = or_calculate_forces(s, energy, NULL);
    double t5 = now_s();
    s->phase_s[0] += t1 - t0;
    s->phase_s[1] += (t2 - t1) + (t4 - t3);
    s->phase_s[2] += t3 - t2;
    s->phase_s[3] += t5 - t4;
    return rc;
}

void or_phase_times(or_system *s, double *out4, int reset) {
    for (int k = 0; k < 4; ++k) {
        out4[k] = s->phase_s[k];
        if (reset) s->phase_s[k] = 0.0;
    }
}

/* ------------------------------------------------------------------------------------------------
 * Spatial topology reactions of the topology-particle enzymatic kind
 * (kernels/cpu/src/actions/CPUEvaluateTopologyReactions.cpp:88-475, SpatialTopologyReaction.h:64-236).
 *   gatherEvents (:200-367): every particle of a topology, with every neighbour of the cell list that
 *     belongs to no topology, for each reaction registered under (its type, its topology's type, the
 *     neighbour's type) whose radius exceeds the distance (distSquared < radius^2), is a candidate of
 *     rate `rate`; a reaction is found from the topology particle's side only, as the registry keys it.
 *   performEvents (algorithm.h:62-136) with shouldEval = 1 - exp(-rate dt) and eventsDependent
 *     (:455-464): events of one topology depend on each other, and so do topology-particle events that
 *     share the free particle.  As for the Gillespie handler, the selection is realised by
 *     Efraimidis-Spirakis keys -ln(u) / rate over the candidates whose Bernoulli draw succeeds (Philox
 *     counter (topology particle id, step, 6 | r << 8, free particle id)), greedily in key order.
 *   handleTopologyParticleReaction (:369-404): the topology particle takes type_to1, the free particle
 *     type_to2, the topology the type top_type_to1.  (The reference also writes the topology's index into
 *     the free particle's record, which only matters for later topology-topology reactions; free
 *     particles here keep no topology.)  A fusion (TP_FUSION) also appends the free particle to the
 *     topology, bonded to the topology particle (GraphTopology::appendParticle): it becomes a member of
 *     the topology under the next id of the reserve (or_set_topology_reserve), taken in the order the
 *     events are performed, and the call lists (topology, topology particle id, old id, new id).  The
 *     reference keeps the particle's id; the MI355X kernel renames it into its topology handle range.
 * --------------------------------------------------------------------------------------------- */
int or_add_spatial_reaction(or_system *s, int type1, int top_type1, int type2, int type_to1, int top_type_to1,
                            int type_to2, double rate, double radius) {
    if (s->n_str >= OR_MAX_STR) return fail("too many spatial topology reactions");
    if (type1 < 0 || type1 >= s->n_types || type2 < 0 || type2 >= s->n_types || type_to1 < 0 ||
        type_to1 >= s->n_types || type_to2 < 0 || type_to2 >= s->n_types)
        return fail("spatial topology reaction species out of range");
    if (!(rate > 0)) return fail("The rate of an structural topology reaction should always be positive");
    str_t *r = &s->str[s->n_str++];
    r->type1 = type1;
    r->top_type1 = top_type1;
    r->type2 = type2;
    r->type_to1 = type_to1;
    r->top_type_to1 = top_type_to1;
    r->type_to2 = type_to2;
    r->rate = rate;
    r->radius = radius;
    r->radius2 = radius * radius;
    r->fusion = 0;
    r->tt = 0;
    r->self_fusion = 0;
    r->top_type2 = r->top_type_to2 = -1;
    return 0;
}

int or_add_spatial_fusion(or_system *s, int type1, int top_type1, int type2, int type_to1, int top_type_to1,
                          int type_to2, double rate, double radius) {
    if (or_add_spatial_reaction(s, type1, top_type1, type2, type_to1, top_type_to1, type_to2, rate, radius)) return -1;
    s->str[s->n_str - 1].fusion = 1;
    return 0;
}

int or_add_spatial_tt(or_system *s, int type1, int top_type1, int type2, int top_type2, int type_to1, int top_type_to1,
                      int type_to2, int top_type_to2, double rate, double radius) {
    if (or_add_spatial_reaction(s, type1, top_type1, type2, type_to1, top_type_to1, type_to2, rate, radius)) return -1;
    str_t *r = &s->str[s->n_str - 1];
    r->tt = 1;
    r->top_type2 = top_type2;
    r->top_type_to2 = top_type_to2;
    return 0;
}

int or_add_spatial_tt_fusion(or_system *s, int type1, int top_type1, int type2, int top_type2, int type_to1,
                             int type_to2, int top_type_to, double rate, double radius) {
    if (or_add_spatial_tt(s, type1, top_type1, type2, top_type2, type_to1, top_type_to, type_to2, top_type_to, rate,
                          radius))
        return -1;
    s->str[s->n_str - 1].fusion = 1;
    return 0;
}

int or_allow_self_fusion(or_system *s) {
    if (!s->n_str || !s->str[s->n_str - 1].tt || !s->str[s->n_str - 1].fusion)
        return fail("self fusion applies to a topology-topology fusion");
    s->str[s->n_str - 1].self_fusion = 1;
    return 0;
}

const uint64_t *or_topology_merges(const or_system *s, size_t *n) {
    if (n) *n = s->n_merges;
    return s->merges;
}

int or_set_topology_reserve(or_system *s, uint64_t next_id, uint64_t end_id) {
    if (end_id < next_id) return fail("topology reserve: end before start");
    s->fuse_next = next_id;
    s->fuse_end = end_id;
    return 0;
}

const uint64_t *or_topology_fusions(const or_system *s, size_t *n) {
    if (n) *n = s->n_fusions;
    return s->fusions;
}

int or_set_topologies(or_system *s, size_t n_ids, const int *topology_of_id, size_t n_tops, const int *top_types) {
    free(s->topology_of);
    free(s->top_types);
    s->topology_of = (int *)malloc((n_ids ? n_ids : 1) * sizeof(int));
    s->top_types = (int *)malloc((n_tops ? n_tops : 1) * sizeof(int));
    for (size_t i = 0; i < n_ids; ++i) {
        if (topology_of_id[i] >= (int)n_tops) return fail("particle assigned to a topology that does not exist");
        s->topology_of[i] = topology_of_id[i];
    }
    for (size_t k = 0; k < n_tops; ++k) s->top_types[k] = top_types[k];
    s->topology_of_len = n_ids;
    s->n_tops = n_tops;
    return 0;
}

const int *or_topology_types(const or_system *s, size_t *n) {
    if (n) *n = s->n_tops;
    return s->top_types;
}

const uint64_t *or_topology_reaction_counts(const or_system *s, int *n) {
    if (n) *n = s->n_str;
    return s->str_counts;
}

static int topology_of_entry(const or_system *s, const entry_t *e) {
    return e->id < s->topology_of_len ? s->topology_of[e->id] : -1;
}

typedef struct {
    size_t p, q; /* slots: topology particle, free particle (topology-topology: the pair, lower id first) */
    int top, r;
    int top2, flip; /* topology-topology: the partner's topology; the reaction matched with its sides swapped */
    double gkey;
    uint64_t ip, iq;
} tevent_t;

typedef struct {
    tevent_t *ev;
    size_t n, cap;
    double dt;
    uint32_t step;
} tev_ud;

/* whether a bond joins the particles of slots p and q (graph distance 1) */
static int slots_bonded(const or_system *s, size_t p, size_t q) {
    for (size_t b = 0; b < s->n_bonds; ++b) {
        const size_t a = s->bonds[b].i[0], c = s->bonds[b].i[1];
        if ((a == p && c == q) || (a == q && c == p)) return 1;
    }
    return 0;
}

static void str_visit(or_system *s, size_t p, size_t q, void *ud) {
    tev_ud *u = (tev_ud *)ud;
    const entry_t *ep = &s->e[p], *eq = &s->e[q];
    if (eq->deactivated) return;
    const int top = topology_of_entry(s, ep), top2 = topology_of_entry(s, eq);
    if (top < 0) return;
    /* a partner in a topology: each pair once, from the lower id (CPUEvaluateTopologyReactions.cpp:244-268); a
     * pair within one topology reacts only by self-fusion, and only when the two are not bonded: graph distance
     * above 1, the default minimum distance (:303-318) */
    if (top2 >= 0 && !(ep->id < eq->id)) return;
    const int same = top2 >= 0 && top2 == top;
    int bonded = -1;
    double d2 = -1.0;
    for (int r = 0; r < s->n_str; ++r) {
        const str_t *x = &s->str[r];
        int flip = 0;
        if (top2 < 0) {
            if (x->tt || x->type1 != ep->type || x->top_type1 != s->top_types[top] || x->type2 != eq->type) continue;
        } else { /* the registry keys a topology-topology reaction under both orders of its sides */
            if (!x->tt) continue;
            if (same) {
                if (!x->self_fusion) continue;
                if (bonded < 0) bonded = slots_bonded(s, p, q);
                if (bonded) continue;
            }
            const int T = s->top_types[top], T2 = s->top_types[top2];
            if (x->type1 == ep->type && x->top_type1 == T && x->type2 == eq->type && x->top_type2 == T2) flip = 0;
            else if (x->type2 == ep->type && x->top_type2 == T && x->type1 == eq->type && x->top_type1 == T2) flip = 1;
            else continue;
        }
        if (d2 < 0) d2 = or_dist2(ep->pos, eq->pos, s->box, s->pbc);
        if (!(d2 < x->radius2)) continue;
        uint32_t w[4];
        philox_draw(s->seed, (uint32_t)ep->id, u->step, 6u | ((uint32_t)r << 8), (uint32_t)eq->id, w);
        if (!(or_u53(w[0], w[1]) < accept_p(x->rate, u->dt))) continue;
        if (u->n == u->cap) {
            u->cap = u->cap ? 2 * u->cap : 256;
            u->ev = (tevent_t *)realloc(u->ev, u->cap * sizeof(tevent_t));
        }
        tevent_t *e = &u->ev[u->n++];
        e->p = p;
        e->q = q;
        e->top = top;
        e->r = r;
        e->top2 = top2;
        e->flip = flip;
        e->gkey = -log(or_u53(w[2], w[3]) + 0x1p-54) / x->rate;
        e->ip = ep->id;
        e->iq = eq->id;
    }
}

static int cmp_tev(const void *a, const void *b) {
    const tevent_t *x = (const tevent_t *)a, *y = (const tevent_t *)b;
    if (x->gkey != y->gkey) return x->gkey < y->gkey ? -1 : 1;
    if (x->ip != y->ip) return x->ip < y->ip ? -1 : 1;
    if (x->iq != y->iq) return x->iq < y->iq ? -1 : 1;
    return x->r - y->r;
}

int or_topology_reactions(or_system *s, double dt, uint32_t step, uint64_t *n_performed) {
    memset(s->str_counts, 0, sizeof s->str_counts); /* CPUEvaluateTopologyReactions.cpp:92-94 */
    if (n_performed) *n_performed = 0;
    s->n_fusions = 0;
    s->n_merges = 0;
    if (!s->n_str || !s->n_tops) return 0;
    if (!s->nl_set) return fail("topology reactions need a neighbour list");
    tev_ud u;
    memset(&u, 0, sizeof u);
    u.dt = dt;
    u.step = step;
    for (size_t c = 0; c < s->n_cells; ++c)
        for (size_t p1 = s->head[c]; p1 != 0; p1 = s->list[p1]) {
            const entry_t *ep = &s->e[p1 - 1];
            if (ep->deactivated || topology_of_entry(s, ep) < 0) continue;
            for_each_neighbor(s, p1 - 1, c, str_visit, &u);
        }
    qsort(u.ev, u.n, sizeof(tevent_t), cmp_tev);
    unsigned char *top_used = (unsigned char *)calloc(s->n_tops, 1);
    unsigned char *p_used = (unsigned char *)calloc(s->n ? s->n : 1, 1);
    uint64_t done = 0;
    for (size_t i = 0; i < u.n; ++i) {
        const tevent_t *e = &u.ev[i];
        const str_t *x = &s->str[e->r];
        if (x->tt) { /* handleTopologyTopologyReaction (:406-475), enzymatic: both topologies used */
            if (top_used[e->top] || top_used[e->top2]) continue;
            top_used[e->top] = top_used[e->top2] = 1;
            s->e[e->p].type = e->flip ? x->type_to2 : x->type_to1;
            s->e[e->q].type = e->flip ? x->type_to1 : x->type_to2;
            if (x->fusion) { /* appendTopology into the first side's topology, of type top_type_to1 (:433-437) */
                s->top_types[e->top] = x->top_type_to1;
                for (size_t k = 0; k < s->topology_of_len; ++k)
                    if (s->topology_of[k] == e->top2) s->topology_of[k] = e->top;
                s->merges = (uint64_t *)realloc(s->merges, 4 * (s->n_merges + 1) * sizeof(uint64_t));
                uint64_t *row = s->merges + 4 * s->n_merges++;
                row[0] = (uint64_t)e->top;
                row[1] = (uint64_t)e->top2;
                row[2] = e->ip;
                row[3] = e->iq;
            } else {
                s->top_types[e->top] = e->flip ? x->top_type_to2 : x->top_type_to1;
                s->top_types[e->top2] = e->flip ? x->top_type_to1 : x->top_type_to2;
            }
            s->str_counts[e->r]++;
            ++done;
            continue;
        }
        if (top_used[e->top] || p_used[e->q]) continue;
        if (x->fusion && s->fuse_next >= s->fuse_end) {
            free(top_used);
            free(p_used);
            free(u.ev);
            return fail("topology-particle fusion beyond the topology reserve");
        }
        top_used[e->top] = 1;
        p_used[e->q] = 1;
        s->e[e->p].type = x->type_to1;
        s->e[e->q].type = x->type_to2;
        if (x->fusion) { /* appendParticle(idx2, idx1): the free particle joins under the next reserve id */
            const uint64_t nid = s->fuse_next++;
            if (nid >= s->topology_of_len) {
                int *t = (int *)realloc(s->topology_of, (size_t)(nid + 1) * sizeof(int));
                if (!t) return fail("out of memory");
                for (size_t k = s->topology_of_len; k <= nid; ++k) t[k] = -1;
                s->topology_of = t;
                s->topology_of_len = (size_t)nid + 1;
            }
            s->topology_of[nid] = e->top;
            s->fusions = (uint64_t *)realloc(s->fusions, 4 * (s->n_fusions + 1) * sizeof(uint64_t));
            uint64_t *row = s->fusions + 4 * s->n_fusions++;
            row[0] = (uint64_t)e->top;
            row[1] = e->ip;
            row[2] = e->iq;
            row[3] = nid;
            s->e[e->q].id = nid;
        }
        s->top_types[e->top] = x->top_type_to1;
        s->str_counts[e->r]++;
        ++done;
    }
    free(top_used);
    free(p_used);
    free(u.ev);
    if (n_performed) *n_performed = done;
    return 0;
}
