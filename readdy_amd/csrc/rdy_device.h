// Device-side building blocks of the MI355X ReaDDy engine: particle record, Philox4x32-10,
// periodic-boundary arithmetic, pair potentials and bonded terms.
//
// Every translation unit that includes this header is compiled with -ffp-contract=off:
// the FP64 pair gate d2 < rc^2 and the integrator follow the reference's operation order
// without fused multiply-add, so pair sets and candidate sets are bit-identical to an FP64
// CPU evaluation of the same inputs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rdy {

// ---- particle record: 32 bytes, the unit of the cell bins -------------------------------------
// tw packs the species (bits 0-9), flags (bits 10-13) and the replica index (bits 14-31).
struct alignas(16) State {
    double x, y, z;
    uint32_t tw;
    uint32_t h;  // replica-local particle handle (the ParticleId of this engine)
};
static_assert(sizeof(State) == 32, "State must stay 32 bytes");

constexpr uint32_t TW_TYPE_MASK = 0x3FFu;
constexpr uint32_t TW_OUTSIDE = 1u << 10;  // outside the box on a non-periodic axis: never binned as a neighbour
constexpr uint32_t TW_DEAD = 1u << 11;     // removed by a reaction in this step
constexpr uint32_t TW_TOPO = 1u << 12;     // member of a topology (has bonded terms)
constexpr uint32_t TW_NEW = 1u << 13;      // inserted or retyped by a reaction of this step (fused step only)
constexpr uint32_t TW_REPLICA_SHIFT = 14;
constexpr uint32_t MAX_TYPES = 1024;
constexpr uint32_t MAX_REPLICAS = 1u << 18;

__host__ __device__ inline uint32_t tw_type(uint32_t tw) { return tw & TW_TYPE_MASK; }
__host__ __device__ inline uint32_t tw_replica(uint32_t tw) { return tw >> TW_REPLICA_SHIFT; }
__host__ __device__ inline uint32_t make_tw(uint32_t type, uint32_t flags, uint32_t replica) {
    return type | flags | (replica << TW_REPLICA_SHIFT);
}

// ---- per-replica record (device resident) -----------------------------------------------------
struct ReplicaDev {
    double box[3];
    double cs[3];        // cell size = box / dims (CellLinkedList.cpp:64-70)
    int32_t pbc[3];
    int32_t dims[3];
    uint32_t cell_base;  // first global cell of this replica
    uint32_t n_cells;
    uint32_t ev_base;    // first event of this replica's event segment
    uint32_t ev_cap;
    uint32_t handle_next;  // next fresh particle handle (Particle::nextId)
    uint32_t step;         // Philox step counter (advanced before each integration)
    uint32_t active;       // steps in the current launch
    uint32_t in_use;       // replica slot holds a system
    uint64_t seed;
    uint64_t steps_left;   // remaining step budget (serving)
    uint64_t steps_done;
    uint32_t topo_base, n_topo;
    uint32_t tt_base, n_tt;  // this replica's topologies: their types at tt_base .. tt_base + n_tt
    int32_t handler;       // reaction handler of this replica (-1: the one the step is launched with)
    uint32_t integ;        // fused step: the walk of this step also integrates the next one
    unsigned long long energy;  // two's-complement int64, fixed point 2^-24
    uint32_t topo_range;   // topology handles of the replica: n_topo in use, the rest reserved for fusions
    uint32_t topo_pad;
};

// ---- potentials and reaction tables -----------------------------------------------------------
enum : int32_t { POT_HR = 0, POT_WEAK = 1, POT_LJ = 2, POT_SE = 3 };
enum : int32_t { RX_CONVERSION = 0, RX_FUSION = 1, RX_FISSION = 2, RX_ENZYMATIC = 3, RX_DECAY = 4 };

struct PotDev {
    double cut2;   // FP64 gate (reference: distSquared < getCutoffRadiusSquared, CPUCalculateForces.cpp:243)
    int32_t kind;
    int32_t ia, ib;  // integer exponents (LJ m, n; SE exponent)
    float p[7];
};

struct PairIdx {   // potentials of an ordered type pair: pots[start .. start + count)
    double maxcut2;    // largest cutoff^2 of the list (exact FP64 gate)
    float maxcut2_f;   // the same in FP32 (prefilter threshold before the rounding margin)
    uint16_t start, count;
};

struct RxDev {
    double rate, p_accept, educt_dist2, product_distance, w1, w2;
    int32_t kind, order;
    uint32_t educts[2], products[2];
};

struct RxPairIdx {  // order-2 reactions of an ordered type pair: rxlist[start .. start + count)
    float maxr2_f;     // largest educt distance^2 of the list (FP32 prefilter threshold before margin)
    uint16_t start, count;
};

// ---- Philox4x32-10, from Salmon, Moraes, Dror & Shaw, SC'11 ------------------------------------
__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint64_t seed,
                                       uint32_t out[4]) {
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // the 64-bit products give both halves from one multiply (v_mad_u64_u32)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// 53-bit uniform in [0, 1) from two words: ((a >> 5) 2^26 + (b >> 6)) / 2^53 (exact in FP64)
__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// three N(0,1) draws by Box-Muller from one Philox block (v_k = (x_k + 1/2) / 2^32)
__device__ __forceinline__ void normals3(const uint32_t x[4], double n[3]) {
    double v0 = ((double)x[0] + 0.5) * (1.0 / 4294967296.0);
    double v1 = ((double)x[1] + 0.5) * (1.0 / 4294967296.0);
    double v2 = ((double)x[2] + 0.5) * (1.0 / 4294967296.0);
    double v3 = ((double)x[3] + 0.5) * (1.0 / 4294967296.0);
    double r0 = sqrt(-2.0 * log(v0));
    double r1 = sqrt(-2.0 * log(v2));
    // sin/cos(2 pi v) as sinpi/cospi(2 v): the same values up to the last bit, without the
    // large-argument reduction path (v < 1)
    double s1, c1;
    sincospi(2.0 * v1, &s1, &c1);
    const double c3 = cospi(2.0 * v3);
    n[0] = r0 * c1;
    n[1] = r0 * s1;
    n[2] = r1 * c3;
}

// three N(0,1) draws for the Brownian-dynamics noise by Box-Muller in FP32 from one Philox block:
// 24-bit uniforms v_k = ((x_k >> 8) + 1/2) 2^-24 (exact in FP32), r = sqrt(-2 ln v), angle 2 pi v.
// The reference draws its noise from std::normal_distribution (RandomProvider.h:64-69); any exact
// N(0,1) sampler is equivalent, and an FP32 draw moves a particle by sqrt(2 D dt) times a few 2^-24,
// far below every other error of the step, for a fifth of the instructions of the FP64 transform.
__device__ __forceinline__ void bd_normals3(const uint32_t x[4], double n[3]) {
    const float s24 = 5.9604644775390625e-08f;  // 2^-24
    const float v0 = ((float)(x[0] >> 8) + 0.5f) * s24;
    const float v1 = ((float)(x[1] >> 8) + 0.5f) * s24;
    const float v2 = ((float)(x[2] >> 8) + 0.5f) * s24;
    const float v3 = ((float)(x[3] >> 8) + 0.5f) * s24;
    const float r0 = sqrtf(-2.0f * logf(v0));
    const float r1 = sqrtf(-2.0f * logf(v2));
    float s1, c1;
    sincospif(2.0f * v1, &s1, &c1);
    const float c3 = cospif(2.0f * v3);
    n[0] = (double)(r0 * c1);
    n[1] = (double)(r0 * s1);
    n[2] = (double)(r1 * c3);
}

// ---- periodic boundary arithmetic (reference semantics, FP64, no FMA) --------------------------
// shortest difference rhs - lhs: "> L/2" subtracts, "<= -L/2" adds (boundary_condition_operations.h:70-80)
__device__ __forceinline__ double min_image(double dv, double L, int32_t periodic) {
    if (periodic) {
        if (dv > L * .5) dv -= L;
        else if (dv <= -L * .5) dv += L;
    }
    return dv;
}

// while-loop wrap (boundary_condition_operations.h:53-61); non-finite values are left as they are,
// and a coordinate more than 1024 box lengths out is first reduced with fmod so the loops stay short
// fmod(v, L) for finite v and L > 0, by binary long division: every subtraction a - m has
// m <= a < 2m, so it is exact (Sterbenz) and the result equals the exactly rounded fmod
__device__ __forceinline__ double fmod_exact(double v, double L) {
    double a = fabs(v);
    if (a >= L) {
        int e = ilogb(a) - ilogb(L);
        double m = ldexp(L, e);
        for (; e >= 0; --e, m *= 0.5)
            if (a >= m) a -= m;
    }
    return copysign(a, v);
}

__device__ __forceinline__ double wrap(double v, double L, int32_t periodic) {
    if (periodic && isfinite(v)) {
        if (!(fabs(v) < 1024.0 * L)) v = fmod_exact(v, L);
        while (v >= .5 * L) v -= L;
        while (v < -.5 * L) v += L;
    }
    return v;
}

__device__ __forceinline__ double dist2(double dx, double dy, double dz) {
    double s = dx * dx;
    s = s + dy * dy;
    s = s + dz * dz;
    return s;
}

// cell coordinate floor((x + L/2) / cs) (CellLinkedList.cpp:181-185), clamped into the grid
__device__ __forceinline__ int32_t cell_coord(double x, double L, double cs, int32_t dims) {
    double q = floor((x + .5 * L) / cs);
    int32_t c = !(q >= 0.0) ? 0 : (q >= (double)dims ? dims - 1 : (int32_t)q);  // NaN -> 0
    return c;
}

__device__ __forceinline__ bool in_box_axis(double x, double L) { return -.5 * L <= x && .5 * L > x; }

// ---- pair potentials in FP32 (x = neighbour - self); the caller has passed the FP64 gate -----
__device__ __forceinline__ float ipowf(float b, int32_t e) {
    float r = 1.0f;
    while (e > 0) {
        if (e & 1) r *= b;
        b *= b;
        e >>= 1;
    }
    return r;
}

// r = |x| and inv_r = 1/r of a pair (0 and 0 for coincident particles) from the hardware reciprocal
// square root (1 ulp): every potential of the pair shares them
struct PairDist {
    float r, inv_r;
};
__device__ __forceinline__ PairDist pair_dist(float d2) {
    const float inv_r = d2 > 0.0f ? __builtin_amdgcn_rsqf(d2) : 0.0f;
    return PairDist{d2 * inv_r, inv_r};
}

// adds the force on "self" to (fx, fy, fz) and returns the pair energy (PotentialsOrder2.h:54-323)
__device__ __forceinline__ float pair_force_energy(const PotDev &pt, float d2, const PairDist &pd, float x, float y,
                                                   float z, float &fx, float &fy, float &fz) {
    switch (pt.kind) {
    case POT_HR: {  // E = k/2 (r - d)^2, F = k (r - d)/r x  (PotentialsOrder2.h:72-91)
        const float k = pt.p[0], d = pt.p[1];
        const float r = pd.r;
        float dr = r - d;
        if (d2 > 0.0f) {
            float fac = k * dr * pd.inv_r;
            fx += fac * x;
            fy += fac * y;
            fz += fac * z;
        }
        return 0.5f * k * dr * dr;
    }
    case POT_WEAK: {  // three harmonic pieces (PotentialsOrder2.h:131-183)
        const float k = pt.p[0], d0 = pt.p[1], h = pt.p[2], rc = pt.p[3], c = pt.p[4];  // c = (2/(rc-d0))^2
        const float r = pd.r;
        float e, fac;
        if (r < d0) {
            e = 0.5f * k * (r - d0) * (r - d0) - h;
            fac = -k * (d0 - r);
        } else if (r < pt.p[5]) {  // p[5] = d0 + (rc - d0)/2
            e = 0.5f * h * c * (r - d0) * (r - d0) - h;
            fac = -h * c * (d0 - r);
        } else {
            e = -0.5f * h * c * (r - rc) * (r - rc);
            fac = h * c * (rc - r);
        }
        if (r > 0.0f && fac != 0.0f) {
            float q = fac * pd.inv_r;
            fx += q * x;
            fy += q * y;
            fz += q * z;
        }
        return e;
    }
    case POT_LJ: {  // E = k[(s/r)^m - (s/r)^n] (- shift), F = -k/s^2 (m (s/r)^(m+2) - n (s/r)^(n+2)) x
        const float k = pt.p[0], sigma = pt.p[1], eshift = pt.p[2];
        const float inv_r = pd.inv_r;
        float q = sigma * inv_r;
        float qm = ipowf(q, pt.ia), qn = ipowf(q, pt.ib), q2 = q * q;
        float fac = -k / (sigma * sigma) * ((float)pt.ia * qm * q2 - (float)pt.ib * qn * q2);
        fx += fac * x;
        fy += fac * y;
        fz += fac * z;
        return k * (qm - qn) - eshift;
    }
    case POT_SE: {  // E = C e^(-kr)/r + D (s/r)^n  (PotentialsOrder2.h:299-315)
        const float C = pt.p[0], kappa = pt.p[1], D = pt.p[2], sigma = pt.p[3];
        const float r = pd.r, inv_r = pd.inv_r;
        float ex = expf(-kappa * r);
        float qn = ipowf(sigma * inv_r, pt.ia);
        float ff = C * ex * (kappa * inv_r + inv_r * inv_r) + D * (float)pt.ia / sigma * qn * (sigma * inv_r);
        float q = -ff * inv_r;
        fx += q * x;
        fy += q * y;
        fz += q * z;
        return C * ex * inv_r + D * qn;
    }
    default:
        return 0.0f;
    }
}

// ---- fixed-point accumulation (order-independent sums => run-to-run and batch-invariant) -----
constexpr double FORCE_SCALE = 4294967296.0;       // 2^32
constexpr double FORCE_INV = 1.0 / 4294967296.0;
constexpr double ENERGY_SCALE = 16777216.0;        // 2^24
constexpr double ENERGY_INV = 1.0 / 16777216.0;

// Each fixed-point term is clamped to +-2^56 (a pair force of 1.68e7 at 2^-32), so that sums of up to
// 128 terms never wrap in int64; a clamped term (or a NaN, which the clamp turns into -2^56) is
// detected by fixed_saturated and reported as a numeric error instead of being summed silently.
constexpr float FIXED_LIM = 72057594037927936.0f;  // 2^56
constexpr long long FIXED_LIM_I = 72057594037927936ll;

__device__ __forceinline__ long long to_fixed(float v, float scale) {
    float s = v * scale;
    s = fminf(fmaxf(s, -FIXED_LIM), FIXED_LIM);
    return (long long)s;
}

// a fixed-point force term or sum at the clamp (a pair force beyond the range)
__device__ __forceinline__ bool fixed_saturated(long long v) { return v >= FIXED_LIM_I || v <= -FIXED_LIM_I; }

// The force on a particle from one neighbour, summed over the potentials of their ordered type
// pair, in 2^-32 fixed point (fx, fy, fz), and the pair energy in 2^-24 fixed point (e).  dx, dy, dz
// is the FP64 minimum-image displacement neighbour - particle and d2 its FP64 square; the pair
// counts only when d2 < the pair's largest cutoff^2 (CPUCalculateForces.cpp:243), else every output
// is zero and the result is false.  Every pair-force sum of the engine goes through this function,
// so equal inputs give equal integers wherever a sum is formed or corrected.  ENERGY = false
// leaves e zero (the force terms are the same).
template <bool ENERGY = true>
__device__ __forceinline__ bool pair_fixed(const PotDev *pots, const PairIdx &pi, double dx, double dy, double dz,
                                           double d2, long long &fx, long long &fy, long long &fz, long long &e) {
    fx = fy = fz = e = 0;
    if (!(d2 < pi.maxcut2)) return false;
    float pfx = 0.f, pfy = 0.f, pfz = 0.f, pe = 0.f;
    const float xf = (float)dx, yf = (float)dy, zf = (float)dz, d2f = (float)d2;
    const PairDist pd = pair_dist(d2f);
    for (uint32_t p = 0; p < pi.count; ++p) {
        const PotDev pot = pots[pi.start + p];
        if (d2 < pot.cut2) pe += pair_force_energy(pot, d2f, pd, xf, yf, zf, pfx, pfy, pfz);
    }
    fx = to_fixed(pfx, 4294967296.0f);
    fy = to_fixed(pfy, 4294967296.0f);
    fz = to_fixed(pfz, 4294967296.0f);
    if (ENERGY) e = to_fixed(pe, 16777216.0f);
    return true;
}

// The virial of one pair as CPUCalculateForces.cpp:247-249 forms it, -x (x) f, with x the displacement
// neighbour - particle and f the force on the particle summed over the pair's potentials (FP32 pair math,
// as pair_fixed), each of the nine products added to v in 2^-24 fixed point.  Seen from the other side
// x and f both change sign exactly, so the product is the same integers: summing both sides and halving
// counts every unordered pair once.
__device__ __forceinline__ void pair_virial(const PotDev *pots, const PairIdx &pi, double dx, double dy, double dz,
                                            double d2, long long v[9]) {
    if (!(d2 < pi.maxcut2)) return;
    float f[3] = {0.f, 0.f, 0.f};
    const float x[3] = {(float)dx, (float)dy, (float)dz};
    const float d2f = (float)d2;
    const PairDist pd = pair_dist(d2f);
    for (uint32_t p = 0; p < pi.count; ++p) {
        const PotDev pot = pots[pi.start + p];
        if (d2 < pot.cut2) (void)pair_force_energy(pot, d2f, pd, x[0], x[1], x[2], f[0], f[1], f[2]);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) v[3 * a + b] += to_fixed(-x[a] * f[b], 16777216.0f);
}

__device__ __forceinline__ long long to_fixed_d(double v, double scale) {
    double s = v * scale;
    const double lim = 72057594037927936.0;  // 2^56, as to_fixed
    s = fmin(fmax(s, -lim), lim);
    return (long long)s;
}

}  // namespace rdy
