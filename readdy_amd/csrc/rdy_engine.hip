// Host side of the MI355X ReaDDy engine: device pools, replica admission, the per-step launch
// sequence (eager and hipGraph), and the extern "C" ABI declared in include/readdy_amd.h.
#include "rdy_observe.hip"
#include "rdy_topo.hip"

#include "../../include/readdy_amd.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

using namespace rdy;

namespace {

thread_local std::string g_err;

int set_err(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(x)                                                                                       \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) return set_err(RDY_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct HostReplica {
    bool in_use = false;
    uint64_t cell_base = 0, n_cells = 0;
    uint32_t topo_base = 0, n_topo = 0;
    uint32_t topo_range = 0;    // topology handle range claimed at upload: n_topo of the upload + the reserve
    uint32_t topo_reserve = 0;  // handles reserved for topology-particle fusions (rdy_reserve_topology_particles)
    uint32_t tt_base = 0, n_tt = 0;  // topologies (their types) for spatial topology reactions
    bool has_particles = false;  // particles were uploaded since the replica was added
    ReplicaDev dev{};
};

template <class T>
hipError_t dalloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void **)p, n * sizeof(T));
}

// Workspace of rdy_histogram_along_axis_noalloc, one per GPU for the life of the process: the bin
// borders (up to RDF_MAX_BORDERS doubles), the species mask (MAX_TYPES bits) and the counts of one
// call.  Host code never passes the array's name to a kernel or a copy: it uses the device address
// that hipGetSymbolAddress returns for the current GPU (g_noalloc_ptr), taken once per device when the
// first engine is created there.
constexpr size_t NOALLOC_OFF_BORDERS = 0;
constexpr size_t NOALLOC_OFF_MASK = 33024;   // 8 * 4097 rounded up to 256
constexpr size_t NOALLOC_OFF_COUNTS = 33280; // + 4 * (MAX_TYPES / 32), rounded up to 256
constexpr size_t NOALLOC_BYTES = 66048;      // + 8 * (RDF_MAX_BORDERS - 1)
static_assert(NOALLOC_OFF_MASK >= 8 * RDF_MAX_BORDERS && NOALLOC_OFF_COUNTS >= NOALLOC_OFF_MASK + 4 * (MAX_TYPES / 32) &&
                  NOALLOC_BYTES >= NOALLOC_OFF_COUNTS + 8 * (RDF_MAX_BORDERS - 1),
              "no-alloc workspace layout");
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// the current device's address of the workspace (device must be current)
int init_noalloc_workspace(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return set_err(RDY_ERR_INVALID, "device index beyond 64");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return RDY_OK;
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_err(RDY_ERR_HIP, "hipGetSymbolAddress returned no address for the workspace");
    g_noalloc_ptr[device] = p;
    return RDY_OK;
}

__global__ void k_zero_u64(unsigned long long *p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

}  // namespace

struct rdy_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    // chemistry
    uint32_t n_types = 0, n_rx = 0;
    double kbt = 1.0;
    std::vector<double> D;
    std::vector<uint8_t> topo_flavor;
    std::vector<rdy_pair_potential> pots_in;
    std::vector<rdy_reaction> rx_in;
    // configuration
    double nl_cutoff = 0.0;
    int radius = 1;
    uint32_t cap = 16;
    uint32_t block = 256;
    uint32_t max_reps = 1;
    uint64_t max_cells = 0, max_particles = 0;
    uint32_t ev_per_rep = RESOLVE_MAX_EVENTS;
    bool pools_ready = false;
    std::vector<std::pair<uint64_t, uint64_t>> free_cells;  // (start, length)
    std::vector<HostReplica> reps_h;
    uint32_t n_reps_used = 0;  // 1 + highest replica index ever used
    uint64_t version = 0;      // bumps on any change that invalidates a captured graph
    // device pools
    ReplicaDev *reps = nullptr;
    State *bins[2] = {nullptr, nullptr};
    uint32_t *counts[2] = {nullptr, nullptr};
    uint32_t *wl[2] = {nullptr, nullptr};
    uint32_t *wl_count = nullptr;  // [2]
    uint32_t *ovf_cell[2] = {nullptr, nullptr};
    uint32_t *ovf_count = nullptr;  // [2]
    uint32_t ovf_cap = 0;
    bool auto_cap = false;
    uint32_t *tile_sums = nullptr;
    // LDS-staged bricks
    uint32_t brick[3] = {0, 0, 0};
    uint32_t max_states = 0, max_halo = 0;
    size_t brick_smem = 0;
    size_t tables_lds = 0;  // bytes of brick_smem reserved for the per-block type tables
    size_t lds_limit = 65536;
    std::vector<BrickDev> bricks_h;
    BrickDev *d_bricks = nullptr;
    size_t d_bricks_cap = 0;
    uint32_t *d_slow = nullptr;  // two lists of bricks left to the bin-reading kernels (forces, search)
    uint32_t n_bricks_ah = 0;    // leading bricks of bricks_h in halo mode on every axis
    bool bricks_dirty = false;   // replicas changed since the last brick upload
    bool wl_valid[2] = {true, true};
    uint64_t cells_hwm = 0;  // end of the highest cell range handed out (the scan covers [0, hwm))
    Event *events = nullptr;
    uint32_t *ev_count = nullptr;
    RxRecord *recs = nullptr;  // reaction records of each replica's last reaction step (when recording)
    uint32_t *rec_count = nullptr;
    bool recording = false;
    uint32_t break_calls = 0;  // Philox counter of rdy_break_bonds
    unsigned long long *rx_step = nullptr, *rx_total = nullptr, *e_bonded = nullptr;
    unsigned long long *virial = nullptr;  // [max_reps][9] 2^-24 fixed point, both sides of every pair
    bool record_virial = false;            // Context::recordVirial (Context.h:168-173)
    bool virial_valid = false;             // the last force evaluation formed the virial
    uint32_t *err = nullptr;
    double *force_out = nullptr;
    // fused step (rdy_step.hip): the walk's per-brick force rows (2^-32 fixed point), the records
    // changed by reactions, the records moved again and the copies withdrawn; force_out and the
    // energies are formed on demand after a run (export_valid)
    Edit *edits = nullptr;
    uint32_t *edit_count = nullptr;  // [3]: changed records, records to move again, copies to withdraw
    uint32_t edit_cap = 0;
    uint32_t *aff_bits = nullptr;
    uint32_t *aff_list = nullptr;
    uint32_t aff_cap = 0;
    uint2 *wd_list = nullptr;
    uint32_t wd_cap = 0;
    bool export_valid = true;
    // tables
    PotDev *d_pots = nullptr;
    PairIdx *d_pidx = nullptr;
    RxDev *d_rx = nullptr;
    uint32_t *d_o1idx = nullptr, *d_o1list = nullptr, *d_o2list = nullptr;
    // DetailedBalance tables, built on the handler's first call (they do not depend on the time step)
    DbPot *d_dbpots = nullptr;
    uint32_t *d_dbpidx = nullptr, *d_dbtable = nullptr;
    int32_t *d_dbkind = nullptr;
    double *d_dbpref = nullptr, *d_dbradii = nullptr, *d_dbcdf = nullptr;
    bool db_ready = false;
    DbTables db{};
    RxPairIdx *d_o2idx = nullptr;
    float2 *d_rowmax = nullptr;
    double *d_D = nullptr;
    double *d_bd_scale = nullptr;
    uint8_t *d_flavor = nullptr;
    std::vector<RxDev> rx_h;
    double tables_dt = -1.0;
    Tables T{};
    // topology
    long long *fb = nullptr;
    uint32_t *loc = nullptr;
    uint64_t topo_cap = 0, topo_used = 0;
    std::vector<std::pair<uint64_t, uint64_t>> free_topo;  // released (start, length) ranges of loc / fb
    uint32_t *topo_of = nullptr;     // [topo_cap] topology index (within the replica) of each topology particle
    uint32_t *topo_owner = nullptr;  // [topo_cap] replica of each topology particle slot, UINT32_MAX if unused
    uint32_t *loc_cur = nullptr;     // [topo_cap] fused step: topology particles' slots in the buffer it reads
    // spatial topology reactions (topology-particle enzymatic) and the topology types they read and write
    std::vector<rdy_spatial_reaction> sr_in;
    SpatialDev *d_sr = nullptr;
    uint32_t *ttype = nullptr;
    uint32_t *treact = nullptr;      // [tt_cap] serial of the evaluation each topology last reacted in
    uint32_t topo_serial = 0;        // serial of the last separate topology-reaction evaluation
    uint64_t tt_cap = 0, tt_used = 0;
    std::vector<std::pair<uint64_t, uint64_t>> free_tt;
    TopoEvent *tev = nullptr;
    uint32_t *tev_count = nullptr;
    uint32_t *tfus = nullptr, *tfus_count = nullptr;  // topology-particle fusions of the last evaluation
    uint32_t *tmrg = nullptr, *tmrg_count = nullptr;  // topology-topology fusions (merges) of the last evaluation
    uint32_t *adj_off = nullptr, *adj = nullptr;       // bond adjacency of the topology particle slots (CSR)
    uint64_t adj_n = 0, adj_cap = 0, adj_off_cap = 0;
    unsigned long long *str_step = nullptr, *str_total = nullptr;
    std::vector<BondTerm> terms_h[3];
    BondTerm *d_terms[3] = {nullptr, nullptr, nullptr};
    uint64_t d_terms_cap[3] = {0, 0, 0};
    // scratch
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    double *d_noise = nullptr;
    size_t noise_cap = 0;
    int cur = 0;
    // graph
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    uint64_t g_version = ~0ull;
    int g_handler = -1;
    double g_dt = -1.0;
    // a run taken a few steps at a time (rdy_run_begin / rdy_run_enqueue / rdy_run_end)
    bool in_session = false;
    uint64_t session_left = 0;
    int session_handler = 0;
    double session_dt = 0.0;
    // profiling of the fused kernel
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_ev;
    size_t prof_used = 0;
    double prof_ms = 0.0;
    uint64_t prof_n = 0;

    Grid grid(int b) const {
        return Grid{bins[b], counts[b], wl[b], wl_count + b, ovf_cell[b], ovf_count + b, (uint32_t)(max_cells * cap),
                    ovf_cap};
    }
    Topo topo() const { return Topo{fb, loc}; }
};

namespace {

bool is_order2(uint32_t kind) { return kind == RDY_RX_FUSION || kind == RDY_RX_ENZYMATIC; }

double pot_cutoff(const rdy_pair_potential &p) {
    switch (p.kind) {
    case RDY_POT_HARMONIC_REPULSION: return p.params[1];
    case RDY_POT_WEAK_INTERACTION: return p.params[3];
    case RDY_POT_LENNARD_JONES: return p.params[2];
    case RDY_POT_SCREENED_ELECTROSTATICS: return p.params[5];
    default: return 0.0;
    }
}

int ensure_scratch(rdy_engine *e, size_t bytes) {
    if (bytes <= e->scratch_bytes) return RDY_OK;
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (e->scratch) HIP_TRY(hipFree(e->scratch));
    e->scratch = nullptr;
    size_t b = std::max(bytes, (size_t)1 << 20);
    HIP_TRY(hipMalloc(&e->scratch, b));
    e->scratch_bytes = b;
    return RDY_OK;
}

int check_errors(rdy_engine *e) {
    hipError_t le = hipGetLastError();  // launch-time errors (e.g. a launch configuration the device rejects)
    if (le != hipSuccess) return set_err(RDY_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(le));
    uint32_t h[N_ERR];
    HIP_TRY(hipMemcpyAsync(h, e->err, sizeof h, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    bool any = false;
    for (int i = 0; i < N_ERR; ++i) any = any || h[i];
    if (!any) return RDY_OK;
    HIP_TRY(hipMemsetAsync(e->err, 0, sizeof(uint32_t) * N_ERR, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (h[ERR_FORCE_RANGE])
        return set_err(RDY_ERR_NUMERIC, "a force beyond the fixed-point range (|F| >= 1.68e7 per pair or particle, " +
                                            std::to_string(h[ERR_FORCE_RANGE]) +
                                            " terms): particles nearly coincide; reduce the time step");
    if (h[ERR_NONFINITE])
        return set_err(RDY_ERR_NUMERIC, "integrator produced non-finite positions (" + std::to_string(h[ERR_NONFINITE]) +
                                            " components); reduce the time step");
    std::string m;
    if (h[ERR_BIN_OVERFLOW])
        m += "cell bin overflow (" + std::to_string(h[ERR_BIN_OVERFLOW]) + " particles): raise bin_capacity; ";
    if (h[ERR_EVENT_OVERFLOW])
        m += "reaction event segment overflow (" + std::to_string(h[ERR_EVENT_OVERFLOW]) +
             "): too many accepted reactions per step, reduce the time step; ";
    if (h[ERR_LDS_LAYOUT]) return set_err(RDY_ERR_STATE, "brick kernels: LDS layout exceeds the reserved size");
    if (h[ERR_EVENT_LDS]) m += "more than " + std::to_string(RESOLVE_MAX_EVENTS) + " accepted events in one replica step; ";
    if (h[ERR_WL_OVERFLOW]) m += "particle pool exhausted: raise max_particles; ";
    if (h[ERR_TOPO_RESERVE])
        m += "topology-particle fusions beyond the replica's topology reserve (" + std::to_string(h[ERR_TOPO_RESERVE]) +
             " skipped): raise it with rdy_reserve_topology_particles; ";
    if (h[ERR_EDIT_OVERFLOW])
        m += "changed-record list of the fused step overflowed (" + std::to_string(h[ERR_EDIT_OVERFLOW]) +
             "): too many reactions per step, reduce the time step; ";
    return set_err(RDY_ERR_CAPACITY, m);
}

// order-dependent tables (p_accept depends on dt)
int upload_tables(rdy_engine *e, double dt) {
    if (e->tables_dt == dt && e->d_rx) return RDY_OK;
    HIP_TRY(hipStreamSynchronize(e->stream));  // old tables may still be read by queued kernels
    const uint32_t nt = e->n_types;
    // pair potentials: symmetric type-pair lists in registration order (PotentialRegistry.h:65-70)
    std::vector<std::vector<uint32_t>> lists((size_t)nt * nt);
    std::vector<PotDev> pots;
    for (size_t k = 0; k < e->pots_in.size(); ++k) {
        const rdy_pair_potential &p = e->pots_in[k];
        PotDev d{};
        d.kind = (int32_t)p.kind;
        const double *q = p.params;
        switch (p.kind) {
        case RDY_POT_HARMONIC_REPULSION:
            d.cut2 = q[1] * q[1];
            d.p[0] = (float)q[0];
            d.p[1] = (float)q[1];
            break;
        case RDY_POT_WEAK_INTERACTION: {
            d.cut2 = q[3] * q[3];
            double len = q[3] - q[1];
            d.p[0] = (float)q[0];
            d.p[1] = (float)q[1];
            d.p[2] = (float)q[2];
            d.p[3] = (float)q[3];
            d.p[4] = (float)((1. / (.5 * len)) * (1. / (.5 * len)));
            d.p[5] = (float)(q[1] + .5 * len);
            break;
        }
        case RDY_POT_LENNARD_JONES: {  // Potentials.cpp:157-171
            double m = q[0], n = q[1], rc = q[2], eps = q[4], sigma = q[5];
            double r_min = sigma * std::pow(n / m, 1. / (n - m));
            double k = -eps / (std::pow(sigma / r_min, m) - std::pow(sigma / r_min, n));
            d.cut2 = rc * rc;
            d.ia = (int32_t)m;
            d.ib = (int32_t)n;
            d.p[0] = (float)k;
            d.p[1] = (float)sigma;
            d.p[2] = q[3] != 0.0 ? (float)(k * (std::pow(sigma / rc, m) - std::pow(sigma / rc, n))) : 0.0f;
            break;
        }
        case RDY_POT_SCREENED_ELECTROSTATICS:
            d.cut2 = q[5] * q[5];
            d.ia = (int32_t)q[4];
            d.p[0] = (float)q[0];
            d.p[1] = (float)q[1];
            d.p[2] = (float)q[2];
            d.p[3] = (float)q[3];
            break;
        }
        uint32_t id = (uint32_t)pots.size();
        pots.push_back(d);
        lists[(size_t)p.type1 * nt + p.type2].push_back(id);
        if (p.type1 != p.type2) lists[(size_t)p.type2 * nt + p.type1].push_back(id);
    }
    std::vector<PotDev> flat;
    std::vector<PairIdx> pidx((size_t)nt * nt);
    for (size_t k = 0; k < lists.size(); ++k) {
        pidx[k].start = (uint16_t)flat.size();
        pidx[k].count = (uint16_t)lists[k].size();
        pidx[k].maxcut2 = 0.0;
        for (uint32_t id : lists[k]) {
            flat.push_back(pots[id]);
            pidx[k].maxcut2 = std::max(pidx[k].maxcut2, pots[id].cut2);
        }
        pidx[k].maxcut2_f = (float)pidx[k].maxcut2;
    }
    if (flat.size() > 65535) return set_err(RDY_ERR_INVALID, "too many pair potentials");
    // reactions
    e->rx_h.assign(e->n_rx, RxDev{});
    std::vector<std::vector<uint32_t>> o1(nt), o2((size_t)nt * nt);
    for (uint32_t i = 0; i < e->n_rx; ++i) {
        const rdy_reaction &x = e->rx_in[i];
        RxDev &d = e->rx_h[i];
        d.kind = (int32_t)x.kind;
        d.rate = x.rate;
        d.p_accept = 1 - std::exp(-x.rate * dt);  // ReactionUtils.h:68-74, approximateRate = false
        d.educt_dist2 = x.educt_distance * x.educt_distance;
        d.product_distance = x.product_distance;
        double w1 = x.weight1, w2 = x.weight2, sum = w1 + w2;
        if (sum != 1) {  // Fusion.h:61-68, Fission.h:63-70
            w1 /= sum;
            w2 /= sum;
        }
        d.w1 = w1;
        d.w2 = w2;
        d.educts[0] = x.educts[0];
        d.educts[1] = x.educts[1];
        d.products[0] = x.products[0];
        d.products[1] = x.products[1];
        d.order = is_order2(x.kind) ? 2 : 1;
        if (d.order == 1) {
            o1[x.educts[0]].push_back(i);
        } else {
            o2[(size_t)x.educts[0] * nt + x.educts[1]].push_back(i);
            if (x.educts[0] != x.educts[1]) o2[(size_t)x.educts[1] * nt + x.educts[0]].push_back(i);
        }
    }
    std::vector<uint32_t> o1idx(nt + 1), o1list, o2list;
    std::vector<RxPairIdx> o2idx((size_t)nt * nt);
    for (uint32_t t = 0; t < nt; ++t) {
        o1idx[t] = (uint32_t)o1list.size();
        o1list.insert(o1list.end(), o1[t].begin(), o1[t].end());
    }
    o1idx[nt] = (uint32_t)o1list.size();
    bool has_o2 = false;
    for (size_t k = 0; k < o2.size(); ++k) {
        o2idx[k].start = (uint16_t)o2list.size();
        o2idx[k].count = (uint16_t)o2[k].size();
        double m2 = 0.0;
        for (uint32_t rx : o2[k]) m2 = std::max(m2, e->rx_h[rx].educt_dist2);
        o2idx[k].maxr2_f = (float)m2;
        has_o2 = has_o2 || !o2[k].empty();
        o2list.insert(o2list.end(), o2[k].begin(), o2[k].end());
    }
    std::vector<float2> rowmax(nt);
    for (uint32_t t = 0; t < nt; ++t) {
        float mf = -1.f, mr = -1.f;
        for (uint32_t u = 0; u < nt; ++u) {
            if (pidx[(size_t)t * nt + u].count) mf = std::max(mf, pidx[(size_t)t * nt + u].maxcut2_f);
            if (o2idx[(size_t)t * nt + u].count) mr = std::max(mr, o2idx[(size_t)t * nt + u].maxr2_f);
        }
        rowmax[t] = make_float2(mf, mr);
    }
    auto up = [&](auto **dst, const auto &v) -> int {
        using T = std::remove_pointer_t<std::remove_reference_t<decltype(*dst)>>;
        if (*dst) HIP_TRY(hipFree(*dst));
        HIP_TRY(dalloc(dst, v.size()));
        if (!v.empty()) HIP_TRY(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return RDY_OK;
    };
    int rc;
    if ((rc = up(&e->d_pots, flat))) return rc;
    if ((rc = up(&e->d_pidx, pidx))) return rc;
    if ((rc = up(&e->d_rx, e->rx_h))) return rc;
    if ((rc = up(&e->d_o1idx, o1idx))) return rc;
    if ((rc = up(&e->d_o1list, o1list))) return rc;
    if ((rc = up(&e->d_o2idx, o2idx))) return rc;
    if ((rc = up(&e->d_o2list, o2list))) return rc;
    if ((rc = up(&e->d_rowmax, rowmax))) return rc;
    if ((rc = up(&e->d_D, e->D))) return rc;
    {  // sqrt(2 D dt) per species, formed as CPUEulerBDIntegrator.cpp:63-73 forms it (FP64, correctly rounded)
        std::vector<double> sc(nt);
        for (uint32_t t = 0; t < nt; ++t) sc[t] = std::sqrt(2. * e->D[t] * dt);
        if ((rc = up(&e->d_bd_scale, sc))) return rc;
    }
    if (!e->d_flavor) {
        if ((rc = up(&e->d_flavor, e->topo_flavor))) return rc;
    }
    e->T.pots = e->d_pots;
    e->T.pidx = e->d_pidx;
    e->T.rx = e->d_rx;
    e->T.o1idx = e->d_o1idx;
    e->T.o1list = e->d_o1list;
    e->T.o2idx = e->d_o2idx;
    e->T.o2list = e->d_o2list;
    e->T.diffusion = e->d_D;
    e->T.bd_scale = e->d_bd_scale;
    e->T.rowmax = e->d_rowmax;
    e->T.n_types = nt;
    e->T.n_rx = e->n_rx;
    e->T.has_o2 = has_o2 ? 1u : 0u;
    e->T.n_pots = (uint32_t)flat.size();
    e->T.kbt = e->kbt;
    // the brick kernels place the type tables ahead of the staged records in the dynamic LDS that
    // configure_bricks reserved (from the registered potentials, before these tables existed)
    if (e->brick_smem && table_bytes(std::min(nt, MAX_STAGED_TYPES), e->T.n_pots) > e->tables_lds)
        return set_err(RDY_ERR_STATE, "type tables exceed the LDS reserved for them");
    if (!e->sr_in.empty()) {  // spatial topology reactions: acceptance 1 - exp(-rate dt) of this time step
        std::vector<SpatialDev> sr;
        for (const auto &x : e->sr_in) {
            SpatialDev d{};
            d.type1 = x.type1;
            d.top_type1 = x.top_type1;
            d.type2 = x.type2;
            d.type_to1 = x.type_to1;
            d.top_type_to1 = x.top_type_to1;
            d.type_to2 = x.type_to2;
            d.rate = x.rate;
            d.p_accept = 1 - std::exp(-x.rate * dt);
            d.radius2 = x.radius * x.radius;
            d.fusion = x.fusion;
            d.tt = x.topology_topology;
            d.top_type2 = x.top_type2;
            d.top_type_to2 = x.top_type_to2;
            d.self_fusion = x.self_fusion;
            sr.push_back(d);
        }
        if ((rc = up(&e->d_sr, sr))) return rc;
    }
    e->tables_dt = dt;
    e->version++;
    return RDY_OK;
}

// returns [start, start + len) to a free list of ranges, merging neighbours
void release_range(std::vector<std::pair<uint64_t, uint64_t>> &fl, uint64_t start, uint64_t len) {
    if (len == 0) return;
    fl.push_back({start, len});
    std::sort(fl.begin(), fl.end());
    std::vector<std::pair<uint64_t, uint64_t>> merged;
    for (auto &fc : fl) {
        if (!merged.empty() && merged.back().first + merged.back().second == fc.first) merged.back().second += fc.second;
        else merged.push_back(fc);
    }
    fl.swap(merged);
}

// first-fit range of len from a free list; false if none is long enough
bool take_range(std::vector<std::pair<uint64_t, uint64_t>> &fl, uint64_t len, uint64_t *start) {
    size_t f = 0;
    while (f < fl.size() && fl[f].second < len) ++f;
    if (f == fl.size()) return false;
    *start = fl[f].first;
    fl[f].first += len;
    fl[f].second -= len;
    if (fl[f].second == 0) fl.erase(fl.begin() + f);
    return true;
}

// uploads the host list of bonded terms of one kind (device array grown as needed)
// the bonds as adjacency lists of the topology particle slots (global: topo_base + handle), for the
// self-fusion test of the topology reactions (a bonded pair does not self-fuse)
int upload_adjacency(rdy_engine *e) {
    const uint64_t n = e->topo_used;
    std::vector<uint32_t> off(n + 1, 0), adj;
    const auto &v = e->terms_h[0];
    for (const BondTerm &t : v) {
        const HostReplica &H = e->reps_h[t.r];
        const uint64_t a = (uint64_t)H.topo_base + t.h[0], b = (uint64_t)H.topo_base + t.h[1];
        if (a < n && b < n) {
            ++off[a + 1];
            ++off[b + 1];
        }
    }
    for (uint64_t i = 0; i < n; ++i) off[i + 1] += off[i];
    adj.resize(off[n]);
    std::vector<uint32_t> fill(off.begin(), off.end() - 1);
    for (const BondTerm &t : v) {
        const HostReplica &H = e->reps_h[t.r];
        const uint64_t a = (uint64_t)H.topo_base + t.h[0], b = (uint64_t)H.topo_base + t.h[1];
        if (a < n && b < n) {
            adj[fill[a]++] = (uint32_t)b;
            adj[fill[b]++] = (uint32_t)a;
        }
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (off.size() > e->adj_off_cap) {
        if (e->adj_off) HIP_TRY(hipFree(e->adj_off));
        e->adj_off = nullptr;
        e->adj_off_cap = std::max<uint64_t>(off.size(), 2 * e->adj_off_cap);
        HIP_TRY(dalloc(&e->adj_off, e->adj_off_cap));
    }
    if (adj.size() > e->adj_cap || !e->adj) {
        if (e->adj) HIP_TRY(hipFree(e->adj));
        e->adj = nullptr;
        e->adj_cap = std::max<uint64_t>(std::max<uint64_t>(adj.size(), 1), 2 * e->adj_cap);
        HIP_TRY(dalloc(&e->adj, e->adj_cap));
    }
    HIP_TRY(hipMemcpyAsync(e->adj_off, off.data(), off.size() * 4, hipMemcpyHostToDevice, e->stream));
    if (!adj.empty()) HIP_TRY(hipMemcpyAsync(e->adj, adj.data(), adj.size() * 4, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->adj_n = off.size();
    return RDY_OK;
}

int upload_terms(rdy_engine *e, int kind) {
    const std::vector<BondTerm> &v = e->terms_h[kind];
    if (v.size() > e->d_terms_cap[kind]) {
        const uint64_t cap = std::max<uint64_t>(v.size(), 2 * e->d_terms_cap[kind]);
        HIP_TRY(hipStreamSynchronize(e->stream));
        if (e->d_terms[kind]) HIP_TRY(hipFree(e->d_terms[kind]));
        e->d_terms[kind] = nullptr;
        e->d_terms_cap[kind] = 0;
        HIP_TRY(dalloc(&e->d_terms[kind], cap));
        e->d_terms_cap[kind] = cap;
    }
    if (!v.empty())
        HIP_TRY(hipMemcpyAsync(e->d_terms[kind], v.data(), v.size() * sizeof(BondTerm), hipMemcpyHostToDevice,
                               e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->version++;
    if (kind == 0) return upload_adjacency(e);
    return RDY_OK;
}

// removes every bonded term of replica r (host lists and device arrays)
int drop_terms(rdy_engine *e, uint32_t r) {
    for (int kind = 0; kind < 3; ++kind) {
        std::vector<BondTerm> &v = e->terms_h[kind];
        const size_t n0 = v.size();
        v.erase(std::remove_if(v.begin(), v.end(), [r](const BondTerm &t) { return t.r == r; }), v.end());
        if (v.size() != n0) {
            int rc = upload_terms(e, kind);
            if (rc) return rc;
        }
    }
    return RDY_OK;
}

// gives replica r's range of topology particle slots (loc / fb) back to the pool
void release_topo(rdy_engine *e, HostReplica &H) {
    if (H.n_tt) release_range(e->free_tt, H.tt_base, H.n_tt);
    H.tt_base = 0;
    H.n_tt = 0;
    if (H.topo_range == 0) return;
    if (e->topo_owner)
        (void)hipMemsetAsync(e->topo_owner + H.topo_base, 0xFF, H.topo_range * sizeof(uint32_t), e->stream);
    release_range(e->free_topo, H.topo_base, H.topo_range);
    // a free range at the end of the used part shrinks it
    if (!e->free_topo.empty() && e->free_topo.back().first + e->free_topo.back().second == e->topo_used) {
        e->topo_used = e->free_topo.back().first;
        e->free_topo.pop_back();
    }
    H.topo_base = 0;
    H.n_topo = 0;
    H.topo_range = 0;
}

int alloc_pools(rdy_engine *e) {
    if (e->pools_ready) return RDY_OK;
    e->ovf_cap = (uint32_t)std::max<uint64_t>(4096, e->max_particles / 8);
    if ((unsigned __int128)e->max_cells * e->cap + e->ovf_cap >= ((unsigned __int128)1 << 32))
        return set_err(RDY_ERR_CAPACITY, "max_cells * bin_capacity must stay below 2^32 slots");
    const uint64_t slots = e->max_cells * e->cap + e->ovf_cap;
    for (int b = 0; b < 2; ++b) {
        HIP_TRY(dalloc(&e->bins[b], slots));
        HIP_TRY(dalloc(&e->ovf_cell[b], e->ovf_cap));
        HIP_TRY(dalloc(&e->counts[b], e->max_cells));
        HIP_TRY(hipMemsetAsync(e->counts[b], 0, e->max_cells * sizeof(uint32_t), e->stream));
        HIP_TRY(dalloc(&e->wl[b], e->max_particles));
    }
    HIP_TRY(dalloc(&e->wl_count, 2));
    HIP_TRY(hipMemsetAsync(e->wl_count, 0, 2 * sizeof(uint32_t), e->stream));
    HIP_TRY(dalloc(&e->tile_sums, (e->max_cells + SCAN_TILE - 1) / SCAN_TILE + 2));
    HIP_TRY(dalloc(&e->ovf_count, 2));
    HIP_TRY(hipMemsetAsync(e->ovf_count, 0, 2 * sizeof(uint32_t), e->stream));
    HIP_TRY(dalloc(&e->reps, e->max_reps));
    HIP_TRY(hipMemsetAsync(e->reps, 0, e->max_reps * sizeof(ReplicaDev), e->stream));
    HIP_TRY(dalloc(&e->events, (size_t)e->max_reps * e->ev_per_rep));
    HIP_TRY(dalloc(&e->ev_count, e->max_reps));
    HIP_TRY(hipMemsetAsync(e->ev_count, 0, e->max_reps * sizeof(uint32_t), e->stream));
    const size_t nrx = std::max<uint32_t>(e->n_rx, 1);
    HIP_TRY(dalloc(&e->rx_step, (size_t)e->max_reps * nrx));
    HIP_TRY(dalloc(&e->rx_total, (size_t)e->max_reps * nrx));
    HIP_TRY(hipMemsetAsync(e->rx_step, 0, (size_t)e->max_reps * nrx * 8, e->stream));
    HIP_TRY(hipMemsetAsync(e->rx_total, 0, (size_t)e->max_reps * nrx * 8, e->stream));
    HIP_TRY(dalloc(&e->e_bonded, e->max_reps));
    HIP_TRY(hipMemsetAsync(e->e_bonded, 0, e->max_reps * 8, e->stream));
    HIP_TRY(dalloc(&e->virial, 9ull * e->max_reps));
    HIP_TRY(hipMemsetAsync(e->virial, 0, 9ull * e->max_reps * 8, e->stream));
    HIP_TRY(dalloc(&e->err, N_ERR));
    HIP_TRY(hipMemsetAsync(e->err, 0, N_ERR * sizeof(uint32_t), e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->free_cells.clear();
    e->free_cells.push_back({0, e->max_cells});
    e->reps_h.assign(e->max_reps, HostReplica{});
    e->pools_ready = true;
    return RDY_OK;
}

uint32_t n_bricks(const rdy_engine *e);

int ensure_step_pools(rdy_engine *e) {
    if (e->edits) return RDY_OK;
    const uint64_t slots = e->max_cells * e->cap + e->ovf_cap;
    HIP_TRY(dalloc(&e->aff_bits, slots / 32 + 1));
    HIP_TRY(hipMemsetAsync(e->aff_bits, 0, (slots / 32 + 1) * sizeof(uint32_t), e->stream));
    // at most three changed records per accepted event (a fusion kills two and inserts one); each
    // changes the forces of its neighbours within the cutoff
    e->edit_cap = (uint32_t)std::min<uint64_t>(3ull * e->max_reps * e->ev_per_rep, 1ull << 22);
    e->aff_cap = (uint32_t)std::min<uint64_t>(32ull * e->edit_cap, std::max<uint64_t>(slots, 1ull << 16));
    e->wd_cap = e->aff_cap;
    HIP_TRY(dalloc(&e->edits, e->edit_cap));
    HIP_TRY(dalloc(&e->aff_list, e->aff_cap));
    HIP_TRY(dalloc(&e->wd_list, e->wd_cap));
    HIP_TRY(dalloc(&e->edit_count, 3));
    HIP_TRY(hipMemsetAsync(e->edit_count, 0, 3 * sizeof(uint32_t), e->stream));
    return RDY_OK;
}

int ensure_force_out(rdy_engine *e) {
    if (e->force_out) return RDY_OK;
    const uint64_t slots = e->max_cells * e->cap + e->ovf_cap;
    HIP_TRY(dalloc(&e->force_out, 3 * slots));
    HIP_TRY(hipMemsetAsync(e->force_out, 0, 3 * slots * sizeof(double), e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return RDY_OK;
}

void cell_dims(const rdy_engine *e, const double *box, int32_t *dims, double *cs) {
    // CellLinkedList.cpp:64-70
    double width = e->nl_cutoff / (double)e->radius;
    for (int d = 0; d < 3; ++d) {
        double q = std::floor(box[d] / width);
        dims[d] = (int32_t)std::max(1.0, q);
        cs[d] = box[d] / (double)dims[d];
    }
}

uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)std::max<uint64_t>(1, (n + bs - 1) / bs); }

template <class T>
int put_field(rdy_engine *e, uint32_t r, size_t offset, const T &v) {
    HIP_TRY(hipMemcpyAsync((char *)(e->reps + r) + offset, &v, sizeof(T), hipMemcpyHostToDevice, e->stream));
    return RDY_OK;
}

template <class T>
int get_field(rdy_engine *e, uint32_t r, size_t offset, T *v) {
    HIP_TRY(hipMemcpyAsync(v, (char *)(e->reps + r) + offset, sizeof(T), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return RDY_OK;
}

// the replica's topology particle count as the device's fusions left it
int sync_topo_count(rdy_engine *e, uint32_t r) {
    HostReplica &H = e->reps_h[r];
    if (H.topo_range <= H.n_topo) return RDY_OK;  // no reserve: fusions cannot change the count
    uint32_t n = 0;
    int rc = get_field(e, r, offsetof(ReplicaDev, n_topo), &n);
    if (rc) return rc;
    H.n_topo = n;
    return RDY_OK;
}

int check_replica(rdy_engine *e, uint32_t r) {
    if (!e || !e->pools_ready || r >= e->max_reps || !e->reps_h[r].in_use)
        return set_err(RDY_ERR_INVALID, "no such replica");
    return RDY_OK;
}

// splits replica r's grid into bricks (halo mode where the axis is long enough, whole-axis otherwise)
void make_bricks(rdy_engine *e, uint32_t r, std::vector<BrickDev> &out) {
    const ReplicaDev &R = e->reps_h[r].dev;
    std::vector<std::pair<int32_t, int32_t>> seg[3];
    uint32_t whole = 0, halo[3];
    for (int a = 0; a < 3; ++a) {
        const int32_t D = R.dims[a], B = (int32_t)e->brick[a], h = e->radius;
        const bool w = D <= B || (R.pbc[a] && D < B + 2 * h);
        if (w) {
            whole |= 1u << a;
            seg[a].push_back({0, D});
            halo[a] = 0;
        } else {
            // ceil(D / B) segments of nearly equal length (at most B), so that no brick is a thin sliver
            const int32_t n = (D + B - 1) / B;
            for (int32_t k = 0, o = 0; k < n; ++k) {
                const int32_t len = D / n + (k < D % n ? 1 : 0);
                seg[a].push_back({o, len});
                o += len;
            }
            halo[a] = (uint32_t)h;
        }
    }
    for (auto &x : seg[0])
        for (auto &y : seg[1])
            for (auto &z : seg[2]) {
                BrickDev b{};
                b.o[0] = x.first;
                b.o[1] = y.first;
                b.o[2] = z.first;
                b.e[0] = x.second;
                b.e[1] = y.second;
                b.e[2] = z.second;
                for (int a = 0; a < 3; ++a) b.h[a] = (int32_t)halo[a];
                b.r = r;
                b.whole = whole;
                for (int a = 0; a < 3; ++a) {
                    b.dims[a] = R.dims[a];
                    if (R.pbc[a]) b.pbc |= 1u << a;
                }
                b.cell_base = R.cell_base;
                const uint32_t H1 = (uint32_t)(b.e[1] + 2 * b.h[1]), H2 = (uint32_t)(b.e[2] + 2 * b.h[2]);
                b.m_hh = div_magic(H1 * H2);
                b.m_h2 = div_magic(H2);
                b.m_e1 = div_magic((uint32_t)b.e[1]);
                out.push_back(b);
            }
}

int upload_bricks(rdy_engine *e) {
    e->bricks_h.clear();
    for (uint32_t r = 0; r < e->max_reps; ++r)
        if (e->reps_h[r].in_use) make_bricks(e, r, e->bricks_h);
    // bricks in halo mode on every axis first: they run in their own (specialised) launches
    auto ah_end = std::stable_partition(e->bricks_h.begin(), e->bricks_h.end(),
                                        [](const BrickDev &b) { return b.whole == 0; });
    e->n_bricks_ah = (uint32_t)(ah_end - e->bricks_h.begin());
    if (e->bricks_h.size() > e->d_bricks_cap) {
        HIP_TRY(hipStreamSynchronize(e->stream));
        if (e->d_bricks) HIP_TRY(hipFree(e->d_bricks));
        if (e->d_slow) HIP_TRY(hipFree(e->d_slow));
        e->d_bricks_cap = std::max<size_t>(e->bricks_h.size(), 2 * e->d_bricks_cap);
        HIP_TRY(dalloc(&e->d_bricks, e->d_bricks_cap));
        HIP_TRY(dalloc(&e->d_slow, 2 * (e->d_bricks_cap + 2)));
        HIP_TRY(hipMemsetAsync(e->d_slow, 0, 2 * (e->d_bricks_cap + 2) * sizeof(uint32_t), e->stream));
    }
    if (!e->bricks_h.empty())
        HIP_TRY(hipMemcpyAsync(e->d_bricks, e->bricks_h.data(), e->bricks_h.size() * sizeof(BrickDev),
                               hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return RDY_OK;
}

uint32_t n_bricks(const rdy_engine *e) { return (uint32_t)std::max<size_t>(e->bricks_h.size(), 1); }

// brick tables of the current replicas on the device (before any brick launch or graph capture)
int ensure_bricks(rdy_engine *e) {
    if (!e->bricks_dirty) return RDY_OK;
    int rc = upload_bricks(e);
    if (rc == RDY_OK) e->bricks_dirty = false;
    return rc;
}

// launch(all_halo, blocks, first_brick) for the all-halo bricks, then for the others
template <class F>
void for_brick_groups(const rdy_engine *e, F &&launch) {
    const uint32_t n = (uint32_t)e->bricks_h.size(), na = e->n_bricks_ah;
    if (na) launch(std::true_type{}, na, 0u);
    if (n > na) launch(std::false_type{}, n - na, na);
}

// list of bricks for the bin-reading forces (0) or search (1) kernel
uint32_t *slow_list(const rdy_engine *e, int which) { return e->d_slow + (size_t)which * (e->d_bricks_cap + 2); }

// the record pool handed to the resolve kernels (null unless reactions are being recorded)
RxRecord *rec_ptr(const rdy_engine *e) { return e->recording ? e->recs : nullptr; }

// the spatial topology reaction tables of the engine
TopoTables topo_tables(const rdy_engine *e) {
    TopoTables t{};
    t.sr = e->d_sr;
    t.n_sr = (uint32_t)e->sr_in.size();
    t.topo_of = e->topo_of;
    t.loc = e->loc;
    t.fus = e->tfus;
    t.fus_count = e->tfus_count;
    t.mrg = e->tmrg;
    t.mrg_count = e->tmrg_count;
    t.adj_off = e->adj_off;
    t.adj = e->adj;
    t.adj_n = (uint32_t)e->adj_n;
    t.topo_owner = e->topo_owner;
    t.ttype = e->ttype;
    t.treact = e->treact;
    t.serial = 0;
    t.ev = e->tev;
    t.ev_count = e->tev_count;
    t.ev_cap = TOPO_MAX_EVENTS;
    t.cnt_step = e->str_step;
    t.cnt_total = e->str_total;
    return t;
}

// true when topology reactions can happen (reactions registered and some replica has topologies)
bool topo_reactions_on(const rdy_engine *e) { return !e->sr_in.empty() && e->tt_used > 0 && e->tev; }

// the topology reactions of one step (search, then the greedy resolve); edits: the fused step's list
void launch_topo_reactions(rdy_engine *e, int require_active, Edit *edits, uint32_t *edit_count, uint32_t serial = 0) {
    const uint32_t nr = std::max<uint32_t>(e->n_reps_used, 1);
    const uint32_t n_slots = (uint32_t)e->topo_used;
    TopoTables tt = topo_tables(e);
    tt.serial = serial;
    const uint32_t *slots = edits ? e->loc_cur : e->loc;  // the fused step's walk has moved loc on
    hipLaunchKernelGGL(k_topo_search, dim3(blocks_for(n_slots, 256)), dim3(256), 0, e->stream, e->grid(e->cur),
                       (const ReplicaDev *)e->reps, slots, tt, n_slots, e->cap, e->radius, require_active, e->err);
    hipLaunchKernelGGL(k_topo_resolve, dim3(nr), dim3(256), 0, e->stream, e->grid(e->cur), e->reps, tt, require_active,
                       edits, edit_count, e->edit_cap, e->err);
}

// blocks of the bin-reading kernels (they loop over the listed bricks; none listed: they just exit)
uint32_t slow_blocks(const rdy_engine *e) { return std::min<uint32_t>(n_bricks(e), 256u); }

// brick shape and LDS budget from the expected particles per cell (overridable in the config)
// Brick shape for mean cell occupancy lam on the grid dims: the brick's own particles fill one round
// of the block's lanes (mean + 3 sigma <= BRICK_THREADS; two rounds above 3 particles per cell,
// where one-round bricks would be tiny), since a brick whose particles spill into a further round
// keeps its whole block waiting on one wave.  Among those shapes the one with the fewest staged cells
// over the whole grid (halos included) plus a fixed cost per brick (its round trips, counted as
// 64 / lam cells) wins; ties go to the larger brick.
void choose_brick(double lam, const int32_t *dims, int radius, uint32_t out[3]) {
    lam = std::max(lam, 1.0 / 64.0);
    const double x = (-3.0 + std::sqrt(9.0 + 4.0 * BRICK_THREADS)) / 2.0;  // sqrt of the one-round mean
    const double vmax = (lam >= 3.0 ? 2.0 : 1.0) * x * x / lam;
    double best = 0.0;
    bool found = false;
    for (uint32_t ex = 1; ex <= 16; ++ex)
        for (uint32_t ey = 1; ey <= 16; ++ey)
            for (uint32_t ez = 1; ez <= 16; ++ez) {
                const double V = (double)ex * ey * ez;
                const double halo = (double)(ex + 2 * radius) * (ey + 2 * radius) * (ez + 2 * radius);
                if (V > vmax || halo * std::max(lam, 0.25) > 1500.0) continue;
                double staged = 1.0, n_bricks = 1.0;
                const uint32_t ee[3] = {ex, ey, ez};
                for (int a = 0; a < 3; ++a) {
                    const double D = (double)dims[a];
                    if ((double)ee[a] + 2.0 * radius > D) {
                        staged *= D;  // whole-axis brick: no halo
                    } else {
                        const double n = std::ceil(D / ee[a]);
                        staged *= D + 2.0 * radius * n;
                        n_bricks *= n;
                    }
                }
                const double cost = staged + n_bricks * 64.0 / lam;
                const uint32_t vol = ex * ey * ez, bvol = out[0] * out[1] * out[2];
                if (!found || cost < best * (1.0 - 1e-9) || (cost <= best * (1.0 + 1e-9) && vol > bvol)) {
                    best = cost;
                    out[0] = ex, out[1] = ey, out[2] = ez;
                    found = true;
                }
            }
    if (!found) out[0] = out[1] = out[2] = 2;
}

int configure_bricks(rdy_engine *e, double lam, const int32_t *dims) {
    if (e->brick[0] == 0 || e->brick[1] == 0 || e->brick[2] == 0) choose_brick(lam, dims, e->radius, e->brick);
    e->max_halo = 1;
    for (int a = 0; a < 3; ++a) e->max_halo *= e->brick[a] + 2u * (uint32_t)e->radius;
    if (e->max_halo > 60000) return set_err(RDY_ERR_INVALID, "brick plus halo exceeds 60000 cells");
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, e->device) == hipSuccess && v > 0)
        e->lds_limit = (size_t)v;
    // static LDS of the brick kernels comes out of the same per-workgroup budget
    size_t static_lds = 0;
    const void *kernels[] = {(const void *)k_brick_step<true>,         (const void *)k_brick_step<false>,
                             (const void *)k_brick<MODE_FORCES, true>, (const void *)k_brick<MODE_FORCES, false>,
                             (const void *)k_brick<MODE_REBIN, false>, (const void *)k_brick_search<true>,
                             (const void *)k_brick_search<false>};
    for (const void *k : kernels) {
        hipFuncAttributes fa;
        if (hipFuncGetAttributes(&fa, k) == hipSuccess) static_lds = std::max(static_lds, (size_t)fa.sharedSizeBytes);
    }
    if (static_lds == 0) static_lds = 16384;
    // per-block type tables precede the staged records in dynamic LDS; they are uploaded later, so
    // size them from the registered potentials (each lands in at most two ordered type-pair lists)
    const size_t tables =
        table_bytes(std::min(e->n_types, MAX_STAGED_TYPES), (uint32_t)(2 * e->pots_in.size()));
    e->tables_lds = tables;
    if (e->lds_limit < static_lds + tables + 8192) return set_err(RDY_ERR_INVALID, "not enough LDS per workgroup");
    const size_t budget = e->lds_limit - static_lds - tables - 256;
    // staged records: the expected halo occupancy plus ten standard deviations of a uniform (Poisson)
    // fill; fuller halos (clustered systems) take the slow path
    uint32_t want = e->max_states;
    if (want == 0) {
        const double mean = std::max(lam, 0.25) * e->max_halo;
        want = (uint32_t)std::max(512.0, mean + 10.0 * std::sqrt(mean) + 64.0);
    }
    want = std::min<uint32_t>(want, 60000);
    while (want > 256 && brick_smem_bytes(want, e->max_halo) > budget) want -= 128;
    if (brick_smem_bytes(want, e->max_halo) > budget)
        return set_err(RDY_ERR_INVALID, "brick does not fit in LDS: use a smaller brick");
    e->max_states = want;
    e->brick_smem = tables + brick_smem_bytes(want, e->max_halo);
    if (e->brick_smem > 65536)
        for (const void *k : kernels)
            (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e->brick_smem);
    return RDY_OK;
}

// launches the bonded-term kernels on the current buffer
void launch_bonded(rdy_engine *e) {
    BondTerm *t;
    uint32_t n;
    Grid g = e->grid(e->cur);
    if ((n = (uint32_t)e->terms_h[0].size()) && (t = e->d_terms[0]))
        hipLaunchKernelGGL(k_bonds, dim3(blocks_for(n, 256)), dim3(256), 0, e->stream, g, e->reps, e->topo(), t, n,
                           e->e_bonded);
    if ((n = (uint32_t)e->terms_h[1].size()) && (t = e->d_terms[1]))
        hipLaunchKernelGGL(k_angles, dim3(blocks_for(n, 256)), dim3(256), 0, e->stream, g, e->reps, e->topo(), t, n,
                           e->e_bonded);
    if ((n = (uint32_t)e->terms_h[2].size()) && (t = e->d_terms[2]))
        hipLaunchKernelGGL(k_dihedrals, dim3(blocks_for(n, 256)), dim3(256), 0, e->stream, g, e->reps, e->topo(), t, n,
                           e->e_bonded);
}

uint32_t wl_blocks(const rdy_engine *e) { return blocks_for(e->max_particles, e->block); }

// rebuilds buffer b's work list in cell order from its cell counters
void launch_wl_rebuild(rdy_engine *e, int b) {
    Grid g = e->grid(b);
    const uint32_t nc = (uint32_t)std::max<uint64_t>(e->cells_hwm, 1);
    const uint32_t tiles = (nc + SCAN_TILE - 1) / SCAN_TILE;
    hipLaunchKernelGGL(k_wl_tiles, dim3(tiles), dim3(SCAN_THREADS), 0, e->stream, g, e->cap, nc, e->tile_sums);
    hipLaunchKernelGGL(k_wl_tile_scan, dim3(1), dim3(1024), 0, e->stream, e->tile_sums, tiles, g,
                       (uint32_t)e->max_particles, e->err);
    hipLaunchKernelGGL(k_wl_write, dim3(tiles), dim3(SCAN_THREADS), 0, e->stream, g, e->cap, nc,
                       (const uint32_t *)e->tile_sums, tiles, (uint32_t)e->max_particles);
    e->wl_valid[b] = true;
}

// the work list is only needed by per-particle helpers (gather, exports, the separate integrator)
void ensure_wl(rdy_engine *e) {
    if (!e->wl_valid[e->cur]) launch_wl_rebuild(e, e->cur);
}

// clears buffer b's cell counters (brick-wise) and its work-list / overflow lengths
void launch_clear_bricks(rdy_engine *e, int b, int grid_of_bricks_buffer) {
    if (!e->bricks_h.empty())
        hipLaunchKernelGGL(k_brick_search<false>, dim3(n_bricks(e)), dim3(BRICK_THREADS), 0, e->stream,
                           e->grid(grid_of_bricks_buffer), e->grid(b), 1, e->reps, e->T, (const BrickDev *)e->d_bricks,
                           0u, e->cap, e->radius, 2, e->events, e->ev_count, e->max_states, e->max_halo, 0u,
                           slow_list(e, 1), e->err);
    else
        hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, e->stream, e->grid(b));
}

// workgroups of the fix-up kernels (their waves / threads stride over the change lists)
constexpr uint32_t FIX_BLOCKS = 1024;

// one walk of k_brick_step from cur into the other buffer (prologue: the first walk of a run)
// the walk of a step (or the prologue); slow_later: the bricks listed for the bin-reading path are
// walked by k_resolve_step instead of a follow-up launch
void launch_walk(rdy_engine *e, int prologue, int handler, double dt, bool slow_later = false) {
    const int c = e->cur, n = 1 - c;
    if (e->prof && !prologue && e->prof_used < e->prof_ev.size())
        (void)hipEventRecord(e->prof_ev[e->prof_used].first, e->stream);
    for_brick_groups(e, [&](auto ah, uint32_t nb, uint32_t b0) {
        hipLaunchKernelGGL((k_brick_step<decltype(ah)::value>), dim3(nb), dim3(BRICK_THREADS), e->brick_smem,
                           e->stream, e->grid(c), e->grid(n), prologue, e->reps, e->T, e->topo(),
                           (const BrickDev *)e->d_bricks, b0, e->cap, e->radius, handler, dt, e->events, e->ev_count,
                           e->max_states, e->max_halo, (uint32_t)e->brick_smem, slow_list(e, 0), e->err);
    });
    if (e->prof && !prologue && e->prof_used < e->prof_ev.size()) {
        (void)hipEventRecord(e->prof_ev[e->prof_used].second, e->stream);
        e->prof_used++;
    }
    if (slow_later) return;
    hipLaunchKernelGGL(k_brick_step_slow, dim3(slow_blocks(e)), dim3(BRICK_THREADS), 0, e->stream, e->grid(c),
                       e->grid(n), prologue, e->reps, e->T, e->topo(), (const BrickDev *)e->d_bricks, e->cap,
                       e->radius, handler, dt, e->events, e->ev_count, slow_list(e, 0), e->err);
}

// the first walk of a run: the forces of the current state and step 1's move (bonded terms first)
void launch_prologue(rdy_engine *e, double dt) {
    launch_bonded(e);
    launch_walk(e, 1, (int)RDY_HANDLER_NONE, dt);
    e->wl_valid[1 - e->cur] = false;
    e->cur = 1 - e->cur;
}

// the end of a run: the buffer the last walk read is cleared, and so are the bonded forces
void launch_epilogue(rdy_engine *e) {
    const uint64_t nc = std::max<uint64_t>(e->cells_hwm, 1);
    const uint64_t nfb = e->fb ? 3 * e->topo_used : 0;
    hipLaunchKernelGGL(k_clear_grid, dim3(blocks_for(std::max(nc, nfb), 256)), dim3(256), 0, e->stream,
                       e->grid(1 - e->cur), (uint32_t)nc, e->fb, (uint32_t)nfb);
}

// one fused step over all replicas (SimulationLoop.h:280-300, rdy_step.hip): the walk over x'(t)
// yields step t's candidates and, optimistically, step t+1's move; the reactions, the exact
// corrections and the redone moves follow.  last: the launch's final step copies instead of moving.
void launch_step(rdy_engine *e, int handler, double dt, int last) {
    const uint32_t nr = std::max<uint32_t>(e->n_reps_used, 1);
    const int c = e->cur, n = 1 - c;
    const uint64_t nc = std::max<uint64_t>(e->cells_hwm, 1);
    const uint64_t nfb = e->fb ? 3 * e->topo_used : 0;
    const uint64_t nloc = topo_reactions_on(e) ? e->topo_used : 0;
    const uint64_t nthreads = std::max<uint64_t>(std::max<uint64_t>(std::max<uint64_t>(nr, (nc + 3) / 4), nfb), nloc);
    hipLaunchKernelGGL(k_step_begin, dim3(blocks_for(nthreads, 256)), dim3(256), 0, e->stream, e->reps, nr, last,
                       e->grid(n), (uint32_t)nc, e->fb, (uint32_t)nfb, e->edit_count, e->e_bonded,
                       (const uint32_t *)e->loc, e->loc_cur, (uint32_t)nloc);
    launch_bonded(e);
    const int h = (handler != RDY_HANDLER_NONE && e->n_rx > 0) ? handler : RDY_HANDLER_NONE;
    launch_walk(e, 0, h, dt, h != RDY_HANDLER_NONE);
    const bool topo_rx = topo_reactions_on(e);
    if (h != RDY_HANDLER_NONE || topo_rx) {
        if (h != RDY_HANDLER_NONE)
        hipLaunchKernelGGL(k_resolve_step, dim3(nr), dim3(RESOLVE_THREADS), 0, e->stream, e->grid(c), e->grid(n),
                           e->reps, e->T, e->topo(), (const BrickDev *)e->d_bricks, e->cap, e->radius, h, dt,
                           e->events, e->ev_count, slow_list(e, 0), e->rx_step, e->rx_total,
                           (uint32_t)e->max_particles, e->edits, e->edit_count, e->edit_cap, rec_ptr(e),
                           e->recording ? e->rec_count : nullptr, e->err);
        // SimulationLoop.h:286-287: the topology reactions follow the reactions, on their result
        if (topo_rx) launch_topo_reactions(e, 1, e->edits, e->edit_count);
        hipLaunchKernelGGL(k_fix_sites, dim3(FIX_BLOCKS), dim3(256), 0, e->stream, e->grid(c),
                           (const ReplicaDev *)e->reps, e->T, e->cap, e->radius, (const Edit *)e->edits,
                           (const uint32_t *)e->edit_count, e->edit_cap, e->aff_bits, e->aff_list, e->edit_count + 1,
                           e->aff_cap, e->wd_list, e->edit_count + 2, e->wd_cap, e->err);
        hipLaunchKernelGGL(k_fix_withdraw, dim3(FIX_BLOCKS), dim3(256), 0, e->stream, e->grid(c), e->grid(n),
                           (const ReplicaDev *)e->reps, e->cap, (const uint2 *)e->wd_list,
                           (const uint32_t *)(e->edit_count + 2), e->wd_cap);
        hipLaunchKernelGGL(k_fix_advance, dim3(FIX_BLOCKS), dim3(256), 0, e->stream, e->grid(c), e->grid(n),
                           (const ReplicaDev *)e->reps, e->T, e->topo(), e->cap, e->radius, dt, e->aff_bits,
                           (const uint32_t *)e->aff_list, (const uint32_t *)(e->edit_count + 1), e->aff_cap, e->err);
    }
    e->wl_valid[n] = false;
    e->cur = n;
}

void launch_forces(rdy_engine *e, double *force_out) {
    const uint32_t nr = std::max<uint32_t>(e->n_reps_used, 1);
    unsigned long long *vir = e->record_virial ? e->virial : nullptr;
    hipLaunchKernelGGL(k_advance, dim3(blocks_for(nr, 64)), dim3(64), 0, e->stream, e->reps, nr, 0, e->e_bonded,
                       (uint32_t *)nullptr, vir);
    launch_bonded(e);
    const int c = e->cur;
    for_brick_groups(e, [&](auto ah, uint32_t nb, uint32_t b0) {
        hipLaunchKernelGGL((k_brick<MODE_FORCES, decltype(ah)::value>), dim3(nb), dim3(BRICK_THREADS), e->brick_smem,
                           e->stream, e->grid(c), e->grid(c), e->reps, e->T, e->topo(), (const BrickDev *)e->d_bricks,
                           b0, e->cap, e->radius, 0.0, (const double *)nullptr, force_out, e->max_states, e->max_halo,
                           (uint32_t)e->brick_smem, slow_list(e, 0), e->err, vir);
    });
    hipLaunchKernelGGL(k_brick_slow<MODE_FORCES>, dim3(slow_blocks(e)), dim3(BRICK_THREADS), 0, e->stream,
                       e->grid(c), e->grid(c), e->reps, e->T, e->topo(), (const BrickDev *)e->d_bricks, e->cap,
                       e->radius, 0.0, (const double *)nullptr, force_out, slow_list(e, 0), e->err, vir);
    e->virial_valid = e->record_virial;
}

// force_out and the energies after a fused run are formed when first asked for.  The type tables and
// the brick tables (with the bin-reading kernel's brick list) exist only once an action has built them:
// an export or an integration straight after an upload builds them here first (the tables' time step
// only sets reaction probabilities, which a force evaluation does not read).
int ensure_export(rdy_engine *e) {
    if (e->export_valid) return RDY_OK;
    int rc = upload_tables(e, e->tables_dt > 0 ? e->tables_dt : 1.0);
    if (rc) return rc;
    if ((rc = ensure_bricks(e))) return rc;
    if ((rc = ensure_force_out(e))) return rc;
    launch_forces(e, e->force_out);
    e->export_valid = true;
    return RDY_OK;
}

int flush_prof(rdy_engine *e) {
    if (!e->prof) return RDY_OK;
    HIP_TRY(hipStreamSynchronize(e->stream));
    for (size_t i = 0; i < e->prof_used; ++i) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e->prof_ev[i].first, e->prof_ev[i].second));
        e->prof_ms += ms;
        e->prof_n++;
    }
    e->prof_used = 0;
    return RDY_OK;
}

}  // namespace

extern "C" {

const char *rdy_last_error(void) { return g_err.c_str(); }

int rdy_device_info(char *name, size_t name_len, uint64_t *hbm_bytes, int32_t *n_cus) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, dev));
    if (name && name_len) {  // the marketing name can be empty on some runtimes: report the CU count too
        std::snprintf(name, name_len, "%s (%s, %d CUs, %.0f GiB)", p.name[0] ? p.name : "AMD GPU", p.gcnArchName,
                      p.multiProcessorCount, (double)p.totalGlobalMem / (1024.0 * 1024.0 * 1024.0));
    }
    if (hbm_bytes) *hbm_bytes = p.totalGlobalMem;
    if (n_cus) *n_cus = p.multiProcessorCount;
    return RDY_OK;
}

int rdy_engine_create(const rdy_context_desc *ctx, const rdy_engine_config *cfg, rdy_engine **out) {
    if (!ctx || !out) return set_err(RDY_ERR_INVALID, "null argument");
    *out = nullptr;
    if (ctx->n_types == 0 || ctx->n_types > MAX_TYPES) return set_err(RDY_ERR_INVALID, "n_types out of range");
    if (!ctx->diffusion) return set_err(RDY_ERR_INVALID, "diffusion constants missing");
    for (uint32_t i = 0; i < ctx->n_potentials; ++i) {
        const rdy_pair_potential &p = ctx->potentials[i];
        if (p.type1 >= ctx->n_types || p.type2 >= ctx->n_types || p.kind > RDY_POT_SCREENED_ELECTROSTATICS)
            return set_err(RDY_ERR_INVALID, "invalid pair potential " + std::to_string(i));
        if (p.kind == RDY_POT_LENNARD_JONES && !(p.params[0] > p.params[1]))
            return set_err(RDY_ERR_INVALID, "When constructing the LJ potential, the first exponent m was not greater "
                                            "than the second exponent n");
        if (!(pot_cutoff(p) > 0)) return set_err(RDY_ERR_INVALID, "pair potential cutoff must be positive");
    }
    for (uint32_t i = 0; i < ctx->n_reactions; ++i) {
        const rdy_reaction &x = ctx->reactions[i];
        if (x.kind > RDY_RX_DECAY) return set_err(RDY_ERR_INVALID, "invalid reaction kind");
        int ne = is_order2(x.kind) ? 2 : 1;
        for (int k = 0; k < ne; ++k)
            if (x.educts[k] >= ctx->n_types) return set_err(RDY_ERR_INVALID, "reaction educt type out of range");
        int np = x.kind == RDY_RX_DECAY ? 0 : ((x.kind == RDY_RX_FISSION) ? 2 : 1);
        for (int k = 0; k < np; ++k)
            if (x.products[k] >= ctx->n_types) return set_err(RDY_ERR_INVALID, "reaction product type out of range");
        if (ctx->topology_flavor) {
            // products of these reactions would leave topologies inconsistent: topology reactions are out of scope
            bool bad = false;
            if (x.kind == RDY_RX_ENZYMATIC) bad = ctx->topology_flavor[x.educts[0]] || ctx->topology_flavor[x.products[0]];
            else {
                for (int k = 0; k < ne; ++k) bad = bad || ctx->topology_flavor[x.educts[k]];
                for (int k = 0; k < np; ++k) bad = bad || ctx->topology_flavor[x.products[k]];
            }
            if (bad)
                return set_err(RDY_ERR_INVALID, "reaction " + std::to_string(i) +
                                                    " consumes or creates topology particles (only an enzymatic "
                                                    "catalyst may be a topology species)");
        }
    }
    if (ctx->n_reactions > 255) return set_err(RDY_ERR_INVALID, "at most 255 reactions");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (int rw = init_noalloc_workspace(dev)) return rw;
    rdy_engine *e = new rdy_engine();
    e->device = dev;
    HIP_TRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    e->n_types = ctx->n_types;
    e->kbt = ctx->kbt;
    e->D.assign(ctx->diffusion, ctx->diffusion + ctx->n_types);
    e->topo_flavor.assign(ctx->n_types, 0);
    if (ctx->topology_flavor) e->topo_flavor.assign(ctx->topology_flavor, ctx->topology_flavor + ctx->n_types);
    e->pots_in.assign(ctx->potentials, ctx->potentials + ctx->n_potentials);
    e->rx_in.assign(ctx->reactions, ctx->reactions + ctx->n_reactions);
    e->n_rx = ctx->n_reactions;
    // Context::calculateMaxCutoff (Context.cpp:126-144)
    double maxc = 0.0;
    for (const auto &p : e->pots_in) maxc = std::max(maxc, pot_cutoff(p));
    for (const auto &x : e->rx_in)
        if (is_order2(x.kind)) maxc = std::max(maxc, x.educt_distance);
    rdy_engine_config c{};
    if (cfg) c = *cfg;
    e->nl_cutoff = c.nl_cutoff > 0 ? c.nl_cutoff : maxc;
    if (!(e->nl_cutoff > 0)) e->nl_cutoff = 1.0;  // no interactions: any grid works
    e->radius = c.cell_radius > 0 ? c.cell_radius : 1;
    if (e->radius > 3) {
        delete e;
        return set_err(RDY_ERR_INVALID, "cell_radius must be 1, 2 or 3");
    }
    e->cap = c.bin_capacity ? c.bin_capacity : 16;
    e->auto_cap = c.bin_capacity == 0;
    for (int a = 0; a < 3; ++a) e->brick[a] = c.brick[a];
    e->max_states = c.lds_states;
    e->block = c.block_size ? c.block_size : 256;
    if (e->block % 64 || e->block > 1024) {
        delete e;
        return set_err(RDY_ERR_INVALID, "block_size must be a multiple of 64 up to 1024");
    }
    e->max_reps = c.max_replicas ? c.max_replicas : 1;
    if (e->max_reps > MAX_REPLICAS) {
        delete e;
        return set_err(RDY_ERR_INVALID, "too many replicas");
    }
    e->max_cells = c.max_cells;
    e->max_particles = c.max_particles;
    e->ev_per_rep = c.events_per_replica ? std::min<uint32_t>(c.events_per_replica, RESOLVE_MAX_EVENTS)
                                         : (uint32_t)RESOLVE_MAX_EVENTS;
    *out = e;
    return RDY_OK;
}

int rdy_engine_destroy(rdy_engine *e) {
    if (!e) return RDY_OK;
    (void)hipStreamSynchronize(e->stream);
    for (int b = 0; b < 2; ++b) {
        if (e->gexec[b]) (void)hipGraphExecDestroy(e->gexec[b]);
        (void)hipFree(e->bins[b]);
        (void)hipFree(e->counts[b]);
        (void)hipFree(e->wl[b]);
        (void)hipFree(e->ovf_cell[b]);
    }
    void *ptrs[] = {e->wl_count, e->ovf_count, e->tile_sums, e->d_bricks, e->d_slow, e->reps, e->events, e->ev_count, e->recs, e->rec_count, e->rx_step, e->rx_total, e->e_bonded, e->virial, e->err,
                    e->force_out, e->edits, e->edit_count, e->aff_bits, e->aff_list, e->wd_list, e->d_pots, e->d_pidx, e->d_rx, e->d_o1idx, e->d_o1list, e->d_o2list, e->d_o2idx,
                    e->d_rowmax, e->d_dbpots, e->d_dbpidx, e->d_dbtable, e->d_dbkind, e->d_dbpref, e->d_dbradii, e->d_dbcdf, e->d_D, e->d_bd_scale, e->d_flavor, e->topo_of, e->topo_owner, e->treact, e->loc_cur, e->d_sr, e->ttype, e->tev, e->tev_count, e->tfus, e->tfus_count, e->tmrg, e->tmrg_count, e->adj_off, e->adj, e->str_step, e->str_total, e->fb, e->loc, e->scratch, e->d_noise, e->d_terms[0], e->d_terms[1],
                    e->d_terms[2]};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto &pe : e->prof_ev) {
        (void)hipEventDestroy(pe.first);
        (void)hipEventDestroy(pe.second);
    }
    (void)hipStreamDestroy(e->stream);
    delete e;
    return RDY_OK;
}

int rdy_replica_footprint(const rdy_engine *e, const rdy_replica_desc *d, uint64_t *bytes) {
    if (!e || !d || !bytes) return set_err(RDY_ERR_INVALID, "null argument");
    int32_t dims[3];
    double cs[3];
    cell_dims(e, d->box, dims, cs);
    uint64_t cells = (uint64_t)dims[0] * dims[1] * dims[2];
    *bytes = cells * (2ull * e->cap * sizeof(State) + 2ull * sizeof(uint32_t)) + d->max_particles * 2ull * 4ull +
             (uint64_t)e->ev_per_rep * sizeof(Event);
    return RDY_OK;
}

int rdy_replica_add(rdy_engine *e, const rdy_replica_desc *d, uint32_t *replica_id) {
    if (!e || !d || !replica_id) return set_err(RDY_ERR_INVALID, "null argument");
    for (int k = 0; k < 3; ++k)
        if (!(d->box[k] > 0)) return set_err(RDY_ERR_INVALID, "box size must be positive");
    int32_t dims[3];
    double cs[3];
    cell_dims(e, d->box, dims, cs);
    const uint64_t cells = (uint64_t)dims[0] * dims[1] * dims[2];
    if (!e->pools_ready) {
        // automatic pools: max_replicas copies of the first replica's grid and particle budget
        if (e->max_cells == 0) e->max_cells = cells * e->max_reps;
        if (e->max_particles == 0)
            e->max_particles = std::max<uint64_t>(4096, (2 * d->max_particles + 1024) * e->max_reps);
        const double lam_raw = (double)d->max_particles / (double)cells;
        if (e->auto_cap) {
            // mean occupancy of the first replica's grid: Poisson mean + 7 sigma + 6, a multiple of 4
            double lam = std::max(1.0, lam_raw);
            uint32_t c = (uint32_t)std::ceil(lam + 7.0 * std::sqrt(lam) + 6.0);
            e->cap = std::max<uint32_t>(8, (c + 3) / 4 * 4);
        }
        int rcc = configure_bricks(e, lam_raw, dims);
        if (rcc) return rcc;
        int rc = alloc_pools(e);
        if (rc) return rc;
    }
    uint32_t r = 0;
    while (r < e->max_reps && e->reps_h[r].in_use) ++r;
    if (r == e->max_reps) return set_err(RDY_ERR_CAPACITY, "replica table full");
    // first-fit cell range
    uint64_t base = 0;
    if (!take_range(e->free_cells, cells, &base)) return set_err(RDY_ERR_CAPACITY, "cell pool exhausted");
    HostReplica &H = e->reps_h[r];
    H = HostReplica{};
    H.in_use = true;
    H.cell_base = base;
    H.n_cells = cells;
    ReplicaDev &R = H.dev;
    for (int k = 0; k < 3; ++k) {
        R.box[k] = d->box[k];
        R.cs[k] = cs[k];
        R.pbc[k] = d->pbc[k] ? 1 : 0;
        R.dims[k] = dims[k];
    }
    R.cell_base = (uint32_t)base;
    R.n_cells = (uint32_t)cells;
    R.ev_base = r * e->ev_per_rep;
    R.ev_cap = e->ev_per_rep;
    R.handle_next = 0;
    R.step = 0xFFFFFFFFu;
    R.active = 1;
    R.in_use = 1;
    R.seed = d->seed;
    R.steps_left = ~0ull;
    R.steps_done = 0;
    R.handler = -1;
    HIP_TRY(hipMemcpyAsync(e->reps + r, &R, sizeof(ReplicaDev), hipMemcpyHostToDevice, e->stream));
    const size_t nrx = std::max<uint32_t>(e->n_rx, 1);
    HIP_TRY(hipMemsetAsync(e->rx_step + (size_t)r * nrx, 0, nrx * 8, e->stream));
    HIP_TRY(hipMemsetAsync(e->rx_total + (size_t)r * nrx, 0, nrx * 8, e->stream));
    if (e->rec_count)  // a reused slot holds no records of its previous replica
        HIP_TRY(hipMemsetAsync(e->rec_count + r, 0, sizeof(uint32_t), e->stream));
    e->n_reps_used = std::max(e->n_reps_used, r + 1);
    e->cells_hwm = std::max<uint64_t>(e->cells_hwm, base + cells);
    e->version++;
    e->bricks_dirty = true;  // uploaded before the next brick launch (adding many replicas stays linear)
    *replica_id = r;
    return RDY_OK;
}

int rdy_replica_remove(rdy_engine *e, uint32_t r) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    ensure_wl(e);
    hipLaunchKernelGGL(k_kill_replica, dim3(wl_blocks(e)), dim3(e->block), 0, e->stream, e->grid(e->cur), r,
                       (uint32_t)e->max_particles);
    uint32_t zero = 0;
    if ((rc = put_field(e, r, offsetof(ReplicaDev, in_use), zero))) return rc;
    if ((rc = put_field(e, r, offsetof(ReplicaDev, active), zero))) return rc;
    HostReplica &H = e->reps_h[r];
    for (int b = 0; b < 2; ++b)
        hipLaunchKernelGGL(k_release_cells, dim3(blocks_for(H.n_cells, 256)), dim3(256), 0, e->stream, e->grid(b),
                           (uint32_t)H.cell_base, (uint32_t)H.n_cells);
    release_range(e->free_cells, H.cell_base, H.n_cells);
    // the replica's bonded terms and topology slots go with it (a later replica may reuse slot r)
    if ((rc = drop_terms(e, r))) return rc;
    release_topo(e, H);
    H.in_use = false;
    e->version++;
    e->wl_valid[0] = e->wl_valid[1] = false;
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->bricks_dirty = true;
    return RDY_OK;
}

int rdy_set_particles(rdy_engine *e, uint32_t r, uint64_t n, const double *pos, const uint32_t *types) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    if (n && (!pos || !types)) return set_err(RDY_ERR_INVALID, "null particle arrays");
    uint32_t n_topo = 0;
    bool seen_free = false;
    for (uint64_t i = 0; i < n; ++i) {
        if (types[i] >= e->n_types) return set_err(RDY_ERR_INVALID, "particle type out of range");
        if (e->topo_flavor[types[i]]) {
            if (seen_free) return set_err(RDY_ERR_INVALID, "topology particles must precede free particles");
            ++n_topo;
        } else {
            seen_free = true;
        }
    }
    // replace: drop the replica's current particles (a replica that never held any has none to drop,
    // which keeps filling many fresh replicas linear in their particles)
    HostReplica &H = e->reps_h[r];
    if (H.has_particles) {
        ensure_wl(e);
        hipLaunchKernelGGL(k_kill_replica, dim3(wl_blocks(e)), dim3(e->block), 0, e->stream, e->grid(e->cur), r,
                           (uint32_t)e->max_particles);
    }
    H.has_particles = true;
    // the bonded terms name the replaced particles' handles: the caller sets them again
    if ((rc = drop_terms(e, r))) return rc;
    release_topo(e, H);
    // the topology handles: the upload's topology particles, then the reserve for fusions
    const uint32_t spare = n_topo ? H.topo_reserve : 0u;
    if ((uint64_t)n + spare > 0xFFFFFFF0ull) return set_err(RDY_ERR_INVALID, "particle handles beyond 32 bits");
    const uint32_t range = n_topo + spare;
    uint64_t tbase = 0;
    if (range && take_range(e->free_topo, range, &tbase)) {
        H.topo_base = (uint32_t)tbase;
        H.n_topo = n_topo;
        H.topo_range = range;
    } else if (range) {
        if (e->topo_used + range > e->topo_cap) {
            uint64_t ncap = std::max<uint64_t>(2 * e->topo_cap, e->topo_used + range);
            long long *fb;
            uint32_t *loc, *tof, *town, *lcur;
            HIP_TRY(dalloc(&fb, 3 * ncap));
            HIP_TRY(dalloc(&loc, ncap));
            HIP_TRY(dalloc(&lcur, ncap));
            HIP_TRY(dalloc(&tof, ncap));
            HIP_TRY(dalloc(&town, ncap));
            HIP_TRY(hipMemsetAsync(fb, 0, 3 * ncap * sizeof(long long), e->stream));
            HIP_TRY(hipMemsetAsync(tof, 0xFF, ncap * sizeof(uint32_t), e->stream));
            HIP_TRY(hipMemsetAsync(town, 0xFF, ncap * sizeof(uint32_t), e->stream));
            if (e->loc) {
                HIP_TRY(hipMemcpyAsync(loc, e->loc, e->topo_used * 4, hipMemcpyDeviceToDevice, e->stream));
                HIP_TRY(hipMemcpyAsync(tof, e->topo_of, e->topo_used * 4, hipMemcpyDeviceToDevice, e->stream));
                HIP_TRY(hipMemcpyAsync(town, e->topo_owner, e->topo_used * 4, hipMemcpyDeviceToDevice, e->stream));
                HIP_TRY(hipStreamSynchronize(e->stream));
                HIP_TRY(hipFree(e->loc));
                HIP_TRY(hipFree(e->fb));
                HIP_TRY(hipFree(e->topo_of));
                HIP_TRY(hipFree(e->topo_owner));
                HIP_TRY(hipFree(e->loc_cur));
            }
            e->loc_cur = lcur;
            e->fb = fb;
            e->loc = loc;
            e->topo_of = tof;
            e->topo_owner = town;
            e->topo_cap = ncap;
        }
        H.topo_base = (uint32_t)e->topo_used;
        H.n_topo = n_topo;
        H.topo_range = range;
        e->topo_used += range;
    } else {
        H.topo_base = 0;
        H.n_topo = 0;
        H.topo_range = 0;
    }
    if (H.topo_range) {  // the new range belongs to replica r; its particles have no topology yet
        std::vector<uint32_t> own(H.topo_range, r), none(H.topo_range, 0xFFFFFFFFu);
        HIP_TRY(hipMemcpyAsync(e->topo_owner + H.topo_base, own.data(), H.topo_range * 4, hipMemcpyHostToDevice,
                               e->stream));
        HIP_TRY(hipMemcpyAsync(e->topo_of + H.topo_base, none.data(), H.topo_range * 4, hipMemcpyHostToDevice,
                               e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    if ((rc = put_field(e, r, offsetof(ReplicaDev, tt_base), H.tt_base))) return rc;
    if ((rc = put_field(e, r, offsetof(ReplicaDev, n_tt), H.n_tt))) return rc;
    uint32_t hn = (uint32_t)n + spare;
    if ((rc = put_field(e, r, offsetof(ReplicaDev, topo_base), H.topo_base))) return rc;
    if ((rc = put_field(e, r, offsetof(ReplicaDev, n_topo), H.n_topo))) return rc;
    if ((rc = put_field(e, r, offsetof(ReplicaDev, topo_range), H.topo_range))) return rc;
    if ((rc = put_field(e, r, offsetof(ReplicaDev, handle_next), hn))) return rc;
    if (n) {
        size_t bytes = n * (3 * sizeof(double) + sizeof(uint32_t));
        if ((rc = ensure_scratch(e, bytes))) return rc;
        double *dpos = (double *)e->scratch;
        uint32_t *dty = (uint32_t *)(dpos + 3 * n);
        HIP_TRY(hipMemcpyAsync(dpos, pos, 3 * n * sizeof(double), hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipMemcpyAsync(dty, types, n * sizeof(uint32_t), hipMemcpyHostToDevice, e->stream));
        if (!e->d_flavor) {
            HIP_TRY(dalloc(&e->d_flavor, e->n_types));
            HIP_TRY(hipMemcpyAsync(e->d_flavor, e->topo_flavor.data(), e->n_types, hipMemcpyHostToDevice, e->stream));
        }
        hipLaunchKernelGGL(k_upload, dim3(blocks_for(n, 256)), dim3(256), 0, e->stream, e->grid(e->cur), e->reps, r,
                           (uint32_t)n, n_topo, spare, dpos, dty, (const uint8_t *)e->d_flavor, e->topo(), e->cap,
                           (uint32_t)e->max_particles, e->err);
    }
    e->wl_valid[e->cur] = false;
    e->version++;
    return check_errors(e);
}

uint64_t rdy_particle_count(rdy_engine *e, uint32_t r) {
    uint64_t n = 0;
    if (rdy_get_particles(e, r, 0, nullptr, nullptr, nullptr, &n)) return 0;
    return n;
}

static int gather(rdy_engine *e, uint32_t r, uint64_t cap, double *pos, uint32_t *types, uint32_t *handles,
                  double *forces, uint64_t *n) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    const uint64_t ncap = e->max_particles;
    size_t bytes = 64 + ncap * (6 * sizeof(double) + 2 * sizeof(uint32_t));
    if ((rc = ensure_scratch(e, bytes))) return rc;
    uint32_t *cnt = (uint32_t *)e->scratch;
    double *dpos = (double *)((char *)e->scratch + 64);
    double *dfor = dpos + 3 * ncap;
    uint32_t *dty = (uint32_t *)(dfor + 3 * ncap);
    uint32_t *dh = dty + ncap;
    HIP_TRY(hipMemsetAsync(cnt, 0, 4, e->stream));
    ensure_wl(e);
    hipLaunchKernelGGL(k_gather, dim3(wl_blocks(e)), dim3(e->block), 0, e->stream, e->grid(e->cur), r, cnt,
                       (uint32_t)ncap, dpos, dty, dh, (const double *)e->force_out, forces ? dfor : (double *)nullptr,
                       (uint32_t)e->max_particles);
    uint32_t hc = 0;
    HIP_TRY(hipMemcpyAsync(&hc, cnt, 4, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    *n = hc;
    uint64_t m = std::min<uint64_t>(hc, cap);
    if (m) {
        if (pos) HIP_TRY(hipMemcpyAsync(pos, dpos, m * 3 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        if (forces) HIP_TRY(hipMemcpyAsync(forces, dfor, m * 3 * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        if (types) HIP_TRY(hipMemcpyAsync(types, dty, m * sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
        if (handles) HIP_TRY(hipMemcpyAsync(handles, dh, m * sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    return RDY_OK;
}

int rdy_get_particles(rdy_engine *e, uint32_t r, uint64_t cap, double *pos, uint32_t *types, uint32_t *handles,
                      uint64_t *n) {
    uint64_t dummy;
    return gather(e, r, cap, pos, types, handles, nullptr, n ? n : &dummy);
}

int rdy_get_forces(rdy_engine *e, uint32_t r, uint64_t cap, double *forces, uint32_t *handles, uint64_t *n) {
    if (!e) return set_err(RDY_ERR_STATE, "no forces have been calculated");
    int rc = ensure_export(e);
    if (rc) return rc;
    if (!e->force_out) return set_err(RDY_ERR_STATE, "no forces have been calculated");
    uint64_t dummy;
    return gather(e, r, cap, nullptr, nullptr, handles, forces, n ? n : &dummy);
}

static int set_terms(rdy_engine *e, uint32_t r, int kind, uint64_t n, const uint32_t *idx, int arity, const double *k,
                     const double *p0, const uint32_t *mult) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    if (n && (!idx || !k || !p0)) return set_err(RDY_ERR_INVALID, "null bonded term arrays");
    if ((rc = sync_topo_count(e, r))) return rc;
    const HostReplica &H = e->reps_h[r];
    // every index is checked before anything changes: a rejected call leaves the terms as they were
    std::vector<BondTerm> fresh;
    fresh.reserve(n);
    for (uint64_t i = 0; i < n; ++i) {
        BondTerm t{};
        t.r = r;
        for (int a = 0; a < arity; ++a) {
            t.h[a] = idx[arity * i + a];
            if (t.h[a] >= H.n_topo) return set_err(RDY_ERR_INVALID, "bonded term index is not a topology particle");
        }
        t.k = k[i];
        t.p0 = p0[i];
        t.mult = mult ? mult[i] : 0;
        fresh.push_back(t);
    }
    if (kind == 0) {
        // a bond is an unordered edge: (min, max) handles, and every configuration of one edge in one
        // consecutive run (BreakBonds sums an edge's configurations, Actions.h:291-341); the harmonic
        // bond's force and energy do not depend on the order of its ends
        for (BondTerm &t : fresh)
            if (t.h[0] > t.h[1]) std::swap(t.h[0], t.h[1]);
        std::stable_sort(fresh.begin(), fresh.end(), [](const BondTerm &a, const BondTerm &b) {
            return a.h[0] != b.h[0] ? a.h[0] < b.h[0] : a.h[1] < b.h[1];
        });
    }
    std::vector<BondTerm> &v = e->terms_h[kind];
    v.erase(std::remove_if(v.begin(), v.end(), [r](const BondTerm &t) { return t.r == r; }), v.end());
    v.insert(v.end(), fresh.begin(), fresh.end());
    return upload_terms(e, kind);
}

int rdy_set_bonds(rdy_engine *e, uint32_t r, uint64_t n, const uint32_t *idx2, const double *k, const double *length) {
    return set_terms(e, r, 0, n, idx2, 2, k, length, nullptr);
}

int rdy_set_angles(rdy_engine *e, uint32_t r, uint64_t n, const uint32_t *idx3, const double *k,
                   const double *theta0) {
    return set_terms(e, r, 1, n, idx3, 3, k, theta0, nullptr);
}

int rdy_set_dihedrals(rdy_engine *e, uint32_t r, uint64_t n, const uint32_t *idx4, const double *k,
                      const uint32_t *multiplicity, const double *phi0) {
    return set_terms(e, r, 2, n, idx4, 4, k, phi0, multiplicity);
}

int rdy_update_neighbor_list(rdy_engine *e) {
    if (!e || !e->pools_ready) return set_err(RDY_ERR_STATE, "engine has no replicas");
    if (int rb = ensure_bricks(e)) return rb;
    int rc = upload_tables(e, e->tables_dt > 0 ? e->tables_dt : 1.0);
    if (rc) return rc;
    const int c = e->cur, n = 1 - c;
    hipLaunchKernelGGL((k_brick<MODE_REBIN, false>), dim3(n_bricks(e)), dim3(BRICK_THREADS), 0, e->stream,
                       e->grid(c), e->grid(n), e->reps, e->T, e->topo(), (const BrickDev *)e->d_bricks, 0u, e->cap,
                       e->radius, 0.0, (const double *)nullptr, (double *)nullptr, e->max_states, e->max_halo, 0u,
                       slow_list(e, 0), e->err, (unsigned long long *)nullptr);
    launch_clear_bricks(e, c, n);
    e->wl_valid[n] = false;
    e->cur = n;
    return check_errors(e);
}

int rdy_calculate_forces(rdy_engine *e) {
    if (!e || !e->pools_ready) return set_err(RDY_ERR_STATE, "engine has no replicas");
    if (int rb = ensure_bricks(e)) return rb;
    int rc = upload_tables(e, e->tables_dt > 0 ? e->tables_dt : 1.0);
    if (rc) return rc;
    if ((rc = ensure_force_out(e))) return rc;
    launch_forces(e, e->force_out);
    e->export_valid = true;
    return check_errors(e);
}


int rdy_get_energy(rdy_engine *e, uint32_t r, double *energy) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    if ((rc = ensure_export(e))) return rc;
    unsigned long long pe = 0, be = 0;
    if ((rc = get_field(e, r, offsetof(ReplicaDev, energy), &pe))) return rc;
    HIP_TRY(hipMemcpyAsync(&be, e->e_bonded + r, 8, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    *energy = 0.5 * (double)(long long)pe * ENERGY_INV + (double)(long long)be * ENERGY_INV;
    return RDY_OK;
}

int rdy_record_virial(rdy_engine *e, int on) {
    if (!e) return set_err(RDY_ERR_INVALID, "null engine");
    e->record_virial = on != 0;
    if (e->record_virial && !e->virial_valid) e->export_valid = false;  // formed with the next export
    return RDY_OK;
}

int rdy_get_virial(rdy_engine *e, uint32_t r, double *virial9) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    if (!virial9) return set_err(RDY_ERR_INVALID, "null virial array");
    if (!e->record_virial) return set_err(RDY_ERR_STATE, "the virial is not being recorded (rdy_record_virial)");
    if ((rc = ensure_export(e))) return rc;
    if (!e->virial_valid) {  // switched on after the last force evaluation of a separate-action sequence
        launch_forces(e, e->force_out);
        if ((rc = check_errors(e))) return rc;
    }
    unsigned long long v[9];
    HIP_TRY(hipMemcpyAsync(v, e->virial + 9ull * r, sizeof v, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    for (int c = 0; c < 9; ++c) virial9[c] = 0.5 * (double)(long long)v[c] * ENERGY_INV;
    return RDY_OK;
}

int rdy_integrate(rdy_engine *e, double dt, const double *noise, uint64_t noise_handles) {
    if (!e || !e->pools_ready) return set_err(RDY_ERR_STATE, "engine has no replicas");
    int rc = upload_tables(e, dt);
    if (rc) return rc;
    const double *dn = nullptr;
    if (noise) {
        if (e->n_reps_used != 1) return set_err(RDY_ERR_INVALID, "injected noise needs a single-replica engine");
        uint32_t hn = 0;
        if ((rc = get_field(e, 0, offsetof(ReplicaDev, handle_next), &hn))) return rc;
        if (noise_handles < hn) return set_err(RDY_ERR_INVALID, "noise buffer shorter than the particle handles");
        if (noise_handles * 3 > e->noise_cap) {
            HIP_TRY(hipStreamSynchronize(e->stream));
            if (e->d_noise) HIP_TRY(hipFree(e->d_noise));
            e->noise_cap = noise_handles * 3;
            HIP_TRY(dalloc(&e->d_noise, e->noise_cap));
        }
        HIP_TRY(hipMemcpyAsync(e->d_noise, noise, noise_handles * 3 * sizeof(double), hipMemcpyHostToDevice, e->stream));
        dn = e->d_noise;
    }
    const uint32_t nr = std::max<uint32_t>(e->n_reps_used, 1);
    if ((rc = ensure_export(e))) return rc;
    ensure_wl(e);
    hipLaunchKernelGGL(k_advance, dim3(blocks_for(nr, 64)), dim3(64), 0, e->stream, e->reps, nr, 1, e->e_bonded,
                       (uint32_t *)nullptr, (unsigned long long *)nullptr);
    hipLaunchKernelGGL(k_integrate, dim3(wl_blocks(e)), dim3(e->block), 0, e->stream, e->grid(e->cur), e->reps, e->T,
                       e->cap, dt, dn, (const double *)e->force_out, (uint32_t)e->max_particles, e->err);
    return check_errors(e);
}

// the reactions and potentials of a context as the reversible-pair search reads them (weights normalised
// as Fusion.h:61-68 and Fission.h:63-70 do)
static void rev_inputs(const rdy_reaction *rx, uint32_t n_rx, const rdy_pair_potential *pots, uint32_t n_pots,
                       std::vector<rdy_rev::Reaction> &R, std::vector<rdy_rev::PairPotential> &P) {
    R.clear();
    P.clear();
    for (uint32_t i = 0; i < n_rx; ++i) {
        const rdy_reaction &x = rx[i];
        rdy_rev::Reaction r{};
        r.kind = (int)x.kind;
        for (int j = 0; j < 2; ++j) {
            r.educts[j] = j < r.n_educts() ? x.educts[j] : 0u;
            r.products[j] = j < r.n_products() ? x.products[j] : 0u;
        }
        r.rate = x.rate;
        r.educt_distance = x.educt_distance;
        r.product_distance = x.product_distance;
        double w1 = x.weight1, w2 = x.weight2;
        if (x.kind == RDY_RX_FUSION || x.kind == RDY_RX_FISSION) {
            const double sum = w1 + w2;
            if (sum != 1) {
                w1 /= sum;
                w2 /= sum;
            }
        }
        r.weight1 = w1;
        r.weight2 = w2;
        R.push_back(r);
    }
    for (uint32_t i = 0; i < n_pots; ++i) {
        rdy_rev::PairPotential q{};
        q.kind = (int)pots[i].kind;
        q.type1 = pots[i].type1;
        q.type2 = pots[i].type2;
        std::memcpy(q.params, pots[i].params, sizeof q.params);
        P.push_back(q);
    }
}

int rdy_reversible_reactions(const rdy_context_desc *ctx, double total_volume, uint64_t cap,
                             rdy_reversible_reaction *out, uint64_t *n) {
    if (!ctx || !n) return set_err(RDY_ERR_INVALID, "null argument");
    if ((ctx->n_reactions && !ctx->reactions) || (ctx->n_potentials && !ctx->potentials))
        return set_err(RDY_ERR_INVALID, "null reaction or potential list");
    *n = 0;
    std::vector<rdy_rev::Reaction> R;
    std::vector<rdy_rev::PairPotential> P;
    rev_inputs(ctx->reactions, ctx->n_reactions, ctx->potentials, ctx->n_potentials, R, P);
    std::vector<rdy_rev::Config> cfg;
    try {
        cfg = rdy_rev::search(R, P, ctx->kbt, total_volume);
    } catch (const std::exception &x) {
        return set_err(RDY_ERR_INVALID, x.what());
    }
    for (size_t i = 0; i < cfg.size() && i < cap && out; ++i) {
        const rdy_rev::Config &c = cfg[i];
        rdy_reversible_reaction &o = out[i];
        o.kind = (int32_t)c.kind;
        o.forward = c.forward;
        o.backward = c.backward;
        o.n_lhs = (uint32_t)c.n_lhs;
        o.n_rhs = (uint32_t)c.n_rhs;
        for (int j = 0; j < 2; ++j) {
            o.lhs[j] = c.lhs[j];
            o.rhs[j] = c.rhs[j];
        }
        o.reaction_radius = c.reaction_radius;
        o.lhs_interaction_radius = c.lhs_radius;
        o.rhs_interaction_radius = c.rhs_radius;
        o.lhs_interaction_volume = c.lhs_volume;
        o.rhs_interaction_volume = c.rhs_volume;
        o.eff_lhs_interaction_volume = c.eff_lhs_interaction;
        o.eff_rhs_interaction_volume = c.eff_rhs_interaction;
        o.eff_lhs_reaction_volume = c.eff_lhs_reaction;
        o.eff_rhs_reaction_volume = c.eff_rhs_reaction;
        o.acceptance_prefactor = c.acceptance_prefactor;
        o.total_volume = c.total_volume;
        o.kbt = c.kbt;
        o.equilibrium_constant = c.equilibrium_constant;
        o.macro_forward_rate = c.macro_forward;
        o.macro_backward_rate = c.macro_backward;
        o.micro_forward_rate = c.micro_forward;
        o.micro_backward_rate = c.micro_backward;
    }
    *n = cfg.size();
    return RDY_OK;
}

// DetailedBalance device tables: FP64 pair potentials per ordered type pair, and per reaction its
// reversible pair's kind, Metropolis prefactor and fission-distance table
static int ensure_db_tables(rdy_engine *e) {
    if (e->db_ready) return RDY_OK;
    std::vector<rdy_rev::Reaction> R;
    std::vector<rdy_rev::PairPotential> P;
    rev_inputs(e->rx_in.data(), e->n_rx, e->pots_in.data(), (uint32_t)e->pots_in.size(), R, P);
    std::vector<rdy_rev::Config> cfg;
    try {
        cfg = rdy_rev::search(R, P, e->kbt, 0.0);  // the volume enters only the inferred macro rates
    } catch (const std::exception &x) {
        return set_err(RDY_ERR_INVALID, x.what());
    }
    const uint32_t nt = e->n_types;
    std::vector<std::vector<uint32_t>> lists((size_t)nt * nt);
    for (uint32_t k = 0; k < (uint32_t)P.size(); ++k) {
        lists[(size_t)P[k].type1 * nt + P[k].type2].push_back(k);
        if (P[k].type1 != P[k].type2) lists[(size_t)P[k].type2 * nt + P[k].type1].push_back(k);
    }
    std::vector<DbPot> pots;
    std::vector<uint32_t> pidx((size_t)nt * nt);
    for (size_t k = 0; k < lists.size(); ++k) {
        pidx[k] = (uint32_t)pots.size() | ((uint32_t)lists[k].size() << 16);
        for (uint32_t id : lists[k]) {
            DbPot d{};
            const double c = rdy_rev::cutoff_of(P[id]);
            d.cut2 = c * c;
            std::memcpy(d.q, P[id].params, sizeof d.q);
            d.kind = P[id].kind;
            pots.push_back(d);
        }
    }
    if (pots.size() > 65535) return set_err(RDY_ERR_INVALID, "too many pair potentials");
    std::vector<int32_t> kind(std::max<uint32_t>(e->n_rx, 1), -1);
    std::vector<double> pref(kind.size(), 1.0), radii, cdf;
    std::vector<uint32_t> table(kind.size(), 0u);
    for (const rdy_rev::Config &c : cfg) {
        uint32_t t0 = 0;
        if (c.kind == rdy_rev::FUSION_FISSION) {
            t0 = (uint32_t)radii.size();
            radii.insert(radii.end(), c.fission_radii.begin(), c.fission_radii.end());
            cdf.insert(cdf.end(), c.fission_cdf.begin(), c.fission_cdf.end());
        }
        for (uint32_t rx : {c.forward, c.backward}) {
            kind[rx] = (int32_t)c.kind;
            table[rx] = t0;
            // SCPUReactionImpls.cpp:482-489: the prefactor for educts == lhs types, in order; the
            // backward branch repeats the same test and is never taken
            if (c.kind == rdy_rev::ENZYMATIC_ENZYMATIC && R[rx].educts[0] == c.lhs[0] && R[rx].educts[1] == c.lhs[1])
                pref[rx] = c.acceptance_prefactor;
        }
    }
    if (radii.empty()) {
        radii.push_back(0.0);
        cdf.push_back(1.0);
    }
    auto up = [&](auto **dst, const auto &v) -> int {
        using T = std::remove_pointer_t<std::remove_reference_t<decltype(*dst)>>;
        if (*dst) HIP_TRY(hipFree(*dst));
        HIP_TRY(dalloc(dst, std::max<size_t>(v.size(), 1)));
        if (!v.empty()) HIP_TRY(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return RDY_OK;
    };
    int rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    if ((rc = up(&e->d_dbpots, pots))) return rc;
    if ((rc = up(&e->d_dbpidx, pidx))) return rc;
    if ((rc = up(&e->d_dbkind, kind))) return rc;
    if ((rc = up(&e->d_dbpref, pref))) return rc;
    if ((rc = up(&e->d_dbtable, table))) return rc;
    if ((rc = up(&e->d_dbradii, radii))) return rc;
    if ((rc = up(&e->d_dbcdf, cdf))) return rc;
    e->db.pots = e->d_dbpots;
    e->db.pidx = e->d_dbpidx;
    e->db.kind = e->d_dbkind;
    e->db.prefactor = e->d_dbpref;
    e->db.table = e->d_dbtable;
    e->db.radii = e->d_dbradii;
    e->db.cdf = e->d_dbcdf;
    e->db.n_table = (uint32_t)(rdy_rev::FISSION_POINTS + 1);
    e->db.radius = e->radius;
    e->db_ready = true;
    return RDY_OK;
}

int rdy_react(rdy_engine *e, int handler, double dt) {
    if (!e || !e->pools_ready) return set_err(RDY_ERR_STATE, "engine has no replicas");
    if (int rb = ensure_bricks(e)) return rb;
    if (handler != RDY_HANDLER_UNCONTROLLED_APPROXIMATION && handler != RDY_HANDLER_GILLESPIE &&
        handler != RDY_HANDLER_DETAILED_BALANCE)
        return set_err(RDY_ERR_INVALID, "unknown reaction handler");
    int rc = upload_tables(e, dt);
    if (rc) return rc;
    if (e->n_rx == 0) return RDY_OK;
    const bool db = handler == RDY_HANDLER_DETAILED_BALANCE;
    if (db && (rc = ensure_db_tables(e))) return rc;
    const uint32_t nr = std::max<uint32_t>(e->n_reps_used, 1);
    const int c = e->cur;
    for_brick_groups(e, [&](auto ah, uint32_t nb, uint32_t b0) {
        hipLaunchKernelGGL((k_brick_search<decltype(ah)::value>), dim3(nb), dim3(BRICK_THREADS), e->brick_smem,
                           e->stream, e->grid(c), e->grid(c), 0, e->reps, e->T, (const BrickDev *)e->d_bricks, b0,
                           e->cap, e->radius, handler, e->events, e->ev_count, e->max_states, e->max_halo,
                           (uint32_t)e->brick_smem, slow_list(e, 1), e->err);
    });
    hipLaunchKernelGGL(k_brick_search_slow, dim3(slow_blocks(e)), dim3(BRICK_THREADS), 0, e->stream, e->grid(c),
                       e->reps, e->T, (const BrickDev *)e->d_bricks, e->cap, e->radius, handler, e->events,
                       e->ev_count, slow_list(e, 1), e->err);
    if (db)
        hipLaunchKernelGGL(k_detailed_balance, dim3(nr), dim3(RESOLVE_THREADS), 0, e->stream, e->grid(c), e->reps,
                           e->T, e->cap, e->events, e->ev_count, e->rx_step, e->rx_total, (uint32_t)e->max_particles,
                           rec_ptr(e), e->recording ? e->rec_count : nullptr, e->err, e->db);
    else
        hipLaunchKernelGGL(k_resolve, dim3(nr), dim3(RESOLVE_THREADS), 0, e->stream, e->grid(c), e->reps, e->T,
                           e->cap, e->events, e->ev_count, e->rx_step, e->rx_total, 0, (uint32_t)e->max_particles,
                           (Edit *)nullptr, (uint32_t *)nullptr, 0u, rec_ptr(e),
                           e->recording ? e->rec_count : nullptr, e->err);
    return check_errors(e);
}

int rdy_get_reaction_counts(rdy_engine *e, uint32_t r, uint64_t *counts) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    if (e->n_rx == 0) return RDY_OK;
    HIP_TRY(hipMemcpyAsync(counts, e->rx_step + (size_t)r * e->n_rx, e->n_rx * 8, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return RDY_OK;
}

int rdy_record_reactions(rdy_engine *e, int on) {
    if (!e || !e->pools_ready) return set_err(RDY_ERR_STATE, "engine has no replicas");
    if (on && !e->recs) {
        HIP_TRY(dalloc(&e->recs, (size_t)e->max_reps * RESOLVE_MAX_EVENTS));
        HIP_TRY(dalloc(&e->rec_count, e->max_reps));
        HIP_TRY(hipMemsetAsync(e->rec_count, 0, e->max_reps * sizeof(uint32_t), e->stream));
    }
    e->recording = on != 0;
    e->version++;  // captured steps carry the record pointers
    return RDY_OK;
}

int rdy_get_reaction_records(rdy_engine *e, uint32_t r, uint64_t cap, rdy_reaction_record *out, uint64_t *n) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    if (!n) return set_err(RDY_ERR_INVALID, "n is NULL");
    *n = 0;
    if (!e->recs || !e->recording) return set_err(RDY_ERR_STATE, "reaction records are not being kept");
    uint32_t ne = 0;
    HIP_TRY(hipMemcpyAsync(&ne, e->rec_count + r, sizeof ne, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    ne = std::min<uint32_t>(ne, RESOLVE_MAX_EVENTS);
    std::vector<RxRecord> h(ne);
    if (ne) {
        HIP_TRY(hipMemcpyAsync(h.data(), e->recs + (size_t)r * RESOLVE_MAX_EVENTS, ne * sizeof(RxRecord),
                               hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    uint64_t k = 0;
    for (const RxRecord &x : h) {
        if (!x.valid) continue;
        if (k < cap && out) {
            rdy_reaction_record &o = out[k];
            o.reaction = x.rx;
            o.type = (int32_t)x.kind;
            for (int j = 0; j < 2; ++j) {
                o.educts[j] = x.educts[j];
                o.products[j] = x.products[j];
                o.types_from[j] = x.types_from[j];
            }
            for (int d = 0; d < 3; ++d) o.position[d] = x.where[d];
        }
        ++k;
    }
    *n = k;
    return RDY_OK;
}

int rdy_get_reaction_totals(rdy_engine *e, uint32_t r, uint64_t *counts) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    if (e->n_rx == 0) return RDY_OK;
    HIP_TRY(hipMemcpyAsync(counts, e->rx_total + (size_t)r * e->n_rx, e->n_rx * 8, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return RDY_OK;
}

int rdy_run(rdy_engine *e, int handler, double dt, uint64_t n_steps) {
    if (!e || !e->pools_ready) return set_err(RDY_ERR_STATE, "engine has no replicas");
    if (e->in_session) return set_err(RDY_ERR_STATE, "a run is open: finish it with rdy_run_end");
    if (int rb = ensure_bricks(e)) return rb;
    if (handler < 0 || handler > RDY_HANDLER_NONE) return set_err(RDY_ERR_INVALID, "unknown reaction handler");
    int rc = upload_tables(e, dt);
    if (rc) return rc;
    if (n_steps == 0) return RDY_OK;
    if ((rc = ensure_step_pools(e))) return rc;
    e->export_valid = false;
    launch_prologue(e, dt);
    for (uint64_t s = 0; s < n_steps; ++s) {
        launch_step(e, handler, dt, s + 1 == n_steps);
        if (e->prof && e->prof_used == e->prof_ev.size())
            if ((rc = flush_prof(e))) return rc;
        if ((s & 1023) == 1023)
            if ((rc = check_errors(e))) return rc;
    }
    launch_epilogue(e);
    if ((rc = flush_prof(e))) return rc;
    return check_errors(e);
}

int rdy_run_begin(rdy_engine *e, int handler, double dt, uint64_t n_steps) {
    if (!e || !e->pools_ready) return set_err(RDY_ERR_STATE, "engine has no replicas");
    if (e->in_session) return set_err(RDY_ERR_STATE, "a run is open: finish it with rdy_run_end");
    if (int rb = ensure_bricks(e)) return rb;
    if (handler < 0 || handler > RDY_HANDLER_NONE) return set_err(RDY_ERR_INVALID, "unknown reaction handler");
    if (n_steps == 0) return set_err(RDY_ERR_INVALID, "a run needs at least one step");
    int rc = upload_tables(e, dt);
    if (rc) return rc;
    if ((rc = ensure_step_pools(e))) return rc;
    e->export_valid = false;
    launch_prologue(e, dt);
    e->in_session = true;
    e->session_left = n_steps;
    e->session_handler = handler;
    e->session_dt = dt;
    return RDY_OK;
}

int rdy_run_enqueue(rdy_engine *e, uint64_t k) {
    if (!e || !e->in_session) return set_err(RDY_ERR_STATE, "no open run (rdy_run_begin)");
    if (k > e->session_left) return set_err(RDY_ERR_INVALID, "more steps than the run has left");
    for (uint64_t s = 0; s < k; ++s) {
        const bool last = --e->session_left == 0;
        launch_step(e, e->session_handler, e->session_dt, last ? 1 : 0);
        if (last) launch_epilogue(e);
    }
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return set_err(RDY_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(le));
    return RDY_OK;
}

int rdy_run_end(rdy_engine *e) {
    if (!e || !e->in_session) return set_err(RDY_ERR_STATE, "no open run (rdy_run_begin)");
    if (e->session_left) return set_err(RDY_ERR_STATE, "the run has steps left to enqueue");
    e->in_session = false;
    return check_errors(e);
}

int rdy_run_graph(rdy_engine *e, int handler, double dt, uint64_t n_steps) {
    if (!e || !e->pools_ready) return set_err(RDY_ERR_STATE, "engine has no replicas");
    if (int rb = ensure_bricks(e)) return rb;
    if (handler < 0 || handler > RDY_HANDLER_NONE) return set_err(RDY_ERR_INVALID, "unknown reaction handler");
    if (e->in_session) return set_err(RDY_ERR_STATE, "a run is open: finish it with rdy_run_end");
    int rc = upload_tables(e, dt);
    if (rc) return rc;
    if (n_steps == 0) return RDY_OK;
    if ((rc = ensure_step_pools(e))) return rc;
    e->export_valid = false;
    launch_prologue(e, dt);
    const bool prof = e->prof;
    e->prof = false;
    if (e->g_version != e->version || e->g_handler != handler || e->g_dt != dt) {
        for (int b = 0; b < 2; ++b) {
            if (e->gexec[b]) HIP_TRY(hipGraphExecDestroy(e->gexec[b]));
            e->gexec[b] = nullptr;
        }
        // capture two steps per parity so that every replay starts and ends on the same buffer
        const int keep = e->cur;
        for (int b = 0; b < 2; ++b) {
            e->cur = b;
            hipGraph_t g;
            HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            launch_step(e, handler, dt, 0);
            launch_step(e, handler, dt, 0);
            HIP_TRY(hipStreamEndCapture(e->stream, &g));
            HIP_TRY(hipGraphInstantiate(&e->gexec[b], g, nullptr, nullptr, 0));
            HIP_TRY(hipGraphDestroy(g));
        }
        e->cur = keep;
        e->g_version = e->version;
        e->g_handler = handler;
        e->g_dt = dt;
    }
    // every step but the last from the graphs (the last one copies instead of moving)
    const uint64_t generic = n_steps - 1, pairs = generic / 2;
    for (uint64_t p = 0; p < pairs; ++p) {
        HIP_TRY(hipGraphLaunch(e->gexec[e->cur], e->stream));
        if ((p & 511) == 511)
            if ((rc = check_errors(e))) return rc;
    }
    if (generic & 1) launch_step(e, handler, dt, 0);
    launch_step(e, handler, dt, 1);
    launch_epilogue(e);
    e->wl_valid[0] = e->wl_valid[1] = false;  // replays bypass launch_step's bookkeeping
    e->prof = prof;
    return check_errors(e);
}

int rdy_replica_set_steps(rdy_engine *e, uint32_t r, uint64_t steps) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    return put_field(e, r, offsetof(ReplicaDev, steps_left), steps);
}

int rdy_replica_steps_done(rdy_engine *e, uint32_t r, uint64_t *steps) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    return get_field(e, r, offsetof(ReplicaDev, steps_done), steps);
}

int rdy_replica_set_handler(rdy_engine *e, uint32_t r, int handler) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    if (handler < -1 || handler > RDY_HANDLER_NONE) return set_err(RDY_ERR_INVALID, "unknown reaction handler");
    const int32_t h = handler;
    return put_field(e, r, offsetof(ReplicaDev, handler), h);
}

int rdy_synchronize(rdy_engine *e) {
    if (!e) return set_err(RDY_ERR_INVALID, "null engine");
    HIP_TRY(hipStreamSynchronize(e->stream));
    return check_errors(e);
}

int rdy_neighbor_pairs(rdy_engine *e, uint32_t r, double cutoff, uint64_t cap, uint32_t *pairs, uint64_t *n) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    size_t bytes = 64 + 2 * cap * sizeof(uint32_t);
    if ((rc = ensure_scratch(e, bytes))) return rc;
    uint32_t *cnt = (uint32_t *)e->scratch;
    uint32_t *out = cnt + 16;
    HIP_TRY(hipMemsetAsync(cnt, 0, 4, e->stream));
    ensure_wl(e);
    hipLaunchKernelGGL(k_pairs, dim3(wl_blocks(e)), dim3(e->block), 0, e->stream, e->grid(e->cur), e->reps, e->cap,
                       e->radius, r, cutoff * cutoff, out, (uint32_t)cap, cnt, (uint32_t)e->max_particles);
    uint32_t hc = 0;
    HIP_TRY(hipMemcpyAsync(&hc, cnt, 4, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    *n = hc;
    uint64_t m = std::min<uint64_t>(hc, cap);
    if (m && pairs) {
        HIP_TRY(hipMemcpyAsync(pairs, out, m * 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    return RDY_OK;
}

int rdy_reaction_candidates(rdy_engine *e, uint32_t r, uint64_t cap, uint32_t *rows, uint64_t *n) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    if ((rc = upload_tables(e, e->tables_dt > 0 ? e->tables_dt : 1.0))) return rc;
    size_t bytes = 64 + 4 * cap * sizeof(uint32_t);
    if ((rc = ensure_scratch(e, bytes))) return rc;
    uint32_t *cnt = (uint32_t *)e->scratch;
    uint32_t *out = cnt + 16;
    HIP_TRY(hipMemsetAsync(cnt, 0, 4, e->stream));
    ensure_wl(e);
    hipLaunchKernelGGL(k_search, dim3(wl_blocks(e)), dim3(e->block), 0, e->stream, e->grid(e->cur), e->reps, e->T,
                       e->cap, e->radius, 0, 1, e->events, cnt, out, (uint32_t)cap, r, (uint32_t)e->max_particles,
                       e->err);
    uint32_t hc = 0;
    HIP_TRY(hipMemcpyAsync(&hc, cnt, 4, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    *n = hc;
    uint64_t m = std::min<uint64_t>(hc, cap);
    if (m && rows) {
        HIP_TRY(hipMemcpyAsync(rows, out, m * 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    return check_errors(e);
}

int rdy_break_bonds(rdy_engine *e, double dt, uint32_t n_pairs, const rdy_break_pair *pairs, uint64_t cap,
                    uint32_t *broken, uint64_t *n_broken) {
    if (!e || !e->pools_ready) return set_err(RDY_ERR_STATE, "engine has no replicas");
    if (!n_broken || (n_pairs && !pairs) || (cap && !broken)) return set_err(RDY_ERR_INVALID, "null argument");
    if (!(dt > 0.0)) return set_err(RDY_ERR_INVALID, "break bonds: time step must be positive");
    *n_broken = 0;
    const uint32_t nt = e->n_types;
    // an unordered type pair -> (threshold, rate); the first entry of a pair wins, as map::emplace
    std::vector<double> tab(2ull * nt * nt, 0.0);
    for (uint32_t k = 0; k < n_pairs; ++k) {
        const rdy_break_pair &q = pairs[k];
        if (q.type1 >= nt || q.type2 >= nt) return set_err(RDY_ERR_INVALID, "break bonds: no such particle type");
        if (!(q.rate > 0.0)) return set_err(RDY_ERR_INVALID, "Rate of breaking a bond must be positive");
        for (const auto ab : {std::make_pair(q.type1, q.type2), std::make_pair(q.type2, q.type1)}) {
            double *c = &tab[2ull * (ab.first * nt + ab.second)];
            if (c[1] > 0.0) continue;
            c[0] = q.threshold_energy;
            c[1] = q.rate;
        }
    }
    const uint32_t n = (uint32_t)e->terms_h[0].size();
    const uint32_t call = e->break_calls++;
    if (n == 0 || !e->d_terms[0]) return RDY_OK;
    const size_t off_c = (8 * tab.size() + 15) & ~(size_t)15, off_o = off_c + 16;
    const uint32_t ocap = (uint32_t)std::min<uint64_t>(n, (uint64_t)UINT32_MAX / 3);
    int rc = ensure_scratch(e, off_o + 12ull * ocap);
    if (rc) return rc;
    char *base = (char *)e->scratch;
    HIP_TRY(hipMemcpyAsync(base, tab.data(), 8 * tab.size(), hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemsetAsync(base + off_c, 0, 4, e->stream));
    hipLaunchKernelGGL(k_break_bonds, dim3(blocks_for(n, 256)), dim3(256), 0, e->stream, e->grid(e->cur),
                       (const ReplicaDev *)e->reps, e->topo(), (const BondTerm *)e->d_terms[0], n,
                       (const double2 *)base, nt, dt, call, (uint32_t *)(base + off_o), ocap,
                       (uint32_t *)(base + off_c));
    uint32_t got = 0;
    HIP_TRY(hipMemcpyAsync(&got, base + off_c, 4, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    got = std::min(got, ocap);  // at most one entry per edge, and edges <= terms
    std::vector<uint32_t> h(3ull * got);
    if (got) {
        HIP_TRY(hipMemcpyAsync(h.data(), base + off_o, 12ull * got, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    std::vector<std::array<uint32_t, 3>> v(got);
    for (uint32_t k = 0; k < got; ++k) v[k] = {h[3 * k], h[3 * k + 1], h[3 * k + 2]};
    std::sort(v.begin(), v.end());
    for (uint64_t k = 0; k < v.size() && k < cap; ++k)
        for (int j = 0; j < 3; ++j) broken[3 * k + j] = v[k][j];
    *n_broken = v.size();
    return check_errors(e);
}

int rdy_count_types(rdy_engine *e, uint64_t *counts, uint32_t n_reps) {
    if (!e || !e->pools_ready) return set_err(RDY_ERR_STATE, "engine has no replicas");
    size_t words = (size_t)n_reps * e->n_types;
    int rc = ensure_scratch(e, words * 8 + 64);
    if (rc) return rc;
    unsigned long long *d = (unsigned long long *)e->scratch;
    HIP_TRY(hipMemsetAsync(d, 0, words * 8, e->stream));
    ensure_wl(e);
    hipLaunchKernelGGL(k_count_types, dim3(wl_blocks(e)), dim3(e->block), 0, e->stream, e->grid(e->cur), e->n_types, d,
                       n_reps, (uint32_t)e->max_particles);
    HIP_TRY(hipMemcpyAsync(counts, d, words * 8, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return RDY_OK;
}

int rdy_rdf_counts(rdy_engine *e, uint32_t r, const double *borders, uint32_t n_borders, const uint32_t *from_types,
                   uint32_t n_from_types, const uint32_t *to_types, uint32_t n_to_types, uint64_t *counts,
                   uint64_t *n_from) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    if (!borders || !counts || !n_from || n_borders < 2 || n_borders > RDF_MAX_BORDERS)
        return set_err(RDY_ERR_INVALID, "radial distribution: 2 to 4097 bin borders and output arrays required");
    for (uint32_t b = 0; b + 1 < n_borders; ++b)
        if (!(borders[b] <= borders[b + 1])) return set_err(RDY_ERR_INVALID, "bin borders must ascend");
    const uint32_t words = (e->n_types + 31) / 32;
    std::vector<uint32_t> masks(2 * (size_t)words, 0u);
    auto fill = [&](const uint32_t *t, uint32_t n, uint32_t *m) -> int {
        for (uint32_t k = 0; k < n; ++k) {
            if (t[k] >= e->n_types) return set_err(RDY_ERR_INVALID, "radial distribution: no such species");
            m[t[k] >> 5] |= 1u << (t[k] & 31u);
        }
        return RDY_OK;
    };
    if ((rc = fill(from_types, n_from_types, masks.data()))) return rc;
    if ((rc = fill(to_types, n_to_types, masks.data() + words))) return rc;
    // stencil radius per axis: every cell within the last border of the particle's cell
    const ReplicaDev &R = e->reps_h[r].dev;
    const double rmax = borders[n_borders - 1];
    int rad[3];
    for (int a = 0; a < 3; ++a) {
        const double c = std::ceil(rmax / R.cs[a]);
        rad[a] = (int)std::min<double>(std::max(c, 1.0), (double)R.dims[a]);
    }
    const uint32_t outside_cap = (uint32_t)std::min<uint64_t>(e->max_particles, 1u << 20);
    const size_t nbins = n_borders - 1;
    size_t off_b = 0, off_m = off_b + 8 * (size_t)n_borders, off_c = (off_m + 8 * words + 15) & ~(size_t)15;
    size_t off_n = off_c + 8 * nbins, off_o = off_n + 16, bytes = off_o + 4 * (size_t)outside_cap;
    if ((rc = ensure_scratch(e, bytes))) return rc;
    char *base = (char *)e->scratch;
    double *d_b = (double *)(base + off_b);
    uint32_t *d_m = (uint32_t *)(base + off_m);
    unsigned long long *d_c = (unsigned long long *)(base + off_c);
    unsigned long long *d_n = (unsigned long long *)(base + off_n);
    uint32_t *d_no = (uint32_t *)(d_n + 1);
    uint32_t *d_out = (uint32_t *)(base + off_o);
    HIP_TRY(hipMemcpyAsync(d_b, borders, 8 * (size_t)n_borders, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(d_m, masks.data(), 8 * (size_t)words, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemsetAsync(d_c, 0, 8 * nbins + 16, e->stream));
    ensure_wl(e);
    const Grid g = e->grid(e->cur);
    hipLaunchKernelGGL(k_rdf, dim3(blocks_for(e->max_particles, 256)), dim3(256), 0, e->stream, g,
                       (const ReplicaDev *)e->reps, r, e->cap, (const double *)d_b, n_borders, (const uint32_t *)d_m,
                       (const uint32_t *)(d_m + words), rad[0], rad[1], rad[2], d_c, d_n, d_out, outside_cap, d_no,
                       (uint32_t)e->max_particles);
    uint32_t no = 0;
    HIP_TRY(hipMemcpyAsync(&no, d_no, 4, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (no > outside_cap) return set_err(RDY_ERR_CAPACITY, "radial distribution: too many particles outside the box");
    if (no) {
        const uint64_t pairs = (uint64_t)no * e->max_particles;
        if ((pairs + 255) / 256 > 0x7FFFFFFFull)
            return set_err(RDY_ERR_CAPACITY, "radial distribution: too many particles outside the box");
        hipLaunchKernelGGL(k_rdf_outside, dim3((uint32_t)((pairs + 255) / 256)), dim3(256), 0, e->stream, g,
                           (const ReplicaDev *)e->reps, r, (const double *)d_b, n_borders, (const uint32_t *)d_m,
                           (const uint32_t *)(d_m + words), d_c, (const uint32_t *)d_out, (const uint32_t *)d_no,
                           (uint32_t)e->max_particles);
    }
    HIP_TRY(hipMemcpyAsync(counts, d_c, 8 * nbins, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipMemcpyAsync(n_from, d_n, 8, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return check_errors(e);
}

int rdy_histogram_along_axis(rdy_engine *e, uint32_t r, const double *borders, uint32_t n_borders,
                             const uint32_t *types, uint32_t n_types, uint32_t axis, uint64_t *counts) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    if (!borders || !counts || n_borders < 2 || n_borders > RDF_MAX_BORDERS || axis > 2)
        return set_err(RDY_ERR_INVALID, "histogram: 2 to 4097 bin borders, an axis 0..2 and an output array required");
    for (uint32_t b = 0; b + 1 < n_borders; ++b)
        if (!(borders[b] <= borders[b + 1])) return set_err(RDY_ERR_INVALID, "bin borders must ascend");
    const uint32_t words = (e->n_types + 31) / 32;
    std::vector<uint32_t> mask(words, 0u);
    for (uint32_t k = 0; k < n_types; ++k) {
        if (types[k] >= e->n_types) return set_err(RDY_ERR_INVALID, "histogram: no such species");
        mask[types[k] >> 5] |= 1u << (types[k] & 31u);
    }
    const size_t nbins = n_borders - 1;
    const size_t off_m = 8 * (size_t)n_borders, off_c = (off_m + 4 * words + 15) & ~(size_t)15;
    if ((rc = ensure_scratch(e, off_c + 8 * nbins))) return rc;
    char *base = (char *)e->scratch;
    HIP_TRY(hipMemcpyAsync(base, borders, 8 * (size_t)n_borders, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(base + off_m, mask.data(), 4 * (size_t)words, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemsetAsync(base + off_c, 0, 8 * nbins, e->stream));
    ensure_wl(e);
    hipLaunchKernelGGL(k_hist_axis, dim3(blocks_for(e->max_particles, 256)), dim3(256), 0, e->stream,
                       e->grid(e->cur), r, (const double *)base, n_borders, (const uint32_t *)(base + off_m),
                       (int)axis, (unsigned long long *)(base + off_c), (uint32_t)e->max_particles);
    HIP_TRY(hipMemcpyAsync(counts, base + off_c, 8 * nbins, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return check_errors(e);
}

int rdy_histogram_along_axis_noalloc(rdy_engine *e, uint32_t r, const double *borders, uint32_t n_borders,
                                     const uint32_t *types, uint32_t n_types, uint32_t axis, uint64_t *counts) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    if (!borders || !counts || axis > 2) return set_err(RDY_ERR_INVALID, "histogram: bin borders, an axis 0..2 and an output array required");
    if (n_borders < 2 || n_borders > RDF_MAX_BORDERS)
        return set_err(RDY_ERR_INVALID, "histogram (no-alloc): 2 to 4097 bin borders (the workspace holds at most "
                                        "4097 borders and 4096 counts)");
    for (uint32_t b = 0; b + 1 < n_borders; ++b)
        if (!(borders[b] <= borders[b + 1])) return set_err(RDY_ERR_INVALID, "bin borders must ascend");
    uint32_t mask[MAX_TYPES / 32] = {};
    const uint32_t words = (e->n_types + 31) / 32;
    for (uint32_t k = 0; k < n_types; ++k) {
        if (!types || types[k] >= e->n_types) return set_err(RDY_ERR_INVALID, "histogram: no such species");
        mask[types[k] >> 5] |= 1u << (types[k] & 31u);
    }
    char *base = (char *)g_noalloc_ptr[e->device];
    if (!base) return set_err(RDY_ERR_STATE, "no workspace on this device");
    const uint32_t nbins = n_borders - 1;
    double *d_b = (double *)(base + NOALLOC_OFF_BORDERS);
    uint32_t *d_m = (uint32_t *)(base + NOALLOC_OFF_MASK);
    unsigned long long *d_c = (unsigned long long *)(base + NOALLOC_OFF_COUNTS);
    HIP_TRY(hipMemcpyAsync(d_b, borders, 8 * (size_t)n_borders, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(d_m, mask, 4 * (size_t)words, hipMemcpyHostToDevice, e->stream));
    hipLaunchKernelGGL(k_zero_u64, dim3(blocks_for(nbins, 256)), dim3(256), 0, e->stream, d_c, nbins);
    ensure_wl(e);
    hipLaunchKernelGGL(k_hist_axis, dim3(blocks_for(e->max_particles, 256)), dim3(256), 0, e->stream,
                       e->grid(e->cur), r, (const double *)d_b, n_borders, (const uint32_t *)d_m, (int)axis, d_c,
                       (uint32_t)e->max_particles);
    HIP_TRY(hipMemcpyAsync(counts, d_c, 8 * (size_t)nbins, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return check_errors(e);
}

int rdy_set_spatial_reactions(rdy_engine *e, uint32_t n, const rdy_spatial_reaction *sr) {
    if (!e) return set_err(RDY_ERR_INVALID, "null engine");
    if (n && !sr) return set_err(RDY_ERR_INVALID, "null spatial reactions");
    if (e->in_session) return set_err(RDY_ERR_STATE, "a run is open: finish it with rdy_run_end");
    if (n > 64) return set_err(RDY_ERR_INVALID, "at most 64 spatial topology reactions");
    for (uint32_t i = 0; i < n; ++i) {
        const rdy_spatial_reaction &x = sr[i];
        if (x.type1 >= e->n_types || x.type2 >= e->n_types || x.type_to1 >= e->n_types || x.type_to2 >= e->n_types)
            return set_err(RDY_ERR_INVALID, "spatial topology reaction species out of range");
        if (x.fusion > 1 || x.topology_topology > 1 || x.self_fusion > 1 || x.reserved ||
            (x.self_fusion && !(x.fusion && x.topology_topology)))
            return set_err(RDY_ERR_INVALID, "spatial topology reaction: fusion, topology_topology and self_fusion are 0 "
                                            "or 1, self_fusion only for a topology-topology fusion");
        if (!e->topo_flavor[x.type1] || !e->topo_flavor[x.type_to1])
            return set_err(RDY_ERR_INVALID, "the topology side of a topology-particle reaction must stay a topology species");
        if (x.topology_topology) {
            if (!e->topo_flavor[x.type2] || !e->topo_flavor[x.type_to2])
                return set_err(RDY_ERR_INVALID, "both sides of a topology-topology reaction must stay topology species");
            if (!(x.rate > 0))
                return set_err(RDY_ERR_INVALID, "The rate of an structural topology reaction should always be positive");
            if (!(x.radius > 0) || x.radius > e->nl_cutoff)
                return set_err(RDY_ERR_INVALID, "a spatial topology reaction radius must be positive and at most the "
                                                "engine's neighbour-list cutoff (Context::calculateMaxCutoff includes "
                                                "these radii)");
            continue;
        }
        if (e->topo_flavor[x.type2])
            return set_err(RDY_ERR_INVALID, "the particle side of a topology-particle reaction must be a normal species");
        if (!x.fusion && e->topo_flavor[x.type_to2])
            return set_err(RDY_ERR_INVALID, "the particle side of a topology-particle enzymatic reaction must stay a "
                                            "normal species");
        if (x.fusion && !e->topo_flavor[x.type_to2])
            return set_err(RDY_ERR_INVALID, "a topology-particle fusion's joining particle must become a topology species");
        if (!(x.rate > 0))
            return set_err(RDY_ERR_INVALID, "The rate of an structural topology reaction should always be positive");
        if (!(x.radius > 0) || x.radius > e->nl_cutoff)
            return set_err(RDY_ERR_INVALID, "a spatial topology reaction radius must be positive and at most the engine's "
                                            "neighbour-list cutoff (Context::calculateMaxCutoff includes these radii)");
    }
    if (!e->pools_ready) return set_err(RDY_ERR_STATE, "add a replica first");
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->sr_in.assign(sr, sr + n);
    if (n && !e->tev) {
        const size_t nr = e->max_reps;
        HIP_TRY(dalloc(&e->tev, nr * TOPO_MAX_EVENTS));
        HIP_TRY(dalloc(&e->tev_count, nr));
        HIP_TRY(hipMemsetAsync(e->tev_count, 0, nr * sizeof(uint32_t), e->stream));
    }
    bool fusion = false, merge = false;
    for (uint32_t i = 0; i < n; ++i) {
        fusion = fusion || (sr[i].fusion && !sr[i].topology_topology);
        merge = merge || (sr[i].fusion && sr[i].topology_topology);
    }
    if (merge && !e->tmrg) {
        const size_t nr = e->max_reps;
        HIP_TRY(dalloc(&e->tmrg, nr * TOPO_MAX_EVENTS * 4));
        HIP_TRY(dalloc(&e->tmrg_count, nr));
        HIP_TRY(hipMemsetAsync(e->tmrg_count, 0, nr * sizeof(uint32_t), e->stream));
    }
    if (fusion && !e->tfus) {
        const size_t nr = e->max_reps;
        HIP_TRY(dalloc(&e->tfus, nr * TOPO_MAX_EVENTS * 4));
        HIP_TRY(dalloc(&e->tfus_count, nr));
        HIP_TRY(hipMemsetAsync(e->tfus_count, 0, nr * sizeof(uint32_t), e->stream));
    }
    if (e->str_step) HIP_TRY(hipFree(e->str_step));
    if (e->str_total) HIP_TRY(hipFree(e->str_total));
    e->str_step = e->str_total = nullptr;
    const size_t cn = (size_t)e->max_reps * std::max<uint32_t>(n, 1);
    HIP_TRY(dalloc(&e->str_step, cn));
    HIP_TRY(dalloc(&e->str_total, cn));
    HIP_TRY(hipMemsetAsync(e->str_step, 0, cn * 8, e->stream));
    HIP_TRY(hipMemsetAsync(e->str_total, 0, cn * 8, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->tables_dt = -1.0;  // the acceptance probabilities go up with the next tables
    e->version++;
    return RDY_OK;
}

int rdy_set_topologies(rdy_engine *e, uint32_t r, uint32_t n_topologies, const uint32_t *topology_types,
                       const uint32_t *topology_of) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    if (e->in_session) return set_err(RDY_ERR_STATE, "a run is open: finish it with rdy_run_end");
    if ((rc = sync_topo_count(e, r))) return rc;
    HostReplica &H = e->reps_h[r];
    if (n_topologies && !topology_types) return set_err(RDY_ERR_INVALID, "null topology types");
    if (H.n_topo && !topology_of) return set_err(RDY_ERR_INVALID, "null topology assignment");
    for (uint32_t h = 0; h < H.n_topo; ++h)
        if (topology_of[h] != 0xFFFFFFFFu && topology_of[h] >= n_topologies)
            return set_err(RDY_ERR_INVALID, "a topology particle's topology does not exist");
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (H.n_tt) release_range(e->free_tt, H.tt_base, H.n_tt);
    H.tt_base = 0;
    H.n_tt = 0;
    if (n_topologies) {
        uint64_t base = 0;
        if (!take_range(e->free_tt, n_topologies, &base)) {
            if (e->tt_used + n_topologies > e->tt_cap) {
                const uint64_t ncap = std::max<uint64_t>(2 * e->tt_cap, e->tt_used + n_topologies);
                uint32_t *nt;
                HIP_TRY(dalloc(&nt, ncap));
                if (e->ttype) {
                    HIP_TRY(hipMemcpyAsync(nt, e->ttype, e->tt_used * 4, hipMemcpyDeviceToDevice, e->stream));
                    HIP_TRY(hipStreamSynchronize(e->stream));
                    HIP_TRY(hipFree(e->ttype));
                }
                e->ttype = nt;
                uint32_t *nr_;
                HIP_TRY(dalloc(&nr_, ncap));
                HIP_TRY(hipMemsetAsync(nr_, 0, ncap * 4, e->stream));
                if (e->treact) {
                    HIP_TRY(hipMemcpyAsync(nr_, e->treact, e->tt_used * 4, hipMemcpyDeviceToDevice, e->stream));
                    HIP_TRY(hipStreamSynchronize(e->stream));
                    HIP_TRY(hipFree(e->treact));
                }
                e->treact = nr_;
                e->tt_cap = ncap;
            }
            base = e->tt_used;
            e->tt_used += n_topologies;
        }
        H.tt_base = (uint32_t)base;
        H.n_tt = n_topologies;
        HIP_TRY(hipMemcpyAsync(e->ttype + base, topology_types, n_topologies * 4ull, hipMemcpyHostToDevice, e->stream));
    }
    if (H.n_topo)
        HIP_TRY(hipMemcpyAsync(e->topo_of + H.topo_base, topology_of, H.n_topo * 4ull, hipMemcpyHostToDevice, e->stream));
    if ((rc = put_field(e, r, offsetof(ReplicaDev, tt_base), H.tt_base))) return rc;
    if ((rc = put_field(e, r, offsetof(ReplicaDev, n_tt), H.n_tt))) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->version++;
    return RDY_OK;
}

int rdy_get_topology_types(rdy_engine *e, uint32_t r, uint32_t cap, uint32_t *types, uint32_t *n) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    const HostReplica &H = e->reps_h[r];
    if (n) *n = H.n_tt;
    const uint32_t m = std::min(cap, H.n_tt);
    if (m && types) {
        HIP_TRY(hipMemcpyAsync(types, e->ttype + H.tt_base, m * 4ull, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    return RDY_OK;
}

int rdy_reserve_topology_particles(rdy_engine *e, uint32_t r, uint32_t n) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    if (n > (1u << 30)) return set_err(RDY_ERR_INVALID, "topology reserve beyond 2^30 handles");
    e->reps_h[r].topo_reserve = n;
    return RDY_OK;
}

int rdy_get_topology_fusions(rdy_engine *e, uint32_t r, uint32_t cap, uint32_t *rows, uint32_t *n) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    uint32_t m = 0;
    if (e->tfus_count) {
        HIP_TRY(hipMemcpyAsync(&m, e->tfus_count + r, 4, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    if (n) *n = m;
    const uint32_t k = std::min(cap, m);
    if (k && rows) {
        HIP_TRY(hipMemcpyAsync(rows, e->tfus + 4ull * r * TOPO_MAX_EVENTS, 16ull * k, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    return sync_topo_count(e, r);
}

int rdy_get_topology_merges(rdy_engine *e, uint32_t r, uint32_t cap, uint32_t *rows, uint32_t *n) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    uint32_t m = 0;
    if (e->tmrg_count) {
        HIP_TRY(hipMemcpyAsync(&m, e->tmrg_count + r, 4, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    if (n) *n = m;
    const uint32_t k = std::min(cap, m);
    if (k && rows) {
        HIP_TRY(hipMemcpyAsync(rows, e->tmrg + 4ull * r * TOPO_MAX_EVENTS, 16ull * k, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    return RDY_OK;
}

int rdy_evaluate_topology_reactions(rdy_engine *e, double dt) {
    if (!e || !e->pools_ready) return set_err(RDY_ERR_STATE, "engine has no replicas");
    if (e->in_session) return set_err(RDY_ERR_STATE, "a run is open: finish it with rdy_run_end");
    int rc = upload_tables(e, dt);
    if (rc) return rc;
    if (!topo_reactions_on(e)) return RDY_OK;
    if (++e->topo_serial == 0) ++e->topo_serial;  // 0 marks evaluations that are not recorded
    launch_topo_reactions(e, 0, nullptr, nullptr, e->topo_serial);
    e->export_valid = false;  // species changed: forces and energies of the state are formed anew
    return check_errors(e);
}

int rdy_get_topology_reacted(rdy_engine *e, uint32_t r, uint32_t cap, uint8_t *reacted, uint32_t *n) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    const HostReplica &H = e->reps_h[r];
    if (n) *n = H.n_tt;
    const uint32_t m = std::min(cap, H.n_tt);
    if (m && reacted) {
        std::vector<uint32_t> v(m);
        HIP_TRY(hipMemcpyAsync(v.data(), e->treact + H.tt_base, m * 4ull, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
        for (uint32_t k = 0; k < m; ++k) reacted[k] = e->topo_serial != 0 && v[k] == e->topo_serial ? 1 : 0;
    }
    return RDY_OK;
}

int rdy_set_particle_types(rdy_engine *e, uint32_t r, uint64_t n, const uint32_t *handles, const uint32_t *types) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    if (e->in_session) return set_err(RDY_ERR_STATE, "a run is open: finish it with rdy_run_end");
    if (n == 0) return RDY_OK;
    if (!handles || !types) return set_err(RDY_ERR_INVALID, "null handles or types");
    if ((rc = sync_topo_count(e, r))) return rc;
    const HostReplica &H = e->reps_h[r];
    for (uint64_t i = 0; i < n; ++i) {
        if (handles[i] >= H.n_topo)
            return set_err(RDY_ERR_INVALID, "only topology particles (handles below the replica's topology count) "
                                            "can change species here");
        if (types[i] >= e->n_types) return set_err(RDY_ERR_INVALID, "particle type out of range");
    }
    if ((rc = upload_tables(e, e->tables_dt > 0 ? e->tables_dt : 1.0))) return rc;
    if ((rc = ensure_scratch(e, (size_t)(2 * n + 1) * sizeof(uint32_t)))) return rc;
    uint32_t *d = (uint32_t *)e->scratch;
    HIP_TRY(hipMemcpyAsync(d, handles, n * 4, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(d + n, types, n * 4, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemsetAsync(d + 2 * n, 0, 4, e->stream));
    hipLaunchKernelGGL(k_set_topo_types, dim3(blocks_for((uint32_t)n, 256)), dim3(256), 0, e->stream,
                       e->grid(e->cur), (const ReplicaDev *)e->reps, r, (const uint32_t *)e->loc, (const uint32_t *)d,
                       (const uint32_t *)(d + n), (const uint8_t *)e->d_flavor, (uint32_t)n, d + 2 * n);
    uint32_t bad = 0;
    HIP_TRY(hipMemcpyAsync(&bad, d + 2 * n, 4, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (bad) return set_err(RDY_ERR_STATE, "a topology particle's record was not found");
    e->export_valid = false;  // species changed: forces and energies are formed anew
    e->version++;
    return RDY_OK;
}

int rdy_set_particle_positions(rdy_engine *e, uint32_t r, uint64_t n, const uint32_t *handles, const double *pos) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    if (e->in_session) return set_err(RDY_ERR_STATE, "a run is open: finish it with rdy_run_end");
    if (n == 0) return RDY_OK;
    if (!handles || !pos) return set_err(RDY_ERR_INVALID, "null handles or positions");
    if ((rc = sync_topo_count(e, r))) return rc;
    const HostReplica &H = e->reps_h[r];
    for (uint64_t i = 0; i < n; ++i) {
        if (handles[i] >= H.n_topo)
            return set_err(RDY_ERR_INVALID, "only topology particles (handles below the replica's topology count) "
                                            "can be moved here");
        for (int d = 0; d < 3; ++d)
            if (!std::isfinite(pos[3 * i + d])) return set_err(RDY_ERR_INVALID, "non-finite particle position");
    }
    if ((rc = ensure_scratch(e, (size_t)n * (4 + 24) + 16))) return rc;
    double *dp = (double *)e->scratch;
    uint32_t *dh = (uint32_t *)(dp + 3 * n), *bad = dh + n;
    HIP_TRY(hipMemcpyAsync(dp, pos, n * 24, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(dh, handles, n * 4, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemsetAsync(bad, 0, 4, e->stream));
    hipLaunchKernelGGL(k_set_topo_positions, dim3(blocks_for((uint32_t)n, 256)), dim3(256), 0, e->stream,
                       e->grid(e->cur), (const ReplicaDev *)e->reps, r, (const uint32_t *)e->loc, (const uint32_t *)dh,
                       (const double *)dp, (uint32_t)n, bad);
    uint32_t hb = 0;
    HIP_TRY(hipMemcpyAsync(&hb, bad, 4, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (hb) return set_err(RDY_ERR_STATE, "a topology particle's record was not found");
    e->export_valid = false;
    e->version++;
    return rdy_update_neighbor_list(e);  // the moved records into the cells of their new positions
}

int rdy_append_topology_particles(rdy_engine *e, uint32_t r, uint64_t n, const uint32_t *types, const double *pos,
                                  uint32_t *handles) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    if (e->in_session) return set_err(RDY_ERR_STATE, "a run is open: finish it with rdy_run_end");
    if (n == 0) return RDY_OK;
    if (!types || !pos || !handles) return set_err(RDY_ERR_INVALID, "null types, positions or handles");
    if ((rc = sync_topo_count(e, r))) return rc;
    HostReplica &H = e->reps_h[r];
    for (uint64_t i = 0; i < n; ++i) {
        if (types[i] >= e->n_types || !e->topo_flavor[types[i]])
            return set_err(RDY_ERR_INVALID, "appended particles must be of a topology species");
        for (int d = 0; d < 3; ++d)
            if (!std::isfinite(pos[3 * i + d])) return set_err(RDY_ERR_INVALID, "non-finite particle position");
    }
    if (H.n_topo + n > H.topo_range)
        return set_err(RDY_ERR_CAPACITY, "appended particles beyond the replica's topology reserve: raise it with "
                                         "rdy_reserve_topology_particles before the upload");
    if ((rc = upload_tables(e, e->tables_dt > 0 ? e->tables_dt : 1.0))) return rc;
    if ((rc = ensure_scratch(e, (size_t)n * (4 + 24)))) return rc;
    double *dp = (double *)e->scratch;
    uint32_t *dt = (uint32_t *)(dp + 3 * n);
    HIP_TRY(hipMemcpyAsync(dp, pos, n * 24, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(dt, types, n * 4, hipMemcpyHostToDevice, e->stream));
    const uint32_t h0 = H.n_topo;
    hipLaunchKernelGGL(k_append_topo, dim3(blocks_for((uint32_t)n, 256)), dim3(256), 0, e->stream, e->grid(e->cur),
                       e->reps, r, h0, (const uint32_t *)dt, (const double *)dp, (uint32_t)n, e->loc, e->cap, e->err);
    H.n_topo += (uint32_t)n;
    if ((rc = put_field(e, r, offsetof(ReplicaDev, n_topo), H.n_topo))) return rc;
    for (uint64_t i = 0; i < n; ++i) handles[i] = h0 + (uint32_t)i;
    e->wl_valid[e->cur] = false;
    e->export_valid = false;
    e->version++;
    return check_errors(e);
}

int rdy_get_topology_reaction_counts(rdy_engine *e, uint32_t r, int totals, uint64_t *counts) {
    int rc = check_replica(e, r);
    if (rc) return rc;
    const size_t n = e->sr_in.size();
    if (!n) return RDY_OK;
    if (!counts) return set_err(RDY_ERR_INVALID, "null counts");
    HIP_TRY(hipMemcpyAsync(counts, (totals ? e->str_total : e->str_step) + (size_t)r * n, n * 8,
                           hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return RDY_OK;
}

int rdy_profile_reset(rdy_engine *e) {
    if (!e) return set_err(RDY_ERR_INVALID, "null engine");
    if (e->prof_ev.empty()) {
        e->prof_ev.resize(256);
        for (auto &pe : e->prof_ev) {
            HIP_TRY(hipEventCreate(&pe.first));
            HIP_TRY(hipEventCreate(&pe.second));
        }
    }
    e->prof = true;
    e->prof_used = 0;
    e->prof_ms = 0.0;
    e->prof_n = 0;
    return RDY_OK;
}

int rdy_profile_get(rdy_engine *e, double *avg_us_fused, uint64_t *launches) {
    if (!e) return set_err(RDY_ERR_INVALID, "null engine");
    int rc = flush_prof(e);
    if (rc) return rc;
    if (avg_us_fused) *avg_us_fused = e->prof_n ? 1000.0 * e->prof_ms / (double)e->prof_n : 0.0;
    if (launches) *launches = e->prof_n;
    return RDY_OK;
}

}  // extern "C"
