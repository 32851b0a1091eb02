// Kernels of the MI355X ReaDDy engine (see rdy_kernels.h for the data layout).
#include "rdy_kernels.h"
#include "rdy_reversible.h"

namespace rdy {

// -------------------------------------------------------------------------------------------------
// helpers
// -------------------------------------------------------------------------------------------------

// XCD-aware block order: the 8 XCDs of an MI355X receive blocks round-robin, so consecutive
// hardware block ids are remapped to contiguous chunks of the work list per XCD.  Neighbouring
// particles (and their shared neighbour cells) then stay within one XCD's L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nblocks) {
    const uint32_t nx = 8;
    uint32_t per = nblocks / nx;
    if (per == 0 || b >= per * nx) return b;
    return (b % nx) * per + b / nx;
}

__device__ __forceinline__ int axis_cells(int i, int dims, int periodic, int radius, int *out) {
    if (periodic && dims <= 2 * radius) {
        for (int j = 0; j < dims; ++j) out[j] = j;
        return dims;
    }
    int n = 0;
    for (int o = -radius; o <= radius; ++o) {
        int j = i + o;
        if (periodic) {
            j = (j % dims + dims) % dims;
        } else if (j < 0 || j >= dims) {
            continue;
        }
        out[n++] = j;
    }
    return n;
}

struct RepGeom {
    double box[3], cs[3];
    int32_t pbc[3], dims[3];
    uint32_t cell_base;
};

__device__ __forceinline__ RepGeom load_geom(const ReplicaDev &R) {
    RepGeom g;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        g.box[d] = R.box[d];
        g.cs[d] = R.cs[d];
        g.pbc[d] = R.pbc[d];
        g.dims[d] = R.dims[d];
    }
    g.cell_base = R.cell_base;
    return g;
}

// visits every slot of the stencil around local cell `cl` (own cell included; the caller skips
// itself): the (2r+1)^3 adjacency of CellLinkedList.cpp:74-137 with the same de-duplication
__device__ __forceinline__ uint32_t cell_of_slot(const Grid &g, uint32_t s, uint32_t cap) {
    return s < g.nbin ? s / cap : g.ovf_cell[s - g.nbin];
}

template <class F>
__device__ __forceinline__ void for_each_stencil_slot(const RepGeom &g, const Grid &grid, uint32_t cap, int radius,
                                                      uint32_t cl, F &&f) {
    const uint32_t *counts = grid.counts;
    const int nz = g.dims[2], ny = g.dims[1];
    int iz = (int)(cl % (uint32_t)nz);
    int iy = (int)((cl / (uint32_t)nz) % (uint32_t)ny);
    int ix = (int)(cl / ((uint32_t)nz * (uint32_t)ny));
    int ax[7], ay[7], az[7];
    int nax = axis_cells(ix, g.dims[0], g.pbc[0], radius, ax);
    int nay = axis_cells(iy, g.dims[1], g.pbc[1], radius, ay);
    int naz = axis_cells(iz, g.dims[2], g.pbc[2], radius, az);
    for (int a = 0; a < nax; ++a)
        for (int b = 0; b < nay; ++b) {
            uint32_t row = g.cell_base + ((uint32_t)ax[a] * (uint32_t)ny + (uint32_t)ay[b]) * (uint32_t)nz;
            for (int c = 0; c < naz; ++c) {
                uint32_t cell = row + (uint32_t)az[c];
                const uint32_t n = counts[cell];
                const uint32_t m = n < cap ? n : cap;
                const uint32_t base = cell * cap;
                for (uint32_t k = 0; k < m; ++k) f(base + k);
                if (n > cap) {  // rare: a crowded cell spilled into the overflow region
                    uint32_t no = *grid.ovf_count;
                    no = no < grid.ovf_cap ? no : grid.ovf_cap;
                    for (uint32_t o = 0; o < no; ++o)
                        if (grid.ovf_cell[o] == cell) f(grid.nbin + o);
                }
            }
        }
}

__device__ __forceinline__ unsigned long long lanemask_lt() {
    return (1ull << __lane_id()) - 1ull;
}

// set bits of a wave mask in the lanes below this one (v_mbcnt_lo / v_mbcnt_hi)
__device__ __forceinline__ uint32_t rank_in_mask(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// squared length of an FP32 local-frame displacement for the conservative prefilters: fused
// multiply-adds round no worse than the separate products and sums the prefilter margin allows for
__device__ __forceinline__ float prefilter_d2(float dx, float dy, float dz) {
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// wave-aggregated append: returns the position for lanes with want == true
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool want) {
    unsigned long long mask = __ballot(want);
    uint32_t cnt = (uint32_t)__popcll(mask);
    uint32_t base = 0;
    if (mask) {
        int leader = __ffsll((long long)mask) - 1;
        if ((int)__lane_id() == leader) base = atomicAdd(counter, cnt);
        base = __shfl(base, leader);
    }
    return base + (uint32_t)__popcll(mask & lanemask_lt());
}

// per-wave reduction of a fixed-point value into reps[r].<field>; lanes may differ in replica
__device__ __forceinline__ void wave_add_energy(ReplicaDev *reps, bool valid, uint32_t r, long long e) {
    unsigned long long vm = __ballot(valid);
    if (!vm) return;
    int leader = __ffsll((long long)vm) - 1;
    uint32_t rl = __shfl(r, leader);
    bool same = !valid || r == rl;
    if (__all(same)) {
        long long v = valid ? e : 0;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if ((int)__lane_id() == leader && v != 0) atomicAdd(&reps[rl].energy, (unsigned long long)v);
    } else if (valid && e != 0) {
        atomicAdd(&reps[r].energy, (unsigned long long)e);
    }
}

// places `st` (position already wrapped) into buffer g; returns the slot or UINT32_MAX
__device__ __forceinline__ uint32_t bin_insert(const Grid &g, const RepGeom &geo, uint32_t cap, State st,
                                               uint32_t *err, bool append_wl, uint32_t wl_cap) {
    const double p[3] = {st.x, st.y, st.z};
    int32_t c[3];
    bool outside = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (!in_box_axis(p[d], geo.box[d])) outside = true;
        c[d] = cell_coord(p[d], geo.box[d], geo.cs[d], geo.dims[d]);
    }
    st.tw = (st.tw & ~TW_OUTSIDE) | (outside ? TW_OUTSIDE : 0u);
    uint32_t cell = geo.cell_base + ((uint32_t)c[0] * (uint32_t)geo.dims[1] + (uint32_t)c[1]) * (uint32_t)geo.dims[2] +
                    (uint32_t)c[2];
    uint32_t k = atomicAdd(&g.counts[cell], 1u);
    uint32_t slot;
    if (k < cap) {
        slot = cell * cap + k;
    } else {
        uint32_t o = atomicAdd(g.ovf_count, 1u);
        if (o >= g.ovf_cap) {
            atomicAdd(&err[ERR_BIN_OVERFLOW], 1u);
            return 0xFFFFFFFFu;
        }
        g.ovf_cell[o] = cell;
        slot = g.nbin + o;
    }
    g.bins[slot] = st;
    if (append_wl) {
        uint32_t pos = atomicAdd(g.wl_count, 1u);
        if (pos < wl_cap) g.wl[pos] = slot;
        else atomicAdd(&err[ERR_WL_OVERFLOW], 1u);
    }
    return slot;
}

// -------------------------------------------------------------------------------------------------
// per-step bookkeeping
// -------------------------------------------------------------------------------------------------
// mode 0: zero the energy accumulators (a force evaluation follows); mode 1: advance the Philox
// step of replicas with budget left (separate integrator action); mode 2: both (fused step)
__global__ void k_advance(ReplicaDev *reps, uint32_t n_reps, int mode, unsigned long long *e_bonded,
                          uint32_t *edit_count, unsigned long long *virial) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0 && edit_count) *edit_count = 0u;
    if (r >= n_reps) return;
    ReplicaDev &R = reps[r];
    if (mode != 1) {
        R.energy = 0ull;
        e_bonded[r] = 0ull;
        if (virial)
            for (int c = 0; c < 9; ++c) virial[9ull * r + c] = 0ull;
    }
    if (mode >= 1) {
        if (R.in_use && R.steps_left > 0) {
            R.active = 1u;
            R.step += 1u;
            R.steps_left -= 1u;
            R.steps_done += 1u;
        } else {
            R.active = 0u;
        }
    }
}

__global__ void k_reset_counter(Grid old) {
    *old.wl_count = 0u;
    *old.ovf_count = 0u;
}

// -------------------------------------------------------------------------------------------------
// work-list rebuild in cell order: an exclusive scan of the per-cell counters, then every cell writes
// the slots of its members.  Neighbouring threads then hold neighbouring particles, so the stencil
// reads of a wave (and of the blocks an XCD runs) hit the same L2 lines.
//   pass 1: tile sums (SCAN_TILE cells per block)      pass 2: scan of the tile sums (one block)
//   pass 3: per-cell offsets + slot writes; the overflow records follow the cell-ordered slots
// -------------------------------------------------------------------------------------------------
constexpr uint32_t SCAN_THREADS = 256;
constexpr uint32_t SCAN_PER_THREAD = 16;
constexpr uint32_t SCAN_TILE = SCAN_THREADS * SCAN_PER_THREAD;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t warp_sums[SCAN_THREADS / 64];
    const uint32_t lane = __lane_id(), wid = threadIdx.x / 64;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63) warp_sums[wid] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
    for (uint32_t w = 0; w < SCAN_THREADS / 64; ++w) {
        if (w < wid) base += warp_sums[w];
        tot += warp_sums[w];
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_wl_tiles(Grid g, uint32_t cap, uint32_t n_cells,
                                                           uint32_t *tile_sums) {
    const uint32_t c0 = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER_THREAD;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < SCAN_PER_THREAD; ++k) {
        uint32_t c = c0 + k;
        if (c < n_cells) {
            uint32_t n = g.counts[c];
            sum += n < cap ? n : cap;
        }
    }
    uint32_t total;
    block_exclusive_scan(sum, &total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_wl_tile_scan(uint32_t *tile_sums, uint32_t n_tiles, Grid g,
                                                        uint32_t max_wl, uint32_t *err) {
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n_tiles; base += blockDim.x) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < n_tiles ? tile_sums[i] : 0u;
        // inclusive scan across the block with wave shuffles + LDS
        __shared__ uint32_t ws[16];
        uint32_t lane = __lane_id(), wid = threadIdx.x / 64;
        uint32_t x = v;
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t y = __shfl_up(x, off);
            if (lane >= (uint32_t)off) x += y;
        }
        if (lane == 63) ws[wid] = x;
        __syncthreads();
        uint32_t pre = 0, tot = 0;
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
            if (w < wid) pre += ws[w];
            tot += ws[w];
        }
        if (i < n_tiles) tile_sums[i] = carry + pre + x - v;
        __syncthreads();
        if (threadIdx.x == 0) carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        uint32_t no = *g.ovf_count;
        no = no < g.ovf_cap ? no : g.ovf_cap;
        uint32_t n = carry + no;
        if (n > max_wl) {
            atomicAdd(&err[ERR_WL_OVERFLOW], 1u);
            n = max_wl;
        }
        *g.wl_count = n;
        tile_sums[n_tiles] = carry;  // start of the overflow records in the work list
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_wl_write(Grid g, uint32_t cap, uint32_t n_cells,
                                                           const uint32_t *tile_sums, uint32_t n_tiles,
                                                           uint32_t max_wl) {
    const uint32_t c0 = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER_THREAD;
    uint32_t cnt[SCAN_PER_THREAD];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < SCAN_PER_THREAD; ++k) {
        uint32_t c = c0 + k;
        uint32_t n = c < n_cells ? g.counts[c] : 0u;
        n = n < cap ? n : cap;
        cnt[k] = n;
        sum += n;
    }
    uint32_t total;
    uint32_t off = tile_sums[blockIdx.x] + block_exclusive_scan(sum, &total);
#pragma unroll
    for (uint32_t k = 0; k < SCAN_PER_THREAD; ++k) {
        const uint32_t c = c0 + k;
        for (uint32_t q = 0; q < cnt[k]; ++q) {
            if (off < max_wl) g.wl[off] = c * cap + q;
            ++off;
        }
    }
    // overflow records (rare) go after the cell-ordered slots
    if (blockIdx.x == 0) {
        uint32_t no = *g.ovf_count;
        no = no < g.ovf_cap ? no : g.ovf_cap;
        const uint32_t start = tile_sums[n_tiles];
        for (uint32_t o = threadIdx.x; o < no; o += blockDim.x)
            if (start + o < max_wl) g.wl[start + o] = g.nbin + o;
    }
}

// One Euler-Maruyama move (CPUEulerBDIntegrator.cpp:63-73): per axis x += sqrt(2 D dt) xi + F D dt / kBT,
// then fixPosition.  sqrt(2 D dt) comes from the tables; the drift keeps the reference's operation order
// (F D dt, then / kBT, skipped when kBT is 1, which leaves the value unchanged).  Every move of the
// engine (separate integrator, fused step, fix-up) goes through this function, so they agree bit for bit.
__device__ __forceinline__ void em_move(const Tables &T, uint32_t type, double dt, const double F[3],
                                        const double xi[3], const double *box, const int32_t *pbc, double p[3],
                                        uint32_t *err) {
    const double D = T.diffusion[type];
    const double sc = T.bd_scale[type];
    const double kbt = T.kbt;
    const bool unit_kbt = kbt == 1.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double det = F[d] * D * dt;
        if (!unit_kbt) det = det / kbt;
        const double rnd = xi[d] * sc;
        p[d] += rnd + det;
        p[d] = wrap(p[d], box[d], pbc[d]);
        if (!isfinite(p[d])) atomicAdd(&err[ERR_NONFINITE], 1u);
    }
}

// the Brownian-dynamics noise of handle h at Philox step `step`: the injected buffer when given,
// else the counter (h, step, 0, 0) of the replica's key
__device__ __forceinline__ void bd_noise(const double *noise, uint32_t h, uint32_t step, uint64_t seed, double xi[3]) {
    if (noise) {
        xi[0] = noise[3ull * h];
        xi[1] = noise[3ull * h + 1];
        xi[2] = noise[3ull * h + 2];
    } else {
        uint32_t x4[4];
        philox(h, step, 0u, 0u, seed, x4);
        bd_normals3(x4, xi);
    }
}

// separate integrator action: in place on the current buffer, forces from force_out
__global__ __launch_bounds__(256) void k_integrate(Grid cur, ReplicaDev *reps, Tables T, uint32_t cap, double dt,
                                                    const double *noise, const double *force_in, uint32_t max_wl,
                                                    uint32_t *err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = *cur.wl_count;
    n = n < max_wl ? n : max_wl;
    if (i >= n) return;
    const uint32_t s = cur.wl[i];
    State me = cur.bins[s];
    if (me.tw & TW_DEAD) return;
    const ReplicaDev &R = reps[tw_replica(me.tw)];
    if (!R.active) return;
    double xi[3];
    bd_noise(noise, me.h, R.step, R.seed, xi);
    const double F[3] = {force_in ? force_in[3ull * s] : 0.0, force_in ? force_in[3ull * s + 1] : 0.0,
                         force_in ? force_in[3ull * s + 2] : 0.0};
    double p[3] = {me.x, me.y, me.z};
    em_move(T, tw_type(me.tw), dt, F, xi, R.box, R.pbc, p, err);
    me.x = p[0];
    me.y = p[1];
    me.z = p[2];
    cur.bins[s] = me;
}

// -------------------------------------------------------------------------------------------------
// bonded terms (CPUTopologyActions.cpp:55-123): FP64, fixed-point atomics into the per-particle
// bonded force; energies into e_bonded[replica] (fixed point 2^-24)
// -------------------------------------------------------------------------------------------------
struct BondTerm {
    uint32_t r, h[4];
    uint32_t mult;
    double k, p0;
};

__device__ __forceinline__ void pos_of(const Grid &g, const Topo &topo, const ReplicaDev &R, uint32_t h, double *p) {
    const State st = g.bins[topo.loc[R.topo_base + h]];
    p[0] = st.x;
    p[1] = st.y;
    p[2] = st.z;
}

__device__ __forceinline__ void sdiff(const double *a, const double *b, const ReplicaDev &R, double *o) {
    for (int d = 0; d < 3; ++d) o[d] = min_image(b[d] - a[d], R.box[d], R.pbc[d]);
}

__device__ __forceinline__ double dot(const double *a, const double *b) {
    double s = a[0] * b[0];
    s = s + a[1] * b[1];
    s = s + a[2] * b[2];
    return s;
}

__device__ __forceinline__ void cross(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __forceinline__ void add_fb(const Topo &topo, const ReplicaDev &R, uint32_t h, const double *f) {
    long long *fb = topo.fb + 3ull * (R.topo_base + h);
    for (int d = 0; d < 3; ++d) atomicAdd((unsigned long long *)&fb[d], (unsigned long long)to_fixed_d(f[d], FORCE_SCALE));
}

__global__ void k_bonds(Grid cur, ReplicaDev *reps, Topo topo, const BondTerm *terms, uint32_t n,
                        unsigned long long *e_bonded) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const BondTerm t = terms[i];
    const ReplicaDev &R = reps[t.r];
    if (!R.in_use || t.k == 0) return;
    double a[3], b[3], x[3];
    pos_of(cur, topo, R, t.h[0], a);
    pos_of(cur, topo, R, t.h[1], b);
    sdiff(a, b, R, x);  // x_ij = x_j - x_i
    double norm = sqrt(dot(x, x));
    double c = (2. * t.k * (norm - t.p0) / norm);
    double fi[3] = {c * x[0], c * x[1], c * x[2]}, fj[3] = {-fi[0], -fi[1], -fi[2]};
    add_fb(topo, R, t.h[0], fi);
    add_fb(topo, R, t.h[1], fj);
    double e = t.k * (norm - t.p0) * (norm - t.p0);
    atomicAdd(&e_bonded[t.r], (unsigned long long)to_fixed_d(e, ENERGY_SCALE));
}

// BreakBonds (Actions.h:235-290): one thread per bond term.  The first term of each run of terms on
// the same edge (same replica and handle pair, consecutive as the configuration lists an edge's bond
// configurations) sums the run's harmonic energies (evaluateEdgeEnergy, Actions.h:291-341); when the
// end types form a breakable pair (brk[t1 * nt + t2] = {threshold, rate}, rate > 0) and the energy
// exceeds the threshold, the edge breaks with probability 1 - exp(-rate dt): a Philox uniform keyed by
// the edge's handles and the call counter.  Broken edges are listed as (replica, h1, h2).
__global__ void k_break_bonds(Grid cur, const ReplicaDev *reps, Topo topo, const BondTerm *terms, uint32_t n,
                              const double2 *brk, uint32_t nt, double dt, uint32_t call, uint32_t *out,
                              uint32_t cap, uint32_t *count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const BondTerm t = terms[i];
    auto same_edge = [&](const BondTerm &u) { return u.r == t.r && u.h[0] == t.h[0] && u.h[1] == t.h[1]; };
    if (i > 0 && same_edge(terms[i - 1])) return;
    const ReplicaDev &R = reps[t.r];
    if (!R.in_use) return;
    const State s0 = cur.bins[topo.loc[R.topo_base + t.h[0]]];
    const State s1 = cur.bins[topo.loc[R.topo_base + t.h[1]]];
    const double2 cfg = brk[tw_type(s0.tw) * nt + tw_type(s1.tw)];
    if (!(cfg.y > 0.0)) return;
    const double a[3] = {s0.x, s0.y, s0.z}, b[3] = {s1.x, s1.y, s1.z};
    double x[3];
    sdiff(a, b, R, x);
    const double norm = sqrt(dot(x, x));
    double energy = 0.0;
    for (uint32_t j = i; j < n; ++j) {
        const BondTerm u = terms[j];
        if (!same_edge(u)) break;
        energy += u.k * (norm - u.p0) * (norm - u.p0);
    }
    if (!(energy > cfg.x)) return;
    uint32_t x4[4];
    philox(t.h[0], call, 5u, t.h[1], R.seed, x4);
    if (!(u53(x4[0], x4[1]) < 1.0 - exp(-cfg.y * dt))) return;
    const uint32_t p = atomicAdd(count, 1u);
    if (p < cap) {
        out[3ull * p] = t.r;
        out[3ull * p + 1] = t.h[0];
        out[3ull * p + 2] = t.h[1];
    }
}

__global__ void k_angles(Grid cur, ReplicaDev *reps, Topo topo, const BondTerm *terms, uint32_t n,
                         unsigned long long *e_bonded) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const BondTerm t = terms[i];
    const ReplicaDev &R = reps[t.r];
    if (!R.in_use) return;
    const double SMALL = 1e-10;  // AnglePotential.cpp:49
    double pi_[3], pj[3], pk[3], xji[3], xjk[3];
    pos_of(cur, topo, R, t.h[0], pi_);
    pos_of(cur, topo, R, t.h[1], pj);
    pos_of(cur, topo, R, t.h[2], pk);
    sdiff(pj, pi_, R, xji);
    sdiff(pj, pk, R, xjk);
    double sp = dot(xji, xjk);
    double e = t.k * (acos(sp / (sqrt(dot(xji, xji)) * sqrt(dot(xjk, xjk)))) - t.p0) *
               (acos(sp / (sqrt(dot(xji, xji)) * sqrt(dot(xjk, xjk)))) - t.p0);
    double nji2 = dot(xji, xji);
    if (nji2 < SMALL) nji2 = SMALL;
    double njk2 = dot(xjk, xjk);
    if (njk2 < SMALL) njk2 = SMALL;
    double nprod = sqrt(nji2) * sqrt(njk2);
    if (nprod < SMALL) nprod = SMALL;
    double inv = 1. / nprod;
    double ct = fmin(fmax(inv * sp, -1.), 1.);
    double si = sqrt(1. - ct * ct);
    if (si < SMALL) si = SMALL;
    si = 1. / si;
    double c = 2. * t.k * (acos(ct) - t.p0) * si;
    double f_i[3], f_j[3], f_k[3];
    for (int d = 0; d < 3; ++d) {
        double force_i = c * ct * (1. / nji2) * xji[d] - c * inv * xjk[d];
        double force_k = -c * inv * xji[d] + c * ct * (1. / njk2) * xjk[d];
        f_i[d] = -force_i;
        f_j[d] = force_i + force_k;
        f_k[d] = -force_k;
    }
    add_fb(topo, R, t.h[0], f_i);
    add_fb(topo, R, t.h[1], f_j);
    add_fb(topo, R, t.h[2], f_k);
    atomicAdd(&e_bonded[t.r], (unsigned long long)to_fixed_d(e, ENERGY_SCALE));
}

// cosine dihedral: energy as TorsionPotential.cpp:53-70; force = analytic gradient of the same
// angle (closed form of the IUPAC dihedral gradient, phi_ref = -psi - pi), zero when |cos phi| < 1e-4
__global__ void k_dihedrals(Grid cur, ReplicaDev *reps, Topo topo, const BondTerm *terms, uint32_t n,
                            unsigned long long *e_bonded) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const BondTerm t = terms[i];
    const ReplicaDev &R = reps[t.r];
    if (!R.in_use) return;
    const double SMALL = 1e-4;
    double pi_[3], pj[3], pk[3], pl[3], xji[3], xkj[3], xkl[3], xjk[3];
    pos_of(cur, topo, R, t.h[0], pi_);
    pos_of(cur, topo, R, t.h[1], pj);
    pos_of(cur, topo, R, t.h[2], pk);
    pos_of(cur, topo, R, t.h[3], pl);
    sdiff(pj, pi_, R, xji);
    sdiff(pk, pj, R, xkj);
    sdiff(pk, pl, R, xkl);
    for (int d = 0; d < 3; ++d) xjk[d] = -1. * xkj[d];
    double m[3], nn[3], mxn[3];
    cross(xji, xkj, m);
    cross(xkl, xjk, nn);
    cross(m, nn, mxn);
    double jkn = sqrt(dot(xjk, xjk));
    if (jkn < SMALL) jkn = SMALL;
    double mnn = sqrt(dot(m, m)) * sqrt(dot(nn, nn));
    if (mnn < SMALL) mnn = SMALL;
    double phi_e = -atan2(dot(mxn, xjk) / (mnn * jkn), dot(m, nn) / mnn);
    double e = t.k * (1 + cos((double)t.mult * phi_e - t.p0));
    atomicAdd(&e_bonded[t.r], (unsigned long long)to_fixed_d(e, ENERGY_SCALE));

    double x2 = dot(xjk, xjk);
    if (x2 < SMALL) x2 = SMALL;
    double m2 = dot(m, m);
    if (m2 < SMALL) m2 = SMALL;
    double n2 = dot(nn, nn);
    if (n2 < SMALL) n2 = SMALL;
    double mn = sqrt(m2) * sqrt(n2);
    if (mn < SMALL) mn = SMALL;
    double cos_phi = dot(nn, m) / mn;
    if (fabs(cos_phi) < SMALL) return;
    double sin_phi = dot(mxn, xjk) / (mn * sqrt(x2));
    double phi = -atan2(sin_phi, cos_phi);
    double dV = -t.k * (double)t.mult * sin((double)t.mult * phi - t.p0);
    double b1[3], b2[3], b3[3], A[3], B[3];
    for (int d = 0; d < 3; ++d) {
        b1[d] = -xji[d];
        b2[d] = xjk[d];
        b3[d] = xkl[d];
    }
    cross(b1, b2, A);
    cross(b2, b3, B);
    double A2 = dot(A, A), B2 = dot(B, B);
    if (A2 < SMALL) A2 = SMALL;
    if (B2 < SMALL) B2 = SMALL;
    double b2n = sqrt(x2);
    double p12 = dot(b1, b2) / x2, p32 = dot(b3, b2) / x2;
    double fi[3], fj[3], fk[3], fl[3];
    for (int d = 0; d < 3; ++d) {
        double gi = -b2n * A[d] / A2, gl = b2n * B[d] / B2;
        fi[d] = dV * gi;
        fj[d] = dV * (-(1.0 + p12) * gi + p32 * gl);
        fk[d] = dV * (p12 * gi - (1.0 + p32) * gl);
        fl[d] = dV * gl;
    }
    add_fb(topo, R, t.h[0], fi);
    add_fb(topo, R, t.h[1], fj);
    add_fb(topo, R, t.h[2], fk);
    add_fb(topo, R, t.h[3], fl);
}

// -------------------------------------------------------------------------------------------------
// reaction candidate search (CPUUncontrolledApproximation.cpp:70-133, ReactionUtils.h:83-144)
// Each unordered pair is found once, by the particle with the smaller (type, bin slot).  Acceptance
// u < 1 - exp(-rate dt) is drawn per candidate from Philox keyed by the pair's handles, so the
// candidate set and the draws do not depend on slots, thread order or batching.
// -------------------------------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long event_key(int handler, const uint32_t x4[4], double rate) {
    if (handler == 0) return ((unsigned long long)x4[2] << 32) | x4[3];
    double u = u53(x4[2], x4[3]) + 0x1p-54;
    double k = -log(u) / rate;  // Efraimidis-Spirakis key, positive: bit order == numeric order
    return (unsigned long long)__double_as_longlong(k);
}

// candidates of the particle in slot s (searching from the smaller (type, bin slot) of every pair)
__device__ __forceinline__ void search_one(const Grid &cur, ReplicaDev *reps, const Tables &T, uint32_t cap, int radius,
                           int handler, int candidates_only, Event *events, uint32_t *ev_count, uint32_t *cand_rows,
                           uint32_t cand_cap, uint32_t s, const State &me, uint32_t *err) {
    const uint32_t r = tw_replica(me.tw);
    const ReplicaDev &R = reps[r];
    const uint32_t t = tw_type(me.tw);
    const uint32_t step = R.step;
    const uint64_t seed = R.seed;
    if (handler == HANDLER_DETAILED_BALANCE) handler = 1;  // Gillespie keys for every replica
    else if (handler <= 1 && R.handler >= 0) handler = R.handler;  // the replica's own handler

    auto emit = [&](uint32_t rx, uint32_t s1, uint32_t s2, uint32_t h1, uint32_t h2, unsigned long long key,
                    uint32_t ne) {
        if (candidates_only) {
            uint32_t p = atomicAdd(&ev_count[0], 1u);
            if (p < cand_cap) {
                cand_rows[4 * p] = ne;
                cand_rows[4 * p + 1] = h1;
                cand_rows[4 * p + 2] = h2;
                cand_rows[4 * p + 3] = rx;
            }
            return;
        }
        uint32_t p = atomicAdd(&ev_count[r], 1u);
        if (p < R.ev_cap) {
            Event ev;
            ev.key = key;
            ev.s1 = s1;
            ev.s2 = s2;
            ev.h1 = h1;
            ev.h2 = h2;
            ev.rx = rx;
            ev.pad = 0;
            events[R.ev_base + p] = ev;
        } else {
            atomicAdd(&err[ERR_EVENT_OVERFLOW], 1u);
        }
    };

    // order 1
    for (uint32_t k = T.o1idx[t]; k < T.o1idx[t + 1]; ++k) {
        const uint32_t rx = T.o1list[k];
        const RxDev X = T.rx[rx];
        if (!(X.rate > 0)) continue;
        uint32_t x4[4];
        philox(me.h, step, 1u | (rx << 8), 0u, seed, x4);
        if (!candidates_only && !(u53(x4[0], x4[1]) < X.p_accept)) continue;
        emit(rx, s, s, me.h, me.h, event_key(handler, x4, X.rate), 1u);
    }
    // order 2
    if (!T.has_o2 || (me.tw & TW_OUTSIDE)) return;
    const RepGeom geo = load_geom(R);
    const RxPairIdx *orow = T.o2idx + (size_t)t * T.n_types;
    const uint32_t cl = cell_of_slot(cur, s, cap) - geo.cell_base;
    for_each_stencil_slot(geo, cur, cap, radius, cl, [&](uint32_t s2) {
        if (s2 == s) return;
        const State o = cur.bins[s2];
        if (o.tw & (TW_DEAD | TW_OUTSIDE)) return;
        const uint32_t t2 = tw_type(o.tw);
        if (t2 < t || (t2 == t && s2 < s)) return;
        const RxPairIdx pi = orow[t2];
        if (pi.count == 0) return;
        // distSquared(neighbor.pos, entry.pos) (CPUUncontrolledApproximation.cpp:108)
        const double dx = min_image(me.x - o.x, geo.box[0], geo.pbc[0]);
        const double dy = min_image(me.y - o.y, geo.box[1], geo.pbc[1]);
        const double dz = min_image(me.z - o.z, geo.box[2], geo.pbc[2]);
        const double d2 = dist2(dx, dy, dz);
        const uint32_t lo = me.h < o.h ? me.h : o.h, hi = me.h < o.h ? o.h : me.h;
        for (uint32_t q = 0; q < pi.count; ++q) {
            const uint32_t rx = T.o2list[pi.start + q];
            const RxDev X = T.rx[rx];
            if (!(X.rate > 0 && d2 < X.educt_dist2)) continue;
            uint32_t x4[4];
            philox(lo, step, 2u | (rx << 8), hi, seed, x4);
            if (!candidates_only && !(u53(x4[0], x4[1]) < X.p_accept)) continue;
            // entry1: the educt of type educts[0]; equal educt types: the lower handle
            bool me_first;
            if (X.educts[0] != X.educts[1]) me_first = (t == X.educts[0]);
            else me_first = me.h < o.h;
            if (me_first) emit(rx, s, s2, me.h, o.h, event_key(handler, x4, X.rate), 2u);
            else emit(rx, s2, s, o.h, me.h, event_key(handler, x4, X.rate), 2u);
        }
    });
}

__global__ __launch_bounds__(256) void k_search(Grid cur, ReplicaDev *reps, Tables T, uint32_t cap, int radius,
                                                 int handler, int candidates_only, Event *events,
                                                 uint32_t *ev_count, uint32_t *cand_rows, uint32_t cand_cap,
                                                 uint32_t only_replica, uint32_t max_wl, uint32_t *err) {
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t i = b * blockDim.x + threadIdx.x;
    uint32_t n = *cur.wl_count;
    n = n < max_wl ? n : max_wl;
    if (i >= n) return;
    const uint32_t s = cur.wl[i];
    const State me = cur.bins[s];
    if (me.tw & TW_DEAD) return;
    const uint32_t r = tw_replica(me.tw);
    if (candidates_only && r != only_replica) return;
    if (!candidates_only && !reps[r].active) return;
    search_one(cur, reps, T, cap, radius, handler, candidates_only, events, ev_count, cand_rows, cand_cap, s, me, err);
}

// -------------------------------------------------------------------------------------------------
// conflict resolution + reaction execution, one workgroup per replica.
// Both handlers reduce to a greedy maximal independent set over the accepted candidates in
// ascending key order (UA: uniform keys = the shuffle of CPUUncontrolledApproximation.cpp:198-253;
// Gillespie: Efraimidis-Spirakis keys = the weighted selection order of algorithm.h:62-136).  The
// greedy result is computed with parallel rounds (an event wins when it holds the smallest index
// on all of its particles among undecided events), which equals the sequential scan.
// Products follow performReaction (ReactionUtils.h:146-227); fresh handles are assigned in key order.
// -------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t hslot(uint32_t key, uint32_t mask) { return (key * 2654435761u) & mask; }

// With edits (the fused step), resolve lists every record it kills, inserts or retypes, and flags
// inserted and retyped records TW_NEW; k_fix_sites / k_fix_advance (rdy_step.hip) then bring the
// other buffer to the post-reaction state.
enum : uint32_t { EDIT_NONE = 0, EDIT_KILLED = 1, EDIT_INSERTED = 2, EDIT_RETYPED = 3 };

__device__ __forceinline__ void fixed_add3(long long *g, long long x, long long y, long long z) {
    atomicAdd((unsigned long long *)g, (unsigned long long)x);
    atomicAdd((unsigned long long *)g + 1, (unsigned long long)y);
    atomicAdd((unsigned long long *)g + 2, (unsigned long long)z);
}

// the k-th cell of the de-duplicated stencil axis of cell i (the set axis_cells lists)
__device__ __forceinline__ int axis_count(int i, int D, int p, int r) {
    if (p) return D <= 2 * r ? D : 2 * r + 1;
    return min(D - 1, i + r) - max(0, i - r) + 1;
}
__device__ __forceinline__ int axis_elem(int k, int i, int D, int p, int r) {
    if (p) {
        if (D <= 2 * r) return k;
        const int j = i - r + k;  // |overhang| <= r < D
        return j < 0 ? j + D : (j >= D ? j - D : j);
    }
    return max(0, i - r) + k;
}

// -------------------------------------------------------------------------------------------------
// DetailedBalance (SCPUReactionImpls.cpp:407-555): the events the Gillespie selection keeps are
// performed one after another in key order, each on the state the previous ones left.  An event
// outside every reversible pair is performed as the Gillespie handler performs it.  An event of a
// reversible pair is a Metropolis move: its products are proposed (a fission at a distance drawn
// from the pair's fission-distance table), the workgroup sums the change of the pair-potential
// energy, and the move is accepted with probability
//   min(1, prefactor exp(-(dE - dU_int) / kBT))
// where dU_int is the educts' (fusion, negated) or products' (fission) mutual interaction energy for
// FusionFission and zero otherwise, and prefactor is the acceptance prefactor for the forward
// reaction of an EnzymaticEnzymatic pair and one otherwise (the reference tests the forward educt
// order twice, so the backward reaction gets one).  A rejected move puts the educts back as new
// particles: their positions and types are kept and they get the ids after the products' ids.
// Energies are the FP64 pair potentials gated by d^2 < cutoff^2, between particles inside the box.
// Random numbers: the fission direction on stream 3 (as every fission), the fission distance on
// stream 10, the Metropolis uniform on stream 11 | reaction << 8 keyed by the educts' handles.
// -------------------------------------------------------------------------------------------------
template <class Perform>
__device__ __forceinline__ void detailed_balance_events(const Grid &cur, const ReplicaDev &R, const Tables &T,
                                                        const DbTables &D, uint32_t cap, uint32_t ne,
                                                        const uint8_t *status, const uint32_t *sidx,
                                                        const uint32_t *prefix, const Event *ev, uint32_t hbase,
                                                        const RepGeom &geo, RxRecord *rrec, double *red,
                                                        Perform &perform) {
    const int tid = threadIdx.x;
    const uint32_t nt = T.n_types;
    auto pair_u = [&](uint32_t ta, uint32_t tb, double d2) -> double {
        const uint32_t pi = D.pidx[(size_t)ta * nt + tb];
        double u = 0.0;
        for (uint32_t p = pi & 0xFFFFu, pe = (pi & 0xFFFFu) + (pi >> 16); p < pe; ++p) {
            const DbPot &P = D.pots[p];
            if (d2 < P.cut2) u += rdy_rev::pair_energy_fp64(P.kind, P.q, sqrt(d2));
        }
        return u;
    };
    auto inside = [&](const double *q) {
        return in_box_axis(q[0], geo.box[0]) && in_box_axis(q[1], geo.box[1]) && in_box_axis(q[2], geo.box[2]);
    };
    const int rad = D.radius;
    const uint32_t side = 2u * (uint32_t)rad + 1u, S = side * side * side;  // stencil cells per center
    for (uint32_t k = 0; k < ne; ++k) {
        if (status[k] != 1) {
            if (rrec && tid == 0) rrec[k].valid = 0u;
            continue;
        }
        const Event e = ev[sidx[k]];
        const RxDev X = T.rx[e.rx];
        const int32_t rk = D.kind[e.rx];
        if (rk < 0) {
            if (tid == 0) perform(k, e, X.product_distance);
            __syncthreads();
            continue;
        }
        const State a = cur.bins[e.s1], b = cur.bins[e.s2];
        // centers 0, 1: educts that leave; 2, 3: proposed products
        double cp[4][3];
        uint32_t ct[4];
        bool cv[4] = {false, false, false, false};
        uint32_t ex0 = e.s1, ex1 = e.s1, educt_slot = e.s1;
        double fdist = X.product_distance, du_int = 0.0;
        auto set = [&](int c, double x, double y, double z, uint32_t t, bool in) {
            cp[c][0] = x;
            cp[c][1] = y;
            cp[c][2] = z;
            ct[c] = t;
            cv[c] = in;
        };
        switch (X.kind) {
        case RX_FISSION: {  // SCPUReactionImpls.cpp:600-633
            uint32_t x4[4];
            double nv[3];
            philox(a.h, R.step, 3u, 0u, R.seed, x4);
            normals3(x4, nv);
            const double nrm = sqrt(dist2(nv[0], nv[1], nv[2]));
            for (int d = 0; d < 3; ++d) nv[d] /= nrm;
            philox(a.h, R.step, 10u, 0u, R.seed, x4);
            const double u = u53(x4[0], x4[1]);
            const double *cdf = D.cdf + D.table[e.rx];
            uint32_t lo = 0, hi = D.n_table;  // lower_bound (DetailedBalance.h:133-138)
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (cdf[mid] < u) lo = mid + 1;
                else hi = mid;
            }
            if (lo >= D.n_table) lo = D.n_table - 1;
            fdist = D.radii[D.table[e.rx] + lo];
            du_int = pair_u(X.products[0], X.products[1], fdist * fdist);
            set(0, a.x, a.y, a.z, tw_type(a.tw), !(a.tw & TW_OUTSIDE));
            double q2[3], q1[3];
            const double p[3] = {a.x, a.y, a.z};
            for (int d = 0; d < 3; ++d) {
                q2[d] = wrap(p[d] - X.w2 * fdist * nv[d], geo.box[d], geo.pbc[d]);
                q1[d] = wrap(p[d] + X.w1 * fdist * nv[d], geo.box[d], geo.pbc[d]);
            }
            set(2, q2[0], q2[1], q2[2], X.products[1], inside(q2));
            set(3, q1[0], q1[1], q1[2], X.products[0], inside(q1));
            break;
        }
        case RX_FUSION: {  // SCPUReactionImpls.cpp:634-681
            const double dx = min_image(b.x - a.x, geo.box[0], geo.pbc[0]);
            const double dy = min_image(b.y - a.y, geo.box[1], geo.pbc[1]);
            const double dz = min_image(b.z - a.z, geo.box[2], geo.pbc[2]);
            du_int = -pair_u(X.educts[0], X.educts[1], dist2(dx, dy, dz));
            const double w = (X.educts[0] == tw_type(a.tw)) ? X.w1 : X.w2;
            double q[3] = {wrap(a.x + w * dx, geo.box[0], geo.pbc[0]), wrap(a.y + w * dy, geo.box[1], geo.pbc[1]),
                           wrap(a.z + w * dz, geo.box[2], geo.pbc[2])};
            set(0, a.x, a.y, a.z, tw_type(a.tw), !(a.tw & TW_OUTSIDE));
            set(1, b.x, b.y, b.z, tw_type(b.tw), !(b.tw & TW_OUTSIDE));
            set(2, q[0], q[1], q[2], X.products[0], inside(q));
            ex1 = e.s2;
            break;
        }
        case RX_ENZYMATIC: {  // SCPUReactionImpls.cpp:703-735: the educt that is not the catalyst converts
            const bool a_cat = tw_type(a.tw) == X.educts[1];
            const State &c = a_cat ? b : a;
            educt_slot = ex0 = ex1 = a_cat ? e.s2 : e.s1;
            set(0, c.x, c.y, c.z, tw_type(c.tw), !(c.tw & TW_OUTSIDE));
            set(2, c.x, c.y, c.z, X.products[0], !(c.tw & TW_OUTSIDE));
            break;
        }
        default: {  // RX_CONVERSION, SCPUReactionImpls.cpp:683-702
            set(0, a.x, a.y, a.z, tw_type(a.tw), !(a.tw & TW_OUTSIDE));
            set(2, a.x, a.y, a.z, X.products[0], !(a.tw & TW_OUTSIDE));
            break;
        }
        }
        // energy change: each center against every other particle inside the box within the pair cutoffs
        // (the leaving educts excluded), plus the educts' and the products' mutual pair once
        double acc = 0.0;
        for (uint32_t w = tid; w < 4u * S; w += blockDim.x) {
            const uint32_t c = w / S, j = w - c * S;
            if (!cv[c]) continue;
            int ci[3], na[3];
            for (int d = 0; d < 3; ++d) {
                ci[d] = cell_coord(cp[c][d], geo.box[d], geo.cs[d], geo.dims[d]);
                na[d] = axis_count(ci[d], geo.dims[d], geo.pbc[d], rad);
            }
            if (j >= (uint32_t)(na[0] * na[1] * na[2])) continue;
            const int jz = (int)(j % (uint32_t)na[2]), jy = (int)((j / (uint32_t)na[2]) % (uint32_t)na[1]),
                      jx = (int)(j / ((uint32_t)na[2] * (uint32_t)na[1]));
            const uint32_t cell =
                geo.cell_base +
                ((uint32_t)axis_elem(jx, ci[0], geo.dims[0], geo.pbc[0], rad) * (uint32_t)geo.dims[1] +
                 (uint32_t)axis_elem(jy, ci[1], geo.dims[1], geo.pbc[1], rad)) *
                    (uint32_t)geo.dims[2] +
                (uint32_t)axis_elem(jz, ci[2], geo.dims[2], geo.pbc[2], rad);
            const double sign = c < 2 ? -1.0 : 1.0;
            auto visit = [&](uint32_t slot) {
                if (slot == ex0 || slot == ex1) return;
                const State o = cur.bins[slot];
                if (o.tw & (TW_DEAD | TW_OUTSIDE)) return;
                const double dx = min_image(o.x - cp[c][0], geo.box[0], geo.pbc[0]);
                const double dy = min_image(o.y - cp[c][1], geo.box[1], geo.pbc[1]);
                const double dz = min_image(o.z - cp[c][2], geo.box[2], geo.pbc[2]);
                const double u = pair_u(ct[c], tw_type(o.tw), dist2(dx, dy, dz));
                if (u != 0.0) acc += sign * u;
            };
            const uint32_t n = cur.counts[cell], m = n < cap ? n : cap;
            for (uint32_t q = 0; q < m; ++q) visit(cell * cap + q);
            if (n > cap) {
                uint32_t no = *cur.ovf_count;
                no = no < cur.ovf_cap ? no : cur.ovf_cap;
                for (uint32_t o = 0; o < no; ++o)
                    if (cur.ovf_cell[o] == cell) visit(cur.nbin + o);
            }
        }
        if (tid == 0) {
            for (int g = 0; g < 4; g += 2)
                if (cv[g] && cv[g + 1]) {
                    const double dx = min_image(cp[g + 1][0] - cp[g][0], geo.box[0], geo.pbc[0]);
                    const double dy = min_image(cp[g + 1][1] - cp[g][1], geo.box[1], geo.pbc[1]);
                    const double dz = min_image(cp[g + 1][2] - cp[g][2], geo.box[2], geo.pbc[2]);
                    acc += (g == 0 ? -1.0 : 1.0) * pair_u(ct[g], ct[g + 1], dist2(dx, dy, dz));
                }
        }
        red[tid] = acc;
        __syncthreads();
        for (uint32_t h = blockDim.x >> 1; h > 0; h >>= 1) {
            if ((uint32_t)tid < h) red[tid] += red[tid + h];
            __syncthreads();
        }
        const double dE = red[0];
        const double boltzmann = exp(-1.0 / T.kbt * (dE - du_int));
        const double acceptance = fmin(1.0, D.prefactor[e.rx] * boltzmann);
        uint32_t x4[4];
        const uint32_t lo = e.h1 < e.h2 ? e.h1 : e.h2, hi = e.h1 < e.h2 ? e.h2 : e.h1;
        philox(lo, R.step, 11u | (e.rx << 8), hi, R.seed, x4);
        const bool accept = u53(x4[0], x4[1]) < acceptance;
        __syncthreads();  // red[0] read by every thread before the next event writes red
        if (tid == 0) {
            if (accept) {
                perform(k, e, fdist);
                if (rrec) {  // the reversible path fills no products (SCPUReactionImpls.cpp:579-742)
                    RxRecord &rc = rrec[k];
                    rc.products[0] = rc.products[1] = 0u;
                    if (X.kind == RX_ENZYMATIC && educt_slot == e.s2) {  // educt first, then the catalyst
                        rc.educts[0] = b.h;
                        rc.educts[1] = a.h;
                        rc.types_from[0] = tw_type(b.tw);
                        rc.types_from[1] = tw_type(a.tw);
                    }
                }
            } else {
                if (rrec) rrec[k].valid = 0u;
                const uint32_t h0 = hbase + prefix[k] + (X.kind == RX_FISSION ? 2u : 1u);
                if (X.kind == RX_FUSION) {
                    cur.bins[e.s1].h = h0;
                    cur.bins[e.s2].h = h0 + 1u;
                } else {
                    cur.bins[educt_slot].h = h0;
                }
            }
        }
        __syncthreads();
    }
}

// the reactions of replica r = blockIdx.x (one workgroup of RESOLVE_THREADS).  DB: the DetailedBalance
// handler (tables D), which performs the selected events one after another.
template <bool DB = false>
__device__ __forceinline__ void resolve_replica(const Grid &cur, ReplicaDev *reps, const Tables &T, uint32_t cap,
                                                Event *events, uint32_t *ev_count, unsigned long long *rx_step,
                                                unsigned long long *rx_total, int require_active, uint32_t max_wl,
                                                Edit *edits, uint32_t *edit_count, uint32_t edit_cap,
                                                RxRecord *recs, uint32_t *rec_count, uint32_t *err,
                                                const DbTables D = DbTables{}) {
    constexpr int E = RESOLVE_MAX_EVENTS;
    constexpr int HT = 4 * E;  // hash table entries (power of two, > 2 slots per event)
    __shared__ unsigned long long skey[E];
    __shared__ uint32_t sidx[E];
    __shared__ uint32_t hkeys[HT];
    __shared__ uint32_t hmin[HT];
    __shared__ uint8_t htaken[HT];
    __shared__ uint8_t status[E];   // 0 undecided, 1 won, 2 lost
    __shared__ uint32_t ent1[E], ent2[E];
    uint32_t *prefix = hmin;  // free once the independent set is decided
    __shared__ uint32_t s_undecided, s_total;
    __shared__ uint32_t rxcount[256];

    const uint32_t r = blockIdx.x;
    ReplicaDev &R = reps[r];
    const int tid = threadIdx.x;
    if (!R.in_use || (require_active && !R.active)) {
        if (tid == 0) ev_count[r] = 0;
        return;
    }
    uint32_t ne = ev_count[r];
    __syncthreads();
    if (tid < (int)T.n_rx && tid < 256) {
        rx_step[(size_t)r * T.n_rx + tid] = 0ull;
        rxcount[tid] = 0;
    }
    if (ne == 0) {
        if (tid == 0) {
            ev_count[r] = 0;
            if (recs) rec_count[r] = 0;
        }
        return;
    }
    if (ne > R.ev_cap) ne = R.ev_cap;
    if (ne > (uint32_t)E) {
        if (tid == 0) atomicAdd(&err[ERR_EVENT_LDS], 1u);
        ne = E;
    }
    const Event *ev = events + R.ev_base;
    // load and pad to a power of two
    uint32_t np2 = 1;
    while (np2 < ne) np2 <<= 1;
    for (uint32_t k = tid; k < np2; k += blockDim.x) {
        skey[k] = k < ne ? ev[k].key : ~0ull;
        sidx[k] = k;
    }
    __syncthreads();
    // sort of (key, rx, h1, h2) ascending.  Keys (Philox words) are distinct but for a vanishing
    // chance: the rank of each key among all keys (every thread reads the same key at once) places
    // it directly; equal keys, detected by a doubly taken rank, fall back to a bitonic network.
    for (uint32_t k = tid; k < ne; k += blockDim.x) hkeys[k] = 0u;
    __syncthreads();
    for (uint32_t k = tid; k < ne; k += blockDim.x) {
        const unsigned long long key = skey[k];
        uint32_t rank = 0;
        for (uint32_t j = 0; j < ne; ++j) rank += skey[j] < key ? 1u : 0u;
        hmin[k] = rank;
        atomicAdd(&hkeys[rank], 1u);
    }
    if (tid == 0) s_undecided = 0;
    __syncthreads();
    for (uint32_t k = tid; k < ne; k += blockDim.x)
        if (hkeys[k] != 1u) s_undecided = 1;  // an equal key
    __syncthreads();
    const bool ranked = s_undecided == 0;  // block-uniform
    if (ranked) {
        for (uint32_t k = tid; k < ne; k += blockDim.x) hkeys[hmin[k]] = k;
        __syncthreads();
        for (uint32_t k = tid; k < ne; k += blockDim.x) sidx[k] = hkeys[k];
        __syncthreads();
    }
    auto less = [&](uint32_t a, uint32_t b) -> bool {
        if (skey[a] != skey[b]) return skey[a] < skey[b];
        uint32_t ia = sidx[a], ib = sidx[b];
        if (ia >= ne || ib >= ne) return ia < ib;
        const Event &ea = ev[ia], &eb = ev[ib];
        if (ea.rx != eb.rx) return ea.rx < eb.rx;
        if (ea.h1 != eb.h1) return ea.h1 < eb.h1;
        return ea.h2 < eb.h2;
    };
    for (uint32_t kk = 2; !ranked && kk <= np2; kk <<= 1) {
        for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
            for (uint32_t a = tid; a < np2; a += blockDim.x) {
                uint32_t bb = a ^ j;
                if (bb > a) {
                    bool up = (a & kk) == 0;
                    bool sw = up ? less(bb, a) : less(a, bb);
                    if (sw) {
                        unsigned long long tk = skey[a];
                        skey[a] = skey[bb];
                        skey[bb] = tk;
                        uint32_t ti = sidx[a];
                        sidx[a] = sidx[bb];
                        sidx[bb] = ti;
                    }
                }
            }
            __syncthreads();
        }
    }
    // hash table of particle slots
    for (uint32_t k = tid; k < (uint32_t)HT; k += blockDim.x) {
        hkeys[k] = 0xFFFFFFFFu;
        htaken[k] = 0;
    }
    __syncthreads();
    auto insert = [&](uint32_t slot) -> uint32_t {
        uint32_t h = hslot(slot, HT - 1);
        while (true) {
            uint32_t prev = atomicCAS(&hkeys[h], 0xFFFFFFFFu, slot);
            if (prev == 0xFFFFFFFFu || prev == slot) return h;
            h = (h + 1) & (HT - 1);
        }
    };
    for (uint32_t k = tid; k < ne; k += blockDim.x) {
        const Event &e = ev[sidx[k]];
        ent1[k] = insert(e.s1);
        ent2[k] = (e.s2 == e.s1) ? ent1[k] : insert(e.s2);
        status[k] = 0;
    }
    __syncthreads();
    // rounds of the parallel greedy MIS (index in sorted order = priority)
    for (int round = 0; round < 4096; ++round) {
        for (uint32_t k = tid; k < (uint32_t)HT; k += blockDim.x) hmin[k] = 0xFFFFFFFFu;
        if (tid == 0) s_undecided = 0;
        __syncthreads();
        for (uint32_t k = tid; k < ne; k += blockDim.x)
            if (status[k] == 0) {
                atomicMin(&hmin[ent1[k]], k);
                atomicMin(&hmin[ent2[k]], k);
            }
        __syncthreads();
        for (uint32_t k = tid; k < ne; k += blockDim.x)
            if (status[k] == 0 && hmin[ent1[k]] == k && hmin[ent2[k]] == k) {
                status[k] = 1;
                htaken[ent1[k]] = 1;
                htaken[ent2[k]] = 1;
            }
        __syncthreads();
        for (uint32_t k = tid; k < ne; k += blockDim.x)
            if (status[k] == 0) {
                if (htaken[ent1[k]] || htaken[ent2[k]]) status[k] = 2;
                else atomicAdd(&s_undecided, 1u);
            }
        __syncthreads();
        if (s_undecided == 0) break;
        __syncthreads();
    }
    // fresh handles in key order: exclusive scan of products needing a new id
    for (uint32_t k = tid; k < ne; k += blockDim.x) {
        uint32_t c = 0;
        if (status[k] == 1) {
            const uint32_t rx = ev[sidx[k]].rx;
            const int kind = T.rx[rx].kind;
            c = (kind == RX_FISSION) ? 2u : (kind == RX_DECAY ? 0u : 1u);
            // a reversible event takes ids for its products and then for the educts' backup particles,
            // whether or not it is accepted (SCPUReactionImpls.cpp:447-461, 557-577)
            if (DB && D.kind[rx] >= 0) c = (kind == RX_FISSION || kind == RX_FUSION) ? 3u : 2u;
        }
        prefix[k] = c;
    }
    __syncthreads();
    {  // exclusive scan: each thread a contiguous chunk of events
        const uint32_t per = (ne + blockDim.x - 1) / blockDim.x, k0 = tid * per;
        uint32_t sum = 0;
        for (uint32_t k = k0; k < k0 + per && k < ne; ++k) sum += prefix[k];
        uint32_t tot;
        uint32_t acc = block_exclusive_scan(sum, &tot);
        for (uint32_t k = k0; k < k0 + per && k < ne; ++k) {
            const uint32_t c = prefix[k];
            prefix[k] = acc;
            acc += c;
        }
        if (tid == 0) s_total = tot;
    }
    __syncthreads();
    const uint32_t hbase = R.handle_next;
    const RepGeom geo = load_geom(R);
    // the fused step lists the changed records (edits) and flags inserted / retyped ones TW_NEW
    const bool fix = edits != nullptr;
    const uint32_t tw_new = fix ? TW_NEW : 0u;
    auto edit = [&](uint32_t slot, uint32_t kind, uint32_t old_type, uint32_t old_handle) {
        if (!fix || slot == 0xFFFFFFFFu) return;
        const uint32_t p = atomicAdd(edit_count, 1u);
        if (p < edit_cap) edits[p] = Edit{slot, kind | (old_type << 2), r, old_handle};
        else atomicAdd(&err[ERR_EDIT_OVERFLOW], 1u);
    };
    RxRecord *rrec = recs ? recs + (size_t)r * E : nullptr;
    // performs sorted event k (fission products fdist apart)
    auto perform = [&](uint32_t k, const Event &e, double fdist) {
        const RxDev X = T.rx[e.rx];
        atomicAdd(&rxcount[e.rx & 255u], 1u);
        State a = cur.bins[e.s1];
        const uint32_t flags_keep = a.tw & TW_TOPO;
        const uint32_t hnew = hbase + prefix[k];
        if (rrec) {
            const State b = cur.bins[e.s2];
            RxRecord rc;
            rc.rx = e.rx;
            rc.kind = (uint32_t)X.kind;
            rc.educts[0] = a.h;
            rc.educts[1] = b.h;
            rc.types_from[0] = tw_type(a.tw);
            rc.types_from[1] = tw_type(b.tw);
            rc.where[0] = wrap((a.x + b.x) / 2., geo.box[0], geo.pbc[0]);
            rc.where[1] = wrap((a.y + b.y) / 2., geo.box[1], geo.pbc[1]);
            rc.where[2] = wrap((a.z + b.z) / 2., geo.box[2], geo.pbc[2]);
            // products as the record holds them after the reaction (ReactionUtils.h:170-224)
            rc.products[0] = rc.products[1] = 0u;
            if (X.kind == RX_CONVERSION || X.kind == RX_FUSION) rc.products[0] = hnew;
            if (X.kind == RX_FISSION) {
                rc.products[0] = hnew + 1u;  // the educt's record, renamed after the new particle's id
                rc.products[1] = hnew;
            }
            if (X.kind == RX_ENZYMATIC) {
                const bool a_is_catalyst = tw_type(a.tw) == X.educts[1];
                rc.products[0] = a_is_catalyst ? a.h : hnew;
                rc.products[1] = a_is_catalyst ? hnew : b.h;
            }
            rc.valid = 1u;
            rc.pad = 0u;
            rrec[k] = rc;
        }
        switch (X.kind) {
        case RX_DECAY:
            cur.bins[e.s1].tw = a.tw | TW_DEAD;
            edit(e.s1, EDIT_KILLED, 0, 0);
            break;
        case RX_CONVERSION: {
            State nw = a;
            nw.tw = make_tw(X.products[0], flags_keep, r) | (a.tw & TW_OUTSIDE) | tw_new;
            nw.h = hnew;
            cur.bins[e.s1] = nw;
            edit(e.s1, EDIT_RETYPED, tw_type(a.tw), a.h);
            break;
        }
        case RX_ENZYMATIC: {
            // ReactionUtils.h:174-190: the educt that is not the catalyst converts
            State bst = cur.bins[e.s2];
            if (tw_type(a.tw) == X.educts[1]) {
                edit(e.s2, EDIT_RETYPED, tw_type(bst.tw), bst.h);
                bst.tw = make_tw(X.products[0], bst.tw & TW_TOPO, r) | (bst.tw & TW_OUTSIDE) | tw_new;
                bst.h = hnew;
                cur.bins[e.s2] = bst;
            } else {
                edit(e.s1, EDIT_RETYPED, tw_type(a.tw), a.h);
                a.tw = make_tw(X.products[0], flags_keep, r) | (a.tw & TW_OUTSIDE) | tw_new;
                a.h = hnew;
                cur.bins[e.s1] = a;
            }
            break;
        }
        case RX_FISSION: {
            uint32_t x4[4];
            double nv[3];
            philox(a.h, R.step, 3u, 0u, R.seed, x4);
            normals3(x4, nv);
            double nrm = sqrt(dist2(nv[0], nv[1], nv[2]));
            for (int d = 0; d < 3; ++d) nv[d] /= nrm;
            const double p[3] = {a.x, a.y, a.z};
            State p2, p1;
            double q2[3], q1[3];
            for (int d = 0; d < 3; ++d) {
                q2[d] = wrap(p[d] - X.w2 * fdist * nv[d], geo.box[d], geo.pbc[d]);
                q1[d] = wrap(p[d] + X.w1 * fdist * nv[d], geo.box[d], geo.pbc[d]);
            }
            p2.x = q2[0]; p2.y = q2[1]; p2.z = q2[2];
            p2.tw = make_tw(X.products[1], tw_new, r);
            p2.h = hnew;
            p1.x = q1[0]; p1.y = q1[1]; p1.z = q1[2];
            p1.tw = make_tw(X.products[0], tw_new, r);
            p1.h = hnew + 1u;
            cur.bins[e.s1].tw = a.tw | TW_DEAD;
            edit(e.s1, EDIT_KILLED, 0, 0);
            edit(bin_insert(cur, geo, cap, p2, err, true, max_wl), EDIT_INSERTED, 0, 0);
            edit(bin_insert(cur, geo, cap, p1, err, true, max_wl), EDIT_INSERTED, 0, 0);
            break;
        }
        case RX_FUSION: {
            const State bst = cur.bins[e.s2];
            const double dx = min_image(bst.x - a.x, geo.box[0], geo.pbc[0]);
            const double dy = min_image(bst.y - a.y, geo.box[1], geo.pbc[1]);
            const double dz = min_image(bst.z - a.z, geo.box[2], geo.pbc[2]);
            const double w = (X.educts[0] == tw_type(a.tw)) ? X.w1 : X.w2;
            State pr;
            pr.x = wrap(a.x + w * dx, geo.box[0], geo.pbc[0]);
            pr.y = wrap(a.y + w * dy, geo.box[1], geo.pbc[1]);
            pr.z = wrap(a.z + w * dz, geo.box[2], geo.pbc[2]);
            pr.tw = make_tw(X.products[0], tw_new, r);
            pr.h = hnew;
            cur.bins[e.s1].tw = a.tw | TW_DEAD;
            cur.bins[e.s2].tw = bst.tw | TW_DEAD;
            edit(e.s1, EDIT_KILLED, 0, 0);
            edit(e.s2, EDIT_KILLED, 0, 0);
            edit(bin_insert(cur, geo, cap, pr, err, true, max_wl), EDIT_INSERTED, 0, 0);
            break;
        }
        }
    };
    if constexpr (!DB) {
        for (uint32_t k = tid; k < ne; k += blockDim.x) {
            if (status[k] != 1) {
                if (rrec) rrec[k].valid = 0u;
                continue;
            }
            const Event e = ev[sidx[k]];
            perform(k, e, T.rx[e.rx].product_distance);
        }
    } else {
        detailed_balance_events(cur, R, T, D, cap, ne, status, sidx, prefix, ev, hbase, geo, rrec,
                                reinterpret_cast<double *>(skey), perform);
    }
    __syncthreads();
    if (tid < (int)T.n_rx && tid < 256) {
        rx_step[(size_t)r * T.n_rx + tid] = rxcount[tid];
        rx_total[(size_t)r * T.n_rx + tid] += rxcount[tid];
    }
    if (tid == 0) {
        R.handle_next = hbase + s_total;
        ev_count[r] = 0;
        if (recs) rec_count[r] = ne;
    }
}

__global__ __launch_bounds__(RESOLVE_THREADS) void k_resolve(Grid cur, ReplicaDev *reps, Tables T, uint32_t cap,
                                                             Event *events, uint32_t *ev_count,
                                                             unsigned long long *rx_step, unsigned long long *rx_total,
                                                             int require_active, uint32_t max_wl, Edit *edits,
                                                             uint32_t *edit_count, uint32_t edit_cap, RxRecord *recs,
                                                             uint32_t *rec_count, uint32_t *err) {
    resolve_replica(cur, reps, T, cap, events, ev_count, rx_step, rx_total, require_active, max_wl, edits, edit_count,
                    edit_cap, recs, rec_count, err);
}

// SCPUDetailedBalance::perform's event loop, one workgroup per replica (the candidates come from the
// search kernels with the Gillespie keys)
__global__ __launch_bounds__(RESOLVE_THREADS) void k_detailed_balance(Grid cur, ReplicaDev *reps, Tables T,
                                                                      uint32_t cap, Event *events, uint32_t *ev_count,
                                                                      unsigned long long *rx_step,
                                                                      unsigned long long *rx_total, uint32_t max_wl,
                                                                      RxRecord *recs, uint32_t *rec_count,
                                                                      uint32_t *err, DbTables D) {
    resolve_replica<true>(cur, reps, T, cap, events, ev_count, rx_step, rx_total, 0, max_wl, (Edit *)nullptr,
                          (uint32_t *)nullptr, 0u, recs, rec_count, err, D);
}

// marks every particle of replica r dead in the current buffer
__global__ void k_kill_replica(Grid cur, uint32_t r, uint32_t max_wl) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = *cur.wl_count;
    n = n < max_wl ? n : max_wl;
    if (i >= n) return;
    uint32_t s = cur.wl[i];
    State st = cur.bins[s];
    if (tw_replica(st.tw) == r) cur.bins[s].tw = st.tw | TW_DEAD;
}

// frees the cells of a removed replica in one buffer: counters to zero, overflow records detached
__global__ void k_release_cells(Grid g, uint32_t base, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) g.counts[base + i] = 0u;
    uint32_t no = *g.ovf_count;
    no = no < g.ovf_cap ? no : g.ovf_cap;
    for (uint32_t o = i; o < no; o += gridDim.x * blockDim.x) {
        uint32_t c = g.ovf_cell[o];
        if (c >= base && c < base + n) g.ovf_cell[o] = 0xFFFFFFFFu;
    }
}

// -------------------------------------------------------------------------------------------------
// upload, gather, observables, test exports
// -------------------------------------------------------------------------------------------------
// handles: the topology particles (the first n_topo) 0 .. n_topo - 1, the free ones from n_topo + spare on
__global__ void k_upload(Grid g, ReplicaDev *reps, uint32_t r, uint32_t n, uint32_t n_topo, uint32_t spare,
                         const double *pos, const uint32_t *types, const uint8_t *topo_flavor, Topo topo, uint32_t cap,
                         uint32_t max_wl, uint32_t *err) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ReplicaDev &R = reps[r];
    const RepGeom geo = load_geom(R);
    State st;
    st.x = pos[3ull * i];
    st.y = pos[3ull * i + 1];
    st.z = pos[3ull * i + 2];
    const uint32_t t = types[i];
    const bool is_topo = topo_flavor && topo_flavor[t];
    st.tw = make_tw(t, is_topo ? TW_TOPO : 0u, r);
    st.h = i < n_topo ? i : i + spare;
    uint32_t slot = bin_insert(g, geo, cap, st, err, false, max_wl);
    if (is_topo && topo.loc && slot != 0xFFFFFFFFu) topo.loc[R.topo_base + i] = slot;
}

// mode 0: particles (pos, type, handle [, force]); only replica `r`
__global__ void k_gather(Grid cur, uint32_t r, uint32_t *out_count, uint32_t out_cap, double *pos, uint32_t *types,
                         uint32_t *handles, const double *force_in, double *forces, uint32_t max_wl) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = *cur.wl_count;
    n = n < max_wl ? n : max_wl;
    if (i >= n) return;
    const uint32_t s = cur.wl[i];
    const State st = cur.bins[s];
    if ((st.tw & TW_DEAD) || tw_replica(st.tw) != r) return;
    uint32_t p = atomicAdd(out_count, 1u);
    if (p >= out_cap) return;
    if (pos) {
        pos[3ull * p] = st.x;
        pos[3ull * p + 1] = st.y;
        pos[3ull * p + 2] = st.z;
    }
    if (types) types[p] = tw_type(st.tw);
    if (handles) handles[p] = st.h;
    if (forces) {
        for (int d = 0; d < 3; ++d) forces[3ull * p + d] = force_in ? force_in[3ull * s + d] : 0.0;
    }
}

__global__ void k_count_types(Grid cur, uint32_t n_types, unsigned long long *counts, uint32_t n_reps,
                              uint32_t max_wl) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = *cur.wl_count;
    n = n < max_wl ? n : max_wl;
    if (i >= n) return;
    const State st = cur.bins[cur.wl[i]];
    if (st.tw & TW_DEAD) return;
    uint32_t r = tw_replica(st.tw);
    if (r >= n_reps) return;
    atomicAdd(&counts[(size_t)r * n_types + tw_type(st.tw)], 1ull);
}

__global__ void k_pairs(Grid cur, ReplicaDev *reps, uint32_t cap, int radius, uint32_t r, double cut2,
                        uint32_t *out, uint32_t out_cap, uint32_t *out_count, uint32_t max_wl) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = *cur.wl_count;
    n = n < max_wl ? n : max_wl;
    if (i >= n) return;
    const uint32_t s = cur.wl[i];
    const State me = cur.bins[s];
    if ((me.tw & (TW_DEAD | TW_OUTSIDE)) || tw_replica(me.tw) != r) return;
    const RepGeom geo = load_geom(reps[r]);
    const uint32_t cl = cell_of_slot(cur, s, cap) - geo.cell_base;
    for_each_stencil_slot(geo, cur, cap, radius, cl, [&](uint32_t s2) {
        if (s2 == s) return;
        const State o = cur.bins[s2];
        if (o.tw & (TW_DEAD | TW_OUTSIDE)) return;
        const double dx = min_image(o.x - me.x, geo.box[0], geo.pbc[0]);
        const double dy = min_image(o.y - me.y, geo.box[1], geo.pbc[1]);
        const double dz = min_image(o.z - me.z, geo.box[2], geo.pbc[2]);
        if (dist2(dx, dy, dz) < cut2) {
            uint32_t p = atomicAdd(out_count, 1u);
            if (p < out_cap) {
                out[2 * p] = me.h;
                out[2 * p + 1] = o.h;
            }
        }
    });
}

}  // namespace rdy
