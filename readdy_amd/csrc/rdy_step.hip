// The fused simulation step: one stencil walk per step.
//
// SimulationLoop.h:280-300 runs integrate -> NL -> reactions -> NL -> forces, and the forces of the
// post-reaction state drive the next integration.  Here the state between two walks is x'(t), the
// positions after step t's integration and before its reactions, and a step is
//   k_step_begin    per-replica bookkeeping; clears the buffer the walk writes
//   (bonded terms)  on x'(t); topology particles are never moved by reactions
//   k_brick_step    one LDS-staged stencil walk over x'(t): the pair forces G (fixed point), the
//                   reaction candidates of step t, and, optimistically, step t+1's Euler-Maruyama
//                   move with G, re-binned into the other buffer
//   k_resolve       step t's reactions on x'(t) (in place), listing every record it changes
//   k_fix_sites     per changed record: lists the moved copies to withdraw (of killed and retyped
//                   records, and of every surviving neighbour whose pair force with it changed) and
//                   the records to move again (those neighbours, inserted and retyped records)
//   k_fix_withdraw  withdraws the listed copies from the other buffer (found by handle)
//   k_fix_advance   forms each listed record's pair force afresh on the post-reaction state and
//                   moves it again
// A reaction changes the pair set only around the records it changes; every other record's G is
// already the force of the post-reaction state, bit for bit (an integer sum of per-pair integers,
// pair_fixed, over the same pairs), so its optimistic move stands, as the reference's order needs.
// The first walk of a run (k_brick_step with prologue set) forms the forces of the current state
// and makes step 1's move; the last step of a run (and of a replica's budget) copies instead of
// moving, so that a run ends on the post-reaction state of its last step.
#pragma once

#include "rdy_bricks.hip"

namespace rdy {

// Moves a record one Euler-Maruyama step (CPUEulerBDIntegrator.cpp:63-73) with the pair force
// (fx, fy, fz, 2^-32 fixed point) plus its bonded force, Philox counter `step`, or copies it
// (integ false), and re-bins it into nxt.  Returns the new slot (UINT32_MAX on a full pool).
__device__ __forceinline__ uint32_t advance_particle(const Grid &nxt, const ReplicaDev &R, const RepGeom &geo,
                                                     const Tables &T, const Topo &topo, uint32_t cap, double dt,
                                                     uint32_t step, bool integ, State me, long long fx, long long fy,
                                                     long long fz, uint32_t *err) {
    State nw = me;
    nw.tw &= ~TW_NEW;
    if (integ) {
        if ((me.tw & TW_TOPO) && topo.fb) {
            const long long *fb = topo.fb + 3ull * (R.topo_base + me.h);
            fx += fb[0];
            fy += fb[1];
            fz += fb[2];
        }
        if (fixed_saturated(fx) || fixed_saturated(fy) || fixed_saturated(fz)) atomicAdd(&err[ERR_FORCE_RANGE], 1u);
        const double Fv[3] = {(double)fx * FORCE_INV, (double)fy * FORCE_INV, (double)fz * FORCE_INV};
        double xi[3];
        bd_noise(nullptr, me.h, step, R.seed, xi);
        double p[3] = {me.x, me.y, me.z};
        em_move(T, tw_type(me.tw), dt, Fv, xi, geo.box, geo.pbc, p, err);
        nw.x = p[0];
        nw.y = p[1];
        nw.z = p[2];
    }
    const uint32_t slot = bin_insert(nxt, geo, cap, nw, err, false, 0);
    if (slot != 0xFFFFFFFFu && (me.tw & TW_TOPO) && topo.loc) topo.loc[R.topo_base + me.h] = slot;
    return slot;
}

// Step t's bookkeeping before its walk: replicas with budget left take the step (Philox step t) and
// integrate step t+1 in the same walk unless this is their last step or the launch's (last);
// clears buffer `clear` (the walk's target; thread i its cell counters 4i .. 4i + 3), the change lists
// and the bonded forces.
// With topology reactions on, it also keeps each topology particle's slot in the buffer the walk reads
// (loc_cur): the walk's moves point loc at the other buffer, and the topology reactions run after it.
__global__ void k_step_begin(ReplicaDev *reps, uint32_t n_reps, int last, Grid clear, uint32_t n_cells,
                             long long *fb, uint32_t n_fb, uint32_t *counters, unsigned long long *e_bonded,
                             const uint32_t *loc, uint32_t *loc_cur, uint32_t n_loc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_loc) loc_cur[i] = loc[i];
    if (i < n_reps) {
        ReplicaDev &R = reps[i];
        R.energy = 0ull;
        e_bonded[i] = 0ull;
        if (R.in_use && R.steps_left > 0) {
            R.active = 1u;
            R.step += 1u;
            R.steps_left -= 1u;
            R.steps_done += 1u;
            R.integ = (R.steps_left > 0 && !last) ? 1u : 0u;
        } else {
            R.active = 0u;
            R.integ = 0u;
        }
    }
    const uint32_t c4 = 4u * i;  // the counters array starts 256-byte aligned (hipMalloc)
    if (c4 + 3u < n_cells) reinterpret_cast<uint4 *>(clear.counts)[i] = make_uint4(0u, 0u, 0u, 0u);
    else
        for (uint32_t c = c4; c < n_cells; ++c) clear.counts[c] = 0u;
    if (i < n_fb) fb[i] = 0;
    if (i == 0) {
        *clear.wl_count = 0u;
        *clear.ovf_count = 0u;
        counters[0] = 0u;  // changed records
        counters[1] = 0u;  // records to move again
        counters[2] = 0u;  // copies to withdraw
    }
}

// clears a buffer and the bonded forces after a run (between runs only the current buffer holds
// records, and the bonded forces are formed where they are consumed)
__global__ void k_clear_grid(Grid g, uint32_t n_cells, long long *fb, uint32_t n_fb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_cells) g.counts[i] = 0u;
    if (i < n_fb) fb[i] = 0;
    if (i == 0) {
        *g.wl_count = 0u;
        *g.ovf_count = 0u;
    }
}

// The type tables of the walk, read into registers before the brick and its replica are known (one
// entry per thread: n_types <= MAX_STAGED_TYPES gives n_types^2 <= BRICK_THREADS), so that their loads
// overlap the brick's, then placed in LDS with the thresholds of the brick's frame: the force tables
// (ft) and the order-2 reaction index (st.o2); st.thr / st.reach hold the larger of the force and
// reaction thresholds / reaches.  Every entry is formed from the placing thread's own words (the
// reaches from the per-type row maxima), so the placement needs no barrier of its own: the staging
// that follows has barriers before anything reads the tables.
struct TablePrefetch {  // raw words (registers, not a stack object)
    uint4 pi;
    uint2 ri;
    uint4 pot0, pot1, pot2;
    float2 rm;
};
static_assert(sizeof(PairIdx) == 16 && sizeof(RxPairIdx) == 8 && sizeof(PotDev) == 48, "table entry sizes");

__device__ __forceinline__ TablePrefetch prefetch_tables(const Tables &T) {
    TablePrefetch f{};
    const uint32_t k = threadIdx.x, n2 = T.n_types * T.n_types;
    if (k < n2) {
        f.pi = reinterpret_cast<const uint4 *>(T.pidx)[k];
        f.ri = reinterpret_cast<const uint2 *>(T.o2idx)[k];
    }
    if (k < T.n_pots) {
        const uint4 *q = reinterpret_cast<const uint4 *>(T.pots) + 3ull * k;
        f.pot0 = q[0];
        f.pot1 = q[1];
        f.pot2 = q[2];
    }
    if (k < T.n_types) f.rm = T.rowmax[k];
    return f;
}

__device__ __forceinline__ void place_tables(char *base, const Tables &T, const TablePrefetch &f, float margin,
                                             ForceTables &ft, SearchTables &st) {
    const uint32_t nt = T.n_types, n2 = nt * nt, k = threadIdx.x;
    char *p = base;
    ft.pidx = (PairIdx *)p;
    p += align16(sizeof(PairIdx) * n2);
    ft.pots = (PotDev *)p;
    p += align16(sizeof(PotDev) * T.n_pots);
    ft.thr = (float *)p;
    p += align16(4 * n2);
    ft.reach = (float *)p;
    p = base + force_table_bytes(nt, T.n_pots);
    st.o2 = (RxPairIdx *)p;
    p += align16(sizeof(RxPairIdx) * n2);
    st.thr = (float *)p;
    p += align16(4 * n2);
    st.reach = (float *)p;
    if (k < n2) {
        reinterpret_cast<uint4 *>(ft.pidx)[k] = f.pi;
        reinterpret_cast<uint2 *>(st.o2)[k] = f.ri;
        const PairIdx pi = ft.pidx[k];
        const RxPairIdx ri = st.o2[k];
        const float tf = prefilter_thr(pi.maxcut2_f, pi.count != 0, margin);
        ft.thr[k] = tf;
        st.thr[k] = fmaxf(tf, prefilter_thr(ri.maxr2_f, ri.count != 0, margin));
    }
    if (k < T.n_pots) {
        uint4 *q = reinterpret_cast<uint4 *>(ft.pots) + 3ull * k;
        q[0] = f.pot0;
        q[1] = f.pot1;
        q[2] = f.pot2;
    }
    for (uint32_t q = k + BRICK_THREADS; q < T.n_pots; q += BRICK_THREADS) ft.pots[q] = T.pots[q];
    if (k < nt) {
        const float rf = reach_of(f.rm.x, f.rm.x >= 0.f, margin);
        ft.reach[k] = rf;
        st.reach[k] = fmaxf(rf, reach_of(f.rm.y, f.rm.y >= 0.f, margin));
    }
}

// Orders the brick's own records so that those of the longest-reach types come first (a stable
// partition of m.inner): a wave then holds particles of like stencil reach, and the short-reach waves
// finish early instead of idling behind their long-reach lanes.  tmp: n_inner halves of free LDS.
// block_exclusive_scan keeps one sum per wave of a SCAN_THREADS block.
static_assert(BRICK_THREADS == SCAN_THREADS, "order_inner_by_reach scans a brick block with block_exclusive_scan");
__device__ void order_inner_by_reach(const BrickSmem &m, uint32_t n_inner, const float *reach_tab, uint32_t nt,
                                     uint16_t *tmp) {
    __shared__ float s_rmax;
    if (threadIdx.x == 0) {
        float r = -1.f;
        for (uint32_t t = 0; t < nt; ++t) r = fmaxf(r, reach_tab[t]);
        s_rmax = r;
    }
    __syncthreads();
    const float rmax = s_rmax;
    auto longest = [&](uint32_t j) -> bool {
        const uint32_t tw = __float_as_uint(m.P[j].w);
        return !(tw & TW_DEAD) && reach_tab[tw_type(tw)] >= rmax;
    };
    uint32_t n0 = 0;
    for (uint32_t c0 = 0; c0 < n_inner; c0 += BRICK_THREADS) {  // count the long-reach records
        const uint32_t ii = c0 + threadIdx.x;
        uint32_t tot;
        block_exclusive_scan(ii < n_inner && longest(m.inner[ii]) ? 1u : 0u, &tot);
        n0 += tot;
    }
    uint32_t o0 = 0, o1 = n0;
    for (uint32_t c0 = 0; c0 < n_inner; c0 += BRICK_THREADS) {
        const uint32_t ii = c0 + threadIdx.x;
        uint32_t j = 0;
        bool f = false;
        if (ii < n_inner) {
            j = m.inner[ii];
            f = longest(j);
        }
        uint32_t t0;
        const uint32_t r0 = block_exclusive_scan(f ? 1u : 0u, &t0);
        if (ii < n_inner) tmp[f ? o0 + r0 : o1 + (ii - c0 - r0)] = (uint16_t)j;
        o0 += t0;
        o1 += min(n_inner - c0, (uint32_t)BRICK_THREADS) - t0;
    }
    __syncthreads();
    for (uint32_t ii = threadIdx.x; ii < n_inner; ii += BRICK_THREADS) m.inner[ii] = tmp[ii];
    __syncthreads();
}

// The walk of one brick over cur = x'(t): the pair forces of the brick's own records (2^-32 fixed
// point, summed in LDS per wave), step t's reaction candidates (order 1 and 2, the draws of
// k_brick_search), and each particle moved to step t+1 (or copied) into nxt as soon as its force is
// complete.  prologue: the first walk of a run (no candidates; moves replicas with steps left, Philox
// step R.step + 1).  Replicas neither searching nor moving are copied without a walk.  Four waves per
// SIMD are asked of the register allocator (the move inside the walk loop would otherwise take the
// kernel to three).
template <bool AH>
__global__ __launch_bounds__(BRICK_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_brick_step(
    Grid cur, Grid nxt, int prologue, ReplicaDev *reps, Tables T, Topo topo, const BrickDev *bricks, uint32_t brick0,
    uint32_t cap, int radius, int handler, double dt, Event *events, uint32_t *ev_count, uint32_t max_states,
    uint32_t max_halo, uint32_t smem_bytes, uint32_t *slow, uint32_t *err) {
    extern __shared__ __align__(16) char smem[];
    __shared__ ForceWave wave_lds[BRICK_WAVES];
    const bool staged_types = T.n_types <= MAX_STAGED_TYPES;
    TablePrefetch tpf;
    if (staged_types) tpf = prefetch_tables(T);
    const uint32_t bidx = brick0 + xcd_block(blockIdx.x, gridDim.x);
    const BrickDev B = bricks[bidx];
    // the halo counters load from the brick record alone, alongside the replica's record
    const Halo Hs = make_halo(B);
    const HaloCounts hcn = load_halo_counts(cur, B, Hs);
    const ReplicaDev &R = reps[B.r];
    if (!R.in_use) return;
    const RepGeom geo = load_geom(R);
    const bool search = !prologue && R.active;                          // block-uniform
    const bool integ = prologue ? (R.steps_left > 0) : (R.integ != 0u);  // block-uniform
    const uint32_t step_next = R.step + 1u;
    if (!search && !integ) {  // copy only: no forces needed
        for_each_own_record(cur, B, geo, cap, [&](uint32_t s, const State &me) {
            if (!(me.tw & TW_DEAD)) advance_particle(nxt, R, geo, T, topo, cap, dt, step_next, false, me, 0, 0, 0, err);
        });
        return;
    }
    // a launch without reactions (handler > 1) covers every replica; else a replica's own choice wins
    if (handler <= 1 && R.handler >= 0) handler = R.handler;
    const bool rx_on = search && handler <= 1 && T.n_rx > 0;  // block-uniform
    const bool o2_on = rx_on && T.has_o2;
    const uint32_t nt = T.n_types;
    const Frame F = make_frame(B, geo);
    uint32_t total = 0xFFFFFFFFu, n_inner = 0;
    BrickSmem m;
    ForceTables ft;
    SearchTables st;
    if (staged_types) {
        if (table_bytes(nt, T.n_pots) + brick_smem_bytes(max_states, max_halo) > smem_bytes) {  // block-uniform
            if (threadIdx.x == 0) atomicAdd(&err[ERR_LDS_LAYOUT], 1u);
            return;
        }
        // with reactions on, the walk reaches the larger of the force and reaction cutoffs
        place_tables(smem, T, tpf, F.margin, ft, st);
        m = carve(smem + table_bytes(nt, T.n_pots), max_states, max_halo);
        total = stage_brick(cur, B, geo, F, cap, m, max_states, Hs, hcn, &n_inner);
    }
    if (total == 0xFFFFFFFFu) {
        list_slow_brick(slow, bidx);
        return;
    }
    const WalkCtx wc = make_walk(B, geo, Hs, radius);
    if (AH != (wc.whole == 0)) {  // launch and brick disagree (host invariant): refuse the brick
        if (threadIdx.x == 0) atomicAdd(&err[ERR_LDS_LAYOUT], 1u);
        return;
    }
    const float *thr_tab = o2_on ? st.thr : ft.thr;
    const float *reach_tab = o2_on ? st.reach : ft.reach;
    if (n_inner > BRICK_THREADS && 2 * n_inner <= transient_bytes(max_halo))  // block-uniform
        order_inner_by_reach(m, n_inner, reach_tab, nt, (uint16_t *)m.rings);
    const uint32_t r = B.r, step = R.step;
    const uint64_t seed = R.seed;
    auto emit = [&](uint32_t rx, uint32_t s1, uint32_t s2, uint32_t h1, uint32_t h2, unsigned long long key) {
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
    ForceWave &W = wave_lds[threadIdx.x / 64];
    const uint32_t lane = __lane_id();
    // a surviving pair (entry = owner lane << 16 | staged neighbour): the exact FP64 gates, the
    // pair force on the owner, and (each unordered pair once) the order-2 draws
    State me{};  // this lane's own record in the current round (the walk's owner records)
    auto eval = [&](uint32_t ent) {
        // the owner's record from the owner lane's registers (every lane takes part in the shuffles)
        const uint32_t ol = ent == NO_ENTRY ? 0u : ent >> 16, q = ent & 0xFFFFu;
        State me_o;
        me_o.x = __shfl(me.x, (int)ol);
        me_o.y = __shfl(me.y, (int)ol);
        me_o.z = __shfl(me.z, (int)ol);
        me_o.tw = __shfl(me.tw, (int)ol);
        me_o.h = __shfl(me.h, (int)ol);
        if (ent == NO_ENTRY) return;
        const State &me = me_o;
        const uint32_t s1 = W.s[ol], s2 = m.slot[q];
        const State o = cur.bins[s2];
        const uint32_t t1 = tw_type(me.tw), t2 = tw_type(o.tw);
        // CPUCalculateForces.cpp:226-259: x_ij = shortestDifference(x_i, x_j), gate d2 < rc^2.
        // The reaction gate measures distSquared(x_j, x_i) (CPUUncontrolledApproximation.cpp:108):
        // the same square, since a - b = -(b - a) and the minimum image is symmetric in magnitude.
        const double ddx = min_image(o.x - me.x, geo.box[0], geo.pbc[0]);
        const double ddy = min_image(o.y - me.y, geo.box[1], geo.pbc[1]);
        const double ddz = min_image(o.z - me.z, geo.box[2], geo.pbc[2]);
        const double d2 = dist2(ddx, ddy, ddz);
        long long px, py, pz, pe;
        if (pair_fixed<false>(ft.pots, ft.pidx[t1 * nt + t2], ddx, ddy, ddz, d2, px, py, pz, pe)) {
            if (fixed_saturated(px) || fixed_saturated(py) || fixed_saturated(pz)) atomicAdd(&err[ERR_FORCE_RANGE], 1u);
            atomicAdd(&W.acc[0][ol], (unsigned long long)px);
            atomicAdd(&W.acc[1][ol], (unsigned long long)py);
            atomicAdd(&W.acc[2][ol], (unsigned long long)pz);
        }
        // each unordered pair once: the lower type finds it; equal types: the lower bin slot
        if (o2_on && ((t2 > t1) || (t2 == t1 && s2 > s1))) {
            const RxPairIdx ri = st.o2[t1 * nt + t2];
            const uint32_t mh = me.h;
            const uint32_t lo = mh < o.h ? mh : o.h, hi = mh < o.h ? o.h : mh;
            for (uint32_t p = 0; p < ri.count; ++p) {
                const uint32_t rx = T.o2list[ri.start + p];
                const RxDev X = T.rx[rx];
                if (!(X.rate > 0 && d2 < X.educt_dist2)) continue;
                uint32_t x4[4];
                philox(lo, step, 2u | (rx << 8), hi, seed, x4);
                if (!(u53(x4[0], x4[1]) < X.p_accept)) continue;
                // entry1: the educt of type educts[0]; equal educt types: the lower handle
                const bool me_first = (X.educts[0] != X.educts[1]) ? (t1 == X.educts[0]) : (mh < o.h);
                if (me_first) emit(rx, s1, s2, mh, o.h, event_key(handler, x4, X.rate));
                else emit(rx, s2, s1, o.h, mh, event_key(handler, x4, X.rate));
            }
        }
    };
    for (uint32_t base = threadIdx.x & ~63u; base < n_inner; base += BRICK_THREADS) {  // wave-uniform
        const uint32_t ii = base + lane;
        bool valid = ii < n_inner;
        uint32_t j = 0;
        float4 pj = make_float4(0.f, 0.f, 0.f, 0.f);
        if (valid) {
            j = m.inner[ii];
            pj = m.P[j];
            valid = !(__float_as_uint(pj.w) & TW_DEAD);
        }
        me = State{};
        uint32_t sj = 0;
        if (valid) {
            sj = m.slot[j];
            me = cur.bins[sj];
        }
        W.s[lane] = sj;
        W.acc[0][lane] = 0ull;
        W.acc[1][lane] = 0ull;
        W.acc[2][lane] = 0ull;
        wave_lds_sync();
        const uint32_t t = tw_type(me.tw);
        if (rx_on && valid) {  // order 1 (CPUUncontrolledApproximation.cpp:81-95)
            for (uint32_t k = T.o1idx[t]; k < T.o1idx[t + 1]; ++k) {
                const uint32_t rx = T.o1list[k];
                const RxDev X = T.rx[rx];
                if (!(X.rate > 0)) continue;
                uint32_t x4[4];
                philox(me.h, step, 1u | (rx << 8), 0u, seed, x4);
                if (!(u53(x4[0], x4[1]) < X.p_accept)) continue;
                emit(rx, sj, sj, me.h, me.h, event_key(handler, x4, X.rate));
            }
        }
        const float reach = reach_tab[t];
        const bool active = valid && !(me.tw & TW_OUTSIDE) && reach >= 0.f;
        const float *thr_row = thr_tab + t * nt;
        // prefilter, branch-free: one LDS read of the candidate, one of the threshold
        auto keep = [&](uint32_t q, const float4 &pq) -> bool {
            const uint32_t tw2 = __float_as_uint(pq.w);
            float dx = pq.x - pj.x, dy = pq.y - pj.y, dz = pq.z - pj.z;
            if constexpr (!AH) frame_min_image(F, dx, dy, dz);
            const float d2p = prefilter_d2(dx, dy, dz);
            const float thr = thr_row[tw_type(tw2)];
            return (q != j) & (d2p < thr);  // dead / outside records hold NaN positions
        };
        pair_pipeline<AH>(wc, m.hstart, m.P, m.rings + (threadIdx.x / 64) * RING, active, m.hc_of[j], pj, reach, keep,
                          eval);
        wave_lds_sync();
        // the move with the finished pair force, from the owner's registers
        const long long fx = (long long)W.acc[0][lane], fy = (long long)W.acc[1][lane], fz = (long long)W.acc[2][lane];
        wave_lds_sync();
        if (valid) advance_particle(nxt, R, geo, T, topo, cap, dt, step_next, integ, me, fx, fy, fz, err);
    }
}

// the listed bricks of k_brick_step, neighbours read straight from the bins
__global__ __launch_bounds__(BRICK_THREADS) void k_brick_step_slow(Grid cur, Grid nxt, int prologue, ReplicaDev *reps,
                                                                   Tables T, Topo topo, const BrickDev *bricks,
                                                                   uint32_t cap, int radius, int handler, double dt,
                                                                   Event *events, uint32_t *ev_count, uint32_t *slow,
                                                                   uint32_t *err) {
    const uint32_t n = slow_count(slow);
    if (n == 0) return;  // grid-uniform
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {  // block-uniform
        const BrickDev B = bricks[slow[2 + i]];
        const ReplicaDev &R = reps[B.r];
        const RepGeom geo = load_geom(R);
        const bool search = !prologue && R.active;
        const bool integ = prologue ? (R.steps_left > 0) : (R.integ != 0u);
        const uint32_t step_next = R.step + 1u;
        for_each_own_record(cur, B, geo, cap, [&](uint32_t s, const State &me) {
            if (me.tw & TW_DEAD) return;
            long long fx = 0, fy = 0, fz = 0, e = 0;
            if (!(me.tw & TW_OUTSIDE)) global_forces(cur, geo, cap, radius, T, s, me, fx, fy, fz, e);
            if (search) search_one(cur, reps, T, cap, radius, handler, 0, events, ev_count, nullptr, 0, s, me, err);
            advance_particle(nxt, R, geo, T, topo, cap, dt, step_next, integ, me, fx, fy, fz, err);
        });
    }
    release_slow_list(slow);
}

// A fused step's reactions (resolve_replica, one workgroup per replica), preceded by the replica's
// bricks that k_brick_step listed for the bin-reading path (their candidates join the replica's list
// first).  The walk then needs no follow-up launch, which costs as much as a short kernel's work.
__global__ __launch_bounds__(RESOLVE_THREADS) void k_resolve_step(Grid cur, Grid nxt, ReplicaDev *reps, Tables T,
                                                                  Topo topo, const BrickDev *bricks, uint32_t cap,
                                                                  int radius, int handler, double dt,
                                                                  Event *events, uint32_t *ev_count, uint32_t *slow,
                                                                  unsigned long long *rx_step,
                                                                  unsigned long long *rx_total, uint32_t max_wl,
                                                                  Edit *edits, uint32_t *edit_count,
                                                                  uint32_t edit_cap, RxRecord *recs,
                                                                  uint32_t *rec_count, uint32_t *err) {
    const uint32_t n = slow_count(slow);  // the same for every block: the walk has ended
    for (uint32_t i = 0; i < n; ++i) {    // block-uniform
        const BrickDev B = bricks[slow[2 + i]];
        if (B.r != blockIdx.x) continue;
        const ReplicaDev &R = reps[B.r];
        const RepGeom geo = load_geom(R);
        const bool integ = R.integ != 0u;
        const uint32_t step_next = R.step + 1u;
        for_each_own_record(cur, B, geo, cap, [&](uint32_t s, const State &me) {
            if (me.tw & TW_DEAD) return;
            long long fx = 0, fy = 0, fz = 0, e = 0;
            if (!(me.tw & TW_OUTSIDE)) global_forces(cur, geo, cap, radius, T, s, me, fx, fy, fz, e);
            if (R.active) search_one(cur, reps, T, cap, radius, handler, 0, events, ev_count, nullptr, 0, s, me, err);
            advance_particle(nxt, R, geo, T, topo, cap, dt, step_next, integ, me, fx, fy, fz, err);
        });
    }
    if (n) {
        __syncthreads();
        __threadfence();  // the listed bricks' candidates before the replica's list is read
    }
    resolve_replica(cur, reps, T, cap, events, ev_count, rx_step, rx_total, 1, max_wl, edits, edit_count, edit_cap,
                    recs, rec_count, err);
    if (n) release_slow_list(slow);
}

// the cells of the (2r+1)^3 stencil of local cell cl, de-duplicated as axis_cells does, spread over
// the lanes of a wave: fn(cell) for each cell of this lane
template <class Fn>
__device__ __forceinline__ void wave_stencil_cells(const RepGeom &geo, uint32_t cl, int radius, Fn &&fn) {
    const int nz = geo.dims[2], ny = geo.dims[1];
    const int iz = (int)(cl % (uint32_t)nz), iy = (int)((cl / (uint32_t)nz) % (uint32_t)ny);
    const int ix = (int)(cl / ((uint32_t)nz * (uint32_t)ny));
    const int nax = axis_count(ix, geo.dims[0], geo.pbc[0], radius);
    const int nay = axis_count(iy, geo.dims[1], geo.pbc[1], radius);
    const int naz = axis_count(iz, geo.dims[2], geo.pbc[2], radius);
    const uint32_t ncell = (uint32_t)(nax * nay * naz);
    for (uint32_t c = __lane_id(); c < ncell; c += 64) {
        const int a = (int)(c / (uint32_t)(nay * naz)), rem = (int)(c % (uint32_t)(nay * naz));
        const int b = rem / naz, cc = rem % naz;
        fn(geo.cell_base +
           ((uint32_t)axis_elem(a, ix, geo.dims[0], geo.pbc[0], radius) * (uint32_t)ny +
            (uint32_t)axis_elem(b, iy, geo.dims[1], geo.pbc[1], radius)) *
               (uint32_t)nz +
           (uint32_t)axis_elem(cc, iz, geo.dims[2], geo.pbc[2], radius));
    }
}

// fn(slot) for every record of one cell of buffer g (overflow records included)
template <class Fn>
__device__ __forceinline__ void cell_records(const Grid &g, uint32_t cap, uint32_t cell, Fn &&fn) {
    const uint32_t cnt = g.counts[cell];
    const uint32_t mm = cnt < cap ? cnt : cap;
    for (uint32_t q = 0; q < mm; ++q) fn(cell * cap + q);
    if (cnt > cap) {  // rare: a crowded cell spilled into the overflow region
        uint32_t no = *g.ovf_count;
        no = no < g.ovf_cap ? no : g.ovf_cap;
        for (uint32_t o = 0; o < no; ++o)
            if (g.ovf_cell[o] == cell) fn(g.nbin + o);
    }
}

// Per changed record of step t's reactions, one wave each, its lanes spread over the record's
// stencil cells; cur is x'(t) with the reactions applied, nxt holds the walk's optimistic moves.
//   killed k:   its moved copy is withdrawn;
//   inserted p: p is listed for a move;
//   retyped r:  its moved copy (under its old handle) is withdrawn and r is listed;
// and, when the replica moves this step, every surviving old neighbour n (a live record without
// TW_NEW) whose pair force from the changed record differs between before and after is listed
// once (bit in aff_bits) for a fresh force and a new move, its moved copy withdrawn.
__global__ __launch_bounds__(256) void k_fix_sites(Grid cur, const ReplicaDev *reps, Tables T, uint32_t cap,
                                                   int radius, const Edit *edits, const uint32_t *edit_count,
                                                   uint32_t edit_cap, uint32_t *aff_bits, uint32_t *aff_list,
                                                   uint32_t *aff_count, uint32_t aff_cap, uint2 *wd_list,
                                                   uint32_t *wd_count, uint32_t wd_cap, uint32_t *err) {
    const uint32_t lane = __lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
    uint32_t n_edit = *edit_count;
    n_edit = n_edit < edit_cap ? n_edit : edit_cap;
    const uint32_t nt = T.n_types;
    auto list = [&](uint32_t s) {
        const uint32_t p = atomicAdd(aff_count, 1u);
        if (p < aff_cap) aff_list[p] = s;
        else atomicAdd(&err[ERR_EDIT_OVERFLOW], 1u);
    };
    auto withdraw = [&](uint32_t s, uint32_t h) {
        const uint32_t p = atomicAdd(wd_count, 1u);
        if (p < wd_cap) wd_list[p] = make_uint2(s, h);
        else atomicAdd(&err[ERR_EDIT_OVERFLOW], 1u);
    };
    for (uint32_t k = wave; k < n_edit; k += n_waves) {  // wave-uniform
        const Edit ed = edits[k];
        const uint32_t kind = ed.info & 3u, t_old = ed.info >> 2;
        const uint32_t slot = ed.slot;
        const ReplicaDev &R = reps[ed.r];
        const State st = cur.bins[slot];
        if (lane == 0) {
            if (kind == EDIT_KILLED) withdraw(slot, st.h);
            if (kind == EDIT_RETYPED) withdraw(slot, ed.pad);  // the copy carries the old handle
            if (kind != EDIT_KILLED) list(slot);
        }
        if (!R.integ || (st.tw & TW_OUTSIDE)) continue;  // copies only, or no neighbours
        const RepGeom geo = load_geom(R);
        const uint32_t ts = tw_type(st.tw);
        wave_stencil_cells(geo, cell_of_slot(cur, slot, cap) - geo.cell_base, radius, [&](uint32_t cell) {
            cell_records(cur, cap, cell, [&](uint32_t s2) {
                if (s2 == slot) return;
                const State n = cur.bins[s2];
                if (n.tw & (TW_DEAD | TW_OUTSIDE | TW_NEW)) return;
                const uint32_t tn = tw_type(n.tw);
                // the pair seen from n: x = site - n
                const double dx = min_image(st.x - n.x, geo.box[0], geo.pbc[0]);
                const double dy = min_image(st.y - n.y, geo.box[1], geo.pbc[1]);
                const double dz = min_image(st.z - n.z, geo.box[2], geo.pbc[2]);
                const double d2 = dist2(dx, dy, dz);
                long long ax, ay, az, bx = 0, by = 0, bz = 0, pe;
                pair_fixed<false>(T.pots, T.pidx[tn * nt + ts], dx, dy, dz, d2, ax, ay, az, pe);
                if (kind == EDIT_RETYPED)
                    pair_fixed<false>(T.pots, T.pidx[tn * nt + t_old], dx, dy, dz, d2, bx, by, bz, pe);
                if (((ax - bx) | (ay - by) | (az - bz)) == 0) return;  // n's sum is unchanged
                const uint32_t bit = 1u << (s2 & 31u);
                if (atomicOr(&aff_bits[s2 >> 5], bit) & bit) return;  // listed by another site
                list(s2);
                withdraw(s2, n.h);
            });
        });
    }
}

// Withdraws the listed moved copies from nxt: the live record of the listed handle, searched in the
// cells around the record's cell (one wave each; a move longer than one cell widens the search,
// up to the replica's whole grid)
__global__ __launch_bounds__(256) void k_fix_withdraw(Grid cur, Grid nxt, const ReplicaDev *reps, uint32_t cap,
                                                      const uint2 *wd_list, const uint32_t *wd_count,
                                                      uint32_t wd_cap) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
    uint32_t n = *wd_count;
    n = n < wd_cap ? n : wd_cap;
    for (uint32_t k = wave; k < n; k += n_waves) {  // wave-uniform
        const uint2 w = wd_list[k];
        const State st = cur.bins[w.x];
        const ReplicaDev &R = reps[tw_replica(st.tw)];
        const RepGeom geo = load_geom(R);
        const uint32_t cl = cell_of_slot(cur, w.x, cap) - geo.cell_base;
        bool found = false;
        auto probe = [&](uint32_t cell) {
            cell_records(nxt, cap, cell, [&](uint32_t s2) {
                const State o = nxt.bins[s2];
                if (o.h == w.y && !(o.tw & TW_DEAD)) {
                    nxt.bins[s2].tw = o.tw | TW_DEAD;
                    found = true;
                }
            });
        };
        wave_stencil_cells(geo, cl, 1, probe);
        if (__ballot(found)) continue;
        wave_stencil_cells(geo, cl, 2, probe);
        if (__ballot(found)) continue;
        for (uint32_t c = __lane_id(); c < R.n_cells; c += 64) probe(geo.cell_base + c);
    }
}

// Moves each listed record again (one wave each): its pair force formed afresh over the live
// records of its stencil in the post-reaction state cur (pair_fixed, the sum a walk forms), then
// the move (or the copy) into nxt; an old record's bit in aff_bits is cleared for the next step
__global__ __launch_bounds__(256) void k_fix_advance(Grid cur, Grid nxt, const ReplicaDev *reps, Tables T, Topo topo,
                                                     uint32_t cap, int radius, double dt, uint32_t *aff_bits,
                                                     const uint32_t *aff_list, const uint32_t *aff_count,
                                                     uint32_t aff_cap, uint32_t *err) {
    const uint32_t lane = __lane_id();
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
    uint32_t n = *aff_count;
    n = n < aff_cap ? n : aff_cap;
    const uint32_t nt = T.n_types;
    for (uint32_t k = wave; k < n; k += n_waves) {  // wave-uniform
        const uint32_t s = aff_list[k];
        const State me = cur.bins[s];
        const ReplicaDev &R = reps[tw_replica(me.tw)];
        const RepGeom geo = load_geom(R);
        const bool integ = R.integ != 0u;
        long long fx = 0, fy = 0, fz = 0;
        if (integ && !(me.tw & TW_OUTSIDE)) {
            const uint32_t t = tw_type(me.tw);
            wave_stencil_cells(geo, cell_of_slot(cur, s, cap) - geo.cell_base, radius, [&](uint32_t cell) {
                cell_records(cur, cap, cell, [&](uint32_t s2) {
                    if (s2 == s) return;
                    const State o = cur.bins[s2];
                    if (o.tw & (TW_DEAD | TW_OUTSIDE)) return;
                    // CPUCalculateForces.cpp:226-259: x = shortestDifference(x_i, x_j)
                    const double dx = min_image(o.x - me.x, geo.box[0], geo.pbc[0]);
                    const double dy = min_image(o.y - me.y, geo.box[1], geo.pbc[1]);
                    const double dz = min_image(o.z - me.z, geo.box[2], geo.pbc[2]);
                    long long px, py, pz, pe;
                    if (pair_fixed<false>(T.pots, T.pidx[t * nt + tw_type(o.tw)], dx, dy, dz, dist2(dx, dy, dz), px,
                                          py, pz, pe)) {
                        fx += px;
                        fy += py;
                        fz += pz;
                    }
                });
            });
            for (int off = 32; off > 0; off >>= 1) {
                fx += __shfl_xor(fx, off);
                fy += __shfl_xor(fy, off);
                fz += __shfl_xor(fz, off);
            }
        }
        if (lane == 0) {
            if (!(me.tw & TW_NEW)) atomicAnd(&aff_bits[s >> 5], ~(1u << (s & 31u)));
            advance_particle(nxt, R, geo, T, topo, cap, dt, R.step + 1u, integ, me, fx, fy, fz, err);
        }
    }
}

}  // namespace rdy
