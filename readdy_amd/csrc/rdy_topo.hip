// Spatial topology reactions (SpatialTopologyReaction.h:64-236; CPUEvaluateTopologyReactions.cpp:88-475):
// topology-particle enzymatic "T1(p1) + (p2) -> T2(p3) + (p4)" and fusion "T1(p1) + (p2) -> T2(p3--p4)",
// topology-topology enzymatic "T1(p1) + T2(p2) -> T3(p3) + T4(p4)" and fusion "T1(p1) + T2(p2) -> T3(p3--p4)".
//
//   k_topo_search   one thread per topology particle: every live partner within the reaction's radius in
//                   the particle's stencil (distSquared < radius^2) that is a particle of no topology, or a
//                   topology particle of another topology with a higher handle (each such pair once; a
//                   topology-topology reaction matches with its sides in either order), for each matching
//                   reaction draws the acceptance 1 - exp(-rate dt) (Philox counter (topology particle
//                   handle, step, 6 | reaction << 8, partner handle)) and appends the accepted candidates
//                   with their Efraimidis-Spirakis keys -ln(u) / rate to the replica's segment
//   k_topo_resolve  one workgroup per replica: the greedy pass over the accepted candidates in key order
//                   (performEvents, algorithm.h:62-136, with eventsDependent, :455-464: events of one
//                   topology depend on each other, and so do topology-particle events sharing the free
//                   particle), then the winners' changes: each particle and topology takes the product of
//                   its side (handleTopologyParticleReaction, :369-404; handleTopologyTopologyReaction,
//                   :406-475); a topology-particle fusion's free particle joins the topology under the next
//                   handle of the replica's topology reserve (allocated in key order) and is listed for the
//                   host (appendParticle); a topology-topology fusion moves every particle of the absorbed
//                   topology into the surviving one and is listed for the host (appendTopology)
//
// In the fused step both run after the step's reactions, on the post-reaction state, and list the
// records they retype for the fix-up (their moves, and those of their neighbours whose pair force
// changed, are redone).  Particles keep their handles, but for a topology-particle fusion's joining
// particle.  The bonded terms are not reconfigured here: the host sets them again when a reaction changes
// a topology's bonded configuration (every fusion does).
#pragma once

#include "rdy_step.hip"

namespace rdy {

struct SpatialDev {
    uint32_t type1, top_type1, type2, type_to1, top_type_to1, type_to2;
    double rate, p_accept, radius2;
    uint32_t fusion, tt;             // tt: topology-topology enzymatic (the partner in another topology)
    uint32_t top_type2, top_type_to2;
    uint32_t self_fusion, pad;       // topology-topology fusion within one topology (unbonded pairs)
};

struct TopoEvent {
    unsigned long long key;
    uint32_t s_top, s_free;  // bin slots of the topology particle and of the free particle (topology-topology:
                             // of the particle that found the pair and of its partner)
    uint32_t h_top, h_free;  // their handles
    uint32_t top;            // topology index within the replica
    uint32_t sr;             // reaction index; bit 31: topology-topology pair matched in reverse order
    uint32_t top2;           // topology-topology: the partner's topology (else UINT32_MAX)
    uint32_t pad;
};
static_assert(sizeof(TopoEvent) == 40, "TopoEvent layout");
constexpr uint32_t SR_FLIP = 0x80000000u;

struct TopoTables {
    const SpatialDev *sr;
    uint32_t n_sr;
    uint32_t *topo_of;           // [topology particle slot] topology index within its replica
    uint32_t *loc;               // [topology particle slot] the record's bin slot (written for joining particles)
    uint32_t *fus;               // [replica][TOPO_MAX_EVENTS][4] the last evaluation's fusions
    uint32_t *fus_count;         // [replica]
    uint32_t *mrg;               // [replica][TOPO_MAX_EVENTS][4] the last evaluation's topology merges
    uint32_t *mrg_count;         // [replica]
    const uint32_t *adj_off;     // [adj_n] bond adjacency of the topology particle slots (CSR offsets)
    const uint32_t *adj;         // the bonded partners' slots
    uint32_t adj_n;
    const uint32_t *topo_owner;  // [topology particle slot] replica, UINT32_MAX when unused
    uint32_t *ttype;             // topology types; replica r's at reps[r].tt_base
    uint32_t *treact;            // per topology (as ttype): the serial of the evaluation it last reacted in
    uint32_t serial;             // this evaluation's serial (0: not recorded, the fused step)
    TopoEvent *ev;               // [replica][ev_cap]
    uint32_t *ev_count;          // [replica]
    uint32_t ev_cap;
    unsigned long long *cnt_step, *cnt_total;  // [replica][n_sr]: last evaluation, since the replica was added
};

constexpr uint32_t TOPO_MAX_EVENTS = 1024;  // accepted candidates per replica per evaluation (LDS resident)
constexpr uint32_t TOPO_HASH = 4096;        // LDS hash-set entries for the used topologies and partners

// slots: each topology particle's slot in cur (loc, or its copy from the start of a fused step)
__global__ __launch_bounds__(256) void k_topo_search(Grid cur, const ReplicaDev *reps, const uint32_t *slots,
                                                     TopoTables TT, uint32_t n_slots, uint32_t cap, int radius,
                                                     int require_active, uint32_t *err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_slots) return;
    const uint32_t r = TT.topo_owner[i];
    if (r == 0xFFFFFFFFu) return;
    const ReplicaDev &R = reps[r];
    if (!R.in_use || (require_active && !R.active)) return;
    if (i < R.topo_base || i - R.topo_base >= R.n_topo) return;
    const uint32_t k = TT.topo_of[i];
    if (k >= R.n_tt) return;
    const uint32_t slot = slots[i];
    const State me = cur.bins[slot];
    if (me.tw & (TW_DEAD | TW_OUTSIDE)) return;
    const uint32_t t = tw_type(me.tw), T = TT.ttype[R.tt_base + k];
    bool any = false;
    for (uint32_t j = 0; j < TT.n_sr; ++j) {
        const SpatialDev &X = TT.sr[j];
        any = any || (X.type1 == t && X.top_type1 == T) || (X.tt && X.type2 == t && X.top_type2 == T);
    }
    if (!any) return;
    const RepGeom geo = load_geom(R);
    for_each_stencil_slot(geo, cur, cap, radius, cell_of_slot(cur, slot, cap) - geo.cell_base, [&](uint32_t s2) {
        if (s2 == slot) return;
        const State o = cur.bins[s2];
        if (o.tw & (TW_DEAD | TW_OUTSIDE)) return;
        // a partner in a topology: the topology-topology kind, each pair once (from the lower handle), the
        // two particles in different topologies (CPUEvaluateTopologyReactions.cpp:244-268)
        uint32_t k2 = 0xFFFFFFFFu, T2 = 0xFFFFFFFFu;
        if (o.tw & TW_TOPO) {
            if (!(me.h < o.h) || o.h >= R.n_topo) return;
            k2 = TT.topo_of[R.topo_base + o.h];
            if (k2 >= R.n_tt) return;
            T2 = TT.ttype[R.tt_base + k2];
        }
        // a pair within one topology reacts only by self-fusion and only when the two are not bonded (graph
        // distance above 1, CPUEvaluateTopologyReactions.cpp:303-318 with the default minimum distance)
        const bool same = k2 == k;
        int bonded = -1;
        auto is_bonded = [&]() -> bool {
            if (bonded < 0) {
                bonded = 0;
                const uint32_t g = R.topo_base + me.h, go = R.topo_base + o.h;
                if (TT.adj_off && g + 1 < TT.adj_n)
                    for (uint32_t a = TT.adj_off[g]; a < TT.adj_off[g + 1]; ++a) bonded |= TT.adj[a] == go ? 1 : 0;
            }
            return bonded != 0;
        };
        const uint32_t t2 = tw_type(o.tw);
        // distSquared(entry.pos, neighbor.pos) (CPUEvaluateTopologyReactions.cpp:278)
        const double dx = min_image(o.x - me.x, geo.box[0], geo.pbc[0]);
        const double dy = min_image(o.y - me.y, geo.box[1], geo.pbc[1]);
        const double dz = min_image(o.z - me.z, geo.box[2], geo.pbc[2]);
        const double d2 = dist2(dx, dy, dz);
        for (uint32_t j = 0; j < TT.n_sr; ++j) {
            const SpatialDev X = TT.sr[j];
            uint32_t flip = 0;
            if (k2 == 0xFFFFFFFFu) {
                if (X.tt || X.type1 != t || X.top_type1 != T || X.type2 != t2) continue;
            } else {  // the registry holds a topology-topology reaction under both orders of its sides
                if (!X.tt) continue;
                if (same && (!X.self_fusion || is_bonded())) continue;
                if (X.type1 == t && X.top_type1 == T && X.type2 == t2 && X.top_type2 == T2) flip = 0;
                else if (X.type2 == t && X.top_type2 == T && X.type1 == t2 && X.top_type1 == T2) flip = SR_FLIP;
                else continue;
            }
            if (!(d2 < X.radius2)) continue;
            uint32_t x4[4];
            philox(me.h, R.step, 6u | (j << 8), o.h, R.seed, x4);
            if (!(u53(x4[0], x4[1]) < X.p_accept)) continue;
            const uint32_t p = atomicAdd(&TT.ev_count[r], 1u);
            if (p >= TT.ev_cap) {
                atomicAdd(&err[ERR_EVENT_OVERFLOW], 1u);
                continue;
            }
            TopoEvent ev;
            ev.key = event_key(1, x4, X.rate);
            ev.s_top = slot;
            ev.s_free = s2;
            ev.h_top = me.h;
            ev.h_free = o.h;
            ev.top = k;
            ev.sr = j | flip;
            ev.top2 = k2;
            ev.pad = 0;
            TT.ev[(size_t)r * TT.ev_cap + p] = ev;
        }
    });
}

// inserts v into an LDS open-addressing set; false if it was there already
__device__ __forceinline__ bool hash_insert(uint32_t *set, uint32_t v) {
    uint32_t h = (v * 2654435761u) & (TOPO_HASH - 1);
    for (;;) {
        const uint32_t cur = set[h];
        if (cur == v) return false;
        if (cur == 0xFFFFFFFFu) {
            set[h] = v;
            return true;
        }
        h = (h + 1) & (TOPO_HASH - 1);
    }
}

__global__ __launch_bounds__(256) void k_topo_resolve(Grid cur, ReplicaDev *reps, TopoTables TT, int require_active,
                                                      Edit *edits, uint32_t *edit_count, uint32_t edit_cap,
                                                      uint32_t *err) {
    __shared__ unsigned long long skey[TOPO_MAX_EVENTS];
    __shared__ uint32_t order[TOPO_MAX_EVENTS];
    __shared__ uint8_t won[TOPO_MAX_EVENTS];
    __shared__ uint32_t used_top[TOPO_HASH], used_free[TOPO_HASH];
    __shared__ uint32_t fidx[TOPO_MAX_EVENTS];  // a won fusion's number: its joining particle's reserve handle
    __shared__ uint32_t s_nf, s_nm;              // fusions and topology merges performed
    const uint32_t r = blockIdx.x, tid = threadIdx.x;
    ReplicaDev &R = reps[r];
    if (!R.in_use || (require_active && !R.active)) {
        if (tid == 0) TT.ev_count[r] = 0u;
        return;
    }
    if (tid < TT.n_sr) TT.cnt_step[(size_t)r * TT.n_sr + tid] = 0ull;
    if (tid == 0 && TT.fus_count) TT.fus_count[r] = 0u;
    if (tid == 0 && TT.mrg_count) TT.mrg_count[r] = 0u;
    uint32_t ne = TT.ev_count[r];
    ne = ne < TT.ev_cap ? ne : TT.ev_cap;
    if (ne > TOPO_MAX_EVENTS) {
        if (tid == 0) atomicAdd(&err[ERR_EVENT_LDS], 1u);
        ne = TOPO_MAX_EVENTS;
    }
    __syncthreads();  // every thread has read the count before thread 0 clears it below
    if (ne == 0) {
        if (tid == 0) TT.ev_count[r] = 0u;
        return;
    }
    const TopoEvent *ev = TT.ev + (size_t)r * TT.ev_cap;
    for (uint32_t k = tid; k < ne; k += blockDim.x) {
        skey[k] = ev[k].key;
        won[k] = 0;
    }
    for (uint32_t k = tid; k < TOPO_HASH; k += blockDim.x) {
        used_top[k] = 0xFFFFFFFFu;
        used_free[k] = 0xFFFFFFFFu;
    }
    __syncthreads();
    // rank of each event under (key, topology particle handle, free particle handle, reaction): a strict
    // total order on distinct candidates, so the ranks are a permutation
    for (uint32_t k = tid; k < ne; k += blockDim.x) {
        const TopoEvent a = ev[k];
        uint32_t rank = 0;
        for (uint32_t j = 0; j < ne; ++j) {
            const unsigned long long kj = skey[j];
            if (kj < a.key) {
                ++rank;
            } else if (kj == a.key && j != k) {
                const TopoEvent b = ev[j];
                if (b.h_top != a.h_top ? b.h_top < a.h_top : (b.h_free != a.h_free ? b.h_free < a.h_free : b.sr < a.sr))
                    ++rank;
            }
        }
        order[rank] = k;
    }
    __syncthreads();
    const uint32_t n_topo0 = R.n_topo, reserve = R.topo_range > n_topo0 ? R.topo_range - n_topo0 : 0u;
    if (tid == 0) {  // the greedy pass in key order (few accepted events per step)
        uint32_t nf = 0, nm = 0;
        for (uint32_t q = 0; q < ne; ++q) {
            const TopoEvent a = ev[order[q]];
            bool top_free = true, partner_free = true;
            auto probe = [&](const uint32_t *set, uint32_t v) -> bool {  // v in the set?
                uint32_t h = (v * 2654435761u) & (TOPO_HASH - 1);
                for (;;) {
                    const uint32_t c = set[h];
                    if (c == v) return true;
                    if (c == 0xFFFFFFFFu) return false;
                    h = (h + 1) & (TOPO_HASH - 1);
                }
            };
            // an event is performed only when its topologies (and a topology-particle event's free particle)
            // are unused (eventsDependent, :455-464)
            const bool tt = a.top2 != 0xFFFFFFFFu;
            top_free = !probe(used_top, a.top) && !(tt && probe(used_top, a.top2));
            partner_free = tt || !probe(used_free, a.s_free);
            if (!top_free || !partner_free) continue;
            if (TT.sr[a.sr & ~SR_FLIP].fusion) {
                if (tt) {  // a topology merge: its row number, in key order
                    if (!TT.mrg) continue;
                    fidx[order[q]] = nm++;
                } else {  // the joining particle takes the next reserve handle, in key order
                    if (nf >= reserve || !TT.fus) {
                        atomicAdd(&err[ERR_TOPO_RESERVE], 1u);
                        continue;
                    }
                    fidx[order[q]] = nf++;
                }
            }
            hash_insert(used_top, a.top);
            if (tt) hash_insert(used_top, a.top2);
            else hash_insert(used_free, a.s_free);
            won[order[q]] = 1;
        }
        s_nf = nf;
        s_nm = nm;
    }
    __syncthreads();
    const bool fix = edits != nullptr;
    auto edit = [&](uint32_t slot, uint32_t old_type, uint32_t handle) {
        const uint32_t p = atomicAdd(edit_count, 1u);
        if (p < edit_cap) edits[p] = Edit{slot, EDIT_RETYPED | (old_type << 2), r, handle};
        else atomicAdd(&err[ERR_EDIT_OVERFLOW], 1u);
    };
    for (uint32_t k = tid; k < ne; k += blockDim.x) {
        if (!won[k]) continue;
        const TopoEvent a = ev[k];
        const uint32_t sr = a.sr & ~SR_FLIP;
        const SpatialDev X = TT.sr[sr];
        const uint32_t tw1 = cur.bins[a.s_top].tw, tw2 = cur.bins[a.s_free].tw;
        // a record the step's reactions changed already carries its change for the fix-up (TW_NEW), which
        // redoes its move on the final state; every other retyped record is listed once here
        const uint32_t n1 = fix && !(tw1 & TW_NEW) ? TW_NEW : 0u, n2 = fix && !(tw2 & TW_NEW) ? TW_NEW : 0u;
        if (X.tt) {  // handleTopologyTopologyReaction (:406-475): each side takes its products
            const bool flip = (a.sr & SR_FLIP) != 0u;
            cur.bins[a.s_top].tw = make_tw(flip ? X.type_to2 : X.type_to1, tw1 & (TW_TOPO | TW_OUTSIDE | TW_NEW), r) | n1;
            cur.bins[a.s_free].tw = make_tw(flip ? X.type_to1 : X.type_to2, tw2 & (TW_TOPO | TW_OUTSIDE | TW_NEW), r) | n2;
            if (X.fusion) {  // the partner's topology merges into this one, which takes top_type_to1 (:433-437)
                TT.ttype[R.tt_base + a.top] = X.top_type_to1;
                uint32_t *row = TT.mrg + 4ull * ((size_t)r * TOPO_MAX_EVENTS + fidx[k]);
                row[0] = a.top;
                row[1] = a.top2;
                row[2] = a.h_top;
                row[3] = a.h_free;
            } else {
                TT.ttype[R.tt_base + a.top] = flip ? X.top_type_to2 : X.top_type_to1;
                TT.ttype[R.tt_base + a.top2] = flip ? X.top_type_to1 : X.top_type_to2;
            }
            if (TT.serial) TT.treact[R.tt_base + a.top2] = TT.serial;
        } else {
            cur.bins[a.s_top].tw = make_tw(X.type_to1, tw1 & (TW_TOPO | TW_OUTSIDE | TW_NEW), r) | n1;
            if (X.fusion) {  // appendParticle: the free particle joins topology a.top under a reserve handle
                const uint32_t hn = n_topo0 + fidx[k], ti = R.topo_base + hn;
                cur.bins[a.s_free].tw = make_tw(X.type_to2, TW_TOPO | (tw2 & (TW_OUTSIDE | TW_NEW)), r) | n2;
                cur.bins[a.s_free].h = hn;
                TT.loc[ti] = a.s_free;
                TT.topo_of[ti] = a.top;
                uint32_t *row = TT.fus + 4ull * ((size_t)r * TOPO_MAX_EVENTS + fidx[k]);
                row[0] = a.top;
                row[1] = a.h_top;
                row[2] = a.h_free;
                row[3] = hn;
            } else {
                cur.bins[a.s_free].tw = make_tw(X.type_to2, tw2 & (TW_OUTSIDE | TW_NEW), r) | n2;
            }
            TT.ttype[R.tt_base + a.top] = X.top_type_to1;
        }
        if (TT.serial) TT.treact[R.tt_base + a.top] = TT.serial;
        atomicAdd(&TT.cnt_step[(size_t)r * TT.n_sr + sr], 1ull);
        atomicAdd(&TT.cnt_total[(size_t)r * TT.n_sr + sr], 1ull);
        if (n1) edit(a.s_top, tw_type(tw1), a.h_top);
        if (n2) edit(a.s_free, tw_type(tw2), a.h_free);  // (its moved copy carries the handle it had)
    }
    if (s_nm) {  // the absorbed topologies' particles join the surviving ones (the merge rows are written)
        __syncthreads();
        const uint32_t *mrows = TT.mrg + 4ull * (size_t)r * TOPO_MAX_EVENTS;
        const uint32_t nt = n_topo0 + s_nf;
        for (uint32_t i = tid; i < nt; i += blockDim.x) {
            const uint32_t t = TT.topo_of[R.topo_base + i];
            for (uint32_t m = 0; m < s_nm; ++m)
                if (t == mrows[4 * m + 1]) TT.topo_of[R.topo_base + i] = mrows[4 * m];
        }
    }
    if (tid == 0) {
        TT.ev_count[r] = 0u;
        if (s_nf) {  // every thread read n_topo0 before the greedy pass's barrier
            R.n_topo = n_topo0 + s_nf;
            TT.fus_count[r] = s_nf;
        }
        if (s_nm) TT.mrg_count[r] = s_nm;
    }
}

// sets the species of topology particles of replica r (by handle, found through loc) and their
// topology flag by the new species' flavour; a handle whose record is not live at loc counts in *bad
__global__ void k_set_topo_types(Grid cur, const ReplicaDev *reps, uint32_t r, const uint32_t *loc,
                                 const uint32_t *handles, const uint32_t *types, const uint8_t *flavor, uint32_t n,
                                 uint32_t *bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ReplicaDev &R = reps[r];
    const uint32_t h = handles[i];
    const uint32_t s = loc[R.topo_base + h];
    const State st = cur.bins[s];
    if (st.h != h || (st.tw & TW_DEAD) || tw_replica(st.tw) != r) {
        atomicAdd(bad, 1u);
        return;
    }
    const uint32_t t = types[i];
    cur.bins[s].tw = make_tw(t, (flavor[t] ? TW_TOPO : 0u) | (st.tw & TW_OUTSIDE), r);
}

// moves topology particles of replica r (by handle, found through loc) to new positions in place; the caller
// re-bins afterwards.  A handle whose record is not live at loc counts in *bad.
__global__ void k_set_topo_positions(Grid cur, const ReplicaDev *reps, uint32_t r, const uint32_t *loc,
                                     const uint32_t *handles, const double *pos, uint32_t n, uint32_t *bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ReplicaDev &R = reps[r];
    const uint32_t h = handles[i];
    const uint32_t s = loc[R.topo_base + h];
    const State st = cur.bins[s];
    if (st.h != h || (st.tw & TW_DEAD) || tw_replica(st.tw) != r || !(st.tw & TW_TOPO)) {
        atomicAdd(bad, 1u);
        return;
    }
    cur.bins[s].x = pos[3ull * i];
    cur.bins[s].y = pos[3ull * i + 1];
    cur.bins[s].z = pos[3ull * i + 2];
}

// new topology particles of replica r under handles h0, h0 + 1, ...: binned into cur, their slots in loc
__global__ void k_append_topo(Grid cur, ReplicaDev *reps, uint32_t r, uint32_t h0, const uint32_t *types,
                              const double *pos, uint32_t n, uint32_t *loc, uint32_t cap, uint32_t *err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ReplicaDev &R = reps[r];
    const RepGeom geo = load_geom(R);
    State st;
    st.x = pos[3ull * i];
    st.y = pos[3ull * i + 1];
    st.z = pos[3ull * i + 2];
    st.tw = make_tw(types[i], TW_TOPO, r);
    st.h = h0 + i;
    const uint32_t slot = bin_insert(cur, geo, cap, st, err, false, 0);
    if (slot != 0xFFFFFFFFu) loc[R.topo_base + h0 + i] = slot;
}

}  // namespace rdy
