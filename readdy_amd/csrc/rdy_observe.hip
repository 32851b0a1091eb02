// Device observables.
//
// Radial distribution (RadialDistribution::evaluate, readdy/main/model/observables/
// RadialDistribution.cpp:75-124).  The reference loops over every ordered pair of particles on the
// host (O(N^2)); here each "from" particle visits the cells within the largest border of its own
// cell (per axis ceil(r_max / cell size) cells, the de-duplicated stencil of the cell list), so the
// work is O(N * density * r_max^3).  Counts are integers, binned exactly as the reference bins them:
// dist = sqrt(distSquared(from, to)) with the minimum image, bin = upper_bound(borders, dist) - 1,
// nothing counted at or beyond the last border.  The normalisation (shell volume, number of "from"
// particles, particleToDensity) is applied by the caller, in FP64, as the reference does.
#pragma once

#include "rdy_step.hip"

namespace rdy {

constexpr uint32_t RDF_MAX_BORDERS = 4097;  // 4096 bins: one LDS histogram per block

__device__ __forceinline__ bool in_mask(const uint32_t *mask, uint32_t t) { return (mask[t >> 5] >> (t & 31u)) & 1u; }

// std::upper_bound over borders[0 .. nb): the first index whose border exceeds dist (nb if none)
__device__ __forceinline__ uint32_t upper_bound_border(const double *b, uint32_t nb, double dist) {
    uint32_t lo = 0, hi = nb;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (dist < b[mid]) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// One thread per live record of the work list.  Records of replica r whose type is in from_mask are
// counted into *n_from; those inside the box count their pairs with the in-box "to" records of the
// stencil of radius rad[a] cells; records outside the box (never binned as neighbours) are listed
// for k_rdf_outside.  Per-block histogram in LDS, added to counts[nb - 1] (64-bit) at the end.
__global__ __launch_bounds__(256) void k_rdf(Grid cur, const ReplicaDev *reps, uint32_t r, uint32_t cap,
                                             const double *borders, uint32_t nb, const uint32_t *from_mask,
                                             const uint32_t *to_mask, int rad0, int rad1, int rad2,
                                             unsigned long long *counts, unsigned long long *n_from,
                                             uint32_t *outside, uint32_t outside_cap, uint32_t *n_outside,
                                             uint32_t max_wl) {
    __shared__ double sb[RDF_MAX_BORDERS];
    __shared__ uint32_t hist[RDF_MAX_BORDERS - 1];
    for (uint32_t k = threadIdx.x; k < nb; k += blockDim.x) sb[k] = borders[k];
    for (uint32_t k = threadIdx.x; k + 1 < nb; k += blockDim.x) hist[k] = 0u;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = *cur.wl_count;
    n = n < max_wl ? n : max_wl;
    bool from = false;
    if (i < n) {
        const uint32_t s = cur.wl[i];
        const State me = cur.bins[s];
        if (!(me.tw & TW_DEAD) && tw_replica(me.tw) == r) {
            from = in_mask(from_mask, tw_type(me.tw));
            if (me.tw & TW_OUTSIDE) {
                const uint32_t p = atomicAdd(n_outside, 1u);
                if (p < outside_cap) outside[p] = s;
            } else if (from) {
                const RepGeom geo = load_geom(reps[r]);
                const uint32_t cl = cell_of_slot(cur, s, cap) - geo.cell_base;
                const int nz = geo.dims[2], ny = geo.dims[1];
                const int c[3] = {(int)(cl / (uint32_t)(nz * ny)), (int)((cl / (uint32_t)nz) % (uint32_t)ny),
                                  (int)(cl % (uint32_t)nz)};
                const int rad[3] = {rad0, rad1, rad2};
                int cnt[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) cnt[a] = axis_count(c[a], geo.dims[a], geo.pbc[a], rad[a]);
                for (int ka = 0; ka < cnt[0]; ++ka) {
                    const int x = axis_elem(ka, c[0], geo.dims[0], geo.pbc[0], rad[0]);
                    for (int kb = 0; kb < cnt[1]; ++kb) {
                        const int y = axis_elem(kb, c[1], geo.dims[1], geo.pbc[1], rad[1]);
                        const uint32_t row = geo.cell_base + ((uint32_t)x * (uint32_t)ny + (uint32_t)y) * (uint32_t)nz;
                        for (int kc = 0; kc < cnt[2]; ++kc) {
                            const uint32_t cell = row + (uint32_t)axis_elem(kc, c[2], nz, geo.pbc[2], rad[2]);
                            cell_records(cur, cap, cell, [&](uint32_t s2) {
                                if (s2 == s) return;
                                const State o = cur.bins[s2];
                                if ((o.tw & (TW_DEAD | TW_OUTSIDE)) || !in_mask(to_mask, tw_type(o.tw))) return;
                                // distSquared(from, to): the difference to - from, minimum image per axis
                                const double dx = min_image(o.x - me.x, geo.box[0], geo.pbc[0]);
                                const double dy = min_image(o.y - me.y, geo.box[1], geo.pbc[1]);
                                const double dz = min_image(o.z - me.z, geo.box[2], geo.pbc[2]);
                                const uint32_t u = upper_bound_border(sb, nb, sqrt(dist2(dx, dy, dz)));
                                if (u > 0 && u < nb) atomicAdd(&hist[u - 1], 1u);
                            });
                        }
                    }
                }
            }
        }
    }
    const unsigned long long fm = __ballot(from);
    if ((threadIdx.x & 63u) == 0u && fm) atomicAdd(n_from, (unsigned long long)__popcll(fm));
    __syncthreads();
    for (uint32_t k = threadIdx.x; k + 1 < nb; k += blockDim.x)
        if (hist[k]) atomicAdd(&counts[k], (unsigned long long)hist[k]);
}

// The pairs that involve a record outside the box: thread (o, i) pairs listed record o with work-list
// record i of the replica; o -> i when o is "from" and i is "to", i -> o when i is an in-box "from"
// record and o is "to" (a pair of two outside records is counted from each side by its own row).
__global__ __launch_bounds__(256) void k_rdf_outside(Grid cur, const ReplicaDev *reps, uint32_t r,
                                                     const double *borders, uint32_t nb, const uint32_t *from_mask,
                                                     const uint32_t *to_mask, unsigned long long *counts,
                                                     const uint32_t *outside, const uint32_t *n_outside,
                                                     uint32_t max_wl) {
    uint32_t n = *cur.wl_count;
    n = n < max_wl ? n : max_wl;
    const uint32_t no = *n_outside;
    const unsigned long long idx = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n == 0 || idx >= (unsigned long long)no * n) return;
    const uint32_t s = outside[idx / n], s2 = cur.wl[idx % n];
    if (s2 == s) return;
    const State a = cur.bins[s], b = cur.bins[s2];
    if ((b.tw & TW_DEAD) || tw_replica(b.tw) != r) return;
    const RepGeom geo = load_geom(reps[r]);
    auto count = [&](const State &f, const State &t) {
        const double dx = min_image(t.x - f.x, geo.box[0], geo.pbc[0]);
        const double dy = min_image(t.y - f.y, geo.box[1], geo.pbc[1]);
        const double dz = min_image(t.z - f.z, geo.box[2], geo.pbc[2]);
        const uint32_t u = upper_bound_border(borders, nb, sqrt(dist2(dx, dy, dz)));
        if (u > 0 && u < nb) atomicAdd(&counts[u - 1], 1ull);
    };
    if (in_mask(from_mask, tw_type(a.tw)) && in_mask(to_mask, tw_type(b.tw))) count(a, b);
    if (!(b.tw & TW_OUTSIDE) && in_mask(from_mask, tw_type(b.tw)) && in_mask(to_mask, tw_type(a.tw))) count(b, a);
}

// HistogramAlongAxis as the CPU kernel evaluates it (kernels/cpu/src/observables/CPUObservables.cpp:
// 85-136): each live record of replica r whose type is in mask adds one to bin u - 1, u =
// upper_bound(borders, position[axis]), when 1 <= u < nb - 1 (the CPU kernel compares u with the bin
// count, so the last bin stays empty).  One thread per work-list record, per-block LDS histogram.
__global__ __launch_bounds__(256) void k_hist_axis(Grid cur, uint32_t r, const double *borders, uint32_t nb,
                                                   const uint32_t *mask, int axis, unsigned long long *counts,
                                                   uint32_t max_wl) {
    __shared__ double sb[RDF_MAX_BORDERS];
    __shared__ uint32_t hist[RDF_MAX_BORDERS - 1];
    for (uint32_t k = threadIdx.x; k < nb; k += blockDim.x) sb[k] = borders[k];
    for (uint32_t k = threadIdx.x; k + 1 < nb; k += blockDim.x) hist[k] = 0u;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = *cur.wl_count;
    n = n < max_wl ? n : max_wl;
    if (i < n) {
        const State me = cur.bins[cur.wl[i]];
        if (!(me.tw & TW_DEAD) && tw_replica(me.tw) == r && in_mask(mask, tw_type(me.tw))) {
            const double x = axis == 0 ? me.x : (axis == 1 ? me.y : me.z);
            const uint32_t u = upper_bound_border(sb, nb, x);
            if (u >= 1u && u + 1u < nb) atomicAdd(&hist[u - 1], 1u);
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k + 1 < nb; k += blockDim.x)
        if (hist[k]) atomicAdd(&counts[k], (unsigned long long)hist[k]);
}

}  // namespace rdy
