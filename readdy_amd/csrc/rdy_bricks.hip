// LDS-staged stencil kernels: one workgroup per brick of cells.
//
// A brick is a box of cells of one replica.  Its workgroup copies the brick plus a halo of
// `radius` cells (periodic images wrapped, non-periodic ones clipped) from the cell bins into LDS
// once, packs every cell's members contiguously, and then evaluates the (2r+1)^3 stencil of every
// particle of the brick against LDS.  Each cache line of the bins is read once per brick instead of
// once per neighbouring particle.  A brick whose halo holds more particles than the LDS budget
// takes a slow path that reads neighbours straight from the bins.
//
// Per axis a brick is either in halo mode (extent e, halo r on both sides; the host guarantees
// e + 2r <= dims on periodic axes, so halo cells are distinct) or in whole-axis mode (the brick
// spans the axis, no halo; neighbour coordinates wrap inside the brick), which also reproduces the
// de-duplicated adjacency of small grids (CellLinkedList.cpp:74-137).
#pragma once

#include "rdy_kernels.hip"

namespace rdy {

struct BrickDev {
    int32_t o[3];  // origin cell
    int32_t e[3];  // extent in cells
    int32_t h[3];  // halo width (0 in whole-axis mode)
    uint32_t r;    // replica
    uint32_t whole;  // bit a: axis a spans the grid
    uint32_t pad;
};

constexpr int BRICK_THREADS = 256;

struct BrickSmem {  // carved out of dynamic LDS
    State *S;          // [max_states]
    uint32_t *slot;    // [max_states]
    uint16_t *hc_of;   // [max_states]
    uint16_t *inner;   // [max_states]
    uint32_t *hcnt;    // [max_halo]
    uint32_t *hstart;  // [max_halo + 1]
    uint32_t *hcell;   // [max_halo]
};

__device__ __forceinline__ BrickSmem carve(char *base, uint32_t max_states, uint32_t max_halo) {
    BrickSmem m;
    char *p = base;
    m.S = (State *)p;
    p += sizeof(State) * max_states;
    m.slot = (uint32_t *)p;
    p += 4 * max_states;
    m.hcnt = (uint32_t *)p;
    p += 4 * max_halo;
    m.hstart = (uint32_t *)p;
    p += 4 * (max_halo + 1);
    m.hcell = (uint32_t *)p;
    p += 4 * max_halo;
    m.hc_of = (uint16_t *)p;
    p += 2 * max_states;
    m.inner = (uint16_t *)p;
    return m;
}

size_t brick_smem_bytes(uint32_t max_states, uint32_t max_halo) {
    return sizeof(State) * (size_t)max_states + 4 * (size_t)max_states + 4 * (size_t)max_halo +
           4 * ((size_t)max_halo + 1) + 4 * (size_t)max_halo + 2 * (size_t)max_states + 2 * (size_t)max_states + 64;
}

// stages buffer g around brick B into LDS; returns the number of staged records, or UINT32_MAX if
// they do not fit (the caller then takes the slow path).  Also returns the number of inner records.
__device__ uint32_t stage_brick(const Grid &g, const BrickDev &B, const RepGeom &geo, uint32_t cap, BrickSmem &m,
                                uint32_t max_states, uint32_t *n_inner_out, uint32_t *H) {
    __shared__ uint32_t s_total, s_inner, s_wave[BRICK_THREADS / 64];
    const uint32_t tid = threadIdx.x;
    H[0] = B.e[0] + 2 * B.h[0];
    H[1] = B.e[1] + 2 * B.h[1];
    H[2] = B.e[2] + 2 * B.h[2];
    const uint32_t nh = H[0] * H[1] * H[2];
    // 1. cell counters of the halo box
    for (uint32_t hc = tid; hc < nh; hc += BRICK_THREADS) {
        const uint32_t hz = hc % H[2], hy = (hc / H[2]) % H[1], hx = hc / (H[2] * H[1]);
        const uint32_t hh[3] = {hx, hy, hz};
        int32_t c[3];
        bool inside = true;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            int32_t x = B.o[a] - B.h[a] + (int32_t)hh[a];
            if (x < 0 || x >= geo.dims[a]) {
                if (geo.pbc[a]) x = (x % geo.dims[a] + geo.dims[a]) % geo.dims[a];
                else inside = false;
            }
            c[a] = x;
        }
        uint32_t n = 0, cell = 0xFFFFFFFFu;
        if (inside) {
            cell = geo.cell_base + ((uint32_t)c[0] * (uint32_t)geo.dims[1] + (uint32_t)c[1]) * (uint32_t)geo.dims[2] +
                   (uint32_t)c[2];
            n = g.counts[cell];
        }
        m.hcnt[hc] = n;
        m.hcell[hc] = cell;
    }
    __syncthreads();
    // 2. exclusive prefix over the halo cells (each thread a contiguous chunk)
    const uint32_t per = (nh + BRICK_THREADS - 1) / BRICK_THREADS;
    const uint32_t c0 = tid * per;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < per; ++k)
        if (c0 + k < nh) sum += m.hcnt[c0 + k];
    uint32_t total;
    uint32_t off = block_exclusive_scan(sum, &total);
    for (uint32_t k = 0; k < per; ++k)
        if (c0 + k < nh) {
            m.hstart[c0 + k] = off;
            off += m.hcnt[c0 + k];
        }
    if (tid == 0) {
        m.hstart[nh] = total;
        s_total = total;
    }
    __syncthreads();
    if (s_total > max_states) return 0xFFFFFFFFu;
    total = s_total;
    // 3. copy the records, one per thread, cell found by binary search over hstart
    for (uint32_t j = tid; j < total; j += BRICK_THREADS) {
        uint32_t lo = 0, hi = nh;  // largest hc with hstart[hc] <= j
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (m.hstart[mid] <= j) lo = mid;
            else hi = mid;
        }
        const uint32_t hc = lo, k = j - m.hstart[hc];
        m.hc_of[j] = (uint16_t)hc;
        if (k < cap) {
            const uint32_t s = m.hcell[hc] * cap + k;
            m.S[j] = g.bins[s];
            m.slot[j] = s;
        }
    }
    // overflow records of crowded cells (rare): scan the overflow list
    for (uint32_t hc = tid; hc < nh; hc += BRICK_THREADS) {
        const uint32_t n = m.hcnt[hc];
        if (n <= cap) continue;
        const uint32_t cell = m.hcell[hc];
        uint32_t no = *g.ovf_count;
        no = no < g.ovf_cap ? no : g.ovf_cap;
        uint32_t found = 0, base = m.hstart[hc] + cap;
        for (uint32_t o = 0; o < no && found < n - cap; ++o)
            if (g.ovf_cell[o] == cell) {
                m.S[base + found] = g.bins[g.nbin + o];
                m.slot[base + found] = g.nbin + o;
                ++found;
            }
        for (; found < n - cap; ++found) {  // counters ahead of a truncated overflow list: mark dead
            m.S[base + found].tw = TW_DEAD;
            m.slot[base + found] = 0xFFFFFFFFu;
        }
    }
    if (tid == 0) s_inner = 0;
    __syncthreads();
    // 4. compact the indices of the brick's own records
    for (uint32_t base = 0; base < total; base += BRICK_THREADS) {
        const uint32_t j = base + tid;
        bool in = false;
        if (j < total) {
            const uint32_t hc = m.hc_of[j];
            const uint32_t hz = hc % H[2], hy = (hc / H[2]) % H[1], hx = hc / (H[2] * H[1]);
            in = hx >= (uint32_t)B.h[0] && hx < (uint32_t)(B.h[0] + B.e[0]) && hy >= (uint32_t)B.h[1] &&
                 hy < (uint32_t)(B.h[1] + B.e[1]) && hz >= (uint32_t)B.h[2] && hz < (uint32_t)(B.h[2] + B.e[2]);
        }
        const unsigned long long mask = __ballot(in);
        const uint32_t w = tid / 64;
        if (__lane_id() == 0) s_wave[w] = (uint32_t)__popcll(mask);
        __syncthreads();
        uint32_t wbase = s_inner;
        for (uint32_t q = 0; q < w; ++q) wbase += s_wave[q];
        if (in) m.inner[wbase + (uint32_t)__popcll(mask & lanemask_lt())] = (uint16_t)j;
        __syncthreads();
        if (tid == 0)
            for (uint32_t q = 0; q < BRICK_THREADS / 64; ++q) s_inner += s_wave[q];
        __syncthreads();
    }
    *n_inner_out = s_inner;
    return total;
}

// neighbour halo cells of halo cell hc: (2r+1)^3 offsets in halo mode, wrapped axis lists in
// whole-axis mode (all cells when the axis has at most 2r cells)
template <class F>
__device__ __forceinline__ void for_each_halo_neighbor(const BrickDev &B, const RepGeom &geo, const uint32_t *H,
                                                       int radius, uint32_t hc, F &&f) {
    const uint32_t hz = hc % H[2], hy = (hc / H[2]) % H[1], hx = hc / (H[2] * H[1]);
    const uint32_t hh[3] = {hx, hy, hz};
    int lst[3][7];
    int nl[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (B.whole & (1u << a)) {
            nl[a] = axis_cells((int)hh[a], geo.dims[a], geo.pbc[a], radius, lst[a]);
        } else {
            int n = 0;
            for (int d = -radius; d <= radius; ++d) lst[a][n++] = (int)hh[a] + d;
            nl[a] = n;
        }
    }
    for (int a = 0; a < nl[0]; ++a)
        for (int b = 0; b < nl[1]; ++b) {
            const uint32_t row = ((uint32_t)lst[0][a] * H[1] + (uint32_t)lst[1][b]) * H[2];
            for (int c = 0; c < nl[2]; ++c) f(row + (uint32_t)lst[2][c]);
        }
}

// accumulates the pair forces/energy on staged record j from staged neighbours
__device__ __forceinline__ void staged_forces(const BrickSmem &m, const BrickDev &B, const RepGeom &geo,
                                              const uint32_t *H, int radius, const Tables &T, uint32_t j,
                                              const State &me, long long &fx, long long &fy, long long &fz,
                                              long long &e) {
    const PairIdx *prow = T.pidx + (size_t)tw_type(me.tw) * T.n_types;
    for_each_halo_neighbor(B, geo, H, radius, m.hc_of[j], [&](uint32_t hc2) {
        const uint32_t q1 = m.hstart[hc2 + 1];
        for (uint32_t q = m.hstart[hc2]; q < q1; ++q) {
            if (q == j) continue;
            const State o = m.S[q];
            if (o.tw & (TW_DEAD | TW_OUTSIDE)) continue;
            const PairIdx pi = prow[tw_type(o.tw)];
            if (pi.count == 0) continue;
            const double dx = min_image(o.x - me.x, geo.box[0], geo.pbc[0]);
            const double dy = min_image(o.y - me.y, geo.box[1], geo.pbc[1]);
            const double dz = min_image(o.z - me.z, geo.box[2], geo.pbc[2]);
            const double d2 = dist2(dx, dy, dz);
            if (!(d2 < pi.maxcut2)) continue;
            float pfx = 0.f, pfy = 0.f, pfz = 0.f, pe = 0.f;
            const float xf = (float)dx, yf = (float)dy, zf = (float)dz, d2f = (float)d2;
            for (uint32_t p = 0; p < pi.count; ++p) {
                const PotDev pt = T.pots[pi.start + p];
                if (d2 < pt.cut2) pe += pair_force_energy(pt, d2f, xf, yf, zf, pfx, pfy, pfz);
            }
            fx += to_fixed(pfx, 4294967296.0f);
            fy += to_fixed(pfy, 4294967296.0f);
            fz += to_fixed(pfz, 4294967296.0f);
            e += to_fixed(pe, 16777216.0f);
        }
    });
}

// the same sum reading neighbours from the bins (slow path for bricks that do not fit in LDS)
__device__ __forceinline__ void global_forces(const Grid &cur, const RepGeom &geo, uint32_t cap, int radius,
                                              const Tables &T, uint32_t s, const State &me, long long &fx,
                                              long long &fy, long long &fz, long long &e) {
    const PairIdx *prow = T.pidx + (size_t)tw_type(me.tw) * T.n_types;
    const uint32_t cl = cell_of_slot(cur, s, cap) - geo.cell_base;
    for_each_stencil_slot(geo, cur, cap, radius, cl, [&](uint32_t s2) {
        if (s2 == s) return;
        const State o = cur.bins[s2];
        if (o.tw & (TW_DEAD | TW_OUTSIDE)) return;
        const PairIdx pi = prow[tw_type(o.tw)];
        if (pi.count == 0) return;
        const double dx = min_image(o.x - me.x, geo.box[0], geo.pbc[0]);
        const double dy = min_image(o.y - me.y, geo.box[1], geo.pbc[1]);
        const double dz = min_image(o.z - me.z, geo.box[2], geo.pbc[2]);
        const double d2 = dist2(dx, dy, dz);
        if (!(d2 < pi.maxcut2)) return;
        float pfx = 0.f, pfy = 0.f, pfz = 0.f, pe = 0.f;
        const float xf = (float)dx, yf = (float)dy, zf = (float)dz, d2f = (float)d2;
        for (uint32_t p = 0; p < pi.count; ++p) {
            const PotDev pt = T.pots[pi.start + p];
            if (d2 < pt.cut2) pe += pair_force_energy(pt, d2f, xf, yf, zf, pfx, pfy, pfz);
        }
        fx += to_fixed(pfx, 4294967296.0f);
        fy += to_fixed(pfy, 4294967296.0f);
        fz += to_fixed(pfz, 4294967296.0f);
        e += to_fixed(pe, 16777216.0f);
    });
}

// block-wide sum of a fixed-point energy, one atomic per block
__device__ __forceinline__ void block_add_energy(ReplicaDev *reps, uint32_t r, long long e) {
    __shared__ long long s_e[BRICK_THREADS / 64];
    for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off);
    if (__lane_id() == 0) s_e[threadIdx.x / 64] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int w = 0; w < BRICK_THREADS / 64; ++w) t += s_e[w];
        if (t != 0) atomicAdd(&reps[r].energy, (unsigned long long)t);
    }
}

// one inner record: bonded force, optional force output, integration and re-binning
template <int MODE>
__device__ __forceinline__ void finish_particle(const Grid &nxt, ReplicaDev *reps, const ReplicaDev &R,
                                                const RepGeom &geo, const Tables &T, const Topo &topo, uint32_t cap,
                                                double dt, const double *noise, double *force_out, uint32_t s,
                                                State me, bool do_forces, long long fx, long long fy, long long fz,
                                                uint32_t *err) {
    const uint32_t r = tw_replica(me.tw);
    if (do_forces && (me.tw & TW_TOPO) && topo.fb) {
        long long *fb = topo.fb + 3ull * (R.topo_base + me.h);
        fx += fb[0];
        fy += fb[1];
        fz += fb[2];
        fb[0] = 0;
        fb[1] = 0;
        fb[2] = 0;
    }
    const double Fx = (double)fx * FORCE_INV, Fy = (double)fy * FORCE_INV, Fz = (double)fz * FORCE_INV;
    if (MODE != MODE_REBIN && force_out && do_forces) {
        force_out[3ull * s] = Fx;
        force_out[3ull * s + 1] = Fy;
        force_out[3ull * s + 2] = Fz;
    }
    if (MODE == MODE_FORCES) return;
    State nw = me;
    if (MODE == MODE_STEP && R.active) {
        // CPUEulerBDIntegrator.cpp:63-73
        const double D = T.diffusion[tw_type(me.tw)];
        double xi[3];
        if (noise) {
            xi[0] = noise[3ull * me.h];
            xi[1] = noise[3ull * me.h + 1];
            xi[2] = noise[3ull * me.h + 2];
        } else {
            uint32_t x4[4];
            philox(me.h, R.step, 0u, 0u, R.seed, x4);
            normals3(x4, xi);
        }
        const double sc = sqrt(2. * D * dt);
        const double kbt = T.kbt;
        const double F[3] = {Fx, Fy, Fz};
        double p[3] = {me.x, me.y, me.z};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double det = F[d] * D * dt / kbt;
            double rnd = xi[d] * sc;
            p[d] += rnd + det;
            p[d] = wrap(p[d], geo.box[d], geo.pbc[d]);
            if (!isfinite(p[d])) atomicAdd(&err[ERR_NONFINITE], 1u);
        }
        nw.x = p[0];
        nw.y = p[1];
        nw.z = p[2];
    }
    const uint32_t slot = bin_insert(nxt, geo, cap, nw, err, false, 0);
    if (slot != 0xFFFFFFFFu && (me.tw & TW_TOPO) && topo.loc) topo.loc[R.topo_base + me.h] = slot;
    (void)r;
}

// forces (+ Euler-Maruyama + re-bin) of every particle of one brick; MODE as in k_fused
template <int MODE>
__global__ __launch_bounds__(BRICK_THREADS) void k_brick(Grid cur, Grid nxt, ReplicaDev *reps, Tables T, Topo topo,
                                                         const BrickDev *bricks, uint32_t cap, int radius, double dt,
                                                         const double *noise, double *force_out, uint32_t max_states,
                                                         uint32_t max_halo, uint32_t *err) {
    extern __shared__ __align__(16) char smem[];
    const BrickDev B = bricks[xcd_block(blockIdx.x, gridDim.x)];
    const ReplicaDev &R = reps[B.r];
    if (!R.in_use) return;
    const RepGeom geo = load_geom(R);
    const bool forces_on = MODE == MODE_FORCES || (MODE == MODE_STEP && R.active);
    BrickSmem m = carve(smem, max_states, max_halo);
    uint32_t H[3], n_inner = 0;
    const uint32_t total = stage_brick(cur, B, geo, cap, m, max_states, &n_inner, H);
    long long e_blk = 0;
    if (total != 0xFFFFFFFFu) {
        for (uint32_t ii = threadIdx.x; ii < n_inner; ii += BRICK_THREADS) {
            const uint32_t j = m.inner[ii];
            const State me = m.S[j];
            if (me.tw & TW_DEAD) continue;
            long long fx = 0, fy = 0, fz = 0, e = 0;
            const bool df = MODE != MODE_REBIN && forces_on;
            if (df && !(me.tw & TW_OUTSIDE)) staged_forces(m, B, geo, H, radius, T, j, me, fx, fy, fz, e);
            e_blk += e;
            finish_particle<MODE>(nxt, reps, R, geo, T, topo, cap, dt, noise, force_out, m.slot[j], me, df, fx, fy,
                                  fz, err);
        }
    } else {
        // slow path: walk the brick's own cells in the bins
        const uint32_t ncell = (uint32_t)(B.e[0] * B.e[1] * B.e[2]);
        for (uint32_t ci = threadIdx.x; ci < ncell; ci += BRICK_THREADS) {
            const uint32_t lz = ci % B.e[2], ly = (ci / B.e[2]) % B.e[1], lx = ci / (B.e[2] * B.e[1]);
            const uint32_t cell = geo.cell_base +
                                  ((uint32_t)(B.o[0] + lx) * (uint32_t)geo.dims[1] + (uint32_t)(B.o[1] + ly)) *
                                      (uint32_t)geo.dims[2] +
                                  (uint32_t)(B.o[2] + lz);
            const uint32_t n = cur.counts[cell];
            const uint32_t no = n > cap ? (*cur.ovf_count < cur.ovf_cap ? *cur.ovf_count : cur.ovf_cap) : 0u;
            for (uint32_t k = 0; k < n && k < cap + no; ++k) {
                uint32_t s;
                if (k < cap) {
                    s = cell * cap + k;
                } else {  // k-th overflow record of this cell
                    uint32_t want = k - cap, seen = 0;
                    s = 0xFFFFFFFFu;
                    for (uint32_t o = 0; o < no; ++o)
                        if (cur.ovf_cell[o] == cell) {
                            if (seen == want) {
                                s = cur.nbin + o;
                                break;
                            }
                            ++seen;
                        }
                    if (s == 0xFFFFFFFFu) break;
                }
                const State me = cur.bins[s];
                if (me.tw & TW_DEAD) continue;
                long long fx = 0, fy = 0, fz = 0, e = 0;
                const bool df = MODE != MODE_REBIN && forces_on;
                if (df && !(me.tw & TW_OUTSIDE)) global_forces(cur, geo, cap, radius, T, s, me, fx, fy, fz, e);
                e_blk += e;
                finish_particle<MODE>(nxt, reps, R, geo, T, topo, cap, dt, noise, force_out, s, me, df, fx, fy, fz,
                                      err);
            }
        }
    }
    if (MODE != MODE_REBIN) block_add_energy(reps, B.r, e_blk);
}

// reaction candidate search of one brick on the current buffer (see k_search for the rules), fused
// with clearing the previous buffer's counters of the brick's cells
__global__ __launch_bounds__(BRICK_THREADS) void k_brick_search(Grid cur, uint32_t *old_counts, ReplicaDev *reps,
                                                                Tables T, const BrickDev *bricks, uint32_t cap,
                                                                int radius, int handler, Event *events,
                                                                uint32_t *ev_count, uint32_t max_states,
                                                                uint32_t max_halo, uint32_t *err) {
    extern __shared__ __align__(16) char smem[];
    const BrickDev B = bricks[xcd_block(blockIdx.x, gridDim.x)];
    const ReplicaDev &R = reps[B.r];
    const RepGeom geo = load_geom(R);
    // clear the old buffer's counters of this brick's cells (the old buffer is not read again)
    if (old_counts) {
        const uint32_t ncell = (uint32_t)(B.e[0] * B.e[1] * B.e[2]);
        for (uint32_t ci = threadIdx.x; ci < ncell; ci += BRICK_THREADS) {
            const uint32_t lz = ci % B.e[2], ly = (ci / B.e[2]) % B.e[1], lx = ci / (B.e[2] * B.e[1]);
            old_counts[geo.cell_base +
                       ((uint32_t)(B.o[0] + lx) * (uint32_t)geo.dims[1] + (uint32_t)(B.o[1] + ly)) *
                           (uint32_t)geo.dims[2] +
                       (uint32_t)(B.o[2] + lz)] = 0u;
        }
    }
    if (handler > 1 || !R.in_use || !R.active) return;
    BrickSmem m = carve(smem, max_states, max_halo);
    uint32_t H[3], n_inner = 0;
    const uint32_t total = stage_brick(cur, B, geo, cap, m, max_states, &n_inner, H);
    if (total == 0xFFFFFFFFu) {  // slow path: search from the brick's own cells straight from the bins
        const uint32_t ncell = (uint32_t)(B.e[0] * B.e[1] * B.e[2]);
        for (uint32_t ci = threadIdx.x; ci < ncell; ci += BRICK_THREADS) {
            const uint32_t lz = ci % B.e[2], ly = (ci / B.e[2]) % B.e[1], lx = ci / (B.e[2] * B.e[1]);
            const uint32_t cell = geo.cell_base +
                                  ((uint32_t)(B.o[0] + lx) * (uint32_t)geo.dims[1] + (uint32_t)(B.o[1] + ly)) *
                                      (uint32_t)geo.dims[2] +
                                  (uint32_t)(B.o[2] + lz);
            const uint32_t n = cur.counts[cell];
            const uint32_t no = n > cap ? (*cur.ovf_count < cur.ovf_cap ? *cur.ovf_count : cur.ovf_cap) : 0u;
            uint32_t seen = 0;
            for (uint32_t k = 0; k < n && k < cap; ++k) {
                const State me = cur.bins[cell * cap + k];
                if (!(me.tw & TW_DEAD))
                    search_one(cur, reps, T, cap, radius, handler, 0, events, ev_count, nullptr, 0, cell * cap + k, me,
                               err);
            }
            for (uint32_t o = 0; o < no && seen < n - cap; ++o)
                if (cur.ovf_cell[o] == cell) {
                    ++seen;
                    const State me = cur.bins[cur.nbin + o];
                    if (!(me.tw & TW_DEAD))
                        search_one(cur, reps, T, cap, radius, handler, 0, events, ev_count, nullptr, 0, cur.nbin + o,
                                   me, err);
                }
        }
        return;
    }
    const uint32_t r = B.r, step = R.step;
    const uint64_t seed = R.seed;
    for (uint32_t ii = threadIdx.x; ii < n_inner; ii += BRICK_THREADS) {
        const uint32_t j = m.inner[ii];
        const State me = m.S[j];
        if (me.tw & TW_DEAD) continue;
        const uint32_t t = tw_type(me.tw);
        const uint32_t s = m.slot[j];
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
        for (uint32_t k = T.o1idx[t]; k < T.o1idx[t + 1]; ++k) {
            const uint32_t rx = T.o1list[k];
            const RxDev X = T.rx[rx];
            if (!(X.rate > 0)) continue;
            uint32_t x4[4];
            philox(me.h, step, 1u | (rx << 8), 0u, seed, x4);
            if (!(u53(x4[0], x4[1]) < X.p_accept)) continue;
            emit(rx, s, s, me.h, me.h, event_key(handler, x4, X.rate));
        }
        if (!T.has_o2 || (me.tw & TW_OUTSIDE)) continue;
        const RxPairIdx *orow = T.o2idx + (size_t)t * T.n_types;
        for_each_halo_neighbor(B, geo, H, radius, m.hc_of[j], [&](uint32_t hc2) {
            const uint32_t q1 = m.hstart[hc2 + 1];
            for (uint32_t q = m.hstart[hc2]; q < q1; ++q) {
                if (q == j) continue;
                const State o = m.S[q];
                if (o.tw & (TW_DEAD | TW_OUTSIDE)) continue;
                const uint32_t t2 = tw_type(o.tw);
                if (t2 < t || (t2 == t && o.h < me.h)) continue;
                const RxPairIdx pi = orow[t2];
                if (pi.count == 0) continue;
                const double dx = min_image(me.x - o.x, geo.box[0], geo.pbc[0]);
                const double dy = min_image(me.y - o.y, geo.box[1], geo.pbc[1]);
                const double dz = min_image(me.z - o.z, geo.box[2], geo.pbc[2]);
                const double d2 = dist2(dx, dy, dz);
                const uint32_t lo = me.h < o.h ? me.h : o.h, hi = me.h < o.h ? o.h : me.h;
                for (uint32_t p = 0; p < pi.count; ++p) {
                    const uint32_t rx = T.o2list[pi.start + p];
                    const RxDev X = T.rx[rx];
                    if (!(X.rate > 0 && d2 < X.educt_dist2)) continue;
                    uint32_t x4[4];
                    philox(lo, step, 2u | (rx << 8), hi, seed, x4);
                    if (!(u53(x4[0], x4[1]) < X.p_accept)) continue;
                    bool me_first = (X.educts[0] != X.educts[1]) ? (t == X.educts[0]) : (me.h < o.h);
                    if (me_first) emit(rx, s, m.slot[q], me.h, o.h, event_key(handler, x4, X.rate));
                    else emit(rx, m.slot[q], s, o.h, me.h, event_key(handler, x4, X.rate));
                }
            }
        });
    }
}

}  // namespace rdy
