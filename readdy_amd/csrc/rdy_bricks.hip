// LDS-staged stencil kernels: one workgroup per brick of cells.
//
// A brick is a box of cells of one replica.  Its workgroup copies the brick plus a halo of
// `radius` cells (periodic images wrapped, non-periodic ones clipped) from the cell bins into LDS
// once, packs every cell's members contiguously (z-fastest, so the three z-neighbours of a cell
// form one contiguous run), and evaluates the (2r+1)^3 stencil of every particle of the brick
// against LDS.  Each staged record carries its position twice: FP64 (the exact reference gate
// d2 < rc^2 and the min-image displacement that feeds the pair math) and FP32 in the brick's
// local frame with periodic images already shifted, used as a conservative prefilter: a candidate
// is dropped in FP32 only when it is farther than the cutoff by more than the FP32 rounding bound,
// so the exact pair set is unchanged.  A brick whose halo holds more particles than the LDS budget
// takes a slow path that reads neighbours straight from the bins.
//
// Per axis a brick is either in halo mode (extent e, halo r on both sides; the host guarantees
// e + 2r <= dims on periodic axes, so halo cells are distinct and the local displacement is the
// minimum image) or in whole-axis mode (the brick spans the axis, no halo; neighbour coordinates
// wrap inside the brick), which also reproduces the de-duplicated adjacency of small grids
// (CellLinkedList.cpp:74-137).
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
constexpr uint32_t LDS_TABLE = 256;  // type pairs whose tables are staged in LDS (n_types <= 16)
constexpr uint32_t LDS_POTS = 64;

struct BrickSmem {  // carved out of dynamic LDS (24 bytes per staged record)
    float4 *P;         // [max_states] local-frame FP32 position, w = tw bits
    uint32_t *slot;    // [max_states] bin slot (the FP64 record is read from the bins when needed)
    uint16_t *hc_of;   // [max_states]
    uint16_t *inner;   // [max_states]
    uint32_t *hcnt;    // [max_halo]
    uint32_t *hstart;  // [max_halo + 1]
    uint32_t *hcell;   // [max_halo]
};

__device__ __forceinline__ BrickSmem carve(char *base, uint32_t max_states, uint32_t max_halo) {
    BrickSmem m;
    char *p = base;
    m.P = (float4 *)p;
    p += sizeof(float4) * max_states;
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
    return (sizeof(float4) + 4 + 2 + 2) * (size_t)max_states + 12 * (size_t)max_halo + 4 + 64;
}

struct Frame {       // local frame of a brick
    double org[3];   // real coordinate of local 0
    float L[3];      // box lengths (FP32) for whole-axis min image
    uint32_t whole;
    float margin;    // 16 * 2^-23 * largest local coordinate (FP32 rounding bound, 4x safety)
};

__device__ __forceinline__ Frame make_frame(const BrickDev &B, const RepGeom &geo) {
    Frame f;
    float M = 0.f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const bool w = B.whole & (1u << a);
        f.org[a] = w ? -.5 * geo.box[a] : (-.5 * geo.box[a] + (double)B.o[a] * geo.cs[a]);
        f.L[a] = (float)geo.box[a];
        const float ext = w ? (float)geo.box[a] : (float)((B.e[a] + 2 * B.h[a]) * geo.cs[a]);
        M = fmaxf(M, ext);
    }
    f.whole = B.whole;
    f.margin = 16.0f * M * 1.1920929e-07f;
    return f;
}

// stages buffer g around brick B into LDS; returns the number of staged records, or UINT32_MAX if
// they do not fit (the caller then takes the slow path).  Also returns the number of inner records.
__device__ uint32_t stage_brick(const Grid &g, const BrickDev &B, const RepGeom &geo, const Frame &F, uint32_t cap,
                                BrickSmem &m, uint32_t max_states, uint32_t *n_inner_out, uint32_t *H) {
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
    auto put = [&](uint32_t j, uint32_t hc, uint32_t s) {
        const State st = g.bins[s];
        m.slot[j] = s;
        const uint32_t hz = hc % H[2], hy = (hc / H[2]) % H[1], hx = hc / (H[2] * H[1]);
        const uint32_t hh[3] = {hx, hy, hz};
        const double p[3] = {st.x, st.y, st.z};
        float l[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double v = p[a] - F.org[a];
            if (!(F.whole & (1u << a))) {  // unwrap halo images into the local frame
                const int32_t u = B.o[a] - B.h[a] + (int32_t)hh[a];
                if (u < 0) v -= geo.box[a];
                else if (u >= geo.dims[a]) v += geo.box[a];
            }
            l[a] = (float)v;
        }
        m.P[j] = make_float4(l[0], l[1], l[2], __uint_as_float(st.tw));
    };
    for (uint32_t j = tid; j < total; j += BRICK_THREADS) {
        uint32_t lo = 0, hi = nh;  // largest hc with hstart[hc] <= j
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (m.hstart[mid] <= j) lo = mid;
            else hi = mid;
        }
        const uint32_t hc = lo, k = j - m.hstart[hc];
        m.hc_of[j] = (uint16_t)hc;
        const uint32_t cell = m.hcell[hc];
        if (k < cap && cell != 0xFFFFFFFFu) put(j, hc, cell * cap + k);
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
                put(base + found, hc, g.nbin + o);
                ++found;
            }
        for (; found < n - cap; ++found) {  // counters ahead of a truncated overflow list: dead records
            m.P[base + found] = make_float4(0.f, 0.f, 0.f, __uint_as_float(TW_DEAD));
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

// stencil around halo cell hc as contiguous runs of staged records: (x, y) rows times z-ranges.
// Halo-mode axes use offsets -r..r; whole-axis mode wraps inside the grid (all cells if dims <= 2r).
struct StencilIter {
    int nx, ny, nzr;
    int bx, by, bz;   // base coordinates
    int mx, my, mz;   // axis modes: 0 halo, 1 whole (all cells), 2 whole (wrapped offsets)
    int Dx, Dy, Dz;
    uint32_t H1, H2;
    int ia, ib, iz;
    uint32_t q, qe;
};

__device__ __forceinline__ void axis_mode(bool whole, int D, int radius, int hcoord, int &n, int &base, int &mode) {
    if (!whole) {
        n = 2 * radius + 1;
        base = hcoord - radius;
        mode = 0;
    } else if (D <= 2 * radius) {
        n = D;
        base = 0;
        mode = 1;
    } else {
        n = 2 * radius + 1;
        base = hcoord - radius;
        mode = 2;
    }
}

__device__ __forceinline__ int axis_elem(int k, int base, int mode, int D) {
    return mode == 2 ? ((base + k) % D + D) % D : base + k;
}

__device__ __forceinline__ void stencil_range(StencilIter &it, const uint32_t *hstart) {
    const int x = axis_elem(it.ia, it.bx, it.mx, it.Dx);
    const int y = axis_elem(it.ib, it.by, it.my, it.Dy);
    const uint32_t row = ((uint32_t)x * it.H1 + (uint32_t)y) * it.H2;
    int z0, z1;
    if (it.mz == 2) {
        z0 = axis_elem(it.iz, it.bz, 2, it.Dz);
        z1 = z0 + 1;
    } else if (it.mz == 1) {
        z0 = 0;
        z1 = it.Dz;
    } else {
        z0 = it.bz;
        z1 = z0 + it.Dz;  // halo mode: Dz holds the run length 2r+1
    }
    it.q = hstart[row + (uint32_t)z0];
    it.qe = hstart[row + (uint32_t)z1];
}

__device__ __forceinline__ StencilIter stencil_begin(const BrickDev &B, const RepGeom &geo, const uint32_t *H,
                                                     int radius, uint32_t hc, const uint32_t *hstart) {
    StencilIter it;
    const uint32_t hz = hc % H[2], hy = (hc / H[2]) % H[1], hx = hc / (H[2] * H[1]);
    axis_mode(B.whole & 1u, geo.dims[0], radius, (int)hx, it.nx, it.bx, it.mx);
    axis_mode(B.whole & 2u, geo.dims[1], radius, (int)hy, it.ny, it.by, it.my);
    int nz;
    axis_mode(B.whole & 4u, geo.dims[2], radius, (int)hz, nz, it.bz, it.mz);
    it.Dx = geo.dims[0];
    it.Dy = geo.dims[1];
    if (it.mz == 0) {  // one contiguous z-run of 2r+1 cells
        it.nzr = 1;
        it.Dz = 2 * radius + 1;
    } else if (it.mz == 1) {
        it.nzr = 1;
        it.Dz = geo.dims[2];
    } else {
        it.nzr = nz;
        it.Dz = geo.dims[2];
    }
    it.H1 = H[1];
    it.H2 = H[2];
    it.ia = it.ib = it.iz = 0;
    stencil_range(it, hstart);
    return it;
}

// advances to the next non-empty run; false when the stencil is exhausted
__device__ __forceinline__ bool stencil_next(StencilIter &it, const uint32_t *hstart) {
    while (true) {
        if (++it.iz >= it.nzr) {
            it.iz = 0;
            if (++it.ib >= it.ny) {
                it.ib = 0;
                if (++it.ia >= it.nx) return false;
            }
        }
        stencil_range(it, hstart);
        if (it.q < it.qe) return true;
    }
}

// pair tables staged in LDS for small type counts (block-uniform choice)
struct PairLds {
    PairIdx pidx[LDS_TABLE];
    float thr[LDS_TABLE];
    PotDev pots[LDS_POTS];
};

__device__ __forceinline__ bool load_pair_tables(PairLds &t, const Tables &T, float margin) {
    const uint32_t np = T.n_types * T.n_types;
    const bool fits = np <= LDS_TABLE;
    if (fits) {
        for (uint32_t k = threadIdx.x; k < np; k += BRICK_THREADS) {
            const PairIdx pi = T.pidx[k];
            t.pidx[k] = pi;
            // prefilter threshold: cutoff^2 widened by the FP32 rounding bound (negative: no potentials)
            t.thr[k] = pi.count ? pi.maxcut2_f * (1.0f + 3.0e-6f) + 2.0f * sqrtf(pi.maxcut2_f) * margin + margin * margin
                                : -1.0f;
        }
        for (uint32_t k = threadIdx.x; k < T.n_pots && k < LDS_POTS; k += BRICK_THREADS) t.pots[k] = T.pots[k];
    }
    __syncthreads();
    return fits;
}

// accumulates the pair forces/energy on staged record j from staged neighbours
__device__ __forceinline__ void staged_forces(const BrickSmem &m, const State *bins, const BrickDev &B,
                                              const RepGeom &geo, const Frame &F, const uint32_t *H, int radius,
                                              const Tables &T,
                                              const PairLds &pt, bool lds_tables, uint32_t j, const State &me,
                                              long long &fx, long long &fy, long long &fz, long long &e) {
    const uint32_t t = tw_type(me.tw);
    const uint32_t nt = T.n_types;
    const float4 pj = m.P[j];
    StencilIter it = stencil_begin(B, geo, H, radius, m.hc_of[j], m.hstart);
    if (!(it.q < it.qe) && !stencil_next(it, m.hstart)) return;
    while (true) {
        const uint32_t q = it.q;
        const float4 pq = m.P[q];
        it.q++;
        const uint32_t tw2 = __float_as_uint(pq.w);
        if (q != j && !(tw2 & (TW_DEAD | TW_OUTSIDE))) {
            const uint32_t k = t * nt + tw_type(tw2);
            float dx = pq.x - pj.x, dy = pq.y - pj.y, dz = pq.z - pj.z;
            if (F.whole) {
                if (F.whole & 1u) dx = dx > .5f * F.L[0] ? dx - F.L[0] : (dx <= -.5f * F.L[0] ? dx + F.L[0] : dx);
                if (F.whole & 2u) dy = dy > .5f * F.L[1] ? dy - F.L[1] : (dy <= -.5f * F.L[1] ? dy + F.L[1] : dy);
                if (F.whole & 4u) dz = dz > .5f * F.L[2] ? dz - F.L[2] : (dz <= -.5f * F.L[2] ? dz + F.L[2] : dz);
            }
            const float d2p = dx * dx + dy * dy + dz * dz;
            float thr;
            PairIdx pi;
            if (lds_tables) {
                thr = pt.thr[k];
            } else {
                pi = T.pidx[k];
                thr = pi.count ? pi.maxcut2_f * (1.0f + 3.0e-6f) + 2.0f * sqrtf(pi.maxcut2_f) * F.margin +
                                     F.margin * F.margin
                               : -1.0f;
            }
            if (d2p < thr) {
                if (lds_tables) pi = pt.pidx[k];
                const State o = bins[m.slot[q]];
                const double ddx = min_image(o.x - me.x, geo.box[0], geo.pbc[0]);
                const double ddy = min_image(o.y - me.y, geo.box[1], geo.pbc[1]);
                const double ddz = min_image(o.z - me.z, geo.box[2], geo.pbc[2]);
                const double d2 = dist2(ddx, ddy, ddz);
                if (d2 < pi.maxcut2) {
                    float pfx = 0.f, pfy = 0.f, pfz = 0.f, pe = 0.f;
                    const float xf = (float)ddx, yf = (float)ddy, zf = (float)ddz, d2f = (float)d2;
                    for (uint32_t p = 0; p < pi.count; ++p) {
                        const uint32_t pidx = pi.start + p;
                        const PotDev pot = (lds_tables && pidx < LDS_POTS) ? pt.pots[pidx] : T.pots[pidx];
                        if (d2 < pot.cut2) pe += pair_force_energy(pot, d2f, xf, yf, zf, pfx, pfy, pfz);
                    }
                    fx += to_fixed(pfx, 4294967296.0f);
                    fy += to_fixed(pfy, 4294967296.0f);
                    fz += to_fixed(pfz, 4294967296.0f);
                    e += to_fixed(pe, 16777216.0f);
                }
            }
        }
        if (!(it.q < it.qe) && !stencil_next(it, m.hstart)) break;
    }
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

// one particle: bonded force, optional force output, integration and re-binning
template <int MODE>
__device__ __forceinline__ void finish_particle(const Grid &nxt, const ReplicaDev &R, const RepGeom &geo,
                                                const Tables &T, const Topo &topo, uint32_t cap, double dt,
                                                const double *noise, double *force_out, uint32_t s, State me,
                                                bool do_forces, long long fx, long long fy, long long fz,
                                                uint32_t *err) {
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
        const double Fv[3] = {Fx, Fy, Fz};
        double p[3] = {me.x, me.y, me.z};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double det = Fv[d] * D * dt / kbt;
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
}

// visits the records of the brick's own cells straight from the bins: fn(slot, state)
template <class Fn>
__device__ __forceinline__ void for_each_own_record(const Grid &cur, const BrickDev &B, const RepGeom &geo,
                                                    uint32_t cap, Fn &&fn) {
    const uint32_t ncell = (uint32_t)(B.e[0] * B.e[1] * B.e[2]);
    for (uint32_t ci = threadIdx.x; ci < ncell; ci += BRICK_THREADS) {
        const uint32_t lz = ci % B.e[2], ly = (ci / B.e[2]) % B.e[1], lx = ci / (B.e[2] * B.e[1]);
        const uint32_t cell = geo.cell_base +
                              ((uint32_t)(B.o[0] + lx) * (uint32_t)geo.dims[1] + (uint32_t)(B.o[1] + ly)) *
                                  (uint32_t)geo.dims[2] +
                              (uint32_t)(B.o[2] + lz);
        const uint32_t n = cur.counts[cell];
        const uint32_t m = n < cap ? n : cap;
        for (uint32_t k = 0; k < m; ++k) fn(cell * cap + k, cur.bins[cell * cap + k]);
        if (n > cap) {
            uint32_t no = *cur.ovf_count;
            no = no < cur.ovf_cap ? no : cur.ovf_cap;
            uint32_t seen = 0;
            for (uint32_t o = 0; o < no && seen < n - cap; ++o)
                if (cur.ovf_cell[o] == cell) {
                    ++seen;
                    fn(cur.nbin + o, cur.bins[cur.nbin + o]);
                }
        }
    }
}

// forces (+ Euler-Maruyama + re-bin) of every particle of one brick; MODE as in k_fused
template <int MODE>
__global__ __launch_bounds__(BRICK_THREADS) void k_brick(Grid cur, Grid nxt, ReplicaDev *reps, Tables T, Topo topo,
                                                         const BrickDev *bricks, uint32_t cap, int radius, double dt,
                                                         const double *noise, double *force_out, uint32_t max_states,
                                                         uint32_t max_halo, uint32_t *err) {
    extern __shared__ __align__(16) char smem[];
    __shared__ PairLds ptab;
    const BrickDev B = bricks[xcd_block(blockIdx.x, gridDim.x)];
    const ReplicaDev &R = reps[B.r];
    if (!R.in_use) return;
    const RepGeom geo = load_geom(R);
    const bool forces_on = MODE == MODE_FORCES || (MODE == MODE_STEP && R.active);
    if (!forces_on) {
        // re-bin only (neighbour-list update, or a replica whose step budget is spent): no staging
        for_each_own_record(cur, B, geo, cap, [&](uint32_t s, const State &me) {
            if (!(me.tw & TW_DEAD))
                finish_particle<MODE_REBIN>(nxt, R, geo, T, topo, cap, dt, noise, force_out, s, me, false, 0, 0, 0,
                                            err);
        });
        return;
    }
    const Frame F = make_frame(B, geo);
    const bool lds_tables = load_pair_tables(ptab, T, F.margin);
    BrickSmem m = carve(smem, max_states, max_halo);
    uint32_t H[3], n_inner = 0;
    const uint32_t total = stage_brick(cur, B, geo, F, cap, m, max_states, &n_inner, H);
    long long e_blk = 0;
    if (total != 0xFFFFFFFFu) {
        for (uint32_t ii = threadIdx.x; ii < n_inner; ii += BRICK_THREADS) {
            const uint32_t j = m.inner[ii];
            if (__float_as_uint(m.P[j].w) & TW_DEAD) continue;
            const State me = cur.bins[m.slot[j]];
            long long fx = 0, fy = 0, fz = 0, e = 0;
            if (!(me.tw & TW_OUTSIDE))
                staged_forces(m, cur.bins, B, geo, F, H, radius, T, ptab, lds_tables, j, me, fx, fy, fz, e);
            e_blk += e;
            finish_particle<MODE>(nxt, R, geo, T, topo, cap, dt, noise, force_out, m.slot[j], me, true, fx, fy, fz,
                                  err);
        }
    } else {
        for_each_own_record(cur, B, geo, cap, [&](uint32_t s, const State &me) {
            if (me.tw & TW_DEAD) return;
            long long fx = 0, fy = 0, fz = 0, e = 0;
            if (!(me.tw & TW_OUTSIDE)) global_forces(cur, geo, cap, radius, T, s, me, fx, fy, fz, e);
            e_blk += e;
            finish_particle<MODE>(nxt, R, geo, T, topo, cap, dt, noise, force_out, s, me, true, fx, fy, fz, err);
        });
    }
    block_add_energy(reps, B.r, e_blk);
}

// reaction candidate search of one brick on the current buffer (see search_one for the rules),
// fused with clearing the previous buffer's counters of the brick's cells
__global__ __launch_bounds__(BRICK_THREADS) void k_brick_search(Grid cur, uint32_t *old_counts, ReplicaDev *reps,
                                                                Tables T, const BrickDev *bricks, uint32_t cap,
                                                                int radius, int handler, Event *events,
                                                                uint32_t *ev_count, uint32_t max_states,
                                                                uint32_t max_halo, uint32_t *err) {
    extern __shared__ __align__(16) char smem[];
    __shared__ float s_thr[LDS_TABLE];
    __shared__ RxPairIdx s_o2[LDS_TABLE];
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
    const uint32_t nt = T.n_types;
    const bool lds_tables = nt * nt <= LDS_TABLE;
    const Frame F = make_frame(B, geo);
    if (lds_tables) {
        for (uint32_t k = threadIdx.x; k < nt * nt; k += BRICK_THREADS) {
            const RxPairIdx pi = T.o2idx[k];
            s_o2[k] = pi;
            s_thr[k] = pi.count ? pi.maxr2_f * (1.0f + 3.0e-6f) + 2.0f * sqrtf(pi.maxr2_f) * F.margin +
                                      F.margin * F.margin
                                : -1.0f;
        }
    }
    __syncthreads();
    BrickSmem m = carve(smem, max_states, max_halo);
    uint32_t H[3], n_inner = 0;
    const uint32_t total = (T.has_o2 || T.o1idx[nt] > 0) ? stage_brick(cur, B, geo, F, cap, m, max_states, &n_inner, H)
                                                          : 0u;
    if (total == 0xFFFFFFFFu) {  // slow path: search from the brick's own cells straight from the bins
        for_each_own_record(cur, B, geo, cap, [&](uint32_t s, const State &me) {
            if (!(me.tw & TW_DEAD))
                search_one(cur, reps, T, cap, radius, handler, 0, events, ev_count, nullptr, 0, s, me, err);
        });
        return;
    }
    const uint32_t r = B.r, step = R.step;
    const uint64_t seed = R.seed;
    for (uint32_t ii = threadIdx.x; ii < n_inner; ii += BRICK_THREADS) {
        const uint32_t j = m.inner[ii];
        if (__float_as_uint(m.P[j].w) & TW_DEAD) continue;
        const uint32_t s = m.slot[j];
        const State me = cur.bins[s];
        const uint32_t t = tw_type(me.tw);
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
        const float4 pj = m.P[j];
        StencilIter it = stencil_begin(B, geo, H, radius, m.hc_of[j], m.hstart);
        if (!(it.q < it.qe) && !stencil_next(it, m.hstart)) continue;
        while (true) {
            const uint32_t q = it.q;
            const float4 pq = m.P[q];
            it.q++;
            const uint32_t tw2 = __float_as_uint(pq.w);
            const uint32_t t2 = tw_type(tw2);
            if (q != j && !(tw2 & (TW_DEAD | TW_OUTSIDE)) && (t2 > t || (t2 == t && cur.bins[m.slot[q]].h > me.h))) {
                const uint32_t kk = t * nt + t2;
                float dx = pq.x - pj.x, dy = pq.y - pj.y, dz = pq.z - pj.z;
                if (F.whole) {
                    if (F.whole & 1u) dx = dx > .5f * F.L[0] ? dx - F.L[0] : (dx <= -.5f * F.L[0] ? dx + F.L[0] : dx);
                    if (F.whole & 2u) dy = dy > .5f * F.L[1] ? dy - F.L[1] : (dy <= -.5f * F.L[1] ? dy + F.L[1] : dy);
                    if (F.whole & 4u) dz = dz > .5f * F.L[2] ? dz - F.L[2] : (dz <= -.5f * F.L[2] ? dz + F.L[2] : dz);
                }
                const float d2p = dx * dx + dy * dy + dz * dz;
                RxPairIdx pi;
                float thr;
                if (lds_tables) {
                    thr = s_thr[kk];
                    pi = s_o2[kk];
                } else {
                    pi = T.o2idx[kk];
                    thr = pi.count ? pi.maxr2_f * (1.0f + 3.0e-6f) + 2.0f * sqrtf(pi.maxr2_f) * F.margin +
                                         F.margin * F.margin
                                   : -1.0f;
                }
                if (d2p < thr) {
                    const State o = cur.bins[m.slot[q]];
                    // distSquared(neighbor.pos, entry.pos) (CPUUncontrolledApproximation.cpp:108)
                    const double ddx = min_image(me.x - o.x, geo.box[0], geo.pbc[0]);
                    const double ddy = min_image(me.y - o.y, geo.box[1], geo.pbc[1]);
                    const double ddz = min_image(me.z - o.z, geo.box[2], geo.pbc[2]);
                    const double d2 = dist2(ddx, ddy, ddz);
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
            }
            if (!(it.q < it.qe) && !stencil_next(it, m.hstart)) break;
        }
    }
}

}  // namespace rdy
