// LDS-staged stencil kernels: one workgroup per brick of cells.
//
// A brick is a box of cells of one replica.  Its workgroup copies the brick plus a halo of
// `radius` cells (periodic images wrapped, non-periodic ones clipped) from the cell bins into LDS
// once, packs every cell's members contiguously (z-fastest, so the z-neighbours of a cell form one
// contiguous run), and walks the (2r+1)^3 stencil of every particle of the brick against LDS.
// Each staged record carries an FP32 position in the brick's local frame with periodic images
// already shifted, used as a conservative prefilter: a candidate is dropped in FP32 only when it
// is farther than the cutoff by more than the FP32 rounding bound, so the exact pair set is
// unchanged.  The exact work (FP64 min-image displacement, the reference gate d2 < rc^2, the pair
// math, the reaction draws) runs on the survivors only, and not per lane: each wave streams its
// survivors through a ring in LDS and evaluates them 64 at a time with every lane busy, because a
// per-lane evaluation would make the whole wave execute the exact branch whenever any one lane
// had a survivor (almost always).  Forces are gathered per particle as 64-bit fixed-point sums
// (LDS atomics; integer addition is order-free, so the result stays deterministic).
// A brick whose halo holds more particles than the LDS budget takes a slow path that reads
// neighbours straight from the bins.
//
// Per axis a brick is either in halo mode (extent e, halo r on both sides; the host guarantees
// e + 2r <= dims on periodic axes, so halo cells are distinct and the local displacement is the
// minimum image) or in whole-axis mode (the brick spans the axis, no halo; neighbour coordinates
// wrap inside the brick), which also reproduces the de-duplicated adjacency of small grids
// (CellLinkedList.cpp:74-137).
#pragma once

#include "rdy_kernels.hip"

#include <type_traits>

namespace rdy {

struct BrickDev {
    int32_t o[3];    // origin cell
    int32_t e[3];    // extent in cells
    int32_t h[3];    // halo width (0 in whole-axis mode)
    uint32_t r;      // replica
    uint32_t whole;  // bit a: axis a spans the grid
    // ceil(2^32 / d) for d = H1 * H2, H2 and e1 (H = e + 2h; 0 when d == 1): exact n / d as a high
    // multiply for n, d < 2^16
    uint32_t m_hh, m_h2, m_e1;
    // the replica's grid, so that the halo's cells are known before the replica's record arrives
    uint32_t cell_base;
    int32_t dims[3];
    uint32_t pbc;    // bit a: the replica's axis a is periodic
    uint32_t pad;
};
static_assert(sizeof(BrickDev) == 80, "brick record size");

constexpr int BRICK_THREADS = 256;
constexpr int BRICK_WAVES = BRICK_THREADS / 64;
constexpr uint32_t BATCH = 4;       // candidates a lane takes from its current run per iteration
constexpr uint32_t RING = 512;      // per-wave ring of prefilter survivors (power of two, >= 63 + 64 * BATCH)
constexpr uint32_t NO_ENTRY = 0xFFFFFFFFu;  // an idle lane's entry in the last batch of the ring

__host__ __device__ inline uint32_t div_magic(uint32_t d) {
    return d > 1 ? (uint32_t)((0x100000000ull + d - 1) / d) : 0u;
}
__device__ __forceinline__ uint32_t fdiv(uint32_t n, uint32_t m) { return m ? __umulhi(n, m) : n; }

struct Halo {  // halo box of a brick, z fastest: hc = (hx * H1 + hy) * H2 + hz
    uint32_t H0, H1, H2, HH, nh, m_hh, m_h2;
    __device__ __forceinline__ void decode(uint32_t hc, uint32_t &hx, uint32_t &hy, uint32_t &hz) const {
        hx = fdiv(hc, m_hh);
        const uint32_t rem = hc - hx * HH;
        hy = fdiv(rem, m_h2);
        hz = rem - hy * H2;
    }
};

__device__ __forceinline__ Halo make_halo(const BrickDev &B) {
    Halo s;
    s.H0 = (uint32_t)(B.e[0] + 2 * B.h[0]);
    s.H1 = (uint32_t)(B.e[1] + 2 * B.h[1]);
    s.H2 = (uint32_t)(B.e[2] + 2 * B.h[2]);
    s.HH = s.H1 * s.H2;
    s.nh = s.H0 * s.HH;
    s.m_hh = B.m_hh;
    s.m_h2 = B.m_h2;
    return s;
}

struct BrickSmem {  // carved out of dynamic LDS
    float4 *P;         // [max_states] local-frame FP32 position, w = tw bits
    uint32_t *slot;    // [max_states] bin slot (the FP64 record is read from the bins when needed)
    uint32_t *hstart;  // [max_halo + 1]
    uint16_t *hc_of;   // [max_states]
    uint16_t *inner;   // [max_states]
    // transient: the halo cells' counters and ids while staging, then the waves' survivor rings
    uint32_t *hcnt;    // [max_halo]
    uint32_t *hcell;   // [max_halo]
    uint32_t *rings;   // [BRICK_WAVES][RING]
};

__host__ __device__ inline size_t transient_bytes(uint32_t max_halo) {
    const size_t staging = 8 * (size_t)max_halo, rings = 4 * (size_t)BRICK_WAVES * RING;
    return staging > rings ? staging : rings;
}

__device__ __forceinline__ BrickSmem carve(char *base, uint32_t max_states, uint32_t max_halo) {
    BrickSmem m;
    char *p = base;
    m.P = (float4 *)p;
    p += sizeof(float4) * max_states;
    m.slot = (uint32_t *)p;
    p += 4 * max_states;
    m.hstart = (uint32_t *)p;
    p += 4 * (max_halo + 1);
    m.hc_of = (uint16_t *)p;
    p += 2 * max_states;
    m.inner = (uint16_t *)p;
    p += 2 * max_states;
    p = base + ((p - base + 15) & ~(ptrdiff_t)15);
    m.hcnt = (uint32_t *)p;
    m.hcell = m.hcnt + max_halo;
    m.rings = (uint32_t *)p;
    return m;
}

__host__ __device__ inline size_t brick_smem_bytes(uint32_t max_states, uint32_t max_halo) {
    const size_t fixed = (sizeof(float4) + 4 + 2 + 2) * (size_t)max_states + 4 * ((size_t)max_halo + 1);
    return ((fixed + 15) & ~(size_t)15) + transient_bytes(max_halo);
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

// prefilter threshold: cutoff^2 widened by the FP32 rounding bound (negative: nothing to find)
__device__ __forceinline__ float prefilter_thr(float cut2_f, bool any, float margin) {
    return any ? cut2_f * (1.0f + 3.0e-6f) + 2.0f * sqrtf(cut2_f) * margin + margin * margin : -1.0f;
}

// the global cell of halo cell hc of brick B (periodic images wrapped), UINT32_MAX outside a
// non-periodic edge; from the brick record alone
__device__ __forceinline__ uint32_t halo_global_cell(const BrickDev &B, const Halo &Hs, uint32_t hc) {
    uint32_t hh[3];
    Hs.decode(hc, hh[0], hh[1], hh[2]);
    int32_t c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int32_t x = B.o[a] - B.h[a] + (int32_t)hh[a];
        if (x < 0 || x >= B.dims[a]) {
            if (!(B.pbc & (1u << a))) return 0xFFFFFFFFu;
            x = x < 0 ? x + B.dims[a] : x - B.dims[a];  // |overhang| <= r < dims
        }
        c[a] = x;
    }
    return B.cell_base + ((uint32_t)c[0] * (uint32_t)B.dims[1] + (uint32_t)c[1]) * (uint32_t)B.dims[2] + (uint32_t)c[2];
}

// Staging a brick: the counters of each thread's contiguous chunk of halo cells (one global round
// trip, issued from the brick record alone, so that they travel with the replica's record), one
// block scan that places every halo cell's records and lists the brick's own (inner) records, then
// the records themselves, spread evenly over the threads (the second round trip).
constexpr uint32_t STAGE_UNROLL = 4;  // loads of one thread in flight together

struct HaloCounts {  // this thread's chunk of halo cells (registers): global cell and counter
    uint32_t cl[STAGE_UNROLL], n[STAGE_UNROLL];
};

// chunk of thread tid: halo cells tid * per .. tid * per + per - 1 (per <= STAGE_UNROLL; larger halos
// load their counters inside stage_brick)
__device__ __forceinline__ HaloCounts load_halo_counts(const Grid &g, const BrickDev &B, const Halo &Hs) {
    HaloCounts h;
    const uint32_t nh = Hs.nh, per = (nh + BRICK_THREADS - 1) / BRICK_THREADS, c0 = threadIdx.x * per;
#pragma unroll
    for (uint32_t u = 0; u < STAGE_UNROLL; ++u) {
        h.cl[u] = 0xFFFFFFFFu;
        h.n[u] = 0;
        if (per <= STAGE_UNROLL && u < per && c0 + u < nh) {
            h.cl[u] = halo_global_cell(B, Hs, c0 + u);
            if (h.cl[u] != 0xFFFFFFFFu) h.n[u] = g.counts[h.cl[u]];
        }
    }
    return h;
}

// inclusive-exclusive pair scan: the exclusive prefixes of a and b over the block and their totals
// (two values, one pair of barriers)
__device__ __forceinline__ uint2 block_exclusive_scan2(uint32_t a, uint32_t b, uint2 *total) {
    __shared__ uint2 wave_sums[BRICK_WAVES];
    const uint32_t lane = __lane_id(), wid = threadIdx.x / 64;
    uint32_t x = a, y = b;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t xa = __shfl_up(x, off), ya = __shfl_up(y, off);
        if (lane >= (uint32_t)off) {
            x += xa;
            y += ya;
        }
    }
    if (lane == 63) wave_sums[wid] = make_uint2(x, y);
    __syncthreads();
    uint2 base = make_uint2(0u, 0u), tot = make_uint2(0u, 0u);
#pragma unroll
    for (uint32_t w = 0; w < BRICK_WAVES; ++w) {
        const uint2 ws = wave_sums[w];
        if (w < wid) {
            base.x += ws.x;
            base.y += ws.y;
        }
        tot.x += ws.x;
        tot.y += ws.y;
    }
    __syncthreads();
    *total = tot;
    return make_uint2(base.x + x - a, base.y + y - b);
}

// stages buffer g around brick B into LDS; returns the number of staged records, or UINT32_MAX if
// they do not fit (the caller then takes the slow path).  Also returns the number of inner records
// (the brick's own cells), listed in m.inner in halo order.  hc: load_halo_counts of this brick.
__device__ uint32_t stage_brick(const Grid &g, const BrickDev &B, const RepGeom &geo, const Frame &F, uint32_t cap,
                                const BrickSmem &m, uint32_t max_states, const Halo &Hs, const HaloCounts &hcn,
                                uint32_t *n_inner_out) {
    __shared__ uint32_t s_total, s_inner;
    const uint32_t tid = threadIdx.x;
    const uint32_t nh = Hs.nh;
    const uint32_t per = (nh + BRICK_THREADS - 1) / BRICK_THREADS;
    const uint32_t c0 = tid * per;
    // a halo cell is inner when every coordinate lies in the brick's own range
    auto inner_cell = [&](uint32_t hc) -> bool {
        uint32_t hh[3];
        Hs.decode(hc, hh[0], hh[1], hh[2]);
        bool in = true;
#pragma unroll
        for (int a = 0; a < 3; ++a) in = in && hh[a] >= (uint32_t)B.h[a] && hh[a] < (uint32_t)(B.h[a] + B.e[a]);
        return in;
    };
    // 1. counters of this thread's halo cells c0 .. c0 + per - 1 (already loaded for small halos)
    uint32_t sum = 0, sum_in = 0;
    if (per <= STAGE_UNROLL) {
#pragma unroll
        for (uint32_t u = 0; u < STAGE_UNROLL; ++u) {
            const uint32_t hc = c0 + u;
            if (u < per && hc < nh) {
                m.hcnt[hc] = hcn.n[u];
                m.hcell[hc] = hcn.cl[u];
                sum += hcn.n[u];
                if (inner_cell(hc)) sum_in += hcn.n[u];
            }
        }
    } else {
        for (uint32_t hc = c0; hc < c0 + per && hc < nh; ++hc) {
            const uint32_t cl = halo_global_cell(B, Hs, hc);
            const uint32_t n = cl != 0xFFFFFFFFu ? g.counts[cl] : 0u;
            m.hcnt[hc] = n;
            m.hcell[hc] = cl;
            sum += n;
            if (inner_cell(hc)) sum_in += n;
        }
    }
    // 2. exclusive prefixes over the halo cells: every record slot learns its halo cell, and the
    // inner records are listed (in halo order)
    uint2 tot;
    const uint2 off2 = block_exclusive_scan2(sum, sum_in, &tot);
    uint32_t off = off2.x, off_in = off2.y;
    for (uint32_t hc = c0; hc < c0 + per && hc < nh; ++hc) {
        m.hstart[hc] = off;
        const uint32_t n = m.hcnt[hc];
        for (uint32_t k = 0; k < n && off + k < max_states; ++k) m.hc_of[off + k] = (uint16_t)hc;
        if (inner_cell(hc)) {
            for (uint32_t k = 0; k < n && off_in + k < max_states; ++k) m.inner[off_in + k] = (uint16_t)(off + k);
            off_in += n;
        }
        off += n;
    }
    if (tid == 0) {
        m.hstart[nh] = tot.x;
        s_total = tot.x;
        s_inner = tot.y;
    }
    __syncthreads();
    const uint32_t total = s_total;
    if (total > max_states) return 0xFFFFFFFFu;
    // 3. the records: thread tid copies records tid, tid + 256, ...
    auto fetch_slot = [&](uint32_t j, uint32_t &hc) -> uint32_t {
        hc = m.hc_of[j];
        const uint32_t k = j - m.hstart[hc], cell = m.hcell[hc];
        if (k < cap) return cell * cap + k;
        // the (k - cap)-th overflow record of a crowded cell (rare); beyond a truncated overflow
        // list the counter runs ahead of the records: a dead placeholder
        uint32_t no = *g.ovf_count;
        no = no < g.ovf_cap ? no : g.ovf_cap;
        for (uint32_t o = 0, seen = 0; o < no; ++o)
            if (g.ovf_cell[o] == cell && seen++ == k - cap) return g.nbin + o;
        return 0xFFFFFFFFu;
    };
    auto put = [&](uint32_t j, uint32_t hc, uint32_t s, const State &st) {
        if (s == 0xFFFFFFFFu) {
            m.P[j] = make_float4(__uint_as_float(0x7FC00000u), 0.f, 0.f, __uint_as_float(TW_DEAD));
            m.slot[j] = 0xFFFFFFFFu;
            return;
        }
        uint32_t hh[3];
        Hs.decode(hc, hh[0], hh[1], hh[2]);
        double shift[3];  // unwraps halo images into the local frame
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            shift[a] = 0.0;
            if (!(F.whole & (1u << a))) {
                const int32_t u = B.o[a] - B.h[a] + (int32_t)hh[a];
                if (u < 0) shift[a] = -geo.box[a];
                else if (u >= geo.dims[a]) shift[a] = geo.box[a];
            }
        }
        m.slot[j] = s;
        // a dead or outside record is never a neighbour: its NaN position fails every distance test
        const float nan = __uint_as_float(0x7FC00000u);
        const bool gone = (st.tw & (TW_DEAD | TW_OUTSIDE)) != 0u;
        const float lx = gone ? nan : (float)((st.x - F.org[0]) + shift[0]);
        const float ly = (float)((st.y - F.org[1]) + shift[1]);
        const float lz = (float)((st.z - F.org[2]) + shift[2]);
        m.P[j] = make_float4(lx, ly, lz, __uint_as_float(st.tw));
    };
    for (uint32_t j0 = 0; j0 < total; j0 += STAGE_UNROLL * BRICK_THREADS) {
        uint32_t sl[STAGE_UNROLL], hcs[STAGE_UNROLL];
        State st[STAGE_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < STAGE_UNROLL; ++u) {
            const uint32_t j = j0 + u * BRICK_THREADS + tid;
            sl[u] = 0xFFFFFFFFu;
            hcs[u] = 0;
            if (j < total) {
                sl[u] = fetch_slot(j, hcs[u]);
                if (sl[u] != 0xFFFFFFFFu) st[u] = g.bins[sl[u]];
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < STAGE_UNROLL; ++u) {
            const uint32_t j = j0 + u * BRICK_THREADS + tid;
            if (j < total) put(j, hcs[u], sl[u], st[u]);
        }
    }
    __syncthreads();
    *n_inner_out = s_inner;
    return total;
}

// ---- stencil walkers: visit the staged records of the (2r+1)^3 stencil of halo cell hc as
// contiguous runs [q, qe) -------------------------------------------------------------------------
struct WalkCtx {
    Halo Hs;
    uint32_t whole;
    int32_t dims[3];
    int radius;
    uint32_t n;         // 2r + 1
    int32_t h[3];       // halo widths
    float cs[3], inv_cs[3];
};

__device__ __forceinline__ WalkCtx make_walk(const BrickDev &B, const RepGeom &geo, const Halo &Hs, int radius) {
    WalkCtx c;
    c.Hs = Hs;
    c.whole = B.whole;
    c.dims[0] = geo.dims[0];
    c.dims[1] = geo.dims[1];
    c.dims[2] = geo.dims[2];
    c.radius = radius;
    c.n = 2u * (uint32_t)radius + 1u;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        c.h[a] = B.h[a];
        c.cs[a] = (float)geo.cs[a];
        c.inv_cs[a] = (float)(1.0 / geo.cs[a]);
    }
    return c;
}

template <bool ALL_HALO>
struct Walk;

// every axis in halo mode: the cells within `reach` of the particle at local position p in halo
// cell hc, per axis offsets lo..hi inside -r..r (cells that overlap [p - reach, p + reach]; reach
// carries the FP32 rounding slack, so the range is a superset of the exact one), walked as
// (hi0 - lo0 + 1) * (hi1 - lo1 + 1) z-runs of hi2 - lo2 + 1 cells
template <>
struct Walk<true> {
    uint32_t q, qe, cur, ia, ib, nx, ny, nz, row_step;
    __device__ __forceinline__ bool begin(const WalkCtx &c, uint32_t hc, const float4 &p, float reach,
                                          const uint32_t *hs) {
        uint32_t hh[3];
        c.Hs.decode(hc, hh[0], hh[1], hh[2]);
        const float pp[3] = {p.x, p.y, p.z};
        int lo[3], hi[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float u = pp[a] - (float)((int)hh[a] - c.h[a]) * c.cs[a];  // position inside its cell
            // clamped in FP32 first (a huge reach or a NaN position selects the full stencil)
            lo[a] = min(0, (int)fmaxf(floorf((u - reach) * c.inv_cs[a]), -(float)c.radius));
            hi[a] = max(0, (int)fminf(floorf((u + reach) * c.inv_cs[a]), (float)c.radius));
        }
        nx = (uint32_t)(hi[0] - lo[0] + 1);
        ny = (uint32_t)(hi[1] - lo[1] + 1);
        nz = (uint32_t)(hi[2] - lo[2] + 1);
        row_step = c.Hs.HH - (ny - 1u) * c.Hs.H2;
        cur = (uint32_t)((int)hc + lo[0] * (int)c.Hs.HH + lo[1] * (int)c.Hs.H2 + lo[2]);
        ia = ib = 0;
        q = hs[cur];
        qe = hs[cur + nz];
        return q < qe || advance(c, hs);
    }
    __device__ __forceinline__ bool advance(const WalkCtx &c, const uint32_t *hs) {
        do {
            if (++ib == ny) {
                ib = 0;
                if (++ia == nx) return false;
                cur += row_step;
            } else {
                cur += c.Hs.H2;
            }
            q = hs[cur];
            qe = hs[cur + nz];
        } while (q >= qe);
        return true;
    }
};

// general case: per axis halo mode (offsets -r..r), whole-axis mode with all cells (dims <= 2r) or
// whole-axis mode with wrapped offsets; z as one run unless wrapped
template <>
struct Walk<false> {
    int nx, ny, nzr;
    int bx, by, bz;  // base coordinates
    int mx, my, mz;  // axis modes: 0 halo, 1 whole (all cells), 2 whole (wrapped offsets)
    int Dz;
    int ia, ib, iz;
    uint32_t q, qe;

    __device__ __forceinline__ static void axis_mode(bool whole, int D, int radius, int hcoord, int &n, int &base,
                                                     int &mode) {
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
    __device__ __forceinline__ static int elem(int k, int base, int mode, int D) {
        if (mode != 2) return base + k;
        int v = base + k;  // |overhang| <= r < D
        return v < 0 ? v + D : (v >= D ? v - D : v);
    }
    __device__ __forceinline__ void range(const WalkCtx &c, const uint32_t *hs) {
        const int x = elem(ia, bx, mx, c.dims[0]);
        const int y = elem(ib, by, my, c.dims[1]);
        const uint32_t row = ((uint32_t)x * c.Hs.H1 + (uint32_t)y) * c.Hs.H2;
        int z0, z1;
        if (mz == 2) {
            z0 = elem(iz, bz, 2, c.dims[2]);
            z1 = z0 + 1;
        } else {
            z0 = bz;
            z1 = bz + Dz;
        }
        q = hs[row + (uint32_t)z0];
        qe = hs[row + (uint32_t)z1];
    }
    __device__ __forceinline__ bool begin(const WalkCtx &c, uint32_t hc, const float4 &, float,
                                          const uint32_t *hs) {
        uint32_t hx, hy, hz;
        c.Hs.decode(hc, hx, hy, hz);
        axis_mode(c.whole & 1u, c.dims[0], c.radius, (int)hx, nx, bx, mx);
        axis_mode(c.whole & 2u, c.dims[1], c.radius, (int)hy, ny, by, my);
        int nz;
        axis_mode(c.whole & 4u, c.dims[2], c.radius, (int)hz, nz, bz, mz);
        if (mz == 0) {  // one contiguous z-run of 2r+1 cells
            nzr = 1;
            Dz = 2 * c.radius + 1;
        } else if (mz == 1) {  // the whole axis
            nzr = 1;
            Dz = c.dims[2];
        } else {
            nzr = nz;
            Dz = 1;
        }
        ia = ib = iz = 0;
        range(c, hs);
        return q < qe || advance(c, hs);
    }
    __device__ __forceinline__ bool advance(const WalkCtx &c, const uint32_t *hs) {
        do {
            if (++iz >= nzr) {
                iz = 0;
                if (++ib >= ny) {
                    ib = 0;
                    if (++ia >= nx) return false;
                }
            }
            range(c, hs);
        } while (q >= qe);
        return true;
    }
};

// makes LDS writes of this wave visible to its other lanes (and orders the compiler's accesses)
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Streams candidate pairs through the wave's ring: every active lane walks the stencil of its own
// staged record (halo cell hc, local position p, reach), taking up to BATCH candidates of its
// current run per iteration (independent LDS reads), and offers each to keep(q, P[q]); survivors are
// appended as (lane << 16 | q) and evaluated by eval(entry) 64 at a time with all lanes of the
// wave.  A lane's survivors of one iteration take consecutive ring places after those of the lower
// lanes (one prefix over the lanes' survivor counts per iteration).  The last, partial batch too runs
// eval on every lane, idle lanes with NO_ENTRY (so that eval may read other lanes' registers).  Must
// be called by all 64 lanes of a wave (wave-uniform control flow around it).
template <bool ALL_HALO, class Keep, class Eval>
__device__ __forceinline__ void pair_pipeline(const WalkCtx &c, const uint32_t *hs, const float4 *P, uint32_t *ring,
                                              bool active, uint32_t hc, const float4 &p, float reach, Keep &&keep,
                                              Eval &&eval) {
    static_assert(BATCH <= 7, "survivor counts are ranked as three bit planes");
    const uint32_t lane = __lane_id();
    Walk<ALL_HALO> w;
    w.q = w.qe = 0;
    bool more = false;
    if (active) more = w.begin(c, hc, p, reach, hs);
    uint32_t head = 0, tail = 0;  // wave-uniform
    while (__ballot(more)) {
        uint32_t q0 = 0, nq = 0;
        if (more) {
            q0 = w.q;
            nq = min(w.qe - w.q, BATCH);
            w.q += nq;
            if (w.q >= w.qe) more = w.advance(c, hs);
        }
        // candidates q0 .. q0 + nq - 1; reads past the batch repeat its last record (masked out)
        const uint32_t last = nq ? q0 + nq - 1u : q0;
        float4 pq[BATCH];  // the batch's records first (independent LDS reads), then the tests
#pragma unroll
        for (uint32_t i = 0; i < BATCH; ++i) pq[i] = P[min(q0 + i, last)];
        uint32_t bits = 0;
#pragma unroll
        for (uint32_t i = 0; i < BATCH; ++i) bits |= (keep(min(q0 + i, last), pq[i]) & (i < nq)) ? (1u << i) : 0u;
        const uint32_t cnt = (uint32_t)__popc(bits);
        const unsigned long long b0 = __ballot(cnt & 1u), b1 = __ballot(cnt & 2u), b2 = __ballot(cnt & 4u);
        uint32_t pos = tail + rank_in_mask(b0) + 2u * rank_in_mask(b1) + 4u * rank_in_mask(b2);
        tail += (uint32_t)__popcll(b0) + 2u * (uint32_t)__popcll(b1) + 4u * (uint32_t)__popcll(b2);
        while (bits) {  // at most BATCH passes, usually one
            const uint32_t i = (uint32_t)__ffs(bits) - 1u;
            bits &= bits - 1u;
            ring[pos & (RING - 1)] = (lane << 16) | (q0 + i);
            ++pos;
        }
        while (tail - head >= 64) {
            wave_lds_sync();
            eval(ring[(head + lane) & (RING - 1)]);
            head += 64;
            wave_lds_sync();
        }
    }
    if (tail != head) {
        wave_lds_sync();
        eval(lane < tail - head ? ring[(head + lane) & (RING - 1)] : NO_ENTRY);
        wave_lds_sync();
    }
}

// ---- per-block type tables, carved from dynamic LDS ahead of the staged records -----------------
// Staged bricks need n_types <= MAX_STAGED_TYPES (larger type counts take the slow path).
constexpr uint32_t MAX_STAGED_TYPES = 16;

__host__ __device__ inline size_t align16(size_t v) { return (v + 15) & ~(size_t)15; }

// bytes of the force tables for nt types and np potentials
__host__ __device__ inline size_t force_table_bytes(uint32_t nt, uint32_t np) {
    const size_t n2 = (size_t)nt * nt;
    return align16(sizeof(PairIdx) * n2) + align16(sizeof(PotDev) * np) + align16(4 * n2) + align16(4 * nt);
}

// bytes of the table region for nt types and np potentials: the force tables followed by the
// search tables (the fused forces + search kernel holds both; the others use one of them)
__host__ __device__ inline size_t table_bytes(uint32_t nt, uint32_t np) {
    const size_t n2 = (size_t)nt * nt;
    return force_table_bytes(nt, np) + align16(sizeof(RxPairIdx) * n2) + align16(4 * n2) + align16(4 * nt);
}

struct ForceTables {
    PairIdx *pidx;  // [nt * nt]
    PotDev *pots;   // [n_pots]
    float *thr;     // [nt * nt] prefilter thresholds
    float *reach;   // [nt] largest cutoff of the type's row plus rounding slack (< 0: no potentials)
};

struct SearchTables {
    RxPairIdx *o2;  // [nt * nt]
    float *thr;     // [nt * nt]
    float *reach;   // [nt] (< 0: no pair reactions)
};

// reach of a type from the FP32 squared cutoffs of its row, widened like the prefilter
__device__ __forceinline__ float reach_of(float max_cut2_f, bool any, float margin) {
    return any ? sqrtf(max_cut2_f) * (1.0f + 1.0e-6f) + 2.0f * margin : -1.0f;
}

__device__ __forceinline__ ForceTables load_force_tables(char *base, const Tables &T, float margin) {
    const uint32_t nt = T.n_types, n2 = nt * nt;
    ForceTables t;
    char *p = base;
    t.pidx = (PairIdx *)p;
    p += align16(sizeof(PairIdx) * n2);
    t.pots = (PotDev *)p;
    p += align16(sizeof(PotDev) * T.n_pots);
    t.thr = (float *)p;
    p += align16(4 * n2);
    t.reach = (float *)p;
    for (uint32_t k = threadIdx.x; k < n2; k += BRICK_THREADS) {
        const PairIdx pi = T.pidx[k];
        t.pidx[k] = pi;
        t.thr[k] = prefilter_thr(pi.maxcut2_f, pi.count != 0, margin);
    }
    for (uint32_t k = threadIdx.x; k < T.n_pots; k += BRICK_THREADS) t.pots[k] = T.pots[k];
    if (threadIdx.x < nt) {
        float m2 = 0.f;
        bool any = false;
        for (uint32_t u = 0; u < nt; ++u) {
            const PairIdx pi = T.pidx[threadIdx.x * nt + u];
            if (pi.count) {
                any = true;
                m2 = fmaxf(m2, pi.maxcut2_f);
            }
        }
        t.reach[threadIdx.x] = reach_of(m2, any, margin);
    }
    __syncthreads();
    return t;
}

__device__ __forceinline__ SearchTables load_search_tables(char *base, const Tables &T, float margin) {
    const uint32_t nt = T.n_types, n2 = nt * nt;
    SearchTables t;
    char *p = base;
    t.o2 = (RxPairIdx *)p;
    p += align16(sizeof(RxPairIdx) * n2);
    t.thr = (float *)p;
    p += align16(4 * n2);
    t.reach = (float *)p;
    for (uint32_t k = threadIdx.x; k < n2; k += BRICK_THREADS) {
        const RxPairIdx pi = T.o2idx[k];
        t.o2[k] = pi;
        t.thr[k] = prefilter_thr(pi.maxr2_f, pi.count != 0, margin);
    }
    if (threadIdx.x < nt) {
        float m2 = 0.f;
        bool any = false;
        for (uint32_t u = 0; u < nt; ++u) {
            const RxPairIdx pi = T.o2idx[threadIdx.x * nt + u];
            if (pi.count) {
                any = true;
                m2 = fmaxf(m2, pi.maxr2_f);
            }
        }
        t.reach[threadIdx.x] = reach_of(m2, any, margin);
    }
    __syncthreads();
    return t;
}


// FP32 min image of a local-frame displacement on whole-axis axes (halo axes need none)
__device__ __forceinline__ void frame_min_image(const Frame &F, float &dx, float &dy, float &dz) {
    if (F.whole & 1u) dx = dx > .5f * F.L[0] ? dx - F.L[0] : (dx <= -.5f * F.L[0] ? dx + F.L[0] : dx);
    if (F.whole & 2u) dy = dy > .5f * F.L[1] ? dy - F.L[1] : (dy <= -.5f * F.L[1] ? dy + F.L[1] : dy);
    if (F.whole & 4u) dz = dz > .5f * F.L[2] ? dz - F.L[2] : (dz <= -.5f * F.L[2] ? dz + F.L[2] : dz);
}

// the pair sum reading neighbours from the bins (slow path for bricks that do not fit in LDS)
__device__ __forceinline__ void global_forces(const Grid &cur, const RepGeom &geo, uint32_t cap, int radius,
                                              const Tables &T, uint32_t s, const State &me, long long &fx,
                                              long long &fy, long long &fz, long long &e, long long *vir = nullptr) {
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
        long long px, py, pz, pe;
        if (!pair_fixed(T.pots, pi, dx, dy, dz, d2, px, py, pz, pe)) return;
        fx += px;
        fy += py;
        fz += pz;
        e += pe;
        if (vir) pair_virial(T.pots, pi, dx, dy, dz, d2, vir);
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
    __syncthreads();  // s_e may be written again by a following call
}

// block-wide sums of the nine fixed-point virial components, one atomic each per block
__device__ __forceinline__ void block_add_virial(unsigned long long *virial, uint32_t r, const long long v[9]) {
    __shared__ long long s_v[9][BRICK_THREADS / 64];
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        long long x = v[c];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
        if (__lane_id() == 0) s_v[c][threadIdx.x / 64] = x;
    }
    __syncthreads();
    if (threadIdx.x < 9) {
        long long t = 0;
        for (int w = 0; w < BRICK_THREADS / 64; ++w) t += s_v[threadIdx.x][w];
        if (t != 0) atomicAdd(&virial[9ull * r + threadIdx.x], (unsigned long long)t);
    }
    __syncthreads();
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
    if (do_forces && (fixed_saturated(fx) || fixed_saturated(fy) || fixed_saturated(fz)))
        atomicAdd(&err[ERR_FORCE_RANGE], 1u);
    const double Fx = (double)fx * FORCE_INV, Fy = (double)fy * FORCE_INV, Fz = (double)fz * FORCE_INV;
    if (MODE != MODE_REBIN && force_out && do_forces) {
        force_out[3ull * s] = Fx;
        force_out[3ull * s + 1] = Fy;
        force_out[3ull * s + 2] = Fz;
    }
    if (MODE == MODE_FORCES) return;
    State nw = me;
    nw.tw &= ~TW_NEW;
    if (MODE == MODE_STEP && R.active) {
        double xi[3];
        bd_noise(noise, me.h, R.step, R.seed, xi);
        const double Fv[3] = {Fx, Fy, Fz};
        double p[3] = {me.x, me.y, me.z};
        em_move(T, tw_type(me.tw), dt, Fv, xi, geo.box, geo.pbc, p, err);
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

struct ForceWave {  // per-wave LDS of k_brick
    unsigned long long acc[3][64];  // fixed-point force of the wave's 64 particles
    uint32_t s[64];                 // their bin slots
};

// Bricks whose halo does not fit in LDS are listed for a second kernel that reads neighbours
// straight from the bins.  list: [0] count, [1] blocks of the second kernel done, [2..] bricks.
__device__ __forceinline__ void list_slow_brick(uint32_t *list, uint32_t b) {
    if (threadIdx.x == 0) list[2 + atomicAdd(&list[0], 1u)] = b;
}

// count of listed bricks; with none listed the second kernel's blocks exit at once (and nothing
// needs clearing)
__device__ __forceinline__ uint32_t slow_count(const uint32_t *list) {
    return __atomic_load_n(list, __ATOMIC_RELAXED);
}

// the last block of the second kernel clears the list for the next launch (every block has read
// the nonzero count before it counts itself done)
__device__ __forceinline__ void release_slow_list(uint32_t *list) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&list[1], 1u) == gridDim.x - 1) {
            list[0] = 0u;
            list[1] = 0u;
            __threadfence();
        }
    }
}

// forces (+ Euler-Maruyama + re-bin) of every particle of one brick; MODE as in k_fused.
// AH: the launch covers bricks brick0.. that are in halo mode on every axis (else: none is).
template <int MODE, bool AH>
__global__ __launch_bounds__(BRICK_THREADS) void k_brick(Grid cur, Grid nxt, ReplicaDev *reps, Tables T, Topo topo,
                                                         const BrickDev *bricks, uint32_t brick0, uint32_t cap,
                                                         int radius, double dt,
                                                         const double *noise, double *force_out, uint32_t max_states,
                                                         uint32_t max_halo, uint32_t smem_bytes, uint32_t *slow,
                                                         uint32_t *err, unsigned long long *virial) {
    extern __shared__ __align__(16) char smem[];
    __shared__ ForceWave wave_lds[BRICK_WAVES];
    const uint32_t bidx = brick0 + xcd_block(blockIdx.x, gridDim.x);
    const BrickDev B = bricks[bidx];
    const ReplicaDev &R = reps[B.r];
    if (!R.in_use) return;
    if (MODE != MODE_FORCES) virial = nullptr;
    long long vir[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
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
    long long e_blk = 0;
    uint32_t total = 0xFFFFFFFFu, n_inner = 0;
    const uint32_t nt = T.n_types;
    BrickSmem m;
    ForceTables ft;
    const Frame F = make_frame(B, geo);
    const Halo Hs = make_halo(B);
    if (nt <= MAX_STAGED_TYPES) {
        if (table_bytes(nt, T.n_pots) + brick_smem_bytes(max_states, max_halo) > smem_bytes) {  // block-uniform
            if (threadIdx.x == 0) atomicAdd(&err[ERR_LDS_LAYOUT], 1u);
            return;
        }
        ft = load_force_tables(smem, T, F.margin);
        m = carve(smem + table_bytes(nt, T.n_pots), max_states, max_halo);
        total = stage_brick(cur, B, geo, F, cap, m, max_states, Hs, load_halo_counts(cur, B, Hs), &n_inner);
    }
    if (total != 0xFFFFFFFFu) {
        const WalkCtx wc = make_walk(B, geo, Hs, radius);
        if (AH != (wc.whole == 0)) {  // launch and brick disagree (host invariant): refuse the brick
            if (threadIdx.x == 0) atomicAdd(&err[ERR_LDS_LAYOUT], 1u);
            return;
        }
        ForceWave &W = wave_lds[threadIdx.x / 64];
        const uint32_t lane = __lane_id();
        // the exact pair force of a survivor (entry = owner lane << 16 | staged neighbour)
        auto eval = [&](uint32_t ent) {
            if (ent == NO_ENTRY) return;
            const uint32_t ol = ent >> 16, q = ent & 0xFFFFu;
            const State me = cur.bins[W.s[ol]];
            const State o = cur.bins[m.slot[q]];
            const PairIdx pi = ft.pidx[tw_type(me.tw) * nt + tw_type(o.tw)];
            // CPUCalculateForces.cpp:226-259: x_ij = shortestDifference(x_i, x_j), gate d2 < rc^2
            const double ddx = min_image(o.x - me.x, geo.box[0], geo.pbc[0]);
            const double ddy = min_image(o.y - me.y, geo.box[1], geo.pbc[1]);
            const double ddz = min_image(o.z - me.z, geo.box[2], geo.pbc[2]);
            const double d2 = dist2(ddx, ddy, ddz);
            long long px, py, pz, pe;
            if (!pair_fixed(ft.pots, pi, ddx, ddy, ddz, d2, px, py, pz, pe)) return;
            if (fixed_saturated(px) || fixed_saturated(py) || fixed_saturated(pz)) atomicAdd(&err[ERR_FORCE_RANGE], 1u);
            atomicAdd(&W.acc[0][ol], (unsigned long long)px);
            atomicAdd(&W.acc[1][ol], (unsigned long long)py);
            atomicAdd(&W.acc[2][ol], (unsigned long long)pz);
            e_blk += pe;
            if (virial) pair_virial(ft.pots, pi, ddx, ddy, ddz, d2, vir);
        };
        auto rounds = [&]() {
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
                State me{};
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
                const float reach = ft.reach[t];
                const bool active = valid && !(me.tw & TW_OUTSIDE) && reach >= 0.f;
                const float *thr_row = ft.thr + t * nt;
                // prefilter, branch-free: one LDS read of the candidate, one of the threshold
                auto keep = [&](uint32_t q, const float4 &pq) -> bool {
                    const uint32_t tw2 = __float_as_uint(pq.w);
                    float dx = pq.x - pj.x, dy = pq.y - pj.y, dz = pq.z - pj.z;
                    if constexpr (!AH) frame_min_image(F, dx, dy, dz);
                    const float d2p = prefilter_d2(dx, dy, dz);
                    const float thr = thr_row[tw_type(tw2)];
                    return (q != j) & (d2p < thr);  // dead / outside records hold NaN positions
                };
                pair_pipeline<AH>(wc, m.hstart, m.P, m.rings + (threadIdx.x / 64) * RING, active, m.hc_of[j], pj, reach, keep, eval);
                wave_lds_sync();
                if (valid)
                    finish_particle<MODE>(nxt, R, geo, T, topo, cap, dt, noise, force_out, sj, me, true,
                                          (long long)W.acc[0][lane], (long long)W.acc[1][lane],
                                          (long long)W.acc[2][lane], err);
                wave_lds_sync();
            }
        };
        rounds();
        block_add_energy(reps, B.r, e_blk);
        if (virial) block_add_virial(virial, B.r, vir);
    } else {
        list_slow_brick(slow, bidx);
    }
}

// the listed bricks of k_brick, neighbours read straight from the bins
template <int MODE>
__global__ __launch_bounds__(BRICK_THREADS) void k_brick_slow(Grid cur, Grid nxt, ReplicaDev *reps, Tables T,
                                                              Topo topo, const BrickDev *bricks, uint32_t cap,
                                                              int radius, double dt, const double *noise,
                                                              double *force_out, uint32_t *slow, uint32_t *err,
                                                              unsigned long long *virial) {
    const uint32_t n = slow_count(slow);
    if (n == 0) return;  // grid-uniform
    if (MODE != MODE_FORCES) virial = nullptr;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {  // block-uniform
        const BrickDev B = bricks[slow[2 + i]];
        const ReplicaDev &R = reps[B.r];
        const RepGeom geo = load_geom(R);
        long long e_blk = 0;
        long long vir[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for_each_own_record(cur, B, geo, cap, [&](uint32_t s, const State &me) {
            if (me.tw & TW_DEAD) return;
            long long fx = 0, fy = 0, fz = 0, e = 0;
            if (!(me.tw & TW_OUTSIDE))
                global_forces(cur, geo, cap, radius, T, s, me, fx, fy, fz, e, virial ? vir : nullptr);
            e_blk += e;
            finish_particle<MODE>(nxt, R, geo, T, topo, cap, dt, noise, force_out, s, me, true, fx, fy, fz, err);
        });
        block_add_energy(reps, B.r, e_blk);
        if (virial) block_add_virial(virial, B.r, vir);
    }
    release_slow_list(slow);
}

struct SearchWave {  // per-wave LDS of k_brick_search
    uint32_t s[64];  // bin slots of the wave's 64 particles
};

// reaction candidate search of one brick on the current buffer (see search_one for the rules),
// fused with clearing the previous buffer (clear_old): the counters of the brick's cells, and (block
// 0) its work-list and overflow lengths
template <bool AH>
__global__ __launch_bounds__(BRICK_THREADS) void k_brick_search(Grid cur, Grid old, int clear_old, ReplicaDev *reps,
                                                                Tables T, const BrickDev *bricks, uint32_t brick0,
                                                                uint32_t cap,
                                                                int radius, int handler, Event *events,
                                                                uint32_t *ev_count, uint32_t max_states,
                                                                uint32_t max_halo, uint32_t smem_bytes, uint32_t *slow,
                                                                uint32_t *err) {
    extern __shared__ __align__(16) char smem[];
    __shared__ SearchWave wave_lds[BRICK_WAVES];
    const uint32_t bidx = brick0 + xcd_block(blockIdx.x, gridDim.x);
    const BrickDev B = bricks[bidx];
    const ReplicaDev &R = reps[B.r];
    const RepGeom geo = load_geom(R);
    // clear the old buffer (it is not read again before it becomes the next target)
    if (clear_old) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            *old.wl_count = 0u;
            *old.ovf_count = 0u;
        }
        const uint32_t ncell = (uint32_t)(B.e[0] * B.e[1] * B.e[2]);
        for (uint32_t ci = threadIdx.x; ci < ncell; ci += BRICK_THREADS) {
            const uint32_t lz = ci % B.e[2], ly = (ci / B.e[2]) % B.e[1], lx = ci / (B.e[2] * B.e[1]);
            old.counts[geo.cell_base +
                       ((uint32_t)(B.o[0] + lx) * (uint32_t)geo.dims[1] + (uint32_t)(B.o[1] + ly)) *
                           (uint32_t)geo.dims[2] +
                       (uint32_t)(B.o[2] + lz)] = 0u;
        }
    }
    // a launch without reactions (handler 2) covers every replica; DetailedBalance searches every replica
    // with the Gillespie keys; else a replica's own choice wins
    if (handler == HANDLER_DETAILED_BALANCE) handler = 1;
    else if (handler <= 1 && R.handler >= 0) handler = R.handler;
    if (handler > 1 || !R.in_use || !R.active) return;
    const uint32_t nt = T.n_types;
    if (!T.has_o2 && T.o1idx[nt] == 0) return;  // no reactions
    const Frame F = make_frame(B, geo);
    const Halo Hs = make_halo(B);
    uint32_t total = 0xFFFFFFFFu, n_inner = 0;
    BrickSmem m;
    SearchTables st;
    if (nt <= MAX_STAGED_TYPES) {
        if (table_bytes(nt, T.n_pots) + brick_smem_bytes(max_states, max_halo) > smem_bytes) {  // block-uniform
            if (threadIdx.x == 0) atomicAdd(&err[ERR_LDS_LAYOUT], 1u);
            return;
        }
        st = load_search_tables(smem, T, F.margin);
        m = carve(smem + table_bytes(nt, T.n_pots), max_states, max_halo);
        total = stage_brick(cur, B, geo, F, cap, m, max_states, Hs, load_halo_counts(cur, B, Hs), &n_inner);
    }
    if (total == 0xFFFFFFFFu) {
        list_slow_brick(slow, bidx);
        return;
    }
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
    const WalkCtx wc = make_walk(B, geo, Hs, radius);
    if (AH != (wc.whole == 0)) {  // launch and brick disagree (host invariant): refuse the brick
        if (threadIdx.x == 0) atomicAdd(&err[ERR_LDS_LAYOUT], 1u);
        return;
    }
    SearchWave &W = wave_lds[threadIdx.x / 64];
    const uint32_t lane = __lane_id();
    // the exact test and draws of a surviving pair (entry = owner lane << 16 | staged neighbour)
    auto eval = [&](uint32_t ent) {
        if (ent == NO_ENTRY) return;
        const uint32_t ol = ent >> 16, q = ent & 0xFFFFu;
        const uint32_t s1 = W.s[ol], s2 = m.slot[q];
        const State me = cur.bins[s1];
        const State o = cur.bins[s2];
        const uint32_t mh = me.h, t = tw_type(me.tw);
        const RxPairIdx pi = st.o2[t * nt + tw_type(o.tw)];
        // distSquared(neighbor.pos, entry.pos) (CPUUncontrolledApproximation.cpp:108)
        const double ddx = min_image(me.x - o.x, geo.box[0], geo.pbc[0]);
        const double ddy = min_image(me.y - o.y, geo.box[1], geo.pbc[1]);
        const double ddz = min_image(me.z - o.z, geo.box[2], geo.pbc[2]);
        const double d2 = dist2(ddx, ddy, ddz);
        const uint32_t lo = mh < o.h ? mh : o.h, hi = mh < o.h ? o.h : mh;
        for (uint32_t p = 0; p < pi.count; ++p) {
            const uint32_t rx = T.o2list[pi.start + p];
            const RxDev X = T.rx[rx];
            if (!(X.rate > 0 && d2 < X.educt_dist2)) continue;
            uint32_t x4[4];
            philox(lo, step, 2u | (rx << 8), hi, seed, x4);
            if (!(u53(x4[0], x4[1]) < X.p_accept)) continue;
            // entry1: the educt of type educts[0]; equal educt types: the lower handle
            const bool me_first = (X.educts[0] != X.educts[1]) ? (t == X.educts[0]) : (mh < o.h);
            if (me_first) emit(rx, s1, s2, mh, o.h, event_key(handler, x4, X.rate));
            else emit(rx, s2, s1, o.h, mh, event_key(handler, x4, X.rate));
        }
    };
    auto rounds = [&]() {
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
            State me{};
            uint32_t sj = 0;
            if (valid) {
                sj = m.slot[j];
                me = cur.bins[sj];
            }
            const uint32_t t = tw_type(me.tw);
            if (valid) {  // order 1 (CPUUncontrolledApproximation.cpp:81-95)
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
            if (!T.has_o2) continue;  // block-uniform
            W.s[lane] = sj;
            wave_lds_sync();
            const float reach = st.reach[t];
            const bool active = valid && !(me.tw & TW_OUTSIDE) && reach >= 0.f;
            const float *thr_row = st.thr + t * nt;
            // each unordered pair once: the lower type finds it; equal types: the lower bin slot
            // (which also excludes the particle itself); branch-free
            auto keep = [&](uint32_t q, const float4 &pq) -> bool {
                const uint32_t tw2 = __float_as_uint(pq.w);
                const uint32_t t2 = tw_type(tw2);
                float dx = pq.x - pj.x, dy = pq.y - pj.y, dz = pq.z - pj.z;
                if constexpr (!AH) frame_min_image(F, dx, dy, dz);
                const float d2p = prefilter_d2(dx, dy, dz);
                const float thr = thr_row[t2];
                const uint32_t sq = m.slot[q];
                return ((tw2 & (TW_DEAD | TW_OUTSIDE)) == 0u) & ((t2 > t) | ((t2 == t) & (sq > sj))) & (d2p < thr);
            };
            pair_pipeline<AH>(wc, m.hstart, m.P, m.rings + (threadIdx.x / 64) * RING, active, m.hc_of[j], pj, reach, keep, eval);
            wave_lds_sync();
        }
    };
    rounds();
}

// the listed bricks of k_brick_search, searched straight from the bins
__global__ __launch_bounds__(BRICK_THREADS) void k_brick_search_slow(Grid cur, ReplicaDev *reps, Tables T,
                                                                     const BrickDev *bricks, uint32_t cap, int radius,
                                                                     int handler, Event *events, uint32_t *ev_count,
                                                                     uint32_t *slow, uint32_t *err) {
    const uint32_t n = slow_count(slow);
    if (n == 0) return;  // grid-uniform
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const BrickDev B = bricks[slow[2 + i]];
        const RepGeom geo = load_geom(reps[B.r]);
        for_each_own_record(cur, B, geo, cap, [&](uint32_t s, const State &me) {
            if (!(me.tw & TW_DEAD))
                search_one(cur, reps, T, cap, radius, handler, 0, events, ev_count, nullptr, 0, s, me, err);
        });
    }
    release_slow_list(slow);
}

}  // namespace rdy
