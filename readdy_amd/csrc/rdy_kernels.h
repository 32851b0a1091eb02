// Kernels of the MI355X ReaDDy engine.  Shared by the engine (rdy_engine.hip); declared here so
// the host code can launch them by name.
//
// Data layout in HBM (one engine):
//   bins[2][n_cells * cap]  State (32 B): the particle records themselves, grouped by cell.  A
//                           particle's slot is cell * cap + rank; the record moves to its new
//                           cell every step (double buffered), so no separate sort pass exists.
//   counts[2][n_cells]      particles per cell (atomic insertion counters).  A cell holding more
//                           than cap particles spills into an overflow region after the cell
//                           blocks (slot >= nbin); its records carry their cell in ovf_cell.
//   wl[2][max_particles]    the work list: slots of live particles of the buffer, appended in
//                           wave-aggregated blocks (keeps neighbouring particles in neighbouring
//                           threads across steps).
//   events[...]             per-replica segments of accepted reaction candidates.
#pragma once

#include "rdy_device.h"

namespace rdy {

struct Tables {
    const PotDev *pots;
    const PairIdx *pidx;     // [n_types * n_types]
    const RxDev *rx;         // [n_rx]
    const uint32_t *o1idx;   // [n_types + 1] CSR into o1list
    const uint32_t *o1list;
    const RxPairIdx *o2idx;  // [n_types * n_types]
    const uint32_t *o2list;
    const double *diffusion; // [n_types]
    const double *bd_scale;  // [n_types] sqrt(2 D dt) of the tables' time step
    const float2 *rowmax;    // [n_types] largest FP32 squared cutoff of the type's potentials (x) and of
                             // its order-2 reactions (y); -1 where it has none
    uint32_t n_types;
    uint32_t n_rx;
    uint32_t has_o2;         // any order-2 reaction
    uint32_t n_pots;         // entries of pots
    double kbt;
};

struct Grid {  // one bin buffer
    State *bins;        // [nbin + ovf_cap]: cell blocks, then the overflow region
    uint32_t *counts;
    uint32_t *wl;
    uint32_t *wl_count;
    uint32_t *ovf_cell;   // [ovf_cap] cell of each overflow record
    uint32_t *ovf_count;
    uint32_t nbin;        // n_cells * cap: first overflow slot
    uint32_t ovf_cap;
};

struct Topo {
    long long *fb;      // [3 * n_topo_total] bonded forces, fixed point 2^-32
    uint32_t *loc;      // [n_topo_total] current slot of each topology particle
};

struct Event {
    unsigned long long key;
    uint32_t s1, s2;  // slots of entry1 / entry2 (equal for order-1)
    uint32_t h1, h2;
    uint32_t rx;
    uint32_t pad;
};

// a record changed by a reaction of the fused step: slot in the current buffer, kind (bits 0-1:
// killed 1, inserted 2, retyped 3) | type before the change << 2, replica, handle before the change
// (retyped records)
struct Edit {
    uint32_t slot, info, r, pad;
};

// one performed reaction as ReactionUtils.h:155-225 fills a ReactionRecord: the educts' handles and
// types before the reaction, the products' handles, where = (pos1 + pos2) / 2 wrapped into the box
// (the plain mean, no minimum image).  Slot k of a replica's region holds sorted event k; valid is 0
// for events that did not fire.
struct RxRecord {
    uint32_t rx, kind;
    uint32_t educts[2], products[2], types_from[2];
    double where[3];
    uint32_t valid, pad;
};
static_assert(sizeof(RxRecord) == 64, "RxRecord layout");

// Tables of the DetailedBalance handler.  The pair potentials again, with their FP64 parameters (the
// rdy_pair_potential parameters and the squared cutoff), listed per ordered type pair as for the
// force tables; per reaction the kind of its reversible pair (-1: none), its Metropolis prefactor and,
// for a FusionFission pair, the first entry of the pair's fission-distance table (radii and
// cumulative probabilities, n_table entries).
struct DbPot {
    double cut2;
    double q[8];
    int32_t kind, pad;
};

struct DbTables {
    const DbPot *pots;
    const uint32_t *pidx;   // [n_types * n_types]: start | count << 16
    const int32_t *kind;    // [n_rx]
    const double *prefactor;
    const uint32_t *table;
    const double *radii, *cdf;
    uint32_t n_table;
    int32_t radius;         // cell stencil radius of the grid
};

// error flags, all counters
enum : int { ERR_BIN_OVERFLOW = 0, ERR_EVENT_OVERFLOW = 1, ERR_WL_OVERFLOW = 2, ERR_NONFINITE = 3, ERR_EVENT_LDS = 4,
             ERR_LDS_SEARCH = 5, ERR_LDS_LAYOUT = 6, ERR_EDIT_OVERFLOW = 7, ERR_FORCE_RANGE = 8, ERR_TOPO_RESERVE = 9,
             N_ERR = 10 };

constexpr int MODE_FORCES = 0;  // forces + energy of the current buffer
constexpr int MODE_STEP = 1;    // forces, Euler-Maruyama step, re-bin into the other buffer
constexpr int MODE_REBIN = 2;   // re-bin only (neighbour-list update)

constexpr int HANDLER_DETAILED_BALANCE = 3;  // RDY_HANDLER_DETAILED_BALANCE: Gillespie search, sequential resolve

constexpr int RESOLVE_THREADS = 256;
constexpr int RESOLVE_MAX_EVENTS = 1024;  // accepted events per replica per step (LDS resident)

}  // namespace rdy
