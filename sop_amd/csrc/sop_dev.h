// sop_dev.h -- device-resident state shared by the HIP kernels and the host
// engine.  One SysDev per resident system (replica); per-bead and per-entry
// arrays live in pooled arenas addressed through DevPtrs.
#pragma once

#include <stdint.h>
#include <hip/hip_runtime_api.h>

namespace sopk {

// Beads per tile.  Every system's bead range starts on a tile boundary, so a
// tile never spans two systems and tile t covers arena beads [64t, 64t+64).
constexpr int TILE = 64;

// SysDev::err bits (sticky until the host clears them)
constexpr int ERR_NL_OVF = 1;    // a bead's neighbor list exceeded nl_cap
constexpr int ERR_PL_OVF = 2;    // a bead's pair list exceeded pl_cap
constexpr int ERR_PBC = 4;       // extent too close to boxl/2 for the open cell grid
constexpr int ERR_NAN = 8;       // non-finite coordinate after a step
constexpr int ERR_BOND = 16;     // FENE bond stretched to |r - r0| >= R0
constexpr int ERR_INTERNAL = 32; // cell counts inconsistent

// kernel modes of the fused force kernel
constexpr int MODE_STEP = 0;        // forces + random force + move (overdamped or underdamped)
constexpr int MODE_FORCE = 1;       // deterministic forces -> fout
constexpr int MODE_FORCE_NOISE = 2; // forces + random force -> fout

struct SysDev {
  // arena placement
  int32_t bead_off;  // arena index of bead 0 (multiple of TILE)
  int32_t n;         // beads
  int32_t nl_cap;    // per-bead capacity of the non-native neighbor list
  int32_t pl_cap;    // per-bead capacity of the non-native pair list
  int32_t idx16;     // 1: pair-list rows hold 16-bit local indices (n <= 65536)
  int64_t nl_base;   // first entry of this system's neighbor-list rows in the scratch (materialized on request)
  int64_t pl_base;   // first entry of this system's pair-list segment
  int32_t cell_base; // first cell of this system's cell segment (cell_cap+1 slots)
  int32_t cell_cap;
  int32_t dmax;      // max cells per axis, dmax^3 <= cell_cap
  int32_t parity;    // which coordinate buffer holds the current coordinates
  // schedule
  int64_t step;      // completed steps
  int64_t step_end;  // the step kernel moves the system only while step < step_end
  int32_t n_nl, n_pl;
  // dynamics
  int32_t dyn;       // 0 overdamped, 1 underdamped
  int32_t vupd;      // underdamped: velocities are updated (T >= 1e-5)
  float a5, var;     // h/zeta and sqrt(2 T zeta / h)
  float a1, a2, a3, a4; // underdamped coefficients
  float boxl, inv_boxl;
  uint32_t seed_lo, seed_hi;
  // force field (type-pair tables indexed (ti-1)*2 + (tj-1))
  float k_bnd, inv_R0sq, half_k_R0sq;
  float ss_e6, ss_f6; // eps_ss*sigma_ss^6 and 6*eps_ss*sigma_ss^6
  float eps_att[4];
  float eps_rep[4], sig2_rep[4], rep_guard[4], rep_m12[4], rep_m6[4];
  float rep_a[4], rep_b[4]; // rep_m12 * sigma^12 and rep_m6 * sigma^6 (step kernel's form)
  float c12, c6;
  float rc2_nl_rep[4], rc2_pl_rep[4];
  double nl_factor, pl_factor;
  float nl_abs, pl_abs;
  float cell_min_edge;    // >= largest neighbor-list cutoff, with a 0.1% margin
  float skin;             // smallest (rc_nl - rc_pl) over pair kinds, less a 1e-5 margin
  // pair-list shell: shell rows hold the master pairs within rc_pl + shell of
  // pos_shell; while every bead has moved less than shell_ok / 2 since then,
  // the pair list is the shell rows filtered at x (BUILD_SHELL); shell_ok is
  // the shell width less a 1e-4 margin
  float shell_ok;
  float rc2_sl_rep[4];    // shell-row cutoffs squared, (rc_pl + shell)^2
  float cell_min_edge_sl; // >= largest shell-row cutoff, with a 0.1% margin
  int32_t sl_valid;       // 1: the shell rows match pos_shell
  int32_t nbuild[3];      // pair-list rebuilds: grid (BUILD_DIRECT), shell, neighbor list
  float wrap_margin;
  // cell grid, written by k_bbox at each neighbor-list rebuild
  float org[3];
  float inv_edge[3];
  int32_t dims[3];
  int32_t ncells;
  int32_t ncells_used; // cells holding at least one bead
  int32_t sort_lds;    // this rebuild's cell sort runs in one block's LDS
  int32_t wrap_free; // 1: every pair separation is below boxl/2 (minimum image is the identity)
  // list state, device-written.  The neighbor list is defined by the snapshot
  // pos_ref taken at its last rebuild; it is materialized only when needed.
  int32_t build;     // this rebuild point: BUILD_NONE / BUILD_DIRECT / BUILD_NL / BUILD_NL_ONLY
  int32_t nl_valid;  // 1: the materialized neighbor list matches pos_ref
  int32_t type_mask; // bit t-1 set when some bead has type t
  // status, device-written
  int32_t err;
  int32_t row_max;   // longest shell row (or neighbor-list build row) since the host last read it
};

// SysDev::build
constexpr int BUILD_NONE = 0;    // no list work at this step
constexpr int BUILD_DIRECT = 1;  // pair list straight from a pair-cutoff cell grid
constexpr int BUILD_NL = 2;      // pair list = neighbor list of pos_ref filtered at x, in one pass
constexpr int BUILD_NL_ONLY = 3; // materialize the neighbor list only
constexpr int BUILD_SHELL = 4;   // pair list = shell rows filtered at x (no cell grid)

// k_skin flags
constexpr int FORCE_NL = 1;      // rebuild the neighbor list now (snapshot)
constexpr int FORCE_PL = 2;      // rebuild the pair list now
constexpr int MATERIALIZE_NL = 4;// materialize the neighbor list of the current snapshot

struct DevPtrs {
  SysDev *sys;
  float4 *pos0, *pos1;  // double-buffered unwrapped coordinates; w = bits (chain<<2 | type)
  float4 *pos_ref;      // coordinates at the last neighbor-list rebuild
  float4 *pos_sorted;   // coordinates in cell order (copy made by the cell sort)
  float *bond_r0;       // r0 of bond (i, i+1)
  int32_t *nat_start, *nat_cnt;   // native CSR per bead (two-sided, sorted by partner)
  int32_t *nat_j;       // partner (arena index)
  float *nat_r0;
  int32_t *nlatt_j; float *nlatt_r0; int32_t *nlatt_cnt; // native neighbor list (same slots)
  int2 *platt; int32_t *platt_cnt;  // native pair list in the native slots: {partner, bits of r0}
  int32_t *platt_off;   // per bead: first entry of its native pair list in platt.  A tile's lists lie
                        // one after another from the tile's first native slot, in the tile's bead
                        // order (neighbor-list builds: each in the bead's own native slots)
  int32_t *nlrep_j; int32_t *nlrep_cnt;                  // non-native neighbor list (nlrep_j: scratch of the
                                                         // last system materialized)
  int32_t *plrep_j; int32_t *plrep_cnt;                  // non-native pair list, entries int32 arena indices
                                                         // or uint16 local indices (idx16), rows packed per
                                                         // tile (below)
  int32_t *plrep_off;   // per bead: first entry of its pair-list row, counted from the system's segment
                        // start (pl_base).  Tile k of a system owns entries [64 k pl_cap, 64 (k + 1) pl_cap)
                        // of the segment; its rows lie one after another from the region's start in the
                        // tile's bead order (tile_ord), each padded to a multiple of ROW_PAD entries.
  uint2 *plhead;        // per bead, what the step kernel reads of the two lists, written by the shell
                        // filter: {native pair-list length | pair-list length << 16, (plrep_off - the
                        // tile's region start) / ROW_PAD | (platt_off - nat_start of the tile's first
                        // bead) << 16}; all four fit 16 bits (pl_cap <= 4095 and natives per tile
                        // <= 65535, both checked on the host)
  int32_t *slrep_j; int32_t *slrep_cnt;                  // shell rows: row i at (i - bead_off) * pl_cap of the
                                                         // segment, same entry format
  float4 *pos_shell;                                     // coordinates at the last shell-row build
  int32_t *cell_cnt, *cell_start, *cell_cursor;          // per cell
  int32_t *cell_of, *cell_tmp, *cell_sorted;             // per bead
  int32_t *cell_pos;    // per bead: its position in the system's cell-sorted order
  int2 *nat_sorted;     // per sorted position: {nat_start, nat_cnt} of the bead there
  float4 *vel;          // underdamped state per bead: w = 1: xyz = last increment;
                        // w = 0: xyz = velocity entering the next step
  float4 *fout;         // forces (force-eval modes)
  double *epart;        // per bead: 6 energy partials
  double *sysout;       // per system slot: SYSOUT_N reduced values
  int32_t *tile_sys;    // system slot of each tile
  uint8_t *tile_ord;    // per tile: TILE bead offsets in the order the step kernel's lane groups take them
  int32_t *work;        // tile ids to process (-1 = padding), XCD-interleaved
  int32_t *nwork;
  int32_t *syslist;     // system slots of the current batch
  int32_t *nsys;
  int32_t *cellwork;    // per batch position: first home cell of the direct pair-list build
  int32_t *ncellwork;   // total home cells of the direct pair-list build
  int32_t *nlbuild;     // 1 if a system of the batch materializes its neighbor list at this rebuild
  int32_t *gsort;       // nonzero if a grid build of this rebuild sorts in global memory (too big for LDS)
  int2 *cellitems;      // home cells of the direct build: {batch position, cell}
  int4 *items;          // step-kernel work: {first tile, tiles | waves per tile << 8, first staged
                        // bead or -1, staged beads}; the waves of a tile split its rounds
  int32_t *nitems;       // items that stage their systems in LDS: items[0 .. *nitems)
  int32_t *nitems_big;   // items staged in a whole-CU LDS block: items[*nitems .. + *nitems_big)
  int32_t *nitems_far;   // items that gather from global memory: the *nitems_far after those
  int4 *fitems;         // shell-filter work, same form with one wave per tile and STEP_WAVES
                        // tiles per item: fitems[0 .. *nfitems) LDS, then whole-CU, then far
  int32_t *nfitems, *nfitems_big, *nfitems_far;
  // epart and sysout are indexed from these (0 for the engine's arrays; the
  // first bead and the slot of the one system when they point into scratch)
  int32_t epart_base, sysout_base;
};

// step kernel: blocks of STEP_WAVES waves, one 64-bead tile per wave; an item
// is {first tile, tiles <= STEP_WAVES, first staged bead or -1, staged beads}
constexpr int STEP_WAVES = 16;
constexpr int STEP_THREADS = 64 * STEP_WAVES;

// Grids of the step kernel's three launches: items staged in LDS at two
// blocks per compute unit, items staged in a whole compute unit's LDS, items
// that gather from global memory.
struct StepGrids {
  int grid, grid_big, grid_far;
  size_t lds, lds_big;
};

// A second stream and two events: the step launches of the whole-CU and
// global-gather tiers run beside the two-blocks-per-unit LDS launch (a fork
// and a join in the captured graph) instead of after it.
struct StepFork {
  hipStream_t side;
  hipEvent_t fork, join;
};

// Grids of the list kernels: per tile, per system, and the shell filter's
// work items (staged in LDS like the step kernel's).
struct ListGrids {
  int grid_tiles, grid_sys;
  StepGrids fg;
};

// step-kernel item: tiles in the low byte of .y, waves per tile above it
__host__ __device__ inline int item_tiles(int y) { return y & 0xFF; }
__host__ __device__ inline int item_split(int y) { return y >> 8; }

// systems whose cell sort fits one block's LDS
constexpr int SORT_N_MAX = 5120;
constexpr int SORT_C_MAX = 4096;

constexpr int EPART_N = 6;   // bond, angle, native, rep, virial, contacts
constexpr int SYSOUT_N = 16; // e[0..5], contacts, rgsq, sum x,y,z, n

// Scratch of the allocation-free energy entry point: one system of up to
// NOALLOC_BEADS beads.  Layout in bytes: work table (one entry per tile),
// counters (the engine's nwork layout: [0] work entries, [16] systems),
// system list, per-bead energy partials, the system's reduced values.
constexpr int NOALLOC_BEADS = 65536;
constexpr size_t NOALLOC_OFF_WORK = 0;
constexpr size_t NOALLOC_OFF_CNT = NOALLOC_OFF_WORK + 4 * (NOALLOC_BEADS / TILE) + 256;
constexpr size_t NOALLOC_OFF_SYSLIST = NOALLOC_OFF_CNT + 256;
constexpr size_t NOALLOC_OFF_EPART = NOALLOC_OFF_SYSLIST + 256;
constexpr size_t NOALLOC_OFF_SYSOUT = NOALLOC_OFF_EPART + 8 * EPART_N * (size_t)NOALLOC_BEADS;
constexpr size_t NOALLOC_BYTES = NOALLOC_OFF_SYSOUT + 8 * SYSOUT_N;

} // namespace sopk
