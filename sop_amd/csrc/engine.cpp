// engine.cpp -- libsop host engine: device pools, systems, batch work tables,
// hipGraph step blocks and the C ABI of include/sop.h (device part).
#include "engine.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <sstream>
#include <unordered_map>
#include <utility>

using namespace sopk;

namespace sopk {

// ---------------------------------------------------------------- allocator

void RangeAlloc::init(int64_t cap) {
  cap_ = cap;
  used_ = 0;
  free_.clear();
  if (cap > 0) free_[0] = cap;
}

int64_t RangeAlloc::alloc(int64_t n, int64_t align) {
  if (n <= 0) n = 1;
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    const int64_t off = it->first, len = it->second;
    const int64_t a = (off + align - 1) / align * align;
    if (a + n > off + len) continue;
    free_.erase(it);
    if (a > off) free_[off] = a - off;
    if (a + n < off + len) free_[a + n] = off + len - (a + n);
    used_ += n;
    return a;
  }
  return -1;
}

void RangeAlloc::release(int64_t off, int64_t n) {
  if (n <= 0) n = 1;
  used_ -= n;
  auto it = free_.emplace(off, n).first;
  auto nx = std::next(it);
  if (nx != free_.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    free_.erase(nx);
  }
  if (it != free_.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      free_.erase(it);
    }
  }
}

void RangeAlloc::grow(int64_t new_cap) {
  if (new_cap <= cap_) return;
  release(cap_, new_cap - cap_);
  used_ += new_cap - cap_;  // release() subtracted it
  cap_ = new_cap;
}

sop_status set_error(sop_engine *e, sop_status st, const std::string &msg) {
  if (e) e->err = msg;
  return st;
}

sop_status hip_check(sop_engine *e, hipError_t h, const char *what) {
  if (h == hipSuccess) return SOP_OK;
  return set_error(e, SOP_ERR_HIP, std::string(what) + ": " + hipGetErrorString(h));
}

int grid_tiles(const sop_engine *e) {
  int g = e->opts.grid_blocks > 0 ? e->opts.grid_blocks : 8 * e->num_cu;
  g = std::max(8, g / 8 * 8);
  if (e->fixed_grids) return g;
  const int need = (e->cur_nwork + 7) / 8 * 8;
  return std::max(8, std::min(g, need));
}

int grid_sys(const sop_engine *e) {
  if (e->fixed_grids) return 1024;
  return std::max(1, std::min(e->cur_nsys, 2048));
}

// Launch grid of a persistent item kernel: `g` blocks (a multiple of 8, so
// that item k runs on XCD k mod 8), fewer when the batch has fewer items, none
// when it has none; fixed grids (serving) always launch g.
static int items_grid(int g, bool fixed, int count) {
  g = std::max(8, g / 8 * 8);
  if (fixed) return g;
  if (count == 0) return 0;  // no launch
  return std::max(8, std::min(g, (count + 7) / 8 * 8));
}

// The step kernel's LDS launch: `step_blocks`, by default four blocks per
// compute unit for a batch (about one block per item of a large batch: two
// are resident per unit and the dispatcher starts each on whichever unit
// frees first; A/B on one box, 512 replicas: 6.61e9 -> 6.90e9 bead-steps/s
// against two per unit) and two while serving (fixed grids).
int default_step_blocks(const sop_engine *e) { return (e->fixed_grids ? 2 : 4) * e->num_cu; }

int grid_items(const sop_engine *e) {
  return items_grid(e->step_blocks > 0 ? e->step_blocks : default_step_blocks(e), e->fixed_grids, e->cur_nitems);
}

int grid_items_far(const sop_engine *e) { return items_grid(2 * e->num_cu, e->fixed_grids, e->cur_nitems_far); }

// Whole-CU LDS items: one block per compute unit (the block takes all of
// the unit's LDS).  While serving, a fixed grid whenever the batch has such
// items and no launch when it has none (two graph variants; a light load of
// small jobs then pays for no empty launch).
static int items_grid_big(const sop_engine *e, int count) {
  if (e->fixed_grids) return count > 0 ? items_grid(e->num_cu, true, count) : 0;
  return items_grid(e->num_cu, false, count);
}
int grid_items_big(const sop_engine *e) { return items_grid_big(e, e->cur_nitems_big); }

StepGrids step_grids(const sop_engine *e) {
  return StepGrids{grid_items(e), grid_items_big(e), grid_items_far(e), e->cur_lds, e->cur_lds_big};
}

// The shell filter's LDS launch: one block per item (a block leaves its
// compute unit after each item, so the cell-grid chain on the other stream
// finds units between the items); fixed grids (serving) as the step kernel's.
StepGrids filter_grids(const sop_engine *e) {
  const int g = e->step_blocks > 0 ? e->step_blocks : default_step_blocks(e);
  return StepGrids{items_grid(e->fixed_grids ? g : std::max(g, 8 * e->num_cu), e->fixed_grids, e->cur_nfitems),
                   items_grid_big(e, e->cur_nfitems_big),
                   items_grid(2 * e->num_cu, e->fixed_grids, e->cur_nfitems_far), e->cur_flds, e->cur_flds_big};
}

ListGrids list_grids(const sop_engine *e) {
  return ListGrids{grid_tiles(e), grid_sys(e), filter_grids(e)};
}

} // namespace sopk

#define HIPCK(expr)                                            \
  do {                                                         \
    sop_status _s = hip_check(e, (expr), #expr);               \
    if (_s != SOP_OK) return _s;                               \
  } while (0)

// ----------------------------------------------------------------- pools

static void bind_arrays(sop_engine *e) {
  DevPtrs &p = e->p;
  e->bead_arrays = {
      {(void **)&p.pos0, 16},      {(void **)&p.pos1, 16},       {(void **)&p.bond_r0, 4},
      {(void **)&p.pos_ref, 16},   {(void **)&p.pos_sorted, 16},
      {(void **)&p.nat_start, 4},  {(void **)&p.nat_cnt, 4},     {(void **)&p.nlatt_cnt, 4},
      {(void **)&p.platt_cnt, 4},  {(void **)&p.nlrep_cnt, 4},   {(void **)&p.plrep_cnt, 4},
      {(void **)&p.cell_of, 4},    {(void **)&p.cell_tmp, 4},    {(void **)&p.cell_sorted, 4},
      {(void **)&p.cell_pos, 4},   {(void **)&p.nat_sorted, 8},
      {(void **)&p.fout, 16},      {(void **)&p.epart, 8 * EPART_N}, {(void **)&p.vel, 16},
      {(void **)&p.tile_ord, 1},   {(void **)&p.slrep_cnt, 4},   {(void **)&p.pos_shell, 16},
      {(void **)&p.plrep_off, 4},  {(void **)&p.platt_off, 4},  {(void **)&p.plhead, 8}};
  e->nat_arrays = {{(void **)&p.nat_j, 4},    {(void **)&p.nat_r0, 4},  {(void **)&p.nlatt_j, 4},
                   {(void **)&p.nlatt_r0, 4}, {(void **)&p.platt, 8}};
  e->cell_arrays = {{(void **)&p.cell_cnt, 4}, {(void **)&p.cell_start, 4}, {(void **)&p.cell_cursor, 4},
                    {(void **)&p.cellitems, 8}};
}

static sop_status grow_array(sop_engine *e, void **ptr, size_t old_bytes, size_t new_bytes) {
  void *np = nullptr;
  HIPCK(hipMalloc(&np, std::max<size_t>(new_bytes, 256)));
  if (*ptr && old_bytes) HIPCK(hipMemcpy(np, *ptr, old_bytes, hipMemcpyDeviceToDevice));
  if (*ptr) HIPCK(hipFree(*ptr));
  *ptr = np;
  return SOP_OK;
}

static void invalidate_graphs(sop_engine *e) {
  for (auto &kv : e->graphs) (void)hipGraphExecDestroy(kv.second);
  e->graphs.clear();
}

// Grow pools so that the requested extra capacity fits (doubling).
static sop_status ensure_capacity(sop_engine *e, int64_t beads, int64_t nat, int64_t nl, int64_t pl,
                                  int64_t cells, int64_t slots) {
  Pools &c = e->cap;
  bool changed = false;
  auto next = [](int64_t cur, int64_t need) {
    int64_t v = std::max<int64_t>(cur, 1);
    while (v < need) v *= 2;
    return v;
  };
  HIPCK(hipStreamSynchronize(e->stream));
  if (beads > c.beads) {
    const int64_t nb = next(c.beads, beads);
    for (auto &a : e->bead_arrays) {
      sop_status s = grow_array(e, a.first, a.second * (size_t)c.beads, a.second * (size_t)nb);
      if (s) return s;
    }
    sop_status s = grow_array(e, (void **)&e->p.tile_sys, 4 * (size_t)(c.beads / TILE), 4 * (size_t)(nb / TILE));
    if (s) return s;
    e->a_beads.grow(nb);
    c.beads = nb;
    changed = true;
  }
  if (nat > c.nat) {
    const int64_t nn = next(c.nat, nat);
    for (auto &a : e->nat_arrays) {
      sop_status s = grow_array(e, a.first, a.second * (size_t)c.nat, a.second * (size_t)nn);
      if (s) return s;
    }
    e->a_nat.grow(nn);
    c.nat = nn;
    changed = true;
  }
  if (nl > c.nl) {
    const int64_t nn = next(c.nl, nl);
    sop_status s = grow_array(e, (void **)&e->p.nlrep_j, 4 * (size_t)c.nl, 4 * (size_t)nn);
    if (s) return s;
    c.nl = nn;
    changed = true;
  }
  if (pl > c.pl) {
    const int64_t nn = next(c.pl, pl);
    sop_status s = grow_array(e, (void **)&e->p.plrep_j, 4 * (size_t)c.pl, 4 * (size_t)nn);
    if (s) return s;
    s = grow_array(e, (void **)&e->p.slrep_j, 4 * (size_t)c.pl, 4 * (size_t)nn);
    if (s) return s;
    e->a_pl.grow(nn);
    c.pl = nn;
    changed = true;
  }
  if (cells > c.cells) {
    const int64_t nn = next(c.cells, cells);
    for (auto &a : e->cell_arrays) {
      sop_status s = grow_array(e, a.first, a.second * (size_t)c.cells, a.second * (size_t)nn);
      if (s) return s;
    }
    e->a_cells.grow(nn);
    c.cells = nn;
    changed = true;
  }
  if (slots > c.slots) {
    const int64_t nn = next(c.slots, slots);
    sop_status s = grow_array(e, (void **)&e->p.sys, sizeof(SysDev) * (size_t)c.slots, sizeof(SysDev) * (size_t)nn);
    if (s) return s;
    s = grow_array(e, (void **)&e->p.sysout, 8 * SYSOUT_N * (size_t)c.slots, 8 * SYSOUT_N * (size_t)nn);
    if (s) return s;
    s = grow_array(e, (void **)&e->p.syslist, 4 * (size_t)c.slots, 4 * (size_t)nn);
    if (s) return s;
    s = grow_array(e, (void **)&e->p.cellwork, 4 * (size_t)(c.slots + 1), 4 * (size_t)(nn + 1));
    if (s) return s;
    c.slots = nn;
    e->sys.resize((size_t)nn);
    changed = true;
  }
  if (changed) invalidate_graphs(e);
  return SOP_OK;
}

// Pool growth on demand: try to allocate, grow and retry once.
static int64_t alloc_or_grow(sop_engine *e, RangeAlloc &a, int64_t n, int64_t align, int which,
                             sop_status *st) {
  *st = SOP_OK;
  int64_t off = a.alloc(n, align);
  if (off >= 0) return off;
  if (e->fixed_pools) {
    *st = set_error(e, SOP_ERR_CAPACITY, "memory pool full");
    return -1;
  }
  const int64_t need = a.cap() + n + align;
  int64_t req[5] = {0, 0, 0, 0, 0};
  req[which] = need;
  *st = ensure_capacity(e, req[0], req[1], req[2], req[3], req[4], 0);
  if (*st) return -1;
  off = a.alloc(n, align);
  if (off < 0) *st = set_error(e, SOP_ERR_CAPACITY, "pool allocation failed after growth");
  return off;
}

// --------------------------------------------------------------- helpers

static float cutoff_host(double factor, float absval, float base) {
  return factor > 0.0 ? (float)(factor * (double)base) : absval;
}

static inline int32_t round_up(int32_t v, int32_t a) { return (v + a - 1) / a * a; }

// Row capacity of a system's shell rows and pair-list tiles (entries per
// bead): it starts at PL_CAP_START, is fitted to the longest row of the first
// build, and grows when a rebuild's longest row comes within an eighth of it,
// up to min(n - 1, pl_cap_max).  Growth moves the rows and changes no result.
constexpr int32_t PL_CAP_START = 384;

static int32_t pl_cap_limit(const sop_engine *e, int32_t n) {
  return round_up(std::max(1, std::min(n - 1, e->opts.pl_cap_max)), 16);
}
static int32_t pl_cap_start(const sop_engine *e, int32_t n) { return std::min(pl_cap_limit(e, n), PL_CAP_START); }
static int32_t pl_cap_target(const sop_engine *e, int32_t n, int32_t longest) {
  return std::min(pl_cap_limit(e, n), round_up(longest + longest / 4 + 32, 16));
}
static bool pl_cap_short(const sop_engine *e, const SysHost &h, int32_t longest) {
  return h.pl_cap < pl_cap_limit(e, h.n) && longest > h.pl_cap - h.pl_cap / 8;
}

namespace sopk {

// Pair-list shell of a system: rows within rc_pl + shell of pos_shell, valid
// while 2 max|x - pos_shell| < shell_ok (the shell less a 1e-4 margin, so the
// rounding of the shell cutoffs cannot matter); shell 0: no shell rows.  The
// rows are rebuilt from the cell grid at the next pair-list rebuild.
void shell_fields(const sop_params &P, float shell, SysDev &S) {
  float max_rcp = 0.0f;
  for (int q = 0; q < 4; ++q) {
    const int ti = q / 2 + 1, tj = q % 2 + 1;
    max_rcp = std::max(max_rcp, cutoff_host(P.pl_factor, P.pl_abs, P.sigma_rep[ti][tj]));
  }
  S.shell_ok = shell * (1.0f - 1.0e-4f);
  for (int q = 0; q < 4; ++q) {
    const int ti = q / 2 + 1, tj = q % 2 + 1;
    const double rs = (double)cutoff_host(P.pl_factor, P.pl_abs, P.sigma_rep[ti][tj]) + (double)shell;
    S.rc2_sl_rep[q] = (float)(rs * rs);
  }
  S.cell_min_edge_sl = std::max(1.0e-3f, (max_rcp + shell) * 1.001f);
  S.sl_valid = 0;
}

// Size bucket of a system of n beads in the tuned table (-1: none).
int engine_bucket(const sop_engine *e, int32_t n) {
  for (int b = 0; b < e->tune.nb; ++b)
    if (n <= e->tune.max_beads[b] || b == e->tune.nb - 1) return b;
  return -1;
}

float engine_shell_for(const sop_engine *e, int32_t n) {
  const int b = engine_bucket(e, n);
  return b >= 0 ? e->tune.shell[b] : e->opts.pl_shell;
}

sop_status engine_read_sys(sop_engine *e, int32_t slot, SysDev *out) {
  HIPCK(hipMemcpyAsync(out, e->p.sys + slot, sizeof(SysDev), hipMemcpyDeviceToHost, e->stream));
  HIPCK(hipStreamSynchronize(e->stream));
  return SOP_OK;
}

int64_t job_bytes_estimate(const sop_engine *e, const sop_system_desc *d) {
  const int64_t n = d->n_beads;
  const int64_t pl_cap = pl_cap_start(e, d->n_beads);
  const int64_t beads = (n + TILE - 1) / TILE * TILE;
  const int64_t cell_cap = std::min<int64_t>(e->opts.cell_cap_max, std::max<int64_t>(64, 2 * n)) + 1;
  int64_t per_bead = 4;  // tile table share
  for (const auto &a : e->bead_arrays) per_bead += (int64_t)a.second;
  // pair-list and shell-row segments (16-bit entries up to 65,536 beads); the
  // neighbor list is materialized only on request, in a scratch segment
  const int64_t esz = n <= 65536 ? 2 : 4;
  int64_t per_cell = 0;
  for (const auto &a : e->cell_arrays) per_cell += (int64_t)a.second;
  return beads * per_bead + 24 * 2 * (int64_t)d->n_native + 2 * esz * n * pl_cap + per_cell * cell_cap;
}

// Validate and upload one system, without its first lists (engine_init_lists).
sop_status engine_add_system_nolists(sop_engine *e, const sop_system_desc *d, int64_t step_end, int32_t *slot_out,
                                     const Upload *upl) {
  hipStream_t up = upl ? upl->stream : nullptr;
  if (!d || d->n_beads < 1 || !d->xyz) return set_error(e, SOP_ERR_INVALID, "system: need n_beads >= 1 and xyz");
  const int32_t n = d->n_beads;
  const sop_params &P = d->params;
  if (P.n_nl < 1 || P.n_pl < 1) return set_error(e, SOP_ERR_INVALID, "n_nl and n_pl must be >= 1");
  if (P.dynamics != SOP_DYN_OVERDAMPED && P.dynamics != SOP_DYN_UNDERDAMPED)
    return set_error(e, SOP_ERR_INVALID, "dynamics must be SOP_DYN_OVERDAMPED or SOP_DYN_UNDERDAMPED");
  const bool under = P.dynamics == SOP_DYN_UNDERDAMPED;
  if (!(under ? P.zeta >= 0.0f : P.zeta > 0.0f) || !(P.h > 0.0f) || !(P.T >= 0.0f) || !(P.R0 > 0.0f))
    return set_error(e, SOP_ERR_INVALID, "need zeta > 0 (>= 0 underdamped), h > 0, T >= 0, R0 > 0");
  if (!(P.nl_factor > 0.0) && !(P.nl_abs > 0.0f)) return set_error(e, SOP_ERR_INVALID, "neighbor-list cutoff");
  if (!(P.pl_factor > 0.0) && !(P.pl_abs > 0.0f)) return set_error(e, SOP_ERR_INVALID, "pair-list cutoff");
  // chains and types
  std::vector<int32_t> chain(n, 0), type(n, 1);
  for (int i = 0; i < n; ++i) {
    if (d->chain) chain[i] = d->chain[i];
    if (d->type) type[i] = d->type[i];
    if (type[i] != 1 && type[i] != 2) return set_error(e, SOP_ERR_INVALID, "bead types must be 1 or 2");
    if (chain[i] < 0 || chain[i] >= (1 << 28)) return set_error(e, SOP_ERR_INVALID, "chain id out of range");
    if (i > 0 && chain[i] < chain[i - 1]) return set_error(e, SOP_ERR_INVALID, "chain ids must be non-decreasing");
  }
  // natives: validate, two-sided CSR sorted by partner
  const int32_t nn = d->n_native;
  if (nn < 0 || (nn > 0 && (!d->nat_i || !d->nat_j || !d->nat_r0)))
    return set_error(e, SOP_ERR_INVALID, "native arrays");
  std::vector<std::vector<std::pair<int32_t, float>>> nat(n);
  for (int k = 0; k < nn; ++k) {
    const int i = d->nat_i[k], j = d->nat_j[k];
    if (i < 0 || j < 0 || i >= n || j >= n || i >= j)
      return set_error(e, SOP_ERR_INVALID, "native pair must satisfy 0 <= i < j < n");
    if (chain[i] == chain[j] && j - i <= 2)
      return set_error(e, SOP_ERR_INVALID, "native pair inside the |i-j| <= 2 exclusion");
    if (!(d->nat_r0[k] > 0.0f)) return set_error(e, SOP_ERR_INVALID, "native r0 must be > 0");
    nat[i].push_back({j, d->nat_r0[k]});
    nat[j].push_back({i, d->nat_r0[k]});
  }
  for (int i = 0; i < n; ++i) {
    std::sort(nat[i].begin(), nat[i].end());
    for (size_t q = 1; q < nat[i].size(); ++q)
      if (nat[i][q].first == nat[i][q - 1].first) return set_error(e, SOP_ERR_INVALID, "duplicate native pair");
    if (nat[i].size() > 65535) return set_error(e, SOP_ERR_CAPACITY, "more than 65535 native partners of one bead");
  }
  // the step kernel holds a native list's offset within its tile's native
  // slots in 16 bits
  for (int t0 = 0; t0 < n; t0 += TILE) {
    size_t tn = 0;
    for (int i = t0; i < std::min(n, t0 + TILE); ++i) tn += nat[i].size();
    if (tn > 65535) return set_error(e, SOP_ERR_CAPACITY, "more than 65535 native partners in one 64-bead tile");
  }
  // capacities
  const int32_t nl_cap = round_up(std::max(1, std::min(n - 1, e->opts.nl_cap_max)), 16);
  const int32_t pl_cap = pl_cap_start(e, n);
  // the step kernel addresses a system's pair-list rows with 32-bit byte
  // offsets, at any capacity they may grow to
  if ((int64_t)n * pl_cap_limit(e, n) * (n <= 65536 ? 2 : 4) > (int64_t)0xFFFFFFFF)
    return set_error(e, SOP_ERR_CAPACITY, "system's pair-list rows exceed 4 GiB (lower pl_cap_max)");
  const int32_t cell_cap = std::min(e->opts.cell_cap_max, std::max(64, 2 * n));
  int32_t dmax = 1;
  while ((int64_t)(dmax + 1) * (dmax + 1) * (dmax + 1) <= cell_cap) ++dmax;
  SysHost h;
  h.used = true;
  h.n = n;
  h.n_nat = nn;
  h.bead_len = round_up(n, TILE);
  h.nat_len = std::max(1, 2 * nn);
  h.nl_len = (int64_t)n * nl_cap;  // entries when materialized (scratch)
  h.pl_len = (int64_t)n * pl_cap / (n <= 65536 ? 2 : 1);  // 32-bit words: 16-bit entries pack two a word
  h.cell_len = cell_cap + 1;
  h.nl_cap = nl_cap;
  h.pl_cap = pl_cap;
  h.cell_cap = cell_cap;
  h.step = d->step0;
  h.parity = 0;
  h.params = P;
  // with an upload stream, every upload goes through the pinned staging
  // buffer (a copy from pageable memory would wait for the engine stream)
  const size_t bl = (size_t)h.bead_len, nl_b = (size_t)h.nat_len;
  const size_t stage_need = 32 * 256 + 16 * bl * 3 + 4 * bl * 9 + 8 * nl_b + bl + 4 * (bl / TILE) + sizeof(SysDev);
  if (up && stage_need > upl->stage_bytes) return SOP_ERR_BUSY;
  sop_status st;
  // with an upload stream, free ranges only: growing a pool reallocates
  // arrays that work in flight on the engine stream may be using
  auto alloc = [&](RangeAlloc &a, int64_t len, int64_t align, int which) -> int64_t {
    if (!up) return alloc_or_grow(e, a, len, align, which, &st);
    const int64_t off = a.alloc(len, align);
    st = off < 0 ? SOP_ERR_BUSY : SOP_OK;
    return off;
  };
  h.bead_off = alloc(e->a_beads, h.bead_len, TILE, 0);
  if (st) return st;
  h.nat_off = alloc(e->a_nat, h.nat_len, 1, 1);
  if (st) { e->a_beads.release(h.bead_off, h.bead_len); return st; }
  if ((h.nat_off + h.nat_len) * 8 > (int64_t)0xFFFFFFFF) {  // 32-bit byte offsets of the native pair list
    e->a_beads.release(h.bead_off, h.bead_len); e->a_nat.release(h.nat_off, h.nat_len);
    return set_error(e, SOP_ERR_CAPACITY, "native pool beyond 4 GiB");
  }
  h.nl_off = 0;
  h.pl_off = alloc(e->a_pl, h.pl_len, 16, 3);
  if (st) { e->a_beads.release(h.bead_off, h.bead_len); e->a_nat.release(h.nat_off, h.nat_len); return st; }
  h.cell_off = alloc(e->a_cells, h.cell_len, 1, 4);
  if (st) {
    e->a_beads.release(h.bead_off, h.bead_len); e->a_nat.release(h.nat_off, h.nat_len);
    e->a_pl.release(h.pl_off, h.pl_len);
    return st;
  }
  // slot
  int32_t slot = -1;
  for (size_t s = 0; s < e->sys.size(); ++s)
    if (!e->sys[s].used) { slot = (int32_t)s; break; }
  if (slot < 0 && up) {
    e->a_beads.release(h.bead_off, h.bead_len); e->a_nat.release(h.nat_off, h.nat_len);
    e->a_pl.release(h.pl_off, h.pl_len);
    e->a_cells.release(h.cell_off, h.cell_len);
    return SOP_ERR_BUSY;
  }
  if (slot < 0) {
    if (e->fixed_pools && (int64_t)e->sys.size() >= e->cap.slots && e->cap.slots > 0) {
      slot = -1;
    } else {
      const int64_t want = e->cap.slots + 1;
      st = ensure_capacity(e, 0, 0, 0, 0, 0, want);
      if (st) return st;
      for (size_t s = 0; s < e->sys.size(); ++s)
        if (!e->sys[s].used) { slot = (int32_t)s; break; }
    }
    if (slot < 0) return set_error(e, SOP_ERR_CAPACITY, "no free system slot");
  }
  // host staging
  const int64_t b0 = h.bead_off;
  std::vector<float4> pos((size_t)h.bead_len, make_float4(0, 0, 0, 0));
  std::vector<float> bond((size_t)h.bead_len, 0.0f);
  std::vector<int32_t> nstart((size_t)h.bead_len, (int32_t)h.nat_off), ncnt((size_t)h.bead_len, 0);
  std::vector<int32_t> zeros((size_t)h.bead_len, 0);
  std::vector<int32_t> natj((size_t)h.nat_len, 0);
  std::vector<float> natr((size_t)h.nat_len, 0.0f);
  std::vector<float4> v((size_t)h.bead_len, make_float4(0, 0, 0, 0));
  std::vector<uint8_t> ord((size_t)h.bead_len);
  std::vector<int32_t> ts((size_t)(h.bead_len / TILE), slot);
  SysDev S;
  // host buffers above stay alive until the uploads that read them are done
  struct UploadDrain {
    hipStream_t s;
    ~UploadDrain() { if (s) (void)hipStreamSynchronize(s); }
  } drain{up};
  size_t staged = 0;
  auto put = [&](void *dst, const void *src, size_t bytes) {
    if (!up) return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    char *q = static_cast<char *>(upl->stage) + staged;
    staged += (bytes + 255) / 256 * 256;
    std::memcpy(q, src, bytes);
    return hipMemcpyAsync(dst, q, bytes, hipMemcpyHostToDevice, up);
  };
  h.wbits.resize((size_t)n);
  int32_t cur = 0;
  for (int i = 0; i < n; ++i) {
    const int32_t wb = (chain[i] << 2) | type[i];
    h.wbits[i] = wb;
    float wf;
    std::memcpy(&wf, &wb, 4);
    pos[i] = make_float4(d->xyz[3 * i], d->xyz[3 * i + 1], d->xyz[3 * i + 2], wf);
    bond[i] = d->bond_r0 ? d->bond_r0[i] : 3.8f;
    nstart[i] = (int32_t)h.nat_off + cur;
    ncnt[i] = (int32_t)nat[i].size();
    for (auto &pr : nat[i]) {
      natj[cur] = (int32_t)b0 + pr.first;
      natr[cur] = pr.second;
      ++cur;
    }
  }
  const DevPtrs &p = e->p;
  HIPCK(put(p.pos0 + b0, pos.data(), 16 * h.bead_len));
  HIPCK(put(p.pos1 + b0, pos.data(), 16 * h.bead_len));
  HIPCK(put(p.bond_r0 + b0, bond.data(), 4 * h.bead_len));
  HIPCK(put(p.nat_start + b0, nstart.data(), 4 * h.bead_len));
  HIPCK(put(p.nat_cnt + b0, ncnt.data(), 4 * h.bead_len));
  for (int32_t *arr : {p.nlatt_cnt, p.platt_cnt, p.nlrep_cnt, p.plrep_cnt, p.plrep_off, p.platt_off})
    HIPCK(put(arr + b0, zeros.data(), 4 * h.bead_len));
  // underdamped state: the velocities entering the first step (w = 0)
  if (d->vel)
    for (int i = 0; i < n; ++i) v[i] = make_float4(d->vel[3 * i], d->vel[3 * i + 1], d->vel[3 * i + 2], 0.0f);
  HIPCK(put(p.vel + b0, v.data(), 16 * h.bead_len));
  HIPCK(put(p.nat_j + h.nat_off, natj.data(), 4 * h.nat_len));
  HIPCK(put(p.nat_r0 + h.nat_off, natr.data(), 4 * h.nat_len));
  // identity bead order per tile until the first pair list orders it
  for (size_t k = 0; k < ord.size(); ++k) ord[k] = (uint8_t)(k % TILE);
  HIPCK(put(p.tile_ord + b0, ord.data(), ord.size()));
  HIPCK(put(p.tile_sys + b0 / TILE, ts.data(), 4 * ts.size()));
  // device descriptor
  std::memset(&S, 0, sizeof(S));
  S.bead_off = (int32_t)b0;
  S.n = n;
  S.nl_cap = nl_cap;
  S.pl_cap = pl_cap;
  S.idx16 = n <= 65536 ? 1 : 0;
  S.nl_base = 0;  // the neighbor-list scratch, when materialized
  S.pl_base = h.pl_off;
  S.cell_base = (int32_t)h.cell_off;
  S.cell_cap = cell_cap;
  S.dmax = dmax;
  S.parity = 0;
  S.step = d->step0;
  S.step_end = step_end > d->step0 ? step_end : d->step0;
  S.n_nl = P.n_nl;
  S.n_pl = P.n_pl;
  float a5, var;
  sop_params_derived(&P, &a5, &var);
  S.a5 = a5;
  S.dyn = under ? 1 : 0;
  S.vupd = P.T >= 1.0e-5f ? 1 : 0;  // sop.cu:475 (no velocity update for steepest descent)
  {
    float a[4];
    sop_params_underdamped(&P, a);
    S.a1 = a[0];
    S.a2 = a[1];
    S.a3 = a[2];
    S.a4 = a[3];
  }
  S.var = var;
  S.boxl = P.boxl;
  S.inv_boxl = P.boxl > 0.0f ? 1.0f / P.boxl : 0.0f;
  S.seed_lo = (uint32_t)P.seed;
  S.seed_hi = (uint32_t)(P.seed >> 32);
  S.k_bnd = P.k_bnd;
  S.inv_R0sq = 1.0f / (P.R0 * P.R0);
  S.half_k_R0sq = 0.5f * P.k_bnd * P.R0 * P.R0;
  const double s6 = std::pow((double)P.sigma_ss, 6.0);
  S.ss_e6 = (float)((double)P.eps_ss * s6);
  S.ss_f6 = (float)(6.0 * (double)P.eps_ss * s6);
  float max_rc = 0.0f;
  for (int ti = 1; ti <= 2; ++ti)
    for (int tj = 1; tj <= 2; ++tj) {
      const int q = (ti - 1) * 2 + (tj - 1);
      const float sg = P.sigma_rep[ti][tj], ep = P.eps_rep[ti][tj];
      S.eps_att[q] = P.eps_att[ti][tj];
      S.eps_rep[q] = ep;
      S.sig2_rep[q] = sg * sg;
      S.rep_guard[q] = sg * 1.0e-7f;
      S.rep_m12[q] = -12.0f * ep * P.rep_c12;
      S.rep_m6[q] = -6.0f * ep * P.rep_c6;
      const double s6q = std::pow((double)S.sig2_rep[q], 3.0);
      S.rep_a[q] = (float)((double)S.rep_m12[q] * s6q * s6q);
      S.rep_b[q] = (float)((double)S.rep_m6[q] * s6q);
      const float rn = cutoff_host(P.nl_factor, P.nl_abs, sg);
      const float rp = cutoff_host(P.pl_factor, P.pl_abs, sg);
      S.rc2_nl_rep[q] = rn * rn;
      S.rc2_pl_rep[q] = rp * rp;
      max_rc = std::max(max_rc, rn);
    }
  S.c12 = P.rep_c12;
  S.c6 = P.rep_c6;
  S.nl_factor = P.nl_factor;
  S.pl_factor = P.pl_factor;
  S.nl_abs = P.nl_abs;
  S.pl_abs = P.pl_abs;
  S.cell_min_edge = std::max(1.0e-3f, max_rc * 1.001f);
  // pair-list grid and list skin (the direct pair-list build is exact while
  // every bead stays within skin/2 of its neighbor-list snapshot)
  float skin = INFINITY;
  for (int q = 0; q < 4; ++q) {
    const int ti = q / 2 + 1, tj = q % 2 + 1;
    const float sg = P.sigma_rep[ti][tj];
    const float rn = cutoff_host(P.nl_factor, P.nl_abs, sg), rp = cutoff_host(P.pl_factor, P.pl_abs, sg);
    skin = std::min(skin, rn - rp);
  }
  for (int k = 0; k < nn; ++k) {
    const float r0 = d->nat_r0[k];
    skin = std::min(skin, cutoff_host(P.nl_factor, P.nl_abs, r0) - cutoff_host(P.pl_factor, P.pl_abs, r0));
  }
  S.skin = skin > 0.0f ? skin * (1.0f - 1.0e-5f) : 0.0f;
  h.shell = engine_shell_for(e, n);
  shell_fields(P, h.shell, S);
  S.build = BUILD_NONE;
  S.nl_valid = 0;
  int tmask = 0;
  for (int i = 0; i < n; ++i) tmask |= 1 << (type[i] - 1);
  S.type_mask = tmask;
  S.wrap_margin = 10.0f;
  S.dims[0] = S.dims[1] = S.dims[2] = 1;
  S.ncells = 1;
  S.wrap_free = 0;
  HIPCK(put(p.sys + slot, &S, sizeof(SysDev)));
  if (up) HIPCK(hipStreamSynchronize(up));
  e->sys[slot] = std::move(h);
  *slot_out = slot;
  return SOP_OK;
}

static sop_status read_sys_all(sop_engine *e, std::vector<SysDev> &sd) {
  sd.resize((size_t)e->cap.slots);
  HIPCK(hipMemcpy(sd.data(), e->p.sys, sizeof(SysDev) * sd.size(), hipMemcpyDeviceToHost));
  return SOP_OK;
}

// Resident bytes of one system (what job_bytes_estimate predicts for it).
int64_t engine_system_bytes(const sop_engine *e, int32_t slot) {
  const SysHost &h = e->sys[slot];
  int64_t per_bead = 4;  // tile table share
  for (const auto &a : e->bead_arrays) per_bead += (int64_t)a.second;
  int64_t per_cell = 0;
  for (const auto &a : e->cell_arrays) per_cell += (int64_t)a.second;
  return h.bead_len * per_bead + 24 * 2 * (int64_t)h.n_nat + 2 * 4 * h.pl_len + per_cell * h.cell_len;
}

// Move one system's shell rows and pair-list tiles to a segment of row
// capacity cap: with keep, the current rows are copied over (k_relayout);
// otherwise the caller builds them again.  Leaves the engine stream idle.
sop_status engine_resize_rows(sop_engine *e, int32_t slot, int32_t cap, bool keep) {
  SysHost &h = e->sys[slot];
  if (cap == h.pl_cap) return SOP_OK;
  const bool i16 = h.n <= 65536;
  const int64_t len = (int64_t)h.n * cap / (i16 ? 2 : 1);
  sop_status st = SOP_OK;
  const int64_t off = alloc_or_grow(e, e->a_pl, len, 16, 3, &st);
  if (st) return st;
  if (keep)
    HIPCK(launch_relayout(e->p, (int)h.bead_off, h.n, i16 ? 1 : 0, h.pl_off, h.pl_cap, off, cap, e->stream));
  const int64_t base = off;
  const int32_t c32 = cap, zero = 0;
  HIPCK(hipMemcpyAsync(&e->p.sys[slot].pl_base, &base, sizeof(base), hipMemcpyHostToDevice, e->stream));
  HIPCK(hipMemcpyAsync(&e->p.sys[slot].pl_cap, &c32, sizeof(c32), hipMemcpyHostToDevice, e->stream));
  if (!keep)  // no shell rows to filter: the next pair-list build comes from the cell grid
    HIPCK(hipMemcpyAsync(&e->p.sys[slot].sl_valid, &zero, sizeof(zero), hipMemcpyHostToDevice, e->stream));
  HIPCK(hipStreamSynchronize(e->stream));
  e->a_pl.release(h.pl_off, h.pl_len);
  h.pl_off = off;
  h.pl_len = len;
  h.pl_cap = cap;
  return SOP_OK;
}

// Systems whose longest row came within an eighth of their row capacity get
// a larger one; their rows move and no result changes.  With fixed pools a
// capacity that finds no room stays as it is (a row that outgrows it is
// reported as an overflow).  The engine stream must be idle; sd is the
// systems' current state.  *grown receives each system's added bytes.
sop_status engine_grow_rows(sop_engine *e, const int32_t *slots, int32_t n, const std::vector<SysDev> &sd,
                            std::vector<int64_t> *grown) {
  if (grown) grown->assign((size_t)n, 0);
  if (!e->grow_rows) return SOP_OK;
  for (int32_t q = 0; q < n; ++q) {
    const int32_t s = slots[q];
    const SysHost &h = e->sys[s];
    const int32_t m = sd[s].row_max;
    if (!pl_cap_short(e, h, m)) continue;
    const int64_t before = engine_system_bytes(e, s);
    const sop_status st = engine_resize_rows(e, s, pl_cap_target(e, h.n, m), true);
    if (st == SOP_ERR_CAPACITY && e->fixed_pools) continue;
    if (st) return st;
    if (grown) (*grown)[q] = engine_system_bytes(e, s) - before;
  }
  return SOP_OK;
}

// Before a system's shell widens to `shell`: a row capacity that the longest
// row so far, scaled by the volume of the wider shell (and a tenth more),
// would come within an eighth of grows now (rows moved, results unchanged),
// so the first rebuild with the new shell does not overflow.
sop_status engine_fit_shell(sop_engine *e, int32_t slot, float shell) {
  SysHost &h = e->sys[slot];
  if (!(shell > h.shell)) return SOP_OK;
  SysDev S;
  sop_status st = engine_read_sys(e, slot, &S);
  if (st) return st;
  float rc2 = 0.0f;
  for (int q = 0; q < 4; ++q) rc2 = std::max(rc2, S.rc2_pl_rep[q]);
  const double rc = std::sqrt((double)rc2);
  const double ratio = (rc + shell) / (rc + h.shell);
  const int32_t m = (int32_t)std::ceil((double)S.row_max * ratio * ratio * ratio * 1.1);
  if (S.row_max <= 0 || !pl_cap_short(e, h, m)) return SOP_OK;
  return engine_resize_rows(e, slot, pl_cap_target(e, h.n, m), true);
}

// After the first lists of newly added systems: a system whose longest row
// overflowed its starting capacity or comes within an eighth of it gets the
// fitted capacity, and its first lists are built again from the same
// coordinates (the lists a fitted first build gives); its build counts are
// set back to one build.  Capacities only grow, so a system never holds less
// than job_bytes_estimate charged for it.
sop_status engine_fit_new(sop_engine *e, const int32_t *slots, int32_t n) {
  std::vector<SysDev> sd;
  sop_status st = read_sys_all(e, sd);
  if (st) return st;
  std::vector<int32_t> redo;
  for (int32_t q = 0; q < n; ++q) {
    const int32_t s = slots[q];
    const SysHost &h = e->sys[s];
    const SysDev &S = sd[s];
    if (S.err & ~ERR_PL_OVF) continue;  // failed for another reason: reported as it is
    const int32_t m = S.row_max, want = pl_cap_target(e, h.n, m);
    const bool over = (S.err & ERR_PL_OVF) != 0;
    if (!over && !pl_cap_short(e, h, m)) continue;
    if (want == h.pl_cap || want < m) continue;  // at the limit: an overflow stands
    st = engine_resize_rows(e, s, want, false);
    if (st) return st;
    redo.push_back(s);
  }
  if (redo.empty()) return SOP_OK;
  const int32_t zero = 0;
  for (int32_t s : redo) {
    HIPCK(hipMemcpy(&e->p.sys[s].err, &zero, sizeof(zero), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(&e->p.sys[s].row_max, &zero, sizeof(zero), hipMemcpyHostToDevice));
  }
  st = engine_set_batch(e, redo.data(), (int32_t)redo.size());
  if (st) return st;
  HIPCK(launch_lists(e->p, list_grids(e), 0, FORCE_NL | FORCE_PL, e->stream));
  HIPCK(hipStreamSynchronize(e->stream));
  for (int32_t s : redo)
    HIPCK(hipMemcpy(e->p.sys[s].nbuild, sd[s].nbuild, sizeof(sd[s].nbuild), hipMemcpyHostToDevice));
  return SOP_OK;
}

// First neighbor and pair lists of newly added systems, in one launch
// sequence (the reference builds both before its step loop), with row
// capacities fitted.  status[q] receives each system's result; systems that
// fail are removed.
sop_status engine_init_lists(sop_engine *e, const int32_t *slots, int32_t n, sop_status *status) {
  if (n <= 0) return SOP_OK;
  sop_status st = engine_set_batch(e, slots, n);
  if (st) return st;
  HIPCK(launch_lists(e->p, list_grids(e), 0, FORCE_NL | FORCE_PL, e->stream));
  HIPCK(hipStreamSynchronize(e->stream));
  st = engine_fit_new(e, slots, n);
  if (st) return st;
  for (int32_t q = 0; q < n; ++q) {
    status[q] = engine_check_errors(e, &slots[q], 1);
    if (status[q]) engine_remove_system(e, slots[q]);
  }
  return SOP_OK;
}

sop_status engine_add_system(sop_engine *e, const sop_system_desc *d, int32_t *slot_out) {
  int32_t slot = -1;
  sop_status st = engine_add_system_nolists(e, d, 0, &slot);
  if (st) return st;
  sop_status one = SOP_OK;
  st = engine_init_lists(e, &slot, 1, &one);
  if (st) return st;
  if (one) return one;
  *slot_out = slot;
  return SOP_OK;
}

sop_status engine_remove_system(sop_engine *e, int32_t slot) {
  if (slot < 0 || slot >= (int32_t)e->sys.size() || !e->sys[slot].used)
    return set_error(e, SOP_ERR_INVALID, "no such system");
  SysHost &h = e->sys[slot];
  e->a_beads.release(h.bead_off, h.bead_len);
  e->a_nat.release(h.nat_off, h.nat_len);
  e->a_pl.release(h.pl_off, h.pl_len);
  if (e->nl_owner == slot) e->nl_owner = -1;
  e->a_cells.release(h.cell_off, h.cell_len);
  h = SysHost();
  return SOP_OK;
}

// Upload the work tables for a batch of systems.  Systems are spread over the
// 8 XCDs by bead count; entry k of the work list is processed by a block on
// XCD k % 8 (round-robin dispatch with a grid that is a multiple of 8), so all
// tiles of a system share one XCD's L2.
sop_status engine_set_batch(sop_engine *e, const int32_t *slots, int32_t nslots) {
  constexpr int NX = 8;
  std::vector<int32_t> order(slots, slots + nslots);
  std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    if (e->sys[a].bead_len != e->sys[b].bead_len) return e->sys[a].bead_len > e->sys[b].bead_len;
    return a < b;
  });
  std::vector<std::vector<int32_t>> bins(NX), bsys(NX);
  std::vector<int64_t> load(NX, 0);
  for (int32_t s : order) {
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    const SysHost &h = e->sys[s];
    for (int64_t t = h.bead_off / TILE; t < (h.bead_off + h.n + TILE - 1) / TILE; ++t) {
      bins[x].push_back((int32_t)t);
      bsys[x].push_back(s);
    }
    load[x] += (h.n + TILE - 1) / TILE;
  }
  size_t mx = 0;
  for (auto &b : bins) mx = std::max(mx, b.size());
  e->h_work.assign(mx * NX, -1);
  for (int x = 0; x < NX; ++x)
    for (size_t m = 0; m < bins[x].size(); ++m) e->h_work[m * NX + x] = bins[x][m];
  // Work items of the step kernel and of the shell filter: {first tile, tiles
  // (| waves per tile << 8 for the step kernel), first staged bead or -1,
  // staged beads}.  Systems that fit the LDS budget are staged whole:
  // arena-adjacent small systems share an item (one staged range), a larger
  // system gets one item per `per` tiles, each staging the whole system.
  // Larger systems gather from global memory and keep their items on one XCD
  // (runs of one XCD's share on several XCDs when they have more items than
  // that).  The shell filter takes one wave per tile, STEP_WAVES tiles per item.
  // The step kernel's items hold `per` tiles and run each tile on
  // min(lanes_per_bead, STEP_WAVES / tiles) waves (power of two), each wave
  // taking that share of the tile's rounds: an item's duration shrinks with
  // its tile count, so small items do not leave waves idle while the big
  // ones finish.
  const int64_t lds_cap = e->opts.lds_beads_max > 0
                              ? std::min<int64_t>(e->opts.lds_beads_max, (int64_t)(e->lds_limit / 16))
                              : 0;
  std::vector<int32_t> by_arena(slots, slots + nslots);
  std::sort(by_arena.begin(), by_arena.end(),
            [&](int32_t a2, int32_t b2) { return e->sys[a2].bead_off < e->sys[b2].bead_off; });
  // tiles per step item: the option, or 16 while the batch's tiles fill the
  // wave slots of the LDS launch (two blocks of STEP_WAVES per compute unit),
  // 8 or 4 below that, so that a small batch still fills the device
  int64_t total_tiles = 0;
  for (int32_t q = 0; q < nslots; ++q) total_tiles += e->sys[slots[q]].bead_len / TILE;
  const int64_t slots_w = (int64_t)e->num_cu * STEP_WAVES;
  // systems above lds_cap that fit one compute unit's whole LDS take the
  // whole-CU tier (one block per unit) instead of global gathers.  While
  // serving only under load (more tiles than two LDS launches' wave slots):
  // there it raises the stream's capacity (3.8e9 -> 4.9e9 bead-steps/s at its
  // stated parameters), while under a light load the global-gather tier,
  // spread over more blocks, gives the shorter step (served p50 0.25 s
  // against 0.30 s)
  const int64_t big_cap = e->fixed_grids && total_tiles <= 2 * slots_w ? 0 : e->lds_big_cap;
  const int lpb = e->opts.lanes_per_bead;
  int32_t per_step = e->opts.step_tiles;
  if (per_step <= 0) per_step = total_tiles > slots_w ? STEP_WAVES : (2 * total_tiles > slots_w ? 8 : 4);
  e->cur_step_tiles = per_step;
  struct Plan {
    std::vector<int4> items;
    int32_t n_lds = 0, n_big = 0, n_far = 0;
    size_t lds = 0, lds_big = 0;
  };
  // tiles per item of a system: its size bucket's tuned value, else per_step
  auto per_of = [&](const SysHost &h) {
    const int b = engine_bucket(e, h.n);
    return b >= 0 && e->tune.tiles[b] > 0 ? e->tune.tiles[b] : per_step;
  };
  auto plan = [&](int32_t per_fixed, bool split, Plan &out) {
    // waves per tile of an item of nt tiles
    auto waves = [&](int32_t nt) {
      int32_t w = 1;
      if (split)
        while (w * 2 <= lpb && nt * w * 2 <= STEP_WAVES) w *= 2;
      return w;
    };
    auto mk = [&](int32_t t0, int32_t nt, int32_t z, int32_t wlen) {
      return make_int4(t0, nt | (waves(nt) << 8), z, wlen);
    };
    std::vector<int4> lds_items, big_items;
    std::vector<std::vector<int4>> far_items;  // per non-LDS system
    int4 cur = make_int4(-1, 0, -1, 0);        // .y: tiles while the item grows
    int32_t cur_per = 0;                       // tiles per item of the systems in cur
    auto flush = [&]() {
      if (cur.x >= 0) lds_items.push_back(mk(cur.x, cur.y, cur.z, cur.w));
      cur = make_int4(-1, 0, -1, 0);
    };
    for (int32_t sl : by_arena) {
      const SysHost &h = e->sys[sl];
      const int32_t t0 = (int32_t)(h.bead_off / TILE), nt = (int32_t)(h.bead_len / TILE);
      const int32_t per = per_fixed > 0 ? per_fixed : per_of(h);
      if (h.bead_len <= lds_cap) {
        out.lds = std::max(out.lds, (size_t)h.bead_len * 16);
        if (nt > per) {
          flush();
          for (int32_t k = 0; k < nt; k += per)
            lds_items.push_back(mk(t0 + k, std::min(per, nt - k), (int32_t)h.bead_off, (int32_t)h.bead_len));
          continue;
        }
        if (cur.x >= 0 && cur_per == per && (int64_t)cur.z + cur.w == h.bead_off && cur.y + nt <= per &&
            (int64_t)cur.w + h.bead_len <= lds_cap) {
          cur.y += nt;
          cur.w += (int32_t)h.bead_len;
          out.lds = std::max(out.lds, (size_t)cur.w * 16);
        } else {
          flush();
          cur = make_int4(t0, nt, (int32_t)h.bead_off, (int32_t)h.bead_len);
          cur_per = per;
        }
      } else if (h.bead_len <= big_cap) {
        flush();
        out.lds_big = std::max(out.lds_big, (size_t)h.bead_len * 16);
        for (int32_t k = 0; k < nt; k += per)
          big_items.push_back(mk(t0 + k, std::min(per, nt - k), (int32_t)h.bead_off, (int32_t)h.bead_len));
      } else {
        flush();
        std::vector<int4> v;
        for (int32_t k = 0; k < nt; k += per) v.push_back(mk(t0 + k, std::min(per, nt - k), -1, 0));
        far_items.push_back(v);
      }
    }
    flush();
    // XCD bins by load (item k runs on XCD k mod 8): the LDS items, longest
    // first (an item lasts about one tile's rounds / its waves per tile), then
    // whole non-LDS systems (largest first) so each keeps its gathers in one L2
    std::vector<std::vector<int4>> ibins(NX), bbins(NX), fbins(NX);
    std::vector<double> iload(NX, 0), bload(NX, 0), fload(NX, 0);
    auto by_dur = [&](const int4 &u, const int4 &v) {
      const double du = 1.0 / item_split(u.y), dv = 1.0 / item_split(v.y);
      return du != dv ? du > dv : item_tiles(u.y) > item_tiles(v.y);
    };
    std::stable_sort(lds_items.begin(), lds_items.end(), by_dur);
    // the tail: items that start only after a first round of full-length
    // items (two LDS blocks per compute unit) are cut into pieces of at most
    // STEP_WAVES / tail tiles, each with more waves per tile, so that the last
    // round is short
    if (split && e->tune.tail > 1) {
      const size_t slots = 2 * (size_t)e->num_cu;
      size_t nfull = 0;
      while (nfull < lds_items.size() && item_split(lds_items[nfull].y) == 1) ++nfull;
      if (nfull > slots) {
        std::vector<int4> cut(lds_items.begin(), lds_items.begin() + slots);
        const int32_t pmax = std::max(1, STEP_WAVES / e->tune.tail);
        for (size_t k = slots; k < lds_items.size(); ++k) {
          const int4 it = lds_items[k];
          const int32_t nt = item_tiles(it.y);
          for (int32_t a = 0; a < nt; a += pmax)
            cut.push_back(mk(it.x + a, std::min(pmax, nt - a), it.z, it.w));
        }
        std::stable_sort(cut.begin() + slots, cut.end(), by_dur);
        lds_items.swap(cut);
      }
    }
    // the items that stage one range (one system's items) go to one XCD, so
    // that their stagings after the first read the range from that XCD's L2;
    // ranges by load, largest first, each to the least-loaded XCD, then each
    // XCD's items longest first
    auto bin_by_range = [&](const std::vector<int4> &items, std::vector<std::vector<int4>> &bins,
                            std::vector<double> &load) {
      std::vector<std::pair<double, std::vector<int4>>> ranges;
      std::unordered_map<int32_t, size_t> at;
      for (const int4 &it : items) {
        auto f = at.find(it.z);
        if (f == at.end()) {
          f = at.emplace(it.z, ranges.size()).first;
          ranges.emplace_back(0.0, std::vector<int4>());
        }
        ranges[f->second].first += (double)item_tiles(it.y);
        ranges[f->second].second.push_back(it);
      }
      std::stable_sort(ranges.begin(), ranges.end(),
                       [](const auto &u, const auto &v) { return u.first > v.first; });
      for (const auto &r : ranges) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        bins[x].insert(bins[x].end(), r.second.begin(), r.second.end());
        load[x] += r.first;
      }
      for (auto &bn : bins) std::stable_sort(bn.begin(), bn.end(), by_dur);
    };
    bin_by_range(lds_items, ibins, iload);
    std::stable_sort(big_items.begin(), big_items.end(), by_dur);
    bin_by_range(big_items, bbins, bload);
    std::sort(far_items.begin(), far_items.end(), [](const std::vector<int4> &u, const std::vector<int4> &v) {
      return u.size() != v.size() ? u.size() > v.size() : u[0].x < v[0].x;
    });
    // a system with more items than one XCD's share of the global-gather
    // launch (2 blocks per compute unit) goes out in contiguous runs of that
    // share, each run to the least-loaded XCD: one XCD would leave the rest of
    // the device idle, and so large a system outgrows one L2 anyway
    const size_t share = (size_t)std::max(1, 2 * e->num_cu / NX);
    for (const auto &v : far_items) {
      for (size_t k0 = 0; k0 < v.size(); k0 += share) {
        const int x = (int)(std::min_element(fload.begin(), fload.end()) - fload.begin());
        for (size_t k = k0; k < std::min(v.size(), k0 + share); ++k) {
          fbins[x].push_back(v[k]);
          fload[x] += (double)item_tiles(v[k].y);
        }
      }
    }
    size_t imx = 0, bmx = 0, fmx = 0;
    for (auto &bn : ibins) imx = std::max(imx, bn.size());
    for (auto &bn : bbins) bmx = std::max(bmx, bn.size());
    for (auto &bn : fbins) fmx = std::max(fmx, bn.size());
    out.items.assign((imx + bmx + fmx) * NX, make_int4(-1, 0, -1, 0));
    for (int x = 0; x < NX; ++x) {
      for (size_t m = 0; m < ibins[x].size(); ++m) out.items[m * NX + x] = ibins[x][m];
      for (size_t m = 0; m < bbins[x].size(); ++m) out.items[(imx + m) * NX + x] = bbins[x][m];
      for (size_t m = 0; m < fbins[x].size(); ++m) out.items[(imx + bmx + m) * NX + x] = fbins[x][m];
    }
    out.n_lds = (int32_t)(imx * NX);
    out.n_big = (int32_t)(bmx * NX);
    out.n_far = (int32_t)(fmx * NX);
    // fixed grids (serving): the launches' LDS sizes too, so that one captured
    // graph serves every batch
    if (e->fixed_grids && lds_cap > 0) out.lds = (size_t)lds_cap * 16;
    if (e->fixed_grids && big_cap > 0) out.lds_big = (size_t)big_cap * 16;
  };
  Plan sp, fp;
  plan(0, true, sp);
  plan(STEP_WAVES, false, fp);
  e->h_items.swap(sp.items);
  e->h_fitems.swap(fp.items);
  e->cur_lds = sp.lds;
  e->cur_lds_big = sp.lds_big;
  e->cur_flds = fp.lds;
  e->cur_flds_big = fp.lds_big;
  e->h_syslist.assign(slots, slots + nslots);
  if ((int64_t)e->h_work.size() + 2 > e->work_cap) {
    HIPCK(hipStreamSynchronize(e->stream));
    int64_t nc = std::max<int64_t>(1024, (int64_t)e->h_work.size() * 2 + 2);
    if (e->p.work) HIPCK(hipFree(e->p.work));
    HIPCK(hipMalloc((void **)&e->p.work, 4 * (size_t)nc));
    e->work_cap = nc;
    invalidate_graphs(e);
  }
  if (!e->p.nwork) {
    HIPCK(hipMalloc((void **)&e->p.nwork, 256));
    e->p.nsys = e->p.nwork + 16;
    e->p.ncellwork = e->p.nwork + 32;
    e->p.nlbuild = e->p.nwork + 40;
    e->p.gsort = e->p.nwork + 41;
    e->p.nfitems = e->p.nwork + 44;
    e->p.nfitems_big = e->p.nwork + 45;
    e->p.nfitems_far = e->p.nwork + 46;
    e->p.nitems = e->p.nwork + 48;
    e->p.nitems_far = e->p.nwork + 56;
    e->p.nitems_big = e->p.nwork + 60;
    invalidate_graphs(e);
  }
  if ((int64_t)e->h_items.size() + 1 > e->items_cap) {
    HIPCK(hipStreamSynchronize(e->stream));
    const int64_t nc = std::max<int64_t>(1024, (int64_t)e->h_items.size() * 2 + 1);
    if (e->p.items) HIPCK(hipFree(e->p.items));
    HIPCK(hipMalloc((void **)&e->p.items, sizeof(int4) * (size_t)nc));
    e->items_cap = nc;
    invalidate_graphs(e);
  }
  if ((int64_t)e->h_fitems.size() + 1 > e->fitems_cap) {
    HIPCK(hipStreamSynchronize(e->stream));
    const int64_t nc = std::max<int64_t>(1024, (int64_t)e->h_fitems.size() * 2 + 1);
    if (e->p.fitems) HIPCK(hipFree(e->p.fitems));
    HIPCK(hipMalloc((void **)&e->p.fitems, sizeof(int4) * (size_t)nc));
    e->fitems_cap = nc;
    invalidate_graphs(e);
  }
  if (nslots > e->cap.slots) return set_error(e, SOP_ERR_INVALID, "batch larger than slot capacity");
  const int32_t nw = (int32_t)e->h_work.size(), ns = nslots;
  if (nw > 0) HIPCK(hipMemcpy(e->p.work, e->h_work.data(), 4 * (size_t)nw, hipMemcpyHostToDevice));
  if (ns > 0) HIPCK(hipMemcpy(e->p.syslist, e->h_syslist.data(), 4 * (size_t)ns, hipMemcpyHostToDevice));
  const int32_t ni = (int32_t)e->h_items.size(), nfi = (int32_t)e->h_fitems.size();
  if (ni > 0) HIPCK(hipMemcpy(e->p.items, e->h_items.data(), sizeof(int4) * (size_t)ni, hipMemcpyHostToDevice));
  if (nfi > 0) HIPCK(hipMemcpy(e->p.fitems, e->h_fitems.data(), sizeof(int4) * (size_t)nfi, hipMemcpyHostToDevice));
  // the counters in one copy: nwork[0] = work entries, [16] batch systems,
  // [44..46] filter items, [48] / [56] / [60] step items (LDS / far / whole-CU)
  int32_t cnt[64];
  HIPCK(hipMemcpy(cnt, e->p.nwork, sizeof(cnt), hipMemcpyDeviceToHost));
  cnt[0] = nw;
  cnt[16] = ns;
  cnt[44] = fp.n_lds;
  cnt[45] = fp.n_big;
  cnt[46] = fp.n_far;
  cnt[48] = sp.n_lds;
  cnt[56] = sp.n_far;
  cnt[60] = sp.n_big;
  HIPCK(hipMemcpy(e->p.nwork, cnt, sizeof(cnt), hipMemcpyHostToDevice));
  e->cur_nwork = nw;
  e->cur_nsys = ns;
  e->cur_nitems = sp.n_lds;
  e->cur_nitems_far = sp.n_far;
  e->cur_nitems_big = sp.n_big;
  e->cur_nfitems = fp.n_lds;
  e->cur_nfitems_far = fp.n_far;
  e->cur_nfitems_big = fp.n_big;
  return SOP_OK;
}

sop_status engine_error_status(sop_engine *e, int32_t s, int32_t err) {
  if (!err) return SOP_OK;
  std::ostringstream os;
  os << "system " << s << ":";
  sop_status st = SOP_ERR_STATE;
  if (err & ERR_NAN) { os << " non-finite coordinates;"; st = SOP_ERR_NUMERIC; }
  if (err & ERR_BOND) { os << " FENE bond stretched to |r-r0| >= R0;"; st = SOP_ERR_NUMERIC; }
  if (err & ERR_NL_OVF) { os << " neighbor list over capacity;"; st = SOP_ERR_CAPACITY; }
  if (err & ERR_PL_OVF) { os << " pair list over capacity;"; st = SOP_ERR_CAPACITY; }
  if (err & ERR_PBC) { os << " extent within 10 A of boxl/2 (open cell grid needs the minimum image to be the identity);"; st = SOP_ERR_PBC; }
  if (err & ERR_INTERNAL) { os << " inconsistent cell histogram;"; st = SOP_ERR_STATE; }
  return set_error(e, st, os.str());
}

sop_status engine_check_errors(sop_engine *e, const int32_t *slots, int32_t nslots) {
  std::vector<SysDev> sd((size_t)e->cap.slots);
  HIPCK(hipMemcpy(sd.data(), e->p.sys, sizeof(SysDev) * sd.size(), hipMemcpyDeviceToHost));
  for (int q = 0; q < nslots; ++q) {
    const sop_status st = engine_error_status(e, slots[q], sd[slots[q]].err);
    if (st) return st;
  }
  return SOP_OK;
}

static std::string graph_key(const sop_engine *e, int G, int n_nl, int n_pl) {
  std::ostringstream os;
  os << G << '|' << n_nl << '|' << n_pl << '|' << e->opts.lanes_per_bead << '|' << grid_tiles(e) << '|'
     << grid_sys(e) << '|' << grid_items(e) << '|' << grid_items_big(e) << '|' << grid_items_far(e) << '|'
     << e->cur_lds << '|' << e->cur_lds_big;
  const StepGrids fg = filter_grids(e);
  os << '|' << fg.grid << '|' << fg.grid_big << '|' << fg.grid_far << '|' << fg.lds << '|' << fg.lds_big;
  return os.str();
}

constexpr int64_t GROW_BLOCKS = 20;

// One hipGraph replay = G steps of an aligned batch (every system's step is a
// multiple of n_nl): NL rebuild at k % n_nl == 0, PL filter at k % n_pl == 0,
// fused step kernel every k, step-counter advance at the end.
sop_status engine_run_block_graph(sop_engine *e, int32_t G) {
  const SysHost &h0 = e->sys[e->h_syslist[0]];
  const int n_nl = h0.params.n_nl, n_pl = h0.params.n_pl;
  const std::string key = graph_key(e, G, n_nl, n_pl);
  auto it = e->graphs.find(key);
  hipGraphExec_t exec;
  if (it == e->graphs.end()) {
    hipGraph_t graph;
    const int lpb = e->opts.lanes_per_bead;
    HIPCK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < G; ++k) {
      if (k % n_pl == 0 || k % n_nl == 0) (void)launch_lists(e->p, list_grids(e), k, 0, e->stream, &e->fork);
      (void)launch_step(e->p, MODE_STEP, lpb, step_grids(e), k, e->stream, &e->fork);
    }
    (void)launch_advance(e->p, G, e->stream);
    HIPCK(hipStreamEndCapture(e->stream, &graph));
    HIPCK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    HIPCK(hipGraphDestroy(graph));
    e->graphs[key] = exec;
  } else {
    exec = it->second;
  }
  HIPCK(hipGraphLaunch(exec, e->stream));
  return SOP_OK;
}

// Host-driven steps for the current batch: launches exactly the rebuilds some
// system is due for (the kernels check each system's own schedule).
sop_status engine_run_hostloop(sop_engine *e, int64_t nsteps) {
  if (nsteps <= 0) return SOP_OK;
  const int lpb = e->opts.lanes_per_bead;
  std::vector<int64_t> st;
  std::vector<int32_t> nnl, npl;
  for (int32_t s : e->h_syslist) {
    st.push_back(e->sys[s].step);
    nnl.push_back(e->sys[s].params.n_nl);
    npl.push_back(e->sys[s].params.n_pl);
  }
  for (int64_t k = 0; k < nsteps; ++k) {
    bool any_nl = false, any_pl = false;
    for (size_t q = 0; q < st.size(); ++q) {
      any_nl |= ((st[q] + k) % nnl[q]) == 0;
      any_pl |= ((st[q] + k) % npl[q]) == 0;
    }
    if (any_nl || any_pl) HIPCK(launch_lists(e->p, list_grids(e), (int)k, 0, e->stream, &e->fork));
    HIPCK(launch_step(e->p, MODE_STEP, lpb, step_grids(e), (int)k, e->stream, &e->fork));
  }
  HIPCK(launch_advance(e->p, (int)nsteps, e->stream));
  return SOP_OK;
}

static void mirror_advance(sop_engine *e, const int32_t *sys, int32_t nsys, int64_t nsteps) {
  for (int q = 0; q < nsys; ++q) {
    e->sys[sys[q]].step += nsteps;
    e->sys[sys[q]].parity = (int32_t)((e->sys[sys[q]].parity + nsteps) & 1);
  }
}

// nsteps of the listed systems in one batch: hipGraph blocks of n_nl steps
// while the schedule is uniform and aligned, host-driven launches for the
// rest.  With events, ev0 is recorded after the batch's tables are set and
// ev1 after the last launch (device time of the steps).
sop_status engine_run(sop_engine *e, const int32_t *sys, int32_t nsys, int64_t nsteps, hipEvent_t ev0,
                      hipEvent_t ev1) {
  if (nsteps == 0) return SOP_OK;
  sop_status st = engine_set_batch(e, sys, nsys);
  if (st) return st;
  HIPCK(launch_begin(e->p, nsteps, e->stream));
  if (ev0) HIPCK(hipEventRecord(ev0, e->stream));
  // graph blocks need a uniform schedule and phase
  const sop_params &P0 = e->sys[sys[0]].params;
  bool uniform = e->opts.use_graphs != 0 && P0.n_nl % P0.n_pl == 0;
  const int64_t G = P0.n_nl;
  const int64_t r0 = e->sys[sys[0]].step % G;
  for (int q = 1; q < nsys && uniform; ++q) {
    const SysHost &h = e->sys[sys[q]];
    uniform = h.params.n_nl == P0.n_nl && h.params.n_pl == P0.n_pl && h.step % G == r0;
  }
  int64_t left = nsteps;
  if (uniform && G > 1) {
    const int64_t pre = std::min(left, (G - r0) % G);
    if (pre > 0) {
      st = engine_run_hostloop(e, pre);
      if (st) return st;
      mirror_advance(e, sys, nsys, pre);
      left -= pre;
    }
    int64_t blocks = 0;
    while (left >= G) {
      st = engine_run_block_graph(e, (int32_t)G);
      if (st) return st;
      mirror_advance(e, sys, nsys, G);
      left -= G;
      // row capacities checked every GROW_BLOCKS blocks of a long run
      if (++blocks % GROW_BLOCKS == 0 && left >= G && !ev1) {
        std::vector<SysDev> sd;
        HIPCK(hipStreamSynchronize(e->stream));
        st = read_sys_all(e, sd);
        if (st) return st;
        st = engine_grow_rows(e, sys, nsys, sd, nullptr);
        if (st) return st;
      }
    }
  }
  if (left > 0) {
    st = engine_run_hostloop(e, left);
    if (st) return st;
    mirror_advance(e, sys, nsys, left);
  }
  if (ev1) HIPCK(hipEventRecord(ev1, e->stream));
  HIPCK(hipStreamSynchronize(e->stream));
  std::vector<SysDev> sd;
  st = read_sys_all(e, sd);
  if (st) return st;
  st = engine_grow_rows(e, sys, nsys, sd, nullptr);
  if (st) return st;
  for (int32_t q = 0; q < nsys; ++q) {
    st = engine_error_status(e, sys[q], sd[sys[q]].err);
    if (st) return st;
  }
  return SOP_OK;
}

} // namespace sopk

// ================================================================ C ABI

extern "C" {

const char *sop_version(void) { return "sop-mi355x 0.1 (gfx950)"; }

double sop_clock_seconds(void) {
  using namespace std::chrono;
  return duration_cast<duration<double>>(steady_clock::now().time_since_epoch()).count();
}

void sop_engine_opts_default(sop_engine_opts *o) {
  std::memset(o, 0, sizeof(*o));
  o->nl_cap_max = 4096;
  o->pl_cap_max = 1024;
  o->cell_cap_max = 32768;
  o->lanes_per_bead = 4;
  o->grid_blocks = 0;
  o->use_graphs = 1;
  o->lds_beads_max = 5056;  // two 1024-thread step blocks per compute unit at 80,896 B each
  o->pl_shell = 3.0f;
  o->pool_bytes = 0;
  o->step_tail = 2;  // late step items cut to 8 tiles (6.12e9 against 6.02e9 bead-steps/s on 512 replicas)
}

sop_status sop_engine_create(int32_t device, const sop_engine_opts *opts, sop_engine **out) {
  if (!out) return SOP_ERR_INVALID;
  *out = nullptr;
  // the options first (refused before any device call)
  sop_engine_opts eo;
  sop_engine_opts_default(&eo);
  if (opts) {
    eo = *opts;
    sop_engine_opts def;
    sop_engine_opts_default(&def);
    if (eo.nl_cap_max <= 0) eo.nl_cap_max = def.nl_cap_max;
    if (eo.pl_cap_max <= 0) eo.pl_cap_max = def.pl_cap_max;
    if (eo.cell_cap_max <= 0) eo.cell_cap_max = def.cell_cap_max;
    if (eo.lanes_per_bead <= 0) eo.lanes_per_bead = def.lanes_per_bead;
    if (eo.lds_beads_max == 0) eo.lds_beads_max = def.lds_beads_max;
    if (eo.pl_shell == 0.0f) eo.pl_shell = def.pl_shell;
    if (!(eo.pl_shell > 0.0f) || eo.pl_shell > 8.0f) eo.pl_shell = 0.0f;
  }
  const int lpb = eo.lanes_per_bead;
  const int stt = eo.step_tiles;
  // the step kernel holds a bead's pair-list and native pair-list lengths in
  // 16 bits each, and a row's offset within its tile's region in 16 bits
  // (units of ROW_PAD entries): pl_cap_max <= 4095 (natives per bead and per
  // tile are checked at add)
  if ((lpb != 4 && lpb != 8 && lpb != 16 && lpb != 32) ||
      (stt != 0 && stt != 1 && stt != 2 && stt != 4 && stt != 8 && stt != 16) || eo.pl_cap_max > 4095)
    return SOP_ERR_INVALID;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return SOP_ERR_NO_DEVICE;
  if (device < 0 || device >= count) return SOP_ERR_NO_DEVICE;
  (void)hipGetLastError();  // do not inherit an error left by earlier calls
  sop_engine *e = new sop_engine();
  e->opts = eo;
  e->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete e;
    return SOP_ERR_NO_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete e;
    return SOP_ERR_HIP;
  }
  e->num_cu = prop.multiProcessorCount;
  // the allocation-free entry point's scratch on this device (the device is current)
  if (noalloc_scratch(device, &e->scratch) != hipSuccess) {
    (void)hipGetLastError();
    e->scratch = nullptr;
  }
  e->step_blocks = e->opts.step_blocks > 0 ? e->opts.step_blocks : 0;
  e->tune.tail = e->opts.step_tail == 2 || e->opts.step_tail == 4 ? e->opts.step_tail : 0;
  // the step kernel keeps 1 KB of static LDS for the tile's forces
  e->lds_limit = prop.sharedMemPerBlock > 2048 ? prop.sharedMemPerBlock - 2048 : 0;
  if (step_lds_limit(e->lds_limit) != hipSuccess) {
    (void)hipGetLastError();
    e->lds_limit = 65536 - 2048;
  }
  // the shell filter's launches need their LDS sizes: no fallback
  if (filter_lds_init() != hipSuccess) {
    (void)hipGetLastError();
    delete e;
    return SOP_ERR_HIP;
  }
  {
    const int64_t whole = (int64_t)(e->lds_limit / 16) / TILE * TILE;
    e->lds_big_cap = e->opts.lds_beads_max > 0 && whole > e->opts.lds_beads_max ? whole : 0;
  }
  // A blocking stream: synchronous hipMemcpy calls on the null stream order
  // after the engine's kernels, so host reads never race device writes.
  if (hipStreamCreate(&e->stream) != hipSuccess) {
    delete e;
    return SOP_ERR_HIP;
  }
  if (hipStreamCreate(&e->fork.side) != hipSuccess ||
      hipEventCreateWithFlags(&e->fork.fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&e->fork.join, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    sop_engine_destroy(e);
    return SOP_ERR_HIP;
  }
  bind_arrays(e);
  // Lane-layout self test of the DPP / swizzle / ballot helpers.
  int *d_out = nullptr;
  float *d_f = nullptr;
  int h_out[64] = {0};
  hipError_t h1 = hipMalloc((void **)&d_out, 64 * 4);
  hipError_t h2 = h1 == hipSuccess ? hipMalloc((void **)&d_f, 64 * 12) : h1;
  hipError_t h3 = h2 == hipSuccess ? launch_selftest(d_out, d_f, e->stream) : h2;
  hipError_t h3s = h3 == hipSuccess ? hipStreamSynchronize(e->stream) : h3;
  hipError_t h4 = h3s == hipSuccess ? hipMemcpy(h_out, d_out, 64 * 4, hipMemcpyDeviceToHost) : h3s;
  bool ok = h4 == hipSuccess;
  for (int k = 0; ok && k < 64; ++k) ok = h_out[k] == 0;
  if (d_out) (void)hipFree(d_out);
  if (d_f) (void)hipFree(d_f);
  if (!ok) {
    std::fprintf(stderr, "sop_engine_create: self test failed (%s / %s / %s / %s, lane0 bits %d)\n",
                 hipGetErrorString(h1), hipGetErrorString(h2), hipGetErrorString(h3), hipGetErrorString(h4),
                 h_out[0]);
    sop_engine_destroy(e);
    return SOP_ERR_HIP;
  }
  if (e->opts.pool_bytes > 0) {
    // fixed pools sized from the budget (serving): split by typical use (per
    // bead ~200 B + 2 * pl_cap entries of 2 B; the pair-list and shell-row
    // segments dominate); the neighbor-list scratch grows on request
    const int64_t B = e->opts.pool_bytes;
    const int64_t beads = std::max<int64_t>(4096, B / 8 / 200) / TILE * TILE;
    const int64_t nat = std::max<int64_t>(4096, B / 64 / 24);
    const int64_t pl = std::max<int64_t>(65536, B * 3 / 4 / 8);
    const int64_t cells = std::max<int64_t>(65536, B / 64 / 12);
    sop_status st = ensure_capacity(e, beads, nat, 0, pl, cells, 4096);
    if (st) {
      sop_engine_destroy(e);
      return st;
    }
    e->fixed_pools = true;
  }
  *out = e;
  return SOP_OK;
}

void sop_engine_destroy(sop_engine *e) {
  if (!e) return;
  if (e->serve) sop_serve_stop(e);
  hipError_t first = hipSuccess;
  const char *what = "";
  auto note = [&](hipError_t h, const char *w) {
    if (h != hipSuccess && first == hipSuccess) { first = h; what = w; }
  };
  note(hipSetDevice(e->device), "hipSetDevice");
  if (e->stream) note(hipStreamSynchronize(e->stream), "hipStreamSynchronize");
  for (auto &kv : e->graphs) note(hipGraphExecDestroy(kv.second), "hipGraphExecDestroy");
  for (auto &a : e->bead_arrays) if (*a.first) note(hipFree(*a.first), "hipFree(bead pool)");
  for (auto &a : e->nat_arrays) if (*a.first) note(hipFree(*a.first), "hipFree(native pool)");
  for (auto &a : e->cell_arrays) if (*a.first) note(hipFree(*a.first), "hipFree(cell pool)");
  void *others[] = {e->p.nlrep_j, e->p.plrep_j, e->p.slrep_j,  e->p.tile_sys, e->p.sys,  e->p.sysout,
                    e->p.syslist, e->p.work,    e->p.nwork,    e->p.cellwork, e->p.items};
  for (void *q : others) if (q) note(hipFree(q), "hipFree");
  if (e->fork.side) note(hipStreamSynchronize(e->fork.side), "hipStreamSynchronize");
  if (e->fork.fork) note(hipEventDestroy(e->fork.fork), "hipEventDestroy");
  if (e->fork.join) note(hipEventDestroy(e->fork.join), "hipEventDestroy");
  if (e->fork.side) note(hipStreamDestroy(e->fork.side), "hipStreamDestroy");
  if (e->stream) note(hipStreamDestroy(e->stream), "hipStreamDestroy");
  if (first != hipSuccess) std::fprintf(stderr, "sop_engine_destroy: %s: %s\n", what, hipGetErrorString(first));
  (void)hipGetLastError();
  delete e;
}

const char *sop_last_error(const sop_engine *e) { return e ? e->err.c_str() : "no engine"; }

sop_status sop_synchronize(sop_engine *e) {
  if (!e) return SOP_ERR_INVALID;
  return hip_check(e, hipStreamSynchronize(e->stream), "hipStreamSynchronize");
}

static bool valid_sys(sop_engine *e, int32_t s) {
  return e && s >= 0 && s < (int32_t)e->sys.size() && e->sys[s].used;
}

sop_status sop_system_add(sop_engine *e, const sop_system_desc *d, int32_t *sys) {
  if (!e || !sys) return SOP_ERR_INVALID;
  if (e->serve) return set_error(e, SOP_ERR_STATE, "engine is serving");
  return engine_add_system(e, d, sys);
}

sop_status sop_system_remove(sop_engine *e, int32_t sys) {
  if (!e) return SOP_ERR_INVALID;
  HIPCK(hipStreamSynchronize(e->stream));
  return engine_remove_system(e, sys);
}

sop_status sop_system_info(sop_engine *e, int32_t sys, int32_t *n_beads, int64_t *step) {
  if (!valid_sys(e, sys)) return set_error(e, SOP_ERR_INVALID, "no such system");
  if (n_beads) *n_beads = e->sys[sys].n;
  if (step) *step = e->sys[sys].step;
  return SOP_OK;
}

// The neighbor list is written only on request, into one scratch segment
// (grown to fit) that holds the rows of the last system materialized.
static sop_status nl_scratch_for(sop_engine *e, int32_t sys) {
  sop_status st = ensure_capacity(e, 0, 0, e->sys[sys].nl_len, 0, 0, 0);
  if (st) return st;
  if (e->nl_owner != sys) {  // the scratch holds another system's rows
    const int32_t zero = 0;
    HIPCK(hipMemcpy(&e->p.sys[sys].nl_valid, &zero, sizeof(zero), hipMemcpyHostToDevice));
  }
  e->nl_owner = sys;
  return SOP_OK;
}

sop_status sop_update_neighbor_list(sop_engine *e, int32_t sys) {
  if (!valid_sys(e, sys)) return set_error(e, SOP_ERR_INVALID, "no such system");
  sop_status st = nl_scratch_for(e, sys);
  if (st) return st;
  st = engine_set_batch(e, &sys, 1);
  if (st) return st;
  HIPCK(launch_lists(e->p, list_grids(e), 0, FORCE_NL, e->stream));
  HIPCK(hipStreamSynchronize(e->stream));
  return engine_check_errors(e, &sys, 1);
}

sop_status sop_update_pair_list(sop_engine *e, int32_t sys) {
  if (!valid_sys(e, sys)) return set_error(e, SOP_ERR_INVALID, "no such system");
  sop_status st = engine_set_batch(e, &sys, 1);
  if (st) return st;
  HIPCK(launch_lists(e->p, list_grids(e), 0, FORCE_PL, e->stream));
  HIPCK(hipStreamSynchronize(e->stream));
  return engine_check_errors(e, &sys, 1);
}

// The neighbor list is materialized on demand from its snapshot.
static sop_status materialize_nl(sop_engine *e, int32_t sys) {
  sop_status st = nl_scratch_for(e, sys);
  if (st) return st;
  st = engine_set_batch(e, &sys, 1);
  if (st) return st;
  HIPCK(launch_lists(e->p, list_grids(e), 0, MATERIALIZE_NL, e->stream));
  HIPCK(hipStreamSynchronize(e->stream));
  return engine_check_errors(e, &sys, 1);
}

sop_status sop_get_list(sop_engine *e, int32_t sys, int32_t which, int32_t *ij, int64_t cap, int64_t *n) {
  if (!valid_sys(e, sys) || !n) return set_error(e, SOP_ERR_INVALID, "bad arguments");
  HIPCK(hipStreamSynchronize(e->stream));
  if (which == SOP_LIST_NL_NATIVE || which == SOP_LIST_NL_REP) {
    sop_status st = materialize_nl(e, sys);
    if (st) return st;
  }
  const SysHost &h = e->sys[sys];
  const int32_t nb = h.n;
  std::vector<int32_t> cnt((size_t)nb);
  std::vector<int32_t> ent;
  std::vector<int32_t> start((size_t)nb);
  int32_t *cnt_d = nullptr;
  int64_t seg_len = 0;
  if (which == SOP_LIST_NL_NATIVE || which == SOP_LIST_PL_NATIVE) {
    cnt_d = which == SOP_LIST_NL_NATIVE ? e->p.nlatt_cnt : e->p.platt_cnt;
    // the neighbor list's natives sit in each bead's native slots; the native
    // pair list is packed per tile (platt_off)
    HIPCK(hipMemcpy(start.data(), (which == SOP_LIST_NL_NATIVE ? e->p.nat_start : e->p.platt_off) + h.bead_off,
                    4 * (size_t)nb, hipMemcpyDeviceToHost));
    ent.resize((size_t)h.nat_len);
    if (which == SOP_LIST_NL_NATIVE) {
      HIPCK(hipMemcpy(ent.data(), e->p.nlatt_j + h.nat_off, 4 * (size_t)h.nat_len, hipMemcpyDeviceToHost));
    } else {
      std::vector<int2> pe((size_t)h.nat_len);
      HIPCK(hipMemcpy(pe.data(), e->p.platt + h.nat_off, 8 * (size_t)h.nat_len, hipMemcpyDeviceToHost));
      for (size_t k = 0; k < pe.size(); ++k) ent[k] = pe[k].x;
    }
    for (int i = 0; i < nb; ++i) start[i] -= (int32_t)h.nat_off;
  } else if (which == SOP_LIST_NL_REP || which == SOP_LIST_PL_REP) {
    const bool nl = which == SOP_LIST_NL_REP;
    cnt_d = nl ? e->p.nlrep_cnt : e->p.plrep_cnt;
    seg_len = nl ? h.nl_len : (int64_t)h.n * h.pl_cap;  // entries
    ent.resize((size_t)seg_len);
    if (!nl && h.n <= 65536) {
      // 16-bit local indices from the segment start
      std::vector<uint16_t> e16((size_t)seg_len);
      HIPCK(hipMemcpy(e16.data(), e->p.plrep_j + h.pl_off, 2 * (size_t)seg_len, hipMemcpyDeviceToHost));
      for (int64_t q = 0; q < seg_len; ++q) ent[(size_t)q] = (int32_t)h.bead_off + e16[(size_t)q];
    } else {
      HIPCK(hipMemcpy(ent.data(), (nl ? e->p.nlrep_j + h.nl_off : e->p.plrep_j + h.pl_off), 4 * (size_t)seg_len,
                      hipMemcpyDeviceToHost));
    }
    if (nl) {
      for (int i = 0; i < nb; ++i) start[i] = i * h.nl_cap;
    } else {  // pair-list rows are packed per tile: each bead's row start is stored
      HIPCK(hipMemcpy(start.data(), e->p.plrep_off + h.bead_off, 4 * (size_t)nb, hipMemcpyDeviceToHost));
    }
  } else {
    return set_error(e, SOP_ERR_INVALID, "unknown list");
  }
  HIPCK(hipMemcpy(cnt.data(), cnt_d + h.bead_off, 4 * (size_t)nb, hipMemcpyDeviceToHost));
  int64_t m = 0;
  for (int i = 0; i < nb; ++i)
    for (int k = 0; k < cnt[i]; ++k) {
      const int32_t j = ent[(size_t)start[i] + k] - (int32_t)h.bead_off;
      if (j > i) {
        if (ij && m < cap) {
          ij[2 * m] = i;
          ij[2 * m + 1] = j;
        }
        ++m;
      }
    }
  *n = m;
  return SOP_OK;
}

sop_status sop_system_rows(sop_engine *e, int32_t sys, int32_t *capacity, int32_t *longest) {
  if (!valid_sys(e, sys)) return set_error(e, SOP_ERR_INVALID, "no such system");
  HIPCK(hipStreamSynchronize(e->stream));
  SysDev S;
  const sop_status st = engine_read_sys(e, sys, &S);
  if (st) return st;
  if (capacity) *capacity = e->sys[sys].pl_cap;
  if (longest) *longest = S.row_max;
  return SOP_OK;
}

sop_status sop_system_reserve_rows(sop_engine *e, int32_t sys, int32_t capacity) {
  if (!valid_sys(e, sys) || capacity < 1) return set_error(e, SOP_ERR_INVALID, "bad arguments");
  if (e->serve) return set_error(e, SOP_ERR_STATE, "engine is serving");
  HIPCK(hipStreamSynchronize(e->stream));
  SysDev S;
  const sop_status st = engine_read_sys(e, sys, &S);
  if (st) return st;
  const SysHost &h = e->sys[sys];
  const int32_t cap = std::min(pl_cap_limit(e, h.n), std::max(round_up(capacity, 16), round_up(S.row_max, 16)));
  return engine_resize_rows(e, sys, cap, true);
}

sop_status sop_list_stats_get(sop_engine *e, int32_t sys, sop_list_stats *stt) {
  if (!valid_sys(e, sys) || !stt) return set_error(e, SOP_ERR_INVALID, "bad arguments");
  HIPCK(hipStreamSynchronize(e->stream));
  sop_status st = materialize_nl(e, sys);
  if (st) return st;
  const SysHost &h = e->sys[sys];
  std::vector<int32_t> c((size_t)h.n);
  int64_t tot[4];
  int32_t mx[4];
  int32_t *arrs[4] = {e->p.nlatt_cnt, e->p.nlrep_cnt, e->p.platt_cnt, e->p.plrep_cnt};
  for (int q = 0; q < 4; ++q) {
    HIPCK(hipMemcpy(c.data(), arrs[q] + h.bead_off, 4 * (size_t)h.n, hipMemcpyDeviceToHost));
    tot[q] = 0;
    mx[q] = 0;
    for (int32_t v : c) { tot[q] += v; mx[q] = std::max(mx[q], v); }
  }
  stt->nl_native = tot[0] / 2;
  stt->nl_rep = tot[1] / 2;
  stt->pl_native = tot[2] / 2;
  stt->pl_rep = tot[3] / 2;
  stt->nl_max_per_bead = mx[1];
  stt->pl_max_per_bead = mx[3];
  SysDev S;
  st = engine_read_sys(e, sys, &S);
  if (st) return st;
  stt->grid_builds = S.nbuild[0];
  stt->shell_builds = S.nbuild[1];
  stt->nl_builds = S.nbuild[2];
  return SOP_OK;
}

sop_status sop_force_eval(sop_engine *e, int32_t sys, int32_t with_noise) {
  if (!valid_sys(e, sys)) return set_error(e, SOP_ERR_INVALID, "no such system");
  sop_status st = engine_set_batch(e, &sys, 1);
  if (st) return st;
  HIPCK(launch_step(e->p, with_noise ? MODE_FORCE_NOISE : MODE_FORCE, e->opts.lanes_per_bead, step_grids(e), 0,
                    e->stream));
  HIPCK(hipStreamSynchronize(e->stream));
  return engine_check_errors(e, &sys, 1);
}

sop_status sop_get_forces(sop_engine *e, int32_t sys, float *f3n) {
  if (!valid_sys(e, sys) || !f3n) return set_error(e, SOP_ERR_INVALID, "bad arguments");
  HIPCK(hipStreamSynchronize(e->stream));
  const SysHost &h = e->sys[sys];
  std::vector<float4> f((size_t)h.n);
  HIPCK(hipMemcpy(f.data(), e->p.fout + h.bead_off, 16 * (size_t)h.n, hipMemcpyDeviceToHost));
  for (int i = 0; i < h.n; ++i) {
    f3n[3 * i] = f[i].x;
    f3n[3 * i + 1] = f[i].y;
    f3n[3 * i + 2] = f[i].z;
  }
  return SOP_OK;
}

static sop_status eval_energy(sop_engine *e, int32_t sys, double *out) {
  sop_status st = engine_set_batch(e, &sys, 1);
  if (st) return st;
  HIPCK(launch_energy(e->p, grid_tiles(e), grid_sys(e), e->stream));
  HIPCK(hipMemcpyAsync(out, e->p.sysout + (int64_t)sys * SYSOUT_N, 8 * SYSOUT_N, hipMemcpyDeviceToHost, e->stream));
  HIPCK(hipStreamSynchronize(e->stream));
  return SOP_OK;
}

sop_status sop_energy_eval(sop_engine *e, int32_t sys, double *energies) {
  if (!valid_sys(e, sys) || !energies) return set_error(e, SOP_ERR_INVALID, "bad arguments");
  double o[SYSOUT_N];
  sop_status st = eval_energy(e, sys, o);
  if (st) return st;
  for (int k = 0; k < SOP_E_COUNT; ++k) energies[k] = o[k];
  return SOP_OK;
}

sop_status sop_energy_eval_noalloc(sop_engine *e, int32_t sys, double *energies) {
  if (!valid_sys(e, sys) || !energies) return set_error(e, SOP_ERR_INVALID, "bad arguments");
  if (!e->scratch) return set_error(e, SOP_ERR_HIP, "sop_energy_eval_noalloc: no scratch address on this device");
  const SysHost &h = e->sys[sys];
  if (h.n > NOALLOC_BEADS) {
    std::ostringstream os;
    os << "sop_energy_eval_noalloc: system " << sys << " has " << h.n << " beads, above the scratch limit of "
       << NOALLOC_BEADS;
    return set_error(e, SOP_ERR_CAPACITY, os.str());
  }
  // the engine's arrays, with the work table, counters and partials in scratch
  DevPtrs q = e->p;
  char *b = static_cast<char *>(e->scratch);
  q.work = reinterpret_cast<int32_t *>(b + NOALLOC_OFF_WORK);
  q.nwork = reinterpret_cast<int32_t *>(b + NOALLOC_OFF_CNT);
  q.nsys = q.nwork + 16;
  q.syslist = reinterpret_cast<int32_t *>(b + NOALLOC_OFF_SYSLIST);
  q.epart = reinterpret_cast<double *>(b + NOALLOC_OFF_EPART);
  q.epart_base = (int32_t)h.bead_off;
  q.sysout = reinterpret_cast<double *>(b + NOALLOC_OFF_SYSOUT);
  q.sysout_base = sys;
  const int t0 = (int)(h.bead_off / TILE), nt = (int)(h.bead_len / TILE);
  const int grid = std::max(8, std::min(8 * e->num_cu, (nt + 7) / 8 * 8));
  HIPCK(launch_noalloc_energy(q, t0, nt, sys, grid, e->stream));
  double o[SYSOUT_N];
  HIPCK(hipMemcpyAsync(o, q.sysout, sizeof(o), hipMemcpyDeviceToHost, e->stream));
  HIPCK(hipStreamSynchronize(e->stream));
  for (int k = 0; k < SOP_E_COUNT; ++k) energies[k] = o[k];
  return SOP_OK;
}

sop_status sop_observables(sop_engine *e, int32_t sys, double *Q, int32_t *contacts, double *rgsq) {
  if (!valid_sys(e, sys)) return set_error(e, SOP_ERR_INVALID, "no such system");
  double o[SYSOUT_N];
  sop_status st = eval_energy(e, sys, o);
  if (st) return st;
  const int32_t c = (int32_t)std::llround(o[6]);
  if (contacts) *contacts = c;
  if (Q) *Q = e->sys[sys].n_nat > 0 ? (double)c / (double)e->sys[sys].n_nat : 0.0;
  if (rgsq) *rgsq = o[7];
  return SOP_OK;
}

static sop_status check_run_args(sop_engine *e, const int32_t *sys, int32_t nsys, int64_t nsteps) {
  if (!e || !sys || nsys <= 0 || nsteps < 0) return set_error(e, SOP_ERR_INVALID, "bad run arguments");
  if (e->serve) return set_error(e, SOP_ERR_STATE, "engine is serving");
  for (int q = 0; q < nsys; ++q) {
    if (!valid_sys(e, sys[q])) return set_error(e, SOP_ERR_INVALID, "no such system");
    for (int r = 0; r < q; ++r)
      if (sys[r] == sys[q]) return set_error(e, SOP_ERR_INVALID, "system listed twice");
  }
  return SOP_OK;
}

sop_status sop_run(sop_engine *e, const int32_t *sys, int32_t nsys, int64_t nsteps) {
  sop_status st = check_run_args(e, sys, nsys, nsteps);
  if (st) return st;
  return engine_run(e, sys, nsys, nsteps, nullptr, nullptr);
}

sop_status sop_run_plain(sop_engine *e, const int32_t *sys, int32_t nsys, int64_t nsteps) {
  sop_status st = check_run_args(e, sys, nsys, nsteps);
  if (st) return st;
  for (int q = 0; q < nsys; ++q) {
    st = engine_set_batch(e, &sys[q], 1);
    if (st) return st;
    HIPCK(launch_begin(e->p, nsteps, e->stream));
    st = engine_run_hostloop(e, nsteps);
    if (st) return st;
    mirror_advance(e, &sys[q], 1, nsteps);
  }
  HIPCK(hipStreamSynchronize(e->stream));
  return engine_check_errors(e, sys, nsys);
}

static sop_status read_coords(sop_engine *e, int32_t sys, std::vector<float4> &x) {
  SysDev S;
  sop_status st = engine_read_sys(e, sys, &S);
  if (st) return st;
  const SysHost &h = e->sys[sys];
  x.resize((size_t)h.n);
  HIPCK(hipMemcpy(x.data(), (S.parity & 1 ? e->p.pos1 : e->p.pos0) + h.bead_off, 16 * (size_t)h.n,
                  hipMemcpyDeviceToHost));
  return SOP_OK;
}

sop_status sop_get_coords(sop_engine *e, int32_t sys, float *xyz, int64_t *step) {
  if (!valid_sys(e, sys)) return set_error(e, SOP_ERR_INVALID, "no such system");
  std::vector<float4> x;
  sop_status st = read_coords(e, sys, x);
  if (st) return st;
  if (xyz)
    for (size_t i = 0; i < x.size(); ++i) {
      xyz[3 * i] = x[i].x;
      xyz[3 * i + 1] = x[i].y;
      xyz[3 * i + 2] = x[i].z;
    }
  if (step) {
    SysDev S;
    st = engine_read_sys(e, sys, &S);
    if (st) return st;
    *step = S.step;
  }
  return SOP_OK;
}

sop_status sop_get_wrapped_coords(sop_engine *e, int32_t sys, float *xyz) {
  if (!valid_sys(e, sys) || !xyz) return set_error(e, SOP_ERR_INVALID, "bad arguments");
  std::vector<float4> x;
  sop_status st = read_coords(e, sys, x);
  if (st) return st;
  const float boxl = e->sys[sys].params.boxl;
  for (size_t i = 0; i < x.size(); ++i) {
    float v[3] = {x[i].x, x[i].y, x[i].z};
    for (int d = 0; d < 3; ++d) {
      if (boxl > 0.0f) v[d] = v[d] - boxl * rintf(v[d] / boxl);
      xyz[3 * i + d] = v[d];
    }
  }
  return SOP_OK;
}

// Forces with noise of the current step (kstep 0) or of the last step taken
// (kstep -1: the other coordinate buffer and step - 1) into fout, copied out.
static sop_status noisy_forces(sop_engine *e, int32_t sys, int kstep, std::vector<float4> &f) {
  sop_status st = engine_set_batch(e, &sys, 1);
  if (st) return st;
  HIPCK(launch_step(e->p, MODE_FORCE_NOISE, e->opts.lanes_per_bead, step_grids(e), kstep, e->stream));
  HIPCK(hipStreamSynchronize(e->stream));
  st = engine_check_errors(e, &sys, 1);
  if (st) return st;
  const SysHost &h = e->sys[sys];
  f.resize((size_t)h.n);
  HIPCK(hipMemcpy(f.data(), e->p.fout + h.bead_off, 16 * (size_t)h.n, hipMemcpyDeviceToHost));
  return SOP_OK;
}

static sop_status read_vel_state(sop_engine *e, int32_t sys, std::vector<float4> &v) {
  HIPCK(hipStreamSynchronize(e->stream));
  const SysHost &h = e->sys[sys];
  v.resize((size_t)h.n);
  HIPCK(hipMemcpy(v.data(), e->p.vel + h.bead_off, 16 * (size_t)h.n, hipMemcpyDeviceToHost));
  return SOP_OK;
}

static bool incr_flag(const float4 &s) {
  int32_t w;
  std::memcpy(&w, &s.w, 4);
  return w == 1;
}

sop_status sop_get_velocities(sop_engine *e, int32_t sys, float *v3n) {
  if (!valid_sys(e, sys) || !v3n) return set_error(e, SOP_ERR_INVALID, "bad arguments");
  const SysHost &h = e->sys[sys];
  if (h.params.dynamics != SOP_DYN_UNDERDAMPED) {
    std::fill(v3n, v3n + 3 * (size_t)h.n, 0.0f);
    return SOP_OK;
  }
  std::vector<float4> v, f;
  sop_status st = read_vel_state(e, sys, v);
  if (st) return st;
  bool any = false;
  for (const auto &s : v) any |= incr_flag(s);
  if (any && (st = noisy_forces(e, sys, 0, f))) return st;
  float a[4];
  sop_params_underdamped(&h.params, a);
  for (int i = 0; i < h.n; ++i) {
    const float4 s = v[i];
    if (incr_flag(s)) {
      v3n[3 * i] = a[2] * s.x + a[3] * f[i].x;
      v3n[3 * i + 1] = a[2] * s.y + a[3] * f[i].y;
      v3n[3 * i + 2] = a[2] * s.z + a[3] * f[i].z;
    } else {
      v3n[3 * i] = s.x;
      v3n[3 * i + 1] = s.y;
      v3n[3 * i + 2] = s.z;
    }
  }
  return SOP_OK;
}

sop_status sop_get_increments(sop_engine *e, int32_t sys, float *d3n) {
  if (!valid_sys(e, sys) || !d3n) return set_error(e, SOP_ERR_INVALID, "bad arguments");
  const SysHost &h = e->sys[sys];
  std::vector<float4> v, f;
  sop_status st;
  if (h.params.dynamics == SOP_DYN_UNDERDAMPED) {
    if ((st = read_vel_state(e, sys, v))) return st;
    bool all = true;
    for (const auto &s : v) all &= incr_flag(s);
    if (!all && (st = noisy_forces(e, sys, -1, f))) return st;
    float a[4];
    sop_params_underdamped(&h.params, a);
    for (int i = 0; i < h.n; ++i) {
      const float4 s = v[i];
      if (incr_flag(s)) {
        d3n[3 * i] = s.x;
        d3n[3 * i + 1] = s.y;
        d3n[3 * i + 2] = s.z;
      } else {
        d3n[3 * i] = a[0] * s.x + a[1] * f[i].x;
        d3n[3 * i + 1] = a[0] * s.y + a[1] * f[i].y;
        d3n[3 * i + 2] = a[0] * s.z + a[1] * f[i].z;
      }
    }
    return SOP_OK;
  }
  if ((st = noisy_forces(e, sys, -1, f))) return st;
  float a5, var;
  sop_params_derived(&h.params, &a5, &var);
  for (int i = 0; i < h.n; ++i) {
    d3n[3 * i] = a5 * f[i].x;
    d3n[3 * i + 1] = a5 * f[i].y;
    d3n[3 * i + 2] = a5 * f[i].z;
  }
  return SOP_OK;
}

sop_status sop_set_rng(sop_engine *e, int32_t sys, uint64_t seed, int64_t step) {
  if (!valid_sys(e, sys)) return set_error(e, SOP_ERR_INVALID, "no such system");
  if (step < 0) return set_error(e, SOP_ERR_INVALID, "step must be >= 0");
  SysDev S;
  sop_status st = engine_read_sys(e, sys, &S);
  if (st) return st;
  S.seed_lo = (uint32_t)seed;
  S.seed_hi = (uint32_t)(seed >> 32);
  S.step = step;
  S.step_end = step;
  HIPCK(hipMemcpy(e->p.sys + sys, &S, sizeof(SysDev), hipMemcpyHostToDevice));
  SysHost &h = e->sys[sys];
  h.params.seed = seed;
  h.step = step;
  return SOP_OK;
}

sop_status sop_set_coords(sop_engine *e, int32_t sys, const float *xyz) {
  if (!valid_sys(e, sys) || !xyz) return set_error(e, SOP_ERR_INVALID, "bad arguments");
  SysDev S;
  sop_status st = engine_read_sys(e, sys, &S);
  if (st) return st;
  const SysHost &h = e->sys[sys];
  std::vector<float4> x((size_t)h.n);
  for (int i = 0; i < h.n; ++i) {
    float wf;
    std::memcpy(&wf, &h.wbits[i], 4);
    x[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], wf);
  }
  HIPCK(hipMemcpy((S.parity & 1 ? e->p.pos1 : e->p.pos0) + h.bead_off, x.data(), 16 * (size_t)h.n,
                  hipMemcpyHostToDevice));
  // new coordinates clear the system's sticky error bits
  const int32_t zero = 0;
  HIPCK(hipMemcpy((char *)(e->p.sys + sys) + offsetof(SysDev, err), &zero, 4, hipMemcpyHostToDevice));
  return SOP_OK;
}

sop_status sop_time_step_kernel(sop_engine *e, const int32_t *sys, int32_t nsys, int32_t reps, double *avg_ms) {
  sop_status st = check_run_args(e, sys, nsys, reps);
  if (st) return st;
  if (reps <= 0 || !avg_ms) return set_error(e, SOP_ERR_INVALID, "reps > 0 and avg_ms required");
  st = engine_set_batch(e, sys, nsys);
  if (st) return st;
  // Timing leaves no trace: coordinates and system descriptors are restored.
  const size_t pos_bytes = 16 * (size_t)e->cap.beads, sys_bytes = sizeof(SysDev) * (size_t)e->cap.slots;
  void *save = nullptr;
  HIPCK(hipMalloc(&save, 2 * pos_bytes + sys_bytes));
  char *sv = (char *)save;
  HIPCK(hipMemcpyAsync(sv, e->p.pos0, pos_bytes, hipMemcpyDeviceToDevice, e->stream));
  HIPCK(hipMemcpyAsync(sv + pos_bytes, e->p.pos1, pos_bytes, hipMemcpyDeviceToDevice, e->stream));
  HIPCK(hipMemcpyAsync(sv + 2 * pos_bytes, e->p.sys, sys_bytes, hipMemcpyDeviceToDevice, e->stream));
  HIPCK(launch_begin(e->p, reps, e->stream));
  hipEvent_t a, b;
  HIPCK(hipEventCreate(&a));
  HIPCK(hipEventCreate(&b));
  const int lpb = e->opts.lanes_per_bead;
  HIPCK(hipEventRecord(a, e->stream));
  for (int r = 0; r < reps; ++r) HIPCK(launch_step(e->p, MODE_STEP, lpb, step_grids(e), r, e->stream));
  HIPCK(hipEventRecord(b, e->stream));
  HIPCK(hipEventSynchronize(b));
  float ms = 0.0f;
  HIPCK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  HIPCK(hipMemcpyAsync(e->p.pos0, sv, pos_bytes, hipMemcpyDeviceToDevice, e->stream));
  HIPCK(hipMemcpyAsync(e->p.pos1, sv + pos_bytes, pos_bytes, hipMemcpyDeviceToDevice, e->stream));
  HIPCK(hipMemcpyAsync(e->p.sys, sv + 2 * pos_bytes, sys_bytes, hipMemcpyDeviceToDevice, e->stream));
  HIPCK(hipStreamSynchronize(e->stream));
  HIPCK(hipFree(save));
  *avg_ms = ms / reps;
  return SOP_OK;
}

// Device self test for the tests: per-lane failure bits of the wave helpers
// and 64 x 3 device normals for (seed 0x9abcdef012345678, bead = lane,
// step = 1000 + lane).
sop_status sop_selftest(sop_engine *e, int32_t *bad64, float *normals192) {
  if (!e) return SOP_ERR_INVALID;
  int *d_out = nullptr;
  float *d_f = nullptr;
  HIPCK(hipMalloc((void **)&d_out, 64 * 4));
  HIPCK(hipMalloc((void **)&d_f, 64 * 12));
  HIPCK(launch_selftest(d_out, d_f, e->stream));
  HIPCK(hipStreamSynchronize(e->stream));
  HIPCK(hipMemcpy(bad64, d_out, 64 * 4, hipMemcpyDeviceToHost));
  HIPCK(hipMemcpy(normals192, d_f, 64 * 12, hipMemcpyDeviceToHost));
  (void)hipFree(d_out);
  (void)hipFree(d_f);
  return SOP_OK;
}

} // extern "C"
