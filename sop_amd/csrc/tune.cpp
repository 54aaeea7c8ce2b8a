// tune.cpp -- sop_autotune: launch shapes and list settings of a batch picked
// by timing whole hipGraph blocks of its own steps (n_nl steps each, list
// rebuilds included).  Every trial starts from a device snapshot of the
// engine, and the snapshot is restored at the end, so tuning leaves the
// systems' coordinates, lists and counters exactly as it found them; only
// the chosen settings remain.
#include <algorithm>
#include <cstring>

#include "engine.h"

using namespace sopk;

#define HIPCK_TUNE(expr)                                         \
  do {                                                         \
    const sop_status _s = hip_check(e, (expr), #expr);           \
    if (_s != SOP_OK) return _s;                               \
  } while (0)

namespace {

// Every device array the list and step kernels write, whole, plus the host
// mirrors of the systems.
struct Snapshot {
  char *buf = nullptr;
  std::vector<std::pair<void *, size_t>> parts;
  std::vector<SysHost> host;
};

void snapshot_parts(sop_engine *e, std::vector<std::pair<void *, size_t>> &parts) {
  const Pools &c = e->cap;
  for (auto &a : e->bead_arrays) parts.push_back({*a.first, a.second * (size_t)c.beads});
  for (auto &a : e->nat_arrays) parts.push_back({*a.first, a.second * (size_t)c.nat});
  for (auto &a : e->cell_arrays) parts.push_back({*a.first, a.second * (size_t)c.cells});
  parts.push_back({e->p.nlrep_j, 4 * (size_t)c.nl});
  parts.push_back({e->p.plrep_j, 4 * (size_t)c.pl});
  parts.push_back({e->p.slrep_j, 4 * (size_t)c.pl});
  parts.push_back({e->p.sys, sizeof(SysDev) * (size_t)c.slots});
  parts.push_back({e->p.sysout, 8 * SYSOUT_N * (size_t)c.slots});
}

sop_status snapshot_take(sop_engine *e, Snapshot &s) {
  snapshot_parts(e, s.parts);
  size_t total = 0;
  for (auto &pt : s.parts) total += (pt.second + 255) / 256 * 256;
  sop_status st = hip_check(e, hipStreamSynchronize(e->stream), "autotune: synchronize");
  if (st) return st;
  if (hipMalloc((void **)&s.buf, std::max<size_t>(total, 256)) != hipSuccess) {
    (void)hipGetLastError();
    s.buf = nullptr;
    return set_error(e, SOP_ERR_CAPACITY, "autotune: no device memory for a snapshot of the engine");
  }
  size_t off = 0;
  for (auto &pt : s.parts) {
    if (pt.first && pt.second)
      if ((st = hip_check(e, hipMemcpyAsync(s.buf + off, pt.first, pt.second, hipMemcpyDeviceToDevice, e->stream),
                          "autotune: snapshot")))
        return st;
    off += (pt.second + 255) / 256 * 256;
  }
  s.host = e->sys;
  return hip_check(e, hipStreamSynchronize(e->stream), "autotune: snapshot");
}

sop_status snapshot_restore(sop_engine *e, const Snapshot &s) {
  size_t off = 0;
  sop_status st;
  for (auto &pt : s.parts) {
    if (pt.first && pt.second)
      if ((st = hip_check(e, hipMemcpyAsync(pt.first, s.buf + off, pt.second, hipMemcpyDeviceToDevice, e->stream),
                          "autotune: restore")))
        return st;
    off += (pt.second + 255) / 256 * 256;
  }
  e->sys = s.host;
  return hip_check(e, hipStreamSynchronize(e->stream), "autotune: restore");
}

struct Setting {
  int32_t lanes = 4, blocks = 0, lds_beads = 5056;
  sop_engine::Tune tune;
};

// The engine as `st` describes it: launch options, the size buckets' shell
// widths and item sizes.  A system whose shell changes gets its shell fields
// rewritten and its shell rows marked stale (the next pair-list rebuild
// makes new ones from the cell grid).
sop_status apply(sop_engine *e, const int32_t *sys, int32_t nsys, const Setting &st) {
  e->opts.lanes_per_bead = st.lanes;
  e->step_blocks = st.blocks;
  e->opts.lds_beads_max = st.lds_beads;
  e->tune = st.tune;
  std::vector<int32_t> changed;
  for (int32_t q = 0; q < nsys; ++q) {
    SysHost &h = e->sys[sys[q]];
    const float w = engine_shell_for(e, h.n);
    if (w != h.shell) {
      if (e->grow_rows) {  // room for the wider shell's rows (not inside timed trials)
        const sop_status rs = engine_fit_shell(e, sys[q], w);
        if (rs) return rs;
      }
      h.shell = w;
      changed.push_back(sys[q]);
    }
  }
  if (changed.empty()) return SOP_OK;
  std::vector<SysDev> sd((size_t)e->cap.slots);
  sop_status rc = hip_check(e, hipMemcpy(sd.data(), e->p.sys, sizeof(SysDev) * sd.size(), hipMemcpyDeviceToHost),
                            "autotune: read systems");
  if (rc) return rc;
  for (int32_t s : changed) shell_fields(e->sys[s].params, e->sys[s].shell, sd[s]);
  return hip_check(e, hipMemcpy(e->p.sys, sd.data(), sizeof(SysDev) * sd.size(), hipMemcpyHostToDevice),
                   "autotune: write systems");
}

// Device time of `blocks` graph blocks after one warm-up block, from the
// snapshot, under setting st.
sop_status trial(sop_engine *e, const int32_t *sys, int32_t nsys, const Snapshot &snap, const Setting &st,
                 int64_t G, int blocks, hipEvent_t ev0, hipEvent_t ev1, double *ms) {
  sop_status rc = snapshot_restore(e, snap);
  if (!rc) rc = apply(e, sys, nsys, st);
  if (!rc) rc = engine_run(e, sys, nsys, G, nullptr, nullptr);
  if (!rc) rc = engine_run(e, sys, nsys, G * blocks, ev0, ev1);
  if (rc) return rc;
  float t = 0.0f;
  rc = hip_check(e, hipEventElapsedTime(&t, ev0, ev1), "autotune: timing");
  *ms = t;
  return rc;
}

}  // namespace

extern "C" {

sop_status sop_autotune_ex(sop_engine *e, const int32_t *sys, int32_t nsys, sop_tune_result *out) {
  if (!e || !sys || nsys <= 0) return set_error(e, SOP_ERR_INVALID, "autotune: bad arguments");
  if (e->serve) return set_error(e, SOP_ERR_STATE, "engine is serving");
  for (int32_t q = 0; q < nsys; ++q)
    if (sys[q] < 0 || sys[q] >= (int32_t)e->sys.size() || !e->sys[sys[q]].used)
      return set_error(e, SOP_ERR_INVALID, "autotune: no such system");
  const sop_params &P0 = e->sys[sys[0]].params;
  const int64_t G = std::max(1, P0.n_nl);
  Setting base;
  base.lanes = e->opts.lanes_per_bead;
  // the search starts from two blocks per compute unit, the serving grid:
  // a served mix tuned on a resident batch keeps it unless more blocks gain
  // (fixed serving grids ran 1 % below with the batch default of four)
  base.blocks = e->step_blocks > 0 ? e->step_blocks : 2 * e->num_cu;
  base.lds_beads = e->opts.lds_beads_max;
  base.tune = e->tune;
  if (base.tune.nb == 0) {
    // size buckets: up to one 16-tile item, up to 3,072 beads, larger
    base.tune.nb = 3;
    const int32_t mb[3] = {1024, 3072, 1 << 30};
    for (int b = 0; b < 3; ++b) {
      base.tune.max_beads[b] = mb[b];
      base.tune.shell[b] = e->opts.pl_shell;
      base.tune.tiles[b] = e->opts.step_tiles;
    }
  }
  bool bucket_used[3] = {false, false, false};
  int32_t nmax = 0;
  for (int32_t q = 0; q < nsys; ++q) {
    const int32_t n = e->sys[sys[q]].n;
    nmax = std::max(nmax, n);
    for (int b = 0; b < base.tune.nb; ++b)
      if (n <= base.tune.max_beads[b] || b == base.tune.nb - 1) {
        bucket_used[b] = true;
        break;
      }
  }
  // row capacities with room for the widest shell tried, then no growth while
  // trials run (a trial must not move rows the snapshot restores)
  constexpr float SHELL_MAX_TRIED = 4.0f;
  for (int32_t q = 0; q < nsys; ++q) {
    const sop_status rs = engine_fit_shell(e, sys[q], SHELL_MAX_TRIED);
    if (rs) return rs;
  }
  e->grow_rows = false;
  Snapshot snap;
  sop_status rc = snapshot_take(e, snap);
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  if (!rc) rc = hip_check(e, hipEventCreate(&ev0), "autotune: event");
  if (!rc) rc = hip_check(e, hipEventCreate(&ev1), "autotune: event");
  const int blocks = 4;
  int trials = 0;
  double best_ms = 0.0, base_ms = 0.0;
  Setting best = base;
  if (!rc) rc = trial(e, sys, nsys, snap, base, G, blocks, ev0, ev1, &best_ms);
  base_ms = best_ms;
  ++trials;
  // one parameter at a time, keeping a change only when it gains more than
  // 1 % twice: a candidate that wins its trial is timed again beside the
  // setting it would replace, so that one lucky trial (clock or placement
  // noise of a few per cent) does not decide
  auto consider = [&](const Setting &cand) {
    if (rc) return;
    double ms = 0.0;
    rc = trial(e, sys, nsys, snap, cand, G, blocks, ev0, ev1, &ms);
    ++trials;
    if (rc || !(ms < 0.99 * best_ms)) return;
    double ms_inc = 0.0, ms2 = 0.0;
    rc = trial(e, sys, nsys, snap, best, G, blocks, ev0, ev1, &ms_inc);
    if (!rc) rc = trial(e, sys, nsys, snap, cand, G, blocks, ev0, ev1, &ms2);
    trials += 2;
    if (rc) return;
    if (ms2 < 0.99 * ms_inc) {
      best_ms = std::min(ms, ms2);
      best = cand;
    } else {
      best_ms = std::min(best_ms, ms_inc);
    }
  };
  for (int b = 0; b < best.tune.nb; ++b) {
    if (!bucket_used[b]) continue;
    for (int32_t t : {16, 8, 4}) {
      Setting c = best;
      c.tune.tiles[b] = t;
      if (t != best.tune.tiles[b]) consider(c);
    }
  }
  for (int32_t l : {4, 8}) {
    Setting c = best;
    c.lanes = l;
    if (l != best.lanes) consider(c);
  }
  for (int32_t m : {2, 3, 4}) {
    Setting c = best;
    c.blocks = m * e->num_cu;
    if (c.blocks != best.blocks) consider(c);
  }
  for (int b = 0; b < best.tune.nb; ++b) {
    if (!bucket_used[b]) continue;
    for (float w : {2.0f, 3.0f, SHELL_MAX_TRIED}) {
      Setting c = best;
      c.tune.shell[b] = w;
      if (w != best.tune.shell[b]) consider(c);
    }
  }
  for (int32_t t : {0, 2, 4}) {
    Setting c = best;
    c.tune.tail = t;
    if (t != best.tune.tail) consider(c);
  }
  if (nmax > 3008 && best.lds_beads > 0) {
    for (int32_t l : {5056, 4032, 3008}) {
      Setting c = best;
      c.lds_beads = l;
      if (l != best.lds_beads && l <= (int32_t)(e->lds_limit / 16)) consider(c);
    }
  }
  // the systems as they were, under the chosen settings (the starting ones
  // when a trial failed)
  if (snap.buf) {
    const sop_status r2 = snapshot_restore(e, snap);
    e->grow_rows = true;
    const sop_status r3 = apply(e, sys, nsys, rc ? base : best);
    if (!rc) rc = r2 ? r2 : r3;
  }
  e->grow_rows = true;
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (snap.buf) (void)hipFree(snap.buf);
  if (rc) return rc;
  if (out) {
    std::memset(out, 0, sizeof(*out));
    out->lanes_per_bead = best.lanes;
    out->step_blocks = best.blocks;
    out->lds_beads_max = best.lds_beads;
    out->buckets = best.tune.nb;
    for (int b = 0; b < best.tune.nb && b < 3; ++b) {
      out->bucket_max_beads[b] = best.tune.max_beads[b];
      out->step_tiles[b] = best.tune.tiles[b];
      out->pl_shell[b] = best.tune.shell[b];
    }
    out->block_steps = (int32_t)G;
    out->block_ms_start = base_ms / blocks;
    out->block_ms_best = best_ms / blocks;
    out->trials = trials;
    out->step_tail = best.tune.tail;
  }
  return SOP_OK;
}

sop_status sop_tune_set(sop_engine *e, const sop_tune_result *r) {
  if (!e || !r) return SOP_ERR_INVALID;
  if (e->serve) return set_error(e, SOP_ERR_STATE, "engine is serving");
  if ((r->lanes_per_bead != 4 && r->lanes_per_bead != 8 && r->lanes_per_bead != 16 && r->lanes_per_bead != 32) ||
      r->buckets < 0 || r->buckets > 3 || r->step_blocks < 0 ||
      (r->step_tail != 0 && r->step_tail != 2 && r->step_tail != 4))
    return set_error(e, SOP_ERR_INVALID, "tune: bad settings");
  Setting st;
  st.lanes = r->lanes_per_bead;
  st.blocks = r->step_blocks;
  st.lds_beads = r->lds_beads_max;
  st.tune.nb = r->buckets;
  for (int b = 0; b < r->buckets; ++b) {
    const int32_t t = r->step_tiles[b];
    if (t != 0 && t != 1 && t != 2 && t != 4 && t != 8 && t != 16) return set_error(e, SOP_ERR_INVALID, "tune: step_tiles");
    if (!(r->pl_shell[b] >= 0.0f) || r->pl_shell[b] > 8.0f) return set_error(e, SOP_ERR_INVALID, "tune: pl_shell");
    st.tune.max_beads[b] = r->bucket_max_beads[b];
    st.tune.tiles[b] = t;
    st.tune.shell[b] = r->pl_shell[b];
  }
  st.tune.tail = r->step_tail;
  HIPCK_TUNE(hipStreamSynchronize(e->stream));
  std::vector<int32_t> all;
  for (size_t s = 0; s < e->sys.size(); ++s)
    if (e->sys[s].used) all.push_back((int32_t)s);
  return apply(e, all.data(), (int32_t)all.size(), st);
}

sop_status sop_autotune(sop_engine *e, const int32_t *sys, int32_t nsys, int32_t *lanes, int32_t *grid) {
  sop_tune_result r;
  const sop_status st = sop_autotune_ex(e, sys, nsys, &r);
  if (st) return st;
  if (lanes) *lanes = r.lanes_per_bead;
  if (grid) *grid = r.step_blocks;
  return SOP_OK;
}

}  // extern "C"
