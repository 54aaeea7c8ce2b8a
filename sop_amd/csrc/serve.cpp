// serve.cpp -- continuous-batching request loop.
//
// Jobs are submitted from any thread and queued; one serve thread owns the
// device.  Each iteration it
//   1. admits queued jobs in arrival order while the HBM budget (job state
//      bytes) and the job slots allow: upload, initial lists, step_end set;
//   2. runs one hipGraph block of G steps over every resident job (a job that
//      reaches its step count inside the block stops there: its coordinates
//      are carried unchanged through the rest of the block);
//   3. retires finished jobs: FP64 energies and coordinates are copied back,
//      the job's device memory is freed and its latency stamps are set.
// Jobs therefore join and leave the running batch at block boundaries while
// others keep stepping: continuous batching, not a queue served one by one.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>

#include "engine.h"

using namespace sopk;

struct sop_engine::Job {
  int64_t id = 0;
  std::vector<float> xyz, bond_r0, nat_r0;
  std::vector<int32_t> chain, type, nat_i, nat_j;
  sop_system_desc desc{};
  int64_t nsteps = 0;
  int64_t done_steps = 0;  // steps of the blocks launched so far (capped at nsteps)
  int64_t bytes = 0;
  double t_submit = 0, t_admit = 0, t_done = 0;
  int32_t slot = -1;
  int32_t status = SOP_OK;
  int32_t done = 0;
  double energies[SOP_E_COUNT] = {0};
};

struct sop_engine::Serve {
  std::thread th;
  std::mutex mu;
  std::condition_variable cv;
  bool stop = false;
  std::deque<Job *> queue;
  std::unordered_map<int64_t, std::unique_ptr<Job>> jobs;
  std::vector<Job *> running;
  int64_t next_id = 0;
  int64_t budget = 0, used = 0;
  int32_t max_jobs = 0, block = 0, n_nl = 0, n_pl = 0;
  bool schedule_set = false;
  std::string fatal;
  sop_serve_stats stats{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  Upload up{nullptr, nullptr, 0};  // uploads of admitted jobs beside the block in flight
};

namespace {

// After a block: retire the finished jobs and start the admitted ones as one
// batch with one synchronization.  The energies of the finished jobs (the
// admitted ones' lists are still empty: their share is cheap and unused),
// their coordinates and energies back to the host, then the admitted jobs'
// first neighbor and pair lists; the finished jobs' status comes from their
// error bits before the list launch, the admitted jobs' from theirs after it.
sop_status retire_and_start(sop_engine *e, sop_engine::Serve *sv, const std::vector<sop_engine::Job *> &added,
                            const std::vector<int32_t> &added_slots) {
  std::vector<SysDev> sd((size_t)e->cap.slots);
  sop_status st = hip_check(e, hipMemcpy(sd.data(), e->p.sys, sizeof(SysDev) * sd.size(), hipMemcpyDeviceToHost),
                            "retire");
  if (st) return st;
  std::vector<sop_engine::Job *> keep, fin;
  std::vector<int32_t> slots;
  for (sop_engine::Job *j : sv->running) {
    const SysDev &S = sd[j->slot];
    if (S.step >= S.step_end || S.err) {
      fin.push_back(j);
      slots.push_back(j->slot);
    } else {
      keep.push_back(j);
    }
  }
  if (fin.empty() && added.empty()) return SOP_OK;
  slots.insert(slots.end(), added_slots.begin(), added_slots.end());
  st = engine_set_batch(e, slots.data(), (int32_t)slots.size());
  if (st) return st;
  std::vector<double> out(fin.size() * SYSOUT_N);
  std::vector<std::vector<float4>> xs(fin.size());
  if (!fin.empty()) {
    st = hip_check(e, launch_energy(e->p, grid_tiles(e), grid_sys(e), e->stream), "energy");
    if (st) return st;
  }
  for (size_t q = 0; q < fin.size(); ++q) {
    const SysDev &S = sd[fin[q]->slot];
    st = hip_check(e, hipMemcpyAsync(&out[q * SYSOUT_N], e->p.sysout + (int64_t)fin[q]->slot * SYSOUT_N,
                                     8 * SYSOUT_N, hipMemcpyDeviceToHost, e->stream), "energy copy");
    if (st) return st;
    xs[q].resize((size_t)S.n);
    st = hip_check(e, hipMemcpyAsync(xs[q].data(), (S.parity & 1 ? e->p.pos1 : e->p.pos0) + S.bead_off,
                                     16 * (size_t)S.n, hipMemcpyDeviceToHost, e->stream), "coords copy");
    if (st) return st;
  }
  if (!added.empty()) {
    st = hip_check(e, launch_lists(e->p, list_grids(e), 0, FORCE_NL | FORCE_PL, e->stream), "first lists");
    if (st) return st;
  }
  st = hip_check(e, hipStreamSynchronize(e->stream), "retire sync");
  if (st) return st;
  // row capacities: the admitted jobs' fitted to their first lists, the
  // running ones' grown where their rows came close (the job's charge follows)
  if (!added.empty()) {
    std::vector<int64_t> before(added.size());
    for (size_t q = 0; q < added.size(); ++q) before[q] = engine_system_bytes(e, added_slots[q]);
    st = engine_fit_new(e, added_slots.data(), (int32_t)added_slots.size());
    if (st) return st;
    std::lock_guard<std::mutex> lk(sv->mu);
    for (size_t q = 0; q < added.size(); ++q) {
      const int64_t d = engine_system_bytes(e, added_slots[q]) - before[q];
      added[q]->bytes += d;
      sv->used += d;
    }
  }
  if (!keep.empty()) {
    std::vector<int32_t> ks;
    for (sop_engine::Job *j : keep) ks.push_back(j->slot);
    std::vector<int64_t> grown;
    st = engine_grow_rows(e, ks.data(), (int32_t)ks.size(), sd, &grown);
    if (st) return st;
    std::lock_guard<std::mutex> lk(sv->mu);
    for (size_t q = 0; q < keep.size(); ++q) {
      keep[q]->bytes += grown[q];
      sv->used += grown[q];
    }
  }
  std::vector<SysDev> sa;
  if (!added.empty()) {
    sa.resize((size_t)e->cap.slots);
    st = hip_check(e, hipMemcpy(sa.data(), e->p.sys, sizeof(SysDev) * sa.size(), hipMemcpyDeviceToHost),
                   "admitted status");
    if (st) return st;
  }
  const double now = sop_clock_seconds();
  for (size_t q = 0; q < fin.size(); ++q) {
    sop_engine::Job *j = fin[q];
    const int32_t status = engine_error_status(e, j->slot, sd[j->slot].err);
    engine_remove_system(e, j->slot);
    std::lock_guard<std::mutex> lk(sv->mu);
    for (size_t i = 0; i < xs[q].size(); ++i) {
      j->xyz[3 * i] = xs[q][i].x;
      j->xyz[3 * i + 1] = xs[q][i].y;
      j->xyz[3 * i + 2] = xs[q][i].z;
    }
    for (int k = 0; k < SOP_E_COUNT; ++k) j->energies[k] = out[q * SYSOUT_N + k];
    j->status = status;
    j->t_done = now;
    j->done = 1;
    j->slot = -1;
    sv->used -= j->bytes;
  }
  for (sop_engine::Job *j : added) {
    const sop_status status = engine_error_status(e, j->slot, sa[j->slot].err);
    if (status != SOP_OK) engine_remove_system(e, j->slot);
    std::lock_guard<std::mutex> lk(sv->mu);
    j->t_admit = now;
    if (status != SOP_OK) {
      j->status = status;
      j->done = 1;
      j->t_done = now;
      j->slot = -1;
      sv->used -= j->bytes;
      continue;
    }
    keep.push_back(j);
  }
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->running.swap(keep);
  return SOP_OK;
}

// Admission: queued jobs in arrival order while the running-job and HBM
// budget limits allow.  With `up`, the uploads overlap the block in flight on
// the engine stream; a job that would need a pool to grow goes back to the
// queue front with the ones behind it, *deferred counts them, and the caller
// admits them once the engine is idle.
void admit(sop_engine *e, sop_engine::Serve *sv, const Upload *up, std::vector<sop_engine::Job *> &added,
           std::vector<int32_t> &added_slots, size_t *deferred) {
  std::vector<sop_engine::Job *> take;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    while (!sv->queue.empty()) {
      sop_engine::Job *j = sv->queue.front();
      const size_t held = sv->running.size() + added.size() + take.size();
      if ((int)held >= sv->max_jobs) break;
      if (sv->used + j->bytes > sv->budget && held > 0) break;
      sv->queue.pop_front();
      sv->used += j->bytes;
      take.push_back(j);
    }
  }
  for (size_t q = 0; q < take.size(); ++q) {
    sop_engine::Job *j = take[q];
    int32_t slot = -1;
    const sop_status st = engine_add_system_nolists(e, &j->desc, j->desc.step0 + j->nsteps, &slot, up);
    if (st == SOP_ERR_BUSY && up) {
      std::lock_guard<std::mutex> lk(sv->mu);
      for (size_t r = take.size(); r-- > q;) {
        sv->queue.push_front(take[r]);
        sv->used -= take[r]->bytes;
      }
      *deferred += take.size() - q;
      return;
    }
    if (st != SOP_OK) {
      std::lock_guard<std::mutex> lk(sv->mu);
      j->t_admit = sop_clock_seconds();
      j->status = st;
      j->done = 1;
      j->t_done = j->t_admit;
      sv->used -= j->bytes;
      continue;
    }
    j->slot = slot;
    added.push_back(j);
    added_slots.push_back(slot);
  }
}

// First lists of the admitted jobs (one launch sequence), then into the
// running batch.
sop_status start_jobs(sop_engine *e, sop_engine::Serve *sv, const std::vector<sop_engine::Job *> &added,
                      const std::vector<int32_t> &added_slots) {
  if (added.empty()) return SOP_OK;
  std::vector<sop_status> ok(added.size(), SOP_OK);
  const sop_status st = engine_init_lists(e, added_slots.data(), (int32_t)added.size(), ok.data());
  if (st != SOP_OK) return st;
  std::lock_guard<std::mutex> lk(sv->mu);
  const double now = sop_clock_seconds();
  for (size_t q = 0; q < added.size(); ++q) {
    sop_engine::Job *j = added[q];
    j->t_admit = now;
    if (ok[q] != SOP_OK) {
      j->status = ok[q];
      j->done = 1;
      j->t_done = now;
      j->slot = -1;
      sv->used -= j->bytes;
      continue;
    }
    const int64_t b = engine_system_bytes(e, j->slot);  // after its row capacity was fitted
    sv->used += b - j->bytes;
    j->bytes = b;
    sv->running.push_back(j);
  }
  return SOP_OK;
}

void serve_loop(sop_engine *e) {
  sop_engine::Serve *sv = e->serve;
  (void)hipSetDevice(e->device);
  for (;;) {
    const double t0 = sop_clock_seconds();
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      if (sv->running.empty() && sv->queue.empty() && !sv->stop)
        sv->cv.wait_for(lk, std::chrono::milliseconds(5));
      if (sv->stop) break;
    }
    // 1. one block of steps over the running batch, in flight on the engine stream
    sop_status st = SOP_OK;
    const bool run = !sv->running.empty();
    double t2 = t0, t3 = t0;
    int64_t block_bead_steps = 0;
    if (run) {
      std::vector<int32_t> slots;
      for (sop_engine::Job *j : sv->running) {
        slots.push_back(j->slot);
        const int64_t adv = std::min<int64_t>(sv->block, j->nsteps - j->done_steps);
        j->done_steps += adv;
        block_bead_steps += (int64_t)j->desc.n_beads * adv;
      }
      st = engine_set_batch(e, slots.data(), (int32_t)slots.size());
      t2 = sop_clock_seconds();
      if (st == SOP_OK && sv->ev0) (void)hipEventRecord(sv->ev0, e->stream);
      if (st == SOP_OK) st = engine_run_block_graph(e, sv->block);
      if (st == SOP_OK && sv->ev1) (void)hipEventRecord(sv->ev1, e->stream);
      t3 = sop_clock_seconds();
    }
    // 2. admission, uploading beside the block (free pool ranges only)
    std::vector<sop_engine::Job *> added;
    std::vector<int32_t> added_slots;
    size_t deferred = 0;
    if (st == SOP_OK) admit(e, sv, run && sv->up.stream ? &sv->up : nullptr, added, added_slots, &deferred);
    const size_t beside = run && sv->up.stream ? added.size() : 0;
    const double ta = sop_clock_seconds();
    // 3. retire finished jobs and start the admitted ones (waits for the block)
    const size_t before = sv->running.size();
    size_t retired = 0;
    if (st == SOP_OK && run) {
      st = retire_and_start(e, sv, added, added_slots);
      retired = before + added.size() - sv->running.size();
    } else if (st == SOP_OK) {
      st = start_jobs(e, sv, added, added_slots);
    }
    const double t4 = sop_clock_seconds();
    // 4. jobs that needed a pool to grow, now that the engine is idle
    const size_t n_deferred = deferred;
    std::vector<sop_engine::Job *> later;
    std::vector<int32_t> later_slots;
    if (st == SOP_OK && deferred) {
      admit(e, sv, nullptr, later, later_slots, &deferred);
      st = start_jobs(e, sv, later, later_slots);
    }
    const double t5 = sop_clock_seconds();
    if (st == SOP_OK) {
      float ms = 0.0f;
      if (run && sv->ev0 && hipEventElapsedTime(&ms, sv->ev0, sv->ev1) != hipSuccess) ms = 0.0f;
      std::lock_guard<std::mutex> lk(sv->mu);
      sop_serve_stats &c = sv->stats;
      c.host_admit_s += (ta - t3) + (t5 - t4);
      c.jobs_admitted += (int64_t)(added.size() + later.size());
      c.jobs_uploaded_beside += (int64_t)beside;
      c.jobs_deferred += (int64_t)n_deferred;
      if (run) {
        c.mean_running = (c.mean_running * (double)c.iterations + (double)before) / (double)(c.iterations + 1);
        c.iterations += 1;
        c.jobs_retired += (int64_t)retired;
        c.steps_run += sv->block;
        c.bead_steps += block_bead_steps;
        c.gpu_block_s += 1e-3 * ms;
        c.host_batch_s += t2 - t0;
        c.host_block_s += t3 - t2;
        c.host_retire_s += t4 - ta;
      }
    }
    if (st != SOP_OK) {
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->fatal = e->err;
      for (sop_engine::Job *j : sv->running) {
        j->status = st;
        j->done = 1;
        j->t_done = sop_clock_seconds();
      }
      sv->running.clear();
      return;
    }
  }
}

} // namespace

extern "C" {

sop_status sop_serve_start(sop_engine *e, const sop_serve_opts *o) {
  if (!e) return SOP_ERR_INVALID;
  if (e->serve) return set_error(e, SOP_ERR_STATE, "already serving");
  auto *sv = new sop_engine::Serve();
  sv->budget = (o && o->hbm_budget_bytes > 0) ? o->hbm_budget_bytes : (int64_t)8 << 30;
  sv->max_jobs = (o && o->max_jobs > 0) ? o->max_jobs : 4096;
  sv->block = (o && o->block_steps > 0) ? o->block_steps : 0;
  e->serve = sv;
  e->fixed_grids = true;
  if (hipEventCreate(&sv->ev0) != hipSuccess || hipEventCreate(&sv->ev1) != hipSuccess) sv->ev0 = sv->ev1 = nullptr;
  // uploads beside the running block: a non-blocking stream and 64 MiB of
  // pinned staging (jobs whose uploads exceed it are admitted between blocks)
  sv->up.stage_bytes = (size_t)64 << 20;
  if (hipHostMalloc(&sv->up.stage, sv->up.stage_bytes, hipHostMallocDefault) != hipSuccess) {
    sv->up.stage = nullptr;
  } else if (hipStreamCreateWithFlags(&sv->up.stream, hipStreamNonBlocking) != hipSuccess) {
    sv->up.stream = nullptr;
  }
  sv->th = std::thread(serve_loop, e);
  return SOP_OK;
}

sop_status sop_serve_submit(sop_engine *e, const sop_system_desc *d, int64_t nsteps, int64_t *job) {
  if (!e || !d || !job || nsteps < 0 || d->n_beads < 1 || !d->xyz) return SOP_ERR_INVALID;
  sop_engine::Serve *sv = e->serve;
  if (!sv) return SOP_ERR_STATE;
  auto j = std::make_unique<sop_engine::Job>();
  const int n = d->n_beads, m = d->n_native;
  j->xyz.assign(d->xyz, d->xyz + 3 * (size_t)n);
  if (d->chain) j->chain.assign(d->chain, d->chain + n);
  if (d->type) j->type.assign(d->type, d->type + n);
  if (d->bond_r0) j->bond_r0.assign(d->bond_r0, d->bond_r0 + n);
  if (m > 0) {
    j->nat_i.assign(d->nat_i, d->nat_i + m);
    j->nat_j.assign(d->nat_j, d->nat_j + m);
    j->nat_r0.assign(d->nat_r0, d->nat_r0 + m);
  }
  j->desc = *d;
  j->desc.xyz = j->xyz.data();
  j->desc.chain = d->chain ? j->chain.data() : nullptr;
  j->desc.type = d->type ? j->type.data() : nullptr;
  j->desc.bond_r0 = d->bond_r0 ? j->bond_r0.data() : nullptr;
  j->desc.nat_i = m > 0 ? j->nat_i.data() : nullptr;
  j->desc.nat_j = m > 0 ? j->nat_j.data() : nullptr;
  j->desc.nat_r0 = m > 0 ? j->nat_r0.data() : nullptr;
  j->nsteps = nsteps;
  j->bytes = job_bytes_estimate(e, d);
  j->t_submit = sop_clock_seconds();
  std::lock_guard<std::mutex> lk(sv->mu);
  if (!sv->fatal.empty()) return SOP_ERR_STATE;
  // one schedule per served batch: graph blocks hold the list rebuilds at
  // fixed offsets, so every job shares n_nl / n_pl and starts on a block boundary
  const sop_params &P = d->params;
  if (!sv->schedule_set) {
    if (P.n_pl < 1 || P.n_nl < 1 || P.n_nl % P.n_pl != 0)
      return SOP_ERR_INVALID;
    sv->n_nl = P.n_nl;
    sv->n_pl = P.n_pl;
    if (sv->block <= 0) sv->block = P.n_nl;
    if (sv->block % P.n_pl != 0) return SOP_ERR_INVALID;
    sv->schedule_set = true;
  }
  if (P.n_nl != sv->n_nl || P.n_pl != sv->n_pl)
    return SOP_ERR_INVALID;
  if (d->step0 % sv->block != 0) return SOP_ERR_INVALID;
  j->id = sv->next_id++;
  *job = j->id;
  sv->queue.push_back(j.get());
  sv->jobs[j->id] = std::move(j);
  sv->cv.notify_one();
  return SOP_OK;
}

sop_status sop_serve_poll(sop_engine *e, int64_t job, sop_job_result *r, float *xyz) {
  if (!e || !r) return SOP_ERR_INVALID;
  sop_engine::Serve *sv = e->serve;
  if (!sv) return SOP_ERR_STATE;
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->jobs.find(job);
  if (it == sv->jobs.end()) return SOP_ERR_INVALID;
  sop_engine::Job *j = it->second.get();
  std::memset(r, 0, sizeof(*r));
  r->done = j->done;
  r->status = j->status;
  r->n_beads = j->desc.n_beads;
  r->steps = j->nsteps;
  r->t_submit = j->t_submit;
  r->t_admit = j->t_admit;
  r->t_done = j->t_done;
  if (!j->done) return SOP_ERR_BUSY;
  for (int k = 0; k < SOP_E_COUNT; ++k) r->energies[k] = j->energies[k];
  if (xyz) std::memcpy(xyz, j->xyz.data(), sizeof(float) * j->xyz.size());
  return SOP_OK;
}

int64_t sop_serve_job_bytes(const sop_engine *e, const sop_system_desc *d) {
  if (!e || !d || d->n_beads < 1) return -1;
  return job_bytes_estimate(e, d);
}

sop_status sop_serve_stats_get(sop_engine *e, sop_serve_stats *out) {
  if (!e || !out) return SOP_ERR_INVALID;
  sop_engine::Serve *sv = e->serve;
  if (!sv) return SOP_ERR_STATE;
  std::lock_guard<std::mutex> lk(sv->mu);
  *out = sv->stats;
  return SOP_OK;
}

sop_status sop_serve_stop(sop_engine *e) {
  if (!e) return SOP_ERR_INVALID;
  sop_engine::Serve *sv = e->serve;
  if (!sv) return SOP_OK;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->cv.notify_all();
  if (sv->th.joinable()) sv->th.join();
  // release whatever is still resident
  for (sop_engine::Job *j : sv->running)
    if (j->slot >= 0) engine_remove_system(e, j->slot);
  const bool ok = sv->fatal.empty();
  if (sv->ev0) (void)hipEventDestroy(sv->ev0);
  if (sv->ev1) (void)hipEventDestroy(sv->ev1);
  if (sv->up.stream) (void)hipStreamDestroy(sv->up.stream);
  if (sv->up.stage) (void)hipHostFree(sv->up.stage);
  delete sv;
  e->serve = nullptr;
  e->fixed_grids = false;
  return ok ? SOP_OK : SOP_ERR_STATE;
}

} // extern "C"
