// engine.h -- internal declarations of the libsop host engine.
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/sop.h"
#include "sop_dev.h"

namespace sopk {

hipError_t launch_lists(const DevPtrs &p, const ListGrids &lg, int kstep, int flags, hipStream_t st,
                        const StepFork *fk = nullptr);
hipError_t launch_step(const DevPtrs &p, int mode, int lpb, const StepGrids &sg, int kstep, hipStream_t st,
                       const StepFork *fk = nullptr);
hipError_t step_lds_limit(size_t bytes);
hipError_t filter_lds_init();
hipError_t launch_energy(const DevPtrs &p, int grid_tiles, int grid_sys, hipStream_t st);
hipError_t launch_advance(const DevPtrs &p, int nsteps, hipStream_t st);
hipError_t launch_relayout(const DevPtrs &p, int bo, int n, int idx16, int64_t old_base, int old_cap,
                           int64_t new_base, int new_cap, hipStream_t st);
hipError_t launch_begin(const DevPtrs &p, int64_t nsteps, hipStream_t st);
hipError_t launch_selftest(int *out, float *fout, hipStream_t st);
hipError_t noalloc_scratch(int device, void **ptr);
hipError_t launch_noalloc_energy(const DevPtrs &p, int t0, int nt, int slot, int grid_tiles, hipStream_t st);

// First-fit range allocator over [0, cap).
class RangeAlloc {
 public:
  void init(int64_t cap);
  int64_t alloc(int64_t n, int64_t align); // -1 when full
  void release(int64_t off, int64_t n);
  void grow(int64_t new_cap);
  int64_t cap() const { return cap_; }
  int64_t used() const { return used_; }

 private:
  int64_t cap_ = 0, used_ = 0;
  std::map<int64_t, int64_t> free_;
};

// Device array that can grow, keeping its contents.
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
};

struct SysHost {
  bool used = false;
  int32_t n = 0, n_nat = 0;
  int64_t bead_off = 0, nat_off = 0, nl_off = 0, pl_off = 0, cell_off = 0;
  int64_t bead_len = 0, nat_len = 0, nl_len = 0, pl_len = 0, cell_len = 0;
  int32_t nl_cap = 0, pl_cap = 0, cell_cap = 0;
  float shell = 0.0f; // pair-list shell width (A) of this system
  int64_t step = 0;  // host mirror (exact outside serving)
  int32_t parity = 0;
  sop_params params{};
  std::vector<int32_t> wbits;  // w component of each bead (chain<<2|type)
};

struct Pools {
  int64_t beads = 0, nat = 0, nl = 0, pl = 0, cells = 0, slots = 0;
};

} // namespace sopk

struct sop_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  sopk::StepFork fork{nullptr, nullptr, nullptr};  // whole-CU and global-gather launches beside the LDS ones
  void *scratch = nullptr;  // this device's copy of the allocation-free entry point's scratch
  int num_cu = 0;
  sop_engine_opts opts{};
  std::string err;
  bool fixed_pools = false;
  bool grow_rows = true;  // row capacities grow after runs (off while the autotuner times trials)

  sopk::Pools cap;
  sopk::RangeAlloc a_beads, a_nat, a_pl, a_cells;
  int32_t nl_owner = -1;  // system whose neighbor list the scratch holds
  std::vector<sopk::SysHost> sys;  // by slot
  sopk::DevPtrs p{};
  // raw buffers in allocation order (for growth/free)
  std::vector<std::pair<void **, size_t>> bead_arrays, nat_arrays, cell_arrays;

  // work tables (host staging)
  std::vector<int32_t> h_work, h_syslist;
  std::vector<int4> h_items, h_fitems;
  int64_t work_cap = 0, syslist_cap = 0, items_cap = 0, fitems_cap = 0;
  int32_t cur_nwork = 0, cur_nsys = 0, cur_nitems = 0, cur_nitems_big = 0, cur_nitems_far = 0;
  int32_t cur_nfitems = 0, cur_nfitems_big = 0, cur_nfitems_far = 0;  // shell-filter items
  size_t cur_lds = 0;     // dynamic LDS of the step kernel for the current batch
  size_t cur_lds_big = 0; // dynamic LDS of the whole-CU tier
  size_t cur_flds = 0, cur_flds_big = 0;  // the same for the shell filter's items
  size_t lds_limit = 0;   // largest dynamic LDS the step kernel may use
  int64_t lds_big_cap = 0; // largest system (beads) of the whole-CU LDS tier; 0: no such tier
  int32_t step_blocks = 0; // step-kernel grid (1024-thread blocks); 0 = 2 per compute unit
  int32_t cur_step_tiles = 16; // current batch: tiles per step-kernel item
  bool fixed_grids = false; // serving: launch shapes independent of the batch, so
                            // one captured graph serves every batch composition

  // settings by system size picked by sop_autotune: systems of up to
  // max_beads[b] beads (the last bucket: any size) use shell[b] and items of
  // tiles[b] tiles (0: by batch size); nb = 0: the engine options
  struct Tune {
    int32_t nb = 0;
    int32_t max_beads[3] = {0, 0, 0};
    float shell[3] = {0, 0, 0};
    int32_t tiles[3] = {0, 0, 0};
    int32_t tail = 0;  // > 1: step items after the first round cut to STEP_WAVES / tail tiles
  } tune;

  // graphs keyed by block shape
  std::map<std::string, hipGraphExec_t> graphs;

  // serving
  struct Job;
  struct Serve;
  Serve *serve = nullptr;
};

namespace sopk {
// engine internals shared with serve.cpp
sop_status set_error(sop_engine *e, sop_status st, const std::string &msg);
sop_status hip_check(sop_engine *e, hipError_t h, const char *what);
sop_status engine_add_system(sop_engine *e, const sop_system_desc *d, int32_t *slot);
// A stream for uploads beside the engine stream's work, and a pinned host
// buffer every upload is staged through.
struct Upload {
  hipStream_t stream;
  void *stage;
  size_t stage_bytes;
};
// With `upl`, the system goes into free pool ranges and a free slot only
// (SOP_ERR_BUSY, nothing changed, when it does not fit without growing a
// pool, or its uploads do not fit the staging buffer) and is uploaded on
// upl->stream, which the call drains before returning; work in flight on the
// engine stream is not waited for and not touched.
sop_status engine_add_system_nolists(sop_engine *e, const sop_system_desc *d, int64_t step_end, int32_t *slot,
                                     const Upload *upl = nullptr);
sop_status engine_init_lists(sop_engine *e, const int32_t *slots, int32_t n, sop_status *status);
sop_status engine_remove_system(sop_engine *e, int32_t slot);
sop_status engine_set_batch(sop_engine *e, const int32_t *slots, int32_t nslots);
sop_status engine_run_block_graph(sop_engine *e, int32_t G);
sop_status engine_run_hostloop(sop_engine *e, int64_t nsteps);
sop_status engine_run(sop_engine *e, const int32_t *sys, int32_t nsys, int64_t nsteps, hipEvent_t ev0,
                      hipEvent_t ev1);
sop_status engine_check_errors(sop_engine *e, const int32_t *slots, int32_t nslots);
// Status (and error message) of system slot s from its error bits, on the host.
sop_status engine_error_status(sop_engine *e, int32_t s, int32_t err);
sop_status engine_read_sys(sop_engine *e, int32_t slot, SysDev *out);
void shell_fields(const sop_params &P, float shell, SysDev &S);
int engine_bucket(const sop_engine *e, int32_t n);
float engine_shell_for(const sop_engine *e, int32_t n);
int grid_tiles(const sop_engine *e);
int grid_sys(const sop_engine *e);
int grid_items(const sop_engine *e);
int grid_items_far(const sop_engine *e);
int grid_items_big(const sop_engine *e);
int default_step_blocks(const sop_engine *e);
StepGrids step_grids(const sop_engine *e);
StepGrids filter_grids(const sop_engine *e);
ListGrids list_grids(const sop_engine *e);
int64_t job_bytes_estimate(const sop_engine *e, const sop_system_desc *d);
int64_t engine_system_bytes(const sop_engine *e, int32_t slot);
sop_status engine_fit_new(sop_engine *e, const int32_t *slots, int32_t n);
sop_status engine_fit_shell(sop_engine *e, int32_t slot, float shell);
sop_status engine_grow_rows(sop_engine *e, const int32_t *slots, int32_t n, const std::vector<SysDev> &sd,
                            std::vector<int64_t> *grown);
} // namespace sopk
