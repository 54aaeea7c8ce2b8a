/*
 * sop.h -- C ABI of libsop, the MI355X (gfx950) SOP overdamped dynamics engine.
 *
 * Plain C types only: pointers, sizes, status codes.  One engine owns one HIP
 * device, its memory pools and streams; systems (replicas) live on the device
 * until removed.  An engine is driven by one host thread and is not
 * thread-safe, except that the serve loop runs an internal thread behind
 * sop_serve_submit / sop_serve_poll.
 *
 * Reference interface replaced (samchogroup/SOP):
 *   force/energy entry points  energy.h:13-24 (set_potential, clear_forces,
 *                              energy_eval, force_eval, random_force,
 *                              fene_energy, soft_sphere_angular_energy,
 *                              vdw_energy)
 *   list entry points          sop.h:32-35 (update_neighbor_list,
 *                              update_pair_list, update_cell_list,
 *                              update_hybrid_list)
 *   integrators                sop.h:27-28 (underdamped_iteration,
 *                              overdamped_iteration)
 *   device state               GPUvars.h:150-153 (alloc_GPU_arrays,
 *                              alloc_cudpp, setup_rng)
 * The reference passes all data through globals (global.h:83-233); here every
 * entry point takes the engine and a system id instead, and returns a status
 * code instead of calling exit(-1).
 */
#ifndef SOP_H_
#define SOP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t sop_status;

#define SOP_OK 0
#define SOP_ERR_INVALID (-1)   /* bad argument / topology */
#define SOP_ERR_NO_DEVICE (-2) /* no HIP device (the engine has no CPU fallback) */
#define SOP_ERR_HIP (-3)       /* HIP runtime error */
#define SOP_ERR_CAPACITY (-4)  /* list / cell capacity overflow or pool full */
#define SOP_ERR_NUMERIC (-5)   /* non-finite coordinates or overstretched bond */
#define SOP_ERR_STATE (-6)     /* call not valid in the current state */
#define SOP_ERR_IO (-7)        /* file read / write failure */
#define SOP_ERR_BUSY (-8)      /* serve: job not finished / queue full */
#define SOP_ERR_PBC (-9)       /* system extent too close to half the periodic box */

/* energy slots returned by sop_energy_eval (FP64 sums of FP32 terms) */
#define SOP_E_BOND 0   /* e_bnd      (energy.cu:194-225) */
#define SOP_E_ANGLE 1  /* e_ang_ss   (energy.cu:278-307) */
#define SOP_E_NATIVE 2 /* e_vdw_rr_att (energy.cu:381-410) */
#define SOP_E_REP 3    /* e_vdw_rr_rep (energy.cu:412-438) */
#define SOP_E_TOTAL 4  /* rna_etot = sum of the above (energy.cu:76) */
#define SOP_E_VIRIAL 5 /* sum over pairs of r_ij . f_ij (no reference counterpart) */
#define SOP_E_COUNT 6

/* lists for sop_get_list: half lists, pairs (i<j) of 0-based bead indices */
#define SOP_LIST_NL_NATIVE 0 /* idx_neighbor_list_att (global.h:165) */
#define SOP_LIST_NL_REP 1    /* idx_neighbor_list_rep (global.h:167) */
#define SOP_LIST_PL_NATIVE 2 /* idx_pair_list_att     (global.h:170) */
#define SOP_LIST_PL_REP 3    /* idx_pair_list_rep     (global.h:172) */

/* dynamics (sim_type, param.cpp:239-254) */
#define SOP_DYN_OVERDAMPED 0  /* sim_type 2: x += (h/zeta)(F + noise) (sop.cu:487-520) */
#define SOP_DYN_UNDERDAMPED 1 /* sim_type 1: Langevin leapfrog with a1..a4 (sop.cu:421-484) */

typedef struct sop_engine sop_engine;

/* Force field and run parameters of one system.  Bead types are 1 or 2 as in
 * the reference's 3x3 tables (global.cpp:51-92); row/column 0 is unused. */
typedef struct sop_params {
  float k_bnd;    /* FENE spring constant, kcal/mol/A^2 (param.cpp:14) */
  float R0;       /* FENE max deviation, A (param.cpp:15) */
  float eps_ss;   /* i,i+2 soft-sphere epsilon (param.cpp:37) */
  float sigma_ss; /* i,i+2 soft-sphere sigma (param.cpp:35) */
  float eps_att[3][3];   /* native LJ well depth, one table for energy and force */
  float eps_rep[3][3];   /* non-native repulsion epsilon (global.cpp:57-61) */
  float sigma_rep[3][3]; /* non-native sigma (global.cpp:75-79) */
  float rep_c12;  /* repulsion E = eps*(c12*(s/r)^12 + c6*(s/r)^6); reference (1,1) */
  float rep_c6;
  /* Cutoffs: factor > 0 gives rc = (float)(factor * (double)base) with base = r0
   * for natives and sigma_rep for non-natives (sop.cu:746,815,861,910);
   * factor <= 0 gives rc = abs (absolute cutoff in A). */
  double nl_factor;
  float nl_abs;
  double pl_factor;
  float pl_abs;
  int32_t n_nl;   /* neighbor-list rebuild interval, steps (nnlup, param.cpp:103) */
  int32_t n_pl;   /* pair-list rebuild interval, steps (reference: every step) */
  float T;        /* temperature, kcal/mol (param.cpp:92) */
  float zeta;     /* friction, 1/tau_L (param.cpp:100) */
  float h;        /* time step, tau_L (param.cpp:104; overdamped 0.02, param.cpp:252) */
  float boxl;     /* periodic box edge, A; <= 0 disables the minimum image */
  uint64_t seed;  /* Philox4x32-10 key of this system's random force stream */
  int32_t dynamics; /* SOP_DYN_OVERDAMPED or SOP_DYN_UNDERDAMPED */
  int32_t reserved; /* 0 */
} sop_params;

/* One system: linear chains of beads.  Beads of a chain are contiguous and
 * consecutive beads of a chain are bonded (reference bonds file, io.cpp:126-148,
 * with ibead = i, jbead = i+1); the i,i+2 soft-sphere term acts inside a chain
 * (angles file, io.cpp:149-173).  Non-native repulsion acts on every pair that
 * is not native and not (same chain and |i-j| <= 2): the reference's master
 * non-native list (io.cpp:190-226).  Indices are 0-based. */
typedef struct sop_system_desc {
  int32_t n_beads;
  const float *xyz;       /* 3*n unwrapped coordinates, A */
  const int32_t *chain;   /* n chain ids, non-decreasing; NULL = one chain */
  const int32_t *type;    /* n bead types in {1,2}; NULL = all 1 */
  const float *bond_r0;   /* n: equilibrium length of bond (i,i+1); NULL = 3.8 */
  int32_t n_native;
  const int32_t *nat_i;   /* native pairs, nat_i[k] < nat_j[k] */
  const int32_t *nat_j;
  const float *nat_r0;    /* native distance (lj_nat_pdb_dist, io.cpp:208) */
  sop_params params;
  int64_t step0;          /* initial value of the step counter */
  const float *vel;       /* underdamped: 3*n velocities entering the first step
                             (vel, global.h:176); NULL = 0 */
} sop_system_desc;

typedef struct sop_engine_opts {
  int32_t nl_cap_max;     /* per-bead neighbor-list capacity bound (default 4096) */
  int32_t pl_cap_max;     /* per-bead pair-list capacity bound (default 1024, at most 4095) */
  int32_t cell_cap_max;   /* per-system cell-grid capacity bound (default 32768) */
  int32_t lanes_per_bead; /* force-kernel lanes per bead: 4, 8, 16 or 32 (default 4; sop_autotune picks one) */
  int32_t grid_blocks;    /* persistent grid size; 0 = 8 per compute unit */
  int32_t use_graphs;     /* capture step blocks as hipGraphs (default 1) */
  int32_t lds_beads_max;  /* systems up to this size run the step with LDS-resident
                             coordinates (default 5056; < 0 disables) */
  float pl_shell;         /* pair-list shell width in A (default 3; < 0 or > 8: none): the
                             cell grid lists pairs within rc_pl + pl_shell, and later
                             pair-list rebuilds filter those rows while every bead has
                             moved less than pl_shell / 2 (same lists, less work) */
  int64_t pool_bytes;     /* serve: HBM budget for all pools; 0 = grow on demand */
  int32_t step_tiles;     /* 64-bead tiles per work item of the step kernel: 16, 8, 4, 2 or 1;
                             an item of t tiles runs each tile on min(lanes_per_bead, 16 / t)
                             waves (0 = 16 while the batch fills the device, fewer below) */
  int32_t step_blocks;    /* step-kernel grid in 1024-thread blocks (0 = 4 per compute unit for
                             a batch, 2 while serving; sop_autotune picks one) */
  int32_t step_tail;      /* 2 or 4: step items that start after a first round of full items
                             are cut to 16 / step_tail tiles with more waves per tile (a
                             shorter last round); 0 = off (default 2) */
  int32_t reserved_opts;  /* 0 */
} sop_engine_opts;

/* --- parameters ---------------------------------------------------------- */
/* Reference defaults (param.cpp:14-110,252, global.cpp:51-92; one eps table). */
sop_status sop_params_reference(sop_params *p);
/* Defaults of the benchmark configurations: eps_h = eps_l = 1, sigma = 3.8,
 * sigma_ss = 5.4, absolute cutoffs 32 A / 16 A every 50 / 10 steps, T = 0.6,
 * dt = 0.05, zeta = 50, no periodic box. */
sop_status sop_params_config(sop_params *p);
/* Derived FP32 coefficients exactly as the engine uses them:
 * a5 = h/zeta (param.cpp:110,253), var = sqrt(2 T zeta / h) (energy.cu:158). */
sop_status sop_params_derived(const sop_params *p, float *a5, float *var);
/* Underdamped coefficients a[0..3] = a1..a4 with the reference's float/double
 * evaluation order (param.cpp:104-108): a1 = h(1 - zeta h/2), a2 = h h/2,
 * a3 = (1 - h zeta/2 + (h zeta)^2/4)/h, a4 = (h/2)(1 - h zeta/2). */
sop_status sop_params_underdamped(const sop_params *p, float *a4);

/* --- engine -------------------------------------------------------------- */
void sop_engine_opts_default(sop_engine_opts *o);
/* device >= 0: HIP device ordinal.  Fails with SOP_ERR_NO_DEVICE without one. */
sop_status sop_engine_create(int32_t device, const sop_engine_opts *opts, sop_engine **out);
void sop_engine_destroy(sop_engine *e);
const char *sop_last_error(const sop_engine *e);
const char *sop_version(void);
sop_status sop_synchronize(sop_engine *e);

/* --- systems (alloc_GPU_arrays / load, GPUvars.cu:99-228, io.cpp:112-259) -- */
sop_status sop_system_add(sop_engine *e, const sop_system_desc *d, int32_t *sys);
sop_status sop_system_remove(sop_engine *e, int32_t sys);
sop_status sop_system_info(sop_engine *e, int32_t sys, int32_t *n_beads, int64_t *step);

/* --- lists (update_neighbor_list / update_pair_list, sop.cu:928-1823) ----- */
/* Rebuild the neighbor list with the cell hash (counting sort + 27-cell
 * stencil + wave-level compaction); also covers update_cell_list and
 * update_hybrid_list, whose intent (a spatial candidate filter) it replaces. */
sop_status sop_update_neighbor_list(sop_engine *e, int32_t sys);
sop_status sop_update_pair_list(sop_engine *e, int32_t sys);
/* Copy a half list (i<j pairs, 2*n int32 as i0,j0,i1,j1,...) to ij; n is the
 * total count even when it exceeds cap. */
sop_status sop_get_list(sop_engine *e, int32_t sys, int32_t which, int32_t *ij, int64_t cap,
                        int64_t *n);

/* --- forces and energies (energy.h:13-24) -------------------------------- */
/* force_eval(): forces at the current coordinates over the current pair
 * list.  with_noise = 0: deterministic forces only; 1: plus the random force
 * of the next step (random_force, energy.cu:156-167). */
sop_status sop_force_eval(sop_engine *e, int32_t sys, int32_t with_noise);
sop_status sop_get_forces(sop_engine *e, int32_t sys, float *f3n);
/* energy_eval(): e[SOP_E_COUNT] over the current pair list. */
sop_status sop_energy_eval(sop_engine *e, int32_t sys, double *energies);
/* sop_energy_eval without any device allocation: no hipMalloc, hipFree or
 * allocating library call.  No counterpart in the reference; for callers that
 * cannot allocate (code under HIP graph capture, serving loops, harnesses
 * that forbid it).  The work table and the energy partials live in a
 * file-scope device array of the library, ~3 MiB per device, held until the
 * process ends; systems of up to 65,536 beads (SOP_ERR_CAPACITY above).
 * Same kernels and summation order as sop_energy_eval: identical results.
 * Every call on one device shares that device's array: calls that overlap on
 * one device (engines on the same device driven from different threads) race. */
sop_status sop_energy_eval_noalloc(sop_engine *e, int32_t sys, double *energies);
/* calculate_observables() (sop.cu:528-611): Q, native contact count, Rg^2. */
sop_status sop_observables(sop_engine *e, int32_t sys, double *Q, int32_t *contacts,
                           double *rgsq);

/* --- dynamics (overdamped_ctrl / overdamped_iteration, sop.cu:287-520) ----- */
/* Advance the listed systems by nsteps each, in one batch (one launch per
 * phase for all systems; hipGraph-captured blocks when aligned). */
sop_status sop_run(sop_engine *e, const int32_t *sys, int32_t nsys, int64_t nsteps);
/* Same dynamics, one system per launch, no graphs: the plain per-system path. */
sop_status sop_run_plain(sop_engine *e, const int32_t *sys, int32_t nsys, int64_t nsteps);
sop_status sop_get_coords(sop_engine *e, int32_t sys, float *xyz, int64_t *step);
sop_status sop_set_coords(sop_engine *e, int32_t sys, const float *xyz);
/* Philox key and the step counter of the next step (setup_rng,
 * GPUvars.cu:430-453; the counter also sets the list schedule).  With the
 * coordinates (and velocities) it is all a restart needs. */
sop_status sop_set_rng(sop_engine *e, int32_t sys, uint64_t seed, int64_t step);
/* Wrapped coordinates pos = unc - boxl*rint(unc/boxl) (sop.cu:510-512). */
sop_status sop_get_wrapped_coords(sop_engine *e, int32_t sys, float *xyz);
/* Velocities as the reference holds them after its last iteration (vel,
 * update_velocities_kernel, sop.cu:455-467): underdamped a3*incr + a4*F with F
 * the force (with noise) at the current coordinates and step, or the frozen
 * velocities when T < 1e-5; overdamped: zeros (sop.cu:296-305). */
sop_status sop_get_velocities(sop_engine *e, int32_t sys, float *v3n);
/* The position increment of the last step (dev_incr, used by kinT in
 * calculate_observables, sop.cu:596-606).  Valid after at least one step. */
sop_status sop_get_increments(sop_engine *e, int32_t sys, float *d3n);

/* --- measurement --------------------------------------------------------- */
typedef struct sop_list_stats {
  int64_t nl_native, nl_rep, pl_native, pl_rep; /* half-list entry counts */
  int32_t nl_max_per_bead, pl_max_per_bead;
  /* pair-list rebuilds since the system was added, by how they were made: from
   * the cell grid (with new shell rows), by filtering the shell rows, or from
   * the neighbor-list snapshot's cell grid (beads had left the skin) */
  int64_t grid_builds, shell_builds, nl_builds;
} sop_list_stats;
sop_status sop_list_stats_get(sop_engine *e, int32_t sys, sop_list_stats *st);
/* Row capacity of a system's shell rows and pair-list tiles (entries per
 * bead) and the longest row built so far.  The capacity starts at
 * min(n - 1, pl_cap_max, 384), is fitted to the first lists, and grows after
 * runs whose rows came within an eighth of it; growth moves the rows and
 * changes no result.  No counterpart in the reference (whose lists are sized
 * by the O(N^2) master list). */
sop_status sop_system_rows(sop_engine *e, int32_t sys, int32_t *capacity, int32_t *longest);
/* Set that capacity ahead of a change the engine cannot foresee (a system
 * about to compact): rounded up to a multiple of 16, kept within
 * [longest row, min(n - 1, pl_cap_max)].  Results do not change. */
sop_status sop_system_reserve_rows(sop_engine *e, int32_t sys, int32_t capacity);
/* Time `reps` launches of the fused force+integrate kernel over the listed
 * systems with HIP events on the engine stream; avg_ms per launch.  The
 * systems' coordinates and counters are restored afterwards. */
sop_status sop_time_step_kernel(sop_engine *e, const int32_t *sys, int32_t nsys, int32_t reps,
                                double *avg_ms);
/* Settings picked by sop_autotune_ex for the listed systems, by timing whole
 * hipGraph blocks of their own steps (n_nl steps each, list rebuilds
 * included), one setting at a time.  Systems fall in size buckets (up to
 * bucket_max_beads[b] beads; the last bucket takes any size) with their own
 * pair-list shell width and step-kernel item size. */
typedef struct sop_tune_result {
  int32_t lanes_per_bead;      /* 4 or 8 */
  int32_t step_blocks;         /* step-kernel grid in 1024-thread blocks */
  int32_t lds_beads_max;       /* largest system staged in LDS at two blocks per compute unit */
  int32_t buckets;             /* size buckets in use (3) */
  int32_t bucket_max_beads[3];
  int32_t step_tiles[3];       /* tiles per step-kernel item (0: by batch size) */
  float pl_shell[3];           /* pair-list shell width, A */
  int32_t block_steps;         /* steps of one timed block (n_nl) */
  double block_ms_start;       /* device time of one block under the starting settings */
  double block_ms_best;        /* the same under the chosen ones */
  int32_t trials;
  int32_t step_tail;           /* as sop_engine_opts.step_tail */
} sop_tune_result;
/* The systems' coordinates, lists and counters are restored after the
 * trials (a device snapshot of the engine, which needs that much free
 * memory); the chosen settings stay in the engine, and systems added later
 * take their size bucket's shell width.  Not while serving. */
sop_status sop_autotune_ex(sop_engine *e, const int32_t *sys, int32_t nsys, sop_tune_result *out);
/* Apply settings (from sop_autotune_ex, possibly on another engine with a
 * representative batch) to this engine: launch shapes, and the size buckets'
 * shell widths and item sizes for its systems and those added later.  Not
 * while serving; call it before sop_serve_start to serve with tuned settings. */
sop_status sop_tune_set(sop_engine *e, const sop_tune_result *r);
/* sop_autotune_ex, returning the lanes per bead and the step-kernel grid. */
sop_status sop_autotune(sop_engine *e, const int32_t *sys, int32_t nsys, int32_t *lanes,
                        int32_t *grid);
/* Device self test of the wave-level helpers (DPP scan and group sums,
 * ds_swizzle, ballot prefix): bad64[lane] != 0 flags a mismatch; normals192
 * receives 64 x 3 device normals for key 0x9abcdef012345678, bead = lane,
 * step = 1000 + lane. */
sop_status sop_selftest(sop_engine *e, int32_t *bad64, float *normals192);

/* --- serving: continuous batching ---------------------------------------- */
typedef struct sop_serve_opts {
  int64_t hbm_budget_bytes; /* admission budget for job state (0 = 8 GiB) */
  int32_t max_jobs;         /* resident job slots (0 = 4096) */
  int32_t block_steps;      /* steps per graph replay (0 = n_nl of the first job) */
} sop_serve_opts;

typedef struct sop_job_result {
  int32_t done;          /* 1 when finished and copied back */
  int32_t status;        /* SOP_OK or the job's error */
  int32_t n_beads;
  int64_t steps;
  double t_submit, t_admit, t_done; /* seconds on the engine clock */
  double energies[SOP_E_COUNT];
} sop_job_result;

sop_status sop_serve_start(sop_engine *e, const sop_serve_opts *o);
/* Copies the description; the job runs nsteps from its step0.  All jobs of a
 * serve session share n_nl and n_pl, and step0 is a multiple of the block.
 * Called from any thread; returns a status code only (sop_last_error is the
 * serve thread's). */
sop_status sop_serve_submit(sop_engine *e, const sop_system_desc *d, int64_t nsteps,
                            int64_t *job);
/* Non-blocking; when done, xyz (3n floats, may be NULL) receives the final
 * unwrapped coordinates. SOP_ERR_BUSY while the job is queued or running. */
sop_status sop_serve_poll(sop_engine *e, int64_t job, sop_job_result *r, float *xyz);
sop_status sop_serve_stop(sop_engine *e);
/* Device bytes a job holds while resident (the admission budget's unit). */
int64_t sop_serve_job_bytes(const sop_engine *e, const sop_system_desc *d);
/* Counters of the serve loop since sop_serve_start (no reference counterpart):
 * where the loop's time goes, to tell a device-bound server from a host-bound
 * one. */
typedef struct sop_serve_stats {
  int64_t iterations;     /* admission + block + retirement rounds */
  int64_t jobs_admitted, jobs_retired;
  int64_t steps_run;      /* sum over iterations of the block length */
  double gpu_block_s;     /* device time of the step blocks (HIP events) */
  double host_admit_s;    /* uploads (beside the block in flight) and first lists of admitted jobs */
  double host_batch_s;    /* work tables of the running batch */
  double host_block_s;    /* wall time around the step blocks */
  double host_retire_s;   /* energies, copies back, frees */
  double mean_running;    /* running jobs per iteration, averaged */
  int64_t jobs_uploaded_beside; /* admissions uploaded while a block ran */
  int64_t jobs_deferred;  /* admissions put off until the engine was idle (a pool had to grow) */
  int64_t bead_steps;     /* bead-steps run: sum over blocks and running jobs of beads x steps
                             the job advanced in the block (finished or not) */
} sop_serve_stats;
sop_status sop_serve_stats_get(sop_engine *e, sop_serve_stats *out);
double sop_clock_seconds(void);

/* --- host-only utilities (no device needed) ------------------------------ */
typedef struct sop_gen_opts {
  float bond;         /* walk step, A (3.8) */
  float min_dist;     /* reject beads closer than this to a non-bonded bead (4.0) */
  float volume_per_bead; /* compact target: confining sphere of n*v volume (140 A^3) */
  int32_t relax_steps;   /* steepest-descent steps after the walk (0) */
} sop_gen_opts;
void sop_gen_opts_default(sop_gen_opts *o);
/* Seeded compact self-avoiding walk: xyz (3n), bond_r0 (n; r0 of (i,i+1)). */
sop_status sop_generate_chain(uint64_t seed, int32_t n, const sop_gen_opts *o, float *xyz,
                              float *bond_r0);
/* Native map: pairs with |i-j| > 2 in a chain (and, unless intra_only, pairs
 * on different chains) closer than cutoff; r0 = their distance.  Pairs come
 * out sorted (i, then j).  Two-call: pass cap = 0 to get *n_nat. */
sop_status sop_native_contacts(int32_t n, const float *xyz, const int32_t *chain, float cutoff,
                               int32_t intra_only, int32_t *nat_i, int32_t *nat_j, float *nat_r0,
                               int64_t cap, int64_t *n_nat);

/* Reference input files (io.cpp:112-259): bonds, angles, vdw, init.  The
 * topology object owns the arrays; sop_topology_desc points into it. */
typedef struct sop_topology sop_topology;
sop_status sop_topology_load(const char *bonds, const char *angles, const char *vdw,
                             const char *init, sop_topology **out);
sop_status sop_topology_desc(const sop_topology *t, sop_system_desc *d);
void sop_topology_free(sop_topology *t);
/* Writers in the reference formats: record_traj appends "%f %f %f" lines
 * (io.cpp:261-282); save_coords writes "%d %f %f %f" (io.cpp:284-305). */
sop_status sop_write_traj(const char *path, int32_t n, const float *xyz);
sop_status sop_save_coords(const char *path, int32_t n, const float *xyz);
sop_status sop_load_coords(const char *path, int32_t n, float *xyz);

#ifdef __cplusplus
}
#endif

#endif /* SOP_H_ */
