"""ctypes declarations of libsop's C ABI (include/sop.h).

Loading fails loudly when libsop.so has not been built: there is no CPU
fallback for any device entry point.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsop.so")

SOP_OK = 0
SOP_ERR_INVALID = -1
SOP_ERR_NO_DEVICE = -2
SOP_ERR_HIP = -3
SOP_ERR_CAPACITY = -4
SOP_ERR_NUMERIC = -5
SOP_ERR_STATE = -6
SOP_ERR_IO = -7
SOP_ERR_BUSY = -8
SOP_ERR_PBC = -9

E_BOND, E_ANGLE, E_NATIVE, E_REP, E_TOTAL, E_VIRIAL = range(6)
E_COUNT = 6
LIST_NL_NATIVE, LIST_NL_REP, LIST_PL_NATIVE, LIST_PL_REP = range(4)
DYN_OVERDAMPED, DYN_UNDERDAMPED = 0, 1

F3x3 = (C.c_float * 3) * 3


class Params(C.Structure):
    _fields_ = [
        ("k_bnd", C.c_float), ("R0", C.c_float), ("eps_ss", C.c_float), ("sigma_ss", C.c_float),
        ("eps_att", F3x3), ("eps_rep", F3x3), ("sigma_rep", F3x3),
        ("rep_c12", C.c_float), ("rep_c6", C.c_float),
        ("nl_factor", C.c_double), ("nl_abs", C.c_float),
        ("pl_factor", C.c_double), ("pl_abs", C.c_float),
        ("n_nl", C.c_int32), ("n_pl", C.c_int32),
        ("T", C.c_float), ("zeta", C.c_float), ("h", C.c_float), ("boxl", C.c_float),
        ("seed", C.c_uint64),
        ("dynamics", C.c_int32), ("reserved", C.c_int32),
    ]


class SystemDesc(C.Structure):
    _fields_ = [
        ("n_beads", C.c_int32),
        ("xyz", C.POINTER(C.c_float)),
        ("chain", C.POINTER(C.c_int32)),
        ("type", C.POINTER(C.c_int32)),
        ("bond_r0", C.POINTER(C.c_float)),
        ("n_native", C.c_int32),
        ("nat_i", C.POINTER(C.c_int32)),
        ("nat_j", C.POINTER(C.c_int32)),
        ("nat_r0", C.POINTER(C.c_float)),
        ("params", Params),
        ("step0", C.c_int64),
        ("vel", C.POINTER(C.c_float)),
    ]


class EngineOpts(C.Structure):
    _fields_ = [
        ("nl_cap_max", C.c_int32), ("pl_cap_max", C.c_int32), ("cell_cap_max", C.c_int32),
        ("lanes_per_bead", C.c_int32), ("grid_blocks", C.c_int32), ("use_graphs", C.c_int32),
        ("lds_beads_max", C.c_int32), ("pl_shell", C.c_float), ("pool_bytes", C.c_int64),
        ("step_tiles", C.c_int32), ("step_blocks", C.c_int32),
        ("step_tail", C.c_int32), ("reserved_opts", C.c_int32),
    ]


class ListStats(C.Structure):
    _fields_ = [
        ("nl_native", C.c_int64), ("nl_rep", C.c_int64), ("pl_native", C.c_int64), ("pl_rep", C.c_int64),
        ("nl_max_per_bead", C.c_int32), ("pl_max_per_bead", C.c_int32),
        ("grid_builds", C.c_int64), ("shell_builds", C.c_int64), ("nl_builds", C.c_int64),
    ]


class ServeOpts(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_jobs", C.c_int32), ("block_steps", C.c_int32)]


class JobResult(C.Structure):
    _fields_ = [
        ("done", C.c_int32), ("status", C.c_int32), ("n_beads", C.c_int32), ("steps", C.c_int64),
        ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
        ("energies", C.c_double * E_COUNT),
    ]


class ServeStats(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("jobs_admitted", C.c_int64), ("jobs_retired", C.c_int64),
                ("steps_run", C.c_int64), ("gpu_block_s", C.c_double), ("host_admit_s", C.c_double),
                ("host_batch_s", C.c_double), ("host_block_s", C.c_double), ("host_retire_s", C.c_double),
                ("mean_running", C.c_double), ("jobs_uploaded_beside", C.c_int64), ("jobs_deferred", C.c_int64),
                ("bead_steps", C.c_int64)]


class TuneResult(C.Structure):
    _fields_ = [("lanes_per_bead", C.c_int32), ("step_blocks", C.c_int32), ("lds_beads_max", C.c_int32),
                ("buckets", C.c_int32), ("bucket_max_beads", C.c_int32 * 3), ("step_tiles", C.c_int32 * 3),
                ("pl_shell", C.c_float * 3), ("block_steps", C.c_int32), ("block_ms_start", C.c_double),
                ("block_ms_best", C.c_double), ("trials", C.c_int32), ("step_tail", C.c_int32)]


class GenOpts(C.Structure):
    _fields_ = [("bond", C.c_float), ("min_dist", C.c_float), ("volume_per_bead", C.c_float),
                ("relax_steps", C.c_int32)]


P = C.POINTER
i32, i64, u64, f32, f64 = C.c_int32, C.c_int64, C.c_uint64, C.c_float, C.c_double
vp = C.c_void_p

# name: (restype, argtypes)
SIGNATURES = {
    "sop_params_reference": (i32, [P(Params)]),
    "sop_params_config": (i32, [P(Params)]),
    "sop_params_derived": (i32, [P(Params), P(f32), P(f32)]),
    "sop_params_underdamped": (i32, [P(Params), P(f32)]),
    "sop_engine_opts_default": (None, [P(EngineOpts)]),
    "sop_engine_create": (i32, [i32, P(EngineOpts), P(vp)]),
    "sop_engine_destroy": (None, [vp]),
    "sop_last_error": (C.c_char_p, [vp]),
    "sop_version": (C.c_char_p, []),
    "sop_synchronize": (i32, [vp]),
    "sop_system_add": (i32, [vp, P(SystemDesc), P(i32)]),
    "sop_system_remove": (i32, [vp, i32]),
    "sop_system_info": (i32, [vp, i32, P(i32), P(i64)]),
    "sop_update_neighbor_list": (i32, [vp, i32]),
    "sop_update_pair_list": (i32, [vp, i32]),
    "sop_get_list": (i32, [vp, i32, i32, P(i32), i64, P(i64)]),
    "sop_force_eval": (i32, [vp, i32, i32]),
    "sop_get_forces": (i32, [vp, i32, P(f32)]),
    "sop_energy_eval": (i32, [vp, i32, P(f64)]),
    "sop_observables": (i32, [vp, i32, P(f64), P(i32), P(f64)]),
    "sop_run": (i32, [vp, P(i32), i32, i64]),
    "sop_run_plain": (i32, [vp, P(i32), i32, i64]),
    "sop_get_coords": (i32, [vp, i32, P(f32), P(i64)]),
    "sop_set_coords": (i32, [vp, i32, P(f32)]),
    "sop_get_wrapped_coords": (i32, [vp, i32, P(f32)]),
    "sop_get_velocities": (i32, [vp, i32, P(f32)]),
    "sop_set_rng": (i32, [vp, i32, C.c_uint64, C.c_int64]),
    "sop_get_increments": (i32, [vp, i32, P(f32)]),
    "sop_list_stats_get": (i32, [vp, i32, P(ListStats)]),
    "sop_system_rows": (i32, [vp, i32, P(C.c_int32), P(C.c_int32)]),
    "sop_system_reserve_rows": (i32, [vp, i32, i32]),
    "sop_time_step_kernel": (i32, [vp, P(i32), i32, i32, P(f64)]),
    "sop_autotune": (i32, [vp, P(i32), i32, P(i32), P(i32)]),
    "sop_autotune_ex": (i32, [vp, P(i32), i32, P(TuneResult)]),
    "sop_tune_set": (i32, [vp, P(TuneResult)]),
    "sop_energy_eval_noalloc": (i32, [vp, i32, P(f64)]),
    "sop_selftest": (i32, [vp, P(i32), P(f32)]),
    "sop_serve_start": (i32, [vp, P(ServeOpts)]),
    "sop_serve_submit": (i32, [vp, P(SystemDesc), i64, P(i64)]),
    "sop_serve_poll": (i32, [vp, i64, P(JobResult), P(f32)]),
    "sop_serve_stop": (i32, [vp]),
    "sop_serve_job_bytes": (i64, [vp, P(SystemDesc)]),
    "sop_serve_stats_get": (i32, [vp, P(ServeStats)]),
    "sop_clock_seconds": (f64, []),
    "sop_gen_opts_default": (None, [P(GenOpts)]),
    "sop_generate_chain": (i32, [u64, i32, P(GenOpts), P(f32), P(f32)]),
    "sop_native_contacts": (i32, [i32, P(f32), P(i32), f32, i32, P(i32), P(i32), P(f32), i64, P(i64)]),
    "sop_topology_load": (i32, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, P(vp)]),
    "sop_topology_desc": (i32, [vp, P(SystemDesc)]),
    "sop_topology_free": (None, [vp]),
    "sop_write_traj": (i32, [C.c_char_p, i32, P(f32)]),
    "sop_save_coords": (i32, [C.c_char_p, i32, P(f32)]),
    "sop_load_coords": (i32, [C.c_char_p, i32, P(f32)]),
}

_lib = None


def load():
    """Load libsop.so; raises ImportError if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.environ.get("SOP_LIBSOP", LIB_PATH)  # another build of libsop.so, for comparisons
    if not os.path.exists(path):
        raise ImportError(f"libsop.so not built at {path}: run __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
