"""sop_amd -- Python mirror of libsop (the MI355X SOP overdamped dynamics engine).

The reference (samchogroup/SOP) drives one system through globals and the
entry points of energy.h:13-24 and sop.h:27-35.  Here the same operations are
methods of an Engine bound to one HIP device, acting on systems (replicas)
identified by an integer handle:

    reference                         here
    load bonds/angles/vdw/init        System.from_reference_files / System(...)
    alloc_GPU_arrays, setup_rng       Engine.add(system)
    update_neighbor_list / _cell_list Engine.update_neighbor_list(sys)
    update_pair_list                  Engine.update_pair_list(sys)
    force_eval (+ random_force)       Engine.force_eval(sys, with_noise)
    energy_eval                       Engine.energy_eval(sys)
    calculate_observables             Engine.observables(sys)
    overdamped_ctrl loop              Engine.run(systems, nsteps)

Every device call goes through libsop.so; if it is missing, importing the
engine raises ImportError (there is no CPU fallback).
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _abi
from ._abi import (E_ANGLE, E_BOND, E_COUNT, E_NATIVE, E_REP, E_TOTAL, E_VIRIAL, LIST_NL_NATIVE,
                   LIST_NL_REP, LIST_PL_NATIVE, LIST_PL_REP, Params)

__all__ = [
    "Engine", "System", "SopError", "Params", "params_reference", "params_config", "params_derived",
    "generate_chain", "native_contacts", "make_chain_system", "E_BOND", "E_ANGLE", "E_NATIVE",
    "E_REP", "E_TOTAL", "E_VIRIAL", "LIST_NL_NATIVE", "LIST_NL_REP", "LIST_PL_NATIVE", "LIST_PL_REP",
    "lib",
]


def lib():
    return _abi.load()


class SopError(RuntimeError):
    def __init__(self, status, msg=""):
        super().__init__(f"libsop status {status}: {msg}")
        self.status = status


def _check(st, eng=None):
    if st != _abi.SOP_OK:
        msg = ""
        if eng is not None and eng._h:
            msg = lib().sop_last_error(eng._h).decode(errors="replace")
        raise SopError(st, msg)


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


def params_reference() -> Params:
    p = Params()
    _check(lib().sop_params_reference(C.byref(p)))
    return p


def params_config() -> Params:
    p = Params()
    _check(lib().sop_params_config(C.byref(p)))
    return p


def params_derived(p: Params):
    a5, var = C.c_float(), C.c_float()
    _check(lib().sop_params_derived(C.byref(p), C.byref(a5), C.byref(var)))
    return np.float32(a5.value), np.float32(var.value)


def params_underdamped(p: Params):
    """a1..a4 of the underdamped integrator as float32 (param.cpp:104-108)."""
    a = (C.c_float * 4)()
    _check(lib().sop_params_underdamped(C.byref(p), a))
    return np.array(a[:], np.float32)


def copy_params(p: Params) -> Params:
    q = Params()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(Params))
    return q


@dataclass
class System:
    """One system: chains of beads with a native map (0-based indices)."""
    xyz: np.ndarray                      # (n,3) float32, unwrapped
    params: Params
    chain: Optional[np.ndarray] = None   # (n,) int32
    type: Optional[np.ndarray] = None    # (n,) int32 in {1,2}
    bond_r0: Optional[np.ndarray] = None # (n,) float32, r0 of bond (i,i+1)
    nat_i: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    nat_j: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    nat_r0: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    step0: int = 0
    vel: Optional[np.ndarray] = None     # (n,3) float32 underdamped initial velocities

    def __post_init__(self):
        self.xyz = np.ascontiguousarray(self.xyz, dtype=np.float32).reshape(-1, 3)
        n = self.xyz.shape[0]
        self.chain = np.zeros(n, np.int32) if self.chain is None else np.ascontiguousarray(self.chain, np.int32)
        self.type = np.ones(n, np.int32) if self.type is None else np.ascontiguousarray(self.type, np.int32)
        if self.bond_r0 is None:
            self.bond_r0 = np.full(n, 3.8, np.float32)
        self.bond_r0 = np.ascontiguousarray(self.bond_r0, np.float32)
        self.nat_i = np.ascontiguousarray(self.nat_i, np.int32)
        self.nat_j = np.ascontiguousarray(self.nat_j, np.int32)
        self.nat_r0 = np.ascontiguousarray(self.nat_r0, np.float32)
        if self.vel is not None:
            self.vel = np.ascontiguousarray(self.vel, np.float32).reshape(-1, 3)

    @property
    def n(self):
        return self.xyz.shape[0]

    def desc(self) -> _abi.SystemDesc:
        d = _abi.SystemDesc()
        d.n_beads = self.n
        d.xyz = _ptr(self.xyz, C.c_float)
        d.chain = _ptr(self.chain, C.c_int32)
        d.type = _ptr(self.type, C.c_int32)
        d.bond_r0 = _ptr(self.bond_r0, C.c_float)
        d.n_native = len(self.nat_i)
        d.nat_i = _ptr(self.nat_i, C.c_int32)
        d.nat_j = _ptr(self.nat_j, C.c_int32)
        d.nat_r0 = _ptr(self.nat_r0, C.c_float)
        d.params = self.params
        d.step0 = self.step0
        d.vel = _ptr(self.vel, C.c_float) if self.vel is not None else None
        return d

    @staticmethod
    def from_reference_files(bonds, angles, vdw, init):
        """Read the reference's input files (io.cpp:112-259)."""
        h = C.c_void_p()
        _check(lib().sop_topology_load(bonds.encode(), angles.encode() if angles else None, vdw.encode(),
                                       init.encode(), C.byref(h)))
        try:
            d = _abi.SystemDesc()
            _check(lib().sop_topology_desc(h, C.byref(d)))
            n, m = d.n_beads, d.n_native
            arr = lambda p, k, dt: np.ctypeslib.as_array(p, shape=(k,)).astype(dt).copy() if k else np.zeros(0, dt)
            return System(xyz=arr(d.xyz, 3 * n, np.float32).reshape(n, 3), params=copy_params(d.params),
                          chain=arr(d.chain, n, np.int32), type=arr(d.type, n, np.int32),
                          bond_r0=arr(d.bond_r0, n, np.float32), nat_i=arr(d.nat_i, m, np.int32),
                          nat_j=arr(d.nat_j, m, np.int32), nat_r0=arr(d.nat_r0, m, np.float32))
        finally:
            lib().sop_topology_free(h)


def generate_chain(seed: int, n: int, volume_per_bead=140.0, bond=3.8, min_dist=4.0):
    """Seeded compact self-avoiding walk; returns (xyz (n,3) float32, bond_r0 (n,))."""
    o = _abi.GenOpts()
    lib().sop_gen_opts_default(C.byref(o))
    o.volume_per_bead, o.bond, o.min_dist = volume_per_bead, bond, min_dist
    xyz = np.zeros((n, 3), np.float32)
    b = np.zeros(n, np.float32)
    _check(lib().sop_generate_chain(seed, n, C.byref(o), _ptr(xyz, C.c_float), _ptr(b, C.c_float)))
    return xyz, b


def native_contacts(xyz, chain=None, cutoff=8.0, intra_only=False):
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    n = xyz.shape[0]
    ch = None if chain is None else np.ascontiguousarray(chain, np.int32)
    cnt = C.c_int64()
    _check(lib().sop_native_contacts(n, _ptr(xyz, C.c_float), _ptr(ch, C.c_int32), cutoff, int(intra_only),
                                     None, None, None, 0, C.byref(cnt)))
    m = cnt.value
    ni, nj, nr = np.zeros(m, np.int32), np.zeros(m, np.int32), np.zeros(m, np.float32)
    _check(lib().sop_native_contacts(n, _ptr(xyz, C.c_float), _ptr(ch, C.c_int32), cutoff, int(intra_only),
                                     _ptr(ni, C.c_int32), _ptr(nj, C.c_int32), _ptr(nr, C.c_float), m,
                                     C.byref(cnt)))
    return ni, nj, nr


def make_chain_system(seed: int, n: int, params: Optional[Params] = None, native_cutoff=8.0, **gen):
    """Synthetic workload unit: compact walk + native map of its own state."""
    xyz, b = generate_chain(seed, n, **gen)
    ni, nj, nr = native_contacts(xyz, None, native_cutoff)
    p = copy_params(params if params is not None else params_config())
    return System(xyz=xyz, params=p, bond_r0=b, nat_i=ni, nat_j=nj, nat_r0=nr)


class Engine:
    """One HIP device's engine (libsop).  Not thread-safe."""

    def __init__(self, device=0, lanes_per_bead=4, grid_blocks=0, use_graphs=True, nl_cap_max=0,
                 pl_cap_max=0, cell_cap_max=0, pool_bytes=0, lds_beads_max=0, pl_shell=None, step_tiles=0,
                 step_blocks=0, step_tail=None):
        self._h = None
        o = _abi.EngineOpts()
        lib().sop_engine_opts_default(C.byref(o))
        o.lanes_per_bead = lanes_per_bead
        o.grid_blocks = grid_blocks
        o.use_graphs = int(use_graphs)
        if nl_cap_max:
            o.nl_cap_max = nl_cap_max
        if pl_cap_max:
            o.pl_cap_max = pl_cap_max
        if cell_cap_max:
            o.cell_cap_max = cell_cap_max
        o.pool_bytes = pool_bytes
        if lds_beads_max:
            o.lds_beads_max = lds_beads_max
        if pl_shell is not None:
            o.pl_shell = pl_shell  # < 0: no shell (every pair-list rebuild scans the cell grid)
        o.step_tiles = step_tiles  # tiles per step-kernel item (0: by batch size)
        o.step_blocks = step_blocks  # step-kernel grid in blocks (0: 4 per compute unit, 2 while serving)
        if step_tail is not None:
            o.step_tail = step_tail  # 2 or 4: late step items cut finer (0: off; default 2)
        h = C.c_void_p()
        _check(lib().sop_engine_create(device, C.byref(o), C.byref(h)))
        self._h = h
        self._n = {}

    def close(self):
        if self._h:
            lib().sop_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, st):
        _check(st, self)

    def add(self, system: System) -> int:
        d = system.desc()
        s = C.c_int32()
        self._c(lib().sop_system_add(self._h, C.byref(d), C.byref(s)))
        self._n[s.value] = system.n
        return s.value

    def remove(self, sys: int):
        self._c(lib().sop_system_remove(self._h, sys))
        self._n.pop(sys, None)

    def step_of(self, sys):
        n, st = C.c_int32(), C.c_int64()
        self._c(lib().sop_system_info(self._h, sys, C.byref(n), C.byref(st)))
        return st.value

    def update_neighbor_list(self, sys):
        self._c(lib().sop_update_neighbor_list(self._h, sys))

    def update_pair_list(self, sys):
        self._c(lib().sop_update_pair_list(self._h, sys))

    def get_list(self, sys, which) -> np.ndarray:
        n = C.c_int64()
        self._c(lib().sop_get_list(self._h, sys, which, None, 0, C.byref(n)))
        out = np.zeros((n.value, 2), np.int32)
        self._c(lib().sop_get_list(self._h, sys, which, _ptr(out, C.c_int32), n.value, C.byref(n)))
        return out

    def force_eval(self, sys, with_noise=False) -> np.ndarray:
        self._c(lib().sop_force_eval(self._h, sys, int(with_noise)))
        f = np.zeros((self._n[sys], 3), np.float32)
        self._c(lib().sop_get_forces(self._h, sys, _ptr(f, C.c_float)))
        return f

    def energy_eval(self, sys) -> np.ndarray:
        e = np.zeros(E_COUNT, np.float64)
        self._c(lib().sop_energy_eval(self._h, sys, _ptr(e, C.c_double)))
        return e

    def observables(self, sys):
        q, c, r = C.c_double(), C.c_int32(), C.c_double()
        self._c(lib().sop_observables(self._h, sys, C.byref(q), C.byref(c), C.byref(r)))
        return q.value, c.value, r.value

    def run(self, systems: Sequence[int], nsteps: int):
        a = np.ascontiguousarray(np.atleast_1d(systems), np.int32)
        self._c(lib().sop_run(self._h, _ptr(a, C.c_int32), len(a), nsteps))

    def run_plain(self, systems: Sequence[int], nsteps: int):
        a = np.ascontiguousarray(np.atleast_1d(systems), np.int32)
        self._c(lib().sop_run_plain(self._h, _ptr(a, C.c_int32), len(a), nsteps))

    def coords(self, sys) -> np.ndarray:
        x = np.zeros((self._n[sys], 3), np.float32)
        st = C.c_int64()
        self._c(lib().sop_get_coords(self._h, sys, _ptr(x, C.c_float), C.byref(st)))
        return x

    def wrapped_coords(self, sys) -> np.ndarray:
        x = np.zeros((self._n[sys], 3), np.float32)
        self._c(lib().sop_get_wrapped_coords(self._h, sys, _ptr(x, C.c_float)))
        return x

    def set_rng(self, sys, seed, step):
        """Philox key and step counter of the next step (an exact restart)."""
        self._c(lib().sop_set_rng(self._h, sys, seed, step))

    def velocities(self, sys) -> np.ndarray:
        v = np.zeros((self._n[sys], 3), np.float32)
        self._c(lib().sop_get_velocities(self._h, sys, _ptr(v, C.c_float)))
        return v

    def increments(self, sys) -> np.ndarray:
        d = np.zeros((self._n[sys], 3), np.float32)
        self._c(lib().sop_get_increments(self._h, sys, _ptr(d, C.c_float)))
        return d

    def set_coords(self, sys, xyz):
        x = np.ascontiguousarray(xyz, np.float32)
        self._c(lib().sop_set_coords(self._h, sys, _ptr(x, C.c_float)))

    def list_stats(self, sys) -> _abi.ListStats:
        s = _abi.ListStats()
        self._c(lib().sop_list_stats_get(self._h, sys, C.byref(s)))
        return s

    def rows(self, sys):
        """(row capacity, longest row so far) of a system's shell rows and
        pair-list tiles, in entries per bead."""
        cap, longest = C.c_int32(), C.c_int32()
        self._c(lib().sop_system_rows(self._h, sys, C.byref(cap), C.byref(longest)))
        return cap.value, longest.value

    def reserve_rows(self, sys, capacity: int):
        """Set a system's row capacity (kept within [longest row, min(n - 1,
        pl_cap_max)]); results do not change."""
        self._c(lib().sop_system_reserve_rows(self._h, sys, capacity))

    def time_step_kernel(self, systems, reps=20) -> float:
        a = np.ascontiguousarray(np.atleast_1d(systems), np.int32)
        ms = C.c_double()
        self._c(lib().sop_time_step_kernel(self._h, _ptr(a, C.c_int32), len(a), reps, C.byref(ms)))
        return ms.value

    def autotune_ex(self, systems) -> dict:
        """Settings chosen by timing whole graph blocks of these systems
        (sop_autotune_ex); the systems' state is left as it was."""
        a = np.ascontiguousarray(np.atleast_1d(systems), np.int32)
        r = _abi.TuneResult()
        self._c(lib().sop_autotune_ex(self._h, _ptr(a, C.c_int32), len(a), C.byref(r)))
        return self._tune_dict(r)

    def tune_set(self, tuned: dict):
        """Apply settings returned by autotune_ex (sop_tune_set)."""
        r = _abi.TuneResult()
        r.lanes_per_bead, r.step_blocks, r.lds_beads_max = (tuned["lanes_per_bead"], tuned["step_blocks"],
                                                            tuned["lds_beads_max"])
        nb = len(tuned["bucket_max_beads"])
        r.buckets = nb
        for b in range(nb):
            r.bucket_max_beads[b] = tuned["bucket_max_beads"][b]
            r.step_tiles[b] = tuned["step_tiles"][b]
            r.pl_shell[b] = tuned["pl_shell"][b]
        r.step_tail = tuned.get("step_tail", 0)
        self._c(lib().sop_tune_set(self._h, C.byref(r)))

    @staticmethod
    def _tune_dict(r) -> dict:
        nb = r.buckets
        return {"lanes_per_bead": r.lanes_per_bead, "step_blocks": r.step_blocks, "lds_beads_max": r.lds_beads_max,
                "bucket_max_beads": list(r.bucket_max_beads[:nb]), "step_tiles": list(r.step_tiles[:nb]),
                "pl_shell": [round(v, 3) for v in r.pl_shell[:nb]], "block_steps": r.block_steps,
                "block_ms_start": r.block_ms_start, "block_ms_best": r.block_ms_best, "trials": r.trials,
                "step_tail": r.step_tail}

    def autotune(self, systems):
        a = np.ascontiguousarray(np.atleast_1d(systems), np.int32)
        l, g = C.c_int32(), C.c_int32()
        self._c(lib().sop_autotune(self._h, _ptr(a, C.c_int32), len(a), C.byref(l), C.byref(g)))
        return l.value, g.value

    def selftest(self):
        bad = np.zeros(64, np.int32)
        nrm = np.zeros((64, 3), np.float32)
        self._c(lib().sop_selftest(self._h, _ptr(bad, C.c_int32), _ptr(nrm, C.c_float)))
        return bad, nrm

    def synchronize(self):
        self._c(lib().sop_synchronize(self._h))
