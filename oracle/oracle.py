"""Python wrapper of the CPU oracle (liboracle.so built from sop_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product (sop_amd/).  See the header
of sop_oracle.c for the reference file:line each function restates.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        P = C.POINTER
        _lib.orc_philox.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
        _lib.orc_normals.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, P(C.c_double)]
        _lib.orc_min_image_f.argtypes = [C.c_float, C.c_float]
        _lib.orc_min_image_f.restype = C.c_float
        _lib.orc_cutoff.argtypes = [C.c_double, C.c_float, C.c_float]
        _lib.orc_cutoff.restype = C.c_float
        _lib.orc_dist2_f.argtypes = [P(C.c_float), C.c_int32, C.c_int32, C.c_float]
        _lib.orc_dist2_f.restype = C.c_float
        _lib.orc_nl.argtypes = [P(OrcSys), P(C.c_float), P(C.c_int32), C.c_int64, P(C.c_int64),
                                P(C.c_int32), P(C.c_int32), C.c_int64, P(C.c_int64)]
        _lib.orc_nl_grid.argtypes = _lib.orc_nl.argtypes
        _lib.orc_pl.argtypes = [P(OrcSys), P(C.c_float), P(C.c_int32), C.c_int64, P(C.c_int32), P(C.c_int32),
                                C.c_int64, P(C.c_int32), P(C.c_int64), P(C.c_int32), P(C.c_int32), P(C.c_int64)]
        _lib.orc_forces.argtypes = [P(OrcSys), P(C.c_double), P(C.c_int32), C.c_int64, P(C.c_int32),
                                    P(C.c_int32), C.c_int64, P(C.c_double), P(C.c_double)]
        _lib.orc_observables.argtypes = [P(OrcSys), P(C.c_float), P(C.c_double), P(C.c_double),
                                         P(C.c_int32), P(C.c_double)]
        _lib.orc_run.argtypes = [P(OrcSys), P(C.c_float), C.c_int64, C.c_int64, C.c_uint64, C.c_float,
                                 C.c_float, C.c_int32, C.c_int32, P(C.c_double)]
        _lib.orc_run_grid.argtypes = _lib.orc_run.argtypes
        _lib.orc_run_sample.argtypes = [P(OrcSys), P(C.c_float), C.c_int64, C.c_int64, C.c_uint64, C.c_float,
                                        C.c_float, C.c_int32, C.c_int32, C.c_int32, P(C.c_double), P(C.c_double)]
        _lib.orc_port_create.argtypes = [P(OrcSys)]
        _lib.orc_run_ud.argtypes = [P(OrcSys), P(C.c_float), P(C.c_float), C.c_int64, C.c_int64, C.c_uint64,
                                    P(C.c_float), C.c_float, C.c_int32, C.c_int32, C.c_int32,
                                    P(C.c_double), P(C.c_double)]
        _lib.orc_port_create.restype = C.c_void_p
        _lib.orc_port_destroy.argtypes = [C.c_void_p]
        _lib.orc_port_run.argtypes = [C.c_void_p, P(C.c_float), C.c_int64, C.c_int64, C.c_uint64, C.c_float,
                                      C.c_float, C.c_int32, C.c_int32]
        _lib.orc_port_master_size.argtypes = [C.c_void_p]
        _lib.orc_port_master_size.restype = C.c_int64
    return _lib


class OrcSys(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("chain", C.POINTER(C.c_int32)), ("type", C.POINTER(C.c_int32)), ("bond_r0", C.POINTER(C.c_float)),
        ("n_nat", C.c_int32),
        ("nat_i", C.POINTER(C.c_int32)), ("nat_j", C.POINTER(C.c_int32)), ("nat_r0", C.POINTER(C.c_float)),
        ("k_bnd", C.c_float), ("R0", C.c_float), ("eps_ss", C.c_float), ("sigma_ss", C.c_float),
        ("eps_att", C.c_float * 9), ("eps_rep", C.c_float * 9), ("sigma_rep", C.c_float * 9),
        ("rep_c12", C.c_float), ("rep_c6", C.c_float),
        ("nl_factor", C.c_double), ("nl_abs", C.c_float),
        ("pl_factor", C.c_double), ("pl_abs", C.c_float),
        ("boxl", C.c_float),
    ]


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(out, C.c_uint32))
    return out


def normals(seed, bead, step):
    out = np.zeros(3, np.float64)
    lib().orc_normals(seed, bead, step, _p(out, C.c_double))
    return out


def min_image_f(d, boxl):
    return np.float32(lib().orc_min_image_f(d, boxl))


def cutoff(factor, absval, base):
    return np.float32(lib().orc_cutoff(factor, absval, base))


class Oracle:
    """Oracle view of one system.  `system` is any object with the fields of
    sop_amd.System (xyz, chain, type, bond_r0, nat_i, nat_j, nat_r0, params)."""

    def __init__(self, system):
        self.n = int(system.xyz.shape[0])
        self.chain = np.ascontiguousarray(system.chain, np.int32)
        self.type = np.ascontiguousarray(system.type, np.int32)
        self.bond_r0 = np.ascontiguousarray(system.bond_r0, np.float32)
        self.nat_i = np.ascontiguousarray(system.nat_i, np.int32)
        self.nat_j = np.ascontiguousarray(system.nat_j, np.int32)
        self.nat_r0 = np.ascontiguousarray(system.nat_r0, np.float32)
        p = system.params
        s = OrcSys()
        s.n = self.n
        s.chain, s.type, s.bond_r0 = _p(self.chain, C.c_int32), _p(self.type, C.c_int32), _p(self.bond_r0, C.c_float)
        s.n_nat = len(self.nat_i)
        s.nat_i, s.nat_j, s.nat_r0 = _p(self.nat_i, C.c_int32), _p(self.nat_j, C.c_int32), _p(self.nat_r0, C.c_float)
        s.k_bnd, s.R0, s.eps_ss, s.sigma_ss = p.k_bnd, p.R0, p.eps_ss, p.sigma_ss
        for a in range(3):
            for b in range(3):
                s.eps_att[a * 3 + b] = p.eps_att[a][b]
                s.eps_rep[a * 3 + b] = p.eps_rep[a][b]
                s.sigma_rep[a * 3 + b] = p.sigma_rep[a][b]
        s.rep_c12, s.rep_c6 = p.rep_c12, p.rep_c6
        s.nl_factor, s.nl_abs, s.pl_factor, s.pl_abs = p.nl_factor, p.nl_abs, p.pl_factor, p.pl_abs
        s.boxl = p.boxl
        self.s = s
        self.params = p

    def nl(self, xyz, grid=False):
        """Neighbor list: (native indices k, rep pairs (m,2) with i<j, row-major).
        grid=False: brute force over all pairs (orc_nl); grid=True: the same
        arrays from candidate cells (orc_nl_grid, open boundaries only)."""
        x = np.ascontiguousarray(xyz, np.float32).reshape(-1)
        cap_a = max(1, len(self.nat_i))
        cap_r = max(1, 192 * self.n) if grid else max(1, self.n * (self.n - 1) // 2)
        fn = lib().orc_nl_grid if grid else lib().orc_nl
        while True:
            ak = np.zeros(cap_a, np.int32)
            ri = np.zeros(cap_r, np.int32)
            rj = np.zeros(cap_r, np.int32)
            na, nr = C.c_int64(), C.c_int64()
            st = fn(C.byref(self.s), _p(x, C.c_float), _p(ak, C.c_int32), cap_a, C.byref(na),
                    _p(ri, C.c_int32), _p(rj, C.c_int32), cap_r, C.byref(nr))
            if st == -1 and grid and nr.value > cap_r:
                cap_r = nr.value
                continue
            assert st == 0, st
            return ak[:na.value].copy(), np.stack([ri[:nr.value], rj[:nr.value]], 1)

    def pl(self, xyz, nl):
        x = np.ascontiguousarray(xyz, np.float32).reshape(-1)
        ak, rp = nl
        ak = np.ascontiguousarray(ak, np.int32)
        ri = np.ascontiguousarray(rp[:, 0], np.int32)
        rj = np.ascontiguousarray(rp[:, 1], np.int32)
        oa = np.zeros(max(1, len(ak)), np.int32)
        oi = np.zeros(max(1, len(ri)), np.int32)
        oj = np.zeros(max(1, len(ri)), np.int32)
        na, nr = C.c_int64(), C.c_int64()
        lib().orc_pl(C.byref(self.s), _p(x, C.c_float), _p(ak, C.c_int32), len(ak), _p(ri, C.c_int32),
                     _p(rj, C.c_int32), len(ri), _p(oa, C.c_int32), C.byref(na), _p(oi, C.c_int32),
                     _p(oj, C.c_int32), C.byref(nr))
        return oa[:na.value].copy(), np.stack([oi[:nr.value], oj[:nr.value]], 1)

    def lists(self, xyz, grid=False):
        nl = self.nl(xyz, grid)
        return nl, self.pl(xyz, nl)

    def forces(self, xyz, pl):
        """FP64 forces (n,3) and energies [bond, angle, native, rep, total, virial]."""
        x = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        ak, rp = pl
        ak = np.ascontiguousarray(ak, np.int32)
        ri = np.ascontiguousarray(rp[:, 0], np.int32)
        rj = np.ascontiguousarray(rp[:, 1], np.int32)
        f = np.zeros(3 * self.n, np.float64)
        e = np.zeros(6, np.float64)
        st = lib().orc_forces(C.byref(self.s), _p(x, C.c_double), _p(ak, C.c_int32), len(ak), _p(ri, C.c_int32),
                              _p(rj, C.c_int32), len(ri), _p(f, C.c_double), _p(e, C.c_double))
        return f.reshape(-1, 3), e, st

    def observables(self, xyz):
        xf = np.ascontiguousarray(xyz, np.float32).reshape(-1)
        xd = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        q, r = C.c_double(), C.c_double()
        c = C.c_int32()
        lib().orc_observables(C.byref(self.s), _p(xf, C.c_float), _p(xd, C.c_double), C.byref(q), C.byref(c),
                              C.byref(r))
        return q.value, c.value, r.value

    def run(self, x0, step0, nsteps, seed, a5, var, n_nl, n_pl, grid=False):
        """FP64 trajectory; returns final coordinates (n,3) float64.  grid=True
        builds the neighbor lists with orc_nl_grid (open boundaries only)."""
        x = np.ascontiguousarray(x0, np.float32).reshape(-1)
        out = np.zeros(3 * self.n, np.float64)
        fn = lib().orc_run_grid if grid else lib().orc_run
        st = fn(C.byref(self.s), _p(x, C.c_float), step0, nsteps, seed, a5, var, n_nl, n_pl,
                _p(out, C.c_double))
        assert st == 0, st
        return out.reshape(-1, 3)

    def run_sample(self, x0, step0, nsteps, seed, a5, var, n_nl, n_pl, nup):
        """orc_run recording (E_total, Rg^2, Q, contacts) after every nup-th
        step; returns (final coordinates (n,3) float64, samples (m,4))."""
        x = np.ascontiguousarray(x0, np.float32).reshape(-1)
        out = np.zeros(3 * self.n, np.float64)
        m = (step0 + nsteps) // nup - step0 // nup
        smp = np.zeros((max(1, m), 4), np.float64)
        st = lib().orc_run_sample(C.byref(self.s), _p(x, C.c_float), step0, nsteps, seed, a5, var, n_nl, n_pl,
                                  nup, _p(out, C.c_double), _p(smp, C.c_double))
        assert st == 0, st
        return out.reshape(-1, 3), smp[:m]

    def run_ud(self, x0, step0, nsteps, seed, a, var, vupd, n_nl, n_pl, v0=None):
        """FP64 underdamped trajectory; returns (coordinates, velocities after
        the last step), each (n,3) float64."""
        x = np.ascontiguousarray(x0, np.float32).reshape(-1)
        v = None if v0 is None else np.ascontiguousarray(v0, np.float32).reshape(-1)
        aa = np.ascontiguousarray(a, np.float32)
        out = np.zeros(3 * self.n, np.float64)
        vout = np.zeros(3 * self.n, np.float64)
        st = lib().orc_run_ud(C.byref(self.s), _p(x, C.c_float), _p(v, C.c_float) if v is not None else None,
                              step0, nsteps, seed, _p(aa, C.c_float), var, int(vupd), n_nl, n_pl,
                              _p(out, C.c_double), _p(vout, C.c_double))
        assert st == 0, st
        return out.reshape(-1, 3), vout.reshape(-1, 3)


class Port:
    """FP32 restatement of the reference's host method (CPU baseline)."""

    def __init__(self, system):
        self.o = Oracle(system)
        self.h = lib().orc_port_create(C.byref(self.o.s))
        assert self.h, "port allocation failed"
        self.x = np.ascontiguousarray(system.xyz, np.float32).reshape(-1).copy()

    def master_size(self):
        return lib().orc_port_master_size(self.h)

    def run(self, step0, nsteps, seed, a5, var, n_nl, n_pl):
        return lib().orc_port_run(self.h, _p(self.x, C.c_float), step0, nsteps, seed, a5, var, n_nl, n_pl)

    def close(self):
        if self.h:
            lib().orc_port_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
