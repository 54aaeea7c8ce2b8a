"""The reference's entry points (libsop_refshim) and the sop2 driver on device 0.

shim   a driver written like sop.cu runs the reference's control loop
       (overdamped_ctrl / underdamped_ctrl, sop.cu:104-417) through
       force_eval, update_*_list and *_iteration.  The oracle replays the same
       loop in FP64: lists rebuilt where the reference rebuilds them, forces of
       step k+1 on the pair list of step k, one Philox counter per force_eval.
       Coordinates within 1e-4 A RMSD after 300 steps; list sizes exact;
       energies within 1e-4 relative of the sum of term magnitudes.
sop2   a command file run end to end: update.out rows, trajectory frames,
       final coordinates within 1e-4 A of the oracle (pair list every step),
       and a restart from the saved files continues the same trajectory.
"""
import os
import subprocess

import numpy as np
import pytest

import sop_amd as S
from oracle import oracle as O
from oracle.oracle import Oracle
from reffiles import SOP_DIR, build_shim_driver, write_command_file, write_reference_files

pytestmark = pytest.mark.gpu

SOP2 = os.path.join(SOP_DIR, "bin", "sop2")


def load_system(d):
    return S.System.from_reference_files(*(os.path.join(d, f) for f in ("bonds.dat", "angles.dat", "vdw.dat",
                                                                       "init.dat")))


def reference_loop(s, sim_type, nstep, seed, nnlup, h, zeta, T):
    """FP64 replay of the reference's control loop on the oracle's lists and forces."""
    o = Oracle(s)
    n = s.n
    a5 = np.float64(np.float32(np.float32(h) / np.float32(zeta)))
    p = S.copy_params(s.params)
    p.h, p.zeta, p.T = h, zeta, T
    a = S.params_underdamped(p).astype(np.float64)
    _, var = S.params_derived(p)
    counter = [0]

    def force(x, pl):
        f, _, st = o.forces(x, pl)
        assert st == 0
        xi = np.array([O.normals(seed, i, counter[0]) for i in range(n)])
        counter[0] += 1
        return f + np.float64(var) * xi

    x = s.xyz.astype(np.float64)
    v = np.zeros_like(x)
    nl = o.nl(np.float32(x))
    pl = o.pl(np.float32(x), nl)
    f = force(x, pl)
    lup = 1 if sim_type == 1 else 0
    for _ in range(nstep):
        if lup % nnlup == 0:
            nl = o.nl(np.float32(x))
            lup = 0
        lup += 1
        pl = o.pl(np.float32(x), nl)
        if sim_type == 2:
            x = x + a5 * f
            f = force(x, pl)
        else:
            incr = a[0] * v + a[1] * f
            x = x + incr
            f = force(x, pl)
            if T >= 1e-5:
                v = a[2] * incr + a[3] * f
    e = o.forces(x, pl)[1]
    return x, v, (len(nl[0]), len(nl[1]), len(pl[0]), len(pl[1])), e


@pytest.mark.parametrize("sim_type,h,zeta", [(2, 0.02, 50.0), (1, 2.5e-3, 0.05)])
def test_shim_runs_the_reference_loop(tmp_path, sim_type, h, zeta):
    d = str(tmp_path)
    write_reference_files(d, S.make_chain_system(70 + sim_type, 100))
    s = load_system(d)
    exe = build_shim_driver(tmp_path)
    nstep, seed, nnlup, T = 300, 4321, 50, 0.6
    out = tmp_path / "out.txt"
    print(f"shim driver sim_type={sim_type}", flush=True)
    subprocess.run([exe, d, str(sim_type), str(nstep), str(zeta), str(h), str(T), str(nnlup), str(seed), str(out)],
                   check=True, timeout=300)
    rows = out.read_text().splitlines()
    x = np.array([r.split() for r in rows[:s.n]], np.float64)
    e = np.array(rows[s.n].split(), np.float64)
    counts = tuple(int(c) for c in rows[s.n + 1].split())
    v = np.array([r.split() for r in rows[s.n + 2:2 * s.n + 2]], np.float64)
    xr, vr, cr, er = reference_loop(s, sim_type, nstep, seed, nnlup, h, zeta, T)
    rmsd = float(np.sqrt(((x - xr) ** 2).sum(1).mean()))
    assert rmsd <= 1e-4, rmsd
    assert counts == cr
    scale = np.abs(er[:4]).sum()
    assert np.all(np.abs(e[:4] - er[:4]) <= 1e-4 * scale), (e, er)
    assert e[4] == pytest.approx(e[:4].sum(), rel=1e-5)
    if sim_type == 1:
        assert np.abs(v - vr).max() <= 1e-3 * np.abs(vr).max()
    else:
        assert np.all(v == 0)
    # the COMBINE'd pair list entries decode to pairs of the oracle's list
    ij = [(int(a) & 0x3FFF, int(b) & 0x3FFF) for a, b in (r.split() for r in rows[2 * s.n + 2:])]
    assert len(ij) == 5 and all(1 <= i < j <= s.n for i, j in ij)


def cli_files(d, nstep, nup, dyn="overdamped", extra=()):
    return [f"load bonds {d}/bonds.dat", f"load angles {d}/angles.dat", f"load vdw {d}/vdw.dat",
            f"load init {d}/init.dat", f"set dynamics {dyn}", "set temp 0.6", "set zeta 50",
            f"set nstep {nstep}", f"set nup {nup}", "set cutofftype neighborlist", "set nnlup 50",
            "set seed 99", *extra, "run"]


def test_sop2_run_and_restart(tmp_path):
    d = str(tmp_path)
    write_reference_files(d, S.make_chain_system(80, 128))
    s = load_system(d)
    full = tmp_path / "full"
    full.mkdir()
    write_command_file(full / "input", cli_files(d, 500, 100))
    print("sop2 500 steps", flush=True)
    subprocess.run([SOP2, "input"], cwd=str(full), check=True, timeout=300, capture_output=True)
    upd = (full / "update.out").read_text().splitlines()
    assert len(upd) == 5 and all(len(r.split()) == 10 for r in upd)
    assert [int(r.split()[0]) for r in upd] == [100, 200, 300, 400, 500]
    frames = (full / "traj.bin").read_text().splitlines()
    assert len(frames) == 6 * s.n and len((full / "traj_uncorrected.bin").read_text().splitlines()) == 6 * s.n
    x = np.loadtxt(full / "unccoord.out")[:, 1:]
    p = s.params
    p.T, p.zeta, p.h, p.n_nl, p.n_pl, p.seed = 0.6, 50.0, 0.02, 50, 1, 99
    a5, var = S.params_derived(p)
    o = Oracle(s)
    xr = o.run(s.xyz, 0, 500, 99, a5, var, 50, 1)
    assert np.sqrt(((x - xr) ** 2).sum(1).mean()) <= 1e-4
    last = np.array(upd[-1].split(), np.float64)
    er = o.forces(np.float32(x), o.lists(np.float32(x))[1])[1]
    assert last[3] == pytest.approx(er[0], abs=1e-3) and last[6] == pytest.approx(er[4], abs=0.05)
    assert 0.3 < last[2] < 1.2  # kinT of the overdamped increments near T
    # restart from the files of step 250
    half = tmp_path / "half"
    half.mkdir()
    write_command_file(half / "input", cli_files(d, 250, 250))
    subprocess.run([SOP2, "input"], cwd=str(half), check=True, timeout=300, capture_output=True)
    write_command_file(half / "input2", cli_files(d, 500, 250, extra=("set restart on", "set istep_restart 250")))
    print("sop2 restart", flush=True)
    subprocess.run([SOP2, "input2"], cwd=str(half), check=True, timeout=300, capture_output=True)
    xh = np.loadtxt(half / "unccoord.out")[:, 1:]
    assert np.sqrt(((xh - x) ** 2).sum(1).mean()) <= 1e-4
    assert int((half / "update.out").read_text().splitlines()[-1].split()[0]) == 500


def test_sop2_underdamped_rows(tmp_path):
    d = str(tmp_path)
    write_reference_files(d, S.make_chain_system(81, 64))
    write_command_file(tmp_path / "input", cli_files(d, 200, 50, dyn="underdamped", extra=("set zeta 0.05",)))
    subprocess.run([SOP2, "input"], cwd=d, check=True, timeout=300, capture_output=True)
    upd = (tmp_path / "update.out").read_text().splitlines()
    assert len(upd) == 4 and all(len(r.split()) == 12 for r in upd)
    v = np.loadtxt(tmp_path / "veloc.out")[:, 1:]
    assert v.shape == (64, 3) and np.abs(v).max() > 0
