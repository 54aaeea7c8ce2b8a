"""libsop_refshim (the reference's energy.h / sop.h / GPUvars.h entry points)
and the sop2 command-file driver, without a device:

- the shim exports every function include/sop_refshim.h declares and leaves
  the reference's globals to the driver (undefined in the library);
- a driver written like sop.cu links against it and, with no GPU, stops with
  the reference's exit(-1) and a message (no CPU fallback);
- sop2 parses the reference's command grammar with its order-dependent
  `set` semantics (param.cpp:233-394) and refuses to run without a GPU.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import sop_amd as S
from reffiles import ROOT, SOP_DIR, build_shim_driver, write_command_file, write_reference_files

SHIM = os.path.join(SOP_DIR, "libsop_refshim.so")
SOP2 = os.path.join(SOP_DIR, "bin", "sop2")


def nm(lib, flag):
    out = subprocess.run(["nm", "-D", flag, lib], check=True, capture_output=True, text=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


def test_shim_exports_the_reference_entry_points():
    src = open(os.path.join(ROOT, "include", "sop_refshim.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    declared = set(re.findall(r"^void\s+(\w+)\s*\(", src, flags=re.M))
    assert {"force_eval", "energy_eval", "update_pair_list", "overdamped_iteration", "underdamped_iteration",
            "alloc_GPU_arrays", "setup_rng"} <= declared
    defined = nm(SHIM, "--defined-only")
    assert declared <= defined, declared - defined
    undefined = nm(SHIM, "--undefined-only")
    assert {"unc_pos", "force", "nbead", "idx_bead_lj_nat", "a5", "e_bnd"} <= undefined


def test_driver_links_and_fails_loudly_without_a_device(tmp_path):
    exe = build_shim_driver(tmp_path)
    write_reference_files(str(tmp_path), S.make_chain_system(4, 40))
    r = subprocess.run([exe, str(tmp_path), "2", "5", "50", "0.02", "0.6", "50", "1234", str(tmp_path / "o")],
                       capture_output=True, text=True, env={**os.environ, "HIP_VISIBLE_DEVICES": "-1"})
    assert r.returncode == 255, (r.returncode, r.stderr)
    assert "sop_refshim: sop_engine_create" in r.stderr


def dry_run(tmp_path, lines):
    cf = tmp_path / "input"
    write_command_file(cf, lines)
    r = subprocess.run([SOP2, str(cf), "--dry-run"], capture_output=True, text=True, check=True)
    return dict(ln.split(" ", 1) for ln in r.stdout.splitlines())


def test_sop2_parses_reference_commands(tmp_path):
    p = dry_run(tmp_path, ["load bonds b.dat", "load angles a.dat", "load vdw v.dat", "load init i.dat",
                           "set dynamics overdamped", "set temp 0.75", "set nstep 2000", "set nup 100",
                           "set cutofftype neighborlist", "set nnlup 25", "set zeta 50", "set boxl 400",
                           "set ufname u.out", "run"])
    assert p["sim_type"] == "2" and float(p["T"]) == np.float32(0.75)
    assert float(p["h"]) == np.float32(0.02)  # the overdamped default time step
    assert float(p["zeta"]) == 50 and p["nnlup"] == "25" and p["nup"] == "100" and p["nstep"] == "2000"
    assert p["cutofftype"] == "neighborlist" and p["ufname"] == "u.out" and p["bonds"] == "b.dat"
    assert p["npl"] == "1" and p["seed"] == "1234"  # pair list every step, sop.cu:138 seed


def test_sop2_zeta_resets_the_time_step(tmp_path):
    # param.cpp:377-386: `set zeta` puts h back to the dynamics default;
    # `set t_step` after it sticks
    p = dry_run(tmp_path, ["set dynamics underdamped", "set t_step 0.01", "set zeta 0.5", "run"])
    assert float(p["h"]) == np.float32(2.5e-3)
    p = dry_run(tmp_path, ["set dynamics underdamped", "set zeta 0.5", "set t_step 0.01", "run"])
    assert float(p["h"]) == np.float32(0.01)
    # several commands on one line, a command over several lines
    cf = tmp_path / "input2"
    cf.write_text("set temp 0.9; set nup\n77;\nrun;\n")
    r = subprocess.run([SOP2, str(cf), "--dry-run"], capture_output=True, text=True, check=True)
    p = dict(ln.split(" ", 1) for ln in r.stdout.splitlines())
    assert float(p["T"]) == np.float32(0.9) and p["nup"] == "77"


def test_sop2_without_a_device_fails(tmp_path):
    write_reference_files(str(tmp_path), S.make_chain_system(4, 40))
    cf = tmp_path / "input"
    write_command_file(cf, [f"load bonds {tmp_path}/bonds.dat", f"load angles {tmp_path}/angles.dat",
                            f"load vdw {tmp_path}/vdw.dat", f"load init {tmp_path}/init.dat",
                            "set dynamics overdamped", "set nstep 10", "run"])
    r = subprocess.run([SOP2, str(cf)], capture_output=True, text=True, cwd=str(tmp_path),
                       env={**os.environ, "HIP_VISIBLE_DEVICES": "-1"})
    assert r.returncode != 0 and "sop_engine_create" in r.stderr
