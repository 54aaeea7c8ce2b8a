"""C ABI and host-side logic of libsop (CPU only: no compute call needs a GPU).

- libsop.so loads and exports every function include/sop.h declares;
- without a device the engine refuses to start (no CPU fallback);
- parameter sets, the seeded chain generator, native maps, and the readers /
  writers of the reference's file formats (io.cpp:112-408).
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import sop_amd as S
from reffiles import write_reference_files
from sop_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header="sop.h"):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sop_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(_abi.LIB_PATH)
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_covers_the_header():
    assert set(declared_functions()) == set(_abi.SIGNATURES)


def test_struct_sizes_match_c_layout(tmp_path):
    # the C compiler's sizes and offsets for include/sop.h against the ctypes mirror
    src = os.path.join(ROOT, "tests", "helpers", "abi_layout.c")
    exe = str(tmp_path / "abi_layout")
    subprocess.run(["gcc", "-std=c11", "-o", exe, src], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    structs = {"sop_params": _abi.Params, "sop_system_desc": _abi.SystemDesc, "sop_engine_opts": _abi.EngineOpts,
               "sop_serve_opts": _abi.ServeOpts, "sop_job_result": _abi.JobResult,
               "sop_list_stats": _abi.ListStats, "sop_gen_opts": _abi.GenOpts,
               "sop_serve_stats": _abi.ServeStats, "sop_tune_result": _abi.TuneResult}
    checked = 0
    for line in filter(None, out):
        name, val = line.split()
        if "." in name:
            st, fld = name.split(".")
            assert getattr(structs[st], fld).offset == int(val), name
        else:
            assert C.sizeof(structs[name]) == int(val), name
        checked += 1
    assert checked >= 20


def test_engine_without_device_fails_loudly():
    h = C.c_void_p()
    st = S.lib().sop_engine_create(0, None, C.byref(h))
    if st == _abi.SOP_OK:  # running on a GPU host
        S.lib().sop_engine_destroy(h)
        pytest.skip("a device is present")
    assert st == _abi.SOP_ERR_NO_DEVICE
    with pytest.raises(S.SopError):
        S.Engine(0)


@pytest.mark.parametrize("field,value", [("pl_cap_max", 4096), ("lanes_per_bead", 2), ("step_tiles", 3)])
def test_options_out_of_range_are_refused_before_the_device(field, value):
    # the step kernel holds a row's offset in its tile's region in 16 bits
    # (units of 4 entries): capacities above 4095 entries are refused
    o = _abi.EngineOpts()
    S.lib().sop_engine_opts_default(C.byref(o))
    setattr(o, field, value)
    h = C.c_void_p()
    assert S.lib().sop_engine_create(0, C.byref(o), C.byref(h)) == _abi.SOP_ERR_INVALID
    o2 = _abi.EngineOpts()
    S.lib().sop_engine_opts_default(C.byref(o2))
    o2.pl_cap_max = 4095
    st = S.lib().sop_engine_create(0, C.byref(o2), C.byref(h))
    if st == _abi.SOP_OK:
        S.lib().sop_engine_destroy(h)
    assert st in (_abi.SOP_OK, _abi.SOP_ERR_NO_DEVICE)


def test_parameter_sets():
    r = S.params_reference()
    assert (r.k_bnd, r.R0, r.sigma_ss, r.n_nl, r.n_pl, r.nl_factor, r.pl_factor) == (20, 2, 3.5, 50, 1, 3.2, 2.5)
    assert r.eps_att[1][1] == pytest.approx(1.0) and r.eps_att[2][2] == pytest.approx(0.7)
    assert r.sigma_rep[1][2] == pytest.approx(5.4) and r.boxl == 500 and r.seed == 1234
    c = S.params_config()
    assert (c.nl_abs, c.pl_abs, c.n_nl, c.n_pl, c.zeta, c.h, c.T) == (32, 16, 50, 10, 50, pytest.approx(0.05),
                                                                      pytest.approx(0.6))
    assert c.sigma_ss == pytest.approx(5.4) and c.eps_att[1][1] == 1 and c.boxl == 0
    a5, var = S.params_derived(c)
    assert a5 == np.float32(np.float32(0.05) / np.float32(50.0))
    assert var == np.float32(np.sqrt(2.0 * np.float64(np.float32(0.6)) * 50.0 / np.float64(np.float32(0.05))))


@pytest.mark.parametrize("n", [10, 128, 2000])
def test_generator_is_seeded_compact_and_self_avoiding(n):
    from scipy.spatial import cKDTree
    x1, b1 = S.generate_chain(5, n)
    x2, b2 = S.generate_chain(5, n)
    x3, _ = S.generate_chain(6, n)
    assert np.array_equal(x1, x2) and np.array_equal(b1, b2) and not np.array_equal(x1, x3)
    d = np.linalg.norm(np.diff(x1.astype(np.float64), axis=0), axis=1)
    assert np.allclose(d, 3.8, atol=1e-5) and np.allclose(b1[:-1], d, atol=1e-6)
    pairs = cKDTree(x1).query_pairs(4.0 - 1e-4, output_type="ndarray")
    far = pairs[np.abs(pairs[:, 0] - pairs[:, 1]) >= 2]
    assert len(far) == 0
    rg = np.sqrt(((x1 - x1.mean(0)) ** 2).sum(1).mean())
    R = (3 * n * 140.0 / (4 * np.pi)) ** (1 / 3) + 3.8
    assert rg < R


def test_assembly_chains_keep_their_distance():
    # the 1,000-chain lattice assembly: no two beads of different chains (nor
    # non-bonded beads of one chain) within 4 A, bonds at their r0, and
    # natives only inside chains
    from scipy.spatial import cKDTree
    from sop_amd import workloads as W
    s = W.c5()
    assert s.n == 500000 and len(np.unique(s.chain)) == 1000
    pairs = cKDTree(s.xyz).query_pairs(4.0 - 1e-4, output_type="ndarray")
    bonded = (s.chain[pairs[:, 0]] == s.chain[pairs[:, 1]]) & (np.abs(pairs[:, 0] - pairs[:, 1]) == 1)
    assert bonded.all(), pairs[~bonded][:5]
    x = s.xyz.astype(np.float64)
    same = s.chain[1:] == s.chain[:-1]
    d = np.linalg.norm(x[1:] - x[:-1], axis=1)
    assert np.abs(d[same] - s.bond_r0[:-1][same]).max() < 1e-5
    assert (s.chain[s.nat_i] == s.chain[s.nat_j]).all()


def test_native_contacts_match_brute_force():
    x, _ = S.generate_chain(3, 300)
    ch = np.repeat([0, 1, 2], 100).astype(np.int32)
    for intra in (False, True):
        ni, nj, nr = S.native_contacts(x, ch, 8.0, intra_only=intra)
        d = np.linalg.norm(x[:, None, :].astype(np.float64) - x[None, :, :], axis=2)
        i, j = np.triu_indices(300, 1)
        ok = d[i, j] < 8.0
        ok &= ~((ch[i] == ch[j]) & (j - i <= 2))
        if intra:
            ok &= ch[i] == ch[j]
        want = set(zip(i[ok].tolist(), j[ok].tolist()))
        assert set(zip(ni.tolist(), nj.tolist())) == want
        assert np.allclose(nr, d[ni, nj].astype(np.float32))


def test_reference_input_files_round_trip(tmp_path):
    sysd = S.make_chain_system(4, 60)
    write_reference_files(str(tmp_path), sysd)
    t = S.System.from_reference_files(*(str(tmp_path / f) for f in ("bonds.dat", "angles.dat", "vdw.dat", "init.dat")))
    assert t.n == 60 and np.all(t.chain == 0) and np.all(t.type == 1)
    assert np.allclose(t.xyz, sysd.xyz, atol=1e-5)
    assert np.allclose(t.bond_r0[:-1], sysd.bond_r0[:-1], atol=1e-5)
    got = sorted(zip(t.nat_i.tolist(), t.nat_j.tolist()))
    assert got == sorted(zip(sysd.nat_i.tolist(), sysd.nat_j.tolist()))
    assert t.params.nl_factor == 3.2 and t.params.n_pl == 1


def test_reference_reader_rejects_partial_master_list(tmp_path):
    sysd = S.make_chain_system(4, 30)
    write_reference_files(str(tmp_path), sysd)
    p = tmp_path / "vdw.dat"
    lines = p.read_text().splitlines()
    head = lines[0].split()
    head[3] = str(int(head[3]) - 1)
    p.write_text(" ".join(head) + "\n" + "\n".join(lines[1:-1]) + "\n")
    with pytest.raises(S.SopError):
        S.System.from_reference_files(*(str(tmp_path / f) for f in ("bonds.dat", "angles.dat", "vdw.dat", "init.dat")))


def test_trajectory_and_coordinate_writers(tmp_path):
    x = np.array([[1.5, -2.25, 3.0], [4.0, 5.0, -6.125]], np.float32)
    tr = str(tmp_path / "traj.bin")
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    assert S.lib().sop_write_traj(tr.encode(), 2, p(x)) == 0
    assert S.lib().sop_write_traj(tr.encode(), 2, p(x)) == 0  # appends (io.cpp:270)
    assert open(tr).read() == "1.500000 -2.250000 3.000000\n4.000000 5.000000 -6.125000\n" * 2
    cf = str(tmp_path / "coord.out")
    assert S.lib().sop_save_coords(cf.encode(), 2, p(x)) == 0
    assert open(cf).read() == "1 1.500000 -2.250000 3.000000\n2 4.000000 5.000000 -6.125000\n"
    y = np.zeros_like(x)
    assert S.lib().sop_load_coords(cf.encode(), 2, p(y)) == 0
    assert np.array_equal(x, y)
