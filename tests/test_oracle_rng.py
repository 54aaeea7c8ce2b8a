"""Pins of the oracle's random-force stream (CPU only).

Philox4x32-10 is pinned by the Random123 known-answer vectors and bitwise by
rocRAND's open-source philox4x32_10 engine; the Box-Muller normals by their
moments, a Kolmogorov-Smirnov test and component independence.  The
reference's cuRAND XORWOW stream (energy.cu:597-599) is closed code and is not
reproduced: the noise stream itself is this engine's own.
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import oracle as O

# Random123 kat_vectors, philox4x32_10: counter words, key words -> output.
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(ctr, key, want):
    assert tuple(int(v) for v in O.philox(ctr, key)) == want


def test_philox_matches_rocrand():
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    src = os.path.join(os.path.dirname(__file__), "helpers", "rocrand_xcheck.cpp")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "xcheck")
        r = subprocess.run([hipcc, "-O1", "-w", src, "-o", exe], capture_output=True, text=True)
        if r.returncode != 0:
            pytest.skip("could not build the rocRAND cross-check: " + r.stderr[-300:])
        rng = np.random.default_rng(7)
        cases = []
        for _ in range(200):
            seed = int(rng.integers(0, 2**63)) * 2 + int(rng.integers(0, 2))
            c_lo = int(rng.integers(0, 2**62))
            c_hi = int(rng.integers(0, 2**63))
            cases.append((seed, c_lo, c_hi))
        cases += [(0, 0, 0), (2**64 - 1, 2**62 - 1, 2**64 - 1), (0x9ABCDEF012345678, 1000, 0)]
        inp = "".join(f"{s:x} {hi:x} {4 * lo:x}\n" for s, lo, hi in cases)
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
        for (s, lo, hi), line in zip(cases, out):
            ctr = (lo & 0xFFFFFFFF, lo >> 32, hi & 0xFFFFFFFF, hi >> 32)
            key = (s & 0xFFFFFFFF, s >> 32)
            got = O.philox(ctr, key)
            assert " ".join(f"{int(v):08x}" for v in got) == line.strip(), (s, lo, hi)


def _sample(nb=400, ns=500, seed=0x1234):
    v = np.array([O.normals(seed, b, s) for b in range(nb) for s in range(ns)])
    return v


def test_normals_moments_and_independence():
    v = _sample()
    assert v.shape == (200000, 3)
    assert np.all(np.abs(v.mean(0)) < 0.01)
    assert np.all(np.abs(v.var(0) - 1.0) < 0.015)
    kurt = ((v - v.mean(0)) ** 4).mean(0) / v.var(0) ** 2
    assert np.all(np.abs(kurt - 3.0) < 0.06)
    c = np.corrcoef(v.T)
    assert np.all(np.abs(c[np.triu_indices(3, 1)]) < 0.01)


def test_normals_ks():
    from scipy import stats
    v = _sample(200, 250, seed=99)
    for k in range(3):
        assert stats.kstest(v[:, k], "norm").pvalue > 1e-3


def test_normals_depend_on_every_counter_field():
    a = O.normals(5, 10, 20)
    assert np.array_equal(a, O.normals(5, 10, 20))
    for b in (O.normals(6, 10, 20), O.normals(5, 11, 20), O.normals(5, 10, 21), O.normals(5, 10, 20 + 2**32)):
        assert not np.allclose(a, b)
