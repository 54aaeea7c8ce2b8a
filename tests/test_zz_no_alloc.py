"""sop_energy_eval_noalloc through the C ABI against sop_energy_eval (collected
last: the file name sorts after every other test file).

The allocation-free entry point keeps its work table and energy partials in
the library's file-scope device array.  Two calls in a row, the second on a
smaller system (the scratch keeps the first call's contents), each give
energies identical to the allocating entry point's; a system above the
scratch limit is refused with SOP_ERR_CAPACITY and a message.
"""
import ctypes as C

import numpy as np
import pytest

import sop_amd as S
from sop_amd import _abi

pytestmark = pytest.mark.gpu


def noalloc(eng, sid):
    e = np.zeros(S.E_COUNT, np.float64)
    st = S.lib().sop_energy_eval_noalloc(eng._h, sid, e.ctypes.data_as(C.POINTER(C.c_double)))
    return st, e


def test_noalloc_energies_equal_the_allocating_entry_point():
    big, small = S.make_chain_system(901, 3000), S.make_chain_system(902, 300)
    with S.Engine(0) as eng:
        ib, i_s = eng.add(big), eng.add(small)
        eng.run([ib, i_s], 20)
        for sid in (ib, i_s):  # larger first, then smaller: stale scratch must not leak
            st, e = noalloc(eng, sid)
            assert st == _abi.SOP_OK, S.lib().sop_last_error(eng._h)
            ref = eng.energy_eval(sid)
            assert np.array_equal(e, ref), (sid, e, ref)
        # again after both, in the other order
        for sid in (i_s, ib):
            st, e = noalloc(eng, sid)
            assert st == _abi.SOP_OK
            assert np.array_equal(e, eng.energy_eval(sid))


def test_noalloc_refuses_systems_above_its_limit():
    huge = S.make_chain_system(903, 65600)
    with S.Engine(0) as eng:
        sid = eng.add(huge)
        st, _ = noalloc(eng, sid)
        assert st == _abi.SOP_ERR_CAPACITY
        msg = S.lib().sop_last_error(eng._h).decode()
        assert "65536" in msg, msg
