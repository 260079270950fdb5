"""The drop-in contract on the device: kernels written only against include/ouro/instances.hpp with the
reference's names call mm->malloc / mm->free from data-dependent branches (odd lanes allocating while
even lanes free in the else branch, every third lane, frees from both arms of a nested branch), on
pools of all six kinds created through the C ABI and on one created by the reference's own host
sequence (initialize, getDeviceMemoryManager)."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import ouroboros_amd as ob

pytestmark = pytest.mark.gpu

SO = Path(__file__).resolve().parent / "dropin" / "libdropin.so"


@pytest.fixture(scope="module")
def dropin():
    lib = C.CDLL(str(SO))
    lib.dropin_run.argtypes = [C.c_int, C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_ulonglong)]
    lib.dropin_own_manager.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(C.c_ulonglong),
                                       C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    return lib


@pytest.mark.parametrize("kind", ["PQ", "CQ", "VAPQ", "VACQ", "VLPQ", "VLCQ"])
def test_divergent_calls_on_every_kind(dropin, kind):
    print(f"drop-in divergent {kind}", flush=True)
    with ob.MemoryManager(kind, 1 << 30) as mm:
        errs = (C.c_ulonglong * 2)()
        assert dropin.dropin_run(ob.VARIANTS[kind], mm.getDeviceMemoryManager(), 65536, 8, errs) == 0
        assert list(errs) == [0, 0]
        st = mm.query()
        assert st.error == 0, ob.decode_errors(st.error)
        chk = mm.check_quiescent()
        assert chk.violations == 0, chk.message.decode()


def test_reference_host_sequence(dropin):
    print("drop-in own manager", flush=True)
    errs = (C.c_ulonglong * 2)()
    err, nfc = C.c_uint32(), C.c_uint32()
    assert dropin.dropin_own_manager(1 << 30, 65536, 6, errs, C.byref(err), C.byref(nfc)) == 0
    assert list(errs) == [0, 0] and err.value == 0 and nfc.value > 12
