"""ems_energy_noalloc, the allocation-free variant of ems_energy: through the
C ABI, its value equals ems_energy's bit for bit on two calls in a row, the
second on a smaller mesh, and on a third call back on the first engine (the
shared device scratch keeps the previous call's partial sums, which every
call overwrites).  The result lands in caller-owned device memory, which the
test allocates through the same HIP runtime the library uses."""
import ctypes as C

import pytest

from openems_amd import scenes
from openems_amd._native import check, lib
from openems_amd.engine import EngineHIP
from openems_amd.operator import Operator

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the engine library is linked against (already loaded)."""
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return hip


def _read(hip, dptr):
    v = C.c_double()
    assert hip.hipMemcpy(C.byref(v), dptr, 8, 2) == 0      # hipMemcpyDeviceToHost
    return v.value


def _energy_pair(hip, eng, dptr):
    """(ems_energy, ems_energy_noalloc) of one engine; dptr: a device double."""
    L = lib()
    e = C.c_double()
    check(L.ems_energy(eng.h, C.byref(e)))
    assert hip.hipMemset(dptr, 0xFF, 8) == 0
    check(L.ems_energy_noalloc(eng.h, dptr))
    eng.sync()
    return e.value, _read(hip, dptr)


def test_energy_noalloc_matches_energy(gpu):
    hip = _hip()
    big = EngineHIP(Operator(scenes.upml_bricks(48, 60, nbricks=6, dump=False)))
    small = EngineHIP(Operator(scenes.pec_cavity(24, 40)))
    dptr = C.c_void_p()
    assert hip.hipMalloc(C.byref(dptr), 8) == 0
    try:
        big.IterateTS(60)
        small.IterateTS(40)
        e1, n1 = _energy_pair(hip, big, dptr)
        e2, n2 = _energy_pair(hip, small, dptr)
        e3, n3 = _energy_pair(hip, big, dptr)
        print("energies", e1, n1, e2, n2, e3, n3, flush=True)
        assert e1 > 0 and e2 > 0 and e1 != e2
        assert n1 == e1 and n2 == e2 and n3 == e3
        # enqueued repeatedly without a synchronisation in between, the last
        # value is the energy of the mesh as it stands
        for _ in range(20):
            check(lib().ems_energy_noalloc(small.h, dptr))
        small.sync()
        assert _read(hip, dptr) == e2
    finally:
        hip.hipFree(dptr)
        big.close()
        small.close()


def test_energy_noalloc_rejects_null(gpu):
    from openems_amd._native import NativeError
    eng = EngineHIP(Operator(scenes.pec_cavity(16, 4)))
    try:
        with pytest.raises(NativeError, match="null"):
            check(lib().ems_energy_noalloc(eng.h, None))
    finally:
        eng.close()
