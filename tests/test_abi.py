"""The C ABI library: loads, exports every entry point include/goqu_hip.h declares, and turns
host-forceable failures into status codes (never an abort).  No compute call needs a GPU."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from goqu_amd import _native as N

ROOT = Path(__file__).resolve().parent.parent


def declared():
    text = (ROOT / "include" / "goqu_hip.h").read_text()
    return sorted(set(re.findall(r"\b(goqu_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    lib = C.CDLL(str(N.LIB_PATH))
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every declared symbol has a ctypes signature in the binding
    assert set(names) <= set(N._SIGS), set(names) - set(N._SIGS)


def test_op_struct_layout_matches_header():
    # int32 x4, int32[16], pointer, int32 x2 -> 4*4 + 64 + 8 + 8 = 96 bytes on LP64
    assert C.sizeof(N.GoquOp) == 96
    assert N.GoquOp.matrix.offset == 80
    assert C.sizeof(N.GoquDeriv) == 24


def test_invalid_arguments_are_status_codes():
    L = N.lib()
    h = C.c_void_p()
    assert L.goqu_sv_create(0, N.C128, C.byref(h)) == N.ERR_INVALID_ARG
    assert L.goqu_sv_create(35, N.C128, C.byref(h)) == N.ERR_INVALID_ARG
    assert L.goqu_sv_create(4, 7, C.byref(h)) == N.ERR_INVALID_ARG
    assert L.goqu_dm_create(17, C.byref(h)) == N.ERR_INVALID_ARG
    assert b"out of range" in L.goqu_last_error()
    assert L.goqu_sv_create(4, N.C128, None) == N.ERR_INVALID_ARG
    assert b"null" in L.goqu_last_error()
    assert L.goqu_sv_apply(None, None, 0) == N.ERR_INVALID_ARG
    assert L.goqu_sv_destroy(None) == N.OK


def test_no_device_is_an_error_not_an_abort():
    if N.device_count() > 0:
        pytest.skip("a HIP device is visible")
    h = C.c_void_p()
    assert N.lib().goqu_sv_create(4, N.C128, C.byref(h)) == N.ERR_NO_DEVICE
    assert N.lib().goqu_dm_create(2, C.byref(h)) == N.ERR_NO_DEVICE
    from goqu_amd import statevector

    with pytest.raises(N.GoquError):
        statevector.New(3)


def _small_buffer(n=5):
    from goqu_amd.circuit import NewBuilder
    from oracle import oracle as O

    b = NewBuilder("p", n)
    for q in range(n):
        b.H(q).RZ(0.3 * q, q)
    b.CNOT(0, n - 1).CZ(1, 2)
    return O.sv_marshal(b.Build())


def test_host_planning_needs_no_device():
    stages, instrs = N.plan_stats(5, _small_buffer())
    assert stages >= 1 and instrs >= 1
    with pytest.raises(N.GoquError) as e:
        N.plan_stats(40, _small_buffer())
    assert e.value.code == N.ERR_INVALID_ARG


def test_host_allocation_failure_is_a_status_code():
    """A std::bad_alloc inside the library (forced by the fault-injection hook before the
    lowering) comes back as GOQU_ERR_OUT_OF_MEMORY, not as an abort of the host process; the hook
    fires once, so the next call succeeds."""
    L = N.lib()
    assert L.goqu_debug_inject_fault(N.FAULT_HOST_ALLOC, 1) == N.OK
    with pytest.raises(N.GoquError) as e:
        N.plan_stats(5, _small_buffer())
    assert e.value.code == N.ERR_OOM
    assert b"host memory" in L.goqu_last_error()
    N.plan_stats(5, _small_buffer())
    assert L.goqu_debug_inject_fault(7, 1) == N.ERR_INVALID_ARG
    assert L.goqu_debug_inject_fault(0, 0) == N.OK


def test_oversize_op_records_are_rejected():
    """Records with too many qubits or qubits out of range are argument errors."""
    from goqu_amd.circuit import NewBuilder
    from oracle import oracle as O

    buf = _small_buffer()
    arr = buf.array()
    arr[0].num_qubits = 17
    st, ni = C.c_int64(0), C.c_int64(0)
    assert N.lib().goqu_plan_stats(5, 0, arr, len(buf), 0, C.byref(st), C.byref(ni)) == N.ERR_INVALID_ARG
    arr[0].num_qubits = 1
    arr[0].qubits[0] = 9
    assert N.lib().goqu_plan_stats(5, 0, arr, len(buf), 0, C.byref(st), C.byref(ni)) == N.ERR_INVALID_ARG
    assert N.lib().goqu_plan_stats(5, 0, arr, len(buf), 21, C.byref(st), C.byref(ni)) == N.ERR_INVALID_ARG
