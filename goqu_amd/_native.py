"""ctypes binding of libgoqu_hip (include/goqu_hip.h) and marshalling of circuits into the
C ABI's op records.  There is no fallback: if the library is missing, importing the
simulators raises."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import gate as G

OK, ERR_INVALID_ARG, ERR_OOM, ERR_UNSUPPORTED, ERR_DEVICE, ERR_NO_DEVICE, ERR_NO_MATRIX = range(7)
C128, C64 = 0, 1
OP_GATE, OP_CONTROLLED, OP_RESET, OP_SKIP, OP_CHANNEL = range(5)
MAX_OP_QUBITS = 16

LIB_PATH = Path(__file__).resolve().parent / "libgoqu_hip.so"


class GoquOp(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("gate_id", C.c_int32), ("num_qubits", C.c_int32), ("num_controls", C.c_int32),
        ("qubits", C.c_int32 * MAX_OP_QUBITS), ("matrix", C.POINTER(C.c_double)),
        ("num_kraus", C.c_int32), ("reserved", C.c_int32),
    ]


# GoquOp as a numpy record (the layout of goqu_op_t)
OP_DTYPE = np.dtype([("kind", "<i4"), ("gate_id", "<i4"), ("num_qubits", "<i4"), ("num_controls", "<i4"),
                     ("qubits", "<i4", (MAX_OP_QUBITS,)), ("matrix", "<u8"), ("num_kraus", "<i4"),
                     ("reserved", "<i4")])
assert OP_DTYPE.itemsize == C.sizeof(GoquOp)


def op_array(records: np.ndarray):
    """A ctypes GoquOp array viewing a contiguous OP_DTYPE numpy array (no copy)."""
    records = np.ascontiguousarray(records)
    return (GoquOp * len(records)).from_buffer(records)


class GoquDeriv(C.Structure):
    _fields_ = [("op", C.c_int32), ("param", C.c_int32), ("scale", C.c_double), ("dmatrix", C.POINTER(C.c_double))]


class GoquDynInstr(C.Structure):
    _fields_ = [("kind", C.c_int32), ("a", C.c_int32), ("b", C.c_int32), ("value", C.c_int32)]


DYN_GATES, DYN_MEASURE, DYN_RESET, DYN_BRANCH, DYN_CASE, DYN_DEFAULT, DYN_END, DYN_WHILE, DYN_FOR = range(9)


class GoquError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"goqu_hip error {code}: {msg}")
        self.code = code


_lib = None

_P = C.c_void_p
_D = C.POINTER(C.c_double)
_U64 = C.POINTER(C.c_uint64)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)

_SIGS = {
    "goqu_last_error": ([], C.c_char_p),
    "goqu_device_count": ([_I32], C.c_int),
    "goqu_device_name": ([C.c_char_p, C.c_int32], C.c_int),
    "goqu_sv_create": ([C.c_int32, C.c_int32, C.POINTER(_P)], C.c_int),
    "goqu_sv_destroy": ([_P], C.c_int),
    "goqu_sv_reset": ([_P], C.c_int),
    "goqu_sv_apply": ([_P, C.POINTER(GoquOp), C.c_int32], C.c_int),
    "goqu_sv_set_plan": ([_P, C.c_int32, C.c_int32], C.c_int),
    "goqu_sv_get_state": ([_P, _D, C.c_uint64, C.c_uint64], C.c_int),
    "goqu_sv_set_state": ([_P, _D, C.c_uint64, C.c_uint64], C.c_int),
    "goqu_sv_sample": ([_P, _D, C.c_int64, _U64], C.c_int),
    "goqu_sv_sample_noalloc": ([_P, _D, C.c_int64, _U64], C.c_int),
    "goqu_sv_expect_paulis": ([_P, C.c_int32, _U64, _U64, _I32, _D, _D], C.c_int),
    "goqu_sv_expectation_z": ([_P, C.c_uint64, _D], C.c_int),
    "goqu_sv_measure_qubit": ([_P, C.c_int32, C.c_double, _I32], C.c_int),
    "goqu_sv_norm2": ([_P, _D], C.c_int),
    "goqu_dm_create": ([C.c_int32, C.POINTER(_P)], C.c_int),
    "goqu_dm_destroy": ([_P], C.c_int),
    "goqu_dm_reset": ([_P], C.c_int),
    "goqu_dm_apply": ([_P, C.POINTER(GoquOp), C.c_int32], C.c_int),
    "goqu_dm_get": ([_P, _D, C.c_uint64, C.c_uint64], C.c_int),
    "goqu_dm_set": ([_P, _D, C.c_uint64, C.c_uint64], C.c_int),
    "goqu_dm_probs": ([_P, _D, _D, _D], C.c_int),
    "goqu_dm_expect_paulis": ([_P, C.c_int32, _U64, _U64, _I32, _D, _D], C.c_int),
    "goqu_dm_expectation_z": ([_P, C.c_uint64, _D], C.c_int),
    "goqu_dm_purity": ([_P, _D], C.c_int),
    "goqu_dm_fidelity": ([_P, _D, _D], C.c_int),
    "goqu_dm_sample": ([_P, _D, _D, C.c_int64, _U64], C.c_int),
    "goqu_sv_adjoint": ([_P, C.POINTER(GoquOp), C.c_int32, C.POINTER(GoquDeriv), C.c_int32, C.c_int32, C.c_int32,
                         _U64, _U64, _I32, _D, _D, _D], C.c_int),
    "goqu_dm_adjoint": ([_P, C.POINTER(GoquOp), C.c_int32, C.POINTER(GoquDeriv), C.c_int32, C.c_int32, C.c_int32,
                         _U64, _U64, _I32, _D, C.c_int32, _D, _D], C.c_int),
    "goqu_qaoa_batch": ([C.c_int32, _I32, _I32, _I64, _I32, _I64, _D, _D, _D], C.c_int),
    "goqu_sv_run_dynamic": ([_P, C.POINTER(GoquOp), C.c_int32, C.POINTER(GoquDynInstr), C.c_int32, _I32, C.c_int32,
                             C.c_int32, C.c_int64, _D, C.c_int32, C.POINTER(C.c_uint8)], C.c_int),
    "goqu_dm_run_dynamic": ([_P, C.POINTER(GoquOp), C.c_int32, C.POINTER(GoquDynInstr), C.c_int32, _I32, C.c_int32,
                             C.c_int32, C.c_int64, _D, C.c_int32, C.POINTER(C.c_uint8)], C.c_int),
    "goqu_sv_batch_expect": ([C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(GoquOp), _I64,
                              C.c_int32, _U64, _U64, _I32, _D, _D], C.c_int),
    "goqu_sv_batch_sample": ([C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(GoquOp), _I64,
                              C.c_int64, _D, _U64], C.c_int),
    "goqu_dm_batch_expect": ([C.c_int32, C.c_int32, C.POINTER(GoquOp), _I64, C.c_int32, _U64, _U64, _I32, _D, _D],
                             C.c_int),
    "goqu_dm_batch_sample": ([C.c_int32, C.c_int32, C.POINTER(GoquOp), _I64, _D, _D, C.c_int64, _D, _U64], C.c_int),
    "goqu_plan_stats": ([C.c_int32, C.c_int32, C.POINTER(GoquOp), C.c_int32, C.c_int32, _I64, _I64], C.c_int),
    "goqu_debug_inject_fault": ([C.c_int32, C.c_int32], C.c_int),
    "goqu_profile_enable": ([C.c_int32], C.c_int),
    "goqu_profile_reset": ([], C.c_int),
    "goqu_profile_read": ([_D, _I64, _D], C.c_int),
    "goqu_profile_read_kernel": ([C.c_int32, _D, _I64, _D], C.c_int),
}


def lib():
    """Load libgoqu_hip.so (built in-tree by goqu_amd.build).  Raises if it is missing."""
    global _lib
    if _lib is None:
        path = Path(os.environ.get("GOQU_HIP_LIB", LIB_PATH))
        if not path.exists():
            raise ImportError(f"{path} is missing: build it with `python -m goqu_amd.build`")
        L = C.CDLL(str(path))
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != OK:
        msg = lib().goqu_last_error().decode(errors="replace")
        raise GoquError(rc, msg)


def device_count() -> int:
    n = C.c_int32(0)
    check(lib().goqu_device_count(C.byref(n)))
    return n.value


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_D)


def u64ptr(a: np.ndarray):
    return a.ctypes.data_as(_U64)


def i32ptr(a: np.ndarray):
    return a.ctypes.data_as(_I32)


def i64ptr(a: np.ndarray):
    return a.ctypes.data_as(_I64)


class OpBuffer:
    """An array of GoquOp records plus the matrices they borrow (kept alive here)."""

    def __init__(self):
        self.records: List[Tuple] = []
        self.keep: List[np.ndarray] = []

    def add(self, kind: int, qubits: Sequence[int], matrix: Optional[np.ndarray] = None, gate_id: int = 0,
            num_controls: int = 0, num_kraus: int = 0) -> int:
        m = None
        if matrix is not None:
            m = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128).reshape(-1)).view(np.float64)
            self.keep.append(m)
        qs = list(qubits)
        if len(qs) > MAX_OP_QUBITS:
            raise ValueError(f"op acts on {len(qs)} qubits (max {MAX_OP_QUBITS})")
        self.records.append((kind, gate_id, qs, num_controls, m, num_kraus, m.ctypes.data if m is not None else 0))
        return len(self.records) - 1

    def __len__(self):
        return len(self.records)

    def structured(self) -> np.ndarray:
        """The records as a numpy array of OP_DTYPE (matrix pointers into self.keep)."""
        n = len(self.records)
        out = np.zeros(max(1, n), dtype=OP_DTYPE)
        if not n:
            return out
        R = self.records
        out["kind"][:n] = [r[0] for r in R]
        out["gate_id"][:n] = [r[1] for r in R]
        out["num_qubits"][:n] = [len(r[2]) for r in R]
        out["num_controls"][:n] = [r[3] for r in R]
        out["num_kraus"][:n] = [r[5] for r in R]
        out["matrix"][:n] = [r[6] for r in R]
        q = out["qubits"]
        for k in range(MAX_OP_QUBITS):
            rows = [i for i, r in enumerate(R) if len(r[2]) > k]
            if not rows:
                break
            q[rows, k] = [R[i][2][k] for i in rows]
        return out

    def array(self):
        """A ctypes GoquOp array over the records (it keeps its numpy storage alive)."""
        st = self.structured()
        return (GoquOp * len(st)).from_buffer(st)


def add_sv_gate(buf: OpBuffer, g: G.Gate, qubits: Sequence[int]) -> int:
    """One ir.Operation -> op record, with the reference's dispatch tiers
    (sim/statevector/sim.go:98-155, kernel2q.go:24-114, kernel3q.go:10-45)."""
    name = g.Name()
    if name in ("barrier", "delay"):
        return buf.add(OP_SKIP, qubits[:1] or [0])
    if name == "reset":
        return buf.add(OP_RESET, qubits)
    k = g.Qubits()
    if k >= 4 and isinstance(g, G.ControlledGate):
        inner = g.Inner().Matrix()
        return buf.add(OP_CONTROLLED, qubits, inner, num_controls=g.NumControls())
    m = g.Matrix()
    return buf.add(OP_GATE, qubits, m, gate_id=g.tier)


FAULT_DEVICE_ALLOC, FAULT_HOST_ALLOC = 1, 2


def plan_stats(nq: int, buf: OpBuffer, density: bool = False, tile_bits: int = 0) -> Tuple[int, int]:
    """(stage sweeps, instructions) of the planned program, computed on the host only."""
    st, ni = C.c_int64(0), C.c_int64(0)
    check(lib().goqu_plan_stats(nq, 1 if density else 0, buf.array(), len(buf), tile_bits, C.byref(st),
                                C.byref(ni)))
    return st.value, ni.value


def profile(enable: bool) -> None:
    check(lib().goqu_profile_enable(1 if enable else 0))
    check(lib().goqu_profile_reset())


PROF_STAGE, PROF_STAGE_DM, PROF_STAGE_PAIR, PROF_HAMILTONIAN, PROF_WIDE, PROF_QAOA = range(6)


def profile_read_kernel(kernel: int):
    """(summed ms, launches, algorithmic bytes or flops) of one profiled kernel."""
    ms, n, b = C.c_double(0), C.c_int64(0), C.c_double(0)
    check(lib().goqu_profile_read_kernel(kernel, C.byref(ms), C.byref(n), C.byref(b)))
    return ms.value, n.value, b.value


def profile_read():
    """(stage-kernel milliseconds, launches, algorithmic bytes) since the last reset."""
    ms, n, b = C.c_double(0), C.c_int64(0), C.c_double(0)
    check(lib().goqu_profile_read(C.byref(ms), C.byref(n), C.byref(b)))
    return ms.value, n.value, b.value
