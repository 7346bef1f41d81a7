"""The adjoint program marshalling: the vectorised refill for rotation-only circuits gives the
records, matrices and derivative records of the generic per-gate path (gate.RX/RY/RZ and
gate_derivative), at several parameter points in a row."""
import ctypes as C
import math

import numpy as np

from goqu_amd import _native as N
from goqu_amd import gradient as GR
from goqu_amd import qaoa as Q
from goqu_amd.circuit import Add, FreeParameters, Literal, Mul, NewBuilder, NewVector


def _records(buf):
    arr = buf.array()
    out = []
    for i in range(len(buf)):
        r = arr[i]
        nm = {0: 0, 1: 4, 2: 16}.get(r.num_qubits, 0) if r.matrix else 0
        m = np.ctypeslib.as_array(C.cast(r.matrix, C.POINTER(C.c_double)), shape=(2 * nm,)).copy() if nm else None
        out.append((r.kind, r.gate_id, r.num_qubits, tuple(r.qubits[:r.num_qubits]), r.num_controls, m))
    return out


def _derivs(darr, nd):
    out = []
    for i in range(nd):
        d = darr[i]
        m = np.ctypeslib.as_array(C.cast(d.dmatrix, C.POINTER(C.c_double)), shape=(8,)).copy()
        out.append((d.op, d.param, d.scale, m))
    return out


def _compare(c, names, xs):
    fast = GR.AdjointProgram(c, names)
    for x in xs:
        bf, df, nf, _ = fast.build(x)
        assert fast._fast, "the rotation-only circuit should take the vectorised path"
        bg, dg, ng, _ = GR.AdjointProgram(c, names)._build_generic(x)
        rf, rg = _records(bf), _records(bg)
        assert len(rf) == len(rg)
        for a, b in zip(rf, rg):
            assert a[:5] == b[:5]
            assert (a[5] is None) == (b[5] is None)
            if a[5] is not None:
                np.testing.assert_allclose(a[5], b[5], rtol=0, atol=4e-16)
        assert nf == ng
        for a, b in zip(_derivs(df, nf), _derivs(dg, ng)):
            assert a[:3] == b[:3]
            np.testing.assert_allclose(a[3], b[3], rtol=0, atol=4e-16)


def test_efficient_su2_refill_matches_generic_build():
    c, npar = Q.EfficientSU2(6, 2)
    names = FreeParameters(c)
    rng = np.random.default_rng(4)
    _compare(c, names, [rng.uniform(0, 2 * math.pi, npar) for _ in range(3)])


def test_rx_ry_rz_with_shared_and_scaled_parameters():
    th = NewVector("t", 3)
    t0, t1, t2 = th.At(0), th.At(1), th.At(2)
    b = NewBuilder("rot", 3)
    b.H(0)
    b.SymRX(t0, 0).SymRY(Mul(Literal(2.0), t1), 1).CNOT(0, 1).SymRZ(Add(t2, t0), 2).SymRX(t0, 2).CZ(1, 2)
    c = b.Build()
    names = FreeParameters(c)
    rng = np.random.default_rng(8)
    _compare(c, names, [rng.uniform(-4, 4, len(names)) for _ in range(3)])


def test_other_parameterised_gates_keep_the_generic_path():
    th = NewVector("t", 2)
    b = NewBuilder("mixed", 2)
    b.SymRX(th.At(0), 0).SymRZZ(th.At(1), 0, 1)
    c = b.Build()
    prog = GR.AdjointProgram(c, FreeParameters(c))
    prog.build(np.array([0.1, 0.2]))
    assert prog._fast is False
