"""The reference's QASM corpus (qasm/testdata/*.qasm, copied as fixtures under tests/golden/qasm)
through this package's QASM reader (a restatement of qasm/parser/parser.go): every file parses
(TestParseQASMFiles, parser_test.go:294-317), the lowering keeps the reference's conventions, and
the oracle's evolution of each circuit's gates equals a brute-force product of full 2^n x 2^n
gate matrices built with numpy."""
from pathlib import Path

import numpy as np
import pytest

from goqu_amd import _native as N
from goqu_amd import qasm
from oracle import oracle as O

CORPUS = sorted((Path(__file__).parent / "golden" / "qasm").glob("*.qasm"))


def full_matrix(m, qubits, n):
    """A gate matrix on qubits (qubits[0] the most significant bit of its index) as a 2^n x 2^n
    matrix over basis indices whose bit k is qubit k."""
    k = len(qubits)
    m = np.asarray(m, dtype=np.complex128).reshape(1 << k, 1 << k)
    dim = 1 << n
    U = np.zeros((dim, dim), dtype=np.complex128)
    for col in range(dim):
        sub = sum(((col >> q) & 1) << (k - 1 - j) for j, q in enumerate(qubits))
        rest = col & ~sum(1 << q for q in qubits)
        for r in range(1 << k):
            row = rest | sum(((r >> (k - 1 - j)) & 1) << q for j, q in enumerate(qubits))
            U[row, col] += m[r, sub]
    return U


def brute_force(c):
    n = c.NumQubits()
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1
    for op in c.Ops():
        g = op.Gate
        if g is None or g.Name() in ("barrier", "delay", "reset"):
            continue
        psi = full_matrix(g.Matrix(), op.Qubits, n) @ psi
    return psi


def test_every_corpus_file_parses():
    assert len(CORPUS) == 15
    for f in CORPUS:
        c = qasm.ParseFile(str(f))
        assert c.NumQubits() > 0, f.name


def _parse(name):
    return qasm.ParseFile(str(Path(__file__).parent / "golden" / "qasm" / name))


def test_lowering_conventions():
    """Register-wide measure and reset expand per qubit; `measure q -> c` and `c = measure q` pair
    bits by index; a user gate is inlined; registers are laid out in declaration order; `if (c ==
    1)` on a whole register tests its first bit; an `if` without else conditions its body's ops."""
    names = lambda c: [op.Gate.Name() if op.Gate is not None else "measure" for op in c.Ops()]
    assert names(_parse("custom_gate.qasm")) == ["H", "CNOT", "measure", "measure"]
    assert [(op.Qubits, op.Clbits) for op in _parse("legacy_syntax.qasm").Ops()[2:]] == [([0], [0]), ([1], [1])]
    mr = _parse("multi_register.qasm")
    assert [op.Qubits for op in mr.Ops()[:3]] == [[0], [0, 1], [0, 2]]
    cond = _parse("conditional.qasm").Ops()[2]
    assert cond.Gate.Name() == "X" and cond.Condition.Clbit == 0 and cond.Condition.Value == 1
    tp = _parse("teleport.qasm")
    z = [op for op in tp.Ops() if op.Gate is not None and op.Gate.Name() == "Z"][0]
    assert z.Qubits == [2] and z.Condition.Clbit == 0 and z.Condition.Value == 1
    g3 = _parse("ghz3.qasm")
    assert names(g3)[:3] == ["reset"] * 3 and g3.IsDynamic()
    rc = _parse("rotation_chain.qasm")
    assert [op.Gate.Name() for op in rc.Ops()[:6]] == ["RX(0.3927)", "RY(0.7854)", "RZ(1.5708)", "RX(3.1416)",
                                                      "RY(2.3562)", "RZ(-1.0472)"]


def test_expression_and_error_cases():
    c = qasm.ParseString('OPENQASM 3.0; qubit[1] q; rx(2 * pi / 8 + tau - tau) q[0]; ry(sqrt(4) ** 2 % 3) q;'
                         ' rz(-(pi)) q[0]; p(euler) q[0];')
    m = [np.asarray(op.Gate.Matrix()) for op in c.Ops()]
    assert np.allclose(m[0], np.asarray(qasm.G.RX(np.pi / 4).Matrix()))
    assert np.allclose(m[1], np.asarray(qasm.G.RY(1.0).Matrix()))
    assert np.allclose(m[2], np.asarray(qasm.G.RZ(-np.pi).Matrix()))
    assert np.allclose(m[3], np.asarray(qasm.G.Phase(np.e).Matrix()))
    for bad in ["qubit[1] q; foo q[0];", "qubit[1] q; h q[1];", "qubit[1] q; rx q[0];",
                "qubit[2] q; bit[1] c; measure q;", "qubit[1] q; rx(1/0) q[0];", "qubit[0] q;"]:
        with pytest.raises(qasm.QASMError):
            qasm.ParseString(bad)


def test_modifiers_and_control_flow_forms():
    c = qasm.ParseString("""OPENQASM 3.0; qubit[3] q; bit[2] c;
        ctrl @ x q[0], q[1]; negctrl @ x q[0], q[2]; inv @ s q[1]; pow(2) @ t q[2]; ctrl(2) @ h q[0], q[1], q[2];
        c[0] = measure q[0];
        if (c[0] == 1) { x q[1]; } else { h q[1]; }
        while (c[0] == 1) { h q[0]; c[0] = measure q[0]; }
        for int[32] i in [0:4:2] { x q[2]; }
        switch (c) { case 0: { x q[0]; } case 1 { y q[0]; } default: { z q[0]; } }""")
    ops = c.Ops()
    assert ops[0].Gate.Name() == "CNOT" and ops[0].Qubits == [0, 1]
    assert [op.Gate.Name() for op in ops[1:4]] == ["X", "CNOT", "X"] and ops[1].Qubits == [0]
    assert ops[6].Gate.Qubits() == 3 and ops[6].Gate.NumControls() == 2
    assert np.allclose(np.asarray(ops[4].Gate.Matrix()).reshape(2, 2), np.diag([1, -1j]))  # inv @ s = sdg
    assert np.allclose(np.asarray(ops[5].Gate.Matrix()).reshape(2, 2), np.diag([1, 1j]))   # t^2 = s
    cf = [op.ControlFlow for op in ops if op.ControlFlow is not None]
    assert [x.Type for x in cf] == ["if_else", "while", "for", "switch"]
    assert cf[2].ForRange == (0, 4, 2) and cf[3].SwitchClbits == [0, 1] and cf[3].SwitchCases == [0, 1]
    assert len(cf[3].Bodies) == 3


@pytest.mark.parametrize("f", [f for f in CORPUS], ids=lambda f: f.name)
def test_oracle_matches_brute_force(f):
    """The oracle's kernels against full-matrix products on every circuit's gates (resets at the
    start of a register are no-ops on |0...0>, so they are skipped on both sides)."""
    c = qasm.ParseFile(str(f))
    want = brute_force(c)
    if c.IsDynamic() and any(op.Condition is not None or op.ControlFlow is not None for op in c.Ops()):
        pytest.skip("feed-forward: covered by the dynamic parity tests")
    buf = N.OpBuffer()
    for op in c.Ops():
        if op.Gate is not None and op.Gate.Name() not in ("reset",):
            N.add_sv_gate(buf, op.Gate, op.Qubits)
    psi = O.zero_state(c.NumQubits())
    assert O.sv_apply_ops(psi, c.NumQubits(), buf) == 0
    assert np.max(np.abs(psi - want)) < 1e-12


def test_known_answers():
    p = lambda name: np.abs(brute_force(_parse(name))) ** 2
    assert np.allclose(p("bell.qasm"), [0.5, 0, 0, 0.5])
    assert np.allclose(p("custom_gate.qasm"), [0.5, 0, 0, 0.5])
    assert np.allclose(p("ghz3.qasm")[[0, 7]], [0.5, 0.5])
    assert np.allclose(p("grover2.qasm"), [0, 0, 0, 1])
    assert np.allclose(p("toffoli.qasm")[7], 1)
    assert np.allclose(p("swap_test.qasm")[2], 1)  # q1 = 1: bitstring "10"
    assert np.allclose(p("multi_register.qasm")[[0, 7]], [0.5, 0.5])
