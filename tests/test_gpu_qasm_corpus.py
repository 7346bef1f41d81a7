"""The reference's QASM corpus (qasm/testdata/*.qasm) on the device: for every file, the
statevector and density-matrix simulators against the oracle (amplitudes / rho within 1e-12,
Run's sampled indices identical for the same uniforms), and the files with feed-forward or resets
through the batched dynamic path, every shot's clbits identical to the oracle shot loop's."""
from pathlib import Path

import numpy as np
import pytest

from goqu_amd import densitymatrix as D
from goqu_amd import qasm
from goqu_amd import statevector
from oracle import oracle as O
from test_oracle_dynamic import OracleDynamic

pytestmark = pytest.mark.gpu

CORPUS = sorted((Path(__file__).parent / "golden" / "qasm").glob("*.qasm"))


@pytest.mark.parametrize("f", CORPUS, ids=lambda f: f.name)
def test_corpus_statevector_and_density_matrix(f):
    c = qasm.ParseFile(str(f))
    n = c.NumQubits()
    print(f"qasm {f.name} n={n} dynamic={c.IsDynamic()}", flush=True)
    if c.IsDynamic():
        U = np.random.default_rng(len(f.name)).random((300, 64))
        sv = statevector.New(n)
        got = sv.RunDynamicShots(c, 300, uniforms=U)
        want, psi = OracleDynamic(n).run_shots(c, U)
        assert np.array_equal(got, want)
        assert np.max(np.abs(sv.StateVector() - psi)) < 1e-12
        sv.Close()
        dm = D.New(n)
        got = dm.RunDynamicShots(c, 300, uniforms=U)
        want, rho = OracleDynamic(n, density=True).run_shots(c, U)
        assert np.array_equal(got, want)
        assert np.max(np.abs(dm.DensityMatrix() - rho)) < 1e-12
        dm.Close()
        return
    psi = O.sv_evolve(c)
    sv = statevector.New(n)
    sv.Evolve(c)
    assert np.max(np.abs(sv.StateVector() - psi)) < 1e-12
    u = np.random.default_rng(7).random(1000)
    assert np.array_equal(sv.Sample(u), O.sv_sample(psi, u))
    sv.Close()
    dm = D.New(n)
    dm.Evolve(c)
    rho = O.dm_zero(n)
    if O.dm_apply_ops(rho, n, D.marshal(c)) == 0:
        assert np.max(np.abs(dm.DensityMatrix() - rho)) < 1e-12
    # the oracle's density matrix stops at 2-qubit gates (the reference decomposes larger controlled
    # gates, sim.go:321-518; this package applies them exactly): noiseless, rho = |psi><psi| in any case
    assert np.max(np.abs(dm.DensityMatrix() - np.outer(psi, psi.conj()).reshape(-1))) < 1e-12
    dm.Close()


def test_corpus_reference_outcomes():
    """Outcomes the circuits are written to produce."""
    run = lambda name, shots=400: statevector.New(qasm.ParseFile(str(Path(__file__).parent / "golden" / "qasm" /
                                                                      name)).NumQubits()).Run(
        qasm.ParseFile(str(Path(__file__).parent / "golden" / "qasm" / name)), shots, rng=np.random.default_rng(3))
    assert set(run("bell.qasm")) == {"00", "11"}
    assert set(run("ghz3.qasm")) == {"000", "111"}
    assert run("grover2.qasm") == {"11": 400}
    assert run("toffoli.qasm") == {"111": 400}
    assert run("swap_test.qasm") == {"10": 400}
    assert set(run("conditional.qasm")) == {"00", "11"}
