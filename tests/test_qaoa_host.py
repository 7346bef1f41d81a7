"""Host-side checks of the batched QAOA path: the Hamiltonian and circuit constructors
against the reference's definitions, and the tile index maps of the large-instance sweeps
(every amplitude visited exactly once per phase, all indices inside the state)."""
import numpy as np
import pytest

from goqu_amd import qaoa as Q
from goqu_amd.circuit import FreeParameters
from goqu_amd.pauli import Z


def tile_global(n, phase_a, tile, l, tb=12):
    if phase_a:
        return (tile << tb) | l
    lb = 2 * tb - n
    lo, hi = l & ((1 << lb) - 1), l >> lb
    return lo | (tile << lb) | (hi << tb)


@pytest.mark.parametrize("n", [13, 16, 19, 22])
def test_tile_maps_are_bijections(n):
    tiles = 1 << (n - 12)
    for phase_a in (True, False):
        t = np.arange(tiles, dtype=np.int64)[:, None]
        l = np.arange(4096, dtype=np.int64)[None, :]
        if phase_a:
            g = (t << 12) | l
        else:
            lb = 24 - n
            g = (l & ((1 << lb) - 1)) | (t << lb) | ((l >> lb) << 12)
        g = g.reshape(-1)
        assert g.min() == 0 and g.max() == (1 << n) - 1
        assert len(np.unique(g)) == 1 << n
        if not phase_a:  # mixer qubits of phase B are 12..n-1 (local bits lb..11)
            lb = 24 - n
            for b in range(lb, 12):
                assert tile_global(n, False, 0, 1 << b) == 1 << (12 + b - lb)


def test_maxcut_hamiltonian_and_circuit():
    H = Q.MaxCutHamiltonian([(0, 1), (1, 2)], 3)
    terms = H.Terms()
    assert len(terms) == 4
    assert terms[0].IsIdentity() and terms[0].Coeff() == -0.5
    assert terms[1].Op(0) == Z and terms[1].Op(1) == Z and terms[1].Coeff() == 0.5
    c = Q.QAOACircuit(H, 2)
    assert FreeParameters(c) == ["γ[0]", "β[0]", "γ[1]", "β[1]"]
    names = [op.Gate.Name() for op in c.Ops()]
    assert names[:3] == ["H", "H", "H"]
    assert names.count("RZZ") == 4 and names.count("RX") == 6
    with pytest.raises(ValueError):
        Q.MaxCutHamiltonian([], 3)


def test_efficient_su2_param_count():
    c, npar = Q.EfficientSU2(24, 7)
    assert npar == 384
    assert len(FreeParameters(c)) == 384
    assert sum(1 for op in c.Ops() if op.Gate.Name() == "CNOT") == 7 * 24
