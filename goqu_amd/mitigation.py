"""Executors of goqu's mitigation package (algorithm/mitigation/executor.go:13-41), plus batch
forms that evaluate many circuits in one device pass (goqu_amd.batch): the folded circuits of
zero-noise extrapolation or the twirled copies of a circuit are evaluated together."""
from __future__ import annotations

from typing import Callable, List, Optional, Sequence

import numpy as np

from . import batch as B
from .circuit import Circuit
from .pauli import PauliSum

Executor = Callable[[Optional[dict], Circuit], float]


def StatevectorExecutor(hamiltonian: PauliSum) -> Executor:
    """<psi|H|psi> by ideal statevector simulation (executor.go:17-29)."""
    def run(_ctx, circuit: Circuit) -> float:
        return float(B.energies([circuit], hamiltonian)[0])

    return run


def DensityMatrixExecutor(hamiltonian: PauliSum, nm=None) -> Executor:
    """Tr(rho H) by density-matrix simulation with a noise model (executor.go:31-41)."""
    def run(_ctx, circuit: Circuit) -> float:
        return float(B.dm_energies([circuit], hamiltonian, nm)[0])

    return run


def StatevectorBatchExecutor(hamiltonian: PauliSum) -> Callable[[Optional[dict], Sequence[Circuit]], np.ndarray]:
    """The statevector executor over many circuits at once (one device batch)."""
    def run(_ctx, circuits: Sequence[Circuit]) -> np.ndarray:
        return B.energies(list(circuits), hamiltonian)

    return run


def DensityMatrixBatchExecutor(hamiltonian: PauliSum, nm=None):
    """The density-matrix executor over many circuits at once (one device batch)."""
    def run(_ctx, circuits: Sequence[Circuit]) -> np.ndarray:
        return B.dm_energies(list(circuits), hamiltonian, nm)

    return run
