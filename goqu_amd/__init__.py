"""goqu_amd — an MI355X (gfx950) simulation backend for goqu.

The hot path lives in libgoqu_hip.so (HIP kernels + C ABI, include/goqu_hip.h).  The Python
modules mirror the slice of goqu's Go API that drives it: gates (circuit/gate), circuits and
parameters (circuit/ir, circuit/param, circuit/builder), Pauli observables (sim/pauli), noise
(sim/noise), the statevector and density-matrix simulators (sim/statevector,
sim/densitymatrix) and gradients (algorithm/gradient).
"""
from . import gate, circuit, pauli, noise  # noqa: F401

__all__ = ["gate", "circuit", "pauli", "noise", "statevector", "densitymatrix", "gradient"]
