"""Kraus channels and the noise model — host-side mirror of goqu's sim/noise
(noise.go:9-90, channel.go:27-298, readout.go)."""
from __future__ import annotations

import math
from typing import Dict, List, Optional, Sequence

import numpy as np


class Channel:
    def __init__(self, name: str, nq: int, kraus: List[np.ndarray]):
        self._name = name
        self._nq = nq
        self._kraus = [np.asarray(k, dtype=np.complex128).reshape(-1) for k in kraus]

    def Name(self) -> str:
        return self._name

    def Qubits(self) -> int:
        return self._nq

    def Kraus(self) -> List[np.ndarray]:
        return list(self._kraus)


def _chk(name, p, what="p"):
    if p < 0 or p > 1:
        raise ValueError(f"noise.{name}: {what}={p:f} out of range [0,1]")


def Depolarizing1Q(p: float) -> Channel:
    """channel.go:27-43"""
    _chk("Depolarizing1Q", p)
    s0, sp = math.sqrt(1 - p), math.sqrt(p / 3)
    return Channel(f"depolarizing1q({p:.4f})", 1, [
        [s0, 0, 0, s0], [0, sp, sp, 0], [0, -sp * 1j, sp * 1j, 0], [sp, 0, 0, -sp]])


def _tensor2(a, b):
    """tensorProduct2x2 (channel.go:88-103): a on the MSB (qubits[0])."""
    r = np.zeros(16, dtype=np.complex128)
    for i in range(2):
        for j in range(2):
            for k in range(2):
                for l in range(2):
                    r[(i * 2 + k) * 4 + (j * 2 + l)] = a[i * 2 + j] * b[k * 2 + l]
    return r


def Depolarizing2Q(p: float) -> Channel:
    """channel.go:50-86"""
    _chk("Depolarizing2Q", p)
    paulis = [np.array(v, dtype=np.complex128) for v in ([1, 0, 0, 1], [0, 1, 1, 0], [0, -1j, 1j, 0], [1, 0, 0, -1])]
    s0, sp = math.sqrt(1 - p), math.sqrt(p / 15)
    ks = []
    for i in range(4):
        for j in range(4):
            ks.append(_tensor2(paulis[i], paulis[j]) * (s0 if i == 0 and j == 0 else sp))
    return Channel(f"depolarizing2q({p:.4f})", 2, ks)


def AmplitudeDamping(gamma: float) -> Channel:
    """channel.go:114-128"""
    _chk("AmplitudeDamping", gamma, "gamma")
    return Channel(f"amplitude_damping({gamma:.4f})", 1, [[1, 0, 0, math.sqrt(1 - gamma)], [0, math.sqrt(gamma), 0, 0]])


def PhaseDamping(lam: float) -> Channel:
    """channel.go:139-153"""
    _chk("PhaseDamping", lam, "lambda")
    return Channel(f"phase_damping({lam:.4f})", 1, [[1, 0, 0, math.sqrt(1 - lam)], [0, 0, 0, math.sqrt(lam)]])


def BitFlip(p: float) -> Channel:
    _chk("BitFlip", p)
    s0, sp = math.sqrt(1 - p), math.sqrt(p)
    return Channel(f"bit_flip({p:.4f})", 1, [[s0, 0, 0, s0], [0, sp, sp, 0]])


def PhaseFlip(p: float) -> Channel:
    _chk("PhaseFlip", p)
    s0, sp = math.sqrt(1 - p), math.sqrt(p)
    return Channel(f"phase_flip({p:.4f})", 1, [[s0, 0, 0, s0], [sp, 0, 0, -sp]])


def GeneralizedAmplitudeDamping(p: float, gamma: float) -> Channel:
    """channel.go:217-238"""
    _chk("GeneralizedAmplitudeDamping", p)
    _chk("GeneralizedAmplitudeDamping", gamma, "gamma")
    sp, s1p, sg, s1g = math.sqrt(p), math.sqrt(1 - p), math.sqrt(gamma), math.sqrt(1 - gamma)
    return Channel(f"generalized_amplitude_damping({p:.4f},{gamma:.4f})", 1, [
        [sp, 0, 0, sp * s1g], [0, sp * sg, 0, 0], [s1p * s1g, 0, 0, s1p], [0, 0, s1p * sg, 0]])


def ThermalRelaxation(t1: float, t2: float, time: float) -> Channel:
    """channel.go:246-298"""
    if t2 > 2 * t1:
        raise ValueError(f"noise.ThermalRelaxation: t2={t2:f} > 2*t1={2 * t1:f} is unphysical")
    if t1 <= 0 or t2 <= 0 or time < 0:
        raise ValueError("noise.ThermalRelaxation: t1, t2 must be positive, time must be non-negative")
    gamma = 1 - math.exp(-time / t1)
    e_l = math.exp(-time / t2)
    lam = 0.0
    if gamma < 1:
        ideal = math.sqrt(1 - gamma)
        if e_l < ideal:
            ratio = e_l / ideal
            lam = 1 - ratio * ratio
    sg, s1g, sl, s1l = math.sqrt(gamma), math.sqrt(1 - gamma), math.sqrt(lam), math.sqrt(1 - lam)
    ks = [[1, 0, 0, s1l * s1g], [0, sg, 0, 0]]
    if lam > 1e-15:
        ks.append([0, 0, 0, sl * s1g])
    return Channel(f"thermal_relaxation({t1:.4g},{t2:.4g},{time:.4g})", 1, ks)


class ReadoutError:
    """Per-qubit readout flip probabilities (sim/noise/readout.go)."""

    def __init__(self, p01: float, p10: float):
        self.P01 = p01
        self.P10 = p10


def NewReadoutError(p01: float, p10: float) -> ReadoutError:
    return ReadoutError(p01, p10)


class NoiseModel:
    """noise.NoiseModel (noise.go:19-90): gate+qubits > gate name > qubit-count default."""

    def __init__(self):
        self.by_gate_qubits: Dict[str, Channel] = {}
        self.by_gate: Dict[str, Channel] = {}
        self.by_qubits: Dict[int, Channel] = {}
        self.readout: Dict[int, ReadoutError] = {}

    @staticmethod
    def _key(name: str, qubits: Sequence[int]) -> str:
        return name + ":" + ",".join(str(q) for q in qubits)

    def AddGateError(self, name: str, ch: Channel):
        self.by_gate[name] = ch

    def AddGateQubitError(self, name: str, qubits: Sequence[int], ch: Channel):
        self.by_gate_qubits[self._key(name, qubits)] = ch

    def AddDefaultError(self, nq: int, ch: Channel):
        self.by_qubits[nq] = ch

    def AddReadoutError(self, q: int, re: ReadoutError):
        self.readout[q] = re

    def Lookup(self, name: str, qubits: Sequence[int]) -> Optional[Channel]:
        k = self._key(name, qubits)
        if k in self.by_gate_qubits:
            return self.by_gate_qubits[k]
        if name in self.by_gate:
            return self.by_gate[name]
        return self.by_qubits.get(len(qubits))

    def ReadoutFor(self, q: int) -> Optional[ReadoutError]:
        return self.readout.get(q)


def New() -> NoiseModel:
    return NoiseModel()
