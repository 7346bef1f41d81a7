"""The simulator plugin point: a mirror of goqu's local backend (backend/local/local.go:21-160)
with this package's device simulator as the factory, the wiring a Go host writes as
`local.New(local.WithMaxQubits(34), local.WithSimulator(rocm.NewSimulator))`.

NewSimulator is the factory (local.SimFactory, local.go:23-25); New builds the backend with the
reference's defaults (28 qubits at most, a statevector factory) and its options.  Submit runs the
circuit synchronously inside the context's WrapSim hook (local.go:95-123) and stores the counts
under a fresh job id.  Pulse programs (local.go:163-) are not part of this backend."""
from __future__ import annotations

import secrets
import threading
from dataclasses import dataclass, field
from typing import Any, Callable, Dict, Optional

from . import observe
from . import statevector
from .circuit import Circuit

StateCompleted = "completed"


@dataclass
class SubmitRequest:
    Circuit: Optional[Circuit] = None
    Shots: int = 0
    PulseProgram: Any = None
    Options: Dict[str, Any] = field(default_factory=dict)


@dataclass
class Job:
    ID: str
    Backend: str
    State: str


@dataclass
class JobStatus:
    ID: str
    State: str
    Progress: float


@dataclass
class Result:
    Counts: Dict[str, int]
    Shots: int


def NewSimulator(num_qubits: int):
    """The device statevector as a local.SimFactory (sim.Simulator, sim/sim.go:9-22)."""
    return statevector.New(num_qubits)


class Backend:
    def __init__(self):
        self.maxQubits = 28  # local.go:57
        self.simFactory: Callable[[int], Any] = statevector.New
        self._results: Dict[str, Result] = {}
        self._lock = threading.Lock()

    def Name(self) -> str:
        return "local.simulator"

    def Submit(self, ctx, req: SubmitRequest) -> Job:
        if req.PulseProgram is not None:
            raise NotImplementedError("local: pulse programs are not simulated by this backend")
        if req.Circuit is None:
            raise ValueError("local: nil circuit")
        nq = req.Circuit.NumQubits()
        if nq > self.maxQubits:
            raise ValueError(f"local: circuit has {nq} qubits, max is {self.maxQubits}")
        if req.Shots <= 0:
            raise ValueError("local: shots must be positive")
        hooks = observe.FromContext(ctx)
        done = None
        if hooks is not None and hooks.WrapSim is not None:
            ctx, done = hooks.WrapSim(ctx, observe.SimInfo(nq, _gate_count(req.Circuit), req.Shots))
        try:
            s = self.simFactory(nq)
        except Exception as e:
            if done is not None:
                done(e)
            raise
        err = None
        try:
            counts = s.Run(req.Circuit, req.Shots)
        except Exception as e:
            err = e
            raise
        finally:
            s.Close()
            if done is not None:
                done(err)
        jid = secrets.token_hex(8)
        with self._lock:
            self._results[jid] = Result(counts, req.Shots)
        return Job(jid, self.Name(), StateCompleted)

    def Status(self, ctx, job_id: str) -> JobStatus:
        with self._lock:
            if job_id not in self._results:
                raise KeyError(f"local: unknown job {job_id}")
        return JobStatus(job_id, StateCompleted, 1.0)

    def Result(self, ctx, job_id: str) -> Result:
        with self._lock:
            if job_id not in self._results:
                raise KeyError(f"local: unknown job {job_id}")
            return self._results[job_id]

    def Cancel(self, ctx, job_id: str) -> None:
        return None


def _gate_count(c: Circuit) -> int:
    """Stats().GateCount (circuit/ir/circuit.go:134-166): every op, control-flow bodies included."""
    def count(ops):
        k = len(ops)
        for op in ops:
            if op.ControlFlow is not None:
                for body in op.ControlFlow.Bodies:
                    k += count(body)
        return k

    return count(c.Ops())


def WithMaxQubits(n: int):
    def opt(b: Backend):
        b.maxQubits = n
    return opt


def WithSimulator(factory: Callable[[int], Any]):
    def opt(b: Backend):
        b.simFactory = factory
    return opt


def New(*opts) -> Backend:
    """local.New (local.go:55-67)."""
    b = Backend()
    for o in opts:
        o(b)
    return b
