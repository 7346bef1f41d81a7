"""Tile-shape autotuner for the stage engine.

The stage planner fuses every run of gates on the qubits of one tile into one sweep of the state
(2^tile_bits amplitudes per tile, of which the low_bits lowest qubits are in every tile).  The
tile width is the fusion width: wider tiles fuse more gates per sweep (fewer sweeps) but cost more
LDS traffic, registers and remap work per sweep; low_bits trades contiguous HBM runs against
fusable low qubits.  The best shape depends on the register size, the precision and the circuit
(its mix of low/high qubits and gate kinds), so it is measured: `tune` times the caller's circuit
under each candidate shape on the device (goqu_sv_set_plan) and keeps the fastest; the choice is
persisted as JSON keyed by device, precision, qubit count and a workload label, and reused by
`plan_for` (statevector.New(n, plan="tuned"), the batch entry points' plan argument).

The cache file is GOQU_TUNE_CACHE, or ~/.cache/goqu_amd/tuning.json.
"""
from __future__ import annotations

import json
import os
import time
from pathlib import Path
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

from . import _native as N

Shape = Tuple[int, int]

# tile_bits x low_bits the stage kernel accepts (goqu_sv_set_plan: tile 8..13, low 2..min(6, tile-4))
DEFAULT_CANDIDATES: List[Shape] = [(t, l) for t in (10, 11, 12, 13) for l in (3, 4, 5) if l <= t - 4]


def cache_path() -> Path:
    p = os.environ.get("GOQU_TUNE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "goqu_amd" / "tuning.json"


def _load() -> Dict[str, dict]:
    p = cache_path()
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return {}


def _store(db: Dict[str, dict]) -> None:
    p = cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        tmp = p.with_suffix(".tmp")
        tmp.write_text(json.dumps(db, indent=1, sort_keys=True))
        os.replace(tmp, p)
    except OSError:
        pass  # a read-only home: the choice still applies to this process


def _device() -> str:
    import ctypes as C

    buf = C.create_string_buffer(128)
    N.check(N.lib().goqu_device_name(buf, 128))
    return buf.value.decode(errors="replace").split(":")[0]


def key(n: int, precision: int, workload: str = "generic") -> str:
    return f"{_device()}|sv|{'c128' if precision == N.C128 else 'c64'}|n={n}|{workload}"


def plan_for(n: int, precision: int = N.C128, workload: str = "generic") -> Shape:
    """The persisted choice, or (0, 0) (the library's default shape) when none was tuned."""
    e = _load().get(key(n, precision, workload))
    if not e:
        return (0, 0)
    return (int(e["tile_bits"]), int(e["low_bits"]))


def _valid(shape: Shape, nq_internal: int) -> bool:
    t, l = shape
    return 8 <= t <= 13 and 2 <= l <= min(6, t - 4) and t <= max(nq_internal, 8)


def tune(circuit, precision: int = N.C128, workload: str = "generic", candidates: Optional[Iterable[Shape]] = None,
         reps: int = 2, persist: bool = True) -> Tuple[Shape, Dict[str, float]]:
    """Times Evolve(circuit) under every candidate shape (after one warm-up each) and returns the
    fastest with the per-shape seconds; persists it under key(n, precision, workload)."""
    from . import statevector

    n = circuit.NumQubits()
    sim = statevector.New(n, precision)
    buf = N.OpBuffer()
    for op in circuit.Ops():
        if op.Gate is not None:
            N.add_sv_gate(buf, op.Gate, op.Qubits)
    arr, nops = buf.array(), len(buf)
    lib = N.lib()
    times: Dict[str, float] = {}
    best, best_t = (0, 0), float("inf")
    try:
        for shape in (candidates or DEFAULT_CANDIDATES):
            if not _valid(shape, max(n, 6)):
                continue
            N.check(lib.goqu_sv_set_plan(sim._h, shape[0], shape[1]))
            N.check(lib.goqu_sv_reset(sim._h))
            N.check(lib.goqu_sv_apply(sim._h, arr, nops))  # warm-up (also pages the state in)
            t0 = time.perf_counter()
            for _ in range(reps):
                N.check(lib.goqu_sv_reset(sim._h))
                N.check(lib.goqu_sv_apply(sim._h, arr, nops))
            dt = (time.perf_counter() - t0) / reps
            times[f"{shape[0]},{shape[1]}"] = dt
            if dt < best_t:
                best, best_t = shape, dt
    finally:
        sim.Close()
    if persist and best != (0, 0):
        db = _load()
        db[key(n, precision, workload)] = {"tile_bits": best[0], "low_bits": best[1], "seconds": best_t,
                                           "candidates": times}
        _store(db)
    return best, times
