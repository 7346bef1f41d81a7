"""Parameter sweeps: a mirror of goqu's sweep package (sweep/sweep.go, product.go, zip.go,
run.go) whose RunSim / RunDensitySim evaluate all points in one device batch (goqu_amd.batch)
instead of one fresh simulator per point on GOMAXPROCS goroutines (run.go:27-171).

Counts are those of Run on each bound circuit (sim.go:40-58, densitymatrix sim.go:51-67): the
point's shots sample its final state; uniforms come from `rng` (or per point from `uniforms`).
Dynamic circuits run point by point through RunDynamic, whose shots are batched on the device."""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import batch as B
from . import observe
from .circuit import Bind, Circuit, FreeParameters


class Sweep:
    """sweep.Sweep (sweep.go:5-13)."""

    def Len(self) -> int:
        raise NotImplementedError

    def Params(self) -> List[str]:
        raise NotImplementedError

    def Resolve(self) -> List[Dict[str, float]]:
        raise NotImplementedError


@dataclass
class Linspace(Sweep):
    """Count points from Start to Stop inclusive (sweep.go:17-44)."""

    Key: str
    Start: float
    Stop: float
    Count: int

    def Len(self) -> int:
        return max(self.Count, 0)

    def Params(self) -> List[str]:
        return [self.Key]

    def Resolve(self) -> List[Dict[str, float]]:
        n = self.Len()
        if n == 0:
            return []
        if n == 1:
            return [{self.Key: self.Start}]
        step = (self.Stop - self.Start) / (n - 1)
        return [{self.Key: self.Start + i * step} for i in range(n)]


class Points(Sweep):
    """Explicit values of one parameter (sweep.go:46-70)."""

    def __init__(self, key: str, values: Sequence[float]):
        self._key = key
        self._values = [float(v) for v in values]

    def Len(self) -> int:
        return len(self._values)

    def Params(self) -> List[str]:
        return [self._key]

    def Resolve(self) -> List[Dict[str, float]]:
        return [{self._key: v} for v in self._values]


def NewPoints(key: str, values: Sequence[float]) -> Points:
    return Points(key, values)


class UnitSweep(Sweep):
    """One point with fixed bindings (sweep.go:72-103)."""

    def __init__(self, bindings: Dict[str, float]):
        self._b = dict(bindings)

    def Len(self) -> int:
        return 1

    def Params(self) -> List[str]:
        return list(self._b)

    def Resolve(self) -> List[Dict[str, float]]:
        return [dict(self._b)]


def Single(bindings: Dict[str, float]) -> UnitSweep:
    return UnitSweep(bindings)


class _Resolved(Sweep):
    def __init__(self, params: List[str], points: List[Dict[str, float]]):
        self._params = params
        self._points = points

    def Len(self) -> int:
        return len(self._points)

    def Params(self) -> List[str]:
        return list(self._params)

    def Resolve(self) -> List[Dict[str, float]]:
        return [dict(p) for p in self._points]


def _params_of(sweeps: Sequence[Sweep]) -> List[str]:
    out: List[str] = []
    for s in sweeps:
        for p in s.Params():
            if p not in out:
                out.append(p)
    return out


def Product(*sweeps: Sweep) -> Sweep:
    """Cartesian product, the first sweep outermost (product.go:12-53); an empty factor empties it."""
    params = _params_of(sweeps)
    if not sweeps:
        return _Resolved([], [])
    results: List[Dict[str, float]] = [{}]
    for s in sweeps:
        resolved = s.Resolve()
        if not resolved:
            return _Resolved(params, [])
        results = [{**e, **p} for e in results for p in resolved]
    return _Resolved(params, results)


def Zip(*sweeps: Sweep) -> Sweep:
    """Point-wise merge of equal-length sweeps (zip.go:14-56)."""
    if not sweeps:
        return _Resolved([], [])
    n = sweeps[0].Len()
    for i, s in enumerate(sweeps[1:], start=1):
        if s.Len() != n:
            raise ValueError(f"sweep.Zip: length mismatch: sweep 0 has {n} points, sweep {i} has {s.Len()}")
    resolved = [s.Resolve() for s in sweeps]
    points = []
    for i in range(n):
        merged: Dict[str, float] = {}
        for r in resolved:
            merged.update(r[i])
        points.append(merged)
    return _Resolved(_params_of(sweeps), points)


MustZip = Zip


@dataclass
class PointResult:
    """sweep.PointResult (run.go:17-23)."""

    Index: int
    Bindings: Dict[str, float]
    Counts: Dict[str, int] = field(default_factory=dict)
    Err: Optional[BaseException] = None


def validate_params(c: Circuit, sw: Sweep) -> None:
    """validateParams (run.go:173-190)."""
    have = set(sw.Params())
    for f in FreeParameters(c):
        if f not in have:
            raise ValueError(f"sweep: circuit has free parameter {f!r} not covered by sweep")


def _counts(idx: np.ndarray, n: int) -> Dict[str, int]:
    vals, cnt = np.unique(idx, return_counts=True)
    return {format(int(v), f"0{n}b"): int(k) for v, k in zip(vals, cnt)}


def _run(ctx, c: Circuit, shots: int, sw: Sweep, density: bool, nm, rng, uniforms) -> List[PointResult]:
    bindings = sw.Resolve()
    validate_params(c, sw)
    nq = c.NumQubits()
    hooks = observe.FromContext(ctx)
    done = None
    if hooks is not None and hooks.WrapSweep is not None:
        ctx, done = hooks.WrapSweep(ctx, observe.SweepInfo(len(bindings), len(sw.Params()), shots, nq))
    results = [PointResult(i, b) for i, b in enumerate(bindings)]
    try:
        bound = []
        for r in results:
            try:
                bound.append(Bind(c, r.Bindings))
            except Exception as e:  # noqa: BLE001 - the reference stores bind errors per point
                r.Err = e
                bound.append(None)
        live = [i for i, b in enumerate(bound) if b is not None]
        sim_done = []
        if hooks is not None and hooks.WrapSim is not None:
            for i in live:  # per-point WrapSim, as runStatevectorPoint fires it (run.go:82-90)
                _, d = hooks.WrapSim(ctx, observe.SimInfo(nq, len(bound[i].Ops()), shots))
                sim_done.append(d)
        rng = rng if rng is not None else np.random.default_rng()
        err: Optional[BaseException] = None
        try:
            if live and bound[live[0]].IsDynamic():
                for i in live:
                    if density:
                        from .densitymatrix import New as DM

                        s = DM(nq)
                        if nm is not None:
                            s.WithNoise(nm)
                    else:
                        from .statevector import New as SV

                        s = SV(nq)
                    results[i].Counts = s.RunDynamic(bound[i], shots, rng=rng)
                    s.Close()
            elif live and shots > 0:
                u = uniforms[live] if uniforms is not None else rng.random((len(live), shots))
                if density:
                    idx = B.dm_sample([bound[i] for i in live], shots, nm, uniforms=u)
                else:  # the points' op records straight from the parameter columns
                    names = FreeParameters(c)
                    params = np.array([[results[i].Bindings[k] for k in names] for i in live],
                                      dtype=np.float64).reshape(len(live), len(names))
                    idx = B.param_sample(c, params, shots, names, uniforms=u)
                for k, i in enumerate(live):
                    results[i].Counts = _counts(idx[k], nq)
        except Exception as e:  # noqa: BLE001
            err = e
            for i in live:
                results[i].Err = e
        for d in sim_done:
            d(err)
    finally:
        if done is not None:
            done(None)
    return results


def RunSim(ctx, c: Circuit, shots: int, sw: Sweep, rng=None, uniforms=None) -> List[PointResult]:
    """sweep.RunSim (run.go:27-73): every point's counts, in sweep order."""
    return _run(ctx, c, shots, sw, False, None, rng, uniforms)


def RunDensitySim(ctx, c: Circuit, shots: int, sw: Sweep, nm=None, rng=None, uniforms=None) -> List[PointResult]:
    """sweep.RunDensitySim (run.go:100-146) with an optional noise model."""
    return _run(ctx, c, shots, sw, True, nm, rng, uniforms)
