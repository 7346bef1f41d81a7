"""Observability hooks: a mirror of goqu's observe package (observe/observe.go:1-120) for the
callers here that emit them (sweep.RunSim / RunDensitySim, the local backend's Submit).

A context is a plain dict; WithHooks returns a copy carrying the hooks, FromContext reads them.
Each Wrap* hook is called before the work with (ctx, info) and returns (ctx, done); done(err) is
called when the work ends, err None on success."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Callable, Dict, Optional, Tuple

Context = Dict[str, Any]
_KEY = "goqu.observe.hooks"


@dataclass
class SimInfo:
    NumQubits: int
    GateCount: int
    Shots: int


@dataclass
class SweepInfo:
    NumPoints: int
    NumParams: int
    Shots: int
    NumQubits: int


@dataclass
class Hooks:
    """observe.Hooks (observe.go:7-31): any hook left None is skipped."""

    WrapSim: Optional[Callable[[Context, SimInfo], Tuple[Context, Callable[[Optional[BaseException]], None]]]] = None
    WrapSweep: Optional[Callable[[Context, SweepInfo], Tuple[Context, Callable[[Optional[BaseException]], None]]]] = None


def Background() -> Context:
    return {}


def WithHooks(ctx: Optional[Context], h: Hooks) -> Context:
    """observe.WithHooks (observe.go:104-107)."""
    out = dict(ctx or {})
    out[_KEY] = h
    return out


def FromContext(ctx: Optional[Context]) -> Optional[Hooks]:
    """observe.FromContext (observe.go:109-113)."""
    if not ctx:
        return None
    h = ctx.get(_KEY)
    return h if isinstance(h, Hooks) else None
