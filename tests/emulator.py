"""ctypes binding of tests/cpp/stage_emulator.cpp (built with g++ on demand)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from goqu_amd import _native as N

HERE = Path(__file__).resolve().parent / "cpp"
CSRC = HERE.parent.parent / "goqu_amd" / "csrc"
SRC = [HERE / "stage_emulator.cpp", CSRC / "planner.cpp", CSRC / "adjoint_plan.cpp"]
LIB = HERE / "_build" / "libstage_emu.so"
_lib = None


def lib():
    global _lib
    if _lib is None:
        deps = SRC + [HERE.parent.parent / "goqu_amd" / "csrc" / "engine.h"]
        if not LIB.exists() or any(p.stat().st_mtime > LIB.stat().st_mtime for p in deps):
            LIB.parent.mkdir(exist_ok=True)
            # build under a private name, then rename: parallel test workers may build at once
            tmp = LIB.with_name(f"{LIB.name}.{os.getpid()}.tmp")
            subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", str(tmp), *map(str, SRC)],
                           check=True)
            os.replace(tmp, LIB)
        L = C.CDLL(str(LIB))
        L.emu_apply.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_int, C.POINTER(N.GoquOp), C.c_int, C.c_int,
                                C.c_int, C.c_int, C.POINTER(C.c_int64)]
        L.emu_apply.restype = C.c_int
        L.emu_error.restype = C.c_char_p
        L.emu_sv_reverse.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.c_int,
                                     C.POINTER(N.GoquOp), C.c_int, C.POINTER(N.GoquDeriv), C.c_int, C.c_int,
                                     C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.emu_sv_reverse.restype = C.c_int
        L.emu_dm_reverse.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.POINTER(N.GoquOp),
                                     C.c_int, C.c_int, C.POINTER(N.GoquDeriv), C.c_int, C.c_int, C.c_int,
                                     C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.emu_dm_reverse.restype = C.c_int
        L.emu_coverage.argtypes = [C.POINTER(C.c_int64), C.c_int, C.c_int]
        L.emu_coverage.restype = C.c_int
        _lib = L
    return _lib


def apply(state: np.ndarray, nq_user: int, buf: "N.OpBuffer", tile_bits: int = 12, low_bits: int = 3,
          dm: bool = False, nq_internal: int | None = None):
    """Runs the planned program on `state` (complex128, 2^nq_internal entries) in place."""
    nqi = nq_internal if nq_internal is not None else int(np.log2(len(state)))
    ns = C.c_int64(0)
    rc = lib().emu_apply(state.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), nqi, nq_user, buf.array(),
                         len(buf), tile_bits, low_bits, int(dm), C.byref(ns))
    if rc:
        raise RuntimeError(f"emulator status {rc}: {lib().emu_error().decode()}")
    return ns.value


def sv_reverse(psi: np.ndarray, lam: np.ndarray, nq_user: int, buf: "N.OpBuffer", derivs, nd: int,
               tile_bits: int = 12, low_bits: int = 3):
    """The statevector adjoint's fused reverse sweep on psi and lam (complex128, padded to the
    internal width) in place.  Returns (per-derivative Re <lam|D|psi>, stage count)."""
    nqi = int(np.log2(len(psi)))
    vals = np.zeros(max(nd, 1))
    ns = C.c_int64(0)
    dp = lambda a: a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    rc = lib().emu_sv_reverse(dp(psi), dp(lam), nqi, nq_user, buf.array(), len(buf), derivs, nd, tile_bits,
                              low_bits, vals.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ns))
    if rc:
        raise RuntimeError(f"emulator status {rc}: {lib().emu_error().decode()}")
    return vals[:nd], ns.value


def dm_reverse(rho: np.ndarray, h: np.ndarray, n: int, ops, k0: int, k1: int, derivs, nd: int, tile_bits: int = 8,
               low_bits: int = 3):
    """The density-matrix adjoint's fused reverse sweep over ops [k0, k1) on rho and h (complex128
    vectors of 2n qubits, padded to the internal width) in place.  Returns (per-derivative
    Re <h|D rho>, stage count); raises when the ops do not allow the fused sweep."""
    vals = np.zeros(max(nd, 1))
    ns = C.c_int64(0)
    dp = lambda a: a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    rc = lib().emu_dm_reverse(dp(rho), dp(h), n, ops, k0, k1, derivs, nd, tile_bits, low_bits,
                              vals.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ns))
    if rc:
        raise RuntimeError(f"emulator status {rc}: {lib().emu_error().decode()}")
    return vals[:nd], ns.value


COVERAGE = ["dense", "dense_tctrl", "dense_pat", "dense_tile", "dense_big", "diag_uniform", "diag_kl1",
            "diag_kl2", "diag_tctrl", "diag_tile", "diag_regs", "anti_pure", "anti_phased", "anti_tctrl",
            "anti_pat", "swap_pure", "swap_phased", "swap_tctrl", "swap_pat", "remap", "grad", "grad_ctrl",
            "reset", "parity2", "parity4", "parity_real", "shear2", "shear3", "xswap_perm",
            "xswap_dpp", "diag_full", "diag_slot", "diag_tslot", "shears"]


def coverage(reset: bool = True) -> dict:
    """Instruction categories (stage-kernel code paths) executed since the last reset."""
    out = (C.c_int64 * 64)()
    n = lib().emu_coverage(out, 64, int(reset))
    return {COVERAGE[i]: out[i] for i in range(min(n, len(COVERAGE)))}
