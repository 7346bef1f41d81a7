"""Benchmark: 30-qubit complex128 gate application on one MI355X.

Workload (one "step"): Evolve of the depth-40 seeded random circuit on 30 qubits —
U3(theta, phi, lambda ~ U[0, 2pi)) on every qubit per layer, then 15 disjoint random pairs
(a quarter of them spanning a low and a high qubit) each carrying CZ, CNOT, SWAP or CP.
The state (16 GiB) stays resident in HBM; the timed region is the K Evolves.

Printed JSON line:
  value     = HBM sweep throughput of the whole step: stage sweeps x 32 B x 2^30 / step time
  roofline  = the stage kernel (k_stage): algorithmic bytes per launch (read + write of the
              whole state, 32 B x 2^30) / its mean launch duration from HIP events recorded on
              its stream, against the 8 TB/s HBM peak; traffic = HBM bytes per launch from the
              committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary under profiles/
  cpu_baseline = the oracle (a C restatement of goqu's CPU statevector with its worker policy)
              applying the first gates of the same circuit at 30 qubits on the host cores;
              bytes per gate as for the GPU metric (32 B x 2^n, a quarter for CZ/CP, half for
              CNOT/SWAP), divided by its wall time.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import numpy as np  # noqa: E402

HBM_PEAK_GBPS = 8000.0
TRAFFIC_PROFILE = ROOT / "profiles" / "k_stage_statevector_30q_c128.md"


def measured_traffic(n: int, layers: int):
    """HBM bytes per k_stage launch from the committed FETCH_SIZE / WRITE_SIZE summary (calibrated
    against sweeps of known size), for the workload that summary covers; None otherwise."""
    if (n, layers) != (30, 40) or not TRAFFIC_PROFILE.exists():
        return None
    for line in TRAFFIC_PROFILE.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def gate_bytes(op, n: int) -> float:
    """Algorithmic bytes of one gate applied on its own (complex128)."""
    full = 32.0 * (1 << n)
    name = op.Gate.Name()
    if name in ("CZ",) or name.startswith("CP("):
        return full / 4
    if name in ("CNOT", "SWAP"):
        return full / 2
    return full


def cpu_baseline(circ, n: int, budget_s: float) -> dict:
    from oracle import oracle as O

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cores = min(cores, 16)
    O.set_workers(cores)
    from goqu_amd import _native as N

    state = O.zero_state(n)
    ops = [op for op in circ.Ops() if op.Gate is not None]
    done, nbytes = 0, 0.0
    t0 = time.perf_counter()
    for op in ops:
        buf = N.OpBuffer()
        N.add_sv_gate(buf, op.Gate, op.Qubits)
        O.sv_apply_ops(state, n, buf)
        done += 1
        nbytes += gate_bytes(op, n)
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    del state
    return {"value": round(nbytes / dt / 1e9, 3), "unit": "GB/s", "cores": cores, "kind": "port",
            "sample": f"first {done} of {len(ops)} gates of the same {n}-qubit circuit, one sweep per gate"}


def adjoint_section(qubits: int, reps: int, terms: int, ps_params: int, seed: int = 3) -> dict:
    """Adjoint gradients vs the reference's parameter-shift rule on the same GPU.

    Workload: EfficientSU2(qubits, reps, Circular) with theta ~ U[0, 2pi) and H = `terms` random
    Pauli strings of weight 1..4 with N(0,1) coefficients.  Parameter shift (2 Evolves +
    <H> per parameter, gradient.go:49-66) is timed on the first `ps_params` parameters with a
    reused simulator and extrapolated to all of them; both paths run on the GPU."""
    import math

    from goqu_amd import gradient as GR
    from goqu_amd import qaoa as Q
    from goqu_amd import statevector
    from goqu_amd.circuit import FreeParameters
    from goqu_amd.pauli import NewPauliString, NewPauliSum, X, Y, Z

    rng = np.random.default_rng(seed)
    c, npar = Q.EfficientSU2(qubits, reps)
    names = FreeParameters(c)
    hs = []
    for _ in range(terms):
        w = int(rng.integers(1, 5))
        ops = {int(q): [X, Y, Z][rng.integers(3)] for q in rng.choice(qubits, w, replace=False)}
        hs.append(NewPauliString(rng.normal(), ops, qubits))
    H = NewPauliSum(hs)
    x = rng.uniform(0, 2 * math.pi, npar)
    sim = statevector.New(qubits)
    adj = GR.Adjoint(c, H, names, sim=sim)
    adj.value_and_grad(x)  # warm-up
    t0 = time.perf_counter()
    e, g = adj.value_and_grad(x)
    t_adj = time.perf_counter() - t0
    f = GR.CostFunc(c, H, names, sim=sim)
    f(x)
    k = min(ps_params, npar)
    t0 = time.perf_counter()
    ps = np.zeros(k)
    for i in range(k):
        xp, xm = x.copy(), x.copy()
        xp[i] += 0.5 * math.pi
        xm[i] -= 0.5 * math.pi
        ps[i] = (f(xp) - f(xm)) / 2
    t_ps = (time.perf_counter() - t0) / k * npar
    sim.Close()
    return {"workload": f"EfficientSU2({qubits},{reps},Circular) {npar} params, H={terms} Pauli strings, c128",
            "adjoint_s": round(t_adj, 4), "adjoint_grads_per_s": round(1.0 / t_adj, 3),
            "param_shift_s": round(t_ps, 3), "param_shift_grads_per_s": round(1.0 / t_ps, 5),
            "param_shift_sample": f"first {k} of {npar} parameters, extrapolated",
            "speedup": round(t_ps / t_adj, 1), "max_abs_diff_vs_param_shift": float(np.max(np.abs(g[:k] - ps)))}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--layers", type=int, default=40)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adj-qubits", type=int, default=24)
    ap.add_argument("--ps-params", type=int, default=16)
    ap.add_argument("--calib", action="store_true",
                    help="also run one read-only sweep (norm) and one write-only sweep (reset) of the state, "
                         "whose HBM bytes are known, to scale PMC byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        return 2

    from circuits import config2_circuit
    from goqu_amd import _native as N
    from goqu_amd import statevector

    n = args.qubits
    circ = config2_circuit(n, args.layers, seed=2)
    buf = N.OpBuffer()
    for op in circ.Ops():
        if op.Gate is not None:
            N.add_sv_gate(buf, op.Gate, op.Qubits)
    arr = buf.array()
    ngates = len(buf)
    sim = statevector.New(n)
    lib = N.lib()

    def step():
        N.check(lib.goqu_sv_reset(sim._h))
        N.check(lib.goqu_sv_apply(sim._h, arr, ngates))

    for _ in range(args.warmup):
        step()
    N.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    dt = time.perf_counter() - t0
    kms, launches, kbytes = N.profile_read()
    N.profile(False)
    norm = sim.Norm2()
    if args.calib:  # k_prob0 reads 16 B x 2^n, k_fill_basis writes 16 B x 2^n
        sim.Norm2()
        sim.Reset()
    sim.Close()

    ms_step = dt / args.steps * 1e3
    sweeps_per_step = launches / args.steps
    sweep_bytes = 32.0 * (1 << n)
    value = sweeps_per_step * sweep_bytes / (ms_step / 1e3) / 1e9
    gate_equiv = sum(gate_bytes(op, n) for op in circ.Ops() if op.Gate is not None) / (ms_step / 1e3) / 1e9
    avg_launch_ms = kms / max(launches, 1)
    achieved = sweep_bytes / (avg_launch_ms / 1e3) / 1e9
    out = {
        "metric": "30q c128 gate-apply GB/s vs HBM peak; adjoint grads/s vs param-shift",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {"workload": f"statevector {n}q complex128 depth-{args.layers} U3 + CZ/CNOT/SWAP/CP random circuit",
                   "qubits": n, "layers": args.layers, "gates": ngates, "stage_sweeps": sweeps_per_step,
                   "gate_equiv_GBps": round(gate_equiv, 1), "norm2": norm},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": measured_traffic(n, args.layers),
                     "kernel": "k_stage<double>", "avg_launch_ms": round(avg_launch_ms, 4)},
        "cpu_baseline": None,
    }
    if not args.no_adjoint:
        out["adjoint"] = adjoint_section(args.adj_qubits, 7, 200, args.ps_params)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(circ, n, args.cpu_budget)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
