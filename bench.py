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


def sampling_section(n: int = 20, layers: int = 20, shots: int = 1024) -> dict:
    """20-qubit layered circuit (RX/RY/RZ per qubit + CNOT brickwork, 1,580 gates), MeasureAll,
    1,024 shots: Evolve + device CDF + sampling, absolute time (the 16 MiB state is cache-resident,
    so no HBM roofline applies)."""
    from circuits import config1_circuit
    from goqu_amd import statevector

    c = config1_circuit(n, layers)
    sim = statevector.New(n)
    rng = np.random.default_rng(11)
    sim.Run(c, shots, rng=rng)  # warm-up
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        counts = sim.Run(c, shots, rng=rng)
    dt = (time.perf_counter() - t0) / reps
    sim.Close()
    return {"workload": f"{n}q {layers}-layer RX/RY/RZ + CNOT brickwork, MeasureAll, {shots} shots, c128",
            "ms_per_run": round(dt * 1e3, 3), "distinct_outcomes": len(counts)}


def qaoa_batch_section(num: int = 4096, seed: int = 5) -> dict:
    """4,096 MaxCut instances, n ~ U{10..22}, G(n, 0.3), p ~ U{1..4}, gamma/beta ~ U[0, pi):
    energy and all 2p gradient entries of every instance in one captured launch sequence.
    Device bytes count what the tile kernels move for instances above 12 qubits (192 B x 2^n per
    layer: two forward sweeps of psi, two reverse sweeps of psi and lambda); smaller instances
    stay in LDS."""
    import math

    from goqu_amd import qaoa as Q

    rng = np.random.default_rng(seed)
    insts = []
    for _ in range(num):
        n = int(rng.integers(10, 23))
        while True:
            edges = [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < 0.3]
            if edges:
                break
        p = int(rng.integers(1, 5))
        insts.append(Q.MaxCutInstance(n, edges, p, rng.uniform(0, math.pi, 2 * p)))
    Q.BatchMaxCut(insts[:64])  # warm-up (module load, graph machinery)
    t0 = time.perf_counter()
    energies, grads = Q.BatchMaxCut(insts)
    dt = time.perf_counter() - t0
    big = [it for it in insts if it.num_qubits > 12]
    dev_bytes = sum(192.0 * (1 << it.num_qubits) * it.layers for it in big)
    return {"workload": f"{num} MaxCut instances, n~U{{10..22}}, G(n,0.3), p~U{{1..4}}, energy + 2p gradients, c128",
            "s": round(dt, 4), "instances_per_s": round(num / dt, 1),
            "large_instance_GBps": round(dev_bytes / dt / 1e9, 1),
            "mean_energy": float(np.mean(energies))}


def dm_section(n: int = 14, layers: int = 20, seed: int = 9) -> dict:
    """14-qubit density matrix (4 GiB): per layer U3 on every qubit and 7 CNOTs on random disjoint
    pairs, Depolarizing1Q(1e-3) / Depolarizing2Q(1e-2) default errors and an AmplitudeDamping(1e-3)
    Delay per qubit at each layer end.  Forward: Evolve + <sum Z_i>, value = stage sweeps x
    32 B x 4^n / time.  Gradient: <sum Z_i> and all 3 x 14 x layers U3 angles by the
    checkpointed density-matrix adjoint."""
    from goqu_amd import densitymatrix
    from goqu_amd import gradient as GR
    from goqu_amd import noise as NZ
    from goqu_amd import _native as N
    from goqu_amd.circuit import Bind, FreeParameters, NewBuilder, NewVector
    from goqu_amd.pauli import NewPauliString, NewPauliSum, Z

    rng = np.random.default_rng(seed)
    th = NewVector("t", 3 * n * layers)
    b = NewBuilder("dm", n)
    k = 0
    for _ in range(layers):
        for q in range(n):
            b.SymU3(th.At(k), th.At(k + 1), th.At(k + 2), q)
            k += 3
        perm = rng.permutation(n)
        for i in range(0, n - 1, 2):
            b.CNOT(int(perm[i]), int(perm[i + 1]))
        for q in range(n):
            b.Delay(q, 1.0)
    c = b.Build()
    nm = NZ.New()
    nm.AddDefaultError(1, NZ.Depolarizing1Q(1e-3))
    nm.AddDefaultError(2, NZ.Depolarizing2Q(1e-2))
    nm.AddGateError("delay", NZ.AmplitudeDamping(1e-3))
    H = NewPauliSum([NewPauliString(1.0, {q: Z}, n) for q in range(n)])
    names = FreeParameters(c)
    x = rng.uniform(0, 2 * np.pi, len(names))
    bound = Bind(c, dict(zip(names, x)))
    dm = densitymatrix.New(n).WithNoise(nm)
    dm.Evolve(bound)  # warm-up
    N.profile_read()
    N.profile(True)
    t0 = time.perf_counter()
    dm.Evolve(bound)
    e = dm.ExpectPauliSum(H)
    t_fwd = time.perf_counter() - t0
    kms, launches, _ = N.profile_read()
    N.profile(False)
    dm.Close()
    sweep = 32.0 * (1 << (2 * n))
    out = {"workload": f"{n}q density matrix, {layers} layers U3 + 7 CNOT + depolarizing/amplitude-damping noise, c128",
           "forward_ms": round(t_fwd * 1e3, 2), "stage_sweeps": launches,
           "forward_GBps": round(launches * sweep / t_fwd / 1e9, 1),
           "k_stage_avg_ms": round(kms / max(launches, 1), 3), "energy": e}
    g = GR.AdjointDensity(c, H, names, noise=nm, checkpoint_every=16)
    t0 = time.perf_counter()
    e2, grad = g.value_and_grad(x)
    out["gradient_s"] = round(time.perf_counter() - t0, 3)
    out["gradient_params"] = len(names)
    out["gradient_energy_matches_forward"] = bool(abs(e2 - e) < 1e-9)
    return out


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
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the 20q sampling, batched QAOA and density-matrix sections")
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
    if not args.no_extra:
        out["sampling"] = sampling_section()
        out["qaoa_batch"] = qaoa_batch_section()
        out["density_matrix"] = dm_section()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(circ, n, args.cpu_budget)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
