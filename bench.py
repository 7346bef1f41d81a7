"""Benchmark: 30-qubit complex128 gate application on one MI355X, plus the other workloads of the
path (adjoint gradients, sampling, batched QAOA, density matrix, wide MFMA ops, dynamic circuits,
circuit batches).

Workload (one "step"): Evolve of the depth-40 seeded random circuit on 30 qubits —
U3(theta, phi, lambda ~ U[0, 2pi)) on every qubit per layer, then 15 disjoint random pairs
(a quarter of them spanning a low and a high qubit) each carrying CZ, CNOT, SWAP or CP.
The state (16 GiB) stays resident in HBM; the timed region is the K Evolves.  The tile shape is
the autotuner's choice for this workload (timed before the warm-up, outside the timed region).

Printed JSON line:
  value     = HBM sweep throughput of the whole step: stage sweeps x 32 B x 2^30 / step time
  roofline  = the stage kernel (k_stage): algorithmic bytes per launch (read + write of the
              whole state, 32 B x 2^30) / its mean launch duration from HIP events recorded on
              its stream, against the 8 TB/s HBM peak; traffic = HBM bytes per launch from the
              committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary under profiles/
  cpu_baseline = the oracle (a C restatement of goqu's CPU statevector with its worker policy)
              applying the first gates of the same circuit at 30 qubits on the host cores (the
              reference itself cannot run 30 qubits), and the same at 28 qubits beside the GPU's
              28-qubit time; bytes per gate as for the GPU metric (32 B x 2^n, a quarter for CZ/CP,
              half for CNOT/SWAP), divided by its wall time.
  each further section carries its own roofline object(s) from the same HIP-event profiler.
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


def measured_traffic(n: int, layers: int, tile_bits: int):
    """HBM bytes per k_stage launch from the committed FETCH_SIZE / WRITE_SIZE summary (calibrated
    against sweeps of known size), for the workload and tile shape that summary covers; None
    otherwise.  The 2^12-tile figure is the plain line, other tile sizes carry [tile_bits]."""
    if (n, layers) != (30, 40) or not TRAFFIC_PROFILE.exists():
        return None
    key = "traffic_bytes_per_launch:" if tile_bits == 12 else f"traffic_bytes_per_launch[{tile_bits}]:"
    for line in TRAFFIC_PROFILE.read_text().splitlines():
        if line.startswith(key):
            return int(line.split(":")[1])
    return None


def section_traffic(profile: str, kernel: str):
    """HBM bytes per launch of `kernel` from a committed profile summary under profiles/ (the
    calibrated FETCH_SIZE / WRITE_SIZE measurement of the same section workload), else None."""
    f = ROOT / "profiles" / profile
    key = f"traffic_bytes_per_launch[{kernel}]:"
    if f.exists():
        for line in f.read_text().splitlines():
            if line.startswith(key):
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


def _cpu_info() -> dict:
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    # the host cores this process may use: its affinity set, within the OpenMP thread budget the
    # machine gives one GPU's job when it sets one
    omp = os.environ.get("OMP_NUM_THREADS")
    cores = min(aff, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else aff
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cores": aff, "omp_num_threads": omp,
            "cores": cores}


def _cpu_gate_rate(circ, n: int, budget_s: float, cores: int) -> tuple:
    from oracle import oracle as O
    from goqu_amd import _native as N

    O.set_workers(cores)
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
    return nbytes / dt / 1e9, done, len(ops), dt


def cpu_baseline(circ, n: int, budget_s: float, layers: int) -> dict:
    from circuits import config2_circuit
    from oracle import oracle as O

    info = _cpu_info()
    cores = info["cores"]
    gbps, done, total, _ = _cpu_gate_rate(circ, n, budget_s, cores)
    out = {"value": round(gbps, 3), "unit": "GB/s", "cores": cores, "kind": "port",
           "sample": f"first {done} of {total} gates of the same {n}-qubit circuit, one sweep per gate",
           "reference_runs": n <= 28,
           "note": ("the reference statevector cannot run this size: statevector.New panics above 28 qubits "
                    "(sim/statevector/sim.go:28-30); timed on the oracle's restatement of its kernels")
           if n > 28 else "", **info,
           "host_stream_copy_GBps": round(O.stream_copy_gbps(1 << 27, cores, 3), 1)}
    # the largest size the reference runs: CPU beside the GPU on the same 28-qubit circuit
    c28 = config2_circuit(28, layers, seed=2)
    g28, done28, total28, _ = _cpu_gate_rate(c28, 28, budget_s / 2, cores)
    out["q28"] = {"cpu_GBps": round(g28, 3), "cpu_sample": f"first {done28} of {total28} gates",
                  "gpu_GBps": round(_gpu_rate(c28, 28), 1), "unit": "GB/s (gate-sweep bytes / time)"}
    return out


def _gpu_rate(circ, n: int) -> float:
    """Gate-sweep bytes of circ / GPU Evolve time (for the CPU comparison at the same size)."""
    from goqu_amd import _native as N
    from goqu_amd import statevector

    sim = statevector.New(n)
    buf = N.OpBuffer()
    for op in circ.Ops():
        if op.Gate is not None:
            N.add_sv_gate(buf, op.Gate, op.Qubits)
    arr = buf.array()
    N.check(N.lib().goqu_sv_apply(sim._h, arr, len(buf)))
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        N.check(N.lib().goqu_sv_reset(sim._h))
        N.check(N.lib().goqu_sv_apply(sim._h, arr, len(buf)))
    dt = (time.perf_counter() - t0) / reps
    sim.Close()
    return sum(gate_bytes(op, n) for op in circ.Ops() if op.Gate is not None) / dt / 1e9


def _roof(kernel: str, bin_id: int, bound: str = "hbm", profile: str = None):
    """Roofline object of one profiled kernel over everything recorded since the last reset;
    `traffic` from the committed profile summary of the section (`profile`), when it has one."""
    from goqu_amd import _native as N

    ms, launches, amount = N.profile_read_kernel(bin_id)
    if not launches or ms <= 0:
        return None
    if bound == "mfma":
        achieved = amount / (ms / 1e3) / 1e12
        peak, unit = 78.6, "TFLOP/s"  # MI355X dense FP64 MFMA
    else:
        achieved = amount / (ms / 1e3) / 1e9
        peak, unit = HBM_PEAK_GBPS, "GB/s"
    return {"kernel": kernel, "bound": bound, "achieved": round(achieved, 1), "peak": peak, "unit": unit,
            "frac": round(achieved / peak, 4), "launches": launches, "avg_launch_ms": round(ms / launches, 4),
            "traffic": section_traffic(profile, kernel) if profile else None}


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
    from goqu_amd import _native as N

    sim = statevector.New(qubits)
    adj = GR.Adjoint(c, H, names, sim=sim)
    adj.value_and_grad(x)  # warm-up
    N.profile(True)
    t0 = time.perf_counter()
    e, g = adj.value_and_grad(x)
    t_adj = time.perf_counter() - t0
    prof = "adjoint_efficientsu2_24q_c128.md" if (qubits, reps, terms) == (24, 7, 200) else None
    roofs = [r for r in (_roof("k_stage_pair<double>", N.PROF_STAGE_PAIR, profile=prof),
                         _roof("k_pauli_apply_tile<double>", N.PROF_HAMILTONIAN, profile=prof),
                         _roof("k_stage<double>", N.PROF_STAGE, profile=prof)) if r]
    N.profile(False)
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
            "speedup": round(t_ps / t_adj, 1), "max_abs_diff_vs_param_shift": float(np.max(np.abs(g[:k] - ps))),
            "roofline": roofs}


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
    Device bytes count the amplitudes the tile kernels move for instances above 12 qubits:
    (96 p + 32) B x 2^n (p + 1 forward passes over psi, the first writing only; p + 1 reverse
    passes over psi and lambda, the first reading psi alone, the last writing nothing); smaller
    instances stay in LDS.  The timed call follows
    one untimed call of the same batch.  `s` is the whole Python call; large_instance_GBps divides the device bytes by the goqu_qaoa_batch call alone (inputs
    marshalled into flat arrays first, as a Go caller passes them), which includes the upload, the
    graph and the download."""
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
    from goqu_amd import _native as N

    # warm-up: one call of the same batch (module load, graph machinery, and the device pool the
    # library retains between calls); the timed call is the steady state of repeated evaluations
    Q.BatchMaxCut(insts)
    N.profile(True)
    t0 = time.perf_counter()
    m = Q.MarshalledBatch(insts)
    t1 = time.perf_counter()
    energies, _ = Q.RunMarshalled(m)
    t2 = time.perf_counter()
    dt, dt_abi = t2 - t0, t2 - t1
    gms, _, _ = N.profile_read_kernel(N.PROF_QAOA)
    N.profile(False)
    big = [it for it in insts if it.num_qubits > 12]
    dev_bytes = sum((96.0 * it.layers + 32.0) * (1 << it.num_qubits) for it in big)
    roof = None
    if gms > 0:
        a = dev_bytes / (gms / 1e3) / 1e9
        roof = {"kernel": "batched QAOA launch graph (k_qaoa_small, k_qaoa_pass, k_qaoa_finish)", "bound": "hbm",
                "achieved": round(a, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(a / HBM_PEAK_GBPS, 4),
                "graph_ms": round(gms, 3), "traffic": None}
    return {"workload": f"{num} MaxCut instances, n~U{{10..22}}, G(n,0.3), p~U{{1..4}}, energy + 2p gradients, c128",
            "s": round(dt, 4), "instances_per_s": round(num / dt, 1),
            "abi_call_s": round(dt_abi, 4), "python_marshal_s": round(dt - dt_abi, 4),
            "large_instance_GBps": round(dev_bytes / dt_abi / 1e9, 1),
            "mean_energy": float(np.mean(energies)), "roofline": roof}


def dm_section(n: int = 14, layers: int = 20, seed: int = 9) -> dict:
    """14-qubit density matrix (4 GiB): per layer U3 on every qubit and 7 CNOTs on random disjoint
    pairs, Depolarizing1Q(1e-3) / Depolarizing2Q(1e-2) default errors and an AmplitudeDamping(1e-3)
    Delay per qubit at each layer end.  Forward: Evolve + <sum Z_i>, value = stage sweeps x
    32 B x 4^n / time.  Gradient: <sum Z_i> and all 3 x 14 x layers U3 angles by the
    density-matrix adjoint (a forward pass, then paired reverse sweeps over rho and h = vec(H)
    that uncompute rho through the inverse channels); its roofline objects cover the paired
    kernel (read + write of both states per sweep) and the forward sweeps."""
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
    prof = "k_stage_dm_densitymatrix_14q.md" if (n, layers) == (14, 20) else None
    roof = _roof("k_stage_dm<double>", N.PROF_STAGE_DM, profile=prof)
    N.profile(False)
    dm.Close()
    sweep = 32.0 * (1 << (2 * n))
    out = {"workload": f"{n}q density matrix, {layers} layers U3 + 7 CNOT + depolarizing/amplitude-damping noise, c128",
           "forward_ms": round(t_fwd * 1e3, 2), "stage_sweeps": launches,
           "forward_GBps": round(launches * sweep / t_fwd / 1e9, 1),
           "k_stage_avg_ms": round(kms / max(launches, 1), 3), "energy": e, "roofline": roof}
    g = GR.AdjointDensity(c, H, names, noise=nm, checkpoint_every=16)
    N.profile_read()
    N.profile(True)
    t0 = time.perf_counter()
    e2, grad = g.value_and_grad(x)
    out["gradient_s"] = round(time.perf_counter() - t0, 3)
    out["gradient_roofline"] = [r for r in (_roof("k_stage_pair_dm<double>", N.PROF_STAGE_PAIR, profile=prof),
                                            _roof("k_stage_dm<double>", N.PROF_STAGE_DM, profile=prof)) if r]
    N.profile(False)
    out["gradient_params"] = len(names)
    out["gradient_energy_matches_forward"] = bool(abs(e2 - e) < 1e-9)
    return out


def wide_section(n: int = 26, targets=(5, 7, 9), seed: int = 12) -> dict:
    """Singly controlled dense gates on 5, 7 and 9 target qubits of a 26-qubit complex128 state
    (1 GiB): FP64 MFMA contractions (k_wide_mfma), 2^k complex multiply-adds per amplitude of the
    controlled half."""
    from scipy.stats import unitary_group

    from goqu_amd import _native as N
    from goqu_amd import gate as G
    from goqu_amd import statevector
    from goqu_amd.circuit import NewBuilder

    rng = np.random.default_rng(seed)
    out = {"workload": f"{n}q c128 state, one dense gate on each of {list(targets)} target qubits", "ops": []}
    sim = statevector.New(n)
    for k in targets:
        U = unitary_group.rvs(1 << k, random_state=int(rng.integers(1 << 30)))
        qs = [int(q) for q in rng.choice(n, k + 1, replace=False)]
        b = NewBuilder("wide", n)
        b.Apply(G.Controlled(G.Unitary("U", U.reshape(-1)), 1), *qs)
        c = b.Build()
        sim.Apply(c)  # warm-up
        N.profile(True)
        t0 = time.perf_counter()
        for _ in range(3):
            sim.Apply(c)
        dt = (time.perf_counter() - t0) / 3
        roof = _roof(f"k_wide_mfma<double> k={k}", N.PROF_WIDE, bound="mfma")
        N.profile(False)
        out["ops"].append({"targets": k, "ms": round(dt * 1e3, 3), "roofline": roof})
    sim.Close()
    return out


def dynamic_section(n: int = 12, shots: int = 20000, seed: int = 13) -> dict:
    """Dynamic circuits with the shots batched on the device: a 12-qubit circuit with mid-circuit
    measurements, resets, if/else feed-forward and a repeat-until-success while loop, 20,000 shots."""
    from circuits import random_dynamic_circuit
    from goqu_amd import statevector

    c = random_dynamic_circuit(n, 4, 6, np.random.default_rng(seed))
    sim = statevector.New(n)
    sim.RunDynamicShots(c, 256, rng=np.random.default_rng(1))  # warm-up
    t0 = time.perf_counter()
    sim.RunDynamicShots(c, shots, rng=np.random.default_rng(2))
    dt = time.perf_counter() - t0
    sim.Close()
    return {"workload": f"{n}q dynamic circuit ({c.NumOps()} top-level ops: measure, reset, if/else, switch, for, "
                        f"while), {shots} shots, c128", "s": round(dt, 4), "shots_per_s": round(shots / dt, 1)}


def batch_section(n: int = 12, reps: int = 4, points: int = 2048, terms: int = 50, seed: int = 14) -> dict:
    """VQE-style parameter batch: EfficientSU2(12, 4) at 2,048 parameter sets, H = 50 Pauli strings,
    one device batch (op records built per gate for all points at once); against binding and
    evaluating one point at a time on one simulator (timed on 64 points, extrapolated)."""
    from circuits import random_pauli_sum
    from goqu_amd import batch as B
    from goqu_amd import qaoa as Q
    from goqu_amd import statevector

    from goqu_amd.circuit import Bind, FreeParameters

    rng = np.random.default_rng(seed)
    c, npar = Q.EfficientSU2(n, reps)
    H = random_pauli_sum(n, terms, rng)
    params = rng.uniform(0, 2 * np.pi, (points, npar))
    B.vqe_energies(c, H, params[:32])  # warm-up
    t0 = time.perf_counter()
    e = B.vqe_energies(c, H, params)
    dt = time.perf_counter() - t0
    # one simulator at a time, binding each point as sweep's runStatevectorPoint does (run.go:75-97)
    names = FreeParameters(c)
    sim = statevector.New(n)
    k = 64
    t0 = time.perf_counter()
    seq = []
    for row in params[:k]:
        sim.Evolve(Bind(c, dict(zip(names, row))))
        seq.append(sim.ExpectPauliSum(H))
    t_seq = (time.perf_counter() - t0) / k * points
    sim.Close()
    return {"workload": f"EfficientSU2({n},{reps}) at {points} parameter sets, H={terms} Pauli strings, c128",
            "s": round(dt, 4), "circuits_per_s": round(points / dt, 1),
            "one_at_a_time_s": round(t_seq, 3), "one_at_a_time_sample": f"first {k} points, extrapolated",
            "speedup": round(t_seq / dt, 1), "max_abs_diff": float(np.max(np.abs(e[:k] - np.array(seq))))}


SECTIONS = ("adjoint", "sampling", "qaoa", "dm", "wide", "dynamic", "batch")
TUNE_CANDIDATES = [(12, 3), (12, 4), (12, 5), (13, 3), (13, 4), (11, 4)]


def _log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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
    ap.add_argument("--no-extra", action="store_true", help="only the headline and the adjoint section")
    ap.add_argument("--sections", default=",".join(SECTIONS),
                    help=f"comma-separated subset of {','.join(SECTIONS)} (or 'none')")
    ap.add_argument("--no-tune", action="store_true", help="use the library's default tile shape")
    ap.add_argument("--calib", action="store_true",
                    help="also run one read-only sweep (norm) and one write-only sweep (reset) of the state, "
                         "whose HBM bytes are known, to scale PMC byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        return 2
    sections = set() if args.sections == "none" else set(args.sections.split(","))
    if args.no_extra:
        sections &= {"adjoint"}
    if args.no_adjoint:
        sections.discard("adjoint")

    from circuits import config2_circuit
    from goqu_amd import _native as N
    from goqu_amd import autotune
    from goqu_amd import statevector

    n = args.qubits
    circ = config2_circuit(n, args.layers, seed=2)
    plan, tune_times = (0, 0), {}
    if not args.no_tune:
        _log("autotuning the tile shape")
        plan, tune_times = autotune.tune(circ, workload=f"config2-depth{args.layers}", candidates=TUNE_CANDIDATES,
                                         reps=1)
    buf = N.OpBuffer()
    for op in circ.Ops():
        if op.Gate is not None:
            N.add_sv_gate(buf, op.Gate, op.Qubits)
    arr = buf.array()
    ngates = len(buf)
    sim = statevector.New(n, plan=plan)
    lib = N.lib()

    def step():
        N.check(lib.goqu_sv_reset(sim._h))
        N.check(lib.goqu_sv_apply(sim._h, arr, ngates))

    _log("headline")
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
    tile_bits = plan[0] or int(os.environ.get("GOQU_TILE_BITS", "12"))  # (0, 0): the library default
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
                   "gate_equiv_GBps": round(gate_equiv, 1), "norm2": norm,
                   "tile_shape": {"tile_bits": plan[0], "low_bits": plan[1], "autotuned": not args.no_tune,
                                  "candidates_s": {k: round(v, 4) for k, v in tune_times.items()}}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": measured_traffic(n, args.layers, tile_bits),
                     "kernel": "k_stage<double>", "avg_launch_ms": round(avg_launch_ms, 4)},
        "cpu_baseline": None,
    }
    runners = [
        ("adjoint", lambda: adjoint_section(args.adj_qubits, 7, 200, args.ps_params)),
        ("sampling", sampling_section),
        ("qaoa", qaoa_batch_section),
        ("dm", dm_section),
        ("wide", wide_section),
        ("dynamic", dynamic_section),
        ("batch", batch_section),
    ]
    keys = {"qaoa": "qaoa_batch", "dm": "density_matrix", "batch": "circuit_batch", "dynamic": "dynamic_circuits",
            "wide": "wide_mfma"}
    for name, fn in runners:
        if name in sections:
            _log(name)
            out[keys.get(name, name)] = fn()
    if not args.no_cpu:
        _log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(circ, n, args.cpu_budget, args.layers)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
