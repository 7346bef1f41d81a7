"""Seeded circuit generators for the parity tests and the benchmark workloads."""
from __future__ import annotations

import math

import numpy as np

from goqu_amd import gate as G
from goqu_amd.circuit import NewBuilder, NewVector

TWO_PI = 2 * math.pi

ONE_Q_FIXED = [G.I, G.H, G.X, G.Y, G.Z, G.S, G.Sdg, G.T, G.Tdg, G.SX]
TWO_Q_FIXED = [G.CNOT, G.CZ, G.SWAP, G.CY, G.ISWAP, G.ECR, G.DCX, G.CH, G.CSX, G.Sycamore]
THREE_Q_FIXED = [G.CCX, G.CSWAP, G.CCZ]


def rand_1q(rng):
    k = rng.integers(0, 12)
    a = rng.uniform(0, TWO_PI, 3)
    return [lambda: ONE_Q_FIXED[rng.integers(len(ONE_Q_FIXED))], lambda: G.RX(a[0]), lambda: G.RY(a[0]),
            lambda: G.RZ(a[0]), lambda: G.Phase(a[0]), lambda: G.U3(*a), lambda: G.U1(a[0]),
            lambda: G.U2(a[0], a[1]), lambda: G.Rot(*a), lambda: G.GPI(a[0]), lambda: G.GPI2(a[0]),
            lambda: G.PhasedXZ(*a)][k]()


def rand_2q(rng):
    k = rng.integers(0, 13)
    a = rng.uniform(0, TWO_PI, 2)
    return [lambda: TWO_Q_FIXED[rng.integers(len(TWO_Q_FIXED))], lambda: G.CP(a[0]), lambda: G.CRZ(a[0]),
            lambda: G.CRX(a[0]), lambda: G.CRY(a[0]), lambda: G.RXX(a[0]), lambda: G.RYY(a[0]),
            lambda: G.RZZ(a[0]), lambda: G.ZZ(a[0]), lambda: G.MS(a[0], a[1]), lambda: G.FSim(a[0], a[1]),
            lambda: G.PSwap(a[0]), lambda: G.Controlled(G.H, 1)][k]()


def random_circuit(n: int, depth: int, rng, three_q: bool = True, controlled: bool = True,
                   resets: float = 0.0):
    """Every layer: a random 1q gate per qubit and random multi-qubit gates on random qubits.
    resets > 0: each qubit is also reset (the Evolve path's deterministic reset) with that
    probability per layer, between its 1q gate and the multi-qubit gates."""
    b = NewBuilder("random", n)
    for _ in range(depth):
        for q in range(n):
            b.Apply(rand_1q(rng), q)
            if resets and rng.random() < resets:
                b.Reset(q)
        for _ in range(max(1, n // 2)):
            if n >= 2:
                qs = rng.choice(n, 2, replace=False)
                b.Apply(rand_2q(rng), int(qs[0]), int(qs[1]))
            if three_q and n >= 3 and rng.random() < 0.3:
                qs = rng.choice(n, 3, replace=False)
                if rng.random() < 0.5:
                    b.Apply(THREE_Q_FIXED[rng.integers(3)], *map(int, qs))
                else:
                    b.Apply(G.Controlled(G.RY(rng.uniform(0, TWO_PI)), 2), *map(int, qs))
            if controlled and n >= 4 and rng.random() < 0.2:
                nc = int(rng.integers(3, min(n, 6)))
                qs = [int(x) for x in rng.choice(n, nc + 1, replace=False)]
                inner = [G.X, G.Z, G.H, G.RX(rng.uniform(0, TWO_PI)), G.Phase(rng.uniform(0, TWO_PI))][rng.integers(5)]
                b.Apply(G.Controlled(inner, nc), *qs)
    return b.Build()


def config1_circuit(n: int = 20, layers: int = 20, seed: int = 1):
    """20 layers of RX/RY/RZ per qubit (angles U[0,2pi)) then CNOT brickwork (even, then odd
    pairs), MeasureAll."""
    rng = np.random.default_rng(seed)
    b = NewBuilder("config1", n)
    for _ in range(layers):
        for q in range(n):
            a = rng.uniform(0, TWO_PI, 3)
            b.RX(a[0], q).RY(a[1], q).RZ(a[2], q)
        for q in range(0, n - 1, 2):
            b.CNOT(q, q + 1)
        for q in range(1, n - 1, 2):
            b.CNOT(q, q + 1)
    b.MeasureAll()
    return b.Build()


def config2_circuit(n: int = 30, layers: int = 40, seed: int = 2):
    """Depth-40 random circuit: U3 (angles U[0,2pi)) on every qubit, then floor(n/2) disjoint
    random pairs, at least a quarter of them spanning a low (<10) and a high (>=n-10) qubit;
    each pair gets CZ, CNOT, SWAP or CP(U[0,2pi))."""
    rng = np.random.default_rng(seed)
    b = NewBuilder("config2", n)
    lo_n = min(10, n // 2)
    for _ in range(layers):
        for q in range(n):
            b.U3(*rng.uniform(0, TWO_PI, 3), q)
        free = list(range(n))
        pairs = []
        n_span = max(1, (n // 2) // 4)
        for _ in range(n_span):
            lows = [q for q in free if q < lo_n]
            highs = [q for q in free if q >= n - lo_n]
            if not lows or not highs:
                break
            a, c = int(rng.choice(lows)), int(rng.choice(highs))
            free.remove(a)
            free.remove(c)
            pairs.append((a, c))
        rng.shuffle(free)
        while len(free) >= 2:
            pairs.append((free.pop(), free.pop()))
        for a, c in pairs:
            if rng.random() < 0.5:
                a, c = c, a
            k = rng.integers(0, 4)
            if k == 0:
                b.CZ(a, c)
            elif k == 1:
                b.CNOT(a, c)
            elif k == 2:
                b.SWAP(a, c)
            else:
                b.CP(rng.uniform(0, TWO_PI), a, c)
    return b.Build()


def efficient_su2(n: int, reps: int):
    """algorithm/ansatz/efficientsu2.go:19-50 with Circular entanglement
    (realamplitudes.go:51-70): reps+1 layers of SymRY(theta[2j]) SymRZ(theta[2j+1]) per qubit,
    CNOT(i,i+1)...CNOT(n-1,0) between layers."""
    nparams = 2 * n * (reps + 1)
    th = NewVector("θ", nparams)
    b = NewBuilder("efficientsu2", n)
    p = 0
    for layer in range(reps + 1):
        for q in range(n):
            b.SymRY(th.At(p), q)
            b.SymRZ(th.At(p + 1), q)
            p += 2
        if layer < reps:
            for q in range(n - 1):
                b.CNOT(q, q + 1)
            if n > 1:
                b.CNOT(n - 1, 0)
    return b.Build(), nparams


def interleaved_product_circuit(n: int, layers: int, seed: int):
    """Two independent config2_circuit halves on interleaved qubits: half A on the even qubits
    (local q -> 2q), half B on the odd ones (q -> 2q + 1), ops alternating between the halves.
    The gates span low and high qubits of the whole register, yet the final state is exactly
    psi_A (x) psi_B, so every amplitude and every Z-string expectation follows from two small
    oracle runs.  Returns (circuit, circuit_A, circuit_B)."""
    na, nb = (n + 1) // 2, n // 2
    ca, cb = config2_circuit(na, layers, seed), config2_circuit(nb, layers, seed + 1)
    b = NewBuilder("c2product", n)
    oa, ob = ca.Ops(), cb.Ops()
    for i in range(max(len(oa), len(ob))):
        if i < len(oa):
            b.Apply(oa[i].Gate, *[2 * q for q in oa[i].Qubits])
        if i < len(ob):
            b.Apply(ob[i].Gate, *[2 * q + 1 for q in ob[i].Qubits])
    return b.Build(), ca, cb


def split_bits(idx: np.ndarray, parity: int, nbits: int) -> np.ndarray:
    """The bits of idx at positions 2j + parity, packed (j = 0..nbits-1)."""
    out = np.zeros_like(idx)
    for j in range(nbits):
        out |= ((idx >> np.uint64(2 * j + parity)) & np.uint64(1)) << np.uint64(j)
    return out


def random_z_masks(n: int, count: int, rng) -> list:
    """`count` random non-empty Z masks over n qubits."""
    out = []
    while len(out) < count:
        m = int(rng.integers(1, 1 << min(n, 62))) if n <= 62 else int(rng.integers(1, 1 << 62))
        m &= (1 << n) - 1
        if m:
            out.append(m)
    return out


def qaoa_instances(num: int, seed: int, nmin: int = 10, nmax: int = 22):
    """MaxCut instances: n ~ U{nmin..nmax}, G(n, 0.3) unit-weight edges (resampled if empty),
    p ~ U{1..4}, gamma, beta ~ U[0, pi), params in FreeParameters order (g0, b0, g1, ...)."""
    from goqu_amd import qaoa as Q

    rng = np.random.default_rng(seed)
    insts = []
    for _ in range(num):
        n = int(rng.integers(nmin, nmax + 1))
        while True:
            edges = [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < 0.3]
            if edges:
                break
        p = int(rng.integers(1, 5))
        insts.append(Q.MaxCutInstance(n, edges, p, rng.uniform(0, math.pi, 2 * p)))
    return insts


def noisy_dm_workload(n: int = 14, layers: int = 20, seed: int = 9):
    """Per layer U3 (symbolic angles) on every qubit and n/2 CNOTs on random disjoint pairs, then
    a Delay on every qubit; noise: Depolarizing1Q(1e-3) / Depolarizing2Q(1e-2) defaults and
    AmplitudeDamping(1e-3) on "delay"; H = sum_i Z_i.  Returns (circuit, noise model, H, names,
    angles ~ U[0, 2pi))."""
    from goqu_amd import noise as NZ
    from goqu_amd.circuit import FreeParameters
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
    return c, nm, H, names, x


def random_pauli_sum(n: int, terms: int, rng, wmax: int = 4):
    """`terms` random Pauli strings, weight ~ U{1..wmax} on distinct qubits, letters ~ U{X,Y,Z},
    coefficients ~ N(0, 1)."""
    from goqu_amd.pauli import NewPauliString, NewPauliSum, X, Y, Z

    out = []
    for _ in range(terms):
        w = int(rng.integers(1, min(wmax, n) + 1))
        ops = {int(q): [X, Y, Z][rng.integers(3)] for q in rng.choice(n, w, replace=False)}
        out.append(NewPauliString(rng.normal(), ops, n))
    return NewPauliSum(out)


def random_dynamic_circuit(n: int, nclbits: int, depth: int, rng, density: bool = False):
    """Mid-circuit measurements, resets, conditioned gates and every control-flow form (if/else,
    switch with and without a default, for, while loops that end with probability 1/2 per
    iteration), nested up to two deep, between layers of random gates.  density=True keeps to
    1- and 2-qubit gates and adds delays (noise hooks)."""
    b = NewBuilder("dynamic", n).WithClbits(nclbits)

    def gates(bb, k):
        for _ in range(k):
            if n >= 2 and rng.random() < 0.4:
                qs = rng.choice(n, 2, replace=False)
                bb.Apply(rand_2q(rng), int(qs[0]), int(qs[1]))
            else:
                bb.Apply(rand_1q(rng), int(rng.integers(n)))
            if density and rng.random() < 0.2:
                bb.Delay(int(rng.integers(n)), 1.0)

    def measure(bb):
        bb.Measure(int(rng.integers(n)), int(rng.integers(nclbits)))

    def block(bb, level):
        k = int(rng.integers(9))
        c0 = int(rng.integers(nclbits))
        if k == 0:
            measure(bb)
        elif k == 1:
            bb.Reset(int(rng.integers(n)))
        elif k == 2:
            bb.If(c0, int(rng.integers(2)), rand_1q(rng), int(rng.integers(n)))
        elif k == 3 and level < 2:
            bb.IfElseBlock(c0, int(rng.integers(2)), lambda x: (gates(x, 2), block(x, level + 1)),
                           (lambda x: (gates(x, 1), measure(x))) if rng.random() < 0.7 else None)
        elif k == 4 and level < 2 and nclbits >= 2:
            bits = [int(x) for x in rng.choice(nclbits, 2, replace=False)]
            vals = [int(v) for v in rng.choice(4, int(rng.integers(1, 4)), replace=False)]
            cases = {v: (lambda x, v=v: (gates(x, 1 + v % 2), block(x, level + 1))) for v in vals}
            bb.Switch(bits, cases, (lambda x: gates(x, 2)) if rng.random() < 0.5 else None)
        elif k == 5 and level < 2:
            bb.For(c0, 0, int(rng.integers(1, 4)), 1, lambda x: (gates(x, 1), block(x, level + 1)))
        elif k == 6 and level < 2:
            q = int(rng.integers(n))
            bb.Measure(q, c0)  # loop while the clbit reads 1: H then measure again
            bb.While(c0, 1, lambda x: (x.H(q), x.Measure(q, c0), gates(x, 1)))
        else:
            gates(bb, 2)

    for _ in range(depth):
        gates(b, n)
        block(b, 0)
    for q in range(min(n, nclbits)):
        b.Measure(q, q)
    return b.Build()
