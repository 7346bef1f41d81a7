"""goqu_sv_sample_noalloc, the sampler that allocates nothing (its buffers live in a static
device array), against goqu_sv_sample through the C ABI: identical indices for the same
uniforms on two calls in a row, the second on a smaller input; device allocation is armed to
fail during the calls, so any allocation would show as an error; oversize requests are
rejected."""
import ctypes as C

import numpy as np
import pytest

from circuits import random_circuit
from goqu_amd import _native as N
from goqu_amd import statevector

pytestmark = pytest.mark.gpu


def _sample(fn, sim, u):
    out = np.zeros(len(u), dtype=np.uint64)
    rc = fn(sim._h, N.dptr(u), len(u), N.u64ptr(out))
    return rc, out


@pytest.mark.parametrize("precision", [N.C128, N.C64])
def test_sample_noalloc_matches_sample(precision):
    L = N.lib()
    for n, shots, seed in [(14, 3000, 1), (9, 257, 2)]:  # the second call: a smaller input
        rng = np.random.default_rng(600 + seed)
        c = random_circuit(n, 4, rng)
        sim = statevector.New(n, precision)
        sim.Evolve(c)
        u = rng.random(shots)
        rc, want = _sample(L.goqu_sv_sample, sim, u)
        assert rc == N.OK
        # any device allocation during the next call fails (and would surface as an error)
        assert L.goqu_debug_inject_fault(N.FAULT_DEVICE_ALLOC, 1) == N.OK
        rc, got = _sample(L.goqu_sv_sample_noalloc, sim, u)
        assert L.goqu_debug_inject_fault(0, 0) == N.OK
        assert rc == N.OK, L.goqu_last_error()
        assert np.array_equal(got, want), (n, shots)
        sim.Close()


def test_sample_noalloc_rejects_oversize_requests():
    L = N.lib()
    sim = statevector.New(10)
    u = np.random.default_rng(3).random(600_000)  # 16 B per shot: beyond the 8 MiB array
    rc, _ = _sample(L.goqu_sv_sample_noalloc, sim, u)
    assert rc == N.ERR_INVALID_ARG
    assert b"static scratch" in L.goqu_last_error()
    rc, _ = _sample(L.goqu_sv_sample_noalloc, sim, u[:1000])
    assert rc == N.OK
    sim.Close()
