"""Host logic of the sweep, observe and local-backend mirrors (no device): the reference's
sweep_test.go known answers (sweep/sweep_test.go:8-260), parameter validation (run.go:173-190),
the backend's request checks (backend/local/local.go:70-90) and hook plumbing
(observe/observe.go:104-113)."""
import math

import pytest

from goqu_amd import backend as BK
from goqu_amd import observe
from goqu_amd import sweep as SW
from goqu_amd.circuit import New, NewBuilder


def test_linspace():
    assert SW.Linspace("x", 0, 1, 0).Len() == 0 and SW.Linspace("x", 0, 1, 0).Resolve() == []
    assert SW.Linspace("x", 0, 1, -3).Len() == 0
    assert SW.Linspace("x", 3.14, 6.28, 1).Resolve() == [{"x": 3.14}]
    r = SW.Linspace("theta", 0, math.pi, 5).Resolve()
    assert len(r) == 5 and r[0]["theta"] == 0 and abs(r[4]["theta"] - math.pi) < 1e-12
    assert abs(r[2]["theta"] - math.pi / 2) < 1e-12
    assert SW.Linspace("theta", 0, math.pi, 5).Params() == ["theta"]


def test_points_and_single_copy_their_inputs():
    assert SW.NewPoints("x", []).Resolve() == []
    vals = [1.0, 2.0, 3.0]
    p = SW.NewPoints("x", vals)
    vals[0] = 99
    assert p.Resolve() == [{"x": 1.0}, {"x": 2.0}, {"x": 3.0}] and p.Params() == ["x"]
    b = {"a": 1.0, "b": 2.0}
    s = SW.Single(b)
    b["a"] = 5
    assert s.Len() == 1 and s.Resolve() == [{"a": 1.0, "b": 2.0}] and sorted(s.Params()) == ["a", "b"]


def test_product():
    p = SW.Product(SW.NewPoints("a", [1, 2]), SW.NewPoints("b", [10, 20, 30]))
    assert p.Len() == 6 and p.Params() == ["a", "b"]
    assert p.Resolve()[0] == {"a": 1, "b": 10} and p.Resolve()[1] == {"a": 1, "b": 20}
    assert p.Resolve()[3] == {"a": 2, "b": 10}
    assert SW.Product(SW.NewPoints("a", [1, 2]), SW.NewPoints("b", [])).Len() == 0
    assert SW.Product(SW.NewPoints("a", [1, 2])).Len() == 2
    assert SW.Product(SW.NewPoints("a", [1, 2]), SW.NewPoints("b", [1, 2]), SW.NewPoints("c", [1, 2])).Len() == 8
    o = SW.Product(SW.NewPoints("a", [1]), SW.NewPoints("a", [2]))
    assert o.Params() == ["a"] and o.Resolve() == [{"a": 2}]  # later factors win
    assert SW.Product().Len() == 0


def test_zip():
    z = SW.Zip(SW.NewPoints("a", [1, 2, 3]), SW.NewPoints("b", [4, 5, 6]))
    assert z.Len() == 3 and z.Resolve()[1] == {"a": 2, "b": 5} and z.Params() == ["a", "b"]
    with pytest.raises(ValueError):
        SW.Zip(SW.NewPoints("a", [1, 2]), SW.NewPoints("b", [1]))
    assert SW.Zip().Len() == 0
    o = SW.Zip(SW.NewPoints("a", [1, 2]), SW.NewPoints("a", [3, 4]))
    assert o.Params() == ["a"] and o.Resolve() == [{"a": 3}, {"a": 4}]


def test_unbound_parameter_is_an_error():
    """TestRunSim_UnboundParam (run_test.go:87-95): checked before any simulation."""
    c = NewBuilder("ry", 1).SymRY(New("theta"), 0).MeasureAll().Build()
    with pytest.raises(ValueError):
        SW.RunSim(None, c, 100, SW.NewPoints("wrong_name", [1.0]))


def test_backend_request_checks_and_options():
    """local.New defaults (28 qubits), WithMaxQubits, and Submit's checks before simulating."""
    big = NewBuilder("big", 30).H(0).MeasureAll().Build()
    b = BK.New()
    assert b.maxQubits == 28 and b.Name() == "local.simulator"
    with pytest.raises(ValueError, match="max is 28"):
        b.Submit(None, BK.SubmitRequest(big, 100))
    with pytest.raises(ValueError):
        b.Submit(None, BK.SubmitRequest(None, 100))
    small = NewBuilder("s", 1).X(0).MeasureAll().Build()
    with pytest.raises(ValueError):
        b.Submit(None, BK.SubmitRequest(small, 0))
    b34 = BK.New(BK.WithMaxQubits(34), BK.WithSimulator(BK.NewSimulator))
    assert b34.maxQubits == 34 and b34.simFactory is BK.NewSimulator


def test_hooks_in_context():
    h = observe.Hooks(WrapSim=lambda ctx, info: (ctx, lambda e: None))
    ctx = observe.WithHooks(observe.Background(), h)
    assert observe.FromContext(ctx) is h
    assert observe.FromContext(observe.Background()) is None and observe.FromContext(None) is None


def test_gate_count_counts_control_flow_bodies():
    b = NewBuilder("cf", 2).WithClbits(1).H(0).Measure(0, 0)
    b.IfElseBlock(0, 1, lambda x: x.X(1).Z(1), lambda x: x.Y(1))
    assert BK._gate_count(b.Build()) == 3 + 3


def test_autotune_cache_roundtrip(tmp_path, monkeypatch):
    """The tuner's JSON cache (GOQU_TUNE_CACHE): a stored choice reads back; a corrupt file
    reads as empty; shapes outside the kernel's limits are never candidates."""
    from goqu_amd import autotune as AT

    monkeypatch.setenv("GOQU_TUNE_CACHE", str(tmp_path / "t.json"))
    assert AT._load() == {}
    AT._store({"k": {"tile_bits": 11, "low_bits": 3}})
    assert AT._load()["k"]["tile_bits"] == 11
    (tmp_path / "t.json").write_text("{not json")
    assert AT._load() == {}
    assert all(8 <= t <= 13 and 2 <= l <= min(6, t - 4) for t, l in AT.DEFAULT_CANDIDATES)
    assert not AT._valid((13, 4), 10) and AT._valid((10, 4), 10) and not AT._valid((12, 9), 30)


def test_param_records_equal_bound_circuit_records():
    """The vectorised parameter-batch marshalling gives, row by row, the op records (kinds, tiers,
    qubits) and matrices of binding each row and marshalling the bound circuit."""
    import ctypes as C

    import numpy as np

    from goqu_amd import batch as B
    from goqu_amd import qaoa as Q
    from goqu_amd.circuit import Add, FreeParameters, Literal, Mul

    c, npar = Q.EfficientSU2(5, 2)
    th = New("extra")
    b = NewBuilder("more", 5)
    for op in c.Ops():
        b.ops.append(op)
    b.SymCP(Mul(Literal(2.0), th), 0, 3).SymRXX(Add(th, Literal(0.5)), 1, 2).SymU3(th, th, th, 4)
    b.SymRot(th, th, th, 2).SymFSim(th, th, 0, 1).H(2).SymRZZ(th, 3, 4).SymRYY(th, 0, 4).SymPhase(th, 1)
    c2 = b.Build()
    names = FreeParameters(c2)
    params = np.random.default_rng(1).uniform(0, 6.28, (4, len(names)))
    R = B.ParamRecords(c2, params, names)
    buf, _ = B._marshal_sv(B.bind_all(c2, params, names))
    st = buf.structured()
    assert len(st) == len(R.records)
    for a, r in zip(st, R.records):
        assert a["kind"] == r["kind"] and a["gate_id"] == r["gate_id"] and (a["qubits"] == r["qubits"]).all()
        k = (1 << a["num_qubits"]) ** 2
        ma = np.ctypeslib.as_array((C.c_double * (2 * k)).from_address(int(a["matrix"])))
        mr = np.ctypeslib.as_array((C.c_double * (2 * k)).from_address(int(r["matrix"])))
        assert np.max(np.abs(ma - mr)) < 1e-15
