"""Planned executor (pq_plan_*): whole plans over all slices with the FP64
slice sum, against brute-force state vectors, the committed config-2
amplitudes and the reference driver path on the same backend.

Tolerances (worst relative amplitude error):
  ComplexF64 ...................................... 1e-12
  complex64 everywhere (gate tensors rounded too) .. 5e-5
  mixed: FP64 leaves and small steps, complex64 on FP32 MFMA for the
  large steps, FP64 slice sum (the bench path) ..... 1e-5"""
import json

import numpy as np
import pytest

from oracle import pq_oracle as O
import picoquant_hip as P
from picoquant_hip import circuits as C
from picoquant_hip import planner as PL
from picoquant_hip.executor import compile_plan

pytestmark = pytest.mark.gpu

RTOL = {np.complex128: 1e-12, np.complex64: 5e-5, "mixed": 1e-5}


@pytest.fixture(scope="module", params=[np.complex128, np.complex64], ids=["c128", "c64"])
def dtype(request):
    return request.param


def grid_net(rows, cols, depth, seed, cfg, dtype, backend=None):
    n = rows * cols
    tn = P.network_from_gates(n, C.grid_circuit_gates(rows, cols, depth, seed),
                              backend=backend or P.HIPBackend(dtype), decompose=True)
    P.add_input(tn, "0" * n)
    P.add_output(tn, cfg)
    return tn


def batch_amps(sv, cfg):
    open_q = [i for i, ch in enumerate(cfg) if ch == "x"]
    out = []
    for j in range(2 ** len(open_q)):
        bits = [int(ch) if ch != "x" else 0 for ch in cfg]
        for t, q in enumerate(open_q):
            bits[q] = (j >> t) & 1
        out.append(sv[sum(b << q for q, b in enumerate(bits))])
    return np.array(out)


def rel(got, want):
    """Worst relative error; results are column-major (first open qubit fastest)."""
    g = np.asarray(got).reshape(-1, order="F")
    return float(np.max(np.abs(g - np.asarray(want).reshape(-1, order="F")) / np.abs(want).reshape(-1, order="F")))


@pytest.mark.parametrize("carry", [True, False], ids=["carry", "no-carry"])
@pytest.mark.parametrize("nslice,group_bits", [(0, 0), (4, 0), (4, 2), (4, 4), (6, 3)])
def test_grid3x3_sliced(dtype, nslice, group_bits, carry):
    """Slice groups of every size; with small intermediates keeping their
    sliced bonds (read at per-slice and per-group offsets) and without."""
    cfg = "1x0x11x01"
    tn = grid_net(3, 3, 6, 21, cfg, dtype)
    sv = O.statevector(9, O.grid_circuit_gates(3, 3, 6, 21))
    open_q = [i + 1 for i, ch in enumerate(cfg) if ch == "x"]
    plan = PL.sweep_plan(tn, PL.grid_site_order(3, 3, open_last=open_q))
    sliced = PL.choose_slices(tn, plan, nslice)[0] if nslice else []
    kw = {} if carry else {"carry_elems": 0}
    p = compile_plan(tn, plan, sliced, group_bits=group_bits, **kw)
    if carry and nslice:
        assert any(s.carry for s in p.slots)
    want = batch_amps(sv, cfg)
    for use_graph in (False, True, True, True):
        p.execute(use_graph)
        assert rel(p.result(), want) < RTOL[dtype]
    p.close()


def test_greedy_plan_qft(dtype):
    """A bgreedy (tau -> 0) plan of QFT-7 from a basis state, all outputs open."""
    n, x = 7, 45
    tn = P.network_from_gates(n, C.qft_gates(n), backend=P.HIPBackend(dtype))
    P.add_input(tn, "".join(str((x >> q) & 1) for q in range(n)))
    plan, _, _ = PL.bgreedy_single(tn, 1.0, 0.0)
    p = compile_plan(tn, plan)
    p.execute(True)
    assert rel(p.result(), O.qft_closed_form(n, x)) < RTOL[dtype]


def test_grid5x5_d12_batch(golden_dir, dtype):
    """Config 2 through the planned executor: the open-column batch of 32
    amplitudes with 6 sliced bonds, replayed as a graph."""
    g = json.load(open(golden_dir / "grid5x5_d12.json"))
    cfg = list(g["batch_base"])
    for q in g["open_qubits_1based"]:
        cfg[q - 1] = "x"
    cfg = "".join(cfg)
    tn = grid_net(5, 5, 12, g["seed"], cfg, dtype)
    plan = PL.sweep_plan(tn, PL.grid_site_order(5, 5, open_last=g["open_qubits_1based"]))
    sliced, _, peak = PL.choose_slices(tn, plan, 6)
    p = compile_plan(tn, plan, sliced, group_bits=3)
    want = np.array([complex(a, b) for a, b in g["batch_amplitudes"]])
    for _ in range(3):
        p.execute(True)
        assert rel(p.result(), want) < RTOL[dtype]
    st = p.stats()
    assert st["out_elems"] == 32 and st["arena_bytes"] > 0
    ms = p.profile()
    assert ms.shape == (len(plan),) and np.all(ms >= 0)


def test_plan_matches_reference_driver_per_slice():
    """Slice-sum of the reference drivers (slice_tensor_network +
    contract_network! per slice on HIPBackend) equals the planned executor."""
    cfg = "0x1x01"
    be = P.HIPBackend(np.complex128)
    tn = grid_net(2, 3, 5, 6, cfg, np.complex128, be)
    open_q = [i + 1 for i, ch in enumerate(cfg) if ch == "x"]
    plan = PL.sweep_plan(tn, PL.grid_site_order(2, 3, open_last=open_q))
    sliced = PL.choose_slices(tn, plan, 3)[0]
    p = compile_plan(tn, plan, sliced, group_bits=1)
    p.execute(True)
    got = p.result().reshape(-1, order="F")
    total = 0
    for s in range(8):
        t2 = grid_net(2, 3, 5, 6, cfg, np.complex128)
        P.slice_tensor_network(t2, sliced, [((s >> j) & 1) + 1 for j in range(3)])
        # the slice views are new nodes: plan by bgreedy on this slice
        plan2, _, _ = PL.bgreedy_single(t2, 1.0, 0.0)
        P.contract_network(t2, plan2, "vector")
        total = total + t2.backend.load_tensor_data("result")
    assert rel(got, total) < 1e-12


def test_plan_rejects_bad_descriptors():
    from picoquant_hip import _lib
    tn = grid_net(2, 2, 3, 1, "0x10", np.complex128)
    plan = PL.sweep_plan(tn, PL.grid_site_order(2, 2, open_last=[2]))
    with pytest.raises(ValueError):
        compile_plan(tn, plan[:-1])                    # leaves two tensors
    with pytest.raises(_lib.PQError):
        compile_plan(tn, plan, group_bits=0, kc=[_lib.KC_GEMM] * len(plan))   # M, N < 16


@pytest.mark.parametrize("nslice,group_bits", [(0, 0), (6, 3), (6, 6)])
def test_grid5x5_d12_mixed_precision(golden_dir, nslice, group_bits):
    """Config 2 on the bench's precision split: FP64 world lines, complex64
    sweep on FP32 MFMA, FP64 slice sum -- within 1e-5 of the FP64 amplitudes."""
    g = json.load(open(golden_dir / "grid5x5_d12.json"))
    cfg = list(g["batch_base"])
    for q in g["open_qubits_1based"]:
        cfg[q - 1] = "x"
    tn = grid_net(5, 5, 12, g["seed"], "".join(cfg), np.complex128)
    plan = PL.sweep_plan(tn, PL.grid_site_order(5, 5, open_last=g["open_qubits_1based"]))
    sliced = PL.choose_slices(tn, plan, nslice)[0] if nslice else []
    # FP64 only up to 2^16 elements, so the sweep's large steps run in complex64
    p = compile_plan(tn, plan, sliced, group_bits=group_bits, precision="mixed", small_elems=1 << 16)
    assert {d for d in p.dtypes} == {0, 1}                 # both element types in use
    assert any(k == 1 for k, *_ in p.steps)                # with conversions
    want = np.array([complex(a, b) for a, b in g["batch_amplitudes"]])
    for _ in range(2):
        p.execute(True)
        assert rel(p.result(), want) < RTOL["mixed"]
    p.close()


def test_plan_keeps_its_leaves_after_they_are_freed():
    """A compiled plan holds its leaves' device buffers: deleting leaf
    tensors (as contract_tensors does with its operands) or reshaping them,
    then allocating over the freed memory, leaves every later run -- eager
    and graph -- with the same amplitudes."""
    cfg = "1x0x11x01"
    tn = grid_net(3, 3, 6, 21, cfg, np.complex128)
    open_q = [i + 1 for i, ch in enumerate(cfg) if ch == "x"]
    plan = PL.sweep_plan(tn, PL.grid_site_order(3, 3, open_last=open_q))
    sliced = PL.choose_slices(tn, plan, 2)[0]
    p = compile_plan(tn, plan, sliced, group_bits=1)
    p.execute(False)
    want = p.result().copy()
    be, ctx = tn.backend, tn.backend.ctx
    labels = list(tn.nodes)
    for lab in labels[: len(labels) // 2]:
        be.delete_tensor(lab)
    for lab in labels[len(labels) // 2:]:
        if len(tn.nodes[lab].dims) > 1:
            ctx.reshape(be.tensors[lab], [int(np.prod(tn.nodes[lab].dims))])
    junk = [ctx.upload(np.full(64, 7 + 7j, dtype=np.complex128)) for _ in range(256)]
    for use_graph in (False, True, True):
        p.execute(use_graph)
        np.testing.assert_array_equal(p.result(), want)
    for h in junk:
        ctx.free(h)
    p.close()


def test_plan_on_a_network_sliced_by_the_reference_driver():
    """compile_plan on a network after slice_tensor_network: the views have
    extent 1 on the sliced bond while their Nodes keep the old dims (as in
    the reference, src/layer2/slicing.jl:67-93); the plan lowers the stored
    shapes and equals contract_network! on the same slice."""
    cfg = "0x1x01"
    for s in range(4):
        vals = [(s & 1) + 1, ((s >> 1) & 1) + 1]
        t1 = grid_net(2, 3, 5, 6, cfg, np.complex128)
        open_q = [i + 1 for i, ch in enumerate(cfg) if ch == "x"]
        plan = PL.sweep_plan(t1, PL.grid_site_order(2, 3, open_last=open_q))
        sliced = PL.choose_slices(t1, plan, 2)[0]
        from picoquant_hip.slicing import slice_with_map, translate_plan
        old = t1.counters["node"]
        m = slice_with_map(t1, sliced, vals)
        plan1 = translate_plan(plan, m, old, t1.counters["node"])
        p = compile_plan(t1, plan1)
        p.execute(True)
        got = p.result().reshape(-1, order="F")
        p.close()
        t2 = grid_net(2, 3, 5, 6, cfg, np.complex128)
        P.slice_tensor_network(t2, sliced, vals)
        P.contract_network(t2, plan1, "vector")
        want = np.asarray(t2.backend.load_tensor_data("result")).reshape(-1, order="F")
        # a slice can close off amplitudes exactly: compare against the largest
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


@pytest.mark.skip(reason="concurrent plan streams are switched off: a run of this test's build with four "
                         "streams failed mixed-precision plans and ended in a host segmentation fault inside "
                         "pq_plan_execute, cause not found")
@pytest.mark.parametrize("case", ["grid5x5_sliced", "grid4x4_chains_bitstrings"])
def test_concurrent_schedule_matches_one_stream(golden_dir, case):
    """pq_set_plan_streams: a plan whose independent launches run on four
    streams (parallel branches of the captured graph) gives the same bits as
    the same plan on one stream, eager and replayed, and the right amplitudes."""
    if case == "grid5x5_sliced":
        g = json.load(open(golden_dir / "grid5x5_d12.json"))
        cfg = list(g["batch_base"])
        for q in g["open_qubits_1based"]:
            cfg[q - 1] = "x"
        tn = grid_net(5, 5, 12, g["seed"], "".join(cfg), np.complex128)
        plan = PL.reconfigure(tn, PL.sweep_plan(tn, PL.grid_site_order(5, 5, open_last=g["open_qubits_1based"])))
        sliced = PL.choose_slices(tn, plan, 6)[0]
        kw = dict(sliced=sliced, group_bits=3, precision="mixed")
        want = np.array([complex(a, b) for a, b in g["batch_amplitudes"]])
        tol = RTOL["mixed"]
    else:
        cfg0 = "0110" + "x" * 4 + "1001" + "01" + "x" * 2
        rng = np.random.default_rng(5)
        batch = ["".join("x" if c == "x" else str(int(rng.integers(0, 2))) for c in cfg0) for _ in range(4)]
        tn = grid_net(4, 4, 8, 7, batch[0], np.complex128)
        plan = PL.reconfigure(tn, PL.sweep_plan(tn, PL.grid_site_order(4, 4, "col")), windows=(8,))
        kw = dict(precision="mixed", small_elems=16, bitstrings=batch,
                  fuse={"tile_bits": 10, "min_elems": 64, "op_elems": 256})
        sv = O.statevector(16, O.grid_circuit_gates(4, 4, 8, 7))
        open_q = [i for i, ch in enumerate(cfg0) if ch == "x"]
        want = []
        for bits in batch:
            for j in range(2 ** len(open_q)):
                full = list(bits)
                for t, q in enumerate(open_q):
                    full[q] = str((j >> t) & 1)
                want.append(sv[sum(int(ch) << q for q, ch in enumerate(full))])
        want = np.array(want)
        tol = 2e-6
    ctx = tn.backend.ctx
    out = {}
    for ns in (1, 4):
        ctx.set_plan_streams(ns)
        try:
            p = compile_plan(tn, plan, **kw)
        finally:
            ctx.set_plan_streams(4)
        runs = []
        for use_graph in (False, True, True, True):
            p.execute(use_graph)
            runs.append(p.result().copy())
        p.close()
        for r in runs[1:]:
            np.testing.assert_array_equal(r, runs[0])
        out[ns] = runs[0]
    np.testing.assert_array_equal(out[1], out[4])
    got = out[4].reshape(-1) if case == "grid5x5_sliced" else np.concatenate(
        [out[4][b].reshape(-1, order="F") for b in range(out[4].shape[0])])
    if case == "grid5x5_sliced":
        got = out[4].reshape(-1, order="F")
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < tol
