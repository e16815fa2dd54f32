"""The benchmark workloads on the GPU, pinned to the oracle.

6x6 grid, 14 cycles (36 qubits, seed 2026), last column open (64
amplitudes), the reconfigured column sweep, 10 sliced bonds (1024 slices):
  * slices 0 and 1023 through the reference's drivers (slice_tensor_network
    + contract_network!, src/layer2.jl:294-321) on HIPBackend{ComplexF64}
    equal the oracle's contraction of the same slice .............. 1e-12
  * the plan executor in FP64 on slice 0 equals the oracle ......... 1e-12
  * the shipping configuration on slice 0 -- FP64 leaves and small
    steps, complex64 on FP32 MFMA, fused chains with 2^13-element tiles
    and operators up to 32 x 32 (4x4, 8->32, 16x16 and 32->8 stages all
    present) -- against the oracle, norm-wise ....................... 1e-5
  * all 1024 slices in FP64: equal to the unsliced contraction (slice-sum
    additivity, test/layer2_tests.jl:419-455) and to a second plan (the
    reconfigured row sweep) ......................................... 1e-12
  * all 1024 slices in the shipping configuration against FP64, worst
    single amplitude ................................................. 1e-5

7x7 grid, 12 cycles (49 qubits, seed 2027), one seeded amplitude, the
reconfigured column sweep with 8 bonds sliced under a per-slice cap of 2^22
elements (the unsliced plan peaks at 2^24, so the cap binds):
  * slices 0 and 255: the plan executor in FP64 on the pre-sliced network
    equals the oracle ................................................ 1e-12
  * all 256 slices in FP64 = the unsliced contraction ................ 1e-12
  * complex64 on FP32 MFMA and on the FP16 hi/lo split against FP64 .. 1e-5
"""
import numpy as np
import pytest

from oracle import pq_oracle as O
import picoquant_hip as P
from picoquant_hip import _lib, circuits as C, planner as PL
from picoquant_hip.executor import chain_stages, compile_plan
from picoquant_hip.slicing import slice_with_map, translate_plan

pytestmark = pytest.mark.gpu


def _cfg(rows, cols, seed, open_last_column):
    n = rows * cols
    bits = C.seeded_bits(n, seed + 1)
    open_q = [r * cols + cols for r in range(rows)] if open_last_column else []
    return "".join("x" if (q + 1) in open_q else str(bits[q]) for q in range(n)), open_q


def _net(rows, cols, depth, seed, cfg):
    n = rows * cols
    tn = P.network_from_gates(n, C.grid_circuit_gates(rows, cols, depth, seed),
                              backend=P.HIPBackend(np.complex128), decompose=True)
    P.add_input(tn, "0" * n)
    P.add_output(tn, cfg)
    return tn


def _values(s, nb):
    return [((s >> j) & 1) + 1 for j in range(nb)]


def _oracle_slice(rows, cols, depth, seed, cfg, plan, sliced, s):
    """The oracle's contraction of slice s (bond j at value bit j of s)."""
    n = rows * cols
    to = O.network_from_gates(n, O.grid_circuit_gates(rows, cols, depth, seed), decompose=True)
    O.add_input(to, "0" * n)
    O.add_output(to, cfg)
    old = to.counters["node"]
    label_map = {}
    for b, v in zip(sliced, _values(s, len(sliced))):
        for nd in (to.edges[b].src, to.edges[b].dst):
            view = f"node_{to.counters['node'] + 1}"
            O.replace_with_view(to, nd, b, (v, v))
            label_map[nd] = view

    def cur(x):
        while x in label_map:
            x = label_map[x]
        return x
    views = set(label_map.values())
    plan2 = translate_plan(plan, {k: cur(k) for k in label_map if k not in views}, old, to.counters["node"])
    O.contract_network(to, plan2)
    return np.asarray(to.backend.load_tensor_data("result"), dtype=np.complex128).reshape(-1, order="F")


def _sliced_net(rows, cols, depth, seed, cfg, plan, sliced, s):
    """A fresh network on the GPU with slice s applied (the reference's
    slice_tensor_network) and the plan renamed for it."""
    tn = _net(rows, cols, depth, seed, cfg)
    old = tn.counters["node"]
    m = slice_with_map(tn, sliced, _values(s, len(sliced)))
    return tn, translate_plan(plan, m, old, tn.counters["node"])


def _worst(got, want):
    """Worst relative error over the nonzero amplitudes; an amplitude that is
    exactly zero (a slice can close off part of an open column) must stay
    below 1e-14 of the largest."""
    got, want = np.asarray(got).reshape(-1), np.asarray(want).reshape(-1)
    nz = want != 0
    assert nz.any()
    if not nz.all():
        assert np.max(np.abs(got[~nz])) <= 1e-14 * np.max(np.abs(want)), "a zero amplitude is not zero"
    return float(np.max(np.abs(got[nz] - want[nz]) / np.abs(want[nz])))


def _normwise(got, want):
    got, want = np.asarray(got).reshape(-1), np.asarray(want).reshape(-1)
    return float(np.linalg.norm(got - want) / np.linalg.norm(want))


# --------------------------------------------------------------------------- #
# 6x6, depth 14
# --------------------------------------------------------------------------- #

G6 = (6, 6, 14, 2026)


@pytest.fixture(scope="module")
def grid6():
    rows, cols, depth, seed = G6
    cfg, open_q = _cfg(rows, cols, seed, True)
    tn = _net(rows, cols, depth, seed, cfg)
    plan = PL.reconfigure(tn, PL.sweep_plan(tn, PL.grid_site_order(rows, cols, "col", open_last=open_q)))
    sliced, _, peak = PL.choose_slices(tn, plan, 10)
    assert len(sliced) == 10 and peak <= 2 ** 28
    oracle = {s: _oracle_slice(rows, cols, depth, seed, cfg, plan, sliced, s) for s in (0, 1023)}
    return {"tn": tn, "cfg": cfg, "open_q": open_q, "plan": plan, "sliced": sliced, "oracle": oracle}


@pytest.mark.parametrize("s", [0, 1023])
def test_grid6_slice_reference_driver_matches_oracle(grid6, s):
    tn, plan2 = _sliced_net(*G6, grid6["cfg"], grid6["plan"], grid6["sliced"], s)
    P.contract_network(tn, plan2, "vector")
    got = np.asarray(tn.backend.load_tensor_data("result")).reshape(-1, order="F")
    assert got.shape == (64,)
    assert _worst(got, grid6["oracle"][s]) < 1e-12


def test_grid6_slice_executor_fp64_and_shipping_chains(grid6):
    tn, plan2 = _sliced_net(*G6, grid6["cfg"], grid6["plan"], grid6["sliced"], 0)
    want = grid6["oracle"][0]
    p64 = compile_plan(tn, plan2)
    p64.execute(True)
    assert _worst(p64.result().reshape(-1, order="F"), want) < 1e-12
    p64.close()
    # the bench's configuration: the executor's default fusion limits
    pm = compile_plan(tn, plan2, precision="mixed")
    shapes = set()
    for i, (kind, *_rest) in enumerate(pm.steps):
        if kind == 2:
            assert pm.step_info(i)["kc"] == _lib.KC_CHAIN
            shapes |= set(chain_stages(pm.chain_programs[i]))
    assert {(2, 2), (3, 5), (4, 4), (5, 3)} <= shapes, shapes
    tiles = {pm.chain_programs[i][1] for i, st in enumerate(pm.steps) if st[0] == 2}
    assert 13 in tiles
    ctx = tn.backend.ctx
    ctx.profile_enable(True)
    for use_graph in (False, True, True):
        pm.execute(use_graph)
        got = pm.result().reshape(-1, order="F")
        assert _normwise(got, want) < 1e-5
    assert ctx.profile_read(_lib.KC_CHAIN)["launches"] >= 2
    ctx.profile_enable(False)
    pm.close()


def test_grid6_all_slices_sum_plans_and_precision(grid6):
    tn, plan, sliced = grid6["tn"], grid6["plan"], grid6["sliced"]
    p = compile_plan(tn, plan, sliced, group_bits=10)
    p.execute(True)
    total = p.result().reshape(-1, order="F")
    p.close()
    # slice-sum additivity: the unsliced contraction
    u = compile_plan(tn, plan)
    u.execute(True)
    assert _worst(total, u.result().reshape(-1, order="F")) < 1e-12
    u.close()
    # a second plan of the same amplitudes (the reconfigured row sweep, its own slicing)
    rows, cols, _, _ = G6
    plan_r = PL.reconfigure(tn, PL.sweep_plan(tn, PL.grid_site_order(rows, cols, "row",
                                                                     open_last=grid6["open_q"])))
    sl_r = PL.choose_slices(tn, plan_r, 10)[0]
    r = compile_plan(tn, plan_r, sl_r, group_bits=10)
    r.execute(True)
    assert _worst(total, r.result().reshape(-1, order="F")) < 1e-12
    r.close()
    # the shipping configuration over all 1024 slices
    m = compile_plan(tn, plan, sliced, group_bits=10, precision="mixed")
    assert any(st[0] == 2 for st in m.steps)
    for use_graph in (True, True):
        m.execute(use_graph)
        assert _worst(m.result().reshape(-1, order="F"), total) < 1e-5
    m.close()


# --------------------------------------------------------------------------- #
# 7x7, depth 12
# --------------------------------------------------------------------------- #

G7 = (7, 7, 12, 2027)
CAP7 = 2 ** 22


@pytest.fixture(scope="module")
def grid7():
    rows, cols, depth, seed = G7
    cfg, _ = _cfg(rows, cols, seed, False)
    tn = _net(rows, cols, depth, seed, cfg)
    plan = PL.reconfigure(tn, PL.sweep_plan(tn, PL.grid_site_order(rows, cols, "col")))
    assert PL.plan_ledger(tn, plan)["max_size"] > CAP7          # the cap binds
    sliced, _, peak = PL.choose_slices(tn, plan, 8, max_elems=CAP7)
    assert peak <= CAP7
    return {"tn": tn, "cfg": cfg, "plan": plan, "sliced": sliced}


@pytest.mark.parametrize("s", [0, 255])
def test_grid7_slice_executor_matches_oracle(grid7, s):
    tn, plan2 = _sliced_net(*G7, grid7["cfg"], grid7["plan"], grid7["sliced"], s)
    p = compile_plan(tn, plan2)
    p.execute(True)
    got = p.result().reshape(-1)
    p.close()
    want = _oracle_slice(*G7, grid7["cfg"], grid7["plan"], grid7["sliced"], s)
    assert _worst(got, want) < 1e-12


def test_grid7_sliced_sum_and_precisions(grid7):
    tn, plan, sliced = grid7["tn"], grid7["plan"], grid7["sliced"]
    p = compile_plan(tn, plan, sliced)
    p.execute(True)
    ref = p.result().reshape(-1)
    p.close()
    u = compile_plan(tn, plan)
    u.execute(True)
    assert _worst(ref, u.result().reshape(-1)) < 1e-12
    u.close()
    for prec in ("mixed", "mixed16"):
        m = compile_plan(tn, plan, sliced, precision=prec)
        m.execute(True)
        assert _worst(m.result().reshape(-1), ref) < 1e-5, prec
        m.close()
