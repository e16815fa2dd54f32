"""Subtree reconfiguration (planner.reconfigure) on the CPU.

Pins: with a window covering every tensor, the result reaches the oracle's
exhaustive optimum over all contraction trees; reconfigured plans contract
to the same amplitudes as the original plan through the oracle's
contract_network (src/layer2.jl:250-321), and the oracle's own ledger
(Metrics.flops) counts exactly the MACs the planner claims; the cost never
rises."""
import math
import os
import pathlib

import numpy as np
import pytest

from oracle import pq_oracle as O
import picoquant_hip as P
from picoquant_hip import circuits as C, planner as PL
from test_host import RecordingBackend


def _pair(gates, n, inp, out, decompose=True):
    tp = P.network_from_gates(n, gates, backend=RecordingBackend(), decompose=decompose)
    P.add_input(tp, inp)
    P.add_output(tp, out)
    to = O.network_from_gates(n, gates, decompose=decompose)
    O.add_input(to, inp)
    O.add_output(to, out)
    return tp, to


def _tensors(tn):
    return [(list(v.indices), list(v.dims)) for v in tn.nodes.values()]


@pytest.mark.parametrize("seed", range(4))
def test_full_window_reaches_the_exhaustive_optimum(seed):
    gates = [(C.H, [0]), (C.CX, [0, 1]), (C.CX, [1, 2])]
    tn, _ = _pair(gates, 3, "000", "0x1", decompose=False)
    # a near-random pair plan: Boltzmann-greedy at a huge temperature
    plan = PL.bgreedy_single(tn, 1.0, 1e6, np.random.default_rng(seed))[0]
    before = PL.plan_ledger(tn, plan)["macs"]
    new = PL.reconfigure(tn, plan, windows=(16,))
    after = PL.plan_ledger(tn, new)["macs"]
    assert after <= before
    assert after == O.optimal_contraction_cost(_tensors(tn), outer=False)


@pytest.mark.parametrize("rows,cols,depth,out", [(2, 3, 6, "01x0x1"), (3, 3, 5, "01x0x1x10")])
def test_reconfigured_plan_same_amplitudes_and_ledger(rows, cols, depth, out):
    n = rows * cols
    gates = C.grid_circuit_gates(rows, cols, depth, seed=11)
    tp, to = _pair(gates, n, "0" * n, out)
    site = PL.grid_site_order(rows, cols, "col")
    plan = PL.sweep_plan(tp, site)
    new = PL.reconfigure(tp, plan, windows=(6, 10))
    assert len(new) == len(plan)
    led_old, led_new = PL.plan_ledger(tp, plan), PL.plan_ledger(tp, new)
    assert led_new["macs"] <= led_old["macs"]
    # the oracle contracts both plans on fresh copies of the same network
    results, flops = [], []
    for p in (plan, new):
        _, to = _pair(gates, n, "0" * n, out)
        O.contract_network(to, p)
        results.append(np.asarray(to.backend.load_tensor_data("result")))
        flops.append(to.backend.metrics.flops)
    np.testing.assert_allclose(results[1], results[0], rtol=1e-12, atol=1e-14)
    assert flops == [led_old["macs"], led_new["macs"]]


def test_reconfigure_cuts_the_grid_sweep_cost():
    """A 4x4 grid at depth 8: the sweep's cost drops, and a second call on
    the result changes nothing (a fixed point of the same windows)."""
    gates = C.grid_circuit_gates(4, 4, 8, seed=3)
    tn = P.network_from_gates(16, gates, backend=RecordingBackend(), decompose=True)
    P.add_input(tn, "0" * 16)
    P.add_output(tn, "0" * 12 + "xxxx")
    plan = PL.sweep_plan(tn, PL.grid_site_order(4, 4, "col", open_last=[4, 8, 12, 16]))
    new = PL.reconfigure(tn, plan, windows=(8,))
    m0, m1 = PL.plan_ledger(tn, plan)["macs"], PL.plan_ledger(tn, new)["macs"]
    assert m1 < m0
    again = PL.reconfigure(tn, new, windows=(8,))
    assert PL.plan_ledger(tn, again)["macs"] == m1


def test_translate_plan_when_one_node_loses_two_bonds():
    """slice_with_map + translate_plan: a node with two sliced bonds is
    replaced twice; the plan must name its last view, and the plan's own
    intermediates must only shift (they may share numbers with the
    intermediate views).  The sliced contraction equals the oracle's."""
    from picoquant_hip.slicing import slice_with_map, translate_plan
    gates = C.grid_circuit_gates(2, 3, 6, seed=5)
    tp, to = _pair(gates, 6, "000000", "0x1x00")
    plan = PL.reconfigure(tp, PL.sweep_plan(tp, PL.grid_site_order(2, 3, "col")), windows=(6,))
    def closed(x):
        return tp.edges[x].src is not None and tp.edges[x].dst is not None
    node = next(k for k, v in tp.nodes.items() if sum(1 for x in v.indices if closed(x)) >= 3)
    bonds = [x for x in tp.nodes[node].indices if closed(x)][:2]
    old = tp.counters["node"]
    m = slice_with_map(tp, bonds, [1, 0])
    assert len(m) == 3 and all(int(v.split("_")[1]) > old for v in m.values())
    plan2 = translate_plan(plan, m, old, tp.counters["node"])
    O.slice_tensor_network(to, bonds, [1, 0])
    O.contract_network(to, plan2)
    _, ref = _pair(gates, 6, "000000", "0x1x00")
    O.slice_tensor_network(ref, bonds, [1, 0])
    O.contract_network(ref, O.greedy_plan(ref)[0])
    np.testing.assert_allclose(np.asarray(to.backend.load_tensor_data("result")),
                               np.asarray(ref.backend.load_tensor_data("result")), rtol=1e-12, atol=1e-14)


def test_reconfigure_rejects_incomplete_plans():
    tn, _ = _pair([(C.H, [0]), (C.CX, [0, 1])], 2, "00", "00", decompose=False)
    plan = PL.bgreedy_single(tn, 1.0, 0.0)[0]
    with pytest.raises(ValueError):
        PL.reconfigure(tn, plan[:-1])


def test_partition_plan_with_one_leaf_is_the_exhaustive_optimum():
    """With `leaf` above the tensor count the top-down planner solves the
    whole network exactly (pq_netcon): the oracle's exhaustive optimum."""
    gates = [(C.H, [0]), (C.CX, [0, 1]), (C.CX, [1, 2])]
    tn, _ = _pair(gates, 3, "000", "0x1", decompose=False)
    plan = PL.partition_plan(tn, leaf=64)
    assert PL.plan_ledger(tn, plan)["macs"] == O.optimal_contraction_cost(_tensors(tn), outer=False)


@pytest.mark.parametrize("leaf", [4, 8])
def test_partition_plan_same_amplitudes_and_ledger(leaf):
    """Recursive bisection down to parts of < leaf tensors on a 3 x 3 grid
    circuit: a complete plan; the oracle contracts it to the amplitudes of a
    greedy plan, and its ledger counts exactly the MACs the planner claims."""
    gates = C.grid_circuit_gates(3, 3, 5, seed=11)
    out = "01x0x1x10"
    tp, _ = _pair(gates, 9, "0" * 9, out)
    plan = PL.partition_plan(tp, leaf=leaf)
    assert len(plan) == len(tp.nodes) - 1
    results, flops = [], []
    for p in (plan, None):
        _, to = _pair(gates, 9, "0" * 9, out)
        p = p if p is not None else O.greedy_plan(to)[0]
        O.contract_network(to, p)
        results.append(np.asarray(to.backend.load_tensor_data("result")))
        flops.append(to.backend.metrics.flops)
    np.testing.assert_allclose(results[0], results[1], rtol=1e-12, atol=1e-14)
    assert flops[0] == PL.plan_ledger(tp, plan)["macs"]


def test_bisection_is_balanced():
    gates = C.grid_circuit_gates(3, 3, 5, seed=2)
    tp, _ = _pair(gates, 9, "0" * 9, "0" * 9)
    from picoquant_hip.planner import _bisect
    nodes = list(tp.nodes)
    adj = {v: {} for v in nodes}
    owners = {}
    for lab, node in tp.nodes.items():
        for x in node.indices:
            owners.setdefault(x, []).append(lab)
    for labs in owners.values():
        if len(labs) == 2:
            a, b = labs
            adj[a][b] = adj[a].get(b, 0) + 1.0
            adj[b][a] = adj[b].get(a, 0) + 1.0
    a, b = _bisect(nodes, adj, 0.1, np.random.default_rng(0))
    assert sorted(a + b) == sorted(nodes)
    assert abs(len(a) - len(b)) <= 0.2 * len(nodes) + 1


def test_sampled_greedy_plan_is_reproducible_across_processes():
    """A Boltzmann-greedy plan (tau > 0, seeded rng) on a network large
    enough for the heap variant is the same in processes with different
    string-hash seeds: no draw order depends on set iteration."""
    import subprocess
    import sys
    code = (
        "import sys, hashlib, numpy as np\n"
        "from picoquant_hip import circuits as C, planner as PL\n"
        "import picoquant_hip as P\n"
        "from test_host import RecordingBackend\n"
        "tn = P.network_from_gates(30, C.grid_circuit_gates(5, 6, 12, 5), backend=RecordingBackend(), decompose=True)\n"
        "P.add_input(tn, '0' * 30); P.add_output(tn, '0' * 30)\n"
        "plan, tc, sc = PL.bgreedy_single(tn, tau=0.5, rng=np.random.default_rng(3), exact_limit=10)\n"
        "print(hashlib.md5(repr(plan).encode()).hexdigest(), tc)\n")
    root = pathlib.Path(__file__).resolve().parent.parent
    paths = [str(root), str(root / "picoquant.jl_amd"), str(root / "tests")]
    outs = []
    for seed in ("1", "2"):
        env = dict(os.environ, PYTHONHASHSEED=seed, PYTHONPATH=":".join(paths))
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=300)
        assert r.returncode == 0, r.stderr
        outs.append(r.stdout.strip())
    assert outs[0] == outs[1]
