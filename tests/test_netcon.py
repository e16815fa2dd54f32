"""netcon (src/layer2/netcon_contraction.jl) on the CPU: the native search
in libpq_hip (pq_netcon, host code) against the oracle's exhaustive
optimum over all contraction trees, plus the reference's plan conversion.
The oracle's optimum is pinned by hand-counted chains."""
import math

import numpy as np
import pytest

from oracle import pq_oracle as O
import picoquant_hip as P
from picoquant_hip import _lib, circuits as C, optimal_order as NC, planner as PL
from test_host import RecordingBackend


def test_oracle_optimum_hand_counted():
    # A(2x3) B(3x4) C(4x5): (AB)C = 24 + 40 = 64; A(BC) = 60 + 30 = 90
    chain = [(["a", "i"], [2, 3]), (["i", "j"], [3, 4]), (["j", "b"], [4, 5])]
    assert O.optimal_contraction_cost(chain) == 64
    # vector-matrix-matrix-vector, all extents 10: 100 + 100 + 10 = 210
    vmmv = [(["i"], [10]), (["i", "j"], [10, 10]), (["j", "k"], [10, 10]), (["k"], [10])]
    assert O.optimal_contraction_cost(vmmv) == 210
    # two disconnected vectors: one outer product of 2 x 3
    assert O.optimal_contraction_cost([(["a"], [2]), (["b"], [3])]) == 6


def _random_network(rng, n):
    """Connected random network: a random spanning tree plus extra bonds,
    extents 2-5, some open legs."""
    tensors = [([], []) for _ in range(n)]
    lab = 0
    def bond(a, b, d):
        nonlocal lab
        tensors[a][0].append(f"x{lab}")
        tensors[a][1].append(d)
        tensors[b][0].append(f"x{lab}")
        tensors[b][1].append(d)
        lab += 1
    for t in range(1, n):
        bond(int(rng.integers(t)), t, int(rng.integers(2, 6)))
    for _ in range(int(rng.integers(0, n))):
        a, b = rng.choice(n, 2, replace=False)
        bond(int(a), int(b), int(rng.integers(2, 5)))
    for t in range(n):
        if rng.random() < 0.4:
            tensors[t][0].append(f"o{t}")
            tensors[t][1].append(int(rng.integers(2, 4)))
    return tensors


@pytest.mark.parametrize("seed", range(12))
def test_netcon_is_optimal_on_random_networks(seed):
    rng = np.random.default_rng(seed)
    tensors = _random_network(rng, int(rng.integers(3, 8)))
    idx = [t[0] for t in tensors]
    dims = {i: d for t in tensors for i, d in zip(*t)}
    tree, cost = NC.optimal_tree(idx, dims)
    assert cost == O.optimal_contraction_cost(tensors, outer=False)
    assert cost >= O.optimal_contraction_cost(tensors)
    # the tree really costs that much
    def walk(t):
        if isinstance(t, int):
            return (list(tensors[t][0]), list(tensors[t][1])), 0
        (ia, da), ca = walk(t[0])
        (ib, db), cb = walk(t[1])
        dd = dict(zip(ia, da))
        dd.update(zip(ib, db))
        union = list(dict.fromkeys(ia + ib))
        macs = math.prod(dd[i] for i in union)
        keep = [i for i in union if (i in ia) != (i in ib)]
        return (keep, [dd[i] for i in keep]), ca + cb + macs
    assert walk(tree)[1] == cost


def _circuit(gates, n, cfg_in, cfg_out=None):
    tn = P.network_from_gates(n, gates, backend=RecordingBackend())
    P.add_input(tn, cfg_in)
    if cfg_out is not None:
        P.add_output(tn, cfg_out)
    return tn


def _as_tensors(tn):
    return [(list(v.indices), list(v.dims)) for v in tn.nodes.values()]


def test_netcon_reference_network_and_plan():
    """The reference's netcon test network (H, CX, CX on 3 qubits, input and
    output 000): optimal cost, a valid plan whose ledger matches, and the
    tree / raw return forms."""
    gates = [(C.H, [0]), (C.CX, [0, 1]), (C.CX, [0, 2])]
    tn = _circuit(gates, 3, "000", "000")
    plan = NC.netcon(tn)
    assert len(plan) == len(tn.nodes) - 1
    led = PL.plan_ledger(tn, plan)
    _, cost = NC.optimal_tree(*NC.create_netcon_input(tn)[:2])
    assert led["macs"] == cost == O.optimal_contraction_cost(_as_tensors(tn), outer=False)
    tree = NC.netcon(tn, "tree")
    flat = []
    def leaves(t):
        if isinstance(t, str):
            flat.append(t)
        else:
            leaves(t[0]); leaves(t[1])
    leaves(tree)
    assert sorted(flat) == sorted(tn.nodes)
    raw, labels = NC.netcon(tn, "raw")
    assert NC.replace_leaf_integers_with_labels(raw, labels) == tree


@pytest.mark.parametrize("n", [3, 4])
def test_netcon_beats_or_ties_greedy_on_qft(n):
    tn = _circuit(C.qft_gates(n), n, "0" * n, "1" * n)
    tn_oracle = _as_tensors(tn)
    if len(tn_oracle) <= 14:
        plan = NC.netcon(tn)
        g, _, _ = PL.bgreedy_single(tn, 1.0, 0.0)
        assert PL.plan_ledger(tn, plan)["macs"] <= PL.plan_ledger(tn, g)["macs"]


def test_netcon_rejects_bad_input_and_limits():
    with pytest.raises(_lib.PQError):       # an index on three tensors
        NC.optimal_tree([["a"], ["a"], ["a"]], {"a": 2})
    with pytest.raises(_lib.PQError):       # a trace inside one tensor
        NC.optimal_tree([["a", "a"]], {"a": 2})
    rng = np.random.default_rng(5)
    tensors = _random_network(rng, 9)
    with pytest.raises(_lib.PQError):       # work cap
        NC.optimal_tree([t[0] for t in tensors], {i: d for t in tensors for i, d in zip(*t)}, max_work=3)
    tn = _circuit(C.qft_gates(6), 6, "0" * 6)
    with pytest.raises(ValueError):
        big = _circuit(C.qft_gates(8), 8, "0" * 8)
        NC.netcon_contraction(big)
