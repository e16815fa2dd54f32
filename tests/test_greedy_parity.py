"""bgreedy in the tau -> 0 limit (src/layer2/bgreedy_contraction.jl:17-45,
:76-131): hand-worked known answers for the oracle's greedy_plan, the
product's planner equal to the oracle on real networks, and the product's
heap path (networks above 160 tensors) equal to its all-pairs scan.

Known-answer network (five tensors, in insertion order):
    T1(a:2, b:8)  T2(b:8, c:4)  T3(c:4, d:3)  T4(d:3)  T5(a:2, e:5)
Energy E = |C| - alpha (|A| + |B|) of every pair, alpha = 1, worked by hand:
  step 1  T1T2 -40  T1T3 164  T1T4 29  T1T5 14  T2T3 -20  T2T4 61  T2T5 278
          T3T4 -11  T3T5 98  T4T5 17                  -> T1,T2 = N6(a, c)
  step 2  T3T4 -11  T3T5 98  T3N6 -14  T4T5 17  T4N6 13  T5N6 2
                                                      -> T3,N6 = N7(d, a)
  step 3  T4T5 17  T4N7 -7  T5N7 -1                   -> T4,N7 = N8(a)
  step 4  T5N8 -7                                     -> T5,N8
  time cost sqrt(16*32*8) + sqrt(12*8*6) + sqrt(3*6*2) + sqrt(10*2*5)
          = 64 + 24 + 6 + 10 = 104; space cost max |C| = 8.
With alpha = 0 (E = |C|): T3,T4 (4); then T1,T2 (8); then N6,N7 (2);
then T5,N8: time cost 12 + 64 + 8 + 10 = 94, space 8.
A flipped alpha sign starts with T3,T4 (E = 19 against 56 for T1,T2)."""
import math
from collections import OrderedDict
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import pq_oracle as O
import picoquant_hip as P
from picoquant_hip import circuits as C, planner as PL
from test_host import RecordingBackend

KAT = [("node_1", ["a", "b"], [2, 8]), ("node_2", ["b", "c"], [8, 4]), ("node_3", ["c", "d"], [4, 3]),
       ("node_4", ["d"], [3]), ("node_5", ["a", "e"], [2, 5])]


def _kat_network(spec=KAT):
    nodes = OrderedDict((lab, O.Node(list(ix), list(d), lab)) for lab, ix, d in spec)
    return SimpleNamespace(nodes=nodes, counters={"node": len(spec)})


def test_oracle_greedy_known_answer_alpha_1():
    plan, tcost, scost = O.greedy_plan(_kat_network(), alpha=1.0)
    assert plan == [["node_1", "node_2"], ["node_3", "node_6"], ["node_4", "node_7"], ["node_5", "node_8"]]
    assert tcost == pytest.approx(104.0)
    assert scost == 8


def test_oracle_greedy_known_answer_alpha_0():
    plan, tcost, scost = O.greedy_plan(_kat_network(), alpha=0.0)
    assert plan == [["node_3", "node_4"], ["node_1", "node_2"], ["node_6", "node_7"], ["node_5", "node_8"]]
    assert tcost == pytest.approx(94.0)
    assert scost == 8


def test_greedy_tie_takes_the_first_pair_in_scan_order():
    """T3(c:4, d:2), T4(d:2), T5(a:2, e:5): at step 3, T4N7 and T5N7 both
    have E = -4.  The reference's tau -> 0 limit draws uniformly among tied
    pairs (StatsBase.sample with equal weights); this build takes the first
    pair in scan order -- a convention, not a parity claim."""
    spec = [("node_1", ["a", "b"], [2, 8]), ("node_2", ["b", "c"], [8, 4]), ("node_3", ["c", "d"], [4, 2]),
            ("node_4", ["d"], [2]), ("node_5", ["a", "e"], [2, 5])]
    plan, _, _ = O.greedy_plan(_kat_network(spec))
    assert plan[2] == ["node_4", "node_7"]
    tn = _kat_network(spec)
    assert PL.bgreedy_single(tn, 1.0, 0.0)[0] == plan
    assert PL.bgreedy_single(tn, 1.0, 0.0, exact_limit=0)[0] == plan


@pytest.mark.parametrize("mistake", ["flipped_alpha", "no_alpha", "outer_size"])
def test_known_answer_catches_planted_mistakes(monkeypatch, mistake):
    """A plausible mistake planted in the oracle's energy changes the
    known-answer plan."""
    real = O.contraction_cost

    def planted(A, B, alpha):
        e, ci, cd, costs = real(A, B, alpha)
        sA, sB = math.prod(A.dims), math.prod(B.dims)
        sC = math.prod(cd) if cd else 1
        if mistake == "flipped_alpha":
            e = sC + alpha * (sA + sB)
        elif mistake == "no_alpha":
            e = sC
        else:                                    # |C| taken as |A||B|
            e = sA * sB - alpha * (sA + sB)
        return e, ci, cd, costs

    monkeypatch.setattr(O, "contraction_cost", planted)
    plan, _, _ = O.greedy_plan(_kat_network(), alpha=1.0)
    assert plan != [["node_1", "node_2"], ["node_3", "node_6"], ["node_4", "node_7"], ["node_5", "node_8"]]


@pytest.mark.parametrize("alpha", [1.0, 0.0])
def test_product_greedy_known_answer(alpha):
    want = O.greedy_plan(_kat_network(), alpha=alpha)
    for limit in (160, 0):                        # the scan and the heap path
        plan, tcost, scost = PL.bgreedy_single(_kat_network(), alpha, 0.0, exact_limit=limit)
        assert plan == want[0]
        assert tcost == pytest.approx(want[1]) and scost == want[2]


def _grid(rows, cols, depth, seed, cfg):
    n = rows * cols
    tn = P.network_from_gates(n, C.grid_circuit_gates(rows, cols, depth, seed), backend=RecordingBackend(),
                              decompose=True)
    P.add_input(tn, "0" * n)
    P.add_output(tn, cfg)
    to = O.network_from_gates(n, O.grid_circuit_gates(rows, cols, depth, seed), decompose=True)
    O.add_input(to, "0" * n)
    O.add_output(to, cfg)
    return tn, to


def _qft(n, x):
    bits = "".join(str((x >> q) & 1) for q in range(n))
    tn = P.network_from_gates(n, C.qft_gates(n), backend=RecordingBackend())
    P.add_input(tn, bits)
    to = O.network_from_gates(n, O.qft_gates(n))
    O.add_input(to, bits)
    return tn, to


NETS = {
    "qft6": lambda: _qft(6, 37),
    "qft8": lambda: _qft(8, 201),
    "grid2x3_d5": lambda: _grid(2, 3, 5, 6, "0x1x01"),
    "grid3x3_d6": lambda: _grid(3, 3, 6, 21, "1x0x11x01"),
    "grid3x4_d6": lambda: _grid(3, 4, 6, 9, "0" * 12),
}


@pytest.mark.parametrize("name", sorted(NETS))
def test_product_greedy_equals_oracle(name):
    tn, to = NETS[name]()
    assert len(tn.nodes) <= 160
    want, wt, ws = O.greedy_plan(to)
    for limit in (160, 0):
        plan, t, s = PL.bgreedy_single(tn, 1.0, 0.0, exact_limit=limit)
        assert plan == want, (name, limit)
        assert t == pytest.approx(wt) and s == ws


def test_heap_path_takes_an_outer_product_when_it_is_cheapest():
    """Two disconnected pieces: a pair of large tensors sharing a small
    bond, and two scalars-to-be.  Their outer product (E = 2*2 - 4 = 0)
    beats every connected pair of the large piece, as the all-pairs scan
    finds; the heap path must pick it too."""
    spec = [("node_1", ["a", "b", "c"], [16, 16, 2]), ("node_2", ["c", "d", "e"], [2, 16, 16]),
            ("node_3", ["x"], [2]), ("node_4", ["y"], [2])]
    scan = PL.bgreedy_single(_kat_network(spec), 1.0, 0.0)
    heap = PL.bgreedy_single(_kat_network(spec), 1.0, 0.0, exact_limit=0)
    assert scan[0][0] == ["node_3", "node_4"]
    assert heap == scan
    assert O.greedy_plan(_kat_network(spec))[0] == scan[0]


@pytest.mark.parametrize("rows,cols,depth", [(4, 4, 8), (5, 5, 8)])
def test_heap_path_equals_scan_on_grids(rows, cols, depth):
    """Grid networks of 250-400 tensors: the heap path (the one bgreedy_single
    takes above 160 tensors) gives the scan's plan."""
    n = rows * cols
    tn = P.network_from_gates(n, C.grid_circuit_gates(rows, cols, depth, 2026), backend=RecordingBackend(),
                              decompose=True)
    P.add_input(tn, "0" * n)
    P.add_output(tn, "0" * (n - 3) + "xxx")
    heap = PL.bgreedy_single(tn, 1.0, 0.0, exact_limit=0)
    scan = PL.bgreedy_single(tn, 1.0, 0.0, exact_limit=10 ** 6)
    assert heap[0] == scan[0]
    assert heap[1] == pytest.approx(scan[1]) and heap[2] == scan[2]


def test_outer_energy_bound_is_a_lower_bound():
    from collections import Counter
    rng = np.random.default_rng(5)
    for _ in range(200):
        sizes = [int(2 ** rng.integers(0, 12)) for _ in range(int(rng.integers(2, 9)))]
        alpha = float(rng.choice([0.0, 0.5, 1.0, 2.0, 1e3]))
        brute = min(a * b - alpha * (a + b) for i, a in enumerate(sizes) for b in sizes[i + 1:])
        assert PL._outer_energy_bound(Counter(sizes), alpha) == pytest.approx(brute)
