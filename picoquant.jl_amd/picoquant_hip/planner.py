"""Contraction planners producing the reference's pair plans [[A, B], ...]
(consumed by contract_network!, src/layer2.jl:294-321).

  bgreedy        -- Boltzmann-greedy with the Gray-Kourtis cost |C| - alpha(|A|+|B|)
                    (src/layer2/bgreedy_contraction.jl:17-131).  tau = 0 is the
                    deterministic limit (argmin over all pairs, first pair in
                    scan order on ties, where the reference's limit would draw
                    uniformly among the tied pairs); tau > 0 samples through the
                    Gumbel-max form of the same Boltzmann weights, which cannot
                    overflow (the reference's exp(-E/tau) does at 2^30-element
                    tensors).
  sweep_plan     -- world lines first (each qubit's tensors contracted along
                    time into one site tensor), then the sites absorbed one by
                    one into a growing boundary in a given site order.
  reconfigure    -- subtree reconfiguration: every window of up to W subtrees
                    of a plan's contraction tree is re-ordered optimally
                    (pq_netcon, the exact search behind netcon) when that
                    lowers the MAC count.
  plan_ledger    -- exact per-step shapes, MACs and bytes of a plan.
"""
from __future__ import annotations

import heapq
import math
from collections import OrderedDict, defaultdict

import numpy as np

from .contraction import sort_indices


def _pair_cost(A_idx, A_dims, B_idx, B_dims, alpha):
    bset = set(B_idx)
    aset = set(A_idx)
    sc = 1
    for x, d in zip(A_idx, A_dims):
        if x not in bset:
            sc *= d
    for x, d in zip(B_idx, B_dims):
        if x not in aset:
            sc *= d
    sa = math.prod(A_dims)
    sb = math.prod(B_dims)
    return sc - alpha * (sa + sb), sa, sb, sc


def bgreedy_single(network, alpha=1.0, tau=0.0, rng=None, exact_limit=160):
    """One bgreedy plan (src/layer2/bgreedy_contraction.jl:17-45).  Returns
    (plan, time_cost, space_cost) with time_cost = sum sqrt(|A||B||C|)."""
    tensors = OrderedDict((k, (list(v.indices), list(v.dims))) for k, v in network.nodes.items())
    counter = network.counters["node"]
    if len(tensors) <= exact_limit:
        return _bgreedy_scan(tensors, counter, alpha, tau, rng)
    return _bgreedy_heap(tensors, counter, alpha, tau, rng)


def _gumbel(rng):
    return -math.log(-math.log(max(rng.random(), 1e-300)))


def _merge(A, B):
    aset, bset = set(A[0]), set(B[0])
    idx = [x for x in A[0] if x not in bset] + [x for x in B[0] if x not in aset]
    d = dict(zip(A[0] + B[0], A[1] + B[1]))
    return idx, [d[x] for x in idx]


def _bgreedy_scan(tensors, counter, alpha, tau, rng):
    plan, tcost, scost = [], 0.0, 0
    while len(tensors) > 1:
        items = list(tensors.items())
        best = None
        for i in range(len(items)):
            ai, ad = items[i][1]
            for j in range(i + 1, len(items)):
                bi, bd = items[j][1]
                e, sa, sb, sc = _pair_cost(ai, ad, bi, bd, alpha)
                key = e if tau <= 0 else e / tau - _gumbel(rng)
                if best is None or key < best[0]:
                    best = (key, items[i][0], items[j][0], sa, sb, sc)
        _, a, b, sa, sb, sc = best
        counter += 1
        c = f"node_{counter}"
        C = _merge(tensors.pop(a), tensors.pop(b))
        tensors[c] = C
        plan.append([a, b])
        tcost += math.sqrt(sa * sb * sc)
        scost = max(scost, sc)
    return plan, tcost, scost


def _outer_energy_bound(sizes, alpha):
    """A lower bound on |A||B| - alpha(|A| + |B|) over all pairs of distinct
    tensors whose sizes are the multiset `sizes` (Counter size -> count):
    the energy of any pair that shares no index.  g(a, b) = ab - alpha(a + b)
    is linear in each argument, so its minimum over pairs lies at a pair of
    extremes: the two smallest, the smallest and the largest, or the two
    largest sizes."""
    ks = sorted(sizes)
    s1, sn = ks[0], ks[-1]
    s2 = s1 if sizes[s1] > 1 else ks[1]
    sm = sn if sizes[sn] > 1 else ks[-2]

    def g(a, b):
        return a * b - alpha * (a + b)

    return min(g(s1, s2), g(s1, sn), g(sm, sn))


def _bgreedy_heap(tensors, counter, alpha, tau, rng):
    """The selection of _bgreedy_scan without the O(N^2) scan per step.
    Pairs that share an index wait in a heap keyed by (energy, position of
    A, position of B) -- positions follow creation order, so ties break as
    the scan's first pair does.  tau = 0: the heap's top is the all-pairs
    argmin whenever its energy is below _outer_energy_bound, which bounds
    every pair that shares nothing; otherwise (on grid circuits: at most
    once per plan) that step scans all pairs.  tau > 0: samples among the
    pairs that share an index (Gumbel-max over the heap), not among all
    pairs."""
    order = {k: i for i, k in enumerate(tensors)}
    ix2t = defaultdict(set)
    sizes = defaultdict(int)
    for k, (idx, dims) in tensors.items():
        sizes[math.prod(dims)] += 1
        for x in idx:
            ix2t[x].add(k)
    heap = []

    def push(a, b):
        if order[a] > order[b]:
            a, b = b, a
        e, _, _, _ = _pair_cost(*tensors[a], *tensors[b], alpha)
        key = e if tau <= 0 else e / tau - _gumbel(rng)
        heapq.heappush(heap, (key, order[a], order[b], a, b))

    for x, s in ix2t.items():
        if len(s) == 2:
            push(*sorted(s, key=lambda k: order[k]))
    plan, tcost, scost = [], 0.0, 0
    nxt = len(order)
    while len(tensors) > 1:
        while heap and (heap[0][3] not in tensors or heap[0][4] not in tensors):
            heapq.heappop(heap)
        if heap and (tau > 0 or heap[0][0] < _outer_energy_bound(sizes, alpha)):
            _, _, _, a, b = heapq.heappop(heap)
        elif tau <= 0:
            a, b = _scan_argmin(tensors, alpha)
        else:
            it = iter(tensors)
            a, b = next(it), next(it)
        A, B = tensors.pop(a), tensors.pop(b)
        sa, sb = math.prod(A[1]), math.prod(B[1])
        C = _merge(A, B)
        sc = math.prod(C[1])
        for v in (sa, sb):
            sizes[v] -= 1
            if not sizes[v]:
                del sizes[v]
        sizes[sc] += 1
        counter += 1
        c = f"node_{counter}"
        for x in A[0]:
            ix2t[x].discard(a)
        for x in B[0]:
            ix2t[x].discard(b)
        tensors[c] = C
        order[c] = nxt
        nxt += 1
        nb = set()
        for x in C[0]:
            ix2t[x].add(c)
            nb |= ix2t[x]
        nb.discard(c)
        # creation order: with tau > 0 each push draws from rng, so the
        # order must not depend on set iteration (string hashing varies
        # between processes)
        for o in sorted(nb, key=order.__getitem__):
            push(o, c)
        plan.append([a, b])
        tcost += math.sqrt(sa * sb * sc)
        scost = max(scost, sc)
    return plan, tcost, scost


def _scan_argmin(tensors, alpha):
    """(A, B): the minimum-energy pair of all, first in scan order on ties."""
    items = list(tensors.items())
    best = None
    for i in range(len(items)):
        ai, ad = items[i][1]
        for j in range(i + 1, len(items)):
            e = _pair_cost(ai, ad, *items[j][1], alpha)[0]
            if best is None or e < best[0]:
                best = (e, items[i][0], items[j][0])
    return best[1], best[2]


def bgreedy(network, alpha=1.0, tau=1.0, N=10, rng=None):
    """src/layer2/bgreedy_contraction.jl:53-66: best of N samples by time cost."""
    rng = rng if rng is not None else np.random.default_rng()
    best = None
    for _ in range(N):
        plan, t, _ = bgreedy_single(network, alpha, tau, rng)
        if best is None or t < best[1]:
            best = (plan, t)
    return best


# --------------------------------------------------------------------------- #
# World-line / sweep planner
# --------------------------------------------------------------------------- #


def world_lines(network):
    """Each qubit's nodes in time order (input, gates, output), following the
    qubit's edge chain.  Requires every node to sit on a single qubit line
    (two-qubit gates added with decompose=True)."""
    lines = {}
    owner = {}
    for q in range(1, network.number_qubits + 1):
        e_lab = network.input_qubits[q - 1]
        seq = []
        e = network.edges[e_lab]
        if e.src is not None:
            seq.append(e.src)
        while True:
            dst = network.edges[e_lab].dst
            if dst is None:
                break
            seq.append(dst)
            nxt = None
            for x in network.nodes[dst].indices:
                ex = network.edges[x]
                if ex.qubit == q and ex.src == dst and not ex.virtual:
                    nxt = x
                    break
            if nxt is None:
                break
            e_lab = nxt
        for k in seq:
            if k in owner:
                raise ValueError(f"node {k} sits on qubits {owner[k]} and {q}: use decompose=True")
            owner[k] = q
        lines[q] = seq
    missing = set(network.nodes) - set(owner)
    if missing:
        raise ValueError(f"nodes not on any qubit line: {sorted(missing)[:5]}")
    return lines


def sweep_plan(network, site_order):
    """Pair plan: contract each world line into a site tensor (in time order),
    then absorb the sites into one boundary tensor in `site_order` (1-based
    qubits)."""
    lines = world_lines(network)
    counter = network.counters["node"]
    plan = []
    site = {}
    for q in sorted(lines):
        seq = lines[q]
        cur = seq[0]
        for k in seq[1:]:
            plan.append([cur, k])
            counter += 1
            cur = f"node_{counter}"
        site[q] = cur
    cur = site[site_order[0]]
    for q in site_order[1:]:
        plan.append([cur, site[q]])
        counter += 1
        cur = f"node_{counter}"
    return plan


# --------------------------------------------------------------------------- #
# Subtree reconfiguration
# --------------------------------------------------------------------------- #


class _Tree:
    """Contraction-tree node: a leaf (network node label) or a pair.  mask =
    open indices (bitset over index ids); cost = MACs of the subtree."""

    __slots__ = ("left", "right", "label", "mask", "cost")

    def __init__(self, left=None, right=None, label=None, mask=0, sizer=None):
        self.left, self.right, self.label = left, right, label
        if label is not None:
            self.mask, self.cost = mask, 0
        else:
            self.mask = left.mask ^ right.mask
            self.cost = left.cost + right.cost + sizer(left.mask | right.mask)


def _bits(mask):
    out = []
    while mask:
        low = mask & -mask
        out.append(low.bit_length() - 1)
        mask ^= low
    return out


def reconfigure(network, plan, windows=(12, 16), rounds=12, max_work=1 << 24):
    """A pair plan computing the same network with no more MACs.

    The plan's contraction tree is walked from the root; at each pair node
    the subtree is cut into a frontier of at most W subtrees (expanding the
    one with the largest result first), the frontier's optimal contraction
    order is found with pq_netcon, and it replaces the node's top of the
    tree when it costs fewer MACs.  Passes repeat until one improves
    nothing (at most `rounds` per window size), once per W in `windows`.
    The MAC count is the reference's Metrics.flops for the plan.  Windows
    whose exact search exceeds max_work pair evaluations are left as they
    are."""
    from .optimal_order import optimal_tree
    from ._lib import PQError

    ids, ext = {}, {}
    leaves = {}
    for lab, node in network.nodes.items():
        m = 0
        for x, d in zip(node.indices, node.dims):
            if x not in ids:
                ids[x] = len(ids)
                ext[ids[x]] = int(d)
            m |= 1 << ids[x]
        leaves[lab] = m

    def size(mask):
        s = 1
        for b in _bits(mask):
            s *= ext[b]
        return s

    cur = {lab: _Tree(label=lab, mask=m) for lab, m in leaves.items()}
    counter = network.counters["node"]
    for a, b in plan:
        t = _Tree(cur.pop(a), cur.pop(b), sizer=size)
        counter += 1
        cur[f"node_{counter}"] = t
    if len(cur) != 1:
        raise ValueError(f"plan leaves {len(cur)} tensors uncontracted")
    (root,) = cur.values()

    def window(t, W):
        frontier = [t.left, t.right]
        while len(frontier) < W:
            inner = [i for i, f in enumerate(frontier) if f.label is None]
            if not inner:
                break
            i = max(inner, key=lambda j: (size(frontier[j].mask), -j))
            f = frontier.pop(i)
            frontier += [f.left, f.right]
        return frontier

    def improve(t, W):
        frontier = window(t, W)
        old = t.cost - sum(f.cost for f in frontier)
        idx = [_bits(f.mask) for f in frontier]
        try:
            tree, new = optimal_tree(idx, {b: ext[b] for bs in idx for b in bs}, max_work)
        except PQError:
            return t, False
        if not new < old * (1 - 1e-12):
            return t, False

        def build(x):
            if isinstance(x, int):
                return frontier[x]
            return _Tree(build(x[0]), build(x[1]), sizer=size)
        return build(tree), True

    def sweep(t, W):
        changed = 0

        # pre-order rebuild: improve a node, then its children
        def rec(t):
            nonlocal changed
            if t.label is not None:
                return t
            t, ok = improve(t, W)
            changed += ok
            return _Tree(rec(t.left), rec(t.right), sizer=size)
        return rec(t), changed

    import sys
    limit = sys.getrecursionlimit()
    sys.setrecursionlimit(max(limit, 20 * len(leaves) + 1000))
    try:
        for W in windows:
            for _ in range(rounds):
                root, changed = sweep(root, W)
                if not changed:
                    break
    finally:
        sys.setrecursionlimit(limit)

    out = []
    counter = network.counters["node"]

    def emit(t):
        nonlocal counter
        if t.label is not None:
            return t.label
        a = emit(t.left)
        b = emit(t.right)
        out.append([a, b])
        counter += 1
        return f"node_{counter}"
    emit(root)
    return out


def _bisect(nodes, adj, imbalance, rng):
    """Split `nodes` into two parts of balanced size (each within
    (1/2 +- imbalance) of the total) with a small cut: the sign of the
    Fiedler vector of the bond-weighted Laplacian (weight log2 of the bond
    dimension) split at the median, then Kernighan-Lin passes moving single
    boundary nodes while the cut shrinks and the balance holds."""
    n = len(nodes)
    pos = {v: i for i, v in enumerate(nodes)}
    L = np.zeros((n, n))
    for v in nodes:
        i = pos[v]
        for u, w in adj[v].items():
            j = pos.get(u)
            if j is not None:
                L[i, j] -= w
                L[i, i] += w
    L += 1e-9 * np.eye(n)
    _, vecs = np.linalg.eigh(L)
    f = vecs[:, 1] + 1e-12 * rng.standard_normal(n)
    order = np.argsort(f)
    side = np.zeros(n, dtype=bool)
    side[order[n // 2:]] = True
    lo, hi = math.ceil((0.5 - imbalance) * n), math.floor((0.5 + imbalance) * n)

    def gain(i):
        # cut decrease when node i changes side
        g = 0.0
        for u, w in adj[nodes[i]].items():
            j = pos.get(u)
            if j is not None:
                g += w if side[j] != side[i] else -w
        return g
    for _ in range(8):
        moved = False
        for i in sorted(range(n), key=lambda i: -gain(i)):
            g = gain(i)
            if g <= 1e-12:
                continue
            cnt = int(side.sum())
            new = cnt - 1 if side[i] else cnt + 1
            if lo <= new <= hi:
                side[i] = not side[i]
                moved = True
        if not moved:
            break
    a = [nodes[i] for i in range(n) if not side[i]]
    b = [nodes[i] for i in range(n) if side[i]]
    return a, b


def partition_plan(network, imbalance=0.1, leaf=20, max_work=1 << 22, seed=0):
    """Top-down plan: the network is bisected (balanced parts with a small
    bond-weighted cut, _bisect), each part planned recursively, and parts of
    fewer than `leaf` tensors contracted in their optimal order (pq_netcon);
    a part whose exact search exceeds max_work is ordered greedily.  The
    idea of the reference's KaHyPar_contraction_plan
    (src/layer2/KaHyPar_contraction.jl:31-93: k-way partition, netcon below
    V nodes), with a spectral / Kernighan-Lin bisection in place of the
    KaHyPar library (absent here; that file is not part of the reference's
    package).  Returns a pair plan."""
    from .optimal_order import optimal_tree
    from ._lib import PQError

    rng = np.random.default_rng(seed)
    ids, ext, leaves = {}, {}, {}
    for lab, node in network.nodes.items():
        m = 0
        for x, d in zip(node.indices, node.dims):
            if x not in ids:
                ids[x] = len(ids)
                ext[ids[x]] = int(d)
            m |= 1 << ids[x]
        leaves[lab] = m
    owners = {}
    for lab, node in network.nodes.items():
        for x in node.indices:
            owners.setdefault(x, []).append(lab)
    adj = {lab: {} for lab in network.nodes}
    for x, labs in owners.items():
        if len(labs) == 2:
            a, b = labs
            w = math.log2(max(ext[ids[x]], 2))
            adj[a][b] = adj[a].get(b, 0.0) + w
            adj[b][a] = adj[b].get(a, 0.0) + w

    def size(mask):
        s = 1
        for bb in _bits(mask):
            s *= ext[bb]
        return s

    def greedy(ts):
        ts = list(ts)
        while len(ts) > 1:
            best = None
            for i in range(len(ts)):
                for j in range(i + 1, len(ts)):
                    if ts[i].mask & ts[j].mask:
                        c = size(ts[i].mask ^ ts[j].mask)
                        if best is None or c < best[0]:
                            best = (c, i, j)
            if best is None:   # disconnected: outer product of the two smallest
                order = sorted(range(len(ts)), key=lambda k: size(ts[k].mask))
                best = (0, min(order[:2]), max(order[:2]))
            _, i, j = best
            t = _Tree(ts[i], ts[j], sizer=size)
            ts = [x for k, x in enumerate(ts) if k not in (i, j)] + [t]
        return ts[0]

    def solve(part):
        ts = [_Tree(label=lab, mask=leaves[lab]) for lab in part]
        if len(ts) == 1:
            return ts[0]
        idx = [_bits(t.mask) for t in ts]
        try:
            tree, _ = optimal_tree(idx, {bb: ext[bb] for bs in idx for bb in bs}, max_work)
        except PQError:
            return greedy(ts)

        def build(x):
            if isinstance(x, int):
                return ts[x]
            return _Tree(build(x[0]), build(x[1]), sizer=size)
        return build(tree)

    def rec(part):
        if len(part) < leaf:
            return solve(part)
        a, b = _bisect(part, adj, imbalance, rng)
        if not a or not b:
            return solve(part)
        return _Tree(rec(a), rec(b), sizer=size)

    root = rec(list(network.nodes))
    out = []
    counter = network.counters["node"]

    def emit(t):
        nonlocal counter
        if t.label is not None:
            return t.label
        x = emit(t.left)
        y = emit(t.right)
        out.append([x, y])
        counter += 1
        return f"node_{counter}"
    emit(root)
    return out


def grid_site_order(rows, cols, axis="col", open_last=()):
    """Snake order over grid sites (1-based qubits, q = r*cols + c + 1); sites
    listed in open_last go to the end (keeps open output legs out of the
    boundary until the last steps)."""
    seq = []
    if axis == "col":
        for c in range(cols):
            rs = range(rows) if c % 2 == 0 else range(rows - 1, -1, -1)
            seq += [r * cols + c + 1 for r in rs]
    else:
        for r in range(rows):
            cs = range(cols) if r % 2 == 0 else range(cols - 1, -1, -1)
            seq += [r * cols + c + 1 for c in cs]
    ol = [q for q in seq if q in set(open_last)]
    return [q for q in seq if q not in set(open_last)] + ol


# --------------------------------------------------------------------------- #
# Ledger
# --------------------------------------------------------------------------- #


def plan_ledger(network, plan):
    """Per-step (|A|, |B|, |C|, K) for a pair plan, without touching data,
    plus totals: MACs (the reference's Metrics.flops), max |C|."""
    nodes = {k: (list(v.indices), list(v.dims)) for k, v in network.nodes.items()}
    counter = network.counters["node"]
    steps = []
    for a, b in plan:
        A, B = nodes.pop(a), nodes.pop(b)
        aset, bset = set(A[0]), set(B[0])
        dA = dict(zip(A[0], A[1]))
        K = math.prod(d for x, d in dA.items() if x in bset)
        C = _merge(A, B)
        counter += 1
        nodes[f"node_{counter}"] = C
        steps.append((math.prod(A[1]), math.prod(B[1]), math.prod(C[1]), K))
    macs = sum(c * k for _, _, c, k in steps)
    return {"steps": steps, "macs": macs, "max_size": max((s[2] for s in steps), default=0)}


# --------------------------------------------------------------------------- #
# Slicing choice
# --------------------------------------------------------------------------- #


def _plan_masks(network, plan):
    """Per step: bitmasks (over index ids) of C's indices, the summed indices
    and every index in the step's subtree, plus log2 extents per index."""
    ids = {}
    ext = {}

    def iid(x):
        if x not in ids:
            ids[x] = len(ids)
        return ids[x]

    tens = {}
    for k, v in network.nodes.items():
        mask = 0
        for x, d in zip(v.indices, v.dims):
            mask |= 1 << iid(x)
            ext[ids[x]] = d
        tens[k] = (mask, mask)        # (open indices, subtree indices)
    counter = network.counters["node"]
    steps = []
    for a, b in plan:
        (ma, sa), (mb, sb) = tens.pop(a), tens.pop(b)
        k = ma & mb
        c = ma ^ mb
        counter += 1
        tens[f"node_{counter}"] = (c, sa | sb)
        steps.append((c, k, sa | sb, ma, mb))
    return steps, ids, ext


def _mask_size(mask, ext):
    s = 1
    while mask:
        low = mask & -mask
        s *= ext[low.bit_length() - 1]
        mask ^= low
    return s


# Roofline time model used to rank slicings (seconds): complex64 on MI355X
# with effective rates below the datasheet peaks, plus a per-launch cost.
ROOFLINE = {"flops": 100e12, "bytes": 4e12, "launch": 4e-6, "esize": 8}


def sliced_cost(steps, ext, smask, nslices, model="macs"):
    """Total cost over all slices and the largest per-slice intermediate.  A
    step runs once per value of the sliced bonds in its subtree (its result
    depends on those only; steps without any run once).  model="macs":
    complex multiply-adds; model="time": roofline seconds (ROOFLINE).
    nslices is kept for the callers' signature (= 2^popcount(smask))."""
    total = 0
    peak = 0
    R = ROOFLINE
    for c, k, sub, ma, mb in steps:
        size_c = _mask_size(c & ~smask, ext)
        size_k = _mask_size(k & ~smask, ext)
        mult = 1 << bin(sub & smask).count("1")
        if model == "macs":
            total += size_c * size_k * mult
        else:
            size_a = _mask_size(ma & ~smask, ext)
            size_b = _mask_size(mb & ~smask, ext)
            t = max(8.0 * size_c * size_k / R["flops"], R["esize"] * (size_a + size_b + size_c) / R["bytes"])
            total += t * mult + R["launch"]
        peak = max(peak, size_c)
    return total, peak


def _bond_matrices(steps, ext, cand):
    """Boolean [steps x candidates] incidence of candidate bonds in C, K, A,
    B and the subtree of each step, plus log2 sizes of C, K, A, B."""
    ns, nc = len(steps), len(cand)
    inC = np.zeros((ns, nc), dtype=bool)
    inK = np.zeros((ns, nc), dtype=bool)
    inA = np.zeros((ns, nc), dtype=bool)
    inB = np.zeros((ns, nc), dtype=bool)
    inS = np.zeros((ns, nc), dtype=bool)
    lg = {j: math.log2(ext[j]) for j in ext}
    lc = np.zeros(ns)
    lk = np.zeros(ns)
    la = np.zeros(ns)
    lb = np.zeros(ns)

    def lsize(mask):
        t = 0.0
        while mask:
            low = mask & -mask
            t += lg[low.bit_length() - 1]
            mask ^= low
        return t

    for i, (c, k, sub, ma, mb) in enumerate(steps):
        lc[i], lk[i], la[i], lb[i] = lsize(c), lsize(k), lsize(ma), lsize(mb)
        for q, j in enumerate(cand):
            bit = 1 << j
            inC[i, q] = bool(c & bit)
            inK[i, q] = bool(k & bit)
            inA[i, q] = bool(ma & bit)
            inB[i, q] = bool(mb & bit)
            inS[i, q] = bool(sub & bit)
    return inC, inK, inA, inB, inS, lc, lk, la, lb


def _sliced_costs(mats, chosen, nslices, model):
    """Cost of adding each candidate to the chosen set (vector over candidates)."""
    inC, inK, inA, inB, inS, lc, lk, la, lb = mats
    ch = np.zeros(inC.shape[1], dtype=np.int64)
    ch[chosen] = 1
    bc = inC @ ch
    bk = inK @ ch
    ba = inA @ ch
    bb = inB @ ch
    bs = inS @ ch
    # every candidate bond has extent 2: each sliced bond halves a tensor
    c = lc[:, None] - bc[:, None] - inC
    k = lk[:, None] - bk[:, None] - inK
    # a step runs once per value of the sliced bonds in its subtree
    mult = np.exp2(bs[:, None] + inS)
    if model == "macs":
        cost = (np.exp2(c + k) * mult).sum(0)
    else:
        R = ROOFLINE
        a = la[:, None] - ba[:, None] - inA
        b = lb[:, None] - bb[:, None] - inB
        t = np.maximum(8.0 * np.exp2(c + k) / R["flops"],
                       R["esize"] * (np.exp2(a) + np.exp2(b) + np.exp2(c)) / R["bytes"])
        cost = (t * mult).sum(0) + R["launch"] * len(lc)
    peak = np.exp2(c.max(0))
    return cost, peak


def _bond_pool(network, steps, ids, ext, exclude, candidates):
    """Extent-2 bonds worth slicing: those of the largest steps first (open
    outputs and `exclude` left out), at most `candidates` of them."""
    inv = {v: k for k, v in ids.items()}
    open_out = set()
    for k, e in network.edges.items():
        if e.src is None or e.dst is None:
            open_out.add(k)
    excl = set(exclude) | open_out
    order = sorted(range(len(steps)), key=lambda i: -_mask_size(steps[i][0], ext) * _mask_size(steps[i][1], ext))
    cand = []
    seen = set()
    for i in order:
        c, k = steps[i][0], steps[i][1]
        m = c | k
        while m:
            low = m & -m
            j = low.bit_length() - 1
            m ^= low
            if j in seen:
                continue
            seen.add(j)
            if ext[j] == 2 and inv[j] not in excl:
                cand.append(j)
        if len(cand) >= candidates:
            break
    return cand, inv


def choose_slices(network, plan, n_bonds, max_elems=2 ** 28, exclude=(), candidates=200, model="time"):
    """Greedy choice of `n_bonds` extent-2 bonds to slice: each round adds the
    bond that minimises the total sliced cost (steps that do not depend on a
    sliced bond run once, the others once per slice), preferring choices that
    meet the per-slice size cap.  model: "time" (roofline seconds) or "macs".
    Returns (bonds, total_macs, peak)."""
    steps, ids, ext = _plan_masks(network, plan)
    if n_bonds == 0:
        tot, peak = sliced_cost(steps, ext, 0, 1, "macs")
        return [], tot, peak
    cand, inv = _bond_pool(network, steps, ids, ext, exclude, candidates)
    if len(cand) < n_bonds:
        raise ValueError("not enough bonds to slice")
    mats = _bond_matrices(steps, ext, cand)
    chosen = []
    for r in range(n_bonds):
        cost, peak = _sliced_costs(mats, chosen, 2 ** (r + 1), model)
        cost = cost.astype(float)
        cost[chosen] = np.inf
        over = peak > max_elems
        key = np.where(over, np.inf, cost)
        if np.all(np.isinf(key)):
            # no bond meets the cap yet: lower the largest per-slice tensor
            # first (a bond that misses it cannot help later rounds either),
            # the cheapest such bond on ties
            pk = peak.astype(float)
            pk[chosen] = np.inf
            key = np.where(pk <= pk.min(), cost, np.inf)
        chosen.append(int(np.argmin(key)))
    smask = 0
    for q in chosen:
        smask |= 1 << cand[q]
    tot, peak = sliced_cost(steps, ext, smask, 2 ** n_bonds, "macs")
    return [inv[cand[q]] for q in chosen], tot, peak


def slicing_candidates(network, plan, n_bonds, max_elems=2 ** 28, n_random=24, seed=0, pool=40):
    """[(bonds, sliced_macs, peak)]: the greedy slicings under the roofline
    and MAC objectives, plus random exchanges of one or two bonds of the
    roofline-greedy set for bonds of the same pool, keeping only sets whose
    per-slice peak is within max_elems (no duplicates)."""
    out, seen = [], set()

    def add(bonds, tot, peak):
        key = frozenset(bonds)
        if key not in seen and peak <= max_elems:
            seen.add(key)
            out.append((list(bonds), tot, peak))

    base = None
    for model in ("time", "macs"):
        try:
            b, tot, peak = choose_slices(network, plan, n_bonds, max_elems, model=model)
        except ValueError:
            continue
        base = base or b
        add(b, tot, peak)
    if base is None or n_bonds == 0:
        return out
    steps, ids, ext = _plan_masks(network, plan)
    cand, inv = _bond_pool(network, steps, ids, ext, (), pool)
    rng = np.random.default_rng(seed)
    names = [inv[j] for j in cand]
    tries = 0
    while len(out) < n_random + 2 and tries < 20 * n_random:
        tries += 1
        b = list(base)
        for _ in range(1 + int(rng.integers(2))):
            free = [x for x in names if x not in b]
            if not free:
                break
            b[int(rng.integers(len(b)))] = free[int(rng.integers(len(free)))]
        smask = 0
        for x in b:
            smask |= 1 << ids[x]
        tot, peak = sliced_cost(steps, ext, smask, 2 ** n_bonds, "macs")
        add(b, tot, peak)
    return out
