"""Learned cost model for pairwise contractions.

A 3-layer perceptron (16 -> 64 -> 64 -> 1, FP32) predicts log(microseconds)
of one contraction step from its shape features; it is trained on timings
that libpq_hip's own kernels produce (pq_time_step), with the library's HIP
forward / backward / Adam kernels (pq_mlp_*).  The plan selector sums the
predicted step times of a lowered plan to rank candidate orders and
slicings; the reference ranks plans by FLOP counts only
(src/layer2/bgreedy_contraction.jl:53-66).
"""
from __future__ import annotations

import ctypes as ct
import math

import numpy as np

from . import _lib
from .executor import SMALL_ELEMS, assign_precision, fuse_chains, lower_plan

N_FEATURES = 16
# Fused chains (PQ_KC_CHAIN) are priced at this rate of complex MACs: 8 real
# flops each at 0.44 of the FP32 MFMA peak, the kernel's measured rate on
# the 36-qubit depth-14 plan (0.42-0.45); the step model never sees them.
CHAIN_MACS_PER_US = 0.44 * 157.3e12 / 8 / 1e6


def merge(modes):
    """The runtime's mode merge (drop extent 1, fuse neighbours contiguous in
    both strides) so features describe what the kernel sees."""
    out = []
    for e, s0, s1 in modes:
        if e == 1:
            continue
        if out and s0 == out[-1][1] * out[-1][0] and s1 == out[-1][2] * out[-1][0]:
            out[-1] = (out[-1][0] * e, out[-1][1], out[-1][2])
        else:
            out.append((e, s0, s1))
    return out


def step_features(M, N, K, batch=1, dtype=_lib.PQ_C64):
    M, N, K = merge(M), merge(N), merge(K)
    Mt = math.prod(m[0] for m in M) if M else 1
    Nt = math.prod(m[0] for m in N) if N else 1
    Kt = math.prod(m[0] for m in K) if K else 1
    es = 16 if dtype == _lib.PQ_C128 else 8
    lm, ln, lk, lb = math.log2(Mt), math.log2(Nt), math.log2(Kt), math.log2(batch)
    return [
        lm, ln, lk, lb,
        lm + ln + lk + lb,
        math.log2(es * (Mt * Kt + Kt * Nt + Mt * Nt) * batch),
        len(M), len(N), len(K),
        1.0 if K and K[0][1] == 1 else 0.0,
        1.0 if K and K[0][2] == 1 else 0.0,
        1.0 if any(m[2] == 1 for m in M) else 0.0,
        math.log2(K[0][0]) if K else 0.0,
        1.0 if dtype == _lib.PQ_C128 else 0.0,
        min(lm, ln),
        max(lm, ln),
    ]


def plan_samples(network, plan, sliced=(), group_bits=None, precision="mixed", max_elems=1 << 27, fused=None):
    """(features, (M, N, K, batch, dtype)) for every contraction step of a
    lowered plan, with slice-dependent steps batched over 2^group_bits.
    fused: a list that receives the MAC count of every fused chain the
    executor would form (fuse_chains); those chains' steps are left out."""
    slots, steps, out_slot = lower_plan(network, plan, sliced)
    dts, steps, slots = assign_precision(slots, steps, _lib.PQ_C128, precision)
    nsl = len(list(sliced))
    gb = nsl if group_bits is None else group_bits
    skip = set()
    if fused is not None:
        _, progs, members = fuse_chains(slots, steps, dts, group_bits=gb)
        for prog, mem in zip(progs, members):
            if prog is None:
                continue
            skip |= set(mem)
            fused.append(sum(math.prod(m[0] for m in steps[j][4] + steps[j][5] + steps[j][6]) for j in mem))
    out = []
    for i, (kind, sa, sb, sc, M, N, K) in enumerate(steps):
        if kind != 0 or i in skip:
            continue
        if max(slots[x].elems for x in (sa, sb, sc)) > max_elems:
            continue
        batch = 1 << sum(1 for j in slots[sc].deps if j < gb)
        if batch * slots[sc].elems > 4 * max_elems:
            batch = max(1, (4 * max_elems) // slots[sc].elems)
        out.append((step_features(M, N, K, batch, dts[sc]), (M, N, K, batch, dts[sc])))
    return out


class MLPCostModel:
    """Perceptron on standardised step features, target log(us)."""

    def __init__(self, ctx, hidden=64, seed=0):
        self.ctx = ctx
        self.lib = ctx.lib
        self.hidden = hidden
        p = ct.c_void_p()
        ctx.check(self.lib.pq_mlp_create(ctx.ptr, N_FEATURES, hidden, seed, ct.byref(p)))
        self.ptr = p
        self.mu = np.zeros(N_FEATURES, dtype=np.float32)
        self.sd = np.ones(N_FEATURES, dtype=np.float32)
        self.trained = False

    def close(self):
        if getattr(self, "ptr", None):
            self.lib.pq_mlp_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def n_params(self):
        return self.lib.pq_mlp_num_params(self.ptr)

    def get_params(self):
        out = np.empty(self.n_params(), dtype=np.float32)
        self.ctx.check(self.lib.pq_mlp_get_params(self.ctx.ptr, self.ptr, out.ctypes.data_as(ct.POINTER(ct.c_float))))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self.ctx.check(self.lib.pq_mlp_set_params(self.ctx.ptr, self.ptr, p.ctypes.data_as(ct.POINTER(ct.c_float))))

    def _std(self, X):
        return np.ascontiguousarray((np.asarray(X, dtype=np.float32) - self.mu) / self.sd, dtype=np.float32)

    def loss_grad(self, Xs, y):
        """Loss and gradient on already-standardised inputs (gradient checks)."""
        Xs = np.ascontiguousarray(Xs, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = ct.c_float()
        g = np.empty(self.n_params(), dtype=np.float32)
        self.ctx.check(self.lib.pq_mlp_loss_grad(self.ctx.ptr, self.ptr, Xs.ctypes.data_as(ct.POINTER(ct.c_float)),
                                                 y.ctypes.data_as(ct.POINTER(ct.c_float)), len(y), ct.byref(loss),
                                                 g.ctypes.data_as(ct.POINTER(ct.c_float))))
        return loss.value, g

    def fit(self, X, us, epochs=150, batch=256, lr=2e-3, seed=0):
        X = np.asarray(X, dtype=np.float32)
        self.mu = X.mean(0)
        self.sd = X.std(0) + 1e-3
        Xs = self._std(X)
        y = np.ascontiguousarray(np.log(np.asarray(us, dtype=np.float64)), dtype=np.float32)
        loss = ct.c_float()
        self.ctx.check(self.lib.pq_mlp_train(self.ctx.ptr, self.ptr, Xs.ctypes.data_as(ct.POINTER(ct.c_float)),
                                             y.ctypes.data_as(ct.POINTER(ct.c_float)), len(y), epochs, batch, lr,
                                             seed, ct.byref(loss)))
        self.trained = True
        return loss.value

    def predict_log_us(self, X):
        Xs = self._std(X)
        out = np.empty(len(Xs), dtype=np.float32)
        self.ctx.check(self.lib.pq_mlp_predict(self.ctx.ptr, self.ptr, Xs.ctypes.data_as(ct.POINTER(ct.c_float)),
                                               len(Xs), out.ctypes.data_as(ct.POINTER(ct.c_float))))
        return out

    def predict_us(self, X):
        return np.exp(self.predict_log_us(X).astype(np.float64))

    def plan_time(self, network, plan, sliced=(), precision="mixed"):
        """Predicted microseconds of one planned run as compile_plan would
        run it: runs of steps the executor fuses into chains at the chain
        kernel's rate (CHAIN_MACS_PER_US), the other individually launched
        steps at their predicted times, slice-dependent steps once per slice
        group; small steps share grouped launches (one per dependency level),
        charged at their largest predicted step."""
        chains = []
        samples = plan_samples(network, plan, sliced, None, precision, max_elems=1 << 34, fused=chains)
        total = sum(chains) / CHAIN_MACS_PER_US
        if not samples:
            return total
        X = np.array([f for f, _ in samples], dtype=np.float32)
        t = self.predict_us(X)
        small = []
        for (f, (M, N, K, batch, dt)), ti in zip(samples, t):
            mt = 2 ** (f[0] + f[1])
            if mt <= SMALL_ELEMS and 2 ** f[2] <= 64:
                small.append(ti)
            else:
                total += ti
        if small:
            # the world lines of all sites advance together: about
            # len(small) / sites grouped launches; each costs about its largest step
            small.sort()
            groups = max(1, int(round(len(small) ** 0.5)))
            total += sum(small[-groups:])
        return total


def roofline_us(f):
    """Lower-bound microseconds of a step from its features: FP32 MFMA and HBM
    peaks of MI355X (157 TFLOP/s, 8 TB/s)."""
    return max(8 * 2.0 ** f[4] / 157e12, 2.0 ** f[5] / 8e12) * 1e6


def shape_pool(network_plans, batches=(1,), max_est_us=5000.0):
    """Unique contraction shapes of the given [(network, plan, sliced)], each
    also at the extra slice-group batch sizes for slice-dependent steps;
    shapes whose roofline time exceeds max_est_us are left out (no plan worth
    running contains them, and timing them would dominate collection)."""
    pool = {}
    for net, plan, sliced in network_plans:
        for gb in sorted({len(list(sliced))} | {b for b in batches if b <= len(list(sliced))}):
            for f, args in plan_samples(net, plan, sliced, gb):
                if roofline_us(f) > max_est_us:
                    continue
                M, N, K, batch, dt = args
                key = (tuple(merge(M)), tuple(merge(N)), tuple(merge(K)), batch, dt)
                pool.setdefault(key, (f, args))
    return list(pool.values())


def collect_dataset(ctx, network_plans, n_samples=20000, seed=0, reps=2, big_weight=10.0, progress=None,
                    batches=(0, 4, 8), max_est_us=2000.0):
    """Time contraction shapes drawn from [(network, plan, sliced)] with
    pq_time_step: unique shapes, sampled with replacement, large steps
    (>= 2^16 elements in some operand) weighted big_weight.  Returns
    (X, us, keys) -- keys index the unique shape of each sample."""
    rng = np.random.default_rng(seed)
    pool = shape_pool(network_plans, batches, max_est_us)
    if not pool:
        raise ValueError("no samples")
    w = np.array([big_weight if max(f[0] + f[2], f[1] + f[2], f[0] + f[1]) >= 16 else 1.0 for f, _ in pool])
    idx = rng.choice(len(pool), size=n_samples, replace=True, p=w / w.sum())
    X, us, keys = [], [], []
    for c, i in enumerate(idx):
        f, (M, N, K, batch, dt) = pool[i]
        try:
            t = ctx.time_step(M, N, K, batch=batch, reps=reps, dtype=dt)
        except _lib.PQError:
            continue
        X.append(f)
        us.append(max(t, 1e-2))
        keys.append(int(i))
        if progress and c % 5000 == 0:
            progress(c, len(idx))
    return np.array(X, dtype=np.float32), np.array(us), np.array(keys)


def grid_training_plans(specs=((4, 5, 10, 11), (5, 5, 12, 12), (5, 6, 12, 13), (6, 6, 14, 14)),
                        per_circuit=16, slicings=(0, 6, 10), max_elems=1 << 27, greedy=2):
    """[(network, plan, sliced)] for the training set: candidate plans
    (sweeps, perturbed sweeps, Boltzmann-greedy samples) of random grid
    circuits, each with several slicings.  Networks carry shapes only."""
    from . import circuits as C, learned, planner as PL
    from .backend import AbstractBackend, Metrics
    from .network import add_input, add_output, network_from_gates

    class _Shapes(AbstractBackend):
        def __init__(self):
            self.metrics = Metrics()

        def save_tensor_data(self, label, data):
            pass

    out = []
    for (r, c, d, s) in specs:
        n = r * c
        tn = network_from_gates(n, C.grid_circuit_gates(r, c, d, s), backend=_Shapes(), decompose=True)
        add_input(tn, "0" * n)
        oq = [rr * c + c for rr in range(r)]
        add_output(tn, "".join("x" if (q + 1) in oq else "0" for q in range(n)))
        for name, plan in learned.candidate_plans(tn, r, c, oq, per_circuit, seed=s, n_greedy=greedy):
            if PL.plan_ledger(tn, plan)["max_size"] > max_elems:
                continue
            for nsl in slicings:
                try:
                    sl = PL.choose_slices(tn, plan, nsl)[0] if nsl else []
                except ValueError:
                    continue
                out.append((tn, plan, sl))
    return out
