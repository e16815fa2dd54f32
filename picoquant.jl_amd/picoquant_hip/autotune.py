"""Autotuner: per-step kernel class and MFMA tile for a compiled plan.

Each large contraction step of a lowered plan is timed (pq_time_step) with
every applicable option -- the MFMA GEMM with each workgroup tile, the
bandwidth kernel, the automatic choice -- and the fastest is kept.  Results
are cached by shape, so repeated shapes (the sweep absorbs many alike sites)
are timed once.
"""
from __future__ import annotations

from . import _lib
from .costmodel import merge
from .executor import MIXED_FP64_ELEMS, SMALL_ELEMS, assign_precision, lower_plan

TILES = {_lib.PQ_C64: list(range(15)), _lib.PQ_C128: list(range(11)) + [12, 14]}


def tune_step(ctx, M, N, K, batch, dtype, reps=3, cache=None, split16=False):
    """(kclass, tile, us) of the fastest option for one contraction shape;
    split16 adds the FP16 hi/lo GEMM to the options."""
    key = (tuple(merge(M)), tuple(merge(N)), tuple(merge(K)), batch, dtype, split16)
    if cache is not None and key in cache:
        return cache[key]
    options = [(-1, -1), (_lib.KC_STREAM, -1)] + [(_lib.KC_GEMM, t) for t in TILES[dtype]]
    if split16 and dtype == _lib.PQ_C64:
        options += [(_lib.KC_GEMM16, t) for t in TILES[dtype]]
    best = None
    for kc, tile in options:
        try:
            us = ctx.time_step(M, N, K, batch=batch, kclass=kc, tile=tile, reps=reps, dtype=dtype)
        except _lib.PQError:
            continue
        if best is None or us < best[2]:
            best = (kc, tile, us)
    if cache is not None:
        cache[key] = best
    return best


def autotune_plan(ctx, network, plan, sliced=(), group_bits=None, precision="mixed", reps=3,
                  min_elems=SMALL_ELEMS, cache=None, small_elems=MIXED_FP64_ELEMS):
    """Per contraction step of `plan`: (kc list, tile list, report).  Small
    steps (grouped launches) keep the automatic choice; small_elems is the
    plan's FP64 size limit (as in compile_plan)."""
    slots, steps, _ = lower_plan(network, plan, sliced)
    split16 = precision == "mixed16"
    if split16:
        precision = "mixed"
    code = _lib.PQ_C128 if (precision == "mixed" or ctx.code == _lib.PQ_C128) else _lib.PQ_C64
    dts, steps, slots = assign_precision(slots, steps, code, precision, small_elems=small_elems)
    n_sliced = len(list(sliced))
    gb = n_sliced if group_bits is None else group_bits
    cache = {} if cache is None else cache
    kcs, tiles, report = [], [], []
    for kind, sa, sb, sc, M, N, K in steps:
        if kind != 0:
            continue
        if max(slots[x].elems for x in (sa, sb, sc)) <= min_elems:
            kcs.append(-1)
            tiles.append(-1)
            continue
        batch = 1 << sum(1 for j in slots[sc].deps if j < gb)
        best = tune_step(ctx, M, N, K, batch, dts[sc], reps, cache, split16)
        if best is None:
            kcs.append(-1)
            tiles.append(-1)
            continue
        kcs.append(best[0])
        tiles.append(best[1])
        report.append(best)
    return kcs, tiles, report
