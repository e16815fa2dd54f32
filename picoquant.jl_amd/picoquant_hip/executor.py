"""Planned execution of a whole contraction (all slices, FP64 slice sum) on
libpq_hip's plan executor (pq_plan_* in include/pq_hip.h).

compile_plan() lowers (network, pair plan, sliced bonds) to slots and steps:
  * leaves are the network's device tensors; a sliced bond stays in the leaf
    and is selected per slice through its stride;
  * small intermediates keep the sliced bonds of their operands as ordinary
    indices (computed once, for both values of each bond) when that keeps
    them small; the first large step that reads them selects the slice
    through the bond's stride, as for leaves;
  * an intermediate is laid out with the indices its consumer sums first
    (so GEMM operand loads run along K), the rest after; the final tensor is
    laid out in output-qubit order, which is the reference's final permute
    (src/layer2.jl:306-316);
  * a slot that depends on a sliced bond is batched over the slices of a
    slice group; everything else runs once.
The result is the sum over slices, accumulated in FP64 on the device.
"""
from __future__ import annotations

import ctypes as ct
import math
from collections import defaultdict

import numpy as np

from . import _lib


class PlanDesc(ct.Structure):
    _fields_ = [
        ("n_slots", ct.c_int32),
        ("slot_elems", ct.POINTER(ct.c_int64)),
        ("slot_batch_mask", ct.POINTER(ct.c_int64)),
        ("slot_leaf", ct.POINTER(ct.c_int64)),
        ("n_sliced", ct.c_int32),
        ("slot_slice_stride", ct.POINTER(ct.c_int64)),
        ("group_bits", ct.c_int32),
        ("n_steps", ct.c_int32),
        ("step_slots", ct.POINTER(ct.c_int32)),
        ("step_kc", ct.POINTER(ct.c_int32)),
        ("step_tile", ct.POINTER(ct.c_int32)),
        ("step_nmodes", ct.POINTER(ct.c_int32)),
        ("step_modes", ct.POINTER(ct.c_int64)),
        ("out_slot", ct.c_int32),
        ("slot_dtype", ct.POINTER(ct.c_int32)),
        ("step_kind", ct.POINTER(ct.c_int32)),
        ("step_arith", ct.POINTER(ct.c_int32)),
        ("step_chain", ct.POINTER(ct.c_int64)),
        ("n_chain_words", ct.c_int64),
        ("chain_words", ct.POINTER(ct.c_int32)),
        ("n_out_bits", ct.c_int32),
        ("n_out_qubits", ct.c_int32),
        ("out_qubit_slot", ct.POINTER(ct.c_int32)),
        ("out_qubit_bit", ct.POINTER(ct.c_int32)),
    ]


def _arr(ctype, xs):
    return (ctype * max(1, len(xs)))(*xs)


SMALL_ELEMS = 1 << 16   # small intermediates: carried sliced bonds, partner re-layout
# precision="mixed": steps whose operands and result have at most this many
# elements run in FP64 (on the 6x6 depth-14 plan this halves the worst
# amplitude error of the 2^16 limit at no measurable cost)
MIXED_FP64_ELEMS = 1 << 20


class SlotInfo:
    __slots__ = ("idx", "ext", "stride", "leaf", "deps", "slice_stride", "carry", "elems")

    def __init__(self):
        self.idx = []          # per-slice indices, in layout order
        self.ext = {}          # index -> extent
        self.stride = {}       # index -> per-slice stride
        self.leaf = 0
        self.deps = []         # sliced bonds whose values select a copy of this slot (batched)
        self.slice_stride = [] # per sliced bond: its stride when kept as an index, else 0
        self.carry = []        # intermediates: sliced bonds kept as indices (laid out last)
        self.elems = 1

    @property
    def dep(self):
        """Depends on (is batched over) at least one sliced bond."""
        return bool(self.deps)


def _assign_strides(C):
    """Column-major strides of an intermediate in its layout order; kept
    sliced bonds last."""
    st = 1
    for x in C.idx:
        C.stride[x] = st
        st *= C.ext[x]
    for j in C.carry:
        C.slice_stride[j] = st
        st *= 2
    C.elems = st


def lower_plan(network, plan, sliced=(), out_order=None, leaf_handles=None, leaf_strides=None,
               carry_elems=SMALL_ELEMS, extra_bonds=None, leaf_dims=None):
    """Pure-host lowering (no device calls): returns (slots, steps, out_slot)
    where steps = [(a, b, c, M, N, K)] with modes as (ext, s0, s1).

    leaf_dims: {leaf label: dims of its stored tensor}, when they differ
    from the Node's (a view made by slice_tensor_network has extent 1 on
    the sliced bond, while its Node keeps the old dims, as in the reference).

    extra_bonds: {leaf label: [stride of bond k, k = 0..n-1]} -- n more
    extent-2 bonds after the sliced ones that are no edge of the network
    but that these leaves keep as indices at the given strides (a batch of
    output bitstrings: each closed output qubit's leaf holds one column per
    bitstring, and bond k selects bit k of the bitstring's number).

    carry_elems: an intermediate whose operands are slice-independent keeps
    their sliced bonds as indices (runs once, for every value of those
    bonds) when it, with them, has at most this many elements and no sliced
    bond joins its two operands; 0 turns this off (every intermediate that
    touches a sliced bond is then batched over the slices)."""
    sliced = list(sliced)
    sset = set(sliced)
    extra_bonds = extra_bonds or {}
    n_extra = max((len(v) for v in extra_bonds.values()), default=0)
    nsl = len(sliced) + n_extra
    slots: list[SlotInfo] = []
    label_slot = {}
    for lab, node in network.nodes.items():
        s = SlotInfo()
        dims = list(leaf_dims[lab]) if leaf_dims is not None and lab in leaf_dims else list(node.dims)
        if leaf_strides is not None:
            strides = list(leaf_strides[lab])
        else:
            strides = [int(np.prod(dims[:i])) for i in range(len(dims))]
        s.leaf = leaf_handles[lab] if leaf_handles is not None else 0
        s.slice_stride = [0] * nsl
        for k, st in enumerate(extra_bonds.get(lab, ())):
            s.slice_stride[len(sliced) + k] = st
        for ix, d, st in zip(node.indices, dims, strides):
            if ix in sset:
                if d != 2:
                    raise ValueError(f"sliced bond {ix} has extent {d}; only extent 2 is supported")
                s.slice_stride[sliced.index(ix)] = st
            else:
                s.idx.append(ix)
                s.ext[ix] = d
                s.stride[ix] = st
        s.elems = (int(np.prod(dims)) if dims else 1) << len(extra_bonds.get(lab, ()))
        label_slot[lab] = len(slots)
        slots.append(s)
    n_leaves = len(slots)
    full_deps = {}
    # Pass 1: index sets of every intermediate, in reference order.
    counter = network.counters["node"]
    raw = []
    for a, b in plan:
        sa, sb = label_slot.pop(a), label_slot.pop(b)
        A, B = slots[sa], slots[sb]
        bset, aset = set(B.idx), set(A.idx)
        c_idx = [x for x in A.idx if x not in bset] + [x for x in B.idx if x not in aset]
        C = SlotInfo()
        C.idx = c_idx
        C.ext = {x: (A.ext.get(x) or B.ext.get(x)) for x in c_idx}
        # sliced bonds each operand keeps as indices (leaves: strides are
        # known; intermediates: laid out in pass 2)
        ca = set(A.carry) | {j for j, v in enumerate(A.slice_stride) if v}
        cb = set(B.carry) | {j for j, v in enumerate(B.slice_stride) if v}
        size = math.prod(C.ext.values()) << len(ca | cb)
        # a result depends on the sliced bonds its operands depend on or
        # keep as indices -- only those select its copies (partial batching)
        full_deps[len(slots)] = sorted(set(A.deps) | set(B.deps) | ca | cb)
        if (ca or cb) and not (A.dep or B.dep or ca & cb) and max(size, A.elems, B.elems) <= carry_elems:
            C.carry = sorted(ca | cb)
        else:
            C.deps = full_deps[len(slots)]
        C.slice_stride = [0] * nsl
        C.elems = size if C.carry else math.prod(C.ext.values())
        counter += 1
        label_slot[f"node_{counter}"] = len(slots)
        raw.append((sa, sb, len(slots)))
        slots.append(C)
    if len(label_slot) != 1:
        raise ValueError(f"plan leaves {len(label_slot)} tensors uncontracted")
    (out_slot,) = label_slot.values()
    if slots[out_slot].carry:
        # the result is per slice (summed over slices afterwards)
        slots[out_slot].carry = []
        slots[out_slot].deps = full_deps[out_slot]
    if nsl and sorted(slots[out_slot].deps) != list(range(nsl)):
        raise ValueError("the result does not depend on every sliced bond")
    # Pass 2: layouts.  Consumer of each intermediate -> indices summed there.
    consumer = {}
    for sa, sb, sc in raw:
        consumer[sa] = sb
        consumer[sb] = sa
    for sa, sb, sc in raw:
        C = slots[sc]
        if sc == out_slot:
            order = out_order if out_order is not None else [
                x for x in network.output_qubits if x in C.ext]
            if sorted(order) != sorted(C.idx):
                raise ValueError("output order does not match the open indices")
            C.idx = list(order)
        elif sc in consumer:
            p = consumer[sc]
            pslot = slots[p]
            cset = set(C.idx)
            pset = set(pslot.idx)
            adapt = (p < sc and not pslot.leaf and p >= n_leaves and pslot.elems <= SMALL_ELEMS
                     and pslot.elems < C.elems)
            if p < sc and not adapt:
                # partner already laid out: take the summed indices in its
                # order, so K is contiguous in the same order in both operands
                first = [x for x in pslot.idx if x in cset]
            else:
                # own order (this step's A-side indices before its B-side
                # ones); a small partner laid out earlier is laid out again
                # to match, as only this step reads it
                first = [x for x in C.idx if x in pset]
            fset = set(first)
            C.idx = first + [x for x in C.idx if x not in fset]
            if adapt:
                pslot.idx = first + [x for x in pslot.idx if x not in fset]
                _assign_strides(pslot)
        _assign_strides(C)
    # Pass 3: modes.
    steps = []
    for sa, sb, sc in raw:
        A, B, C = slots[sa], slots[sb], slots[sc]
        bset, aset = set(B.idx), set(A.idx)
        M = [(A.ext[x], A.stride[x], C.stride[x]) for x in A.idx if x not in bset]
        N = [(B.ext[x], B.stride[x], C.stride[x]) for x in B.idx if x not in aset]
        K = sorted([(A.ext[x], A.stride[x], B.stride[x]) for x in A.idx if x in bset], key=lambda m: m[1])
        for j in C.carry:      # kept sliced bonds are open indices of this step
            if A.slice_stride[j]:
                M.append((2, A.slice_stride[j], C.slice_stride[j]))
            else:
                N.append((2, B.slice_stride[j], C.slice_stride[j]))
        M.sort(key=lambda m: m[2])
        N.sort(key=lambda m: m[2])
        steps.append((sa, sb, sc, M, N, K))
    return slots, steps, out_slot




def assign_precision(slots, steps, ctx_code, precision, small_elems=MIXED_FP64_ELEMS, fp64_steps=()):
    """Element type per slot and the step list with conversions:
    [(kind, a, b, c, M, N, K)], kind 0 = contract, 1 = convert a -> c.

    "uniform": every slot in the context's type.  "mixed" (context type
    complex128): a step whose operands and result all have at most
    small_elems elements per slice runs in FP64 while its operands are FP64;
    every other step runs in complex64, after a conversion of FP64 operands.
    Leaves stay FP64, so a complex64 step that reads a leaf runs in FP64
    instead (its other operand converted up).  fp64_steps: indices of
    contraction steps that run in FP64 whatever their size."""
    n0 = len(slots)
    dt = [ctx_code] * n0
    out = []
    if precision == "uniform":
        return dt, [(0, sa, sb, sc, M, N, K) for sa, sb, sc, M, N, K in steps], slots
    if ctx_code != _lib.PQ_C128:
        raise ValueError('precision="mixed" needs a ComplexF64 backend (FP64 leaves)')
    slots = list(slots)
    force = set(fp64_steps)
    for i, (sa, sb, sc, M, N, K) in enumerate(steps):
        small = all(slots[x].elems <= small_elems for x in (sa, sb, sc))
        want = _lib.PQ_C128 if (small and dt[sa] == _lib.PQ_C128 and dt[sb] == _lib.PQ_C128) else _lib.PQ_C64
        if i in force or (want == _lib.PQ_C64 and (slots[sa].leaf or slots[sb].leaf)):
            want = _lib.PQ_C128
        ops = []
        for x in (sa, sb):
            if dt[x] != want:
                src = slots[x]
                y = SlotInfo()
                y.idx, y.ext, y.stride = list(src.idx), dict(src.ext), dict(src.stride)
                y.deps, y.elems = list(src.deps), src.elems
                y.slice_stride, y.carry = list(src.slice_stride), list(src.carry)
                slots.append(y)
                dt.append(want)
                out.append((1, x, x, len(slots) - 1, [], [], []))
                ops.append(len(slots) - 1)
            else:
                ops.append(x)
        dt[sc] = want
        out.append((0, ops[0], ops[1], sc, M, N, K))
    return dt, out, slots


# --------------------------------------------------------------------------- #
# Fused chains: runs of steps that each apply a small operator to the
# previous step's (large) result, executed as one kernel (PQ_KC_CHAIN)
# --------------------------------------------------------------------------- #

CHAIN_TILE_BITS = 13     # tile of 2^13 complex64 = 64 KiB of LDS per workgroup (fewer, longer chains
                         # beat 2^12 tiles with more workgroups per CU on the 6x6 d14 plan)
CHAIN_MIN_ELEMS = 1 << 16
CHAIN_OP_ELEMS = 1024


def _bits(slot, labels):
    """(label, j) bits of labels in a slot, j-th factor 2 of the extent, with
    their strides; None when an extent is not a power of two."""
    out = []
    for x in labels:
        e = slot.ext[x]
        if e & (e - 1):
            return None
        for j in range(e.bit_length() - 1):
            out.append(((x, j), slot.stride[x] << j))
    return out


def chain_program(slots, steps, group, tile_bits=CHAIN_TILE_BITS):
    """Program words (include/pq_hip.h, PQ_CHAIN_*) for the contraction steps
    `group` (indices into steps, each applying its small operand to the
    previous one's result), or None when a limit is exceeded."""
    first = steps[group[0]]
    xs = [first[1], first[2]]
    X0 = max(xs, key=lambda s: slots[s].elems)
    XL = steps[group[-1]][3]
    stages = []
    touched = set()
    x = X0
    for g in group:
        kind, a, b, c = steps[g][:4]
        big, op = (a, b) if a == x else (b, a)
        if big != x:
            return None
        K = [l for l in slots[op].idx if l in set(slots[big].idx)]
        M = [l for l in slots[op].idx if l not in set(slots[big].idx)]
        stages.append((op, K, M))
        touched |= set(K) | set(M)
        x = c
    X, Y = slots[X0], slots[XL]
    spect = [l for l in X.idx if l not in touched]
    if sorted(spect) != sorted(l for l in Y.idx if l not in touched):
        return None
    tile0 = _bits(X, [l for l in X.idx if l in touched])
    if tile0 is None or _bits(Y, Y.idx) is None or _bits(X, X.idx) is None:
        return None
    if any(_bits(slots[op], slots[op].idx) is None for op, _, _ in stages):
        return None

    # tile size over the stages for a given set of passenger labels
    def size_with(pass_labels):
        n = len(tile0) + sum(len(_bits(X, [l])) for l in pass_labels)
        peak = n
        for op, K, M in stages:
            n += len(_bits(slots[op], M)) - len(_bits(slots[op], K))
            peak = max(peak, n)
        return peak

    if size_with([]) > tile_bits:
        return None
    # passengers: spectators with the smallest output (then input) strides,
    # so that the tile's stores (and loads) run over >= 16 adjacent elements
    passengers = []
    for ref in (Y, X):
        low = sorted(ref.idx, key=lambda l: ref.stride[l])
        run = 0
        for l in low:
            if run >= 4:
                break
            if l in spect and l not in passengers:
                if size_with(passengers + [l]) > tile_bits:
                    break
                passengers.append(l)
            run += len(_bits(ref, [l]))
    spect = [l for l in spect if l not in passengers]
    # stage 0: tile bits in input-stride order take positions 0, 1, ...
    gather = sorted(_bits(X, [l for l in X.idx if l in touched or l in passengers]), key=lambda t: t[1])
    pos = {bit: i for i, (bit, _) in enumerate(gather)}
    free = []
    top = len(gather)
    words_st = []
    for op, K, M in stages:
        S = slots[op]
        kb, mb = _bits(S, K), _bits(S, M)
        if kb is None or mb is None or len(kb) > _lib.PQ_CHAIN_MAX_OP_BITS or len(mb) > _lib.PQ_CHAIN_MAX_OP_BITS:
            return None
        kpos = [(pos.pop(bit), s) for bit, s in kb]
        free = sorted(free + [p for p, _ in kpos])
        rbits = sorted(pos.values())
        mpos = []
        for bit, s in mb:
            if free:
                p = free.pop(0)
            else:
                p = top
                top += 1
            pos[bit] = p
            mpos.append((p, s))
        words_st.append((op, kpos, mpos, rbits))
    if top > tile_bits:
        return None
    # operators and per-stage offsets staged in LDS beside the tile
    staged = sum(8 * (1 << (len(kp) + len(mp)))
                 + 4 * ((1 << len(kp)) + (1 << len(mp)) + (1 << min(len(rb), 4)) + (1 << (len(rb) - min(len(rb), 4))))
                 for _, kp, mp, rb in words_st)
    if staged > _lib.PQ_CHAIN_STAGE_BYTES:
        return None
    final = {bit[0] for bit in pos}
    scatter = sorted(((pos[bit], s) for bit, s in _bits(Y, [l for l in Y.idx if l in final])), key=lambda t: t[1])
    if len(scatter) != len(pos):
        return None
    words = [len(stages), top, len(gather), len(scatter), len(spect)]
    for i, (bit, s) in enumerate(gather):
        words += [i, s]
    for p, s in scatter:
        words += [p, s]
    for l in spect:
        words += [X.ext[l], X.stride[l], Y.stride[l]]
    for op, kpos, mpos, rbits in words_st:
        words += [op, len(kpos), len(mpos), len(rbits)]
        for p, s in kpos + mpos:
            words += [p, s]
        words += rbits
    return words


def chain_stages(words):
    """(K bits, M bits) of every stage of a fused-chain program: 2^K x 2^M
    is the operator shape the stage applies (e.g. (5, 3): 32 -> 8)."""
    L, tb, nin, nout, nspec = words[:5]
    q = 5 + 2 * nin + 2 * nout + 3 * nspec
    out = []
    for _ in range(L):
        _, nk, nm, nr = words[q:q + 4]
        out.append((nk, nm))
        q += 4 + 2 * (nk + nm) + nr
    return out


def fuse_chains(slots, steps, dts, tile_bits=CHAIN_TILE_BITS, min_elems=CHAIN_MIN_ELEMS,
                op_elems=CHAIN_OP_ELEMS, group_bits=None):
    """Replace runs of complex64 steps that each contract a small operand
    (<= op_elems) with the previous step's large (>= min_elems) result by
    fused-chain steps (kind 2), as many consecutive steps per fused step as
    fit a tile of 2^tile_bits.  Only steps that run once per slice group
    take part: their tensors depend on no sliced bond below group_bits
    (None: on none at all) and keep none as an index -- except operators,
    which may keep bonds above the group (read at each group's offset).
    Returns (steps, programs, members): programs[i] are the words of step i
    (None for other steps), members[i] the indices (into the input
    `steps`) of the contraction steps it stands for ([] for a
    conversion)."""
    reads = defaultdict(int)
    for kind, a, b, c, *_ in steps:
        reads[a] += 1
        if kind == 0 and b != a:
            reads[b] += 1

    gb = len(slots[0].slice_stride) if group_bits is None else group_bits

    def plain(s, operator):
        sl = slots[s]
        if any(j < gb for j in sl.deps):
            return False
        if operator:
            return not any(v for j, v in enumerate(sl.slice_stride) if j < gb)
        return not sl.carry and not any(sl.slice_stride)

    cand = {}
    for i, (kind, a, b, c, M, N, K) in enumerate(steps):
        if kind != 0 or not (dts[a] == dts[b] == dts[c] == _lib.PQ_C64):
            continue
        big, op = (a, b) if slots[a].elems >= slots[b].elems else (b, a)
        if slots[op].elems > op_elems or slots[big].elems < min_elems or slots[c].elems < min_elems:
            continue
        if not (plain(big, False) and plain(op, True) and plain(c, False)):
            continue
        cand[i] = big
    # chains: a candidate continues the chain whose last result it reads
    chains, end_of = [], {}
    for i in sorted(cand):
        big = cand[i]
        if big in end_of and reads[big] == 1:
            ch = end_of.pop(big)
        else:
            ch = []
            chains.append(ch)
        ch.append(i)
        end_of[steps[i][3]] = ch
    groups = []
    for ch in chains:
        i = 0
        while i < len(ch):
            j, best = i + 1, None
            while j < len(ch) and j - i < _lib.PQ_CHAIN_MAX_STAGES:
                w = chain_program(slots, steps, ch[i:j + 1], tile_bits)
                if w is None:
                    break
                best = (j, w)
                j += 1
            if best is None:
                i += 1
                continue
            groups.append((ch[i:best[0] + 1], best[1]))
            i = best[0] + 1
    if not groups:
        return steps, [None] * len(steps), [[i] if s[0] == 0 else [] for i, s in enumerate(steps)]
    drop, at = set(), {}
    for members, words in groups:
        drop |= set(members[:-1])
        at[members[-1]] = (members, words)
    out, progs, mem = [], [], []
    for i, st in enumerate(steps):
        if i in drop:
            continue
        if i in at:
            members, words = at[i]
            first = steps[members[0]]
            x0 = first[1] if slots[first[1]].elems >= slots[first[2]].elems else first[2]
            op0 = first[2] if x0 == first[1] else first[1]
            out.append((2, x0, op0, st[3], [], [], []))
            progs.append(words)
            mem.append(list(members))
        else:
            out.append(st)
            progs.append(None)
            mem.append([i] if st[0] == 0 else [])
    return out, progs, mem


class Plan:
    """A compiled plan on a context.  execute() enqueues one full run."""

    def __init__(self, ctx, ptr, slots, steps, out_slot, n_sliced, group_bits, dtypes=None):
        self.ctx = ctx
        self.ptr = ptr
        self.slots = slots
        self.steps = steps          # [(kind, a, b, c, M, N, K)]
        self.out_slot = out_slot
        self.n_sliced = n_sliced
        self.group_bits = group_bits
        self.dtypes = dtypes
        self.chain_programs = [None] * len(steps)   # fused-chain program words per step
        out = slots[out_slot]
        self.out_shape = tuple(out.ext[x] for x in out.idx)
        self.network = None
        self.batch_handles = {}    # closed output leaf -> its bitstring-batch tensor
        self.n_out_bits = 0
        self.n_bitstrings = None

    def execute(self, use_graph=True):
        self.ctx.check(self.ctx.lib.pq_plan_execute(self.ctx.ptr, self.ptr, 1 if use_graph else 0))

    def result(self):
        """The FP64 result in the output qubits' order; with a bitstring
        batch, shape (number of bitstrings,) + that shape."""
        buf = self._result_buffer()
        self.ctx.check(self.ctx.lib.pq_plan_result(self.ctx.ptr, self.ptr,
                                                   buf.ctypes.data_as(ct.POINTER(ct.c_double))))
        return self._shaped(buf)

    def _result_buffer(self):
        n = int(np.prod(self.out_shape)) if self.out_shape else 1
        return np.empty(2 * n * (1 << self.n_out_bits), dtype=np.float64)

    def _shaped(self, buf):
        n = int(np.prod(self.out_shape)) if self.out_shape else 1
        z = buf[0::2] + 1j * buf[1::2]
        if self.n_bitstrings is None:
            return z.reshape(self.out_shape, order="F")
        return np.stack([z[b * n:(b + 1) * n].reshape(self.out_shape, order="F")
                         for b in range(self.n_bitstrings)])

    def set_bitstrings(self, bitstrings):
        """Rewrite the output batch in place (pq_plan_set_bitstrings): the
        next runs, graph replays included, compute these bitstrings.  Same
        count as the plan was compiled with."""
        words = self._words(bitstrings)
        self.ctx.check(self.ctx.lib.pq_plan_set_bitstrings(self.ctx.ptr, self.ptr, len(words), words))

    def _words(self, bitstrings):
        if self.n_bitstrings is None:
            raise ValueError("the plan was compiled without a bitstring batch")
        bitstrings = list(bitstrings)
        if len(bitstrings) != self.n_bitstrings:
            raise ValueError(f"the plan computes {self.n_bitstrings} bitstrings per run")
        data, _ = output_batch(self.network, bitstrings)
        if set(data) != set(self.batch_handles):
            raise ValueError("the bitstrings close other outputs than the plan's")
        return bitstring_words(self.network, bitstrings)

    def run(self, bitstrings=None):
        """pq_plan_run: set the bitstrings (if given), run once (graph) and
        return the result as result() does."""
        words = self._words(bitstrings) if bitstrings is not None else None
        buf = self._result_buffer()
        self.ctx.check(self.ctx.lib.pq_plan_run(self.ctx.ptr, self.ptr, 0 if words is None else len(words), words,
                                                buf.ctypes.data_as(ct.POINTER(ct.c_double))))
        return self._shaped(buf)

    def cost(self):
        """pq_plan_cost: exact MACs and algorithmic bytes of one run, in
        total and per step."""
        tot = _lib.Cost()
        m = np.zeros(len(self.steps), dtype=np.float64)
        b = np.zeros(len(self.steps), dtype=np.float64)
        self.ctx.check(self.ctx.lib.pq_plan_cost(self.ptr, ct.byref(tot), m.ctypes.data_as(ct.POINTER(ct.c_double)),
                                                 b.ctypes.data_as(ct.POINTER(ct.c_double))))
        return {"macs": tot.macs, "bytes": tot.bytes, "steps": tot.steps, "launches": tot.launches,
                "step_macs": m, "step_bytes": b}

    def stats(self):
        a, l, o = ct.c_int64(), ct.c_int64(), ct.c_int64()
        f, b = ct.c_double(), ct.c_double()
        self.ctx.check(self.ctx.lib.pq_plan_stats(self.ptr, ct.byref(a), ct.byref(l), ct.byref(f), ct.byref(b),
                                                  ct.byref(o)))
        return {"arena_bytes": a.value, "launches": l.value, "alg_flops": f.value, "alg_bytes": b.value,
                "out_elems": o.value}

    def step_info(self, i):
        kc, tile, dep = ct.c_int(), ct.c_int(), ct.c_int()
        f, b = ct.c_double(), ct.c_double()
        self.ctx.check(self.ctx.lib.pq_plan_step_info(self.ptr, i, ct.byref(kc), ct.byref(tile), ct.byref(dep),
                                                      ct.byref(f), ct.byref(b)))
        return {"kc": kc.value, "tile": tile.value, "dep": bool(dep.value), "flops": f.value, "bytes": b.value}

    def profile(self):
        ms = np.zeros(len(self.steps), dtype=np.float64)
        self.ctx.check(self.ctx.lib.pq_plan_profile(self.ctx.ptr, self.ptr, ms.ctypes.data_as(ct.POINTER(ct.c_double))))
        return ms

    def close(self):
        if self.ptr:
            self.ctx.lib.pq_plan_destroy(self.ctx.ptr, self.ptr)
            self.ptr = None
        for h in self.batch_handles.values():
            self.ctx.free(h)
        self.batch_handles = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def step_choices(pre_steps, members, kc=None, tiles=None):
    """Kernel class and GEMM tile of every step after fusion.  kc / tiles
    are per contraction step of the plan, i.e. per kind-0 step of
    pre_steps (the step list before fusion) in order; members[i] names the
    pre_steps a fused step stands for (fuse_chains), so a step between two
    members of a chain keeps its own choice.  Chains and conversions: -1."""
    plan_step, ci = {}, 0
    for i, st in enumerate(pre_steps):
        if st[0] == 0:
            plan_step[i] = ci
            ci += 1
    kcs, tls = [], []
    for mem in members:
        if len(mem) == 1:
            kcs.append(kc[plan_step[mem[0]]] if kc is not None else -1)
            tls.append(tiles[plan_step[mem[0]]] if tiles is not None else -1)
        else:
            kcs.append(-1)
            tls.append(-1)
    return kcs, tls


def _leaf_layout(backend):
    """Handles, strides and dims of the backend's device tensors."""
    handles, strides, dims = {}, {}, {}
    ctx = backend.ctx
    for lab, h in backend.tensors.items():
        base = ct.c_void_p()
        off = ct.c_int64()
        st = (ct.c_int64 * _lib.PQ_MAX_RANK)()
        ctx.check(ctx.lib.pq_tensor_device_ptr(ctx.ptr, h, ct.byref(base), ct.byref(off), st))
        shp = ctx.shape(h)
        handles[lab] = h
        strides[lab] = [int(st[i]) for i in range(len(shp))]
        dims[lab] = list(shp)
    return handles, strides, dims


def choose_group_bits(slots, n_sliced, budget_bytes, esize):
    """Largest slice group whose batched intermediates fit the budget (a slot
    holds one copy per value of its sliced bonds within the group)."""
    gb = n_sliced

    def peak(g):
        return max((s.elems << sum(1 for j in s.deps if j < g) for s in slots if s.dep and not s.leaf), default=1)

    while gb > 0 and 3 * peak(gb) * esize > budget_bytes:
        gb -= 1
    return min(gb, 16)


def output_batch(network, bitstrings):
    """(leaf data, n_bits) of a batch of output bitstrings: for every closed
    output qubit (a node made by add_output), a [2, 2^n_bits] array whose
    column b is the qubit's output vector for bitstring b ('0' -> [1, 0],
    '1' -> [0, 1], src/layer3.jl:302); the batch is padded to a power of two
    with copies of the first bitstring.  bitstrings: strings over all qubits
    in add_output's order ('x' or any character at an open qubit)."""
    bitstrings = list(bitstrings)
    if not bitstrings:
        raise ValueError("empty bitstring batch")
    n = network.number_qubits
    if any(len(b) != n for b in bitstrings):
        raise ValueError(f"every bitstring needs {n} characters")
    n_bits = max(0, math.ceil(math.log2(len(bitstrings))))
    padded = bitstrings + [bitstrings[0]] * ((1 << n_bits) - len(bitstrings))
    data = {}
    for i in range(n):
        pos = network.qubit_ordering[i]
        node = network.edges[network.output_qubits[pos - 1]].dst
        if node is None or network.node_layers.get(node) != -1:
            continue                       # open output
        col = np.zeros((2, len(padded)), dtype=np.complex128)
        for b, bits in enumerate(padded):
            if bits[i] not in "01":
                raise ValueError(f"bitstring {b} has {bits[i]!r} at closed qubit {i + 1}")
            col[int(bits[i]), b] = 1.0
        data[node] = col
    if not data:
        raise ValueError("the network has no closed outputs to batch over")
    return data, n_bits


def closed_output_qubits(network):
    """[(node label, qubit index)] of the closed output qubits (the nodes
    add_output made), qubit index = position in add_output's string."""
    out = []
    for i in range(network.number_qubits):
        pos = network.qubit_ordering[i]
        node = network.edges[network.output_qubits[pos - 1]].dst
        if node is not None and network.node_layers.get(node) == -1:
            out.append((node, i))
    return out


def bitstring_words(network, bitstrings):
    """uint64 words of a bitstring batch for pq_plan_set_bitstrings: bit i
    of word b = the value of closed output qubit i in bitstring b."""
    closed = closed_output_qubits(network)
    if any(i > 63 for _, i in closed):
        raise ValueError("bitstring words hold at most 64 qubits")
    words = (ct.c_uint64 * len(bitstrings))()
    for b, bits in enumerate(bitstrings):
        w = 0
        for _, i in closed:
            if bits[i] not in "01":
                raise ValueError(f"bitstring {b} has {bits[i]!r} at closed qubit {i + 1}")
            w |= int(bits[i]) << i
        words[b] = w
    return words


def compile_plan(network, plan, sliced=(), out_order=None, group_bits=None, kc=None, tiles=None,
                 budget_bytes=64 << 30, precision="uniform", fp64_steps=(), carry_elems=SMALL_ELEMS,
                 small_elems=MIXED_FP64_ELEMS, fuse=True, bitstrings=None):
    """Compile a pair plan of a network whose backend is a HIPBackend.

    precision: "uniform" (every step in the backend's type), "mixed" (FP64
    backend: small steps in FP64, large ones in complex64 on FP32 MFMA; see
    assign_precision with small_elems; fp64_steps forces contraction steps
    to FP64) or
    "mixed16" (as "mixed", with the complex64 GEMMs on FP16 hi/lo matrix
    cores where their shape allows).  kc / tiles, when given, are per contraction step of
    `plan` (kernel class and GEMM tile, -1 = automatic).  fuse: runs of
    complex64 steps that apply small operators to one large tensor become
    fused-chain steps (fuse_chains; a dict passes its limits).
    bitstrings: a batch of output bitstrings (output_batch) computed in one
    run, one result per bitstring; the closed output qubits' leaves are
    replaced by device tensors with a column per bitstring, which
    Plan.set_bitstrings rewrites in place for the next runs."""
    backend = network.backend
    ctx = backend.ctx
    handles, strides, dims = _leaf_layout(backend)
    missing = [lab for lab in network.nodes if lab not in handles]
    if missing:
        raise KeyError(f"leaf tensors missing on the device: {missing[:4]}")
    extra, batch_handles, n_out_bits, n_bits_in = {}, {}, 0, None
    if bitstrings is not None:
        data, n_out_bits = output_batch(network, bitstrings)
        n_bits_in = len(list(bitstrings))
        for lab, col in data.items():
            h = ctx.upload(col)
            batch_handles[lab] = h
            handles[lab], strides[lab], dims[lab] = h, [1], [2]
            extra[lab] = [2 << k for k in range(n_out_bits)]
    for lab, node in network.nodes.items():
        if len(dims[lab]) != len(node.dims):
            raise ValueError(f"leaf {lab}: the stored tensor has rank {len(dims[lab])}, the node {len(node.dims)}")
    slots, steps, out_slot = lower_plan(network, plan, sliced, out_order, handles, strides, carry_elems,
                                        extra_bonds=extra, leaf_dims=dims)
    n_real = len(list(sliced))
    n_sliced = n_real + n_out_bits
    split16 = precision == "mixed16"
    dts, steps, slots = assign_precision(slots, steps, ctx.code, "mixed" if split16 else precision,
                                         small_elems=small_elems, fp64_steps=fp64_steps)
    if group_bits is None:
        esize = 8 if precision in ("mixed", "mixed16") or ctx.code == _lib.PQ_C64 else 16
        group_bits = choose_group_bits(slots, n_real, budget_bytes, esize)
    if group_bits > n_real:
        raise ValueError("the slice group cannot include output-batch bonds")
    pre_steps = steps
    if fuse:
        steps, progs, members = fuse_chains(slots, steps, dts, group_bits=group_bits,
                                            **(fuse if isinstance(fuse, dict) else {}))
    else:
        progs, members = [None] * len(steps), [[i] if st[0] == 0 else [] for i, st in enumerate(steps)]
    slot_elems = [s.elems for s in slots]
    slot_mask = [0 if s.leaf else sum(1 << j for j in s.deps) for s in slots]
    slot_leaf = [s.leaf for s in slots]
    lss = []
    for s in slots:
        lss += (list(s.slice_stride) if s.slice_stride else [0] * n_sliced)
    step_slots, nmodes, modes, kinds, ariths = [], [], [], [], []
    kcs, tls = step_choices(pre_steps, members, kc, tiles)
    chain_off, words = [], []
    for (kind, sa, sb, sc, M, N, K), prog in zip(steps, progs):
        step_slots += [sa, sb, sc]
        nmodes += [len(M), len(N), len(K)]
        kinds.append(kind)
        ariths.append(1 if (split16 and kind == 0 and dts[sc] == _lib.PQ_C64) else 0)
        for grp in (M, N, K):
            for m in grp:
                modes += list(m)
        chain_off.append(len(words) if prog is not None else -1)
        if prog is not None:
            words += prog
    keep = dict(
        slot_elems=_arr(ct.c_int64, slot_elems), slot_mask=_arr(ct.c_int64, slot_mask),
        slot_leaf=_arr(ct.c_int64, slot_leaf), lss=_arr(ct.c_int64, lss or [0]),
        step_slots=_arr(ct.c_int32, step_slots), kc=_arr(ct.c_int32, kcs), tile=_arr(ct.c_int32, tls),
        nmodes=_arr(ct.c_int32, nmodes), modes=_arr(ct.c_int64, modes or [0]),
        dt=_arr(ct.c_int32, dts), kind=_arr(ct.c_int32, kinds), arith=_arr(ct.c_int32, ariths),
        chain_off=_arr(ct.c_int64, chain_off), words=_arr(ct.c_int32, words or [0]))
    # closed output qubits of a bitstring batch: their leaf slots and bits
    oq_slot, oq_bit = [], []
    if batch_handles:
        slot_of = {s.leaf: i for i, s in enumerate(slots) if s.leaf}
        for node, i in closed_output_qubits(network):
            if node in batch_handles:
                oq_slot.append(slot_of[batch_handles[node]])
                oq_bit.append(i)
    keep["oq_slot"] = _arr(ct.c_int32, oq_slot)
    keep["oq_bit"] = _arr(ct.c_int32, oq_bit)
    d = PlanDesc(len(slots), keep["slot_elems"], keep["slot_mask"], keep["slot_leaf"], n_sliced, keep["lss"],
                 group_bits, len(steps), keep["step_slots"], keep["kc"], keep["tile"], keep["nmodes"],
                 keep["modes"], out_slot, keep["dt"], keep["kind"], keep["arith"], keep["chain_off"],
                 len(words), keep["words"], n_out_bits, len(oq_slot), keep["oq_slot"], keep["oq_bit"])
    p = ct.c_void_p()
    try:
        ctx.check(ctx.lib.pq_plan_create(ctx.ptr, ct.cast(ct.pointer(d), ct.c_void_p), ct.byref(p)))
    except Exception:
        for h in batch_handles.values():
            ctx.free(h)
        raise
    out = Plan(ctx, p, slots, steps, out_slot, n_sliced, group_bits, dts)
    out.chain_programs = progs
    out.network = network
    out.batch_handles = batch_handles
    out.n_out_bits = n_out_bits
    out.n_bitstrings = n_bits_in
    return out
