"""pq_contract / permute / view / reshape through the C ABI on the GPU versus
the oracle's contraction (numpy tensordot in complex128).

Tolerances: ComplexF64 -- 1e-12 relative to sum_k |a||b| (the natural scale of
each output element); ComplexF32 -- 2e-6 of the same scale."""
import numpy as np
import pytest

from oracle import pq_oracle as O
from picoquant_hip import _lib

pytestmark = pytest.mark.gpu

TOL = {np.complex128: 1e-12, np.complex64: 2e-6}


@pytest.fixture(scope="module", params=[np.complex128, np.complex64], ids=["c128", "c64"])
def ctx(request):
    c = _lib.Context(request.param)
    yield c
    c.close()


def rand(rng, shape, dtype=np.complex128):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)


def check(ctx, A, ia, B, ib, got):
    want = O.contract_tensors_data(A.astype(np.complex128), ia, B.astype(np.complex128), ib)
    scale = O.contract_tensors_data(np.abs(A).astype(np.complex128), ia,
                                    np.abs(B).astype(np.complex128), ib).real
    assert got.shape == want.shape
    err = np.abs(got - want) / np.maximum(scale, 1e-300)
    assert err.max() <= TOL[ctx.np_dtype.type], err.max()


def run(ctx, A, ia, B, ib, kc=-1):
    ha, hb = ctx.upload(A), ctx.upload(B)
    ctx.set_kernel_override(kc)
    try:
        hc = ctx.contract(ha, ia, hb, ib)
    finally:
        ctx.set_kernel_override(-1)
    got = ctx.download(hc)
    for h in (ha, hb, hc):
        ctx.free(h)
    return got


def random_case(rng, max_rank=6):
    ra = int(rng.integers(0, max_rank + 1))
    rb = int(rng.integers(0, max_rank + 1))
    nc = int(rng.integers(0, min(ra, rb) + 1))
    dims_c = [int(rng.integers(1, 5)) for _ in range(nc)]
    da = [int(rng.integers(1, 5)) for _ in range(ra - nc)]
    db = [int(rng.integers(1, 5)) for _ in range(rb - nc)]
    la = list(range(1, nc + 1)) + [-(i + 1) for i in range(ra - nc)]
    lb = list(range(1, nc + 1)) + [-(ra - nc + i + 1) for i in range(rb - nc)]
    dimsA = dims_c + da
    dimsB = dims_c + db
    pa = rng.permutation(ra)
    pb = rng.permutation(rb)
    ia = [la[i] for i in pa]
    ib = [lb[i] for i in pb]
    sa = [dimsA[i] for i in pa]
    sb = [dimsB[i] for i in pb]
    return rand(rng, sa), ia, rand(rng, sb), ib


def test_random_label_patterns(ctx):
    rng = np.random.default_rng(0)
    for _ in range(60):
        A, ia, B, ib = random_case(rng)
        check(ctx, A, ia, B, ib, run(ctx, A, ia, B, ib))


@pytest.mark.parametrize("kc", [_lib.KC_GENERIC, _lib.KC_STREAM])
def test_forced_kernel_classes(ctx, kc):
    rng = np.random.default_rng(1)
    n = 0
    while n < 25:
        A, ia, B, ib = random_case(rng, 5)
        try:
            got = run(ctx, A, ia, B, ib, kc)
        except _lib.PQError as e:
            assert e.code == _lib.PQ_ERR_UNSUPPORTED
            continue
        check(ctx, A, ia, B, ib, got)
        n += 1


@pytest.mark.parametrize("pos", [0, 3, 9, 16])
def test_gate_on_state(ctx, pos):
    """One- and two-qubit gates applied to a rank-17 state (2^17 elements)."""
    rng = np.random.default_rng(pos)
    n = 17
    psi = rand(rng, (2,) * n)
    ia = [-(i + 1) for i in range(n)]
    ia[pos] = 1
    G = rand(rng, (2, 2))
    check(ctx, psi, ia, G, [1, -(n + 1)], run(ctx, psi, ia, G, [1, -(n + 1)]))
    ia2 = list(ia)
    q2 = (pos + 5) % n
    ia2[q2] = 2
    G2 = rand(rng, (2, 2, 2, 2))
    ib2 = [1, 2, -(n + 1), -(n + 2)]
    check(ctx, psi, ia2, G2, ib2, run(ctx, psi, ia2, G2, ib2))
    # small operand first: C order starts with the gate's open index
    check(ctx, G, [-1, 1], psi, [1 if i == pos else -(i + 2) for i in range(n)],
          run(ctx, G, [-1, 1], psi, [1 if i == pos else -(i + 2) for i in range(n)]))


def test_reduce_to_scalar(ctx):
    rng = np.random.default_rng(5)
    A = rand(rng, (2,) * 18)
    B = rand(rng, (2,) * 18)
    ia = list(range(1, 19))
    ib = [int(p) + 1 for p in rng.permutation(18)]      # B's axes in another order
    check(ctx, A, ia, B, ib, run(ctx, A, ia, B, ib))
    check(ctx, A, ia, B, ib, run(ctx, A, ia, B, ib, _lib.KC_REDUCE))
    # A's first/last and B's matching axes stay open: 2x2x2x2 = 16 outputs
    ia2 = [-1] + ia[1:-1] + [-2]
    ib2 = [-3 if x == 1 else (-4 if x == 18 else x) for x in ib]
    check(ctx, A, ia2, B, ib2, run(ctx, A, ia2, B, ib2))
    check(ctx, A, ia2, B, ib2, run(ctx, A, ia2, B, ib2, _lib.KC_REDUCE))


def test_dense_pair(ctx):
    """Two rank-10 tensors sharing 5 indices (GEMM-like, M=N=K=32)."""
    rng = np.random.default_rng(7)
    A = rand(rng, (2,) * 10)
    B = rand(rng, (2,) * 10)
    ia = [1, -1, 2, -2, 3, -3, 4, -4, 5, -5]
    ib = [-6, 5, -7, 4, -8, 3, -9, 2, -10, 1]
    check(ctx, A, ia, B, ib, run(ctx, A, ia, B, ib))


def test_outer_product(ctx):
    rng = np.random.default_rng(8)
    A = rand(rng, (2, 3, 2))
    B = rand(rng, (4, 2))
    check(ctx, A, [-1, -2, -3], B, [-4, -5], run(ctx, A, [-1, -2, -3], B, [-4, -5]))
    s = rand(rng, ())
    check(ctx, s, [], B, [-1, -2], run(ctx, s, [], B, [-1, -2]))


def test_permute_view_reshape(ctx):
    rng = np.random.default_rng(9)
    T = rand(rng, (2, 3, 4, 5))
    h = ctx.upload(T)
    hp = ctx.permute(h, [3, 1, 4, 2])
    np.testing.assert_array_equal(ctx.download(hp), np.transpose(T, (2, 0, 3, 1)).astype(ctx.np_dtype))
    hv = ctx.view(hp, 3, 2, 4)
    want_v = np.transpose(T, (2, 0, 3, 1))[:, :, 1:4, :].astype(ctx.np_dtype)
    np.testing.assert_array_equal(ctx.download(hv), want_v)
    # contraction straight from the strided view
    G = rand(rng, (3, 2))
    got = run_handles(ctx, hv, [-1, -2, 1, -3], ctx.upload(G), [1, -4])
    check(ctx, want_v, [-1, -2, 1, -3], G, [1, -4], got)
    ctx.reshape(hv, [4 * 2, 3 * 3])
    np.testing.assert_array_equal(ctx.download(hv), np.reshape(want_v, (8, 9), order="F"))
    assert ctx.shape(hv) == (8, 9)
    for x in (h, hp, hv):
        ctx.free(x)


def run_handles(ctx, ha, ia, hb, ib):
    hc = ctx.contract(ha, ia, hb, ib)
    out = ctx.download(hc)
    ctx.free(hc)
    ctx.free(hb)
    return out


def test_error_paths(ctx):
    rng = np.random.default_rng(10)
    h = ctx.upload(rand(rng, (2, 2)))
    with pytest.raises(KeyError):
        ctx.download(987654321)
    with pytest.raises(_lib.PQError):
        ctx.contract(h, [1, 1], h, [-1, -2])          # repeated label
    with pytest.raises(_lib.PQError):
        ctx.contract(h, [1, -1], h, [2, -2])          # summed label missing from B
    with pytest.raises(_lib.PQError):
        ctx.contract(h, [1], h, [1, -1])              # rank mismatch
    with pytest.raises(_lib.PQError):
        ctx.permute(h, [1, 1])
    with pytest.raises(_lib.PQError):
        ctx.view(h, 3, 1, 1)
    with pytest.raises(_lib.PQError):
        ctx.reshape(h, [3])
    ctx.free(h)
    with pytest.raises(KeyError):
        ctx.free(h)
