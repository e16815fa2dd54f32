"""MFMA GEMM path (PQ_KC_GEMM) against the oracle: every tile shape, both
operand access orders, permuted and sliced operands, and the boundary x site
shape of the grid sweep plans.  Tolerances as in test_gpu_contract."""
import numpy as np
import pytest

from oracle import pq_oracle as O
from picoquant_hip import _lib
from test_gpu_contract import check, rand

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[np.complex128, np.complex64], ids=["c128", "c64"])
def ctx(request):
    c = _lib.Context(request.param)
    yield c
    c.close()


def gemm(ctx, A, ia, B, ib, tile=-1, orient=-1):
    ha, hb = ctx.upload(A), ctx.upload(B)
    ctx.set_kernel_override(_lib.KC_GEMM)
    ctx.set_gemm_tile(tile)
    ctx.set_gemm_orientation(orient)
    try:
        hc = ctx.contract(ha, ia, hb, ib)
    finally:
        ctx.set_kernel_override(-1)
        ctx.set_gemm_tile(-1)
        ctx.set_gemm_orientation(-1)
    out = ctx.download(hc)
    for h in (ha, hb, hc):
        ctx.free(h)
    return out


@pytest.mark.parametrize("orient", [0, 1], ids=["as-given", "transposed"])
@pytest.mark.parametrize("tile", list(range(_lib.PQ_GEMM_TILES)))
def test_tiles_and_access_orders(ctx, tile, orient):
    """C[m,n] = A[m,k] B[k,n] with K fastest / slowest in each operand, the
    product launched as given and transposed (C^T = B^T A^T)."""
    rng = np.random.default_rng(tile)
    M, N, K = 80, 72, 40          # ragged against every tile
    for a_kfast in (False, True):
        for b_kfast in (False, True):
            A = rand(rng, (K, M)) if a_kfast else rand(rng, (M, K))
            ia = [1, -1] if a_kfast else [-1, 1]
            B = rand(rng, (K, N)) if b_kfast else rand(rng, (N, K))
            ib = [1, -2] if b_kfast else [-2, 1]
            check(ctx, A, ia, B, ib, gemm(ctx, A, ia, B, ib, tile, orient))


def test_multimode_groups(ctx):
    """Rank-12 x rank-10 over 2-dims: open and summed modes interleaved."""
    rng = np.random.default_rng(11)
    A = rand(rng, (2,) * 12)
    B = rand(rng, (2,) * 10)
    ia = [-1, 1, -2, -3, 2, 3, -4, 4, -5, 5, -6, -7]
    ib = [5, -8, 3, 1, -9, 2, -10, 4, -11, -12]
    check(ctx, A, ia, B, ib, gemm(ctx, A, ia, B, ib))


def test_strided_views(ctx):
    rng = np.random.default_rng(12)
    T = rand(rng, (3, 20, 2, 24))
    h = ctx.upload(T)
    hp = ctx.permute(h, [4, 2, 1, 3])            # (24, 20, 3, 2)
    hv = ctx.view(hp, 4, 2, 2)                   # (24, 20, 3, 1)
    want = np.transpose(T, (3, 1, 0, 2))[:, :, :, 1:2]
    G = rand(rng, (20, 3, 17))
    hg = ctx.upload(G)
    ctx.set_kernel_override(_lib.KC_GEMM)
    try:
        hc = ctx.contract(hv, [-1, 1, 2, -2], hg, [1, 2, -3])
    finally:
        ctx.set_kernel_override(-1)
    got = ctx.download(hc)
    check(ctx, want, [-1, 1, 2, -2], G, [1, 2, -3], got)
    for x in (h, hp, hv, hg, hc):
        ctx.free(x)


def test_boundary_times_site(ctx):
    """Sweep-plan shape: boundary (M=2^12 x K=2^6, K fastest) x site (K x N=2^6)."""
    rng = np.random.default_rng(13)
    A = rand(rng, (2,) * 18)
    B = rand(rng, (2,) * 12)
    ia = list(range(1, 7)) + [-(i + 1) for i in range(12)]
    ib = [-(13 + i) for i in range(3)] + list(range(1, 7)) + [-(16 + i) for i in range(3)]
    check(ctx, A, ia, B, ib, gemm(ctx, A, ia, B, ib))
    # automatic dispatch picks the GEMM for this shape
    ha, hb = ctx.upload(A), ctx.upload(B)
    ctx.profile_enable(True)
    hc = ctx.contract(ha, ia, hb, ib)
    got = ctx.download(hc)
    st = ctx.profile_read(_lib.KC_GEMM)
    ctx.profile_enable(False)
    assert st["launches"] == 1
    check(ctx, A, ia, B, ib, got)
    for x in (ha, hb, hc):
        ctx.free(x)


@pytest.mark.parametrize("kfast", [True, False], ids=["contiguous-K", "strided-K"])
def test_long_k_beyond_one_offset_table(ctx, kfast):
    """K = 4608 > 2048: the any-stride kernel walks K in windows of offset
    tables, the contiguous fast path decodes tile offsets on the fly; the
    automatic dispatch takes the GEMM for this shape (rank-2 output 32 x 48,
    the dep-chain shape of reconfigured grid plans)."""
    rng = np.random.default_rng(21)
    M, N, K = 32, 48, 4608
    A = rand(rng, (K, M)) if kfast else rand(rng, (M, K))
    ia = [1, -1] if kfast else [-1, 1]
    B = rand(rng, (K, N))
    ib = [1, -2]
    check(ctx, A, ia, B, ib, gemm(ctx, A, ia, B, ib))
    # split over two K modes with the slow one first in A (strided windows)
    A2 = rand(rng, (72, M, 64))
    B2 = rand(rng, (64, 72, N))
    check(ctx, A2, [2, -1, 1], B2, [1, 2, -2], gemm(ctx, A2, [2, -1, 1], B2, [1, 2, -2]))
    ha, hb = ctx.upload(A), ctx.upload(B)
    ctx.profile_enable(True)
    hc = ctx.contract(ha, ia, hb, ib)
    got = ctx.download(hc)
    st = ctx.profile_read(_lib.KC_GEMM)
    ctx.profile_enable(False)
    assert st["launches"] == 1
    check(ctx, A, ia, B, ib, got)
    for x in (ha, hb, hc):
        ctx.free(x)


@pytest.mark.parametrize("orient", [0, 1], ids=["as-given", "transposed"])
@pytest.mark.parametrize("M,N,K", [(32, 64, 16384), (64, 16, 8192), (40, 24, 6144)])
def test_split_k_partial_products(ctx, M, N, K, orient):
    """Output tiles far fewer than the device's CUs with K >= 6144: the GEMM
    runs as K-parts written to scratch and summed in FP64 (at least 4 parts
    for each shape here: M*N spans at most a few 32x32 tiles, and every part
    keeps >= 512 of K).  Contiguous and strided K, ragged M/N, both launch
    orientations, against the oracle; then the same split K through two K
    modes (the slow one cut) and through the automatic dispatch."""
    rng = np.random.default_rng(M + N + K)
    A = rand(rng, (K, M))
    B = rand(rng, (N, K))
    check(ctx, A, [1, -1], B, [-2, 1], gemm(ctx, A, [1, -1], B, [-2, 1], orient=orient))
    A2 = rand(rng, (M, K // 64, 64))
    B2 = rand(rng, (64, N, K // 64))
    check(ctx, A2, [-1, 2, 1], B2, [1, -2, 2], gemm(ctx, A2, [-1, 2, 1], B2, [1, -2, 2], orient=orient))
    ha, hb = ctx.upload(A), ctx.upload(B)
    hc = ctx.contract(ha, [1, -1], hb, [-2, 1])
    check(ctx, A, [1, -1], B, [-2, 1], ctx.download(hc))
    for x in (ha, hb, hc):
        ctx.free(x)


GEMM16_TOL = 1e-5    # relative to sum |a||b|: hi/lo split drops lo*lo (~2^-22) and rounds lo


@pytest.mark.parametrize("orient", [0, 1], ids=["as-given", "transposed"])
@pytest.mark.parametrize("tile", list(range(_lib.PQ_GEMM_TILES)))
def test_gemm16_split_against_oracle(tile, orient):
    """complex64 GEMM on FP16 hi/lo matrix cores (PQ_KC_GEMM16): K contiguous
    in both operands and a multiple of 32; operands spanning many binades
    (per-row scaling) and ragged M, N."""
    ctx = _lib.Context(np.complex64)
    try:
        rng = np.random.default_rng(100 + tile)
        M, N, K = 80, 72, 96
        A = rand(rng, (K, M)) * np.exp2(rng.integers(-30, 30, size=(1, M)))      # row scales 2^+-30
        B = rand(rng, (K, N)) * np.exp2(rng.integers(-20, 20, size=(1, N)))
        A, B = A.astype(np.complex64), B.astype(np.complex64)
        ha, hb = ctx.upload(A), ctx.upload(B)
        ctx.set_kernel_override(_lib.KC_GEMM16)
        ctx.set_gemm_tile(tile)
        ctx.set_gemm_orientation(orient)
        ctx.profile_enable(True)
        try:
            hc = ctx.contract(ha, [1, -1], hb, [1, -2])
        finally:
            ctx.set_kernel_override(-1)
            ctx.set_gemm_tile(-1)
            ctx.set_gemm_orientation(-1)
        got = ctx.download(hc)
        assert ctx.profile_read(_lib.KC_GEMM16)["launches"] == 1
        ctx.profile_enable(False)
        want = O.contract_tensors_data(A.astype(np.complex128), [1, -1], B.astype(np.complex128), [1, -2])
        scale = O.contract_tensors_data(np.abs(A).astype(np.complex128), [1, -1],
                                        np.abs(B).astype(np.complex128), [1, -2]).real
        err = np.abs(got - want) / scale
        assert err.max() <= GEMM16_TOL, err.max()
        assert err.max() > 1e-9          # really ran in reduced precision
    finally:
        ctx.close()


def test_gemm16_long_k():
    """K = 4160 (130 tiles of 32, beyond the 64-entry tile table): tile
    offsets decoded on the fly."""
    ctx = _lib.Context(np.complex64)
    try:
        rng = np.random.default_rng(31)
        M, N, K = 48, 40, 4160
        A = rand(rng, (K, M)).astype(np.complex64)
        B = rand(rng, (K, N)).astype(np.complex64)
        ha, hb = ctx.upload(A), ctx.upload(B)
        ctx.set_kernel_override(_lib.KC_GEMM16)
        try:
            hc = ctx.contract(ha, [1, -1], hb, [1, -2])
        finally:
            ctx.set_kernel_override(-1)
        got = ctx.download(hc)
        want = O.contract_tensors_data(A.astype(np.complex128), [1, -1], B.astype(np.complex128), [1, -2])
        scale = O.contract_tensors_data(np.abs(A).astype(np.complex128), [1, -1],
                                        np.abs(B).astype(np.complex128), [1, -2]).real
        assert (np.abs(got - want) / scale).max() <= GEMM16_TOL
        for h in (ha, hb, hc):
            ctx.free(h)
    finally:
        ctx.close()


def test_gemm16_arithmetic_setting_and_fallback():
    """pq_set_arithmetic(1) routes qualifying complex64 GEMMs to the split
    kernel; a K that is not a multiple of 32 stays on FP32."""
    ctx = _lib.Context(np.complex64)
    try:
        rng = np.random.default_rng(7)
        ctx.set_arithmetic(1)
        ctx.profile_enable(True)
        for K, cls in ((64, _lib.KC_GEMM16), (40, _lib.KC_GEMM)):
            A = rand(rng, (K, 96)).astype(np.complex64)
            B = rand(rng, (K, 64)).astype(np.complex64)
            check(ctx, A, [1, -1], B, [1, -2], gemm(ctx, A, [1, -1], B, [1, -2]))
            assert ctx.profile_read(cls)["launches"] >= 1
        with pytest.raises(_lib.PQError):
            ctx.set_arithmetic(2)
    finally:
        ctx.close()
