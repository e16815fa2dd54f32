"""MFMA GEMM path (PQ_KC_GEMM) against the oracle: every tile shape, both
operand access orders, permuted and sliced operands, and the boundary x site
shape of the grid sweep plans.  Tolerances as in test_gpu_contract."""
import numpy as np
import pytest

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
