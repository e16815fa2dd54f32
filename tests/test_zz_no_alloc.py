"""pq_contract_noalloc through the C ABI: the same results as pq_contract
(bit for bit: the same kernels and launch shapes), on contractions whose
kernels use scratch -- split-K GEMMs (tile-starved outputs, long K) and
split reductions (tiny outputs, long K) -- and on one that does not; two
calls in a row on every case, the second on a smaller input, so the shared
device scratch keeps the first call's partials when the second runs."""
import ctypes as ct

import numpy as np
import pytest

from picoquant_hip import _lib

pytestmark = pytest.mark.gpu


def _rand(rng, shape, dtype):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)


# (A shape, A labels, B shape, B labels), second call smaller
CASES = {
    "splitk_gemm": [((16384, 32), [1, -1], (64, 16384), [-2, 1]),
                    ((8192, 24), [1, -1], (40, 8192), [-2, 1])],
    "splitk_two_k_modes": [((64, 256, 32), [-1, 2, 1], (32, 16, 256), [1, -2, 2]),
                           ((48, 128, 32), [-1, 2, 1], (32, 16, 128), [1, -2, 2])],
    "reduce": [((3, 40000), [-1, 1], (40000, 4), [1, -2]),
               ((2, 5000), [-1, 1], (5000, 3), [1, -2])],
    "stream": [((2, 2, 4096), [1, 2, -1], (2, 2), [2, 1]),
               ((2, 2, 256), [1, 2, -1], (2, 2), [2, 1])],
}


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64], ids=["c128", "c64"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_noalloc_matches_contract(case, dtype):
    ctx = _lib.Context(dtype)
    lib = ctx.lib
    rng = np.random.default_rng(len(case))
    try:
        for sa, ia, sb, ib in CASES[case]:
            A, B = _rand(rng, sa, dtype), _rand(rng, sb, dtype)
            ha, hb = ctx.upload(A), ctx.upload(B)
            want_h = ct.c_int64()
            ctx.check(lib.pq_contract(ctx.ptr, ha, len(ia), _lib.i32arr(ia), hb, len(ib), _lib.i32arr(ib),
                                      ct.byref(want_h)))
            want = ctx.download(want_h.value)
            hc = ctx.upload(np.full(want.shape, 3 + 4j, dtype=dtype))     # stale values to overwrite
            ctx.check(lib.pq_contract_noalloc(ctx.ptr, ha, len(ia), _lib.i32arr(ia), hb, len(ib),
                                              _lib.i32arr(ib), hc))
            got = ctx.download(hc)
            np.testing.assert_array_equal(got, want)
            for h in (ha, hb, hc, want_h.value):
                ctx.free(h)
    finally:
        ctx.close()


def test_noalloc_rejects_a_wrong_result_tensor():
    ctx = _lib.Context(np.complex128)
    lib = ctx.lib
    try:
        ha = ctx.upload(np.ones((4, 8), dtype=np.complex128))
        hb = ctx.upload(np.ones((8, 2), dtype=np.complex128))
        ia, ib = [-1, 1], [1, -2]
        for bad in (np.zeros((4, 3), dtype=np.complex128), np.zeros((8,), dtype=np.complex128)):
            hc = ctx.upload(bad)
            rc = lib.pq_contract_noalloc(ctx.ptr, ha, 2, _lib.i32arr(ia), hb, 2, _lib.i32arr(ib), hc)
            assert rc == _lib.PQ_ERR_ARG
            ctx.free(hc)
        # the result may not share an operand's buffer
        rc = lib.pq_contract_noalloc(ctx.ptr, ha, 2, _lib.i32arr(ia), hb, 2, _lib.i32arr(ib), ha)
        assert rc == _lib.PQ_ERR_ARG
        hc = ctx.upload(np.zeros((4, 2), dtype=np.complex128))
        ctx.contract_into(ha, ia, hb, ib, hc)
        np.testing.assert_array_equal(ctx.download(hc), np.full((4, 2), 8.0))
    finally:
        ctx.close()
