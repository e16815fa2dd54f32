"""ctypes binding to libpq_hip.so (the C ABI in include/pq_hip.h).

There is no fallback: if the library is missing or no device is present the
calls raise, so nothing silently runs on the CPU.
"""
from __future__ import annotations

import ctypes as ct
import os
import pathlib

import numpy as np

# PQ_HIP_LIB names another build of the library (same C ABI).
LIB_PATH = pathlib.Path(os.environ.get("PQ_HIP_LIB") or pathlib.Path(__file__).resolve().parent / "libpq_hip.so")

PQ_OK = 0
PQ_ERR_ARG = 1
PQ_ERR_HIP = 2
PQ_ERR_NOT_FOUND = 3
PQ_ERR_OOM = 4
PQ_ERR_UNSUPPORTED = 5

PQ_C128 = 0
PQ_C64 = 1
PQ_MAX_RANK = 64
PQ_GEMM_TILES = 15

KC_STREAM, KC_GEMM, KC_GENERIC, KC_REDUCE, KC_COPY, KC_BATCHED, KC_ACCUM, KC_GEMM16, KC_CHAIN = range(9)
KC_NAMES = {KC_STREAM: "stream", KC_GEMM: "gemm", KC_GENERIC: "generic", KC_REDUCE: "reduce",
            KC_COPY: "copy", KC_BATCHED: "batched", KC_ACCUM: "accum", KC_GEMM16: "gemm16",
            KC_CHAIN: "chain"}
PQ_CHAIN_MAX_STAGES = 8
PQ_CHAIN_MAX_TILE_BITS = 13
PQ_CHAIN_MAX_OP_BITS = 5
PQ_CHAIN_STAGE_BYTES = 16384

class Cost(ct.Structure):
    """pq_cost (include/pq_hip.h)."""
    _fields_ = [("macs", ct.c_double), ("bytes", ct.c_double), ("steps", ct.c_int64), ("launches", ct.c_int64)]


# Every exported symbol, with (restype, argtypes).  tests/test_boundary.py
# checks this list against include/pq_hip.h.
_P = ct.c_void_p
_I64P = ct.POINTER(ct.c_int64)
_I32P = ct.POINTER(ct.c_int32)
SIGNATURES = {
    "pq_ctx_create": (ct.c_int, [ct.c_int, ct.c_int, ct.POINTER(_P)]),
    "pq_ctx_destroy": (ct.c_int, [_P]),
    "pq_last_error": (ct.c_char_p, [_P]),
    "pq_ctx_dtype": (ct.c_int, [_P]),
    "pq_synchronize": (ct.c_int, [_P]),
    "pq_tensor_upload": (ct.c_int, [_P, ct.c_int, _I64P, _P, ct.c_int, _I64P]),
    "pq_tensor_shape": (ct.c_int, [_P, ct.c_int64, ct.POINTER(ct.c_int), _I64P]),
    "pq_tensor_download": (ct.c_int, [_P, ct.c_int64, _P]),
    "pq_tensor_free": (ct.c_int, [_P, ct.c_int64]),
    "pq_tensor_write": (ct.c_int, [_P, ct.c_int64, _P, ct.c_int]),
    "pq_contract": (ct.c_int, [_P, ct.c_int64, ct.c_int, _I32P, ct.c_int64, ct.c_int, _I32P, _I64P]),
    "pq_contract_noalloc": (ct.c_int, [_P, ct.c_int64, ct.c_int, _I32P, ct.c_int64, ct.c_int, _I32P, ct.c_int64]),
    "pq_set_plan_streams": (ct.c_int, [_P, ct.c_int]),
    "pq_permute": (ct.c_int, [_P, ct.c_int64, _I32P, _I64P]),
    "pq_decompose": (ct.c_int, [_P, ct.c_int64, ct.c_int, _I32P, ct.c_int, _I32P, ct.c_double, ct.c_int,
                                _I64P, _I64P, ct.POINTER(ct.c_int)]),
    "pq_reshape": (ct.c_int, [_P, ct.c_int64, ct.c_int, _I64P]),
    "pq_view": (ct.c_int, [_P, ct.c_int64, ct.c_int, ct.c_int64, ct.c_int64, _I64P]),
    "pq_materialize": (ct.c_int, [_P, ct.c_int64, _I64P]),
    "pq_tensor_device_ptr": (ct.c_int, [_P, ct.c_int64, ct.POINTER(_P), _I64P, _I64P]),
    "pq_profile_enable": (ct.c_int, [_P, ct.c_int]),
    "pq_profile_read": (ct.c_int, [_P, ct.c_int, _I64P, ct.POINTER(ct.c_double),
                                   ct.POINTER(ct.c_double), ct.POINTER(ct.c_double)]),
    "pq_set_kernel_override": (ct.c_int, [_P, ct.c_int]),
    "pq_set_gemm_tile": (ct.c_int, [_P, ct.c_int]),
    "pq_set_gemm_orientation": (ct.c_int, [_P, ct.c_int]),
    "pq_set_chain_pairs": (ct.c_int, [_P, ct.c_int]),
    "pq_set_arithmetic": (ct.c_int, [_P, ct.c_int]),
    "pq_plan_create": (ct.c_int, [_P, _P, ct.POINTER(_P)]),
    "pq_plan_execute": (ct.c_int, [_P, _P, ct.c_int]),
    "pq_plan_result": (ct.c_int, [_P, _P, ct.POINTER(ct.c_double)]),
    "pq_plan_stats": (ct.c_int, [_P, _I64P, _I64P, ct.POINTER(ct.c_double), ct.POINTER(ct.c_double), _I64P]),
    "pq_plan_step_info": (ct.c_int, [_P, ct.c_int, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int),
                                     ct.POINTER(ct.c_int), ct.POINTER(ct.c_double), ct.POINTER(ct.c_double)]),
    "pq_plan_profile": (ct.c_int, [_P, _P, ct.POINTER(ct.c_double)]),
    "pq_plan_destroy": (ct.c_int, [_P, _P]),
    "pq_plan_cost": (ct.c_int, [_P, _P, ct.POINTER(ct.c_double), ct.POINTER(ct.c_double)]),
    "pq_plan_set_bitstrings": (ct.c_int, [_P, _P, ct.c_int, ct.POINTER(ct.c_uint64)]),
    "pq_plan_run": (ct.c_int, [_P, _P, ct.c_int, ct.POINTER(ct.c_uint64), ct.POINTER(ct.c_double)]),
    "pq_time_step": (ct.c_int, [_P, ct.c_int, ct.c_int, ct.c_int, ct.c_int, _I64P, ct.c_int, ct.c_int, ct.c_int,
                                ct.c_int, ct.POINTER(ct.c_double)]),
    "pq_mlp_create": (ct.c_int, [_P, ct.c_int, ct.c_int, ct.c_uint64, ct.POINTER(_P)]),
    "pq_mlp_destroy": (ct.c_int, [_P]),
    "pq_mlp_num_params": (ct.c_int, [_P]),
    "pq_mlp_get_params": (ct.c_int, [_P, _P, ct.POINTER(ct.c_float)]),
    "pq_mlp_set_params": (ct.c_int, [_P, _P, ct.POINTER(ct.c_float)]),
    "pq_mlp_predict": (ct.c_int, [_P, _P, ct.POINTER(ct.c_float), ct.c_int, ct.POINTER(ct.c_float)]),
    "pq_mlp_loss_grad": (ct.c_int, [_P, _P, ct.POINTER(ct.c_float), ct.POINTER(ct.c_float), ct.c_int,
                                    ct.POINTER(ct.c_float), ct.POINTER(ct.c_float)]),
    "pq_mlp_train": (ct.c_int, [_P, _P, ct.POINTER(ct.c_float), ct.POINTER(ct.c_float), ct.c_int, ct.c_int,
                                ct.c_int, ct.c_float, ct.c_uint64, ct.POINTER(ct.c_float)]),
    "pq_netcon": (ct.c_int, [ct.c_int, ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int64),
                             ct.c_int64, ct.POINTER(ct.c_int32), ct.POINTER(ct.c_double)]),
}

_lib = None


class PQError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libpq_hip error {code}: {msg}")
        self.code = code


class PQKeyError(PQError, KeyError):
    pass


def load(required_symbols=True):
    """Load libpq_hip.so and bind every symbol in SIGNATURES."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with `python picoquant.jl_amd/build_lib.py`")
    lib = ct.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def i64arr(xs):
    a = (ct.c_int64 * max(1, len(xs)))(*[int(x) for x in xs])
    return a


def i32arr(xs):
    return (ct.c_int32 * max(1, len(xs)))(*[int(x) for x in xs])


class Context:
    """One device context (InteractiveBackend{T}'s storage + one HIP stream)."""

    def __init__(self, dtype=np.complex128, device=0):
        self.lib = load()
        self.np_dtype = np.dtype(dtype)
        if self.np_dtype == np.complex128:
            self.code = PQ_C128
        elif self.np_dtype == np.complex64:
            self.code = PQ_C64
        else:
            raise TypeError(f"unsupported element type {dtype}")
        p = ct.c_void_p()
        rc = self.lib.pq_ctx_create(int(device), self.code, ct.byref(p))
        if rc != PQ_OK:
            raise PQError(rc, "pq_ctx_create failed (no HIP device?)")
        self.ptr = p

    def close(self):
        if getattr(self, "ptr", None):
            self.lib.pq_ctx_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != PQ_OK:
            msg = self.lib.pq_last_error(self.ptr).decode()
            if rc == PQ_ERR_NOT_FOUND:
                raise PQKeyError(rc, msg)
            raise PQError(rc, msg)

    # ---- tensors
    def upload(self, data):
        a = np.asarray(data)
        if a.dtype == np.complex64:
            host, code = np.asfortranarray(a), PQ_C64
        else:
            host, code = np.asfortranarray(a.astype(np.complex128, copy=False)), PQ_C128
        dims = i64arr(a.shape)
        h = ct.c_int64()
        self.check(self.lib.pq_tensor_upload(self.ptr, a.ndim, dims, host.ctypes.data_as(ct.c_void_p),
                                             code, ct.byref(h)))
        return h.value

    def shape(self, h):
        r = ct.c_int()
        dims = (ct.c_int64 * PQ_MAX_RANK)()
        self.check(self.lib.pq_tensor_shape(self.ptr, h, ct.byref(r), dims))
        return tuple(int(dims[i]) for i in range(r.value))

    def download(self, h):
        shp = self.shape(h)
        out = np.empty(int(np.prod(shp)) if shp else 1, dtype=self.np_dtype)
        self.check(self.lib.pq_tensor_download(self.ptr, h, out.ctypes.data_as(ct.c_void_p)))
        return out.reshape(shp, order="F")

    def write(self, h, data):
        """Overwrite tensor h's elements in place (pq_tensor_write)."""
        a = np.asarray(data)
        if tuple(a.shape) != self.shape(h):
            raise ValueError(f"shape {a.shape} does not match the tensor's {self.shape(h)}")
        if a.dtype == np.complex64:
            host, code = np.asfortranarray(a), PQ_C64
        else:
            host, code = np.asfortranarray(a.astype(np.complex128, copy=False)), PQ_C128
        self.check(self.lib.pq_tensor_write(self.ptr, h, host.ctypes.data_as(ct.c_void_p), code))

    def free(self, h):
        self.check(self.lib.pq_tensor_free(self.ptr, h))

    def contract(self, ha, ia, hb, ib):
        out = ct.c_int64()
        self.check(self.lib.pq_contract(self.ptr, ha, len(ia), i32arr(ia), hb, len(ib), i32arr(ib),
                                        ct.byref(out)))
        return out.value

    def set_plan_streams(self, n):
        """Streams that plans created afterwards spread their launches over."""
        self.check(self.lib.pq_set_plan_streams(self.ptr, int(n)))

    def contract_into(self, ha, ia, hb, ib, hc):
        """pq_contract_noalloc: the contraction written into the existing
        contiguous tensor hc (no device allocation during the call)."""
        self.check(self.lib.pq_contract_noalloc(self.ptr, ha, len(ia), i32arr(ia), hb, len(ib), i32arr(ib), hc))

    def decompose(self, h, left, right, threshold=1e-13, max_rank=0):
        """SVD split (pq_decompose): (B, C, chi) for 1-based axis lists."""
        b, c, chi = ct.c_int64(), ct.c_int64(), ct.c_int()
        self.check(self.lib.pq_decompose(self.ptr, h, len(left), i32arr(left or [0]), len(right),
                                         i32arr(right or [0]), float(threshold), int(max_rank),
                                         ct.byref(b), ct.byref(c), ct.byref(chi)))
        return b.value, c.value, chi.value

    def permute(self, h, perm):
        out = ct.c_int64()
        self.check(self.lib.pq_permute(self.ptr, h, i32arr(perm), ct.byref(out)))
        return out.value

    def reshape(self, h, dims):
        self.check(self.lib.pq_reshape(self.ptr, h, len(dims), i64arr(dims)))

    def view(self, h, axis, lo, hi):
        out = ct.c_int64()
        self.check(self.lib.pq_view(self.ptr, h, int(axis), int(lo), int(hi), ct.byref(out)))
        return out.value

    def materialize(self, h):
        out = ct.c_int64()
        self.check(self.lib.pq_materialize(self.ptr, h, ct.byref(out)))
        return out.value

    def synchronize(self):
        self.check(self.lib.pq_synchronize(self.ptr))

    def set_kernel_override(self, kc):
        self.check(self.lib.pq_set_kernel_override(self.ptr, int(kc)))

    def set_gemm_tile(self, tile):
        self.check(self.lib.pq_set_gemm_tile(self.ptr, int(tile)))

    def set_arithmetic(self, mode):
        """complex64 GEMM arithmetic: 0 FP32 matrix cores, 1 FP16 hi/lo split."""
        self.check(self.lib.pq_set_arithmetic(self.ptr, int(mode)))

    def set_gemm_orientation(self, mode):
        """-1 automatic, 0 as given, 1 transposed (C^T = B^T A^T)."""
        self.check(self.lib.pq_set_gemm_orientation(self.ptr, int(mode)))

    def set_chain_pairs(self, mode):
        """Fused chains: -1 automatic 16-byte element-pair moves, 0 always two 8-byte moves."""
        self.check(self.lib.pq_set_chain_pairs(self.ptr, int(mode)))

    def time_step(self, M, N, K, batch=1, kclass=-1, tile=-1, reps=3, dtype=None):
        """Median microseconds of one contraction shape (modes as (ext, s0, s1))."""
        code = self.code if dtype is None else dtype
        modes = [x for grp in (M, N, K) for m in grp for x in m]
        us = ct.c_double()
        self.check(self.lib.pq_time_step(self.ptr, code, len(M), len(N), len(K), i64arr(modes or [0]), batch,
                                         kclass, tile, reps, ct.byref(us)))
        return us.value

    def profile_enable(self, on=True):
        self.check(self.lib.pq_profile_enable(self.ptr, 1 if on else 0))

    def profile_read(self, kc):
        n = ct.c_int64()
        ms, b, f = ct.c_double(), ct.c_double(), ct.c_double()
        self.check(self.lib.pq_profile_read(self.ptr, int(kc), ct.byref(n), ct.byref(ms), ct.byref(b),
                                            ct.byref(f)))
        return {"launches": n.value, "ms": ms.value, "bytes": b.value, "flops": f.value}
