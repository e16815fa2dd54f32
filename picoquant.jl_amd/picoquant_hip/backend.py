"""HIPBackend: PicoQuant's AbstractBackend verbs on libpq_hip.

Mirrors src/backends.jl (AbstractBackend, Metrics, record_compute_costs!,
record_memory_costs!) and src/backends/interactive.jl (the nine verbs), with
the tensor storage on the GPU.  Julia's `name!` verbs are spelled `name` here.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import _lib


class AbstractBackend:
    """src/backends.jl:3.  Concrete backends carry a `metrics` field."""
    metrics: "Metrics"


@dataclass
class Metrics:
    """src/backends.jl:8-20: analytic ledger (elements / complex multiplications)."""
    max_tensor_size: int = 0
    total_space_allocated: int = 0
    flops: int = 0


def record_memory_costs(backend, tensor_memory: int):
    """src/backends.jl:50-57."""
    backend.metrics.total_space_allocated += tensor_memory
    if tensor_memory > backend.metrics.max_tensor_size:
        backend.metrics.max_tensor_size = tensor_memory


def record_compute_costs(backend, open_dims, contracted_dims):
    """src/backends.jl:30-41: space = prod(open), flops += space * prod(contracted)."""
    space = math.prod(open_dims)
    record_memory_costs(backend, space)
    fl = space * math.prod(contracted_dims)
    backend.metrics.flops += fl
    return fl, space


class HIPBackend(AbstractBackend):
    """InteractiveBackend{T} with device storage (src/backends/interactive.jl:10-172).

    dtype: np.complex128 (ComplexF64, the plain drop-in method) or np.complex64.
    Tensors live on the GPU under their Symbol label; load_tensor_data copies
    back to the host (the only synchronising verb).
    """

    def __init__(self, dtype=np.complex128, device=0, ctx: _lib.Context | None = None):
        self.ctx = ctx if ctx is not None else _lib.Context(dtype, device)
        self.dtype = self.ctx.np_dtype
        self.tensors: dict[str, int] = {}
        self.metrics = Metrics()

    # -- storage verbs ------------------------------------------------------
    def _set(self, label, handle):
        old = self.tensors.get(label)
        self.tensors[label] = handle
        if old is not None:
            self.ctx.free(old)

    def save_tensor_data(self, label, data):
        """src/backends/interactive.jl:32-36: store a converted copy."""
        self._set(str(label), self.ctx.upload(np.asarray(data)))

    def load_tensor_data(self, label):
        """src/backends/interactive.jl:44-49: the tensor, or None if absent."""
        h = self.tensors.get(str(label))
        if h is None:
            return None
        return self.ctx.download(h)

    def tensor_shape(self, label):
        """Shape without a data transfer (what slicing needs)."""
        return self.ctx.shape(self.tensors[str(label)])

    def delete_tensor(self, label):
        """src/backends/interactive.jl:159-161."""
        h = self.tensors.pop(str(label), None)
        if h is not None:
            self.ctx.free(h)

    # -- arithmetic verbs ---------------------------------------------------
    def contract_tensors(self, a, ia, b, ib, c):
        """src/backends/interactive.jl:60-75: C = contract(A, B); A and B deleted."""
        ha = self.tensors[str(a)]
        hb = self.tensors[str(b)]
        hc = self.ctx.contract(ha, ia, hb, ib)
        self._set(str(c), hc)
        self.delete_tensor(a)
        self.delete_tensor(b)

    def save_output(self, node, name="result"):
        """src/backends/interactive.jl:84-88: `name` refers to the node's tensor."""
        h = self.tensors[str(node)]
        n = len(self.ctx.shape(h))
        self._set(str(name), self.ctx.permute(h, list(range(1, n + 1))))

    def reshape_tensor(self, label, groups):
        """src/backends/interactive.jl:97-102: dims = prod of each 1-based group."""
        h = self.tensors[str(label)]
        shp = self.ctx.shape(h)
        dims = [math.prod(shp[g - 1] for g in grp) for grp in groups]
        self.ctx.reshape(h, dims)

    def permute_tensor(self, label, axes):
        """src/backends/interactive.jl:111-115 (permutedims; zero-copy strided view)."""
        h = self.tensors[str(label)]
        self._set(str(label), self.ctx.permute(h, list(axes)))

    def view_tensor(self, view_node, node, bond_idx, bond_range):
        """src/backends/interactive.jl:169-172: restrict axis bond_idx to lo:hi."""
        h = self.tensors[str(node)]
        lo, hi = bond_range
        self._set(str(view_node), self.ctx.view(h, bond_idx, lo, hi))

    def decompose_tensor(self, tensor, left_positions, right_positions, threshold=1e-13,
                         max_rank=0, left_label=None, right_label=None):
        """src/backends/interactive.jl:130-152 -> src/layer1.jl:146-190.

        On the GPU (pq_decompose: rocSOLVER gesdd, B = U sqrt(S), C = sqrt(S) V^H);
        returns chi, the new bond's dimension.
        """
        h = self.tensors[str(tensor)]
        hb, hc, chi = self.ctx.decompose(h, list(left_positions), list(right_positions), threshold, max_rank)
        self._set(str(left_label), hb)
        self._set(str(right_label), hc)
        if str(tensor) not in (str(left_label), str(right_label)):
            self.delete_tensor(tensor)
        return chi

