"""DSL record / replay (src/backends/dsl.jl, src/layer1.jl:211-315).

DSLBackend records the backend verbs a driver issues as the reference's
text commands (one per line, a `.tl` file):

    tensor <label> <data label>            load from the tensor store
    ncon <C> <A> <a1,a2,..> <B> <b1,..>    contract (ncon labels)
    del <label>
    save <label> <file> <name>             write a tensor to an output store
    reshape <label> <g1,g2;g3;..>          1-based index groups
    permute <label> <p1,p2,..>
    view <new> <label> <axis> <lo,hi>
    decompose <A> <B> <b-axes> <C> <c-axes> {"threshold":..,"max_rank":..}

execute_dsl_file replays such a file on the GPU through libpq_hip: an
offline, Julia-free driver for a recorded contraction.  The command text is
the reference's format.  The tensor store is chosen by the file's suffix:
`.h5` / `.hdf5` is HDF5 in the layout the reference's HDF5.jl writes
(h5store.H5TensorStore, through libhdf5), so stores recorded by the
reference replay here; any other name is a numpy `.npz` archive, one array
per label, axis i = the tensor's index i.
"""
from __future__ import annotations

import json
import math
import pathlib

import numpy as np

from . import _lib
from .backend import AbstractBackend, Metrics


class TensorStore:
    """Labelled arrays in one .npz file (loaded without pickle)."""

    def __init__(self, path, overwrite=False):
        self.path = pathlib.Path(path)
        if overwrite or not self.path.exists():
            self._write({})

    def _read(self):
        if not self.path.exists():
            return {}
        with np.load(self.path, allow_pickle=False) as f:
            return {k: f[k] for k in f.files}

    def _write(self, d):
        tmp = self.path.with_name(self.path.name + ".tmp.npz")
        np.savez(tmp, **d)
        tmp.replace(self.path)

    def write(self, label, data):
        d = self._read()
        d[str(label)] = np.asarray(data)
        self._write(d)

    def read(self, label):
        return self._read().get(str(label))

    def labels(self):
        return list(self._read())


def open_store(path, overwrite=False):
    """The tensor store for `path`: HDF5 for .h5 / .hdf5, else .npz."""
    if pathlib.Path(path).suffix.lower() in (".h5", ".hdf5"):
        from .h5store import H5TensorStore
        return H5TensorStore(path, overwrite)
    return TensorStore(path, overwrite)


class DSLBackend(AbstractBackend):
    """src/backends/dsl.jl:14-44: the verbs append commands to `dsl_filename`;
    tensor data goes to the store `tensor_data_filename`; results are written
    by the replay to `output_data_filename` (the tensor store when empty)."""

    def __init__(self, dsl_filename="contract_network.tl", tensor_data_filename="tensor_data.npz",
                 output_data_filename="", overwrite=False):
        self.dsl_filename = str(dsl_filename)
        self.tensor_data_filename = str(tensor_data_filename)
        self.output_data_filename = str(output_data_filename or tensor_data_filename)
        p = pathlib.Path(self.dsl_filename)
        if overwrite or not p.exists():
            p.write_text("")
        self.store = open_store(self.tensor_data_filename, overwrite)
        self.metrics = Metrics()

    def push(self, instruction):
        """push!(dsl, instruction) (src/backends/dsl.jl:52-56)."""
        with open(self.dsl_filename, "a") as f:
            f.write(instruction + "\n")

    def save_tensor_data(self, label, data):
        """src/backends/dsl.jl:68-84."""
        self.store.write(label, data)
        self.push(f"tensor {label} {label}")

    def load_tensor_data(self, label):
        """src/backends/dsl.jl:92-106: `result` is read from the output store."""
        if str(label) == "result":
            return open_store(self.output_data_filename).read(label)
        return self.store.read(label)

    def contract_tensors(self, a, ia, b, ib, c):
        """src/backends/dsl.jl:117-131."""
        self.push(f"ncon {c} {a} {','.join(map(str, ia))} {b} {','.join(map(str, ib))}")
        self.push(f"del {a}")
        self.push(f"del {b}")

    def save_output(self, node, name="result"):
        """src/backends/dsl.jl:139-142."""
        self.push(f"save {node} {self.output_data_filename} {name}")

    def reshape_tensor(self, label, groups):
        """src/backends/dsl.jl:151-156."""
        self.push(f"reshape {label} " + ";".join(",".join(map(str, g)) for g in groups))

    def permute_tensor(self, label, axes):
        """src/backends/dsl.jl:164-169."""
        self.push(f"permute {label} " + ",".join(map(str, axes)))

    def decompose_tensor(self, tensor, left_positions, right_positions, threshold=1e-13, max_rank=0,
                         left_label=None, right_label=None):
        """src/backends/dsl.jl:185-197: the rank is known only at run time (0)."""
        self.push(f"decompose {tensor} {left_label} {','.join(map(str, left_positions))} "
                  f"{right_label} {','.join(map(str, right_positions))} "
                  + json.dumps({"threshold": threshold, "max_rank": max_rank}))
        return 0

    def delete_tensor(self, label):
        """src/backends/dsl.jl:202-204."""
        self.push(f"del {label}")

    def view_tensor(self, view_node, node, bond_idx, bond_range):
        """src/backends/dsl.jl:211-214."""
        self.push(f"view {view_node} {node} {bond_idx} {','.join(map(str, bond_range))}")


def execute_dsl_file(dtype=np.complex64, dsl_filename="contract_network.tl", tensor_data_filename="tensor_data.npz",
                     output_data_filename="", device=0, ctx=None):
    """execute_dsl_file (src/layer1.jl:211-315) on the GPU: every command
    runs through libpq_hip; `save` downloads into the output store (the
    file named by the command unless output_data_filename is given).
    Returns the context's remaining tensor labels (handles are freed)."""
    own = ctx is None
    ctx = ctx if ctx is not None else _lib.Context(dtype, device)
    store = open_store(tensor_data_filename)
    tensors = {}

    def put(label, h):
        old = tensors.get(label)
        tensors[label] = h
        if old is not None and old != h:
            ctx.free(old)

    try:
        with open(dsl_filename) as f:
            for lineno, line in enumerate(f, 1):
                cmd = line.split()
                if not cmd:
                    continue
                op = cmd[0]
                if op == "ncon":
                    c, a, ia, b, ib = cmd[1:6]
                    hc = ctx.contract(tensors[a], [int(x) for x in ia.split(",")],
                                      tensors[b], [int(x) for x in ib.split(",")])
                    put(c, hc)
                elif op == "del":
                    h = tensors.pop(cmd[1], None)
                    if h is not None:
                        ctx.free(h)
                elif op == "tensor":
                    data = store.read(cmd[2])
                    if data is None:
                        raise KeyError(f"line {lineno}: no tensor {cmd[2]!r} in {tensor_data_filename}")
                    put(cmd[1], ctx.upload(np.asarray(data)))
                elif op == "save":
                    out = output_data_filename or cmd[2]
                    name = cmd[3] if len(cmd) > 3 else cmd[1]
                    open_store(out).write(name, ctx.download(tensors[cmd[1]]))
                elif op == "reshape":
                    h = tensors[cmd[1]]
                    shp = ctx.shape(h)
                    groups = [[int(x) for x in g.split(",")] for g in cmd[2].split(";")]
                    ctx.reshape(h, [math.prod(shp[g - 1] for g in grp) for grp in groups])
                elif op == "permute":
                    put(cmd[1], ctx.permute(tensors[cmd[1]], [int(x) for x in cmd[2].split(",")]))
                elif op == "view":
                    lo, hi = (int(x) for x in cmd[4].split(","))
                    put(cmd[1], ctx.view(tensors[cmd[2]], int(cmd[3]), lo, hi))
                elif op == "decompose":
                    a, bl, bi, cl, ci = cmd[1:6]
                    opts = json.loads(" ".join(cmd[6:]) or "{}")
                    hb, hc, _ = ctx.decompose(tensors[a], [int(x) for x in bi.split(",")],
                                              [int(x) for x in ci.split(",")], float(opts.get("threshold", 1e-13)),
                                              int(opts.get("max_rank", 0)))
                    put(bl, hb)
                    put(cl, hc)
                else:
                    raise ValueError(f"line {lineno}: unknown DSL command {op!r}")
        return sorted(tensors)
    finally:
        for h in tensors.values():
            ctx.free(h)
        if own:
            ctx.close()
