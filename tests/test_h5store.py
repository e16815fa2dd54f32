"""HDF5 tensor store (picoquant_hip.h5store) in the layout HDF5.jl gives the
reference's DSL tensor files (src/backends/dsl.jl:69-101, src/layer1.jl:
28-63): round trips, and the on-disk layout read back by h5dump (HDF5's
own tool, an independent reader): dataspace dims are the array's dims
reversed, elements a compound {"r", "i"} of IEEE floats, the values in
the array's column-major order."""
import ctypes
import pathlib
import re
import shutil
import subprocess

import numpy as np
import pytest

h5store = pytest.importorskip("picoquant_hip.h5store")
try:
    h5store._load()
except h5store.HDF5Error as e:       # pragma: no cover - the image ships libhdf5
    pytest.skip(str(e), allow_module_level=True)

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if pathlib.Path("/opt/conda/bin/h5dump").exists()
                                    else None)


def rand(shape, dtype, seed=0):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("shape", [(), (2,), (2, 3, 4), (2, 2, 2, 2, 1), (5, 1, 3)])
def test_round_trip(tmp_path, dtype, shape):
    s = h5store.H5TensorStore(tmp_path / "t.h5", overwrite=True)
    a = rand(shape, dtype)
    s.write("node_1", a)
    b = s.read("node_1")
    assert b.shape == a.shape
    np.testing.assert_array_equal(b, a.astype(np.complex128))


def test_overwrite_missing_and_labels(tmp_path):
    p = tmp_path / "t.h5"
    s = h5store.H5TensorStore(p, overwrite=True)
    assert s.read("x") is None and s.labels() == []
    s.write("node_2", rand((2, 2), np.complex128, 1))
    s.write("node_1", rand((4,), np.complex128, 2))
    new = rand((3, 2), np.complex128, 3)
    s.write("node_2", new)            # the label is replaced, shape and all
    assert sorted(s.labels()) == ["node_1", "node_2"]
    np.testing.assert_array_equal(h5store.H5TensorStore(p).read("node_2"), new)   # reopened
    assert h5store.H5TensorStore(tmp_path / "absent.h5").read("node_1") is None
    h5store.H5TensorStore(p, overwrite=True)
    assert h5store.H5TensorStore(p).labels() == []


def test_real_dataset_reads_as_complex(tmp_path):
    """A dataset of plain doubles (as a writer without complex support makes)
    loads as complex with zero imaginary parts, as the reference's
    convert(T, read(...)) does."""
    lib = h5store._load()
    p = tmp_path / "real.h5"
    h5store.H5TensorStore(p, overwrite=True)
    f = lib.H5Fopen(str(p).encode(), h5store.H5F_ACC_RDWR, 0)
    a = np.arange(6, dtype=np.float64).reshape(2, 3)          # the "Julia" array
    buf = np.ascontiguousarray(a.T)
    dims = (h5store.hsize_t * 2)(*buf.shape)
    sp = lib.H5Screate_simple(2, dims, None)
    dbl = h5store.hid_t.in_dll(lib, "H5T_NATIVE_DOUBLE_g").value
    d = lib.H5Dcreate2(f, b"gate", dbl, sp, 0, 0, 0)
    assert d >= 0 and lib.H5Dwrite(d, dbl, 0, 0, 0, buf.ctypes.data) >= 0
    lib.H5Dclose(d), lib.H5Sclose(sp), lib.H5Fclose(f)
    got = h5store.H5TensorStore(p).read("gate")
    assert got.dtype == np.complex128
    np.testing.assert_array_equal(got, a)


def test_bad_compound_is_rejected(tmp_path):
    lib = h5store._load()
    p = tmp_path / "bad.h5"
    h5store.H5TensorStore(p, overwrite=True)
    f = lib.H5Fopen(str(p).encode(), h5store.H5F_ACC_RDWR, 0)
    dbl = h5store.hid_t.in_dll(lib, "H5T_NATIVE_DOUBLE_g").value
    t = lib.H5Tcreate(h5store.H5T_COMPOUND, 16)
    lib.H5Tinsert(t, b"re", 0, dbl), lib.H5Tinsert(t, b"im", 8, dbl)
    dims = (h5store.hsize_t * 1)(2)
    sp = lib.H5Screate_simple(1, dims, None)
    d = lib.H5Dcreate2(f, b"x", t, sp, 0, 0, 0)
    buf = np.zeros(2, dtype=np.complex128)
    assert lib.H5Dwrite(d, t, 0, 0, 0, buf.ctypes.data) >= 0
    lib.H5Dclose(d), lib.H5Sclose(sp), lib.H5Tclose(t), lib.H5Fclose(f)
    with pytest.raises(h5store.HDF5Error, match="compound members"):
        h5store.H5TensorStore(p).read("x")


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
@pytest.mark.parametrize("dtype,ftype", [(np.complex128, "H5T_IEEE_F64LE"), (np.complex64, "H5T_IEEE_F32LE")])
def test_layout_read_by_h5dump(tmp_path, dtype, ftype):
    p = tmp_path / "lay.h5"
    a = (np.arange(24).reshape(2, 3, 4) + 1j * (100 + np.arange(24).reshape(2, 3, 4))).astype(dtype)
    h5store.H5TensorStore(p, overwrite=True).write("node_7", a)
    out = subprocess.run([H5DUMP, "-d", "/node_7", str(p)], capture_output=True, text=True, check=True).stdout
    assert re.search(r'H5T_COMPOUND\s*{\s*' + ftype + r'\s+"r";\s*' + ftype + r'\s+"i";\s*}', out), out
    assert re.search(r"DATASPACE\s+SIMPLE\s*{\s*\(\s*4,\s*3,\s*2\s*\)", out), out
    body = out[out.index("DATA {"):]
    pairs = [(float(r), float(i)) for r, i in re.findall(r"{\s*([-\d.e+]+),\s*([-\d.e+]+)\s*}", body)]
    want = a.reshape(-1, order="F")
    assert pairs == [(w.real, w.imag) for w in want]
