"""HDF5 tensor store for DSL record / replay (src/backends/dsl.jl:55-105,
src/layer1.jl:15-64), read and written through libhdf5's C API (ctypes).

The layout is the one HDF5.jl gives the reference's `write(file, label,
data)` / `read(file, label)`:

* one dataset per label at the file's root;
* the dataspace lists a Julia array's dims in reverse (HDF5 is row-major,
  Julia column-major), so the bytes are the array's column-major storage;
* complex elements are a compound type with float members "r" and "i"
  (HDF5.jl's complex support), real datasets are read as complex with a
  zero imaginary part (the reference converts on load, src/layer1.jl:28-30);
* a 0-dimensional array is a scalar dataspace.

The library is located by PQ_HDF5_LIB, then the loader's search path, then
the conda prefix that ships it with this image; without it, opening an
HDF5 store raises an error naming what was searched.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import os
import pathlib

import numpy as np

hid_t = ctypes.c_int64
herr_t = ctypes.c_int
htri_t = ctypes.c_int
hsize_t = ctypes.c_uint64

H5F_ACC_RDONLY = 0x0000
H5F_ACC_RDWR = 0x0001
H5F_ACC_TRUNC = 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0
H5T_INTEGER = 0
H5T_FLOAT = 1
H5T_COMPOUND = 6
H5E_DEFAULT = 0

_CANDIDATES = ("/opt/conda/lib/libhdf5.so", "/opt/conda/lib/libhdf5.so.103")
_lib = None


class HDF5Error(RuntimeError):
    pass


def _load():
    global _lib
    if _lib is not None:
        return _lib
    tried = []
    names = [os.environ.get("PQ_HDF5_LIB"), ctypes.util.find_library("hdf5"), *_CANDIDATES]
    for name in names:
        if not name:
            continue
        tried.append(name)
        try:
            lib = ctypes.CDLL(name)
        except OSError:
            continue
        break
    else:
        raise HDF5Error("libhdf5 not found (tried " + ", ".join(tried) + "; set PQ_HDF5_LIB)")
    sig = {
        "H5open": (herr_t, []),
        "H5Eset_auto2": (herr_t, [hid_t, ctypes.c_void_p, ctypes.c_void_p]),
        "H5Fcreate": (hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t, hid_t]),
        "H5Fopen": (hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t]),
        "H5Fclose": (herr_t, [hid_t]),
        "H5Lexists": (htri_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Ldelete": (herr_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Dopen2": (hid_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Dcreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
        "H5Dget_space": (hid_t, [hid_t]),
        "H5Dget_type": (hid_t, [hid_t]),
        "H5Dread": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
        "H5Dwrite": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
        "H5Dclose": (herr_t, [hid_t]),
        "H5Screate": (hid_t, [ctypes.c_int]),
        "H5Screate_simple": (hid_t, [ctypes.c_int, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]),
        "H5Sget_simple_extent_ndims": (ctypes.c_int, [hid_t]),
        "H5Sget_simple_extent_dims": (ctypes.c_int, [hid_t, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]),
        "H5Sclose": (herr_t, [hid_t]),
        "H5Tcreate": (hid_t, [ctypes.c_int, ctypes.c_size_t]),
        "H5Tinsert": (herr_t, [hid_t, ctypes.c_char_p, ctypes.c_size_t, hid_t]),
        "H5Tget_class": (ctypes.c_int, [hid_t]),
        "H5Tget_nmembers": (ctypes.c_int, [hid_t]),
        "H5Tget_member_name": (ctypes.c_void_p, [hid_t, ctypes.c_uint]),
        "H5Tget_member_class": (ctypes.c_int, [hid_t, ctypes.c_uint]),
        "H5Tclose": (herr_t, [hid_t]),
        "H5free_memory": (herr_t, [ctypes.c_void_p]),
        "H5Gget_info": (herr_t, [hid_t, ctypes.c_void_p]),
        "H5Lget_name_by_idx": (ctypes.c_ssize_t, [hid_t, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, hsize_t,
                                                  ctypes.c_char_p, ctypes.c_size_t, hid_t]),
    }
    for fn, (res, args) in sig.items():
        f = getattr(lib, fn)
        f.restype = res
        f.argtypes = args
    if lib.H5open() < 0:
        raise HDF5Error("H5open failed")
    lib.H5Eset_auto2(H5E_DEFAULT, None, None)   # errors are reported through return codes
    _lib = lib
    return lib


def _native(lib, name):
    return hid_t.in_dll(lib, name).value


def _check(v, what):
    if v < 0:
        raise HDF5Error(what + " failed")
    return v


def _complex_type(lib, double):
    """Compound {r, i} of native double (or float), as HDF5.jl writes complex."""
    base = _native(lib, "H5T_NATIVE_DOUBLE_g" if double else "H5T_NATIVE_FLOAT_g")
    size = 8 if double else 4
    t = _check(lib.H5Tcreate(H5T_COMPOUND, 2 * size), "H5Tcreate")
    _check(lib.H5Tinsert(t, b"r", 0, base), "H5Tinsert")
    _check(lib.H5Tinsert(t, b"i", size, base), "H5Tinsert")
    return t


class H5TensorStore:
    """Labelled tensors in one HDF5 file, in HDF5.jl's layout (see module
    docstring).  Same interface as the .npz TensorStore."""

    def __init__(self, path, overwrite=False):
        self.path = pathlib.Path(path)
        self.lib = _load()
        if overwrite or not self.path.exists():
            f = _check(self.lib.H5Fcreate(str(self.path).encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT),
                       f"creating {self.path}")
            self.lib.H5Fclose(f)

    def _open(self, write):
        if not self.path.exists():
            if not write:
                return None
            self.__init__(self.path, overwrite=True)
        return _check(self.lib.H5Fopen(str(self.path).encode(), H5F_ACC_RDWR if write else H5F_ACC_RDONLY,
                                       H5P_DEFAULT), f"opening {self.path}")

    def write(self, label, data):
        """write(file, label, data) after deleting an existing `label`
        (src/backends/dsl.jl:69-78)."""
        lib = self.lib
        a = np.asarray(data)
        double = a.dtype not in (np.complex64, np.float32)
        a = a.astype(np.complex128 if double else np.complex64)
        buf = np.ascontiguousarray(a.T)        # column-major bytes, dims reversed
        f = self._open(True)
        try:
            name = str(label).encode()
            if lib.H5Lexists(f, name, H5P_DEFAULT) > 0:
                _check(lib.H5Ldelete(f, name, H5P_DEFAULT), "H5Ldelete")
            if a.ndim == 0:
                space = _check(lib.H5Screate(H5S_SCALAR), "H5Screate")
            else:
                dims = (hsize_t * a.ndim)(*buf.shape)
                space = _check(lib.H5Screate_simple(a.ndim, dims, None), "H5Screate_simple")
            t = _complex_type(lib, double)
            try:
                d = _check(lib.H5Dcreate2(f, name, t, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                           f"creating dataset {label!r}")
                try:
                    _check(lib.H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.ctypes.data), "H5Dwrite")
                finally:
                    lib.H5Dclose(d)
            finally:
                lib.H5Tclose(t)
                lib.H5Sclose(space)
        finally:
            lib.H5Fclose(f)

    def read(self, label):
        """read(file, label) (src/backends/dsl.jl:96-101): the array, or None
        when the label is absent."""
        lib = self.lib
        f = self._open(False)
        if f is None:
            return None
        try:
            name = str(label).encode()
            if lib.H5Lexists(f, name, H5P_DEFAULT) <= 0:
                return None
            d = _check(lib.H5Dopen2(f, name, H5P_DEFAULT), f"opening dataset {label!r}")
            try:
                space = _check(lib.H5Dget_space(d), "H5Dget_space")
                try:
                    nd = _check(lib.H5Sget_simple_extent_ndims(space), "H5Sget_simple_extent_ndims")
                    dims = (hsize_t * max(nd, 1))()
                    if nd:
                        _check(lib.H5Sget_simple_extent_dims(space, dims, None), "H5Sget_simple_extent_dims")
                    hdims = tuple(int(dims[i]) for i in range(nd))
                finally:
                    lib.H5Sclose(space)
                ft = _check(lib.H5Dget_type(d), "H5Dget_type")
                try:
                    cls = lib.H5Tget_class(ft)
                    if cls == H5T_COMPOUND:
                        names = []
                        for k in range(lib.H5Tget_nmembers(ft)):
                            p = lib.H5Tget_member_name(ft, k)
                            names.append(ctypes.string_at(p).decode())
                            lib.H5free_memory(p)
                        if names != ["r", "i"]:
                            raise HDF5Error(f"dataset {label!r}: compound members {names}, expected ['r', 'i']")
                        kind = "complex"
                    elif cls in (H5T_FLOAT, H5T_INTEGER):
                        kind = "real"
                    else:
                        raise HDF5Error(f"dataset {label!r}: unsupported element class {cls}")
                finally:
                    lib.H5Tclose(ft)
                if kind == "complex":
                    out = np.empty(hdims, dtype=np.complex128)
                    mt = _complex_type(lib, True)
                    try:
                        _check(lib.H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data), "H5Dread")
                    finally:
                        lib.H5Tclose(mt)
                else:
                    re = np.empty(hdims, dtype=np.float64)
                    _check(lib.H5Dread(d, _native(lib, "H5T_NATIVE_DOUBLE_g"), H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                       re.ctypes.data), "H5Dread")
                    out = re.astype(np.complex128)
            finally:
                lib.H5Dclose(d)
        finally:
            lib.H5Fclose(f)
        return out.T      # Julia dims order; column-major in memory

    def labels(self):
        lib = self.lib
        f = self._open(False)
        if f is None:
            return []
        try:
            info = (ctypes.c_uint8 * 64)()         # H5G_info_t: storage type, nlinks, max order, mounted
            _check(lib.H5Gget_info(f, info), "H5Gget_info")
            nlinks = int.from_bytes(bytes(info[8:16]), "little")
            out = []
            for k in range(nlinks):
                n = lib.H5Lget_name_by_idx(f, b".", 0, 0, k, None, 0, H5P_DEFAULT)   # by name, increasing
                buf = ctypes.create_string_buffer(_check(n, "H5Lget_name_by_idx") + 1)
                lib.H5Lget_name_by_idx(f, b".", 0, 0, k, buf, n + 1, H5P_DEFAULT)
                out.append(buf.value.decode())
            return out
        finally:
            lib.H5Fclose(f)
