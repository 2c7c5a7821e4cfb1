"""HDF5 (.mat) files in SeisCL's conventions, without h5py.

ctypes binding of the HDF5 helpers in libseiscl_cli.so (include/seiscl_cli.h).
SeisCL.py stores every array transposed (np.transpose, SeisCL/SeisCL.py:
658-763) so that the C reader sees the reference's host layout;
``savemat``/``loadmat`` apply the same transposition, ``write``/``read`` do
not.  Python ints are stored as int64 and floats as float64 scalars, the
types h5py gives them; float32 arrays keep their type.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._lib import SeisError

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libseiscl_cli.so")
BIN_PATH = os.path.join(_HERE, "bin", "SeisCL_MPI")

F32, F64, I32, I64 = 0, 1, 2, 3
_CODE = {np.dtype(np.float32): F32, np.dtype(np.float64): F64,
         np.dtype(np.int32): I32, np.dtype(np.int64): I64}
_DTYPE = {v: k for k, v in _CODE.items()}
MAXDIM = 8

_lib = None


def load():
    """The CLI/HDF5 library; raises SeisError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SeisError(f"{LIB_PATH} is missing: build it with make -C seiscl_amd/cli")
    lib = ctypes.CDLL(LIB_PATH)
    c, P, I, LL = ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong
    LLP = ctypes.POINTER(LL)
    for name, res, args in [
        ("seis_h5_last_error", c, []),
        ("seis_h5_file_create", I, [c]),
        ("seis_h5_file_write", I, [c, c, I, P, I, LLP]),
        ("seis_h5_file_dims", I, [c, c, ctypes.POINTER(I), LLP]),
        ("seis_h5_file_read", I, [c, c, I, P, LL]),
        ("seis_h5_file_list", I, [c, c, ctypes.c_size_t]),
        ("seis_cli_main", I, [I, ctypes.POINTER(c)]),
    ]:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(r):
    if r != 0:
        raise SeisError(load().seis_h5_last_error().decode())


def _b(s):
    return os.fsencode(s)


def create(path):
    """New, empty file with the MATLAB 7.3 user block."""
    _check(load().seis_h5_file_create(_b(path)))


def write(path, name, value):
    """One dataset, array axes as given (C order)."""
    a = np.asarray(value)
    if a.dtype.kind in "iub" and a.dtype not in (np.dtype(np.int32),):
        a = a.astype(np.int64)
    elif a.dtype.kind == "f" and a.dtype != np.dtype(np.float32):
        a = a.astype(np.float64)
    a = np.require(a, requirements="C")   # keeps 0-d arrays 0-d (scalars)
    if a.ndim > MAXDIM:
        raise SeisError(f"{name}: rank {a.ndim} exceeds {MAXDIM}")
    dims = (ctypes.c_longlong * max(a.ndim, 1))(*a.shape)
    _check(load().seis_h5_file_write(_b(path), _b(name), _CODE[a.dtype],
                                     a.ctypes.data_as(ctypes.c_void_p), a.ndim, dims))


def names(path):
    buf = ctypes.create_string_buffer(1 << 16)
    _check(load().seis_h5_file_list(_b(path), buf, len(buf)))
    return [n for n in buf.value.decode().split("\n") if n]


def shape(path, name):
    nd = ctypes.c_int()
    dims = (ctypes.c_longlong * MAXDIM)()
    _check(load().seis_h5_file_dims(_b(path), _b(name), ctypes.byref(nd), dims))
    return tuple(int(dims[i]) for i in range(nd.value))


def read(path, name, dtype=np.float32):
    """One dataset as stored (C order), converted to dtype."""
    dt = np.dtype(dtype)
    a = np.empty(shape(path, name), dt)
    _check(load().seis_h5_file_read(_b(path), _b(name), _CODE[dt],
                                    a.ctypes.data_as(ctypes.c_void_p), a.size))
    return a


def savemat(path, data):
    """Write {name: value} as SeisCL.py does: a new file, arrays transposed."""
    create(path)
    for k, v in data.items():
        write(path, k, np.transpose(np.asarray(v)))


def loadmat(path, keys=None, dtype=np.float32):
    """Read datasets back in SeisCL.py's orientation (transposed)."""
    keys = names(path) if keys is None else keys
    return {k: np.transpose(read(path, k, dtype)) for k in keys}


def cli_main(args):
    """Run the SeisCL_MPI program's body in this process (no exec); returns
    its exit status."""
    argv = [b"SeisCL_MPI"] + [_b(a) for a in args]
    arr = (ctypes.c_char_p * (len(argv) + 1))(*argv, None)
    return load().seis_cli_main(len(argv), arr)
