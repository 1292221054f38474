"""HDF5 datasets of the reference's output files, through libqss.so.

The reference writes its checkpoints with libhdf5 (myhdfshell.cpp:
WriteHDF5Recorddouble / RecordInt / RecordLong, and the particle compounds of
save.cpp:225-251).  libqss.so carries its own writer and reader of that
format subset (no libhdf5): these helpers move whole datasets between such
files and numpy arrays.  Host-only; no GPU is touched.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List

import numpy as np

from . import _lib
from ._lib import QssH5Set, check

F64, I32, I64, PARTICLE, ION = 0, 1, 2, 3, 4

#: numpy records equal to qss_saved_particle / qss_saved_ion (particles.h:111-153)
SAVED_PARTICLE = np.dtype({"names": ["cell", "x", "y", "z", "px", "py", "pz", "weight", "q2m"],
                           "formats": ["<i4"] + ["<f8"] * 8,
                           "offsets": [0, 8, 16, 24, 32, 40, 48, 56, 64], "itemsize": 72})
SAVED_ION = np.dtype({"names": ["cell", "x", "y", "z", "px", "py", "pz", "weight", "q2m",
                                "i_z", "sort"],
                      "formats": ["<i4"] + ["<f8"] * 8 + ["<i4", "<i4"],
                      "offsets": [0, 8, 16, 24, 32, 40, 48, 56, 64, 72, 76], "itemsize": 80})
_DTYPE = {F64: np.dtype("<f8"), I32: np.dtype("<i4"), I64: np.dtype("<i8"),
          PARTICLE: SAVED_PARTICLE, ION: SAVED_ION}


def _kind_of(a: np.ndarray) -> int:
    if a.dtype == SAVED_ION:
        return ION
    if a.dtype == SAVED_PARTICLE:
        return PARTICLE
    if a.dtype.kind == "f":
        return F64
    if a.dtype.kind in "iu":
        return I32 if a.dtype.itemsize <= 4 else I64
    raise TypeError("no HDF5 kind for dtype %s" % a.dtype)


def write(path: str, datasets: Dict[str, np.ndarray]) -> None:
    """Create `path` with one dataset per entry (rank 1..3; float -> F64LE,
    int32 -> I32LE, int64 -> I64LE, the record dtypes -> their compounds)."""
    keep, sets = [], (QssH5Set * len(datasets))()
    for i, (name, arr) in enumerate(datasets.items()):
        a = np.asarray(arr)
        kind = _kind_of(a)
        a = np.ascontiguousarray(a, dtype=_DTYPE[kind])
        if not 1 <= a.ndim <= 3:
            raise ValueError("dataset %r: rank must be 1..3" % name)
        keep.append(a)
        sets[i].name = name.encode()
        sets[i].kind = kind
        sets[i].rank = a.ndim
        for r in range(a.ndim):
            sets[i].dims[r] = a.shape[r]
        sets[i].data = a.ctypes.data_as(ctypes.c_void_p)
    check(_lib.load().qss_h5_write(path.encode(), len(datasets), sets), "qss_h5_write")


def names(path: str) -> List[str]:
    """Dataset names in the file's root group."""
    L = _lib.load()
    cap = 1 << 16
    buf = ctypes.create_string_buffer(cap)
    check(L.qss_h5_list(path.encode(), buf, cap), "qss_h5_list")
    return [n for n in buf.value.decode().split("\n") if n]


def info(path: str, name: str):
    """(kind, shape) of a dataset; kind -1 for types other than these five."""
    kind, rank = ctypes.c_int32(), ctypes.c_int32()
    dims = (ctypes.c_int64 * 3)()
    check(_lib.load().qss_h5_info(path.encode(), name.encode(), ctypes.byref(kind),
                                  ctypes.byref(rank), dims), "qss_h5_info")
    return kind.value, tuple(dims[r] for r in range(rank.value))


def read(path: str, name: str, kind: int = None) -> np.ndarray:
    """The whole dataset as a numpy array of its shape, converted to `kind`
    (default: the stored kind)."""
    k, shape = info(path, name)
    kind = k if kind is None else kind
    if kind not in _DTYPE:
        raise TypeError("dataset %r has a type these helpers do not convert" % name)
    out = np.empty(shape, dtype=_DTYPE[kind])
    check(_lib.load().qss_h5_read(path.encode(), name.encode(), kind, 0, out.size,
                                  out.ctypes.data_as(ctypes.c_void_p)), "qss_h5_read")
    return out
