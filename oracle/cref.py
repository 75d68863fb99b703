"""ctypes access to the C restatement in ``patternbuild_ref.c`` (test infrastructure only).

``build()`` compiles it with ``make`` into ``oracle/_build``; the loaders here only load
what is built (``__graft_entry__.build`` builds it with the rest).
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libhm_oracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        L.hm_ref_build.restype = ctypes.c_int64
        L.hm_ref_build.argtypes = [P, P, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int64, ctypes.c_int, P, P, P]
        L.hm_ref_bincount.restype = None
        L.hm_ref_bincount.argtypes = [P, P, ctypes.c_int64, P, ctypes.c_int64]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def pattern_tables(tdofs, cdofs, nrow, ncol, diag_first):
    """``(indptr int32, indices int32, slot, nnz)`` from the sort-free restatement."""
    tdofs = np.ascontiguousarray(tdofs, dtype=np.int64)
    cdofs = np.ascontiguousarray(cdofs, dtype=np.int64)
    ne, ndt = tdofs.shape
    ndr = cdofs.shape[1]
    indptr = np.zeros(nrow + 1, np.int64)
    slot = np.empty(ne * ndt * ndr, np.int64)
    nnz_max = ne * ndt * ndr
    indices = np.empty(max(nnz_max, 1), np.int32)
    nnz = lib().hm_ref_build(_p(tdofs), _p(cdofs), ne, ndt, ndr, nrow, int(bool(diag_first)),
                             _p(indptr), _p(indices), _p(slot))
    if nnz < 0:
        raise ValueError("a row holds more entries than the packed per-row key allows")
    idx_t = np.int32 if nnz <= np.iinfo(np.int32).max else np.intp
    return indptr.astype(np.int32), indices[:nnz].copy(), slot.astype(idx_t), int(nnz)


def bincount(slot, w, nnz):
    slot = np.ascontiguousarray(slot, dtype=np.int64)
    w = np.ascontiguousarray(w, dtype=np.float64).reshape(-1)
    out = np.empty(nnz)
    lib().hm_ref_bincount(_p(slot), _p(w), slot.size, _p(out), nnz)
    return out
