"""The assembled block: hypre ParCSR layout on one rank (``hippymfem/fem/parmat.py``).

On one rank a block is its diag block alone (the offd block is empty): ``indptr``/``indices``
int32 (``HYPRE_Int``), ``data`` fp64 (fp32 in the precision form), rows of a square block
leading with their diagonal, which is the order hypre keeps a square matrix in so it can
alias these arrays (``parmat.py:60-168``).  The arrays stay on the device.
"""

import numpy as np
import torch

from . import _lib


class ParCSRMatrix:
    """One rank's ``HypreParMatrix``: the diag block's CSR on the device."""

    def __init__(self, pattern, data, row_starts=None, col_starts=None):
        self.pattern = pattern
        dev = data.device
        self.indptr = (pattern.dev("indptr") if dev.type == "cuda"
                       else torch.as_tensor(pattern.indptr))
        self._indices = None
        self.data = data
        self.shape = (pattern.nrow, pattern.ncol)
        self.row_starts = np.array([0, pattern.nrow]) if row_starts is None else row_starts
        self.col_starts = np.array([0, pattern.ncol]) if col_starts is None else col_starts
        #: for single-precision values: what their rounding took from each row's sum
        #: (``RowDefect``, ``singlesolve.py:682-700``), or None
        self.row_defect = None

    @property
    def indices(self):
        if self._indices is None:
            self._indices = (self.pattern.dev("indices") if self.data.device.type == "cuda"
                             else torch.as_tensor(self.pattern.indices))
        return self._indices

    @property
    def nnz(self):
        return self.pattern.nnz

    def mult(self, x, y=None, stream=None):
        """``y = A x`` on the device (``HypreParMatrix::Mult``)."""
        dev = self.data.device
        x = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x, np.float64))
        x = x.to(dev, torch.float64).contiguous()
        if y is None:
            y = torch.empty(self.shape[0], dtype=torch.float64, device=dev)
        code = _lib.FP64 if self.data.dtype == torch.float64 else _lib.FP32
        _lib.check(_lib.lib().hm_csr_matvec(self.shape[0], _lib.ptr(self.indptr),
                                            _lib.ptr(self.indices), _lib.ptr(self.data), code,
                                            _lib.ptr(x), _lib.ptr(y), _lib.stream_handle(stream)),
                   "csr matvec")
        return y

    Mult = mult

    def host_csr(self):
        """``(indptr, indices, data)`` as numpy arrays."""
        return (np.asarray(self.pattern.indptr), np.asarray(self.pattern.indices),
                self.data.double().cpu().numpy())

    def to_scipy(self):
        import scipy.sparse as sp

        ip, ix, d = self.host_csr()
        return sp.csr_matrix((d, ix, ip), shape=self.shape)

    def __repr__(self):
        return "ParCSRMatrix(%dx%d, nnz=%d, %s)" % (self.shape[0], self.shape[1], self.nnz,
                                                   self.data.dtype)


def local_par_matrix(pattern, acc, test_space=None, trial_space=None):
    """The one-rank matrix over the scattered values (``local_par_matrix``,
    ``parmat.py:106-168``): it aliases ``acc``."""
    return ParCSRMatrix(pattern, acc)
