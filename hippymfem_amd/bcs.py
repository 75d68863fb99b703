"""Dirichlet conditions (``hippymfem/fem/bcs.py``): the essential dofs of a space and the
zeroing of essential entries."""

import numpy as np
import torch


class DirichletBC:
    """Essential dofs of ``space`` where ``where`` holds (``DirichletBC``, ``bcs.py:105``).

    ``where``: ``"x0x1"`` (the faces x = 0 and x = 1, the benchmarks' condition), ``"all"``
    (the whole boundary), or a callable ``f(X) -> bool mask`` over dof coordinates
    ``(ndof, 3)``.  ``ess_tdof`` is the sorted int64 array of those dofs (one rank: true dofs
    are the local dofs)."""

    def __init__(self, space, value=None, where="x0x1", tol=1e-12):
        self.space = space
        X = space.dof_coordinates()
        if callable(where):
            mask = np.asarray(where(X), dtype=bool)
        elif where == "x0x1":
            mask = (np.abs(X[:, 0]) < tol) | (np.abs(X[:, 0] - 1.0) < tol)
        elif where == "all":
            mask = np.any((np.abs(X) < tol) | (np.abs(X - 1.0) < tol), axis=1)
        else:
            raise ValueError("unknown boundary selection %r" % (where,))
        self.ess_tdof = np.flatnonzero(mask).astype(np.int64)
        self.ess = self.ess_tdof
        self.mask = mask
        self._dev = {}

    def row_flags(self, device):
        """int8 flag per dof (1 on essential dofs) on ``device``, made once."""
        key = str(device)
        if key not in self._dev:
            self._dev[key] = torch.as_tensor(self.mask.astype(np.int8)).to(device)
        return self._dev[key]


def zero_essential(v, ess):
    """``v[ess] = 0`` where ``v`` is (``zero_essential``, ``bcs.py:80-102``)."""
    if ess is None or not len(ess):
        return v
    if isinstance(v, torch.Tensor):
        v.index_fill_(0, torch.as_tensor(np.asarray(ess, dtype=np.int64)).to(v.device), 0.0)
    else:
        v[np.asarray(ess, dtype=np.int64)] = 0.0
    return v


def ess_mask(n, ess):
    """Boolean marker of the essential dofs (``_ess_ldof_mask``, ``prolongation.py:184-209``)."""
    m = np.zeros(int(n), dtype=bool)
    if ess is not None and len(ess):
        m[np.asarray(ess, dtype=np.int64)] = True
    return m
