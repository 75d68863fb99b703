"""Orchestration of a block's assembly (``hippymfem/fem/csrassemble.py``): the
value-independent plan, the scatter (whole or in chunks, several blocks from one pass) and
the finished matrix, with the essential-dof elimination folded into the scatter.
"""

import numpy as np
import torch

from . import _lib
from .bcs import ess_mask
from .parmat import local_par_matrix
from .pattern import get_pattern, get_vector_pattern

_MASKS = {}


def _mask(space, ess):
    """The boolean marker of ``ess`` on ``space``, cached on the identity of the two, as
    the reference caches its ldof masks (``prolongation.py:184-209``)."""
    if ess is None:
        return None
    key = (id(space), id(ess))
    got = _MASKS.get(key)
    if got is not None and got[0] is space and got[1] is ess:
        return got[2]
    m = ess_mask(space.ndof, ess)
    _MASKS[key] = (space, ess, m)
    return m


class BlockPlan:
    """What :func:`plan_block` decides (``BlockPlan``, ``csrassemble.py:141-161``)."""

    __slots__ = ("test_space", "trial_space", "pattern", "same", "sym", "test_ess",
                 "trial_ess", "diag_policy", "fold", "row_mask", "col_mask", "diag", "device",
                 "_kill")

    @property
    def kill(self):
        """Host slots the elimination zeroes (``masked_slots``): for checks and callers
        that scatter without the folded plan."""
        if self._kill is None and self.fold:
            self._kill = self.pattern.masked_slots(self.row_mask, self.col_mask)
        return self._kill

    @property
    def zero(self):
        return self.kill


def plan_block(test_space, trial_space, groups, test_ess=None, trial_ess=None,
               diag_policy="one", symmetric=None):
    """The value-independent part of assembling a block (``plan_block``,
    ``csrassemble.py:164-226``).  ``symmetric``: whether a block whose two sides are one
    space is eliminated on both sides (the default); ``False`` for a parameter's block,
    whose columns stay."""
    if diag_policy not in ("one", "zero", "keep"):
        raise ValueError("diag_policy must be 'one', 'zero' or 'keep', not %r" % (diag_policy,))
    p = BlockPlan()
    p.test_space, p.trial_space = test_space, trial_space
    p.test_ess, p.trial_ess, p.diag_policy = test_ess, trial_ess, diag_policy
    p.pattern = get_pattern(test_space, trial_space, groups)
    p.same = test_space is trial_space
    p.sym = p.same and (symmetric is None or bool(symmetric))
    p.fold = test_ess is not None or trial_ess is not None
    p._kill = None
    p.row_mask = _mask(test_space, test_ess)
    if p.sym and test_ess is not None:
        p.col_mask = p.row_mask
    else:
        p.col_mask = _mask(trial_space, trial_ess)
    p.diag = None
    if p.fold and p.sym and test_ess is not None:
        p.diag = 0.0 if diag_policy == "zero" else 1.0
    p.device = None
    return p


def _device_plan(p):
    if p.device is None:
        if p.diag_policy == "keep" and p.diag is not None:
            p.device = p.pattern.device_plan()          # the elimination comes after
        else:
            p.device = p.pattern.device_plan(p.row_mask, p.col_mask, p.diag)
    return p.device


def _keep_diagonal(p, acc):
    """``diag_policy="keep"``: eliminate, then put the assembled diagonal back
    (``elimination._eliminate``, ``elimination.py:75-117``)."""
    rows = torch.as_tensor(np.flatnonzero(p.row_mask)).to(acc.device)
    first = torch.as_tensor(p.pattern.indptr[:-1].astype(np.int64)).to(acc.device)[rows]
    keep = acc[first].clone()
    acc.index_fill_(0, torch.as_tensor(p.kill).to(acc.device), 0.0)
    acc[first] = keep
    return acc


def finish_block(p, acc):
    """The matrix of a planned block from its scattered values (``finish_block``,
    ``csrassemble.py:229-270``); one rank: the diag block over ``acc``."""
    if p.diag_policy == "keep" and p.diag is not None:
        acc = _keep_diagonal(p, acc)
    return local_par_matrix(p.pattern, acc, p.test_space, p.trial_space)


def scatter_many(plans, chunks, dtype=torch.float64):
    """One pass over chunks carrying several blocks, into every plan's target
    (``scatter_many``, ``csrassemble.py:286-301``).  ``chunks`` yields
    ``(group, start, stop, {key: array})``; returns ``{key: accumulator}``."""
    state = {}
    for k, p in plans.items():
        acc, maps = p.pattern.fused_begin(dtype)
        state[k] = (acc, maps)
    for g, a, bnd, mats in chunks:
        for k, p in plans.items():
            acc, maps = state[k]
            p.pattern.fused_add(acc, maps, g, a, bnd, mats[k], plan=_device_plan(p))
    return {k: state[k][0] for k in plans}


def assemble_matrix_csr(test_space, trial_space, groups, element_matrices, nelem,
                        test_ess=None, trial_ess=None, diag_policy="one", boundary=None,
                        symmetric=None, dtype=None):
    """Assemble a block from element matrices (``assemble_matrix_csr``,
    ``csrassemble.py:364-399``).  ``element_matrices``: one ``(ne, nd_test, nd_trial)``
    array per group, or a callable returning an iterator of ``(group, start, stop, array)``
    chunks, each scattered as it comes.  ``dtype``: precision of the values (that of the
    element matrices by default)."""
    if boundary is not None:
        raise NotImplementedError("boundary terms are not part of this path")
    p = plan_block(test_space, trial_space, groups, test_ess, trial_ess, diag_policy,
                   symmetric=symmetric)
    plan = _device_plan(p)
    if callable(element_matrices):
        acc = p.pattern.data_fused(element_matrices(), plan=plan,
                                   dtype=torch.float64 if dtype is None else dtype)
    else:
        acc = p.pattern.data(element_matrices, plan=plan, dtype=dtype)
    return finish_block(p, acc)


def assemble_vector_csr(space, groups, element_vectors, nelem, ess=None, out=None):
    """Assemble a dual vector (``assemble_vector_csr``, ``csrassemble.py:402-424``): the
    element vectors summed by dof in entry order, ``P^T`` the identity on one rank, the
    essential entries zeroed."""
    pat = get_vector_pattern(space, groups)
    zero_rows = None
    if ess is not None and len(ess):
        m = _mask(space, ess)
        dev = _lib.require_device()
        key = ("rows", id(space), id(ess))
        got = _MASKS.get(key)
        if got is None or got[0] is not m:
            got = _MASKS[key] = (m, torch.as_tensor(m.astype(np.int8)).to(dev))
        zero_rows = got[1]
    return pat.scatter(element_vectors, zero_rows=zero_rows, out=out)
