"""Every block at its full size, through ``PDEVariationalProblem`` on the fused route (the kernel
of these box meshes): 64^3 linear (C, W_um, W_mm against the oracle on an element sample), 96^3 nonlinear
and 128^3 perturbed (size-independent properties).

Properties that fix a mistake the oracle comparison at small sizes could miss at these sizes:
* the closed-form nnz of a box stencil: Q2 rows x Q2 columns (8n+1)^3, Q2 x Q1 (5n+1)^3,
  Q1 x Q1 (3n+1)^3 (a lattice point couples to the trial points of the elements around it);
* the eliminated rows: zero in C and W_um (``csrassemble.py:164-226``, rectangular blocks),
  the identity row in A (``elimination.py:53-72``);
* the two routes (fused kernel; element matrices then the scatter) give the same bits;
* the Hessian blocks W_uu, W_mm (and A of a linear density) are symmetric:
  ``y . (B x) == x . (B y)`` to round-off;
* ``C`` at (u, p) = (a, b) is ``W_um`` at (u, p) = (b, a) for the linear density (both are
  ``int exp(m) m_hat grad a . grad v``);
* the FP32 form of every block is within 1e-6 of the FP64 one (largest entry).
"""

import numpy as np
import pytest
import torch

from oracle import fe_oracle as O

from hippymfem_amd import ADJOINT, PARAMETER, STATE
from hippymfem_amd.problem import PDEVariationalProblem
from hippymfem_amd.workloads import Workload

pytestmark = pytest.mark.gpu

TOL = 1e-13
NAMES = {(ADJOINT, STATE): "A", (ADJOINT, PARAMETER): "C", (STATE, PARAMETER): "W_um",
         (PARAMETER, PARAMETER): "W_mm", (STATE, STATE): "W_uu"}


def closed_nnz(n, i, j):
    w = {STATE: 8, ADJOINT: 8, PARAMETER: 3}
    if (i == PARAMETER) != (j == PARAMETER):
        return (5 * n + 1) ** 3
    return (w[i] * n + 1) ** 3


def rel(a, b):
    a = torch.as_tensor(a, dtype=torch.float64)
    b = torch.as_tensor(b, dtype=torch.float64)
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))


def problem(W, route="auto"):
    return PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc, route=route)


def point(W):
    return [W.u, W.m, W.p]


def check_eliminated_rows(M, ij, ess, gpu):
    """Rows of the essential dofs: zero (C, W_um) or the identity row (A, W_uu: zero, with
    the policy's diagonal first)."""
    if ij == (PARAMETER, PARAMETER) or len(ess) == 0:
        return
    ip = torch.as_tensor(np.asarray(M.pattern.indptr), dtype=torch.int64, device=gpu)
    r = torch.as_tensor(np.asarray(ess), dtype=torch.int64, device=gpu)
    lo, hi = ip[r], ip[r + 1]
    idx = torch.repeat_interleave(lo, hi - lo)
    idx = idx + torch.arange(idx.numel(), device=gpu) - torch.repeat_interleave(
        torch.cumsum(hi - lo, 0) - (hi - lo), hi - lo)
    vals = M.data[idx].double()
    first = torch.zeros_like(vals, dtype=torch.bool)
    first[torch.cumsum(hi - lo, 0) - (hi - lo)] = True
    if ij in ((ADJOINT, STATE), (STATE, STATE)):
        diag = 1.0 if ij == (ADJOINT, STATE) else 0.0
        assert bool((vals[first] == diag).all()), NAMES[ij]
        assert bool((vals[~first] == 0.0).all()), NAMES[ij]
    else:
        assert bool((vals == 0.0).all()), NAMES[ij]


def symmetric(M, n, seed):
    rng = np.random.default_rng(seed)
    x = torch.as_tensor(rng.standard_normal(n), device=M.data.device)
    y = torch.as_tensor(rng.standard_normal(n), device=M.data.device)
    a = float(torch.dot(y, M.mult(x).double()))
    b = float(torch.dot(x, M.mult(y).double()))
    return abs(a - b) <= 1e-11 * max(abs(a), abs(b))


def test_64cube_hessian_blocks(gpu):
    """C, W_um, W_mm at 64^3 P2/P1: closed-form nnz, eliminated rows, the two routes' bits,
    and 512 elements against the oracle."""
    n = 64
    W = Workload(n)
    fu = problem(W, "fused")
    tp = problem(W, "two-pass")
    P = O.Problem(n)
    elems = np.sort(np.random.default_rng(3).choice(W.mesh.ne, 512, replace=False))
    for ij in [(ADJOINT, PARAMETER), (STATE, PARAMETER), (PARAMETER, PARAMETER)]:
        print("64^3 block %s" % NAMES[ij], flush=True)
        M = fu.block(ij[0], ij[1], point(W))
        assert fu.block_assembly(*ij).fused_plan() is not None
        assert M.nnz == closed_nnz(n, *ij), (NAMES[ij], M.nnz)
        assert int(M.pattern.indptr[-1]) == M.nnz
        check_eliminated_rows(M, ij, W.ess, gpu)
        T = tp.block(ij[0], ij[1], point(W))
        assert torch.equal(M.data, T.data), NAMES[ij]
        E = tp.block_assembly(*ij).E
        Eg = E[torch.as_tensor(elems, device=gpu)].cpu().numpy()
        assert rel(Eg, P.block(ij[0], ij[1], elems)) <= TOL, NAMES[ij]
        del M, T, E
        torch.cuda.empty_cache()


def _properties(W, gpu, blocks_expected):
    n = W.n
    prob = problem(W, "fused")
    got = prob.set_linearization_point(point(W))
    assert sorted(got) == sorted(blocks_expected)
    got32 = prob.set_linearization_point(point(W), precision="fp32")
    for ij, M in got.items():
        print("%d^3 block %s: nnz %d" % (n, NAMES[ij], M.nnz), flush=True)
        assert prob.block_assembly(*ij).fused_plan() is not None
        assert M.nnz == closed_nnz(n, *ij), (NAMES[ij], M.nnz)
        check_eliminated_rows(M, ij, W.ess, gpu)
        e32 = rel(got32[ij].data, M.data)
        assert e32 <= 1e-6, (NAMES[ij], e32)
        assert bool(torch.isfinite(M.data).all())
    return prob, got


def test_96cube_nonlinear_properties(gpu):
    W = Workload(96, nonlinear=True)
    prob, got = _properties(W, gpu, [(ADJOINT, STATE), (ADJOINT, PARAMETER), (STATE, STATE),
                                     (STATE, PARAMETER), (PARAMETER, PARAMETER)])
    assert symmetric(got[(STATE, STATE)], W.Vu.ndof, 1)
    assert symmetric(got[(PARAMETER, PARAMETER)], W.Vm.ndof, 2)


def test_128cube_perturbed_properties(gpu):
    W = Workload(128, perturb=0.2)
    prob, got = _properties(W, gpu, [(ADJOINT, STATE), (ADJOINT, PARAMETER),
                                     (STATE, PARAMETER), (PARAMETER, PARAMETER)])
    assert symmetric(got[(ADJOINT, STATE)], W.Vu.ndof, 1)
    assert symmetric(got[(PARAMETER, PARAMETER)], W.Vm.ndof, 2)
    # C at (u, p) = (a, b) is W_um at (u, p) = (b, a)
    C = got[(ADJOINT, PARAMETER)]
    del got
    torch.cuda.empty_cache()
    Wum = prob.block(STATE, PARAMETER, [W.p, W.m, W.u])
    assert np.array_equal(np.asarray(C.pattern.indptr), np.asarray(Wum.pattern.indptr))
    assert rel(Wum.data, C.data) <= TOL
