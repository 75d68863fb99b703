"""Caller-supplied densities (``PointwiseDensity``: torch differentiates the density at the
points, ``hm_element_matrices_pointwise`` / ``hm_element_vectors_pointwise`` contract).

Pinned two ways: the closed-form density restated as a caller density gives the closed-form
kernels' element arrays and assembled blocks (and so the oracle's), and a density the library
does not know has element vectors that are the derivatives of its functional and element
matrices that are the derivatives of its vectors (central differences).
"""

import numpy as np
import pytest
import torch

from oracle import fe_oracle as O

from hippymfem_amd import ADJOINT, PARAMETER, STATE
from hippymfem_amd.density import PointwiseDensity
from hippymfem_amd.kernel import QuadratureKernel
from hippymfem_amd.problem import PDEVariationalProblem
from hippymfem_amd.workloads import Workload

pytestmark = pytest.mark.gpu

TOL = 1e-13


def rel(a, b):
    a = torch.as_tensor(a, dtype=torch.float64).cpu()
    b = torch.as_tensor(b, dtype=torch.float64).cpu()
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))


def exp_diffusion(nonlinear, f=1.0):
    def fn(u, gu, m, gm, p, gp, x):
        g = 1.0 + u * u if nonlinear else torch.ones_like(u)
        return torch.exp(m) * g * (gu * gp).sum(-1) - f * p
    return PointwiseDensity(fn, is_fwd_linear=not nonlinear,
                            name="exp diffusion%s" % (" (nonlinear)" if nonlinear else ""))


@pytest.mark.parametrize("nonlinear", [False, True])
def test_caller_density_equals_closed_form(gpu, nonlinear):
    """Every block, the vectors, the products and the values of the closed-form density,
    restated as a caller density, on a perturbed mesh: within 1e-13 of the closed-form
    kernels, and the blocks within 1e-13 of the oracle."""
    W = Workload(4, nonlinear=nonlinear, perturb=0.2)
    P = O.Problem(4, perturb=0.2, density="nonlinear" if nonlinear else "linear")
    K = QuadratureKernel(exp_diffusion(nonlinear), W.spaces, W.batches)
    loc = W.locals()
    for i in range(3):
        for j in range(3):
            got = K.element_matrices(i, j, loc)[0]
            ref = W.kernel.element_matrices(i, j, loc)[0]
            if float(ref.abs().max()) == 0.0:
                assert float(got.abs().max()) <= 1e-12, (i, j)
                continue
            assert rel(got, ref) <= TOL, (i, j, rel(got, ref))
            assert rel(got, P.block(i, j)) <= TOL, (i, j)
        v = K.element_vectors(i, loc)[0]
        assert rel(v, W.kernel.element_vectors(i, loc)[0]) <= TOL, i
    d = W.direction()
    assert rel(K.element_hvp(ADJOINT, STATE, loc, d)[0],
               W.kernel.element_hvp(ADJOINT, STATE, loc, d)[0]) <= TOL
    assert rel(K.element_values(loc)[0], P.values()) <= TOL
    # FP32 element matrices: the FP64 sums rounded once
    got32 = K.element_matrices(ADJOINT, STATE, loc, precision="fp32")[0]
    assert got32.dtype == torch.float32
    assert rel(got32, W.kernel.element_matrices(ADJOINT, STATE, loc)[0]) <= 1e-6


def test_caller_density_through_the_problem(gpu):
    """``PDEVariationalProblem`` with a caller density: the linearization point's blocks and
    the residual are the closed-form problem's (two-pass route; the fused route is for the
    closed-form densities)."""
    W = Workload(5, nonlinear=True, perturb=0.2)
    x = [W.u, W.m, W.p]
    mine = PDEVariationalProblem(W.spaces, exp_diffusion(True), W.batches, W.bc)
    ref = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc)
    a = mine.set_linearization_point(x)
    b = ref.set_linearization_point(x)
    assert sorted(a) == sorted(b)
    for ij in a:
        assert mine.block_assembly(*ij).fused_plan() is None
        assert np.array_equal(np.asarray(a[ij].pattern.indptr), np.asarray(b[ij].pattern.indptr))
        assert rel(a[ij].data, b[ij].data) <= TOL, ij
    assert rel(mine.residual(x), ref.residual(x)) <= TOL
    assert abs(mine.functional(x) - ref.functional(x)) <= TOL * abs(ref.functional(x))


def _new_density():
    """A density the library has no closed form for: every slot pair coupled, values and
    gradients, the parameter's gradient included."""
    def fn(u, gu, m, gm, p, gp, x):
        return (0.5 * torch.exp(0.3 * m) * (1.0 + u * u) * (gu * gu).sum(-1)
                + torch.sin(u) * p + 0.1 * (gm * gu).sum(-1) * p + m * m * p
                + 0.2 * (gu * gp).sum(-1) * (1.0 + x[..., 0]))
    return PointwiseDensity(fn, name="test density")


def test_new_density_derivatives_are_consistent(gpu):
    W = Workload(3, perturb=0.2)
    K = QuadratureKernel(_new_density(), W.spaces, W.batches)
    rng = np.random.default_rng(4)
    x = [0.5 * rng.standard_normal(W.Vu.ndof), 0.3 * rng.standard_normal(W.Vm.ndof),
         rng.standard_normal(W.Vu.ndof)]
    spaces = W.spaces
    eps = 1e-6

    def total(xx):
        return float(K.element_values(xx)[0].sum())

    def vec(i, xx):
        v = K.element_vectors(i, xx)[0].cpu().numpy()
        return np.bincount(spaces[i].edofs.reshape(-1), weights=v.reshape(-1),
                           minlength=spaces[i].ndof)

    for i in range(3):
        d = rng.standard_normal(spaces[i].ndof)
        xp = [a.copy() for a in x]
        xm = [a.copy() for a in x]
        xp[i] = x[i] + eps * d
        xm[i] = x[i] - eps * d
        fd = (total(xp) - total(xm)) / (2 * eps)
        g = vec(i, x) @ d
        assert abs(fd - g) <= 1e-7 * max(1.0, abs(g)), (i, fd, g)
        for j in range(3):
            dj = rng.standard_normal(spaces[j].ndof)
            xp = [a.copy() for a in x]
            xm = [a.copy() for a in x]
            xp[j] = x[j] + eps * dj
            xm[j] = x[j] - eps * dj
            fdv = (K.element_vectors(i, xp)[0] - K.element_vectors(i, xm)[0]).cpu().numpy() / (2 * eps)
            Ed = K.element_hvp(i, j, x, dj)[0].cpu().numpy()
            assert np.abs(fdv - Ed).max() <= 1e-6 * max(1.0, np.abs(Ed).max()), (i, j)
