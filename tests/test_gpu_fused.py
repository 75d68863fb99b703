"""The fused route (element kernel and scatter in one kernel, ``hm_assemble_fused``) against
the two-pass route and the oracle.

The fused values are the two-pass values bit for bit (every entry has the element kernel's
bits and every CSR value sums its entries in ascending element order, as ``np.bincount``
does), for every block, numbering, mesh, rule and tile edge; and within 1e-13 of the oracle.
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
BLOCKS = [(ADJOINT, STATE), (ADJOINT, PARAMETER), (STATE, PARAMETER), (PARAMETER, PARAMETER)]


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def both_routes(W, ij, precision="fp64", nonlinear=False):
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc,
                                 is_fwd_linear=not nonlinear)
    loc = [torch.as_tensor(v).cuda() for v in (W.u, W.m, W.p)]
    f = prob.block_assembly(ij[0], ij[1], precision, route="fused")
    t = prob.block_assembly(ij[0], ij[1], precision, route="two-pass")
    A = f.assemble(loc)
    B = t.assemble(loc)
    assert f.fused_plan() is not None and t.fused_plan() is None
    return A, B, f, loc


CASES = [
    # n, numbering, nonlinear, perturb, quadrature degree
    (4, "lattice", False, 0.0, 4),
    (5, "mfem", True, 0.2, 4),
    (7, "shuffled", False, 0.2, 4),
    (6, "lattice", True, 0.2, 6),
    (9, "lattice", False, 0.0, 4),
]


@pytest.mark.parametrize("n,numbering,nonlinear,perturb,degree", CASES)
def test_fused_equals_two_pass_and_oracle(gpu, n, numbering, nonlinear, perturb, degree):
    W = Workload(n, numbering=numbering, nonlinear=nonlinear, perturb=perturb,
                 quadrature_degree=degree)
    P = O.Problem(n, perturb=perturb, numbering=numbering,
                  density="nonlinear" if nonlinear else "linear", n1d=degree // 2 + 1)
    for ij in BLOCKS + ([(STATE, STATE)] if nonlinear else []):
        print("case n=%d %s nl=%s perturb=%g deg=%d block %s" % (n, numbering, nonlinear,
                                                                 perturb, degree, ij), flush=True)
        A, B, f, _ = both_routes(W, ij, nonlinear=nonlinear)
        assert torch.equal(A.data, B.data), (ij, (A.data - B.data).abs().max().item())
        ip, ix, d = A.host_csr()
        ipo, ixo, do = P.assembled(*ij)
        assert np.array_equal(ip, ipo) and np.array_equal(ix, ixo), ij
        assert rel(d, do) <= TOL, (ij, rel(d, do))


@pytest.mark.parametrize("nonlinear", [False, True])
def test_fused_fp32_equals_two_pass(gpu, nonlinear):
    W = Workload(6, nonlinear=nonlinear, perturb=0.2)
    for ij in BLOCKS:
        A, B, _, _ = both_routes(W, ij, precision="fp32", nonlinear=nonlinear)
        assert A.data.dtype == torch.float32
        assert torch.equal(A.data, B.data), ij


def test_fused_repeat_bit_identical_and_linearization_point(gpu):
    """Repeated fused assemblies give the same bits; a linearization point on the fused route
    equals the one on the two-pass route."""
    W = Workload(8, perturb=0.2)
    A, _, f, loc = both_routes(W, (ADJOINT, STATE))
    ref = A.data.clone()
    for _ in range(3):
        assert torch.equal(f.assemble(loc).data, ref)
    x = [W.u, W.m, W.p]
    fu = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc, route="fused")
    tp = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc, route="two-pass")
    a = fu.set_linearization_point(x)
    b = tp.set_linearization_point(x)
    assert sorted(a) == sorted(b)
    for ij in a:
        assert torch.equal(a[ij].data, b[ij].data), ij


def test_fused_nonfinite_raises(gpu):
    """A parameter whose exponential overflows makes non-finite element values: the fused
    route flags them and ``block`` raises RuntimeError, as on the two-pass route."""
    W = Workload(4)
    m = W.m.copy()
    m[W.Vm.ndof // 2:] = 800.0                             # exp(800) overflows FP64
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc, route="fused")
    with pytest.raises(RuntimeError):
        prob.block(ADJOINT, STATE, [W.u, m, W.p])
    prob.block(ADJOINT, STATE, [W.u, W.m, W.p])            # finite again


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_aligned_form_equals_general_form(gpu, precision):
    """On a mesh of axis-aligned elements the plan carries HM_FUSED_ALIGNED and the kernel takes
    its aligned form (no per-element alignment test, slots for the diagonal reference-gradient
    pairs only); its values are the general form's bits, every block, nonlinear density."""
    W = Workload(5, nonlinear=True)
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc, is_fwd_linear=False)
    loc = [torch.as_tensor(v).cuda() for v in (W.u, W.m, W.p)]
    for ij in BLOCKS + [(STATE, STATE)]:
        f = prob.block_assembly(ij[0], ij[1], precision, route="fused")
        plan = f.fused_plan()
        assert plan is not None and plan.struct.flags == 1, ij
        A = f.assemble(loc).data.clone()
        plan.struct.flags = 0
        try:
            B = f.assemble(loc).data.clone()
        finally:
            plan.struct.flags = 1
        assert torch.equal(A, B), (ij, precision)


def test_perturbed_mesh_takes_general_form(gpu):
    W = Workload(4, perturb=0.2)
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc)
    f = prob.block_assembly(ADJOINT, STATE, route="fused")
    assert f.fused_plan().struct.flags == 0
