"""The fused route (element kernel and scatter in one kernel, ``hm_assemble_fused``) against
the two-pass route and the oracle.

The FP64 fused values are the two-pass values bit for bit (every entry has the element
kernel's bits and every CSR value sums its entries in ascending element order, as
``np.bincount`` does), for every block, numbering, mesh, rule and tile edge; and within 1e-13
of the oracle.  The FP32 form sums the same FP32 entries in FP32 in the same order.
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


def fp32_sums_in_element_order(t):
    """The FP32 form's CSR values restated on the host: each value the FP32 sum, from 0, of its
    entries in ascending element order, the entries those of the element kernel (the two-pass
    block ``t``'s FP32 element matrices); the slots of the elimination as ``t`` writes them."""
    pat = t.pattern
    E = t.E.cpu().numpy()
    ne, nt, nc = E.shape
    rows = np.repeat(pat.tdofs[:, :, None], nc, axis=2).reshape(-1).astype(np.int64)
    cols = np.repeat(pat.cdofs[:, None, :], nt, axis=1).reshape(-1).astype(np.int64)
    indptr = pat.indptr.astype(np.int64)
    row_of = np.repeat(np.arange(pat.nrow, dtype=np.int64), np.diff(indptr))
    key = row_of * pat.ncol + pat.indices.astype(np.int64)
    order = np.argsort(key)
    want = rows * pat.ncol + cols
    pos = order[np.searchsorted(key[order], want)]
    assert np.array_equal(key[pos], want)
    srt = np.argsort(pos, kind="stable")            # entries are element-major: stable keeps e order
    p, v = pos[srt], E.reshape(-1)[srt]
    start = np.r_[0, np.flatnonzero(np.diff(p)) + 1]
    rank = np.arange(p.size) - np.repeat(start, np.diff(np.r_[start, p.size]))
    out = np.zeros(pat.nnz, dtype=np.float32)
    for r in range(int(rank.max()) + 1):
        sel = rank == r
        out[p[sel]] = out[p[sel]] + v[sel]
    kill = t.plan.kill
    if kill is not None and len(kill):
        out[kill] = t.data.cpu().numpy()[kill]
    return out


@pytest.mark.parametrize("n,nonlinear,perturb", [(6, False, 0.2), (6, True, 0.2), (9, False, 0.0),
                                                 (5, True, 0.0)])
def test_fused_fp32_sums_in_fp32_in_element_order(gpu, n, nonlinear, perturb):
    """The FP32 form sums its FP32 entries in FP32, in ascending element order: its values are
    the host's FP32 sums of the element kernel's entries in that order, bit for bit, on
    perturbed meshes and box meshes (two blocks a compute unit, the 9^3 case in chunks of layers;
    the perturbed nonlinear blocks of many pair slots one); and within 1e-6 of the FP64 values."""
    W = Workload(n, nonlinear=nonlinear, perturb=perturb)
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc, is_fwd_linear=not nonlinear)
    loc = [torch.as_tensor(v).cuda() for v in (W.u, W.m, W.p)]
    for ij in BLOCKS + ([(STATE, STATE)] if nonlinear else []):
        print("case n=%d nl=%s perturb=%g block %s" % (n, nonlinear, perturb, ij), flush=True)
        f = prob.block_assembly(ij[0], ij[1], "fp32", route="fused")
        t = prob.block_assembly(ij[0], ij[1], "fp32", route="two-pass")
        r = prob.block_assembly(ij[0], ij[1], "fp64", route="two-pass")
        A = f.assemble(loc).data.clone()
        t.assemble(loc)
        assert f.fused_plan() is not None and t.fused_plan() is None
        assert A.dtype == torch.float32
        want = fp32_sums_in_element_order(t)
        got = A.cpu().numpy()
        assert np.array_equal(got, want), (ij, int((got != want).sum()))
        ref = r.assemble(loc).data
        assert rel(got, ref.cpu().numpy()) <= 1e-6, ij
        assert torch.equal(f.assemble(loc).data, A), ij


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


def test_fp32_route_choice_and_blocks_a_compute_unit(gpu):
    """``hm_fused_blocks``: the FP32 form of A runs two blocks a compute unit on box and perturbed
    meshes (its tile's LDS fits half the unit's), FP64 one; the nonlinear blocks of the perturbed
    mesh hold more pair slots and some run one. The auto route takes the fused kernel for FP32
    exactly where it runs two."""
    from hippymfem_amd import fused

    seen = set()
    for n, nonlinear, perturb in ((8, False, 0.0), (8, False, 0.2), (6, True, 0.2)):
        W = Workload(n, nonlinear=nonlinear, perturb=perturb)
        prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc, is_fwd_linear=not nonlinear)
        k = prob.kernel
        aligned = perturb == 0.0
        assert fused.aligned_box(W.mesh) is aligned
        for ij in BLOCKS + ([(STATE, STATE)] if nonlinear else []):
            if ij[1] == PARAMETER:
                continue                       # Q1 trial spaces: the two-pass route
            assert fused.blocks_of(k, ij[0], ij[1], "fp64", aligned) == 1
            nb = fused.blocks_of(k, ij[0], ij[1], "fp32", aligned)
            seen.add(nb)
            if not nonlinear and ij == (ADJOINT, STATE):
                assert nb == 2, (n, perturb)
            auto = prob.block_assembly(ij[0], ij[1], "fp32")
            assert (auto.fused_plan() is not None) == (nb == 2), (ij, auto.why_two_pass)
    assert seen == {1, 2}


def test_fp64_auto_route_by_density_and_mesh(gpu):
    """The auto route of the FP64 Q2 blocks: fused on box meshes (either density) and for the
    linear density on perturbed meshes; two-pass for the nonlinear density on perturbed meshes,
    where the fused kernel's tiles are smaller."""
    for nonlinear, perturb, want in ((False, 0.0, True), (False, 0.2, True), (True, 0.0, True),
                                     (True, 0.2, False)):
        W = Workload(6, nonlinear=nonlinear, perturb=perturb)
        prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc, is_fwd_linear=not nonlinear)
        for ij in [(ADJOINT, STATE)] + ([(STATE, STATE)] if nonlinear else []):
            b = prob.block_assembly(ij[0], ij[1], "fp64")
            assert (b.fused_plan() is not None) is want, (nonlinear, perturb, ij, b.why_two_pass)
            if not want:
                assert "nonlinear density" in b.why_two_pass
