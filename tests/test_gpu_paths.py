"""Caller paths of the assembly against the oracle: results that own their values,
``apply_ij_at`` for every block of the linearization point, column elimination by
``trial_ess``, and the Gauss-Newton linearization point.

Bar as in ``test_gpu_parity``: CSR structure bit-exact, values within 1e-13 of the block's
largest entry.
"""

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import fe_oracle as O

from hippymfem_amd import ADJOINT, PARAMETER, STATE
from hippymfem_amd.assemble import assemble_matrix
from hippymfem_amd.problem import PDEVariationalProblem
from hippymfem_amd.workloads import Workload

pytestmark = pytest.mark.gpu

TOL = 1e-13


def pair(n, **kw):
    W = Workload(n, **kw)
    P = O.Problem(n, perturb=kw.get("perturb", 0.0),
                  density="nonlinear" if kw.get("nonlinear") else "linear")
    return W, P


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def csr(ip, ix, d, shape):
    return sp.csr_matrix((d, ix, ip), shape=shape)


@pytest.mark.parametrize("nonlinear", [False, True])
def test_linearization_point_blocks_survive_later_assemblies(gpu, nonlinear):
    """The blocks of ``set_linearization_point(x0)`` own their values: a Jacobian at another
    point and a later ``block`` call leave them as they were (the reference keeps the
    linearization point's A through a line search, ``PDEVariationalProblem.py:562-568``)."""
    W, P = pair(4, perturb=0.2, nonlinear=nonlinear)
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc,
                                 is_fwd_linear=not nonlinear)
    x0 = [W.u.copy(), W.m.copy(), W.p.copy()]
    got = prob.set_linearization_point(x0)
    before = {ij: got[ij].data.clone() for ij in got}
    x1 = [1.5 * W.u, W.m + 0.3, -W.p]
    A1 = prob.jacobian(x1)
    for ij in got:
        prob.block(ij[0], ij[1], x1)
    for ij in got:
        assert torch.equal(got[ij].data, before[ij]), ij
        ip, ix, d = got[ij].host_csr()
        ipo, ixo, do = P.assembled(*ij)
        assert np.array_equal(ip, ipo) and np.array_equal(ix, ixo), ij
        assert rel(d, do) <= TOL, (ij, rel(d, do))
    # and the Jacobian at x1 is the oracle's at x1
    P.u, P.m, P.p = x1
    ip, ix, d = A1.host_csr()
    assert rel(d, P.assembled(ADJOINT, STATE)[2]) <= TOL
    assert not torch.equal(A1.data, got[(ADJOINT, STATE)].data)


#: every block ``apply_ij`` serves (``PDEVariationalProblem.py:1641-1650``): the assembled
#: block and whether it is applied transposed
KKT = {(ADJOINT, STATE): ((ADJOINT, STATE), False), (STATE, ADJOINT): ((ADJOINT, STATE), True),
       (ADJOINT, PARAMETER): ((ADJOINT, PARAMETER), False),
       (PARAMETER, ADJOINT): ((ADJOINT, PARAMETER), True),
       (STATE, STATE): ((STATE, STATE), False),
       (STATE, PARAMETER): ((STATE, PARAMETER), False),
       (PARAMETER, STATE): ((STATE, PARAMETER), True),
       (PARAMETER, PARAMETER): ((PARAMETER, PARAMETER), False)}


@pytest.mark.parametrize("nonlinear", [False, True])
def test_apply_ij_at_equals_assembled_blocks(gpu, nonlinear):
    """``apply_ij_at(i, j, x, d)`` is what the assembled block of the linearization point at
    ``x`` gives ``d`` (``apply_ij_at``, ``PDEVariationalProblem.py:1672-1718``): the
    direction loses its essential state/adjoint entries, the product has them zeroed, and the
    Jacobian's pair puts ``d`` back on the essential rows (the eliminated block's unit
    diagonal).  Checked for every block against the oracle's eliminated CSR times ``d``."""
    W, P = pair(4, perturb=0.2, nonlinear=nonlinear)
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc,
                                 is_fwd_linear=not nonlinear)
    x = [W.u.copy(), W.m.copy(), W.p.copy()]
    rng = np.random.default_rng(21)
    for (i, j), (blk, transpose) in KKT.items():
        Vi, Vj = W.spaces[blk[0]], W.spaces[blk[1]]
        if blk == (STATE, STATE) and not nonlinear:
            M = sp.csr_matrix((Vi.ndof, Vj.ndof))          # W_uu = 0 for the linear density
        else:
            ip, ix, d = P.assembled(*blk)
            M = csr(ip, ix, d, (Vi.ndof, Vj.ndof))
        if transpose:
            M = M.T.tocsr()
        dvec = rng.standard_normal(W.spaces[j].ndof)
        want = M @ dvec
        got = prob.apply_ij_at(i, j, x, dvec).cpu().numpy()
        assert got.shape == want.shape, (i, j)
        scale = max(np.abs(want).max(), 1.0)
        err = np.abs(got - want).max() / scale
        assert err <= TOL, ((i, j), err)
        if {i, j} == {STATE, ADJOINT}:
            ess = W.ess
            assert np.array_equal(got[ess], dvec[ess]), (i, j)      # the identity rows
        elif i in (STATE, ADJOINT):
            assert np.all(got[W.ess] == 0.0), (i, j)
    # a device direction and an output array
    d = torch.as_tensor(rng.standard_normal(W.Vu.ndof)).to(gpu)
    out = torch.empty(W.Vu.ndof, dtype=torch.float64, device=gpu)
    res = prob.apply_ij_at(ADJOINT, STATE, x, d, out=out)
    assert res is out
    ip, ix, dd = P.assembled(ADJOINT, STATE)
    want = csr(ip, ix, dd, (W.Vu.ndof, W.Vu.ndof)) @ d.cpu().numpy()
    assert rel(out.cpu().numpy(), want) <= TOL


def _dense_block(P, i, j, row_ess=None, col_ess=None, diag=None):
    """The block's dense matrix from the oracle's element matrices (no pattern: a plain sum
    over elements), rows ``row_ess`` and columns ``col_ess`` zeroed, ``diag`` on the
    diagonal of the eliminated rows."""
    Vi, Vj = P.space(i), P.space(j)
    E = P.block(i, j)
    D = np.zeros((Vi.ndof, Vj.ndof))
    for e in range(E.shape[0]):
        D[np.ix_(Vi.edofs[e], Vj.edofs[e])] += E[e]
    if row_ess is not None:
        D[np.asarray(row_ess), :] = 0.0
    if col_ess is not None:
        D[:, np.asarray(col_ess)] = 0.0
    if diag is not None:
        r = np.asarray(row_ess)
        D[r, r] = diag
    return D


def test_trial_ess_column_elimination(gpu):
    """``trial_ess`` eliminates the columns of a block without touching its rows or its
    diagonal (``plan_block``, ``csrassemble.py:164-226``): W_um with its state rows and a set
    of parameter columns eliminated, and A with different row and column sets and
    ``symmetric=False``, against the dense restatement; the pattern keeps its explicit
    zeros."""
    W, P = pair(3, perturb=0.2)
    mcols = np.flatnonzero(W.Vm.dof_coordinates()[:, 1] < 1e-12)     # m dofs on y = 0
    E = W.kernel.element_matrices(STATE, PARAMETER, W.locals())
    A = assemble_matrix(W.Vu, W.Vm, W.groups, E, W.mesh.ne, test_ess=W.ess, trial_ess=mcols,
                        symmetric=False)
    ipo, ixo, _ = P.assembled(STATE, PARAMETER)
    ip, ix, d = A.host_csr()
    assert np.array_equal(ip, ipo) and np.array_equal(ix, ixo)
    D = _dense_block(P, STATE, PARAMETER, row_ess=W.ess, col_ess=mcols)
    got = csr(ip, ix, d, D.shape).toarray()
    assert rel(got, D) <= TOL
    assert np.all(got[:, mcols] == 0.0) and np.all(got[W.ess] == 0.0)
    # the square block, rows of x = 0, 1 and columns of z = 0 (no symmetric elimination)
    zcols = np.flatnonzero(W.Vu.dof_coordinates()[:, 2] < 1e-12)
    E = W.kernel.element_matrices(ADJOINT, STATE, W.locals())
    A = assemble_matrix(W.Vu, W.Vu, W.groups, E, W.mesh.ne, test_ess=W.ess, trial_ess=zcols,
                        symmetric=False)
    ip, ix, d = A.host_csr()
    D = _dense_block(P, ADJOINT, STATE, row_ess=W.ess, col_ess=zcols)
    got = csr(ip, ix, d, D.shape).toarray()
    assert rel(got, D) <= TOL
    assert np.all(got[:, zcols] == 0.0)
    # symmetric (the default): columns follow the rows, the diagonal is the policy's
    A = assemble_matrix(W.Vu, W.Vu, W.groups, E, W.mesh.ne, test_ess=W.ess)
    ip, ix, d = A.host_csr()
    D = _dense_block(P, ADJOINT, STATE, row_ess=W.ess, col_ess=W.ess, diag=1.0)
    assert rel(csr(ip, ix, d, D.shape).toarray(), D) <= TOL


@pytest.mark.parametrize("nonlinear", [False, True])
def test_gauss_newton_linearization_point(gpu, nonlinear):
    """``gauss_newton_approx=True`` assembles A and C only (``setLinearizationPoint``,
    ``PDEVariationalProblem.py:1372-1482``), to the bits of the full point's A and C."""
    W, P = pair(4, perturb=0.2, nonlinear=nonlinear)
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc,
                                 is_fwd_linear=not nonlinear)
    x = [W.u.copy(), W.m.copy(), W.p.copy()]
    gn = prob.set_linearization_point(x, gauss_newton_approx=True)
    assert sorted(gn) == sorted([(ADJOINT, STATE), (ADJOINT, PARAMETER)])
    full = prob.set_linearization_point(x)
    for ij in gn:
        assert torch.equal(gn[ij].data, full[ij].data), ij
        assert rel(gn[ij].host_csr()[2], P.assembled(*ij)[2]) <= TOL


@pytest.mark.parametrize("degree", [4, 6])
@pytest.mark.parametrize("numbering", ["lattice", "mfem"])
@pytest.mark.parametrize("nonlinear", [False, True])
def test_element_values_and_functional(gpu, nonlinear, numbering, degree):
    """``element_values`` (``hm_element_scalars``) against the oracle's element scalars, and
    ``functional`` (their ``assemble_scalar``) against their sum, on a perturbed mesh."""
    n = 5
    W = Workload(n, nonlinear=nonlinear, perturb=0.2, numbering=numbering,
                 quadrature_degree=degree)
    P = O.Problem(n, perturb=0.2, numbering=numbering,
                  density="nonlinear" if nonlinear else "linear", n1d=degree // 2 + 1)
    print("case nl=%s %s degree %d" % (nonlinear, numbering, degree), flush=True)
    got = W.kernel.element_values(W.locals())[0].cpu().numpy()
    want = P.values()
    assert got.shape == (W.mesh.ne,)
    assert rel(got, want) <= TOL
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc)
    total = prob.functional([W.u, W.m, W.p])
    assert abs(total - float(np.sum(want))) <= TOL * float(np.abs(want).sum())
    # a range of elements gives those elements' values
    part = W.kernel.element_values(W.locals(), elements=(7, 40))[0].cpu().numpy()
    assert np.array_equal(part, got[7:40])


def test_element_values_closed_form_and_nonfinite(gpu):
    """u = p = x, m = 0 on the perturbed unit cube: the functional is 1/2 (the oracle's
    closed form); a parameter whose exponential overflows raises RuntimeError."""
    W = Workload(4, perturb=0.2)
    x = W.Vu.dof_coordinates()[:, 0]
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc)
    assert abs(prob.functional([x, np.zeros(W.Vm.ndof), x]) - 0.5) <= 1e-13
    m = W.m.copy()
    m[W.Vm.ndof // 2:] = 800.0
    with pytest.raises(RuntimeError):
        prob.functional([W.u, m, W.p])


@pytest.mark.parametrize("nonlinear", [False, True])
def test_element_batch_over_budget_streams_in_chunks(gpu, nonlinear, monkeypatch):
    """An element batch above the (faked) device budget is assembled chunk by chunk, with a
    warning: every block of the linearization point has the whole batch's FP64 bits, the FP32
    values are within round-off, a non-finite entry still raises RuntimeError."""
    import hippymfem_amd.problem as PR

    W = Workload(8, nonlinear=nonlinear, perturb=0.2)
    x = [W.u, W.m, W.p]
    whole = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc, route="two-pass")
    ref = {ij: M.data.clone() for ij, M in whole.set_linearization_point(x).items()}
    ref32 = {ij: M.data.clone()
             for ij, M in whole.set_linearization_point(x, precision="fp32").items()}
    monkeypatch.setattr(PR, "ELEMENT_BUDGET", 1 << 20)   # 1 MiB: A's 3.0 MB does not fit
    before = PR.RETRIES
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc, route="two-pass")
    with pytest.warns(RuntimeWarning, match="did not fit"):
        got = prob.set_linearization_point(x)
    assert PR.RETRIES > before
    blk = prob.block_assembly(ADJOINT, STATE)
    assert blk.streamed and 1 <= blk.chunk < W.mesh.ne // 2
    assert len(blk.chunks()) > 2
    for ij, M in got.items():
        assert torch.equal(M.data, ref[ij]), ij
    got32 = prob.set_linearization_point(x, precision="fp32")
    for ij, M in got32.items():
        assert rel(M.data.cpu().numpy(), ref32[ij].cpu().numpy()) <= 1e-6, ij
    # repeated: the same bits
    again = prob.block(ADJOINT, STATE, x)
    assert torch.equal(again.data, ref[(ADJOINT, STATE)])
    m = W.m.copy()
    m[W.Vm.ndof // 2:] = 800.0
    with pytest.raises(RuntimeError):
        prob.block(ADJOINT, STATE, [W.u, m, W.p])


def test_element_groups_assemble_as_one(gpu):
    """Two element groups (contiguous element ranges, one kernel each): their element
    matrices are the whole batch's rows, and the block and the residual assembled from the
    two groups' arrays (entries in group order, ``pattern.py:437-546``) are the one-group
    ones bit for bit; a chunked pass over the groups gives the same bits."""
    from hippymfem_amd.assemble import assemble_vector
    from hippymfem_amd.elementbatch import MeshBatches
    from hippymfem_amd.kernel import QuadratureKernel

    W = Workload(4, perturb=0.2, nonlinear=True)
    B2 = MeshBatches(W.mesh, 4, groups=2)
    assert [g.ne for g in B2.groups] == [32, 32]
    with pytest.raises(NotImplementedError):
        QuadratureKernel(W.density, W.spaces, B2)
    kernels = [QuadratureKernel(W.density, W.spaces, B2.subset(g)) for g in range(2)]
    loc = W.locals()
    whole = W.kernel.element_matrices(ADJOINT, STATE, loc)[0]
    E = [k.element_matrices(ADJOINT, STATE, loc)[0] for k in kernels]
    assert torch.equal(E[0], whole[:32]) and torch.equal(E[1], whole[32:])
    A1 = assemble_matrix(W.Vu, W.Vu, W.groups, [whole], W.mesh.ne, test_ess=W.ess)
    A2 = assemble_matrix(W.Vu, W.Vu, B2.groups, E, W.mesh.ne, test_ess=W.ess)
    assert np.array_equal(np.asarray(A1.pattern.indptr), np.asarray(A2.pattern.indptr))
    assert np.array_equal(np.asarray(A1.pattern.indices), np.asarray(A2.pattern.indices))
    assert torch.equal(A1.data, A2.data)

    def chunks():
        for g, k in enumerate(kernels):
            for a in range(0, 32, 12):
                b = min(a + 12, 32)
                yield g, a, b, k.element_matrices(ADJOINT, STATE, loc, elements=(a, b))[0]

    A3 = assemble_matrix(W.Vu, W.Vu, B2.groups, chunks, W.mesh.ne, test_ess=W.ess)
    assert torch.equal(A1.data, A3.data)
    v1 = assemble_vector(W.Vu, W.groups, W.kernel.element_vectors(ADJOINT, loc), W.mesh.ne,
                         ess=W.ess)
    v2 = assemble_vector(W.Vu, B2.groups, [k.element_vectors(ADJOINT, loc)[0] for k in kernels],
                         W.mesh.ne, ess=W.ess)
    assert torch.equal(v1, v2)
