"""Parity of the HIP path with the CPU oracle (the reference's algorithm restated).

Bar (the reference's own, ``hippymfem/test/_device.py:19-49``): CSR structure bit-exact,
values within 1e-13 of the block's largest entry; given identical element matrices the
scatter is bit-exact with ``np.bincount``.
"""

import numpy as np
import pytest
import torch

from oracle import fe_oracle as O

import hippymfem_amd as hm
from hippymfem_amd import ADJOINT, PARAMETER, STATE
from hippymfem_amd.assemble import assemble_matrix, assemble_vector
from hippymfem_amd.workloads import Workload

pytestmark = pytest.mark.gpu

TOL = 1e-13

ELIM = {(ADJOINT, STATE): dict(test=True, diag_policy="one", symmetric=None),
        (STATE, STATE): dict(test=True, diag_policy="zero", symmetric=None),
        (ADJOINT, PARAMETER): dict(test=True, diag_policy="one", symmetric=False),
        (STATE, PARAMETER): dict(test=True, diag_policy="one", symmetric=False),
        (PARAMETER, PARAMETER): dict(test=False, diag_policy="one", symmetric=False)}


def pair(n, **kw):
    state = kw.pop("state", "uniform")
    W = Workload(n, state=state, **kw)
    P = O.Problem(n, order=kw.get("order", 2), perturb=kw.get("perturb", 0.0),
                  numbering=kw.get("numbering", "lattice"), state=state,
                  density="nonlinear" if kw.get("nonlinear") else "linear",
                  n1d=(kw.get("quadrature_degree") or 4) // 2 + 1)
    return W, P


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


def assemble_gpu(W, i, j, mats=None):
    rule = ELIM[(i, j)]
    mats = W.kernel.element_matrices(i, j, W.locals()) if mats is None else mats
    return assemble_matrix(W.spaces[i], W.spaces[j], W.groups, mats, W.mesh.ne,
                           test_ess=W.ess if rule["test"] else None,
                           diag_policy=rule["diag_policy"], symmetric=rule["symmetric"])


BLOCKS = [(ADJOINT, STATE), (ADJOINT, PARAMETER), (STATE, PARAMETER), (PARAMETER, PARAMETER)]


@pytest.mark.parametrize("numbering", ["lattice", "mfem", "shuffled"])
@pytest.mark.parametrize("nonlinear", [False, True])
def test_element_matrices(gpu, numbering, nonlinear):
    W, P = pair(3, numbering=numbering, nonlinear=nonlinear)
    blocks = BLOCKS + ([(STATE, STATE)] if nonlinear else [])
    for i, j in blocks:
        E = W.kernel.element_matrices(i, j, W.locals())[0].cpu().numpy()
        Eo = P.block(i, j)
        assert E.shape == Eo.shape
        assert rel(E, Eo) <= TOL, ((i, j), rel(E, Eo))


@pytest.mark.parametrize("perturb", [0.0, 0.2])
def test_element_matrices_perturbed(gpu, perturb):
    W, P = pair(4, perturb=perturb, nonlinear=True, state="fourier")
    for i, j in BLOCKS + [(STATE, STATE), (STATE, ADJOINT), (PARAMETER, STATE)]:
        E = W.kernel.element_matrices(i, j, W.locals())[0].cpu().numpy()
        assert rel(E, P.block(i, j)) <= TOL, (i, j)


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("nonlinear", [False, True])
def test_element_vectors(gpu, nonlinear, order, monkeypatch):
    """The element vectors on the sum-factorized route (tensor problems) and on the generic
    one agree with the oracle."""
    W, P = pair(4, perturb=0.2, nonlinear=nonlinear, order=order)
    for route in ("sumfact", "valu"):
        monkeypatch.setenv("HM_CONTRACTION", route)
        for i in (STATE, PARAMETER, ADJOINT):
            V = W.kernel.element_vectors(i, W.locals())[0].cpu().numpy()
            assert rel(V, P.vector(i)) <= TOL, (route, i)


def test_element_vector_and_product_ranges(gpu):
    """Element vectors and products over a sub-range of the elements are that slice of the
    whole batch, bit for bit (both routes of the vectors)."""
    W, _ = pair(5, perturb=0.2, nonlinear=True)
    K = W.kernel
    d = np.random.default_rng(3).standard_normal(W.Vu.ndof)
    for a, b in ((0, 1), (7, 61), (100, 125)):
        for i in (STATE, PARAMETER, ADJOINT):
            whole = K.element_vectors(i, W.locals())[0]
            part = K.element_vectors(i, W.locals(), elements=(a, b))[0]
            assert torch.equal(part, whole[a:b]), (a, b, i)
        whole = K.element_hvp(ADJOINT, STATE, W.locals(), d)[0]
        part = K.element_hvp(ADJOINT, STATE, W.locals(), d, elements=(a, b))[0]
        assert torch.equal(part, whole[a:b]), (a, b)


def test_element_hvp(gpu):
    W, P = pair(4, perturb=0.2, nonlinear=True)
    rng = np.random.default_rng(9)
    for i, j in BLOCKS + [(STATE, STATE)]:
        d = rng.standard_normal(W.spaces[j].ndof)
        y = W.kernel.element_hvp(i, j, W.locals(), d)[0].cpu().numpy()
        yo = O.element_hvp(P.data, P.density, i, j, P.u, P.m, P.p, d)
        assert rel(y, yo) <= TOL, (i, j)


def test_config0_jacobian_and_residual(gpu):
    """16^3 P2, linear density: A and the residual against the oracle."""
    W, P = pair(16)
    A = assemble_gpu(W, ADJOINT, STATE)
    ip, ix, d = A.host_csr()
    ipo, ixo, do = P.assembled(ADJOINT, STATE)
    assert A.nnz == 129 ** 3 == ipo[-1]
    assert np.array_equal(ip, ipo) and ip.dtype == np.int32
    assert np.array_equal(ix, ixo) and ix.dtype == np.int32
    assert rel(d, do) <= TOL
    r = assemble_vector(W.Vu, W.groups, W.kernel.element_vectors(ADJOINT, W.locals()),
                        W.mesh.ne, ess=W.ess).cpu().numpy()
    assert rel(r, P.residual(ADJOINT)) <= TOL


@pytest.mark.parametrize("numbering", ["lattice", "mfem"])
@pytest.mark.parametrize("nonlinear", [False, True])
def test_all_blocks_assembled(gpu, numbering, nonlinear):
    W, P = pair(4, numbering=numbering, nonlinear=nonlinear, perturb=0.2)
    blocks = BLOCKS + ([(STATE, STATE)] if nonlinear else [])
    for i, j in blocks:
        A = assemble_gpu(W, i, j)
        ip, ix, d = A.host_csr()
        ipo, ixo, do = P.assembled(i, j)
        assert np.array_equal(ip, ipo) and np.array_equal(ix, ixo), (i, j)
        assert rel(d, do) <= TOL, (i, j)


def test_scatter_bitexact_with_bincount(gpu):
    """Given the oracle's element matrices, the device scatter equals np.bincount bit for
    bit (ascending entry order per slot), eliminated slots and diagonal included."""
    W, P = pair(6, perturb=0.2, nonlinear=True)
    for i, j in BLOCKS + [(STATE, STATE)]:
        Eo = P.block(i, j)
        A = assemble_gpu(W, i, j, mats=[torch.as_tensor(Eo).to(gpu)])
        _, _, d = A.host_csr()
        _, _, do = P.assembled(i, j, E=Eo)
        assert np.array_equal(d, do), (i, j, np.abs(d - do).max())


@pytest.mark.parametrize("out_dtype", [torch.float32, torch.float64])
def test_scatter_fp32_entries_in_entry_order(gpu, out_dtype):
    """FP32 element matrices of 27-entry rows: every value is the FP64 sum of its entries in
    ascending entry order, rounded once, i.e.
    np.bincount of the FP64 entries.  Entries spread over 2^+-40, so that any other order of
    the sums shows in the result."""
    from hippymfem_amd.elementbatch import group_tables
    from hippymfem_amd.pattern import ScatterPattern

    W = Workload(6)
    g = W.groups[0]
    rng = np.random.default_rng(11)
    for T, C in ((W.Vu, W.Vu), (W.Vm, W.Vu)):
        S = ScatterPattern([group_tables(T, g)], [group_tables(C, g)], T, C)
        E = (rng.choice([-1.0, 1.0], S.nent) * rng.uniform(1, 2, S.nent) *
             np.exp2(rng.integers(-40, 41, S.nent))).astype(np.float32)
        got = S.data([torch.as_tensor(E).to(gpu)], dtype=out_dtype).cpu().numpy()
        ref = np.bincount(S.slot, weights=E.astype(np.float64), minlength=S.nnz)
        ref = ref.astype(np.float32) if out_dtype == torch.float32 else ref
        assert got.dtype == ref.dtype and np.array_equal(got, ref), (T.ndof, C.ndof)


def test_row_defect_of_single_values(gpu):
    """FP32 values of the square block A keep, when asked, the defect of their row sums:
    the FP64 sums (np.bincount of the FP32 element matrices) minus their FP32 roundings,
    summed over each row; values plus diag(defect) have the FP64 row sums.  Eliminated rows
    have none, and FP64 or non-square blocks (or a block not asking) keep no defect."""
    from hippymfem_amd.problem import PDEVariationalProblem

    W = Workload(6, perturb=0.2)
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc)
    loc = [torch.as_tensor(v).to(gpu) for v in W.locals()]
    blk = prob.block_assembly(ADJOINT, STATE, precision="fp32", row_sums=True)
    A = blk.assemble(loc)
    assert A.row_defect is not None and A.row_defect.dtype == torch.float64
    S = blk.pattern
    E = blk.E.cpu().numpy().astype(np.float64).reshape(-1)
    v64 = np.bincount(S.slot, weights=E, minlength=S.nnz)
    mask = O.ess_mask(W.Vu.ndof, W.ess)
    v64[O.masked_slots(S.indptr, S.indices, mask, mask)] = 0.0
    ip = S.indptr.astype(np.int64)
    rows = np.flatnonzero(mask)
    v64[ip[rows]] = 1.0
    v32 = v64.astype(np.float32)
    got = A.data.cpu().numpy()
    assert np.array_equal(got, v32)
    res = v64 - v32.astype(np.float64)
    row = np.repeat(np.arange(S.nrow), np.diff(ip))
    d_ref = np.bincount(row, weights=res, minlength=S.nrow)
    scale = np.bincount(row, weights=np.abs(res), minlength=S.nrow)
    d = A.row_defect.cpu().numpy()
    assert np.all(np.abs(d - d_ref) <= 1e-12 * scale + 1e-300)
    assert np.all(d[rows] == 0.0) and np.any(d != 0.0)
    sums64 = np.bincount(row, weights=v64, minlength=S.nrow)
    sums32 = np.bincount(row, weights=v32.astype(np.float64), minlength=S.nrow)
    size = np.bincount(row, weights=np.abs(v64), minlength=S.nrow)
    assert np.all(np.abs(sums32 + d - sums64) <= 1e-14 * size)
    assert prob.block_assembly(ADJOINT, STATE, row_sums=True).assemble(loc).row_defect is None
    assert prob.block_assembly(ADJOINT, PARAMETER, precision="fp32",
                               row_sums=True).assemble(loc).row_defect is None
    assert prob.block_assembly(ADJOINT, STATE, precision="fp32").assemble(loc).row_defect is None


def test_non_finite_raises(gpu):
    """A parameter whose exp overflows inside the density gives non-finite element values:
    the residual and the blocks raise RuntimeError, as the reference's callers do
    (PDEVariationalProblem.py:1844-1890); a non-finite parameter is refused up front."""
    from hippymfem_amd.problem import PDEVariationalProblem

    W = Workload(3)
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc)
    x = [W.u.copy(), W.m.copy(), W.p.copy()]
    prob.block(ADJOINT, STATE, x)                      # finite: no error
    x[PARAMETER][:] = 800.0                            # exp(800) overflows
    with pytest.raises(RuntimeError):
        prob.block(ADJOINT, STATE, x)
    with pytest.raises(RuntimeError):
        prob.residual(x)
    x[PARAMETER][0] = np.nan
    with pytest.raises(RuntimeError):
        prob.jacobian(x)


@pytest.mark.parametrize("degree", [4, 6])
@pytest.mark.parametrize("perturb", [0.0, 0.2])
@pytest.mark.parametrize("nonlinear", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_pair_pass_equals_single_blocks(gpu, precision, nonlinear, perturb, degree):
    """Two blocks of one shape from one pass over the elements have the bits of the two
    single-block launches, for every pair of blocks that shares a pass; pairs that do not
    (other shapes, a block that is zero for the density, BF16) are refused and write nothing.
    On the 3- and the 4-point rule."""
    W, P = pair(3, perturb=perturb, nonlinear=nonlinear, quadrature_degree=degree)
    loc = W.locals()
    K = W.kernel
    single = {(i, j): K.element_matrices(i, j, loc, precision=precision)[0].clone()
              for i in range(3) for j in range(3)}
    shared = set()
    for a in range(9):
        for b in range(9):
            ij1, ij2 = divmod(a, 3), divmod(b, 3)
            if a == b:
                continue
            got = K.element_matrices_pair(ij1, ij2, loc, precision=precision)
            if got is None:
                continue
            shared.add((ij1, ij2))
            assert (K.nd[ij1[0]], K.nd[ij1[1]]) == (K.nd[ij2[0]], K.nd[ij2[1]])
            assert bool(single[ij1].abs().max() > 0) and bool(single[ij2].abs().max() > 0)
            assert torch.equal(got[0], single[ij1]), (ij1, ij2)
            assert torch.equal(got[1], single[ij2]), (ij1, ij2)
    # C and W_um (and their transposed counterparts) share a pass at either density, and so
    # do A and W_uu of the nonlinear density (19 pointwise pairs together)
    assert ((ADJOINT, PARAMETER), (STATE, PARAMETER)) in shared
    assert ((PARAMETER, ADJOINT), (PARAMETER, STATE)) in shared
    if nonlinear:
        assert ((ADJOINT, STATE), (STATE, STATE)) in shared
    # a refused pair leaves the outputs as they were
    o1 = torch.full_like(single[(ADJOINT, STATE)], 7.0)
    o2 = torch.full_like(single[(ADJOINT, STATE)], 7.0)
    assert K.element_matrices_pair((ADJOINT, STATE), (ADJOINT, PARAMETER), loc,
                                   precision=precision, outs=(o1, o2)) is None
    assert K.element_matrices_pair((ADJOINT, STATE), (STATE, STATE), loc, precision="bf16",
                                   outs=(o1, o2)) is None
    assert bool((o1 == 7.0).all()) and bool((o2 == 7.0).all())
    many = K.element_matrices_many([(ADJOINT, PARAMETER), (STATE, PARAMETER),
                                    (ADJOINT, STATE)], loc, precision=precision)
    for ij, E in many.items():
        assert torch.equal(E[0], single[ij]), ij


@pytest.mark.parametrize("nonlinear", [False, True])
def test_linearization_point_blocks(gpu, nonlinear):
    """The blocks of a linearization point, taken two at a time from one element pass where
    they share one, are the blocks ``block`` assembles one by one, bit for bit, and match the
    oracle's assembled blocks."""
    from hippymfem_amd.problem import PDEVariationalProblem

    W, P = pair(4, perturb=0.2, nonlinear=nonlinear)
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc,
                                 is_fwd_linear=not nonlinear)
    x = [W.u.copy(), W.m.copy(), W.p.copy()]
    got = prob.set_linearization_point(x)
    want = [(ADJOINT, STATE), (ADJOINT, PARAMETER), (STATE, PARAMETER), (PARAMETER, PARAMETER)]
    if nonlinear:
        want.append((STATE, STATE))
    assert sorted(got) == sorted(want)
    # host copies first: nothing assembled later may change what is compared
    host = {ij: got[ij].host_csr() for ij in want}
    for ij in want:
        ref = prob.block(ij[0], ij[1], x)
        ip, ix, d = host[ij]
        ip1, ix1, d1 = ref.host_csr()
        assert np.array_equal(ip, ip1) and np.array_equal(ix, ix1)
        assert np.array_equal(d, d1), ij
        ipo, ixo, do = P.assembled(*ij)
        assert np.array_equal(ip, ipo) and np.array_equal(ix, ixo)
        assert rel(d, do) <= TOL, (ij, rel(d, do))


@pytest.mark.parametrize("route", ["sumfact", "mfma", "valu"])
@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16"])
def test_nonfinite_flag(gpu, route, precision, monkeypatch):
    """The element-matrix kernels set the caller's flag exactly when a value they write is
    not finite (checked against torch.isfinite on the same output), on every contraction
    route; the two-block pass too; and a linearization point raises RuntimeError."""
    from hippymfem_amd.problem import PDEVariationalProblem

    if route == "valu" and precision != "fp64" or route != "sumfact" and precision == "bf16":
        pytest.skip("the route has no such precision")
    monkeypatch.setenv("HM_CONTRACTION", route)
    W = Workload(3)
    K = W.kernel
    flag = torch.zeros(1, dtype=torch.int32, device=gpu)
    for m_value, want in ((None, 0), (800.0, 1)):
        loc = list(W.locals())
        if m_value is not None:
            loc[PARAMETER] = np.full_like(loc[PARAMETER], m_value)    # exp(800) overflows
        for ij in ((ADJOINT, STATE), (ADJOINT, PARAMETER), (PARAMETER, PARAMETER)):
            flag.zero_()
            E = K.element_matrices(*ij, loc, precision=precision, nonfinite=flag)[0]
            torch.cuda.synchronize()
            assert int(flag.item()) == int(not bool(torch.isfinite(E).all())), (ij, m_value)
            if ij == (ADJOINT, STATE):
                assert int(flag.item()) == want, (route, precision, m_value)
        if route == "sumfact" and precision != "bf16":
            flag.zero_()
            got = K.element_matrices_pair((ADJOINT, PARAMETER), (STATE, PARAMETER), loc,
                                          precision=precision, nonfinite=flag)
            torch.cuda.synchronize()
            bad = not (bool(torch.isfinite(got[0]).all()) and bool(torch.isfinite(got[1]).all()))
            assert int(flag.item()) == int(bad)
    monkeypatch.delenv("HM_CONTRACTION")
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc)
    x = [W.u.copy(), W.m.copy(), W.p.copy()]
    prob.set_linearization_point(x)
    x[PARAMETER][:] = 800.0
    with pytest.raises(RuntimeError):
        prob.set_linearization_point(x)


def test_batch_that_does_not_fit_raises_memory_error(gpu):
    """An element batch whose result does not fit raises MemoryError (the reference's
    message when no chunk fits, kernel.py:1096-1101): a result of 10^8 Q2 element matrices
    (583 GB) on a 288 GB device; a batch that fits still comes out."""
    from hippymfem_amd.kernel import _result_array

    with pytest.raises(MemoryError):
        _result_array((10 ** 8, 27, 27), torch.float64, gpu)
    W, _ = pair(4)
    assert W.kernel.element_matrices(ADJOINT, STATE, W.locals())[0].shape == (64, 27, 27)


def test_fused_chunks_equal_whole(gpu):
    W, _ = pair(6)
    whole = assemble_gpu(W, ADJOINT, STATE).data.cpu().numpy()
    for chunk in (1, 7, 50, 216):
        def chunks(chunk=chunk):
            return W.kernel.element_matrix_chunks(ADJOINT, STATE, W.locals(), chunk=chunk)
        A = assemble_matrix(W.Vu, W.Vu, W.groups, chunks, W.mesh.ne, test_ess=W.ess)
        assert np.array_equal(A.data.cpu().numpy(), whole), chunk


@pytest.mark.parametrize("nonlinear", [False, True])
def test_chunks_of_several_blocks(gpu, nonlinear):
    """element_matrix_chunks_many (C and W_um, and at the nonlinear density A and W_uu, from
    one pass a chunk) scattered chunk by chunk with scatter_many gives every block the bits of
    its whole-batch assembly."""
    from hippymfem_amd.csrassemble import finish_block, plan_block, scatter_many

    W, _ = pair(4, perturb=0.2, nonlinear=nonlinear)
    blocks = BLOCKS + ([(STATE, STATE)] if nonlinear else [])
    plans = {}
    for ij in blocks:
        rule = ELIM[ij]
        plans[ij] = plan_block(W.spaces[ij[0]], W.spaces[ij[1]], W.groups,
                               test_ess=W.ess if rule["test"] else None,
                               diag_policy=rule["diag_policy"], symmetric=rule["symmetric"])
    for chunk in (9, 64):
        chunks = W.kernel.element_matrix_chunks_many(blocks, W.locals(), chunk=chunk)
        accs = scatter_many(plans, chunks)
        for ij in blocks:
            got = finish_block(plans[ij], accs[ij]).data.cpu().numpy()
            want = assemble_gpu(W, *ij).data.cpu().numpy()
            assert np.array_equal(got, want), (chunk, ij)


def test_repeat_bit_identical(gpu):
    W, _ = pair(8)
    ref = assemble_gpu(W, ADJOINT, STATE).data.cpu().numpy()
    for _ in range(4):
        assert np.array_equal(assemble_gpu(W, ADJOINT, STATE).data.cpu().numpy(), ref)


def test_matvec_and_matrix_free(gpu):
    W, P = pair(6)
    A = assemble_gpu(W, ADJOINT, STATE)
    x = W.direction()
    y = A.mult(x).cpu().numpy()
    ip, ix, d = P.assembled(ADJOINT, STATE)
    assert rel(y, O.csr_matvec(ip, ix, d, x)) <= TOL


def test_zero_slots_and_keep_policy(gpu):
    W, P = pair(4)
    E = W.kernel.element_matrices(ADJOINT, STATE, W.locals())
    pat = hm.get_pattern(W.Vu, W.Vu, W.groups)
    mask = np.zeros(W.Vu.ndof, bool)
    mask[W.ess] = True
    kill = pat.masked_slots(mask, mask)
    d = pat.data(E, zero_slots=kill).cpu().numpy()
    ip, ix, _ = P.assembled(ADJOINT, STATE)
    Eo = P.block(ADJOINT, STATE)
    slot = O.pattern_tables_sort(P.Vu.edofs, P.Vu.edofs, P.Vu.ndof, P.Vu.ndof, True)[2]
    do = O.scatter(slot, Eo, len(ix), O.masked_slots(ip, ix, mask, mask))
    assert rel(d, do) <= TOL
    A = assemble_matrix(W.Vu, W.Vu, W.groups, E, W.mesh.ne, test_ess=W.ess,
                        diag_policy="keep")
    raw = O.scatter(slot, Eo, len(ix))
    dk = A.data.cpu().numpy()
    first = ip[W.ess]
    assert rel(dk[first], raw[first]) <= TOL
    others = np.setdiff1d(kill, first)
    assert np.all(dk[others] == 0.0)


def test_errors(gpu):
    W, _ = pair(2)
    pat = hm.get_pattern(W.Vu, W.Vu, W.groups)
    with pytest.raises(ValueError):
        pat.data([torch.zeros(3, device=gpu, dtype=torch.float64)])
    with pytest.raises(ValueError):
        W.kernel.element_matrices(ADJOINT, STATE, W.locals(), elements=(0, W.mesh.ne + 1))


def test_full_size_properties(gpu):
    """64^3 P2 (config 1): structure closed form, symmetry of A, agreement of the assembled
    product with the matrix-free one, and the oracle on a sample of elements."""
    W = Workload(64)
    A = assemble_gpu(W, ADJOINT, STATE)
    assert A.nnz == 513 ** 3
    ip = A.pattern.indptr
    assert ip[-1] == 513 ** 3 and ip.dtype == np.int32
    x = W.direction()
    y = A.mult(x)
    xz = x.copy()
    xz[W.ess] = 0.0
    yf = assemble_vector(W.Vu, W.groups, W.kernel.element_hvp(ADJOINT, STATE, W.locals(), xz),
                         W.mesh.ne, ess=W.ess)
    yf[torch.as_tensor(W.ess).to(gpu)] = torch.as_tensor(x[W.ess]).to(gpu)
    assert rel(y.cpu().numpy(), yf.cpu().numpy()) <= TOL
    # symmetric: y . A z == z . A y
    z = np.random.default_rng(5).standard_normal(W.Vu.ndof)
    az = A.mult(z).cpu().numpy()
    ay = y.cpu().numpy()
    assert abs(ay @ z - az @ x) <= 1e-11 * abs(ay @ z)
    # a sample of elements against the oracle
    P = O.Problem(64)
    elems = np.sort(np.random.default_rng(3).choice(W.mesh.ne, 512, replace=False))
    E = W.kernel.element_matrices(ADJOINT, STATE, W.locals())[0]
    Eg = E[torch.as_tensor(elems).to(gpu)].cpu().numpy()
    assert rel(Eg, P.block(ADJOINT, STATE, elems)) <= TOL


@pytest.mark.parametrize("numbering", ["lattice", "mfem"])
@pytest.mark.parametrize("nonlinear", [False, True])
def test_contraction_routes_agree(gpu, numbering, nonlinear, monkeypatch):
    """The sum-factorized, matrix-core and plain FMA contractions agree to round-off, and
    with the oracle."""
    W, P = pair(4, perturb=0.2, nonlinear=nonlinear, numbering=numbering)
    blocks = BLOCKS + ([(STATE, STATE)] if nonlinear else [])
    for i, j in blocks:
        Eo = P.block(i, j)
        for route in ("sumfact", "mfma", "valu"):
            monkeypatch.setenv("HM_CONTRACTION", route)
            E = W.kernel.element_matrices(i, j, W.locals())[0].cpu().numpy()
            assert rel(E, Eo) <= TOL, (route, i, j, rel(E, Eo))
        monkeypatch.delenv("HM_CONTRACTION")


@pytest.mark.parametrize("nonlinear", [False, True])
@pytest.mark.parametrize("route", ["sumfact", "mfma"])
def test_fp32_contraction_error(gpu, nonlinear, route, monkeypatch):
    """The FP32 form (sum-factorized or matrix-core) is within 1e-6 of the FP64 oracle,
    element matrices and assembled blocks (values accumulated from FP32 entries)."""
    monkeypatch.setenv("HM_CONTRACTION", route)
    W, P = pair(6, perturb=0.2, nonlinear=nonlinear)
    blocks = BLOCKS + ([(STATE, STATE)] if nonlinear else [])
    for i, j in blocks:
        E = W.kernel.element_matrices(i, j, W.locals(), precision="fp32")[0]
        assert E.dtype == torch.float32
        assert rel(E.cpu().numpy(), P.block(i, j)) <= 1e-6, (i, j)
        A = assemble_gpu(W, i, j, mats=[E])
        assert A.data.dtype == torch.float32
        ip, ix, d = A.host_csr()
        ipo, ixo, do = P.assembled(i, j)
        assert np.array_equal(ip, ipo) and np.array_equal(ix, ixo)
        assert rel(d, do) <= 1e-6, (i, j, rel(d, do))


# Error bounds of the low-precision forms against the FP64 oracle, relative to the largest
# entry: operands rounded to 8 (BF16) or 11 (FP16) significant bits, about three roundings
# along each product chain, FP32 sums.
LOWP_BOUND = {"bf16": 2e-2, "fp16": 3e-3}


@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_lowp_contraction_error(gpu, precision):
    """BF16 / FP16 operands with FP32 accumulation and per-element scaling: every block of
    a perturbed mesh within the bound of the FP64 oracle; the error is reported."""
    W, P = pair(6, perturb=0.2, nonlinear=True)
    for i, j in BLOCKS + [(STATE, STATE)]:
        E = W.kernel.element_matrices(i, j, W.locals(), precision=precision)[0]
        assert E.dtype == torch.float32
        err = rel(E.cpu().numpy(), P.block(i, j))
        print("%s block (%d, %d): %.2e" % (precision, i, j, err), flush=True)
        assert err <= LOWP_BOUND[precision], (i, j, err)
        assert err > 0.0                          # the operands were rounded


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("nonlinear", [False, True])
def test_default_quadrature_degree(gpu, nonlinear, order, monkeypatch):
    """The reference's default rule for P2 (degree 6, 4 x 4 x 4 points), sum-factorized with
    four points a direction: element matrices (FP64 within 1e-13, FP32 within 1e-6, BF16 and
    FP16 within their bounds) and element vectors against the oracle on the same rule; the
    matrix-core route agrees too.  P1 state fields on the same rule take the four-point
    kernels with two nodes a direction."""
    W = Workload(3, order=order, perturb=0.2, nonlinear=nonlinear, quadrature_degree=6)
    P = O.Problem(3, order=order, perturb=0.2, density="nonlinear" if nonlinear else "linear", n1d=4)
    assert W.kernel.group.nq == 64
    blocks = BLOCKS + ([(STATE, STATE)] if nonlinear else [])
    for i, j in blocks:
        Eo = P.block(i, j)
        E = W.kernel.element_matrices(i, j, W.locals())[0].cpu().numpy()
        assert rel(E, Eo) <= TOL, (i, j, rel(E, Eo))
        E32 = W.kernel.element_matrices(i, j, W.locals(), precision="fp32")[0].cpu().numpy()
        assert rel(E32, Eo) <= 1e-6, (i, j, rel(E32, Eo))
        for lowp in ("bf16", "fp16"):
            El = W.kernel.element_matrices(i, j, W.locals(), precision=lowp)[0].cpu().numpy()
            assert rel(El, Eo) <= LOWP_BOUND[lowp], (lowp, i, j, rel(El, Eo))
    for i in (STATE, PARAMETER, ADJOINT):
        V = W.kernel.element_vectors(i, W.locals())[0].cpu().numpy()
        assert rel(V, P.vector(i)) <= TOL, i
    # element ranges that split the wavefronts' element groups: the same bits as the whole pass
    full = W.kernel.element_matrices(ADJOINT, STATE, W.locals())[0].cpu()
    for a, b in ((0, 5), (5, 6), (6, 27)):
        part = W.kernel.element_matrices(ADJOINT, STATE, W.locals(), elements=(a, b))[0].cpu()
        assert torch.equal(part, full[a:b]), (a, b)
    monkeypatch.setenv("HM_CONTRACTION", "mfma")
    for i, j in blocks:
        E = W.kernel.element_matrices(i, j, W.locals())[0].cpu().numpy()
        assert rel(E, P.block(i, j)) <= TOL, ("mfma", i, j)


def test_lowp_needs_tensor_structure(gpu):
    W = Workload(2, quadrature_degree=2)            # 2 points a direction: no sum factorization
    with pytest.raises(NotImplementedError):
        W.kernel.element_matrices(ADJOINT, STATE, W.locals(), precision="bf16")


@pytest.mark.parametrize("nonlinear", [False, True])
def test_consistent_block(gpu, nonlinear):
    """The consistent single-precision block (FP32 contraction, rank-one FP64 correction):
    its products with the constants are the FP64 ones (rows, and columns when two-sided);
    the symmetric form is symmetric to the last bit; the entries stay within the FP32 form's
    error of the oracle; the plain FP32 matrix misses the constants by far more."""
    W, P = pair(4, perturb=0.2, nonlinear=nonlinear)
    blocks = [((ADJOINT, STATE), not nonlinear), ((ADJOINT, PARAMETER), False),
              ((STATE, PARAMETER), False)] + ([((STATE, STATE), True)] if nonlinear else [])
    for (i, j), sym in blocks:
        kind = "symmetric" if sym else "two-sided"
        Ec = W.kernel.element_matrices(i, j, W.locals(), consistent=kind, precision="fp32")[0]
        assert Ec.dtype == torch.float64
        Ec = Ec.cpu().numpy()
        Eo = P.block(i, j)
        K = W.kernel.element_matrices(i, j, W.locals(), precision="fp32")[0].cpu().numpy()
        scale = np.abs(Eo).max() * Eo.shape[2]
        err_rows = np.abs(Ec.sum(axis=2) - Eo.sum(axis=2)).max() / scale
        err_rows_fp32 = np.abs(K.astype(np.float64).sum(axis=2) - Eo.sum(axis=2)).max() / scale
        print("%s (%d, %d): rows %.1e (fp32 %.1e)" % (kind, i, j, err_rows, err_rows_fp32),
              flush=True)
        assert err_rows <= 1e-14, (i, j, err_rows)
        assert err_rows_fp32 > 10 * err_rows
        if sym:
            assert np.array_equal(Ec, Ec.transpose(0, 2, 1)), (i, j)
        else:
            err_cols = np.abs(Ec.sum(axis=1) - Eo.sum(axis=1)).max() / (np.abs(Eo).max() * Eo.shape[1])
            assert err_cols <= 1e-14, (i, j, err_cols)
        assert rel(Ec, Eo) <= 1e-6, (i, j)
