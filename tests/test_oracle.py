"""Pins the CPU oracle (CPU only).

Each check is chosen so that a plausible mistake in the oracle (a dropped term, a wrong sign
or index, a transposed operand) fails it: closed forms (the Gauss rule, the textbook Q2
stiffness of a cube, the domain's volume), finite differences of the functional written
directly from the density, invariants (symmetry, row sums, affinity, the u/p exchange),
brute force in extended precision, and two independent routes for the pattern.
"""

import numpy as np
import pytest

from oracle import cref
from oracle import fe_oracle as O

STATE, PARAMETER, ADJOINT = O.STATE, O.PARAMETER, O.ADJOINT


# ------------------------------------------------------------------- quadrature
def test_gauss_rule_closed_form():
    x, w = O.gauss_legendre_01(3)
    s = 0.5 * np.sqrt(3.0 / 5.0)
    assert np.allclose(x, [0.5 - s, 0.5, 0.5 + s], atol=1e-15, rtol=0)
    assert np.allclose(w, [5 / 18, 8 / 18, 5 / 18], atol=1e-15, rtol=0)
    pts, wts = O.integration_rule(3)
    # x fastest: point 1 moves in x only
    assert np.allclose(pts[1] - pts[0], [x[1] - x[0], 0, 0])
    assert np.allclose(pts[3] - pts[0], [0, x[1] - x[0], 0])
    # exact for degree <= 5 in each variable
    for a, b, c in [(5, 0, 0), (2, 3, 4), (5, 5, 5), (1, 4, 2)]:
        approx = np.sum(wts * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c)
        assert abs(approx - 1.0 / ((a + 1) * (b + 1) * (c + 1))) < 1e-15


def test_four_point_rule():
    """The 4-point rule (degree 6, the default for P2 fields): the closed-form Gauss-Legendre
    nodes 1/2 -+ 1/2 sqrt(3/7 +- 2/7 sqrt(6/5)) and weights (18 -+ sqrt 30)/72 on [0, 1]; exact
    through degree 7 in each variable and not at degree 8."""
    x, w = O.gauss_legendre_01(4)
    r = np.sqrt(6.0 / 5.0)
    inner, outer = np.sqrt(3 / 7 - 2 / 7 * r), np.sqrt(3 / 7 + 2 / 7 * r)
    assert np.allclose(x, [0.5 - 0.5 * outer, 0.5 - 0.5 * inner, 0.5 + 0.5 * inner, 0.5 + 0.5 * outer],
                       atol=1e-15, rtol=0)
    c = np.sqrt(30.0)
    assert np.allclose(w, [(18 - c) / 72, (18 + c) / 72, (18 + c) / 72, (18 - c) / 72], atol=1e-15, rtol=0)
    pts, wts = O.integration_rule(4)
    assert pts.shape == (64, 3)
    for a, b, cc in [(7, 0, 0), (2, 7, 4), (7, 7, 7), (6, 1, 5)]:
        approx = np.sum(wts * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** cc)
        assert abs(approx - 1.0 / ((a + 1) * (b + 1) * (cc + 1))) < 1e-15
    assert abs(np.sum(wts * pts[:, 0] ** 8) - 1.0 / 9) > 1e-6


# ------------------------------------------------------------------------ basis
@pytest.mark.parametrize("n1d", [3, 4])
@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("local", ["lattice", "mfem"])
def test_basis_invariants(order, local, n1d):
    nodes = O.lattice_local_nodes(order) if local == "lattice" else O.mfem_local_nodes(order)
    pts, _ = O.integration_rule(n1d)
    N, G = O.basis_tables(order, pts, nodes)
    assert np.allclose(N.sum(axis=1), 1.0, atol=1e-14)
    assert np.allclose(G.sum(axis=1), 0.0, atol=1e-13)
    Nn, _ = O.basis_tables(order, nodes, nodes)
    assert np.allclose(Nn, np.eye(len(nodes)), atol=1e-14)
    # gradients reproduce linear functions: sum_a x_a dN_a/dxi_k = e_k
    lin = np.einsum("ad,qak->qdk", nodes, G)
    assert np.allclose(lin, np.broadcast_to(np.eye(3), lin.shape), atol=1e-13)


def test_mfem_local_order():
    """Vertices first in MFEM's cube order, edge midpoints, face centres, centre."""
    nodes = O.mfem_local_nodes(2)
    assert np.array_equal(nodes[:8], O.MFEM_HEX_VERTS)
    assert np.allclose(nodes[8], [0.5, 0, 0]) and np.allclose(nodes[9], [1, 0.5, 0])
    assert np.allclose(nodes[20], [0.5, 0.5, 0]) and np.allclose(nodes[26], [0.5, 0.5, 0.5])


# --------------------------------------------------------------------- geometry
def test_geometry_cartesian():
    mesh = O.BoxMesh(4)
    pts, w = O.integration_rule(3)
    Jinv, det, wdet = O.geometry(mesh, pts, w)
    h = 0.25
    assert Jinv.shape == (64, 1, 3, 3)                   # affine: one per element
    assert np.allclose(Jinv, np.eye(3) / h, atol=1e-12)
    assert np.allclose(det, h ** 3, rtol=1e-14)
    assert abs(wdet.sum() - 1.0) < 1e-14


def test_geometry_perturbed_volume():
    mesh = O.BoxMesh(5, perturb=0.2)
    pts, w = O.integration_rule(3)
    Jinv, det, wdet = O.geometry(mesh, pts, w)
    assert Jinv.shape == (125, 27, 3, 3)                 # not affine
    assert (det > 0).all()
    # det of a trilinear map is degree <= 2 in each variable: the rule is exact, and the
    # boundary vertices stay, so the elements tile the unit cube
    assert abs(wdet.sum() - 1.0) < 1e-13


# ------------------------------------------------------------------ element arrays
def _data(n=2, perturb=0.0, numbering="lattice", n1d=3):
    mesh = O.BoxMesh(n, perturb=perturb)
    Vu = O.H1Space(mesh, 2, numbering)
    Vm = O.H1Space(mesh, 1, numbering)
    return O.ElementData(mesh, Vu, Vm, n1d), Vu, Vm


@pytest.mark.parametrize("n1d", [3, 4])
def test_constant_coefficient_is_textbook_q2(n1d):
    """kappa = 1 on a cube of side h: A_e = h (K1 x M1 x M1 + M1 x K1 x M1 + M1 x M1 x K1)
    with the 1-D P2 matrices on [0, 1]; both rules integrate these products exactly."""
    data, Vu, Vm = _data(3, n1d=n1d)
    M1 = np.array([[4, 2, -1], [2, 16, 2], [-1, 2, 4]]) / 30.0
    K1 = np.array([[7, -8, 1], [-8, 16, -8], [1, -8, 7]]) / 3.0
    Kref = (np.kron(M1, np.kron(M1, K1)) + np.kron(M1, np.kron(K1, M1))
            + np.kron(K1, np.kron(M1, M1)))
    u = np.zeros(Vu.ndof)
    m = np.zeros(Vm.ndof)
    E = O.element_matrices(data, O.Density("linear"), ADJOINT, STATE, u, m, u)
    h = 1.0 / 3
    assert np.abs(E - h * Kref).max() < 1e-14
    # and the mass-like block W_mm with u = p = x: grad u . grad p = 1
    X = Vu.dof_coordinates()[:, 0]
    Wmm = O.element_matrices(data, O.Density("linear"), PARAMETER, PARAMETER, X, m, X)
    M1q1 = np.array([[2, 1], [1, 2]]) / 6.0
    assert np.abs(Wmm - h ** 3 * np.kron(M1q1, np.kron(M1q1, M1q1))).max() < 1e-15


def test_brute_force_extended_precision():
    """One Cartesian element, variable coefficient, in long double with explicit loops."""
    data, Vu, Vm = _data(2)
    rng = np.random.default_rng(0)
    m = rng.standard_normal(Vm.ndof)
    z = np.zeros(Vu.ndof)
    E = O.element_matrices(data, O.Density("linear"), ADJOINT, STATE, z, m, z)
    e = 5
    h = np.longdouble(0.5)
    pts, w = O.integration_rule(3)
    N1, _ = O.basis_tables(1, pts, O.lattice_local_nodes(1))
    _, G2 = O.basis_tables(2, pts, O.lattice_local_nodes(2))
    me = m[Vm.edofs[e]].astype(np.longdouble)
    ref = np.zeros((27, 27), dtype=np.longdouble)
    for q in range(27):
        mq = sum(np.longdouble(N1[q, v]) * me[v] for v in range(8))
        c = np.longdouble(w[q]) * h ** 3 * np.exp(mq) / h ** 2
        for a in range(27):
            for b in range(27):
                ref[a, b] += c * sum(np.longdouble(G2[q, a, d]) * np.longdouble(G2[q, b, d])
                                     for d in range(3))
    assert np.abs(E[e] - ref.astype(np.float64)).max() <= 1e-14 * np.abs(ref).max()


def _functional(data, density, u, m, p):
    """R = sum_e R_e, the oracle's element values summed."""
    return float(np.sum(O.element_values(data, density, u, m, p)))


@pytest.mark.parametrize("perturb", [0.0, 0.2])
@pytest.mark.parametrize("kind", ["linear", "nonlinear"])
def test_element_values_closed_forms(kind, perturb):
    """On the unit cube with m = 0 (kappa = 1), f = 1: u = p = x gives
    int (g(x) - x) dx = 1/2 (g = 1) or 5/6 (g = 1 + x^2); u = 0, p = 1 gives -1 (the volume).
    x is in the Q2 space of the trilinear elements and the integrands have degree <= 4 a
    direction with det J, so the 3- and 4-point rules are exact, perturbed or not."""
    for n1d in (3, 4):
        data, Vu, Vm = _data(2, perturb=perturb, n1d=n1d)
        dens = O.Density(kind)
        x = Vu.dof_coordinates()[:, 0]
        m0 = np.zeros(Vm.ndof)
        want = 0.5 if kind == "linear" else 5.0 / 6.0
        got = float(np.sum(O.element_values(data, dens, x, m0, x)))
        assert abs(got - want) <= 1e-13, (n1d, got)
        vol = float(np.sum(O.element_values(data, dens, np.zeros(Vu.ndof), m0,
                                            np.ones(Vu.ndof))))
        assert abs(vol + 1.0) <= 1e-13, (n1d, vol)


@pytest.mark.parametrize("n1d", [3, 4])
@pytest.mark.parametrize("kind", ["linear", "nonlinear"])
def test_vectors_are_gradient_of_functional(kind, n1d):
    data, Vu, Vm = _data(2, perturb=0.2, n1d=n1d)
    dens = O.Density(kind)
    rng = np.random.default_rng(1)
    u, m, p = rng.standard_normal(Vu.ndof) * 0.5, rng.standard_normal(Vm.ndof) * 0.3, \
        rng.standard_normal(Vu.ndof)
    x = [u, m, p]
    for i, V in ((STATE, Vu), (PARAMETER, Vm), (ADJOINT, Vu)):
        g = O.assemble_vector(V.edofs, V.ndof, O.element_vectors(data, dens, i, u, m, p))
        d = rng.standard_normal(V.ndof)
        eps = 1e-5
        xp = [v.copy() for v in x]
        xm = [v.copy() for v in x]
        xp[i] = x[i] + eps * d
        xm[i] = x[i] - eps * d
        fd = (_functional(data, dens, *xp) - _functional(data, dens, *xm)) / (2 * eps)
        assert abs(fd - g @ d) <= 1e-8 * max(1.0, abs(g @ d)), (i, fd, g @ d)


@pytest.mark.parametrize("n1d", [3, 4])
@pytest.mark.parametrize("kind", ["linear", "nonlinear"])
def test_blocks_are_derivatives_of_vectors(kind, n1d):
    """Block (i, j) = d(element vectors of slot i) / d(slot j), by central differences."""
    data, Vu, Vm = _data(2, perturb=0.2, n1d=n1d)
    dens = O.Density(kind)
    rng = np.random.default_rng(2)
    u, m, p = rng.standard_normal(Vu.ndof) * 0.5, rng.standard_normal(Vm.ndof) * 0.3, \
        rng.standard_normal(Vu.ndof)
    x = [u, m, p]
    spaces = [Vu, Vm, Vu]
    for i in range(3):
        for j in range(3):
            E = O.element_matrices(data, dens, i, j, u, m, p)
            d = rng.standard_normal(spaces[j].ndof)
            Ed = np.einsum("eab,eb->ea", E, d[spaces[j].edofs])
            eps = 1e-6
            xp = [v.copy() for v in x]
            xm = [v.copy() for v in x]
            xp[j] = x[j] + eps * d
            xm[j] = x[j] - eps * d
            fd = (O.element_vectors(data, dens, i, *xp) - O.element_vectors(data, dens, i, *xm)) \
                / (2 * eps)
            scale = max(np.abs(Ed).max(), 1e-12)
            if np.abs(Ed).max() == 0.0:
                assert np.abs(fd).max() < 1e-8, (i, j)
            else:
                assert np.abs(fd - Ed).max() <= 1e-7 * scale, (i, j, np.abs(fd - Ed).max() / scale)


def test_block_invariants():
    data, Vu, Vm = _data(3, perturb=0.2)
    dens = O.Density("linear")
    rng = np.random.default_rng(3)
    u, m, p = rng.standard_normal(Vu.ndof), rng.standard_normal(Vm.ndof), \
        rng.standard_normal(Vu.ndof)
    A = O.element_matrices(data, dens, ADJOINT, STATE, u, m, p)
    assert np.abs(A - A.transpose(0, 2, 1)).max() < 1e-14 * np.abs(A).max()
    assert np.abs(A.sum(axis=2)).max() < 1e-13 * np.abs(A).max()          # A 1 = 0
    Wmm = O.element_matrices(data, dens, PARAMETER, PARAMETER, u, m, p)
    assert np.abs(Wmm - Wmm.transpose(0, 2, 1)).max() < 1e-14 * np.abs(Wmm).max()
    # W_um(u, p) is C(p, u): the two differ by exchanging the state and the adjoint
    Wum = O.element_matrices(data, dens, STATE, PARAMETER, u, m, p)
    C_swapped = O.element_matrices(data, dens, ADJOINT, PARAMETER, p, m, u)
    assert np.abs(Wum - C_swapped).max() < 1e-14 * np.abs(Wum).max()
    # (m, u) is the transpose of (u, m)
    Wmu = O.element_matrices(data, dens, PARAMETER, STATE, u, m, p)
    assert np.abs(Wmu - Wum.transpose(0, 2, 1)).max() < 1e-14 * np.abs(Wum).max()
    # W_uu vanishes for the linear density
    assert np.abs(O.element_matrices(data, dens, STATE, STATE, u, m, p)).max() == 0.0


def test_residual_affine_in_state():
    """R(u) = A u + R(0) for the linear density (``test_kernels.py:261-307``)."""
    data, Vu, Vm = _data(3, perturb=0.2)
    dens = O.Density("linear")
    rng = np.random.default_rng(4)
    u, m = rng.standard_normal(Vu.ndof), rng.standard_normal(Vm.ndof)
    z = np.zeros(Vu.ndof)
    r_u = O.assemble_vector(Vu.edofs, Vu.ndof, O.element_vectors(data, dens, ADJOINT, u, m, z))
    r_0 = O.assemble_vector(Vu.edofs, Vu.ndof, O.element_vectors(data, dens, ADJOINT, z, m, z))
    E = O.element_matrices(data, dens, ADJOINT, STATE, u, m, z)
    ip, ix, d = O.assemble_block(Vu.edofs, Vu.edofs, Vu.ndof, Vu.ndof, E, same=True)
    Au = O.csr_matvec(ip, ix, d, u)
    assert np.abs(r_u - (Au + r_0)).max() < 1e-13 * np.abs(r_u).max()


# ---------------------------------------------------------------------- pattern
@pytest.mark.parametrize("numbering", ["lattice", "mfem", "shuffled"])
@pytest.mark.parametrize("n", [1, 3])
def test_pattern_routes_agree_and_closed_form(numbering, n):
    mesh = O.BoxMesh(n)
    Vu = O.H1Space(mesh, 2, numbering)
    Vm = O.H1Space(mesh, 1, "lattice" if numbering == "shuffled" else numbering)
    for (T, C, want) in ((Vu, Vu, (8 * n + 1) ** 3 if n > 1 else 27 * 27),
                         (Vu, Vm, (5 * n + 1) ** 3 if n > 1 else 27 * 8),
                         (Vm, Vm, (3 * n + 1) ** 3 if n > 1 else 64)):
        same = T is C
        a = O.pattern_tables_sort(T.edofs, C.edofs, T.ndof, C.ndof, same)
        b = cref.pattern_tables(T.edofs, C.edofs, T.ndof, C.ndof, same)
        assert a[3] == b[3]
        for x, y in zip(a[:3], b[:3]):
            assert x.dtype == y.dtype and np.array_equal(x, y)
        ip, ix, slot, nnz = a
        if n > 1:
            assert nnz == want
        # diagonal first on a square pattern, the rest ascending
        for r in range(0, T.ndof, max(1, T.ndof // 50)):
            row = ix[ip[r]:ip[r + 1]]
            if same:
                assert row[0] == r and np.all(np.diff(row[1:]) > 0)
            else:
                assert np.all(np.diff(row) > 0)
        # every entry's slot holds its column and lies in its row
        k = np.arange(slot.size)
        e, rem = np.divmod(k, T.nd * C.nd)
        a_, b_ = np.divmod(rem, C.nd)
        assert np.array_equal(ix[slot], C.edofs[e, b_])
        assert np.all((slot >= ip[T.edofs[e, a_]]) & (slot < ip[T.edofs[e, a_] + 1]))


def test_pattern_numbering_free():
    """The same couplings in space whatever the numbering."""
    ref = None
    for numbering in ("lattice", "mfem", "shuffled"):
        mesh = O.BoxMesh(3)
        V = O.H1Space(mesh, 2, numbering)
        ip, ix, _, _ = O.pattern_tables_sort(V.edofs, V.edofs, V.ndof, V.ndof, True)
        X = V.dof_coordinates()
        pairs = O.coordinate_pairs(ip, ix, X, X)
        if ref is None:
            ref = pairs
        assert np.array_equal(pairs, ref)


def test_pattern_row_too_long():
    # 2**16 entries in one row: ValueError (patternbuild.py:233-236)
    tdofs = np.zeros((2500, 27), np.int64)
    cdofs = np.tile(np.arange(27), (2500, 1))
    with pytest.raises(ValueError):
        cref.pattern_tables(tdofs, cdofs, 1, 27, False)


# ---------------------------------------------------------------------- scatter
def test_bincount_is_sequential_sum():
    rng = np.random.default_rng(6)
    slot = rng.integers(0, 50, 5000)
    w = rng.standard_normal(5000) * 10.0 ** rng.integers(-8, 8, 5000)
    a = O.scatter(slot, w, 50)
    b = cref.bincount(slot, w, 50)
    assert np.array_equal(a, b)


def test_elimination_against_dense():
    data, Vu, Vm = _data(2, perturb=0.2)
    dens = O.Density("nonlinear")
    rng = np.random.default_rng(7)
    u, m, p = rng.standard_normal(Vu.ndof) * 0.5, rng.standard_normal(Vm.ndof), \
        rng.standard_normal(Vu.ndof)
    ess = O.boundary_dofs_x(Vu)
    for (i, j, sym, diag) in ((ADJOINT, STATE, True, 1.0), (STATE, STATE, True, 0.0),
                              (ADJOINT, PARAMETER, False, None), (PARAMETER, PARAMETER, None, None)):
        T, C = ([Vu, Vm, Vu][i], [Vu, Vm, Vu][j])
        E = O.element_matrices(data, dens, i, j, u, m, p)
        dense = np.zeros((T.ndof, C.ndof))
        for e in range(E.shape[0]):
            for a in range(T.nd):
                for b in range(C.nd):
                    dense[T.edofs[e, a], C.edofs[e, b]] += E[e, a, b]
        test_ess = ess if i != PARAMETER else None
        if test_ess is not None:
            dense[test_ess, :] = 0.0
            if sym:
                dense[:, test_ess] = 0.0
                dense[test_ess, test_ess] = diag
        ip, ix, d = O.assemble_block(T.edofs, C.edofs, T.ndof, C.ndof, E, test_ess=test_ess,
                                     diag_policy="zero" if diag == 0.0 else "one",
                                     symmetric=bool(sym), same=T is C)
        got = np.zeros_like(dense)
        rows = np.repeat(np.arange(T.ndof), np.diff(ip))
        got[rows, ix] = d
        assert np.abs(got - dense).max() <= 1e-14 * np.abs(dense).max(), (i, j)


def test_fields_and_boundary():
    P = O.Problem(4)
    assert abs(P.m.mean()) < 1e-14 and abs(P.m.std() - 1.0) < 1e-14
    assert P.u.min() >= -1.0 and P.u.max() <= 1.0
    X = P.Vu.dof_coordinates()
    assert np.all((X[P.ess, 0] == 0.0) | (X[P.ess, 0] == 1.0))
    assert P.ess.size == 2 * 9 ** 2
    Q = O.Problem(4, state="fourier")
    assert abs(np.abs(Q.u).max() - 0.5) < 1e-15
