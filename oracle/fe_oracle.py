"""numpy restatement of the reference's assembly hot path (test infrastructure only).

Every function cites the reference file:line it restates.  The reference computes its
element arrays with JAX and its mesh, numbering and geometric factors with MFEM; neither
is importable here, so the pieces that come from MFEM (mesh, dof numbering, the basis in
MFEM's local dof order) are stated here for the unit-cube hexahedral meshes of this
path and are "parity unpinned" with respect to MFEM's own numbering (the structure is
compared through the numbering-free coordinate-pair check instead).

Conventions (shared with the product, which receives the same ``edofs``):

* slots ``STATE=0`` (u, the Q2 state), ``PARAMETER=1`` (m, Q1), ``ADJOINT=2`` (p, Q2),
  as ``hippymfem/modeling/variables.py:8-14``;
* block ``(i, j)`` is ``d2R / d(slot i dofs) d(slot j dofs)``: rows test slot ``i``,
  columns trial slot ``j`` (MFEM ``elmat`` convention, ``kernel.py:1707-1709``);
* element-matrix entry ``k = (e * nd_i + a) * nd_j + b`` (``_patternkernels.py:103-105``).
"""

import numpy as np

STATE, PARAMETER, ADJOINT = 0, 1, 2


# ------------------------------------------------------------------ a1: quadrature
def gauss_legendre_01(n):
    """n-point Gauss-Legendre rule on [0, 1] (MFEM ``IntRules.Get(CUBE, deg)`` takes
    floor(deg/2)+1 of them per direction; ``elementbatch.py:97-145``)."""
    x, w = np.polynomial.legendre.leggauss(int(n))
    return 0.5 * (x + 1.0), 0.5 * w


def integration_rule(n1d=3):
    """Tensor rule on the unit cube, point ``(i, j, k)`` at ``i + n (j + n k)``
    (``elementbatch.py:126-131``, ``_tensor_places``).  Returns ``(pts (nq,3), w (nq,))``."""
    x, w = gauss_legendre_01(n1d)
    n = len(x)
    q = np.arange(n ** 3)
    ix, iy, iz = q % n, (q // n) % n, q // (n * n)
    pts = np.stack([x[ix], x[iy], x[iz]], axis=1)
    wts = w[ix] * w[iy] * w[iz]
    return pts, wts


def quadrature_points_1d_for_degree(deg):
    """``default_quadrature_degree`` (``elementbatch.py:148-156``) gives 2*order+2; MFEM
    takes floor(deg/2)+1 Gauss points per direction for it."""
    return int(deg) // 2 + 1


# --------------------------------------------------------------- a2: basis tables
def _lagrange_1d(order, x):
    """Values and derivatives of the 1-D Lagrange basis on equispaced nodes of [0, 1]
    (MFEM's GaussLobatto nodes coincide with them for order <= 2)."""
    nodes = np.linspace(0.0, 1.0, order + 1)
    x = np.asarray(x, dtype=np.float64)
    V = np.ones((x.size, order + 1))
    D = np.zeros((x.size, order + 1))
    for a in range(order + 1):
        others = [nodes[c] for c in range(order + 1) if c != a]
        den = np.prod([nodes[a] - c for c in others])
        V[:, a] = np.prod([x - c for c in others], axis=0) / den if others else 1.0
        d = np.zeros(x.size)
        for s in range(len(others)):
            t = np.ones(x.size)
            for r, c in enumerate(others):
                if r != s:
                    t = t * (x - c)
            d += t
        D[:, a] = d / den
    return V, D


def basis_tables(order, pts, local_nodes):
    """Reference shape values ``N (nq, nd)`` and gradients ``G (nq, nd, 3)`` at ``pts``
    for the tensor Lagrange element whose local dof ``a`` sits at ``local_nodes[a]``
    (``elementbatch.py:509-525``, ``_basis_tables``)."""
    nodes1 = np.linspace(0.0, 1.0, order + 1)
    idx = np.rint(local_nodes * order).astype(int)            # (nd, 3) 1-D node indices
    assert np.allclose(nodes1[idx], local_nodes)
    Vx, Dx = _lagrange_1d(order, pts[:, 0])
    Vy, Dy = _lagrange_1d(order, pts[:, 1])
    Vz, Dz = _lagrange_1d(order, pts[:, 2])
    ax, ay, az = idx[:, 0], idx[:, 1], idx[:, 2]
    N = Vx[:, ax] * Vy[:, ay] * Vz[:, az]
    G = np.stack([Dx[:, ax] * Vy[:, ay] * Vz[:, az],
                  Vx[:, ax] * Dy[:, ay] * Vz[:, az],
                  Vx[:, ax] * Vy[:, ay] * Dz[:, az]], axis=2)
    return np.ascontiguousarray(N), np.ascontiguousarray(G)


# ----------------------------------------------------------------------- meshes
def lattice_local_nodes(order):
    """Local nodes in tensor (lexicographic) order: ``a = ax + (p+1)(ay + (p+1) az)``."""
    p1 = order + 1
    a = np.arange(p1 ** 3)
    return np.stack([a % p1, (a // p1) % p1, a // (p1 * p1)], axis=1) / float(order)


#: MFEM's reference hexahedron (MFEM 4.x ``Geometry::CUBE``): vertices, edges (vertex
#: pairs) and faces (vertex quadruples).  Unpinned here (MFEM is not installed); used by
#: the ``"mfem"`` numbering hypothesis only.
MFEM_HEX_VERTS = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                           (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], dtype=float)
MFEM_HEX_EDGES = [(0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7),
                  (0, 4), (1, 5), (2, 6), (3, 7)]
MFEM_HEX_FACES = [(3, 2, 1, 0), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7),
                  (4, 5, 6, 7)]


def mfem_local_nodes(order):
    """Local dof positions in MFEM's H1 hexahedron order for order 1 or 2: vertices,
    then one dof per edge, per face and the interior (hypothesis, unpinned)."""
    if order == 1:
        return MFEM_HEX_VERTS.copy()
    if order != 2:
        raise NotImplementedError("the mfem numbering hypothesis covers orders 1 and 2")
    V = MFEM_HEX_VERTS
    pts = [v for v in V]
    pts += [0.5 * (V[a] + V[b]) for a, b in MFEM_HEX_EDGES]
    pts += [V[list(f)].mean(axis=0) for f in MFEM_HEX_FACES]
    pts += [np.array([0.5, 0.5, 0.5])]
    return np.array(pts)


class BoxMesh:
    """Unit cube split into ``n^3`` hexahedra, elements lexicographic
    ``e = ex + n (ey + n ez)``, vertices on the ``(n+1)^3`` lattice, element vertex
    ``v = vx + 2 vy + 4 vz``.  ``perturb`` moves interior vertices by
    ``U[-perturb*h, perturb*h]^3`` (seeded); the map of each element is trilinear."""

    def __init__(self, n, perturb=0.0, seed=4):
        self.n = n = int(n)
        self.h = 1.0 / n
        n1 = n + 1
        i = np.arange(n1 ** 3)
        I, J, K = i % n1, (i // n1) % n1, i // (n1 * n1)
        X = np.stack([I / n, J / n, K / n], axis=1).astype(np.float64)
        if perturb:
            rng = np.random.default_rng(seed)
            d = rng.uniform(-perturb * self.h, perturb * self.h, size=X.shape)
            interior = ((I > 0) & (I < n) & (J > 0) & (J < n) & (K > 0) & (K < n))
            X = X + d * interior[:, None]
        self.verts = X
        e = np.arange(n ** 3)
        ex, ey, ez = e % n, (e // n) % n, e // (n * n)
        v = np.arange(8)
        vx, vy, vz = v % 2, (v // 2) % 2, v // 4
        self.elem_verts = ((ex[:, None] + vx) + n1 * ((ey[:, None] + vy)
                           + n1 * (ez[:, None] + vz))).astype(np.int64)
        self.elem_ijk = np.stack([ex, ey, ez], axis=1)
        self.ne = n ** 3

    def map_points(self, ref):
        """Physical coordinates ``(ne, npts, 3)`` of reference points ``ref (npts, 3)``."""
        N1, _ = basis_tables(1, ref, lattice_local_nodes(1))
        x = self.verts[self.elem_verts]                      # (ne, 8, 3)
        return np.einsum("pv,evd->epd", N1, x)


class H1Space:
    """H1 space of order 1 or 2 on a :class:`BoxMesh`: ``edofs (ne, nd)``, the number of
    dofs, the local node table and the (lattice) position of every dof.

    ``numbering``: ``"lattice"`` numbers the dofs on the ``(p n + 1)^3`` lattice
    (x fastest) with tensor local order; ``"mfem"`` is the hypothesis of MFEM's order
    (vertices, then edges, faces and interiors by first appearance over the elements;
    MFEM local order) -- unpinned; ``"shuffled"`` is the lattice numbering under a seeded
    random permutation (to show that nothing depends on the numbering)."""

    def __init__(self, mesh, order, numbering="lattice", seed=11):
        self.mesh, self.order, self.numbering = mesh, int(order), numbering
        n, p = mesh.n, int(order)
        L = p * n + 1
        self.lattice_n = L
        if numbering == "mfem":
            self.local_nodes = mfem_local_nodes(p)
        else:
            self.local_nodes = lattice_local_nodes(p)
        self.nd = self.local_nodes.shape[0]
        loc = np.rint(self.local_nodes * p).astype(np.int64)     # (nd, 3)
        ijk = mesh.elem_ijk * p                                   # (ne, 3)
        lat = ijk[:, None, :] + loc[None, :, :]                   # (ne, nd, 3)
        lat_index = lat[..., 0] + L * (lat[..., 1] + L * lat[..., 2])
        nlat = L ** 3
        if numbering == "lattice":
            to_dof = np.arange(nlat)
        elif numbering == "shuffled":
            to_dof = np.random.default_rng(seed).permutation(nlat)
        elif numbering == "mfem":
            to_dof = self._mfem_order(lat_index, loc, p, L, n)
        else:
            raise ValueError(numbering)
        self.lattice_to_dof = to_dof.astype(np.int64)
        self.edofs = np.ascontiguousarray(to_dof[lat_index]).astype(np.int64)
        self.ndof = nlat
        inv = np.empty(nlat, np.int64)
        inv[to_dof] = np.arange(nlat)
        self.dof_to_lattice = inv

    @staticmethod
    def _mfem_order(lat_index, loc, p, L, n):
        """Vertex dofs first (vertex lattice order), then the other dofs by first
        appearance over the elements in element order and MFEM's local order."""
        nlat = L ** 3
        to_dof = np.full(nlat, -1, np.int64)
        i = np.arange(nlat)
        I, J, K = i % L, (i // L) % L, i // (L * L)
        isv = (I % p == 0) & (J % p == 0) & (K % p == 0)
        vlat = i[isv]
        to_dof[vlat] = np.arange(vlat.size)
        nxt = vlat.size
        # edges, then faces, then interiors, each in order of first appearance
        kinds = np.count_nonzero(loc % p != 0, axis=1)            # 0 vertex 1 edge 2 face 3 int
        for kind in (1, 2, 3):
            cols = np.flatnonzero(kinds == kind)
            seq = lat_index[:, cols].reshape(-1)
            _, first = np.unique(seq, return_index=True)
            order = seq[np.sort(first)]
            to_dof[order] = nxt + np.arange(order.size)
            nxt += order.size
        assert nxt == nlat and (to_dof >= 0).all()
        return to_dof

    def dof_coordinates(self):
        """Physical coordinates of every dof, ``(ndof, 3)``, through the element maps."""
        X = self.mesh.map_points(self.local_nodes)                # (ne, nd, 3)
        out = np.empty((self.ndof, 3))
        out[self.edofs.reshape(-1)] = X.reshape(-1, 3)
        return out

    def lattice_field(self, values_on_lattice):
        """A dof vector from values given in lattice order."""
        out = np.empty(self.ndof)
        out[self.lattice_to_dof] = values_on_lattice
        return out


def boundary_dofs_x(space, tol=1e-12):
    """Dofs on x = 0 and x = 1 (Dirichlet u = 0 there), selected geometrically."""
    X = space.dof_coordinates()
    return np.flatnonzero((np.abs(X[:, 0]) < tol) | (np.abs(X[:, 0] - 1.0) < tol))


# ------------------------------------------------------------- synthetic fields
def smoothed_field(n, seed=0, passes=3):
    """Parameter field on the ``(n+1)^3`` vertex lattice: seeded N(0,1) white noise,
    ``passes`` of the 7-point average (over the neighbours that exist), then shifted
    and scaled to mean 0, std 1.  Lattice order."""
    n1 = n + 1
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((n1, n1, n1))                     # [k, j, i]
    for _ in range(passes):
        s = f.copy()
        c = np.ones_like(f)
        for ax in range(3):
            for sh in (1, -1):
                sl_dst = [slice(None)] * 3
                sl_src = [slice(None)] * 3
                if sh == 1:
                    sl_dst[ax], sl_src[ax] = slice(1, None), slice(None, -1)
                else:
                    sl_dst[ax], sl_src[ax] = slice(None, -1), slice(1, None)
                s[tuple(sl_dst)] += f[tuple(sl_src)]
                c[tuple(sl_dst)] += 1.0
        f = s / c
    f = f.reshape(-1)
    f = (f - f.mean()) / f.std()
    return f


def uniform_field(nlat, seed=1, lo=-1.0, hi=1.0):
    return np.random.default_rng(seed).uniform(lo, hi, nlat)


def normal_field(nlat, seed=2):
    return np.random.default_rng(seed).standard_normal(nlat)


def fourier_field(coords, seed=1, modes=8, amplitude=0.5):
    """Smooth state: ``sum_k a_k sin(2 pi kvec_k . x + phi_k)`` over ``modes`` seeded
    integer wave vectors in ``{0..3}^3`` (not all zero), scaled so that its largest
    magnitude over the given points is ``amplitude``."""
    rng = np.random.default_rng(seed)
    kv = rng.integers(0, 4, size=(modes, 3))
    kv[np.all(kv == 0, axis=1)] = 1
    a = rng.standard_normal(modes)
    ph = rng.uniform(0.0, 2.0 * np.pi, modes)
    u = np.sin(2.0 * np.pi * coords @ kv.T + ph) @ a
    return amplitude * u / np.abs(u).max()


# ------------------------------------------------------------------ a4: geometry
AFFINE_TOL = 1e-12


def geometry(mesh, pts, w, compact_affine=True):
    """``(Jinv, detJ, wdet)`` of every element at every point.

    Restates ``ElementGroup._geometry_sliced`` (``elementbatch.py:296-352``): ``J`` is
    one product of the element's nodal coordinates with the nodal basis gradients,
    ``det``/``inv`` by numpy, ``wdet = w |det J|``; then ``_compact_affine``
    (``elementbatch.py:206-223``): where every element's inverse Jacobians agree across
    the points to ``AFFINE_TOL`` of their largest entry, one (the mean) is kept per
    element, shape ``(ne, 1, 3, 3)``."""
    _, G1 = basis_tables(1, pts, lattice_local_nodes(1))       # (nq, 8, 3)
    x = mesh.verts[mesh.elem_verts]                              # (ne, 8, 3)
    nq = pts.shape[0]
    Gm = np.ascontiguousarray(np.transpose(G1, (1, 0, 2)).reshape(8, nq * 3))
    xm = np.ascontiguousarray(np.transpose(x, (0, 2, 1)).reshape(-1, 8))   # (ne*3, 8)
    J = (xm @ Gm).reshape(mesh.ne, 3, nq, 3).transpose(0, 2, 1, 3)        # [e,q,d,k]
    det = np.linalg.det(J)
    Jinv = np.linalg.inv(J)                                      # [e,q,k,d] = dxi_k/dx_d
    wdet = w[None, :] * np.abs(det)
    if compact_affine:
        mean = Jinv.mean(axis=1, keepdims=True)
        off = np.abs(Jinv - mean).max(axis=(1, 2, 3))
        if not (off > AFFINE_TOL * np.abs(mean).max(axis=(1, 2, 3))).any():
            Jinv = mean
    return Jinv, det, wdet


def full_jinv(Jinv, nq):
    if Jinv.shape[1] == nq:
        return Jinv
    return np.broadcast_to(Jinv, (Jinv.shape[0], nq, 3, 3))


# ------------------------------------------------------------------ a5/a6: fields
def eval_field(dofs_e, N, G, Jinv):
    """Values and physical gradients at the points (``kernel.py:378-394``):
    ``val_q = sum_i N[q,i] d_i``; ``grad_q = sum_i sum_k G[q,i,k] d_i Jinv[q,k,:]``."""
    val = dofs_e @ N.T                                           # (ne, nq)
    gref = np.einsum("qik,ei->eqk", G, dofs_e)                   # (ne, nq, 3)
    Jf = full_jinv(Jinv, N.shape[0])
    grad = np.einsum("eqk,eqkd->eqd", gref, Jf)
    return val, grad


def physical_gradients(G, Jinv):
    """``G @ Jinv`` per point: ``(ne, nq, nd, 3)`` (``kernel.py:1465-1468``)."""
    return np.einsum("qil,eqlk->eqik", G, full_jinv(Jinv, G.shape[0]))


# ---------------------------------------------------------------- densities
class Density:
    """``r = exp(m) g(u) grad u . grad p - f p`` with ``g = 1`` (``"linear"``) or
    ``g = 1 + u^2`` (``"nonlinear"``): the two densities of this path
    (``benchmarks/bench_assembly.py:123-131`` has the first; the second is the
    nonlinear coefficient of the 96^3 configuration)."""

    def __init__(self, kind="linear", f=1.0):
        if kind not in ("linear", "nonlinear"):
            raise ValueError(kind)
        self.kind, self.f = kind, float(f)

    def g(self, u):
        if self.kind == "linear":
            return np.ones_like(u), np.zeros_like(u), np.zeros_like(u)
        return 1.0 + u * u, 2.0 * u, 2.0 * np.ones_like(u)


# ------------------------------------------------------------- a8, a10: element arrays
class ElementData:
    """Everything the element arrays of one configuration need, on the host."""

    def __init__(self, mesh, Vu, Vm, n1d=3, compact_affine=True):
        self.mesh, self.Vu, self.Vm = mesh, Vu, Vm
        self.pts, self.w = integration_rule(n1d)
        self.Nu, self.Gu = basis_tables(Vu.order, self.pts, Vu.local_nodes)
        self.Nm, self.Gm = basis_tables(Vm.order, self.pts, Vm.local_nodes)
        self.Jinv, self.detJ, self.wdet = geometry(mesh, self.pts, self.w, compact_affine)

    def tables(self, slot):
        if slot == PARAMETER:
            return self.Nm, self.Gm, self.Vm.edofs
        return self.Nu, self.Gu, self.Vu.edofs

    def fields(self, u, m, p, elems=slice(None)):
        """Per point ``(u, grad u, m, p, grad p)`` for the elements ``elems``."""
        Jinv = self.Jinv[elems]
        uv, ug = eval_field(u[self.Vu.edofs[elems]], self.Nu, self.Gu, Jinv)
        mv, _ = eval_field(m[self.Vm.edofs[elems]], self.Nm, self.Gm, Jinv)
        pv, pg = eval_field(p[self.Vu.edofs[elems]], self.Nu, self.Gu, Jinv)
        return uv, ug, mv, pv, pg


def point_blocks(density, s, j, F):
    """The density's second derivatives at every point for test slot ``s`` and trial
    slot ``j``: ``(vv, vg, gv, gg)`` = d2r / d(s value|grad) d(j value|grad), the
    ``D[s][r][c]`` of ``kernel.py:1581-1604`` stated in closed form.  ``vg``/``gv`` are
    vectors over the 3 components of the gradient, ``gg`` a scalar times the identity.
    ``None`` for a part that is identically zero (``_zero_point_blocks``)."""
    uv, ug, mv, pv, pg = F
    kap = np.exp(mv)
    g, g1, g2 = density.g(uv)
    dot = np.einsum("eqd,eqd->eq", ug, pg)
    key = (s, j)
    if key == (STATE, STATE):
        return (kap * g2 * dot, (kap * g1)[..., None] * pg, (kap * g1)[..., None] * pg, None)
    if key == (STATE, PARAMETER):
        return (kap * g1 * dot, None, (kap * g)[..., None] * pg, None)
    if key == (PARAMETER, STATE):
        return (kap * g1 * dot, (kap * g)[..., None] * pg, None, None)
    if key == (STATE, ADJOINT):
        return (None, (kap * g1)[..., None] * ug, None, kap * g)
    if key == (ADJOINT, STATE):
        return (None, None, (kap * g1)[..., None] * ug, kap * g)
    if key == (PARAMETER, PARAMETER):
        return (kap * g * dot, None, None, None)
    if key == (PARAMETER, ADJOINT):
        return (None, (kap * g)[..., None] * ug, None, None)
    if key == (ADJOINT, PARAMETER):
        return (None, None, (kap * g)[..., None] * ug, None)
    if key == (ADJOINT, ADJOINT):
        return (None, None, None, None)
    raise ValueError(key)


def _contract(wdet, parts, Nt, Gt, Nc, Gc):
    """``E = sum_q Psi_t^T (wdet D) Psi_c`` with ``Psi = (N, G Jinv)``
    (``kernel.py:1452-1472``); ``Gt``/``Gc`` are physical gradients (ne, nq, nd, 3)."""
    vv, vg, gv, gg = parts
    ne = wdet.shape[0]
    E = np.zeros((ne, Nt.shape[1], Nc.shape[1]))
    if vv is not None:
        E += np.einsum("eq,qa,qb->eab", wdet * vv, Nt, Nc, optimize=True)
    if vg is not None:
        E += np.einsum("eq,qa,eqd,eqbd->eab", wdet, Nt, vg, Gc, optimize=True)
    if gv is not None:
        E += np.einsum("eq,eqad,eqd,qb->eab", wdet, Gt, gv, Nc, optimize=True)
    if gg is not None:
        E += np.einsum("eq,eqad,eqbd->eab", wdet * gg, Gt, Gc, optimize=True)
    return E


def element_matrices(data, density, i, j, u, m, p, elems=slice(None)):
    """Block ``(i, j)`` element matrices ``(ne, nd_i, nd_j)``: the quadrature route
    (``_hess_slot_quadrature_fn``, ``kernel.py:1431-1485``) with the pointwise blocks in closed
    form: the second derivatives of ``r = exp(m) g(u) grad u . grad p - f p`` by slot pair, the
    ones the reference's ``_point_blocks`` forms by automatic differentiation
    (``kernel.py:1581-1646``)."""
    F = data.fields(u, m, p, elems)
    Jinv = data.Jinv[elems]
    Nt, Gt, _ = data.tables(i)
    Nc, Gc, _ = data.tables(j)
    parts = point_blocks(density, i, j, F)
    return _contract(data.wdet[elems], parts, Nt, physical_gradients(Gt, Jinv),
                     Nc, physical_gradients(Gc, Jinv))


def element_values(data, density, u, m, p, elems=slice(None)):
    """Element scalars ``R_e = sum_q w |det J| (exp(m) g(u) grad u . grad p - f p)``: the
    functional every derivative is taken of (``_functional``, ``kernel.py:977-1009``;
    ``residual_value``/``element_values``, ``kernel.py:1859-1865``, ``:2129-2136``)."""
    uv, ug, mv, pv, pg = data.fields(u, m, p, elems)
    g, _, _ = density.g(uv)
    r = np.exp(mv) * g * np.einsum("eqd,eqd->eq", ug, pg) - density.f * pv
    return np.sum(data.wdet[elems] * r, axis=1)


def element_vectors(data, density, i, u, m, p, elems=slice(None)):
    """``dR_e / d(slot i dofs)`` (``grad_block``, ``kernel.py:1343-1355``), closed form."""
    uv, ug, mv, pv, pg = data.fields(u, m, p, elems)
    kap = np.exp(mv)
    g, g1, _ = density.g(uv)
    dot = np.einsum("eqd,eqd->eq", ug, pg)
    wdet = data.wdet[elems]
    Jinv = data.Jinv[elems]
    Nt, Gt, _ = data.tables(i)
    Gp = physical_gradients(Gt, Jinv)
    if i == ADJOINT:
        val = -density.f * np.ones_like(uv)
        grad = (kap * g)[..., None] * ug
    elif i == STATE:
        val = kap * g1 * dot
        grad = (kap * g)[..., None] * pg
    else:
        val = kap * g * dot
        grad = None
    out = np.einsum("eq,qa->ea", wdet * val, Nt)
    if grad is not None:
        out += np.einsum("eq,eqad,eqd->ea", wdet, Gp, grad)
    return out


def element_hvp(data, density, i, j, u, m, p, d, elems=slice(None)):
    """Block ``(i, j)`` applied to the slot-``j`` direction ``d`` element by element
    (``hvp_block``, ``kernel.py:1756-1782``)."""
    E = element_matrices(data, density, i, j, u, m, p, elems)
    _, _, ed = data.tables(j)
    return np.einsum("eab,eb->ea", E, d[ed[elems]])


# --------------------------------------------------------------- a14: the pattern
def pattern_sort(rows, cols, nrow, ncol, diag_first):
    """The sort route of ``ScatterPattern._build`` (``pattern.py:548-663``): one packed
    key per entry (column coded 0 for the diagonal of a square pattern, else col+1),
    a stable argsort, ``cumsum`` of key changes for the slots.  Returns
    ``(indptr int32, indices int32, slot, nnz)``."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    n = rows.size
    if max(nrow, ncol) > (1 << 32) - 1:
        raise ValueError("a rank holds %d rows and %d columns, which does not fit the "
                         "packed (row, column) key" % (nrow, ncol))
    if n == 0:
        return (np.zeros(nrow + 1, np.int32), np.zeros(0, np.int32),
                np.zeros(0, np.intp), 0)
    c = cols.copy()
    if diag_first:
        on = c == rows
        c += 1
        c[on] = 0
    key = (rows << 32) | c
    order = np.argsort(key, kind="stable")
    skey = key[order]
    first = np.empty(n, dtype=bool)
    first[0] = True
    np.not_equal(skey[1:], skey[:-1], out=first[1:])
    slot_sorted = np.cumsum(first) - 1
    nnz = int(slot_sorted[-1]) + 1
    uk = skey[first]
    ucol = uk & ((1 << 32) - 1)
    urow = uk >> 32
    if diag_first:
        ucol = np.where(ucol == 0, urow, ucol - 1)
    counts = np.bincount(urow, minlength=nrow)
    indptr = np.zeros(nrow + 1, np.int64)
    np.cumsum(counts, out=indptr[1:])
    if indptr[-1] > np.iinfo(np.int32).max:
        raise ValueError("the sparsity pattern has %d entries, beyond int32" % indptr[-1])
    idx_t = np.int32 if nnz <= np.iinfo(np.int32).max else np.intp
    slot = np.empty(n, dtype=idx_t)
    slot[order] = slot_sorted
    return indptr.astype(np.int32), ucol.astype(np.int32), slot, nnz


def pattern_tables_sort(tdofs, cdofs, nrow, ncol, diag_first):
    """The sort route over the flat rows/columns of one group's tables
    (``ScatterPattern.__init__``, ``pattern.py:488-546``)."""
    ne, a = tdofs.shape
    b = cdofs.shape[1]
    rows = np.broadcast_to(tdofs[:, :, None], (ne, a, b)).reshape(-1)
    cols = np.broadcast_to(cdofs[:, None, :], (ne, a, b)).reshape(-1)
    return pattern_sort(rows, cols, nrow, ncol, diag_first)


def runs_of_rows(tdofs, nrow):
    """``(rcount, runs_of_row)`` of ``_patternkernels.group_table`` (``:110-127``):
    element rows ``t = e*nd_t + a`` grouped by their dof, ascending within a dof."""
    t_row = np.asarray(tdofs, dtype=np.int64).reshape(-1)
    order = np.argsort(t_row, kind="stable")
    rcount = np.zeros(nrow + 1, np.int64)
    np.cumsum(np.bincount(t_row, minlength=nrow), out=rcount[1:])
    return rcount, order.astype(np.int64)


def coordinate_pairs(indptr, indices, Xrow, Xcol):
    """The numbering-free form of a pattern: the sorted set of (row coordinates, column
    coordinates) pairs, rounded to 1e-9."""
    rows = np.repeat(np.arange(indptr.size - 1), np.diff(indptr))
    P = np.concatenate([Xrow[rows], Xcol[indices]], axis=1)
    P = np.round(P * 1e9).astype(np.int64)
    return P[np.lexsort(P.T[::-1])]


# ----------------------------------------------------------- a16: the scatter
def scatter(slot, flat, nnz, zero_slots=None):
    """``np.bincount(slot, weights, minlength=nnz)``, then the eliminated slots zeroed
    (``ScatterPattern.data``, ``pattern.py:749-791``)."""
    out = np.bincount(slot, weights=np.asarray(flat, dtype=np.float64).reshape(-1),
                      minlength=nnz)
    if zero_slots is not None and len(zero_slots):
        out[zero_slots] = 0.0
    return out


# -------------------------------------------------------- a18: folded elimination
def masked_slots(indptr, indices, row_mask, col_mask):
    """CSR slots in masked rows or columns (``ScatterPattern.masked_slots``,
    ``pattern.py:693-713``)."""
    nnz = int(indptr[-1])
    hit = (np.repeat(row_mask, np.diff(indptr.astype(np.int64)))
           if row_mask is not None else np.zeros(nnz, dtype=bool))
    if col_mask is not None:
        hit |= col_mask[indices]
    return np.flatnonzero(hit)


def ess_mask(n, ess):
    m = np.zeros(n, dtype=bool)
    if ess is not None:
        m[np.asarray(ess, dtype=np.int64)] = True
    return m


def assemble_block(tdofs, cdofs, nrow, ncol, E, test_ess=None, diag_policy="one",
                   symmetric=None, same=None):
    """One block as ``(indptr, indices, data)``: pattern, scatter, folded elimination and
    the eliminated diagonal (``csrassemble.plan_block``/``finish_block``,
    ``csrassemble.py:164-270``; ``elimination._set_eliminated_diagonal``,
    ``elimination.py:53-72``).  ``same``: both sides are one space (diagonal first);
    ``symmetric``: columns eliminated too (``PDEVariationalProblem._symmetric``)."""
    if same is None:
        same = nrow == ncol and tdofs is cdofs
    indptr, indices, slot, nnz = pattern_tables_sort(tdofs, cdofs, nrow, ncol, same)
    sym = same and (symmetric is None or bool(symmetric))
    kill = None
    if test_ess is not None:
        rmask = ess_mask(nrow, test_ess)
        cmask = rmask if sym else None
        kill = masked_slots(indptr, indices, rmask, cmask)
    data = scatter(slot, E, nnz, kill)
    if sym and test_ess is not None and len(test_ess):
        r = np.asarray(test_ess, dtype=np.int64)
        pos = indptr[r].astype(np.int64)
        if not np.all(indices[pos] == r):
            raise RuntimeError("an eliminated row has no diagonal entry")
        data[pos] = 0.0 if diag_policy == "zero" else 1.0
    return indptr, indices, data


def assemble_vector(edofs, n, vecs, ess=None):
    """``np.bincount(edofs, flat, minlength=n)`` (``VectorPattern.scatter``,
    ``pattern.py:1186-1209``), ``P^T`` the identity on one rank, then
    ``zero_essential`` (``bcs.py:80-102``)."""
    out = np.bincount(np.asarray(edofs).reshape(-1),
                      weights=np.asarray(vecs, dtype=np.float64).reshape(-1), minlength=n)
    if ess is not None and len(ess):
        out[np.asarray(ess, dtype=np.int64)] = 0.0
    return out


def csr_matvec(indptr, indices, data, x):
    rows = np.repeat(np.arange(indptr.size - 1), np.diff(indptr))
    return np.bincount(rows, weights=data * x[indices], minlength=indptr.size - 1)


# -------------------------------------------------------------- configurations
class Problem:
    """One configuration of the path: mesh, spaces, geometry, seeded fields.

    ``m``: smoothed white noise (seed 0); ``u``: U[-1,1] (seed 1) or the smooth Fourier
    field (``state="fourier"``); ``p``: N(0,1) (seed 2); Dirichlet dofs of the state on
    x = 0 and x = 1."""

    def __init__(self, n, order=2, perturb=0.0, numbering="lattice", state="uniform",
                 density="linear", n1d=3, compact_affine=True):
        self.mesh = BoxMesh(n, perturb=perturb)
        self.Vu = H1Space(self.mesh, order, numbering)
        self.Vm = H1Space(self.mesh, 1, "lattice" if numbering == "shuffled" else numbering)
        self.density = Density(density)
        self.data = ElementData(self.mesh, self.Vu, self.Vm, n1d, compact_affine)
        nl = self.Vu.ndof
        self.m = self.Vm.lattice_field(smoothed_field(n, seed=0))
        if state == "fourier":
            self.u = fourier_field(self.Vu.dof_coordinates(), seed=1)
        else:
            self.u = self.Vu.lattice_field(uniform_field(nl, seed=1))
        self.p = self.Vu.lattice_field(normal_field(nl, seed=2))
        self.ess = boundary_dofs_x(self.Vu)

    def space(self, slot):
        return self.Vm if slot == PARAMETER else self.Vu

    def block(self, i, j, elems=slice(None)):
        return element_matrices(self.data, self.density, i, j, self.u, self.m, self.p, elems)

    def vector(self, i, elems=slice(None)):
        return element_vectors(self.data, self.density, i, self.u, self.m, self.p, elems)

    def values(self, elems=slice(None)):
        return element_values(self.data, self.density, self.u, self.m, self.p, elems)

    #: elimination of each block (``PDEVariationalProblem.py:1448-1451``, ``:597-598``)
    ELIMINATION = {(ADJOINT, STATE): dict(test=True, diag_policy="one"),
                   (STATE, STATE): dict(test=True, diag_policy="zero"),
                   (ADJOINT, PARAMETER): dict(test=True, diag_policy="one"),
                   (STATE, PARAMETER): dict(test=True, diag_policy="one"),
                   (PARAMETER, PARAMETER): dict(test=False, diag_policy="one")}

    def assembled(self, i, j, E=None):
        Vi, Vj = self.space(i), self.space(j)
        E = self.block(i, j) if E is None else E
        rule = self.ELIMINATION[(i, j)]
        sym = PARAMETER not in (i, j)
        same = Vi is Vj
        return assemble_block(Vi.edofs, Vj.edofs, Vi.ndof, Vj.ndof, E,
                              test_ess=self.ess if rule["test"] else None,
                              diag_policy=rule["diag_policy"], symmetric=sym, same=same)

    def residual(self, i=ADJOINT):
        V = self.space(i)
        return assemble_vector(V.edofs, V.ndof, self.vector(i),
                               self.ess if i != PARAMETER else None)
