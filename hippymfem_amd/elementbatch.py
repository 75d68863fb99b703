"""Element batches: the quadrature rule, the reference basis tables and the device arrays
the element kernels read (``hippymfem/fem/elementbatch.py``).

The reference precomputes per element and point ``Jinv`` and ``wdet`` on the host and
streams them to the device (``elementbatch.py:232-352``).  Here only the vertex
coordinates and element tables go to the device once; the element map's Jacobian, its
determinant and inverse are computed in FP64 inside the kernels at every point, which reads
24 doubles an element instead of 27 x 10.
"""

import numpy as np
import torch

from . import _lib


def gauss_legendre(n):
    """``n``-point Gauss-Legendre rule on [0, 1]."""
    x, w = np.polynomial.legendre.leggauss(int(n))
    return 0.5 * (x + 1.0), 0.5 * w


def integration_rule(degree):
    """MFEM's Gauss-Legendre rule of ``degree`` on the cube (``floor(degree/2)+1`` points
    per direction), points ``(i, j, k)`` at ``i + n (j + n k)`` (``elementbatch.py:97-145``).
    Returns ``(points (nq, 3), weights (nq,))``."""
    n = int(degree) // 2 + 1
    x, w = gauss_legendre(n)
    q = np.arange(n ** 3)
    ix, iy, iz = q % n, (q // n) % n, q // (n * n)
    return np.stack([x[ix], x[iy], x[iz]], axis=1), w[ix] * w[iy] * w[iz]


def default_quadrature_degree(orders, extra=2):
    """``2 * max(order) + extra`` (``elementbatch.py:148-156``)."""
    return 2 * max(orders) + extra


def _lagrange(order, x):
    """1-D Lagrange basis on equispaced nodes of [0, 1] and its derivative, ``(len(x), p+1)``."""
    nodes = np.linspace(0.0, 1.0, order + 1)
    x = np.asarray(x, dtype=np.float64)
    V = np.ones((x.size, order + 1))
    D = np.zeros((x.size, order + 1))
    for a in range(order + 1):
        for c in range(order + 1):
            if c == a:
                continue
            # product rule over the factors (x - x_c) / (x_a - x_c)
            s = 1.0 / (nodes[a] - nodes[c])
            D[:, a] = D[:, a] * (x - nodes[c]) * s + V[:, a] * s
            V[:, a] = V[:, a] * (x - nodes[c]) * s
    return V, D


def basis_tables(order, points, local_nodes):
    """Reference values ``N (nq, nd)`` and gradients ``G (nq, nd, 3)`` of the tensor
    Lagrange basis whose dof ``a`` sits at ``local_nodes[a]`` (``_basis_tables``,
    ``elementbatch.py:509-525``)."""
    idx = np.rint(np.asarray(local_nodes) * order).astype(np.int64)
    V = [None] * 3
    D = [None] * 3
    for d in range(3):
        V[d], D[d] = _lagrange(order, points[:, d])
    vx, vy, vz = (V[d][:, idx[:, d]] for d in range(3))
    dx, dy, dz = (D[d][:, idx[:, d]] for d in range(3))
    N = vx * vy * vz
    G = np.stack([dx * vy * vz, vx * dy * vz, vx * vy * dz], axis=2)
    return np.ascontiguousarray(N), np.ascontiguousarray(G)


def vertex_basis(points):
    """Trilinear basis of the element map (vertex ``v = vx + 2 vy + 4 vz``) at ``points``."""
    nodes = np.array([(v & 1, (v >> 1) & 1, v >> 2) for v in range(8)], dtype=float)
    return basis_tables(1, np.asarray(points, dtype=np.float64), nodes)


def _dev(a, dtype):
    dev = _lib.require_device()
    return torch.as_tensor(np.ascontiguousarray(a)).to(device=dev, dtype=dtype)


class ElementGroup:
    """The hexahedra of a box mesh with one quadrature rule (``ElementGroup``,
    ``elementbatch.py:159-230``).  ``w``/``points`` on the host; the device copies of the
    vertex coordinates, element vertices, trilinear-basis gradients and weights are made on
    first use and kept."""

    def __init__(self, mesh, quadrature_degree, elems=None):
        self.mesh = mesh
        self.degree = int(quadrature_degree)
        self.points, self.w = integration_rule(self.degree)
        self.nq = self.points.shape[0]
        self.nq1 = int(self.degree) // 2 + 1
        self.points1, _ = gauss_legendre(self.nq1)
        #: the mesh elements of the group (ascending), all of them by default
        self.elems = np.arange(mesh.ne) if elems is None else np.asarray(elems, dtype=np.int64)
        self.whole = elems is None or (self.elems.size == mesh.ne and
                                       np.array_equal(self.elems, np.arange(mesh.ne)))
        self.ne = int(self.elems.size)
        _, self.Gx = vertex_basis(self.points)
        self._dev = None

    def device(self):
        if self._dev is None:
            ev = self.mesh.elem_verts if self.whole else self.mesh.elem_verts[self.elems]
            self._dev = dict(
                verts=_dev(self.mesh.verts, torch.float64),
                elem_verts=_dev(ev, torch.int32),
                tab_x=_dev(self.Gx, torch.float64),
                wq=_dev(self.w, torch.float64))
        return self._dev

    def __repr__(self):
        return "ElementGroup(ne=%d, nq=%d)" % (self.ne, self.nq)


class SpaceTables:
    """Basis tables and element dof map of one space on one group (``SpaceTables``,
    ``elementbatch.py:413-507``): ``N``, ``G``, ``edofs``; the device copies of the table
    ``[N, G]`` (fp64) and of ``edofs`` (int32) are made on first use."""

    def __init__(self, space, group):
        self.space, self.group = space, group
        self.nd = space.nd
        self.nd_total = space.nd
        self.vdim = 1
        self.N, self.G = basis_tables(space.order, group.points, space.local_nodes)
        self.edofs = space.edofs if getattr(group, "whole", True) else \
            np.ascontiguousarray(space.edofs[group.elems])
        # tensor structure: the 1-D basis at the 1-D points and the local dof of every
        # tensor index a1 + n1 (a2 + n1 a3)
        self.n1 = space.order + 1
        V, D = _lagrange(space.order, group.points1)                 # (nq1, n1)
        self.tab1 = np.ascontiguousarray(np.stack([V.T, D.T]))        # (2, n1, nq1)
        idx = np.rint(np.asarray(space.local_nodes) * space.order).astype(np.int64)
        tensor = idx[:, 0] + self.n1 * (idx[:, 1] + self.n1 * idx[:, 2])
        self.lex = np.empty(self.nd, np.int32)
        self.lex[tensor] = np.arange(self.nd, dtype=np.int32)
        self.signs = np.ones((1, 1))
        self._dev = None

    def device(self):
        if self._dev is None:
            tab = np.concatenate([self.N.reshape(-1), self.G.reshape(-1)])
            self._dev = dict(tab=_dev(tab, torch.float64),
                             edofs=_dev(self.edofs, torch.int32),
                             tab1=_dev(self.tab1, torch.float64),
                             lex=_dev(self.lex, torch.int32))
        return self._dev


class MeshBatches:
    """The element groups of a mesh (``MeshBatches``, ``elementbatch.py:600-698``): one group
    of every hexahedron by default, or ``groups`` contiguous ranges of elements (the
    reference's groups by geometry or attribute; here all hexahedra, one quadrature rule).
    A kernel takes one group: ``subset(g)`` is the batches of group ``g`` alone."""

    def __init__(self, mesh, quadrature_degree, groups=1):
        self.mesh = mesh
        if int(groups) == 1:
            self.groups = [ElementGroup(mesh, quadrature_degree)]
        else:
            self.groups = [ElementGroup(mesh, quadrature_degree, elems=r)
                           for r in np.array_split(np.arange(mesh.ne), int(groups))]
        self._tables = {}

    def subset(self, g):
        """The batches of group ``g`` alone (same mesh, group and tables)."""
        sub = MeshBatches.__new__(MeshBatches)
        sub.mesh, sub.groups, sub._tables = self.mesh, [self.groups[g]], {}
        return sub

    def tables(self, space):
        key = id(space)
        got = self._tables.get(key)
        if got is None:
            got = self._tables[key] = [SpaceTables(space, g) for g in self.groups]
        return got


_TABLES = {}


def group_tables(space, group):
    """Cached :class:`SpaceTables` per (space, group) (``group_tables``,
    ``elementbatch.py:568-582``)."""
    key = (id(space), id(group))
    got = _TABLES.get(key)
    if got is None or got.space is not space or got.group is not group:
        got = _TABLES[key] = SpaceTables(space, group)
    return got
