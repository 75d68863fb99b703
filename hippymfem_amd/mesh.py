"""Hexahedral box meshes and H1 spaces on them.

The reference takes its mesh, dof numbering and element dof tables from MFEM
(``Mesh::MakeCartesian3D``, ``ParFiniteElementSpace``; ``benchmarks/bench_assembly.py:
109-121``, ``hippymfem/fem/elementbatch.py:528-565``).  MFEM is not part of this path; the
meshes of the assembly benchmarks (the unit cube in ``n^3`` hexahedra, optionally with its
interior vertices moved) and their H1 spaces are built here, and every kernel takes the
element dof tables as plain arrays, so any other numbering is handled the same way.

Numberings:

* ``"lattice"`` (default): the dofs of an order-``p`` space are the points of the
  ``(p n + 1)^3`` lattice, x fastest; the local order in an element is the tensor order
  ``a = ax + (p+1) (ay + (p+1) az)``.
* ``"mfem"``: MFEM's order as far as it can be stated without MFEM (vertex dofs first in
  vertex order, then edge, face and interior dofs by first appearance over the elements;
  local order vertices, edges, faces, interior of MFEM's reference cube).
* ``"shuffled"``: the lattice numbering under a seeded permutation.
"""

import numpy as np

#: MFEM's reference cube: vertices, edges and faces (as vertex lists).
_CUBE_V = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                    (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], dtype=float)
_CUBE_E = ((0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7),
           (0, 4), (1, 5), (2, 6), (3, 7))
_CUBE_F = ((3, 2, 1, 0), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7), (4, 5, 6, 7))


class BoxMesh:
    """The unit cube in ``n^3`` hexahedra (``MakeCartesian3D``'s domain).

    Elements ``e = ex + n (ey + n ez)``; vertices on the ``(n+1)^3`` lattice; element vertex
    ``v = vx + 2 vy + 4 vz``.  ``perturb > 0`` moves every interior vertex by a seeded
    ``U[-perturb h, perturb h]^3`` (``numpy.random.default_rng(seed)``), which makes the
    element maps trilinear rather than affine."""

    def __init__(self, n, perturb=0.0, seed=4):
        n = int(n)
        if n < 1:
            raise ValueError("a box mesh needs n >= 1, got %d" % n)
        self.n, self.h, self.perturb = n, 1.0 / n, float(perturb)
        m = n + 1
        lat = np.arange(m ** 3)
        i, j, k = lat % m, (lat // m) % m, lat // (m * m)
        X = np.stack([i / n, j / n, k / n], axis=1).astype(np.float64)
        if perturb:
            shift = np.random.default_rng(seed).uniform(-perturb * self.h, perturb * self.h,
                                                        size=X.shape)
            inside = (i > 0) & (i < n) & (j > 0) & (j < n) & (k > 0) & (k < n)
            X = X + shift * inside[:, None]
        self.verts = X
        e = np.arange(n ** 3)
        self.elem_ijk = np.stack([e % n, (e // n) % n, e // (n * n)], axis=1)
        corner = np.array([(v & 1, (v >> 1) & 1, v >> 2) for v in range(8)])
        vi = self.elem_ijk[:, None, :] + corner[None, :, :]
        self.elem_verts = (vi[..., 0] + m * (vi[..., 1] + m * vi[..., 2])).astype(np.int64)
        self.ne = n ** 3
        self.nvert = m ** 3

    def __repr__(self):
        return "BoxMesh(n=%d, perturb=%g)" % (self.n, self.perturb)


def _tensor_nodes(order):
    p1 = order + 1
    a = np.arange(p1 ** 3)
    return np.stack([a % p1, (a // p1) % p1, a // (p1 * p1)], axis=1) / float(order)


def _mfem_nodes(order):
    if order == 1:
        return _CUBE_V.copy()
    if order != 2:
        raise NotImplementedError("the mfem numbering covers orders 1 and 2")
    V = _CUBE_V
    return np.array(list(V) + [0.5 * (V[a] + V[b]) for a, b in _CUBE_E]
                    + [V[list(f)].mean(axis=0) for f in _CUBE_F] + [np.full(3, 0.5)])


class H1Space:
    """An H1 space of order 1 or 2 on a :class:`BoxMesh` (the part of
    ``FunctionSpace.H1``, ``hippymfem/fem/spaces.py:58``, that assembly reads).

    ``edofs (ne, nd)`` int64 is the element dof table, ``local_nodes (nd, 3)`` the
    reference position of each local dof (which fixes the basis tables), ``ndof`` the
    number of dofs.  ``lattice_to_dof`` maps the lattice point of a dof to its number."""

    def __init__(self, mesh, order, numbering="lattice", seed=11):
        order = int(order)
        if order not in (1, 2):
            raise NotImplementedError("H1 orders 1 and 2 are supported, not %d" % order)
        self.mesh, self.order, self.numbering = mesh, order, numbering
        self.local_nodes = _mfem_nodes(order) if numbering == "mfem" else _tensor_nodes(order)
        self.nd = self.local_nodes.shape[0]
        L = order * mesh.n + 1
        self.ndof = L ** 3
        off = np.rint(self.local_nodes * order).astype(np.int64)
        pt = mesh.elem_ijk[:, None, :] * order + off[None, :, :]
        lat = pt[..., 0] + L * (pt[..., 1] + L * pt[..., 2])
        if numbering == "lattice":
            to_dof = np.arange(self.ndof, dtype=np.int64)
        elif numbering == "shuffled":
            to_dof = np.random.default_rng(seed).permutation(self.ndof).astype(np.int64)
        elif numbering == "mfem":
            to_dof = self._first_appearance(lat, off, order, L)
        else:
            raise ValueError("unknown numbering %r" % (numbering,))
        self.lattice_to_dof = to_dof
        self.edofs = np.ascontiguousarray(to_dof[lat])
        self.vdim = 1
        self._coords = None

    def _first_appearance(self, lat, off, order, L):
        to_dof = np.full(self.ndof, -1, np.int64)
        idx = np.arange(self.ndof)
        i, j, k = idx % L, (idx // L) % L, idx // (L * L)
        vert = idx[(i % order == 0) & (j % order == 0) & (k % order == 0)]
        to_dof[vert] = np.arange(vert.size)
        nxt = vert.size
        kind = np.count_nonzero(off % order != 0, axis=1)
        for want in (1, 2, 3):                          # edges, faces, interiors
            seq = lat[:, kind == want].reshape(-1)
            if seq.size == 0:
                continue
            _, first = np.unique(seq, return_index=True)
            fresh = seq[np.sort(first)]
            to_dof[fresh] = nxt + np.arange(fresh.size)
            nxt += fresh.size
        return to_dof

    @property
    def nd_total(self):
        return self.nd

    def dof_coordinates(self):
        """Physical coordinates of the dofs ``(ndof, 3)``: the reference nodes through each
        element's trilinear map."""
        if self._coords is None:
            from .elementbatch import vertex_basis
            Nv, _ = vertex_basis(self.local_nodes)
            X = np.einsum("pv,evd->epd", Nv, self.mesh.verts[self.mesh.elem_verts])
            out = np.empty((self.ndof, 3))
            out[self.edofs.reshape(-1)] = X.reshape(-1, 3)
            self._coords = out
        return self._coords

    def from_lattice(self, values):
        """A dof vector from values in lattice order."""
        out = np.empty(self.ndof)
        out[self.lattice_to_dof] = np.asarray(values, dtype=np.float64)
        return out

    def vector(self):
        return np.zeros(self.ndof)

    def __repr__(self):
        return "H1Space(order=%d, ndof=%d, %s)" % (self.order, self.ndof, self.numbering)


def FunctionSpace(mesh, order, numbering="lattice"):
    """``FunctionSpace.H1(mesh, order)`` of the reference, for box meshes."""
    return H1Space(mesh, order, numbering)
