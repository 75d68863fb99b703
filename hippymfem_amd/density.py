"""Caller-supplied densities (``QuadratureKernel(density, ...)``, ``kernel.py:1917-1937``).

The reference takes any JAX-traceable density ``r(u, m, p, x, *params)`` of :class:`Field`
values and gradients and differentiates it at every quadrature point (``_point_blocks``,
``kernel.py:1581-1646``).  Here a caller density is a torch function of the fields at the
points; the library evaluates the fields on the device (``hm_point_fields``), torch takes the
density's pointwise first and second derivatives by automatic differentiation, and the
library contracts them with the basis (``hm_element_matrices_pointwise``,
``hm_element_vectors_pointwise``), with the geometry (w |det J|, J^-1) in FP64 as for the
closed-form densities.

A slot's pointwise variables are its value and physical gradient, ``(v, dv/dx_0..2)``: the
four rows (or columns) of the 4 x 4 block ``D[r][c] = d2 r / d(slot i var r) d(slot j var c)``.
"""

import numpy as np
import torch

from . import _lib
from ._lib import ADJOINT, PARAMETER, STATE
from .elementbatch import vertex_basis

#: elements a chunk of the pointwise derivatives (the second derivatives take 12 x 12
#: doubles a point)
POINT_CHUNK = 8192


class PointwiseDensity:
    """``fn(u, gu, m, gm, p, gp, x) -> r``: values ``(...)``, gradients and coordinates
    ``(..., 3)``, torch operations only (differentiated by ``torch.func``).  ``f`` is unused
    (the closed-form densities' source); ``is_fwd_linear``: the residual is linear in the
    state (as the reference's ``PDEVariationalProblem`` is told)."""

    pointwise = True
    code = 0
    f = 0.0

    def __init__(self, fn, is_fwd_linear=False, name=None):
        self.fn = fn
        self.is_fwd_linear = bool(is_fwd_linear)
        self.nonlinear = not self.is_fwd_linear
        self.name = name or getattr(fn, "__name__", "density")

    def __repr__(self):
        return "PointwiseDensity(%s)" % self.name

    # ---------------------------------------------------------------- at points
    def _point(self, v, x):
        return self.fn(v[0], v[1:4], v[4], v[5:8], v[8], v[9:12], x)

    def values(self, v, x):
        """``r`` at points: ``v (n, 12)``, ``x (n, 3)``."""
        return self.fn(v[:, 0], v[:, 1:4], v[:, 4], v[:, 5:8], v[:, 8], v[:, 9:12], x)

    def first(self, v, x, i):
        """``d r / d(slot i value, gradient)`` at points ``(n, 4)``."""
        g = torch.func.vmap(torch.func.grad(self._point))(v, x)
        return g[:, 4 * i:4 * i + 4].contiguous()

    def second(self, v, x, i, j):
        """``d2 r / d(slot i var) d(slot j var)`` at points ``(n, 4, 4)``."""
        h = torch.func.vmap(torch.func.hessian(self._point))(v, x)
        return h[:, 4 * i:4 * i + 4, 4 * j:4 * j + 4].contiguous()


def point_fields(kernel, local_arrays, a, b):
    """The fields at the points of elements ``[a, b)``: ``w |det J|`` ``(b-a, nq)``, the slot
    variables ``v (b-a, nq, 12)`` (u, grad u, m, grad m, p, grad p) and the coordinates
    ``x (b-a, nq, 3)`` (``hm_point_fields`` and the element map)."""
    P, keep = kernel.problem(local_arrays)
    nq = kernel.group.nq
    F = torch.empty((b - a, nq, 13), dtype=torch.float64, device=keep[0].device)
    _lib.check(_lib.lib().hm_point_fields(_lib.ctypes.byref(P), int(a), int(b), _lib.ptr(F),
                                          _lib.stream_handle(None)), "point fields")
    g = kernel.group.device()
    Nx = getattr(kernel.group, "_Nx_dev", None)
    if Nx is None:
        Nx, _ = vertex_basis(kernel.group.points)
        Nx = kernel.group._Nx_dev = torch.as_tensor(Nx).to(keep[0].device)
    X = g["verts"][g["elem_verts"][a:b].long()]                        # (b-a, 8, 3)
    x = torch.einsum("qv,evd->eqd", Nx, X)
    return F[..., 0], F[..., 1:], x


def _chunks(a, b):
    return [(s, min(s + POINT_CHUNK, b)) for s in range(a, b, POINT_CHUNK)]


def element_matrices(kernel, i, j, local_arrays, a, b, out, precision="fp64", nonfinite=None):
    """Block ``(i, j)`` element matrices of elements ``[a, b)`` into ``out``."""
    dens = kernel.density
    code = _lib.PRECISIONS[precision]
    if code not in (_lib.FP64, _lib.FP32):
        raise NotImplementedError("a caller density takes FP64 or FP32 element matrices")
    for s, t in _chunks(a, b):
        _, v, x = point_fields(kernel, local_arrays, s, t)
        nq = v.shape[1]
        D = dens.second(v.reshape(-1, 12), x.reshape(-1, 3), i, j).reshape(t - s, nq, 16)
        P, keep = kernel.problem(local_arrays, nonfinite)
        o = out[s - a:t - a]
        _lib.check(_lib.lib().hm_element_matrices_pointwise(
            _lib.ctypes.byref(P), int(i), int(j), _lib.ptr(D.contiguous()), int(s), int(t), code,
            _lib.ptr(o), _lib.stream_handle(None)), "element matrices (%d, %d), caller density"
            % (i, j))
    return out


def element_vectors(kernel, i, local_arrays, a, b, out):
    dens = kernel.density
    for s, t in _chunks(a, b):
        _, v, x = point_fields(kernel, local_arrays, s, t)
        nq = v.shape[1]
        V = dens.first(v.reshape(-1, 12), x.reshape(-1, 3), i).reshape(t - s, nq, 4)
        P, keep = kernel.problem(local_arrays)
        _lib.check(_lib.lib().hm_element_vectors_pointwise(
            _lib.ctypes.byref(P), int(i), _lib.ptr(V.contiguous()), int(s), int(t),
            _lib.ptr(out[s - a:t - a]), _lib.stream_handle(None)),
            "element vectors (%d), caller density" % i)
    return out


def element_values(kernel, local_arrays, a, b, out):
    dens = kernel.density
    for s, t in _chunks(a, b):
        w, v, x = point_fields(kernel, local_arrays, s, t)
        nq = v.shape[1]
        r = dens.values(v.reshape(-1, 12), x.reshape(-1, 3)).reshape(t - s, nq)
        out[s - a:t - a] = torch.sum(w * r, dim=1)
    return out
