"""Element kernels on the GPU (``hippymfem/fem/kernel.py``).

The reference builds its element arrays by automatic differentiation of a user density
(JAX); this path serves the densities of the assembly benchmarks, each stated in closed form
inside the HIP kernels:

* :class:`ExpDiffusion` ``r = exp(m) grad u . grad p - f p`` (``nonlinear=False``,
  ``benchmarks/bench_assembly.py:123-125``), and
* ``r = exp(m) (1 + u^2) grad u . grad p - f p`` (``nonlinear=True``).

Slots are ``STATE=0`` (u), ``PARAMETER=1`` (m), ``ADJOINT=2`` (p); ``spaces`` is
``[Vu, Vm, Vu]``.  Arrays come back as torch tensors on the device.
"""

import contextlib

import numpy as np
import torch

from . import _lib
from . import density as _density
from ._lib import ADJOINT, PARAMETER, STATE

#: Largest number of elements handed to one launch when chunks are asked for
#: (``ELEMENT_CHUNK``, ``kernel.py:402``); 0 means the whole batch.
ELEMENT_CHUNK = 0


class ExpDiffusion:
    """The density ``exp(m) g(u) grad u . grad p - f p``, ``g = 1`` or ``1 + u^2``."""

    def __init__(self, nonlinear=False, f=1.0):
        self.nonlinear = bool(nonlinear)
        self.f = float(f)

    @property
    def code(self):
        return 1 if self.nonlinear else 0

    @property
    def is_fwd_linear(self):
        return not self.nonlinear

    def __repr__(self):
        return "ExpDiffusion(nonlinear=%s, f=%g)" % (self.nonlinear, self.f)


class QuadratureKernel:
    """Element arrays of a density on a mesh's element groups (``QuadratureKernel``,
    ``kernel.py:1917-2090``).  ``local_arrays`` is ``[u, m, p]``: dof vectors (numpy or
    device tensors) of the three slots."""

    def __init__(self, density, spaces, batches, nparams=0):
        if len(spaces) != 3 or spaces[STATE] is not spaces[ADJOINT] and \
                spaces[STATE].order != spaces[ADJOINT].order:
            raise ValueError("spaces must be [state, parameter, adjoint], state and adjoint alike")
        self.density = density
        #: a caller-supplied density (``density.PointwiseDensity``): its pointwise derivatives
        #: by automatic differentiation, the contraction by the library
        self.pointwise = bool(getattr(density, "pointwise", False))
        self.spaces = list(spaces)
        self.batches = batches
        if len(batches.groups) != 1:
            raise NotImplementedError("a kernel takes one element group; use one kernel a group "
                                      "(MeshBatches.subset) and assemble their arrays together")
        self.group = batches.groups[0]
        self.tables = [batches.tables(s)[0] for s in spaces]
        self.nd = [t.nd for t in self.tables]
        self._keep = None

    # ------------------------------------------------------------- plumbing
    def problem(self, local_arrays, nonfinite=None):
        """The ``hm_problem`` of the current fields; keeps the device arrays alive.
        ``nonfinite``: an int32 device flag the element-matrix kernels set to 1 when they
        write a value that is not finite."""
        dev = _lib.require_device()
        g = self.group.device()
        tu = self.tables[STATE].device()
        tm = self.tables[PARAMETER].device()
        u, m, p = (_vec(a, dev) for a in local_arrays[:3])
        P = _lib.HmProblem(
            ne=self.group.ne, nq=self.group.nq, nd_u=self.nd[STATE], nd_m=self.nd[PARAMETER],
            density=self.density.code, f=self.density.f,
            verts=_lib.ptr(g["verts"]), elem_verts=_lib.ptr(g["elem_verts"]),
            edofs_u=_lib.ptr(tu["edofs"]), edofs_m=_lib.ptr(tm["edofs"]),
            tab_u=_lib.ptr(tu["tab"]), tab_m=_lib.ptr(tm["tab"]), tab_x=_lib.ptr(g["tab_x"]),
            wq=_lib.ptr(g["wq"]), u=_lib.ptr(u), m=_lib.ptr(m), p=_lib.ptr(p),
            nq1=self.group.nq1 if self.group.nq1 ** 3 == self.group.nq else 0,
            n1_u=self.tables[STATE].n1, n1_m=self.tables[PARAMETER].n1,
            tab1_u=_lib.ptr(tu["tab1"]), tab1_m=_lib.ptr(tm["tab1"]),
            lex_u=_lib.ptr(tu["lex"]), lex_m=_lib.ptr(tm["lex"]),
            nonfinite=None if nonfinite is None else _lib.ptr(nonfinite))
        self._keep = (u, m, p, g, tu, tm)
        return P, (u, m, p)

    @staticmethod
    def _hold(tensors, stream):
        """Keep the temporaries a launch on ``stream`` reads alive until that stream has used
        them: the caching allocator orders a free against the current stream only, so a
        launch on another stream records its use (``Tensor.record_stream``)."""
        if stream is None or stream == torch.cuda.current_stream():
            return
        for t in tensors:
            if isinstance(t, torch.Tensor) and t.is_cuda:
                t.record_stream(stream)

    # ------------------------------------------------------------ element arrays
    def element_matrices(self, i, j, local_arrays, params=(), consistent=None,
                         precision="fp64", elements=None, out=None, stream=None, nonfinite=None):
        """Block ``(i, j)`` element matrices ``[(ne, nd_i, nd_j)]`` (``element_matrices``,
        ``kernel.py:1964-2006``).  ``precision``: ``"fp64"`` (default), or ``"fp32"``,
        ``"bf16"``, ``"fp16"`` for the precision form (fp32 results).  ``consistent``
        (``"symmetric"`` or ``"two-sided"``) with a precision form: the consistent block of
        ``_block_fn`` (``kernel.py:1573-1580``), fp64 matrices exact on the constants.
        ``nonfinite``: an int32 device flag set to 1 when an entry written is not finite (the
        caller clears it)."""
        if consistent and precision != "fp64":
            return [self.consistent_block(i, j, local_arrays, consistent == "symmetric",
                                          precision=precision, elements=elements, out=out,
                                          stream=stream)]
        a, b = (0, self.group.ne) if elements is None else elements
        code = _lib.PRECISIONS[precision]
        dtype = torch.float64 if code == _lib.FP64 else torch.float32
        if self.pointwise:
            if out is None:
                out = _result_array((b - a, self.nd[i], self.nd[j]), dtype,
                                    _lib.require_device())
            with _on(stream):
                _density.element_matrices(self, i, j, local_arrays, a, b, out, precision,
                                          nonfinite)
            return [out]
        P, keep = self.problem(local_arrays, nonfinite)
        if out is None:
            out = _result_array((b - a, self.nd[i], self.nd[j]), dtype, keep[0].device)
        _lib.check(_lib.lib().hm_element_matrices(
            _lib.ctypes.byref(P), int(i), int(j), int(a), int(b), code, _lib.ptr(out),
            _lib.stream_handle(stream)), "element matrices (%d, %d)" % (i, j))
        self._hold(keep, stream)
        return [out]

    def consistent_block(self, i, j, local_arrays, symmetric=False, precision="fp32",
                         elements=None, out=None, stream=None):
        """Block ``(i, j)`` in single precision, exact on the constants
        (``consistent_block``, ``kernel.py:1503-1571``): the precision-form element matrices
        corrected by rank-one terms so that their products with the constants of slot ``j``
        (and, transposed, of slot ``i``) are the FP64 ones, which come from the matrix-free
        element products.  ``symmetric``: symmetrize first; the result is symmetric to the
        last bit.  FP64 ``(ne, nd_i, nd_j)``."""
        a, b = (0, self.group.ne) if elements is None else elements
        K = self.element_matrices(i, j, local_arrays, precision=precision, elements=(a, b),
                                  stream=stream)[0]
        dev = K.device
        ones_j = torch.full((self.spaces[j].ndof,), 1.0 / np.sqrt(self.nd[j]), dtype=torch.float64,
                            device=dev)
        Rj = self.element_hvp(i, j, local_arrays, ones_j, elements=(a, b), stream=stream)[0]
        Ri = None
        if not symmetric:
            ones_i = torch.full((self.spaces[i].ndof,), 1.0 / np.sqrt(self.nd[i]),
                                dtype=torch.float64, device=dev)
            Ri = self.element_hvp(j, i, local_arrays, ones_i, elements=(a, b), stream=stream)[0]
        if out is None:
            out = torch.empty((b - a, self.nd[i], self.nd[j]), dtype=torch.float64, device=dev)
        _lib.check(_lib.lib().hm_consistent_block(
            _lib.ptr(K), _lib.ptr(Rj), _lib.ptr(Ri), b - a, self.nd[i], self.nd[j],
            1 if symmetric else 0, _lib.ptr(out), _lib.stream_handle(stream)), "consistent block")
        self._hold((K, Rj, Ri), stream)
        self._keep = (self._keep, K, Rj, Ri)
        return out

    def element_matrices_pair(self, ij1, ij2, local_arrays, precision="fp64", elements=None,
                              outs=(None, None), stream=None, nonfinite=None):
        """Blocks ``ij1`` and ``ij2`` of one shape from one pass over the elements: the
        geometry, the fields and both blocks' pointwise values at the points once, then the two
        contractions.  Each result has the bits ``element_matrices`` gives it.  ``None``, with
        nothing written, where the two do not share a pass (another shape, a block that is
        zero for the density, the BF16/FP16 forms)."""
        if precision not in ("fp64", "fp32") or self.pointwise:
            return None
        (i1, j1), (i2, j2) = ij1, ij2
        if (self.nd[i1], self.nd[j1]) != (self.nd[i2], self.nd[j2]):
            return None
        P, keep = self.problem(local_arrays, nonfinite)
        a, b = (0, self.group.ne) if elements is None else elements
        code = _lib.PRECISIONS[precision]
        dtype = torch.float64 if code == _lib.FP64 else torch.float32
        o1, o2 = (o if o is not None else _result_array((b - a, self.nd[i1], self.nd[j1]), dtype,
                                                        keep[0].device) for o in outs)
        rc = _lib.lib().hm_element_matrices_pair(
            _lib.ctypes.byref(P), int(i1), int(j1), int(i2), int(j2), int(a), int(b), code,
            _lib.ptr(o1), _lib.ptr(o2), _lib.stream_handle(stream))
        if rc == _lib.HM_ERR_UNSUPPORTED:
            return None
        _lib.check(rc, "element matrices (%d, %d) and (%d, %d)" % (i1, j1, i2, j2))
        self._hold(keep, stream)
        return o1, o2

    def element_matrices_many(self, pairs, local_arrays, precision="fp64"):
        """Several blocks from one call (``element_matrices_many``, ``kernel.py:2048-2068``):
        blocks of one shape two at a time from one pass where the kernel allows, the others
        one by one."""
        pairs = list(dict.fromkeys(tuple(ij) for ij in pairs))
        got = self._many(pairs, local_arrays, precision, None)
        return {ij: [got[ij]] for ij in pairs}

    def element_matrix_chunks(self, i, j, local_arrays, chunk=None, precision="fp64"):
        """A generator of ``(group, start, stop, array)`` chunks (``element_matrix_chunks``,
        ``kernel.py:2008-2024``)."""
        chunk = int(chunk or ELEMENT_CHUNK or self.group.ne)
        for a in range(0, self.group.ne, chunk):
            b = min(a + chunk, self.group.ne)
            yield 0, a, b, self.element_matrices(i, j, local_arrays, precision=precision,
                                                 elements=(a, b))[0]

    def element_matrix_chunks_many(self, pairs, local_arrays, chunk=None, precision="fp64"):
        """A generator of ``(group, start, stop, {(i, j): array})`` chunks of several blocks
        (``element_matrix_chunks_many``, ``kernel.py:2026-2046``), each chunk's blocks as
        ``element_matrices_many`` makes them (two of one shape from one pass), so the scatter
        can take each chunk and nothing larger than a chunk is resident."""
        chunk = int(chunk or ELEMENT_CHUNK or self.group.ne)
        pairs = list(dict.fromkeys(tuple(ij) for ij in pairs))
        for a in range(0, self.group.ne, chunk):
            b = min(a + chunk, self.group.ne)
            got = self._many(pairs, local_arrays, precision, (a, b))
            yield 0, a, b, {ij: got[ij] for ij in pairs}

    def _many(self, pairs, local_arrays, precision, elements):
        out, todo = {}, list(pairs)
        while todo:
            ij = todo.pop(0)
            for k, other in enumerate(todo):
                got = self.element_matrices_pair(ij, other, local_arrays, precision=precision,
                                                 elements=elements)
                if got is not None:
                    out[ij], out[other] = got
                    del todo[k]
                    break
            else:
                out[ij] = self.element_matrices(ij[0], ij[1], local_arrays, precision=precision,
                                                elements=elements)[0]
        return out

    def element_vectors(self, i, local_arrays, params=(), elements=None, out=None, stream=None):
        """``dR_e / d(slot i)`` ``[(ne, nd_i)]`` (``element_vectors``, ``kernel.py:2070-2078``)."""
        a, b = (0, self.group.ne) if elements is None else elements
        if self.pointwise:
            if out is None:
                out = torch.empty((b - a, self.nd[i]), dtype=torch.float64,
                                  device=_lib.require_device())
            with _on(stream):
                _density.element_vectors(self, i, local_arrays, a, b, out)
            return [out]
        P, keep = self.problem(local_arrays)
        if out is None:
            out = torch.empty((b - a, self.nd[i]), dtype=torch.float64, device=keep[0].device)
        _lib.check(_lib.lib().hm_element_vectors(_lib.ctypes.byref(P), int(i), int(a), int(b),
                                                 _lib.ptr(out), _lib.stream_handle(stream)),
                   "element vectors (%d)" % i)
        self._hold(keep, stream)
        return [out]

    def element_values(self, local_arrays, params=(), elements=None, out=None, stream=None):
        """Per-element integrals of the density ``[(ne,)]`` (``element_values``,
        ``kernel.py:2129-2136``; the functional of ``_functional``, ``kernel.py:977-1009``):
        ``sum_q w_q |det J| (exp(m) g(u) grad u . grad p - f p)``, summed in point order.
        ``RuntimeError`` where one is not finite."""
        a, b = (0, self.group.ne) if elements is None else elements
        if self.pointwise:
            if out is None:
                out = torch.empty(b - a, dtype=torch.float64, device=_lib.require_device())
            with _on(stream):
                _density.element_values(self, local_arrays, a, b, out)
            if not bool(torch.isfinite(out).all()):
                raise RuntimeError("element values: a value is not finite")
            return [out]
        flag = torch.zeros(1, dtype=torch.int32, device=_lib.require_device())
        P, keep = self.problem(local_arrays, flag)
        if out is None:
            out = torch.empty(b - a, dtype=torch.float64, device=keep[0].device)
        _lib.check(_lib.lib().hm_element_scalars(_lib.ctypes.byref(P), int(a), int(b),
                                                 _lib.ptr(out), _lib.stream_handle(stream)),
                   "element values")
        self._hold(keep + (flag,), stream)
        if int(flag.item()) != 0:
            raise RuntimeError("element values: a value is not finite")
        return [out]

    def element_hvp(self, i, j, local_arrays, jdir_local, params=(), elements=None, out=None,
                    stream=None):
        """Block ``(i, j)`` applied to the slot-``j`` direction element by element
        ``[(ne, nd_i)]`` (``element_hvp``, ``kernel.py:2080-2090``)."""
        a, b = (0, self.group.ne) if elements is None else elements
        if self.pointwise:
            # the element matrices of the caller density times the direction's element dofs
            E = self.element_matrices(i, j, local_arrays, elements=(a, b), stream=stream)[0]
            d = _vec(jdir_local, E.device)
            ed = self.tables[j].device()["edofs"][a:b].long()
            with _on(stream):
                y = torch.bmm(E, d[ed].unsqueeze(2)).squeeze(2)
                if out is not None:
                    out.copy_(y)
                    y = out
            return [y]
        P, keep = self.problem(local_arrays)
        d = _vec(jdir_local, keep[0].device)
        if out is None:
            out = torch.empty((b - a, self.nd[i]), dtype=torch.float64, device=keep[0].device)
        _lib.check(_lib.lib().hm_element_hvp(_lib.ctypes.byref(P), int(i), int(j), _lib.ptr(d),
                                             int(a), int(b), _lib.ptr(out),
                                             _lib.stream_handle(stream)),
                   "element products (%d, %d)" % (i, j))
        self._hold(keep + (d,), stream)
        return [out]


def _result_array(shape, dtype, dev):
    """The result of a batch, or ``MemoryError`` when it does not fit on the device (as the
    reference's batch planner says when no chunk size fits, ``kernel.py:1084-1101``): the
    kernels hold nothing per element beyond their result, so only a smaller batch helps."""
    try:
        return torch.empty(shape, dtype=dtype, device=dev)
    except torch.cuda.OutOfMemoryError as exc:
        need = int(np.prod(shape)) * torch.empty((), dtype=dtype).element_size()
        free = torch.cuda.mem_get_info(dev)[0] if dev.type == "cuda" else 0
        raise MemoryError(
            "element batch of %d does not fit on %s: its result is %.2f GB against %.2f GB "
            "free. Assemble it in chunks (element_matrix_chunks, or a callable for "
            "assemble_matrix)." % (shape[0], dev, need / 2 ** 30, free / 2 ** 30)) from exc


def _on(stream):
    return torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext()


def _vec(a, dev):
    if isinstance(a, torch.Tensor):
        return a.to(device=dev, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
