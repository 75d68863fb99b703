"""The caller of the assembly path: ``PDEVariationalProblem``'s assembly methods
(``hippymfem/modeling/PDEVariationalProblem.py``).

It keeps, per block, the plan (pattern, folded elimination) and device buffers for the
element matrices and the CSR values, so a repeated assembly (a Newton step, a serving
request, a benchmark step) launches kernels into resident memory and allocates nothing.
"""

import contextlib
import os
import warnings

import numpy as np
import torch

from . import _lib
from ._lib import ADJOINT, PARAMETER, STATE
from . import fused
from .csrassemble import _device_plan, assemble_vector_csr, finish_block, plan_block
from .kernel import QuadratureKernel

NVAR = 3

#: elimination per block (``PDEVariationalProblem.py:1448-1451``, ``:597-598``); a block
#: with the parameter on a side keeps its columns (``_symmetric``, ``:118-124``)
ELIMINATION = {
    (ADJOINT, STATE): dict(test=True, diag_policy="one"),
    (STATE, ADJOINT): dict(test=True, diag_policy="one"),
    (STATE, STATE): dict(test=True, diag_policy="zero"),
    (ADJOINT, PARAMETER): dict(test=True, diag_policy="one"),
    (STATE, PARAMETER): dict(test=True, diag_policy="one"),
    (PARAMETER, PARAMETER): dict(test=False, diag_policy="one"),
}


def require_finite_arrays(arrays, what):
    """Raise ``RuntimeError`` if an element array has a non-finite entry
    (``require_finite_arrays``, ``PDEVariationalProblem.py:1844-1866``): a parameter may
    be finite while ``exp(m)`` overflows inside the density, so only the element arrays
    show it.  Reduced on the device; one flag comes back."""
    ok = True
    for a in arrays:
        if isinstance(a, torch.Tensor):
            ok = ok and bool(torch.isfinite(a).all())
        else:
            ok = ok and bool(np.isfinite(np.asarray(a)).all())
        if not ok:
            break
    if not ok:
        raise RuntimeError("%s produced non-finite element values; the residual density "
                           "overflowed at this point (a line search should reject it)" % what)
    return arrays


def require_finite(v, what):
    """Raise ``RuntimeError`` if the vector ``v`` has a non-finite entry
    (``require_finite``, ``PDEVariationalProblem.py:1868-1890``)."""
    ok = (bool(torch.isfinite(v).all()) if isinstance(v, torch.Tensor)
          else bool(np.isfinite(np.asarray(v)).all()))
    if not ok:
        raise RuntimeError("%s contains non-finite values; the step that produced it is too "
                           "large for this problem (a line search should reject it)" % what)
    return v


#: Bytes an element batch's resident matrices may take on the device (``None``: what the
#: device has free).  A batch above it, or one whose allocation fails, is assembled in
#: chunks (``BlockAssembly``); ``HM_ELEMENT_BUDGET`` in the environment sets it too.
ELEMENT_BUDGET = None

#: How many element batches of this process did not fit and were assembled in chunks (the
#: reference's ``RETRIES``, ``kernel.py:608-612``): 0 where every batch fit.
RETRIES = 0


def _budget():
    if ELEMENT_BUDGET is not None:
        return int(ELEMENT_BUDGET)
    env = os.environ.get("HM_ELEMENT_BUDGET")
    return int(float(env)) if env else None


def _fits(ne, want, free):
    """Elements of a chunk, from the size the whole batch asked for: a third of what is free
    (the failed request is the largest buffer, not the whole working set), at least halved
    (``_fits``, ``kernel.py:629-656``)."""
    if free <= 0 or want <= 0:
        return max(1, ne // 2)
    return max(1, min(ne // 2, int(ne * (0.33 * free) / want)))


def _symmetric(i, j):
    """Whether block (i, j) is eliminated on both sides where they share a space
    (``PDEVariationalProblem._symmetric``)."""
    return PARAMETER not in (i, j)


class BlockAssembly:
    """Resident buffers and plan of one block.  ``assemble`` is one launch of the fused
    kernel (element contraction and scatter in one pass, no element matrix in memory) where
    the fused route applies (``route="auto"``, the default, or ``"fused"``: a box mesh,
    tensor elements, FP64 or FP32), else one element-kernel launch into ``E`` and one scatter
    launch into ``data`` (``route="two-pass"``); both on one stream, with the same bits.

    With ``row_sums``, a square block with single-precision values keeps the defect of its
    row sums beside them (``row_defect``, fp64: what the rounding of each row's FP64 sums
    took from the row's sum), as the reference's single-precision matrices do for their
    solver (``tdofassemble.SINGLE_ROW_SUMS``, ``tdofassemble.py:107-125``).  It costs the
    scatter a reduction and a store a row (about 0.06 ms on A at 64^3), so it is asked for
    rather than made by default here, where no solver reads it."""

    def __init__(self, kernel, i, j, test_ess=None, diag_policy="one", precision="fp64",
                 value_dtype=None, row_sums=False, route="auto"):
        if route not in ("auto", "fused", "two-pass"):
            raise ValueError("route is 'auto', 'fused' or 'two-pass', not %r" % (route,))
        self.kernel, self.i, self.j = kernel, i, j
        self.route = route
        self._fused = None
        Vi, Vj = kernel.spaces[i], kernel.spaces[j]
        self.plan = plan_block(Vi, Vj, kernel.batches.groups, test_ess=test_ess,
                               diag_policy=diag_policy, symmetric=_symmetric(i, j))
        self.pattern = self.plan.pattern
        self._dplan = None
        self.precision = precision
        dev = _lib.require_device()
        e_dtype = torch.float64 if precision == "fp64" else torch.float32
        if value_dtype is None:
            value_dtype = e_dtype
        self.e_dtype = e_dtype
        self._E = None
        self.chunk = None   # elements a chunk when the whole batch does not fit
        self.data = torch.empty(self.pattern.nnz, dtype=value_dtype, device=dev)
        self.row_defect = None
        if row_sums and Vi is Vj and value_dtype != torch.float64:
            self.row_defect = torch.empty(self.pattern.nrow, dtype=torch.float64, device=dev)
        # set by the element kernel when it writes a value that is not finite
        self.nonfinite = torch.zeros(1, dtype=torch.int32, device=dev)

    @property
    def E(self):
        """The resident element matrices ``(ne, nd_i, nd_j)``, made at the first use; where
        the whole batch does not fit (the allocation fails, or it is above
        ``ELEMENT_BUDGET``), a buffer of one chunk (``chunk`` elements) and a warning: the
        block is then assembled chunk by chunk, every chunk's entries added to the values in
        order (the reference's retry in chunks sized from the failed request,
        ``kernel.py:1080-1113``)."""
        if self._E is None:
            global RETRIES
            k = self.kernel
            ne, row = k.group.ne, k.nd[self.i] * k.nd[self.j]
            size = torch.empty((), dtype=self.e_dtype).element_size()
            want = ne * row * size
            budget = _budget()
            try:
                if budget is not None and want > budget:
                    raise torch.cuda.OutOfMemoryError(
                        "allocate %d bytes: above the element budget of %d" % (want, budget))
                self._E = torch.empty((ne, k.nd[self.i], k.nd[self.j]), dtype=self.e_dtype,
                                      device=self.data.device)
            except torch.cuda.OutOfMemoryError:
                RETRIES += 1
                free = budget if budget is not None else torch.cuda.mem_get_info(
                    self.data.device)[0]
                self.chunk = _fits(ne, want, free)
                warnings.warn(
                    "element batch of %d (%.3g GB of element matrices) did not fit on %s; "
                    "assembling block (%d, %d) in chunks of %d. FP64 values are the same bits "
                    "as a whole-batch assembly; single-precision values differ at round-off. "
                    "Set ELEMENT_BUDGET (HM_ELEMENT_BUDGET) to pin the budget."
                    % (ne, want / 2 ** 30, self.data.device, self.i, self.j, self.chunk),
                    RuntimeWarning, stacklevel=2)
                self._E = torch.empty((self.chunk, k.nd[self.i], k.nd[self.j]),
                                      dtype=self.e_dtype, device=self.data.device)
        return self._E

    def _e_fits(self):
        """Whether the whole batch's element matrices fit in half of the free device memory
        (and the element budget)."""
        k = self.kernel
        size = torch.empty((), dtype=self.e_dtype).element_size()
        want = k.group.ne * k.nd[self.i] * k.nd[self.j] * size
        budget = _budget()
        if budget is not None and want > budget:
            return False
        return want <= 0.5 * torch.cuda.mem_get_info(self.data.device)[0]

    def _fp32_two_blocks(self):
        """Whether the fused kernel of this block runs two blocks a compute unit (its FP32 form,
        where the tile's LDS fits half the unit's, in the form of the mesh's kind)."""
        if self.precision != "fp32" or fused.supported(self.kernel, self.i, self.j,
                                                       self.precision) is not None:
            return False
        aligned = fused.aligned_box(self.kernel.batches.mesh)
        return fused.blocks_of(self.kernel, self.i, self.j, self.precision, aligned) == 2

    @property
    def dplan(self):
        """The scatter's device plan (codes in run order, row modes), made at the first
        two-pass assembly: the fused route does not read it."""
        if self._dplan is None:
            self._dplan = _device_plan(self.plan)
        return self._dplan

    @property
    def streamed(self):
        """Whether the two-pass route of this block runs chunk by chunk (its batch did not
        fit)."""
        self.E
        return self.chunk is not None

    def fused_plan(self):
        """The plan of the fused route (``fused.FusedPlan``), made at the first use, or None
        where the block takes the two-pass route."""
        if self._fused is None:
            self._fused = False
            why = None
            if self.route == "two-pass":
                why = "two-pass asked for"
            elif self.row_defect is not None:
                why = "row sums asked for"
            elif self.data.dtype != self.e_dtype:
                why = "values of another precision than the element matrices"
            elif self.plan.diag_policy == "keep" and self.plan.diag is not None:
                why = "diag_policy 'keep'"
            elif self.route == "auto" and self.precision != "fp64" and self._e_fits() and \
                    not self._fp32_two_blocks():
                # single precision: FP32 element matrices halve the two-pass route's traffic;
                # the fused kernel gains on it only where two of its blocks share a compute
                # unit (64^3 P2 A: box 0.55 ms fused against 0.83 ms two-pass, perturbed 0.81
                # against 1.05; with one block, perturbed: 1.10 ms fused)
                why = "single precision with one fused block a compute unit (the two-pass " \
                      "route is faster)"
            elif self.route == "auto" and self.precision == "fp64" and \
                    getattr(self.kernel.density, "nonlinear", False) and \
                    not fused.aligned_box(self.kernel.batches.mesh):
                # a nonlinear density on a mesh of non-aligned elements: the fields' dofs in
                # the element data and the general form's pair slots leave the fused kernel
                # smaller tiles (96^3 perturbed: A 6.05 ms fused against 5.10 ms two-pass,
                # W_uu 5.29 against 4.47; on box meshes the fused route is faster)
                why = "a nonlinear density on a mesh of non-aligned elements (the two-pass " \
                      "route is faster)"
            elif self.route == "auto" and self.kernel.group.nq1 > 3:
                # the 4-point rule: 64 points' pairs in LDS leave the fused kernel small tiles
                # and blocks of 512 threads (64^3 P2 A, degree 6: 2.01 ms fused on 4 x 3 tiles
                # against 1.38 ms two-pass)
                why = "a 4-point rule (the two-pass route is faster)"
            elif self.route == "auto" and self.kernel.spaces[self.j].order < 2:
                # a Q1 trial space: the element matrices are small against the point work
                # the fused kernel repeats for its halo, and the two-pass route takes two such
                # blocks from one element pass (64^3 P2: C 1.04 ms fused against 0.46 ms a
                # block for C and W_um together, W_mm 0.78 against 0.36 ms)
                why = "a Q1 trial space (the two-pass route is faster)"
            else:
                why = fused.supported(self.kernel, self.i, self.j, self.precision)
            tile = None
            if why is None:
                tile = fused.tile_of(self.kernel, self.i, self.j, self.precision)
                if tile is None:
                    why = "no tile of the fused kernel for this block"
            if why is None:
                p = self.plan
                try:
                    self._fused = fused.FusedPlan(p.pattern, p.test_space, p.trial_space,
                                                  self.kernel.batches.mesh, tile,
                                                  row_mask=p.row_mask, col_mask=p.col_mask,
                                                  diag=p.diag)
                except NotImplementedError as exc:
                    why = str(exc)
            self.why_two_pass = why
            if why is not None and self.route == "fused":
                raise NotImplementedError("block (%d, %d) cannot take the fused route: %s"
                                          % (self.i, self.j, why))
        return self._fused or None

    def fresh_values(self):
        """An uninitialized values array (and row defect) of this block's shape, for a result
        that owns its values."""
        d = torch.empty_like(self.data)
        rd = None if self.row_defect is None else torch.empty_like(self.row_defect)
        return d, rd

    def clear_flag(self, stream=None):
        with torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext():
            self.nonfinite.zero_()

    def element_matrices(self, local_arrays, stream=None):
        if self.streamed:
            raise MemoryError("block (%d, %d): the element batch does not fit; it is assembled "
                              "in chunks of %d (assemble_values)" % (self.i, self.j, self.chunk))
        self.clear_flag(stream)
        return self.kernel.element_matrices(self.i, self.j, local_arrays, precision=self.precision,
                                            out=self.E, stream=stream, nonfinite=self.nonfinite)[0]

    def chunks(self):
        """The ``[a, b)`` element ranges of a streamed assembly."""
        ne = self.kernel.group.ne
        return [(a, min(a + self.chunk, ne)) for a in range(0, ne, self.chunk)]

    def _assemble_streamed(self, local_arrays, stream, out, row_defect):
        """Chunk after chunk: its element matrices into the chunk buffer, then added to the
        values (the scatter's rows take the chunk's entries after the ones already summed, so
        the FP64 sums are those of the whole batch)."""
        self.clear_flag(stream)
        if out is None:
            out, row_defect = self.data, self.row_defect
        if row_defect is not None:
            raise NotImplementedError("row sums of a block assembled in chunks")
        k = self.kernel
        p = self.dplan
        for n, (a, b) in enumerate(self.chunks()):
            E = self.E[:b - a]
            k.element_matrices(self.i, self.j, local_arrays, precision=self.precision,
                               elements=(a, b), out=E, stream=stream, nonfinite=self.nonfinite)
            _lib.check(_lib.lib().hm_scatter_matrix(
                _lib.ctypes.byref(p["plan"]), _lib.ptr(E), _lib.FP64 if E.dtype == torch.float64
                else _lib.FP32, int(a), int(b), _lib.ptr(out),
                _lib.FP64 if out.dtype == torch.float64 else _lib.FP32, 0 if n == 0 else 1,
                _lib.stream_handle(stream)), "scatter of elements [%d, %d)" % (a, b))
        return out

    def finite(self):
        """Whether the last element pass wrote finite values only (one int32 read back)."""
        return not bool(self.nonfinite.item())

    def scatter(self, stream=None, out=None, row_defect=None):
        """The resident element matrices into CSR values: into the resident ``data`` (and
        ``row_defect``), or into ``out`` (and ``row_defect``) when given."""
        if out is None:
            out, row_defect = self.data, self.row_defect
        return self.pattern.data([self.E], plan=self.dplan, out=out, stream=stream,
                                 row_defect=row_defect)

    def finish(self, data=None, row_defect=None):
        """The matrix over ``data`` (the resident values by default)."""
        if data is None:
            data, row_defect = self.data, self.row_defect
        A = finish_block(self.plan, data)
        A.row_defect = row_defect
        return A

    def assemble_values(self, local_arrays, stream=None, out=None, row_defect=None):
        """The block's CSR values at the fields ``local_arrays``, into the resident ``data``
        or into ``out``: the fused kernel where it applies, else the element kernel then the
        scatter.  The non-finite flag is set when an element-matrix entry is not finite."""
        F = self.fused_plan()
        if F is not None:
            self.clear_flag(stream)
            return F.assemble(self.kernel, self.i, self.j, local_arrays, self.precision,
                              self.data if out is None else out, stream=stream,
                              nonfinite=self.nonfinite)
        if self.streamed:
            return self._assemble_streamed(local_arrays, stream, out, row_defect)
        self.element_matrices(local_arrays, stream)
        return self.scatter(stream, out=out, row_defect=row_defect)

    def assemble(self, local_arrays, stream=None, out=None, row_defect=None):
        """The block at the fields ``local_arrays``.  The matrix aliases the resident values
        unless ``out`` (and ``row_defect``) is given: a repeated assembly into resident
        memory allocates nothing, and the next assembly of this block overwrites it."""
        self.assemble_values(local_arrays, stream, out=out, row_defect=row_defect)
        return self.finish(out, row_defect) if out is not None else self.finish()

    def check_finite(self, local_arrays, what):
        """``RuntimeError`` where the last pass wrote a non-finite element-matrix entry
        (``require_finite_arrays``); the element matrices are made for the check when the
        fused route left none."""
        if self.finite():
            return
        if self.streamed:
            for a, b in self.chunks():
                E = self.E[:b - a]
                self.kernel.element_matrices(self.i, self.j, local_arrays,
                                             precision=self.precision, elements=(a, b), out=E)
                require_finite_arrays([E], what)
            return
        if self.fused_plan() is not None:
            self.element_matrices(local_arrays)
        require_finite_arrays([self.E], what)


class PDEVariationalProblem:
    """Assembly methods of the reference's forward problem for a density on ``Vh =
    [Vu, Vm, Vu]`` with essential conditions ``bc0`` (``PDEVariationalProblem``,
    ``PDEVariationalProblem.py:127-197``)."""

    def __init__(self, Vh, density, batches, bc0, is_fwd_linear=None, route="auto"):
        """``route``: how the blocks are assembled, ``"auto"`` (the fused kernel where it
        applies), ``"fused"`` or ``"two-pass"`` (``BlockAssembly``)."""
        self.route = route
        self.Vh = list(Vh)
        self.density = density
        self.batches = batches
        self.bc0 = bc0
        self.is_fwd_linear = density.is_fwd_linear if is_fwd_linear is None else is_fwd_linear
        self.kernel = QuadratureKernel(density, self.Vh, batches)
        self.nelem = batches.groups[0].ne
        self._blocks = {}

    def _locals(self, x):
        return [x[STATE], x[PARAMETER], x[ADJOINT] if len(x) > ADJOINT and x[ADJOINT] is not None
                else np.zeros(self.Vh[ADJOINT].ndof)]

    def block_assembly(self, i, j, precision="fp64", value_dtype=None, row_sums=False,
                       route=None):
        route = self.route if route is None else route
        key = (i, j, precision, value_dtype, route) + (("row_sums",) if row_sums else ())
        got = self._blocks.get(key)
        if got is None:
            rule = ELIMINATION[(i, j)]
            got = self._blocks[key] = BlockAssembly(
                self.kernel, i, j, test_ess=self.bc0.ess_tdof if rule["test"] else None,
                diag_policy=rule["diag_policy"], precision=precision, value_dtype=value_dtype,
                row_sums=row_sums, route=route)
        return got

    # ------------------------------------------------------------------ assembly
    def residual(self, x, var=ADJOINT, ess=None):
        """``dR/d(var)`` at ``x`` (``_residual``, ``PDEVariationalProblem.py:423-446``)."""
        vecs = self.kernel.element_vectors(var, self._locals(x))
        require_finite_arrays(vecs, "the residual")
        return assemble_vector_csr(self.Vh[var], self.batches.groups, vecs, self.nelem, ess=ess)

    def block(self, i, j, x, precision="fp64"):
        """Block ``(i, j)`` at ``x`` with its elimination (``_block``/``_domain_block``,
        ``PDEVariationalProblem.py:448-525``); ``RuntimeError`` when an element matrix is
        not finite (the check of ``_domain_block``, ``:495-525``).  The matrix owns its
        values: a later assembly at another point leaves it as it is, as the reference's
        matrices are (a line search or a finite-difference check keeps the linearization
        point's blocks, ``PDEVariationalProblem.py:562-568``)."""
        blk = self.block_assembly(i, j, precision)
        out, rd = blk.fresh_values()
        loc = self._locals(x)
        A = blk.assemble(loc, out=out, row_defect=rd)
        blk.check_finite(loc, "block (%d, %d)" % (i, j))
        return A

    def blocks(self, pairs, x, precision="fp64"):
        """Blocks ``pairs`` at ``x``, each as ``block`` gives it, with two blocks of one shape
        from one pass over the elements wherever the kernel allows (the reference takes every
        linearization point's blocks from one differentiation pass, ``element_matrices_many``,
        ``kernel.py:2048-2068``)."""
        loc = self._locals(x)
        todo = list(dict.fromkeys(tuple(ij) for ij in pairs))
        passes = []
        results = {}
        # the blocks the fused route takes (one launch each, no element matrix in memory) and
        # the blocks whose element batch does not fit (a chunked pass each)
        for ij in list(todo):
            b = self.block_assembly(ij[0], ij[1], precision)
            if b.fused_plan() is not None or b.streamed:
                out, rd = b.fresh_values()
                b.assemble_values(loc, out=out, row_defect=rd)
                results[ij] = (out, rd)
                passes.append((b, [ij]))
                todo.remove(ij)
        while todo:
            ij = todo.pop(0)
            blk = self.block_assembly(ij[0], ij[1], precision)
            blk.clear_flag()
            done = [ij]
            for k, other in enumerate(todo):
                ob = self.block_assembly(other[0], other[1], precision)
                if self.kernel.element_matrices_pair(ij, other, loc, precision=precision,
                                                     outs=(blk.E, ob.E),
                                                     nonfinite=blk.nonfinite) is not None:
                    done.append(todo.pop(k))
                    break
            else:
                blk.element_matrices(loc)
            for d in done:
                b = self.block_assembly(d[0], d[1], precision)
                out, rd = b.fresh_values()
                b.scatter(out=out, row_defect=rd)
                results[d] = (out, rd)
            passes.append((blk, done))
        # every pass launched before the first read of a flag (one flag a pass, the first
        # block's); a set flag is traced to its block
        for blk, done in passes:
            if not blk.finite():
                if blk.fused_plan() is not None or blk.streamed:
                    blk.check_finite(loc, "block (%d, %d)" % done[0])
                    continue
                for d in done:
                    require_finite_arrays([self.block_assembly(d[0], d[1], precision).E],
                                          "block (%d, %d)" % d)
        # every matrix owns its values (``block``)
        return {d: self.block_assembly(d[0], d[1], precision).finish(*results[d])
                for _, done in passes for d in done}

    def jacobian(self, x, precision="fp64"):
        """The forward Jacobian ``A`` (``_jacobian``, ``PDEVariationalProblem.py:542-659``):
        the parameter (and, for a nonlinear residual, the state) must be finite
        (``:555-557``)."""
        require_finite(x[PARAMETER], "the parameter")
        if not self.is_fwd_linear:
            require_finite(x[STATE], "the state")
        return self.block(ADJOINT, STATE, x, precision)

    def set_linearization_point(self, x, gauss_newton_approx=False, precision="fp64"):
        """``A`` and the Hessian blocks C, W_uu (nonlinear state), W_um, W_mm
        (``setLinearizationPoint``, ``PDEVariationalProblem.py:1372-1482``)."""
        require_finite(x[PARAMETER], "the parameter")
        if not self.is_fwd_linear:
            require_finite(x[STATE], "the state")
        need = [(ADJOINT, STATE), (ADJOINT, PARAMETER)]
        if not gauss_newton_approx:
            if not self.is_fwd_linear:
                need.append((STATE, STATE))
            need += [(STATE, PARAMETER), (PARAMETER, PARAMETER)]
        return self.blocks(need, x, precision)

    def functional(self, x):
        """The integral of the residual density itself, ``R(u, m, p)`` (``functional``,
        ``PDEVariationalProblem.py:1802-1817``, domain part): the element values summed
        (``assemble_scalar``)."""
        from .assemble import assemble_scalar

        return assemble_scalar(None, self.kernel.element_values(self._locals(x)))

    def apply_ij_at(self, i, j, x, dir, out=None):
        """Block ``(i, j)`` at ``x`` applied to ``dir`` without assembling it
        (``apply_ij_at``, ``PDEVariationalProblem.py:1672-1718``)."""
        ess = self.bc0.ess_tdof
        dvec = np.array(dir, dtype=np.float64, copy=True) if not isinstance(dir, torch.Tensor) \
            else dir.detach().double().clone()
        if j in (STATE, ADJOINT) and len(ess):
            dvec[ess] = 0.0
        vecs = self.kernel.element_hvp(i, j, self._locals(x), dvec)
        res = assemble_vector_csr(self.Vh[i], self.batches.groups, vecs, self.nelem,
                                  ess=ess if i in (STATE, ADJOINT) else None)
        if i in (STATE, ADJOINT) and {i, j} == {STATE, ADJOINT} and len(ess):
            src = dir if isinstance(dir, torch.Tensor) else torch.as_tensor(np.asarray(dir))
            idx = torch.as_tensor(ess).to(res.device)
            res[idx] = src.to(res.device, torch.float64)[idx]
        if out is not None:
            out.copy_(res)
            return out
        return res
