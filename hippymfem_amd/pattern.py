"""Sparsity patterns and the scatter of element arrays into them
(``hippymfem/fem/pattern.py``).

:class:`ScatterPattern` is the value-independent part of a matrix assembly: the CSR graph
(``indptr``, ``indices``; rows of a square pattern lead with their diagonal, the order hypre
keeps), the element rows of every CSR row (``rptr``/``runs``) and every entry's place in its
row (``pos``), from the native library's pattern build (on the GPU when one is visible, else
on the host's threads; both give the same arrays bit for bit).  On the device it keeps one
scatter plan per elimination: the rows, the runs and a one-byte code per entry (two where a
row is longer than 254 entries).  :class:`VectorPattern` is the dual vectors' counterpart.
"""

import numpy as np
import torch

from . import _lib
from .elementbatch import group_tables

_HYPRE_INT = np.int32


def _threads():
    import os
    try:
        return max(len(os.sched_getaffinity(0)), 1)
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def build_tables(tdofs, cdofs, nrow, ncol, diag_first, nthreads=0):
    """``(indptr int64, rptr int64, runs int32, pos uint16, nnz, max_row)`` of the
    sort-free build (``patternbuild.build_tables``, ``patternbuild.py:209-253``), in the
    native library."""
    L = _lib.lib()
    tdofs = np.ascontiguousarray(tdofs, dtype=np.int32)
    cdofs = np.ascontiguousarray(cdofs, dtype=np.int32)
    ne, nd_t = tdofs.shape
    nd_c = cdofs.shape[1]
    if cdofs.shape[0] != ne:
        raise ValueError("the two dof tables hold %d and %d elements" % (ne, cdofs.shape[0]))
    rptr = np.zeros(nrow + 1, np.int64)
    runs = np.empty(max(ne * nd_t, 1), np.int32)
    pos = np.empty(max(ne * nd_t * nd_c, 1), np.uint16)
    indptr = np.zeros(nrow + 1, np.int64)
    nnz = _lib.I64(0)
    max_row = _lib.I64(0)
    _lib.check(L.hm_pattern_build(_lib.ptr(tdofs), _lib.ptr(cdofs), ne, nd_t, nd_c, nrow, ncol,
                                  int(bool(diag_first)), int(nthreads or _threads()),
                                  _lib.ptr(rptr), _lib.ptr(runs), _lib.ptr(pos),
                                  _lib.ptr(indptr), _lib.ctypes.byref(nnz),
                                  _lib.ctypes.byref(max_row)), "pattern build")
    return indptr, rptr, runs[:ne * nd_t], pos[:ne * nd_t * nd_c], int(nnz.value), int(max_row.value)


def build_tables_device(tdofs, cdofs, nrow, ncol, diag_first, stream=None):
    """The same build on the device (``hm_pattern_build_device``, then
    ``hm_pattern_place_device``): a dict of device tensors ``indptr``, ``rptr``, ``runs``
    (int32), ``pos`` (uint16 words held as int16), ``indices`` (int32), and ``nnz``,
    ``max_row``, ``repeats``.  ``tdofs``/``cdofs`` are (ne, nd) int32 arrays, host or device."""
    dev = _lib.require_device()
    td = torch.as_tensor(np.ascontiguousarray(tdofs, dtype=np.int32)
                         if isinstance(tdofs, np.ndarray) else tdofs).to(dev, torch.int32)
    cd = torch.as_tensor(np.ascontiguousarray(cdofs, dtype=np.int32)
                         if isinstance(cdofs, np.ndarray) else cdofs).to(dev, torch.int32)
    td, cd = td.contiguous(), cd.contiguous()
    if td.dim() != 2 or cd.dim() != 2 or td.shape[0] != cd.shape[0]:
        raise ValueError("dof tables of shapes %s and %s" % (tuple(td.shape), tuple(cd.shape)))
    ne, nd_t = td.shape
    nd_c = cd.shape[1]
    L = _lib.lib()
    rptr = torch.empty(nrow + 1, dtype=torch.int32, device=dev)
    indptr = torch.empty(nrow + 1, dtype=torch.int32, device=dev)
    runs = torch.empty(max(ne * nd_t, 1), dtype=torch.int32, device=dev)
    pos = torch.empty(max(ne * nd_t * nd_c, 1), dtype=torch.int16, device=dev)
    work = torch.empty(int(L.hm_pattern_work_bytes(nrow)), dtype=torch.uint8, device=dev)
    info = (_lib.I64 * 3)()
    st = _lib.stream_handle(stream)
    _lib.check(L.hm_pattern_build_device(_lib.ptr(td), _lib.ptr(cd), ne, nd_t, nd_c, nrow, ncol,
                                         int(bool(diag_first)), _lib.ptr(rptr), _lib.ptr(runs),
                                         _lib.ptr(pos), _lib.ptr(indptr), _lib.ptr(work), info,
                                         st), "device pattern build")
    nnz = int(info[0])
    indices = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    _lib.check(L.hm_pattern_place_device(_lib.ptr(td), _lib.ptr(cd), ne, nd_t, nd_c, nrow,
                                         _lib.ptr(indptr), _lib.ptr(pos), _lib.ptr(indices), st),
               "device pattern place")
    return dict(indptr=indptr, rptr=rptr, runs=runs[:ne * nd_t], pos=pos[:ne * nd_t * nd_c],
                indices=indices[:nnz], nnz=nnz, max_row=int(info[1]), repeats=bool(info[2]))


class ScatterPattern:
    """CSR graph of the ``test x trial`` element-assembly map (``ScatterPattern``,
    ``pattern.py:375-1090``).

    ``indptr``/``indices`` int32 (hypre's ``HYPRE_Int``), ``nnz``, ``nent`` (element-matrix
    entries); ``slot`` (int32, the CSR place of entry ``k = (e nd_t + a) nd_c + b``) is made
    when first asked for.  Raises ``ValueError`` for a row of 2**16 entries or more, and
    ``NotImplementedError`` for an element row that repeats a column.

    ``build="device"`` builds on the GPU (:func:`build_tables_device`, the same arrays bit for
    bit) and keeps the device arrays for the scatter plans; the host copies serve the host
    paths.  ``"host"`` builds on the host's threads.  ``"auto"`` (the default) builds on the
    GPU when one is visible, and on the host where the device build does not take the
    pattern (rows of more than 2048 entries) or there is no GPU."""

    def __init__(self, test_tables, trial_tables, test_space=None, trial_space=None,
                 nthreads=0, build="auto"):
        if len(test_tables) != len(trial_tables) or not test_tables:
            raise ValueError("test and trial tables must cover the same groups")
        tt, tr = test_tables[0], trial_tables[0]
        test_space = test_space if test_space is not None else tt.space
        trial_space = trial_space if trial_space is not None else tr.space
        self.nrow, self.ncol = int(test_space.ndof), int(trial_space.ndof)
        # groups: their entries in group order (entry k of group g at its offset), the
        # elements of all groups as one run of elements
        self.shapes = [(int(a.group.ne), int(a.nd_total), int(b.nd_total))
                       for a, b in zip(test_tables, trial_tables)]
        if len({(a, b) for _, a, b in self.shapes}) != 1:
            raise NotImplementedError("element groups of different dof counts (element "
                                      "geometries) in one pattern")
        self.sizes = [ne * a * b for ne, a, b in self.shapes]
        self.nent = int(sum(self.sizes))
        self.ne = int(sum(ne for ne, _, _ in self.shapes))
        _, self.nd_t, self.nd_c = self.shapes[0]
        self.group_start = np.concatenate([[0], np.cumsum([ne for ne, _, _ in self.shapes])])
        self.sign = None
        self.perm = None
        self.tdofs = np.ascontiguousarray(np.concatenate([np.asarray(t.edofs).reshape(
            -1, self.nd_t) for t in test_tables]), dtype=np.int32)
        self.cdofs = np.ascontiguousarray(np.concatenate([np.asarray(t.edofs).reshape(
            -1, self.nd_c) for t in trial_tables]), dtype=np.int32)
        self.diag_first = self.nrow == self.ncol
        if max(self.nrow, self.ncol) > (1 << 31) - 1:
            raise ValueError("%d rows and %d columns do not fit 32-bit local indices"
                             % (self.nrow, self.ncol))
        self._plans = {}
        if build == "auto":
            build = "device" if torch.cuda.is_available() else "host"
            if build == "device":
                try:
                    repeats = self._build_device()
                except NotImplementedError:
                    self._plans = {}
                    build = "host"
            if build == "host":
                repeats = self._build_host(nthreads)
        elif build == "device":
            repeats = self._build_device()
        elif build == "host":
            repeats = self._build_host(nthreads)
        else:
            raise ValueError("build is 'auto', 'host' or 'device', not %r" % (build,))
        self.build = build
        if repeats:
            raise NotImplementedError("an element row repeats a column; the device scatter "
                                      "adds one element row's entries at once")
        self._slot = None
        self._masked = {}

    def _build_host(self, nthreads):
        (indptr, rptr, runs, pos, self.nnz,
         self.max_row) = build_tables(self.tdofs, self.cdofs, self.nrow, self.ncol,
                                      self.diag_first, nthreads)
        if self.nnz > np.iinfo(np.int32).max:
            raise ValueError("the sparsity pattern has %d entries, beyond int32" % self.nnz)
        indices = np.empty(self.nnz, _HYPRE_INT)
        L = _lib.lib()
        _lib.check(L.hm_pattern_place(_lib.ptr(self.cdofs), self.ne, self.nd_t, self.nd_c,
                                      self.nrow, _lib.ptr(rptr), _lib.ptr(runs),
                                      _lib.ptr(indptr), _lib.ptr(pos), _threads(),
                                      _lib.ptr(indices), None), "pattern place")
        self._h = dict(indptr=indptr.astype(np.int32), indptr64=indptr, rptr=rptr, runs=runs,
                       pos=pos, indices=indices)
        self._d = {}
        rep = _lib.I32(0)
        _lib.check(L.hm_pattern_repeats(self.ne, self.nd_t, self.nd_c, _lib.ptr(pos),
                                        _threads(), _lib.ctypes.byref(rep)), "pattern check")
        return bool(rep.value)

    def _build_device(self):
        d = build_tables_device(self.tdofs, self.cdofs, self.nrow, self.ncol, self.diag_first)
        self.nnz, self.max_row = d["nnz"], d["max_row"]
        self._d = {k: d[k] for k in ("indptr", "rptr", "runs", "pos", "indices")}
        self._h = {}
        return d["repeats"]

    # ------------------------------------------------------- the arrays, either side
    def dev(self, name):
        """A device tensor of the pattern: ``indptr``, ``rptr``, ``runs``, ``indices`` (int32)
        or ``pos`` (uint16 words held as int16).  A device build keeps its own; a host build's
        are copied once."""
        got = self._d.get(name)
        if got is None:
            h = self._h[name]
            if name == "rptr":
                h = h.astype(np.int32)
            elif name == "pos":
                h = h.view(np.int16)
            got = self._d[name] = torch.as_tensor(np.ascontiguousarray(h)).to(
                _lib.require_device())
        return got

    def _host(self, name):
        got = self._h.get(name)
        if got is None:
            if name == "indptr64":
                got = self._host("indptr").astype(np.int64)
            else:
                got = self._d[name].cpu().numpy()
                if name == "rptr":
                    got = got.astype(np.int64)
                elif name == "pos":
                    got = got.view(np.uint16)
            self._h[name] = got
        return got

    #: host arrays (numpy), copied from the device build when first asked for
    indptr = property(lambda self: self._host("indptr"))
    indices = property(lambda self: self._host("indices"))
    rptr = property(lambda self: self._host("rptr"))
    runs = property(lambda self: self._host("runs"))
    pos = property(lambda self: self._host("pos"))
    _indptr64 = property(lambda self: self._host("indptr64"))

    def dev_cdofs(self):
        """The trial dof table ``(ne, nd_c)`` int32 on the device."""
        got = self._d.get("cdofs")
        if got is None:
            got = self._d["cdofs"] = torch.as_tensor(self.cdofs).to(_lib.require_device())
        return got

    # ---------------------------------------------------------------- host arrays
    @property
    def slot(self):
        """CSR place of every element-matrix entry (``pattern.py:437-446``), int32."""
        if self._slot is None:
            slot = np.empty(self.nent, np.int32)
            _lib.check(_lib.lib().hm_pattern_place(
                _lib.ptr(self.cdofs), self.ne, self.nd_t, self.nd_c, self.nrow,
                _lib.ptr(self.rptr), _lib.ptr(self.runs), _lib.ptr(self._indptr64),
                _lib.ptr(self.pos), _threads(), _lib.ptr(self.indices), _lib.ptr(slot)),
                "pattern slots")
            self._slot = slot
        return self._slot

    def masked_slots(self, row_mask, col_mask):
        """CSR slots in masked rows or columns (``masked_slots``, ``pattern.py:693-713``)."""
        key = (id(row_mask), id(col_mask))
        got = self._masked.get(key)
        if got is not None and got[0] is row_mask and got[1] is col_mask:
            return got[2]
        if self.nnz == 0:
            kill = np.zeros(0, dtype=np.intp)
        else:
            hit = (np.repeat(row_mask, np.diff(self._indptr64)) if row_mask is not None
                   else np.zeros(self.nnz, dtype=bool))
            if col_mask is not None:
                hit |= col_mask[self.indices]
            kill = np.flatnonzero(hit).astype(np.intp)
        self._masked[key] = (row_mask, col_mask, kill)
        return kill

    # ------------------------------------------------------------- device plans
    def device_plan(self, row_mask=None, col_mask=None, diag=None):
        """The device scatter plan with the elimination folded in: rows in ``row_mask``
        come out zero (with ``diag`` on their diagonal when it is given, which needs a
        square pattern), entries in columns of ``col_mask`` are dropped.  Cached per
        (masks, diagonal)."""
        key = (id(row_mask), id(col_mask), diag)
        got = self._plans.get(key)
        if got is not None and got["row_mask"] is row_mask and got["col_mask"] is col_mask:
            return got
        dev = _lib.require_device()
        base = dict(indptr=self.dev("indptr"), rptr=self.dev("rptr"), runs=self.dev("runs"))
        wide = self.max_row > 254
        codes_key = ("codes", id(col_mask))
        codes = self._plans.get(codes_key)
        if codes is None or codes[0] is not col_mask:
            # on the device: the place codes (uint8 for rows below 255 entries, else the
            # uint16 words as int16), the dropped columns all ones, in run order (the codes
            # of the run at position p at p nd_c, a row's together)
            pos = self.dev("pos")
            c = pos.clone() if wide else pos.to(torch.uint8)
            if col_mask is not None:
                cm = torch.as_tensor(np.asarray(col_mask, dtype=bool)).to(dev)
                cm = cm[self.dev_cdofs().long()]                                 # (ne, nd_c)
                c = c.view(self.ne, self.nd_t, self.nd_c)
                c.masked_fill_(cm[:, None, :].expand(self.ne, self.nd_t, self.nd_c),
                               -1 if wide else 255)
            c = c.view(-1, self.nd_c)[base["runs"].long()].reshape(-1).contiguous()
            codes = self._plans[codes_key] = (col_mask, c)
        rowmode = None
        if row_mask is not None:
            rm = torch.as_tensor(np.asarray(row_mask, dtype=bool)).to(dev)
            if diag is not None:
                if not self.diag_first:
                    raise ValueError("a diagonal is written on square patterns only")
                rows = torch.nonzero(rm).reshape(-1)
                first = base["indptr"].long()[rows]
                bad = int((self.dev("indices")[first].long() != rows).sum()) if rows.numel() else 0
                if bad:
                    raise RuntimeError("%d of %d eliminated rows have no diagonal entry"
                                       % (bad, rows.numel()))
            rowmode = rm.to(torch.int8) * (2 if diag is not None else 1)
        head = self._row_headers(base, rowmode) if self.nd_c > 8 else None
        plan = _lib.HmScatterPlan(
            nrow=self.nrow, nnz=self.nnz, ne=self.ne, nd_t=self.nd_t, nd_c=self.nd_c,
            code_bytes=2 if wide else 1, max_row=self.max_row,
            indptr=_lib.ptr(base["indptr"]), rptr=_lib.ptr(base["rptr"]),
            runs=_lib.ptr(base["runs"]), codes=_lib.ptr(codes[1]),
            rowmode=_lib.ptr(rowmode), diag=0.0 if diag is None else float(diag),
            rowhead=_lib.ptr(head))
        got = dict(plan=plan, keep=(base, codes, rowmode, head), row_mask=row_mask,
                   col_mask=col_mask)
        self._plans[key] = got
        return got

    def _row_headers(self, base, rowmode):
        """Every row's scatter tables packed in one 48-byte record (``hm_scatter_plan``'s
        ``rowhead``): offset; length, mode and run count in one word; first run position; the
        first eight run ids."""
        ip = base["indptr"].long()
        rp = base["rptr"].long()
        runs = base["runs"]
        nrun = rp[1:] - rp[:-1]
        if self.max_row >= 1 << 16 or int(nrun.max()) >= 1 << 7:
            return None
        head = torch.full((self.nrow, _lib.ROWHEAD), -1, dtype=torch.int32, device=ip.device)
        mode = rowmode.long() if rowmode is not None else torch.zeros_like(nrun)
        head[:, 0] = ip[:-1].int()
        head[:, 1] = ((ip[1:] - ip[:-1]) | (mode << 16) | (nrun << 24)).int()
        head[:, 2] = rp[:-1].int()
        for k in range(8):
            has = nrun > k
            idx = torch.where(has, rp[:-1] + k, torch.zeros_like(rp[:-1]))
            head[:, 3 + k] = torch.where(has, runs[idx], torch.full_like(runs[idx], -1))
        return head

    # ------------------------------------------------------------------ values
    def data(self, element_matrices, zero_slots=None, host=False, dtype=None, plan=None,
             out=None, stream=None, row_defect=None):
        """CSR ``data`` for one batch of element matrices (``ScatterPattern.data``,
        ``pattern.py:749-791``): the entries of every slot summed in ascending entry order
        on the device, then ``zero_slots`` cleared.  ``plan`` is a folded plan of
        :meth:`device_plan` (then ``zero_slots`` is not needed).  ``dtype``: the values'
        precision (that of the element matrices by default).  ``row_defect`` (a device
        fp64 array of ``nrow``): also what the rounding of each row's FP64 sums to ``dtype``
        takes from the row's sum (``RowDefect``, ``singlesolve.py:682-700``)."""
        if len(element_matrices) != len(self.shapes):
            raise ValueError("expected %d element arrays, got %d"
                             % (len(self.shapes), len(element_matrices)))
        # the groups' arrays as one run of elements (one scatter: every sum in entry order,
        # rounded once)
        E = _as_device(element_matrices[0]) if len(element_matrices) == 1 else torch.cat(
            [_as_device(a).reshape(-1) for a in element_matrices])
        if E.numel() != self.nent:
            raise ValueError("element arrays hold %d entries, pattern expects %d"
                             % (E.numel(), self.nent))
        dtype = E.dtype if dtype is None else dtype
        if out is None:
            out = torch.empty(self.nnz, dtype=dtype, device=E.device)
        p = plan if plan is not None else self.device_plan()
        if row_defect is not None:
            if (row_defect.dtype != torch.float64 or row_defect.numel() != self.nrow
                    or row_defect.device != out.device):
                raise ValueError("row_defect must be a device float64 array of %d" % self.nrow)
            _lib.check(_lib.lib().hm_scatter_matrix_defect(
                _lib.ctypes.byref(p["plan"]), _lib.ptr(E), _dtype_code(E.dtype), 0, self.ne,
                _lib.ptr(out), _dtype_code(out.dtype), _lib.ptr(row_defect),
                _lib.stream_handle(stream)), "scatter")
        else:
            _lib.check(_lib.lib().hm_scatter_matrix(
                _lib.ctypes.byref(p["plan"]), _lib.ptr(E), _dtype_code(E.dtype), 0, self.ne,
                _lib.ptr(out), _dtype_code(out.dtype), 0, _lib.stream_handle(stream)),
                "scatter")
        if zero_slots is not None and len(zero_slots):
            out.index_fill_(0, _index_tensor(zero_slots, out.device), 0.0)
        return out.cpu().numpy() if host else out

    def fused_begin(self, dtype=torch.float64):
        """A zeroed accumulator and the maps a chunked scatter feeds
        (``fused_begin``, ``pattern.py:843-862``)."""
        dev = _lib.require_device()
        return torch.zeros(self.nnz, dtype=dtype, device=dev), {}

    def fused_add(self, acc, maps, g, a, bnd, arr, plan=None, stream=None):
        """Add the element matrices of elements ``[a, bnd)`` of group ``g``, in order
        (``fused_add``, ``pattern.py:864-886``)."""
        off = int(self.group_start[g])
        a, bnd = a + off, bnd + off
        E = _as_device(arr)
        if E.numel() != (bnd - a) * self.nd_t * self.nd_c:
            raise ValueError("chunk [%d, %d) of group %d holds %d entries, pattern expects %d"
                             % (a, bnd, g, E.numel(), (bnd - a) * self.nd_t * self.nd_c))
        p = plan if plan is not None else self.device_plan()
        _lib.check(_lib.lib().hm_scatter_matrix(
            _lib.ctypes.byref(p["plan"]), _lib.ptr(E), _dtype_code(E.dtype), int(a), int(bnd),
            _lib.ptr(acc), _dtype_code(acc.dtype), 1, _lib.stream_handle(stream)), "fused scatter")
        return acc

    def fused_end(self, acc, zero_slots=None, host=False):
        if zero_slots is not None and len(zero_slots):
            acc.index_fill_(0, _index_tensor(zero_slots, acc.device), 0.0)
        return acc.cpu().numpy() if host else acc

    def data_fused(self, chunks, zero_slots=None, host=False, dtype=torch.float64, plan=None):
        acc, maps = self.fused_begin(dtype)
        for g, a, bnd, arr in chunks:
            acc = self.fused_add(acc, maps, g, a, bnd, arr, plan=plan)
        return self.fused_end(acc, zero_slots, host=host)

    def __repr__(self):
        return "ScatterPattern(%dx%d, nnz=%d, entries=%d)" % (self.nrow, self.ncol, self.nnz,
                                                            self.nent)


class VectorPattern:
    """The dof of every element-vector entry, grouped by dof (``VectorPattern``,
    ``pattern.py:1156-1233``)."""

    def __init__(self, tables, space=None):
        t = tables[0]
        self.n = int((space if space is not None else t.space).ndof)
        if len({int(x.nd_total) for x in tables}) != 1:
            raise NotImplementedError("element groups of different dof counts in one pattern")
        self.nd = int(t.nd_total)
        self.ne = int(sum(int(x.group.ne) for x in tables))
        self.ngroups = len(tables)
        self.edofs = np.ascontiguousarray(np.concatenate(
            [np.asarray(x.edofs).reshape(-1, self.nd) for x in tables]), dtype=np.int32)
        zero = np.zeros((self.ne, 1), np.int32)
        _, rptr, runs, _, _, _ = build_tables(self.edofs, zero, self.n, 1, False)
        self.rptr, self.runs = rptr, runs
        self._dev = None

    def device(self):
        if self._dev is None:
            dev = _lib.require_device()
            self._dev = (torch.as_tensor(self.rptr.astype(np.int32)).to(dev),
                         torch.as_tensor(self.runs).to(dev))
        return self._dev

    def scatter(self, element_vectors, host=False, zero_rows=None, out=None, accumulate=False,
                elements=None, stream=None):
        """The local dof array of summed element vectors, in ascending entry order
        (``VectorPattern.scatter``, ``pattern.py:1186-1209``); ``zero_rows`` (int8, one a
        dof) clears the essential entries (``zero_essential``, ``bcs.py:80-102``)."""
        V = _as_device(element_vectors[0]) if len(element_vectors) == 1 else torch.cat(
            [_as_device(v).reshape(-1) for v in element_vectors])
        a, bnd = (0, self.ne) if elements is None else elements
        if V.numel() != (bnd - a) * self.nd:
            raise ValueError("element vectors hold %d entries, pattern expects %d"
                             % (V.numel(), (bnd - a) * self.nd))
        if V.dtype != torch.float64:
            V = V.double()
        rptr, runs = self.device()
        if out is None:
            out = torch.zeros(self.n, dtype=torch.float64, device=V.device)
        _lib.check(_lib.lib().hm_scatter_vector(
            self.n, self.nd, _lib.ptr(rptr), _lib.ptr(runs), _lib.ptr(V), int(a), int(bnd),
            _lib.ptr(zero_rows), _lib.ptr(out), int(bool(accumulate)),
            _lib.stream_handle(stream)), "vector scatter")
        return out.cpu().numpy() if host else out


def _as_device(a):
    if isinstance(a, torch.Tensor):
        if not a.is_cuda:
            a = a.to(_lib.require_device())
        return a.contiguous()
    return torch.as_tensor(np.ascontiguousarray(a)).to(_lib.require_device())


def _dtype_code(dt):
    if dt == torch.float64:
        return _lib.FP64
    if dt == torch.float32:
        return _lib.FP32
    raise ValueError("element arrays must be float64 or float32, not %s" % dt)


def _index_tensor(idx, dev):
    if isinstance(idx, torch.Tensor):
        return idx.to(dev, torch.int64)
    return torch.as_tensor(np.asarray(idx, dtype=np.int64)).to(dev)


_PATTERNS = {}
_VECTORS = {}


def get_pattern(test_space, trial_space, groups):
    """Cached :class:`ScatterPattern` per (test space, trial space, groups)
    (``get_pattern``, ``pattern.py:1094-1127``)."""
    key = (id(test_space), id(trial_space)) + tuple(id(g) for g in groups)
    got = _PATTERNS.get(key)
    if got is not None and got[0] is test_space and got[1] is trial_space:
        return got[2]
    pat = ScatterPattern([group_tables(test_space, g) for g in groups],
                         [group_tables(trial_space, g) for g in groups],
                         test_space=test_space, trial_space=trial_space)
    _PATTERNS[key] = (test_space, trial_space, pat, tuple(groups))
    return pat


def get_vector_pattern(space, groups):
    key = (id(space),) + tuple(id(g) for g in groups)
    got = _VECTORS.get(key)
    if got is not None and got[0] is space:
        return got[1]
    pat = VectorPattern([group_tables(space, g) for g in groups], space=space)
    _VECTORS[key] = (space, pat, tuple(groups))
    return pat


def clear_pattern_cache():
    _PATTERNS.clear()
    _VECTORS.clear()
