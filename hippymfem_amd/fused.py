"""Fused assembly of a block on a box mesh: the plan of ``hm_assemble_fused``.

The fused kernel computes the element matrices and sums them into CSR values in one pass, so
no element matrix goes to memory (the two-pass route, ``QuadratureKernel.element_matrices``
then ``ScatterPattern.data``, writes and reads them back).  Its plan is value-independent and
made once per block and elimination, on the device:

* the rows in tile order: lattice planes of the test space from bottom to top, in each the
  tiles of ``bx x by`` element columns (x fastest), in each tile its lattice points (x
  fastest); for every row its CSR offset and row mode (0 assemble, 1 zero row, 2 zero row with
  the policy's diagonal);
* for every CSR entry, in the same order, its place in the row's dense stencil (the box of
  trial-lattice points the row couples to, x fastest), bit 7 set where the elimination drops
  its column (``pattern.py:693-713``);
* the start of every (plane, tile) in both.

The values are those of the two-pass route bit for bit: every entry has the element-matrix
kernel's bits and every CSR value is summed in ascending element order (``np.bincount``'s
order, ``pattern.py:788``).
"""

import numpy as np
import torch

from . import _lib

#: where the fused route applies: a box mesh with lattice-numbered spaces (any dof numbering of
#: the lattice points), tensor-product elements and the 3- or 4-point Gauss rule
PRECISIONS = ("fp64", "fp32")


def supported(kernel, i, j, precision):
    """Whether block ``(i, j)`` of ``kernel`` can be assembled by the fused route; ``None``
    when it can, else the reason."""
    mesh = kernel.batches.mesh
    if kernel.pointwise:
        return "a caller-supplied density"
    if precision not in PRECISIONS:
        return "precision %s" % precision
    if not hasattr(mesh, "elem_ijk") or not hasattr(mesh, "n"):
        return "not a box mesh"
    if not getattr(kernel.group, "whole", True):
        return "an element group that is not the whole box"
    if kernel.group.nq1 ** 3 != kernel.group.nq or kernel.group.nq1 not in (3, 4):
        return "no tensor rule of 3 or 4 points a direction"
    for s in (kernel.spaces[i], kernel.spaces[j]):
        if not hasattr(s, "lattice_to_dof") or s.order not in (1, 2):
            return "a space without a lattice numbering"
    if kernel.spaces[i].order < kernel.spaces[j].order:
        return "a test order below the trial order"
    return None


def tile_of(kernel, i, j, precision):
    """``(bx, by)``: the tile the kernel takes for this block (``hm_fused_tile``), or None."""
    P, _ = kernel.problem([_zeros(kernel, s) for s in range(3)])
    bx, by = _lib.I32(0), _lib.I32(0)
    rc = _lib.lib().hm_fused_tile(_lib.ctypes.byref(P), int(i), int(j),
                                  _lib.PRECISIONS[precision], _lib.ctypes.byref(bx),
                                  _lib.ctypes.byref(by))
    if rc == _lib.HM_ERR_UNSUPPORTED:
        return None
    _lib.check(rc, "fused tile")
    return int(bx.value), int(by.value)


def blocks_of(kernel, i, j, precision, aligned):
    """Blocks of the fused kernel that share a compute unit for this block on a mesh of the
    kind ``aligned`` (``hm_fused_blocks``): 2 where the FP32 form's tile fits half the unit's
    LDS, else 1."""
    P, _ = kernel.problem([_zeros(kernel, s) for s in range(3)])
    nb = _lib.I32(0)
    rc = _lib.lib().hm_fused_blocks(_lib.ctypes.byref(P), int(i), int(j),
                                    _lib.PRECISIONS[precision],
                                    _lib.FUSED_ALIGNED if aligned else 0, _lib.ctypes.byref(nb))
    if rc == _lib.HM_ERR_UNSUPPORTED:
        return 1
    _lib.check(rc, "fused blocks")
    return int(nb.value)


def _zeros(kernel, s):
    return np.zeros(1)


def _pattern_array(pattern, name, dev):
    """The pattern's ``indptr``/``indices`` on ``dev``: its device arrays on the GPU (no host
    round trip), its host arrays for a plan built on the CPU."""
    if dev.type == "cuda":
        return pattern.dev(name).to(dev)
    return torch.as_tensor(np.asarray(getattr(pattern, name)))


def _widths(L, p, n, pc):
    """Stencil width of every lattice point ``0..L-1`` of order ``p`` on ``n`` elements."""
    l = np.arange(L)
    shared = (l % p == 0) & (l > 0) & (l < p * n)
    return pc * np.where(shared, 2, 1) + 1


def aligned_box(mesh):
    """Whether every element of ``mesh`` is an axis-aligned box: each vertex coordinate equals
    that of the vertex with the same index along its axis (the test the element kernels make
    element by element, which then take J as the diagonal of the edge lengths).  Made once a
    mesh and kept on it."""
    got = getattr(mesh, "_hm_aligned", None)
    if got is None:
        V = np.asarray(mesh.verts)
        ev = np.asarray(mesh.elem_verts)
        got = True
        for lo in range(0, ev.shape[0], 1 << 18):
            blk = ev[lo:lo + (1 << 18)]
            for d, bit in ((0, 1), (1, 2), (2, 4)):
                C = V[blk, d]
                if not np.array_equal(C, C[:, [v & bit for v in range(8)]]):
                    got = False
                    break
            if not got:
                break
        mesh._hm_aligned = got
    return got


def _first(L, p, n):
    """The first element holding each lattice point (the stencil's origin, in elements)."""
    l = np.arange(L)
    f = np.where((l % p == 0) & (l > 0), l // p - 1, l // p)
    return np.minimum(f, n - 1)


class FusedPlan:
    """The plan of one block's fused assembly (``hm_fused_plan``).  ``row_mask`` (bool per
    test dof) and ``diag`` the folded elimination of the rows, ``col_mask`` (bool per trial
    dof) the dropped columns; raises ``NotImplementedError`` where a row of the pattern is not
    its stencil (a mesh that is not the box the spaces describe)."""

    def __init__(self, pattern, test_space, trial_space, mesh, tile, row_mask=None,
                 col_mask=None, diag=None, device=None):
        dev = _lib.require_device() if device is None else torch.device(device)
        self.tile = bx, by = tile
        if int(pattern.nnz) >= 2 ** 31:
            raise NotImplementedError("CSR offsets past 2^31 (the row headers are 32-bit)")
        nx = ny = nz = int(mesh.n)
        self.dims = (nx, ny, nz)
        pt, pc = int(test_space.order), int(trial_space.order)
        L, C = pt * nx + 1, pc * nx + 1
        ntx, nty = -(-nx // bx), -(-ny // by)
        ntile = ntx * nty
        # the rows of one plane in tile order: tile (Y, X), then the tile's points y, x
        lat = np.arange(L)
        tx = np.minimum(lat // (pt * bx), ntx - 1)
        ty = np.minimum(lat // (pt * by), nty - 1)
        lx2, ly2 = np.meshgrid(lat, lat, indexing="xy")               # [ly, lx]
        lx2, ly2 = lx2.reshape(-1), ly2.reshape(-1)
        order2 = np.lexsort((lx2, ly2, tx[lx2], ty[ly2]))
        plane_pts = (lx2 + L * ly2)[order2]                            # lattice index in plane
        # rows a tile, and the first row of every (plane, tile)
        cnt_x = np.bincount(tx, minlength=ntx)
        cnt_y = np.bincount(ty, minlength=nty)
        rows_tile = (cnt_y[:, None] * cnt_x[None, :]).reshape(-1)     # tile = X + ntx Y
        wx = _widths(L, pt, nx, pc)
        wz = _widths(L, pt, nz, pc)
        sx_tile = np.bincount(tx, weights=wx, minlength=ntx).astype(np.int64)
        sy_tile = np.bincount(ty, weights=wx, minlength=nty).astype(np.int64)
        s2d_tile = (sy_tile[:, None] * sx_tile[None, :]).reshape(-1)   # stencil entries / wz
        hdr_off = (np.arange(L)[:, None] * (L * L)
                   + np.concatenate([[0], np.cumsum(rows_tile)[:-1]])[None, :])
        chunk = s2d_tile[None, :] * wz[:, None]                        # (plane, tile) entries
        padded = (chunk + 15) // 16 * 16
        sig_off = np.concatenate([[0], np.cumsum(padded.reshape(-1))[:-1]]).reshape(L, ntile)
        self.sig_bytes = int(padded.sum())
        # every row in tile order, its CSR offset, length and mode
        ptsd = torch.as_tensor(plane_pts, dtype=torch.int64, device=dev)
        zs = torch.arange(L, dtype=torch.int64, device=dev)
        rows_lat = (zs[:, None] * (L * L) + ptsd[None, :]).reshape(-1)
        l2d_t = torch.as_tensor(np.asarray(test_space.lattice_to_dof), dtype=torch.int64,
                                device=dev)
        r = l2d_t[rows_lat]
        indptr = _pattern_array(pattern, "indptr", dev).long()
        lo = indptr[r]
        ln = indptr[r + 1] - lo
        lxr, lyr, lzr = rows_lat % L, (rows_lat // L) % L, rows_lat // (L * L)
        wxt = torch.as_tensor(wx, device=dev)
        wzt = torch.as_tensor(wz, device=dev)
        want = wxt[lxr] * wxt[lyr] * wzt[lzr]
        if not torch.equal(ln, want):
            raise NotImplementedError("the pattern's rows are not the box's stencils")
        mode = torch.zeros_like(r)
        if row_mask is not None:
            rm = torch.as_tensor(np.asarray(row_mask, dtype=np.int64), device=dev)
            mode = rm[r] * (2 if diag is not None else 1)
        self.hdr = torch.stack([lo, mode], dim=1).to(torch.int32).contiguous()
        # every CSR entry in tile order: its stencil place
        total = int(ln.sum())
        if total != pattern.nnz:
            raise NotImplementedError("the tiles do not cover the pattern")
        starts = torch.cumsum(ln, 0) - ln
        idx = torch.repeat_interleave(lo - starts, ln) + torch.arange(total, device=dev)
        cols = _pattern_array(pattern, "indices", dev)[idx].long()
        d2l_c = torch.empty(C ** 3, dtype=torch.int64, device=dev)
        d2l_c[torch.as_tensor(np.asarray(trial_space.lattice_to_dof), dtype=torch.int64,
                              device=dev)] = torch.arange(C ** 3, device=dev)
        cl = d2l_c[cols]
        cx, cy, cz = cl % C, (cl // C) % C, cl // (C * C)
        first = torch.as_tensor(_first(L, pt, nx), device=dev)
        ox = cx - pc * torch.repeat_interleave(first[lxr], ln)
        oy = cy - pc * torch.repeat_interleave(first[lyr], ln)
        oz = cz - pc * torch.repeat_interleave(first[lzr], ln)
        ww = torch.repeat_interleave(wxt[lxr], ln)
        hh = torch.repeat_interleave(wxt[lyr], ln)
        dd = torch.repeat_interleave(wzt[lzr], ln)
        if not (bool((ox >= 0).all()) and bool((ox < ww).all()) and bool((oy >= 0).all())
                and bool((oy < hh).all()) and bool((oz >= 0).all()) and bool((oz < dd).all())):
            raise NotImplementedError("a column outside its row's stencil")
        place = ox + ww * (oy + hh * oz)
        del ox, oy, oz, ww, hh, dd, cx, cy, cz, cl
        if col_mask is not None:
            cm = torch.as_tensor(np.asarray(col_mask, dtype=bool), device=dev)
            place = place + 128 * cm[cols].long()
        del cols
        # destination of every entry: its (plane, tile) chunk, then its place in the chunk
        tile_row = torch.as_tensor(np.repeat(np.arange(ntile), rows_tile), device=dev)
        plane_row = zs.repeat_interleave(len(plane_pts))
        chunk_of_row = plane_row * ntile + tile_row.repeat(L)
        sig_off_t = torch.as_tensor(sig_off.reshape(-1), device=dev)
        # first entry of every row within its chunk: rows are contiguous in tile order
        row_first = starts - torch.as_tensor(
            np.concatenate([[0], np.cumsum(chunk.reshape(-1))[:-1]]), device=dev)[chunk_of_row]
        dst_row = sig_off_t[chunk_of_row] + row_first
        dst = torch.repeat_interleave(dst_row - starts, ln) + torch.arange(total, device=dev)
        sig = torch.zeros(self.sig_bytes + 16, dtype=torch.uint8, device=dev)
        sig[dst] = place.to(torch.uint8)
        self.sigma = sig
        self.hdr_off = torch.as_tensor(hdr_off.reshape(-1), dtype=torch.int64, device=dev)
        self.sig_off = sig_off_t.contiguous()
        self.nnz, self.nrow = int(pattern.nnz), int(r.numel())
        self.chunk_len = chunk                  # (plane, tile) entries before padding
        self.hdr_off_np, self.sig_off_np = hdr_off, sig_off
        self.struct = _lib.HmFusedPlan(
            nx=nx, ny=ny, nz=nz, bx=bx, by=by,
            flags=_lib.FUSED_ALIGNED if aligned_box(mesh) else 0, hdr=_lib.ptr(self.hdr),
            hdr_off=_lib.ptr(self.hdr_off), sigma=_lib.ptr(self.sigma),
            sig_off=_lib.ptr(self.sig_off), diag=0.0 if diag is None else float(diag))

    def assemble(self, kernel, i, j, local_arrays, precision, out, stream=None, nonfinite=None):
        """Block ``(i, j)`` at the fields ``local_arrays`` into ``out`` (every value written)."""
        P, keep = kernel.problem(local_arrays, nonfinite)
        code = _lib.PRECISIONS[precision]
        _lib.check(_lib.lib().hm_assemble_fused(
            _lib.ctypes.byref(P), int(i), int(j), _lib.ctypes.byref(self.struct), code,
            _lib.ptr(out), _lib.FP64 if out.dtype == torch.float64 else _lib.FP32,
            _lib.stream_handle(stream)), "fused assembly (%d, %d)" % (i, j))
        kernel._hold(keep, stream)
        return out
