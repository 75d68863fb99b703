"""The fused assembly on the CPU: its plan (``fused.FusedPlan``) and an emulation of the
kernel's indexing and order of additions (``hm_fused.hip``: tiles, halo, chunks of layers,
contraction tasks, stencil places, the sub-steps of the adds, the plane buffers, the staged
row headers and stencil maps), run over the oracle's element matrices.

Checked at every access: the plane-buffer and staged-map indices stay inside the sizes the
launch gives them, no two adds of one sub-step touch one place (so they need no atomics),
every CSR value is written exactly once, and the values equal the oracle's ``np.bincount``
assembly bit for bit (every value summed in ascending element order).
"""

import numpy as np
import pytest

from oracle import fe_oracle as O

from hippymfem_amd import ADJOINT, PARAMETER, STATE
from hippymfem_amd.csrassemble import plan_block
from hippymfem_amd.fused import FusedPlan, _widths
from hippymfem_amd.workloads import Workload

ELIM = {(ADJOINT, STATE): dict(test=True, diag_policy="one", symmetric=None),
        (STATE, STATE): dict(test=True, diag_policy="zero", symmetric=None),
        (ADJOINT, PARAMETER): dict(test=True, diag_policy="one", symmetric=False),
        (STATE, PARAMETER): dict(test=True, diag_policy="one", symmetric=False),
        (PARAMETER, PARAMETER): dict(test=False, diag_policy="one", symmetric=False)}


def sizes(nt, nc, bx, by):
    """The plane-buffer and staged-map sizes of ``fu_sizes`` (hm_fused.hip)."""
    pt, pc = nt - 1, nc - 1
    sx = bx * (2 * pc + 1) + (pt - 1) * bx * (pc + 1) + (pc + 1)
    sy = by * (2 * pc + 1) + (pt - 1) * by * (pc + 1) + (pc + 1)
    s2d = sx * sy
    p0 = (s2d * (2 * pc + 1) + 1) & ~1
    p1 = (s2d * (pc + 1 if pt == 2 else 2 * pc + 1) + 1) & ~1
    sig = (s2d * (2 * pc + 1) + 15) & ~15
    rows_max = (pt * bx + 1) * (pt * by + 1)
    t_max = nc * (bx * by * nt * nt + (bx + by) * nt + 1)
    return p0, p1, sig, rows_max, t_max


def emulate(W, i, j, E_t, plan, tile, lz_chunk):
    """The CSR values of k_fused over element matrices ``E_t`` (ne, nd_t, nd_c) in tensor
    order, with every index checked."""
    Vi, Vj = W.spaces[i], W.spaces[j]
    nt, nc = Vi.order + 1, Vj.order + 1
    pt, pc = nt - 1, nc - 1
    n = W.mesh.n
    nx = ny = nz = n
    bx, by = tile
    ntx, nty = -(-nx // bx), -(-ny // by)
    ntile = ntx * nty
    p0w, p1w, sigw, rows_max, t_max = sizes(nt, nc, bx, by)
    hdr = plan.hdr.numpy()
    sigma = plan.sigma.numpy()
    hdr_off = plan.hdr_off.numpy()
    sig_off = plan.sig_off.numpy()
    diag = plan.struct.diag
    nnz = plan.nnz
    data = np.full(nnz, np.nan)
    written = np.zeros(nnz, np.int32)
    nchunk = -(-nz // lz_chunk)
    for blk in range(ntile * nchunk):
        tile_id, chunk = blk % ntile, blk // ntile
        X, Y = tile_id % ntx, tile_id // ntx
        x0, x1 = X * bx, min(X * bx + bx, nx)
        y0, y1 = Y * by, min(Y * by + by, ny)
        z0, z1 = chunk * lz_chunk, min(chunk * lz_chunk + lz_chunk, nz)
        xs, ys, zs = max(x0 - 1, 0), max(y0 - 1, 0), max(z0 - 1, 0)
        cx, cy = x1 - xs, y1 - ys
        M = cx * cy
        lx_lo, lx_hi = pt * x0, pt * x1 + (1 if x1 == nx else 0)
        ly_lo, ly_hi = pt * y0, pt * y1 + (1 if y1 == ny else 0)
        lz_lo, lz_hi = pt * z0, pt * z1 + (1 if z1 == nz else 0)
        cxl, cyl = lx_hi - lx_lo, ly_hi - ly_lo
        nrows = cxl * cyl
        assert nrows <= rows_max
        WX = _widths(pt * nx + 1, pt, nx, pc)[lx_lo:lx_hi]
        WY = _widths(pt * ny + 1, pt, ny, pc)[ly_lo:ly_hi]
        px = np.concatenate([[0], np.cumsum(WX)[:-1]])
        sxt = int(WX.sum())
        py = np.concatenate([[0], np.cumsum(WY)[:-1]])
        B2D = np.array([[sxt * py[iy] + WY[iy] * px[ix] for ix in range(cxl)]
                        for iy in range(cyl)]).reshape(-1)
        s2d = sxt * int(WY.sum())
        tasks = []
        for m in range(M):
            ex, ey = xs + m % cx, ys + m // cx
            for a2 in range(max(ly_lo - pt * ey, 0), min(ly_hi - pt * ey, nt)):
                for a1 in range(max(lx_lo - pt * ex, 0), min(lx_hi - pt * ex, nt)):
                    for b1 in range(nc):
                        tasks.append((m, a1, b1, a2))
        assert len(tasks) <= t_max
        bufs = [np.zeros(p0w), np.zeros(p1w)]

        def wz_of(lz):
            return int(_widths(pt * nz + 1, pt, nz, pc)[lz])

        for ez in range(zs, z1):
            retire = [pt * ez + a3 for a3 in range(pt) if lz_lo <= pt * ez + a3 < lz_hi]
            top = pt * (ez + 1)
            if ez == z1 - 1 and lz_lo <= top < lz_hi:
                retire.append(top)
            assert len([lz for lz in retire if lz != top]) <= 2      # two staging slots
            staged = []
            for k, lz in enumerate(retire):
                h0 = hdr_off[lz * ntile + tile_id]
                assert h0 + nrows <= hdr.shape[0]
                nbytes = s2d * wz_of(lz)
                assert nbytes <= sigw and (nbytes + 15) // 16 * 16 <= sigw
                s0 = sig_off[lz * ntile + tile_id]
                assert s0 % 16 == 0 and s0 + (nbytes + 15) // 16 * 16 <= sigma.size
                staged.append((hdr[h0:h0 + nrows], sigma[s0:s0 + nbytes]))
            rk = 0
            na = min(nt, 2)                      # phase A: a3 < 2; phase B: a3 = 2
            for a3lo, a3hi in ((0, na), (na, nt)):
                owned = [a3 for a3 in range(a3lo, a3hi) if lz_lo <= pt * ez + a3 < lz_hi]
                if not owned:
                    continue
                for g in range(4):
                    touched = set()
                    for (m, a1, b1, a2) in tasks:
                        grp = (0 if a2 >= 1 else 2) + (0 if a1 >= 1 else 1)
                        if grp != g:
                            continue
                        ex, ey = xs + m % cx, ys + m // cx
                        e = ex + nx * (ey + ny * ez)
                        ix, iy = pt * ex + a1 - lx_lo, pt * ey + a2 - ly_lo
                        assert 0 <= ix < cxl and 0 <= iy < cyl
                        wx, wy = WX[ix], WY[iy]
                        ox = pc * (1 if (a1 == 0 and ex > 0) else 0) + b1
                        oy0 = pc * (1 if (a2 == 0 and ey > 0) else 0)
                        for a3 in owned:
                            lz = pt * ez + a3
                            buf = bufs[lz & 1]
                            wz = wz_of(lz)
                            oz0 = pc * (1 if (a3 == 0 and ez > 0) else 0)
                            row = B2D[iy * cxl + ix] * wz
                            for b2 in range(nc):
                                for b3 in range(nc):
                                    idx = row + ox + wx * ((oy0 + b2) + wy * (oz0 + b3))
                                    assert 0 <= idx < buf.size and idx < s2d * wz
                                    assert (lz & 1, idx) not in touched
                                    touched.add((lz & 1, idx))
                                    a = a1 + nt * (a2 + nt * a3)
                                    b = b1 + nc * (b2 + nc * b3)
                                    buf[idx] = buf[idx] + E_t[e, a, b]
                for a3 in owned:
                    lz = pt * ez + a3
                    if not (a3 < pt or (ez == nz - 1 and z1 == nz)):
                        continue
                    buf = bufs[lz & 1]
                    wz = wz_of(lz)
                    hds, sg = staged[rk]
                    assert retire[rk] == lz
                    rk += 1
                    rowof = np.repeat(np.arange(nrows), WX[np.arange(nrows) % cxl]
                                      * WY[np.arange(nrows) // cxl])
                    assert rowof.size == s2d and nrows <= 256 and s2d <= 2048
                    mul = (65536 + wz - 1) // wz
                    for kk in range(s2d * wz):
                        assert (kk * mul) >> 16 == kk // wz           # the kernel's division
                        rho = int(rowof[kk // wz])
                        rx, ry = rho % cxl, rho // cxl
                        lo, mode = int(hds[rho, 0]), int(hds[rho, 1])
                        ln = WX[rx] * WY[ry] * wz
                        rb = B2D[rho] * wz
                        t = kk - rb
                        assert 0 <= t < ln
                        if True:
                            c = int(sg[rb + t])
                            pos = c & 0x7F
                            assert pos < ln
                            v = buf[rb + pos]
                            buf[rb + pos] = 0.0
                            o = 0.0 if c & 0x80 else v
                            if mode != 0:
                                o = diag if (mode == 2 and t == 0) else 0.0
                            assert 0 <= lo + t < nnz
                            data[lo + t] = o
                            written[lo + t] += 1
            assert rk == len(retire)
        assert not np.any(bufs[0]) and not np.any(bufs[1])   # every buffer cleared again
    assert np.all(written == 1)
    return data


CASES = [
    # (n, tile, layers a chunk, numbering, nonlinear, perturb)
    (4, (2, 2), 4, "lattice", False, 0.0),
    (5, (2, 3), 2, "mfem", True, 0.2),
    (3, (4, 4), 1, "shuffled", False, 0.2),
    (4, (3, 2), 3, "lattice", True, 0.0),
]


@pytest.mark.parametrize("n,tile,lz_chunk,numbering,nonlinear,perturb", CASES)
def test_fused_emulation_matches_bincount(n, tile, lz_chunk, numbering, nonlinear, perturb):
    W = Workload(n, numbering=numbering, nonlinear=nonlinear, perturb=perturb)
    P = O.Problem(n, perturb=perturb, numbering=numbering,
                  density="nonlinear" if nonlinear else "linear")
    blocks = [(ADJOINT, STATE), (ADJOINT, PARAMETER), (STATE, PARAMETER),
              (PARAMETER, PARAMETER)] + ([(STATE, STATE)] if nonlinear else [])
    for ij in blocks:
        i, j = ij
        rule = ELIM[ij]
        p = plan_block(W.spaces[i], W.spaces[j], W.groups,
                       test_ess=W.ess if rule["test"] else None,
                       diag_policy=rule["diag_policy"], symmetric=rule["symmetric"])
        plan = FusedPlan(p.pattern, W.spaces[i], W.spaces[j], W.mesh, tile,
                         row_mask=p.row_mask, col_mask=p.col_mask, diag=p.diag, device="cpu")
        E = P.block(i, j)                                  # local order of the spaces
        lt = W.kernel.tables[i].lex                        # tensor index -> local dof
        lc = W.kernel.tables[j].lex
        E_t = E[:, lt][:, :, lc]
        got = emulate(W, i, j, E_t, plan, tile, lz_chunk)
        ip, ix, want = P.assembled(i, j, E=E)
        assert np.array_equal(np.asarray(p.pattern.indptr), ip)
        assert np.array_equal(np.asarray(p.pattern.indices), ix)
        assert np.array_equal(got, want), (ij, np.abs(got - want).max())


def test_fused_plan_refuses_a_mesh_that_is_not_its_box():
    """A pattern whose rows are not the stencils of the box (here a shuffled element order
    of the dof tables) is refused, so that the route falls back to the two-pass one."""
    W = Workload(3)
    p = plan_block(W.Vu, W.Vu, W.groups, test_ess=W.ess)
    plan = FusedPlan(p.pattern, W.Vu, W.Vu, W.mesh, (2, 2), row_mask=p.row_mask,
                     col_mask=p.col_mask, diag=p.diag, device="cpu")
    assert plan.nnz == p.pattern.nnz

    class Skewed:
        n = 2                                              # a box of the wrong size
    with pytest.raises((NotImplementedError, IndexError)):
        FusedPlan(p.pattern, W.Vu, W.Vu, Skewed(), (2, 2), device="cpu")


def test_aligned_box_flag():
    """``aligned_box``: every element's vertex coordinates those of its axis neighbours (the
    element kernels' per-element test), True on the unit-cube box, False once interior vertices
    move, and kept on the mesh."""
    from hippymfem_amd.fused import aligned_box
    from hippymfem_amd.mesh import BoxMesh
    assert aligned_box(BoxMesh(5))
    moved = BoxMesh(5, perturb=0.2)
    assert not aligned_box(moved)
    assert moved._hm_aligned is False
    one = BoxMesh(3)
    one.verts = one.verts.copy()
    one.verts[one.elem_verts[13, 7], 1] += 1e-12          # one interior vertex, one coordinate
    assert not aligned_box(one)
