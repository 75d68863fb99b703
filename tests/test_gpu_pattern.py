"""The device pattern build against the host build and the oracle: every array bit for bit
(``patternbuild.build_tables``, ``patternbuild.py:209-253``; the sort route of
``pattern.py:548-663`` gives the same CSR graph), its error behaviour, and an assembly on a
device-built pattern equal to one on the host-built pattern."""

import numpy as np
import pytest
import torch

from oracle import fe_oracle as O

from hippymfem_amd import ADJOINT, STATE
from hippymfem_amd.elementbatch import group_tables
from hippymfem_amd.pattern import ScatterPattern, build_tables, build_tables_device
from hippymfem_amd.workloads import Workload

pytestmark = pytest.mark.gpu


def _host_and_device(tdofs, cdofs, nrow, ncol, diag_first):
    ip, rp, runs, pos, nnz, mx = build_tables(tdofs, cdofs, nrow, ncol, diag_first)
    d = build_tables_device(tdofs, cdofs, nrow, ncol, diag_first)
    assert d["nnz"] == nnz and d["max_row"] == mx
    assert np.array_equal(d["indptr"].cpu().numpy(), ip)
    assert np.array_equal(d["rptr"].cpu().numpy(), rp)
    assert np.array_equal(d["runs"].cpu().numpy(), runs)
    assert np.array_equal(d["pos"].cpu().numpy().view(np.uint16), pos)
    return d, (ip, rp, runs, pos, nnz, mx)


@pytest.mark.parametrize("numbering", ["lattice", "mfem", "shuffled"])
def test_device_build_equals_host_and_sort_route(gpu, numbering):
    W = Workload(5, numbering=numbering, perturb=0.1)
    for T, C in ((W.Vu, W.Vu), (W.Vu, W.Vm), (W.Vm, W.Vu), (W.Vm, W.Vm)):
        square = T is C
        d, _ = _host_and_device(T.edofs, C.edofs, T.ndof, C.ndof, square)
        ip, ix, slot, nnz = O.pattern_tables_sort(T.edofs, C.edofs, T.ndof, C.ndof, square)
        assert np.array_equal(d["indptr"].cpu().numpy(), ip)
        assert np.array_equal(d["indices"].cpu().numpy(), ix)
        # place of every entry: indptr[row] + pos == the sort route's slot
        rows = np.repeat(np.asarray(T.edofs, np.int64).reshape(-1), C.edofs.shape[1])
        got = ip[rows] + d["pos"].cpu().numpy().view(np.uint16)
        assert np.array_equal(got, slot)
        rc, runs = O.runs_of_rows(T.edofs, T.ndof)
        assert np.array_equal(d["rptr"].cpu().numpy(), rc)
        assert np.array_equal(d["runs"].cpu().numpy(), runs)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_device_build_random_tables(gpu, seed):
    """Random dof tables: empty rows, rows shared by many elements, repeated columns inside
    an element row (flagged by both builds alike), square and rectangular."""
    rng = np.random.default_rng(seed)
    for ne, nd_t, nd_c, nrow, ncol, square in ((300, 4, 6, 97, 120, False),
                                               (500, 8, 8, 400, 400, True),
                                               (64, 27, 27, 50, 50, True),
                                               (1000, 3, 5, 2000, 11, False)):
        td = rng.integers(0, nrow, size=(ne, nd_t), dtype=np.int32)
        cd = rng.integers(0, ncol, size=(ne, nd_c), dtype=np.int32)
        d, host = _host_and_device(td, cd, nrow, ncol, square)
        ip, rp, runs, pos, nnz, mx = host
        ix = np.empty(nnz, np.int32)
        from hippymfem_amd import _lib
        _lib.check(_lib.lib().hm_pattern_place(_lib.ptr(cd), ne, nd_t, nd_c, nrow, _lib.ptr(rp),
                                               _lib.ptr(runs), _lib.ptr(ip), _lib.ptr(pos), 0,
                                               _lib.ptr(ix), None))
        assert np.array_equal(d["indices"].cpu().numpy(), ix)
        rep = _lib.I32(0)
        _lib.check(_lib.lib().hm_pattern_repeats(ne, nd_t, nd_c, _lib.ptr(pos), 0,
                                                 _lib.ctypes.byref(rep)))
        assert d["repeats"] == bool(rep.value)


def test_device_build_edges_and_errors(gpu):
    # no elements: every row empty
    d = build_tables_device(np.zeros((0, 4), np.int32), np.zeros((0, 4), np.int32), 7, 7, True)
    assert d["nnz"] == 0 and np.array_equal(d["indptr"].cpu().numpy(), np.zeros(8, np.int32))
    # a row dof and a column dof out of range
    with pytest.raises(ValueError):
        build_tables_device(np.full((1, 2), 5, np.int32), np.zeros((1, 2), np.int32), 3, 3, False)
    with pytest.raises(ValueError):
        build_tables_device(np.zeros((1, 2), np.int32), np.full((1, 2), -1, np.int32), 3, 3, False)
    # a row of 2**16 entries or more: ValueError, as the host build
    tdofs = np.zeros((2500, 27), np.int32)
    cdofs = np.tile(np.arange(27, dtype=np.int32), (2500, 1))
    with pytest.raises(ValueError):
        build_tables_device(tdofs, cdofs, 1, 27, False)
    # a row above the device build's 2048 entries (and below the key limit): refused
    with pytest.raises(NotImplementedError):
        build_tables_device(np.zeros((100, 1), np.int32), cdofs[:100], 1, 27, False)
    # a device pattern with a repeated column is refused like a host one
    with pytest.raises(NotImplementedError):
        _pattern_with_repeat()


def _pattern_with_repeat():
    class _G:
        ne = 1

    class _T:
        group = _G()
        nd_total = 2
        edofs = np.array([[0, 0]], np.int32)

    class _S:
        ndof = 2

    return ScatterPattern([_T()], [_T()], _S(), _S(), build="device")


def test_device_built_pattern_assembles_identically(gpu):
    """64^3 P2 Jacobian A on a device-built pattern: the graph and the assembled values
    bit for bit those of the host-built pattern."""
    W = Workload(64)
    g = W.groups[0]
    tu = group_tables(W.Vu, g)
    host = ScatterPattern([tu], [tu], W.Vu, W.Vu, build="host")
    torch.cuda.synchronize()
    dev = ScatterPattern([tu], [tu], W.Vu, W.Vu, build="device")
    assert ScatterPattern([tu], [tu], W.Vu, W.Vu).build == "device"    # the default on a GPU
    assert dev.nnz == host.nnz == 513 ** 3
    for name in ("indptr", "indices", "rptr", "runs", "pos"):
        assert np.array_equal(getattr(dev, name), getattr(host, name)), name
    E = W.kernel.element_matrices(ADJOINT, STATE, W.locals())[0]
    mask = O.ess_mask(W.Vu.ndof, W.ess)
    a = host.data([E], zero_slots=None, plan=host.device_plan(mask, mask, 1.0))
    b = dev.data([E], zero_slots=None, plan=dev.device_plan(mask, mask, 1.0))
    assert torch.equal(a, b)
