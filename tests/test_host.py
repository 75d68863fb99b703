"""CPU tests of the host side: the golden fixtures, the product's host logic against the
oracle (pattern build, tables, fields, boundary dofs), and the C ABI's exports."""

import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle import cref
from oracle import fe_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")

CASES = {"hex2_p1_linear": (2, 1, 0.0, "linear", "lattice"),
         "hex2_p2_nonlinear_perturbed": (2, 2, 0.2, "nonlinear", "lattice"),
         "hex4_p2_linear": (4, 2, 0.0, "linear", "lattice"),
         "hex3_p2_linear_mfem": (3, 2, 0.0, "linear", "mfem")}
BLOCKS = {"A": (2, 0), "C": (2, 1), "Wum": (0, 1), "Wmm": (1, 1), "Wuu": (0, 0)}


@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_fixtures(name):
    """The oracle still gives the committed fixtures (structure exactly, values to 1e-14)."""
    g = np.load(os.path.join(GOLDEN, name + ".npz"))
    n, order, perturb, density, numbering = CASES[name]
    P = O.Problem(n, order=order, perturb=perturb, density=density, numbering=numbering)
    assert np.array_equal(g["edofs_u"], P.Vu.edofs) and np.array_equal(g["ess"], P.ess)
    assert np.array_equal(g["m"], P.m) and np.array_equal(g["u"], P.u)
    r = P.residual(2)
    assert np.abs(r - g["residual"]).max() <= 1e-14 * np.abs(g["residual"]).max()
    for key, (i, j) in BLOCKS.items():
        if key + "_data" not in g:
            continue
        ip, ix, d = P.assembled(i, j)
        assert np.array_equal(ip, g[key + "_indptr"]) and np.array_equal(ix, g[key + "_indices"])
        assert np.abs(d - g[key + "_data"]).max() <= 1e-14 * np.abs(d).max(), key


# ------------------------------------------------------------- product host logic
hm_lib = pytest.importorskip("hippymfem_amd._lib")


def _built():
    return os.path.exists(hm_lib.LIB_PATH)


needs_lib = pytest.mark.skipif(not _built(), reason="native library not built")


@needs_lib
def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "hippymfem_amd.h")).read()
    declared = set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(hm_\w+)\s*\(", header, re.M))
    assert declared == set(hm_lib.EXPORTS)
    L = ctypes.CDLL(hm_lib.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name
    assert hm_lib.lib().hm_version() >= 1


def test_ctypes_structs_match_header(tmp_path):
    """The ctypes mirrors of hm_problem and hm_scatter_plan have the header's layout: every
    field's offset and the size, as gcc lays the C structs out."""
    structs = {"hm_problem": hm_lib.HmProblem, "hm_scatter_plan": hm_lib.HmScatterPlan,
               "hm_fused_plan": hm_lib.HmFusedPlan}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "hippymfem_amd.h"',
             "int main(void) {"]
    for cname, cls in structs.items():
        lines.append('printf("%%s size %%zu\\n", "%s", sizeof(%s));' % (cname, cname))
        for f, _ in cls._fields_:
            lines.append('printf("%%s %%s %%zu\\n", "%s", "%s", offsetof(%s, %s));'
                         % (cname, f, cname, f))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            parts = line.split()
            got[(parts[0], parts[1])] = int(parts[2])
    for cname, cls in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert got[(cname, f)] == getattr(cls, f).offset, (cname, f)


@needs_lib
@pytest.mark.parametrize("numbering", ["lattice", "mfem", "shuffled"])
def test_product_host_matches_oracle(numbering):
    from hippymfem_amd.elementbatch import group_tables
    from hippymfem_amd.pattern import ScatterPattern, VectorPattern
    from hippymfem_amd.workloads import Workload

    W = Workload(3, numbering=numbering, perturb=0.2)
    P = O.Problem(3, numbering=numbering, perturb=0.2)
    assert np.array_equal(W.mesh.verts, P.mesh.verts)
    assert np.array_equal(W.Vu.edofs, P.Vu.edofs) and np.array_equal(W.Vm.edofs, P.Vm.edofs)
    assert np.array_equal(W.m, P.m) and np.array_equal(W.u, P.u) and np.array_equal(W.p, P.p)
    assert np.array_equal(W.ess, P.ess)
    g = W.groups[0]
    tu = group_tables(W.Vu, g)
    assert np.abs(tu.N - P.data.Nu).max() < 1e-15 and np.abs(tu.G - P.data.Gu).max() < 1e-14
    assert np.array_equal(g.w, P.data.w)
    for (T, C) in ((W.Vu, W.Vu), (W.Vu, W.Vm), (W.Vm, W.Vu), (W.Vm, W.Vm)):
        S = ScatterPattern([group_tables(T, g)], [group_tables(C, g)], T, C)
        ip, ix, slot, nnz = cref.pattern_tables(T.edofs, C.edofs, T.ndof, C.ndof, T is C)
        assert S.nnz == nnz
        assert S.indptr.dtype == np.int32 and np.array_equal(S.indptr, ip)
        assert S.indices.dtype == np.int32 and np.array_equal(S.indices, ix)
        assert np.array_equal(S.slot, slot)
        # runs: element rows per CSR row, ascending
        rc, runs = O.runs_of_rows(T.edofs, T.ndof)
        assert np.array_equal(S.rptr, rc) and np.array_equal(S.runs, runs)
    V = VectorPattern([tu], W.Vu)
    rc, runs = O.runs_of_rows(W.Vu.edofs, W.Vu.ndof)
    assert np.array_equal(V.rptr, rc) and np.array_equal(V.runs, runs)


@needs_lib
def test_product_pattern_errors():
    from hippymfem_amd.pattern import build_tables

    tdofs = np.zeros((2500, 27), np.int32)
    cdofs = np.tile(np.arange(27, dtype=np.int32), (2500, 1))
    with pytest.raises(ValueError):
        build_tables(tdofs, cdofs, 1, 27, False)
    with pytest.raises(ValueError):
        build_tables(np.full((1, 2), 5, np.int32), np.zeros((1, 2), np.int32), 3, 3, False)


@needs_lib
def test_product_masked_slots_match_oracle():
    from hippymfem_amd.elementbatch import group_tables
    from hippymfem_amd.pattern import ScatterPattern
    from hippymfem_amd.workloads import Workload

    W = Workload(3)
    g = W.groups[0]
    S = ScatterPattern([group_tables(W.Vu, g)], [group_tables(W.Vu, g)], W.Vu, W.Vu)
    mask = O.ess_mask(W.Vu.ndof, W.ess)
    assert np.array_equal(S.masked_slots(mask, mask), O.masked_slots(S.indptr, S.indices, mask, mask))
    assert np.array_equal(S.masked_slots(mask, None), O.masked_slots(S.indptr, S.indices, mask, None))


def test_bench_refuses_several_gpus():
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode != 0 and "one GPU" in res.stderr


def test_product_fails_loudly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    if not _built():
        pytest.skip("native library not built")
    from hippymfem_amd.workloads import Workload

    W = Workload(2)
    with pytest.raises(RuntimeError):
        W.kernel.element_matrices(2, 0, W.locals())


@needs_lib
@pytest.mark.parametrize("nd_c,longest,lanes", [(8, 257, 8), (8, 300, 8), (16, 520, 16),
                                                (27, 1100, 32)])
def test_scatter_refuses_rows_beyond_its_buffers(nd_c, longest, lanes):
    """The scatter holds 256 / LPR row buffers of the longest row in LDS, LPR the least
    power of two of at least nd_c lanes; a longer row is refused before any launch with
    HM_ERR_UNSUPPORTED (a Q1-column row of more than 256 entries, with 8 lanes a row)."""
    L = hm_lib.lib()
    assert (256 // lanes) * longest * 8 > 64 * 1024
    plan = hm_lib.HmScatterPlan(nrow=4, nnz=4 * longest, ne=1, nd_t=8, nd_c=nd_c,
                                code_bytes=2, max_row=longest)
    rc = L.hm_scatter_matrix(ctypes.byref(plan), None, hm_lib.FP64, 0, 1, None, hm_lib.FP64,
                             0, None)
    assert rc == hm_lib.HM_ERR_UNSUPPORTED
    msg = L.hm_last_error().decode()
    assert "row of %d entries" % longest in msg and str(256 // lanes) in msg
    with pytest.raises(NotImplementedError):
        hm_lib.check(rc, "scatter")


def test_fused_item_slot_division_is_exact():
    """The fused kernel takes an element-data item's slot it // NI as the truncated FP32 product
    (it + 1/2) * (1 / NI), both roundings to nearest; exact over the kernel's range (it below
    2^13, NI at most 128), checked here in the same FP32 arithmetic."""
    import numpy as np

    it = np.arange(1 << 13, dtype=np.int64)
    for ni in range(1, 129):
        r = np.float32(1.0) / np.float32(ni)
        got = ((it.astype(np.float32) + np.float32(0.5)) * r).astype(np.int64)
        assert np.array_equal(got, it // ni), ni
