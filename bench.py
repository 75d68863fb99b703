"""Assembly benchmark: assembled CSR nnz/s of the forward Jacobian at 64^3 P2 on one GPU.

A step is one full assembly of the block A = d2R/dp du of ``exp(m) grad u . grad p - f p``
on the unit cube in 64^3 hexahedra, state Q2 / parameter Q1, 3x3x3 Gauss rule, Dirichlet
u = 0 on x = 0 and x = 1 (folded elimination, diagonal 1) into hypre-layout CSR values.  On
this box mesh the step is one launch of the fused kernel (geometry, fields, pointwise blocks,
contraction and the sums into CSR rows, with no element matrix in memory); the two-pass
route (element kernel into resident element matrices, then the scatter) of the same block is
timed beside it and gives the same bits.  Inputs (mesh, fields, pattern, plan) are resident
in HBM before the timed region; the pattern build and the fused plan are one-time set-up, as
in the reference's warm-up (``benchmarks/bench_assembly.py:10-17``).

Prints one JSON line.  ``--gpus`` above 1 is refused: this path is measured on one GPU.
``--copy-calibration`` runs a plain device copy of known size (the reference point for the
HBM byte counters) and prints its size.
"""

import argparse
import json
import os
import re
import sys
import time

import numpy as np

METRIC = ("assembled CSR nnz/s + % HBM roofline at 64³ P2 (FP64 vs FP32-MFMA); "
          "p50/p99 latency")
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6    # MI355X dense FP64 (matrix and vector)
HERE = os.path.dirname(os.path.abspath(__file__))
TRAFFIC_FILE = os.path.join(HERE, "profiles", "hbm_traffic_A_64cube.md")


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=64, help="hexahedra per direction")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving sample")
    ap.add_argument("--no-default-rule", action="store_true",
                    help="skip the assembly on the 4x4x4 rule (so a profile's averages are the headline's)")
    ap.add_argument("--cpu-n", type=int, default=12, help="mesh of the CPU baseline sample")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain 1 GiB device copy five times and exit")
    return ap.parse_args()


def cpu_baseline(n):
    """The oracle (numpy restatement of the reference's host path) on an n^3 sample:
    element kernel in 2048-element windows (``kernel.py:659-672``), ``np.bincount``
    scatter (``pattern.py:788``), folded elimination; pattern build excluded.  Timed on one
    core (BLAS limited to one thread), then on every core the process may use."""
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:                                             # pragma: no cover
        threadpool_limits = None
    from oracle import fe_oracle as O

    P = O.Problem(n)
    ip, ix, slot, nnz = O.pattern_tables_sort(P.Vu.edofs, P.Vu.edofs, P.Vu.ndof, P.Vu.ndof,
                                              True)
    mask = O.ess_mask(P.Vu.ndof, P.ess)
    kill = O.masked_slots(ip, ix, mask, mask)

    def once():
        E = np.empty((P.mesh.ne, 27, 27))
        for a in range(0, P.mesh.ne, 2048):
            el = np.arange(a, min(a + 2048, P.mesh.ne))
            E[a:a + el.size] = P.block(2, 0, el)
        d = O.scatter(slot, E, nnz, kill)
        d[ip[P.ess]] = 1.0
        return d

    import contextlib

    def timed(seconds):
        once()
        t0 = time.perf_counter()
        reps = 0
        while True:
            once()
            reps += 1
            if time.perf_counter() - t0 > seconds or reps >= 20:
                break
        return (time.perf_counter() - t0) / reps, reps

    with (threadpool_limits(limits=1) if threadpool_limits else contextlib.nullcontext()):
        dt, reps = timed(10.0)
    try:
        cores = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):                               # pragma: no cover
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 15))
    dta, repsa = _parallel_cpu(P, ip, slot, nnz, kill, cores)
    return {"value": nnz / dt, "unit": "nnz/s", "cores": 1, "kind": "port",
            "sample": "%d^3 P2 Jacobian A assembly (%d elements, %d nnz), %d reps, "
                      "%.1f us/elem" % (n, P.mesh.ne, nnz, reps, 1e6 * dt / P.mesh.ne),
            "all_cores": {"value": nnz / dta, "unit": "nnz/s", "cores": cores,
                          "sample": "the same, %d reps, element windows over %d worker "
                                    "processes of one thread each, %.1f us/elem"
                                    % (repsa, cores, 1e6 * dta / P.mesh.ne)}}


_CPU_PROBLEM = None


def _cpu_window(el):
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:                                             # pragma: no cover
        return _CPU_PROBLEM.block(2, 0, el)
    with threadpool_limits(limits=1):
        return _CPU_PROBLEM.block(2, 0, el)


def _parallel_cpu(P, ip, slot, nnz, kill, workers):
    """The same assembly with the element matrices of equal element windows computed by
    ``workers`` forked processes (one BLAS thread each) and scattered in the parent."""
    import multiprocessing as mp

    from oracle import fe_oracle as O

    global _CPU_PROBLEM
    _CPU_PROBLEM = P
    windows = np.array_split(np.arange(P.mesh.ne), workers)
    with mp.get_context("fork").Pool(workers) as pool:
        def once():
            E = np.concatenate(pool.map(_cpu_window, windows))
            d = O.scatter(slot, E, nnz, kill)
            d[ip[P.ess]] = 1.0
            return d

        once()
        t0 = time.perf_counter()
        reps = 0
        while True:
            once()
            reps += 1
            if time.perf_counter() - t0 > 5.0 or reps >= 20:
                break
        return (time.perf_counter() - t0) / reps, reps


def timed_steps(blk, loc, steps, warmup, stream):
    """Warm-up, then ``steps`` assemblies bracketed by synchronizations; HIP events on the
    stream the kernels run on time every launch: the fused kernel (one launch a step) on the
    fused route, else the element kernel and the scatter.  Returns (wall time a step, first
    launch, second launch or 0)."""
    import gc

    import torch

    fused = blk.fused_plan() is not None
    for _ in range(warmup):
        blk.assemble_values(loc, stream=stream)
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    # no garbage collection pause inside the timed region (a full collection after building
    # a block's plan takes tens of milliseconds of host time)
    gc.collect()
    gc.disable()
    try:
        t0 = time.perf_counter()
        for k in range(steps):
            ev[k][0].record(stream)
            if fused:
                blk.assemble_values(loc, stream=stream)
                ev[k][1].record(stream)
            else:
                blk.element_matrices(loc, stream=stream)
                ev[k][1].record(stream)
                blk.scatter(stream=stream)
            ev[k][2].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    finally:
        gc.enable()
    tk = float(np.mean([ev[k][0].elapsed_time(ev[k][1]) for k in range(steps)])) * 1e-3
    ts = float(np.mean([ev[k][1].elapsed_time(ev[k][2]) for k in range(steps)])) * 1e-3
    return wall / steps, tk, (0.0 if fused else ts)


def fused_bytes(blk, W):
    """HBM bytes the fused launch moves at the least: the values written (the single-pass
    floor, B_min = vsize nnz), the plan read (the stencil map, one byte an entry; the row
    headers, 8 bytes a row) and the element data (vertex and dof ids, vertex coordinates and
    dof values, the fields only where the block reads them)."""
    pat = blk.pattern
    k = blk.kernel
    vsize = blk.data.element_size()
    ne = pat.ne
    Vu, Vm = k.spaces[0], k.spaces[1]
    fields = not (blk.i == 2 and blk.j == 0 and not k.density.nonlinear)
    b_min = vsize * pat.nnz
    b_plan = pat.nnz + 8 * pat.nrow
    b_inputs = (4 * 8 * ne + 24 * W.mesh.nvert + 4 * ne * Vm.nd + 8 * Vm.ndof +
                ((4 * ne * Vu.nd + 16 * Vu.ndof) if fields else 0))
    return b_min, b_min + b_plan + b_inputs


def algorithmic_bytes(blk, W):
    """HBM bytes a launch must move at the least (every operand read once, every result
    written once).

    element kernel: the element matrices written (8 or 4 bytes an entry), the element vertex
    and dof tables and the vertex coordinates and dof values read (the fields only where the
    block reads them); scatter: the element matrices read, one code byte an entry, the row
    tables (a 48-byte row header with the first eight runs where the element rows have more
    than 8 entries; otherwise the runs, 4 bytes an element row, and indptr, rptr and rowmode,
    9 bytes a row) and the values written."""
    pat = blk.pattern
    k = blk.kernel
    esize = 8 if blk.precision == "fp64" else 4
    vsize = blk.data.element_size()
    ne, nd_t, nd_c = pat.ne, pat.nd_t, pat.nd_c
    nent = ne * nd_t * nd_c
    Vu, Vm = k.spaces[0], k.spaces[1]
    fields = not (blk.i == 2 and blk.j == 0 and not k.density.nonlinear)
    b_elem = (esize * nent + 4 * 8 * ne + 24 * W.mesh.nvert + 4 * ne * Vm.nd + 8 * Vm.ndof +
              ((4 * ne * Vu.nd + 16 * Vu.ndof) if fields else 0))
    tables = 48 * pat.nrow if nd_c > 8 else 4 * ne * nd_t + 9 * pat.nrow
    b_scatter = esize * nent + nent + tables + vsize * pat.nnz
    return b_elem, b_scatter


def measured_traffic():
    """HBM bytes a launch moved, read by the byte counters (``FETCH_SIZE`` + ``WRITE_SIZE``,
    scaled by a plain copy's reading) from the committed profile summary; None without one."""
    try:
        text = open(TRAFFIC_FILE).read()
    except OSError:
        return {}
    out = {}
    for key in ("element_kernel", "scatter", "fused"):
        m = re.search(r"^%s_traffic_bytes:\s*([0-9.eE+]+)" % key, text, re.M)
        if m:
            out[key] = float(m.group(1))
    return out


def fp32_form(prob, loc, blk64, W, args, stream):
    """The precision form: the Jacobian, its determinant and inverse in FP64; the fields, the
    pointwise pairs, the contraction and the values in FP32 (the fused route, two blocks a
    compute unit on this box mesh, sums the FP32 entries in FP32 in ascending element order;
    the two-pass route sums them in FP64).  Error against the FP64 values of the same step."""
    import torch

    from hippymfem_amd import ADJOINT, STATE

    blk = prob.block_assembly(ADJOINT, STATE, precision="fp32")
    fused = blk.fused_plan() is not None
    step, tk, ts = timed_steps(blk, loc, args.steps, args.warmup, stream)
    d64 = blk64.data
    d32 = blk.data.double()
    err_max = float((d32 - d64).abs().max() / d64.abs().max())
    err_fro = float(torch.linalg.vector_norm(d32 - d64) / torch.linalg.vector_norm(d64))
    if fused:
        _, b_e2e = fused_bytes(blk, W)
        split = {"fused_kernel_ms": 1e3 * tk}
    else:
        b_elem, b_scatter = algorithmic_bytes(blk, W)
        b_e2e = b_elem + b_scatter
        split = {"element_kernel_ms": 1e3 * tk, "scatter_ms": 1e3 * ts}
    return dict({"ms_per_step": 1e3 * step, "value": blk.pattern.nnz / step, "unit": "nnz/s",
                 "route": "fused" if fused else "two-pass", "speedup_vs_fp64": None,
                 "e2e": {"bytes": b_e2e, "GBs": b_e2e / step / 1e9,
                         "frac": b_e2e / step / 1e9 / HBM_PEAK_GBS},
                 "max_rel_err_vs_fp64": err_max, "fro_rel_err_vs_fp64": err_fro}, **split)


def linearization_point(prob, loc, args):
    """A and the Hessian blocks C, W_um, W_mm of the same problem at the seeded state,
    parameter and adjoint (``set_linearization_point``; W_uu is zero for this linear density
    and not assembled, as in the reference), FP64, with every block's elimination and
    non-finite check; C and W_um from one element pass.  Host wall time a point."""
    import gc

    import torch

    from hippymfem_amd import ADJOINT, PARAMETER, STATE

    x = [loc[STATE], loc[PARAMETER], loc[ADJOINT]]
    for _ in range(args.warmup):
        got = prob.set_linearization_point(x)
    torch.cuda.synchronize()
    gc.collect()
    gc.disable()
    try:
        t0 = time.perf_counter()
        for _ in range(args.steps):
            got = prob.set_linearization_point(x)
        torch.cuda.synchronize()
        step = (time.perf_counter() - t0) / args.steps
    finally:
        gc.enable()
    names = {(ADJOINT, STATE): "A", (ADJOINT, PARAMETER): "C", (STATE, PARAMETER): "W_um",
             (PARAMETER, PARAMETER): "W_mm", (STATE, STATE): "W_uu"}
    nnz = int(sum(prob.block_assembly(*ij).pattern.nnz for ij in got))
    return {"blocks": [names[ij] for ij in got], "nnz": nnz, "ms_per_point": 1e3 * step,
            "nnz_s": nnz / step}


def default_rule(args, loc, stream):
    """The same A assembly (FP64) on the reference's default rule for P2 fields: degree
    2 * 2 + 2 = 6 (``elementbatch.py:148-156``), 4 x 4 x 4 Gauss points, against the
    3 x 3 x 3 rule of the headline.  Same mesh, fields and CSR pattern."""
    import torch

    from hippymfem_amd import ADJOINT, STATE
    from hippymfem_amd.problem import PDEVariationalProblem
    from hippymfem_amd.workloads import Workload

    W = Workload(args.n, quadrature_degree=6)
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc)
    blk = prob.block_assembly(ADJOINT, STATE)
    fused = blk.fused_plan() is not None
    step, tk, ts = timed_steps(blk, loc, args.steps, args.warmup, stream)
    out = {"quadrature": "4x4x4 Gauss (degree 6)", "ms_per_step": 1e3 * step,
           "value": blk.pattern.nnz / step, "unit": "nnz/s",
           "route": "fused" if fused else "two-pass"}
    out.update({"fused_kernel_ms": 1e3 * tk} if fused else
               {"element_kernel_ms": 1e3 * tk, "scatter_ms": 1e3 * ts})
    del blk, prob
    torch.cuda.empty_cache()
    return out


def block_setup(args):
    """One-time set-up of the headline block on fresh spaces (nothing cached): the CSR
    pattern built on the device (``hm_pattern_build_device``, ``hm_pattern_place_device``),
    the elimination plan, then the fused kernel's plan (row headers and stencil maps), and the
    two-pass route's scatter plan (codes in run order, row headers) for comparison.  Host wall
    time, synchronized."""
    import torch

    from hippymfem_amd import ADJOINT, STATE
    from hippymfem_amd.problem import PDEVariationalProblem
    from hippymfem_amd.workloads import Workload

    W = Workload(args.n)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc)
    blk = prob.block_assembly(ADJOINT, STATE)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    fused = blk.fused_plan() is not None
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    blk.dplan
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    out = {"block_setup_ms": 1e3 * (t2 - t0), "pattern_ms": 1e3 * (t1 - t0),
           "fused_plan_ms": 1e3 * (t2 - t1) if fused else None,
           "two_pass_scatter_plan_ms": 1e3 * (t3 - t2),
           "route": "fused" if fused else "two-pass"}
    del blk, prob, W
    torch.cuda.empty_cache()
    return out


def copy_calibration():
    import torch

    n = 1 << 27                                                     # 2^27 doubles = 1 GiB
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "launches": 5}))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures this path on one GPU; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    # the CPU baseline first: its worker processes are forked before this process opens the GPU
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_n)
    import torch

    from hippymfem_amd import ADJOINT, STATE
    from hippymfem_amd.problem import PDEVariationalProblem
    from hippymfem_amd.workloads import Workload

    dev = torch.device("cuda", 0)
    W = Workload(args.n)
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc)
    blk = prob.block_assembly(ADJOINT, STATE)
    loc = [torch.as_tensor(v).to(dev) for v in (W.u, W.m, W.p)]
    stream = torch.cuda.current_stream()
    fused = blk.fused_plan() is not None
    step, t_kernel, _ = timed_steps(blk, loc, args.steps, args.warmup, stream)
    pat = blk.pattern
    traffic = measured_traffic()
    # the contraction's algorithmic flops: the dense formulation E = Psi^T (w D) Psi of the
    # stiffness-type block, 2 * 27 * 27 * 81 an element (the sum-factorized kernel does fewer)
    flops = 2.0 * pat.nd_t * pat.nd_c * 81 * pat.ne
    # the two-pass route of the same block (element matrices in memory, then the scatter)
    blk2 = prob.block_assembly(ADJOINT, STATE, route="two-pass")
    step2, t2_kernel, t2_scatter = timed_steps(blk2, loc, args.steps, args.warmup, stream)
    b_elem, b_scatter = algorithmic_bytes(blk2, W)
    two_pass = {
        "ms_per_step": 1e3 * step2, "value": pat.nnz / step2, "unit": "nnz/s",
        "kernels": {
            "k_sumfact": {"ms": 1e3 * t2_kernel, "bytes": b_elem,
                          "GBs": b_elem / t2_kernel / 1e9,
                          "traffic": traffic.get("element_kernel"),
                          "algorithmic_flops": flops, "TFLOPs": flops / t2_kernel / 1e12},
            "k_scatter_rows": {"ms": 1e3 * t2_scatter, "bytes": b_scatter,
                               "GBs": b_scatter / t2_scatter / 1e9,
                               "traffic": traffic.get("scatter")}},
        "e2e": {"bytes": b_elem + b_scatter,
                "frac": (b_elem + b_scatter) / step2 / 1e9 / HBM_PEAK_GBS},
        "same_bits_as_fused": bool(torch.equal(blk.data, blk2.data)) if fused else None}
    if fused:
        b_min, b_alg = fused_bytes(blk, W)
        kernels = {"k_fused": {"ms": 1e3 * t_kernel, "bytes_min": b_min, "bytes": b_alg,
                               "GBs_min": b_min / t_kernel / 1e9, "GBs": b_alg / t_kernel / 1e9,
                               "traffic": traffic.get("fused"),
                               "algorithmic_flops": flops, "TFLOPs": flops / t_kernel / 1e12,
                               "frac_fp64_peak": flops / t_kernel / 1e12 / FP64_PEAK_TFLOPS,
                               "tile": list(blk.fused_plan().tile)}}
        d = kernels["k_fused"]
        roof = {"kernel": "k_fused", "bound": "hbm", "achieved": d["GBs_min"],
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": d["GBs_min"] / HBM_PEAK_GBS,
                "traffic": d["traffic"],
                "basis": "single-pass floor B_min = 8 nnz (the values written); the launch is "
                         "bound by its FP64 arithmetic, see kernels.k_fused.TFLOPs"}
        b_e2e = b_alg
    else:
        kernels = two_pass["kernels"]
        dom = max(kernels, key=lambda k: kernels[k]["ms"])
        d = kernels[dom]
        roof = {"kernel": dom, "bound": "hbm", "achieved": d["GBs"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": d["GBs"] / HBM_PEAK_GBS, "traffic": d["traffic"]}
        b_e2e = b_elem + b_scatter
    out = {
        "metric": METRIC, "value": pat.nnz / step, "unit": "nnz/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded fields, unit-cube mesh)",
        "config": {"workload": "%d^3 hex P2 Jacobian A, exp(m) diffusion, FP64, 3x3x3 Gauss, "
                               "Dirichlet x=0,1 folded" % args.n,
                   "elements": pat.ne, "nnz": pat.nnz, "entries": pat.ne * pat.nd_t * pat.nd_c,
                   "route": "fused" if fused else "two-pass", "parallelism": "1 GPU"},
        "roofline": roof,
        "kernels": kernels,
        "e2e": {"bytes": b_e2e, "GBs": b_e2e / step / 1e9,
                "frac": b_e2e / step / 1e9 / HBM_PEAK_GBS},
        "two_pass": two_pass,
    }
    out["linearization_point"] = linearization_point(prob, loc, args)
    if not args.no_default_rule:
        out["default_rule"] = default_rule(args, loc, stream)
    out["fp32_form"] = fp32_form(prob, loc, blk, W, args, stream)
    out["fp32_form"]["speedup_vs_fp64"] = step / (out["fp32_form"]["ms_per_step"] * 1e-3)
    out["setup"] = block_setup(args)
    if not args.no_serving:
        from hippymfem_amd import serving

        out["serving"] = serving.bench_sample(seconds=20.0)
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    sys.path.insert(0, HERE)
    main()
