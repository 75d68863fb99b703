"""The assembly configurations beside the headline one, each printed as JSON lines.

* ``precision``: the precision form on a perturbed mesh (interior vertices moved by up to
  0.2 h, so det J varies): every block assembled in FP64 and in the FP32, BF16 and FP16
  contraction forms; the error of each form against the FP64 values (largest entry error
  over the largest entry, and relative Frobenius norm) and the time of each assembly.
* ``nonlinear``: ``exp(m) (1 + u^2) grad u . grad p - f p`` at a smooth state (8 Fourier
  modes, amplitude 0.5): the residual and every block in FP64 and FP32, times and errors.

* ``linearization``: the blocks of a linearization point (A, C, W_um, W_mm, and W_uu at the
  nonlinear density) assembled block by block (``block``), and with the blocks of one shape
  two at a time from one element pass (``blocks``); FP64, both with the finiteness check of
  every element matrix; the two times and whether the values are the same bits.

* ``contraction``: A's element matrices alone, on the sum-factorized route (vector units)
  and on the dense matrix-core route (``v_mfma_f64_16x16x4f64`` / ``v_mfma_f32_16x16x4f32``),
  FP64 and FP32: the time of a launch; the counters of the same launches give each route's
  flop rate and matrix-core busy share against the peaks.

``python -m hippymfem_amd.configs precision --n 128`` / ``nonlinear --n 96`` /
``linearization --n 64 [--nonlinear]`` / ``contraction --n 64``.  ``--degree 6`` runs
``precision`` and ``linearization`` on the reference's default rule for P2 fields (4 x 4 x 4
Gauss points, ``elementbatch.py:148-156``) instead of the 3 x 3 x 3 rule.
"""

import argparse
import json
import time

import numpy as np
import torch

from ._lib import ADJOINT, PARAMETER, STATE
from .problem import PDEVariationalProblem
from .workloads import Workload

NAMES = {(ADJOINT, STATE): "A", (ADJOINT, PARAMETER): "C", (STATE, PARAMETER): "W_um",
         (PARAMETER, PARAMETER): "W_mm", (STATE, STATE): "W_uu"}


def _time(fn, steps):
    import gc

    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    gc.collect()
    gc.disable()                # no collection pause between the events
    try:
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
    finally:
        gc.enable()
    return e0.elapsed_time(e1) / steps


def _errors(d, ref):
    d = d.double()
    diff = d - ref
    return {"max_rel": float(diff.abs().max() / ref.abs().max()),
            "fro_rel": float(torch.linalg.vector_norm(diff) / torch.linalg.vector_norm(ref))}


def _drop(prob, ij, prec):
    """Free the cached assembly of block ij in precision prec (every route and option)."""
    for key in [k for k in prob._blocks if k[:3] == (ij[0], ij[1], prec)]:
        del prob._blocks[key]


def _route(blk):
    return "fused" if blk.fused_plan() is not None else "two-pass"


def _blocks(prob, loc, blocks, precisions, steps, n):
    out = []
    for ij in blocks:
        ref_blk = prob.block_assembly(*ij, precision="fp64")
        t64 = _time(lambda: ref_blk.assemble(loc), steps)
        ref = ref_blk.data.clone()
        nnz = ref_blk.pattern.nnz
        line = {"n": n, "block": NAMES[ij], "nnz": nnz, "fp64_ms": t64,
                "fp64_nnz_s": nnz / (t64 * 1e-3), "fp64_route": _route(ref_blk)}
        _drop(prob, ij, "fp64")
        del ref_blk
        for prec in precisions:
            blk = prob.block_assembly(*ij, precision=prec)
            t = _time(lambda: blk.assemble(loc), steps)
            line[prec] = dict(ms=t, nnz_s=nnz / (t * 1e-3), speedup_vs_fp64=t64 / t,
                              route=_route(blk), **_errors(blk.data, ref))
            _drop(prob, ij, prec)
            del blk
        del ref
        torch.cuda.empty_cache()
        print(json.dumps(line), flush=True)
        out.append(line)
    return out


def precision_form(n=128, perturb=0.2, steps=3, degree=None):
    W = Workload(n, perturb=perturb, quadrature_degree=degree)
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc)
    loc = [torch.as_tensor(v).cuda() for v in W.locals()]
    blocks = [(ADJOINT, STATE), (ADJOINT, PARAMETER), (STATE, PARAMETER), (PARAMETER, PARAMETER)]
    return _blocks(prob, loc, blocks, ("fp32", "bf16", "fp16"), steps, n)


def nonlinear(n=96, steps=3):
    W = Workload(n, nonlinear=True, state="fourier")
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc)
    loc = [torch.as_tensor(v).cuda() for v in W.locals()]
    x = [loc[0], loc[1], loc[2]]
    t_res = _time(lambda: prob.residual(x), steps)
    print(json.dumps({"n": n, "residual_fp64_ms": t_res, "dofs": W.Vu.ndof}), flush=True)
    blocks = [(ADJOINT, STATE), (ADJOINT, PARAMETER), (STATE, STATE), (STATE, PARAMETER),
              (PARAMETER, PARAMETER)]
    return _blocks(prob, loc, blocks, ("fp32",), steps, n)


def linearization(n=64, nonlinear=False, steps=3, degree=None):
    W = Workload(n, perturb=0.2, nonlinear=nonlinear, state="fourier" if nonlinear else "uniform",
                 quadrature_degree=degree)
    prob = PDEVariationalProblem(W.spaces, W.density, W.batches, W.bc,
                                 is_fwd_linear=not nonlinear)
    loc = [torch.as_tensor(v).cuda() for v in W.locals()]
    x = [loc[0], loc[1], loc[2]]
    need = [(ADJOINT, STATE), (ADJOINT, PARAMETER), (STATE, PARAMETER), (PARAMETER, PARAMETER)]
    if nonlinear:
        need.insert(2, (STATE, STATE))
    blks = [prob.block_assembly(*ij) for ij in need]
    t_sep = _time(lambda: [prob.block(ij[0], ij[1], x) for ij in need], steps)
    sep = [b.data.clone() for b in blks]
    t_shr = _time(lambda: prob.blocks(need, x), steps)
    same = all(torch.equal(b.data, d) for b, d in zip(blks, sep))
    line = {"n": n, "density": "nonlinear" if nonlinear else "linear", "degree": W.batches.groups[0].degree,
            "blocks": [NAMES[ij] for ij in need],
            "nnz": int(sum(b.pattern.nnz for b in blks)),
            "separate_ms": t_sep, "shared_ms": t_shr, "speedup": t_sep / t_shr,
            "same_bits": bool(same)}
    print(json.dumps(line), flush=True)
    return line


def contraction(n=64, steps=10):
    import os

    W = Workload(n)
    loc = [torch.as_tensor(v).cuda() for v in W.locals()]
    K = W.kernel
    ne = K.group.ne
    out = []
    old = os.environ.get("HM_CONTRACTION")
    try:
        for route in ("sumfact", "mfma"):
            os.environ["HM_CONTRACTION"] = route
            for prec in ("fp64", "fp32"):
                dt = torch.float64 if prec == "fp64" else torch.float32
                E = torch.empty((ne, K.nd[ADJOINT], K.nd[STATE]), dtype=dt, device="cuda")
                t = _time(lambda: K.element_matrices(ADJOINT, STATE, loc, precision=prec, out=E),
                          steps)
                line = {"n": n, "block": "A", "route": route, "precision": prec, "ms": t,
                        "elements_per_s": ne / (t * 1e-3)}
                print(json.dumps(line), flush=True)
                out.append(line)
                del E
    finally:
        if old is None:
            os.environ.pop("HM_CONTRACTION", None)
        else:
            os.environ["HM_CONTRACTION"] = old
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("which", choices=["precision", "nonlinear", "linearization", "contraction"])
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--nonlinear", action="store_true")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--degree", type=int, default=None,
                    help="quadrature degree of precision / linearization (default 4: 3x3x3 points)")
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    if a.which == "precision":
        precision_form(a.n or 128, steps=a.steps, degree=a.degree)
    elif a.which == "linearization":
        linearization(a.n or 64, nonlinear=a.nonlinear, steps=a.steps, degree=a.degree)
    elif a.which == "contraction":
        contraction(a.n or 64, steps=a.steps)
    else:
        nonlinear(a.n or 96, steps=a.steps)
    print(json.dumps({"wall_s": time.perf_counter() - t0}), flush=True)


if __name__ == "__main__":
    main()
