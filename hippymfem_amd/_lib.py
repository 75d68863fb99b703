"""Loader of the native library (``_lib/libhippymfem_amd.so``) and its C structures.

torch is imported first so that the process has one HIP runtime, the one torch loaded,
which the library then binds to.  A missing library is an error: there is no fallback.
"""

import ctypes
import os

import torch  # noqa: F401  (binds the HIP runtime before the library is loaded)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libhippymfem_amd.so")

STATE, PARAMETER, ADJOINT = 0, 1, 2
FP64, FP32, BF16, FP16 = 0, 1, 2, 3
PRECISIONS = {"fp64": FP64, "fp32": FP32, "bf16": BF16, "fp16": FP16}

HM_OK = 0
#: the largest lattice edge of ``hm_serving_field_noalloc`` (its static workspace)
SERVING_FIELD_NOALLOC_MAX_N1 = 65
HM_ERR_ARG, HM_ERR_HIP, HM_ERR_ROW_LONG, HM_ERR_DUP, HM_ERR_NODIAG, HM_ERR_UNSUPPORTED = (
    -1, -2, -3, -4, -5, -6)

#: every symbol ``include/hippymfem_amd.h`` declares
EXPORTS = ("hm_last_error", "hm_version", "hm_element_matrices", "hm_element_matrices_pair",
           "hm_consistent_block", "hm_element_vectors", "hm_element_hvp", "hm_element_scalars",
           "hm_point_fields", "hm_element_matrices_pointwise", "hm_element_vectors_pointwise",
           "hm_pattern_build",
           "hm_pattern_place", "hm_pattern_repeats", "hm_pattern_work_bytes",
           "hm_pattern_build_device", "hm_pattern_place_device", "hm_scatter_matrix",
           "hm_scatter_matrix_defect", "hm_scatter_vector", "hm_csr_matvec", "hm_fused_tile",
           "hm_fused_blocks", "hm_assemble_fused", "hm_serving_normal", "hm_serving_field",
           "hm_serving_field_noalloc")

P = ctypes.c_void_p
I32, I64, DBL = ctypes.c_int32, ctypes.c_int64, ctypes.c_double


class HmProblem(ctypes.Structure):
    """``hm_problem`` of ``include/hippymfem_amd.h``."""
    _fields_ = [("ne", I64), ("nq", I32), ("nd_u", I32), ("nd_m", I32), ("density", I32),
                ("f", DBL), ("verts", P), ("elem_verts", P), ("edofs_u", P), ("edofs_m", P),
                ("tab_u", P), ("tab_m", P), ("tab_x", P), ("wq", P), ("u", P), ("m", P),
                ("p", P), ("nq1", I32), ("n1_u", I32), ("n1_m", I32), ("reserved", I32),
                ("tab1_u", P), ("tab1_m", P), ("lex_u", P), ("lex_m", P), ("nonfinite", P)]


class HmScatterPlan(ctypes.Structure):
    """``hm_scatter_plan`` of ``include/hippymfem_amd.h``."""
    _fields_ = [("nrow", I64), ("nnz", I64), ("ne", I64), ("nd_t", I32), ("nd_c", I32),
                ("code_bytes", I32), ("max_row", I32), ("indptr", P), ("rptr", P),
                ("runs", P), ("codes", P), ("rowmode", P), ("diag", DBL), ("rowhead", P)]


class HmFusedPlan(ctypes.Structure):
    """``hm_fused_plan`` of ``include/hippymfem_amd.h``."""
    _fields_ = [("nx", I32), ("ny", I32), ("nz", I32), ("bx", I32), ("by", I32),
                ("flags", I32), ("hdr", P), ("hdr_off", P), ("sigma", P), ("sig_off", P),
                ("diag", DBL)]


#: ``hm_fused_plan.flags``: every element an axis-aligned box (``HM_FUSED_ALIGNED``)
FUSED_ALIGNED = 1

#: int32 words of a row header (``HM_ROWHEAD``)
ROWHEAD = 12


_LIB = None


def lib():
    """The loaded library, its signatures declared.  Raises when it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("the native library %s is not built; run "
                           "`python -m hippymfem_amd.build`" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "hm_last_error": (ctypes.c_char_p, []),
        "hm_version": (ctypes.c_int, []),
        "hm_element_matrices": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, I64, I64,
                                               ctypes.c_int, P, P]),
        "hm_element_matrices_pair": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                    ctypes.c_int, I64, I64, ctypes.c_int, P, P,
                                                    P]),
        "hm_consistent_block": (ctypes.c_int, [P, P, P, I64, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, P, P]),
        "hm_element_vectors": (ctypes.c_int, [P, ctypes.c_int, I64, I64, P, P]),
        "hm_element_hvp": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, P, I64, I64, P, P]),
        "hm_element_scalars": (ctypes.c_int, [P, I64, I64, P, P]),
        "hm_point_fields": (ctypes.c_int, [P, I64, I64, P, P]),
        "hm_element_matrices_pointwise": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, P, I64,
                                                         I64, ctypes.c_int, P, P]),
        "hm_element_vectors_pointwise": (ctypes.c_int, [P, ctypes.c_int, P, I64, I64, P, P]),
        "hm_pattern_build": (ctypes.c_int, [P, P, I64, ctypes.c_int, ctypes.c_int, I64, I64,
                                            ctypes.c_int, ctypes.c_int, P, P, P, P, P, P]),
        "hm_pattern_place": (ctypes.c_int, [P, I64, ctypes.c_int, ctypes.c_int, I64, P, P, P,
                                            P, ctypes.c_int, P, P]),
        "hm_pattern_repeats": (ctypes.c_int, [I64, ctypes.c_int, ctypes.c_int, P,
                                              ctypes.c_int, P]),
        "hm_pattern_work_bytes": (I64, [I64]),
        "hm_pattern_build_device": (ctypes.c_int, [P, P, I64, ctypes.c_int, ctypes.c_int, I64,
                                                   I64, ctypes.c_int, P, P, P, P, P, P, P]),
        "hm_pattern_place_device": (ctypes.c_int, [P, P, I64, ctypes.c_int, ctypes.c_int, I64,
                                                   P, P, P, P]),
        "hm_scatter_matrix": (ctypes.c_int, [P, P, ctypes.c_int, I64, I64, P, ctypes.c_int,
                                             ctypes.c_int, P]),
        "hm_scatter_matrix_defect": (ctypes.c_int, [P, P, ctypes.c_int, I64, I64, P,
                                                    ctypes.c_int, P, P]),
        "hm_scatter_vector": (ctypes.c_int, [I64, ctypes.c_int, P, P, P, I64, I64, P, P,
                                             ctypes.c_int, P]),
        "hm_csr_matvec": (ctypes.c_int, [I64, P, P, P, ctypes.c_int, P, P, P]),
        "hm_fused_tile": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P]),
        "hm_fused_blocks": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, I32, P]),
        "hm_assemble_fused": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, P, ctypes.c_int, P,
                                             ctypes.c_int, P]),
        "hm_serving_normal": (ctypes.c_int, [P, I64, I64, P, P]),
        "hm_serving_field": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, P, P, P, P]),
        "hm_serving_field_noalloc": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, P, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def check(rc, what=""):
    """Raise the Python error of a library return code (the reference's error types)."""
    if rc == HM_OK:
        return
    msg = lib().hm_last_error().decode(errors="replace")
    text = "%s: %s" % (what, msg) if what else msg
    if rc in (HM_ERR_ARG, HM_ERR_ROW_LONG):
        raise ValueError(text)
    if rc in (HM_ERR_DUP, HM_ERR_UNSUPPORTED):
        raise NotImplementedError(text)
    raise RuntimeError(text)


def ptr(t):
    """Address of a torch tensor or numpy array (``None`` gives NULL)."""
    if t is None:
        return None
    if isinstance(t, torch.Tensor):
        return ctypes.c_void_p(t.data_ptr())
    return ctypes.c_void_p(t.ctypes.data)


def stream_handle(stream=None):
    """The ``hipStream_t`` of a torch stream (the current one by default)."""
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def require_device():
    """The product path runs on the GPU only: say so when there is none."""
    if not torch.cuda.is_available():
        raise RuntimeError("hippymfem_amd assembles on a GPU (gfx950); none is visible")
    return torch.device("cuda", torch.cuda.current_device())
