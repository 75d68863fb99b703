"""hippymfem_amd: the finite-element assembly path of hIPPyMFEM on MI355X (gfx950).

Quadrature-point residual, Jacobian and Hessian blocks (A, C, W_uu, W_um, W_mm) of a
pointwise density, contracted per element and scattered into hypre-layout CSR by HIP
kernels behind a C ABI (``include/hippymfem_amd.h``).  The Python layer mirrors the
reference's interfaces: :func:`assemble_matrix`, :func:`assemble_vector`,
:class:`QuadratureKernel`, :class:`ScatterPattern`, :class:`PDEVariationalProblem`.
"""

from ._lib import ADJOINT, PARAMETER, STATE  # noqa: F401
from .assemble import assemble_matrix, assemble_vector  # noqa: F401
from .bcs import DirichletBC, zero_essential  # noqa: F401
from .elementbatch import MeshBatches, integration_rule  # noqa: F401
from .kernel import ExpDiffusion, QuadratureKernel  # noqa: F401
from .mesh import BoxMesh, FunctionSpace, H1Space  # noqa: F401
from .pattern import ScatterPattern, VectorPattern, get_pattern  # noqa: F401

__version__ = "0.1.0"
