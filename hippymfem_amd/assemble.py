"""Public assembly API (``hippymfem/fem/assemble.py``): the drop-in boundary of the path."""

from .csrassemble import assemble_matrix_csr, assemble_vector_csr


def assemble_matrix(test_space, trial_space, groups, element_matrices, nelem,
                    test_ess=None, trial_ess=None, diag_policy="one", symmetric=None,
                    dtype=None):
    """Assemble a block from per-group element matrices (``assemble_matrix``,
    ``assemble.py:34-69``).

    Parameters are the reference's: ``element_matrices`` is a sequence of
    ``(ne, nd_test, nd_trial)`` arrays (torch device tensors or numpy) or a callable
    returning ``(group, start, stop, array)`` chunks; ``test_ess``/``trial_ess`` the
    essential dofs eliminated on each side; ``diag_policy`` ``"one"``, ``"zero"`` or
    ``"keep"``; ``symmetric=False`` keeps a same-space block's columns.  ``dtype``
    (an addition) chooses the precision of the values.  Returns a
    :class:`~hippymfem_amd.parmat.ParCSRMatrix` (hypre's diag-block layout, on the device).
    """
    return assemble_matrix_csr(test_space, trial_space, groups, element_matrices, nelem,
                               test_ess=test_ess, trial_ess=trial_ess,
                               diag_policy=diag_policy, symmetric=symmetric, dtype=dtype)


def assemble_vector(space, groups, element_vectors, nelem, ess=None, out=None):
    """Assemble a dual (residual) vector from per-group element vectors
    (``assemble_vector``, ``assemble.py:72-79``).  Returns a device tensor."""
    return assemble_vector_csr(space, groups, element_vectors, nelem, ess=ess, out=out)


def assemble_scalar(comm, element_values):
    """Sum per-element integrals into one global scalar (``assemble_scalar``,
    ``assemble.py:82-98``): each group's values summed where they are (one device reduction
    a group, so one scalar comes back rather than one value an element), the groups in
    order.  ``comm`` is the reference's communicator; on one rank the local sum is the
    global one, and ``None`` is accepted."""
    import numpy as np
    import torch

    loc = 0.0
    for v in element_values:
        if isinstance(v, torch.Tensor):
            loc += float(torch.sum(v.double()))
        else:
            loc += float(np.sum(np.asarray(v, dtype=np.float64)))
    if comm is not None and getattr(comm, "Get_size", lambda: 1)() > 1:
        raise NotImplementedError("assemble_scalar: one rank (the local sum is the global one)")
    return loc
