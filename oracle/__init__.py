"""CPU oracle for the finite-element assembly path: TEST INFRASTRUCTURE ONLY.

A plain restatement, in numpy and C, of the reference's algorithm for the assembly hot
path (quadrature rule and basis tables, element geometry, the quadrature-route element
blocks, the sparsity pattern, the ``np.bincount`` scatter, folded elimination and the
residual scatter).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or as the
timed CPU baseline: the product path in ``hippymfem_amd`` never calls into it.
"""
