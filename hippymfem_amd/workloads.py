"""Synthetic inputs of the assembly benchmarks: meshes, spaces and seeded fields.

* ``m``: N(0,1) white noise on the ``(n+1)^3`` vertex lattice (seed 0), three passes of the
  7-point average (over the neighbours that exist), then mean 0 and std 1;
* ``u``: U[-1, 1] on every state dof (seed 1), or a smooth field, a sum of 8 seeded
  Fourier modes of amplitude 0.5 (seed 1);
* ``p``: N(0, 1) on every state dof (seed 2);  ``x`` (serving products): N(0,1) (seed 3);
* node perturbation of the interior vertices by U[-0.2h, 0.2h]^3 (seed 4);
* ``f = 1``; Dirichlet u = 0 on x = 0 and x = 1.
Random values are drawn in lattice order and then numbered, so the fields do not depend on
the dof numbering.
"""

import numpy as np

from .bcs import DirichletBC
from .elementbatch import MeshBatches
from .kernel import ExpDiffusion, QuadratureKernel
from .mesh import BoxMesh, H1Space


def smoothed_noise(n, seed=0, passes=3):
    """The parameter field on the vertex lattice, lattice order."""
    m1 = n + 1
    f = np.random.default_rng(seed).standard_normal((m1, m1, m1))
    for _ in range(passes):
        acc = f.copy()
        cnt = np.ones_like(f)
        for ax in range(3):
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[ax], hi[ax] = slice(None, -1), slice(1, None)
            acc[tuple(hi)] += f[tuple(lo)]
            cnt[tuple(hi)] += 1.0
            acc[tuple(lo)] += f[tuple(hi)]
            cnt[tuple(lo)] += 1.0
        f = acc / cnt
    f = f.reshape(-1)
    return (f - f.mean()) / f.std()


def fourier_state(coords, seed=1, modes=8, amplitude=0.5):
    """``sum_k a_k sin(2 pi kvec_k . x + phi_k)``, integer wave vectors in ``{0..3}^3`` (an
    all-zero one replaced by ones), scaled to largest magnitude ``amplitude``."""
    rng = np.random.default_rng(seed)
    kv = rng.integers(0, 4, size=(modes, 3))
    kv[np.all(kv == 0, axis=1)] = 1
    a = rng.standard_normal(modes)
    ph = rng.uniform(0.0, 2.0 * np.pi, modes)
    u = np.sin(2.0 * np.pi * coords @ kv.T + ph) @ a
    return amplitude * u / np.abs(u).max()


class Workload:
    """One configuration: mesh, spaces ``[Vu, Vm, Vu]``, batches, kernel, fields, BC."""

    def __init__(self, n, order=2, perturb=0.0, nonlinear=False, state="uniform",
                 numbering="lattice", quadrature_degree=None, f=1.0):
        self.n, self.order = int(n), int(order)
        self.mesh = BoxMesh(n, perturb=perturb)
        self.Vu = H1Space(self.mesh, order, numbering)
        self.Vm = H1Space(self.mesh, 1, "lattice" if numbering == "shuffled" else numbering)
        # degree 4: the 3x3x3 Gauss rule of the benchmarks (2 * order + 2 = 4 is also the
        # reference's default for P1, elementbatch.py:148-156)
        deg = 4 if quadrature_degree is None else quadrature_degree
        self.batches = MeshBatches(self.mesh, deg)
        self.density = ExpDiffusion(nonlinear=nonlinear, f=f)
        self.spaces = [self.Vu, self.Vm, self.Vu]
        self.kernel = QuadratureKernel(self.density, self.spaces, self.batches)
        nl = self.Vu.ndof
        self.m = self.Vm.from_lattice(smoothed_noise(self.n, seed=0))
        if state == "fourier":
            self.u = fourier_state(self.Vu.dof_coordinates(), seed=1)
        else:
            self.u = self.Vu.from_lattice(np.random.default_rng(1).uniform(-1.0, 1.0, nl))
        self.p = self.Vu.from_lattice(np.random.default_rng(2).standard_normal(nl))
        self.bc = DirichletBC(self.Vu, where="x0x1")
        self.ess = self.bc.ess_tdof

    @property
    def groups(self):
        return self.batches.groups

    def locals(self):
        return [self.u, self.m, self.p]

    def direction(self, seed=3):
        return self.Vu.from_lattice(np.random.default_rng(seed).standard_normal(self.Vu.ndof))
