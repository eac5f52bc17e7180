"""Shared scene builders for the tests (inputs only; no expected values)."""
from __future__ import annotations

import numpy as np

import titsolver_amd as ta


def perturbed_dam_break(dim, dr, seed, jitter=0.15, vel=0.5, drho=2.0,
                        tank=None, column=None):
    """Dam break with jittered fluid positions, random velocities and density
    perturbations: a generic state with walls, corners and free surfaces."""
    if tank is None:
        tank = [1.2, 1.0] if dim == 2 else [0.6, 0.5, 0.3 + dr]
    if column is None:
        column = [0.5, 0.6] if dim == 2 else [0.25, 0.3, 0.3]
    pa = ta.dam_break(dim, dr, tank, column)
    rng = np.random.default_rng(seed)
    nf = pa.n_fluid
    pa.r[:nf] += rng.uniform(-jitter, jitter, size=(nf, dim)) * dr
    pa.v = np.zeros_like(pa.r)
    pa.v[:nf] = rng.normal(0.0, vel, size=(nf, dim))
    pa.rho[:nf] += rng.uniform(-drho, drho, size=nf)
    return pa


def random_cloud(dim, n, seed, box=1.0):
    """n fluid particles uniformly in [0, box]^dim (no fixed particles)."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(0.0, box, size=(n, dim))
    return ta.ParticleArray(dim, r, np.full(n, 1000.0), n, np.zeros((0, dim)),
                            tuple([0.0] * dim), tuple([box] * dim))


def lattice(dim, n_axis, dr):
    """Regular lattice dr*(i+1, ...) of n_axis^dim fluid particles: many pairs
    sit exactly on the support radius 2h when h = 2 dr."""
    g = np.meshgrid(*[np.arange(n_axis)] * dim, indexing="ij")
    r = np.stack([dr * (gi.ravel() + 1.0) for gi in g], axis=1)
    n = r.shape[0]
    return ta.ParticleArray(dim, r, np.full(n, 1000.0), n, np.zeros((0, dim)),
                            tuple([0.0] * dim), tuple([dr * (n_axis + 1)] * dim))


def csr_equal(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
