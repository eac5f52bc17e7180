"""The threaded CPU baseline (oracle/cpu_step.cpp) against the oracle.

The baseline runs the oracle's step with the pair scatter split into
x-slab blocks on a thread pool; per pair the arithmetic is the same, so it
may differ from the oracle only by the order of the sums: density within
1e-13 relative, positions within 1e-11 dr, dt within 1e-14, over 10 steps,
for any thread count.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import titsolver_amd as ta
from helpers import perturbed_dam_break


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("flavour", [ta.FLAVOUR_CFG, ta.FLAVOUR_REF])
@pytest.mark.parametrize("kernel", [ta.KERNEL_WENDLAND_C2, ta.KERNEL_WENDLAND_C4])
@pytest.mark.parametrize("threads", [1, 3])
def test_cpu_baseline_matches_oracle(dim, flavour, kernel, threads):
    dr = 0.03 if dim == 2 else 0.025
    pa = perturbed_dam_break(dim, dr, seed=5)
    eq = ta.cfg_equations(dim, dr, 0.6, kernel=kernel, flavour=flavour,
                          h_factor=2.0 if dim == 2 else 1.3)
    if flavour == ta.FLAVOUR_REF:
        eq.mu, eq.cfl = 1e-3, 0.4
    prm = oracle.params(eq)
    a = oracle.State.from_particles(pa)
    b = oracle.State.from_particles(pa)
    assert oracle.cpu_set_threads(threads) == threads
    for _ in range(10):
        dt_a = oracle.ssprk_step(prm, a, 3)
        dt_b = oracle.cpu_ssprk_step(prm, b, 3)
        assert abs(dt_a - dt_b) <= 1e-14 * dt_a
    assert np.max(np.abs(a.rho - b.rho) / a.rho) < 1e-13
    assert np.max(np.abs(a.r - b.r)) < 1e-11 * dr
    assert np.max(np.abs(a.v - b.v)) < 1e-9


def test_cpu_baseline_slab_colouring_covers_every_pair():
    """A tank far wider than the support (hundreds of slabs): the threaded
    scatter must visit each unique pair once, which the order-2 step from rest
    reproduces to summation rounding."""
    dr = 0.01
    pa = ta.dam_break(2, dr, [3.22, 0.5], [2.5, 0.3])
    eq = ta.cfg_equations(2, dr, 0.3)
    prm = oracle.params(eq)
    a = oracle.State.from_particles(pa)
    b = oracle.State.from_particles(pa)
    oracle.cpu_set_threads(4)
    for _ in range(3):
        oracle.ssprk_step(prm, a, 2)
        oracle.cpu_ssprk_step(prm, b, 2)
    assert np.max(np.abs(a.rho - b.rho) / a.rho) < 1e-13


def test_cpu_baseline_rejects_walls_and_post_integration():
    pa = perturbed_dam_break(2, 0.04, seed=1)
    eq = ta.cfg_equations(2, 0.04, 0.6)
    prm = oracle.params(eq)
    prm.post_integrate = 1
    with pytest.raises(RuntimeError):
        oracle.cpu_ssprk_step(prm, oracle.State.from_particles(pa), 3)
