"""Pins of the oracle's per-particle masses (ParticleArray's m,
source/tit/sph/fluid_equations.hpp:41-48), CPU.

Each pair term of FluidEquations carries the neighbour's mass: m_b on the
a side and m_a on the b side (fluid_equations.hpp:257-258, 302-303), and the
boundary extrapolation and post-integration weigh neighbours by m_b / rho_b
(:153, :353-354, :500-501).  What fixes this without the reference:
  - momentum conservation: sum_a m_a dv_a (pair forces only) vanishes for any
    masses only if the a side carries m_b and the b side m_a;
  - splitting a particle into two halves of its mass, a hair apart, leaves
    every other particle's right-hand sides unchanged: the terms are linear
    in the neighbour's mass and indexed by the neighbour;
  - scaling every mass by s scales the pair parts of both right-hand sides
    by s.
The planted mistakes 14 and 15 (the target's own mass in the continuity and
momentum sums) fail the first two.
"""
import numpy as np
import pytest

import oracle
import titsolver_amd as ta
from helpers import perturbed_dam_break


def _eq(flavour, kernel=ta.KERNEL_WENDLAND_C2, dr=0.02):
    eq = ta.cfg_equations(2, dr, 0.6, kernel=kernel, flavour=flavour)
    if flavour == ta.FLAVOUR_REF:
        eq.mu = 1e-3
    return eq


def _masses(pa, eq, seed):
    rng = np.random.default_rng(seed)
    return eq.m * rng.uniform(0.6, 1.4, size=pa.n)


def _rhs(eq, r, rho, n_fluid, normal, v, mass, boundary=True):
    prm = oracle.params(eq)
    st = oracle.State(r, rho, n_fluid, normal, v, mass=mass)
    off, idx = oracle.adjacency(st, prm)
    if boundary:
        oracle.setup_boundary(prm, st, off, idx)
    oracle.compute_continuity(prm, st, off, idx)
    oracle.compute_momentum(prm, st, off, idx)
    return st


def _case(flavour, seed=3, kernel=ta.KERNEL_WENDLAND_C2):
    pa = perturbed_dam_break(2, 0.02, seed=seed)
    eq = _eq(flavour, kernel)
    return pa, eq, _masses(pa, eq, seed + 100)


def _weighted_momentum_residual(st, eq, mass):
    dv = st.dv_dt.copy()
    dv[: st.n_fluid, 1] += eq.g
    f = mass[:, None] * dv
    return np.abs(f.sum(0)).max() / np.abs(f).sum()


@pytest.mark.parametrize("flavour", [ta.FLAVOUR_CFG, ta.FLAVOUR_REF])
@pytest.mark.parametrize("kernel", [ta.KERNEL_WENDLAND_C2, ta.KERNEL_WENDLAND_C4])
def test_mass_weighted_momentum_is_conserved(flavour, kernel):
    pa, eq, m = _case(flavour, kernel=kernel)
    st = _rhs(eq, pa.r, pa.rho, pa.n_fluid, pa.normal, pa.v, m, boundary=False)
    assert _weighted_momentum_residual(st, eq, m) < 1e-13


@pytest.mark.parametrize("flavour", [ta.FLAVOUR_CFG, ta.FLAVOUR_REF])
def test_planted_mass_index_breaks_momentum_conservation(flavour):
    pa, eq, m = _case(flavour)
    oracle.set_mutation(15)
    try:
        st = _rhs(eq, pa.r, pa.rho, pa.n_fluid, pa.normal, pa.v, m, boundary=False)
    finally:
        oracle.set_mutation(0)
    assert _weighted_momentum_residual(st, eq, m) > 1e-3


def _split(pa, m, b, eps):
    """Particle b (fluid) replaced by two halves at r_b +- eps along x; the
    second half is inserted right after the fluid particles' end."""
    nf = pa.n_fluid
    r = np.insert(pa.r, nf, pa.r[b], axis=0)
    r[b, 0] += eps
    r[nf, 0] -= eps
    ins = lambda x: np.insert(x, nf, x[b], axis=0)  # noqa: E731
    mass = ins(m)
    mass[b] *= 0.5
    mass[nf] *= 0.5
    return r, ins(pa.rho), nf + 1, ins(pa.v), mass


def _split_errors(flavour, mutate=0):
    pa, eq, m = _case(flavour, seed=9)
    nf = pa.n_fluid
    # an interior fluid particle: the one nearest the column's centre
    centre = pa.r[:nf].mean(0)
    b = int(np.argmin(np.linalg.norm(pa.r[:nf] - centre, axis=1)))
    oracle.set_mutation(mutate)
    try:
        a = _rhs(eq, pa.r, pa.rho, nf, pa.normal, pa.v, m)
        r2, rho2, nf2, v2, m2 = _split(pa, m, b, 1e-9)
        s = _rhs(eq, r2, rho2, nf2, pa.normal, v2, m2)
    finally:
        oracle.set_mutation(0)
    keep = np.ones(pa.n, dtype=bool)
    keep[b] = False
    # particle i of the original is particle i (i < nf) or i + 1 (i >= nf)
    new_of = np.where(np.arange(pa.n) < nf, np.arange(pa.n), np.arange(pa.n) + 1)
    near = np.linalg.norm(pa.r - pa.r[b], axis=1) <= 2.0 * eq.h
    assert near[keep].sum() > 20
    e_rho = np.abs(s.drho_dt[new_of][keep] - a.drho_dt[keep]).max() / np.abs(a.drho_dt).max()
    e_v = np.abs(s.dv_dt[new_of][keep] - a.dv_dt[keep]).max() / np.abs(a.dv_dt).max()
    return e_rho, e_v


@pytest.mark.parametrize("flavour", [ta.FLAVOUR_CFG, ta.FLAVOUR_REF])
def test_splitting_a_particle_leaves_the_others_unchanged(flavour):
    e_rho, e_v = _split_errors(flavour)
    assert e_rho < 1e-6 and e_v < 1e-6, (e_rho, e_v)


@pytest.mark.parametrize("flavour", [ta.FLAVOUR_CFG, ta.FLAVOUR_REF])
@pytest.mark.parametrize("mutate,field", [(14, 0), (15, 1)])
def test_planted_mass_index_breaks_splitting(flavour, mutate, field):
    errs = _split_errors(flavour, mutate)
    assert errs[field] > 1e-3, errs


@pytest.mark.parametrize("flavour", [ta.FLAVOUR_CFG, ta.FLAVOUR_REF])
def test_scaling_the_masses_scales_the_pair_terms(flavour):
    pa, eq, m = _case(flavour, seed=5)
    a = _rhs(eq, pa.r, pa.rho, pa.n_fluid, pa.normal, pa.v, m, boundary=False)
    b = _rhs(eq, pa.r, pa.rho, pa.n_fluid, pa.normal, pa.v, 2.5 * m, boundary=False)
    nf = pa.n_fluid
    dva, dvb = a.dv_dt.copy(), b.dv_dt.copy()
    dva[:nf, 1] += eq.g
    dvb[:nf, 1] += eq.g
    np.testing.assert_allclose(b.drho_dt, 2.5 * a.drho_dt, rtol=0,
                               atol=1e-12 * np.abs(b.drho_dt).max())
    np.testing.assert_allclose(dvb, 2.5 * dva, rtol=0, atol=1e-12 * np.abs(dvb).max())


def test_uniform_masses_equal_the_default():
    pa, eq, _ = _case(ta.FLAVOUR_REF, seed=6)
    a = _rhs(eq, pa.r, pa.rho, pa.n_fluid, pa.normal, pa.v, None)
    b = _rhs(eq, pa.r, pa.rho, pa.n_fluid, pa.normal, pa.v, np.full(pa.n, eq.m))
    assert np.array_equal(a.drho_dt, b.drho_dt)
    assert np.array_equal(a.dv_dt, b.dv_dt)
    assert np.array_equal(a.rho, b.rho)


def test_boundary_extrapolation_weighs_by_neighbour_volume():
    """Fixed particles' densities: scaling every fluid mass leaves them
    unchanged (Shepard-normalised V_b = m_b / rho_b), doubling one fluid
    neighbour's mass moves them."""
    pa, eq, m = _case(ta.FLAVOUR_CFG, seed=8)
    a = _rhs(eq, pa.r, pa.rho, pa.n_fluid, pa.normal, pa.v, m)
    b = _rhs(eq, pa.r, pa.rho, pa.n_fluid, pa.normal, pa.v, 3.0 * m)
    nf = pa.n_fluid
    np.testing.assert_allclose(b.rho[nf:], a.rho[nf:], rtol=1e-12, atol=0)
    m2 = m.copy()
    m2[:nf] *= np.where(np.arange(nf) % 2 == 0, 4.0, 1.0)
    c = _rhs(eq, pa.r, pa.rho, pa.n_fluid, pa.normal, pa.v, m2)
    assert np.abs(c.rho[nf:] - a.rho[nf:]).max() > 1e-9 * a.rho[nf:].max()
