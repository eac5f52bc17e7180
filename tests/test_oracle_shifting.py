"""Pins of the oracle's post-integration step (particle shifting and the
free-surface density correction, FluidEquations::post_integrate,
source/tit/sph/fluid_equations.hpp:313-512), on the CPU.

The reference has no test of this step, so each check is a property the
method fixes, chosen so that a plausible mistake fails it; the planted
mistakes 5-8 (orc_set_mutation) show that each check catches one:
  - the renormalised gradients reproduce linear velocity and density fields
    exactly, wherever the renormalisation matrix is invertible (5: grad v
    not renormalised);
  - on a lattice block with free surfaces all round, the outer layer is
    marked free surface, particles deeper than 2h are far, and the layer in
    between gets phi = depth / 2h (6: visibility cone reversed);
  - a particle displaced from its lattice site is shifted back toward it,
    and free-surface particles are not shifted (7: every particle shifted);
  - the free-surface correction keeps a uniform block's surface density at
    rho0 (8: blend exponent with the wrong sign).
"""
import numpy as np
import pytest

import oracle
import titsolver_amd as ta

DR = 0.01


def _block(nx, ny, jitter, seed, dr=DR):
    """nx x ny lattice of fluid particles (no walls), positions jittered by
    jitter * dr."""
    g = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    r = np.stack([dr * (gi.ravel() + 1.0) for gi in g], axis=1)
    rng = np.random.default_rng(seed)
    r = r + rng.uniform(-jitter, jitter, size=r.shape) * dr
    return r, g[0].ravel(), g[1].ravel()


def _run(r, rho=None, v=None, dr=DR, flavour=ta.FLAVOUR_REF):
    eq = ta.cfg_equations(2, dr, 0.5, flavour=flavour)
    prm = oracle.params(eq)
    n = r.shape[0]
    rho = np.full(n, eq.rho0) if rho is None else rho
    st = oracle.State(r, rho, n, np.zeros((0, 2)), v)
    off, idx = oracle.adjacency(st, prm)
    out = oracle.post_integrate(prm, st, off, idx)
    return eq, st, out, (off, idx)


@pytest.fixture(autouse=True)
def _faithful():
    oracle.set_mutation(0)
    yield
    oracle.set_mutation(0)


def _linear_fields_error():
    r, _, _ = _block(24, 20, 0.2, seed=1)
    A = np.array([[0.3, -1.2], [0.7, 0.25]])
    c = np.array([0.1, -0.4])
    gvec = np.array([35.0, -80.0])
    v = r @ A.T + c
    rho = 1000.0 + r @ gvec
    _, _, out, _ = _run(r, rho=rho, v=v)
    gv, gr = out["grad_v"], out["grad_rho"]
    # gradients of particles whose renormalisation matrix was singular stay raw;
    # a block has none, so every particle is checked
    err_v = np.abs(gv - A).max() / np.abs(A).max()
    err_r = np.abs(gr - gvec).max() / np.abs(gvec).max()
    return err_v, err_r


def test_renormalised_gradients_are_exact_for_linear_fields():
    err_v, err_r = _linear_fields_error()
    assert err_v < 1e-9 and err_r < 1e-9, (err_v, err_r)


def test_planted_unrenormalised_gradient_fails():
    oracle.set_mutation(5)
    err_v, _ = _linear_fields_error()
    assert err_v > 1e-3


def _classification(jitter=0.01):
    nx, ny = 30, 26
    r, i, j = _block(nx, ny, jitter, seed=2)
    eq, _, out, _ = _run(r)
    R = 2 * eq.h
    phi = out["phi"]
    edge = (i == 0) | (i == nx - 1) | (j == 0) | (j == ny - 1)
    depth = np.minimum.reduce([i, nx - 1 - i, j, ny - 1 - j]) * DR
    return phi, edge, depth, R, i, j, nx, ny


def test_free_surface_indicator_on_a_block():
    phi, edge, depth, R, i, j, nx, ny = _classification()
    tiny = np.finfo(np.float64).tiny
    assert np.all(phi[edge] == tiny)
    assert np.all(phi[depth > R + 0.5 * DR] == 1.0)
    near = (~edge) & (depth < R - 0.5 * DR)
    assert np.all((phi[near] > 0) & (phi[near] < 1))
    # along the middle of the top face the nearest surface particle is the
    # one straight above: phi = its normal component of r_ab over 2h = depth / 2h
    mid = near & (np.abs(i - nx // 2) <= 3) & (j > ny // 2)
    assert mid.sum() >= 7
    assert np.abs(phi[mid] - depth[mid] / R).max() < 0.05


def test_planted_reversed_visibility_cone_fails():
    oracle.set_mutation(6)
    phi, edge, _, _, _, _, _, _ = _classification()
    assert not np.all(phi[edge] == np.finfo(np.float64).tiny)


def _displaced_particle():
    nx, ny = 26, 26
    r, i, j = _block(nx, ny, 1e-3, seed=3)
    k = int(np.flatnonzero((i == nx // 2) & (j == ny // 2))[0])
    site = r[k].copy()
    delta = np.array([0.1, 0.05]) * DR
    r[k] += delta
    eq, st, out, _ = _run(r)
    return st, out, k, site, delta


def test_shift_moves_a_displaced_particle_back_and_spares_the_surface():
    st, out, k, site, delta = _displaced_particle()
    dr = out["dr"]
    assert out["phi"][k] == 1.0
    assert dr[k] @ delta < 0
    assert np.linalg.norm(st.r[k] - site) < np.linalg.norm(delta)
    surf = out["phi"] != 1.0
    assert surf.sum() > 0
    assert np.all(dr[surf] == 0.0)


def test_planted_shift_of_every_particle_fails():
    oracle.set_mutation(7)
    _, out, _, _, _ = _displaced_particle()
    surf = out["phi"] != 1.0
    assert np.any(out["dr"][surf] != 0.0)


def _surface_density():
    r, i, j = _block(28, 24, 0.01, seed=4)
    eq, st, out, _ = _run(r)
    return eq, st, out


def test_free_surface_correction_keeps_rho0_on_a_uniform_block():
    eq, st, out = _surface_density()
    surf = out["phi"] != 1.0
    # a uniform block's kernel deficit alpha at the surface is far below 0.97,
    # where the blend weight beta = exp(-K (alpha - 1)^2) < 2e-12: the
    # corrected density m sum W / alpha is rho0 itself
    assert np.abs(st.rho[surf] / eq.rho0 - 1.0).max() < 1e-9
    assert np.all(st.rho[~surf] == eq.rho0)  # uniform: grad rho = 0, no change


def test_planted_blend_sign_fails():
    oracle.set_mutation(8)
    eq, st, out = _surface_density()
    surf = out["phi"] != 1.0
    # the faithful check (every surface density within 1e-9 of rho0) fails
    assert not np.all(np.abs(st.rho[surf] / eq.rho0 - 1.0) < 1e-9)


def test_post_integrate_runs_inside_the_step():
    """With post_integrate on, a step equals the plain step followed by
    prepare + post_integrate on the result."""
    pa = ta.dam_break(2, 0.05, [1.2, 1.0], [0.5, 0.6])
    eq = ta.cfg_equations(2, 0.05, 0.6, flavour=ta.FLAVOUR_REF)
    eq.mu = 1e-3
    prm = oracle.params(eq)
    a = oracle.State.from_particles(pa)
    b = oracle.State.from_particles(pa)
    for _ in range(3):
        oracle.ssprk_step(prm, a)
    prm.post_integrate = 1
    oracle.ssprk_step(prm, a)
    prm.post_integrate = 0
    for _ in range(3):
        oracle.ssprk_step(prm, b)
    oracle.ssprk_step(prm, b)
    off, idx = oracle.adjacency(b, prm)
    oracle.setup_boundary(prm, b, off, idx)
    oracle.post_integrate(prm, b, off, idx)
    for k in ("r", "v", "rho"):
        assert np.array_equal(getattr(a, k)[: pa.n_fluid], getattr(b, k)[: pa.n_fluid]), k
