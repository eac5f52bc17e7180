"""HIP path with the semi-analytic wall (2D, reference terms) against the FP64
oracle, device 0.

Tolerances: gamma, grad gamma, the face terms, the boundary extrapolation
and post-integration run in FP64 on both sides (fixed-particle densities and
the post-integrated state agree to 1e-12, the free-surface classification
exactly); the pair sums run in FP32 with Kahan accumulation as in the
wall-free tests, so the right-hand sides are compared at 2e-5 of the field's
largest value, and 20-step runs at the bars of the wall-free step test.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import titsolver_amd as ta

pytestmark = pytest.mark.gpu

KERNELS = [ta.KERNEL_WENDLAND_C2, ta.KERNEL_WENDLAND_C4]


def _case(dr, kernel, seed=None):
    """Dam break in a closed 1.0 x 0.8 pool with one wall particle per wall
    vertex; with a seed, jittered positions, random velocities and densities
    so that particles approach the walls."""
    pa, walls = ta.dam_break_walls(dr, [1.0, 0.8], [0.4, 0.5])
    eq = ta.cfg_equations(2, dr, 0.5, kernel=kernel, flavour=ta.FLAVOUR_REF)
    eq.h, eq.mu, eq.cfl = 2.0 * dr, 1e-3, 0.4
    if seed is not None:
        rng = np.random.default_rng(seed)
        nf = pa.n_fluid
        pa.r[:nf] += rng.uniform(-0.1, 0.1, size=(nf, 2)) * dr
        pa.v = np.zeros_like(pa.r)
        pa.v[:nf] = rng.normal(0.0, 0.3, size=(nf, 2))
        pa.rho[:nf] *= 1.0 + rng.uniform(-2e-3, 2e-3, size=nf)
    return pa, walls, eq


def _oracle(pa, walls, eq):
    prm = oracle.params(eq, walls)
    st = oracle.State.from_particles(pa)
    oracle.initialize(prm, st)
    return prm, st


@pytest.mark.parametrize("kernel", KERNELS)
def test_rhs_with_walls_vs_oracle(kernel):
    dr = 0.02
    pa, walls, eq = _case(dr, kernel, seed=11)
    ctx = ta.HipContext(eq, walls=walls)
    ctx.upload([pa])
    ctx.prepare()
    st_g = ctx.download()
    drho_g, dv_g = ctx.compute_rhs()

    prm, st = _oracle(pa, walls, eq)
    off, idx = oracle.adjacency(st, prm)
    oracle.compute_gamma(prm, st)
    oracle.setup_boundary(prm, st, off, idx)
    nf = pa.n_fluid
    np.testing.assert_allclose(st_g["rho"][nf:], st.rho[nf:], rtol=1e-12, atol=0)
    oracle.compute_continuity(prm, st, off, idx)
    oracle.compute_momentum(prm, st, off, idx)
    near = np.abs(st.gamma[:nf] - 1.0) > 1e-12
    scale_rho = np.abs(st.drho_dt[:nf]).max()
    scale_v = np.abs(st.dv_dt[:nf]).max()
    err_rho = np.abs(drho_g[:nf] - st.drho_dt[:nf]).max() / scale_rho
    err_v = np.abs(dv_g[:nf] - st.dv_dt[:nf]).max() / scale_v
    print(f"kernel={kernel}: {near.sum()} of {nf} fluid particles near a wall; "
          f"drho err {err_rho:.2e} dv err {err_v:.2e}")
    assert near.sum() > 50
    assert err_rho < 2e-5
    assert err_v < 2e-5


def test_steps_with_walls_vs_oracle():
    """20 SSPRK3 steps of the dam break in the closed pool (Wendland C4,
    h = 2 dr, as titwcsph runs it)."""
    dr = 0.02
    pa, walls, eq = _case(dr, ta.KERNEL_WENDLAND_C4)
    ctx = ta.HipContext(eq, walls=walls)
    ctx.upload([pa])
    prm, st = _oracle(pa, walls, eq)
    nf = pa.n_fluid
    for s in range(20):
        dt_g = ctx.step(1)
        dt_o = oracle.ssprk_step(prm, st, 3)
        assert abs(dt_g - dt_o) <= 1e-7 * dt_o, (s, dt_g, dt_o)
    out = ctx.download()
    rho_err = np.max(np.abs(out["rho"][:nf] - st.rho[:nf]) / st.rho[:nf])
    r_err = np.max(np.abs(out["r"][:nf] - st.r[:nf])) / dr
    print(f"20 steps with walls: rho rel {rho_err:.2e}, r/dr {r_err:.2e}")
    assert rho_err < 1e-6
    assert r_err < 1e-5
    np.testing.assert_allclose(out["rho"][nf:], st.rho[nf:], rtol=1e-6, atol=0)


def test_walls_graph_equals_eager():
    dr = 0.025
    pa, walls, eq = _case(dr, ta.KERNEL_WENDLAND_C2, seed=5)
    res = []
    for eager in (False, True):
        ctx = ta.HipContext(eq, walls=walls)
        ctx.upload([pa])
        ctx.set_profiling(eager)
        ctx.step(6)
        res.append(ctx.download())
        ctx.close()
    for k in ("r", "v", "rho"):
        assert np.array_equal(res[0][k], res[1][k]), k


def test_post_integrate_with_walls_matches_oracle():
    """FluidEquations::post_integrate with the semi-analytic wall: the
    free-surface classification identical to the oracle's, r, v, rho to
    1e-12 (FP64 on both sides)."""
    dr = 0.02
    pa, walls, eq = _case(dr, ta.KERNEL_WENDLAND_C4, seed=4)
    ctx = ta.HipContext(eq, post_integrate=True, walls=walls)
    ctx.upload([pa])
    phi = ctx.post_integrate()
    got = ctx.download()
    prm, st = _oracle(pa, walls, eq)
    off, idx = oracle.adjacency(st, prm)
    oracle.compute_gamma(prm, st)
    oracle.setup_boundary(prm, st, off, idx)
    ref = oracle.post_integrate(prm, st, off, idx)
    nf = pa.n_fluid
    tiny = np.finfo(np.float64).tiny
    near = np.abs(st.gamma[:nf] - 1.0) > 1e-12
    shifted = np.any(ref["dr"][:nf] != 0.0, axis=1)
    print(f"{near.sum()} near a wall, {(ref['phi'][:nf] == tiny).sum()} on the surface, "
          f"{shifted.sum()} shifted, {(shifted & near).sum()} shifted near a wall")
    assert (shifted & near).sum() > 10
    assert np.array_equal(phi[:nf] == tiny, ref["phi"][:nf] == tiny)
    assert np.array_equal(phi[:nf] == 1.0, ref["phi"][:nf] == 1.0)
    for k, want in (("r", st.r), ("v", st.v)):
        err = np.abs(got[k][:nf] - want[:nf]).max() / max(np.abs(want[:nf]).max(), 1e-30)
        assert err < 1e-12, (k, err)
    assert np.abs(got["rho"][:nf] / st.rho[:nf] - 1.0).max() < 1e-12


def test_steps_with_walls_and_post_integrate_vs_oracle():
    """20 SSPRK3 steps, each ending with post_integrate, in the closed pool:
    the reference's full step on titwcsph's set-up."""
    dr = 0.02
    pa, walls, eq = _case(dr, ta.KERNEL_WENDLAND_C4)
    ctx = ta.HipContext(eq, post_integrate=True, walls=walls)
    ctx.upload([pa])
    prm, st = _oracle(pa, walls, eq)
    prm.post_integrate = 1
    nf = pa.n_fluid
    for s in range(20):
        dt_g = ctx.step(1)
        dt_o = oracle.ssprk_step(prm, st, 3)
        assert abs(dt_g - dt_o) <= 1e-7 * dt_o, (s, dt_g, dt_o)
    out = ctx.download()
    rho_err = np.max(np.abs(out["rho"][:nf] - st.rho[:nf]) / st.rho[:nf])
    r_err = np.max(np.abs(out["r"][:nf] - st.r[:nf])) / dr
    print(f"20 steps with walls and post_integrate: rho rel {rho_err:.2e}, r/dr {r_err:.2e}")
    assert rho_err < 1e-6
    assert r_err < 1e-5


def test_wall_arguments_rejected():
    dr = 0.05
    pa, walls, eq = _case(dr, ta.KERNEL_WENDLAND_C4)
    cfg = ta.cfg_equations(2, dr, 0.5)
    with pytest.raises(ta.TitError, match="reference flavour"):
        ta.HipContext(cfg, walls=walls)
    ctx = ta.HipContext(eq, walls=walls)
    moved = pa.copy()
    moved.r[pa.n_fluid] += 1e-3
    with pytest.raises(ta.TitError, match="vertex"):
        ctx.upload([moved])
    ctx.upload([pa])
    with pytest.raises(ta.TitError, match="without walls"):
        ta.LearnedStep(ctx)
