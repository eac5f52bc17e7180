"""Per-particle masses (ParticleArray's m, source/tit/sph/fluid_equations.hpp:41-48)
on the device against the FP64 oracle, device 0.

The masses enter the uploaded host view (tit_hip_host_view::m).  Bars as in
the uniform-mass tests: right-hand sides at 2e-5 of the field's largest value
(FP32 pair sums, Kahan-compensated), the boundary extrapolation and
post-integration at 1e-12 (FP64 on both sides), 20-step runs at 1e-6 in
density and 1e-5 dr in position.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import titsolver_amd as ta
from helpers import perturbed_dam_break

pytestmark = pytest.mark.gpu

KERNELS = [ta.KERNEL_WENDLAND_C2, ta.KERNEL_WENDLAND_C4]
FLAVOURS = [ta.FLAVOUR_CFG, ta.FLAVOUR_REF]


def _eq(dim, dr, kernel=ta.KERNEL_WENDLAND_C2, flavour=ta.FLAVOUR_CFG):
    eq = ta.cfg_equations(dim, dr, 0.6, kernel=kernel, flavour=flavour)
    if flavour == ta.FLAVOUR_REF:
        eq.mu, eq.cfl = 1e-3, 0.4
    return eq


def _with_masses(pa, eq, seed, lo=0.6, hi=1.4):
    rng = np.random.default_rng(seed)
    pa = pa.copy()
    pa.m = eq.m * rng.uniform(lo, hi, size=pa.n)
    return pa


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("flavour", FLAVOURS)
def test_rhs_with_particle_masses_vs_oracle(kernel, flavour):
    dr = 0.02
    eq = _eq(2, dr, kernel, flavour)
    pa = _with_masses(perturbed_dam_break(2, dr, seed=7), eq, seed=17)
    ctx = ta.HipContext(eq)
    ctx.upload([pa])
    ctx.prepare()
    st_g = ctx.download()
    drho_g, dv_g = ctx.compute_rhs()

    prm = oracle.params(eq)
    st = oracle.State.from_particles(pa)
    off, idx = oracle.adjacency(st, prm)
    oracle.setup_boundary(prm, st, off, idx)
    nf = pa.n_fluid
    np.testing.assert_allclose(st_g["rho"][nf:], st.rho[nf:], rtol=1e-12, atol=0)
    oracle.compute_continuity(prm, st, off, idx)
    oracle.compute_momentum(prm, st, off, idx)
    err_rho = np.abs(drho_g[:nf] - st.drho_dt[:nf]).max() / np.abs(st.drho_dt[:nf]).max()
    err_v = np.abs(dv_g[:nf] - st.dv_dt[:nf]).max() / np.abs(st.dv_dt[:nf]).max()
    # the same state with uniform masses differs by far more than the bar
    uni = oracle.State.from_particles(perturbed_dam_break(2, dr, seed=7))
    oracle.setup_boundary(prm, uni, off, idx)
    oracle.compute_continuity(prm, uni, off, idx)
    gap = np.abs(uni.drho_dt[:nf] - st.drho_dt[:nf]).max() / np.abs(st.drho_dt[:nf]).max()
    print(f"kernel={kernel} flavour={flavour}: drho err {err_rho:.2e} dv err {err_v:.2e} "
          f"(uniform masses off by {gap:.2e})")
    assert gap > 1e-2
    assert err_rho < 2e-5
    assert err_v < 2e-5


@pytest.mark.parametrize("flavour", FLAVOURS)
def test_steps_with_particle_masses_vs_oracle(flavour):
    """20 SSPRK3 steps of the 2D dam break (dr = 0.02) with fluid masses
    varying by +-20 % and the wall particles' by +-40 %."""
    dr = 0.02
    eq = _eq(2, dr, flavour=flavour)
    pa = ta.dam_break(2, dr, [3.22, 2.4], [1.0, 2.0])
    rng = np.random.default_rng(23)
    pa.m = np.full(pa.n, eq.m)
    pa.m[: pa.n_fluid] *= rng.uniform(0.8, 1.2, size=pa.n_fluid)
    pa.m[pa.n_fluid:] *= rng.uniform(0.6, 1.4, size=pa.n_fixed)
    ctx = ta.HipContext(eq)
    ctx.upload([pa])
    prm = oracle.params(eq)
    st = oracle.State.from_particles(pa)
    nf = pa.n_fluid
    for s in range(20):
        dt_g = ctx.step(1)
        dt_o = oracle.ssprk_step(prm, st, 3)
        assert abs(dt_g - dt_o) <= 1e-7 * dt_o, (s, dt_g, dt_o)
    out = ctx.download()
    rho_err = np.max(np.abs(out["rho"][:nf] - st.rho[:nf]) / st.rho[:nf])
    r_err = np.max(np.abs(out["r"][:nf] - st.r[:nf])) / dr
    print(f"flavour={flavour}, 20 steps: rho rel {rho_err:.2e}, r/dr {r_err:.2e}")
    assert rho_err < 1e-6
    assert r_err < 1e-5


def test_post_integrate_with_particle_masses_vs_oracle():
    dr = 0.02
    eq = _eq(2, dr, ta.KERNEL_WENDLAND_C4, ta.FLAVOUR_REF)
    pa = _with_masses(perturbed_dam_break(2, dr, seed=7, jitter=0.1), eq, seed=29)
    ctx = ta.HipContext(eq, post_integrate=True)
    ctx.upload([pa])
    phi = ctx.post_integrate()
    got = ctx.download()
    prm = oracle.params(eq)
    st = oracle.State.from_particles(pa)
    off, idx = oracle.adjacency(st, prm)
    oracle.setup_boundary(prm, st, off, idx)
    ref = oracle.post_integrate(prm, st, off, idx)
    nf = pa.n_fluid
    tiny = np.finfo(np.float64).tiny
    assert np.array_equal(phi[:nf] == tiny, ref["phi"][:nf] == tiny)
    assert np.array_equal(phi[:nf] == 1.0, ref["phi"][:nf] == 1.0)
    r_err = np.abs(got["r"][:nf] - st.r[:nf]).max() / dr
    rho_err = np.abs(got["rho"][:nf] / st.rho[:nf] - 1.0).max()
    print(f"post_integrate with masses: r/dr {r_err:.2e}, rho {rho_err:.2e}")
    assert r_err < 1e-12 and rho_err < 1e-12


@pytest.mark.parametrize("kernel", KERNELS)
def test_rhs_with_walls_and_particle_masses_vs_oracle(kernel):
    """Semi-analytic wall: the fixed particles' masses are their uploaded
    masses times gamma (fluid_equations.hpp:87-88)."""
    dr = 0.02
    pa, walls = ta.dam_break_walls(dr, [1.0, 0.8], [0.4, 0.5])
    eq = ta.cfg_equations(2, dr, 0.5, kernel=kernel, flavour=ta.FLAVOUR_REF)
    eq.h, eq.mu, eq.cfl = 2.0 * dr, 1e-3, 0.4
    rng = np.random.default_rng(31)
    nf = pa.n_fluid
    pa.r[:nf] += rng.uniform(-0.1, 0.1, size=(nf, 2)) * dr
    pa.v = np.zeros_like(pa.r)
    pa.v[:nf] = rng.normal(0.0, 0.3, size=(nf, 2))
    pa.m = eq.m * rng.uniform(0.7, 1.3, size=pa.n)
    ctx = ta.HipContext(eq, walls=walls)
    ctx.upload([pa])
    ctx.prepare()
    st_g = ctx.download()
    drho_g, dv_g = ctx.compute_rhs()
    prm = oracle.params(eq, walls)
    st = oracle.State.from_particles(pa)
    oracle.initialize(prm, st)
    off, idx = oracle.adjacency(st, prm)
    oracle.setup_boundary(prm, st, off, idx)
    np.testing.assert_allclose(st_g["rho"][nf:], st.rho[nf:], rtol=1e-12, atol=0)
    oracle.compute_continuity(prm, st, off, idx)
    oracle.compute_momentum(prm, st, off, idx)
    err_rho = np.abs(drho_g[:nf] - st.drho_dt[:nf]).max() / np.abs(st.drho_dt[:nf]).max()
    err_v = np.abs(dv_g[:nf] - st.dv_dt[:nf]).max() / np.abs(st.dv_dt[:nf]).max()
    print(f"walls, kernel={kernel}: drho err {err_rho:.2e} dv err {err_v:.2e}")
    assert err_rho < 2e-5
    assert err_v < 2e-5


@pytest.mark.parametrize("flavour", FLAVOURS)
def test_uniform_masses_equal_no_masses_bitwise(flavour):
    dr = 0.02
    eq = _eq(2, dr, flavour=flavour)
    pa = perturbed_dam_break(2, dr, seed=4)
    res = []
    for m in (None, np.full(pa.n, eq.m)):
        q = pa.copy()
        q.m = m
        ctx = ta.HipContext(eq)
        ctx.upload([q])
        ctx.step(5)
        res.append(ctx.download())
        ctx.close()
    for k in ("r", "v", "rho"):
        assert np.array_equal(res[0][k], res[1][k]), k


def test_batched_tanks_with_and_without_masses():
    """A batch of a tank with per-particle masses and one without: the first
    equals its own run bit for bit; the second, whose particles take the
    uniform m on the batch's per-particle path, agrees with its own
    uniform-mass run to the FP32 pair sums."""
    dr = 0.02
    eq = _eq(2, dr)
    a = _with_masses(perturbed_dam_break(2, dr, seed=2), eq, seed=41)
    b = perturbed_dam_break(2, dr, seed=3)
    both = ta.HipContext(eq)
    both.upload([a, b])
    both.step(5)
    out = both.download()
    both.close()
    for k, (pa, sl) in enumerate([(a, slice(0, a.n)), (b, slice(a.n, a.n + b.n))]):
        one = ta.HipContext(eq)
        one.upload([pa])
        one.step(5)
        o = one.download()
        one.close()
        if k == 0:
            for f in ("r", "v", "rho"):
                assert np.array_equal(o[f], out[f][sl]), f
        else:
            nf = pa.n_fluid
            assert np.abs(o["rho"][:nf] / out["rho"][sl][:nf] - 1.0).max() < 1e-6
            assert np.abs(o["r"][:nf] - out["r"][sl][:nf]).max() / dr < 1e-5


@pytest.mark.parametrize("flavour", FLAVOURS)
def test_3d_rhs_with_particle_masses_vs_oracle(flavour):
    """3D: the neighbours' volumes m_b / rho_b come from the volume array
    (the pair record's third slot holds v_z)."""
    dr = 0.03
    eq = _eq(3, dr, flavour=flavour)
    pa = _with_masses(perturbed_dam_break(3, dr, seed=5), eq, seed=19)
    ctx = ta.HipContext(eq)
    ctx.upload([pa])
    ctx.prepare()
    st_g = ctx.download()
    drho_g, dv_g = ctx.compute_rhs()
    prm = oracle.params(eq)
    st = oracle.State.from_particles(pa)
    off, idx = oracle.adjacency(st, prm)
    oracle.setup_boundary(prm, st, off, idx)
    nf = pa.n_fluid
    np.testing.assert_allclose(st_g["rho"][nf:], st.rho[nf:], rtol=1e-12, atol=0)
    oracle.compute_continuity(prm, st, off, idx)
    oracle.compute_momentum(prm, st, off, idx)
    err_rho = np.abs(drho_g[:nf] - st.drho_dt[:nf]).max() / np.abs(st.drho_dt[:nf]).max()
    err_v = np.abs(dv_g[:nf] - st.dv_dt[:nf]).max() / np.abs(st.dv_dt[:nf]).max()
    print(f"3D flavour={flavour}: drho err {err_rho:.2e} dv err {err_v:.2e}")
    assert err_rho < 2e-5
    assert err_v < 2e-5


def test_3d_steps_with_particle_masses_vs_oracle():
    """10 SSPRK3 steps of a 3D dam break (dr = 0.03, h = 1.3 dr) with masses
    varying by +-20 %, through the cell rebuilds and the LDS block kernel."""
    dr = 0.03
    eq = ta.cfg_equations(3, dr, 0.6, h_factor=1.3)
    pa = ta.dam_break(3, dr, [1.2, 1.0, 0.3 + dr], [0.45, 0.6, 0.3])
    pa = _with_masses(pa, eq, seed=37, lo=0.8, hi=1.2)
    ctx = ta.HipContext(eq)
    ctx.upload([pa])
    prm = oracle.params(eq)
    st = oracle.State.from_particles(pa)
    nf = pa.n_fluid
    for s in range(10):
        dt_g = ctx.step(1)
        dt_o = oracle.ssprk_step(prm, st, 3)
        assert abs(dt_g - dt_o) <= 1e-7 * dt_o, (s, dt_g, dt_o)
    out = ctx.download()
    rho_err = np.max(np.abs(out["rho"][:nf] - st.rho[:nf]) / st.rho[:nf])
    r_err = np.max(np.abs(out["r"][:nf] - st.r[:nf])) / dr
    print(f"3D, 10 steps: rho rel {rho_err:.2e}, r/dr {r_err:.2e}")
    assert rho_err < 1e-6
    assert r_err < 1e-5


def test_masses_rejected_by_the_learned_step_and_when_not_positive():
    eq = _eq(2, 0.02)
    pa = _with_masses(perturbed_dam_break(2, 0.02, seed=1), eq, seed=3)
    ctx = ta.HipContext(eq)
    ctx.upload([pa])
    with pytest.raises(ta.TitError, match="uniform masses"):
        ta.LearnedStep(ctx, seed=0)
    ctx.close()
    ctx = ta.HipContext(eq)
    pa.m[5] = -1.0
    with pytest.raises(ta.TitError, match="positive"):
        ctx.upload([pa])
    ctx.close()
