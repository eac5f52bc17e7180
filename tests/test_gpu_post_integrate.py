"""Post-integration on the device (particle shifting and the free-surface
density correction, FluidEquations::post_integrate,
source/tit/sph/fluid_equations.hpp:313-512) against the FP64 oracle.

Both sides run in FP64 and differ only in summation order, so the
free-surface indicator (a discrete classification) must match exactly and
positions, velocities and densities to 1e-12 relative.  Over steps, the
trajectory carries the FP32 pair sums of the stages: there the bars are
those of the plain step.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import titsolver_amd as ta
from helpers import perturbed_dam_break

pytestmark = pytest.mark.gpu


def _eq(dim, dr, kernel=ta.KERNEL_WENDLAND_C2, flavour=ta.FLAVOUR_REF):
    eq = ta.cfg_equations(dim, dr, 0.6, kernel=kernel, flavour=flavour)
    if flavour == ta.FLAVOUR_REF:
        eq.mu, eq.cfl = 1e-3, 0.4
    return eq


@pytest.mark.parametrize("dim,kernel", [(2, ta.KERNEL_WENDLAND_C2), (2, ta.KERNEL_WENDLAND_C4),
                                        (3, ta.KERNEL_WENDLAND_C2), (3, ta.KERNEL_WENDLAND_C4)])
def test_post_integrate_matches_oracle(dim, kernel):
    dr = 0.02 if dim == 2 else 0.03
    pa = perturbed_dam_break(dim, dr, seed=7, jitter=0.1)
    eq = _eq(dim, dr, kernel)
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
    n_surf, n_far = int((ref["phi"][:nf] == tiny).sum()), int((ref["phi"][:nf] == 1.0).sum())
    shifted = np.any(ref["dr"][:nf] != 0.0, axis=1)
    print(f"dim={dim} kernel={kernel}: {nf} fluid, {n_surf} on the surface, {n_far} far, "
          f"{shifted.sum()} shifted")
    assert n_surf > 0 and n_far > 0 and shifted.sum() > 0
    # classification: identical (equal bits for the surface and far values;
    # near-surface values are products of FP64 sums)
    assert np.array_equal(phi[:nf] == tiny, ref["phi"][:nf] == tiny)
    assert np.array_equal(phi[:nf] == 1.0, ref["phi"][:nf] == 1.0)
    near = (ref["phi"][:nf] != tiny) & (ref["phi"][:nf] != 1.0)
    assert np.allclose(phi[:nf][near], ref["phi"][:nf][near], rtol=1e-12, atol=0)
    # state after the shifts and the correction
    r_err = np.abs(got["r"][:nf] - st.r[:nf]).max() / dr
    v_err = np.abs(got["v"][:nf] - st.v[:nf]).max() / max(np.abs(st.v[:nf]).max(), 1e-30)
    rho_err = np.abs(got["rho"][:nf] / st.rho[:nf] - 1.0).max()
    print(f"  r/dr {r_err:.2e}, v {v_err:.2e}, rho {rho_err:.2e}")
    assert r_err < 1e-12 and v_err < 1e-12 and rho_err < 1e-12


def test_steps_with_post_integrate_match_oracle():
    """20 SSPRK3 steps of the 2D dam break (dr = 0.02) with post_integrate,
    reference flavour, against the oracle's step with post_integrate."""
    dr = 0.02
    pa = ta.dam_break(2, dr, [3.22, 2.4], [1.0, 2.0])
    eq = _eq(2, dr)
    ctx = ta.HipContext(eq, post_integrate=True)
    ctx.upload([pa])
    prm = oracle.params(eq)
    prm.post_integrate = 1
    st = oracle.State.from_particles(pa)
    nf = pa.n_fluid
    for s in range(20):
        dt_g = ctx.step(1)
        dt_o = oracle.ssprk_step(prm, st, 3)
        assert abs(dt_g - dt_o) <= 1e-7 * dt_o, (s, dt_g, dt_o)
    out = ctx.download()
    rho_err = np.max(np.abs(out["rho"][:nf] - st.rho[:nf]) / st.rho[:nf])
    r_err = np.max(np.abs(out["r"][:nf] - st.r[:nf])) / dr
    print(f"20 steps with post_integrate: rho rel {rho_err:.2e}, r/dr {r_err:.2e}")
    assert rho_err < 1e-6
    assert r_err < 1e-5


def test_shifts_past_the_skin_rebuild_the_index():
    """A strongly jittered start at rest (shifts of 0.04 dr, motion of
    0.002 dr per step) with a skin of half the largest shift: post-
    integration's shifts carry particles past skin/2, and the next step
    rebuilds the cell index instead of failing; 20 steps with post_integrate
    still match the oracle."""
    dr = 0.02
    pa = perturbed_dam_break(2, dr, seed=7, jitter=0.25, vel=0.0, drho=0.0)
    eq = _eq(2, dr)
    prm = oracle.params(eq)
    st0 = oracle.State.from_particles(pa)
    off, idx = oracle.adjacency(st0, prm)
    oracle.setup_boundary(prm, st0, off, idx)
    shift = np.linalg.norm(oracle.post_integrate(prm, st0, off, idx)["dr"], axis=1).max()
    skin = 0.5 * shift / eq.h  # skin/2 = skin * h = half the largest shift
    print(f"largest shift {shift / dr:.3f} dr, skin {skin:.4f} of the support radius")
    assert shift > 0.0
    ctx = ta.HipContext(eq, post_integrate=True, skin=skin)
    ctx.upload([pa])
    prm.post_integrate = 1
    st = oracle.State.from_particles(pa)
    nf = pa.n_fluid
    for s in range(20):
        dt_g = ctx.step(1)
        dt_o = oracle.ssprk_step(prm, st, 3)
        assert abs(dt_g - dt_o) <= 1e-7 * dt_o, (s, dt_g, dt_o)
    _, rebuilds, steps = ctx.counts()
    out = ctx.download()
    rho_err = np.max(np.abs(out["rho"][:nf] - st.rho[:nf]) / st.rho[:nf])
    r_err = np.max(np.abs(out["r"][:nf] - st.r[:nf])) / dr
    print(f"{rebuilds} rebuilds in {steps} steps; rho rel {rho_err:.2e}, r/dr {r_err:.2e}")
    assert rebuilds >= 2
    assert rho_err < 1e-6
    assert r_err < 1e-5


def test_batched_tanks_with_post_integrate_match_oracle():
    """Three tanks of different sizes in one context, 10 steps with
    post_integrate: each tank against the oracle run on its own (per-tank
    shifts, displacement bounds and time steps)."""
    dr = 0.02
    specs = [([1.6, 1.2], [0.4, 0.8]), ([2.0, 1.5], [0.7, 1.0]), ([1.2, 1.0], [0.3, 0.5])]
    arrays = [ta.dam_break(2, dr, t, c) for t, c in specs]
    eq = _eq(2, dr)
    ctx = ta.HipContext(eq, post_integrate=True)
    ctx.upload(arrays)
    ctx.step(10)
    out = ctx.download()
    t_g, _ = ctx.tank_clock()
    prm = oracle.params(eq)
    prm.post_integrate = 1
    o = 0
    for k, pa in enumerate(arrays):
        st = oracle.State.from_particles(pa)
        t_o = sum(oracle.ssprk_step(prm, st, 3) for _ in range(10))
        nf = pa.n_fluid
        sl = slice(o, o + nf)
        rho_err = np.max(np.abs(out["rho"][sl] - st.rho[:nf]) / st.rho[:nf])
        r_err = np.max(np.abs(out["r"][sl] - st.r[:nf])) / dr
        print(f"tank {k}: rho rel {rho_err:.2e}, r/dr {r_err:.2e}, clock {t_g[k]:.6e} vs {t_o:.6e}")
        assert rho_err < 1e-6 and r_err < 1e-5
        assert abs(t_g[k] - t_o) <= 1e-7 * t_o
        o += pa.n


def test_post_integrate_graph_equals_eager():
    """The captured two-step graph and eager launches give the same bits."""
    dr = 0.03
    pa = perturbed_dam_break(2, dr, seed=3, jitter=0.05)
    eq = _eq(2, dr)
    res = []
    for eager in (False, True):
        ctx = ta.HipContext(eq, post_integrate=True)
        ctx.upload([pa])
        ctx.set_profiling(eager)
        ctx.step(6)
        res.append(ctx.download())
        ctx.close()
    for k in ("r", "v", "rho"):
        assert np.array_equal(res[0][k], res[1][k]), k
