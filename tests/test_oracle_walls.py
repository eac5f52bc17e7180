"""Pins of the oracle's semi-analytic wall (2D): the kernel's face integrals,
gamma and grad gamma (sph/kernel.hpp:193-314, fluid_equations.hpp:168-193),
and the wall terms of the right-hand sides (fluid_equations.hpp:232-305).

The reference has no test of these, and its face integrals are generated
code that is not in its source tree, so each check is a property that the
mathematics fixes:
  - unit_flux and unit_antigrad_flux equal numerical quadrature of their
    defining integrals (f(rho) and eta M(rho)/rho^2 along the face, with the
    tail moment M itself a quadrature of s f(s));
  - gamma = 1 - sum of antigradient fluxes equals the kernel mass inside the
    domain, a direct polar area integral of W over the box (divergence
    theorem), at corners and along walls;
  - grad gamma equals a central difference of that area integral (planted
    mistake 9, face normals reversed, fails it);
  - with a uniform pressure and no gravity, the face terms cancel most of
    the truncated pair sum's push on the particles next to the walls
    (planted mistake 10, face terms dropped, fails it);
  - a short dam break with the walls keeps every particle in the box.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate

import oracle
import titsolver_amd as ta

KERNELS = [ta.KERNEL_WENDLAND_C2, ta.KERNEL_WENDLAND_C4]


@pytest.fixture(autouse=True)
def _faithful():
    oracle.set_mutation(0)
    yield
    oracle.set_mutation(0)


def _f(kernel):
    lib = oracle.load()
    return lambda q: lib.orc_kernel_unit_value(kernel, q)


def _quad_flux(kernel, eta, z0, z1):
    f = _f(kernel)
    if eta >= 2.0:
        return 0.0
    zc = math.sqrt(4.0 - eta * eta)
    lo, hi = max(z0, -zc), min(z1, zc)
    if lo >= hi:
        return 0.0
    pts = [0.0] if lo < 0.0 < hi else None
    return integrate.quad(lambda z: f(math.hypot(z, eta)), lo, hi, points=pts,
                          epsabs=1e-14, epsrel=1e-13, limit=200)[0]


def _quad_antigrad(kernel, eta, z0, z1):
    f = _f(kernel)
    if eta >= 2.0:
        return 0.0
    zc = math.sqrt(4.0 - eta * eta)
    lo, hi = max(z0, -zc), min(z1, zc)
    if lo >= hi:
        return 0.0

    def M(q):  # tail moment int_q^2 s f(s) ds
        return 0.0 if q >= 2.0 else integrate.quad(lambda s: s * f(s), q, 2.0, epsabs=1e-15,
                                                   epsrel=1e-14)[0]

    # eta M0 / rho^2 integrates to M0 atan(z / eta) exactly; the rest,
    # eta (M(rho) - M0) / rho^2, is bounded (M(q) - M0 ~ -f(0) q^2 / 2)
    M0 = M(0.0)
    pts = [0.0] if lo < 0.0 < hi else None
    rest = integrate.quad(lambda z: (M(math.hypot(z, eta)) - M0) / (z * z + eta * eta), lo, hi,
                          points=pts, epsabs=1e-14, epsrel=1e-13, limit=400)[0]
    return M0 * (math.atan2(hi, eta) - math.atan2(lo, eta)) + eta * rest


CASES = [(0.0, -1.0, 1.5), (1e-9, -0.7, 0.2), (0.05, -3.0, 3.0), (0.3, 0.1, 0.9),
         (0.8, -1.9, -0.2), (1.2, -0.4, 0.4), (1.7, -2.5, 0.3), (1.99, -1.0, 1.0),
         (2.0, -1.0, 1.0), (2.5, -1.0, 1.0), (0.5, 1.95, 2.5), (0.4, -0.5, -0.5)]


@pytest.mark.parametrize("kernel", KERNELS)
def test_unit_flux_equals_quadrature(kernel):
    for eta, z0, z1 in CASES:
        got = oracle.unit_flux(kernel, eta, z0, z1)
        want = _quad_flux(kernel, eta, z0, z1)
        assert abs(got - want) <= 1e-12 * max(1.0, abs(want)), (eta, z0, z1, got, want)


@pytest.mark.parametrize("kernel", KERNELS)
def test_unit_antigrad_flux_equals_quadrature(kernel):
    for eta, z0, z1 in CASES:
        if eta == 0.0:
            continue  # the limit eta -> 0 is checked below
        got = oracle.unit_flux(kernel, eta, z0, z1, antigrad=True)
        want = _quad_antigrad(kernel, eta, z0, z1)
        assert abs(got - want) <= 1e-10, (eta, z0, z1, got, want)
    # eta -> 0: half the point source's total flux 2 pi M0 = 1 / weight
    w = oracle.load().orc_kernel_weight(kernel, 2)
    assert abs(oracle.unit_flux(kernel, 0.0, -1.0, 1.0, antigrad=True) * w - 0.5) < 1e-14
    assert abs(oracle.unit_flux(kernel, 1e-12, -1.0, 1.0, antigrad=True) -
               oracle.unit_flux(kernel, 0.0, -1.0, 1.0, antigrad=True)) < 1e-11


# ---- gamma against the area integral of W over a box ----------------------

W_BOX, H_BOX, DR = 0.3, 0.2, 0.01


def _gamma_area(kernel, x, h):
    """int over the box [0, W] x [0, H] of W(x - y) dy, polar around x."""
    f = _f(kernel)
    w = oracle.load().orc_kernel_weight(kernel, 2)
    R = 2.0 * h

    def exit_dist(th):
        c, s = math.cos(th), math.sin(th)
        t = math.inf
        if c > 0:
            t = min(t, (W_BOX - x[0]) / c)
        if c < 0:
            t = min(t, -x[0] / c)
        if s > 0:
            t = min(t, (H_BOX - x[1]) / s)
        if s < 0:
            t = min(t, -x[1] / s)
        return t

    def radial(th):
        smax = min(2.0, exit_dist(th) / h)
        return integrate.quad(lambda q: f(q) * q, 0.0, smax, epsabs=1e-15, epsrel=1e-14)[0]

    pts = []
    for cx, cy in [(0, 0), (W_BOX, 0), (W_BOX, H_BOX), (0, H_BOX)]:
        pts.append(math.atan2(cy - x[1], cx - x[0]) % (2 * math.pi))
    for d, nrm in [(x[0], math.pi), (W_BOX - x[0], 0.0), (x[1], 1.5 * math.pi),
                   (H_BOX - x[1], 0.5 * math.pi)]:
        if d < R:
            a = math.acos(d / R)
            pts += [(nrm - a) % (2 * math.pi), (nrm + a) % (2 * math.pi)]
    pts = sorted(p for p in pts if 0.0 < p < 2 * math.pi)
    return w * integrate.quad(radial, 0.0, 2 * math.pi, points=pts, epsabs=1e-13,
                              epsrel=1e-12, limit=500)[0]


def _gamma_state(kernel, pts, h=1.3 * DR):
    walls = ta.box_walls(W_BOX, H_BOX, DR)
    eq = ta.cfg_equations(2, DR, 0.6, kernel=kernel, flavour=ta.FLAVOUR_REF)
    eq.h = h
    prm = oracle.params(eq, walls)
    pts = np.asarray(pts, dtype=np.float64)
    r = np.concatenate([pts, walls.verts])
    st = oracle.State(r, np.full(len(r), eq.rho0), len(pts), np.zeros((0, 2)))
    oracle.compute_gamma(prm, st)
    return st, h


POINTS = [(0.15, 0.1), (0.15, 0.002), (0.15, 0.009), (0.15, 0.021), (0.004, 0.006),
          (0.011, 0.013), (0.296, 0.15), (0.2913, 0.1937), (0.05, 0.0155)]


@pytest.mark.parametrize("kernel", KERNELS)
def test_gamma_equals_area_integral(kernel):
    st, h = _gamma_state(kernel, POINTS)
    for i, x in enumerate(POINTS):
        g = st.grad_gamma[i]
        ng = np.linalg.norm(g)
        xs = np.asarray(x) + (g / ng) * h * h if ng > 1e-5 else np.asarray(x)
        want = _gamma_area(kernel, xs, h)
        assert abs(st.gamma[i] - want) < 1e-9, (x, st.gamma[i], want)
    assert st.gamma[0] == 1.0  # deep inside: no face within 2h
    # the fixed particles: 1/2 along a wall, 1/4 at the corners (the tiny
    # shift along the normal moves both by about |grad gamma| h^2)
    nf = len(POINTS)
    corner = st.gamma[nf:nf + 4]
    edge = st.gamma[nf + 4:]
    assert np.all(np.abs(corner - 0.25) < 0.05) and np.all(np.abs(edge - 0.5) < 0.05)


def _grad_fd_error(kernel):
    pts = [(0.15, 0.006), (0.007, 0.012), (0.28, 0.17)]
    st, h = _gamma_state(kernel, pts)
    err = 0.0
    eps = 1e-4 * h
    for i, x in enumerate(pts):
        fd = np.zeros(2)
        for k in range(2):
            xp, xm = np.array(x, dtype=float), np.array(x, dtype=float)
            xp[k] += eps
            xm[k] -= eps
            fd[k] = (_gamma_area(kernel, xp, h) - _gamma_area(kernel, xm, h)) / (2 * eps)
        err = max(err, np.abs(st.grad_gamma[i] - fd).max() / np.abs(fd).max())
    return err


def test_grad_gamma_equals_area_gradient():
    assert _grad_fd_error(ta.KERNEL_WENDLAND_C4) < 1e-6


def test_planted_reversed_face_normal_fails():
    oracle.set_mutation(9)
    assert _grad_fd_error(ta.KERNEL_WENDLAND_C4) > 1.0


# ---- right-hand sides with walls -------------------------------------------

def _filled_box(kernel, mutation=0):
    """The box filled with a lattice at uniform density 1.001 rho0 (so a
    uniform pressure), at rest, without gravity."""
    dr = DR
    walls = ta.box_walls(W_BOX, H_BOX, dr)
    nx, ny = round(W_BOX / dr) - 1, round(H_BOX / dr) - 1
    g = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    fl = np.stack([dr * (g[0].ravel() + 1.0), dr * (g[1].ravel() + 1.0)], axis=1)
    eq = ta.cfg_equations(2, dr, 0.6, kernel=kernel, flavour=ta.FLAVOUR_REF)
    eq.h, eq.g, eq.mu = 2.0 * dr, 0.0, 1e-3
    prm = oracle.params(eq, walls)
    r = np.concatenate([fl, walls.verts])
    st = oracle.State(r, np.full(len(r), 1.001 * eq.rho0), len(fl), np.zeros((0, 2)))
    oracle.initialize(prm, st)
    off, idx = oracle.adjacency(st, prm)
    oracle.compute_gamma(prm, st)
    oracle.setup_boundary(prm, st, off, idx)
    oracle.set_mutation(mutation)
    oracle.compute_momentum(prm, st, off, idx)
    oracle.set_mutation(0)
    return st, fl


def test_face_terms_balance_a_uniform_pressure():
    st, fl = _filled_box(ta.KERNEL_WENDLAND_C4)
    st10, _ = _filled_box(ta.KERNEL_WENDLAND_C4, mutation=10)
    nf = len(fl)
    a = np.abs(st.dv_dt[:nf]).max()
    a10 = np.abs(st10.dv_dt[:nf]).max()
    print(f"uniform pressure: max |dv/dt| with face terms {a:.3e}, without {a10:.3e}")
    assert a < 0.15 * a10
    # deep inside, the pressure gradient of a uniform field vanishes
    deep = ((fl[:, 0] > 0.06) & (fl[:, 0] < W_BOX - 0.06) & (fl[:, 1] > 0.06) &
            (fl[:, 1] < H_BOX - 0.06))
    assert np.abs(st.dv_dt[:nf][deep]).max() < 1e-6 * a10


def test_dam_break_with_walls_stays_in_the_box():
    dr = 0.02
    pa, walls = ta.dam_break_walls(dr, [1.0, 0.8], [0.4, 0.4])
    eq = ta.cfg_equations(2, dr, 0.4, kernel=ta.KERNEL_WENDLAND_C4, flavour=ta.FLAVOUR_REF)
    eq.h, eq.mu = 2.0 * dr, 1e-3
    prm = oracle.params(eq, walls)
    st = oracle.State.from_particles(pa)
    oracle.initialize(prm, st)
    t = 0.0
    for _ in range(60):
        t += oracle.ssprk_step(prm, st, 3)
    nf = pa.n_fluid
    print(f"t = {t:.3f} s, y min {st.r[:nf, 1].min():.4f}, x max {st.r[:nf, 0].max():.4f}")
    assert np.all(np.isfinite(st.r[:nf]))
    assert st.r[:nf, 0].min() > 0.0 and st.r[:nf, 1].min() > 0.0
    assert st.r[:nf, 0].max() < 1.0
    assert np.abs(st.rho[:nf] / eq.rho0 - 1.0).max() < 0.05


def test_box_walls_tessellation_and_orientation():
    """geom::tessellate's rule (n = ceil(length / d_max) equal pieces per
    face, new vertices after the originals) and the pool's orientation: face
    normals cross(b - a) point into the fluid, the containment polygon winds
    once around the interior."""
    W, H, d = 1.0, 0.8, 0.03
    w = ta.box_walls(W, H, d)
    n_expect = 2 * math.ceil(W / d) + 2 * math.ceil(H / d)
    assert len(w.faces) == n_expect and len(w.verts) == n_expect
    assert np.all(w.verts[:4] == [[0.0, H], [W, H], [W, 0.0], [0.0, 0.0]])
    a, b = w.verts[w.faces[:, 0]], w.verts[w.faces[:, 1]]
    lengths = np.linalg.norm(b - a, axis=1)
    assert lengths.max() <= d + 1e-15 and lengths.min() > 0.5 * d
    assert np.all(np.bincount(w.faces.ravel(), minlength=len(w.verts)) == 2)
    ba = b - a
    nrm = np.stack([ba[:, 1], -ba[:, 0]], axis=1) / lengths[:, None]
    probe = 0.5 * (a + b) + 1e-3 * nrm
    assert np.all((probe > 0.0) & (probe < [W, H]))
    cw = oracle.Walls(w)
    inside = np.array([0.3, 0.2])
    outside = np.array([1.3, 0.2])
    lib = oracle.load()
    assert abs(lib.orc_winding(C.byref(cw.c), oracle._dp(inside)) - 1.0) < 1e-12
    assert abs(lib.orc_winding(C.byref(cw.c), oracle._dp(outside))) < 1e-12


def test_face_grid_sums_equal_a_scan_of_every_face():
    """The oracle's face grid selects the faces that a scan over every face
    with the exact predicate selects, in the same order: grad gamma equal bits."""
    rng = np.random.default_rng(3)
    pts = np.stack([rng.uniform(-0.01, W_BOX + 0.01, 400), rng.uniform(-0.01, H_BOX + 0.01, 400)],
                   axis=1)
    pts[:100, 1] = rng.uniform(0.0, 0.03, 100)
    st, h = _gamma_state(ta.KERNEL_WENDLAND_C4, pts)
    walls = ta.box_walls(W_BOX, H_BOX, DR)
    lib = oracle.load()
    R = 2.0 * h
    for i, x in enumerate(pts):
        g = np.zeros(2)
        for fa, fb in walls.faces:
            a = np.ascontiguousarray(walls.verts[fa])
            b = np.ascontiguousarray(walls.verts[fb])
            xx = np.ascontiguousarray(x)
            if lib.orc_face_intersects(oracle._dp(a), oracle._dp(b), oracle._dp(xx), R):
                g = g + oracle.face_flux(ta.KERNEL_WENDLAND_C4, a, b, xx, h)
        assert np.array_equal(g, st.grad_gamma[i]), (x, g, st.grad_gamma[i])


def _linear_fields_with_walls():
    """Every particle, the wall particles included, carries v = A r + c and
    rho = 1000 + g . r: the renormalised gradients of apply_shifts, face
    terms included, reproduce A and g wherever L is invertible."""
    dr = DR
    walls = ta.box_walls(W_BOX, H_BOX, dr)
    nx, ny = round(W_BOX / dr) - 1, round(H_BOX / dr) - 1
    g = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    rng = np.random.default_rng(8)
    fl = np.stack([dr * (g[0].ravel() + 1.0), dr * (g[1].ravel() + 1.0)], axis=1)
    fl += rng.uniform(-0.1, 0.1, size=fl.shape) * dr
    r = np.concatenate([fl, walls.verts])
    A = np.array([[0.3, -1.2], [0.7, 0.25]])
    c = np.array([0.1, -0.4])
    gvec = np.array([35.0, -80.0])
    v = r @ A.T + c
    rho = 1000.0 + r @ gvec
    eq = ta.cfg_equations(2, dr, 0.6, kernel=ta.KERNEL_WENDLAND_C4, flavour=ta.FLAVOUR_REF)
    eq.h = 2.0 * dr
    prm = oracle.params(eq, walls)
    st = oracle.State(r, rho, len(fl), np.zeros((0, 2)), v)
    oracle.initialize(prm, st)
    off, idx = oracle.adjacency(st, prm)
    oracle.compute_gamma(prm, st)
    near = np.abs(st.gamma[: len(fl)] - 1.0) > 1e-9
    out = oracle.post_integrate(prm, st, off, idx)
    nf = len(fl)
    err_v = np.abs(out["grad_v"][:nf] - A).max(axis=(1, 2)) / np.abs(A).max()
    err_r = np.abs(out["grad_rho"][:nf] - gvec).max(axis=1) / np.abs(gvec).max()
    return near, err_v, err_r


def test_apply_shifts_with_walls_exact_for_linear_fields():
    near, err_v, err_r = _linear_fields_with_walls()
    assert near.sum() > 100
    assert err_v.max() < 1e-9 and err_r.max() < 1e-9, (err_v.max(), err_r.max())


def _shift_ratio():
    """Mean shift of the wall-adjacent (gamma < 1) particles over the mean shift
    of the others, on a box filled with a jittered lattice.  The face terms of
    apply_shifts complete the truncated kernel sums at the walls, so both are
    at the level of the jitter; without them the kernel deficit pushes every
    wall-adjacent particle."""
    walls = ta.box_walls(W_BOX, H_BOX, DR)
    nx, ny = round(W_BOX / DR) - 1, round(H_BOX / DR) - 1
    g = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    rng = np.random.default_rng(5)
    fl = np.stack([DR * (g[0].ravel() + 1.0), DR * (g[1].ravel() + 1.0)], axis=1)
    fl += rng.uniform(-0.05, 0.05, size=fl.shape) * DR
    r = np.concatenate([fl, walls.verts])
    eq = ta.cfg_equations(2, DR, 0.6, kernel=ta.KERNEL_WENDLAND_C4, flavour=ta.FLAVOUR_REF)
    eq.h = 2.0 * DR
    prm = oracle.params(eq, walls)
    st = oracle.State(r, np.full(len(r), eq.rho0), len(fl), np.zeros((0, 2)))
    oracle.initialize(prm, st)
    off, idx = oracle.adjacency(st, prm)
    oracle.compute_gamma(prm, st)
    nf = len(fl)
    near = st.gamma[:nf] < 1.0 - 1e-9
    out = oracle.post_integrate(prm, st, off, idx)
    d = np.linalg.norm(out["dr"][:nf], axis=1)
    far = out["phi"][:nf] == 1.0
    assert (far & near).sum() > 100 and (far & ~near).sum() > 100
    return d[far & near].mean() / d[far & ~near].mean()


def test_shifts_at_the_walls_stay_at_the_jitter_level():
    assert _shift_ratio() < 2.0


def test_planted_dropped_shift_face_terms_fail():
    oracle.set_mutation(11)
    assert _shift_ratio() > 5.0
