"""Pins of the oracle's smoothing kernels and Tait EOS (CPU).

Kernel checks follow source/tit/sph/kernel.test.cpp: normalisation over the
support sphere and over a box (:33-91) and the gradient against a derivative
of the value (:95-122), plus the closed-form weights of the theory guide's
table (manual/.../02-kernel-functions.md:145-152).  EOS checks use identities
that the Tait law fixes: p(rho0) = 0, inverse pairs, cs^2 = dp/drho and
dH/drho = cs^2/rho.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate

import oracle

KERNELS = [oracle.KERNEL_WENDLAND_C2, oracle.KERNEL_WENDLAND_C4]
TINY = 6.0554544523933395e-06  # eps^(1/3), the reference's approx tolerance


def W(kernel, x, h):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return oracle.load().orc_kernel_value(kernel, x.size, oracle._dp(x), h)


def gradW(kernel, x, h):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    oracle.load().orc_kernel_grad(kernel, x.size, oracle._dp(x), h, oracle._dp(out))
    return out


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("h", [1.0, 0.1, 0.01])
def test_kernel_normalised_over_sphere(kernel, dim, h):
    R = oracle.load().orc_kernel_radius(kernel, h)
    assert R == 2.0 * h

    def radial(r):
        x = np.zeros(dim)
        x[0] = r
        return W(kernel, x, h) * (2 * math.pi * r if dim == 2 else 4 * math.pi * r * r)

    val, _ = integrate.quad(radial, 0.0, R, epsabs=1e-13, epsrel=1e-12)
    assert abs(val - 1.0) <= TINY


@pytest.mark.parametrize("kernel", KERNELS)
def test_kernel_normalised_over_box_2d(kernel):
    # Tensor Gauss-Legendre over [-R, R]^2: also checks W = 0 outside the
    # sphere and the symmetry of W.
    h = 0.1
    R = 2.0 * h
    xs, ws = np.polynomial.legendre.leggauss(400)
    xs, ws = xs * R, ws * R
    tot = 0.0
    for xi, wi in zip(xs, ws):
        for yj, wj in zip(xs, ws):
            tot += wi * wj * W(kernel, [xi, yj], h)
    assert abs(tot - 1.0) < 2e-4  # the kink at |x| = R limits the rule


@pytest.mark.parametrize("kernel", KERNELS)
def test_kernel_weights_closed_form(kernel):
    lib = oracle.load()
    # 1 / (area_d * int_0^2 w(q) q^(d-1) dq) computed by quadrature
    for dim in (2, 3):
        mom, _ = integrate.quad(lambda q: lib.orc_kernel_unit_value(kernel, q) * q ** (dim - 1),
                                0.0, 2.0, epsabs=1e-14, epsrel=1e-13)
        area = 2 * math.pi if dim == 2 else 4 * math.pi
        assert lib.orc_kernel_weight(kernel, dim) == pytest.approx(1.0 / (area * mom), rel=1e-12)
    table = {oracle.KERNEL_WENDLAND_C2: (7 / (4 * math.pi), 21 / (16 * math.pi)),
             oracle.KERNEL_WENDLAND_C4: (9 / (4 * math.pi), 495 / (256 * math.pi))}
    assert lib.orc_kernel_weight(kernel, 2) == pytest.approx(table[kernel][0], rel=1e-15)
    assert lib.orc_kernel_weight(kernel, 3) == pytest.approx(table[kernel][1], rel=1e-15)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("h", [1.0, 0.1, 0.01])
def test_kernel_grad_is_derivative_of_value(kernel, h):
    rng = np.random.default_rng(11)
    for dim in (2, 3):
        for _ in range(20):
            x = rng.uniform(-1.9, 1.9, size=dim) * h / math.sqrt(dim)
            g = gradW(kernel, x, h)
            eps = 1e-6 * h
            fd = np.zeros(dim)
            for k in range(dim):
                e = np.zeros(dim)
                e[k] = eps
                fd[k] = (W(kernel, x + e, h) - W(kernel, x - e, h)) / (2 * eps)
            scale = abs(gradW(kernel, np.full(dim, 0.5 * h), h)).max()
            assert np.abs(g - fd).max() <= 1e-7 * scale


@pytest.mark.parametrize("kernel", KERNELS)
def test_kernel_support_and_origin(kernel):
    lib = oracle.load()
    assert lib.orc_kernel_unit_value(kernel, 2.0) == 0.0
    assert lib.orc_kernel_unit_value(kernel, 2.5) == 0.0
    assert lib.orc_kernel_unit_deriv(kernel, 2.0) == 0.0
    assert lib.orc_kernel_unit_value(kernel, 0.0) == 1.0
    assert np.all(gradW(kernel, [0.0, 0.0], 0.3) == 0.0)  # normalize(0) = 0
    # w'(q) is the derivative of w(q)
    for q in np.linspace(0.05, 1.95, 12):
        d = 1e-6
        fd = (lib.orc_kernel_unit_value(kernel, q + d) - lib.orc_kernel_unit_value(kernel, q - d)) / (2 * d)
        assert lib.orc_kernel_unit_deriv(kernel, q) == pytest.approx(fd, abs=1e-8)


def _prm(xi=7.0, cs0=62.64, rho0=1000.0):
    p = oracle.OrcParams()
    p.rho0, p.cs0, p.xi = rho0, cs0, xi
    return p


@pytest.mark.parametrize("xi", [7.0, 3.0, 1.5])
def test_tait_identities(xi):
    lib = oracle.load()
    prm = _prm(xi)
    b = C.byref(prm)
    assert lib.orc_eos_pressure(b, 1000.0) == 0.0
    assert lib.orc_eos_potential(b, 1000.0) == 0.0
    assert lib.orc_eos_sound_speed(b, 1000.0) == 62.64
    for rho in (990.0, 999.9, 1000.3, 1007.0, 1050.0):
        p = lib.orc_eos_pressure(b, rho)
        assert lib.orc_eos_density_from_pressure(b, p) == pytest.approx(rho, rel=1e-14)
        H = lib.orc_eos_potential(b, rho)
        assert lib.orc_eos_density_from_potential(b, H) == pytest.approx(rho, rel=1e-14)
        d = 1e-4
        dp = (lib.orc_eos_pressure(b, rho + d) - lib.orc_eos_pressure(b, rho - d)) / (2 * d)
        cs = lib.orc_eos_sound_speed(b, rho)
        assert cs * cs == pytest.approx(dp, rel=1e-8)
        dH = (lib.orc_eos_potential(b, rho + d) - lib.orc_eos_potential(b, rho - d)) / (2 * d)
        assert dH == pytest.approx(cs * cs / rho, rel=1e-8)


def test_tait_pressure_array_matches_scalar():
    prm = _prm()
    rho = np.linspace(995.0, 1010.0, 17)
    p = oracle.pressure(prm, rho)
    lib = oracle.load()
    assert all(p[i] == lib.orc_eos_pressure(C.byref(prm), rho[i]) for i in range(rho.size))
