"""Host-side checks (CPU): the C ABI library loads and exports every entry
point declared in include/tit_hip.h, and the dam-break set-up builds the
lattice and hydrostatic state of titwcsph (wcsph.cpp:104-142)."""
import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

import titsolver_amd as ta
from titsolver_amd import _abi

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "tit_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tit_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ta.load()
    names = _declared()
    assert len(names) >= 17
    for n in names:
        assert hasattr(lib, n), n
        assert n in _abi.SIGNATURES, f"{n} missing from the ctypes table"


def test_no_entry_point_without_declaration():
    assert set(_abi.SIGNATURES) == set(_declared())


def test_invalid_arguments_rejected_without_device():
    lib = ta.load()
    h = C.c_void_p()
    p = ta.cfg_equations(2, 0.01, 1.0).to_params()
    p.dim = 4
    assert lib.tit_hip_create(0, C.byref(p), C.byref(h)) == _abi.TIT_E_ARG
    assert not h.value
    n1, n2 = C.c_size_t(), C.c_size_t()
    tank = (C.c_double * 2)(1.0, 1.0)
    col = (C.c_double * 2)(2.0, 0.5)  # column wider than the tank
    assert lib.tit_dam_break_count(2, 0.01, tank, col, 3, C.byref(n1), C.byref(n2)) == _abi.TIT_E_ARG


@pytest.mark.parametrize("dim", [2, 3])
def test_dam_break_counts_and_lattice(dim):
    dr = 0.02
    tank = [3.22, 2.4] if dim == 2 else [1.0, 1.0, 0.52]
    col = [1.0, 2.0] if dim == 2 else [0.4, 0.6, 0.5]
    pa = ta.dam_break(dim, dr, tank, col)
    m, n = round(col[0] / dr), round(col[1] / dr)
    k = 1 if dim == 2 else round(col[2] / dr)
    assert pa.n_fluid == m * n * k
    L = 3
    nx, ny = round(tank[0] / dr), round(tank[1] / dr)
    if dim == 2:
        assert pa.n_fixed == L * (nx + 1 + 2 * (L - 1)) + 2 * L * ny
    # fluid on dr*(i+1, j+1[, k+1]), i outer (wcsph.cpp:104-109)
    r = pa.r[: pa.n_fluid]
    assert np.array_equal(r[0], dr * np.ones(dim))
    assert r[1, dim - 1] == 2 * dr and r[1, 0] == dr  # last axis innermost
    # fixed particles lie on or outside the walls, never inside the tank
    f = pa.r[pa.n_fluid:]
    inside = np.all((f > 1e-12) & (f < np.array(tank) - 1e-12), axis=1)
    assert not inside.any()
    # wall normals are unit vectors pointing into the tank
    nrm = pa.normal
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0)
    assert np.all(nrm[f[:, 1] < 0, 1] > 0)


def test_hydrostatic_initial_pressure():
    """The reference's Fourier series (wcsph.cpp:129-141) is harmonic, vanishes
    on the free surface y = H and on the free side x = L, and rho = rho0 +
    p/cs0^2."""
    dr, cs0 = 0.01, 40.0
    H, L = 0.6, 0.4
    pa = ta.dam_break(2, dr, [1.0, 1.0], [L, H], cs0=cs0)
    nf = pa.n_fluid
    r = pa.r[:nf]
    p = (pa.rho[:nf] - 1000.0) * cs0 ** 2
    M, N = round(L / dr), round(H / dr)
    P = p.reshape(M, N)
    lap = (P[2:, 1:-1] + P[:-2, 1:-1] + P[1:-1, 2:] + P[1:-1, :-2] - 4 * P[1:-1, 1:-1]) / dr ** 2
    # every term is harmonic; the 5-point stencil's O(dr^2 a^4) error is small
    # wherever the high-k terms have decayed (x <= L - 0.1)
    far = r[:, 0].reshape(M, N)[1:-1, 1:-1] <= L - 0.1 + 1e-12
    assert np.abs(lap[far]).max() < 1e-3 * 1000 * 9.81 / dr
    # top row y = H: p = 0 exactly up to rounding of cos(k pi/2)
    assert np.abs(P[:, -1]).max() < 1e-9 * 1000 * 9.81 * H
    # free side x = L: the cosine series of rho0 g (H - y), truncated at k < 100
    assert np.abs(P[-1, :]).max() < 2e-2 * 1000 * 9.81 * H
    # away from the free side the pressure grows with depth
    assert np.all(np.diff(P[0]) < 0)
