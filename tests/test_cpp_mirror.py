"""The C++ host mirror (include/tit_amd/sph.hpp): titwcsph's dam break written
against FluidEquations / ParticleArray / SSPRKIntegrator (tests/cpp/
dam_break_mirror.cpp), stepped on the GPU and compared with the FP64 oracle
from the same initial state, at the bars of the Python-side step tests (step
size 1e-7, density 1e-6, position 1e-5 dr after 20 steps)."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
import titsolver_amd as ta

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "tests" / "cpp" / "build" / "dam_break_mirror"


def _exe():
    # make rebuilds the program when the headers or the library changed (a
    # program built against an older tit_hip_host_view would misread it)
    subprocess.run(["make", "-s", "-C", str(ROOT / "tests" / "cpp")], check=True)
    return str(EXE)


def test_mirror_selftest_and_error_behaviour(tmp_path):
    """Host-only checks of the mirror; and where no device is visible, a step
    fails with the mirror's Exception carrying TIT_E_HIP (exit 3)."""
    r = subprocess.run([_exe(), "selftest"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "selftest ok" in r.stdout, r.stdout + r.stderr
    r = subprocess.run([_exe(), "walls", "0.05", "1", str(tmp_path / "w.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode in (0, 3), r.stdout + r.stderr
    if r.returncode == 3:
        assert "Exception (status 2): tit_hip:" in r.stderr, r.stderr


def _load(path):
    raw = Path(path).read_bytes()
    dim, n, nf, nx, steps = (int(x) for x in np.frombuffer(raw[:40], dtype=np.int64))
    f = np.frombuffer(raw[40:], dtype=np.float64)
    pos = [2]

    def take(k, shape=None):
        a = f[pos[0]:pos[0] + k]
        pos[0] += k
        return a.reshape(shape) if shape else a.copy()

    out = {"dim": dim, "n": n, "nf": nf, "steps": steps, "h": f[0], "m": f[1]}
    out["r0"], out["v0"], out["rho0"] = take(n * dim, (n, dim)), take(n * dim, (n, dim)), take(n)
    out["normal"] = take(nx * dim, (nx, dim)) if nx else np.zeros((0, dim))
    out["r1"], out["v1"], out["rho1"] = take(n * dim, (n, dim)), take(n * dim, (n, dim)), take(n)
    out["dt"] = take(steps)
    assert pos[0] == f.size
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["walls", "layers"])
def test_mirror_dam_break_vs_oracle(case, tmp_path):
    dr, steps = 0.02, 20
    path = tmp_path / f"{case}.bin"
    print(f"mirror case {case}, dr {dr}, {steps} steps", flush=True)
    r = subprocess.run([_exe(), case, str(dr), str(steps), str(path)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    d = _load(path)
    nf = d["nf"]
    if case == "walls":
        eq = ta.ref_equations(2, dr, 0.5)
        walls = ta.box_walls(1.0, 0.8, dr)
        # the C++ tessellation put the wall particles where the Python one does
        assert np.array_equal(walls.verts, d["r0"][nf:])
        prm = oracle.params(eq, walls)
        prm.post_integrate = 1  # the mirror's default, as the reference's step
    else:
        eq = ta.cfg_equations(2, dr, 0.5)
        prm = oracle.params(eq)
    assert d["h"] == eq.h and d["m"] == eq.m
    st = oracle.State(d["r0"], d["rho0"], nf, d["normal"], d["v0"])
    oracle.initialize(prm, st)
    for s in range(steps):
        dt_o = oracle.ssprk_step(prm, st, 3)
        assert abs(d["dt"][s] - dt_o) <= 1e-7 * dt_o, (s, d["dt"][s], dt_o)
    rho_err = np.max(np.abs(d["rho1"][:nf] - st.rho[:nf]) / st.rho[:nf])
    r_err = np.max(np.abs(d["r1"][:nf] - st.r[:nf])) / dr
    print(f"{case}: rho rel {rho_err:.2e}, r/dr {r_err:.2e}")
    assert rho_err < 1e-6 and r_err < 1e-5
