"""Particle output of a device run into a .ttdb file (titwcsph's loop,
wcsph.cpp:158-189): a frame at t = 0 and every 100 steps, written on a
worker thread while the device keeps stepping.  Every frame reads back equal
to the state downloaded at its step, and the run itself is unchanged by the
output (the same bits as a run without it)."""
from __future__ import annotations

import math

import numpy as np
import pytest

import titsolver_amd as ta
from titsolver_amd import ttdb

pytestmark = pytest.mark.gpu


def test_dam_break_frames_every_100_steps(tmp_path):
    dr = 0.02
    pa = ta.dam_break(2, dr, [3.22, 2.4], [1.0, 2.0])
    eq = ta.cfg_equations(2, dr, 2.0)
    H, g = 2.0, 9.81
    ctx = ta.HipContext(eq)
    ctx.upload([pa])
    st = ttdb.Storage(tmp_path / "particles.ttdb")
    st.set_max_series(1)
    writer = ttdb.FrameWriter(st.create_series())
    kept = []
    s0 = ctx.download()
    writer.submit(0.0, ttdb.particle_fields(s0, 2, eq.m, eq))
    kept.append(s0)
    for _ in range(3):
        ctx.step(100)
        t, _ = ctx.tank_clock()
        s = ctx.download()
        kept.append(s)
        writer.submit(float(t[0]) * math.sqrt(g / H), ttdb.particle_fields(s, 2, eq.m, eq))
    writer.close()
    ref = ta.HipContext(eq)
    ref.upload([pa])
    ref.step(300)
    assert np.array_equal(ref.download()["rho"], kept[-1]["rho"])
    st.close()

    ro = ttdb.Storage(tmp_path / "particles.ttdb", read_only=True)
    frames = ro.last_series().frames()
    assert len(frames) == 4
    times = [f.time for f in frames]
    assert times[0] == 0.0 and all(b > a for a, b in zip(times, times[1:]))
    for f, s in zip(frames, kept):
        assert f.array_names() == ["r", "v", "rho", "p", "m", "cs"]
        for k in ("r", "v", "rho", "p"):
            assert np.array_equal(f.read_array(k), s[k]), k
        assert f.read_array("r").shape == (pa.n, 2)
    ro.close()
