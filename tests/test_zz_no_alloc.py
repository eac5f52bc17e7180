"""tit_hip_neighbors_noalloc (no device allocation; a static scratch array
per GPU) against tit_hip_neighbors, called through the C ABI: two calls in a
row on different contexts, the second with fewer particles, then a third on
the first context after a few steps; every result equals the allocating
entry point's bit for bit, and the sizes beyond the scratch are rejected."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import titsolver_amd as ta
from helpers import perturbed_dam_break

pytestmark = pytest.mark.gpu


def _noalloc(ctx):
    lib = ctx.lib
    u64 = C.POINTER(C.c_uint64)
    u32 = C.POINTER(C.c_uint32)
    off = np.zeros(ctx.n + 1, dtype=np.uint64)
    total = C.c_size_t()
    st = lib.tit_hip_neighbors_noalloc(ctx.h, off.ctypes.data_as(u64), C.cast(None, u32), 0,
                                       C.byref(total))
    assert st == ta._abi.TIT_E_ARG  # capacity 0: offsets and total only
    idx = np.zeros(max(total.value, 1), dtype=np.uint32)
    st = lib.tit_hip_neighbors_noalloc(ctx.h, off.ctypes.data_as(u64), idx.ctypes.data_as(u32),
                                       idx.size, C.byref(total))
    assert st == 0, lib.tit_hip_last_error(ctx.h).decode()
    return off, idx[: total.value]


def test_noalloc_neighbors_equal_the_allocating_entry_point():
    big = perturbed_dam_break(2, 0.005, seed=31)
    small = perturbed_dam_break(3, 0.03, seed=32)
    a = ta.HipContext(ta.cfg_equations(2, 0.005, 0.6))
    a.upload([big])
    b = ta.HipContext(ta.cfg_equations(3, 0.03, 0.6, h_factor=1.3))
    b.upload([small])
    assert b.n < a.n
    for ctx in (a, b):
        got = _noalloc(ctx)
        want = ctx.neighbors()
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    a.step(5)
    got = _noalloc(a)
    want = a.neighbors()
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    a.close()
    b.close()


def test_noalloc_rejects_more_particles_than_its_scratch():
    r = np.random.default_rng(0).uniform(0.0, 1.0, size=((1 << 20) + 1, 2))
    n = len(r)
    pa = ta.ParticleArray(2, r, np.full(n, 1000.0), n, np.zeros((0, 2)), (0.0, 0.0), (1.0, 1.0))
    eq = ta.cfg_equations(2, 0.002, 0.6)
    ctx = ta.HipContext(eq)
    ctx.upload([pa])
    off = np.zeros(n + 1, dtype=np.uint64)
    total = C.c_size_t()
    st = ctx.lib.tit_hip_neighbors_noalloc(ctx.h, off.ctypes.data_as(C.POINTER(C.c_uint64)),
                                           C.cast(None, C.POINTER(C.c_uint32)), 0,
                                           C.byref(total))
    assert st == ta._abi.TIT_E_ARG
    assert "1048576" in ctx.lib.tit_hip_last_error(ctx.h).decode()
    ctx.close()
