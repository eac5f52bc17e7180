"""HIP path vs the FP64 oracle (device 0).

Tolerances: neighbour sets are integer data and must match bit-exactly.
Boundary extrapolation and time-step selection run in FP64 on the device and
must agree to 1e-12 relative.  The pair sums run in FP32 on relative
coordinates with Kahan accumulation; their tolerance is stated per test
relative to the largest oracle value of the same field.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import titsolver_amd as ta
from helpers import csr_equal, lattice, perturbed_dam_break, random_cloud

pytestmark = pytest.mark.gpu

KERNELS = [ta.KERNEL_WENDLAND_C2, ta.KERNEL_WENDLAND_C4]
FLAVOURS = [ta.FLAVOUR_CFG, ta.FLAVOUR_REF]


def _eq(dim, dr, kernel=ta.KERNEL_WENDLAND_C2, flavour=ta.FLAVOUR_CFG, h_factor=2.0):
    eq = ta.cfg_equations(dim, dr, 0.6, kernel=kernel, h_factor=h_factor,
                          flavour=flavour)
    if flavour == ta.FLAVOUR_REF:
        eq.mu, eq.cfl = 1e-3, 0.4
    return eq


def _ctx(eq, arrays, order=3):
    ctx = ta.HipContext(eq, order=order)
    ctx.upload(arrays)
    return ctx


@pytest.mark.parametrize("dim", [2, 3])
def test_neighbors_dam_break_bit_exact(dim):
    dr = 0.02 if dim == 2 else 0.025
    pa = perturbed_dam_break(dim, dr, seed=1)
    eq = _eq(dim, dr)
    ctx = _ctx(eq, [pa])
    got = ctx.neighbors()
    want = oracle.grid_search(pa.r, 2.0 * eq.h, eq.h)
    assert csr_equal(got, want)


@pytest.mark.parametrize("dim,n_axis", [(2, 40), (3, 14)])
def test_neighbors_lattice_ties_bit_exact(dim, n_axis):
    # h = 2 dr: offsets (4, 0) etc. lie exactly on the support radius.
    dr = 0.01
    pa = lattice(dim, n_axis, dr)
    eq = _eq(dim, dr)
    ctx = _ctx(eq, [pa])
    got = ctx.neighbors()
    want = oracle.naive_search(pa.r, 2.0 * eq.h)
    assert csr_equal(got, want)
    counts = np.diff(got[0])
    assert counts.max() == (49 if dim == 2 else 257)


@pytest.mark.parametrize("dim,radius", [(2, 0.01), (2, 0.1), (3, 0.1), (3, 0.5)])
def test_neighbors_random_cloud_bit_exact(dim, radius):
    pa = random_cloud(dim, 2000, seed=123)
    eq = _eq(dim, radius / 2.0)
    eq.h = radius / 2.0
    ctx = _ctx(eq, [pa])
    got = ctx.neighbors()
    want = oracle.naive_search(pa.r, radius)
    assert csr_equal(got, want)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("flavour", FLAVOURS)
def test_prepare_and_rhs_vs_oracle(dim, kernel, flavour):
    dr = 0.02 if dim == 2 else 0.025
    pa = perturbed_dam_break(dim, dr, seed=7)
    eq = _eq(dim, dr, kernel, flavour)
    ctx = _ctx(eq, [pa])
    ctx.prepare()
    st_g = ctx.download()
    drho_g, dv_g = ctx.compute_rhs()

    prm = oracle.params(eq)
    st = oracle.State.from_particles(pa)
    off, idx = oracle.adjacency(st, prm)
    oracle.setup_boundary(prm, st, off, idx)
    nf = pa.n_fluid
    # boundary extrapolation: FP64 on both sides
    np.testing.assert_allclose(st_g["rho"][nf:], st.rho[nf:], rtol=1e-12, atol=0)
    oracle.compute_continuity(prm, st, off, idx)
    oracle.compute_momentum(prm, st, off, idx)
    # FP32 pair sums, Kahan-compensated: 2e-5 of the field's largest value.
    scale_rho = np.abs(st.drho_dt[:nf]).max()
    scale_v = np.abs(st.dv_dt[:nf]).max()
    err_rho = np.abs(drho_g[:nf] - st.drho_dt[:nf]).max() / scale_rho
    err_v = np.abs(dv_g[:nf] - st.dv_dt[:nf]).max() / scale_v
    print(f"dim={dim} kernel={kernel} flavour={flavour}: "
          f"drho err {err_rho:.2e} dv err {err_v:.2e}")
    assert err_rho < 2e-5
    assert err_v < 2e-5
    assert np.all(drho_g[nf:] == 0.0) and np.all(dv_g[nf:] == 0.0)


@pytest.mark.parametrize("flavour", FLAVOURS)
def test_time_step_vs_oracle(flavour):
    pa = perturbed_dam_break(2, 0.02, seed=3)
    rng = np.random.default_rng(5)
    pa.dv_dt = rng.normal(0.0, 50.0, size=pa.r.shape)
    eq = _eq(2, 0.02, flavour=flavour)
    ctx = _ctx(eq, [pa])
    dt_g = ctx.compute_time_step()[0]
    st = oracle.State.from_particles(pa)
    dt_o = oracle.compute_time_step(oracle.params(eq), st)
    assert abs(dt_g - dt_o) <= 1e-14 * dt_o


@pytest.mark.parametrize("order", [2, 3])
def test_dam_break_steps_vs_oracle(order):
    """50 steps of the configuration-1 dam break (dr = 0.02, 6218 particles)."""
    dr = 0.02
    pa = ta.dam_break(2, dr, [3.22, 2.4], [1.0, 2.0])
    eq = ta.cfg_equations(2, dr, 2.0)
    ctx = _ctx(eq, [pa], order=order)
    st = oracle.State.from_particles(pa)
    prm = oracle.params(eq)
    nf = pa.n_fluid
    for s in range(50):
        dt_g = ctx.step(1)
        dt_o = oracle.ssprk_step(prm, st, order)
        # the force criterion sqrt(h / |a|) takes the acceleration from FP32
        # pair terms (each rounded to ~6e-8), halved by the square root
        assert abs(dt_g - dt_o) <= 1e-7 * dt_o, (s, dt_g, dt_o)
    out = ctx.download()
    rho_err = np.max(np.abs(out["rho"][:nf] - st.rho[:nf]) / st.rho[:nf])
    p_o = oracle.pressure(prm, st.rho[:nf])
    p_err = np.max(np.abs(out["p"][:nf] - p_o)) / np.abs(p_o).max()
    r_err = np.max(np.abs(out["r"][:nf] - st.r[:nf])) / dr
    print(f"order {order}: rho rel {rho_err:.2e}, p/pmax {p_err:.2e}, r/dr {r_err:.2e}")
    assert rho_err < 1e-6
    assert p_err < 1e-4
    assert r_err < 1e-5
    # fixed particles never move
    assert np.array_equal(out["r"][nf:], pa.r[nf:])


def test_batched_tanks_equal_single_runs():
    dr = 0.02
    specs = [([1.6, 1.2], [0.4, 0.8]), ([2.0, 1.5], [0.7, 1.0]),
             ([1.2, 1.0], [0.3, 0.5])]
    arrays = [ta.dam_break(2, dr, t, c) for t, c in specs]
    eq = ta.cfg_equations(2, dr, 1.0)
    batch = _ctx(eq, arrays)
    batch.step(6)
    out_b = batch.download()
    t_b, _ = batch.tank_clock()
    o = 0
    for k, pa in enumerate(arrays):
        single = _ctx(eq, [pa])
        single.step(6)
        out_s = single.download()
        t_s, _ = single.tank_clock()
        sl = slice(o, o + pa.n)
        for f in ("r", "v", "rho"):
            assert np.array_equal(out_b[f][sl], out_s[f]), (k, f)
        assert t_b[k] == t_s[0]
        o += pa.n


def test_graph_replay_equals_eager():
    pa = ta.dam_break(2, 0.02, [3.22, 2.4], [1.0, 2.0])
    eq = ta.cfg_equations(2, 0.02, 2.0)
    a = _ctx(eq, [pa])
    a.step(8)  # graph of two steps, replayed
    b = _ctx(eq, [pa])
    b.set_profiling(True)  # eager launches with event timing
    b.step(8)
    ra, rb = a.download(), b.download()
    for f in ("r", "v", "rho"):
        assert np.array_equal(ra[f], rb[f]), f
    stats = b.kernel_stats()
    assert stats["stage"][0] == 8 * 3


def test_rebuild_interleaved_with_hooks():
    """neighbors() between graph replays leaves the trajectory unchanged: the
    hook reuses the valid cell index (or runs the pending rebuild the next
    step would run at the same positions)."""
    pa = ta.dam_break(2, 0.02, [3.22, 2.4], [1.0, 2.0])
    eq = ta.cfg_equations(2, 0.02, 2.0)
    a = _ctx(eq, [pa])
    a.step(2)
    a.neighbors()
    a.step(4)
    b = _ctx(eq, [pa])
    b.step(6)
    ra, rb = a.download(), b.download()
    for f in ("r", "v", "rho"):
        np.testing.assert_array_equal(ra[f], rb[f])


@pytest.mark.parametrize("dim", [2, 3])
def test_schedules_are_bitwise_identical(dim):
    """Any chunk length, and LDS staging on or off, gives the same bits: each
    target walks its own stencil in the same order on every path.  With a
    small LDS capacity most chunks overflow to the global-memory kernel (2D:
    column chunks; 3D: block chunks); capacity 0 is the column-chunk
    global-memory path."""
    dr = 0.02 if dim == 2 else 0.03
    pa = perturbed_dam_break(dim, dr, seed=9)
    eq = _eq(dim, dr)
    outs = []
    base = _ctx(eq, [pa]).get_schedule()
    variants = [(0, None), (1, None), (37, None), (0, 0), (64, 0)]
    variants.append((0, 60) if dim == 2 else (0, 300))
    for chunk, lds in variants:
        ctx = _ctx(eq, [pa])
        ctx.set_schedule(chunk, lds)
        ctx.step(4)
        outs.append(ctx.download())
        ctx.close()
    for o in outs[1:]:
        for f in ("r", "v", "rho"):
            assert np.array_equal(o[f], outs[0][f]), f
    assert base["chunk_targets"] >= 1 and base["occupancy"] > 1.0


def test_autotune_keeps_the_trajectory():
    """The autotuner's candidate runs are ordinary steps: after it, the state
    equals a plain run of the same number of steps, bit for bit."""
    dr = 0.02
    pa = perturbed_dam_break(2, dr, seed=3)
    eq = _eq(2, dr)
    ta.HipContext.clear_autotune_cache()
    a = _ctx(eq, [pa])
    best = a.autotune(2)
    n_a, _, steps = a.counts()
    assert not best["cached"] and steps > 0
    assert best["chunk_targets"] in (32, 48, 64, 96, 128)
    assert best["ms_per_step"] > 0
    b = _ctx(eq, [pa])
    b.step(steps)
    ga, gb = a.download(), b.download()
    for k in ("r", "v", "rho"):
        assert np.array_equal(ga[k], gb[k]), k
    # and it keeps stepping identically under the chosen schedule
    a.step(3)
    b.step(3)
    assert np.array_equal(a.download()["rho"], b.download()["rho"])
    a.close()
    b.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_autotune_cache_reuses_the_schedule(dim):
    """A second context of the same kind and size takes the first one's
    tuned schedule from the cache without timing (no steps taken); a context
    four times larger is tuned afresh.  In 3D the candidates include both
    schedule families (block chunks with LDS staging, column chunks from
    global memory)."""
    dr = 0.02 if dim == 2 else 0.03
    ta.HipContext.clear_autotune_cache()
    pa = perturbed_dam_break(dim, dr, seed=4)
    eq = _eq(dim, dr)
    a = _ctx(eq, [pa])
    first = a.autotune(2)
    assert not first["cached"]
    b = _ctx(eq, [pa])
    second = b.autotune(2)
    assert second["cached"] and b.counts()[2] == 0
    assert (second["chunk_targets"], second["lds_capacity"]) == \
        (first["chunk_targets"], first["lds_capacity"])
    assert b.get_schedule()["lds_capacity"] == first["lds_capacity"]
    big = perturbed_dam_break(dim, dr / (2.0 if dim == 2 else 1.6), seed=4)
    eq_big = _eq(dim, dr / (2.0 if dim == 2 else 1.6))
    c = _ctx(eq_big, [big])
    assert big.n >= 2 * pa.n
    assert not c.autotune(2)["cached"]
    for x in (a, b, c):
        x.close()


@pytest.mark.parametrize("dim,dr,schedule", [(2, 0.02, (0, None)), (2, 0.02, (0, 0)),
                                             (2, 0.01, (64, None)), (3, 0.025, (0, None))])
def test_stage_kernel_sums_exactly_the_neighbour_sets(dim, dr, schedule):
    """The fused stage kernel (with its per-target stencil trimming) sums every
    neighbour of the adjacency and nothing else: per-particle pair counts equal
    the neighbour-list sizes minus self, on a moved, jittered state."""
    pa = perturbed_dam_break(dim, dr, seed=11)
    eq = _eq(dim, dr)
    ctx = _ctx(eq, [pa])
    ctx.step(5)
    if schedule != (0, None):
        ctx.set_schedule(*schedule)
    counts = ctx.pair_counts().astype(np.int64)
    off, idx = ctx.neighbors()
    sizes = np.diff(off.astype(np.int64)) - 1
    fluid = np.zeros(ctx.n, dtype=bool)
    fluid[: pa.n_fluid] = True
    assert np.array_equal(counts[fluid], sizes[fluid])
    assert np.all(counts[~fluid] == 0)
    ctx.close()


@pytest.mark.parametrize("dim,n,radius", [(2, 3000, 0.06), (3, 3000, 0.15)])
def test_escaped_particles_keep_exact_sets(dim, n, radius):
    """Particles outside the tank box (here 3/8 of the cloud per axis) are
    clamped into its border cells: the index's neighbour sets and the stage
    kernel's accepted pairs stay the brute-force sets."""
    rng = np.random.default_rng(17)
    r = rng.uniform(-0.3, 1.3, size=(n, dim))
    pa = ta.ParticleArray(dim, r, np.full(n, 1000.0), n, np.zeros((0, dim)),
                          tuple([0.0] * dim), tuple([1.0] * dim))
    eq = _eq(dim, radius / 2.0)
    eq.h = radius / 2.0
    ctx = _ctx(eq, [pa])
    want = oracle.naive_search(pa.r, radius)
    assert csr_equal(ctx.neighbors(), want)
    counts = ctx.pair_counts().astype(np.int64)
    assert np.array_equal(counts, np.diff(want[0].astype(np.int64)) - 1)
    ctx.close()


def test_gapped_columns_are_cut_and_sum_exact_sets():
    """A tall tank (300 cell rows) with a fluid block at the bottom and a thin
    layer 2.6 m above it: each column's 140 particles fall into two equal
    parts of the chunk schedule, the second spanning the gap, which the
    builder cuts where it passes 92 cells.  Every particle is still summed
    over exactly its brute-force neighbour set (a lost or doubled piece would
    leave targets with no pairs or twice their pairs)."""
    dr = 0.005
    xs = (np.arange(40) + 0.5) * dr
    low = np.stack(np.meshgrid(xs, (np.arange(60) + 1.0) * dr, indexing="ij"), -1).reshape(-1, 2)
    high = np.stack(np.meshgrid(xs, 2.9 + np.arange(10) * dr, indexing="ij"), -1).reshape(-1, 2)
    r = np.concatenate([low, high])
    n = len(r)
    pa = ta.ParticleArray(2, r, np.full(n, 1000.0), n, np.zeros((0, 2)), (0.0, 0.0), (0.2, 3.0))
    eq = _eq(2, dr)
    ctx = _ctx(eq, [pa])
    assert ctx.get_schedule()["chunk_targets"] == 128
    want = oracle.naive_search(pa.r, 2.0 * eq.h)
    assert csr_equal(ctx.neighbors(), want)
    counts = ctx.pair_counts().astype(np.int64)
    assert np.array_equal(counts, np.diff(want[0].astype(np.int64)) - 1)
    ctx.close()


def test_empty_upload_rejected():
    eq = _eq(2, 0.02)
    ctx = ta.HipContext(eq)
    empty = ta.ParticleArray(2, np.zeros((0, 2)), np.zeros(0), 0, np.zeros((0, 2)),
                             (0.0, 0.0), (1.0, 1.0))
    with pytest.raises(ta.TitError, match="no particles"):
        ctx.upload([empty])
    ctx.close()


def test_cell_index_rebuilds_only_when_needed():
    """The cell index is rebuilt on the first step and then only when a tank's
    particles could leave the skin: far fewer rebuilds than steps, and the
    result stays within the RHS tolerance of a run that rebuilds on every
    step (the oracle's per-stage search)."""
    dr = 0.02
    pa = ta.dam_break(2, dr, [3.22, 2.4], [1.0, 2.0])
    eq = _eq(2, dr)
    ctx = _ctx(eq, [pa])
    ctx.step(60)
    _, rebuilds, steps = ctx.counts()
    print(f"{rebuilds} rebuilds in {steps} steps")
    assert steps == 60
    assert 1 <= rebuilds < steps
    prm = oracle.params(eq)
    st = oracle.State.from_particles(pa)
    for _ in range(60):
        oracle.ssprk_step(prm, st, 3)
    g = ctx.download()
    nf = pa.n_fluid
    assert np.max(np.abs(g["rho"][:nf] - st.rho[:nf]) / st.rho[:nf]) < 1e-6
    ctx.close()
