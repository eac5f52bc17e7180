"""GPU tests of the learned step: the BF16 MFMA GEMM against numpy, the
k-nearest graph bit-exact against the oracle, features / forward / loss /
gradients against the FP64 network oracle (oracle/surrogate.py), training,
the judge/fallback controller, and what the trained network gains."""
from __future__ import annotations

import numpy as np
import pytest

import titsolver_amd as ta
from oracle import surrogate as S

pytestmark = pytest.mark.gpu


def bf16(x):
    """Round-to-nearest-even FP32 -> BF16 (returned as FP32)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def hswish(x):
    """Hard-swish, the network's activation."""
    return x * np.clip(x / 6.0 + 0.5, 0.0, 1.0)


@pytest.mark.parametrize("M,N,K,tA,tB,bias,act,splits", [
    (64, 64, 16, False, False, False, 0, 1),
    (100, 70, 50, False, False, True, 1, 1),
    (37, 129, 80, True, False, True, 0, 1),
    (300, 80, 128, False, True, False, 0, 1),
    (128, 16, 1000, True, False, False, 0, 8),
    (160, 128, 3000, True, False, False, 0, 5),
])
def test_gemm_vs_numpy(M, N, K, tA, tB, bias, act, splits):
    rng = np.random.default_rng(M * 7 + N)
    A = rng.normal(size=(K, M) if tA else (M, K)).astype(np.float32)
    B = rng.normal(size=(N, K) if tB else (K, N)).astype(np.float32)
    b = rng.normal(size=N).astype(np.float32) if bias else None
    C0 = rng.normal(size=(M, N)).astype(np.float32) if splits > 1 else None
    got = ta.nn_gemm(A, B, tA, tB, b, act, splits, C0)
    Ar = bf16(A).astype(np.float64)
    Br = bf16(B).astype(np.float64)
    ref = (Ar.T if tA else Ar) @ (Br.T if tB else Br)
    if b is not None:
        ref = ref + b
    if act:
        ref = hswish(ref)
    if C0 is not None:
        ref = ref + C0
    scale = np.sqrt(K) * max(1.0, np.abs(ref).max() / np.sqrt(K))
    err = np.abs(got - ref).max() / scale
    print(f"gemm {M}x{N}x{K} tA={tA} tB={tB} splits={splits}: max err / scale {err:.2e}")
    assert err < 1e-5


def _dam(dr=0.05, steps=30, tank=(3.2, 2.4), column=(1.0, 2.0)):
    pa = ta.dam_break(2, dr, list(tank), list(column))
    eq = ta.cfg_equations(2, dr, column[1])
    ctx = ta.HipContext(eq)
    ctx.upload([pa])
    if steps:
        ctx.step(steps)
    return ctx, eq, pa


def _fixed_and_normals(pa):
    n = pa.n
    fixed = np.zeros(n, dtype=bool)
    fixed[pa.n_fluid:] = True
    normal = np.zeros((n, pa.dim))
    normal[pa.n_fluid:] = pa.normal
    return fixed, normal


def test_graph_bit_exact_and_features_vs_oracle():
    ctx, eq, pa = _dam()
    L = ta.LearnedStep(ctx, seed=1)
    nbr, nf, ef = L.graph()
    g = ctx.download()
    nbr_o = S.knn_graph(g["r"], eq.h)
    assert np.array_equal(nbr.astype(np.int64), nbr_o)
    sc = L.scales()
    fixed, normal = _fixed_and_normals(pa)
    nf_o = S.node_features(g["v"], g["rho"], fixed, normal, eq.rho0, sc[0], sc[1])
    ef_o = S.edge_features(g["r"], g["v"], g["rho"], nbr_o, eq.h, eq.kernel, eq.rho0,
                           sc[0], sc[1])
    # BF16 storage: relative 2^-8 plus FP32 rounding of the inputs
    assert np.all(np.abs(nf - nf_o) <= 2.0 ** -8 * np.abs(nf_o) + 1e-5)
    assert np.all(np.abs(ef - ef_o) <= 2.0 ** -8 * np.abs(ef_o) + 1e-5)
    assert ef[..., 9].sum() > 0.9 * 8 * pa.n_fluid  # most edges are real
    ctx.close()


def test_forward_vs_oracle():
    ctx, eq, pa = _dam()
    L = ta.LearnedStep(ctx, seed=3)
    w = L.params()
    out = L.forward()
    nbr, nf, ef = L.graph()
    u_o, _ = S.forward(w, nf, ef, nbr)
    rel = np.linalg.norm(out - u_o) / np.linalg.norm(u_o)
    print(f"forward: relative L2 error vs FP64 oracle {rel:.2e}")
    assert rel < 3e-2
    ctx.close()


def test_loss_and_gradient_vs_oracle():
    ctx, eq, pa = _dam()
    L = ta.LearnedStep(ctx, seed=5)
    k = L.record(4)
    smp = L.sample(k)
    loss, grad = L.loss_grad(k)
    w = L.params()
    sc = L.scales()
    loss_o, grad_o = S.loss_and_grad(w, smp["node_feat"], smp["edge_feat"], smp["nbr"],
                                     smp["target"], smp["fluid"], sc[2:6], eq.dim)
    rel_loss = abs(loss - loss_o) / loss_o
    rel = np.linalg.norm(grad - grad_o) / np.linalg.norm(grad_o)
    cos = float(grad @ grad_o / (np.linalg.norm(grad) * np.linalg.norm(grad_o)))
    print(f"loss {loss:.6e} vs {loss_o:.6e} (rel {rel_loss:.2e}); gradient rel L2 {rel:.2e}, cos {cos:.6f}")
    assert rel_loss < 2e-2
    assert rel < 5e-2
    assert cos > 0.998
    # per block: no block may be far off (a transposed or dropped term shows here)
    o = 0
    for r in range(S.ROUNDS):
        for name, rr, cc in S.LAYOUT:
            gb, ob = grad[o:o + rr * cc], grad_o[o:o + rr * cc]
            o += rr * cc
            nb = np.linalg.norm(ob)
            if nb > 1e-6 * np.linalg.norm(grad_o):
                assert np.linalg.norm(gb - ob) / nb < 0.15, (r, name)
    ctx.close()


def _batched(seed, n_tanks, dr):
    rng = np.random.default_rng(seed)
    arrays = []
    for _ in range(n_tanks):
        col = [float(rng.uniform(0.3, 1.2)), float(rng.uniform(0.5, 2.0))]
        arrays.append(ta.dam_break(2, dr, [3.22, 2.4], col))
    eq = ta.cfg_equations(2, dr, 2.0)
    ctx = ta.HipContext(eq)
    ctx.upload(arrays)
    return ctx, eq, arrays


def test_training_reduces_loss_and_beats_untrained():
    ctx, eq, _ = _batched(11, 4, 0.04)
    L = ta.LearnedStep(ctx, seed=7)
    ctx.step(40)
    for _ in range(6):
        L.record(4)
        ctx.step(16)
    losses = L.train(1500, lr=2e-3)
    first, last = float(np.mean(losses[:6])), float(np.mean(losses[-12:]))
    print(f"training loss {first:.4e} -> {last:.4e}")
    assert np.all(np.isfinite(losses))
    assert last < 0.6 * first
    w, sc = L.params(), L.scales()
    ctx.close()

    # held-out tanks: one judged learned step; error of the velocity change
    # against four plain steps, trained vs untrained network
    errs = {}
    for name in ("trained", "untrained"):
        hctx, _, _ = _batched(99, 2, 0.04)
        hctx.step(60)
        H = ta.LearnedStep(hctx, seed=8)
        H.set_scales(sc)
        if name == "trained":
            H.set_params(w)
        rep = H.run(1, judge_every=1, tol=1e9, dv_tol=1e9)
        errs[name] = (rep["max_dv_err"], rep["max_rho_err"])
        hctx.close()
    print(f"held-out (velocity-change error, density error) of one 4 dt step: {errs}")
    assert errs["trained"][0] < 0.8 * errs["untrained"][0]
    assert errs["trained"][1] < 0.01


def test_untrained_network_is_rejected_at_the_default_tolerances():
    """An untrained network passes the 1 % density bound of one 4 dt step
    (the density barely changes over 4 dt near rest), so the judge also bounds
    the error of the step's velocity change: at the default tolerances every
    judged group of an untrained network is rejected and covered by plain
    steps."""
    ctx, eq, _ = _batched(99, 2, 0.04)
    ctx.step(60)
    L = ta.LearnedStep(ctx, seed=8)
    L.record(4)  # output scales from this sample, as training would set them
    rep = L.run(20, judge_every=5)
    print(f"untrained: {rep}")
    assert rep["judged"] == 4
    assert rep["max_rho_err"] < ta.LearnedStep.RHO_TOL  # density alone would accept
    assert rep["max_dv_err"] > ta.LearnedStep.DV_TOL
    assert rep["rejected"] == rep["judged"]
    ctx.close()


def test_run_with_zero_tolerance_is_the_plain_trajectory():
    ctx, eq, pa = _dam(steps=10)
    L = ta.LearnedStep(ctx, seed=2)
    rep = L.run(6, judge_every=3, tol=0.0)
    assert rep["judged"] == 2 and rep["rejected"] == 2
    assert rep["plain_steps"] == 2 * 4 + 6 * 4
    ref, _, _ = _dam(steps=10 + 24)
    a, b = ctx.download(), ref.download()
    for k in ("r", "v", "rho"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(ctx.tank_clock()[0], ref.tank_clock()[0])
    ctx.close()
    ref.close()


def test_run_accepting_advances_four_dt_per_step():
    ctx, eq, pa = _dam(steps=10)
    L = ta.LearnedStep(ctx, seed=2)
    t0, dt0 = ctx.tank_clock()
    rep = L.run(3, judge_every=3, tol=1e9, dv_tol=1e9)
    assert rep["judged"] == 1 and rep["rejected"] == 0 and rep["learned_steps"] == 3
    assert rep["plain_steps"] == 4
    t1, _ = ctx.tank_clock()
    # two learned steps of 4 dt0 each, then four plain steps from the judged state
    assert t1[0] > t0[0] + 8 * dt0[0]
    g = ctx.download()
    assert np.all(np.isfinite(g["rho"])) and np.all(np.isfinite(g["r"]))
    ctx.close()


@pytest.mark.parametrize("n_tanks", [1, 3])
def test_fused_forward_matches_unfused_and_oracle(n_tanks):
    ctx, eq, _ = _batched(5, n_tanks, 0.03)
    ctx.step(20)
    L = ta.LearnedStep(ctx, seed=4)
    w = L.params()
    P = S.unpack(w)
    # make both rounds matter: larger output weights
    for p in P:
        p["W5"] = p["W5"] * 8.0
    L.set_params(S.pack(P).astype(np.float32))
    w = L.params()
    L.set_fused(True)
    fused = L.forward()
    L.set_fused(False)
    unfused = L.forward()
    nbr, nf, ef = L.graph()
    u_o, _ = S.forward(w, nf, ef, nbr)
    d_fu = np.linalg.norm(fused - unfused) / np.linalg.norm(unfused)
    d_fo = np.linalg.norm(fused - u_o) / np.linalg.norm(u_o)
    print(f"n={ctx.n}: fused vs unfused {d_fu:.2e}, fused vs FP64 oracle {d_fo:.2e}")
    assert d_fu < 1e-2
    assert d_fo < 3e-2
    ctx.close()
