"""Pins of the FP64 network oracle (oracle/surrogate.py): backward vs central
finite differences (with planted mistakes that must fail), the k-nearest
graph vs a direct search, features and outputs in closed form."""
import numpy as np
import pytest

from oracle import surrogate as S


@pytest.fixture(autouse=True)
def _no_mutation():
    S.set_mutation(0)
    yield
    S.set_mutation(0)


def small_problem(seed=0, n=14, dim=2):
    rng = np.random.default_rng(seed)
    nf = rng.normal(size=(n, S.NODE_FEAT))
    ef = rng.normal(size=(n, S.N_KNN, S.EDGE_FEAT))
    nbr = rng.integers(0, n, size=(n, S.N_KNN))
    target = rng.normal(size=(n, 4))
    fluid = rng.random(n) < 0.7
    fluid[0] = True
    w = S.init_params(seed + 1)
    # larger output weights so every block has a visible gradient
    P = S.unpack(w)
    for p in P:
        p["W5"] = p["W5"] * 10.0
        p["b1"] = rng.normal(size=S.HIDDEN) * 0.1
        p["b3"] = rng.normal(size=S.HIDDEN) * 0.1
    return S.pack(P), nf, ef, nbr, target, fluid, np.array([1.5, 1.5, 1.5, 2.0]), dim


def fd_check(args, n_probe=80, seed=3):
    w, nf, ef, nbr, target, fluid, osc, dim = args
    L, g = S.loss_and_grad(w, nf, ef, nbr, target, fluid, osc, dim)
    rng = np.random.default_rng(seed)
    # probe every block of both rounds
    offs, o = [], 0
    for _ in range(S.ROUNDS):
        for name, r, c in S.LAYOUT:
            offs += list(o + rng.choice(r * c, size=min(r * c, 4), replace=False))
            o += r * c
    probe = np.array(offs[:n_probe] if len(offs) > n_probe else offs)
    eps = 1e-6
    fd = np.empty(probe.size)
    for k, i in enumerate(probe):
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        Lp, _ = S.loss_fn(S.forward(wp, nf, ef, nbr)[0], target, fluid, osc, dim)
        Lm, _ = S.loss_fn(S.forward(wm, nf, ef, nbr)[0], target, fluid, osc, dim)
        fd[k] = (Lp - Lm) / (2 * eps)
    err = np.abs(fd - g[probe]).max() / np.abs(fd).max()
    return err, g, fd, probe


def test_parameter_count_and_layout():
    assert S.ROUND_SIZE == 80 * 128 + 128 + 128 * 128 + 128 + 160 * 128 + 128 + 128 * 128 + 128 + 128 * 16 + 16
    w = np.arange(S.N_PARAMS, dtype=np.float64)
    assert np.array_equal(S.pack(S.unpack(w)), w)


@pytest.mark.parametrize("dim", [2, 3])
def test_backward_matches_finite_differences(dim):
    err, g, fd, probe = fd_check(small_problem(dim=dim))
    assert np.abs(fd).max() > 1e-3
    assert err < 1e-6


@pytest.mark.parametrize("mutation", [1, 2, 3])
def test_planted_backward_mistakes_fail(mutation):
    args = small_problem()
    S.set_mutation(mutation)
    err, *_ = fd_check(args)
    assert err > 1e-3


def test_zero_weights_give_output_bias():
    w = np.zeros(S.N_PARAMS)
    P = S.unpack(w)
    P[-1]["b5"] = np.arange(S.OUT, dtype=np.float64)
    args = small_problem()
    u, _ = S.forward(S.pack(P), args[1], args[2], args[3])
    assert np.array_equal(u, np.tile(np.arange(S.OUT, dtype=np.float64), (u.shape[0], 1)))


def test_single_path_closed_form():
    """One-hot weights route one input feature through the layers; the output
    is the activation's composition of that feature (plus the summed messages)."""
    n = 3
    nf = np.zeros((n, S.NODE_FEAT))
    nf[:, 0] = [0.5, -1.0, 2.0]
    ef = np.zeros((n, S.N_KNN, S.EDGE_FEAT))
    nbr = np.repeat(np.arange(n)[:, None], S.N_KNN, axis=1)
    P = S.unpack(np.zeros(S.N_PARAMS))
    for p in P:
        p["W3"][0, 0] = 1.0   # y[0] = H[0] = nf[0]
        p["W4"][0, 0] = 1.0
        p["W5"][0, 0] = 1.0
    u, _ = S.forward(S.pack(P), nf, ef, nbr)
    # messages: act(act(0)) = 0 for zero edge weights, so M = 0; the
    # hard-swish in closed form: 0.5 -> 0.5 * 0.5833.., -1 -> -1/3, 2 -> 5/3
    expect = np.array([0.5 * (0.5 / 6 + 0.5), -1.0 * (-1.0 / 6 + 0.5), 2.0 * (2.0 / 6 + 0.5)])
    expect = expect * np.clip(expect / 6 + 0.5, 0, 1)
    assert np.allclose(u[:, 0], expect, rtol=0, atol=1e-15)
    assert np.all(u[:, 1:] == 0)


def brute_knn(r, h):
    n = len(r)
    out = np.empty((n, S.N_KNN), dtype=np.int64)
    for a in range(n):
        c = []
        for b in range(n):
            if b == a:
                continue
            d = [r[b][k] - r[a][k] for k in range(r.shape[1])]
            s = d[0] * d[0]
            for k in range(1, len(d)):
                s = s + d[k] * d[k]
            if s <= h * h:
                c.append((s, b))
        c.sort()
        row = [b for _, b in c[:S.N_KNN]]
        out[a] = row + [a] * (S.N_KNN - len(row))
    return out


@pytest.mark.parametrize("seed", [0, 1])
def test_knn_graph_vs_direct_search(seed):
    rng = np.random.default_rng(seed)
    dr = 0.1
    ij = np.stack(np.meshgrid(np.arange(9), np.arange(7), indexing="ij"), -1).reshape(-1, 2)
    r = dr * (ij + 1.0)
    r = r + rng.normal(scale=0.02 * dr, size=r.shape) * (rng.random(len(r)) < 0.3)[:, None]
    h = 2 * dr
    assert np.array_equal(S.knn_graph(r, h, block=7), brute_knn(r, h))


def test_knn_lattice_interior_is_axis_then_diagonal():
    dr = 0.01
    ij = np.stack(np.meshgrid(np.arange(7), np.arange(7), indexing="ij"), -1).reshape(-1, 2)
    r = dr * (ij + 1.0)
    nbr = S.knn_graph(r, 2 * dr)
    a = 3 * 7 + 3  # centre
    got = [tuple(ij[b] - ij[a]) for b in nbr[a]]
    assert set(got[:4]) == {(1, 0), (-1, 0), (0, 1), (0, -1)}
    assert set(got[4:]) == {(1, 1), (1, -1), (-1, 1), (-1, -1)}
    # corner particle: 2 axis + 1 diagonal + 2 at distance 2 dr, then self padding
    c = 0
    d = [tuple(ij[b] - ij[c]) for b in nbr[c]]
    assert d[:2] == [(0, 1), (1, 0)] or d[:2] == [(1, 0), (0, 1)]
    assert d[2] == (1, 1)
    assert set(d[3:5]) == {(0, 2), (2, 0)}
    assert all(b == c for b in nbr[c][5:])


def test_edge_features_two_particles():
    h, rho0, vs, rs = 0.2, 1000.0, 3.0, 0.01
    r = np.array([[0.0, 0.0], [0.12, 0.16]])
    v = np.array([[1.0, 0.0], [0.0, 3.0]])
    rho = np.array([1000.0, 1010.0])
    nbr = np.array([[1] + [0] * 7, [0] + [1] * 7])
    ef = S.edge_features(r, v, rho, nbr, h, 0, rho0, vs, rs)
    e = ef[0, 0]
    q = 0.2 / h
    w = (1 + 2 * q) * (1 - q / 2) ** 4
    assert np.allclose(e[:5], [0.6, 0.8, 0.0, q, w], atol=1e-15)
    assert np.allclose(e[5:10], [-1 / 3, 1.0, 0.0, 1.0, 1.0], atol=1e-12)
    assert np.all(ef[0, 1:] == 0)
    assert np.allclose(ef[1, 0, :2], [-0.6, -0.8])


def test_node_features():
    v = np.array([[3.0, -6.0], [0.0, 0.0]])
    rho = np.array([1010.0, 990.0])
    nf = S.node_features(v, rho, [False, True], np.array([[0, 0], [0, 2.0]]), 1000.0, 3.0, 0.01)
    assert np.allclose(nf[0, :6], [1, -2, 0, 1.0, 0, 0])
    assert np.allclose(nf[1, :7], [0, 0, 0, -1.0, 1, 0, 1])
