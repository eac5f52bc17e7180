"""FP64 restatement of the learned step's network (test infrastructure only).

Only the tests, the smoke check and the CPU-baseline leg of the benchmark use
this module, as the checker of the HIP kernels; the product path never calls
it.  The network has no counterpart in the reference: it stands in for the
pair sums of FluidEquations (source/tit/sph/fluid_equations.hpp:232-305)
inside SSPRKIntegrator::step (source/tit/sph/time_integrator.hpp:159-184).

Definition (same parameter layout as the device code):

* graph: for every particle a, the 8 nearest particles b != a with
  |r_b - r_a|^2 <= h^2 (FP64, d = r_b - r_a, s = d0^2 + d1^2 [+ d2^2] with no
  fused multiply-add, the predicate form of geom/bsphere.hpp:52-54), ordered
  by (s, index of b); missing edges point back to a;
* node features nf (16): v / v_s (3), (rho / rho0 - 1) / rho_s, fixed flag,
  unit wall normal (3), zeros;
* edge features ef (16): (r_b - r_a) / h (3), q = |r_b - r_a| / h,
  w(q) / w(0) of the smoothing kernel, (v_b - v_a) / v_s (3),
  (rho_b - rho_a) / (rho0 rho_s), valid flag, zeros (all zero for a missing
  edge);
* per round r with H_r = [nf | u_{r-1}] (u_{-1} = 0):
  x_e = [ef_e | H_a | H_b], m_e = act(act(x_e W1 + b1) W2 + b2),
  M_a = sum of the 8 m_e of a, y_a = [H_a | M_a],
  u_a = act(act(y_a W3 + b3) W4 + b4) W5 + b5, with the hard-swish
  activation act(x) = x min(max(x / 6 + 1/2, 0), 1);
* loss: mean over fluid particles and used channels (acceleration 0..dim-1
  and density rate 3) of (u_c - target_c / scale_c)^2.

Pins (tests/test_oracle_surrogate.py): the backward pass against central
finite differences of the loss (FP64, relative error <= 1e-6); the graph
against a plain O(N^2) neighbour search with ties broken by index; features
and outputs against closed forms on hand-made configurations (zero weights
give the bias, a single path gives the activation's composition, a lattice gives
8 neighbours at dr and sqrt(2) dr).
"""
from __future__ import annotations

import numpy as np

N_KNN = 8
NODE_FEAT = 16
LATENT = 16
NODE_IN = NODE_FEAT + LATENT      # 32
EDGE_FEAT = 16
EDGE_IN = EDGE_FEAT + 2 * NODE_IN  # 80
HIDDEN = 128
NODE_CAT = NODE_IN + HIDDEN       # 160
OUT = LATENT                      # 16
ROUNDS = 2

# (name, rows, cols) of one round's parameter blocks, in storage order
LAYOUT = [("W1", EDGE_IN, HIDDEN), ("b1", 1, HIDDEN),
          ("W2", HIDDEN, HIDDEN), ("b2", 1, HIDDEN),
          ("W3", NODE_CAT, HIDDEN), ("b3", 1, HIDDEN),
          ("W4", HIDDEN, HIDDEN), ("b4", 1, HIDDEN),
          ("W5", HIDDEN, OUT), ("b5", 1, OUT)]
ROUND_SIZE = sum(r * c for _, r, c in LAYOUT)
N_PARAMS = ROUNDS * ROUND_SIZE


def unpack(w: np.ndarray) -> list[dict]:
    """Flat parameter vector -> per-round dicts of FP64 arrays."""
    w = np.asarray(w, dtype=np.float64)
    assert w.size == N_PARAMS
    out, o = [], 0
    for _ in range(ROUNDS):
        d = {}
        for name, r, c in LAYOUT:
            blk = w[o:o + r * c].reshape(r, c)
            d[name] = blk[0] if name.startswith("b") else blk
            o += r * c
        out.append(d)
    return out


def pack(rounds: list[dict]) -> np.ndarray:
    parts = []
    for d in rounds:
        for name, r, c in LAYOUT:
            parts.append(np.asarray(d[name], dtype=np.float64).reshape(r * c))
    return np.concatenate(parts)


# Planted mistakes for the tests that show the finite-difference check can
# fail: 1 drops the neighbour (H_b) path of the latent gradient, 2 uses the
# gate min(max(x / 6 + 1/2, 0), 1) in place of the activation's derivative
# (the product rule's second term dropped), 3 drops the bias gradients.
_mutation = 0


def set_mutation(m: int):
    global _mutation
    _mutation = int(m)


def act(x):
    """Hard-swish x min(max(x / 6 + 1/2, 0), 1)."""
    return x * np.clip(x / 6.0 + 0.5, 0.0, 1.0)


def act_grad(x):
    """Derivative of act: 0 below -3, x / 3 + 1/2 on [-3, 3], 1 above 3."""
    gate = np.clip(x / 6.0 + 0.5, 0.0, 1.0)
    if _mutation == 2:
        return gate
    inner = (x > -3.0) & (x < 3.0)
    return gate + np.where(inner, x / 6.0, 0.0)


# ---- graph and features --------------------------------------------------------

def knn_graph(r: np.ndarray, h: float, block: int = 512) -> np.ndarray:
    """n x 8 neighbour indices (self for a missing edge), brute force."""
    r = np.asarray(r, dtype=np.float64)
    n, dim = r.shape
    nbr = np.repeat(np.arange(n, dtype=np.int64)[:, None], N_KNN, axis=1)
    h2 = h * h
    idx = np.arange(n)
    for a0 in range(0, n, block):
        a1 = min(n, a0 + block)
        d = r[None, :, :] - r[a0:a1, None, :]          # r_b - r_a
        s = d[..., 0] * d[..., 0]
        for k in range(1, dim):
            s = s + d[..., k] * d[..., k]
        for row, a in enumerate(range(a0, a1)):
            ok = (s[row] <= h2) & (idx != a)
            cand = idx[ok]
            order = np.lexsort((cand, s[row][ok]))[:N_KNN]
            nbr[a, :order.size] = cand[order]
    return nbr


def kernel_w(kernel: int, q: np.ndarray) -> np.ndarray:
    """Unit Wendland profile w(q): C2 (1+2q)(1-q/2)^4, C4 (1+3q+35/12 q^2)(1-q/2)^6
    (source/tit/sph/kernel.gen.cpp:539-546), zero for q >= 2."""
    t = np.clip(1.0 - 0.5 * q, 0.0, None)
    if kernel == 0:
        return (1.0 + 2.0 * q) * t ** 4
    return (1.0 + 3.0 * q + 35.0 / 12.0 * q * q) * t ** 6


def node_features(v, rho, fixed, normal, rho0, v_s, rho_s):
    """normal: n x dim (fixed particles' wall normals, any length)."""
    v = np.asarray(v, dtype=np.float64)
    n, dim = v.shape
    f = np.zeros((n, NODE_FEAT))
    f[:, :dim] = v / v_s
    f[:, 3] = (np.asarray(rho) / rho0 - 1.0) / rho_s
    fixed = np.asarray(fixed, dtype=bool)
    f[:, 4] = fixed
    nr = np.asarray(normal, dtype=np.float64)
    ln = np.sqrt((nr * nr).sum(axis=1))
    unit = np.where(ln[:, None] > 0, nr / np.where(ln > 0, ln, 1.0)[:, None], 0.0)
    f[fixed, 5:5 + dim] = unit[fixed]
    return f


def edge_features(r, v, rho, nbr, h, kernel, rho0, v_s, rho_s):
    r = np.asarray(r, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    rho = np.asarray(rho, dtype=np.float64)
    n, dim = r.shape
    a = np.repeat(np.arange(n), N_KNN)
    b = np.asarray(nbr, dtype=np.int64).reshape(-1)
    f = np.zeros((n * N_KNN, EDGE_FEAT))
    valid = b != a
    d = (r[b] - r[a]) / h
    q = np.sqrt((d * d).sum(axis=1))
    f[:, :dim] = d
    f[:, 3] = q
    f[:, 4] = kernel_w(kernel, q) / kernel_w(kernel, np.zeros(1))[0]
    f[:, 5:5 + dim] = (v[b] - v[a]) / v_s
    f[:, 8] = (rho[b] - rho[a]) / (rho0 * rho_s)
    f[:, 9] = 1.0
    f[~valid] = 0.0
    return f.reshape(n, N_KNN, EDGE_FEAT)


# ---- network -------------------------------------------------------------------

def forward(w, nf, ef, nbr):
    """Returns (u of the last round, cache for backward)."""
    P = unpack(w)
    nf = np.asarray(nf, dtype=np.float64)
    n = nf.shape[0]
    ef = np.asarray(ef, dtype=np.float64).reshape(n * N_KNN, EDGE_FEAT)
    b = np.asarray(nbr, dtype=np.int64).reshape(-1)
    a = np.repeat(np.arange(n), N_KNN)
    H = np.concatenate([nf, np.zeros((n, LATENT))], axis=1)
    cache = []
    u = None
    for r in range(ROUNDS):
        p = P[r]
        X = np.concatenate([ef, H[a], H[b]], axis=1)
        z1 = X @ p["W1"] + p["b1"]
        h1 = act(z1)
        z2 = h1 @ p["W2"] + p["b2"]
        m = act(z2)
        M = m.reshape(n, N_KNN, HIDDEN).sum(axis=1)
        Y = np.concatenate([H, M], axis=1)
        z3 = Y @ p["W3"] + p["b3"]
        h3 = act(z3)
        z4 = h3 @ p["W4"] + p["b4"]
        h4 = act(z4)
        u = h4 @ p["W5"] + p["b5"]
        cache.append(dict(X=X, z1=z1, h1=h1, z2=z2, Y=Y, z3=z3, h3=h3, z4=z4, h4=h4))
        H = np.concatenate([nf, u], axis=1)
    return u, dict(P=P, rounds=cache, a=a, b=b, n=n)


def loss_fn(u, target, fluid, out_scale, dim):
    """Mean squared error and its gradient with respect to u."""
    u = np.asarray(u, dtype=np.float64)
    fluid = np.asarray(fluid, dtype=bool)
    chans = list(range(dim)) + [3]
    count = max(int(fluid.sum()), 1) * len(chans)
    du = np.zeros_like(u)
    y = np.asarray(target, dtype=np.float64) / np.asarray(out_scale, dtype=np.float64)[None, :4]
    L = 0.0
    for c in chans:
        res = np.where(fluid, u[:, c] - y[:, c], 0.0)
        L += float((res * res).sum())
        du[:, c] = 2.0 * res / count
    return L / count, du


def backward(cache, du):
    """Gradient of the loss with respect to the flat parameter vector."""
    P, a, b, n = cache["P"], cache["a"], cache["b"], cache["n"]
    grads = [None] * ROUNDS
    for r in range(ROUNDS - 1, -1, -1):
        p, c = P[r], cache["rounds"][r]
        g = {}
        g["W5"] = c["h4"].T @ du
        g["b5"] = du.sum(axis=0)
        dz4 = (du @ p["W5"].T) * act_grad(c["z4"])
        g["W4"] = c["h3"].T @ dz4
        g["b4"] = dz4.sum(axis=0)
        dz3 = (dz4 @ p["W4"].T) * act_grad(c["z3"])
        g["W3"] = c["Y"].T @ dz3
        g["b3"] = dz3.sum(axis=0)
        dY = dz3 @ p["W3"].T
        dM = dY[:, NODE_IN:]
        dz2 = np.repeat(dM, N_KNN, axis=0) * act_grad(c["z2"])
        g["W2"] = c["h1"].T @ dz2
        g["b2"] = dz2.sum(axis=0)
        dz1 = (dz2 @ p["W2"].T) * act_grad(c["z1"])
        g["W1"] = c["X"].T @ dz1
        g["b1"] = dz1.sum(axis=0)
        if _mutation == 3:
            for k in ("b1", "b2", "b3", "b4", "b5"):
                g[k] = np.zeros_like(g[k])
        grads[r] = g
        if r > 0:
            dX = dz1 @ p["W1"].T
            dH = dY[:, :NODE_IN].copy()
            np.add.at(dH, a, dX[:, EDGE_FEAT:EDGE_FEAT + NODE_IN])
            if _mutation != 1:
                np.add.at(dH, b, dX[:, EDGE_FEAT + NODE_IN:])
            du = dH[:, NODE_FEAT:]
    return pack(grads)


def loss_and_grad(w, nf, ef, nbr, target, fluid, out_scale, dim):
    u, cache = forward(w, nf, ef, nbr)
    L, du = loss_fn(u, target, fluid, out_scale, dim)
    return L, backward(cache, du)


def init_params(seed: int) -> np.ndarray:
    """Glorot-uniform weights, zero biases (numpy generator; any weights do
    for the checks)."""
    rng = np.random.default_rng(seed)
    rounds = []
    for _ in range(ROUNDS):
        d = {}
        for name, rr, cc in LAYOUT:
            if name.startswith("b"):
                d[name] = np.zeros(cc)
            else:
                a = np.sqrt(6.0 / (rr + cc)) * (0.1 if name == "W5" else 1.0)
                d[name] = rng.uniform(-a, a, size=(rr, cc))
        rounds.append(d)
    return pack(rounds)
