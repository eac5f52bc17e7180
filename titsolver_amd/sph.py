"""Host-side mirror of TitSolver's tit::sph objects over the C ABI.

`FluidEquations` carries the constructor arguments of the reference's
FluidEquations + TaitEquationOfState + kernel (source/tit/sph/
fluid_equations.hpp:61-69, equation_of_state.hpp:27-34); `ParticleArray`
holds particles in the reference layout (fluid first, then fixed;
particle_array.hpp:138-293); `SSPRKIntegrator.step` advances a device-resident
array the way SSPRKIntegrator::step does (time_integrator.hpp:159-184) and
returns the step size.  Every computation runs in libtit_hip.so.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from ._abi import (FLAVOUR_CFG, FLAVOUR_REF, KERNEL_WENDLAND_C2,
                   KERNEL_WENDLAND_C4)


class TitError(RuntimeError):
    """Raised for a non-OK tit_status (the role of tit::Exception)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{_abi.STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def _dp(a):
    if a is None:
        return C.cast(None, C.POINTER(C.c_double))
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


@dataclass
class FluidEquations:
    """Physical parameters of the WCSPH right-hand sides."""

    dim: int = 2
    kernel: int = KERNEL_WENDLAND_C2
    flavour: int = FLAVOUR_CFG
    h: float = 0.04
    m: float = 0.4
    rho0: float = 1000.0
    cs0: float = 62.64
    xi: float = 7.0
    g: float = 9.81
    mu: float = 0.0
    alpha: float = 0.02
    delta: float = 0.1
    cfl: float = 0.25
    c_force: float = 0.25
    c_visc: float = 0.125

    def to_params(self, order: int = 3, skin: float = 0.0,
                  post_integrate: bool = False) -> _abi.Params:
        p = _abi.Params()
        p.dim, p.kernel, p.flavour, p.ssprk_order = (self.dim, self.kernel,
                                                    self.flavour, order)
        for name in ("h", "m", "rho0", "cs0", "xi", "g", "mu", "alpha",
                     "delta", "cfl", "c_force", "c_visc"):
            setattr(p, name, float(getattr(self, name)))
        p.skin = float(skin)
        p.post_integrate = 1 if post_integrate else 0
        return p


@dataclass
class ParticleArray:
    """Particles of one tank in the reference layout (fluid, then fixed)."""

    dim: int
    r: np.ndarray            # (n, dim)
    rho: np.ndarray          # (n,)
    n_fluid: int
    normal: np.ndarray       # (n_fixed, dim) wall normals of fixed particles
    box_lo: tuple = ()
    box_hi: tuple = ()
    v: np.ndarray | None = None
    dv_dt: np.ndarray | None = None
    m: np.ndarray | None = None   # (n,) per-particle masses; None: FluidEquations.m

    @property
    def n(self) -> int:
        return self.r.shape[0]

    @property
    def n_fixed(self) -> int:
        return self.n - self.n_fluid

    def fluid(self) -> slice:
        return slice(0, self.n_fluid)

    def fixed(self) -> slice:
        return slice(self.n_fluid, self.n)

    def copy(self) -> "ParticleArray":
        return ParticleArray(self.dim, self.r.copy(), self.rho.copy(), self.n_fluid,
                             self.normal.copy(), self.box_lo, self.box_hi,
                             None if self.v is None else self.v.copy(),
                             None if self.dv_dt is None else self.dv_dt.copy(),
                             None if self.m is None else self.m.copy())


def dam_break(dim: int, dr: float, tank, column, layers: int = 3,
              rho0: float = 1000.0, g: float = 9.81, cs0: float = 62.64,
              margin: float | None = None) -> ParticleArray:
    """Lattice dam break built by libtit_hip's host-side set-up."""
    lib = _abi.load()
    tank_a = np.asarray(tank, dtype=np.float64)
    col_a = np.asarray(column, dtype=np.float64)
    nf, nx = C.c_size_t(), C.c_size_t()
    st = lib.tit_dam_break_count(dim, dr, _dp(tank_a), _dp(col_a), layers,
                                 C.byref(nf), C.byref(nx))
    if st != 0:
        raise TitError(st, "dam break set-up rejected its arguments")
    n = nf.value + nx.value
    r = np.zeros((n, dim))
    rho = np.zeros(n)
    normal = np.zeros((max(nx.value, 1), dim))
    st = lib.tit_dam_break_fill(dim, dr, _dp(tank_a), _dp(col_a), layers, rho0,
                                g, cs0, _dp(r), _dp(rho), _dp(normal))
    if st != 0:
        raise TitError(st, "dam break fill failed")
    if margin is None:
        margin = 4.0 * dr
    lo = [-(layers * dr) - margin] * dim
    hi = [float(tank_a[k]) + layers * dr + margin for k in range(dim)]
    return ParticleArray(dim, r, rho, nf.value, normal[: nx.value], tuple(lo),
                         tuple(hi))


def seeded_tanks(seed: int, dr: float, n_tanks: int = 0, min_fluid: int = 0,
                 tank=(3.22, 2.4)) -> list[ParticleArray]:
    """Dam breaks of the batched configuration: columns of width U[0.3, 1.2] m
    and height U[0.5, 2.0] m drawn from numpy's default generator with `seed`,
    in the open-top tank; n_tanks of them, or as many as reach min_fluid
    fluid particles."""
    rng = np.random.default_rng(seed)
    arrays, nf = [], 0
    while (n_tanks and len(arrays) < n_tanks) or (not n_tanks and nf < min_fluid):
        col = [float(rng.uniform(0.3, 1.2)), float(rng.uniform(0.5, 2.0))]
        arrays.append(dam_break(2, dr, list(tank), col))
        nf += arrays[-1].n_fluid
    return arrays


def surge_front(r: np.ndarray, n_fluid: int, off: np.ndarray) -> float:
    """Surge-front position: the largest x over fluid particles whose
    neighbour list (self included) holds more than 8 entries in 2D, 26 in 3D,
    i.e. that post_integrate would not mark as a splash
    (fluid_equations.hpp:418-425).  `off` is the CSR offset array of the
    particles' neighbour lists."""
    dim = r.shape[1]
    cutoff = 8 if dim == 2 else 26
    sizes = np.diff(np.asarray(off, dtype=np.int64))[:n_fluid]
    keep = sizes > cutoff
    return float(r[:n_fluid][keep, 0].max()) if keep.any() else float("nan")


@dataclass
class Walls:
    """Semi-analytic wall boundary, 2D: the wall surface (vertices, faces as
    vertex pairs, each face's normal normalize(cross(b - a)) pointing into the
    fluid) and a closed polygon whose winding number decides containment, as
    FluidEquations takes them (fluid_equations.hpp:58-69)."""
    verts: np.ndarray   # (n_verts, 2)
    faces: np.ndarray   # (n_faces, 2) uint32
    cverts: np.ndarray  # (n_cverts, 2)
    cfaces: np.ndarray  # (n_cfaces, 2) uint32


def tessellate(verts, faces, d_max: float):
    """geom::tessellate (geom/tessellation.hpp:28-62): every face split into
    n = max(1, ceil(length / d_max)) equal segments; the new vertices follow
    the original ones, face by face."""
    if not d_max > 0.0:
        raise ValueError("maximum segment length must be positive")
    out_v = [tuple(map(float, v)) for v in np.asarray(verts, dtype=np.float64)]
    out_f = []
    for ia, ib in np.asarray(faces, dtype=np.int64):
        a, b = np.asarray(out_v[ia]), np.asarray(out_v[ib])
        ba = b - a
        length = math.sqrt(ba[1] * ba[1] + ba[0] * ba[0])  # norm(cross(ba))
        n = max(1, int(math.ceil(length / d_max)))
        prev = int(ia)
        for i in range(1, n):
            t = float(i) / float(n)
            out_v.append(tuple(a + t * ba))
            out_f.append((prev, len(out_v) - 1))
            prev = len(out_v) - 1
        out_f.append((prev, int(ib)))
    return np.asarray(out_v, dtype=np.float64), np.asarray(out_f, dtype=np.uint32)


def box_walls(width: float, height: float, dr: float) -> Walls:
    """The closed pool of titwcsph (wcsph.cpp:55-86): the wall surface
    (0, H) -> (W, H) -> (W, 0) -> (0, 0) -> (0, H), tessellated to dr, and the
    counter-clockwise box (0, 0), (W, 0), (W, H), (0, H) for containment."""
    sv = [(0.0, height), (width, height), (width, 0.0), (0.0, 0.0)]
    sq = [(0, 1), (1, 2), (2, 3), (3, 0)]
    verts, faces = tessellate(sv, sq, dr)
    cverts = np.asarray([(0.0, 0.0), (width, 0.0), (width, height), (0.0, height)])
    return Walls(verts, faces, cverts, np.asarray(sq, dtype=np.uint32))


def dam_break_walls(dr: float, tank, column, rho0: float = 1000.0, g: float = 9.81,
                    cs0: float = 62.64, margin: float | None = None):
    """The dam break of titwcsph with the semi-analytic wall: the fluid
    lattice r = dr (i + 1, j + 1) with the hydrostatic start, and one fixed
    particle at each wall vertex (wcsph.cpp:103-142).  Returns
    (ParticleArray, Walls)."""
    fluid = dam_break(2, dr, tank, column, layers=0, rho0=rho0, g=g, cs0=cs0)
    walls = box_walls(float(tank[0]), float(tank[1]), dr)
    nf, nv = fluid.n_fluid, len(walls.verts)
    r = np.concatenate([fluid.r[:nf], walls.verts])
    rho = np.concatenate([fluid.rho[:nf], np.full(nv, rho0)])
    if margin is None:
        margin = 4.0 * dr
    lo = (-margin, -margin)
    hi = (float(tank[0]) + margin, float(tank[1]) + margin)
    return ParticleArray(2, r, rho, nf, np.zeros((nv, 2)), lo, hi), walls


class HipContext:
    """One tit_hip_ctx: device buffers, stream and captured step graph.  With
    walls (a Walls), the semi-analytic wall boundary (tit_hip_set_walls)."""

    def __init__(self, equations: FluidEquations, order: int = 3,
                 device: int = 0, skin: float = 0.0, post_integrate: bool = False,
                 walls: Walls | None = None):
        self.lib = _abi.load()
        self.eq = equations
        self.params = equations.to_params(order, skin, post_integrate)
        h = C.c_void_p()
        st = self.lib.tit_hip_create(device, C.byref(self.params), C.byref(h))
        if st != 0:
            raise TitError(st, "tit_hip_create failed (is a GPU visible?)")
        self.h = h
        self.sizes: list[tuple[int, int]] = []
        if walls is not None:
            self._walls = [np.ascontiguousarray(walls.verts, dtype=np.float64),
                           np.ascontiguousarray(walls.faces, dtype=np.uint32),
                           np.ascontiguousarray(walls.cverts, dtype=np.float64),
                           np.ascontiguousarray(walls.cfaces, dtype=np.uint32)]
            v, f, cv, cf = self._walls
            u32 = C.POINTER(C.c_uint32)
            w = _abi.Walls(_dp(v), len(v), len(f), f.ctypes.data_as(u32), _dp(cv), len(cv),
                           len(cf), cf.ctypes.data_as(u32))
            self._check(self.lib.tit_hip_set_walls(self.h, C.byref(w)))

    def _check(self, st: int):
        if st != 0:
            raise TitError(st, self.lib.tit_hip_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib.tit_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, arrays: list[ParticleArray]):
        views = (_abi.HostView * len(arrays))()
        tanks = (_abi.Tank * len(arrays))()
        keep = []
        for t, pa in enumerate(arrays):
            assert pa.dim == self.eq.dim
            r = np.ascontiguousarray(pa.r, dtype=np.float64)
            rho = np.ascontiguousarray(pa.rho, dtype=np.float64)
            v = None if pa.v is None else np.ascontiguousarray(pa.v, dtype=np.float64)
            dv = None if pa.dv_dt is None else np.ascontiguousarray(pa.dv_dt, dtype=np.float64)
            nrm = np.ascontiguousarray(pa.normal, dtype=np.float64)
            m = None if pa.m is None else np.ascontiguousarray(pa.m, dtype=np.float64)
            assert m is None or m.shape == (pa.n,)
            keep += [r, rho, v, dv, nrm, m]
            views[t].r, views[t].rho, views[t].v = _dp(r), _dp(rho), _dp(v)
            views[t].dv_dt, views[t].normal, views[t].m = _dp(dv), _dp(nrm), _dp(m)
            views[t].n_fluid, views[t].n_fixed = pa.n_fluid, pa.n_fixed
            for k in range(pa.dim):
                tanks[t].lo[k] = pa.box_lo[k]
                tanks[t].hi[k] = pa.box_hi[k]
        self._check(self.lib.tit_hip_upload_batch(self.h, views, tanks, len(arrays)))
        self.sizes = [(pa.n_fluid, pa.n_fixed) for pa in arrays]
        self.n = sum(a + b for a, b in self.sizes)

    def step(self, n_steps: int = 1) -> float:
        dt = C.c_double()
        self._check(self.lib.tit_hip_step(self.h, n_steps, C.byref(dt)))
        return dt.value

    def download(self) -> dict:
        d = self.eq.dim
        r, v = np.empty((self.n, d)), np.empty((self.n, d))
        rho, p = np.empty(self.n), np.empty(self.n)
        self._check(self.lib.tit_hip_download(self.h, _dp(r), _dp(v), _dp(rho), _dp(p)))
        return {"r": r, "v": v, "rho": rho, "p": p}

    def tank_clock(self):
        t = np.empty(len(self.sizes))
        dt = np.empty(len(self.sizes))
        self._check(self.lib.tit_hip_tank_clock(self.h, _dp(t), _dp(dt)))
        return t, dt

    def prepare(self):
        self._check(self.lib.tit_hip_prepare(self.h))

    def compute_time_step(self) -> np.ndarray:
        dt = np.empty(len(self.sizes))
        self._check(self.lib.tit_hip_compute_time_step(self.h, _dp(dt)))
        return dt

    def compute_rhs(self):
        drho = np.empty(self.n)
        dv = np.empty((self.n, self.eq.dim))
        self._check(self.lib.tit_hip_compute_rhs(self.h, _dp(drho), _dp(dv)))
        return drho, dv

    def post_integrate(self) -> np.ndarray:
        """FluidEquations::post_integrate on the current state (context made
        with post_integrate=True); returns the free-surface indicator phi."""
        phi = np.empty(self.n)
        self._check(self.lib.tit_hip_post_integrate(self.h, _dp(phi)))
        return phi

    def neighbors(self):
        off = np.zeros(self.n + 1, dtype=np.uint64)
        total = C.c_size_t()
        st = self.lib.tit_hip_neighbors(
            self.h, off.ctypes.data_as(C.POINTER(C.c_uint64)),
            C.cast(None, C.POINTER(C.c_uint32)), 0, C.byref(total))
        if st not in (0, _abi.TIT_E_ARG):
            self._check(st)
        idx = np.zeros(max(total.value, 1), dtype=np.uint32)
        self._check(self.lib.tit_hip_neighbors(
            self.h, off.ctypes.data_as(C.POINTER(C.c_uint64)),
            idx.ctypes.data_as(C.POINTER(C.c_uint32)), idx.size, C.byref(total)))
        return off, idx[: total.value]

    def neighbor_offsets(self) -> np.ndarray:
        """CSR offsets of the sorted neighbour lists (self included), without
        materialising the lists."""
        off = np.zeros(self.n + 1, dtype=np.uint64)
        total = C.c_size_t()
        st = self.lib.tit_hip_neighbors(
            self.h, off.ctypes.data_as(C.POINTER(C.c_uint64)),
            C.cast(None, C.POINTER(C.c_uint32)), 0, C.byref(total))
        if st not in (0, _abi.TIT_E_ARG):
            self._check(st)
        return off

    def pair_counts(self) -> np.ndarray:
        """Neighbours summed per particle by the stage kernel (test hook)."""
        out = np.empty(self.n, dtype=np.uint32)
        self._check(self.lib.tit_hip_pair_counts(self.h, out.ctypes.data_as(C.POINTER(C.c_uint32))))
        return out

    def set_profiling(self, on: bool):
        self._check(self.lib.tit_hip_set_profiling(self.h, 1 if on else 0))

    def kernel_stats(self) -> dict:
        cap = 64
        arr = (_abi.KernelStat * cap)()
        n = C.c_int()
        self._check(self.lib.tit_hip_kernel_stats(self.h, arr, cap, C.byref(n)))
        return {arr[k].name.decode(): (arr[k].launches, arr[k].total_ms)
                for k in range(min(n.value, cap))}

    def get_schedule(self) -> dict:
        sch = _abi.Schedule()
        self._check(self.lib.tit_hip_get_schedule(self.h, C.byref(sch)))
        return {"chunk_targets": sch.chunk_targets, "lds_capacity": sch.lds_capacity,
                "occupancy": sch.occupancy}

    def set_schedule(self, chunk_targets: int = 0, lds_capacity: int | None = None):
        cur = self.get_schedule()
        sch = _abi.Schedule(chunk_targets, cur["lds_capacity"] if lds_capacity is None
                            else lds_capacity, 0.0)
        self._check(self.lib.tit_hip_set_schedule(self.h, C.byref(sch)))

    def autotune(self, steps_per_candidate: int = 4) -> dict:
        """Times every stage schedule over real steps, keeps the fastest (or
        takes the process cache's schedule for this kind and size of
        context: "cached" True, no steps taken)."""
        sch = _abi.Schedule()
        ms = C.c_double()
        self._check(self.lib.tit_hip_autotune(self.h, steps_per_candidate, C.byref(sch),
                                              C.byref(ms)))
        return {"chunk_targets": sch.chunk_targets, "lds_capacity": sch.lds_capacity,
                "occupancy": sch.occupancy, "ms_per_step": ms.value,
                "cached": bool(sch.cached)}

    @staticmethod
    def clear_autotune_cache():
        _abi.load().tit_hip_autotune_clear_cache()

    def counts(self):
        """(particles, cell-index rebuilds run, plain steps taken)."""
        n = C.c_size_t()
        rb = C.c_uint64()
        st = C.c_uint64()
        self._check(self.lib.tit_hip_counts(self.h, C.byref(n), C.byref(rb), C.byref(st)))
        return n.value, rb.value, st.value


def _fp(a):
    if a is None:
        return C.cast(None, C.POINTER(C.c_float))
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_float))


class LearnedStep:
    """The learned 4 dt step of one HipContext (tit_hip_learned_* entry
    points): network parameters, recorded training samples, training, and the
    judged run with plain-step fallback."""

    N_KNN = 8
    NODE_FEAT = 16
    EDGE_FEAT = 16
    OUT = 16

    def __init__(self, ctx: HipContext, seed: int = 0):
        self.ctx = ctx
        self.lib = ctx.lib
        ctx._check(self.lib.tit_hip_learned_init(ctx.h, seed))

    def params(self) -> np.ndarray:
        n = C.c_size_t()
        self.ctx._check(self.lib.tit_hip_learned_params(self.ctx.h, _fp(None), 0, C.byref(n)))
        w = np.empty(n.value, dtype=np.float32)
        self.ctx._check(self.lib.tit_hip_learned_params(self.ctx.h, _fp(w), w.size, C.byref(n)))
        return w

    def set_params(self, w: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self.ctx._check(self.lib.tit_hip_learned_set_params(self.ctx.h, _fp(w), w.size))

    def scales(self) -> np.ndarray:
        s = np.empty(6)
        self.ctx._check(self.lib.tit_hip_learned_scales(self.ctx.h, _dp(s)))
        return s

    def set_scales(self, s):
        s = np.ascontiguousarray(s, dtype=np.float64)
        assert s.size == 6
        self.ctx._check(self.lib.tit_hip_learned_set_scales(self.ctx.h, _dp(s)))

    def record(self, plain_steps: int = 4) -> int:
        k = C.c_int()
        self.ctx._check(self.lib.tit_hip_learned_record(self.ctx.h, plain_steps, C.byref(k)))
        return k.value

    def clear_samples(self):
        self.ctx._check(self.lib.tit_hip_learned_clear_samples(self.ctx.h))

    def sample(self, index: int) -> dict:
        n = C.c_size_t()
        null_u32 = C.cast(None, C.POINTER(C.c_uint32))
        null_u8 = C.cast(None, C.POINTER(C.c_uint8))
        self.ctx._check(self.lib.tit_hip_learned_sample(
            self.ctx.h, index, C.byref(n), null_u32, _fp(None), _fp(None), _fp(None), null_u8))
        n = n.value
        nbr = np.empty((n, self.N_KNN), dtype=np.uint32)
        nf = np.empty((n, self.NODE_FEAT), dtype=np.float32)
        ef = np.empty((n, self.N_KNN, self.EDGE_FEAT), dtype=np.float32)
        y = np.empty((n, 4), dtype=np.float32)
        fl = np.empty(n, dtype=np.uint8)
        self.ctx._check(self.lib.tit_hip_learned_sample(
            self.ctx.h, index, C.byref(C.c_size_t()),
            nbr.ctypes.data_as(C.POINTER(C.c_uint32)), _fp(nf), _fp(ef), _fp(y),
            fl.ctypes.data_as(C.POINTER(C.c_uint8))))
        return {"nbr": nbr, "node_feat": nf, "edge_feat": ef, "target": y,
                "fluid": fl.astype(bool)}

    def loss_grad(self, index: int):
        loss = C.c_double()
        g = np.empty(self.params().size, dtype=np.float32)
        self.ctx._check(self.lib.tit_hip_learned_loss_grad(self.ctx.h, index, C.byref(loss), _fp(g)))
        return loss.value, g

    def train(self, iters: int, lr: float = 1e-3) -> np.ndarray:
        losses = np.zeros(max(iters, 1))
        self.ctx._check(self.lib.tit_hip_learned_train(self.ctx.h, iters, lr, _dp(losses)))
        return losses[:iters]

    def graph(self):
        n = self.ctx.n
        nbr = np.empty((n, self.N_KNN), dtype=np.uint32)
        nf = np.empty((n, self.NODE_FEAT), dtype=np.float32)
        ef = np.empty((n, self.N_KNN, self.EDGE_FEAT), dtype=np.float32)
        self.ctx._check(self.lib.tit_hip_learned_graph(
            self.ctx.h, nbr.ctypes.data_as(C.POINTER(C.c_uint32)), _fp(nf), _fp(ef)))
        return nbr, nf, ef

    def forward(self) -> np.ndarray:
        out = np.empty((self.ctx.n, self.OUT), dtype=np.float32)
        self.ctx._check(self.lib.tit_hip_learned_forward(self.ctx.h, _fp(out)))
        return out

    def set_fused(self, on: bool):
        self.ctx._check(self.lib.tit_hip_learned_set_fused(self.ctx.h, 1 if on else 0))

    # Default judge tolerances: 1 % in density, and half of the velocity change
    # of the judged step (an untrained network's error is about 1).
    RHO_TOL = 0.01
    DV_TOL = 0.5

    def run(self, n_steps: int, judge_every: int = 10, tol: float = RHO_TOL,
            dv_tol: float = DV_TOL) -> dict:
        rep = _abi.LearnedReport()
        self.ctx._check(self.lib.tit_hip_learned_run(self.ctx.h, n_steps, judge_every, tol,
                                                     dv_tol, C.byref(rep)))
        return {k: getattr(rep, k) for k, _ in rep._fields_}


def nn_gemm(A, B, transA=False, transB=False, bias=None, act=0, k_splits=1, C0=None):
    """The network's BF16 MFMA GEMM on the current device (test hook)."""
    lib = _abi.load()
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    M, K = (A.shape[1], A.shape[0]) if transA else A.shape
    N = B.shape[0] if transB else B.shape[1]
    out = np.zeros((M, N), dtype=np.float32) if C0 is None else np.array(C0, dtype=np.float32)
    b = None if bias is None else np.ascontiguousarray(bias, dtype=np.float32)
    st = lib.tit_hip_nn_gemm(M, N, K, _fp(A), int(transA), _fp(B), int(transB), _fp(b),
                             act, k_splits, _fp(out))
    if st != 0:
        raise TitError(st, "tit_hip_nn_gemm failed")
    return out


class SSPRKIntegrator:
    """SSPRKIntegrator::step on device-resident particles."""

    def __init__(self, equations: FluidEquations, order: int = 3):
        self.equations = equations
        self.order = order

    def step(self, mesh: HipContext, n_steps: int = 1) -> float:
        return mesh.step(n_steps)


def cfg_equations(dim: int, dr: float, column_height: float,
                  kernel: int = KERNEL_WENDLAND_C2, h_factor: float = 2.0,
                  flavour: int = FLAVOUR_CFG) -> FluidEquations:
    """Parameters of the dam-break configurations: Wendland C2, h = h_factor dr,
    Tait rho0 = 1000, xi = 7, cs0 = 10 sqrt(2 g H), alpha = 0.02, delta = 0.1,
    CFL 0.25."""
    g = 9.81
    return FluidEquations(dim=dim, kernel=kernel, flavour=flavour,
                          h=h_factor * dr, m=1000.0 * dr ** dim, rho0=1000.0,
                          cs0=10.0 * math.sqrt(2.0 * g * column_height), xi=7.0,
                          g=g, mu=0.0, alpha=0.02, delta=0.1, cfl=0.25,
                          c_force=0.25, c_visc=0.125)


def ref_equations(dim: int, dr: float, column_height: float) -> FluidEquations:
    """Parameters of titwcsph (wcsph.cpp:47-52, :87-89): Wendland C4, h = 2 dr,
    cs0 = 20 sqrt(g H), mu = 1e-3, Ferrari diffusion, CFL 0.4."""
    g = 9.81
    return FluidEquations(dim=dim, kernel=KERNEL_WENDLAND_C4, flavour=FLAVOUR_REF,
                          h=2.0 * dr, m=1000.0 * dr ** dim, rho0=1000.0,
                          cs0=20.0 * math.sqrt(g * column_height), xi=7.0, g=g,
                          mu=1e-3, alpha=0.0, delta=0.0, cfl=0.4, c_force=0.25,
                          c_visc=0.125)
