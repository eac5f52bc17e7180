"""ctypes wrapper of the FP64 CPU oracle (tit_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's CPU-baseline leg.  The product library never
uses it.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libtit_oracle.so"

KERNEL_WENDLAND_C2, KERNEL_WENDLAND_C4 = 0, 1
FLAVOUR_REF, FLAVOUR_CFG = 0, 1


_DP = C.POINTER(C.c_double)
_U32P = C.POINTER(C.c_uint32)


class OrcWalls(C.Structure):
    _fields_ = [("n_verts", C.c_size_t), ("verts", _DP), ("n_faces", C.c_size_t),
                ("faces", _U32P), ("n_cverts", C.c_size_t), ("cverts", _DP),
                ("n_cfaces", C.c_size_t), ("cfaces", _U32P)]


class OrcParams(C.Structure):
    _fields_ = [("dim", C.c_int), ("kernel", C.c_int), ("flavour", C.c_int),
                ("post_integrate", C.c_int)] + [
        (n, C.c_double) for n in ("h", "m", "rho0", "cs0", "xi", "g", "mu",
                                  "alpha", "delta", "cfl", "c_force", "c_visc")] + [
        ("walls", C.POINTER(OrcWalls))]


class OrcShiftOut(C.Structure):
    _fields_ = [("N", _DP), ("grad_v", _DP), ("grad_rho", _DP), ("phi", _DP), ("dr", _DP)]


class OrcState(C.Structure):
    _fields_ = [("n_fluid", C.c_size_t), ("n_fixed", C.c_size_t),
                ("r", _DP), ("v", _DP), ("rho", _DP), ("drho_dt", _DP),
                ("dv_dt", _DP), ("p", _DP), ("cs", _DP), ("normal", _DP),
                ("gamma", _DP), ("grad_gamma", _DP), ("mass", _DP)]


_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", os.fspath(HERE)], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} missing; run `make -C oracle`")
    lib = C.CDLL(os.fspath(LIB_PATH))
    P = C.POINTER(OrcParams)
    S = C.POINTER(OrcState)
    U64 = C.POINTER(C.c_uint64)
    U32 = C.POINTER(C.c_uint32)
    sig = {
        "orc_kernel_weight": (C.c_double, [C.c_int, C.c_int]),
        "orc_kernel_unit_value": (C.c_double, [C.c_int, C.c_double]),
        "orc_kernel_unit_deriv": (C.c_double, [C.c_int, C.c_double]),
        "orc_kernel_radius": (C.c_double, [C.c_int, C.c_double]),
        "orc_kernel_value": (C.c_double, [C.c_int, C.c_int, _DP, C.c_double]),
        "orc_kernel_grad": (None, [C.c_int, C.c_int, _DP, C.c_double, _DP]),
        "orc_eos_pressure": (C.c_double, [P, C.c_double]),
        "orc_eos_density_from_pressure": (C.c_double, [P, C.c_double]),
        "orc_eos_sound_speed": (C.c_double, [P, C.c_double]),
        "orc_eos_potential": (C.c_double, [P, C.c_double]),
        "orc_eos_density_from_potential": (C.c_double, [P, C.c_double]),
        "orc_grid_search": (C.c_longlong, [C.c_int, C.c_size_t, _DP, C.c_double,
                                           C.c_double, U64, U32, C.c_size_t]),
        "orc_naive_search": (C.c_longlong, [C.c_int, C.c_size_t, _DP, C.c_double,
                                            C.c_int, U64, U32, C.c_size_t]),
        "orc_setup_boundary": (C.c_int, [P, S, U64, U32]),
        "orc_compute_continuity": (C.c_int, [P, S, U64, U32]),
        "orc_compute_momentum": (C.c_int, [P, S, U64, U32]),
        "orc_compute_time_step": (C.c_double, [P, S]),
        "orc_ssprk_step": (C.c_int, [P, S, C.c_int, _DP]),
        "orc_post_integrate": (C.c_int, [P, S, U64, U32, C.POINTER(OrcShiftOut)]),
        "orc_ssprk_generic": (C.c_int, [C.c_int, C.c_double, C.c_size_t, _DP,
                                        C.c_void_p, C.c_void_p]),
        "orc_set_mutation": (None, [C.c_int]),
        "orc_kernel_unit_flux": (C.c_double, [C.c_int, C.c_double, C.c_double, C.c_double]),
        "orc_kernel_unit_antigrad_flux": (C.c_double, [C.c_int, C.c_double, C.c_double,
                                                       C.c_double]),
        "orc_face_flux": (None, [C.c_int, _DP, _DP, _DP, C.c_double, _DP]),
        "orc_face_antigrad_flux": (C.c_double, [C.c_int, _DP, _DP, _DP, C.c_double]),
        "orc_face_intersects": (C.c_int, [_DP, _DP, _DP, C.c_double]),
        "orc_winding": (C.c_double, [C.POINTER(OrcWalls), _DP]),
        "orc_compute_gamma": (C.c_int, [P, S]),
        "orc_initialize": (C.c_int, [P, S]),
        "orc_eos_pressure_array": (None, [P, C.c_size_t, _DP, _DP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], a.dtype
    return a.ctypes.data_as(_DP)


class Walls:
    """Wall surface and containment polygon (2D), owned by numpy arrays; any
    object with verts, faces, cverts, cfaces (titsolver_amd.Walls) converts."""

    def __init__(self, walls):
        self.verts = np.ascontiguousarray(walls.verts, dtype=np.float64).reshape(-1, 2).copy()
        self.faces = np.ascontiguousarray(walls.faces, dtype=np.uint32).reshape(-1, 2).copy()
        self.cverts = np.ascontiguousarray(walls.cverts, dtype=np.float64).reshape(-1, 2).copy()
        self.cfaces = np.ascontiguousarray(walls.cfaces, dtype=np.uint32).reshape(-1, 2).copy()
        u32 = lambda a: a.ctypes.data_as(_U32P)  # noqa: E731
        self.c = OrcWalls(len(self.verts), _dp(self.verts), len(self.faces), u32(self.faces),
                          len(self.cverts), _dp(self.cverts), len(self.cfaces), u32(self.cfaces))


def params(eq, walls=None) -> OrcParams:
    """OrcParams from a titsolver_amd.FluidEquations (same field names); with
    walls, the semi-analytic wall form."""
    p = OrcParams()
    p.dim, p.kernel, p.flavour = eq.dim, eq.kernel, eq.flavour
    for n in ("h", "m", "rho0", "cs0", "xi", "g", "mu", "alpha", "delta", "cfl",
              "c_force", "c_visc"):
        setattr(p, n, float(getattr(eq, n)))
    if walls is not None:
        w = walls if isinstance(walls, Walls) else Walls(walls)
        p._walls = w  # keeps the arrays alive
        p.walls = C.pointer(w.c)
    return p


def winding(cverts, cfaces, pts) -> np.ndarray:
    """orc_winding at each point: the exact winding sum of the polygon
    (cverts, cfaces)."""
    from types import SimpleNamespace
    w = Walls(SimpleNamespace(verts=np.zeros((0, 2)), faces=np.zeros((0, 2)),
                              cverts=cverts, cfaces=cfaces))
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 2)
    fn = load().orc_winding
    return np.array([fn(C.byref(w.c), _dp(pts[i])) for i in range(len(pts))])


def pressure(prm, rho):
    """Tait pressure of each density (FP64)."""
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    p = np.empty_like(rho)
    load().orc_eos_pressure_array(C.byref(prm), rho.size, _dp(rho), _dp(p))
    return p


def set_mutation(m: int):
    load().orc_set_mutation(int(m))


def _search(fn, *args, n, dim, pts):
    lib = load()
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(n, dim)
    off = np.zeros(n + 1, dtype=np.uint64)
    cap = max(64 * n, 1)
    while True:
        idx = np.zeros(cap, dtype=np.uint32)
        r = fn(dim, n, _dp(pts), *args,
               off.ctypes.data_as(C.POINTER(C.c_uint64)),
               idx.ctypes.data_as(C.POINTER(C.c_uint32)), cap)
        if r == -1:
            raise ValueError("bad search arguments")
        if r < 0:
            cap = -r
            continue
        return off, idx[:r]


def grid_search(pts, radius, size_hint):
    pts = np.asarray(pts, dtype=np.float64)
    n, dim = pts.shape
    return _search(load().orc_grid_search, float(radius), float(size_hint),
                   n=n, dim=dim, pts=pts)


def naive_search(pts, radius, strict=False):
    pts = np.asarray(pts, dtype=np.float64)
    n, dim = pts.shape
    return _search(load().orc_naive_search, float(radius), 1 if strict else 0,
                   n=n, dim=dim, pts=pts)


class State:
    """Particle state in the reference layout, owned by numpy arrays."""

    def __init__(self, r, rho, n_fluid, normal, v=None, dv_dt=None, mass=None):
        r = np.ascontiguousarray(r, dtype=np.float64)
        self.n, self.dim = r.shape
        self.n_fluid = int(n_fluid)
        self.r = r.copy()
        self.v = np.zeros_like(r) if v is None else np.ascontiguousarray(v, dtype=np.float64).copy()
        self.rho = np.ascontiguousarray(rho, dtype=np.float64).copy()
        self.drho_dt = np.zeros(self.n)
        self.dv_dt = np.zeros_like(r) if dv_dt is None else np.ascontiguousarray(dv_dt, dtype=np.float64).copy()
        self.p = np.zeros(self.n)
        self.cs = np.zeros(self.n)
        self.normal = np.ascontiguousarray(normal, dtype=np.float64).reshape(-1, self.dim).copy()
        if self.normal.shape[0] == 0:
            self.normal = np.zeros((1, self.dim))
        self.gamma = np.ones(self.n)
        self.grad_gamma = np.zeros_like(r)
        # per-particle masses; zero entries mean the parameters' uniform m
        self.mass = (np.zeros(self.n) if mass is None
                     else np.ascontiguousarray(mass, dtype=np.float64).copy())
        assert self.mass.shape == (self.n,)
        self.c = OrcState(self.n_fluid, self.n - self.n_fluid, _dp(self.r), _dp(self.v),
                          _dp(self.rho), _dp(self.drho_dt), _dp(self.dv_dt),
                          _dp(self.p), _dp(self.cs), _dp(self.normal), _dp(self.gamma),
                          _dp(self.grad_gamma), _dp(self.mass))

    @classmethod
    def from_particles(cls, pa):
        return cls(pa.r, pa.rho, pa.n_fluid, pa.normal, pa.v, pa.dv_dt,
                   getattr(pa, "m", None))


def _csr(off, idx):
    return (off.ctypes.data_as(C.POINTER(C.c_uint64)),
            idx.ctypes.data_as(C.POINTER(C.c_uint32)))


def adjacency(st: State, prm: OrcParams):
    """Sorted neighbour lists as ParticleMesh builds them (hint h, radius 2h)."""
    return grid_search(st.r, 2.0 * prm.h, prm.h)


def initialize(prm, st):
    """FluidEquations::initialize: masses, and with walls gamma and the fixed
    particles' mass fractions."""
    if load().orc_initialize(C.byref(prm), C.byref(st.c)) != 0:
        raise RuntimeError("oracle initialize failed")


def compute_gamma(prm, st):
    if load().orc_compute_gamma(C.byref(prm), C.byref(st.c)) != 0:
        raise RuntimeError("oracle compute_gamma failed")


def unit_flux(kernel, eta, z0, z1, antigrad=False):
    fn = load().orc_kernel_unit_antigrad_flux if antigrad else load().orc_kernel_unit_flux
    return fn(int(kernel), float(eta), float(z0), float(z1))


def face_flux(kernel, a, b, x, h):
    a, b, x = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, x))
    out = np.zeros(2)
    load().orc_face_flux(int(kernel), _dp(a), _dp(b), _dp(x), float(h), _dp(out))
    return out


def face_antigrad_flux(kernel, a, b, x, h):
    a, b, x = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, x))
    return load().orc_face_antigrad_flux(int(kernel), _dp(a), _dp(b), _dp(x), float(h))


def setup_boundary(prm, st, off, idx):
    load().orc_setup_boundary(C.byref(prm), C.byref(st.c), *_csr(off, idx))


def compute_continuity(prm, st, off, idx):
    load().orc_compute_continuity(C.byref(prm), C.byref(st.c), *_csr(off, idx))


def compute_momentum(prm, st, off, idx):
    load().orc_compute_momentum(C.byref(prm), C.byref(st.c), *_csr(off, idx))


def post_integrate(prm, st, off, idx) -> dict:
    """FluidEquations::apply_shifts + apply_free_surface_correction on the
    given adjacency; updates st and returns the renormalised N, grad v,
    grad rho, the free-surface indicator phi and the applied shift dr."""
    n, dim = st.n, st.dim
    res = {"N": np.zeros((n, dim)), "grad_v": np.zeros((n, dim, dim)),
           "grad_rho": np.zeros((n, dim)), "phi": np.zeros(n), "dr": np.zeros((n, dim))}
    out = OrcShiftOut(*(_dp(res[k]) for k in ("N", "grad_v", "grad_rho", "phi", "dr")))
    rc = load().orc_post_integrate(C.byref(prm), C.byref(st.c), *_csr(off, idx), C.byref(out))
    if rc != 0:
        raise RuntimeError("oracle post_integrate failed")
    return res


def compute_time_step(prm, st) -> float:
    return load().orc_compute_time_step(C.byref(prm), C.byref(st.c))


def ssprk_step(prm, st, order=3) -> float:
    dt = C.c_double()
    rc = load().orc_ssprk_step(C.byref(prm), C.byref(st.c), order, C.byref(dt))
    if rc != 0:
        raise RuntimeError("oracle step failed")
    return dt.value


RHS_FN = C.CFUNCTYPE(None, C.c_void_p, _DP, _DP, C.c_size_t)


def ssprk_generic(order, dt, u, rhs):
    """u' = rhs(u) stepped once with the oracle's SSPRK combination."""
    u = np.ascontiguousarray(u, dtype=np.float64).copy()
    n = u.size

    def cb(_ctx, up, dup, nn):
        uu = np.ctypeslib.as_array(up, shape=(nn,))
        du = np.ctypeslib.as_array(dup, shape=(nn,))
        du[:] = rhs(uu.copy())

    f = RHS_FN(cb)
    rc = load().orc_ssprk_generic(order, dt, n, _dp(u), C.cast(f, C.c_void_p), None)
    if rc != 0:
        raise RuntimeError("ssprk_generic failed")
    return u


# ---------------------------------------------------------------------------
# Threaded CPU baseline (cpu_step.cpp): the same step on a std::thread pool.

CPU_LIB_PATH = HERE / "libtit_cpu.so"
_cpu = None


def _load_cpu():
    global _cpu
    if _cpu is None:
        if not CPU_LIB_PATH.exists():
            raise RuntimeError(f"{CPU_LIB_PATH} missing; run `make -C oracle`")
        lib = C.CDLL(os.fspath(CPU_LIB_PATH))
        lib.cpu_set_threads.restype = C.c_int
        lib.cpu_set_threads.argtypes = [C.c_int]
        lib.cpu_ssprk_step.restype = C.c_int
        lib.cpu_ssprk_step.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcState), C.c_int, _DP]
        _cpu = lib
    return _cpu


def cpu_set_threads(n: int) -> int:
    """Thread count of the CPU baseline's pool (1: the caller only)."""
    return _load_cpu().cpu_set_threads(int(n))


def cpu_ssprk_step(prm, st, order=3) -> float:
    """One step of the threaded CPU baseline (wall-free form, no
    post-integration); the same contract as ssprk_step."""
    dt = C.c_double()
    rc = _load_cpu().cpu_ssprk_step(C.byref(prm), C.byref(st.c), order, C.byref(dt))
    if rc != 0:
        raise RuntimeError("CPU baseline step rejected its arguments")
    return dt.value
