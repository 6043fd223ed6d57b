"""Test infrastructure only: ctypes access to the CPU oracle (liboracle.so).

The oracle restates the reference MLS-MPM step (src/nclr.h:80-84 and the
helpers of src/nclr_math.h) in plain C++; see nclr_oracle.hpp for what each
function follows and what pins it.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this module, always as the checker.

Arrays are particle-major numpy arrays: x (n,d), v (n,d), F (n,d,d), C (n,d,d),
Jp (n,), mass (n,), volume (n,).  Grids are flat x-major (the reference's
node order): vel (nodes,d), mass (nodes,).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SNOW, JELLY, LIQUID = 0, 1, 2

_libs: dict[str, ctypes.CDLL] = {}


def build(force: bool = False) -> None:
    """Compile liboracle.so / liboracle_fast.so with the Makefile in this folder."""
    args = ["make", "-C", HERE]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


def _lib(fast: bool = False) -> ctypes.CDLL:
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name not in _libs:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        lib.oracle_last_error.restype = ctypes.c_char_p
        lib.oracle_time_steps_f32.restype = ctypes.c_double
        _libs[name] = lib
    return _libs[name]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _ty(fp64: bool):
    return (np.float64, ctypes.c_double, "f64") if fp64 else (np.float32, ctypes.c_float, "f32")


class OracleOutOfRange(IndexError):
    """The oracle's restatement of std::out_of_range from std::vector::at."""


def _check(rc: int, lib) -> None:
    if rc == 0:
        return
    msg = lib.oracle_last_error().decode()
    if rc == 1:
        raise OracleOutOfRange(msg)
    raise RuntimeError(msg)


@dataclass
class State:
    x: np.ndarray
    v: np.ndarray
    F: np.ndarray
    C: np.ndarray
    Jp: np.ndarray
    mass: np.ndarray
    volume: np.ndarray

    @property
    def n(self) -> int:
        return self.x.shape[0]

    @property
    def dim(self) -> int:
        return self.x.shape[1]

    def astype(self, dtype) -> "State":
        return State(*(np.array(getattr(self, k), dtype=dtype, order="C", copy=True) for k in
                       ("x", "v", "F", "C", "Jp", "mass", "volume")))

    def copy(self) -> "State":
        return State(*(getattr(self, k).copy() for k in ("x", "v", "F", "C", "Jp", "mass", "volume")))


def run(state: State, material: int, res: int = 64, dt: float = 1e-4, E: float = 1e4, nu: float = 0.2,
        gravity: float = -100.0, steps: int = 1, identity_mode: int = 0, fp64: bool = False,
        return_grid: bool = False):
    """Advance `steps` steps of the restated reference step; returns a new State (and the grid)."""
    dt_, cty, suf = _ty(fp64)
    s = state.astype(dt_)
    d = s.dim
    nodes = (res + 1) ** d
    gv = np.zeros((nodes, d), dt_) if return_grid else None
    gm = np.zeros(nodes, dt_) if return_grid else None
    lib = _lib()
    fn = getattr(lib, f"oracle_run_{suf}")
    rc = fn(ctypes.c_int(d), ctypes.c_int(material), ctypes.c_int(res), cty(dt), cty(E), cty(nu), cty(gravity),
            ctypes.c_int(identity_mode), ctypes.c_int64(s.n), ctypes.c_int(steps), _ptr(s.x), _ptr(s.v), _ptr(s.F),
            _ptr(s.C), _ptr(s.Jp), _ptr(s.mass), _ptr(s.volume), _ptr(gv), _ptr(gm))
    _check(rc, lib)
    if return_grid:
        return s, gv, gm
    return s


def p2g(state: State, material: int, res: int = 64, dt: float = 1e-4, E: float = 1e4, nu: float = 0.2,
        identity_mode: int = 0, fp64: bool = False):
    """P2G only (src/nclr.h:104-165): returns (momentum (nodes,d), mass (nodes,))."""
    dt_, cty, suf = _ty(fp64)
    s = state.astype(dt_)
    d = s.dim
    nodes = (res + 1) ** d
    mom = np.zeros((nodes, d), dt_)
    m = np.zeros(nodes, dt_)
    lib = _lib()
    rc = getattr(lib, f"oracle_p2g_{suf}")(
        ctypes.c_int(d), ctypes.c_int(material), ctypes.c_int(res), cty(dt), cty(E), cty(nu),
        ctypes.c_int(identity_mode), ctypes.c_int64(s.n), _ptr(s.x), _ptr(s.v), _ptr(s.F), _ptr(s.C), _ptr(s.Jp),
        _ptr(s.mass), _ptr(s.volume), _ptr(mom), _ptr(m))
    _check(rc, lib)
    return mom, m


def grid_op(mom: np.ndarray, mass: np.ndarray, dim: int, res: int, dt: float = 1e-4, gravity: float = -100.0,
            fp64: bool = False):
    """Grid normalisation + sticky box (src/nclr.h:263-310) on copies of the inputs."""
    dt_, cty, suf = _ty(fp64)
    v = np.ascontiguousarray(mom, dtype=dt_).copy()
    m = np.ascontiguousarray(mass, dtype=dt_).copy()
    lib = _lib()
    rc = getattr(lib, f"oracle_grid_op_{suf}")(ctypes.c_int(dim), ctypes.c_int(res), cty(dt), cty(gravity),
                                               _ptr(v), _ptr(m))
    _check(rc, lib)
    return v, m


def g2p(state: State, grid_v: np.ndarray, material: int, res: int, dt: float = 1e-4, identity_mode: int = 0,
        fp64: bool = False) -> State:
    """G2P only (src/nclr.h:167-261) against a given grid velocity field."""
    dt_, cty, suf = _ty(fp64)
    s = state.astype(dt_)
    gv = np.ascontiguousarray(grid_v, dtype=dt_)
    lib = _lib()
    rc = getattr(lib, f"oracle_g2p_{suf}")(
        ctypes.c_int(s.dim), ctypes.c_int(material), ctypes.c_int(res), cty(dt), ctypes.c_int(identity_mode),
        ctypes.c_int64(s.n), _ptr(gv), _ptr(s.x), _ptr(s.v), _ptr(s.F), _ptr(s.C), _ptr(s.Jp))
    _check(rc, lib)
    return s


def svd(A: np.ndarray, fp64: bool = False):
    """nclr_svd (src/nclr_math.h:50-74) over a batch (n,d,d): returns U, sigma, V."""
    dt_, _, suf = _ty(fp64)
    A = np.ascontiguousarray(A, dtype=dt_)
    n, d, _ = A.shape
    U = np.empty_like(A)
    V = np.empty_like(A)
    S = np.empty((n, d), dt_)
    getattr(_lib(), f"oracle_svd_{suf}")(ctypes.c_int(d), ctypes.c_int64(n), _ptr(A), _ptr(U), _ptr(S), _ptr(V))
    return U, S, V


def polar(A: np.ndarray, fp64: bool = False) -> np.ndarray:
    """Rotation factor of nclr_polar (src/nclr_math.h:76-98) over a batch (n,d,d)."""
    dt_, _, suf = _ty(fp64)
    A = np.ascontiguousarray(A, dtype=dt_)
    n, d, _ = A.shape
    R = np.empty_like(A)
    getattr(_lib(), f"oracle_polar_{suf}")(ctypes.c_int(d), ctypes.c_int64(n), _ptr(A), _ptr(R))
    return R


def stress(F, C, Jp, mass, volume, material: int, res: int = 64, dt: float = 1e-4, E: float = 1e4,
           nu: float = 0.2, fp64: bool = False) -> np.ndarray:
    """The fused affine matrix of first_piola_kirchoff_stress (src/nclr.h:313-337)."""
    dt_, cty, suf = _ty(fp64)
    F = np.ascontiguousarray(F, dtype=dt_)
    n, d, _ = F.shape
    C = np.ascontiguousarray(C, dtype=dt_)
    Jp, mass, volume = (np.ascontiguousarray(a, dtype=dt_) for a in (Jp, mass, volume))
    out = np.empty_like(F)
    getattr(_lib(), f"oracle_stress_{suf}")(ctypes.c_int(d), ctypes.c_int(material), ctypes.c_int(res), cty(dt),
                                            cty(E), cty(nu), ctypes.c_int64(n), _ptr(F), _ptr(C), _ptr(Jp),
                                            _ptr(mass), _ptr(volume), _ptr(out))
    return out


def linspaced(n: int, low: float, high: float, fp64: bool = False) -> np.ndarray:
    """Eigen LinSpaced(n, low, high) as used by cube<dim> (src/nclr_math.h:103-116)."""
    dt_, cty, suf = _ty(fp64)
    out = np.empty(n, dt_)
    getattr(_lib(), f"oracle_linspaced_{suf}")(ctypes.c_int(n), cty(low), cty(high), _ptr(out))
    return out


def time_steps(state: State, material: int, res: int, dt: float = 1e-4, E: float = 1e4, nu: float = 0.2,
               gravity: float = -100.0, warmup: int = 1, steps: int = 1, identity_mode: int = 0,
               threads: int = 1) -> float:
    """Seconds for `steps` FP32 steps of the -O3 -ffast-math oracle build (the CPU baseline), serial
    (threads = 1, as the reference runs) or on `threads` OpenMP threads."""
    s = state.astype(np.float32)
    lib = _lib(fast=True)
    secs = lib.oracle_time_steps_f32(
        ctypes.c_int(s.dim), ctypes.c_int(material), ctypes.c_int(res), ctypes.c_float(dt), ctypes.c_float(E),
        ctypes.c_float(nu), ctypes.c_float(gravity), ctypes.c_int(identity_mode), ctypes.c_int64(s.n),
        ctypes.c_int(warmup), ctypes.c_int(steps), _ptr(s.x), _ptr(s.v), _ptr(s.F), _ptr(s.C), _ptr(s.Jp),
        _ptr(s.mass), _ptr(s.volume), ctypes.c_int(threads))
    if secs < 0:
        raise RuntimeError(lib.oracle_last_error().decode())
    return secs


def run_threads(state: State, material: int, res: int = 64, dt: float = 1e-4, E: float = 1e4, nu: float = 0.2,
                gravity: float = -100.0, steps: int = 1, identity_mode: int = 0, threads: int = 1,
                fast: bool = True, fp64: bool = False) -> State:
    """Steps of the oracle on `threads` OpenMP threads: FP32 (by default on the fast build, the CPU
    baseline) or FP64 (always on the exact build).  The threaded step is bitwise the serial one:
    every node receives its contributions in particle order."""
    dt_, cty, suf = _ty(fp64)
    s = state.astype(dt_)
    lib = _lib(fast=fast and not fp64)
    rc = getattr(lib, f"oracle_run_threads_{suf}")(
        ctypes.c_int(s.dim), ctypes.c_int(material), ctypes.c_int(res), cty(dt), cty(E),
        cty(nu), cty(gravity), ctypes.c_int(identity_mode), ctypes.c_int64(s.n),
        ctypes.c_int(steps), ctypes.c_int(threads), _ptr(s.x), _ptr(s.v), _ptr(s.F), _ptr(s.C), _ptr(s.Jp),
        _ptr(s.mass), _ptr(s.volume))
    _check(rc, lib)
    return s


@dataclass
class Grads:
    """Cotangents of one step's particle state (and of Young's modulus on the input side)."""
    x: np.ndarray
    v: np.ndarray
    F: np.ndarray
    C: np.ndarray
    E: float = 0.0
    E_abs: float = 0.0  # sum of the particles' |contributions| to E (the scale of its rounding)
    Jp: np.ndarray = None  # None = zero


def adjoint(state: State, grads_out: Grads, material: int, res: int = 64, dt: float = 1e-4, E: float = 1e4,
            nu: float = 0.2, gravity: float = -100.0, identity_mode: int = 0) -> Grads:
    """FP64 vector-Jacobian product of one step of any material (nclr_adjoint.hpp): given dL/d(state
    after the step) for x, v, F, C and Jp, returns dL/d(state before the step) and dL/dE."""
    s = state.astype(np.float64)
    n, d = s.n, s.dim
    g = [np.ascontiguousarray(getattr(grads_out, k), dtype=np.float64) for k in ("x", "v", "F", "C")]
    gj = np.zeros(n) if grads_out.Jp is None else np.ascontiguousarray(grads_out.Jp, dtype=np.float64)
    outs = [np.zeros_like(a) for a in g]
    oj = np.zeros(n)
    e = np.zeros(2, np.float64)
    lib = _lib()
    rc = lib.oracle_adjoint_f64(
        ctypes.c_int(d), ctypes.c_int(material), ctypes.c_int(res), ctypes.c_double(dt), ctypes.c_double(E),
        ctypes.c_double(nu), ctypes.c_double(gravity), ctypes.c_int(identity_mode), ctypes.c_int64(n), _ptr(s.x),
        _ptr(s.v), _ptr(s.F), _ptr(s.C), _ptr(s.Jp), _ptr(s.mass), _ptr(s.volume), *(_ptr(a) for a in g), _ptr(gj),
        *(_ptr(a) for a in outs), _ptr(oj), _ptr(e))
    _check(rc, lib)
    return Grads(*outs, E=float(e[0]), E_abs=float(e[1]), Jp=oj)


def adjoint_jelly(state: State, grads_out: Grads, res: int = 64, dt: float = 1e-4, E: float = 1e4, nu: float = 0.2,
                  gravity: float = -100.0, identity_mode: int = 0) -> Grads:
    """adjoint() for the jelly material."""
    return adjoint(state, grads_out, JELLY, res, dt, E, nu, gravity, identity_mode)


def rollout_vjp(state: State, grads_out: Grads, steps: int, material: int = JELLY, res: int = 64, dt: float = 1e-4,
                E: float = 1e4, nu: float = 0.2, gravity: float = -100.0, identity_mode: int = 0) -> Grads:
    """FP64 vector-Jacobian product of `steps` steps: the FP64 trajectory is stored and the
    one-step adjoint is chained backwards through it; dL/dE sums over the steps."""
    kw = dict(res=res, dt=dt, E=E, nu=nu, gravity=gravity, identity_mode=identity_mode)
    states = [state.astype(np.float64)]
    for _ in range(steps - 1):
        states.append(run(states[-1], material, steps=1, fp64=True, **kw))
    g = Grads(*(np.asarray(getattr(grads_out, k), dtype=np.float64) for k in ("x", "v", "F", "C")),
              Jp=None if grads_out.Jp is None else np.asarray(grads_out.Jp, dtype=np.float64))
    total_E = total_abs = 0.0
    for s in reversed(states):
        g = adjoint(s, g, material, **kw)
        total_E += g.E
        total_abs += g.E_abs
    g.E, g.E_abs = total_E, total_abs
    return g


def rollout_vjp_jelly(state: State, grads_out: Grads, steps: int, res: int = 64, dt: float = 1e-4, E: float = 1e4,
                      nu: float = 0.2, gravity: float = -100.0, identity_mode: int = 0) -> Grads:
    """rollout_vjp() for the jelly material."""
    return rollout_vjp(state, grads_out, steps, JELLY, res, dt, E, nu, gravity, identity_mode)


def polar_vjp(A: np.ndarray, Rbar: np.ndarray) -> np.ndarray:
    """FP64 d<Rbar, R(A)>/dA for the rotation factor R of nclr_polar, over a batch (n,d,d)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    Rbar = np.ascontiguousarray(Rbar, dtype=np.float64)
    n, d, _ = A.shape
    out = np.empty_like(A)
    _check(_lib().oracle_polar_vjp_f64(ctypes.c_int(d), ctypes.c_int64(n), _ptr(A), _ptr(Rbar), _ptr(out)), _lib())
    return out


def default_state(x: np.ndarray, identity_mode: int = 0) -> State:
    """Particles as the reference's Particle ctor makes them (src/nclr.h:46-47): v=0, F=diag<dim>(1), C=0,
    Jp=1, mass=volume=1."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, d = x.shape
    F = np.zeros((n, d, d), np.float32)
    F[:, 0, 0] = 1
    F[:, 1, 1] = 1
    if identity_mode == 1 and d == 3:
        F[:, 2, 2] = 1
    return State(x=x, v=np.zeros_like(x), F=F, C=np.zeros_like(F), Jp=np.ones(n, np.float32),
                 mass=np.ones(n, np.float32), volume=np.ones(n, np.float32))
