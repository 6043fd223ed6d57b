"""Python mirror of the reference's engine API (src/nclr.h:20-87) over the HIP engine.

    from nuclearmpm_amd import MPMSimulation, MaterialModel, Particle, cube
    parts = [Particle(p, 0xED553B) for p in cube(2, 50, 0.4, 0.6)]
    sim = MPMSimulation(parts, MaterialModel.kSnow)
    sim.advance()
    state = sim.particles()          # numpy arrays, original particle order

MPMBatch advances many independent scenes (uneven particle counts, any materials, per-scene
res/dt/E/nu/gravity) with one launch sequence per step; scene s behaves exactly like an
MPMSimulation built from the same arguments.

Differences from the reference, all deliberate:
  * particles() returns a Particles record of numpy arrays (x (n,d), v, F (n,d,d), C, Jp,
    mass, volume, c) instead of a vector of structs;
  * advance(steps) takes an optional step count (queued on the device without waiting);
  * a stencil leaving the grid raises IndexError (the reference's std::out_of_range) at the next
    call that reads results, and is raised for every axis, where the reference aliases silently
    into the neighbouring row on the faster axes.  The scene it happens in stops at its state
    before that step (what the reference's throwing advance() leaves); in a batch the other scenes
    keep stepping, and scene_status() lists the stopped scenes without raising.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from enum import IntEnum
from typing import List, Optional, Sequence, Union

import numpy as np

from ._lib import check, lib, mpm_options, mpm_scene, ptr


class MaterialModel(IntEnum):
    """src/nclr.h:57-61"""
    kSnow = 0
    kJelly = 1
    kLiquid = 2


def diag(dim: int, value: float, identity_mode: int = 0) -> np.ndarray:
    """diag<dim>(value) of src/nclr_math.h:13-19: only (0,0) and (1,1) are set in reference mode."""
    m = np.zeros((dim, dim), np.float32)
    m[0, 0] = value
    m[1, 1] = value
    if identity_mode == 1:
        np.fill_diagonal(m, value)
    return m


@dataclass
class Particle:
    """One particle as built by the reference ctor (src/nclr.h:46-47): F = diag<dim>(1), C = 0, Jp = 1."""
    x: np.ndarray
    c: int = 0
    v: Optional[np.ndarray] = None
    mass: float = 1.0
    volume: float = 1.0


@dataclass
class Particles:
    """Structure-of-arrays particle set (the reference's std::vector<Particle<dim>>)."""
    x: np.ndarray
    v: np.ndarray
    F: np.ndarray
    C: np.ndarray
    Jp: np.ndarray
    mass: np.ndarray
    volume: np.ndarray
    c: np.ndarray = field(default=None)

    def __post_init__(self):
        for k in ("x", "v", "F", "C", "Jp", "mass", "volume"):
            setattr(self, k, np.ascontiguousarray(getattr(self, k), dtype=np.float32))
        if self.c is None:
            self.c = np.zeros(self.n, np.int32)

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @property
    def dim(self) -> int:
        return int(self.x.shape[1])

    def __len__(self) -> int:
        return self.n

    @staticmethod
    def from_positions(x, v=None, mass=None, volume=None, c=None, identity_mode: int = 0) -> "Particles":
        """Particles with the ctor defaults of src/nclr.h:46-47 at positions x (n, d)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        n, d = x.shape
        F = np.broadcast_to(diag(d, 1.0, identity_mode), (n, d, d)).copy()
        return Particles(
            x=x,
            v=np.zeros_like(x) if v is None else np.broadcast_to(np.asarray(v, np.float32), (n, d)).copy(),
            F=F,
            C=np.zeros((n, d, d), np.float32),
            Jp=np.ones(n, np.float32),
            mass=np.ones(n, np.float32) if mass is None else np.broadcast_to(np.asarray(mass, np.float32), (n,)).copy(),
            volume=np.ones(n, np.float32) if volume is None else np.broadcast_to(np.asarray(volume, np.float32),
                                                                                  (n,)).copy(),
            c=None if c is None else np.broadcast_to(np.asarray(c, np.int32), (n,)).copy(),
        )

    @staticmethod
    def from_list(parts: Sequence[Particle], identity_mode: int = 0) -> "Particles":
        if len(parts) == 0:
            raise ValueError("empty particle list: pass Particles with an explicit dimension")
        x = np.stack([np.asarray(p.x, np.float32) for p in parts])
        v = np.stack([np.zeros_like(x[0]) if p.v is None else np.asarray(p.v, np.float32) for p in parts])
        return Particles.from_positions(x, v=v, mass=np.array([p.mass for p in parts], np.float32),
                                        volume=np.array([p.volume for p in parts], np.float32),
                                        c=np.array([p.c for p in parts], np.int32), identity_mode=identity_mode)

    @staticmethod
    def concatenate(parts: Sequence["Particles"]) -> "Particles":
        return Particles(*(np.concatenate([getattr(p, k) for p in parts]) for k in
                           ("x", "v", "F", "C", "Jp", "mass", "volume", "c")))

    def slice(self, a: int, b: int) -> "Particles":
        return Particles(*(getattr(self, k)[a:b].copy() for k in ("x", "v", "F", "C", "Jp", "mass", "volume", "c")))

    def copy(self) -> "Particles":
        return Particles(*(getattr(self, k).copy() for k in ("x", "v", "F", "C", "Jp", "mass", "volume", "c")))


def _image_dim(image) -> int:
    rows = int(image.shape[0])
    for d in (2, 3):
        if rows == 2 * d + 2 * d * d + 1:
            return d
    raise ValueError(f"a state image has 13 (2D) or 25 (3D) rows, not {rows}")


def _placeholder(n: int, dim: int) -> Particles:
    # positions are overwritten from the image before any step
    return Particles.from_positions(np.zeros((n, dim), np.float32))


SVD_MODES = {"exact": 0, "fast": 1, "accurate": 2}


def svd_mode(svd: Optional[str], fast_svd: bool = False) -> int:
    """Engine option fast_svd from svd = "exact" (nclr_svd's FP32 operation order, bitwise the
    reference's arithmetic; the default), "fast" (FP32 Jacobi of A^T A) or "accurate" (FP64 Newton
    polar factor and FP64 Jacobi SVD); fast_svd=True is the older spelling of "fast"."""
    if svd is None:
        return 1 if fast_svd else 0
    if svd not in SVD_MODES:
        raise ValueError(f"svd must be one of {sorted(SVD_MODES)}")
    return SVD_MODES[svd]


def default_options(**overrides) -> mpm_options:
    o = mpm_options()
    lib().mpm_default_options(ctypes.byref(o))
    for k, v in overrides.items():
        if not hasattr(o, k):
            raise AttributeError(f"unknown engine option {k!r}")
        setattr(o, k, int(v))
    return o


class _Engine:
    """Handle management shared by MPMSimulation and MPMBatch."""

    _h = None

    def _init_common(self, dim: int, n: int, colors, identity_mode, p2g_chunk, g2p_threads, graph, autotune, timing,
                     device, fast_svd=False, svd=None, compact=0):
        self._dim = dim
        self._n = n
        self._colors = colors
        self.options = default_options(identity_mode=identity_mode, p2g_chunk=p2g_chunk, g2p_threads=g2p_threads,
                                       graph=int(graph), autotune=int(autotune), timing=int(timing), device=device,
                                       fast_svd=svd_mode(svd, fast_svd), compact=int(compact))
        self._cache = None

    def set_state_image(self, image) -> None:
        """Sets x, v, F, C, Jp of every particle from a device state image: a float32 tensor of shape
        (2d + 2d^2 + 1, n) on the engine's device, rows x, v, F, C (row-major), Jp, particle order
        (the format of nuclearmpm_amd.generators).  Mass and volume are kept."""
        import torch

        rows = 2 * self._dim + 2 * self._dim * self._dim + 1
        if (not isinstance(image, torch.Tensor) or image.dtype != torch.float32 or image.device.type != "cuda"
                or tuple(image.shape) != (rows, self._n)):
            raise ValueError(f"state image must be a float32 cuda tensor of shape ({rows}, {self._n})")
        image = image.contiguous()
        torch.cuda.synchronize(image.device)
        self._cache = None
        check(lib().mpm_set_state_device(self._handle(), image.data_ptr() if self._n else None))
        check(lib().mpm_synchronize(self._handle()))

    def state_image(self):
        """The current x, v, F, C, Jp as a device state image (particle order)."""
        import torch

        rows = 2 * self._dim + 2 * self._dim * self._dim + 1
        out = torch.empty((rows, self._n), dtype=torch.float32, device=f"cuda:{self.options.device}")
        check(lib().mpm_synchronize(self._handle()))
        check(lib().mpm_get_state_device(self._handle(), out.data_ptr() if self._n else None))
        check(lib().mpm_synchronize(self._handle()))
        return out

    def advance(self, steps: int = 1) -> None:
        """advance() (src/nclr.h:80-84), `steps` times for every scene; queued without waiting."""
        self._cache = None
        check(lib().mpm_advance(self._handle(), int(steps)))

    def _all_particles(self) -> Particles:
        if self._cache is None:
            n, d = self._n, self._dim
            out = Particles(x=np.empty((n, d), np.float32), v=np.empty((n, d), np.float32),
                            F=np.empty((n, d, d), np.float32), C=np.empty((n, d, d), np.float32),
                            Jp=np.empty(n, np.float32), mass=np.empty(n, np.float32),
                            volume=np.empty(n, np.float32), c=self._colors.copy())
            check(lib().mpm_get_particles(self._handle(), ptr(out.x), ptr(out.v), ptr(out.F), ptr(out.C),
                                          ptr(out.Jp), ptr(out.mass), ptr(out.volume)))
            self._cache = out
        return self._cache

    def _grid(self, scene: int, res: int):
        nodes = (res + 1) ** self._dim
        vel = np.empty((nodes, self._dim), np.float32)
        mass = np.empty(nodes, np.float32)
        check(lib().mpm_get_grid_scene(self._handle(), int(scene), ptr(vel), ptr(mass)))
        shape = (res + 1,) * self._dim
        return vel.reshape(shape + (self._dim,)), mass.reshape(shape)

    def set_particles(self, p: Particles) -> None:
        """Overwrite the particle state of all scenes (same count and dimension)."""
        if p.n != self._n or p.dim != self._dim:
            raise ValueError("particle count/dimension mismatch")
        self._cache = None
        check(lib().mpm_set_particles(self._handle(), ptr(p.x), ptr(p.v), ptr(p.F), ptr(p.C), ptr(p.Jp),
                                      ptr(p.mass), ptr(p.volume)))

    def synchronize(self) -> None:
        check(lib().mpm_synchronize(self._handle()))

    def scene_status(self) -> dict:
        """Stopped scenes without raising: {scene: {"flags": status bits, "step": the step that
        stopped it}} (bit 1: a particle's stencil left the grid, the reference's std::out_of_range;
        bit 2: a particle crossed more than one tile).  Waits for queued steps.  A scene stops on its
        own (its particles stay as they were before that step) while the others keep stepping."""
        ns = self.scenes
        flags = np.zeros(ns, np.uint32)
        steps = np.zeros(ns, np.int64)
        check(lib().mpm_scene_status(self._handle(), ptr(flags), ptr(steps)))
        return {s: {"flags": int(flags[s]), "step": int(steps[s])} for s in range(ns) if flags[s] or steps[s] >= 0}

    def kernel_time(self, name: str):
        """(average ms per launch, launches) of kernel group `name` ("sort", "p2g", "grid", "g2p",
        "step"); needs timing on."""
        avg = ctypes.c_double()
        cnt = ctypes.c_int64()
        check(lib().mpm_kernel_time(self._handle(), name.encode(), ctypes.byref(avg), ctypes.byref(cnt)))
        return avg.value, cnt.value

    def set_timing(self, enable: bool) -> None:
        check(lib().mpm_set_timing(self._handle(), int(bool(enable))))

    def reset_kernel_times(self) -> None:
        check(lib().mpm_reset_kernel_times(self._handle()))

    def set_tuning(self, p2g_chunk: int, g2p_threads: int) -> None:
        check(lib().mpm_set_tuning(self._handle(), int(p2g_chunk), int(g2p_threads)))

    def tuning(self):
        """The launch shape: P2G work-item size, G2P workgroup size and the compact step's
        particles per thread (0: the tiled step)."""
        c = ctypes.c_int()
        t = ctypes.c_int()
        k = ctypes.c_int()
        check(lib().mpm_get_tuning(self._handle(), ctypes.byref(c), ctypes.byref(t)))
        check(lib().mpm_get_compact(self._handle(), ctypes.byref(k)))
        return {"p2g_chunk": c.value, "g2p_threads": t.value, "compact": k.value}

    def set_compact(self, particles_per_thread: int) -> None:
        """The compact step (single scenes): all steps of an advance in one cooperative launch of
        ceil(n / (256 particles_per_thread)) workgroups, no sort, the grid box in LDS; 0 turns it
        off.  Same parity bound as the tiled step, not bitwise equal to it."""
        check(lib().mpm_set_compact(self._handle(), int(particles_per_thread)))

    def autotune(self, steps_per_trial: int = 3):
        """Time every launch shape (and the compact step, on single scenes) on the current state
        (restored afterwards) and keep the fastest."""
        check(lib().mpm_autotune(self._handle(), int(steps_per_trial)))
        return self.tuning()

    def set_elasticity(self, scene: int, E: float, nu: Optional[float] = None) -> None:
        """Young's modulus (and Poisson's ratio) of one scene; mu_0 and lambda_0 follow in float."""
        nu = self._nus[scene] if nu is None else float(nu)
        check(lib().mpm_set_scene_elasticity(self._handle(), int(scene), float(E), nu))
        self._nus[scene] = nu

    @property
    def scenes(self) -> int:
        """Number of scenes the handle advances (1 for a plain simulation)."""
        return len(self._nus)

    def active_nodes(self, scene: int = -1) -> int:
        c = ctypes.c_int64()
        check(lib().mpm_active_nodes_scene(self._handle(), int(scene), ctypes.byref(c)))
        return c.value

    @property
    def steps_taken(self) -> int:
        return int(lib().mpm_steps_taken(self._handle()))

    @property
    def dim(self) -> int:
        return self._dim

    @property
    def n(self) -> int:
        return self._n

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().mpm_destroy(self._h)
            self._h = None

    def _handle(self):
        if not self._h:
            raise ValueError("simulation is closed")
        return self._h

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class MPMSimulation(_Engine):
    """MPMSimulation<dim> (src/nclr.h:63-87) running on the MI355X engine."""

    kBoundary = 3
    kSnowHardening = np.float32(10.0)
    kJellyHardening = np.float32(0.3)
    kLiquidHardening = np.float32(1.0)

    def __init__(self, particles: Union[Particles, Sequence[Particle]], model: MaterialModel, res: int = 64,
                 dt: float = 1e-4, E: float = 1e4, nu: float = 0.2, gravity: float = -100.0, *,
                 identity_mode: int = 0, p2g_chunk: int = 0, g2p_threads: int = 0, graph: bool = False,
                 autotune: bool = False, timing: bool = False, device: int = 0, fast_svd: bool = False,
                 svd: Optional[str] = None, compact: int = 0):
        if not isinstance(particles, Particles):
            particles = Particles.from_list(list(particles), identity_mode=identity_mode)
        if particles.dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        self._res = int(res)
        self._model = MaterialModel(model)
        self._nus = [float(nu)]
        self._init_common(particles.dim, particles.n, particles.c.copy(), identity_mode, p2g_chunk, g2p_threads,
                          graph, autotune, timing, device, fast_svd, svd, compact)
        h = ctypes.c_void_p()
        check(lib().mpm_create(self._dim, int(self._model), self._res, dt, E, nu, gravity, self._n,
                               ptr(particles.x), ptr(particles.v), ptr(particles.F), ptr(particles.C),
                               ptr(particles.Jp), ptr(particles.mass), ptr(particles.volume),
                               ctypes.byref(self.options), ctypes.byref(h)))
        self._h = h
        mu, lam, dx, inv_dx = (ctypes.c_float() for _ in range(4))
        check(lib().mpm_get_params(self._h, ctypes.byref(mu), ctypes.byref(lam), ctypes.byref(dx),
                                   ctypes.byref(inv_dx)))
        self.mu_0 = np.float32(mu.value)
        self.lambda_0 = np.float32(lam.value)
        self.dx = np.float32(dx.value)
        self.inv_dx = np.float32(inv_dx.value)

    @classmethod
    def from_image(cls, image, model: MaterialModel, res: int = 64, dt: float = 1e-4, E: float = 1e4,
                   nu: float = 0.2, gravity: float = -100.0, *, autotune: bool = False, **kw) -> "MPMSimulation":
        """A simulation whose particles come from a device state image (nuclearmpm_amd.generators);
        Jp, mass and volume take the Particle ctor defaults (1)."""
        dim = _image_dim(image)
        n = int(image.shape[1])
        sim = cls(_placeholder(n, dim), model, res, dt, E, nu, gravity, **kw)
        sim.set_state_image(image)
        if autotune:
            sim.autotune()
        return sim

    def particles(self) -> Particles:
        """particles() (src/nclr.h:86): the current state, original particle order."""
        return self._all_particles()

    def grid(self):
        """grid() (src/nclr.h:87): (velocity (res+1,)*d + (d,), mass (res+1,)*d) after the last grid update."""
        return self._grid(0, self._res)

    @property
    def res(self) -> int:
        return self._res


@dataclass
class SceneSpec:
    """One scene of a batch: the reference ctor's arguments."""
    particles: Particles
    material: MaterialModel
    res: int = 64
    dt: float = 1e-4
    E: float = 1e4
    nu: float = 0.2
    gravity: float = -100.0


class MPMBatch(_Engine):
    """Many independent scenes advanced together: one launch sequence per step for all of them."""

    def __init__(self, scenes: Sequence, *, identity_mode: int = 0, p2g_chunk: int = 0, g2p_threads: int = 0,
                 graph: bool = False, autotune: bool = False, timing: bool = False, device: int = 0,
                 fast_svd: bool = False, svd: Optional[str] = None):
        specs: List[SceneSpec] = []
        for s in scenes:
            if isinstance(s, SceneSpec):
                specs.append(s)
            else:  # scenes.Scene or anything with the same attributes
                specs.append(SceneSpec(s.particles, MaterialModel(s.material), s.res, s.dt, s.E, s.nu, s.gravity))
        if not specs:
            raise ValueError("a batch needs at least one scene")
        dims = {sp.particles.dim for sp in specs}
        if len(dims) != 1:
            raise ValueError("all scenes of a batch must have the same dimension")
        self._specs = specs
        self._nus = [float(sp.nu) for sp in specs]
        self._offsets = np.cumsum([0] + [sp.particles.n for sp in specs])
        allp = Particles.concatenate([sp.particles for sp in specs])
        self._init_common(allp.dim, allp.n, allp.c.copy(), identity_mode, p2g_chunk, g2p_threads, graph, autotune,
                          timing, device, fast_svd, svd)
        table = (mpm_scene * len(specs))()
        for i, sp in enumerate(specs):
            table[i] = mpm_scene(int(sp.material), int(sp.res), sp.dt, sp.E, sp.nu, sp.gravity, sp.particles.n)
        h = ctypes.c_void_p()
        check(lib().mpm_create_batch(self._dim, len(specs), table, ptr(allp.x), ptr(allp.v), ptr(allp.F),
                                     ptr(allp.C), ptr(allp.Jp), ptr(allp.mass), ptr(allp.volume),
                                     ctypes.byref(self.options), ctypes.byref(h)))
        self._h = h

    @classmethod
    def from_image(cls, image, scenes: Sequence, *, autotune: bool = False, **kw) -> "MPMBatch":
        """A batch whose particles come from a device state image (scenes concatenated in order).
        Each scene gives its particle count (`n`, or len(particles)) and the ctor arguments
        material, res, dt, E, nu, gravity (defaults as in SceneSpec)."""
        dim = _image_dim(image)
        specs = []
        for s in scenes:
            n = int(s.n) if hasattr(s, "n") and not hasattr(s, "particles") else int(s.particles.n)
            specs.append(SceneSpec(_placeholder(n, dim), MaterialModel(s.material), getattr(s, "res", 64),
                                   getattr(s, "dt", 1e-4), getattr(s, "E", 1e4), getattr(s, "nu", 0.2),
                                   getattr(s, "gravity", -100.0)))
        if sum(sp.particles.n for sp in specs) != int(image.shape[1]):
            raise ValueError("the scenes' particle counts do not add up to the image's columns")
        batch = cls(specs, **kw)
        batch.set_state_image(image)
        if autotune:
            batch.autotune()
        return batch

    @property
    def num_scenes(self) -> int:
        return len(self._specs)

    def scene_range(self, s: int):
        return int(self._offsets[s]), int(self._offsets[s + 1])

    def particles(self, scene: Optional[int] = None) -> Particles:
        """All particles (scenes concatenated in order), or those of one scene."""
        allp = self._all_particles()
        if scene is None:
            return allp
        a, b = self.scene_range(scene)
        return allp.slice(a, b)

    def grid(self, scene: int):
        return self._grid(scene, self._specs[scene].res)

    def params(self, scene: int):
        mu, lam, dx, inv_dx = (ctypes.c_float() for _ in range(4))
        check(lib().mpm_get_scene_params(self._handle(), int(scene), ctypes.byref(mu), ctypes.byref(lam),
                                         ctypes.byref(dx), ctypes.byref(inv_dx)))
        return {"mu_0": np.float32(mu.value), "lambda_0": np.float32(lam.value), "dx": np.float32(dx.value),
                "inv_dx": np.float32(inv_dx.value)}
