"""Scene generators on the device: the reference's scene helpers (src/nclr_math.h:22-39, 100-147)
as HIP kernels that write particle state straight into HBM (torch tensors on a ROCm device).

A generator returns a *state image*: a float32 tensor of shape (2d + 2d^2 + 1, n), rows x[d], v[d],
F[d*d], C[d*d] (matrices row-major), Jp, column i = particle i, with F = diag<d>(1) (diag(1, 1, 0)
in 3D for identity_mode 0, as the reference's Particle ctor builds it; I for identity_mode 1),
C = 0 and Jp = 1.  ``MPMSimulation.from_image`` / ``MPMBatch.from_image`` start an engine from one.

    img = generators.cubes(3, scenes.four_cubes_3d().extra["cubes"])   # the benchmark scene
    sim = MPMSimulation.from_image(img, MaterialModel.kSnow, res=128)

Random draws reproduce the reference's global nc_rand stream (xorshift128 seeded
123456789, 362436069, 521288629, 88675123; uniform = u32 * 2^-32 in float) through an explicit
``NcRand`` state: successive calls continue the stream as successive calls of the reference's
randvec / sample do.
"""
from __future__ import annotations

import ctypes
from typing import Iterable, Optional, Sequence

import numpy as np

from ._lib import check, lib, mpm_cube

REFERENCE_SEED = (123456789, 362436069, 521288629, 88675123)


def _torch():
    import torch

    return torch


def channels(dim: int) -> int:
    """Rows of a state image: x, v (dim each), F, C (dim^2 each), Jp."""
    return 2 * dim + 2 * dim * dim + 1


def _device_index(device) -> int:
    torch = _torch()
    d = torch.device("cuda", 0) if device is None else torch.device(device)
    if d.type != "cuda":
        raise ValueError("generators write device memory: pass a cuda device")
    return 0 if d.index is None else d.index


def _ready(device_index: int) -> None:
    # the kernels run on the device's null stream; inputs written by torch must be complete
    _torch().cuda.synchronize(device_index)


def _cube_struct(spec, dim: int) -> mpm_cube:
    if isinstance(spec, mpm_cube):
        return spec
    res, lo, hi = (spec.res, spec.lo, spec.hi) if hasattr(spec, "res") else (spec[0], spec[1], spec[2])
    v = getattr(spec, "v", None) if hasattr(spec, "res") else (spec[3] if len(spec) > 3 else None)
    res = [int(res)] * dim if np.isscalar(res) else [int(r) for r in res]
    lo = [float(lo)] * dim if np.isscalar(lo) else [float(a) for a in lo]
    hi = [float(hi)] * dim if np.isscalar(hi) else [float(a) for a in hi]
    v = [0.0] * dim if v is None else [float(a) for a in v]
    if not (len(res) == len(lo) == len(hi) == len(v) == dim):
        raise ValueError(f"cube {spec!r} does not have {dim} axes")
    c = mpm_cube()
    for a in range(dim):
        c.res[a], c.lo[a], c.hi[a], c.v[a] = res[a], lo[a], hi[a], v[a]
    for a in range(dim, 3):
        c.res[a] = 1
    return c


def cubes(dim: int, specs: Iterable, identity_mode: int = 0, device=None):
    """Lattice cubes one after another: each spec is a scenes.CubeSpec or (res, lo, hi[, v]) with
    scalars (a cube<dim>(res, min, max) of src/nclr_math.h:100-129) or per-axis lists."""
    torch = _torch()
    dev = _device_index(device)
    table = [_cube_struct(s, dim) for s in specs]
    n = int(sum(int(np.prod([c.res[a] for a in range(dim)])) for c in table))
    img = torch.empty((channels(dim), n), dtype=torch.float32, device=f"cuda:{dev}")
    arr = (mpm_cube * max(1, len(table)))(*table)
    _ready(dev)
    check(lib().mpm_gen_cubes(dim, len(table), arr, int(identity_mode), n, img.data_ptr() if n else None, dev))
    return img


def cube(dim: int, res: int, lo: float, hi: float, v: Optional[Sequence[float]] = None, identity_mode: int = 0,
         device=None):
    """cube<dim>(res, lo, hi) (src/nclr_math.h:100-129) as a state image of res^dim particles."""
    return cubes(dim, [(res, lo, hi, v)], identity_mode, device)


def scene_image(scene, identity_mode: int = 0, device=None):
    """The particles of a nuclearmpm_amd.scenes scene (or a list of scenes, concatenated) rebuilt on
    the device from the cube descriptors the scene records in extra["cubes"]."""
    scenes = scene if isinstance(scene, (list, tuple)) else [scene]
    specs = []
    for s in scenes:
        if "cubes" not in s.extra:
            raise ValueError(f"scene {s.name!r} records no cube descriptors")
        specs.extend(s.extra["cubes"])
    return cubes(scenes[0].particles.dim, specs, identity_mode, device)


class NcRand:
    """State of the reference's nc_rand_int stream (src/nclr_math.h:22-29), default-seeded like it."""

    def __init__(self, state: Sequence[int] = REFERENCE_SEED):
        self._state = (ctypes.c_uint32 * 4)(*[int(w) & 0xFFFFFFFF for w in state])

    @property
    def state(self):
        return tuple(int(w) for w in self._state)

    def advance(self, k: int) -> None:
        """Skips k draws."""
        check(lib().mpm_rand_advance(self._state, ctypes.c_uint64(int(k))))

    def uniform(self, count: int, device=None):
        """The next `count` nc_rand() values as a float32 device tensor."""
        torch = _torch()
        dev = _device_index(device)
        out = torch.empty(int(count), dtype=torch.float32, device=f"cuda:{dev}")
        _ready(dev)
        check(lib().mpm_gen_rand(int(count), self._state, out.data_ptr() if count else None, dev))
        return out


def sample(dim: int, centers, rng: Optional[NcRand] = None, per_center: int = 1000, half_width: float = 0.08,
           identity_mode: int = 0, device=None):
    """sample<dim>(center) (src/nclr_math.h:141-147) for every centre, concatenated: per_center
    particles at (randvec<dim>() * 2 - 1) * half_width + center, drawn from `rng` (a fresh
    reference-seeded stream if None)."""
    torch = _torch()
    dev = _device_index(device)
    rng = NcRand() if rng is None else rng
    c = np.ascontiguousarray(np.asarray(centers, np.float32).reshape(-1, dim))
    n = c.shape[0] * int(per_center)
    img = torch.empty((channels(dim), n), dtype=torch.float32, device=f"cuda:{dev}")
    _ready(dev)
    check(lib().mpm_gen_sample(dim, c.shape[0], c.ctypes.data, int(per_center), float(half_width), rng._state,
                               int(identity_mode), img.data_ptr() if n else None, dev))
    return img


def gyroid(pos, k: float, t: float):
    """gyroid(k, t, pos) (src/nclr_math.h:131-139) of a (n, 3) float32 device tensor of positions."""
    torch = _torch()
    pos = pos.contiguous()
    if pos.dtype != torch.float32 or pos.dim() != 2 or pos.shape[1] != 3 or pos.device.type != "cuda":
        raise ValueError("pos must be a (n, 3) float32 cuda tensor")
    out = torch.empty(pos.shape[0], dtype=torch.float32, device=pos.device)
    dev = pos.device.index or 0
    _ready(dev)
    check(lib().mpm_gyroid(pos.shape[0], pos.data_ptr() if pos.shape[0] else None, float(k), float(t),
                           out.data_ptr() if pos.shape[0] else None, dev))
    return out


def gyroid_solid(res, lo, hi, k: float, t: float = 0.0, thickness: float = 0.0, v=None, identity_mode: int = 0,
                 device=None):
    """The points of the 3D lattice (res, lo, hi) inside the gyroid: gyroid(k, t, p) <= 0, or
    |gyroid(k, t, p)| <= thickness for a sheet of that half-width; lattice order."""
    torch = _torch()
    dev = _device_index(device)
    box = _cube_struct((res, lo, hi, v), 3)
    n = ctypes.c_int64()
    check(lib().mpm_gen_gyroid(ctypes.byref(box), float(k), float(t), float(thickness), int(identity_mode), None, 0,
                               ctypes.byref(n), dev))
    img = torch.empty((channels(3), n.value), dtype=torch.float32, device=f"cuda:{dev}")
    if n.value:
        _ready(dev)
        check(lib().mpm_gen_gyroid(ctypes.byref(box), float(k), float(t), float(thickness), int(identity_mode),
                                   img.data_ptr(), n.value, ctypes.byref(n), dev))
    return img


def positions(img, dim: int):
    """The (n, dim) positions of a state image (a view)."""
    return img[:dim].t()


def velocities(img, dim: int):
    return img[dim:2 * dim].t()


def concatenate(images: Sequence):
    """State images of one dimension side by side (particles of the first come first)."""
    return _torch().cat(list(images), dim=1)
