"""Scene generators: the reference's cube<dim> lattice and the benchmark workloads.

cube(dim, res, lo, hi) follows src/nclr_math.h:100-129 (a res^dim lattice, last axis
fastest, coordinates from an Eigen LinSpaced(res, lo, hi) sequence in float32).  The
workloads are:

  two_cubes_2d        two 50x50 snow cubes in [0.4,0.6]^2, the second shifted -0.35 in y
                      (src/example.cpp:29-46), res 64
  four_cubes_3d       four 64^3 snow cubes, side 0.2, seeded centres in [0.2,0.8]^3 rejected on
                      overlap, one uniform [-1,1]^3 velocity per cube, res 128
  batched_2d          256 independent 2D scenes of 1-4 cubes (side resolution 20..80, side 0.2,
                      corner in [0.1,0.7]^2), material uniform over snow/jelly/liquid, res 128
  jelly_2d_diff       one 140x140 jelly cube in [0.3,0.5]^2, res 128, E = 5e3
  jelly_3d_batched    8 scenes of one 64^3 jelly cube, side 0.25, corner in [0.2,0.6]^3,
                      E log-uniform in [5e2, 5e4], res 128

Random draws use Marsaglia's xorshift128 (the generator family of src/nclr_math.h:22-31),
seeded through splitmix64, uniform = u32 * 2^-32.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .sim import MaterialModel, Particles


class XorShift128:
    """Marsaglia (2003) xorshift128 on four uint32 words, seeded by splitmix64."""

    def __init__(self, seed: int):
        s = seed & 0xFFFFFFFFFFFFFFFF
        words = []
        for _ in range(2):
            s = (s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
            z ^= z >> 31
            words += [z & 0xFFFFFFFF, z >> 32]
        if not any(words):
            words[0] = 1
        self.x, self.y, self.z, self.w = words

    def next_u32(self) -> int:
        t = (self.x ^ (self.x << 11)) & 0xFFFFFFFF
        self.x, self.y, self.z = self.y, self.z, self.w
        self.w = (self.w ^ (self.w >> 19) ^ (t ^ (t >> 8))) & 0xFFFFFFFF
        return self.w

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * (self.next_u32() * (1.0 / 4294967296.0))

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1)) if hi > lo else lo


def linspaced(n: int, low: float, high: float) -> np.ndarray:
    """Eigen LinSpaced(n, low, high) in float32: low + i*step with step = (high-low)/(n-1), the last
    value exactly `high` (built from the high end when |high| < |low|)."""
    low = np.float32(low)
    high = np.float32(high)
    if n <= 0:
        return np.zeros(0, np.float32)
    if n == 1:
        return np.array([high], np.float32)
    size1 = n - 1
    step = np.float32((high - low) / np.float32(size1))
    i = np.arange(n, dtype=np.float32)
    if abs(high) < abs(low):
        out = (high - np.float32(size1 - np.arange(n)).astype(np.float32) * step).astype(np.float32)
        out[0] = low
    else:
        out = (low + i * step).astype(np.float32)
        out[-1] = high
    return out


def lattice(res_per_axis, lo_per_axis, hi_per_axis) -> np.ndarray:
    """Regular lattice, one LinSpaced range per axis, last axis fastest (src/nclr_math.h:107-122)."""
    axes = [linspaced(r, lo, hi) for r, lo, hi in zip(res_per_axis, lo_per_axis, hi_per_axis)]
    grids = np.meshgrid(*axes, indexing="ij")
    return np.stack(grids, -1).reshape(-1, len(axes)).astype(np.float32)


def cube(dim: int, res: int, lo: float, hi: float) -> np.ndarray:
    """cube<dim>(res, min, max) (src/nclr_math.h:100-129): (res^dim, dim) float32 positions."""
    if dim not in (2, 3):
        raise ValueError("dim must be 2 or 3")
    return lattice([res] * dim, [lo] * dim, [hi] * dim)


@dataclass
class CubeSpec:
    """One lattice cube of a scene: res[a] points on axis a over [lo[a], hi[a]] (converted to float32
    as lattice() does), every particle with velocity v.  The device generator
    (nuclearmpm_amd.generators) rebuilds a scene's particles from these bit for bit."""
    res: List[int]
    lo: List[float]
    hi: List[float]
    v: Optional[List[float]] = None


@dataclass
class Scene:
    """One simulation's inputs."""
    particles: Particles
    material: MaterialModel
    res: int
    dt: float = 1e-4
    E: float = 1e4
    nu: float = 0.2
    gravity: float = -100.0
    steps: int = 1
    name: str = ""
    extra: dict = field(default_factory=dict)


def two_cubes_2d(material: MaterialModel = MaterialModel.kSnow) -> Scene:
    """src/example.cpp:29-46: two cube<2>(50, 0.4, 0.6), the second moved by pos(1) -= 0.35."""
    a = cube(2, 50, 0.4, 0.6)
    b = cube(2, 50, 0.4, 0.6)
    b[:, 1] = (b[:, 1].astype(np.float64) - 0.35).astype(np.float32)  # float -= double
    x = np.concatenate([a, b])
    return Scene(Particles.from_positions(x, c=0xED553B), material, res=64, steps=4000, name="two_cubes_2d")


def _cube_centres(rng: XorShift128, count: int, dim: int, lo: float, hi: float, min_sep: float):
    centres: List[np.ndarray] = []
    tries = 0
    while len(centres) < count:
        c = np.array([rng.uniform(lo, hi) for _ in range(dim)])
        tries += 1
        if all(np.max(np.abs(c - o)) >= min_sep for o in centres):
            centres.append(c)
        if tries > 100000:
            raise RuntimeError("could not place non-overlapping cubes")
    return centres


def four_cubes_3d(seed: int = 0, side_res: int = 64, cubes: int = 4, side: float = 0.2) -> Scene:
    """Four 64^3 snow cubes of side 0.2, centres uniform in [0.2,0.8]^3 rejected on overlap,
    one velocity uniform in [-1,1]^3 per cube; res 128, dt 1e-4, 1000 steps."""
    rng = XorShift128(seed)
    centres = _cube_centres(rng, cubes, 3, 0.2, 0.8, side)
    xs, vs, cubes = [], [], []
    for c in centres:
        v = np.array([rng.uniform(-1.0, 1.0) for _ in range(3)], np.float32)
        x = lattice([side_res] * 3, list(c - side / 2), list(c + side / 2))
        xs.append(x)
        vs.append(np.broadcast_to(v, x.shape))
        cubes.append(CubeSpec([side_res] * 3, list(c - side / 2), list(c + side / 2), list(v)))
    x = np.concatenate(xs)
    p = Particles.from_positions(x, v=np.concatenate(vs))
    return Scene(p, MaterialModel.kSnow, res=128, steps=1000, name="four_cubes_3d",
                 extra={"centres": np.array(centres), "cubes": cubes})


def batched_2d(seed: int = 1, scenes: int = 256, side: float = 0.2) -> List[Scene]:
    """256 scenes of 1-4 cubes each (side resolution uniform in 20..80, side 0.2, corner in
    [0.1, 0.9-side]^2), material uniform over {liquid, snow, jelly}; res 128, 2000 steps."""
    rng = XorShift128(seed)
    out = []
    mats = [MaterialModel.kLiquid, MaterialModel.kSnow, MaterialModel.kJelly]
    for s in range(scenes):
        ncubes = rng.randint(1, 4)
        xs, cubes = [], []
        for _ in range(ncubes):
            r = rng.randint(20, 80)
            corner = [rng.uniform(0.1, 0.9 - side) for _ in range(2)]
            xs.append(lattice([r, r], corner, [c + side for c in corner]))
            cubes.append(CubeSpec([r, r], corner, [c + side for c in corner]))
        mat = mats[min(2, int(rng.uniform() * 3))]
        out.append(Scene(Particles.from_positions(np.concatenate(xs)), mat, res=128, steps=2000,
                         name=f"batched_2d[{s}]", extra={"cubes": cubes}))
    return out


def jelly_2d_diff() -> Scene:
    """One cube<2>(140, 0.3, 0.5) jelly cube, res 128, E = 5e3; 512 steps, checkpoint every 64;
    loss = |centre of mass - (0.7, 0.3)|^2."""
    x = cube(2, 140, 0.3, 0.5)
    return Scene(Particles.from_positions(x), MaterialModel.kJelly, res=128, E=5e3, steps=512,
                 name="jelly_2d_diff", extra={"target": np.array([0.7, 0.3], np.float32), "checkpoint_every": 64,
                                              "cubes": [CubeSpec([140, 140], [0.3, 0.3], [0.5, 0.5])]})


def jelly_3d_batched(seed: int = 2, scenes: int = 8, side_res: int = 64, side: float = 0.25) -> List[Scene]:
    """8 scenes of one 64^3 jelly cube of side 0.25, corner in [0.2,0.6]^3, E log-uniform in
    [5e2, 5e4], nu 0.2, res 128; 256 steps, checkpoint every 32; targets in [0.2,0.8]^3."""
    rng = XorShift128(seed)
    out = []
    for s in range(scenes):
        corner = np.array([rng.uniform(0.2, 0.6) for _ in range(3)])
        E = float(np.exp(rng.uniform(np.log(5e2), np.log(5e4))))
        target = np.array([rng.uniform(0.2, 0.8) for _ in range(3)], np.float32)
        x = lattice([side_res] * 3, list(corner), list(corner + side))
        out.append(Scene(Particles.from_positions(x), MaterialModel.kJelly, res=128, E=E, steps=256,
                         name=f"jelly_3d_batched[{s}]",
                         extra={"target": target, "checkpoint_every": 32,
                                "cubes": [CubeSpec([side_res] * 3, list(corner), list(corner + side))]}))
    return out
