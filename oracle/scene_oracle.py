"""Test infrastructure only: CPU restatements of the reference's scene helpers and viewer
projection (src/nclr_math.h:22-39, 100-167, 173-195), the checker of the device generators in
nuclearmpm_amd/csrc/mpm_scene.hip.  Only tests/ import this module.

What pins each function (tests/test_scene_oracle.py):
  nc_rand_int   Marsaglia's published xor128 outputs for the reference's seeds
                (3701687786, 458299110, 2500872618, ...).
  linspaced     endpoints exact, constant step to rounding, the flipped branch for |hi| < |lo|,
                and Eigen's documented LinSpaced(1, lo, hi) = hi.
  cube          its point count, order (last axis fastest) and the linspaced axes.
  sample        the stream above applied as (u * 2 - 1) * 0.08 + centre.
  gyroid        closed forms: gyroid(k, t, 0) = -t, gyroid(k, t, (k/4, 0, 0)) = 1 - t,
                periodicity with period k on every axis.
  pt_3d_to_2d   (0,0,0) -> (0.5, 0.5); the (x, z) rotation preserves x^2 + z^2; the angles
                from glibc's cosf / sinf, the functions std::cos(float) / std::sin(float) call.
  igl_grid      libigl's documented example layout (first coordinate fastest).
"""
from __future__ import annotations

import ctypes
import ctypes.util
import math
from typing import List, Sequence

import numpy as np

SEED = (123456789, 362436069, 521288629, 88675123)
_M32 = 0xFFFFFFFF


class NcRand:
    """nc_rand_int / nc_rand (src/nclr_math.h:22-31): xorshift128, static state seeded SEED."""

    def __init__(self, state: Sequence[int] = SEED):
        self.x, self.y, self.z, self.w = (int(s) & _M32 for s in state)

    @property
    def state(self):
        return (self.x, self.y, self.z, self.w)

    def next_u32(self) -> int:
        t = (self.x ^ (self.x << 11)) & _M32
        self.x, self.y, self.z = self.y, self.z, self.w
        self.w = ((self.w ^ (self.w >> 19)) ^ (t ^ (t >> 8))) & _M32
        return self.w

    def u32(self, count: int) -> np.ndarray:
        return np.array([self.next_u32() for _ in range(count)], dtype=np.uint32)

    def uniform(self, count: int) -> np.ndarray:
        """nc_rand(): float(u32) * (1.0f / 4294967296.0f), in float32."""
        return self.u32(count).astype(np.float32) * np.float32(1.0 / 4294967296.0)


def linspaced(n: int, lo: float, hi: float) -> np.ndarray:
    """Eigen::Vector<float,-1>::LinSpaced(n, lo, hi) (used by cube, src/nclr_math.h:103,110)."""
    lo, hi = np.float32(lo), np.float32(hi)
    if n == 1:
        return np.array([hi], np.float32)
    size1 = n - 1
    step = np.float32(np.float32(hi - lo) / np.float32(size1))
    out = np.empty(n, np.float32)
    for i in range(n):
        if abs(hi) < abs(lo):
            out[i] = lo if i == 0 else np.float32(hi - np.float32(np.float32(size1 - i) * step))
        else:
            out[i] = hi if i == size1 else np.float32(lo + np.float32(np.float32(i) * step))
    return out


def cube(dim: int, res: Sequence[int], lo: Sequence[float], hi: Sequence[float]) -> np.ndarray:
    """cube<dim> (src/nclr_math.h:100-129) with per-axis counts and bounds: last axis fastest."""
    axes = [linspaced(int(res[a]), lo[a], hi[a]) for a in range(dim)]
    mesh = np.meshgrid(*axes, indexing="ij")
    return np.stack(mesh, -1).reshape(-1, dim).astype(np.float32)


def sample(dim: int, centers: np.ndarray, rng: NcRand, per_center: int = 1000, half: float = 0.08) -> np.ndarray:
    """sample<dim>(center) (src/nclr_math.h:141-147) per centre, with the points returned:
    (randvec<dim>() * 2 - 1) * half + center, each operation rounded to float32."""
    centers = np.asarray(centers, np.float32).reshape(-1, dim)
    out = []
    for c in centers:
        u = rng.uniform(per_center * dim).reshape(per_center, dim)
        out.append(((u * np.float32(2.0) - np.float32(1.0)) * np.float32(half) + c).astype(np.float32))
    return np.concatenate(out) if out else np.zeros((0, dim), np.float32)


def gyroid(k: float, t: float, pos: np.ndarray) -> np.ndarray:
    """gyroid(k, t, pos) (src/nclr_math.h:131-139): double sin / cos (the C library's, as std::sin
    on a double calls), the sum in double, rounded to float."""
    two_pi = (2.0 * math.pi) / float(np.float32(k))
    out = np.empty(len(pos), np.float32)
    for i, p in enumerate(np.asarray(pos, np.float32)):
        x, y, z = (two_pi * float(v) for v in p)
        g = math.sin(x) * math.cos(y) + math.sin(y) * math.cos(z) + math.sin(z) * math.cos(x) - float(np.float32(t))
        out[i] = np.float32(g)
    return out


def _libm():
    return ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")


def projection_constants():
    """cos / sin of to_radians(28) and to_radians(32) in float (src/nclr_math.h:149-160)."""
    m = _libm()
    for f in ("cosf", "sinf"):
        getattr(m, f).argtypes = [ctypes.c_float]
        getattr(m, f).restype = ctypes.c_float
    conv = np.float32(math.pi / 180.0)
    phi, theta = np.float32(np.float32(28.0) * conv), np.float32(np.float32(32.0) * conv)
    return (np.float32(m.cosf(phi)), np.float32(m.sinf(phi)), np.float32(m.cosf(theta)), np.float32(m.sinf(theta)))


def pt_3d_to_2d(points: np.ndarray, centered: bool = False) -> np.ndarray:
    """pt_3d_to_2d (src/nclr_math.h:154-167) in float32; centered applies it to point - 0.5."""
    c, s, C, S = projection_constants()
    p = np.asarray(points, np.float32).reshape(-1, 3)
    if centered:
        p = (p - np.float32(0.5)).astype(np.float32)
    x = (p[:, 0] * c).astype(np.float32) + (p[:, 2] * s).astype(np.float32)
    z = (p[:, 2] * c).astype(np.float32) - (p[:, 0] * s).astype(np.float32)
    v = (p[:, 1] * C).astype(np.float32) + (z * S).astype(np.float32)
    return np.stack([x + np.float32(0.5), v + np.float32(0.5)], -1).astype(np.float32)


def igl_grid(res: Sequence[int]) -> np.ndarray:
    """nclr::grid / igl::grid (src/nclr_math.h:173-195): prod(res) rows, first coordinate fastest,
    GV(i, c) = sub_c / (res_c - 1)."""
    res = [int(r) for r in res]
    rows: List[List[float]] = []
    sub = [0] * len(res)
    for _ in range(int(np.prod(res))):
        for c in range(len(res) - 1):
            if sub[c] >= res[c]:
                sub[c] = 0
                sub[c + 1] += 1
        rows.append([sub[c] / (res[c] - 1) for c in range(len(res))])
        sub[0] += 1
    return np.array(rows, np.float64)
