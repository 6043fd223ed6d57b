"""Viewer hooks: the reference's pt_3d_to_2d projection (src/nclr_math.h:149-167) on the device,
its igl-style lattice helper nclr::grid (src/nclr_math.h:173-195) and a grid export for plotting.
The GUI itself (src/taichi.h) is not part of this package.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from ._lib import check, lib


def pt_3d_to_2d(points, centered: bool = False, device=None):
    """pt_3d_to_2d of (n, 3) points on the device: (x0 c + x2 s + 0.5, x1 C + (x2 c - x0 s) S + 0.5)
    with c, s = cos, sin of 28 degrees and C, S of 32 degrees, in FP32.  centered=False is what
    the reference computes (its centred copy of the point is never used); centered=True rotates
    about (0.5, 0.5, 0.5).  A numpy input returns numpy; a cuda tensor returns a cuda tensor."""
    import torch

    as_numpy = not isinstance(points, torch.Tensor)
    if as_numpy:
        dev = torch.device("cuda", 0) if device is None else torch.device(device)
        pts = torch.as_tensor(np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3), device=dev)
    else:
        if points.device.type != "cuda":
            raise ValueError("points must be a cuda tensor or a numpy array")
        pts = points.to(torch.float32).reshape(-1, 3).contiguous()
    out = torch.empty((pts.shape[0], 2), dtype=torch.float32, device=pts.device)
    torch.cuda.synchronize(pts.device)
    n = pts.shape[0]
    check(lib().mpm_project_points(n, pts.data_ptr() if n else None, int(bool(centered)),
                                   out.data_ptr() if n else None, pts.device.index or 0))
    return out.cpu().numpy() if as_numpy else out


def grid_points(res: Sequence[int]) -> np.ndarray:
    """nclr::grid (src/nclr_math.h:173-195, after igl::grid): prod(res) points of the unit box,
    row i = (sub_0 / (res_0 - 1), sub_1 / (res_1 - 1), ...) with the first coordinate fastest."""
    res = [int(r) for r in res]
    axes = [np.arange(r, dtype=np.float64) / (r - 1) for r in res]
    mesh = np.meshgrid(*axes[::-1], indexing="ij")  # last axis slowest -> reverse, then flip back
    return np.stack([m.reshape(-1) for m in mesh[::-1]], -1)


def export_grid(sim, scene: Optional[int] = None, active_only: bool = True):
    """The grid after the last step of a simulation (or of one scene of a batch) for plotting:
    node positions (m, d) = index * dx, mass (m,) and velocity (m, d), x-major node order; with
    active_only only nodes with mass > 0."""
    if scene is None:
        vel, mass = sim.grid()
    else:
        vel, mass = sim.grid(scene)
    d = mass.ndim
    side = mass.shape[0]
    dx = 1.0 / (side - 1)
    idx = np.stack(np.meshgrid(*[np.arange(side)] * d, indexing="ij"), -1).reshape(-1, d)
    pos = (idx * dx).astype(np.float32)
    m = mass.reshape(-1)
    v = vel.reshape(-1, d)
    if active_only:
        keep = m > 0
        return {"position": pos[keep], "mass": m[keep], "velocity": v[keep]}
    return {"position": pos, "mass": m, "velocity": v}
