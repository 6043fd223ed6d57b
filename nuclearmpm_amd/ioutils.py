"""Loader for the headless solver's per-step dumps (python/ioutils.py:1-102 of the reference).

The solver (nuclearmpm_amd/solver/nclr_solver.cpp) writes one file per step and field into a
folder: ``{step}_{field}.txt`` in the reference's text format (src/solver.cpp:66-131) or
``{step}_{field}.npy`` with ``--dump-format npy``.  ``process_tmp`` mirrors the reference's
function of the same name (python/ioutils.py:32-102): it returns ``{"{step}_": SimResult}``.
Unlike the reference it reads 3D as well as 2D dumps and any grid resolution (the reference
hard-codes 2 components and a 65 x 65 grid, python/ioutils.py:40-70): the dimension comes from
the line counts of ``x`` and ``Jp`` (or ``{step}_meta.npy``) and the resolution from the node
count.  It writes the reference's ``results.pickle`` only when ``pickle_path`` is given.
``load_trajectory`` stacks all steps into arrays of shape ``[T, ...]`` (numpy or torch).

Text files are parsed with ``numpy.fromfile(sep=" ")``, which reads whitespace-separated
numbers (newlines included) in one pass; a field's values come out in file order, i.e. one
particle (or node) after another, matrix rows first.
"""
from __future__ import annotations

import os
import pickle
import re
from collections import defaultdict
from dataclasses import dataclass, field
from typing import Dict, Iterable, List, Optional, Tuple

import numpy as np

PARTICLE_FIELDS = ("timestep", "x", "v", "F", "C", "Jp", "lame")
GRID_FIELDS = ("mass", "velocity")
FIELDS = PARTICLE_FIELDS + GRID_FIELDS

_NAME = re.compile(r"^(\d+)_([A-Za-z]+)\.(txt|npy)$")


def to_3d(idx: int, max_: int) -> Tuple[int, int]:
    """python/ioutils.py:12-16: (x, y) of the flat 2D index ``idx`` with row length ``max_``."""
    x = idx // max_
    idx -= x * max_
    return x, idx % max_


def to_nd(idx: int, side: int, dim: int) -> Tuple[int, ...]:
    """Node coordinates of the flat, x-major grid index ``idx`` (``side`` = res + 1 nodes per axis)."""
    out = []
    for _ in range(dim):
        out.append(idx % side)
        idx //= side
    return tuple(reversed(out))


class SimResult:
    """One step of a dump (python/ioutils.py:19-29).  Fields present in the folder are set:
    x, v (n, d); F, C (n, d, d); Jp (n,); lame (n, 2) = (mu, lambda) per particle; timestep (n,)
    = the step's time per particle; mass (side,)*d and velocity (side,)*d + (d,) over the grid
    nodes, side = res + 1, indexed [i, j(, k)] by node coordinates.  ``time`` is the step's time
    as a scalar, ``dim`` and ``res`` what the loader inferred."""

    def __init__(self):
        self.x: np.ndarray
        self.v: np.ndarray
        self.F: np.ndarray
        self.C: np.ndarray
        self.Jp: np.ndarray
        self.lame: np.ndarray
        self.velocity: np.ndarray
        self.mass: np.ndarray
        self.timestep: np.ndarray
        self.time: float
        self.dim: int
        self.res: int

    def __repr__(self) -> str:
        shapes = {k: getattr(v, "shape", v) for k, v in self.__dict__.items()}
        return f"SimResult({shapes})"


def _read(path: str) -> np.ndarray:
    if path.endswith(".npy"):
        return np.load(path, allow_pickle=False)
    return np.fromfile(path, dtype=np.float64, sep=" ")


def _infer_dim(files: Dict[str, str]) -> Optional[int]:
    if "meta" in files:
        return int(_read(files["meta"])[2])
    if "x" in files and "Jp" in files:
        x, jp = _read(files["x"]), _read(files["Jp"])
        if jp.size:
            d = x.size // jp.size
            if d in (2, 3) and d * jp.size == x.size:
                return d
    if "F" in files and "x" in files:
        x, F = _read(files["x"]), _read(files["F"])
        if x.size:
            d = F.size // x.size
            if d in (2, 3):
                return d
    return None


def _grid_side(nodes: int, dim: int) -> int:
    side = int(round(nodes ** (1.0 / dim)))
    for s in (side - 1, side, side + 1):
        if s > 0 and s ** dim == nodes:
            return s
    raise ValueError(f"{nodes} grid nodes are not a {dim}D cube of nodes")


def load_step(files: Dict[str, str], dim: Optional[int] = None, res: Optional[int] = None) -> SimResult:
    """Reads one step from ``{field: path}`` (the files of that step)."""
    out = SimResult()
    d = dim if dim is not None else _infer_dim(files)
    if d is None:
        raise ValueError("cannot infer the dimension: pass dim=2 or dim=3")
    out.dim = d
    meta = _read(files["meta"]) if "meta" in files else None
    for key, path in files.items():
        if key == "meta":
            continue
        a = _read(path)
        npy = path.endswith(".npy")
        if key in ("x", "v"):
            a = a.reshape(-1, d)
        elif key in ("F", "C"):
            a = a.reshape(-1, d, d)
        elif key == "lame":
            a = a.reshape(-1, 2)
        elif key in ("Jp", "timestep"):
            a = a.reshape(-1)
        elif key in ("mass", "velocity"):
            nodes = a.size if key == "mass" else a.size // d
            side = res + 1 if res is not None else _grid_side(nodes, d)
            if side ** d != nodes:
                raise ValueError(f"{path}: {nodes} nodes do not match res={side - 1} in {d}D")
            a = a.reshape((side,) * d + ((d,) if key == "velocity" else ()))
            out.res = side - 1
        else:
            raise ValueError(f"ValueKey {key} is invalid")
        if npy:
            a = np.ascontiguousarray(a)
        setattr(out, key, a)
    if meta is not None:
        out.time = float(meta[1])
        out.res = int(meta[3])
        if "lame" not in files:
            n = int(meta[4])
            out.lame = np.tile(np.asarray(meta[5:7], dtype=np.float32), (n, 1))
        if "timestep" not in files:
            out.timestep = np.full(int(meta[4]), meta[1], dtype=np.float32)
    elif "timestep" in files and out.timestep.size:
        out.time = float(out.timestep[0])
    return out


def scan(tmp: str) -> Dict[int, Dict[str, str]]:
    """``{step: {field: path}}`` of the dump files in ``tmp`` (other files are ignored)."""
    steps: Dict[int, Dict[str, str]] = defaultdict(dict)
    for name in os.listdir(tmp):
        m = _NAME.match(name)
        if m is None:
            continue
        step, key, _ = m.groups()
        if key not in FIELDS and key != "meta":
            continue
        steps[int(step)][key] = os.path.join(tmp, name)
    return dict(sorted(steps.items()))


def process_tmp(tmp: str, dim: Optional[int] = None, res: Optional[int] = None,
                pickle_path: Optional[str] = None) -> Dict[str, SimResult]:
    """python/ioutils.py:32-102: every step of the dump folder ``tmp`` as ``{"{step}_": SimResult}``,
    in step order.  ``dim``/``res`` override the inferred dimension and grid resolution."""
    results: Dict[str, SimResult] = {}
    for step, files in scan(tmp).items():
        results[f"{step}_"] = load_step(files, dim, res)
    if pickle_path is not None:
        with open(pickle_path, "wb") as f:
            pickle.dump(results, f)
    return results


@dataclass
class Trajectory:
    """All steps of a dump stacked on a leading axis: ``fields[name]`` has shape [T, ...]."""

    steps: List[int]
    time: np.ndarray
    dim: int
    res: Optional[int]
    fields: Dict[str, object] = field(default_factory=dict)

    def __getattr__(self, name):
        f = self.__dict__.get("fields", {})
        if name in f:
            return f[name]
        raise AttributeError(name)


def load_trajectory(tmp: str, fields: Iterable[str] = FIELDS, dim: Optional[int] = None,
                    res: Optional[int] = None, as_torch: bool = False, device=None) -> Trajectory:
    """Every step of ``tmp`` stacked: ``traj.x`` is [T, n, d], ``traj.mass`` [T, side, side(, side)]
    and so on for each requested field present in every step; ``as_torch`` returns torch tensors
    (float32, on ``device`` if given)."""
    table = scan(tmp)
    if not table:
        raise FileNotFoundError(f"no dump files in {tmp}")
    steps = list(table)
    wanted = [f for f in fields if all(f in table[s] for s in steps) or
              (f in ("lame", "timestep") and all("meta" in table[s] for s in steps))]
    per = []
    for s in steps:
        files = {k: v for k, v in table[s].items() if k in wanted or k == "meta"}
        per.append(load_step(files, dim, res))
    d = per[0].dim
    traj = Trajectory(steps=steps, time=np.array([getattr(p, "time", np.nan) for p in per]), dim=d,
                      res=getattr(per[0], "res", res))
    for f in wanted:
        a = np.stack([np.asarray(getattr(p, f), dtype=np.float32) for p in per])
        if as_torch:
            import torch

            t = torch.from_numpy(a)
            a = t.to(device) if device is not None else t
        traj.fields[f] = a
    return traj
