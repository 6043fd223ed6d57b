"""The headless solver (nuclearmpm_amd/solver/nclr_solver.cpp, the command line of
src/solver.cpp:25-217) on the GPU: its dumps, read back with nuclearmpm_amd.ioutils, must hold the
reference's cube scene, step-to-step states that match the oracle's one-step results, the grid
of the previous step and the Lame values of src/solver.cpp:100-105."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from parity import assert_parity

from nuclearmpm_amd import ioutils, scenes

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "nuclearmpm_amd", "solver",
                   "nclr_solver")
MATERIAL = {"jelly": O.JELLY, "snow": O.SNOW, "liquid": O.LIQUID}
HARDENING = {"jelly": 0.3, "snow": 10.0, "liquid": 1.0}


def solve(out, *args):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    cmd = [EXE, *map(str, args), "--out", str(out)]
    print("[solver]", " ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    return r.stdout


def state_of(r):
    n = r.x.shape[0]
    return O.State(x=r.x, v=r.v, F=r.F, C=r.C, Jp=r.Jp, mass=np.ones(n, np.float32), volume=np.ones(n, np.float32))


def check_run(res, dim, material, cubes, E, nu, steps, grid_res=64, identity_mode=0):
    assert len(res) == steps
    lat = np.concatenate([scenes.cube(dim, cube_res, lo, hi) for cube_res, lo, hi in cubes])
    r0 = res["0_"]
    assert np.array_equal(r0.x, lat), "step 0 is not the cube<dim> lattice of the flags"
    F0 = np.eye(dim, dtype=np.float32)
    if dim == 3 and identity_mode == 0:
        F0[2, 2] = 0.0  # diag<3>(1) of the reference is diag(1, 1, 0) (src/nclr_math.h:40-48)
    assert not r0.v.any() and np.array_equal(r0.F, np.broadcast_to(F0, r0.F.shape))
    assert not r0.mass.any() and not r0.velocity.any(), "the grid dump of step 0 must be empty"
    mu0 = np.float32(E) / (np.float32(2) * (np.float32(1) + np.float32(nu)))
    lam0 = np.float32(E) * np.float32(nu) / ((np.float32(1) + np.float32(nu)) * (np.float32(1) - np.float32(2) * np.float32(nu)))
    e = np.float32(HARDENING[material])
    np.testing.assert_allclose(r0.lame[0], [mu0 * e, lam0 * e], rtol=1e-6)
    n = lat.shape[0]
    for s in range(steps - 1):
        a, b = res[f"{s}_"], res[f"{s + 1}_"]
        ref, gv, gm = O.run(state_of(a), MATERIAL[material], res=grid_res, E=E, nu=nu, steps=1, return_grid=True,
                            identity_mode=identity_mode)
        errs = assert_parity(state_of(b), ref, inv_dx=float(grid_res), label=f"solver {dim}D {material} step {s}")
        side = grid_res + 1
        gm = gm.reshape((side,) * dim)
        gv = gv.reshape((side,) * dim + (dim,))
        merr = np.abs(b.mass - gm).max() / np.abs(gm).max()
        verr = np.abs(b.velocity - gv).max() / np.abs(gv).max()
        assert merr <= 1e-5 and verr <= 1e-5, (s, merr, verr)
        assert b.mass.sum() == pytest.approx(n, rel=1e-5)
        print(f"[solver] {dim}D {material} step {s}: {errs} grid {merr:.2e} {verr:.2e}", flush=True)


def test_solver_2d_jelly_npy_and_text(tmp_path):
    args = ["--steps", 4, "--cubes", 2, "--cube-res", 12, "--cube0-x", 0.4, "--cube0-y", 0.55,
            "--cube1-x", 0.2, "--cube1-y", 0.32, "--dump"]
    out = solve(tmp_path / "npy", *args, "--dump-format", "npy")
    assert "Simulation done" in out
    res = ioutils.process_tmp(str(tmp_path / "npy"))
    check_run(res, 2, "jelly", [(12, 0.4, 0.55), (12, 0.2, 0.32)], 1000.0, 0.3, 4)
    solve(tmp_path / "txt", *args)
    assert sorted(os.listdir(tmp_path / "txt"))[:3] == ["0_C.txt", "0_F.txt", "0_Jp.txt"]
    check_text(ioutils.process_tmp(str(tmp_path / "txt")), res)


def check_text(txt, npy):
    """The text dump of a run holds the .npy dump's values to 6 significant digits."""
    assert list(txt) == list(npy)
    for key in npy:
        for f in ("x", "v", "F", "C", "Jp", "mass", "velocity", "lame", "timestep"):
            a, b = getattr(txt[key], f), getattr(npy[key], f)
            assert a.shape == b.shape, (key, f)
            np.testing.assert_allclose(a, b, rtol=6e-6, atol=0, err_msg=f"{key}{f}")


def test_solver_3d_snow(tmp_path):
    """--dim 3 runs (src/solver.cpp:211-216 builds an empty 3D simulation); min/max spellings."""
    solve(tmp_path, "--dim", 3, "--material-model", "snow", "--E", 10000, "--nu", 0.2, "--steps", 3,
          "--cube-res", 10, "--cube0-min", 0.3, "--cube0-max", 0.5, "--dump", "--dump-format", "npy")
    res = ioutils.process_tmp(str(tmp_path))
    check_run(res, 3, "snow", [(10, 0.3, 0.5)], 10000.0, 0.2, 3)
    traj = ioutils.load_trajectory(str(tmp_path))
    assert traj.x.shape == (3, 1000, 3) and traj.mass.shape == (3, 65, 65, 65)


def test_solver_3d_jelly_identity_text(tmp_path):
    """--identity-mode 1 (F = I in 3D), two cubes, text and .npy dumps of 3D states."""
    args = ["--dim", 3, "--identity-mode", 1, "--steps", 3, "--cubes", 2, "--cube-res", 8,
            "--cube0-x", 0.25, "--cube0-y", 0.4, "--cube1-x", 0.5, "--cube1-y", 0.7, "--dump"]
    solve(tmp_path / "npy", *args, "--dump-format", "npy")
    res = ioutils.process_tmp(str(tmp_path / "npy"))
    check_run(res, 3, "jelly", [(8, 0.25, 0.4), (8, 0.5, 0.7)], 1000.0, 0.3, 3, identity_mode=1)
    solve(tmp_path / "txt", *args)
    check_text(ioutils.process_tmp(str(tmp_path / "txt")), res)


def test_solver_no_dump_and_errors(tmp_path):
    out = solve(tmp_path, "--steps", 50, "--cube-res", 20, "--cube0-x", 0.3, "--cube0-y", 0.6,
                "--material-model", "liquid")
    assert "Simulation done" in out and not os.path.exists(tmp_path / "0_x.txt")
    r = subprocess.run([EXE, "--steps", "1", "--cubes", "2", "--cube0-x", "0.3", "--cube0-y", "0.6"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Cube: 1 is missing coordinates" in r.stderr
    r = subprocess.run([EXE, "--material-model", "sand"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Invalid Option: sand" in r.stderr
