"""The C++ drop-in header (include/nclr/nclr.h) driven the way the reference's example.cpp and
solver.cpp drive MPMSimulation<dim>: the program's own API checks must pass, its scene must equal
the Python scene generator bit for bit, and its one-step result must match the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from parity import assert_parity

from nuclearmpm_amd import scenes

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "dropin_example")


def run_exe(steps, out):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    print(f"[dropin] running {steps} steps", flush=True)
    r = subprocess.run([EXE, str(steps), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    data = np.loadtxt(out, skiprows=1, dtype=np.float64)
    n = data.shape[0]
    return O.State(x=data[:, 0:2], v=data[:, 2:4], F=data[:, 4:8].reshape(n, 2, 2),
                   C=data[:, 8:12].reshape(n, 2, 2), Jp=data[:, 12], mass=np.ones(n), volume=np.ones(n))


def test_dropin_scene_and_one_step(tmp_path):
    s0 = run_exe(0, tmp_path / "s0.txt")
    sc = scenes.two_cubes_2d()
    assert np.array_equal(s0.x.astype(np.float32), sc.particles.x), "cube<2> lattice differs from the scene generator"
    s1 = run_exe(1, tmp_path / "s1.txt")
    ref = O.run(O.State(x=sc.particles.x, v=sc.particles.v, F=sc.particles.F, C=sc.particles.C, Jp=sc.particles.Jp,
                        mass=sc.particles.mass, volume=sc.particles.volume), O.SNOW, res=64, steps=1)
    errs = assert_parity(s1, ref, inv_dx=64.0, label="drop-in C++ one step")
    print(f"[dropin] errors {errs}", flush=True)
