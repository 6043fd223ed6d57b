"""Dump format (include/nclr/nclr_dump.h) and its loader (nuclearmpm_amd/ioutils.py), on the CPU.

A small C++ program (tests/cpp/dump_writer.cpp) writes known values through the same writer the
headless solver uses; the loader must give them back: to the 6 significant digits of the
reference's text format (src/solver.cpp:66-131 prints with the stream's default precision) and
exactly for the .npy format.
"""
import os
import pickle
import subprocess

import numpy as np
import pytest

from nuclearmpm_amd import ioutils

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def value(s, f, i):
    """dump_writer.cpp's value(): element i of field f at step s."""
    h = (i * 2654435761 + f * 97 + s * 7919) % 20011
    mag = 10.0 ** (float(h % 9) - 4.0)
    return np.float32((h / 20011.0 - 0.5) * mag)


def field(s, f, count):
    return np.array([value(s, f, i) for i in range(count)], dtype=np.float32)


@pytest.fixture(scope="module")
def writer(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("bin") / "dump_writer")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include", "nclr"), "-o", exe,
                    os.path.join(HERE, "cpp", "dump_writer.cpp")], check=True)
    return exe


def test_text_format_matches_eigen_stream_output(writer):
    """Eigen's operator<< with the default IOFormat: precision 6, every coefficient right-aligned
    to the widest one of the matrix, one space between columns, one row per line."""
    out = subprocess.run([writer, "format"], check=True, capture_output=True, text=True).stdout
    blocks = out.split("|\n")[:-1]
    assert blocks == [
        "  0.5\n-0.25",
        "   1.5     -2\n 1e-07 100000",
        "1.23457e+06\n        0.1\n         -3",
        "384.615 576.923",
    ]


def expected(s, d, n, res):
    nodes = (res + 1) ** d
    return dict(x=field(s, 0, n * d).reshape(n, d), v=field(s, 1, n * d).reshape(n, d),
                F=field(s, 2, n * d * d).reshape(n, d, d), C=field(s, 3, n * d * d).reshape(n, d, d),
                Jp=field(s, 4, n), mass=field(s, 5, nodes).reshape((res + 1,) * d),
                velocity=field(s, 6, nodes * d).reshape((res + 1,) * d + (d,)))


@pytest.mark.parametrize("fmt", ["text", "npy"])
@pytest.mark.parametrize("d", [2, 3])
def test_round_trip(writer, tmp_path, fmt, d):
    n, res, steps = 7, 4, 3
    subprocess.run([writer, str(tmp_path), fmt, str(d), str(n), str(res), str(steps)], check=True)
    results = ioutils.process_tmp(str(tmp_path))
    assert list(results) == [f"{s}_" for s in range(steps)]
    mu, lam = np.float32(np.float32(384.615) * np.float32(0.3)), np.float32(np.float32(576.923) * np.float32(0.3))
    for s in range(steps):
        r = results[f"{s}_"]
        assert r.dim == d and r.res == res
        for k, want in expected(s, d, n, res).items():
            got = getattr(r, k)
            assert got.shape == want.shape, k
            if fmt == "npy":
                assert got.dtype == np.float32 and np.array_equal(got, want), k
            else:
                np.testing.assert_allclose(got, want, rtol=6e-6, atol=0, err_msg=k)
        t = np.float32(1e-4) * np.float32(max(s, 1))
        assert r.timestep.shape == (n,) and np.allclose(r.timestep, t, rtol=6e-6)
        assert r.time == pytest.approx(float(t), rel=6e-6)
        assert r.lame.shape == (n, 2)
        np.testing.assert_allclose(r.lame, np.tile([mu, lam], (n, 1)), rtol=6e-6)


def test_text_matches_reference_style_parse(writer, tmp_path):
    """numpy.loadtxt + reshape, as python/ioutils.py:38-70 reads 2D dumps, gives the same arrays."""
    subprocess.run([writer, str(tmp_path), "text", "2", "5", "64", "1"], check=True)
    r = ioutils.process_tmp(str(tmp_path))["0_"]
    x = np.loadtxt(tmp_path / "0_x.txt")
    np.testing.assert_array_equal(r.x, x.reshape(len(x) // 2, 2))
    F = np.loadtxt(tmp_path / "0_F.txt")
    np.testing.assert_array_equal(r.F, F.reshape(F.shape[0] // 2, 2, 2))
    np.testing.assert_array_equal(r.mass, np.loadtxt(tmp_path / "0_mass.txt").reshape(65, 65))
    np.testing.assert_array_equal(r.velocity, np.loadtxt(tmp_path / "0_velocity.txt").reshape(65, 65, 2))
    np.testing.assert_array_equal(r.lame, np.loadtxt(tmp_path / "0_lame.txt"))
    # node (i, j) is flat index i * 65 + j
    for idx in (0, 1, 64, 65, 4224):
        i, j = ioutils.to_3d(idx, 65)
        assert (i, j) == ioutils.to_nd(idx, 65, 2)
        assert r.mass[i, j] == np.loadtxt(tmp_path / "0_mass.txt")[idx]


def test_trajectory_and_pickle(writer, tmp_path):
    out = tmp_path / "dump"
    subprocess.run([writer, str(out), "npy", "3", "6", "3", "4"], check=True)
    traj = ioutils.load_trajectory(str(out))
    assert traj.steps == [0, 1, 2, 3] and traj.dim == 3 and traj.res == 3
    assert traj.x.shape == (4, 6, 3) and traj.F.shape == (4, 6, 3, 3)
    assert traj.mass.shape == (4, 4, 4, 4) and traj.velocity.shape == (4, 4, 4, 4, 3)
    for s in range(4):
        np.testing.assert_array_equal(traj.C[s], expected(s, 3, 6, 3)["C"])
    np.testing.assert_allclose(traj.time, np.float32(1e-4) * np.maximum(np.arange(4), 1).astype(np.float32))
    torch = pytest.importorskip("torch")
    tt = ioutils.load_trajectory(str(out), fields=("x", "Jp"), as_torch=True)
    assert isinstance(tt.x, torch.Tensor) and tuple(tt.x.shape) == (4, 6, 3) and set(tt.fields) == {"x", "Jp"}
    pk = tmp_path / "results.pickle"
    res = ioutils.process_tmp(str(out), pickle_path=str(pk))
    with open(pk, "rb") as f:  # a file this test just wrote
        back = pickle.load(f)
    np.testing.assert_array_equal(back["2_"].x, res["2_"].x)


def test_errors(tmp_path):
    with pytest.raises(FileNotFoundError):
        ioutils.load_trajectory(str(tmp_path))
    (tmp_path / "0_mass.txt").write_text("1\n2\n3\n")
    with pytest.raises(ValueError):
        ioutils.process_tmp(str(tmp_path), dim=2)
