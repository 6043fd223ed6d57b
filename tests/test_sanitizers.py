"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (g++, CPU only).

* The CPU oracle (oracle/oracle_capi.cpp with nclr_oracle.hpp and nclr_adjoint.hpp), driven by
  tests/cpp/oracle_sanitize.cpp through every step, adjoint and decomposition entry point for both
  dimensions, every material and both identity modes, on scenes in the interior and against a wall,
  plus a particle outside the grid (the error path).  The reference's own P2G/G2P index the grid
  through std::vector::at only after flattening (src/nclr.h:141-142, 163-164); the restatement checks
  every axis, and the sanitizers confirm no access leaves its arrays.
* The dump writer of the headless solver (include/nclr/nclr_dump.h), through tests/cpp/dump_writer.cpp
  in both formats.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


def _build(tmp_path, name, src, extra=()):
    exe = str(tmp_path / name)
    subprocess.run(["g++", "-std=c++17", *SAN, *extra, "-o", exe, src], check=True)
    return exe


def test_oracle_under_sanitizers(tmp_path):
    exe = _build(tmp_path, "oracle_sanitize", os.path.join(HERE, "cpp", "oracle_sanitize.cpp"), ["-fopenmp"])
    r = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=600)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout[-2000:] + r.stderr[-4000:]


@pytest.mark.parametrize("fmt,dim", [("text", 2), ("npy", 3)])
def test_dump_writer_under_sanitizers(tmp_path, fmt, dim):
    exe = _build(tmp_path, "dump_writer", os.path.join(HERE, "cpp", "dump_writer.cpp"),
                 ["-I", os.path.join(ROOT, "include", "nclr")])
    out = tmp_path / "dump"
    out.mkdir()
    r = subprocess.run([exe, str(out), fmt, str(dim), "7", "4", "3"], capture_output=True, text=True, env=ENV,
                       timeout=120)
    assert r.returncode == 0, r.stderr[-4000:]
    assert any(out.iterdir())
    r = subprocess.run([exe, "format"], capture_output=True, text=True, env=ENV, timeout=60)
    assert r.returncode == 0, r.stderr[-4000:]
