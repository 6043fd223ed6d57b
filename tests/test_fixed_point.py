"""P2G's fixed-point conversion (fixed_point in nuclearmpm_amd/csrc/mpm_device.hpp), emulated on the
CPU with the GPU's truncating, saturating conversions (tests/cpp/fixed_point_check.c).

P2G and the G2P adjoint sum every quantised contribution as a 64-bit integer, so any deterministic
quantisation within one unit of y keeps the sums exact and independent of the order of the LDS
atomics.  The conversion is floor(y) for every float below 2^62 in magnitude except four values in
(-4, -1], where the fma's rounding of 2^32 + y gives floor(y) + 1; this test pins that set.
"""
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# y just above -n by one or two float ulps: 2^32 + y rounds up to an integer in double
OFF_BY_ONE = {-(1 + 2.0 ** -23): -1, -(1 + 2.0 ** -22): -1, -(2 + 2.0 ** -22): -2, -(3 + 2.0 ** -22): -3}


def test_conversion_is_floor_up_to_four_values(tmp_path):
    exe = str(tmp_path / "fixed_point_check")
    subprocess.run(["gcc", "-std=c11", "-O2", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "cpp", "fixed_point_check.c"), "-lm"], check=True)
    out = subprocess.run([exe, "61"], check=True, capture_output=True, text=True).stdout.splitlines()
    summary = out[-1].split()
    assert summary[0] == "checked" and int(summary[1]) > 10 ** 8
    odd = {}
    for line in out[:-1]:
        y, _, got, _, want = line.split()
        y = float.fromhex(y[2:])
        odd[y] = (int(got), int(want))
    assert int(summary[3]) == len(odd) == 4
    for y, got in OFF_BY_ONE.items():
        assert np.float32(y) == y
        assert odd[y] == (got, got - 1)
