"""The fixed-point conversion of P2G and the G2P adjoint (fixed_point in
nuclearmpm_amd/csrc/mpm_device.hpp), emulated on the CPU (tests/cpp/fixed_point_check.c).

Both kernels sum every quantised contribution as a 64-bit integer, so any deterministic
quantisation within one unit of y keeps the sums exact and independent of the order of the LDS
atomics.  The conversion adds 1.5 * 2^52 in double and reads round(y) off the bit pattern: it is
round-to-nearest-even of y for every float below 2^51 in magnitude, which this test checks over a
strided sweep of all float bit patterns, every float in [-4, 4] and the top binade below 2^51.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_conversion_is_round_to_nearest_even_below_2_51(tmp_path):
    exe = str(tmp_path / "fixed_point_check")
    subprocess.run(["gcc", "-std=c11", "-O2", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "cpp", "fixed_point_check.c"), "-lm"], check=True)
    out = subprocess.run([exe, "61"], check=True, capture_output=True, text=True).stdout.splitlines()
    summary = out[-1].split()
    assert summary[0] == "checked" and int(summary[1]) > 10 ** 8
    assert int(summary[3]) == 0, out[:-1]
