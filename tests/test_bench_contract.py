"""bench.py's command-line contract that needs no GPU: one GPU only, and the algorithmic byte
counts it prices the kernels with (SoA FP32 minimum I/O of each kernel)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_algorithmic_bytes():
    import bench

    for d in (2, 3):
        ng = 2 * d + 2 * d * d + 1  # x, v, F, C, Jp
        # P2G reads x, v, F, C, (Jp for snow), mass, volume; G2P reads x, F, (Jp) and writes x, v, C, F, (Jp)
        assert bench.P2G_BYTES_PER_PARTICLE[(d, 0)] == 4 * (ng + 2)
        assert bench.P2G_BYTES_PER_PARTICLE[(d, 1)] == 4 * (ng + 1)
        assert bench.G2P_BYTES_PER_PARTICLE[(d, 0)] == 4 * ((d + d * d + 1) + (2 * d + 2 * d * d + 1))
        assert bench.G2P_BYTES_PER_PARTICLE[(d, 1)] == 4 * ((d + d * d) + (2 * d + 2 * d * d))
        assert bench.P2G_BYTES_PER_NODE[d] == 4 * (d + 1)
        assert bench.GRID_BYTES_PER_NODE[d] == 4 * 2 * (d + 1)
        assert bench.G2P_BYTES_PER_NODE[d] == 4 * d
        adj = bench.adjoint_bytes(d)
        assert adj["g2p_adjoint"] == (4 * (d + d * d + ng + d + d * d + 1), 4 * 2 * d)
        assert adj["grid_adjoint"] == (0, 4 * (2 * (d + 1) + d))
        assert adj["p2g_adjoint"] == (4 * ((ng + 1) + (d + d * d + 1) + ng + 1), 4 * (d + 1))
    # the 3D snow figures quoted for the headline workload
    assert bench.P2G_BYTES_PER_PARTICLE[(3, 0)] == 108 and bench.G2P_BYTES_PER_PARTICLE[(3, 0)] == 152
