"""The per-step particle sort is a stable counting sort: its permutation must equal numpy's stable
argsort bit for bit (integer work, no tolerance), on adversarial key sets."""
import numpy as np
import pytest

from nuclearmpm_amd._lib import sort_keys

pytestmark = pytest.mark.gpu


def expect(keys, bits):
    masked = keys.astype(np.uint64) & ((1 << bits) - 1) if bits < 32 else keys.astype(np.uint64)
    return np.argsort(masked, kind="stable").astype(np.uint32)


CASES = {
    "random13": lambda rng, n: rng.integers(0, 1 << 13, n),
    "all_equal": lambda rng, n: np.full(n, 77),
    "sorted": lambda rng, n: np.sort(rng.integers(0, 4097, n)),
    "reversed": lambda rng, n: np.sort(rng.integers(0, 4097, n))[::-1],
    "nearly_sorted": lambda rng, n: np.clip(np.sort(rng.integers(0, 4096, n)) + rng.integers(-1, 2, n), 0, 4096),
    "two_values": lambda rng, n: rng.integers(0, 2, n) * 4096,
    "wide24": lambda rng, n: rng.integers(0, 1 << 24, n),
    "full32": lambda rng, n: rng.integers(0, 1 << 32, n, dtype=np.uint64),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("n", [1, 63, 4096, 5000, 70001])
def test_stable_sort_matches_numpy(name, n):
    rng = np.random.default_rng(n)
    keys = np.asarray(CASES[name](rng, n)).astype(np.uint32)
    bits = 32 if name == "full32" else (24 if name == "wide24" else 13)
    print(f"[sort] {name} n={n} bits={bits}", flush=True)
    perm = sort_keys(keys, bits)
    assert np.array_equal(perm, expect(keys, bits))


def test_sort_ignores_bits_above_request_and_handles_large_n():
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 1 << 20, 1 << 20).astype(np.uint32)
    print("[sort] masked bits n=2^20", flush=True)
    perm = sort_keys(keys, 13)
    assert np.array_equal(perm, expect(keys, 13))
    assert sort_keys(np.zeros(0, np.uint32), 8).shape == (0,)


def slot_order(sim):
    """Particle index of every storage slot: the engine's particle -> slot permutation of state
    images (mpm_grad_from_particle_order) applied to the particle indices."""
    import torch

    from nuclearmpm_amd._lib import check, lib

    d, n = sim.dim, sim.n
    src = torch.arange(n, dtype=torch.float32, device="cuda").repeat(lib().mpm_grad_channels(d), 1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    check(lib().mpm_grad_from_particle_order(sim._handle(), src.data_ptr(), dst.data_ptr()))
    check(lib().mpm_synchronize(sim._handle()))
    return dst[0].cpu().numpy().astype(np.int64)


def tile_keys(x, res, dim):
    """Tile of every particle's stencil base, x-major (8^3 cells per tile in 3D, 16^2 in 2D)."""
    log = 4 if dim == 2 else 3
    nt = ((res - 2) >> log) + 1
    s = (x.astype(np.float32) * np.float32(res)).astype(np.float32)
    base = (s - np.float32(0.5)).astype(np.float32).astype(np.int64)
    key = np.zeros(len(x), np.int64)
    for a in range(dim):
        key = key * nt + (base[:, a] >> log)
    return key


@pytest.mark.parametrize("dim,res", [(2, 64), (3, 48), (2, 1100), (3, 140)])
def test_incremental_sort_is_the_stable_sort_by_tile(dim, res):
    """Every step's storage order must be the stable sort, by the tile of the state before the
    step, of the previous step's order (the first step sorts the particle order).  res 1100 in 2D
    (4,761 tiles) and 140 in 3D (5,832 tiles) take the tile scans' one-workgroup-per-round path."""
    from nuclearmpm_amd import MaterialModel, MPMSimulation
    from nuclearmpm_amd.sim import Particles

    rng = np.random.default_rng(10 + dim)
    n = 6000
    x = rng.uniform(0.15, 0.85, (n, dim)).astype(np.float32)
    # up to 0.6 cells per step on each axis: many tile crossings every step
    v = (rng.uniform(-1.0, 1.0, (n, dim)) * (0.6 / (res * 1e-4))).astype(np.float32)
    with MPMSimulation(Particles.from_positions(x, v=v), MaterialModel.kJelly, res=res) as sim:
        prev = np.arange(n)
        moved = 0
        for step in range(20):
            key = tile_keys(sim.particles().x, res, dim)
            sim.advance(1)
            order = slot_order(sim)
            want = prev[np.argsort(key[prev], kind="stable")]
            assert np.array_equal(order, want), f"step {step}"
            if step:
                moved += int(np.count_nonzero(key != last_key))
            prev, last_key = order, key
        print(f"[sort] {dim}D: {moved} tile changes over 19 steps", flush=True)
        assert moved > 100
