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
