"""GPU parity of the primitives (scan, radix sort, binning, Philox) through the
C ABI, against the oracle and textbook routines.  Integer work: bit-exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from tests.test_oracle_primitives import PHILOX_KAT


def test_philox_device_kat(flip2):
    rows = np.array([list(c) + list(k) for c, k, _ in PHILOX_KAT], dtype=np.uint32)
    out = flip2.philox4x32_10(rows)
    for r, (_, _, expect) in zip(out, PHILOX_KAT):
        assert tuple(int(x) for x in r) == expect


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 63, 64, 65, 255, 1000, 4095, 4096, 4097, 8191, 65536 + 17,
                               1_000_003])
@pytest.mark.parametrize("exclusive", [False, True])
def test_scan_matches_oracle(flip2, oracle, n, exclusive):
    rng = np.random.default_rng(n + exclusive)
    v = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    got, total = flip2.scan_u32(v, exclusive)
    assert np.array_equal(got, oracle.scan_u32(v, exclusive))
    assert total == int(np.sum(v.astype(np.uint64)) & 0xFFFFFFFF)


def test_scan_empty(flip2):
    got, total = flip2.scan_u32(np.zeros(0, np.uint32))
    assert got.size == 0 and total == 0


def test_scan_16m_counts(flip2):
    # exclusive scan over 16,777,216 per-cell counts (primitive microbench size)
    rng = np.random.default_rng(11)
    counts = rng.poisson(1.0, 1 << 24).astype(np.uint32)
    got, total = flip2.scan_u32(counts, exclusive=True)
    expect = np.concatenate([[0], np.cumsum(counts.astype(np.uint64))[:-1]]).astype(np.uint32)
    assert np.array_equal(got, expect)
    assert total == int(counts.sum())


@pytest.mark.parametrize("case", ["random", "dups", "sorted", "reversed", "equal", "high", "tiny"])
@pytest.mark.parametrize("n", [1, 100, 4096, 4097, 300_001])
def test_sort_matches_stable_sort(flip2, oracle, case, n):
    rng = np.random.default_rng(n)
    keys = {
        "random": rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32),
        "dups": rng.integers(0, 17, n).astype(np.uint32),
        "sorted": np.arange(n, dtype=np.uint32),
        "reversed": np.arange(n, dtype=np.uint32)[::-1].copy(),
        "equal": np.full(n, 0xDEADBEEF, np.uint32),
        "high": (rng.integers(0, 4, n).astype(np.uint32) << np.uint32(30)) | np.uint32(5),
        "tiny": rng.integers(0, 2, n).astype(np.uint32),
    }[case]
    k, perm = flip2.sort_pairs_u32(keys)
    ref = np.argsort(keys, kind="stable").astype(np.uint32)
    assert np.array_equal(perm, ref)
    assert np.array_equal(k, keys[ref])
    if n <= 5000:
        ko, po = oracle.radix_sort_u32(keys)
        assert np.array_equal(po, perm) and np.array_equal(ko, k)


@pytest.mark.parametrize("bits", [(0, 1), (0, 8), (0, 13), (3, 19), (0, 24), (8, 32)])
def test_sort_bit_ranges(flip2, bits):
    b0, b1 = bits
    rng = np.random.default_rng(b1)
    keys = rng.integers(0, 2**32, 50_000, dtype=np.uint64).astype(np.uint32)
    vals = rng.integers(0, 2**32, 50_000, dtype=np.uint64).astype(np.uint32)
    k, v = flip2.sort_pairs_u32(keys, vals, b0, b1)
    field = (keys >> np.uint32(b0)) & np.uint32((1 << (b1 - b0)) - 1 if b1 - b0 < 32 else 0xFFFFFFFF)
    ref = np.argsort(field, kind="stable")
    assert np.array_equal(v, vals[ref]) and np.array_equal(k, keys[ref])


def test_sort_16m_cell_keys(flip2):
    # 2^24 keys of a 256^3 grid (24 significant bits) with a u32 payload
    from flip2_amd import scenes
    n = 1 << 24
    x, y, z = scenes.uniform_random(n, seed=2)
    ix, iy, iz = (np.minimum((a * 256).astype(np.uint32), 255) for a in (x, y, z))
    keys = ix + iy * np.uint32(256) + iz * np.uint32(65536)
    k, perm = flip2.sort_pairs_u32(keys, None, 0, 24)
    assert np.all(k[1:] >= k[:-1])
    ref = np.argsort(keys, kind="stable").astype(np.uint32)
    assert np.array_equal(perm, ref)


def test_find_grid_cell_matches_oracle(flip2, oracle):
    rng = np.random.default_rng(4)
    n = 200_000
    g = (0.0, 0.0, 0.0, 64, 64, 64, 1.0 / 64)
    px, py, pz = (rng.uniform(-0.1, 1.1, n) for _ in range(3))
    # plant points on and next to cell faces
    px[:64] = np.arange(64) / 64.0
    px[64:128] = np.nextafter(np.arange(64) / 64.0, -1)
    got = flip2.find_grid_cell(px, py, pz, g)
    exp = oracle.find_grid_cell(px, py, pz, g)
    for a, b in zip(got, exp):
        assert np.array_equal(a, b)
    g2 = (-10.0, -10.0, -10.0, 100, 100, 100, 0.25)
    qx, qy, qz = (rng.uniform(-12, 17, n) for _ in range(3))
    got = flip2.find_grid_cell(qx, qy, qz, g2)
    exp = oracle.find_grid_cell(qx, qy, qz, g2)
    for a, b in zip(got, exp):
        assert np.array_equal(a, b)
