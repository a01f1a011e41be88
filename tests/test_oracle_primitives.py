"""Pins the oracle's primitives against published vectors, textbook routines
and hand-derived cases (CPU only)."""
import numpy as np
import pytest

from flip2_amd import scenes

# Random123 known-answer vectors for philox4x32-10 (kat_vectors).
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_kat(oracle, ctr, key, expect):
    assert tuple(int(x) for x in oracle.philox4x32_10(ctr, key)) == expect
    w = scenes.philox4x32_10(*[[c] for c in ctr], *key)
    assert tuple(int(x[0]) for x in w) == expect


@pytest.mark.parametrize("n", [0, 1, 7, 4096, 100003])
@pytest.mark.parametrize("exclusive", [False, True])
def test_scan_matches_cumsum(oracle, n, exclusive):
    rng = np.random.default_rng(n)
    v = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    cs = np.cumsum(v.astype(np.uint64)) & 0xFFFFFFFF
    expect = cs if not exclusive else np.concatenate([[0], cs[:-1]]) if n else cs
    assert np.array_equal(oracle.scan_u32(v, exclusive), expect.astype(np.uint32))


@pytest.mark.parametrize("case", ["random", "dups", "sorted", "reversed", "equal", "one", "empty", "high"])
def test_sort_is_stable_sort(oracle, case):
    rng = np.random.default_rng(5)
    n = 5000
    keys = {
        "random": rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32),
        "dups": rng.integers(0, 17, n).astype(np.uint32),
        "sorted": np.arange(n, dtype=np.uint32),
        "reversed": np.arange(n, dtype=np.uint32)[::-1].copy(),
        "equal": np.full(n, 12345, np.uint32),
        "one": np.array([7], np.uint32),
        "empty": np.zeros(0, np.uint32),
        "high": (rng.integers(0, 4, n).astype(np.uint32) << np.uint32(30)) | np.uint32(5),
    }[case]
    k, perm = oracle.radix_sort_u32(keys)
    ref = np.argsort(keys, kind="stable")
    assert np.array_equal(perm, ref.astype(np.uint32))
    assert np.array_equal(k, keys[ref])


def test_sort_partial_bits_sorts_low_bits_only(oracle):
    rng = np.random.default_rng(9)
    keys = rng.integers(0, 2**32, 3000, dtype=np.uint64).astype(np.uint32)
    k, perm = oracle.radix_sort_u32(keys, nbits=12)
    ref = np.argsort(keys & np.uint32(0xFFF), kind="stable")
    assert np.array_equal(perm, ref.astype(np.uint32))


def test_sort_pairs_carries_values(oracle):
    rng = np.random.default_rng(2)
    keys = rng.integers(0, 100, 1000).astype(np.uint32)
    vals = rng.integers(0, 2**32, 1000, dtype=np.uint64).astype(np.uint32)
    k, v = oracle.radix_sort_pairs_u32(keys, vals)
    ref = np.argsort(keys, kind="stable")
    assert np.array_equal(v, vals[ref]) and np.array_equal(k, keys[ref])


def test_binning_cells_and_clamping(oracle):
    # grid (0,0,0), 8^3 cells of 0.125
    g = (0.0, 0.0, 0.0, 8, 8, 8, 0.125)
    px = np.array([0.0, 0.125 * 3, 0.125 * 3 + 1e-9, -0.5, 1.5, 0.999, 0.125 * 5 - 1e-12])
    py = np.array([0.0, 0.3, 0.3, 0.3, 0.3, 0.999, 0.3])
    pz = np.array([0.0, 0.3, 0.3, 0.3, 0.3, 0.999, 0.3])
    x, y, z, keys = oracle.find_grid_cell(px, py, pz, g)
    # y = z = 0.3 -> cell 2 -> row offset 2*8 + 2*64 = 144
    assert keys[0] == 0
    assert keys[1] == 3 + 144 and keys[2] == 3 + 144
    assert x[3] == 0.0 and keys[3] == 0 + 144            # clamped low, written back
    assert x[4] == 1.0 and keys[4] == 8 + 144            # clamped high: index 8 aliases the next row
    assert keys[5] == 7 + 7 * 8 + 7 * 64
    # quotient 5 - 8e-12 narrowed to float rounds to 5.0 -> cell 5, not 4
    assert keys[6] == 5 + 144


def test_binning_offset_domain(oracle):
    # engine/main.cu:16 domain: corner -10, 100^3 cells of 0.25
    g = (-10.0, -10.0, -10.0, 100, 100, 100, 0.25)
    px = np.array([-10 + 0.25 * 37 + 0.1])
    py = np.array([-10 + 0.25 * 5 + 0.2])
    pz = np.array([-10 + 0.25 * 99 + 0.2])
    _, _, _, keys = oracle.find_grid_cell(px, py, pz, g)
    assert keys[0] == 37 + 5 * 100 + 99 * 10000


def test_sampled_absmax_hand_cases(oracle):
    a = np.zeros(10); a[0] = 1.0; a[5] = 100.0
    assert oracle.sampled_absmax(a) == 1.0            # only a[0] and (absent) a[256] are sampled
    a = np.zeros(300); a[0] = -3.0; a[256] = 2.0; a[100] = 50.0
    assert oracle.sampled_absmax(a) == -3.0           # signed value of the larger magnitude
    a = np.zeros(300); a[0] = -2.0; a[256] = 2.0
    assert oracle.sampled_absmax(a) == 2.0            # ties take the second element
    a = np.zeros(600); a[0] = 1.0; a[256] = -5.0; a[512] = 1000.0
    assert oracle.sampled_absmax(a) == -5.0           # tile 1 is dropped by the second level
    a = np.zeros(512 * 300); a[0] = 1; a[256] = 2; a[131072] = 3; a[131328] = -7; a[1000] = 99
    assert oracle.sampled_absmax(a) == -7.0
    assert oracle.sampled_absmax(np.zeros(5000)) == 0.0


def test_jittered_scene_particles_lie_in_their_cells():
    x, y, z = scenes.jittered_cells((4, 4, 4), 0.25, seed=3)
    assert x.size == 4 ** 3 * 8
    cells = (np.floor(x / 0.25) + 4 * np.floor(y / 0.25) + 16 * np.floor(z / 0.25)).astype(int)
    counts = np.bincount(cells, minlength=64)
    assert np.all(counts == 8)


def test_dam_break_particle_count():
    x, y, z = scenes.dam_break(64, seed=1)
    # 8 ppc over x < 0.4, y < 0.6 of the unit cube: 0.24 * 64^3 * 8 = 503,316.5
    assert abs(x.size - 503316) < 2000
    assert x.max() < 0.4 and y.max() < 0.6
