"""flip2_sort_pairs_u32_noalloc (no allocation: fixed per-GPU workspace
array) against flip2_sort_pairs_u32 and a stable numpy sort, bit for bit,
on two calls in a row (the second smaller, over fewer bits, sorting the
permutation), and its refusal of a request larger than the array."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_vp, _u32 = ctypes.c_void_p, ctypes.c_uint32


def _noalloc(F):
    fn = F.lib().flip2_sort_pairs_u32_noalloc
    fn.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _u32, _vp]
    fn.restype = ctypes.c_int
    return fn


def _run_both(F, keys, vals, end_bit):
    n = keys.size
    print(f"noalloc sort: n={n} end_bit={end_bit} values={'yes' if vals is not None else 'permutation'}",
          flush=True)
    dk = F.DeviceBuffer.from_array(keys)
    dv = F.DeviceBuffer.from_array(vals) if vals is not None else None
    outs = []
    for which in ("noalloc", "workspace"):
        ko, vo, ka, va = (F.DeviceBuffer(4 * n) for _ in range(4))
        if which == "noalloc":
            F.check(_noalloc(F)(dk.ptr, dv.ptr if dv else None, ko.ptr, vo.ptr, ka.ptr, va.ptr, n, 0, end_bit, None))
        else:
            ws = F.lib().flip2_sort_workspace_bytes(n, 0, end_bit)
            dws = F.DeviceBuffer(ws)
            F.check(F.lib().flip2_sort_pairs_u32(dk.ptr, dv.ptr if dv else None, ko.ptr, vo.ptr, ka.ptr, va.ptr, n,
                                                 0, end_bit, dws.ptr, ws, None))
        F.synchronize()
        outs.append((ko.to_array(np.uint32, n), vo.to_array(np.uint32, n)))
    return outs


def test_noalloc_sort_matches_workspace_sort(flip2):
    F = flip2
    rng = np.random.default_rng(11)
    # first call: 3M keys over 32 bits with values
    k1 = rng.integers(0, 2**32, size=3_000_001, dtype=np.uint64).astype(np.uint32)
    v1 = rng.integers(0, 2**32, size=k1.size, dtype=np.uint64).astype(np.uint32)
    (a_k, a_v), (b_k, b_v) = _run_both(F, k1, v1, 32)
    order = np.argsort(k1, kind="stable")
    assert np.array_equal(a_k, b_k) and np.array_equal(a_v, b_v)
    assert np.array_equal(a_k, k1[order]) and np.array_equal(a_v, v1[order])
    # second call, smaller: 70,000 keys over 24 bits, the permutation
    k2 = rng.integers(0, 2**24, size=70_000, dtype=np.uint64).astype(np.uint32)
    (a_k, a_p), (b_k, b_p) = _run_both(F, k2, None, 24)
    order = np.argsort(k2, kind="stable").astype(np.uint32)
    assert np.array_equal(a_k, b_k) and np.array_equal(a_p, b_p)
    assert np.array_equal(a_p, order) and np.array_equal(a_k, k2[order])


def test_noalloc_sort_refuses_more_than_its_array(flip2):
    F = flip2
    n = 40_000_000  # needs more than FLIP2_NOALLOC_SORT_SCRATCH_BYTES over 32 bits
    assert F.lib().flip2_sort_workspace_bytes(n, 0, 32) > (34 << 20)
    small = F.DeviceBuffer(64)
    rc = _noalloc(F)(small.ptr, None, small.ptr, small.ptr, small.ptr, small.ptr, n, 0, 32, None)
    assert rc == -5  # FLIP2_ERR_RANGE, refused on the host before any launch
    assert b"without allocation" in F.lib().flip2_last_error()
