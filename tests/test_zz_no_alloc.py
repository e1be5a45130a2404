"""The no-allocation DEC split (ns_dec_split_device_noalloc) against the allocating one (ns_dec_split).

The no-allocation entry point keeps its counters in a static device array shared by every call on
a GPU, so consecutive calls must each start from cleared counters: the calls below go from all
planes non-zero to only the low planes non-zero, from a 16-byte-aligned length to a ragged one,
and through a rejected out-of-range input, and each result must equal the allocating entry
point's bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
D = 54


def _noalloc(ns, zmix: np.ndarray, k_dec: int):
    """Upload zmix, let the entry point split it into device planes, read the planes back."""
    L = ns.lib()
    fn = L.ns_dec_split_device_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p, C.c_void_p]
    zmix = np.ascontiguousarray(zmix, np.int32).reshape(-1)
    m = len(zmix) // D
    zd = ns.DeviceBuffer.from_array(zmix)
    planes = [ns.DeviceBuffer(D * m) for _ in range(k_dec)]
    ptrs = (C.c_void_p * k_dec)(*[p.data_ptr() for p in planes])
    nz = np.full(k_dec, -1, np.int32)
    rc = fn(zd.data_ptr(), m, k_dec, ptrs, nz.ctypes.data)
    if rc != 0:
        raise ns.NsError(rc, L.ns_last_error().decode())
    return [p.to_array(np.int8, D * m).reshape(D, m) for p in planes], [bool(x) for x in nz]


def _check_equal(ns, zmix, k_dec):
    want_planes, want_nz = ns.dec_split(zmix, k_dec)
    got_planes, got_nz = _noalloc(ns, zmix, k_dec)
    assert got_nz == want_nz
    for i in range(k_dec):
        assert np.array_equal(got_planes[i], want_planes[i]), f"plane {i}"
    return got_nz


def test_noalloc_dec_split_matches_allocating(ns):
    rng = np.random.default_rng(2027)
    print("no-alloc DEC split: m=4096 full range, then m=999 small values", flush=True)
    # first call: the full k_dec = 12 range, every plane non-zero
    big = rng.integers(-(1 << 12) + 1, 1 << 12, size=(D, 4096), dtype=np.int32)
    nz1 = _check_equal(ns, big, 12)
    assert all(nz1)
    # second call, smaller and ragged (54 * 999 entries, not a multiple of 16): values in [-7, 7],
    # so planes 3.. are zero; stale counters from the first call would show here
    small = rng.integers(-7, 8, size=(D, 999), dtype=np.int32)
    nz2 = _check_equal(ns, small, 12)
    assert nz2[:3] == [True, True, True] and not any(nz2[3:])


def test_noalloc_dec_split_rejects_then_recovers(ns):
    rng = np.random.default_rng(7)
    bad = np.zeros((D, 256), np.int32)
    bad[5, 17] = 1 << 12
    with pytest.raises(ns.NsError):
        _noalloc(ns, bad, 12)
    # the range counter of the rejected call must not leak into the next one
    ok = rng.integers(-100, 101, size=(D, 128), dtype=np.int32)
    _check_equal(ns, ok, 12)
    # k_dec beyond the scratch's planes is refused before any launch
    with pytest.raises(ns.NsError):
        _noalloc(ns, ok, 31)
