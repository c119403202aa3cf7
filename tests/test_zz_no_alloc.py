"""GPU: injectdata_float_noalloc, the injection entry point that allocates nothing, against the
existing injectdata_float on device arrays (its grouped device path), through the C-ABI.

* two calls in a row on the same device arrays, the second with fewer points and partitions:
  bit for bit equal to injectdata_float (the shared staging array keeps no stale state that
  matters);
* partitions not sorted by grid index (the one-thread-per-partition form): equal to the
  reference's loop on the host, bit for bit;
* inputs beyond the stated maxima, or host arrays, are rejected with a message and leave the
  field unchanged.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32 = np.float32


def _points(w, rng, n, ncell, nr):
    pts = np.zeros(n, w.SourcePoint32)
    # repeated grid indices (several taps on one cell) and repeated traces
    pts["iu"] = rng.integers(1, ncell + 1, n)
    pts["iu"][: n // 4] = pts["iu"][n // 4: n // 2]
    pts["ir"] = rng.integers(1, nr + 1, n)
    pts["c"] = rng.standard_normal(n).astype(F32)
    return pts


def _noalloc(field, partitions, data, it):
    from wavefd_amd import _lib
    from wavefd_amd.spacetime import _addr, _partition_arrays
    keep, pp, nparts, npts = _partition_arrays(partitions)
    nt = data.shape[-1]
    _lib.load("spacetime").injectdata_float_noalloc(_addr(field), _addr(data), it, nt, pp, nparts, npts, 8)
    return keep


def test_noalloc_injection_equals_injectdata_twice():
    import wavefd_amd as w
    from tests.devmem import DeviceArray
    rng = np.random.default_rng(5)
    nz, ny, nx, nr, nt = 40, 36, 44, 37, 50
    ncell = nz * ny * nx
    field = rng.standard_normal(ncell).astype(F32)
    data = rng.standard_normal((nr, nt)).astype(F32)
    dd = DeviceArray(data)
    a, b = DeviceArray(field), DeviceArray(field)
    # call 1: 5000 points in 8 partitions; call 2: 300 points in 3 partitions
    for n, nparts, it in ((5000, 8, 17), (300, 3, 50)):
        pts = _points(w, rng, n, ncell, nr)
        parts = w.source_blocking(pts, nparts)
        w.injectdata(a, parts, dd, it)
        _noalloc(b, parts, dd, it)
        got, want = b.numpy(), a.numpy()
        assert np.array_equal(got, want), (n, int(np.sum(got != want)))
        assert not np.array_equal(got, field)


def test_noalloc_injection_unsorted_partitions_equal_reference_loop():
    import wavefd_amd as w
    from tests.devmem import DeviceArray
    rng = np.random.default_rng(6)
    ncell, nr, nt = 30 * 31 * 32, 9, 40
    field = rng.standard_normal(ncell).astype(F32)
    data = rng.standard_normal((nr, nt)).astype(F32)
    pts = _points(w, rng, 2000, ncell, nr)
    # race-free (each grid index in one partition) but not sorted inside a partition
    key = pts["iu"] % 5
    parts = [pts[key == k] for k in range(5)]
    host = field.copy()
    w.injectdata(host, parts, data, 7)  # the reference's loop, on the host
    b, dd = DeviceArray(field), DeviceArray(data)
    _noalloc(b, parts, dd, 7)
    assert np.array_equal(b.numpy(), host)


def test_noalloc_injection_rejects_oversize_and_host_arrays():
    import wavefd_amd as w
    from tests.devmem import DeviceArray
    from wavefd_amd import _lib
    field = np.zeros(1000, F32)
    data = np.ones((2, 5), F32)
    b, dd = DeviceArray(field), DeviceArray(data)
    one = np.zeros(1, w.SourcePoint32)
    one["iu"], one["ir"], one["c"] = 3, 1, 1.0
    cases = [([one] * 4097, dd, "partitions"), ([one], data, "device memory"),
             ([np.array([(1001, 1, 1.0)], w.SourcePoint32)], dd, "outside")]
    for parts, d, msg in cases:
        _noalloc(b, parts, d, 1)
        assert msg in _lib.last_error("spacetime")
        assert float(np.abs(b.numpy()).sum()) == 0.0
    # a valid call still works after the rejected ones
    _noalloc(b, [one], dd, 2)
    out = b.numpy()
    assert out[2] == 1.0 and float(np.abs(out).sum()) == 1.0
