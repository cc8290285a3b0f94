"""The allocation-free scan (rr_hip_record_scan_static: block sums in the
library's static device array) against rr_hip_record_scan with a caller
buffer, through the C-ABI: two calls in a row, the second smaller, so the
second reads an array the first left full."""
import numpy as np
import pytest

import rr_hip as R

pytestmark = pytest.mark.gpu


def _input(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    x[: n // 3] = rng.integers(0, 16, n // 3).astype(np.uint32)
    return x


def _scan_with_buffer(be, x):
    n = len(x)
    d = be.to_device(x)
    bs = be.alloc(4 * ((n + 1023) // 1024))
    try:
        be.begin_page()
        be.record_scan(d, bs, n)
        be.wait_page(be.submit_page())
        return be.download(d, np.zeros(n, np.uint32))
    finally:
        be.free(d)
        be.free(bs)


def _scan_static(be, x):
    n = len(x)
    d = be.to_device(x)
    try:
        be.begin_page()
        be.record_scan_static(d, n)
        be.wait_page(be.submit_page())
        return be.download(d, np.zeros(n, np.uint32))
    finally:
        be.free(d)


@pytest.mark.parametrize("sizes", [(R.SCAN_STATIC_MAX, 100_003), (5000, 1), (1025, 1024)])
def test_static_scan_matches_buffered_scan(backend, sizes):
    for k, n in enumerate(sizes):
        x = _input(n, 1000 * k + n)
        want = _scan_with_buffer(backend, x)
        got = _scan_static(backend, x)
        assert np.array_equal(got, want), f"call {k} (n={n}): first diff {np.nonzero(got != want)[0][:5]}"
        assert got[0] == 0


def test_static_scan_rejects_sizes_beyond_the_array(backend):
    d = backend.alloc(64)
    try:
        backend.begin_page()
        for n in (0, R.SCAN_STATIC_MAX + 1, 0xFFFFFFFF):
            with pytest.raises(R.RrError) as ei:
                backend.record_scan_static(d, n)
            assert ei.value.code == R.RR_E_INVARIANT
        backend.wait_page(backend.submit_page())
    finally:
        backend.free(d)
