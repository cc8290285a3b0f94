"""Device exclusive scan (the reference's record_scan, three phases in one
call) against numpy on uint32 with wrap-around."""
import numpy as np
import pytest

import rr_hip as R

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2, 63, 64, 1023, 1024, 1025, 4096, 100_003, (1 << 20) + 7])
def test_scan_matches_cumsum(backend, n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    if n > 4:
        x[: n // 4] = rng.integers(0, 16, n // 4).astype(np.uint32)
    want = np.zeros(n, np.uint64)
    want[1:] = np.cumsum(x.astype(np.uint64))[:-1]
    want = (want & 0xFFFFFFFF).astype(np.uint32)
    d = backend.to_device(x)
    bs = backend.alloc(4 * ((n + 1023) // 1024))
    try:
        backend.begin_page()
        backend.record_scan(d, bs, n)
        backend.wait_page(backend.submit_page())
        got = backend.download(d, np.zeros(n, np.uint32))
    finally:
        backend.free(d)
        backend.free(bs)
    assert np.array_equal(got, want), np.nonzero(got != want)[0][:5]


def test_scan_rejects_empty(backend):
    d = backend.alloc(64)
    try:
        backend.begin_page()
        with pytest.raises(R.RrError):
            backend.record_scan(d, d, 0)
        backend.wait_page(backend.submit_page())
    finally:
        backend.free(d)
