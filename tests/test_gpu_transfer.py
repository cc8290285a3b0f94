"""GpuBackend items beyond the page ops, through the C-ABI: buffer lengths,
the VRAM budget, pinned host buffers, fenced asynchronous transfers, per-page
error status carried by the page's fence, and the callers' CPU fallback when
a page reports an error (crates/pdf_interp/src/renderer/page/vk_ops.rs:62-70;
the oracle stands in for the caller's CPU raster)."""
import numpy as np
import pytest

import rr_hip as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_buffer_len_and_zero_size(backend):
    b = backend.alloc(12345)
    try:
        assert backend.buffer_len(b) == 12345
    finally:
        backend.free(b)
    with pytest.raises(R.RrError) as ei:
        backend.buffer_len(b)  # freed: no longer a buffer of the context
    assert ei.value.code == R.RR_E_INVARIANT
    with pytest.raises(R.RrError) as ei:
        backend.alloc(0)
    assert ei.value.code == R.RR_E_ZERO_SIZE_ALLOC
    with pytest.raises(R.RrError) as ei:
        backend.alloc_host_pinned(0)
    assert ei.value.code == R.RR_E_ZERO_SIZE_ALLOC


def test_vram_budget(backend):
    total, usable = backend.vram_budget()
    assert total > 100 << 30          # an MI355X carries 288 GB of HBM3E
    assert 0 < usable <= total


def test_pinned_async_round_trip(backend):
    n = (1 << 22) + 77
    rng = np.random.default_rng(3)
    src_p, src = backend.alloc_host_pinned(n)
    dst_p, dst = backend.alloc_host_pinned(n)
    d = backend.alloc(n)
    try:
        src[:] = rng.integers(0, 256, n, dtype=np.uint8)
        dst[:] = 0
        f_up = backend.upload_async(d, src_p, n)
        f_dn = backend.download_async(dst_p, d, n)
        assert f_dn > f_up
        backend.wait_transfer(f_dn)
        assert np.array_equal(dst, src)
        backend.wait_transfer(f_up)        # an earlier fence stays waitable
        f = backend.submit_transfer()
        backend.wait_transfer(f)
        with pytest.raises(R.RrError):
            backend.upload_async(d, src_p, n + 1)   # longer than the buffer
        with pytest.raises(R.RrError):
            backend.wait_page(f)                    # a transfer fence is not a page
        with pytest.raises(R.RrError):
            backend.wait_transfer(f + 1000)         # never handed out
    finally:
        backend.free(d)
        backend.free_host_pinned(src_p)
        backend.free_host_pinned(dst_p)


def test_page_status_belongs_to_its_fence(backend):
    # Page A reports an error; page B, begun and submitted before A is
    # waited, is clean.  Each wait returns its own page's status.
    backend.begin_page()
    backend.inject_error(1)
    fa = backend.submit_page()
    backend.begin_page()
    fb = backend.submit_page()
    backend.wait_page(fb)
    with pytest.raises(R.RrError) as ei:
        backend.wait_page(fa)
    assert ei.value.code == R.RR_E_CAPACITY
    with pytest.raises(R.RrError):
        backend.wait_page(fa)              # still A's status on a second wait
    backend.wait_page(fb)


def test_error_bits_survive_a_pool_regrow():
    # An error set earlier in the page must survive the batched recording's
    # grow-and-redo of its span pass (which clears only its own bit).
    pages = R.workload(3, 0xC0F10003, 2)
    be = R.Backend(0)
    db = R.DeviceBatch(be, pages)
    try:
        db.reset()
        be.begin_page()
        be.inject_error(1)
        be.record_fill_pages(db.fill_pages(launch=R.LAUNCH_BATCHED, hashes=True, variant=2))
        f = be.submit_page()
        with pytest.raises(R.RrError) as ei:
            be.wait_page(f)
        assert ei.value.code == R.RR_E_CAPACITY
        assert be.last_batch_stats()[1]["reruns"] >= 1
    finally:
        db.free()
        be.close()


def test_cpu_fallback_after_a_page_error(backend):
    # The caller's contract: any error -> discard the page's device output and
    # render it on the CPU.  The fallback page equals the device page of a
    # clean run, so a failure never changes the pixels.
    page = R.workload(1, 0xC0F10001)[0].subset(range(0, 500, 9))
    dp = R.DevicePage(backend, page)
    a = page.attrs
    try:
        clean, _ = dp.render(launch=R.LAUNCH_BATCHED)
        dp.reset_canvas()
        backend.begin_page()
        backend.record_fill_paths(dp.batch(launch=R.LAUNCH_BATCHED))
        backend.inject_error(1)
        try:
            backend.wait_page(backend.submit_page())
            out = dp.read_canvas()
        except R.RrError as e:
            assert e.code == R.RR_E_CAPACITY
            out = np.full((page.height, page.width), 255, np.uint8)
            O.fill_page(out, page.edges, page.edge_off, a["eo"], a["color"], a["a_input"])
    finally:
        dp.free()
    assert np.array_equal(out, clean)
