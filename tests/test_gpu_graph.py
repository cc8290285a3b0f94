"""The batched fill captured as one HIP graph gives, on every replay, the
pixels and coverage hashes of the eager recording (replays interleaved with
eager work and synchronisation)."""
import numpy as np
import pytest

import rr_hip as R

pytestmark = pytest.mark.gpu


def _state(db):
    return [db.read_canvas(i) for i in range(len(db.pages))], db.read_hashes()


@pytest.mark.parametrize("cfg,seed,n", [(3, 0xC0F10003, 3), (2, 0xC0F10002, 1)])
def test_graph_replays_match_eager(backend, cfg, seed, n):
    pages = R.workload(cfg, seed, n)
    db = R.DeviceBatch(backend, pages)
    g = None
    try:
        db.run(launch=R.LAUNCH_BATCHED, hashes=True, variant=2)
        want_c, want_h = _state(db)
        g = db.capture(hashes=True, variant=2)
        for rep in range(3):
            db.run_graph(g)
            got_c, got_h = _state(db)
            for i, (a, b) in enumerate(zip(got_c, want_c)):
                assert np.array_equal(a, b), f"replay {rep} page {i}"
            assert np.array_equal(got_h, want_h), f"replay {rep} hashes"
            # eager work between replays: a plain fill of the first page
            db.run(launch=R.LAUNCH_PLAIN, variant=0)
    finally:
        if g:
            db.destroy_graph(g)
        db.free()


def test_graph_refuses_after_workspace_growth():
    # a fresh context sized for a small batch: a larger eager recording must
    # reallocate its workspace, after which the graph refuses to replay
    be = R.Backend(0)
    try:
        small = R.workload(1, 0xC0F10001, 1)
        db = R.DeviceBatch(be, [small[0].subset(np.arange(50))])
        g = db.capture(variant=2)
        big = R.DeviceBatch(be, R.workload(3, 0xC0F10003, 2))
        big.run(launch=R.LAUNCH_BATCHED, variant=2)
        big.free()
        be.begin_page()
        with pytest.raises(R.RrError):
            be._check(be.L.rr_hip_graph_launch(be.h, g), "graph_launch")
        be.wait_page(be.submit_page())
        db.destroy_graph(g)
        db.free()
    finally:
        be.close()


def test_capture_first_on_a_fresh_context_with_occlusion():
    # The capture is the context's first recording: its sizing pass must grow
    # the interval pool itself (whole mixed pages need more than the initial
    # pool) so that replays, occlusion skipping on (no coverage outputs), fit
    # and match the oracle on every replay.
    from test_gpu_batched import oracle_pages
    pages = R.workload(3, 0xC0F10003, 2)
    oc, _ = oracle_pages(pages)
    be = R.Backend(0)
    db = R.DeviceBatch(be, pages)
    g = None
    try:
        g = db.capture(hashes=False, variant=0)
        for rep in range(3):
            db.run_graph(g)
            for i, o in enumerate(oc):
                assert np.array_equal(db.read_canvas(i), o), f"replay {rep} page {i}"
    finally:
        if g:
            db.destroy_graph(g)
        db.free()
        be.close()
