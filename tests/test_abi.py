"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
header declares, and its host path ingestion matches the oracle's XPath
restatement bit for bit.  No GPU calls."""
import ctypes

import numpy as np
import pytest

import rr_hip as R
from oracle import oracle as O


def test_library_exports_every_header_symbol():
    names = R.exported_symbols()
    assert "rr_hip_record_fill_paths" in names and len(names) >= 15
    L = R.lib()
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/rr_hip.h but not exported"


def test_no_device_is_reported_not_crashed():
    # In a container without a GPU, create must fail with RR_E_NO_DEVICE.
    h = ctypes.c_void_p()
    rc = R.lib().rr_hip_create(0, ctypes.byref(h))
    if rc == 0:
        R.lib().rr_hip_destroy(h)
        pytest.skip("a HIP device is present")
    assert rc == -4


def test_invalid_arguments():
    n = ctypes.c_int64(0)
    assert R.lib().rr_xpath_edges(None, None, 3, None, 0.1, None, 0, ctypes.byref(n)) == -2
    assert R.lib().rr_hip_begin_page(None) == -2
    h = ctypes.c_void_p()
    assert R.lib().rr_workload_create(99, 1, 0, ctypes.byref(h)) == -2


@pytest.mark.parametrize("cfg,seed", [(1, 0xC0F10001), (102, 0xC0F10002)])
def test_host_edges_match_oracle_xpath(cfg, seed):
    page = R.workload(cfg, seed)[0]
    step = 1 if cfg == 1 else 37
    for i in range(0, page.n_paths, step):
        xy, fl = page.path_points(i)
        e_o = O.xpath_edges(xy, fl, flatness=0.1)
        e_p = R.xpath_edges(xy, fl, flatness=0.1)
        assert e_o.shape == e_p.shape
        assert np.array_equal(e_o.view(np.uint64), e_p.view(np.uint64)), f"path {i}"
        assert np.array_equal(e_p.view(np.uint64), page.path_edges(i).view(np.uint64))


def test_host_edges_with_matrix():
    rng = np.random.default_rng(5)
    for _ in range(50):
        p = O.Path()
        p.move_to(*rng.uniform(0, 100, 2))
        for _ in range(rng.integers(1, 6)):
            if rng.random() < 0.5:
                p.line_to(*rng.uniform(0, 100, 2))
            else:
                p.curve_to(*rng.uniform(0, 100, 6))
        p.close(bool(rng.random() < 0.5))
        xy, fl = p.arrays()
        m = np.array([rng.uniform(0.5, 2), rng.uniform(-.3, .3), rng.uniform(-.3, .3),
                      rng.uniform(0.5, 2), rng.uniform(-10, 10), rng.uniform(-10, 10)])
        e_o = O.xpath_edges(xy, fl, m, flatness=0.1)
        e_p = R.xpath_edges(xy, fl, m, flatness=0.1)
        assert np.array_equal(e_o.view(np.uint64), e_p.view(np.uint64))


def test_cov_hash_term_matches_oracle():
    for pix, sh in [(0, 1), (12345, 255), (2550 * 3300 - 1, 17)]:
        assert R.cov_hash_term(pix, sh) == O.cov_hash_term(pix, sh)


def test_workload_shapes():
    p1 = R.workload(1, 0xC0F10001)[0]
    assert (p1.width, p1.height, p1.n_paths) == (2550, 3300, 500)
    assert set(np.unique(p1.attrs["eo"])) == {0, 1} and p1.attrs["eo"].sum() == 250
    p2 = R.workload(2, 0xC0F10002)[0]
    assert p2.n_paths == 6000 and (p2.attrs["color"] == 0).all() and (p2.attrs["eo"] == 0).all()
    pages = R.workload(3, 0xC0F10003, 3)
    assert [(p.width, p.height) for p in pages] == [(2550, 3300)] * 3
    for p in pages:
        assert 200 <= p.n_paths <= 20000
