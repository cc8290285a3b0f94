"""The dispatcher: every route gives the oracle's pixels and coverage hashes;
the in-kernel MLP forward pass matches a float32 numpy restatement; the HIP
training kernel fits a known function."""
import numpy as np
import pytest

import rr_hip as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _oracle(pages):
    canv, hs = [], []
    for pg in pages:
        c = np.full((pg.height, pg.width), 255, np.uint8)
        a = pg.attrs
        _, h = O.fill_page(c, pg.edges, pg.edge_off, a["eo"], a["color"], a["a_input"])
        canv.append(c)
        hs.append(h)
    return canv, np.concatenate(hs)


def _np_forward(P, F):
    mu, sd = P[0:12], P[12:24]
    W1 = P[24:792].reshape(64, 12); b1 = P[792:856]
    W2 = P[856:4952].reshape(64, 64); b2 = P[4952:5016]
    W3 = P[5016:5336].reshape(5, 64); b3 = P[5336:5341]
    x = ((F - mu) / sd).astype(np.float32)
    h1 = np.maximum(x @ W1.T + b1, 0).astype(np.float32)
    h2 = np.maximum(h1 @ W2.T + b2, 0).astype(np.float32)
    return (h2 @ W3.T + b3).astype(np.float32)


def _random_params(rng):
    P = rng.normal(0, 0.3, R.MLP_PARAMS).astype(np.float32)
    P[0:12] = rng.normal(0, 1, 12)
    P[12:24] = rng.uniform(0.5, 2.0, 12)
    return P


@pytest.fixture(scope="module")
def pages():
    glyph = R.workload(2, 0xC0F10002, 1)[0].subset(range(0, 6000, 7))
    mixed = [pg.subset(range(0, pg.n_paths, 5)) for pg in R.workload(3, 0xC0F10003, 2)]
    return [glyph] + mixed


@pytest.mark.parametrize("route", [R.ROUTE_SPAN, R.ROUTE_TILE16, R.ROUTE_TILE32, R.ROUTE_TILE64, R.ROUTE_CPU,
                                   R.ROUTE_AUTO])
def test_every_route_is_bit_exact(backend, pages, route):
    oc, oh = _oracle(pages)
    db = R.DeviceBatch(backend, pages)
    try:
        backend.set_dispatcher(None)
        db.set_routes(route)
        db.run(launch=R.LAUNCH_BATCHED, hashes=True, variant=2)
        for i, o in enumerate(oc):
            assert np.array_equal(db.read_canvas(i), o), f"page {i}"
        assert np.array_equal(db.read_hashes(), oh)
        _, cls = backend.last_dispatch(db.n_paths)
        if route == R.ROUTE_AUTO:  # static rule: the reference's bbox area < 256 -> CPU
            want_cpu = np.concatenate([[O.ref_bbox_area(pg.path_edges(i), pg.width, pg.height) < 256
                                        for i in range(pg.n_paths)] for pg in pages])
            assert np.array_equal(cls == R.ROUTE_CPU, want_cpu)
            assert set(np.unique(cls)) <= {0, 1, 2, 3, 4}
        elif route in (R.ROUTE_SPAN, R.ROUTE_CPU):
            assert np.all(cls == route)
    finally:
        db.free()


def _rect(x0, y0, x1, y1):
    return [[x0, y0, x1, y0], [x1, y0, x1, y1], [x1, y1, x0, y1], [x0, y1, x0, y0]]


def test_static_rule_around_the_threshold(backend):
    # Rectangles whose reference bbox area (gpu_ops.rs:52-84) falls just
    # below, at and above 256 px: fractional corners (floor / ceil), parts
    # off the page (clamped), a NaN endpoint (skipped) and an infinite one
    # (no bbox: CPU).  Without an MLP, AUTO takes the CPU class exactly when
    # the area is below 256 (crates/gpu/src/lib.rs:112,120), and every class
    # still gives the oracle's pixels.
    from test_gpu_batched import edge_page
    rng = np.random.default_rng(31)
    W, H = 300, 200
    paths = []
    for k in range(160):
        x0, y0 = rng.uniform(-6, W - 20), rng.uniform(-6, H - 20)
        w = rng.choice([14.0, 15.0, 15.5, 16.0, 16.2, 17.0]) + rng.choice([0.0, 0.3, 0.75])
        h = rng.choice([14.0, 15.0, 15.9, 16.0, 17.0])
        e = _rect(x0, y0, x0 + w, y0 + h)
        if k % 40 == 7:
            e[1][2] = float("nan")
        if k % 40 == 9:
            e[0][2] = float("inf")
        paths.append((e, k % 2, int(rng.integers(0, 256)), 255))
    pg = edge_page(W, H, paths)
    areas = np.array([O.ref_bbox_area(pg.path_edges(i), W, H) for i in range(pg.n_paths)])
    assert (areas < 256).any() and (areas >= 256).any() and (areas == 256).any()
    db = R.DeviceBatch(backend, [pg])
    try:
        backend.set_dispatcher(None)
        db.set_routes(R.ROUTE_AUTO)
        db.run(launch=R.LAUNCH_BATCHED, hashes=True, variant=0)
        _, cls = backend.last_dispatch(db.n_paths)
        assert np.array_equal(cls == R.ROUTE_CPU, areas < 256), np.nonzero((cls == R.ROUTE_CPU) != (areas < 256))
        oc, oh = _oracle([pg])
        assert np.array_equal(db.read_canvas(0), oc[0])
        assert np.array_equal(db.read_hashes(), oh)
    finally:
        db.free()


def test_mlp_forward_matches_numpy(backend):
    rng = np.random.default_rng(7)
    P = _random_params(rng)
    F = rng.normal(0, 2, (1000, 12)).astype(np.float32)
    backend.set_dispatcher(P)
    try:
        got = backend.mlp_forward(F)
    finally:
        backend.set_dispatcher(None)
    want = _np_forward(P, F)
    assert np.allclose(got, want, rtol=1e-4, atol=1e-4), np.abs(got - want).max()


def test_learned_routes_are_bit_exact(backend, pages):
    # an MLP that prefers the tile classes where valid: AUTO then executes them
    rng = np.random.default_rng(3)
    P = _random_params(rng)
    P[5016:5336] = 0.0
    P[5336:5341] = [5.0, 1.0, 2.0, 3.0, 9.0]
    oc, oh = _oracle(pages)
    db = R.DeviceBatch(backend, pages)
    try:
        backend.set_dispatcher(P)
        db.set_routes(R.ROUTE_AUTO)
        db.run(launch=R.LAUNCH_BATCHED, hashes=True, variant=2)
        feats, cls = backend.last_dispatch(db.n_paths)
        bands = feats[:, 5]
        want = np.where(bands <= 1, 1, np.where(bands <= 2, 2, np.where(bands <= 4, 3, 0)))
        assert np.array_equal(cls, want)
        for i, o in enumerate(oc):
            assert np.array_equal(db.read_canvas(i), o), f"page {i}"
        assert np.array_equal(db.read_hashes(), oh)
    finally:
        backend.set_dispatcher(None)
        db.free()


def test_training_fits_a_known_function(backend):
    # targets: a smooth function of two features per class, one class masked
    rng = np.random.default_rng(11)
    F = rng.normal(0, 1, (4096, 12)).astype(np.float32)
    T = np.stack([F[:, 0] + 0.5 * F[:, 1], -F[:, 2], np.abs(F[:, 3]), F[:, 4] * 0.25 + 1.0,
                  np.full(len(F), np.nan)], 1).astype(np.float32)
    P, loss = backend.train_dispatcher(F, T, steps=3000, batch=64, lr=3e-3, seed=5)
    assert np.all(np.isfinite(P))
    assert loss < 0.05, loss
    backend.set_dispatcher(P)
    try:
        got = backend.mlp_forward(F[:512])
    finally:
        backend.set_dispatcher(None)
    err = np.abs(got[:, :4] - T[:512, :4])
    assert np.mean(err) < 0.15, np.mean(err)
    assert np.allclose(got, _np_forward(P, F[:512]), rtol=1e-4, atol=1e-4)
