"""Full-page GPU renders against the oracle fingerprints in tests/golden/."""
import hashlib
import json
import os

import numpy as np
import pytest

import rr_hip as R

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "page_hashes.json")))


@pytest.mark.parametrize("name,cfg,seed", [("glyph-page", 2, 0xC0F10002),
                                           ("polygon-page", 1, 0xC0F10001)])
@pytest.mark.parametrize("mode", ["documented", "literal"])
def test_full_page_matches_oracle_fingerprint(backend, name, cfg, seed, mode):
    pages = R.workload(cfg, seed)
    m = R.MODE_LITERAL if mode == "literal" else R.MODE_DOCUMENTED
    for launch in (R.LAUNCH_PLAIN, R.LAUNCH_BATCHED):
        db = R.DeviceBatch(backend, pages)
        try:
            db.run(m, R.FLAVOUR_GRAY8, launch, hashes=True)
            d = hashlib.sha256()
            for i in range(len(pages)):
                d.update(db.read_canvas(i).tobytes())
            d.update(db.read_hashes().astype("<u8").tobytes())
        finally:
            db.free()
        assert d.hexdigest() == GOLD[f"{name}:{mode}"], f"launch {launch}"


def _canvas_digest(db, n):
    d = hashlib.sha256()
    for i in range(n):
        d.update(db.read_canvas(i).tobytes())
    return d.hexdigest()


@pytest.mark.parametrize("mode", ["documented", "literal"])
def test_mixed64_whole_batch_in_the_timed_configuration(backend, mode):
    # All 64 mixed pages (Letter, A3, A4@600, thumbnails) in one batched
    # recording: occlusion on (no coverage outputs, as timed), eager and as a
    # replayed graph, then occlusion off with coverage hashes; every run's
    # canvases match the oracle's, and the last its coverage hashes too.
    pages = R.workload(3, 0xC0F10003)
    m = R.MODE_LITERAL if mode == "literal" else R.MODE_DOCUMENTED
    db = R.DeviceBatch(backend, pages)
    g = None
    try:
        db.run(m, R.FLAVOUR_GRAY8, R.LAUNCH_BATCHED, hashes=False, variant=4)
        assert _canvas_digest(db, len(pages)) == GOLD[f"mixed-64:{mode}:canvas"], "eager, occlusion on"
        g = db.capture(m, R.FLAVOUR_GRAY8, variant=4)
        db.run_graph(g)
        assert _canvas_digest(db, len(pages)) == GOLD[f"mixed-64:{mode}:canvas"], "graph replay"
        db.run(m, R.FLAVOUR_GRAY8, R.LAUNCH_BATCHED, hashes=True, variant=4)
        d = hashlib.sha256()
        for i in range(len(pages)):
            d.update(db.read_canvas(i).tobytes())
        d.update(db.read_hashes().astype("<u8").tobytes())
        assert d.hexdigest() == GOLD[f"mixed-64:{mode}"], "occlusion off, coverage hashes"
    finally:
        if g:
            db.destroy_graph(g)
        db.free()


@pytest.mark.parametrize("index,kind", [(40, "A3@300"), (52, "A4@600"), (60, "thumbnail@150")])
def test_mixed64_single_pages_every_route(backend, index, kind):
    # One page of each size class alone: batched with occlusion on and off,
    # and the plain one-path-per-launch route, against the oracle's page.
    pg = R.workload(3, 0xC0F10003)[index]
    want = GOLD["mixed-64:documented:pages"][index]
    db = R.DeviceBatch(backend, [pg])
    try:
        for launch, hashes in ((R.LAUNCH_BATCHED, False), (R.LAUNCH_BATCHED, True), (R.LAUNCH_PLAIN, False)):
            db.run(R.MODE_DOCUMENTED, R.FLAVOUR_GRAY8, launch, hashes=hashes, variant=4)
            assert _canvas_digest(db, 1) == want, f"{kind}: launch {launch}, hashes {hashes}"
    finally:
        db.free()
