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
