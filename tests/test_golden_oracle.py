"""The oracle reproduces the committed tiny fixtures exactly (regression pin)."""
import os

import numpy as np

from oracle import oracle as O

HERE = os.path.join(os.path.dirname(__file__), "golden")


def test_tiny_masks_reproduce():
    z = np.load(os.path.join(HERE, "tiny_masks.npz"))
    names = sorted({k.rsplit("_", 1)[0] for k in z.files})
    assert len(names) == 6
    for n in names:
        e = z[n + "_edges"]
        eo = int(n[-1])
        canvas = np.full((16, 16), 255, np.uint8)
        _, _, shape = O.fill(canvas, e, eo, 40, want_shape=True)
        assert np.array_equal(shape, z[n + "_shape"]), n
        assert np.array_equal(canvas, z[n + "_canvas"]), n
        assert shape.max() == 255 and (shape > 0).sum() > 20
