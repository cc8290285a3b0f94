"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

page_hashes.json: for each named page workload and AA row mode, the sha256 of
  the oracle's final canvas bytes (H*W, row-major, page after page) followed
  by all per-path coverage hashes (uint64 little-endian) -- the same
  fingerprint bench.py computes from the GPU output; "<name>:<mode>:canvas"
  the sha256 of the canvas bytes alone (what the timed run, which computes no
  coverage hashes, is checked against); "<name>:<mode>:pages" the sha256 of
  each page's canvas of a multi-page workload; "images-32" the canvases of the
  image workload (32 images, 4 per page).
tiny_masks.npz: full coverage masks (gamma-mapped shape per pixel) and final
  canvases of a few hand-built paths, for exact regression of the oracle.

Run: python tests/golden/make_golden.py
"""
import hashlib
import json
import multiprocessing as mp
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "pdf-raster_amd"))

import numpy as np  # noqa: E402

import rr_hip as R  # noqa: E402
from oracle import oracle as O  # noqa: E402

PAGES = {"glyph-page": (2, 0xC0F10002, 0), "polygon-page": (1, 0xC0F10001, 0),
         "mixed-64": (3, 0xC0F10003, 0)}
MODES = {"documented": O.MODE_DOCUMENTED, "literal": O.MODE_LITERAL}


_JOB = None  # (pages, mode) or image set, shared with forked workers


def _render(i):
    pages, mode = _JOB
    pg = pages[i]
    c = np.full((pg.height, pg.width), 255, np.uint8)
    a = pg.attrs
    _, hs = O.fill_page(c, pg.edges, pg.edge_off, a["eo"], a["color"], a["a_input"], mode)
    return c.tobytes(), hs


def fingerprint(pages, mode):
    """(sha256 over every page's canvas bytes in order, then all coverage
    hashes; sha256 over the canvas bytes alone), pages rendered in parallel."""
    global _JOB
    _JOB = (pages, mode)
    d, dc = hashlib.sha256(), hashlib.sha256()
    hashes, per_page = [], []
    with mp.get_context("fork").Pool(min(8, os.cpu_count() or 1)) as pool:
        for cb, hs in pool.imap(_render, range(len(pages))):
            d.update(cb)
            dc.update(cb)
            per_page.append(hashlib.sha256(cb).hexdigest())
            hashes.append(hs)
    d.update(np.concatenate(hashes).astype("<u8").tobytes())
    return d.hexdigest(), dc.hexdigest(), per_page


def _place_page(i):
    iset = _JOB
    c = np.full((iset.page_h, iset.page_w), 255, np.uint8)
    for im in iset.images:
        if im.page == i:
            grid, g = (iset.grid4, iset.grid4_n) if im.channels == 4 else (iset.grid3, iset.grid3_n)
            O.place_image(c, im.pixels, im.smask, im.ctm, im.blend, grid, g)
    return c.tobytes()


def image_fingerprint(iset):
    """sha256 over the image workload's page canvases in order."""
    global _JOB
    _JOB = iset
    d = hashlib.sha256()
    with mp.get_context("fork").Pool(min(8, os.cpu_count() or 1)) as pool:
        for cb in pool.imap(_place_page, range(iset.n_pages)):
            d.update(cb)
    return d.hexdigest()


def tiny_cases():
    out = {}
    def tri(p):
        p.move_to(1.3, 0.7)
        p.line_to(14.6, 5.2)
        p.line_to(4.1, 11.9)
        p.close(False)
    def star(p):
        pts = [(8, 0.5), (10, 6), (15.5, 6), (11, 9.5), (13, 15), (8, 11.5), (3, 15), (5, 9.5), (0.5, 6), (6, 6)]
        p.move_to(*pts[0])
        for q in pts[1:]:
            p.line_to(*q)
        p.close(False)
    def blob(p):
        p.move_to(8, 1)
        p.curve_to(15, 1, 15, 15, 8, 15)
        p.curve_to(1, 15, 1, 1, 8, 1)
        p.close(False)
    for name, fn in (("tri", tri), ("star", star), ("blob", blob)):
        for eo in (0, 1):
            p = O.Path()
            fn(p)
            e = O.xpath_edges(*p.arrays(), flatness=0.1)
            canvas = np.full((16, 16), 255, np.uint8)
            _, _, shape = O.fill(canvas, e, eo, 40, want_shape=True)
            out[f"{name}_eo{eo}_shape"] = shape
            out[f"{name}_eo{eo}_canvas"] = canvas
            out[f"{name}_eo{eo}_edges"] = e
    return out


def main():
    res = {}
    for name, (cfg, seed, max_pages) in PAGES.items():
        pages = R.workload(cfg, seed, max_pages)
        for mname, mode in MODES.items():
            full, canvas, per_page = fingerprint(pages, mode)
            res[f"{name}:{mname}"] = full
            res[f"{name}:{mname}:canvas"] = canvas
            if len(pages) > 1:
                res[f"{name}:{mname}:pages"] = per_page
            print(name, mname, full[:16], canvas[:16], flush=True)
        del pages
    res["images-32"] = image_fingerprint(R.images(0xC0F10004, 32, 4))
    print("images-32", res["images-32"][:16], flush=True)
    with open(os.path.join(HERE, "page_hashes.json"), "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
    np.savez_compressed(os.path.join(HERE, "tiny_masks.npz"), **tiny_cases())


if __name__ == "__main__":
    main()
