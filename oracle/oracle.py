"""ctypes binding of the CPU oracle (oracle/rr_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the reference the HIP results are
compared with or the timed CPU baseline, never as part of the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "librr_oracle.so")

MODE_DOCUMENTED = 0
MODE_LITERAL = 1
FLAVOUR_GRAY8 = 0
FLAVOUR_RGB8_FAST = 1
FLAVOUR_RGB8 = 2
FILL_EVEN_ODD, FILL_MULTIPLY, FILL_SOFT_MASK = 1, 2, 4

PATH_FIRST, PATH_LAST, PATH_CLOSED, PATH_CURVE = 1, 2, 4, 8
SEG_HORIZ, SEG_VERT, SEG_FLIPPED = 1, 2, 4


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class Seg(ctypes.Structure):
    _fields_ = [("x0", ctypes.c_double), ("y0", ctypes.c_double), ("x1", ctypes.c_double),
                ("y1", ctypes.c_double), ("dxdy", ctypes.c_double), ("dxdy_fp", ctypes.c_int32),
                ("flags", ctypes.c_uint32)]


class Intersect(ctypes.Structure):
    _fields_ = [("x0", ctypes.c_int32), ("x1", ctypes.c_int32), ("count", ctypes.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


def mutant_path(m):
    """Path of the oracle build with planted mistake m (see RRO_MUTANT in rr_oracle.c)."""
    return os.path.join(_HERE, "build", f"librr_oracle_m{m}.so")


class using:
    """Context manager: route every function of this module through the oracle
    library at `path` (a planted-mistake build) while the block runs."""

    def __init__(self, path):
        self.L = load(path)

    def __enter__(self):
        global _lib
        self.saved = _lib
        _lib = self.L
        return self.L

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved
        return False


def load(path):
    """A CDLL of an oracle build with the argument and result types bound."""
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    i64, i32, u8, u32, dbl = (ctypes.c_int64, ctypes.c_int32, ctypes.c_uint8,
                              ctypes.c_uint32, ctypes.c_double)
    sig = {
        "rro_div255": (u8, [u32]),
        "rro_splash_floor": (i32, [dbl]),
        "rro_splash_ceil": (i32, [dbl]),
        "rro_splash_round": (i32, [dbl]),
        "rro_fp_from_f64": (i32, [dbl]),
        "rro_path_new": (P, []),
        "rro_path_free": (None, [P]),
        "rro_path_move_to": (ctypes.c_int, [P, dbl, dbl]),
        "rro_path_line_to": (ctypes.c_int, [P, dbl, dbl]),
        "rro_path_curve_to": (ctypes.c_int, [P, dbl, dbl, dbl, dbl, dbl, dbl]),
        "rro_path_close": (ctypes.c_int, [P, ctypes.c_int]),
        "rro_path_len": (i64, [P]),
        "rro_path_get": (None, [P, P, P]),
        "rro_conic_to_cubic": (None, [P, P, P, P, P]),
        "rro_flatten_curve": (i64, [P, dbl, P, i64]),
        "rro_xpath_edges": (i64, [P, P, i64, P, dbl, ctypes.c_int, P, i64]),
        "rro_add_segments": (None, [P, i64, P]),
        "rro_aa_scale": (None, [P, i64]),
        "rro_scanner_new": (P, [P, i64, ctypes.c_int, i32, i32]),
        "rro_scanner_free": (None, [P]),
        "rro_scanner_bbox": (None, [P, P]),
        "rro_scanner_row": (i64, [P, i32, P, i64]),
        "rro_scanner_test": (ctypes.c_int, [P, i32, i32]),
        "rro_scanner_render_aa_line": (ctypes.c_int, [P, P, i32, P, P, i32]),
        "rro_scanner_stats": (None, [P, P]),
        "rro_aabuf_set_span": (None, [P, i32, i32, i32]),
        "rro_aa_coverage_span": (None, [P, i32, i32, i32, P]),
        "rro_composite_gray8": (u8, [u8, u8, u8, u8]),
        "rro_composite_rgb8_fast": (u8, [u8, u8, u8, u8]),
        "rro_rgb_to_gray": (i32, [i32, i32, i32]),
        "rro_fill": (i64, [P, i32, i32, i32, P, i64, ctypes.c_int, u8, u8, ctypes.c_int,
                           ctypes.c_int, P, P, P]),
        "rro_fill_page": (i64, [P, i32, i32, i32, P, P, i64, P, P, P, ctypes.c_int,
                                ctypes.c_int, P]),
        "rro_fill_page_timed": (i64, [P, i32, i32, i32, P, P, i64, P, P, P, ctypes.c_int, P]),
        "rro_cov_hash_term": (ctypes.c_uint64, [i64, u8]),
        "rro_compute_a_src": (u32, [u32, u32, u32, ctypes.c_int, ctypes.c_int]),
        "rro_composite_general": (u8, [u8, u8, u8, u8, u8, ctypes.c_int, ctypes.c_int]),
        "rro_fill_ex": (i64, [P, i32, i32, i32, P, i64, u32, P, u8, ctypes.c_int, ctypes.c_int, P, P, i32,
                              P, P]),
        "rro_fill_page_ex": (i64, [P, i32, i32, i32, P, P, i64, P, P, P, ctypes.c_int, ctypes.c_int, P, i32,
                                   P]),
        "rro_rgb_to_gray_page": (None, [P, i32, P, i32, i32, i32]),
        "rro_ref_bbox_area": (i64, [P, i64, i32, i32]),
        "rro_blend_u8": (u8, [u8, u8, u32]),
        "rro_blend_multiply": (u8, [u8, u8]),
        "rro_composite_rgba8": (None, [P, P, i64]),
        "rro_apply_soft_mask": (None, [P, P, i64]),
        "rro_cmyk_matrix": (None, [P, P, i64]),
        "rro_icc_clut_ref": (None, [P, P, i64, P, i32]),
        "rro_clut_gray": (u8, [P, i32, P, i32]),
        "rro_place_image": (i64, [P, i32, i32, i32, P, i32, i32, i32, P, P, i32, P, i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


AA_GAMMA = None


def aa_gamma():
    return np.frombuffer((ctypes.c_uint8 * 17).in_dll(lib(), "RRO_AA_GAMMA"), dtype=np.uint8).copy()


class Path:
    """PathBuilder restatement (crates/raster/src/path/mod.rs:291-451)."""

    def __init__(self):
        self.h = lib().rro_path_new()

    def __del__(self):
        if getattr(self, "h", None):
            lib().rro_path_free(self.h)
            self.h = None

    def move_to(self, x, y):
        return lib().rro_path_move_to(self.h, x, y)

    def line_to(self, x, y):
        return lib().rro_path_line_to(self.h, x, y)

    def curve_to(self, x1, y1, x2, y2, x3, y3):
        return lib().rro_path_curve_to(self.h, x1, y1, x2, y2, x3, y3)

    def close(self, force=False):
        return lib().rro_path_close(self.h, int(force))

    def arrays(self):
        n = lib().rro_path_len(self.h)
        xy = np.zeros((max(n, 1), 2), np.float64)
        fl = np.zeros(max(n, 1), np.uint8)
        lib().rro_path_get(self.h, _p(xy), _p(fl))
        return xy[:n].copy(), fl[:n].copy()


IDENTITY = np.array([1.0, 0.0, 0.0, 1.0, 0.0, 0.0])


def xpath_edges(xy, flags, matrix=IDENTITY, flatness=0.1, close_subpaths=True):
    xy = np.ascontiguousarray(xy, np.float64)
    flags = np.ascontiguousarray(flags, np.uint8)
    m = np.ascontiguousarray(matrix, np.float64)
    n = len(flags)
    cap = max(16, 8 * n)
    while True:
        out = np.zeros((cap, 4), np.float64)
        k = lib().rro_xpath_edges(_p(xy), _p(flags), n, _p(m), flatness, int(close_subpaths),
                                  _p(out), cap)
        if k <= cap:
            return out[:k].copy()
        cap = int(k)


def flatten_curve(p, flatness_sq):
    p = np.ascontiguousarray(p, np.float64)
    out = np.zeros((1100, 2), np.float64)
    k = lib().rro_flatten_curve(_p(p), flatness_sq, _p(out), 1100)
    return out[:k].copy()


def add_segments(edges):
    edges = np.ascontiguousarray(edges, np.float64).reshape(-1, 4)
    out = (Seg * max(len(edges), 1))()
    lib().rro_add_segments(_p(edges), len(edges), ctypes.cast(out, ctypes.c_void_p))
    return [out[i] for i in range(len(edges))]


class Scanner:
    """XPathScanner over add_segment'ed, optionally aa_scale'd, edges."""

    def __init__(self, edges, eo, clip_y_min, clip_y_max, aa=False):
        edges = np.ascontiguousarray(edges, np.float64).reshape(-1, 4)
        n = len(edges)
        self._segs = (Seg * max(n, 1))()
        sp = ctypes.cast(self._segs, ctypes.c_void_p)
        lib().rro_add_segments(_p(edges), n, sp)
        if aa:
            lib().rro_aa_scale(sp, n)
        self.h = lib().rro_scanner_new(sp, n, int(eo), clip_y_min, clip_y_max)

    def __del__(self):
        if getattr(self, "h", None):
            lib().rro_scanner_free(self.h)
            self.h = None

    def bbox(self):
        b = np.zeros(4, np.int32)
        lib().rro_scanner_bbox(self.h, _p(b))
        return tuple(int(v) for v in b)

    def row(self, y):
        n = lib().rro_scanner_row(self.h, y, None, 0)
        out = (Intersect * max(n, 1))()
        lib().rro_scanner_row(self.h, y, ctypes.cast(out, ctypes.c_void_p), n)
        return [(out[i].x0, out[i].x1, out[i].count) for i in range(n)]

    def test(self, x, y):
        return bool(lib().rro_scanner_test(self.h, x, y))

    def stats(self):
        s = np.zeros(4, np.int64)
        lib().rro_scanner_stats(self.h, _p(s))
        return [int(v) for v in s]

    def render_aa_line(self, aa_width, y):
        rb = (aa_width + 7) // 8
        buf = np.zeros((4, rb), np.uint8)
        x0 = ctypes.c_int32(0)
        x1 = ctypes.c_int32(0)
        unb = lib().rro_scanner_render_aa_line(self.h, _p(buf), aa_width, ctypes.byref(x0),
                                               ctypes.byref(x1), y)
        return buf, x0.value, x1.value, unb


def page_clip(w, h):
    return np.array([0.0, 0.0, w - 0.001, h - 0.001])


def fill(canvas, edges, eo, color, a_input=255, mode=MODE_DOCUMENTED, flavour=FLAVOUR_GRAY8,
         clip=None, want_shape=False):
    """rro_fill on a 2-D uint8 canvas (modified in place).

    Returns (drawn_pixels, cov_hash, shape-or-None)."""
    h, w = canvas.shape
    assert canvas.dtype == np.uint8 and canvas.flags.c_contiguous
    edges = np.ascontiguousarray(edges, np.float64).reshape(-1, 4)
    clip = page_clip(w, h) if clip is None else np.ascontiguousarray(clip, np.float64)
    shape = np.zeros((h, w), np.uint8) if want_shape else None
    hsh = ctypes.c_uint64(0)
    d = lib().rro_fill(_p(canvas), w, h, w, _p(edges), len(edges), int(eo), color, a_input,
                       mode, flavour, _p(clip), _p(shape), ctypes.byref(hsh))
    return int(d), hsh.value, shape


def fill_page(canvas, edges, edge_off, eo, color, a_input, mode=MODE_DOCUMENTED,
              flavour=FLAVOUR_GRAY8, hashes=True):
    """Render one page of paths in order.  Returns (drawn_pixels, hashes)."""
    h, w = canvas.shape
    stride = canvas.strides[0]
    edges = np.ascontiguousarray(edges, np.float64)
    edge_off = np.ascontiguousarray(edge_off, np.int64)
    n = len(edge_off) - 1
    eo = np.ascontiguousarray(eo, np.uint8)
    color = np.ascontiguousarray(color, np.uint8)
    a_input = np.ascontiguousarray(a_input, np.uint8)
    hs = np.zeros(n, np.uint64) if hashes else None
    d = lib().rro_fill_page(_p(canvas), w, h, stride, _p(edges), _p(edge_off), n, _p(eo),
                            _p(color), _p(a_input), mode, flavour, _p(hs))
    return int(d), hs


def fill_ex(canvas, edges, flags, color, a_input=255, mode=MODE_DOCUMENTED, flavour=FLAVOUR_GRAY8,
            clip=None, soft_mask=None, want_shape=False):
    """rro_fill_ex on a (H, W) grey or (H, W, 3) Rgb8 uint8 canvas (in place).
    color: a grey value or an (r, g, b) triple.  Returns (drawn, cov_hash, shape)."""
    h, w = canvas.shape[:2]
    assert canvas.dtype == np.uint8 and canvas.flags.c_contiguous
    assert (canvas.ndim == 3) == (flavour == FLAVOUR_RGB8)
    edges = np.ascontiguousarray(edges, np.float64).reshape(-1, 4)
    clip = page_clip(w, h) if clip is None else np.ascontiguousarray(clip, np.float64)
    c3 = np.ascontiguousarray(np.broadcast_to(np.asarray(color, np.uint8), (3,)))
    sm = None if soft_mask is None else np.ascontiguousarray(soft_mask, np.uint8)
    shape = np.zeros((h, w), np.uint8) if want_shape else None
    hsh = ctypes.c_uint64(0)
    d = lib().rro_fill_ex(_p(canvas), w, h, canvas.strides[0], _p(edges), len(edges), int(flags), _p(c3),
                          a_input, mode, flavour, _p(clip), _p(sm), 0 if sm is None else sm.strides[0],
                          _p(shape), ctypes.byref(hsh))
    return int(d), hsh.value, shape


def fill_page_ex(canvas, edges, edge_off, flags, colors, a_input, mode=MODE_DOCUMENTED,
                 flavour=FLAVOUR_GRAY8, soft_mask=None, hashes=True):
    """A page of paths in order on a grey or Rgb8 canvas.  colors: (n, 3)
    uint8 (a grey page reads column 0).  Returns (drawn, hashes)."""
    h, w = canvas.shape[:2]
    edges = np.ascontiguousarray(edges, np.float64)
    edge_off = np.ascontiguousarray(edge_off, np.int64)
    n = len(edge_off) - 1
    flags = np.ascontiguousarray(flags, np.uint8)
    colors = np.ascontiguousarray(colors, np.uint8).reshape(n, 3)
    a_input = np.ascontiguousarray(a_input, np.uint8)
    sm = None if soft_mask is None else np.ascontiguousarray(soft_mask, np.uint8)
    hs = np.zeros(n, np.uint64) if hashes else None
    d = lib().rro_fill_page_ex(_p(canvas), w, h, canvas.strides[0], _p(edges), _p(edge_off), n, _p(flags),
                               _p(colors), _p(a_input), mode, flavour, _p(sm),
                               0 if sm is None else sm.strides[0], _p(hs))
    return int(d), hs


def ref_bbox_area(edges, w, h):
    """The reference dispatcher's bbox area of one path (-1: stays on the CPU)."""
    edges = np.ascontiguousarray(edges, np.float64).reshape(-1, 4)
    return int(lib().rro_ref_bbox_area(_p(edges), len(edges), w, h))


def rgb_to_gray_page(rgb):
    """Gray8 of an (H, W, 3) Rgb8 bitmap (render.rs:1306-1324)."""
    h, w = rgb.shape[:2]
    rgb = np.ascontiguousarray(rgb, np.uint8)
    out = np.zeros((h, w), np.uint8)
    lib().rro_rgb_to_gray_page(_p(rgb), rgb.strides[0], _p(out), w, w, h)
    return out


def cov_hash_term(pix, shape):
    return lib().rro_cov_hash_term(pix, shape)


# ---- images and colour (rr_oracle_img.c) ----

def blend_u8(src, dst, alpha):
    return lib().rro_blend_u8(src, dst, alpha)


def blend_multiply(src, dst):
    return lib().rro_blend_multiply(src, dst)


def composite_rgba8(src, dst):
    """In-place source-over of RGBA8 src onto dst (both (n, 4) uint8)."""
    src = np.ascontiguousarray(src, np.uint8)
    assert dst.flags.c_contiguous and dst.dtype == np.uint8 and dst.size == src.size
    lib().rro_composite_rgba8(_p(src), _p(dst), src.size // 4)
    return dst


def apply_soft_mask(px, mask):
    mask = np.ascontiguousarray(mask, np.uint8)
    assert px.flags.c_contiguous and px.dtype == np.uint8 and px.size == 4 * mask.size
    lib().rro_apply_soft_mask(_p(px), _p(mask), mask.size)
    return px


def cmyk_matrix(cmyk):
    cmyk = np.ascontiguousarray(cmyk, np.uint8).reshape(-1, 4)
    out = np.zeros((len(cmyk), 3), np.uint8)
    lib().rro_cmyk_matrix(_p(cmyk), _p(out), len(cmyk))
    return out


def icc_clut(cmyk, table, g):
    """icc_cmyk_to_rgb_cpu's CLUT arm: table is g^4*3 u8, K-major."""
    cmyk = np.ascontiguousarray(cmyk, np.uint8).reshape(-1, 4)
    table = np.ascontiguousarray(table, np.uint8)
    assert table.size == g ** 4 * 3
    out = np.zeros((len(cmyk), 3), np.uint8)
    lib().rro_icc_clut_ref(_p(cmyk), _p(out), len(cmyk), _p(table), g)
    return out


def clut_gray(px, grid, g):
    px = np.ascontiguousarray(px, np.uint8)
    grid = np.ascontiguousarray(grid, np.uint16)
    assert grid.size == g ** len(px)
    return lib().rro_clut_gray(_p(px), len(px), _p(grid), g)


def place_image(canvas, pixels, smask, ctm, blend, grid, g):
    """Place one image on a grey canvas ((h, stride) view, width = shape[1]).
    Returns the number of pixels written, or -1 for an axis-aligned CTM."""
    h, w = canvas.shape
    stride = canvas.strides[0]
    pixels = np.ascontiguousarray(pixels, np.uint8)
    ih, iw, ch = pixels.shape
    sm = None if smask is None else np.ascontiguousarray(smask, np.uint8)
    m = np.ascontiguousarray(ctm, np.float64)
    grid = np.ascontiguousarray(grid, np.uint16)
    return lib().rro_place_image(_p(canvas), w, h, stride, _p(pixels), iw, ih, ch, _p(sm), _p(m),
                                 blend, _p(grid), g)
