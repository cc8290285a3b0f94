"""ctypes binding of the HIP backend's C-ABI (include/rr_hip.h).

This is the host-side harness the tests and bench use; it holds no compute of
its own.  Every fill runs in librr_hip.so's gfx950 kernels, and loading fails
loudly when the library is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librr_hip.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "rr_hip.h")

RR_OK = 0
MODE_DOCUMENTED, MODE_LITERAL = 0, 1
FLAVOUR_GRAY8, FLAVOUR_RGB8_FAST, FLAVOUR_RGB8 = 0, 1, 2
FILL_EVEN_ODD, FILL_MULTIPLY, FILL_SOFT_MASK = 1, 2, 4
LAUNCH_PLAIN, LAUNCH_BATCHED = 0, 1
SCAN_STATIC_MAX = 1 << 24  # elements per rr_hip_record_scan_static call
RR_OK, RR_E_ZERO_SIZE_ALLOC, RR_E_INVARIANT, RR_E_DEVICE, RR_E_NO_DEVICE, RR_E_CAPACITY = 0, -1, -2, -3, -4, -5
ROUTE_SPAN, ROUTE_TILE16, ROUTE_TILE32, ROUTE_TILE64, ROUTE_CPU, ROUTE_AUTO = 0, 1, 2, 3, 4, 255
MLP_FEATURES, MLP_CLASSES, MLP_PARAMS = 12, 5, 5341

PATH_ATTR_DTYPE = np.dtype([("eo", np.uint8), ("color", np.uint8), ("a_input", np.uint8),
                            ("route", np.uint8)])


class RrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rr_hip error {code}: {msg}")
        self.code = code


class FillBatch(ctypes.Structure):
    _fields_ = [("edges", ctypes.c_void_p), ("edge_off", ctypes.c_void_p),
                ("attrs", ctypes.c_void_p), ("host_rows", ctypes.c_void_p),
                ("n_paths", ctypes.c_int64), ("n_edges", ctypes.c_int64),
                ("canvas", ctypes.c_void_p), ("width", ctypes.c_int32),
                ("height", ctypes.c_int32), ("stride", ctypes.c_int32), ("mode", ctypes.c_int32),
                ("flavour", ctypes.c_int32), ("launch", ctypes.c_int32),
                ("clip", ctypes.c_double * 4), ("cov_hash", ctypes.c_void_p),
                ("cov_count", ctypes.c_void_p), ("variant", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("colors", ctypes.c_void_p), ("soft_mask", ctypes.c_void_p),
                ("soft_mask_stride", ctypes.c_int64), ("gray", ctypes.c_void_p),
                ("gray_stride", ctypes.c_int64)]


class PageDesc(ctypes.Structure):
    _fields_ = [("canvas", ctypes.c_void_p), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("stride", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("path_begin", ctypes.c_int64), ("path_end", ctypes.c_int64),
                ("clip", ctypes.c_double * 4), ("soft_mask", ctypes.c_void_p),
                ("soft_mask_stride", ctypes.c_int64), ("gray", ctypes.c_void_p),
                ("gray_stride", ctypes.c_int64)]


class FillPages(ctypes.Structure):
    _fields_ = [("edges", ctypes.c_void_p), ("edge_off", ctypes.c_void_p),
                ("attrs", ctypes.c_void_p), ("host_rows", ctypes.c_void_p),
                ("n_paths", ctypes.c_int64), ("n_edges", ctypes.c_int64),
                ("pages", ctypes.c_void_p), ("n_pages", ctypes.c_int32),
                ("mode", ctypes.c_int32), ("flavour", ctypes.c_int32), ("launch", ctypes.c_int32),
                ("cov_hash", ctypes.c_void_p), ("cov_count", ctypes.c_void_p),
                ("variant", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("colors", ctypes.c_void_p)]


class ImageDesc(ctypes.Structure):
    _fields_ = [("pixels", ctypes.c_void_p), ("smask", ctypes.c_void_p),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("channels", ctypes.c_int32), ("blend", ctypes.c_int32),
                ("ctm", ctypes.c_double * 6), ("page", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class ImageBatch(ctypes.Structure):
    _fields_ = [("images", ctypes.c_void_p), ("n_images", ctypes.c_int32),
                ("pages", ctypes.c_void_p), ("n_pages", ctypes.c_int32),
                ("grid4", ctypes.c_void_p), ("grid4_n", ctypes.c_int32),
                ("grid3", ctypes.c_void_p), ("grid3_n", ctypes.c_int32)]


class XpathBatch(ctypes.Structure):
    _fields_ = [("xy", ctypes.c_void_p), ("flags", ctypes.c_void_p), ("pt_off", ctypes.c_void_p),
                ("matrix", ctypes.c_void_p), ("n_paths", ctypes.c_int64), ("n_pts", ctypes.c_int64),
                ("flatness", ctypes.c_double), ("edges", ctypes.c_void_p), ("edge_cap", ctypes.c_int64),
                ("edge_off", ctypes.c_void_p)]


_lib = None


def lib():
    """Load librr_hip.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        P, i32, i64, u64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64
        sig = {
            "rr_hip_create": (i32, [i32, ctypes.POINTER(ctypes.c_void_p)]),
            "rr_hip_destroy": (i32, [P]),
            "rr_hip_last_error": (ctypes.c_char_p, [P]),
            "rr_hip_alloc": (i32, [P, ctypes.c_size_t, i32, ctypes.POINTER(ctypes.c_void_p)]),
            "rr_hip_free": (i32, [P, P]),
            "rr_hip_upload": (i32, [P, P, P, ctypes.c_size_t]),
            "rr_hip_download": (i32, [P, P, P, ctypes.c_size_t]),
            "rr_hip_memset": (i32, [P, P, i32, ctypes.c_size_t]),
            "rr_hip_copy": (i32, [P, P, P, ctypes.c_size_t]),
            "rr_hip_buffer_len": (i32, [P, P, ctypes.POINTER(ctypes.c_size_t)]),
            "rr_hip_alloc_host_pinned": (i32, [P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
            "rr_hip_free_host_pinned": (i32, [P, P]),
            "rr_hip_vram_budget": (i32, [P, ctypes.POINTER(u64), ctypes.POINTER(u64)]),
            "rr_hip_upload_async": (i32, [P, P, P, ctypes.c_size_t, ctypes.POINTER(u64)]),
            "rr_hip_download_async": (i32, [P, P, P, ctypes.c_size_t, ctypes.POINTER(u64)]),
            "rr_hip_submit_transfer": (i32, [P, ctypes.POINTER(u64)]),
            "rr_hip_wait_transfer": (i32, [P, u64]),
            "rr_hip_begin_page": (i32, [P]),
            "rr_hip_record_fill_paths": (i32, [P, ctypes.POINTER(FillBatch)]),
            "rr_hip_record_fill_pages": (i32, [P, ctypes.POINTER(FillPages)]),
            "rr_hip_submit_page": (i32, [P, ctypes.POINTER(u64)]),
            "rr_hip_wait_page": (i32, [P, u64]),
            "rr_hip_inject_error": (i32, [P, i32]),
            "rr_hip_set_profiling": (i32, [P, i32]),
            "rr_hip_set_trace": (i32, [P, P, i64]),
            "rr_hip_variant_count": (i32, []),
            "rr_hip_autotune_fill": (i32, [P, ctypes.POINTER(FillPages), i32,
                                           ctypes.POINTER(i32), P]),
            "rr_hip_variant_info": (i32, [i32, ctypes.POINTER(i32), ctypes.POINTER(i32)]),
            "rr_hip_trace_count": (i32, [P, ctypes.POINTER(i64)]),
            "rr_hip_last_page_stats": (i32, [P, ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(i64)]),
            "rr_xpath_edges": (i32, [P, P, i64, P, ctypes.c_double, P, i64,
                                     ctypes.POINTER(i64)]),
            "rr_hip_xpath_edges": (i32, [P, ctypes.POINTER(XpathBatch), ctypes.POINTER(i64)]),
            "rr_hip_last_xpath_stats": (i32, [P, P]),
            "rr_cov_hash_term": (u64, [i64, ctypes.c_uint8]),
            "rr_workload_create": (i32, [i32, u64, i32, ctypes.POINTER(ctypes.c_void_p)]),
            "rr_workload_destroy": (None, [P]),
            "rr_workload_n_pages": (i32, [P]),
            "rr_workload_page_info": (i32, [P, i32, P]),
            "rr_workload_page_data": (i32, [P, i32] + [ctypes.POINTER(ctypes.c_void_p)] * 7),
            "rr_hip_record_place_images": (i32, [P, ctypes.POINTER(ImageBatch)]),
            "rr_hip_record_icc_clut": (i32, [P, P, P, P, ctypes.c_size_t, i64]),
            "rr_hip_record_composite": (i32, [P, P, P, i64]),
            "rr_hip_record_apply_soft_mask": (i32, [P, P, P, i64]),
            "rr_hip_record_rgb_to_gray": (i32, [P, P, i64, P, i64, i32, i32]),
            "rr_hip_record_scan": (i32, [P, P, P, ctypes.c_uint32]),
            "rr_hip_record_scan_static": (i32, [P, P, ctypes.c_uint32]),
            "rr_hip_last_batch_stats": (i32, [P, P, P]),
            "rr_hip_debug_batch_table": (i32, [P, i32, P, i64, ctypes.POINTER(i64)]),
            "rr_hip_set_dispatcher": (i32, [P, P, i32]),
            "rr_hip_last_dispatch": (i32, [P, P, P, i64]),
            "rr_hip_mlp_forward": (i32, [P, P, P, i64]),
            "rr_hip_train_dispatcher": (i32, [P, P, P, i64, i32, i32, ctypes.c_float, u64, P, P]),
            "rr_hip_graph_create": (i32, [P, ctypes.POINTER(FillPages), ctypes.POINTER(ctypes.c_void_p)]),
            "rr_hip_graph_launch": (i32, [P, P]),
            "rr_hip_graph_destroy": (i32, [P, P]),
            "rr_images_create": (i32, [u64, i32, i32, ctypes.POINTER(ctypes.c_void_p)]),
            "rr_images_destroy": (None, [P]),
            "rr_images_info": (i32, [P, P]),
            "rr_images_get": (i32, [P, i32, ctypes.POINTER(ImageDesc)]),
            "rr_images_grids": (i32, [P, ctypes.POINTER(ctypes.c_void_p),
                                      ctypes.POINTER(ctypes.c_void_p)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    """Names declared in include/rr_hip.h (functions only)."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(rr_\w+)\s*\(", txt, re.M)))


def variants():
    """[(lds_entries, chunk_px)] of the band-fill variants built into the library."""
    L = lib()
    out = []
    for v in range(L.rr_hip_variant_count()):
        k, ch = ctypes.c_int32(), ctypes.c_int32()
        L.rr_hip_variant_info(v, ctypes.byref(k), ctypes.byref(ch))
        out.append((k.value, ch.value))
    return out


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def xpath_edges(xy, flags, matrix=(1.0, 0.0, 0.0, 1.0, 0.0, 0.0), flatness=0.1):
    xy = np.ascontiguousarray(xy, np.float64)
    flags = np.ascontiguousarray(flags, np.uint8)
    m = np.ascontiguousarray(matrix, np.float64)
    n = ctypes.c_int64(0)
    rc = lib().rr_xpath_edges(_p(xy), _p(flags), len(flags), _p(m), flatness, None, 0,
                              ctypes.byref(n))
    if rc:
        raise RrError(rc, "rr_xpath_edges")
    out = np.zeros((max(n.value, 1), 4), np.float64)
    rc = lib().rr_xpath_edges(_p(xy), _p(flags), len(flags), _p(m), flatness, _p(out), n.value,
                              ctypes.byref(n))
    if rc:
        raise RrError(rc, "rr_xpath_edges")
    return out[:n.value].copy()


def cov_hash_term(pix, shape):
    return lib().rr_cov_hash_term(pix, shape)


class Page:
    """One page of a synthetic workload (host arrays, numpy views copied out)."""

    def __init__(self, width, height, edges, edge_off, attrs, host_rows, pts, pt_flags, pt_off):
        self.width, self.height = width, height
        self.edges, self.edge_off, self.attrs, self.host_rows = edges, edge_off, attrs, host_rows
        self.pts, self.pt_flags, self.pt_off = pts, pt_flags, pt_off

    @property
    def n_paths(self):
        return len(self.edge_off) - 1

    @property
    def n_edges(self):
        return len(self.edges)

    def path_points(self, i):
        a, b = self.pt_off[i], self.pt_off[i + 1]
        return self.pts[a:b], self.pt_flags[a:b]

    def path_edges(self, i):
        return self.edges[self.edge_off[i]:self.edge_off[i + 1]]

    def subset(self, idx):
        """A page holding only paths idx (in the given order)."""
        idx = list(idx)
        e = [self.path_edges(i) for i in idx]
        off = np.zeros(len(idx) + 1, np.int64)
        off[1:] = np.cumsum([len(x) for x in e]) if idx else []
        pts = [self.path_points(i) for i in idx]
        poff = np.zeros(len(idx) + 1, np.int64)
        poff[1:] = np.cumsum([len(f) for _, f in pts]) if idx else []
        return Page(self.width, self.height,
                    np.concatenate(e) if idx else np.zeros((0, 4)), off,
                    self.attrs[idx].copy(), self.host_rows[idx].copy(),
                    np.concatenate([p for p, _ in pts]) if idx else np.zeros((0, 2)),
                    np.concatenate([f for _, f in pts]) if idx else np.zeros(0, np.uint8), poff)


def workload(config, seed, max_pages=0):
    """Generate a seeded workload (configs 1, 2, 102, 3); returns a list of Page."""
    L = lib()
    h = ctypes.c_void_p()
    rc = L.rr_workload_create(config, seed, max_pages, ctypes.byref(h))
    if rc:
        raise RrError(rc, f"rr_workload_create({config})")
    try:
        pages = []
        for p in range(L.rr_workload_n_pages(h)):
            info = np.zeros(5, np.int64)
            L.rr_workload_page_info(h, p, _p(info))
            w, hh, n_paths, n_edges, n_pts = (int(v) for v in info)
            ptrs = [ctypes.c_void_p() for _ in range(7)]
            L.rr_workload_page_data(h, p, *[ctypes.byref(q) for q in ptrs])

            def arr(ptr, ctype, count, dtype):
                if count == 0:
                    return np.zeros(0, dtype)
                buf = (ctype * count).from_address(ptr.value)
                return np.frombuffer(buf, dtype=dtype).copy()

            edges = arr(ptrs[0], ctypes.c_double, 4 * n_edges, np.float64).reshape(-1, 4)
            edge_off = arr(ptrs[1], ctypes.c_int64, n_paths + 1, np.int64)
            attrs = arr(ptrs[2], ctypes.c_uint8, 4 * n_paths, np.uint8).view(PATH_ATTR_DTYPE)
            host_rows = arr(ptrs[3], ctypes.c_int32, 2 * n_paths, np.int32).reshape(-1, 2)
            pts = arr(ptrs[4], ctypes.c_double, 2 * n_pts, np.float64).reshape(-1, 2)
            pt_flags = arr(ptrs[5], ctypes.c_uint8, n_pts, np.uint8)
            pt_off = arr(ptrs[6], ctypes.c_int64, n_paths + 1, np.int64)
            pages.append(Page(w, hh, edges, edge_off, attrs, host_rows, pts, pt_flags, pt_off))
        return pages
    finally:
        L.rr_workload_destroy(h)


class Image:
    """One host image of the image workload (see rr_images_create)."""

    def __init__(self, pixels, smask, channels, blend, ctm, page):
        self.pixels, self.smask = pixels, smask
        self.height, self.width = pixels.shape[:2]
        self.channels, self.blend, self.ctm, self.page = channels, blend, ctm, page


class ImageSet:
    """Seeded image workload: images, page geometry and the two grey grids."""

    def __init__(self, images, n_pages, page_w, page_h, grid4, grid4_n, grid3, grid3_n):
        self.images, self.n_pages, self.page_w, self.page_h = images, n_pages, page_w, page_h
        self.grid4, self.grid4_n, self.grid3, self.grid3_n = grid4, grid4_n, grid3, grid3_n


def images(seed, n_images, images_per_page):
    """Generate the seeded image workload with the library's own generator."""
    L = lib()
    h = ctypes.c_void_p()
    rc = L.rr_images_create(seed, n_images, images_per_page, ctypes.byref(h))
    if rc:
        raise RrError(rc, "rr_images_create")
    try:
        info = np.zeros(6, np.int32)
        L.rr_images_info(h, _p(info))
        n, n_pages, pw, ph, g4, g3 = (int(v) for v in info)
        out = []
        for i in range(n):
            d = ImageDesc()
            L.rr_images_get(h, i, ctypes.byref(d))
            npx = d.width * d.height
            px = np.frombuffer((ctypes.c_uint8 * (npx * d.channels)).from_address(d.pixels),
                               np.uint8).copy().reshape(d.height, d.width, d.channels)
            sm = None
            if d.smask:
                sm = np.frombuffer((ctypes.c_uint8 * npx).from_address(d.smask),
                                   np.uint8).copy().reshape(d.height, d.width)
            out.append(Image(px, sm, d.channels, d.blend, tuple(d.ctm), d.page))
        p4, p3 = ctypes.c_void_p(), ctypes.c_void_p()
        L.rr_images_grids(h, ctypes.byref(p4), ctypes.byref(p3))
        grid4 = np.frombuffer((ctypes.c_uint16 * g4 ** 4).from_address(p4.value), np.uint16).copy()
        grid3 = np.frombuffer((ctypes.c_uint16 * g3 ** 3).from_address(p3.value), np.uint16).copy()
        return ImageSet(out, n_pages, pw, ph, grid4, g4, grid3, g3)
    finally:
        L.rr_images_destroy(h)


class Backend:
    """Mirror of GpuBackend (crates/gpu/src/backend/mod.rs:216-503) over the C-ABI."""

    def __init__(self, device=0):
        self.L = lib()
        self.h = ctypes.c_void_p()
        rc = self.L.rr_hip_create(device, ctypes.byref(self.h))
        if rc:
            raise RrError(rc, "rr_hip_create: no usable HIP device")
        self._bufs = []

    def close(self):
        if self.h:
            for b in self._bufs:
                self.L.rr_hip_free(self.h, b)
            self._bufs = []
            self.L.rr_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            raise RrError(rc, f"{what}: {self.L.rr_hip_last_error(self.h).decode()}")

    def alloc(self, nbytes, zeroed=False):
        b = ctypes.c_void_p()
        self._check(self.L.rr_hip_alloc(self.h, nbytes, int(zeroed), ctypes.byref(b)), "alloc")
        self._bufs.append(b.value)
        return b.value

    def free(self, buf):
        self._check(self.L.rr_hip_free(self.h, buf), "free")
        self._bufs.remove(buf)

    def upload(self, dst, arr):
        arr = np.ascontiguousarray(arr)
        self._check(self.L.rr_hip_upload(self.h, dst, _p(arr), arr.nbytes), "upload")

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        buf = self.alloc(max(arr.nbytes, 1))
        if arr.nbytes:
            self.upload(buf, arr)
        return buf

    def buffer_len(self, buf):
        n = ctypes.c_size_t()
        self._check(self.L.rr_hip_buffer_len(self.h, buf, ctypes.byref(n)), "buffer_len")
        return n.value

    def vram_budget(self):
        t, u = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self.L.rr_hip_vram_budget(self.h, ctypes.byref(t), ctypes.byref(u)), "vram_budget")
        return t.value, u.value

    def alloc_host_pinned(self, nbytes):
        """A page-locked host buffer of nbytes, as a numpy uint8 view (free with free_host_pinned)."""
        b = ctypes.c_void_p()
        self._check(self.L.rr_hip_alloc_host_pinned(self.h, nbytes, ctypes.byref(b)), "alloc_host_pinned")
        arr = np.ctypeslib.as_array((ctypes.c_uint8 * nbytes).from_address(b.value))
        return b.value, arr

    def free_host_pinned(self, ptr):
        self._check(self.L.rr_hip_free_host_pinned(self.h, ptr), "free_host_pinned")

    def upload_async(self, dst, src_ptr, nbytes):
        f = ctypes.c_uint64()
        self._check(self.L.rr_hip_upload_async(self.h, dst, src_ptr, nbytes, ctypes.byref(f)), "upload_async")
        return f.value

    def download_async(self, dst_ptr, src, nbytes):
        f = ctypes.c_uint64()
        self._check(self.L.rr_hip_download_async(self.h, dst_ptr, src, nbytes, ctypes.byref(f)), "download_async")
        return f.value

    def submit_transfer(self):
        f = ctypes.c_uint64()
        self._check(self.L.rr_hip_submit_transfer(self.h, ctypes.byref(f)), "submit_transfer")
        return f.value

    def wait_transfer(self, fence):
        self._check(self.L.rr_hip_wait_transfer(self.h, fence), "wait_transfer")

    def copy(self, dst, src, nbytes):
        self._check(self.L.rr_hip_copy(self.h, dst, src, nbytes), "copy")

    def download(self, src, arr):
        self._check(self.L.rr_hip_download(self.h, _p(arr), src, arr.nbytes), "download")
        return arr

    def begin_page(self):
        self._check(self.L.rr_hip_begin_page(self.h), "begin_page")

    def record_fill_paths(self, batch):
        self._check(self.L.rr_hip_record_fill_paths(self.h, ctypes.byref(batch)),
                    "record_fill_paths")

    def record_fill_pages(self, batch):
        self._check(self.L.rr_hip_record_fill_pages(self.h, ctypes.byref(batch)),
                    "record_fill_pages")

    def record_place_images(self, batch):
        self._check(self.L.rr_hip_record_place_images(self.h, ctypes.byref(batch)),
                    "record_place_images")

    def record_icc_clut(self, cmyk, rgb, clut, clut_len, n_pixels):
        self._check(self.L.rr_hip_record_icc_clut(self.h, cmyk, rgb, clut, clut_len, n_pixels),
                    "record_icc_clut")

    def record_composite(self, src, dst, n_pixels):
        self._check(self.L.rr_hip_record_composite(self.h, src, dst, n_pixels), "record_composite")

    def record_rgb_to_gray(self, rgb, rgb_stride, gray, gray_stride, width, height):
        self._check(self.L.rr_hip_record_rgb_to_gray(self.h, rgb, rgb_stride, gray, gray_stride, width, height),
                    "record_rgb_to_gray")

    def record_apply_soft_mask(self, pixels, mask, n_pixels):
        self._check(self.L.rr_hip_record_apply_soft_mask(self.h, pixels, mask, n_pixels),
                    "record_apply_soft_mask")

    def last_batch_stats(self):
        """({span_ms, wide_ms, compose_ms}, {items, intervals, composite_items, path_tiles,
        tiles, reruns, skipped, deferred, built_paths, built_edges, filled_tiles,
        intervals_read}) of the last batched recording."""
        ms = np.zeros(3, np.float64)
        cnt = np.zeros(12, np.int64)
        self._check(self.L.rr_hip_last_batch_stats(self.h, _p(ms), _p(cnt)), "last_batch_stats")
        keys = ("items", "intervals", "composite_items", "path_tiles", "tiles", "reruns", "skipped", "deferred",
                "built_paths", "built_edges", "filled_tiles", "intervals_read")
        return ({"span_ms": float(ms[0]), "wide_ms": float(ms[1]), "compose_ms": float(ms[2])},
                dict(zip(keys, (int(v) for v in cnt))))

    def debug_batch_table(self, table):
        """One internal table of the last batched recording (see rr_hip.h)."""
        dtypes = {0: np.int32, 1: np.int64, 2: np.dtype((np.int32, 4)), 3: np.dtype(
            [("off", "<i8"), ("fx0", "<i2"), ("fx1", "<i2"), ("cx0", "<i2"), ("cx1", "<i2")]),
            4: np.int64, 5: np.dtype((np.int32, 16)), 6: np.dtype((np.uint16, 64)), 7: np.dtype((np.int32, 2))}
        n = ctypes.c_int64(0)
        self._check(self.L.rr_hip_debug_batch_table(self.h, table, None, 0, ctypes.byref(n)), "debug_batch_table")
        out = np.zeros(max(n.value, 1), dtypes[table])
        self._check(self.L.rr_hip_debug_batch_table(self.h, table, _p(out), n.value, ctypes.byref(n)),
                    "debug_batch_table")
        return out[:n.value]

    def set_dispatcher(self, params):
        """Load the dispatcher MLP (MLP_PARAMS float32) or clear it (None)."""
        if params is None:
            self._check(self.L.rr_hip_set_dispatcher(self.h, None, 0), "set_dispatcher")
            return
        params = np.ascontiguousarray(params, np.float32)
        self._check(self.L.rr_hip_set_dispatcher(self.h, _p(params), params.size), "set_dispatcher")

    def last_dispatch(self, n_paths):
        """(features (n, 12) float32, classes (n,) uint8) of the last batched recording."""
        f = np.zeros((n_paths, MLP_FEATURES), np.float32)
        r = np.zeros(n_paths, np.uint8)
        self._check(self.L.rr_hip_last_dispatch(self.h, _p(f), _p(r), n_paths), "last_dispatch")
        return f, r

    def mlp_forward(self, features):
        features = np.ascontiguousarray(features, np.float32).reshape(-1, MLP_FEATURES)
        n = len(features)
        d_f = self.to_device(features)
        d_o = self.alloc(max(4 * MLP_CLASSES * n, 4))
        try:
            self._check(self.L.rr_hip_mlp_forward(self.h, d_f, d_o, n), "mlp_forward")
            return self.download(d_o, np.zeros((n, MLP_CLASSES), np.float32))
        finally:
            self.free(d_f)
            self.free(d_o)

    def xpath_edges(self, xy, flags, pt_off, n_paths, n_pts, edges, edge_cap, edge_off, matrix=None,
                    flatness=0.1, check=True):
        """Device path ingestion (XPath::new + flatten_curve on the GPU) over
        device buffers; returns (return code, total edges)."""
        b = XpathBatch(xy, flags, pt_off, matrix, n_paths, n_pts, flatness, edges, edge_cap, edge_off)
        n = ctypes.c_int64(0)
        rc = self.L.rr_hip_xpath_edges(self.h, ctypes.byref(b), ctypes.byref(n))
        if check:
            self._check(rc, "xpath_edges")
        return rc, n.value

    def last_xpath_stats(self):
        """Kernel ms of the last xpath_edges call: (walk, count, emit)."""
        ms = np.zeros(3, np.float64)
        self._check(self.L.rr_hip_last_xpath_stats(self.h, _p(ms)), "last_xpath_stats")
        return tuple(float(v) for v in ms)

    def train_dispatcher(self, features, targets, steps=4000, batch=64, lr=2e-3, seed=1):
        """Train the MLP on the device; targets (n, 5) log2 microseconds, NaN = invalid.
        Returns (params, mean masked loss of the last 100 steps)."""
        features = np.ascontiguousarray(features, np.float32).reshape(-1, MLP_FEATURES)
        targets = np.ascontiguousarray(targets, np.float32).reshape(-1, MLP_CLASSES)
        d_f, d_t = self.to_device(features), self.to_device(targets)
        params = np.zeros(MLP_PARAMS, np.float32)
        loss = ctypes.c_float(0)
        try:
            self._check(self.L.rr_hip_train_dispatcher(self.h, d_f, d_t, len(features), steps, batch, lr, seed,
                                                        _p(params), ctypes.byref(loss)), "train_dispatcher")
        finally:
            self.free(d_f)
            self.free(d_t)
        return params, loss.value

    def record_scan(self, data, block_sums, n):
        self._check(self.L.rr_hip_record_scan(self.h, data, block_sums, n), "record_scan")

    def record_scan_static(self, data, n):
        """record_scan with the library's own static block-sum array: no
        allocation; at most SCAN_STATIC_MAX elements."""
        self._check(self.L.rr_hip_record_scan_static(self.h, data, n), "record_scan_static")

    def inject_error(self, bits=1):
        self._check(self.L.rr_hip_inject_error(self.h, bits), "inject_error")

    def submit_page(self):
        f = ctypes.c_uint64(0)
        self._check(self.L.rr_hip_submit_page(self.h, ctypes.byref(f)), "submit_page")
        return f.value

    def wait_page(self, fence):
        self._check(self.L.rr_hip_wait_page(self.h, fence), "wait_page")

    def set_trace(self, buf, capacity):
        self._check(self.L.rr_hip_set_trace(self.h, buf, capacity), "set_trace")

    def trace_count(self):
        n = ctypes.c_int64(0)
        self._check(self.L.rr_hip_trace_count(self.h, ctypes.byref(n)), "trace_count")
        return n.value

    def set_profiling(self, on):
        self._check(self.L.rr_hip_set_profiling(self.h, int(on)), "set_profiling")

    def last_page_stats(self):
        """(span_ms, fill_kernel_ms, fill_launches) of the last waited page."""
        span = ctypes.c_double(0)
        fill = ctypes.c_double(0)
        n = ctypes.c_int64(0)
        self._check(self.L.rr_hip_last_page_stats(self.h, ctypes.byref(span), ctypes.byref(fill),
                                                  ctypes.byref(n)), "last_page_stats")
        return span.value, fill.value, n.value


def stride_for(width):
    return (width + 3) & ~3


class DevicePage:
    """A page's edges, offsets and attributes resident on the device."""

    def __init__(self, be, page, attrs=None):
        self.be, self.page = be, page
        self.edges = be.to_device(page.edges.astype(np.float64))
        self.edge_off = be.to_device(page.edge_off.astype(np.int64))
        self.attrs_host = (page.attrs if attrs is None else attrs).copy()
        self.attrs = be.to_device(self.attrs_host)
        self.host_rows = np.ascontiguousarray(page.host_rows, np.int32)
        self.stride = stride_for(page.width)
        self.canvas = be.alloc(self.stride * page.height)
        self.hashes = be.alloc(8 * max(page.n_paths, 1))
        self.counts = be.alloc(16 * max(page.n_paths, 1))

    def free(self):
        for b in (self.edges, self.edge_off, self.attrs, self.canvas, self.hashes, self.counts):
            self.be.free(b)

    def batch(self, mode=MODE_DOCUMENTED, flavour=FLAVOUR_GRAY8, launch=LAUNCH_PLAIN,
              hashes=False, clip=None, counts=False):
        pg = self.page
        b = FillBatch()
        b.edges, b.edge_off, b.attrs = self.edges, self.edge_off, self.attrs
        b.host_rows = self.host_rows.ctypes.data
        b.n_paths, b.n_edges = pg.n_paths, pg.n_edges
        b.canvas = self.canvas
        b.width, b.height, b.stride = pg.width, pg.height, self.stride
        b.mode, b.flavour, b.launch = mode, flavour, launch
        c = clip if clip is not None else (0.0, 0.0, pg.width - 0.001, pg.height - 0.001)
        for i in range(4):
            b.clip[i] = c[i]
        b.cov_hash = self.hashes if hashes else None
        b.cov_count = self.counts if counts else None
        return b

    def reset_canvas(self, value=255):
        self.be._check(self.be.L.rr_hip_memset(self.be.h, self.canvas, value,
                                               self.stride * self.page.height), "memset")
        self.be._check(self.be.L.rr_hip_memset(self.be.h, self.hashes, 0,
                                               8 * max(self.page.n_paths, 1)), "memset")
        self.be._check(self.be.L.rr_hip_memset(self.be.h, self.counts, 0,
                                               16 * max(self.page.n_paths, 1)), "memset")

    def read_canvas(self):
        pg = self.page
        out = np.zeros((pg.height, self.stride), np.uint8)
        self.be.download(self.canvas, out)
        return out[:, :pg.width].copy()

    def read_hashes(self):
        out = np.zeros(max(self.page.n_paths, 1), np.uint64)
        self.be.download(self.hashes, out)
        return out[:self.page.n_paths].copy()

    def read_counts(self):
        out = np.zeros((max(self.page.n_paths, 1), 2), np.uint64)
        self.be.download(self.counts, out)
        return out[:self.page.n_paths].copy()

    def render(self, mode=MODE_DOCUMENTED, flavour=FLAVOUR_GRAY8, launch=LAUNCH_PLAIN,
               hashes=True, clip=None, background=255, counts=False):
        """Fill all paths in order on a fresh canvas; returns (canvas, hashes)."""
        self.reset_canvas(background)
        b = self.batch(mode, flavour, launch, hashes, clip, counts)
        self.be.begin_page()
        self.be.record_fill_paths(b)
        self.be.wait_page(self.be.submit_page())
        return self.read_canvas(), (self.read_hashes() if hashes else None)


class DeviceBatch:
    """Several pages resident on the device: edges, offsets and attributes
    concatenated with global path indices, one canvas per page.

    rgb: 3-byte Rgb8 canvases with per-path colours (colors: n_paths x 3
    uint8); gray: a Gray8 output per page (Rgb8 only); soft_masks: one
    (H, W) uint8 page-space mask per page, or None."""

    def __init__(self, be, pages, attrs=None, rgb=False, colors=None, gray=False, soft_masks=None):
        self.be, self.pages = be, pages
        self.rgb, self.bpp = rgb, (3 if rgb else 1)
        offs = [0]
        e_base = 0
        edge_off = []
        for pg in pages:
            edge_off.append(pg.edge_off[:-1] + e_base)
            e_base += pg.n_edges
        edge_off.append(np.array([e_base], np.int64))
        self.n_paths = sum(pg.n_paths for pg in pages)
        self.n_edges = e_base
        self.path_begin = np.cumsum([0] + [pg.n_paths for pg in pages])
        self.edges = be.to_device(np.concatenate([pg.edges for pg in pages]).astype(np.float64)
                                  if self.n_edges else np.zeros((1, 4)))
        self.edge_off = be.to_device(np.concatenate(edge_off).astype(np.int64))
        a = np.concatenate([pg.attrs for pg in pages]) if attrs is None else attrs
        self.attrs_host = a.copy()
        self.attrs = be.to_device(self.attrs_host)
        self.host_rows = np.ascontiguousarray(np.concatenate([pg.host_rows for pg in pages]), np.int32)
        self.strides = [self.bpp * stride_for(pg.width) for pg in pages]
        self.canvases = [be.alloc(st * pg.height) for st, pg in zip(self.strides, pages)]
        self.hashes = be.alloc(8 * max(self.n_paths, 1))
        self.counts = be.alloc(16 * max(self.n_paths, 1))
        self.extra = []
        self.colors = None
        if rgb:
            c4 = np.zeros((max(self.n_paths, 1), 4), np.uint8)
            if colors is not None:
                c4[:self.n_paths, :3] = colors
            self.colors = be.to_device(c4)
            self.extra.append(self.colors)
        self.grays = [be.alloc(stride_for(pg.width) * pg.height) for pg in pages] if gray else None
        self.descs = (PageDesc * len(pages))()
        for i, pg in enumerate(pages):
            d = self.descs[i]
            d.canvas, d.width, d.height, d.stride = self.canvases[i], pg.width, pg.height, self.strides[i]
            d.path_begin, d.path_end = int(self.path_begin[i]), int(self.path_begin[i + 1])
            for k, v in enumerate((0.0, 0.0, pg.width - 0.001, pg.height - 0.001)):
                d.clip[k] = v
            if soft_masks is not None and soft_masks[i] is not None:
                sm = np.zeros((pg.height, stride_for(pg.width)), np.uint8)
                sm[:, :pg.width] = soft_masks[i]
                d.soft_mask = be.to_device(sm)
                d.soft_mask_stride = sm.shape[1]
                self.extra.append(d.soft_mask)
            if gray:
                d.gray, d.gray_stride = self.grays[i], stride_for(pg.width)

    def free(self):
        for b in [self.edges, self.edge_off, self.attrs, self.hashes, self.counts] + self.canvases + self.extra:
            self.be.free(b)
        for b in self.grays or []:
            self.be.free(b)

    def fill_pages(self, mode=MODE_DOCUMENTED, flavour=FLAVOUR_GRAY8, launch=LAUNCH_BATCHED,
                   hashes=False, counts=False, variant=0):
        b = FillPages()
        b.edges, b.edge_off, b.attrs = self.edges, self.edge_off, self.attrs
        b.host_rows = self.host_rows.ctypes.data
        b.n_paths, b.n_edges = self.n_paths, self.n_edges
        b.pages = ctypes.cast(self.descs, ctypes.c_void_p)
        b.n_pages = len(self.pages)
        b.mode, b.flavour, b.launch = mode, flavour, launch
        b.cov_hash = self.hashes if hashes else None
        b.cov_count = self.counts if counts else None
        b.variant = variant
        b.colors = self.colors
        return b

    def reset(self, background=255):
        L, h = self.be.L, self.be.h
        for c, st, pg in zip(self.canvases, self.strides, self.pages):
            self.be._check(L.rr_hip_memset(h, c, background, st * pg.height), "memset")
        self.be._check(L.rr_hip_memset(h, self.hashes, 0, 8 * max(self.n_paths, 1)), "memset")
        self.be._check(L.rr_hip_memset(h, self.counts, 0, 16 * max(self.n_paths, 1)), "memset")

    def run(self, mode=MODE_DOCUMENTED, flavour=FLAVOUR_GRAY8, launch=LAUNCH_BATCHED,
            hashes=False, counts=False, background=255, variant=0):
        self.reset(background)
        self.be.begin_page()
        self.be.record_fill_pages(self.fill_pages(mode, flavour, launch, hashes, counts, variant))
        self.be.wait_page(self.be.submit_page())

    def set_routes(self, routes):
        """Set every path's route (a scalar RR_ROUTE_* or one per path) and upload."""
        self.attrs_host["route"] = np.asarray(routes, np.uint8)
        self.be.upload(self.attrs, self.attrs_host)

    def capture(self, mode=MODE_DOCUMENTED, flavour=FLAVOUR_GRAY8, hashes=False, counts=False, variant=0):
        """Capture the batched fill of this batch as a HIP graph; returns a handle."""
        b = self.fill_pages(mode, flavour, LAUNCH_BATCHED, hashes, counts, variant)
        g = ctypes.c_void_p()
        self.be._check(self.be.L.rr_hip_graph_create(self.be.h, ctypes.byref(b), ctypes.byref(g)),
                       "graph_create")
        return g.value

    def run_graph(self, graph, background=255):
        self.reset(background)
        self.be.begin_page()
        self.be._check(self.be.L.rr_hip_graph_launch(self.be.h, graph), "graph_launch")
        self.be.wait_page(self.be.submit_page())

    def destroy_graph(self, graph):
        self.be._check(self.be.L.rr_hip_graph_destroy(self.be.h, graph), "graph_destroy")

    def autotune(self, mode=MODE_DOCUMENTED, flavour=FLAVOUR_GRAY8, launch=LAUNCH_BATCHED, reps=2):
        """Time every band-fill variant on this batch; returns (best, [ms per variant])."""
        n = self.be.L.rr_hip_variant_count()
        ms = np.zeros(n, np.float64)
        best = ctypes.c_int32(0)
        b = self.fill_pages(mode, flavour, launch)
        self.be._check(self.be.L.rr_hip_autotune_fill(self.be.h, ctypes.byref(b), reps,
                                                       ctypes.byref(best), _p(ms)), "autotune")
        return best.value, ms.tolist()

    def read_canvas(self, i):
        pg = self.pages[i]
        out = np.zeros((pg.height, self.strides[i]), np.uint8)
        self.be.download(self.canvases[i], out)
        if self.rgb:
            return out[:, :3 * pg.width].reshape(pg.height, pg.width, 3).copy()
        return out[:, :pg.width].copy()

    def upload_canvas(self, i, img):
        """Set page i's canvas to img ((H, W) or, Rgb8, (H, W, 3) uint8)."""
        pg = self.pages[i]
        buf = np.zeros((pg.height, self.strides[i]), np.uint8)
        buf[:, :self.bpp * pg.width] = np.asarray(img, np.uint8).reshape(pg.height, -1)
        self.be.upload(self.canvases[i], buf)

    def read_gray(self, i):
        pg = self.pages[i]
        out = np.zeros((pg.height, stride_for(pg.width)), np.uint8)
        self.be.download(self.grays[i], out)
        return out[:, :pg.width].copy()

    def read_hashes(self):
        out = np.zeros(max(self.n_paths, 1), np.uint64)
        self.be.download(self.hashes, out)
        return out[:self.n_paths].copy()

    def read_counts(self):
        out = np.zeros((max(self.n_paths, 1), 2), np.uint64)
        self.be.download(self.counts, out)
        return out[:self.n_paths].copy()


class DeviceImages:
    """An image set resident on the device: pixels, masks, grids and one grey
    canvas per page."""

    def __init__(self, be, iset, n_pages=None):
        self.be, self.iset = be, iset
        self.n_pages = iset.n_pages if n_pages is None else n_pages
        self.stride = stride_for(iset.page_w)
        self.canvases = [be.alloc(self.stride * iset.page_h) for _ in range(self.n_pages)]
        self.grid4 = be.to_device(iset.grid4)
        self.grid3 = be.to_device(iset.grid3)
        self.bufs = []
        self.descs = (ImageDesc * max(len(iset.images), 1))()
        for i, im in enumerate(iset.images):
            d = self.descs[i]
            d.pixels = be.to_device(im.pixels)
            self.bufs.append(d.pixels)
            if im.smask is not None:
                d.smask = be.to_device(im.smask)
                self.bufs.append(d.smask)
            d.width, d.height, d.channels, d.blend = im.width, im.height, im.channels, im.blend
            for k, v in enumerate(im.ctm):
                d.ctm[k] = v
            d.page = im.page
        self.pages = (PageDesc * max(self.n_pages, 1))()
        for i in range(self.n_pages):
            pd = self.pages[i]
            pd.canvas, pd.width, pd.height, pd.stride = (self.canvases[i], iset.page_w,
                                                         iset.page_h, self.stride)

    def free(self):
        for b in self.bufs + self.canvases + [self.grid4, self.grid3]:
            self.be.free(b)

    def batch(self, first=0, count=None):
        b = ImageBatch()
        n = len(self.iset.images) - first if count is None else count
        b.images = ctypes.addressof(self.descs) + first * ctypes.sizeof(ImageDesc)
        b.n_images = n
        b.pages = ctypes.cast(self.pages, ctypes.c_void_p)
        b.n_pages = self.n_pages
        b.grid4, b.grid4_n = self.grid4, self.iset.grid4_n
        b.grid3, b.grid3_n = self.grid3, self.iset.grid3_n
        return b

    def reset(self, background=255):
        L, h = self.be.L, self.be.h
        for c in self.canvases:
            self.be._check(L.rr_hip_memset(h, c, background, self.stride * self.iset.page_h),
                           "memset")

    def run(self, background=255):
        self.reset(background)
        self.be.begin_page()
        self.be.record_place_images(self.batch())
        self.be.wait_page(self.be.submit_page())

    def read_canvas(self, i):
        out = np.zeros((self.iset.page_h, self.stride), np.uint8)
        self.be.download(self.canvases[i], out)
        return out[:, :self.iset.page_w].copy()
