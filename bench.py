#!/usr/bin/env python3
"""Benchmark of the MI355X AA fill + composite hot path.

Prints ONE JSON line.  A step renders one batch of the workload (every page:
canvas reset to 255, all paths filled and composited in order) with the edge
lists already resident in HBM.  `value` is the reference's pixel-composite
count (the summed draw_aa_line widths, crates/raster/src/fill/mod.rs:344-401)
divided by the wall time of the timed steps, in Mpx/s.  The same pages are
also run through the plain one-path-per-launch route for the batched/plain
ratio, and through the CPU oracle on a time-bounded sample.
"""
import argparse
import ctypes
import hashlib
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pdf-raster_amd"))
# The plain route spreads independent paths over 20 HIP streams; HIP maps
# streams onto GPU_MAX_HW_QUEUES hardware queues (its default is 4), so the
# bench asks for 24 (--hw-queues) before the runtime starts (it reads the
# variable once, at initialisation): a queue for each of those streams, the
# context stream and HIP's own.
_hwq = [a.split("=", 1)[1] for a in sys.argv if a.startswith("--hw-queues=")]
os.environ["GPU_MAX_HW_QUEUES"] = _hwq[0] if _hwq else "24"

import numpy as np  # noqa: E402

import rr_hip as R  # noqa: E402

METRIC = "AA-filled Mpx/s + pages/s @300dpi, batched+learned vs plain HIP; % HBM peak"
HBM_PEAK_GBS = 8000.0


def batched_kernel_bytes(ms, st, ch, n_paths):
    """Algorithmic bytes and achieved bandwidth of the batched span and
    composite passes from one recording's statistics."""
    # span pass, compulsory bytes: the edge records of every path with a built
    # band read once, the work items and path records read, the intervals
    # written, a 144-byte header per built band and a 16-byte one per band
    # skipped as occluded
    built = st["items"] - st["skipped"]
    a_bytes = (32 * st["built_edges"] + 16 * st["items"] + 64 * n_paths + 8 * st["intervals"]
               + 144 * built + 16 * st["skipped"])
    # composite pass: each touched tile written once and read unless it starts
    # from an opaque cover's colour, the intervals it reads (per tile), a bin
    # entry and a header per composited path-tile
    b_bytes = (16 * ch * (2 * st["tiles"] - st["filled_tiles"]) + 8 * st["intervals_read"]
               + (16 + 144) * st["path_tiles"])
    # the bands the main span pass deferred (rows with more entries than its
    # LDS slots) are built by the wide pass: their share of the span bytes,
    # in proportion to the bands, goes with the wide pass's time
    wide_share = st["deferred"] / built if built > 0 else 0.0
    kernels = {}
    for name, t_ms, nbytes in (("k_spans", ms["span_ms"], a_bytes * (1.0 - wide_share)),
                               ("k_spans_wide", ms["wide_ms"], a_bytes * wide_share),
                               ("k_compose", ms["compose_ms"], b_bytes)):
        gbs = nbytes / (t_ms / 1e3) / 1e9 if t_ms > 0 else 0.0
        kernels[name] = {"ms": round(t_ms, 3), "alg_bytes": int(nbytes), "achieved_gbs": round(gbs, 3),
                         "frac": round(gbs / HBM_PEAK_GBS, 6)}
    return kernels

WORKLOADS = {
    # name: (config id, seed, max pages, primary launch shape, description)
    "mixed-64": (3, 0xC0F10003, 0, R.LAUNCH_BATCHED,
                 "64 pages (40 Letter@300, 12 A3@300, 8 A4@600, 4 thumbnails@150), 200-20000 "
                 "paths per page log-uniform, bbox sides 4-4096 px, 3-2000 edges, 80/20 NZ/EO, "
                 "grey U{0..255}; one persistent batched launch"),
    "glyph-page": (2, 0xC0F10002, 0, R.LAUNCH_PLAIN,
                   "Letter@300 2550x3300 u8, 6000 synthetic glyphs (1-3 contours of 12-40 "
                   "quadratic/cubic segments, flatness 0.1), NZ black, one path per launch"),
    "polygon-page": (1, 0xC0F10001, 0, R.LAUNCH_PLAIN,
                     "Letter@300 2550x3300 u8, 500 random 3-64 vertex polygons, half NZ half EO, "
                     "one path per launch"),
    "images-32": (4, 0xC0F10004, 0, None,
                  "32 images 1024-4096 px a side (CMYK via a 17^4 u16 grid / RGB via a 33^3 u16 "
                  "grid to grey), rotation 0-30 deg, scale 0.5-2, 8-bit smooth soft mask, Normal "
                  "and Multiply alternating, 4 per 2550x3300 grey page; one launch per image"),
}
IMAGES_PER_PAGE = 4
WORKLOADS["dispatch"] = (5, 0xC0F10005, 0, None,
                         "learned dispatch over a glyph-page stream (6000 glyphs per Letter@300 page); "
                         "MLP trained on span-pass traces of 8 pages, evaluated on 32 held-out pages, "
                         "a quarter of them also carrying 4 CMYK/RGB images each; the stream re-renders "
                         "the 32 held-out pages (images included) in recordings of 32 pages")
WORKLOADS["dispatch-mixed"] = (3, 0xC0F10003, 0, None,
                               "learned dispatch on the mixed 64-page batch: MLP trained on span-pass traces "
                               "and CPU times of 8 other mixed pages (seed + 7), evaluated on the 64 pages")
DISPATCH_TRAIN_PAGES, DISPATCH_EVAL_PAGES = 8, 32
DISPATCH_STREAM_PAGES = 10000
WORKLOADS["ingest-glyphs"] = (6, 0xC0F10005, 8, None,
                              "device path ingestion (XPath::new + flatten_curve, flatness 0.1) of 8 glyph "
                              "pages: 48,000 outlines of 1-3 contours of 12-40 quadratic/cubic segments; "
                              "points resident in HBM, edges written to HBM")
GPU_ROUTES = (R.ROUTE_SPAN, R.ROUTE_TILE16, R.ROUTE_TILE32, R.ROUTE_TILE64)


def image_bbox_px(im, page_w, page_h):
    """Canvas pixels the general affine path visits for one image: its clamped
    device-space bounding box (crates/pdf_interp/src/renderer/page/mod.rs:1607-1647).
    The corner products are exact, so plain f64 arithmetic matches the fused form."""
    import math
    a, b, c, d, e, f = im.ctm
    xs = [e, a + e, c + e, a + c + e]
    ys = [page_h - f, page_h - (b + f), page_h - (d + f), page_h - (b + d + f)]
    x0 = min(max(math.floor(min(xs)), 0), page_w)
    x1 = min(max(math.ceil(max(xs)), 0), page_w)
    y0 = min(max(math.floor(min(ys)), 0), page_h)
    y1 = min(max(math.ceil(max(ys)), 0), page_h)
    return max(x1 - x0, 0) * max(y1 - y0, 0)


def measured_traffic(workload, variant, kernel):
    """HBM bytes per launch of `kernel` on this workload and variant from the
    committed rocprofv3 counter runs (profiles/hbm_traffic.json), or None.
    The plain kernel's variants differ only in the LDS slots of the rare
    one-wave fallback, so its entry stands for every variant."""
    p = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        for e in json.load(f)["entries"]:
            if e["workload"] == workload and e["kernel"] == kernel and (
                    e["variant"] == variant or kernel == "k_fill_plain"):
                return e["bytes"]
    return None


def load_golden(name):
    p = os.path.join(ROOT, "tests", "golden", "page_hashes.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get(name)


def host_cpu():
    """CPUs this process may run on, the machine's count, and its model name."""
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return allowed, os.cpu_count() or 1, model


# Worker threads of the CPU baseline: every CPU the process may use, at most
# 16 -- the host CPU share of one GPU on the machines this runs on.
CPU_THREADS_CAP = 16


def cpu_threads():
    allowed, total, model = host_cpu()
    return min(CPU_THREADS_CAP, allowed), {"host_cpus": total, "allowed_cpus": allowed, "cpu_model": model}


def cpu_baseline(pages, w_per_path, budget_s, threads, mode):
    """Time the oracle (the reference CPU raster restated) on the same paths:
    thread t renders page t mod len(pages) in order, 32 paths per call into the
    C oracle, until the budget is spent; Mpx/s counts the reference composites
    of the paths done."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O  # the oracle, timed here as the CPU reference

    base = np.cumsum([0] + [pg.n_paths for pg in pages])
    done_w = [0] * threads
    done_paths = [0] * threads

    def worker(t):
        pi = t % len(pages)
        pg = pages[pi]
        canvas = np.full((pg.height, pg.width), 255, np.uint8)
        a = pg.attrs
        eo = np.ascontiguousarray(a["eo"])
        col = np.ascontiguousarray(a["color"])
        ain = np.ascontiguousarray(a["a_input"])
        wp = np.asarray(w_per_path[base[pi]:base[pi + 1]], np.int64)
        t_end = time.perf_counter() + budget_s
        p = 0
        while time.perf_counter() < t_end:
            if p == pg.n_paths:
                canvas[:] = 255
                p = 0
            q = min(p + 32, pg.n_paths)
            O.lib().rro_fill_page(O._p(canvas), pg.width, pg.height, pg.width, O._p(pg.edges),
                                  O._p(pg.edge_off[p:q + 1]), q - p, O._p(eo[p:q]), O._p(col[p:q]),
                                  O._p(ain[p:q]), mode, 0, None)
            done_w[t] += int(wp[p:q].sum())
            done_paths[t] += q - p
            p = q

    t0 = time.perf_counter()
    ths = [threading.Thread(target=worker, args=(t,)) for t in range(threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    dt = time.perf_counter() - t0
    return sum(done_w), sum(done_paths), dt


def cpu_baseline_images(iset, budget_s, threads):
    """Time the oracle's image placement: thread t places images t, t+threads, ...
    onto its own page canvas, cycling, until the budget is spent."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O  # the oracle, timed here as the CPU reference

    done_px = [0] * threads
    done_img = [0] * threads

    def worker(t):
        canvas = np.full((iset.page_h, iset.page_w), 255, np.uint8)
        mine = list(range(t % len(iset.images), len(iset.images), threads)) or [0]
        t_end = time.perf_counter() + budget_s
        k = 0
        while time.perf_counter() < t_end:
            im = iset.images[mine[k % len(mine)]]
            grid, g = (iset.grid4, iset.grid4_n) if im.channels == 4 else (iset.grid3, iset.grid3_n)
            O.place_image(canvas, im.pixels, im.smask, im.ctm, im.blend, grid, g)
            done_px[t] += image_bbox_px(im, iset.page_w, iset.page_h)
            done_img[t] += 1
            k += 1

    t0 = time.perf_counter()
    ths = [threading.Thread(target=worker, args=(t,)) for t in range(threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    return sum(done_px), sum(done_img), time.perf_counter() - t0


def bench_images(args, desc):
    """Image placement + soft mask + Normal/Multiply through the grey CLUT.  A
    step resets every page canvas to 255 and places all images in order; value
    is canvas pixels visited (the images' clamped bounding boxes) per second."""
    _, seed, _, _, _ = WORKLOADS[args.workload]
    t_gen = time.perf_counter()
    iset = R.images(seed, 32, IMAGES_PER_PAGE)
    print(f"bench: {args.workload}: {len(iset.images)} images on {iset.n_pages} pages generated in "
          f"{time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)
    be = R.Backend(0)
    di = R.DeviceImages(be, iset)
    bbox_px = sum(image_bbox_px(im, iset.page_w, iset.page_h) for im in iset.images)
    img_bytes = sum(min(image_bbox_px(im, iset.page_w, iset.page_h), im.width * im.height) *
                    (im.channels + 1) for im in iset.images)

    di.run()
    digest = hashlib.sha256()
    for i in range(iset.n_pages):
        digest.update(di.read_canvas(i).tobytes())
    fingerprint = digest.hexdigest()
    golden = load_golden(f"{args.workload}")
    parity = "no-fixture" if golden is None else ("match" if golden == fingerprint else "MISMATCH")

    for _ in range(args.warmup):
        di.run()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        di.run()
    sec = (time.perf_counter() - t0) / args.steps
    print(f"bench: images {sec * 1e3:.2f} ms/step", file=sys.stderr, flush=True)

    be.set_profiling(True)
    di.run()
    _, k_ms, launches = be.last_page_stats()
    be.set_profiling(False)
    alg_bytes = 2 * bbox_px + img_bytes
    per_launch = alg_bytes / max(launches, 1)
    avg_s = k_ms / 1e3 / max(launches, 1)
    achieved = per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    out = {
        "metric": METRIC, "value": round(bbox_px / sec / 1e6, 3), "unit": "Mpx/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": args.workload, "description": desc, "images": len(iset.images),
                   "pages": iset.n_pages, "canvas_px_per_step": bbox_px},
        "pages_per_s": round(iset.n_pages / sec, 3), "launches_per_step": launches,
        "parity": parity, "fingerprint": fingerprint[:16],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None,
                     "kernel": "k_place_image", "avg_launch_us": round(avg_s * 1e6, 3),
                     "alg_bytes_per_launch": round(per_launch, 1)},
    }
    di.free()
    be.close()
    if not args.no_cpu_baseline:
        threads, host = cpu_threads()
        px, n_img, cdt = cpu_baseline_images(iset, args.cpu_budget, threads)
        out["cpu_baseline"] = {
            "value": round(px / cdt / 1e6, 3), "unit": "Mpx/s", "cores": threads, "kind": "port",
            "sample": f"{n_img} image placements by the C oracle over {threads} threads "
                      f"(thread t cycles images t, t+{threads}, ...), {cdt:.1f} s", **host}
    print(json.dumps(out), flush=True)


def cpu_baseline_ingest(pages, budget_s, threads):
    """Time the oracle's XPath::new + flatten_curve: thread t walks paths t,
    t+threads, ... of the pages, cycling, until the budget is spent."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O  # the oracle, timed here as the CPU reference

    paths = [(pg, i) for pg in pages for i in range(pg.n_paths)]
    done_edges = [0] * threads
    done_paths = [0] * threads

    def worker(t):
        t_end = time.perf_counter() + budget_s
        k = t
        while time.perf_counter() < t_end:
            pg, i = paths[k % len(paths)]
            xy, fl = pg.path_points(i)
            done_edges[t] += len(O.xpath_edges(xy, fl, flatness=0.1))
            done_paths[t] += 1
            k += threads

    t0 = time.perf_counter()
    ths = [threading.Thread(target=worker, args=(t,)) for t in range(threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    return sum(done_edges), sum(done_paths), time.perf_counter() - t0


def bench_ingest(args, desc):
    """Device path ingestion.  A step turns every path's PathBuilder points
    (resident in HBM) into its add_segment edge list in HBM with one
    rr_hip_xpath_edges call (walk, count, scan, emit, two host syncs); value is
    edges written per second."""
    _, seed, max_pages, _, _ = WORKLOADS[args.workload]
    t_gen = time.perf_counter()
    pages = R.workload(5, seed, max_pages)
    n_paths = sum(pg.n_paths for pg in pages)
    n_pts = sum(len(pg.pt_flags) for pg in pages)
    n_edges = sum(pg.n_edges for pg in pages)
    print(f"bench: {args.workload}: {len(pages)} pages, {n_paths} paths, {n_pts} points, {n_edges} edges "
          f"generated in {time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)
    pts = np.concatenate([pg.pts for pg in pages]).astype(np.float64)
    flags = np.concatenate([pg.pt_flags for pg in pages]).astype(np.uint8)
    off = [np.zeros(1, np.int64)]
    base = 0
    for pg in pages:
        off.append(pg.pt_off[1:] + base)
        base += len(pg.pt_flags)
    pt_off = np.concatenate(off).astype(np.int64)
    want = np.concatenate([pg.edges.reshape(-1, 4) for pg in pages])
    be = R.Backend(0)
    d_xy, d_fl, d_off = be.to_device(pts.reshape(-1)), be.to_device(flags), be.to_device(pt_off)
    d_e, d_eoff = be.alloc(32 * n_edges), be.alloc(8 * (n_paths + 1))

    def step():
        be.xpath_edges(d_xy, d_fl, d_off, n_paths, n_pts, d_e, n_edges, d_eoff)

    step()
    got = be.download(d_e, np.zeros((n_edges, 4), np.float64))
    parity = "match-host" if np.array_equal(got.view(np.uint64), want.view(np.uint64)) else "MISMATCH"
    for _ in range(args.warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    sec = (time.perf_counter() - t0) / args.steps
    print(f"bench: ingest {sec * 1e3:.3f} ms/step", file=sys.stderr, flush=True)
    be.set_profiling(True)
    step()
    walk_ms, count_ms, emit_ms = be.last_xpath_stats()
    be.set_profiling(False)
    # emit: reads 4 points (16 B) + a flag per point, its role and offset, and
    # writes 32 B per edge; the dominant kernel
    emit_bytes = n_pts * (16 + 1 + 1 + 8 + 8) + 32 * n_edges
    achieved = emit_bytes / (emit_ms / 1e3) / 1e9 if emit_ms > 0 else 0.0
    out = {
        "metric": METRIC, "value": round(n_edges / sec / 1e6, 3), "unit": "Medges/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "description": desc, "pages": len(pages), "paths": n_paths,
                   "points": n_pts, "edges": n_edges},
        "paths_per_s": round(n_paths / sec, 1), "kernel_ms": {"k_xpath_walk": round(walk_ms, 4),
                                                               "k_xpath_count": round(count_ms, 4),
                                                               "k_xpath_emit": round(emit_ms, 4)},
        "parity": parity,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None, "kernel": "k_xpath_emit",
                     "avg_launch_us": round(emit_ms * 1e3, 3), "alg_bytes_per_launch": emit_bytes},
    }
    for b in (d_xy, d_fl, d_off, d_e, d_eoff):
        be.free(b)
    be.close()
    if not args.no_cpu_baseline:
        threads, host = cpu_threads()
        ne, npth, cdt = cpu_baseline_ingest(pages, args.cpu_budget, threads)
        out["cpu_baseline"] = {
            "value": round(ne / cdt / 1e6, 3), "unit": "Medges/s", "cores": threads, "kind": "port",
            "sample": f"{npth} paths ({ne} edges) through the C oracle's XPath::new over {threads} threads, "
                      f"{cdt:.1f} s", **host}
    print(json.dumps(out), flush=True)


def route_valid(bands):
    """(n, 5) validity of each class for paths with `bands` 16-row bands."""
    b = np.asarray(bands)
    return np.stack([b > 0, (b > 0) & (b <= 1), (b > 0) & (b <= 2), (b > 0) & (b <= 4), b > 0], 1)


def traced_costs(be, db, route, variant, reps=2):
    """Span-pass cost of every path with all paths on `route`: the wave time of
    its work items scaled by (kernel time / total wave time), i.e. its share of
    the span kernel's wall time, in microseconds (min over reps)."""
    n = db.n_paths
    db.set_routes(R.ROUTE_SPAN)
    db.run(launch=R.LAUNCH_BATCHED, variant=variant)
    _, st = be.last_batch_stats()
    cap = max(8 * n, int(st["items"])) + 4096  # at most one record per span item
    buf = be.alloc(cap * 32)
    best = np.full(n, np.inf)
    db.set_routes(route)
    try:
        for _ in range(reps):
            be.set_trace(buf, cap)
            be.set_profiling(True)
            db.run(launch=R.LAUNCH_BATCHED, variant=variant)
            ms, _ = be.last_batch_stats()
            be.set_profiling(False)
            k = min(be.trace_count(), cap)
            tr = np.zeros((k, 4), np.int64)
            be.download(buf, tr)
            wave = np.bincount(tr[:, 0], weights=(tr[:, 3] - tr[:, 2]) / 100.0, minlength=n)
            scale = ms["span_ms"] * 1e3 / max(wave.sum(), 1e-9)
            best = np.minimum(best, wave * scale)
    finally:
        be.set_trace(None, 0)
        be.free(buf)
    return best


def cpu_path_costs(pages, threads, mode=R.MODE_DOCUMENTED):
    """Reference CPU raster (the C oracle) time of every path, one page per
    thread, each path timed inside the C loop (rro_fill_page_timed), divided
    by the threads used: each path's share of the CPU pool, in microseconds."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O  # the oracle, timed here as the CPU reference

    out = [None] * len(pages)

    def worker(t):
        for pi in range(t, len(pages), threads):
            pg = pages[pi]
            canvas = np.full((pg.height, pg.width), 255, np.uint8)
            a = pg.attrs
            eo = np.ascontiguousarray(a["eo"])
            col = np.ascontiguousarray(a["color"])
            ain = np.ascontiguousarray(a["a_input"])
            ns = np.zeros(max(pg.n_paths, 1), np.int64)
            O.lib().rro_fill_page_timed(O._p(canvas), pg.width, pg.height, pg.width, O._p(pg.edges),
                                        O._p(pg.edge_off), pg.n_paths, O._p(eo), O._p(col), O._p(ain), mode,
                                        O._p(ns))
            out[pi] = ns[:pg.n_paths] / 1e3

    t0 = time.perf_counter()
    ths = [threading.Thread(target=worker, args=(t,)) for t in range(threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    wall = time.perf_counter() - t0
    return np.concatenate(out) / threads, wall


def bench_dispatch(args, desc):
    """Learned dispatcher: per-path classes from an FP32 MLP trained on the
    bench's own span-pass timing traces and per-path CPU times, against the
    static area rule.  dispatch: glyph pages (+ images) and a 10,000-page
    stream; dispatch-mixed: the mixed 64-page batch."""
    cfg, seed, _, _, _ = WORKLOADS[args.workload]
    mixed = cfg == 3
    t_gen = time.perf_counter()
    if mixed:
        train_pages = R.workload(3, seed + 7, DISPATCH_TRAIN_PAGES)
        eval_pages = R.workload(3, seed, args.pages if args.pages else 0)
        img_pages = []
    else:
        train_pages = R.workload(5, seed, DISPATCH_TRAIN_PAGES)
        eval_pages = R.workload(5, seed + 1, DISPATCH_EVAL_PAGES)
        img_pages = list(range(0, DISPATCH_EVAL_PAGES, 4))  # a quarter of the pages
    iset = R.images(0xC0F10004, IMAGES_PER_PAGE * len(img_pages), IMAGES_PER_PAGE) if img_pages else None
    print(f"bench: {args.workload}: {len(train_pages)} train + {len(eval_pages)} eval pages, "
          f"{len(iset.images) if iset else 0} images, generated in {time.perf_counter() - t_gen:.1f} s",
          file=sys.stderr, flush=True)
    be = R.Backend(0)
    dtr, dev = R.DeviceBatch(be, train_pages), R.DeviceBatch(be, eval_pages)
    di, ib = None, None
    if iset:
        di = R.DeviceImages(be, iset, n_pages=0)
        # images of eval page k*4 go onto that page's canvas
        img_descs = (R.ImageDesc * len(iset.images))()
        for i, im in enumerate(iset.images):
            img_descs[i] = di.descs[i]
            img_descs[i].page = img_pages[im.page]
        ib = R.ImageBatch()
        ib.images = ctypes.addressof(img_descs)
        ib.n_images = len(iset.images)
        ib.pages = ctypes.cast(dev.descs, ctypes.c_void_p)
        ib.n_pages = len(eval_pages)
        ib.grid4, ib.grid4_n, ib.grid3, ib.grid3_n = di.grid4, iset.grid4_n, di.grid3, iset.grid3_n

    dev.set_routes(R.ROUTE_SPAN)
    variant = args.variant if args.variant >= 0 else dev.autotune(reps=1)[0]

    def features(db):
        db.set_routes(R.ROUTE_SPAN)
        db.run(launch=R.LAUNCH_BATCHED, variant=variant)
        return be.last_dispatch(db.n_paths)[0]

    def cost_table(db):
        cols = [traced_costs(be, db, r, variant, reps=3) for r in GPU_ROUTES]
        return np.stack(cols + [np.full(db.n_paths, np.nan)], 1)

    f_tr, f_ev = features(dtr), features(dev)
    c_tr, c_ev = cost_table(dtr), cost_table(dev)
    print(f"bench: span-pass traces done", file=sys.stderr, flush=True)
    cpu_note = "cpu class not measured (--no-cpu-baseline)"
    cpu_wall = None
    threads, host = cpu_threads()
    if not args.no_cpu_baseline:
        c_tr[:, 4], _ = cpu_path_costs(train_pages, threads)
        c_ev[:, 4], cpu_wall = cpu_path_costs(eval_pages, threads)
        cpu_note = (f"cpu class = reference CPU raster time per path (C oracle, timed per path inside the C loop) "
                    f"/ {threads} threads")
        print(f"bench: CPU path times done ({cpu_wall:.1f} s for the eval pages)", file=sys.stderr, flush=True)
    v_tr, v_ev = route_valid(f_tr[:, 5]), route_valid(f_ev[:, 5])
    c_tr[~v_tr] = np.nan
    c_ev[~v_ev] = np.nan
    keep = v_tr[:, 0]
    targets = np.log2(np.maximum(c_tr[keep], 1e-3)).astype(np.float32)
    t0 = time.perf_counter()
    params, loss = be.train_dispatcher(f_tr[keep], targets, steps=6000, batch=64, lr=2e-3, seed=seed)
    t_train = time.perf_counter() - t0
    print(f"bench: dispatcher trained on {int(keep.sum())} paths in {t_train:.2f} s, loss {loss:.4f}",
          file=sys.stderr, flush=True)

    def step(routes_auto, mlp):
        be.set_dispatcher(params if mlp else None)
        dev.set_routes(R.ROUTE_AUTO if routes_auto else R.ROUTE_SPAN)
        dev.reset(255)
        be.begin_page()
        be.record_fill_pages(dev.fill_pages(launch=R.LAUNCH_BATCHED, variant=variant))
        if ib is not None:
            be.record_place_images(ib)
        be.wait_page(be.submit_page())

    results = {}
    fingerprints = {}
    for name, auto, mlp in (("span", False, False), ("static", True, False), ("learned", True, True)):
        step(auto, mlp)
        _, cls = be.last_dispatch(dev.n_paths)
        h = hashlib.sha256()
        for i in range(len(eval_pages)):
            h.update(dev.read_canvas(i).tobytes())
        fingerprints[name] = h.hexdigest()
        for _ in range(args.warmup):
            step(auto, mlp)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(auto, mlp)
        sec = (time.perf_counter() - t0) / args.steps
        cost = np.where(v_ev[:, 0], c_ev[np.arange(len(cls)), cls], 0.0)
        results[name] = {"ms_per_step": round(sec * 1e3, 3), "pages_per_s": round(len(eval_pages) / sec, 3),
                         "classes": np.bincount(cls, minlength=5).tolist(), "cost_us": float(np.nansum(cost))}
    best = np.where(v_ev[:, 0], np.nanmin(np.where(np.isnan(c_ev), np.inf, c_ev), 1), 0.0)
    best_total = float(best.sum())
    for name in results:
        results[name]["regret_us"] = round(results[name]["cost_us"] - best_total, 3)
        results[name]["cost_us"] = round(results[name]["cost_us"], 3)
    parity = "match" if len(set(fingerprints.values())) == 1 else "MISMATCH"
    golden = load_golden("mixed-64:documented:canvas") if mixed and not args.pages else None
    if golden is not None and fingerprints["learned"] != golden:
        parity = "MISMATCH"
    be.set_dispatcher(params)
    dev.set_routes(R.ROUTE_AUTO)
    dev.reset(255)
    be.begin_page()
    be.record_fill_pages(dev.fill_pages(launch=R.LAUNCH_BATCHED, counts=True, variant=variant))
    be.wait_page(be.submit_page())
    w_ref = int(dev.read_counts()[:, 1].sum())
    learned_sec = results["learned"]["ms_per_step"] / 1e3
    # the stream: the held-out pages re-rendered, recording after recording,
    # with the learned dispatch (images included), timed as a whole
    stream = {}
    n_stream = args.stream_pages if args.stream_pages >= 0 else (0 if mixed else DISPATCH_STREAM_PAGES)
    if n_stream:
        batches = -(-n_stream // len(eval_pages))
        step(True, True)
        t0 = time.perf_counter()
        for _ in range(batches):
            step(True, True)
        st_sec = time.perf_counter() - t0
        stream = {"pages": batches * len(eval_pages), "recordings": batches, "seconds": round(st_sec, 3),
                  "pages_per_s": round(batches * len(eval_pages) / st_sec, 3),
                  "note": f"{batches} recordings of the {len(eval_pages)} held-out pages (the stream is drawn "
                          f"from them), learned dispatch in-kernel, images on every fourth page"}
        print(f"bench: stream of {stream['pages']} pages in {st_sec:.2f} s", file=sys.stderr, flush=True)
    be.set_profiling(True)
    step(True, True)
    ms, st = be.last_batch_stats()
    be.set_profiling(False)
    ch = R.variants()[variant][1]
    kernels = batched_kernel_bytes(ms, st, ch, sum(p.n_paths for p in eval_pages))
    kern = max(("k_spans", "k_compose"), key=lambda k: kernels[k]["ms"])
    plain = {}
    if mixed and not args.no_plain:
        dev.set_routes(R.ROUTE_SPAN)
        dev.reset(255)
        t0 = time.perf_counter()
        be.begin_page()
        be.record_fill_pages(dev.fill_pages(launch=R.LAUNCH_PLAIN, variant=variant))
        be.wait_page(be.submit_page())
        p_sec = time.perf_counter() - t0
        plain = {"plain_ms_per_step": round(p_sec * 1e3, 3), "plain_pages_per_s": round(len(eval_pages) / p_sec, 3),
                 "learned_over_plain": round(p_sec / learned_sec, 3)}
    out = {
        "metric": METRIC, "value": round(w_ref / learned_sec / 1e6, 3), "unit": "Mpx/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": results["learned"]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8+f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "description": desc, "train_pages": len(train_pages),
                   "eval_pages": len(eval_pages), "image_pages": len(img_pages),
                   "images": len(iset.images) if iset else 0,
                   "eval_paths": dev.n_paths, "variant": {"id": variant, "lds_entries": R.variants()[variant][0],
                                                          "chunk_px": ch}},
        "pages_per_s": results["learned"]["pages_per_s"],
        "dispatch": {"learned": results["learned"], "static": results["static"], "span_only": results["span"],
                     "best_cost_us": round(best_total, 3),
                     "cost_definition": "span-pass share of kernel wall time per path (traced wave time x "
                                        "kernel time / total wave time, min of 3 runs); " + cpu_note,
                     "classes_order": ["span", "tile16", "tile32", "tile64", "cpu"],
                     "train_loss": round(loss, 5), "train_s": round(t_train, 3)},
        **({"stream": stream} if stream else {}), **plain,
        "parity": parity, "fingerprint": fingerprints["learned"][:16],
        "roofline": {"bound": "hbm", "achieved": kernels[kern]["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": kernels[kern]["frac"], "traffic": None, "kernel": kern,
                     "avg_launch_us": round(kernels[kern]["ms"] * 1e3, 3),
                     "alg_bytes_per_launch": kernels[kern]["alg_bytes"]},
        "kernels": kernels,
    }
    if cpu_wall is not None:
        n_ev = sum(pg.n_paths for pg in eval_pages)
        out["cpu_baseline"] = {"value": round(len(eval_pages) / cpu_wall, 3), "unit": "pages/s", "cores": threads,
                               "kind": "port",
                               "sample": f"all {n_ev} paths of the {len(eval_pages)} eval pages filled by the C "
                                         f"oracle, one page per thread, {cpu_wall:.1f} s (no images)", **host}
    be.set_dispatcher(None)
    dtr.free()
    dev.free()
    if di is not None:
        di.free()
    be.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="mixed-64", choices=sorted(WORKLOADS))
    ap.add_argument("--pages", type=int, default=0, help="limit the page count (0 = all)")
    ap.add_argument("--path-stride", type=int, default=1,
                    help="keep every n-th path of each page (a smaller run for counter passes)")
    ap.add_argument("--mode", default="documented", choices=["documented", "literal"])
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU oracle work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-path-per-launch comparison")
    ap.add_argument("--variant", type=int, default=-1, help="band-fill variant (-1 = autotune)")
    ap.add_argument("--hw-queues", type=int, default=24,
                    help="GPU_MAX_HW_QUEUES for this process (pass as --hw-queues=N)")
    ap.add_argument("--stream-pages", type=int, default=-1,
                    help="dispatch workload: pages streamed through the learned dispatch (-1 = 10000 for the "
                         "glyph stream, 0 for the mixed batch)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="end with a 1 GiB device-to-device copy (FETCH/WRITE counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this library runs on one GPU only; --gpus must be 1 "
              "(multi-GPU sharding is a design note, not a measured configuration)", file=sys.stderr)
        sys.exit(2)

    cfg, seed, max_pages, launch, desc = WORKLOADS[args.workload]
    if cfg == 4:
        return bench_images(args, desc)
    if cfg == 5 or args.workload == "dispatch-mixed":
        return bench_dispatch(args, desc)
    if cfg == 6:
        return bench_ingest(args, desc)
    if args.pages:
        max_pages = args.pages
    mode = R.MODE_LITERAL if args.mode == "literal" else R.MODE_DOCUMENTED
    t_gen = time.perf_counter()
    pages = R.workload(cfg, seed, max_pages)
    if args.path_stride > 1:
        pages = [pg.subset(range(0, pg.n_paths, args.path_stride)) for pg in pages]
    t_gen = time.perf_counter() - t_gen
    print(f"bench: {args.workload}: {len(pages)} pages generated in {t_gen:.1f} s", file=sys.stderr,
          flush=True)
    be = R.Backend(0)
    db = R.DeviceBatch(be, pages)
    n_paths, n_edges = db.n_paths, db.n_edges

    # autotuner: time every band-fill variant on this batch, keep the fastest
    if args.variant >= 0:
        variant, tune_ms = args.variant, []
    else:
        variant, tune_ms = db.autotune(mode, R.FLAVOUR_GRAY8, launch, reps=2)
    print(f"bench: variant {variant} {R.variants()[variant]} autotune ms={tune_ms}", file=sys.stderr,
          flush=True)

    def step(shape, hashes=False, counts=False, v=None):
        db.run(mode, R.FLAVOUR_GRAY8, shape, hashes=hashes, counts=counts, variant=variant if v is None else v)

    # work counts + parity fingerprint (outside the timed region)
    step(launch, hashes=True, counts=True)
    cnt = db.read_counts()
    w_cov = int(cnt[:, 0].sum())
    w_ref = int(cnt[:, 1].sum())
    digest = hashlib.sha256()
    for i in range(len(pages)):
        digest.update(db.read_canvas(i).tobytes())
    digest.update(db.read_hashes().astype("<u8").tobytes())
    fingerprint = digest.hexdigest()
    golden = load_golden(f"{args.workload}:{args.mode}")
    parity_cov = "no-fixture" if golden is None else ("match" if golden == fingerprint else "MISMATCH")

    def canvas_parity():
        """The canvases as the last run left them against the oracle's."""
        d = hashlib.sha256()
        for i in range(len(pages)):
            d.update(db.read_canvas(i).tobytes())
        g = load_golden(f"{args.workload}:{args.mode}:canvas")
        return d.hexdigest(), ("no-fixture" if g is None else ("match" if g == d.hexdigest() else "MISMATCH"))

    def timed(shape, steps, warmup, v=None):
        for _ in range(warmup):
            step(shape, v=v)
        t0 = time.perf_counter()
        for _ in range(steps):
            step(shape, v=v)
        return (time.perf_counter() - t0) / steps

    sec = timed(launch, args.steps, args.warmup)
    print(f"bench: primary {sec * 1e3:.2f} ms/step", file=sys.stderr, flush=True)
    timed_fp, parity_timed = canvas_parity()  # the eager timed configuration
    graph_sec = None
    eager_sec = sec
    if launch == R.LAUNCH_BATCHED:
        # the same batch captured once as a HIP graph, then replayed
        g = db.capture(mode, R.FLAVOUR_GRAY8, variant=variant)
        for _ in range(args.warmup):
            db.run_graph(g)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            db.run_graph(g)
        graph_sec = (time.perf_counter() - t0) / args.steps
        db.destroy_graph(g)
        fp_g, par_g = canvas_parity()  # the replayed graph's output
        if par_g != "match" or parity_timed != "match":
            parity_timed = par_g if par_g != "match" else parity_timed
        print(f"bench: graph replay {graph_sec * 1e3:.2f} ms/step", file=sys.stderr, flush=True)
        if graph_sec < sec:
            sec = graph_sec
    nocull = None
    if launch == R.LAUNCH_BATCHED:
        # the same batch with every path at a_input 254: no path is an opaque
        # cover, so nothing is culled -- every (path, band) item is built and
        # every path composited on every tile it touches, with the alpha
        # composite (the rate on pages without opaque covers)
        a_saved = db.attrs_host["a_input"].copy()
        db.attrs_host["a_input"] = 254
        be.upload(db.attrs, db.attrs_host)
        step(launch)
        t0 = time.perf_counter()
        for _ in range(3):
            step(launch)
        nc_sec = (time.perf_counter() - t0) / 3
        db.attrs_host["a_input"] = a_saved
        be.upload(db.attrs, db.attrs_host)
        nocull = {"ms_per_step": round(nc_sec * 1e3, 3), "pages_per_s": round(len(pages) / nc_sec, 3),
                  "ref_mpx_per_s": round(w_ref / nc_sec / 1e6, 3),
                  "note": "every path at a_input 254 (no opaque covers, so no occlusion culling), eager, 3 steps "
                          "after 1 warmup"}
        print(f"bench: without opaque covers {nc_sec * 1e3:.2f} ms/step", file=sys.stderr, flush=True)
    value = w_ref / sec / 1e6
    pages_per_s = len(pages) / sec
    # the whole step against the compulsory bytes: every canvas pixel read and
    # written once, every edge's four f64 endpoints read once (one recording
    # of the batch; "effective" counts the pixels with coverage instead)
    canvas_px = sum(pg.width * pg.height for pg in pages)
    comp = 2 * canvas_px + 32 * n_edges
    eff = 2 * w_cov + 32 * n_edges
    pipeline = {"compulsory_bytes": comp, "achieved_gbs": round(comp / sec / 1e9, 3),
                "frac": round(comp / sec / 1e9 / HBM_PEAK_GBS, 6), "effective_bytes": eff,
                "effective_gbs": round(eff / sec / 1e9, 3),
                "basis": "2 B per canvas pixel + 32 B per edge, over ms_per_step"}

    # roofline: every fill launch bracketed by its own HIP event pair
    be.set_profiling(True)
    db.reset(255)
    be.begin_page()
    be.record_fill_pages(db.fill_pages(mode, R.FLAVOUR_GRAY8, launch, variant=variant))
    be.wait_page(be.submit_page())
    _, fill_ms, launches = be.last_page_stats()
    kernels = {}
    if launch == R.LAUNCH_BATCHED:
        ms, st = be.last_batch_stats()
        kernels = batched_kernel_bytes(ms, st, R.variants()[variant][1], n_paths)
        kernels["counts"] = st
        kernel = max(("k_spans", "k_compose"), key=lambda k: kernels[k]["ms"])
        per_launch_bytes = kernels[kernel]["alg_bytes"]
        avg_launch_s = kernels[kernel]["ms"] / 1e3
    else:
        alg_bytes = 2 * w_cov + 32 * n_edges + 8 * n_paths
        kernel = "k_fill_plain"
        per_launch_bytes = alg_bytes / max(launches, 1)
        avg_launch_s = fill_ms / 1e3 / max(launches, 1)
    be.set_profiling(False)
    achieved = per_launch_bytes / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpx/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": args.workload, "description": desc, "pages": len(pages),
                   **({"path_stride": args.path_stride} if args.path_stride > 1 else {}),
                   "paths": n_paths, "edges": n_edges, "aa_row_mode": args.mode,
                   "canvas": "gray8-exact",
                   "launch": "batched" if launch == R.LAUNCH_BATCHED else "plain",
                   "variant": {"id": variant, "lds_entries": R.variants()[variant][0],
                               "chunk_px": R.variants()[variant][1]}},
        "autotune_ms": [round(m, 3) for m in tune_ms],
        "pages_per_s": round(pages_per_s, 3),
        "eager_ms_per_step": round(eager_sec * 1e3, 3),
        "graph_ms_per_step": None if graph_sec is None else round(graph_sec * 1e3, 3),
        "covered_px_per_step": w_cov, "ref_composites_per_step": w_ref,
        "fill_launches_per_step": launches,
        "value_basis": "reference-equivalent pixel composites: the summed draw_aa_line widths "
                       "(crates/raster/src/fill/mod.rs:344-401) of every path, as the CPU raster "
                       "performs them; pixels hidden under a later opaque cover are not recomputed "
                       "on the GPU (see gpu_work)",
        "parity": "match" if parity_cov == parity_timed == "match" else
                  ("no-fixture" if "no-fixture" in (parity_cov, parity_timed) else "MISMATCH"),
        "parity_detail": {"coverage_run": parity_cov, "timed_run_canvases": parity_timed},
        "fingerprint": fingerprint[:16], "timed_fingerprint": timed_fp[:16],
        "pipeline": pipeline,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6),
                     "traffic": measured_traffic(args.workload, variant, kernel),
                     "kernel": kernel, "avg_launch_us": round(avg_launch_s * 1e6, 3),
                     "alg_bytes_per_launch": round(per_launch_bytes, 1)},
    }
    if kernels:
        out["kernels"] = kernels
        st = kernels["counts"]
        ch = R.variants()[variant][1]
        out["gpu_work"] = {
            "path_tile_composites": st["path_tiles"],
            "composited_px_upper_bound": st["path_tiles"] * 16 * ch,
            "span_items": st["items"], "span_items_occluded": st["skipped"],
            "tiles_started_from_cover": st["filled_tiles"]}
    if nocull:
        out["without_occlusion_culling"] = nocull
    if not args.no_plain:
        other = R.LAUNCH_PLAIN if launch == R.LAUNCH_BATCHED else R.LAUNCH_BATCHED
        key = "plain" if other == R.LAUNCH_PLAIN else "batched"
        if key == "batched":
            # the batched launch with its own best variant (the plain route's
            # variants differ only in the fallback's LDS slots)
            v_o = variant if args.variant >= 0 else db.autotune(mode, R.FLAVOUR_GRAY8, other, reps=1)[0]
            sec_o = timed(other, 3, 1, v=v_o)
            out["batched_variant"] = v_o
        else:
            sec_o = timed(other, 1, 0)
        out[f"{key}_ms_per_step"] = round(sec_o * 1e3, 3)
        out[f"{key}_pages_per_s"] = round(len(pages) / sec_o, 3)
        out[f"{key}_mpx_per_s"] = round(w_ref / sec_o / 1e6, 3)
        out["batched_over_plain"] = round((sec_o / sec) if key == "plain" else (sec / sec_o), 3)
    db.free()
    if args.calibrate_copy:
        n = 1 << 30
        a, b = be.alloc(n), be.alloc(n)
        be.copy(b, a, n)
        be.free(a)
        be.free(b)
    be.close()
    if not args.no_cpu_baseline:
        threads, host = cpu_threads()
        w_per_path = cnt[:, 1]
        w_done, p_done, cdt = cpu_baseline(pages, w_per_path, args.cpu_budget, threads, mode)
        out["cpu_baseline"] = {
            "value": round(w_done / cdt / 1e6, 3), "unit": "Mpx/s", "cores": threads,
            "kind": "port",
            "sample": f"{p_done} paths of the first {min(threads, len(pages))} page(s) rendered "
                      f"in order by the C oracle, one page per thread over {threads} threads, "
                      f"{cdt:.1f} s",
            "pages_per_s": round(w_done / cdt / (w_ref / len(pages)), 3), **host}
        # the same on one thread (a quarter of the budget)
        w1, p1, dt1 = cpu_baseline(pages, w_per_path, args.cpu_budget / 4, 1, mode)
        out["cpu_baseline"]["single_thread"] = {"value": round(w1 / dt1 / 1e6, 3), "unit": "Mpx/s",
                                                "sample": f"{p1} paths of page 0, {dt1:.1f} s"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
