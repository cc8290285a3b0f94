"""Host-side logic of bench.py (no GPU): the one-GPU guard, the lookup of the
committed counter bytes and the algorithmic-byte model of the batched passes."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_measured_traffic_lookup():
    # the batched kernels' entries are per variant; the plain kernel's stand
    # for every variant (they differ only in the fallback's LDS slots)
    assert bench.measured_traffic("mixed-64", 0, "k_spans") > 0
    assert bench.measured_traffic("mixed-64", 7, "k_spans") is None
    plain = bench.measured_traffic("glyph-page", 12, "k_fill_plain")
    assert plain and bench.measured_traffic("glyph-page", 3, "k_fill_plain") == plain
    assert bench.measured_traffic("no-such-workload", 0, "k_spans") is None


def test_batched_kernel_bytes_model():
    st = {"items": 10, "skipped": 4, "built_edges": 1000, "intervals": 50, "tiles": 3, "filled_tiles": 1,
          "intervals_read": 70, "path_tiles": 5, "deferred": 0}
    ms = {"span_ms": 1.0, "wide_ms": 0.0, "compose_ms": 2.0}
    k = bench.batched_kernel_bytes(ms, st, 64, 7)
    span = 32 * 1000 + 16 * 10 + 64 * 7 + 8 * 50 + 144 * 6 + 16 * 4
    comp = 16 * 64 * (2 * 3 - 1) + 8 * 70 + 160 * 5
    assert k["k_spans"]["alg_bytes"] == span and k["k_spans_wide"]["alg_bytes"] == 0
    assert k["k_compose"]["alg_bytes"] == comp
    assert abs(k["k_spans"]["achieved_gbs"] - span / 1e-3 / 1e9) < 1e-3
    # deferred bands move their share of the span bytes to the wide pass
    st["deferred"] = 3
    ms["wide_ms"] = 0.5
    k = bench.batched_kernel_bytes(ms, st, 64, 7)
    assert k["k_spans"]["alg_bytes"] + k["k_spans_wide"]["alg_bytes"] in (span - 1, span)
    assert k["k_spans_wide"]["alg_bytes"] == int(span * 0.5)
