"""The bench's HBM traffic constants (rocprofv3 FETCH_SIZE / WRITE_SIZE
readings) belong to one state of the kernel sources: the digest recorded
beside them must be the digest of the sources in the tree, so a kernel
change that is not followed by new counter passes fails here instead of
pricing the new kernels with old counters.  The bench itself reports no
traffic figure in that case (checked with a planted digest)."""
import importlib
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    return importlib.import_module("bench")


def test_traffic_constants_match_the_kernel_sources():
    b = _bench()
    assert b.kernel_sources_digest() == b.TRAFFIC_SOURCES_DIGEST
    assert b.traffic_current()
    # per thread slot of the sweep: the counters scaled by the copy kernel's
    # factors, 2 x 114.394 + 55.865 B
    assert abs(b.traffic_bytes("k_sweep3d<float>") - (2 * 114.394 + 55.865)) < 1e-9


def test_stale_digest_gives_no_traffic(monkeypatch):
    b = _bench()
    monkeypatch.setattr(b, "TRAFFIC_SOURCES_DIGEST", "0" * 16)
    assert not b.traffic_current()
    for k in b.TRAFFIC_KIB:
        assert b.traffic_bytes(k) is None
    assert b.TRAFFIC_STALE.startswith("null:")


def test_digest_follows_the_sources(tmp_path, monkeypatch):
    b = _bench()
    for name in ("a.h", "b.hip", "notes.txt"):
        (tmp_path / name).write_text(name)
    monkeypatch.setattr(b, "CSRC", str(tmp_path))
    d0 = b.kernel_sources_digest()
    (tmp_path / "notes.txt").write_text("changed")  # not a source: same digest
    assert b.kernel_sources_digest() == d0
    (tmp_path / "b.hip").write_text("changed")
    assert b.kernel_sources_digest() != d0
