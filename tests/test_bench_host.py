"""bench.py's use of the committed rocprofv3 counter summaries: they apply only to the batch size and
the kernel sources they were measured on."""
import importlib.util
import json
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_summaries_record_shape_and_sources():
    for name in ("hbm_traffic.json", "valu_counters.json"):
        with open(os.path.join(ROOT, "profiles", name)) as f:
            data = json.load(f)
        assert data["count"] == 128
        assert len(data["source_sha256"]) == 64
        assert tuple(data["sources"]) == _bench().PROFILED_SOURCES


def test_summaries_rejected_on_other_count_or_sources(monkeypatch):
    b = _bench()
    with open(os.path.join(ROOT, "profiles", "valu_counters.json")) as f:
        recorded = json.load(f)["source_sha256"]
    monkeypatch.setattr(b, "profiled_sources_sha256", lambda: recorded)
    assert b.traffic_from_profiles("ntt_mul_dim3", 128) is not None
    assert b.valu_issue_from_profiles("ntt_mul_dim3", 0.56, 128)["frac"] > 0
    # another batch size: the per-launch counts do not apply
    assert b.traffic_from_profiles("ntt_mul_dim3", 64) is None
    assert b.valu_issue_from_profiles("ntt_mul_dim3", 0.56, 256) is None
    # changed kernel sources: neither applies
    monkeypatch.setattr(b, "profiled_sources_sha256", lambda: "0" * 64)
    assert b.traffic_from_profiles("ntt_mul_dim3", 128) is None
    assert b.valu_issue_from_profiles("ntt_mul_dim3", 0.56, 128) is None
