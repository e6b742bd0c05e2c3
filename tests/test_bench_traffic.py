"""bench.py reports roofline.traffic only from a profile taken on the current
kernel source: the committed profile's traffic record is found and used
while gpusharesat_amd/csrc/kernels.hip hashes to the recorded SHA-256, and
a record of another source or kernel gives null.  No GPU is used."""
import hashlib
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_traffic_record_is_tied_to_the_kernel_source(tmp_path):
    b = _bench()
    src = hashlib.sha256(open(b.KERNEL_SOURCE, "rb").read()).hexdigest()
    rec = ("traffic-record: kernel=k_sweep<16,1,4,true> source_sha256={} fetch_kib=1000 "
           "fetch_factor=2.00 write_kib=500 write_factor=1.00 date=2026-01-01\n")
    p = tmp_path / "prof.md"
    p.write_text("# profile\n" + rec.format(src))
    b.TRAFFIC_PROFILE = str(p)
    out = b.measured_traffic("k_sweep<16,1,4,true>")
    assert out["bytes"] == (1000 * 2 + 500) * 1024
    assert out["bytes"] and out["date"] == "2026-01-01"
    assert b.measured_traffic("k_sweep<16,1,2,true>")["bytes"] is None  # another kernel
    p.write_text(rec.format("0" * 64))  # another kernel source
    out = b.measured_traffic("k_sweep<16,1,4,true>")
    assert out["bytes"] is None and "another" in out["note"]
    p.write_text("no record here\n")
    assert b.measured_traffic("k_sweep<16,1,4,true>")["bytes"] is None


def test_committed_profile_has_a_traffic_record():
    b = _bench()
    out = b.measured_traffic("k_sweep<16,1,4,true>")
    assert "profiled_source_sha256" in out, out
