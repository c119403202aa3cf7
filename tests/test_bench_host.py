"""bench.py's host-side contract (CPU only): one GPU only, and roofline.traffic taken from the
committed PMC measurement only while it was measured on the kernel sources being run."""
import json
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and not r.stdout.strip()


def _profile(tmp_path, doc):
    (tmp_path / "profiles").mkdir(exist_ok=True)
    (tmp_path / "profiles" / "k_wide_step_traffic.json").write_text(doc if isinstance(doc, str) else json.dumps(doc))


def test_traffic_only_for_the_measured_kernel_sources(tmp_path, monkeypatch):
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    monkeypatch.setattr(bench, "kernel_source_hash", lambda: "0123456789abcdef")
    rec = {"fetch_kib": 1000.0, "fetch_scale": 2.0, "write_kib": 300.0, "write_scale": 1.0}
    _profile(tmp_path, {"kernel_source_hash": "0123456789abcdef", "tti_800": rec})
    # FETCH_SIZE and WRITE_SIZE in KiB, each scaled by its copy-kernel calibration
    assert bench.measured_traffic("tti", 800) == (1000.0 * 2.0 + 300.0) * 1024
    assert bench.measured_traffic("tti", 512) is None  # no measurement of that workload
    _profile(tmp_path, {"kernel_source_hash": "fedcba9876543210", "tti_800": rec})
    assert bench.measured_traffic("tti", 800) is None  # measured on other kernel sources
    _profile(tmp_path, "{not json")
    assert bench.measured_traffic("tti", 800) is None


def test_kernel_source_hash_follows_the_sources(tmp_path, monkeypatch):
    src = tmp_path / "wavefd_amd" / "csrc"
    src.mkdir(parents=True)
    for name in bench.KERNEL_SOURCES:
        (src / name).write_text("// " + name + "\n")
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    h0 = bench.kernel_source_hash()
    assert len(h0) == 16 and h0 == bench.kernel_source_hash()
    (src / bench.KERNEL_SOURCES[-1]).write_text("// changed\n")
    assert bench.kernel_source_hash() != h0
