"""Host-side checks of bench.py (CPU only): the one-GPU guard, the lookup of
measured HBM traffic by workload / kernel / algorithmic bytes / stripe count,
and the consistency of the committed bench lines with the traffic records
and with the output contract (every key present, roofline arithmetic)."""
import json
import os
import subprocess
import sys

import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH_DIR = os.path.join(ROOT, "profiles", "bench")
CONTRACT_KEYS = {"metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
                 "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
                 "roofline", "cpu_baseline"}
ROOFLINE_KEYS = {"bound", "achieved", "peak", "unit", "frac", "traffic"}
CPU_KEYS = {"value", "unit", "cores", "kind", "sample"}


def load_line(name):
    with open(os.path.join(BENCH_DIR, "bench_%s.json" % name)) as f:
        return json.load(f)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_measured_traffic_lookup(tmp_path, monkeypatch):
    recs = {"workloads": {
        "w": [{"kernel": "k_persistent", "algorithmic_bytes": 1000, "stripes": 1,
               "traffic_bytes": 900.0},
              {"kernel": "k_persistent", "algorithmic_bytes": 1000, "stripes": 64,
               "traffic_bytes": 250.0},
              {"kernel": "k_plain", "algorithmic_bytes": 1000, "traffic_bytes": 1010.0}]}}
    (tmp_path / "profiles").mkdir()
    (tmp_path / "profiles" / "hbm_traffic.json").write_text(json.dumps(recs))
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    assert bench.measured_traffic("w", "k_persistent", 1000, 64) == 250.0
    assert bench.measured_traffic("w", "k_persistent", 1000, 1) == 900.0
    assert bench.measured_traffic("w", "k_persistent", 1000, 32) is None
    assert bench.measured_traffic("w", "k_persistent", 999, 64) is None
    assert bench.measured_traffic("w", "k_plain", 1000) == 1010.0
    assert bench.measured_traffic("other", "k_plain", 1000) is None
    (tmp_path / "profiles" / "hbm_traffic.json").write_text("{not json")
    assert bench.measured_traffic("w", "k_plain", 1000) is None


def test_cpu_model_is_a_name():
    name = bench.cpu_model()
    assert isinstance(name, str) and name


@pytest.mark.parametrize("name", ["batched", "plain", "large"])
def test_committed_lines_keep_the_contract(name):
    line = load_line(name)
    assert CONTRACT_KEYS <= set(line)
    assert line["n_gpus"] == 1 and line["higher_is_better"] is True
    assert line["steps"] > 0 and line["value"] > 0 and "workload" in line["config"]
    roof = line["roofline"]
    assert ROOFLINE_KEYS <= set(roof)
    assert roof["bound"] == "hbm" and roof["unit"] == "GB/s" and roof["peak"] == 8000.0
    # achieved = algorithmic bytes of one launch / its average duration
    gbs = roof["algorithmic_bytes_per_launch"] / (roof["avg_kernel_ms"] * 1e-3) / 1e9
    assert roof["achieved"] == pytest.approx(gbs, rel=1e-9)
    assert roof["frac"] == pytest.approx(roof["achieved"] / roof["peak"], rel=1e-12)
    cpu = line["cpu_baseline"]
    assert CPU_KEYS <= set(cpu) and cpu["kind"] in ("port", "reference")
    assert cpu["cores"] >= 1 and cpu["unit"] == line["unit"]


@pytest.mark.parametrize("name", ["batched", "plain", "large"])
def test_committed_traffic_matches_records(name):
    """The roofline's traffic in each committed line is the rocprofv3 record
    of the same kernel, input and stripe count."""
    line = load_line(name)
    roof = line["roofline"]
    stripes = (line["config"]["stripes"] if roof["kernel"] == "wmh_persistent_kernel"
               else None)
    traffic = bench.measured_traffic(name, roof["kernel"],
                                     roof["algorithmic_bytes_per_launch"], stripes)
    assert traffic is not None and traffic == roof["traffic"]
