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
              {"kernel": "k_persistent", "algorithmic_bytes": 1000, "stripes": 64,
               "seeds": True, "traffic_bytes": 200.0},
              {"kernel": "k_plain", "algorithmic_bytes": 1000, "traffic_bytes": 1010.0}]}}
    (tmp_path / "profiles").mkdir()
    (tmp_path / "profiles" / "hbm_traffic.json").write_text(json.dumps(recs))
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    assert bench.measured_traffic("w", "k_persistent", 1000, 64) == 250.0
    assert bench.measured_traffic("w", "k_persistent", 1000, 64, seeds=True) == 200.0
    assert bench.measured_traffic("w", "k_persistent", 1000, 1, seeds=True) is None
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


def test_roofline_picks_the_binding_resource(tmp_path, monkeypatch):
    """bound = the ceiling closer to its peak among VALU issue (SQ_INSTS_VALU x
    64 lanes against 157.3e12 / 4 lane operations per second) and HBM (fetched +
    written bytes against 8 TB/s); without a record, the algorithmic bytes."""
    recs = {"workloads": {"w": [
        {"kernel": "wmh_persistent_kernel", "algorithmic_bytes": 1e12, "stripes": 48, "traffic_bytes": 4e9,
         "valu_insts": 3e9, "profile": "p.md"},
        {"kernel": "wmh_persistent_kernel", "algorithmic_bytes": 1e12, "stripes": 7, "traffic_bytes": 7e9,
         "valu_insts": 1e8}]}}
    (tmp_path / "profiles").mkdir()
    (tmp_path / "profiles" / "hbm_traffic.json").write_text(json.dumps(recs))
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    r = bench.roofline("w", "wmh_persistent_kernel", 1e12, 1.0, 48)  # 1 ms
    valu = 3e9 * 64 / 1e-3 / bench.VALU_PEAK_LANE_OPS
    hbm = 4e9 / 1e-3 / 8e12
    assert r["bound"] == "valu" and r["frac"] == pytest.approx(valu)
    assert r["valu_frac"] == pytest.approx(valu) and r["hbm_traffic_frac"] == pytest.approx(hbm)
    assert r["achieved"] == pytest.approx(3e9 * 64 / 1e-3 / 1e12)
    assert r["hbm_algorithmic_x"] == pytest.approx(1e12 / 1e-3 / 8e12)
    r = bench.roofline("w", "wmh_persistent_kernel", 1e12, 1.0, 7)
    assert r["bound"] == "hbm" and r["frac"] == pytest.approx(7e9 / 1e-3 / 8e12)
    assert r["achieved"] == pytest.approx(7e9 / 1e-3 / 1e9) and r["unit"] == "GB/s"
    r = bench.roofline("w", "wmh_persistent_kernel", 1e12, 1.0, 9)
    assert r["bound"] == "hbm" and r["traffic"] is None
    assert r["frac"] == pytest.approx(1e12 / 1e-3 / 8e12)


def test_serve_summary_picks_the_highest_stable_rate():
    sweep = [{"offered_qps": 1000.0, "achieved_qps": 1001.0, "p50_ms": 1, "p99_ms": 2,
              "nnz_samples_per_sec": 1e10},
             {"offered_qps": 4000.0, "achieved_qps": 3950.0, "p50_ms": 3, "p99_ms": 9,
              "nnz_samples_per_sec": 4e10},
             {"offered_qps": 8000.0, "achieved_qps": 6000.0, "p50_ms": 90, "p99_ms": 300,
              "nnz_samples_per_sec": 6e10}]
    s = bench.serve_summary(sweep)
    assert s["stable_offered_qps"] == 4000.0 and (s["p50_ms"], s["p99_ms"]) == (3, 9)
    assert s["saturation_nnz_samples_per_sec"] == 6e10 and s["saturation_p99_ms"] == 300


@pytest.mark.parametrize("name", ["batched", "plain", "large"])
def test_committed_lines_keep_the_contract(name):
    line = load_line(name)
    assert CONTRACT_KEYS <= set(line)
    assert line["n_gpus"] == 1 and line["higher_is_better"] is True
    assert line["steps"] > 0 and line["value"] > 0 and "workload" in line["config"]
    roof = line["roofline"]
    assert ROOFLINE_KEYS <= set(roof)
    assert roof["bound"] in ("valu", "hbm") and 0 < roof["frac"] <= 1
    # the line's roofline is what bench.roofline computes from the committed
    # counter record of the same kernel, input, stripe count and seed setting
    stripes = line["config"].get("stripes")
    again = bench.roofline(name, roof["kernel"], roof["algorithmic_bytes_per_launch"],
                           roof["avg_kernel_ms"], stripes,
                           line["config"].get("heavy_seeds", False))
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert again[key] == pytest.approx(roof[key]) if isinstance(roof[key], float) \
            else again[key] == roof[key], key
    cpu = line["cpu_baseline"]
    assert CPU_KEYS <= set(cpu) and cpu["kind"] in ("port", "reference")
    assert cpu["cores"] >= 1 and cpu["unit"] == line["unit"]
    assert cpu["sample_bit_exact_vs_gpu"] is True


def test_default_line_carries_the_large_share_and_serving():
    line = load_line("batched")
    assert line["batched_over_plain"] >= 3.0 and line["plain"]["identical_to_batched"]
    # a second draw of the batch's layout, hashed with the same tuning
    seed2 = line["seed2"]
    assert seed2["layout_seed"] == 2 and seed2["nnz"] != line["config"]["nnz"]
    assert 0.8 <= seed2["value"] / line["value"] <= 1.25
    large = line["large"]
    assert large["value"] > 0 and 0 < large["roofline"]["frac"] <= 1
    assert large["end_to_end"]["identical_to_resident_chunk"] is True
    serve = line["serve"]
    assert serve["stable_offered_qps"] and serve["p50_ms"] <= serve["p99_ms"]
    assert serve["saturation_nnz_samples_per_sec"] > 0


@pytest.mark.parametrize("name", ["batched", "plain", "large"])
def test_committed_traffic_matches_records(name):
    """The roofline's traffic in each committed line is the rocprofv3 record
    of the same kernel, input, stripe count and heavy-seed setting."""
    line = load_line(name)
    roof = line["roofline"]
    persistent = roof["kernel"] == "wmh_persistent_kernel"
    stripes = line["config"]["stripes"] if persistent else None
    seeds = persistent and line["config"].get("heavy_seeds", False)
    traffic = bench.measured_traffic(name, roof["kernel"],
                                     roof["algorithmic_bytes_per_launch"], stripes, seeds)
    assert traffic is not None and traffic == roof["traffic"]
