"""The benchmark's contract: one GPU only, the PMC records it quotes, and (on a GPU) one
JSON line with the fields a caller of bench.py reads."""
import importlib.util
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench_module():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_pmc_records_are_quoted_only_for_their_instance_and_workload():
    b = _bench_module()
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        rec = json.load(f)["w3_native"]
    cfg = dict(rec["config"])
    assert b.measured_traffic("native", cfg, 128, 4096) == rec["traffic_bytes"]
    assert b.measured_traffic("native", cfg, 64, 4096) is None
    assert b.measured_traffic("native", cfg, 128, 1024) is None
    assert b.measured_traffic("native", dict(cfg, ranks=cfg["ranks"] + 1), 128, 4096) is None
    assert b.measured_traffic("split_fullk", cfg, 128, 4096) is None
    # traffic within 0.1 % of the algorithmic bytes: A and B read once, C written once
    assert abs(rec["traffic_bytes"] / rec["algorithmic_bytes"] - 1.0) < 1e-3
    lds = b.measured_lds("native", cfg, 128, 4096)
    assert lds["bytes_per_flop"] == pytest.approx(0.3125)
    assert lds["bank_conflict_cycles"] == 0
    assert b.measured_lds("native", cfg, 128, 1024) is None


def test_cpu_model_is_read_from_the_host():
    model = _bench_module().cpu_model()
    assert model is None or (isinstance(model, str) and model)


def test_graft_entry_points_exist():
    spec = importlib.util.spec_from_file_location("graft_entry",
                                                  os.path.join(ROOT, "__graft_entry__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    assert callable(mod.build) and callable(mod.smoke)


@pytest.mark.gpu
def test_bench_prints_one_json_line_with_the_contract_fields():
    print("\n[case] bench.py --quick --no-cpu --steps 5 --warmup 2", flush=True)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--quick", "--no-cpu",
                        "--steps", "5", "--warmup", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
                "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
                "roofline", "cpu_baseline"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["steps"] == 5 and d["warmup"] == 2
    assert d["value"] > 0 and d["unit"] == "TFLOPS" and d["higher_is_better"] is True
    roof = d["roofline"]
    assert roof["peak"] > 0 and roof["frac"] == pytest.approx(roof["achieved"] / roof["peak"])
    assert d["max_rel_err"]["componentwise"] <= 1e-13
    assert "workload" in d["config"]
    assert list(d)[-1] == "targets" and d["targets"]["bound_componentwise"] == 1e-13


def test_targets_summarise_the_paths_and_workloads():
    b = _bench_module()
    res = {"native": {"available": True, "tflops": 40.0},
           "split_i8": {"available": True, "tflops": 60.0,
                        "max_rel_err": {"componentwise": 3e-15, "normwise": 1e-16}},
           "split": {"available": False}}
    w = {"w3r_resident_f64_64cubed_x4096": {
             "split_i8_over_native": 1.5, "split_f16_over_native": 1.1,
             "split_i8": {"max_rel_err_vs_torch_fp64": {"componentwise": 2e-15}},
             "split_f16": {"max_rel_err_vs_torch_fp64": {"componentwise": 5e-14}}},
         "w2_f16_block_gemm": {"64x64x64": {"1d": {"resident_frac_f16_peak": 0.5},
                                            "2d": {"resident_frac_f16_peak": 0.74},
                                            "3d": {"available": False}}}}
    t = b.targets(res, w)
    assert t["w3_split_i8_over_native"] == pytest.approx(1.5)
    assert "w3_split_over_native" not in t
    assert t["w3r_split_i8_over_native"] == 1.5 and t["w3r_split_f16_componentwise"] == 5e-14
    assert t["w2_best_resident_frac_f16_peak"] == {"64x64x64": [0.74, "2d"]}
