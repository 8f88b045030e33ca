"""The byte and flop accounting bench.py reports (CPU): the survey's per-element-stage bytes, the
per-kernel algorithmic bytes from their array-by-array definition, the roofline fractions and the
matrix-core utilisation from a synthetic run record."""
import pytest

import bench


def test_stage_bytes_are_the_survey_accounting():
    # B32 = 80 Np + 53 and B64 = 117.33 Np + 91 bytes per element-stage; 853 and 1264.3 at p = 3
    assert bench.stage_bytes_min(10, True) == pytest.approx(853.0)
    assert bench.stage_bytes_min(10, False) == pytest.approx(1264.3)


def test_kernel_bytes_count_every_array_once():
    np_, ne_pts, nel, nedge = 10, 5, 1000, 1500
    n = 4 * np_
    kb = bench.kernel_bytes(np_, ne_pts, 4, nel, nedge)
    # surface: each element's modes once, per edge two ids + info, normal and half length, and the point fluxes
    assert kb["surface"] == nel * n * 4 + nedge * (12 + 12 + ne_pts * 16)
    # element (FP32-bulk): state, metric + 2/J, three edge ids, the RK traffic averaged over the three stages,
    # and every edge's point fluxes, normal and half length once
    rk = (2 * n * 4 + 3 * n * 4 + (n * 4 + n * 8 + n * 8 + n * 4)) / 3.0
    assert kb["element"] == pytest.approx(nel * (n * 4 + 20 + 12 + rk) + nedge * (ne_pts * 16 + 12))
    kb64 = bench.kernel_bytes(np_, ne_pts, 8, nel, nedge)
    assert kb64["surface"] == nel * n * 8 + nedge * (12 + 24 + ne_pts * 32)


def test_summary_fractions_and_matrix_core_utilisation():
    r = dict(ne=128000, np=10, nq=16, ng=5, nedge=192640, wall=0.1714 * 2000 / 1e3, t=0.0, dt=0.0,
             troubled=[100, 100, 100], tuning={"element_waves": 4},
             kt=dict(element_ms=0.0227, surface_ms=0.0143, detect_ms=0.011, rebuild_ms=0.006, apply_ms=0.0,
                     time_step_ms=0.0, control_ms=0.0))
    s = bench.summarise(r, 2000, True)
    roof = s["roofline"]
    # step: 853 B x 128000 x 3 per 0.1714 ms
    assert roof["achieved"] == pytest.approx(853.0 * 128000 * 3 / 0.1714e-3 / 1e9, rel=1e-9)
    assert roof["frac"] == pytest.approx(roof["achieved"] / 8000.0)
    dom = roof["dominant_kernel"]
    assert dom["name"] == "k_element_mfma"
    assert dom["achieved"] == pytest.approx(853.0 * 128000 / 22.7e-6 / 1e9, rel=1e-9)
    m = s["mfma"]
    # 64 MFMA of 2048 flops per 16 elements per launch
    assert m["gflop_issued_per_launch"] == pytest.approx(128000 / 16 * 64 * 2048 / 1e9)
    assert m["issued_frac"] == pytest.approx(m["issued_tflops"] / 157.3)
    assert 0 < m["useful_frac"] < m["issued_frac"]
    assert s["value"] == pytest.approx(128000 * 40 * 3 * 2000 / r["wall"] / 1e9)


def test_more_than_one_gpu_is_refused_before_any_gpu_work():
    """`bench.py --gpus N` with N > 1 exits with status 2 and a message naming the one-GPU limit, before it
    loads the library or touches a device."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stdout + r.stderr
    assert r.stdout.strip() == "" or not r.stdout.strip().startswith("{")
