"""The rocprofv3 summary writer: kernel keys, the counter scaling by the device copy, the traffic column
that bench.py reads back, and the trace-only table. Inputs are small synthetic CSVs in rocprofv3's
column layout."""
import csv
import os

import pytest

from dgm_in_rust_amd import rocprof_summary as rs

SURF = "void dgm::(anonymous namespace)::k_surface<float, 10, 16, 5>(dgm::DevMesh, float const*, float*)"
WALK = "void dgm::(anonymous namespace)::k_limit_walk<double, 2, 6, 3, 0>(dgm::DevMesh, int const*)"
COPY = "__amd_rocclr_copyBuffer"


def write_csv(path, header, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def stats(out, rows):
    write_csv(os.path.join(out, "trace", "kt_kernel_stats.csv"),
              ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs", "StdDev"], rows)


def pmc(out, sub, counter, rows):
    write_csv(os.path.join(out, sub, "pmc_counter_collection.csv"),
              ["Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"], [[k, g, counter, v] for k, g, v in rows])


def test_kernel_key():
    assert rs.kernel_key(SURF) == "k_surface<float,10,16,5>"
    assert rs.kernel_key(COPY) == COPY


def test_summary_scales_counters_by_the_copy(tmp_path):
    out = str(tmp_path)
    stats(out, [[SURF, 10, 150000, 15000, 50, 14000, 16000, 1], [COPY, 2, 200000, 100000, 50, 1000, 199000, 1]])
    kib = 256 * 1024
    # the largest copy reads half its true size on FETCH_SIZE and its true size on WRITE_SIZE
    pmc(out, "fetch", "FETCH_SIZE", [[COPY, 1024, 10.0], [COPY, 1 << 26, kib / 2], [SURF, 4096, 1000.0]])
    pmc(out, "write", "WRITE_SIZE", [[COPY, 1024, 10.0], [COPY, 1 << 26, kib], [SURF, 4096, 500.0]])
    text = rs.summarise(out, "t", "--steps 1")
    assert "x2.00" in text and "x1.00" in text
    path = tmp_path / "s.txt"
    path.write_text(text)
    traffic, us = rs.read_traffic(str(path), "k_surface<float")
    assert traffic == pytest.approx((1000.0 * 2 + 500.0) * 1024.0, rel=1e-6)
    assert us == pytest.approx(15.0)
    assert rs.read_traffic(str(path), "k_element") is None
    assert rs.read_traffic(str(tmp_path / "absent.txt"), "k_surface") is None


def test_trace_only_summary(tmp_path):
    out = str(tmp_path)
    stats(out, [[SURF, 4, 60000, 15000, 10, 14000, 16000, 1], [WALK, 3, 2640000, 880000, 90, 870000, 890000, 1]])
    text = rs.trace_summary(out, "O-grid", "python3 -m dgm_in_rust_amd.studies --ogrid --steps 100")
    lines = [l for l in text.splitlines() if not l.startswith("#")]
    assert lines[0].split()[:2] == ["kernel", "calls"]
    # sorted by total time: the walk first
    assert lines[1].split()[0] == "k_limit_walk<double,2,6,3,0>"
    assert float(lines[1].split()[2]) == pytest.approx(880.0)
    assert "--ogrid --steps 100" in text


def test_instruction_table_per_wave(tmp_path):
    """The SQ run's counters become wavefronts per dispatch and instructions per wavefront, in rows that
    read_traffic() does not take for kernel rows."""
    out = str(tmp_path)
    stats(out, [[SURF, 2, 30000, 15000, 50, 14000, 16000, 1], [COPY, 1, 100000, 100000, 50, 100000, 100000, 1]])
    kib = 256 * 1024
    pmc(out, "fetch", "FETCH_SIZE", [[COPY, 1 << 26, kib / 2], [SURF, 4096, 1000.0]])
    pmc(out, "write", "WRITE_SIZE", [[COPY, 1 << 26, kib], [SURF, 4096, 500.0]])
    rows = []
    for waves, valu, mfma, salu in ((3000, 3.6e6, 0, 1.2e6), (3000, 3.9e6, 0, 1.5e6)):
        rows += [[SURF, 4096, "SQ_WAVES", waves], [SURF, 4096, "SQ_INSTS_VALU", valu],
                 [SURF, 4096, "SQ_INSTS_VALU_MFMA_F32", mfma], [SURF, 4096, "SQ_INSTS_SALU", salu]]
    write_csv(os.path.join(out, "sq", "pmc_counter_collection.csv"),
              ["Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"], rows)
    text = rs.summarise(out, "t", "--steps 1")
    row = [l for l in text.splitlines() if l.startswith("sq k_surface")]
    assert len(row) == 1
    f = row[0].split()
    assert float(f[2]) == 3000 and float(f[3]) == pytest.approx(1250.0) and float(f[4]) == 0.0
    assert float(f[5]) == pytest.approx(450.0)
    path = tmp_path / "s.txt"
    path.write_text(text)
    traffic, us = rs.read_traffic(str(path), "k_surface<float")
    assert traffic == pytest.approx((1000.0 * 2 + 500.0) * 1024.0, rel=1e-6) and us == pytest.approx(15.0)
