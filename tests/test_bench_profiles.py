"""bench.py's measurement helpers on the CPU: the rocprofv3 summary of the
timed window (synthetic kernel-trace and counter CSVs laid out as rocprofv3
writes them) and the parsing of reyn_profile_summary lines."""

import csv
import importlib.util
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def name(k):
    return f"void reyn::(anonymous namespace)::{k}<1, false, true, false>(reyn::SceneArg, float*)"


def steps(warm, timed, iters):
    """Dispatches of warm + timed steps: k_vmax, k_density, then per
    iteration k_iisph_pacc and k_iisph_jacobi (plus one launch of each that
    finds the loop converged), then k_bounds; returns (kernel, duration ns)."""
    out = []
    for s in range(warm + timed):
        out += [("k_vmax", 10_000), ("k_density", 480_000)]
        for _ in range(iters):
            out += [("k_iisph_pacc", 120_000), ("k_iisph_jacobi", 200_000)]
        out += [("k_iisph_pacc", 5_000), ("k_iisph_jacobi", 6_000), ("k_bounds", 8_000)]
    out += [("k_reduce_final", 7_000)]  # after the timed steps
    return out


def write_trace(d, disp, gap=2_000):
    d.mkdir(parents=True)
    t = 1_000_000
    with open(d / "run_kernel_trace.csv", "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Kind", "Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        for i, (k, dur) in enumerate(disp, 1):
            w.writerow(["KERNEL_DISPATCH", i, name(k), t, t + dur])
            t += dur + gap


def write_counters(d, disp, counter, per_kernel):
    d.mkdir(parents=True)
    with open(d / "run_counter_collection.csv", "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value", "VGPR_Count", "SGPR_Count",
                    "LDS_Block_Size"])
        for i, (k, dur) in enumerate(disp, 1):
            # a launch that returns at once counts almost nothing
            w.writerow([i, name(k), counter, per_kernel.get(k, 1.0) * (dur > 20_000 or k in ("k_vmax", "k_bounds")),
                        40, 32, 0])


def test_trace_summary_covers_the_timed_window(bench, tmp_path):
    disp = steps(warm=3, timed=2, iters=4)
    write_trace(tmp_path / "kt", disp)
    out = bench.rocprof_summary("trace", [str(tmp_path / "kt")], 3)
    k = out["kernels"]
    assert k["k_density"]["launches"] == 2
    assert k["k_iisph_jacobi"]["launches"] == 10 and k["k_iisph_jacobi"]["working_launches"] == 8
    assert k["k_iisph_jacobi"]["working_avg_us"] == pytest.approx(200.0)
    assert k["k_iisph_pacc"]["total_ms"] == pytest.approx((8 * 120 + 2 * 5) / 1e3)
    assert "k_reduce_final" not in k  # after the last k_bounds
    per_step = 10 + 480 + 4 * 320 + 5 + 6 + 8
    assert out["kernels_ms"] == pytest.approx(2 * per_step / 1e3)


def test_counter_summaries_average_working_launches(bench, tmp_path):
    disp = steps(warm=2, timed=2, iters=3)
    for nm, c in (("f", "FETCH_SIZE"), ("w", "WRITE_SIZE"), ("v", "SQ_INSTS_VALU")):
        write_counters(tmp_path / nm, disp, c, {"k_iisph_jacobi": 1000.0, "k_density": 500.0})
    # the copy calibration: one k_gather reading 512 MiB and writing 256 MiB, counted as half and all
    for nm, c, v in (("cf", "FETCH_SIZE", 262144.0), ("cw", "WRITE_SIZE", 262144.0)):
        write_counters(tmp_path / nm, [("k_gather", 100_000)], c, {"k_gather": v})
    t = bench.rocprof_summary("traffic", [str(tmp_path / x) for x in ("f", "w", "cf", "cw")], 2)
    assert t["fetch_scale"] == pytest.approx(2.0) and t["write_scale"] == pytest.approx(1.0)
    j = t["kernels"]["k_iisph_jacobi"]
    assert j["working_launches"] == 6 and j["fetch_kib_raw"] == pytest.approx(1000.0)
    assert t["bytes_per_launch"]["k_iisph_jacobi"] == pytest.approx(1024 * (2000.0 + 1000.0))
    v = bench.rocprof_summary("valu", [str(tmp_path / "v")], 2)
    assert v["valu_insts_per_launch"]["k_density"] == 500


def test_address_summary(bench, tmp_path):
    disp = steps(warm=2, timed=2, iters=3)
    write_counters(tmp_path / "a", disp, "TA_TA_BUSY_sum", {"k_iisph_jacobi": 4000.0, "k_density": 900.0})
    a = bench.rocprof_summary("address", [str(tmp_path / "a")], 2)
    assert a["ta_busy_cycles_per_launch"]["k_iisph_jacobi"] == 4000
    assert a["kernels"]["k_iisph_jacobi"]["working_launches"] == 6
    assert a["cycles_per_s"] == pytest.approx(256 * 2.4e9)


def test_committed_address_profile_is_a_fraction_of_the_launch(bench):
    """The committed TA busy cycles of the dominant kernel, over its launch
    time in the committed kernel trace, are a fraction of one TA per CU."""
    import json

    ta = bench.measured_ta_busy("k_iisph_jacobi")
    assert ta is not None
    trace = (ROOT / "profiles" / "iisph_dambreak_2M_kernel_trace.md").read_text()
    row = next(line for line in trace.splitlines() if line.startswith("| k_iisph_jacobi |"))
    avg_us = float(row.split("|")[4])
    frac = ta / (avg_us * 1e-6) / bench.TA_CYCLES_PER_S
    assert 0.3 < frac <= 1.0, frac
    assert json.loads((ROOT / "profiles" / "address_path.json").read_text())["commands"]


def test_window_needs_the_warmup_steps(bench, tmp_path):
    write_trace(tmp_path / "kt", steps(warm=1, timed=1, iters=2))
    with pytest.raises(SystemExit):
        bench.rocprof_summary("trace", [str(tmp_path / "kt")], 5)


def test_profile_summary_lines(bench):
    got = bench.parse_profile("k_iisph_jacobi 12 2.5 10 2.4\nk_density 2 0.9 2 0.9\n")
    assert got == {"k_iisph_jacobi": (12, 2.5, 10, 2.4), "k_density": (2, 0.9, 2, 0.9)}
