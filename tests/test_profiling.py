"""CPU checks of the rocprofv3 summary tool (memphys_amd/profiling.py) on
synthetic result databases with the rocpd tables it reads: the working-launch
rule, the iteration extraction and the counter scaling of the HBM traffic."""
import sqlite3

import pytest

import memphys_amd.profiling as P


def _kernels_db(path, rows):
    con = sqlite3.connect(path)
    con.execute("create table kernels(name text, start int, end int, grid_x int)")
    con.executemany("insert into kernels values(?,?,?,?)", rows)
    con.commit()
    con.close()
    return str(path)


def test_stats_working_launches(tmp_path):
    # 6 launches of 100 us and 4 of 2 us (a converged system's early returns)
    rows, t = [], 0
    for d in [100_000] * 6 + [2_000] * 4:
        rows.append(("void mp::k_spmv(int)", t, t + d, 999424))
        t += d + 1000
    out = P.stats(_kernels_db(tmp_path / "a.db", rows), 10)
    line = [ln for ln in out.splitlines() if ln.startswith("mp::k_spmv")][0].split()
    calls, working, work_us = int(line[2]), int(line[3]), float(line[6])
    assert (calls, working) == (10, 6) and work_us == pytest.approx(100.0)


def test_iteration_picks_the_median_iteration(tmp_path):
    rows, t = [], 0

    def add(name, d, grid):
        nonlocal t
        rows.append((name, t, t + d, grid))
        t += d + 500

    for it, spmv in enumerate([60_000, 70_000, 65_000, 1_000]):  # the last: a stopped system
        add("mp::k_gm_scale<double, float>(x)", 6_000, 998912)
        add("mp::k_amg_op<0, 3, double, float>(x)", 58_000, 999424)
        add("mp::k_gm_spmv<float, float, float>(x)", spmv, 999424)
        add("mp::k_gm_dots<float>(x)", 12_000, 998912)
    add("mp::k_gm_scale<double, float>(x)", 6_000, 998912)
    add("mp::k_amg_op<0, 3, double, float>(x)", 58_000, 999424)
    out = P.iteration(_kernels_db(tmp_path / "b.db", rows), "k_gm_scale", "k_amg_op",
                      "k_gm_spmv").splitlines()
    assert out[0].startswith("iterations 3,")  # the stopped system's stretch is left out
    spmv = [ln for ln in out if "k_gm_spmv" in ln][0]
    assert "dur    65.0 us" in spmv  # median of the three working iterations


def test_traffic_scales_counters_by_the_copy_kernel(tmp_path):
    con = sqlite3.connect(tmp_path / "f.db")
    con.execute("create table counters_collection(kernel_name text, grid_size int, "
                "counter_name text, value real, duration int)")
    copy_kib = 1 << 20  # a 1 GiB copy
    con.execute("insert into counters_collection values(?,?,?,?,?)",
                ("neg_kernel", 64, "FETCH_SIZE", 0.5 * copy_kib, 1000))
    for _ in range(3):
        con.execute("insert into counters_collection values(?,?,?,?,?)",
                    ("mp::k_gm_spmv", 999424, "FETCH_SIZE", 0.5 * 1000.0, 70_000))
    con.commit()
    con.close()
    con = sqlite3.connect(tmp_path / "w.db")
    con.execute("create table counters_collection(kernel_name text, grid_size int, "
                "counter_name text, value real, duration int)")
    con.execute("insert into counters_collection values(?,?,?,?,?)",
                ("neg_kernel", 64, "WRITE_SIZE", 1.0 * copy_kib, 1000))
    for _ in range(3):
        con.execute("insert into counters_collection values(?,?,?,?,?)",
                    ("mp::k_gm_spmv", 999424, "WRITE_SIZE", 8.0, 70_000))
    con.commit()
    con.close()
    out = P.traffic(str(tmp_path / "f.db"), str(tmp_path / "w.db"), "k_gm_spmv", "neg",
                    1 << 30, 999424).splitlines()
    vals = {ln.split(" = ")[0]: float(ln.split(" = ")[1]) for ln in out if " = " in ln}
    assert vals["fetch_bytes_per_launch"] == pytest.approx(1000.0 * 1024)
    assert vals["write_bytes_per_launch"] == pytest.approx(8.0 * 1024)
    assert vals["traffic_bytes_per_launch"] == pytest.approx(1008.0 * 1024)


def test_bench_reads_committed_traffic_summaries():
    # The bench's `traffic` fields come from the committed rocprofv3 summaries;
    # each must be the measured figure of its own kernel, within 1 % of that
    # kernel's algorithmic bytes per launch (SpMV FP32 279,638,428; FP64
    # 545,879,136; multigrid level-0 post-smoothing of the GMRES form, which
    # reads the FP32 basis vector and writes the FP32 direction, 190,656,204),
    # and absent on other workloads and for the BiCGSTAB form's product.
    import importlib.util
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(root, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    for got, alg in [(bench.measured_traffic(100, 0), 279_638_428),
                     (bench.measured_traffic(100, 1), 545_879_136),
                     (bench.smooth_roofline((381, 0.06, 190_656_204.0, 393), 6000.0,
                                            100)["traffic"], 190_656_204)]:
        assert got is not None and abs(got / alg - 1) < 0.01
    assert bench.measured_traffic(50, 0) is None
    assert bench.smooth_roofline((381, 0.06, 2e8, 393), 6000.0, 50)["traffic"] is None
    assert bench.smooth_roofline((381, 0.06, 2e8, 393), 6000.0, 100, gmres=False)["traffic"] is None


def test_steps_splits_time_and_idle(tmp_path):
    # two whole timesteps (three predictor launches), each: predictor 500 us,
    # a 20 us gap, one level-0 smoothing product 50 us, a coarse product 5 us,
    # a GMRES SpMV 60 us, a 100 us gap, the corrector 300 us
    rows, t = [], 0
    for step in range(3):
        seq = [("void mp::predictor_kernel<3>(int)", 500_000, 998_912, 20_000),
               ("void mp::k_amg_op<0, 3, float, float>(int)", 50_000, 999_424, 0),
               ("void mp::k_amg_op<8, 1, float, float>(int)", 5_000, 655_360, 0),
               ("void mp::k_gm_spmv<float, float, float>(int)", 60_000, 999_424, 100_000),
               ("void mp::corrector_kernel<3>(int)", 300_000, 998_912, 0)]
        for name, d, grid, gap in seq:
            rows.append((name, t, t + d, grid))
            t += d + gap
    out = P.steps(_kernels_db(tmp_path / "s.db", rows)).splitlines()
    assert out[0].startswith("timesteps 2")
    cells = out[2].split()
    wall, busy, idle, launches = float(cells[1]), float(cells[2]), float(cells[3]), int(cells[4])
    assert (wall, busy, idle) == pytest.approx((1.035, 0.915, 0.12), abs=0.006)  # (2 decimals)
    assert launches == 5
    pred, rhs, corr, guess, gmres, bicg, l0, mg, other = map(float, cells[5:14])
    assert (pred, corr, gmres, l0, mg) == (0.5, 0.3, 0.06, 0.05, 0.005)
    assert rhs == guess == bicg == other == 0.0
    idle_line = [ln for ln in out if ln.startswith("largest idle")][0]
    assert "0.120 ms over 2 stretches" in idle_line
    assert "100.0 us after mp::k_gm_spmv<float, float, float>" in "\n".join(out)
