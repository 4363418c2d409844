"""krabmaga_amd.profsum: rocprofv3 CSVs grouped into schedule steps."""
import csv

from krabmaga_amd import profsum

TRACE_COLS = ["Kernel_Name", "Grid_Size", "Workgroup_Size", "Start_Timestamp", "End_Timestamp"]
PMC_COLS = ["Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"]
# a counter pass's own launch columns, as rocprofv3 writes them beside the counter
PMC_TIMED_COLS = ["Kernel_Name", "Grid_Size", "Workgroup_Size", "Start_Timestamp", "End_Timestamp",
                  "Counter_Name", "Counter_Value"]
RT = "void kg::k_rt_step<256>(kg::RepDev const*, int)"
RT512 = "void kg::k_rt_step<512>(kg::RepDev const*, int)"


def write(path, cols, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(cols)
        w.writerows(rows)


def trace_rows():
    rows = []
    # three steps: class 0 (the largest 256-thread grid), class 1 started
    # just before class 0, and the deferred-tile launch after both
    for k in range(3):
        t = 1_000_000 * (k + 1)
        rows.append([RT512, 4096, 512, t - 2_000, t + 150_000])
        rows.append([RT, 1235200, 256, t, t + 120_000])
        rows.append([RT512, 131072, 512, t + 160_000, t + 200_000])
    rows.append(["kg::k_add_u64(unsigned long*, unsigned long)", 64, 64, 5_000_000, 5_001_000])
    return rows


def test_launches_group_into_steps(tmp_path):
    p = tmp_path / "t.csv"
    write(p, TRACE_COLS, trace_rows())
    groups = profsum.steps_of(profsum.load_trace(str(p)))
    assert len(groups) == 3
    for g in groups:
        assert len(g) == 3
        span = (max(r["end"] for r in g) - min(r["start"] for r in g)) / 1000.0
        assert span == 202.0


def test_summary_reports_span_and_traffic(tmp_path):
    t = tmp_path / "t.csv"
    write(t, TRACE_COLS, trace_rows())
    copy = "__amd_rocclr_copyBuffer"
    # a 1 GiB copy reads half in this counter
    fetch = [[copy, 131072, 512, 9_000_000, 9_100_000, "FETCH_SIZE", 524288.0]]
    write_ = [[copy, 131072, "WRITE_SIZE", 1048576.0]]
    # the counter pass ran two steps, the traced run three: traffic is per
    # step of the counter pass
    for k in range(2):
        t0 = 1_000_000 * (k + 1)
        fetch += [[RT, 1235200, 256, t0, t0 + 120_000, "FETCH_SIZE", 1000.0],
                  [RT512, 4096, 512, t0 - 2_000, t0 + 150_000, "FETCH_SIZE", 24.0]]
        write_ += [[RT, 1235200, "WRITE_SIZE", 800.0], [RT512, 4096, "WRITE_SIZE", 0.0]]
    fp, wp = tmp_path / "f.csv", tmp_path / "w.csv"
    write(fp, PMC_TIMED_COLS, fetch)
    write(wp, PMC_COLS, write_)

    class A:
        pass

    a = A()
    a.trace, a.fetch, a.write = str(t), str(fp), str(wp)
    a.agents, a.skip, a.count = 1000, 0, 0
    a.title, a.cmd, a.pmc_cmd = "test", "python3 bench.py", None
    a.counters = None
    out = profsum.summarize(a)
    assert "step_span_us: 202.0" in out
    # per step: (1000 + 24) KiB x 2 fetched + 800 KiB written
    assert f"traffic_bytes_per_step: {(1024 * 2 + 800) * 1024}" in out


def test_counter_table_shares(tmp_path):
    rows = []
    for d in range(2):
        rows += [[RT, 1235200, "SQ_WAVES", 100.0, d], [RT, 1235200, "SQ_WAVE_CYCLES", 1000.0, d],
                 [RT, 1235200, "SQ_WAIT_ANY", 550.0, d], [RT, 1235200, "SQ_INSTS_VALU", 2000.0, d],
                 [RT, 1235200, "SQ_ACTIVE_INST_VALU", 400.0, d],
                 [RT512, 4096, "SQ_WAVES", 8.0, 10 + d]]
    p = tmp_path / "c.csv"
    write(p, PMC_COLS + ["Dispatch_Id"], rows)
    out = "\n".join(profsum.counter_table(str(p)))
    assert "| 1235200 | 2 |" in out
    assert "55.0 % of their cycles" in out
    assert "VALU busy 40.0 %" in out
    assert "20 VALU instructions per wave" in out


FLK = "kg::k_flock(kg::RepDev const*, int)"
STEP = [FLK, "kg::k_scan_reduce(unsigned int const*, unsigned int, unsigned int*)",
        "kg::k_scan_down(unsigned int*)", "kg::k_scatter_aligned(kg::RepDev const*)",
        "kg::k_sort_bins(kg::RepDev const*)"]


def three_kernel_rows():
    # three steps of five kernels, 100 + 10 + 10 + 20 + 20 us back to back,
    # plus an unrelated kernel between steps 1 and 2
    rows = []
    for k in range(3):
        t = 1_000_000 * (k + 1)
        for name, dur in zip(STEP, (100_000, 10_000, 10_000, 20_000, 20_000)):
            rows.append([name, 642304, 256, t, t + dur])
            t += dur
    rows.append(["kg::k_rt_step<512>(kg::RepDev const*)", 4096, 512, 2_500_000, 2_600_000])
    # another field's work after the last step ends that step
    rows.append(["kg::k_count_append(kg::Buf)", 1048576, 256, 3_500_000, 3_900_000])
    return rows


def test_three_kernel_steps_span_and_traffic(tmp_path):
    t = tmp_path / "t.csv"
    write(t, TRACE_COLS, three_kernel_rows())
    groups = profsum.three_kernel_steps(profsum.load_trace(str(t)), "kg::k_flock")
    # neither the k_rt_step launch nor the k_count_append one is a step kernel
    assert [len(g) for g in groups] == [5, 5, 5]
    assert all(g[-1]["end"] - g[0]["start"] == 160_000 for g in groups)
    copy = "__amd_rocclr_copyBuffer"
    fetch = [[copy, 131072, "FETCH_SIZE", 524288.0]]
    write_ = [[copy, 131072, "WRITE_SIZE", 1048576.0]]
    for _ in range(3):
        for name, kib in zip(STEP, (300.0, 10.0, 20.0, 100.0, 50.0)):
            fetch.append([name, 642304, "FETCH_SIZE", kib])
            write_.append([name, 642304, "WRITE_SIZE", kib / 2])
    # a dispatch of one step kernel at another grid (another field) is left out
    fetch.append([STEP[1], 1024, "FETCH_SIZE", 9999.0])
    fp, wp = tmp_path / "f.csv", tmp_path / "w.csv"
    write(fp, PMC_COLS, fetch)
    write(wp, PMC_COLS, write_)

    class A:
        pass

    a = A()
    a.trace, a.fetch, a.write = str(t), str(fp), str(wp)
    a.agents, a.skip, a.count = 1000, 1, 0
    a.title, a.cmd, a.pmc_cmd = "test", "python3 bench.py", None
    a.counters = None
    a.step_kernel, a.step_grid = "kg::k_flock", 642304
    out = profsum.summarize(a)
    assert "step_span_us: 160.0" in out
    assert "flock_us: 100.0" in out
    # per step: 480 KiB fetched (counted twice by FETCH_SIZE) + 240 KiB written
    assert f"traffic_bytes_per_step: {(480 * 2 + 240) * 1024}" in out
    assert "kernel_us_per_step: 160.0" in out
