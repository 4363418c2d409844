"""Summaries of rocprofv3 output for the flockers schedule step.

Two device paths run a step.  The three-kernel path (the default for large
fields) launches k_flock (plus k_flock_global for crowded tiles), the scan
(k_scan_reduce and k_scan_down, or k_scan_partials, or none when the gather
does it), then k_rebin_gather and k_rebin_fallback, or k_scatter_aligned and
k_sort_bins, once each per step, one after another on one stream: with
--step-kernel kg::k_flock a
step runs from one k_flock launch to the next, and --fetch / --write sum
those kernels' bytes per step.  The tile-resident path is described next.

One schedule step of the tile-resident path is several launches of
k_rt_step: one per tile class (side by side on their own streams) and one
for the tiles deferred in that step.  This module groups the launches of a
kernel trace into steps -- each step has exactly one class-0 launch (one
work-group per tile: the most work-groups of any launch); a launch belongs
to the last class-0 launch that started before it, give or take `slack_us` -- and
reports per-step spans (first start to last end) beside per-launch
durations.  With FETCH_SIZE / WRITE_SIZE passes it reports the HBM bytes per
step, scaled by what the same counters read for a device-to-device copy of
known size (the copy the bench issues with --calibrate-copy).

    python -m krabmaga_amd.profsum --trace T.csv [--fetch F.csv --write W.csv]
        [--counters C.csv] --agents N --skip S --count K --title "..." --cmd "..." > out.md

Algorithmic bytes are 76 per agent-update (a step reads and writes each
agent's 20 bytes once to bin it, reads them again and writes the 16 bytes of
new position and heading), whatever the kernel split; the tile-resident
kernel itself moves 40 (one read, one write).  --counters tabulates any
other rocprofv3 --pmc counters (SQ_WAVES, SQ_WAIT_ANY, ...) per k_rt_step
launch class.
"""
import argparse
import collections
import csv
import statistics

KERNEL = "kg::k_rt_step"
FLOCK = "kg::k_flock"
# the kernels of one three-kernel schedule step
STEP_KERNELS = ("kg::k_flock", "kg::k_flock_global", "kg::k_scan_reduce", "kg::k_scan_down",
                "kg::k_rebin_gather", "kg::k_rebin_fallback", "kg::k_scatter_aligned",
                "kg::k_sort_bins")
# launched within a three-kernel step on some layouts: the tile-sum prefix
# pass of large scans, the device step counter
STEP_AUX_KERNELS = ("kg::k_scan_partials", "kg::k_add_u64")
COPY = "__amd_rocclr_copyBuffer"
BYTES_PER_AGENT = 76  # per agent-update of a whole schedule step
KERNEL_BYTES_PER_AGENT = 40  # the tile-resident kernel: read 20 bytes, write 20 back
FLOCK_BYTES_PER_AGENT = 36  # k_flock: read 20 bytes, write 16 (position, heading)


def _name(r):
    return r["Kernel_Name"].split("(")[0].replace("void ", "").split("<")[0]


def _grid(r):
    return int(r.get("Grid_Size") or r.get("Grid_Size_X"))


def _wg(r):
    return int(r.get("Workgroup_Size") or r.get("Workgroup_Size_X"))


def load_trace(path):
    rows = list(csv.DictReader(open(path)))
    out = []
    for r in rows:
        out.append({"name": _name(r), "grid": _grid(r), "wg": _wg(r),
                    "start": int(r["Start_Timestamp"]), "end": int(r["End_Timestamp"])})
    out.sort(key=lambda r: r["start"])
    return out


def steps_of(trace, slack_us=20.0):
    """Launch groups of k_rt_step, one per schedule step, in order."""
    rt = [r for r in trace if r["name"] == KERNEL]
    if not rt:
        return []
    c0_wgs = max(r["grid"] // r["wg"] for r in rt)
    is_c0 = [r["grid"] // r["wg"] == c0_wgs for r in rt]
    c0 = [r for r, c in zip(rt, is_c0) if c]
    starts = [r["start"] for r in c0]
    groups = [[r] for r in c0]
    slack = int(slack_us * 1000)
    j = 0
    for r, c in zip(rt, is_c0):
        if c:
            continue
        while j + 1 < len(starts) and starts[j + 1] <= r["start"] + slack:
            j += 1
        # launches before the first class-0 launch belong to it
        groups[j].append(r)
    return groups


def pmc_totals(path, counter):
    """Per-kernel sums and dispatch counts of one counter (KiB)."""
    tot = collections.defaultdict(float)
    n = collections.Counter()
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        k = (_name(r), _grid(r))
        tot[k] += float(r["Counter_Value"])
        n[k] += 1
    return tot, n


def three_kernel_steps(trace, step_kernel, grid=0):
    """Launch groups of the three-kernel path: each launch of `step_kernel`
    (of `grid` threads, if given) starts a step, which runs through the step's
    kernels that follow it and ends at the first launch of any other kernel
    (a flag read's copy, another field's work) or at the next step."""
    starts = [i for i, r in enumerate(trace)
              if r["name"] == step_kernel and (not grid or r["grid"] == grid)]
    groups = []
    for a, b in zip(starts, starts[1:] + [len(trace)]):
        g = [trace[a]]
        for r in trace[a + 1:b]:
            if r["name"] not in STEP_KERNELS + STEP_AUX_KERNELS:
                break
            g.append(r)
        groups.append(g)
    return groups


def summarize(a):
    tr = load_trace(a.trace)
    three = bool(getattr(a, "step_kernel", None))
    kname = a.step_kernel if three else KERNEL
    lines = [f"# {a.title}\n", "Commands (repository root, MI355X, ROCm 7.2):\n", "```"]
    lines.append(f"rocprofv3 --kernel-trace --stats --output-format csv -d <dir> -o trace -- {a.cmd}")
    if a.fetch:
        lines.append(f"rocprofv3 --pmc FETCH_SIZE --kernel-trace --output-format csv -d <dir> -o pmc -- "
                     f"{a.pmc_cmd or a.cmd}")
        lines.append(f"rocprofv3 --pmc WRITE_SIZE --kernel-trace --output-format csv -d <dir> -o pmc -- "
                     f"{a.pmc_cmd or a.cmd}")
    if a.counters:
        names = sorted({r["Counter_Name"] for r in csv.DictReader(open(a.counters))})
        lines.append(f"rocprofv3 --pmc {' '.join(names)} --kernel-trace --output-format csv "
                     f"-d <dir> -o pmc -- {a.pmc_cmd or a.cmd}")
    step = (f"--step-kernel {a.step_kernel} --step-grid {getattr(a, 'step_grid', 0) or 0} "
            if three else "")
    lines.append(f"python -m krabmaga_amd.profsum --trace <trace_kernel_trace.csv> ... {step}"
                 f"--agents {a.agents} --skip {a.skip} --count {a.count}")
    lines.append("```\n")
    # per-launch table
    d = collections.defaultdict(list)
    for r in tr:
        d[(r["name"], r["grid"], r["wg"])].append((r["end"] - r["start"]) / 1000.0)
    lines.append("Launches (kernel trace):\n")
    lines.append("| kernel | grid (threads) | work-group | launches | avg us | min us | max us |")
    lines.append("|---|---|---|---|---|---|---|")
    for k in sorted(d, key=lambda k: -sum(d[k])):
        v = d[k]
        if sum(v) < 50:
            continue
        lines.append(f"| {k[0]} | {k[1]} | {k[2]} | {len(v)} | {statistics.mean(v):.1f} | "
                     f"{min(v):.1f} | {max(v):.1f} |")
    # per-step spans
    groups = (three_kernel_steps(tr, a.step_kernel, getattr(a, "step_grid", 0) or 0) if three
              else steps_of(tr))
    sel = groups[a.skip:a.skip + a.count] if a.count else groups[a.skip:]
    spans = [(max(r["end"] for r in g) - min(r["start"] for r in g)) / 1000.0 for g in sel]
    per_step = None
    if spans:
        per_step = statistics.mean(spans)
        nl = statistics.mean(len(g) for g in sel)
        what = ("three-kernel step (launches from one k_flock to the next)" if three
                else "k_rt_step per schedule step (launches of one step grouped")
        lines.append(f"\n{what}; steps "
                     f"{a.skip}..{a.skip + len(sel) - 1} of {len(groups)} in the trace):\n")
        lines.append(f"- launches per step: {nl:.2f}")
        lines.append(f"- step span (first start to last end): mean {per_step:.1f} us, "
                     f"median {statistics.median(spans):.1f}, min {min(spans):.1f}, "
                     f"max {max(spans):.1f}")
        region = (max(r["end"] for r in sel[-1]) - min(r["start"] for r in sel[0])) / 1000.0
        lines.append(f"- the steps' region, first start to last end, per step: {region / len(sel):.1f} us "
                     f"(with the gaps between steps: flag reads, re-plans)")
        alg = BYTES_PER_AGENT * a.agents
        own = (f"" if three else
               f" (the kernel's own {KERNEL_BYTES_PER_AGENT} B: "
               f"{KERNEL_BYTES_PER_AGENT * a.agents / per_step / 1e3 / 8000 * 100:.2f} %)")
        lines.append(f"- algorithmic bytes: {BYTES_PER_AGENT} B x {a.agents} agents = {alg / 1e6:.2f} MB "
                     f"-> {alg / per_step / 1e3:.1f} GB/s = {alg / per_step / 1e3 / 8000 * 100:.2f} % "
                     f"of 8 TB/s{own}")
        if three:
            fl = [(r["end"] - r["start"]) / 1000.0 for g in sel for r in g if r["name"] == FLOCK]
            if fl:
                fb = FLOCK_BYTES_PER_AGENT * a.agents
                lines.append(f"- k_flock: mean {statistics.mean(fl):.1f} us per launch; its own "
                             f"{FLOCK_BYTES_PER_AGENT} B x {a.agents} agents -> "
                             f"{fb / statistics.mean(fl) / 1e3:.1f} GB/s")
                lines.append(f"flock_us: {statistics.mean(fl):.1f}")
        if len(groups) > len(sel) and not three:
            allspans = [(max(r["end"] for r in g) - min(r["start"] for r in g)) / 1000.0
                        for g in groups[1:]]
            lines.append(f"- every step of the trace after the first ({len(allspans)} steps): "
                         f"mean span {statistics.mean(allspans):.1f} us")
        lines.append(f"step_span_us: {per_step:.1f}")
    if a.fetch and a.write:
        f, fn = pmc_totals(a.fetch, "FETCH_SIZE")
        w, wn = pmc_totals(a.write, "WRITE_SIZE")
        cf = [k for k in f if k[0] == COPY and k[1] == max(g for (nm, g) in f if nm == COPY)]
        cw = [k for k in w if k[0] == COPY and k[1] == max(g for (nm, g) in w if nm == COPY)]
        kf = (1 << 30) / (f[cf[0]] / fn[cf[0]] * 1024)
        kw = (1 << 30) / (w[cw[0]] / wn[cw[0]] * 1024)
        lines.append("\nPMC byte counters, KiB per dispatch (mean):\n")
        lines.append("| kernel | grid | FETCH_SIZE KiB | WRITE_SIZE KiB |")
        lines.append("|---|---|---|---|")
        for k in sorted(set(f) | set(w)):
            fv = f.get(k, 0.0) / max(fn.get(k, 0), 1)
            wv = w.get(k, 0.0) / max(wn.get(k, 0), 1)
            if fv < 1 and wv < 1:
                continue
            lines.append(f"| {k[0]} | {k[1]} | {fv:.1f} | {wv:.1f} |")
        lines.append(f"\nCalibration: the 1 GiB copy reads FETCH_SIZE = {f[cf[0]] / fn[cf[0]]:.0f} KiB and "
                     f"WRITE_SIZE = {w[cw[0]] / wn[cw[0]]:.0f} KiB, so bytes = FETCH_SIZE x {kf:.3f} x 1024 "
                     f"and WRITE_SIZE x {kw:.3f} x 1024.\n")
        alg = BYTES_PER_AGENT * a.agents
        if three:
            # each step kernel's mean per dispatch times its dispatches per
            # step in the trace's steps (a kernel some steps leave out
            # counts in proportion)
            launches = collections.Counter((r["name"], r["grid"]) for g in sel for r in g)
            fetch = write = 0.0
            for k, cnt in launches.items():
                if k in f and fn[k]:
                    fetch += f[k] / fn[k] * 1024 * kf * cnt / len(sel)
                if k in w and wn[k]:
                    write += w[k] / wn[k] * 1024 * kw * cnt / len(sel)
            what = "three-kernel step (its kernels' dispatches per step)"
        else:
            # per step: all k_rt_step dispatches of a counter pass over the
            # steps that pass ran (its own launches, grouped as the trace's)
            ktot_f = sum(v for k, v in f.items() if k[0] == KERNEL)
            ktot_w = sum(v for k, v in w.items() if k[0] == KERNEL)
            nsteps_f = len(steps_of(load_trace(a.fetch)))
            fetch = ktot_f / nsteps_f * 1024 * kf
            write = ktot_w / nsteps_f * 1024 * kw
            what = f"k_rt_step, all launches of the step, {nsteps_f} steps"
        lines.append(f"HBM traffic per schedule step ({what}): fetch {fetch / 1e6:.2f} MB + "
                     f"write {write / 1e6:.2f} MB = "
                     f"{(fetch + write) / 1e6:.2f} MB ({(fetch + write) / alg:.2f}x the algorithmic "
                     f"{alg / 1e6:.2f} MB)")
        lines.append(f"traffic_bytes_per_step: {int(fetch + write)}")
    if a.counters:
        lines.extend(counter_table(a.counters, kname))
    if three:
        lines.extend(per_step_table(tr, a.step_kernel, a.skip, a.agents,
                                    getattr(a, "step_grid", 0) or 0, a.count))
    return "\n".join(lines) + "\n"


def per_step_table(tr, step_kernel, skip, agents, grid=0, count=0):
    """For the three-kernel path: each kernel's time summed per step over the
    steps (three_kernel_steps) from number `skip`, `count` of them (0: all)."""
    groups = three_kernel_steps(tr, step_kernel, grid)[skip:]
    if count:
        groups = groups[:count]
    if not groups:
        return []
    tot = collections.defaultdict(float)
    for g in groups:
        for r in g:
            tot[r["name"]] += (r["end"] - r["start"]) / 1000.0
    n = len(groups)
    out = [f"\nKernel time per schedule step ({step_kernel} launches from number {skip}: "
           f"{n} steps; the kernels run one after another, so the times add up):\n",
           "| kernel | us per step |", "|---|---|"]
    for k in sorted(tot, key=lambda k: -tot[k]):
        out.append(f"| {k} | {tot[k] / n:.1f} |")
    step_us = sum(tot.values()) / n
    out.append(f"\n- sum: {step_us:.1f} us per step; {BYTES_PER_AGENT} B x {agents} agents -> "
               f"{BYTES_PER_AGENT * agents / step_us / 1e3:.1f} GB/s = "
               f"{BYTES_PER_AGENT * agents / step_us / 1e3 / 8000 * 100:.2f} % of 8 TB/s")
    out.append(f"kernel_us_per_step: {step_us:.1f}")
    return out


def counter_table(path, kernel=KERNEL):
    """Per launch class (grid size) of `kernel`: the mean of every counter
    per dispatch, plus the derived shares the step kernel is judged by."""
    vals = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(set)
    for r in csv.DictReader(open(path)):
        if _name(r) != kernel:
            continue
        k = _grid(r)
        vals[k][r["Counter_Name"]] += float(r["Counter_Value"])
        disp[k].add(r.get("Dispatch_Id") or r.get("Correlation_Id") or len(disp[k]))
    if not vals:
        return []
    names = sorted({c for v in vals.values() for c in v})
    short = kernel.split("::")[-1]
    out = [f"\nHardware counters per {short} dispatch (mean over dispatches of each grid):\n",
           "| grid | dispatches | " + " | ".join(names) + " |",
           "|---|---|" + "---|" * len(names)]
    for k in sorted(vals, key=lambda k: -k):
        nd = max(len(disp[k]), 1)
        out.append(f"| {k} | {nd} | " + " | ".join(f"{vals[k][c] / nd:.4g}" for c in names) + " |")
    big = max(vals, key=lambda k: k)
    v = vals[big]
    if v.get("SQ_WAVE_CYCLES"):
        wc = v["SQ_WAVE_CYCLES"]
        out.append("")
        if "SQ_WAIT_ANY" in v:
            out.append(f"- largest launches: waves wait on memory or dependencies "
                       f"{v['SQ_WAIT_ANY'] / wc * 100:.1f} % of their cycles (SQ_WAIT_ANY / SQ_WAVE_CYCLES)")
        if "SQ_ACTIVE_INST_VALU" in v:
            out.append(f"- VALU busy {v['SQ_ACTIVE_INST_VALU'] / wc * 100:.1f} % of wave cycles "
                       f"(SQ_ACTIVE_INST_VALU / SQ_WAVE_CYCLES)")
        if "SQ_ACTIVE_INST_LDS" in v:
            out.append(f"- LDS instructions issuing {v['SQ_ACTIVE_INST_LDS'] / wc * 100:.1f} % of wave "
                       f"cycles")
    if v.get("SQ_WAVES") and v.get("SQ_INSTS_VALU"):
        out.append(f"- {v['SQ_INSTS_VALU'] / v['SQ_WAVES']:.0f} VALU instructions per wave"
                   + (f", {v['SQ_INSTS_SALU'] / v['SQ_WAVES']:.0f} SALU" if v.get("SQ_INSTS_SALU")
                      else "")
                   + (f", {v['SQ_INSTS_LDS'] / v['SQ_WAVES']:.0f} LDS" if v.get("SQ_INSTS_LDS")
                      else ""))
    if v.get("SQ_LDS_BANK_CONFLICT") is not None and v.get("SQ_INSTS_LDS"):
        out.append(f"- LDS bank-conflict cycles per LDS instruction: "
                   f"{v['SQ_LDS_BANK_CONFLICT'] / v['SQ_INSTS_LDS']:.3f}")
    return out


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--trace", required=True)
    p.add_argument("--fetch")
    p.add_argument("--write")
    p.add_argument("--counters", help="a rocprofv3 --pmc counter CSV with SQ_* counters")
    p.add_argument("--step-kernel", help="three-kernel path: the kernel that starts each step")
    p.add_argument("--step-grid", type=int, default=0,
                   help="only launches of --step-kernel with this grid size (threads)")
    p.add_argument("--agents", type=int, required=True)
    p.add_argument("--skip", type=int, default=0, help="steps to skip (warm-up)")
    p.add_argument("--count", type=int, default=0, help="steps to summarise (0: all)")
    p.add_argument("--title", required=True)
    p.add_argument("--cmd", required=True, help="the profiled command, for the header")
    p.add_argument("--pmc-cmd", help="the command of the counter passes, if different")
    print(summarize(p.parse_args()), end="")


if __name__ == "__main__":
    main()
