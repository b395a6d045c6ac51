"""Summaries of rocprofv3 output for the step's kernels, restricted to a timed window.

    python -m physecs_amd.profile_summary --dir <rocprofv3 -d folder> --window-steps K \
        [--marker k_scene_bounds_init] [--fetch-scale F --write-scale W] --out summary.md

A step of the library starts with one k_scene_bounds_init launch (the broadphase's scene
bounds), so the last K steps of a run are the dispatches from the K-th last
k_scene_bounds_init on. The summary gives, over the whole run and over that window, each
kernel's dispatch count, total and average duration from the kernel trace
(*kernel_trace.csv), and, where a counter run's *counter_collection.csv is in the folder,
each counter's mean per dispatch over the window. FETCH_SIZE and WRITE_SIZE (KiB per
dispatch) are also given in bytes after scaling by the factors a copy of known size reads
(--fetch-scale, --write-scale: true bytes per counted byte). Where a run collected the SQ
counters for them, each kernel also gets, over the window:
- occupancy: mean resident waves per CU while the kernel ran, 4 * SQ_WAVE_CYCLES (quad
  cycles) / (GRBM_GUI_ACTIVE / 8, the counter is summed over the 8 XCDs) / 256 CUs, and
  that as a share of the 32 wave slots of a CU;
- LDS bank conflicts per LDS access cycle, SQ_LDS_BANK_CONFLICT / (SQ_LDS_IDX_ACTIVE -
  SQ_LDS_BANK_CONFLICT);
- the share of wave time spent waiting to issue, SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES.
"""
import argparse
import collections
import csv
import glob
import os
import sys


def _find(folder, suffix):
    hits = sorted(glob.glob(os.path.join(folder, "**", "*" + suffix), recursive=True))
    return hits[-1] if hits else None


def _short(name):
    name = name.split("(")[0]
    if name.startswith("void "):
        name = name[5:]
    if "rocprim" in name:
        for part in ("merge_sort_block_merge", "merge_sort_block_sort", "radix_sort_block_sort", "onesweep",
                     "transform", "radix_sort", "scan", "reduce"):
            if part in name:
                return "rocprim " + part
    return name


def read_trace(path):
    rows = []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Dispatch_Id"]), _short(r["Kernel_Name"]), int(r["Start_Timestamp"]),
                         int(r["End_Timestamp"])))
    rows.sort(key=lambda x: x[2])
    return rows


def window_start(rows, marker, steps):
    idx = [i for i, r in enumerate(rows) if r[1] == marker]
    if len(idx) < steps:
        raise SystemExit(f"only {len(idx)} dispatches of {marker}, fewer than {steps} steps")
    return idx[-steps]


def table(rows, steps):
    agg = collections.OrderedDict()
    for _, name, t0, t1 in rows:
        a = agg.setdefault(name, [0, 0])
        a[0] += 1
        a[1] += t1 - t0
    total = sum(a[1] for a in agg.values())
    out = ["| kernel | calls | calls/step | total (us) | average (us) | share |", "|---|---:|---:|---:|---:|---:|"]
    for name, (n, ns) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        out.append(f"| `{name}` | {n} | {n / steps:.1f} | {ns / 1e3:.1f} | {ns / n / 1e3:.2f} | "
                   f"{100.0 * ns / max(total, 1):.1f}% |")
    return out, total, sum(a[0] for a in agg.values())


def read_counters(path, first_dispatch):
    """{kernel: {counter: [values per dispatch]}} over dispatches >= first_dispatch."""
    per = collections.defaultdict(lambda: collections.defaultdict(dict))
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            d = int(r["Dispatch_Id"])
            if d < first_dispatch:
                continue
            per[_short(r["Kernel_Name"])][r["Counter_Name"]][d] = float(r["Counter_Value"])
    return per


CUS, XCDS, WAVE_SLOTS = 256, 8, 32


def derived(per):
    """Table lines of occupancy, LDS conflict rate and issue-wait share per kernel."""
    out = []
    for kern in sorted(per):
        c = {name: sum(v.values()) for name, v in per[kern].items()}
        n = max(len(v) for v in per[kern].values())
        occ = lds = wait = ""
        if c.get("SQ_WAVE_CYCLES") and c.get("GRBM_GUI_ACTIVE"):
            waves = 4.0 * c["SQ_WAVE_CYCLES"] / (c["GRBM_GUI_ACTIVE"] / XCDS) / CUS
            occ = f"{waves:.2f} ({100.0 * waves / WAVE_SLOTS:.1f}%)"
        if "SQ_LDS_BANK_CONFLICT" in c and c.get("SQ_LDS_IDX_ACTIVE"):
            busy = c["SQ_LDS_IDX_ACTIVE"] - c["SQ_LDS_BANK_CONFLICT"]
            lds = f"{c['SQ_LDS_BANK_CONFLICT'] / busy:.3f}" if busy > 0 else "0"
        if c.get("SQ_WAIT_INST_ANY") is not None and c.get("SQ_WAVE_CYCLES"):
            wait = f"{100.0 * c['SQ_WAIT_INST_ANY'] / c['SQ_WAVE_CYCLES']:.1f}%"
        if occ or lds or wait:
            waves_per = c.get("SQ_WAVES", 0) / n
            out.append(f"| `{kern}` | {n} | {waves_per:.0f} | {occ} | {lds} | {wait} |")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--dir", required=True, help="rocprofv3 output folder (-d)")
    ap.add_argument("--window-steps", type=int, required=True)
    ap.add_argument("--marker", default="pgx::k_scene_bounds_init")
    ap.add_argument("--title", default="Kernel trace")
    ap.add_argument("--command", default="")
    ap.add_argument("--window", default="", help="the window as the bench names it (config.window)")
    ap.add_argument("--fetch-scale", type=float, default=None)
    ap.add_argument("--write-scale", type=float, default=None)
    ap.add_argument("--traffic-kernel", action="append", default=[],
                    help="kernel whose scaled FETCH+WRITE bytes per launch become the traffic_bytes_per_launch "
                    "line; given more than once, the mean over the launches of all of them")
    ap.add_argument("--counters-dir", action="append", default=[], help="folders of counter runs")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)

    trace = _find(a.dir, "kernel_trace.csv")
    if not trace:
        raise SystemExit(f"no kernel_trace.csv under {a.dir}")
    rows = read_trace(trace)
    marker = a.marker
    if not any(r[1] == marker for r in rows):
        marker = _short(marker)
    w0 = window_start(rows, marker, a.window_steps)
    win = rows[w0:]
    first_dispatch = win[0][0]
    n_steps_all = sum(1 for r in rows if r[1] == marker)
    lines = [f"# {a.title}", ""]
    if a.command:
        lines += ["Command (from the repository root, on one MI355X):", "", "```", a.command, "```", ""]
    if a.window:
        lines += [f"window: {a.window}", ""]
    wt, wtotal, wcalls = table(win, a.window_steps)
    span = (win[-1][3] - win[0][2]) / 1e3
    lines += [f"window_steps: {a.window_steps}",
              f"dispatches_per_step: {wcalls / a.window_steps:.1f}",
              f"kernel_us_per_step: {wtotal / 1e3 / a.window_steps:.1f}",
              f"span_us_per_step: {span / a.window_steps:.1f}",
              f"first_window_dispatch: {first_dispatch}", ""]
    counter_lines = []
    derived_lines = []
    traffic = None
    fetch_vals, write_vals = [], []
    for cdir in a.counters_dir:
        cpath = _find(cdir, "counter_collection.csv")
        if not cpath:
            counter_lines.append(f"(no counter_collection.csv under {cdir})")
            continue
        # the counter run's own trace gives its window (dispatch ids differ between runs)
        ctrace = _find(cdir, "kernel_trace.csv")
        crows = read_trace(ctrace) if ctrace else None
        cfirst = crows[window_start(crows, marker, a.window_steps)][0] if crows else None
        if cfirst is None:
            # without a trace, the window is the dispatches after the K-th last marker by id
            ids = collections.defaultdict(list)
            with open(cpath, newline="") as f:
                for r in csv.DictReader(f):
                    if _short(r["Kernel_Name"]) == marker:
                        ids[int(r["Dispatch_Id"])] = 1
            ks = sorted(ids)
            cfirst = ks[-a.window_steps]
        per = read_counters(cpath, cfirst)
        derived_lines += derived(per)
        for kern in sorted(per):
            for cname, vals in sorted(per[kern].items()):
                v = list(vals.values())
                mean = sum(v) / len(v)
                extra = ""
                if cname.startswith("FETCH_SIZE") and a.fetch_scale:
                    extra = f" | {mean * 1024 * a.fetch_scale:.0f} B |"
                elif cname.startswith("WRITE_SIZE") and a.write_scale:
                    extra = f" | {mean * 1024 * a.write_scale:.0f} B |"
                counter_lines.append(f"| `{kern}` | {cname} | {len(v)} | {mean:.4g}{extra or ' | |'}")
        for name in a.traffic_kernel:
            # FETCH_SIZE and WRITE_SIZE come from separate runs: gather them over the runs
            tk = per.get(_short(name)) or per.get(name) or {}
            fetch_vals += [v for c, d in tk.items() if c.startswith("FETCH_SIZE") for v in d.values()]
            write_vals += [v for c, d in tk.items() if c.startswith("WRITE_SIZE") for v in d.values()]
    if fetch_vals and write_vals and a.fetch_scale and a.write_scale:
        traffic = (sum(fetch_vals) / len(fetch_vals) * 1024 * a.fetch_scale,
                   sum(write_vals) / len(write_vals) * 1024 * a.write_scale)
    if traffic:
        lines += [f"traffic_fetch_bytes_per_launch: {traffic[0]:.0f}",
                  f"traffic_write_bytes_per_launch: {traffic[1]:.0f}",
                  f"traffic_bytes_per_launch: {traffic[0] + traffic[1]:.0f}", ""]
    lines += [f"## Last {a.window_steps} steps (the timed window)", ""] + wt + [""]
    if counter_lines:
        lines += ["## Counters over the window (mean per dispatch)", "",
                  "| kernel | counter | dispatches | mean | scaled |", "|---|---|---:|---:|---:|"] + counter_lines + [""]
    if derived_lines:
        lines += ["## Occupancy, LDS bank conflicts and issue waits over the window", "",
                  "| kernel | dispatches | waves/dispatch | resident waves per CU (of 32) | LDS conflicts per access "
                  "| wave time waiting to issue |", "|---|---:|---:|---:|---:|---:|"] + derived_lines + [""]
    at, atotal, acalls = table(rows, max(1, n_steps_all))
    lines += [f"## Whole run ({n_steps_all} steps, {acalls} dispatches, {atotal / 1e3:.1f} us of kernels)", ""] + at[:32]
    with open(a.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {a.out}: window {a.window_steps} steps from dispatch {first_dispatch}, "
          f"{wcalls / a.window_steps:.1f} dispatches/step, {wtotal / 1e3 / a.window_steps:.1f} us kernels/step",
          file=sys.stderr)


if __name__ == "__main__":
    main()
