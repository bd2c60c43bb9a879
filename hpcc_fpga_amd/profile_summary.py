"""Summaries of rocprofv3 output for the profiles/ folder.

    python -m hpcc_fpga_amd.profile_summary traffic --fetch DIR --write DIR --bench JSON
        [--bench-key SECTION] --kernels PAT[,PAT] [--min-grid G] --copy-bytes N
        --out profiles/X.json --md profiles/X.md
    python -m hpcc_fpga_amd.profile_summary mfma --pmc DIR --kernels PAT[,PAT] [--min-grid G] --md M
    python -m hpcc_fpga_amd.profile_summary trace --trace CSV --md M [--kernels PAT --flops F]
    python -m hpcc_fpga_amd.profile_summary bandwidth --fetch DIR --write DIR --copy-bytes N
        --kernels PAT[,PAT] --md M

DIR is a rocprofv3 ``--output-format csv -d DIR -o run`` folder; counters are
per dispatch.  FETCH_SIZE and WRITE_SIZE (KiB) are calibrated against the
largest ``__amd_rocclr_copyBuffer`` dispatch of the same run, a device copy of
known size (--copy-bytes, read once and written once): the scale factors turn
the counters into HBM bytes.  Every markdown file starts with the commands
given in --command (repeatable), so the summary names how it was made.
"""
import argparse
import collections
import csv
import json
import os


def _rows(d, name):
    with open(os.path.join(d, name)) as f:
        return list(csv.DictReader(f))


def _match(name, pats):
    return any(p in name for p in pats)


def _counter(d, counter):
    """{dispatch id: (kernel, value, duration ns, grid size)} for one counter."""
    out = {}
    for r in _rows(d, "run_counter_collection.csv"):
        if r["Counter_Name"] != counter:
            continue
        did = int(r["Dispatch_Id"])
        dur = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        out[did] = (r["Kernel_Name"], float(r["Counter_Value"]), dur, int(r["Grid_Size"]))
    return out


def _scale(rows, counter, copy_bytes):
    copies = [v for v in rows.values() if "copyBuffer" in v[0]]
    if not copies:
        raise SystemExit(f"no copyBuffer dispatch to calibrate {counter}")
    kib = max(v[1] for v in copies)
    return copy_bytes / (kib * 1024.0), kib


def _header(title, commands, lines):
    out = [f"# {title}", "", "Commands (repo root, TMPDIR=/tmp):", ""]
    out += [f"    {c}" for c in commands]
    out += [""] + lines
    return "\n".join(out) + "\n"


def _select(rows, pats, min_grid):
    return {k: v for k, v in rows.items() if _match(v[0], pats) and v[3] >= min_grid}


def cmd_traffic(a):
    pats = a.kernels.split(",")
    f = _counter(a.fetch, "FETCH_SIZE")
    w = _counter(a.write, "WRITE_SIZE")
    fs, fkib = _scale(f, "FETCH_SIZE", a.copy_bytes)
    ws, wkib = _scale(w, "WRITE_SIZE", a.copy_bytes)
    fk = _select(f, pats, a.min_grid)
    wk = _select(w, pats, a.min_grid)
    n = len(fk)
    fetch = sum(v[1] for v in fk.values()) * 1024.0 * fs / max(1, n)
    write = sum(v[1] for v in wk.values()) * 1024.0 * ws / max(1, len(wk))
    dur = sum(v[2] for v in fk.values()) / max(1, n)
    with open(a.bench) as fh:
        line = [ln for ln in fh if ln.strip().startswith("{")][-1]
    bench = json.loads(line)
    rf = bench[a.bench_key]["roofline"] if a.bench_key else bench["roofline"]
    alg = rf["algorithmic_bytes_per_launch"]
    res = {"kernels": pats, "launches_counted": n, "launches_expected": rf["launches"],
           "fetch_scale": fs, "write_scale": ws, "copy_fetch_kib": fkib, "copy_write_kib": wkib,
           "copy_bytes": a.copy_bytes, "hbm_bytes_per_launch": fetch + write,
           "fetch_bytes_per_launch": fetch, "write_bytes_per_launch": write,
           "algorithmic_bytes_per_launch": alg,
           "algorithmic_flops_per_launch": rf["algorithmic_flops_per_launch"],
           "avg_kernel_ns_counters": dur, "shapes_match": n == rf["launches"]}
    with open(a.out, "w") as fh:
        json.dump(res, fh, indent=1)
        fh.write("\n")
    lines = ["| quantity | value |", "|---|---|",
             f"| copy dispatch: bytes / FETCH_SIZE KiB / WRITE_SIZE KiB | {a.copy_bytes} / "
             f"{fkib:.1f} / {wkib:.1f} |",
             f"| scale factors fetch / write | {fs:.4f} / {ws:.4f} |",
             f"| update dispatches counted / expected | {n} / {rf['launches']} |",
             f"| HBM read per launch (scaled) | {fetch / 1e6:.1f} MB |",
             f"| HBM write per launch (scaled) | {write / 1e6:.1f} MB |",
             f"| algorithmic bytes per launch | {alg / 1e6:.1f} MB |",
             f"| traffic / algorithmic | {(fetch + write) / alg:.3f} |",
             f"| average dispatch (counter run, serialised) | {dur / 1e3:.1f} us |"]
    with open(a.md, "w") as fh:
        fh.write(_header(a.title, a.command, lines))


def cmd_mfma(a):
    pats = a.kernels.split(",")
    m = _select(_counter(a.pmc, "SQ_VALU_MFMA_BUSY_CYCLES"), pats, a.min_grid)
    g = _select(_counter(a.pmc, "GRBM_GUI_ACTIVE"), pats, a.min_grid)
    busy = sum(v[1] for v in m.values())
    act = sum(v[1] for v in g.values())
    lines = [f"MFMA utilisation = SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE / {a.xcds} x "
             f"{a.simds} SIMDs) over the {len(m)} matching dispatches (GRBM_GUI_ACTIVE is summed "
             f"over the {a.xcds} XCDs).", "",
             "| quantity | value |", "|---|---|",
             f"| SQ_VALU_MFMA_BUSY_CYCLES | {busy:.4e} |",
             f"| GRBM_GUI_ACTIVE | {act:.4e} |",
             f"| MFMA utilisation | {100.0 * busy / (act / a.xcds * a.simds):.1f} % |"]
    for extra in ("SQ_LDS_BANK_CONFLICT", "SQ_LDS_IDX_ACTIVE", "SQ_BUSY_CU_CYCLES"):
        e = _select(_counter(a.pmc, extra), pats, a.min_grid)
        if e:
            lines.append(f"| {extra} | {sum(v[1] for v in e.values()):.4e} |")
    with open(a.md, "w") as fh:
        fh.write(_header(a.title, a.command, lines))


def _union(iv):
    tot, cur = 0, None
    for s, e in sorted(iv):
        if cur is None:
            cur = [s, e]
        elif s <= cur[1]:
            cur[1] = max(cur[1], e)
        else:
            tot += cur[1] - cur[0]
            cur = [s, e]
    if cur:
        tot += cur[1] - cur[0]
    return tot


def cmd_trace(a):
    rows = _rows(os.path.dirname(a.trace) or ".", os.path.basename(a.trace))
    agg = collections.defaultdict(lambda: [0, 0])
    ev = []
    for r in rows:
        s, e = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        agg[r["Kernel_Name"]][0] += 1
        agg[r["Kernel_Name"]][1] += e - s
        ev.append((r["Kernel_Name"], s, e, int(r["Grid_Size_X"]) if "Grid_Size_X" in r else 0))
    total = sum(v[1] for v in agg.values())
    lines = ["| kernel | calls | total ms | avg us | % |", "|---|---|---|---|---|"]
    for k, (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:a.top]:
        name = k if len(k) <= 90 else k[:87] + "..."
        lines.append(f"| `{name}` | {c} | {t / 1e6:.2f} | {t / c / 1e3:.2f} | "
                     f"{100.0 * t / total:.2f} |")
    if a.kernels:
        pats = a.kernels.split(",")
        sel = [(s, e) for k, s, e, g in ev if _match(k, pats) and g >= a.min_grid]
        summed = sum(e - s for s, e in sel)
        un = _union(sel)
        lines += ["", f"Update launches matching {pats} (grid >= {a.min_grid} work items): "
                  f"{len(sel)}, average {summed / max(1, len(sel)) / 1e3:.1f} us; union of their "
                  f"intervals {un / 1e6:.2f} ms against {summed / 1e6:.2f} ms summed."]
        if a.flops:
            lines.append(f"{a.flops / 1e12:.2f} TFLOP of updates over the union: "
                         f"{a.flops / (un * 1e-9) / 1e12:.1f} TFLOP/s; per launch (summed "
                         f"durations): {a.flops / (summed * 1e-9) / 1e12:.1f} TFLOP/s.")
    with open(a.md, "w") as fh:
        fh.write(_header(a.title, a.command, lines))


def cmd_bandwidth(a):
    pats = a.kernels.split(",")
    f = _counter(a.fetch, "FETCH_SIZE")
    w = _counter(a.write, "WRITE_SIZE")
    fs, _ = _scale(f, "FETCH_SIZE", a.copy_bytes)
    ws, _ = _scale(w, "WRITE_SIZE", a.copy_bytes)
    lines = ["Per kernel: HBM bytes (scaled FETCH_SIZE + WRITE_SIZE) over the dispatch time of "
             "the counter run (kernels serialised by the profiler).", "",
             "| kernel | dispatches | avg us | read MB | write MB | HBM GB/s |",
             "|---|---|---|---|---|---|"]
    for p in pats:
        fk = {k: v for k, v in f.items() if p in v[0]}
        wk = {k: v for k, v in w.items() if p in v[0]}
        if not fk:
            continue
        n = len(fk)
        rd = sum(v[1] for v in fk.values()) * 1024.0 * fs / n
        wr = sum(v[1] for v in wk.values()) * 1024.0 * ws / max(1, len(wk))
        dur = sum(v[2] for v in fk.values()) / n
        lines.append(f"| `{p}` | {n} | {dur / 1e3:.1f} | {rd / 1e6:.2f} | {wr / 1e6:.2f} | "
                     f"{(rd + wr) / dur:.1f} |")
    with open(a.md, "w") as fh:
        fh.write(_header(a.title, a.command, lines))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    for name in ("traffic", "mfma", "trace", "bandwidth"):
        p = sub.add_parser(name)
        p.add_argument("--md", required=True)
        p.add_argument("--title", required=True)
        p.add_argument("--command", action="append", default=[])
        p.add_argument("--kernels", default="")
        p.add_argument("--min-grid", type=int, default=0)
    sub.choices["traffic"].add_argument("--fetch", required=True)
    sub.choices["traffic"].add_argument("--write", required=True)
    sub.choices["traffic"].add_argument("--bench", required=True)
    sub.choices["traffic"].add_argument("--bench-key", default="",
                                        help="section of the bench JSON holding the roofline")
    sub.choices["traffic"].add_argument("--copy-bytes", type=float, required=True)
    sub.choices["traffic"].add_argument("--out", required=True)
    sub.choices["mfma"].add_argument("--pmc", required=True)
    sub.choices["mfma"].add_argument("--xcds", type=int, default=8)
    sub.choices["mfma"].add_argument("--simds", type=int, default=1024)
    sub.choices["trace"].add_argument("--trace", required=True)
    sub.choices["trace"].add_argument("--top", type=int, default=18)
    sub.choices["trace"].add_argument("--flops", type=float, default=0.0)
    sub.choices["bandwidth"].add_argument("--fetch", required=True)
    sub.choices["bandwidth"].add_argument("--write", required=True)
    sub.choices["bandwidth"].add_argument("--copy-bytes", type=float, required=True)
    a = ap.parse_args(argv)
    {"traffic": cmd_traffic, "mfma": cmd_mfma, "trace": cmd_trace,
     "bandwidth": cmd_bandwidth}[a.cmd](a)


if __name__ == "__main__":
    main()
