"""Summaries of rocprofv3 output for the profiles kept with this package.

  python -m birdwing_amd.rocprof_summary stats DIR        kernel-trace --stats table (markdown)
  python -m birdwing_amd.rocprof_summary pmc DIR          per-kernel average of each PMC counter
  python -m birdwing_amd.rocprof_summary last DIR K       the last K kernel dispatches of a kernel
        trace in start order (the runtime's own copy and fill kernels left out): duration,
        gap since the previous one ended, grid
  python -m birdwing_amd.rocprof_summary traffic FETCH_DIR WRITE_DIR CALIB_MIB LABEL=REGEX ...
        HBM bytes per dispatch of each labelled kernel (kernels matching REGEX), from a
        FETCH_SIZE run and a WRITE_SIZE run (KiB per dispatch), each scaled by what the
        same counter reads for the calibration kernel (one elementwise kernel reading and
        writing CALIB_MIB MiB) recorded in the same runs (JSON).
"""
import csv
import glob
import json
import os
import re
import sys


def _rows(d, pattern):
    files = sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True))
    if not files:
        raise SystemExit("no %s under %s" % (pattern, d))
    out = []
    for f in files:
        with open(f, newline="") as fh:
            out.extend(csv.DictReader(fh))
    return out


def _get(row, *names):
    low = {k.lower(): v for k, v in row.items()}
    for n in names:
        if n.lower() in low:
            return low[n.lower()]
    raise KeyError(names)


def short(name, width=90):
    name = re.sub(r"\s+", " ", name)
    return name if len(name) <= width else name[:width - 3] + "..."


def stats(d):
    rows = _rows(d, "*kernel_stats.csv")
    lines = ["| kernel | calls | total ms | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for r in rows:
        lines.append("| %s | %s | %.3f | %.1f | %.1f | %.1f | %.1f |" % (
            short(_get(r, "Name", "KernelName")), _get(r, "Calls"),
            float(_get(r, "TotalDurationNs")) / 1e6, float(_get(r, "AverageNs")) / 1e3,
            float(_get(r, "MinNs")) / 1e3, float(_get(r, "MaxNs")) / 1e3, float(_get(r, "Percentage"))))
    return "\n".join(lines)


def pmc_table(d):
    rows = _rows(d, "*counter_collection.csv")
    acc = {}
    for r in rows:
        key = (_get(r, "Kernel_Name", "KernelName"), _get(r, "Counter_Name"))
        v = float(_get(r, "Counter_Value"))
        disp = _get(r, "Dispatch_Id", "Dispatch_ID")
        a = acc.setdefault(key, {})
        a[disp] = a.get(disp, 0.0) + v        # sum over agents / dimensions of one dispatch
    out = {}
    for (k, c), per in acc.items():
        out.setdefault(k, {})[c] = (sum(per.values()) / len(per), len(per))
    return out


def pmc(d):
    t = pmc_table(d)
    lines = ["| kernel | counter | mean per dispatch | dispatches |", "|---|---|---|---|"]
    for k in sorted(t):
        for c, (m, n) in sorted(t[k].items()):
            lines.append("| %s | %s | %.1f | %d |" % (short(k), c, m, n))
    return "\n".join(lines)


def _match(t, regex, counter):
    vals = [(m, n) for k, cs in t.items() if re.search(regex, k) for c, (m, n) in cs.items() if c == counter]
    if not vals:
        raise SystemExit("no %s for kernels matching %r" % (counter, regex))
    tot = sum(m * n for m, n in vals)
    cnt = sum(n for _, n in vals)
    return tot / cnt, cnt


def traffic(fetch_dir, write_dir, calib_mib, pairs):
    """pairs: [(label, regex)]; HBM bytes per dispatch of each labelled kernel."""
    tf, tw = pmc_table(fetch_dir), pmc_table(write_dir)
    calib = r"bitwise_not"
    cf, _ = _match(tf, calib, "FETCH_SIZE")
    cw, _ = _match(tw, calib, "WRITE_SIZE")
    kib = float(calib_mib) * 1024.0
    ff, fw = kib / cf, kib / cw           # true KiB per counted KiB
    out = {"calibration": {"kernel": "one elementwise kernel reading and writing %g MiB" % calib_mib,
                           "fetch_kib": cf, "write_kib": cw, "fetch_factor": ff, "write_factor": fw},
           "kernels": {}}
    for label, regex in pairs:
        f, nf = _match(tf, regex, "FETCH_SIZE")
        w, nw = _match(tw, regex, "WRITE_SIZE")
        out["kernels"][label] = {"regex": regex, "fetch_kib_per_dispatch": f, "write_kib_per_dispatch": w,
                                 "dispatches": [nf, nw], "hbm_bytes_per_launch": (f * ff + w * fw) * 1024.0}
    return out


def last(d, k):
    rows = [r for r in _rows(d, "*kernel_trace.csv") if not _get(r, "Kernel_Name").startswith("__amd_rocclr")]
    rows.sort(key=lambda r: int(_get(r, "Start_Timestamp")))
    rows = rows[-k:]
    lines = ["| # | kernel | us | gap us | grid |", "|---|---|---|---|---|"]
    prev = None
    for i, r in enumerate(rows):
        s, e = int(_get(r, "Start_Timestamp")), int(_get(r, "End_Timestamp"))
        gap = (s - prev) / 1e3 if prev is not None else 0.0
        prev = e
        grid = "%s x %s x %s / %s" % (_get(r, "Grid_Size_X"), _get(r, "Grid_Size_Y"), _get(r, "Grid_Size_Z"),
                                      _get(r, "Workgroup_Size_X"))
        lines.append("| %d | %s | %.1f | %.1f | %s |" % (i, short(_get(r, "Kernel_Name"), 60), (e - s) / 1e3, gap, grid))
    span = (int(_get(rows[-1], "End_Timestamp")) - int(_get(rows[0], "Start_Timestamp"))) / 1e3
    busy = sum(int(_get(r, "End_Timestamp")) - int(_get(r, "Start_Timestamp")) for r in rows) / 1e3
    lines.append("")
    lines.append("%d dispatches: %.1f us from the first start to the last end, %.1f us inside kernels" %
                 (len(rows), span, busy))
    return "\n".join(lines)


def main(argv):
    if len(argv) >= 2 and argv[0] == "stats":
        print(stats(argv[1]))
    elif len(argv) >= 2 and argv[0] == "pmc":
        print(pmc(argv[1]))
    elif len(argv) >= 3 and argv[0] == "last":
        print(last(argv[1], int(argv[2])))
    elif len(argv) >= 5 and argv[0] == "traffic":
        pairs = [a.rsplit("=", 1) for a in argv[4:]]   # labels may hold "=", regexes do not
        print(json.dumps(traffic(argv[1], argv[2], float(argv[3]), pairs), indent=1))
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
