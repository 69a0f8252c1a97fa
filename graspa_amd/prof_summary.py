"""Markdown summaries of rocprofv3 CSV output (``--output-format csv``).

    python -m graspa_amd.prof_summary DIR [--kernel SUBSTRING] [--copy-bytes N]

DIR is a rocprofv3 output directory (searched recursively).  Printed:
- per kernel, dispatches and durations from ``*kernel_trace.csv``;
- per kernel, counter values averaged per dispatch from
  ``*counter_collection.csv`` (every counter the runs collected), and, when
  --copy-bytes gives the size of a plain device-to-device copy that ran in the
  same program (bench.py --copy-calibration), the FETCH_SIZE / WRITE_SIZE
  scale factors that copy implies.
Only the standard library is used.
"""
import argparse
import collections
import csv
import os
import sys


def _rows(root, suffix):
    for d, _, files in os.walk(root):
        for f in sorted(files):
            if f.endswith(suffix):
                with open(os.path.join(d, f), newline="") as fh:
                    yield from csv.DictReader(fh)


def _get(row, *names):
    for n in names:
        if n in row and row[n] != "":
            return row[n]
    return None


def kernel_durations(root):
    """{kernel name: [duration ns, ...]} from the kernel traces."""
    out = collections.defaultdict(list)
    for r in _rows(root, "kernel_trace.csv"):
        name = _get(r, "Kernel_Name", "kernel_name", "Name")
        t0 = _get(r, "Start_Timestamp", "start", "Start")
        t1 = _get(r, "End_Timestamp", "end", "End")
        if name and t0 and t1:
            out[name].append(int(t1) - int(t0))
    return out


def counters(root):
    """{kernel name: {counter: [value per dispatch, ...]}} (values of one
    dispatch summed over the instances the profiler reports)."""
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    names = {}
    for r in _rows(root, "counter_collection.csv"):
        name = _get(r, "Kernel_Name", "kernel_name")
        disp = _get(r, "Dispatch_Id", "dispatch_id", "Correlation_Id")
        cn = _get(r, "Counter_Name", "counter_name")
        cv = _get(r, "Counter_Value", "counter_value")
        if name is None or disp is None or cn is None or cv is None:
            continue
        per[(name, disp)][cn] += float(cv)
        names[(name, disp)] = name
    out = collections.defaultdict(lambda: collections.defaultdict(list))
    for key, vals in per.items():
        for cn, v in vals.items():
            out[names[key]][cn].append(v)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("root")
    ap.add_argument("--kernel", default=None, help="only kernels whose name contains this")
    ap.add_argument("--copy-bytes", type=int, default=0)
    a = ap.parse_args(argv)
    keep = (lambda n: a.kernel in n) if a.kernel else (lambda n: True)
    dur = kernel_durations(a.root)
    if dur:
        tot = sum(sum(v) for v in dur.values())
        print("| kernel | calls | total us | average us | % |\n|---|---|---|---|---|")
        for n, v in sorted(dur.items(), key=lambda kv: -sum(kv[1])):
            if keep(n):
                print("| %s | %d | %.1f | %.1f | %.2f |" % (n, len(v), sum(v) / 1e3, sum(v) / len(v) / 1e3,
                                                           100.0 * sum(v) / tot))
    cnt = counters(a.root)
    if cnt:
        cols = sorted({c for v in cnt.values() for c in v})
        print("\n| kernel | dispatches | " + " | ".join(cols) + " |")
        print("|---|---|" + "---|" * len(cols))
        for n, v in sorted(cnt.items()):
            if not keep(n) and "copyBuffer" not in n:
                continue
            nd = max(len(x) for x in v.values())
            cells = ["%.4g" % (sum(v[c]) / len(v[c])) if v.get(c) else "" for c in cols]
            print("| %s | %d | %s |" % (n, nd, " | ".join(cells)))
        if a.copy_bytes:
            for n, v in cnt.items():
                if "copyBuffer" in n:
                    kib = a.copy_bytes / 1024.0
                    for c in ("FETCH_SIZE", "WRITE_SIZE"):
                        if v.get(c):
                            print("\n%s of the %d-byte copy: %.1f KiB per dispatch (max), scale factor %.3f"
                                  % (c, a.copy_bytes, max(v[c]), kib / max(v[c])))
    if not dur and not cnt:
        print("no kernel_trace.csv or counter_collection.csv under %s" % a.root, file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
