"""Digest of rocprofv3 CSV output for one kernel (kernel trace and PMC runs).

    python -m gpufs_amd.profsum trace   <dir> <kernel-substring> [--skip N]
    python -m gpufs_amd.profsum counter <dir> <kernel-substring> [--skip N]

`trace` reads <dir>/*_kernel_trace.csv and prints every dispatch of the
kernel with its duration, then the mean over the dispatches after the first
N (the cold or cache-filling ones).  `counter` reads
<dir>/*_counter_collection.csv, sums each counter over a dispatch's rows
(one per XCD / TCC instance), and scales it by what the same counter reads
for the device-to-device copy kernel in the same run (torch's 1 GiB copy,
`__amd_rocclr_copyBuffer`): the factor is the copy's known bytes over the
counter's reading, so a kernel's scaled value is in true bytes.
"""
import argparse
import csv
import glob
import os
import sys

COPY_KERNEL = "__amd_rocclr_copyBuffer"
COPY_BYTES = 1 << 30


def _one(pattern):
    files = sorted(glob.glob(pattern))
    if not files:
        raise SystemExit(f"no file matches {pattern}")
    return files[0]


def trace(d, kernel, skip):
    rows = list(csv.DictReader(open(_one(os.path.join(d, "*kernel_trace.csv")))))
    ks = [r for r in rows if kernel in r["Kernel_Name"]]
    ks.sort(key=lambda r: int(r["Start_Timestamp"]))
    out = []
    for i, r in enumerate(ks):
        ns = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        out.append(ns)
        print(f"dispatch {r['Dispatch_Id']:>5}  {ns / 1e6:10.4f} ms{'  (skipped)' if i < skip else ''}")
    warm = out[skip:]
    if warm:
        print(f"{kernel}: {len(warm)} dispatches after the first {skip}: mean {sum(warm) / len(warm) / 1e6:.4f} ms, "
              f"min {min(warm) / 1e6:.4f} ms, max {max(warm) / 1e6:.4f} ms")
    return warm


def counter(d, kernel, skip):
    rows = list(csv.DictReader(open(_one(os.path.join(d, "*counter_collection.csv")))))
    per = {}
    names = {}
    for r in rows:
        key = (int(r["Dispatch_Id"]), r["Counter_Name"])
        per[key] = per.get(key, 0.0) + float(r["Counter_Value"])
        names[int(r["Dispatch_Id"])] = r["Kernel_Name"]
    counters = sorted({c for _, c in per})
    res = {}
    for c in counters:
        copies = [v for (di, cn), v in per.items() if cn == c and names[di].startswith(COPY_KERNEL) and v > 1e5]
        if not copies:
            print(f"{c}: no {COPY_KERNEL} dispatch to calibrate against", file=sys.stderr)
            continue
        cal = sum(copies) / len(copies)   # KiB read/written by a 1 GiB copy
        factor = (COPY_BYTES / 1024) / cal
        print(f"{c}: {COPY_KERNEL} (1 GiB) reads {cal:.0f} KiB per dispatch over {len(copies)} dispatches -> "
              f"factor {factor:.4f}")
        ds = sorted(di for (di, cn) in per if cn == c and kernel in names[di])
        vals = []
        for i, di in enumerate(ds):
            v = per[(di, c)] * 1024 * factor
            vals.append(v)
            print(f"  dispatch {di:>5} {kernel}: {per[(di, c)]:.0f} KiB raw = {v / 2**30:.4f} GiB scaled"
                  f"{'  (skipped)' if i < skip else ''}")
        warm = vals[skip:]
        if warm:
            res[c] = sum(warm) / len(warm)
            print(f"  {c} per launch after the first {skip}: {res[c]:.0f} bytes ({res[c] / 2**30:.4f} GiB)")
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["trace", "counter"])
    ap.add_argument("dir")
    ap.add_argument("kernel")
    ap.add_argument("--skip", type=int, default=0)
    a = ap.parse_args(argv)
    (trace if a.mode == "trace" else counter)(a.dir, a.kernel, a.skip)


if __name__ == "__main__":
    main()
