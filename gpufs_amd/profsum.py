"""Digest of rocprofv3 CSV output for one kernel (kernel trace and PMC runs).

    python -m gpufs_amd.profsum trace   <dir> <kernel-substring> [--skip N | --last N]
    python -m gpufs_amd.profsum counter <dir> <kernel-substring> [--skip N]
    python -m gpufs_amd.profsum sq      <dir> <kernel-substring> [--skip N]

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


def trace(d, kernel, skip, last=0):
    rows = list(csv.DictReader(open(_one(os.path.join(d, "*kernel_trace.csv")))))
    ks = [r for r in rows if kernel in r["Kernel_Name"]]
    ks.sort(key=lambda r: int(r["Start_Timestamp"]))
    if last:   # only the last N dispatches count (the earlier ones are listed as skipped)
        skip = max(0, len(ks) - last)
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


def sq(d, kernel, skip, xcds=8, cus=256):
    """Shader counters of each dispatch (summed over instances), then the mean
    over the dispatches after the first `skip` with the ratios they give:
    elapsed cycles per XCD (GRBM_GUI_ACTIVE / xcds), busy CUs, VALU and LDS
    instructions per wave, waves' share of time waiting, and bank-conflict
    cycles over LDS-active cycles."""
    rows = list(csv.DictReader(open(_one(os.path.join(d, "*counter_collection.csv")))))
    per, order = {}, {}
    for r in rows:
        if kernel not in r["Kernel_Name"]:
            continue
        di = int(r["Dispatch_Id"])
        order[di] = int(r["Start_Timestamp"])
        per.setdefault(di, {})
        per[di][r["Counter_Name"]] = per[di].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    ds = sorted(per, key=lambda k: order[k])
    for i, di in enumerate(ds):
        vals = "  ".join(f"{c}={v:.0f}" for c, v in sorted(per[di].items()))
        print(f"dispatch {di:>5}{' (skipped)' if i < skip else ''}: {vals}")
    keep = ds[skip:]
    if not keep:
        return {}
    m = {c: sum(per[di].get(c, 0.0) for di in keep) / len(keep) for c in per[keep[0]]}
    print(f"mean over {len(keep)} dispatches after the first {skip}:")
    for c, v in sorted(m.items()):
        print(f"  {c} = {v:.0f}")
    g = m.get
    if g("GRBM_GUI_ACTIVE"):
        el = g("GRBM_GUI_ACTIVE") / xcds
        print(f"  elapsed cycles per XCD = {el:.0f}")
        if g("SQ_BUSY_CU_CYCLES"):
            print(f"  SQ_BUSY_CU_CYCLES / (elapsed x {cus} CUs) = {g('SQ_BUSY_CU_CYCLES') / (el * cus):.3f}")
    if g("SQ_WAVES"):
        for c in ("SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_INSTS_SALU", "SQ_INSTS_VMEM_RD"):
            if g(c):
                print(f"  {c} per wave = {g(c) / g('SQ_WAVES'):.0f}")
    if g("SQ_WAIT_ANY") and g("SQ_WAVE_CYCLES"):
        print(f"  SQ_WAIT_ANY / SQ_WAVE_CYCLES = {g('SQ_WAIT_ANY') / g('SQ_WAVE_CYCLES'):.3f}")
    if g("SQ_LDS_BANK_CONFLICT") and g("SQ_LDS_IDX_ACTIVE"):
        print(f"  SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE = {g('SQ_LDS_BANK_CONFLICT') / g('SQ_LDS_IDX_ACTIVE'):.3f}")
    return m


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["trace", "counter", "sq"])
    ap.add_argument("dir")
    ap.add_argument("kernel")
    ap.add_argument("--skip", type=int, default=0)
    ap.add_argument("--last", type=int, default=0, help="trace: count only the last N dispatches")
    a = ap.parse_args(argv)
    if a.mode == "trace":
        trace(a.dir, a.kernel, a.skip, a.last)
    else:
        {"counter": counter, "sq": sq}[a.mode](a.dir, a.kernel, a.skip)


if __name__ == "__main__":
    main()
