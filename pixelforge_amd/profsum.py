"""Summaries of rocprofv3 output for profiles/.

    python -m pixelforge_amd.profsum trace DB          kernel-trace db -> per (kernel, grid) durations
    python -m pixelforge_amd.profsum pmc DIR...        --pmc csv dirs -> per (kernel, grid) counter means
    python -m pixelforge_amd.profsum md DB TITLE CMD [GRID...]   markdown table of a kernel trace
    python -m pixelforge_amd.profsum traffic FETCH_DIR WRITE_DIR CAL_FETCH_DIR CAL_WRITE_DIR BYTES KERNEL GRID
        HBM bytes per dispatch of KERNEL (name substring) at GRID from
        FETCH_SIZE / WRITE_SIZE runs, each scaled by the same counter of the
        copy kernel (pixelforge_amd/copy_probe.py) moving BYTES per dispatch.
"""
import collections
import csv
import glob
import json
import os
import re
import sqlite3
import sys


def _short(name):
    name = name.replace("(anonymous namespace)::", "").replace("void ", "")
    return re.sub(r"\((?:[^()]|\([^()]*\))*\)$", "", name).strip()


def trace(db):
    con = sqlite3.connect(db)
    rows = con.execute("select name, grid_x, duration from kernels").fetchall()
    g = collections.defaultdict(list)
    for n, gx, d in rows:
        g[(_short(n), gx)].append(d / 1e3)
    out = []
    for (n, gx), ds in sorted(g.items(), key=lambda kv: -sum(kv[1])):
        ds.sort()
        out.append((n, gx, len(ds), sum(ds) / len(ds), ds[len(ds) // 2], sum(ds)))
    return out


def pmc(dirs):
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for d in dirs:
        for f in glob.glob(os.path.join(d, "**", "*_counter_collection.csv"), recursive=True):
            for r in csv.DictReader(open(f)):
                agg[(_short(r["Kernel_Name"]), int(r["Grid_Size"]))][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in d.items()} for k, d in agg.items()}


def _pick(table, kernel, grid, counter):
    vals = [d[counter] for (n, g), d in table.items() if kernel in n and (grid is None or g == grid) and counter in d]
    if not vals:
        raise SystemExit(f"no {counter} for {kernel} grid {grid}")
    return sum(vals) / len(vals)


def traffic(fetch_dir, write_dir, cal_fetch, cal_write, nbytes, kernel, grid):
    f = _pick(pmc([fetch_dir]), kernel, grid, "FETCH_SIZE")
    w = _pick(pmc([write_dir]), kernel, grid, "WRITE_SIZE")
    cf = _pick(pmc([cal_fetch]), "copyBuffer", None, "FETCH_SIZE")
    cw = _pick(pmc([cal_write]), "copyBuffer", None, "WRITE_SIZE")
    kf, kw = nbytes / cf, nbytes / cw                 # bytes per counter unit
    return {"kernel": kernel, "grid": grid, "fetch_counter": f, "write_counter": w,
            "copy_fetch_counter": cf, "copy_write_counter": cw, "copy_bytes": nbytes,
            "fetch_bytes": f * kf, "write_bytes": w * kw, "traffic_bytes": f * kf + w * kw}


def main():
    if sys.argv[1] == "trace":
        print(f"{'kernel':60s} {'grid':>10s} {'calls':>6s} {'mean us':>10s} {'median us':>10s}")
        for n, gx, c, mean, med, tot in trace(sys.argv[2]):
            print(f"{n[:60]:60s} {gx:10d} {c:6d} {mean:10.2f} {med:10.2f}")
    elif sys.argv[1] == "pmc":
        for (n, g), d in sorted(pmc(sys.argv[2:]).items()):
            print(f"{n}  grid {g}")
            for c, v in sorted(d.items()):
                print(f"    {c:28s} {v:14.6g}")
    elif sys.argv[1] == "md":
        # md DB TITLE COMMAND [GRID...]: markdown table of the kernel trace (optionally only some grids)
        db, title, cmd = sys.argv[2], sys.argv[3], sys.argv[4]
        grids = {int(g) for g in sys.argv[5:]}
        print(f"# {title}\n\nCommand (GPU box, repository root; `cd /tmp && export TMPDIR=/tmp` first):\n")
        print(f"    {cmd}\n")
        print("Per kernel and grid size (work-items): mean and median duration over the dispatches of the run.\n")
        print("| kernel | grid | calls | mean µs | median µs |\n|---|---:|---:|---:|---:|")
        for n, gx, c, mean, med, tot in trace(db):
            if grids and gx not in grids:
                continue
            print(f"| `{n}` | {gx} | {c} | {mean:.2f} | {med:.2f} |")
    elif sys.argv[1] == "traffic":
        a = sys.argv[2:]
        print(json.dumps(traffic(a[0], a[1], a[2], a[3], float(a[4]), a[5], int(a[6])), indent=1))


if __name__ == "__main__":
    main()
