"""HBM traffic per kernel launch from rocprofv3 counter runs.

Two ``rocprofv3 --pmc`` passes over the benchmark (one with FETCH_SIZE, one
with WRITE_SIZE; the counters report KiB per dispatch) and the same two
passes over a plain device copy of known size (``bench.py --copy-probe``)
give, per kernel, the bytes read from and written to HBM per launch.  The
copy fixes the scale of each counter: a copy of B bytes reads B and writes B.

    python3 -m partmc_amd.pmc_traffic COPY_FETCH.csv COPY_WRITE.csv \\
        FETCH.csv WRITE.csv COPY_BYTES > summary.json
"""
from __future__ import annotations

import csv
import json
import re
import sys
from collections import defaultdict


def per_dispatch(path, counter):
    """{kernel: [value per dispatch]} for one counter of a counter_collection.csv."""
    sums = defaultdict(float)
    names = {}
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            if row.get("Counter_Name") != counter:
                continue
            d = row["Dispatch_Id"]
            sums[d] += float(row["Counter_Value"])
            names[d] = re.sub(r"\(.*", "", row["Kernel_Name"]).replace("pmc::", "").strip()
    out = defaultdict(list)
    for d, v in sums.items():
        out[names[d]].append(v)
    return out


def copy_scale(path, counter, nbytes):
    """bytes per counter unit, from the largest dispatch of the copy probe"""
    vals = [v for vs in per_dispatch(path, counter).values() for v in vs]
    return nbytes / max(vals)


def summarize(copy_fetch, copy_write, fetch, write, nbytes):
    sf = copy_scale(copy_fetch, "FETCH_SIZE", nbytes)
    sw = copy_scale(copy_write, "WRITE_SIZE", nbytes)
    f = per_dispatch(fetch, "FETCH_SIZE")
    w = per_dispatch(write, "WRITE_SIZE")
    kernels = {}
    for k in sorted(set(f) | set(w)):
        fk, wk = f.get(k, []), w.get(k, [])
        rd = sf * sum(fk) / len(fk) if fk else None
        wr = sw * sum(wk) / len(wk) if wk else None
        kernels[k] = {"dispatches": max(len(fk), len(wk)), "fetch_bytes": rd, "write_bytes": wr,
                      "bytes": (rd or 0.0) + (wr or 0.0)}
    return {"copy_bytes": nbytes, "fetch_scale": sf, "write_scale": sw, "kernels": kernels}


if __name__ == "__main__":
    a = sys.argv[1:]
    print(json.dumps(summarize(a[0], a[1], a[2], a[3], float(a[4])), indent=1))
