"""HBM traffic per launch from rocprofv3 counter runs.

    python3 -m elmm_amd.traffic GROUP KERNEL_MATCH WORKLOAD FETCH WRITE \\
        COPY_FETCH COPY_WRITE COPY_BYTES OUT_DIR [COMMANDS]

FETCH / WRITE are the outputs of two separate `rocprofv3 --pmc FETCH_SIZE`
and `--pmc WRITE_SIZE` runs of the workload (the `*_results.db` database of
rocprofv3's default output format, or a `*_counter_collection.csv`); COPY_*
the same counters for `bench.py --copy-probe COPY_BYTES` (a plain copy kernel
that reads and writes COPY_BYTES per dispatch).  The counters are
in KiB per dispatch; each is scaled by the ratio the copy kernel shows
between its known bytes and its counter.  Writes
OUT_DIR/traffic_<group>_<workload>.json (read by bench.py for the kernel
group of that name) and a .md summary beside it; KERNEL_MATCH selects the
dispatches by kernel-name substring; COMMANDS (text) is recorded in the
summary as the commands that produced the counter outputs.
"""
import csv
import json
import os
import sys


def _rows(path):
    """(kernel name, counter name, dispatch id, value) of a rocprofv3 counter
    output: the rocpd SQLite database or the CSV format."""
    if path.endswith(".db"):
        import sqlite3
        db = sqlite3.connect(path)
        try:
            yield from db.execute("select kernel_name, counter_name, dispatch_id, value from counters_collection")
        finally:
            db.close()
        return
    with open(path) as f:
        for row in csv.DictReader(f):
            yield (row.get("Kernel_Name", ""), row.get("Counter_Name"),
                   row.get("Dispatch_Id") or row.get("Correlation_Id"), float(row["Counter_Value"]))


def per_dispatch(path, kernel_substr, counter):
    """Mean counter value (summed over the counter's instances) per dispatch
    of the kernels whose name contains kernel_substr."""
    by_dispatch = {}
    for name, cname, d, value in _rows(path):
        if kernel_substr not in name or cname != counter:
            continue
        by_dispatch[d] = by_dispatch.get(d, 0.0) + float(value)
    if not by_dispatch:
        raise SystemExit(f"no {counter} rows for kernels matching {kernel_substr!r} in {path}")
    return sum(by_dispatch.values()) / len(by_dispatch), len(by_dispatch)


def main(argv):
    group, kernel, workload, fetch, write, cfetch, cwrite, copy_bytes, out = argv[:9]
    commands = argv[9] if len(argv) > 9 else ""
    copy_bytes = int(copy_bytes)
    f, nf = per_dispatch(fetch, kernel, "FETCH_SIZE")
    w, nw = per_dispatch(write, kernel, "WRITE_SIZE")
    cf, _ = per_dispatch(cfetch, "copy_probe", "FETCH_SIZE")
    cw, _ = per_dispatch(cwrite, "copy_probe", "WRITE_SIZE")
    sf = copy_bytes / (cf * 1024.0)
    sw = copy_bytes / (cw * 1024.0)
    fetch_b = f * 1024.0 * sf
    write_b = w * 1024.0 * sw
    res = {"group": group, "kernel": kernel, "workload": workload, "dispatches": [nf, nw],
           "fetch_kib_per_dispatch": f, "write_kib_per_dispatch": w,
           "copy_bytes": copy_bytes, "copy_fetch_kib": cf, "copy_write_kib": cw,
           "scale_fetch": sf, "scale_write": sw,
           "hbm_bytes_per_launch": int(fetch_b + write_b), "fetch_bytes": int(fetch_b), "write_bytes": int(write_b)}
    os.makedirs(out, exist_ok=True)
    base = os.path.join(out, f"traffic_{group}_{workload}")
    with open(base + ".json", "w") as fh:
        json.dump(res, fh, indent=1)
    with open(base + ".md", "w") as fh:
        fh.write(f"# HBM traffic of {kernel}, {workload}\n\n")
        fh.write("rocprofv3 `--pmc FETCH_SIZE` and `--pmc WRITE_SIZE` in separate runs (KiB per dispatch), scaled by "
                 f"a plain copy kernel of {copy_bytes} bytes read + written per dispatch.\n\n")
        if commands:
            fh.write(f"Commands (from the repository root on an MI355X):\n\n```\n{commands}\n```\n\n")
        fh.write("| quantity | value |\n|---|---:|\n")
        for k in ("fetch_kib_per_dispatch", "write_kib_per_dispatch", "copy_fetch_kib", "copy_write_kib",
                  "scale_fetch", "scale_write", "fetch_bytes", "write_bytes", "hbm_bytes_per_launch"):
            v = res[k]
            fh.write(f"| {k} | {v:.4g} |\n" if isinstance(v, float) else f"| {k} | {v} |\n")
    print(json.dumps(res))


if __name__ == "__main__":
    main(sys.argv[1:])
