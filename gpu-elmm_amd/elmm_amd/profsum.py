"""Summaries of rocprofv3 runs as short markdown tables.

    python3 -m elmm_amd.profsum RESULTS OUT.md TITLE COMMAND
    python3 -m elmm_amd.profsum --counters OUT.md TITLE COMMAND LABEL=DB[,DB...] ...

RESULTS is the `*_results.db` database that `rocprofv3 --kernel-trace
--stats -d DIR -o run -- ...` writes (or its `*_kernel_stats.csv`); with
--counters, each LABEL names a set of `rocprofv3 --pmc ...` databases whose
per-dispatch counter values are averaged per kernel.  COMMAND is recorded in
the summary as the command that produced the inputs.
"""
import csv
import sys


def kernel_stats(path):
    """[(name, calls, total_ns, average_ns)] sorted by total time."""
    if path.endswith(".db"):
        import sqlite3
        db = sqlite3.connect(path)
        try:
            rows = list(db.execute("select name, count(*), sum(duration), avg(duration) from kernels group by name"))
        finally:
            db.close()
    else:
        with open(path) as f:
            rows = [(r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"]))
                    for r in csv.DictReader(f)]
    return sorted(((n, int(c), float(t), float(a)) for n, c, t, a in rows), key=lambda r: -r[2])


def short(name):
    return name.replace("(anonymous namespace)::", "").replace("void ", "").split("(")[0]


COUNTER_KERNELS = ("k_momentum_zm", "k_scalar_zm", "k_sgs<", "k_projection", "k_fft_x_fwd", "k_fft_x_inv",
                   "k_fft_y", "k_solve_z")


def counters(dst, title, cmd, groups):
    import sqlite3
    rows, cols = [], []
    for g in groups:
        label, dbs = g.split("=", 1)
        vals = {}
        for path in dbs.split(","):
            db = sqlite3.connect(path)
            try:
                for kn, cn, v in db.execute("select kernel_name, counter_name, avg(value) from counters_collection "
                                            "group by kernel_name, counter_name"):
                    vals.setdefault(kn, {})[cn] = v
                    if cn not in cols:
                        cols.append(cn)
            finally:
                db.close()
        for key in COUNTER_KERNELS:
            for kn, d in vals.items():
                if key in kn:
                    rows.append((label, short(kn), d))
    with open(dst, "w") as f:
        f.write(f"# {title}\n\nCommands (from the repository root on an MI355X):\n\n```\n{cmd}\n```\n\n")
        f.write("| form | kernel | " + " | ".join(cols) + " |\n|---|---|" + "---:|" * len(cols) + "\n")
        for label, k, d in rows:
            f.write(f"| {label} | `{k}` | " + " | ".join(f"{d.get(c, 0):.3g}" for c in cols) + " |\n")


def main(argv):
    if argv and argv[0] == "--counters":
        counters(argv[1], argv[2], argv[3], argv[4:])
        return
    src, dst, title, cmd = argv
    rows = kernel_stats(src)
    tot = sum(r[2] for r in rows)
    with open(dst, "w") as f:
        f.write(f"# {title}\n\nCommand (from the repository root on an MI355X):\n\n```\n{cmd}\n```\n\n")
        f.write("| kernel | calls | total ms | avg us | % |\n|---|---:|---:|---:|---:|\n")
        for n, c, t, a in rows:
            f.write(f"| `{short(n)}` | {c} | {t / 1e6:.3f} | {a / 1e3:.1f} | {100 * t / tot:.2f} |\n")
        f.write(f"\nTotal kernel time: {tot / 1e6:.2f} ms\n")


if __name__ == "__main__":
    main(sys.argv[1:])
