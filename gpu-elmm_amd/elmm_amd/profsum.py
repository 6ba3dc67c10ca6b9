"""Kernel-time summary of a rocprofv3 run as a short markdown table.

    python3 -m elmm_amd.profsum RESULTS OUT.md TITLE COMMAND

RESULTS is the `*_results.db` database that `rocprofv3 --kernel-trace
--stats -d DIR -o run -- ...` writes (or its `*_kernel_stats.csv`); COMMAND is
the command that produced it, recorded in the summary.
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


def main(argv):
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
