"""Summaries of rocprofv3 runs as short markdown tables.

    python3 -m elmm_amd.profsum RESULTS OUT.md TITLE COMMAND
    python3 -m elmm_amd.profsum --counters OUT.md TITLE COMMAND LABEL=DB[,DB...] ...

With --counters each LABEL=DB,... group merges the counters of several
`rocprofv3 --pmc` passes over the same program.

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


COUNTER_KERNELS = ("k_momentum_zm", "k_scalar_zm", "k_sgs_zm", "k_sgs<", "k_projection", "k_fft_x_fwd",
                   "k_fft_x_inv", "k_fft_y", "k_solve_z")

# Per dispatch a counter has one value per hardware instance (shader engine,
# XCD...): SQ and TCC counters are summed over instances, GRBM ones (cycles
# of each XCD's graphics engine) averaged; dispatches are then averaged.
SIMDS = 1024  # 256 CUs x 4 SIMDs on MI355X


def _dispatch_values(path):
    import sqlite3
    db = sqlite3.connect(path)
    try:
        per = {}
        for did, name, cn, cv in db.execute("select dispatch_id, name, counter_name, counter_value from pmc_events"):
            d = per.setdefault((name, did), {})
            d.setdefault(cn, []).append(cv)
    finally:
        db.close()
    out = {}
    for (name, _), d in per.items():
        agg = out.setdefault(name, {})
        for cn, vals in d.items():
            v = sum(vals) / len(vals) if cn.startswith("GRBM") else sum(vals)
            agg.setdefault(cn, []).append(v)
    return {n: {c: sum(v) / len(v) for c, v in d.items()} for n, d in out.items()}


def counters(dst, title, cmd, groups):
    rows, cols = [], []
    for g in groups:
        label, dbs = g.split("=", 1)
        vals = {}
        for path in dbs.split(","):
            for kn, d in _dispatch_values(path).items():
                vals.setdefault(kn, {}).update(d)
                for cn in d:
                    if cn not in cols:
                        cols.append(cn)
        for key in COUNTER_KERNELS:
            for kn, d in vals.items():
                if key in kn:
                    rows.append((label, short(kn), d))
    derived = []
    if "SQ_ACTIVE_INST_VALU" in cols and "GRBM_GUI_ACTIVE" in cols:
        derived.append(("VALU busy %", lambda d: 100 * d["SQ_ACTIVE_INST_VALU"] * 4 / SIMDS / d["GRBM_GUI_ACTIVE"]))
    if "SQ_INSTS_VALU" in cols and "GRBM_GUI_ACTIVE" in cols:
        derived.append(("VALU instr / SIMD-cycle", lambda d: d["SQ_INSTS_VALU"] / SIMDS / d["GRBM_GUI_ACTIVE"]))
    if "SQ_INSTS_VALU" in cols and "SQ_WAVES" in cols:
        derived.append(("VALU instr / wave", lambda d: d["SQ_INSTS_VALU"] / d["SQ_WAVES"]))
    with open(dst, "w") as f:
        f.write(f"# {title}\n\nCommands (from the repository root on an MI355X):\n\n```\n{cmd}\n```\n\n")
        f.write("Per dispatch: SQ counters summed over instances, GRBM_GUI_ACTIVE averaged over XCDs; "
                "dispatches averaged.\n\n")
        names = [c for c in cols] + [n for n, _ in derived]
        f.write("| form | kernel | " + " | ".join(names) + " |\n|---|---|" + "---:|" * len(names) + "\n")
        for label, k, d in rows:
            vals = [f"{d.get(c, 0):.4g}" for c in cols]
            for _, fn in derived:
                try:
                    vals.append(f"{fn(d):.3g}")
                except (KeyError, ZeroDivisionError):
                    vals.append("")
            f.write(f"| {label} | `{k}` | " + " | ".join(vals) + " |\n")


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
