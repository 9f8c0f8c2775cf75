"""Per-run averages of a kernel's launches in a rocprofv3 kernel-trace
database (rocpd SQLite).

    python3 warp-reactor_amd/prof/trace_phases.py OUT/..._results.db [KERNEL_PREFIX]

bench.py uploads the cells (k_scatter) before every run, so the launches of a
kernel between two uploads belong to one run; prints, per run, the number of
launches, their mean and total duration.  Default prefix: the production
queue-fed advance instance."""
import bisect
import sqlite3
import sys

PROD = "void wr::k_advance_q<false, wr::QPlainNet, false, false>"


def phases(db, prefix=PROD):
    c = sqlite3.connect(db)
    rows = list(c.execute("select name, start, end from kernels order by start"))
    uploads = [s for n, s, _ in rows if n.startswith("wr::k_scatter")]
    seg = {}
    for n, s, e in rows:
        if n.startswith(prefix):
            seg.setdefault(bisect.bisect(uploads, s), []).append(e - s)
    return [(k, len(v), sum(v) / len(v) / 1e6, sum(v) / 1e9) for k, v in sorted(seg.items())]


def main(argv):
    prefix = argv[1] if len(argv) > 1 else PROD
    print(f"# {prefix}")
    print("# run (uploads before it)  launches  mean ms  total s")
    for k, n, mean, tot in phases(argv[0], prefix):
        print(f"{k:24d} {n:9d} {mean:8.3f} {tot:8.3f}")


if __name__ == "__main__":
    main(sys.argv[1:])
