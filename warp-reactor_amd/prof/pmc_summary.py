"""Per-kernel averages of a rocprofv3 counter-collection CSV.

    python3 warp-reactor_amd/prof/pmc_summary.py OUT/..._counter_collection.csv

Prints, for every kernel and counter: dispatches and mean value per
dispatch (FETCH_SIZE / WRITE_SIZE are in KiB)."""
import collections
import csv
import sys


def summarize(path):
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    with open(path) as f:
        for r in csv.DictReader(f):
            agg[r["Kernel_Name"]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    out = {}
    for k, cs in agg.items():
        out[k] = {c: (len(v), sum(v) / len(v)) for c, v in cs.items()}
    return out


def main(argv):
    for path in argv:
        print(f"# {path}")
        for k, cs in sorted(summarize(path).items()):
            for c, (n, mean) in sorted(cs.items()):
                print(f"{c:>22s} {n:6d} dispatches  mean {mean:16.3f}  {k}")


if __name__ == "__main__":
    main(sys.argv[1:])
