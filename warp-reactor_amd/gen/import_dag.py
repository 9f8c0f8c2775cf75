"""Import the primordial network DAG from the reference's Mojo source into IR.

Usage: python import_dag.py [PATH/reproducer.mojo] [OUT.ir]

Reads the three generated functions ``rhs_specie``, ``rhs_eint`` and
``jac_nuc`` (``reproducer.mojo:388-2321``), checks the pinned manifest
(119/14, 162/1, 1394/225 definitions/outputs, ``tools/slice_dag.py:26-30``),
hash-conses the three into one DAG and writes the line-oriented IR that the
emitters in ``emit.py`` consume.  The IR records the sha256 of the source it
came from.  This is a one-time import step; the build reads the committed IR.
"""
import hashlib
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import dag as dagmod  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_SRC = "/root/reference/reproducer.mojo"
DEFAULT_OUT = os.path.join(HERE, "primordial_network.ir")


def main(argv):
    sys.setrecursionlimit(100000)
    src = argv[1] if len(argv) > 1 else DEFAULT_SRC
    out = argv[2] if len(argv) > 2 else DEFAULT_OUT
    raw = open(src, "rb").read()
    sha = hashlib.sha256(raw).hexdigest()
    d = dagmod.parse_network(raw.decode())
    digest = dagmod.write_ir(d, out, sha)
    d2 = dagmod.read_ir(out)
    print(f"source {src} sha256 {sha}")
    print(f"IR {out}: {len(d2.nodes)} nodes, sha256 {digest}")
    print("op counts:", dict(sorted(d2.op_counts().items())))


if __name__ == "__main__":
    main(sys.argv)
