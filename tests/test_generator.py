"""The network generator as a tool (gen/dag.py, gen/import_dag.py, gen/emit.py,
gen/stage.py, gen/coop.py, gen/table.py), restating the reference's own slicer tests
(tests/test_slice_dag.py:27-78 parser cases, 82-161 manifest and deterministic
output) for this build's parser and emitters:

* the parser accepts multi-line definitions and names with any suffix, and
  rejects duplicate definitions, uses before definition, unsupported
  statements and a manifest that differs from the pinned counts
  (119/14, 162/1, 1394/225 definitions/outputs, tools/slice_dag.py:26-30);
* re-importing the reference's generated functions (reproducer.mojo:388-2321)
  reproduces the committed IR byte for byte, source checksum included, when
  the reference source is present;
* every emitter writes the same bytes on every run, and the op counts the
  roofline uses are pinned.
"""
import hashlib
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(REPO, "warp-reactor_amd", "gen")
IR = os.path.join(GEN, "primordial_network.ir")
sys.path.insert(0, GEN)

import dag as dagmod  # noqa: E402

REFERENCE_MOJO = "/root/reference/reproducer.mojo"


def _source(rhs_body, eint="    var T = state.T\n    var x0_0 = T\n    return x0_0\n",
            jac="    var T = state.T\n    var x0_0 = T\n    mset(jac, 0, 0, x0_0)\n"):
    return ("def rhs_specie(state, mut ydot, X, z):\n"
            + "".join("    " + l + "\n" for l in rhs_body.splitlines())
            + "\ndef rhs_eint(state, X, z):\n" + eint
            + "\ndef jac_nuc(state, mut jac, X, z):\n" + jac)


def test_multiline_definition_and_nonzero_suffix():
    d = dagmod.Dag()
    with pytest.raises(ValueError, match="manifest drift"):
        # a well-formed but small network parses and then fails the manifest pin
        dagmod.parse_network(_source("var T = state.T\nvar x1_7 = ((((\n   T\n))))\n"
                                     "var x2_3 = x1_7 + 1\nvset(ydot, 0, x2_3)"), d)
    # the multi-line definition was read as one statement: x2_3 = T + 1
    T = d.index[dagmod.Node("T", (), None)]
    one = d.const(1.0)
    assert dagmod.Node("add", (T, one), None) in d.index
    assert d.outputs[("ydot", 0)] == d.index[dagmod.Node("add", (T, one), None)]


@pytest.mark.parametrize("body,message", [
    ("var T = state.T\nvar x0_0 = T\nvar x0_0 = T\nvset(ydot, 0, x0_0)", "duplicate"),
    ("var T = state.T\nvar x0_0 = x1_0\nvar x1_0 = T\nvset(ydot, 0, x0_0)", "use before"),
    ("var T = state.T\nprint(T)", "unsupported"),
    ("var T = state.T\nvar x0_0 = (T\nvset(ydot, 0, x0_0)", "unbalanced|expected|unexpected"),
])
def test_parser_rejects_bad_sources(body, message):
    with pytest.raises(SyntaxError, match=message):
        dagmod.parse_network(_source(body))


def test_manifest_drift_is_rejected():
    with pytest.raises(ValueError, match="manifest drift"):
        dagmod.parse_network(_source("var T = state.T\nvar x0_0 = T\nvset(ydot, 0, x0_0)"))


def test_committed_ir_pins():
    d = dagmod.read_ir(IR)
    assert d.meta["manifest"] == {"rhs_specie": (119, 14), "rhs_eint": (162, 1),
                                  "jac_nuc": (1394, 225)}
    assert len(d.nodes) == 8038
    assert len(d.outputs) == 14 + 1 + 225
    assert d.meta["source_sha256"] == \
        "be2adf68ec04b3f7854f7d87a173e0f346ea8fea49987cc5412c525db9416f98"


@pytest.mark.skipif(not os.path.exists(REFERENCE_MOJO), reason="reference source not present")
def test_reimport_reproduces_the_committed_ir(tmp_path):
    out = tmp_path / "network.ir"
    subprocess.run([sys.executable, os.path.join(GEN, "import_dag.py"), REFERENCE_MOJO, str(out)],
                   check=True, capture_output=True)
    assert out.read_bytes() == open(IR, "rb").read()


@pytest.mark.parametrize("script,outputs", [
    ("emit.py", ["chem_network.h", "chem_counts.json"]),
    ("coop.py", ["coop_network.h", "coop_stats.json"]),
    ("table.py", ["chem_table.h", "chem_table.json"]),
])
def test_emitters_are_deterministic(tmp_path, script, outputs):
    digests = []
    for run in ("a", "b"):
        d = tmp_path / run
        d.mkdir()
        subprocess.run([sys.executable, os.path.join(GEN, script)] + [str(d / o) for o in outputs],
                       check=True, capture_output=True)
        digests.append([hashlib.sha256((d / o).read_bytes()).hexdigest() for o in outputs])
    assert digests[0] == digests[1]


def test_emitted_op_counts_are_pinned(tmp_path):
    import json
    subprocess.run([sys.executable, os.path.join(GEN, "emit.py"), str(tmp_path / "n.h"),
                    str(tmp_path / "c.json")], check=True, capture_output=True)
    c = json.load(open(tmp_path / "c.json"))
    flops = ("add", "sub", "mul", "div", "sqrt", "exp", "log")
    assert sum(c["chem_fused"].get(op, 0) for op in flops) == 5966
    assert sum(c["chem_rhs"].get(op, 0) for op in flops) == 1536
    assert c["chem_fused"]["exp"] == 278 and c["chem_fused"]["log"] == 69
    assert c["staged"]["fused"]["R"] == 128


def test_table_frontier_is_smooth_and_temperature_only():
    """The rate-table split (gen/table.py): every frontier value depends on T
    and constants only, through no selection or test, and everything the
    network consumes from the T-only part that is not a frontier value is a
    selection, a test or a leaf (so the split network computes it from T)."""
    import emit
    import table
    d = emit.specialise(dagmod.read_ir(IR))
    roots = list(d.outputs.values())
    fr = table.frontier(d, roots)
    sm = table.smooth(d)
    assert len(fr) == 441 and len(set(fr)) == len(fr)
    for nid in fr:
        assert sm[nid] and d.nodes[nid].op not in table.LEAF_OPS
    live = table.live_with_leaves(d, roots, set(fr))
    for nid in live:
        n = d.nodes[nid]
        if sm[nid] and nid not in fr:
            assert n.op in ("const", "T"), n.op  # smooth values only enter through tv
    ops = {}
    for nid in d.live_nodes(fr):
        op = d.nodes[nid].op
        ops[op] = ops.get(op, 0) + 1
    assert ops.get("exp") == 213 and ops.get("log") == 24 and "sel" not in ops
