"""The reproducer-compatible CLI on the GPU: stdout lines of
reproducer.cpp:8611-8703 for the gridwide-independent device policy, the
packed final-state file (8357-8378) and the comparison flow (8416-8578).

The run is bit-identical to the oracle built with the device libm, so every
printed summary number and every byte of the written file is checked
against the oracle's grid.  The golden file tests/golden/final_states_grid2.bin
was written by the reference's own writer from its CPU arithmetic; the
device run must pass the reference comparator against it."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(REPO, "warp-reactor_amd", "bin", "wr_reproducer")
GOLD2 = os.path.join(REPO, "tests", "golden", "final_states_grid2.bin")


def run(args, cwd):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=300, cwd=cwd)


def _summary(lines, label):
    line = next(l for l in lines if l.startswith(label + ": "))
    return [float(x) for x in re.findall(r"[-+0-9.eE]+", line[len(label) + 2:])]


def _med(v):
    v = np.sort(np.asarray(v, dtype=np.float64))
    n = len(v)
    return v[n // 2] if n % 2 else (v[n // 2 - 1] + v[n // 2]) / 2


@pytest.fixture(scope="module")
def oracle_grid2(wr, oracle_wr):
    cells = oracle_wr.make_cells(8, wr.CELL_DTYPE)
    r, completed = oracle_wr.run_gridwide(cells, True)
    assert r == 1
    return cells, completed


def test_cli_grid2_stdout_file_and_pass(tmp_path, wr, oracle_grid2):
    ocells, ocompleted = oracle_grid2
    r = run(["--grid", "2", "--compare-final-state", GOLD2], tmp_path)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[:5] == ["Primordial chemistry collapse grid with ROS2S", "grid: 2^3 (8 cells)",
                         "perturbations: enabled", "driver policy: gridwide-independent",
                         f"completed global collapse steps: {ocompleted}"]
    st = ocells["stats"]
    for label, v in (("cell completed steps", ocells["completed_steps"]),
                     ("cell physical time", ocells["time"]),
                     ("cell density driver", ocells["density_driver"]),
                     ("ROS2S internal steps", st["internal_steps"]),
                     ("ROS2S rhs calls", st["rhs_calls"]),
                     ("ROS2S jacobian calls", st["jacobian_calls"]),
                     ("ROS2S decompositions", st["decompositions"]),
                     ("ROS2S linear solves", st["linear_solves"])):
        assert _summary(lines, label) == [float(np.min(v)), float(_med(v)), float(np.max(v))], label
    acc = _summary(lines, "ROS2S accepted/rejected")
    a, j = st["accepted_steps"], st["rejected_steps"]
    assert acc == [a.min(), j.min(), _med(a), _med(j), a.max(), j.max()]
    assert any(l.startswith("wall time: ") and l.endswith(" s") for l in lines)
    assert f"final-state comparison file: {GOLD2}" in lines
    assert any(l.startswith("final-state comparison: PASS (max deuterium-bearing") for l in lines)
    assert "final states: final_states_grid2_hip.bin (8 packed records, 172 bytes each)" in lines
    assert lines[-1] == "PASS" and lines[-2] == ""
    data = (tmp_path / "final_states_grid2_hip.bin").read_bytes()
    assert data == wr.pack_final_states(ocells, 2).tobytes()


def test_cli_rotates_and_reports_failures(tmp_path, wr, oracle_grid2):
    """A second run moves the existing file to .old.NNNNNN, NNNNNN a random
    six-digit suffix (8307-8355); a
    comparison against a file with one perturbed temperature fails, prints
    FAIL, names the field and writes the failures CSV (8100-8142)."""
    ocells, _ = oracle_grid2
    assert run(["--grid", "2", "--no-compare-final-state"], tmp_path).returncode == 0
    bad = wr.pack_final_states(ocells, 2)
    bad["T"][3] *= 1.01
    bad_path = tmp_path / "bad.bin"
    bad_path.write_bytes(bad.tobytes())
    r = run(["--grid", "2", f"--compare-final-state={bad_path}"], tmp_path)
    assert r.returncode == 1
    lines = r.stdout.splitlines()
    assert lines[-1] == "FAIL"
    assert any(l.startswith("final-state comparison: FAIL") for l in lines)
    assert any(l.startswith("first comparison failure:") and "T" in l for l in lines)
    old = [p.name for p in tmp_path.iterdir() if p.name.startswith("final_states_grid2_hip.bin.old.")]
    assert len(old) == 1 and re.fullmatch(r"final_states_grid2_hip\.bin\.old\.[0-9]{6}", old[0])
    assert any(l == f"existing final states moved to: final_states_grid2_hip.bin.old.{old[0][-6:]}"
               for l in lines)
    assert (tmp_path / "final_states_grid2_hip.bin").exists()
    csv = (tmp_path / "final_state_comparison_failures_hip.csv").read_text().splitlines()
    assert len(csv) >= 2 and "cell" in csv[0]


def test_cli_legacy_local_policy(tmp_path, wr, oracle_wr):
    """--driver-policy legacy-local: the reference CPU build's loop on the
    device; its summary line, the oracle's step count, the file bytes of the
    oracle's grid and a PASS against the reference CPU build's own file
    (tests/golden/final_states_grid2_legacy_local.bin)."""
    gold = os.path.join(REPO, "tests", "golden", "final_states_grid2_legacy_local.bin")
    ocells = oracle_wr.make_cells(8, wr.CELL_DTYPE)
    orr, odone, ofailed = oracle_wr.run_legacy(ocells, True)
    assert orr == 1 and ofailed == -1
    r = run(["--grid", "2", "--driver-policy", "legacy-local", "--compare-final-state", gold],
            tmp_path)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[3:5] == ["driver policy: legacy-local", f"completed global collapse steps: {odone}"]
    assert any(l.startswith("final-state comparison: PASS (max deuterium-bearing") for l in lines)
    assert lines[-1] == "PASS"
    data = (tmp_path / "final_states_grid2_hip_legacy_local.bin").read_bytes()
    assert data == wr.pack_final_states(ocells, 2).tobytes()


def test_cli_gridwide_cta32_policy(tmp_path, wr, oracle_wr):
    """--driver-policy gridwide-cta32 (the Mojo front end's policy): its
    summary lines -- the policy, the oracle's step count and the limiting
    cell, printed as reproducer.mojo:3655-3664 does -- the file bytes of the
    oracle's grid, and the Mojo manifest's tile size and policy tag."""
    ocells = oracle_wr.make_cells(27, wr.CELL_DTYPE)
    orr, odone, olim = oracle_wr.run_cta(ocells, True, 1000, tile=32)
    assert orr == 1
    out = tmp_path / "cta.bin"
    r = run(["--grid", "3", "--driver-policy", "gridwide-cta32", "--no-compare-final-state",
             "--write-final-state", str(out)], tmp_path)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    lim = olim if olim >= 0 else 2 ** 63 - 1
    assert lines[3:6] == ["driver policy: gridwide-cta32",
                          f"completed global collapse steps: {odone}", f"limiting cell: {lim}"]
    data = (tmp_path / "final_states_grid3_hip_cta32.bin").read_bytes()
    assert data == wr.pack_final_states(ocells, 3).tobytes() == out.read_bytes()
    manifest = dict(l.split("=", 1) for l in (tmp_path / "cta.bin.manifest").read_text().splitlines())
    assert manifest["policy"] == "gridwide-cta32-v1" and manifest["tile_cells"] == "32"


def test_write_final_state_file_and_manifest(tmp_path):
    """--write-final-state FILE (the Mojo front end's option,
    reproducer.mojo:3318-3340, 3529-3535): the packed records go to FILE --
    the same bytes as the default final_states_grid<N>_hip.bin -- with a
    FILE.manifest sidecar, and an existing FILE or manifest is never
    overwritten."""
    out = tmp_path / "states.bin"
    r = run(["--grid", "2", "--no-compare-final-state", "--write-final-state", str(out)], tmp_path)
    assert r.returncode == 0, r.stderr
    assert f"final-state output: {out} (8 packed records" in r.stdout
    assert out.read_bytes() == (tmp_path / "final_states_grid2_hip.bin").read_bytes()
    manifest = dict(l.split("=", 1) for l in (tmp_path / "states.bin.manifest").read_text().splitlines())
    assert manifest["format"] == "packed-final-state-v1" and manifest["cell_count"] == "8"
    assert manifest["policy"] == "gridwide-independent" and manifest["grid_dim"] == "2"
    assert manifest["network_source_sha256"] == \
        "be2adf68ec04b3f7854f7d87a173e0f346ea8fea49987cc5412c525db9416f98"
    before = out.read_bytes()
    r = run(["--grid", "2", "--no-compare-final-state", "--write-final-state", str(out)], tmp_path)
    assert r.returncode == 1
    assert "refusing to overwrite final-state output" in r.stderr
    assert out.read_bytes() == before


SHIM = os.path.join(REPO, "oracle", "_ref", "ref_shim")


@pytest.mark.skipif(not os.path.exists(SHIM), reason="oracle/_ref/ref_shim not built (needs the "
                    "reference source where build() runs)")
def test_reference_side_shim_equals_the_cli(tmp_path):
    """The run_cells_cuda shim of INTEGRATION.md, compiled against the
    reference's own CollapseState, make_collapse_state and final-state writer
    (oracle/ref_shim.cpp), drives the library through wr_run_cells and writes
    the same bytes as the library's CLI for the same grid."""
    out = tmp_path / "shim_grid4.bin"
    r = subprocess.run([SHIM, "4", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    c = run(["--grid", "4", "--no-compare-final-state"], tmp_path)
    assert c.returncode == 0, c.stderr
    steps = next(l for l in c.stdout.splitlines() if l.startswith("completed global"))
    assert steps in r.stdout.splitlines()
    assert out.read_bytes() == (tmp_path / "final_states_grid4_hip.bin").read_bytes()
