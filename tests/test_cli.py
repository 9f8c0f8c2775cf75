"""The reproducer-compatible CLI (warp-reactor_amd/bin/wr_reproducer): argument
handling that needs no GPU (reproducer.cpp:8229-8277, 8591-8610)."""
import os
import subprocess

import pytest

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "warp-reactor_amd", "bin", "wr_reproducer")

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")


def run(*args, cwd=None):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60, cwd=cwd)


def test_help_prints_usage_and_succeeds():
    r = run("--help")
    assert r.returncode == 0
    assert r.stderr.startswith("usage: ")
    assert "--grid N" in r.stderr and "--compare-final-state FILE" in r.stderr
    assert "--write-final-state FILE" in r.stderr
    assert r.stdout == ""


@pytest.mark.parametrize("args", [["--grid"], ["--grid", "0"], ["--grid", "-3"], ["--grid", "x"],
                                  ["--bogus"], ["--compare-final-state"],
                                  ["--write-final-state"], ["--write-final-state", ""]])
def test_bad_arguments_fail_with_usage(args, tmp_path):
    r = run(*args, cwd=tmp_path)
    assert r.returncode == 1
    assert "usage: " in r.stderr
    assert r.stdout == ""
    assert not os.listdir(tmp_path)  # nothing written


def test_driver_policy_option(tmp_path):
    assert "--driver-policy gridwide-independent|legacy-local|gridwide-cta32" in \
        run("--help").stderr
    for args in (["--driver-policy"], ["--driver-policy", "gridwide-cta64"]):
        r = run(*args, cwd=tmp_path)
        assert r.returncode == 1 and "usage: " in r.stderr and r.stdout == ""
    assert not os.listdir(tmp_path)
