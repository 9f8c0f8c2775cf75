"""The C-ABI library loads and exports every function include/warp_reactor.h
declares; host-only entry points (no GPU needed) behave as the reference's."""
import os
import re

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(REPO, "include", "warp_reactor.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(wr_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(wr):
    lib = wr.library()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(wr.runtime.EXPORTED)
    assert lib.wr_abi_version() == 2


def test_initial_state_matches_oracle(wr, oracle):
    a = wr.make_collapse_states(3)
    b = oracle.make_cells(3, wr.CELL_DTYPE)
    assert a.tobytes() == b.tobytes()
