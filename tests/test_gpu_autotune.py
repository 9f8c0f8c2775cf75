"""Launch shapes of the queue-fed kernel and the autotuner that picks one:
every shape gives the same grid bit for bit (the per-cell work does not depend
on which lane takes a cell), and the autotuner leaves its choice set."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bytes(cells):
    mask = np.ones(216, bool)
    mask[156:160] = False
    return cells.view(np.uint8).reshape(len(cells), 216)[:, mask]


def test_launch_shapes_give_identical_grids(wr):
    n = 6 ** 3
    init = wr.make_collapse_states(n)
    ref = None
    with wr.Context(n) as ctx:
        for shape in [(256, 0, 64), (64, 0, 0), (128, 1, 64), (64, 1, 64), (192, 0, 0)]:
            print(f"launch shape {shape}", flush=True)
            ctx.set_queue_launch(*shape)
            cells = init.copy()
            ctx.upload(cells)
            info = ctx.run(perturb=True, max_steps=120)
            ctx.download(cells)
            assert info["completed_global_steps"] == 120
            assert ctx.queue_launch()[:3] == shape
            if ref is None:
                ref = cells
            else:
                assert np.array_equal(_bytes(cells), _bytes(ref)), shape


def test_bad_launch_shapes_are_rejected(wr):
    with wr.Context(8) as ctx:
        for shape in [(32, 0, 0), (320, 0, 0), (100, 0, 0), (64, -1, 0), (64, 0, 32), (64, 0, 128)]:
            with pytest.raises(wr.WarpReactorError):
                ctx.set_queue_launch(*shape)


def test_autotuner_sets_its_choice(wr):
    autotune = importlib.import_module("warp-reactor_amd.autotune")
    n = 4 ** 3
    init = wr.make_collapse_states(n)
    with wr.Context(n) as ctx:
        chosen, table = autotune.autotune(ctx, init, steps=20)
        assert chosen in [(r["threads"], r["blocks_per_cu"], r["solo_lanes"], r["coop_cells"])
                          for r in table]
        assert len(table) == len(list(autotune.candidates()))
        assert ctx.queue_launch()[:3] == chosen[:3]
        assert ctx.coop()[:2] == (chosen[3], chosen[3])
