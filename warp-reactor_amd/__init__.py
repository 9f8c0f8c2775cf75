"""MI355X (gfx950) primordial-chemistry ROS2S grid advance.

Python mirror of the reference's device-runner interface
(``run_cells_cuda``, reproducer.cpp:7791) over the C ABI in
``include/warp_reactor.h``.  The compute runs in the HIP library
``lib/libwarp_reactor.so``; there is no CPU fallback: loading fails loudly
when the library is missing.

Import with ``importlib.import_module("warp-reactor_amd")`` (the directory
name is not a Python identifier).
"""
from .runtime import (  # noqa: F401
    CELL_DTYPE,
    PACKED_DTYPE,
    STATS_FIELDS,
    SPECIES_NAMES,
    Context,
    WarpReactorError,
    adjoint_ntheta,
    reaction_table,
    compare_final_states,
    library,
    library_path,
    RECIPES,
    make_collapse_states,
    make_recipe_states,
    pack_final_states,
    read_final_states,
    run_cells,
    write_final_states,
)
