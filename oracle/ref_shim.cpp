// Test infrastructure, not product code.
//
// The drop-in shim for run_cells_cuda (reproducer.cpp:7791-7975) that a
// maintainer of the reference would add, compiled against the reference's own
// types and helpers: this translation unit includes reproducer.cpp where it
// lies (built by oracle/Makefile with PRIMORDIAL_ROS2S_NO_MAIN, nothing
// copied) and links libwarp_reactor.so.  main() does what the reference's
// GPU main does around the call (reproducer.cpp:8617-8628: redshift,
// make_collapse_state for every cell, run_cells_cuda) and writes the packed
// final states with the reference's own writer (8357), so a test can compare
// the file with the one the library's CLI writes.
//
//   ref_shim GRID OUT_FILE    exit 0 on success
#include "reproducer.cpp"

#include <cstdlib>

#include "warp_reactor.h"

static_assert(sizeof(CollapseState) == sizeof(wr_collapse_state),
              "CollapseState and wr_collapse_state differ in size");

// The shim itself (the text of INTEGRATION.md, section 1).
integrators::IntegratorResult run_cells_cuda(std::vector<CollapseState>& cells, bool perturb,
                                             int& completed_global_steps) {
    const int r = wr_run_cells(reinterpret_cast<wr_collapse_state*>(cells.data()),
                               static_cast<int>(cells.size()), perturb ? 1 : 0,
                               &completed_global_steps);
    return static_cast<integrators::IntegratorResult>(r);
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::cerr << "usage: " << argv[0] << " GRID OUT_FILE\n";
        return 2;
    }
    const int grid = std::atoi(argv[1]);
    if (grid < 2 || grid > 64) {
        std::cerr << "grid must be 2..64\n";
        return 2;
    }
    pc::set_redshift(30.0);
    std::vector<CollapseState> cells(static_cast<std::size_t>(grid) * grid * grid);
    for (auto& cell : cells) cell = make_collapse_state();
    int completed_global_steps = 0;
    const integrators::IntegratorResult failure =
        run_cells_cuda(cells, true, completed_global_steps);
    if (failure != integrators::IntegratorResult::SUCCESS) {
        std::cerr << "run_cells_cuda failed: " << static_cast<int>(failure) << "\n";
        return 1;
    }
    std::cout << "completed global collapse steps: " << completed_global_steps << "\n";
    return write_final_states(cells, grid, argv[2]) ? 0 : 1;
}
