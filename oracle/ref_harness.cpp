// Test infrastructure, not product code.
//
// C-callable harness around the reference's own CPU implementation.  It is
// compiled (by oracle/Makefile) together with the reference source file where
// it lies, `reproducer.cpp`, with PRIMORDIAL_ROS2S_NO_MAIN so that the
// reference's anonymous-namespace driver functions are visible in this
// translation unit.  Nothing from the reference is copied here: every entry
// point below calls the reference's functions.  The output library
// oracle/_ref/libwr_ref.so is used only by tests/ and the golden-fixture
// script to pin the oracle restatement (oracle/oracle.c) and the HIP path.
//
// The gridwide-independent outer loop exists in the reference only inside the
// GPU build (run_cells_cuda, reproducer.cpp:7791-7975, kernels 7696-7785); the
// CPU build cannot compile it.  ref_run_gridwide() below therefore restates
// that host loop and the two kernel bodies on the CPU, calling the reference's
// apply_perturbation, collapse_timestep, burn_ros2s and EOS helpers for all
// arithmetic.
#include "reproducer.cpp"

#include <omp.h>

#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace {

using integrators::IntegratorResult;
using integrators::Real;
namespace ref_pc = integrators::primordial_chem;

static_assert(sizeof(CollapseState) == 216, "CollapseState layout changed");
static_assert(offsetof(CollapseState, time) == 136, "CollapseState layout changed");
static_assert(offsetof(CollapseState, density_driver) == 144, "CollapseState layout changed");
static_assert(offsetof(CollapseState, completed_steps) == 152, "CollapseState layout changed");
static_assert(offsetof(CollapseState, stats) == 160, "CollapseState layout changed");

// Body of prepare_grid_timestep_kernel (reproducer.cpp:7696-7734) for one cell.
// Returns the candidate dt (DBL_MAX when the cell offers none) and sets
// *failure when the cell is inconsistent.
Real prepare_cell(CollapseState& state, int cell, int completed_global_steps,
                  Real grid_time, int step, bool perturb, int* failure) {
    const Real none = std::numeric_limits<Real>::max();
    if (state.completed_steps < completed_global_steps) {
        return none;
    }
    if (state.completed_steps != completed_global_steps || state.time != grid_time) {
        *failure = static_cast<int>(IntegratorResult::BAD_INPUTS);
        return none;
    }
    apply_perturbation(state, cell, step, perturb);
    const Real dt = collapse_timestep(state);
    if (!valid_positive(dt) || !valid_positive(state.density_driver)) {
        *failure = static_cast<int>(IntegratorResult::BAD_INPUTS);
        return none;
    }
    if (dt < 10.0) {
        const Real tff = dt / tff_reduc;
        state.density_driver += dt * (state.density_driver / tff);
        return none;
    }
    return dt;
}

// Body of advance_collapse_gridwide_kernel (reproducer.cpp:7736-7785) for one
// cell.  Returns 1 when the cell integrated, 0 otherwise; sets *failure on a
// failed burn.
int advance_cell(CollapseState& state, int completed_global_steps, Real next_grid_time,
                 Real dt_grid, int* failure) {
    if (state.completed_steps != completed_global_steps) {
        return 0;
    }
    const Real local_dt = collapse_timestep(state);
    if (local_dt < 10.0) {
        return 0;
    }
    const Real old_density = state.density_driver;
    const Real tff = local_dt / tff_reduc;
    state.density_driver += dt_grid * (state.density_driver / tff);
    if (!valid_positive(state.density_driver) || state.density_driver > 2.0e-6) {
        return 0;
    }
    const Real density_ratio = state.density_driver / old_density;
    for (auto& xn : state.current.xn) {
        xn *= density_ratio;
    }
    state.current.rho *= density_ratio;
    const auto result = burn_ros2s(state.current, dt_grid, state.stats);
    if (result != IntegratorResult::SUCCESS) {
        *failure = static_cast<int>(result);
        return 0;
    }
    ref_pc::floor_and_normalize_number_densities(state.current);
    ref_pc::balance_charge(state.current);
    ref_pc::floor_and_normalize_number_densities(state.current);
    ref_pc::eos_re(state.current);
    state.time = next_grid_time;
    state.completed_steps = completed_global_steps + 1;
    return 1;
}

std::array<Real, 15> to_arr15(const double* p) {
    std::array<Real, 15> a{};
    std::memcpy(a.data(), p, sizeof(a));
    return a;
}

}  // namespace

extern "C" {

int ref_sizeof_collapse_state() { return static_cast<int>(sizeof(CollapseState)); }
int ref_sizeof_packed_state() { return static_cast<int>(sizeof(PackedFinalState)); }

void ref_rhs(const double* y, double* dydt) {
    ref_pc::PrimordialChem::state_type ys = to_arr15(y);
    ref_pc::PrimordialChem::rhs_type out{};
    ref_pc::PrimordialChem::rhs(0.0, ys, out);
    std::memcpy(dydt, out.data(), sizeof(out));
}

void ref_jac(const double* y, double* jac) {
    ref_pc::PrimordialChem::state_type ys = to_arr15(y);
    ref_pc::PrimordialChem::jacobian_type j{};
    ref_pc::PrimordialChem::jacobian(0.0, ys, j);
    for (int r = 0; r < 15; ++r) {
        for (int c = 0; c < 15; ++c) {
            jac[r * 15 + c] = j[r][c];
        }
    }
}

// T from a 15-vector y via burn_state_from_y (reproducer.cpp:7029-7038).
double ref_temperature_from_y(const double* y) {
    return ref_pc::PrimordialChem::burn_state_from_y(to_arr15(y)).T;
}

int ref_lu(double* a, int* ipvt) {
    std::array<std::array<Real, 15>, 15> m{};
    for (int r = 0; r < 15; ++r) {
        for (int c = 0; c < 15; ++c) {
            m[r][c] = a[r * 15 + c];
        }
    }
    std::array<int, 15> ip{};
    const int info = integrators::linalg::lu_decomposition<15>(m, ip);
    for (int r = 0; r < 15; ++r) {
        for (int c = 0; c < 15; ++c) {
            a[r * 15 + c] = m[r][c];
        }
        ipvt[r] = ip[r];
    }
    return info;
}

void ref_lu_solve(const double* lu, const int* ipvt, double* x) {
    std::array<std::array<Real, 15>, 15> m{};
    for (int r = 0; r < 15; ++r) {
        for (int c = 0; c < 15; ++c) {
            m[r][c] = lu[r * 15 + c];
        }
    }
    std::array<int, 15> ip{};
    for (int r = 0; r < 15; ++r) {
        ip[r] = ipvt[r];
    }
    std::array<Real, 15> v = to_arr15(x);
    integrators::linalg::lu_solve<15>(m, ip, v);
    std::memcpy(x, v.data(), sizeof(v));
}

// RODAS<PrimordialChem>::integrate (reproducer.cpp:7205-7347) on y[15] from
// t=0 to tout with the collapse tolerances (configure_ros2s, 7582-7590) and a
// caller-chosen max_steps (fault injection for TOO_MANY_STEPS).
// counters[7] = n_step, n_rhs, n_jac, n_decomp, n_solve, n_accept, n_reject.
int ref_integrate(double* y, double tout, double dt0, int max_steps, int* counters,
                  double* t_end, double* dt_end) {
    Ros2sIntegrator integrator;
    Ros2sIntegrator::State s;
    configure_ros2s(s);
    if (max_steps > 0) {
        s.max_steps = max_steps;
    }
    s.t = 0.0;
    s.tout = tout;
    s.dt = dt0;
    s.y = to_arr15(y);
    const auto result = integrator.integrate(s);
    std::memcpy(y, s.y.data(), 15 * sizeof(double));
    counters[0] = s.n_step;
    counters[1] = s.n_rhs;
    counters[2] = s.n_jac;
    counters[3] = s.n_decomp;
    counters[4] = s.n_solve;
    counters[5] = s.n_accept;
    counters[6] = s.n_reject;
    *t_end = s.t;
    *dt_end = s.dt;
    return static_cast<int>(result);
}

// burn_ros2s (reproducer.cpp:7592-7627) on a burn_t laid out as
// rho, T, e, xn[14]; stats[7] accumulate.
int ref_burn(double* burn, double dt, std::uint64_t* stats) {
    ref_pc::burn_t b;
    std::memcpy(&b, burn, sizeof(b));
    IntegratorStats st{};
    std::memcpy(&st, stats, sizeof(st));
    const auto r = burn_ros2s(b, dt, st);
    std::memcpy(burn, &b, sizeof(b));
    std::memcpy(stats, &st, sizeof(st));
    return static_cast<int>(r);
}

void ref_make_collapse_state(void* cell) {
    const CollapseState s = make_collapse_state();
    std::memcpy(cell, &s, sizeof(s));
}

std::uint64_t ref_splitmix64(std::uint64_t v) { return splitmix64(v); }
double ref_perturbation_factor(int cell, int step) { return perturbation_factor(cell, step); }

void ref_apply_perturbation(void* cell, int index, int step, int enabled) {
    apply_perturbation(*static_cast<CollapseState*>(cell), index, step, enabled != 0);
}

double ref_collapse_timestep(const void* cell) {
    return collapse_timestep(*static_cast<const CollapseState*>(cell));
}

void ref_eos_rt(double* burn) {
    ref_pc::burn_t b;
    std::memcpy(&b, burn, sizeof(b));
    ref_pc::eos_rt(b);
    std::memcpy(burn, &b, sizeof(b));
}

void ref_post_step(double* burn) {
    ref_pc::burn_t b;
    std::memcpy(&b, burn, sizeof(b));
    ref_pc::floor_and_normalize_number_densities(b);
    ref_pc::balance_charge(b);
    ref_pc::floor_and_normalize_number_densities(b);
    ref_pc::eos_re(b);
    std::memcpy(burn, &b, sizeof(b));
}

double ref_prepare_cell(void* cell, int index, int completed, double grid_time, int step,
                        int perturb, int* failure) {
    return prepare_cell(*static_cast<CollapseState*>(cell), index, completed, grid_time, step,
                        perturb != 0, failure);
}

int ref_advance_cell(void* cell, int completed, double next_grid_time, double dt_grid,
                     int* failure) {
    return advance_cell(*static_cast<CollapseState*>(cell), completed, next_grid_time, dt_grid,
                        failure);
}

// Restatement of run_cells_cuda's host loop (reproducer.cpp:7859-7959) on the
// CPU, cells in parallel within each kernel-equivalent phase.  max_steps<=0
// means the reference's 1000.  Returns the IntegratorResult code.
int ref_run_gridwide(void* cells_v, int num_cells, int perturb, int max_steps,
                     int* completed_global_steps) {
    auto* cells = static_cast<CollapseState*>(cells_v);
    const int limit = max_steps > 0 ? max_steps : max_collapse_steps;
    const int success = static_cast<int>(IntegratorResult::SUCCESS);
    std::vector<Real> candidates(static_cast<std::size_t>(num_cells));
    Real grid_time = 0.0;
    int completed = *completed_global_steps;
    int result = success;
    for (int step = 0; result == success && step < limit; ++step) {
        int failure = success;
#pragma omp parallel for schedule(dynamic, 1)
        for (int c = 0; c < num_cells; ++c) {
            int f = success;
            candidates[static_cast<std::size_t>(c)] =
                prepare_cell(cells[c], c, completed, grid_time, step, perturb != 0, &f);
            if (f != success) {
#pragma omp critical
                if (failure == success) failure = f;
            }
        }
        if (failure != success) {
            result = failure;
            break;
        }
        Real dt_grid = std::numeric_limits<Real>::max();
        for (const Real cand : candidates) {
            if (cand < dt_grid) {
                dt_grid = cand;
            }
        }
        if (dt_grid == std::numeric_limits<Real>::max()) {
            break;
        }
        const Real next_grid_time = grid_time + dt_grid;
        if (!valid_positive(dt_grid) || !std::isfinite(next_grid_time)) {
            result = static_cast<int>(IntegratorResult::BAD_INPUTS);
            break;
        }
        int integrated = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : integrated)
        for (int c = 0; c < num_cells; ++c) {
            int f = success;
            integrated += advance_cell(cells[c], completed, next_grid_time, dt_grid, &f);
            if (f != success) {
#pragma omp critical
                if (failure == success) failure = f;
            }
        }
        if (failure != success) {
            result = failure;
            break;
        }
        if (integrated == 0) {
            break;
        }
        grid_time = next_grid_time;
        completed += 1;
    }
    *completed_global_steps = completed;
    return result;
}

// Cells of a gridwide run replayed one by one through the run's dt_grid of
// every outer step: with the grid-wide minimum given, each cell's trajectory
// is the reference's prepare + advance kernel bodies (7706-7784) in the host
// loop's order (7859-7959), independent of the other cells.  index[k] is cell
// k's grid index; status[k] its first failure code (1 = success).  Runs on
// `threads` OpenMP threads (<= 0: the OpenMP default); returns the ROS2S
// attempts made (sum of internal_steps added).
long long ref_replay_cells(void* cells_v, const int* index, int m, const double* dt_hist, int nsteps,
                           int perturb, int* status, int threads) {
    auto* cells = static_cast<CollapseState*>(cells_v);
    const int success = static_cast<int>(IntegratorResult::SUCCESS);
    long long attempts = 0;
    const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) reduction(+ : attempts)
    for (int k = 0; k < m; ++k) {
        CollapseState& s = cells[k];
        const unsigned long long before = s.stats.internal_steps;
        Real t = 0.0;
        int f = success;
        for (int step = 0; step < nsteps && f == success; ++step) {
            prepare_cell(s, index[k], step, t, step, perturb != 0, &f);
            if (f != success) break;
            const Real next = t + dt_hist[step];
            advance_cell(s, step, next, dt_hist[step], &f);
            t = next;
        }
        status[k] = f;
        attempts += static_cast<long long>(s.stats.internal_steps - before);
    }
    return attempts;
}

// The reference's CPU driver loop (legacy-local policy, reproducer.cpp:8630-8648).
int ref_run_legacy(void* cells_v, int num_cells, int perturb, int* completed_global_steps) {
    auto* cells = static_cast<CollapseState*>(cells_v);
    IntegratorResult failure = IntegratorResult::SUCCESS;
    int completed = 0;
    for (int step = 0; step < max_collapse_steps; ++step) {
        bool all_stopped = true;
        for (int cell = 0; cell < num_cells; ++cell) {
            auto& state = cells[cell];
            if (state.completed_steps < completed) {
                continue;
            }
            const bool stopped = advance_collapse_step(state, cell, step, perturb != 0, failure);
            if (failure != IntegratorResult::SUCCESS) {
                break;
            }
            all_stopped = all_stopped && stopped;
        }
        if (failure != IntegratorResult::SUCCESS || all_stopped) {
            break;
        }
        completed += 1;
    }
    *completed_global_steps = completed;
    return static_cast<int>(failure);
}

void ref_make_packed_final_state(const void* cell, int index, int grid_dim, void* out) {
    const auto p = make_packed_final_state(*static_cast<const CollapseState*>(cell), index, grid_dim);
    std::memcpy(out, &p, sizeof(p));
}

// write_final_states / compare_final_states_from_file (reproducer.cpp:8357-8378,
// 8551-8578) on an array of CollapseState.
int ref_write_final_states(const void* cells_v, int num_cells, int grid_dim, const char* path) {
    const auto* cells = static_cast<const CollapseState*>(cells_v);
    std::vector<CollapseState> v(cells, cells + num_cells);
    return write_final_states(v, grid_dim, path) ? 1 : 0;
}

int ref_compare_final_states(const void* cells_v, int num_cells, int grid_dim, const char* path) {
    const auto* cells = static_cast<const CollapseState*>(cells_v);
    std::vector<CollapseState> v(cells, cells + num_cells);
    return compare_final_states_from_file(v, grid_dim, path) ? 1 : 0;
}

}  // extern "C"
