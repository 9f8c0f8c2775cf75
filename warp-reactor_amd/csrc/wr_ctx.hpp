// Device state shared by the library's translation units: the SoA cell
// state, the per-run control block on the device and the host-side context.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/warp_reactor.h"

namespace wr {

// SoA cell state on the device.
struct DevCells {
    double* xn;       // [14][cap]
    double* rho;      // [cap]
    double* T;
    double* e;
    double* time;
    double* driver;
    int* completed;
    unsigned long long* stats;  // [7][cap]
    int* cost;                  // ROS2S attempts of the cell's last advance
    int* order;                 // queue position -> cell (most attempts first)
    int n;
    int cap;
};

struct DevWs {
    double* J;   // [225][cap]
    double* LU;  // [225][cap]
    int cap;
};

// Device control block of one run.
struct Ctl {
    double grid_time;
    double dt_grid;
    double next_grid_time;
    unsigned long long cell_advances;
    int failure;
    int integrated;
    int completed;
    int step;
    int step_limit;  // the run stops once step reaches this
    int done;
    int perturb;
    int stop_reason;
    int failed_step;
    int failed_phase;
    int limiting_cell;
    int failed_cell;
    int step0;          // step index at the start of the run
    int hist_cap;       // capacity of dt_hist
    double* dt_hist;    // dt_grid of every outer step of the run (device)
    int queue;          // next queue position to hand out in the queue-fed advance
    int queue_solo;     // next of the positions reserved for the solo wave
    int queue_coop;     // next of the positions reserved for the cooperative waves
    int coop_cells;     // queue positions (most expensive cells first) advanced by whole waves
    int coop_waves;     // blocks whose wave 0 advances those cells
    int integ_max_steps;  // integrator step limit (configure_ros2s: 1e7)
    int lu_sweeps;        // refinement sweeps per stage solve (FP32-factorisation form)
    int solo_lanes;       // queue positions reserved for block 0's first wave (0: none)
    int jac_age_max;      // Jacobian reuse: accepted steps per J (1: the reference)
    unsigned long long mixed_lu_fallbacks;      // attempts whose FP32 factors were singular
    unsigned long long mixed_refine_fallbacks;  // stage solves whose refinement did not converge
    // legacy-local and gridwide-cta32 policies: (failing cell << 32 | code) of
    // the step's lowest failing cell (or tile's first cell), all ones when none
    // failed
    unsigned long long local_fail_key;
    // gridwide-cta32 policy: density-driver updates as (dt * driver) / tff
    int mojo_order;
    int bin_count[16];  // cells per attempt-count bin, for the queue order
    int bin_used[16];   // queue positions already handed to blocks, per bin
};

// First failure wins (reproducer.cpp:7716, 7724, 7782); the winner records its cell.
__device__ __forceinline__ void set_failure(Ctl* ctl, int code, int cell) {
    if (atomicCAS(&ctl->failure, 1, code) == 1) ctl->failed_cell = cell;
}

constexpr int kMaxStepsDefault = 10000000;  // configure_ros2s (reproducer.cpp:7589)

struct AdjointState;  // tapes and workspaces of the adjoint (wr_adjoint.hip)

}  // namespace wr

namespace wr {
namespace detail {

extern thread_local std::string g_last_error;

inline int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

inline bool hip_ok(hipError_t st, const char* action, std::string* msg = nullptr) {
    if (st == hipSuccess) return true;
    const std::string m = std::string(action) + " failed: " + hipGetErrorString(st);
    g_last_error = m;
    if (msg) *msg = m;
    return false;
}

inline int blocks_for(int n, int b) { return (n + b - 1) / b; }

}  // namespace detail
}  // namespace wr

#define WR_HIP(call, action)                                          \
    do {                                                              \
        if (!::wr::detail::hip_ok((call), (action))) return WR_ERR_HIP; \
    } while (0)

struct wr_context {
    int device = 0;
    int cap = 0;
    int n = 0;
    hipStream_t stream = nullptr;
    wr::DevCells cells{};
    wr::DevWs ws{};
    wr::Ctl* ctl = nullptr;         // device
    wr::Ctl* ctl_host = nullptr;    // pinned mirror
    static constexpr int kHistCap = 4096;
    double* dt_hist = nullptr;  // dt_grid per outer step of the last run
    int hist_len = 0;
    double* part_dt = nullptr;  // per-prepare-block partials
    int* part_cell = nullptr;
    wr_collapse_state* staging = nullptr;  // device AoS staging buffer
    void* cells_mem = nullptr;
    void* ws_mem = nullptr;
    // advance kernel: 1 (default) queue-fed persistent kernel, 0 one thread per cell
    int advance_kernel = 1;
    // factorisation of the queue-fed kernel: 0 FP64, 1 FP32 + FP64 refinement
    int lu_mode = 0;
    int lu_sweeps = 2;
    // Jacobian reuse of the queue-fed kernel: accepted steps per J (1: the reference)
    int jac_age_max = 1;
    int integ_max_steps = wr::kMaxStepsDefault;
    // launch shape of the queue-fed kernel (wr_set_queue_launch): threads per
    // block, resident blocks per CU (0: the occupancy limit), solo-wave lanes
    int q_threads = 256;
    int q_per_cu = 0;
    int q_solo = 64;
    int q_blocks = 0;
    // wave-cooperative advance of the most expensive cells (wr_set_coop):
    // queue positions given to whole waves (-1: by grid size), and how many
    // waves take them
    int coop_cells = -1;
    int coop_waves = 0;
    // HIP-event timing of the advance launches of wr_run (when enabled)
    static constexpr int kChunk = 8;
    bool timing = false;
    // one chunk of outer steps captured as a HIP graph (used when not timing)
    bool use_graph = true;
    hipGraphExec_t chunk_exec = nullptr;
    int graph_n = -1;             // what the captured launches were built for
    long long graph_kernel = -1;
    hipEvent_t adv_start[kChunk] = {}, adv_stop[kChunk] = {};
    double adv_ms = 0.0;
    int adv_launches = 0;
    // driver policy of wr_run: WR_POLICY_GRIDWIDE (default), WR_POLICY_LEGACY_LOCAL
    // or WR_POLICY_GRIDWIDE_CTA32
    int policy = 0;
    int cta_blocks = 0;  // persistent grid of the gridwide-cta32 kernel (0: not sized yet)
    int adj_perturb = 0;  // subsequent adjoint recordings apply the perturbations
    size_t cell_bytes = 0;           // bytes of cells_mem
    void* local_backup = nullptr;    // cells at the start of a legacy-local step
    int q_blocks_local = 0;          // persistent grid of the legacy-local kernel
    wr_run_info info{};
    wr::AdjointState* adj = nullptr;  // created on first use by the adjoint entry points
};

namespace wr {
namespace detail {
// Outer-step pieces of wr_kernels.hip used by the adjoint's recording run.
int ctl_init(wr_context* ctx, int perturb, int max_steps, int completed, double grid_time);
int ctl_read(wr_context* ctx);
int ctl_prepare(wr_context* ctx);  // k_prepare + k_plan of one outer step
int ctl_queue_order(wr_context* ctx);  // queue order of the queue-fed advance
int ctl_commit(wr_context* ctx);   // k_commit of one outer step
void ctl_fill_info(wr_context* ctx);
void adjoint_destroy(wr_context* ctx);  // wr_adjoint.hip
}  // namespace detail
}  // namespace wr


