// gfx950 kernels of the gridwide-independent collapse driver and the host-side
// context that runs them.  Replaces run_cells_cuda and its kernels
// (reproducer.cpp:7687-7975).
//
// Outer step (one HIP stream, no host round trip inside a step):
//   k_prepare  one thread per cell: perturbation, free-fall dt candidate,
//              terminal update, consistency checks (reproducer.cpp:7696-7734);
//              each block reduces its candidates to a (dt, lowest cell) pair.
//   k_plan     one block: grid-wide (dt, cell) minimum over the block pairs
//              (the host loop at 7888-7896, same lowest-index tie break),
//              next_grid_time and the termination tests (7894-7901).
//   k_advance  one thread per cell: common-dt density update, stop tests and
//              the cell's full adaptive ROS2S advance (7736-7785).
//   k_commit   one thread: failure / zero-integrated termination, grid time
//              and global step bookkeeping (7943-7958).
// The production advance is the queue-fed k_advance_q (wr_queue.hpp, with
// wave-cooperative cells from wr_coop.hpp); the Mojo front end's
// gridwide-cta32 policy replaces the advance by k_advance_cta (wr_cta.hpp).
// Every kernel returns immediately once the device control block says the
// run is done, so the host enqueues several steps ahead and polls one flag.
#include "wr_ctx.hpp"
#include "wr_device.hpp"
#include "wr_queue.hpp"
#include "wr_coop.hpp"
#include "wr_cta.hpp"

#include <float.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/warp_reactor.h"

namespace wr {

constexpr int kPrepBlock = 256;
constexpr int kAdvBlock = 64;
constexpr int kPlanBlock = 1024;


// (dt, cell) pair order: smaller dt, then smaller cell.
__device__ __forceinline__ bool pair_less(double da, int ca, double db, int cb) {
    return da < db || (da == db && ca < cb);
}

__device__ __forceinline__ void block_min_pair(double& dt, int& cell, double* sdt, int* scell) {
    // wave-level then block-level reduction through LDS
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double odt = __shfl_down(dt, off, 64);
        const int oc = __shfl_down(cell, off, 64);
        if (pair_less(odt, oc, dt, cell)) {
            dt = odt;
            cell = oc;
        }
    }
    if (lane == 0) {
        sdt[wave] = dt;
        scell[wave] = cell;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 1; w < nw; ++w) {
            if (pair_less(sdt[w], scell[w], dt, cell)) {
                dt = sdt[w];
                cell = scell[w];
            }
        }
    }
}

// ---- AoS <-> SoA ------------------------------------------------------------------
__global__ void k_scatter(const wr_collapse_state* __restrict__ in, DevCells s) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= s.n) return;
    const wr_collapse_state& a = in[c];
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) s.xn[n * s.cap + c] = a.current.xn[n];
    s.rho[c] = a.current.rho;
    s.T[c] = a.current.T;
    s.e[c] = a.current.e;
    s.time[c] = a.time;
    s.driver[c] = a.density_driver;
    s.completed[c] = a.completed_steps;
    s.cost[c] = 0;
    const unsigned long long* st = reinterpret_cast<const unsigned long long*>(&a.stats);
#pragma unroll
    for (int k = 0; k < 7; ++k) s.stats[k * s.cap + c] = st[k];
}

__global__ void k_gather(wr_collapse_state* __restrict__ out, DevCells s) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= s.n) return;
    wr_collapse_state& a = out[c];
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) a.current.xn[n] = s.xn[n * s.cap + c];
    a.current.rho = s.rho[c];
    a.current.T = s.T[c];
    a.current.e = s.e[c];
    a.time = s.time[c];
    a.density_driver = s.driver[c];
    a.completed_steps = s.completed[c];
    a.reserved_ = 0;
    unsigned long long* st = reinterpret_cast<unsigned long long*>(&a.stats);
#pragma unroll
    for (int k = 0; k < 7; ++k) st[k] = s.stats[k * s.cap + c];
}

// ---- outer step kernels ---------------------------------------------------------------

// prepare_grid_timestep_kernel (reproducer.cpp:7696-7734) + per-block (dt, cell) min.
__global__ __launch_bounds__(kPrepBlock) void k_prepare(DevCells s, Ctl* ctl, double* part_dt,
                                                        int* part_cell) {
    __shared__ double sdt[kPrepBlock / 64];
    __shared__ int scell[kPrepBlock / 64];
    if (ctl->done) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    double cand = DBL_MAX;
    int cand_cell = 0x7fffffff;
    const int completed = ctl->completed;
    if (c < s.n && ctl->failure == kSuccess) {
        const int done_steps = s.completed[c];
        if (done_steps >= completed) {
            if (done_steps != completed || s.time[c] != ctl->grid_time) {
                set_failure(ctl, kBadInputs, c);
            } else {
                double xn[kNumSpec];
#pragma unroll
                for (int n = 0; n < kNumSpec; ++n) xn[n] = s.xn[n * s.cap + c];
                double driver = s.driver[c];
                const int step = ctl->step;
                // apply_perturbation (reproducer.cpp:7564-7580)
                if (ctl->perturb && step != 0 && step % 20 == 0) {
                    const double f = perturbation_factor(c, step);
                    driver *= f;
                    double rho = s.rho[c] * f;
#pragma unroll
                    for (int n = 0; n < kNumSpec; ++n) xn[n] *= f;
                    const double T = post_step(xn, rho, s.e[c]);
#pragma unroll
                    for (int n = 0; n < kNumSpec; ++n) s.xn[n * s.cap + c] = xn[n];
                    s.rho[c] = rho;
                    s.T[c] = T;
                }
                const double dt = collapse_timestep(xn);
                if (!valid_positive(dt) || !valid_positive(driver)) {
                    set_failure(ctl, kBadInputs, c);
                } else if (dt < 10.0) {
                    const double tff = dt / 1.0e-1;
                    // gridwide-cta32 evaluates (dt * driver) / tff (reproducer.mojo:2858)
                    if (ctl->mojo_order)
                        driver += dt * driver / tff;
                    else
                        driver += dt * (driver / tff);
                } else {
                    cand = dt;
                    cand_cell = c;
                }
                s.driver[c] = driver;
            }
        }
    }
    block_min_pair(cand, cand_cell, sdt, scell);
    if (threadIdx.x == 0) {
        part_dt[blockIdx.x] = cand;
        part_cell[blockIdx.x] = cand_cell;
    }
}

// grid-wide minimum and the host-loop checks of reproducer.cpp:7881-7901.
__global__ __launch_bounds__(kPlanBlock) void k_plan(Ctl* ctl, const double* part_dt,
                                                     const int* part_cell, int nparts) {
    __shared__ double sdt[kPlanBlock / 64];
    __shared__ int scell[kPlanBlock / 64];
    if (ctl->done) return;
    double dt = DBL_MAX;
    int cell = 0x7fffffff;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        if (pair_less(part_dt[i], part_cell[i], dt, cell)) {
            dt = part_dt[i];
            cell = part_cell[i];
        }
    }
    block_min_pair(dt, cell, sdt, scell);
    if (threadIdx.x != 0) return;
    if (ctl->failure != kSuccess) {
        ctl->done = 1;
        ctl->stop_reason = WR_STOP_FAILURE;
        ctl->failed_step = ctl->step;
        ctl->failed_phase = 1;
        return;
    }
    if (dt == DBL_MAX) {
        ctl->done = 1;
        ctl->stop_reason = WR_STOP_NO_CANDIDATE;
        ctl->limiting_cell = -1;
        return;
    }
    const double next = ctl->grid_time + dt;
    if (!valid_positive(dt) || !isfinite(next)) {
        ctl->failure = kBadInputs;
        ctl->done = 1;
        ctl->stop_reason = WR_STOP_FAILURE;
        ctl->failed_step = ctl->step;
        ctl->failed_phase = 1;
        return;
    }
    ctl->dt_grid = dt;
    ctl->next_grid_time = next;
    ctl->limiting_cell = cell;
    ctl->integrated = 0;
    ctl->queue = 0;
    ctl->queue_solo = 0;
    ctl->queue_coop = 0;
    ctl->local_fail_key = ~0ull;
    const int k = ctl->step - ctl->step0;
    if (ctl->dt_hist && k >= 0 && k < ctl->hist_cap) ctl->dt_hist[k] = dt;
}

// The integrator as a real (non-inlined) function: its register allocation
// is separate from the callers', which keeps the kernels' own control flow
// and state out of the integrator's spill code.
__device__ __noinline__ int integrate_cell(double (&y)[kNeqs], double tout, double dt0,
                                           int max_steps, Counters& c, double* __restrict__ J,
                                           double* __restrict__ LU, int stride, double& t_end,
                                           double& dt_end) {
    return ros2s_integrate(y, tout, dt0, max_steps, c, J, LU, stride, t_end, dt_end);
}

// advance_collapse_gridwide_kernel (reproducer.cpp:7736-7785), one thread per cell.
__global__ __launch_bounds__(kAdvBlock) void k_advance(DevCells s, DevWs ws, Ctl* ctl) {
    if (ctl->done) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= s.n || ctl->failure != kSuccess) return;
    const int completed = ctl->completed;
    if (s.completed[c] != completed) return;
    double xn[kNumSpec];
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) xn[n] = s.xn[n * s.cap + c];
    const double local_dt = collapse_timestep(xn);
    if (local_dt < 10.0) return;
    const double dt_grid = ctl->dt_grid;
    const double old_density = s.driver[c];
    const double tff = local_dt / 1.0e-1;
    const double driver = old_density + dt_grid * (old_density / tff);
    s.driver[c] = driver;
    if (!valid_positive(driver) || driver > 2.0e-6) return;

    const double ratio = driver / old_density;
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) xn[n] *= ratio;
    const double rho = s.rho[c] * ratio;
    s.rho[c] = rho;

    // burn_ros2s (reproducer.cpp:7592-7627): eos_rt, y = (xn, e), integrate.
    const double e0 = eos_energy(xn, s.T[c]);
    double y[kNeqs];
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) y[n] = xn[n];
    y[kNumSpec] = e0;
    Counters cnt;
    double t_end, dt_end;
    const int result = integrate_cell(y, dt_grid, dt_grid, ctl->integ_max_steps, cnt, ws.J + c,
                                       ws.LU + c, ws.cap, t_end, dt_end);
    s.cost[c] = cnt.n_step;
    const int cv[7] = {cnt.n_step, cnt.n_rhs, cnt.n_jac, cnt.n_decomp, cnt.n_solve,
                       cnt.n_accept, cnt.n_reject};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        s.stats[k * s.cap + c] += static_cast<unsigned long long>(cv[k] > 0 ? cv[k] : 0);
    }
    if (result != kSuccess) {
        // burn_ros2s leaves the state untouched except e (eos_rt) on failure
        s.e[c] = e0;
#pragma unroll
        for (int n = 0; n < kNumSpec; ++n) s.xn[n * s.cap + c] = xn[n];
        set_failure(ctl, result, c);
        return;
    }
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) xn[n] = y[n];
    const double e = y[kNumSpec];
    const double T = post_step(xn, rho, e);
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) s.xn[n * s.cap + c] = xn[n];
    s.e[c] = e;
    s.T[c] = T;
    s.time[c] = ctl->next_grid_time;
    s.completed[c] = completed + 1;
    atomicAdd(&ctl->integrated, 1);
}

// ---- queue order ------------------------------------------------------------------
// Cells are handed out most-expensive first (longest-processing-time order):
// the attempt count of a cell's last advance predicts its next one, and the
// one or few cells that need tens of attempts then start at the beginning of
// the step instead of finishing long after the rest.  Counting sort over 16
// bins (15 = 15 or more attempts); the order within a bin is arbitrary.
// Results do not depend on the order: every cell's advance is independent.
constexpr int kCostBins = 16;
constexpr int kOrderBlock = 256;

__device__ __forceinline__ int cost_bin(int attempts) {
    return attempts < 0 ? 0 : (attempts < kCostBins - 1 ? attempts : kCostBins - 1);
}

__global__ __launch_bounds__(kOrderBlock) void k_order_count(DevCells s, Ctl* ctl) {
    __shared__ int h[kCostBins];
    if (ctl->done) return;
    if (threadIdx.x < kCostBins) h[threadIdx.x] = 0;
    __syncthreads();
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < s.n) atomicAdd(&h[cost_bin(s.cost[c])], 1);
    __syncthreads();
    if (threadIdx.x < kCostBins && h[threadIdx.x] != 0)
        atomicAdd(&ctl->bin_count[threadIdx.x], h[threadIdx.x]);
}

__global__ __launch_bounds__(kOrderBlock) void k_order_scatter(DevCells s, Ctl* ctl) {
    __shared__ int h[kCostBins];
    __shared__ int base[kCostBins];
    if (ctl->done) return;
    if (threadIdx.x < kCostBins) h[threadIdx.x] = 0;
    __syncthreads();
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    int b = 0, rank = 0;
    if (c < s.n) {
        b = cost_bin(s.cost[c]);
        rank = atomicAdd(&h[b], 1);
    }
    __syncthreads();
    if (threadIdx.x < kCostBins) {
        const int bb = threadIdx.x;
        int start = 0;  // higher bins come first
        for (int k = kCostBins - 1; k > bb; --k) start += ctl->bin_count[k];
        base[bb] = (h[bb] != 0) ? start + atomicAdd(&ctl->bin_used[bb], h[bb]) : 0;
    }
    __syncthreads();
    if (c < s.n) s.order[base[b] + rank] = c;
}

// ---- queue-fed advance: wr_queue.hpp -------------------------------------------------
// The production network: the generated staged J + f and f, no tape.
struct QPlainNet {
    static constexpr int kSlots = kStageSlots;
    // the most expensive cells of a step may be advanced by whole waves
    static constexpr bool kCoop = true;
    __device__ static void coop(DevCells s, Ctl* ctl, int lane) { coop_run(s, ctl, lane); }
    __device__ __forceinline__ void jac_rhs(const ThreadWs& w) { ws_jac_rhs(w); }
    __device__ __forceinline__ void rhs(const ThreadWs& w) { ws_rhs(w); }
    __device__ __forceinline__ void begin(int, const double (&)[kNumSpec], double, double, double,
                                          double) {}
    __device__ __forceinline__ void skipped(int) {}
    __device__ __forceinline__ void accepted(int, double) {}
    __device__ __forceinline__ void end(int) {}
};

__global__ void k_commit(Ctl* ctl) {
    if (ctl->done) return;
    for (int b = 0; b < kCostBins; ++b) {
        ctl->bin_count[b] = 0;
        ctl->bin_used[b] = 0;
    }
    // gridwide-cta32: the lowest failing tile (its first cell) is reported
    const unsigned long long key = ctl->local_fail_key;
    if (key != ~0ull) {
        ctl->failure = static_cast<int>(key & 0xffffffffull);
        ctl->failed_cell = static_cast<int>(key >> 32);
    }
    if (ctl->failure != kSuccess) {
        ctl->done = 1;
        ctl->stop_reason = WR_STOP_FAILURE;
        ctl->failed_step = ctl->step;
        ctl->failed_phase = 2;
        return;
    }
    if (ctl->integrated == 0) {
        ctl->done = 1;
        ctl->stop_reason = WR_STOP_NONE_INTEGRATED;
        return;
    }
    ctl->cell_advances += static_cast<unsigned long long>(ctl->integrated);
    ctl->grid_time = ctl->next_grid_time;
    ctl->completed += 1;
    ctl->step += 1;
    if (ctl->step >= ctl->step_limit) {
        ctl->done = 1;
        ctl->stop_reason = WR_STOP_STEP_LIMIT;
    }
}

// ---- legacy-local policy (reproducer.cpp:8628-8648) ----------------------------------
// One outer step: k_local_begin, a copy of the cells, the queue order,
// k_advance_q<.., kLocal>, k_local_restore, k_local_commit.
__global__ void k_local_begin(Ctl* ctl) {
    if (ctl->done) return;
    ctl->integrated = 0;
    ctl->queue = 0;
    ctl->queue_solo = 0;
    ctl->local_fail_key = ~0ull;
}

// The reference's cell loop stops at the first failing cell (8636-8638): the
// cells after it keep their state of the step's start, taken from the copy b.
__global__ __launch_bounds__(kOrderBlock) void k_local_restore(DevCells s, DevCells b,
                                                                const Ctl* ctl) {
    if (ctl->done) return;
    const unsigned long long key = ctl->local_fail_key;
    if (key == ~0ull) return;
    const int f = static_cast<int>(key >> 32);
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c <= f || c >= s.n) return;
    for (int n = 0; n < kNumSpec; ++n) s.xn[n * s.cap + c] = b.xn[n * b.cap + c];
    s.rho[c] = b.rho[c];
    s.T[c] = b.T[c];
    s.e[c] = b.e[c];
    s.time[c] = b.time[c];
    s.driver[c] = b.driver[c];
    s.completed[c] = b.completed[c];
    for (int k = 0; k < 7; ++k) s.stats[k * s.cap + c] = b.stats[k * b.cap + c];
}

// End of the outer step (8639-8642): failure, every cell stopped, or next step.
__global__ void k_local_commit(Ctl* ctl) {
    if (ctl->done) return;
    for (int b = 0; b < kCostBins; ++b) {
        ctl->bin_count[b] = 0;
        ctl->bin_used[b] = 0;
    }
    const unsigned long long key = ctl->local_fail_key;
    if (key != ~0ull) {
        ctl->failure = static_cast<int>(key & 0xffffffffull);
        ctl->failed_cell = static_cast<int>(key >> 32);
        ctl->done = 1;
        ctl->stop_reason = WR_STOP_FAILURE;
        ctl->failed_step = ctl->step;
        ctl->failed_phase = 2;
        return;
    }
    if (ctl->integrated == 0) {
        ctl->done = 1;
        ctl->stop_reason = WR_STOP_NONE_INTEGRATED;
        return;
    }
    ctl->cell_advances += static_cast<unsigned long long>(ctl->integrated);
    ctl->completed += 1;
    ctl->step += 1;
    if (ctl->step >= ctl->step_limit) {
        ctl->done = 1;
        ctl->stop_reason = WR_STOP_STEP_LIMIT;
    }
}

// ---- building-block kernels (verification entry points) ------------------------------
__global__ void k_eval_rhs(const double* __restrict__ y, int n, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double yy[kNeqs], f[kNeqs];
#pragma unroll
    for (int k = 0; k < kNeqs; ++k) yy[k] = y[i * kNeqs + k];
    rhs(yy, f);
#pragma unroll
    for (int k = 0; k < kNeqs; ++k) out[i * kNeqs + k] = f[k];
}

__global__ void k_eval_jac(const double* __restrict__ y, int n, DevWs ws, double* __restrict__ jac,
                           double* __restrict__ f_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double yy[kNeqs], f[kNeqs];
#pragma unroll
    for (int k = 0; k < kNeqs; ++k) yy[k] = y[i * kNeqs + k];
    jac_rhs(yy, f, ws.J + i, ws.cap);
    for (int k = 0; k < kNeqs * kNeqs; ++k) jac[i * kNeqs * kNeqs + k] = ws.J[k * ws.cap + i];
#pragma unroll
    for (int k = 0; k < kNeqs; ++k) f_out[i * kNeqs + k] = f[k];
}

__global__ void k_eval_lu(double* __restrict__ a, int n, DevWs ws, int* __restrict__ ipvt_out,
                          int* __restrict__ info_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < kNeqs * kNeqs; ++k) ws.LU[k * ws.cap + i] = a[i * kNeqs * kNeqs + k];
    int ipvt[kNeqs];
    info_out[i] = lu_factor(ws.LU + i, ws.cap, ipvt);
    for (int k = 0; k < kNeqs * kNeqs; ++k) a[i * kNeqs * kNeqs + k] = ws.LU[k * ws.cap + i];
#pragma unroll
    for (int k = 0; k < kNeqs; ++k) ipvt_out[i * kNeqs + k] = ipvt[k];
}

__global__ void k_eval_solve(const double* __restrict__ lu, const int* __restrict__ ipvt_in, int n,
                             DevWs ws, double* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < kNeqs * kNeqs; ++k) ws.LU[k * ws.cap + i] = lu[i * kNeqs * kNeqs + k];
    int ipvt[kNeqs];
    double xx[kNeqs];
#pragma unroll
    for (int k = 0; k < kNeqs; ++k) {
        ipvt[k] = ipvt_in[i * kNeqs + k];
        xx[k] = x[i * kNeqs + k];
    }
    lu_solve(ws.LU + i, ws.cap, ipvt, xx);
#pragma unroll
    for (int k = 0; k < kNeqs; ++k) x[i * kNeqs + k] = xx[k];
}

// One stage solve of the FP32-factorisation form per thread (the queue-fed
// kernel's functions): E = fac I - J equilibrated and factorised in FP32, then
// x solved with `sweeps` refinement sweeps.  conv: -1 FP32 factors singular,
// else whether the refinement converged; lu_out: the FP32 factors.
__global__ void k_eval_solve_mixed(const double* __restrict__ jac, const double* __restrict__ fac,
                                   const double* __restrict__ y, int n, int sweeps, int zero,
                                   double* __restrict__ x, int* __restrict__ conv,
                                   float* __restrict__ lu_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double priv[kNeqs * kNeqs + kVecLen];
    double* const p = priv + zero;
    float priv32[kNeqs * kNeqs];
    float* const lu32 = priv32 + zero;
    ThreadWs w;
    w.J = p;
    w.LU = nullptr;
    w.vec = p + kNeqs * kNeqs;
    w.stage = nullptr;
    w.stride = 1;
    for (int k = 0; k < kNeqs * kNeqs; ++k) w.J[k] = jac[static_cast<size_t>(i) * kNeqs * kNeqs + k];
    for (int k = 0; k < kNeqs; ++k) w.vec[kVecY + k] = y[i * kNeqs + k];
    int ks[kNeqs];
    mixed_scales(w, ks);
    uint64_t piv32 = 0;
    const int info = lu_left_t<float>(w.J, lu32, 1, fac[i], piv32, ks);
    for (int k = 0; k < kNeqs * kNeqs; ++k) lu_out[static_cast<size_t>(i) * kNeqs * kNeqs + k] = lu32[k];
    double v[kNeqs];
#pragma unroll
    for (int k = 0; k < kNeqs; ++k) v[k] = x[i * kNeqs + k];
    int c = -1;
    if (info == 0) c = solve_mixed(w, lu32, piv32, ks, fac[i], sweeps, v) ? 1 : 0;
#pragma unroll
    for (int k = 0; k < kNeqs; ++k) x[i * kNeqs + k] = v[k];
    conv[i] = c;
}

// Device libm on a batch: 0 exp, 1 log, 2 cbrt, 3 sqrt as the network uses
// them (wr_math.h / IEEE sqrt), 4..6 the ROCm device library's exp, log, cbrt.
__global__ void k_eval_math(int op, const double* __restrict__ x, int n, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = x[i];
    double r;
    switch (op) {
        case 0: r = wr_exp(v); break;
        case 1: r = wr_log(v); break;
        case 2: r = wr_cbrt(v); break;
        case 3: r = sqrt(v); break;
        case 4: r = exp(v); break;
        case 5: r = log(v); break;
        default: r = cbrt(v); break;
    }
    out[i] = r;
}

__global__ __launch_bounds__(kAdvBlock) void k_eval_integrate(double* __restrict__ y,
                                                              const double* __restrict__ tout,
                                                              const double* __restrict__ dt0, int n,
                                                              int max_steps, DevWs ws,
                                                              int* __restrict__ counters,
                                                              double* __restrict__ t_end,
                                                              double* __restrict__ dt_end,
                                                              int* __restrict__ result) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double yy[kNeqs];
#pragma unroll
    for (int k = 0; k < kNeqs; ++k) yy[k] = y[i * kNeqs + k];
    Counters c;
    double te, de;
    const int r = integrate_cell(yy, tout[i], dt0[i], max_steps, c, ws.J + i, ws.LU + i, ws.cap,
                                  te, de);
#pragma unroll
    for (int k = 0; k < kNeqs; ++k) y[i * kNeqs + k] = yy[k];
    const int cv[7] = {c.n_step, c.n_rhs, c.n_jac, c.n_decomp, c.n_solve, c.n_accept, c.n_reject};
#pragma unroll
    for (int k = 0; k < 7; ++k) counters[i * 7 + k] = cv[k];
    t_end[i] = te;
    dt_end[i] = de;
    result[i] = r;
}

}  // namespace wr

// =====================================================================================
// Host side: context and C ABI
// =====================================================================================
using namespace wr;

namespace wr {
namespace detail {
thread_local std::string g_last_error;
}  // namespace detail
}  // namespace wr
using namespace wr::detail;

extern "C" {

int wr_abi_version(void) { return WR_ABI_VERSION; }

const char* wr_last_error(void) { return g_last_error.c_str(); }

void wr_context_destroy(wr_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->cells_mem);
    (void)hipFree(ctx->local_backup);
    (void)hipFree(ctx->ws_mem);
    (void)hipFree(ctx->ctl);
    (void)hipFree(ctx->part_dt);
    (void)hipFree(ctx->part_cell);
    (void)hipFree(ctx->dt_hist);
    (void)hipFree(ctx->staging);
    if (ctx->ctl_host) (void)hipHostFree(ctx->ctl_host);
    if (ctx->chunk_exec) (void)hipGraphExecDestroy(ctx->chunk_exec);
    for (int k = 0; k < wr_context::kChunk; ++k) {
        if (ctx->adv_start[k]) (void)hipEventDestroy(ctx->adv_start[k]);
        if (ctx->adv_stop[k]) (void)hipEventDestroy(ctx->adv_stop[k]);
    }
    wr::detail::adjoint_destroy(ctx);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int wr_context_create(int device, int capacity_cells, wr_context** out) {
    if (!out || capacity_cells <= 0) return fail(WR_ERR_ARGUMENT, "wr_context_create: bad argument");
    *out = nullptr;
    auto* ctx = new wr_context();
    ctx->device = device;
    // keep every SoA column 256-byte aligned
    ctx->cap = (capacity_cells + 63) / 64 * 64;
    const size_t cap = static_cast<size_t>(ctx->cap);
    auto cleanup = [&](int code) {
        wr_context_destroy(ctx);
        return code;
    };
    if (!hip_ok(hipSetDevice(device), "hipSetDevice")) return cleanup(WR_ERR_HIP);
    if (!hip_ok(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate"))
        return cleanup(WR_ERR_HIP);
    const size_t cell_bytes = cap * (sizeof(double) * (kNumSpec + 5) + sizeof(int) * 3 +
                                     sizeof(unsigned long long) * 7) + 4096;
    ctx->cell_bytes = cell_bytes;
    if (!hip_ok(hipMalloc(&ctx->cells_mem, cell_bytes), "hipMalloc(cells)"))
        return cleanup(WR_ERR_HIP);
    char* p = static_cast<char*>(ctx->cells_mem);
    auto take = [&](size_t bytes) {
        char* q = p;
        p += (bytes + 255) / 256 * 256;
        return q;
    };
    ctx->cells.xn = reinterpret_cast<double*>(take(cap * kNumSpec * sizeof(double)));
    ctx->cells.rho = reinterpret_cast<double*>(take(cap * sizeof(double)));
    ctx->cells.T = reinterpret_cast<double*>(take(cap * sizeof(double)));
    ctx->cells.e = reinterpret_cast<double*>(take(cap * sizeof(double)));
    ctx->cells.time = reinterpret_cast<double*>(take(cap * sizeof(double)));
    ctx->cells.driver = reinterpret_cast<double*>(take(cap * sizeof(double)));
    ctx->cells.stats = reinterpret_cast<unsigned long long*>(take(cap * 7 * sizeof(unsigned long long)));
    ctx->cells.completed = reinterpret_cast<int*>(take(cap * sizeof(int)));
    ctx->cells.cost = reinterpret_cast<int*>(take(cap * sizeof(int)));
    ctx->cells.order = reinterpret_cast<int*>(take(cap * sizeof(int)));
    ctx->cells.cap = ctx->cap;
    ctx->cells.n = 0;
    const size_t ws_bytes = cap * kNeqs * kNeqs * sizeof(double) * 2;
    if (!hip_ok(hipMalloc(&ctx->ws_mem, ws_bytes), "hipMalloc(workspace)")) return cleanup(WR_ERR_HIP);
    ctx->ws.J = static_cast<double*>(ctx->ws_mem);
    ctx->ws.LU = ctx->ws.J + cap * kNeqs * kNeqs;
    ctx->ws.cap = ctx->cap;
    const int nparts = blocks_for(ctx->cap, kPrepBlock);
    if (!hip_ok(hipMalloc(&ctx->ctl, sizeof(Ctl)), "hipMalloc(ctl)") ||
        !hip_ok(hipHostMalloc(&ctx->ctl_host, sizeof(Ctl)), "hipHostMalloc(ctl)") ||
        !hip_ok(hipMalloc(&ctx->part_dt, nparts * sizeof(double)), "hipMalloc(partials)") ||
        !hip_ok(hipMalloc(&ctx->part_cell, nparts * sizeof(int)), "hipMalloc(partials)") ||
        !hip_ok(hipMalloc(&ctx->dt_hist, wr_context::kHistCap * sizeof(double)), "hipMalloc(dt history)") ||
        !hip_ok(hipMalloc(&ctx->staging, cap * sizeof(wr_collapse_state)), "hipMalloc(staging)"))
        return cleanup(WR_ERR_HIP);
    for (int k = 0; k < wr_context::kChunk; ++k) {
        if (!hip_ok(hipEventCreate(&ctx->adv_start[k]), "hipEventCreate") ||
            !hip_ok(hipEventCreate(&ctx->adv_stop[k]), "hipEventCreate"))
            return cleanup(WR_ERR_HIP);
    }
    memset(ctx->ctl_host, 0, sizeof(Ctl));
    *out = ctx;
    return WR_SUCCESS;
}

int wr_upload_cells(wr_context* ctx, const wr_collapse_state* cells, int num_cells) {
    if (!ctx || (!cells && num_cells > 0) || num_cells < 0)
        return fail(WR_ERR_ARGUMENT, "wr_upload_cells: bad argument");
    if (num_cells > ctx->cap) return fail(WR_ERR_CAPACITY, "wr_upload_cells: more cells than capacity");
    WR_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    ctx->n = num_cells;
    ctx->cells.n = num_cells;
    if (num_cells == 0) return WR_SUCCESS;
    // host or device memory: the copy kind follows from unified addressing
    WR_HIP(hipMemcpyAsync(ctx->staging, cells, num_cells * sizeof(wr_collapse_state),
                          hipMemcpyDefault, ctx->stream),
           "hipMemcpy(cells to device)");
    hipLaunchKernelGGL(k_scatter, dim3(blocks_for(num_cells, 256)), dim3(256), 0, ctx->stream,
                       ctx->staging, ctx->cells);
    WR_HIP(hipGetLastError(), "k_scatter launch");
    WR_HIP(hipStreamSynchronize(ctx->stream), "k_scatter synchronize");
    return WR_SUCCESS;
}

int wr_download_cells(wr_context* ctx, wr_collapse_state* cells, int num_cells) {
    if (!ctx || (!cells && num_cells > 0) || num_cells < 0 || num_cells > ctx->n)
        return fail(WR_ERR_ARGUMENT, "wr_download_cells: bad argument");
    if (num_cells == 0) return WR_SUCCESS;
    WR_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    hipLaunchKernelGGL(k_gather, dim3(blocks_for(num_cells, 256)), dim3(256), 0, ctx->stream,
                       ctx->staging, ctx->cells);
    WR_HIP(hipGetLastError(), "k_gather launch");
    WR_HIP(hipMemcpyAsync(cells, ctx->staging, num_cells * sizeof(wr_collapse_state),
                          hipMemcpyDefault, ctx->stream),
           "hipMemcpy(cells to host)");
    WR_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return WR_SUCCESS;
}

// Cooperative cells and waves of the context's grid: the setting, or (cells
// = -1, the default) one cooperative cell per 2,048 cells, at most 128, with
// as many waves.  On the 64^3 run 8 to 128 cooperative cells all take
// 2.71-2.81 s against 4.39 s without; grids under 2,048 cells keep every
// cell on the single-lane path.
static void coop_resolved(const wr_context* ctx, int* cells, int* waves) {
    if (ctx->coop_cells >= 0) {
        *cells = ctx->coop_cells;
        *waves = ctx->coop_waves;
        return;
    }
    const int c = ctx->n / 2048 < kCoopAutoMax ? ctx->n / 2048 : kCoopAutoMax;
    *cells = c;
    *waves = c;
}

// Whether the wave-cooperative advance applies to subsequent launches: the
// queue-fed FP64 kernel with the reference's Jacobian rule, gridwide policy.
static bool coop_active(const wr_context* ctx) {
    int cells = 0, waves = 0;
    coop_resolved(ctx, &cells, &waves);
    return cells > 0 && waves > 0 && ctx->advance_kernel == 1 && ctx->lu_mode == 0 &&
           ctx->jac_age_max == 1 && ctx->policy == WR_POLICY_GRIDWIDE;
}

// first_step: the outer-loop index of the run's first step, which the
// perturbation schedule reads (reproducer.cpp:7859 starts it at 0 whatever
// completed count the caller passes); it advances with the completed count.
static int init_ctl(wr_context* ctx, int perturb, int max_steps, int completed, double grid_time,
                    int first_step) {
    Ctl h{};
    h.grid_time = grid_time;
    h.failure = kSuccess;
    h.completed = completed;
    h.step = first_step;
    h.step_limit = first_step + max_steps;
    h.failed_cell = -1;
    h.step0 = first_step;
    h.hist_cap = wr_context::kHistCap;
    h.dt_hist = ctx->dt_hist;
    h.done = (max_steps <= 0) ? 1 : 0;
    h.perturb = perturb ? 1 : 0;
    h.stop_reason = (max_steps <= 0) ? WR_STOP_STEP_LIMIT : WR_STOP_NONE;
    h.failed_step = -1;
    h.limiting_cell = -1;
    h.integ_max_steps = ctx->integ_max_steps;
    h.lu_sweeps = ctx->lu_sweeps;
    h.solo_lanes = ctx->q_solo;
    h.jac_age_max = ctx->jac_age_max;
    h.local_fail_key = ~0ull;
    h.mojo_order = ctx->policy == WR_POLICY_GRIDWIDE_CTA32 ? 1 : 0;
    if (coop_active(ctx)) coop_resolved(ctx, &h.coop_cells, &h.coop_waves);
    *ctx->ctl_host = h;
    WR_HIP(hipMemcpyAsync(ctx->ctl, ctx->ctl_host, sizeof(Ctl), hipMemcpyHostToDevice, ctx->stream),
           "hipMemcpy(control to device)");
    WR_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return WR_SUCCESS;
}

static int read_ctl(wr_context* ctx) {
    WR_HIP(hipMemcpyAsync(ctx->ctl_host, ctx->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, ctx->stream),
           "hipMemcpy(control to host)");
    WR_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return WR_SUCCESS;
}

static int launch_prepare(wr_context* ctx) {
    const int nb = blocks_for(ctx->n, kPrepBlock);
    hipLaunchKernelGGL(k_prepare, dim3(nb), dim3(kPrepBlock), 0, ctx->stream, ctx->cells, ctx->ctl,
                       ctx->part_dt, ctx->part_cell);
    WR_HIP(hipGetLastError(), "prepare_grid_timestep_kernel launch");
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(kPlanBlock), 0, ctx->stream, ctx->ctl, ctx->part_dt,
                       ctx->part_cell, nb);
    WR_HIP(hipGetLastError(), "grid timestep reduction launch");
    return WR_SUCCESS;
}

// Workspace of the queue-fed kernel, sized for the persistent grid: the
// resident blocks of k_advance_q on every CU (at most the blocks that the
// cells fill).
// Persistent grid of the queue-fed kernel: its resident blocks on every CU
// (at most the blocks the cells fill).
static int queue_grid(wr_context* ctx) {
    if (ctx->q_blocks > 0) return WR_SUCCESS;
    int cus = 0, per_cu = 0;
    WR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device),
           "hipDeviceGetAttribute(CU count)");
    WR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
               &per_cu, ctx->lu_mode ? k_advance_q<true, QPlainNet> : k_advance_q<false, QPlainNet>,
               ctx->q_threads, 0),
           "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (per_cu < 1) per_cu = 1;
    if (ctx->q_per_cu > 0) per_cu = ctx->q_per_cu;
    int blocks = cus * per_cu;
    const int need = blocks_for(ctx->cap, ctx->q_threads);
    if (blocks > need) blocks = need;
    ctx->q_blocks = blocks < 1 ? 1 : blocks;
    return WR_SUCCESS;
}

// Persistent grid of the gridwide-cta32 kernel: every resident wave slot (at
// most one wave per 64 cells).
static int cta_grid(wr_context* ctx) {
    if (ctx->cta_blocks > 0) return WR_SUCCESS;
    int cus = 0, per_cu = 0;
    WR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device),
           "hipDeviceGetAttribute(CU count)");
    WR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_advance_cta, kCtaBlock, 0),
           "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    const int blocks = cus * (per_cu < 1 ? 1 : per_cu);
    const int need = blocks_for(ctx->cap, kCtaBlock);
    ctx->cta_blocks = blocks > need ? need : (blocks < 1 ? 1 : blocks);
    return WR_SUCCESS;
}

static int launch_advance(wr_context* ctx) {
    if (ctx->lu_mode != 0 && ctx->advance_kernel != 1)
        return fail(WR_ERR_ARGUMENT,
                    "FP32 factorisation (wr_set_lu_precision) needs the queue-fed advance kernel");
    if (ctx->jac_age_max > 1 && (ctx->advance_kernel != 1 || ctx->lu_mode != 0))
        return fail(WR_ERR_ARGUMENT,
                    "Jacobian reuse (wr_set_jacobian_reuse) needs the queue-fed advance kernel "
                    "with FP64 factors");
    if (ctx->policy == WR_POLICY_GRIDWIDE_CTA32) {
        if (ctx->lu_mode != 0 || ctx->jac_age_max > 1)
            return fail(WR_ERR_ARGUMENT,
                        "the gridwide-cta32 policy runs the reference integrator "
                        "(wr_set_lu_precision 0, wr_set_jacobian_reuse 1)");
        const int st = cta_grid(ctx);
        if (st != WR_SUCCESS) return st;
        hipLaunchKernelGGL(k_advance_cta, dim3(ctx->cta_blocks), dim3(kCtaBlock), 0, ctx->stream,
                           ctx->cells, ctx->ctl, 0);
    } else if (ctx->advance_kernel == 1) {
        const int st = queue_grid(ctx);
        if (st != WR_SUCCESS) return st;
        const int nb = blocks_for(ctx->n, kOrderBlock);
        hipLaunchKernelGGL(k_order_count, dim3(nb), dim3(kOrderBlock), 0, ctx->stream, ctx->cells,
                           ctx->ctl);
        hipLaunchKernelGGL(k_order_scatter, dim3(nb), dim3(kOrderBlock), 0, ctx->stream, ctx->cells,
                           ctx->ctl);
        if (ctx->lu_mode)
            hipLaunchKernelGGL((k_advance_q<true, QPlainNet>), dim3(ctx->q_blocks), dim3(ctx->q_threads), 0,
                               ctx->stream, ctx->cells, ctx->ctl, 0, QPlainNet{});
        else if (ctx->jac_age_max > 1)
            hipLaunchKernelGGL((k_advance_q<false, QPlainNet, false, true>), dim3(ctx->q_blocks),
                               dim3(ctx->q_threads), 0, ctx->stream, ctx->cells, ctx->ctl, 0,
                               QPlainNet{});
        else
            hipLaunchKernelGGL((k_advance_q<false, QPlainNet>), dim3(ctx->q_blocks), dim3(ctx->q_threads),
                               0, ctx->stream, ctx->cells, ctx->ctl, 0, QPlainNet{});
    } else {
        hipLaunchKernelGGL(k_advance, dim3(blocks_for(ctx->n, kAdvBlock)), dim3(kAdvBlock), 0,
                           ctx->stream, ctx->cells, ctx->ws, ctx->ctl);
    }
    WR_HIP(hipGetLastError(), "advance_collapse_gridwide_kernel launch");
    return WR_SUCCESS;
}

static void fill_info(wr_context* ctx) {
    const Ctl& h = *ctx->ctl_host;
    wr_run_info& i = ctx->info;
    i.status = h.failure;
    i.stop_reason = h.stop_reason;
    i.completed_global_steps = h.completed;
    i.failed_step = h.failed_step;
    i.failed_phase = h.failed_phase;
    i.failed_cell = h.failed_cell;
    i.limiting_cell = h.limiting_cell;
    i.reserved_ = 0;
    i.grid_time = h.grid_time;
    i.cell_advances = h.cell_advances;
    // steps whose dt_grid was planned: completed steps, plus one if the run
    // ended in an advance phase
    int len = h.step - h.step0;
    if (h.stop_reason == WR_STOP_NONE_INTEGRATED ||
        (h.stop_reason == WR_STOP_FAILURE && h.failed_phase == 2))
        len += 1;
    ctx->hist_len = len < 0 ? 0 : (len > wr_context::kHistCap ? wr_context::kHistCap : len);
}

// Enqueue `todo` outer steps: prepare, plan, (order,) advance, commit each.
static int enqueue_steps(wr_context* ctx, int todo, bool timed) {
    int st;
    for (int k = 0; k < todo; ++k) {
        if ((st = launch_prepare(ctx)) != WR_SUCCESS) return st;
        if (timed) WR_HIP(hipEventRecord(ctx->adv_start[k], ctx->stream), "hipEventRecord");
        if ((st = launch_advance(ctx)) != WR_SUCCESS) return st;
        if (timed) WR_HIP(hipEventRecord(ctx->adv_stop[k], ctx->stream), "hipEventRecord");
        hipLaunchKernelGGL(k_commit, dim3(1), dim3(1), 0, ctx->stream, ctx->ctl);
        WR_HIP(hipGetLastError(), "commit launch");
    }
    return WR_SUCCESS;
}

// A full chunk of outer steps as one HIP graph: the kernels' arguments are
// fixed for a context (the control block carries everything that changes),
// except the cell count and the advance kernel, so the graph is captured
// again only when those change.  The graph holds kernel nodes only.
static int chunk_graph(wr_context* ctx) {
    const long long kernel_key =
        ctx->advance_kernel + 2 * ctx->lu_mode + 4 * (ctx->jac_age_max > 1) +
        8LL * (ctx->q_threads + 1024LL * (ctx->q_per_cu + 64LL * (ctx->q_solo + 128LL * coop_active(ctx)))) +
        (1LL << 40) * ctx->policy;
    if (ctx->chunk_exec && ctx->graph_n == ctx->n && ctx->graph_kernel == kernel_key)
        return WR_SUCCESS;
    if (ctx->chunk_exec) {
        WR_HIP(hipGraphExecDestroy(ctx->chunk_exec), "hipGraphExecDestroy");
        ctx->chunk_exec = nullptr;
    }
    int st = queue_grid(ctx);  // occupancy queries stay outside the capture
    if (st == WR_SUCCESS && ctx->policy == WR_POLICY_GRIDWIDE_CTA32) st = cta_grid(ctx);
    if (st != WR_SUCCESS) return st;
    hipGraph_t g = nullptr;
    WR_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal),
           "hipStreamBeginCapture");
    st = enqueue_steps(ctx, wr_context::kChunk, false);
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (st != WR_SUCCESS) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    WR_HIP(e, "hipStreamEndCapture");
    const hipError_t ie = hipGraphInstantiate(&ctx->chunk_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    WR_HIP(ie, "hipGraphInstantiate");
    ctx->graph_n = ctx->n;
    ctx->graph_kernel = kernel_key;
    return WR_SUCCESS;
}

// The cells' arrays at the same offsets in another allocation of cells_mem's size.
static DevCells cells_at(const wr_context* ctx, void* base) {
    const auto move = [&](auto* q) {
        const char* p = reinterpret_cast<const char*>(q);
        return reinterpret_cast<decltype(q)>(static_cast<char*>(base) +
                                             (p - static_cast<const char*>(ctx->cells_mem)));
    };
    DevCells b = ctx->cells;
    b.xn = move(b.xn);
    b.rho = move(b.rho);
    b.T = move(b.T);
    b.e = move(b.e);
    b.time = move(b.time);
    b.driver = move(b.driver);
    b.completed = move(b.completed);
    b.stats = move(b.stats);
    b.cost = move(b.cost);
    b.order = move(b.order);
    return b;
}

// Outer loop of the legacy-local policy, 8 steps per host poll (no graph: the
// step copies the cells with hipMemcpyAsync).
static int run_legacy(wr_context* ctx, int max_steps) {
    if (!ctx->local_backup &&
        !hip_ok(hipMalloc(&ctx->local_backup, ctx->cell_bytes), "hipMalloc(legacy-local copy)"))
        return WR_ERR_HIP;
    if (ctx->q_blocks_local == 0) {
        int cus = 0, per_cu = 0;
        WR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device),
               "hipDeviceGetAttribute(CU count)");
        WR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                   &per_cu, k_advance_q<false, QPlainNet, true>, ctx->q_threads, 0),
               "hipOccupancyMaxActiveBlocksPerMultiprocessor");
        if (per_cu < 1) per_cu = 1;
        if (ctx->q_per_cu > 0) per_cu = ctx->q_per_cu;
        int blocks = cus * per_cu;
        const int need = blocks_for(ctx->cap, ctx->q_threads);
        if (blocks > need) blocks = need;
        ctx->q_blocks_local = blocks < 1 ? 1 : blocks;
    }
    const DevCells backup = cells_at(ctx, ctx->local_backup);
    const int nb = blocks_for(ctx->n, kOrderBlock);
    const int chunk = wr_context::kChunk;
    for (int issued = 0; issued < max_steps;) {
        const int todo = (max_steps - issued < chunk) ? max_steps - issued : chunk;
        const int steps_before = ctx->ctl_host->step;
        for (int k = 0; k < todo; ++k) {
            hipLaunchKernelGGL(k_local_begin, dim3(1), dim3(1), 0, ctx->stream, ctx->ctl);
            WR_HIP(hipMemcpyAsync(ctx->local_backup, ctx->cells_mem, ctx->cell_bytes,
                                  hipMemcpyDeviceToDevice, ctx->stream),
                   "hipMemcpyAsync(legacy-local copy)");
            hipLaunchKernelGGL(k_order_count, dim3(nb), dim3(kOrderBlock), 0, ctx->stream,
                               ctx->cells, ctx->ctl);
            hipLaunchKernelGGL(k_order_scatter, dim3(nb), dim3(kOrderBlock), 0, ctx->stream,
                               ctx->cells, ctx->ctl);
            if (ctx->timing)
                WR_HIP(hipEventRecord(ctx->adv_start[k], ctx->stream), "hipEventRecord");
            hipLaunchKernelGGL((k_advance_q<false, QPlainNet, true>), dim3(ctx->q_blocks_local),
                               dim3(ctx->q_threads), 0, ctx->stream, ctx->cells, ctx->ctl, 0,
                               QPlainNet{});
            if (ctx->timing)
                WR_HIP(hipEventRecord(ctx->adv_stop[k], ctx->stream), "hipEventRecord");
            hipLaunchKernelGGL(k_local_restore, dim3(nb), dim3(kOrderBlock), 0, ctx->stream,
                               ctx->cells, backup, ctx->ctl);
            hipLaunchKernelGGL(k_local_commit, dim3(1), dim3(1), 0, ctx->stream, ctx->ctl);
            WR_HIP(hipGetLastError(), "legacy-local step launch");
        }
        issued += todo;
        int st;
        if ((st = read_ctl(ctx)) != WR_SUCCESS) return st;
        if (ctx->timing) {
            const Ctl& h = *ctx->ctl_host;
            int ran = h.step - steps_before;
            if (h.done && (h.stop_reason == WR_STOP_NONE_INTEGRATED ||
                           (h.stop_reason == WR_STOP_FAILURE && h.failed_phase == 2)))
                ran += 1;
            for (int k = 0; k < ran && k < todo; ++k) {
                float ms = 0.0f;
                WR_HIP(hipEventElapsedTime(&ms, ctx->adv_start[k], ctx->adv_stop[k]),
                       "hipEventElapsedTime");
                ctx->adv_ms += ms;
                ctx->adv_launches += 1;
            }
        }
        if (ctx->ctl_host->done) break;
    }
    fill_info(ctx);
    ctx->hist_len = 0;
    return ctx->ctl_host->failure;
}

int wr_run(wr_context* ctx, int perturb, int max_steps, int completed_global_steps,
           double grid_time) {
    return wr_run_from(ctx, perturb, max_steps, completed_global_steps, grid_time,
                       completed_global_steps);
}

int wr_run_from(wr_context* ctx, int perturb, int max_steps, int completed_global_steps,
                double grid_time, int first_step) {
    if (!ctx) return fail(WR_ERR_ARGUMENT, "wr_run: null context");
    if (first_step < 0) return fail(WR_ERR_ARGUMENT, "wr_run_from: negative first step");
    if (max_steps <= 0) max_steps = 1000;
    WR_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    if (ctx->policy == WR_POLICY_LEGACY_LOCAL && (ctx->lu_mode != 0 || ctx->jac_age_max > 1))
        return fail(WR_ERR_ARGUMENT,
                    "wr_run: the legacy-local policy runs the reference integrator "
                    "(wr_set_lu_precision 0, wr_set_jacobian_reuse 1)");
    int st = init_ctl(ctx, perturb, ctx->n > 0 ? max_steps : 0, completed_global_steps, grid_time,
                      first_step);
    if (st != WR_SUCCESS) return st;
    ctx->adv_ms = 0.0;
    ctx->adv_launches = 0;
    if (ctx->n == 0) {
        ctx->ctl_host->stop_reason = WR_STOP_NO_CANDIDATE;
        fill_info(ctx);
        return WR_SUCCESS;
    }
    if (ctx->policy == WR_POLICY_LEGACY_LOCAL) return run_legacy(ctx, max_steps);
    // Steps are enqueued in chunks; kernels of steps past the end return at once.
    const int chunk = wr_context::kChunk;
    for (int issued = 0; issued < max_steps;) {
        const int todo = (max_steps - issued < chunk) ? max_steps - issued : chunk;
        const int steps_before = ctx->ctl_host->step;
        if (ctx->use_graph && !ctx->timing) {
            // steps past max_steps in the last chunk return at once (step_limit)
            if ((st = chunk_graph(ctx)) != WR_SUCCESS) return st;
            WR_HIP(hipGraphLaunch(ctx->chunk_exec, ctx->stream), "hipGraphLaunch");
        } else if ((st = enqueue_steps(ctx, todo, ctx->timing)) != WR_SUCCESS) {
            return st;
        }
        issued += todo;
        if ((st = read_ctl(ctx)) != WR_SUCCESS) return st;
        if (ctx->timing) {
            // advance launches that did work: completed steps of this chunk, plus
            // the last one when the run ended in its advance phase
            const Ctl& h = *ctx->ctl_host;
            int ran = h.step - steps_before;
            if (h.done && (h.stop_reason == WR_STOP_NONE_INTEGRATED ||
                           (h.stop_reason == WR_STOP_FAILURE && h.failed_phase == 2)))
                ran += 1;
            for (int k = 0; k < ran && k < todo; ++k) {
                float ms = 0.0f;
                WR_HIP(hipEventElapsedTime(&ms, ctx->adv_start[k], ctx->adv_stop[k]),
                       "hipEventElapsedTime");
                ctx->adv_ms += ms;
                ctx->adv_launches += 1;
            }
        }
        if (ctx->ctl_host->done) break;
    }
    fill_info(ctx);
    return ctx->ctl_host->failure;
}

int wr_dt_history(const wr_context* ctx, double* out, int max_entries, int* count) {
    if (!ctx || !count || (max_entries > 0 && !out))
        return fail(WR_ERR_ARGUMENT, "wr_dt_history: bad argument");
    const int n = ctx->hist_len < max_entries ? ctx->hist_len : max_entries;
    *count = ctx->hist_len;
    if (n > 0) {
        WR_HIP(hipSetDevice(ctx->device), "hipSetDevice");
        WR_HIP(hipMemcpy(out, ctx->dt_hist, n * sizeof(double), hipMemcpyDeviceToHost),
               "hipMemcpy(dt history)");
    }
    return WR_SUCCESS;
}

int wr_set_policy(wr_context* ctx, int policy) {
    if (!ctx || (policy != WR_POLICY_GRIDWIDE && policy != WR_POLICY_LEGACY_LOCAL &&
                 policy != WR_POLICY_GRIDWIDE_CTA32))
        return fail(WR_ERR_ARGUMENT,
                    "wr_set_policy: policy must be 0 (gridwide), 1 (legacy-local) or 2 "
                    "(gridwide-cta32)");
    ctx->policy = policy;
    return WR_SUCCESS;
}

int wr_set_jacobian_reuse(wr_context* ctx, int age_max) {
    if (!ctx || age_max < 1 || age_max > 64)
        return fail(WR_ERR_ARGUMENT, "wr_set_jacobian_reuse: age_max must be 1..64");
    ctx->jac_age_max = age_max;
    return WR_SUCCESS;
}

int wr_set_advance_kernel(wr_context* ctx, int kernel) {
    if (!ctx || kernel < 0 || kernel > 1)
        return fail(WR_ERR_ARGUMENT, "wr_set_advance_kernel: kernel must be 0 or 1");
    ctx->advance_kernel = kernel;
    return WR_SUCCESS;
}

int wr_set_lu_precision(wr_context* ctx, int mode, int sweeps) {
    if (!ctx || mode < 0 || mode > 1 || (mode == 1 && (sweeps < 1 || sweeps > 4)))
        return fail(WR_ERR_ARGUMENT,
                    "wr_set_lu_precision: mode must be 0 or 1, sweeps 1..4 for mode 1");
    if (mode != ctx->lu_mode) ctx->q_blocks = 0;  // occupancy of the other kernel
    ctx->lu_mode = mode;
    if (mode == 1) ctx->lu_sweeps = sweeps;
    return WR_SUCCESS;
}

int wr_set_queue_launch(wr_context* ctx, int threads_per_block, int blocks_per_cu,
                        int solo_lanes) {
    if (!ctx || threads_per_block < 64 || threads_per_block > kQBlock || threads_per_block % 64 ||
        blocks_per_cu < 0 || blocks_per_cu > 16 || (solo_lanes != 0 && solo_lanes != 64) ||
        solo_lanes > threads_per_block)
        return fail(WR_ERR_ARGUMENT,
                    "wr_set_queue_launch: threads 64..256 in steps of 64, blocks per CU 0..16, "
                    "solo lanes 0 or 64");
    ctx->q_threads = threads_per_block;
    ctx->q_per_cu = blocks_per_cu;
    ctx->q_solo = solo_lanes;
    ctx->q_blocks = 0;
    ctx->q_blocks_local = 0;
    return WR_SUCCESS;
}

int wr_set_coop(wr_context* ctx, int cells, int waves) {
    if (!ctx || cells < -1 || waves < 0 || cells > (1 << 20) || waves > 4096)
        return fail(WR_ERR_ARGUMENT, "wr_set_coop: cells -1..2^20, waves 0..4096");
    ctx->coop_cells = cells;
    ctx->coop_waves = waves;
    ctx->q_blocks = 0;  // occupancy with the cooperative LDS
    return WR_SUCCESS;
}

int wr_get_coop(const wr_context* ctx, int* cells, int* waves, int* lds_bytes) {
    if (!ctx || !cells || !waves) return fail(WR_ERR_ARGUMENT, "wr_get_coop: bad argument");
    coop_resolved(ctx, cells, waves);
    if (lds_bytes) *lds_bytes = COOP_LDS_BYTES;
    return WR_SUCCESS;
}

int wr_get_queue_launch(const wr_context* ctx, int* threads_per_block, int* blocks_per_cu,
                        int* solo_lanes, int* blocks) {
    if (!ctx || !threads_per_block || !blocks_per_cu || !solo_lanes || !blocks)
        return fail(WR_ERR_ARGUMENT, "wr_get_queue_launch: bad argument");
    *threads_per_block = ctx->q_threads;
    *blocks_per_cu = ctx->q_per_cu;
    *solo_lanes = ctx->q_solo;
    *blocks = ctx->q_blocks;
    return WR_SUCCESS;
}

int wr_mixed_fallbacks(const wr_context* ctx, unsigned long long* lu_fallbacks,
                       unsigned long long* refine_fallbacks) {
    if (!ctx || !lu_fallbacks || !refine_fallbacks)
        return fail(WR_ERR_ARGUMENT, "wr_mixed_fallbacks: bad argument");
    *lu_fallbacks = ctx->ctl_host->mixed_lu_fallbacks;
    *refine_fallbacks = ctx->ctl_host->mixed_refine_fallbacks;
    return WR_SUCCESS;
}

int wr_set_integrator_max_steps(wr_context* ctx, int max_steps) {
    if (!ctx) return fail(WR_ERR_ARGUMENT, "wr_set_integrator_max_steps: null context");
    ctx->integ_max_steps = max_steps > 0 ? max_steps : kMaxStepsDefault;
    return WR_SUCCESS;
}

int wr_set_graph(wr_context* ctx, int enable) {
    if (!ctx) return fail(WR_ERR_ARGUMENT, "wr_set_graph: null context");
    ctx->use_graph = enable != 0;
    return WR_SUCCESS;
}

int wr_set_timing(wr_context* ctx, int enable) {
    if (!ctx) return fail(WR_ERR_ARGUMENT, "wr_set_timing: null context");
    ctx->timing = enable != 0;
    return WR_SUCCESS;
}

int wr_run_info_get(const wr_context* ctx, wr_run_info* out) {
    if (!ctx || !out) return fail(WR_ERR_ARGUMENT, "wr_run_info_get: bad argument");
    *out = ctx->info;
    return WR_SUCCESS;
}

int wr_last_advance_timing(const wr_context* ctx, double* total_ms, int* launches) {
    if (!ctx) return fail(WR_ERR_ARGUMENT, "wr_last_advance_timing: null context");
    if (total_ms) *total_ms = ctx->adv_ms;
    if (launches) *launches = ctx->adv_launches;
    return WR_SUCCESS;
}

int wr_prepare(wr_context* ctx, int completed_global_steps, double grid_time, int step, int perturb,
               double* dt_grid, int* limiting_cell) {
    if (!ctx || !dt_grid) return fail(WR_ERR_ARGUMENT, "wr_prepare: bad argument");
    WR_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    int st = init_ctl(ctx, perturb, 1, completed_global_steps, grid_time, step);
    if (st != WR_SUCCESS) return st;
    if (ctx->n == 0) {
        *dt_grid = DBL_MAX;
        if (limiting_cell) *limiting_cell = -1;
        return WR_SUCCESS;
    }
    if ((st = launch_prepare(ctx)) != WR_SUCCESS) return st;
    if ((st = read_ctl(ctx)) != WR_SUCCESS) return st;
    const Ctl& h = *ctx->ctl_host;
    if (h.failure != kSuccess) {
        *dt_grid = DBL_MAX;
        if (limiting_cell) *limiting_cell = -1;
        return h.failure;
    }
    if (h.stop_reason == WR_STOP_NO_CANDIDATE) {
        *dt_grid = DBL_MAX;
        if (limiting_cell) *limiting_cell = -1;
        return WR_SUCCESS;
    }
    *dt_grid = h.dt_grid;
    if (limiting_cell) *limiting_cell = h.limiting_cell;
    return WR_SUCCESS;
}

int wr_advance(wr_context* ctx, int completed_global_steps, double next_grid_time, double dt_grid,
               int* integrated) {
    if (!ctx) return fail(WR_ERR_ARGUMENT, "wr_advance: null context");
    WR_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    int st = init_ctl(ctx, 0, 1, completed_global_steps, 0.0, 0);
    if (st != WR_SUCCESS) return st;
    ctx->ctl_host->dt_grid = dt_grid;
    ctx->ctl_host->next_grid_time = next_grid_time;
    ctx->ctl_host->integrated = 0;
    WR_HIP(hipMemcpyAsync(ctx->ctl, ctx->ctl_host, sizeof(Ctl), hipMemcpyHostToDevice, ctx->stream),
           "hipMemcpy(control to device)");
    if (ctx->n > 0 && (st = launch_advance(ctx)) != WR_SUCCESS) return st;
    if ((st = read_ctl(ctx)) != WR_SUCCESS) return st;
    if (integrated) *integrated = ctx->ctl_host->integrated;
    // gridwide-cta32: the lowest failing tile's code
    const unsigned long long key = ctx->ctl_host->local_fail_key;
    if (key != ~0ull) return static_cast<int>(key & 0xffffffffull);
    return ctx->ctl_host->failure;
}

// Drop-in for run_cells_cuda (reproducer.cpp:7791-7975).
int wr_run_cells(wr_collapse_state* cells, int num_cells, int perturb, int* completed_global_steps) {
    if (!cells || num_cells <= 0 || !completed_global_steps) {
        fprintf(stderr, "wr_run_cells: bad argument\n");
        return WR_BAD_INPUTS;
    }
    wr_context* ctx = nullptr;
    int st = wr_context_create(0, num_cells, &ctx);
    if (st == WR_SUCCESS) st = wr_upload_cells(ctx, cells, num_cells);
    if (st != WR_SUCCESS) {
        fprintf(stderr, "%s\n", g_last_error.c_str());
        wr_context_destroy(ctx);
        return WR_BAD_INPUTS;
    }
    // the reference's loop index starts at 0 (reproducer.cpp:7859)
    st = wr_run_from(ctx, perturb, 1000, *completed_global_steps, 0.0, 0);
    int result = WR_SUCCESS;
    if (st <= WR_ERR_HIP) {
        fprintf(stderr, "%s\n", g_last_error.c_str());
        result = WR_BAD_INPUTS;
    } else {
        const wr_run_info& i = ctx->info;
        *completed_global_steps = i.completed_global_steps;
        if (i.status != WR_SUCCESS) {
            result = i.status;
            if (i.failed_phase == 1)
                fprintf(stderr, "grid timestep preparation failed on hip collapse step %d with code %d\n",
                        i.failed_step, i.status);
            else
                fprintf(stderr, "ROS2S failed on hip collapse step %d with code %d\n", i.failed_step,
                        i.status);
        }
    }
    const int dl = wr_download_cells(ctx, cells, num_cells);
    if (dl != WR_SUCCESS) {
        fprintf(stderr, "%s\n", g_last_error.c_str());
        if (result == WR_SUCCESS) result = WR_BAD_INPUTS;
    }
    wr_context_destroy(ctx);
    return result;
}

// ---- batched verification entry points ---------------------------------------------------
namespace {

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

int eval_common(wr_context* ctx, int n) {
    if (!ctx || n < 0) return fail(WR_ERR_ARGUMENT, "eval: bad argument");
    if (n > ctx->cap) return fail(WR_ERR_CAPACITY, "eval: more items than context capacity");
    WR_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    return WR_SUCCESS;
}

}  // namespace

int wr_eval_rhs(wr_context* ctx, const double* y, int n, double* dydt) {
    int st = eval_common(ctx, n);
    if (st != WR_SUCCESS || n == 0) return st;
    DevBuf dy, dout;
    const size_t b = static_cast<size_t>(n) * kNeqs * sizeof(double);
    WR_HIP(hipMalloc(&dy.p, b), "hipMalloc");
    WR_HIP(hipMalloc(&dout.p, b), "hipMalloc");
    WR_HIP(hipMemcpy(dy.p, y, b, hipMemcpyHostToDevice), "hipMemcpy");
    hipLaunchKernelGGL(k_eval_rhs, dim3(blocks_for(n, 64)), dim3(64), 0, ctx->stream,
                       static_cast<double*>(dy.p), n, static_cast<double*>(dout.p));
    WR_HIP(hipGetLastError(), "k_eval_rhs launch");
    WR_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    WR_HIP(hipMemcpy(dydt, dout.p, b, hipMemcpyDeviceToHost), "hipMemcpy");
    return WR_SUCCESS;
}

int wr_eval_jac(wr_context* ctx, const double* y, int n, double* jac) {
    int st = eval_common(ctx, n);
    if (st != WR_SUCCESS || n == 0) return st;
    DevBuf dy, dj, df;
    const size_t b = static_cast<size_t>(n) * kNeqs * sizeof(double);
    WR_HIP(hipMalloc(&dy.p, b), "hipMalloc");
    WR_HIP(hipMalloc(&df.p, b), "hipMalloc");
    WR_HIP(hipMalloc(&dj.p, b * kNeqs), "hipMalloc");
    WR_HIP(hipMemcpy(dy.p, y, b, hipMemcpyHostToDevice), "hipMemcpy");
    hipLaunchKernelGGL(k_eval_jac, dim3(blocks_for(n, 64)), dim3(64), 0, ctx->stream,
                       static_cast<double*>(dy.p), n, ctx->ws, static_cast<double*>(dj.p),
                       static_cast<double*>(df.p));
    WR_HIP(hipGetLastError(), "k_eval_jac launch");
    WR_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    WR_HIP(hipMemcpy(jac, dj.p, b * kNeqs, hipMemcpyDeviceToHost), "hipMemcpy");
    return WR_SUCCESS;
}

int wr_eval_math(wr_context* ctx, int op, const double* x, int n, double* out) {
    int st = eval_common(ctx, n);
    if (st != WR_SUCCESS || n == 0) return st;
    if (op < 0 || op > 6) return fail(WR_ERR_ARGUMENT, "wr_eval_math: op must be 0..6");
    DevBuf dx, dout;
    const size_t b = static_cast<size_t>(n) * sizeof(double);
    WR_HIP(hipMalloc(&dx.p, b), "hipMalloc");
    WR_HIP(hipMalloc(&dout.p, b), "hipMalloc");
    WR_HIP(hipMemcpy(dx.p, x, b, hipMemcpyHostToDevice), "hipMemcpy");
    hipLaunchKernelGGL(k_eval_math, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, op,
                       static_cast<double*>(dx.p), n, static_cast<double*>(dout.p));
    WR_HIP(hipGetLastError(), "k_eval_math launch");
    WR_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    WR_HIP(hipMemcpy(out, dout.p, b, hipMemcpyDeviceToHost), "hipMemcpy");
    return WR_SUCCESS;
}

int wr_eval_lu(wr_context* ctx, double* a, int n, int32_t* ipvt, int32_t* info) {
    int st = eval_common(ctx, n);
    if (st != WR_SUCCESS || n == 0) return st;
    DevBuf da, dp, di;
    const size_t b = static_cast<size_t>(n) * kNeqs * kNeqs * sizeof(double);
    WR_HIP(hipMalloc(&da.p, b), "hipMalloc");
    WR_HIP(hipMalloc(&dp.p, static_cast<size_t>(n) * kNeqs * sizeof(int)), "hipMalloc");
    WR_HIP(hipMalloc(&di.p, static_cast<size_t>(n) * sizeof(int)), "hipMalloc");
    WR_HIP(hipMemcpy(da.p, a, b, hipMemcpyHostToDevice), "hipMemcpy");
    hipLaunchKernelGGL(k_eval_lu, dim3(blocks_for(n, 64)), dim3(64), 0, ctx->stream,
                       static_cast<double*>(da.p), n, ctx->ws, static_cast<int*>(dp.p),
                       static_cast<int*>(di.p));
    WR_HIP(hipGetLastError(), "k_eval_lu launch");
    WR_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    WR_HIP(hipMemcpy(a, da.p, b, hipMemcpyDeviceToHost), "hipMemcpy");
    WR_HIP(hipMemcpy(ipvt, dp.p, static_cast<size_t>(n) * kNeqs * sizeof(int), hipMemcpyDeviceToHost),
           "hipMemcpy");
    WR_HIP(hipMemcpy(info, di.p, static_cast<size_t>(n) * sizeof(int), hipMemcpyDeviceToHost),
           "hipMemcpy");
    return WR_SUCCESS;
}

int wr_eval_lu_solve(wr_context* ctx, const double* lu, const int32_t* ipvt, int n, double* x) {
    int st = eval_common(ctx, n);
    if (st != WR_SUCCESS || n == 0) return st;
    DevBuf dl, dp, dx;
    const size_t b = static_cast<size_t>(n) * kNeqs * kNeqs * sizeof(double);
    const size_t bx = static_cast<size_t>(n) * kNeqs * sizeof(double);
    WR_HIP(hipMalloc(&dl.p, b), "hipMalloc");
    WR_HIP(hipMalloc(&dp.p, static_cast<size_t>(n) * kNeqs * sizeof(int)), "hipMalloc");
    WR_HIP(hipMalloc(&dx.p, bx), "hipMalloc");
    WR_HIP(hipMemcpy(dl.p, lu, b, hipMemcpyHostToDevice), "hipMemcpy");
    WR_HIP(hipMemcpy(dp.p, ipvt, static_cast<size_t>(n) * kNeqs * sizeof(int), hipMemcpyHostToDevice),
           "hipMemcpy");
    WR_HIP(hipMemcpy(dx.p, x, bx, hipMemcpyHostToDevice), "hipMemcpy");
    hipLaunchKernelGGL(k_eval_solve, dim3(blocks_for(n, 64)), dim3(64), 0, ctx->stream,
                       static_cast<double*>(dl.p), static_cast<int*>(dp.p), n, ctx->ws,
                       static_cast<double*>(dx.p));
    WR_HIP(hipGetLastError(), "k_eval_solve launch");
    WR_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    WR_HIP(hipMemcpy(x, dx.p, bx, hipMemcpyDeviceToHost), "hipMemcpy");
    return WR_SUCCESS;
}

int wr_eval_solve_mixed(wr_context* ctx, const double* jac, const double* fac, const double* y,
                        int n, int sweeps, double* x, int32_t* conv, float* lu32) {
    int st = eval_common(ctx, n);
    if (st != WR_SUCCESS || n == 0) return st;
    if (!jac || !fac || !y || !x || !conv || !lu32 || sweeps < 0 || sweeps > 4)
        return fail(WR_ERR_ARGUMENT, "wr_eval_solve_mixed: bad argument");
    DevBuf dj, df, dy, dx, dc, dl;
    const size_t bj = static_cast<size_t>(n) * kNeqs * kNeqs * sizeof(double);
    const size_t bv = static_cast<size_t>(n) * kNeqs * sizeof(double);
    WR_HIP(hipMalloc(&dj.p, bj), "hipMalloc");
    WR_HIP(hipMalloc(&df.p, static_cast<size_t>(n) * sizeof(double)), "hipMalloc");
    WR_HIP(hipMalloc(&dy.p, bv), "hipMalloc");
    WR_HIP(hipMalloc(&dx.p, bv), "hipMalloc");
    WR_HIP(hipMalloc(&dc.p, static_cast<size_t>(n) * sizeof(int)), "hipMalloc");
    WR_HIP(hipMalloc(&dl.p, static_cast<size_t>(n) * kNeqs * kNeqs * sizeof(float)), "hipMalloc");
    WR_HIP(hipMemcpy(dj.p, jac, bj, hipMemcpyHostToDevice), "hipMemcpy");
    WR_HIP(hipMemcpy(df.p, fac, static_cast<size_t>(n) * sizeof(double), hipMemcpyHostToDevice),
           "hipMemcpy");
    WR_HIP(hipMemcpy(dy.p, y, bv, hipMemcpyHostToDevice), "hipMemcpy");
    WR_HIP(hipMemcpy(dx.p, x, bv, hipMemcpyHostToDevice), "hipMemcpy");
    hipLaunchKernelGGL(k_eval_solve_mixed, dim3(blocks_for(n, 64)), dim3(64), 0, ctx->stream,
                       static_cast<double*>(dj.p), static_cast<double*>(df.p),
                       static_cast<double*>(dy.p), n, sweeps, 0, static_cast<double*>(dx.p),
                       static_cast<int*>(dc.p), static_cast<float*>(dl.p));
    WR_HIP(hipGetLastError(), "k_eval_solve_mixed launch");
    WR_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    WR_HIP(hipMemcpy(x, dx.p, bv, hipMemcpyDeviceToHost), "hipMemcpy");
    WR_HIP(hipMemcpy(conv, dc.p, static_cast<size_t>(n) * sizeof(int), hipMemcpyDeviceToHost),
           "hipMemcpy");
    WR_HIP(hipMemcpy(lu32, dl.p, static_cast<size_t>(n) * kNeqs * kNeqs * sizeof(float),
                     hipMemcpyDeviceToHost),
           "hipMemcpy");
    return WR_SUCCESS;
}

int wr_eval_integrate(wr_context* ctx, double* y, const double* tout, const double* dt0, int n,
                      int max_steps, int32_t* counters, double* t_end, double* dt_end,
                      int32_t* result) {
    int st = eval_common(ctx, n);
    if (st != WR_SUCCESS || n == 0) return st;
    if (max_steps <= 0) max_steps = 10000000;
    DevBuf dy, dto, ddt, dc, dte, dde, dr;
    const size_t bn = static_cast<size_t>(n) * sizeof(double);
    WR_HIP(hipMalloc(&dy.p, bn * kNeqs), "hipMalloc");
    WR_HIP(hipMalloc(&dto.p, bn), "hipMalloc");
    WR_HIP(hipMalloc(&ddt.p, bn), "hipMalloc");
    WR_HIP(hipMalloc(&dc.p, static_cast<size_t>(n) * 7 * sizeof(int)), "hipMalloc");
    WR_HIP(hipMalloc(&dte.p, bn), "hipMalloc");
    WR_HIP(hipMalloc(&dde.p, bn), "hipMalloc");
    WR_HIP(hipMalloc(&dr.p, static_cast<size_t>(n) * sizeof(int)), "hipMalloc");
    WR_HIP(hipMemcpy(dy.p, y, bn * kNeqs, hipMemcpyHostToDevice), "hipMemcpy");
    WR_HIP(hipMemcpy(dto.p, tout, bn, hipMemcpyHostToDevice), "hipMemcpy");
    WR_HIP(hipMemcpy(ddt.p, dt0, bn, hipMemcpyHostToDevice), "hipMemcpy");
    hipLaunchKernelGGL(k_eval_integrate, dim3(blocks_for(n, kAdvBlock)), dim3(kAdvBlock), 0,
                       ctx->stream, static_cast<double*>(dy.p), static_cast<double*>(dto.p),
                       static_cast<double*>(ddt.p), n, max_steps, ctx->ws, static_cast<int*>(dc.p),
                       static_cast<double*>(dte.p), static_cast<double*>(dde.p),
                       static_cast<int*>(dr.p));
    WR_HIP(hipGetLastError(), "k_eval_integrate launch");
    WR_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    WR_HIP(hipMemcpy(y, dy.p, bn * kNeqs, hipMemcpyDeviceToHost), "hipMemcpy");
    WR_HIP(hipMemcpy(counters, dc.p, static_cast<size_t>(n) * 7 * sizeof(int), hipMemcpyDeviceToHost),
           "hipMemcpy");
    WR_HIP(hipMemcpy(t_end, dte.p, bn, hipMemcpyDeviceToHost), "hipMemcpy");
    WR_HIP(hipMemcpy(dt_end, dde.p, bn, hipMemcpyDeviceToHost), "hipMemcpy");
    WR_HIP(hipMemcpy(result, dr.p, static_cast<size_t>(n) * sizeof(int), hipMemcpyDeviceToHost),
           "hipMemcpy");
    return WR_SUCCESS;
}

}  // extern "C"

// ---- outer-step pieces for the adjoint's recording run (wr_adjoint.hip) ----------------
namespace wr {
namespace detail {
int ctl_init(wr_context* ctx, int perturb, int max_steps, int completed, double grid_time) {
    return ::init_ctl(ctx, perturb, max_steps, completed, grid_time, completed);
}
int ctl_read(wr_context* ctx) { return ::read_ctl(ctx); }
int ctl_prepare(wr_context* ctx) { return ::launch_prepare(ctx); }
int ctl_queue_order(wr_context* ctx) {
    const int nb = blocks_for(ctx->n, kOrderBlock);
    hipLaunchKernelGGL(k_order_count, dim3(nb), dim3(kOrderBlock), 0, ctx->stream, ctx->cells,
                       ctx->ctl);
    hipLaunchKernelGGL(k_order_scatter, dim3(nb), dim3(kOrderBlock), 0, ctx->stream, ctx->cells,
                       ctx->ctl);
    WR_HIP(hipGetLastError(), "queue order launch");
    return WR_SUCCESS;
}
int ctl_commit(wr_context* ctx) {
    hipLaunchKernelGGL(k_commit, dim3(1), dim3(1), 0, ctx->stream, ctx->ctl);
    WR_HIP(hipGetLastError(), "commit launch");
    return WR_SUCCESS;
}
void ctl_fill_info(wr_context* ctx) { ::fill_info(ctx); }
}  // namespace detail
}  // namespace wr
