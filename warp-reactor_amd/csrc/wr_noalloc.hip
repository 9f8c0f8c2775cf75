// The prepare phase of one outer step -- prepare_grid_timestep_kernel
// (reproducer.cpp:7696-7734) on every cell, then the host loop's grid-wide
// minimum of the candidates (7888-7893) -- for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  It works on the caller's device array of
// CollapseState records (the layout the reference's kernels take) and writes
// (dt_grid, lowest cell holding it, failure code) to caller-given device
// memory; the per-block partial minima and the failure word live in a
// file-scope device array of fixed size instead of context memory.  The
// results equal wr_prepare's on the same cells.
#include <hip/hip_runtime.h>

#include <float.h>

#include <mutex>
#include <string>

#include "wr_ctx.hpp"
#include "wr_eos.hpp"

using namespace wr;

namespace {

constexpr int kNaBlock = 256;                // cells per prepare block
constexpr int kNaMaxCells = 1 << 24;         // largest grid a call accepts
constexpr int kNaMaxParts = kNaMaxCells / kNaBlock;
constexpr size_t kNaPartOff = 256;           // failure word, then the partials
constexpr size_t kNaBytes = kNaPartOff + size_t(kNaMaxParts) * (sizeof(double) + sizeof(int));
constexpr int kNaMaxDevices = 64;

// The library's no-allocation scratch (768 KiB + 256 B per GPU, held until the
// process ends): [0, 4) failure word, [256, ...) per-block dt partials, then
// per-block cell partials.  Host code uses only the addresses that
// hipGetSymbolAddress returns for it (scratch_ptr).
__device__ __align__(256) unsigned char scratch[kNaBytes];

std::mutex g_na_mutex;
unsigned char* g_na_ptr[kNaMaxDevices] = {};

__device__ __forceinline__ bool na_less(double da, int ca, double db, int cb) {
    return da < db || (da == db && ca < cb);
}

// (dt, cell) minimum over the block, the result on thread 0.
__device__ __forceinline__ void na_block_min(double& dt, int& cell) {
    __shared__ double sdt[kNaBlock / 64];
    __shared__ int scell[kNaBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double odt = __shfl_down(dt, off, 64);
        const int oc = __shfl_down(cell, off, 64);
        if (na_less(odt, oc, dt, cell)) {
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
        for (int w = 1; w < static_cast<int>(blockDim.x >> 6); ++w) {
            if (na_less(sdt[w], scell[w], dt, cell)) {
                dt = sdt[w];
                cell = scell[w];
            }
        }
    }
}

// The failure word starts each call at success (a kernel, not a memset, so
// that the call can be captured in a graph).
__global__ void k_na_begin(unsigned char* scr) { *reinterpret_cast<int*>(scr) = kSuccess; }

// prepare_grid_timestep_kernel on record c, and the block's (dt, cell) minimum.
__global__ __launch_bounds__(kNaBlock) void k_na_prepare(wr_collapse_state* __restrict__ cells,
                                                         int n, int completed, double grid_time,
                                                         int step, int perturb,
                                                         unsigned char* scr) {
    int* failure = reinterpret_cast<int*>(scr);
    double* part_dt = reinterpret_cast<double*>(scr + kNaPartOff);
    int* part_cell = reinterpret_cast<int*>(scr + kNaPartOff + sizeof(double) * kNaMaxParts);
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    double cand = DBL_MAX;
    int cand_cell = 0x7fffffff;
    if (c < n && *reinterpret_cast<volatile int*>(failure) == kSuccess) {
        wr_collapse_state& s = cells[c];
        if (s.completed_steps >= completed) {
            if (s.completed_steps != completed || s.time != grid_time) {
                atomicCAS(failure, kSuccess, kBadInputs);
            } else {
                double xn[kNumSpec];
#pragma unroll
                for (int k = 0; k < kNumSpec; ++k) xn[k] = s.current.xn[k];
                double driver = s.density_driver;
                if (perturb && step != 0 && step % 20 == 0) {  // apply_perturbation (7564-7580)
                    const double f = perturbation_factor(c, step);
                    driver *= f;
                    const double rho = s.current.rho * f;
#pragma unroll
                    for (int k = 0; k < kNumSpec; ++k) xn[k] *= f;
                    const double T = post_step(xn, rho, s.current.e);
#pragma unroll
                    for (int k = 0; k < kNumSpec; ++k) s.current.xn[k] = xn[k];
                    s.current.rho = rho;
                    s.current.T = T;
                }
                const double dt = collapse_timestep(xn);
                if (!valid_positive(dt) || !valid_positive(driver)) {
                    atomicCAS(failure, kSuccess, kBadInputs);
                } else if (dt < 10.0) {
                    const double tff = dt / 1.0e-1;
                    driver += dt * (driver / tff);
                } else {
                    cand = dt;
                    cand_cell = c;
                }
                s.density_driver = driver;
            }
        }
    }
    na_block_min(cand, cand_cell);
    if (threadIdx.x == 0) {
        part_dt[blockIdx.x] = cand;
        part_cell[blockIdx.x] = cand_cell;
    }
}

// The grid-wide minimum of the partials and the outputs.
__global__ __launch_bounds__(kNaBlock) void k_na_reduce(const unsigned char* scr, int nparts,
                                                        double* dt_grid, int* limiting_cell,
                                                        int* failure_out) {
    const double* part_dt = reinterpret_cast<const double*>(scr + kNaPartOff);
    const int* part_cell =
        reinterpret_cast<const int*>(scr + kNaPartOff + sizeof(double) * kNaMaxParts);
    double dt = DBL_MAX;
    int cell = 0x7fffffff;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        if (na_less(part_dt[i], part_cell[i], dt, cell)) {
            dt = part_dt[i];
            cell = part_cell[i];
        }
    }
    na_block_min(dt, cell);
    if (threadIdx.x != 0) return;
    const int f = *reinterpret_cast<const int*>(scr);
    *failure_out = f;
    if (f != kSuccess || dt == DBL_MAX) {
        *dt_grid = DBL_MAX;
        *limiting_cell = -1;
    } else {
        *dt_grid = dt;
        *limiting_cell = cell;
    }
}

// The scratch array's device address on the current device, looked up once
// per device.
int scratch_ptr(unsigned char** out) {
    int dev = 0;
    WR_HIP(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kNaMaxDevices)
        return detail::fail(WR_ERR_ARGUMENT, "wr_prepare_cells_noalloc: device index too large");
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (!g_na_ptr[dev]) {
        void* p = nullptr;
        WR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_na_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_na_ptr[dev];
    return WR_SUCCESS;
}

}  // namespace

int wr_prepare_cells_noalloc(wr_collapse_state* cells, int num_cells, int completed_global_steps,
                             double grid_time, int step, int perturb, double* dt_grid,
                             int* limiting_cell, int* failure, void* stream) {
    if (num_cells < 0 || num_cells > kNaMaxCells)
        return detail::fail(WR_ERR_ARGUMENT, "wr_prepare_cells_noalloc: num_cells must be 0.." +
                                                 std::to_string(kNaMaxCells));
    if ((num_cells > 0 && !cells) || !dt_grid || !limiting_cell || !failure)
        return detail::fail(WR_ERR_ARGUMENT, "wr_prepare_cells_noalloc: null pointer");
    unsigned char* scr = nullptr;
    const int st = scratch_ptr(&scr);
    if (st != WR_SUCCESS) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int nparts = num_cells > 0 ? (num_cells + kNaBlock - 1) / kNaBlock : 0;
    hipLaunchKernelGGL(k_na_begin, dim3(1), dim3(1), 0, s, scr);
    WR_HIP(hipGetLastError(), "k_na_begin launch");
    if (nparts > 0) {
        hipLaunchKernelGGL(k_na_prepare, dim3(nparts), dim3(kNaBlock), 0, s, cells, num_cells,
                           completed_global_steps, grid_time, step, perturb ? 1 : 0, scr);
        WR_HIP(hipGetLastError(), "prepare_grid_timestep_kernel launch");
    }
    hipLaunchKernelGGL(k_na_reduce, dim3(1), dim3(kNaBlock), 0, s, scr, nparts, dt_grid,
                       limiting_cell, failure);
    WR_HIP(hipGetLastError(), "grid timestep reduction launch");
    return WR_SUCCESS;
}
