// gridwide-cta32 policy of the reference's Mojo front end
// (reproducer.mojo:2874-3048): setup_gridwide_cells and integrate_gridwide_tile.
//
// The Mojo policy advances tiles of 32 consecutive cells with one ROS2S step
// controller per tile.  Here a tile is one half of a 64-lane wave, one lane
// per cell.  Every lane runs the trial of its own cell; the tile's decision
// (the lowest singular cell, the largest error, whether any error is not
// finite) comes from ballots and a 32-lane ds_swizzle butterfly, so every lane
// of a tile holds the same controller state, bit for bit.  The trial is the
// queue-fed kernel's attempt: register-budgeted network code, lane-private
// workspace (J, LU, stage vectors, staging slots), left-looking LU with the
// factor columns loaded in batches.  Waves are persistent and take the next
// pair of tiles (64 cells) from an atomic counter, so a wave whose tiles
// finish early does not idle until the step ends.
#pragma once

#include <float.h>

#include "wr_ctx.hpp"
#include "wr_device.hpp"
#include "wr_queue.hpp"

namespace wr {

constexpr int kTileCells = 32;  // TileCells, reproducer.mojo:40
constexpr int kCtaBlock = 64;   // one wave: two tiles

// v of lane (lane ^ M) within the same 32-lane half: ds_swizzle in its bit
// mode (and_mask 0x1F, or_mask 0, xor_mask M), one exchange per 32-bit half.
template <int M>
__device__ __forceinline__ double swizzle_xor(double v) {
    constexpr int kPattern = 0x1F | (M << 10);
    const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(v), kPattern);
    const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(v), kPattern);
    return __hiloint2double(hi, lo);
}

// Largest value over a tile's 32 lanes (no NaN operands: the maximum is exact
// and every lane gets the same bits).
__device__ __forceinline__ double tile_max(double v) {
    v = rmax(v, swizzle_xor<16>(v));
    v = rmax(v, swizzle_xor<8>(v));
    v = rmax(v, swizzle_xor<4>(v));
    v = rmax(v, swizzle_xor<2>(v));
    v = rmax(v, swizzle_xor<1>(v));
    return v;
}

// One ROS2S trial of step h (ros2s_trial, reproducer.mojo:2460-2560) from y =
// the workspace's kVecY: returns the factorisation's info; when 0, the
// candidate is in kVecYnew and err holds the error norm.  J and f(y) are
// evaluated unless `fresh` says they are those of this y already: a rejected
// or singular trial leaves y unchanged, so its retry reuses them (the Mojo
// code evaluates them again and gets the same bits; the Jacobian counter
// counts every trial either way).  Same operations in the same order as the
// queue-fed kernel's attempt.
__device__ __forceinline__ int cta_trial(const ThreadWs& w, double h, bool& fresh, double& err) {
    using C = Ros2sCoef;
    const int st = w.stride;
    if (!fresh) {
        ws_jac_rhs(w);
        fresh = true;
    }
    const double fac = 1.0 / (h * C::gamma);
    uint64_t piv = 0;
    const int info = lu_left(w, fac, piv);
    if (info != 0) return info;
    double v[kNeqs], a[kNeqs], b[kNeqs];
    load_vec(w, kVecF0, v);
    lu_solve_ws(w, piv, v);  // ak1
    load_vec(w, kVecY, a);
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) {
        WR_AT(w.vec, kVecAk1 + i, st) = v[i];
        WR_AT(w.vec, kVecYnew + i, st) = a[i] + C::a21 * v[i];
        WR_AT(w.vec, kVecAk2 + i, st) = (C::c21 / h) * v[i];
    }
    ws_rhs(w);
    load_vec(w, kVecAk2, a);
    load_vec(w, kVecDy, b);
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) v[i] = a[i] + b[i];
    lu_solve_ws(w, piv, v);  // ak2
    load_vec(w, kVecAk1, a);
    load_vec(w, kVecY, b);
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) {
        const double a1 = a[i];
        WR_AT(w.vec, kVecAk2 + i, st) = v[i];
        WR_AT(w.vec, kVecYnew + i, st) = b[i] + C::a31 * a1 + C::a32 * v[i];
        WR_AT(w.vec, kVecAk3 + i, st) = (C::c31 * a1 + C::c32 * v[i]) / h;
    }
    ws_rhs(w);
    load_vec(w, kVecAk3, a);
    load_vec(w, kVecDy, b);
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) v[i] = a[i] + b[i];
    lu_solve_ws(w, piv, v);  // ak3
    double e = 0.0;
    load_vec(w, kVecAk1, a);
    load_vec(w, kVecAk2, b);
    double yv[kNeqs];
    load_vec(w, kVecY, yv);
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) {
        const double yi = yv[i];
        const double yn = yi + C::b1 * a[i] + C::b2 * b[i] + C::b3 * v[i];
        WR_AT(w.vec, kVecYnew + i, st) = yn;
        const double wv = C::e1 * a[i] + C::e2 * b[i] + C::e3 * v[i];
        // error_norm (reproducer.mojo:2448-2457)
        const double sk = atol_of(i) + rtol_of(i) * rmax(fabs(yi), fabs(yn));
        const double term = wv / sk;
        e += term * term;
    }
    err = sqrt(e / static_cast<double>(kNeqs));
    return 0;
}

// One step of the policy over the cells [64 k, 64 k + 64) for every k a wave
// takes from ctl->queue (reset by the plan kernel each step).  `zero` is
// always 0: it keeps the workspace array in private memory.
__global__ __launch_bounds__(kCtaBlock) void k_advance_cta(DevCells s, Ctl* ctl, int zero) {
    if (ctl->done || ctl->failure != kSuccess) return;
    double priv[2 * kNeqs * kNeqs + kVecLen + kStageSlots];
    double* const p = priv + zero;
    ThreadWs w;
    w.J = p;
    w.LU = p + kNeqs * kNeqs;
    w.vec = p + 2 * kNeqs * kNeqs;
    w.stage = p + 2 * kNeqs * kNeqs + kVecLen;
    w.stride = 1;
    const int st = 1;
    const int lane = threadIdx.x;
    const int half = lane >> 5, tl = lane & (kTileCells - 1);
    const unsigned long long tmask = 0xffffffffull << (32 * half);
    const int completed = ctl->completed;
    const double dt_grid = ctl->dt_grid;
    const double next_time = ctl->next_grid_time;
    const int max_steps = ctl->integ_max_steps;
    const int nchunks = (s.n + kCtaBlock - 1) / kCtaBlock;
    for (;;) {
        int chunk = 0;
        if (lane == 0) chunk = ctl->failure != kSuccess ? nchunks : atomicAdd(&ctl->queue, 1);
        chunk = __shfl(chunk, 0, kCtaBlock);
        if (chunk >= nchunks) break;
        const int c = chunk * kCtaBlock + lane;
        // setup_gridwide_cells: the density update as (dt_grid * driver) / tff, eos_rt
        bool act = false;
        double rho = 0.0;
        if (c < s.n && s.completed[c] == completed) {
            double xn[kNumSpec];
#pragma unroll
            for (int n = 0; n < kNumSpec; ++n) xn[n] = s.xn[n * s.cap + c];
            const double local_dt = collapse_timestep(xn);
            if (local_dt >= 10.0) {
                const double old_density = s.driver[c];
                const double tff = local_dt / 1.0e-1;
                const double driver = old_density + dt_grid * old_density / tff;
                s.driver[c] = driver;
                if (valid_positive(driver) && driver <= 2.0e-6) {
                    const double ratio = driver / old_density;
#pragma unroll
                    for (int n = 0; n < kNumSpec; ++n) xn[n] *= ratio;
                    rho = s.rho[c] * ratio;
                    s.rho[c] = rho;
#pragma unroll
                    for (int n = 0; n < kNumSpec; ++n) WR_AT(w.vec, kVecY + n, st) = xn[n];
                    WR_AT(w.vec, kVecY + kNumSpec, st) = eos_energy(xn, s.T[c]);
                    act = true;
                }
            }
        }
        const unsigned long long amask = __ballot(act) & tmask;
        if (amask == 0) continue;  // no participating cell in this tile

        // integrate_gridwide_tile: one controller per tile, held by every lane
        double h = dt_grid, x = 0.0, hacc = 0.0, erracc = 1.0;
        bool reject = false, fresh = false;
        int nsing = 0, n_step = 0, n_accept = 0, status = kSuccess;
        // this step's counter increments: internal, rhs, jacobian,
        // decompositions, solves, accepted, rejected
        int inc[7] = {0, 0, 0, 0, 0, 0, 0};
        while (x < dt_grid) {
            if (n_step > max_steps) {
                status = kTooManySteps;
                break;
            }
            if (0.1 * fabs(h) <= fabs(x) * kUround) {
                status = kDtUnderflow;
                break;
            }
            bool final_trial = false;
            if (x + h * 1.0001 >= dt_grid) {
                h = dt_grid - x;
                final_trial = true;
            }
            int info = 0;
            double err = 0.0;
            if (act) info = cta_trial(w, h, fresh, err);
            // the Mojo loop visits the tile's cells in order and stops at the
            // first singular factorisation: the cells after it run no trial
            const unsigned long long smask = __ballot(act && info != 0) & tmask;
            const int first_sing =
                smask ? (__ffsll(static_cast<long long>(smask)) - 1 - 32 * half) : kTileCells;
            if (act && tl <= first_sing) {
                inc[2] += 1;
                if (tl < first_sing) {
                    inc[0] += 1;
                    inc[1] += 3;
                    inc[3] += 1;
                    inc[4] += 3;
                }
            }
            if (smask) {
                nsing += 1;
                if (nsing >= 5) {
                    status = kLuError;
                    break;
                }
                h *= 0.5;
                reject = true;
                continue;
            }
            const bool nonfinite = act && !isfinite(err);
            const bool any_nonfinite = (__ballot(nonfinite) & tmask) != 0;
            const double emax = tile_max((act && !nonfinite) ? err : 0.0);
            const double err_tile = any_nonfinite ? DBL_MAX : emax;
            n_step += 1;
            const double raw_fac = WR_CBRT(err_tile) / kSafe;
            const double fac_step = rmax(1.0 / kFacMax, rmin(1.0 / kFacMin, raw_fac));
            double hnew = h / fac_step;
            if (err_tile <= 1.0) {
                n_accept += 1;
                if (n_accept > 1) {
                    const double facgus = rmax(
                        1.0 / kFacMax,
                        rmin(1.0 / kFacMin,
                             (hacc / h) * WR_CBRT((err_tile * err_tile) / erracc) / kSafe));
                    hnew = h / rmax(fac_step, facgus);
                }
                hacc = h;
                erracc = rmax(1.0e-2, err_tile);
                if (act) {
                    double yn[kNeqs];
                    load_vec(w, kVecYnew, yn);
#pragma unroll
                    for (int i = 0; i < kNeqs; ++i) WR_AT(w.vec, kVecY + i, st) = yn[i];
                    fresh = false;
                    inc[5] += 1;
                }
                x = final_trial ? dt_grid : x + h;
                if (fabs(hnew) > dt_grid) hnew = dt_grid;
                if (reject) hnew = rmin(fabs(hnew), fabs(h));
                reject = false;
                h = hnew;
                continue;
            }
            reject = true;
            h = hnew;
            if (n_accept >= 1 && act) inc[6] += 1;
        }
        double y[kNeqs];
        if (act) {
            load_vec(w, kVecY, y);
#pragma unroll
            for (int k = 0; k < 7; ++k)
                s.stats[k * s.cap + c] += static_cast<unsigned long long>(inc[k]);
            s.cost[c] = inc[0];
        }
        if (status != kSuccess) {
            // the Mojo driver raises on a failed step: the cells keep the
            // last accepted trial of their tile (no post-step); the lowest
            // failing tile is reported
            if (act) {
#pragma unroll
                for (int n = 0; n < kNumSpec; ++n) s.xn[n * s.cap + c] = y[n];
                s.e[c] = y[kNumSpec];
            }
            if (tl == 0) {
                atomicMin(&ctl->local_fail_key, (static_cast<unsigned long long>(c) << 32) |
                                                    static_cast<unsigned int>(status));
                atomicCAS(&ctl->failure, kSuccess, status);
            }
            continue;
        }
        if (act) {
            double xn[kNumSpec];
#pragma unroll
            for (int n = 0; n < kNumSpec; ++n) xn[n] = y[n];
            const double e = y[kNumSpec];
            const double T = post_step(xn, rho, e);
#pragma unroll
            for (int n = 0; n < kNumSpec; ++n) s.xn[n * s.cap + c] = xn[n];
            s.e[c] = e;
            s.T[c] = T;
            s.time[c] = next_time;
            s.completed[c] = completed + 1;
        }
        if (tl == 0) atomicAdd(&ctl->integrated, __popcll(amask));
    }
}

}  // namespace wr
