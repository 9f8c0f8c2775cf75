// The queue-fed persistent advance kernel, shared by the library's translation
// units: the production instance (wr_kernels.hip, network = the generated
// staged J + f) and the adjoint's recording instance (wr_adjoint.hip, network
// = f_theta, with checkpoints and the accepted-step tape).  The Net policy
// supplies the two network evaluations and the tape hooks; the production
// policy's hooks are empty.
#pragma once

#include "wr_ctx.hpp"
#include "wr_device.hpp"

namespace wr {

// ---- queue-fed advance ---------------------------------------------------------------
// The same per-cell work as k_advance (reproducer.cpp:7736-7785 with
// RODAS::integrate 7205-7347), reorganised for the GPU:
//  * a fixed grid of persistent threads takes cells from an atomic queue;
//  * the integrator is a flat state machine -- one loop iteration is one
//    ROS2S attempt (or one singular-matrix retry) of the cell a lane holds --
//    so a lane whose cell finishes takes the next cell at the next iteration
//    instead of idling until its wave's slowest cell is done;
//  * J, LU and the stage vectors live in a per-thread SoA workspace; the
//    network code is register-budgeted (staging slots in the same workspace),
//    so nothing is left to compiler spilling.
// Counters, results and every arithmetic operation are those of the
// reference; only where values are kept differs.
constexpr int kQBlock = 256;

// The per-thread workspace is the lane's private memory: the hardware
// interleaves private (scratch) memory by lane, so element e of all 64 lanes
// of a wave is one contiguous 512-byte line, and an element is addressed by
// an immediate offset.  `zero` (always 0) is an index the compiler cannot
// resolve, which keeps the array in memory instead of registers.
// Net::kSlots staging slots follow J, LU and the stage vectors.

// The first kSoloLanes queue positions -- the cells that took the most
// attempts last time -- go to the first wave of block 0, which runs alone on
// its CU (the block's other waves leave at once): the one cell that needs
// tens of attempts then does not share instruction issue and fetch with
// three other waves while the bulk of the cells is being advanced.
constexpr int kSoloLanes = 64;  // default of Ctl::solo_lanes
// Most cooperative cells of the default setting (one per 2,048 cells).
constexpr int kCoopAutoMax = 128;

// kMixed: the FP32-factorisation form (E = fac I - J factorised in FP32, FP64
// residual and refinement sweeps per stage solve, FP64 factors when the FP32
// ones are singular or a refinement does not converge); the RHS, energy
// equation, error norm and controller stay FP64.
// Stage solve of the queue-fed kernel: FP64 factors, or (kMixed) FP32 factors
// with refinement and the FP64 factors from the first non-convergence on.
// One stage vector of the lane workspace into registers, its loads issued
// together (the scheduling barrier keeps them ahead of their uses).
__device__ __forceinline__ void load_vec(const ThreadWs& w, int off, double (&out)[kNeqs]) {
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) out[i] = WR_AT(w.vec, off + i, w.stride);
    __builtin_amdgcn_sched_barrier(0);
}

template <bool kMixed>
__device__ __forceinline__ void stage_solve(const ThreadWs& w, const float* lu32, uint64_t piv32,
                                            const int (&ks)[kNeqs], double fac, int sweeps,
                                            Ctl* ctl, bool& use64, uint64_t& piv,
                                            double (&v)[kNeqs]) {
    if constexpr (kMixed) {
        if (!use64) {
            double b[kNeqs];
#pragma unroll
            for (int i = 0; i < kNeqs; ++i) b[i] = v[i];
            if (solve_mixed(w, lu32, piv32, ks, fac, sweeps, v)) return;
            atomicAdd(&ctl->mixed_refine_fallbacks, 1ull);
            if (lu_left(w, fac, piv) != 0) return;  // keep the refined FP32 result
            use64 = true;
#pragma unroll
            for (int i = 0; i < kNeqs; ++i) v[i] = b[i];
        }
    }
    lu_solve_ws(w, piv, v);
}

// kLocal: the legacy-local policy (the reference CPU driver, reproducer.cpp:
// 7646-7685): each cell perturbs itself and steps with its own free-fall dt;
// a failure is recorded as (cell, code) by atomic minimum, so that the lowest
// failing cell wins as in the reference's sequential cell loop.
// kJacReuse: Jacobian reuse (no counterpart in the reference, which evaluates
// J at the head of every outer iteration, 7236-7260): J is evaluated at a
// burn's first attempt and again once Ctl::jac_age_max steps have been
// accepted with it; an attempt rejected with a J from an earlier state
// evaluates J at the current state before the retry.  Other attempts evaluate
// f alone.
template <bool kMixed, class Net, bool kLocal = false, bool kJacReuse = false>
__global__ __launch_bounds__(kQBlock) void k_advance_q(DevCells s, Ctl* ctl, int zero, Net net) {
    using C = Ros2sCoef;
    if (ctl->done) return;
    // wave-cooperative cells: the first coop_cells queue positions, taken by
    // wave 0 of the first coop_waves blocks, one cell per wave (wr_coop.hpp);
    // those waves then join the per-lane queue below
    constexpr bool kCoop = Net::kCoop && !kMixed && !kLocal && !kJacReuse;
    const int coop0 = kCoop ? ctl->coop_cells : 0;
    if constexpr (kCoop) {
        if (coop0 > 0 && threadIdx.x < 64 && static_cast<int>(blockIdx.x) < ctl->coop_waves)
            Net::coop(s, ctl, static_cast<int>(threadIdx.x));
    }
    const int solo_lanes = ctl->solo_lanes;
    bool solo = blockIdx.x == 0 && solo_lanes > 0;
    if (solo && static_cast<int>(threadIdx.x) >= solo_lanes) return;
    double priv[2 * kNeqs * kNeqs + kVecLen + Net::kSlots];
    double* const p = priv + zero;
    float priv32[kMixed ? kNeqs * kNeqs : 1];
    float* const lu32 = priv32 + zero;
    const int sweeps = ctl->lu_sweeps;
    ThreadWs w;
    w.J = p;
    w.LU = p + kNeqs * kNeqs;
    w.vec = p + 2 * kNeqs * kNeqs;
    w.stage = p + 2 * kNeqs * kNeqs + kVecLen;
    w.stride = 1;
    const int st = 1;
    const int completed = ctl->completed;
    const double dt_grid = ctl->dt_grid;
    const double next_time = ctl->next_grid_time;
    const int step = ctl->step;
    double tout = dt_grid;  // the burn's end time: dt_grid, or (kLocal) the cell's own dt
    const int max_steps = ctl->integ_max_steps;
    const int jac_age_max = kJacReuse ? ctl->jac_age_max : 1;
    int jac_age = -1;  // (kJacReuse) accepted steps since J was evaluated; -1: none yet
    int c = -1;
    double x = 0.0, h = 0.0, hopt = 0.0, hacc = 0.0, erracc = 1.0, rho = 0.0, e0 = 0.0;
    int nsing = 0, n_step = 0, n_accept = 0;
    bool reject = false, last = false, top = true;
    Counters cnt{0, 0, 0, 0, 0, 0, 0};
    uint64_t piv = 0, piv32 = 0;
    for (;;) {
        if (c < 0) {
            // ---- take the next cell: advance_collapse_gridwide_kernel's preamble
            int q;
            if (solo) {
                q = atomicAdd(&ctl->queue_solo, 1);
                if (q >= solo_lanes) {
                    solo = false;  // reserved positions used up: join the common queue
                    continue;
                }
                q += coop0;
            } else {
                q = coop0 + solo_lanes + atomicAdd(&ctl->queue, 1);
            }
            if (q >= s.n || ctl->failure != kSuccess) break;
            c = s.order[q];
            if constexpr (kLocal) {
                // advance_collapse_step (reproducer.cpp:7646-7665)
                if (s.completed[c] < completed) {  // stopped in an earlier step
                    c = -1;
                    continue;
                }
                double xn[kNumSpec];
#pragma unroll
                for (int n = 0; n < kNumSpec; ++n) xn[n] = s.xn[n * s.cap + c];
                double driver = s.driver[c];
                double rho0 = s.rho[c];
                double T0 = s.T[c];
                if (ctl->perturb && step != 0 && step % 20 == 0) {  // apply_perturbation (7564-7580)
                    const double f = perturbation_factor(c, step);
                    driver *= f;
                    rho0 *= f;
#pragma unroll
                    for (int n = 0; n < kNumSpec; ++n) xn[n] *= f;
                    T0 = post_step(xn, rho0, s.e[c]);
                    s.T[c] = T0;
                }
                const double old_density = driver;
                const double ldt = collapse_timestep(xn);
                const double tff = ldt / 1.0e-1;
                driver += ldt * (driver / tff);
                s.driver[c] = driver;
                if (ldt < 10.0 || driver > 2.0e-6) {
#pragma unroll
                    for (int n = 0; n < kNumSpec; ++n) s.xn[n * s.cap + c] = xn[n];
                    s.rho[c] = rho0;
                    c = -1;
                    continue;
                }
                const double ratio = driver / old_density;
#pragma unroll
                for (int n = 0; n < kNumSpec; ++n) {
                    xn[n] *= ratio;
                    s.xn[n * s.cap + c] = xn[n];
                }
                rho = rho0 * ratio;
                s.rho[c] = rho;
                e0 = eos_energy(xn, T0);
#pragma unroll
                for (int n = 0; n < kNumSpec; ++n) WR_AT(w.vec, kVecY + n, st) = xn[n];
                WR_AT(w.vec, kVecY + kNumSpec, st) = e0;
                tout = ldt;
            } else {
            if (s.completed[c] != completed) {
                net.skipped(c);
                c = -1;
                continue;
            }
            double xn[kNumSpec];
#pragma unroll
            for (int n = 0; n < kNumSpec; ++n) xn[n] = s.xn[n * s.cap + c];
            net.begin(c, xn, s.T[c], s.e[c], s.rho[c], s.driver[c]);
            const double local_dt = collapse_timestep(xn);
            if (local_dt < 10.0) {
                net.skipped(c);
                c = -1;
                continue;
            }
            const double old_density = s.driver[c];
            const double tff = local_dt / 1.0e-1;
            const double driver = old_density + dt_grid * (old_density / tff);
            s.driver[c] = driver;
            if (!valid_positive(driver) || driver > 2.0e-6) {
                net.skipped(c);
                c = -1;
                continue;
            }
            const double ratio = driver / old_density;
#pragma unroll
            for (int n = 0; n < kNumSpec; ++n) {
                xn[n] *= ratio;
                s.xn[n * s.cap + c] = xn[n];  // the kernel scales the state in place
            }
            rho = s.rho[c] * ratio;
            s.rho[c] = rho;
            // burn_ros2s (7592-7606): eos_rt, y = (xn, e)
            e0 = eos_energy(xn, s.T[c]);
#pragma unroll
            for (int n = 0; n < kNumSpec; ++n) WR_AT(w.vec, kVecY + n, st) = xn[n];
            WR_AT(w.vec, kVecY + kNumSpec, st) = e0;
            }
            // integrate() prologue, t = 0, tout = dt (> 0: checked by k_plan, or >= 10)
            h = tout;
            if (fabs(h) <= 10.0 * kUround) h = 1.0e-6;
            h = rmin(fabs(h), tout - 0.0);
            x = 0.0;
            hopt = h;
            hacc = 0.0;
            erracc = 1.0;
            nsing = 0;
            n_step = 0;
            n_accept = 0;
            reject = false;
            last = false;
            top = true;
            cnt = Counters{0, 0, 0, 0, 0, 0, 0};
            jac_age = -1;
        }
        int result = 0;  // 0: the cell goes on
        if (top) {
            // head of the outer loop of integrate (7236-7258)
            if (n_step > max_steps) {
                result = kTooManySteps;
            } else if (0.1 * fabs(h) <= fabs(x) * kUround) {
                result = kDtUnderflow;
            } else if (last) {
                result = kSuccess;
            } else {
                hopt = h;
                if (x + h * 1.0001 >= tout) {
                    h = tout - x;
                    last = true;
                }
                if (!kJacReuse || jac_age < 0 || jac_age >= jac_age_max) {
                    net.jac_rhs(w);
                    cnt.n_jac += 1;
                    jac_age = 0;
                } else {  // f(y) alone, through the stage evaluation's slots
#pragma unroll
                    for (int i = 0; i < kNeqs; ++i)
                        WR_AT(w.vec, kVecYnew + i, st) = WR_AT(w.vec, kVecY + i, st);
                    net.rhs(w);
#pragma unroll
                    for (int i = 0; i < kNeqs; ++i)
                        WR_AT(w.vec, kVecF0 + i, st) = WR_AT(w.vec, kVecDy + i, st);
                }
                top = false;
            }
        }
        if (result == 0) {
            // retry loop body (7260-7345): factorize, three stages, controller
            const double fac = 1.0 / (h * C::gamma);
            bool use64 = !kMixed;
            int info;
            int ks[kNeqs];
            if constexpr (kMixed) {
                mixed_scales(w, ks);
                info = lu_left_t<float>(w.J, lu32, st, fac, piv32, ks);
                if (info != 0) {  // singular in FP32: FP64 factors for this attempt
                    atomicAdd(&ctl->mixed_lu_fallbacks, 1ull);
                    info = lu_left(w, fac, piv);
                    use64 = true;
                }
            } else {
                info = lu_left(w, fac, piv);
            }
            if (info != 0) {
                nsing += 1;
                if (nsing >= 5) {
                    result = kLuError;
                } else {
                    h *= 0.5;
                    reject = true;
                    last = false;
                    if (kJacReuse && jac_age > 0) {  // J from an earlier state
                        net.jac_rhs(w);
                        cnt.n_jac += 1;
                        jac_age = 0;
                    }
                }
            } else {
                cnt.n_decomp += 1;
                // Stage vectors are read from the workspace in whole vectors
                // (load_vec) before they are used, so their loads overlap.
                double v[kNeqs], a[kNeqs], b[kNeqs];
                load_vec(w, kVecF0, v);
                stage_solve<kMixed>(w, lu32, piv32, ks, fac, sweeps, ctl, use64, piv, v);  // ak1
                cnt.n_solve += 1;
                load_vec(w, kVecY, a);
#pragma unroll
                for (int i = 0; i < kNeqs; ++i) {
                    WR_AT(w.vec, kVecAk1 + i, st) = v[i];
                    WR_AT(w.vec, kVecYnew + i, st) = a[i] + C::a21 * v[i];
                    WR_AT(w.vec, kVecAk2 + i, st) = (C::c21 / h) * v[i];
                }
                net.rhs(w);
                load_vec(w, kVecAk2, a);
                load_vec(w, kVecDy, b);
#pragma unroll
                for (int i = 0; i < kNeqs; ++i) v[i] = a[i] + b[i];
                stage_solve<kMixed>(w, lu32, piv32, ks, fac, sweeps, ctl, use64, piv, v);  // ak2
                cnt.n_solve += 1;
                load_vec(w, kVecAk1, a);
                load_vec(w, kVecY, b);
#pragma unroll
                for (int i = 0; i < kNeqs; ++i) {
                    const double a1 = a[i];
                    WR_AT(w.vec, kVecAk2 + i, st) = v[i];
                    WR_AT(w.vec, kVecYnew + i, st) = b[i] + C::a31 * a1 + C::a32 * v[i];
                    WR_AT(w.vec, kVecAk3 + i, st) = (C::c31 * a1 + C::c32 * v[i]) / h;
                }
                net.rhs(w);
                load_vec(w, kVecAk3, a);
                load_vec(w, kVecDy, b);
#pragma unroll
                for (int i = 0; i < kNeqs; ++i) v[i] = a[i] + b[i];
                stage_solve<kMixed>(w, lu32, piv32, ks, fac, sweeps, ctl, use64, piv, v);  // ak3
                cnt.n_solve += 1;
                double err = 0.0;
                load_vec(w, kVecAk1, a);
                load_vec(w, kVecAk2, b);
                double yv[kNeqs];
                load_vec(w, kVecY, yv);
#pragma unroll
                for (int i = 0; i < kNeqs; ++i) {
                    const double a1 = a[i];
                    const double a2 = b[i];
                    const double yi = yv[i];
                    const double yn = yi + C::b1 * a1 + C::b2 * a2 + C::b3 * v[i];
                    WR_AT(w.vec, kVecYnew + i, st) = yn;
                    const double wv = C::e1 * a1 + C::e2 * a2 + C::e3 * v[i];
                    const double sk = atol_of(i) + rtol_of(i) * rmax(fabs(yi), fabs(yn));
                    const double term = wv / sk;
                    err += term * term;
                }
                cnt.n_rhs += 3;
                n_step += 1;
                cnt.n_step = n_step;
                err = sqrt(err / static_cast<double>(kNeqs));
                const double raw_fac = WR_CBRT(err) / kSafe;
                const double fac_step = rmax(1.0 / kFacMax, rmin(1.0 / kFacMin, raw_fac));
                double hnew = h / fac_step;
                if (err <= 1.0) {
                    n_accept += 1;
                    cnt.n_accept = n_accept;
                    if (n_accept > 1) {
                        const double facgus = rmax(
                            1.0 / kFacMax,
                            rmin(1.0 / kFacMin, (hacc / h) * WR_CBRT((err * err) / erracc) / kSafe));
                        hnew = h / rmax(fac_step, facgus);
                    }
                    hacc = h;
                    erracc = rmax(1.0e-2, err);
                    load_vec(w, kVecYnew, a);
#pragma unroll
                    for (int i = 0; i < kNeqs; ++i) WR_AT(w.vec, kVecY + i, st) = a[i];
                    net.accepted(c, h);
                    jac_age += 1;
                    x += h;
                    if (fabs(hnew) > tout) hnew = tout;
                    if (reject) hnew = rmin(fabs(hnew), fabs(h));
                    reject = false;
                    h = hnew;
                    top = true;
                } else {
                    reject = true;
                    last = false;
                    h = hnew;
                    if (n_accept >= 1) cnt.n_reject += 1;
                    if (kJacReuse && jac_age > 0) {  // J from an earlier state
                        net.jac_rhs(w);
                        cnt.n_jac += 1;
                        jac_age = 0;
                    }
                }
            }
        }
        if (result != 0) {
            // ---- the cell is done: burn_ros2s epilogue and the kernel's commit
            const int cv[7] = {cnt.n_step, cnt.n_rhs, cnt.n_jac, cnt.n_decomp, cnt.n_solve,
                               cnt.n_accept, cnt.n_reject};
#pragma unroll
            for (int k = 0; k < 7; ++k)
                s.stats[k * s.cap + c] += static_cast<unsigned long long>(cv[k] > 0 ? cv[k] : 0);
            s.cost[c] = cnt.n_step;
            net.end(c);
            if (result == kSuccess) {
                double xn[kNumSpec];
#pragma unroll
                for (int n = 0; n < kNumSpec; ++n) xn[n] = WR_AT(w.vec, kVecY + n, st);
                const double e = WR_AT(w.vec, kVecY + kNumSpec, st);
                const double T = post_step(xn, rho, e);
#pragma unroll
                for (int n = 0; n < kNumSpec; ++n) s.xn[n * s.cap + c] = xn[n];
                s.e[c] = e;
                s.T[c] = T;
                s.time[c] = kLocal ? s.time[c] + tout : next_time;
                s.completed[c] = completed + 1;
                atomicAdd(&ctl->integrated, 1);
            } else {
                s.e[c] = e0;  // eos_rt ran before the failed burn
                if constexpr (kLocal)
                    atomicMin(&ctl->local_fail_key,
                              (static_cast<unsigned long long>(c) << 32) |
                                  static_cast<unsigned int>(result));
                else
                    set_failure(ctl, result, c);
            }
            c = -1;
        }
    }
}

}  // namespace wr
