// Wave-cooperative ROS2S advance of one cell: all 64 lanes of a wavefront
// work on the same cell.  The queue-fed kernel (wr_queue.hpp) hands the
// cells that took the most attempts in the last outer step -- the one or few
// cells whose serial chain of attempts otherwise bounds the step -- to such
// waves, one cell per wave, and the rest to single lanes.
//
// Per attempt (the operations of RODAS<PrimordialChem>::integrate,
// reproducer.cpp:7205-7347, in the same order, so every value is
// bit-identical to the single-lane path):
//  * the network (fused J + f, and f at the two stage states) runs as the
//    generated cooperative programs of gen/coop.py over the wave's LDS
//    region: vector steps with per-lane operand addresses, uniform register
//    code for the thin chains;
//  * LU of E = fac I - J: lane i holds one physical row in registers and its
//    current logical row index; a row exchange swaps two labels, the pivot
//    search scans the column in LDS in logical order (first strict maximum,
//    LINPACK, 96-151), the pivot row is broadcast with v_readlane, and each
//    element receives the updates a_ij += a_kj * m_ik in the reference's k
//    order.  The factors go to LDS in LINPACK storage (multipliers negated,
//    stored at their row's position when their column is eliminated);
//  * stage solves: lane i holds x_i and row i of the factors; interchanges and
//    the column updates of lu_solve (154-181) with v_readlane broadcasts;
//  * stage vectors are elementwise (lane i owns element i); the error norm
//    sums its 15 terms in index order on every lane; the step-size controller
//    is scalar code every lane runs identically.
// No block barrier is used: a cooperative wave owns its LDS region, and LDS
// operations of one wave complete in issue order.
#pragma once

#include "wr_ctx.hpp"
#include "wr_device.hpp"

namespace wr {


__device__ __forceinline__ double coop_flip(double x, uint32_t w) {
    return __longlong_as_double(__double_as_longlong(x) ^
                                (static_cast<long long>(w & 0x80000000u) << 32));
}
__device__ __forceinline__ double coop_sgn(double x, uint32_t w) {
    long long b = __double_as_longlong(x);
    if (w & 0x40000000u) b &= 0x7fffffffffffffffll;
    return __longlong_as_double(b ^ (static_cast<long long>(w & 0x80000000u) << 32));
}

// Every access to the cooperative LDS region goes through may_alias types:
// the region is a byte array holding doubles, 32-bit booleans and 64-bit
// constant patterns, and the order of a load and a later store to the same
// bytes by other lanes is the programs' only synchronisation.  Plain typed
// accesses would let type-based alias analysis reorder a 32-bit access and a
// 64-bit access to the same bytes.
typedef double __attribute__((may_alias)) coop_f64;
typedef uint32_t __attribute__((may_alias)) coop_u32;
typedef uint64_t __attribute__((may_alias)) coop_u64;

#define COOP_FN __device__ __forceinline__
#define COOP_PARAMS const int lane
#define COOP_DATA __device__ const
#define COOP_DU(x) uint32_t x
#define COOP_DD(x) double x
#define COOP_V(x) x
#define COOP_L(...) \
    { __VA_ARGS__ }
#define COOP_TW(tab, row) (tab)[(row) * 64 + lane]
#define COOP_LD(a) (*reinterpret_cast<const coop_f64*>(wr_coop_lds + (a)))
#define COOP_LDU(a) (*reinterpret_cast<const coop_u32*>(wr_coop_lds + (a)))
#define COOP_ST(a, v) (*reinterpret_cast<coop_f64*>(wr_coop_lds + (a)) = (v))
#define COOP_STU(a, v) (*reinterpret_cast<coop_u32*>(wr_coop_lds + (a)) = (v))
#define COOP_LAYOUT_ONLY
#include "_gen/coop_network.h"
#undef COOP_LAYOUT_ONLY

// The cooperative wave's LDS region (the block's wave 0; the kernel instance
// that can run cooperative waves is the only one that references it, so no
// other kernel allocates it).  Byte layout from the generator: constants,
// trash slot, inputs T and X, value slots, J, stage vectors, factors,
// pivot-search column.
__shared__ __attribute__((aligned(16))) char wr_coop_lds[COOP_LDS_BYTES];

#include "_gen/coop_network.h"

// LDS accessors of the hand-written part (byte offsets from the generator).
__device__ __forceinline__ coop_f64& coop_at(int off) {
    return *reinterpret_cast<coop_f64*>(wr_coop_lds + off);
}
constexpr int kCoopVecY = 0, kCoopVecF0 = 15, kCoopVecYnew = 75, kCoopVecDy = 90;
__device__ __forceinline__ coop_f64& coop_vec(int v, int i) { return coop_at(COOP_OFF_VEC + 8 * (v + i)); }

// A double of lane `src` (wave-uniform) on every lane.
__device__ __forceinline__ double coop_readlane(double v, int src) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffll), src);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), src);
    return __longlong_as_double((static_cast<long long>(hi) << 32) |
                                static_cast<long long>(static_cast<unsigned int>(lo)));
}

__device__ __forceinline__ int coop_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// burn_state_from_y (reproducer.cpp:7029-7038) of the vector `v` in LDS into
// the programs' input slots: X = max(y, 1e-100), T = e / (...).  Every lane
// computes the same values; lanes 0-14 store them.
__device__ __forceinline__ void coop_state_in(int v, int lane) {
    double y[kNeqs], X[kNumSpec];
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) y[i] = coop_vec(v, i);
    const double T = state_from_y(y, X);
    double mine = T;
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) mine = (lane == n + 1) ? X[n] : mine;
    if (lane < kNeqs) coop_at(COOP_OFF_IN + 8 * lane) = mine;
}

// LU of E = fac I - J (J in LDS).  Returns info (0: nonsingular); pivots 4 bits
// per column in *piv; factors in LDS (LINPACK storage).
__device__ __forceinline__ int coop_lu(double fac, int lane, uint64_t& piv) {
    const bool real = lane < kNeqs;
    const int r = real ? lane : 0;
    double a[kNeqs];
#pragma unroll
    for (int j = 0; j < kNeqs; ++j) {
        const double v = -coop_at(COOP_OFF_J + 8 * (r * kNeqs + j));
        a[j] = (r == j) ? v + fac : v;
    }
    int lam = real ? lane : 64;  // logical row of this lane's physical row
    int info = 0;
    piv = 0;
#pragma unroll
    for (int k = 0; k < kNeqs - 1; ++k) {
        // pivot search over logical rows k..14 in order (first strict maximum)
        if (real && lam >= k) coop_at(COOP_OFF_COL + 8 * lam) = a[k];
        double col[kNeqs];
#pragma unroll
        for (int i = k; i < kNeqs; ++i) col[i] = coop_at(COOP_OFF_COL + 8 * i);
        int p = k;
        double best = fabs(col[k]);
#pragma unroll
        for (int i = k + 1; i < kNeqs; ++i) {
            const double v = fabs(col[i]);
            if (v > best) {
                best = v;
                p = i;
            }
        }
        double pv = col[k];
#pragma unroll
        for (int i = k + 1; i < kNeqs; ++i) pv = (p == i) ? col[i] : pv;
        p = coop_uniform(p);
        piv |= static_cast<uint64_t>(p) << (4 * k);
        if (pv == 0.0) {
            // singular column: no exchange, scaling or update (the column's
            // current values are its factors)
            info = k + 1;
            if (real && lam > k) coop_at(COOP_OFF_LU + 8 * (lam * kNeqs + k)) = a[k];
            continue;
        }
        if (p != k) lam = (lam == k) ? p : ((lam == p) ? k : lam);
        const double m = -1.0 / pv;
        const int P = coop_uniform(__ffsll(static_cast<unsigned long long>(__ballot(real && lam == k))) - 1);
        const bool below = real && lam > k;
        if (below) {
            a[k] *= m;
            coop_at(COOP_OFF_LU + 8 * (lam * kNeqs + k)) = a[k];
        }
#pragma unroll
        for (int j = k + 1; j < kNeqs; ++j) {
            const double t = coop_readlane(a[j], P);
            if (below) a[j] += t * a[k];
        }
    }
    piv |= static_cast<uint64_t>(kNeqs - 1) << (4 * (kNeqs - 1));
    // the upper part of each logical row: its physical row from its own column on
#pragma unroll
    for (int j = 0; j < kNeqs; ++j)
        if (real && j >= lam) coop_at(COOP_OFF_LU + 8 * (lam * kNeqs + j)) = a[j];
    if (coop_at(COOP_OFF_LU + 8 * (kNeqs * kNeqs - 1)) == 0.0) info = kNeqs;
    return info;
}

// Row `lane` of the factors into registers (lanes >= 15: row 0, unused).
__device__ __forceinline__ void coop_lu_row(int lane, double (&R)[kNeqs]) {
    const int r = lane < kNeqs ? lane : 0;
#pragma unroll
    for (int j = 0; j < kNeqs; ++j) R[j] = coop_at(COOP_OFF_LU + 8 * (r * kNeqs + j));
}

// lu_solve (reproducer.cpp:154-181): x is element `lane` of the vector.
__device__ __forceinline__ double coop_solve(double x, const double (&R)[kNeqs], uint64_t piv,
                                             int lane) {
#pragma unroll
    for (int k = 0; k < kNeqs - 1; ++k) {
        const int p = piv_get(piv, k);
        const double xk = coop_readlane(x, k);
        const double xp = coop_readlane(x, p);
        if (lane == p) x = xk;
        if (lane == k) x = xp;
        if (lane > k && lane < kNeqs) x += xp * R[k];
    }
#pragma unroll
    for (int kb = 0; kb < kNeqs; ++kb) {
        const int k = kNeqs - 1 - kb;
        const double q = coop_readlane(x, k) / coop_readlane(R[k], k);
        if (lane == k) x = q;
        const double t = -q;
        if (lane < k) x += t * R[k];
    }
    return x;
}

// The sum over i of v_i^2 in index order (v_i on lane i), on every lane.
__device__ __forceinline__ double coop_sumsq(double v, int lane) {
    if (lane < kNeqs) coop_at(COOP_OFF_COL + 8 * lane) = v;
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) {
        const double t = coop_at(COOP_OFF_COL + 8 * i);
        s += t * t;
    }
    return s;
}

// One cell's advance (the body of advance_collapse_gridwide_kernel,
// reproducer.cpp:7745-7784, with the integrator of 7205-7347), all lanes of
// the wave.  Returns after committing the cell (or leaving it untouched when
// it does not integrate this step).
__device__ __noinline__ void coop_cell(DevCells s, Ctl* ctl, int c, int lane) {
    using C = Ros2sCoef;
    const int completed = ctl->completed;
    const double dt_grid = ctl->dt_grid;
    const double next_time = ctl->next_grid_time;
    const int max_steps = ctl->integ_max_steps;
    const bool lead = lane == 0;
    if (s.completed[c] != completed) return;
    double xn[kNumSpec];
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) xn[n] = s.xn[n * s.cap + c];
    const double local_dt = collapse_timestep(xn);
    if (local_dt < 10.0) return;
    const double old_density = s.driver[c];
    const double tff = local_dt / 1.0e-1;
    const double driver = old_density + dt_grid * (old_density / tff);
    if (lead) s.driver[c] = driver;
    if (!valid_positive(driver) || driver > 2.0e-6) return;
    const double ratio = driver / old_density;
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) xn[n] *= ratio;
    if (lane < kNumSpec) {
        double mine = xn[0];
#pragma unroll
        for (int n = 1; n < kNumSpec; ++n) mine = (lane == n) ? xn[n] : mine;
        s.xn[lane * s.cap + c] = mine;
    }
    const double rho = s.rho[c] * ratio;
    const double e0 = eos_energy(xn, s.T[c]);
    if (lead) s.rho[c] = rho;
    // y = (xn, e): lane i owns y_i
    double y = e0;
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) y = (lane == n) ? xn[n] : y;
    if (lane < kNeqs) coop_vec(kCoopVecY, lane) = y;

    const double tout = dt_grid;
    double h = tout;
    if (fabs(h) <= 10.0 * kUround) h = 1.0e-6;
    h = rmin(fabs(h), tout - 0.0);
    double x = 0.0, hopt = h, hacc = 0.0, erracc = 1.0;
    int nsing = 0, n_step = 0, n_accept = 0;
    bool reject = false, last = false;
    Counters cnt{0, 0, 0, 0, 0, 0, 0};
    const int st = lane < kNeqs ? lane : 0;
    const double atol = atol_of(st), rtol = rtol_of(st);
    int result = 0;
    for (;;) {
        // head of the outer loop of integrate (7236-7258)
        if (n_step > max_steps) {
            result = kTooManySteps;
            break;
        }
        if (0.1 * fabs(h) <= fabs(x) * kUround) {
            result = kDtUnderflow;
            break;
        }
        if (last) {
            result = kSuccess;
            break;
        }
        hopt = h;
        if (x + h * 1.0001 >= tout) {
            h = tout - x;
            last = true;
        }
        coop_state_in(kCoopVecY, lane);
        coop_fused(lane);
        cnt.n_jac += 1;
        const double f0 = coop_vec(kCoopVecF0, st);
        for (;;) {
            // retry loop body (7260-7345): factorize, three stages, controller
            const double fac = 1.0 / (h * C::gamma);
            uint64_t piv;
            const int info = coop_lu(fac, lane, piv);
            if (info != 0) {
                nsing += 1;
                if (nsing >= 5) {
                    result = kLuError;
                    break;
                }
                h *= 0.5;
                reject = true;
                last = false;
                continue;
            }
            cnt.n_decomp += 1;
            double R[kNeqs];
            coop_lu_row(lane, R);
            const double ak1 = coop_solve(f0, R, piv, lane);
            cnt.n_solve += 1;
            if (lane < kNeqs) coop_vec(kCoopVecYnew, lane) = y + C::a21 * ak1;
            const double ak2p = (C::c21 / h) * ak1;
            coop_state_in(kCoopVecYnew, lane);
            coop_rhs(lane);
            const double ak2 = coop_solve(ak2p + coop_vec(kCoopVecDy, st), R, piv, lane);
            cnt.n_solve += 1;
            if (lane < kNeqs) coop_vec(kCoopVecYnew, lane) = y + C::a31 * ak1 + C::a32 * ak2;
            const double ak3p = (C::c31 * ak1 + C::c32 * ak2) / h;
            coop_state_in(kCoopVecYnew, lane);
            coop_rhs(lane);
            const double ak3 = coop_solve(ak3p + coop_vec(kCoopVecDy, st), R, piv, lane);
            cnt.n_solve += 1;
            const double yn = y + C::b1 * ak1 + C::b2 * ak2 + C::b3 * ak3;
            const double wv = C::e1 * ak1 + C::e2 * ak2 + C::e3 * ak3;
            const double sk = atol + rtol * rmax(fabs(y), fabs(yn));
            const double term = wv / sk;
            double err = coop_sumsq(term, lane);
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
                y = yn;
                if (lane < kNeqs) coop_vec(kCoopVecY, lane) = y;
                x += h;
                if (fabs(hnew) > tout) hnew = tout;
                if (reject) hnew = rmin(fabs(hnew), fabs(h));
                reject = false;
                h = hnew;
                break;
            }
            reject = true;
            last = false;
            h = hnew;
            if (n_accept >= 1) cnt.n_reject += 1;
        }
        if (result != 0) break;
    }
    // burn_ros2s epilogue and the kernel's commit (7607-7627, 7760-7784)
    const int cv[7] = {cnt.n_step, cnt.n_rhs, cnt.n_jac, cnt.n_decomp, cnt.n_solve, cnt.n_accept,
                       cnt.n_reject};
    if (lane < 7) {
        int mine = cv[0];
#pragma unroll
        for (int k = 1; k < 7; ++k) mine = (lane == k) ? cv[k] : mine;
        s.stats[lane * s.cap + c] += static_cast<unsigned long long>(mine > 0 ? mine : 0);
    }
    if (lead) s.cost[c] = cnt.n_step;
    if (result == kSuccess) {
        double yy[kNeqs];
#pragma unroll
        for (int i = 0; i < kNeqs; ++i) yy[i] = coop_vec(kCoopVecY, i);
        double xo[kNumSpec];
#pragma unroll
        for (int n = 0; n < kNumSpec; ++n) xo[n] = yy[n];
        const double e = yy[kNumSpec];
        const double T = post_step(xo, rho, e);
        if (lane < kNumSpec) {
            double mine = xo[0];
#pragma unroll
            for (int n = 1; n < kNumSpec; ++n) mine = (lane == n) ? xo[n] : mine;
            s.xn[lane * s.cap + c] = mine;
        }
        if (lead) {
            s.e[c] = e;
            s.T[c] = T;
            s.time[c] = next_time;
            s.completed[c] = completed + 1;
            atomicAdd(&ctl->integrated, 1);
        }
    } else {
        if (lead) {
            s.e[c] = e0;  // eos_rt ran before the failed burn
            set_failure(ctl, result, c);
        }
    }
}

// The cooperative waves' loop: take cells from the first ctl->coop_cells
// queue positions (the most expensive cells of the last step), one at a time.
__device__ __forceinline__ void coop_run(DevCells s, Ctl* ctl, int lane) {
    // the programs' constant slots
    coop_u64* kc = reinterpret_cast<coop_u64*>(wr_coop_lds + COOP_OFF_CONST);
    for (int i = lane; i < COOP_NCONST; i += 64) kc[i] = coop_const_bits[i];
    const int ncoop = ctl->coop_cells;
    for (;;) {
        int q = 0;
        if (lane == 0) q = atomicAdd(&ctl->queue_coop, 1);
        q = coop_uniform(__shfl(q, 0));
        if (q >= ncoop || q >= s.n || ctl->failure != kSuccess) break;
        coop_cell(s, ctl, s.order[q], lane);
    }
}

}  // namespace wr
