// Discrete adjoint of the gridwide-independent collapse advance on gfx950
// (no counterpart in the reference, whose forward path is
// reproducer.cpp:7736-7785 with RODAS::integrate 7205-7347).
//
// Differentiated map: every cell's outer steps with the controller decisions
// of a recorded run held fixed -- dt_grid of each step and the step sizes of
// the accepted ROS2S attempts (rejected attempts leave y unchanged).  Per
// step a cell maps (xn, T, rho, driver) -pre-> (y0, rho1, driver1) -accepted
// ROS2S steps-> y_n -post-> (xn', T'), with e' = y_n[14].  Parameters: the
// state at the first advance and the rate multipliers M_r = exp(theta_r) of
// the network's 50 reaction flux terms (the generated f_theta).
//
//   k_adj_record   the advance kernel of the recording run: one thread per
//                  cell, the reference integrator on f_theta, writing the
//                  cell's state at the start of the advance (checkpoint) and
//                  the accepted step sizes (tape);
//   k_adj_replay   the frozen map forward from any state / multipliers;
//   k_adj_backward per cell, steps in reverse: recompute the step from its
//                  checkpoint storing y at each accepted substep, then reverse
//                  the post-step sequence, the substeps (recomputing J, LU
//                  and the stage vectors inside each; transposed LU solves;
//                  generated adjoints of f_theta and J), and the density
//                  update.
// The derivative code is generated from the network DAG (gen/ad.py); staged
// (register-budgeted) forms write their staging slots and outputs to a
// per-thread workspace that is structure of arrays over the cells.
#include "wr_ctx.hpp"
#include "wr_device.hpp"

#include <string.h>

#include <vector>

namespace wr {

// ---- generated adjoint code -------------------------------------------------------------
#define CHEM_FN __device__ __forceinline__
#define AD_STAGE_IN(a) const double* __restrict__ a, int a##_st
#define AD_IN(a, i) (a[(i) * a##_st])
#define AD_STAGE_PARAMS , double* __restrict__ stg, int sstride, double* __restrict__ o1, \
                          double* __restrict__ o2
#define AD_OUT(kind, i, v) AD_OUT_##kind(i, v)
#define AD_OUT_F(i, v) (o1[(i) * sstride] = (v))
#define AD_OUT_J(i, v) (o2[(i) * sstride] = (v))
#define AD_OUT_Y(i, v) (o1[(i) * sstride] = (v))
#define AD_OUT_MB(i, v) (o2[(i) * sstride] = (v))
#include "_gen/chem_adjoint_dev.h"
#undef CHEM_FN

constexpr int kAdjBlock = 64;
constexpr int kNF = 18;  // state fields: xn[14], T, e, rho, driver
constexpr int kR = AD_NTHETA;

// Per-thread workspace (element e of thread c at ws[e * stride + c]).
constexpr int kWJ = 0, kWLU = 225, kWW = 450, kWV = 675;
constexpr int kVF = 0, kVK1 = 15, kVK2 = 30, kVK3 = 45, kVU2 = 60, kVU3 = 75, kVDY = 90,
              kVUB = 105, kVLen = 120;
constexpr int kWMB = kWV + kVLen;        // accumulated dL/dM of the cell
constexpr int kWMT = kWMB + kR;          // dL/dM of one f adjoint
constexpr int kWSTG = kWMT + kR;         // staging slots of the generated code
constexpr int kWYS = kWSTG + AD_NSLOT;   // y at each accepted substep of one step
constexpr int kWsFixed = kWYS;

struct AdjDev {
    double* ckpt;    // [nsteps][18][cap]
    int* hoff;       // [nsteps + 1][cap]: first tape index of each step
    double* htape;   // [hcap][cap]: accepted step sizes
    double* dts;     // [nsteps]
    const double* M; // [R]
    double* ws;      // [kWsFixed + ys_cap * 15][cap]
    int* err;        // first error: -100 a cell did not integrate, -101 tape full
    int hcap, nsteps, cap, ys_cap;
};

__device__ __forceinline__ void adj_error(int* err, int code) { atomicCAS(err, 0, code); }

// The generated pieces as real (non-inlined) functions: each body is compiled
// once, whatever the number of call sites.  w: this thread's workspace column.
__device__ __noinline__ void theta_fj(const double* y, const double* M, double* w, int st,
                                      double* J) {
    ad_fj_st(y, 1, M, 1, w + kWSTG * st, st, w + (kWV + kVF) * st, J);
}
__device__ __noinline__ void theta_f(const double* y, const double* M, double* w, int st,
                                     double* out) {
    ad_f_st(y, 1, M, 1, w + kWSTG * st, st, out, nullptr);
}
__device__ __noinline__ void theta_f_vjp(const double* u, int ust, const double* rb,
                                         const double* M, double* w, int st) {
    ad_f_vjp_st(u, ust, M, 1, rb, 1, w + kWSTG * st, st, w + (kWV + kVUB) * st, w + kWMT * st);
}
__device__ __noinline__ void theta_jac_vjp(const double* y, double* w, int st) {
    ad_jac_vjp_st(y, 1, w + kWW * st, st, w + kWSTG * st, st, w + (kWV + kVUB) * st, nullptr);
}

// ---- the recording run's network: f_theta and J, accepted step sizes to the tape ----
struct ThetaNet {
    const double* M;
    double* w;  // this thread's workspace column
    int st;     // workspace stride
    double* htape;
    int hpos, hcap, cap, c;
    bool overflow;
    __device__ __forceinline__ void jac_rhs(const double (&y)[kNeqs], double (&f)[kNeqs],
                                            double* __restrict__ J, int stride) {
        theta_fj(y, M, w, st, J);
#pragma unroll
        for (int i = 0; i < kNeqs; ++i) f[i] = w[(kWV + kVF + i) * st];
    }
    __device__ __forceinline__ void rhs(const double (&y)[kNeqs], double (&f)[kNeqs]) {
        theta_f(y, M, w, st, w + (kWV + kVF) * st);
#pragma unroll
        for (int i = 0; i < kNeqs; ++i) f[i] = w[(kWV + kVF + i) * st];
    }
    __device__ __forceinline__ void accepted(double h) {
        if (hpos < hcap)
            htape[static_cast<size_t>(hpos) * cap + c] = h;
        else
            overflow = true;
        hpos += 1;
    }
};

__device__ __noinline__ int integrate_theta(double (&y)[kNeqs], double dt, int max_steps,
                                            Counters& cnt, ThetaNet& net) {
    double t_end, dt_end;
    return ros2s_integrate_net(y, dt, dt, max_steps, cnt, net.w + kWJ * net.st,
                               net.w + kWLU * net.st, net.st, t_end, dt_end, net);
}

// advance_collapse_gridwide_kernel (reproducer.cpp:7736-7785) on f_theta, with the
// checkpoint and the tape.  Every cell must integrate.
__global__ __launch_bounds__(kAdjBlock) void k_adj_record(DevCells s, Ctl* ctl, AdjDev a, int step) {
    if (ctl->done) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= s.n || ctl->failure != kSuccess) return;
    const int completed = ctl->completed;
    const double dt_grid = ctl->dt_grid;
    if (threadIdx.x == 0 && blockIdx.x == 0) a.dts[step] = dt_grid;
    if (s.completed[c] != completed) {
        adj_error(a.err, -100);
        return;
    }
    double xn[kNumSpec];
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) xn[n] = s.xn[n * s.cap + c];
    double* ck = a.ckpt + static_cast<size_t>(step) * kNF * a.cap + c;
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) ck[n * a.cap] = xn[n];
    ck[14 * a.cap] = s.T[c];
    ck[15 * a.cap] = s.e[c];
    ck[16 * a.cap] = s.rho[c];
    ck[17 * a.cap] = s.driver[c];
    const double local_dt = collapse_timestep(xn);
    if (local_dt < 10.0) {
        adj_error(a.err, -100);
        return;
    }
    const double old_density = s.driver[c];
    const double tff = local_dt / 1.0e-1;
    const double driver = old_density + dt_grid * (old_density / tff);
    s.driver[c] = driver;
    if (!valid_positive(driver) || driver > 2.0e-6) {
        adj_error(a.err, -100);
        return;
    }
    const double ratio = driver / old_density;
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) {
        xn[n] *= ratio;
        s.xn[n * s.cap + c] = xn[n];
    }
    const double rho = s.rho[c] * ratio;
    s.rho[c] = rho;
    const double e0 = eos_energy(xn, s.T[c]);
    double y[kNeqs];
#pragma unroll
    for (int n = 0; n < kNumSpec; ++n) y[n] = xn[n];
    y[kNumSpec] = e0;
    ThetaNet net{a.M, a.ws + c, a.cap, a.htape, a.hoff[static_cast<size_t>(step) * a.cap + c],
                 a.hcap, a.cap, c, false};
    Counters cnt;
    const int result = integrate_theta(y, dt_grid, ctl->integ_max_steps, cnt, net);
    a.hoff[static_cast<size_t>(step + 1) * a.cap + c] = net.hpos;
    if (net.overflow) adj_error(a.err, -101);
    const int cv[7] = {cnt.n_step, cnt.n_rhs, cnt.n_jac, cnt.n_decomp, cnt.n_solve,
                       cnt.n_accept, cnt.n_reject};
#pragma unroll
    for (int k = 0; k < 7; ++k)
        s.stats[k * s.cap + c] += static_cast<unsigned long long>(cv[k] > 0 ? cv[k] : 0);
    if (result != kSuccess) {
        s.e[c] = e0;
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

// ---- one accepted ROS2S step at a fixed h -------------------------------------------------
// The attempt of the integrator (reproducer.cpp:7260-7302), same operations in the
// same order; J, LU, k1..k3, u2, u3 stay in the workspace for the adjoint.
__device__ __noinline__ int adj_substep(double (&y)[kNeqs], double h, const double* M, double* w,
                                       int st, uint64_t& piv) {
    using C = Ros2sCoef;
    theta_fj(y, M, w, st, w + kWJ * st);
    ThreadWs tw;
    tw.J = w + kWJ * st;
    tw.LU = w + kWLU * st;
    tw.vec = nullptr;
    tw.stage = nullptr;
    tw.stride = st;
    const double fac = 1.0 / (h * C::gamma);
    if (lu_left(tw, fac, piv) != 0) return kLuError;
    double v[kNeqs], u[kNeqs], k1[kNeqs];
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) v[i] = w[(kWV + kVF + i) * st];
    lu_solve_ws(tw, piv, v);
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) {
        k1[i] = v[i];
        w[(kWV + kVK1 + i) * st] = v[i];
        u[i] = y[i] + C::a21 * v[i];
        w[(kWV + kVU2 + i) * st] = u[i];
        v[i] = (C::c21 / h) * v[i];
    }
    theta_f(u, M, w, st, w + (kWV + kVDY) * st);
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) v[i] += w[(kWV + kVDY + i) * st];
    lu_solve_ws(tw, piv, v);
    double k2[kNeqs];
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) {
        k2[i] = v[i];
        w[(kWV + kVK2 + i) * st] = v[i];
        u[i] = y[i] + C::a31 * k1[i] + C::a32 * k2[i];
        w[(kWV + kVU3 + i) * st] = u[i];
        v[i] = (C::c31 * k1[i] + C::c32 * k2[i]) / h;
    }
    theta_f(u, M, w, st, w + (kWV + kVDY) * st);
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) v[i] += w[(kWV + kVDY + i) * st];
    lu_solve_ws(tw, piv, v);
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) {
        w[(kWV + kVK3 + i) * st] = v[i];
        y[i] = y[i] + C::b1 * k1[i] + C::b2 * k2[i] + C::b3 * v[i];
    }
    return kSuccess;
}

// E^T x = b from the LINPACK factors (dgesl with job != 0): U^T, then each L_k^T
// and interchange in reverse elimination order.
__device__ __forceinline__ void lu_solve_t_ws(const double* __restrict__ LU, int st, uint64_t piv,
                                              double (&x)[kNeqs]) {
#pragma unroll
    for (int k = 0; k < kNeqs; ++k) {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < kNeqs; ++j)
            if (j < k) t += WR_AT(LU, j * kNeqs + k, st) * x[j];
        x[k] = (x[k] - t) / WR_AT(LU, k * kNeqs + k, st);
    }
#pragma unroll
    for (int kb = 0; kb < kNeqs - 1; ++kb) {
        const int k = kNeqs - 2 - kb;
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < kNeqs; ++j)
            if (j > k) t += WR_AT(LU, j * kNeqs + k, st) * x[j];
        x[k] += t;
        swap_rows(x, k, piv_get(piv, k));
    }
}

// adjoint of f_theta at u (workspace slot) with seed rb: yb += u_bar, Mbar += M_bar
__device__ __forceinline__ void adj_f(const double* u, int ust, const double (&rb)[kNeqs],
                                      const double* M, double* w, int st, double (&ub)[kNeqs]) {
    theta_f_vjp(u, ust, rb, M, w, st);
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) ub[i] = w[(kWV + kVUB + i) * st];
    for (int r = 0; r < kR; ++r) w[(kWMB + r) * st] += w[(kWMT + r) * st];
}

// adjoint of one accepted step (after adj_substep at the same y): yb in/out
__device__ __forceinline__ void adj_substep_reverse(const double (&y)[kNeqs], double h,
                                                   const double* M, double* w, int st,
                                                   uint64_t piv, double (&yb)[kNeqs]) {
    using C = Ros2sCoef;
    const double* LU = w + kWLU * st;
    double* W = w + kWW * st;
    double kb1[kNeqs], kb2[kNeqs], rb[kNeqs], ub[kNeqs];
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) {
        kb1[i] = C::b1 * yb[i];
        kb2[i] = C::b2 * yb[i];
        rb[i] = C::b3 * yb[i];
    }
    // stage 3
    lu_solve_t_ws(LU, st, piv, rb);
    for (int i = 0; i < kNeqs; ++i)
#pragma unroll
        for (int j = 0; j < kNeqs; ++j) W[(i * kNeqs + j) * st] = rb[i] * w[(kWV + kVK3 + j) * st];
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) {
        kb1[i] += (C::c31 / h) * rb[i];
        kb2[i] += (C::c32 / h) * rb[i];
    }
    adj_f(w + (kWV + kVU3) * st, st, rb, M, w, st, ub);
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) {
        yb[i] += ub[i];
        kb1[i] += C::a31 * ub[i];
        kb2[i] += C::a32 * ub[i];
    }
    // stage 2
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) rb[i] = kb2[i];
    lu_solve_t_ws(LU, st, piv, rb);
    for (int i = 0; i < kNeqs; ++i)
#pragma unroll
        for (int j = 0; j < kNeqs; ++j) W[(i * kNeqs + j) * st] += rb[i] * w[(kWV + kVK2 + j) * st];
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) kb1[i] += (C::c21 / h) * rb[i];
    adj_f(w + (kWV + kVU2) * st, st, rb, M, w, st, ub);
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) {
        yb[i] += ub[i];
        kb1[i] += C::a21 * ub[i];
    }
    // stage 1
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) rb[i] = kb1[i];
    lu_solve_t_ws(LU, st, piv, rb);
    for (int i = 0; i < kNeqs; ++i)
#pragma unroll
        for (int j = 0; j < kNeqs; ++j) W[(i * kNeqs + j) * st] += rb[i] * w[(kWV + kVK1 + j) * st];
    adj_f(y, 1, rb, M, w, st, ub);
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) yb[i] += ub[i];
    // J(y)
    theta_jac_vjp(y, w, st);
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) yb[i] += w[(kWV + kVUB + i) * st];
}

// ---- the frozen map forward ------------------------------------------------------------
__global__ __launch_bounds__(kAdjBlock) void k_adj_replay(AdjDev a, const double* __restrict__ s0,
                                                          double* __restrict__ sN, int n) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double* w = a.ws + c;
    const int st = a.cap;
    double sv[kNF];
#pragma unroll
    for (int f = 0; f < kNF; ++f) sv[f] = s0[static_cast<size_t>(f) * a.cap + c];
    for (int step = 0; step < a.nsteps; ++step) {
        const double pin[4] = {sv[14], sv[16], sv[17], a.dts[step]};
        double y[kNeqs], s1[2];
        ad_pre(sv, pin, y, s1);
        const int k0 = a.hoff[static_cast<size_t>(step) * a.cap + c];
        const int k1 = a.hoff[static_cast<size_t>(step + 1) * a.cap + c];
        for (int k = k0; k < k1; ++k) {
            uint64_t piv;
            if (adj_substep(y, a.htape[static_cast<size_t>(k) * a.cap + c], a.M, w, st, piv) !=
                kSuccess)
                adj_error(a.err, kLuError);
        }
        const double ps[4] = {0.0, s1[0], 0.0, 0.0};
        double xn[kNumSpec], T;
        ad_post(y, ps, xn, &T);
#pragma unroll
        for (int i = 0; i < kNumSpec; ++i) sv[i] = xn[i];
        sv[14] = T;
        sv[15] = y[kNumSpec];
        sv[16] = s1[0];
        sv[17] = s1[1];
    }
#pragma unroll
    for (int f = 0; f < kNF; ++f) sN[static_cast<size_t>(f) * a.cap + c] = sv[f];
}

// ---- the adjoint -------------------------------------------------------------------------
__global__ __launch_bounds__(kAdjBlock) void k_adj_backward(AdjDev a, const double* __restrict__ gN,
                                                            double* __restrict__ g0,
                                                            double* __restrict__ mbar, int n) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double* w = a.ws + c;
    const int st = a.cap;
    double lam[kNF];
#pragma unroll
    for (int f = 0; f < kNF; ++f) lam[f] = gN[static_cast<size_t>(f) * a.cap + c];
    for (int r = 0; r < kR; ++r) w[(kWMB + r) * st] = 0.0;
    for (int step = a.nsteps - 1; step >= 0; --step) {
        double sv[kNF];
        const double* ck = a.ckpt + static_cast<size_t>(step) * kNF * a.cap + c;
#pragma unroll
        for (int f = 0; f < kNF; ++f) sv[f] = ck[f * a.cap];
        const double pin[4] = {sv[14], sv[16], sv[17], a.dts[step]};
        double y[kNeqs], s1[2];
        ad_pre(sv, pin, y, s1);
        const int k0 = a.hoff[static_cast<size_t>(step) * a.cap + c];
        const int k1 = a.hoff[static_cast<size_t>(step + 1) * a.cap + c];
        const int nk = k1 - k0;
        if (nk + 1 > a.ys_cap) {
            adj_error(a.err, -101);
            return;
        }
        double* ys = w + kWYS * st;
#pragma unroll
        for (int i = 0; i < kNeqs; ++i) ys[i * st] = y[i];
        for (int k = 0; k < nk; ++k) {
            uint64_t piv;
            adj_substep(y, a.htape[static_cast<size_t>(k0 + k) * a.cap + c], a.M, w, st, piv);
#pragma unroll
            for (int i = 0; i < kNeqs; ++i) ys[((k + 1) * kNeqs + i) * st] = y[i];
        }
        // post-step sequence
        const double ps[4] = {0.0, s1[0], 0.0, 0.0};
        const double wT[1] = {lam[14]};
        double yb[kNeqs], rb[1];
        ad_post_vjp(y, ps, lam, wT, yb, rb);
        yb[kNumSpec] += lam[15];
        const double rho1b = lam[16] + rb[0];
        // substeps in reverse, each recomputed from its stored y
        for (int k = nk - 1; k >= 0; --k) {
            double yk[kNeqs], yt[kNeqs];
#pragma unroll
            for (int i = 0; i < kNeqs; ++i) {
                yk[i] = ys[(k * kNeqs + i) * st];
                yt[i] = yk[i];
            }
            const double h = a.htape[static_cast<size_t>(k0 + k) * a.cap + c];
            uint64_t piv;
            adj_substep(yt, h, a.M, w, st, piv);
            adj_substep_reverse(yk, h, a.M, w, st, piv, yb);
        }
        // density update and eos_rt
        const double ws2[2] = {rho1b, lam[17]};
        double xb[kNumSpec], sb[3];
        ad_pre_vjp(sv, pin, yb, ws2, xb, sb);
#pragma unroll
        for (int i = 0; i < kNumSpec; ++i) lam[i] = xb[i];
        lam[14] = sb[0];
        lam[15] = 0.0;
        lam[16] = sb[1];
        lam[17] = sb[2];
    }
#pragma unroll
    for (int f = 0; f < kNF; ++f) g0[static_cast<size_t>(f) * a.cap + c] = lam[f];
    for (int r = 0; r < kR; ++r) mbar[static_cast<size_t>(r) * a.cap + c] = w[(kWMB + r) * st];
}

// ---- host state ------------------------------------------------------------------------------
struct AdjointState {
    int nsteps = 0, hcap = 0, ys_cap = 0, cap = 0, n = 0;
    bool recorded = false;
    double* ckpt = nullptr;
    int* hoff = nullptr;
    double* htape = nullptr;
    double* dts = nullptr;
    double* M = nullptr;   // multipliers of the recording
    double* Mr = nullptr;  // multipliers of a replay
    double* ws = nullptr;
    size_t ws_elems = 0;
    int* err = nullptr;
    double* io = nullptr;  // device staging for states / seeds: 3 x [18][cap] + [R][cap]
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
};

}  // namespace wr

using namespace wr;
using namespace wr::detail;

void wr::detail::adjoint_destroy(wr_context* ctx) {
    AdjointState* a = ctx->adj;
    if (!a) return;
    (void)hipFree(a->ckpt);
    (void)hipFree(a->hoff);
    (void)hipFree(a->htape);
    (void)hipFree(a->dts);
    (void)hipFree(a->M);
    (void)hipFree(a->Mr);
    (void)hipFree(a->ws);
    (void)hipFree(a->err);
    (void)hipFree(a->io);
    if (a->ev0) (void)hipEventDestroy(a->ev0);
    if (a->ev1) (void)hipEventDestroy(a->ev1);
    delete a;
    ctx->adj = nullptr;
}

namespace {

AdjDev dev_view(const AdjointState* a, const double* M) {
    AdjDev d;
    d.ckpt = a->ckpt;
    d.hoff = a->hoff;
    d.htape = a->htape;
    d.dts = a->dts;
    d.M = M;
    d.ws = a->ws;
    d.err = a->err;
    d.hcap = a->hcap;
    d.nsteps = a->nsteps;
    d.cap = a->cap;
    d.ys_cap = a->ys_cap;
    return d;
}

int ensure_ws(wr_context* ctx, AdjointState* a, int ys_cap) {
    const size_t need = static_cast<size_t>(kWsFixed + ys_cap * kNeqs) * a->cap;
    if (a->ws && a->ws_elems >= need) {
        a->ys_cap = static_cast<int>((a->ws_elems / a->cap - kWsFixed) / kNeqs);
        return WR_SUCCESS;
    }
    (void)hipFree(a->ws);
    a->ws = nullptr;
    WR_HIP(hipMalloc(&a->ws, need * sizeof(double)), "hipMalloc(adjoint workspace)");
    a->ws_elems = need;
    a->ys_cap = ys_cap;
    return WR_SUCCESS;
}

int read_err(wr_context* ctx, AdjointState* a, int* code) {
    WR_HIP(hipMemcpyAsync(code, a->err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream),
           "hipMemcpy(adjoint status)");
    WR_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return WR_SUCCESS;
}

int timed_begin(wr_context* ctx, AdjointState* a) {
    WR_HIP(hipEventRecord(a->ev0, ctx->stream), "hipEventRecord");
    return WR_SUCCESS;
}

int timed_end(wr_context* ctx, AdjointState* a) {
    WR_HIP(hipEventRecord(a->ev1, ctx->stream), "hipEventRecord");
    WR_HIP(hipEventSynchronize(a->ev1), "hipEventSynchronize");
    float ms = 0.0f;
    WR_HIP(hipEventElapsedTime(&ms, a->ev0, a->ev1), "hipEventElapsedTime");
    a->last_ms = ms;
    return WR_SUCCESS;
}

}  // namespace

extern "C" {

int wr_adjoint_ntheta(void) { return kR; }

int wr_adjoint_record(wr_context* ctx, int nsteps, const double* multipliers, int hcap,
                      int* completed_global_steps) {
    if (!ctx || nsteps <= 0 || !multipliers || hcap <= 0 || !completed_global_steps)
        return fail(WR_ERR_ARGUMENT, "wr_adjoint_record: bad argument");
    if (ctx->n <= 0) return fail(WR_ERR_ARGUMENT, "wr_adjoint_record: no cells uploaded");
    WR_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    if (!ctx->adj) ctx->adj = new AdjointState();
    AdjointState* a = ctx->adj;
    a->recorded = false;
    if (a->cap != ctx->cap || a->nsteps < nsteps || a->hcap < hcap || !a->ckpt) {
        (void)hipFree(a->ckpt);
        (void)hipFree(a->hoff);
        (void)hipFree(a->htape);
        (void)hipFree(a->dts);
        a->ckpt = nullptr;
        a->hoff = nullptr;
        a->htape = nullptr;
        a->dts = nullptr;
        const size_t cap = static_cast<size_t>(ctx->cap);
        WR_HIP(hipMalloc(&a->ckpt, static_cast<size_t>(nsteps) * kNF * cap * sizeof(double)),
               "hipMalloc(adjoint checkpoints)");
        WR_HIP(hipMalloc(&a->hoff, static_cast<size_t>(nsteps + 1) * cap * sizeof(int)),
               "hipMalloc(adjoint tape index)");
        WR_HIP(hipMalloc(&a->htape, static_cast<size_t>(hcap) * cap * sizeof(double)),
               "hipMalloc(adjoint tape)");
        WR_HIP(hipMalloc(&a->dts, static_cast<size_t>(nsteps) * sizeof(double)),
               "hipMalloc(adjoint dt)");
        if (!a->M) WR_HIP(hipMalloc(&a->M, kR * sizeof(double)), "hipMalloc(multipliers)");
        if (!a->Mr) WR_HIP(hipMalloc(&a->Mr, kR * sizeof(double)), "hipMalloc(multipliers)");
        if (!a->err) WR_HIP(hipMalloc(&a->err, sizeof(int)), "hipMalloc(adjoint status)");
        if (!a->io)
            WR_HIP(hipMalloc(&a->io, (3 * kNF + kR) * cap * sizeof(double)), "hipMalloc(adjoint io)");
        if (!a->ev0) WR_HIP(hipEventCreate(&a->ev0), "hipEventCreate");
        if (!a->ev1) WR_HIP(hipEventCreate(&a->ev1), "hipEventCreate");
        a->cap = ctx->cap;
        a->nsteps = nsteps;
        a->hcap = hcap;
    }
    a->nsteps = nsteps;
    a->n = ctx->n;
    int st = ensure_ws(ctx, a, 1);
    if (st != WR_SUCCESS) return st;
    WR_HIP(hipMemcpyAsync(a->M, multipliers, kR * sizeof(double), hipMemcpyHostToDevice,
                          ctx->stream),
           "hipMemcpy(multipliers)");
    WR_HIP(hipMemsetAsync(a->err, 0, sizeof(int), ctx->stream), "hipMemset");
    WR_HIP(hipMemsetAsync(a->hoff, 0, static_cast<size_t>(a->cap) * sizeof(int), ctx->stream),
           "hipMemset");
    if ((st = ctl_init(ctx, 0, nsteps, 0, 0.0)) != WR_SUCCESS) return st;
    const AdjDev d = dev_view(a, a->M);
    if ((st = timed_begin(ctx, a)) != WR_SUCCESS) return st;
    for (int step = 0; step < nsteps; ++step) {
        if ((st = ctl_prepare(ctx)) != WR_SUCCESS) return st;
        hipLaunchKernelGGL(k_adj_record, dim3(blocks_for(ctx->n, kAdjBlock)), dim3(kAdjBlock), 0,
                           ctx->stream, ctx->cells, ctx->ctl, d, step);
        WR_HIP(hipGetLastError(), "k_adj_record launch");
        if ((st = ctl_commit(ctx)) != WR_SUCCESS) return st;
    }
    if ((st = timed_end(ctx, a)) != WR_SUCCESS) return st;
    if ((st = ctl_read(ctx)) != WR_SUCCESS) return st;
    ctl_fill_info(ctx);
    int code = 0;
    if ((st = read_err(ctx, a, &code)) != WR_SUCCESS) return st;
    *completed_global_steps = ctx->ctl_host->completed;
    if (ctx->ctl_host->failure != kSuccess) return ctx->ctl_host->failure;
    if (code == -101) return fail(WR_ERR_CAPACITY, "wr_adjoint_record: tape capacity exceeded");
    if (code != 0 || ctx->ctl_host->completed != nsteps)
        return fail(WR_ERR_ARGUMENT,
                    "wr_adjoint_record: every cell must integrate in every outer step");
    // largest number of accepted steps of a cell in one outer step
    std::vector<int> hoff(static_cast<size_t>(nsteps + 1) * a->cap);
    WR_HIP(hipMemcpy(hoff.data(), a->hoff, hoff.size() * sizeof(int), hipMemcpyDeviceToHost),
           "hipMemcpy(adjoint tape index)");
    int most = 0;
    for (int s = 0; s < nsteps; ++s)
        for (int c = 0; c < ctx->n; ++c) {
            const int k = hoff[static_cast<size_t>(s + 1) * a->cap + c] -
                          hoff[static_cast<size_t>(s) * a->cap + c];
            if (k > most) most = k;
        }
    if ((st = ensure_ws(ctx, a, most + 1)) != WR_SUCCESS) return st;
    a->recorded = true;
    return kSuccess;
}

int wr_adjoint_tape(const wr_context* ctx, int32_t* hoff, double* htape, double* dts,
                    double* ckpt) {
    if (!ctx || !ctx->adj || !ctx->adj->recorded)
        return fail(WR_ERR_ARGUMENT, "wr_adjoint_tape: nothing recorded");
    const AdjointState* a = ctx->adj;
    WR_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    const size_t cap = static_cast<size_t>(a->cap), n = static_cast<size_t>(a->n);
    for (int s = 0; s <= a->nsteps; ++s)
        if (hoff)
            WR_HIP(hipMemcpy(hoff + s * n, a->hoff + s * cap, n * sizeof(int), hipMemcpyDeviceToHost),
                   "hipMemcpy");
    if (htape)
        for (int k = 0; k < a->hcap; ++k)
            WR_HIP(hipMemcpy(htape + k * n, a->htape + k * cap, n * sizeof(double),
                             hipMemcpyDeviceToHost),
                   "hipMemcpy");
    if (dts) WR_HIP(hipMemcpy(dts, a->dts, a->nsteps * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    if (ckpt)
        for (int f = 0; f < a->nsteps * kNF; ++f)
            WR_HIP(hipMemcpy(ckpt + f * n, a->ckpt + f * cap, n * sizeof(double), hipMemcpyDeviceToHost),
                   "hipMemcpy");
    return WR_SUCCESS;
}

static int put_fields(wr_context* ctx, double* dst, const double* src, int rows) {
    const size_t cap = static_cast<size_t>(ctx->adj->cap), n = static_cast<size_t>(ctx->adj->n);
    for (int f = 0; f < rows; ++f)
        WR_HIP(hipMemcpyAsync(dst + f * cap, src + f * n, n * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream),
               "hipMemcpy(to device)");
    return WR_SUCCESS;
}

static int get_fields(wr_context* ctx, double* dst, const double* src, int rows) {
    const size_t cap = static_cast<size_t>(ctx->adj->cap), n = static_cast<size_t>(ctx->adj->n);
    for (int f = 0; f < rows; ++f)
        WR_HIP(hipMemcpyAsync(dst + f * n, src + f * cap, n * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream),
               "hipMemcpy(to host)");
    WR_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return WR_SUCCESS;
}

int wr_adjoint_replay(wr_context* ctx, const double* s0, const double* multipliers, double* sN) {
    if (!ctx || !ctx->adj || !ctx->adj->recorded || !sN)
        return fail(WR_ERR_ARGUMENT, "wr_adjoint_replay: nothing recorded or bad argument");
    AdjointState* a = ctx->adj;
    WR_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    const size_t cap = static_cast<size_t>(a->cap);
    double* d_s0 = a->io;
    double* d_sN = a->io + kNF * cap;
    int st;
    if (s0) {
        if ((st = put_fields(ctx, d_s0, s0, kNF)) != WR_SUCCESS) return st;
    } else {
        WR_HIP(hipMemcpyAsync(d_s0, a->ckpt, kNF * cap * sizeof(double), hipMemcpyDeviceToDevice,
                              ctx->stream),
               "hipMemcpy(checkpoint)");
    }
    const double* M = a->M;
    if (multipliers) {
        WR_HIP(hipMemcpyAsync(a->Mr, multipliers, kR * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream),
               "hipMemcpy(multipliers)");
        M = a->Mr;
    }
    WR_HIP(hipMemsetAsync(a->err, 0, sizeof(int), ctx->stream), "hipMemset");
    if ((st = timed_begin(ctx, a)) != WR_SUCCESS) return st;
    hipLaunchKernelGGL(k_adj_replay, dim3(blocks_for(a->n, kAdjBlock)), dim3(kAdjBlock), 0,
                       ctx->stream, dev_view(a, M), d_s0, d_sN, a->n);
    WR_HIP(hipGetLastError(), "k_adj_replay launch");
    if ((st = timed_end(ctx, a)) != WR_SUCCESS) return st;
    if ((st = get_fields(ctx, sN, d_sN, kNF)) != WR_SUCCESS) return st;
    int code = 0;
    if ((st = read_err(ctx, a, &code)) != WR_SUCCESS) return st;
    return code == 0 ? kSuccess : kLuError;
}

int wr_adjoint_backward(wr_context* ctx, const double* gN, double* g0, double* mbar) {
    if (!ctx || !ctx->adj || !ctx->adj->recorded || !gN || !g0 || !mbar)
        return fail(WR_ERR_ARGUMENT, "wr_adjoint_backward: nothing recorded or bad argument");
    AdjointState* a = ctx->adj;
    WR_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    const size_t cap = static_cast<size_t>(a->cap);
    double* d_gN = a->io + 2 * kNF * cap;
    double* d_g0 = a->io;
    double* d_mb = a->io + 3 * kNF * cap;
    int st;
    if ((st = put_fields(ctx, d_gN, gN, kNF)) != WR_SUCCESS) return st;
    WR_HIP(hipMemsetAsync(a->err, 0, sizeof(int), ctx->stream), "hipMemset");
    if ((st = timed_begin(ctx, a)) != WR_SUCCESS) return st;
    hipLaunchKernelGGL(k_adj_backward, dim3(blocks_for(a->n, kAdjBlock)), dim3(kAdjBlock), 0,
                       ctx->stream, dev_view(a, a->M), d_gN, d_g0, d_mb, a->n);
    WR_HIP(hipGetLastError(), "k_adj_backward launch");
    if ((st = timed_end(ctx, a)) != WR_SUCCESS) return st;
    if ((st = get_fields(ctx, g0, d_g0, kNF)) != WR_SUCCESS) return st;
    if ((st = get_fields(ctx, mbar, d_mb, kR)) != WR_SUCCESS) return st;
    int code = 0;
    if ((st = read_err(ctx, a, &code)) != WR_SUCCESS) return st;
    if (code == -101) return fail(WR_ERR_CAPACITY, "wr_adjoint_backward: substep workspace full");
    return code == 0 ? kSuccess : kLuError;
}

int wr_adjoint_last_ms(const wr_context* ctx, double* ms) {
    if (!ctx || !ctx->adj || !ms) return fail(WR_ERR_ARGUMENT, "wr_adjoint_last_ms: bad argument");
    *ms = ctx->adj->last_ms;
    return WR_SUCCESS;
}

}  // extern "C"
