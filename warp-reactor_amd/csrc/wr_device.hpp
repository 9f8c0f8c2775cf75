// Device-side building blocks of the primordial-chemistry ROS2S grid advance
// for gfx950: EOS and species helpers, the generated network (RHS + Jacobian),
// dense 15x15 LU with partial pivoting, the ROS2S integrator and the per-cell
// prepare/advance bodies.  All arithmetic is FP64 and every operation follows
// the reference's operand order (compiled with -ffp-contract=off), so results
// differ from the reference CPU build only through the device libm
// (exp/log/cbrt).
//
// Layout: per-cell state is structure-of-arrays (one array per field, cell
// index fastest), so a wavefront's 64 lanes touch 512 contiguous bytes per
// field.  The per-cell Jacobian and LU workspaces are SoA too: element (i, j)
// of cell c lives at base[(i * 15 + j) * stride + c].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wr_eos.hpp"
#include "wr_math.h"

namespace wr {

// ---- libm ---------------------------------------------------------------------
// exp/log/cbrt from wr_math.h: only correctly rounded operations, so a host
// build of the same algorithm reproduces the device bits.  Building with
// -DWR_DEVICE_LIBM_OCML selects the ROCm device library instead.
#if defined(WR_DEVICE_LIBM_OCML)
#define CHEM_EXP(x) exp(x)
#define CHEM_LOG(x) log(x)
#define WR_CBRT(x) cbrt(x)
#else
#define CHEM_EXP(x) wr_exp(x)
#define CHEM_LOG(x) wr_log(x)
#define WR_CBRT(x) wr_cbrt(x)
#endif

// ---- generated network --------------------------------------------------------
#define CHEM_FN __device__ __forceinline__
#define CHEM_SINK_PARAMS , double* __restrict__ jsink, int jstride
#define CHEM_JAC_STORE(r, c, v) (jsink[((r) * kNeqs + (c)) * jstride] = (v))
// register-budgeted variants: staging slots and outputs in a per-thread SoA
// workspace with the given stride
#define CHEM_STAGE_PARAMS                                                                   \
    , double* __restrict__ stg, int sstride, double* __restrict__ jout,                    \
        double* __restrict__ fout
#define STAGE_STORE(s, v) (stg[(s) * sstride] = (v))
#define STAGE_LOAD(s) (stg[(s) * sstride])
#define CHEM_OUT(kind, ...) CHEM_OUT_##kind(__VA_ARGS__)
#define CHEM_OUT_J(r, c, v) (jout[((r) * kNeqs + (c)) * sstride] = (v))
#define CHEM_OUT_F(i, v) (fout[(i) * sstride] = (v))
#include "_gen/chem_network.h"
#undef CHEM_JAC_STORE
#undef CHEM_SINK_PARAMS
#undef CHEM_FN

// f(y) (PrimordialChem::rhs, reproducer.cpp:7040-7050)
__device__ __forceinline__ void rhs(const double (&y)[kNeqs], double (&f)[kNeqs]) {
    double X[kNumSpec];
    const double T = state_from_y(y, X);
    chem_rhs(T, X, f);
}

// J(y) into the SoA workspace and f(y) (reproducer.cpp:7052-7060 + 7040-7050)
__device__ __forceinline__ void jac_rhs(const double (&y)[kNeqs], double (&f)[kNeqs],
                                        double* __restrict__ J, int stride) {
    double X[kNumSpec];
    const double T = state_from_y(y, X);
    chem_fused(T, X, f, J, stride);
}

// ---- LU, LINPACK conventions (reproducer.cpp:96-151) ----------------------------
// A: this cell's matrix in the SoA workspace, element (i, j) at A[(i*15+j)*stride].
// Column k is held in registers for the pivot search, exchange and scaling;
// the trailing update streams rows through memory.  The pivot is the first
// strict maximum of |a_ik|; rows are exchanged from column k on (multipliers
// of earlier columns stay in place) and multipliers are stored negated.
__device__ __forceinline__ int lu_factor(double* __restrict__ A, int stride, int (&ipvt)[kNeqs]) {
    int info = 0;
#pragma unroll
    for (int k = 0; k < kNeqs - 1; ++k) {
        double col[kNeqs];
#pragma unroll
        for (int i = k; i < kNeqs; ++i) col[i] = A[(i * kNeqs + k) * stride];
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
        ipvt[k] = p;
        double piv = col[k];
#pragma unroll
        for (int i = k + 1; i < kNeqs; ++i) piv = (p == i) ? col[i] : piv;
        if (piv == 0.0) {
            info = k + 1;
            continue;
        }
        if (p != k) {
            // exchange rows k and p from column k+1 on; column k in registers
#pragma unroll
            for (int j = k + 1; j < kNeqs; ++j) {
                double* rp = A + (p * kNeqs + j) * stride;
                double* rk = A + (k * kNeqs + j) * stride;
                const double tp = *rp;
                *rp = *rk;
                *rk = tp;
            }
            const double ck = col[k];
#pragma unroll
            for (int i = k + 1; i < kNeqs; ++i) col[i] = (p == i) ? ck : col[i];
            col[k] = piv;
            A[(k * kNeqs + k) * stride] = piv;
        }
        const double m = -1.0 / col[k];
#pragma unroll
        for (int i = k + 1; i < kNeqs; ++i) {
            col[i] *= m;
            A[(i * kNeqs + k) * stride] = col[i];
        }
#pragma unroll
        for (int j = k + 1; j < kNeqs; ++j) {
            const double t = A[(k * kNeqs + j) * stride];
#pragma unroll
            for (int i = k + 1; i < kNeqs; ++i) {
                double* a = A + (i * kNeqs + j) * stride;
                *a += t * col[i];
            }
        }
    }
    ipvt[kNeqs - 1] = kNeqs - 1;
    if (A[(kNeqs * kNeqs - 1) * stride] == 0.0) info = kNeqs;
    return info;
}

// lu_solve (reproducer.cpp:154-181).  Row exchanges of x are done with selects
// so x stays in registers (no dynamically indexed private array).
__device__ __forceinline__ void lu_solve(const double* __restrict__ LU, int stride,
                                         const int (&ipvt)[kNeqs], double (&x)[kNeqs]) {
#pragma unroll
    for (int k = 0; k < kNeqs - 1; ++k) {
        const int p = ipvt[k];
        double t = x[k];
#pragma unroll
        for (int i = k + 1; i < kNeqs; ++i) t = (p == i) ? x[i] : t;
        const double xk = x[k];
#pragma unroll
        for (int i = k + 1; i < kNeqs; ++i) x[i] = (p == i) ? xk : x[i];
        x[k] = t;
#pragma unroll
        for (int j = k + 1; j < kNeqs; ++j) x[j] += t * LU[(j * kNeqs + k) * stride];
    }
#pragma unroll
    for (int kb = 0; kb < kNeqs; ++kb) {
        const int k = kNeqs - 1 - kb;
        x[k] /= LU[(k * kNeqs + k) * stride];
        const double t = -x[k];
#pragma unroll
        for (int j = 0; j < k; ++j) x[j] += t * LU[(j * kNeqs + k) * stride];
    }
}

// ---- ROS2S (reproducer.cpp:7066-7347) --------------------------------------------
struct Ros2sCoef {
    static constexpr double gamma = 0.292893218813452;
    static constexpr double a21 = 2.0000000000000036;
    static constexpr double a31 = 6.828427124746214;
    static constexpr double a32 = 3.4142135623731007;
    static constexpr double c21 = -6.828427124746214;
    static constexpr double c31 = -10.949747468305889;
    static constexpr double c32 = -7.535533905932761;
    static constexpr double b1 = 6.828427124746214;
    static constexpr double b2 = 3.414213562373101;
    static constexpr double b3 = 1.0;
    static constexpr double e1 = -0.23570226039551292;
    static constexpr double e2 = -0.23570226039551567;
    static constexpr double e3 = -0.13807118745769906;
};
constexpr double kUround = 1.0e-16, kFacMin = 0.2, kFacMax = 6.0, kSafe = 0.9;

// Collapse tolerances (configure_ros2s, reproducer.cpp:7409-7412, 7582-7590).
__device__ __forceinline__ constexpr double rtol_of(int i) { return i < kNumSpec ? 1.0e-4 : 1.0e-6; }
__device__ __forceinline__ constexpr double atol_of(int i) { return i < kNumSpec ? 1.0e-4 : 1.0e-6; }

struct Counters {
    int n_step, n_rhs, n_jac, n_decomp, n_solve, n_accept, n_reject;
};

// The network an integration evaluates: the generated RHS and Jacobian, and a
// hook that sees the step size of every accepted attempt.
struct PlainNet {
    __device__ __forceinline__ void jac_rhs(const double (&y)[kNeqs], double (&f)[kNeqs],
                                            double* __restrict__ J, int stride) {
        wr::jac_rhs(y, f, J, stride);
    }
    __device__ __forceinline__ void rhs(const double (&y)[kNeqs], double (&f)[kNeqs]) {
        wr::rhs(y, f);
    }
    __device__ __forceinline__ void accepted(double) {}
};

// RODAS<PrimordialChem>::integrate from t = 0 to tout (reproducer.cpp:7205-7347).
// J and LU point at this cell's SoA workspace columns.  The RHS at the start
// of an attempt, f(y), depends only on y, so it is evaluated together with
// the Jacobian and reused by every retry of the same step: the reference
// recomputes it per retry and gets the same bits.
template <class Net>
__device__ __forceinline__ int ros2s_integrate_net(double (&y)[kNeqs], double tout, double dt0,
                                                   int max_steps, Counters& c,
                                                   double* __restrict__ J, double* __restrict__ LU,
                                                   int stride, double& t_end, double& dt_end,
                                                   Net& net) {
    using C = Ros2sCoef;
    c = Counters{0, 0, 0, 0, 0, 0, 0};
    if (tout <= 0.0) {
        t_end = 0.0;
        dt_end = dt0;
        return kBadInputs;
    }
    const double hmaxn = tout - 0.0;
    double h = dt0;
    if (fabs(h) <= 10.0 * kUround) h = 1.0e-6;
    h = rmin(fabs(h), hmaxn);

    bool reject = false, last = false;
    int nsing = 0, n_step = 0, n_accept = 0;
    double hacc = 0.0, erracc = 1.0, hopt = h, x = 0.0;
    double f0[kNeqs];
    int ipvt[kNeqs];

    for (;;) {
        if (n_step > max_steps) {
            t_end = x;
            dt_end = h;
            return kTooManySteps;
        }
        if (0.1 * fabs(h) <= fabs(x) * kUround) {
            t_end = x;
            dt_end = h;
            return kDtUnderflow;
        }
        if (last) {
            t_end = x;
            dt_end = hopt;
            return kSuccess;
        }
        hopt = h;
        if (x + h * 1.0001 >= tout) {
            h = tout - x;
            last = true;
        }
        net.jac_rhs(y, f0, J, stride);
        c.n_jac += 1;

        for (;;) {
            const double fac = 1.0 / (h * C::gamma);
            // E = -J + fac I into the LU workspace
#pragma unroll
            for (int i = 0; i < kNeqs; ++i) {
#pragma unroll
                for (int j = 0; j < kNeqs; ++j) {
                    const double v = -J[(i * kNeqs + j) * stride];
                    LU[(i * kNeqs + j) * stride] = (i == j) ? v + fac : v;
                }
            }
            if (lu_factor(LU, stride, ipvt) != 0) {
                nsing += 1;
                if (nsing >= 5) {
                    t_end = x;
                    dt_end = h;
                    return kLuError;
                }
                h *= 0.5;
                reject = true;
                last = false;
                continue;
            }
            c.n_decomp += 1;

            double ak1[kNeqs], ak2[kNeqs], ak3[kNeqs], ynew[kNeqs], dy[kNeqs];
#pragma unroll
            for (int i = 0; i < kNeqs; ++i) ak1[i] = f0[i];
            lu_solve(LU, stride, ipvt, ak1);
            c.n_solve += 1;
#pragma unroll
            for (int i = 0; i < kNeqs; ++i) {
                ynew[i] = y[i] + C::a21 * ak1[i];
                ak2[i] = (C::c21 / h) * ak1[i];
            }
            net.rhs(ynew, dy);
#pragma unroll
            for (int i = 0; i < kNeqs; ++i) ak2[i] += dy[i];
            lu_solve(LU, stride, ipvt, ak2);
            c.n_solve += 1;
#pragma unroll
            for (int i = 0; i < kNeqs; ++i) {
                ynew[i] = y[i] + C::a31 * ak1[i] + C::a32 * ak2[i];
                ak3[i] = (C::c31 * ak1[i] + C::c32 * ak2[i]) / h;
            }
            net.rhs(ynew, dy);
#pragma unroll
            for (int i = 0; i < kNeqs; ++i) ak3[i] += dy[i];
            lu_solve(LU, stride, ipvt, ak3);
            c.n_solve += 1;
            double err = 0.0;
#pragma unroll
            for (int i = 0; i < kNeqs; ++i) {
                ynew[i] = y[i] + C::b1 * ak1[i] + C::b2 * ak2[i] + C::b3 * ak3[i];
                const double w = C::e1 * ak1[i] + C::e2 * ak2[i] + C::e3 * ak3[i];
                // error_norm (reproducer.cpp:7194-7202)
                const double sk = atol_of(i) + rtol_of(i) * rmax(fabs(y[i]), fabs(ynew[i]));
                const double term = w / sk;
                err += term * term;
            }
            c.n_rhs += 3;
            n_step += 1;
            c.n_step = n_step;
            err = sqrt(err / static_cast<double>(kNeqs));

            const double raw_fac = WR_CBRT(err) / kSafe;
            const double fac_step = rmax(1.0 / kFacMax, rmin(1.0 / kFacMin, raw_fac));
            double hnew = h / fac_step;
            if (err <= 1.0) {
                n_accept += 1;
                c.n_accept = n_accept;
                if (n_accept > 1) {
                    const double facgus =
                        rmax(1.0 / kFacMax,
                             rmin(1.0 / kFacMin, (hacc / h) * WR_CBRT((err * err) / erracc) / kSafe));
                    hnew = h / rmax(fac_step, facgus);
                }
                hacc = h;
                erracc = rmax(1.0e-2, err);
                net.accepted(h);
#pragma unroll
                for (int i = 0; i < kNeqs; ++i) y[i] = ynew[i];
                x += h;
                if (fabs(hnew) > hmaxn) hnew = hmaxn;
                if (reject) hnew = rmin(fabs(hnew), fabs(h));
                reject = false;
                h = hnew;
                break;
            }
            reject = true;
            last = false;
            h = hnew;
            if (n_accept >= 1) c.n_reject += 1;
        }
    }
}

__device__ __forceinline__ int ros2s_integrate(double (&y)[kNeqs], double tout, double dt0,
                                               int max_steps, Counters& c,
                                               double* __restrict__ J, double* __restrict__ LU,
                                               int stride, double& t_end, double& dt_end) {
    PlainNet net;
    return ros2s_integrate_net(y, tout, dt0, max_steps, c, J, LU, stride, t_end, dt_end, net);
}

// =====================================================================================
// Queue-fed integrator: per-thread workspace, staged network, left-looking LU
// =====================================================================================

// Per-thread workspace, structure of arrays over the threads of a launch:
// element e of thread t at base[e * stride + t] (a wave touches 512
// contiguous bytes per element).
struct ThreadWs {
    double* J;      // [225]  Jacobian at y (kept for retries)
    double* LU;     // [225]  factors of E = fac I - J
    double* vec;    // [kVecLen] y, f0, ak1, ak2, ak3, ynew, dy
    double* stage;  // [kStageSlots] staging slots of the network code
    int stride;
};
constexpr int kVecY = 0, kVecF0 = 15, kVecAk1 = 30, kVecAk2 = 45, kVecAk3 = 60, kVecYnew = 75,
              kVecDy = 90, kVecLen = 105;
constexpr int kStageSlots = (CHEM_FUSED_NSLOT > CHEM_RHS_NSLOT) ? CHEM_FUSED_NSLOT : CHEM_RHS_NSLOT;

#define WR_AT(base, e, stride) ((base)[(e) * (stride)])

// J(y) and f(y) with the register-budgeted network code; y is read from and
// J, f0 written to the workspace.  Inlined into the advance kernel; the build
// turns off machine-level loop-invariant code motion, which would otherwise
// hoist the network's ~1100 FP64 constants out of the persistent loop and
// spill them (over 1700 SGPR spills).  -DWR_NET_NOINLINE keeps the two
// network wrappers as separate functions instead.
#ifdef WR_NET_NOINLINE
#define WR_NET_FN __device__ __noinline__
#else
#define WR_NET_FN __device__ __forceinline__
#endif
WR_NET_FN void ws_jac_rhs(const ThreadWs& w) {
    double X[kNumSpec];
    double yy[kNeqs];
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) yy[i] = WR_AT(w.vec, kVecY + i, w.stride);
    const double T = state_from_y(yy, X);
    double* __restrict__ stg = w.stage;
    double* __restrict__ J = w.J;
    double* __restrict__ F = w.vec + kVecF0 * w.stride;
    const int stride = w.stride;
    chem_fused_staged(T, X, stg, stride, J, F);
}

// f(ynew) into dy.
WR_NET_FN void ws_rhs(const ThreadWs& w) {
    double X[kNumSpec];
    double yy[kNeqs];
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) yy[i] = WR_AT(w.vec, kVecYnew + i, w.stride);
    const double T = state_from_y(yy, X);
    double* __restrict__ stg = w.stage;
    double* __restrict__ F = w.vec + kVecDy * w.stride;
    const int stride = w.stride;
    chem_rhs_staged(T, X, stg, stride, nullptr, F);
}

__device__ __forceinline__ int piv_get(uint64_t piv, int k) { return (int)((piv >> (4 * k)) & 15u); }

// Row exchange k <-> p of a register vector (p >= k, dynamic), with selects;
// skipped when no lane of the wave exchanges.
template <typename R>
__device__ __forceinline__ void swap_rows(R (&v)[kNeqs], int k, int p) {
    if (__ballot(p != k) == 0) return;
    const R vk = v[k];
    R vp = vk;
#pragma unroll
    for (int i = 0; i < kNeqs; ++i)
        if (i > k) vp = (p == i) ? v[i] : vp;
#pragma unroll
    for (int i = 0; i < kNeqs; ++i)
        if (i > k) v[i] = (p == i) ? vk : v[i];
    v[k] = vp;
}

// LU of E = fac I - J, left-looking (column by column), with exactly the
// operations of the reference's right-looking LINPACK elimination
// (reproducer.cpp:97-151): each element receives the same updates
// a_ij += a_kj * m_ik in the same k order, pivots are the first strict
// maximum, rows are exchanged from column k on and multipliers are stored
// negated; a zero pivot skips its exchange, scaling and updates.  Column j
// lives in registers; the multipliers of the first kCache columns stay in
// registers too, the rest are re-read from LU.  Returns info; pivots packed
// 4 bits per column in *piv.
constexpr int kCache = 5;

// R = double: E itself.  R = float: E rounded to float after the exact
// power-of-two equilibration E_ij 2^(ks_j - ks_i) (ks: mixed_scales).
template <typename R>
__device__ __forceinline__ int lu_left_t(const double* __restrict__ J, R* __restrict__ LU, int st,
                                         double fac, uint64_t& piv, const int* ks = nullptr) {
    int info = 0;
    uint32_t sing = 0;
    piv = 0;
    R mc[kCache][kNeqs];
#pragma unroll
    for (int j = 0; j < kNeqs; ++j) {
        R col[kNeqs];
#pragma unroll
        for (int i = 0; i < kNeqs; ++i) {
            const double v = -WR_AT(J, i * kNeqs + j, st);
            const double eij = (i == j) ? v + fac : v;
            col[i] = ks ? static_cast<R>(ldexp(eij, ks[j] - ks[i])) : static_cast<R>(eij);
        }
        // updates from columns 0..j-1 in order: the first kCache from registers,
        // the rest from LU with the multipliers of 4 columns loaded in one batch
        // (the scheduling barrier keeps the loads together, ahead of their uses)
#pragma unroll
        for (int k = 0; k < kCache; ++k) {
            if (k < j && !((sing >> k) & 1u)) {
                swap_rows(col, k, piv_get(piv, k));
                const R t = col[k];
#pragma unroll
                for (int i = 0; i < kNeqs; ++i)
                    if (i > k) col[i] += t * mc[k][i];
            }
        }
#pragma unroll
        for (int kb = kCache; kb < kNeqs - 1; kb += 4) {
            if (kb < j) {
                R mm[4][kNeqs];
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int i = 0; i < kNeqs; ++i)
                        if (kb + q < j && i > kb + q) mm[q][i] = WR_AT(LU, i * kNeqs + kb + q, st);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int k = kb + q;
                    if (k < j && !((sing >> k) & 1u)) {
                        swap_rows(col, k, piv_get(piv, k));
                        const R t = col[k];
#pragma unroll
                        for (int i = 0; i < kNeqs; ++i)
                            if (i > k) col[i] += t * mm[q][i];
                    }
                }
            }
        }
        if (j < kNeqs - 1) {
            int p = j;
            R best = fabs(col[j]);
#pragma unroll
            for (int i = 0; i < kNeqs; ++i) {
                if (i <= j) continue;
                const R v = fabs(col[i]);
                if (v > best) {
                    best = v;
                    p = i;
                }
            }
            piv |= (uint64_t)p << (4 * j);
            R pv = col[j];
#pragma unroll
            for (int i = 0; i < kNeqs; ++i)
                if (i > j) pv = (p == i) ? col[i] : pv;
            if (pv == R(0)) {
                info = j + 1;
                sing |= 1u << j;
            } else {
                swap_rows(col, j, p);
                const R mult = R(-1) / col[j];
#pragma unroll
                for (int i = 0; i < kNeqs; ++i)
                    if (i > j) col[i] *= mult;
            }
            if (j < kCache) {
#pragma unroll
                for (int i = 0; i < kNeqs; ++i) mc[j < kCache ? j : 0][i] = col[i];
            }
        } else {
            piv |= (uint64_t)(kNeqs - 1) << (4 * j);
            if (col[kNeqs - 1] == R(0)) info = kNeqs;
        }
#pragma unroll
        for (int i = 0; i < kNeqs; ++i) WR_AT(LU, i * kNeqs + j, st) = col[i];
    }
    return info;
}

__device__ __forceinline__ int lu_left(const ThreadWs& w, double fac, uint64_t& piv) {
    return lu_left_t<double>(w.J, w.LU, w.stride, fac, piv);
}

// Forward elimination over columns K0..K1-1: their multipliers are loaded
// in one batch (the scheduling barrier keeps the loads together, ahead of
// their uses), so the loads overlap instead of each update waiting for its
// own load.
template <typename R, int K0, int K1>
__device__ __forceinline__ void solve_fwd_cols(const R* __restrict__ LU, int st, uint64_t piv,
                                               R (&x)[kNeqs]) {
    R m[K1 - K0][kNeqs];
#pragma unroll
    for (int k = K0; k < K1; ++k)
#pragma unroll
        for (int j = 0; j < kNeqs; ++j)
            if (j > k) m[k - K0][j] = WR_AT(LU, j * kNeqs + k, st);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = K0; k < K1; ++k) {
        swap_rows(x, k, piv_get(piv, k));
        const R t = x[k];
#pragma unroll
        for (int j = 0; j < kNeqs; ++j)
            if (j > k) x[j] += t * m[k - K0][j];
    }
}

// Back substitution over columns K1-1 down to K0, their upper parts loaded
// in one batch.
template <typename R, int K0, int K1>
__device__ __forceinline__ void solve_bwd_cols(const R* __restrict__ LU, int st, R (&x)[kNeqs]) {
    R m[K1 - K0][kNeqs];
#pragma unroll
    for (int k = K0; k < K1; ++k)
#pragma unroll
        for (int j = 0; j < kNeqs; ++j)
            if (j <= k) m[k - K0][j] = WR_AT(LU, j * kNeqs + k, st);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = K1 - 1; k >= K0; --k) {
        x[k] /= m[k - K0][k];
        const R t = -x[k];
#pragma unroll
        for (int j = 0; j < kNeqs; ++j)
            if (j < k) x[j] += t * m[k - K0][j];
    }
}

// lu_solve (reproducer.cpp:154-181) from the workspace factors, x in registers:
// the same operations in the same order, the factors read in batches of
// 21-50 values (forward columns 0-3, 4-7, 8-13; backward 14-12, 11-8, 7-4, 3-0).
template <typename R>
__device__ __forceinline__ void lu_solve_t(const R* __restrict__ LU, int st, uint64_t piv,
                                           R (&x)[kNeqs]) {
    static_assert(kNeqs == 15, "batch boundaries are chosen for 15 unknowns");
    solve_fwd_cols<R, 0, 4>(LU, st, piv, x);
    solve_fwd_cols<R, 4, 8>(LU, st, piv, x);
    solve_fwd_cols<R, 8, 14>(LU, st, piv, x);
    solve_bwd_cols<R, 12, 15>(LU, st, x);
    solve_bwd_cols<R, 8, 12>(LU, st, x);
    solve_bwd_cols<R, 4, 8>(LU, st, x);
    solve_bwd_cols<R, 0, 4>(LU, st, x);
}

__device__ __forceinline__ void lu_solve_ws(const ThreadWs& w, uint64_t piv, double (&x)[kNeqs]) {
    lu_solve_t<double>(w.LU, w.stride, piv, x);
}

// ---- mixed precision: FP32 factors, FP64 residual, iterative refinement ------------
// The unknowns are equilibrated by powers of two taken from the controller's
// weights, s_i = 2^ks_i with ks_i = ilogb(atol_i + rtol_i |y_i|) (exact scaling:
// the species are number densities of 1e-40..1, the energy ~1e10): the FP32
// factors are those of D^-1 E D, D = diag(s), each element built in FP64
// exactly as the FP64 path builds E, scaled, then rounded.  A stage solve
// E x = b takes x = D E32^-1 D^-1 b and then `sweeps` refinement sweeps
// x += D E32^-1 D^-1 (b - E x), the residual in FP64 from J:
// r_i = b_i - sum_j E_ij x_j, j ascending.  The solve has converged when the
// last correction is at most 1e-6 s-weights (atol_i + rtol_i |y_i|) in every
// component; a non-finite correction never converges.
constexpr double kRefineTol = 1.0e-6;

__device__ __forceinline__ void mixed_scales(const ThreadWs& w, int (&ks)[kNeqs]) {
#pragma unroll
    for (int i = 0; i < kNeqs; ++i)
        ks[i] = ilogb(atol_of(i) + rtol_of(i) * fabs(WR_AT(w.vec, kVecY + i, w.stride)));
}

__device__ __forceinline__ void residual_e(const double* __restrict__ J, int st, double fac,
                                           const double (&b)[kNeqs], const double (&x)[kNeqs],
                                           double (&r)[kNeqs]) {
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < kNeqs; ++j) {
            const double v = -WR_AT(J, i * kNeqs + j, st);
            const double eij = (i == j) ? v + fac : v;
            acc += eij * x[j];
        }
        r[i] = b[i] - acc;
    }
}

// x (in: b, out: the refined solution); returns whether the refinement converged.
__device__ __forceinline__ bool solve_mixed(const ThreadWs& w, const float* __restrict__ LU32,
                                            uint64_t piv32, const int (&ks)[kNeqs], double fac,
                                            int sweeps, double (&x)[kNeqs]) {
    const int st = w.stride;
    double b[kNeqs];
    float xf[kNeqs];
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) {
        b[i] = x[i];
        xf[i] = static_cast<float>(ldexp(x[i], -ks[i]));
    }
    lu_solve_t<float>(LU32, st, piv32, xf);
#pragma unroll
    for (int i = 0; i < kNeqs; ++i) x[i] = ldexp(static_cast<double>(xf[i]), ks[i]);
    bool conv = true;
    for (int sw = 0; sw < sweeps; ++sw) {
        double r[kNeqs];
        residual_e(w.J, st, fac, b, x, r);
#pragma unroll
        for (int i = 0; i < kNeqs; ++i) xf[i] = static_cast<float>(ldexp(r[i], -ks[i]));
        lu_solve_t<float>(LU32, st, piv32, xf);
        conv = true;
#pragma unroll
        for (int i = 0; i < kNeqs; ++i) {
            const double d = ldexp(static_cast<double>(xf[i]), ks[i]);
            x[i] += d;
            const double sk = atol_of(i) + rtol_of(i) * fabs(WR_AT(w.vec, kVecY + i, st));
            conv = conv && (fabs(d) <= kRefineTol * sk);
        }
    }
    return conv;
}

}  // namespace wr
