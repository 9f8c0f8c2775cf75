// CPU emulation of one cooperative wave advancing one cell: the hand-written
// part of warp-reactor_amd/csrc/wr_coop.hpp (coop_state_in, coop_lu,
// coop_lu_row, coop_solve, coop_sumsq, coop_cell) restated statement by
// statement with every per-lane statement a loop over the 64 lanes and
// v_readlane / ballot as array reads, on the generated cooperative programs
// (gen/coop.py) built as C.  Test infrastructure: tests/test_coop.py runs it
// on the cells of the 64^3 run that take the most attempts and compares the
// advanced cells with the oracle's advance, bit for bit.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "wr_math.h"
#include "wr_eos.hpp"

#define CHEM_EXP(x) wr_exp(x)
#define CHEM_LOG(x) wr_log(x)

static char coop_smem[1 << 18];

static inline double coop_flip(double x, uint32_t w) {
    return wr_from_bits(wr_bits_of(x) ^ ((uint64_t)(w & 0x80000000u) << 32));
}
static inline double coop_sgn(double x, uint32_t w) {
    uint64_t b = wr_bits_of(x);
    if (w & 0x40000000u) b &= ~(1ull << 63);
    return wr_from_bits(b ^ ((uint64_t)(w & 0x80000000u) << 32));
}

#define COOP_FN static
#define COOP_PARAMS void
#define COOP_DATA static const
#define COOP_DU(x) static uint32_t x[64]
#define COOP_DD(x) static double x[64]
#define COOP_V(x) x[ln]
#define COOP_L(...)                   \
    for (int ln = 0; ln < 64; ++ln) { \
        __VA_ARGS__                   \
    }
#define COOP_TW(tab, row) (tab)[(row) * 64 + ln]
#define COOP_LD(a) (*(const double *)(coop_smem + (a)))
#define COOP_LDU(a) (*(const uint32_t *)(coop_smem + (a)))
#define COOP_ST(a, v) (*(double *)(coop_smem + (a)) = (v))
#define COOP_STU(a, v) (*(uint32_t *)(coop_smem + (a)) = (v))
#include "coop_network.h"

using namespace wr;

namespace {

constexpr int NL = 64;
// ROS2S coefficients and controller constants (reproducer.cpp:7096-7128),
// the values wr_device.hpp's Ros2sCoef holds
constexpr double gamma_ = 0.292893218813452, a21 = 2.0000000000000036, a31 = 6.828427124746214,
                 a32 = 3.4142135623731007, c21 = -6.828427124746214, c31 = -10.949747468305889,
                 c32 = -7.535533905932761, b1 = 6.828427124746214, b2 = 3.414213562373101,
                 b3 = 1.0, e1 = -0.23570226039551292, e2 = -0.23570226039551567,
                 e3 = -0.13807118745769906;
constexpr double kUround = 1.0e-16, kFacMin = 0.2, kFacMax = 6.0, kSafe = 0.9;
constexpr int VY = 0, VF0 = 15, VYNEW = 75, VDY = 90;

double &at(int off) { return *(double *)(coop_smem + off); }
double &vec(int v, int i) { return at(COOP_OFF_VEC + 8 * (v + i)); }
double atol_of(int i) { return i < kNumSpec ? 1.0e-4 : 1.0e-6; }
double rtol_of(int i) { return i < kNumSpec ? 1.0e-4 : 1.0e-6; }
int piv_get(uint64_t piv, int k) { return (int)((piv >> (4 * k)) & 15u); }

void state_in(int v) {
    double y[kNeqs], X[kNumSpec];
    for (int i = 0; i < kNeqs; ++i) y[i] = vec(v, i);
    const double T = state_from_y(y, X);
    for (int ln = 0; ln < kNeqs; ++ln) at(COOP_OFF_IN + 8 * ln) = ln == 0 ? T : X[ln - 1];
}

int lu(double fac, uint64_t *piv) {
    static double a[NL][kNeqs];
    int lam[NL];
    for (int ln = 0; ln < NL; ++ln) {
        const int r = ln < kNeqs ? ln : 0;
        for (int j = 0; j < kNeqs; ++j) {
            const double v = -at(COOP_OFF_J + 8 * (r * kNeqs + j));
            a[ln][j] = (r == j) ? v + fac : v;
        }
        lam[ln] = ln < kNeqs ? ln : 64;
    }
    int info = 0;
    *piv = 0;
    for (int k = 0; k < kNeqs - 1; ++k) {
        for (int ln = 0; ln < kNeqs; ++ln)
            if (lam[ln] >= k) at(COOP_OFF_COL + 8 * lam[ln]) = a[ln][k];
        double col[kNeqs];
        for (int i = k; i < kNeqs; ++i) col[i] = at(COOP_OFF_COL + 8 * i);
        int p = k;
        double best = fabs(col[k]);
        for (int i = k + 1; i < kNeqs; ++i) {
            const double v = fabs(col[i]);
            if (v > best) {
                best = v;
                p = i;
            }
        }
        double pv = col[k];
        for (int i = k + 1; i < kNeqs; ++i) pv = (p == i) ? col[i] : pv;
        *piv |= (uint64_t)p << (4 * k);
        if (pv == 0.0) {
            info = k + 1;
            for (int ln = 0; ln < kNeqs; ++ln)
                if (lam[ln] > k) at(COOP_OFF_LU + 8 * (lam[ln] * kNeqs + k)) = a[ln][k];
            continue;
        }
        if (p != k)
            for (int ln = 0; ln < NL; ++ln)
                lam[ln] = (lam[ln] == k) ? p : ((lam[ln] == p) ? k : lam[ln]);
        const double m = -1.0 / pv;
        int P = -1;
        for (int ln = 0; ln < kNeqs; ++ln)
            if (lam[ln] == k && P < 0) P = ln;
        for (int ln = 0; ln < kNeqs; ++ln)
            if (lam[ln] > k) {
                a[ln][k] *= m;
                at(COOP_OFF_LU + 8 * (lam[ln] * kNeqs + k)) = a[ln][k];
            }
        for (int j = k + 1; j < kNeqs; ++j) {
            const double t = a[P][j];
            for (int ln = 0; ln < kNeqs; ++ln)
                if (lam[ln] > k) a[ln][j] += t * a[ln][k];
        }
    }
    *piv |= (uint64_t)(kNeqs - 1) << (4 * (kNeqs - 1));
    for (int j = 0; j < kNeqs; ++j)
        for (int ln = 0; ln < kNeqs; ++ln)
            if (j >= lam[ln]) at(COOP_OFF_LU + 8 * (lam[ln] * kNeqs + j)) = a[ln][j];
    if (at(COOP_OFF_LU + 8 * (kNeqs * kNeqs - 1)) == 0.0) info = kNeqs;
    return info;
}

void lu_rows(double (*R)[kNeqs]) {
    for (int ln = 0; ln < NL; ++ln) {
        const int r = ln < kNeqs ? ln : 0;
        for (int j = 0; j < kNeqs; ++j) R[ln][j] = at(COOP_OFF_LU + 8 * (r * kNeqs + j));
    }
}

void solve(double *x, double (*R)[kNeqs], uint64_t piv) {
    for (int k = 0; k < kNeqs - 1; ++k) {
        const int p = piv_get(piv, k);
        const double xk = x[k], xp = x[p];
        for (int ln = 0; ln < NL; ++ln) {
            if (ln == p) x[ln] = xk;
            if (ln == k) x[ln] = xp;
            if (ln > k && ln < kNeqs) x[ln] += xp * R[ln][k];
        }
    }
    for (int kb = 0; kb < kNeqs; ++kb) {
        const int k = kNeqs - 1 - kb;
        const double q = x[k] / R[k][k];
        for (int ln = 0; ln < NL; ++ln) {
            if (ln == k) x[ln] = q;
            const double t = -q;
            if (ln < k) x[ln] += t * R[ln][k];
        }
    }
}

double sumsq(const double *v) {
    for (int ln = 0; ln < kNeqs; ++ln) at(COOP_OFF_COL + 8 * ln) = v[ln];
    double s = 0.0;
    for (int i = 0; i < kNeqs; ++i) {
        const double t = at(COOP_OFF_COL + 8 * i);
        s += t * t;
    }
    return s;
}

struct Cell {  // CollapseState (reproducer.cpp:7455-7461)
    double rho, T, e, xn[14], time, driver;
    int32_t completed, reserved;
    uint64_t stats[7];
};
static_assert(sizeof(Cell) == 216, "CollapseState layout");

}  // namespace

extern "C" {

// coop_cell of wr_coop.hpp on one cell; returns 1 if it integrated, 0 if
// not, or the failure code.  max_steps <= 0: 1e7.
int coop_emul_advance(void *cell_v, int completed, double next_time, double dt_grid,
                      int max_steps) {
    Cell &s = *(Cell *)cell_v;
    if (max_steps <= 0) max_steps = 10000000;
    memcpy(coop_smem + COOP_OFF_CONST, coop_const_bits, sizeof(coop_const_bits));
    if (s.completed != completed) return 0;
    double xn[kNumSpec];
    for (int n = 0; n < kNumSpec; ++n) xn[n] = s.xn[n];
    const double local_dt = collapse_timestep(xn);
    if (local_dt < 10.0) return 0;
    const double old_density = s.driver;
    const double tff = local_dt / 1.0e-1;
    const double driver = old_density + dt_grid * (old_density / tff);
    s.driver = driver;
    if (!valid_positive(driver) || driver > 2.0e-6) return 0;
    const double ratio = driver / old_density;
    for (int n = 0; n < kNumSpec; ++n) xn[n] *= ratio;
    for (int n = 0; n < kNumSpec; ++n) s.xn[n] = xn[n];
    const double rho = s.rho * ratio;
    const double e0 = eos_energy(xn, s.T);
    s.rho = rho;
    double y[NL];
    for (int ln = 0; ln < NL; ++ln) y[ln] = ln < kNumSpec ? xn[ln] : e0;
    for (int ln = 0; ln < kNeqs; ++ln) vec(VY, ln) = y[ln];

    const double tout = dt_grid;
    double h = tout;
    if (fabs(h) <= 10.0 * kUround) h = 1.0e-6;
    h = rmin(fabs(h), tout - 0.0);
    double x = 0.0, hopt = h, hacc = 0.0, erracc = 1.0;
    int nsing = 0, n_step = 0, n_accept = 0;
    bool reject = false, last = false;
    int cnt[7] = {0, 0, 0, 0, 0, 0, 0};  // step, rhs, jac, decomp, solve, accept, reject
    int result = 0;
    static double R[NL][kNeqs];
    double f0[NL], ak1[NL], ak2[NL], ak3[NL], tmp[NL], term[NL];
    (void)hopt;
    for (;;) {
        if (n_step > max_steps) {
            result = -4;
            break;
        }
        if (0.1 * fabs(h) <= fabs(x) * kUround) {
            result = -2;
            break;
        }
        if (last) {
            result = 1;
            break;
        }
        hopt = h;
        if (x + h * 1.0001 >= tout) {
            h = tout - x;
            last = true;
        }
        state_in(VY);
        coop_fused();
        cnt[2] += 1;
        for (int ln = 0; ln < NL; ++ln) f0[ln] = vec(VF0, ln < kNeqs ? ln : 0);
        for (;;) {
            const double fac = 1.0 / (h * gamma_);
            uint64_t piv;
            const int info = lu(fac, &piv);
            if (info != 0) {
                nsing += 1;
                if (nsing >= 5) {
                    result = -7;
                    break;
                }
                h *= 0.5;
                reject = true;
                last = false;
                continue;
            }
            cnt[3] += 1;
            lu_rows(R);
            for (int ln = 0; ln < NL; ++ln) ak1[ln] = f0[ln];
            solve(ak1, R, piv);
            cnt[4] += 1;
            for (int ln = 0; ln < kNeqs; ++ln) vec(VYNEW, ln) = y[ln] + a21 * ak1[ln];
            for (int ln = 0; ln < NL; ++ln) tmp[ln] = (c21 / h) * ak1[ln];
            state_in(VYNEW);
            coop_rhs();
            for (int ln = 0; ln < NL; ++ln) ak2[ln] = tmp[ln] + vec(VDY, ln < kNeqs ? ln : 0);
            solve(ak2, R, piv);
            cnt[4] += 1;
            for (int ln = 0; ln < kNeqs; ++ln) vec(VYNEW, ln) = y[ln] + a31 * ak1[ln] + a32 * ak2[ln];
            for (int ln = 0; ln < NL; ++ln) tmp[ln] = (c31 * ak1[ln] + c32 * ak2[ln]) / h;
            state_in(VYNEW);
            coop_rhs();
            for (int ln = 0; ln < NL; ++ln) ak3[ln] = tmp[ln] + vec(VDY, ln < kNeqs ? ln : 0);
            solve(ak3, R, piv);
            cnt[4] += 1;
            double yn[NL];
            for (int ln = 0; ln < NL; ++ln) {
                const int st = ln < kNeqs ? ln : 0;
                yn[ln] = y[ln] + b1 * ak1[ln] + b2 * ak2[ln] + b3 * ak3[ln];
                const double wv = e1 * ak1[ln] + e2 * ak2[ln] + e3 * ak3[ln];
                const double sk = atol_of(st) + rtol_of(st) * rmax(fabs(y[ln]), fabs(yn[ln]));
                term[ln] = wv / sk;
            }
            double err = sumsq(term);
            cnt[1] += 3;
            n_step += 1;
            cnt[0] = n_step;
            err = sqrt(err / 15.0);
            const double raw_fac = wr_cbrt(err) / kSafe;
            const double fac_step = rmax(1.0 / kFacMax, rmin(1.0 / kFacMin, raw_fac));
            double hnew = h / fac_step;
            if (err <= 1.0) {
                n_accept += 1;
                cnt[5] = n_accept;
                if (n_accept > 1) {
                    const double facgus = rmax(
                        1.0 / kFacMax,
                        rmin(1.0 / kFacMin, (hacc / h) * wr_cbrt((err * err) / erracc) / kSafe));
                    hnew = h / rmax(fac_step, facgus);
                }
                hacc = h;
                erracc = rmax(1.0e-2, err);
                for (int ln = 0; ln < NL; ++ln) y[ln] = yn[ln];
                for (int ln = 0; ln < kNeqs; ++ln) vec(VY, ln) = y[ln];
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
            if (n_accept >= 1) cnt[6] += 1;
        }
        if (result != 0) break;
    }
    for (int k = 0; k < 7; ++k) s.stats[k] += (uint64_t)(cnt[k] > 0 ? cnt[k] : 0);
    if (result == 1) {
        double xo[kNumSpec];
        for (int n = 0; n < kNumSpec; ++n) xo[n] = vec(VY, n);
        const double e = vec(VY, kNumSpec);
        const double T = post_step(xo, rho, e);
        for (int n = 0; n < kNumSpec; ++n) s.xn[n] = xo[n];
        s.e = e;
        s.T = T;
        s.time = next_time;
        s.completed = completed + 1;
        return 1;
    }
    s.e = e0;
    return result;
}

}  // extern "C"
