/*
 * ORACLE -- test infrastructure, not product code.
 *
 * Plain-C CPU restatement of the reference's primordial-chemistry ROS2S grid
 * advance.  It is the checker for the HIP path: only tests/, the smoke check
 * and bench.py's CPU-baseline leg load it.  The product never links it.
 *
 * Every function states the reference code it follows (reproducer.cpp:LINE).
 * The network right-hand side and Jacobian are the generated straight-line
 * code emitted from the committed DAG IR (_build/chem_network.h).
 *
 * Pinning (tests/test_oracle.py): each function below is compared bit for bit
 * with the reference's own compiled CPU code (oracle/_ref/libwr_ref.so, built
 * from reproducer.cpp by oracle/Makefile) and with committed golden vectors
 * written from it (tests/golden/): RHS/J at several hundred states spanning
 * every temperature branch, LU on the reference's LU test matrices and on real
 * E matrices, single ROS2S integrations with all seven counters, prepare and
 * advance semantics, and the grid-1 known answer (438 global steps).  Closed
 * forms pin the rest: SplitMix64 against its published test vector, the
 * free-fall time formula, mass/charge conservation of the RHS, LU against
 * brute-force permutation checks.
 *
 * or_run_cta restates the Mojo front end's gridwide-cta32 policy
 * (reproducer.mojo), for which no toolchain is here: it is pinned by what
 * the policy's definition fixes (tests/test_cta.py: one-cell tiles equal the
 * pinned gridwide run, identical cells share one path, tile invariants, the
 * reference comparator), except the operation order of its two
 * density-driver updates, (dt * driver) / tff, which follows the Mojo source
 * by reading: "parity unpinned" for that order.  Every other function is
 * pinned as above.
 *
 * Build: oracle/Makefile, gcc -O2 -ffp-contract=off (no FMA contraction; x86-64
 * SSE2 arithmetic is IEEE double).
 */
#include <math.h>
#include <float.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NSPEC 14
#define NEQ 15

typedef struct {
    double rho, T, e;
    double xn[NSPEC];
} or_burn_t;

typedef struct {
    uint64_t internal_steps, rhs_calls, jacobian_calls, decompositions,
        linear_solves, accepted_steps, rejected_steps;
} or_stats;

/* Same 216-byte layout as CollapseState (reproducer.cpp:7455-7461). */
typedef struct {
    or_burn_t current;
    double time;
    double density_driver;
    int32_t completed_steps;
    int32_t pad_;
    or_stats stats;
} or_cell;

_Static_assert(sizeof(or_cell) == 216, "or_cell must match CollapseState");

enum { OR_SUCCESS = 1, OR_BAD_INPUTS = -1, OR_DT_UNDERFLOW = -2, OR_TOO_MANY_STEPS = -4,
       OR_TOO_MUCH_ACCURACY = -5, OR_LU_ERROR = -7 };

/* ---- generated network ------------------------------------------------- */
/* libm: the host's (glibc, as the reference build) by default; with
 * -DOR_WRMATH the deterministic exp/log/cbrt the GPU path uses
 * (warp-reactor_amd/csrc/wr_math.h), so that this build and the GPU compute
 * the same bits and differ from the default build only through libm. */
#ifdef OR_WRMATH
#include "../warp-reactor_amd/csrc/wr_math.h"
#define CHEM_EXP(x) wr_exp(x)
#define CHEM_LOG(x) wr_log(x)
#define OR_CBRT(x) wr_cbrt(x)
#else
#define CHEM_EXP(x) exp(x)
#define CHEM_LOG(x) log(x)
#define OR_CBRT(x) cbrt(x)
#endif
#define CHEM_FN static inline
#define CHEM_SINK_PARAMS , double *jsink
#define CHEM_JAC_STORE(r, c, v) (jsink[(r) * NEQ + (c)] = (v))
/* staged (register-budgeted) variants: a local staging array, outputs to jac/f */
#define CHEM_STAGE_PARAMS , double *stg, double *jout, double *fout
#define STAGE_STORE(s, v) (stg[(s)] = (v))
#define STAGE_LOAD(s) (stg[(s)])
#define CHEM_OUT(kind, ...) CHEM_OUT_##kind(__VA_ARGS__)
#define CHEM_OUT_J(r, c, v) (jout[(r) * NEQ + (c)] = (v))
#define CHEM_OUT_F(i, v) (fout[(i)] = (v))
#include "_build/chem_network.h"
#ifdef OR_RATE_TABLE
#include "_build/chem_table.h"
#include "rate_table.c"
#endif

/* ---- network constants and EOS (reproducer.cpp:199-339) ----------------- */
static const double k_mass[NSPEC] = {
    9.10938188e-28, 1.67262158e-24, 1.67353251819e-24, 1.67444345638e-24,
    3.34512158e-24, 3.34603251819e-24, 3.34615409819e-24, 3.34694345638e-24,
    3.34706503638e-24, 5.01865409819e-24, 5.01956503638e-24, 6.69024316e-24,
    6.69115409819e-24, 6.69206503638e-24};
static const double k_nA = 6.02214076e+23, k_kB = 1.3806490000000002e-16,
                    k_mp = 1.67262192595e-24;
static const double k_floor = 1.0e-100;
static const double k_G = 6.674e-8;
static const double k_pi = 3.141592653589793238462643383279502884;

static double gamma_of(int n) { return (n == 6 || n == 8 || n == 9 || n == 10) ? 1.4 : 5.0 / 3.0; }

/* density, reproducer.cpp:270-276 */
static double density_of(const double *xn) {
    double rho = 0.0;
    for (int n = 0; n < NSPEC; ++n) rho += xn[n] * k_mass[n];
    return rho;
}

/* eos_sums_from_number_densities, reproducer.cpp:284-304: returns the
 * denominator product used by eos_re/eos_rt pieces. */
static void eos_sums(const double *xn, double *abarinv, double *gammasinv) {
    double rhotot = 0.0, sa = 0.0, sg = 0.0;
    for (int n = 0; n < NSPEC; ++n) rhotot += xn[n] * k_mass[n];
    for (int n = 0; n < NSPEC; ++n) {
        sa += xn[n];
        sg += (xn[n] * k_mp / rhotot) * (1.0 / (gamma_of(n) - 1.0));
    }
    sa *= k_mp / rhotot;
    sg /= sa;
    *abarinv = sa;
    *gammasinv = sg;
}

/* eos_rt, reproducer.cpp:306-309 */
static void eos_rt(or_burn_t *s) {
    double sa, sg;
    eos_sums(s->xn, &sa, &sg);
    s->e = sg * sa * (k_nA * k_kB) * s->T;
}

/* eos_re, reproducer.cpp:311-314 */
static void eos_re(or_burn_t *s) {
    double sa, sg;
    eos_sums(s->xn, &sa, &sg);
    s->T = s->e / (sg * (k_nA * k_kB) * sa);
}

/* balance_charge, reproducer.cpp:316-319 */
static void balance_charge(or_burn_t *s) {
    const double *x = s->xn;
    s->xn[0] = -x[3] - x[7] + x[1] + x[12] + x[6] + x[4] + x[9] + 2.0 * x[11];
}

/* normalize_number_densities_to_density, reproducer.cpp:321-332 */
static void normalize(or_burn_t *s) {
    double mf[NSPEC], sum = 0.0;
    for (int n = 0; n < NSPEC; ++n) {
        mf[n] = k_mass[n] * s->xn[n] / s->rho;
        sum += mf[n];
    }
    for (int n = 0; n < NSPEC; ++n) {
        mf[n] /= sum;
        s->xn[n] = mf[n] * s->rho / k_mass[n];
    }
}

/* std::max(a, b) as the reference evaluates it: (a < b) ? b : a */
static double ref_max(double a, double b) { return (a < b) ? b : a; }
static double ref_min(double a, double b) { return (b < a) ? b : a; }

/* floor_and_normalize_number_densities, reproducer.cpp:334-339 */
static void floor_and_normalize(or_burn_t *s) {
    for (int n = 0; n < NSPEC; ++n) s->xn[n] = ref_max(s->xn[n], k_floor);
    normalize(s);
}

/* burn_state_from_y, reproducer.cpp:7029-7038: X = max(y, floor), T from e. */
static double temperature_from_y(const double *y, double *X) {
    or_burn_t b;
    for (int n = 0; n < NSPEC; ++n) {
        b.xn[n] = ref_max(y[n], k_floor);
        X[n] = b.xn[n];
    }
    b.e = y[NSPEC];
    eos_re(&b);
    return b.T;
}

/* PrimordialChem::rhs, reproducer.cpp:7040-7050 */
void or_rhs(const double *y, double *dydt) {
    double X[NSPEC];
    const double T = temperature_from_y(y, X);
#ifdef OR_RATE_TABLE
    if (g_rt_mode) {
        double tv[CHEM_NTV];
        rt_values(T, tv);
        chem_rhs_tv(tv, T, X, dydt);
        return;
    }
#endif
    chem_rhs(T, X, dydt);
}

/* PrimordialChem::jacobian, reproducer.cpp:7052-7060 (row-major 15x15) */
void or_jac(const double *y, double *jac) {
    double X[NSPEC], f[NEQ];
    const double T = temperature_from_y(y, X);
#ifdef OR_RATE_TABLE
    if (g_rt_mode) {
        double tv[CHEM_NTV];
        rt_values(T, tv);
        chem_fused_tv(tv, T, X, f, jac);
        return;
    }
#endif
    chem_fused(T, X, f, jac);
}

/* J and f(y) at one state, as the fused evaluation the GPU path uses. */
void or_jac_rhs(const double *y, double *jac, double *dydt) {
    double X[NSPEC];
    const double T = temperature_from_y(y, X);
#ifdef OR_RATE_TABLE
    if (g_rt_mode) {
        double tv[CHEM_NTV];
        rt_values(T, tv);
        chem_fused_tv(tv, T, X, dydt, jac);
        return;
    }
#endif
    chem_fused(T, X, dydt, jac);
}

double or_temperature_from_y(const double *y) {
    double X[NSPEC];
    return temperature_from_y(y, X);
}

/* ---- dense LU, LINPACK conventions (reproducer.cpp:96-181) ------------- */

/* lu_decomposition<15>: first strict maximum of |a_ik| is the pivot, rows are
 * exchanged from column k on (earlier multipliers stay in place), multipliers
 * are stored negated; info = k+1 for a zero pivot, N if the last is zero. */
int or_lu(double *a, int *ipvt) {
    int info = 0;
    for (int k = 0; k < NEQ - 1; ++k) {
        int p = k;
        double best = fabs(a[k * NEQ + k]);
        for (int i = k + 1; i < NEQ; ++i) {
            const double v = fabs(a[i * NEQ + k]);
            if (v > best) {
                best = v;
                p = i;
            }
        }
        ipvt[k] = p;
        if (a[p * NEQ + k] == 0.0) {
            info = k + 1;
            continue;
        }
        if (p != k) {
            for (int j = k; j < NEQ; ++j) {
                const double t = a[p * NEQ + j];
                a[p * NEQ + j] = a[k * NEQ + j];
                a[k * NEQ + j] = t;
            }
        }
        const double m = -1.0 / a[k * NEQ + k];
        for (int i = k + 1; i < NEQ; ++i) a[i * NEQ + k] *= m;
        for (int j = k + 1; j < NEQ; ++j) {
            const double t = a[k * NEQ + j];
            for (int i = k + 1; i < NEQ; ++i) a[i * NEQ + j] += t * a[i * NEQ + k];
        }
    }
    ipvt[NEQ - 1] = NEQ - 1;
    if (a[(NEQ - 1) * NEQ + NEQ - 1] == 0.0) info = NEQ;
    return info;
}

/* lu_solve<15>: interchanges and L applied in elimination order, then U. */
void or_lu_solve(const double *lu, const int *ipvt, double *x) {
    for (int k = 0; k < NEQ - 1; ++k) {
        const int p = ipvt[k];
        const double t = x[p];
        if (p != k) {
            x[p] = x[k];
            x[k] = t;
        }
        for (int j = k + 1; j < NEQ; ++j) x[j] += t * lu[j * NEQ + k];
    }
    for (int k = NEQ - 1; k >= 0; --k) {
        x[k] /= lu[k * NEQ + k];
        const double t = -x[k];
        for (int j = 0; j < k; ++j) x[j] += t * lu[j * NEQ + k];
    }
}

/* ---- FP32 factorisation with FP64 refinement (no reference counterpart) --
 * The same LINPACK elimination in single precision on E rounded to float. */
int or_lu32(float *a, int *ipvt) {
    int info = 0;
    for (int k = 0; k < NEQ - 1; ++k) {
        int p = k;
        float best = fabsf(a[k * NEQ + k]);
        for (int i = k + 1; i < NEQ; ++i) {
            const float v = fabsf(a[i * NEQ + k]);
            if (v > best) {
                best = v;
                p = i;
            }
        }
        ipvt[k] = p;
        if (a[p * NEQ + k] == 0.0f) {
            info = k + 1;
            continue;
        }
        if (p != k) {
            for (int j = k; j < NEQ; ++j) {
                const float t = a[p * NEQ + j];
                a[p * NEQ + j] = a[k * NEQ + j];
                a[k * NEQ + j] = t;
            }
        }
        const float m = -1.0f / a[k * NEQ + k];
        for (int i = k + 1; i < NEQ; ++i) a[i * NEQ + k] *= m;
        for (int j = k + 1; j < NEQ; ++j) {
            const float t = a[k * NEQ + j];
            for (int i = k + 1; i < NEQ; ++i) a[i * NEQ + j] += t * a[i * NEQ + k];
        }
    }
    ipvt[NEQ - 1] = NEQ - 1;
    if (a[(NEQ - 1) * NEQ + NEQ - 1] == 0.0f) info = NEQ;
    return info;
}

void or_lu_solve32(const float *lu, const int *ipvt, float *x) {
    for (int k = 0; k < NEQ - 1; ++k) {
        const int p = ipvt[k];
        const float t = x[p];
        if (p != k) {
            x[p] = x[k];
            x[k] = t;
        }
        for (int j = k + 1; j < NEQ; ++j) x[j] += t * lu[j * NEQ + k];
    }
    for (int k = NEQ - 1; k >= 0; --k) {
        x[k] /= lu[k * NEQ + k];
        const float t = -x[k];
        for (int j = 0; j < k; ++j) x[j] += t * lu[j * NEQ + k];
    }
}

static int g_lu_mode = 0, g_lu_sweeps = 2;
static uint64_t g_lu_fallbacks = 0, g_refine_fallbacks = 0;

/* mode 0: FP64 factors (the reference); mode 1: FP32 factors, `sweeps`
 * refinement sweeps with FP64 residuals per stage solve; the unknowns are
 * equilibrated by the powers of two 2^ilogb(atol_i + rtol_i |y_i|). */
void or_set_lu_precision(int mode, int sweeps) {
    g_lu_mode = mode;
    g_lu_sweeps = sweeps;
}
void or_mixed_fallbacks(uint64_t *lu_fb, uint64_t *refine_fb) {
    *lu_fb = g_lu_fallbacks;
    *refine_fb = g_refine_fallbacks;
}
void or_reset_mixed_fallbacks(void) {
    g_lu_fallbacks = 0;
    g_refine_fallbacks = 0;
}

/* Jacobian reuse (a study with no counterpart in the reference, which
 * evaluates J at the head of every outer iteration of integrate and reuses it
 * only on rejection, reproducer.cpp:7236-7260): J is evaluated at the first
 * attempt of a burn and again once `age_max` steps have been accepted with it;
 * a rejection of an attempt that used a J from an earlier state evaluates J
 * at the current state before the retry.  age_max = 1 is the reference. */
static int g_jac_age_max = 1;
void or_set_jacobian_reuse(int age_max) { g_jac_age_max = age_max > 1 ? age_max : 1; }

#define OR_REFINE_TOL 1.0e-6

/* Stage solve of the FP32-factorisation form: x (in: b) from the FP32
 * factors, then sweeps of x += E32^-1 (b - E x) with E = fac I - J rebuilt in
 * FP64 (E_ij = -J_ij, plus fac on the diagonal) and the row sums taken in
 * ascending column order; FP32 quantities are scaled by 2^-ks_i (unknown i)
 * before rounding.  Converged when the last correction is at most
 * 1e-6 (atol_i + rtol_i |y_i|) in every component. */
static int solve_mixed(const float *lu32, const int *ip32, const double *fjac, double fac,
                       const double *y, int sweeps, double *x);

/* ---- ROS2S (reproducer.cpp:7066-7347) ---------------------------------- */
static const double c_gamma = 0.292893218813452, c_ct2 = 0.585786437626905,
                    c_a21 = 2.0000000000000036, c_a31 = 6.828427124746214,
                    c_a32 = 3.4142135623731007, c_c21 = -6.828427124746214,
                    c_c31 = -10.949747468305889, c_c32 = -7.535533905932761,
                    c_b1 = 6.828427124746214, c_b2 = 3.414213562373101, c_b3 = 1.0,
                    c_e1 = -0.23570226039551292, c_e2 = -0.23570226039551567,
                    c_e3 = -0.13807118745769906;
static const double k_uround = 1.0e-16, k_fac_min = 0.2, k_fac_max = 6.0, k_safe = 0.9;

/* Tolerances of configure_ros2s (reproducer.cpp:7409-7412, 7582-7590). */
static double rtol_of(int i) { return i < NSPEC ? 1.0e-4 : 1.0e-6; }
static double atol_of(int i) { return i < NSPEC ? 1.0e-4 : 1.0e-6; }

typedef struct {
    int n_step, n_rhs, n_jac, n_decomp, n_solve, n_accept, n_reject;
} or_counters;

static void mixed_scales(const double *y, int *ks) {
    for (int i = 0; i < NEQ; ++i) ks[i] = ilogb(atol_of(i) + rtol_of(i) * fabs(y[i]));
}

static int solve_mixed(const float *lu32, const int *ip32, const double *fjac, double fac,
                       const double *y, int sweeps, double *x) {
    double b[NEQ];
    float xf[NEQ];
    int ks[NEQ];
    mixed_scales(y, ks);
    for (int i = 0; i < NEQ; ++i) {
        b[i] = x[i];
        xf[i] = (float)ldexp(x[i], -ks[i]);
    }
    or_lu_solve32(lu32, ip32, xf);
    for (int i = 0; i < NEQ; ++i) x[i] = ldexp((double)xf[i], ks[i]);
    int conv = 1;
    for (int sw = 0; sw < sweeps; ++sw) {
        for (int i = 0; i < NEQ; ++i) {
            double acc = 0.0;
            for (int j = 0; j < NEQ; ++j) {
                const double v = -fjac[i * NEQ + j];
                const double eij = (i == j) ? v + fac : v;
                acc += eij * x[j];
            }
            xf[i] = (float)ldexp(b[i] - acc, -ks[i]);
        }
        or_lu_solve32(lu32, ip32, xf);
        conv = 1;
        for (int i = 0; i < NEQ; ++i) {
            const double d = ldexp((double)xf[i], ks[i]);
            x[i] += d;
            const double sk = atol_of(i) + rtol_of(i) * fabs(y[i]);
            conv = conv && (fabs(d) <= OR_REFINE_TOL * sk);
        }
    }
    return conv;
}

/* One stage solve of the FP32-factorisation form on its own (tests): E = fac I
 * - J equilibrated and factorised in FP32, then x (in: b) solved with `sweeps`
 * refinement sweeps.  Returns -1 if the FP32 factors are singular, else
 * whether the refinement converged. */
/* The FP32 factors of the equilibrated E = fac I - J (tests). */
int or_mixed_lu(const double *fjac, double fac, const double *y, float *e32, int *ip32) {
    int ks[NEQ];
    mixed_scales(y, ks);
    for (int i = 0; i < NEQ; ++i)
        for (int j = 0; j < NEQ; ++j) {
            const double v = -fjac[i * NEQ + j];
            e32[i * NEQ + j] = (float)ldexp((i == j) ? v + fac : v, ks[j] - ks[i]);
        }
    return or_lu32(e32, ip32);
}

int or_solve_mixed(const double *fjac, double fac, const double *y, int sweeps, double *x) {
    float e32[NEQ * NEQ];
    int ip32[NEQ], ks[NEQ];
    mixed_scales(y, ks);
    for (int i = 0; i < NEQ; ++i)
        for (int j = 0; j < NEQ; ++j) {
            const double v = -fjac[i * NEQ + j];
            e32[i * NEQ + j] = (float)ldexp((i == j) ? v + fac : v, ks[j] - ks[i]);
        }
    if (or_lu32(e32, ip32) != 0) return -1;
    return solve_mixed(e32, ip32, fjac, fac, y, sweeps, x);
}

/* Hooks for the adjoint oracle (adjoint.c, built into the OR_ADJOINT
 * library): with rate multipliers set, the RHS and Jacobian come from the
 * generated f_theta code (identical bits when every multiplier is 1), and the
 * step sizes of accepted attempts are recorded.  Per thread. */
typedef void (*or_fj_fn)(const double *y, const double *M, double *f, double *J);
typedef void (*or_f_fn)(const double *y, const double *M, double *f);
static _Thread_local const double *t_M = NULL;
static _Thread_local or_fj_fn t_fj = NULL;
static _Thread_local or_f_fn t_f = NULL;
static _Thread_local double *t_hrec = NULL;
static _Thread_local int t_nh = 0, t_hcap = 0;

void or_set_theta_hooks(const double *M, or_fj_fn fj, or_f_fn f) {
    t_M = M;
    t_fj = fj;
    t_f = f;
}
void or_set_h_recorder(double *hrec, int hcap) {
    t_hrec = hrec;
    t_hcap = hcap;
    t_nh = 0;
}
int or_h_recorded(void) { return t_nh; }

static void hook_jac(const double *y, double *jac) {
    if (t_M) {
        double f[NEQ];
        t_fj(y, t_M, f, jac);
    } else {
        or_jac(y, jac);
    }
}
static void hook_rhs(const double *y, double *f) {
    if (t_M)
        t_f(y, t_M, f);
    else
        or_rhs(y, f);
}

/* RODAS<PrimordialChem>::integrate from t=0 to tout; y updated in place.
 * The checks on safe/fac/tolerances (7208-7218) cannot fail for the
 * collapse configuration and are kept for completeness. */
int or_integrate(double *y, double tout, double dt0, int max_steps, int *counters,
                 double *t_end, double *dt_end) {
    or_counters c = {0, 0, 0, 0, 0, 0, 0};
    double fjac[NEQ * NEQ], e[NEQ * NEQ];
    double ynew[NEQ], ak1[NEQ], ak2[NEQ], work[NEQ], dy[NEQ];
    int ip[NEQ];
    float e32[NEQ * NEQ];
    int ip32[NEQ];
    double t = 0.0;
    int result;
    if (max_steps <= 0) max_steps = 10000000;

    if (tout <= t) {
        result = OR_BAD_INPUTS;
        *t_end = t;
        *dt_end = dt0;
        goto done;
    }
    for (int i = 0; i < NEQ; ++i) {
        if (atol_of(i) <= 0.0 || rtol_of(i) <= 10.0 * k_uround) {
            result = OR_TOO_MUCH_ACCURACY;
            *t_end = t;
            *dt_end = dt0;
            goto done;
        }
    }
    {
        const double hmaxn = tout - t;
        double h = dt0;
        if (fabs(h) <= 10.0 * k_uround) h = 1.0e-6;
        h = ref_min(fabs(h), hmaxn);
        int reject = 0, last = 0, nsing = 0, n_step = 0, n_accept = 0;
        double hacc = 0.0, erracc = 1.0, hopt = h, x = t;
        int jac_age = -1; /* accepted steps since J was evaluated; -1: none yet */
        for (;;) {
            if (n_step > max_steps) {
                *t_end = x; *dt_end = h; result = OR_TOO_MANY_STEPS; goto done;
            }
            if (0.1 * fabs(h) <= fabs(x) * k_uround) {
                *t_end = x; *dt_end = h; result = OR_DT_UNDERFLOW; goto done;
            }
            if (last) {
                *t_end = x; *dt_end = hopt; result = OR_SUCCESS; goto done;
            }
            hopt = h;
            if (x + h * 1.0001 >= tout) {
                h = tout - x;
                last = 1;
            }
            if (jac_age < 0 || jac_age >= g_jac_age_max) {
                hook_jac(y, fjac);
                c.n_jac += 1;
                jac_age = 0;
            }
            for (;;) {
                const double fac = 1.0 / (h * c_gamma);
                for (int i = 0; i < NEQ; ++i) {
                    for (int j = 0; j < NEQ; ++j) e[i * NEQ + j] = -fjac[i * NEQ + j];
                    e[i * NEQ + i] += fac;
                }
                int use64 = (g_lu_mode == 0), info;
                if (!use64) {
                    int ks[NEQ];
                    mixed_scales(y, ks);
                    for (int i = 0; i < NEQ; ++i)
                        for (int j = 0; j < NEQ; ++j)
                            e32[i * NEQ + j] = (float)ldexp(e[i * NEQ + j], ks[j] - ks[i]);
                    info = or_lu32(e32, ip32);
                    if (info != 0) { /* singular in FP32: FP64 factors for this attempt */
                        __atomic_fetch_add(&g_lu_fallbacks, 1, __ATOMIC_RELAXED);
                        info = or_lu(e, ip);
                        use64 = 1;
                    }
                } else {
                    info = or_lu(e, ip);
                }
#define OR_STAGE_SOLVE(v)                                                              \
    do {                                                                               \
        if (!use64) {                                                                  \
            double b_[NEQ];                                                            \
            memcpy(b_, (v), sizeof(b_));                                               \
            if (solve_mixed(e32, ip32, fjac, fac, y, g_lu_sweeps, (v))) break;         \
            __atomic_fetch_add(&g_refine_fallbacks, 1, __ATOMIC_RELAXED);             \
            for (int i_ = 0; i_ < NEQ; ++i_) {                                         \
                for (int j_ = 0; j_ < NEQ; ++j_) e[i_ * NEQ + j_] = -fjac[i_ * NEQ + j_];  \
                e[i_ * NEQ + i_] += fac;                                               \
            }                                                                          \
            if (or_lu(e, ip) != 0) break; /* keep the refined FP32 result */           \
            use64 = 1;                                                                 \
            memcpy((v), b_, sizeof(b_));                                               \
        }                                                                              \
        or_lu_solve(e, ip, (v));                                                       \
    } while (0)
                if (info != 0) {
                    nsing += 1;
                    if (nsing >= 5) {
                        *t_end = x; *dt_end = h; result = OR_LU_ERROR; goto done;
                    }
                    h *= 0.5;
                    reject = 1;
                    last = 0;
                    if (jac_age > 0) { /* J from an earlier state: evaluate it here */
                        hook_jac(y, fjac);
                        c.n_jac += 1;
                        jac_age = 0;
                    }
                    continue;
                }
                c.n_decomp += 1;

                hook_rhs(y, ak1);
                OR_STAGE_SOLVE(ak1);
                c.n_solve += 1;
                for (int i = 0; i < NEQ; ++i) {
                    ynew[i] = y[i] + c_a21 * ak1[i];
                    ak2[i] = (c_c21 / h) * ak1[i];
                }
                hook_rhs(ynew, dy);
                for (int i = 0; i < NEQ; ++i) ak2[i] += dy[i];
                OR_STAGE_SOLVE(ak2);
                c.n_solve += 1;
                for (int i = 0; i < NEQ; ++i) {
                    ynew[i] = y[i] + c_a31 * ak1[i] + c_a32 * ak2[i];
                    work[i] = (c_c31 * ak1[i] + c_c32 * ak2[i]) / h;
                }
                hook_rhs(ynew, dy);
                for (int i = 0; i < NEQ; ++i) work[i] += dy[i];
                OR_STAGE_SOLVE(work);
                c.n_solve += 1;
                for (int i = 0; i < NEQ; ++i) {
                    const double ak3 = work[i];
                    ynew[i] = y[i] + c_b1 * ak1[i] + c_b2 * ak2[i] + c_b3 * ak3;
                    work[i] = c_e1 * ak1[i] + c_e2 * ak2[i] + c_e3 * ak3;
                }
                c.n_rhs += 3;
                n_step += 1;
                c.n_step = n_step;

                /* error_norm, reproducer.cpp:7194-7202 */
                double err = 0.0;
                for (int i = 0; i < NEQ; ++i) {
                    const double sk = atol_of(i) + rtol_of(i) * ref_max(fabs(y[i]), fabs(ynew[i]));
                    const double term = work[i] / sk;
                    err += term * term;
                }
                err = sqrt(err / (double)NEQ);

                const double raw_fac = OR_CBRT(err) / k_safe;
                const double fac_step = ref_max(1.0 / k_fac_max, ref_min(1.0 / k_fac_min, raw_fac));
                double hnew = h / fac_step;
                if (err <= 1.0) {
                    n_accept += 1;
                    c.n_accept = n_accept;
                    if (n_accept > 1) {
                        const double facgus = ref_max(
                            1.0 / k_fac_max,
                            ref_min(1.0 / k_fac_min,
                                    (hacc / h) * OR_CBRT((err * err) / erracc) / k_safe));
                        hnew = h / ref_max(fac_step, facgus);
                    }
                    hacc = h;
                    erracc = ref_max(1.0e-2, err);
                    if (t_hrec) {
                        if (t_nh < t_hcap) t_hrec[t_nh] = h;
                        t_nh += 1;
                    }
                    memcpy(y, ynew, sizeof(ynew));
                    jac_age += 1;
                    x += h;
                    if (fabs(hnew) > hmaxn) hnew = hmaxn;
                    if (reject) hnew = ref_min(fabs(hnew), fabs(h));
                    reject = 0;
                    h = hnew;
                    break;
                }
                reject = 1;
                last = 0;
                h = hnew;
                if (n_accept >= 1) c.n_reject += 1;
                if (jac_age > 0) { /* J from an earlier state: evaluate it here */
                    hook_jac(y, fjac);
                    c.n_jac += 1;
                    jac_age = 0;
                }
            }
        }
    }
done:
    counters[0] = c.n_step;
    counters[1] = c.n_rhs;
    counters[2] = c.n_jac;
    counters[3] = c.n_decomp;
    counters[4] = c.n_solve;
    counters[5] = c.n_accept;
    counters[6] = c.n_reject;
    return result;
}

/* ---- cell driver (reproducer.cpp:7406-7785) ----------------------------- */

static uint64_t clamp_u64(int v) { return v > 0 ? (uint64_t)v : 0u; }

/* Integrator step limit of burn_ros2s: configure_ros2s sets 1e7
 * (reproducer.cpp:7589); tests lower it to force TOO_MANY_STEPS. */
static int g_integ_max_steps = 10000000;
void or_set_integrator_max_steps(int m) { g_integ_max_steps = m > 0 ? m : 10000000; }

/* burn_ros2s, reproducer.cpp:7592-7627 */
int or_burn_ros2s(or_burn_t *s, double dt, or_stats *st) {
    double y[NEQ], t_end, dt_end;
    int cnt[7];
    eos_rt(s);
    for (int n = 0; n < NSPEC; ++n) y[n] = s->xn[n];
    y[NSPEC] = s->e;
    const int r = or_integrate(y, dt, dt, g_integ_max_steps, cnt, &t_end, &dt_end);
    st->internal_steps += clamp_u64(cnt[0]);
    st->rhs_calls += clamp_u64(cnt[1]);
    st->jacobian_calls += clamp_u64(cnt[2]);
    st->decompositions += clamp_u64(cnt[3]);
    st->linear_solves += clamp_u64(cnt[4]);
    st->accepted_steps += clamp_u64(cnt[5]);
    st->rejected_steps += clamp_u64(cnt[6]);
    if (r != OR_SUCCESS) return r;
    for (int n = 0; n < NSPEC; ++n) s->xn[n] = y[n];
    s->e = y[NSPEC];
    return r;
}

int or_burn(double *burn, double dt, uint64_t *stats) {
    return or_burn_ros2s((or_burn_t *)burn, dt, (or_stats *)stats);
}

/* make_initial_state / make_collapse_state, reproducer.cpp:7424-7426, 7533-7546 */
void or_make_collapse_state(or_cell *c) {
    static const double n0[NSPEC] = {1.0e-4, 1.0e-4, 1.0e0, 1.0e-40, 1.0e-40, 1.0e-40, 1.0e-40,
                                     1.0e-40, 1.0e-6, 1.0e-40, 1.0e-40, 1.0e-40, 1.0e-40, 0.0775};
    memset(c, 0, sizeof(*c));
    c->current.T = 1.0e2;
    memcpy(c->current.xn, n0, sizeof(n0));
    c->current.rho = density_of(c->current.xn);
    normalize(&c->current);
    eos_rt(&c->current);
    c->density_driver = c->current.rho;
}

/* Benchmark-recipe initial state of cell `index` (the build's own recipe; the
 * reference has none): SplitMix64 chain h1 = sm(seed ^ index), h2 = sm(h1),
 * h3 = sm(h2); Box-Muller z = sqrt(-2 ln u1) cos(2 pi u2), u1 in (0, 1];
 * n_H = exp(sigma z); T = t_min, or t_min exp(u3 ln(t_max / t_min)) when
 * t_max > t_min; H+ 1e-4, H2 1e-6, H the rest of n_H, He Y = 0.24, D/H 3e-5,
 * other species 1e-40 (per n_H), electrons by charge balance; then rho,
 * normalise and eos_rt as make_initial_state (reproducer.cpp:7533-7541). */
uint64_t or_splitmix64(uint64_t v);
void or_make_recipe_state(or_cell *c, int index, uint64_t seed, double sigma, double t_min,
                          double t_max) {
    const uint64_t h1 = or_splitmix64(seed ^ (uint64_t)index);
    const uint64_t h2 = or_splitmix64(h1);
    const uint64_t h3 = or_splitmix64(h2);
    const double u1 = (double)((h1 >> 11) + 1) * 0x1p-53;
    const double u2 = (double)(h2 >> 11) * 0x1p-53;
    const double u3 = (double)(h3 >> 11) * 0x1p-53;
    const double z = sqrt(-2.0 * log(u1)) * cos((2.0 * k_pi) * u2);
    const double nh = exp(sigma * z);
    const double T = (t_max > t_min) ? t_min * exp(u3 * log(t_max / t_min)) : t_min;
    memset(c, 0, sizeof(*c));
    for (int n = 0; n < NSPEC; ++n) c->current.xn[n] = 1.0e-40 * nh;
    c->current.xn[1] = 1.0e-4 * nh;
    c->current.xn[8] = 1.0e-6 * nh;
    c->current.xn[2] = nh - c->current.xn[1] - 2.0 * c->current.xn[8];
    c->current.xn[5] = 3.0e-5 * nh;
    c->current.xn[13] = (0.24 / (4.0 * (1.0 - 0.24))) * nh;
    balance_charge(&c->current);
    c->current.T = T;
    c->current.rho = density_of(c->current.xn);
    normalize(&c->current);
    eos_rt(&c->current);
    c->density_driver = c->current.rho;
}

/* splitmix64 (reproducer.cpp:7548-7553; Steele, Lea & Flood 2014). */
uint64_t or_splitmix64(uint64_t v) {
    v += 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31);
}

/* perturbation_factor, reproducer.cpp:7555-7562 */
double or_perturbation_factor(int cell, int step) {
    const uint64_t seed = ((uint64_t)(uint32_t)cell << 32) ^ ((uint64_t)(uint32_t)step << 16);
    const uint64_t bits = or_splitmix64(seed) >> 11;
    const double unit = (double)bits * (1.0 / (double)(1ULL << 53));
    return 1.0 + 0.2 * (2.0 * unit - 1.0);
}

/* apply_perturbation, reproducer.cpp:7564-7580 */
void or_apply_perturbation(or_cell *c, int cell, int step, int enabled) {
    if (!enabled || step == 0 || step % 20 != 0) return;
    const double f = or_perturbation_factor(cell, step);
    c->density_driver *= f;
    c->current.rho *= f;
    for (int n = 0; n < NSPEC; ++n) c->current.xn[n] *= f;
    floor_and_normalize(&c->current);
    balance_charge(&c->current);
    floor_and_normalize(&c->current);
    eos_re(&c->current);
}

static int valid_positive(double v) { return v > 0.0 && isfinite(v); }

/* collapse_timestep, reproducer.cpp:7629-7644 */
double or_collapse_timestep(const or_cell *c) {
    const double rho = density_of(c->current.xn);
    if (!valid_positive(rho)) return -1.0;
    const double tff = sqrt(k_pi * 3.0 / (32.0 * rho * k_G));
    if (!valid_positive(tff)) return -1.0;
    const double dt = 1.0e-1 * tff;
    return valid_positive(dt) ? dt : -1.0;
}

/* post-step sequence, reproducer.cpp:7773-7776 */
void or_post_step(double *burn) {
    or_burn_t *b = (or_burn_t *)burn;
    floor_and_normalize(b);
    balance_charge(b);
    floor_and_normalize(b);
    eos_re(b);
}

/* prepare_grid_timestep_kernel body, reproducer.cpp:7706-7733 */
double or_prepare_cell(or_cell *s, int cell, int completed, double grid_time, int step,
                       int perturb, int *failure) {
    const double none = DBL_MAX;
    if (s->completed_steps < completed) return none;
    if (s->completed_steps != completed || s->time != grid_time) {
        *failure = OR_BAD_INPUTS;
        return none;
    }
    or_apply_perturbation(s, cell, step, perturb);
    const double dt = or_collapse_timestep(s);
    if (!valid_positive(dt) || !valid_positive(s->density_driver)) {
        *failure = OR_BAD_INPUTS;
        return none;
    }
    if (dt < 10.0) {
        const double tff = dt / 1.0e-1;
        s->density_driver += dt * (s->density_driver / tff);
        return none;
    }
    return dt;
}

/* The prepare phase of one outer step over a whole grid (the kernel body on
 * every cell, reproducer.cpp:7706-7733) and the host loop's minimum of the
 * candidates (7888-7893, `candidate < dt_grid`, so the lowest cell index
 * holding the minimum is the one found first).  *cell_out: that cell, or -1
 * when no cell offers a candidate (*dt_out = DBL_MAX); *failure_out: the
 * lowest failing cell's code (OR_SUCCESS if none). */
void or_prepare_grid(or_cell *cells, int n, int completed, double grid_time, int step, int perturb,
                     double *dt_out, int *cell_out, int *failure_out) {
    double best = DBL_MAX;
    int best_cell = -1, failure = OR_SUCCESS;
    for (int c = 0; c < n; ++c) {
        int f = OR_SUCCESS;
        const double cand = or_prepare_cell(&cells[c], c, completed, grid_time, step, perturb, &f);
        if (f != OR_SUCCESS && failure == OR_SUCCESS) failure = f;
        if (cand < best) {
            best = cand;
            best_cell = c;
        }
    }
    *dt_out = best;
    *cell_out = best_cell;
    *failure_out = failure;
}

/* advance_collapse_gridwide_kernel body, reproducer.cpp:7745-7784 */
int or_advance_cell(or_cell *s, int completed, double next_grid_time, double dt_grid,
                    int *failure) {
    if (s->completed_steps != completed) return 0;
    const double local_dt = or_collapse_timestep(s);
    if (local_dt < 10.0) return 0;
    const double old_density = s->density_driver;
    const double tff = local_dt / 1.0e-1;
    s->density_driver += dt_grid * (s->density_driver / tff);
    if (!valid_positive(s->density_driver) || s->density_driver > 2.0e-6) return 0;
    const double ratio = s->density_driver / old_density;
    for (int n = 0; n < NSPEC; ++n) s->current.xn[n] *= ratio;
    s->current.rho *= ratio;
    const int r = or_burn_ros2s(&s->current, dt_grid, &s->stats);
    if (r != OR_SUCCESS) {
        *failure = r;
        return 0;
    }
    or_post_step((double *)&s->current);
    s->time = next_grid_time;
    s->completed_steps = completed + 1;
    return 1;
}

/* Host loop of run_cells_cuda, reproducer.cpp:7859-7959 (gridwide-independent
 * policy), cells in parallel inside each phase.  max_steps<=0 means 1000. */
int or_run_gridwide(or_cell *cells, int n, int perturb, int max_steps, int *completed_out) {
    const int limit = max_steps > 0 ? max_steps : 1000;
    double grid_time = 0.0;
    int completed = *completed_out;
    int result = OR_SUCCESS;
    for (int step = 0; result == OR_SUCCESS && step < limit; ++step) {
        int failure = OR_SUCCESS;
        double dt_grid = DBL_MAX;
#pragma omp parallel for schedule(dynamic, 16) reduction(min : dt_grid)
        for (int c = 0; c < n; ++c) {
            int f = OR_SUCCESS;
            const double cand = or_prepare_cell(&cells[c], c, completed, grid_time, step, perturb, &f);
            if (cand < dt_grid) dt_grid = cand;
            if (f != OR_SUCCESS) {
#pragma omp critical
                if (failure == OR_SUCCESS) failure = f;
            }
        }
        if (failure != OR_SUCCESS) {
            result = failure;
            break;
        }
        if (dt_grid == DBL_MAX) break;
        const double next = grid_time + dt_grid;
        if (!valid_positive(dt_grid) || !isfinite(next)) {
            result = OR_BAD_INPUTS;
            break;
        }
        int integrated = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : integrated)
        for (int c = 0; c < n; ++c) {
            int f = OR_SUCCESS;
            integrated += or_advance_cell(&cells[c], completed, next, dt_grid, &f);
            if (f != OR_SUCCESS) {
#pragma omp critical
                if (failure == OR_SUCCESS) failure = f;
            }
        }
        if (failure != OR_SUCCESS) {
            result = failure;
            break;
        }
        if (integrated == 0) break;
        grid_time = next;
        completed += 1;
    }
    *completed_out = completed;
    return result;
}

/* The reference's CPU driver, legacy-local policy (reproducer.cpp:8630-8648):
 * every outer step each cell that has not stopped takes one step of its own
 * free-fall dt (advance_collapse_step, 7646-7685); the cell loop stops at the
 * first failing cell (cells after it keep their state), the run stops on a
 * failure or when every cell stopped.  max_steps <= 0 means 1000. */
static int or_advance_local(or_cell *s, int cell, int step, int perturb, int *failure) {
    or_apply_perturbation(s, cell, step, perturb);
    const double old_density = s->density_driver;
    const double dt = or_collapse_timestep(s);
    const double tff = dt / 1.0e-1;
    s->density_driver += dt * (s->density_driver / tff);
    if (dt < 10.0 || s->density_driver > 2.0e-6) return 1;
    const double ratio = s->density_driver / old_density;
    for (int n = 0; n < NSPEC; ++n) s->current.xn[n] *= ratio;
    s->current.rho *= ratio;
    const int r = or_burn_ros2s(&s->current, dt, &s->stats);
    if (r != OR_SUCCESS) {
        *failure = r;
        return 1;
    }
    or_post_step((double *)&s->current);
    s->time += dt;
    s->completed_steps += 1;
    return 0;
}

int or_run_legacy(or_cell *cells, int n, int perturb, int max_steps, int *completed_out,
                  int *failed_cell) {
    const int limit = max_steps > 0 ? max_steps : 1000;
    int failure = OR_SUCCESS, completed = 0;
    *failed_cell = -1;
    for (int step = 0; step < limit; ++step) {
        int all_stopped = 1;
        for (int c = 0; c < n; ++c) {
            if (cells[c].completed_steps < completed) continue;
            const int stopped = or_advance_local(&cells[c], c, step, perturb, &failure);
            if (failure != OR_SUCCESS) {
                *failed_cell = c;
                break;
            }
            all_stopped = all_stopped && stopped;
        }
        if (failure != OR_SUCCESS || all_stopped) break;
        completed += 1;
    }
    *completed_out = completed;
    return failure;
}

/* ---- gridwide-cta32 policy of the reference's Mojo front end ---------------
 * reproducer.mojo:2829-3093 with the driver loop of main (3604-3621).  The
 * outer step is the gridwide-independent one (prepare, grid minimum of
 * (dt, cell), density update, advance, commit), with these differences:
 *  - both density-driver updates evaluate (dt * driver) / tff
 *    (reproducer.mojo:2858-2860, 2891-2893), where reproducer.cpp:7730 and
 *    7759 evaluate dt * (driver / tff);
 *  - cells advance in tiles of `tile` consecutive cells (TileCells = 32,
 *    reproducer.mojo:40), one ROS2S step controller per tile
 *    (integrate_gridwide_tile, 2910-3048): each trial runs on the tile's
 *    participating cells in index order; a singular factorisation stops the
 *    trial there and halves h for the tile; otherwise the tile error is the
 *    largest cell error (DBL_MAX if any is not finite) and the Gustafsson
 *    controller acts on it; an accepted final trial sets x = dt_grid exactly;
 *  - the Jacobian is evaluated, and counted, on every trial (ros2s_trial,
 *    2460-2560: eval_jacobian 2420-2424 before the decomposition);
 *  - the first failing tile in index order ends the step and the run
 *    (advance_gridwide_step 3079-3088; main raises, so no state is written).
 * With tile = 1 and cpp_order = 1 (the density updates evaluated in the
 * order of reproducer.cpp) each cell has its own controller, and the run
 * equals or_run_gridwide in every state bit and every counter except
 * jacobian_calls (one per trial here, one per accepted step there). */

/* ros2s_trial, reproducer.mojo:2460-2560: one trial of step h from the cell's
 * (xn, e); ynew and the error norm out, the cell unchanged. */
static int cta_trial(const or_burn_t *s, double h, double *ynew, double *err_out, or_stats *st) {
    double y[NEQ], fjac[NEQ * NEQ], e[NEQ * NEQ], ak1[NEQ], ak2[NEQ], work[NEQ], dy[NEQ];
    int ip[NEQ];
    for (int n = 0; n < NSPEC; ++n) y[n] = s->xn[n];
    y[NSPEC] = s->e;
    hook_jac(y, fjac);
    st->jacobian_calls += 1;
    const double fac = 1.0 / (h * c_gamma);
    for (int i = 0; i < NEQ; ++i) {
        for (int j = 0; j < NEQ; ++j) e[i * NEQ + j] = -fjac[i * NEQ + j];
        e[i * NEQ + i] += fac;
    }
    if (or_lu(e, ip) != 0) return OR_LU_ERROR;
    st->decompositions += 1;
    hook_rhs(y, ak1);
    or_lu_solve(e, ip, ak1);
    st->linear_solves += 1;
    for (int i = 0; i < NEQ; ++i) {
        ynew[i] = y[i] + c_a21 * ak1[i];
        ak2[i] = (c_c21 / h) * ak1[i];
    }
    hook_rhs(ynew, dy);
    for (int i = 0; i < NEQ; ++i) ak2[i] += dy[i];
    or_lu_solve(e, ip, ak2);
    st->linear_solves += 1;
    for (int i = 0; i < NEQ; ++i) {
        ynew[i] = y[i] + c_a31 * ak1[i] + c_a32 * ak2[i];
        work[i] = (c_c31 * ak1[i] + c_c32 * ak2[i]) / h;
    }
    hook_rhs(ynew, dy);
    for (int i = 0; i < NEQ; ++i) work[i] += dy[i];
    or_lu_solve(e, ip, work);
    st->linear_solves += 1;
    for (int i = 0; i < NEQ; ++i) {
        const double ak3 = work[i];
        ynew[i] = y[i] + c_b1 * ak1[i] + c_b2 * ak2[i] + c_b3 * ak3;
        work[i] = c_e1 * ak1[i] + c_e2 * ak2[i] + c_e3 * ak3;
    }
    st->rhs_calls += 3;
    st->internal_steps += 1;
    /* error_norm, reproducer.mojo:2448-2457 */
    double err = 0.0;
    for (int i = 0; i < NEQ; ++i) {
        const double sk = atol_of(i) + rtol_of(i) * ref_max(fabs(y[i]), fabs(ynew[i]));
        const double term = work[i] / sk;
        err += term * term;
    }
    *err_out = sqrt(err / (double)NEQ);
    return OR_SUCCESS;
}

/* integrate_gridwide_tile, reproducer.mojo:2910-3048: cells [b, e) of which
 * those with act[c] participate; returns OR_SUCCESS or the tile's failure. */
static int cta_tile(or_cell *cells, const unsigned char *act, int b, int e, double dt_grid,
                    double next_grid_time, int completed) {
    double cand[64][NEQ];
    int participants = 0;
    for (int c = b; c < e; ++c) participants += act[c] ? 1 : 0;
    if (participants == 0) return OR_SUCCESS;
    double h = dt_grid, x = 0.0, hacc = 0.0, erracc = 1.0;
    int reject = 0, nsing = 0, n_step = 0, n_accept = 0;
    while (x < dt_grid) {
        if (n_step > g_integ_max_steps) return OR_TOO_MANY_STEPS;
        if (0.1 * fabs(h) <= fabs(x) * k_uround) return OR_DT_UNDERFLOW;
        int final_trial = 0;
        if (x + h * 1.0001 >= dt_grid) {
            h = dt_grid - x;
            final_trial = 1;
        }
        int singular = 0, nonfinite = 0;
        double err_tile = 0.0;
        for (int c = b; c < e; ++c) {
            if (!act[c] || singular) continue;
            double err = 0.0;
            if (cta_trial(&cells[c].current, h, cand[c - b], &err, &cells[c].stats) != OR_SUCCESS)
                singular = 1;
            else if (!isfinite(err))
                nonfinite = 1;
            else
                err_tile = ref_max(err_tile, err);
        }
        if (singular) {
            nsing += 1;
            if (nsing >= 5) return OR_LU_ERROR;
            h *= 0.5;
            reject = 1;
            continue;
        }
        if (nonfinite) err_tile = DBL_MAX;
        n_step += 1;
        const double raw_fac = OR_CBRT(err_tile) / k_safe;
        const double fac_step = ref_max(1.0 / k_fac_max, ref_min(1.0 / k_fac_min, raw_fac));
        double hnew = h / fac_step;
        if (err_tile <= 1.0) {
            n_accept += 1;
            if (n_accept > 1) {
                const double facgus = ref_max(
                    1.0 / k_fac_max,
                    ref_min(1.0 / k_fac_min,
                            (hacc / h) * OR_CBRT((err_tile * err_tile) / erracc) / k_safe));
                hnew = h / ref_max(fac_step, facgus);
            }
            hacc = h;
            erracc = ref_max(1.0e-2, err_tile);
            for (int c = b; c < e; ++c) {
                if (!act[c]) continue;
                for (int n = 0; n < NSPEC; ++n) cells[c].current.xn[n] = cand[c - b][n];
                cells[c].current.e = cand[c - b][NSPEC];
                cells[c].stats.accepted_steps += 1;
            }
            x = final_trial ? dt_grid : x + h;
            if (fabs(hnew) > dt_grid) hnew = dt_grid;
            if (reject) hnew = ref_min(fabs(hnew), fabs(h));
            reject = 0;
            h = hnew;
            continue;
        }
        reject = 1;
        h = hnew;
        if (n_accept >= 1)
            for (int c = b; c < e; ++c)
                if (act[c]) cells[c].stats.rejected_steps += 1;
    }
    for (int c = b; c < e; ++c) {
        if (!act[c]) continue;
        or_post_step((double *)&cells[c].current);
        cells[c].time = next_grid_time;
        cells[c].completed_steps = completed + 1;
    }
    return OR_SUCCESS;
}

/* The driver loop (reproducer.mojo:3604-3621) over advance_gridwide_step
 * (3051-3093): prepare_grid_timestep (2829-2871), setup_gridwide_cells
 * (2874-2907), the tiles in index order.  tile in [1, 64]; max_steps <= 0
 * means 1000; cpp_order: evaluate the density updates as reproducer.cpp
 * does (a pinning aid; the policy is cpp_order = 0).  *limiting_cell: the last step's (dt, lowest cell) holder. */
int or_run_cta(or_cell *cells, int n, int perturb, int max_steps, int tile, int cpp_order,
               int *completed_out, int *limiting_cell) {
    const int limit = max_steps > 0 ? max_steps : 1000;
    if (tile < 1 || tile > 64) return OR_BAD_INPUTS;
    unsigned char *act = (unsigned char *)calloc(n > 0 ? (size_t)n : 1, 1);
    int *rc = (int *)calloc(n / tile + 1, sizeof(int));
    double grid_time = 0.0;
    int completed = *completed_out, result = OR_SUCCESS;
    *limiting_cell = -1;
    for (int step = 0; step < limit; ++step) {
        /* prepare_grid_timestep: the lowest failing cell's code wins (the
         * loop returns at the first) */
        double dt_grid = DBL_MAX;
        int cell_min = -1, active_count = 0, failure = OR_SUCCESS;
        for (int c = 0; c < n && failure == OR_SUCCESS; ++c) {
            or_cell *s = &cells[c];
            act[c] = 0;
            if (s->completed_steps < completed) continue;
            if (s->completed_steps != completed || s->time != grid_time) {
                failure = OR_BAD_INPUTS;
                break;
            }
            or_apply_perturbation(s, c, step, perturb);
            const double dt = or_collapse_timestep(s);
            if (dt <= 0.0 || !valid_positive(s->density_driver)) {
                failure = OR_BAD_INPUTS;
                break;
            }
            if (dt < 10.0) {
                const double tff = dt / 1.0e-1;
                if (cpp_order)
                    s->density_driver += dt * (s->density_driver / tff);
                else
                    s->density_driver += dt * s->density_driver / tff;
                continue;
            }
            act[c] = 1;
            active_count += 1;
            if (dt < dt_grid || (dt == dt_grid && c < cell_min)) {
                dt_grid = dt;
                cell_min = c;
            }
        }
        *limiting_cell = cell_min;
        if (failure != OR_SUCCESS) {
            result = failure;
            break;
        }
        if (active_count == 0) break;
        const double next = grid_time + dt_grid;
        if (!valid_positive(dt_grid) || !isfinite(next)) {
            result = OR_BAD_INPUTS;
            break;
        }
        /* setup_gridwide_cells */
        int participants = 0;
        for (int c = 0; c < n; ++c) {
            if (!act[c]) continue;
            or_cell *s = &cells[c];
            const double local_dt = or_collapse_timestep(s);
            if (local_dt < 10.0) {
                act[c] = 0;
                continue;
            }
            const double tff = local_dt / 1.0e-1;
            const double old_density = s->density_driver;
            if (cpp_order)
                s->density_driver += dt_grid * (s->density_driver / tff);
            else
                s->density_driver += dt_grid * s->density_driver / tff;
            if (!valid_positive(s->density_driver) || s->density_driver > 2.0e-6) {
                act[c] = 0;
                continue;
            }
            const double ratio = s->density_driver / old_density;
            for (int k = 0; k < NSPEC; ++k) s->current.xn[k] *= ratio;
            s->current.rho *= ratio;
            eos_rt(&s->current);
            participants += 1;
        }
        if (participants == 0) break;
        /* tiles are independent until the first failure: run them in
         * parallel, then take the lowest failing tile's code */
        const int ntiles = (n + tile - 1) / tile;
#pragma omp parallel for schedule(dynamic, 1)
        for (int t = 0; t < ntiles; ++t) {
            const int b = t * tile, e = (b + tile < n) ? b + tile : n;
            rc[t] = cta_tile(cells, act, b, e, dt_grid, next, completed);
        }
        for (int t = 0; t < ntiles && failure == OR_SUCCESS; ++t)
            if (rc[t] != OR_SUCCESS) failure = rc[t];
        if (failure != OR_SUCCESS) {
            result = failure;
            break;
        }
        grid_time = next;
        completed += 1;
    }
    free(act);
    free(rc);
    *completed_out = completed;
    return result;
}

int or_sizeof_cell(void) { return (int)sizeof(or_cell); }

/* eos_rt on a burn_t laid out as rho, T, e, xn[14] (reproducer.cpp:306-309) */
void or_eos_rt(double *burn) { eos_rt((or_burn_t *)burn); }

/* The libm this build evaluates the network with (for the libm harness). */
double or_libm_exp(double x) { return CHEM_EXP(x); }
double or_libm_log(double x) { return CHEM_LOG(x); }
double or_libm_cbrt(double x) { return OR_CBRT(x); }

/* Replay cells of a gridwide run one by one: given the run's dt_grid of every
 * outer step (the grid-wide minimum the host loop took, reproducer.cpp:7888-
 * 7897), each cell's trajectory no longer depends on the other cells.  Cell k
 * has grid index index[k]; status[k] receives its first failure code or
 * OR_SUCCESS.  Steps start at completed0 with grid time time0. */
void or_replay_cells(or_cell *cells, const int *index, int m, const double *dt_hist, int nsteps,
                     int perturb, int completed0, double time0, int *status) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < m; ++k) {
        double t = time0;
        int f = OR_SUCCESS;
        for (int s = 0; s < nsteps && f == OR_SUCCESS; ++s) {
            const int step = completed0 + s;
            or_prepare_cell(&cells[k], index[k], step, t, step, perturb, &f);
            if (f != OR_SUCCESS) break;
            const double next = t + dt_hist[s];
            or_advance_cell(&cells[k], step, next, dt_hist[s], &f);
            t = next;
        }
        status[k] = f;
    }
}

/* As or_replay_cells, also recording the ROS2S attempts each cell makes in
 * each step: attempts[k * nsteps + s]. */
void or_replay_cells_profile(or_cell *cells, const int *index, int m, const double *dt_hist,
                             int nsteps, int perturb, int completed0, double time0, int *status,
                             int *attempts) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < m; ++k) {
        double t = time0;
        int f = OR_SUCCESS;
        for (int s = 0; s < nsteps; ++s) attempts[(size_t)k * nsteps + s] = 0;
        for (int s = 0; s < nsteps && f == OR_SUCCESS; ++s) {
            const int step = completed0 + s;
            or_prepare_cell(&cells[k], index[k], step, t, step, perturb, &f);
            if (f != OR_SUCCESS) break;
            const double next = t + dt_hist[s];
            const uint64_t before = cells[k].stats.internal_steps;
            or_advance_cell(&cells[k], step, next, dt_hist[s], &f);
            attempts[(size_t)k * nsteps + s] = (int)(cells[k].stats.internal_steps - before);
            t = next;
        }
        status[k] = f;
    }
}

/* The register-budgeted (staged) network functions the GPU kernels compile,
 * evaluated here with a local staging array: they must reproduce chem_fused /
 * chem_rhs bit for bit (only the evaluation order differs). */
void or_jac_rhs_staged(const double *y, double *jac, double *dydt) {
    double X[NSPEC];
    const double T = temperature_from_y(y, X);
    double stg[CHEM_FUSED_NSLOT];
    chem_fused_staged(T, X, stg, jac, dydt);
}

void or_rhs_staged(const double *y, double *dydt) {
    double X[NSPEC];
    const double T = temperature_from_y(y, X);
    double stg[CHEM_RHS_NSLOT];
    chem_rhs_staged(T, X, stg, (double *)0, dydt);
}

#ifdef OR_ADJOINT
#include "adjoint.c"
#endif
