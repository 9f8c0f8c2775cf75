/* Rate-table study: test infrastructure, included by oracle.c when built with
 * -DOR_RATE_TABLE (_build/liboracle_table.so); never part of the product.
 *
 * The reference evaluates every rate fit at every call
 * (reproducer.cpp:341-6994); its performance notes list tabulated or fitted
 * rates as an experiment gated by the accuracy budget
 * (docs/performance/optimization-ideas.md:152-169).  This measures that gate
 * on the oracle: the network's temperature-only frontier values (chem_tvals,
 * generated by gen/table.py: smooth values, the selections between fits stay
 * exact) are taken from a table on a uniform grid in ln T and interpolated
 * with the 4-point Lagrange cubic: in ln |v| for values of one sign at every
 * node (values that underflow to 0 held at ln 1e-300), in v otherwise.
 *
 *   mode 0  chem_fused / chem_rhs as generated
 *   mode 1  chem_*_tv with exact tv from chem_tvals (equals mode 0 bit for bit)
 *   mode 2  chem_*_tv with tv from the table; T outside [tlo, thi] evaluates
 *           the fits (counted)
 *   mode 3  as mode 2 with every value interpolated in v (no exp per value)
 *
 * The mode is process-wide: set it before a run, not during one. */

static int g_rt_mode = 0;
static int g_rt_n = 0;
static double g_rt_a = 0.0, g_rt_b = 0.0, g_rt_h = 1.0; /* ln tlo, ln thi, spacing */
static double g_rt_tlo = 0.0, g_rt_thi = 0.0;
static double *g_rt_tab = NULL;             /* [node][CHEM_NTV] */
/* 0: the table holds v; 1: ln v; -1: ln(-v) */
static signed char g_rt_inlog[CHEM_NTV];
#define RT_TINY 1e-300
static long long g_rt_outside = 0;

int or_rt_ntv(void) { return CHEM_NTV; }
long long or_rt_outside(void) { return g_rt_outside; }

static void rt_interp(double T, double *tv) {
    const double u = (log(T) - g_rt_a) / g_rt_h;
    int i = (int)floor(u);
    if (i < 1) i = 1;
    if (i > g_rt_n - 3) i = g_rt_n - 3;
    const double t = u - i;
    const double w0 = -t * (t - 1.0) * (t - 2.0) / 6.0;
    const double w1 = (t + 1.0) * (t - 1.0) * (t - 2.0) / 2.0;
    const double w2 = -(t + 1.0) * t * (t - 2.0) / 2.0;
    const double w3 = (t + 1.0) * t * (t - 1.0) / 6.0;
    const double *r0 = g_rt_tab + (size_t)(i - 1) * CHEM_NTV;
    const double *r1 = r0 + CHEM_NTV, *r2 = r1 + CHEM_NTV, *r3 = r2 + CHEM_NTV;
    for (int k = 0; k < CHEM_NTV; ++k) {
        const double s = w0 * r0[k] + w1 * r1[k] + w2 * r2[k] + w3 * r3[k];
        tv[k] = g_rt_inlog[k] == 0 ? s : g_rt_inlog[k] > 0 ? exp(s) : -exp(s);
    }
}

static void rt_values(double T, double *tv) {
    if (g_rt_mode >= 2 && T >= g_rt_tlo && T <= g_rt_thi) {
        rt_interp(T, tv);
        return;
    }
    if (g_rt_mode >= 2) {
#pragma omp atomic
        g_rt_outside++;
    }
    chem_tvals(T, tv);
}

/* Select the mode; modes 2 and 3 build the table over [tlo, thi] with n >= 4
 * nodes.  Returns 0, or -1 on bad arguments or allocation failure. */
int or_rt_set(int mode, double tlo, double thi, int n) {
    if (mode < 0 || mode > 3) return -1;
    g_rt_outside = 0;
    if (mode < 2) {
        g_rt_mode = mode;
        return 0;
    }
    if (!(tlo > 0.0) || !(thi > tlo) || n < 4) return -1;
    double *tab = (double *)malloc((size_t)n * CHEM_NTV * sizeof(double));
    if (!tab) return -1;
    free(g_rt_tab);
    g_rt_tab = tab;
    g_rt_n = n;
    g_rt_tlo = tlo;
    g_rt_thi = thi;
    g_rt_a = log(tlo);
    g_rt_b = log(thi);
    g_rt_h = (g_rt_b - g_rt_a) / (n - 1);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) chem_tvals(exp(g_rt_a + i * g_rt_h), tab + (size_t)i * CHEM_NTV);
    for (int k = 0; k < CHEM_NTV; ++k) {
        int nonneg = 1, nonpos = 1;
        for (int i = 0; i < n; ++i) {
            const double v = tab[(size_t)i * CHEM_NTV + k];
            nonneg &= v >= 0.0;
            nonpos &= v <= 0.0;
        }
        g_rt_inlog[k] = (signed char)(mode == 3 ? 0 : nonneg ? 1 : nonpos ? -1 : 0);
        if (g_rt_inlog[k] != 0)
            for (int i = 0; i < n; ++i) {
                double *v = &tab[(size_t)i * CHEM_NTV + k];
                *v = log(fmax(fabs(*v), RT_TINY));
            }
    }
    g_rt_mode = mode;
    return 0;
}

/* Interpolated against exact frontier values at the given temperatures:
 * err[k] = max over the samples of |interp - exact| / |exact|, where
 * |exact| > 1e-250. */
void or_rt_errors(const double *T, int m, double *err) {
    for (int k = 0; k < CHEM_NTV; ++k) err[k] = 0.0;
    if (!g_rt_tab) return;
    double tv[CHEM_NTV], ex[CHEM_NTV];
    for (int j = 0; j < m; ++j) {
        rt_interp(T[j], tv);
        chem_tvals(T[j], ex);
        for (int k = 0; k < CHEM_NTV; ++k) {
            if (!(fabs(ex[k]) > 1e-250)) continue;
            const double e = fabs(tv[k] - ex[k]) / fabs(ex[k]);
            if (e > err[k]) err[k] = e;
        }
    }
}
