/*
 * ORACLE -- test infrastructure, not product code.  Included at the end of
 * oracle.c in the OR_ADJOINT build (_build/liboracle_adj.so), so it shares the
 * oracle's LU, integrator and cell driver.
 *
 * Tangent-linear and adjoint of the gridwide-independent collapse advance
 * with its controller decisions frozen (no reference counterpart; the
 * forward pieces follow reproducer.cpp:7736-7785, 7205-7347, 311-339).  Per
 * outer step a cell maps
 *     (xn, T, rho, driver) -pre-> (y0, rho1, driver1) -ROS2S substeps-> y_n
 *     -post-> (xn', T'),  e' = y_n[14],  rho' = rho1,  driver' = driver1
 * with dt_grid and the accepted step sizes h_k of the recorded run held fixed
 * (rejected attempts leave y unchanged).  One accepted ROS2S step with
 * E = I/(h gamma) - J(y):
 *     k1 = E^-1 f(y)
 *     k2 = E^-1 (c21/h k1 + f(u2)),             u2 = y + a21 k1
 *     k3 = E^-1 ((c31 k1 + c32 k2)/h + f(u3)),  u3 = y + a31 k1 + a32 k2
 *     y' = y + b1 k1 + b2 k2 + b3 k3
 * Its tangent uses dk = E^-1 (dr + dJ k) and its adjoint r_bar = E^-T k_bar,
 * J_bar += r_bar k^T; f's and J's own tangents/adjoints and those of pre/post
 * are generated from the network DAG (warp-reactor_amd/gen/ad.py).  f and J
 * carry the rate multipliers M_r = exp(theta_r) of the network's 45 reactions
 * (J's tangent and adjoint therefore take M, dM and give M-bar too).
 *
 * Pinning (tests/test_adjoint.py): the record run at M = 1 equals the plain
 * oracle run bit for bit; replay of the recorded steps reproduces the
 * recorded checkpoints bit for bit; the tangent agrees with central finite
 * differences of the replay; the adjoint agrees with the tangent through the
 * dot-product identity <w, T v> = <T^T w, v>; the transposed LU solve against
 * numpy.
 */
#include <stdlib.h>

#include "_build/chem_adjoint.h"

#define AD_NF 18 /* state fields: xn[14], T, e, rho, driver */

static void ad_fj_hook(const double *y, const double *M, double *f, double *J) { ad_fj(y, M, f, J); }
static void ad_f_hook(const double *y, const double *M, double *f) { ad_f(y, M, f); }

int or_adj_ntheta(void) { return AD_NTHETA; }

/* E^T x = b with LINPACK factors (dgesl, job != 0): U^T, then the L_k^T and
 * interchanges in reverse elimination order. */
void or_lu_solve_t(const double *lu, const int *ipvt, double *x) {
    for (int k = 0; k < NEQ; ++k) {
        double t = 0.0;
        for (int j = 0; j < k; ++j) t += lu[j * NEQ + k] * x[j];
        x[k] = (x[k] - t) / lu[k * NEQ + k];
    }
    for (int k = NEQ - 2; k >= 0; --k) {
        double t = 0.0;
        for (int j = k + 1; j < NEQ; ++j) t += lu[j * NEQ + k] * x[j];
        x[k] += t;
        const int p = ipvt[k];
        if (p != k) {
            const double v = x[p];
            x[p] = x[k];
            x[k] = v;
        }
    }
}

typedef struct {
    double J[NEQ * NEQ], lu[NEQ * NEQ], fac;
    int ip[NEQ];
    double k1[NEQ], k2[NEQ], k3[NEQ], u2[NEQ], u3[NEQ];
} ad_sub;

/* One accepted attempt of or_integrate at step size h, the same operations in
 * the same order (f0 from the fused J + f evaluation: the same DAG nodes). */
static int ad_substep(const double *y, double h, const double *M, ad_sub *s, double *ynew) {
    double f[NEQ], dy[NEQ];
    ad_fj(y, M, f, s->J);
    s->fac = 1.0 / (h * c_gamma);
    for (int i = 0; i < NEQ; ++i) {
        for (int j = 0; j < NEQ; ++j) s->lu[i * NEQ + j] = -s->J[i * NEQ + j];
        s->lu[i * NEQ + i] += s->fac;
    }
    if (or_lu(s->lu, s->ip) != 0) return OR_LU_ERROR;
    memcpy(s->k1, f, sizeof(f));
    or_lu_solve(s->lu, s->ip, s->k1);
    for (int i = 0; i < NEQ; ++i) {
        s->u2[i] = y[i] + c_a21 * s->k1[i];
        s->k2[i] = (c_c21 / h) * s->k1[i];
    }
    ad_f(s->u2, M, dy);
    for (int i = 0; i < NEQ; ++i) s->k2[i] += dy[i];
    or_lu_solve(s->lu, s->ip, s->k2);
    for (int i = 0; i < NEQ; ++i) {
        s->u3[i] = y[i] + c_a31 * s->k1[i] + c_a32 * s->k2[i];
        s->k3[i] = (c_c31 * s->k1[i] + c_c32 * s->k2[i]) / h;
    }
    ad_f(s->u3, M, dy);
    for (int i = 0; i < NEQ; ++i) s->k3[i] += dy[i];
    or_lu_solve(s->lu, s->ip, s->k3);
    for (int i = 0; i < NEQ; ++i)
        ynew[i] = y[i] + c_b1 * s->k1[i] + c_b2 * s->k2[i] + c_b3 * s->k3[i];
    return OR_SUCCESS;
}

static void matvec_add(const double *J, const double *v, double *out) {
    for (int i = 0; i < NEQ; ++i) {
        double t = 0.0;
        for (int j = 0; j < NEQ; ++j) t += J[i * NEQ + j] * v[j];
        out[i] += t;
    }
}

/* tangent of one substep: dy (in) -> dynew */
static void ad_substep_tl(const double *y, double h, const double *M, const double *dM,
                          const ad_sub *s, const double *dy, double *dynew) {
    double dJ[NEQ * NEQ], dk1[NEQ], dk2[NEQ], dk3[NEQ], du[NEQ], t[NEQ];
    ad_jac_jvp(y, M, dy, dM, dJ);
    ad_f_jvp(y, M, dy, dM, dk1);
    matvec_add(dJ, s->k1, dk1);
    or_lu_solve(s->lu, s->ip, dk1);
    for (int i = 0; i < NEQ; ++i) du[i] = dy[i] + c_a21 * dk1[i];
    ad_f_jvp(s->u2, M, du, dM, t);
    for (int i = 0; i < NEQ; ++i) dk2[i] = (c_c21 / h) * dk1[i] + t[i];
    matvec_add(dJ, s->k2, dk2);
    or_lu_solve(s->lu, s->ip, dk2);
    for (int i = 0; i < NEQ; ++i) du[i] = dy[i] + c_a31 * dk1[i] + c_a32 * dk2[i];
    ad_f_jvp(s->u3, M, du, dM, t);
    for (int i = 0; i < NEQ; ++i) dk3[i] = (c_c31 * dk1[i] + c_c32 * dk2[i]) / h + t[i];
    matvec_add(dJ, s->k3, dk3);
    or_lu_solve(s->lu, s->ip, dk3);
    for (int i = 0; i < NEQ; ++i)
        dynew[i] = dy[i] + c_b1 * dk1[i] + c_b2 * dk2[i] + c_b3 * dk3[i];
}

/* Whether every rate multiplier is exactly 1 (theta = 0). */
static int multipliers_are_one(const double *M) {
    for (int r = 0; r < AD_NTHETA; ++r)
        if (M[r] != 1.0) return 0;
    return 1;
}

/* adjoint of one substep: yb (in: adjoint of y', out: adjoint of y); Mbar += */
static void ad_substep_adj(const double *y, double h, const double *M, const ad_sub *s,
                           double *yb, double *Mbar) {
    double W[NEQ * NEQ], kb1[NEQ], kb2[NEQ], kb3[NEQ], rb[NEQ], ub[NEQ], mb[AD_NTHETA];
    for (int i = 0; i < NEQ; ++i) {
        kb1[i] = c_b1 * yb[i];
        kb2[i] = c_b2 * yb[i];
        kb3[i] = c_b3 * yb[i];
    }
    /* stage 3 */
    memcpy(rb, kb3, sizeof(rb));
    or_lu_solve_t(s->lu, s->ip, rb);
    for (int i = 0; i < NEQ; ++i)
        for (int j = 0; j < NEQ; ++j) W[i * NEQ + j] = rb[i] * s->k3[j];
    for (int i = 0; i < NEQ; ++i) {
        kb1[i] += (c_c31 / h) * rb[i];
        kb2[i] += (c_c32 / h) * rb[i];
    }
    ad_f_vjp(s->u3, M, rb, ub, mb);
    for (int r = 0; r < AD_NTHETA; ++r) Mbar[r] += mb[r];
    for (int i = 0; i < NEQ; ++i) {
        yb[i] += ub[i];
        kb1[i] += c_a31 * ub[i];
        kb2[i] += c_a32 * ub[i];
    }
    /* stage 2 */
    memcpy(rb, kb2, sizeof(rb));
    or_lu_solve_t(s->lu, s->ip, rb);
    for (int i = 0; i < NEQ; ++i)
        for (int j = 0; j < NEQ; ++j) W[i * NEQ + j] += rb[i] * s->k2[j];
    for (int i = 0; i < NEQ; ++i) kb1[i] += (c_c21 / h) * rb[i];
    ad_f_vjp(s->u2, M, rb, ub, mb);
    for (int r = 0; r < AD_NTHETA; ++r) Mbar[r] += mb[r];
    for (int i = 0; i < NEQ; ++i) {
        yb[i] += ub[i];
        kb1[i] += c_a21 * ub[i];
    }
    /* stage 1 */
    memcpy(rb, kb1, sizeof(rb));
    or_lu_solve_t(s->lu, s->ip, rb);
    for (int i = 0; i < NEQ; ++i)
        for (int j = 0; j < NEQ; ++j) W[i * NEQ + j] += rb[i] * s->k1[j];
    ad_f_vjp(y, M, rb, ub, mb);
    for (int r = 0; r < AD_NTHETA; ++r) Mbar[r] += mb[r];
    for (int i = 0; i < NEQ; ++i) yb[i] += ub[i];
    /* J(y, M) = J(y) - D(y, M): the reference Jacobian's adjoint, then D's.
     * D = sum_r (1 - M_r) dG_r/dz is linear in M: with every M_r = 1 its
     * y-adjoint is zero and only its M-adjoint (which does not depend on M)
     * is computed. */
    ad_jac_vjp(y, W, ub);
    for (int i = 0; i < NEQ; ++i) yb[i] += ub[i];
    if (multipliers_are_one(M)) {
        ad_jacd_mbar(y, W, mb);
        for (int r = 0; r < AD_NTHETA; ++r) Mbar[r] += mb[r];
        return;
    }
    ad_jacd_vjp(y, M, W, ub, mb);
    for (int r = 0; r < AD_NTHETA; ++r) Mbar[r] += mb[r];
    for (int i = 0; i < NEQ; ++i) yb[i] += ub[i];
}

static void get_state(const double *a, int n, int c, double *st) {
    for (int f = 0; f < AD_NF; ++f) st[f] = a[(size_t)f * n + c];
}
static void put_state(double *a, int n, int c, const double *st) {
    for (int f = 0; f < AD_NF; ++f) a[(size_t)f * n + c] = st[f];
}

/* Perturbation steps of a run whose loop index starts at 0
 * (apply_perturbation, reproducer.cpp:7564-7580): step != 0, step % 20 == 0.
 * The record keeps the state before each of them (the perturbation is not
 * invertible: it floors and renormalises), at index step / 20 - 1. */
static int ad_perturbed(int perturb, int step) { return perturb && step != 0 && step % 20 == 0; }
int or_adj_npre(int nsteps) { return nsteps > 1 ? (nsteps - 1) / 20 : 0; }

/* apply_perturbation on a state [18] through the generated post-step
 * sequence (floor, normalise, charge, floor, normalise, eos_re: the same
 * operations): xn, rho and the driver scaled by the step's factor, e kept, T
 * recomputed; with its tangent when dst != NULL. */
static void ad_perturb(double *st, double *dst, int cell, int step) {
    const double f = or_perturbation_factor(cell, step);
    double y[NEQ], ps[4] = {0.0, st[16] * f, 0.0, 0.0}, xn[NSPEC], T;
    for (int i = 0; i < NSPEC; ++i) y[i] = st[i] * f;
    y[NSPEC] = st[15];
    if (dst) {
        double dy[NEQ], pds[4] = {0.0, dst[16] * f, 0.0, 0.0}, dxn[NSPEC], dT;
        for (int i = 0; i < NSPEC; ++i) dy[i] = dst[i] * f;
        dy[NSPEC] = dst[15];
        ad_post_jvp(y, ps, dy, pds, dxn, &dT);
        memcpy(dst, dxn, sizeof(dxn));
        dst[14] = dT;
        dst[16] = pds[1];
        dst[17] = dst[17] * f;
    }
    ad_post(y, ps, xn, &T);
    memcpy(st, xn, sizeof(xn));
    st[14] = T;
    st[16] = ps[1];
    st[17] = st[17] * f;
}

/* Adjoint of ad_perturb at the state st0 before it: lam in, the adjoint of
 * the perturbed state; out, of st0 (T of st0 does not enter: eos_re
 * recomputes it). */
static void ad_perturb_adj(const double *st0, double *lam, int cell, int step) {
    const double f = or_perturbation_factor(cell, step);
    double y[NEQ], ps[4] = {0.0, st0[16] * f, 0.0, 0.0}, wT[1] = {lam[14]}, yb[NEQ], rb[1];
    for (int i = 0; i < NSPEC; ++i) y[i] = st0[i] * f;
    y[NSPEC] = st0[15];
    ad_post_vjp(y, ps, lam, wT, yb, rb);
    for (int i = 0; i < NSPEC; ++i) lam[i] = f * yb[i];
    lam[14] = 0.0;
    lam[15] = yb[NSPEC] + lam[15];
    lam[16] = f * (rb[0] + lam[16]);
    lam[17] = f * lam[17];
}

/* Record: the gridwide loop of or_run_gridwide (perturbations when
 * `perturb`), with rate multipliers M, recording per step dt_grid, each
 * cell's state at the start of its advance (ckpt[step][field][cell]), the
 * state before each perturbation step (pre[step / 20 - 1][field][cell]) and
 * the accepted step sizes (htape[cell][k], hoff[step][cell] = first index of
 * the step).  Every cell must integrate in every step.  Returns OR_SUCCESS, a
 * failure code, or -100 (a cell stopped integrating) / -101 (tape capacity
 * hcap exceeded). */
int or_adj_record(or_cell *cells, int n, int nsteps, const double *M, double *dts, double *ckpt,
                  int *hoff, double *htape, int hcap, int perturb, double *pre) {
    double grid_time = 0.0;
    int completed = 0, result = OR_SUCCESS;
    for (int c = 0; c < n; ++c) hoff[c] = 0;
    for (int step = 0; step < nsteps; ++step) {
        double dt_grid = DBL_MAX;
        int failure = OR_SUCCESS;
        if (ad_perturbed(perturb, step))
            for (int c = 0; c < n; ++c) {
                const or_cell *s = &cells[c];
                double st[AD_NF];
                memcpy(st, s->current.xn, sizeof(double) * NSPEC);
                st[14] = s->current.T;
                st[15] = s->current.e;
                st[16] = s->current.rho;
                st[17] = s->density_driver;
                put_state(pre + (size_t)(step / 20 - 1) * AD_NF * n, n, c, st);
            }
        for (int c = 0; c < n; ++c) {
            int f = OR_SUCCESS;
            const double cand =
                or_prepare_cell(&cells[c], c, completed, grid_time, step, perturb, &f);
            if (cand < dt_grid) dt_grid = cand;
            if (f != OR_SUCCESS && failure == OR_SUCCESS) failure = f;
        }
        if (failure != OR_SUCCESS) return failure;
        if (dt_grid == DBL_MAX) return -100;
        const double next = grid_time + dt_grid;
        dts[step] = dt_grid;
#pragma omp parallel for schedule(dynamic, 4) reduction(min : result)
        for (int c = 0; c < n; ++c) {
            or_cell *s = &cells[c];
            double st[AD_NF];
            memcpy(st, s->current.xn, sizeof(double) * NSPEC);
            st[14] = s->current.T;
            st[15] = s->current.e;
            st[16] = s->current.rho;
            st[17] = s->density_driver;
            put_state(ckpt + (size_t)step * AD_NF * n, n, c, st);
            const int h0 = hoff[(size_t)step * n + c];
            or_set_theta_hooks(M, ad_fj_hook, ad_f_hook);
            or_set_h_recorder(htape + (size_t)c * hcap + h0, hcap - h0);
            int f = OR_SUCCESS;
            const int integ = or_advance_cell(s, completed, next, dt_grid, &f);
            const int nh = or_h_recorded();
            or_set_h_recorder(NULL, 0);
            or_set_theta_hooks(NULL, NULL, NULL);
            int r = OR_SUCCESS;
            if (f != OR_SUCCESS)
                r = f;
            else if (!integ)
                r = -100;
            else if (h0 + nh > hcap)
                r = -101;
            hoff[(size_t)(step + 1) * n + c] = h0 + nh;
            if (r < result) result = r;
        }
        if (result != OR_SUCCESS) return result;
        grid_time = next;
        completed += 1;
    }
    return result;
}

/* Frozen map of one cell from state st over steps [s0, s1), with tangent
 * (dst != NULL) or not. */
static int ad_cell_forward(double *st, double *dst, int c, int n, int s0, int s1, const double *dts,
                           const int *hoff, const double *htape, int hcap, const double *M,
                           const double *dM, int perturb) {
    for (int step = s0; step < s1; ++step) {
        if (ad_perturbed(perturb, step)) ad_perturb(st, dst, c, step);
        double sv[4] = {st[14], st[16], st[17], dts[step]};
        double y[NEQ], s1v[2], dy[NEQ], ds1[2] = {0.0, 0.0};
        ad_pre(st, sv, y, s1v);
        if (dst) {
            double dsv[4] = {dst[14], dst[16], dst[17], 0.0};
            ad_pre_jvp(st, sv, dst, dsv, dy, ds1);
        }
        const int a = hoff[(size_t)step * n + c], b = hoff[(size_t)(step + 1) * n + c];
        for (int k = a; k < b; ++k) {
            const double h = htape[(size_t)c * hcap + k];
            ad_sub s;
            double ynew[NEQ];
            if (ad_substep(y, h, M, &s, ynew) != OR_SUCCESS) return OR_LU_ERROR;
            if (dst) {
                double dynew[NEQ];
                ad_substep_tl(y, h, M, dM, &s, dy, dynew);
                memcpy(dy, dynew, sizeof(dy));
            }
            memcpy(y, ynew, sizeof(y));
        }
        double ps[4] = {0.0, s1v[0], 0.0, 0.0}, xn[NSPEC], T;
        if (dst) {
            double pds[4] = {0.0, ds1[0], 0.0, 0.0}, dxn[NSPEC], dT;
            ad_post_jvp(y, ps, dy, pds, dxn, &dT);
            memcpy(dst, dxn, sizeof(dxn));
            dst[14] = dT;
            dst[15] = dy[14];
            dst[16] = ds1[0];
            dst[17] = ds1[1];
        }
        ad_post(y, ps, xn, &T);
        memcpy(st, xn, sizeof(xn));
        st[14] = T;
        st[15] = y[14];
        st[16] = s1v[0];
        st[17] = s1v[1];
    }
    return OR_SUCCESS;
}

/* Replay of the frozen map from s0 [field][cell] over nsteps, into sN;
 * with tangents ds0 -> dsN (dM: multiplier tangent) when ds0 != NULL. */
int or_adj_replay(const double *s0, const double *ds0, int n, int nsteps, const double *dts,
                  const int *hoff, const double *htape, int hcap, const double *M,
                  const double *dM, double *sN, double *dsN, int perturb) {
    int result = OR_SUCCESS;
#pragma omp parallel for schedule(dynamic, 4) reduction(min : result)
    for (int c = 0; c < n; ++c) {
        double st[AD_NF], dst[AD_NF];
        get_state(s0, n, c, st);
        if (ds0) get_state(ds0, n, c, dst);
        const int r = ad_cell_forward(st, ds0 ? dst : NULL, c, n, 0, nsteps, dts, hoff, htape,
                                      hcap, M, dM, perturb);
        if (r < result) result = r;
        put_state(sN, n, c, st);
        if (ds0) put_state(dsN, n, c, dst);
    }
    return result;
}

/* Adjoint: gN [field][cell] = dL/d(final state) -> g0 = dL/d(state at the
 * first advance), Mbar [r][cell] = dL/dM_r per cell.  Recomputes each step
 * from its checkpoint and the substeps inside it from the stored y_k. */
int or_adj_backward(const double *ckpt, int n, int nsteps, const double *dts, const int *hoff,
                    const double *htape, int hcap, const double *M, const double *gN, double *g0,
                    double *Mbar, int perturb, const double *pre) {
    int result = OR_SUCCESS;
#pragma omp parallel for schedule(dynamic, 4) reduction(min : result)
    for (int c = 0; c < n; ++c) {
        double lam[AD_NF], mb[AD_NTHETA];
        get_state(gN, n, c, lam);
        memset(mb, 0, sizeof(mb));
        for (int step = nsteps - 1; step >= 0; --step) {
            double st[AD_NF];
            get_state(ckpt + (size_t)step * AD_NF * n, n, c, st);
            double sv[4] = {st[14], st[16], st[17], dts[step]};
            double y[NEQ], s1v[2];
            ad_pre(st, sv, y, s1v);
            const int a = hoff[(size_t)step * n + c], b = hoff[(size_t)(step + 1) * n + c];
            double *ys = (double *)malloc(sizeof(double) * NEQ * (size_t)(b - a + 1));
            memcpy(ys, y, sizeof(y));
            for (int k = a; k < b; ++k) {
                ad_sub s;
                if (ad_substep(ys + (size_t)(k - a) * NEQ, htape[(size_t)c * hcap + k], M, &s,
                               ys + (size_t)(k - a + 1) * NEQ) != OR_SUCCESS)
                    result = OR_LU_ERROR;
            }
            /* post */
            double ps[4] = {0.0, s1v[0], 0.0, 0.0}, wT[1] = {lam[14]}, yb[NEQ], rb[1];
            ad_post_vjp(ys + (size_t)(b - a) * NEQ, ps, lam, wT, yb, rb);
            yb[14] += lam[15];
            const double rho1b = lam[16] + rb[0];
            for (int k = b - 1; k >= a; --k) {
                ad_sub s;
                double yn[NEQ];
                const double *yk = ys + (size_t)(k - a) * NEQ;
                const double h = htape[(size_t)c * hcap + k];
                ad_substep(yk, h, M, &s, yn);
                ad_substep_adj(yk, h, M, &s, yb, mb);
            }
            free(ys);
            double ws[2] = {rho1b, lam[17]}, xb[NSPEC], sb[3];
            ad_pre_vjp(st, sv, yb, ws, xb, sb);
            memcpy(lam, xb, sizeof(xb));
            lam[14] = sb[0];
            lam[15] = 0.0;
            lam[16] = sb[1];
            lam[17] = sb[2];
            if (ad_perturbed(perturb, step)) {
                double st0[AD_NF];
                get_state(pre + (size_t)(step / 20 - 1) * AD_NF * n, n, c, st0);
                ad_perturb_adj(st0, lam, c, step);
            }
        }
        put_state(g0, n, c, lam);
        for (int r = 0; r < AD_NTHETA; ++r) Mbar[(size_t)r * n + c] = mb[r];
    }
    return result;
}

/* The generated pieces on their own (tests). */
void or_adj_f(const double *y, const double *M, double *f) { ad_f(y, M, f); }
void or_adj_fj(const double *y, const double *M, double *f, double *J) { ad_fj(y, M, f, J); }
void or_adj_f_jvp(const double *y, const double *M, const double *dy, const double *dM, double *df) {
    ad_f_jvp(y, M, dy, dM, df);
}
void or_adj_f_vjp(const double *y, const double *M, const double *w, double *yb, double *mb) {
    ad_f_vjp(y, M, w, yb, mb);
}
void or_adj_jac_jvp(const double *y, const double *M, const double *dy, const double *dM, double *dJ) {
    ad_jac_jvp(y, M, dy, dM, dJ);
}
void or_adj_jacd_vjp(const double *y, const double *M, const double *W, double *yb, double *mb) {
    ad_jacd_vjp(y, M, W, yb, mb);
}
void or_adj_jacd_mbar(const double *y, const double *W, double *mb) { ad_jacd_mbar(y, W, mb); }
void or_adj_jac_vjp(const double *y, const double *M, const double *W, double *yb, double *mb) {
    double ub[NEQ];
    ad_jac_vjp(y, W, yb);
    ad_jacd_vjp(y, M, W, ub, mb);
    for (int i = 0; i < NEQ; ++i) yb[i] += ub[i];
}
