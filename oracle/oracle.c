/*
 * CPU oracle for the FMM/QBX interaction path.  TEST INFRASTRUCTURE ONLY:
 * imported by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, never by the sumpy_amd package.
 *
 * FP64 restatement of the reference's CPU algorithms, in their loop order:
 *   - P2P dense            sumpy/p2p.py:232-258 (pair expression :120-174)
 *   - P2PFromCSR (CPU)     sumpy/p2p.py:626-670
 *   - Laplace derivatives  sumpy/derivative_taker.py:195-262 (Fernando2021
 *                          recurrence; max(mi)<=1 by direct differentiation)
 *   - P2M                  sumpy/expansion/multipole.py:78-114 + transpose
 *                          projection expansion/__init__.py:588-597
 *   - M2M ("slow" form)    sumpy/expansion/multipole.py:363-402
 *   - M2L (non-FFT)        sumpy/expansion/m2l.py:461-558 (class data and
 *                          preprocessing), sumpy/tools.py:689-701 (Toeplitz
 *                          matvec), m2l.py:679-705 (postprocessing)
 *   - L2L                  sumpy/expansion/local.py:348-466
 *   - L2P                  sumpy/expansion/local.py:284-309
 *   - M2P                  sumpy/expansion/multipole.py:131-157
 *   - P2L                  sumpy/expansion/local.py:226-271
 *   - QBX line Taylor 2D   sumpy/qbx.py:262-308, expansion/local.py:132-185
 *
 * Index tables (identifier lists, projection matrices) are built by the
 * oracle's own Python layer (oracle/tables.py) and passed in.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex cplx;

#define KNL_LAP2D 2
#define KNL_LAP3D 3
#define KNL_HELM3D 13

/* {{{ pair kernels: value and gradient w.r.t. d = target - source, unscaled */

static void lap3d_pair(const double d[3], double *g, double grad[3])
{
    double r2 = d[0]*d[0] + d[1]*d[1] + d[2]*d[2];
    double rinv = 1.0/sqrt(r2);
    *g = rinv;
    double r3 = rinv*rinv*rinv;
    for (int i = 0; i < 3; ++i) grad[i] = -d[i]*r3;
}

static void lap2d_pair(const double d[2], double *g, double grad[2])
{
    double r2 = d[0]*d[0] + d[1]*d[1];
    *g = 0.5*log(r2);
    for (int i = 0; i < 2; ++i) grad[i] = d[i]/r2;
}

static void helm3d_pair(const double d[3], cplx k, cplx *g, cplx grad[3])
{
    double r2 = d[0]*d[0] + d[1]*d[1] + d[2]*d[2];
    double r = sqrt(r2);
    cplx e = cexp(I*k*r);
    *g = e/r;
    /* d/dr (e^{ikr}/r) = e^{ikr} (ik r - 1)/r^2, times d_i/r */
    cplx dr = e*(I*k*r - 1.0)/(r2*r);
    for (int i = 0; i < 3; ++i) grad[i] = dr*d[i];
}

/* }}} */

/* {{{ P2P */

/*
 * weights: real kernels use w0[ns] and wd[dim][ns] (wd may be NULL);
 * Helmholtz uses interleaved complex arrays.  Outputs pot[nt] and grad[dim][nt]
 * (either may be NULL), scaled by *scale*; the reference's exclude_self zeroes
 * the pair isrc == t2s[itgt] (p2p.py:162-170).
 */
static void pair_accumulate(int knl, int dim, cplx k, const double *d,
                            double w0r, double w0i, const double *wdr,
                            const double *wdi, cplx *acc_pot, cplx acc_grad[3],
                            int want_pot, int want_grad)
{
    (void)want_pot;
    if (knl == KNL_HELM3D) {
        cplx g, gr[3];
        helm3d_pair(d, k, &g, gr);
        cplx w0 = w0r + I*w0i;
        *acc_pot += w0*g;
        if (wdr) for (int a = 0; a < 3; ++a) *acc_pot += (wdr[a] + I*wdi[a])*gr[a];
        if (want_grad) {
            /* Hessian of e^{ikr}/r: f'(r)/r (delta - dd/r^2) + f''(r) dd/r^2 */
            double r2 = d[0]*d[0] + d[1]*d[1] + d[2]*d[2];
            double r = sqrt(r2);
            cplx e = cexp(I*k*r);
            cplx f1 = e*(I*k*r - 1.0)/r2;                        /* f'(r) */
            cplx f2 = e*(-k*k*r2 - 2.0*I*k*r + 2.0)/(r2*r);      /* f''(r) */
            for (int b = 0; b < 3; ++b) {
                acc_grad[b] += w0*gr[b];
                if (wdr) for (int a = 0; a < 3; ++a) {
                    double dd = d[a]*d[b]/r2;
                    cplx h = f1/r*((a == b ? 1.0 : 0.0) - dd) + f2*dd;
                    acc_grad[b] += (wdr[a] + I*wdi[a])*h;
                }
            }
        }
        return;
    }
    double g, gr[3];
    if (knl == KNL_LAP3D) lap3d_pair(d, &g, gr);
    else lap2d_pair(d, &g, gr);
    *acc_pot += w0r*g;
    if (wdr) for (int a = 0; a < dim; ++a) *acc_pot += wdr[a]*gr[a];
    if (want_grad) {
        double r2 = 0; for (int a = 0; a < dim; ++a) r2 += d[a]*d[a];
        for (int b = 0; b < dim; ++b) {
            acc_grad[b] += w0r*gr[b];
            if (wdr) for (int a = 0; a < dim; ++a) {
                double h;
                if (knl == KNL_LAP3D) {
                    double rinv = 1.0/sqrt(r2);
                    double r5 = rinv*rinv*rinv*rinv*rinv;
                    h = (3.0*d[a]*d[b] - (a == b ? r2 : 0.0))*r5;
                } else {
                    h = ((a == b ? r2 : 0.0) - 2.0*d[a]*d[b])/(r2*r2);
                }
                acc_grad[b] += wdr[a]*h;
            }
        }
    }
}

static void load_weights(int knl, int dim, const double *w0, const double *const *wd,
                         long s, double *w0r, double *w0i, double *wdr, double *wdi,
                         int *has_d)
{
    if (knl == KNL_HELM3D) {
        *w0r = w0[2*s]; *w0i = w0[2*s+1];
        if (wd) for (int a = 0; a < dim; ++a) { wdr[a] = wd[a][2*s]; wdi[a] = wd[a][2*s+1]; }
    } else {
        *w0r = w0[s]; *w0i = 0;
        if (wd) for (int a = 0; a < dim; ++a) { wdr[a] = wd[a][s]; wdi[a] = 0; }
    }
    *has_d = wd != NULL;
}

static void store_out(int knl, int dim, long t, double scale, cplx pot, const cplx *grad,
                      double *out_pot, double *const *out_grad, int accumulate)
{
    if (knl == KNL_HELM3D) {
        if (out_pot) {
            if (!accumulate) { out_pot[2*t] = 0; out_pot[2*t+1] = 0; }
            out_pot[2*t] += scale*creal(pot); out_pot[2*t+1] += scale*cimag(pot);
        }
        if (out_grad) for (int b = 0; b < dim; ++b) {
            if (!accumulate) { out_grad[b][2*t] = 0; out_grad[b][2*t+1] = 0; }
            out_grad[b][2*t] += scale*creal(grad[b]);
            out_grad[b][2*t+1] += scale*cimag(grad[b]);
        }
    } else {
        if (out_pot) {
            if (!accumulate) out_pot[t] = 0;
            out_pot[t] += scale*creal(pot);
        }
        if (out_grad) for (int b = 0; b < dim; ++b) {
            if (!accumulate) out_grad[b][t] = 0;
            out_grad[b][t] += scale*creal(grad[b]);
        }
    }
}

void orc_p2p_dense(int knl, int dim, double k_re, double k_im, double scale,
                   long nt, const double *const *tgt, long ns, const double *const *src,
                   const double *w0, const double *const *wd, const int *t2s,
                   double *out_pot, double *const *out_grad)
{
    cplx k = k_re + I*k_im;
    #pragma omp parallel for schedule(dynamic, 16)
    for (long t = 0; t < nt; ++t) {
        cplx pot = 0, grad[3] = {0, 0, 0};
        double x[3];
        for (int a = 0; a < dim; ++a) x[a] = tgt[a][t];
        for (long s = 0; s < ns; ++s) {
            if (t2s && s == t2s[t]) continue;
            double d[3];
            for (int a = 0; a < dim; ++a) d[a] = x[a] - src[a][s];
            double w0r, w0i, wdr[3], wdi[3]; int hd;
            load_weights(knl, dim, w0, wd, s, &w0r, &w0i, wdr, wdi, &hd);
            pair_accumulate(knl, dim, k, d, w0r, w0i, hd ? wdr : NULL, wdi,
                            &pot, grad, 1, out_grad != NULL);
        }
        store_out(knl, dim, t, scale, pot, grad, out_pot, out_grad, 0);
    }
}

void orc_p2p_csr(int knl, int dim, double k_re, double k_im, double scale,
                 long ntgt_boxes, const int *target_boxes, const int *src_box_starts,
                 const int *src_box_lists, const int *box_src_starts,
                 const int *box_src_counts, const int *box_tgt_starts,
                 const int *box_tgt_counts, const double *const *tgt,
                 const double *const *src, const double *w0, const double *const *wd,
                 const int *t2s, double *out_pot, double *const *out_grad)
{
    cplx k = k_re + I*k_im;
    #pragma omp parallel for schedule(dynamic, 4)
    for (long ib = 0; ib < ntgt_boxes; ++ib) {
        int tb = target_boxes[ib];
        long t0 = box_tgt_starts[tb], t1 = t0 + box_tgt_counts[tb];
        for (long t = t0; t < t1; ++t) {
            cplx pot = 0, grad[3] = {0, 0, 0};
            for (int isb = src_box_starts[ib]; isb < src_box_starts[ib+1]; ++isb) {
                int sb = src_box_lists[isb];
                long s0 = box_src_starts[sb], s1 = s0 + box_src_counts[sb];
                for (long s = s0; s < s1; ++s) {
                    if (t2s && s == t2s[t]) continue;
                    double d[3];
                    for (int a = 0; a < dim; ++a) d[a] = tgt[a][t] - src[a][s];
                    double w0r, w0i, wdr[3], wdi[3]; int hd;
                    load_weights(knl, dim, w0, wd, s, &w0r, &w0i, wdr, wdi, &hd);
                    pair_accumulate(knl, dim, k, d, w0r, w0i, hd ? wdr : NULL, wdi,
                                    &pot, grad, 1, out_grad != NULL);
                }
            }
            store_out(knl, dim, t, scale, pot, grad, out_pot, out_grad, 0);
        }
    }
}

/* }}} */

/* {{{ Laplace 3D scaled derivatives  D(m) = rscale^|m| d^m (1/|x|) */

/*
 * Fill D over the box [0..M0]x[0..M1]x[0..M2] with |m| <= maxorder, in
 * increasing |m|.  For max(m) <= 1 the derivative is taken directly
 * (derivative_taker.py:225-226 falls back to symbolic differentiation);
 * otherwise the 3D recurrence of derivative_taker.py:227-259 with d the first
 * axis where m_d >= 2.
 */
static void lap3d_derivs(const double x[3], double rscale, int maxorder,
                         const int M[3], double *D)
{
    const int S1 = M[2] + 1, S0 = (M[1] + 1)*S1;
    double xs[3] = {x[0]/rscale, x[1]/rscale, x[2]/rscale};
    double r2s = xs[0]*xs[0] + xs[1]*xs[1] + xs[2]*xs[2];
    double r2 = x[0]*x[0] + x[1]*x[1] + x[2]*x[2];
    double r = sqrt(r2);
    for (int n = 0; n <= maxorder; ++n)
    for (int m0 = 0; m0 <= M[0] && m0 <= n; ++m0)
    for (int m1 = 0; m1 <= M[1] && m0 + m1 <= n; ++m1) {
        int m2 = n - m0 - m1;
        if (m2 > M[2]) continue;
        int mi[3] = {m0, m1, m2};
        double val;
        if (m0 <= 1 && m1 <= 1 && m2 <= 1) {
            double rp = pow(rscale, n);
            if (n == 0) val = 1.0/r;
            else if (n == 1) {
                int ax = m0 ? 0 : (m1 ? 1 : 2);
                val = -x[ax]/(r2*r);
            } else if (n == 2) {
                int a = m0 ? 0 : 1, b = m2 ? 2 : 1;
                val = 3.0*x[a]*x[b]/(r2*r2*r);
            } else {
                val = -15.0*x[0]*x[1]*x[2]/(r2*r2*r2*r);
            }
            val *= rp;
        } else {
            int d = mi[0] >= 2 ? 0 : (mi[1] >= 2 ? 1 : 2);
            double e = 0;
            for (int i = 0; i < 3; ++i) {
                int nn = mi[i];
                int m1i[3] = {mi[0], mi[1], mi[2]}; m1i[i] -= 1;
                int m2i[3] = {mi[0], mi[1], mi[2]}; m2i[i] -= 2;
                double v1 = (m1i[i] >= 0) ? D[m1i[0]*S0 + m1i[1]*S1 + m1i[2]] : 0;
                double v2 = (m2i[i] >= 0) ? D[m2i[0]*S0 + m2i[1]*S1 + m2i[2]] : 0;
                if (i == d) {
                    e -= (2*nn - 1)*xs[i]*v1;
                    e -= (double)(nn - 1)*(nn - 1)*v2;
                } else {
                    e -= 2*nn*xs[i]*v1;
                    e -= (double)nn*(nn - 1)*v2;
                }
            }
            val = e/r2s;
        }
        D[m0*S0 + m1*S1 + m2] = val;
    }
}

void orc_lap3d_derivs(const double *x, double rscale, int maxorder, const int *M,
                      double *D)
{
    lap3d_derivs(x, rscale, maxorder, M, D);
}

/* }}} */

/* {{{ expansion helpers */

static double fact(int n) { double f = 1; for (int i = 2; i <= n; ++i) f *= i; return f; }

/* D_stored for stored ids ids[ns][3] from the derivative box */
static void gather_stored(const double *Dbox, const int M[3], int n, const int *ids,
                          double *out)
{
    const int S1 = M[2] + 1, S0 = (M[1] + 1)*S1;
    for (int i = 0; i < n; ++i)
        out[i] = Dbox[ids[3*i]*S0 + ids[3*i+1]*S1 + ids[3*i+2]];
}

/* }}} */

/* {{{ P2M: coeffs[box] = P^T sum_src w (a/r)^mi / mi!,  a = center - source */

void orc_p2m(int nfull, const int *full_ids, int nstored, const double *P /* nfull x nstored */,
             long nboxes, const int *source_boxes, const int *box_src_starts,
             const int *box_src_counts, const double *const *centers,
             const double *const *src, const double *w0, const double *const *wd,
             double rscale, double *tgt_expansions, int tgt_base_ibox)
{
    #pragma omp parallel for schedule(dynamic, 4)
    for (long ib = 0; ib < nboxes; ++ib) {
        int b = source_boxes[ib];
        double *cf = calloc(nfull, sizeof(double));
        long s0 = box_src_starts[b], s1 = s0 + box_src_counts[b];
        for (long s = s0; s < s1; ++s) {
            double a[3];
            for (int i = 0; i < 3; ++i) a[i] = (centers[i][b] - src[i][s])/rscale;
            for (int j = 0; j < nfull; ++j) {
                const int *mi = full_ids + 3*j;
                double mon = 1;
                for (int i = 0; i < 3; ++i) mon *= pow(a[i], mi[i])/fact(mi[i]);
                double c = w0[s]*mon;
                if (wd) for (int ax = 0; ax < 3; ++ax) {
                    if (mi[ax] == 0) continue;
                    /* d/da_ax of a^mi/mi! = a^(mi-e_ax)/(mi-e_ax)!, and the
                       1/rscale from the scaled variable */
                    double m2 = 1;
                    for (int i = 0; i < 3; ++i) {
                        int e = mi[i] - (i == ax);
                        m2 *= pow(a[i], e)/fact(e);
                    }
                    c += wd[ax][s]*m2/rscale;
                }
                cf[j] += c;
            }
        }
        double *out = tgt_expansions + (long)(b - tgt_base_ibox)*nstored;
        for (int k = 0; k < nstored; ++k) {
            double acc = 0;
            for (int j = 0; j < nfull; ++j) acc += P[(long)j*nstored + k]*cf[j];
            out[k] = acc;
        }
        free(cf);
    }
}

/* }}} */

/* {{{ M2M, "slow" binomial form of multipole.py:363-402, then compression */

void orc_m2m(int nfull_t, const int *full_ids_t, int nstored_t, const double *P_t,
             int nstored_s, const int *stored_ids_s,
             long ntgt_boxes, const int *target_boxes, const int *box_child_ids,
             long nboxes_total, const double *const *centers,
             double src_rscale, double tgt_rscale,
             const double *src_expansions, int src_base_ibox,
             double *tgt_expansions, int tgt_base_ibox)
{
    #pragma omp parallel for schedule(dynamic, 4)
    for (long ib = 0; ib < ntgt_boxes; ++ib) {
        int tb = target_boxes[ib];
        double *full = calloc(nfull_t, sizeof(double));
        for (int c = 0; c < 8; ++c) {
            int sb = box_child_ids[(long)c*nboxes_total + tb];
            if (sb == 0) continue;
            double d[3];
            for (int i = 0; i < 3; ++i) d[i] = (centers[i][tb] - centers[i][sb])/tgt_rscale;
            const double *sc = src_expansions + (long)(sb - src_base_ibox)*nstored_s;
            for (int j = 0; j < nfull_t; ++j) {
                const int *tm = full_ids_t + 3*j;
                double acc = 0;
                for (int s = 0; s < nstored_s; ++s) {
                    const int *sm = stored_ids_s + 3*s;
                    if (sm[0] > tm[0] || sm[1] > tm[1] || sm[2] > tm[2]) continue;
                    double contrib = sc[s]*fact(sm[0])*fact(sm[1])*fact(sm[2]);
                    for (int i = 0; i < 3; ++i) {
                        int n = tm[i], kk = sm[i];
                        double binom = fact(n)/(fact(kk)*fact(n - kk));
                        contrib *= binom*pow(d[i], n - kk);
                    }
                    contrib *= pow(src_rscale/tgt_rscale, sm[0] + sm[1] + sm[2]);
                    acc += contrib;
                }
                full[j] += acc/(fact(tm[0])*fact(tm[1])*fact(tm[2]));
            }
        }
        double *out = tgt_expansions + (long)(tb - tgt_base_ibox)*nstored_t;
        for (int k = 0; k < nstored_t; ++k) {
            double acc = 0;
            for (int j = 0; j < nfull_t; ++j) acc += P_t[(long)j*nstored_t + k]*full[j];
            out[k] += acc;
        }
        free(full);
    }
}

/* }}} */

/* {{{ M2L through translation classes (non-FFT form) */

/*
 * Class data for one translation vector d (m2l.py:461-537): stored
 * derivatives of order (p_src + p_tgt) restricted to max_mi, projected to all
 * needed derivatives by the PDE (P2 is the projection of that wrangler), and
 * scattered into the circulant layout at the needed positions (others zero).
 */
void orc_m2l_class_data(const double *d, double src_rscale, int order2, const int *max_mi,
                        int nfull2, const int *full_ids2, int nstored2,
                        const int *stored_ids2, const double *P2,
                        int nneeded, const int *needed_full_idx, const int *needed_circ_idx,
                        int ncirc, double *out)
{
    int M[3] = {max_mi[0] + 2, max_mi[1] + 2, max_mi[2] + 2};
    long nbox = (long)(M[0] + 1)*(M[1] + 1)*(M[2] + 1);
    double *Dbox = calloc(nbox, sizeof(double));
    lap3d_derivs(d, src_rscale, order2, M, Dbox);
    double *st = malloc(sizeof(double)*nstored2);
    gather_stored(Dbox, M, nstored2, stored_ids2, st);
    double *full = calloc(nfull2, sizeof(double));
    for (int j = 0; j < nfull2; ++j) {
        double acc = 0;
        for (int k = 0; k < nstored2; ++k) acc += P2[(long)j*nstored2 + k]*st[k];
        full[j] = acc;
    }
    (void)full_ids2;
    for (int i = 0; i < ncirc; ++i) out[i] = 0;
    for (int i = 0; i < nneeded; ++i) out[needed_circ_idx[i]] = full[needed_full_idx[i]];
    free(full); free(st); free(Dbox);
}

/*
 * One level of M2L with translation classes: preprocess (scatter into the
 * circulant vector, m2l.py:539-558), Toeplitz matvec (tools.py:689-701), then
 * select stored targets and scale (m2l.py:679-705).  Overwrites
 * tgt_expansions rows of the listed target boxes (e2e.py:472-474).
 */
void orc_m2l(int ncirc, int nstored_s, const int *src_circ_idx,
             int nstored_t, const int *tgt_circ_idx, const int *tgt_degree,
             long ntgt_boxes, const int *target_boxes, const int *src_box_starts,
             const int *src_box_lists, const int *class_lists, int class_level_start,
             const double *class_data, double src_rscale, double tgt_rscale,
             const double *src_expansions, int src_base_ibox,
             double *tgt_expansions, int tgt_base_ibox)
{
    #pragma omp parallel for schedule(dynamic, 2)
    for (long ib = 0; ib < ntgt_boxes; ++ib) {
        int tb = target_boxes[ib];
        double *res = calloc(ncirc, sizeof(double));
        double *circ = malloc(sizeof(double)*ncirc);
        for (int is = src_box_starts[ib]; is < src_box_starts[ib+1]; ++is) {
            int sb = src_box_lists[is];
            int cls = class_lists[is] - class_level_start;
            const double *data = class_data + (long)cls*ncirc;
            const double *sc = src_expansions + (long)(sb - src_base_ibox)*nstored_s;
            for (int i = 0; i < ncirc; ++i) circ[i] = 0;
            for (int i = 0; i < nstored_s; ++i) circ[src_circ_idx[i]] = sc[i];
            for (int row = 0; row < ncirc; ++row) {
                double acc = 0;
                for (int col = row; col < ncirc; ++col) acc += circ[col - row]*data[col];
                res[row] += acc;
            }
        }
        double ratio = tgt_rscale/src_rscale;
        double *out = tgt_expansions + (long)(tb - tgt_base_ibox)*nstored_t;
        for (int i = 0; i < nstored_t; ++i)
            out[i] = res[tgt_circ_idx[i]]*pow(ratio, tgt_degree[i]);
        free(circ); free(res);
    }
}

/* }}} */

/* {{{ L2L (local.py:348-466): expand full derivatives, shift along d, select */

void orc_l2l(int nfull_s, const int *full_ids_s, int nstored_s, const double *P_s,
             int nstored_t, const int *stored_ids_t, int order_s,
             long ntgt_boxes, const int *target_boxes, const int *box_parent_ids,
             const double *const *centers, double src_rscale, double tgt_rscale,
             const double *src_expansions, int src_base_ibox,
             double *tgt_expansions, int tgt_base_ibox)
{
    #pragma omp parallel for schedule(dynamic, 8)
    for (long ib = 0; ib < ntgt_boxes; ++ib) {
        int tb = target_boxes[ib];
        int sb = box_parent_ids[tb];
        double d[3];
        for (int i = 0; i < 3; ++i) d[i] = (centers[i][tb] - centers[i][sb])/src_rscale;
        const double *sc = src_expansions + (long)(sb - src_base_ibox)*nstored_s;
        double *full = malloc(sizeof(double)*nfull_s);
        for (int j = 0; j < nfull_s; ++j) {
            double acc = 0;
            for (int k = 0; k < nstored_s; ++k) acc += P_s[(long)j*nstored_s + k]*sc[k];
            full[j] = acc;
        }
        double *out = tgt_expansions + (long)(tb - tgt_base_ibox)*nstored_t;
        for (int i = 0; i < nstored_t; ++i) {
            const int *mi = stored_ids_t + 3*i;
            int deg = mi[0] + mi[1] + mi[2];
            if (deg > order_s) continue;
            double acc = 0;
            for (int j = 0; j < nfull_s; ++j) {
                const int *sm = full_ids_s + 3*j;
                if (sm[0] < mi[0] || sm[1] < mi[1] || sm[2] < mi[2]) continue;
                double c = full[j];
                for (int a = 0; a < 3; ++a) {
                    int q = sm[a] - mi[a];
                    c *= pow(d[a], q)/fact(q);
                }
                acc += c;
            }
            out[i] += acc*pow(tgt_rscale/src_rscale, deg);
        }
        free(full);
    }
}

/* }}} */

/* {{{ L2P (local.py:284-309), overwrite; grad via the target derivative */

void orc_l2p(int nfull, const int *full_ids, int nstored, const double *P,
             long ntgt_boxes, const int *target_boxes, const int *box_tgt_starts,
             const int *box_tgt_counts, const double *const *centers,
             const double *const *tgt, double rscale, double scale,
             const double *src_expansions, int src_base_ibox,
             double *out_pot, double *const *out_grad)
{
    #pragma omp parallel for schedule(dynamic, 8)
    for (long ib = 0; ib < ntgt_boxes; ++ib) {
        int tb = target_boxes[ib];
        const double *sc = src_expansions + (long)(tb - src_base_ibox)*nstored;
        double *full = malloc(sizeof(double)*nfull);
        for (int j = 0; j < nfull; ++j) {
            double acc = 0;
            for (int k = 0; k < nstored; ++k) acc += P[(long)j*nstored + k]*sc[k];
            full[j] = acc;
        }
        long t0 = box_tgt_starts[tb], t1 = t0 + box_tgt_counts[tb];
        for (long t = t0; t < t1; ++t) {
            double b[3];
            for (int i = 0; i < 3; ++i) b[i] = (tgt[i][t] - centers[i][tb])/rscale;
            double pot = 0, grad[3] = {0, 0, 0};
            for (int j = 0; j < nfull; ++j) {
                const int *mi = full_ids + 3*j;
                double mon = 1;
                for (int i = 0; i < 3; ++i) mon *= pow(b[i], mi[i])/fact(mi[i]);
                pot += full[j]*mon;
                for (int a = 0; a < 3; ++a) {
                    if (mi[a] == 0) continue;
                    double m2 = 1;
                    for (int i = 0; i < 3; ++i) {
                        int e = mi[i] - (i == a);
                        m2 *= pow(b[i], e)/fact(e);
                    }
                    grad[a] += full[j]*m2/rscale;
                }
            }
            if (out_pot) out_pot[t] = scale*pot;
            if (out_grad) for (int a = 0; a < 3; ++a) out_grad[a][t] = scale*grad[a];
        }
        free(full);
    }
}

/* }}} */

/* {{{ M2P (multipole.py:131-157), accumulate */

void orc_m2p(int order, int nstored, const int *stored_ids, long ntgt_boxes,
             const int *target_boxes, const int *src_box_starts, const int *src_box_lists,
             const int *box_tgt_starts, const int *box_tgt_counts,
             const double *const *centers, const double *const *tgt, double rscale,
             double scale, const double *src_expansions, int src_base_ibox,
             double *out_pot, double *const *out_grad)
{
    int M[3] = {order + 1, order + 1, order + 1};
    long nbox = (long)(M[0] + 1)*(M[1] + 1)*(M[2] + 1);
    const int S1 = M[2] + 1, S0 = (M[1] + 1)*S1;
    #pragma omp parallel for schedule(dynamic, 4)
    for (long ib = 0; ib < ntgt_boxes; ++ib) {
        int tb = target_boxes[ib];
        double *Dbox = malloc(sizeof(double)*nbox);
        long t0 = box_tgt_starts[tb], t1 = t0 + box_tgt_counts[tb];
        for (long t = t0; t < t1; ++t) {
            double pot = 0, grad[3] = {0, 0, 0};
            for (int is = src_box_starts[ib]; is < src_box_starts[ib+1]; ++is) {
                int sb = src_box_lists[is];
                double b[3];
                for (int i = 0; i < 3; ++i) b[i] = tgt[i][t] - centers[i][sb];
                lap3d_derivs(b, rscale, order + 1, M, Dbox);
                const double *sc = src_expansions + (long)(sb - src_base_ibox)*nstored;
                for (int k = 0; k < nstored; ++k) {
                    const int *mi = stored_ids + 3*k;
                    pot += sc[k]*Dbox[mi[0]*S0 + mi[1]*S1 + mi[2]];
                    for (int a = 0; a < 3; ++a) {
                        int m2[3] = {mi[0], mi[1], mi[2]}; m2[a] += 1;
                        grad[a] += sc[k]*Dbox[m2[0]*S0 + m2[1]*S1 + m2[2]]/rscale;
                    }
                }
            }
            if (out_pot) out_pot[t] += scale*pot;
            if (out_grad) for (int a = 0; a < 3; ++a) out_grad[a][t] += scale*grad[a];
        }
        free(Dbox);
    }
}

/* }}} */

/* {{{ P2L (local.py:226-271), overwrite */

void orc_p2l(int order, int nstored, const int *stored_ids, long ntgt_boxes,
             const int *target_boxes, const int *src_box_starts, const int *src_box_lists,
             const int *box_src_starts, const int *box_src_counts,
             const double *const *centers, const double *const *src,
             const double *w0, const double *const *wd, double rscale,
             double *tgt_expansions, int tgt_base_ibox)
{
    int M[3] = {order + 1, order + 1, order + 1};
    long nbox = (long)(M[0] + 1)*(M[1] + 1)*(M[2] + 1);
    const int S1 = M[2] + 1, S0 = (M[1] + 1)*S1;
    #pragma omp parallel for schedule(dynamic, 4)
    for (long ib = 0; ib < ntgt_boxes; ++ib) {
        int tb = target_boxes[ib];
        double *Dbox = malloc(sizeof(double)*nbox);
        double *acc = calloc(nstored, sizeof(double));
        for (int is = src_box_starts[ib]; is < src_box_starts[ib+1]; ++is) {
            int sb = src_box_lists[is];
            long s0 = box_src_starts[sb], s1 = s0 + box_src_counts[sb];
            for (long s = s0; s < s1; ++s) {
                double a[3];
                for (int i = 0; i < 3; ++i) a[i] = centers[i][tb] - src[i][s];
                lap3d_derivs(a, rscale, order + 1, M, Dbox);
                for (int k = 0; k < nstored; ++k) {
                    const int *mi = stored_ids + 3*k;
                    double v = w0[s]*Dbox[mi[0]*S0 + mi[1]*S1 + mi[2]];
                    if (wd) for (int ax = 0; ax < 3; ++ax) {
                        int m2[3] = {mi[0], mi[1], mi[2]}; m2[ax] += 1;
                        v += wd[ax][s]*Dbox[m2[0]*S0 + m2[1]*S1 + m2[2]]/rscale;
                    }
                    acc[k] += v;
                }
            }
        }
        double *out = tgt_expansions + (long)(tb - tgt_base_ibox)*nstored;
        for (int k = 0; k < nstored; ++k) out[k] = acc[k];
        free(acc); free(Dbox);
    }
}

/* }}} */

/* {{{ QBX, 2D Laplace line-Taylor by real Taylor-jet arithmetic */

/*
 * coeff_i = d^i/dtau^i post_src(G)(a + tau b) at tau = 0 (local.py:132-170) and
 * result = sum_{i<=p} coeff_i/i! (local.py:172-185), i.e. the sum of the
 * Taylor-jet entries in tau.  G = log|x| = 0.5 log q,
 * q(tau) = |a|^2 + 2 (a.b) tau + |b|^2 tau^2.
 *   srcmode 0: SLP G; 1: DLP -n.grad_a G (kernel.py:2474-2490)
 *   tgt_axis -1: no target derivative; 0/1: d/da_axis applied to the
 *   coefficients (qbx.py:131-138 postprocesses with avec).
 */
static void jet_div(int p, const double *num, const double *den, double *out)
{
    for (int n = 0; n <= p; ++n) {
        double v = num[n];
        for (int k = 0; k < n; ++k) v -= out[k]*den[n - k];
        out[n] = v/den[0];
    }
}

static void jet_mul(int p, const double *x, const double *y, double *out)
{
    for (int n = 0; n <= p; ++n) {
        double v = 0;
        for (int k = 0; k <= n; ++k) v += x[k]*y[n - k];
        out[n] = v;
    }
}

void orc_qbx2d(int order, int srcmode, int tgt_axis, long nt, const double *const *tgt,
               const double *const *ctr, const double *radii, long ns,
               const double *const *src, const double *const *nrm, const double *dens,
               double scale, double *out)
{
    (void)radii;
    int p = order;
    #pragma omp parallel for schedule(dynamic, 16)
    for (long t = 0; t < nt; ++t) {
        double q[64], lq[64], tmp[64], h[2][64], num[64], hh[64];
        double res = 0;
        for (long s = 0; s < ns; ++s) {
            double a[2] = {ctr[0][t] - src[0][s], ctr[1][t] - src[1][s]};
            double b[2] = {tgt[0][t] - ctr[0][t], tgt[1][t] - ctr[1][t]};
            for (int n = 0; n <= p + 2; ++n) q[n] = 0;
            q[0] = a[0]*a[0] + a[1]*a[1];
            q[1] = 2*(a[0]*b[0] + a[1]*b[1]);
            q[2] = b[0]*b[0] + b[1]*b[1];
            /* h_j = (dq/da_j)/q,  dq/da_j = 2 a_j + 2 b_j tau */
            for (int j = 0; j < 2; ++j) {
                for (int n = 0; n <= p; ++n) num[n] = 0;
                num[0] = 2*a[j]; if (p >= 1) num[1] = 2*b[j];
                jet_div(p, num, q, h[j]);
            }
            double v = 0;
            if (srcmode == 0 && tgt_axis < 0) {
                /* log q: n lq_n q_0 = n q_n - sum_{k=1}^{n-1} k lq_k q_{n-k} */
                lq[0] = log(q[0]);
                for (int n = 1; n <= p; ++n) {
                    double val = n*q[n];
                    for (int k = 1; k < n; ++k) val -= k*lq[k]*q[n - k];
                    lq[n] = val/(n*q[0]);
                }
                for (int n = 0; n <= p; ++n) v += 0.5*lq[n];
            } else if (srcmode == 0) {
                for (int n = 0; n <= p; ++n) v += 0.5*h[tgt_axis][n];
            } else if (tgt_axis < 0) {
                for (int n = 0; n <= p; ++n)
                    v += -0.5*(nrm[0][s]*h[0][n] + nrm[1][s]*h[1][n]);
            } else {
                /* d/da_i (dq/da_j / q) = 2 delta_ij / q - h_i h_j */
                int i = tgt_axis;
                double inv[64], one[64];
                for (int n = 0; n <= p; ++n) one[n] = 0;
                one[0] = 1;
                jet_div(p, one, q, inv);
                for (int j = 0; j < 2; ++j) {
                    jet_mul(p, h[i], h[j], tmp);
                    for (int n = 0; n <= p; ++n)
                        hh[n] = 2.0*(i == j)*inv[n] - tmp[n];
                    for (int n = 0; n <= p; ++n) v += -0.5*nrm[j][s]*hh[n];
                }
            }
            res += v*dens[s];
        }
        out[t] = scale*res;
    }
}

/* }}} */

/* vim: set foldmethod=marker: */
