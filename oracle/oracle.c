/*
 * oracle.c -- CPU FP64 restatement of the reference GB energy path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the GPU path,
 * never part of the product.  Plain C99 + OpenMP, built with gcc.
 *
 * Every routine restates the reference algorithm of dacase/nabc src/sff
 * (citations at each function), over the same per-atom lower/upper pair
 * lists the reference uses, with thread-private gradient and sumdeijda
 * copies summed after each parallel loop (as eff.c:1813-1835, :2768-2797).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"
#include "oracle_internal.h"

/* GB-neck maxima table (gbneck.h:17-147), shared with the product only as
 * published data: the numbers are the same physical constants. */
#include "../nabc_amd/csrc/neck_tables.h"

#define KSCALE 0.73


static int orc_threads = 0;
void orc_set_threads(int n) { orc_threads = n; }
int orc_nthreads(void)
{
#ifdef _OPENMP
    return orc_threads > 0 ? orc_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

/* sff.c:1153-1162 (dielc), :1187-1333 (gb 2/5/7/8 parameters) */
orc_t *orc_create(PARMSTRUCT_T *prm, const int *frozen, int gb, double cut,
                  double rgbmax, double kappa, double epsext, double dielc,
                  int dield)
{
    orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
    int n = prm->Natom, i, k, e;
    o->prm = prm;
    o->n = n;
    o->gb = gb;
    o->cut = cut;
    o->rgbmax = rgbmax;
    o->kappa = kappa;
    o->epsext = epsext;
    o->dield = dield;
    o->fsmax = prm->Fsmax;
    o->gboffset = 0.09;
    o->neckscale = 0.0;
    o->fs = (double *)malloc(sizeof(double) * n);
    o->q = (double *)malloc(sizeof(double) * n);
    o->alpha = (double *)calloc(n, sizeof(double));
    o->beta = (double *)calloc(n, sizeof(double));
    o->gamma = (double *)calloc(n, sizeof(double));
    o->neck = (int *)calloc(n, sizeof(int));
    o->frozen = (int *)calloc(n, sizeof(int));
    for (i = 0; i < n; i++) {
        o->fs[i] = prm->Fs[i];
        o->q[i] = prm->Charges[i];
        if (frozen) o->frozen[i] = frozen[i] != 0;
    }
    if (dielc != 1.0) {
        double s = 1.0 / sqrt(dielc);
        for (i = 0; i < n; i++) o->q[i] *= s;
    }
    if (gb == 2) {
        for (i = 0; i < n; i++) { o->alpha[i] = 0.8; o->beta[i] = 0.0; o->gamma[i] = 2.909125; }
    } else if (gb == 5) {
        for (i = 0; i < n; i++) { o->alpha[i] = 1.0; o->beta[i] = 0.8; o->gamma[i] = 4.85; }
    } else if (gb == 7) {
        o->neckscale = 0.361825;
        for (i = 0; i < n; i++) {
            o->alpha[i] = 1.09511284; o->beta[i] = 1.90792938; o->gamma[i] = 2.50798245;
            switch (prm->AtomicNum[i]) {
            case 1: o->fs[i] = 1.09085413633e0; break;
            case 6: o->fs[i] = 4.84353823306e-1; break;
            case 7: o->fs[i] = 7.00147318409e-1; break;
            case 8: o->fs[i] = 1.06557401132e0; break;
            case 16: o->fs[i] = 6.02256336067e-1; break;
            default: o->fs[i] = 5.0e-1;
            }
            o->neck[i] = (int)((prm->Rborn[i] - 1.0) * 20.0 + 0.5);
        }
    } else if (gb == 8) {
        for (i = 0; i < n; i++) {
            switch (prm->AtomicNum[i]) {
            case 1: o->alpha[i] = 0.788440; o->beta[i] = 0.798699; o->gamma[i] = 0.437334; o->fs[i] = 1.425952; break;
            case 6: o->alpha[i] = 0.733756; o->beta[i] = 0.506378; o->gamma[i] = 0.205844; o->fs[i] = 1.058554; break;
            case 7: o->alpha[i] = 0.503364; o->beta[i] = 0.316828; o->gamma[i] = 0.192915; o->fs[i] = 0.733599; break;
            case 8: o->alpha[i] = 0.867814; o->beta[i] = 0.876635; o->gamma[i] = 0.387882; o->fs[i] = 1.061039; break;
            case 15: o->alpha[i] = 1.0; o->beta[i] = 0.8; o->gamma[i] = 4.85; o->fs[i] = 0.5; break;
            case 16: o->alpha[i] = 0.867814; o->beta[i] = 0.876635; o->gamma[i] = 0.387882; o->fs[i] = -0.703469; break;
            default: o->alpha[i] = 1.0; o->beta[i] = 0.8; o->gamma[i] = 4.85; o->fs[i] = 0.5;
            }
            o->neck[i] = (int)((prm->Rborn[i] - 1.0) * 20.0 + 0.5);
        }
        o->gboffset = 0.195141;
        o->neckscale = 0.826836;
    }
    /* IexclAt (sff.c:1551-1574): 1-based partner numbers, 0 = none */
    o->iexcl = (int **)calloc(n, sizeof(int *));
    for (i = 0, e = 0; i < n; i++) {
        o->iexcl[i] = (int *)malloc(sizeof(int) * (prm->Iblo[i] + 1));
        for (k = 0; k < prm->Iblo[i]; k++) o->iexcl[i][k] = prm->ExclAt[e + k];
        e += prm->Iblo[i];
    }
    /* N14pearlist (sff.c:1521-1537) */
    o->n14 = (int **)calloc(n, sizeof(int *));
    o->scee14 = (double **)calloc(n, sizeof(double *));
    o->scnb14 = (double **)calloc(n, sizeof(double *));
    for (i = 0, e = 0; i < n; i++) {
        int c = prm->N14pairs[i];
        if (c > 0) {
            o->n14[i] = (int *)malloc(sizeof(int) * c);
            o->scee14[i] = (double *)malloc(sizeof(double) * c);
            o->scnb14[i] = (double *)malloc(sizeof(double) * c);
            for (k = 0; k < c; k++) {
                o->n14[i][k] = prm->N14pairlist[e + k];
                o->scee14[i][k] = prm->N14sceelist[e + k];
                o->scnb14[i][k] = prm->N14scnblist[e + k];
            }
            e += c;
        }
    }
    o->lp = (int *)calloc(n, sizeof(int));
    o->up = (int *)calloc(n, sizeof(int));
    o->list = (int **)calloc(n, sizeof(int *));
    return o;
}

void orc_destroy(orc_t *o)
{
    int i;
    if (!o) return;
    for (i = 0; i < o->n; i++) {
        free(o->list[i]);
        free(o->iexcl[i]);
        free(o->n14[i]);
        free(o->scee14[i]);
        free(o->scnb14[i]);
    }
    free(o->list); free(o->lp); free(o->up); free(o->iexcl); free(o->n14);
    free(o->scee14); free(o->scnb14);
    free(o->fs); free(o->q); free(o->alpha); free(o->beta); free(o->gamma);
    free(o->neck); free(o->frozen);
    orc_hcp_free(o);
    free(o);
}

void orc_gbparams(orc_t *o, double *gboffset, double *neckscale, double *fs,
                  double *alpha, double *beta, double *gamma, int *neckidx,
                  double *charges)
{
    int i;
    if (gboffset) *gboffset = o->gboffset;
    if (neckscale) *neckscale = o->neckscale;
    for (i = 0; i < o->n; i++) {
        if (fs) fs[i] = o->fs[i];
        if (alpha) alpha[i] = o->alpha[i];
        if (beta) beta[i] = o->beta[i];
        if (gamma) gamma[i] = o->gamma[i];
        if (neckidx) neckidx[i] = o->neck[i];
        if (charges) charges[i] = o->q[i];
    }
}

/*
 * Pair list (nblist.c:813-1240): j is on i's list iff j != i, not both
 * frozen, and r^2 < cut^2 (strict) with r^2 = xij^2 + yij^2 + zij^2,
 * xij = x_i - x_j (nblist.c:748-765).  Lower entries (j < i) then upper
 * (j > i), each ascending (heap-sorted, nblist.c:1146-1147, :1189-1196).
 * A k-d tree only prunes; scanning every j gives the same set.
 */
long long orc_nblist(orc_t *o, const double *x)
{
    const int n = o->n;
    const double cut2 = o->cut * o->cut;
    long long tot = 0;
    int i;
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : tot) num_threads(orc_nthreads())
    for (i = 0; i < n; i++) {
        int j, cnt = 0, lo = 0;
        int *buf = (int *)malloc(sizeof(int) * (n > 1 ? n : 1));
        for (j = 0; j < n; j++) {
            double xij, yij, zij, r2;
            if (j == i) continue;
            if (o->frozen[i] && o->frozen[j]) continue;
            xij = x[3 * i] - x[3 * j];
            yij = x[3 * i + 1] - x[3 * j + 1];
            zij = x[3 * i + 2] - x[3 * j + 2];
            r2 = xij * xij + yij * yij + zij * zij;
            if (r2 < cut2) {
                buf[cnt++] = j;
                if (j < i) lo++;
            }
        }
        free(o->list[i]);
        o->list[i] = (int *)realloc(buf, sizeof(int) * (cnt > 0 ? cnt : 1));
        o->lp[i] = lo;
        o->up[i] = cnt - lo;
        tot += cnt;
    }
    o->have_list = 1;
    return tot;
}

void orc_list_counts(orc_t *o, int *lpears, int *upears)
{
    int i;
    for (i = 0; i < o->n; i++) {
        lpears[i] = o->lp[i];
        upears[i] = o->up[i];
    }
}

void orc_list_copy(orc_t *o, int *flat)
{
    int i, k, e = 0;
    for (i = 0; i < o->n; i++)
        for (k = 0; k < o->lp[i] + o->up[i]; k++) flat[e++] = o->list[i][k];
}

#define TA 0.33333333333333333333
#define TB 0.4
#define TC 0.42857142857142857143
#define TD 0.44444444444444444444
#define TDD 0.45454545454545454545
#define TE 1.33333333333333333333
#define TF 2.4
#define TG 3.42857142857142857143
#define TH 4.44444444444444444444
#define THH 5.45454545454545454545

/* egb(), eff.c:1363-2417 (gbsa = 0, dim = 3) */
double orc_egb(orc_t *o, const double *x, double *f, double *enb, double *eel,
               double *reff_out, double *sumdeijda_out)
{
    if (o->hcp) return orc_egb_hcp(o, x, f, enb, eel, reff_out, sumdeijda_out);
    const int n = o->n, gb = o->gb;
    const PARMSTRUCT_T *prm = o->prm;
    const double rgbmax = o->rgbmax, gboffset = o->gboffset;
    const double rgbmax1i = 1.0 / rgbmax, rgbmax2i = rgbmax1i * rgbmax1i;
    const double rgbmaxpsmax2 = (rgbmax + o->fsmax) * (rgbmax + o->fsmax);
    const double kappa = o->kappa, diel_ext = o->epsext;
    const int nt = orc_nthreads();
    double *reff = (double *)calloc(n, sizeof(double));
    double *psi = (double *)calloc(n, sizeof(double));
    double *sumdeijda = (double *)calloc((size_t)n * nt, sizeof(double));
    double *gbuf = (double *)calloc((size_t)3 * n * nt, sizeof(double));
    double epol = 0.0, elec = 0.0, evdw = 0.0;
    int i;

    /* pass 1: effective radii (eff.c:1566-1733) */
#pragma omp parallel for schedule(dynamic, 8) num_threads(nt)
    for (i = 0; i < n; i++) {
        double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
        double ri = prm->Rborn[i] - gboffset, ri1i = 1.0 / ri, sumi = 0.0;
        int k;
        for (k = 0; k < o->lp[i] + o->up[i]; k++) {
            int j = o->list[i][k];
            double xij = xi - x[3 * j], yij = yi - x[3 * j + 1], zij = zi - x[3 * j + 2];
            double r2 = xij * xij + yij * yij + zij * zij, dij1i, dij, sj, sj2;
            if (r2 > rgbmaxpsmax2) continue;
            dij1i = 1.0 / sqrt(r2);
            dij = r2 * dij1i;
            sj = o->fs[j] * (prm->Rborn[j] - gboffset);
            sj2 = sj * sj;
            if (dij > rgbmax + sj) continue;
            if (dij > rgbmax - sj) {
                double uij = 1. / (dij - sj);
                sumi -= 0.125 * dij1i * (1.0 + 2.0 * dij * uij + rgbmax2i * (r2 - 4.0 * rgbmax * dij - sj2) +
                                         2.0 * log((dij - sj) * rgbmax1i));
            } else if (dij > 4.0 * sj) {
                double dij2i = dij1i * dij1i, tmpsd = sj2 * dij2i;
                double dumbo = TA + tmpsd * (TB + tmpsd * (TC + tmpsd * (TD + tmpsd * TDD)));
                sumi -= sj * tmpsd * dij2i * dumbo;
            } else if (dij > ri + sj) {
                sumi -= 0.5 * (sj / (r2 - sj2) + 0.5 * dij1i * log((dij - sj) / (dij + sj)));
            } else if (dij > fabs(ri - sj)) {
                double theta = 0.5 * ri1i * dij1i * (r2 + ri * ri - sj2);
                double uij = 1. / (dij + sj);
                sumi -= 0.25 * (ri1i * (2. - theta) - uij + dij1i * log(ri * uij));
            } else if (ri < sj) {
                sumi -= 0.5 * (sj / (r2 - sj2) + 2. * ri1i + 0.5 * dij1i * log((sj - dij) / (sj + dij)));
            }
            if (gb == 7 || gb == 8) {
                if (dij < prm->Rborn[i] + prm->Rborn[j] + NECK_CUT) {
                    double mdist = dij - NECK_POS[o->neck[i] * NECK_N + o->neck[j]];
                    double mdist2 = mdist * mdist, mdist6 = mdist2 * mdist2 * mdist2;
                    double neck = NECK_VAL[o->neck[i] * NECK_N + o->neck[j]] / (1.0 + mdist2 + 0.3 * mdist6);
                    sumi -= o->neckscale * neck;
                }
            }
        }
        if (gb == 1) {
            reff[i] = 1.0 / (ri1i + sumi);
            if (reff[i] < 0.0) reff[i] = 30.0;
        } else {
            psi[i] = -ri * sumi;
            reff[i] = 1.0 / (ri1i - tanh((o->alpha[i] - o->beta[i] * psi[i] + o->gamma[i] * psi[i] * psi[i]) * psi[i]) /
                                        prm->Rborn[i]);
        }
    }

    /* pass 2: GB, Coulomb, LJ energies and sumdeijda (eff.c:1768-2066) */
#pragma omp parallel num_threads(nt) reduction(+ : epol, elec, evdw)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        double *sd = sumdeijda + (size_t)n * tid, *g = gbuf + (size_t)3 * n * tid;
        int *iexw = (int *)malloc(sizeof(int) * (n + 1)) + 1; /* index -1 .. n-1 */
        int ii;
        for (ii = -1; ii < n; ii++) iexw[ii] = -1;
#pragma omp for schedule(dynamic, 8)
        for (ii = 0; ii < n; ii++) {
            double ri = reff[ii], qi = o->q[ii];
            double xi, yi, zi, daix = 0, daiy = 0, daiz = 0;
            int k, iaci;
            if (!o->frozen[ii]) {
                double expmkf = exp(-KSCALE * kappa * ri) / diel_ext;
                double dielfac = 1.0 - expmkf;
                double qi2h = 0.5 * qi * qi, qid2h = qi2h * dielfac;
                epol += -qid2h / ri;
                sd[ii] += qid2h - KSCALE * kappa * qi2h * expmkf * ri;
            }
            if (o->up[ii] <= 0) continue;
            xi = x[3 * ii]; yi = x[3 * ii + 1]; zi = x[3 * ii + 2];
            iaci = prm->Ntypes * (prm->Iac[ii] - 1);
            for (k = 0; k < prm->Iblo[ii]; k++) iexw[o->iexcl[ii][k] - 1] = ii;
            for (k = o->lp[ii]; k < o->lp[ii] + o->up[ii]; k++) {
                int j = o->list[ii][k], ic;
                double xij = xi - x[3 * j], yij = yi - x[3 * j + 1], zij = zi - x[3 * j + 2];
                double r2 = xij * xij + yij * yij + zij * zij;
                double qiqj = qi * o->q[j], rj = reff[j], rb2 = ri * rj;
                double efac = exp(-r2 / (4.0 * rb2));
                double fgbi = 1.0 / sqrt(r2 + rb2 * efac);
                double fgbk = -kappa * KSCALE / fgbi;
                double expmkf = exp(fgbk) / diel_ext, dielfac = 1.0 - expmkf;
                double temp4, temp6, de, temp5;
                epol += -qiqj * dielfac * fgbi;
                temp4 = fgbi * fgbi * fgbi;
                temp6 = qiqj * temp4 * (dielfac + fgbk * expmkf);
                de = temp6 * (1.0 - 0.25 * efac);
                temp5 = 0.5 * efac * temp6 * (rb2 + 0.25 * r2);
                sd[ii] += ri * temp5;
                sd[j] += rj * temp5;
                if (iexw[j] != ii) {
                    double rinv = 1. / sqrt(r2), r2inv = rinv * rinv, e = qiqj * rinv;
                    elec += e;
                    de -= e * r2inv;
                    ic = prm->Cno[iaci + prm->Iac[j] - 1] - 1;
                    if (ic >= 0) {
                        double r6inv = r2inv * r2inv * r2inv;
                        double f6 = prm->Cn2[ic] * r6inv, f12 = prm->Cn1[ic] * r6inv * r6inv;
                        evdw += f12 - f6;
                        de -= (12. * f12 - 6. * f6) * r2inv;
                    }
                }
                daix += de * xij; daiy += de * yij; daiz += de * zij;
                g[3 * j] -= de * xij; g[3 * j + 1] -= de * yij; g[3 * j + 2] -= de * zij;
            }
            g[3 * ii] += daix; g[3 * ii + 1] += daiy; g[3 * ii + 2] += daiz;
        }
        free(iexw - 1);
    }
    /* reduce the sumdeijda copies (eff.c:2077-2098) */
    {
        int t;
        for (t = 1; t < nt; t++)
            for (i = 0; i < n; i++) sumdeijda[i] += sumdeijda[(size_t)n * t + i];
    }

    /* pass 3: chain rule (eff.c:2131-2405) */
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        double *g = gbuf + (size_t)3 * n * tid;
        int ii;
#pragma omp for schedule(dynamic, 8)
        for (ii = 0; ii < n; ii++) {
            double xi, yi, zi, ri, ri1i, sumda, daix = 0, daiy = 0, daiz = 0;
            int k;
            if (o->frozen[ii] || o->lp[ii] + o->up[ii] <= 0) continue;
            xi = x[3 * ii]; yi = x[3 * ii + 1]; zi = x[3 * ii + 2];
            ri = prm->Rborn[ii] - gboffset;
            ri1i = 1. / ri;
            sumda = sumdeijda[ii];
            if (gb > 1) {
                double a = o->alpha[ii], b = o->beta[ii], c = o->gamma[ii], p = psi[ii];
                double thi = tanh((a - b * p + c * p * p) * p);
                sumda *= (a - 2.0 * b * p + 3.0 * c * p * p) * (1.0 - thi * thi) * ri / prm->Rborn[ii];
            }
            for (k = 0; k < o->lp[ii] + o->up[ii]; k++) {
                int j = o->list[ii][k];
                double xij = xi - x[3 * j], yij = yi - x[3 * j + 1], zij = zi - x[3 * j + 2];
                double r2 = xij * xij + yij * yij + zij * zij;
                double dij1i, dij2i, dij, sj, sj2, datmp, temp1, dij3i;
                if (r2 > rgbmaxpsmax2) continue;
                dij1i = 1.0 / sqrt(r2);
                dij2i = dij1i * dij1i;
                dij = r2 * dij1i;
                sj = o->fs[j] * (prm->Rborn[j] - gboffset);
                sj2 = sj * sj;
                if (dij > rgbmax + sj) continue;
                if (dij > rgbmax - sj) {
                    temp1 = 1. / (dij - sj);
                    dij3i = dij1i * dij2i;
                    datmp = 0.125 * dij3i * ((r2 + sj2) * (temp1 * temp1 - rgbmax2i) - 2.0 * log(rgbmax * temp1));
                } else if (dij > 4.0 * sj) {
                    double tmpsd = sj2 * dij2i;
                    double dumbo = TE + tmpsd * (TF + tmpsd * (TG + tmpsd * (TH + tmpsd * THH)));
                    datmp = tmpsd * sj * dij2i * dij2i * dumbo;
                } else if (dij > ri + sj) {
                    temp1 = 1. / (r2 - sj2);
                    datmp = temp1 * sj * (-0.5 * dij2i + temp1) + 0.25 * dij1i * dij2i * log((dij - sj) / (dij + sj));
                } else if (dij > fabs(ri - sj)) {
                    temp1 = 1. / (dij + sj);
                    dij3i = dij2i * dij1i;
                    datmp = -0.25 * (-0.5 * (r2 - ri * ri + sj2) * dij3i * ri1i * ri1i +
                                     dij1i * temp1 * (temp1 - dij1i) - dij3i * log(ri * temp1));
                } else if (ri < sj) {
                    temp1 = 1. / (r2 - sj2);
                    datmp = -0.5 * (sj * dij2i * temp1 - 2. * sj * temp1 * temp1 -
                                    0.5 * dij2i * dij1i * log((sj - dij) / (sj + dij)));
                } else {
                    datmp = 0.;
                }
                if (gb == 7 || gb == 8) {
                    if (dij < prm->Rborn[ii] + prm->Rborn[j] + NECK_CUT) {
                        double mdist = dij - NECK_POS[o->neck[ii] * NECK_N + o->neck[j]];
                        double mdist2 = mdist * mdist, mdist6 = mdist2 * mdist2 * mdist2;
                        temp1 = 1.0 + mdist2 + 0.3 * mdist6;
                        temp1 = temp1 * temp1 * dij;
                        datmp += ((2.0 * mdist + 1.8 * mdist2 * mdist2 * mdist) *
                                  NECK_VAL[o->neck[ii] * NECK_N + o->neck[j]] * o->neckscale) / temp1;
                    }
                }
                daix += xij * datmp; daiy += yij * datmp; daiz += zij * datmp;
                datmp *= sumda;
                g[3 * j] += xij * datmp; g[3 * j + 1] += yij * datmp; g[3 * j + 2] += zij * datmp;
            }
            g[3 * ii] -= sumda * daix; g[3 * ii + 1] -= sumda * daiy; g[3 * ii + 2] -= sumda * daiz;
        }
    }
    {
        int t;
        for (t = 0; t < nt; t++)
            for (i = 0; i < 3 * n; i++) f[i] += gbuf[(size_t)3 * n * t + i];
    }
    if (reff_out) memcpy(reff_out, reff, sizeof(double) * n);
    if (sumdeijda_out) memcpy(sumdeijda_out, sumdeijda, sizeof(double) * n);
    free(reff); free(psi); free(sumdeijda); free(gbuf);
    *enb = evdw;
    *eel = elec;
    return epol;
}

/* ebond, eff.c:82-168 */
static double o_ebond(int nb, const int *a1, const int *a2, const int *at, const double *Rk,
                      const double *Req, const double *x, double *f)
{
    double e_bond = 0.0;
    int i;
    for (i = 0; i < nb; i++) {
        int i1 = a1[i], i2 = a2[i], ty = at[i] - 1;
        double rx = x[i1] - x[i2], ry = x[i1 + 1] - x[i2 + 1], rz = x[i1 + 2] - x[i2 + 2];
        double s = sqrt(rx * rx + ry * ry + rz * rz), r = 2.0 / s;
        double db = s - Req[ty], df = Rk[ty] * db;
        e_bond += df * db;
        df *= r;
        f[i1] += rx * df; f[i1 + 1] += ry * df; f[i1 + 2] += rz * df;
        f[i2] -= rx * df; f[i2 + 1] -= ry * df; f[i2 + 2] -= rz * df;
    }
    return e_bond;
}

/* eangl, eff.c:168-316 */
static double o_eangl(int na, const int *a1, const int *a2, const int *a3, const int *at,
                      const double *Tk, const double *Teq, const double *x, double *f)
{
    double e_theta = 0.0;
    int i;
    for (i = 0; i < na; i++) {
        int i1 = a1[i], i2 = a2[i], i3 = a3[i], ty = at[i] - 1;
        double dxi = x[i1] - x[i2], dyi = x[i1 + 1] - x[i2 + 1], dzi = x[i1 + 2] - x[i2 + 2];
        double dxj = x[i3] - x[i2], dyj = x[i3 + 1] - x[i2 + 1], dzj = x[i3 + 2] - x[i2 + 2];
        double ri = sqrt(dxi * dxi + dyi * dyi + dzi * dzi), rj = sqrt(dxj * dxj + dyj * dyj + dzj * dzj);
        double rir = 1. / ri, rjr = 1. / rj;
        double dxir = dxi * rir, dyir = dyi * rir, dzir = dzi * rir;
        double dxjr = dxj * rjr, dyjr = dyj * rjr, dzjr = dzj * rjr;
        double cst = dxir * dxjr + dyir * dyjr + dzir * dzjr, at_, da, df;
        if (cst > 1.0) cst = 1.0;
        if (cst < -1.0) cst = -1.0;
        at_ = acos(cst);
        da = at_ - Teq[ty];
        df = da * Tk[ty];
        e_theta += df * da;
        df = df + df;
        at_ = sin(at_);
        if (at_ > 0 && at_ < 1.e-3) at_ = 1.e-3;
        else if (at_ < 0 && at_ > -1.e-3) at_ = -1.e-3;
        df = -df / at_;
        {
            double xt = df * rir * (dxjr - cst * dxir), dxt = df * rjr * (dxir - cst * dxjr);
            double yt = df * rir * (dyjr - cst * dyir), dyt = df * rjr * (dyir - cst * dyjr);
            double zt = df * rir * (dzjr - cst * dzir), dzt = df * rjr * (dzir - cst * dzjr);
            f[i1] += xt; f[i3] += dxt; f[i2] -= xt + dxt;
            f[i1 + 1] += yt; f[i3 + 1] += dyt; f[i2 + 1] -= yt + dyt;
            f[i1 + 2] += zt; f[i3 + 2] += dzt; f[i2 + 2] -= zt + dzt;
        }
    }
    return e_theta;
}

/* ephi, eff.c:316-627 (3D), including multi-term series (Pn < 0) */
static double o_ephi(int nphi, const int *a1, const int *a2, const int *a3, const int *a4,
                     const int *at, const double *Pk, const double *Pn, const double *Phase,
                     const double *x, double *f)
{
    double e_tors = 0.0;
    int i;
    for (i = 0; i < nphi; i++) {
        int i1 = a1[i], i2 = a2[i], i3 = abs(a3[i]), i4 = abs(a4[i]), ty = at[i] - 1;
        double ax = x[i2] - x[i1], ay = x[i2 + 1] - x[i1 + 1], az = x[i2 + 2] - x[i1 + 2];
        double bx = x[i3] - x[i2], by = x[i3 + 1] - x[i2 + 1], bz = x[i3 + 2] - x[i2 + 2];
        double cx = x[i4] - x[i3], cy = x[i4 + 1] - x[i3 + 1], cz = x[i4 + 2] - x[i3 + 2];
        double ab = ax * bx + ay * by + az * bz, bc = bx * cx + by * cy + bz * cz;
        double ac = ax * cx + ay * cy + az * cz, aa = ax * ax + ay * ay + az * az;
        double bb = bx * bx + by * by + bz * bz, cc = cx * cx + cy * cy + cz * cz;
        double uu = aa * bb - ab * ab, vv = bb * cc - bc * bc, uv = ab * bc - ac * bb;
        double den = uu == 0.0 ? 0.0 : 1.0 / sqrt(uu * vv);
        double co = uv * den, co1 = 0.5 * co * den;
        double a0x = -bc * bx + bb * cx, a0y = -bc * by + bb * cy, a0z = -bc * bz + bb * cz;
        double b0x = ab * cx + bc * ax - 2. * ac * bx, b0y = ab * cy + bc * ay - 2. * ac * by;
        double b0z = ab * cz + bc * az - 2. * ac * bz;
        double c0x = ab * bx - bb * ax, c0y = ab * by - bb * ay, c0z = ab * bz - bb * az;
        double a1x = 2. * uu * (-cc * bx + bc * cx), a1y = 2. * uu * (-cc * by + bc * cy);
        double a1z = 2. * uu * (-cc * bz + bc * cz);
        double b1x = 2. * uu * (bb * cx - bc * bx), b1y = 2. * uu * (bb * cy - bc * by);
        double b1z = 2. * uu * (bb * cz - bc * bz);
        double a2x = -2. * vv * (bb * ax - ab * bx), a2y = -2. * vv * (bb * ay - ab * by);
        double a2z = -2. * vv * (bb * az - ab * bz);
        double b2x = 2. * vv * (aa * bx - ab * ax), b2y = 2. * vv * (aa * by - ab * ay);
        double b2z = 2. * vv * (aa * bz - ab * az);
        double dd1x = (a0x - a2x * co1) * den, dd1y = (a0y - a2y * co1) * den, dd1z = (a0z - a2z * co1) * den;
        double dd2x = (-a0x - b0x - (a1x - a2x - b2x) * co1) * den;
        double dd2y = (-a0y - b0y - (a1y - a2y - b2y) * co1) * den;
        double dd2z = (-a0z - b0z - (a1z - a2z - b2z) * co1) * den;
        double dd3x = (b0x - c0x - (-a1x - b1x + b2x) * co1) * den;
        double dd3y = (b0y - c0y - (-a1y - b1y + b2y) * co1) * den;
        double dd3z = (b0z - c0z - (-a1z - b1z + b2z) * co1) * den;
        double dd4x = (c0x - b1x * co1) * den, dd4y = (c0y - b1y * co1) * den, dd4z = (c0z - b1z * co1) * den;
        for (;;) {
            double e, df;
            if (fabs(Phase[ty] - 3.142) > 0.01 && fabs(Phase[ty]) > 0.01) {
                double phi, ux, uy, uz, vx, vy, vz, dx1, dy1, dz1, delta, yy;
                /* the clamp persists into later terms of a series (eff.c:493-494) */
                co = co > 1.0 ? 1.0 : co;
                co = co < -1.0 ? -1.0 : co;
                phi = acos(co);
                ux = ay * bz - az * by; uy = az * bx - ax * bz; uz = ax * by - ay * bx;
                vx = by * cz - bz * cy; vy = bz * cx - bx * cz; vz = bx * cy - by * cx;
                dx1 = uy * vz - uz * vy; dy1 = uz * vx - ux * vz; dz1 = ux * vy - uy * vx;
                dx1 = dx1 * bx + dy1 * by + dz1 * bz;
                if (dx1 < 0.0) phi = -phi;
                delta = Pn[ty] * phi - Phase[ty];
                e = Pk[ty] * (1.0 + cos(delta));
                yy = sin(phi);
                if (fabs(yy) > 0.001) df = Pk[ty] * Pn[ty] * sin(delta) / yy;
                else if (phi > -1.0 && phi < 1.) df = Pk[ty] * Pn[ty] * sin(delta) * 1000.;
                else df = -Pk[ty] * Pn[ty] * sin(delta) * 1000.;
            } else {
                double phase = fabs(Phase[ty] - 3.142) < 0.01 ? -1.0 : 1.0, k = Pk[ty], cosq;
                switch ((int)fabs(Pn[ty])) {
                case 1: e = k * (1.0 + phase * co); df = phase * k; break;
                case 2: e = k * (1.0 + phase * (2. * co * co - 1.)); df = phase * k * 4. * co; break;
                case 3:
                    cosq = co * co;
                    e = k * (1.0 + phase * co * (4. * cosq - 3.));
                    df = phase * k * (12. * cosq - 3.);
                    break;
                case 4:
                    cosq = co * co;
                    e = k * (1.0 + phase * (8. * cosq * (cosq - 1.) + 1.));
                    df = phase * k * co * (32. * cosq - 16.);
                    break;
                case 5:
                    cosq = co * co;
                    e = k * (1.0 + phase * (16. * cosq * cosq * co - 20. * cosq * co + 5. * co));
                    df = phase * k * (80. * cosq * cosq - 60. * cosq + 5.);
                    break;
                case 6:
                    cosq = co * co;
                    e = k * (1.0 + phase * (32. * cosq * cosq * cosq - 48. * cosq * cosq + 18. * cosq - 1.));
                    df = phase * k * co * (192. * cosq * cosq - 192. * cosq + 36.);
                    break;
                default:
                    e = 0.0;
                    df = 0.0;
                }
            }
            e_tors += e;
            f[i1] += df * dd1x; f[i1 + 1] += df * dd1y; f[i1 + 2] += df * dd1z;
            f[i2] += df * dd2x; f[i2 + 1] += df * dd2y; f[i2 + 2] += df * dd2z;
            f[i3] += df * dd3x; f[i3 + 1] += df * dd3y; f[i3 + 2] += df * dd3z;
            f[i4] += df * dd4x; f[i4 + 1] += df * dd4y; f[i4 + 2] += df * dd4z;
            if (Pn[ty] < 0.0) ty++;
            else break;
        }
    }
    return e_tors;
}

/* nbond() with N14 = 1 (eff.c:896-1331), dield 1 (diel=C) or 0 (diel=R) */
static void o_nb14(orc_t *o, const double *x, double *f, double *enb, double *eel)
{
    const PARMSTRUCT_T *prm = o->prm;
    double evdw = 0.0, elec = 0.0;
    int i;
    for (i = 0; i < o->n - 1; i++) {
        int npr = prm->N14pairs[i], k, iaci;
        double cgi = o->q[i], dumx = 0, dumy = 0, dumz = 0;
        if (npr <= 0) continue;
        iaci = prm->Ntypes * (prm->Iac[i] - 1);
        for (k = 0; k < npr; k++) {
            int j = o->n14[i][k], ic;
            double enbfaci = o->scnb14[i][k], eelfaci = o->scee14[i][k];
            double cgj = o->q[j] * eelfaci;
            double xij = x[3 * i] - x[3 * j], yij = x[3 * i + 1] - x[3 * j + 1], zij = x[3 * i + 2] - x[3 * j + 2];
            double r2 = xij * xij + yij * yij + zij * zij, r2inv = 1.0 / r2, r = sqrt(r2), rinv = r * r2inv;
            double rs, df2, df;
            if (o->dield == 0) {
                rs = cgi * cgj * r2inv;
                df2 = -2.0 * rs;
            } else {
                rs = cgi * cgj * rinv;
                df2 = -rs;
            }
            elec += rs;
            ic = prm->Cno[iaci + prm->Iac[j] - 1];
            if (ic > 0 || enbfaci != 1.0) {
                double r6, f1, f2;
                if (ic > 0) ic--;
                else {
                    int ibig = prm->Iac[i] > prm->Iac[j] ? prm->Iac[i] : prm->Iac[j];
                    int isml = prm->Iac[i] > prm->Iac[j] ? prm->Iac[j] : prm->Iac[i];
                    ic = ibig * (ibig - 1) / 2 + isml - 1;
                }
                r6 = r2inv * r2inv * r2inv;
                f2 = prm->Cn2[ic] * r6;
                f1 = prm->Cn1[ic] * r6 * r6;
                evdw += (f1 - f2) * enbfaci;
                df = (df2 + (6.0 * f2 - 12.0 * f1) * enbfaci) * rinv;
            } else {
                int hc = -ic - 1;
                double r10 = r2inv * r2inv * r2inv * r2inv * r2inv;
                double f2 = prm->HB10[hc] * r10, f1 = prm->HB12[hc] * r10 * r2inv;
                evdw += (f1 - f2) * enbfaci;
                df = (df2 + (10.0 * f2 - 12.0 * f1) * enbfaci) * rinv;
            }
            df *= rinv;
            dumx += df * xij; dumy += df * yij; dumz += df * zij;
            f[3 * j] -= df * xij; f[3 * j + 1] -= df * yij; f[3 * j + 2] -= df * zij;
        }
        f[3 * i] += dumx; f[3 * i + 1] += dumy; f[3 * i + 2] += dumz;
    }
    *enb = evdw;
    *eel = elec;
}

/* mme34(), eff.c:2442-2947 */
double orc_mme(orc_t *o, const double *x, double *f, int rebuild, double *ene, double *frms)
{
    const PARMSTRUCT_T *p = o->prm;
    const int n = o->n;
    double e[13], enb, eel, enb14, eel14, s = 0.0;
    int i;
    memset(e, 0, sizeof(e));
    /* HCP uses no pair list (eff.c:2560-2561) */
    if (!o->hcp && (!o->have_list || rebuild)) orc_nblist(o, x);
    memset(f, 0, sizeof(double) * 3 * n);
    e[3] = o_ebond(p->Nbonh, p->BondHAt1, p->BondHAt2, p->BondHNum, p->Rk, p->Req, x, f) +
           o_ebond(p->Mbona, p->BondAt1, p->BondAt2, p->BondNum, p->Rk, p->Req, x, f);
    e[4] = o_eangl(p->Ntheth, p->AngleHAt1, p->AngleHAt2, p->AngleHAt3, p->AngleHNum, p->Tk, p->Teq, x, f) +
           o_eangl(p->Ntheta, p->AngleAt1, p->AngleAt2, p->AngleAt3, p->AngleNum, p->Tk, p->Teq, x, f);
    e[5] = o_ephi(p->Nphih, p->DihHAt1, p->DihHAt2, p->DihHAt3, p->DihHAt4, p->DihHNum, p->Pk, p->Pn, p->Phase, x, f) +
           o_ephi(p->Mphia, p->DihAt1, p->DihAt2, p->DihAt3, p->DihAt4, p->DihNum, p->Pk, p->Pn, p->Phase, x, f);
    o_nb14(o, x, f, &enb14, &eel14);
    e[7] = enb14;
    e[8] = eel14;
    e[10] = orc_egb(o, x, f, &enb, &eel, NULL, NULL);
    e[1] = enb;
    e[2] = eel;
    for (i = 0; i < n; i++)
        if (o->frozen[i]) f[3 * i] = f[3 * i + 1] = f[3 * i + 2] = 0.0;
    for (i = 0; i < 3 * n; i++) s += f[i] * f[i];
    if (frms) *frms = sqrt(s / (3 * n));
    e[0] = 0.0;
    for (i = 1; i <= 12; i++) e[0] += e[i];
    if (ene) memcpy(ene, e, sizeof(e));
    return e[0];
}

/* ---- mme-compatible callback with a call log (for minimizer traces) ---- */

static orc_t *cur = NULL;
static int cur_nsnb = 10;
static int log_n = 0, log_cap = 0;
static double *log_e = NULL;
static int *log_it = NULL;

void orc_set_current(orc_t *o, int nsnb)
{
    cur = o;
    cur_nsnb = nsnb > 0 ? nsnb : 1;
    if (o) o->have_list = 0;
}

double orc_mme_cb(double *x, double *f, int *iter)
{
    double ene[13], frms;
    int rebuild = (*iter > 0 && *iter % cur_nsnb == 0) || *iter == -4;
    if (*iter == -3) return 0.0;
    orc_mme(cur, x, f, rebuild, ene, &frms);
    if (log_n == log_cap) {
        log_cap = log_cap ? 2 * log_cap : 64;
        log_e = (double *)realloc(log_e, sizeof(double) * 14 * log_cap);
        log_it = (int *)realloc(log_it, sizeof(int) * log_cap);
    }
    memcpy(log_e + 14 * log_n, ene, sizeof(ene));
    log_e[14 * log_n + 13] = frms;
    log_it[log_n++] = *iter;
    return ene[0];
}

int orc_log_count(void) { return log_n; }

void orc_log_get(int k, int *iter, double *ene13, double *frms)
{
    if (k < 0 || k >= log_n) return;
    if (iter) *iter = log_it[k];
    if (ene13) memcpy(ene13, log_e + 14 * k, sizeof(double) * 13);
    if (frms) *frms = log_e[14 * k + 13];
}

void orc_log_clear(void) { log_n = 0; }

/* ------------------------------------------------------------------ md() */

/*
 * Langevin noise: N(0,1) of (seed, step, coordinate index) from a counter
 * hash (splitmix64 finaliser of seed, step and index; Box-Muller of two
 * 53-bit uniforms).  This is the noise stream the product documents in
 * place of rand2.c's serial generator (the one deviation of its md());
 * restated here from that description so the oracle's dynamics draw the
 * same deviates: initial velocities at step -1, the step-0 accelerations
 * at step 0, step k's kick at step k.
 */
static unsigned long long o_mix(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double orc_gauss(unsigned long long seed, long long step, long long idx)
{
    unsigned long long h = o_mix(seed ^ o_mix((unsigned long long)step * 0x100000001B3ULL ^ o_mix((unsigned long long)idx)));
    unsigned long long h2 = o_mix(h);
    double u1 = ((double)(h >> 11) + 1.0) * (1.0 / 9007199254740993.0);
    double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/*
 * md(), sff.c:2071-2366, over orc_mme with mme34's rebuild cadence (the
 * list is built at the first call and when the step number is a positive
 * multiple of nsnb, eff.c:2564).  Constants sff.c:2140-2155; initial
 * velocities sff.c:2165-2188 (zerov, tempi > 0, or the given v with the
 * reference's v*v*minv kinetic sum); step-0 accelerations sff.c:2199-2215;
 * the loop sff.c:2223-2352: Berendsen scaling (gamma_ln = 0) or the
 * Langevin half kicks, the vlimit clamp, frozen atoms held.  No RATTLE, no
 * centre-of-mass removal (nscm = 0, the reference's default, sff.c:76).
 * rows (16 per step, may be NULL): ene[0..12], frms, ekin, temperature.
 */
int orc_md(orc_t *o, int maxstep, double *x, double *v, const double *masses, const orc_md_opts *mo,
           double *rows)
{
    const int n = 3 * o->n, dim = 3;
    double *minv = (double *)calloc(n, sizeof(double));
    double *accel = (double *)calloc(n, sizeof(double));
    double *sqrmass = (double *)calloc(n, sizeof(double));
    double *f = (double *)calloc(n, sizeof(double));
    double dtx = mo->dt * 20.455, dt5 = 0.5 * dtx, rndf, ekin0, dttp, gammai, c_implic, c_explic, c_ave, sdfac;
    double ekin = 0.0, tscal, ene[13], frms;
    int i, nstep, nfrozen = 0, nsnb = mo->nsnb > 0 ? mo->nsnb : 1;
    if (!minv || !accel || !sqrmass || !f) {
        free(minv); free(accel); free(sqrmass); free(f);
        return -1;
    }
    for (i = 0; i < o->n; i++) nfrozen += o->frozen[i] != 0;
    rndf = n - 3 * nfrozen;
    ekin0 = mo->boltz2 * rndf * mo->temp0;
    dttp = mo->dt / mo->tautp;
    gammai = mo->gamma_ln / 20.455;
    c_implic = 1.0 / (1.0 + gammai * dt5);
    c_explic = 1.0 - gammai * dt5;
    c_ave = 1.0 + gammai * dt5;
    sdfac = sqrt(4.0 * gammai * mo->boltz2 * mo->temp0 / dtx);
    for (i = 0; i < n; i++) minv[i] = 1.0 / masses[i / dim];
    if (mo->zerov) {
        for (i = 0; i < n; i++) v[i] = 0.0;
    } else if (mo->tempi > 0.0) {
        for (i = 0; i < n; i++) {
            if (o->frozen[i / dim]) {
                v[i] = 0.0;
            } else {
                v[i] = sqrt(2.0 * mo->boltz2 * mo->tempi * minv[i]) * orc_gauss(mo->seed, -1, i);
                ekin += v[i] * v[i] / minv[i];
            }
        }
        ekin *= 0.5;
    } else {
        for (i = 0; i < n; i++) ekin += v[i] * v[i] * minv[i];
        ekin *= 0.5;
    }
    orc_mme(o, x, f, 0, ene, &frms); /* nstep = 0: builds the list when there is none */
    for (i = 0; i < n; i++) {
        if (gammai == 0.0) {
            accel[i] = -f[i] * minv[i] * dt5;
        } else if (o->frozen[i / dim]) {
            accel[i] = 0.0;
        } else {
            sqrmass[i] = 1.0 / sqrt(minv[i]);
            accel[i] = (-f[i] + sdfac * sqrmass[i] * orc_gauss(mo->seed, 0, i)) * minv[i] * dt5;
        }
    }
    for (nstep = 1; nstep <= maxstep; nstep++) {
        tscal = ekin > 0.01 ? sqrt(1.0 + dttp * (ekin0 / ekin - 1.0)) : 1.0;
        ekin = 0.0;
        for (i = 0; i < n; i++) {
            if (gammai == 0.0) v[i] = (v[i] + accel[i]) * tscal;
            else v[i] = c_explic * v[i] + accel[i];
            v[i] = v[i] > mo->vlimit ? mo->vlimit : v[i];
            v[i] = v[i] < -mo->vlimit ? -mo->vlimit : v[i];
            x[i] += v[i] * dtx;
        }
        orc_mme(o, x, f, nstep % nsnb == 0, ene, &frms);
        for (i = 0; i < n; i++) {
            if (gammai == 0.0) {
                accel[i] = -f[i] * minv[i] * dt5;
                v[i] = (v[i] + accel[i]) * tscal;
            } else if (!o->frozen[i / dim]) {
                accel[i] = (-f[i] + sdfac * sqrmass[i] * orc_gauss(mo->seed, nstep, i)) * minv[i] * dt5;
                v[i] = (v[i] + accel[i]) * c_implic;
            }
        }
        for (i = 0; i < n; i++) ekin += v[i] * v[i] / minv[i];
        ekin *= 0.5;
        if (gammai > 0.0) ekin *= c_ave;
        if (rows) {
            double *r = rows + 16 * (size_t)(nstep - 1);
            memcpy(r, ene, sizeof(ene));
            r[13] = frms;
            r[14] = ekin;
            r[15] = ekin / (mo->boltz2 * rndf);
        }
    }
    free(minv); free(accel); free(sqrmass); free(f);
    return 0;
}
