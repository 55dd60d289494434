/*
 * oracle_hcp.c -- CPU restatement of the reference's hierarchical charge
 * partitioning GB (HCP / HCPO), the checker of the product's HCP kernels.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): never part of the product.
 *
 * Reference (dacase/nabc src/sff):
 *   set-up        sff.c:1335-1340 (hcp 4 -> two charges + HCPO),
 *                 sff.c:1607-1679 (symmetric excluded-atom lists),
 *                 hcp_gb.c:45-124 egb_hcp_rbondi (component radii)
 *   charges       hcp.c:337-540 calc_approx_q, hcp.c:543-840
 *                 calc_approx_q_hcpo with OPCA_K2 (hcp.c:150-335) and
 *                 eigen_decomposition (hcp.c:20-82), q2idx (hcp.c:115-146)
 *   Born radii    hcp_gb.c:179-430 egb_hcp_atom_reff, hcp_gb.c:139-160
 *                 (far components: Taylor term only)
 *   averaging     hcp_gb.c:745-858 egb_hcp_reff, hcp_gb.c:872-1026
 *                 egb_hcpo_reff_avg
 *   energy/grad   hcp_gb.c:1795-2112 egb_hcp_nbond with
 *                 egb_calc_energy_approx/atom (:1353-1562) and
 *                 egb_calc_deriv_approx/atom (:1567-1789)
 *   driver        hcp_gb.c:2166-2240 egb_hcp
 *
 * Arithmetic follows the reference's types: the component radii and the
 * HCPO multipoles and optimal point charges are single precision there
 * (float temporaries and arrays), everything else double.  The per-atom
 * loops run in parallel over atoms i; each atom's own sums are formed in
 * the reference's order, gradient scatters go to thread-private copies.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_internal.h"

#define KSCALE 0.73
#define MIN_Q 0.00001

/* Taylor coefficients of the descreening integral and of its derivative */
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

struct orc_hcp {
    int hcp;       /* mm_options value: 1, 2 or 4 */
    int nq;        /* charges per component: 1 or 2 */
    int opca;      /* 1: HCPO optimal point charges (hcp 4) */
    double h1, h2, h3, dhcp;
    int nres, nstrand, ncomplex;
    int *ipres;    /* nres + 1, 0-based */
    int *rfirst;   /* per strand: first residue, nstrand + 1 */
    int *sfirst;   /* per complex: first strand, ncomplex + 1 */
    /* symmetric excluded-atom lists (0-based partners) */
    int *ex_ptr, *ex_j;
    /* per level: centres (3 per component), charges (4 per charge:
     * x, y, z, q), radii and Born radii (one per charge) */
    double *xc[3], *qc[3], *rc[3], *reffc[3];
    int ncomp[3];
    double *reff0, *psi;
};

static int omp_thread(void)
{
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

/* hcp.c:115-146; hcp 3 cannot be selected (mm_options.l:34 takes 0124) */
static int qidx(const struct orc_hcp *H, double q)
{
    return (H->nq == 2 && q > 0.0) ? 1 : 0;
}

static double dist2(double xi, double yi, double zi, const double *p)
{
    double xij = xi - p[0], yij = yi - p[1], zij = zi - p[2];
    return xij * xij + yij * yij + zij * zij;
}

void orc_hcp_free(orc_t *o)
{
    struct orc_hcp *H = o->hcp;
    int l;
    if (!H) return;
    free(H->ipres); free(H->rfirst); free(H->sfirst); free(H->ex_ptr); free(H->ex_j);
    for (l = 0; l < 3; l++) { free(H->xc[l]); free(H->qc[l]); free(H->rc[l]); free(H->reffc[l]); }
    free(H->reff0); free(H->psi);
    free(H);
    o->hcp = NULL;
}

/* egb_hcp_rbondi, hcp_gb.c:45-124: a component's radius is the cube root of
 * the summed cubes of its atoms' scaled radii (each cube in single
 * precision, as the float temporary there), rolled up level by level */
static void hcp_radii(orc_t *o, struct orc_hcp *H)
{
    const PARMSTRUCT_T *p = o->prm;
    const int nq = H->nq;
    int c, s, r, a, q;
    for (c = 0; c < H->ncomplex; c++) {
        for (q = 0; q < nq; q++) H->rc[2][c * nq + q] = 0.0;
        for (s = H->sfirst[c]; s < H->sfirst[c + 1]; s++) {
            for (q = 0; q < nq; q++) H->rc[1][s * nq + q] = 0.0;
            for (r = H->rfirst[s]; r < H->rfirst[s + 1]; r++) {
                for (q = 0; q < nq; q++) H->rc[0][r * nq + q] = 0.0;
                for (a = H->ipres[r]; a < H->ipres[r + 1]; a++) {
                    float t = (float)((p->Rborn[a] - o->gboffset) * o->fs[a]);
                    float t3 = t * t * t;
                    H->rc[0][r * nq + qidx(H, o->q[a])] += (double)t3;
                }
                for (q = 0; q < nq; q++) {
                    H->rc[1][s * nq + q] += H->rc[0][r * nq + q];
                    H->rc[0][r * nq + q] = cbrt(H->rc[0][r * nq + q]);
                }
            }
            for (q = 0; q < nq; q++) {
                H->rc[2][c * nq + q] += H->rc[1][s * nq + q];
                H->rc[1][s * nq + q] = cbrt(H->rc[1][s * nq + q]);
            }
        }
        for (q = 0; q < nq; q++) H->rc[2][c * nq + q] = cbrt(H->rc[2][c * nq + q]);
    }
}

int orc_set_hcp(orc_t *o, const orc_hcp_opts *h)
{
    const PARMSTRUCT_T *p = o->prm;
    struct orc_hcp *H;
    int i, k, e, l, n = o->n;
    orc_hcp_free(o);
    if (!h || h->hcp == 0) return 0;
    if (h->hcp != 1 && h->hcp != 2 && h->hcp != 4) return -1;
    if (h->nres < 1 || h->nstrand < 1 || h->ncomplex < 1 || !h->ipres || !h->ipstrand || !h->ipcomplex) return -1;
    if (h->ipres[0] != 0) return -1;
    for (i = 0; i < h->nres; i++)
        if (h->ipres[i + 1] < h->ipres[i]) return -1;
    if (h->ipres[h->nres] != n) return -1;
    for (i = 0; i < h->nstrand; i++)
        if (h->ipstrand[i] < 1 || h->ipstrand[i] > h->nres || (i > 0 && h->ipstrand[i] <= h->ipstrand[i - 1])) return -1;
    for (i = 0; i < h->ncomplex; i++)
        if (h->ipcomplex[i] < 1 || h->ipcomplex[i] > h->nstrand || (i > 0 && h->ipcomplex[i] <= h->ipcomplex[i - 1]))
            return -1;
    if (h->ipstrand[0] != 1 || h->ipcomplex[0] != 1) return -1;
    H = (struct orc_hcp *)calloc(1, sizeof(*H));
    H->hcp = h->hcp;
    H->nq = h->hcp == 1 ? 1 : 2;
    H->opca = h->hcp == 4;
    H->h1 = h->h1; H->h2 = h->h2; H->h3 = h->h3; H->dhcp = h->dhcp;
    H->nres = h->nres; H->nstrand = h->nstrand; H->ncomplex = h->ncomplex;
    H->ipres = (int *)malloc(sizeof(int) * (h->nres + 1));
    memcpy(H->ipres, h->ipres, sizeof(int) * (h->nres + 1));
    H->rfirst = (int *)malloc(sizeof(int) * (h->nstrand + 1));
    for (i = 0; i < h->nstrand; i++) H->rfirst[i] = h->ipstrand[i] - 1;
    H->rfirst[h->nstrand] = h->nres;
    H->sfirst = (int *)malloc(sizeof(int) * (h->ncomplex + 1));
    for (i = 0; i < h->ncomplex; i++) H->sfirst[i] = h->ipcomplex[i] - 1;
    H->sfirst[h->ncomplex] = h->nstrand;
    H->ncomp[0] = h->nres; H->ncomp[1] = h->nstrand; H->ncomp[2] = h->ncomplex;
    for (l = 0; l < 3; l++) {
        H->xc[l] = (double *)calloc(3 * (size_t)H->ncomp[l], sizeof(double));
        H->qc[l] = (double *)calloc(4 * (size_t)H->ncomp[l] * H->nq, sizeof(double));
        H->rc[l] = (double *)calloc((size_t)H->ncomp[l] * H->nq, sizeof(double));
        H->reffc[l] = (double *)calloc((size_t)H->ncomp[l] * H->nq, sizeof(double));
    }
    H->reff0 = (double *)calloc(n, sizeof(double));
    H->psi = (double *)calloc(n, sizeof(double));
    /* sff.c:1644-1679: every excluded pair listed under both atoms */
    H->ex_ptr = (int *)calloc(n + 1, sizeof(int));
    for (i = 0, e = 0; i < n; i++)
        for (k = 0; k < p->Iblo[i]; k++, e++)
            if (p->ExclAt[e] > 0) { H->ex_ptr[i + 1]++; H->ex_ptr[p->ExclAt[e]]++; }
    for (i = 0; i < n; i++) H->ex_ptr[i + 1] += H->ex_ptr[i];
    H->ex_j = (int *)malloc(sizeof(int) * (H->ex_ptr[n] + 1));
    {
        int *fill = (int *)calloc(n, sizeof(int));
        for (i = 0, e = 0; i < n; i++)
            for (k = 0; k < p->Iblo[i]; k++, e++) {
                int j = p->ExclAt[e] - 1;
                if (j < 0) continue;
                H->ex_j[H->ex_ptr[i] + fill[i]++] = j;
                H->ex_j[H->ex_ptr[j] + fill[j]++] = i;
            }
        free(fill);
    }
    o->hcp = H;
    hcp_radii(o, H);
    return 0;
}

int orc_hcp_components(orc_t *o, int level, double *centres, double *charges, double *radii, double *reff)
{
    const struct orc_hcp *H = o->hcp;
    int l = level - 1, m;
    if (!H || l < 0 || l > 2) return -1;
    m = H->ncomp[l];
    if (centres) memcpy(centres, H->xc[l], sizeof(double) * 3 * m);
    if (charges) memcpy(charges, H->qc[l], sizeof(double) * 4 * m * H->nq);
    if (radii) memcpy(radii, H->rc[l], sizeof(double) * m * H->nq);
    if (reff) memcpy(reff, H->reffc[l], sizeof(double) * m * H->nq);
    return m;
}

/* ------------------------------------------- approximate charges, hcp 1/2 */

/* the position of a component charge: its centre of charge, or the
 * component's geometric centre when the charge is (nearly) zero */
static void place_charge(double *qv, const double *centre)
{
    double q = qv[3];
    if (fabs(q) > MIN_Q) {
        qv[0] /= q; qv[1] /= q; qv[2] /= q;
    } else {
        qv[0] = centre[0]; qv[1] = centre[1]; qv[2] = centre[2];
    }
}

/* hcp.c:337-540: geometric centres and centre-of-charge point charges per
 * charge group; level sums are formed from the unnormalised sums of the
 * level below, in component order */
static void approx_charges(orc_t *o, struct orc_hcp *H, const double *x)
{
    const int nq = H->nq;
    double *x1 = H->xc[0], *x2 = H->xc[1], *x3 = H->xc[2];
    double *q1 = H->qc[0], *q2 = H->qc[1], *q3 = H->qc[2];
    int c, s, r, a, m, k;
    for (c = 0; c < H->ncomplex; c++) {
        int n3 = 0;
        for (k = 0; k < 3; k++) x3[3 * c + k] = 0.0;
        for (m = 0; m < 4 * nq; m++) q3[4 * nq * c + m] = 0.0;
        for (s = H->sfirst[c]; s < H->sfirst[c + 1]; s++) {
            int n2 = 0;
            for (k = 0; k < 3; k++) x2[3 * s + k] = 0.0;
            for (m = 0; m < 4 * nq; m++) q2[4 * nq * s + m] = 0.0;
            for (r = H->rfirst[s]; r < H->rfirst[s + 1]; r++) {
                int n1 = 0;
                for (k = 0; k < 3; k++) x1[3 * r + k] = 0.0;
                for (m = 0; m < 4 * nq; m++) q1[4 * nq * r + m] = 0.0;
                for (a = H->ipres[r]; a < H->ipres[r + 1]; a++) {
                    double q = o->q[a], *qv = q1 + 4 * (nq * r + qidx(H, q));
                    n1++;
                    for (k = 0; k < 3; k++) {
                        x1[3 * r + k] += x[3 * a + k];
                        qv[k] += x[3 * a + k] * q;
                    }
                    qv[3] += q;
                }
                n2 += n1;
                for (k = 0; k < 3; k++) x2[3 * s + k] += x1[3 * r + k];
                if (n1 > 0)
                    for (k = 0; k < 3; k++) x1[3 * r + k] /= n1;
                for (m = 0; m < nq; m++) {
                    for (k = 0; k < 4; k++) q2[4 * (nq * s + m) + k] += q1[4 * (nq * r + m) + k];
                    place_charge(q1 + 4 * (nq * r + m), x1 + 3 * r);
                }
            }
            n3 += n2;
            for (k = 0; k < 3; k++) x3[3 * c + k] += x2[3 * s + k];
            if (n2 > 0)
                for (k = 0; k < 3; k++) x2[3 * s + k] /= n2;
            for (m = 0; m < nq; m++) {
                for (k = 0; k < 4; k++) q3[4 * (nq * c + m) + k] += q2[4 * (nq * s + m) + k];
                place_charge(q2 + 4 * (nq * s + m), x2 + 3 * s);
            }
        }
        if (n3 > 0)
            for (k = 0; k < 3; k++) x3[3 * c + k] /= n3;
        for (m = 0; m < nq; m++) place_charge(q3 + 4 * (nq * c + m), x3 + 3 * c);
    }
}

/* ------------------------------------------------ HCPO (hcp 4): OPCA k=2 */

/* hcp.c:20-82: eigenvalues of a symmetric 3x3 matrix from the roots of its
 * characteristic cubic (trigonometric form, with the reference's 8-digit
 * pi), eigenvectors as products of cofactor combinations; only the first
 * eigenpair is used */
static void sym3_eigen(double m[3][3], double vec[3][3], double val[3])
{
    const double pi = 3.14159265;
    double k0, k1, k2, pp, qq, s, arc, alpha;
    int i;
    k0 = m[0][0] * m[1][1] * m[2][2] + 2.0 * m[0][1] * m[1][2] * m[0][2] - m[0][0] * m[1][2] * m[1][2] -
         m[1][1] * m[0][2] * m[0][2] - m[2][2] * m[0][1] * m[0][1];
    k1 = -m[0][0] * m[1][1] - m[1][1] * m[2][2] - m[0][0] * m[2][2] + m[0][1] * m[0][1] + m[1][2] * m[1][2] +
         m[0][2] * m[0][2];
    k2 = m[0][0] + m[1][1] + m[2][2];
    pp = -k1 - k2 * k2 / 3.0;
    qq = -k0 - 2.0 * k2 * k2 * k2 / 27.0 - k1 * k2 / 3.0;
    s = -2.0 * sqrt(-pp / 3.0);
    arc = -4.0 * qq / (s * s * s);
    alpha = acos(arc);
    val[0] = s * cos(alpha / 3.0) + k2 / 3.0;
    val[1] = -s * cos((alpha + pi) / 3.0) + k2 / 3.0;
    val[2] = -s * cos((alpha - pi) / 3.0) + k2 / 3.0;
    for (i = 0; i < 3; i++) {
        double A = m[0][0] - val[i], B = m[1][1] - val[i], C = m[2][2] - val[i];
        double AA = m[0][2] * m[0][1] - A * m[1][2];
        double BB = m[0][1] * m[1][2] - B * m[0][2];
        double CC = m[0][2] * m[1][2] - C * m[0][1];
        double mag;
        vec[0][i] = BB * CC;
        vec[1][i] = CC * AA;
        vec[2][i] = AA * BB;
        mag = sqrt(vec[0][i] * vec[0][i] + vec[1][i] * vec[1][i] + vec[2][i] * vec[2][i]);
        if (mag != 0) {
            vec[0][i] = vec[0][i] / mag;
            vec[1][i] = vec[1][i] / mag;
            vec[2][i] = vec[2][i] / mag;
        }
    }
}

/* the smooth radial limit of an optimal point charge's offset: unchanged
 * below R1, R2 beyond R2, a cubic blend in between (hcp.c:300-333) */
static float opca_limit(float dis, float R1, float R2)
{
    float R1s = R1 * R1, R2s = R2 * R2;
    float denom = (R1s - R2s) * (R1s - R2s) * (R1s - R2s);
    float dis2 = dis * dis;
    if (dis < R1) return dis;
    if (dis > R2) return R2;
    return (float)((R1s - dis2) * (R1s - dis2) * (R1s + 2.0 * dis * dis - 3.0 * R2s) / denom * (R2 - dis) + dis);
}

/*
 * hcp.c:150-335 (OPCA_K2): two point charges reproducing the monopole and
 * the dipole of a charge distribution (and, for a charged one, lying along
 * the principal axis of its quadrupole about the centre of charge).
 * Charged (|q| > 0.6 e): q1 = q d0 / (rbar^2 q + d0), q2 = q - q1, at
 * +-rbar along the first eigenvector with zero dipole about the centre of
 * charge, the pair taken on the side away from the geometric centre.
 * Neutral: +-|p| / dr at +-p / 2q about the dipole centre dk.  Then each
 * offset from the geometric centre is limited by opca_limit.  Q is
 * modified (as in the reference) in the charged case.
 */
static void opca2(float qtotal, const float p[3], float Q[3][3], float rmax, const float rcog[3], float *qbar1,
                  float rbar1[3], float *qbar2, float rbar2[3], float R1, float R2, float coef2)
{
    int i, j;
    *qbar1 = 0.0f; *qbar2 = 0.0f;
    for (i = 0; i < 3; i++) { rbar1[i] = 0.0f; rbar2[i] = 0.0f; }
    if (fabs(qtotal) > 0.6 * 18.2223) {
        float rcoq[3], rcoq2, prcoq, cod[3], cdotq, rbar;
        const float coef = 1.5f;
        double A[3][3], V[3][3], d[3];
        rcoq[0] = p[0] / qtotal;
        rcoq[1] = p[1] / qtotal;
        rcoq[2] = p[2] / qtotal;
        rcoq2 = rcoq[0] * rcoq[0] + rcoq[1] * rcoq[1] + rcoq[2] * rcoq[2];
        prcoq = p[0] * rcoq[0] + p[1] * rcoq[1] + p[2] * rcoq[2];
        /* quadrupole moved to the centre of charge */
        for (i = 0; i < 3; i++)
            for (j = i; j < 3; j++)
                Q[i][j] = (float)(0.5 * (2.0 * Q[i][j] - (3.0 * p[i] * rcoq[j] + 3.0 * rcoq[i] * p[j]) +
                                         qtotal * (3.0 * rcoq[i] * rcoq[j])));
        Q[1][0] = Q[0][1];
        Q[2][0] = Q[0][2];
        Q[2][1] = Q[1][2];
        for (i = 0; i < 3; i++) Q[i][i] = (float)(Q[i][i] + prcoq - 0.5 * qtotal * rcoq2);
        for (i = 0; i < 3; i++)
            for (j = 0; j < 3; j++) A[i][j] = Q[i][j];
        sym3_eigen(A, V, d);
        rbar = coef * rmax;
        if (((rbar * rbar * qtotal) + d[0]) == 0.0) rbar = 0.0f;
        *qbar1 = (float)(qtotal * d[0] / ((rbar * rbar * qtotal) + d[0]));
        *qbar2 = qtotal - *qbar1;
        for (i = 0; i < 3; i++) rbar1[i] = (float)(rbar * V[i][0]);
        if (*qbar2 != 0.0f) {
            for (i = 0; i < 3; i++) rbar2[i] = -(*qbar1) / (*qbar2) * rbar1[i];
        } else {
            for (i = 0; i < 3; i++) rbar2[i] = 0.0f;
        }
        for (i = 0; i < 3; i++) cod[i] = (float)((rbar1[i] + rbar2[i]) / 2.0);
        cdotq = cod[0] * (rcoq[0] - rcog[0]) + cod[1] * (rcoq[1] - rcog[1]) + cod[2] * (rcoq[2] - rcog[2]);
        for (i = 0; i < 3; i++) {
            if (cdotq > 0) {
                rbar1[i] = -rbar1[i];
                rbar2[i] = -rbar2[i];
            }
            rbar1[i] = rbar1[i] + rcoq[i] - rcog[i];
            rbar2[i] = rbar2[i] + rcoq[i] - rcog[i];
        }
    } else {
        float sumqpp = 0.0f, mpsq, dk[3], sumqp, dr, q;
        for (i = 0; i < 3; i++)
            for (j = 0; j < 3; j++) sumqpp = sumqpp + Q[i][j] * p[i] * p[j];
        mpsq = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
        for (i = 0; i < 3; i++) {
            dk[i] = 0.0f;
            sumqp = 0.0f;
            for (j = 0; j < 3; j++) sumqp = sumqp + Q[i][j] * p[j];
            dk[i] = (float)(dk[i] + 2.0 / (3.0 * mpsq) * (sumqp - (sumqpp / (4.0 * mpsq) * p[i])));
        }
        dr = coef2 * rmax;
        q = (float)(sqrt(pow(p[0], 2) + pow(p[1], 2) + pow(p[2], 2)) / dr);
        *qbar1 = q;
        *qbar2 = -q;
        for (i = 0; i < 3; i++) {
            rbar1[i] = (float)(+p[i] / (2.0 * q) + dk[i] - rcog[i]);
            rbar2[i] = (float)(-p[i] / (2.0 * q) + dk[i] - rcog[i]);
        }
    }
    {
        float a1 = (float)sqrt(rbar1[0] * rbar1[0] + rbar1[1] * rbar1[1] + rbar1[2] * rbar1[2]);
        float a2 = (float)sqrt(rbar2[0] * rbar2[0] + rbar2[1] * rbar2[1] + rbar2[2] * rbar2[2]);
        float m1 = opca_limit(a1, R1, R2), m2 = opca_limit(a2, R1, R2);
        for (i = 0; i < 3; i++) rbar1[i] = m1 / a1 * rbar1[i] + rcog[i];
        for (i = 0; i < 3; i++) rbar2[i] = m2 / a2 * rbar2[i] + rcog[i];
    }
}

/* single-precision monopole, dipole and quadrupole of a distribution */
typedef struct {
    float qt, p[3], Q[3][3];
} mpole;

static void mpole_zero(mpole *m) { memset(m, 0, sizeof(*m)); }

static void mpole_add_atom(mpole *m, const double *xa, double q)
{
    float r2 = (float)(xa[0] * xa[0] + xa[1] * xa[1] + xa[2] * xa[2]);
    double qh = 0.5 * q;
    m->qt = (float)(m->qt + q);
    m->p[0] = (float)(m->p[0] + xa[0] * q);
    m->p[1] = (float)(m->p[1] + xa[1] * q);
    m->p[2] = (float)(m->p[2] + xa[2] * q);
    m->Q[0][0] = (float)(m->Q[0][0] + qh * (3.0 * xa[0] * xa[0] - r2));
    m->Q[1][1] = (float)(m->Q[1][1] + qh * (3.0 * xa[1] * xa[1] - r2));
    m->Q[2][2] = (float)(m->Q[2][2] + qh * (3.0 * xa[2] * xa[2] - r2));
    m->Q[0][1] = (float)(m->Q[0][1] + qh * (3.0 * xa[0] * xa[1]));
    m->Q[0][2] = (float)(m->Q[0][2] + qh * (3.0 * xa[0] * xa[2]));
    m->Q[1][2] = (float)(m->Q[1][2] + qh * (3.0 * xa[1] * xa[2]));
}

static void mpole_sym(mpole *m)
{
    m->Q[1][0] = m->Q[0][1];
    m->Q[2][0] = m->Q[0][2];
    m->Q[2][1] = m->Q[1][2];
}

/* the level above accumulates the level below's moments (upper triangle) */
static void mpole_roll(mpole *up, const mpole *m)
{
    int i, j;
    up->qt += m->qt;
    for (i = 0; i < 3; i++) up->p[i] += m->p[i];
    for (i = 0; i < 3; i++)
        for (j = i; j < 3; j++) up->Q[i][j] += m->Q[i][j];
    mpole_sym(up);
}

/* largest distance (single precision) from the centre over n points */
static float spread(const double *pts, int stride, int n, const float rcog[3])
{
    float best = 0.0f;
    int k;
    for (k = 0; k < n; k++) {
        const double *x = pts + (size_t)stride * k;
        float d2 = (float)((x[0] - rcog[0]) * (x[0] - rcog[0]) + (x[1] - rcog[1]) * (x[1] - rcog[1]) +
                           (x[2] - rcog[2]) * (x[2] - rcog[2]));
        if (d2 > best) best = d2;
    }
    return (float)sqrt(best);
}

static void opca_component(mpole *m, const double *centre, float rmax, float r1f, float r2f, float coef2,
                           double *qv /* 2 x (x, y, z, q) */)
{
    float rcog[3], qb1, qb2, rb1[3], rb2[3];
    int k;
    for (k = 0; k < 3; k++) rcog[k] = (float)centre[k];
    opca2(m->qt, m->p, m->Q, rmax, rcog, &qb1, rb1, &qb2, rb2, r1f * rmax, r2f * rmax, coef2);
    for (k = 0; k < 3; k++) { qv[k] = rb1[k]; qv[4 + k] = rb2[k]; }
    qv[3] = qb1;
    qv[7] = qb2;
}

/* hcp.c:543-840 (calc_approx_q_hcpo) with gb > 0: limits 0.8 / 0.9 of each
 * component's spread at every level */
static void approx_charges_opca(orc_t *o, struct orc_hcp *H, const double *x)
{
    const float lim1 = 0.8f, lim2 = 0.9f, coef2 = (float)H->dhcp;
    double *x1 = H->xc[0], *x2 = H->xc[1], *x3 = H->xc[2];
    int c, s, r, a, k;
    for (c = 0; c < H->ncomplex; c++) {
        int n3 = 0;
        mpole m3;
        float rcog[3];
        mpole_zero(&m3);
        for (k = 0; k < 3; k++) x3[3 * c + k] = 0.0;
        for (s = H->sfirst[c]; s < H->sfirst[c + 1]; s++) {
            int n2 = 0;
            mpole m2;
            mpole_zero(&m2);
            for (k = 0; k < 3; k++) x2[3 * s + k] = 0.0;
            for (r = H->rfirst[s]; r < H->rfirst[s + 1]; r++) {
                int n1 = 0;
                mpole m1;
                mpole_zero(&m1);
                for (k = 0; k < 3; k++) x1[3 * r + k] = 0.0;
                for (a = H->ipres[r]; a < H->ipres[r + 1]; a++) {
                    n1++;
                    for (k = 0; k < 3; k++) x1[3 * r + k] += x[3 * a + k];
                    mpole_add_atom(&m1, x + 3 * a, o->q[a]);
                }
                mpole_sym(&m1);
                mpole_roll(&m2, &m1);
                n2 += n1;
                for (k = 0; k < 3; k++) x2[3 * s + k] += x1[3 * r + k];
                if (n1 > 0)
                    for (k = 0; k < 3; k++) x1[3 * r + k] /= n1;
                for (k = 0; k < 3; k++) rcog[k] = (float)x1[3 * r + k];
                opca_component(&m1, x1 + 3 * r, spread(x + 3 * H->ipres[r], 3, H->ipres[r + 1] - H->ipres[r], rcog),
                               lim1, lim2, coef2, H->qc[0] + 8 * r);
            }
            n3 += n2;
            for (k = 0; k < 3; k++) x3[3 * c + k] += x2[3 * s + k];
            if (n2 > 0)
                for (k = 0; k < 3; k++) x2[3 * s + k] /= n2;
            mpole_roll(&m3, &m2);
            for (k = 0; k < 3; k++) rcog[k] = (float)x2[3 * s + k];
            opca_component(&m2, x2 + 3 * s,
                           spread(x1 + 3 * H->rfirst[s], 3, H->rfirst[s + 1] - H->rfirst[s], rcog), lim1, lim2,
                           coef2, H->qc[1] + 8 * s);
        }
        if (n3 > 0)
            for (k = 0; k < 3; k++) x3[3 * c + k] /= n3;
        for (k = 0; k < 3; k++) rcog[k] = (float)x3[3 * c + k];
        opca_component(&m3, x3 + 3 * c, spread(x2 + 3 * H->sfirst[c], 3, H->sfirst[c + 1] - H->sfirst[c], rcog),
                       lim1, lim2, coef2, H->qc[2] + 8 * c);
    }
}

/* ---------------------------------------------------------- Born radii */

/* far component: the Taylor term of the descreening integral only
 * (hcp_gb.c:139-160) */
static double far_descreen(double r2, double sj)
{
    double dij2i = 1.0 / r2, sj2 = sj * sj, tmpsd = sj2 * dij2i;
    double dumbo = TA + tmpsd * (TB + tmpsd * (TC + tmpsd * (TD + tmpsd * TDD)));
    return sj * tmpsd * dij2i * dumbo;
}

/* an atom's descreening of atom i (hcp_gb.c:345-390): no rgbmax cut-off
 * and no neck term on this path */
static double near_descreen(double r2, double ri, double ri1i, double sj)
{
    double dij1i = 1.0 / sqrt(r2), dij = r2 * dij1i, sj2 = sj * sj;
    if (dij > 4.0 * sj) {
        double dij2i = dij1i * dij1i, tmpsd = sj2 * dij2i;
        double dumbo = TA + tmpsd * (TB + tmpsd * (TC + tmpsd * (TD + tmpsd * TDD)));
        return sj * tmpsd * dij2i * dumbo;
    } else if (dij > ri + sj) {
        return 0.5 * (sj / (r2 - sj2) + 0.5 * dij1i * log((dij - sj) / (dij + sj)));
    } else if (dij > fabs(ri - sj)) {
        double theta = 0.5 * ri1i * dij1i * (r2 + ri * ri - sj2), uij = 1. / (dij + sj);
        return 0.25 * (ri1i * (2. - theta) - uij + dij1i * log(ri * uij));
    } else if (ri < sj) {
        return 0.5 * (sj / (r2 - sj2) + 2. * ri1i + 0.5 * dij1i * log((sj - dij) / (sj + dij)));
    }
    return 0.0;
}

/*
 * The hierarchical walk of atom i (hcp_gb.c:2000-2071 and its two other
 * copies): complexes, then the strands of a near complex, the residues of
 * a near strand and the atoms (except i) of a near residue; a component
 * farther than its threshold (strictly) is visited as its charges.
 */
enum { V_FAR3, V_FAR2, V_FAR1, V_ATOM };
#define WALK(H, X, ...)                                                                                   \
    do {                                                                                                   \
        int c_, s_, r_, a_;                                                                                \
        const double h3sq_ = (H)->h3 * (H)->h3, h2sq_ = (H)->h2 * (H)->h2, h1sq_ = (H)->h1 * (H)->h1;    \
        for (c_ = 0; c_ < (H)->ncomplex; c_++) {                                                           \
            if (dist2((X)[0], (X)[1], (X)[2], (H)->xc[2] + 3 * c_) > h3sq_) {                              \
                const int kind = V_FAR3, comp = c_;                                                        \
                (void)kind; (void)comp;                                                                    \
                { __VA_ARGS__ }                                                                                   \
                continue;                                                                                  \
            }                                                                                              \
            for (s_ = (H)->sfirst[c_]; s_ < (H)->sfirst[c_ + 1]; s_++) {                                   \
                if (dist2((X)[0], (X)[1], (X)[2], (H)->xc[1] + 3 * s_) > h2sq_) {                          \
                    const int kind = V_FAR2, comp = s_;                                                    \
                    (void)kind; (void)comp;                                                                \
                    { __VA_ARGS__ }                                                                               \
                    continue;                                                                              \
                }                                                                                          \
                for (r_ = (H)->rfirst[s_]; r_ < (H)->rfirst[s_ + 1]; r_++) {                               \
                    if (dist2((X)[0], (X)[1], (X)[2], (H)->xc[0] + 3 * r_) > h1sq_) {                      \
                        const int kind = V_FAR1, comp = r_;                                                \
                        (void)kind; (void)comp;                                                            \
                        { __VA_ARGS__ }                                                                           \
                        continue;                                                                          \
                    }                                                                                      \
                    for (a_ = (H)->ipres[r_]; a_ < (H)->ipres[r_ + 1]; a_++) {                             \
                        const int kind = V_ATOM, comp = a_;                                                \
                        (void)kind; (void)comp;                                                            \
                        { __VA_ARGS__ }                                                                           \
                    }                                                                                      \
                }                                                                                          \
            }                                                                                              \
        }                                                                                                  \
    } while (0)

/* level index (0..2) of a far visit kind */
static int lvl(int kind) { return kind == V_FAR1 ? 0 : (kind == V_FAR2 ? 1 : 2); }

/* hcp_gb.c:179-430 (egb_hcp_atom_reff) */
static void hcp_born(orc_t *o, struct orc_hcp *H, const double *x)
{
    const PARMSTRUCT_T *p = o->prm;
    const int n = o->n, nq = H->nq;
    int i;
#pragma omp parallel for schedule(dynamic, 8) num_threads(orc_nthreads())
    for (i = 0; i < n; i++) {
        const double *xi = x + 3 * i;
        double ri = p->Rborn[i] - o->gboffset, ri1i = 1. / ri, sumi = 0.0;
        WALK(H, xi, {
            if (kind == V_ATOM) {
                if (comp != i)
                    sumi -= near_descreen(dist2(xi[0], xi[1], xi[2], x + 3 * comp), ri, ri1i,
                                          o->fs[comp] * (p->Rborn[comp] - o->gboffset));
            } else {
                const int l = lvl(kind);
                int q;
                for (q = 0; q < nq; q++)
                    sumi -= far_descreen(dist2(xi[0], xi[1], xi[2], H->qc[l] + 4 * (nq * comp + q)),
                                         H->rc[l][nq * comp + q]);
            }
        });
        if (o->gb == 1) {
            H->reff0[i] = 1.0 / (ri1i + sumi);
            if (H->reff0[i] < 0.0) H->reff0[i] = 30.0;
        } else {
            double psi = -ri * sumi;
            H->psi[i] = psi;
            H->reff0[i] = 1.0 / (ri1i - tanh((o->alpha[i] - o->beta[i] * psi + o->gamma[i] * psi * psi) * psi) /
                                             p->Rborn[i]);
        }
    }
}

/* a component charge's Born radius from charge-weighted inverse square
 * roots of its members' radii; 2.3 when that weight is below one */
static double avg_radius(double wsum, double qsum, int guard)
{
    if (guard && fabs(wsum) < 1) return 2.3;
    return (qsum * qsum) / (wsum * wsum);
}

/* hcp_gb.c:745-858 (egb_hcp_reff).  As there, each complex sums over every
 * strand of the molecule (the strand loop is not limited to the complex's
 * own strands), and strand and complex averages have no small-weight
 * guard. */
static void hcp_reff_avg(orc_t *o, struct orc_hcp *H)
{
    const int nq = H->nq;
    int c, s, r, a, m;
    for (c = 0; c < H->ncomplex; c++) {
        double w3[2] = {0, 0}, q3[2] = {0, 0};
        for (s = 0; s < H->nstrand; s++) {
            double w2[2] = {0, 0}, q2[2] = {0, 0};
            for (r = H->rfirst[s]; r < H->rfirst[s + 1]; r++) {
                double w1[2] = {0, 0}, q1[2] = {0, 0};
                for (a = H->ipres[r]; a < H->ipres[r + 1]; a++) {
                    m = qidx(H, o->q[a]);
                    w1[m] += o->q[a] / sqrt(H->reff0[a]);
                    q1[m] += o->q[a];
                }
                for (m = 0; m < nq; m++) {
                    w2[m] += w1[m];
                    q2[m] += q1[m];
                    H->reffc[0][nq * r + m] = avg_radius(w1[m], q1[m], 1);
                }
            }
            for (m = 0; m < nq; m++) {
                w3[m] += w2[m];
                q3[m] += q2[m];
                H->reffc[1][nq * s + m] = avg_radius(w2[m], q2[m], 0);
            }
        }
        for (m = 0; m < nq; m++) H->reffc[2][nq * c + m] = avg_radius(w3[m], q3[m], 0);
    }
}

/* hcp_gb.c:872-1026 (egb_hcpo_reff_avg): weights scaled by the inverse
 * distance to the component charge; a charge's radius goes to the charge
 * of the level above with the same sign.  The residue pair is swapped when
 * its first charge is the larger (the strand and complex swaps are
 * commented out there).  Each complex again runs over every strand. */
static void hcp_reff_avg_opca(orc_t *o, struct orc_hcp *H, const double *x)
{
    const double *qc1 = H->qc[0], *qc2 = H->qc[1], *qc3 = H->qc[2];
    int c, s, r, a, m;
    for (c = 0; c < H->ncomplex; c++) {
        double w3[2] = {0, 0}, q3[2] = {0, 0};
        for (s = 0; s < H->nstrand; s++) {
            double w2[2] = {0, 0}, q2[2] = {0, 0};
            for (r = H->rfirst[s]; r < H->rfirst[s + 1]; r++) {
                double w1[2] = {0, 0}, q1[2] = {0, 0};
                double *rr = H->reffc[0] + 2 * r;
                for (a = H->ipres[r]; a < H->ipres[r + 1]; a++) {
                    double q = o->q[a], d2, ql;
                    m = qidx(H, q);
                    d2 = (x[3 * a] - qc1[4 * (2 * r + m)]) * (x[3 * a] - qc1[4 * (2 * r + m)]) +
                         (x[3 * a + 1] - qc1[4 * (2 * r + m) + 1]) * (x[3 * a + 1] - qc1[4 * (2 * r + m) + 1]) +
                         (x[3 * a + 2] - qc1[4 * (2 * r + m) + 2]) * (x[3 * a + 2] - qc1[4 * (2 * r + m) + 2]);
                    ql = fabs(q) / sqrt(d2);
                    w1[m] += ql / sqrt(H->reff0[a]);
                    q1[m] += ql;
                }
                for (m = 0; m < 2; m++) rr[m] = avg_radius(w1[m], q1[m], 1);
                if (qc1[4 * (2 * r) + 3] > qc1[4 * (2 * r + 1) + 3]) {
                    double t = rr[0];
                    rr[0] = rr[1];
                    rr[1] = t;
                }
                for (m = 0; m < 2; m++) {
                    const double *a1 = qc1 + 4 * (2 * r + m);
                    int nn = a1[3] > 0 ? 1 : 0;
                    const double *b = qc2 + 4 * (2 * s + nn);
                    double d2 = (a1[0] - b[0]) * (a1[0] - b[0]) + (a1[1] - b[1]) * (a1[1] - b[1]) +
                                (a1[2] - b[2]) * (a1[2] - b[2]);
                    double ql = a1[3] / sqrt(d2);
                    w2[nn] += ql / sqrt(rr[m]);
                    q2[nn] += ql;
                }
            }
            for (m = 0; m < 2; m++) H->reffc[1][2 * s + m] = avg_radius(w2[m], q2[m], 1);
            for (m = 0; m < 2; m++) {
                const double *a2 = qc2 + 4 * (2 * s + m);
                int nn = a2[3] > 0 ? 1 : 0;
                const double *b = qc3 + 4 * (2 * c + nn);
                double d2 =
                    (a2[0] - b[0]) * (a2[0] - b[0]) + (a2[1] - b[1]) * (a2[1] - b[1]) + (a2[2] - b[2]) * (a2[2] - b[2]);
                double ql = a2[3] / sqrt(d2);
                w3[nn] += ql / sqrt(H->reffc[1][2 * s + m]);
                q3[nn] += ql;
            }
        }
        for (m = 0; m < 2; m++) H->reffc[2][2 * c + m] = avg_radius(w3[m], q3[m], 1);
    }
}

/* ------------------------------------------------------ energy and grad */

/* GB + Coulomb of atom i with one charge (q, position, Born radius rj):
 * adds to epol, elec, i's sumdeijda term and the pair force factor
 * (hcp_gb.c:1369-1436, :1473-1520) */
static double gb_pair(double qi, double ri, double qj, double rj, double r2, double kappa, double diel_ext,
                      double *epol, double *elec, double *sd, int coulomb)
{
    double qiqj = qi * qj, rb2 = ri * rj;
    double efac = exp(-r2 / (4.0 * rb2));
    double fgbi = 1.0 / sqrt(r2 + rb2 * efac);
    double fgbk = -kappa * KSCALE / fgbi;
    double expmkf = exp(fgbk) / diel_ext, dielfac = 1.0 - expmkf;
    double temp4, temp6, de, temp5;
    *epol += -qiqj * dielfac * fgbi;
    temp4 = fgbi * fgbi * fgbi;
    temp6 = qiqj * temp4 * (dielfac + fgbk * expmkf);
    de = temp6 * (1.0 - 0.25 * efac);
    temp5 = 0.5 * efac * temp6 * (rb2 + 0.25 * r2);
    *sd += ri * temp5;
    if (coulomb) {
        double rinv = 1. / sqrt(r2), r2inv = rinv * rinv, eel = qiqj * rinv;
        *elec += eel;
        de -= eel * r2inv;
    }
    return de;
}

/* numerator of dR_i/dd_ij over d_ij (hcp_gb.c:1599-1645, :1712-1764) */
static double descreen_deriv(double r2, double ri, double ri1i, double sj)
{
    double dij1i = 1.0 / sqrt(r2), dij2i = dij1i * dij1i, dij = r2 * dij1i, sj2 = sj * sj, temp1, dij3i;
    if (dij > 4.0 * sj) {
        double tmpsd = sj2 * dij2i, dumbo = TE + tmpsd * (TF + tmpsd * (TG + tmpsd * (TH + tmpsd * THH)));
        return tmpsd * sj * dij2i * dij2i * dumbo;
    } else if (dij > ri + sj) {
        temp1 = 1. / (r2 - sj2);
        return temp1 * sj * (-0.5 * dij2i + temp1) + 0.25 * dij1i * dij2i * log((dij - sj) / (dij + sj));
    } else if (dij > fabs(ri - sj)) {
        temp1 = 1. / (dij + sj);
        dij3i = dij2i * dij1i;
        return -0.25 * (-0.5 * (r2 - ri * ri + sj2) * dij3i * ri1i * ri1i + dij1i * temp1 * (temp1 - dij1i) -
                        dij3i * log(ri * temp1));
    } else if (ri < sj) {
        temp1 = 1. / (r2 - sj2);
        return -0.5 * (sj * dij2i * temp1 - 2. * sj * temp1 * temp1 -
                       0.5 * dij2i * dij1i * log((sj - dij) / (sj + dij)));
    }
    return 0.;
}

/* hcp_gb.c:1795-2112 (egb_hcp_nbond) */
static double hcp_nbond(orc_t *o, struct orc_hcp *H, const double *x, double *f, double *enb, double *eel,
                        double *sumdeijda)
{
    const PARMSTRUCT_T *p = o->prm;
    const int n = o->n, nq = H->nq, nt = orc_nthreads();
    const double kappa = o->kappa, diel_ext = o->epsext;
    double *gbuf = (double *)calloc((size_t)3 * n * nt, sizeof(double));
    double epol = 0.0, elec = 0.0, evdw = 0.0;
    int i;

    /* step 1: energies, sumdeijda and the pair forces on i */
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : epol, elec, evdw) num_threads(nt)
    for (i = 0; i < n; i++) {
        const double *xi = x + 3 * i;
        const double ri = H->reff0[i], qi = o->q[i];
        const int iaci = p->Ntypes * (p->Iac[i] - 1);
        const int *ex = H->ex_j + H->ex_ptr[i], nex = H->ex_ptr[i + 1] - H->ex_ptr[i];
        double sd = 0.0, dx = 0.0, dy = 0.0, dz = 0.0, ep = 0.0, el = 0.0, ev = 0.0;
        if (!o->frozen[i]) {
            double expmkf = exp(-KSCALE * kappa * ri) / diel_ext, dielfac = 1.0 - expmkf;
            double qi2h = 0.5 * qi * qi, qid2h = qi2h * dielfac;
            ep += -qid2h * 2.0 / ri;
            sd += qid2h - KSCALE * kappa * qi2h * expmkf * ri;
        }
        WALK(H, xi, {
            if (kind == V_ATOM) {
                if (comp != i) {
                    const int j = comp;
                    double xij = xi[0] - x[3 * j], yij = xi[1] - x[3 * j + 1], zij = xi[2] - x[3 * j + 2];
                    double r2 = xij * xij + yij * yij + zij * zij, de;
                    int excluded = 0, k;
                    for (k = 0; k < nex; k++)
                        if (ex[k] == j) { excluded = 1; break; }
                    de = gb_pair(qi, ri, o->q[j], H->reff0[j], r2, kappa, diel_ext, &ep, &el, &sd, !excluded);
                    if (!excluded) {
                        int ic = p->Cno[iaci + p->Iac[j] - 1] - 1;
                        if (ic >= 0) {
                            double rinv = 1. / sqrt(r2), r2inv = rinv * rinv, r6inv = r2inv * r2inv * r2inv;
                            double f6 = p->Cn2[ic] * r6inv, f12 = p->Cn1[ic] * r6inv * r6inv;
                            ev += f12 - f6;
                            de -= (12. * f12 - 6. * f6) * r2inv;
                        }
                    }
                    dx += de * xij; dy += de * yij; dz += de * zij;
                }
            } else {
                const int l = lvl(kind);
                int q;
                for (q = 0; q < nq; q++) {
                    const double *qv = H->qc[l] + 4 * (nq * comp + q);
                    if (fabs(qv[3]) > MIN_Q) {
                        double xij = xi[0] - qv[0], yij = xi[1] - qv[1], zij = xi[2] - qv[2];
                        double r2 = xij * xij + yij * yij + zij * zij;
                        double de = gb_pair(qi, ri, qv[3], H->reffc[l][nq * comp + q], r2, kappa, diel_ext, &ep, &el,
                                            &sd, 1);
                        dx += de * xij; dy += de * yij; dz += de * zij;
                    }
                }
            }
        });
        sumdeijda[i] = sd;
        f[3 * i] += dx; f[3 * i + 1] += dy; f[3 * i + 2] += dz;
        epol += ep; elec += el; evdw += ev;
    }

    /* step 2: the chain rule through the Born radii */
#pragma omp parallel for schedule(dynamic, 8) num_threads(nt)
    for (i = 0; i < n; i++) {
        double *g = gbuf + (size_t)3 * n * omp_thread();
        const double *xi = x + 3 * i;
        const double ri = p->Rborn[i] - o->gboffset, ri1i = 1. / ri;
        double sumda = sumdeijda[i], dx = 0.0, dy = 0.0, dz = 0.0;
        if (o->frozen[i]) continue;
        if (o->gb > 1) {
            const double ps = H->psi[i];
            double thi = tanh((o->alpha[i] - o->beta[i] * ps + o->gamma[i] * ps * ps) * ps);
            sumda *= (o->alpha[i] - 2.0 * o->beta[i] * ps + 3.0 * o->gamma[i] * ps * ps) * (1.0 - thi * thi) * ri /
                     p->Rborn[i];
        }
        WALK(H, xi, {
            if (kind == V_ATOM) {
                if (comp != i) {
                    const int j = comp;
                    double xij = xi[0] - x[3 * j], yij = xi[1] - x[3 * j + 1], zij = xi[2] - x[3 * j + 2];
                    double r2 = xij * xij + yij * yij + zij * zij;
                    double datmp = descreen_deriv(r2, ri, ri1i, o->fs[j] * (p->Rborn[j] - o->gboffset)) * sumda;
                    dx += xij * datmp; dy += yij * datmp; dz += zij * datmp;
                    g[3 * j] += xij * datmp; g[3 * j + 1] += yij * datmp; g[3 * j + 2] += zij * datmp;
                }
            } else {
                const int l = lvl(kind);
                int q;
                for (q = 0; q < nq; q++) {
                    const double *qv = H->qc[l] + 4 * (nq * comp + q);
                    double xij = xi[0] - qv[0], yij = xi[1] - qv[1], zij = xi[2] - qv[2];
                    double r2 = xij * xij + yij * yij + zij * zij;
                    double datmp = descreen_deriv(r2, ri, ri1i, H->rc[l][nq * comp + q]) * (sumda * 2);
                    dx += xij * datmp; dy += yij * datmp; dz += zij * datmp;
                }
            }
        });
        g[3 * i] -= dx; g[3 * i + 1] -= dy; g[3 * i + 2] -= dz;
    }
    {
        int t;
        for (t = 0; t < nt; t++)
            for (i = 0; i < 3 * n; i++) f[i] += gbuf[(size_t)3 * n * t + i];
    }
    free(gbuf);
    *eel = elec * 0.5;
    *enb = evdw * 0.5;
    return epol * 0.5;
}

/* hcp_gb.c:2166-2240 (egb_hcp) without the surface-area term */
double orc_egb_hcp(orc_t *o, const double *x, double *f, double *enb, double *eel, double *reff_out,
                   double *sumdeijda_out)
{
    struct orc_hcp *H = o->hcp;
    double *sd = (double *)calloc(o->n, sizeof(double)), e;
    if (H->opca) approx_charges_opca(o, H, x);
    else approx_charges(o, H, x);
    hcp_born(o, H, x);
    if (H->opca) hcp_reff_avg_opca(o, H, x);
    else hcp_reff_avg(o, H);
    e = hcp_nbond(o, H, x, f, enb, eel, sd);
    if (reff_out) memcpy(reff_out, H->reff0, sizeof(double) * o->n);
    if (sumdeijda_out) memcpy(sumdeijda_out, sd, sizeof(double) * o->n);
    free(sd);
    return e;
}
