/*
 * hcp.hip -- hierarchical charge partitioning GB on the GPU (gfx950).
 *
 * The drop-in for egb_hcp() (src/sff/hcp_gb.c:2166-2240), which mme34 calls
 * in place of egb() when mm_options sets hcp (eff.c:2687-2694) and for which
 * it builds no pair list (eff.c:2560-2561).  Atoms keep their original
 * order, grouped into residues (prm->Ipres), strands and complexes
 * (prm->Ipstrand / Ipcomplex, prm.c:1743-1790).  Each atom walks the
 * hierarchy: a complex, strand or residue whose geometric centre lies
 * farther from the atom than that level's threshold acts through one or
 * two point charges (hcp 1: centre of charge; hcp 2: negative and positive
 * atoms apart; hcp 4: two optimal point charges, HCPO, hcp.c:150-840),
 * nearer ones are opened down to their atoms.
 *
 * One evaluation, on the engine stream:
 *   k_hcp_res      per residue: centre, charges (calc_approx_q /
 *                  calc_approx_q_hcpo at the residue level) and the
 *                  unnormalised sums the levels above add up
 *   k_hcp_up       strands, then complexes (one workgroup), in the
 *                  reference's summation order
 *   k_hcp_born     descreening sums (egb_hcp_atom_reff)
 *   k_hcp_radii    per atom: Born radius, diagonal GB term
 *   k_hcp_reff_res per residue: component Born radii (egb_hcp_reff /
 *                  egb_hcpo_reff_avg at the residue level)
 *   k_hcp_reff_up  strands and complexes
 *   k_hcp_energy   GB, Coulomb and LJ energies, sumdeijda and the pair
 *                  forces on each atom (egb_hcp_nbond step 1)
 *   k_hcp_deriv    the chain rule through the Born radii (step 2); forces
 *                  on member atoms summed in LDS, one atomic per atom and
 *                  residue visit into an array in atom order
 *   k_hcp_sum      per atom: the items' partial sums (and those member
 *                  forces) into sd and the gradient
 * then the engine's own energy-slot and gradient kernels.  The walking
 * kernels take the context's precision: FP64, or the mixed variant with
 * the descreening and GB terms in FP32 (Coulomb, LJ, coordinate
 * differences and every sum FP64).  What the reference computes in single
 * precision (component radii, the HCPO moments and point charges) is
 * single precision here in both, compiled without contraction so each
 * operation rounds as it does there.
 *
 * The walking kernels (born, energy, deriv) take work items of one
 * wavefront: the 64 atoms of a tile (one per lane, original order) against
 * one range of consecutive residues, about 256 work items per CU, range
 * after range over the tiles.  An item stores its lanes' sums (Born sums,
 * sumdeijda, forces on the lane's atom) in a buffer of its own, and a pass
 * over the atoms adds them up in range order: atomics from every item of
 * a tile to its 64 atoms had cost about as much as the walk itself at 25k
 * atoms.  The forces on member atoms in the chain rule still go through
 * LDS sums and atomics.  Within an item the hierarchy loops are
 * wavefront-uniform (every lane steps through the same complexes, strands,
 * residues and atoms) and a lane's branch (far charges or the component's
 * members) runs only where some lane needs it.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "engine.h"
#include "gb_math.h"

#define HCP_KSCALE 0.73
#define HCP_MIN_Q 0.00001
#define HCP_TA 0.33333333333333333333
#define HCP_TB 0.4
#define HCP_TC 0.42857142857142857143
#define HCP_TD 0.44444444444444444444
#define HCP_TDD 0.45454545454545454545
#define HCP_TE 1.33333333333333333333
#define HCP_TF 2.4
#define HCP_TG 3.42857142857142857143
#define HCP_TH 4.44444444444444444444
#define HCP_THH 5.45454545454545454545

#define HCP_THREADS 256
#define MP_N 10 /* single-precision moments: q, p[3], Q00 Q01 Q02 Q11 Q12 Q22 */

/* ------------------------------------------------------- device helpers */

__device__ __forceinline__ double hcp_d2(double xi, double yi, double zi, const double *p)
{
    const double xij = xi - p[0], yij = yi - p[1], zij = zi - p[2];
    return xij * xij + yij * yij + zij * zij;
}

__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0ull; }

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

/*
 * Pair arithmetic, R = double (every operation FP64) or float (the mixed
 * variant: the descreening and GB terms in FP32 with the hardware
 * transcendentals, as the pair kernels' mixed variant computes them;
 * squared distances from FP64 coordinates, Coulomb and LJ in FP64, every
 * sum the caller keeps in FP64).  gbm<R> (gb_math.h) supplies exp, log,
 * reciprocal and reciprocal square root.
 */

/* far component: the Taylor term of the descreening integral only
 * (hcp_gb.c:139-160) */
template <typename R> __device__ __forceinline__ double far_descreen(double r2d, double sjd)
{
    const R r2 = (R)r2d, sj = (R)sjd;
    const R rs = gbm<R>::rsqrt_(r2), dij2i = rs * rs, sj2 = sj * sj, tmpsd = sj2 * dij2i;
    const R dumbo = R(HCP_TA) + tmpsd * (R(HCP_TB) + tmpsd * (R(HCP_TC) + tmpsd * (R(HCP_TD) + tmpsd * R(HCP_TDD))));
    return (double)(sj * tmpsd * dij2i * dumbo);
}

/* a member atom's descreening of atom i: the analytic cases without the
 * rgbmax cut-off or the neck term (hcp_gb.c:345-390) */
template <typename R> __device__ __forceinline__ double near_descreen(double r2d, R ri, R ri1i, double sjd)
{
    const R r2 = (R)r2d, sj = (R)sjd;
    const R dij1i = gbm<R>::rsqrt_(r2), dij = r2 * dij1i, sj2 = sj * sj;
    R v = R(0);
    if (dij > R(4) * sj) {
        const R dij2i = dij1i * dij1i, tmpsd = sj2 * dij2i;
        const R dumbo =
            R(HCP_TA) + tmpsd * (R(HCP_TB) + tmpsd * (R(HCP_TC) + tmpsd * (R(HCP_TD) + tmpsd * R(HCP_TDD))));
        v = sj * tmpsd * dij2i * dumbo;
    } else if (dij > ri + sj) {
        v = R(0.5) * (sj * gbm<R>::rcp(r2 - sj2) + R(0.5) * dij1i * gbm<R>::log_((dij - sj) * gbm<R>::rcp(dij + sj)));
    } else if (dij > fabs(ri - sj)) {
        const R theta = R(0.5) * ri1i * dij1i * (r2 + ri * ri - sj2), uij = gbm<R>::rcp(dij + sj);
        v = R(0.25) * (ri1i * (R(2) - theta) - uij + dij1i * gbm<R>::log_(ri * uij));
    } else if (ri < sj) {
        v = R(0.5) * (sj * gbm<R>::rcp(r2 - sj2) + R(2) * ri1i +
                      R(0.5) * dij1i * gbm<R>::log_((sj - dij) * gbm<R>::rcp(sj + dij)));
    }
    return (double)v;
}

/* numerator of dR_i/dd_ij over d_ij (hcp_gb.c:1599-1645, :1712-1764) */
template <typename R> __device__ __forceinline__ double descreen_deriv(double r2d, R ri, R ri1i, double sjd)
{
    const R r2 = (R)r2d, sj = (R)sjd;
    const R dij1i = gbm<R>::rsqrt_(r2), dij2i = dij1i * dij1i, dij = r2 * dij1i, sj2 = sj * sj;
    R v = R(0);
    if (dij > R(4) * sj) {
        const R tmpsd = sj2 * dij2i;
        const R dumbo =
            R(HCP_TE) + tmpsd * (R(HCP_TF) + tmpsd * (R(HCP_TG) + tmpsd * (R(HCP_TH) + tmpsd * R(HCP_THH))));
        v = tmpsd * sj * dij2i * dij2i * dumbo;
    } else if (dij > ri + sj) {
        const R t1 = gbm<R>::rcp(r2 - sj2);
        v = t1 * sj * (R(-0.5) * dij2i + t1) +
            R(0.25) * dij1i * dij2i * gbm<R>::log_((dij - sj) * gbm<R>::rcp(dij + sj));
    } else if (dij > fabs(ri - sj)) {
        const R t1 = gbm<R>::rcp(dij + sj), dij3i = dij2i * dij1i;
        v = R(-0.25) * (R(-0.5) * (r2 - ri * ri + sj2) * dij3i * ri1i * ri1i + dij1i * t1 * (t1 - dij1i) -
                        dij3i * gbm<R>::log_(ri * t1));
    } else if (ri < sj) {
        const R t1 = gbm<R>::rcp(r2 - sj2);
        v = R(-0.5) * (sj * dij2i * t1 - R(2) * sj * t1 * t1 -
                       R(0.5) * dij2i * dij1i * gbm<R>::log_((sj - dij) * gbm<R>::rcp(sj + dij)));
    }
    return (double)v;
}

/* FP64 1/r for Coulomb and LJ: the hardware estimate refined by Newton
 * steps (FP64) or the FP32 estimate refined by one FP64 step (mixed) */
template <typename R> __device__ __forceinline__ double hcp_rinv(double r2);
template <> __device__ __forceinline__ double hcp_rinv<double>(double r2) { return gbm<double>::rsqrt_(r2); }
template <> __device__ __forceinline__ double hcp_rinv<float>(double r2) { return rsqrt_pair<float>(r2); }

/* GB (+ Coulomb) of atom i with one charge qj of Born radius rj at r^2:
 * energies, i's sumdeijda term; returns the pair force factor
 * (hcp_gb.c:1369-1436, :1473-1520).  f_GB from one reciprocal square root
 * (1/f and f = (r^2 + RiRj e) / f), the screening exponential only with
 * salt (KAPPA); kks = -0.73 kappa, inv_eps = 1 / epsext. */
template <typename R, bool KAPPA>
__device__ __forceinline__ double gb_pair(double qi, R ri, double qj, R rj, double r2d, R kks, R inv_eps,
                                          double &epol, double &elec, double &sd, bool coulomb)
{
    const R r2 = (R)r2d;
    const double qiqj_d = qi * qj;
    const R qiqj = (R)qiqj_d, rb2 = ri * rj;
    const R efac = gbm<R>::exp_(-r2 * gbm<R>::rcp(R(4) * rb2));
    const R fsq = r2 + rb2 * efac;
    const R fgbi = gbm<R>::rsqrt_(fsq);
    R expmkf, fgbk = R(0);
    if (KAPPA) {
        fgbk = kks * (fsq * fgbi);
        expmkf = gbm<R>::exp_(fgbk) * inv_eps;
    } else {
        expmkf = inv_eps;
    }
    const R dielfac = R(1) - expmkf;
    epol += (double)(-qiqj * dielfac * fgbi);
    const R temp4 = fgbi * fgbi * fgbi;
    const R temp6 = qiqj * temp4 * (dielfac + fgbk * expmkf);
    double de = (double)(temp6 * (R(1) - R(0.25) * efac));
    const R temp5 = R(0.5) * efac * temp6 * (rb2 + R(0.25) * r2);
    sd += (double)(ri * temp5);
    if (coulomb) {
        const double rinv = hcp_rinv<R>(r2d), r2inv = rinv * rinv, eel = qiqj_d * rinv;
        elec += eel;
        de -= eel * r2inv;
    }
    return de;
}

/* ---------------------------------------------- HCPO point charges (float) */

#pragma clang fp contract(off)

/* hcp.c:20-82: the first eigenpair of a symmetric 3x3 matrix (the smallest
 * eigenvalue, s cos(alpha / 3) + tr / 3 of the trigonometric roots of the
 * characteristic cubic) and its eigenvector from cofactor combinations */
__device__ static void sym3_first_eigen(const double m[3][3], double v0[3], double *val0)
{
    const double k0 = m[0][0] * m[1][1] * m[2][2] + 2.0 * m[0][1] * m[1][2] * m[0][2] - m[0][0] * m[1][2] * m[1][2] -
                      m[1][1] * m[0][2] * m[0][2] - m[2][2] * m[0][1] * m[0][1];
    const double k1 = -m[0][0] * m[1][1] - m[1][1] * m[2][2] - m[0][0] * m[2][2] + m[0][1] * m[0][1] +
                      m[1][2] * m[1][2] + m[0][2] * m[0][2];
    const double k2 = m[0][0] + m[1][1] + m[2][2];
    const double pp = -k1 - k2 * k2 / 3.0;
    const double qq = -k0 - 2.0 * k2 * k2 * k2 / 27.0 - k1 * k2 / 3.0;
    const double s = -2.0 * sqrt(-pp / 3.0);
    const double arc = -4.0 * qq / (s * s * s);
    const double alpha = acos(arc);
    const double val = s * cos(alpha / 3.0) + k2 / 3.0;
    const double A = m[0][0] - val, B = m[1][1] - val, C = m[2][2] - val;
    const double AA = m[0][2] * m[0][1] - A * m[1][2];
    const double BB = m[0][1] * m[1][2] - B * m[0][2];
    const double CC = m[0][2] * m[1][2] - C * m[0][1];
    v0[0] = BB * CC;
    v0[1] = CC * AA;
    v0[2] = AA * BB;
    const double mag = sqrt(v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2]);
    if (mag != 0) {
        v0[0] = v0[0] / mag;
        v0[1] = v0[1] / mag;
        v0[2] = v0[2] / mag;
    }
    *val0 = val;
}

/* offset limit of an optimal point charge (hcp.c:300-333) */
__device__ static float opca_limit(float dis, float R1, float R2)
{
    const float R1s = R1 * R1, R2s = R2 * R2;
    const float denom = (R1s - R2s) * (R1s - R2s) * (R1s - R2s);
    const float dis2 = dis * dis;
    if (dis < R1) return dis;
    if (dis > R2) return R2;
    return (float)((R1s - dis2) * (R1s - dis2) * (R1s + 2.0 * dis * dis - 3.0 * R2s) / denom * (R2 - dis) + dis);
}

/*
 * OPCA with two charges (hcp.c:150-335) from single-precision moments
 * mp = {q, p[3], Q00 Q01 Q02 Q11 Q12 Q22} about the origin; writes
 * qv[0..7] = (x1, y1, z1, q1, x2, y2, z2, q2).
 */
__device__ static void opca2(const float *mp, float rmax, const float rcog[3], float R1, float R2, float coef2,
                             double *qv)
{
    const float qtotal = mp[0];
    const float p[3] = {mp[1], mp[2], mp[3]};
    float Q[3][3] = {{mp[4], mp[5], mp[6]}, {mp[5], mp[7], mp[8]}, {mp[6], mp[8], mp[9]}};
    float qbar1 = 0.0f, qbar2 = 0.0f, rbar1[3] = {0, 0, 0}, rbar2[3] = {0, 0, 0};
    if (fabs((double)qtotal) > 0.6 * 18.2223) {
        float rcoq[3];
        rcoq[0] = p[0] / qtotal;
        rcoq[1] = p[1] / qtotal;
        rcoq[2] = p[2] / qtotal;
        const float rcoq2 = rcoq[0] * rcoq[0] + rcoq[1] * rcoq[1] + rcoq[2] * rcoq[2];
        const float prcoq = p[0] * rcoq[0] + p[1] * rcoq[1] + p[2] * rcoq[2];
        for (int i = 0; i < 3; i++)
            for (int j = i; j < 3; j++)
                Q[i][j] = (float)(0.5 * (2.0 * (double)Q[i][j] -
                                         (3.0 * (double)p[i] * (double)rcoq[j] + 3.0 * (double)rcoq[i] * (double)p[j]) +
                                         (double)qtotal * (3.0 * (double)rcoq[i] * (double)rcoq[j])));
        Q[1][0] = Q[0][1];
        Q[2][0] = Q[0][2];
        Q[2][1] = Q[1][2];
        for (int i = 0; i < 3; i++) {
            const float t = Q[i][i] + prcoq;
            Q[i][i] = (float)((double)t - 0.5 * (double)qtotal * (double)rcoq2);
        }
        double A[3][3], v0[3], d0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) A[i][j] = Q[i][j];
        sym3_first_eigen(A, v0, &d0);
        float rbar = 1.5f * rmax;
        if ((double)(rbar * rbar * qtotal) + d0 == 0.0) rbar = 0.0f;
        qbar1 = (float)((double)qtotal * d0 / ((double)(rbar * rbar * qtotal) + d0));
        qbar2 = qtotal - qbar1;
        for (int i = 0; i < 3; i++) rbar1[i] = (float)((double)rbar * v0[i]);
        if (qbar2 != 0.0f) {
            for (int i = 0; i < 3; i++) rbar2[i] = -qbar1 / qbar2 * rbar1[i];
        }
        float cod[3];
        for (int i = 0; i < 3; i++) cod[i] = (float)((double)(rbar1[i] + rbar2[i]) / 2.0);
        const float cdotq = cod[0] * (rcoq[0] - rcog[0]) + cod[1] * (rcoq[1] - rcog[1]) + cod[2] * (rcoq[2] - rcog[2]);
        for (int i = 0; i < 3; i++) {
            if (cdotq > 0) {
                rbar1[i] = -rbar1[i];
                rbar2[i] = -rbar2[i];
            }
            rbar1[i] = rbar1[i] + rcoq[i] - rcog[i];
            rbar2[i] = rbar2[i] + rcoq[i] - rcog[i];
        }
    } else {
        float sumqpp = 0.0f;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) sumqpp = sumqpp + Q[i][j] * p[i] * p[j];
        const float mpsq = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
        float dk[3];
        for (int i = 0; i < 3; i++) {
            float sumqp = 0.0f;
            for (int j = 0; j < 3; j++) sumqp = sumqp + Q[i][j] * p[j];
            dk[i] = (float)(0.0 + 2.0 / (3.0 * (double)mpsq) *
                                      ((double)sumqp - ((double)sumqpp / (4.0 * (double)mpsq) * (double)p[i])));
        }
        const float dr = coef2 * rmax;
        const double pp = (double)p[0] * (double)p[0] + (double)p[1] * (double)p[1] + (double)p[2] * (double)p[2];
        const float q = (float)(sqrt(pp) / (double)dr);
        qbar1 = q;
        qbar2 = -q;
        for (int i = 0; i < 3; i++) {
            rbar1[i] = (float)((double)p[i] / (2.0 * (double)q) + (double)dk[i] - (double)rcog[i]);
            rbar2[i] = (float)((double)(-p[i]) / (2.0 * (double)q) + (double)dk[i] - (double)rcog[i]);
        }
    }
    const float a1 = (float)sqrt((double)(rbar1[0] * rbar1[0] + rbar1[1] * rbar1[1] + rbar1[2] * rbar1[2]));
    const float a2 = (float)sqrt((double)(rbar2[0] * rbar2[0] + rbar2[1] * rbar2[1] + rbar2[2] * rbar2[2]));
    const float m1 = opca_limit(a1, R1, R2), m2 = opca_limit(a2, R1, R2);
    for (int i = 0; i < 3; i++) {
        qv[i] = (double)(m1 / a1 * rbar1[i] + rcog[i]);
        qv[4 + i] = (double)(m2 / a2 * rbar2[i] + rcog[i]);
    }
    qv[3] = qbar1;
    qv[7] = qbar2;
}

/* moments of one atom added (hcp.c:633-653) */
__device__ __forceinline__ void mp_add_atom(float *m, double x, double y, double z, double q)
{
    const float r2 = (float)(x * x + y * y + z * z);
    const double qh = 0.5 * q;
    m[0] = (float)((double)m[0] + q);
    m[1] = (float)((double)m[1] + x * q);
    m[2] = (float)((double)m[2] + y * q);
    m[3] = (float)((double)m[3] + z * q);
    m[4] = (float)((double)m[4] + qh * (3.0 * x * x - (double)r2));
    m[5] = (float)((double)m[5] + qh * (3.0 * x * y));
    m[6] = (float)((double)m[6] + qh * (3.0 * x * z));
    m[7] = (float)((double)m[7] + qh * (3.0 * y * y - (double)r2));
    m[8] = (float)((double)m[8] + qh * (3.0 * y * z));
    m[9] = (float)((double)m[9] + qh * (3.0 * z * z - (double)r2));
}

/* the level above accumulates the moments of the level below */
__device__ __forceinline__ void mp_roll(float *up, const float *m)
{
    for (int k = 0; k < MP_N; k++) up[k] = up[k] + m[k];
}

/* largest distance of n points (stride 3) from the centre, single precision */
__device__ static float spread(const double *pts, int n, const float rcog[3])
{
    float best = 0.0f;
    for (int k = 0; k < n; k++) {
        const double *x = pts + 3 * (size_t)k;
        const float d2 = (float)((x[0] - (double)rcog[0]) * (x[0] - (double)rcog[0]) +
                                 (x[1] - (double)rcog[1]) * (x[1] - (double)rcog[1]) +
                                 (x[2] - (double)rcog[2]) * (x[2] - (double)rcog[2]));
        if (d2 > best) best = d2;
    }
    return (float)sqrt((double)best);
}

/* two optimal point charges of a component with the given moments and
 * centre; offsets limited to 0.8 / 0.9 of its spread (gb > 0, hcp.c:578-581) */
__device__ static void opca_component(const float *mp, const double *centre, const double *pts, int npts,
                                      float coef2, double *qv)
{
    float rcog[3];
    for (int k = 0; k < 3; k++) rcog[k] = (float)centre[k];
    const float rmax = spread(pts, npts, rcog);
    opca2(mp, rmax, rcog, 0.8f * rmax, 0.9f * rmax, coef2, qv);
}

#pragma clang fp contract(fast)

/* a charge at its centre of charge, or at the component's centre when it
 * is (nearly) zero (hcp.c:428-441) */
__device__ __forceinline__ void place_charge(const double *sum4, const double *centre, double *out4)
{
    const double q = sum4[3];
    if (fabs(q) > HCP_MIN_Q) {
        out4[0] = sum4[0] / q;
        out4[1] = sum4[1] / q;
        out4[2] = sum4[2] / q;
    } else {
        out4[0] = centre[0];
        out4[1] = centre[1];
        out4[2] = centre[2];
    }
    out4[3] = q;
}

/* ------------------------------------------------ component charges */

/* per residue: centre, charges and the sums the strand level adds up */
template <bool OPCA>
__global__ __launch_bounds__(HCP_THREADS) void k_hcp_res(HcpDev H, const double *__restrict__ x)
{
    const int r = blockIdx.x * HCP_THREADS + threadIdx.x;
    if (r >= H.nres) return;
    const int a0 = H.ipres[r], a1 = H.ipres[r + 1], n1 = a1 - a0, nq = H.nq;
    double s[3] = {0.0, 0.0, 0.0};
    double qs[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    float mp[MP_N];
    for (int k = 0; k < MP_N; k++) mp[k] = 0.0f;
    for (int a = a0; a < a1; a++) {
        const double xa = x[3 * (size_t)a], ya = x[3 * (size_t)a + 1], za = x[3 * (size_t)a + 2];
        const double q = H.atom[a].x;
        s[0] += xa;
        s[1] += ya;
        s[2] += za;
        if (OPCA) {
            mp_add_atom(mp, xa, ya, za, q);
        } else {
            const int m = (nq == 2 && q > 0.0) ? 1 : 0;
            qs[m][0] += xa * q;
            qs[m][1] += ya * q;
            qs[m][2] += za * q;
            qs[m][3] += q;
        }
    }
    double *rs = H.rsum + (size_t)r * (3 + 4 * nq);
    for (int k = 0; k < 3; k++) rs[k] = s[k];
    double c[3] = {s[0], s[1], s[2]};
    if (n1 > 0)
        for (int k = 0; k < 3; k++) c[k] /= n1;
    for (int k = 0; k < 3; k++) H.xc1[3 * (size_t)r + k] = c[k];
    if (OPCA) {
        for (int k = 0; k < MP_N; k++) H.rmp[(size_t)r * MP_N + k] = mp[k];
        opca_component(mp, c, x + 3 * (size_t)a0, n1, H.dhcp, H.qc1 + 8 * (size_t)r);
    } else {
        for (int m = 0; m < nq; m++) {
            for (int k = 0; k < 4; k++) rs[3 + 4 * m + k] = qs[m][k];
            place_charge(qs[m], c, H.qc1 + 4 * ((size_t)r * nq + m));
        }
    }
}

/* strands (each summing its residues in order), then complexes (each
 * summing its strands); one workgroup */
template <bool OPCA>
__global__ __launch_bounds__(HCP_THREADS) void k_hcp_up(HcpDev H)
{
    const int nq = H.nq, w = 3 + 4 * nq;
    for (int st = threadIdx.x; st < H.nstrand; st += HCP_THREADS) {
        double s[3 + 8];
        for (int k = 0; k < w; k++) s[k] = 0.0;
        float mp[MP_N];
        for (int k = 0; k < MP_N; k++) mp[k] = 0.0f;
        int n2 = 0;
        const int r0 = H.rfirst[st], r1 = H.rfirst[st + 1];
        for (int r = r0; r < r1; r++) {
            const double *rs = H.rsum + (size_t)r * w;
            n2 += H.ipres[r + 1] - H.ipres[r];
            for (int k = 0; k < 3; k++) s[k] += rs[k];
            if (OPCA) mp_roll(mp, H.rmp + (size_t)r * MP_N);
            else
                for (int k = 3; k < w; k++) s[k] += rs[k];
        }
        double *ss = H.ssum + (size_t)st * w;
        for (int k = 0; k < w; k++) ss[k] = s[k];
        H.scnt[st] = n2;
        double c[3] = {s[0], s[1], s[2]};
        if (n2 > 0)
            for (int k = 0; k < 3; k++) c[k] /= n2;
        for (int k = 0; k < 3; k++) H.xc2[3 * (size_t)st + k] = c[k];
        if (OPCA) {
            for (int k = 0; k < MP_N; k++) H.smp[(size_t)st * MP_N + k] = mp[k];
            opca_component(mp, c, H.xc1 + 3 * (size_t)r0, r1 - r0, H.dhcp, H.qc2 + 8 * (size_t)st);
        } else {
            for (int m = 0; m < nq; m++) place_charge(s + 3 + 4 * m, c, H.qc2 + 4 * ((size_t)st * nq + m));
        }
    }
    __syncthreads();
    for (int cx = threadIdx.x; cx < H.ncomplex; cx += HCP_THREADS) {
        double s[3 + 8];
        for (int k = 0; k < w; k++) s[k] = 0.0;
        float mp[MP_N];
        for (int k = 0; k < MP_N; k++) mp[k] = 0.0f;
        int n3 = 0;
        const int s0 = H.sfirst[cx], s1 = H.sfirst[cx + 1];
        for (int st = s0; st < s1; st++) {
            const double *ss = H.ssum + (size_t)st * w;
            n3 += H.scnt[st];
            for (int k = 0; k < 3; k++) s[k] += ss[k];
            if (OPCA) mp_roll(mp, H.smp + (size_t)st * MP_N);
            else
                for (int k = 3; k < w; k++) s[k] += ss[k];
        }
        double c[3] = {s[0], s[1], s[2]};
        if (n3 > 0)
            for (int k = 0; k < 3; k++) c[k] /= n3;
        for (int k = 0; k < 3; k++) H.xc3[3 * (size_t)cx + k] = c[k];
        if (OPCA) opca_component(mp, c, H.xc2 + 3 * (size_t)s0, s1 - s0, H.dhcp, H.qc3 + 8 * (size_t)cx);
        else
            for (int m = 0; m < nq; m++) place_charge(s + 3 + 4 * m, c, H.qc3 + 4 * ((size_t)cx * nq + m));
    }
}

/* ------------------------------------------------- the hierarchy walk */

/*
 * One work item: the 64 atoms of a tile (one per lane) against the residues
 * r0 .. r1-1 of a residue range (which may span strands and complexes).
 * far(level, comp) runs for lanes to which component comp of that level
 * (0 residue, 1 strand, 2 complex) is far -- a far strand or complex by
 * the range holding its first residue only, so each atom meets it once;
 * member(a, on) runs for every lane at each atom a of a residue that is
 * near to some lane, `on` telling whether it is near to this one (and
 * a != i) -- or, given as near_res(r, in1), once per such residue.  The
 * loops are wavefront-uniform: all lanes reach every member() or
 * near_res() call together, so these may use cross-lane operations and
 * the workgroup barrier (one wavefront per workgroup).
 */
template <typename Far, typename NearRes>
__device__ __forceinline__ void hcp_walk_res(const HcpDev &H, int r0, int r1, bool valid, double xi, double yi,
                                             double zi, Far far, NearRes near_res)
{
    int r = r0;
    while (r < r1) {
        const int s = H.res_s[r], c = H.s_cx[s];
        const int sr0 = H.rfirst[s], sr1 = H.rfirst[s + 1];
        const int cr0 = H.rfirst[H.sfirst[c]], cr1 = H.rfirst[H.sfirst[c + 1]];
        const bool f3 = valid && hcp_d2(xi, yi, zi, H.xc3 + 3 * c) > H.h3sq;
        if (r == cr0 && wave_any(f3)) {
            if (f3) far(2, c);
        }
        const bool in3 = valid && !f3;
        if (!wave_any(in3)) {
            r = min(r1, cr1);
            continue;
        }
        const double ds = hcp_d2(xi, yi, zi, H.xc2 + 3 * s);
        const bool f2 = in3 && ds > H.h2sq;
        if (r == sr0 && wave_any(f2)) {
            if (f2) far(1, s);
        }
        const bool in2 = in3 && !(ds > H.h2sq);
        const int rend = min(r1, sr1);
        if (wave_any(in2)) {
            for (; r < rend; r++) {
                const double dr = hcp_d2(xi, yi, zi, H.xc1 + 3 * r);
                const bool f1 = in2 && dr > H.h1sq;
                if (wave_any(f1)) {
                    if (f1) far(0, r);
                }
                const bool in1 = in2 && !(dr > H.h1sq);
                if (!wave_any(in1)) continue;
                near_res(r, in1);
            }
        }
        r = rend;
    }
}

/* member(a) for each member a != i of each residue near to this lane, in a
 * wavefront-uniform loop over the members (lanes without the residue idle) */
template <typename Far, typename Member>
__device__ __forceinline__ void hcp_walk(const HcpDev &H, int r0, int r1, bool valid, int i, double xi, double yi,
                                         double zi, Far far, Member member)
{
    hcp_walk_res(H, r0, r1, valid, xi, yi, zi, far, [&](int r, bool in1) {
        if (in1)
            for (int a = H.ipres[r]; a < H.ipres[r + 1]; a++)
                if (a != i) member(a);
    });
}

/*
 * The same walk with the near residues taken per lane: while the far terms
 * are evaluated, each lane marks its near residues in a 64-bit mask (bit k
 * = residue base + k); every 64 residues, and at the end, the lanes drain
 * their own masks independently, atom after atom in increasing order,
 * member(a) for each member a != i of each marked residue, all lanes busy
 * until the longest list ends.  A tile's lanes disagree about most of the
 * residues near them (a residue is near to some of the 64 atoms and far
 * from the rest), which leaves a wavefront-uniform atom loop with half its
 * lanes idle; here only the far charges of such residues run on part of
 * the wavefront.  member() must not use cross-lane operations.  Used by the
 * energy pass (FP64 2.99 -> 2.86 ms at 100k atoms, mixed unchanged); the
 * Born pass, whose member terms are short, was slower with it (mixed 0.78
 * -> 0.83 ms: the per-lane gathers of member data cost more than the idle
 * lanes).
 */
template <typename Far, typename Member>
__device__ __forceinline__ void hcp_walk_lanes(const HcpDev &H, int r0, int r1, bool valid, int i, double xi,
                                               double yi, double zi, Far far, Member member)
{
    unsigned long long nm = 0ull;
    int base = r0;
    auto drain = [&]() {
        if (!wave_any(nm != 0ull)) return;
        int a = 0, aend = 0;
        for (;;) {
            while (a == aend && nm != 0ull) {
                const int k = __ffsll((long long)nm) - 1;
                nm &= nm - 1ull;
                a = H.ipres[base + k];
                aend = H.ipres[base + k + 1];
            }
            const bool on = a < aend;
            if (!wave_any(on)) break;
            if (on) {
                if (a != i) member(a);
                a++;
            }
        }
        nm = 0ull;
    };
    int r = r0;
    while (r < r1) {
        const int s = H.res_s[r], c = H.s_cx[s];
        const int sr0 = H.rfirst[s], sr1 = H.rfirst[s + 1];
        const int cr0 = H.rfirst[H.sfirst[c]], cr1 = H.rfirst[H.sfirst[c + 1]];
        const bool f3 = valid && hcp_d2(xi, yi, zi, H.xc3 + 3 * c) > H.h3sq;
        if (r == cr0 && wave_any(f3)) {
            if (f3) far(2, c);
        }
        const bool in3 = valid && !f3;
        if (!wave_any(in3)) {
            r = min(r1, cr1);
            continue;
        }
        const double ds = hcp_d2(xi, yi, zi, H.xc2 + 3 * s);
        const bool f2 = in3 && ds > H.h2sq;
        if (r == sr0 && wave_any(f2)) {
            if (f2) far(1, s);
        }
        const bool in2 = in3 && !(ds > H.h2sq);
        const int rend = min(r1, sr1);
        if (wave_any(in2)) {
            for (; r < rend; r++) {
                if (r - base >= 64) {
                    drain();
                    base = r;
                }
                const double dr = hcp_d2(xi, yi, zi, H.xc1 + 3 * r);
                const bool f1 = in2 && dr > H.h1sq;
                if (wave_any(f1)) {
                    if (f1) far(0, r);
                }
                if (in2 && !(dr > H.h1sq)) nm |= 1ull << (r - base);
            }
        }
        r = rend;
    }
    drain();
}

/* the work item of this wavefront: its tile's atom for this lane and its
 * residue range */
struct HcpItem {
    int i, r0, r1;
    int k, t; /* residue range, and tile * 64 + lane */
    bool valid;
};

__device__ __forceinline__ HcpItem hcp_item(const HcpDev &H)
{
    const int b = blockIdx.x, k = b / H.ntile, tile = b - k * H.ntile;
    HcpItem it;
    const int t = tile * 64 + (int)threadIdx.x;
    it.k = k;
    it.t = t;
    it.valid = t < H.natom;
    it.i = it.valid ? t : H.natom - 1;
    it.r0 = k * H.csize;
    it.r1 = min(H.nres, it.r0 + H.csize);
    return it;
}

/* this lane's partial sum c of its item */
__device__ __forceinline__ double &hcp_part(const HcpDev &H, const HcpItem &it, int c)
{
    return H.part[((size_t)c * H.nchunk + it.k) * ((size_t)H.ntile * 64) + it.t];
}

__device__ __forceinline__ const double *lvl_q(const HcpDev &H, int l)
{
    return l == 0 ? H.qc1 : (l == 1 ? H.qc2 : H.qc3);
}
__device__ __forceinline__ const double *lvl_r(const HcpDev &H, int l)
{
    return l == 0 ? H.rc1 : (l == 1 ? H.rc2 : H.rc3);
}
__device__ __forceinline__ const double *lvl_reff(const HcpDev &H, int l)
{
    return l == 0 ? H.rf1 : (l == 1 ? H.rf2 : H.rf3);
}

__device__ __forceinline__ void ene_add(const DevPtrs &d, int term, double v)
{
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) atomicAdd(&d.eslot[(size_t)term * ENE_SLOTS + (blockIdx.x & (ENE_SLOTS - 1))], v);
}

/* hcp_gb.c:179-430: descreening sums of each atom over one residue range */
template <typename R>
__global__ __launch_bounds__(64) void k_hcp_born(HcpDev H, const double *__restrict__ x)
{
    const HcpItem it = hcp_item(H);
    const int i = it.i;
    const double xi = x[3 * (size_t)i], yi = x[3 * (size_t)i + 1], zi = x[3 * (size_t)i + 2];
    const R ri = (R)H.atom[i].y, ri1i = R(1) / ri;
    const int nq = H.nq;
    double sumi = 0.0;
    hcp_walk(
        H, it.r0, it.r1, it.valid, i, xi, yi, zi,
        [&](int l, int comp) {
            const double *qc = lvl_q(H, l) + 4 * (size_t)nq * comp;
            const double *rc = lvl_r(H, l) + (size_t)nq * comp;
            for (int q = 0; q < nq; q++) sumi -= far_descreen<R>(hcp_d2(xi, yi, zi, qc + 4 * q), rc[q]);
        },
        [&](int a) { sumi -= near_descreen<R>(hcp_d2(xi, yi, zi, x + 3 * (size_t)a), ri, ri1i, H.atom[a].z); });
    if (it.valid) hcp_part(H, it, 0) = sumi;
}

/*
 * per atom: the Born radius from its descreening sum (hcp_gb.c:395-425),
 * and the diagonal GB term with its sumdeijda part (hcp_gb.c:1877-1888),
 * which start the sums the energy kernel adds to
 */
__global__ __launch_bounds__(HCP_THREADS) void k_hcp_radii(HcpDev H, DevPtrs d)
{
    const int t = blockIdx.x * HCP_THREADS + threadIdx.x;
    const bool valid = t < H.natom;
    const int i = valid ? t : H.natom - 1;
    const double4 ai = H.atom[i];
    const size_t np = (size_t)H.ntile * 64;
    double sumi = 0.0;
    for (int k = 0; k < H.nchunk; k++) sumi += H.part[(size_t)k * np + i];
    const double ri = ai.y, ri1i = 1. / ri;
    double reff, psi = 0.0;
    if (H.gb == 1) {
        reff = 1.0 / (ri1i + sumi);
        if (reff < 0.0) reff = 30.0;
    } else {
        const double4 g = H.obc[i];
        psi = -ri * sumi;
        reff = 1.0 / (ri1i - tanh((g.x - g.y * psi + g.z * psi * psi) * psi) / ai.w);
    }
    double ep = 0.0, sd = 0.0;
    if (valid && !H.frz[i]) {
        const double qi = ai.x, kappa = H.kappa;
        const double expmkf = exp(-HCP_KSCALE * kappa * reff) / H.epsext, dielfac = 1.0 - expmkf;
        const double qi2h = 0.5 * qi * qi, qid2h = qi2h * dielfac;
        ep = -qid2h * 2.0 / reff;
        sd = qid2h - HCP_KSCALE * kappa * qi2h * expmkf * reff;
    }
    if (valid) {
        H.reff[i] = reff;
        H.psi[i] = psi;
        H.sd[i] = sd;
        H.gm[i] = 0.0; /* the chain rule's member forces start here */
        H.gm[H.natom + i] = 0.0;
        H.gm[2 * (size_t)H.natom + i] = 0.0;
        const int sl = H.inv[i];
        d.reff[sl] = reff;
        d.psi[sl] = psi;
    }
    ene_add(d, 0, 0.5 * ep);
}

/* ------------------------------------------- component Born radii */

__device__ __forceinline__ double avg_radius(double w, double q, bool guard)
{
    if (guard && fabs(w) < 1) return 2.3;
    return (q * q) / (w * w);
}

/*
 * per residue: its charges' Born radii (hcp_gb.c:780-818 / :913-964) and
 * its terms of the strand averages, rw[r] = (target, weight, charge) for
 * each of its charges in the order the reference adds them
 */
template <bool OPCA>
__global__ __launch_bounds__(HCP_THREADS) void k_hcp_reff_res(HcpDev H, const double *__restrict__ x)
{
    const int r = blockIdx.x * HCP_THREADS + threadIdx.x;
    if (r >= H.nres) return;
    const int nq = H.nq;
    double w1[2] = {0, 0}, q1[2] = {0, 0};
    const double *qc = H.qc1 + 4 * (size_t)nq * r;
    for (int a = H.ipres[r]; a < H.ipres[r + 1]; a++) {
        const double q = H.atom[a].x;
        const int m = (nq == 2 && q > 0.0) ? 1 : 0;
        if (OPCA) {
            const double *c = qc + 4 * m;
            const double dx = x[3 * (size_t)a] - c[0], dy = x[3 * (size_t)a + 1] - c[1],
                         dz = x[3 * (size_t)a + 2] - c[2];
            const double ql = fabs(q) / sqrt(dx * dx + dy * dy + dz * dz);
            w1[m] += ql / sqrt(H.reff[a]);
            q1[m] += ql;
        } else {
            w1[m] += q / sqrt(H.reff[a]);
            q1[m] += q;
        }
    }
    double rr[2];
    for (int m = 0; m < nq; m++) rr[m] = avg_radius(w1[m], q1[m], true);
    if (OPCA) {
        /* the strand terms depend on the strand's charges: k_hcp_reff_up */
        if (qc[3] > qc[7]) {
            const double t = rr[0];
            rr[0] = rr[1];
            rr[1] = t;
        }
    } else {
        double *rw = H.rw + 6 * (size_t)r;
        for (int m = 0; m < nq; m++) {
            rw[3 * m] = m;
            rw[3 * m + 1] = w1[m];
            rw[3 * m + 2] = q1[m];
        }
    }
    for (int m = 0; m < nq; m++) H.rf1[(size_t)nq * r + m] = rr[m];
}

/*
 * strands and complexes (hcp_gb.c:820-858 / :966-1022; one workgroup).  As
 * there, every complex sums over every strand of the molecule, and the
 * strand and complex averages of hcp 1/2 have no small-weight guard.
 */
template <bool OPCA>
__global__ __launch_bounds__(HCP_THREADS) void k_hcp_reff_up(HcpDev H)
{
    const int nq = H.nq;
    for (int st = threadIdx.x; st < H.nstrand; st += HCP_THREADS) {
        double w2[2] = {0, 0}, q2[2] = {0, 0};
        const double *qs = H.qc2 + 4 * (size_t)nq * st;
        for (int r = H.rfirst[st]; r < H.rfirst[st + 1]; r++) {
            if (OPCA) {
                const double *qr = H.qc1 + 8 * (size_t)r;
                for (int m = 0; m < 2; m++) {
                    const double *a1 = qr + 4 * m;
                    const int nn = a1[3] > 0 ? 1 : 0;
                    const double *b = qs + 4 * nn;
                    const double d2 = (a1[0] - b[0]) * (a1[0] - b[0]) + (a1[1] - b[1]) * (a1[1] - b[1]) +
                                      (a1[2] - b[2]) * (a1[2] - b[2]);
                    const double ql = a1[3] / sqrt(d2);
                    w2[nn] += ql / sqrt(H.rf1[2 * (size_t)r + m]);
                    q2[nn] += ql;
                }
            } else {
                const double *rw = H.rw + 6 * (size_t)r;
                for (int m = 0; m < nq; m++) {
                    w2[m] += rw[3 * m + 1];
                    q2[m] += rw[3 * m + 2];
                }
            }
        }
        double *sw = H.sw + 4 * (size_t)st;
        sw[0] = w2[0]; sw[1] = w2[1]; sw[2] = q2[0]; sw[3] = q2[1];
        for (int m = 0; m < nq; m++) H.rf2[(size_t)nq * st + m] = avg_radius(w2[m], q2[m], OPCA);
    }
    __syncthreads();
    for (int cx = threadIdx.x; cx < H.ncomplex; cx += HCP_THREADS) {
        double w3[2] = {0, 0}, q3[2] = {0, 0};
        const double *qcx = H.qc3 + 4 * (size_t)nq * cx;
        for (int st = 0; st < H.nstrand; st++) {
            if (OPCA) {
                for (int m = 0; m < 2; m++) {
                    const double *a2 = H.qc2 + 4 * (2 * (size_t)st + m);
                    const int nn = a2[3] > 0 ? 1 : 0;
                    const double *b = qcx + 4 * nn;
                    const double d2 = (a2[0] - b[0]) * (a2[0] - b[0]) + (a2[1] - b[1]) * (a2[1] - b[1]) +
                                      (a2[2] - b[2]) * (a2[2] - b[2]);
                    const double ql = a2[3] / sqrt(d2);
                    w3[nn] += ql / sqrt(H.rf2[2 * (size_t)st + m]);
                    q3[nn] += ql;
                }
            } else {
                const double *sw = H.sw + 4 * (size_t)st;
                for (int m = 0; m < nq; m++) {
                    w3[m] += sw[m];
                    q3[m] += sw[2 + m];
                }
            }
        }
        for (int m = 0; m < nq; m++) H.rf3[(size_t)nq * cx + m] = avg_radius(w3[m], q3[m], OPCA);
    }
}

/* ------------------------------------------------ energies and forces */

/* hcp_gb.c:1843-1976 (egb_hcp_nbond, step 1) over one residue range:
 * energies, sumdeijda and the pair forces on each atom of the tile */
template <typename R, bool KAPPA>
__global__ __launch_bounds__(64) void k_hcp_energy(HcpDev H, const double *__restrict__ x, DevPtrs d)
{
    const HcpItem it = hcp_item(H);
    const int i = it.i;
    const double xi = x[3 * (size_t)i], yi = x[3 * (size_t)i + 1], zi = x[3 * (size_t)i + 2];
    const double qi = H.atom[i].x;
    const R ri = (R)H.reff[i], kks = (R)(-HCP_KSCALE * H.kappa), inv_eps = (R)(1.0 / H.epsext);
    const int2 lji = H.lj[i];
    const int nq = H.nq;
    int ep = H.ex_ptr[i];
    const int ee = H.ex_ptr[i + 1];
    double sd = 0.0, eg = 0.0, el = 0.0, ev = 0.0, dx = 0.0, dy = 0.0, dz = 0.0;
    hcp_walk_lanes(
        H, it.r0, it.r1, it.valid, i, xi, yi, zi,
        [&](int l, int comp) {
            const double *qc = lvl_q(H, l) + 4 * (size_t)nq * comp;
            const double *rf = lvl_reff(H, l) + (size_t)nq * comp;
            for (int q = 0; q < nq; q++) {
                const double *qv = qc + 4 * q;
                if (fabs(qv[3]) > HCP_MIN_Q) {
                    const double xij = xi - qv[0], yij = yi - qv[1], zij = zi - qv[2];
                    const double r2 = xij * xij + yij * yij + zij * zij;
                    const double de = gb_pair<R, KAPPA>(qi, ri, qv[3], (R)rf[q], r2, kks, inv_eps, eg, el, sd, true);
                    dx += de * xij;
                    dy += de * yij;
                    dz += de * zij;
                }
            }
        },
        [&](int a) {
            /* each lane meets its atoms in increasing order: a pointer into
             * the ascending excluded-partner list finds them */
            while (ep < ee && H.ex_j[ep] < a) ep++;
            const bool excl = ep < ee && H.ex_j[ep] == a;
            const double xij = xi - x[3 * (size_t)a], yij = yi - x[3 * (size_t)a + 1], zij = zi - x[3 * (size_t)a + 2];
            const double r2 = xij * xij + yij * yij + zij * zij;
            double de = gb_pair<R, KAPPA>(qi, ri, H.atom[a].x, (R)H.reff[a], r2, kks, inv_eps, eg, el, sd, !excl);
            if (!excl) {
                const double2 ab = H.ljab[lji.x + H.lj[a].y];
                const double rinv = hcp_rinv<R>(r2), r2inv = rinv * rinv, r6inv = r2inv * r2inv * r2inv;
                const double f6 = ab.y * r6inv, f12 = ab.x * r6inv * r6inv;
                ev += f12 - f6;
                de -= (12. * f12 - 6. * f6) * r2inv;
            }
            dx += de * xij;
            dy += de * yij;
            dz += de * zij;
        });
    if (it.valid) {
        hcp_part(H, it, 0) = sd;
        hcp_part(H, it, 1) = dx;
        hcp_part(H, it, 2) = dy;
        hcp_part(H, it, 3) = dz;
    }
    ene_add(d, 0, it.valid ? 0.5 * eg : 0.0);
    ene_add(d, 1, it.valid ? 0.5 * ev : 0.0);
    ene_add(d, 2, it.valid ? 0.5 * el : 0.0);
}

/*
 * hcp_gb.c:1978-2097 (egb_hcp_nbond, step 2) over one residue range.  A
 * near residue's atoms are staged in LDS, 64 at a time, and the lanes step
 * through them in rotated order (lane l takes member (l + t) mod m at step
 * t), so that each step's partner-side terms go to different LDS sums;
 * each member's sum is added with one atomic per residue visit to a
 * per-atom array in original order (consecutive lanes, consecutive
 * addresses; through the slot map they had been 64 scattered addresses
 * per wavefront), which k_hcp_sum adds to the gradient.
 */
template <typename R>
__global__ __launch_bounds__(64) void k_hcp_deriv(HcpDev H, const double *__restrict__ x, DevPtrs d)
{
    __shared__ double s_pos[4][64]; /* x, y, z, scaled radius of the staged members */
    __shared__ double s_acc[3][64];
    const HcpItem it = hcp_item(H);
    const int i = it.i, lane = threadIdx.x;
    const bool valid = it.valid && !H.frz[i];
    const double xi = x[3 * (size_t)i], yi = x[3 * (size_t)i + 1], zi = x[3 * (size_t)i + 2];
    const double4 ai = H.atom[i];
    const R ri = (R)ai.y, ri1i = R(1) / ri;
    double sumda = H.sd[i];
    if (it.valid && it.r0 == 0) d.sumdeijda[H.inv[i]] = sumda;
    if (H.gb > 1) {
        const double4 g = H.obc[i];
        const double ps = H.psi[i];
        const double thi = tanh((g.x - g.y * ps + g.z * ps * ps) * ps);
        sumda *= (g.x - 2.0 * g.y * ps + 3.0 * g.z * ps * ps) * (1.0 - thi * thi) * ai.y / ai.w;
    }
    const int nq = H.nq;
    double dx = 0.0, dy = 0.0, dz = 0.0;
    hcp_walk_res(
        H, it.r0, it.r1, valid, xi, yi, zi,
        [&](int l, int comp) {
            const double *qc = lvl_q(H, l) + 4 * (size_t)nq * comp;
            const double *rc = lvl_r(H, l) + (size_t)nq * comp;
            for (int q = 0; q < nq; q++) {
                const double *qv = qc + 4 * q;
                const double xij = xi - qv[0], yij = yi - qv[1], zij = zi - qv[2];
                const double r2 = xij * xij + yij * yij + zij * zij;
                const double datmp = descreen_deriv<R>(r2, ri, ri1i, rc[q]) * (sumda * 2);
                dx += xij * datmp;
                dy += yij * datmp;
                dz += zij * datmp;
            }
        },
        [&](int r, bool in1) {
            for (int base = H.ipres[r]; base < H.ipres[r + 1]; base += 64) {
                const int m = min(64, H.ipres[r + 1] - base);
                __syncthreads();
                if (lane < m) {
                    const int a = base + lane;
                    s_pos[0][lane] = x[3 * (size_t)a];
                    s_pos[1][lane] = x[3 * (size_t)a + 1];
                    s_pos[2][lane] = x[3 * (size_t)a + 2];
                    s_pos[3][lane] = H.atom[a].z;
                    s_acc[0][lane] = 0.0;
                    s_acc[1][lane] = 0.0;
                    s_acc[2][lane] = 0.0;
                }
                __syncthreads();
                if (in1) {
                    int k = lane % m;
                    for (int t = 0; t < m; t++) {
                        if (base + k != i) {
                            const double xij = xi - s_pos[0][k], yij = yi - s_pos[1][k], zij = zi - s_pos[2][k];
                            const double r2 = xij * xij + yij * yij + zij * zij;
                            const double datmp = descreen_deriv<R>(r2, ri, ri1i, s_pos[3][k]) * sumda;
                            const double cx = xij * datmp, cy = yij * datmp, cz = zij * datmp;
                            dx += cx;
                            dy += cy;
                            dz += cz;
                            atomicAdd(&s_acc[0][k], cx);
                            atomicAdd(&s_acc[1][k], cy);
                            atomicAdd(&s_acc[2][k], cz);
                        }
                        k = k + 1 == m ? 0 : k + 1;
                    }
                }
                __syncthreads();
                if (lane < m) {
                    /* original order: consecutive lanes, consecutive addresses */
                    const int a = base + lane;
                    const double ax = s_acc[0][lane], ay = s_acc[1][lane], az = s_acc[2][lane];
                    if (ax != 0.0 || ay != 0.0 || az != 0.0) {
                        atomicAdd(&H.gm[a], ax);
                        atomicAdd(&H.gm[H.natom + a], ay);
                        atomicAdd(&H.gm[2 * (size_t)H.natom + a], az);
                    }
                }
            }
        });
    if (it.valid) {
        hcp_part(H, it, 0) = -dx;
        hcp_part(H, it, 1) = -dy;
        hcp_part(H, it, 2) = -dz;
    }
}

/* per atom: the partial sums of the energy (SD: sd, then the force) or
 * chain-rule pass (with the member forces) over its tile's residue ranges,
 * in range order, added to sd and (one atomic per component) to the
 * gradient */
template <bool SD>
__global__ __launch_bounds__(HCP_THREADS) void k_hcp_sum(HcpDev H, DevPtrs d)
{
    const int a = blockIdx.x * HCP_THREADS + threadIdx.x;
    if (a >= H.natom) return;
    const size_t np = (size_t)H.ntile * 64, cs = (size_t)H.nchunk * np;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < H.nchunk; k++) {
        const double *p = H.part + (size_t)k * np + a;
        for (int c = 0; c < (SD ? 4 : 3); c++) s[c] += p[c * cs];
    }
    if (!SD)
        for (int c = 0; c < 3; c++) s[c] += H.gm[(size_t)c * H.natom + a];
    const double *g = SD ? s + 1 : s;
    if (SD) H.sd[a] += s[0];
    const int sl = H.inv[a];
    if (g[0] != 0.0) atomicAdd(&d.gx[sl], g[0]);
    if (g[1] != 0.0) atomicAdd(&d.gy[sl], g[1]);
    if (g[2] != 0.0) atomicAdd(&d.gz[sl], g[2]);
}

/* ------------------------------------------------------------ host side */

int engine_eval_hcp(Engine &E, const double *x_dev, double *g_dev, int flags)
{
    if (!(flags & EV_GATHERED)) {
        int rc = engine_gather(E, x_dev);
        if (rc) return rc;
    }
    const HcpDev &H = E.hcp;
    hipEvent_t *ev = nullptr;
    if (E.tev_used + 5 <= E.tev_cap) {
        ev = &E.tev[E.tev_used];
        E.tev_used += 5;
    }
    const int gres = (H.nres + HCP_THREADS - 1) / HCP_THREADS, gat = (H.natom + HCP_THREADS - 1) / HCP_THREADS;
    if (ev) HIPCHK(hipEventRecord(ev[0], E.stream));
    if (H.opca) {
        hipLaunchKernelGGL(k_hcp_res<true>, dim3(gres), dim3(HCP_THREADS), 0, E.stream, H, x_dev);
        hipLaunchKernelGGL(k_hcp_up<true>, dim3(1), dim3(HCP_THREADS), 0, E.stream, H);
    } else {
        hipLaunchKernelGGL(k_hcp_res<false>, dim3(gres), dim3(HCP_THREADS), 0, E.stream, H, x_dev);
        hipLaunchKernelGGL(k_hcp_up<false>, dim3(1), dim3(HCP_THREADS), 0, E.stream, H);
    }
    const unsigned items = (unsigned)H.ntile * (unsigned)H.nchunk;
    const bool mixed = E.precision == GBHIP_MIXED;
    if (mixed) hipLaunchKernelGGL(k_hcp_born<float>, dim3(items), dim3(64), 0, E.stream, H, x_dev);
    else hipLaunchKernelGGL(k_hcp_born<double>, dim3(items), dim3(64), 0, E.stream, H, x_dev);
    hipLaunchKernelGGL(k_hcp_radii, dim3(gat), dim3(HCP_THREADS), 0, E.stream, H, E.d);
    if (ev) HIPCHK(hipEventRecord(ev[1], E.stream));
    if (H.opca) {
        hipLaunchKernelGGL(k_hcp_reff_res<true>, dim3(gres), dim3(HCP_THREADS), 0, E.stream, H, x_dev);
        hipLaunchKernelGGL(k_hcp_reff_up<true>, dim3(1), dim3(HCP_THREADS), 0, E.stream, H);
    } else {
        hipLaunchKernelGGL(k_hcp_reff_res<false>, dim3(gres), dim3(HCP_THREADS), 0, E.stream, H, x_dev);
        hipLaunchKernelGGL(k_hcp_reff_up<false>, dim3(1), dim3(HCP_THREADS), 0, E.stream, H);
    }
    if (ev) HIPCHK(hipEventRecord(ev[2], E.stream));
    if (mixed && H.kappa != 0.0)
        hipLaunchKernelGGL((k_hcp_energy<float, true>), dim3(items), dim3(64), 0, E.stream, H, x_dev, E.d);
    else if (mixed)
        hipLaunchKernelGGL((k_hcp_energy<float, false>), dim3(items), dim3(64), 0, E.stream, H, x_dev, E.d);
    else if (H.kappa != 0.0)
        hipLaunchKernelGGL((k_hcp_energy<double, true>), dim3(items), dim3(64), 0, E.stream, H, x_dev, E.d);
    else
        hipLaunchKernelGGL((k_hcp_energy<double, false>), dim3(items), dim3(64), 0, E.stream, H, x_dev, E.d);
    /* the energy pass is timed alone; its sums count with the chain rule's */
    if (ev) HIPCHK(hipEventRecord(ev[3], E.stream));
    hipLaunchKernelGGL(k_hcp_sum<true>, dim3(gat), dim3(HCP_THREADS), 0, E.stream, H, E.d);
    if (mixed) hipLaunchKernelGGL(k_hcp_deriv<float>, dim3(items), dim3(64), 0, E.stream, H, x_dev, E.d);
    else hipLaunchKernelGGL(k_hcp_deriv<double>, dim3(items), dim3(64), 0, E.stream, H, x_dev, E.d);
    hipLaunchKernelGGL(k_hcp_sum<false>, dim3(gat), dim3(HCP_THREADS), 0, E.stream, H, E.d);
    if (ev) HIPCHK(hipEventRecord(ev[4], E.stream));
    if (!(flags & EV_NO_POST)) {
        const int nene = 3 * E.nmol, nt = std::max(64 * nene, g_dev ? E.nslots : 0);
        int rc = engine_post(E, g_dev, nene, nt);
        if (rc) return rc;
    }
    LAUNCHCHK();
    E.have_eval = true;
    return 0;
}

/* host cube sums of the scaled radii in single precision, rolled up
 * (egb_hcp_rbondi, hcp_gb.c:45-124) */
static void hcp_radii(const std::vector<int> &ipres, const std::vector<int> &rfirst, const std::vector<int> &sfirst,
                      const std::vector<double> &sj, const std::vector<double> &q, int nq, std::vector<double> &r1,
                      std::vector<double> &r2, std::vector<double> &r3)
{
    const int ncomplex = (int)sfirst.size() - 1;
    for (int c = 0; c < ncomplex; c++) {
        for (int k = 0; k < nq; k++) r3[(size_t)c * nq + k] = 0.0;
        for (int s = sfirst[c]; s < sfirst[c + 1]; s++) {
            for (int k = 0; k < nq; k++) r2[(size_t)s * nq + k] = 0.0;
            for (int r = rfirst[s]; r < rfirst[s + 1]; r++) {
                for (int k = 0; k < nq; k++) r1[(size_t)r * nq + k] = 0.0;
                for (int a = ipres[r]; a < ipres[r + 1]; a++) {
                    const float t = (float)sj[a];
                    const float t3 = t * t * t;
                    r1[(size_t)r * nq + ((nq == 2 && q[a] > 0.0) ? 1 : 0)] += (double)t3;
                }
                for (int k = 0; k < nq; k++) {
                    r2[(size_t)s * nq + k] += r1[(size_t)r * nq + k];
                    r1[(size_t)r * nq + k] = std::cbrt(r1[(size_t)r * nq + k]);
                }
            }
            for (int k = 0; k < nq; k++) {
                r3[(size_t)c * nq + k] += r2[(size_t)s * nq + k];
                r2[(size_t)s * nq + k] = std::cbrt(r2[(size_t)s * nq + k]);
            }
        }
        for (int k = 0; k < nq; k++) r3[(size_t)c * nq + k] = std::cbrt(r3[(size_t)c * nq + k]);
    }
}

extern "C" int gbhip_set_hcp(gbhip_ctx *ctx, const gbhip_hcp *h)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    Engine &E = ctx->E;
    if (!h || h->hcp == 0) {
        if (E.hcp.on) {
            E.hcp.on = 0;
            E.have_list = false;
        }
        return 0;
    }
    if (h->hcp != 1 && h->hcp != 2 && h->hcp != 4)
        return seterr(GBHIP_EINVAL, "hcp must be 0, 1, 2 or 4 (mm_options takes these), got %d", h->hcp);
    if (E.nmol != 1) return seterr(GBHIP_EINVAL, "hcp applies to a single-molecule context");
    if (E.par.det) return seterr(GBHIP_EINVAL, "hcp has no deterministic mode (its forces add with atomics)");
    if (ctx->M.g_step || ctx->M.g_rebuild) return seterr(GBHIP_ESTATE, "hcp cannot be set while dynamics graphs exist");
    if (!(h->h1 > 0.0) || !(h->h2 > 0.0) || !(h->h3 > 0.0)) return seterr(GBHIP_EINVAL, "hcp thresholds must be > 0");
    const int n = E.natoms;
    if (h->nres < 1 || !h->ipres || h->ipres[0] != 0 || h->ipres[h->nres] != n)
        return seterr(GBHIP_EINVAL, "hcp residues must start at atom 0 and end at natom (ipres[nres] = %d)", n);
    for (int r = 0; r < h->nres; r++)
        if (h->ipres[r + 1] < h->ipres[r]) return seterr(GBHIP_EINVAL, "hcp residue pointers must not decrease");
    if (h->nstrand < 1 || !h->ipstrand || h->ipstrand[0] != 1 || h->ncomplex < 1 || !h->ipcomplex ||
        h->ipcomplex[0] != 1)
        return seterr(GBHIP_EINVAL, "hcp strands and complexes must start at 1");
    for (int s = 1; s < h->nstrand; s++)
        if (h->ipstrand[s] <= h->ipstrand[s - 1] || h->ipstrand[s] > h->nres)
            return seterr(GBHIP_EINVAL, "hcp strand pointers must increase within 1..nres");
    for (int c = 1; c < h->ncomplex; c++)
        if (h->ipcomplex[c] <= h->ipcomplex[c - 1] || h->ipcomplex[c] > h->nstrand)
            return seterr(GBHIP_EINVAL, "hcp complex pointers must increase within 1..nstrand");
    if (h->nres > (1 << 28) / 8) return seterr(GBHIP_EINVAL, "too many residues");

    /* per-atom constants from the engine's slot arrays, original order */
    const int NS = E.nslots;
    std::vector<double> q(NS), ri0(NS), sj(NS), rb(NS), ga(NS), gb(NS), gg(NS);
    std::vector<int> frz(NS), ljrow(NS), ljcol(NS), inv(n), exp_(n + 1), exj;
    int rc;
    if ((rc = engine_d2h(E, q.data(), E.d.q, NS * sizeof(double))) ||
        (rc = engine_d2h(E, ri0.data(), E.d.ri0, NS * sizeof(double))) ||
        (rc = engine_d2h(E, sj.data(), E.d.sj, NS * sizeof(double))) ||
        (rc = engine_d2h(E, rb.data(), E.d.rb, NS * sizeof(double))) ||
        (rc = engine_d2h(E, ga.data(), E.d.ga, NS * sizeof(double))) ||
        (rc = engine_d2h(E, gb.data(), E.d.gbb, NS * sizeof(double))) ||
        (rc = engine_d2h(E, gg.data(), E.d.gg, NS * sizeof(double))) ||
        (rc = engine_d2h(E, frz.data(), E.d.frz, NS * sizeof(int))) ||
        (rc = engine_d2h(E, ljrow.data(), E.d.ljrow, NS * sizeof(int))) ||
        (rc = engine_d2h(E, ljcol.data(), E.d.ljcol, NS * sizeof(int))) ||
        (rc = engine_d2h(E, exp_.data(), E.d.ex_ptr, (n + 1) * sizeof(int))))
        return rc;
    exj.resize(std::max(1, exp_[n]));
    if (exp_[n] > 0 && (rc = engine_d2h(E, exj.data(), E.d.ex_j, exp_[n] * sizeof(int)))) return rc;
    std::vector<double4> atom(n), obc(n);
    std::vector<int2> lj(n);
    std::vector<int> afrz(n);
    std::vector<double> sj_a(n), q_a(n);
    for (int s = 0; s < NS; s++) {
        const int a = E.perm_h[s];
        if (a < 0) continue;
        inv[a] = s;
        atom[a] = make_double4(q[s], ri0[s], sj[s], rb[s]);
        obc[a] = make_double4(ga[s], gb[s], gg[s], 0.0);
        lj[a] = make_int2(ljrow[s], ljcol[s]);
        afrz[a] = frz[s];
        sj_a[a] = sj[s];
        q_a[a] = q[s];
    }
    /* excluded pairs in both directions, ascending per atom (sff.c:1644-1679) */
    std::vector<std::vector<int>> ex(n);
    for (int i = 0; i < n; i++)
        for (int k = exp_[i]; k < exp_[i + 1]; k++) {
            ex[i].push_back(exj[k]);
            ex[exj[k]].push_back(i);
        }
    std::vector<int> hex_ptr(n + 1, 0), hex_j;
    for (int i = 0; i < n; i++) {
        std::sort(ex[i].begin(), ex[i].end());
        ex[i].erase(std::unique(ex[i].begin(), ex[i].end()), ex[i].end());
        hex_ptr[i + 1] = hex_ptr[i] + (int)ex[i].size();
        hex_j.insert(hex_j.end(), ex[i].begin(), ex[i].end());
    }
    if (hex_j.empty()) hex_j.push_back(0);
    std::vector<int> ipres(h->ipres, h->ipres + h->nres + 1), rfirst(h->nstrand + 1), sfirst(h->ncomplex + 1);
    for (int s = 0; s < h->nstrand; s++) rfirst[s] = h->ipstrand[s] - 1;
    rfirst[h->nstrand] = h->nres;
    for (int c = 0; c < h->ncomplex; c++) sfirst[c] = h->ipcomplex[c] - 1;
    sfirst[h->ncomplex] = h->nstrand;
    const int nq = h->hcp == 1 ? 1 : 2;
    std::vector<double> r1((size_t)h->nres * nq), r2((size_t)h->nstrand * nq), r3((size_t)h->ncomplex * nq);
    hcp_radii(ipres, rfirst, sfirst, sj_a, q_a, nq, r1, r2, r3);
    std::vector<int> res_s(h->nres), s_cx(h->nstrand);
    for (int s = 0; s < h->nstrand; s++)
        for (int r = rfirst[s]; r < rfirst[s + 1]; r++) res_s[r] = s;
    for (int c = 0; c < h->ncomplex; c++)
        for (int s = sfirst[c]; s < sfirst[c + 1]; s++) s_cx[s] = c;
    /* work items: 64-atom tiles x ranges of csize residues, about 256 per CU
     * (ranges of nearby and of distant residues differ several-fold in
     * work).  Mixed hcp 1, ms per evaluation at 25k / 100k / 200k atoms for
     * 128 / 192 / 256 / 384 / 512 / 1024 items per CU: 1.06 / 3.84 / 7.99,
     * 1.09 / 3.80 / 7.88, 1.13 / 3.72 / 7.86, 1.25 / 3.77 / 7.78,
     * 1.34 / 3.79 / 7.76, 1.34 / 4.24 / 7.90 (the partial-sum buffer and
     * its sums grow with the item count) */
    const int ntile = (n + 63) / 64;
    const long long target = 256LL * std::max(1, E.ncu);
    const long long want_chunks = std::max(1LL, (target + ntile - 1) / ntile);
    int csize = (int)std::max(4LL, ((long long)h->nres + want_chunks - 1) / want_chunks);
    csize = std::min(csize, h->nres);
    const int nchunk = (h->nres + csize - 1) / csize;
    if ((long long)ntile * nchunk > 0x7fffffffLL) return seterr(GBHIP_EINVAL, "too many hcp work items");

    HcpDev H;
    H.nq = nq;
    H.opca = h->hcp == 4;
    H.gb = E.par.gb;
    H.natom = n;
    H.nres = h->nres;
    H.nstrand = h->nstrand;
    H.ncomplex = h->ncomplex;
    H.dhcp = (float)h->dhcp;
    H.h1sq = h->h1 * h->h1;
    H.h2sq = h->h2 * h->h2;
    H.h3sq = h->h3 * h->h3;
    H.kappa = E.par.kappa;
    H.epsext = E.par.epsext;
    H.ljab = E.d.ljab;
    H.ntile = ntile;
    H.nchunk = nchunk;
    H.csize = csize;
    const size_t w = 3 + 4 * (size_t)nq;
    std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
    if ((rc = E.upload(&H.ipres, ipres)) || (rc = E.upload(&H.rfirst, rfirst)) || (rc = E.upload(&H.sfirst, sfirst)) ||
        (rc = E.upload(&H.atom, atom)) || (rc = E.upload(&H.obc, obc)) || (rc = E.upload(&H.lj, lj)) ||
        (rc = E.upload(&H.frz, afrz)) || (rc = E.upload(&H.ex_ptr, hex_ptr)) || (rc = E.upload(&H.ex_j, hex_j)) ||
        (rc = E.upload(&H.inv, inv)) || (rc = E.upload(&H.rc1, r1)) || (rc = E.upload(&H.rc2, r2)) ||
        (rc = E.upload(&H.rc3, r3)) || (rc = E.dalloc(&H.xc1, 3 * (size_t)h->nres)) ||
        (rc = E.dalloc(&H.xc2, 3 * (size_t)h->nstrand)) || (rc = E.dalloc(&H.xc3, 3 * (size_t)h->ncomplex)) ||
        (rc = E.dalloc(&H.qc1, 4 * (size_t)nq * h->nres)) || (rc = E.dalloc(&H.qc2, 4 * (size_t)nq * h->nstrand)) ||
        (rc = E.dalloc(&H.qc3, 4 * (size_t)nq * h->ncomplex)) || (rc = E.dalloc(&H.rf1, (size_t)nq * h->nres)) ||
        (rc = E.dalloc(&H.rf2, (size_t)nq * h->nstrand)) || (rc = E.dalloc(&H.rf3, (size_t)nq * h->ncomplex)) ||
        (rc = E.dalloc(&H.rsum, w * h->nres)) || (rc = E.dalloc(&H.ssum, w * h->nstrand)) ||
        (rc = E.dalloc(&H.scnt, (size_t)h->nstrand)) || (rc = E.dalloc(&H.rmp, (size_t)MP_N * h->nres)) ||
        (rc = E.dalloc(&H.smp, (size_t)MP_N * h->nstrand)) || (rc = E.dalloc(&H.rw, 6 * (size_t)h->nres)) ||
        (rc = E.dalloc(&H.sw, 4 * (size_t)h->nstrand)) || (rc = E.dalloc(&H.reff, (size_t)n)) ||
        (rc = E.dalloc(&H.psi, (size_t)n)) || (rc = E.dalloc(&H.sd, (size_t)n)) ||
        (rc = E.dalloc(&H.part, 4 * (size_t)nchunk * ntile * 64)) || (rc = E.dalloc(&H.gm, 3 * (size_t)n)) || (rc = E.upload(&H.res_s, res_s)) || (rc = E.upload(&H.s_cx, s_cx)))
        return rc;
    if ((rc = E.upload_done())) return rc;
    H.on = 1;
    E.hcp = H;
    E.have_list = true; /* HCP builds no pair list (eff.c:2560-2561) */
    return 0;
}
