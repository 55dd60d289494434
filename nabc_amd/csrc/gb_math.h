/*
 * gb_math.h -- per-pair generalized-Born arithmetic shared by the FP64 and
 * the mixed-precision kernels (template parameter R is double or float).
 *
 * Formulas follow egb() in src/sff/eff.c:
 *   descreen_d() pass 1 term, eff.c:1663-1697 (Schaefer-Froemmel appendix,
 *                Taylor series for d > 4 s, smooth rgbmax switch), and its
 *                pass 3 derivative numerator, eff.c:2289-2334
 *   neck_d()     gb 7/8 neck correction, eff.c:1699-1708, and its
 *                derivative, eff.c:2336-2360
 *   ddescreen(), dneck_term(): the derivative parts alone (pass 3 when it
 *                recomputes them)
 * The value returned by descreen_d() is the amount egb() subtracts from
 * sumi; the derivative is datmp before the multiplication by sumdeijda.
 */
#ifndef NABC_AMD_GB_MATH_H
#define NABC_AMD_GB_MATH_H

#include <hip/hip_runtime.h>

/* Taylor coefficients k/(2k+1) and 2k(k+1)/(2k+1)... of the d >> s series */
#define GB_TA (1.0 / 3.0)
#define GB_TB (2.0 / 5.0)
#define GB_TC (3.0 / 7.0)
#define GB_TD (4.0 / 9.0)
#define GB_TDD (5.0 / 11.0)
#define GB_TE (4.0 / 3.0)
#define GB_TF (12.0 / 5.0)
#define GB_TG (24.0 / 7.0)
#define GB_TH (40.0 / 9.0)
#define GB_THH (60.0 / 11.0)
#define GB_KSCALE 0.73

template <typename R> struct gbm;

/*
 * FP64: exp/log from the device math library; reciprocal and reciprocal
 * square root from the hardware estimates (v_rcp_f64 / v_rsq_f64, ~2^-26
 * relative) refined by Newton steps to within an ulp or two of the IEEE
 * result, instead of the correctly rounded division sequence.
 */
template <> struct gbm<double> {
    static __device__ __forceinline__ double log_(double x) { return log(x); }
    static __device__ __forceinline__ double exp_(double x) { return exp(x); }
    static __device__ __forceinline__ double rcp(double x)
    {
        double y = __builtin_amdgcn_rcp(x);
        double e = fma(-x, y, 1.0);
        y = fma(y, e, y);
        e = fma(-x, y, 1.0);
        return fma(y, e, y);
    }
    static __device__ __forceinline__ double rsqrt_(double x)
    {
        double y = __builtin_amdgcn_rsq(x);
        const double h = 0.5 * x;
        double e = fma(-h * y, y, 0.5);
        y = fma(y, e, y);
        e = fma(-h * y, y, 0.5);
        return fma(y, e, y);
    }
};

/*
 * FP32: the hardware transcendental instructions (v_exp_f32, v_log_f32,
 * v_rcp_f32, v_rsq_f32; ~1 ulp), every sum they feed is accumulated in
 * FP64 by the caller.  log and exp go through base 2 directly: the
 * arguments here (ratios of distances and radii, -r^2/4RiRj, -kappa f) are
 * never denormal, so the library's denormal rescaling is not needed.
 */
template <> struct gbm<float> {
    static __device__ __forceinline__ float log_(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
    static __device__ __forceinline__ float exp_(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float rsqrt_(float x) { return __builtin_amdgcn_rsqf(x); }
};

/*
 * FP64 1/sqrt(r2) for the Coulomb/LJ terms: exact division of the FP64
 * kernel, or (mixed kernel) the FP32 hardware estimate (relative error e0
 * of about 1e-7) refined by one Newton step y <- y (3 - r2 y^2)/2 in FP64,
 * which leaves 1.5 e0^2 ~ 2e-14 relative, far inside the mixed variant's
 * 1e-6 energy bound.
 */
template <typename R> __device__ __forceinline__ double rsqrt_pair(double r2);
template <> __device__ __forceinline__ double rsqrt_pair<double>(double r2) { return 1.0 / sqrt(r2); }
template <> __device__ __forceinline__ double rsqrt_pair<float>(double r2)
{
    const double y = (double)__builtin_amdgcn_rsqf((float)r2);
    return y * fma(-0.5 * r2 * y, y, 1.5);
}

/*
 * The five analytic cases of the Schaefer-Froemmel descreening integral
 * (eff.c:1663-1697 and their derivatives eff.c:2289-2334), in the order the
 * reference tests them:
 *   A  dij > rgbmax - sj     integral cut at rgbmax
 *   B  dij > 4 sj            Taylor series in (sj/dij)^2
 *   C  dij > ri + sj         sphere j wholly outside ri
 *   D  dij > |ri - sj|       partial overlap
 *   E  ri < sj               atom i inside sphere j
 *   -  otherwise (sphere j inside ri): no contribution
 * Every case but B needs one logarithm and at most two reciprocals, so they
 * are computed once, for every lane, before the case split: a wavefront
 * whose lanes fall into several cases then issues one log and two rcp per
 * step instead of one set per case present; the FP64 kernels also skip them
 * when no lane needs them (see descreen_d).  The arguments per case are
 *   u1 = 1/(dij + sj)                       (C, D, E)
 *   u2 = 1/(dij - sj) for A, 1/(r2 - sj2)   (A, C, E)
 *   L  = log((dij - sj)/rgbmax)             A
 *        log(|dij - sj| / (dij + sj))       C, E
 *        log(ri / (dij + sj))               D
 */
struct DsCase {
    bool skip, a, b, c, d, e;
};

/*
 * NEAR = false: the caller guarantees dij > 4 sj for every pair it passes
 * (a far tile pair, see k_born), so the near cases C, D, E cannot occur;
 * they and u1 are compiled out and the log takes case A's argument.  For
 * such pairs the results equal those of NEAR = true.
 */
template <typename R, bool NEAR = true>
__device__ __forceinline__ DsCase ds_case(R ri, R sj, R dij, R rgbmax)
{
    DsCase k;
    k.skip = dij > rgbmax + sj;
    k.a = !k.skip && dij > rgbmax - sj;
    const bool rest = !k.skip && !k.a;
    if constexpr (!NEAR) {
        k.b = rest;
        k.c = k.d = k.e = false;
        return k;
    }
    k.b = rest && dij > R(4) * sj;
    k.c = rest && !k.b && dij > ri + sj;
    k.d = rest && !k.b && !k.c && dij > fabs(ri - sj);
    k.e = rest && !k.b && !k.c && !k.d && ri < sj;
    return k;
}

template <typename R, bool NEAR = true>
__device__ __forceinline__ void ds_trans(const DsCase &k, R ri, R sj, R r2, R dij, R rgbmax1i,
                                         R &u1, R &u2, R &L)
{
    typedef gbm<R> M;
    const R dm = dij - sj;
    if constexpr (!NEAR) {
        /* only case A uses them */
        u1 = R(0);
        u2 = M::rcp(dm);
        L = M::log_(dm * rgbmax1i);
        return;
    }
    u1 = M::rcp(dij + sj);
    u2 = M::rcp(k.a ? dm : r2 - sj * sj);
    R arg;
    if constexpr (sizeof(R) == sizeof(float)) {
        /* FP32: every candidate computed, then selected per lane, instead
         * of the branches the nested conditional compiles to (pass 1 at
         * 25k atoms, 25 A: 0.246 -> 0.233 ms); the FP64 kernel is faster
         * with the conditional */
        const R arg_a = dm * rgbmax1i, arg_d = ri * u1, arg_ce = fabs(dm) * u1;
        arg = k.a ? arg_a : (k.d ? arg_d : arg_ce);
    } else {
        arg = k.a ? dm * rgbmax1i : (k.d ? ri * u1 : fabs(dm) * u1);
    }
    L = M::log_(arg);
}

/*
 * Pass-1 descreening of an atom with offset radius ri (ri1i = 1/ri) by a
 * sphere of scaled radius sj at distance dij (r2 = dij^2, dij1i = 1/dij):
 * returns the amount egb() subtracts from sumi and sets dv to the pass-3
 * derivative numerator datmp (eff.c:2289-2334) of the same pair and
 * direction; callers that need only one of the two let the compiler drop
 * the other.  skip is set when dij > rgbmax + sj (the pair is dropped,
 * neck included, in both passes), taylor when the pair is in case B.
 */
template <typename R, bool NEAR = true>
__device__ __forceinline__ R descreen_dc(R ri, R ri1i, R sj, R r2, R dij, R dij1i,
                                         R rgbmax, R rgbmax1i, R rgbmax2i, bool &skip, bool &taylor, R &dv)
{
    const DsCase k = ds_case<R, NEAR>(ri, sj, dij, rgbmax);
    skip = k.skip;
    taylor = k.b;
    R u1 = R(0), u2 = R(0), L = R(0);
    if (sizeof(R) == sizeof(double)) {
        /* FP64: log and division are long library sequences, so a
         * wavefront step whose lanes are all in case B (or skipped) jumps
         * over them (measured: pass 1 of 1,024 molecules without cutoff
         * 9.7 -> 6.4 ms) */
        if (k.a || k.c || k.d || k.e) ds_trans<R, NEAR>(k, ri, sj, r2, dij, rgbmax1i, u1, u2, L);
    } else {
        /* FP32: the hardware instructions are cheaper than the branch
         * (measured: pass 1 at 25k atoms, 25 A, 4 % slower with it) */
        ds_trans<R, NEAR>(k, ri, sj, r2, dij, rgbmax1i, u1, u2, L);
    }
    const R sj2 = sj * sj;
    const R dij2i = dij1i * dij1i;
    const R dij3i = dij2i * dij1i;
    R v = R(0), w = R(0);
    if (k.a) {
        v = R(0.125) * dij1i * (R(1) + R(2) * dij * u2 + rgbmax2i * (r2 - R(4) * rgbmax * dij - sj2) + R(2) * L);
        /* -2 log(rgbmax/(dij - sj)) = 2 L */
        w = R(0.125) * dij3i * ((r2 + sj2) * (u2 * u2 - rgbmax2i) + R(2) * L);
    } else if (k.b) {
        const R tmpsd = sj2 * dij2i;
        const R dumbo = R(GB_TA) + tmpsd * (R(GB_TB) + tmpsd * (R(GB_TC) + tmpsd * (R(GB_TD) + tmpsd * R(GB_TDD))));
        const R dumbd = R(GB_TE) + tmpsd * (R(GB_TF) + tmpsd * (R(GB_TG) + tmpsd * (R(GB_TH) + tmpsd * R(GB_THH))));
        v = sj * tmpsd * dij2i * dumbo;
        w = tmpsd * sj * dij2i * dij2i * dumbd;
    } else if (k.c) {
        v = R(0.5) * (sj * u2 + R(0.5) * dij1i * L);
        w = u2 * sj * (R(-0.5) * dij2i + u2) + R(0.25) * dij3i * L;
    } else if (k.d) {
        const R theta = R(0.5) * ri1i * dij1i * (r2 + ri * ri - sj2);
        v = R(0.25) * (ri1i * (R(2) - theta) - u1 + dij1i * L);
        w = R(-0.25) * (R(-0.5) * (r2 - ri * ri + sj2) * dij3i * ri1i * ri1i + dij1i * u1 * (u1 - dij1i) - dij3i * L);
    } else if (k.e) {
        v = R(0.5) * (sj * u2 + R(0.5) * dij1i * L) + ri1i;
        w = R(-0.5) * (sj * dij2i * u2 - R(2) * sj * u2 * u2 - R(0.5) * dij3i * L);
    }
    dv = w;
    return v;
}

template <typename R, bool NEAR = true>
__device__ __forceinline__ R descreen_d(R ri, R ri1i, R sj, R r2, R dij, R dij1i,
                                        R rgbmax, R rgbmax1i, R rgbmax2i, bool &skip, R &dv)
{
    bool taylor;
    return descreen_dc<R, NEAR>(ri, ri1i, sj, r2, dij, dij1i, rgbmax, rgbmax1i, rgbmax2i, skip, taylor, dv);
}

/*
 * The case-B (Taylor) derivative numerator alone, for a pair direction
 * known to be in case B without a neck term: descreen_dc's expression
 * tmpsd sj dij2i^2 dumbd with tmpsd = sj^2 dij2i, regrouped as
 * sj^3 dij2i^3 dumbd(sj^2 dij2i) so that a caller can hold sj^2 and sj^3
 * per atom and dij2i^3 per pair (dij2i = 1/dij^2).
 */
template <typename R>
__device__ __forceinline__ R taylor_dv3(R sj2, R sj3, R dij2i, R dij6i)
{
    const R tmpsd = sj2 * dij2i;
    const R dumbd = R(GB_TE) + tmpsd * (R(GB_TF) + tmpsd * (R(GB_TG) + tmpsd * (R(GB_TH) + tmpsd * R(GB_THH))));
    return sj3 * dij6i * dumbd;
}

template <typename R>
__device__ __forceinline__ R ddescreen(R ri, R ri1i, R sj, R r2, R dij, R dij1i,
                                       R rgbmax, R rgbmax1i, R rgbmax2i, bool &skip)
{
    R dv;
    (void)descreen_d<R>(ri, ri1i, sj, r2, dij, dij1i, rgbmax, rgbmax1i, rgbmax2i, skip, dv);
    return dv;
}

/*
 * Neck correction for neck maximum (pos, val) (eff.c:1699-1708): returns
 * the amount subtracted from sumi and sets dn to its pass-3 derivative
 * divided by dij (eff.c:2336-2360); t = 1 + m^2 + 0.3 m^6, m = dij - pos.
 */
template <typename R>
__device__ __forceinline__ R neck_d(R dij, R dij1i, R pos, R val, R neckscale, R &dn)
{
    const R mdist = dij - pos;
    const R mdist2 = mdist * mdist;
    const R mdist6 = mdist2 * mdist2 * mdist2;
    const R rt = gbm<R>::rcp(R(1) + mdist2 + R(0.3) * mdist6);
    const R nv = neckscale * val * rt;
    dn = (R(2) * mdist + R(1.8) * mdist2 * mdist2 * mdist) * nv * rt * dij1i;
    return nv;
}

template <typename R>
__device__ __forceinline__ R dneck_term(R dij, R dij1i, R pos, R val, R neckscale)
{
    R dn;
    (void)neck_d<R>(dij, dij1i, pos, val, neckscale, dn);
    return dn;
}

#endif
