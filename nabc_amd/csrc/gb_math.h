/*
 * gb_math.h -- per-pair generalized-Born arithmetic shared by the FP64 and
 * the mixed-precision kernels (template parameter R is double or float).
 *
 * Formulas follow egb() in src/sff/eff.c:
 *   descreen()   pass 1 term, eff.c:1663-1697 (Schaefer-Froemmel appendix,
 *                Taylor series for d > 4 s, smooth rgbmax switch)
 *   neck_term()  gb 7/8 neck correction, eff.c:1699-1708
 *   ddescreen()  pass 3 derivative numerator, eff.c:2289-2334
 *   dneck_term() derivative of the neck term, eff.c:2336-2360
 * The value returned by descreen() is the amount egb() subtracts from sumi;
 * ddescreen() returns datmp before the multiplication by sumdeijda.
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
 * FP64 by the caller.
 */
template <> struct gbm<float> {
    static __device__ __forceinline__ float log_(float x) { return __logf(x); }
    static __device__ __forceinline__ float exp_(float x) { return __expf(x); }
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float rsqrt_(float x) { return __builtin_amdgcn_rsqf(x); }
};

/*
 * FP64 1/sqrt(r2) for the Coulomb/LJ terms: exact division of the FP64
 * kernel, or (mixed kernel) the FP32 hardware estimate refined by two
 * Newton steps y <- y (3 - r2 y^2)/2 in FP64, ~1 ulp of FP64.
 */
template <typename R> __device__ __forceinline__ double rsqrt_pair(double r2);
template <> __device__ __forceinline__ double rsqrt_pair<double>(double r2) { return 1.0 / sqrt(r2); }
template <> __device__ __forceinline__ double rsqrt_pair<float>(double r2)
{
    double y = (double)rsqrtf((float)r2);
    const double h = 0.5 * r2;
    y = y * (1.5 - h * y * y);
    y = y * (1.5 - h * y * y);
    return y;
}

/*
 * Pass-1 descreening of an atom with offset radius ri (ri1i = 1/ri) by a
 * sphere of scaled radius sj at distance dij (r2 = dij^2, dij1i = 1/dij).
 * skip is set when dij > rgbmax + sj (the pair is dropped, neck included).
 */
template <typename R>
__device__ __forceinline__ R descreen(R ri, R ri1i, R sj, R r2, R dij, R dij1i,
                                      R rgbmax, R rgbmax1i, R rgbmax2i, bool &skip)
{
    typedef gbm<R> M;
    const R sj2 = sj * sj;
    skip = dij > rgbmax + sj;
    if (skip) return R(0);
    if (dij > rgbmax - sj) {
        R uij = M::rcp(dij - sj);
        return R(0.125) * dij1i *
               (R(1) + R(2) * dij * uij + rgbmax2i * (r2 - R(4) * rgbmax * dij - sj2) +
                R(2) * M::log_((dij - sj) * rgbmax1i));
    } else if (dij > R(4) * sj) {
        R dij2i = dij1i * dij1i;
        R tmpsd = sj2 * dij2i;
        R dumbo = R(GB_TA) + tmpsd * (R(GB_TB) + tmpsd * (R(GB_TC) + tmpsd * (R(GB_TD) + tmpsd * R(GB_TDD))));
        return sj * tmpsd * dij2i * dumbo;
    } else if (dij > ri + sj) {
        return R(0.5) * (sj * M::rcp(r2 - sj2) + R(0.5) * dij1i * M::log_((dij - sj) * M::rcp(dij + sj)));
    } else if (dij > fabs(ri - sj)) {
        R theta = R(0.5) * ri1i * dij1i * (r2 + ri * ri - sj2);
        R uij = M::rcp(dij + sj);
        return R(0.25) * (ri1i * (R(2) - theta) - uij + dij1i * M::log_(ri * uij));
    } else if (ri < sj) {
        return R(0.5) * (sj * M::rcp(r2 - sj2) + R(2) * ri1i +
                         R(0.5) * dij1i * M::log_((sj - dij) * M::rcp(sj + dij)));
    }
    return R(0);
}

/* Neck contribution (subtracted from sumi) for neck maximum (pos, val). */
template <typename R>
__device__ __forceinline__ R neck_term(R dij, R pos, R val, R neckscale)
{
    R mdist = dij - pos;
    R mdist2 = mdist * mdist;
    R mdist6 = mdist2 * mdist2 * mdist2;
    return neckscale * val * gbm<R>::rcp(R(1) + mdist2 + R(0.3) * mdist6);
}

/* Pass-3 derivative numerator datmp (eff.c:2289-2334). */
template <typename R>
__device__ __forceinline__ R ddescreen(R ri, R ri1i, R sj, R r2, R dij, R dij1i,
                                       R rgbmax, R rgbmax2i, bool &skip)
{
    typedef gbm<R> M;
    const R sj2 = sj * sj;
    const R dij2i = dij1i * dij1i;
    skip = dij > rgbmax + sj;
    if (skip) return R(0);
    if (dij > rgbmax - sj) {
        R temp1 = M::rcp(dij - sj);
        R dij3i = dij1i * dij2i;
        return R(0.125) * dij3i * ((r2 + sj2) * (temp1 * temp1 - rgbmax2i) - R(2) * M::log_(rgbmax * temp1));
    } else if (dij > R(4) * sj) {
        R tmpsd = sj2 * dij2i;
        R dumbo = R(GB_TE) + tmpsd * (R(GB_TF) + tmpsd * (R(GB_TG) + tmpsd * (R(GB_TH) + tmpsd * R(GB_THH))));
        return tmpsd * sj * dij2i * dij2i * dumbo;
    } else if (dij > ri + sj) {
        R temp1 = M::rcp(r2 - sj2);
        return temp1 * sj * (R(-0.5) * dij2i + temp1) +
               R(0.25) * dij1i * dij2i * M::log_((dij - sj) * M::rcp(dij + sj));
    } else if (dij > fabs(ri - sj)) {
        R temp1 = M::rcp(dij + sj);
        R dij3i = dij2i * dij1i;
        return R(-0.25) * (R(-0.5) * (r2 - ri * ri + sj2) * dij3i * ri1i * ri1i +
                           dij1i * temp1 * (temp1 - dij1i) - dij3i * M::log_(ri * temp1));
    } else if (ri < sj) {
        R temp1 = M::rcp(r2 - sj2);
        return R(-0.5) * (sj * dij2i * temp1 - R(2) * sj * temp1 * temp1 -
                          R(0.5) * dij2i * dij1i * M::log_((sj - dij) * M::rcp(sj + dij)));
    }
    return R(0);
}

/* Derivative of the neck term, divided by dij (eff.c:2336-2360). */
template <typename R>
__device__ __forceinline__ R dneck_term(R dij, R pos, R val, R neckscale)
{
    R mdist = dij - pos;
    R mdist2 = mdist * mdist;
    R mdist6 = mdist2 * mdist2 * mdist2;
    R t = R(1) + mdist2 + R(0.3) * mdist6;
    t = t * t * dij;
    return (R(2) * mdist + R(1.8) * mdist2 * mdist2 * mdist) * val * neckscale * gbm<R>::rcp(t);
}

#endif
