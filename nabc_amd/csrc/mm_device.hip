/*
 * mm_device.hip -- device-resident mme(): bonded and 1-4 terms next to the
 * GB engine, so one energy + gradient never leaves the GPU.
 *
 *   k_bond    ebond, eff.c:82-168    E = Rk (r - Req)^2
 *   k_angle   eangl, eff.c:168-316   E = Tk (theta - Teq)^2, sin clamp 1e-3
 *   k_dihed   ephi,  eff.c:316-627   Fourier terms; a series (Pn < 0) is
 *             expanded on the host into one entry per term
 *   k_nb14    nbond(N14=1), eff.c:896-1331 (diel C or R), scaled by the
 *             per-pair 1/SCEE and 1/SCNB from the prmtop
 * One thread per term, FP64, gradient added with FP64 atomics into the AoS
 * gradient in original atom order; energies reduced per block.
 * The assembled energy vector follows mme34's ene[] (eff.c:2617-2874);
 * frozen atoms get a zero gradient (eff.c:2826-2834).
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "engine.h"

#define MM_THREADS 256

/* sum v over the block and add it to *dst (one atomic per block) */
__device__ __forceinline__ void block_add(double v, double *dst)
{
    __shared__ double s_red[MM_THREADS / 64];
    __syncthreads(); /* an earlier call's reads of s_red are done */
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) s_red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += s_red[k];
        if (s != 0.0) unsafeAtomicAdd(dst, s);
    }
}

__global__ __launch_bounds__(MM_THREADS) void k_fill(double *p, long long n, double v)
{
    long long k = (long long)blockIdx.x * MM_THREADS + threadIdx.x;
    if (k < n) p[k] = v;
}

__global__ __launch_bounds__(MM_THREADS) void k_bond(BondedDev B, const double *__restrict__ x, double *g)
{
    const int t = blockIdx.x * MM_THREADS + threadIdx.x;
    double e = 0.0;
    if (t < B.nb) {
        const int i = 3 * B.b_i[t], j = 3 * B.b_j[t];
        const double rx = x[i] - x[j], ry = x[i + 1] - x[j + 1], rz = x[i + 2] - x[j + 2];
        const double s = sqrt(rx * rx + ry * ry + rz * rz);
        const double db = s - B.b_req[t];
        double df = B.b_rk[t] * db;
        e = df * db;
        df *= 2.0 / s;
        unsafeAtomicAdd(&g[i], rx * df);
        unsafeAtomicAdd(&g[i + 1], ry * df);
        unsafeAtomicAdd(&g[i + 2], rz * df);
        unsafeAtomicAdd(&g[j], -rx * df);
        unsafeAtomicAdd(&g[j + 1], -ry * df);
        unsafeAtomicAdd(&g[j + 2], -rz * df);
    }
    block_add(e, &B.ene[3]);
}

__global__ __launch_bounds__(MM_THREADS) void k_angle(BondedDev B, const double *__restrict__ x, double *g)
{
    const int t = blockIdx.x * MM_THREADS + threadIdx.x;
    double e = 0.0;
    if (t < B.na) {
        const int i1 = 3 * B.a_i[t], i2 = 3 * B.a_j[t], i3 = 3 * B.a_k[t];
        const double dxi = x[i1] - x[i2], dyi = x[i1 + 1] - x[i2 + 1], dzi = x[i1 + 2] - x[i2 + 2];
        const double dxj = x[i3] - x[i2], dyj = x[i3 + 1] - x[i2 + 1], dzj = x[i3 + 2] - x[i2 + 2];
        const double rir = 1.0 / sqrt(dxi * dxi + dyi * dyi + dzi * dzi);
        const double rjr = 1.0 / sqrt(dxj * dxj + dyj * dyj + dzj * dzj);
        const double dxir = dxi * rir, dyir = dyi * rir, dzir = dzi * rir;
        const double dxjr = dxj * rjr, dyjr = dyj * rjr, dzjr = dzj * rjr;
        double cst = dxir * dxjr + dyir * dyjr + dzir * dzjr;
        cst = fmin(1.0, fmax(-1.0, cst));
        const double th = acos(cst);
        const double da = th - B.a_teq[t];
        double df = da * B.a_tk[t];
        e = df * da;
        df = df + df;
        double sn = sin(th);
        if (sn > 0 && sn < 1.e-3) sn = 1.e-3;
        else if (sn < 0 && sn > -1.e-3) sn = -1.e-3;
        df = -df / sn;
        const double xt = df * rir * (dxjr - cst * dxir), dxt = df * rjr * (dxir - cst * dxjr);
        const double yt = df * rir * (dyjr - cst * dyir), dyt = df * rjr * (dyir - cst * dyjr);
        const double zt = df * rir * (dzjr - cst * dzir), dzt = df * rjr * (dzir - cst * dzjr);
        unsafeAtomicAdd(&g[i1], xt);
        unsafeAtomicAdd(&g[i1 + 1], yt);
        unsafeAtomicAdd(&g[i1 + 2], zt);
        unsafeAtomicAdd(&g[i3], dxt);
        unsafeAtomicAdd(&g[i3 + 1], dyt);
        unsafeAtomicAdd(&g[i3 + 2], dzt);
        unsafeAtomicAdd(&g[i2], -(xt + dxt));
        unsafeAtomicAdd(&g[i2 + 1], -(yt + dyt));
        unsafeAtomicAdd(&g[i2 + 2], -(zt + dzt));
    }
    block_add(e, &B.ene[4]);
}

/*
 * One Fourier term of a torsion.  The derivative of cos(phi) with respect
 * to the four atoms is formed from the dot products of the bond vectors
 * a = r2-r1, b = r3-r2, c = r4-r3 (the same decomposition as eff.c:384-461):
 * cos(phi) = uv / sqrt(uu vv) with uu = |a|^2|b|^2 - (a.b)^2,
 * vv = |b|^2|c|^2 - (b.c)^2, uv = (a.b)(b.c) - (a.c)|b|^2.
 */
__global__ __launch_bounds__(MM_THREADS) void k_dihed(BondedDev B, const double *__restrict__ x, double *g)
{
    const int t = blockIdx.x * MM_THREADS + threadIdx.x;
    double e = 0.0;
    if (t < B.nd) {
        const int i1 = 3 * B.d_i[t], i2 = 3 * B.d_j[t], i3 = 3 * B.d_k[t], i4 = 3 * B.d_l[t];
        const double ax = x[i2] - x[i1], ay = x[i2 + 1] - x[i1 + 1], az = x[i2 + 2] - x[i1 + 2];
        const double bx = x[i3] - x[i2], by = x[i3 + 1] - x[i2 + 1], bz = x[i3 + 2] - x[i2 + 2];
        const double cx = x[i4] - x[i3], cy = x[i4 + 1] - x[i3 + 1], cz = x[i4 + 2] - x[i3 + 2];
        const double ab = ax * bx + ay * by + az * bz, bc = bx * cx + by * cy + bz * cz;
        const double ac = ax * cx + ay * cy + az * cz, aa = ax * ax + ay * ay + az * az;
        const double bb = bx * bx + by * by + bz * bz, cc = cx * cx + cy * cy + cz * cz;
        const double uu = aa * bb - ab * ab, vv = bb * cc - bc * bc, uv = ab * bc - ac * bb;
        const double den = uu == 0.0 ? 0.0 : 1.0 / sqrt(uu * vv);
        double co = uv * den;
        const double co1 = 0.5 * co * den;
        /* d cos(phi) / d r_k = dd_k */
        const double a0x = -bc * bx + bb * cx, a0y = -bc * by + bb * cy, a0z = -bc * bz + bb * cz;
        const double b0x = ab * cx + bc * ax - 2. * ac * bx, b0y = ab * cy + bc * ay - 2. * ac * by;
        const double b0z = ab * cz + bc * az - 2. * ac * bz;
        const double c0x = ab * bx - bb * ax, c0y = ab * by - bb * ay, c0z = ab * bz - bb * az;
        const double a1x = 2. * uu * (-cc * bx + bc * cx), a1y = 2. * uu * (-cc * by + bc * cy);
        const double a1z = 2. * uu * (-cc * bz + bc * cz);
        const double b1x = 2. * uu * (bb * cx - bc * bx), b1y = 2. * uu * (bb * cy - bc * by);
        const double b1z = 2. * uu * (bb * cz - bc * bz);
        const double a2x = -2. * vv * (bb * ax - ab * bx), a2y = -2. * vv * (bb * ay - ab * by);
        const double a2z = -2. * vv * (bb * az - ab * bz);
        const double b2x = 2. * vv * (aa * bx - ab * ax), b2y = 2. * vv * (aa * by - ab * ay);
        const double b2z = 2. * vv * (aa * bz - ab * az);
        const double d1x = (a0x - a2x * co1) * den, d1y = (a0y - a2y * co1) * den, d1z = (a0z - a2z * co1) * den;
        const double d2x = (-a0x - b0x - (a1x - a2x - b2x) * co1) * den;
        const double d2y = (-a0y - b0y - (a1y - a2y - b2y) * co1) * den;
        const double d2z = (-a0z - b0z - (a1z - a2z - b2z) * co1) * den;
        const double d3x = (b0x - c0x - (-a1x - b1x + b2x) * co1) * den;
        const double d3y = (b0y - c0y - (-a1y - b1y + b2y) * co1) * den;
        const double d3z = (b0z - c0z - (-a1z - b1z + b2z) * co1) * den;
        const double d4x = (c0x - b1x * co1) * den, d4y = (c0y - b1y * co1) * den, d4z = (c0z - b1z * co1) * den;
        const double pk = B.d_pk[t], pn = B.d_pn[t], ph = B.d_phase[t];
        double df;
        if (fabs(ph - 3.142) > 0.01 && fabs(ph) > 0.01) {
            /* general phase: signed phi from the sign of (u x v).b */
            co = fmin(1.0, fmax(-1.0, co));
            double phi = acos(co);
            const double ux = ay * bz - az * by, uy = az * bx - ax * bz, uz = ax * by - ay * bx;
            const double vx = by * cz - bz * cy, vy = bz * cx - bx * cz, vz = bx * cy - by * cx;
            const double sx = uy * vz - uz * vy, sy = uz * vx - ux * vz, sz = ux * vy - uy * vx;
            if (sx * bx + sy * by + sz * bz < 0.0) phi = -phi;
            const double delta = pn * phi - ph;
            e = pk * (1.0 + cos(delta));
            const double yy = sin(phi);
            if (fabs(yy) > 0.001) df = pk * pn * sin(delta) / yy;
            else if (phi > -1.0 && phi < 1.0) df = pk * pn * sin(delta) * 1000.;
            else df = -pk * pn * sin(delta) * 1000.;
        } else {
            /* phase 0 or pi: polynomials in cos(phi) */
            const double phase = fabs(ph - 3.142) < 0.01 ? -1.0 : 1.0;
            const double c2 = co * co;
            switch ((int)fabs(pn)) {
            case 1: e = pk * (1.0 + phase * co); df = phase * pk; break;
            case 2: e = pk * (1.0 + phase * (2. * c2 - 1.)); df = phase * pk * 4. * co; break;
            case 3: e = pk * (1.0 + phase * co * (4. * c2 - 3.)); df = phase * pk * (12. * c2 - 3.); break;
            case 4: e = pk * (1.0 + phase * (8. * c2 * (c2 - 1.) + 1.)); df = phase * pk * co * (32. * c2 - 16.); break;
            case 5:
                e = pk * (1.0 + phase * (16. * c2 * c2 * co - 20. * c2 * co + 5. * co));
                df = phase * pk * (80. * c2 * c2 - 60. * c2 + 5.);
                break;
            case 6:
                e = pk * (1.0 + phase * (32. * c2 * c2 * c2 - 48. * c2 * c2 + 18. * c2 - 1.));
                df = phase * pk * co * (192. * c2 * c2 - 192. * c2 + 36.);
                break;
            default: e = 0.0; df = 0.0;
            }
        }
        unsafeAtomicAdd(&g[i1], df * d1x); unsafeAtomicAdd(&g[i1 + 1], df * d1y); unsafeAtomicAdd(&g[i1 + 2], df * d1z);
        unsafeAtomicAdd(&g[i2], df * d2x); unsafeAtomicAdd(&g[i2 + 1], df * d2y); unsafeAtomicAdd(&g[i2 + 2], df * d2z);
        unsafeAtomicAdd(&g[i3], df * d3x); unsafeAtomicAdd(&g[i3 + 1], df * d3y); unsafeAtomicAdd(&g[i3 + 2], df * d3z);
        unsafeAtomicAdd(&g[i4], df * d4x); unsafeAtomicAdd(&g[i4 + 1], df * d4y); unsafeAtomicAdd(&g[i4 + 2], df * d4z);
    }
    block_add(e, &B.ene[5]);
}

__global__ __launch_bounds__(MM_THREADS) void k_nb14(BondedDev B, const double *__restrict__ x, double *g)
{
    const int t = blockIdx.x * MM_THREADS + threadIdx.x;
    double ev = 0.0, ee = 0.0;
    if (t < B.n14) {
        const int i = 3 * B.p_i[t], j = 3 * B.p_j[t];
        const double xij = x[i] - x[j], yij = x[i + 1] - x[j + 1], zij = x[i + 2] - x[j + 2];
        const double r2 = xij * xij + yij * yij + zij * zij, r2inv = 1.0 / r2;
        const double rinv = sqrt(r2) * r2inv;
        const double cij = B.q[B.p_i[t]] * B.q[B.p_j[t]] * B.p_scee[t];
        double rs, df2, df;
        if (B.dield == 0) {
            rs = cij * r2inv;
            df2 = -2.0 * rs;
        } else {
            rs = cij * rinv;
            df2 = -rs;
        }
        ee = rs;
        const double fac = B.p_scnb[t];
        if (!B.p_hb[t]) {
            const double r6 = r2inv * r2inv * r2inv;
            const double f2 = B.p_b[t] * r6, f1 = B.p_a[t] * r6 * r6;
            ev = (f1 - f2) * fac;
            df = (df2 + (6.0 * f2 - 12.0 * f1) * fac) * rinv;
        } else {
            const double r10 = r2inv * r2inv * r2inv * r2inv * r2inv;
            const double f2 = B.p_b[t] * r10, f1 = B.p_a[t] * r10 * r2inv;
            ev = (f1 - f2) * fac;
            df = (df2 + (10.0 * f2 - 12.0 * f1) * fac) * rinv;
        }
        df *= rinv;
        unsafeAtomicAdd(&g[i], df * xij);
        unsafeAtomicAdd(&g[i + 1], df * yij);
        unsafeAtomicAdd(&g[i + 2], df * zij);
        unsafeAtomicAdd(&g[j], -df * xij);
        unsafeAtomicAdd(&g[j + 1], -df * yij);
        unsafeAtomicAdd(&g[j + 2], -df * zij);
    }
    block_add(ev, &B.ene[7]);
    block_add(ee, &B.ene[8]);
}

/* zero frozen gradients, copy the GB terms in, total and rms gradient */
__global__ __launch_bounds__(MM_THREADS) void k_finish(BondedDev B, const double *gbene, double *g, int natom,
                                                       double *ene_out)
{
    __shared__ double s_red[MM_THREADS / 64];
    double s = 0.0;
    for (int a = threadIdx.x; a < natom; a += MM_THREADS) {
        if (B.frozen[a]) {
            g[3 * a] = 0.0; g[3 * a + 1] = 0.0; g[3 * a + 2] = 0.0;
        } else {
            s += g[3 * a] * g[3 * a] + g[3 * a + 1] * g[3 * a + 1] + g[3 * a + 2] * g[3 * a + 2];
        }
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = 0.0, e[13];
        for (int k = 0; k < MM_THREADS / 64; k++) tot += s_red[k];
        for (int k = 0; k < 13; k++) e[k] = B.ene[k];
        e[10] = gbene[0];
        e[1] = gbene[1];
        e[2] = gbene[2];
        e[0] = 0.0;
        for (int k = 1; k <= 12; k++) e[0] += e[k];
        for (int k = 0; k < 13; k++) ene_out[k] = e[k];
        ene_out[13] = sqrt(tot / (3.0 * natom));
    }
}

/* ---------------------------------------------------------------- C-ABI */

extern "C" int gbhip_set_bonded(gbhip_ctx *ctx, const gbhip_bonded *b)
{
    if (!ctx || !b) return seterr(GBHIP_EINVAL, "NULL argument");
    Engine &E = ctx->E;
    BondedDev &B = ctx->B;
    if (E.nmol != 1) return seterr(GBHIP_EINVAL, "bonded terms need a single-molecule context");
    const int n = E.natoms;
    auto chk = [n](int k, const int *v) {
        for (int t = 0; t < k; t++)
            if (v[t] < 0 || v[t] >= n) return false;
        return true;
    };
    if (!chk(b->nbond, b->bond_i) || !chk(b->nbond, b->bond_j) || !chk(b->nangle, b->angle_i) ||
        !chk(b->nangle, b->angle_j) || !chk(b->nangle, b->angle_k) || !chk(b->ndihed, b->dihed_i) ||
        !chk(b->ndihed, b->dihed_j) || !chk(b->ndihed, b->dihed_k) || !chk(b->ndihed, b->dihed_l) ||
        !chk(b->npair14, b->p14_i) || !chk(b->npair14, b->p14_j))
        return seterr(GBHIP_EINVAL, "bonded term references an atom outside 0..%d", n - 1);
    if (!b->charge) return seterr(GBHIP_EINVAL, "bonded terms need the charges");
    int rc;
    auto up_i = [&](const int **dst, const int *src, int k) -> int {
        std::vector<int> v(src, src + k);
        return E.upload(dst, v);
    };
    auto up_d = [&](const double **dst, const double *src, int k) -> int {
        std::vector<double> v(src, src + k);
        return E.upload(dst, v);
    };
    B.nb = b->nbond; B.na = b->nangle; B.nd = b->ndihed; B.n14 = b->npair14;
    B.dield = b->dield;
    if ((rc = up_i(&B.b_i, b->bond_i, B.nb)) || (rc = up_i(&B.b_j, b->bond_j, B.nb)) ||
        (rc = up_d(&B.b_rk, b->bond_rk, B.nb)) || (rc = up_d(&B.b_req, b->bond_req, B.nb)) ||
        (rc = up_i(&B.a_i, b->angle_i, B.na)) || (rc = up_i(&B.a_j, b->angle_j, B.na)) ||
        (rc = up_i(&B.a_k, b->angle_k, B.na)) || (rc = up_d(&B.a_tk, b->angle_tk, B.na)) ||
        (rc = up_d(&B.a_teq, b->angle_teq, B.na)) || (rc = up_i(&B.d_i, b->dihed_i, B.nd)) ||
        (rc = up_i(&B.d_j, b->dihed_j, B.nd)) || (rc = up_i(&B.d_k, b->dihed_k, B.nd)) ||
        (rc = up_i(&B.d_l, b->dihed_l, B.nd)) || (rc = up_d(&B.d_pk, b->dihed_pk, B.nd)) ||
        (rc = up_d(&B.d_pn, b->dihed_pn, B.nd)) || (rc = up_d(&B.d_phase, b->dihed_phase, B.nd)) ||
        (rc = up_i(&B.p_i, b->p14_i, B.n14)) || (rc = up_i(&B.p_j, b->p14_j, B.n14)) ||
        (rc = up_i(&B.p_hb, b->p14_hb, B.n14)) || (rc = up_d(&B.p_scee, b->p14_scee, B.n14)) ||
        (rc = up_d(&B.p_scnb, b->p14_scnb, B.n14)) || (rc = up_d(&B.p_a, b->p14_a, B.n14)) ||
        (rc = up_d(&B.p_b, b->p14_b, B.n14)) || (rc = up_d(&B.q, b->charge, n)))
        return rc;
    if (!B.ene && (rc = E.dalloc(&B.ene, 14))) return rc;
    if (!B.frozen) {
        std::vector<int> fz(n, 0);
        for (int s = 0; s < E.nslots; s++) {
            int a = E.perm_h[s];
            if (a >= 0) fz[a] = E.frz_h[s];
        }
        if ((rc = E.upload(&B.frozen, fz))) return rc;
    }
    B.set = true;
    return 0;
}

static int grid_of(long long n) { return (int)((n + MM_THREADS - 1) / MM_THREADS); }

/* full mme34 energy + gradient on the device; d_grad overwritten */
int mme_device(gbhip_ctx *ctx, const double *x_dev, double *d_grad, double *d_ene14)
{
    Engine &E = ctx->E;
    BondedDev &B = ctx->B;
    if (!B.set) return seterr(GBHIP_ESTATE, "gbhip_set_bonded has not been called");
    const long long n3 = 3LL * E.natoms;
    hipLaunchKernelGGL(k_fill, dim3(grid_of(n3)), dim3(MM_THREADS), 0, E.stream, d_grad, n3, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(MM_THREADS), 0, E.stream, B.ene, 14LL, 0.0);
    if (B.nb) hipLaunchKernelGGL(k_bond, dim3(grid_of(B.nb)), dim3(MM_THREADS), 0, E.stream, B, x_dev, d_grad);
    if (B.na) hipLaunchKernelGGL(k_angle, dim3(grid_of(B.na)), dim3(MM_THREADS), 0, E.stream, B, x_dev, d_grad);
    if (B.nd) hipLaunchKernelGGL(k_dihed, dim3(grid_of(B.nd)), dim3(MM_THREADS), 0, E.stream, B, x_dev, d_grad);
    if (B.n14) hipLaunchKernelGGL(k_nb14, dim3(grid_of(B.n14)), dim3(MM_THREADS), 0, E.stream, B, x_dev, d_grad);
    LAUNCHCHK();
    int rc = engine_eval(E, x_dev, d_grad);
    if (rc) return rc;
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(MM_THREADS), 0, E.stream, B, (const double *)E.d.ene, d_grad,
                       E.natoms, d_ene14);
    LAUNCHCHK();
    return 0;
}

extern "C" int gbhip_mme_device(gbhip_ctx *ctx, double *d_grad, double *d_ene14)
{
    if (!ctx || !d_grad || !d_ene14) return seterr(GBHIP_EINVAL, "NULL argument");
    return mme_device(ctx, ctx->E.x_dev, d_grad, d_ene14);
}

extern "C" int gbhip_mme(gbhip_ctx *ctx, double *grad, double *ene14)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    Engine &E = ctx->E;
    int rc = mme_device(ctx, E.x_dev, E.g_dev, E.ene14_dev);
    if (rc) return rc;
    double e[14];
    if (grad)
        HIPCHK(hipMemcpyAsync(grad, E.g_dev, 3 * sizeof(double) * (size_t)E.natoms, hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipMemcpyAsync(e, E.ene14_dev, sizeof(e), hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    if (ene14)
        for (int k = 0; k < 14; k++) ene14[k] = e[k];
    return 0;
}
