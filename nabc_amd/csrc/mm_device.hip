/*
 * mm_device.hip -- device-resident mme(): bonded and 1-4 terms next to the
 * GB engine, so one energy + gradient never leaves the GPU.
 *
 *   d_bond    ebond, eff.c:82-168    E = Rk (r - Req)^2
 *   d_angle   eangl, eff.c:168-316   E = Tk (theta - Teq)^2
 *   d_dihed   ephi,  eff.c:316-627   Fourier terms; a series (Pn < 0) is
 *             expanded on the host into one entry per term
 *   (angle and torsion gradients from the geometry of the plane normals,
 *   see each function)
 *   d_nb14    nbond(N14=1), eff.c:896-1331 (diel C or R), scaled by the
 *             per-pair 1/SCEE and 1/SCNB from the prmtop
 * all four in one launch (k_bonded, a block range per term type).  One
 * thread per term, FP64, gradient added with FP64 atomics into the AoS
 * gradient in original atom order; energies reduced per block.
 * The assembled energy vector follows mme34's ene[] (eff.c:2617-2874);
 * frozen atoms get a zero gradient (eff.c:2826-2834); the rms gradient is
 * reduced over a grid of blocks, then summed in a fixed order (k_mme_post,
 * k_mme_finish).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.h"

#define MM_THREADS 256
#define GNORM_PARTS 1024 /* blocks of the gradient-norm reduction */

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

__device__ __forceinline__ void d_bond(const BondedDev &B, const double *__restrict__ x, double *g, int blk)
{
    const int t = blk * MM_THREADS + threadIdx.x;
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

__device__ __forceinline__ double3 cross3(double3 a, double3 b)
{
    return make_double3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double dot3(double3 a, double3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double3 sub3(const double *__restrict__ x, int i, int j)
{
    return make_double3(x[i] - x[j], x[i + 1] - x[j + 1], x[i + 2] - x[j + 2]);
}
__device__ __forceinline__ void add_grad(double *g, int i, double s, double3 v)
{
    unsafeAtomicAdd(&g[i], s * v.x);
    unsafeAtomicAdd(&g[i + 1], s * v.y);
    unsafeAtomicAdd(&g[i + 2], s * v.z);
}

/*
 * Bond angle i-j-k, E = Tk (theta - Teq)^2.  With u = ri - rj, v = rk - rj
 * and w = u x v (|w| = |u||v| sin theta), theta = atan2(|w|, u.v), and
 * moving ri (rk) changes theta only through its component perpendicular to
 * u (v) in the plane of the angle:
 *   d theta / d ri = (u x w) / (|u|^2 |w|),  d theta / d rk = (w x v) / (|v|^2 |w|),
 *   d theta / d rj = -(d theta / d ri + d theta / d rk)  (translation).
 * The reference bounds sin theta below by 1e-3 in the gradient
 * (eff.c:270-276), which is |w| >= 1e-3 |u||v| here.
 */
__device__ __forceinline__ void d_angle(const BondedDev &B, const double *__restrict__ x, double *g, int blk)
{
    const int t = blk * MM_THREADS + threadIdx.x;
    double e = 0.0;
    if (t < B.na) {
        const int i = 3 * B.a_i[t], j = 3 * B.a_j[t], k = 3 * B.a_k[t];
        const double3 u = sub3(x, i, j), v = sub3(x, k, j);
        const double3 w = cross3(u, v);
        const double uu = dot3(u, u), vv = dot3(v, v), wn = sqrt(dot3(w, w));
        const double theta = atan2(wn, dot3(u, v));
        const double dth = theta - B.a_teq[t];
        e = B.a_tk[t] * dth * dth;
        const double de = 2.0 * B.a_tk[t] * dth / fmax(wn, 1.e-3 * sqrt(uu * vv));
        const double3 gi = cross3(u, w), gk = cross3(w, v);
        const double si = de / uu, sk = de / vv;
        add_grad(g, i, si, gi);
        add_grad(g, k, sk, gk);
        add_grad(g, j, -1.0, make_double3(si * gi.x + sk * gk.x, si * gi.y + sk * gk.y, si * gi.z + sk * gk.z));
    }
    block_add(e, &B.ene[4]);
}

/*
 * One Fourier term of the torsion i-j-k-l, E = Pk (1 + cos(n phi - gamma)).
 * Following Blondel & Karplus (J. Comput. Chem. 17:1132, 1996): with
 * F = ri - rj, G = rj - rk, H = rl - rk and the plane normals A = F x G,
 * B = H x G, phi = atan2((B x A).G / |G|, A.B) (the sign convention of
 * eff.c:582-590: positive when (a x b) x (b x c) points along b), and
 *   d phi / d ri = -|G| A / |A|^2,      d phi / d rl = |G| B / |B|^2,
 *   d phi / d rj = (|G| + F.G/|G|) A / |A|^2 - (H.G/|G|) B / |B|^2,
 *   d phi / d rk = -|G| B / |B|^2 - (F.G/|G|) A / |A|^2 + (H.G/|G|) B / |B|^2,
 * no division by sin phi anywhere.  Energies as the reference evaluates
 * them: a phase of 0 or pi (within 0.01) is the exact sign +-1 and a
 * multiplicity 1 to 6 (eff.c:600-627, others give 0); other phases use the
 * signed Pn as stored (eff.c:591-599).  For those the reference divides by
 * sin phi and replaces 1/sin phi by +-1000 where |sin phi| <= 0.001
 * (eff.c:594-598); the factor `scale` below reproduces that gradient.
 * Collinear i-j-k or j-k-l: cos phi taken as 0 with no gradient, as the
 * reference does (eff.c:412).
 */
__device__ __forceinline__ void d_dihed(const BondedDev &B, const double *__restrict__ x, double *g, int blk)
{
    const int t = blk * MM_THREADS + threadIdx.x;
    double e = 0.0;
    if (t < B.nd) {
        const int i = 3 * B.d_i[t], j = 3 * B.d_j[t], k = 3 * B.d_k[t], l = 3 * B.d_l[t];
        const double3 F = sub3(x, i, j), G = sub3(x, j, k), H = sub3(x, l, k);
        const double3 A = cross3(F, G), Bn = cross3(H, G);
        const double aa = dot3(A, A), bb = dot3(Bn, Bn), gn = sqrt(dot3(G, G));
        const bool degenerate = aa * bb == 0.0;
        const double phi = degenerate ? 1.5707963267948966 : atan2(dot3(cross3(Bn, A), G) / gn, dot3(A, Bn));
        const double pk = B.d_pk[t], pn = B.d_pn[t], ph = B.d_phase[t];
        double dedphi;
        if (fabs(ph - 3.142) > 0.01 && fabs(ph) > 0.01) {
            const double delta = pn * phi - ph;
            e = pk * (1.0 + cos(delta));
            const double sp = sin(phi);
            const double scale = fabs(sp) > 0.001 ? 1.0 : ((phi > -1.0 && phi < 1.0) ? 1000.0 : -1000.0) * sp;
            dedphi = -pk * pn * sin(delta) * scale;
        } else {
            const double sgn = fabs(ph - 3.142) < 0.01 ? -1.0 : 1.0;
            const int n = (int)fabs(pn);
            if (n >= 1 && n <= 6) {
                e = pk * (1.0 + sgn * cos(n * phi));
                dedphi = -pk * sgn * n * sin(n * phi);
            } else {
                e = 0.0;
                dedphi = 0.0;
            }
        }
        if (!degenerate) {
            const double ga = dedphi * gn / aa, gb = dedphi * gn / bb;
            const double fg = dot3(F, G) / gn, hg = dot3(H, G) / gn;
            const double ca = dedphi * fg / aa, cb = dedphi * hg / bb;
            add_grad(g, i, -ga, A);
            add_grad(g, l, gb, Bn);
            add_grad(g, j, 1.0, make_double3((ga + ca) * A.x - cb * Bn.x, (ga + ca) * A.y - cb * Bn.y,
                                             (ga + ca) * A.z - cb * Bn.z));
            add_grad(g, k, 1.0, make_double3((cb - gb) * Bn.x - ca * A.x, (cb - gb) * Bn.y - ca * A.y,
                                             (cb - gb) * Bn.z - ca * A.z));
        }
    }
    block_add(e, &B.ene[5]);
}

__device__ __forceinline__ void d_nb14(const BondedDev &B, const double *__restrict__ x, double *g, int blk)
{
    const int t = blk * MM_THREADS + threadIdx.x;
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

/* all bonded and 1-4 terms in one launch: consecutive block ranges per term
 * type (block-uniform branches, so block_add's barriers are reached by the
 * whole block) */
__global__ __launch_bounds__(MM_THREADS) void k_bonded(BondedDev B, const double *__restrict__ x, double *g,
                                                       int4 nblk)
{
    int b = blockIdx.x;
    if (b < nblk.x) { d_bond(B, x, g, b); return; }
    b -= nblk.x;
    if (b < nblk.y) { d_angle(B, x, g, b); return; }
    b -= nblk.y;
    if (b < nblk.z) { d_dihed(B, x, g, b); return; }
    b -= nblk.z;
    d_nb14(B, x, g, b);
}

/* ---------------------------------------------------------------- C-ABI */

extern "C" int gbhip_set_bonded(gbhip_ctx *ctx, const gbhip_bonded *b)
{
    if (!ctx || !b) return seterr(GBHIP_EINVAL, "NULL argument");
    std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
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
    if (!B.gpart && (rc = E.dalloc(&B.gpart, GNORM_PARTS))) return rc;
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

struct MDPar;
static int mme_launch(gbhip_ctx *ctx, double *x_dev, double *d_grad, double *d_ene14, const MDPar *md, bool rebuild);

/* full mme34 energy + gradient on the device; d_grad overwritten */
int mme_device(gbhip_ctx *ctx, const double *x_dev, double *d_grad, double *d_ene14)
{
    /* x_dev is only read without a dynamics step */
    return mme_launch(ctx, const_cast<double *>(x_dev), d_grad, d_ene14, nullptr, false);
}

extern "C" int gbhip_mme_device(gbhip_ctx *ctx, double *d_grad, double *d_ene14)
{
    if (!ctx || !d_grad || !d_ene14) return seterr(GBHIP_EINVAL, "NULL argument");
    return mme_device(ctx, ctx->E.x_dev, d_grad, d_ene14);
}

/*
 * Everything the launches of mme_device take from the host: device pointers
 * (the derivative store's included), pair-kernel grids, precision, model
 * parameters and the bonded term counts.  A captured evaluation is replayed
 * only while this is unchanged; list contents, coordinates and the store's
 * in-use flag live on the device and need no re-capture.
 */
static std::vector<unsigned char> mme_key(const gbhip_ctx *c)
{
    const Engine &E = c->E;
    std::vector<unsigned char> k;
    auto put = [&k](const void *p, size_t n) {
        const unsigned char *b = (const unsigned char *)p;
        k.insert(k.end(), b, b + n);
    };
    put(&E.d, sizeof(E.d));
    put(&E.par, sizeof(E.par));
    put(E.pass_blocks, sizeof(E.pass_blocks));
    put(&E.precision, sizeof(E.precision));
    put(&E.x_dev, sizeof(E.x_dev));
    put(&E.g_dev, sizeof(E.g_dev));
    put(&E.ene14_dev, sizeof(E.ene14_dev));
    put(&c->B, sizeof(c->B));
    put(&E.hcp, sizeof(E.hcp));
    const bool held = engine_store_holds_list(E); /* decides whether pass 3's recompute kernel is launched */
    put(&held, sizeof(held));
    return k;
}

/*
 * One mme evaluation as a hipGraph launch on the engine stream: captured at
 * the first call and again whenever mme_key changes (a regrown derivative
 * store, new grids, new bonded terms), so a minimizer's calls cost one
 * launch each instead of a dozen.  A context without a stream gets one of
 * its own (capture needs one), as gbhip_md_init does.
 */
static int mme_graph_launch(gbhip_ctx *c)
{
    Engine &E = c->E;
    MmeGraph &G = c->G;
    if (!c->B.set) return seterr(GBHIP_ESTATE, "gbhip_set_bonded has not been called");
    if (!E.have_list) return seterr(GBHIP_ESTATE, "egb called before the pair list was built");
    if (!E.stream) {
        /* work queued on the default stream (a coordinate copy, a list
         * build) is not ordered before a non-blocking stream's: wait for it */
        HIPCHK(hipStreamSynchronize(nullptr));
        hipStream_t s;
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        E.stream = s;
        E.own_stream = true;
    }
    std::vector<unsigned char> key = mme_key(c);
    if (!G.exec || key != G.key) {
        std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
        HIPCHK(hipStreamSynchronize(E.stream));
        if (G.exec) (void)hipGraphExecDestroy(G.exec);
        if (G.graph) (void)hipGraphDestroy(G.graph);
        G.exec = nullptr;
        G.graph = nullptr;
        G.key.clear();
        HIPCHK(hipStreamBeginCapture(E.stream, hipStreamCaptureModeThreadLocal));
        const int rc = mme_device(c, E.x_dev, E.g_dev, E.ene14_dev);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(E.stream, &g);
        if (rc || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
            return rc ? rc : seterr(GBHIP_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e2));
        }
        const hipError_t e3 = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
        if (e3 != hipSuccess) {
            (void)hipGraphDestroy(g);
            G.exec = nullptr;
            return seterr(GBHIP_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e3));
        }
        G.graph = g;
        G.key.swap(key);
        G.captures++;
    }
    HIPCHK(hipGraphLaunch(G.exec, E.stream));
    G.replays++;
    return 0;
}

/* one mme evaluation of E.x_dev into E.g_dev / E.ene14_dev, as gbhip_mme
 * launches it (the captured graph, or kernel by kernel while timing), no
 * copies */
int engine_mme(gbhip_ctx *c)
{
    Engine &E = c->E;
    const bool graph = c->G.on && E.tev_cap == 0;
    return graph ? mme_graph_launch(c) : mme_device(c, E.x_dev, E.g_dev, E.ene14_dev);
}

extern "C" int gbhip_set_mme_graph(gbhip_ctx *ctx, int on)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    if (on != 0 && on != 1) return seterr(GBHIP_EINVAL, "mme graph mode must be 0 or 1");
    ctx->G.on = on != 0;
    return 0;
}

extern "C" int gbhip_mme_graph_stats(gbhip_ctx *ctx, long long *captures, long long *replays)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    if (captures) *captures = ctx->G.captures;
    if (replays) *replays = ctx->G.replays;
    return 0;
}

extern "C" int gbhip_mme(gbhip_ctx *ctx, double *grad, double *ene14)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    Engine &E = ctx->E;
    /* per-pass timing events cannot sit inside a graph: a timing session
     * evaluates with plain launches */
    const bool graph = ctx->G.on && E.tev_cap == 0;
    int rc = graph ? mme_graph_launch(ctx) : mme_device(ctx, E.x_dev, E.g_dev, E.ene14_dev);
    if (rc) return rc;
    /* out through the pinned staging buffer (gradient after the coordinate
     * part, energies last), or directly when there is none */
    const size_t n3 = 3 * (size_t)E.natoms;
    double e14[14];
    const bool staged = engine_stage(E);
    double *hg = staged ? E.h_stage + n3 : grad, *he = staged ? E.h_stage + 2 * n3 : e14;
    if (grad) HIPCHK(hipMemcpyAsync(hg, E.g_dev, sizeof(double) * n3, hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipMemcpyAsync(he, E.ene14_dev, 14 * sizeof(double), hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    if (grad && staged) memcpy(grad, hg, sizeof(double) * n3);
    if (ene14) memcpy(ene14, he, 14 * sizeof(double));
    return 0;
}

/* ------------------------------------------------------ second derivatives */

/* coordinate c of x set to x0[c] + delta (the others are untouched) */
__global__ void k_hess_set(double *x, const double *x0, long long c, double delta)
{
    if (threadIdx.x == 0) x[c] = x0[c] + delta;
}

/* one column of the Hessian: (g(x + h e_c) - g(x - h e_c)) / 2h */
__global__ __launch_bounds__(MM_THREADS) void k_hess_col(double *__restrict__ col, const double *__restrict__ gp,
                                                         const double *__restrict__ gm, double inv2h, long long n)
{
    const long long r = (long long)blockIdx.x * MM_THREADS + threadIdx.x;
    if (r < n) col[r] = (gp[r] - gm[r]) * inv2h;
}

/* H <- (H + H^T) / 2 over 16 x 16 blocks of the upper triangle */
__global__ __launch_bounds__(256) void k_hess_sym(double *H, long long n)
{
    const long long r = (long long)blockIdx.y * 16 + threadIdx.y, c = (long long)blockIdx.x * 16 + threadIdx.x;
    if (r < n && c < n && r < c) {
        const double a = 0.5 * (H[r * n + c] + H[c * n + r]);
        H[r * n + c] = a;
        H[c * n + r] = a;
    }
}

/*
 * Hessian of the mme energy (bonded, 1-4 and GB terms) at the context's
 * coordinates and pair list, by central differences of the analytic
 * gradient: every column is two device evaluations at x +- h e_c (the
 * captured mme graph when it is on), differenced on the device, and the
 * result symmetrised.  The coordinates are restored and evaluated once
 * more, so grad / ene14 are those at x.  The list is not rebuilt.
 */
extern "C" int gbhip_hessian(gbhip_ctx *ctx, double h, double *hess, double *grad, double *ene14)
{
    if (!ctx || !hess) return seterr(GBHIP_EINVAL, "NULL argument");
    if (!(h > 0.0) || !std::isfinite(h)) return seterr(GBHIP_EINVAL, "difference step must be positive");
    Engine &E = ctx->E;
    if (!ctx->B.set) return seterr(GBHIP_ESTATE, "gbhip_set_bonded has not been called");
    if (!E.have_list) return seterr(GBHIP_ESTATE, "hessian called before the pair list was built");
    const long long n3 = 3LL * E.natoms;
    double *x0 = nullptr, *gp = nullptr, *dH = nullptr;
    {
        std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
        if (hipMalloc(&x0, n3 * sizeof(double)) != hipSuccess || hipMalloc(&gp, n3 * sizeof(double)) != hipSuccess ||
            hipMalloc(&dH, (size_t)n3 * (size_t)n3 * sizeof(double)) != hipSuccess) {
            (void)hipGetLastError();
            if (x0) (void)hipFree(x0);
            if (gp) (void)hipFree(gp);
            return seterr(GBHIP_ENOMEM, "hipMalloc of the %lld x %lld Hessian failed", n3, n3);
        }
    }
    int rc = 0;
    auto fail = [&](int code) {
        rc = code;
        return code;
    };
    const unsigned nb = (unsigned)((n3 + MM_THREADS - 1) / MM_THREADS);
    if (hipMemcpyAsync(x0, E.x_dev, n3 * sizeof(double), hipMemcpyDeviceToDevice, E.stream) != hipSuccess)
        fail(seterr(GBHIP_EHIP, "coordinate copy failed"));
    for (long long c = 0; c < n3 && !rc; c++) {
        hipLaunchKernelGGL(k_hess_set, dim3(1), dim3(64), 0, E.stream, E.x_dev, x0, c, h);
        if (engine_mme(ctx)) { fail(GBHIP_EHIP); break; }
        if (hipMemcpyAsync(gp, E.g_dev, n3 * sizeof(double), hipMemcpyDeviceToDevice, E.stream) != hipSuccess) {
            fail(seterr(GBHIP_EHIP, "gradient copy failed"));
            break;
        }
        hipLaunchKernelGGL(k_hess_set, dim3(1), dim3(64), 0, E.stream, E.x_dev, x0, c, -h);
        if (engine_mme(ctx)) { fail(GBHIP_EHIP); break; }
        hipLaunchKernelGGL(k_hess_col, dim3(nb), dim3(MM_THREADS), 0, E.stream, dH + c * n3, gp, E.g_dev,
                           0.5 / h, n3);
        hipLaunchKernelGGL(k_hess_set, dim3(1), dim3(64), 0, E.stream, E.x_dev, x0, c, 0.0);
    }
    if (!rc) {
        const unsigned nt = (unsigned)((n3 + 15) / 16);
        hipLaunchKernelGGL(k_hess_sym, dim3(nt, nt), dim3(16, 16), 0, E.stream, dH, n3);
        if (hipGetLastError() != hipSuccess) fail(seterr(GBHIP_EHIP, "kernel launch failed"));
    }
    /* the gradient and energies at x itself */
    if (!rc && engine_mme(ctx)) fail(GBHIP_EHIP);
    if (!rc && hipMemcpyAsync(hess, dH, (size_t)n3 * (size_t)n3 * sizeof(double), hipMemcpyDeviceToHost,
                              E.stream) != hipSuccess)
        fail(seterr(GBHIP_EHIP, "Hessian copy failed"));
    if (!rc && grad && hipMemcpyAsync(grad, E.g_dev, n3 * sizeof(double), hipMemcpyDeviceToHost, E.stream) != hipSuccess)
        fail(seterr(GBHIP_EHIP, "gradient copy failed"));
    if (!rc && ene14 && hipMemcpyAsync(ene14, E.ene14_dev, 14 * sizeof(double), hipMemcpyDeviceToHost, E.stream) !=
                            hipSuccess)
        fail(seterr(GBHIP_EHIP, "energy copy failed"));
    if (hipStreamSynchronize(E.stream) != hipSuccess && !rc) fail(seterr(GBHIP_EHIP, "stream synchronisation failed"));
    {
        std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
        (void)hipFree(x0);
        (void)hipFree(gp);
        (void)hipFree(dH);
    }
    return rc;
}

/* ------------------------------------------------------------- dynamics */

/*
 * Device-resident md() (sff.c:2071-2366): velocity Verlet with the
 * Langevin (gamma_ln > 0) or Berendsen (gamma_ln = 0) scheme of the
 * reference, vlimit clamp, frozen atoms held.  One step is
 *   drift -> [pair-list rebuild] -> mme (bonded + 1-4 + GB) -> kick ->
 *   kinetic energy and log row
 * (the drift, kick and bookkeeping fused into the evaluation's first and
 * last kernels, see mme_launch) with no host synchronisation; the two variants (with / without the
 * rebuild) are captured once as hipGraphs and replayed.  Noise: one N(0,1)
 * per (step, coordinate) from a counter-based hash (splitmix64 mixing of
 * seed, step and index, then Box-Muller), the same function as the host
 * nabc_md_gauss().
 */

__device__ __forceinline__ unsigned long long md_mix(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double md_gauss(unsigned long long seed, long long step, long long idx)
{
    const unsigned long long h =
        md_mix(seed ^ md_mix((unsigned long long)step * 0x100000001B3ULL ^ md_mix((unsigned long long)idx)));
    const unsigned long long h2 = md_mix(h);
    const double u1 = ((double)(h >> 11) + 1.0) * (1.0 / 9007199254740993.0);
    const double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

struct MDPar {
    double dtx, dt5, c_implic, c_explic, c_ave, sdfac, gammai, ekin0, dttp, vlimit, rndf, boltz2;
    unsigned long long seed;
    int n3, log_cap;
};

__device__ __forceinline__ double md_tscal(const MDPar &P, const double *scal)
{
    const double ekin = scal[0];
    return ekin > 0.01 ? sqrt(1.0 + P.dttp * (P.ekin0 / ekin - 1.0)) : 1.0;
}

/* initial velocities (sff.c:2165-2188) and kinetic energy */
__global__ __launch_bounds__(MM_THREADS) void k_md_init_v(MDState M, MDPar P, const int *frz, double tempi,
                                                          int zerov, int keep)
{
    const int i = blockIdx.x * MM_THREADS + threadIdx.x;
    double e = 0.0;
    if (i < P.n3) {
        double v = M.v[i];
        if (zerov) {
            v = 0.0;
        } else if (!keep) {
            if (frz[i / 3]) {
                v = 0.0;
            } else {
                v = sqrt(2.0 * P.boltz2 * tempi * M.minv[i]) * md_gauss(P.seed, -1, i);
                e = v * v / M.minv[i];
            }
        } else {
            /* the reference sums v^2 * minv here (sff.c:2186) */
            e = v * v * M.minv[i];
        }
        M.v[i] = v;
    }
    block_add(e, &M.scal[1]);
}

__global__ void k_md_init_done(MDState M)
{
    M.scal[0] = 0.5 * M.scal[1];
    M.scal[1] = 0.0;
    *M.step = 1;
}

/* first accelerations after the step-0 force (sff.c:2199-2215) */
__global__ __launch_bounds__(MM_THREADS) void k_md_accel0(MDState M, MDPar P, const int *frz, const double *g)
{
    const int i = blockIdx.x * MM_THREADS + threadIdx.x;
    if (i >= P.n3) return;
    if (P.gammai == 0.0) {
        M.accel[i] = -g[i] * M.minv[i] * P.dt5;
    } else if (frz[i / 3]) {
        M.accel[i] = 0.0;
    } else {
        M.sqrmass[i] = 1.0 / sqrt(M.minv[i]);
        M.accel[i] = (-g[i] + P.sdfac * M.sqrmass[i] * md_gauss(P.seed, 0, i)) * M.minv[i] * P.dt5;
    }
}

/* ------------------------------------------------- fused evaluation kernels
 * One mme evaluation is k_mme_pre -> k_bonded -> the GB passes ->
 * k_mme_post -> k_mme_finish; in a dynamics step the drift is done by
 * k_mme_pre, the second half kick by k_mme_post and the step's bookkeeping
 * by k_mme_finish (MD = true), so a step is as many launches as an
 * evaluation.  Each launch costs a few microseconds however small its
 * work, which is most of a step below a few thousand atoms.
 */

/* per slot: [dynamics: drift of its atom (sff.c:2225-2249)], its gradient
 * cleared, its coordinates gathered (slot_gather); the bonded energies and
 * [dynamics] the kinetic-energy accumulator cleared */
template <bool MD>
__global__ __launch_bounds__(MM_THREADS) void k_mme_pre(DevPtrs d, GBPar p, double *x, double *g, double *bene,
                                                        MDState M, MDPar P)
{
    const int s = blockIdx.x * MM_THREADS + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x < 14) bene[threadIdx.x] = 0.0;
    if (MD && s == 0) M.scal[1] = 0.0;
    if (s >= p.nslots) return;
    const int a = d.perm[s];
    if (a >= 0) {
        for (int c = 0; c < 3; c++) {
            const int i = 3 * a + c;
            if (MD) {
                double v;
                if (P.gammai == 0.0) v = (M.v[i] + M.accel[i]) * md_tscal(P, M.scal);
                else v = P.c_explic * M.v[i] + M.accel[i];
                v = fmin(P.vlimit, fmax(-P.vlimit, v));
                M.v[i] = v;
                x[i] += v * P.dtx;
            }
            g[i] = 0.0;
        }
    }
    slot_gather(d, s, x);
}

/*
 * The energy-slot sums (ene_slot_sums), then grid-stride over (slot,
 * component) pairs, one gradient component per thread (a dynamics step's
 * noise is three FP64 log/sqrt/cos evaluations per atom, so a thread per
 * component triples the parallelism of the kick): the slot's GB gradient
 * added into its atom's, frozen gradients zeroed (eff.c:2826-2834),
 * per-block sums of squared gradient components (written per block, so the
 * rms is summed in a fixed order), and [dynamics] the second half kick with
 * the new forces and 2*ekin (sff.c:2262-2277).
 */
template <bool MD>
__global__ __launch_bounds__(MM_THREADS) void k_mme_post(DevPtrs d, GBPar p, BondedDev B, double *g, int nene,
                                                         MDState M, MDPar P)
{
    __shared__ double s_red[MM_THREADS / 64];
    ene_slot_sums(d, nene);
    double sq = 0.0, ek = 0.0;
    const int n3 = 3 * p.nslots;
    for (int t = blockIdx.x * MM_THREADS + threadIdx.x; t < n3; t += gridDim.x * MM_THREADS) {
        const int s = t / 3, c = t - 3 * s;
        const int a = d.perm[s];
        if (a < 0) continue;
        const int i = 3 * a + c;
        const bool frz = B.frozen[a] != 0;
        const double gs = c == 0 ? d.gx[s] : (c == 1 ? d.gy[s] : d.gz[s]);
        const double gv = frz ? 0.0 : g[i] + gs;
        sq += gv * gv;
        g[i] = gv;
        if (MD) {
            double v = M.v[i];
            if (P.gammai == 0.0) {
                const double acc = -gv * M.minv[i] * P.dt5;
                M.accel[i] = acc;
                v = (v + acc) * md_tscal(P, M.scal);
            } else if (!frz) {
                const double acc = (-gv + P.sdfac * M.sqrmass[i] * md_gauss(P.seed, *M.step, i)) * M.minv[i] * P.dt5;
                M.accel[i] = acc;
                v = (v + acc) * P.c_implic;
            }
            M.v[i] = v;
            ek += v * v / M.minv[i];
        }
    }
    for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = sq;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < MM_THREADS / 64; k++) t += s_red[k];
        B.gpart[blockIdx.x] = t;
    }
    if (MD) block_add(ek, &M.scal[1]);
}

/* the GB terms copied in, the total energy and the rms gradient (one
 * wavefront); [dynamics] kinetic energy, temperature, the per-step log row
 * and the step counter (sff.c:2294-2339) */
template <bool MD>
__global__ __launch_bounds__(64) void k_mme_finish(BondedDev B, const double *gbene, int nparts, int natom,
                                                   double *ene_out, MDState M, MDPar P)
{
    double s = 0.0;
    for (int k = threadIdx.x; k < nparts; k += 64) s += B.gpart[k];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (threadIdx.x != 0) return;
    double e[14];
    for (int k = 0; k < 13; k++) e[k] = B.ene[k];
    e[10] = gbene[0];
    e[1] = gbene[1];
    e[2] = gbene[2];
    e[0] = 0.0;
    for (int k = 1; k <= 12; k++) e[0] += e[k];
    e[13] = sqrt(s / (3.0 * natom));
    for (int k = 0; k < 14; k++) ene_out[k] = e[k];
    if (MD) {
        double ekin = 0.5 * M.scal[1];
        if (P.gammai > 0.0) ekin *= P.c_ave;
        const int st = *M.step;
        double *row = M.log + 16 * (size_t)((st - 1) % P.log_cap);
        for (int k = 0; k < 14; k++) row[k] = e[k];
        row[14] = ekin;
        row[15] = ekin / (P.boltz2 * P.rndf);
        M.scal[0] = ekin;
        *M.step = st + 1;
    }
}

/* grid of k_mme_post (its partial sums fit the GNORM_PARTS entries) */
static int post_blocks(const Engine &E)
{
    return std::min(GNORM_PARTS, std::max(1, grid_of(3LL * E.nslots)));
}

/*
 * The launches of one evaluation on the engine stream (also under
 * capture).  md: a dynamics step around it (drift before, kick and
 * bookkeeping after); rebuild: the pair list is rebuilt from the drifted
 * coordinates first.
 */
static int mme_launch(gbhip_ctx *ctx, double *x_dev, double *d_grad, double *d_ene14, const MDPar *md, bool rebuild)
{
    Engine &E = ctx->E;
    BondedDev &B = ctx->B;
    if (!B.set) return seterr(GBHIP_ESTATE, "gbhip_set_bonded has not been called");
    if (!E.have_list && !rebuild) return seterr(GBHIP_ESTATE, "egb called before the pair list was built");
    const MDPar P = md ? *md : MDPar{};
    const int npre = std::max(1, grid_of(E.nslots)), npost = post_blocks(E);
    if (md)
        hipLaunchKernelGGL(k_mme_pre<true>, dim3(npre), dim3(MM_THREADS), 0, E.stream, E.d, E.par, x_dev, d_grad,
                           B.ene, ctx->M, P);
    else
        hipLaunchKernelGGL(k_mme_pre<false>, dim3(npre), dim3(MM_THREADS), 0, E.stream, E.d, E.par, x_dev, d_grad,
                           B.ene, ctx->M, P);
    LAUNCHCHK();
    int rc;
    if (rebuild && (rc = engine_nblist(E))) return rc;
    const int4 nblk = make_int4(grid_of(B.nb), grid_of(B.na), grid_of(B.nd), grid_of(B.n14));
    const int nbonded = nblk.x + nblk.y + nblk.z + nblk.w;
    if (nbonded)
        hipLaunchKernelGGL(k_bonded, dim3(nbonded), dim3(MM_THREADS), 0, E.stream, B, (const double *)x_dev, d_grad,
                           nblk);
    LAUNCHCHK();
    if ((rc = engine_eval(E, x_dev, d_grad, EV_GATHERED | EV_NO_POST))) return rc;
    if (md) {
        hipLaunchKernelGGL(k_mme_post<true>, dim3(npost), dim3(MM_THREADS), 0, E.stream, E.d, E.par, B, d_grad,
                           3 * E.nmol, ctx->M, P);
        hipLaunchKernelGGL(k_mme_finish<true>, dim3(1), dim3(64), 0, E.stream, B, (const double *)E.d.ene, npost,
                           E.natoms, d_ene14, ctx->M, P);
    } else {
        hipLaunchKernelGGL(k_mme_post<false>, dim3(npost), dim3(MM_THREADS), 0, E.stream, E.d, E.par, B, d_grad,
                           3 * E.nmol, ctx->M, P);
        hipLaunchKernelGGL(k_mme_finish<false>, dim3(1), dim3(64), 0, E.stream, B, (const double *)E.d.ene, npost,
                           E.natoms, d_ene14, ctx->M, P);
    }
    LAUNCHCHK();
    return 0;
}

static MDPar md_par(const MDState &M, int n3)
{
    MDPar P;
    P.dtx = M.dtx; P.dt5 = M.dt5; P.c_implic = M.c_implic; P.c_explic = M.c_explic; P.c_ave = M.c_ave;
    P.sdfac = M.sdfac; P.gammai = M.gammai; P.ekin0 = M.ekin0; P.dttp = M.dttp; P.vlimit = M.vlimit;
    P.rndf = M.rndf; P.boltz2 = M.boltz2; P.seed = M.seed; P.n3 = n3; P.log_cap = M.log_cap;
    return P;
}

/* one dynamics step, launched on the engine stream (also used under capture) */
static int md_step_launch(gbhip_ctx *c, bool rebuild)
{
    Engine &E = c->E;
    MDState &M = c->M;
    const int n3 = 3 * E.natoms;
    const MDPar P = md_par(M, n3);
    return mme_launch(c, E.x_dev, E.g_dev, E.ene14_dev, &P, rebuild);
}

/* bounding-box diagonal of x (3n, AoS) */
static double bbox_diag(const double *x, int n)
{
    double lo[3] = {x[0], x[1], x[2]}, hi[3] = {x[0], x[1], x[2]};
    for (int a = 1; a < n; a++)
        for (int c = 0; c < 3; c++) {
            lo[c] = fmin(lo[c], x[3 * a + c]);
            hi[c] = fmax(hi[c], x[3 * a + c]);
        }
    return sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                (hi[2] - lo[2]) * (hi[2] - lo[2]));
}

/* From coordinates x at step `step`: the last step up to which every pair
 * stays strictly within the cutoff, given the per-step growth bound; steps
 * from `step` on are covered while diag + (s - step) g < cut. */
static void md_static_bound(const Engine &E, MDState &M, const double *x, long long step)
{
    M.bbox_checks++;
    M.static_until = step;
    if (M.static_growth <= 0.0) return;
    const double diag = bbox_diag(x, E.natoms), cut = sqrt(E.par.cut2);
    if (!std::isfinite(diag) || !(diag < cut)) return;
    long long k = (long long)floor((cut - diag) / M.static_growth);
    while (k > 0 && !(diag + (double)k * M.static_growth < cut)) k--;
    M.static_until = step + k;
}

static int md_capture(gbhip_ctx *c, bool rebuild, hipGraph_t *g, hipGraphExec_t *ge)
{
    Engine &E = c->E;
    std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
    /* the graph keeps the derivative-store pointer it was captured with;
     * the steps replay across list rebuilds */
    E.store_frozen = true;
    HIPCHK(hipStreamBeginCapture(E.stream, hipStreamCaptureModeThreadLocal));
    const int saved_cap = E.tev_cap;
    E.tev_cap = 0; /* no timing events inside a graph */
    E.capture_spans_rebuilds = true;
    int rc = md_step_launch(c, rebuild);
    E.capture_spans_rebuilds = false;
    hipError_t e2 = hipStreamEndCapture(E.stream, g);
    E.tev_cap = saved_cap;
    if (rc) return rc;
    if (e2 != hipSuccess) return seterr(GBHIP_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e2));
    HIPCHK(hipGraphInstantiate(ge, *g, nullptr, nullptr, 0));
    return 0;
}

extern "C" int gbhip_md_init(gbhip_ctx *ctx, const double *masses, const double *x, const double *v0,
                             const gbhip_md_options *o, int log_cap)
{
    if (!ctx || !masses || !x || !o) return seterr(GBHIP_EINVAL, "NULL argument");
    Engine &E = ctx->E;
    MDState &M = ctx->M;
    if (!ctx->B.set) return seterr(GBHIP_ESTATE, "gbhip_set_bonded must precede gbhip_md_init");
    if (log_cap < 1 || o->dt <= 0 || o->nsnb < 1) return seterr(GBHIP_EINVAL, "bad dynamics options");
    const int n = E.natoms, n3 = 3 * n;
    int rc;
    if (!E.stream) {
        /* graph capture needs a stream of our own (default priority); work
         * already queued on the default stream is not ordered before it */
        HIPCHK(hipStreamSynchronize(nullptr));
        hipStream_t s;
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        E.stream = s;
        E.own_stream = true;
    }
    {
        std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
        if (!M.v) {
            if ((rc = E.dalloc(&M.v, n3)) || (rc = E.dalloc(&M.accel, n3)) || (rc = E.dalloc(&M.minv, n3)) ||
                (rc = E.dalloc(&M.sqrmass, n3)) || (rc = E.dalloc(&M.scal, 2)) || (rc = E.dalloc(&M.step, 1)))
                return rc;
        }
        if (M.log_cap != log_cap) {
            if ((rc = E.dalloc(&M.log, 16 * (size_t)log_cap))) return rc;
            M.log_cap = log_cap;
        }
    }
    std::vector<double> minv(n3);
    for (int a = 0; a < n; a++) {
        if (!(masses[a] > 0)) return seterr(GBHIP_EINVAL, "atom %d has mass %g", a + 1, masses[a]);
        minv[3 * a] = minv[3 * a + 1] = minv[3 * a + 2] = 1.0 / masses[a];
    }
    int nfrozen = 0;
    for (int s = 0; s < E.nslots; s++)
        if (E.perm_h[s] >= 0 && E.frz_h[s]) nfrozen++;
    if ((rc = E.h2d(M.minv, minv.data(), sizeof(double) * n3))) return rc;
    if (v0) {
        if ((rc = E.h2d(M.v, v0, sizeof(double) * n3))) return rc;
    } else {
        hipLaunchKernelGGL(k_fill, dim3(grid_of(n3)), dim3(MM_THREADS), 0, E.stream, M.v, (long long)n3, 0.0);
    }
    hipLaunchKernelGGL(k_fill, dim3(grid_of(n3)), dim3(MM_THREADS), 0, E.stream, M.sqrmass, (long long)n3, 0.0);
    if ((rc = E.h2d(E.x_dev, x, sizeof(double) * n3))) return rc;
    const double dtx = o->dt * 20.455;
    M.dtx = dtx;
    M.dt5 = 0.5 * dtx;
    M.rndf = n3 - 3.0 * nfrozen;
    M.boltz2 = o->boltz2;
    M.ekin0 = o->boltz2 * M.rndf * o->temp0;
    M.dttp = o->dt / o->tautp;
    M.gammai = o->gamma_ln / 20.455;
    M.c_implic = 1.0 / (1.0 + M.gammai * M.dt5);
    M.c_explic = 1.0 - M.gammai * M.dt5;
    M.c_ave = 1.0 + M.gammai * M.dt5;
    M.sdfac = sqrt(4.0 * M.gammai * o->boltz2 * o->temp0 / dtx);
    M.vlimit = o->vlimit;
    M.nsnb = o->nsnb;
    /* A rebuild can change the list only if some pair crosses the cutoff.
     * Each step moves a coordinate by at most vlimit * dtx (the clamp of
     * the drift), so a pair's separation grows by at most
     * 2 sqrt(3) vlimit dtx per step; while the bounding-box diagonal plus
     * that growth stays below the cutoff, every pair is listed and a
     * rebuild builds the list it replaces: such rebuild steps are skipped
     * (same results).  gbhip_md_run checks the box again when the bound
     * runs out. */
    M.static_growth = (o->vlimit > 0.0 && std::isfinite(o->vlimit)) ? 2.0 * sqrt(3.0) * o->vlimit * dtx : 0.0;
    M.rebuilds_run = M.rebuilds_skipped = M.bbox_checks = 0;
    M.static_until = 0;
    md_static_bound(E, M, x, 0);
    M.seed = o->seed;
    M.steps_done = 0;
    const MDPar P = md_par(M, n3);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(MM_THREADS), 0, E.stream, M.scal, 2LL, 0.0);
    hipLaunchKernelGGL(k_md_init_v, dim3(grid_of(n3)), dim3(MM_THREADS), 0, E.stream, M, P, ctx->B.frozen,
                       o->tempi, o->zerov, (o->tempi > 0.0 || o->zerov) ? 0 : 1);
    hipLaunchKernelGGL(k_md_init_done, dim3(1), dim3(1), 0, E.stream, M);
    LAUNCHCHK();
    /* graphs of an earlier run hold stale launch arguments */
    if (M.g_step) { (void)hipGraphExecDestroy(M.g_step); M.g_step = nullptr; }
    if (M.g_rebuild) { (void)hipGraphExecDestroy(M.g_rebuild); M.g_rebuild = nullptr; }
    if (M.gr_step) { (void)hipGraphDestroy(M.gr_step); M.gr_step = nullptr; }
    if (M.gr_rebuild) { (void)hipGraphDestroy(M.gr_rebuild); M.gr_rebuild = nullptr; }
    E.store_frozen = false;
    /* step 0: list build and forces (md() calls mme with iter 0) */
    if ((rc = engine_nblist(E))) return rc;
    /* the step graphs freeze the store: size it for any list this molecule
     * can have, so that no rebuild during the run falls back to recomputing */
    if ((rc = engine_store_reserve(E, true))) return rc;
    if ((rc = mme_device(ctx, E.x_dev, E.g_dev, E.ene14_dev))) return rc;
    hipLaunchKernelGGL(k_md_accel0, dim3(grid_of(n3)), dim3(MM_THREADS), 0, E.stream, M, P, ctx->B.frozen,
                       (const double *)E.g_dev);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(E.stream));
    M.ready = true;
    return 0;
}

extern "C" int gbhip_md_run(gbhip_ctx *ctx, int nsteps, int use_graph)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    MDState &M = ctx->M;
    Engine &E = ctx->E;
    if (!M.ready) return seterr(GBHIP_ESTATE, "gbhip_md_init has not been called");
    int rc;
    if (use_graph && !M.g_step) {
        if ((rc = md_capture(ctx, false, &M.gr_step, &M.g_step))) return rc;
        if ((rc = md_capture(ctx, true, &M.gr_rebuild, &M.g_rebuild))) return rc;
    }
    bool recheck = M.static_growth > 0.0;
    for (int k = 0; k < nsteps; k++) {
        const int step = M.steps_done + 1;
        /* mme34: iter % nsnb == 0 (eff.c:2564), unless no pair can leave the list */
        bool rebuild = step % M.nsnb == 0;
        if (rebuild && step > M.static_until && recheck) {
            /* the bound ran out: measure the box of the current coordinates
             * (the steps queued so far finish first); once it no longer
             * holds, this run rebuilds as the reference does */
            std::vector<double> xh(3 * (size_t)E.natoms);
            if ((rc = engine_d2h(E, xh.data(), E.x_dev, xh.size() * sizeof(double)))) return rc;
            md_static_bound(E, M, xh.data(), step - 1);
            recheck = M.static_until > step - 1;
        }
        if (rebuild && step <= M.static_until) {
            rebuild = false;
            M.rebuilds_skipped++;
        } else if (rebuild) {
            M.rebuilds_run++;
        }
        if (use_graph) HIPCHK(hipGraphLaunch(rebuild ? M.g_rebuild : M.g_step, E.stream));
        else if ((rc = md_step_launch(ctx, rebuild))) return rc;
        M.steps_done = step;
    }
    return 0;
}

extern "C" int gbhip_md_stats(gbhip_ctx *ctx, long long *rebuilds, long long *skipped, long long *bbox_checks)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    if (rebuilds) *rebuilds = ctx->M.rebuilds_run;
    if (skipped) *skipped = ctx->M.rebuilds_skipped;
    if (bbox_checks) *bbox_checks = ctx->M.bbox_checks;
    return 0;
}

extern "C" int gbhip_md_state(gbhip_ctx *ctx, double *x, double *v, double *grad)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    Engine &E = ctx->E;
    MDState &M = ctx->M;
    if (!M.ready) return seterr(GBHIP_ESTATE, "gbhip_md_init has not been called");
    const size_t nb = 3 * sizeof(double) * (size_t)E.natoms;
    int rc;
    if (x && (rc = engine_d2h(E, x, E.x_dev, nb))) return rc;
    if (v && (rc = engine_d2h(E, v, M.v, nb))) return rc;
    if (grad && (rc = engine_d2h(E, grad, E.g_dev, nb))) return rc;
    HIPCHK(hipStreamSynchronize(E.stream));
    return 0;
}

extern "C" int gbhip_md_log(gbhip_ctx *ctx, double *rows, int max_rows, int *nrows)
{
    if (!ctx || !rows || !nrows) return seterr(GBHIP_EINVAL, "NULL argument");
    MDState &M = ctx->M;
    Engine &E = ctx->E;
    const int n = std::min(std::min(M.steps_done, M.log_cap), max_rows);
    /* rows of the last n steps in step order */
    std::vector<double> all(16 * (size_t)M.log_cap);
    const int rc = engine_d2h(E, all.data(), M.log, all.size() * sizeof(double));
    if (rc) return rc;
    for (int r = 0; r < n; r++) {
        const int step = M.steps_done - n + 1 + r;
        const double *src = all.data() + 16 * (size_t)((step - 1) % M.log_cap);
        for (int k = 0; k < 16; k++) rows[16 * r + k] = src[k];
    }
    *nrows = n;
    return 0;
}
