/*
 * cg_device.hip -- conjgrad() with its vectors resident on the device.
 *
 * The minimiser of conjgrad.c:34-263 (Polak-Ribiere directions with
 * Powell restarts, a cubic line search, the reference's convergence and
 * failure tests), as the host restatement in sff_host.c runs it, but with
 * the coordinates, gradient, search and restart directions, the gradient
 * at the line start and the best point kept on the device: every vector
 * update is a kernel on the context's stream, every mme() is the captured
 * evaluation (engine_mme), and only scalars cross to the host -- the
 * energies of each evaluation and the dot products and step size the
 * line search branches on.  Dot products are reduced in a fixed order
 * (per-block partial sums over a fixed grid, then one block), so a run is
 * reproducible.  Per evaluation the host waits once: the energies and
 * every product of the new gradient the line search and the next
 * direction can need come back together.
 *
 * A line-search step x = xopt + stepch d is evaluated together with the
 * test that it moved x at all (the reference's "work > 0", conjgrad.c:
 * L50): the step, the evaluation and the test go to the device in one
 * batch, and an evaluation whose step did not move x is dropped (x is then
 * xopt, the gradient is put back from the copy of the best point, and a
 * pair-list rebuild it made is made again at the next evaluation, where
 * the reference makes it).
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "engine.h"

#define CG_THREADS 256
#define CG_BLOCKS 256 /* fixed grid of the reductions: fixed summation order */
#define CG_MAXDOT 5

/* slots of CGState::res */
enum { R_DOT0 = 0, R_WORK = 5, R_NRES = 8 };

__device__ __forceinline__ double cg_block_sum(double v, double *s_red)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < CG_THREADS / 64; k++) t += s_red[k];
    __syncthreads();
    return t;
}

__device__ __forceinline__ double cg_block_max(double v, double *s_red)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < CG_THREADS / 64; k++) t = fmax(t, s_red[k]);
    __syncthreads();
    return t;
}

/* nd dot products a[k].b[k] over n entries: per-block partial sums */
struct DotArgs {
    const double *a[CG_MAXDOT], *b[CG_MAXDOT];
};

__global__ __launch_bounds__(CG_THREADS) void k_cg_dots(DotArgs A, int nd, long long n, double *part)
{
    __shared__ double s_red[CG_THREADS / 64];
    double acc[CG_MAXDOT] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (long long i = (long long)blockIdx.x * CG_THREADS + threadIdx.x; i < n; i += (long long)CG_BLOCKS * CG_THREADS)
        for (int k = 0; k < CG_MAXDOT; k++)
            if (k < nd) acc[k] += A.a[k][i] * A.b[k][i];
    for (int k = 0; k < CG_MAXDOT; k++) {
        if (k >= nd) break;
        const double t = cg_block_sum(acc[k], s_red);
        if (threadIdx.x == 0) part[k * CG_BLOCKS + blockIdx.x] = t;
    }
}

/* x = xopt + s d, and per block the largest |x - xopt| */
__global__ __launch_bounds__(CG_THREADS) void k_cg_step(double *x, const double *xopt, const double *d, double s,
                                                        long long n, double *part)
{
    __shared__ double s_red[CG_THREADS / 64];
    double w = 0.0;
    for (long long i = (long long)blockIdx.x * CG_THREADS + threadIdx.x; i < n; i += (long long)CG_BLOCKS * CG_THREADS) {
        const double xi = xopt[i] + s * d[i];
        x[i] = xi;
        w = fmax(w, fabs(xi - xopt[i]));
    }
    const double t = cg_block_max(w, s_red);
    if (threadIdx.x == 0) part[CG_MAXDOT * CG_BLOCKS + blockIdx.x] = t;
}

/* one block: res[k] = sum of the nd rows of partials (fixed order); with
 * want_work, res[R_WORK] = max of the step partials */
__global__ __launch_bounds__(CG_THREADS) void k_cg_finish(const double *part, int nd, int want_work, double *res)
{
    __shared__ double s_red[CG_THREADS / 64];
    static_assert(CG_BLOCKS == CG_THREADS, "one partial per thread");
    for (int k = 0; k < nd; k++) {
        const double t = cg_block_sum(part[k * CG_BLOCKS + threadIdx.x], s_red);
        if (threadIdx.x == 0) res[R_DOT0 + k] = t;
    }
    if (want_work) {
        const double t = cg_block_max(part[CG_MAXDOT * CG_BLOCKS + threadIdx.x], s_red);
        if (threadIdx.x == 0) res[R_WORK] = t;
    }
}

__global__ __launch_bounds__(CG_THREADS) void k_cg_fill(double *p, double v, long long n)
{
    const long long i = (long long)blockIdx.x * CG_THREADS + threadIdx.x;
    if (i < n) p[i] = v;
}

/* d = -g (+ beta d (+ gama r)) */
__global__ __launch_bounds__(CG_THREADS) void k_cg_dir(double *d, const double *g, double beta, const double *r,
                                                       double gama, long long n)
{
    const long long i = (long long)blockIdx.x * CG_THREADS + threadIdx.x;
    if (i >= n) return;
    double v = -g[i];
    if (beta != 0.0) v += beta * d[i];
    if (r) v += gama * r[i];
    d[i] = v;
}

/* Powell restart vectors: rsdx = d, rsdg = g - ginit; then d = -g + beta d */
__global__ __launch_bounds__(CG_THREADS) void k_cg_restart(double *d, double *rsdx, double *rsdg, const double *g,
                                                           const double *ginit, double beta, long long n)
{
    const long long i = (long long)blockIdx.x * CG_THREADS + threadIdx.x;
    if (i >= n) return;
    const double di = d[i], gi = g[i];
    rsdx[i] = di;
    rsdg[i] = gi - ginit[i];
    d[i] = -gi + beta * di;
}

namespace {

struct CGRun {
    gbhip_ctx *c;
    Engine &E;
    long long n;
    int nblk;
    int nsnb;
    bool force_rebuild = false;
    gbhip_cg_report report;
    void *user;

    CGRun(gbhip_ctx *ctx, long long n_, int nsnb_, gbhip_cg_report rep, void *u)
        : c(ctx), E(ctx->E), n(n_), nblk((int)((n_ + CG_THREADS - 1) / CG_THREADS)), nsnb(nsnb_), report(rep),
          user(u)
    {
    }

    int copy(double *dst, const double *src)
    {
        HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToDevice, E.stream));
        return 0;
    }

    int launch_dots(int nd, const double *const *a, const double *const *b)
    {
        DotArgs A;
        for (int k = 0; k < CG_MAXDOT; k++) {
            A.a[k] = k < nd ? a[k] : nullptr;
            A.b[k] = k < nd ? b[k] : nullptr;
        }
        hipLaunchKernelGGL(k_cg_dots, dim3(CG_BLOCKS), dim3(CG_THREADS), 0, E.stream, A, nd, n, c->C.part);
        LAUNCHCHK();
        return 0;
    }

    /* the scalars (and optionally the energies) back to the host: one wait */
    int read(double *res, int nres, double *ene14)
    {
        double *h = E.h_stage + 6 * (size_t)E.natoms; /* the last 64 doubles of the staging */
        HIPCHK(hipMemcpyAsync(h, c->C.res, sizeof(double) * nres, hipMemcpyDeviceToHost, E.stream));
        if (ene14) HIPCHK(hipMemcpyAsync(h + 16, E.ene14_dev, 14 * sizeof(double), hipMemcpyDeviceToHost, E.stream));
        HIPCHK(hipStreamSynchronize(E.stream));
        memcpy(res, h, sizeof(double) * nres);
        if (ene14) memcpy(ene14, h + 16, 14 * sizeof(double));
        return 0;
    }

    int dots(int nd, const double *const *a, const double *const *b, double *out)
    {
        int rc = launch_dots(nd, a, b);
        if (rc) return rc;
        hipLaunchKernelGGL(k_cg_finish, dim3(1), dim3(CG_THREADS), 0, E.stream, c->C.part, nd, 0, c->C.res);
        LAUNCHCHK();
        double r[R_NRES];
        if ((rc = read(r, nd, nullptr))) return rc;
        for (int k = 0; k < nd; k++) out[k] = r[R_DOT0 + k];
        return 0;
    }

    /*
     * Evaluation at the current x (after a step when step != nullptr):
     * the pair list rebuilt at the reference's cadence for iteration it,
     * mme, then the products of the new gradient g the minimiser can need
     * before the next evaluation -- d.g, g.g and, for the next direction,
     * g.ginit, g.rsdg, g.rsdx (with the vectors as they are during this
     * line search); one wait.  Returns in *work the step's largest
     * coordinate change (step evaluations only) and whether it rebuilt.
     */
    int eval(int it, const double *step, double *ene14, double *pr, double *work, bool *rebuilt)
    {
        int rc;
        if (step) {
            hipLaunchKernelGGL(k_cg_step, dim3(CG_BLOCKS), dim3(CG_THREADS), 0, E.stream, E.x_dev, c->C.xopt,
                               c->C.d, *step, n, c->C.part);
            LAUNCHCHK();
        }
        *rebuilt = false;
        if (!E.have_list || (it > 0 && it % nsnb == 0) || force_rebuild) {
            if ((rc = gbhip_nblist(c, nullptr))) return rc;
            *rebuilt = true;
            force_rebuild = false;
        }
        if ((rc = engine_mme(c))) return rc;
        const double *a[5] = {c->C.d, E.g_dev, E.g_dev, E.g_dev, E.g_dev};
        const double *b[5] = {E.g_dev, E.g_dev, c->C.ginit, c->C.rsdg, c->C.rsdx};
        if ((rc = launch_dots(5, a, b))) return rc;
        hipLaunchKernelGGL(k_cg_finish, dim3(1), dim3(CG_THREADS), 0, E.stream, c->C.part, 5, step ? 1 : 0,
                           c->C.res);
        LAUNCHCHK();
        double r[R_NRES];
        if ((rc = read(r, R_NRES, ene14))) return rc;
        for (int k = 0; k < 5; k++) pr[k] = r[R_DOT0 + k];
        if (work) *work = r[R_WORK];
        return 0;
    }
};

enum cg_state { CG_EVAL, CG_RESET_DIR, CG_TEST, CG_NEW_SEARCH, CG_STEP, CG_SPLINE, CG_NEXT_STEP, CG_NEW_DIR,
                CG_FINISH, CG_DONE };

} // namespace

extern "C" int gbhip_conjgrad(gbhip_ctx *ctx, double *x, int n, double *f, double rmsgrad, double dfpred,
                              int maxiter, int nsnb, gbhip_cg_report report, void *user, int *result)
{
    if (!ctx || !x || !f || !result) return seterr(GBHIP_EINVAL, "NULL argument");
    Engine &E = ctx->E;
    if (E.nmol != 1 || n != 3 * E.natoms) return seterr(GBHIP_EINVAL, "n must be 3 * natom of a one-molecule context");
    if (!ctx->B.set) return seterr(GBHIP_ESTATE, "gbhip_set_bonded has not been called");
    if (nsnb < 1) nsnb = 1;
    int rc;
    if (!E.stream) {
        /* graphs and the staging need a stream of the engine's own */
        HIPCHK(hipStreamSynchronize(nullptr));
        hipStream_t s;
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        E.stream = s;
        E.own_stream = true;
    }
    if (!engine_stage(E)) return seterr(GBHIP_ENOMEM, "no pinned staging for the minimiser's scalars");
    CGState &C = ctx->C;
    if (!C.d) {
        std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
        if ((rc = E.dalloc(&C.d, n)) || (rc = E.dalloc(&C.rsdx, n)) || (rc = E.dalloc(&C.rsdg, n)) ||
            (rc = E.dalloc(&C.ginit, n)) || (rc = E.dalloc(&C.xopt, n)) || (rc = E.dalloc(&C.gopt, n)) ||
            (rc = E.dalloc(&C.part, (CG_MAXDOT + 1) * CG_BLOCKS)) || (rc = E.dalloc(&C.res, R_NRES)))
            return rc;
    }
    CGRun R(ctx, n, nsnb, report, user);
    if ((rc = E.h2d(E.x_dev, x, sizeof(double) * n))) return rc;
    /* the direction starts at zero (it is first set from the gradient) */
    hipLaunchKernelGGL(k_cg_fill, dim3(R.nblk), dim3(CG_THREADS), 0, E.stream, C.d, 0.0, (long long)n);
    LAUNCHCHK();

    const int maxlin = 10, mxfcon = 4;
    const double dgrad = rmsgrad * rmsgrad * n;
    double gama = 0, beta = 0, fmin = 0, gmin = 0, dfpr = 0, gnew = 0, step = 0, work = 0, finit = 0, ginit = 0;
    double gsqrd = 0, gspln, stmin = 0, gamden = 0, ddspln = 0, stepch = 0, sbound = 0, fch = 0, sum = 0;
    int niter = 0, nfbeg = 0, iterc = 0, iterfm = 0, iterrs = 0, iretry = 0, nfopt = 0, ier = 0;
    /*
     * Products of the gradient now in g (pc) and of the best point's (po):
     * [0] d.g, [1] g.g, [2] g.ginit, [3] g.rsdg, [4] g.rsdx.  Each set
     * carries the epochs of d ([0]) and of ginit/rsdg/rsdx ([2..4]) it was
     * formed with; a vector update bumps its epoch, d.g of a new direction
     * follows from the old products (d' = -g + beta d + gama rsdx gives
     * d'.g = -g.g + beta d.g + gama g.rsdx), and a product whose epoch is
     * stale when it is needed is formed again on the device.
     */
    struct Prod {
        double v[5];
        long ed, ev;
    } pc{}, po{};
    long ed = 0, ev = 0;
    double ene14[14];
    bool rebuilt = false;
    cg_state st = CG_EVAL;

    auto refresh = [&](bool want_d, bool want_v) -> int {
        if (!(want_d && pc.ed != ed) && !(want_v && pc.ev != ev)) return 0;
        const double *a[5] = {C.d, E.g_dev, E.g_dev, E.g_dev, E.g_dev};
        const double *b[5] = {E.g_dev, E.g_dev, C.ginit, C.rsdg, C.rsdx};
        const int r2 = R.dots(5, a, b, pc.v);
        pc.ed = ed;
        pc.ev = ev;
        return r2;
    };
    auto save_best = [&]() -> int {
        int r2;
        fmin = *f;
        gsqrd = sum;
        nfopt = niter;
        if ((r2 = R.copy(C.xopt, E.x_dev)) || (r2 = R.copy(C.gopt, E.g_dev))) return r2;
        po = pc;
        return 0;
    };

    while (st != CG_DONE) {
        switch (st) {
        case CG_EVAL: { /* L10: the first evaluation (later ones come from CG_STEP) */
            ++niter;
            if ((rc = R.eval(niter - 1, nullptr, ene14, pc.v, nullptr, &rebuilt))) return rc;
            pc.ed = ed;
            pc.ev = ev;
            *f = ene14[0];
            if (report) report(niter - 1, ene14, user);
            st = niter >= 2 ? CG_TEST : CG_RESET_DIR;
            break;
        }
        case CG_RESET_DIR: /* L20 */
            hipLaunchKernelGGL(k_cg_dir, dim3(R.nblk), dim3(CG_THREADS), 0, E.stream, C.d, (const double *)E.g_dev,
                               0.0, nullptr, 0.0, (long long)n);
            LAUNCHCHK();
            ++ed;
            pc.v[0] = -pc.v[1]; /* (-g).g summed in the order g.g is */
            pc.ed = ed;
            iterrs = 0;
            st = iterc > 0 ? CG_NEW_SEARCH : CG_TEST;
            break;
        case CG_TEST: /* L30 */
            if ((rc = refresh(true, false))) return rc;
            gnew = pc.v[0];
            sum = pc.v[1];
            if (niter == 1) {
                if ((rc = save_best())) return rc;
                if (sum <= dgrad) { st = CG_DONE; break; }
            } else {
                fch = *f - fmin;
                if (fch < 0.0 || (fch == 0.0 && gnew / gmin >= -1.0)) {
                    if ((rc = save_best())) return rc;
                    if (sum <= dgrad) { st = CG_DONE; break; }
                }
            }
            if (niter == maxiter + 1) { ier = -3; st = CG_DONE; break; }
            if (niter > 1) { st = CG_SPLINE; break; }
            dfpr = dfpred;
            stmin = dfpred / gsqrd;
            st = CG_NEW_SEARCH;
            break;
        case CG_NEW_SEARCH: /* L40 */
            ++iterc;
            finit = *f;
            if ((rc = refresh(true, false))) return rc;
            if ((rc = R.copy(C.ginit, E.g_dev))) return rc;
            ++ev;
            ginit = pc.v[0];
            if (ginit >= 0.0) { ier = -2; st = CG_FINISH; break; }
            gmin = ginit;
            sbound = -1.0;
            nfbeg = niter;
            iretry = -1;
            {
                const double lim = fabs(dfpr / ginit);
                stepch = stmin <= lim ? stmin : lim;
            }
            stmin = 0.0;
            st = CG_STEP;
            break;
        case CG_STEP: { /* L50, with the evaluation of L10 when the step moved x */
            step = stmin + stepch;
            double e14[14];
            Prod pn;
            if ((rc = R.eval(niter, &stepch, e14, pn.v, &work, &rebuilt))) return rc;
            if (work > 0.0) {
                ++niter;
                pc = pn;
                pc.ed = ed;
                pc.ev = ev;
                memcpy(ene14, e14, sizeof(ene14));
                *f = ene14[0];
                if (report) report(niter - 1, ene14, user);
                st = CG_TEST; /* niter >= 2 here */
                break;
            }
            /* x did not move: the evaluation is dropped, as the reference
             * never makes it; the gradient goes back to the best point's */
            if (rebuilt) R.force_rebuild = true;
            if ((rc = R.copy(E.g_dev, C.gopt))) return rc;
            pc = po;
            if (niter > nfbeg + 1) ier = -1;
            if (fabs(gmin / ginit) > 0.2) ier = -1;
            st = CG_FINISH;
            break;
        }
        case CG_SPLINE: /* L60 */
            work = (fch + fch) / stepch - gnew - gmin;
            ddspln = (gnew - gmin) / stepch;
            if (niter > nfopt) sbound = step;
            if (niter <= nfopt) {
                if (gmin * gnew <= 0.0) sbound = stmin;
                stmin = step;
                gmin = gnew;
                stepch = -stepch;
            }
            if (fch != 0.0) ddspln += (work + work) / stepch;
            if (gmin == 0.0) { st = CG_FINISH; break; }
            if (niter > nfbeg + 1) {
                if (fabs(gmin / ginit) <= 0.2) { st = CG_FINISH; break; }
                if (niter >= nfopt + maxlin) { ier = -1; st = CG_FINISH; break; }
            }
            st = CG_NEXT_STEP;
            break;
        case CG_NEXT_STEP: /* L70 */
            stepch = (sbound - stmin) * 0.5;
            if (sbound < -0.5) stepch = stmin * 9.0;
            gspln = gmin + stepch * ddspln;
            if (gmin * gspln < 0.0) stepch = stepch * gmin / (gmin - gspln);
            st = CG_STEP;
            break;
        case CG_NEW_DIR: { /* L80 */
            if ((rc = refresh(true, true))) return rc;
            sum = pc.v[2];
            beta = (gsqrd - sum) / (gmin - ginit);
            if (fabs(beta * gmin) > gsqrd * 0.2) {
                ++iretry;
                if (iretry <= 0) {
                    if (niter >= nfopt + maxlin) { ier = -1; st = CG_FINISH; }
                    else st = CG_NEXT_STEP;
                    break;
                }
            }
            if (*f < finit) iterfm = iterc;
            if (iterc >= iterfm + mxfcon) { ier = -4; st = CG_DONE; break; }
            dfpr = stmin * ginit;
            if (iretry > 0) { st = CG_RESET_DIR; break; }
            if (iterrs != 0 && iterc - iterrs < n && fabs(sum) < gsqrd * 0.2) {
                gama = pc.v[3] / gamden;
                if (fabs(beta * gmin + gama * pc.v[4]) < gsqrd * 0.2) {
                    hipLaunchKernelGGL(k_cg_dir, dim3(R.nblk), dim3(CG_THREADS), 0, E.stream, C.d,
                                       (const double *)E.g_dev, beta, (const double *)C.rsdx, gama, (long long)n);
                    LAUNCHCHK();
                    ++ed;
                    pc.v[0] = -pc.v[1] + beta * pc.v[0] + gama * pc.v[4];
                    pc.ed = ed;
                    st = CG_NEW_SEARCH;
                    break;
                }
            }
            gamden = gmin - ginit;
            hipLaunchKernelGGL(k_cg_restart, dim3(R.nblk), dim3(CG_THREADS), 0, E.stream, C.d, C.rsdx, C.rsdg,
                               (const double *)E.g_dev, (const double *)C.ginit, beta, (long long)n);
            LAUNCHCHK();
            ++ed;
            ++ev;
            pc.v[0] = -pc.v[1] + beta * pc.v[0];
            pc.ed = ed;
            iterrs = iterc;
            st = CG_NEW_SEARCH;
            break;
        }
        case CG_FINISH: /* L90 */
            if (niter != nfopt) {
                *f = fmin;
                if ((rc = R.copy(E.x_dev, C.xopt)) || (rc = R.copy(E.g_dev, C.gopt))) return rc;
                pc = po;
            }
            st = ier == 0 ? CG_NEW_DIR : CG_DONE;
            break;
        default:
            st = CG_DONE;
        }
    }
    if ((rc = engine_d2h(E, x, E.x_dev, sizeof(double) * n))) return rc;
    *result = ier == 0 ? niter : ier;
    return 0;
}
