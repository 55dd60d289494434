/*
 * sff_host.c -- libsff-compatible host API over the GPU engine.
 *
 *   mme_init_sff  sff.c:858-1846: GB parameters (sff.c:1187-1333 with the
 *                 gb 8 element constants of gbneck.h:150-170), dielc charge
 *                 scaling (sff.c:1153-1162), bonded/1-4 term tables
 *   mme           sff.c:1854 -> mme34 eff.c:2442-2947: list rebuild cadence
 *                 (eff.c:2564), one device evaluation, ff: / per-term output
 *                 (eff.c:2899-2927), gb_debug radii (eff.c:1729-1731)
 *   md            sff.c:2071-2366 velocity Verlet / Langevin; noise from a
 *                 counter-based normal stream (see md_gauss below) instead
 *                 of the serial rand2.c generator
 *   conjgrad      conjgrad.c:34-263 (restated as a state machine below)
 *
 * The GPU context is created at the first mme() call (its coordinates set
 * the spatial tile order) and released by mme(x, f, -3) (eff.c:2504-2521).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gbhip.h"
#include "host_internal.h"

static struct {
    nabc_model m;       /* GB parameters and bonded tables (model.c) */
    gbhip_ctx *ctx;     /* created at the first mme() call */
    int have_list;
    long long nb_pairs;
    double ene[14];     /* last mme(): mme34 energies + rms gradient */
    gbhip_ctx *ctx2;    /* mme2(): an FP64 context of its own */
    nabc_nctraj traj;   /* md() trajectory, when mme_init_sff was given a file name */
} S;

static FILE *out(void) { return nabout ? nabout : stdout; }

static void release_ctx(void)
{
    if (S.ctx) gbhip_ctx_destroy(S.ctx);
    S.ctx = NULL;
    S.have_list = 0;
    if (S.ctx2) gbhip_ctx_destroy(S.ctx2);
    S.ctx2 = NULL;
}

int mme_init_sff(PARMSTRUCT_T *prm, int *frozen, int *constrained, double *x0, char *ncname)
{
    int rc;
    (void)x0;
    (void)constrained;
    release_ctx();
    nabc_model_free(&S.m);
    nabc_nctraj_close(&S.traj);
    rc = nabc_model_build(&S.m, prm, frozen, &nabc_opts);
    /* the trajectory file md() writes every ntwx steps (sff.c:1164-1168) */
    if (rc == 0 && ncname) rc = nabc_nctraj_create(&S.traj, ncname, prm->Natom);
    return rc;
}

static int ensure_ctx(const double *x)
{
    if (S.ctx) return 0;
    return nabc_model_context(&S.m, x, &nabc_opts, NULL, &S.ctx);
}

void *mme_context(double *x)
{
    if (!S.m.prm) {
        nabc_seterr("mme_context: mme_init_sff has not been called");
        return NULL;
    }
    if (ensure_ctx(x) != 0) return NULL;
    return S.ctx;
}

void mme_get_energies(double ene[13], double *frms)
{
    if (ene) memcpy(ene, S.ene, sizeof(double) * 13);
    if (frms) *frms = S.ene[13];
}

/* the ff: line of an evaluation (eff.c:2899-2927): the header at iteration
 * 0, then every ntpr iterations */
static void mme_report(int it, const double *e)
{
    const nabc_opts_t *o = &nabc_opts;
    if (it == 0) fprintf(out(), "      iter    Total       bad      vdW     elect   nonpolar   genBorn      frms\n");
    if (it > -1 && (it == 0 || it % o->ntpr == 0)) {
        fprintf(out(), "ff:%6d %9.2f %9.2f %9.2f %9.2f %9.2f %9.2f %9.2e\n", it, e[0], e[3] + e[4] + e[5],
                e[1] + e[7], e[2] + e[8], e[9] + e[11] + e[12], e[10], e[13]);
        fflush(out());
    }
}

/* mme() / mme34(), sff.c:1854, eff.c:2442-2947 */
double mme(double *x, double *f, int *iter)
{
    const nabc_opts_t *o = &nabc_opts;
    int it = *iter;
    if (it == -3) {
        release_ctx();
        return 0.0;
    }
    if (!S.m.prm) {
        nabc_seterr("mme: mme_init_sff has not been called");
        return NAN;
    }
    if (ensure_ctx(x) != 0) return NAN;
    if (gbhip_set_coords(S.ctx, x) != 0) {
        nabc_seterr("%s", gbhip_last_error());
        return NAN;
    }
    if (!S.have_list || (it > 0 && it % (o->nsnb > 0 ? o->nsnb : 1) == 0) || it == -4) {
        long long np = 0;
        if (gbhip_nblist(S.ctx, &np) != 0) {
            nabc_seterr("%s", gbhip_last_error());
            return NAN;
        }
        S.nb_pairs = np;
        S.have_list = 1;
    }
    if (gbhip_mme(S.ctx, f, S.ene) != 0) {
        nabc_seterr("%s", gbhip_last_error());
        return NAN;
    }
    mme_report(it, S.ene);
    if (o->gb_debug) {
        int i, n = S.m.prm->Natom;
        double *reff = (double *)malloc(sizeof(double) * n);
        if (gbhip_get_reff(S.ctx, reff) == 0) {
            fprintf(out(), "Effective Born radii:\n");
            for (i = 0; i < n; i++) fprintf(out(), "%d\t%15.7f\t%15.7f\n", i + 1, S.m.prm->Rborn[i], reff[i]);
        }
        free(reff);
    }
    if (it == -1) {
        const double *e = S.ene;
        fprintf(out(), "     bond:  %15.9f\n", e[3]);
        fprintf(out(), "    angle:  %15.9f\n", e[4]);
        fprintf(out(), " dihedral:  %15.9f\n", e[5]);
        fprintf(out(), "    enb14:  %15.9f\n", e[7]);
        fprintf(out(), "    eel14:  %15.9f\n", e[8]);
        fprintf(out(), "      enb:  %15.9f\n", e[1]);
        fprintf(out(), "      eel:  %15.9f\n", e[2]);
        fprintf(out(), "      egb:  %15.9f\n", e[10]);
        fprintf(out(), "    econs:  %15.9f\n", e[9]);
        fprintf(out(), "    esurf:  %15.9f\n", e[11]);
        fprintf(out(), "    Total:  %15.9f\n", e[0]);
        fprintf(out(), "     frms:  %15.9f\n", e[13]);
    }
    return S.ene[0];
}

/* ------------------------------------------------------------------ mme2 */

#define HESS_STEP 1.0e-4 /* A, central-difference step of the Hessian columns */

/*
 * mme2() (sff2.c:6176-6835, non-ScaLAPACK): energy, gradient, Hessian,
 * masses and atom-name letters.  The pair list is rebuilt at every call
 * (the Hessian needs it at x) and the evaluation is FP64 whatever gbprec
 * says: the columns are differences of gradients.  The header line at
 * iterations 0 and 1 and the ff: line at iteration 1 and every ntpr-th
 * (sff2.c:6770-6775) go out as in the reference.
 */
double mme2(double *x, double *f, double *g, double *m, double *grad, int *descF_PxQ, int *descF_1x1,
            int *descG_PxQ, int *context_PxQ, int *context_1x1, int *context_Nx1, int *gridim, int *natom,
            int *iter, char *name)
{
    const nabc_opts_t *o = &nabc_opts;
    const int it = iter ? *iter : 0;
    int i, n;
    long long np = 0;
    (void)descF_PxQ; (void)descF_1x1; (void)descG_PxQ;
    (void)context_PxQ; (void)context_1x1; (void)context_Nx1; (void)gridim;
    if (it == -3) {
        if (S.ctx2) gbhip_ctx_destroy(S.ctx2);
        S.ctx2 = NULL;
        return 0.0;
    }
    if (!S.m.prm) {
        nabc_seterr("mme2: mme_init_sff has not been called");
        return NAN;
    }
    n = S.m.prm->Natom;
    if (natom) *natom = n;
    for (i = 0; i < n; i++) {
        if (m) m[i] = S.m.prm->Masses[i];
        if (name) name[i] = S.m.prm->AtomNames ? S.m.prm->AtomNames[i][0] : ' ';
    }
    if (!S.ctx2) {
        nabc_opts_t o64 = *o;
        o64.gbprec = 0;
        if (nabc_model_context(&S.m, x, &o64, NULL, &S.ctx2) != 0) return NAN;
    }
    if (gbhip_set_coords(S.ctx2, x) != 0 || gbhip_nblist(S.ctx2, &np) != 0 ||
        gbhip_hessian(S.ctx2, HESS_STEP, g, f, S.ene) != 0) {
        nabc_seterr("%s", gbhip_last_error());
        return NAN;
    }
    if (grad && grad != f) memcpy(grad, f, sizeof(double) * 3 * (size_t)n);
    if (it == 0 || it == 1)
        fprintf(out(), "      iter    Total       bad      vdW     elect   nonpolar   genBorn      frms\n");
    if (it > -1 && (it == 1 || it % o->ntpr == 0)) {
        const double *e = S.ene;
        printf("ff:%6d %9.2f %9.2f %9.2f %9.2f %9.2f %9.2f %9.2e\n", it, e[0], e[3] + e[4] + e[5], e[1] + e[7],
               e[2] + e[8], e[9] + e[11] + e[12], e[10], e[13]);
        fflush(stdout);
    }
    return S.ene[0];
}

/* ------------------------------------------------------------------- md */

/*
 * Langevin noise: one N(0,1) deviate per (step, coordinate) from a
 * counter-based generator (splitmix64 of seed, step and index, then
 * Box-Muller), so a trajectory does not depend on call order.
 * rand2.c's serial stream is not reproduced (MD parity is statistical).
 */
static unsigned long long mix64(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double nabc_md_gauss(unsigned long long seed, long long step, long long idx)
{
    unsigned long long h = mix64(seed ^ mix64((unsigned long long)step * 0x100000001B3ULL ^ mix64((unsigned long long)idx)));
    unsigned long long h2 = mix64(h);
    double u1 = ((h >> 11) + 1.0) * (1.0 / 9007199254740993.0);
    double u2 = (h2 >> 11) * (1.0 / 9007199254740992.0);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/*
 * md() over this library's own mme(): the whole loop runs on the device
 * (gbhip_md_*), each step a hipGraph launch; the md: lines are printed
 * from the device's per-step log afterwards.
 */
static int md_device(int n, int maxstep, double *x, double *f, double *v)
{
    const nabc_opts_t *o = &nabc_opts;
    gbhip_md_options mo;
    int cap = maxstep > 0 ? maxstep : 1, rows = 0, k, it;
    double *log;
    if (ensure_ctx(x) != 0) return -1;
    memset(&mo, 0, sizeof(mo));
    mo.dt = o->dt;
    mo.temp0 = o->temp0;
    mo.tempi = o->tempi;
    mo.gamma_ln = o->gamma_ln;
    mo.tautp = o->tautp;
    mo.vlimit = o->vlimit;
    mo.boltz2 = o->boltz2;
    mo.nsnb = o->nsnb > 0 ? o->nsnb : 1;
    mo.zerov = o->zerov;
    mo.seed = o->md_seed;
    if (cap > (1 << 20)) cap = 1 << 20;
    if (gbhip_md_init(S.ctx, S.m.prm->Masses, x, v, &mo, cap) != 0) {
        nabc_seterr("%s", gbhip_last_error());
        return -1;
    }
    if (S.traj.fp && o->ntwx > 0) {
        /* frames every ntwx steps (sff.c:2341-2342): the device runs in
         * chunks that end on a frame step */
        int done = 0;
        while (done < maxstep) {
            const int chunk = maxstep - done < o->ntwx ? maxstep - done : o->ntwx;
            if (gbhip_md_run(S.ctx, chunk, 1) != 0 || gbhip_md_state(S.ctx, x, NULL, NULL) != 0) {
                nabc_seterr("%s", gbhip_last_error());
                return -1;
            }
            done += chunk;
            if (done % o->ntwx == 0 && nabc_nctraj_write(&S.traj, x, o->t + done * o->dt) != 0) return -1;
        }
    } else if (gbhip_md_run(S.ctx, maxstep, 1) != 0) {
        nabc_seterr("%s", gbhip_last_error());
        return -1;
    }
    if (gbhip_md_state(S.ctx, x, v, f) != 0) {
        nabc_seterr("%s", gbhip_last_error());
        return -1;
    }
    S.have_list = 1;
    log = (double *)malloc(sizeof(double) * 16 * (size_t)cap);
    if (gbhip_md_log(S.ctx, log, cap, &rows) == 0) {
        for (k = 0; k < rows; k++) {
            const int step = maxstep - rows + 1 + k;
            const double *r = log + 16 * (size_t)k;
            if (step % o->ntpr_md == 0 || step == 1)
                fprintf(out(), "md:       %5d %10.3f %10.2f %10.2f %10.2f %10.2f\n", step, o->t + step * o->dt,
                        r[14], r[0], r[14] + r[0], r[15]);
        }
        if (rows > 0) memcpy(S.ene, log + 16 * (size_t)(rows - 1), sizeof(double) * 14);
    }
    free(log);
    fflush(out());
    nabc_opts.t += maxstep * o->dt;
    it = -3;
    mme(x, f, &it);
    return 0;
}

/* md(), sff.c:2071-2366 (no rattle, no COM removal); trajectory frames
 * every ntwx steps when mme_init_sff was given a file name */
int md(int n, int maxstep, double *x, double *f, double *v, double (*func)(double *, double *, int *))
{
    const nabc_opts_t *o = &nabc_opts;
    double *minv, *accel, *sqrmass;
    double dtx = o->dt * 20.455, dt5 = 0.5 * dtx, rndf, ekin0, dttp, gammai, c_implic, c_explic, c_ave, sdfac;
    double ekin = 0.0, epot, tscal, temp, t = o->t;
    int i, nstep, dim = 3;
    unsigned long long seed = o->md_seed;
    if (!S.m.prm && func == mme) {
        nabc_seterr("md: mme_init_sff has not been called");
        return -1;
    }
    if (func == mme && n == 3 * S.m.prm->Natom) return md_device(n, maxstep, x, f, v);
    minv = (double *)calloc(n, sizeof(double));
    accel = (double *)calloc(n, sizeof(double));
    sqrmass = (double *)calloc(n, sizeof(double));
    rndf = n - 3 * S.m.nfrozen;
    ekin0 = o->boltz2 * rndf * o->temp0;
    dttp = o->dt / o->tautp;
    gammai = o->gamma_ln / 20.455;
    c_implic = 1.0 / (1.0 + gammai * dt5);
    c_explic = 1.0 - gammai * dt5;
    c_ave = 1.0 + gammai * dt5;
    sdfac = sqrt(4.0 * gammai * o->boltz2 * o->temp0 / dtx);
    if (S.m.prm) {
        for (i = 0; i < n; i++) minv[i] = 1.0 / S.m.prm->Masses[i / dim];
    } else {
        for (i = 0; i < n; i++) minv[i] = 1.0 / o->genmass;
    }
    if (o->zerov) {
        for (i = 0; i < n; i++) v[i] = 0.0;
        ekin = 0.0;
    } else if (o->tempi > 0.0) {
        for (i = 0; i < n; i++) {
            if (S.m.frozen && S.m.frozen[i / dim]) v[i] = 0.0;
            else {
                double sd = sqrt(2. * o->boltz2 * o->tempi * minv[i]);
                v[i] = sd * nabc_md_gauss(seed, -1, i);
                ekin += v[i] * v[i] / minv[i];
            }
        }
        ekin *= 0.5;
    } else {
        for (i = 0; i < n; i++) ekin += v[i] * v[i] * minv[i];
        ekin *= 0.5;
    }
    nstep = 0;
    epot = func(x, f, &nstep);
    if (gammai == 0.0) {
        for (i = 0; i < n; i++) accel[i] = -f[i] * minv[i] * dt5;
    } else {
        for (i = 0; i < n; i++) {
            if (S.m.frozen && S.m.frozen[i / dim]) accel[i] = 0.0;
            else {
                sqrmass[i] = 1.0 / sqrt(minv[i]);
                accel[i] = (-f[i] + sdfac * sqrmass[i] * nabc_md_gauss(seed, 0, i)) * minv[i] * dt5;
            }
        }
    }
    for (nstep = 1; nstep <= maxstep; nstep++) {
        tscal = ekin > 0.01 ? sqrt(1. + dttp * (ekin0 / ekin - 1.)) : 1.0;
        ekin = 0.0;
        for (i = 0; i < n; i++) {
            if (gammai == 0.0) v[i] = (v[i] + accel[i]) * tscal;
            else v[i] = c_explic * v[i] + accel[i];
            if (v[i] > o->vlimit) v[i] = o->vlimit;
            if (v[i] < -o->vlimit) v[i] = -o->vlimit;
            x[i] += v[i] * dtx;
        }
        epot = func(x, f, &nstep);
        if (gammai == 0.0) {
            for (i = 0; i < n; i++) {
                accel[i] = -f[i] * minv[i] * dt5;
                v[i] = (v[i] + accel[i]) * tscal;
            }
        } else {
            for (i = 0; i < n; i++) {
                if (!(S.m.frozen && S.m.frozen[i / dim])) {
                    accel[i] = (-f[i] + sdfac * sqrmass[i] * nabc_md_gauss(seed, nstep, i)) * minv[i] * dt5;
                    v[i] = (v[i] + accel[i]) * c_implic;
                }
            }
        }
        for (i = 0; i < n; i++) ekin += v[i] * v[i] / minv[i];
        ekin *= 0.5;
        if (gammai > 0.) ekin *= c_ave;
        temp = ekin / (o->boltz2 * rndf);
        t += o->dt;
        if (o->ntwx > 0 && nstep % o->ntwx == 0 && S.traj.fp && nabc_nctraj_write(&S.traj, x, t) != 0) {
            free(minv);
            free(accel);
            free(sqrmass);
            return -1;
        }
        if (nstep % o->ntpr_md == 0 || nstep == 1) {
            fprintf(out(), "md:       %5d %10.3f %10.2f %10.2f %10.2f %10.2f\n", nstep, t, ekin, epot, ekin + epot,
                    temp);
            fflush(out());
        }
    }
    nstep = -3;
    func(x, f, &nstep);
    nabc_opts.t = t;
    free(minv);
    free(accel);
    free(sqrmass);
    return 0;
}

/* -------------------------------------------------------------- conjgrad */

/*
 * Conjugate gradients with Powell restarts and a cubic-spline line search,
 * the algorithm of conjgrad.c:34-263 (return codes: >0 iterations,
 * -1 bad line search, -2 uphill direction, -3 maxiter, -4 no decrease).
 * The original is a web of gotos; the same steps are arranged here as a
 * loop over named states.
 */
enum cg_state { CG_EVAL, CG_RESET_DIR, CG_TEST, CG_NEW_SEARCH, CG_STEP, CG_SPLINE, CG_NEXT_STEP,
                CG_NEW_DIR, CG_FINISH, CG_DONE };

static int conjgrad_host(double x[], int *n, double *f, double (*func)(double *, double *, int *),
                         double *rmsgrad, double *dfpred, int *maxiter);

/* each evaluation of the device-resident minimiser: what mme() records and prints */
static void cg_device_report(int it, const double *ene14, void *user)
{
    (void)user;
    memcpy(S.ene, ene14, sizeof(S.ene));
    mme_report(it, S.ene);
}

/*
 * conjgrad over this library's own mme(): the whole minimiser runs with its
 * vectors on the device (gbhip_conjgrad), only scalars cross per
 * evaluation; the ff: lines are printed as mme() prints them.
 */
static int conjgrad_device(double x[], int n, double *f, double rmsgrad, double dfpred, int maxiter)
{
    const nabc_opts_t *o = &nabc_opts;
    int ret = 0;
    if (ensure_ctx(x) != 0) return -1;
    if (gbhip_conjgrad(S.ctx, x, n, f, rmsgrad, dfpred, maxiter, o->nsnb > 0 ? o->nsnb : 1, cg_device_report, NULL,
                       &ret) != 0) {
        nabc_seterr("%s", gbhip_last_error());
        return -1;
    }
    S.have_list = 1;
    return ret;
}

int conjgrad(double x[], int *n, double *f, double (*func)(double *, double *, int *), double *rmsgrad,
             double *dfpred, int *maxiter)
{
    if (func == mme && S.m.prm && *n == 3 * S.m.prm->Natom && !nabc_opts.gb_debug && nabc_opts.cgdevice)
        return conjgrad_device(x, *n, f, *rmsgrad, *dfpred, *maxiter);
    return conjgrad_host(x, n, f, func, rmsgrad, dfpred, maxiter);
}

static int conjgrad_host(double x[], int *n, double *f, double (*func)(double *, double *, int *),
                         double *rmsgrad, double *dfpred, int *maxiter)
{
    const int N = *n, maxlin = 10, mxfcon = 4;
    double *d = (double *)calloc(N, sizeof(double));      /* search direction (w[1..n]) */
    double *rsdx = (double *)calloc(N, sizeof(double));   /* restart direction */
    double *rsdg = (double *)calloc(N, sizeof(double));   /* restart gradient change */
    double *ginit_v = (double *)calloc(N, sizeof(double));/* gradient at line start */
    double *xopt = (double *)calloc(N, sizeof(double));
    double *gopt = (double *)calloc(N, sizeof(double));
    double *g = (double *)calloc(N, sizeof(double));
    double dgrad = (*rmsgrad) * (*rmsgrad) * N;
    double gama = 0, beta = 0, fmin = 0, gmin = 0, dfpr = 0, gnew = 0, step = 0, work, finit = 0, ginit = 0;
    double gsqrd = 0, gspln, stmin = 0, gamden = 0, ddspln = 0, stepch = 0, sbound = 0, fch = 0, sum;
    int niter = 0, nfbeg = 0, iterc = 0, iterfm = 0, iterrs = 0, iretry = 0, nfopt = 0, ier = 0, i;
    enum cg_state st = CG_EVAL;

    while (st != CG_DONE) {
        switch (st) {
        case CG_EVAL: { /* L10 */
            int it;
            ++niter;
            it = niter - 1;
            *f = func(x, g, &it);
            st = niter >= 2 ? CG_TEST : CG_RESET_DIR;
            break;
        }
        case CG_RESET_DIR: /* L20 */
            for (i = 0; i < N; i++) d[i] = -g[i];
            iterrs = 0;
            st = iterc > 0 ? CG_NEW_SEARCH : CG_TEST;
            break;
        case CG_TEST: /* L30 */
            gnew = 0.0;
            sum = 0.0;
            for (i = 0; i < N; i++) {
                gnew += d[i] * g[i];
                sum += g[i] * g[i];
            }
            if (niter == 1) {
                fmin = *f;
                gsqrd = sum;
                nfopt = niter;
                memcpy(xopt, x, sizeof(double) * N);
                memcpy(gopt, g, sizeof(double) * N);
                if (sum <= dgrad) { st = CG_DONE; break; }
            } else {
                fch = *f - fmin;
                if (fch < 0.0 || (fch == 0.0 && gnew / gmin >= -1.0)) {
                    fmin = *f;
                    gsqrd = sum;
                    nfopt = niter;
                    memcpy(xopt, x, sizeof(double) * N);
                    memcpy(gopt, g, sizeof(double) * N);
                    if (sum <= dgrad) { st = CG_DONE; break; }
                }
            }
            if (niter == *maxiter + 1) { ier = -3; st = CG_DONE; break; }
            if (niter > 1) { st = CG_SPLINE; break; }
            dfpr = *dfpred;
            stmin = *dfpred / gsqrd;
            st = CG_NEW_SEARCH;
            break;
        case CG_NEW_SEARCH: /* L40 */
            ++iterc;
            finit = *f;
            ginit = 0.0;
            for (i = 0; i < N; i++) {
                ginit_v[i] = g[i];
                ginit += d[i] * g[i];
            }
            if (ginit >= 0.0) { ier = -2; st = CG_FINISH; break; }
            gmin = ginit;
            sbound = -1.0;
            nfbeg = niter;
            iretry = -1;
            {
                double lim = fabs(dfpr / ginit);
                stepch = stmin <= lim ? stmin : lim;
            }
            stmin = 0.0;
            st = CG_STEP;
            break;
        case CG_STEP: /* L50 */
            step = stmin + stepch;
            work = 0.0;
            for (i = 0; i < N; i++) {
                x[i] = xopt[i] + stepch * d[i];
                if (fabs(x[i] - xopt[i]) > work) work = fabs(x[i] - xopt[i]);
            }
            if (work > 0.0) { st = CG_EVAL; break; }
            if (niter > nfbeg + 1) ier = -1;
            if (fabs(gmin / ginit) > 0.2) ier = -1;
            st = CG_FINISH;
            break;
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
        case CG_NEW_DIR: /* L80 */
            sum = 0.0;
            for (i = 0; i < N; i++) sum += g[i] * ginit_v[i];
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
            if (iterrs != 0 && iterc - iterrs < N && fabs(sum) < gsqrd * 0.2) {
                double s2 = 0.0;
                gama = 0.0;
                for (i = 0; i < N; i++) {
                    gama += g[i] * rsdg[i];
                    s2 += g[i] * rsdx[i];
                }
                gama /= gamden;
                if (fabs(beta * gmin + gama * s2) < gsqrd * 0.2) {
                    for (i = 0; i < N; i++) d[i] = -g[i] + beta * d[i] + gama * rsdx[i];
                    st = CG_NEW_SEARCH;
                    break;
                }
            }
            gamden = gmin - ginit;
            for (i = 0; i < N; i++) {
                rsdx[i] = d[i];
                rsdg[i] = g[i] - ginit_v[i];
                d[i] = -g[i] + beta * d[i];
            }
            iterrs = iterc;
            st = CG_NEW_SEARCH;
            break;
        case CG_FINISH: /* L90 */
            if (niter != nfopt) {
                *f = fmin;
                memcpy(x, xopt, sizeof(double) * N);
                memcpy(g, gopt, sizeof(double) * N);
            }
            st = ier == 0 ? CG_NEW_DIR : CG_DONE;
            break;
        default:
            st = CG_DONE;
        }
    }
    free(d); free(rsdx); free(rsdg); free(ginit_v); free(xopt); free(gopt); free(g);
    return ier == 0 ? niter : ier;
}
