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
    PARMSTRUCT_T *prm;
    int *frozen;
    int nfrozen;
    int gb;
    double gboffset, gbneckscale;
    double *alpha, *beta, *gamma;
    int *neckidx;
    gbhip_ctx *ctx;
    int have_list;
    long long nb_pairs;
    double ene[14];
    /* bonded tables (0-based atom indices, one entry per Fourier term) */
    int nb, na, nd, n14;
    int *bi, *bj, *ai, *aj, *ak, *di, *dj, *dk, *dl, *pi, *pj, *phb;
    double *brk, *breq, *atk, *ateq, *dpk, *dpn, *dph, *pscee, *pscnb, *pa, *pb;
} S;

static FILE *out(void) { return nabc_out ? nabc_out : stdout; }

static void free_tables(void)
{
    free(S.alpha); free(S.beta); free(S.gamma); free(S.neckidx);
    free(S.bi); free(S.bj); free(S.brk); free(S.breq);
    free(S.ai); free(S.aj); free(S.ak); free(S.atk); free(S.ateq);
    free(S.di); free(S.dj); free(S.dk); free(S.dl); free(S.dpk); free(S.dpn); free(S.dph);
    free(S.pi); free(S.pj); free(S.phb); free(S.pscee); free(S.pscnb); free(S.pa); free(S.pb);
    free(S.frozen);
    S.alpha = S.beta = S.gamma = NULL;
    S.neckidx = NULL;
    S.bi = S.bj = S.ai = S.aj = S.ak = S.di = S.dj = S.dk = S.dl = S.pi = S.pj = S.phb = NULL;
    S.brk = S.breq = S.atk = S.ateq = S.dpk = S.dpn = S.dph = NULL;
    S.pscee = S.pscnb = S.pa = S.pb = NULL;
    S.frozen = NULL;
}

static void release_ctx(void)
{
    if (S.ctx) gbhip_ctx_destroy(S.ctx);
    S.ctx = NULL;
    S.have_list = 0;
}

/* GB parameter set-up of mme_init_sff, sff.c:1187-1333 */
static int gb_setup(PARMSTRUCT_T *prm, int gb)
{
    int n = prm->Natom, i;
    S.alpha = (double *)calloc(n, sizeof(double));
    S.beta = (double *)calloc(n, sizeof(double));
    S.gamma = (double *)calloc(n, sizeof(double));
    S.neckidx = (int *)calloc(n, sizeof(int));
    S.gboffset = 0.09;
    S.gbneckscale = 0.0;
    for (i = 0; i < n; i++) {
        double a = 0, b = 0, g = 0;
        switch (gb) {
        case 2: a = 0.8; b = 0.0; g = 2.909125; break;
        case 5: a = 1.0; b = 0.8; g = 4.85; break;
        case 7:
            a = 1.09511284; b = 1.90792938; g = 2.50798245;
            switch (prm->AtomicNum[i]) {
            case 1: prm->Fs[i] = 1.09085413633; break;
            case 6: prm->Fs[i] = 0.484353823306; break;
            case 7: prm->Fs[i] = 0.700147318409; break;
            case 8: prm->Fs[i] = 1.06557401132; break;
            case 16: prm->Fs[i] = 0.602256336067; break;
            default: prm->Fs[i] = 0.5;
            }
            break;
        case 8:
            switch (prm->AtomicNum[i]) {
            case 1: a = 0.788440; b = 0.798699; g = 0.437334; prm->Fs[i] = 1.425952; break;
            case 6: a = 0.733756; b = 0.506378; g = 0.205844; prm->Fs[i] = 1.058554; break;
            case 7: a = 0.503364; b = 0.316828; g = 0.192915; prm->Fs[i] = 0.733599; break;
            case 8: a = 0.867814; b = 0.876635; g = 0.387882; prm->Fs[i] = 1.061039; break;
            case 15: a = 1.0; b = 0.8; g = 4.85; prm->Fs[i] = 0.5; break;
            case 16: a = 0.867814; b = 0.876635; g = 0.387882; prm->Fs[i] = -0.703469; break;
            default: a = 1.0; b = 0.8; g = 4.85; prm->Fs[i] = 0.5;
            }
            break;
        default: break;
        }
        S.alpha[i] = a;
        S.beta[i] = b;
        S.gamma[i] = g;
        if (gb == 7 || gb == 8) {
            S.neckidx[i] = (int)((prm->Rborn[i] - 1.0) * 20.0 + 0.5);
            if (S.neckidx[i] < 0 || S.neckidx[i] > 20) {
                nabc_seterr("Atom %d outside the allowed range of 1-2 Angstroms for igb==%d.", i, gb);
                return -1;
            }
        }
    }
    if (gb == 7) S.gbneckscale = 0.361825;
    if (gb == 8) {
        S.gboffset = 0.195141;
        S.gbneckscale = 0.826836;
    }
    return 0;
}

/* bonded tables for the device (eff.c:2617-2653 inputs) */
static void bonded_setup(PARMSTRUCT_T *p)
{
    int k, m, pass;
    S.nb = p->Nbonh + p->Mbona;
    S.bi = (int *)malloc(sizeof(int) * (S.nb + 1));
    S.bj = (int *)malloc(sizeof(int) * (S.nb + 1));
    S.brk = (double *)malloc(sizeof(double) * (S.nb + 1));
    S.breq = (double *)malloc(sizeof(double) * (S.nb + 1));
    for (k = 0, m = 0; k < p->Nbonh; k++, m++) {
        S.bi[m] = p->BondHAt1[k] / 3; S.bj[m] = p->BondHAt2[k] / 3;
        S.brk[m] = p->Rk[p->BondHNum[k] - 1]; S.breq[m] = p->Req[p->BondHNum[k] - 1];
    }
    for (k = 0; k < p->Mbona; k++, m++) {
        S.bi[m] = p->BondAt1[k] / 3; S.bj[m] = p->BondAt2[k] / 3;
        S.brk[m] = p->Rk[p->BondNum[k] - 1]; S.breq[m] = p->Req[p->BondNum[k] - 1];
    }
    S.na = p->Ntheth + p->Ntheta;
    S.ai = (int *)malloc(sizeof(int) * (S.na + 1));
    S.aj = (int *)malloc(sizeof(int) * (S.na + 1));
    S.ak = (int *)malloc(sizeof(int) * (S.na + 1));
    S.atk = (double *)malloc(sizeof(double) * (S.na + 1));
    S.ateq = (double *)malloc(sizeof(double) * (S.na + 1));
    for (k = 0, m = 0; k < p->Ntheth; k++, m++) {
        S.ai[m] = p->AngleHAt1[k] / 3; S.aj[m] = p->AngleHAt2[k] / 3; S.ak[m] = p->AngleHAt3[k] / 3;
        S.atk[m] = p->Tk[p->AngleHNum[k] - 1]; S.ateq[m] = p->Teq[p->AngleHNum[k] - 1];
    }
    for (k = 0; k < p->Ntheta; k++, m++) {
        S.ai[m] = p->AngleAt1[k] / 3; S.aj[m] = p->AngleAt2[k] / 3; S.ak[m] = p->AngleAt3[k] / 3;
        S.atk[m] = p->Tk[p->AngleNum[k] - 1]; S.ateq[m] = p->Teq[p->AngleNum[k] - 1];
    }
    /* dihedrals: one entry per term of a Pn < 0 series (eff.c:619-623) */
    {
        int cap = 2 * (p->Nphih + p->Mphia) + 8;
        S.di = (int *)malloc(sizeof(int) * cap); S.dj = (int *)malloc(sizeof(int) * cap);
        S.dk = (int *)malloc(sizeof(int) * cap); S.dl = (int *)malloc(sizeof(int) * cap);
        S.dpk = (double *)malloc(sizeof(double) * cap); S.dpn = (double *)malloc(sizeof(double) * cap);
        S.dph = (double *)malloc(sizeof(double) * cap);
        m = 0;
        for (pass = 0; pass < 2; pass++) {
            int nd = pass == 0 ? p->Nphih : p->Mphia;
            int *a1 = pass == 0 ? p->DihHAt1 : p->DihAt1, *a2 = pass == 0 ? p->DihHAt2 : p->DihAt2;
            int *a3 = pass == 0 ? p->DihHAt3 : p->DihAt3, *a4 = pass == 0 ? p->DihHAt4 : p->DihAt4;
            int *tp = pass == 0 ? p->DihHNum : p->DihNum;
            for (k = 0; k < nd; k++) {
                int ty = tp[k] - 1;
                for (;;) {
                    if (m == cap) {
                        cap *= 2;
                        S.di = (int *)realloc(S.di, sizeof(int) * cap); S.dj = (int *)realloc(S.dj, sizeof(int) * cap);
                        S.dk = (int *)realloc(S.dk, sizeof(int) * cap); S.dl = (int *)realloc(S.dl, sizeof(int) * cap);
                        S.dpk = (double *)realloc(S.dpk, sizeof(double) * cap);
                        S.dpn = (double *)realloc(S.dpn, sizeof(double) * cap);
                        S.dph = (double *)realloc(S.dph, sizeof(double) * cap);
                    }
                    S.di[m] = a1[k] / 3; S.dj[m] = a2[k] / 3;
                    S.dk[m] = abs(a3[k]) / 3; S.dl[m] = abs(a4[k]) / 3;
                    S.dpk[m] = p->Pk[ty]; S.dpn[m] = p->Pn[ty]; S.dph[m] = p->Phase[ty];
                    m++;
                    if (p->Pn[ty] < 0.0 && ty + 1 < p->Nptra) ty++;
                    else break;
                }
            }
        }
        S.nd = m;
    }
    /* 1-4 pairs (prm.c:1854-1889 list; LJ selection as eff.c:1209-1234) */
    {
        int i, e = 0;
        S.n14 = 0;
        for (i = 0; i < p->Natom - 1; i++) S.n14 += p->N14pairs[i];
        S.pi = (int *)malloc(sizeof(int) * (S.n14 + 1)); S.pj = (int *)malloc(sizeof(int) * (S.n14 + 1));
        S.phb = (int *)malloc(sizeof(int) * (S.n14 + 1));
        S.pscee = (double *)malloc(sizeof(double) * (S.n14 + 1));
        S.pscnb = (double *)malloc(sizeof(double) * (S.n14 + 1));
        S.pa = (double *)malloc(sizeof(double) * (S.n14 + 1));
        S.pb = (double *)malloc(sizeof(double) * (S.n14 + 1));
        m = 0;
        for (i = 0; i < p->Natom - 1; i++) {
            for (k = 0; k < p->N14pairs[i]; k++, e++, m++) {
                int j = p->N14pairlist[e];
                int ic = p->Cno[p->Ntypes * (p->Iac[i] - 1) + p->Iac[j] - 1];
                double fac = p->N14scnblist[e];
                S.pi[m] = i; S.pj[m] = j;
                S.pscee[m] = p->N14sceelist[e];
                S.pscnb[m] = fac;
                if (ic > 0 || fac != 1.0) {
                    if (ic > 0) ic--;
                    else {
                        int big = p->Iac[i] > p->Iac[j] ? p->Iac[i] : p->Iac[j];
                        int sml = p->Iac[i] > p->Iac[j] ? p->Iac[j] : p->Iac[i];
                        ic = big * (big - 1) / 2 + sml - 1;
                    }
                    S.phb[m] = 0;
                    S.pa[m] = p->Cn1[ic];
                    S.pb[m] = p->Cn2[ic];
                } else {
                    int hc = -ic - 1;
                    S.phb[m] = 1;
                    S.pa[m] = p->HB12 ? p->HB12[hc] : 0.0;
                    S.pb[m] = p->HB10 ? p->HB10[hc] : 0.0;
                }
            }
        }
    }
}

int mme_init_sff(PARMSTRUCT_T *prm, int *frozen, int *constrained, double *x0, char *ncname)
{
    const nabc_opts_t *o = &nabc_opts;
    int i, n;
    (void)x0;
    release_ctx();
    free_tables();
    S.prm = NULL;
    if (!prm) {
        nabc_seterr("mme_init_sff: NULL topology");
        return -1;
    }
    if (prm->IfBox && o->gb) {
        nabc_seterr("Error: gb>0 is incompatible with periodic boundary conditions.");
        return -1;
    }
    if (o->gb != 1 && o->gb != 2 && o->gb != 5 && o->gb != 7 && o->gb != 8) {
        nabc_seterr("mme_init_sff: gb=%d is not provided by this library (GB models 1, 2, 5, 7, 8 only)", o->gb);
        return -1;
    }
    if (o->gbsa || o->hcp || o->ips || o->dim != 3 || o->irattle || o->nscm || o->wdssp != 0.0) {
        nabc_seterr("mme_init_sff: gbsa, hcp, ips, dim=4, rattle, nscm and wdssp are not provided by this library");
        return -1;
    }
    if (ncname) {
        nabc_seterr("mme_init_sff: netCDF trajectories are not provided by this library");
        return -1;
    }
    if (o->dield != 1 && o->dield != 0) {
        nabc_seterr("mme_init_sff: only diel=C and diel=R are provided");
        return -1;
    }
    n = prm->Natom;
    S.prm = prm;
    S.gb = o->gb;
    S.frozen = (int *)calloc(n, sizeof(int));
    S.nfrozen = 0;
    for (i = 0; i < n; i++) {
        S.frozen[i] = frozen ? frozen[i] != 0 : 0;
        S.nfrozen += S.frozen[i];
    }
    (void)constrained;
    if (o->dielc != 1.0) {
        double s = 1.0 / sqrt(o->dielc);
        for (i = 0; i < n; i++) prm->Charges[i] *= s;
    }
    if (gb_setup(prm, o->gb) != 0) return -1;
    bonded_setup(prm);
    return 0;
}

static int ensure_ctx(const double *x)
{
    PARMSTRUCT_T *p = S.prm;
    const nabc_opts_t *o = &nabc_opts;
    gbhip_topology t;
    gbhip_options op;
    gbhip_bonded b;
    if (S.ctx) return 0;
    memset(&t, 0, sizeof(t));
    t.natom = p->Natom;
    t.ntypes = p->Ntypes;
    t.charge = p->Charges;
    t.rborn = p->Rborn;
    t.fs = p->Fs;
    t.iac = p->Iac;
    t.cno = p->Cno;
    t.cn1 = p->Cn1;
    t.cn2 = p->Cn2;
    t.iblo = p->Iblo;
    t.exclat = p->ExclAt;
    t.frozen = S.frozen;
    t.gbalpha = S.alpha;
    t.gbbeta = S.beta;
    t.gbgamma = S.gamma;
    t.neckidx = S.neckidx;
    memset(&op, 0, sizeof(op));
    op.gb = S.gb;
    op.precision = o->gbprec ? GBHIP_MIXED : GBHIP_FP64;
    op.cut = o->cut;
    op.rgbmax = o->rgbmax;
    t.fsmax = p->Fsmax;
    op.gboffset = S.gboffset;
    op.gbneckscale = S.gbneckscale;
    op.kappa = o->kappa;
    op.epsext = o->epsext;
    if (gbhip_ctx_create(&t, &op, x, &S.ctx) != 0) {
        nabc_seterr("%s", gbhip_last_error());
        return -1;
    }
    memset(&b, 0, sizeof(b));
    b.nbond = S.nb; b.bond_i = S.bi; b.bond_j = S.bj; b.bond_rk = S.brk; b.bond_req = S.breq;
    b.nangle = S.na; b.angle_i = S.ai; b.angle_j = S.aj; b.angle_k = S.ak; b.angle_tk = S.atk; b.angle_teq = S.ateq;
    b.ndihed = S.nd; b.dihed_i = S.di; b.dihed_j = S.dj; b.dihed_k = S.dk; b.dihed_l = S.dl;
    b.dihed_pk = S.dpk; b.dihed_pn = S.dpn; b.dihed_phase = S.dph;
    b.npair14 = S.n14; b.p14_i = S.pi; b.p14_j = S.pj; b.p14_hb = S.phb;
    b.p14_scee = S.pscee; b.p14_scnb = S.pscnb; b.p14_a = S.pa; b.p14_b = S.pb;
    b.charge = p->Charges;
    b.dield = o->dield;
    if (gbhip_set_bonded(S.ctx, &b) != 0) {
        nabc_seterr("%s", gbhip_last_error());
        release_ctx();
        return -1;
    }
    return 0;
}

void *mme_context(double *x)
{
    if (!S.prm) {
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

/* mme() / mme34(), sff.c:1854, eff.c:2442-2947 */
double mme(double *x, double *f, int *iter)
{
    const nabc_opts_t *o = &nabc_opts;
    int it = *iter;
    if (it == -3) {
        release_ctx();
        return 0.0;
    }
    if (!S.prm) {
        nabc_seterr("mme: mme_init_sff has not been called");
        return NAN;
    }
    if (ensure_ctx(x) != 0) return NAN;
    if (it == 0) fprintf(out(), "      iter    Total       bad      vdW     elect   nonpolar   genBorn      frms\n");
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
    if (o->gb_debug) {
        int i, n = S.prm->Natom;
        double *reff = (double *)malloc(sizeof(double) * n);
        if (gbhip_get_reff(S.ctx, reff) == 0) {
            fprintf(out(), "Effective Born radii:\n");
            for (i = 0; i < n; i++) fprintf(out(), "%d\t%15.7f\t%15.7f\n", i + 1, S.prm->Rborn[i], reff[i]);
        }
        free(reff);
    }
    if (it > -1 && (it == 0 || it % o->ntpr == 0)) {
        const double *e = S.ene;
        fprintf(out(), "ff:%6d %9.2f %9.2f %9.2f %9.2f %9.2f %9.2f %9.2e\n", it, e[0], e[3] + e[4] + e[5],
                e[1] + e[7], e[2] + e[8], e[9] + e[11] + e[12], e[10], e[13]);
        fflush(out());
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
    if (gbhip_md_init(S.ctx, S.prm->Masses, x, v, &mo, cap) != 0 || gbhip_md_run(S.ctx, maxstep, 1) != 0 ||
        gbhip_md_state(S.ctx, x, v, f) != 0) {
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

/* md(), sff.c:2071-2366 (no rattle, no COM removal, no trajectory output) */
int md(int n, int maxstep, double *x, double *f, double *v, double (*func)(double *, double *, int *))
{
    const nabc_opts_t *o = &nabc_opts;
    double *minv, *accel, *sqrmass;
    double dtx = o->dt * 20.455, dt5 = 0.5 * dtx, rndf, ekin0, dttp, gammai, c_implic, c_explic, c_ave, sdfac;
    double ekin = 0.0, epot, tscal, temp, t = o->t;
    int i, nstep, dim = 3;
    unsigned long long seed = o->md_seed;
    if (!S.prm && func == mme) {
        nabc_seterr("md: mme_init_sff has not been called");
        return -1;
    }
    if (func == mme && n == 3 * S.prm->Natom) return md_device(n, maxstep, x, f, v);
    minv = (double *)calloc(n, sizeof(double));
    accel = (double *)calloc(n, sizeof(double));
    sqrmass = (double *)calloc(n, sizeof(double));
    rndf = n - 3 * S.nfrozen;
    ekin0 = o->boltz2 * rndf * o->temp0;
    dttp = o->dt / o->tautp;
    gammai = o->gamma_ln / 20.455;
    c_implic = 1.0 / (1.0 + gammai * dt5);
    c_explic = 1.0 - gammai * dt5;
    c_ave = 1.0 + gammai * dt5;
    sdfac = sqrt(4.0 * gammai * o->boltz2 * o->temp0 / dtx);
    if (S.prm) {
        for (i = 0; i < n; i++) minv[i] = 1.0 / S.prm->Masses[i / dim];
    } else {
        for (i = 0; i < n; i++) minv[i] = 1.0 / o->genmass;
    }
    if (o->zerov) {
        for (i = 0; i < n; i++) v[i] = 0.0;
        ekin = 0.0;
    } else if (o->tempi > 0.0) {
        for (i = 0; i < n; i++) {
            if (S.frozen && S.frozen[i / dim]) v[i] = 0.0;
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
            if (S.frozen && S.frozen[i / dim]) accel[i] = 0.0;
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
                if (!(S.frozen && S.frozen[i / dim])) {
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

int conjgrad(double x[], int *n, double *f, double (*func)(double *, double *, int *), double *rmsgrad,
             double *dfpred, int *maxiter)
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
