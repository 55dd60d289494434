/*
 * model.c -- the per-system state mme_init_sff() derives from a topology:
 * GB parameters (sff.c:1187-1333, gb 8 element constants gbneck.h:150-170),
 * dielc charge scaling (sff.c:1153-1162) and the bonded / 1-4 term tables
 * mme34 evaluates (eff.c:2617-2653), plus the GPU context built from them.
 * mme_init_sff() keeps one such model for the libsff-style global API; the
 * serving loop builds one per job.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gbhip.h"
#include "host_internal.h"

void nabc_model_free(nabc_model *m)
{
    free(m->alpha); free(m->beta); free(m->gamma); free(m->neckidx);
    free(m->bi); free(m->bj); free(m->brk); free(m->breq);
    free(m->ai); free(m->aj); free(m->ak); free(m->atk); free(m->ateq);
    free(m->di); free(m->dj); free(m->dk); free(m->dl); free(m->dpk); free(m->dpn); free(m->dph);
    free(m->pi); free(m->pj); free(m->phb); free(m->pscee); free(m->pscnb); free(m->pa); free(m->pb);
    free(m->frozen);
    memset(m, 0, sizeof(*m));
}

/* GB parameter set-up of mme_init_sff, sff.c:1187-1333 (modifies prm->Fs for gb 7/8) */
static int gb_setup(nabc_model *m, PARMSTRUCT_T *prm, int gb)
{
    int n = prm->Natom, i;
    m->alpha = (double *)calloc(n, sizeof(double));
    m->beta = (double *)calloc(n, sizeof(double));
    m->gamma = (double *)calloc(n, sizeof(double));
    m->neckidx = (int *)calloc(n, sizeof(int));
    m->gboffset = 0.09;
    m->gbneckscale = 0.0;
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
        m->alpha[i] = a;
        m->beta[i] = b;
        m->gamma[i] = g;
        if (gb == 7 || gb == 8) {
            m->neckidx[i] = (int)((prm->Rborn[i] - 1.0) * 20.0 + 0.5);
            if (m->neckidx[i] < 0 || m->neckidx[i] > 20) {
                nabc_seterr("Atom %d outside the allowed range of 1-2 Angstroms for igb==%d.", i, gb);
                return -1;
            }
        }
    }
    if (gb == 7) m->gbneckscale = 0.361825;
    if (gb == 8) {
        m->gboffset = 0.195141;
        m->gbneckscale = 0.826836;
    }
    return 0;
}

/* bonded tables for the device (eff.c:2617-2653 inputs) */
static void bonded_setup(nabc_model *m, PARMSTRUCT_T *p)
{
    int k, e, pass;
    m->nb = p->Nbonh + p->Mbona;
    m->bi = (int *)malloc(sizeof(int) * (m->nb + 1));
    m->bj = (int *)malloc(sizeof(int) * (m->nb + 1));
    m->brk = (double *)malloc(sizeof(double) * (m->nb + 1));
    m->breq = (double *)malloc(sizeof(double) * (m->nb + 1));
    e = 0;
    for (k = 0; k < p->Nbonh; k++, e++) {
        m->bi[e] = p->BondHAt1[k] / 3; m->bj[e] = p->BondHAt2[k] / 3;
        m->brk[e] = p->Rk[p->BondHNum[k] - 1]; m->breq[e] = p->Req[p->BondHNum[k] - 1];
    }
    for (k = 0; k < p->Mbona; k++, e++) {
        m->bi[e] = p->BondAt1[k] / 3; m->bj[e] = p->BondAt2[k] / 3;
        m->brk[e] = p->Rk[p->BondNum[k] - 1]; m->breq[e] = p->Req[p->BondNum[k] - 1];
    }
    m->na = p->Ntheth + p->Ntheta;
    m->ai = (int *)malloc(sizeof(int) * (m->na + 1));
    m->aj = (int *)malloc(sizeof(int) * (m->na + 1));
    m->ak = (int *)malloc(sizeof(int) * (m->na + 1));
    m->atk = (double *)malloc(sizeof(double) * (m->na + 1));
    m->ateq = (double *)malloc(sizeof(double) * (m->na + 1));
    e = 0;
    for (k = 0; k < p->Ntheth; k++, e++) {
        m->ai[e] = p->AngleHAt1[k] / 3; m->aj[e] = p->AngleHAt2[k] / 3; m->ak[e] = p->AngleHAt3[k] / 3;
        m->atk[e] = p->Tk[p->AngleHNum[k] - 1]; m->ateq[e] = p->Teq[p->AngleHNum[k] - 1];
    }
    for (k = 0; k < p->Ntheta; k++, e++) {
        m->ai[e] = p->AngleAt1[k] / 3; m->aj[e] = p->AngleAt2[k] / 3; m->ak[e] = p->AngleAt3[k] / 3;
        m->atk[e] = p->Tk[p->AngleNum[k] - 1]; m->ateq[e] = p->Teq[p->AngleNum[k] - 1];
    }
    /* dihedrals: one entry per term of a Pn < 0 series (eff.c:619-623) */
    {
        int cap = 2 * (p->Nphih + p->Mphia) + 8;
        m->di = (int *)malloc(sizeof(int) * cap); m->dj = (int *)malloc(sizeof(int) * cap);
        m->dk = (int *)malloc(sizeof(int) * cap); m->dl = (int *)malloc(sizeof(int) * cap);
        m->dpk = (double *)malloc(sizeof(double) * cap); m->dpn = (double *)malloc(sizeof(double) * cap);
        m->dph = (double *)malloc(sizeof(double) * cap);
        e = 0;
        for (pass = 0; pass < 2; pass++) {
            int nd = pass == 0 ? p->Nphih : p->Mphia;
            int *a1 = pass == 0 ? p->DihHAt1 : p->DihAt1, *a2 = pass == 0 ? p->DihHAt2 : p->DihAt2;
            int *a3 = pass == 0 ? p->DihHAt3 : p->DihAt3, *a4 = pass == 0 ? p->DihHAt4 : p->DihAt4;
            int *tp = pass == 0 ? p->DihHNum : p->DihNum;
            for (k = 0; k < nd; k++) {
                int ty = tp[k] - 1;
                for (;;) {
                    if (e == cap) {
                        cap *= 2;
                        m->di = (int *)realloc(m->di, sizeof(int) * cap); m->dj = (int *)realloc(m->dj, sizeof(int) * cap);
                        m->dk = (int *)realloc(m->dk, sizeof(int) * cap); m->dl = (int *)realloc(m->dl, sizeof(int) * cap);
                        m->dpk = (double *)realloc(m->dpk, sizeof(double) * cap);
                        m->dpn = (double *)realloc(m->dpn, sizeof(double) * cap);
                        m->dph = (double *)realloc(m->dph, sizeof(double) * cap);
                    }
                    m->di[e] = a1[k] / 3; m->dj[e] = a2[k] / 3;
                    m->dk[e] = abs(a3[k]) / 3; m->dl[e] = abs(a4[k]) / 3;
                    m->dpk[e] = p->Pk[ty]; m->dpn[e] = p->Pn[ty]; m->dph[e] = p->Phase[ty];
                    e++;
                    if (p->Pn[ty] < 0.0 && ty + 1 < p->Nptra) ty++;
                    else break;
                }
            }
        }
        m->nd = e;
    }
    /* 1-4 pairs (prm.c:1854-1889 list; LJ selection as eff.c:1209-1234) */
    {
        int i, f = 0;
        m->n14 = 0;
        for (i = 0; i < p->Natom - 1; i++) m->n14 += p->N14pairs[i];
        m->pi = (int *)malloc(sizeof(int) * (m->n14 + 1)); m->pj = (int *)malloc(sizeof(int) * (m->n14 + 1));
        m->phb = (int *)malloc(sizeof(int) * (m->n14 + 1));
        m->pscee = (double *)malloc(sizeof(double) * (m->n14 + 1));
        m->pscnb = (double *)malloc(sizeof(double) * (m->n14 + 1));
        m->pa = (double *)malloc(sizeof(double) * (m->n14 + 1));
        m->pb = (double *)malloc(sizeof(double) * (m->n14 + 1));
        e = 0;
        for (i = 0; i < p->Natom - 1; i++) {
            for (k = 0; k < p->N14pairs[i]; k++, f++, e++) {
                int j = p->N14pairlist[f];
                int ic = p->Cno[p->Ntypes * (p->Iac[i] - 1) + p->Iac[j] - 1];
                double fac = p->N14scnblist[f];
                m->pi[e] = i; m->pj[e] = j;
                m->pscee[e] = p->N14sceelist[f];
                m->pscnb[e] = fac;
                if (ic > 0 || fac != 1.0) {
                    if (ic > 0) ic--;
                    else {
                        int big = p->Iac[i] > p->Iac[j] ? p->Iac[i] : p->Iac[j];
                        int sml = p->Iac[i] > p->Iac[j] ? p->Iac[j] : p->Iac[i];
                        ic = big * (big - 1) / 2 + sml - 1;
                    }
                    m->phb[e] = 0;
                    m->pa[e] = p->Cn1[ic];
                    m->pb[e] = p->Cn2[ic];
                } else {
                    int hc = -ic - 1;
                    m->phb[e] = 1;
                    m->pa[e] = p->HB12 ? p->HB12[hc] : 0.0;
                    m->pb[e] = p->HB10 ? p->HB10[hc] : 0.0;
                }
            }
        }
    }
}

int nabc_model_build(nabc_model *m, PARMSTRUCT_T *prm, const int *frozen, const nabc_opts_t *o)
{
    int i, n;
    memset(m, 0, sizeof(*m));
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
    if (o->gbsa || o->ips || o->dim != 3 || o->irattle || o->nscm || o->wdssp != 0.0) {
        nabc_seterr("mme_init_sff: gbsa, ips, dim=4, rattle, nscm and wdssp are not provided by this library");
        return -1;
    }
    if (o->dield != 1 && o->dield != 0) {
        nabc_seterr("mme_init_sff: only diel=C and diel=R are provided");
        return -1;
    }
    n = prm->Natom;
    m->prm = prm;
    m->gb = o->gb;
    m->dield = o->dield;
    m->frozen = (int *)calloc(n, sizeof(int));
    for (i = 0; i < n; i++) {
        m->frozen[i] = frozen ? frozen[i] != 0 : 0;
        m->nfrozen += m->frozen[i];
    }
    if (o->dielc != 1.0) {
        double s = 1.0 / sqrt(o->dielc);
        for (i = 0; i < n; i++) prm->Charges[i] *= s;
    }
    if (gb_setup(m, prm, o->gb) != 0) {
        nabc_model_free(m);
        return -1;
    }
    bonded_setup(m, prm);
    return 0;
}

int nabc_model_context(const nabc_model *m, const double *x, const nabc_opts_t *o, void *stream,
                       gbhip_ctx **out)
{
    return nabc_model_context_in(m, x, o, stream, NULL, out);
}

int nabc_model_context_in(const nabc_model *m, const double *x, const nabc_opts_t *o, void *stream,
                          gbhip_slot *slot, gbhip_ctx **out)
{
    const PARMSTRUCT_T *p = m->prm;
    gbhip_topology t;
    gbhip_options op;
    gbhip_bonded b;
    gbhip_ctx *ctx = NULL;
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
    t.frozen = m->frozen;
    t.gbalpha = m->alpha;
    t.gbbeta = m->beta;
    t.gbgamma = m->gamma;
    t.neckidx = m->neckidx;
    t.fsmax = p->Fsmax;
    memset(&op, 0, sizeof(op));
    op.gb = m->gb;
    op.precision = o->gbprec ? GBHIP_MIXED : GBHIP_FP64;
    op.cut = o->cut;
    op.rgbmax = o->rgbmax;
    op.gboffset = m->gboffset;
    op.gbneckscale = m->gbneckscale;
    op.kappa = o->kappa;
    op.epsext = o->epsext;
    if ((slot ? gbhip_ctx_create_in_slot(&t, &op, x, slot, &ctx) : gbhip_ctx_create_on(&t, &op, x, stream, &ctx)) !=
        0) {
        nabc_seterr("%s", gbhip_last_error());
        return -1;
    }
    memset(&b, 0, sizeof(b));
    b.nbond = m->nb; b.bond_i = m->bi; b.bond_j = m->bj; b.bond_rk = m->brk; b.bond_req = m->breq;
    b.nangle = m->na; b.angle_i = m->ai; b.angle_j = m->aj; b.angle_k = m->ak; b.angle_tk = m->atk;
    b.angle_teq = m->ateq;
    b.ndihed = m->nd; b.dihed_i = m->di; b.dihed_j = m->dj; b.dihed_k = m->dk; b.dihed_l = m->dl;
    b.dihed_pk = m->dpk; b.dihed_pn = m->dpn; b.dihed_phase = m->dph;
    b.npair14 = m->n14; b.p14_i = m->pi; b.p14_j = m->pj; b.p14_hb = m->phb;
    b.p14_scee = m->pscee; b.p14_scnb = m->pscnb; b.p14_a = m->pa; b.p14_b = m->pb;
    b.charge = p->Charges;
    b.dield = m->dield;
    if (gbhip_set_bonded(ctx, &b) != 0) {
        nabc_seterr("%s", gbhip_last_error());
        gbhip_ctx_destroy(ctx);
        return -1;
    }
    if (o->hcp) {
        /* egb_hcp over the prmtop's residues, strands and complexes
         * (eff.c:2687-2694; prm.c:1743-1790) */
        gbhip_hcp h;
        memset(&h, 0, sizeof(h));
        h.hcp = o->hcp;
        h.h1 = o->hcp_h1;
        h.h2 = o->hcp_h2;
        h.h3 = o->hcp_h3;
        h.dhcp = o->dhcp;
        h.nres = p->Nres;
        h.ipres = p->Ipres;
        h.nstrand = p->Nstrand;
        h.ipstrand = p->Ipstrand;
        h.ncomplex = p->Ncomplex;
        h.ipcomplex = p->Ipcomplex;
        if (gbhip_set_hcp(ctx, &h) != 0) {
            nabc_seterr("%s", gbhip_last_error());
            gbhip_ctx_destroy(ctx);
            return -1;
        }
    }
    *out = ctx;
    return 0;
}
