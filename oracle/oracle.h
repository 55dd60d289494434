/*
 * oracle.h -- CPU FP64 restatement of the reference's GB energy path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for nabc_amd's GPU
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product never links or calls it.
 *
 * It follows, function by function, src/sff of dacase/nabc:
 *   orc_create   GB parameter set-up of mme_init_sff (sff.c:1153-1162,
 *                :1187-1333, gbneck.h:150-170)
 *   orc_nblist   pair-list semantics of nblist()/searchkdtree()
 *                (nblist.c:748-765, :1146-1147, :1189-1196) by brute force
 *   orc_egb      egb() passes 1-3 (eff.c:1363-2417)
 *   orc_mme      mme34() assembly (eff.c:2442-2947) with ebond/eangl/ephi
 *                (eff.c:82-627) and the 1-4 nbond() (eff.c:896-1331)
 *   orc_md       md() (sff.c:2071-2366) over orc_mme
 *   orc_set_hcp  egb_hcp() and its set-up (hcp_gb.c, hcp.c; oracle_hcp.c)
 */
#ifndef NABC_AMD_ORACLE_H
#define NABC_AMD_ORACLE_H

#include "sff_gb.h"

typedef struct orc orc_t;

/* dield: 1 = constant dielectric (diel=C), 0 = distance-dependent (diel=R) */
orc_t *orc_create(PARMSTRUCT_T *prm, const int *frozen, int gb, double cut,
                  double rgbmax, double kappa, double epsext, double dielc,
                  int dield);
void orc_destroy(orc_t *o);

/* GB parameters derived at creation (for comparison with the product). */
void orc_gbparams(orc_t *o, double *gboffset, double *neckscale, double *fs,
                  double *alpha, double *beta, double *gamma, int *neckidx,
                  double *charges);

/* Build the list from x; returns the number of entries (2 per pair). */
long long orc_nblist(orc_t *o, const double *x);
/* Per-atom lower/upper counts and the flattened lists (sizes from counts). */
void orc_list_counts(orc_t *o, int *lpears, int *upears);
void orc_list_copy(orc_t *o, int *flat);

/* egb() on the current list; f (3N) accumulated into; reff (N) optional. */
double orc_egb(orc_t *o, const double *x, double *f, double *enb, double *eel,
               double *reff, double *sumdeijda);

/*
 * mme34(): builds the list when none exists or when rebuild != 0, zeroes f,
 * adds bonded, 1-4 and GB terms, zeroes frozen gradients; fills ene[13]
 * (mme34's layout) and frms; returns the total.
 */
double orc_mme(orc_t *o, const double *x, double *f, int rebuild, double *ene,
               double *frms);

/*
 * mme-compatible callback over a current oracle, with the reference's
 * rebuild cadence (first call, iter % nsnb == 0, iter == -4); every call's
 * ene[0..12] and frms are appended to a log.
 */
void orc_set_current(orc_t *o, int nsnb);
double orc_mme_cb(double *x, double *f, int *iter);
int orc_log_count(void);
void orc_log_get(int k, int *iter, double *ene13, double *frms);
void orc_log_clear(void);

/*
 * md(), sff.c:2071-2366, over orc_mme (rebuild cadence of mme34: first
 * call, then step % nsnb == 0).  Langevin noise from orc_gauss, the
 * counter-based stream the product uses in place of rand2.c.  x, v (3N)
 * are updated; rows (16 doubles per step, may be NULL) receive ene[0..12],
 * frms, ekin and the temperature of each step.
 */
typedef struct {
    double dt, temp0, tempi, gamma_ln, tautp, vlimit, boltz2;
    int nsnb, zerov;
    unsigned long long seed;
} orc_md_opts;
double orc_gauss(unsigned long long seed, long long step, long long idx);
int orc_md(orc_t *o, int maxstep, double *x, double *v, const double *masses,
           const orc_md_opts *mo, double *rows);

/* threads used by the OpenMP loops (0 = library default) */
void orc_set_threads(int n);

/*
 * Hierarchical charge partitioning (egb_hcp, hcp_gb.c:2166-2240, with the
 * set-up of sff.c:1335-1340 and :1577-1692).  hcp as mm_options takes it:
 * 1 = one charge per component, 2 = two (negative and positive atoms
 * apart), 4 = two-charge HCPO (optimal point charges, hcp.c:150-840).
 * h1, h2, h3: residue, strand and complex threshold distances; dhcp the
 * HCPO neutral-component factor.  Residues are atoms ipres[r] ..
 * ipres[r+1]-1 (0-based, nres+1 entries, as rdparm leaves prm->Ipres);
 * strands and complexes start at the 1-based ipstrand[s] / ipcomplex[c]
 * (prm.c:1743-1790).  Once set, orc_egb evaluates egb_hcp and orc_mme
 * builds no pair list (eff.c:2560-2561, :2687-2694).  h == NULL or
 * h->hcp == 0 switches back.  Returns 0, or -1 on a bad hierarchy.
 */
typedef struct {
    int hcp;
    double h1, h2, h3, dhcp;
    int nres, nstrand, ncomplex;
    const int *ipres, *ipstrand, *ipcomplex;
} orc_hcp_opts;
int orc_set_hcp(orc_t *o, const orc_hcp_opts *h);
/*
 * Component data of the last HCP evaluation, level 1 (residues), 2
 * (strands) or 3 (complexes); nq = 1 or 2 charges per component.  Any
 * pointer may be NULL: centres[3 n], charges[4 n nq] (x, y, z, q),
 * radii[n nq] (egb_hcp_rbondi), reff[n nq] (component Born radii).
 * Returns the component count, or -1.
 */
int orc_hcp_components(orc_t *o, int level, double *centres, double *charges, double *radii,
                       double *reff);

#endif
