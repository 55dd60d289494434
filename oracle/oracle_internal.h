/*
 * oracle_internal.h -- the oracle's state, shared by its C files.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#ifndef NABC_AMD_ORACLE_INTERNAL_H
#define NABC_AMD_ORACLE_INTERNAL_H

#include "oracle.h"

struct orc_hcp;

struct orc {
    PARMSTRUCT_T *prm;
    int n, gb, dield;
    double cut, rgbmax, kappa, epsext, gboffset, neckscale, fsmax;
    double *fs, *q, *alpha, *beta, *gamma;
    int *neck, *frozen;
    /* pair list: per atom, lower (j<i, ascending) then upper (j>i, ascending) */
    int *lp, *up, **list;
    int have_list;
    /* 1-2/1-3/1-4 exclusion lists per atom (IexclAt, sff.c:1551-1574) */
    int **iexcl;
    /* 1-4 list per atom (sff.c:1521-1537) */
    int **n14;
    double **scee14, **scnb14;
    /* hierarchical charge partitioning (oracle_hcp.c); NULL = off */
    struct orc_hcp *hcp;
};

int orc_nthreads(void);
double orc_egb_hcp(orc_t *o, const double *x, double *f, double *enb, double *eel,
                   double *reff_out, double *sumdeijda_out);
void orc_hcp_free(orc_t *o);

#endif
