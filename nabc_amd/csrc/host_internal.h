/*
 * host_internal.h -- state shared by the host-side C files of nabc_amd.
 * Not part of the public API.
 */
#ifndef NABC_AMD_HOST_INTERNAL_H
#define NABC_AMD_HOST_INTERNAL_H

#include <stdio.h>
#include "sff_gb.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Option values set by mm_options().  Defaults are those of the file-scope
 * statics in src/sff/sff.c:68-234; names follow the option keywords.
 */
typedef struct {
    double cut, cutnp, scnb, scee, rgbmax, gboffset, epsext, kappa, dielc;
    double dt, t, tautp, gamma_ln, temp0, boltz2, vlimit, genmass, tempi;
    double wcons, k4d, surften, dradius, deltar, rwater, alphanp, kappanp;
    double min_volume, wdssp;
    double hcp_h1, hcp_h2, hcp_h3, dhcp;
    int ntpr, nscm, nsnb, nsnp, nchk, nchk2, ntwx, ntpr_md, zerov;
    int gb, gbsa, gb_debug, nr_debug, gb2_debug, gbsa_debug, e_debug;
    int dield, irattle, ips, hcp, blocksize, dim, max_set_size;
    int mpi_min_tasks, cull_np_lists, use_lower_tri, dynamic_loops;
    /* nabc_amd extensions */
    int gbprec;          /* 0 = fp64, 1 = mixed fp32-pair/fp64-accumulate */
    unsigned long long md_seed; /* Langevin noise stream seed */
    int cgdevice;        /* conjgrad over mme: 1 = vectors on the device (default), 0 = on the host */
} nabc_opts_t;

extern nabc_opts_t nabc_opts;
extern FILE *nabout;

/* what mme_init_sff derives from a topology (model.c) */
typedef struct {
    PARMSTRUCT_T *prm;
    int *frozen;
    int nfrozen, gb, dield;
    double gboffset, gbneckscale;
    double *alpha, *beta, *gamma;
    int *neckidx;
    /* bonded tables: 0-based atom indices, one entry per Fourier term */
    int nb, na, nd, n14;
    int *bi, *bj, *ai, *aj, *ak, *di, *dj, *dk, *dl, *pi, *pj, *phb;
    double *brk, *breq, *atk, *ateq, *dpk, *dpn, *dph, *pscee, *pscnb, *pa, *pb;
} nabc_model;

struct gbhip_ctx;
struct gbhip_slot;
int nabc_model_build(nabc_model *m, PARMSTRUCT_T *prm, const int *frozen, const nabc_opts_t *o);
void nabc_model_free(nabc_model *m);
int nabc_model_context(const nabc_model *m, const double *x, const nabc_opts_t *o, void *stream,
                       struct gbhip_ctx **out);
/* the same in a slot (gbhip_ctx_create_in_slot) when slot is not NULL */
int nabc_model_context_in(const nabc_model *m, const double *x, const nabc_opts_t *o, void *stream,
                          struct gbhip_slot *slot, struct gbhip_ctx **out);

void nabc_seterr(const char *fmt, ...);

/* dense symmetric solvers (dense.hip) */
int nabc_dense_eigh(int n, double *A, double *w, int vectors);
int nabc_dense_posv(int n, const double *A, double *b);
int nabc_atomic_number_from_mass(double mass);
void nabc_build_14_list(PARMSTRUCT_T *prm);
void nabc_parm_complete(PARMSTRUCT_T *prm);

/* Amber netCDF trajectory writer (nctraj.c) */
typedef struct {
    FILE *fp;
    int natom, nframes;
    long long rec_begin, recsize; /* offset of the first frame, bytes per frame */
} nabc_nctraj;
int nabc_nctraj_create(nabc_nctraj *t, const char *path, int natom);
int nabc_nctraj_write(nabc_nctraj *t, const double *x, double time_ps);
void nabc_nctraj_close(nabc_nctraj *t);

#ifdef __cplusplus
}
#endif

#endif
