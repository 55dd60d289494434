/*
 * A libsff client: the calls a NAB-generated C program makes for a
 * generalized-Born minimisation followed by a short dynamics run, written
 * against libsff's header name (sff.h) and linked with libnabc_amd.so.
 *
 *   sff_driver prmtop rst7 "mm_options string" maxiter mdsteps out.rst7
 *   sff_driver prmtop rst7 --parse
 *
 * --parse stops after rdparm/getxv and prints a few fields (no device).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sff.h"

int main(int argc, char **argv)
{
    PARMSTRUCT_T *prm;
    REAL_T *x, *f, *v, fret = 0.0, dgrad, dfpred = 10.0;
    INT_T natom, n3, maxiter, mdsteps, ier;
    if (argc < 4) {
        fprintf(stderr, "usage: %s prmtop rst7 options maxiter mdsteps out.rst7 | prmtop rst7 --parse\n", argv[0]);
        return 2;
    }
    prm = rdparm(argv[1]);
    if (!prm) return 3;
    natom = prm->Natom;
    x = (REAL_T *)calloc(3 * natom, sizeof(REAL_T));
    f = (REAL_T *)calloc(3 * natom, sizeof(REAL_T));
    v = (REAL_T *)calloc(3 * natom, sizeof(REAL_T));
    if (getxv(argv[2], natom, 0.0, x, v) != 0) return 4;
    if (!strcmp(argv[3], "--parse")) {
        printf("natom %d nres %d ipres_end %d nat3 %d nspm %d fsmax %.6f name0 %.4s res0 %.4s x0 %.4f\n",
               prm->Natom, prm->Nres, prm->Ipres[prm->Nres], prm->Nat3, prm->Nspm, prm->Fsmax,
               prm->AtomNames[0], prm->ResNames[0], x[0]);
        free_prm(prm);
        return 0;
    }
    if (argc < 7) return 2;
    maxiter = atoi(argv[4]);
    mdsteps = atoi(argv[5]);
    nabout = stdout;
    if (mm_options(argv[3]) != 0) return 5;
    if (mme_init_sff(prm, NULL, NULL, x, NULL) != 0) return 6;
    n3 = 3 * natom;
    dgrad = 3 * natom * 0.0001;
    ier = conjgrad(x, &n3, &fret, mme, &dgrad, &dfpred, &maxiter);
    printf("conjgrad returns %d\n", ier);
    if (mdsteps > 0 && md(3 * natom, mdsteps, x, f, v, mme) != 0) return 7;
    if (putxv(argv[6], "sff_driver", natom, 0.0, x, v) != 0) return 8;
    free(x);
    free(f);
    free(v);
    return free_prm(prm);
}
