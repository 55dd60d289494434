/*
 * sff_gb.h -- host-side C API of nabc_amd, mirroring the libsff entry points
 * that sit around the generalized-Born energy path.
 *
 * Every function keeps the name, argument order and argument meaning of the
 * libsff routine it replaces (src/sff/sff.h:124-180 in dacase/nabc):
 *
 *   rdparm        prm.c:584          read an Amber prmtop (%FLAG and old layouts)
 *   mm_options    mm_options.l:276   keyword=value option string
 *   mme_init_sff  sff.c:858          GB parameter set-up, exclusion lists
 *   mme           sff.c:1854         energy + gradient, returns the total
 *   md            sff.c:2071         velocity-Verlet / Langevin dynamics
 *   conjgrad      conjgrad.c:34      Polak-Ribiere CG with Powell restarts
 *   mme2          sff2.c:6176        energy, gradient and Hessian
 *   newton        newton.c:357       Newton-Raphson minimiser with line search
 *   nmode         nmode.c:1596       normal modes and thermochemistry
 *   getxv, putxv  sff.c:651, :733    rst7 coordinate / velocity files
 *
 * The generalized-Born pair work (egb() and nblist() of eff.c/nblist.c) runs
 * on the GPU through the C-ABI in gbhip.h; mme() here orchestrates it.
 *
 * One documented deviation: where libsff prints and calls exit(1), these
 * functions print the same message to stderr and return a negative code
 * (mme returns NaN).  mm_get_error() gives the last message.
 */
#ifndef NABC_AMD_SFF_GB_H
#define NABC_AMD_SFF_GB_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NABC_NAMESIZE 6
/* atom, residue and type names: 4 characters of the prmtop plus padding
 * and the terminating NUL (NAME_T, sff.h:36) */
typedef char NABC_NAME_T[NABC_NAMESIZE];

/*
 * Topology as read from a prmtop: PARMSTRUCT_T of src/sff/sff.h:40-62,
 * field for field in the same order and with the same types (int for
 * INT_T, double for REAL_T), so code compiled against libsff's header
 * reads every field at the offset it expects.  Atom indices in the
 * bond/angle/dihedral arrays are stored as in the file (3*(atom-1),
 * negative third/fourth index flags as in Amber); ExclAt is 1-based with
 * 0 meaning "none".  Charges are in the prmtop's internal unit (electron
 * charge * 18.2223).  Scee/Scnb hold the inverted SCEE/SCNB factors.
 * AtomRes holds each atom's 0-based residue; Boundary/Nspm the molecule
 * boundaries (one molecule of Natom atoms without a box); Box[0..2] the
 * box lengths (999999999 without a box) and Box[3] its angle; Gvdw is
 * allocated and zero (the reference leaves its computation disabled);
 * Nstrand/Ncomplex and Ipstrand/Ipcomplex are 1 / {1} unless the file has
 * the HCP strand and complex pointers.
 */
typedef struct parm {
    char ititl[81];
    int IfBox, Nmxrs, IfCap,
        Natom, Ntypes, Nbonh, Mbona, Ntheth, Mtheta,
        Nphih, Mphia, Nhparm, Nparm, Nnb, Nres,
        Nbona, Ntheta, Nphia, Numbnd, Numang, Nptra,
        Natyp, Nphb, Nat3, Ntype2d, Nttyp, Nspm, Iptres, Nspsol,
        Ipatm, Natcap, Numextra;
    NABC_NAME_T *AtomNames, *ResNames, *AtomType, *AtomTree;
    double *Charges, *Masses, *Rk, *Req, *Tk, *Teq, *Pk, *Pn, *Phase,
        *Solty, *Cn1, *Cn2, *HB12, *HB10, *Rborn, *Fs;
    double Box[4], Cutcap, Xcap, Ycap, Zcap, Fsmax;
    int *Iac, *Iblo, *Cno, *Ipres, *ExclAt, *TreeJoin,
        *AtomRes, *BondHAt1, *BondHAt2, *BondHNum, *BondAt1, *BondAt2,
        *BondNum, *AngleHAt1, *AngleHAt2, *AngleHAt3, *AngleHNum,
        *AngleAt1, *AngleAt2, *AngleAt3, *AngleNum, *DihHAt1,
        *DihHAt2, *DihHAt3, *DihHAt4, *DihHNum, *DihAt1, *DihAt2,
        *DihAt3, *DihAt4, *DihNum, *Boundary, *AtomicNum;
    int *N14pairs, *N14pairlist;
    double *Gvdw, *Scee, *Scnb, *N14sceelist, *N14scnblist;
    int Nstrand, Ncomplex;
    int *Ipstrand, *Ipcomplex;
} PARMSTRUCT_T;

/*
 * prmtop I/O (prm.c:584 rdparm, prm.c:1997 wrparm, free_prm sff.h:178).
 * rdparm reads both layouts the reference reads: %FLAG/%FORMAT files and
 * the older fixed-order files without flags (whose GB radii and screening
 * factors come from the first letter of each atom name, prm.c:1656-1741).
 */
PARMSTRUCT_T *rdparm(char *name);
int wrparm(PARMSTRUCT_T *prm, char *name);
int free_prm(PARMSTRUCT_T *prm);

/* rst7 I/O (sff.c:651 getxv, sff.c:733 putxv). */
int getxv(char *fname, int natom, double start_time, double *x, double *v);
int putxv(char *fname, char *title, int natom, double start_time,
          double *x, double *v);

/* Options (mm_options.l:28-97) and the defaults of sff.c:68-234. */
int mm_options(char *opts);
void mm_reset_options(void);
const char *mm_get_error(void);

/*
 * Initialisation (sff.c:858).  frozen[] and constrained[] are per-atom
 * flags (may be NULL = none).  x0 is kept for restraints (unused by GB).
 * ncname (may be NULL): an Amber netCDF trajectory file is created there
 * and md() appends a frame every ntwx steps (sff.c:1164-1168, :2341-2342;
 * the layout of AmberNetcdf.c's netcdfCreate, plus the frame times).
 */
int mme_init_sff(PARMSTRUCT_T *prm, int *frozen, int *constrained,
                 double *x0, char *ncname);

/*
 * Energy and gradient (sff.c:1854 -> eff.c:2442).  x and f are 3*Natom,
 * f receives the gradient (not the force).  iter semantics: -1 prints the
 * per-term table, -3 releases the device state, -4 forces a pair-list
 * rebuild, other negatives run silently without a rebuild.
 */
double mme(double *x, double *f, int *iter);

/*
 * Per-term energies of the last mme() call, indexed as ene[] in mme34
 * (eff.c:2617-2874): [0] total, [1] enb, [2] eel, [3] bond, [4] angle,
 * [5] dihedral, [7] enb14, [8] eel14, [9] constraint, [10] egb,
 * [11] esurf, [12] evdwnp.  Also the rms gradient of that call.
 */
void mme_get_energies(double ene[13], double *frms);

/*
 * The GPU context behind mme() (a gbhip_ctx *, see gbhip.h), created from
 * coordinates x if needed, with the bonded terms of the mme_init_sff
 * topology attached; NULL on error.  Lets a caller drive the device
 * directly (gbhip_md_*, gbhip_mme_device) on the same state.
 */
void *mme_context(double *x);

/* Dynamics (sff.c:2071) and conjugate gradients (conjgrad.c:34). */
int md(int n, int maxstep, double *x, double *f, double *v,
       double (*func)(double *, double *, int *));
int conjgrad(double x[], int *n, double *f,
             double (*func)(double *, double *, int *),
             double *rmsgrad, double *dfpred, int *maxiter);

/*
 * Second derivatives (sff.h:142-162; the non-ScaLAPACK meaning of the
 * arguments: the descriptor and context pointers are not read).
 *
 * mme2(x, f, g, m, grad, ..., natom, iter, name): energy (returned),
 * gradient f[3N] (also copied to grad when not NULL), Hessian g[3N*3N]
 * (symmetric), atomic masses m[N], *natom, and name[i] the first letter of
 * atom i's name.  iter -3 releases its state.  The Hessian is the device's
 * (gbhip_hessian): central differences (step 1e-4 A) of the analytic FP64
 * gradient of every term, where sff2.c forms it analytically; the
 * reference's normal modes of test/nab/asp_nm1 reproduce at their printed
 * precision.  Whatever gbprec says, mme2 evaluates in FP64.
 *
 * newton(x, n, f, func1, func2, rmsgrad, nradd, maxiter): Newton-Raphson
 * (newton.c:357-1371): Hessian from func2, translations and rotations
 * level-shifted (3500 d d^T), nradd added to the diagonal, the step from a
 * Cholesky solve, then a cubic line search along it
 * with func1 unless the step is already below 1e-4 rms.  Returns the
 * iteration count on convergence (sum g^2 < rmsgrad^2 n), -1 after maxiter
 * iterations, -2 when the Hessian is not positive definite.
 *
 * nmode(x, n, func, eigp, ntrun, eta, hrmax, ioseen): mass-weighted Hessian
 * from func, its eigenvalues (Householder reduction, on the GPU from order
 * 1024, and QR iteration) as frequencies in cm^-1 and the
 * thermochemistry of nmode.c's thermo() at 298.15 K and 1 atm.  ntrun 0
 * and 1 print the full table (and with eigp > 0 write the first eigp modes
 * to the file "vecs"); ntrun 2 prints the lowest eigp frequencies (all
 * eigenvalues are computed, where the reference uses ARPACK).  ntrun 3
 * (Langevin modes) is not provided (returns -1).
 */
double mme2(double *x, double *f, double *g, double *m, double *grad, int *descF_PxQ, int *descF_1x1,
            int *descG_PxQ, int *context_PxQ, int *context_1x1, int *context_Nx1, int *gridim, int *natom,
            int *iter, char *name);
typedef double (*nabc_mme2_fn)(double *, double *, double *, double *, double *, int *, int *, int *, int *, int *,
                               int *, int *, int *, int *, char *);
int newton(double x[], int *n, double *f, double (*func1)(double *, double *, int *), nabc_mme2_fn func2,
           double *rmsgrad, double *nradd, int *maxiter);
int nmode(double *x, int n, nabc_mme2_fn func, int eigp, int ntrun, double eta, double hrmax, int ioseen);

/* Output stream of the ff:/md: lines and mm_options echo (libsff's
 * nabout, sff.h:119; NULL means stdout); mm_set_output sets it too. */
extern FILE *nabout;
void mm_set_output(FILE *fp);

/*
 * Amber netCDF trajectories and restarts (AmberNetcdf.h:22-62, declared in
 * sff.h:240-274): the same structure (field names, types and order) and
 * functions, written and read without the netCDF library.  ncid indexes a
 * table of at most 64 open files; the *VID fields are -1 for a variable
 * the file lacks and 1 otherwise.  netcdfCreate writes trajectories
 * without box information only (isBox must be 0).  Return 0 on success, 1
 * on failure (message via mm_get_error).
 */
struct AmberNetcdf {
    double temp0;       /* temperature of a restart (TempVID != -1) */
    double restartTime; /* simulation time of a restart */
    int isNCrestart;    /* 0 trajectory, 1 restart */
    int ncid;
    int frameDID;
    int ncframe;        /* frames in the file */
    int currentFrame;   /* next frame of netcdfGet/WriteNextFrame */
    int atomDID;
    int ncatom;
    int ncatom3;
    int coordVID;
    int velocityVID;
    int cellAngleVID;
    int cellLengthVID;
    int spatialDID;
    int labelDID;
    int cell_spatialDID;
    int cell_angularDID;
    int spatialVID;
    int timeVID;
    int cell_spatialVID;
    int cell_angularVID;
    int TempVID;
};
int netcdfDebug(struct AmberNetcdf *A);
int netcdfLoad(struct AmberNetcdf *A, char *filename);
int netcdfClose(struct AmberNetcdf *A);
int netcdfWriteRestart(char *filename, int natom, double *X, double *V, double *box, double time,
                       double temp);
int netcdfCreate(struct AmberNetcdf *A, char *filename, int natom, int isBox);
int netcdfGetVelocity(struct AmberNetcdf *A, int set, double *V);
int netcdfGetFrame(struct AmberNetcdf *A, int set, double *X, double *box);
int netcdfGetNextFrame(struct AmberNetcdf *A, double *X, double *box);
int netcdfWriteFrame(struct AmberNetcdf *A, int set, double *X, double *box);
int netcdfWriteNextFrame(struct AmberNetcdf *A, double *X, double *box);
int netcdfInfo(struct AmberNetcdf *A);

#ifdef __cplusplus
}
#endif

#endif
