/*
 * gen.c -- deterministic synthetic polymer systems for the GB workloads.
 *
 * A system of N atoms is a single self-avoiding chain:
 *   - bonds of 1.5 A between consecutive atoms; the bond angle at every
 *     interior atom is drawn with cos(theta) uniform on [cos 130, cos 100];
 *     the torsion is uniform;
 *   - every new atom must lie inside the sphere of radius
 *     R = (3N / (4 pi 0.1))^(1/3) A (mean density 0.1 atoms/A^3) and at least
 *     NABC_GEN_MINDIST = 1.75 A from every atom except its bonded predecessor
 *     (2.0 A self-traps at this density); of GEN_TRIES trial directions the
 *     one with the fewest atoms within GEN_CROWD = 5 A is taken, which steers
 *     growth into free space;
 *     a step that fails GEN_TRIES trial directions backtracks, removing more
 *     atoms the longer the run of failures;
 *   - element classes H, O, N, C, S with probabilities 0.48/0.10/0.09/0.32/0.01,
 *     intrinsic radii 1.2/1.5/1.55/1.7/1.8 A, screening factors
 *     0.85/0.85/0.79/0.72/0.96, atomic numbers 1/8/7/6/16;
 *   - partial charges ~ N(0, 0.35) e, shifted to a net charge of zero and
 *     stored in the prmtop unit (e * 18.2223);
 *   - 5 Lennard-Jones types (one per class), Lorentz-Berthelot combination
 *     of (Rmin/2, epsilon) = H (1.487, 0.0157), O (1.25, 0.15),
 *     N (1.30, 0.15), C (1.35, 0.10), S (1.45, 0.25); the hydrogen values are
 *     those of aliphatic hydrogen in common biomolecular force fields: with
 *     a smaller wall, pairs of opposite partial charges pull together below
 *     1 A (the Coulomb attraction of two 0.35 e charges outweighs a
 *     (1.0, 0.02) wall there) and dynamics bounces them off the r^-12 wall;
 *   - bonds (Rk 300, Req 1.5), angles (Tk 40, Teq 115 deg) and dihedrals
 *     from chain neighbours; three dihedral types, one of them a two-term
 *     series (Pn = -3 then 1) so the multi-term path is exercised;
 *     SCEE 1.2 and SCNB 2.0;
 *   - exclusions: 1-2, 1-3 and 1-4 partners j > i of every atom.
 * The random numbers come from xoshiro256** seeded through splitmix64, so
 * the same (N, seed) gives the same bits on every machine with the same libm.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gbhip.h"
#include "host_internal.h"

#define NABC_GEN_MINDIST 1.75
#define GEN_CROWD 5.0
#define GEN_REACH 3
#define GEN_BOND 1.5
#define GEN_TRIES 24

typedef struct { unsigned long long s[4]; } rng_t;

static unsigned long long splitmix64(unsigned long long *x)
{
    unsigned long long z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, unsigned long long seed)
{
    int i;
    for (i = 0; i < 4; i++) r->s[i] = splitmix64(&seed);
}

static unsigned long long rotl(unsigned long long x, int k)
{
    return (x << k) | (x >> (64 - k));
}

static unsigned long long rng_next(rng_t *r)
{
    unsigned long long *s = r->s;
    unsigned long long res = rotl(s[1] * 5, 7) * 9;
    unsigned long long t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return res;
}

/* uniform in [0,1) with 53 random bits */
static double rng_uniform(rng_t *r)
{
    return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0);
}

static double rng_normal(rng_t *r)
{
    /* Box-Muller */
    double u1 = rng_uniform(r), u2 = rng_uniform(r);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

typedef struct {
    int g;            /* cells per dimension */
    double lo, h;     /* grid origin and cell size */
    int *head, *next, *cell_of;
} grid_t;

static int grid_cell(const grid_t *G, const double *p)
{
    int c[3], k;
    for (k = 0; k < 3; k++) {
        c[k] = (int)floor((p[k] - G->lo) / G->h);
        if (c[k] < 0) c[k] = 0;
        if (c[k] >= G->g) c[k] = G->g - 1;
    }
    return (c[2] * G->g + c[1]) * G->g + c[0];
}

static void grid_push(grid_t *G, const double *x, int i)
{
    int c = grid_cell(G, x + 3 * i);
    G->cell_of[i] = c;
    G->next[i] = G->head[c];
    G->head[c] = i;
}

/* atoms are removed in reverse order of insertion, so each is its cell's head */
static void grid_pop(grid_t *G, int i)
{
    int c = G->cell_of[i];
    G->head[c] = G->next[i];
}

/* -1 if p is closer than NABC_GEN_MINDIST to an atom other than skip,
 * else the number of atoms within GEN_CROWD of p */
static int clash(const grid_t *G, const double *x, const double *p, int skip)
{
    int c[3], k, dx, dy, dz, crowd = 0;
    const double d2min = NABC_GEN_MINDIST * NABC_GEN_MINDIST, d2crowd = GEN_CROWD * GEN_CROWD;
    for (k = 0; k < 3; k++) {
        c[k] = (int)floor((p[k] - G->lo) / G->h);
        if (c[k] < 0) c[k] = 0;
        if (c[k] >= G->g) c[k] = G->g - 1;
    }
    for (dz = -GEN_REACH; dz <= GEN_REACH; dz++) {
        int z = c[2] + dz;
        if (z < 0 || z >= G->g) continue;
        for (dy = -GEN_REACH; dy <= GEN_REACH; dy++) {
            int y = c[1] + dy;
            if (y < 0 || y >= G->g) continue;
            for (dx = -GEN_REACH; dx <= GEN_REACH; dx++) {
                int xx = c[0] + dx, j;
                if (xx < 0 || xx >= G->g) continue;
                for (j = G->head[(z * G->g + y) * G->g + xx]; j >= 0; j = G->next[j]) {
                    double a = p[0] - x[3 * j], b = p[1] - x[3 * j + 1], e = p[2] - x[3 * j + 2];
                    double d2 = a * a + b * b + e * e;
                    if (j == skip) continue;
                    if (d2 < d2min) return -1;
                    crowd += d2 < d2crowd;
                }
            }
        }
    }
    return crowd;
}

static void unit_random(rng_t *r, double *v)
{
    double z = 2.0 * rng_uniform(r) - 1.0, phi = 6.283185307179586 * rng_uniform(r);
    double s = sqrt(1.0 - z * z);
    v[0] = s * cos(phi);
    v[1] = s * sin(phi);
    v[2] = z;
}

/* Chain coordinates; returns 0 on success. */
static int gen_chain(int n, rng_t *r, double *x, long long *work_out)
{
    grid_t G;
    double R = cbrt(3.0 * n / (4.0 * 3.141592653589793 * 0.1));
    const double cmin = cos(130.0 * 3.141592653589793 / 180.0);
    const double cmax = cos(100.0 * 3.141592653589793 / 180.0);
    int i, fails = 0, ncell;
    long long work = 0, limit = 2000LL * n + 100000;

    G.h = NABC_GEN_MINDIST;
    G.lo = -R - G.h;
    G.g = (int)ceil((2.0 * R + 2.0 * G.h) / G.h) + 1;
    ncell = G.g * G.g * G.g;
    G.head = (int *)malloc(sizeof(int) * (size_t)ncell);
    G.next = (int *)malloc(sizeof(int) * (size_t)n);
    G.cell_of = (int *)malloc(sizeof(int) * (size_t)n);
    for (i = 0; i < ncell; i++) G.head[i] = -1;

    x[0] = x[1] = x[2] = 0.0;
    grid_push(&G, x, 0);
    if (n > 1) {
        double v[3];
        unit_random(r, v);
        for (i = 0; i < 3; i++) x[3 + i] = GEN_BOND * v[i];
        grid_push(&G, x, 1);
    }
    i = 2;
    while (i < n) {
        double u[3], e1[3], e2[3], a[3], len;
        int t, ok = 0;
        if (++work > limit) break;
        for (t = 0; t < 3; t++) u[t] = x[3 * (i - 1) + t] - x[3 * (i - 2) + t];
        len = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        for (t = 0; t < 3; t++) u[t] /= len;
        /* any vector not parallel to u */
        if (fabs(u[0]) < 0.9) { a[0] = 1; a[1] = 0; a[2] = 0; }
        else { a[0] = 0; a[1] = 1; a[2] = 0; }
        e1[0] = u[1] * a[2] - u[2] * a[1];
        e1[1] = u[2] * a[0] - u[0] * a[2];
        e1[2] = u[0] * a[1] - u[1] * a[0];
        len = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        for (t = 0; t < 3; t++) e1[t] /= len;
        e2[0] = u[1] * e1[2] - u[2] * e1[1];
        e2[1] = u[2] * e1[0] - u[0] * e1[2];
        e2[2] = u[0] * e1[1] - u[1] * e1[0];
        {
            /* of the trial positions that fit, keep the one with the fewest
             * atoms within GEN_CROWD (growth is steered into free space) */
            int best = 1 << 30;
            for (t = 0; t < GEN_TRIES; t++) {
                /* theta is the bond angle at atom i-1; the new bond makes the
                 * angle pi-theta with u, whose cosine is -cos(theta) */
                double ct = cmin + (cmax - cmin) * rng_uniform(r);
                double c = -ct, s = sqrt(1.0 - c * c);
                double phi = 6.283185307179586 * rng_uniform(r);
                double cp = cos(phi), sp = sin(phi), p[3];
                int k, crowd;
                for (k = 0; k < 3; k++)
                    p[k] = x[3 * (i - 1) + k] + GEN_BOND * (c * u[k] + s * (cp * e1[k] + sp * e2[k]));
                if (p[0] * p[0] + p[1] * p[1] + p[2] * p[2] > R * R) continue;
                crowd = clash(&G, x, p, i - 1);
                if (crowd < 0) continue;
                if (crowd < best) {
                    best = crowd;
                    for (k = 0; k < 3; k++) x[3 * i + k] = p[k];
                    ok = 1;
                }
            }
        }
        if (ok) {
            grid_push(&G, x, i);
            i++;
            if (fails > 0) fails--;
        } else {
            /* backtrack: remove 1 + fails/4 atoms, keeping the first two */
            int k, b = 1 + fails / 4;
            fails += 2;
            for (k = 0; k < b && i > 2; k++) {
                i--;
                grid_pop(&G, i);
            }
        }
    }
    free(G.head);
    free(G.next);
    free(G.cell_of);
    if (work_out) *work_out = work;
    if (i < n) {
        nabc_seterr("gbgen: chain growth did not finish (%d of %d atoms placed)", i, n);
        return -1;
    }
    return 0;
}

static const double cls_prob[5] = {0.48, 0.10, 0.09, 0.32, 0.01};
static const char cls_name[5] = {'H', 'O', 'N', 'C', 'S'};
static const double cls_radius[5] = {1.2, 1.5, 1.55, 1.7, 1.8};
static const double cls_screen[5] = {0.85, 0.85, 0.79, 0.72, 0.96};
static const int cls_atnum[5] = {1, 8, 7, 6, 16};
static const double cls_mass[5] = {1.008, 16.00, 14.01, 12.01, 32.06};
static const double cls_rmin2[5] = {1.487, 1.25, 1.30, 1.35, 1.45};
static const double cls_eps[5] = {0.0157, 0.15, 0.15, 0.10, 0.25};

int gbgen_build(int natom, unsigned long long seed, PARMSTRUCT_T **prm_out, double *x_out)
{
    rng_t r;
    PARMSTRUCT_T *p;
    int n = natom, i, k, nt = 5;
    double qsum = 0.0;
    int *cls;

    if (natom < 4) {
        nabc_seterr("gbgen: need at least 4 atoms");
        return -1;
    }
    rng_seed(&r, seed ^ (0x5851F42D4C957F2DULL * (unsigned long long)natom));
    if (gen_chain(n, &r, x_out, NULL) != 0) return -1;

    p = (PARMSTRUCT_T *)calloc(1, sizeof(PARMSTRUCT_T));
    snprintf(p->ititl, sizeof(p->ititl), "synthetic GB polymer N=%d seed=%llu", natom, seed);
    p->Natom = n;
    p->Ntypes = nt;
    p->Ntype2d = nt * nt;
    p->Nttyp = nt * (nt + 1) / 2;
    p->Natyp = nt;
    p->Nres = (n + 19) / 20;
    p->Numbnd = 1;
    p->Numang = 1;
    p->Nptra = 4;
    p->Nphb = 0;

    cls = (int *)malloc(sizeof(int) * (size_t)n);
    p->AtomNames = (NABC_NAME_T *)calloc((size_t)n, NABC_NAMESIZE);
    p->Charges = (double *)calloc((size_t)n, sizeof(double));
    p->Masses = (double *)calloc((size_t)n, sizeof(double));
    p->Rborn = (double *)calloc((size_t)n, sizeof(double));
    p->Fs = (double *)calloc((size_t)n, sizeof(double));
    p->Iac = (int *)calloc((size_t)n, sizeof(int));
    p->AtomicNum = (int *)calloc((size_t)n, sizeof(int));
    for (i = 0; i < n; i++) {
        double u = rng_uniform(&r), acc = 0.0;
        int c = 4;
        for (k = 0; k < 5; k++) {
            acc += cls_prob[k];
            if (u < acc) { c = k; break; }
        }
        cls[i] = c;
        p->AtomNames[i][0] = cls_name[c];
        p->AtomNames[i][1] = '\0';
        p->Masses[i] = cls_mass[c];
        p->Rborn[i] = cls_radius[c];
        p->Fs[i] = cls_screen[c];
        p->Iac[i] = c + 1;
        p->AtomicNum[i] = cls_atnum[c];
        p->Charges[i] = 0.35 * rng_normal(&r);
        qsum += p->Charges[i];
    }
    for (i = 0; i < n; i++) p->Charges[i] = (p->Charges[i] - qsum / n) * 18.2223;

    /* Lennard-Jones: Cno is 1-based into the packed lower triangle */
    p->Cno = (int *)calloc((size_t)(nt * nt), sizeof(int));
    p->Cn1 = (double *)calloc((size_t)p->Nttyp, sizeof(double));
    p->Cn2 = (double *)calloc((size_t)p->Nttyp, sizeof(double));
    for (i = 0; i < nt; i++) {
        for (k = 0; k <= i; k++) {
            int ic = i * (i + 1) / 2 + k;
            double rm = cls_rmin2[i] + cls_rmin2[k], ep = sqrt(cls_eps[i] * cls_eps[k]);
            double r6 = pow(rm, 6.0);
            p->Cn1[ic] = ep * r6 * r6;
            p->Cn2[ic] = 2.0 * ep * r6;
            p->Cno[i * nt + k] = ic + 1;
            p->Cno[k * nt + i] = ic + 1;
        }
    }

    p->Rk = (double *)malloc(sizeof(double)); p->Rk[0] = 300.0;
    p->Req = (double *)malloc(sizeof(double)); p->Req[0] = 1.5;
    p->Tk = (double *)malloc(sizeof(double)); p->Tk[0] = 40.0;
    p->Teq = (double *)malloc(sizeof(double)); p->Teq[0] = 115.0 * 3.141592653589793 / 180.0;
    p->Pk = (double *)malloc(4 * sizeof(double));
    p->Pn = (double *)malloc(4 * sizeof(double));
    p->Phase = (double *)malloc(4 * sizeof(double));
    p->Scee = (double *)malloc(4 * sizeof(double));
    p->Scnb = (double *)malloc(4 * sizeof(double));
    p->Pk[0] = 0.30; p->Pn[0] = 3.0;  p->Phase[0] = 0.0;
    p->Pk[1] = 0.25; p->Pn[1] = 2.0;  p->Phase[1] = 3.14159400;
    p->Pk[2] = 0.20; p->Pn[2] = -3.0; p->Phase[2] = 0.0;
    p->Pk[3] = 0.15; p->Pn[3] = 1.0;  p->Phase[3] = 0.0;
    for (i = 0; i < 4; i++) {
        p->Scee[i] = 1.0 / 1.2;
        p->Scnb[i] = 1.0 / 2.0;
    }

    /* bonds / angles / dihedrals: terms touching an H go to the "H" lists */
    p->BondHAt1 = (int *)calloc((size_t)n, sizeof(int));
    p->BondHAt2 = (int *)calloc((size_t)n, sizeof(int));
    p->BondHNum = (int *)calloc((size_t)n, sizeof(int));
    p->BondAt1 = (int *)calloc((size_t)n, sizeof(int));
    p->BondAt2 = (int *)calloc((size_t)n, sizeof(int));
    p->BondNum = (int *)calloc((size_t)n, sizeof(int));
    for (i = 0; i + 1 < n; i++) {
        int h = cls[i] == 0 || cls[i + 1] == 0;
        if (h) {
            p->BondHAt1[p->Nbonh] = 3 * i; p->BondHAt2[p->Nbonh] = 3 * (i + 1);
            p->BondHNum[p->Nbonh++] = 1;
        } else {
            p->BondAt1[p->Mbona] = 3 * i; p->BondAt2[p->Mbona] = 3 * (i + 1);
            p->BondNum[p->Mbona++] = 1;
        }
    }
    p->Nbona = p->Mbona;
    p->AngleHAt1 = (int *)calloc((size_t)n, sizeof(int));
    p->AngleHAt2 = (int *)calloc((size_t)n, sizeof(int));
    p->AngleHAt3 = (int *)calloc((size_t)n, sizeof(int));
    p->AngleHNum = (int *)calloc((size_t)n, sizeof(int));
    p->AngleAt1 = (int *)calloc((size_t)n, sizeof(int));
    p->AngleAt2 = (int *)calloc((size_t)n, sizeof(int));
    p->AngleAt3 = (int *)calloc((size_t)n, sizeof(int));
    p->AngleNum = (int *)calloc((size_t)n, sizeof(int));
    for (i = 0; i + 2 < n; i++) {
        int h = cls[i] == 0 || cls[i + 1] == 0 || cls[i + 2] == 0;
        if (h) {
            int m = p->Ntheth++;
            p->AngleHAt1[m] = 3 * i; p->AngleHAt2[m] = 3 * (i + 1);
            p->AngleHAt3[m] = 3 * (i + 2); p->AngleHNum[m] = 1;
        } else {
            int m = p->Mtheta++;
            p->AngleAt1[m] = 3 * i; p->AngleAt2[m] = 3 * (i + 1);
            p->AngleAt3[m] = 3 * (i + 2); p->AngleNum[m] = 1;
        }
    }
    p->Ntheta = p->Mtheta;
    p->DihHAt1 = (int *)calloc((size_t)n, sizeof(int));
    p->DihHAt2 = (int *)calloc((size_t)n, sizeof(int));
    p->DihHAt3 = (int *)calloc((size_t)n, sizeof(int));
    p->DihHAt4 = (int *)calloc((size_t)n, sizeof(int));
    p->DihHNum = (int *)calloc((size_t)n, sizeof(int));
    p->DihAt1 = (int *)calloc((size_t)n, sizeof(int));
    p->DihAt2 = (int *)calloc((size_t)n, sizeof(int));
    p->DihAt3 = (int *)calloc((size_t)n, sizeof(int));
    p->DihAt4 = (int *)calloc((size_t)n, sizeof(int));
    p->DihNum = (int *)calloc((size_t)n, sizeof(int));
    for (i = 0; i + 3 < n; i++) {
        int h = cls[i] == 0 || cls[i + 1] == 0 || cls[i + 2] == 0 || cls[i + 3] == 0;
        int type = 1 + (int)((unsigned)(i * 2654435761u) >> 30) % 3; /* 1, 2 or 3 */
        int *A1 = h ? p->DihHAt1 : p->DihAt1, *A2 = h ? p->DihHAt2 : p->DihAt2;
        int *A3 = h ? p->DihHAt3 : p->DihAt3, *A4 = h ? p->DihHAt4 : p->DihAt4;
        int *T = h ? p->DihHNum : p->DihNum;
        int m = h ? p->Nphih++ : p->Mphia++;
        A1[m] = 3 * i; A2[m] = 3 * (i + 1); A3[m] = 3 * (i + 2); A4[m] = 3 * (i + 3);
        T[m] = type;
    }
    p->Nphia = p->Mphia;

    /* exclusions: partners i+1..i+3; atoms with none get one 0 placeholder */
    p->Iblo = (int *)calloc((size_t)n, sizeof(int));
    p->ExclAt = (int *)calloc((size_t)(3 * n + 3), sizeof(int));
    p->Nnb = 0;
    for (i = 0; i < n; i++) {
        int c = 0;
        for (k = 1; k <= 3; k++)
            if (i + k < n) p->ExclAt[p->Nnb + c++] = i + k + 1;
        if (c == 0) p->ExclAt[p->Nnb + c++] = 0;
        p->Iblo[i] = c;
        p->Nnb += c;
    }
    /* residues of 20 atoms, 0-based first atoms with Ipres[Nres] = Natom */
    p->Ipres = (int *)calloc((size_t)p->Nres + 1, sizeof(int));
    for (i = 0; i < p->Nres; i++) p->Ipres[i] = 20 * i;
    p->Ipres[p->Nres] = n;
    p->HB12 = (double *)calloc(1, sizeof(double));
    p->HB10 = (double *)calloc(1, sizeof(double));

    p->Fsmax = 0.0;
    for (i = 0; i < n; i++) {
        double si = p->Fs[i] * (p->Rborn[i] - 0.09);
        if (si > p->Fsmax) p->Fsmax = si;
    }
    nabc_parm_complete(p);
    nabc_build_14_list(p);
    free(cls);
    *prm_out = p;
    return 0;
}

int gbgen_write(int natom, unsigned long long seed, const char *prmtop_path,
                const char *rst7_path)
{
    PARMSTRUCT_T *p = NULL;
    double *x = (double *)malloc(sizeof(double) * 3 * (size_t)natom);
    int rc = gbgen_build(natom, seed, &p, x);
    if (rc == 0) rc = wrparm(p, (char *)prmtop_path);
    if (rc == 0) rc = putxv((char *)rst7_path, p->ititl, natom, 0.0, x, NULL);
    free_prm(p);
    free(x);
    return rc;
}
