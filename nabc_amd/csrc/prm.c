/*
 * prm.c -- Amber prmtop reader and writer for nabc_amd.
 *
 * Reads the %FLAG/%FORMAT ("new") prmtop layout that rdparm() in
 * src/sff/prm.c:584-1911 reads, with the same post-processing:
 *   - charges stay in the file's internal unit (e * 18.2223), prm.c:777-792;
 *   - SCEE/SCNB are stored inverted, defaults 1/1.2 and 1/2, prm.c:1074-1114;
 *   - ATOMIC_NUMBER is read or, if absent, taken as the element whose
 *     standard atomic weight is nearest to the mass, prm.c:85-99, :1122-1133;
 *   - Fsmax = max_i Fs_i * (Rborn_i - 0.09) from the file's SCREEN values,
 *     prm.c:1621-1656 (gb 7/8 later replace Fs but not Fsmax);
 *   - the 1-4 pair list is built from the dihedrals in file order, skipping
 *     entries whose third or fourth atom index is negative, prm.c:1854-1889.
 * Sections are located by name and parsed by the field width that their
 * %FORMAT line states.  Old-format (pre-%FLAG) files are rejected.
 */
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sff_gb.h"
#include "host_internal.h"

typedef struct {
    const char *name;   /* flag name */
    char *fmt;          /* points into the buffer, after "%FORMAT(" */
    char *data;         /* first data line */
    char *end;          /* one past the last data byte */
} section_t;

typedef struct {
    char *buf;
    size_t len;
    section_t *sec;
    int nsec;
} prmfile_t;

static void prmfile_free(prmfile_t *pf)
{
    free(pf->buf);
    free(pf->sec);
}

static int prmfile_load(const char *name, prmfile_t *pf)
{
    FILE *fp = fopen(name, "rb");
    long n;
    char *p;
    int cap = 64;

    memset(pf, 0, sizeof(*pf));
    if (!fp) {
        nabc_seterr("Cannot read parm file %s", name);
        return -1;
    }
    fseek(fp, 0, SEEK_END);
    n = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    pf->buf = (char *)malloc((size_t)n + 2);
    if (!pf->buf || fread(pf->buf, 1, (size_t)n, fp) != (size_t)n) {
        fclose(fp);
        nabc_seterr("read error on parm file %s", name);
        free(pf->buf);
        pf->buf = NULL;
        return -1;
    }
    fclose(fp);
    pf->buf[n] = '\n';
    pf->buf[n + 1] = '\0';
    pf->len = (size_t)n;
    if (strncmp(pf->buf, "%VERSION", 8) != 0 && strncmp(pf->buf, "%FLAG", 5) != 0) {
        nabc_seterr("parm file %s is not in %%FLAG format (old-format prmtops are not supported)", name);
        free(pf->buf);
        pf->buf = NULL;
        return -1;
    }
    pf->sec = (section_t *)malloc(sizeof(section_t) * cap);
    for (p = pf->buf; *p;) {
        char *eol = strchr(p, '\n');
        if (!eol) break;
        if (strncmp(p, "%FLAG ", 6) == 0) {
            section_t s;
            char *q = p + 6, *e;
            while (*q == ' ') q++;
            e = q;
            while (e < eol && !isspace((unsigned char)*e)) e++;
            *e = '\0';
            s.name = q;
            s.fmt = NULL;
            p = eol + 1;
            /* the %FORMAT line (there may be %COMMENT lines first) */
            while (*p == '%' && strncmp(p, "%FLAG", 5) != 0) {
                char *eol2 = strchr(p, '\n');
                if (strncmp(p, "%FORMAT(", 8) == 0) s.fmt = p + 8;
                p = eol2 + 1;
            }
            s.data = p;
            while (*p && strncmp(p, "%FLAG", 5) != 0) {
                char *eol2 = strchr(p, '\n');
                if (!eol2) { p += strlen(p); break; }
                p = eol2 + 1;
            }
            s.end = p;
            if (pf->nsec == cap) {
                cap *= 2;
                pf->sec = (section_t *)realloc(pf->sec, sizeof(section_t) * cap);
            }
            pf->sec[pf->nsec++] = s;
            continue;
        }
        p = eol + 1;
    }
    return 0;
}

static section_t *prm_find(prmfile_t *pf, const char *name)
{
    int i;
    for (i = 0; i < pf->nsec; i++)
        if (strcmp(pf->sec[i].name, name) == 0) return &pf->sec[i];
    return NULL;
}

/* Field width from a Fortran edit descriptor such as 10I8, 5E16.8, 20a4. */
static int fmt_width(const char *fmt)
{
    const char *p = fmt;
    while (*p && isdigit((unsigned char)*p)) p++;
    if (*p) p++;
    return atoi(p) > 0 ? atoi(p) : 16;
}

/* Parse up to n numbers of a section into out (as doubles). */
static int sec_numbers(section_t *s, int n, double *out)
{
    int w = s->fmt ? fmt_width(s->fmt) : 16;
    int k = 0;
    char *p = s->data;
    char field[64];
    while (k < n && p < s->end) {
        char *eol = strchr(p, '\n');
        char *q = p;
        if (!eol || eol > s->end) eol = s->end;
        while (k < n && q < eol) {
            int len = (int)(eol - q) < w ? (int)(eol - q) : w;
            char *endp;
            memcpy(field, q, (size_t)len);
            field[len] = '\0';
            q += len;
            {
                char *t = field;
                while (*t == ' ') t++;
                if (*t == '\0' || *t == '\r') continue;
                /* Fortran D exponents */
                for (endp = t; *endp; endp++)
                    if (*endp == 'D' || *endp == 'd') *endp = 'E';
                out[k++] = strtod(t, &endp);
            }
        }
        p = eol + 1;
    }
    return k;
}

static int read_dbl(prmfile_t *pf, const char *name, int n, double **dst, int required)
{
    section_t *s = prm_find(pf, name);
    *dst = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    if (!s) {
        if (required) {
            nabc_seterr("rdparm: missing %%FLAG %s", name);
            return -1;
        }
        return 1;
    }
    if (n > 0 && sec_numbers(s, n, *dst) != n) {
        nabc_seterr("rdparm: short section %s (wanted %d values)", name, n);
        return -1;
    }
    return 0;
}

static int read_int(prmfile_t *pf, const char *name, int n, int **dst, int required)
{
    double *tmp;
    int i, rc = read_dbl(pf, name, n, &tmp, required);
    *dst = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
    if (rc == 0)
        for (i = 0; i < n; i++) (*dst)[i] = (int)lround(tmp[i]);
    free(tmp);
    return rc;
}

/*
 * Nearest standard atomic weight (IUPAC values, elements 0..36; index 0 is
 * an extra point of mass 0).  Same selection rule as get_atomic_number,
 * prm.c:85-99.
 */
static const double std_mass[37] = {
    0.0, 1.0079, 4.0026, 6.941, 9.0122, 10.811, 12.0107, 14.0067, 15.9994,
    18.9984, 20.1797, 22.9898, 24.3050, 26.9815, 28.0855, 30.9738, 32.065,
    35.453, 39.948, 39.0983, 40.078, 44.9559, 47.867, 50.9415, 51.9961,
    54.9380, 55.845, 58.9331, 58.6934, 63.546, 65.409, 69.723, 72.64,
    74.9216, 78.96, 79.904, 83.798};

int nabc_atomic_number_from_mass(double mass)
{
    int i, best = 0;
    double dmin = fabs(mass);
    for (i = 0; i < 37; i++) {
        double d = fabs(mass - std_mass[i]);
        if (d < dmin) {
            dmin = d;
            best = i;
        }
    }
    return best;
}

void nabc_build_14_list(PARMSTRUCT_T *prm)
{
    int n = prm->Natom, i, k, idx;
    int *cnt = (int *)calloc((size_t)n, sizeof(int));
    int *cap = (int *)calloc((size_t)n, sizeof(int));
    int **lst = (int **)calloc((size_t)n, sizeof(int *));
    int **typ = (int **)calloc((size_t)n, sizeof(int *));
    int total = 0, pass;

    for (pass = 0; pass < 2; pass++) {
        int nd = pass == 0 ? prm->Nphih : prm->Mphia;
        int *a1 = pass == 0 ? prm->DihHAt1 : prm->DihAt1;
        int *a3 = pass == 0 ? prm->DihHAt3 : prm->DihAt3;
        int *a4 = pass == 0 ? prm->DihHAt4 : prm->DihAt4;
        int *tp = pass == 0 ? prm->DihHNum : prm->DihNum;
        for (i = 0; i < nd; i++) {
            int iat = a1[i] / 3, kat = a3[i] / 3, lat = a4[i] / 3;
            int lo, hi;
            if (kat < 0 || lat < 0) continue;
            lo = iat < lat ? iat : lat;
            hi = iat > lat ? iat : lat;
            if (cnt[lo] == cap[lo]) {
                cap[lo] = cap[lo] ? 2 * cap[lo] : 4;
                lst[lo] = (int *)realloc(lst[lo], sizeof(int) * cap[lo]);
                typ[lo] = (int *)realloc(typ[lo], sizeof(int) * cap[lo]);
            }
            lst[lo][cnt[lo]] = hi;
            typ[lo][cnt[lo]++] = tp[i] - 1;
            total++;
        }
    }
    prm->N14pairs = (int *)calloc((size_t)n, sizeof(int));
    prm->N14pairlist = (int *)calloc((size_t)(total + 1), sizeof(int));
    prm->N14sceelist = (double *)calloc((size_t)(total + 1), sizeof(double));
    prm->N14scnblist = (double *)calloc((size_t)(total + 1), sizeof(double));
    idx = 0;
    for (i = 0; i < n; i++) {
        prm->N14pairs[i] = cnt[i];
        /* the flattened list skips the last atom, as prm.c:1876 does */
        if (i < n - 1) {
            for (k = 0; k < cnt[i]; k++) {
                prm->N14pairlist[idx] = lst[i][k];
                prm->N14sceelist[idx] = prm->Scee[typ[i][k]];
                prm->N14scnblist[idx++] = prm->Scnb[typ[i][k]];
            }
        }
        free(lst[i]);
        free(typ[i]);
    }
    free(cnt);
    free(cap);
    free(lst);
    free(typ);
}

PARMSTRUCT_T *rdparm(char *name)
{
    prmfile_t pf;
    PARMSTRUCT_T *prm;
    double ptr[32];
    section_t *s;
    int i, rc = 0, n;

    if (prmfile_load(name, &pf) != 0) return NULL;
    prm = (PARMSTRUCT_T *)calloc(1, sizeof(PARMSTRUCT_T));

    s = prm_find(&pf, "TITLE");
    if (s) {
        int len = (int)(s->end - s->data);
        char *eol = strchr(s->data, '\n');
        if (eol && eol < s->end) len = (int)(eol - s->data);
        if (len > 80) len = 80;
        memcpy(prm->ititl, s->data, (size_t)len);
        prm->ititl[len] = '\0';
    }
    s = prm_find(&pf, "POINTERS");
    memset(ptr, 0, sizeof(ptr));
    if (!s || sec_numbers(s, 30, ptr) < 30) {
        nabc_seterr("rdparm: bad or missing POINTERS in %s", name);
        prmfile_free(&pf);
        free(prm);
        return NULL;
    }
    prm->Natom = (int)ptr[0];  prm->Ntypes = (int)ptr[1];
    prm->Nbonh = (int)ptr[2];  prm->Mbona = (int)ptr[3];
    prm->Ntheth = (int)ptr[4]; prm->Mtheta = (int)ptr[5];
    prm->Nphih = (int)ptr[6];  prm->Mphia = (int)ptr[7];
    prm->Nhparm = (int)ptr[8]; prm->Nparm = (int)ptr[9];
    prm->Nnb = (int)ptr[10];   prm->Nres = (int)ptr[11];
    prm->Nbona = (int)ptr[12]; prm->Ntheta = (int)ptr[13];
    prm->Nphia = (int)ptr[14]; prm->Numbnd = (int)ptr[15];
    prm->Numang = (int)ptr[16]; prm->Nptra = (int)ptr[17];
    prm->Natyp = (int)ptr[18]; prm->Nphb = (int)ptr[19];
    prm->IfBox = (int)ptr[27]; prm->Nmxrs = (int)ptr[28];
    prm->IfCap = (int)ptr[29];
    prm->Ntype2d = prm->Ntypes * prm->Ntypes;
    prm->Nttyp = prm->Ntypes * (prm->Ntypes + 1) / 2;
    n = prm->Natom;

    prm->AtomNames = (char (*)[NABC_NAMESIZE])calloc((size_t)n, NABC_NAMESIZE);
    s = prm_find(&pf, "ATOM_NAME");
    if (s) {
        char *p = s->data;
        int k = 0;
        while (k < n && p < s->end) {
            char *eol = strchr(p, '\n');
            char *q = p;
            while (k < n && q + 1 <= eol) {
                int len = (int)(eol - q) < 4 ? (int)(eol - q) : 4;
                memset(prm->AtomNames[k], ' ', 5);
                memcpy(prm->AtomNames[k], q, (size_t)len);
                prm->AtomNames[k][5] = '\0';
                prm->AtomNames[k][4] = '\0';
                k++;
                q += 4;
            }
            p = eol + 1;
        }
    }

    rc |= read_dbl(&pf, "CHARGE", n, &prm->Charges, 1) < 0;
    rc |= read_dbl(&pf, "MASS", n, &prm->Masses, 1) < 0;
    rc |= read_int(&pf, "ATOM_TYPE_INDEX", n, &prm->Iac, 1) < 0;
    rc |= read_int(&pf, "NUMBER_EXCLUDED_ATOMS", n, &prm->Iblo, 1) < 0;
    rc |= read_int(&pf, "NONBONDED_PARM_INDEX", prm->Ntype2d, &prm->Cno, 1) < 0;
    rc |= read_int(&pf, "RESIDUE_POINTER", prm->Nres, &prm->Ipres, 1) < 0;
    rc |= read_dbl(&pf, "BOND_FORCE_CONSTANT", prm->Numbnd, &prm->Rk, 1) < 0;
    rc |= read_dbl(&pf, "BOND_EQUIL_VALUE", prm->Numbnd, &prm->Req, 1) < 0;
    rc |= read_dbl(&pf, "ANGLE_FORCE_CONSTANT", prm->Numang, &prm->Tk, 1) < 0;
    rc |= read_dbl(&pf, "ANGLE_EQUIL_VALUE", prm->Numang, &prm->Teq, 1) < 0;
    rc |= read_dbl(&pf, "DIHEDRAL_FORCE_CONSTANT", prm->Nptra, &prm->Pk, 1) < 0;
    rc |= read_dbl(&pf, "DIHEDRAL_PERIODICITY", prm->Nptra, &prm->Pn, 1) < 0;
    rc |= read_dbl(&pf, "DIHEDRAL_PHASE", prm->Nptra, &prm->Phase, 1) < 0;
    if (rc) goto fail;

    /* SCEE / SCNB: stored inverted (prm.c:1074-1114) */
    if (read_dbl(&pf, "SCEE_SCALE_FACTOR", prm->Nptra, &prm->Scee, 0) == 0) {
        for (i = 0; i < prm->Nptra; i++) prm->Scee[i] = 1.0 / prm->Scee[i];
    } else {
        for (i = 0; i < prm->Nptra; i++) prm->Scee[i] = 1.0 / 1.2;
    }
    if (read_dbl(&pf, "SCNB_SCALE_FACTOR", prm->Nptra, &prm->Scnb, 0) == 0) {
        for (i = 0; i < prm->Nptra; i++) prm->Scnb[i] = 1.0 / prm->Scnb[i];
    } else {
        for (i = 0; i < prm->Nptra; i++) prm->Scnb[i] = 1.0 / 2.0;
    }
    if (read_int(&pf, "ATOMIC_NUMBER", n, &prm->AtomicNum, 0) != 0) {
        for (i = 0; i < n; i++)
            prm->AtomicNum[i] = nabc_atomic_number_from_mass(prm->Masses[i]);
    }

    rc |= read_dbl(&pf, "LENNARD_JONES_ACOEF", prm->Nttyp, &prm->Cn1, 1) < 0;
    rc |= read_dbl(&pf, "LENNARD_JONES_BCOEF", prm->Nttyp, &prm->Cn2, 1) < 0;
    {
        double *t;
        int k;
        rc |= read_dbl(&pf, "BONDS_INC_HYDROGEN", 3 * prm->Nbonh, &t, 1) < 0;
        prm->BondHAt1 = (int *)calloc((size_t)prm->Nbonh + 1, sizeof(int));
        prm->BondHAt2 = (int *)calloc((size_t)prm->Nbonh + 1, sizeof(int));
        prm->BondHNum = (int *)calloc((size_t)prm->Nbonh + 1, sizeof(int));
        for (k = 0; k < prm->Nbonh; k++) {
            prm->BondHAt1[k] = (int)t[3 * k];
            prm->BondHAt2[k] = (int)t[3 * k + 1];
            prm->BondHNum[k] = (int)t[3 * k + 2];
        }
        free(t);
        rc |= read_dbl(&pf, "BONDS_WITHOUT_HYDROGEN", 3 * prm->Mbona, &t, 1) < 0;
        prm->BondAt1 = (int *)calloc((size_t)prm->Mbona + 1, sizeof(int));
        prm->BondAt2 = (int *)calloc((size_t)prm->Mbona + 1, sizeof(int));
        prm->BondNum = (int *)calloc((size_t)prm->Mbona + 1, sizeof(int));
        for (k = 0; k < prm->Mbona; k++) {
            prm->BondAt1[k] = (int)t[3 * k];
            prm->BondAt2[k] = (int)t[3 * k + 1];
            prm->BondNum[k] = (int)t[3 * k + 2];
        }
        free(t);
        rc |= read_dbl(&pf, "ANGLES_INC_HYDROGEN", 4 * prm->Ntheth, &t, 1) < 0;
        prm->AngleHAt1 = (int *)calloc((size_t)prm->Ntheth + 1, sizeof(int));
        prm->AngleHAt2 = (int *)calloc((size_t)prm->Ntheth + 1, sizeof(int));
        prm->AngleHAt3 = (int *)calloc((size_t)prm->Ntheth + 1, sizeof(int));
        prm->AngleHNum = (int *)calloc((size_t)prm->Ntheth + 1, sizeof(int));
        for (k = 0; k < prm->Ntheth; k++) {
            prm->AngleHAt1[k] = (int)t[4 * k];
            prm->AngleHAt2[k] = (int)t[4 * k + 1];
            prm->AngleHAt3[k] = (int)t[4 * k + 2];
            prm->AngleHNum[k] = (int)t[4 * k + 3];
        }
        free(t);
        rc |= read_dbl(&pf, "ANGLES_WITHOUT_HYDROGEN", 4 * prm->Mtheta, &t, 1) < 0;
        prm->AngleAt1 = (int *)calloc((size_t)prm->Mtheta + 1, sizeof(int));
        prm->AngleAt2 = (int *)calloc((size_t)prm->Mtheta + 1, sizeof(int));
        prm->AngleAt3 = (int *)calloc((size_t)prm->Mtheta + 1, sizeof(int));
        prm->AngleNum = (int *)calloc((size_t)prm->Mtheta + 1, sizeof(int));
        for (k = 0; k < prm->Mtheta; k++) {
            prm->AngleAt1[k] = (int)t[4 * k];
            prm->AngleAt2[k] = (int)t[4 * k + 1];
            prm->AngleAt3[k] = (int)t[4 * k + 2];
            prm->AngleNum[k] = (int)t[4 * k + 3];
        }
        free(t);
        rc |= read_dbl(&pf, "DIHEDRALS_INC_HYDROGEN", 5 * prm->Nphih, &t, 1) < 0;
        prm->DihHAt1 = (int *)calloc((size_t)prm->Nphih + 1, sizeof(int));
        prm->DihHAt2 = (int *)calloc((size_t)prm->Nphih + 1, sizeof(int));
        prm->DihHAt3 = (int *)calloc((size_t)prm->Nphih + 1, sizeof(int));
        prm->DihHAt4 = (int *)calloc((size_t)prm->Nphih + 1, sizeof(int));
        prm->DihHNum = (int *)calloc((size_t)prm->Nphih + 1, sizeof(int));
        for (k = 0; k < prm->Nphih; k++) {
            prm->DihHAt1[k] = (int)t[5 * k];
            prm->DihHAt2[k] = (int)t[5 * k + 1];
            prm->DihHAt3[k] = (int)t[5 * k + 2];
            prm->DihHAt4[k] = (int)t[5 * k + 3];
            prm->DihHNum[k] = (int)t[5 * k + 4];
        }
        free(t);
        rc |= read_dbl(&pf, "DIHEDRALS_WITHOUT_HYDROGEN", 5 * prm->Mphia, &t, 1) < 0;
        prm->DihAt1 = (int *)calloc((size_t)prm->Mphia + 1, sizeof(int));
        prm->DihAt2 = (int *)calloc((size_t)prm->Mphia + 1, sizeof(int));
        prm->DihAt3 = (int *)calloc((size_t)prm->Mphia + 1, sizeof(int));
        prm->DihAt4 = (int *)calloc((size_t)prm->Mphia + 1, sizeof(int));
        prm->DihNum = (int *)calloc((size_t)prm->Mphia + 1, sizeof(int));
        for (k = 0; k < prm->Mphia; k++) {
            prm->DihAt1[k] = (int)t[5 * k];
            prm->DihAt2[k] = (int)t[5 * k + 1];
            prm->DihAt3[k] = (int)t[5 * k + 2];
            prm->DihAt4[k] = (int)t[5 * k + 3];
            prm->DihNum[k] = (int)t[5 * k + 4];
        }
        free(t);
    }
    rc |= read_int(&pf, "EXCLUDED_ATOMS_LIST", prm->Nnb, &prm->ExclAt, 1) < 0;
    read_dbl(&pf, "HBOND_ACOEF", prm->Nphb, &prm->HB12, 0);
    read_dbl(&pf, "HBOND_BCOEF", prm->Nphb, &prm->HB10, 0);
    if (rc) goto fail;

    if (prm->IfBox) {
        /* GB runs are non-periodic only (sff.c:890-899); the box sections
         * are not needed for that and are not read. */
    }
    rc |= read_dbl(&pf, "RADII", n, &prm->Rborn, 1) < 0;
    rc |= read_dbl(&pf, "SCREEN", n, &prm->Fs, 1) < 0;
    if (rc) goto fail;
    prm->Fsmax = 0.0;
    for (i = 0; i < n; i++) {
        double si = prm->Fs[i] * (prm->Rborn[i] - 0.09);
        if (si > prm->Fsmax) prm->Fsmax = si;
    }
    nabc_build_14_list(prm);
    prmfile_free(&pf);
    return prm;

fail:
    prmfile_free(&pf);
    freeparm(prm);
    return NULL;
}

void freeparm(PARMSTRUCT_T *p)
{
    if (!p) return;
    free(p->AtomNames);
    free(p->Charges); free(p->Masses); free(p->Rk); free(p->Req);
    free(p->Tk); free(p->Teq); free(p->Pk); free(p->Pn); free(p->Phase);
    free(p->Cn1); free(p->Cn2); free(p->HB12); free(p->HB10);
    free(p->Rborn); free(p->Fs); free(p->Scee); free(p->Scnb);
    free(p->Iac); free(p->Iblo); free(p->Cno); free(p->Ipres); free(p->ExclAt);
    free(p->BondHAt1); free(p->BondHAt2); free(p->BondHNum);
    free(p->BondAt1); free(p->BondAt2); free(p->BondNum);
    free(p->AngleHAt1); free(p->AngleHAt2); free(p->AngleHAt3); free(p->AngleHNum);
    free(p->AngleAt1); free(p->AngleAt2); free(p->AngleAt3); free(p->AngleNum);
    free(p->DihHAt1); free(p->DihHAt2); free(p->DihHAt3); free(p->DihHAt4); free(p->DihHNum);
    free(p->DihAt1); free(p->DihAt2); free(p->DihAt3); free(p->DihAt4); free(p->DihNum);
    free(p->AtomicNum);
    free(p->N14pairs); free(p->N14pairlist);
    free(p->N14sceelist); free(p->N14scnblist);
    free(p);
}

/* ---------------------------------------------------------------- writer */

static void w_flag(FILE *fp, const char *flag, const char *fmt)
{
    fprintf(fp, "%%FLAG %-74s\n%%FORMAT(%s)%*s\n", flag, fmt, (int)(72 - strlen(fmt)), "");
}

static void w_ints(FILE *fp, const char *flag, const int *v, int n)
{
    int i;
    w_flag(fp, flag, "10I8");
    for (i = 0; i < n; i++) {
        fprintf(fp, "%8d", v[i]);
        if (i % 10 == 9 || i == n - 1) fputc('\n', fp);
    }
    if (n == 0) fputc('\n', fp);
}

static void w_dbls(FILE *fp, const char *flag, const double *v, int n)
{
    int i;
    w_flag(fp, flag, "5E16.8");
    for (i = 0; i < n; i++) {
        fprintf(fp, "%16.8E", v[i]);
        if (i % 5 == 4 || i == n - 1) fputc('\n', fp);
    }
    if (n == 0) fputc('\n', fp);
}

static void w_names(FILE *fp, const char *flag, char (*names)[NABC_NAMESIZE], int n)
{
    int i;
    w_flag(fp, flag, "20a4");
    for (i = 0; i < n; i++) {
        fprintf(fp, "%-4.4s", names[i]);
        if (i % 20 == 19 || i == n - 1) fputc('\n', fp);
    }
    if (n == 0) fputc('\n', fp);
}

/*
 * Write a %FLAG prmtop holding every section rdparm() of prm.c requires,
 * so the reference can read the generated systems too.  Scee/Scnb are
 * written back un-inverted.
 */
int wrparm(PARMSTRUCT_T *p, char *name)
{
    FILE *fp = fopen(name, "w");
    int ptr[31], i, n = p->Natom;
    int *tmp;
    double *dt;
    int ntmp;

    if (!fp) {
        nabc_seterr("wrparm: cannot open %s", name);
        return -1;
    }
    fprintf(fp, "%%VERSION  VERSION_STAMP = V0001.000  DATE = 01/01/26  00:00:00%17s\n", "");
    w_flag(fp, "TITLE", "20a4");
    fprintf(fp, "%-80s\n", p->ititl);
    memset(ptr, 0, sizeof(ptr));
    ptr[0] = n; ptr[1] = p->Ntypes; ptr[2] = p->Nbonh; ptr[3] = p->Mbona;
    ptr[4] = p->Ntheth; ptr[5] = p->Mtheta; ptr[6] = p->Nphih; ptr[7] = p->Mphia;
    ptr[8] = p->Nhparm; ptr[9] = p->Nparm; ptr[10] = p->Nnb; ptr[11] = p->Nres;
    ptr[12] = p->Nbona; ptr[13] = p->Ntheta; ptr[14] = p->Nphia; ptr[15] = p->Numbnd;
    ptr[16] = p->Numang; ptr[17] = p->Nptra; ptr[18] = p->Natyp; ptr[19] = p->Nphb;
    ptr[27] = p->IfBox; ptr[28] = p->Nmxrs; ptr[29] = p->IfCap;
    w_ints(fp, "POINTERS", ptr, 31);
    w_names(fp, "ATOM_NAME", p->AtomNames, n);
    w_dbls(fp, "CHARGE", p->Charges, n);
    w_ints(fp, "ATOMIC_NUMBER", p->AtomicNum, n);
    w_dbls(fp, "MASS", p->Masses, n);
    w_ints(fp, "ATOM_TYPE_INDEX", p->Iac, n);
    w_ints(fp, "NUMBER_EXCLUDED_ATOMS", p->Iblo, n);
    w_ints(fp, "NONBONDED_PARM_INDEX", p->Cno, p->Ntypes * p->Ntypes);
    {
        char (*rl)[NABC_NAMESIZE] = (char (*)[NABC_NAMESIZE])calloc((size_t)p->Nres + 1, NABC_NAMESIZE);
        for (i = 0; i < p->Nres; i++) strcpy(rl[i], "UNK ");
        w_names(fp, "RESIDUE_LABEL", rl, p->Nres);
        free(rl);
    }
    w_ints(fp, "RESIDUE_POINTER", p->Ipres, p->Nres);
    w_dbls(fp, "BOND_FORCE_CONSTANT", p->Rk, p->Numbnd);
    w_dbls(fp, "BOND_EQUIL_VALUE", p->Req, p->Numbnd);
    w_dbls(fp, "ANGLE_FORCE_CONSTANT", p->Tk, p->Numang);
    w_dbls(fp, "ANGLE_EQUIL_VALUE", p->Teq, p->Numang);
    w_dbls(fp, "DIHEDRAL_FORCE_CONSTANT", p->Pk, p->Nptra);
    w_dbls(fp, "DIHEDRAL_PERIODICITY", p->Pn, p->Nptra);
    w_dbls(fp, "DIHEDRAL_PHASE", p->Phase, p->Nptra);
    dt = (double *)malloc(sizeof(double) * (size_t)(p->Nptra + 1));
    for (i = 0; i < p->Nptra; i++) dt[i] = 1.0 / p->Scee[i];
    w_dbls(fp, "SCEE_SCALE_FACTOR", dt, p->Nptra);
    for (i = 0; i < p->Nptra; i++) dt[i] = 1.0 / p->Scnb[i];
    w_dbls(fp, "SCNB_SCALE_FACTOR", dt, p->Nptra);
    free(dt);
    dt = (double *)calloc((size_t)(p->Natyp + 1), sizeof(double));
    w_dbls(fp, "SOLTY", dt, p->Natyp);
    free(dt);
    w_dbls(fp, "LENNARD_JONES_ACOEF", p->Cn1, p->Nttyp);
    w_dbls(fp, "LENNARD_JONES_BCOEF", p->Cn2, p->Nttyp);

    ntmp = 5 * (p->Nbonh + p->Mbona + p->Ntheth + p->Mtheta + p->Nphih + p->Mphia) + 8;
    tmp = (int *)malloc(sizeof(int) * (size_t)ntmp);
    for (i = 0; i < p->Nbonh; i++) {
        tmp[3 * i] = p->BondHAt1[i]; tmp[3 * i + 1] = p->BondHAt2[i]; tmp[3 * i + 2] = p->BondHNum[i];
    }
    w_ints(fp, "BONDS_INC_HYDROGEN", tmp, 3 * p->Nbonh);
    for (i = 0; i < p->Mbona; i++) {
        tmp[3 * i] = p->BondAt1[i]; tmp[3 * i + 1] = p->BondAt2[i]; tmp[3 * i + 2] = p->BondNum[i];
    }
    w_ints(fp, "BONDS_WITHOUT_HYDROGEN", tmp, 3 * p->Mbona);
    for (i = 0; i < p->Ntheth; i++) {
        tmp[4 * i] = p->AngleHAt1[i]; tmp[4 * i + 1] = p->AngleHAt2[i];
        tmp[4 * i + 2] = p->AngleHAt3[i]; tmp[4 * i + 3] = p->AngleHNum[i];
    }
    w_ints(fp, "ANGLES_INC_HYDROGEN", tmp, 4 * p->Ntheth);
    for (i = 0; i < p->Mtheta; i++) {
        tmp[4 * i] = p->AngleAt1[i]; tmp[4 * i + 1] = p->AngleAt2[i];
        tmp[4 * i + 2] = p->AngleAt3[i]; tmp[4 * i + 3] = p->AngleNum[i];
    }
    w_ints(fp, "ANGLES_WITHOUT_HYDROGEN", tmp, 4 * p->Mtheta);
    for (i = 0; i < p->Nphih; i++) {
        tmp[5 * i] = p->DihHAt1[i]; tmp[5 * i + 1] = p->DihHAt2[i]; tmp[5 * i + 2] = p->DihHAt3[i];
        tmp[5 * i + 3] = p->DihHAt4[i]; tmp[5 * i + 4] = p->DihHNum[i];
    }
    w_ints(fp, "DIHEDRALS_INC_HYDROGEN", tmp, 5 * p->Nphih);
    for (i = 0; i < p->Mphia; i++) {
        tmp[5 * i] = p->DihAt1[i]; tmp[5 * i + 1] = p->DihAt2[i]; tmp[5 * i + 2] = p->DihAt3[i];
        tmp[5 * i + 3] = p->DihAt4[i]; tmp[5 * i + 4] = p->DihNum[i];
    }
    w_ints(fp, "DIHEDRALS_WITHOUT_HYDROGEN", tmp, 5 * p->Mphia);
    free(tmp);
    w_ints(fp, "EXCLUDED_ATOMS_LIST", p->ExclAt, p->Nnb);
    w_dbls(fp, "HBOND_ACOEF", p->HB12, p->Nphb);
    w_dbls(fp, "HBOND_BCOEF", p->HB10, p->Nphb);
    w_dbls(fp, "HBCUT", p->HB10, p->Nphb);
    w_names(fp, "AMBER_ATOM_TYPE", p->AtomNames, n);
    {
        char (*tc)[NABC_NAMESIZE] = (char (*)[NABC_NAMESIZE])calloc((size_t)n + 1, NABC_NAMESIZE);
        for (i = 0; i < n; i++) strcpy(tc[i], "M   ");
        w_names(fp, "TREE_CHAIN_CLASSIFICATION", tc, n);
        free(tc);
    }
    tmp = (int *)calloc((size_t)n + 1, sizeof(int));
    w_ints(fp, "JOIN_ARRAY", tmp, n);
    w_ints(fp, "IROTAT", tmp, n);
    free(tmp);
    w_flag(fp, "RADIUS_SET", "1a80");
    fprintf(fp, "%-80s\n", "synthetic intrinsic radii");
    w_dbls(fp, "RADII", p->Rborn, n);
    w_dbls(fp, "SCREEN", p->Fs, n);
    fclose(fp);
    return 0;
}
