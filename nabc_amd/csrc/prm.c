/*
 * prm.c -- Amber prmtop reader and writer for nabc_amd.
 *
 * Reads the two layouts rdparm() of src/sff/prm.c:584-1911 reads: the
 * %FLAG/%FORMAT ("new") layout, whose sections are found by name, and the
 * older layout without flags, whose sections follow one another in a fixed
 * order (prm.c:617-679: the same order rdparm reads them in), each starting
 * on a new line and holding ceil(count / per-line) lines (one blank line
 * when empty).  The same post-processing as the reference:
 *   - charges stay in the file's internal unit (e * 18.2223), prm.c:777-792;
 *   - residue pointers are stored 0-based with Ipres[Nres] = Natom, and
 *     AtomRes holds each atom's residue, prm.c:898-912, :1513-1520;
 *   - SCEE/SCNB are stored inverted, defaults 1/1.2 and 1/2, prm.c:1074-1144;
 *   - ATOMIC_NUMBER is read or, if absent, taken as the element whose
 *     standard atomic weight is nearest to the mass, prm.c:85-99, :1122-1133;
 *   - Fsmax = max_i Fs_i * (Rborn_i - 0.09) from the file's SCREEN values,
 *     prm.c:1621-1656 (gb 7/8 later replace Fs but not Fsmax); an old-layout
 *     file has no RADII/SCREEN and takes them from the first letter of each
 *     atom name (Bondi radii, hydrogens by the atom they follow), prm.c:1656-1741;
 *   - the 1-4 pair list is built from the dihedrals in file order, skipping
 *     entries whose third or fourth atom index is negative, prm.c:1854-1889.
 * Numeric fields are parsed by the width their format states (%FORMAT line,
 * or 12I6 / 5E16.8 / 20a4 in the old layout).
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
    int nsec, cap;
    int old;            /* no %FLAG lines: fixed section order */
    char *cursor;       /* old layout: start of the next section */
} prmfile_t;

static void prm_add(prmfile_t *pf, section_t s)
{
    if (pf->nsec == pf->cap) {
        pf->cap *= 2;
        pf->sec = (section_t *)realloc(pf->sec, sizeof(section_t) * pf->cap);
    }
    pf->sec[pf->nsec++] = s;
}

/*
 * Old layout: the next section holds `count` values of format fmt ("20a4",
 * "12I6" or "5E16.8", per_line values a line), starting on the next line:
 * ceil(count / per_line) lines, or one blank line when count is 0.
 */
static int old_section(prmfile_t *pf, const char *name, const char *fmt, int per_line, long long count)
{
    section_t s;
    long long lines = count > 0 ? (count + per_line - 1) / per_line : 1, k;
    char *p = pf->cursor;
    s.name = name;
    s.fmt = (char *)fmt;
    s.data = p;
    for (k = 0; k < lines; k++) {
        char *eol = strchr(p, '\n');
        if (!eol || eol >= pf->buf + pf->len + 1) {
            nabc_seterr("rdparm: old-format file ends inside section %s", name);
            return -1;
        }
        p = eol + 1;
    }
    s.end = p;
    pf->cursor = p;
    prm_add(pf, s);
    return 0;
}

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
    pf->old = strncmp(pf->buf, "%VERSION", 8) != 0 && strncmp(pf->buf, "%FLAG", 5) != 0;
    pf->sec = (section_t *)malloc(sizeof(section_t) * cap);
    pf->cap = cap;
    if (pf->old) return 0; /* sections are laid out once POINTERS is known */
    for (p = pf->buf; *p;) {
        char *eol = strchr(p, '\n');
        if (!eol) break;
        if (strncmp(p, "%FLAG ", 6) == 0) {
            section_t s;
            char *q = p + 6, *e;
            while (*q == ' ') q++;
            /* the whole label: HCP files have "POINTERS FOR STRAND AND COMPLEX" */
            e = eol;
            while (e > q && isspace((unsigned char)e[-1])) e--;
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
            prm_add(pf, s);
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

/* names of a 20a4 section: four characters each, then a blank and the NUL
 * (NAME_T of sff.h; prm.c:766-775 reads them the same way) */
static NABC_NAME_T *read_names(prmfile_t *pf, const char *name, int n)
{
    section_t *s = prm_find(pf, name);
    NABC_NAME_T *out = (NABC_NAME_T *)calloc((size_t)(n > 0 ? n : 1), sizeof(NABC_NAME_T));
    int k;
    for (k = 0; k < n; k++) memcpy(out[k], "     ", NABC_NAMESIZE);
    if (!s) return out;
    {
        char *p = s->data;
        k = 0;
        while (k < n && p < s->end) {
            char *eol = strchr(p, '\n');
            char *q = p;
            if (!eol || eol > s->end) eol = s->end;
            while (k < n && q < eol) {
                int len = (int)(eol - q) < 4 ? (int)(eol - q) : 4;
                memcpy(out[k], q, (size_t)len);
                k++;
                q += 4;
            }
            p = eol + 1;
        }
    }
    return out;
}

/* sections of an old-layout file after POINTERS, in file order (the order
 * rdparm reads them, prm.c:766-1610) */
static int old_layout(prmfile_t *pf, const PARMSTRUCT_T *p, char *after_ptr)
{
    const char *I = "12I6", *E = "5E16.8", *A = "20a4";
    int rc = 0;
    pf->cursor = after_ptr;
    rc |= old_section(pf, "ATOM_NAME", A, 20, p->Natom);
    rc |= old_section(pf, "CHARGE", E, 5, p->Natom);
    rc |= old_section(pf, "MASS", E, 5, p->Natom);
    rc |= old_section(pf, "ATOM_TYPE_INDEX", I, 12, p->Natom);
    rc |= old_section(pf, "NUMBER_EXCLUDED_ATOMS", I, 12, p->Natom);
    rc |= old_section(pf, "NONBONDED_PARM_INDEX", I, 12, (long long)p->Ntypes * p->Ntypes);
    rc |= old_section(pf, "RESIDUE_LABEL", A, 20, p->Nres);
    rc |= old_section(pf, "RESIDUE_POINTER", I, 12, p->Nres);
    rc |= old_section(pf, "BOND_FORCE_CONSTANT", E, 5, p->Numbnd);
    rc |= old_section(pf, "BOND_EQUIL_VALUE", E, 5, p->Numbnd);
    rc |= old_section(pf, "ANGLE_FORCE_CONSTANT", E, 5, p->Numang);
    rc |= old_section(pf, "ANGLE_EQUIL_VALUE", E, 5, p->Numang);
    rc |= old_section(pf, "DIHEDRAL_FORCE_CONSTANT", E, 5, p->Nptra);
    rc |= old_section(pf, "DIHEDRAL_PERIODICITY", E, 5, p->Nptra);
    rc |= old_section(pf, "DIHEDRAL_PHASE", E, 5, p->Nptra);
    rc |= old_section(pf, "SOLTY", E, 5, p->Natyp);
    rc |= old_section(pf, "LENNARD_JONES_ACOEF", E, 5, p->Nttyp);
    rc |= old_section(pf, "LENNARD_JONES_BCOEF", E, 5, p->Nttyp);
    rc |= old_section(pf, "BONDS_INC_HYDROGEN", I, 12, 3LL * p->Nbonh);
    rc |= old_section(pf, "BONDS_WITHOUT_HYDROGEN", I, 12, 3LL * p->Nbona);
    rc |= old_section(pf, "ANGLES_INC_HYDROGEN", I, 12, 4LL * p->Ntheth);
    rc |= old_section(pf, "ANGLES_WITHOUT_HYDROGEN", I, 12, 4LL * p->Ntheta);
    rc |= old_section(pf, "DIHEDRALS_INC_HYDROGEN", I, 12, 5LL * p->Nphih);
    rc |= old_section(pf, "DIHEDRALS_WITHOUT_HYDROGEN", I, 12, 5LL * p->Nphia);
    rc |= old_section(pf, "EXCLUDED_ATOMS_LIST", I, 12, p->Nnb);
    rc |= old_section(pf, "HBOND_ACOEF", E, 5, p->Nphb);
    rc |= old_section(pf, "HBOND_BCOEF", E, 5, p->Nphb);
    rc |= old_section(pf, "HBCUT", E, 5, p->Nphb);
    rc |= old_section(pf, "AMBER_ATOM_TYPE", A, 20, p->Natom);
    rc |= old_section(pf, "TREE_CHAIN_CLASSIFICATION", A, 20, p->Natom);
    rc |= old_section(pf, "JOIN_ARRAY", I, 12, p->Natom);
    rc |= old_section(pf, "IROTAT", I, 12, p->Natom);
    if (rc) return -1;
    if (p->IfBox) {
        double sp[3];
        if (old_section(pf, "SOLVENT_POINTERS", I, 12, 3)) return -1;
        if (sec_numbers(prm_find(pf, "SOLVENT_POINTERS"), 3, sp) != 3) {
            nabc_seterr("rdparm: short SOLVENT_POINTERS");
            return -1;
        }
        if (old_section(pf, "ATOMS_PER_MOLECULE", I, 12, (long long)sp[1]) ||
            old_section(pf, "BOX_DIMENSIONS", E, 5, 4))
            return -1;
    }
    if (p->IfCap && (old_section(pf, "CAP_INFO", I, 12, 1) || old_section(pf, "CAP_INFO2", E, 5, 4))) return -1;
    return 0;
}

/*
 * Intrinsic radius and screening factor of an old-layout file's atom from
 * the first letter of its name (prm.c:1656-1741): Bondi radii with the
 * screening factors of Tinker's pairwise method; a hydrogen takes the
 * radius for the element of the nearest preceding non-hydrogen atom within
 * three (oxygen when it is the first atom).  Returns -1 for a letter
 * outside the table.
 */
static int old_gb_radius(const NABC_NAME_T *names, int i, double *rb, double *fs)
{
    static const struct { char c; double rb, fs; } tab[] = {
        {'C', 1.70, 0.72}, {'N', 1.55, 0.79}, {'O', 1.50, 0.85}, {'F', 1.47, 0.88},
        {'P', 1.85, 0.86}, {'S', 1.80, 0.96}, {'L', 1.00, 0.96}, {'Z', 1.40, 0.96},
        {'M', 1.40, 0.96}};
    const char c = names[i][0];
    int k;
    if (c == 'H') {
        char prev = 'O';
        if (i > 0) {
            for (k = 1; k <= 3; k++) {
                prev = names[i - k][0];
                if (prev != 'H') break;
            }
        }
        *fs = 0.85;
        *rb = prev == 'O' ? 0.8 : prev == 'C' ? 1.3 : 1.2; /* N and the rest: 1.2 */
        return 0;
    }
    for (k = 0; k < (int)(sizeof(tab) / sizeof(tab[0])); k++)
        if (tab[k].c == c) {
            *rb = tab[k].rb;
            *fs = tab[k].fs;
            return 0;
        }
    return -1;
}

/*
 * Fields the reference derives after reading (and those a topology built
 * in memory may leave unset): 3N, residue of each atom from Ipres,
 * boundaries and box of a non-periodic system, HCP strand/complex defaults,
 * the zero Gvdw array and default names.
 */
void nabc_parm_complete(PARMSTRUCT_T *p)
{
    int i, n = p->Natom, res;
    p->Nat3 = 3 * n;
    p->Ntype2d = p->Ntypes * p->Ntypes;
    p->Nttyp = p->Ntypes * (p->Ntypes + 1) / 2;
    if (!p->ResNames) {
        p->ResNames = (NABC_NAME_T *)calloc((size_t)p->Nres + 1, sizeof(NABC_NAME_T));
        for (i = 0; i < p->Nres; i++) memcpy(p->ResNames[i], "UNK  ", NABC_NAMESIZE);
    }
    if (!p->AtomType) {
        p->AtomType = (NABC_NAME_T *)calloc((size_t)n + 1, sizeof(NABC_NAME_T));
        for (i = 0; i < n; i++) memcpy(p->AtomType[i], p->AtomNames[i], NABC_NAMESIZE);
    }
    if (!p->AtomTree) {
        p->AtomTree = (NABC_NAME_T *)calloc((size_t)n + 1, sizeof(NABC_NAME_T));
        for (i = 0; i < n; i++) memcpy(p->AtomTree[i], "M    ", NABC_NAMESIZE);
    }
    if (!p->Solty) p->Solty = (double *)calloc((size_t)p->Natyp + 1, sizeof(double));
    if (!p->TreeJoin) p->TreeJoin = (int *)calloc((size_t)n + 1, sizeof(int));
    if (!p->AtomRes) p->AtomRes = (int *)calloc((size_t)n + 1, sizeof(int));
    for (i = 0, res = 0; i < n; i++) {
        while (res + 1 < p->Nres && i >= p->Ipres[res + 1]) res++;
        p->AtomRes[i] = res;
    }
    if (!p->IfBox && !p->Boundary) {
        p->Nspm = 1;
        p->Boundary = (int *)calloc(1, sizeof(int));
        p->Boundary[0] = n;
        p->Box[0] = p->Box[1] = p->Box[2] = 999999999.;
    }
    if (!p->Gvdw) p->Gvdw = (double *)calloc((size_t)n + 1, sizeof(double));
    if (!p->Ipstrand) {
        p->Nstrand = p->Ncomplex = 1;
        p->Ipstrand = (int *)calloc(2, sizeof(int));
        p->Ipcomplex = (int *)calloc(2, sizeof(int));
        p->Ipstrand[0] = p->Ipcomplex[0] = 1;
    }
}

/* per-term index arrays of a bond/angle/dihedral section of k columns */
static int read_terms(prmfile_t *pf, const char *name, int nterm, int k, int **c0, int **c1, int **c2, int **c3,
                      int **c4)
{
    double *t;
    int **cols[5] = {c0, c1, c2, c3, c4}, j, m;
    int rc = read_dbl(pf, name, k * nterm, &t, 1) < 0;
    for (j = 0; j < k; j++) {
        *cols[j] = (int *)calloc((size_t)nterm + 1, sizeof(int));
        if (!rc)
            for (m = 0; m < nterm; m++) (*cols[j])[m] = (int)t[k * m + j];
    }
    free(t);
    return rc;
}

PARMSTRUCT_T *rdparm(char *name)
{
    prmfile_t pf;
    PARMSTRUCT_T *prm;
    double ptr[32];
    section_t *s;
    int i, rc = 0, n;
    char *after_ptr = NULL;

    if (prmfile_load(name, &pf) != 0) return NULL;
    prm = (PARMSTRUCT_T *)calloc(1, sizeof(PARMSTRUCT_T));
    memset(ptr, 0, sizeof(ptr));
    if (pf.old) {
        /* title line, then 30 integers of POINTERS (12I6: three lines) */
        char *eol = strchr(pf.buf, '\n');
        int len = (int)(eol - pf.buf);
        section_t ps;
        if (len > 80) len = 80;
        memcpy(prm->ititl, pf.buf, (size_t)len);
        prm->ititl[len] = '\0';
        pf.cursor = eol + 1;
        if (old_section(&pf, "POINTERS", "12I6", 12, 30) != 0) goto fail_early;
        ps = pf.sec[pf.nsec - 1];
        if (sec_numbers(&ps, 30, ptr) < 30) {
            nabc_seterr("rdparm: bad POINTERS in %s", name);
            goto fail_early;
        }
        after_ptr = pf.cursor;
    } else {
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
        if (!s || sec_numbers(s, 31, ptr) < 30) {
            nabc_seterr("rdparm: bad or missing POINTERS in %s", name);
            goto fail_early;
        }
        prm->Numextra = (int)ptr[30]; /* read in the %FLAG layout only, prm.c:679-685 */
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
    prm->Nat3 = 3 * prm->Natom;
    prm->Ntype2d = prm->Ntypes * prm->Ntypes;
    prm->Nttyp = prm->Ntypes * (prm->Ntypes + 1) / 2;
    n = prm->Natom;
    if (pf.old && old_layout(&pf, prm, after_ptr) != 0) goto fail;

    prm->AtomNames = read_names(&pf, "ATOM_NAME", n);
    rc |= read_dbl(&pf, "CHARGE", n, &prm->Charges, 1) < 0;
    rc |= read_dbl(&pf, "MASS", n, &prm->Masses, 1) < 0;
    rc |= read_int(&pf, "ATOM_TYPE_INDEX", n, &prm->Iac, 1) < 0;
    rc |= read_int(&pf, "NUMBER_EXCLUDED_ATOMS", n, &prm->Iblo, 1) < 0;
    rc |= read_int(&pf, "NONBONDED_PARM_INDEX", prm->Ntype2d, &prm->Cno, 1) < 0;
    prm->ResNames = read_names(&pf, "RESIDUE_LABEL", prm->Nres);
    {
        /* 0-based first atoms, Ipres[Nres] = Natom (prm.c:898-912) */
        int *ip;
        rc |= read_int(&pf, "RESIDUE_POINTER", prm->Nres, &ip, 1) < 0;
        prm->Ipres = (int *)calloc((size_t)prm->Nres + 1, sizeof(int));
        for (i = 0; i < prm->Nres; i++) prm->Ipres[i] = ip[i] - 1;
        prm->Ipres[prm->Nres] = n;
        free(ip);
    }
    rc |= read_dbl(&pf, "BOND_FORCE_CONSTANT", prm->Numbnd, &prm->Rk, 1) < 0;
    rc |= read_dbl(&pf, "BOND_EQUIL_VALUE", prm->Numbnd, &prm->Req, 1) < 0;
    rc |= read_dbl(&pf, "ANGLE_FORCE_CONSTANT", prm->Numang, &prm->Tk, 1) < 0;
    rc |= read_dbl(&pf, "ANGLE_EQUIL_VALUE", prm->Numang, &prm->Teq, 1) < 0;
    rc |= read_dbl(&pf, "DIHEDRAL_FORCE_CONSTANT", prm->Nptra, &prm->Pk, 1) < 0;
    rc |= read_dbl(&pf, "DIHEDRAL_PERIODICITY", prm->Nptra, &prm->Pn, 1) < 0;
    rc |= read_dbl(&pf, "DIHEDRAL_PHASE", prm->Nptra, &prm->Phase, 1) < 0;
    if (rc) goto fail;

    /* SCEE / SCNB: stored inverted; an old-layout file has neither and
     * takes the defaults (prm.c:1070-1144) */
    if (!pf.old && read_dbl(&pf, "SCEE_SCALE_FACTOR", prm->Nptra, &prm->Scee, 0) == 0) {
        for (i = 0; i < prm->Nptra; i++) prm->Scee[i] = 1.0 / prm->Scee[i];
    } else {
        if (!prm->Scee) prm->Scee = (double *)calloc((size_t)prm->Nptra + 1, sizeof(double));
        for (i = 0; i < prm->Nptra; i++) prm->Scee[i] = 1.0 / 1.2;
    }
    if (!pf.old && read_dbl(&pf, "SCNB_SCALE_FACTOR", prm->Nptra, &prm->Scnb, 0) == 0) {
        for (i = 0; i < prm->Nptra; i++) prm->Scnb[i] = 1.0 / prm->Scnb[i];
    } else {
        if (!prm->Scnb) prm->Scnb = (double *)calloc((size_t)prm->Nptra + 1, sizeof(double));
        for (i = 0; i < prm->Nptra; i++) prm->Scnb[i] = 1.0 / 2.0;
    }
    /* atomic numbers: read, else from the masses (the reference leaves them
     * unset in an old-layout file; they only matter for gb 7/8) */
    if (pf.old || read_int(&pf, "ATOMIC_NUMBER", n, &prm->AtomicNum, 0) != 0) {
        if (!prm->AtomicNum) prm->AtomicNum = (int *)calloc((size_t)n + 1, sizeof(int));
        for (i = 0; i < n; i++)
            prm->AtomicNum[i] = nabc_atomic_number_from_mass(prm->Masses[i]);
    }

    read_dbl(&pf, "SOLTY", prm->Natyp, &prm->Solty, 0);
    rc |= read_dbl(&pf, "LENNARD_JONES_ACOEF", prm->Nttyp, &prm->Cn1, 1) < 0;
    rc |= read_dbl(&pf, "LENNARD_JONES_BCOEF", prm->Nttyp, &prm->Cn2, 1) < 0;
    /* the "without hydrogen" sections hold Nbona/Ntheta/Nphia terms
     * (prm.c:1242-1365); the energy terms use Mbona/Mtheta/Mphia of them */
    rc |= read_terms(&pf, "BONDS_INC_HYDROGEN", prm->Nbonh, 3, &prm->BondHAt1, &prm->BondHAt2, &prm->BondHNum,
                     NULL, NULL);
    rc |= read_terms(&pf, "BONDS_WITHOUT_HYDROGEN", prm->Nbona > prm->Mbona ? prm->Nbona : prm->Mbona, 3,
                     &prm->BondAt1, &prm->BondAt2, &prm->BondNum, NULL, NULL);
    rc |= read_terms(&pf, "ANGLES_INC_HYDROGEN", prm->Ntheth, 4, &prm->AngleHAt1, &prm->AngleHAt2,
                     &prm->AngleHAt3, &prm->AngleHNum, NULL);
    rc |= read_terms(&pf, "ANGLES_WITHOUT_HYDROGEN", prm->Ntheta > prm->Mtheta ? prm->Ntheta : prm->Mtheta, 4,
                     &prm->AngleAt1, &prm->AngleAt2, &prm->AngleAt3, &prm->AngleNum, NULL);
    rc |= read_terms(&pf, "DIHEDRALS_INC_HYDROGEN", prm->Nphih, 5, &prm->DihHAt1, &prm->DihHAt2, &prm->DihHAt3,
                     &prm->DihHAt4, &prm->DihHNum);
    rc |= read_terms(&pf, "DIHEDRALS_WITHOUT_HYDROGEN", prm->Nphia > prm->Mphia ? prm->Nphia : prm->Mphia, 5,
                     &prm->DihAt1, &prm->DihAt2, &prm->DihAt3, &prm->DihAt4, &prm->DihNum);
    rc |= read_int(&pf, "EXCLUDED_ATOMS_LIST", prm->Nnb, &prm->ExclAt, 1) < 0;
    read_dbl(&pf, "HBOND_ACOEF", prm->Nphb, &prm->HB12, 0);
    read_dbl(&pf, "HBOND_BCOEF", prm->Nphb, &prm->HB10, 0);
    if (rc) goto fail;
    prm->AtomType = read_names(&pf, "AMBER_ATOM_TYPE", n);
    prm->AtomTree = read_names(&pf, "TREE_CHAIN_CLASSIFICATION", n);
    read_int(&pf, "JOIN_ARRAY", n, &prm->TreeJoin, 0);

    if (prm->IfBox) {
        /* read for the record: GB runs are non-periodic (sff.c:890-899) */
        double sp[3] = {0, 0, 0}, box[4] = {0, 0, 0, 0};
        if ((s = prm_find(&pf, "SOLVENT_POINTERS")) && sec_numbers(s, 3, sp) == 3) {
            prm->Iptres = (int)sp[0];
            prm->Nspm = (int)sp[1];
            prm->Nspsol = (int)sp[2];
        }
        if (prm->Nspm > 0) read_int(&pf, "ATOMS_PER_MOLECULE", prm->Nspm, &prm->Boundary, 0);
        if ((s = prm_find(&pf, "BOX_DIMENSIONS")) && sec_numbers(s, 4, box) == 4) {
            prm->Box[3] = box[0];
            prm->Box[0] = box[1];
            prm->Box[1] = box[2];
            prm->Box[2] = box[3];
        }
        if (prm->Iptres > 0 && prm->Iptres <= prm->Nres) prm->Ipatm = prm->Ipres[prm->Iptres];
    }
    if (prm->IfCap) {
        double cap[4] = {0, 0, 0, 0}, nc = 0;
        if ((s = prm_find(&pf, "CAP_INFO")) && sec_numbers(s, 1, &nc) == 1) prm->Natcap = (int)nc;
        if ((s = prm_find(&pf, "CAP_INFO2")) && sec_numbers(s, 4, cap) == 4) {
            prm->Cutcap = cap[0];
            prm->Xcap = cap[1];
            prm->Ycap = cap[2];
            prm->Zcap = cap[3];
        }
    }

    if (!pf.old) {
        rc |= read_dbl(&pf, "RADII", n, &prm->Rborn, 1) < 0;
        rc |= read_dbl(&pf, "SCREEN", n, &prm->Fs, 1) < 0;
        if (rc) goto fail;
    } else {
        prm->Rborn = (double *)calloc((size_t)n + 1, sizeof(double));
        prm->Fs = (double *)calloc((size_t)n + 1, sizeof(double));
        if (!prm->IfBox) {
            for (i = 0; i < n; i++)
                if (old_gb_radius(prm->AtomNames, i, &prm->Rborn[i], &prm->Fs[i]) != 0) {
                    nabc_seterr("rdparm: bad atom symbol: %d, %c", i, prm->AtomNames[i][0]);
                    goto fail;
                }
        }
    }
    prm->Fsmax = 0.0;
    for (i = 0; i < n; i++) {
        double si = prm->Fs[i] * (prm->Rborn[i] - 0.09);
        if (si > prm->Fsmax) prm->Fsmax = si;
    }
    /* HCP strand and complex pointers, when present (prm.c:1743-1795) */
    if ((s = prm_find(&pf, "POINTERS FOR STRAND AND COMPLEX")) != NULL) {
        double sc[2] = {1, 1};
        if (sec_numbers(s, 2, sc) == 2) {
            prm->Ncomplex = (int)sc[0];
            prm->Nstrand = (int)sc[1];
            read_int(&pf, "STRAND_POINTER", prm->Nstrand, &prm->Ipstrand, 0);
            read_int(&pf, "COMPLEX_POINTER", prm->Ncomplex, &prm->Ipcomplex, 0);
        }
    }
    nabc_parm_complete(prm);
    nabc_build_14_list(prm);
    prmfile_free(&pf);
    return prm;

fail_early:
    prmfile_free(&pf);
    free(prm);
    return NULL;
fail:
    prmfile_free(&pf);
    free_prm(prm);
    return NULL;
}

int free_prm(PARMSTRUCT_T *p)
{
    if (!p) return 0;
    free(p->AtomNames); free(p->ResNames); free(p->AtomType); free(p->AtomTree);
    free(p->Charges); free(p->Masses); free(p->Rk); free(p->Req);
    free(p->Tk); free(p->Teq); free(p->Pk); free(p->Pn); free(p->Phase); free(p->Solty);
    free(p->Cn1); free(p->Cn2); free(p->HB12); free(p->HB10);
    free(p->Rborn); free(p->Fs);
    free(p->Iac); free(p->Iblo); free(p->Cno); free(p->Ipres); free(p->ExclAt);
    free(p->TreeJoin); free(p->AtomRes);
    free(p->BondHAt1); free(p->BondHAt2); free(p->BondHNum);
    free(p->BondAt1); free(p->BondAt2); free(p->BondNum);
    free(p->AngleHAt1); free(p->AngleHAt2); free(p->AngleHAt3); free(p->AngleHNum);
    free(p->AngleAt1); free(p->AngleAt2); free(p->AngleAt3); free(p->AngleNum);
    free(p->DihHAt1); free(p->DihHAt2); free(p->DihHAt3); free(p->DihHAt4); free(p->DihHNum);
    free(p->DihAt1); free(p->DihAt2); free(p->DihAt3); free(p->DihAt4); free(p->DihNum);
    free(p->Boundary); free(p->AtomicNum);
    free(p->N14pairs); free(p->N14pairlist);
    free(p->Gvdw); free(p->Scee); free(p->Scnb);
    free(p->N14sceelist); free(p->N14scnblist);
    free(p->Ipstrand); free(p->Ipcomplex);
    free(p);
    return 0;
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

static void w_names(FILE *fp, const char *flag, NABC_NAME_T *names, int n)
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
    ptr[27] = p->IfBox; ptr[28] = p->Nmxrs; ptr[29] = p->IfCap; ptr[30] = p->Numextra;
    w_ints(fp, "POINTERS", ptr, 31);
    w_names(fp, "ATOM_NAME", p->AtomNames, n);
    w_dbls(fp, "CHARGE", p->Charges, n);
    w_ints(fp, "ATOMIC_NUMBER", p->AtomicNum, n);
    w_dbls(fp, "MASS", p->Masses, n);
    w_ints(fp, "ATOM_TYPE_INDEX", p->Iac, n);
    w_ints(fp, "NUMBER_EXCLUDED_ATOMS", p->Iblo, n);
    w_ints(fp, "NONBONDED_PARM_INDEX", p->Cno, p->Ntypes * p->Ntypes);
    w_names(fp, "RESIDUE_LABEL", p->ResNames, p->Nres);
    tmp = (int *)malloc(sizeof(int) * (size_t)(p->Nres + 1));
    for (i = 0; i < p->Nres; i++) tmp[i] = p->Ipres[i] + 1; /* 1-based in the file */
    w_ints(fp, "RESIDUE_POINTER", tmp, p->Nres);
    free(tmp);
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
    w_dbls(fp, "SOLTY", p->Solty, p->Natyp);
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
    w_names(fp, "AMBER_ATOM_TYPE", p->AtomType, n);
    w_names(fp, "TREE_CHAIN_CLASSIFICATION", p->AtomTree, n);
    w_ints(fp, "JOIN_ARRAY", p->TreeJoin, n);
    tmp = (int *)calloc((size_t)n + 1, sizeof(int));
    w_ints(fp, "IROTAT", tmp, n);
    free(tmp);
    w_flag(fp, "RADIUS_SET", "1a80");
    fprintf(fp, "%-80s\n", "synthetic intrinsic radii");
    w_dbls(fp, "RADII", p->Rborn, n);
    w_dbls(fp, "SCREEN", p->Fs, n);
    fclose(fp);
    return 0;
}
