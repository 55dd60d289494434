/*
 * nctraj.c -- Amber netCDF trajectories and restarts, written and read
 * without the netCDF library (which this image lacks): the netCDF classic
 * format, 64-bit offset variant ("CDF\2"; "CDF\1" files are read too), all
 * integers big-endian, names and values padded to 4 bytes, fixed-size
 * variables after the header, then one record per frame holding every
 * record variable in order.
 *
 * The files have the layout AmberNetcdf.c gives them (the reference's
 * writer, through the netCDF library):
 *   trajectory (netcdfCreate, AmberNetcdf.c:595-735): dimensions frame
 *     (unlimited), spatial 3, atom N, cell_spatial 3, label 5, cell_angular
 *     3; variables time(frame) float "picosecond", spatial(spatial) char
 *     "xyz", coordinates(frame, atom, spatial) float "angstrom",
 *     cell_spatial char "abc", cell_angular(cell_angular, label) char
 *     "alpha", "beta ", "gamma"; global attributes title "AmberNetcdf
 *     trajectory", application "AMBER", program "AmberNetcdf",
 *     programVersion "1.0", Conventions "AMBER", ConventionVersion "1.0".
 *     Frames are coordinates rounded to float (:1000-1011).
 *   restart (netcdfWriteRestart, :366-593): no frame dimension; time
 *     double, coordinates(atom, spatial) double, optional velocities
 *     (atom, spatial) double "angstrom/picosecond" with scale_factor 20.455,
 *     optional cell_lengths/cell_angles double, optional temp0 double
 *     "kelvin"; Conventions "AMBERRESTART".
 * One deviation: trajectory frames carry their time (the reference creates
 * the file without fill and never writes it, :716).
 *
 * md() uses the nabc_nctraj_* functions; the reference's struct AmberNetcdf
 * API (sff.h:240-274) is provided on top of them, with ncid an index into a
 * table of open files.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"

enum { NC_DIMENSION = 10, NC_VARIABLE = 11, NC_ATTRIBUTE = 12 };
enum { NC_BYTE = 1, NC_CHAR = 2, NC_SHORT = 3, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };

/* ------------------------------------------------------------ byte buffer */

typedef struct {
    unsigned char *p;
    size_t n, cap;
} buf_t;

static int put(buf_t *b, const void *src, size_t n)
{
    if (b->n + n > b->cap) {
        size_t c = b->cap ? 2 * b->cap : 1024;
        while (c < b->n + n) c *= 2;
        unsigned char *q = (unsigned char *)realloc(b->p, c);
        if (!q) return -1;
        b->p = q;
        b->cap = c;
    }
    memcpy(b->p + b->n, src, n);
    b->n += n;
    return 0;
}

static int put_u32(buf_t *b, uint32_t v)
{
    unsigned char c[4] = {(unsigned char)(v >> 24), (unsigned char)(v >> 16), (unsigned char)(v >> 8),
                          (unsigned char)v};
    return put(b, c, 4);
}

static int put_u64(buf_t *b, uint64_t v) { return put_u32(b, (uint32_t)(v >> 32)) || put_u32(b, (uint32_t)v); }

static int put_f32(buf_t *b, float f)
{
    union { float f; uint32_t u; } c;
    c.f = f;
    return put_u32(b, c.u);
}

static int put_f64(buf_t *b, double d)
{
    union { double d; uint64_t u; } c;
    c.d = d;
    return put_u64(b, c.u);
}

/* bytes padded with zeros to a multiple of 4 */
static int put_padded(buf_t *b, const void *src, size_t n)
{
    static const unsigned char zero[4] = {0, 0, 0, 0};
    if (put(b, src, n)) return -1;
    return (n & 3) ? put(b, zero, 4 - (n & 3)) : 0;
}

static int put_name(buf_t *b, const char *s)
{
    const size_t n = strlen(s);
    return put_u32(b, (uint32_t)n) || put_padded(b, s, n);
}

/* ------------------------------------------------------ file description */

typedef struct {
    const char *name;
    const char *text; /* text attribute, or NULL for a double */
    double value;
} att_t;

typedef struct {
    const char *name;
    int ndims, dims[3];
    att_t att[2];
    int natt, type;
    uint32_t vsize; /* bytes (per record for record variables), padded to 4 */
    int record;
} var_t;

typedef struct {
    int ndims;
    const char *dname[6];
    uint32_t dlen[6]; /* 0: the unlimited (record) dimension */
    const char *gval[6]; /* title, application, program, programVersion, Conventions, ConventionVersion */
    int nvars;
    var_t v[10];
} desc_t;

static int put_att(buf_t *b, const att_t *a)
{
    if (put_name(b, a->name)) return -1;
    if (a->text) {
        const size_t n = strlen(a->text);
        return put_u32(b, NC_CHAR) || put_u32(b, (uint32_t)n) || put_padded(b, a->text, n);
    }
    return put_u32(b, NC_DOUBLE) || put_u32(b, 1) || put_f64(b, a->value);
}

static int header(buf_t *b, const desc_t *D, uint32_t numrecs, const uint64_t *begin)
{
    static const char *gname[6] = {"title", "application", "program", "programVersion", "Conventions",
                                   "ConventionVersion"};
    int k, d;
    b->n = 0;
    if (put(b, "CDF\002", 4) || put_u32(b, numrecs)) return -1;
    if (put_u32(b, NC_DIMENSION) || put_u32(b, (uint32_t)D->ndims)) return -1;
    for (k = 0; k < D->ndims; k++)
        if (put_name(b, D->dname[k]) || put_u32(b, D->dlen[k])) return -1;
    if (put_u32(b, NC_ATTRIBUTE) || put_u32(b, 6)) return -1;
    for (k = 0; k < 6; k++) {
        const att_t a = {gname[k], D->gval[k], 0.0};
        if (put_att(b, &a)) return -1;
    }
    if (put_u32(b, NC_VARIABLE) || put_u32(b, (uint32_t)D->nvars)) return -1;
    for (k = 0; k < D->nvars; k++) {
        const var_t *v = &D->v[k];
        if (put_name(b, v->name) || put_u32(b, (uint32_t)v->ndims)) return -1;
        for (d = 0; d < v->ndims; d++)
            if (put_u32(b, (uint32_t)v->dims[d])) return -1;
        if (v->natt) {
            if (put_u32(b, NC_ATTRIBUTE) || put_u32(b, (uint32_t)v->natt)) return -1;
            for (d = 0; d < v->natt; d++)
                if (put_att(b, &v->att[d])) return -1;
        } else if (put_u64(b, 0)) { /* no attributes: two zero words */
            return -1;
        }
        if (put_u32(b, (uint32_t)v->type) || put_u32(b, v->vsize) || put_u64(b, begin[k])) return -1;
    }
    return 0;
}

/* header with offsets: fixed-size variables follow the header in order,
 * then the records; *rec_begin, *recsize describe the records */
static int layout_header(buf_t *b, const desc_t *D, uint64_t *begin, long long *rec_begin, long long *recsize)
{
    uint64_t off;
    int k;
    memset(begin, 0, sizeof(uint64_t) * (size_t)D->nvars);
    if (header(b, D, 0, begin)) return -1; /* its size does not depend on the offsets */
    off = b->n;
    for (k = 0; k < D->nvars; k++)
        if (!D->v[k].record) {
            begin[k] = off;
            off += D->v[k].vsize;
        }
    *rec_begin = (long long)off;
    *recsize = 0;
    for (k = 0; k < D->nvars; k++)
        if (D->v[k].record) {
            begin[k] = off + (uint64_t)*recsize;
            *recsize += D->v[k].vsize;
        }
    return header(b, D, 0, begin);
}

static const char ABC_LABELS[15] = {'a', 'l', 'p', 'h', 'a', 'b', 'e', 't', 'a', ' ', 'g', 'a', 'm', 'm', 'a'};

static void traj_desc(desc_t *D, int natom)
{
    const desc_t T = {
        6,
        {"frame", "spatial", "atom", "cell_spatial", "label", "cell_angular"},
        {0, 3, (uint32_t)natom, 3, 5, 3},
        {"AmberNetcdf trajectory", "AMBER", "AmberNetcdf", "1.0", "AMBER", "1.0"},
        5,
        {{"time", 1, {0, 0, 0}, {{"units", "picosecond", 0}}, 1, NC_FLOAT, 4, 1},
         {"spatial", 1, {1, 0, 0}, {{0}}, 0, NC_CHAR, 4, 0},
         {"coordinates", 3, {0, 2, 1}, {{"units", "angstrom", 0}}, 1, NC_FLOAT, (uint32_t)(12 * (size_t)natom), 1},
         {"cell_spatial", 1, {3, 0, 0}, {{0}}, 0, NC_CHAR, 4, 0},
         {"cell_angular", 2, {5, 4, 0}, {{0}}, 0, NC_CHAR, 16, 0}},
    };
    *D = T;
}

/* ------------------------------------------------------- trajectory (md) */

int nabc_nctraj_create(nabc_nctraj *t, const char *path, int natom)
{
    desc_t D;
    uint64_t begin[10];
    buf_t b = {NULL, 0, 0};
    memset(t, 0, sizeof(*t));
    if (!path || natom < 1 || natom > (1 << 28)) {
        nabc_seterr("netCDF trajectory: bad file name or atom count");
        return -1;
    }
    traj_desc(&D, natom);
    if (layout_header(&b, &D, begin, &t->rec_begin, &t->recsize) || put_padded(&b, "xyz", 3) ||
        put_padded(&b, "abc", 3) || put_padded(&b, ABC_LABELS, 15)) {
        free(b.p);
        nabc_seterr("netCDF trajectory: out of memory");
        return -1;
    }
    t->fp = fopen(path, "w+b");
    if (!t->fp || fwrite(b.p, 1, b.n, t->fp) != b.n || fflush(t->fp) != 0) {
        nabc_seterr("netCDF trajectory: cannot write %s", path);
        if (t->fp) fclose(t->fp);
        t->fp = NULL;
        free(b.p);
        return -1;
    }
    free(b.p);
    t->natom = natom;
    t->nframes = 0;
    return 0;
}

/* frame `set` (0-based; frames beyond the last one extend the file) */
static int traj_put_frame(nabc_nctraj *t, int set, const double *x, double time_ps)
{
    buf_t b = {NULL, 0, 0};
    size_t k;
    const size_t n3 = 3 * (size_t)t->natom;
    if (put_f32(&b, (float)time_ps)) goto oom;
    for (k = 0; k < n3; k++)
        if (put_f32(&b, (float)x[k])) goto oom;
    if (fseek(t->fp, (long)(t->rec_begin + (long long)set * t->recsize), SEEK_SET) != 0 ||
        fwrite(b.p, 1, b.n, t->fp) != b.n)
        goto io;
    free(b.p);
    if (set >= t->nframes) {
        /* record count, bytes 4-7 of the header */
        unsigned char nr[4];
        t->nframes = set + 1;
        nr[0] = (unsigned char)(t->nframes >> 24);
        nr[1] = (unsigned char)(t->nframes >> 16);
        nr[2] = (unsigned char)(t->nframes >> 8);
        nr[3] = (unsigned char)t->nframes;
        if (fseek(t->fp, 4, SEEK_SET) != 0 || fwrite(nr, 1, 4, t->fp) != 4) goto io_nofree;
    }
    if (fflush(t->fp) != 0) goto io_nofree;
    return 0;
oom:
    free(b.p);
    nabc_seterr("netCDF trajectory: out of memory");
    return -1;
io:
    free(b.p);
io_nofree:
    nabc_seterr("netCDF trajectory: write failed");
    return -1;
}

/* one frame: time (ps) and coordinates x[3*natom] rounded to float */
int nabc_nctraj_write(nabc_nctraj *t, const double *x, double time_ps)
{
    if (!t->fp) {
        nabc_seterr("netCDF trajectory: file not open");
        return -1;
    }
    return traj_put_frame(t, t->nframes, x, time_ps);
}

void nabc_nctraj_close(nabc_nctraj *t)
{
    if (t->fp) fclose(t->fp);
    t->fp = NULL;
}

/* --------------------------------------------------------------- restart */

int netcdfWriteRestart(char *filename, int natom, double *X, double *V, double *box, double time,
                       double temp)
{
    desc_t D = {
        5,
        {"spatial", "atom", "cell_spatial", "label", "cell_angular"},
        {3, 0, 3, 5, 3},
        {"AmberNetcdf restart", "AMBER", "AmberNetcdf", "1.0", "AMBERRESTART", "1.0"},
        0,
        {{0}},
    };
    uint64_t begin[10];
    long long rb, rs;
    buf_t b = {NULL, 0, 0};
    FILE *fp;
    int k, iv = -1, ib = -1, it = -1;
    const size_t n3 = 3 * (size_t)natom;
    if (!filename || natom < 1 || natom > (1 << 26) || !X) {
        nabc_seterr("netcdfWriteRestart: bad file name, atom count or coordinates");
        return 1;
    }
    D.dlen[1] = (uint32_t)natom;
    /* variables in the reference's order (AmberNetcdf.c:415-512) */
    {
        const var_t tv = {"time", 0, {0}, {{"units", "picosecond", 0}}, 1, NC_DOUBLE, 8, 0};
        const var_t sv = {"spatial", 1, {0}, {{0}}, 0, NC_CHAR, 4, 0};
        const var_t cv = {"coordinates", 2, {1, 0}, {{"units", "angstrom", 0}}, 1, NC_DOUBLE, (uint32_t)(8 * n3), 0};
        const var_t vv = {"velocities", 2, {1, 0}, {{"units", "angstrom/picosecond", 0}, {"scale_factor", NULL, 20.455}},
                          2, NC_DOUBLE, (uint32_t)(8 * n3), 0};
        const var_t csv = {"cell_spatial", 1, {2}, {{0}}, 0, NC_CHAR, 4, 0};
        const var_t cav = {"cell_angular", 2, {4, 3}, {{0}}, 0, NC_CHAR, 16, 0};
        const var_t lv = {"cell_lengths", 1, {2}, {{"units", "angstrom", 0}}, 1, NC_DOUBLE, 24, 0};
        const var_t av = {"cell_angles", 1, {4}, {{"units", "degree", 0}}, 1, NC_DOUBLE, 24, 0};
        const var_t pv = {"temp0", 0, {0}, {{"units", "kelvin", 0}}, 1, NC_DOUBLE, 8, 0};
        D.v[D.nvars++] = tv;
        D.v[D.nvars++] = sv;
        D.v[D.nvars++] = cv;
        if (V) {
            iv = D.nvars;
            D.v[D.nvars++] = vv;
        }
        D.v[D.nvars++] = csv;
        D.v[D.nvars++] = cav;
        if (box) {
            ib = D.nvars;
            D.v[D.nvars++] = lv;
            D.v[D.nvars++] = av;
        }
        if (temp >= 0) {
            it = D.nvars;
            D.v[D.nvars++] = pv;
        }
    }
    if (layout_header(&b, &D, begin, &rb, &rs)) goto oom;
    /* data in variable order */
    for (k = 0; k < D.nvars; k++) {
        size_t i;
        if (k == 0) {
            if (put_f64(&b, time >= 0 ? time : 0.0)) goto oom;
        } else if (k == 1) {
            if (put_padded(&b, "xyz", 3)) goto oom;
        } else if (k == 2) {
            for (i = 0; i < n3; i++)
                if (put_f64(&b, X[i])) goto oom;
        } else if (k == iv) {
            for (i = 0; i < n3; i++)
                if (put_f64(&b, V[i])) goto oom;
        } else if (!strcmp(D.v[k].name, "cell_spatial")) {
            if (put_padded(&b, "abc", 3)) goto oom;
        } else if (!strcmp(D.v[k].name, "cell_angular")) {
            if (put_padded(&b, ABC_LABELS, 15)) goto oom;
        } else if (k == ib) {
            for (i = 0; i < 3; i++)
                if (put_f64(&b, box[i])) goto oom;
        } else if (ib >= 0 && k == ib + 1) {
            for (i = 3; i < 6; i++)
                if (put_f64(&b, box[i])) goto oom;
        } else if (k == it) {
            if (put_f64(&b, temp)) goto oom;
        }
    }
    fp = fopen(filename, "wb");
    if (!fp || fwrite(b.p, 1, b.n, fp) != b.n) {
        if (fp) fclose(fp);
        free(b.p);
        nabc_seterr("netcdfWriteRestart: cannot write %s", filename);
        return 1;
    }
    fclose(fp);
    free(b.p);
    return 0;
oom:
    free(b.p);
    nabc_seterr("netcdfWriteRestart: out of memory");
    return 1;
}

/* ----------------------------------------------------------------- reader */

typedef struct {
    char name[64];
    int type, ndims, dims[4], record;
    uint64_t begin, vsize;
} rvar_t;

typedef struct {
    FILE *fp;
    int writable;      /* created by netcdfCreate */
    nabc_nctraj traj;  /* writable files */
    int restart, natom, nframes, version;
    long long recsize;
    rvar_t coord, velo, time, lengths, angles, temp;
} ncfile_t;

#define NC_MAX_OPEN 64
static ncfile_t g_nc[NC_MAX_OPEN];

static int get(FILE *fp, void *dst, size_t n) { return fread(dst, 1, n, fp) == n ? 0 : -1; }

static int get_u32(FILE *fp, uint32_t *v)
{
    unsigned char c[4];
    if (get(fp, c, 4)) return -1;
    *v = (uint32_t)c[0] << 24 | (uint32_t)c[1] << 16 | (uint32_t)c[2] << 8 | c[3];
    return 0;
}

static int skip_padded(FILE *fp, size_t n) { return fseek(fp, (long)((n + 3) & ~(size_t)3), SEEK_CUR); }

static int get_name(FILE *fp, char *out, size_t cap)
{
    uint32_t n;
    if (get_u32(fp, &n) || n > (1u << 20)) return -1;
    if (n < cap) {
        if (get(fp, out, n)) return -1;
        out[n] = 0;
        return (n & 3) ? fseek(fp, (long)(4 - (n & 3)), SEEK_CUR) : 0;
    }
    out[0] = 0;
    return skip_padded(fp, n);
}

static size_t type_size(int t)
{
    switch (t) {
    case NC_BYTE: case NC_CHAR: return 1;
    case NC_SHORT: return 2;
    case NC_INT: case NC_FLOAT: return 4;
    case NC_DOUBLE: return 8;
    default: return 0;
    }
}

/* an attribute list; the value of a text attribute named want is copied
 * to out */
static int get_atts(FILE *fp, const char *want, char *out, size_t cap)
{
    uint32_t tag, n, k, type, nel;
    char name[64];
    if (get_u32(fp, &tag) || get_u32(fp, &n)) return -1;
    if (tag == 0 && n == 0) return 0;
    if (tag != NC_ATTRIBUTE) return -1;
    for (k = 0; k < n; k++) {
        size_t sz;
        if (get_name(fp, name, sizeof(name)) || get_u32(fp, &type) || get_u32(fp, &nel)) return -1;
        sz = type_size((int)type);
        if (!sz) return -1;
        if (want && type == NC_CHAR && !strcmp(name, want) && nel < cap) {
            if (get(fp, out, nel)) return -1;
            out[nel] = 0;
            if ((nel & 3) && fseek(fp, (long)(4 - (nel & 3)), SEEK_CUR)) return -1;
        } else if (skip_padded(fp, sz * nel)) {
            return -1;
        }
    }
    return 0;
}

static int parse(ncfile_t *F, const char *path)
{
    unsigned char magic[4];
    uint32_t tag, n, k, d, numrecs, len[64];
    char name[64], conv[64] = "";
    int atom_dim = -1, rec_dim = -1;
    F->fp = fopen(path, "rb");
    if (!F->fp) return -1;
    if (get(F->fp, magic, 4) || memcmp(magic, "CDF", 3) || (magic[3] != 1 && magic[3] != 2)) return -1;
    F->version = magic[3];
    if (get_u32(F->fp, &numrecs)) return -1;
    if (get_u32(F->fp, &tag) || get_u32(F->fp, &n) || n > 64) return -1;
    if (tag != 0 && tag != NC_DIMENSION) return -1;
    for (k = 0; k < n; k++) {
        if (get_name(F->fp, name, sizeof(name)) || get_u32(F->fp, &len[k])) return -1;
        if (!strcmp(name, "atom")) atom_dim = (int)k;
        if (len[k] == 0) rec_dim = (int)k;
    }
    if (get_atts(F->fp, "Conventions", conv, sizeof(conv))) return -1;
    F->restart = !strcmp(conv, "AMBERRESTART");
    if (!F->restart && strcmp(conv, "AMBER")) return -1;
    if (atom_dim < 0) return -1;
    F->natom = (int)len[atom_dim];
    F->nframes = F->restart ? 1 : (int)numrecs;
    if (get_u32(F->fp, &tag) || get_u32(F->fp, &n) || n > 64) return -1;
    if (tag != 0 && tag != NC_VARIABLE) return -1;
    F->recsize = 0;
    for (k = 0; k < n; k++) {
        rvar_t v;
        uint32_t nd, dim, type, vsize, b32;
        memset(&v, 0, sizeof(v));
        if (get_name(F->fp, v.name, sizeof(v.name)) || get_u32(F->fp, &nd) || nd > 4) return -1;
        v.ndims = (int)nd;
        for (d = 0; d < nd; d++) {
            if (get_u32(F->fp, &dim)) return -1;
            v.dims[d] = (int)dim;
            if ((int)dim == rec_dim) v.record = 1;
        }
        if (get_atts(F->fp, NULL, NULL, 0) || get_u32(F->fp, &type) || get_u32(F->fp, &vsize)) return -1;
        v.type = (int)type;
        v.vsize = vsize;
        if (F->version == 2) {
            uint32_t hi;
            if (get_u32(F->fp, &hi) || get_u32(F->fp, &b32)) return -1;
            v.begin = (uint64_t)hi << 32 | b32;
        } else {
            if (get_u32(F->fp, &b32)) return -1;
            v.begin = b32;
        }
        if (v.record) F->recsize += v.vsize;
        if (!strcmp(v.name, "coordinates")) F->coord = v;
        else if (!strcmp(v.name, "velocities")) F->velo = v;
        else if (!strcmp(v.name, "time")) F->time = v;
        else if (!strcmp(v.name, "cell_lengths")) F->lengths = v;
        else if (!strcmp(v.name, "cell_angles")) F->angles = v;
        else if (!strcmp(v.name, "temp0")) F->temp = v;
    }
    if (!F->coord.type) return -1;
    return 0;
}

/* count values of variable v at record `set` as doubles */
static int read_var(const ncfile_t *F, const rvar_t *v, int set, double *out, size_t count)
{
    size_t k, sz = type_size(v->type);
    unsigned char c[8];
    const long long off = (long long)v->begin + (v->record ? (long long)set * F->recsize : 0);
    if ((sz != 4 && sz != 8) || fseek(F->fp, (long)off, SEEK_SET) != 0) return -1;
    for (k = 0; k < count; k++) {
        if (get(F->fp, c, sz)) return -1;
        if (sz == 4) {
            union { uint32_t u; float f; } w;
            w.u = (uint32_t)c[0] << 24 | (uint32_t)c[1] << 16 | (uint32_t)c[2] << 8 | c[3];
            out[k] = v->type == NC_FLOAT ? (double)w.f : (double)(int32_t)w.u;
        } else {
            union { uint64_t u; double d; } w;
            int i;
            w.u = 0;
            for (i = 0; i < 8; i++) w.u = w.u << 8 | c[i];
            out[k] = w.d;
        }
    }
    return 0;
}

/* ------------------------------------------- the reference's API (sff.h) */

static void fill_struct(struct AmberNetcdf *A, int id, const ncfile_t *F)
{
    memset(A, 0, sizeof(*A));
    A->ncid = id;
    A->isNCrestart = F->restart;
    A->ncframe = F->nframes;
    A->ncatom = F->natom;
    A->ncatom3 = 3 * F->natom;
    A->frameDID = A->atomDID = A->spatialDID = A->labelDID = A->cell_spatialDID = A->cell_angularDID = -1;
    A->coordVID = 1;
    A->velocityVID = F->velo.type ? 1 : -1;
    A->cellLengthVID = F->lengths.type ? 1 : -1;
    A->cellAngleVID = F->angles.type ? 1 : -1;
    A->timeVID = F->time.type ? 1 : -1;
    A->TempVID = F->temp.type ? 1 : -1;
    A->spatialVID = A->cell_spatialVID = A->cell_angularVID = -1;
}

static ncfile_t *handle(struct AmberNetcdf *A)
{
    if (!A || A->ncid < 0 || A->ncid >= NC_MAX_OPEN || !g_nc[A->ncid].fp) return NULL;
    return &g_nc[A->ncid];
}

static int free_slot(void)
{
    int k;
    for (k = 0; k < NC_MAX_OPEN; k++)
        if (!g_nc[k].fp) return k;
    return -1;
}

int netcdfLoad(struct AmberNetcdf *A, char *filename)
{
    const int id = free_slot();
    ncfile_t *F;
    if (!A || !filename || id < 0) {
        nabc_seterr("netcdfLoad: bad arguments or too many open files");
        return 1;
    }
    F = &g_nc[id];
    memset(F, 0, sizeof(*F));
    if (parse(F, filename)) {
        if (F->fp) fclose(F->fp);
        memset(F, 0, sizeof(*F));
        nabc_seterr("netcdfLoad: %s is not an Amber netCDF trajectory or restart", filename);
        return 1;
    }
    fill_struct(A, id, F);
    if (F->restart) {
        double t = 0.0;
        if (F->time.type && read_var(F, &F->time, 0, &t, 1) == 0) A->restartTime = t;
        if (F->temp.type) (void)read_var(F, &F->temp, 0, &A->temp0, 1);
    }
    return 0;
}

int netcdfCreate(struct AmberNetcdf *A, char *filename, int natom, int isBox)
{
    const int id = free_slot();
    ncfile_t *F;
    if (!A || id < 0) {
        nabc_seterr("netcdfCreate: bad arguments or too many open files");
        return 1;
    }
    if (isBox) {
        nabc_seterr("netcdfCreate: box information is not written (non-periodic GB only)");
        return 1;
    }
    F = &g_nc[id];
    memset(F, 0, sizeof(*F));
    if (nabc_nctraj_create(&F->traj, filename, natom)) return 1;
    F->fp = F->traj.fp;
    F->writable = 1;
    F->natom = natom;
    fill_struct(A, id, F);
    return 0;
}

int netcdfWriteFrame(struct AmberNetcdf *A, int set, double *X, double *box)
{
    ncfile_t *F = handle(A);
    (void)box;
    if (!F || !F->writable || !X || set < 0) {
        nabc_seterr("netcdfWriteFrame: no trajectory open for writing, or bad arguments");
        return 1;
    }
    if (traj_put_frame(&F->traj, set, X, 0.0)) return 1;
    A->ncframe = F->traj.nframes;
    return 0;
}

int netcdfWriteNextFrame(struct AmberNetcdf *A, double *X, double *box)
{
    if (netcdfWriteFrame(A, A ? A->currentFrame : 0, X, box)) return 1;
    A->currentFrame++;
    return 0;
}

int netcdfGetFrame(struct AmberNetcdf *A, int set, double *X, double *box)
{
    ncfile_t *F = handle(A);
    if (!F || F->writable) {
        nabc_seterr("netcdfGetFrame: no file open for reading");
        return 1;
    }
    if (!X || set < 0 || set >= F->nframes) {
        nabc_seterr("netcdfGetFrame: no frame %d", set);
        return 1;
    }
    if (read_var(F, &F->coord, set, X, 3 * (size_t)F->natom)) return 1;
    if (box && F->lengths.type && F->angles.type) {
        if (read_var(F, &F->lengths, set, box, 3) || read_var(F, &F->angles, set, box + 3, 3)) return 1;
    }
    return 0;
}

int netcdfGetNextFrame(struct AmberNetcdf *A, double *X, double *box)
{
    if (!A || netcdfGetFrame(A, A->currentFrame, X, box)) return 1;
    A->currentFrame++;
    return 0;
}

/* restart velocities as stored (internal units; the file's scale_factor
 * 20.455 converts them to angstrom/picosecond) */
int netcdfGetVelocity(struct AmberNetcdf *A, int set, double *V)
{
    ncfile_t *F = handle(A);
    if (!F || !V || !F->velo.type || set < 0 || set >= F->nframes || F->writable) {
        nabc_seterr("netcdfGetVelocity: no velocities");
        return 1;
    }
    return read_var(F, &F->velo, set, V, 3 * (size_t)F->natom) ? 1 : 0;
}

int netcdfClose(struct AmberNetcdf *A)
{
    ncfile_t *F = handle(A);
    if (!F) return 0;
    if (F->writable) nabc_nctraj_close(&F->traj);
    else fclose(F->fp);
    memset(F, 0, sizeof(*F));
    A->ncid = -1;
    return 0;
}

int netcdfInfo(struct AmberNetcdf *A)
{
    ncfile_t *F = handle(A);
    if (!F) {
        fprintf(stdout, "Error: netcdfInfo: file not open.\n");
        return 1;
    }
    fprintf(stdout, "  File is a NetCDF AMBER %s%s%s\n", F->restart ? "restart" : "trajectory",
            F->velo.type ? ", with velocity info" : "", F->lengths.type ? ", with box info" : "");
    fprintf(stdout, "    %i atoms.\n    %i coordinates.\n    %i frames.\n", F->natom, 3 * F->natom,
            F->writable ? F->traj.nframes : F->nframes);
    return 0;
}

int netcdfDebug(struct AmberNetcdf *A)
{
    return netcdfInfo(A);
}
