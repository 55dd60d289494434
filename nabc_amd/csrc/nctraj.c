/*
 * nctraj.c -- Amber netCDF trajectory output, the file md() writes when
 * mme_init_sff() is given a file name (sff.c:1164-1168) and ntwx > 0
 * (sff.c:2341-2342).
 *
 * The reference writes it through the netCDF library (AmberNetcdf.c:595-735
 * netcdfCreate, :963-1050 netcdfWriteFrame); this library writes the same
 * file directly: netCDF classic format, 64-bit offset variant ("CDF\2"),
 * all integers big-endian, every name and value padded to 4 bytes.  The
 * definitions, in the reference's order:
 *   dimensions  frame (unlimited), spatial 3, atom N, cell_spatial 3,
 *               label 5, cell_angular 3
 *   variables   time(frame) float, units "picosecond"
 *               spatial(spatial) char = "xyz"
 *               coordinates(frame, atom, spatial) float, units "angstrom"
 *               cell_spatial(cell_spatial) char = "abc"
 *               cell_angular(cell_angular, label) char = "alpha", "beta ",
 *               "gamma"
 *   attributes  title "AmberNetcdf trajectory", application "AMBER",
 *               program "AmberNetcdf", programVersion "1.0",
 *               Conventions "AMBER", ConventionVersion "1.0"
 * Fixed-size variables follow the header; each record (frame) holds time
 * then coordinates.  The record count in the header is rewritten and the
 * file flushed after every frame, so the file is complete at any point.
 * One deviation: the reference never writes time (the file is created
 * without fill, AmberNetcdf.c:716), this writer stores the frame's time.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"

enum { NC_DIMENSION = 10, NC_VARIABLE = 11, NC_ATTRIBUTE = 12, NC_CHAR = 2, NC_FLOAT = 5 };

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

static int put_u64(buf_t *b, uint64_t v)
{
    return put_u32(b, (uint32_t)(v >> 32)) || put_u32(b, (uint32_t)v);
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

static int put_text_att(buf_t *b, const char *name, const char *val)
{
    const size_t n = strlen(val);
    return put_name(b, name) || put_u32(b, NC_CHAR) || put_u32(b, (uint32_t)n) || put_padded(b, val, n);
}

/* variables: name, dimension ids, optional units, type, size (per record
 * for record variables) */
typedef struct {
    const char *name;
    int ndims, dims[3];
    const char *units;
    int type;
    uint32_t vsize;
    int record;
} var_t;

/* header for natom atoms with the given variable offsets */
static int header(buf_t *b, int natom, uint32_t numrecs, const var_t *v, const uint64_t *begin)
{
    static const char *dname[6] = {"frame", "spatial", "atom", "cell_spatial", "label", "cell_angular"};
    const uint32_t dlen[6] = {0, 3, (uint32_t)natom, 3, 5, 3};
    static const char *gname[6] = {"title", "application", "program", "programVersion", "Conventions",
                                   "ConventionVersion"};
    static const char *gval[6] = {"AmberNetcdf trajectory", "AMBER", "AmberNetcdf", "1.0", "AMBER", "1.0"};
    int k, d;
    b->n = 0;
    if (put(b, "CDF\002", 4) || put_u32(b, numrecs)) return -1;
    if (put_u32(b, NC_DIMENSION) || put_u32(b, 6)) return -1;
    for (k = 0; k < 6; k++)
        if (put_name(b, dname[k]) || put_u32(b, dlen[k])) return -1;
    if (put_u32(b, NC_ATTRIBUTE) || put_u32(b, 6)) return -1;
    for (k = 0; k < 6; k++)
        if (put_text_att(b, gname[k], gval[k])) return -1;
    if (put_u32(b, NC_VARIABLE) || put_u32(b, 5)) return -1;
    for (k = 0; k < 5; k++) {
        if (put_name(b, v[k].name) || put_u32(b, (uint32_t)v[k].ndims)) return -1;
        for (d = 0; d < v[k].ndims; d++)
            if (put_u32(b, (uint32_t)v[k].dims[d])) return -1;
        if (v[k].units) {
            if (put_u32(b, NC_ATTRIBUTE) || put_u32(b, 1) || put_text_att(b, "units", v[k].units)) return -1;
        } else if (put_u64(b, 0)) { /* no attributes: ABSENT = two zero words */
            return -1;
        }
        if (put_u32(b, (uint32_t)v[k].type) || put_u32(b, v[k].vsize) || put_u64(b, begin[k])) return -1;
    }
    return 0;
}

static void layout(int natom, var_t *v)
{
    const var_t def[5] = {
        {"time", 1, {0, 0, 0}, "picosecond", NC_FLOAT, 4, 1},
        {"spatial", 1, {1, 0, 0}, NULL, NC_CHAR, 4, 0},
        {"coordinates", 3, {0, 2, 1}, "angstrom", NC_FLOAT, (uint32_t)(12 * (size_t)natom), 1},
        {"cell_spatial", 1, {3, 0, 0}, NULL, NC_CHAR, 4, 0},
        {"cell_angular", 2, {5, 4, 0}, NULL, NC_CHAR, 16, 0},
    };
    memcpy(v, def, sizeof(def));
}

int nabc_nctraj_create(nabc_nctraj *t, const char *path, int natom)
{
    var_t v[5];
    uint64_t begin[5] = {0, 0, 0, 0, 0}, off;
    buf_t b = {NULL, 0, 0};
    int k, rc = -1;
    memset(t, 0, sizeof(*t));
    if (!path || natom < 1 || natom > (1 << 28)) {
        nabc_seterr("netCDF trajectory: bad file name or atom count");
        return -1;
    }
    layout(natom, v);
    /* the header's size does not depend on the offsets it holds */
    if (header(&b, natom, 0, v, begin)) goto oom;
    off = b.n;
    for (k = 0; k < 5; k++)
        if (!v[k].record) {
            begin[k] = off;
            off += v[k].vsize;
        }
    t->rec_begin = (long long)off;
    t->recsize = 0;
    for (k = 0; k < 5; k++)
        if (v[k].record) {
            begin[k] = off + (uint64_t)t->recsize;
            t->recsize += v[k].vsize;
        }
    if (header(&b, natom, 0, v, begin)) goto oom;
    /* fixed-size variables: spatial, cell_spatial, cell_angular */
    if (put_padded(&b, "xyz", 3) || put_padded(&b, "abc", 3) || put_padded(&b, "alphabeta gamma", 15)) goto oom;
    t->fp = fopen(path, "wb");
    if (!t->fp) {
        nabc_seterr("netCDF trajectory: cannot create %s", path);
        free(b.p);
        return -1;
    }
    if (fwrite(b.p, 1, b.n, t->fp) != b.n || fflush(t->fp) != 0) {
        nabc_seterr("netCDF trajectory: write to %s failed", path);
        fclose(t->fp);
        t->fp = NULL;
        free(b.p);
        return -1;
    }
    t->natom = natom;
    t->nframes = 0;
    rc = 0;
    free(b.p);
    return rc;
oom:
    free(b.p);
    nabc_seterr("netCDF trajectory: out of memory");
    return -1;
}

/* one frame: time (ps) and coordinates x[3*natom] rounded to float, as
 * netcdfWriteFrame does */
int nabc_nctraj_write(nabc_nctraj *t, const double *x, double time_ps)
{
    buf_t b = {NULL, 0, 0};
    union { float f; uint32_t u; } c;
    size_t k;
    const size_t n3 = 3 * (size_t)t->natom;
    if (!t->fp) {
        nabc_seterr("netCDF trajectory: file not open");
        return -1;
    }
    c.f = (float)time_ps;
    if (put_u32(&b, c.u)) goto oom;
    for (k = 0; k < n3; k++) {
        c.f = (float)x[k];
        if (put_u32(&b, c.u)) goto oom;
    }
    if (fseek(t->fp, (long)(t->rec_begin + (long long)t->nframes * t->recsize), SEEK_SET) != 0 ||
        fwrite(b.p, 1, b.n, t->fp) != b.n)
        goto io;
    t->nframes++;
    {
        /* record count, bytes 4-7 of the header */
        const unsigned char nr[4] = {(unsigned char)(t->nframes >> 24), (unsigned char)(t->nframes >> 16),
                                     (unsigned char)(t->nframes >> 8), (unsigned char)t->nframes};
        if (fseek(t->fp, 4, SEEK_SET) != 0 || fwrite(nr, 1, 4, t->fp) != 4 || fflush(t->fp) != 0) goto io;
    }
    free(b.p);
    return 0;
oom:
    free(b.p);
    nabc_seterr("netCDF trajectory: out of memory");
    return -1;
io:
    free(b.p);
    nabc_seterr("netCDF trajectory: write failed");
    return -1;
}

void nabc_nctraj_close(nabc_nctraj *t)
{
    if (t->fp) fclose(t->fp);
    t->fp = NULL;
}
