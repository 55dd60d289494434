/*
 * rst7.c -- Amber restart (rst7) coordinate/velocity files.
 *
 * getxv() follows src/sff/sff.c:651-722: a title line, a line holding the
 * atom count and optionally the time, then 3*natom coordinates in 12-column
 * fields, six per line, then optionally 3*natom velocities in the same
 * layout (zeros if absent).  putxv() follows sff.c:733-777: "%6d%15.5f"
 * header and "%12.7f" fields, six per line.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"

static int read_fields(FILE *fp, int count, double *out)
{
    char line[256], field[13];
    int i = 0;
    while (i < count) {
        int k, len;
        if (!fgets(line, sizeof(line), fp)) return i;
        len = (int)strcspn(line, "\r\n");
        for (k = 0; k < 6 && i < count; k++) {
            if (12 * k >= len) break;
            memcpy(field, line + 12 * k, 12);
            field[12] = '\0';
            out[i++] = atof(field);
        }
    }
    return i;
}

int getxv(char *fname, int natom, double start_time, double *x, double *v)
{
    FILE *fp;
    char line[256];
    int n = natom, i;
    (void)start_time;
    if (!fname || !*fname || !strcmp(fname, "-")) fp = stdin;
    else if (!(fp = fopen(fname, "r"))) {
        nabc_seterr("getxv: can't open file %s", fname);
        return -1;
    }
    if (!fgets(line, sizeof(line), fp) || !fgets(line, sizeof(line), fp)) {
        nabc_seterr("getxv: short file %s", fname);
        if (fp != stdin) fclose(fp);
        return -1;
    }
    {
        double t = 0.0;
        if (sscanf(line, "%d %lf", &n, &t) < 1) {
            nabc_seterr("getxv: bad atom-count line in %s", fname);
            if (fp != stdin) fclose(fp);
            return -1;
        }
    }
    if (n != natom) {
        nabc_seterr("getxv: %s holds %d atoms, expected %d", fname, n, natom);
        if (fp != stdin) fclose(fp);
        return -1;
    }
    if (read_fields(fp, 3 * n, x) != 3 * n) {
        nabc_seterr("getxv: %s ends before all coordinates", fname);
        if (fp != stdin) fclose(fp);
        return -1;
    }
    if (v) {
        if (read_fields(fp, 3 * n, v) != 3 * n)
            for (i = 0; i < 3 * n; i++) v[i] = 0.0;
    }
    if (fp != stdin) fclose(fp);
    return 0;
}

int putxv(char *fname, char *title, int natom, double start_time,
          double *x, double *v)
{
    FILE *fp;
    int i, pass;
    if (!strcmp(fname, "-")) fp = nabout ? nabout : stdout;
    else if (!(fp = fopen(fname, "w"))) {
        nabc_seterr("Can't open file %s", fname);
        return -1;
    }
    fprintf(fp, "%s\n", title);
    fprintf(fp, "%6d%15.5f\n", natom, start_time);
    for (pass = 0; pass < 2; pass++) {
        double *a = pass == 0 ? x : v;
        if (!a) continue;
        for (i = 0; i < 3 * natom; i += 6) {
            int k, m = 3 * natom - i < 6 ? 3 * natom - i : 6;
            for (k = 0; k < m; k++) fprintf(fp, "%12.7f", a[i + k]);
            fputc('\n', fp);
        }
    }
    if (fp != stdout && fp != nabout) fclose(fp);
    return 0;
}
