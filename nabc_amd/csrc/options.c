/*
 * options.c -- mm_options() keyword parser for nabc_amd.
 *
 * Accepts the keyword set of the libsff lexer (src/sff/mm_options.l:28-97):
 * "key=value" or "key value" items separated by blanks, tabs, newlines or
 * commas.  Integer-valued keys with a restricted value set (gb, gb_debug,
 * e_debug, ips, hcp, rattle, gbsa, dim) are checked the same way.  An
 * unknown key or a malformed value is an error, as there (mm_options.l:101);
 * here the message goes to stderr and the call returns -1 instead of exiting.
 *
 * Three keys are additions of this library: gbprec=fp64|mixed selects the
 * GB arithmetic, md_seed=<int> seeds the device Langevin noise stream, and
 * cgdevice=0|1 chooses where conjgrad() over mme() keeps its vectors (1,
 * the default: on the device; 0: on the host, as the reference does).
 */
#include <ctype.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"

nabc_opts_t nabc_opts;
FILE *nabout = NULL;
static char nabc_errbuf[512];

void nabc_seterr(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(nabc_errbuf, sizeof(nabc_errbuf), fmt, ap);
    va_end(ap);
    fprintf(stderr, "%s\n", nabc_errbuf);
}

const char *mm_get_error(void) { return nabc_errbuf; }

void mm_set_output(FILE *fp) { nabout = fp; }

void mm_reset_options(void)
{
    nabc_opts_t *o = &nabc_opts;
    memset(o, 0, sizeof(*o));
    o->cut = 20.0;  o->cutnp = 10.0; o->ntpr = 10; o->nscm = 0;
    o->nsnb = 10;   o->nsnp = 10;    o->nchk = 10000; o->nchk2 = 10000;
    o->scnb = 2.0;  o->scee = 1.2;   o->rgbmax = 15.0; o->gboffset = 0.09;
    o->gb = 0;      o->gbsa = 0;
    o->surften = 0.005; o->dradius = 0.5; o->deltar = 0.5; o->rwater = 1.4;
    o->alphanp = 0.75; o->kappanp = 2.227; o->min_volume = 0.01;
    o->blocksize = 8; o->max_set_size = 20; o->dynamic_loops = 1;
    o->mpi_min_tasks = 8; o->cull_np_lists = 1; o->use_lower_tri = 0;
    o->epsext = 78.5; o->kappa = 0.0; o->dield = 1; o->dielc = 1.0;
    o->t = 0.0; o->dt = 0.001; o->tautp = 999999.0; o->gamma_ln = 0.0;
    o->temp0 = 300.0; o->boltz2 = 9.93595e-4; o->vlimit = 10.0;
    o->genmass = 10.0; o->ntpr_md = 10; o->ntwx = 0; o->zerov = 0;
    o->tempi = 0.0; o->irattle = 0; o->dim = 3;
    o->hcp_h1 = 15.0; o->hcp_h2 = 50.0; o->hcp_h3 = 150.0; o->dhcp = 0.25;
    o->gbprec = 0; o->md_seed = 71277ULL; o->cgdevice = 1;
}

typedef enum { K_DBL, K_INT, K_POSINT, K_NNINT, K_SET, K_DIEL, K_PREC, K_U64 } kind_t;

typedef struct {
    const char *key;
    kind_t kind;
    void *dst;
    const char *allowed; /* for K_SET: allowed single digits */
} keydef_t;

#define O (&nabc_opts)
static const keydef_t keys[] = {
    {"cut", K_DBL, &O->cut, 0}, {"cutnp", K_DBL, &O->cutnp, 0},
    {"scnb", K_DBL, &O->scnb, 0}, {"scee", K_DBL, &O->scee, 0},
    {"ntpr", K_POSINT, &O->ntpr, 0}, {"hcp", K_SET, &O->hcp, "0124"},
    {"hcp_h1", K_DBL, &O->hcp_h1, 0}, {"hcp_h2", K_DBL, &O->hcp_h2, 0},
    {"hcp_h3", K_DBL, &O->hcp_h3, 0}, {"dhcp", K_DBL, &O->dhcp, 0},
    {"nscm", K_NNINT, &O->nscm, 0}, {"nsnb", K_NNINT, &O->nsnb, 0},
    {"nsnp", K_NNINT, &O->nsnp, 0}, {"ntwx", K_NNINT, &O->ntwx, 0},
    {"nchk", K_NNINT, &O->nchk, 0}, {"nchk2", K_NNINT, &O->nchk2, 0},
    {"dt", K_DBL, &O->dt, 0}, {"t", K_DBL, &O->t, 0},
    {"tautp", K_DBL, &O->tautp, 0}, {"temp0", K_DBL, &O->temp0, 0},
    {"tempi", K_DBL, &O->tempi, 0}, {"vlimit", K_DBL, &O->vlimit, 0},
    {"gamma_ln", K_DBL, &O->gamma_ln, 0}, {"genmass", K_DBL, &O->genmass, 0},
    {"wcons", K_DBL, &O->wcons, 0}, {"ntpr_md", K_POSINT, &O->ntpr_md, 0},
    {"zerov", K_NNINT, &O->zerov, 0}, {"k4d", K_DBL, &O->k4d, 0},
    {"ips", K_SET, &O->ips, "0123"}, {"gb", K_SET, &O->gb, "0123578"},
    {"gb_debug", K_SET, &O->gb_debug, "01"}, {"nr_debug", K_SET, &O->nr_debug, "01"},
    {"gb2_debug", K_SET, &O->gb2_debug, "01"}, {"gbsa_debug", K_SET, &O->gbsa_debug, "01"},
    {"e_debug", K_SET, &O->e_debug, "0123"}, {"epsext", K_DBL, &O->epsext, 0},
    {"kappa", K_DBL, &O->kappa, 0}, {"rgbmax", K_DBL, &O->rgbmax, 0},
    {"surften", K_DBL, &O->surften, 0}, {"dradius", K_DBL, &O->dradius, 0},
    {"deltar", K_DBL, &O->deltar, 0}, {"rwater", K_DBL, &O->rwater, 0},
    {"alphanp", K_DBL, &O->alphanp, 0}, {"kappanp", K_DBL, &O->kappanp, 0},
    {"min_volume", K_DBL, &O->min_volume, 0}, {"max_set_size", K_NNINT, &O->max_set_size, 0},
    {"MPI_min_tasks", K_NNINT, &O->mpi_min_tasks, 0},
    {"cull_np_lists", K_SET, &O->cull_np_lists, "01"},
    {"use_lower_tri", K_SET, &O->use_lower_tri, "01"},
    {"dynamic_loops", K_SET, &O->dynamic_loops, "01"},
    {"blocksize", K_NNINT, &O->blocksize, 0}, {"gbsa", K_SET, &O->gbsa, "01"},
    {"dim", K_SET, &O->dim, "34"}, {"diel", K_DIEL, &O->dield, 0},
    {"dielc", K_DBL, &O->dielc, 0}, {"rattle", K_SET, &O->irattle, "012"},
    {"wdssp", K_DBL, &O->wdssp, 0},
    {"gbprec", K_PREC, &O->gbprec, 0}, {"md_seed", K_U64, &O->md_seed, 0},
    {"cgdevice", K_SET, &O->cgdevice, "01"},
};
#undef O

static int is_float(const char *s)
{
    char *e;
    if (!*s) return 0;
    strtod(s, &e);
    return *e == '\0';
}

static int is_int(const char *s)
{
    const char *p = s;
    if (*p == '+' || *p == '-') p++;
    if (!isdigit((unsigned char)*p)) return 0;
    while (isdigit((unsigned char)*p)) p++;
    return *p == '\0';
}

static int set_key(const char *key, const char *val)
{
    size_t i;
    for (i = 0; i < sizeof(keys) / sizeof(keys[0]); i++) {
        const keydef_t *k = &keys[i];
        if (strcmp(k->key, key) != 0) continue;
        switch (k->kind) {
        case K_DBL:
            if (!is_float(val)) goto bad;
            *(double *)k->dst = strtod(val, NULL);
            break;
        case K_INT:
            if (!is_int(val)) goto bad;
            *(int *)k->dst = atoi(val);
            break;
        case K_POSINT:
            if (!is_int(val) || atoi(val) < 1 || val[0] == '-') goto bad;
            *(int *)k->dst = atoi(val);
            break;
        case K_NNINT:
            if (!is_int(val) || val[0] == '-') goto bad;
            *(int *)k->dst = atoi(val);
            break;
        case K_U64:
            if (!is_int(val) || val[0] == '-') goto bad;
            *(unsigned long long *)k->dst = strtoull(val, NULL, 10);
            break;
        case K_SET:
            if (strlen(val) != 1 || !strchr(k->allowed, val[0])) goto bad;
            *(int *)k->dst = val[0] - '0';
            break;
        case K_DIEL:
            if (!strcmp(val, "Yammp")) *(int *)k->dst = -5;
            else if (!strcmp(val, "R94")) *(int *)k->dst = -4;
            else if (!strcmp(val, "RL94")) *(int *)k->dst = -3;
            else if (!strcmp(val, "RL")) *(int *)k->dst = -2;
            else if (!strcmp(val, "R")) *(int *)k->dst = 0;
            else if (!strcmp(val, "C")) *(int *)k->dst = 1;
            else goto bad;
            break;
        case K_PREC:
            if (!strcmp(val, "fp64")) *(int *)k->dst = 0;
            else if (!strcmp(val, "mixed")) *(int *)k->dst = 1;
            else goto bad;
            break;
        }
        if (nabout) fprintf(nabout, "\tmm_options:  %s=%s\n", key, val);
        return 0;
    }
bad:
    nabc_seterr("Error: mm_option: unknown keyword or invalid value: %s=%s", key, val);
    return -1;
}

int mm_options(char *opts)
{
    const char *p = opts;
    char key[64], val[128];
    while (*p) {
        size_t n;
        while (*p == ' ' || *p == '\t' || *p == '\n' || *p == ',') p++;
        if (!*p) break;
        n = 0;
        while (*p && !strchr(" \t\n,=", *p) && n < sizeof(key) - 1) key[n++] = *p++;
        key[n] = '\0';
        /* Assign: " *= *" or " +" */
        while (*p == ' ') p++;
        if (*p == '=') {
            p++;
            while (*p == ' ') p++;
        }
        n = 0;
        while (*p && !strchr(" \t\n,=", *p) && n < sizeof(val) - 1) val[n++] = *p++;
        val[n] = '\0';
        if (set_key(key, val) != 0) return -1;
    }
    return 0;
}

/* make sure defaults are in place even if mm_options() is never called */
__attribute__((constructor)) static void nabc_options_init(void)
{
    mm_reset_options();
}
