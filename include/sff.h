/*
 * sff.h -- the libsff header name, for C programs written against libsff
 * (src/sff/sff.h of dacase/nabc): its basic types (sff.h:18-36) and
 * everything this library provides of it, declared in sff_gb.h with the
 * same names and signatures -- PARMSTRUCT_T field for field, rdparm,
 * wrparm, free_prm, mm_options, mme_init_sff, mme, md, conjgrad, getxv,
 * putxv, the AmberNetcdf functions and the nabout stream.  Such a program
 * compiles against this directory and links with libnabc_amd.so.  The
 * parts of libsff outside the generalized-Born energy path (nmode, newton,
 * xmin, lmod, mme2, mdrat, the distance-geometry types) are not provided.
 */
#ifndef NABC_AMD_SFF_COMPAT_H
#define NABC_AMD_SFF_COMPAT_H

#include <stddef.h>
#include <stdio.h>

#include "sff_gb.h"

#ifndef FALSE
#define FALSE 0
#endif
#ifndef TRUE
#define TRUE 1
#endif
#define UNDEF (-1)

typedef int INT_T;
typedef size_t SIZE_T;
#define NAB_DOUBLE_PRECISION 1
#define REAL_T double
typedef char STRING_T;
typedef FILE FILE_T;
#define NAMESIZE NABC_NAMESIZE
typedef NABC_NAME_T NAME_T;

#endif
