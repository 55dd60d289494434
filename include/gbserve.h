/*
 * gbserve.h -- serving loop for minimisation / dynamics jobs on one GPU.
 *
 * No counterpart in the reference (libsff runs one system per process, its
 * state in file-scope statics, sff.c:68-234).  A job is what a libsff
 * driver does with one molecule: conjgrad() over mme() (conjgrad.c:34) or
 * md() (sff.c:2071); the loop runs many such jobs at once, each with its
 * state resident on the device, admitting new ones as slots and device
 * memory free up.  The device memory is one pool of hbm_budget bytes
 * allocated when the loop starts; each job gets a contiguous block of its
 * footprint (first fit) and its context lives in it, so nothing is
 * allocated, freed or synchronised device-wide while jobs run.
 */
#ifndef NABC_AMD_GBSERVE_H
#define NABC_AMD_GBSERVE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GBSERVE_CG 0
#define GBSERVE_MD 1
#define TILE_ATOMS 64

typedef struct {
    void *prm;               /* PARMSTRUCT_T * of the molecule (read by the job) */
    const double *x;         /* starting coordinates, 3*natom */
    int kind;                /* GBSERVE_CG or GBSERVE_MD */
    int nsteps;              /* CG: maxiter; MD: steps */
    unsigned long long seed; /* MD noise stream */
    double arrival_s;        /* arrival time, seconds after the run starts */
    double *x_out;           /* may be NULL: receives the final coordinates, 3*natom */
} gbserve_job;

typedef struct {
    int slots;               /* jobs resident at once (one HIP stream each) */
    size_t hbm_budget;       /* device bytes the admitted jobs may hold together */
    int precision;           /* GBHIP_MIXED or GBHIP_FP64 for the jobs */
    int check_fp64;          /* re-evaluate each final state in FP64 and in the job precision */
    double cg_rmsgrad;       /* conjgrad rmsgrad argument */
    double cg_dfpred;        /* conjgrad dfpred argument */
    int progress;            /* > 0: a line on stderr every `progress` completed jobs */
    double watchdog_s;       /* a job in one phase longer than this is reported on stderr
                                (0: 30 s; < 0: no watchdog) */
} gbserve_config;

typedef struct {
    double start_s;          /* admission time */
    double end_s;            /* completion time */
    double latency_s;        /* completion - arrival */
    double energy;           /* final energy reported by the job */
    double energy_fp64;      /* final coordinates, fresh list, FP64 (check_fp64) */
    double energy_final_eval;/* same coordinates and list in the job precision */
    size_t footprint;        /* predicted device bytes */
    int status;              /* conjgrad return code, 0 for MD, < 0 error; -2 never fits */
    int slot;                /* worker that ran it */
    int max_active;          /* jobs resident when it was admitted */
    double setup_s;          /* admission to context ready */
    size_t pool_offset;      /* its block of the device pool */
} gbserve_result;

/* Device bytes a context for natom atoms (and an md log of md_steps rows) holds. */
size_t gbhip_footprint_bytes(int natom, int md_steps);

/*
 * Run all jobs to completion.  mm_opts (may be NULL) is an mm_options()
 * string applied on top of the current options for the jobs only (e.g.
 * "gb=5, cut=20, rgbmax=20, dt=0.001, gamma_ln=5, temp0=300, tempi=300").
 * Blocks; returns 0, or -1 on a setup error (message: mm_get_error()).
 */
int gbserve_run(const gbserve_job *jobs, int njobs, const gbserve_config *cfg,
                const char *mm_opts, gbserve_result *res);

/*
 * The same admission and completion loop with the work of a job done by a
 * caller's function instead of the GPU job: run(job, slot, user, &energy)
 * is called on the slot's worker thread (slots threads at once) and returns
 * the job status (0 or a positive code for success, negative for failure).
 * Admission by arrival time, free slot and a free block of
 * gbhip_footprint_bytes in a pool of cfg->hbm_budget bytes is unchanged
 * (pool_offset reports the block); precision and check_fp64 are ignored and
 * no HIP call is made.  For schedulers that run other work per job and for
 * testing the loop itself without a device.  No counterpart in the
 * reference.
 */
typedef int (*gbserve_runner)(const gbserve_job *job, int slot, void *user, double *energy);
int gbserve_run_custom(const gbserve_job *jobs, int njobs, const gbserve_config *cfg, gbserve_runner run,
                       void *user, gbserve_result *res);

#ifdef __cplusplus
}
#endif

#endif
