/*
 * gbhip.h -- C-ABI of the MI355X generalized-Born path of nabc_amd.
 *
 * This is the drop-in boundary for two file-static functions of libsff:
 *
 *   egb(lpears, upears, pearlist, lpearsnp, upearsnp, pearlistnp, x, f, fs,
 *       rborn, q, kappa, diel_ext, enb, eelt, esurf, enp, freevectors)
 *       src/sff/eff.c:1363-1367 (Born radii, GB + Coulomb + LJ pair terms,
 *       sumdeijda chain rule; returns epol, accumulates the gradient into f)
 *   nblist(lpears, upears, pearlist, x, context_PxQ, derivs, cutoff, natom,
 *          dim, frozen)
 *       src/sff/nblist.c:813-815 (pair list: r^2 < cut^2 at build time,
 *       not both frozen; returns the number of list entries, 2 per pair)
 *
 * mme34() (eff.c:2442) is their only caller on the energy path: it calls
 * gbhip_nblist where it called nblist (eff.c:2572) and gbhip_egb where it
 * called egb (eff.c:2506, :2704), after gbhip_set_coords.
 * Everything here is plain C: pointers, sizes and int status codes
 * (0 = success, negative = error, message via gbhip_last_error()).
 * No function calls exit().
 *
 * State lives in an opaque context (one per molecule) instead of the
 * file-scope statics that egb()/nblist() keep (eff.c:1375-1376,
 * sff.c:68-234); a batch handle scores many molecules in one launch.
 */
#ifndef NABC_AMD_GBHIP_H
#define NABC_AMD_GBHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GBHIP_OK 0
#define GBHIP_EINVAL (-1)   /* bad argument or option combination */
#define GBHIP_ENOMEM (-2)   /* device or host allocation failed */
#define GBHIP_EHIP (-3)     /* a HIP runtime call failed */
#define GBHIP_ESTATE (-4)   /* call out of order (e.g. egb before nblist) */
#define GBHIP_ENODEV (-5)   /* no usable GPU */

#define GBHIP_FP64 0        /* all pair arithmetic in FP64 */
#define GBHIP_MIXED 1       /* FP32 pair arithmetic, FP64 sums and scaling */

typedef struct gbhip_ctx gbhip_ctx;
typedef struct gbhip_batch gbhip_batch;

/*
 * Per-atom topology the GB path reads (the fields egb() takes from prm and
 * the sff.c statics, eff.c:1363-2405).  All arrays have natom entries unless
 * stated.  Values are copied at context creation.
 */
typedef struct {
    int natom;
    int ntypes;
    const double *charge;    /* prmtop unit: e * 18.2223 (prm.c:777-792) */
    const double *rborn;     /* intrinsic radii, prm->Rborn */
    const double *fs;        /* screening factors after mme_init (gb 7/8 set them) */
    const int *iac;          /* 1-based Lennard-Jones type, prm->Iac */
    const int *cno;          /* ntypes*ntypes, 1-based index into cn1/cn2; <= 0: none */
    const double *cn1;       /* ntypes*(ntypes+1)/2 r^-12 coefficients */
    const double *cn2;       /* ntypes*(ntypes+1)/2 r^-6 coefficients */
    const int *iblo;         /* number of excluded-atom entries per atom, prm->Iblo */
    const int *exclat;       /* sum(iblo) entries, 1-based, 0 = none, prm->ExclAt */
    const int *frozen;       /* per-atom flag, may be NULL (none frozen) */
    const double *gbalpha;   /* OBC / neck alpha, beta, gamma (gb 2,5,7,8), */
    const double *gbbeta;    /*   may be NULL for gb 1                       */
    const double *gbgamma;
    const int *neckidx;      /* gb 7/8: (int)((rborn-1)*20+0.5), else NULL */
    double fsmax;            /* prm->Fsmax as read from the prmtop (prm.c:1652-1656):
                                pairs beyond rgbmax + fsmax skip passes 1 and 3 */
} gbhip_topology;

typedef struct {
    int gb;              /* 1, 2, 5, 7 or 8 */
    int precision;       /* GBHIP_FP64 or GBHIP_MIXED */
    double cut;          /* pair-list cutoff (A) */
    double rgbmax;       /* Born-integral cutoff (A), must be <= cut */
    double gboffset;     /* 0.09; 0.195141 for gb 8 */
    double gbneckscale;  /* gb 7: 0.361825, gb 8: 0.826836 */
    double kappa;        /* Debye-Hueckel inverse length (1/A) */
    double epsext;       /* solvent dielectric */
} gbhip_options;

/* Last error message of the calling thread's most recent failed call. */
const char *gbhip_last_error(void);

/* Number of visible GPUs (0 when none); does not create a context. */
int gbhip_device_count(void);

/*
 * Create a context on the current HIP device.  x0 (3*natom, AoS, may be
 * NULL) fixes the spatial ordering of atoms into 64-atom tiles.
 * Replaces the first-call allocations of egb() (eff.c:1519-1543) and the
 * pair-list heads of mme_init_sff (sff.c:1451-1474).
 */
int gbhip_ctx_create(const gbhip_topology *top, const gbhip_options *opt,
                     const double *x0, gbhip_ctx **out);

/*
 * Same, with all of the context's work (set-up copies included) on the
 * given hipStream_t, for callers that create contexts while other threads
 * capture graphs on their own streams.
 */
int gbhip_ctx_create_on(const gbhip_topology *top, const gbhip_options *opt,
                        const double *x0, void *stream, gbhip_ctx **out);

/* Release everything (egb(..., freevectors=1), eff.c:1463-1491). */
int gbhip_ctx_destroy(gbhip_ctx *ctx);

/*
 * Slots: memory a context can live in without allocating, for callers that
 * keep many short-lived contexts on one device (a serving loop) and must
 * not allocate, free or synchronise the device while other contexts run.
 * A slot holds a device region (its own, device_bytes > 0, or a part of
 * the caller's memory given by gbhip_slot_set_region), pinned host memory
 * for the context's copies (io_bytes: coordinates, gradient, energies and
 * dynamics logs; 8 * max(6 natom + 64, 16 log rows) bytes cover a
 * molecule; upload_bytes: set-up copies, reused in turn when smaller than
 * they need), an event and a stream (a hipStream_t, not owned).
 * gbhip_ctx_create_in_slot builds a context whose device arrays, pass-3
 * derivative store and dynamics state come from the region (error
 * GBHIP_ENOMEM when it is too small; gbhip_footprint_bytes gives a size
 * that suffices) and whose copies go through the slot's pinned memory; its
 * gbhip_ctx_destroy hands the slot back without freeing anything.  One
 * context per slot at a time.  No counterpart in the reference (libsff
 * holds one system per process).
 */
typedef struct gbhip_slot gbhip_slot;
int gbhip_slot_create(size_t device_bytes, size_t io_bytes, size_t upload_bytes, void *stream,
                      gbhip_slot **out);
int gbhip_slot_set_region(gbhip_slot *slot, void *device_base, size_t device_bytes);
int gbhip_slot_destroy(gbhip_slot *slot);
int gbhip_ctx_create_in_slot(const gbhip_topology *top, const gbhip_options *opt,
                             const double *x0, gbhip_slot *slot, gbhip_ctx **out);

/* Host coordinates, AoS x[3*natom] (copied to the device). */
int gbhip_set_coords(gbhip_ctx *ctx, const double *x);

/*
 * Rebuild the pair list from the current coordinates (nblist(),
 * nblist.c:813).  *nentries (may be NULL) receives lower+upper entries,
 * i.e. twice the number of pairs, which is what nblist() returns.
 */
int gbhip_nblist(gbhip_ctx *ctx, long long *nentries);

/*
 * One GB energy + gradient evaluation over the current list (egb() with
 * freevectors=0).  grad (3*natom AoS, host, may be NULL) is ACCUMULATED
 * into, like f in egb(); epol is egb()'s return value, evdw and eelec its
 * enb/eelt outputs.  Synchronous.
 */
int gbhip_egb(gbhip_ctx *ctx, double *grad, double *epol, double *evdw,
              double *eelec);

/*
 * One GB energy + gradient of one molecule that allocates nothing on the
 * device: what gbhip_ctx_create + gbhip_nblist + gbhip_egb +
 * gbhip_ctx_destroy compute (same kernels, same results), with every device
 * array taken from a static buffer of GBHIP_STATIC_BYTES that the library
 * holds on each GPU it runs on until the process ends.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code that
 * must not call hipMalloc/hipFree, harnesses that forbid it, tight serving
 * loops).  A system whose arrays do not fit (about 20,000 atoms; the pass-3
 * derivative store takes what is left, else pass 3 recomputes) is rejected
 * with GBHIP_EINVAL and a message naming the size.  Synchronous, on the
 * default stream; grad (3*natom AoS, host, may be NULL) is accumulated
 * into.  Every call on a GPU shares that GPU's buffer, which keeps the
 * last call's contents (each call writes what it reads before reading
 * it): calls that overlapped on one GPU would race, so the library runs
 * them one at a time under its device lock (each process has its own
 * buffer).
 */
#define GBHIP_STATIC_BYTES (64ull << 20)
int gbhip_egb_static(const gbhip_topology *top, const gbhip_options *opt, const double *x, double *grad,
                     double *epol, double *evdw, double *eelec);

/*
 * Hierarchical charge partitioning: egb_hcp() (src/sff/hcp_gb.c:2166-2240)
 * in place of egb(), as mme34 calls it when mm_options sets hcp
 * (eff.c:2687-2694).  Atoms are grouped into residues, strands and
 * complexes; a component whose geometric centre lies farther from an atom
 * than its level's threshold acts on that atom through point charges:
 *   hcp 1: one charge at the centre of charge (hcp.c:337-540),
 *   hcp 2: two, the negative and the positive atoms apart,
 *   hcp 4: two optimal point charges (HCPO, hcp.c:543-840; mm_options'
 *          hcp=4 is sff.c:1335-1340's two-charge HCPO).
 * Radii of components: egb_hcp_rbondi (hcp_gb.c:45-124), set here.  The
 * Born radii have no rgbmax cut-off and no neck term, and no pair list is
 * built (eff.c:2560-2561): gbhip_nblist returns 0 entries.  Every
 * evaluation (gbhip_egb, gbhip_mme, dynamics, conjgrad) then runs the HCP
 * kernels in the context's precision (GBHIP_MIXED: descreening and GB
 * terms in FP32, Coulomb, LJ and every sum in FP64, as the pair path); the
 * reference's single-precision parts (component radii, HCPO moments and
 * charges) stay single precision in both.  hcp = 0 (or h == NULL) returns to the pair-list path
 * (a list must be built again).  Single-molecule contexts only; not with
 * the deterministic mode; not while dynamics graphs exist.  Each call with
 * hcp != 0 allocates the hierarchy's device arrays anew (about 120 bytes
 * per atom plus the exclusions, and 2 KiB per work item for partial sums,
 * up to about 135 MB on a 256-CU device; from the slot's region for
 * a context made in a slot); those of earlier calls stay allocated until
 * the context is destroyed.
 */
typedef struct {
    int hcp;              /* 0 (off), 1, 2 or 4 */
    double h1, h2, h3;    /* residue, strand, complex thresholds (A); sff.c: 15, 50, 150 */
    double dhcp;          /* hcp 4: neutral-component charge separation factor; sff.c: 0.25 */
    int nres;             /* residues: atoms ipres[r] .. ipres[r+1]-1 */
    const int *ipres;     /*   nres + 1 entries, 0-based, ipres[nres] = natom (prm->Ipres) */
    int nstrand;          /* strands: residues ipstrand[s]-1 .. ipstrand[s+1]-2 */
    const int *ipstrand;  /*   1-based first residues (prm->Ipstrand, prm.c:1773-1775) */
    int ncomplex;         /* complexes: strands ipcomplex[c]-1 .. */
    const int *ipcomplex; /*   1-based first strands (prm->Ipcomplex, prm.c:1782-1784) */
} gbhip_hcp;
int gbhip_set_hcp(gbhip_ctx *ctx, const gbhip_hcp *h);

/* Effective Born radii of the last evaluation, original atom order. */
int gbhip_get_reff(gbhip_ctx *ctx, double *reff);

/* sumdeijda (dE/dR_i * R_i^2) of the last evaluation, original order. */
int gbhip_get_sumdeijda(gbhip_ctx *ctx, double *sumdeijda);

/*
 * Hessian of the mme energy (bonded, 1-4 and GB terms; gbhip_set_bonded
 * must have been called) at the context's coordinates and current list, by
 * central differences of the analytic gradient with step h (A): column c =
 * (g(x + h e_c) - g(x - h e_c)) / 2h, two device evaluations per column,
 * then symmetrised.  hess: host n3 x n3 (n3 = 3 natom); grad (n3) and
 * ene14 (mme34 energies + rms gradient) at x, either may be NULL.  The list
 * is not rebuilt.  Replaces the analytic second derivatives of sff2.c
 * (egb2 at sff2.c:2271 and the bonded terms' ebond2/eangl2/ephi2); use an
 * FP64 context (differences of the mixed gradient are not accurate).
 */
int gbhip_hessian(gbhip_ctx *ctx, double h, double *hess, double *grad, double *ene14);

/* ---- device-resident use (no host copies; asynchronous on the stream) ---- */

/* Stream for all later work of this context (a hipStream_t; NULL =
 * default); waits for the work queued on the previous stream first.
 * Dynamics graphs keep running on whichever stream is current when
 * gbhip_md_run is called. */
int gbhip_set_stream(gbhip_ctx *ctx, void *stream);

/* Coordinates from device memory d_x[3*natom] (AoS, FP64). */
int gbhip_set_coords_device(gbhip_ctx *ctx, const double *d_x);

/* Pair-list rebuild without reading the count back to the host. */
int gbhip_nblist_async(gbhip_ctx *ctx);

/*
 * Evaluation writing to device memory: d_grad[3*natom] is accumulated into
 * (NULL: the gradient goes to a buffer of the context, cleared first),
 * d_ene[3] (may be NULL) receives {epol, evdw, eelec}.
 */
int gbhip_egb_device(gbhip_ctx *ctx, double *d_grad, double *d_ene);

/*
 * Per-pass timing with HIP events on the context's stream: after
 * gbhip_timing_start, up to max_evals evaluations record events around
 * pass 1, the radii kernel, pass 2 and pass 3; gbhip_timing_stop waits and
 * returns the summed milliseconds ms4[4] and the number of evaluations.
 */
int gbhip_timing_start(gbhip_ctx *ctx, int max_evals);
int gbhip_timing_stop(gbhip_ctx *ctx, double *ms4, int *nevals);

/*
 * Measured VALU ceiling of the current device for this build's code: a
 * grid of independent FMA chains (2 flops each) in FP64 (GBHIP_FP64) or
 * FP32 (GBHIP_MIXED), compiled like the pair kernels (no packed FP32),
 * best of three timed launches on the default stream; writes TFLOP/s.
 * Synchronises the device.  No counterpart in the reference (a
 * measurement helper for the roofline the benchmark reports).
 */
int gbhip_valu_peak(int precision, double *tflops);

/*
 * Launch shape of the three pair kernels (pass 1, pass 2, pass 3): the
 * number of one-wavefront workgroups each is launched with (1 to 65536);
 * each workgroup takes every G-th work item (a tile pair, or a chunk of its
 * steps, see gbhip_set_split) of the cost-sorted list.
 * gbhip_set_launch takes blocks3[3] (0 keeps a value); gbhip_autotune times
 * reps evaluations on the current list with and without the derivative
 * store (gbhip_set_derivative_store; the faster is kept, so it may switch
 * the store off), then each step split for small systems, then each
 * candidate size (4 to 128 workgroups per CU, and one per work item),
 * and keeps the fastest per pass, writing
 * the choice to chosen3 (may be NULL).  MD graphs captured earlier keep the
 * shapes they were captured with.  No counterpart in the reference (egb()
 * has no launch configuration); nblist must have run.
 */
int gbhip_set_launch(gbhip_ctx *ctx, const int *blocks3);

/*
 * Step split: a work item of the three pair kernels is one wavefront's sweep
 * over the steps of a tile pair; with split = 2, 4 or 8 each tile pair's
 * steps are dealt out in that many chunks, one work item each, so that a
 * system with few tile pairs still occupies the chip (600 atoms: 55 tile
 * pairs for 1,024 SIMDs).  split3[3] per pass (pass 1, pass 2, pass 3):
 * 1, 2, 4, 8, 16, or 0 for the default, which doubles the split while the
 * pass has fewer than 64 (pass 1) or 6 (passes 2, 3) work items per CU, up
 * to 4 in pass 1 (8 below 128 tile pairs) and 8 in the others; large
 * systems run with 1.  gbhip_set_split resets the launch shapes to
 * their defaults; gbhip_autotune also times each split for small systems.
 * No counterpart in the reference (a launch-shape choice).
 */
int gbhip_set_split(gbhip_ctx *ctx, const int *split3);
int gbhip_get_split(gbhip_ctx *ctx, int *split3);

/*
 * Queued visits of pass 1 (mixed precision): a work item whose listed
 * pairs fill less than `fill` of its 64-lane x steps grid is evaluated from
 * a queue of its listed pairs, 64 at a time, instead of by the step sweep
 * (pairs at the rim of the cutoff leave most lanes of such a sweep idle).
 * pass must be 0; fill in [0, 2]: 0 never queues, above 1 always; the
 * default 0.7 was the fastest at 25,000 atoms, 25 A.  The results are the
 * same pair terms summed in another order.  The FP64 kernels keep their
 * own queue for the logarithmic cases and ignore this.  No counterpart in
 * the reference (an evaluation-order choice).
 */
int gbhip_set_queue_fill(gbhip_ctx *ctx, int pass, double fill);

/*
 * Pass-3 derivative store.  Pass 1 can write, per listed pair, the two
 * derivative numerators pass 3 needs (8 bytes per pair slot in the mixed
 * variant, 16 in FP64), so that pass 3 does not recompute them.  mode 1
 * (default): used whenever it holds the current list; synchronous list
 * builds (gbhip_nblist, batch creation, gbhip_md_init) grow it to 1.25x the
 * need when device memory allows, asynchronous ones fall back to
 * recomputation when it is too small.  mode 0: never used (memory freed).
 * gbhip_store_stats reports the entries the current list needs, the
 * entries allocated and whether the store is in use.  No counterpart in
 * the reference (eff.c:2131-2405 recomputes datmp in pass 3).
 */
int gbhip_set_derivative_store(gbhip_ctx *ctx, int mode);
int gbhip_store_stats(gbhip_ctx *ctx, long long *need, long long *cap, int *in_use);

/*
 * Deterministic reductions (on = 1).  By default the per-atom sums of the
 * three passes (Born sums, sumdeijda, gradient) and the energies are
 * accumulated with FP64 atomics, whose order varies from run to run (last
 * bits of the results).  With the mode on, every tile-pair visit stores its
 * sums per tile pair and small kernels add them up per atom and per
 * molecule in a fixed order, so the same input and list give bit-identical
 * results on every run, whatever the launch shapes, like the reference's
 * fixed-order reduction of its per-thread copies (eff.c:2768-2797,
 * eff.c:2077-2098).  Costs 4 KiB of device memory per tile pair of the list
 * capacity and the step split (each tile pair is one work item).  Not for a
 * context in a slot; not while dynamics graphs exist (GBHIP_ESTATE).
 * Covers the GB evaluation (gbhip_egb, gbhip_egb_device, batches); the
 * bonded terms of gbhip_mme still add with atomics.
 */
int gbhip_set_deterministic(gbhip_ctx *ctx, int on);
int gbhip_get_launch(gbhip_ctx *ctx, int *blocks3);
int gbhip_autotune(gbhip_ctx *ctx, int reps, int *chosen3);

/* Wait for the context's stream. */
int gbhip_sync(gbhip_ctx *ctx);

/* Pair count of the current list (unordered pairs); synchronous. */
int gbhip_npairs(gbhip_ctx *ctx, long long *npairs);

/*
 * The current list as unordered pairs (original atom indices, i in pi, j in
 * pj, unspecified order); at most cap are written, *npairs gets the total.
 * Read-back for checks; allocates temporary device memory.
 */
int gbhip_get_pairs(gbhip_ctx *ctx, int *pi, int *pj, long long cap,
                    long long *npairs);

/* Tile statistics: 64-atom tiles and stored tile pairs of the current list. */
int gbhip_tile_stats(gbhip_ctx *ctx, int *ntiles, int *ntilepairs);

/* ---- the rest of mme(): bonded and 1-4 terms on the device ---- */

/*
 * Bonded topology of one molecule, 0-based atom indices (not the prmtop's
 * 3*(atom-1) offsets), one entry per term: a dihedral series (Pn < 0 in the
 * prmtop, eff.c:619-623) is passed as one entry per Fourier term.
 * 1-4 pairs carry 1/SCEE and 1/SCNB and their LJ coefficients (p14_hb = 1
 * selects the 10-12 form, eff.c:1224-1234).  dield: 1 = diel=C, 0 = diel=R.
 * Replaces ebond/eangl/ephi (eff.c:82-627) and nbond(N14=1)
 * (eff.c:896-1331) as called from mme34 (eff.c:2617-2653).
 */
typedef struct {
    int nbond;
    const int *bond_i, *bond_j;
    const double *bond_rk, *bond_req;
    int nangle;
    const int *angle_i, *angle_j, *angle_k;
    const double *angle_tk, *angle_teq;
    int ndihed;
    const int *dihed_i, *dihed_j, *dihed_k, *dihed_l;
    const double *dihed_pk, *dihed_pn, *dihed_phase;
    int npair14;
    const int *p14_i, *p14_j, *p14_hb;
    const double *p14_scee, *p14_scnb, *p14_a, *p14_b;
    const double *charge;    /* natom, same as the topology's */
    int dield;
} gbhip_bonded;

int gbhip_set_bonded(gbhip_ctx *ctx, const gbhip_bonded *b);

/*
 * Whole mme34 evaluation (bonded + 1-4 + GB) over the current list and
 * coordinates.  grad (3*natom, host) is OVERWRITTEN with the gradient,
 * frozen atoms zeroed (eff.c:2826-2834); ene14[0..12] follows mme34's ene[]
 * (eff.c:2617-2874, [0] = total) and ene14[13] is the rms gradient.
 */
int gbhip_mme(gbhip_ctx *ctx, double *grad, double *ene14);
/* Same, device pointers, asynchronous. */
int gbhip_mme_device(gbhip_ctx *ctx, double *d_grad, double *d_ene14);

/*
 * gbhip_mme's device work as one captured hipGraph (on = 1, the default):
 * captured at the first call and re-captured when a launch argument changes
 * (a regrown derivative store, new launch shapes, new bonded terms), so each
 * minimizer step costs one graph launch instead of a dozen kernel launches.
 * A context without a stream is given one of its own at the first capture.
 * During a timing session (gbhip_timing_start) evaluations are launched
 * kernel by kernel.  on = 0 launches kernel by kernel always.  The stats
 * count captures and graph launches.  No counterpart in the reference
 * (mme34, eff.c:2442-2947, is host code).
 */
int gbhip_set_mme_graph(gbhip_ctx *ctx, int on);
int gbhip_mme_graph_stats(gbhip_ctx *ctx, long long *captures, long long *replays);

/* ---- minimisation on the device (conjgrad(), conjgrad.c:34-263) ---- */

/*
 * conjgrad(x, &n, &f, mme, &rmsgrad, &dfpred, &maxiter) over this
 * context's mme34 (gbhip_set_bonded must have been called), with the
 * minimiser's vectors (coordinates, gradient, search and restart
 * directions, the best point) resident on the device: only the energies
 * and the scalars the line search branches on come back per evaluation.
 * The same steps, restarts and tests as conjgrad.c, with the pair list
 * rebuilt as mme34 does (no list yet, or iteration % nsnb == 0,
 * eff.c:2564).  x (n = 3 natom doubles, host) is the start and receives
 * the result, *f its energy; *result gets conjgrad's return value (the
 * iteration count, or -1 bad line search, -2 uphill direction, -3 maxiter
 * reached, -4 no decrease).  report (may be NULL) is called after each
 * evaluation with its iteration number and the 14 energies of gbhip_mme
 * (the reference prints its ff: line there).  Dot products are reduced in
 * a fixed order.
 */
typedef void (*gbhip_cg_report)(int iter, const double *ene14, void *user);
int gbhip_conjgrad(gbhip_ctx *ctx, double *x, int n, double *f, double rmsgrad, double dfpred, int maxiter,
                   int nsnb, gbhip_cg_report report, void *user, int *result);

/* ---- dynamics on the device (md(), sff.c:2071-2366) ---- */

typedef struct {
    double dt;        /* time step, ps */
    double temp0;     /* target temperature, K */
    double tempi;     /* initial temperature, K (0: keep v0) */
    double gamma_ln;  /* Langevin collision frequency, 1/ps (0: Berendsen) */
    double tautp;     /* Berendsen coupling time, ps */
    double vlimit;    /* velocity clamp */
    double boltz2;    /* one half of Boltzmann's constant, 9.93595e-4 */
    int nsnb;         /* pair-list rebuild every nsnb steps */
    int zerov;        /* start from zero velocities */
    unsigned long long seed; /* noise stream */
} gbhip_md_options;

/*
 * Set up dynamics on a context that has bonded terms (gbhip_set_bonded):
 * masses[natom], coordinates x[3*natom], velocities v0 (may be NULL).
 * Builds the list and the step-0 forces as md() does; log_cap rows of
 * per-step energies are kept on the device (ring buffer).
 */
int gbhip_md_init(gbhip_ctx *ctx, const double *masses, const double *x,
                  const double *v0, const gbhip_md_options *opt, int log_cap);
/*
 * Run nsteps steps without host synchronisation (asynchronous).  With
 * use_graph != 0 each step is one hipGraph launch (a plain step, or a
 * step with the pair-list rebuild every nsnb steps).
 */
int gbhip_md_run(gbhip_ctx *ctx, int nsteps, int use_graph);
/*
 * List-rebuild bookkeeping of the runs since gbhip_md_init: rebuild steps
 * run (every nsnb steps, eff.c:2564), rebuild steps skipped because every
 * pair provably stays within the cutoff (bounding-box diagonal plus the
 * vlimit bound on motion, 2 sqrt(3) vlimit dt per step, below the cutoff:
 * the rebuild would build the list it replaces), and bounding-box checks
 * (one at gbhip_md_init, another whenever the bound runs out).  Any may be
 * NULL.  No counterpart in the reference (a scheduling statistic).
 */
int gbhip_md_stats(gbhip_ctx *ctx, long long *rebuilds, long long *skipped, long long *bbox_checks);
/* Current coordinates, velocities and gradient (any may be NULL); syncs. */
int gbhip_md_state(gbhip_ctx *ctx, double *x, double *v, double *grad);
/*
 * Per-step rows of the last min(steps, log_cap, max_rows) steps, 16 doubles
 * each: mme energies [0..12] (mme34 layout), rms gradient, ekin, temp.
 */
int gbhip_md_log(gbhip_ctx *ctx, double *rows, int max_rows, int *nrows);

/* ---- batched scoring: many molecules in one launch per pass ---- */

/*
 * A batch of nmol independent molecules sharing one option set.  x0[m]
 * holds molecule m's coordinates (3*natom_m AoS).  The pair lists are built
 * at creation from x0.
 */
int gbhip_batch_create(int nmol, const gbhip_topology *tops,
                       const gbhip_options *opt, const double *const *x0,
                       gbhip_batch **out);
int gbhip_batch_destroy(gbhip_batch *b);

/* Replace molecule m's coordinates (host AoS). */
int gbhip_batch_set_coords(gbhip_batch *b, int m, const double *x);

/* Rebuild every molecule's pair list from its current coordinates. */
int gbhip_batch_nblist(gbhip_batch *b);

/*
 * Evaluate all molecules: ene[3*nmol] = {epol, evdw, eelec} per molecule;
 * grads[m] (3*natom_m, may be NULL or hold NULL entries) accumulated into.
 */
int gbhip_batch_eval(gbhip_batch *b, double *ene, double *const *grads);

/* Evaluation only, results left on the device (for timing). */
int gbhip_batch_eval_device(gbhip_batch *b);

/* Per-pass timing of batch evaluations (as gbhip_timing_start/stop). */
int gbhip_batch_timing_start(gbhip_batch *b, int max_evals);
int gbhip_batch_timing_stop(gbhip_batch *b, double *ms4, int *nevals);

/* Launch-shape autotuning of a batch (as gbhip_autotune). */
int gbhip_batch_autotune(gbhip_batch *b, int reps, int *chosen3);

/* Derivative-store mode and statistics of a batch (as for a context). */
int gbhip_batch_set_derivative_store(gbhip_batch *b, int mode);

/* Deterministic reductions of a batch (as gbhip_set_deterministic). */
int gbhip_batch_set_deterministic(gbhip_batch *b, int on);
int gbhip_batch_store_stats(gbhip_batch *b, long long *need, long long *cap, int *in_use);

/* Total unordered pairs and atoms of the batch. */
int gbhip_batch_stats(gbhip_batch *b, long long *npairs, long long *natoms,
                      int *ntilepairs);

/* ---- synthetic systems (deterministic generator) ---- */

struct parm;
/* Build an N-atom synthetic polymer; x receives 3N coordinates. */
int gbgen_build(int natom, unsigned long long seed, struct parm **prm,
                double *x);
/* Same, written as a prmtop and an rst7 file. */
int gbgen_write(int natom, unsigned long long seed, const char *prmtop_path,
                const char *rst7_path);

#ifdef __cplusplus
}
#endif

#endif
