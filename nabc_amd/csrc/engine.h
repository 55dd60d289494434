/*
 * engine.h -- internal declarations shared by the HIP translation units of
 * nabc_amd (GB kernels in gbhip.hip, bonded terms and dynamics in
 * mm_device.hip).  Not part of the public C-ABI.
 */
#ifndef NABC_AMD_ENGINE_H
#define NABC_AMD_ENGINE_H

#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "gbhip.h"

#define TILE 64
/* energy partial sums per (molecule, term): workgroups add into slot
 * blockIdx mod ENE_SLOTS, so that tens of thousands of workgroups of one
 * molecule do not queue on one address */
#define ENE_SLOTS 64

/*
 * Held while a stream is captured into a graph and while contexts allocate
 * or free device memory: allocation calls synchronise implicitly, which
 * another thread's capture in progress does not allow.
 */
inline std::recursive_mutex &device_alloc_lock()
{
    static std::recursive_mutex m;
    return m;
}

int seterr(int code, const char *fmt, ...);

#define HIPCHK(call)                                                                \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess)                                                       \
            return seterr(GBHIP_EHIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

#define LAUNCHCHK()                                                                    \
    do {                                                                               \
        hipError_t e_ = hipGetLastError();                                             \
        if (e_ != hipSuccess) return seterr(GBHIP_EHIP, "kernel launch failed: %s", hipGetErrorString(e_)); \
    } while (0)

/* ------------------------------------------------------------ device data */

/* per queue position: the tile pair's oriented tiles, index, non-empty
 * columns and derivative-store offset (one scalar load per visit) */
struct QMeta {
    int a, b, t, mol; /* lane tile, column tile, tile pair, molecule */
    unsigned long long cols;
    long long start;
};

struct DevPtrs {
    /* per slot */
    const int *perm;      /* slot -> global atom index, -1 = padding */
    const int *slot_mol;  /* slot -> molecule */
    const double *q, *rb, *ri0, *sj, *ga, *gbb, *gg;
    const double *rgbpsmax2;  /* per molecule: (rgbmax + Fsmax)^2 */
    const float *ri0_f, *sj_f, *rb_f;
    const int *neck, *frz, *ljrow, *ljcol;
    double *xs, *ys, *zs;
    double *sumi, *reff, *psi, *dfac, *sumdeijda, *gx, *gy, *gz;
    double *ene;          /* 3 per molecule: epol, evdw, eelec */
    double *eslot;        /* ENE_SLOTS partial sums per ene entry (ene_slot_sums adds them up) */
    /* tiles and tile pairs */
    int ntiles;
    const int *tile_mol;
    const int *tile_end;  /* one past the molecule's last tile, per tile */
    const long long *rowbase;
    int *cand, *rowcnt, *rowstart;
    int *ntp;             /* device count of tile pairs */
    int2 *tp;             /* (I, J), I <= J, sorted by I then J */
    int2 *tpo;            /* (A, B): lanes take tile A, steps walk tile B */
    int *tflip;           /* 1 when (A, B) = (J, I) */
    unsigned long long *tcol; /* non-empty columns (B atoms) of the bit matrix */
    unsigned long long *mask, *excl; /* bit matrices in (A, B) orientation */
    unsigned long long *sbal; /* per tile pair and step: lanes whose derivative-store entry pass 1 wrote */
    int *tord;            /* non-empty tile pairs, most non-empty columns first */
    int *tbin;            /* counting sort by column count: 64 counts, 64 offsets */
    int *nnz;             /* number of non-empty tile pairs */
    int *tnp;             /* listed pairs per tile pair */
    QMeta *qm;            /* per queue position, see QMeta */
    long long *store_need; /* entries the current list needs */
    int *store_ok;        /* 1 when the store holds the current list */
    void *dstore;         /* pass-3 derivative numerators (float2 / double2) */
    long long store_cap;  /* entries allocated */
    unsigned long long *npairs;
    double *bblo, *bbhi;  /* per tile bounding box */
    /* exclusions, CSR over global atoms, partners j > i */
    const int *ex_ptr, *ex_j, *inv;
    int natoms;
    /* LJ dense tables, concatenated per molecule */
    const double2 *ljab;  /* (A, B) = (cn1, cn2) per type pair */
    const int2 *mol_lj;   /* per molecule: offset of its ntypes x ntypes block in ljab, ntypes */
    const double *neck_pos, *neck_val;
    const float *neck_pos_f, *neck_val_f;
    /* deterministic mode (GBPar::det): per tile pair t and side (0: tile I
     * of (I, J), 1: tile J), DET_NC components of 64 per-atom visit sums;
     * per tile pair its three pass-2 energies; per tile its self energy;
     * per molecule its first tile (nmol + 1 entries); per tile its row's
     * pass-2 energies (3) and the tile pairs (I, T), I < T, of its column
     * by increasing I (count per tile; the list of column T of a molecule
     * whose first tile is I0 starts at its tile-pair base + (T - I0)(T - I0
     * + 1) / 2, a triangle as large as its tile-pair capacity) */
    double *part, *pe, *pself, *prow;
    const int *mol_tile0;
    int *dcol, *dcolcnt;
};

#define DET_NC 4

struct GBPar {
    int gb, has_neck, nslots, nmol;
    /* per pass (born, pairs, chain): each tile pair's steps are dealt out
     * in 1 << split_sh chunks, one work item each (more waves than tile
     * pairs for small systems) */
    int split_sh[3];
    /* 1: deterministic reductions (every visit's sums stored per tile pair
     * and added up in a fixed order instead of by atomics; no step split) */
    int det;
    /* per pass: a visit whose listed pairs fill less than this fraction of
     * its 64 x steps lane-steps is evaluated from a queue of its listed
     * pairs, 64 at a time, instead of by the step sweep */
    float qfill[3];
    int pad_;
    double cut2, rgbmax, rgbmax1i, rgbmax2i;
    double neckscale, kappa, epsext;
};

/*
 * Hierarchical charge partitioning (hcp.hip): atoms in their original
 * order, grouped into residues, strands and complexes; a component farther
 * from an atom than its level's threshold acts through nq point charges.
 * All arrays are device memory of the engine (freed with it).
 */
struct HcpDev {
    int on = 0;        /* 1: evaluations run the HCP kernels */
    int nq = 1;        /* charges per component (1, or 2 for hcp 2 and 4) */
    int opca = 0;      /* hcp 4: optimal point charges */
    int gb = 5;
    int natom = 0, nres = 0, nstrand = 0, ncomplex = 0;
    float dhcp = 0.25f;
    double h1sq = 0, h2sq = 0, h3sq = 0, kappa = 0, epsext = 78.5;
    const int *ipres = nullptr;   /* nres + 1 first atoms */
    const int *rfirst = nullptr;  /* nstrand + 1 first residues */
    const int *sfirst = nullptr;  /* ncomplex + 1 first strands */
    const double4 *atom = nullptr; /* per atom: charge, rborn - gboffset, scaled radius, rborn */
    const double4 *obc = nullptr;  /* per atom: alpha, beta, gamma */
    const int2 *lj = nullptr;     /* per atom: row offset into ljab, column */
    const int *frz = nullptr;
    const int *ex_ptr = nullptr, *ex_j = nullptr; /* excluded partners, both directions, ascending */
    const int *inv = nullptr;     /* atom -> engine slot */
    const double2 *ljab = nullptr;
    const double *rc1 = nullptr, *rc2 = nullptr, *rc3 = nullptr; /* component radii */
    double *xc1 = nullptr, *xc2 = nullptr, *xc3 = nullptr;       /* geometric centres */
    double *qc1 = nullptr, *qc2 = nullptr, *qc3 = nullptr;       /* (x, y, z, q) per charge */
    double *rf1 = nullptr, *rf2 = nullptr, *rf3 = nullptr;       /* component Born radii */
    double *rsum = nullptr;  /* per residue: coordinate sums, charge-group sums (3 + 4 nq) */
    double *ssum = nullptr;  /* per strand: the same */
    int *scnt = nullptr;     /* per strand: atoms */
    float *rmp = nullptr, *smp = nullptr; /* per residue / strand: single-precision moments (10) */
    double *rw = nullptr;    /* per residue: its terms of the strand averages (6) */
    double *sw = nullptr;    /* per strand: its averaging sums (4) */
    double *reff = nullptr, *psi = nullptr, *sd = nullptr; /* per atom */
    /* per work item, its lanes' partial sums, [component][range][tile * 64 + lane]
     * (4 components: Born sum; energy sd, x, y, z; chain rule x, y, z), added
     * up per atom in a fixed order instead of by atomics */
    double *part = nullptr;
    double *gm = nullptr; /* chain rule: forces on member atoms, original order, x | y | z */
    const int *res_s = nullptr;   /* per residue: its strand */
    const int *s_cx = nullptr;    /* per strand: its complex */
    int ntile = 0;                /* 64-atom tiles of the work items */
    int nchunk = 1, csize = 1;    /* residue ranges of csize residues */
};

/*
 * Memory a context can live in without allocating (a serving slot,
 * gbhip_slot): a device region the engine suballocates from (256-byte
 * aligned, released as a whole when the context goes), pinned host memory
 * for copies (an io part for coordinates / gradient / energies / logs, an
 * upload part for set-up copies, reused in turn), an event marking the end
 * of the last copy out of the io part, and the stream.  None of it is
 * allocated or freed by a context that uses it.
 */
struct SlotMem {
    char *dbase = nullptr;
    size_t dsize = 0, dused = 0;
    char *hio = nullptr, *hup = nullptr;   /* pinned */
    size_t hio_size = 0, hup_size = 0, hup_used = 0;
    hipEvent_t ev = nullptr;
    hipStream_t stream = nullptr;
    bool own_device = false;               /* dbase allocated by the slot itself */
    bool in_use = false;                   /* a context lives in it */
};

struct Engine {
    int nmol = 0, natoms = 0, nslots = 0, ntiles = 0;
    long long tp_cap = 0;
    int precision = GBHIP_FP64;
    /* FP64 pass 1 with the logarithmic descreening cases queued (k_born_q) */
    bool born_queue = true;
    int pair_blocks = 0;       /* workgroups of the list-build kernels (16 per CU, at most the tile pairs) */
    int pass_blocks[3] = {0, 0, 0}; /* per pass: born, pairs, chain */
    int ncu = 0;
    GBPar par{};
    DevPtrs d{};
    hipStream_t stream = nullptr;
    bool own_stream = false;   /* created by the engine, destroyed with it */
    /* streams the engine created and then left for a caller's stream: kept
     * until the engine goes, since graphs captured on them still exist */
    std::vector<hipStream_t> retired_streams;
    std::vector<int> mol_atom_off, mol_natom, mol_slot_off;
    std::vector<int> perm_h;  /* slot -> global atom */
    std::vector<int> frz_h;   /* slot -> frozen flag */
    double *ene14_dev = nullptr; /* mme energies + rms gradient (14 doubles) */
    std::vector<void *> allocs;
    double *x_dev = nullptr;   /* concatenated coordinates, AoS */
    double *g_dev = nullptr;   /* concatenated gradient scratch, AoS */
    double *tmp_dev = nullptr; /* natoms doubles */
    bool have_list = false;
    bool have_eval = false;
    /* optional per-pass timing: 5 events per evaluation on the engine stream
     * (before pass 1, after pass 1, after the radii, after pass 2, after pass 3) */
    std::vector<hipEvent_t> tev;
    int tev_cap = 0, tev_used = 0;
    /* pass-3 derivative store (separate from allocs: it is resized) */
    void *store_mem = nullptr;
    /* the deterministic mode's arrays (allocated when the mode is first set) */
    void *det_mem = nullptr;
    int store_mode = 1;        /* 1: used when it fits, 0: always recompute */
    bool store_frozen = false; /* a captured graph holds the current pointer */
    /* the host knows that the store holds the current list (sized for every
     * tile pair full, or checked at the last synchronous list build), so
     * pass 3's recomputing kernel, which would exit at once, is not launched */
    bool store_holds_list = false;
    /* set while a graph that rebuilds the list itself is captured: the
     * recomputing pass-3 kernel is then left out only when the store holds
     * every list the graph can build */
    bool capture_spans_rebuilds = false;
    /* pinned host staging for coordinates in and gradient/energies out
     * (3*natoms + 3*natoms + 64 doubles: energies and minimiser scalars at
     * the end); stage_ev marks the end of the last copy out of the
     * coordinate part */
    double *h_stage = nullptr;
    size_t h_stage_doubles = 0;
    hipEvent_t stage_ev = nullptr;
    bool stage_failed = false; /* no pinned memory: copy directly */
    /* set when the context lives in a slot: device memory, staging and event
     * are the slot's (nothing is allocated or freed here) */
    SlotMem *slot = nullptr;
    /* hierarchical charge partitioning in place of the pair-list passes */
    HcpDev hcp{};

    ~Engine()
    {
        if (slot) {
            /* the slot's memory outlives the context: only wait for this
             * context's work, then hand the region back */
            if (stream) (void)hipStreamSynchronize(stream);
            for (hipEvent_t e : tev) (void)hipEventDestroy(e);
            slot->dused = 0;
            slot->hup_used = 0;
            slot->in_use = false;
            return;
        }
        if (h_stage) {
            if (stream) (void)hipStreamSynchronize(stream);
            (void)hipHostFree(h_stage);
        }
        if (stage_ev) (void)hipEventDestroy(stage_ev);
        for (hipEvent_t e : tev) (void)hipEventDestroy(e);
        if (store_mem) (void)hipFree(store_mem);
        if (det_mem) (void)hipFree(det_mem);
        if (own_stream && stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        for (hipStream_t s : retired_streams) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
        for (void *p : allocs) (void)hipFree(p);
    }

    /* device bytes: from the slot's region, or hipMalloc */
    int dalloc_bytes(void **p, size_t bytes)
    {
        if (bytes == 0) bytes = 1;
        if (slot) {
            const size_t off = (slot->dused + 255) & ~(size_t)255;
            if (off + bytes > slot->dsize)
                return seterr(GBHIP_ENOMEM, "slot device region of %zu bytes is full (%zu more wanted)", slot->dsize,
                              bytes);
            *p = slot->dbase + off;
            slot->dused = off + bytes;
            return 0;
        }
        void *q = nullptr;
        if (hipMalloc(&q, bytes) != hipSuccess) return seterr(GBHIP_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
        allocs.push_back(q);
        *p = q;
        return 0;
    }

    template <typename T> int dalloc(T **p, size_t n)
    {
        void *q = nullptr;
        int rc = dalloc_bytes(&q, (n ? n : 1) * sizeof(T));
        if (rc) return rc;
        *p = (T *)q;
        return 0;
    }

    /*
     * Host-to-device copy on the engine stream.  In a slot it goes through
     * the pinned upload area (asynchronous; when the area is full the
     * stream is waited for and the area reused); otherwise from pageable
     * memory and waited for.  A legacy-stream copy would have to order
     * against other threads' streams, some of them capturing.
     */
    int h2d(void *dst, const void *src, size_t bytes)
    {
        if (bytes == 0) return 0;
        if (slot && slot->hup && bytes <= slot->hup_size) {
            size_t off = (slot->hup_used + 255) & ~(size_t)255;
            if (off + bytes > slot->hup_size) {
                HIPCHK(hipStreamSynchronize(stream));
                off = 0;
            }
            memcpy(slot->hup + off, src, bytes);
            slot->hup_used = off + bytes;
            HIPCHK(hipMemcpyAsync(dst, slot->hup + off, bytes, hipMemcpyHostToDevice, stream));
            return 0;
        }
        HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return 0;
    }

    template <typename T> int upload(const T **dst, const std::vector<T> &src)
    {
        T *p;
        int rc = dalloc(&p, src.size());
        if (rc) return rc;
        if ((rc = h2d(p, src.data(), src.size() * sizeof(T)))) return rc;
        *dst = p;
        return 0;
    }

    /* wait for the set-up copies of a slot's upload area */
    int upload_done()
    {
        HIPCHK(hipStreamSynchronize(stream));
        if (slot) slot->hup_used = 0;
        return 0;
    }
};


/* bonded and 1-4 terms of one molecule, original atom order (mm_device.hip) */
struct BondedDev {
    bool set = false;
    int nb = 0, na = 0, nd = 0, n14 = 0, dield = 1;
    const int *b_i = nullptr, *b_j = nullptr;
    const double *b_rk = nullptr, *b_req = nullptr;
    const int *a_i = nullptr, *a_j = nullptr, *a_k = nullptr;
    const double *a_tk = nullptr, *a_teq = nullptr;
    const int *d_i = nullptr, *d_j = nullptr, *d_k = nullptr, *d_l = nullptr;
    const double *d_pk = nullptr, *d_pn = nullptr, *d_phase = nullptr;
    const int *p_i = nullptr, *p_j = nullptr, *p_hb = nullptr;
    const double *p_scee = nullptr, *p_scnb = nullptr, *p_a = nullptr, *p_b = nullptr;
    const double *q = nullptr;
    const int *frozen = nullptr;  /* original order, always allocated */
    double *ene = nullptr;        /* 13 doubles, mme34 layout */
    double *gpart = nullptr;      /* per-block partial sums of |grad|^2 */
};

/* device-resident dynamics state (mm_device.hip) */
struct MDState {
    bool ready = false;
    double *v = nullptr, *accel = nullptr, *minv = nullptr, *sqrmass = nullptr;
    double *scal = nullptr;   /* [0] ekin of the last step, [1] ekin accumulator */
    int *step = nullptr;      /* device step counter (noise stream index) */
    double *log = nullptr;    /* per step: ene[0..13], ekin, temp */
    int log_cap = 0, steps_done = 0;
    double dtx = 0, dt5 = 0, c_implic = 1, c_explic = 1, c_ave = 1, sdfac = 0, gammai = 0;
    double ekin0 = 0, dttp = 0, vlimit = 10, rndf = 1, boltz2 = 9.93595e-4;
    int nsnb = 10;
    /* Rebuild steps up to static_until can be skipped: every pair is within
     * the cutoff at every step up to it (bounding-box diagonal at the last
     * check + the largest separation growth per step, 2 sqrt(3) vlimit dtx,
     * times the steps since), so a rebuild would build the list it replaces.
     * static_growth > 0 while the bound can apply (vlimit > 0, finite). */
    long long static_until = 0;
    double static_growth = 0.0;
    long long rebuilds_run = 0, rebuilds_skipped = 0, bbox_checks = 0;
    unsigned long long seed = 0;
    hipGraphExec_t g_step = nullptr, g_rebuild = nullptr;
    hipGraph_t gr_step = nullptr, gr_rebuild = nullptr;
};

/* captured mme evaluation (mm_device.hip): replayed while the launch
 * arguments it was captured with (key) are unchanged */
struct MmeGraph {
    bool on = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned char> key;
    long long captures = 0, replays = 0;
};

/* device-resident conjugate-gradient vectors (cg_device.hip), allocated at
 * the first gbhip_conjgrad call: search direction, restart direction and
 * gradient change, gradient at the line start, best point and its
 * gradient (3 natoms each), per-block partial sums and the reduced
 * scalars */
struct CGState {
    double *d = nullptr, *rsdx = nullptr, *rsdg = nullptr, *ginit = nullptr, *xopt = nullptr, *gopt = nullptr;
    double *part = nullptr, *res = nullptr;
};

struct gbhip_ctx {
    Engine E;
    BondedDev B;
    MDState M;
    MmeGraph G;
    CGState C;
    ~gbhip_ctx()
    {
        if (G.exec) (void)hipGraphExecDestroy(G.exec);
        if (G.graph) (void)hipGraphDestroy(G.graph);
        if (M.g_step) (void)hipGraphExecDestroy(M.g_step);
        if (M.g_rebuild) (void)hipGraphExecDestroy(M.g_rebuild);
        if (M.gr_step) (void)hipGraphDestroy(M.gr_step);
        if (M.gr_rebuild) (void)hipGraphDestroy(M.gr_rebuild);
    }
};

struct gbhip_batch {
    Engine E;
};

int engine_build(Engine &E, int nmol, const gbhip_topology *tops, const gbhip_options *o,
                 const double *const *x0);
int engine_gather(Engine &E, const double *x_dev);
int engine_nblist(Engine &E);
int engine_store_reserve(Engine &E, bool upper = false);
bool engine_store_holds_list(const Engine &E);
/* engine_eval flags */
#define EV_GATHERED 1 /* the slots already hold x_dev's coordinates, accumulators cleared */
#define EV_NO_POST 2  /* the caller sums the energy slots and scatters the gradient itself */
int engine_eval(Engine &E, const double *x_dev, double *g_dev, int flags = 0);
/* energy-slot sums and slot gradients into g_dev (may be NULL): k_post */
int engine_post(Engine &E, double *g_dev, int nene, int nthreads);
/* the HCP evaluation (hcp.hip), called by engine_eval when E.hcp.on */
int engine_eval_hcp(Engine &E, const double *x_dev, double *g_dev, int flags);
/* mme34 of E.x_dev into E.g_dev and E.ene14_dev (mm_device.hip) */
int engine_mme(struct gbhip_ctx *c);

/* ------------------------------------------- per-slot steps shared by the
 * evaluation kernels and the fused kernels of the mme and dynamics paths */

/* slot s: coordinates from AoS x (zero for padding slots) and the
 * per-evaluation accumulators cleared */
__device__ __forceinline__ void slot_gather(const DevPtrs &d, int s, const double *__restrict__ x)
{
    const int a = d.perm[s];
    if (a >= 0) {
        d.xs[s] = x[3 * (size_t)a];
        d.ys[s] = x[3 * (size_t)a + 1];
        d.zs[s] = x[3 * (size_t)a + 2];
    } else {
        d.xs[s] = 0.0; d.ys[s] = 0.0; d.zs[s] = 0.0;
    }
    d.sumi[s] = 0.0;
    d.gx[s] = 0.0; d.gy[s] = 0.0; d.gz[s] = 0.0;
}

/* energy entries (3 per molecule): entry e = the sum of its ENE_SLOTS
 * partials, each read by one lane of a wavefront and summed by a fixed
 * shuffle tree (deterministic); the partials are cleared for the next
 * evaluation (they start at zero from allocation).  Called by every lane
 * of the grid; its wavefronts take the entries in turn. */
static_assert(ENE_SLOTS == 64, "one lane per energy slot");
__device__ __forceinline__ void ene_slot_sums(const DevPtrs &d, int nene)
{
    const int lane = threadIdx.x & 63;
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int nwaves = (int)((gridDim.x * blockDim.x) >> 6);
    for (int e = wave; e < nene; e += nwaves) {
        double *p = d.eslot + (size_t)e * ENE_SLOTS;
        double v = p[lane];
        p[lane] = 0.0;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) d.ene[e] = v;
    }
}

bool engine_stage(Engine &E);
/* device-to-host copy on the engine stream, waited for; through the pinned
 * staging when it is large enough (no pageable copy), else directly */
int engine_d2h(Engine &E, void *dst, const void *src, size_t bytes);

#endif
