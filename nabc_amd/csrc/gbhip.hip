/*
 * gbhip.hip -- MI355X (gfx950) generalized-Born energy + gradient engine.
 *
 * One engine serves a single molecule (gbhip_ctx) or a batch (gbhip_batch):
 * molecules are concatenated, each padded to whole 64-atom tiles, so every
 * kernel below runs once per pass for the whole set.
 *
 * Layout in HBM (per atom "slot" = position in the spatially sorted order):
 *   SoA FP64 coordinates xs/ys/zs, charge, intrinsic radius rb, offset
 *   radius ri0 = rb - gboffset, scaled radius sj = fs*ri0, OBC alpha/beta/
 *   gamma, neck index, LJ row/column into a dense per-molecule table.
 *   Pair list = tile pairs (I <= J, same molecule) with an exact 64x64
 *   membership bit matrix (r^2 < cut^2 at build time, not both frozen, one
 *   bit per unordered pair) and a parallel exclusion bit matrix.
 *
 * Per evaluation (egb(), eff.c:1363-2417):
 *   k_gather   sort coordinates into slots, clear accumulators
 *   k_born     pass 1: descreening sums, both directions per pair
 *   k_radii    reff/psi (HCT or OBC/neck tanh rescale, FP64), the diagonal
 *              GB self term, and the chain-rule factor dR/dsum per atom
 *   k_pairs    pass 2: GB + Coulomb + LJ energies, sumdeijda, direct forces
 *   k_chain    pass 3: sumdeijda * dR/dr chain-rule forces; the per-atom
 *              OBC factor is applied as the atoms are loaded, so the
 *              separate sumdeijda reduction of eff.c:2077-2129 disappears
 *   k_post     sum the energy slots; add the slot gradients into the
 *              caller's AoS array
 * Atoms are ordered per molecule by a k-d split into 64-atom leaf boxes.
 * Pair kernels: one 64-lane wavefront per tile pair; lane l holds atom l of
 * tile A, the non-empty columns (atoms of tile B with at least one partner)
 * are staged in LDS and visited in the rotated order k = (l + s) mod m, so
 * that lanes touch different columns each step; the pair is oriented at
 * list build so that the side with fewer non-empty atoms is the one
 * stepped.  j-side sums are accumulated in LDS (ds_add_f64) and flushed
 * with one FP64 atomic per atom per tile pair.  Tile pairs are processed in
 * descending step count, dealt to the grid in boustrophedon rounds (the
 * autotuner picks per pass between a persistent grid and up to one
 * workgroup per tile pair, which the hardware balances as workgroups end).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <type_traits>
#include <vector>

#include "engine.h"
#include "gb_math.h"
#include "neck_tables.h"

#define PAIR_BLOCKS_MAX 4096
#define MAX_LAUNCH (64 * 1024) /* workgroups per pair-kernel launch */
#define LJ_LDS 256 /* LJ table entries pass 2 keeps in LDS */
#define CHAIN_DEPTH 8  /* derivative-store loads in flight per lane in pass 3 */
#define SPLIT_MAX 16   /* step chunks per tile pair, at most (a tile pair has at most 64 steps) */
#define SPLIT_DEFAULT_MAX 8 /* the most the default split uses */

/* ------------------------------------------------------------------ errors */

static thread_local char g_err[512] = "";

int seterr(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char *gbhip_last_error(void) { return g_err; }

extern "C" int gbhip_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

/* ---------------------------------------------------------- pair-list build */

__global__ __launch_bounds__(64) void k_bbox(DevPtrs d)
{
    const int T = blockIdx.x, l = threadIdx.x;
    const int s = T * TILE + l;
    const bool ok = d.perm[s] >= 0;
    double lo[3], hi[3];
    lo[0] = ok ? d.xs[s] : 1e300; hi[0] = ok ? d.xs[s] : -1e300;
    lo[1] = ok ? d.ys[s] : 1e300; hi[1] = ok ? d.ys[s] : -1e300;
    lo[2] = ok ? d.zs[s] : 1e300; hi[2] = ok ? d.zs[s] : -1e300;
    for (int off = 32; off > 0; off >>= 1) {
        for (int k = 0; k < 3; k++) {
            lo[k] = fmin(lo[k], __shfl_xor(lo[k], off));
            hi[k] = fmax(hi[k], __shfl_xor(hi[k], off));
        }
    }
    if (l == 0) {
        for (int k = 0; k < 3; k++) {
            d.bblo[3 * T + k] = lo[k];
            d.bbhi[3 * T + k] = hi[k];
        }
    }
}

/* candidates of tile row I: tiles J >= I of the same molecule whose boxes
 * are closer than the cutoff; written in increasing J */
__global__ __launch_bounds__(256) void k_tilecand(DevPtrs d, GBPar p)
{
    const int I = blockIdx.x;
    const int jend = d.tile_end[I];
    __shared__ int s_cnt;
    __shared__ int s_warp[4];
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    double lo0 = d.bblo[3 * I], lo1 = d.bblo[3 * I + 1], lo2 = d.bblo[3 * I + 2];
    double hi0 = d.bbhi[3 * I], hi1 = d.bbhi[3 * I + 1], hi2 = d.bbhi[3 * I + 2];
    int *row = d.cand + d.rowbase[I];
    for (int base = I; base < jend; base += 256) {
        int J = base + threadIdx.x;
        bool take = false;
        if (J < jend) {
            double a = fmax(0.0, fmax(d.bblo[3 * J] - hi0, lo0 - d.bbhi[3 * J]));
            double b = fmax(0.0, fmax(d.bblo[3 * J + 1] - hi1, lo1 - d.bbhi[3 * J + 1]));
            double c = fmax(0.0, fmax(d.bblo[3 * J + 2] - hi2, lo2 - d.bbhi[3 * J + 2]));
            take = a * a + b * b + c * c < p.cut2;
        }
        unsigned long long bal = __ballot(take);
        int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
        if (ln == 0) s_warp[wv] = __popcll(bal);
        __syncthreads();
        int off = s_cnt;
        for (int w = 0; w < wv; w++) off += s_warp[w];
        off += __popcll(bal & ((1ull << ln) - 1ull));
        if (take) row[off] = J;
        __syncthreads();
        if (threadIdx.x == 0) s_cnt += s_warp[0] + s_warp[1] + s_warp[2] + s_warp[3];
        __syncthreads();
    }
    if (threadIdx.x == 0) d.rowcnt[I] = s_cnt;
}

/* exclusive scan of the row counts (single block) and total */
__global__ __launch_bounds__(1024) void k_rowscan(DevPtrs d)
{
    __shared__ int s_part[1024];
    const int n = d.ntiles, t = threadIdx.x;
    if (t < 2 * TILE) d.tbin[t] = 0;
    const int per = (n + 1023) / 1024;
    const int b = t * per, e = min(n, b + per);
    int sum = 0;
    for (int k = b; k < e; k++) sum += d.rowcnt[k];
    s_part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = t >= off ? s_part[t - off] : 0;
        __syncthreads();
        s_part[t] += v;
        __syncthreads();
    }
    int run = s_part[t] - sum;
    for (int k = b; k < e; k++) {
        d.rowstart[k] = run;
        run += d.rowcnt[k];
    }
    if (t == 1023) {
        *d.ntp = s_part[1023];
        *d.npairs = 0ull;
    }
}

__global__ __launch_bounds__(256) void k_compact(DevPtrs d)
{
    const int I = blockIdx.x;
    const int n = d.rowcnt[I], o = d.rowstart[I];
    const int *row = d.cand + d.rowbase[I];
    for (int k = threadIdx.x; k < n; k += 256) d.tp[o + k] = make_int2(I, row[k]);
}

/* exact membership bits: r^2 < cut^2 with the reference's unfused FP64
 * arithmetic (nblist.c:748-765), i != j, not both frozen */
__global__ __launch_bounds__(64) void k_masks(DevPtrs d, GBPar p)
{
    __shared__ double s_x[TILE], s_y[TILE], s_z[TILE];
    __shared__ int s_ok[TILE], s_frz[TILE];
    const int l = threadIdx.x;
    const int ntp = *d.ntp;
    unsigned long long cnt = 0;
    for (int t = blockIdx.x; t < ntp; t += gridDim.x) {
        const int2 ij = d.tp[t];
        const int i = ij.x * TILE + l, j = ij.y * TILE + l;
        const double xi = d.xs[i], yi = d.ys[i], zi = d.zs[i];
        const bool oki = d.perm[i] >= 0;
        const int fi = d.frz[i];
        __syncthreads();
        s_x[l] = d.xs[j]; s_y[l] = d.ys[j]; s_z[l] = d.zs[j];
        s_ok[l] = d.perm[j] >= 0;
        s_frz[l] = d.frz[j];
        __syncthreads();
        unsigned long long w = 0;
        if (oki) {
            const int b0 = ij.x == ij.y ? l + 1 : 0;
            for (int b = b0; b < TILE; b++) {
                double dx = __dsub_rn(xi, s_x[b]);
                double dy = __dsub_rn(yi, s_y[b]);
                double dz = __dsub_rn(zi, s_z[b]);
                double r2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
                bool in = s_ok[b] && !(fi && s_frz[b]) && r2 < p.cut2;
                w |= (unsigned long long)in << b;
            }
        }
        cnt += __popcll(w);
        /* orientation: the evaluation kernels step over the non-empty
         * columns only, so the side with fewer non-empty atoms is walked
         * and the other is held one atom per lane */
        unsigned long long cols = w;
        for (int off = 32; off > 0; off >>= 1) cols |= __shfl_xor(cols, off);
        const unsigned long long rows = __ballot(w != 0ull);
        const bool flip = __popcll(rows) < __popcll(cols);
        if (flip) {
            unsigned long long wt = 0ull;
            for (int b = 0; b < TILE; b++) {
                const unsigned long long c = __ballot((w >> b) & 1ull);
                wt = l == b ? c : wt;
            }
            w = wt;
        }
        d.mask[(size_t)t * TILE + l] = w;
        d.excl[(size_t)t * TILE + l] = 0ull;
        int np = __popcll(w);
        for (int off = 32; off > 0; off >>= 1) np += __shfl_xor(np, off);
        if (l == 0) {
            d.tnp[t] = np;
            d.tpo[t] = flip ? make_int2(ij.y, ij.x) : ij;
            d.tflip[t] = flip;
            d.tcol[t] = flip ? rows : cols;
        }
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    if (l == 0 && cnt) atomicAdd(d.npairs, cnt);
}

/* exclusion bits from the per-atom lists (only partners j > i count, as in
 * the upper-triangle loop of eff.c:1931-2025) */
__global__ __launch_bounds__(256) void k_excl(DevPtrs d)
{
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (a >= d.natoms) return;
    const int sa = d.inv[a];
    for (int k = d.ex_ptr[a]; k < d.ex_ptr[a + 1]; k++) {
        const int sb = d.inv[d.ex_j[k]];
        int s1 = sa, s2 = sb;
        int T1 = s1 / TILE, T2 = s2 / TILE;
        if (T1 > T2 || (T1 == T2 && (s1 & 63) > (s2 & 63))) {
            int t = s1; s1 = s2; s2 = t;
            t = T1; T1 = T2; T2 = t;
        }
        /* binary search for tile T2 in row T1 (sorted by J) */
        int lo = d.rowstart[T1], hi = lo + d.rowcnt[T1] - 1, found = -1;
        while (lo <= hi) {
            int mid = (lo + hi) >> 1;
            int J = d.tp[mid].y;
            if (J == T2) { found = mid; break; }
            if (J < T2) lo = mid + 1; else hi = mid - 1;
        }
        if (found >= 0) {
            const int r = d.tflip[found] ? (s2 & 63) : (s1 & 63);
            const int c = d.tflip[found] ? (s1 & 63) : (s2 & 63);
            atomicOr(&d.excl[(size_t)found * TILE + r], 1ull << c);
        }
    }
}

/*
 * Membership and exclusion words in step order.  Every evaluation pass has
 * lane l of a tile pair with m non-empty columns visit compacted column
 * (l mod m + s) mod m at step s; bit s of the lane's words is rewritten to
 * hold that column's bit, so the passes test (w >> s) & 1 without looking
 * the column up.  Bits of empty columns hold no listed pair and are dropped.
 */
__global__ __launch_bounds__(64) void k_steporder(DevPtrs d)
{
    __shared__ int s_pos[TILE];
    const int l = threadIdx.x;
    const int ntp = *d.ntp;
    for (int t = blockIdx.x; t < ntp; t += gridDim.x) {
        const unsigned long long cols = d.tcol[t];
        const int m = __popcll(cols);
        if (m == 0) continue;
        __syncthreads();
        if ((cols >> l) & 1ull) s_pos[__popcll(cols & ((1ull << l) - 1ull))] = l;
        __syncthreads();
        const size_t o = (size_t)t * TILE + l;
        const unsigned long long w = d.mask[o], e = d.excl[o];
        unsigned long long sw = 0ull, se = 0ull;
        int k = l % m;
        for (int st = 0; st < m; st++) {
            const int c = s_pos[k];
            sw |= ((w >> c) & 1ull) << st;
            se |= ((e >> c) & 1ull) << st;
            k = k + 1 == m ? 0 : k + 1;
        }
        d.mask[o] = sw;
        d.excl[o] = se;
    }
}

/*
 * Processing order of the evaluation passes: the non-empty tile pairs sorted
 * by their number of non-empty columns (the step count of a visit), largest
 * first, by a counting sort over 64 bins.  Workgroups take tile pairs from a
 * shared queue in this order, so the last ones taken are the shortest.
 */
__global__ __launch_bounds__(256) void k_bincount(DevPtrs d)
{
    __shared__ int h[TILE];
    const int ntp = *d.ntp;
    if (threadIdx.x < TILE) h[threadIdx.x] = 0;
    __syncthreads();
    for (int t = blockIdx.x * 256 + threadIdx.x; t < ntp; t += gridDim.x * 256) {
        const int m = __popcll(d.tcol[t]);
        if (m) atomicAdd(&h[TILE - m], 1);
    }
    __syncthreads();
    if (threadIdx.x < TILE && h[threadIdx.x]) atomicAdd(&d.tbin[threadIdx.x], h[threadIdx.x]);
}

__global__ __launch_bounds__(64) void k_binscan(DevPtrs d)
{
    if (threadIdx.x == 0) {
        int off = 0;
        for (int b = 0; b < TILE; b++) {
            d.tbin[TILE + b] = off;
            off += d.tbin[b];
        }
        *d.nnz = off;
        *d.store_need = 0; /* k_binscatter reserves the store regions from here */
    }
}

/* same grid as k_bincount: each block reserves its slice of every bin once */
__global__ __launch_bounds__(256) void k_binscatter(DevPtrs d)
{
    __shared__ int h[TILE], base[TILE];
    const int ntp = *d.ntp;
    for (int t0 = blockIdx.x * 256; t0 < ntp; t0 += gridDim.x * 256) {
        if (threadIdx.x < TILE) h[threadIdx.x] = 0;
        __syncthreads();
        const int t = t0 + threadIdx.x;
        const int m = t < ntp ? __popcll(d.tcol[t]) : 0;
        int r = 0;
        if (m) r = atomicAdd(&h[TILE - m], 1);
        __syncthreads();
        if (threadIdx.x < TILE && h[threadIdx.x])
            base[threadIdx.x] = atomicAdd(&d.tbin[TILE + threadIdx.x], h[threadIdx.x]);
        __syncthreads();
        /* the tile pair's derivative-store region (one entry per listed
         * pair): a wavefront's tile pairs take consecutive regions reserved
         * by one atomic; any disjoint placement serves the passes */
        const long long np = m ? d.tnp[t] : 0;
        long long incl = np;
        for (int off = 1; off < 64; off <<= 1) {
            const long long v = __shfl_up(incl, off);
            if ((threadIdx.x & 63) >= off) incl += v;
        }
        const long long tot = __shfl(incl, 63);
        long long rbase = 0;
        if ((threadIdx.x & 63) == 63 && tot)
            rbase = (long long)atomicAdd((unsigned long long *)d.store_need, (unsigned long long)tot);
        rbase = __shfl(rbase, 63);
        if (m) {
            const int q = base[TILE - m] + r;
            d.tord[q] = t;
            QMeta mq;
            mq.a = d.tpo[t].x;
            mq.b = d.tpo[t].y;
            mq.t = t;
            mq.mol = d.tile_mol[mq.a];
            mq.cols = d.tcol[t];
            mq.start = rbase + incl - np;
            d.qm[q] = mq;
        }
        __syncthreads();
    }
}

/* whether the store allocated holds the list just built (strictly less:
 * pass 3 may read one entry past a region's last) */
__global__ __launch_bounds__(64) void k_storeok(DevPtrs d)
{
    if (threadIdx.x == 0) *d.store_ok = d.dstore != nullptr && *d.store_need < d.store_cap;
}

__global__ __launch_bounds__(64) void k_setflag(int *flag, int v)
{
    if (threadIdx.x == 0) *flag = v;
}

/*
 * Static assignment of the sorted tile pairs to a grid of G workgroups, in
 * boustrophedon rounds: round r gives workgroup b the item r*G + b for even
 * r and r*G + G-1-b for odd r.  Over a descending cost order this pairs each
 * workgroup's expensive items with cheap ones (longest-processing-time
 * balance) without a shared counter; with G at least the number of items
 * every workgroup has one item or none.
 */
__device__ __forceinline__ int snake_item(int r)
{
    const int G = gridDim.x, b = blockIdx.x;
    return r * G + ((r & 1) ? (G - 1 - b) : b);
}

/*
 * Work item q of a pair pass: the tile pair at queue position q >> sh, and
 * of its m steps the chunk c = q mod 2^sh, steps [m c / 2^sh, m (c+1) / 2^sh)
 * (sh = 0: one item per tile pair, all steps).  Chunks of one tile pair are
 * consecutive items, so they go to different workgroups.
 */
__device__ __forceinline__ void item_steps(int q, int sh, int m, int &sb, int &se)
{
    const int c = q & ((1 << sh) - 1);
    sb = (m * c) >> sh;
    se = (m * (c + 1)) >> sh;
}

/* listed pairs of all lanes at the steps before sb (derivative-store
 * entries that precede the chunk's first step) */
__device__ __forceinline__ int entries_before(unsigned long long mr, int sb)
{
    int n = __popcll(mr & ((1ull << sb) - 1ull));
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off);
    return n;
}

/* ------------------------------------------------------------ evaluation */

__global__ __launch_bounds__(256) void k_gather(DevPtrs d, GBPar p, const double *__restrict__ x)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s < p.nslots) slot_gather(d, s, x);
}

template <typename R> __device__ __forceinline__ R ldri0(const DevPtrs &d, int s);
template <> __device__ __forceinline__ double ldri0<double>(const DevPtrs &d, int s) { return d.ri0[s]; }
template <> __device__ __forceinline__ float ldri0<float>(const DevPtrs &d, int s) { return d.ri0_f[s]; }
template <typename R> __device__ __forceinline__ R ldsj(const DevPtrs &d, int s);
template <> __device__ __forceinline__ double ldsj<double>(const DevPtrs &d, int s) { return d.sj[s]; }
template <> __device__ __forceinline__ float ldsj<float>(const DevPtrs &d, int s) { return d.sj_f[s]; }
template <typename R> __device__ __forceinline__ R ldrb(const DevPtrs &d, int s);
template <> __device__ __forceinline__ double ldrb<double>(const DevPtrs &d, int s) { return d.rb[s]; }
template <> __device__ __forceinline__ float ldrb<float>(const DevPtrs &d, int s) { return d.rb_f[s]; }
template <typename R> __device__ __forceinline__ const R *neckpos(const DevPtrs &d);
template <> __device__ __forceinline__ const double *neckpos<double>(const DevPtrs &d) { return d.neck_pos; }
template <> __device__ __forceinline__ const float *neckpos<float>(const DevPtrs &d) { return d.neck_pos_f; }
template <typename R> __device__ __forceinline__ const R *neckval(const DevPtrs &d);
template <> __device__ __forceinline__ const double *neckval<double>(const DevPtrs &d) { return d.neck_val; }
template <> __device__ __forceinline__ const float *neckval<float>(const DevPtrs &d) { return d.neck_val_f; }

template <typename R> struct Vec2;
template <> struct Vec2<double> { typedef double2 T; };
template <> struct Vec2<float> { typedef float2 T; };

/*
 * Derivative store: pass 1 writes, for a listed pair, the two pass-3
 * derivative numerators (datmp of i by j, datmp of j by i; zero when the
 * pair is beyond rgbmax + Fsmax), and pass 3 reads them back instead of
 * recomputing the descreening case split, logarithms and reciprocals.  A
 * pair whose two directions are both the Taylor case without a neck term
 * gets no entry (pass 3 recomputes those few FMAs), except in a queued
 * tile pair, where every listed pair has one.  Each tile pair owns a
 * region as large as its listed pairs (reserved at list build); a work
 * item packs its entries from the region's start plus the listed pairs of
 * the steps before its first, step by step, the entries of one step in
 * lane order (ballot and lane prefix count), and records per step the
 * 64-bit mask of the lanes that wrote one (sbal).
 */
__device__ __forceinline__ int lanes_below(unsigned long long bal)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
}

/* a 64-bit value of lane n (n the same in every lane) */
__device__ __forceinline__ unsigned long long readlane64(unsigned long long v, int n)
{
    const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)v, n);
    const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(v >> 32), n);
    return (unsigned long long)hi << 32 | lo;
}

/*
 * Deterministic mode: a visit of tile pair t leaves the sums of its lane
 * tile A and of its column tile B in two rows of DET_NC * 64 doubles (side
 * 0 belongs to tile I of t = (I, J), side 1 to tile J; tflip says which of
 * them is A), and k_det_sum adds the rows up per atom in a fixed order.
 * Lane l's element of component c of a side: [c * TILE].
 */
__device__ __forceinline__ double *det_side(const DevPtrs &d, int t, int side)
{
    return d.part + ((size_t)t * 2 + side) * (DET_NC * TILE) + threadIdx.x;
}

/*
 * Kinds of a pass-1 visit.  The descreening's near cases C-E need
 * dij <= 4 sj and the neck term dij < rbi + rbj + 6.8 A (eff.c:1680-1708);
 * most tile pairs at a 25 A cutoff lie farther apart than either bound, and
 * for them the sweep is compiled without those branches and the extra
 * reciprocal and selects they need (BORN_FAR, or BORN_NECK when only the
 * neck range is reachable).  The kind is decided per visit from the current
 * coordinates: the box of the visit's column atoms, and for every lane with
 * a listed pair its distance to that box, a lower bound of its pair
 * distances; a pair can only be near if that bound is.  The margin covers
 * the FP32 rounding of coordinates and of dij; results equal those of the
 * full sweep.
 */
#define BORN_NEAR 0
#define BORN_NECK 1
#define BORN_FAR 2
#define BORN_MARGIN 0.05f
#ifndef BORN_QFILL
#define BORN_QFILL 0.7f /* pass 1: queue visits filled below this */
#endif

/*
 * Wavefront minimum and maximum of a float (every lane active), on keys
 * that order as the floats do when compared as signed integers (the
 * magnitude bits of negative values flipped): DPP min/max within rows of
 * 16 lanes (quad swaps, half-row and row mirrors), then the four row
 * results in scalar registers.
 */
__device__ __forceinline__ int fkey(float f)
{
    const int b = __float_as_int(f);
    return b ^ ((b >> 31) & 0x7fffffff);
}
__device__ __forceinline__ float fkey_value(int k) { return __int_as_float(k ^ ((k >> 31) & 0x7fffffff)); }
template <int CTRL> __device__ __forceinline__ int dpp_i(int v)
{
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}
template <bool MAX> __device__ __forceinline__ float wave_ext(float f)
{
    int v = fkey(f);
    v = MAX ? max(v, dpp_i<0xB1>(v)) : min(v, dpp_i<0xB1>(v));   /* quad_perm [1,0,3,2] */
    v = MAX ? max(v, dpp_i<0x4E>(v)) : min(v, dpp_i<0x4E>(v));   /* quad_perm [2,3,0,1] */
    v = MAX ? max(v, dpp_i<0x141>(v)) : min(v, dpp_i<0x141>(v)); /* row_half_mirror */
    v = MAX ? max(v, dpp_i<0x140>(v)) : min(v, dpp_i<0x140>(v)); /* row_mirror */
    const int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return fkey_value(MAX ? max(max(r0, r1), max(r2, r3)) : min(min(r0, r1), min(r2, r3)));
}

__device__ __forceinline__ int born_visit_kind(bool row, bool col, double xi, double yi, double zi, float si,
                                               float rbi, double xj, double yj, double zj, float sj, float rbj,
                                               bool has_neck)
{
    /* the box of the column atoms and their largest sj and rb */
    const float fxj = (float)xj, fyj = (float)yj, fzj = (float)zj;
    const float lox = wave_ext<false>(col ? fxj : 3e30f), hix = wave_ext<true>(col ? fxj : -3e30f);
    const float loy = wave_ext<false>(col ? fyj : 3e30f), hiy = wave_ext<true>(col ? fyj : -3e30f);
    const float loz = wave_ext<false>(col ? fzj : 3e30f), hiz = wave_ext<true>(col ? fzj : -3e30f);
    const float smax = wave_ext<true>(col ? sj : 0.0f), rbmax = wave_ext<true>(col ? rbj : 0.0f);
    const float gx = fmaxf(0.0f, fmaxf(lox - (float)xi, (float)xi - hix));
    const float gy = fmaxf(0.0f, fmaxf(loy - (float)yi, (float)yi - hiy));
    const float gz = fmaxf(0.0f, fmaxf(loz - (float)zi, (float)zi - hiz));
    const float dist = sqrtf(gx * gx + gy * gy + gz * gz);
    const bool near = row && dist < 4.0f * fmaxf(si, smax) + BORN_MARGIN;
    const bool neck = row && has_neck && dist < rbi + rbmax + (float)NECK_CUT + BORN_MARGIN;
    if (__ballot(near)) return BORN_NEAR;
    return __ballot(neck) ? BORN_NECK : BORN_FAR;
}

/*
 * Pass 1 (eff.c:1566-1733): for every listed pair, the descreening of i by
 * j (added to i's sum) and of j by i (added to j's sum).  The r^2 test
 * against (rgbmax+Fsmax)^2 and the FP64 sums are kept in both variants.
 * When the derivative store is large enough for the current list, the
 * pass-3 numerators of both directions are written as well (see the store
 * above).  Each visit runs one of three compiled sweeps by its kind
 * (born_visit_kind), or, for a tile pair of low fill, the queued form.
 */
template <typename R, bool SPL>
__global__ __launch_bounds__(64) void k_born(DevPtrs d, GBPar p)
{
    typedef typename Vec2<R>::T R2;
    __shared__ double s_x[TILE], s_y[TILE], s_z[TILE], s_acc[TILE];
    __shared__ R s_ri[TILE], s_ri1i[TILE], s_sj[TILE], s_rb[TILE];
    __shared__ int s_nk[TILE]; /* neck index n | (n * NECK_N) << 8 */
    /* the neck tables are read through the cache: in LDS they took the
     * space of a sixth wavefront per SIMD */
    const R *__restrict__ npos = neckpos<R>(d);
    const R *__restrict__ nval = neckval<R>(d);
    /* queued visits: row sums, and a ring of listed pairs (row lane | column
     * << 6 | store offset << 12) */
    __shared__ double s_row[TILE];
    __shared__ int s_q[2 * TILE];
    const int l = threadIdx.x;
    const int nnz = *d.nnz;
    const bool store = *d.store_ok;
    R2 *st = (R2 *)d.dstore;
    const R rgbmax = (R)p.rgbmax, rgbmax1i = (R)p.rgbmax1i, rgbmax2i = (R)p.rgbmax2i;
    const R neckscale = (R)p.neckscale, neckcut = (R)NECK_CUT;
    /* without a split (SPL false) the chunk is the whole sweep, folded at compile time */
    const int sh = SPL ? p.split_sh[0] : 0, nit = nnz << sh;
    QMeta nxt;
    if (snake_item(0) < nit) nxt = d.qm[snake_item(0) >> sh];
    for (int r = 0, q = snake_item(0); q < nit; q = snake_item(++r)) {
        /* this visit's metadata was loaded one visit ahead */
        const QMeta mq = nxt;
        if (snake_item(r + 1) < nit) nxt = d.qm[snake_item(r + 1) >> sh];
        const int t = mq.t;
        const unsigned long long cols = mq.cols;
        const int m = __popcll(cols);
        int sb, se;
        item_steps(q, sh, m, sb, se);
        if (sb == se) continue;
        const int2 ab = make_int2(mq.a, mq.b);
        const int i = ab.x * TILE + l, j = ab.y * TILE + l;
        const double xi = d.xs[i], yi = d.ys[i], zi = d.zs[i];
        const double rp2 = d.rgbpsmax2[mq.mol];
        const R ri = ldri0<R>(d, i), si = ldsj<R>(d, i), rbi = ldrb<R>(d, i);
        const R ri1i = R(1) / ri;
        const int ni = d.neck[i], niN = ni * NECK_N;
        const unsigned long long mr = d.mask[(size_t)t * TILE + l];
        const bool own = (cols >> l) & 1ull;
        const int pos = __popcll(cols & ((1ull << l) - 1ull));
        R2 *sp = store ? st + mq.start + (sb ? entries_before(mr, sb) : 0) : nullptr;
        double xj = 0.0, yj = 0.0, zj = 0.0;
        R sjj = R(0), rbj = R(0);
        __syncthreads();
        if (own) {
            xj = d.xs[j]; yj = d.ys[j]; zj = d.zs[j];
            s_x[pos] = xj; s_y[pos] = yj; s_z[pos] = zj;
            const R rj = ldri0<R>(d, j);
            sjj = ldsj<R>(d, j); rbj = ldrb<R>(d, j);
            s_ri[pos] = rj; s_ri1i[pos] = R(1) / rj;
            s_sj[pos] = sjj; s_rb[pos] = rbj;
            s_nk[pos] = d.neck[j] | (d.neck[j] * NECK_N) << 8;
            s_acc[pos] = 0.0;
        }
        const int kind = born_visit_kind(mr != 0ull, own, xi, yi, zi, (float)si, (float)rbi, xj, yj, zj,
                                         (float)sjj, (float)rbj, p.has_neck != 0);
        /* the tile pair's fill, listed pairs over lane-steps (of the whole
         * tile pair, so that every chunk of it, whatever the split, takes
         * the same form and computes the same FP32 terms) */
        int npt = __popcll(mr);
        for (int off = 32; off > 0; off >>= 1) npt += __shfl_xor(npt, off);
        const bool use_queue = (float)npt < p.qfill[0] * (float)(TILE * m);
        s_row[l] = 0.0;
        __syncthreads();
        double acc = 0.0;
        unsigned long long mybal = 0ull; /* lane l: the store mask of step sb + l */
        /* one pair's terms: row atom (lane li's) against column k; the
         * column sums are kept with the opposite sign (one LDS add without
         * a negation) and negated at the flush */
        auto pair = [&](auto kind_c, double px, double py, double pz, R pri, R pri1i, R psi, R prb, int pni, int k,
                        R &tij, R &aij, R &aji, bool &plain) {
            constexpr int KIND = decltype(kind_c)::value;
            constexpr bool NEAR = KIND == BORN_NEAR, NECK = KIND != BORN_FAR;
            R tji = R(0);
            tij = aij = aji = R(0);
            plain = false;
            const double dx = px - s_x[k], dy = py - s_y[k], dz = pz - s_z[k];
            const double r2d = dx * dx + dy * dy + dz * dz;
            if (r2d <= rp2) {
                const R r2 = (R)r2d;
                const R dij1i = gbm<R>::rsqrt_(r2);
                const R dij = r2 * dij1i;
                const R rj = s_ri[k], sj = s_sj[k];
                bool skip_ij, skip_ji, b_ij, b_ji;
                tij = descreen_dc<R, NEAR>(pri, pri1i, sj, r2, dij, dij1i, rgbmax, rgbmax1i, rgbmax2i, skip_ij, b_ij,
                                           aij);
                tji = descreen_dc<R, NEAR>(rj, s_ri1i[k], psi, r2, dij, dij1i, rgbmax, rgbmax1i, rgbmax2i, skip_ji,
                                           b_ji, aji);
                plain = b_ij && b_ji;
                if (NECK && p.has_neck) {
                    const int nk = s_nk[k], nij = pni * NECK_N + (nk & 255), nji = (nk >> 8) + pni;
                    if (dij < prb + s_rb[k] + neckcut) {
                        R dn;
                        plain = false;
                        if (!skip_ij) {
                            tij += neck_d<R>(dij, dij1i, npos[nij], nval[nij], neckscale, dn);
                            aij += dn;
                        }
                        if (!skip_ji) {
                            tji += neck_d<R>(dij, dij1i, npos[nji], nval[nji], neckscale, dn);
                            aji += dn;
                        }
                    }
                }
                atomicAdd(&s_acc[k], (double)tji);
            }
        };
        /* the step loop, compiled for the three visit kinds */
        auto sweep = [&](auto kind_c) {
        int k = (l + sb) % m;
        for (int s = sb; s < se; s++) {
            R aij = R(0), aji = R(0), tij;
            bool plain = false; /* both directions case B, no neck term: pass 3 recomputes */
            const bool listed = (mr >> s) & 1ull;
            if (listed) {
                pair(kind_c, xi, yi, zi, ri, ri1i, si, rbi, ni, k, tij, aij, aji, plain);
                acc -= (double)tij;
            }
            if (store) {
                const bool need = listed && !plain;
                const unsigned long long bal = __ballot(need);
                if (need) {
                    R2 v;
                    v.x = aij;
                    v.y = aji;
                    sp[lanes_below(bal)] = v;
                }
                sp += __popcll(bal);
                mybal = l == s - sb ? bal : mybal;
            }
            k = k + 1 == m ? 0 : k + 1;
        }
        };
        /* the listed pairs queued in step order and evaluated 64 at a time,
         * one per lane (row data from its lane by cross-lane reads, row sums
         * in LDS); every listed pair has a store entry */
        auto queued = [&](auto kind_c) {
            auto run = [&](int h, int n) {
                const int c = s_q[(h + l) & (2 * TILE - 1)];
                const int li = c & 63, k = (c >> 6) & 63, po = c >> 12;
                const double px = __shfl(xi, li), py = __shfl(yi, li), pz = __shfl(zi, li);
                const R pri = __shfl(ri, li), pri1i = __shfl(ri1i, li), psi = __shfl(si, li), prb = __shfl(rbi, li);
                const int pni = __shfl(ni, li);
                if (l < n) {
                    R tij, aij, aji;
                    bool plain;
                    pair(kind_c, px, py, pz, pri, pri1i, psi, prb, pni, k, tij, aij, aji, plain);
                    atomicAdd(&s_row[li], -(double)tij);
                    if (store) {
                        R2 v;
                        v.x = aij;
                        v.y = aji;
                        sp[po] = v;
                    }
                }
            };
            int k = (l + sb) % m, qh = 0, qn = 0, off = 0;
            for (int s = sb; s < se; s++) {
                const bool listed = (mr >> s) & 1ull;
                const unsigned long long bal = __ballot(listed);
                if (listed) {
                    const int lb = lanes_below(bal);
                    s_q[(qh + qn + lb) & (2 * TILE - 1)] = l | k << 6 | (off + lb) << 12;
                }
                if (store) mybal = l == s - sb ? bal : mybal;
                off += __popcll(bal);
                qn += __popcll(bal);
                if (qn >= TILE) {
                    __syncthreads();
                    run(qh, TILE);
                    __syncthreads();
                    qh = (qh + TILE) & (2 * TILE - 1);
                    qn -= TILE;
                }
                k = k + 1 == m ? 0 : k + 1;
            }
            __syncthreads();
            if (qn) run(qh, qn);
            __syncthreads();
            acc = s_row[l];
        };
        if (use_queue) {
            if (kind == BORN_FAR) queued(std::integral_constant<int, BORN_FAR>{});
            else if (kind == BORN_NECK) queued(std::integral_constant<int, BORN_NECK>{});
            else queued(std::integral_constant<int, BORN_NEAR>{});
        } else {
            if (kind == BORN_FAR) sweep(std::integral_constant<int, BORN_FAR>{});
            else if (kind == BORN_NECK) sweep(std::integral_constant<int, BORN_NECK>{});
            else sweep(std::integral_constant<int, BORN_NEAR>{});
        }
        if (store && l < se - sb) d.sbal[(size_t)t * TILE + sb + l] = mybal;
        __syncthreads();
        if (p.det) {
            const int fl = d.tflip[t];
            *det_side(d, t, fl) = acc;
            *det_side(d, t, 1 - fl) = own ? -s_acc[pos] : 0.0;
        } else {
            if (acc != 0.0) unsafeAtomicAdd(&d.sumi[i], acc);
            if (own && s_acc[pos] != 0.0) unsafeAtomicAdd(&d.sumi[j], -s_acc[pos]);
        }
    }
}

/*
 * Pass 1 in FP64 with the logarithmic cases and the neck term deferred.  In
 * FP64 the log and the divisions of the descreening cases A, C, D and E,
 * and the division of the neck term, are long library sequences; at a 25 A
 * cutoff about one pair direction in five needs one of them (the pairs
 * near rgbmax fall in case A, the near pairs in C-E or within the neck
 * range, the rest are case B, the Taylor series, or skipped), so computed
 * in the sweep nearly every wavefront step would pay them for a few lanes.
 * Here each step computes the case-B directions at once and pushes every
 * other direction that contributes onto a per-workgroup queue in LDS (the
 * lane and column of the pair, r^2, 1/dij and where its derivative-store
 * entry goes); whenever 64 are queued the whole wavefront takes one each
 * and evaluates descreen_d and the neck term with every lane active, and
 * the rest at the end of the visit.  Sums and store entries are the same
 * as k_born's: the pair terms are the same functions of the same operands,
 * only the order of the FP64 additions into the atom sums differs.
 */
#define BQ_CAP 128
template <typename R> struct BornQueue {
    R r2[BQ_CAP], dij1i[BQ_CAP];
    R *out[BQ_CAP];        /* derivative-store component, or null */
    int code[BQ_CAP];      /* row lane | column << 6 | (1 << 12: direction j by i) */
};

/* per-visit LDS copies of the two tiles' atom data the queue needs */
template <typename R> struct BornTiles {
    R ri[TILE], ri1i[TILE], sj[TILE], rb[TILE]; /* column atoms */
    R r_ri[TILE], r_ri1i[TILE], r_sj[TILE], r_rb[TILE]; /* row atoms */
    int nk[TILE], r_nk[TILE]; /* neck indices of the column and row atoms */
};

/* the queued directions at positions [h, h + n) of the ring, one per lane */
template <typename R, bool NEAR = true, bool NECK = true>
__device__ __forceinline__ void born_queue_run(BornQueue<R> &Q, int h, int n, const BornTiles<R> &T, const R *npos,
                                               const R *nval, double *s_row, double *s_col, const GBPar &p)
{
    const int l = threadIdx.x;
    if (l >= n) return;
    const int e = (h + l) & (BQ_CAP - 1);
    const int code = Q.code[e], lane = code & 63, col = (code >> 6) & 63;
    const bool ji = code >> 12;
    /* direction i by j: the row atom descreened by the column atom's
     * sphere; direction j by i the other way round */
    const R ri = ji ? T.ri[col] : T.r_ri[lane], ri1i = ji ? T.ri1i[col] : T.r_ri1i[lane];
    const R sj = ji ? T.r_sj[lane] : T.sj[col];
    const R r2 = Q.r2[e], dij1i = Q.dij1i[e], dij = r2 * dij1i;
    bool skip;
    R dv;
    R v = descreen_d<R, NEAR>(ri, ri1i, sj, r2, dij, dij1i, (R)p.rgbmax, (R)p.rgbmax1i, (R)p.rgbmax2i, skip, dv);
    if (NECK && p.has_neck && !skip && dij < T.r_rb[lane] + T.rb[col] + (R)NECK_CUT) {
        const int nidx = ji ? T.nk[col] * NECK_N + T.r_nk[lane] : T.r_nk[lane] * NECK_N + T.nk[col];
        R dn;
        v += neck_d<R>(dij, dij1i, npos[nidx], nval[nidx], (R)p.neckscale, dn);
        dv += dn;
    }
    atomicAdd(ji ? &s_col[col] : &s_row[lane], -(double)v);
    if (Q.out[e]) *Q.out[e] = dv;
}

template <typename R, bool SPL>
__global__ __launch_bounds__(64) void k_born_q(DevPtrs d, GBPar p)
{
    typedef typename Vec2<R>::T R2;
    __shared__ double s_x[TILE], s_y[TILE], s_z[TILE], s_acc[TILE], s_row[TILE];
    __shared__ BornTiles<R> T;
    __shared__ R npos[NECK_N * NECK_N], nval[NECK_N * NECK_N];
    __shared__ BornQueue<R> Q;
    const int l = threadIdx.x;
    const int nnz = *d.nnz;
    const bool store = *d.store_ok;
    R2 *st = (R2 *)d.dstore;
    const R rgbmax = (R)p.rgbmax;
    if (p.has_neck) {
        for (int k = l; k < NECK_N * NECK_N; k += TILE) {
            npos[k] = neckpos<R>(d)[k];
            nval[k] = neckval<R>(d)[k];
        }
    }
    const int sh = SPL ? p.split_sh[0] : 0, nit = nnz << sh;
    for (int r = 0, q = snake_item(0); q < nit; q = snake_item(++r)) {
        const QMeta mq = d.qm[q >> sh];
        const int t = mq.t;
        const unsigned long long cols = mq.cols;
        const int m = __popcll(cols);
        int sb, se;
        item_steps(q, sh, m, sb, se);
        if (sb == se) continue;
        const int i = mq.a * TILE + l, j = mq.b * TILE + l;
        const double xi = d.xs[i], yi = d.ys[i], zi = d.zs[i];
        const double rp2 = d.rgbpsmax2[mq.mol];
        const R ri = ldri0<R>(d, i), si = ldsj<R>(d, i), rbi = ldrb<R>(d, i);
        const unsigned long long mr = d.mask[(size_t)t * TILE + l];
        const bool own = (cols >> l) & 1ull;
        const int pos = __popcll(cols & ((1ull << l) - 1ull));
        R2 *sp = store ? st + mq.start + (sb ? entries_before(mr, sb) : 0) : nullptr;
        __syncthreads();
        /* 1/Ri: the Newton-refined reciprocal in FP64, the IEEE quotient in
         * FP32 (what the sweep kernel k_born uses) */
        T.r_ri[l] = ri; T.r_ri1i[l] = sizeof(R) == sizeof(double) ? gbm<R>::rcp(ri) : R(1) / ri;
        T.r_sj[l] = si; T.r_rb[l] = rbi;
        T.r_nk[l] = d.neck[i];
        s_row[l] = 0.0;
        double xj = 0.0, yj = 0.0, zj = 0.0;
        R sjj = R(0), rbj = R(0);
        if (own) {
            xj = d.xs[j]; yj = d.ys[j]; zj = d.zs[j];
            s_x[pos] = xj; s_y[pos] = yj; s_z[pos] = zj;
            const R rj = ldri0<R>(d, j);
            sjj = ldsj<R>(d, j); rbj = ldrb<R>(d, j);
            T.ri[pos] = rj; T.ri1i[pos] = sizeof(R) == sizeof(double) ? gbm<R>::rcp(rj) : R(1) / rj;
            T.sj[pos] = sjj; T.rb[pos] = rbj;
            T.nk[pos] = d.neck[j];
            s_acc[pos] = 0.0;
        }
        /* the visit kinds of k_born: no near cases and/or no neck term where
         * the pair distances rule them out */
        const int kind = born_visit_kind(mr != 0ull, own, xi, yi, zi, (float)si, (float)rbi, xj, yj, zj,
                                         (float)sjj, (float)rbj, p.has_neck != 0);
        __syncthreads();
        double acc = 0.0;
        unsigned long long mybal = 0ull; /* lane l: the store mask of step sb + l */
        auto sweep = [&](auto kind_c) {
        constexpr int KIND = decltype(kind_c)::value;
        constexpr bool NEAR = KIND == BORN_NEAR, NECK = KIND != BORN_FAR;
        int qh = 0, qn = 0; /* ring head and length, the same in every lane */
        int k = (l + sb) % m;
        for (int s = sb; s < se; s++) {
            const bool listed = (mr >> s) & 1ull;
            R aij = R(0), aji = R(0), r2 = R(0), dij1i = R(0);
            bool defer_ij = false, defer_ji = false, plain = false;
            if (listed) {
                const double dx = xi - s_x[k], dy = yi - s_y[k], dz = zi - s_z[k];
                const double r2d = dx * dx + dy * dy + dz * dz;
                if (r2d <= rp2) {
                    r2 = (R)r2d;
                    dij1i = gbm<R>::rsqrt_(r2);
                    const R dij = r2 * dij1i;
                    const R sj = T.sj[k];
                    const DsCase cij = ds_case<R, NEAR>(ri, sj, dij, rgbmax);
                    const DsCase cji = ds_case<R, NEAR>(T.ri[k], si, dij, rgbmax);
                    const bool neck = NECK && p.has_neck && dij < rbi + T.rb[k] + (R)NECK_CUT;
                    defer_ij = !cij.skip && (cij.a || cij.c || cij.d || cij.e || neck);
                    defer_ji = !cji.skip && (cji.a || cji.c || cji.d || cji.e || neck);
                    plain = cij.b && cji.b && !defer_ij && !defer_ji;
                    const R dij2i = dij1i * dij1i;
                    R tij = R(0), tji = R(0);
                    if (cij.b && !defer_ij) {
                        const R sj2 = sj * sj, tmpsd = sj2 * dij2i;
                        const R dumbo = R(GB_TA) + tmpsd * (R(GB_TB) + tmpsd * (R(GB_TC) + tmpsd * (R(GB_TD) + tmpsd * R(GB_TDD))));
                        const R dumbd = R(GB_TE) + tmpsd * (R(GB_TF) + tmpsd * (R(GB_TG) + tmpsd * (R(GB_TH) + tmpsd * R(GB_THH))));
                        tij = sj * tmpsd * dij2i * dumbo;
                        aij = tmpsd * sj * dij2i * dij2i * dumbd;
                    }
                    if (cji.b && !defer_ji) {
                        const R si2 = si * si, tmpsd = si2 * dij2i;
                        const R dumbo = R(GB_TA) + tmpsd * (R(GB_TB) + tmpsd * (R(GB_TC) + tmpsd * (R(GB_TD) + tmpsd * R(GB_TDD))));
                        const R dumbd = R(GB_TE) + tmpsd * (R(GB_TF) + tmpsd * (R(GB_TG) + tmpsd * (R(GB_TH) + tmpsd * R(GB_THH))));
                        tji = si * tmpsd * dij2i * dumbo;
                        aji = tmpsd * si * dij2i * dij2i * dumbd;
                    }
                    acc -= (double)tij;
                    atomicAdd(&s_acc[k], -(double)tji);
                }
            }
            /* derivative-store entries: the immediate ones now, the
             * deferred ones when their queue entry runs */
            R2 *slot = nullptr;
            if (store) {
                /* entries of the pairs that are not plain case B; those
                 * pass 3 recomputes */
                const bool need = listed && !plain;
                const unsigned long long bal = __ballot(need);
                if (need) slot = sp + lanes_below(bal);
                sp += __popcll(bal);
                mybal = l == s - sb ? bal : mybal;
                if (need) {
                    if (!defer_ij) slot->x = aij;
                    if (!defer_ji) slot->y = aji;
                }
            }
            /* queue the deferred directions, running 64 at a time */
            for (int dir = 0; dir < 2; dir++) {
                const bool push = dir ? defer_ji : defer_ij;
                const unsigned long long bal = __ballot(push);
                if (push) {
                    const int e = (qh + qn + lanes_below(bal)) & (BQ_CAP - 1);
                    Q.r2[e] = r2;
                    Q.dij1i[e] = dij1i;
                    Q.out[e] = slot ? (dir ? &slot->y : &slot->x) : nullptr;
                    Q.code[e] = l | (k << 6) | (dir << 12);
                }
                qn += __popcll(bal);
                if (qn >= TILE) {
                    __syncthreads();
                    born_queue_run<R, NEAR, NECK>(Q, qh, TILE, T, npos, nval, s_row, s_acc, p);
                    __syncthreads();
                    qh = (qh + TILE) & (BQ_CAP - 1);
                    qn -= TILE;
                }
            }
            k = k + 1 == m ? 0 : k + 1;
        }
        __syncthreads();
        born_queue_run<R, NEAR, NECK>(Q, qh, qn, T, npos, nval, s_row, s_acc, p);
        __syncthreads();
        };
        if (kind == BORN_FAR) sweep(std::integral_constant<int, BORN_FAR>{});
        else if (kind == BORN_NECK) sweep(std::integral_constant<int, BORN_NECK>{});
        else sweep(std::integral_constant<int, BORN_NEAR>{});
        if (store && l < se - sb) d.sbal[(size_t)t * TILE + sb + l] = mybal;
        acc += s_row[l];
        if (p.det) {
            const int fl = d.tflip[t];
            *det_side(d, t, fl) = acc;
            *det_side(d, t, 1 - fl) = own ? s_acc[pos] : 0.0;
        } else {
            if (acc != 0.0) unsafeAtomicAdd(&d.sumi[i], acc);
            if (own && s_acc[pos] != 0.0) unsafeAtomicAdd(&d.sumi[j], s_acc[pos]);
        }
    }
}

/*
 * Per atom (FP64): effective radius (eff.c:1712-1727), the diagonal GB term
 * and its sumdeijda start value (eff.c:1884-1893), and the factor that turns
 * sumdeijda into the pass-3 weight (eff.c:2212-2231; 0 for frozen atoms).
 */
__global__ __launch_bounds__(64) void k_radii(DevPtrs d, GBPar p)
{
    const int s = blockIdx.x * 64 + threadIdx.x;
    const bool ok = s < p.nslots && d.perm[s] >= 0;
    double eself = 0.0;
    if (s < p.nslots) {
        double reff = 1.0, psi = 0.0, dfac = 0.0, sd = 0.0;
        if (ok) {
            const double ri = d.ri0[s], rb = d.rb[s];
            const double ri1i = 1.0 / ri;
            const double sumi = d.sumi[s];
            if (p.gb == 1) {
                reff = 1.0 / (ri1i + sumi);
                if (reff < 0.0) reff = 30.0;
                dfac = 1.0;
            } else {
                const double a = d.ga[s], b = d.gbb[s], g = d.gg[s];
                psi = -ri * sumi;
                const double th = tanh((a - b * psi + g * psi * psi) * psi);
                reff = 1.0 / (ri1i - th / rb);
                dfac = (a - 2.0 * b * psi + 3.0 * g * psi * psi) * (1.0 - th * th) * ri / rb;
            }
            if (!d.frz[s]) {
                const double qi = d.q[s];
                const double expmkf = exp(-GB_KSCALE * p.kappa * reff) / p.epsext;
                const double dielfac = 1.0 - expmkf;
                const double qi2h = 0.5 * qi * qi;
                const double qid2h = qi2h * dielfac;
                eself = -qid2h / reff;
                sd = qid2h - GB_KSCALE * p.kappa * qi2h * expmkf * reff;
            } else {
                dfac = 0.0;
            }
        }
        d.reff[s] = reff;
        d.psi[s] = psi;
        d.dfac[s] = dfac;
        d.sumdeijda[s] = sd;
    }
    /* a wave is one tile, so one molecule */
    for (int off = 32; off > 0; off >>= 1) eself += __shfl_xor(eself, off);
    if (threadIdx.x == 0 && p.det)
        d.pself[blockIdx.x] = eself;
    else if (threadIdx.x == 0 && eself != 0.0)
        unsafeAtomicAdd(&d.eslot[3 * (size_t)d.slot_mol[s] * ENE_SLOTS + (blockIdx.x & (ENE_SLOTS - 1))], eself);
}

/* one pair's pass-2 contributions, zeros when the lane has no listed pair */
struct PairOut {
    bool act;
    double epol, eel, evdw, sdi, sdj, gx, gy, gz;
};

/*
 * The pass-2 terms of k_pairs below for one pair, branch-free.  Instead of
 * a select per output, two operands are selected: the GB terms take a zero
 * charge product on a lane without a listed pair (every GB output is then
 * an exact zero; the GB arithmetic stays finite even at r = 0, the lane's
 * own diagonal entry), and the Coulomb and LJ terms take 1/r = 0 unless the
 * pair is listed and not excluded (their outputs, and their share of the
 * force, are then exact zeros whatever 1/r would have been).
 */
template <typename R>
__device__ __forceinline__ void pair_terms(PairOut &o, bool act, bool coul, double xi, double yi, double zi,
                                           double qi, R ri, R ci, double xj, double yj, double zj, double qj, R rj,
                                           R irj, double2 ab, R epsinv, R kks, bool screened)
{
    const double dxd = xi - xj, dyd = yi - yj, dzd = zi - zj;
    const double r2d = dxd * dxd + dyd * dyd + dzd * dzd;
    const R r2 = (R)r2d;
    const double qiqj_d = qi * qj;
    const R qiqj = act ? (R)qiqj_d : R(0);
    const R rb2 = ri * rj;
    /* exp(-r^2 / 4 Ri Rj) = 2^(r^2 ci / Rj) with ci = -log2(e) / 4 Ri of
     * the row and 1/Rj of the column staged once per visit; 1/fgbi =
     * sqrt(D) = D fgbi: no reciprocal per pair */
    const R efac = __builtin_amdgcn_exp2f(r2 * (ci * irj));
    const R dd = r2 + rb2 * efac;
    const R fgbi = gbm<R>::rsqrt_(dd);
    R expmkf, fgbk;
    if (screened) {
        fgbk = kks * dd * fgbi;
        expmkf = gbm<R>::exp_(fgbk) * epsinv;
    } else {
        fgbk = R(0);
        expmkf = epsinv;
    }
    const R dielfac = R(1) - expmkf;
    const R temp4 = fgbi * fgbi * fgbi;
    const R temp6 = qiqj * temp4 * (dielfac + fgbk * expmkf);
    const R de = temp6 * (R(1) - R(0.25) * efac);
    const R temp5 = R(0.5) * efac * temp6 * (rb2 + R(0.25) * r2);
    /* Coulomb and LJ in FP64 (see k_pairs) */
    const double rinv = coul ? rsqrt_pair<R>(r2d) : 0.0;
    const double r2inv = rinv * rinv;
    const double el = qiqj_d * rinv;
    const double r6inv = r2inv * r2inv * r2inv;
    const double f6 = ab.y * r6inv;
    const double f12 = ab.x * r6inv * r6inv;
    const double dcl = -(el + 12.0 * f12 - 6.0 * f6) * r2inv;
    const double ded = (double)de + dcl;
    o.act = act;
    o.epol = -(double)(qiqj * dielfac * fgbi);
    o.eel = el;
    o.evdw = f12 - f6;
    o.sdi = (double)(ri * temp5);
    o.sdj = (double)(rj * temp5);
    o.gx = ded * dxd;
    o.gy = ded * dyd;
    o.gz = ded * dzd;
}

/*
 * Pass 2 (eff.c:1768-2066): GB pair energy with f = sqrt(r^2 + RiRj
 * exp(-r^2/4RiRj)) and Debye-Hueckel screening, its sumdeijda contributions
 * to both atoms, and for non-excluded pairs Coulomb and 12-6 LJ; the direct
 * pair force de*xij goes to i (+) and j (-).
 */
/* one workgroup's pass-2 energies of molecule mol into its slots */
__device__ __forceinline__ void ene_flush(const DevPtrs &d, int mol, double epol, double evdw, double eel)
{
    double *e = d.eslot + 3 * (size_t)mol * ENE_SLOTS + (blockIdx.x & (ENE_SLOTS - 1));
    unsafeAtomicAdd(e, epol);
    unsafeAtomicAdd(e + ENE_SLOTS, evdw);
    unsafeAtomicAdd(e + 2 * ENE_SLOTS, eel);
}

template <typename R, bool SPL>
__global__ __launch_bounds__(64) void k_pairs(DevPtrs d, GBPar p)
{
    __shared__ double s_x[TILE], s_y[TILE], s_z[TILE];
    __shared__ double s_fx[TILE], s_fy[TILE], s_fz[TILE], s_sd[TILE];
    __shared__ double s_q[TILE];
    /* the molecule's LJ (A, B) table when it has at most LJ_LDS entries
     * (16 types): a per-pair global load leaves the dependency chain */
    __shared__ double2 s_tab[LJ_LDS];
    __shared__ R s_r[TILE], s_ir[TILE];
    __shared__ int s_lj[TILE];
    const int l = threadIdx.x;
    const int nnz = *d.nnz;
    const R epsinv = (R)(1.0 / p.epsext);
    const R kks = (R)(-GB_KSCALE * p.kappa);
    const bool screened = p.kappa != 0.0;
    double epol = 0.0, eel = 0.0, evdw = 0.0;
    int cur_mol = -1, lj0 = 0;
    bool tab_lds = false;
    /* without a split (SPL false) the chunk is the whole sweep, folded at compile time */
    const int sh = SPL ? p.split_sh[1] : 0, nit = nnz << sh;
    QMeta nxt;
    if (snake_item(0) < nit) nxt = d.qm[snake_item(0) >> sh];
    for (int r = 0, q = snake_item(0); q < nit; q = snake_item(++r)) {
        /* this visit's metadata was loaded one visit ahead */
        const QMeta mq = nxt;
        if (snake_item(r + 1) < nit) nxt = d.qm[snake_item(r + 1) >> sh];
        const int t = mq.t;
        const unsigned long long cols = mq.cols;
        const int m = __popcll(cols);
        int sb, se;
        item_steps(q, sh, m, sb, se);
        if (sb == se) continue;
        const int2 ab = make_int2(mq.a, mq.b);
        const int i = ab.x * TILE + l, j = ab.y * TILE + l;
        const int mol = mq.mol;
        if (mol != cur_mol) {
            if (cur_mol >= 0 && !p.det) {
                for (int off = 32; off > 0; off >>= 1) {
                    epol += __shfl_xor(epol, off);
                    eel += __shfl_xor(eel, off);
                    evdw += __shfl_xor(evdw, off);
                }
                if (l == 0) ene_flush(d, cur_mol, epol, evdw, eel);
                epol = eel = evdw = 0.0;
            }
            cur_mol = mol;
            const int2 mlj = d.mol_lj[mol];
            lj0 = mlj.x;
            tab_lds = mlj.y * mlj.y <= LJ_LDS;
            if (tab_lds) {
                __syncthreads(); /* the previous molecule's table is no longer read */
                for (int k = l; k < mlj.y * mlj.y; k += TILE) s_tab[k] = d.ljab[lj0 + k];
            }
        }
        const double xi = d.xs[i], yi = d.ys[i], zi = d.zs[i];
        const double qi = d.q[i];
        const R ri = (R)d.reff[i];
        const R ci = R(-1.44269504088896341) / (R(4) * ri);
        const int row = d.ljrow[i];
        const double2 *gtab = d.ljab;
        const unsigned long long mr = d.mask[(size_t)t * TILE + l];
        const unsigned long long e = d.excl[(size_t)t * TILE + l];
        const bool own = (cols >> l) & 1ull;
        const int pos = __popcll(cols & ((1ull << l) - 1ull));
        __syncthreads();
        if (own) {
            s_x[pos] = d.xs[j]; s_y[pos] = d.ys[j]; s_z[pos] = d.zs[j];
            s_q[pos] = d.q[j]; s_r[pos] = (R)d.reff[j]; s_ir[pos] = R(1) / (R)d.reff[j];
            s_lj[pos] = d.ljcol[j];
            s_fx[pos] = 0.0; s_fy[pos] = 0.0; s_fz[pos] = 0.0; s_sd[pos] = 0.0;
        }
        __syncthreads();
        double fxi = 0.0, fyi = 0.0, fzi = 0.0, sdi = 0.0;
        /* the column sweep, instantiated for the LJ table in LDS and in
         * global memory and chosen per visit (a per-pair choice compiles to
         * one generic-address load, slower than either) */
        auto sweep = [&](auto in_lds, auto scr) {
        constexpr bool TAB = decltype(in_lds)::value;
        constexpr bool SCR = decltype(scr)::value; /* Debye-Hueckel screening (kappa != 0) */
        auto lj = [&](int kk) -> double2 {
            if constexpr (TAB) return s_tab[kk - lj0];
            else return gtab[kk];
        };
        int k = (l + sb) % m;
        if constexpr (sizeof(R) == sizeof(float)) {
            /* mixed: two steps per iteration, branch-free, so the two pairs'
             * dependency chains interleave (the kernel is latency-bound at
             * its register-limited occupancy); lanes without a listed pair
             * contribute zeros chosen by select, never by multiplication (a
             * lane's own diagonal entry has r = 0).  Measured at 25k atoms:
             * 0.32 -> 0.30 ms; the FP64 kernel is faster as written below */
            for (int s = sb; s < se; s += 2) {
                const bool two = s + 1 < se;
                const int k1 = k + 1 == m ? 0 : k + 1;
                const int kk2[2] = {k, k1};
                PairOut P[2];
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const int kq = kk2[u];
                    const int b = (s + u) & 63;
                    const bool act = (u == 0 || two) && ((mr >> b) & 1ull);
                    const bool coul = act && !((e >> b) & 1ull);
                    pair_terms<R>(P[u], act, coul, xi, yi, zi, qi, ri, ci, s_x[kq], s_y[kq], s_z[kq], s_q[kq],
                                  s_r[kq], s_ir[kq], lj(row + s_lj[kq]), epsinv, kks, SCR);
                }
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const int kq = kk2[u];
                    epol += P[u].epol; eel += P[u].eel; evdw += P[u].evdw; sdi += P[u].sdi;
                    fxi += P[u].gx; fyi += P[u].gy; fzi += P[u].gz;
                    if (P[u].act) {
                        atomicAdd(&s_sd[kq], P[u].sdj);
                        atomicAdd(&s_fx[kq], P[u].gx);
                        atomicAdd(&s_fy[kq], P[u].gy);
                        atomicAdd(&s_fz[kq], P[u].gz);
                    }
                }
                k = two ? (k1 + 1 == m ? 0 : k1 + 1) : k1;
            }
        } else {
            for (int s = sb; s < se; s++) {
                const int b = s;
                if ((mr >> b) & 1ull) {
                    const double dxd = xi - s_x[k], dyd = yi - s_y[k], dzd = zi - s_z[k];
                    const double r2d = dxd * dxd + dyd * dyd + dzd * dzd;
                    const R r2 = (R)r2d;
                    const double qiqj_d = qi * s_q[k];
                    const R qiqj = (R)qiqj_d;
                    const R rj = s_r[k];
                    const R rb2 = ri * rj;
                    const R efac = gbm<R>::exp_(-r2 * gbm<R>::rcp(R(4) * rb2));
                    const R fgbi = gbm<R>::rsqrt_(r2 + rb2 * efac);
                    R expmkf, dielfac, fgbk;
                    if (SCR) {
                        fgbk = kks * gbm<R>::rcp(fgbi);
                        expmkf = gbm<R>::exp_(fgbk) * epsinv;
                    } else {
                        fgbk = R(0);
                        expmkf = epsinv;
                    }
                    dielfac = R(1) - expmkf;
                    epol -= (double)(qiqj * dielfac * fgbi);
                    const R temp4 = fgbi * fgbi * fgbi;
                    const R temp6 = qiqj * temp4 * (dielfac + fgbk * expmkf);
                    R de = temp6 * (R(1) - R(0.25) * efac);
                    const R temp5 = R(0.5) * efac * temp6 * (rb2 + R(0.25) * r2);
                    sdi += (double)(ri * temp5);
                    atomicAdd(&s_sd[k], (double)(rj * temp5));
                    double ded = (double)de;
                    if (!((e >> b) & 1ull)) {
                        /* Coulomb and LJ in FP64 in both variants: their molecular
                         * totals cancel strongly (sums of +-100 kcal/mol terms), so
                         * FP32 terms cannot hold 1e-6 of the total; they cost a few
                         * FP64 operations next to the FP32 transcendentals above */
                        const double rinv = rsqrt_pair<R>(r2d);
                        const double r2inv = rinv * rinv;
                        const double el = qiqj_d * rinv;
                        eel += el;
                        ded -= el * r2inv;
                        const double r6inv = r2inv * r2inv * r2inv;
                        const int kk = row + s_lj[k];
                        const double2 ab = lj(kk);
                        const double f6 = ab.y * r6inv;
                        const double f12 = ab.x * r6inv * r6inv;
                        evdw += f12 - f6;
                        ded -= (12.0 * f12 - 6.0 * f6) * r2inv;
                    }
                    const double gxv = ded * dxd, gyv = ded * dyd, gzv = ded * dzd;
                    fxi += gxv; fyi += gyv; fzi += gzv;
                    atomicAdd(&s_fx[k], gxv);
                    atomicAdd(&s_fy[k], gyv);
                    atomicAdd(&s_fz[k], gzv);
                }
                k = k + 1 == m ? 0 : k + 1;
            }
        }
        };
        /* instantiated per LJ table location and per screening, so that
         * neither is a per-pair select */
        auto sweep_scr = [&](auto in_lds) {
            if (screened) sweep(in_lds, std::true_type{});
            else sweep(in_lds, std::false_type{});
        };
        if (tab_lds) sweep_scr(std::true_type{});
        else sweep_scr(std::false_type{});
        __syncthreads();
        if (p.det) {
            /* the visit's energies per tile pair, its sums per side */
            for (int off = 32; off > 0; off >>= 1) {
                epol += __shfl_xor(epol, off);
                eel += __shfl_xor(eel, off);
                evdw += __shfl_xor(evdw, off);
            }
            if (l == 0) {
                d.pe[3 * (size_t)t] = epol;
                d.pe[3 * (size_t)t + 1] = evdw;
                d.pe[3 * (size_t)t + 2] = eel;
            }
            epol = eel = evdw = 0.0;
            const int fl = d.tflip[t];
            double *pa = det_side(d, t, fl), *pb = det_side(d, t, 1 - fl);
            pa[0] = sdi; pa[TILE] = fxi; pa[2 * TILE] = fyi; pa[3 * TILE] = fzi;
            pb[0] = own ? s_sd[pos] : 0.0;
            pb[TILE] = own ? -s_fx[pos] : 0.0;
            pb[2 * TILE] = own ? -s_fy[pos] : 0.0;
            pb[3 * TILE] = own ? -s_fz[pos] : 0.0;
        } else {
            if (mr) {
                unsafeAtomicAdd(&d.gx[i], fxi);
                unsafeAtomicAdd(&d.gy[i], fyi);
                unsafeAtomicAdd(&d.gz[i], fzi);
                unsafeAtomicAdd(&d.sumdeijda[i], sdi);
            }
            /* the column atoms' direct forces were summed with the row's
             * sign (one LDS add without a negation per pair) */
            if (own) {
                unsafeAtomicAdd(&d.gx[j], -s_fx[pos]);
                unsafeAtomicAdd(&d.gy[j], -s_fy[pos]);
                unsafeAtomicAdd(&d.gz[j], -s_fz[pos]);
                unsafeAtomicAdd(&d.sumdeijda[j], s_sd[pos]);
            }
        }
    }
    if (cur_mol >= 0 && !p.det) {
        for (int off = 32; off > 0; off >>= 1) {
            epol += __shfl_xor(epol, off);
            eel += __shfl_xor(eel, off);
            evdw += __shfl_xor(evdw, off);
        }
        if (l == 0) ene_flush(d, cur_mol, epol, evdw, eel);
    }
}

/*
 * Pass 3 (eff.c:2131-2405): chain rule through dR/dr.  For the pair (i,j)
 * the reference adds, from row i, -sumda_i*datmp_ij*xij to i and the
 * opposite to j, and from row j the same with i and j swapped; both rows
 * are done here in one visit: F = xij*(datmp_ij*w_i + datmp_ji*w_j),
 * grad_i -= F, grad_j += F, with w = sumdeijda * dR/dsum (0 if frozen).
 * Two kernels, both launched, exactly one of which does the work, as the
 * device flag of the current list says: k_chain_st reads datmp from the
 * derivative store written by pass 1 where an entry exists and recomputes
 * the Taylor-case ones; k_chain recomputes every datmp (the store could not
 * hold the list).
 */
template <typename R, bool SPL>
__global__ __launch_bounds__(64) void k_chain_st(DevPtrs d, GBPar p)
{
    typedef typename Vec2<R>::T R2;
    __shared__ double s_x[TILE], s_y[TILE], s_z[TILE];
    __shared__ double s_fx[TILE], s_fy[TILE], s_fz[TILE];
    __shared__ R s_w[TILE], s_s2[TILE], s_s3[TILE]; /* column sj^2, sj^3 */
    if (!*d.store_ok) return;
    const int l = threadIdx.x;
    const int nnz = *d.nnz;
    const R2 *st = (const R2 *)d.dstore;
    /* pass 1 packed the entries of each of its chunks (split 1 << sh1)
     * from that chunk's first step; this pass's chunks (split 1 << sh) may
     * be coarser or finer */
    const int sh = SPL ? p.split_sh[2] : 0, nit = nnz << sh;
    const int sh1 = p.split_sh[0];
    QMeta nxt;
    if (snake_item(0) < nit) nxt = d.qm[snake_item(0) >> sh];
    for (int r = 0, q = snake_item(0); q < nit; q = snake_item(++r)) {
        /* this visit's metadata was loaded one visit ahead */
        const QMeta mq = nxt;
        if (snake_item(r + 1) < nit) nxt = d.qm[snake_item(r + 1) >> sh];
        const int t = mq.t;
        const unsigned long long cols = mq.cols;
        const int m = __popcll(cols);
        int sb, se;
        item_steps(q, sh, m, sb, se);
        if (sb == se) continue;
        const int2 ab = make_int2(mq.a, mq.b);
        const int i = ab.x * TILE + l, j = ab.y * TILE + l;
        const double xi = d.xs[i], yi = d.ys[i], zi = d.zs[i];
        const R wi = (R)(d.sumdeijda[i] * d.dfac[i]);
        const R si = ldsj<R>(d, i), si2 = si * si, si3 = si2 * si;
        const unsigned long long mr = d.mask[(size_t)t * TILE + l];
        /* lane l: the store mask of step sb + l (which lanes have an entry) */
        const unsigned long long mybal = l < se - sb ? d.sbal[(size_t)t * TILE + sb + l] : 0ull;
        /* pass 1's chunk c1 holding step sb starts at step p1s; its entries
         * start at the listed pairs before p1s, and those of steps [p1s, sb)
         * precede sb's; nb1: the next pass-1 chunk's first step */
        int c1 = 0, nb1 = m;
        /* c1: the last pass-1 chunk starting at or before step ps (chunks
         * of a split larger than m can be empty and share a start) */
        auto chunk_at = [&](int ps) {
            while (c1 + 1 < (1 << sh1) && ((m * (c1 + 1)) >> sh1) <= ps) c1++;
            nb1 = c1 + 1 < (1 << sh1) ? (m * (c1 + 1)) >> sh1 : m;
        };
        chunk_at(sb);
        const int p1s = (m * c1) >> sh1;
        const R2 *sp = st + mq.start;
        if (p1s > 0 || sb > 0) {
            int skip = l < sb - p1s ? __popcll(d.sbal[(size_t)t * TILE + p1s + l]) : 0;
            for (int off = 32; off > 0; off >>= 1) skip += __shfl_xor(skip, off);
            sp += (p1s ? entries_before(mr, p1s) : 0) + skip;
        }
        const bool own = (cols >> l) & 1ull;
        const int pos = __popcll(cols & ((1ull << l) - 1ull));
        __syncthreads();
        if (own) {
            s_x[pos] = d.xs[j]; s_y[pos] = d.ys[j]; s_z[pos] = d.zs[j];
            s_w[pos] = (R)(d.sumdeijda[j] * d.dfac[j]);
            const R sjc = ldsj<R>(d, j);
            s_s2[pos] = sjc * sjc;
            s_s3[pos] = sjc * sjc * sjc;
            s_fx[pos] = 0.0; s_fy[pos] = 0.0; s_fz[pos] = 0.0;
        }
        __syncthreads();
        /* the stored entries of the next CHAIN_DEPTH steps are in flight
         * while a step is processed: sp walks ahead of k */
        const int k0 = (l + sb) % m;
        R2 win[CHAIN_DEPTH];
        bool act[CHAIN_DEPTH], sto[CHAIN_DEPTH];
#pragma unroll
        for (int u = 0; u < CHAIN_DEPTH; u++) {
            act[u] = sto[u] = false;
            if (sb + u < se) {
                /* every lane loads (lanes without an entry read the step's
                 * first position, in bounds: the store holds more than the
                 * list needs), so no branch separates a load from its use
                 * and the loads stay in flight */
                if (sb + u == nb1) {
                    sp = st + mq.start + entries_before(mr, nb1);
                    chunk_at(nb1);
                }
                const unsigned long long bal = readlane64(mybal, u);
                const bool e = (bal >> l) & 1ull;
                win[u] = sp[e ? lanes_below(bal) : 0];
                act[u] = (mr >> (sb + u)) & 1ull;
                sto[u] = e;
                sp += __popcll(bal);
            }
        }
        double fxi = 0.0, fyi = 0.0, fzi = 0.0;
        int k = k0;
        for (int s0 = sb; s0 < se; s0 += CHAIN_DEPTH) {
#pragma unroll
            for (int u = 0; u < CHAIN_DEPTH; u++) {
                const int s = s0 + u;
                if (s < se) {
                    const R2 a = win[u];
                    const bool on = act[u], stored = sto[u];
                    if (s + CHAIN_DEPTH < se) {
                        if (s + CHAIN_DEPTH == nb1) {
                            sp = st + mq.start + entries_before(mr, nb1);
                            chunk_at(nb1);
                        }
                        const unsigned long long bal = readlane64(mybal, s + CHAIN_DEPTH - sb);
                        const bool e = (bal >> l) & 1ull;
                        win[u] = sp[e ? lanes_below(bal) : 0];
                        act[u] = (mr >> (s + CHAIN_DEPTH)) & 1ull;
                        sto[u] = e;
                        sp += __popcll(bal);
                    }
                    if (on) {
                        const double dx = xi - s_x[k], dy = yi - s_y[k], dz = zi - s_z[k];
                        R aij = a.x, aji = a.y;
                        if (!stored) {
                            /* both directions plain case B (pass 1 wrote no
                             * entry): the Taylor derivative from 1/dij^2 as
                             * pass 1 forms it */
                            const R r2 = (R)(dx * dx + dy * dy + dz * dz);
                            const R dij1i = gbm<R>::rsqrt_(r2);
                            const R dij2i = dij1i * dij1i;
                            const R dij6i = dij2i * dij2i * dij2i;
                            aij = taylor_dv3<R>(s_s2[k], s_s3[k], dij2i, dij6i);
                            aji = taylor_dv3<R>(si2, si3, dij2i, dij6i);
                        }
                        const double f = (double)(aij * wi + aji * s_w[k]);
                        const double gxv = f * dx, gyv = f * dy, gzv = f * dz;
                        fxi -= gxv; fyi -= gyv; fzi -= gzv;
                        atomicAdd(&s_fx[k], gxv);
                        atomicAdd(&s_fy[k], gyv);
                        atomicAdd(&s_fz[k], gzv);
                    }
                    k = k + 1 == m ? 0 : k + 1;
                }
            }
        }
        __syncthreads();
        if (p.det) {
            const int fl = d.tflip[t];
            double *pa = det_side(d, t, fl), *pb = det_side(d, t, 1 - fl);
            pa[0] = fxi; pa[TILE] = fyi; pa[2 * TILE] = fzi;
            pb[0] = own ? s_fx[pos] : 0.0;
            pb[TILE] = own ? s_fy[pos] : 0.0;
            pb[2 * TILE] = own ? s_fz[pos] : 0.0;
        } else {
            if (fxi != 0.0 || fyi != 0.0 || fzi != 0.0) {
                unsafeAtomicAdd(&d.gx[i], fxi);
                unsafeAtomicAdd(&d.gy[i], fyi);
                unsafeAtomicAdd(&d.gz[i], fzi);
            }
            if (own && (s_fx[pos] != 0.0 || s_fy[pos] != 0.0 || s_fz[pos] != 0.0)) {
                unsafeAtomicAdd(&d.gx[j], s_fx[pos]);
                unsafeAtomicAdd(&d.gy[j], s_fy[pos]);
                unsafeAtomicAdd(&d.gz[j], s_fz[pos]);
            }
        }
    }
}

template <typename R, bool SPL>
__global__ __launch_bounds__(64) void k_chain(DevPtrs d, GBPar p)
{
    __shared__ double s_x[TILE], s_y[TILE], s_z[TILE];
    __shared__ double s_fx[TILE], s_fy[TILE], s_fz[TILE];
    __shared__ R s_ri[TILE], s_ri1i[TILE], s_sj[TILE], s_rb[TILE], s_w[TILE];
    __shared__ int s_nk[TILE];
    __shared__ R npos[NECK_N * NECK_N], nval[NECK_N * NECK_N];
    if (*d.store_ok) return;
    const int l = threadIdx.x;
    const int nnz = *d.nnz;
    const R rgbmax = (R)p.rgbmax, rgbmax1i = (R)p.rgbmax1i, rgbmax2i = (R)p.rgbmax2i;
    const R neckscale = (R)p.neckscale, neckcut = (R)NECK_CUT;
    if (p.has_neck) {
        for (int k = l; k < NECK_N * NECK_N; k += TILE) {
            npos[k] = neckpos<R>(d)[k];
            nval[k] = neckval<R>(d)[k];
        }
    }
    /* without a split (SPL false) the chunk is the whole sweep, folded at compile time */
    const int sh = SPL ? p.split_sh[2] : 0, nit = nnz << sh;
    QMeta nxt;
    if (snake_item(0) < nit) nxt = d.qm[snake_item(0) >> sh];
    for (int r = 0, q = snake_item(0); q < nit; q = snake_item(++r)) {
        /* this visit's metadata was loaded one visit ahead */
        const QMeta mq = nxt;
        if (snake_item(r + 1) < nit) nxt = d.qm[snake_item(r + 1) >> sh];
        const int t = mq.t;
        const unsigned long long cols = mq.cols;
        const int m = __popcll(cols);
        int sb, se;
        item_steps(q, sh, m, sb, se);
        if (sb == se) continue;
        const int2 ab = make_int2(mq.a, mq.b);
        const int i = ab.x * TILE + l, j = ab.y * TILE + l;
        const double xi = d.xs[i], yi = d.ys[i], zi = d.zs[i];
        const double rp2 = d.rgbpsmax2[mq.mol];
        const R ri = ldri0<R>(d, i), si = ldsj<R>(d, i), rbi = ldrb<R>(d, i);
        const R ri1i = R(1) / ri;
        const R wi = (R)(d.sumdeijda[i] * d.dfac[i]);
        const int ni = d.neck[i];
        const unsigned long long mr = d.mask[(size_t)t * TILE + l];
        const bool own = (cols >> l) & 1ull;
        const int pos = __popcll(cols & ((1ull << l) - 1ull));
        __syncthreads();
        if (own) {
            s_x[pos] = d.xs[j]; s_y[pos] = d.ys[j]; s_z[pos] = d.zs[j];
            const R rj = ldri0<R>(d, j);
            s_ri[pos] = rj; s_ri1i[pos] = R(1) / rj;
            s_sj[pos] = ldsj<R>(d, j); s_rb[pos] = ldrb<R>(d, j);
            s_w[pos] = (R)(d.sumdeijda[j] * d.dfac[j]);
            s_nk[pos] = d.neck[j];
            s_fx[pos] = 0.0; s_fy[pos] = 0.0; s_fz[pos] = 0.0;
        }
        __syncthreads();
        double fxi = 0.0, fyi = 0.0, fzi = 0.0;
        int k = (l + sb) % m;
        for (int s = sb; s < se; s++) {
            if ((mr >> s) & 1ull) {
                const double dxd = xi - s_x[k], dyd = yi - s_y[k], dzd = zi - s_z[k];
                const double r2d = dxd * dxd + dyd * dyd + dzd * dzd;
                if (r2d <= rp2) {
                    const R r2 = (R)r2d;
                    const R dij1i = gbm<R>::rsqrt_(r2);
                    const R dij = r2 * dij1i;
                    const R rj = s_ri[k], sj = s_sj[k];
                    bool skip_ij, skip_ji;
                    R aij = ddescreen<R>(ri, ri1i, sj, r2, dij, dij1i, rgbmax, rgbmax1i, rgbmax2i, skip_ij);
                    R aji = ddescreen<R>(rj, s_ri1i[k], si, r2, dij, dij1i, rgbmax, rgbmax1i, rgbmax2i, skip_ji);
                    if (p.has_neck) {
                        const int nj = s_nk[k];
                        if (dij < rbi + s_rb[k] + neckcut) {
                            if (!skip_ij) aij += dneck_term<R>(dij, dij1i, npos[ni * NECK_N + nj], nval[ni * NECK_N + nj], neckscale);
                            if (!skip_ji) aji += dneck_term<R>(dij, dij1i, npos[nj * NECK_N + ni], nval[nj * NECK_N + ni], neckscale);
                        }
                    }
                    const double f = (double)(aij * wi + aji * s_w[k]);
                    const double gxv = f * dxd, gyv = f * dyd, gzv = f * dzd;
                    fxi -= gxv; fyi -= gyv; fzi -= gzv;
                    atomicAdd(&s_fx[k], gxv);
                    atomicAdd(&s_fy[k], gyv);
                    atomicAdd(&s_fz[k], gzv);
                }
            }
            k = k + 1 == m ? 0 : k + 1;
        }
        __syncthreads();
        if (p.det) {
            const int fl = d.tflip[t];
            double *pa = det_side(d, t, fl), *pb = det_side(d, t, 1 - fl);
            pa[0] = fxi; pa[TILE] = fyi; pa[2 * TILE] = fzi;
            pb[0] = own ? s_fx[pos] : 0.0;
            pb[TILE] = own ? s_fy[pos] : 0.0;
            pb[2 * TILE] = own ? s_fz[pos] : 0.0;
        } else {
            if (fxi != 0.0 || fyi != 0.0 || fzi != 0.0) {
                unsafeAtomicAdd(&d.gx[i], fxi);
                unsafeAtomicAdd(&d.gy[i], fyi);
                unsafeAtomicAdd(&d.gz[i], fzi);
            }
            if (own && (s_fx[pos] != 0.0 || s_fy[pos] != 0.0 || s_fz[pos] != 0.0)) {
                unsafeAtomicAdd(&d.gx[j], s_fx[pos]);
                unsafeAtomicAdd(&d.gy[j], s_fy[pos]);
                unsafeAtomicAdd(&d.gz[j], s_fz[pos]);
            }
        }
    }
}

/*
 * Deterministic mode, after a list build: the column lists, one wavefront
 * per tile J finding, for every earlier tile I of its molecule, the tile
 * pair (I, J) in row I by binary search (rows are sorted by J) and writing
 * the ones present in increasing I.
 */
__global__ __launch_bounds__(64) void k_det_cols(DevPtrs d)
{
    const int J = blockIdx.x, l = threadIdx.x;
    const int I0 = d.mol_tile0[d.tile_mol[J]];
    int *col = d.dcol + d.rowbase[I0] + (size_t)(J - I0) * (J - I0 + 1) / 2;
    int n = 0;
    for (int b = I0; b < J; b += TILE) {
        const int I = b + l;
        int t = -1;
        if (I < J) {
            const int r0 = d.rowstart[I], r1 = r0 + d.rowcnt[I];
            int lo = r0, hi = r1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (d.tp[mid].y < J) lo = mid + 1;
                else hi = mid;
            }
            if (lo < r1 && d.tp[lo].y == J) t = lo;
        }
        const unsigned long long bal = __ballot(t >= 0);
        if (t >= 0) col[n + lanes_below(bal)] = t;
        n += __popcll(bal);
    }
    if (l == 0) d.dcolcnt[J] = n;
}

/*
 * Deterministic mode, after a pass: per slot of tile T, the visit sums its
 * tile received, added in a fixed order whichever workgroup made them and
 * when.  The contributions are listed as the tile pairs (T, J) of row T by
 * increasing J (side 0), then the tile pairs (I, T) by increasing I (side
 * 1), then the column side of (T, T); four wavefronts take fixed quarters
 * of the list, each adding its quarter in order, and the quarters are
 * joined in order.  mode 0: pass 1, into sumi; mode 1: pass 2, added to
 * sumdeijda (which holds the self term) and gx, gy, gz, and the row's
 * tile-pair energies summed into prow; mode 2: pass 3, added to gx, gy, gz.
 * Tile pairs without listed pairs (not visited) are skipped.
 */
__global__ __launch_bounds__(256) void k_det_sum(DevPtrs d, int mode)
{
    __shared__ double red[4][DET_NC][TILE];
    __shared__ double ered[3][256];
    const int T = blockIdx.x, tid = threadIdx.x, l = tid & 63, w = tid >> 6;
    const int nc = mode == 0 ? 1 : mode == 1 ? 4 : 3;
    const int r0 = d.rowstart[T], nr = d.rowcnt[T];
    const int I0 = d.mol_tile0[d.tile_mol[T]];
    const int *col = d.dcol + d.rowbase[I0] + (size_t)(T - I0) * (T - I0 + 1) / 2;
    const int ncol = d.dcolcnt[T];
    const bool diag = nr > 0 && d.tp[r0].y == T;
    const int n = nr + ncol + (diag ? 1 : 0);
    const int kb = (int)(((long long)n * w) >> 2), ke = (int)(((long long)n * (w + 1)) >> 2);
    double v[DET_NC] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int k = kb; k < ke; k++) {
        int t, side;
        if (k < nr) {
            t = r0 + k;
            side = 0;
        } else if (k < nr + ncol) {
            t = col[k - nr];
            side = 1;
        } else {
            t = r0;
            side = 1;
        }
        const bool vis = d.tcol[t] != 0ull;
        const double *pr = d.part + ((size_t)t * 2 + side) * (DET_NC * TILE) + l;
#pragma unroll
        for (int c = 0; c < DET_NC; c++)
            if (c < nc) {
                const double x = pr[c * TILE];
                v[c] += vis ? x : 0.0;
            }
    }
#pragma unroll
    for (int c = 0; c < DET_NC; c++) red[w][c][l] = v[c];
    if (mode == 1) {
        /* the row's tile-pair energies: fixed strides, fixed tree */
        double e0 = 0.0, e1 = 0.0, e2 = 0.0;
        for (int t = r0 + tid; t < r0 + nr; t += 256)
            if (d.tcol[t]) {
                e0 += d.pe[3 * (size_t)t];
                e1 += d.pe[3 * (size_t)t + 1];
                e2 += d.pe[3 * (size_t)t + 2];
            }
        ered[0][tid] = e0; ered[1][tid] = e1; ered[2][tid] = e2;
    }
    __syncthreads();
    if (mode == 1) {
        for (int h = 128; h > 0; h >>= 1) {
            if (tid < h)
                for (int c = 0; c < 3; c++) ered[c][tid] += ered[c][tid + h];
            __syncthreads();
        }
        if (tid < 3) d.prow[3 * (size_t)T + tid] = ered[tid][0];
    }
    if (w) return;
    const int s = T * TILE + l;
    double u[DET_NC];
#pragma unroll
    for (int c = 0; c < DET_NC; c++) u[c] = ((red[0][c][l] + red[1][c][l]) + red[2][c][l]) + red[3][c][l];
    if (mode == 0) {
        d.sumi[s] = u[0];
    } else if (mode == 1) {
        d.sumdeijda[s] += u[0];
        d.gx[s] += u[1];
        d.gy[s] += u[2];
        d.gz[s] += u[3];
    } else {
        d.gx[s] += u[0];
        d.gy[s] += u[1];
        d.gz[s] += u[2];
    }
}

/*
 * Deterministic mode, after pass 2: per molecule (one workgroup each), the
 * row energies and self energies of its tiles, each thread summing a fixed
 * stride and a fixed tree joining the threads; the totals go to the first
 * energy slot of each entry (the others hold zero), where k_post's slot
 * sums pick them up unchanged.
 */
__global__ __launch_bounds__(256) void k_det_energy(DevPtrs d)
{
    __shared__ double sh[4][256];
    const int m = blockIdx.x, tid = threadIdx.x;
    const int T0 = d.mol_tile0[m], T1 = d.mol_tile0[m + 1];
    double e0 = 0.0, e1 = 0.0, e2 = 0.0, es = 0.0;
    for (int T = T0 + tid; T < T1; T += 256) {
        e0 += d.prow[3 * (size_t)T];
        e1 += d.prow[3 * (size_t)T + 1];
        e2 += d.prow[3 * (size_t)T + 2];
        es += d.pself[T];
    }
    sh[0][tid] = e0; sh[1][tid] = e1; sh[2][tid] = e2; sh[3][tid] = es;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h)
            for (int c = 0; c < 4; c++) sh[c][tid] += sh[c][tid + h];
        __syncthreads();
    }
    if (tid == 0) {
        double *e = d.eslot + 3 * (size_t)m * ENE_SLOTS;
        e[0] = sh[3][0] + sh[0][0];
        e[ENE_SLOTS] = sh[1][0];
        e[2 * ENE_SLOTS] = sh[2][0];
    }
}

/*
 * After pass 3: the energy entries' slot sums (ene_slot_sums), and with a
 * gradient array, thread s < nslots adds slot s's GB gradient into the
 * gradient of its atom (original order, AoS).
 */
__global__ __launch_bounds__(256) void k_post(DevPtrs d, GBPar p, double *__restrict__ g, int nene)
{
    ene_slot_sums(d, nene);
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (!g || s >= p.nslots) return;
    const int a = d.perm[s];
    if (a < 0) return;
    g[3 * (size_t)a] += d.gx[s];
    g[3 * (size_t)a + 1] += d.gy[s];
    g[3 * (size_t)a + 2] += d.gz[s];
}

/* expand the membership bits into (i, j) original-index pairs (read-back) */
__global__ __launch_bounds__(64) void k_listpairs(DevPtrs d, int *pi, int *pj, unsigned long long *cnt,
                                                  long long cap)
{
    __shared__ int s_pos[TILE];
    const int l = threadIdx.x;
    const int ntp = *d.ntp;
    for (int t = blockIdx.x; t < ntp; t += gridDim.x) {
        const int2 ij = d.tpo[t];
        const unsigned long long cols = d.tcol[t];
        const int m = __popcll(cols);
        if (m == 0) continue;
        __syncthreads();
        if ((cols >> l) & 1ull) s_pos[__popcll(cols & ((1ull << l) - 1ull))] = l;
        __syncthreads();
        /* the word is in step order: bit s is column (l mod m + s) mod m */
        unsigned long long w = d.mask[(size_t)t * TILE + l];
        const int a = d.perm[ij.x * TILE + l];
        const int k0 = l % m;
        while (w) {
            const int st = __ffsll((unsigned long long)w) - 1;
            w &= w - 1ull;
            const int b = s_pos[k0 + st < m ? k0 + st : k0 + st - m];
            const unsigned long long k = atomicAdd(cnt, 1ull);
            if ((long long)k < cap) {
                pi[k] = a;
                pj[k] = d.perm[ij.y * TILE + b];
            }
        }
    }
}

/* per-atom values in original order (for reff / sumdeijda read-back) */
__global__ __launch_bounds__(256) void k_unsort(DevPtrs d, GBPar p, const double *src, double *dst)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= p.nslots) return;
    const int a = d.perm[s];
    if (a >= 0) dst[a] = src[s];
}

/* ------------------------------------------------------------------ engine */

/*
 * Atom order within a molecule: a k-d split of the atom set, recursively
 * along the longest extent of the current box, at a multiple of the tile
 * size (64), so that every tile is one leaf box.  Leaf boxes are compact, so
 * the 64x64 pair matrices of two tiles have fewer empty rows and columns
 * than with tiles cut from a space-filling curve.  Ties are broken by atom
 * index, so the order is deterministic.
 */
static void kd_split(const double *x, int *idx, int n)
{
    if (n <= TILE) return;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int k = 0; k < n; k++)
        for (int c = 0; c < 3; c++) {
            lo[c] = std::min(lo[c], x[3 * idx[k] + c]);
            hi[c] = std::max(hi[c], x[3 * idx[k] + c]);
        }
    int ax = 0;
    for (int c = 1; c < 3; c++)
        if (hi[c] - lo[c] > hi[ax] - lo[ax]) ax = c;
    const int ntile = (n + TILE - 1) / TILE;
    const int nleft = (ntile / 2) * TILE;
    std::nth_element(idx, idx + nleft, idx + n, [x, ax](int a, int b) {
        const double xa = x[3 * a + ax], xb = x[3 * b + ax];
        return xa < xb || (xa == xb && a < b);
    });
    kd_split(x, idx, nleft);
    kd_split(x, idx + nleft, n - nleft);
}

static void tile_order(int n, const double *x, std::vector<int> &order)
{
    order.resize(n);
    for (int i = 0; i < n; i++) order[i] = i;
    if (!x || n < 2) return;
    kd_split(x, order.data(), n);
}

static int check_options(const gbhip_options *o)
{
    if (!o) return seterr(GBHIP_EINVAL, "options pointer is NULL");
    if (o->gb != 1 && o->gb != 2 && o->gb != 5 && o->gb != 7 && o->gb != 8)
        return seterr(GBHIP_EINVAL, "gb=%d is not supported (1, 2, 5, 7, 8)", o->gb);
    if (o->precision != GBHIP_FP64 && o->precision != GBHIP_MIXED)
        return seterr(GBHIP_EINVAL, "unknown precision %d", o->precision);
    if (o->rgbmax > o->cut)
        return seterr(GBHIP_EINVAL, "Error in egb: rgbmax = %f is greater than cutoff = %f", o->rgbmax, o->cut);
    if (!(o->cut > 0) || !(o->rgbmax > 0) || !(o->epsext > 0))
        return seterr(GBHIP_EINVAL, "cut, rgbmax and epsext must be positive");
    return 0;
}

static int check_topology(const gbhip_topology *t, int gb)
{
    if (!t || t->natom < 1) return seterr(GBHIP_EINVAL, "empty topology");
    if (!t->charge || !t->rborn || !t->fs || !t->iac || !t->cno || !t->cn1 || !t->cn2 || !t->iblo || !t->exclat)
        return seterr(GBHIP_EINVAL, "topology is missing a required array");
    if (gb != 1 && (!t->gbalpha || !t->gbbeta || !t->gbgamma))
        return seterr(GBHIP_EINVAL, "gb=%d needs gbalpha/gbbeta/gbgamma", gb);
    if ((gb == 7 || gb == 8) && !t->neckidx) return seterr(GBHIP_EINVAL, "gb=%d needs neckidx", gb);
    if (t->ntypes < 1) return seterr(GBHIP_EINVAL, "ntypes must be >= 1");
    for (int i = 0; i < t->natom; i++) {
        if (t->iac[i] < 1 || t->iac[i] > t->ntypes)
            return seterr(GBHIP_EINVAL, "atom %d has LJ type %d outside 1..%d", i + 1, t->iac[i], t->ntypes);
        if ((gb == 7 || gb == 8) && (t->neckidx[i] < 0 || t->neckidx[i] > 20))
            return seterr(GBHIP_EINVAL, "Atom %d outside the allowed range of 1-2 Angstroms for igb==%d", i, gb);
    }
    return 0;
}

/*
 * Step split for a system with few tile pairs: one work item is one
 * wavefront's sweep, so a system with fewer tile pairs than the chip has
 * SIMDs leaves most of them idle, and a pass cannot end before its longest
 * sweep.  Each tile pair's steps are dealt out in chunks, doubling the
 * split while a pass has fewer than split_items_per_cu work items per CU
 * (counted from the tile-pair capacity): 64 for pass 1, whose sweeps are
 * the longest, 6 for the others; pass 1 stops at 4 chunks unless the system
 * has fewer than 128 tile pairs, the others at SPLIT_DEFAULT_MAX.  Fitted to
 * what the autotuner picks (mixed, HIP events): 600 atoms 8 in every pass;
 * 2,500 atoms without cutoff (820 tile pairs) and 3,000 atoms at 20 A
 * (1,108) pass 1 at 4, passes 2 and 3 at 2 (3,000 atoms: 86 -> 74 us per
 * evaluation against the earlier default); 5,000-7,000 atoms at 20 A pass
 * 1 at 4, passes 2 and 3 unsplit (more chunks add per-visit set-up and
 * flush atomics).
 */
static const int split_items_per_cu[3] = {64, 6, 6};

static int default_split(const Engine &E, int pass)
{
    int sh = 0;
    if (E.par.det) return 0;
    const long long cap = std::max<long long>(1, E.tp_cap);
    const int most = pass == 0 && cap >= 128 ? 4 : SPLIT_DEFAULT_MAX;
    while ((1 << sh) < most && (cap << sh) < (long long)split_items_per_cu[pass] * E.ncu) sh++;
    return sh;
}

/* evaluation grids until autotuned (measured at 25k atoms, 25 A, where the
 * tile-pair capacity is 2.6 times the 28,713 non-empty tile pairs): passes
 * 1 and 3 fastest with one workgroup per work item up to 128 per CU, pass
 * 2 near 64 per CU */
static void default_grids(Engine &E)
{
    long long items[3];
    for (int k = 0; k < 3; k++) items[k] = std::max<long long>(1, E.tp_cap) << E.par.split_sh[k];
    const long long wide = std::min<long long>((long long)E.ncu * 128, MAX_LAUNCH);
    E.pass_blocks[0] = (int)std::min<long long>(items[0], wide);
    E.pass_blocks[1] = (int)std::min<long long>(items[1], (long long)E.ncu * 64);
    E.pass_blocks[2] = (int)std::min<long long>(items[2], wide);
}

int engine_build(Engine &E, int nmol, const gbhip_topology *tops, const gbhip_options *o,
                        const double *const *x0)
{
    int rc = check_options(o);
    if (rc) return rc;
    for (int m = 0; m < nmol; m++)
        if ((rc = check_topology(&tops[m], o->gb))) return rc;
    if (gbhip_device_count() < 1) return seterr(GBHIP_ENODEV, "no GPU is visible to the HIP runtime");

    E.nmol = nmol;
    E.precision = o->precision;
    E.mol_atom_off.resize(nmol + 1);
    E.mol_natom.resize(nmol);
    E.mol_slot_off.resize(nmol + 1);
    long long atoms = 0, slots = 0;
    for (int m = 0; m < nmol; m++) {
        E.mol_atom_off[m] = (int)atoms;
        E.mol_slot_off[m] = (int)slots;
        E.mol_natom[m] = tops[m].natom;
        atoms += tops[m].natom;
        slots += (long long)(tops[m].natom + TILE - 1) / TILE * TILE;
    }
    E.mol_atom_off[nmol] = (int)atoms;
    E.mol_slot_off[nmol] = (int)slots;
    if (slots > 0x7fffffffLL / 4) return seterr(GBHIP_EINVAL, "system too large");
    E.natoms = (int)atoms;
    E.nslots = (int)slots;
    E.ntiles = E.nslots / TILE;

    const int NS = E.nslots, NT = E.ntiles;
    std::vector<int> perm(NS, -1), inv(E.natoms), slot_mol(NS), frz(NS, 1), neck(NS, 0), ljrow(NS, 0), ljcol(NS, 0);
    std::vector<double> q(NS, 0.0), rb(NS, 1.0), ri0(NS, 1.0), sj(NS, 0.0), ga(NS, 0.0), gbb(NS, 0.0), gg(NS, 0.0);
    std::vector<int> tile_mol(NT), tile_end(NT);
    std::vector<long long> rowbase(NT);
    std::vector<double2> ljab;
    std::vector<int2> mol_lj(nmol);
    std::vector<double> rgbpsmax2(nmol);
    std::vector<int> ex_ptr(E.natoms + 1, 0), ex_j;
    long long cap = 0;

    for (int m = 0; m < nmol; m++) {
        const gbhip_topology &t = tops[m];
        const int n = t.natom, so = E.mol_slot_off[m], ao = E.mol_atom_off[m];
        std::vector<int> order;
        tile_order(n, x0 ? x0[m] : nullptr, order);
        const int lj0 = (int)ljab.size(), nt = t.ntypes;
        mol_lj[m] = make_int2(lj0, nt);
        rgbpsmax2[m] = (o->rgbmax + t.fsmax) * (o->rgbmax + t.fsmax);
        for (int a = 0; a < nt; a++)
            for (int b = 0; b < nt; b++) {
                int ic = t.cno[a * nt + b] - 1;
                ljab.push_back(make_double2(ic >= 0 ? t.cn1[ic] : 0.0, ic >= 0 ? t.cn2[ic] : 0.0));
            }
        for (int k = 0; k < n; k++) {
            const int a = order[k], s = so + k;
            perm[s] = ao + a;
            inv[ao + a] = s;
            q[s] = t.charge[a];
            rb[s] = t.rborn[a];
            ri0[s] = t.rborn[a] - o->gboffset;
            sj[s] = t.fs[a] * (t.rborn[a] - o->gboffset);
            if (o->gb != 1) {
                ga[s] = t.gbalpha[a];
                gbb[s] = t.gbbeta[a];
                gg[s] = t.gbgamma[a];
            }
            neck[s] = (o->gb == 7 || o->gb == 8) ? t.neckidx[a] : 0;
            frz[s] = t.frozen ? (t.frozen[a] != 0) : 0;
            ljrow[s] = lj0 + (t.iac[a] - 1) * nt;
            ljcol[s] = t.iac[a] - 1;
        }
        const int nst = (E.mol_slot_off[m + 1] - so);
        for (int s = so; s < so + nst; s++) slot_mol[s] = m;
        const int T0 = so / TILE, T1 = (so + nst) / TILE;
        for (int T = T0; T < T1; T++) {
            tile_mol[T] = m;
            tile_end[T] = T1;
            rowbase[T] = cap;
            cap += T1 - T;
        }
        /* exclusions with partner index > own index (eff.c:1920-1923) */
        int e = 0;
        for (int i = 0; i < n; i++) {
            ex_ptr[ao + i] = (int)ex_j.size();
            for (int k = 0; k < t.iblo[i]; k++) {
                int jj = t.exclat[e + k] - 1;
                if (jj > i && jj < n) ex_j.push_back(ao + jj);
            }
            e += t.iblo[i];
        }
    }
    ex_ptr[E.natoms] = (int)ex_j.size();
    E.tp_cap = cap;
    E.perm_h = perm;
    E.frz_h = frz;

    GBPar &P = E.par;
    P.gb = o->gb;
    P.has_neck = (o->gb == 7 || o->gb == 8);
    P.nslots = NS;
    P.nmol = nmol;
    P.cut2 = o->cut * o->cut;
    P.rgbmax = o->rgbmax;
    P.rgbmax1i = 1.0 / o->rgbmax;
    P.rgbmax2i = P.rgbmax1i * P.rgbmax1i;
    P.neckscale = o->gbneckscale;
    P.kappa = o->kappa;
    P.epsext = o->epsext;
    P.qfill[0] = BORN_QFILL;
    P.qfill[1] = 0.0f;
    P.qfill[2] = 0.0f;
    P.pad_ = 0;

    DevPtrs &D = E.d;
    D.ntiles = NT;
    D.natoms = E.natoms;
    if ((rc = E.upload(&D.perm, perm)) || (rc = E.upload(&D.slot_mol, slot_mol)) ||
        (rc = E.upload(&D.q, q)) || (rc = E.upload(&D.rb, rb)) || (rc = E.upload(&D.ri0, ri0)) ||
        (rc = E.upload(&D.sj, sj)) || (rc = E.upload(&D.ga, ga)) || (rc = E.upload(&D.gbb, gbb)) ||
        (rc = E.upload(&D.gg, gg)) || (rc = E.upload(&D.neck, neck)) || (rc = E.upload(&D.frz, frz)) ||
        (rc = E.upload(&D.ljrow, ljrow)) || (rc = E.upload(&D.ljcol, ljcol)) ||
        (rc = E.upload(&D.tile_mol, tile_mol)) || (rc = E.upload(&D.tile_end, tile_end)) ||
        (rc = E.upload(&D.rowbase, rowbase)) || (rc = E.upload(&D.ex_ptr, ex_ptr)) ||
        (rc = E.upload(&D.ex_j, ex_j)) || (rc = E.upload(&D.inv, inv)) ||
        (rc = E.upload(&D.ljab, ljab)) || (rc = E.upload(&D.mol_lj, mol_lj)) ||
        (rc = E.upload(&D.rgbpsmax2, rgbpsmax2)))
        return rc;
    {
        std::vector<float> rf(ri0.begin(), ri0.end()), sf(sj.begin(), sj.end()),
            bf(rb.begin(), rb.end());
        std::vector<double> np(NECK_POS, NECK_POS + NECK_N * NECK_N), nv(NECK_VAL, NECK_VAL + NECK_N * NECK_N);
        std::vector<float> npf(np.begin(), np.end()), nvf(nv.begin(), nv.end());
        if ((rc = E.upload(&D.ri0_f, rf)) || (rc = E.upload(&D.sj_f, sf)) ||
            (rc = E.upload(&D.rb_f, bf)) ||
            (rc = E.upload(&D.neck_pos, np)) || (rc = E.upload(&D.neck_val, nv)) ||
            (rc = E.upload(&D.neck_pos_f, npf)) || (rc = E.upload(&D.neck_val_f, nvf)))
            return rc;
    }
    if ((rc = E.dalloc(&D.xs, NS)) || (rc = E.dalloc(&D.ys, NS)) || (rc = E.dalloc(&D.zs, NS)) ||
        (rc = E.dalloc(&D.sumi, NS)) || (rc = E.dalloc(&D.reff, NS)) || (rc = E.dalloc(&D.psi, NS)) ||
        (rc = E.dalloc(&D.dfac, NS)) || (rc = E.dalloc(&D.sumdeijda, NS)) || (rc = E.dalloc(&D.gx, NS)) ||
        (rc = E.dalloc(&D.gy, NS)) || (rc = E.dalloc(&D.gz, NS)) || (rc = E.dalloc(&D.ene, 3 * (size_t)nmol)) ||
        (rc = E.dalloc(&D.eslot, 3 * (size_t)nmol * ENE_SLOTS)) ||
        (rc = E.dalloc(&D.cand, (size_t)cap)) || (rc = E.dalloc(&D.rowcnt, NT)) ||
        (rc = E.dalloc(&D.rowstart, NT)) || (rc = E.dalloc(&D.ntp, 1)) || (rc = E.dalloc(&D.tp, (size_t)cap)) ||
        (rc = E.dalloc(&D.tpo, (size_t)cap)) || (rc = E.dalloc(&D.tflip, (size_t)cap)) ||
        (rc = E.dalloc(&D.tcol, (size_t)cap)) || (rc = E.dalloc(&D.tord, (size_t)cap)) ||
        (rc = E.dalloc(&D.tbin, 2 * TILE)) || (rc = E.dalloc(&D.nnz, 1)) ||
        (rc = E.dalloc(&D.tnp, (size_t)cap)) ||
        (rc = E.dalloc(&D.qm, (size_t)cap)) ||
        (rc = E.dalloc(&D.store_need, 1)) || (rc = E.dalloc(&D.store_ok, 1)) ||
        (rc = E.dalloc(&D.mask, (size_t)cap * TILE)) || (rc = E.dalloc(&D.excl, (size_t)cap * TILE)) ||
        (rc = E.dalloc(&D.sbal, (size_t)cap * TILE)) || (rc = E.dalloc(&D.npairs, 1)) || (rc = E.dalloc(&D.bblo, 3 * (size_t)NT)) ||
        (rc = E.dalloc(&D.bbhi, 3 * (size_t)NT)) || (rc = E.dalloc(&E.x_dev, 3 * (size_t)E.natoms)) ||
        (rc = E.dalloc(&E.g_dev, 3 * (size_t)E.natoms)) || (rc = E.dalloc(&E.tmp_dev, (size_t)E.natoms)) ||
        (rc = E.dalloc(&E.ene14_dev, 14)))
        return rc;
    HIPCHK(hipMemsetAsync(D.ntp, 0, sizeof(int), E.stream));
    HIPCHK(hipMemsetAsync(D.eslot, 0, 3 * (size_t)nmol * ENE_SLOTS * sizeof(double), E.stream));
    HIPCHK(hipMemsetAsync(D.ene, 0, 3 * (size_t)nmol * sizeof(double), E.stream));
    HIPCHK(hipMemsetAsync(D.nnz, 0, sizeof(int), E.stream));
    HIPCHK(hipMemsetAsync(D.store_ok, 0, sizeof(int), E.stream));
    HIPCHK(hipMemsetAsync(D.store_need, 0, sizeof(long long), E.stream));
    D.dstore = nullptr;
    D.store_cap = 0;
    HIPCHK(hipMemsetAsync(E.x_dev, 0, 3 * sizeof(double) * (size_t)E.natoms, E.stream));
    if (x0) {
        for (int m = 0; m < nmol; m++)
            if (x0[m] && (rc = E.h2d(E.x_dev + 3 * (size_t)E.mol_atom_off[m], x0[m],
                                     3 * sizeof(double) * (size_t)E.mol_natom[m])))
                return rc;
    }
    if ((rc = E.upload_done())) return rc;
    int dev = 0, ncu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    E.pair_blocks = (int)std::min<long long>(std::max<long long>(1, cap), (long long)ncu * 16);
    E.pair_blocks = std::min(E.pair_blocks, PAIR_BLOCKS_MAX * 4);
    E.ncu = ncu;
    for (int k = 0; k < 3; k++) E.par.split_sh[k] = default_split(E, k);
    default_grids(E);
    return 0;
}

int engine_gather(Engine &E, const double *x_dev)
{
    const int nb = (E.nslots + 255) / 256;
    hipLaunchKernelGGL(k_gather, dim3(nb), dim3(256), 0, E.stream, E.d, E.par, x_dev);
    LAUNCHCHK();
    return 0;
}

/* the store is large enough for every tile pair full (k_storeok's test,
 * need < cap, then holds for any list) */
static bool store_covers_all(const Engine &E)
{
    return E.d.dstore && E.d.store_cap > E.tp_cap * TILE * TILE;
}

int engine_nblist(Engine &E)
{
    /* HCP builds no pair list (eff.c:2560-2561) */
    if (E.hcp.on) {
        E.have_list = true;
        return 0;
    }
    const int NT = E.ntiles;
    int rc = engine_gather(E, E.x_dev);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bbox, dim3(NT), dim3(64), 0, E.stream, E.d);
    hipLaunchKernelGGL(k_tilecand, dim3(NT), dim3(256), 0, E.stream, E.d, E.par);
    hipLaunchKernelGGL(k_rowscan, dim3(1), dim3(1024), 0, E.stream, E.d);
    hipLaunchKernelGGL(k_compact, dim3(NT), dim3(256), 0, E.stream, E.d);
    hipLaunchKernelGGL(k_masks, dim3(E.pair_blocks), dim3(64), 0, E.stream, E.d, E.par);
    const int nsb = (int)std::min<long long>(64, (E.tp_cap + 255) / 256);
    hipLaunchKernelGGL(k_bincount, dim3(nsb), dim3(256), 0, E.stream, E.d);
    hipLaunchKernelGGL(k_binscan, dim3(1), dim3(64), 0, E.stream, E.d);
    hipLaunchKernelGGL(k_binscatter, dim3(nsb), dim3(256), 0, E.stream, E.d);
    hipLaunchKernelGGL(k_storeok, dim3(1), dim3(64), 0, E.stream, E.d);
    if (E.natoms > 0)
        hipLaunchKernelGGL(k_excl, dim3((E.natoms + 255) / 256), dim3(256), 0, E.stream, E.d);
    hipLaunchKernelGGL(k_steporder, dim3(E.pair_blocks), dim3(64), 0, E.stream, E.d);
    if (E.par.det) hipLaunchKernelGGL(k_det_cols, dim3(NT), dim3(64), 0, E.stream, E.d);
    LAUNCHCHK();
    E.have_list = true;
    /* a store sized for every tile pair full holds any list; otherwise the
     * host does not know until a synchronous build checks (k_storeok has
     * set the device flag either way) */
    E.store_holds_list = store_covers_all(E);
    return 0;
}

/* the derivative store holds the current list, as the host knows it (and,
 * while a graph that rebuilds the list is captured, every list) */
bool engine_store_holds_list(const Engine &E)
{
    if (E.capture_spans_rebuilds) return E.store_mode == 1 && store_covers_all(E);
    return E.store_mode == 1 && E.d.dstore && E.store_holds_list;
}

/*
 * Size the pass-3 derivative store for the list just built (synchronises the
 * engine stream).  Grows it to 1.25x the need when it is too small, unless a
 * captured graph holds the current pointer or the device lacks the memory;
 * in those cases the list is evaluated with pass 3 recomputing the
 * derivatives (the device flag written at list build says which).  With
 * upper (before dynamics graphs are captured) it is sized for the largest
 * list the molecules can have, every tile pair full, when that takes at
 * most a quarter of the free device memory (25k atoms: 2.5 GB mixed, 5 GB
 * FP64), else as above.
 */
static int store_reserve(Engine &E, bool upper, long long need)
{
    if (E.store_mode == 0 || E.store_frozen) return 0;
    const long long full = E.tp_cap * TILE * TILE + (long long)TILE * TILE;
    if (E.d.dstore && need < E.d.store_cap && (!upper || E.d.store_cap >= full)) return 0;
    const size_t esz = E.precision == GBHIP_FP64 ? 2 * sizeof(double) : 2 * sizeof(float);
    long long cap = need + need / 4 + (long long)TILE * TILE;
    if (E.slot) {
        /* in a slot: taken once from the slot's region, for every tile pair
         * full when that fits (then no list outgrows it), else for what the
         * region has left; never freed or regrown (a list it cannot hold is
         * evaluated with pass 3 recomputing, as the device flag says) */
        if (E.d.dstore) return 0;
        const size_t left = E.slot->dsize - std::min(E.slot->dsize, (E.slot->dused + 255) & ~(size_t)255);
        const long long want = (size_t)full * esz <= left ? full : std::min(cap, (long long)(left / esz));
        void *p = nullptr;
        if (want <= need || E.dalloc_bytes(&p, (size_t)want * esz) != 0) return 0;
        E.d.dstore = p;
        E.d.store_cap = want;
        hipLaunchKernelGGL(k_setflag, dim3(1), dim3(64), 0, E.stream, E.d.store_ok, 1);
        LAUNCHCHK();
        HIPCHK(hipStreamSynchronize(E.stream));
        return 0;
    }
    std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
    if (E.store_mem) {
        HIPCHK(hipStreamSynchronize(E.stream));
        (void)hipFree(E.store_mem);
        E.store_mem = nullptr;
        E.d.dstore = nullptr;
        E.d.store_cap = 0;
        /* the flag the list build set may name the store just freed: clear
         * it until a new store is in place */
        hipLaunchKernelGGL(k_setflag, dim3(1), dim3(64), 0, E.stream, E.d.store_ok, 0);
        LAUNCHCHK();
        HIPCHK(hipStreamSynchronize(E.stream));
    }
    size_t freeb = 0, totb = 0;
    if (hipMemGetInfo(&freeb, &totb) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    if (upper && (size_t)full * esz <= freeb / 4) cap = full;
    if ((size_t)cap * esz > freeb / 2) {
        (void)hipGetLastError();
        return 0;
    }
    void *p = nullptr;
    if (hipMalloc(&p, (size_t)cap * esz) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    E.store_mem = p;
    E.d.dstore = p;
    E.d.store_cap = cap;
    hipLaunchKernelGGL(k_setflag, dim3(1), dim3(64), 0, E.stream, E.d.store_ok, 1);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(E.stream));
    return 0;
}

int engine_store_reserve(Engine &E, bool upper)
{
    if (E.hcp.on) return 0; /* no pair list, no derivative store */
    long long need = 0;
    int rc = engine_d2h(E, &need, E.d.store_need, sizeof(need));
    if (rc) return rc;
    rc = store_reserve(E, upper, need);
    /* the same test k_storeok (or k_setflag after a regrowth) applied on the
     * device */
    E.store_holds_list = E.d.dstore && need < E.d.store_cap;
    return rc;
}

static int set_store_mode(Engine &E, int mode)
{
    if (mode != 0 && mode != 1) return seterr(GBHIP_EINVAL, "derivative store mode must be 0 or 1");
    if (E.store_frozen && mode != E.store_mode)
        return seterr(GBHIP_ESTATE, "derivative store mode cannot change while dynamics graphs exist");
    E.store_mode = mode;
    if (mode == 0) {
        std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
        HIPCHK(hipStreamSynchronize(E.stream));
        if (E.store_mem) (void)hipFree(E.store_mem);
        E.store_mem = nullptr;
        E.d.dstore = nullptr;
        E.d.store_cap = 0;
        hipLaunchKernelGGL(k_setflag, dim3(1), dim3(64), 0, E.stream, E.d.store_ok, 0);
        LAUNCHCHK();
        HIPCHK(hipStreamSynchronize(E.stream));
    } else if (E.have_list) {
        return engine_store_reserve(E);
    }
    return 0;
}

/*
 * One evaluation; coordinates in x_dev (AoS).  flags: EV_GATHERED when the
 * caller's own kernel has done k_gather's work, EV_NO_POST when it does
 * k_post's (the fused kernels of the mme and dynamics paths).
 */
/* the energy-slot sums and the slot gradients added into g_dev (k_post) */
int engine_post(Engine &E, double *g_dev, int nene, int nthreads)
{
    hipLaunchKernelGGL(k_post, dim3((nthreads + 255) / 256), dim3(256), 0, E.stream, E.d, E.par, g_dev, nene);
    LAUNCHCHK();
    return 0;
}

int engine_eval(Engine &E, const double *x_dev, double *g_dev, int flags)
{
    if (E.hcp.on) return engine_eval_hcp(E, x_dev, g_dev, flags);
    if (!E.have_list) return seterr(GBHIP_ESTATE, "egb called before the pair list was built");
    if (!(flags & EV_GATHERED)) {
        int rc = engine_gather(E, x_dev);
        if (rc) return rc;
    }
    const int nb64 = (E.nslots + 63) / 64;
    hipEvent_t *ev = nullptr;
    if (E.tev_used + 5 <= E.tev_cap) {
        ev = &E.tev[E.tev_used];
        E.tev_used += 5;
    }
    if (ev) HIPCHK(hipEventRecord(ev[0], E.stream));
    const bool det = E.par.det != 0;
    const bool sp1 = !det && E.par.split_sh[0] > 0, sp2 = !det && E.par.split_sh[1] > 0,
               sp3 = !det && E.par.split_sh[2] > 0;
    if (E.precision == GBHIP_FP64 && E.born_queue)
        hipLaunchKernelGGL((sp1 ? k_born_q<double, true> : k_born_q<double, false>), dim3(E.pass_blocks[0]), dim3(64),
                           0, E.stream, E.d, E.par);
    else if (E.precision == GBHIP_FP64)
        hipLaunchKernelGGL((sp1 ? k_born<double, true> : k_born<double, false>), dim3(E.pass_blocks[0]), dim3(64), 0,
                           E.stream, E.d, E.par);
    else
        hipLaunchKernelGGL((sp1 ? k_born<float, true> : k_born<float, false>), dim3(E.pass_blocks[0]), dim3(64), 0,
                           E.stream, E.d, E.par);
    if (det) hipLaunchKernelGGL(k_det_sum, dim3(E.ntiles), dim3(256), 0, E.stream, E.d, 0);
    if (ev) HIPCHK(hipEventRecord(ev[1], E.stream));
    hipLaunchKernelGGL(k_radii, dim3(nb64), dim3(64), 0, E.stream, E.d, E.par);
    if (ev) HIPCHK(hipEventRecord(ev[2], E.stream));
    if (E.precision == GBHIP_FP64)
        hipLaunchKernelGGL((sp2 ? k_pairs<double, true> : k_pairs<double, false>), dim3(E.pass_blocks[1]), dim3(64), 0,
                           E.stream, E.d, E.par);
    else
        hipLaunchKernelGGL((sp2 ? k_pairs<float, true> : k_pairs<float, false>), dim3(E.pass_blocks[1]), dim3(64), 0,
                           E.stream, E.d, E.par);
    if (det) {
        hipLaunchKernelGGL(k_det_sum, dim3(E.ntiles), dim3(256), 0, E.stream, E.d, 1);
        hipLaunchKernelGGL(k_det_energy, dim3(E.nmol), dim3(256), 0, E.stream, E.d);
    }
    if (ev) HIPCHK(hipEventRecord(ev[3], E.stream));
    /* k_chain_st and k_chain each read the device flag and one of them
     * works; when the host knows the store holds the list, k_chain (which
     * would exit at once) is not launched */
    const bool recompute = !engine_store_holds_list(E);
    if (E.precision == GBHIP_FP64) {
        hipLaunchKernelGGL((sp3 ? k_chain_st<double, true> : k_chain_st<double, false>), dim3(E.pass_blocks[2]),
                           dim3(64), 0, E.stream, E.d, E.par);
        if (recompute)
            hipLaunchKernelGGL((sp3 ? k_chain<double, true> : k_chain<double, false>), dim3(E.pass_blocks[2]),
                               dim3(64), 0, E.stream, E.d, E.par);
    } else {
        hipLaunchKernelGGL((sp3 ? k_chain_st<float, true> : k_chain_st<float, false>), dim3(E.pass_blocks[2]),
                           dim3(64), 0, E.stream, E.d, E.par);
        if (recompute)
            hipLaunchKernelGGL((sp3 ? k_chain<float, true> : k_chain<float, false>), dim3(E.pass_blocks[2]),
                               dim3(64), 0, E.stream, E.d, E.par);
    }
    if (det) hipLaunchKernelGGL(k_det_sum, dim3(E.ntiles), dim3(256), 0, E.stream, E.d, 2);
    if (ev) HIPCHK(hipEventRecord(ev[4], E.stream));
    if (!(flags & EV_NO_POST)) {
        /* a wavefront per energy entry, a thread per slot */
        const int nene = 3 * E.nmol, nt = std::max(64 * nene, g_dev ? E.nslots : 0);
        hipLaunchKernelGGL(k_post, dim3((nt + 255) / 256), dim3(256), 0, E.stream, E.d, E.par, g_dev, nene);
    }
    LAUNCHCHK();
    E.have_eval = true;
    return 0;
}

/* ----------------------------------------------------------------- timing */

__global__ __launch_bounds__(256) void k_zero(double *p, long long n)
{
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k < n) p[k] = 0.0;
}

static int timing_start(Engine &E, int max_evals)
{
    if (max_evals < 0 || max_evals > 1000000) return seterr(GBHIP_EINVAL, "max_evals out of range");
    HIPCHK(hipStreamSynchronize(E.stream));
    for (hipEvent_t e : E.tev) (void)hipEventDestroy(e);
    E.tev.assign(5 * (size_t)max_evals, nullptr);
    for (auto &e : E.tev) HIPCHK(hipEventCreate(&e));
    E.tev_cap = 5 * max_evals;
    E.tev_used = 0;
    return 0;
}

/* ms4: summed milliseconds of pass 1, radii, pass 2, pass 3 over nevals */
static int timing_stop(Engine &E, double *ms4, int *nevals)
{
    HIPCHK(hipStreamSynchronize(E.stream));
    double acc[4] = {0, 0, 0, 0};
    const int n = E.tev_used / 5;
    for (int k = 0; k < n; k++) {
        for (int s = 0; s < 4; s++) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, E.tev[5 * k + s], E.tev[5 * k + s + 1]));
            acc[s] += ms;
        }
    }
    for (hipEvent_t e : E.tev) (void)hipEventDestroy(e);
    E.tev.clear();
    E.tev_cap = E.tev_used = 0;
    if (ms4)
        for (int s = 0; s < 4; s++) ms4[s] = acc[s];
    if (nevals) *nevals = n;
    return 0;
}

/* ------------------------------------------------------------ launch shapes */

/* per pass 1, 2, 4, 8 or 16 step chunks per tile pair (0: the default for the
 * system's size); the grids return to their defaults for the new split */
static int set_split(Engine &E, const int *split3)
{
    int sh[3];
    for (int k = 0; k < 3; k++) {
        if (split3[k] == 0) {
            sh[k] = default_split(E, k);
        } else {
            sh[k] = 0;
            while ((1 << sh[k]) < split3[k] && (1 << sh[k]) < SPLIT_MAX) sh[k]++;
            if (split3[k] < 0 || (1 << sh[k]) != split3[k])
                return seterr(GBHIP_EINVAL, "step split must be 0 or a power of two up to %d", SPLIT_MAX);
        }
        if ((std::max<long long>(1, E.tp_cap) << sh[k]) > (1LL << 30))
            return seterr(GBHIP_EINVAL, "too many tile pairs for a step split of %d", 1 << sh[k]);
        if (E.par.det && sh[k] != 0)
            return seterr(GBHIP_EINVAL, "the deterministic mode evaluates each tile pair in one piece (step split 1)");
    }
    for (int k = 0; k < 3; k++) E.par.split_sh[k] = sh[k];
    default_grids(E);
    return 0;
}

static int set_launch(Engine &E, const int *blocks3)
{
    for (int k = 0; k < 3; k++) {
        if (blocks3[k] < 0 || blocks3[k] > MAX_LAUNCH) return seterr(GBHIP_EINVAL, "launch size out of range");
    }
    for (int k = 0; k < 3; k++)
        if (blocks3[k] > 0) E.pass_blocks[k] = blocks3[k];
    return 0;
}

/*
 * Time whole evaluations on the current list and coordinates: first with
 * and without the pass-3 derivative store (keeping the faster), then, for a
 * system with few tile pairs, each step split (1, 2, 4, 8; the fastest per
 * pass is kept), then for each candidate
 * grid size (workgroups per CU: 4 .. 128, and one workgroup per work item),
 * keeping per pass the fastest.  Up to the number
 * the CUs hold at once the grid is persistent and the candidates differ in
 * latency hiding; beyond it the hardware hands later workgroups to CUs as
 * earlier ones finish, which balances the uneven tile-pair costs better than
 * the static deal (measured at 25k atoms: pass 1 0.32 ms at 40 per CU, 0.26
 * ms at one tile pair per workgroup) but repeats per-workgroup set-up (pass 2
 * prefers a persistent grid).  Leaves the engine evaluated at the current
 * coordinates; replaces any timing session in progress.
 */
/* summed pass times (ms) of reps evaluations at the current settings, after
 * one untimed evaluation */
static int time_evals(Engine &E, int reps, double *ms3)
{
    int rc;
    if ((rc = engine_eval(E, E.x_dev, nullptr))) return rc;
    if ((rc = timing_start(E, reps))) return rc;
    for (int r = 0; r < reps; r++)
        if ((rc = engine_eval(E, E.x_dev, nullptr))) return rc;
    double ms4[4];
    int n = 0;
    if ((rc = timing_stop(E, ms4, &n))) return rc;
    ms3[0] = ms4[0];
    ms3[1] = ms4[2];
    ms3[2] = ms4[3];
    return 0;
}

static int autotune(Engine &E, int reps, int *chosen3)
{
    static const int per_cu[] = {4, 8, 16, 24, 32, 40, 64, 96, 128};
    const int nper = (int)(sizeof(per_cu) / sizeof(per_cu[0]));
    if (!E.have_list) return seterr(GBHIP_ESTATE, "autotune needs a pair list (call nblist first)");
    if (reps < 1 || reps > 1000) return seterr(GBHIP_EINVAL, "reps out of range");
    int rc;
    double ms[3];
    int nnz = 0;
    HIPCHK(hipMemcpyAsync(&nnz, E.d.nnz, sizeof(nnz), hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    /* 1. derivative store or recomputation in pass 3, at the current launch
     * shapes: the store saves pass 3 the descreening arithmetic but costs
     * 8 (mixed) or 16 (FP64) bytes per pair written and read, which does not
     * pay where that arithmetic is cheap (FP32, no rgbmax shell) */
    if (E.store_mode == 1 && !E.store_frozen && E.d.dstore) {
        int ok = 0;
        HIPCHK(hipMemcpyAsync(&ok, E.d.store_ok, sizeof(ok), hipMemcpyDeviceToHost, E.stream));
        HIPCHK(hipStreamSynchronize(E.stream));
        if (ok) {
            if ((rc = time_evals(E, reps, ms))) return rc;
            const double t_on = ms[0] + ms[1] + ms[2];
            if ((rc = set_store_mode(E, 0))) return rc;
            if ((rc = time_evals(E, reps, ms))) return rc;
            const double t_off = ms[0] + ms[1] + ms[2];
            if (t_on <= t_off && (rc = set_store_mode(E, 1))) return rc;
        }
    }
    /* 2. step split per pass, for systems with few tile pairs (each at one
     * workgroup per work item, the grid that suits short work lists) */
    if (!E.par.det && E.tp_cap < (long long)split_items_per_cu[0] * E.ncu) {
        double tbest[3] = {1e300, 1e300, 1e300};
        int spbest[3] = {1 << E.par.split_sh[0], 1 << E.par.split_sh[1], 1 << E.par.split_sh[2]};
        for (int sh = 0; (1 << sh) <= SPLIT_MAX; sh++) {
            const int sp3[3] = {1 << sh, 1 << sh, 1 << sh};
            if ((rc = set_split(E, sp3))) return rc;
            for (int k = 0; k < 3; k++) E.pass_blocks[k] = std::min(MAX_LAUNCH, std::max(1, nnz << sh));
            if ((rc = time_evals(E, reps, ms))) return rc;
            for (int k = 0; k < 3; k++)
                if (ms[k] < tbest[k]) {
                    tbest[k] = ms[k];
                    spbest[k] = 1 << sh;
                }
        }
        if ((rc = set_split(E, spbest))) return rc;
    }
    std::vector<int> cand;
    for (int c = 0; c < nper; c++) cand.push_back(std::min(MAX_LAUNCH, std::max(1, E.ncu * per_cu[c])));
    for (int k = 0; k < 3; k++) cand.push_back(std::min(MAX_LAUNCH, std::max(1, nnz << E.par.split_sh[k])));
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    /* 3. grid size per pass */
    double best[3] = {1e300, 1e300, 1e300};
    int pick[3] = {E.pass_blocks[0], E.pass_blocks[1], E.pass_blocks[2]};
    for (const int nb : cand) {
        for (int k = 0; k < 3; k++) E.pass_blocks[k] = nb;
        if ((rc = time_evals(E, reps, ms))) return rc;
        for (int k = 0; k < 3; k++)
            if (ms[k] < best[k]) {
                best[k] = ms[k];
                pick[k] = nb;
            }
    }
    for (int k = 0; k < 3; k++) E.pass_blocks[k] = pick[k];
    if (chosen3)
        for (int k = 0; k < 3; k++) chosen3[k] = pick[k];
    return engine_eval(E, E.x_dev, nullptr);
}

/*
 * Deterministic reductions on or off.  The arrays (about 4 KiB per tile
 * pair of the list capacity, 32 bytes per tile) are allocated
 * the first time the mode is set and kept; the step split returns to 1 per
 * pass while the mode is on, and to the default for the system when it is
 * turned off.
 */
static int set_det(Engine &E, int on)
{
    if (on != 0 && on != 1) return seterr(GBHIP_EINVAL, "deterministic mode must be 0 or 1");
    if (on == E.par.det) return 0;
    if (E.store_frozen) return seterr(GBHIP_ESTATE, "deterministic mode cannot change while dynamics graphs exist");
    if (on && !E.det_mem) {
        if (E.slot) return seterr(GBHIP_ESTATE, "deterministic mode is not available to a context in a slot");
        const size_t cap = (size_t)std::max<long long>(1, E.tp_cap);
        const size_t b_part = cap * 2 * DET_NC * TILE * sizeof(double), b_pe = cap * 3 * sizeof(double);
        const size_t b_ps = (size_t)E.ntiles * sizeof(double), b_mt = (size_t)(E.nmol + 1) * sizeof(int);
        const size_t b_pr = (size_t)E.ntiles * 3 * sizeof(double), b_dc = cap * sizeof(int),
                     b_dn = (size_t)E.ntiles * sizeof(int);
        auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
        const size_t total = al(b_part) + al(b_pe) + al(b_ps) + al(b_mt) + al(b_pr) + al(b_dc) + al(b_dn);
        std::vector<int> mt(E.nmol + 1);
        for (int m = 0; m <= E.nmol; m++) mt[m] = E.mol_slot_off[m] / TILE;
        std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
        HIPCHK(hipStreamSynchronize(E.stream));
        void *p = nullptr;
        if (hipMalloc(&p, total) != hipSuccess) {
            (void)hipGetLastError();
            return seterr(GBHIP_ENOMEM, "hipMalloc of %zu bytes for the deterministic mode failed", total);
        }
        char *c = (char *)p;
        E.det_mem = p;
        E.d.part = (double *)c;
        c += al(b_part);
        E.d.pe = (double *)c;
        c += al(b_pe);
        E.d.pself = (double *)c;
        c += al(b_ps);
        int *mtd = (int *)c;
        c += al(b_mt);
        E.d.prow = (double *)c;
        c += al(b_pr);
        E.d.dcol = (int *)c;
        c += al(b_dc);
        E.d.dcolcnt = (int *)c;
        HIPCHK(hipMemcpyAsync(mtd, mt.data(), b_mt, hipMemcpyHostToDevice, E.stream));
        HIPCHK(hipStreamSynchronize(E.stream));
        E.d.mol_tile0 = mtd;
    }
    E.par.det = on;
    for (int k = 0; k < 3; k++) E.par.split_sh[k] = default_split(E, k);
    default_grids(E);
    /* the column lists of the current list */
    if (on && E.have_list) {
        hipLaunchKernelGGL(k_det_cols, dim3(E.ntiles), dim3(64), 0, E.stream, E.d);
        LAUNCHCHK();
    }
    return 0;
}

extern "C" int gbhip_set_deterministic(gbhip_ctx *ctx, int on)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    if (on && ctx->E.hcp.on) return seterr(GBHIP_EINVAL, "hcp has no deterministic mode (its forces add with atomics)");
    return set_det(ctx->E, on);
}

extern "C" int gbhip_batch_set_deterministic(gbhip_batch *b, int on)
{
    if (!b) return seterr(GBHIP_EINVAL, "NULL batch");
    return set_det(b->E, on);
}

extern "C" int gbhip_set_derivative_store(gbhip_ctx *ctx, int mode)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    return set_store_mode(ctx->E, mode);
}

extern "C" int gbhip_batch_set_derivative_store(gbhip_batch *b, int mode)
{
    if (!b) return seterr(GBHIP_EINVAL, "NULL batch");
    return set_store_mode(b->E, mode);
}

static int store_stats(Engine &E, long long *need, long long *cap, int *in_use)
{
    long long nd = 0;
    int ok = 0;
    HIPCHK(hipMemcpyAsync(&nd, E.d.store_need, sizeof(nd), hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipMemcpyAsync(&ok, E.d.store_ok, sizeof(ok), hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    if (need) *need = nd;
    if (cap) *cap = E.d.store_cap;
    if (in_use) *in_use = ok;
    return 0;
}

extern "C" int gbhip_store_stats(gbhip_ctx *ctx, long long *need, long long *cap, int *in_use)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    return store_stats(ctx->E, need, cap, in_use);
}

extern "C" int gbhip_batch_store_stats(gbhip_batch *b, long long *need, long long *cap, int *in_use)
{
    if (!b) return seterr(GBHIP_EINVAL, "NULL batch");
    return store_stats(b->E, need, cap, in_use);
}

extern "C" int gbhip_set_launch(gbhip_ctx *ctx, const int *blocks3)
{
    if (!ctx || !blocks3) return seterr(GBHIP_EINVAL, "NULL argument");
    return set_launch(ctx->E, blocks3);
}

extern "C" int gbhip_get_launch(gbhip_ctx *ctx, int *blocks3)
{
    if (!ctx || !blocks3) return seterr(GBHIP_EINVAL, "NULL argument");
    for (int k = 0; k < 3; k++) blocks3[k] = ctx->E.pass_blocks[k];
    return 0;
}

extern "C" int gbhip_set_split(gbhip_ctx *ctx, const int *split3)
{
    if (!ctx || !split3) return seterr(GBHIP_EINVAL, "NULL argument");
    return set_split(ctx->E, split3);
}

extern "C" int gbhip_set_queue_fill(gbhip_ctx *ctx, int pass, double fill)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    if (pass != 0) return seterr(GBHIP_EINVAL, "only pass 1 (pass index 0) has a queued form");
    if (!(fill >= 0.0 && fill <= 2.0)) return seterr(GBHIP_EINVAL, "queue fill must be in [0, 2]");
    ctx->E.par.qfill[0] = (float)fill;
    return 0;
}

extern "C" int gbhip_get_split(gbhip_ctx *ctx, int *split3)
{
    if (!ctx || !split3) return seterr(GBHIP_EINVAL, "NULL argument");
    for (int k = 0; k < 3; k++) split3[k] = 1 << ctx->E.par.split_sh[k];
    return 0;
}

extern "C" int gbhip_autotune(gbhip_ctx *ctx, int reps, int *chosen3)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    if (ctx->E.hcp.on) return seterr(GBHIP_EINVAL, "an hcp context has no pair-kernel launch shapes to tune");
    return autotune(ctx->E, reps, chosen3);
}

extern "C" int gbhip_batch_autotune(gbhip_batch *b, int reps, int *chosen3)
{
    if (!b) return seterr(GBHIP_EINVAL, "NULL batch");
    return autotune(b->E, reps, chosen3);
}

/* ---------------------------------------------------------- single context */

/*
 * VALU peak probe: eight independent FMA chains per lane (no dependency
 * stalls at full occupancy), 2 flops per FMA; the result is stored so the
 * chains are not removed.  Built with the same flags as the pair kernels
 * (packed FP32 off), so it measures the ceiling those kernels can reach.
 */
template <typename R>
__global__ __launch_bounds__(256) void k_valu_probe(R *out, int iters, R a, R b)
{
    R x[8];
#pragma unroll
    for (int u = 0; u < 8; u++) x[u] = (R)(threadIdx.x + u) * R(1e-3);
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int u = 0; u < 8; u++) x[u] = fma(x[u], a, b);
    }
    R sum = R(0);
#pragma unroll
    for (int u = 0; u < 8; u++) sum += x[u];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = sum;
}

extern "C" int gbhip_valu_peak(int precision, double *tflops)
{
    if (!tflops) return seterr(GBHIP_EINVAL, "NULL argument");
    if (precision != GBHIP_FP64 && precision != GBHIP_MIXED) return seterr(GBHIP_EINVAL, "bad precision");
    int dev = 0, ncu = 256;
    HIPCHK(hipGetDevice(&dev));
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const int blocks = ncu * 8, iters = 8192;
    void *out = nullptr;
    std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
    HIPCHK(hipMalloc(&out, (size_t)blocks * 256 * sizeof(double)));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    float best = 1e30f;
    for (int rep = 0; rep < 4; rep++) {
        HIPCHK(hipEventRecord(e0, nullptr));
        if (precision == GBHIP_FP64)
            hipLaunchKernelGGL(k_valu_probe<double>, dim3(blocks), dim3(256), 0, nullptr, (double *)out, iters,
                               0.9999999, 1e-7);
        else
            hipLaunchKernelGGL(k_valu_probe<float>, dim3(blocks), dim3(256), 0, nullptr, (float *)out, iters,
                               0.9999f, 1e-4f);
        HIPCHK(hipEventRecord(e1, nullptr));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        if (rep > 0 && ms < best) best = ms; /* the first launch warms up */
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(hipFree(out));
    const double flops = 2.0 * 8.0 * iters * (double)blocks * 256.0;
    *tflops = flops / (best * 1e-3) / 1e12;
    return 0;
}

extern "C" int gbhip_timing_start(gbhip_ctx *ctx, int max_evals)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    return timing_start(ctx->E, max_evals);
}

extern "C" int gbhip_timing_stop(gbhip_ctx *ctx, double *ms4, int *nevals)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    return timing_stop(ctx->E, ms4, nevals);
}

extern "C" int gbhip_batch_timing_start(gbhip_batch *b, int max_evals)
{
    if (!b) return seterr(GBHIP_EINVAL, "NULL batch");
    return timing_start(b->E, max_evals);
}

extern "C" int gbhip_batch_timing_stop(gbhip_batch *b, double *ms4, int *nevals)
{
    if (!b) return seterr(GBHIP_EINVAL, "NULL batch");
    return timing_stop(b->E, ms4, nevals);
}


extern "C" int gbhip_ctx_create(const gbhip_topology *top, const gbhip_options *opt, const double *x0,
                                gbhip_ctx **out)
{
    return gbhip_ctx_create_on(top, opt, x0, nullptr, out);
}

extern "C" int gbhip_ctx_create_on(const gbhip_topology *top, const gbhip_options *opt, const double *x0,
                                   void *stream, gbhip_ctx **out)
{
    if (!out) return seterr(GBHIP_EINVAL, "out pointer is NULL");
    *out = nullptr;
    gbhip_ctx *c = new (std::nothrow) gbhip_ctx;
    if (!c) return seterr(GBHIP_ENOMEM, "host allocation failed");
    c->E.stream = (hipStream_t)stream;
    const double *xs[1] = {x0};
    std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
    int rc = engine_build(c->E, 1, top, opt, xs);
    if (rc) {
        delete c;
        return rc;
    }
    *out = c;
    return 0;
}

extern "C" int gbhip_ctx_destroy(gbhip_ctx *ctx)
{
    if (!ctx) return 0;
    if (ctx->E.stream) (void)hipStreamSynchronize(ctx->E.stream);
    std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
    if (!ctx->E.stream) (void)hipDeviceSynchronize();
    delete ctx;
    return 0;
}

/* --------------------------------------------- evaluation without allocation */

/*
 * One GB energy + gradient with every device array in a static buffer of
 * the library: `scratch`, GBHIP_STATIC_BYTES per GPU, held until the
 * process ends.  Host code never passes `scratch` itself to the device (its
 * host-side name is not a device address): the address of each GPU's copy
 * is looked up once per device with hipGetSymbolAddress and kept in
 * g_static_ptr.  The context's arrays (atom data, coordinates, tile pairs
 * and bit matrices, accumulators, energy sums, and the pass-3 derivative
 * store when room is left) are carved from that address by the slot
 * mechanism; everything a call reads is written by the same call first
 * (set-up copies and clears, list build, the gathered coordinates and
 * cleared accumulators of the evaluation), so nothing of an earlier call
 * leaks in.  Within the process the calls hold the device lock, so they
 * run one at a time (and never inside another thread's graph capture);
 * processes sharing a GPU each have their own copy.
 */
static __device__ __align__(256) unsigned char scratch[GBHIP_STATIC_BYTES];

#define STATIC_MAX_DEVICES 64
static void *g_static_ptr[STATIC_MAX_DEVICES];

/* device bytes engine_build takes for one molecule (every array rounded up
 * to the 256-byte alignment of the slot allocator), without the store */
static size_t static_build_bytes(size_t natom, size_t ntypes, size_t nexcl)
{
    const size_t NS = (natom + TILE - 1) / TILE * TILE, NT = NS / TILE, cap = NT * (NT + 1) / 2;
    size_t b = 0;
    auto add = [&b](size_t bytes) { b += (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255; };
    for (int k = 0; k < 6; k++) add(NS * sizeof(int));      /* perm slot_mol neck frz ljrow ljcol */
    for (int k = 0; k < 7; k++) add(NS * sizeof(double));   /* q rb ri0 sj ga gbb gg */
    for (int k = 0; k < 3; k++) add(NS * sizeof(float));    /* ri0_f sj_f rb_f */
    for (int k = 0; k < 11; k++) add(NS * sizeof(double));  /* xs ys zs sumi reff psi dfac sumdeijda gx gy gz */
    add(NT * sizeof(int)); add(NT * sizeof(int)); add(NT * sizeof(long long));   /* tile_mol tile_end rowbase */
    add((natom + 1) * sizeof(int)); add(nexcl * sizeof(int)); add(natom * sizeof(int)); /* ex_ptr ex_j inv */
    add(ntypes * ntypes * sizeof(double2)); add(sizeof(int2)); add(sizeof(double));    /* ljab mol_lj rgbpsmax2 */
    add(NECK_N * NECK_N * sizeof(double)); add(NECK_N * NECK_N * sizeof(double));
    add(NECK_N * NECK_N * sizeof(float)); add(NECK_N * NECK_N * sizeof(float));
    add(3 * sizeof(double)); add(3 * ENE_SLOTS * sizeof(double));                     /* ene eslot */
    add(cap * sizeof(int)); add(NT * sizeof(int)); add(NT * sizeof(int)); add(sizeof(int)); /* cand rowcnt rowstart ntp */
    add(cap * sizeof(int2)); add(cap * sizeof(int2)); add(cap * sizeof(int));          /* tp tpo tflip */
    add(cap * sizeof(unsigned long long)); add(cap * sizeof(int)); add(2 * TILE * sizeof(int)); /* tcol tord tbin */
    add(sizeof(int)); add(cap * sizeof(int)); add(cap * sizeof(QMeta)); /* nnz tnp qm */
    add(sizeof(long long)); add(sizeof(int));                                          /* store_need store_ok */
    for (int k = 0; k < 3; k++) add(cap * TILE * sizeof(unsigned long long)); /* mask excl sbal */
    add(sizeof(unsigned long long)); add(3 * NT * sizeof(double)); add(3 * NT * sizeof(double)); /* npairs bblo bbhi */
    add(3 * natom * sizeof(double)); add(3 * natom * sizeof(double)); add(natom * sizeof(double)); /* x g tmp */
    add(14 * sizeof(double));
    return b;
}

extern "C" int gbhip_egb_static(const gbhip_topology *top, const gbhip_options *opt, const double *x, double *grad,
                                double *epol, double *evdw, double *eelec)
{
    if (!top || !opt || !x) return seterr(GBHIP_EINVAL, "NULL argument");
    if (top->natom < 1) return seterr(GBHIP_EINVAL, "natom must be positive");
    if (!top->iblo) return seterr(GBHIP_EINVAL, "topology without exclusion counts");
    long long nexcl = 0;
    for (int i = 0; i < top->natom; i++) nexcl += top->iblo[i];
    const size_t need = static_build_bytes((size_t)top->natom, (size_t)std::max(0, top->ntypes), (size_t)nexcl);
    if (need > GBHIP_STATIC_BYTES)
        return seterr(GBHIP_EINVAL,
                      "gbhip_egb_static: %d atoms need %zu bytes of device memory, more than the %llu-byte "
                      "static buffer; evaluate this system through a context (gbhip_ctx_create)",
                      top->natom, need, (unsigned long long)GBHIP_STATIC_BYTES);
    if (gbhip_device_count() < 1) return seterr(GBHIP_ENODEV, "no GPU is visible to the HIP runtime");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= STATIC_MAX_DEVICES) return seterr(GBHIP_EINVAL, "device %d out of range", dev);
    std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
    if (!g_static_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return seterr(GBHIP_EHIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_static_ptr[dev] = ptr;
    }
    SlotMem sm;
    sm.dbase = (char *)g_static_ptr[dev];
    sm.dsize = GBHIP_STATIC_BYTES;
    sm.in_use = true;
    std::unique_ptr<gbhip_ctx> c(new (std::nothrow) gbhip_ctx);
    if (!c) return seterr(GBHIP_ENOMEM, "host allocation failed");
    Engine &E = c->E;
    E.slot = &sm;
    E.stream = nullptr;
    const double *xs[1] = {x};
    int rc;
    if ((rc = engine_build(E, 1, top, opt, xs)) || (rc = engine_nblist(E)) || (rc = engine_store_reserve(E)))
        return rc;
    const size_t n3 = 3 * (size_t)E.natoms;
    if (grad) {
        hipLaunchKernelGGL(k_zero, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, E.stream, E.g_dev, (long long)n3);
        LAUNCHCHK();
    }
    if ((rc = engine_eval(E, E.x_dev, grad ? E.g_dev : nullptr))) return rc;
    double ene[3];
    std::vector<double> g(grad ? n3 : 0);
    HIPCHK(hipMemcpyAsync(ene, E.d.ene, sizeof(ene), hipMemcpyDeviceToHost, E.stream));
    if (grad) HIPCHK(hipMemcpyAsync(g.data(), E.g_dev, n3 * sizeof(double), hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    if (grad)
        for (size_t k = 0; k < n3; k++) grad[k] += g[k];
    if (epol) *epol = ene[0];
    if (evdw) *evdw = ene[1];
    if (eelec) *eelec = ene[2];
    return 0;
}

/* ------------------------------------------------------------ slots */

struct gbhip_slot {
    SlotMem m;
};

extern "C" int gbhip_slot_create(size_t device_bytes, size_t io_bytes, size_t upload_bytes, void *stream,
                                 gbhip_slot **out)
{
    if (!out) return seterr(GBHIP_EINVAL, "out pointer is NULL");
    *out = nullptr;
    if (gbhip_device_count() < 1) return seterr(GBHIP_ENODEV, "no GPU is visible to the HIP runtime");
    gbhip_slot *sl = new (std::nothrow) gbhip_slot;
    if (!sl) return seterr(GBHIP_ENOMEM, "host allocation failed");
    SlotMem &m = sl->m;
    std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
    int rc = 0;
    if (device_bytes && hipMalloc((void **)&m.dbase, device_bytes) != hipSuccess)
        rc = seterr(GBHIP_ENOMEM, "hipMalloc of %zu bytes failed", device_bytes);
    else if (device_bytes) {
        m.dsize = device_bytes;
        m.own_device = true;
    }
    if (!rc && io_bytes && hipHostMalloc((void **)&m.hio, io_bytes, hipHostMallocDefault) != hipSuccess)
        rc = seterr(GBHIP_ENOMEM, "hipHostMalloc of %zu bytes failed", io_bytes);
    else if (!rc)
        m.hio_size = io_bytes;
    if (!rc && upload_bytes && hipHostMalloc((void **)&m.hup, upload_bytes, hipHostMallocDefault) != hipSuccess)
        rc = seterr(GBHIP_ENOMEM, "hipHostMalloc of %zu bytes failed", upload_bytes);
    else if (!rc)
        m.hup_size = upload_bytes;
    if (!rc && hipEventCreateWithFlags(&m.ev, hipEventDisableTiming) != hipSuccess)
        rc = seterr(GBHIP_EHIP, "hipEventCreate failed");
    m.stream = (hipStream_t)stream;
    if (rc) {
        (void)hipGetLastError();
        if (m.own_device) (void)hipFree(m.dbase);
        if (m.hio) (void)hipHostFree(m.hio);
        if (m.hup) (void)hipHostFree(m.hup);
        if (m.ev) (void)hipEventDestroy(m.ev);
        delete sl;
        return rc;
    }
    *out = sl;
    return 0;
}

extern "C" int gbhip_slot_set_region(gbhip_slot *sl, void *device_base, size_t device_bytes)
{
    if (!sl) return seterr(GBHIP_EINVAL, "NULL slot");
    if (sl->m.in_use) return seterr(GBHIP_ESTATE, "a context lives in this slot");
    if (sl->m.own_device) return seterr(GBHIP_ESTATE, "the slot owns its device region");
    sl->m.dbase = (char *)device_base;
    sl->m.dsize = device_base ? device_bytes : 0;
    sl->m.dused = 0;
    return 0;
}

extern "C" int gbhip_slot_destroy(gbhip_slot *sl)
{
    if (!sl) return 0;
    if (sl->m.in_use) return seterr(GBHIP_ESTATE, "a context lives in this slot");
    std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
    if (sl->m.stream) (void)hipStreamSynchronize(sl->m.stream);
    if (sl->m.own_device) (void)hipFree(sl->m.dbase);
    if (sl->m.hio) (void)hipHostFree(sl->m.hio);
    if (sl->m.hup) (void)hipHostFree(sl->m.hup);
    if (sl->m.ev) (void)hipEventDestroy(sl->m.ev);
    delete sl;
    return 0;
}

extern "C" int gbhip_ctx_create_in_slot(const gbhip_topology *top, const gbhip_options *opt, const double *x0,
                                        gbhip_slot *sl, gbhip_ctx **out)
{
    if (!out || !sl) return seterr(GBHIP_EINVAL, "NULL argument");
    *out = nullptr;
    if (sl->m.in_use) return seterr(GBHIP_ESTATE, "a context already lives in this slot");
    if (!sl->m.dbase) return seterr(GBHIP_ESTATE, "the slot has no device region");
    gbhip_ctx *c = new (std::nothrow) gbhip_ctx;
    if (!c) return seterr(GBHIP_ENOMEM, "host allocation failed");
    sl->m.dused = 0;
    sl->m.hup_used = 0;
    sl->m.in_use = true;
    c->E.slot = &sl->m;
    c->E.stream = sl->m.stream;
    const double *xs[1] = {x0};
    /* nothing is allocated: the lock is not needed for the set-up itself */
    int rc = engine_build(c->E, 1, top, opt, xs);
    if (rc) {
        delete c; /* hands the region back */
        return rc;
    }
    *out = c;
    return 0;
}

extern "C" int gbhip_set_stream(gbhip_ctx *ctx, void *stream)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    /* work already queued on the old stream finishes before any on the new
     * one can start (the context's calls assume one ordered stream) */
    if ((hipStream_t)stream != ctx->E.stream) {
        HIPCHK(hipStreamSynchronize(ctx->E.stream));
        /* a stream the engine created (for dynamics or the mme graph) is
         * kept until the context goes: the graphs captured on it live on
         * and are replayed on the new stream; the caller's stream is never
         * destroyed by the engine */
        if (ctx->E.own_stream && ctx->E.stream) ctx->E.retired_streams.push_back(ctx->E.stream);
        ctx->E.own_stream = false;
    }
    ctx->E.stream = (hipStream_t)stream;
    return 0;
}

/*
 * Pinned host staging (allocated at first use): copies between pinned
 * memory and the device are plain DMA transfers, where copies from pageable
 * memory go through the runtime's bounce buffers (measured at 600 atoms:
 * about 30 us per copy, most of a minimizer step's host time).  Returns
 * false when pinned memory cannot be had; callers then copy directly.
 */
bool engine_stage(Engine &E)
{
    if (E.h_stage) return true;
    if (E.stage_failed) return false;
    if (E.slot) {
        const size_t need = 6 * (size_t)E.natoms + 64;
        if (!E.slot->hio || E.slot->hio_size < need * sizeof(double)) {
            E.stage_failed = true;
            return false;
        }
        E.h_stage = (double *)E.slot->hio;
        E.h_stage_doubles = E.slot->hio_size / sizeof(double);
        E.stage_ev = E.slot->ev;
        return true;
    }
    std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
    void *p = nullptr;
    hipEvent_t ev = nullptr;
    if (hipHostMalloc(&p, sizeof(double) * (6 * (size_t)E.natoms + 64), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        E.stage_failed = true;
        return false;
    }
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipHostFree(p);
        E.stage_failed = true;
        return false;
    }
    E.h_stage = (double *)p;
    E.h_stage_doubles = 6 * (size_t)E.natoms + 64;
    E.stage_ev = ev;
    return true;
}

int engine_d2h(Engine &E, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return 0;
    if (engine_stage(E) && bytes <= E.h_stage_doubles * sizeof(double)) {
        HIPCHK(hipEventSynchronize(E.stage_ev));
        HIPCHK(hipMemcpyAsync(E.h_stage, src, bytes, hipMemcpyDeviceToHost, E.stream));
        HIPCHK(hipStreamSynchronize(E.stream));
        memcpy(dst, E.h_stage, bytes);
        return 0;
    }
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    return 0;
}

extern "C" int gbhip_set_coords(gbhip_ctx *ctx, const double *x)
{
    if (!ctx || !x) return seterr(GBHIP_EINVAL, "NULL argument");
    Engine &E = ctx->E;
    const size_t nb = 3 * sizeof(double) * (size_t)E.natoms;
    if (!engine_stage(E)) {
        HIPCHK(hipMemcpyAsync(E.x_dev, x, nb, hipMemcpyHostToDevice, E.stream));
        return 0;
    }
    /* the previous copy out of the staging buffer has finished */
    HIPCHK(hipEventSynchronize(E.stage_ev));
    memcpy(E.h_stage, x, nb);
    HIPCHK(hipMemcpyAsync(E.x_dev, E.h_stage, nb, hipMemcpyHostToDevice, E.stream));
    HIPCHK(hipEventRecord(E.stage_ev, E.stream));
    return 0;
}

extern "C" int gbhip_set_coords_device(gbhip_ctx *ctx, const double *d_x)
{
    if (!ctx || !d_x) return seterr(GBHIP_EINVAL, "NULL argument");
    Engine &E = ctx->E;
    HIPCHK(hipMemcpyAsync(E.x_dev, d_x, 3 * sizeof(double) * (size_t)E.natoms, hipMemcpyDeviceToDevice, E.stream));
    return 0;
}

extern "C" int gbhip_nblist_async(gbhip_ctx *ctx)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    return engine_nblist(ctx->E);
}

extern "C" int gbhip_npairs(gbhip_ctx *ctx, long long *npairs)
{
    if (!ctx || !npairs) return seterr(GBHIP_EINVAL, "NULL argument");
    Engine &E = ctx->E;
    if (E.hcp.on) {
        *npairs = 0;
        return 0;
    }
    unsigned long long v = 0;
    const int rc = engine_d2h(E, &v, E.d.npairs, sizeof(v));
    if (rc) return rc;
    *npairs = (long long)v;
    return 0;
}

extern "C" int gbhip_nblist(gbhip_ctx *ctx, long long *nentries)
{
    int rc = gbhip_nblist_async(ctx);
    if (rc) return rc;
    if ((rc = engine_store_reserve(ctx->E))) return rc;
    long long np = 0;
    if ((rc = gbhip_npairs(ctx, &np))) return rc;
    if (nentries) *nentries = 2 * np;
    return 0;
}

extern "C" int gbhip_tile_stats(gbhip_ctx *ctx, int *ntiles, int *ntilepairs)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    Engine &E = ctx->E;
    int n = 0;
    HIPCHK(hipMemcpyAsync(&n, E.d.ntp, sizeof(int), hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    if (ntiles) *ntiles = E.ntiles;
    if (ntilepairs) *ntilepairs = n;
    return 0;
}

extern "C" int gbhip_egb_device(gbhip_ctx *ctx, double *d_grad, double *d_ene)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    Engine &E = ctx->E;
    if (!d_grad) {
        /* gradient into the context's own buffer, cleared first */
        const long long n3 = 3LL * E.natoms;
        hipLaunchKernelGGL(k_zero, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, E.stream, E.g_dev, n3);
        LAUNCHCHK();
    }
    int rc = engine_eval(E, E.x_dev, d_grad ? d_grad : E.g_dev);
    if (rc) return rc;
    if (d_ene) HIPCHK(hipMemcpyAsync(d_ene, E.d.ene, 3 * sizeof(double), hipMemcpyDeviceToDevice, E.stream));
    return 0;
}

extern "C" int gbhip_egb(gbhip_ctx *ctx, double *grad, double *epol, double *evdw, double *eelec)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    Engine &E = ctx->E;
    const size_t nb = 3 * sizeof(double) * (size_t)E.natoms;
    if (grad) HIPCHK(hipMemsetAsync(E.g_dev, 0, nb, E.stream));
    int rc = engine_eval(E, E.x_dev, grad ? E.g_dev : nullptr);
    if (rc) return rc;
    double ene[3];
    HIPCHK(hipMemcpyAsync(ene, E.d.ene, sizeof(ene), hipMemcpyDeviceToHost, E.stream));
    std::vector<double> g;
    if (grad) {
        g.resize(3 * (size_t)E.natoms);
        HIPCHK(hipMemcpyAsync(g.data(), E.g_dev, nb, hipMemcpyDeviceToHost, E.stream));
    }
    HIPCHK(hipStreamSynchronize(E.stream));
    if (grad)
        for (size_t k = 0; k < g.size(); k++) grad[k] += g[k];
    if (epol) *epol = ene[0];
    if (evdw) *evdw = ene[1];
    if (eelec) *eelec = ene[2];
    return 0;
}

static int read_unsorted(Engine &E, const double *src, double *dst)
{
    if (!E.have_eval) return seterr(GBHIP_ESTATE, "no evaluation has run yet");
    hipLaunchKernelGGL(k_unsort, dim3((E.nslots + 255) / 256), dim3(256), 0, E.stream, E.d, E.par, src, E.tmp_dev);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(dst, E.tmp_dev, sizeof(double) * (size_t)E.natoms, hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    return 0;
}

extern "C" int gbhip_get_reff(gbhip_ctx *ctx, double *reff)
{
    if (!ctx || !reff) return seterr(GBHIP_EINVAL, "NULL argument");
    return read_unsorted(ctx->E, ctx->E.d.reff, reff);
}

extern "C" int gbhip_get_sumdeijda(gbhip_ctx *ctx, double *sd)
{
    if (!ctx || !sd) return seterr(GBHIP_EINVAL, "NULL argument");
    return read_unsorted(ctx->E, ctx->E.d.sumdeijda, sd);
}

extern "C" int gbhip_get_pairs(gbhip_ctx *ctx, int *pi, int *pj, long long cap, long long *npairs)
{
    if (!ctx || !npairs || cap < 0 || (cap > 0 && (!pi || !pj))) return seterr(GBHIP_EINVAL, "bad argument");
    Engine &E = ctx->E;
    if (!E.have_list) return seterr(GBHIP_ESTATE, "no pair list has been built");
    int *dpi = nullptr, *dpj = nullptr;
    unsigned long long *dcnt = nullptr, cnt = 0;
    const size_t n = cap > 0 ? (size_t)cap : 1;
    if (hipMalloc(&dpi, n * sizeof(int)) != hipSuccess || hipMalloc(&dpj, n * sizeof(int)) != hipSuccess ||
        hipMalloc(&dcnt, sizeof(unsigned long long)) != hipSuccess) {
        (void)hipFree(dpi);
        (void)hipFree(dpj);
        (void)hipFree(dcnt);
        return seterr(GBHIP_ENOMEM, "read-back buffers");
    }
    int rc = 0;
    if (hipMemsetAsync(dcnt, 0, sizeof(cnt), E.stream) != hipSuccess) rc = seterr(GBHIP_EHIP, "memset");
    if (!rc) {
        hipLaunchKernelGGL(k_listpairs, dim3(E.pair_blocks), dim3(64), 0, E.stream, E.d, dpi, dpj, dcnt, cap);
        if (hipGetLastError() != hipSuccess) rc = seterr(GBHIP_EHIP, "k_listpairs launch");
    }
    if (!rc && hipMemcpyAsync(&cnt, dcnt, sizeof(cnt), hipMemcpyDeviceToHost, E.stream) != hipSuccess)
        rc = seterr(GBHIP_EHIP, "copy");
    if (!rc && hipStreamSynchronize(E.stream) != hipSuccess) rc = seterr(GBHIP_EHIP, "sync");
    const long long m = std::min<long long>((long long)cnt, cap);
    if (!rc && m > 0 &&
        (hipMemcpy(pi, dpi, m * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
         hipMemcpy(pj, dpj, m * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess))
        rc = seterr(GBHIP_EHIP, "copy pairs");
    (void)hipFree(dpi);
    (void)hipFree(dpj);
    (void)hipFree(dcnt);
    *npairs = (long long)cnt;
    return rc;
}

extern "C" int gbhip_sync(gbhip_ctx *ctx)
{
    if (!ctx) return seterr(GBHIP_EINVAL, "NULL context");
    HIPCHK(hipStreamSynchronize(ctx->E.stream));
    return 0;
}

/* ------------------------------------------------------------------ batch */


extern "C" int gbhip_batch_create(int nmol, const gbhip_topology *tops, const gbhip_options *opt,
                                  const double *const *x0, gbhip_batch **out)
{
    if (!out || nmol < 1 || !tops) return seterr(GBHIP_EINVAL, "bad batch arguments");
    *out = nullptr;
    gbhip_batch *b = new (std::nothrow) gbhip_batch;
    if (!b) return seterr(GBHIP_ENOMEM, "host allocation failed");
    std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
    int rc = engine_build(b->E, nmol, tops, opt, x0);
    if (!rc) rc = engine_nblist(b->E);
    if (!rc) rc = engine_store_reserve(b->E);
    if (!rc && hipStreamSynchronize(b->E.stream) != hipSuccess) rc = seterr(GBHIP_EHIP, "pair-list build failed");
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return 0;
}

extern "C" int gbhip_batch_destroy(gbhip_batch *b)
{
    if (!b) return 0;
    std::lock_guard<std::recursive_mutex> lk(device_alloc_lock());
    (void)hipDeviceSynchronize();
    delete b;
    return 0;
}

extern "C" int gbhip_batch_set_coords(gbhip_batch *b, int m, const double *x)
{
    if (!b || !x || m < 0 || m >= b->E.nmol) return seterr(GBHIP_EINVAL, "bad molecule index");
    Engine &E = b->E;
    HIPCHK(hipMemcpyAsync(E.x_dev + 3 * (size_t)E.mol_atom_off[m], x, 3 * sizeof(double) * (size_t)E.mol_natom[m],
                          hipMemcpyHostToDevice, E.stream));
    return 0;
}

extern "C" int gbhip_batch_nblist(gbhip_batch *b)
{
    if (!b) return seterr(GBHIP_EINVAL, "NULL batch");
    return engine_nblist(b->E);
}

extern "C" int gbhip_batch_eval_device(gbhip_batch *b)
{
    if (!b) return seterr(GBHIP_EINVAL, "NULL batch");
    return engine_eval(b->E, b->E.x_dev, nullptr);
}

extern "C" int gbhip_batch_eval(gbhip_batch *b, double *ene, double *const *grads)
{
    if (!b) return seterr(GBHIP_EINVAL, "NULL batch");
    Engine &E = b->E;
    const size_t nb = 3 * sizeof(double) * (size_t)E.natoms;
    bool want_g = false;
    if (grads)
        for (int m = 0; m < E.nmol; m++) want_g |= grads[m] != nullptr;
    if (want_g) HIPCHK(hipMemsetAsync(E.g_dev, 0, nb, E.stream));
    int rc = engine_eval(E, E.x_dev, want_g ? E.g_dev : nullptr);
    if (rc) return rc;
    std::vector<double> g;
    if (want_g) {
        g.resize(3 * (size_t)E.natoms);
        HIPCHK(hipMemcpyAsync(g.data(), E.g_dev, nb, hipMemcpyDeviceToHost, E.stream));
    }
    if (ene) HIPCHK(hipMemcpyAsync(ene, E.d.ene, 3 * sizeof(double) * (size_t)E.nmol, hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    if (want_g)
        for (int m = 0; m < E.nmol; m++) {
            if (!grads[m]) continue;
            const double *src = g.data() + 3 * (size_t)E.mol_atom_off[m];
            for (int k = 0; k < 3 * E.mol_natom[m]; k++) grads[m][k] += src[k];
        }
    return 0;
}

extern "C" int gbhip_batch_stats(gbhip_batch *b, long long *npairs, long long *natoms, int *ntilepairs)
{
    if (!b) return seterr(GBHIP_EINVAL, "NULL batch");
    Engine &E = b->E;
    unsigned long long v = 0;
    int n = 0;
    HIPCHK(hipMemcpyAsync(&v, E.d.npairs, sizeof(v), hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipMemcpyAsync(&n, E.d.ntp, sizeof(int), hipMemcpyDeviceToHost, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    if (npairs) *npairs = (long long)v;
    if (natoms) *natoms = E.natoms;
    if (ntilepairs) *ntilepairs = n;
    return 0;
}
