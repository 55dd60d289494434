/*
 * serve.cpp -- continuous-batching request loop for minimisation and
 * dynamics jobs on one GPU.
 *
 * Jobs arrive at given times (a Poisson stream in the benchmark).  A job is
 * admitted, in arrival order, as soon as it has arrived, one of `slots`
 * workers is free and a contiguous block of its device footprint
 * (gbhip_footprint_bytes) is free in the device pool; otherwise it waits and
 * is admitted the moment a running job hands back its slot and block.  The
 * pool is one device allocation of the HBM budget made when the loop
 * starts; each worker owns a HIP stream (default priority) and a slot
 * (gbhip_slot: pinned staging and an event) made at the start as well, so
 * while jobs run nothing is allocated, freed or synchronised device-wide:
 * a job's context is built inside its pool block (gbhip_ctx_create_in_slot)
 * and handing it back frees nothing.  The device work of all admitted jobs
 * overlaps.  A job keeps its whole state on the device (coordinates, pair
 * list, minimiser vectors or velocities, captured graphs):
 *   - CG: gbhip_conjgrad (conjgrad.c:34 with device-resident vectors) over
 *     the device mme() (bonded + 1-4 + GB), pair list every nsnb calls;
 *   - MD: gbhip_md_init + gbhip_md_run, every step one hipGraph launch.
 * Latency = completion - arrival (queueing included).  Optionally the final
 * coordinates are re-evaluated in FP64 to report the precision error.
 *
 * A watchdog thread reports a job that has stayed in one phase (set-up,
 * minimisation, dynamics, read-back, hand-back) longer than
 * cfg.watchdog_s, naming the job, its worker and the phase.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "gbhip.h"
#include "host_internal.h"
#include "sff_gb.h"

#include "gbserve.h"

namespace {

double now_s(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

/*
 * First-fit allocator over the device pool (offsets; 4 KiB granules): free
 * blocks by offset, adjacent free blocks merged on release.
 */
class Pool {
  public:
    explicit Pool(size_t size) { free_[0] = size; }
    static size_t round(size_t b) { return (b + 4095) & ~(size_t)4095; }
    /* offset of a block of b bytes, or SIZE_MAX */
    size_t take(size_t b)
    {
        b = round(b);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < b) continue;
            const size_t off = it->first, left = it->second - b;
            free_.erase(it);
            if (left) free_[off + b] = left;
            return off;
        }
        return SIZE_MAX;
    }
    void give(size_t off, size_t b)
    {
        b = round(b);
        auto it = free_.emplace(off, b).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

  private:
    std::map<size_t, size_t> free_;
};

enum Phase { PH_IDLE = 0, PH_SETUP, PH_CG, PH_MD_INIT, PH_MD_RUN, PH_READBACK, PH_RELEASE };
const char *phase_name(int p)
{
    static const char *names[] = {"idle", "set-up", "minimisation", "dynamics set-up", "dynamics",
                                  "read-back", "hand-back"};
    return p >= 0 && p <= PH_RELEASE ? names[p] : "?";
}

struct Server {
    const gbserve_job *jobs;
    gbserve_result *res;
    int njobs;
    gbserve_config cfg;
    nabc_opts_t opts;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv;
    std::vector<int> assigned;     /* per worker: job index or -1 */
    std::vector<size_t> block;     /* per job: pool offset */
    std::vector<hipStream_t> streams;
    std::vector<gbhip_slot *> slots;
    char *pool_base = nullptr;     /* device pool (nullptr without a device) */
    Pool *pool = nullptr;
    std::vector<std::vector<double>> xfinal; /* per job, kept for the precision check */
    gbserve_runner runner = nullptr; /* caller's job function (gbserve_run_custom) */
    void *user = nullptr;
    int done = 0;
    bool stop = false;
    size_t mem_in_use = 0;
    int active = 0;
    /* watchdog: per worker, the phase and when it began */
    std::unique_ptr<std::atomic<int>[]> phase;
    std::unique_ptr<std::atomic<double>[]> phase_t;
};

void set_phase(Server &S, int w, int p)
{
    S.phase_t[w].store(now_s(S.t0));
    S.phase[w].store(p);
}

void run_job(Server &S, int w, int k)
{
    const gbserve_job &jb = S.jobs[k];
    gbserve_result &r = S.res[k];
    nabc_opts_t o = S.opts;
    o.gbprec = S.cfg.precision == GBHIP_MIXED ? 1 : 0;
    nabc_model m;
    gbhip_ctx *ctx = nullptr;
    PARMSTRUCT_T *prm = (PARMSTRUCT_T *)jb.prm;
    const int n = prm->Natom;
    std::vector<double> x(jb.x, jb.x + 3 * (size_t)n);
    r.status = -100;
    set_phase(S, w, PH_SETUP);
    if (gbhip_slot_set_region(S.slots[w], S.pool_base + S.block[k], r.footprint) != 0 ||
        nabc_model_build(&m, prm, nullptr, &o) != 0) {
        r.status = -10;
        return;
    }
    if (nabc_model_context_in(&m, x.data(), &o, S.streams[w], S.slots[w], &ctx) != 0) {
        nabc_model_free(&m);
        r.status = -11;
        return;
    }
    r.setup_s = now_s(S.t0) - r.start_s;
    const int nsnb = o.nsnb > 0 ? o.nsnb : 1;
    if (jb.kind == GBSERVE_CG) {
        set_phase(S, w, PH_CG);
        double fret = 0.0;
        int ret = 0;
        if (gbhip_conjgrad(ctx, x.data(), 3 * n, &fret, S.cfg.cg_rmsgrad, S.cfg.cg_dfpred, jb.nsteps, nsnb, nullptr,
                           nullptr, &ret) != 0)
            r.status = -12;
        else
            r.status = ret;
        r.energy = fret;
    } else {
        gbhip_md_options mo;
        std::memset(&mo, 0, sizeof(mo));
        mo.dt = o.dt;
        mo.temp0 = o.temp0;
        mo.tempi = o.tempi;
        mo.gamma_ln = o.gamma_ln;
        mo.tautp = o.tautp;
        mo.vlimit = o.vlimit;
        mo.boltz2 = o.boltz2;
        mo.nsnb = nsnb;
        mo.seed = jb.seed;
        const int cap = std::max(1, jb.nsteps);
        std::vector<double> row(16 * (size_t)cap);
        int nr = 0;
        set_phase(S, w, PH_MD_INIT);
        int rc = gbhip_md_init(ctx, prm->Masses, x.data(), nullptr, &mo, cap);
        if (!rc) {
            set_phase(S, w, PH_MD_RUN);
            rc = gbhip_md_run(ctx, jb.nsteps, 1);
        }
        if (!rc) {
            set_phase(S, w, PH_READBACK);
            rc = gbhip_md_state(ctx, x.data(), nullptr, nullptr);
        }
        if (!rc) rc = gbhip_md_log(ctx, row.data(), jb.nsteps, &nr);
        if (rc || nr < 1) {
            r.status = -13;
        } else {
            r.status = 0;
            r.energy = row[16 * (size_t)(nr - 1)];
        }
    }
    set_phase(S, w, PH_RELEASE);
    r.end_s = now_s(S.t0);
    r.latency_s = r.end_s - jb.arrival_s;
    gbhip_ctx_destroy(ctx);
    nabc_model_free(&m);
    if (jb.x_out && r.status >= -4) std::memcpy(jb.x_out, x.data(), sizeof(double) * 3 * (size_t)n);
    if (S.cfg.check_fp64) S.xfinal[k].swap(x);
}

/* the final coordinates of a job evaluated in FP64 and in the job precision
 * (fresh list each), after the timed run */
void check_job(Server &S, int k)
{
    const gbserve_job &jb = S.jobs[k];
    gbserve_result &r = S.res[k];
    PARMSTRUCT_T *prm = (PARMSTRUCT_T *)jb.prm;
    std::vector<double> &x = S.xfinal[k];
    if (r.status < -4 || x.empty()) return;
    std::vector<double> f(x.size());
    nabc_model m;
    if (nabc_model_build(&m, prm, nullptr, &S.opts) != 0) return;
    for (int prec = 0; prec < 2; prec++) {
        nabc_opts_t o = S.opts;
        o.gbprec = prec == 0 ? 0 : (S.cfg.precision == GBHIP_MIXED ? 1 : 0);
        gbhip_ctx *c = nullptr;
        if (nabc_model_context(&m, x.data(), &o, S.streams[0], &c) != 0) continue;
        double e[14];
        if (gbhip_set_coords(c, x.data()) == 0 && gbhip_nblist(c, nullptr) == 0 && gbhip_mme(c, f.data(), e) == 0) {
            if (prec == 0) r.energy_fp64 = e[0];
            else r.energy_final_eval = e[0];
        }
        gbhip_ctx_destroy(c);
    }
    nabc_model_free(&m);
}

void worker(Server *S, int w)
{
    for (;;) {
        int k;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv.wait(lk, [&] { return S->stop || S->assigned[w] >= 0; });
            if (S->assigned[w] < 0) return;
            k = S->assigned[w];
        }
        if (S->runner) {
            const gbserve_job &jb = S->jobs[k];
            gbserve_result &r = S->res[k];
            double energy = 0.0;
            set_phase(*S, w, PH_SETUP);
            r.status = S->runner(&jb, w, S->user, &energy);
            r.energy = energy;
            r.end_s = now_s(S->t0);
            r.latency_s = r.end_s - jb.arrival_s;
        } else {
            run_job(*S, w, k);
        }
        set_phase(*S, w, PH_IDLE);
        int done;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->assigned[w] = -1;
            S->mem_in_use -= S->res[k].footprint;
            S->pool->give(S->block[k], S->res[k].footprint);
            S->active--;
            done = ++S->done;
        }
        S->cv.notify_all();
        if (S->cfg.progress > 0 && done % S->cfg.progress == 0) {
            std::fprintf(stderr, "gbserve: %d of %d jobs done at %.2f s\n", done, S->njobs, now_s(S->t0));
            std::fflush(stderr);
        }
    }
}

/* reports (once per phase) a worker that has stayed in one phase of a job
 * longer than the limit */
void watchdog(Server *S, double limit_s)
{
    const int nw = (int)S->assigned.size();
    std::vector<double> reported(nw, -1.0);
    std::unique_lock<std::mutex> lk(S->mu);
    while (!S->stop) {
        S->cv.wait_for(lk, std::chrono::milliseconds(250));
        const double t = now_s(S->t0);
        for (int w = 0; w < nw; w++) {
            const int p = S->phase[w].load(), k = S->assigned[w];
            const double since = S->phase_t[w].load();
            if (p == PH_IDLE || k < 0 || t - since < limit_s || reported[w] == since) continue;
            reported[w] = since;
            std::fprintf(stderr, "gbserve watchdog: job %d (%d atoms, %s, %d steps) on worker %d in phase '%s' for %.1f s\n",
                         k, ((const PARMSTRUCT_T *)S->jobs[k].prm)->Natom,
                         S->jobs[k].kind == GBSERVE_CG ? "CG" : "MD", S->jobs[k].nsteps, w, phase_name(p), t - since);
            std::fflush(stderr);
        }
    }
}

} // namespace

extern "C" size_t gbhip_footprint_bytes(int natom, int md_steps)
{
    const size_t NS = ((size_t)natom + TILE_ATOMS - 1) / TILE_ATOMS * TILE_ATOMS, NT = NS / TILE_ATOMS;
    const size_t cap = NT * (NT + 1) / 2;
    size_t b = 0;
    b += NS * (6 * 4 + 7 * 8 + 4 * 4 + 11 * 8);          /* per-slot constants and work arrays */
    b += NT * (4 * 4 + 8 + 2 * 24);                       /* tile arrays */
    b += cap * (4 + 8 + 8 + 4 + 8 + 4 + 3 * TILE_ATOMS * 8); /* candidates, tile pairs (both orders, flags,
                                                              column sets, queue), bit matrices, store
                                                              step masks */
    /* pass-3 derivative store: at most every tile pair full, 16 bytes per
     * entry (FP64), 1.25x headroom */
    b += (cap * TILE_ATOMS * TILE_ATOMS + TILE_ATOMS * TILE_ATOMS) * 20;
    b += (size_t)natom * (4 + 4 + 12 + 2 * 24 + 8);       /* inverse order, exclusions, x, g, scratch */
    b += (size_t)natom * 220;                              /* bonded / 1-4 tables (chain topology bound) */
    b += (size_t)natom * 4 * 24 + 16 * 8 * (size_t)(md_steps > 0 ? md_steps : 0); /* dynamics state + log */
    b += (size_t)natom * 6 * 24 + 6 * 256 * 8 + 64;        /* minimiser vectors and reductions */
    b += 3 * 64 * 8;                                       /* energy partial sums */
    b += 97 * 256;                                         /* 256-byte alignment of each array */
    return b;
}

static int serve_loop(const gbserve_job *jobs, int njobs, const gbserve_config *cfg, const char *mm_opts,
                      gbserve_result *res, gbserve_runner runner, void *user)
{
    if (!jobs || !res || !cfg || njobs < 0 || cfg->slots < 1) {
        nabc_seterr("gbserve_run: bad arguments");
        return -1;
    }
    Server S;
    S.jobs = jobs;
    S.res = res;
    S.njobs = njobs;
    S.cfg = *cfg;
    S.runner = runner;
    S.user = user;
    if (runner) S.cfg.check_fp64 = 0;
    if (!runner) {
        /* options for the jobs: current mm_options state plus the given string */
        nabc_opts_t saved = nabc_opts;
        if (mm_opts && mm_options((char *)mm_opts) != 0) {
            nabc_opts = saved;
            return -1;
        }
        S.opts = nabc_opts;
        nabc_opts = saved;
    }
    std::vector<int> order(njobs);
    int max_atoms = 1, max_steps = 0;
    for (int k = 0; k < njobs; k++) {
        order[k] = k;
        std::memset(&res[k], 0, sizeof(res[k]));
        const PARMSTRUCT_T *p = (const PARMSTRUCT_T *)jobs[k].prm;
        res[k].footprint = gbhip_footprint_bytes(p->Natom, jobs[k].kind == GBSERVE_MD ? jobs[k].nsteps : 0);
        res[k].slot = -1;
        res[k].status = -100;
        if (res[k].footprint <= cfg->hbm_budget) {
            max_atoms = std::max(max_atoms, p->Natom);
            if (jobs[k].kind == GBSERVE_MD) max_steps = std::max(max_steps, jobs[k].nsteps);
        }
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return jobs[a].arrival_s < jobs[b].arrival_s; });
    S.assigned.assign(cfg->slots, -1);
    S.block.assign(njobs, 0);
    S.xfinal.assign(njobs, std::vector<double>());
    S.streams.assign(cfg->slots, nullptr);
    S.slots.assign(cfg->slots, nullptr);
    S.phase.reset(new std::atomic<int>[cfg->slots]);
    S.phase_t.reset(new std::atomic<double>[cfg->slots]);
    for (int w = 0; w < cfg->slots; w++) {
        S.phase[w].store(PH_IDLE);
        S.phase_t[w].store(0.0);
    }
    Pool pool(cfg->hbm_budget);
    S.pool = &pool;
    auto release = [&]() {
        for (gbhip_slot *sl : S.slots)
            if (sl) (void)gbhip_slot_destroy(sl);
        for (hipStream_t s : S.streams)
            if (s) (void)hipStreamDestroy(s);
        if (S.pool_base) (void)hipFree(S.pool_base);
    };
    if (!runner) {
        /* everything the jobs need, made once: the device pool, and per
         * worker a stream and a slot's pinned staging */
        const size_t io = 8 * std::max<size_t>(6 * (size_t)max_atoms + 64, 16 * (size_t)std::max(1, max_steps)) + 4096;
        const size_t up = std::min<size_t>(64 << 20, 1024 * (size_t)max_atoms + (1 << 20));
        if (hipMalloc((void **)&S.pool_base, cfg->hbm_budget) != hipSuccess) {
            (void)hipGetLastError();
            nabc_seterr("gbserve_run: device pool of %zu bytes could not be allocated", cfg->hbm_budget);
            S.pool_base = nullptr;
            return -1;
        }
        for (int w = 0; w < cfg->slots; w++) {
            if (hipStreamCreateWithFlags(&S.streams[w], hipStreamNonBlocking) != hipSuccess ||
                gbhip_slot_create(0, io, up, S.streams[w], &S.slots[w]) != 0) {
                nabc_seterr("gbserve_run: stream or slot creation failed: %s", gbhip_last_error());
                release();
                return -1;
            }
        }
    }
    std::vector<std::thread> threads;
    S.t0 = std::chrono::steady_clock::now();
    for (int w = 0; w < cfg->slots; w++) threads.emplace_back(worker, &S, w);
    std::thread dog;
    if (cfg->watchdog_s >= 0.0) dog = std::thread(watchdog, &S, cfg->watchdog_s > 0.0 ? cfg->watchdog_s : 30.0);
    size_t next = 0;
    {
        std::unique_lock<std::mutex> lk(S.mu);
        for (;;) {
            /* admit in arrival order while a slot and a pool block are free */
            while (next < order.size()) {
                const int k = order[next];
                if (jobs[k].arrival_s > now_s(S.t0)) break;
                if (res[k].footprint > cfg->hbm_budget) {
                    res[k].status = -2; /* can never fit */
                    S.done++;
                    next++;
                    continue;
                }
                int w = -1;
                for (int v = 0; v < cfg->slots; v++)
                    if (S.assigned[v] < 0) { w = v; break; }
                if (w < 0) break;
                const size_t off = pool.take(res[k].footprint);
                if (off == SIZE_MAX) break;
                S.block[k] = off;
                res[k].pool_offset = off;
                S.assigned[w] = k;
                S.mem_in_use += res[k].footprint;
                S.active++;
                if (S.active > res[k].max_active) res[k].max_active = S.active;
                res[k].slot = w;
                res[k].start_s = now_s(S.t0);
                next++;
                S.cv.notify_all();
            }
            if (S.done >= njobs) break;
            if (next < order.size()) {
                const double wait = jobs[order[next]].arrival_s - now_s(S.t0);
                if (wait > 0) S.cv.wait_for(lk, std::chrono::duration<double>(wait));
                else S.cv.wait_for(lk, std::chrono::milliseconds(5));
            } else {
                S.cv.wait(lk);
            }
        }
        S.stop = true;
    }
    S.cv.notify_all();
    for (auto &t : threads) t.join();
    if (dog.joinable()) dog.join();
    /* the precision check creates ordinary contexts: the pool goes first */
    if (S.pool_base) {
        (void)hipStreamSynchronize(S.streams[0]);
        (void)hipFree(S.pool_base);
        S.pool_base = nullptr;
    }
    if (S.cfg.check_fp64)
        for (int k = 0; k < njobs; k++) check_job(S, k);
    release();
    return 0;
}

extern "C" int gbserve_run(const gbserve_job *jobs, int njobs, const gbserve_config *cfg,
                           const char *mm_opts, gbserve_result *res)
{
    return serve_loop(jobs, njobs, cfg, mm_opts, res, nullptr, nullptr);
}

extern "C" int gbserve_run_custom(const gbserve_job *jobs, int njobs, const gbserve_config *cfg,
                                  gbserve_runner run, void *user, gbserve_result *res)
{
    if (!run) {
        nabc_seterr("gbserve_run_custom: NULL runner");
        return -1;
    }
    return serve_loop(jobs, njobs, cfg, nullptr, res, run, user);
}
