/*
 * serve.cpp -- continuous-batching request loop for minimisation and
 * dynamics jobs on one GPU.
 *
 * Jobs arrive at given times (a Poisson stream in the benchmark).  A job is
 * admitted as soon as it has arrived, one of `slots` workers is free and its
 * device footprint (gbhip_footprint_bytes) fits in what the budget has left;
 * otherwise it waits and is admitted the moment a running job releases its
 * slot and memory.  Each worker owns a HIP stream (default priority), so
 * the device work of all admitted jobs overlaps.  A job keeps its whole
 * state on the device (coordinates, pair list, velocities, MD graphs):
 *   - CG: conjgrad() (conjgrad.c:34 restated in sff_host.c) over a device
 *     mme() (bonded + 1-4 + GB), pair list rebuilt every nsnb calls;
 *   - MD: gbhip_md_init + gbhip_md_run, every step one hipGraph launch.
 * Latency = completion - arrival (queueing included).  Optionally the final
 * coordinates are re-evaluated in FP64 to report the precision error.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "gbhip.h"
#include "host_internal.h"
#include "sff_gb.h"

#include "gbserve.h"

namespace {

struct JobState {
    gbhip_ctx *ctx = nullptr;
    int nsnb = 10;
    bool have_list = false;
    double ene[14];
    int fail = 0;
};

thread_local JobState *tl_job = nullptr;

/* mme() for the job running on this thread (conjgrad needs a plain function) */
double job_mme(double *x, double *f, int *iter)
{
    JobState *J = tl_job;
    if (*iter == -3) return 0.0;
    if (gbhip_set_coords(J->ctx, x) != 0) { J->fail = 1; return 1e300; }
    if (!J->have_list || (*iter > 0 && *iter % J->nsnb == 0) || *iter == -4) {
        if (gbhip_nblist(J->ctx, nullptr) != 0) { J->fail = 1; return 1e300; }
        J->have_list = true;
    }
    if (gbhip_mme(J->ctx, f, J->ene) != 0) { J->fail = 1; return 1e300; }
    return J->ene[0];
}

double now_s(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
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
    std::vector<hipStream_t> streams;
    std::vector<std::vector<double>> xfinal; /* per job, kept for the precision check */
    gbserve_runner runner = nullptr; /* caller's job function (gbserve_run_custom) */
    void *user = nullptr;
    int done = 0;
    bool stop = false;
    size_t mem_in_use = 0;
    int active = 0;
};

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
    std::vector<double> x(jb.x, jb.x + 3 * (size_t)n), f(3 * (size_t)n, 0.0);
    r.status = -100;
    if (nabc_model_build(&m, prm, nullptr, &o) != 0) { r.status = -10; return; }
    if (nabc_model_context(&m, x.data(), &o, S.streams[w], &ctx) != 0) {
        nabc_model_free(&m);
        r.status = -11;
        return;
    }
    JobState J;
    J.ctx = ctx;
    J.nsnb = o.nsnb > 0 ? o.nsnb : 1;
    if (jb.kind == GBSERVE_CG) {
        tl_job = &J;
        int nn = 3 * n, maxiter = jb.nsteps;
        double fret = 0.0, rms = S.cfg.cg_rmsgrad, dfpred = S.cfg.cg_dfpred;
        int ret = conjgrad(x.data(), &nn, &fret, job_mme, &rms, &dfpred, &maxiter);
        tl_job = nullptr;
        r.status = J.fail ? -12 : ret;
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
        mo.nsnb = J.nsnb;
        mo.seed = jb.seed;
        std::vector<double> row(16 * (size_t)std::max(1, jb.nsteps));
        int nr = 0;
        if (gbhip_md_init(ctx, prm->Masses, x.data(), nullptr, &mo, std::max(1, jb.nsteps)) != 0 ||
            gbhip_md_run(ctx, jb.nsteps, 1) != 0 || gbhip_md_state(ctx, x.data(), nullptr, nullptr) != 0 ||
            gbhip_md_log(ctx, row.data(), jb.nsteps, &nr) != 0 || nr < 1) {
            r.status = -13;
        } else {
            r.status = 0;
            r.energy = row[16 * (size_t)(nr - 1)];
        }
    }
    r.end_s = now_s(S.t0);
    r.latency_s = r.end_s - jb.arrival_s;
    gbhip_ctx_destroy(ctx);
    nabc_model_free(&m);
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
    if (r.status < -3 || x.empty()) return;
    std::vector<double> f(x.size());
    nabc_model m;
    if (nabc_model_build(&m, prm, nullptr, &S.opts) != 0) return;
    for (int prec = 0; prec < 2; prec++) {
        nabc_opts_t o = S.opts;
        o.gbprec = prec == 0 ? 0 : (S.cfg.precision == GBHIP_MIXED ? 1 : 0);
        gbhip_ctx *c = nullptr;
        if (nabc_model_context(&m, x.data(), &o, S.streams[0], &c) != 0) continue;
        double e[14];
        if (gbhip_set_coords(c, x.data()) == 0 && gbhip_nblist_async(c) == 0 && gbhip_mme(c, f.data(), e) == 0) {
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
            r.status = S->runner(&jb, w, S->user, &energy);
            r.energy = energy;
            r.end_s = now_s(S->t0);
            r.latency_s = r.end_s - jb.arrival_s;
        } else {
            run_job(*S, w, k);
        }
        int done;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->assigned[w] = -1;
            S->mem_in_use -= S->res[k].footprint;
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

} // namespace

extern "C" size_t gbhip_footprint_bytes(int natom, int md_steps)
{
    const size_t NS = ((size_t)natom + TILE_ATOMS - 1) / TILE_ATOMS * TILE_ATOMS, NT = NS / TILE_ATOMS;
    const size_t cap = NT * (NT + 1) / 2;
    size_t b = 0;
    b += NS * (6 * 4 + 7 * 8 + 4 * 4 + 11 * 8);          /* per-slot constants and work arrays */
    b += NT * (4 * 4 + 8 + 2 * 24);                       /* tile arrays */
    b += cap * (4 + 8 + 8 + 4 + 8 + 4 + 2 * TILE_ATOMS * 8); /* candidates, tile pairs (both orders, flags,
                                                              column sets, queue), bit matrices */
    /* pass-3 derivative store: at most every tile pair full, 16 bytes per
     * entry (FP64), 1.25x headroom */
    b += (cap * TILE_ATOMS * TILE_ATOMS + TILE_ATOMS * TILE_ATOMS) * 20;
    b += (size_t)natom * (4 + 4 + 12 + 2 * 24 + 8);       /* inverse order, exclusions, x, g, scratch */
    b += (size_t)natom * 220;                              /* bonded / 1-4 tables (chain topology bound) */
    b += (size_t)natom * 4 * 24 + 16 * 8 * (size_t)(md_steps > 0 ? md_steps : 0); /* dynamics state + log */
    b += 3 * 64 * 8;                                       /* energy partial sums */
    b += 64 * 4096;                                        /* allocation granularity of ~60 buffers */
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
    for (int k = 0; k < njobs; k++) {
        order[k] = k;
        std::memset(&res[k], 0, sizeof(res[k]));
        const PARMSTRUCT_T *p = (const PARMSTRUCT_T *)jobs[k].prm;
        res[k].footprint = gbhip_footprint_bytes(p->Natom, jobs[k].kind == GBSERVE_MD ? jobs[k].nsteps : 0);
        res[k].slot = -1;
        res[k].status = -100;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return jobs[a].arrival_s < jobs[b].arrival_s; });
    S.assigned.assign(cfg->slots, -1);
    S.xfinal.assign(njobs, std::vector<double>());
    S.streams.assign(cfg->slots, nullptr);
    for (int w = 0; w < cfg->slots && !runner; w++) {
        if (hipStreamCreateWithFlags(&S.streams[w], hipStreamNonBlocking) != hipSuccess) {
            nabc_seterr("gbserve_run: stream creation failed");
            for (int v = 0; v < w; v++) (void)hipStreamDestroy(S.streams[v]);
            return -1;
        }
    }
    std::vector<std::thread> threads;
    S.t0 = std::chrono::steady_clock::now();
    for (int w = 0; w < cfg->slots; w++) threads.emplace_back(worker, &S, w);
    size_t next = 0;
    int rejected = 0;
    {
        std::unique_lock<std::mutex> lk(S.mu);
        for (;;) {
            /* admit in arrival order while a slot and memory are free */
            while (next < order.size()) {
                const int k = order[next];
                if (jobs[k].arrival_s > now_s(S.t0)) break;
                if (res[k].footprint > cfg->hbm_budget) {
                    res[k].status = -2; /* can never fit */
                    rejected++;
                    S.done++;
                    next++;
                    continue;
                }
                int w = -1;
                for (int v = 0; v < cfg->slots; v++)
                    if (S.assigned[v] < 0) { w = v; break; }
                if (w < 0 || S.mem_in_use + res[k].footprint > cfg->hbm_budget) break;
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
    if (S.cfg.check_fp64)
        for (int k = 0; k < njobs; k++) check_job(S, k);
    for (hipStream_t s : S.streams)
        if (s) (void)hipStreamDestroy(s);
    (void)rejected;
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
