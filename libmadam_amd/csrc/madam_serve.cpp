// Request loop: destripe jobs of different sizes served on one GPU.
//
// The reference solves one job per destripe() call and, in Monte Carlo mode,
// keeps one configuration's set-up cached between calls (src/smadam.F90:440-448,
// 593-951).  Here many jobs are in flight together:
//   * admission by an HBM budget: a job is admitted when its estimated device
//     memory fits beside the jobs in flight and the resident scans kept for
//     reuse (the least recently used resident scans are evicted to make room);
//     jobs are admitted first-fit in arrival order, so a small job behind a
//     large one that does not fit yet goes first;
//   * device memory is checked as well as estimated: before a group starts,
//     the free HBM (hipMemGetInfo) less the estimates of the groups in flight
//     must hold the group's estimate and a reserve, or idle resident scans
//     are evicted first (the estimates can fall short of what a context holds
//     after a batched solve, and other users of the device are not counted);
//   * resident scans: a job whose scan key matches a kept context skips the
//     configuration and pointing set-up (pixel grouping, projection tables)
//     and, nna being bitwise the same, the filters and preconditioner;
//   * large jobs (an estimate above an eighth of the budget) take at most
//     half of the workers, so that a few long solves cannot hold every worker
//     while short jobs queue behind them;
//   * continuous batching: jobs waiting on the same scan with the same
//     settings are solved together as one batched PCG over their right-hand
//     sides (madam_solve_batch), and jobs of different scans run on their own
//     streams from their own worker threads, so their work overlaps on the
//     device; a job arriving while others run is admitted as soon as a worker
//     and memory are free, without waiting for a batch to drain.
// The jobs' buffers are read in place and must stay valid until the job is
// done; the pointing and the signal may be host or device memory (on_device).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "madam.h"

namespace {

double wall()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Job {
    long id = 0;
    // inputs (the caller's host buffers)
    std::string pars;
    long ndet = 0, nsamp = 0, nnz = 0, nperiod = 0, npsdtot = 0, npsdbin = 0, npsdval = 0;
    const double *detweights = nullptr, *timestamps = nullptr, *psdstarts = nullptr, *psdfreqs = nullptr,
                 *psdvals = nullptr;
    const long *periods = nullptr, *npsd = nullptr;
    const int32_t *pix = nullptr;
    const float *w = nullptr;
    const void *signal = nullptr;
    int tod_f32 = 0, precision = 0, on_device = 0;
    unsigned long long scan_key = 0;
    size_t est_bytes = 0;
    // state
    int state = 0;   // 0 pending, 1 running, 2 done, 3 failed
    std::string error;
    double t_submit = 0, t_start = 0, t_done = 0;
    int batch = 0, resident = 0;
    // results
    long npix = 0, nmap = 0, nbase = 0;
    double niter = 0;
    std::vector<double> map, binmap, aa;
    std::vector<int32_t> hits;
};

struct Resident {
    void *ctx = nullptr;
    unsigned long long key = 0;
    size_t bytes = 0;
    bool busy = false;
    double last_use = 0;
};

// device memory of one job's context: per sample the pointing copies, packed
// records, CSR and compact records and the TOD (~90 B); per map pixel the
// pixel matrices, maps and tables (~300 B); per baseline the PCG vectors,
// band factor and slot tables (~260 B, x R right-hand sides for the vectors);
// plus the FFT batch buffer and a fixed allowance
size_t estimate_bytes(const Job &j, long nside_max, long nbase_est, int nrhs)
{
    size_t n = (size_t)j.ndet * j.nsamp;
    size_t npix = 12ull * nside_max * nside_max;
    size_t nb = (size_t)j.ndet * nbase_est;
    size_t per_rhs = (size_t)(j.tod_f32 ? 4 : 8) * n + nb * 80 + npix * 3 * 8;
    return n * 90 + npix * 300 + nb * 260 + (size_t)nrhs * per_rhs + (256ull << 20);
}

long parse_long(const std::string &pars, const char *key, long dflt)
{
    size_t k = pars.find(key);
    while (k != std::string::npos) {
        bool start = k == 0 || pars[k - 1] == ';' || pars[k - 1] == ' ';
        size_t e = pars.find('=', k);
        if (start && e != std::string::npos && pars.substr(k, e - k).find_first_not_of(" ", strlen(key)) == std::string::npos)
            return atol(pars.c_str() + e + 1);
        k = pars.find(key, k + 1);
    }
    return dflt;
}

double parse_double(const std::string &pars, const char *key, double dflt)
{
    size_t k = pars.find(key);
    while (k != std::string::npos) {
        bool start = k == 0 || pars[k - 1] == ';' || pars[k - 1] == ' ';
        size_t e = pars.find('=', k);
        if (start && e != std::string::npos && pars.substr(k, e - k).find_first_not_of(" ", strlen(key)) == std::string::npos)
            return atof(pars.c_str() + e + 1);
        k = pars.find(key, k + 1);
    }
    return dflt;
}

struct Server {
    int device = 0;
    size_t budget = 0;
    int workers = 1;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::map<long, std::shared_ptr<Job>> jobs;
    std::list<std::shared_ptr<Job>> pending;   // arrival order
    std::list<Resident> resident;
    std::set<unsigned long long> busy_keys;   // scans with a group in flight (later jobs join the next batch)
    size_t inflight_bytes = 0;
    long next_id = 0;
    bool stop = false;
    std::vector<std::thread> threads;
    // statistics
    long n_done = 0, n_failed = 0, n_batches = 0, n_batched_jobs = 0, n_resident_hits = 0, n_evicted = 0;
    size_t peak_bytes = 0;
    int max_concurrent = 0, running = 0, running_large = 0;

    bool is_large(size_t est) const { return est > budget / 8; }
    int max_large() const { return std::max(1, workers / 2); }
    long n_mem_evicted = 0;
    size_t min_free = SIZE_MAX;

    // HBM left free beside the device's own runtime allocations
    static constexpr size_t RESERVE = 6ull << 30;

    size_t device_free()
    {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return SIZE_MAX;   // unknown: the estimate alone decides
        min_free = std::min(min_free, fr);
        return fr;
    }

    size_t resident_bytes() const
    {
        size_t b = 0;
        for (const Resident &r : resident) b += r.bytes;
        return b;
    }

    bool evict_one()
    {
        auto victim = resident.end();
        for (auto it = resident.begin(); it != resident.end(); ++it)
            if (!it->busy && (victim == resident.end() || it->last_use < victim->last_use)) victim = it;
        if (victim == resident.end()) return false;
        madam_destroy(victim->ctx);
        resident.erase(victim);
        n_evicted++;
        return true;
    }

    // evict idle resident scans (least recently used first) until `need` fits
    // the budget and, measured, the device
    bool make_room(size_t need)
    {
        for (;;) {
            size_t used = inflight_bytes + resident_bytes();
            if (used + need <= budget) break;
            if (!evict_one()) return false;
        }
        for (;;) {
            size_t fr = device_free();
            if (fr == SIZE_MAX || fr >= need + inflight_bytes + RESERVE) return true;
            if (!evict_one()) return running == 0;   // nothing runs that could free memory: try anyway
            n_mem_evicted++;
        }
    }

    // Called with mu held: the next group of pending jobs that can start.
    std::vector<std::shared_ptr<Job>> take_group(Resident **res_out)
    {
        *res_out = nullptr;
        for (auto it = pending.begin(); it != pending.end(); ++it) {
            std::shared_ptr<Job> head = *it;
            // a scan with a group in flight serves this job in its next batch
            if (head->scan_key && busy_keys.count(head->scan_key)) continue;
            Resident *res = nullptr;
            if (head->scan_key)
                for (Resident &r : resident)
                    if (r.key == head->scan_key) res = &r;
            // followers: same scan, same settings, while the batch has room
            std::vector<std::shared_ptr<Job>> grp{head};
            if (head->scan_key)
                for (auto jt = std::next(it); jt != pending.end() && grp.size() < 64; ++jt) {
                    const Job &o = **jt;
                    if (o.scan_key == head->scan_key && o.precision == head->precision &&
                        o.tod_f32 == head->tod_f32 && o.on_device == head->on_device && o.pars == head->pars)
                        grp.push_back(*jt);
                }
            size_t need = 0;
            for (auto &j : grp) need += j->est_bytes;
            if (is_large(head->est_bytes) && running_large >= max_large()) continue;
            if (res) need = need > res->bytes ? need - res->bytes : 0;   // the set-up is already resident
            if (res) res->busy = true;   // not a candidate for eviction while room is made
            if (!make_room(need)) {
                if (res) res->busy = false;
                continue;
            }
            for (auto &j : grp) pending.remove(j);
            for (auto &j : grp) j->state = 1;
            if (is_large(head->est_bytes)) running_large++;
            inflight_bytes += need;
            if (head->scan_key) busy_keys.insert(head->scan_key);
            *res_out = res;
            size_t used = inflight_bytes + resident_bytes();
            peak_bytes = std::max(peak_bytes, used);
            return grp;
        }
        return {};
    }

    // a worker's own copy stream and staging buffer for the jobs' signals
    struct Lane {
        hipStream_t stream = nullptr;
        void *stage = nullptr;
        size_t cap = 0;
    };
    static constexpr size_t KEEP_STAGE = 256ull << 20;   // larger staging buffers are freed after the group

    void run_group(std::vector<std::shared_ptr<Job>> &grp, Resident *res, Lane &lane);
    void worker();
};

void Server::run_group(std::vector<std::shared_ptr<Job>> &grp, Resident *res, Lane &lane)
{
    Job &h = *grp[0];
    double t0 = wall();
    for (auto &j : grp) {
        j->t_start = t0;
        j->batch = (int)grp.size();
        j->resident = res ? 1 : 0;
    }
    void *ctx = res ? res->ctx : nullptr;
    std::string err;
    auto check = [&](int rc) {
        if (rc != 0) throw std::runtime_error(madam_last_error());
    };
    try {
        if (!ctx) {
            ctx = madam_create(device);
            if (!ctx) throw std::runtime_error(madam_last_error());
            check(madam_configure(ctx, h.pars.c_str(), h.ndet, "", h.detweights, h.nsamp, h.nnz, h.timestamps,
                                  h.nperiod, h.periods, h.npsd, h.npsdtot, h.psdstarts, h.npsdbin, h.psdfreqs,
                                  h.npsdval, h.psdvals, 0));
            // the context keeps its own device copies of the pointing
            check(madam_set_pointing(ctx, (int32_t *)h.pix, (float *)h.w, h.on_device));
        }
        long sz[6];
        check(madam_get_sizes(ctx, sz));
        long n = h.ndet * h.nsamp;
        int R = (int)grp.size();
        bool single64 = R == 1 && h.precision == 0 && !h.tod_f32;
        size_t es = h.tod_f32 ? 4 : 8;
        if (single64 && !h.on_device) {
            // the FP64 path of destripe() (it reads a host signal and leaves it unchanged here)
            std::vector<double> tod((const double *)h.signal, (const double *)h.signal + n);
            check(madam_solve(ctx, tod.data(), 0));
        } else if (R == 1 && h.on_device && !single64) {
            // one device signal: the batched solve reads it in place
            check(madam_solve_batch(ctx, 1, h.signal, h.tod_f32, 1, h.precision));
        } else {
            auto fail = [&](const char *what, hipError_t e) {
                size_t fr = 0, tot = 0;
                (void)hipMemGetInfo(&fr, &tot);
                throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e) + " (" +
                                         std::to_string(es * n) + " bytes per job, " + std::to_string(fr >> 20) +
                                         " MiB of device memory free)");
            };
            size_t bytes = es * n * R;
            hipError_t e = hipSuccess;
            if (!lane.stream) {
                e = hipStreamCreateWithFlags(&lane.stream, hipStreamNonBlocking);
                if (e != hipSuccess) fail("stream creation failed", e);
            }
            if (bytes > lane.cap) {
                if (lane.stage) (void)hipFree(lane.stage);
                lane.stage = nullptr;
                lane.cap = 0;
                e = hipMalloc(&lane.stage, bytes);
                if (e != hipSuccess) fail("staging allocation failed", e);
                lane.cap = bytes;
            }
            // copies on the worker's own stream: the legacy stream would have
            // to wait for every blocking stream, which fails while another
            // worker captures its PCG step into a graph
            hipMemcpyKind kind = h.on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
            for (int k = 0; k < R && e == hipSuccess; ++k)
                e = hipMemcpyAsync((char *)lane.stage + es * n * k, grp[k]->signal, es * n, kind, lane.stream);
            if (e == hipSuccess) e = hipStreamSynchronize(lane.stream);
            if (e != hipSuccess) fail("staging copy failed", e);
            // (the FP64 path works on a copy of the signal: it may write it)
            if (single64) check(madam_solve(ctx, (double *)lane.stage, 1));
            else check(madam_solve_batch(ctx, R, lane.stage, h.tod_f32, 1, h.precision));
        }
        check(madam_get_sizes(ctx, sz));
        for (int k = 0; k < R; ++k) {
            Job &j = *grp[k];
            j.nmap = sz[0];
            j.npix = sz[1];
            j.nbase = sz[2];
            j.map.assign((size_t)j.npix * j.nmap, 0.0);
            j.binmap.assign((size_t)j.npix * j.nmap, 0.0);
            j.hits.assign(j.npix, 0);
            j.aa.assign((size_t)j.ndet * j.nbase, 0.0);
            if (single64) {
                check(madam_get_maps(ctx, j.map.data(), j.binmap.data(), j.hits.data(), nullptr, nullptr, nullptr));
                check(madam_get_baselines(ctx, j.aa.data(), nullptr, nullptr));
                double st[1];
                check(madam_get_stats(ctx, st, 1));
                j.niter = st[0];
            } else {
                check(madam_get_batch_maps(ctx, k, j.map.data(), j.binmap.data()));
                check(madam_get_batch_baselines(ctx, k, j.aa.data(), nullptr));
                check(madam_get_hits(ctx, j.hits.data()));
                double info[8];
                check(madam_get_batch_info(ctx, k, info, 8));
                j.niter = info[7];
            }
        }
    } catch (const std::exception &e) {
        err = e.what();
    }
    if (lane.cap > KEEP_STAGE) {
        (void)hipFree(lane.stage);
        lane.stage = nullptr;
        lane.cap = 0;
    }
    std::lock_guard<std::mutex> lk(mu);
    double t1 = wall();
    size_t used = 0;
    for (auto &j : grp) used += j->est_bytes;
    if (res) used = used > res->bytes ? used - res->bytes : 0;
    inflight_bytes -= std::min(inflight_bytes, used);
    for (auto &j : grp) {
        j->t_done = t1;
        if (err.empty()) {
            j->state = 2;
            n_done++;
        } else {
            j->state = 3;
            j->error = err;
            n_failed++;
        }
    }
    if (grp.size() > 1) {
        n_batches++;
        n_batched_jobs += (long)grp.size();
    }
    if (res) {
        res->busy = false;
        res->last_use = t1;
        n_resident_hits += (long)grp.size();
        if (!err.empty()) {   // a failed solve leaves the context in an unknown state
            madam_destroy(res->ctx);
            for (auto it = resident.begin(); it != resident.end(); ++it)
                if (&*it == res) { resident.erase(it); break; }
        }
    } else if (ctx) {
        if (err.empty() && h.scan_key) {
            // keep the set-up for later jobs on this scan
            Resident r;
            r.ctx = ctx;
            r.key = h.scan_key;
            r.bytes = h.est_bytes;
            r.last_use = t1;
            resident.push_back(r);
            make_room(0);
        } else {
            madam_destroy(ctx);
        }
    }
    if (h.scan_key) busy_keys.erase(h.scan_key);
    running--;
    if (is_large(h.est_bytes)) running_large--;
    cv_done.notify_all();
    cv_work.notify_all();
}

void Server::worker()
{
    (void)hipSetDevice(device);
    Lane lane;
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        Resident *res = nullptr;
        std::vector<std::shared_ptr<Job>> grp;
        cv_work.wait(lk, [&] {
            if (stop) return true;
            grp = take_group(&res);
            return !grp.empty();
        });
        if (stop && grp.empty()) {
            if (lane.stage) (void)hipFree(lane.stage);
            if (lane.stream) (void)hipStreamDestroy(lane.stream);
            return;
        }
        running++;
        max_concurrent = std::max(max_concurrent, running);
        lk.unlock();
        run_group(grp, res, lane);
        lk.lock();
    }
}

}  // namespace

#define MADAM_API extern "C" __attribute__((visibility("default")))

MADAM_API void *madam_server_create(int device, double hbm_budget_bytes, int max_inflight)
{
    Server *s = new Server();
    s->device = device;
    s->budget = (size_t)hbm_budget_bytes;
    s->workers = std::max(1, max_inflight);
    for (int i = 0; i < s->workers; ++i) s->threads.emplace_back([s] { s->worker(); });
    return s;
}

MADAM_API void madam_server_destroy(void *srv)
{
    Server *s = (Server *)srv;
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->pending.clear();
    }
    s->cv_work.notify_all();
    for (auto &t : s->threads) t.join();
    for (Resident &r : s->resident) madam_destroy(r.ctx);
    delete s;
}

MADAM_API long madam_server_submit(void *srv, const char *parstring, long ndet, const double *detweights, long nsamp,
                                   long nnz, const double *timestamps, const int32_t *pix, const float *pixweights,
                                   const void *signal, int tod_f32, long nperiod, const long *periods,
                                   const long *npsd, long npsdtot, const double *psdstarts, long npsdbin,
                                   const double *psdfreqs, long npsdval, const double *psdvals,
                                   unsigned long long scan_key, int precision, int on_device)
{
    Server *s = (Server *)srv;
    auto j = std::make_shared<Job>();
    j->pars = parstring ? parstring : "";
    j->ndet = ndet;
    j->nsamp = nsamp;
    j->nnz = nnz;
    j->detweights = detweights;
    j->timestamps = timestamps;
    j->pix = pix;
    j->w = pixweights;
    j->signal = signal;
    j->tod_f32 = tod_f32;
    j->nperiod = nperiod;
    j->periods = periods;
    j->npsd = npsd;
    j->npsdtot = npsdtot;
    j->psdstarts = psdstarts;
    j->npsdbin = npsdbin;
    j->psdfreqs = psdfreqs;
    j->npsdval = npsdval;
    j->psdvals = psdvals;
    j->scan_key = scan_key;
    j->precision = precision;
    j->on_device = on_device;
    long nside = std::max(parse_long(j->pars, "nside_map", 512), parse_long(j->pars, "nside_cross", 0));
    double base = parse_double(j->pars, "base_first", 1.0) * parse_double(j->pars, "fsample", 1.0);
    long nbase = (long)(nsamp / std::max(1.0, base)) + nperiod + 1;
    j->est_bytes = estimate_bytes(*j, nside, nbase, 1);
    std::lock_guard<std::mutex> lk(s->mu);
    if (j->est_bytes > s->budget) return -1;
    j->id = s->next_id++;
    j->t_submit = wall();
    s->jobs[j->id] = j;
    s->pending.push_back(j);
    s->cv_work.notify_one();
    return j->id;
}

// 0 done, 1 failed (message from madam_last_error is not set; see madam_server_error), 2 still running at timeout
MADAM_API int madam_server_wait(void *srv, long id, double timeout_s)
{
    Server *s = (Server *)srv;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return 1;
    std::shared_ptr<Job> j = it->second;
    auto ready = [&] { return j->state >= 2; };
    if (timeout_s < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) return 2;
    return j->state == 2 ? 0 : 1;
}

MADAM_API const char *madam_server_error(void *srv, long id)
{
    Server *s = (Server *)srv;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    return it == s->jobs.end() ? "unknown request" : it->second->error.c_str();
}

// sizes6 = {npix, nmap, nbase, ndet, batch size, resident}; times4 = {submit,
// start, done, PCG steps}; NULL buffers are skipped
MADAM_API int madam_server_result(void *srv, long id, double *map, double *binmap, int32_t *hits, double *aa,
                                  long *sizes6, double *times4)
{
    Server *s = (Server *)srv;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end() || it->second->state != 2) return 1;
    Job &j = *it->second;
    if (map) std::copy(j.map.begin(), j.map.end(), map);
    if (binmap) std::copy(j.binmap.begin(), j.binmap.end(), binmap);
    if (hits) std::copy(j.hits.begin(), j.hits.end(), hits);
    if (aa) std::copy(j.aa.begin(), j.aa.end(), aa);
    if (sizes6) {
        sizes6[0] = j.npix;
        sizes6[1] = j.nmap;
        sizes6[2] = j.nbase;
        sizes6[3] = j.ndet;
        sizes6[4] = j.batch;
        sizes6[5] = j.resident;
    }
    if (times4) {
        times4[0] = j.t_submit;
        times4[1] = j.t_start;
        times4[2] = j.t_done;
        times4[3] = j.niter;
    }
    return 0;
}

MADAM_API int madam_server_release(void *srv, long id)
{
    Server *s = (Server *)srv;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end() || it->second->state < 2) return 1;
    s->jobs.erase(it);
    return 0;
}

// out = {done, failed, batches, jobs in batches, resident-scan jobs, evictions,
// peak estimated bytes, most jobs in flight at once, resident scans now,
// evictions forced by the measured free memory, least free device memory seen}
MADAM_API int madam_server_stats(void *srv, double *out, int n)
{
    Server *s = (Server *)srv;
    std::lock_guard<std::mutex> lk(s->mu);
    double v[11] = {(double)s->n_done, (double)s->n_failed, (double)s->n_batches, (double)s->n_batched_jobs,
                    (double)s->n_resident_hits, (double)s->n_evicted, (double)s->peak_bytes,
                    (double)s->max_concurrent, (double)s->resident.size(), (double)s->n_mem_evicted,
                    s->min_free == SIZE_MAX ? -1.0 : (double)s->min_free};
    for (int i = 0; i < n && i < 11; ++i) out[i] = v[i];
    return 0;
}
