// Continuous-batching service of sector solves on one GPU.
//
// No reference counterpart: the reference runs sectors one after another inside
// _ed_scan_sectors! / ed_scan! (src/symmetry_resolved_ed.jl:1199-1356).  Here every
// admitted request runs on its own solver lane (host thread + HIP stream + reusable
// workspace), so the kernels of all admitted requests overlap on the device.  A queued
// request is admitted as soon as a lane is free and its device footprint fits the HBM
// budget; lanes pick up new work the moment they finish, so requests submitted while
// others run are admitted without waiting for the batch to drain.
//
// Memory accounting: the budget covers everything the server allocates.  The packed
// engine's arena and its fixed buffers are charged once, at creation.  Each lane is
// charged what its workspace holds while idle, and while running what it held before
// plus the request's footprint (a workspace buffer only grows to the size the request
// needs, so that sum bounds it).  When a request does not fit, idle lanes' kept
// workspaces are released first; a request larger than the whole lane share of the
// budget fails at once with SED_ENOMEM.  Catalogs and sectors belong to the caller and
// are not counted.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <memory>
#include <unordered_map>

#include "sed_internal.hpp"

namespace sed {
namespace {

using clk = std::chrono::steady_clock;
constexpr size_t kLaneKeepBytes = (size_t)1 << 30;   // idle lane workspace kept for reuse
constexpr int kPackMaxActive = 512;                   // sectors in flight in the packed engine

struct Request {
    int64_t id = 0;
    sed_sector* sec = nullptr;
    int nev = 1, maxiter = 300, precision = SED_PREC_FP64;
    double tol = 1e-10;
    uint64_t seed = 0;
    int status = 1;  // 1 pending, 2 running, 0 done, <0 error code
    std::string error;
    std::vector<double> evals;
    sed_solve_stats stats{};
    int64_t footprint = 0;
    clk::time_point t_submit, t_start, t_end;
};

struct Lane {
    std::thread th;
    Workspace ws;
    int64_t job = -1;   // request id assigned, -1 idle
    int64_t charge = 0; // bytes charged to the budget for this lane
};

}  // namespace
}  // namespace sed

struct sed_server {
    sed_context* ctx = nullptr;
    int64_t budget = 0;       // everything the server allocates
    int64_t fixed = 0;        // packed engine arena + buffers, charged at creation
    sed::PackEngine* pack = nullptr;  // small sectors: packed persistent launches
    std::unordered_map<int64_t, std::unique_ptr<sed::PackJob>> pack_jobs;
    std::mutex mu;
    std::condition_variable cv_lane, cv_done;
    std::deque<int64_t> queue;
    std::unordered_map<int64_t, sed::Request> reqs;
    std::vector<sed::Lane*> lanes;
    int64_t next_id = 1;
    int64_t in_use = 0;       // fixed + the lanes' charges
    int64_t peak = 0;         // largest in_use so far
    int64_t n_done = 0;
    int64_t n_submitted = 0;
    bool stop = false;
    bool counted = false;   // registered as a child of ctx
};

namespace sed {
namespace {

void set_in_use(sed_server* srv) {
    int64_t u = srv->fixed;
    for (const Lane* l : srv->lanes) u += l->charge;
    srv->in_use = u;
    srv->peak = std::max(srv->peak, u);
}

void release_lane(sed_server* srv, Lane* l) {
    (void)hipSetDevice(srv->ctx->device);
    l->ws.release();
    l->charge = 0;
    set_in_use(srv);
}

// Charge request r to the idle lane: what the lane keeps plus r's footprint.  When that
// does not fit, the lane's own kept buffers are released, then other idle lanes' (idle
// lanes have no work on their streams: a solve ends with a stream synchronisation, so
// their buffers can be freed from this thread).  False, with nothing released, if r
// cannot fit while the running requests hold their memory.
bool fit_locked(sed_server* srv, Lane* lane, const Request& r) {
    if (srv->in_use + r.footprint <= srv->budget) return true;
    int64_t idle = 0;
    for (const Lane* l : srv->lanes)
        if (l->job < 0) idle += l->charge;
    if (srv->in_use - idle + r.footprint > srv->budget) return false;
    if (lane->charge > 0) release_lane(srv, lane);
    for (Lane* l : srv->lanes) {
        if (srv->in_use + r.footprint <= srv->budget) break;
        if (l->job < 0 && l->charge > 0) release_lane(srv, l);
    }
    return srv->in_use + r.footprint <= srv->budget;
}

// Assign queued requests to idle lanes while they fit (caller holds the lock).
void admit_locked(sed_server* srv) {
    for (size_t li = 0; li < srv->lanes.size(); ++li) {
        Lane* lane = srv->lanes[li];
        if (lane->job >= 0) continue;
        for (auto it = srv->queue.begin(); it != srv->queue.end(); ++it) {
            Request& r = srv->reqs[*it];
            if (r.sec->busy) continue;   // one solve per sector at a time
            if (!fit_locked(srv, lane, r)) continue;
            r.status = 2;
            r.t_start = clk::now();
            r.sec->busy = 1;
            lane->charge += r.footprint;
            set_in_use(srv);
            lane->job = r.id;
            srv->queue.erase(it);
            break;
        }
    }
    srv->cv_lane.notify_all();
}

void lane_main(sed_server* srv, Lane* lane) {
    (void)hipSetDevice(srv->ctx->device);
    std::unique_lock<std::mutex> lk(srv->mu);
    while (true) {
        srv->cv_lane.wait(lk, [&] { return srv->stop || lane->job >= 0; });
        if (srv->stop) break;
        Request& r = srv->reqs[lane->job];
        sed_sector* sec = r.sec;
        const int nev = r.nev, maxiter = r.maxiter, prec = r.precision;
        const double tol = r.tol;
        const uint64_t seed = r.seed;
        lk.unlock();
        std::vector<double> ev;
        sed_solve_stats st{};
        const int rc = solve_sector(sec, nev, tol, seed, maxiter, prec, ev, nullptr, &st, lane->ws);
        const std::string err = rc == SED_OK ? std::string() : std::string(sed_last_error());
        // keep up to kLaneKeepBytes of workspace for the next request (charged while idle)
        if (lane->ws.device_bytes() > kLaneKeepBytes) lane->ws.release();
        lk.lock();
        Request& r2 = srv->reqs[lane->job];
        r2.evals = ev;
        r2.stats = st;
        r2.status = rc == SED_OK ? 0 : rc;
        r2.error = err;
        r2.t_end = clk::now();
        r2.sec->busy = 0;
        lane->charge = (int64_t)lane->ws.device_bytes();
        set_in_use(srv);
        ++srv->n_done;
        lane->job = -1;
        admit_locked(srv);
        srv->cv_done.notify_all();
    }
}

// completion of a packed request (called from the packed engine's thread)
void pack_done(void* user, PackJob* job) {
    auto* srv = static_cast<sed_server*>(user);
    std::lock_guard<std::mutex> lk(srv->mu);
    Request& r = srv->reqs[job->id];
    r.evals = job->evals;
    r.stats = job->stats;
    r.status = job->status;
    r.error = job->error;
    r.t_end = clk::now();
    r.sec->busy = 0;
    ++srv->n_done;
    admit_locked(srv);   // a queued request on the same sector may start now
    srv->cv_done.notify_all();
}

bool pack_disabled() {
    const char* e = std::getenv("SED_PACKED");
    return e && e[0] == '0';
}

}  // namespace
}  // namespace sed

using namespace sed;

extern "C" {

int sed_server_create(sed_context* ctx, int64_t hbm_budget_bytes, int max_slots, sed_server** out) {
    if (!ctx || !out) return fail(SED_EINVAL, "NULL argument");
    if (max_slots < 1 || max_slots > 64) return fail(SED_EINVAL, "max_slots must be in 1..64");
    *out = nullptr;
    SED_HIP(hipSetDevice(ctx->device));
    if (hbm_budget_bytes <= 0) {
        size_t fr = 0, tot = 0;
        SED_HIP(hipMemGetInfo(&fr, &tot));
        hbm_budget_bytes = (int64_t)(0.8 * (double)fr);
    }
    auto* srv = new sed_server();
    srv->ctx = ctx;
    srv->budget = hbm_budget_bytes;
    for (int i = 0; i < max_slots; ++i) {
        auto* lane = new Lane();
        hipError_t e = hipStreamCreateWithFlags(&lane->ws.stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete lane;
            sed_server_destroy(srv);
            return fail(SED_EDEVICE, std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
        }
        srv->lanes.push_back(lane);
    }
    for (Lane* lane : srv->lanes) lane->th = std::thread(lane_main, srv, lane);
    if (!pack_disabled()) {
        const int64_t arena = std::min<int64_t>(srv->budget / 4, (int64_t)16 << 30);
        srv->pack = pack_create(ctx, arena, kPackMaxActive, pack_done, srv);
        if (!srv->pack) {
            const std::string err = sed_last_error();
            sed_server_destroy(srv);
            return fail(SED_ENOMEM, "packed engine: " + err);
        }
        srv->fixed = pack_device_bytes(srv->pack);
    }
    srv->in_use = srv->peak = srv->fixed;
    *out = srv;
    context_child_added(ctx);
    srv->counted = true;
    return SED_OK;
}

int sed_server_submit(sed_server* srv, sed_sector* sec, int nev, double tol, uint64_t seed, int maxiter, int precision,
                      int64_t* request_id) {
    if (!srv || !sec || !request_id) return fail(SED_EINVAL, "NULL argument");
    if (nev < 1 || !(tol > 0)) return fail(SED_EINVAL, "bad nev / tol");
    if (precision != SED_PREC_FP64 && precision != SED_PREC_FP32 && precision != SED_PREC_FP32_REFINE)
        return fail(SED_EINVAL, "bad precision");
    if (!sec->has_terms) return fail(SED_ESTATE, "submit before sed_sector_set_terms");
    std::unique_lock<std::mutex> lk(srv->mu);
    Request r;
    r.id = srv->next_id++;
    r.sec = sec;
    r.nev = nev;
    r.tol = tol;
    r.seed = seed;
    r.maxiter = maxiter < 1 ? 300 : maxiter;
    r.precision = precision;
    r.t_submit = clk::now();
    ++srv->n_submitted;
    // Small sectors take a free lane when one is idle (lowest latency); when every lane is
    // busy they join the packed persistent launches at the next check point (throughput).
    // A sector already running (lane or packed) waits in the queue for its turn.
    bool lane_idle = false;
    if (srv->queue.empty())
        for (const Lane* l : srv->lanes) lane_idle = lane_idle || l->job < 0;
    if (srv->pack && !lane_idle && !sec->busy && pack_eligible(sec, nev)) {
        r.status = 2;
        r.t_start = r.t_submit;
        r.footprint = pack_footprint(sec, nev, precision);
        sec->busy = 1;
        auto job = std::make_unique<PackJob>();
        job->id = r.id;
        job->sec = sec;
        job->nev = nev;
        job->tol = tol;
        job->seed = seed;
        job->maxiter = r.maxiter;
        job->precision = precision;
        PackJob* jp = job.get();
        srv->pack_jobs[r.id] = std::move(job);
        srv->reqs[r.id] = r;
        *request_id = r.id;
        lk.unlock();
        pack_submit(srv->pack, jp);
        return SED_OK;
    }
    r.footprint = solve_footprint(sec, nev, precision);
    if (r.footprint > srv->budget - srv->fixed) {
        r.status = SED_ENOMEM;
        r.error = "request footprint " + std::to_string(r.footprint) + " B exceeds the server's HBM budget (" +
                  std::to_string(srv->budget - srv->fixed) + " B after the packed engine's arena)";
        r.t_start = r.t_end = r.t_submit;
        ++srv->n_done;
        srv->reqs[r.id] = r;
    } else {
        srv->reqs[r.id] = r;
        srv->queue.push_back(r.id);
        admit_locked(srv);
    }
    *request_id = r.id;
    return SED_OK;
}

int sed_server_run(sed_server* srv, int max_steps, int64_t* n_done) {
    if (!srv) return fail(SED_EINVAL, "NULL server");
    std::unique_lock<std::mutex> lk(srv->mu);
    admit_locked(srv);
    if (max_steps == 0) {
        srv->cv_done.wait(lk, [&] { return srv->n_done >= srv->n_submitted; });
    } else if (max_steps > 0) {
        const int64_t target = std::min<int64_t>(srv->n_done + max_steps, srv->n_submitted);
        srv->cv_done.wait(lk, [&] { return srv->n_done >= target; });
    }
    if (n_done) *n_done = srv->n_done;
    return SED_OK;
}

int sed_server_result(sed_server* srv, int64_t request_id, int* status, double* evals, sed_solve_stats* stats,
                      double* latency_s) {
    if (!srv || !status) return fail(SED_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(request_id);
    if (it == srv->reqs.end()) return fail(SED_EINVAL, "unknown request id");
    const Request& r = it->second;
    *status = r.status;
    if (r.status == 0) {
        if (evals)
            for (size_t i = 0; i < r.evals.size(); ++i) evals[i] = r.evals[i];
        if (stats) *stats = r.stats;
        if (latency_s) *latency_s = std::chrono::duration<double>(r.t_end - r.t_submit).count();
    } else if (r.status < 0) {
        set_error(r.error);
        if (latency_s) *latency_s = std::chrono::duration<double>(r.t_end - r.t_submit).count();
    }
    return SED_OK;
}

int sed_server_memory(sed_server* srv, int64_t* budget, int64_t* fixed, int64_t* in_use, int64_t* peak) {
    if (!srv) return fail(SED_EINVAL, "NULL server");
    std::lock_guard<std::mutex> lk(srv->mu);
    if (budget) *budget = srv->budget;
    if (fixed) *fixed = srv->fixed;
    if (in_use) *in_use = srv->in_use;
    if (peak) *peak = srv->peak;
    return SED_OK;
}

int sed_server_destroy(sed_server* srv) {
    if (!srv) return SED_OK;
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        // let running requests finish; drop the queue
        srv->queue.clear();
        srv->cv_done.wait(lk, [&] {
            for (Lane* l : srv->lanes)
                if (l->job >= 0) return false;
            return true;
        });
        srv->stop = true;
        srv->cv_lane.notify_all();
    }
    if (srv->pack) pack_destroy(srv->pack);   // finishes or cancels its requests
    srv->pack = nullptr;
    for (Lane* lane : srv->lanes) {
        if (lane->th.joinable()) lane->th.join();
        (void)hipSetDevice(srv->ctx->device);
        if (lane->ws.stream) (void)hipStreamDestroy(lane->ws.stream);
        lane->ws.release();
        delete lane;
    }
    sed_context* ctx = srv->ctx;
    const bool counted = srv->counted;
    delete srv;
    if (counted) context_child_removed(ctx);
    return SED_OK;
}

}  // extern "C"
