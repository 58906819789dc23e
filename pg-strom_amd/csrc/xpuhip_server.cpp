/*
 * xpuhip_server.cpp - the GPU service as a request loop (continuous batching).
 *
 * The reference's GPU service runs one XpuTaskExec per worker thread, one
 * kernel per chunk (gpu_service.c:4005-4290, :5912-5993).  Here requests from
 * many backends (a session and a chunk already in device memory) enter one
 * queue; a dispatcher admits them in arrival order as in-flight slots and
 * the HBM budget allow, coalesces queued group-by fragments of one plan shape
 * into a single batched launch (xpuhip_batch_exec), and hands the work to
 * worker threads whose tasks run on their own streams, so admitted requests
 * overlap on the device.  A request that arrives while others run is
 * admitted as soon as a slot and its memory free up.  The same loop with one
 * slot and no coalescing is the one-query-at-a-time service.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "xpuhip.h"

namespace {

typedef std::chrono::steady_clock Clock;

struct Req
{
    xpuhip_request  r;
    uint64_t        id = 0;
    Clock::time_point t_submit, t_admit, t_done;
    bool            done = false;
    int             rc = 0;
    uint32_t        batched = 0;
    kern_exec_results results;
    kern_errorbuf   errbuf;
};

struct Job
{
    std::vector<Req *> reqs;
};

double
ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}   /* namespace */

struct xpuhip_server
{
    std::mutex      m;
    std::condition_variable cv_queue;       /* dispatcher: new request / free slot */
    std::condition_variable cv_jobs;        /* workers: new job */
    std::condition_variable cv_done;        /* waiters */
    std::deque<Req *> queue;
    std::deque<Job *> jobs;
    std::map<uint64_t, Req *> reqs;
    uint64_t        next_id = 1;
    uint64_t        budget = 0;
    uint64_t        used = 0;
    uint32_t        slots = 1;
    uint32_t        inflight = 0;
    uint32_t        max_batch = 1;
    bool            stop = false;
    std::thread     dispatcher;
    std::vector<std::thread> workers;
    /* counters */
    uint64_t        launches = 0;
    uint64_t        coalesced = 0;
};

namespace {

void
run_job(xpuhip_server *srv, Job *job)
{
    if (job->reqs.size() == 1)
    {
        Req *q = job->reqs[0];
        uint64_t pos = 0;
        for (;;)
        {
            q->rc = xpuhip_task_exec_device(q->r.sess, q->r.d_kds_src, q->r.nitems,
                                            q->r.d_kds_dst, q->r.dst_length, &pos,
                                            &q->results, &q->errbuf);
            /* a projection that fills kds_dst restarts at the resume point
             * (the caller sized kds_dst; the served form keeps the last part) */
            if (q->rc != ERRCODE_SUSPEND_NO_SPACE)
                break;
        }
    }
    else
    {
        std::vector<xpuhip_batch_task> tasks(job->reqs.size());
        for (size_t i = 0; i < job->reqs.size(); i++)
        {
            Req *q = job->reqs[i];
            memset(&tasks[i], 0, sizeof(tasks[i]));
            tasks[i].sess = q->r.sess;
            tasks[i].d_kds_src = q->r.d_kds_src;
            tasks[i].nitems = q->r.nitems;
            tasks[i].results = &q->results;
        }
        kern_errorbuf eb;
        memset(&eb, 0, sizeof(eb));
        int rc = xpuhip_batch_exec(tasks.data(), (uint32_t)tasks.size(), 0, &eb);
        for (Req *q : job->reqs)
        {
            q->rc = rc;
            q->errbuf = eb;
            q->batched = (uint32_t)job->reqs.size();
        }
    }
    std::lock_guard<std::mutex> g(srv->m);
    Clock::time_point now = Clock::now();
    for (Req *q : job->reqs)
    {
        q->t_done = now;
        q->done = true;
        srv->used -= q->r.hbm_bytes;
    }
    srv->inflight--;
    srv->launches++;
    if (job->reqs.size() > 1)
        srv->coalesced += job->reqs.size();
    delete job;
    srv->cv_done.notify_all();
    srv->cv_queue.notify_all();
}

void
worker_loop(xpuhip_server *srv)
{
    for (;;)
    {
        Job *job;
        {
            std::unique_lock<std::mutex> lk(srv->m);
            srv->cv_jobs.wait(lk, [&] { return srv->stop || !srv->jobs.empty(); });
            if (srv->jobs.empty())
                return;
            job = srv->jobs.front();
            srv->jobs.pop_front();
        }
        run_job(srv, job);
    }
}

bool
batchable(const xpuhip_request &r)
{
    return (r.flags & XPUHIP_REQ_BATCHABLE) != 0;
}

void
dispatch_loop(xpuhip_server *srv)
{
    std::unique_lock<std::mutex> lk(srv->m);
    for (;;)
    {
        srv->cv_queue.wait(lk, [&] {
            if (srv->stop)
                return true;
            if (srv->queue.empty() || srv->inflight >= srv->slots)
                return false;
            Req *h = srv->queue.front();
            /* arrival order: the head waits for its memory (an empty GPU
             * always admits it, whatever its size) */
            return srv->inflight == 0 || srv->used + h->r.hbm_bytes <= srv->budget;
        });
        if (srv->stop)
            return;
        Job *job = new Job();
        Req *h = srv->queue.front();
        srv->queue.pop_front();
        job->reqs.push_back(h);
        srv->used += h->r.hbm_bytes;
        /* coalesce queued group-by fragments of the head's plan shape */
        if (batchable(h->r) && srv->max_batch > 1)
        {
            for (auto it = srv->queue.begin();
                 it != srv->queue.end() && job->reqs.size() < srv->max_batch; )
            {
                Req *q = *it;
                if (batchable(q->r) && q->r.shape == h->r.shape &&
                    srv->used + q->r.hbm_bytes <= srv->budget)
                {
                    job->reqs.push_back(q);
                    srv->used += q->r.hbm_bytes;
                    it = srv->queue.erase(it);
                }
                else
                    ++it;
            }
        }
        Clock::time_point now = Clock::now();
        for (Req *q : job->reqs)
            q->t_admit = now;
        srv->inflight++;
        srv->jobs.push_back(job);
        srv->cv_jobs.notify_one();
    }
}

}   /* namespace */

extern "C" xpuhip_server *
xpuhip_server_create(const xpuhip_server_config *cfg, kern_errorbuf *eb)
{
    if (!cfg || cfg->slots == 0)
    {
        if (eb)
        {
            memset(eb, 0, sizeof(*eb));
            eb->errcode = ERRCODE_DEVICE_ERROR;
            snprintf(eb->message, sizeof(eb->message), "server needs at least one slot");
        }
        return nullptr;
    }
    xpuhip_server *srv = new xpuhip_server();
    srv->budget = cfg->hbm_budget;
    srv->slots = cfg->slots;
    srv->max_batch = cfg->max_batch ? cfg->max_batch : 1;
    srv->dispatcher = std::thread(dispatch_loop, srv);
    for (uint32_t i = 0; i < cfg->slots; i++)
        srv->workers.emplace_back(worker_loop, srv);
    return srv;
}

extern "C" uint64_t
xpuhip_server_submit(xpuhip_server *srv, const xpuhip_request *req)
{
    Req *q = new Req();
    q->r = *req;
    memset(&q->results, 0, sizeof(q->results));
    memset(&q->errbuf, 0, sizeof(q->errbuf));
    std::lock_guard<std::mutex> g(srv->m);
    q->id = srv->next_id++;
    q->t_submit = Clock::now();
    srv->reqs[q->id] = q;
    srv->queue.push_back(q);
    srv->cv_queue.notify_one();
    return q->id;
}

extern "C" int
xpuhip_server_wait(xpuhip_server *srv, uint64_t id, xpuhip_request_status *st)
{
    std::unique_lock<std::mutex> lk(srv->m);
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end())
        return ERRCODE_DEVICE_ERROR;
    Req *q = it->second;
    srv->cv_done.wait(lk, [&] { return q->done; });
    if (st)
    {
        memset(st, 0, sizeof(*st));
        st->rc = q->rc;
        st->batched = q->batched;
        st->queue_ms = ms_between(q->t_submit, q->t_admit);
        st->exec_ms = ms_between(q->t_admit, q->t_done);
        st->total_ms = ms_between(q->t_submit, q->t_done);
        st->results = q->results;
        st->errbuf = q->errbuf;
    }
    srv->reqs.erase(it);
    delete q;
    return 0;
}

extern "C" void
xpuhip_server_stats(xpuhip_server *srv, uint64_t *launches, uint64_t *coalesced)
{
    std::lock_guard<std::mutex> g(srv->m);
    if (launches)
        *launches = srv->launches;
    if (coalesced)
        *coalesced = srv->coalesced;
}

extern "C" void
xpuhip_server_destroy(xpuhip_server *srv)
{
    if (!srv)
        return;
    {
        std::unique_lock<std::mutex> lk(srv->m);
        /* drain: every submitted request completes */
        srv->cv_done.wait(lk, [&] { return srv->queue.empty() && srv->inflight == 0; });
        srv->stop = true;
        srv->cv_queue.notify_all();
        srv->cv_jobs.notify_all();
    }
    srv->dispatcher.join();
    for (std::thread &t : srv->workers)
        t.join();
    for (auto &kv : srv->reqs)
        delete kv.second;
    delete srv;
}
