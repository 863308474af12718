// qdt_serve.cpp -- serving form: a request loop with continuous batching.
//
// Callers submit point sets at any time.  Worker threads, each owning an
// Engine (its own HIP stream and device workspace, kept resident between
// batches), take every queued request that fits the HBM budget the moment
// they are free and triangulate them together as one batch (Engine::
// run_batch).  Requests that arrive while a batch runs are picked up by the
// next free worker, so batches overlap on the device.  Each request's result
// is exactly what qdt_execute gives for it alone (the batch path keeps every
// set's triangulation independent).  No counterpart in the reference.
#include "../../include/qdt.h"
#include "../../include/qdt_gpu.h"
#include "engine.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

namespace
{

constexpr size_t kMaxPoints = 100000000;  // per request, libqdt/wrapper.hxx:15
constexpr size_t kMaxBatchPoints = 64u << 20;

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Pinned host memory holding one batch's triangles; the batch's requests
// point into it and the last one released returns it to the server's pool.
struct ResultBuf
{
    void* p = nullptr;
    size_t cap = 0;
    int refs = 0;
};

struct Request
{
    uint64_t id = 0;
    std::vector<qdt_vec2d> pts;
    ResultBuf* buf = nullptr;    // triangles: 3 * ntris uint32 at buf->p + tri_off
    size_t tri_off = 0, ntris = 0; // (request-local indices)
    size_t unique = 0;
    qdt_status status = QDT_OK;
    int state = 0; // 0 queued, 1 running, 2 done
    double t_submit = 0, t_start = 0, t_done = 0;
};

struct Worker
{
    qdt::Engine* eng = nullptr;
    void* d_in = nullptr;
    size_t d_in_cap = 0;
    void* h_in = nullptr; // pinned staging of a batch's points
    size_t h_in_cap = 0;
    std::thread th;
    // device bytes charged to this worker against the budget (guarded by the
    // server's mutex): while a batch runs, the larger of what it holds and
    // the batch's estimate; while idle, what its engine and input buffer
    // still hold (kept for the next batch)
    size_t held = 0;
    bool idle = true;
};

bool grow_pinned(void*& p, size_t& cap, size_t bytes)
{
    if (bytes <= cap) return true;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    const size_t c = std::max(bytes, cap * 2);
    if (hipHostMalloc(&p, c, hipHostMallocDefault) != hipSuccess) return false;
    cap = c;
    return true;
}

bool grow_device(void*& p, size_t& cap, size_t bytes)
{
    if (bytes <= cap) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const size_t c = std::max(bytes, cap * 2);
    if (hipMalloc(&p, c) != hipSuccess) return false;
    cap = c;
    return true;
}

} // namespace

struct qdt_gpu_server
{
    size_t budget = 0;      // device bytes all workers together may hold
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, Request*> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    std::vector<Worker*> workers;
    std::vector<ResultBuf*> pool; // free result buffers
    uint64_t batches = 0, batched_requests = 0;
    size_t peak_held = 0; // largest sum of the workers' charges so far
    void note_peak()      // with mu held
    {
        size_t t = 0;
        for (const Worker* v : workers) t += std::max(v->held, v->eng->workspace_bytes() + v->d_in_cap);
        peak_held = std::max(peak_held, t);
    }

    void work(Worker* w);
    size_t held_by_others(const Worker* w) const; // with mu held
    bool fits(Worker* w, size_t pts, size_t k);   // with mu held
    ResultBuf* get_buf(size_t bytes);   // with mu held
    void put_ref(ResultBuf* b);         // with mu held
    ~qdt_gpu_server();
};

ResultBuf* qdt_gpu_server::get_buf(size_t bytes)
{
    ResultBuf* best = nullptr;
    size_t at = 0;
    for (size_t i = 0; i < pool.size(); ++i)
        if (pool[i]->cap >= bytes && (!best || pool[i]->cap < best->cap))
        {
            best = pool[i];
            at = i;
        }
    if (best)
    {
        pool.erase(pool.begin() + (long)at);
        return best;
    }
    ResultBuf* b = new (std::nothrow) ResultBuf();
    if (!b) return nullptr;
    const size_t cap = std::max<size_t>(bytes, 1 << 20);
    if (hipHostMalloc(&b->p, cap, hipHostMallocDefault) != hipSuccess)
    {
        delete b;
        return nullptr;
    }
    b->cap = cap;
    return b;
}

void qdt_gpu_server::put_ref(ResultBuf* b)
{
    if (!b || --b->refs > 0) return;
    pool.push_back(b);
    // keep the pool small: free the smallest buffers beyond eight
    while (pool.size() > 8)
    {
        size_t lo = 0;
        for (size_t i = 1; i < pool.size(); ++i)
            if (pool[i]->cap < pool[lo]->cap) lo = i;
        (void)hipHostFree(pool[lo]->p);
        delete pool[lo];
        pool.erase(pool.begin() + (long)lo);
    }
}

size_t qdt_gpu_server::held_by_others(const Worker* w) const
{
    size_t t = 0;
    for (const Worker* v : workers)
        if (v != w) t += v->held;
    return t;
}

// Device bytes worker w holds after running a batch of `pts` points in `k`
// requests: its engine's regions at the larger of their current size and
// what the batch needs, and the input buffer.
size_t charge(const Worker* w, size_t pts, size_t k)
{
    return w->eng->bytes_after(pts, k) + std::max(w->d_in_cap, 16 * std::max<size_t>(pts, 1));
}

void drop_memory(Worker* v)
{
    v->eng->release();
    if (v->d_in) (void)hipFree(v->d_in);
    v->d_in = nullptr;
    v->d_in_cap = 0;
    v->held = 0;
}

// Whether worker w may run a batch of `pts` points in `k` requests within
// the budget, next to what the other workers hold.  Idle workers give their
// retained memory back when that is what stands in the way (their engines
// are not in use while they wait for work, and the mutex is held), and so
// does w itself when its retained regions are larger than the batch needs.
bool qdt_gpu_server::fits(Worker* w, size_t pts, size_t k)
{
    if (held_by_others(w) + charge(w, pts, k) <= budget) return true;
    for (Worker* v : workers)
    {
        if (v == w || !v->idle || !v->held) continue;
        drop_memory(v);
        if (held_by_others(w) + charge(w, pts, k) <= budget) return true;
    }
    if (w->held) drop_memory(w);
    return held_by_others(w) + charge(w, pts, k) <= budget;
}

void qdt_gpu_server::work(Worker* w)
{
    std::vector<Request*> take;
    std::vector<size_t> off;
    for (;;)
    {
        size_t need = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            w->idle = true;
            cv_work.wait(lk, [&] { return stop || !queue.empty(); });
            if (stop && queue.empty()) return;
            // admission: the queue's head requests while they fit the budget
            take.clear();
            size_t pts = 0;
            while (!queue.empty())
            {
                Request* r = queue.front();
                const size_t np = pts + r->pts.size();
                if (!take.empty() && (np > kMaxBatchPoints || !fits(w, np, take.size() + 1))) break;
                if (take.empty() && !fits(w, np, 1))
                {
                    if (held_by_others(w) == 0 && w->held == 0)
                    {
                        // alone it does not fit: reject it
                        queue.pop_front();
                        r->status = QDT_ERR_OUT_OF_MEMORY;
                        r->state = 2;
                        r->t_start = r->t_done = now_ms();
                        cv_done.notify_all();
                        continue;
                    }
                    break; // wait for memory to free up
                }
                queue.pop_front();
                r->state = 1;
                take.push_back(r);
                pts = np;
            }
            if (take.empty())
            {
                // the head does not fit next to the running batches: wait
                // until a batch ends
                cv_done.wait(lk, [&] { return stop || queue.empty() || fits(w, queue.front()->pts.size(), 1); });
                continue;
            }
            need = charge(w, pts, take.size());
            w->held = std::max(w->held, need);
            w->idle = false;
            note_peak();
            ++batches;
            batched_requests += take.size();
        }
        const double t0 = now_ms();
        for (Request* r : take) r->t_start = t0;
        // one batch: stage the points, run, split the triangles per request
        off.assign(take.size() + 1, 0);
        for (size_t i = 0; i < take.size(); ++i) off[i + 1] = off[i] + take[i]->pts.size();
        const size_t n = off.back();
        qdt_status st = QDT_OK;
        if (!grow_pinned(w->h_in, w->h_in_cap, 16 * std::max<size_t>(n, 1)) ||
            !grow_device(w->d_in, w->d_in_cap, 16 * std::max<size_t>(n, 1)))
            st = QDT_ERR_OUT_OF_MEMORY;
        if (st == QDT_OK)
        {
            for (size_t i = 0; i < take.size(); ++i)
                std::memcpy((qdt_vec2d*)w->h_in + off[i], take[i]->pts.data(), 16 * take[i]->pts.size());
            if (hipMemcpyAsync(w->d_in, w->h_in, 16 * n, hipMemcpyHostToDevice, w->eng->stream()) != hipSuccess)
                st = QDT_ERR_INTERNAL;
        }
        if (st == QDT_OK)
        {
            const qdt::Status s = w->eng->run_batch((const double*)w->d_in, off.data(), (int)take.size());
            if (s != qdt::Status::ok) st = s == qdt::Status::out_of_memory ? QDT_ERR_OUT_OF_MEMORY : QDT_ERR_INTERNAL;
        }
        const qdt::BatchResult& R = w->eng->batch();
        ResultBuf* buf = nullptr;
        if (st == QDT_OK && R.triangles)
        {
            {
                std::lock_guard<std::mutex> lk(mu);
                buf = get_buf(12 * (size_t)R.triangles);
            }
            if (!buf)
                st = QDT_ERR_OUT_OF_MEMORY;
            else if (hipMemcpyAsync(buf->p, R.tris, 12 * (size_t)R.triangles, hipMemcpyDeviceToHost,
                                    w->eng->stream()) != hipSuccess ||
                     hipStreamSynchronize(w->eng->stream()) != hipSuccess)
                st = QDT_ERR_INTERNAL;
        }
        for (Request* r : take)
        {
            r->pts.clear();
            r->pts.shrink_to_fit();
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            const double t1 = now_ms();
            for (size_t i = 0; i < take.size(); ++i)
            {
                Request* r = take[i];
                r->status = st != QDT_OK ? st : (qdt_status)R.status[i];
                if (st == QDT_OK)
                {
                    r->unique = R.unique[i];
                    r->ntris = R.tri_count[i];
                    r->tri_off = R.tri_start[i];
                    if (buf && r->ntris)
                    {
                        r->buf = buf;
                        ++buf->refs;
                    }
                }
                r->t_done = t1;
                r->state = 2;
            }
            if (buf && buf->refs == 0) pool.push_back(buf);
            // from now on the worker is charged what it really keeps
            note_peak();
            w->held = w->eng->workspace_bytes() + w->d_in_cap;
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

qdt_gpu_server::~qdt_gpu_server()
{
    {
        std::lock_guard<std::mutex> lk(mu);
        stop = true;
    }
    cv_work.notify_all();
    cv_done.notify_all();
    for (Worker* w : workers)
    {
        if (w->th.joinable()) w->th.join();
        delete w->eng;
        if (w->d_in) (void)hipFree(w->d_in);
        if (w->h_in) (void)hipHostFree(w->h_in);
        delete w;
    }
    for (auto& kv : reqs)
    {
        put_ref(kv.second->buf);
        delete kv.second;
    }
    for (ResultBuf* b : pool)
    {
        (void)hipHostFree(b->p);
        delete b;
    }
}

namespace
{

Request* find(qdt_gpu_server* S, uint64_t id)
{
    auto it = S->reqs.find(id);
    return it == S->reqs.end() ? nullptr : it->second;
}

} // namespace

extern "C" {

qdt_gpu_server* qdt_gpu_server_create(size_t hbm_budget_bytes, int workers)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return nullptr;
    qdt_gpu_server* S = new (std::nothrow) qdt_gpu_server();
    if (!S) return nullptr;
    S->budget = hbm_budget_bytes;
    const int nw = std::min(std::max(workers, 1), 8);
    try
    {
        for (int i = 0; i < nw; ++i)
        {
            Worker* w = new Worker();
            w->eng = new qdt::Engine();
            S->workers.push_back(w);
            if (!w->eng->stream())
            {
                delete S;
                return nullptr;
            }
        }
        for (Worker* w : S->workers) w->th = std::thread([S, w] { S->work(w); });
    }
    catch (...)
    {
        delete S;
        return nullptr;
    }
    return S;
}

void qdt_gpu_server_destroy(qdt_gpu_server* s) { delete s; }

uint64_t qdt_gpu_server_submit(qdt_gpu_server* S, const qdt_vec2d* pts, size_t n)
{
    if (!S || (n && !pts)) return 0;
    Request* r = new (std::nothrow) Request();
    if (!r) return 0;
    r->t_submit = now_ms();
    if (n > kMaxPoints)
    {
        r->status = QDT_ERR_TOO_MANY_POINTS;
        r->state = 2;
        r->t_start = r->t_done = r->t_submit;
    }
    else
    {
        try
        {
            r->pts.assign(pts, pts + n);
        }
        catch (...)
        {
            delete r;
            return 0;
        }
    }
    std::lock_guard<std::mutex> lk(S->mu);
    r->id = S->next_id++;
    S->reqs[r->id] = r;
    if (r->state == 0) S->queue.push_back(r);
    S->cv_work.notify_one();
    return r->id;
}

int qdt_gpu_server_poll(qdt_gpu_server* S, uint64_t id)
{
    if (!S) return -1;
    std::lock_guard<std::mutex> lk(S->mu);
    Request* r = find(S, id);
    return r ? r->state : -1;
}

qdt_status qdt_gpu_server_wait(qdt_gpu_server* S, uint64_t id, double* latency_ms)
{
    if (!S) return QDT_ERR_NULL_CONTEXT;
    std::unique_lock<std::mutex> lk(S->mu);
    if (!find(S, id)) return QDT_ERR_INVALID_INDEX;
    // the request is looked up again on every wake-up: another thread may
    // release it (and free it) once it is done
    S->cv_done.wait(lk, [&] {
        const Request* q = find(S, id);
        return !q || q->state == 2;
    });
    const Request* r = find(S, id);
    if (!r) return QDT_ERR_INVALID_INDEX;
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    return r->status;
}

int qdt_gpu_server_times(qdt_gpu_server* S, uint64_t id, double* submit_ms, double* start_ms, double* done_ms)
{
    if (!S) return 0;
    std::lock_guard<std::mutex> lk(S->mu);
    Request* r = find(S, id);
    if (!r || r->state != 2) return 0;
    if (submit_ms) *submit_ms = r->t_submit;
    if (start_ms) *start_ms = r->t_start;
    if (done_ms) *done_ms = r->t_done;
    return 1;
}

size_t qdt_gpu_server_triangles_size(qdt_gpu_server* S, uint64_t id)
{
    if (!S) return 0;
    std::lock_guard<std::mutex> lk(S->mu);
    Request* r = find(S, id);
    return (r && r->state == 2) ? r->ntris : 0;
}

size_t qdt_gpu_server_unique_points(qdt_gpu_server* S, uint64_t id)
{
    if (!S) return 0;
    std::lock_guard<std::mutex> lk(S->mu);
    Request* r = find(S, id);
    return (r && r->state == 2) ? r->unique : 0;
}

size_t qdt_gpu_server_copy_triangles(qdt_gpu_server* S, uint64_t id, uint32_t* out)
{
    if (!S || !out) return 0;
    std::lock_guard<std::mutex> lk(S->mu);
    Request* r = find(S, id);
    if (!r || r->state != 2) return 0;
    if (r->ntris) std::memcpy(out, (const uint32_t*)r->buf->p + 3 * r->tri_off, 12 * r->ntris);
    return r->ntris;
}

void qdt_gpu_server_release(qdt_gpu_server* S, uint64_t id)
{
    if (!S) return;
    std::lock_guard<std::mutex> lk(S->mu);
    Request* r = find(S, id);
    if (!r || r->state != 2) return;
    S->reqs.erase(id);
    S->put_ref(r->buf);
    delete r;
    S->cv_done.notify_all();
}

int qdt_gpu_server_memory(qdt_gpu_server* S, size_t* held, size_t* peak)
{
    if (!S) return 0;
    std::lock_guard<std::mutex> lk(S->mu);
    size_t t = 0;
    for (const Worker* v : S->workers) t += v->held;
    if (held) *held = t;
    if (peak) *peak = S->peak_held;
    return 1;
}

int qdt_gpu_server_counters(qdt_gpu_server* S, uint64_t* batches, uint64_t* requests)
{
    if (!S) return 0;
    std::lock_guard<std::mutex> lk(S->mu);
    if (batches) *batches = S->batches;
    if (requests) *requests = S->batched_requests;
    return 1;
}

} // extern "C"
