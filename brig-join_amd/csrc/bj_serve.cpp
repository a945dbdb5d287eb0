// bj_serve.cpp -- the serving form: a request loop over device-resident relations.
//
// Relations are partitioned once (bj_server_add_relation) and kept with their
// filter.  A request (one S against one relation) runs the S side of the join
// and the build + probe (enqueue_s_probe, the same stages bj_join runs) on one
// of the server's slots, each a context with its own stream.  Its workspace and
// rows are carved from one device arena, so admission is by footprint: the
// scheduler thread admits queued requests in arrival order whenever a slot is
// free and the arena has room, and retires them when their completion event
// fires.  No hipMalloc / hipFree happens while requests run.  The reference
// has no serving form.
#include "bj_api.h"
#include "bj_arena.h"
#include "bj_internal.h"
#include "bj_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

using namespace bj;

namespace {

constexpr int SRV_MAX_SLOTS = 8;
constexpr uint64_t ALIGN = 256;

uint64_t align_up(uint64_t x) { return (x + ALIGN - 1) / ALIGN * ALIGN; }

struct Relation {
    uint64_t nr = 0;
    Plan plan;  // plan of (R, R): radix bits and filter geometry of the relation
    DevBuf key, val, start, filter;
};

using Item = std::pair<DevBuf bj_ctx::*, uint64_t>;

struct Request {
    uint64_t id = 0;
    int rel = 0;
    const Relation* relp = nullptr;  // relations are heap objects that stay put
    const uint32_t* sk = nullptr;
    const uint32_t* sv = nullptr;
    Plan pl;
    std::vector<Item> items;  // S-side workspace, laid out back to back
    uint64_t ws_bytes = 0, out_bytes = 0, cap = 1;
    uint64_t ws_off = 0, out_off = 0;
    bool done = false, reported = false;
    bj_request_result res{};
};

struct Slot {
    bj_ctx* c = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr, done = nullptr;
    uint64_t* h_ctr = nullptr;  // pinned
    Request* cur = nullptr;
};

// The S-side buffers a request needs (no R side, no rows).
int request_items(const Plan& pl, std::vector<Item>& items, uint64_t& bytes)
{
    items.clear();
    bytes = 0;
    return workspace_items(pl, false, false, [&](DevBuf bj_ctx::*m, uint64_t b) -> int {
        items.emplace_back(m, align_up(b));
        bytes += align_up(b);
        return BJ_OK;
    });
}

}  // namespace

struct bj_server {
    int device = 0;
    bj_server_opts opts{};
    bj_ctx* main = nullptr;  // partitions relations
    std::vector<Relation*> rels;
    uint64_t rel_bytes = 0;
    DevBuf arena;
    ArenaAlloc alloc;
    std::vector<Slot> slots;
    // guards rels, the queue, reqs and the counters below (rels only grows; its
    // elements are heap objects, so a pointer taken under the lock stays valid)
    mutable std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::map<uint64_t, Request*> reqs;
    uint64_t next_id = 1;
    uint64_t submitted = 0, completed = 0, released = 0;
    int in_flight = 0, max_in_flight = 0;
    bool stop = false;
    bool blocked = false;  // the queue's head waits for rows to be released
    std::thread sched;
    std::chrono::steady_clock::time_point t_origin;

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t_origin).count();
    }
};

namespace {

int make_request_plan(const bj_server* s, const Relation& R, uint64_t ns, Plan& pl)
{
    bj_params p;
    bj_params_default(&p);
    p.r_size = (int64_t)R.nr;
    p.s_size = (int64_t)ns;
    p.radix_bits = R.plan.bits;
    p.use_bf = R.plan.bf ? 1 : 0;
    p.bf_bits_per_key = s->opts.bf_bits_per_key;
    p.num_hashes = s->opts.num_hashes;
    p.inputs_on_device = 1;
    int rc = make_plan(p, pl);
    if (rc) return rc;
    // the filter is the relation's: its geometry depends only on |R|, bits and k
    if (pl.bf && (pl.geom.seg_bits != R.plan.geom.seg_bits || pl.geom.wps != R.plan.geom.wps ||
                  pl.geom.k != R.plan.geom.k))
        return fail(BJ_ERR_INTERNAL, "request filter geometry differs from the relation's");
    return BJ_OK;
}

// The relation with index rel (null if there is none), looked up under the lock.
const Relation* find_rel(const bj_server* s, int32_t rel)
{
    std::lock_guard<std::mutex> lk(s->mu);
    return (rel >= 0 && rel < (int32_t)s->rels.size()) ? s->rels[rel] : nullptr;
}

// Starts a request on a slot (scheduler thread, lock held).
int launch(bj_server* s, Slot& sl, Request* q)
{
    bj_ctx* c = sl.c;
    uint8_t* base = (uint8_t*)s->arena.p;
    uint64_t off = q->ws_off;
    for (const Item& it : q->items) {
        (c->*(it.first)).p = base + off;
        (c->*(it.first)).bytes = it.second;
        off += it.second;
    }
    const Relation& R = *q->relp;
    const RSide rs{P<uint32_t>(R.key), P<uint32_t>(R.val), P<uint64_t>(R.start),
                   R.plan.bf ? P<uint64_t>(R.filter) : nullptr};
    const uint64_t col = align_up(q->cap * 4);
    uint32_t* ok = (uint32_t*)(base + q->out_off);
    uint32_t* orv = (uint32_t*)(base + q->out_off + col);
    uint32_t* osv = (uint32_t*)(base + q->out_off + 2 * col);
    hipStream_t st = c->stream;
    BJ_HIP(hipEventRecord(sl.t0, st));
    tl_reset(c, false);  // only the request's span is timed
    BJ_HIP(launch_zero_u64(P<uint64_t>(c->counters), CTR_N, st));
    int rc = enqueue_s_probe(c, q->pl, rs, q->sk, q->sv, ok, orv, osv, q->cap);
    if (rc) return rc;
    BJ_HIP(hipEventRecord(sl.t1, st));
    BJ_HIP(hipMemcpyAsync(sl.h_ctr, c->counters.p, CTR_N * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    BJ_HIP(hipEventRecord(sl.done, st));
    q->res.key = ok;
    q->res.r_val = orv;
    q->res.s_val = osv;
    return BJ_OK;
}

void detach_views(bj_ctx* c, const std::vector<Item>& items)
{
    for (const Item& it : items) (c->*(it.first)) = DevBuf{};
}

// Completes the request of a slot (scheduler thread, lock held).
void retire(bj_server* s, Slot& sl, int status)
{
    Request* q = sl.cur;
    bj_request_result& r = q->res;
    r.t_end = s->now();
    r.status = status;
    if (status == BJ_OK) {
        const uint64_t* h = sl.h_ctr;
        if (h[CTR_ERR]) {
            r.status = BJ_ERR_INTERNAL;
        } else {
            r.match_count = h[CTR_OUT];
            r.rows_written = std::min(h[CTR_OUT], q->cap);
            r.truncated = h[CTR_OUT] > q->cap ? 1 : 0;
            r.s_survivors = q->pl.bf ? h[CTR_SPRIME] : q->pl.ns;
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, sl.t0, sl.t1) == hipSuccess) r.device_ms = ms;
        }
    }
    if (r.status != BJ_OK) {
        r.key = r.r_val = r.s_val = nullptr;
        r.rows_written = 0;
    }
    detach_views(sl.c, q->items);
    s->alloc.release(q->ws_off, q->ws_bytes);
    if (r.status != BJ_OK) {
        s->alloc.release(q->out_off, q->out_bytes);
        q->out_bytes = 0;
    }
    q->done = true;
    sl.cur = nullptr;
    s->in_flight--;
    s->completed++;
}

void scheduler(bj_server* s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        bool changed = false;
        // retire finished requests
        for (Slot& sl : s->slots) {
            if (!sl.cur) continue;
            const hipError_t e = hipEventQuery(sl.done);
            if (e == hipErrorNotReady) continue;
            retire(s, sl, e == hipSuccess ? BJ_OK : BJ_ERR_HIP);
            changed = true;
        }
        // admit queued requests in arrival order while a slot and arena room are free
        while (!s->stop && !s->queue.empty()) {
            Slot* free_slot = nullptr;
            for (Slot& sl : s->slots)
                if (!sl.cur) {
                    free_slot = &sl;
                    break;
                }
            if (!free_slot) break;
            Request* q = s->queue.front();
            if (!s->alloc.alloc(q->ws_bytes, q->ws_off)) break;
            if (!s->alloc.alloc(q->out_bytes, q->out_off)) {
                s->alloc.release(q->ws_off, q->ws_bytes);
                break;
            }
            s->queue.pop_front();
            free_slot->cur = q;
            s->in_flight++;
            s->max_in_flight = std::max(s->max_in_flight, s->in_flight);
            q->res.t_start = s->now();
            const int rc = launch(s, *free_slot, q);
            if (rc != BJ_OK) {
                (void)hipStreamSynchronize(free_slot->c->stream);
                retire(s, *free_slot, rc);
            }
            changed = true;
        }
        // nothing runs, yet the head of the queue does not fit: the arena holds
        // rows that were never released
        const bool blocked = s->in_flight == 0 && !s->queue.empty();
        if (blocked != s->blocked) {
            s->blocked = blocked;
            changed = true;
        }
        if (changed) s->cv_done.notify_all();
        if (s->stop && s->in_flight == 0) break;  // queued requests are dropped
        if (s->in_flight == 0 && s->queue.empty())
            s->cv_work.wait(lk);
        else
            s->cv_work.wait_for(lk, std::chrono::microseconds(20));
    }
}

// The arena, the slots and the scheduler of a new server (its options checked);
// on an error the caller destroys the half-built server.
int build_server(bj_server* s, int device, const bj_server_opts& opts)
{
    int rc;
    s->t_origin = std::chrono::steady_clock::now();
    uint64_t arena = opts.arena_bytes;
    if (arena == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return fail(BJ_ERR_HIP, "hipMemGetInfo failed");
        arena = fr / 2;
    }
    arena = align_up(arena);
    if ((rc = ensure(s->arena, arena))) return rc;
    s->alloc.reset(arena);
    s->slots.resize(opts.slots);
    for (Slot& sl : s->slots) {
        if ((rc = bj_create(&sl.c, device))) return rc;
        release(sl.c->counters);  // a request brings its counters from the arena
        hipError_t e = hipEventCreate(&sl.t0);
        if (e == hipSuccess) e = hipEventCreate(&sl.t1);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipHostMalloc((void**)&sl.h_ctr, CTR_N * sizeof(uint64_t), hipHostMallocDefault);
        if (e != hipSuccess) return fail(BJ_ERR_HIP, std::string("server slot: ") + hipGetErrorString(e));
    }
    s->sched = std::thread(scheduler, s);
    return BJ_OK;
}

}  // namespace

extern "C" {

void bj_server_opts_default(bj_server_opts* o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->slots = 3;
    o->arena_bytes = 0;
    o->use_bf = 1;
    o->bf_bits_per_key = 10.0;
    o->num_hashes = 3;
    o->radix_bits = -1;
    o->answer_scale = 1.0;
}

int bj_server_destroy(bj_server* s)
{
    if (!s) return BJ_OK;
    if (s->sched.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        s->sched.join();
    }
    (void)hipSetDevice(s->device);
    for (Slot& sl : s->slots) {
        if (sl.c) {
            (void)hipStreamSynchronize(sl.c->stream);
            sl.c->counters = DevBuf{};
            bj_destroy(sl.c);
        }
        if (sl.t0) (void)hipEventDestroy(sl.t0);
        if (sl.t1) (void)hipEventDestroy(sl.t1);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.h_ctr) (void)hipHostFree(sl.h_ctr);
    }
    for (auto& kv : s->reqs) delete kv.second;
    for (Relation* r : s->rels) {
        release(r->key);
        release(r->val);
        release(r->start);
        release(r->filter);
        delete r;
    }
    release(s->arena);
    if (s->main) bj_destroy(s->main);
    delete s;
    return BJ_OK;
}

int bj_server_create(int device, const bj_server_opts* o, bj_server** out)
{
    return guarded([&]() -> int {
        if (!out) return fail(BJ_ERR_INVALID, "null argument");
        *out = nullptr;
        bj_server_opts opts;
        bj_server_opts_default(&opts);
        if (o) opts = *o;
        if (opts.slots < 1 || opts.slots > SRV_MAX_SLOTS) return fail(BJ_ERR_INVALID, "slots must be in [1, 8]");
        if (opts.radix_bits == 0 || opts.radix_bits > 15) return fail(BJ_ERR_INVALID, "radix_bits must be -1 or 1..15");
        if (opts.use_bf && !(opts.bf_bits_per_key > 0.0 && std::isfinite(opts.bf_bits_per_key)))
            return fail(BJ_ERR_INVALID, "bf_bits_per_key must be a finite number > 0");
        if (!std::isfinite(opts.answer_scale)) return fail(BJ_ERR_INVALID, "answer_scale must be a finite number");
        bj_ctx* main = nullptr;
        int rc = bj_create(&main, device);
        if (rc) return rc;
        bj_server* s = nullptr;
        try {
            s = new bj_server();
        } catch (...) {
            bj_destroy(main);
            throw;
        }
        s->device = device;
        s->opts = opts;
        s->main = main;
        // from here on the server owns everything: a throw (slot vector,
        // scheduler thread) tears down what was built so far
        try {
            rc = build_server(s, device, opts);
        } catch (...) {
            bj_server_destroy(s);
            throw;
        }
        if (rc) {
            bj_server_destroy(s);
            return rc;
        }
        *out = s;
        return BJ_OK;
    });
}

int bj_server_add_relation(bj_server* s, const uint32_t* r_key, const uint32_t* r_val, int64_t n_r,
                           int32_t* rel_id)
{
    return guarded([&]() -> int {
        if (!s || !rel_id || n_r < 0 || (n_r > 0 && (!r_key || !r_val))) return fail(BJ_ERR_INVALID, "bad argument");
        bj_params p;
        bj_params_default(&p);
        p.r_size = n_r;
        p.s_size = n_r;
        p.radix_bits = s->opts.radix_bits;
        p.use_bf = s->opts.use_bf && n_r > 0 ? 1 : 0;
        p.bf_bits_per_key = s->opts.bf_bits_per_key;
        p.num_hashes = s->opts.num_hashes;
        p.inputs_on_device = 1;
        std::unique_ptr<Relation> R(new Relation());
        R->nr = (uint64_t)n_r;
        int rc = make_plan(p, R->plan);
        if (rc) {
            return rc;
        }
        Plan pl = R->plan;
        pl.ns = 0;  // R side only
        pl.G_s = 1;
        bj_ctx* c = s->main;
        hipError_t e = hipSetDevice(s->device);
        if (e != hipSuccess) {
            return fail(BJ_ERR_HIP, "hipSetDevice failed");
        }
        if ((rc = ensure_workspace(c, pl))) {
            return rc;
        }
        tl_reset(c, false);
        e = launch_zero_u64(P<uint64_t>(c->counters), CTR_N, c->stream);
        if (e != hipSuccess) {
            return fail(BJ_ERR_HIP, std::string("relation counters: ") + hipGetErrorString(e));
        }
        if ((rc = enqueue_r(c, pl, r_key, r_val))) {
            return rc;
        }
        e = hipMemcpyAsync(c->h_counters, c->counters.p, CTR_N * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            return fail(BJ_ERR_HIP, std::string("relation partition: ") + hipGetErrorString(e));
        }
        if (c->h_counters[CTR_ERR]) {
            return fail(BJ_ERR_INTERNAL, "device check flags " + std::to_string(c->h_counters[CTR_ERR]));
        }
        // the partitioned relation moves from the context into the server
        std::swap(R->key, c->r_key);
        std::swap(R->val, c->r_val);
        std::swap(R->start, c->r_start);
        if (R->plan.bf) std::swap(R->filter, c->filter);
        std::lock_guard<std::mutex> lk(s->mu);
        s->rel_bytes += R->key.bytes + R->val.bytes + R->start.bytes + R->filter.bytes;
        s->rels.push_back(R.get());  // on a throw R still owns the relation and frees it
        R.release();                 // owned by the server from here on
        *rel_id = (int32_t)s->rels.size() - 1;
        return BJ_OK;
    });
}

int bj_server_request_bytes(const bj_server* s, int32_t rel, int64_t n_s, uint64_t* bytes)
{
    if (!s || !bytes || n_s < 0) return fail(BJ_ERR_INVALID, "bad argument");
    const Relation* R = find_rel(s, rel);
    if (!R) return fail(BJ_ERR_INVALID, "no such relation");
    Plan pl;
    int rc = make_request_plan(s, *R, (uint64_t)n_s, pl);
    if (rc) return rc;
    std::vector<Item> items;
    uint64_t ws = 0;
    if ((rc = request_items(pl, items, ws))) return rc;
    uint64_t cap = 0;
    if ((rc = row_capacity((uint64_t)n_s, s->opts.answer_scale, cap))) return rc;
    *bytes = ws + 3 * align_up(cap * 4);
    return BJ_OK;
}

int bj_server_submit(bj_server* s, int32_t rel, const uint32_t* s_key, const uint32_t* s_val, int64_t n_s,
                     uint64_t* id)
{
    return guarded([&]() -> int {
        if (!s || !id || n_s < 0 || (n_s > 0 && (!s_key || !s_val))) return fail(BJ_ERR_INVALID, "bad argument");
        const Relation* R = find_rel(s, rel);
        if (!R) return fail(BJ_ERR_INVALID, "no such relation");
        std::unique_ptr<Request> q(new Request());
        int rc = make_request_plan(s, *R, (uint64_t)n_s, q->pl);
        if (rc == BJ_OK) rc = request_items(q->pl, q->items, q->ws_bytes);
        if (rc) return rc;
        if ((rc = row_capacity((uint64_t)n_s, s->opts.answer_scale, q->cap))) return rc;
        q->out_bytes = 3 * align_up(q->cap * 4);
        if (align_up(q->ws_bytes) + align_up(q->out_bytes) > s->alloc.size())
            return fail(BJ_ERR_SIZE, "request footprint exceeds the server arena");
        q->rel = rel;
        q->relp = R;
        q->sk = s_key;
        q->sv = s_val;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            q->id = s->next_id++;
            q->res.id = q->id;
            q->res.relation = rel;
            q->res.s_size = (uint64_t)n_s;
            q->res.capacity = q->cap;
            q->res.t_submit = s->now();
            auto it = s->reqs.emplace(q->id, q.get()).first;
            try {
                s->queue.push_back(q.get());
            } catch (...) {
                s->reqs.erase(it);
                throw;
            }
            *id = q.release()->id;  // owned by the server from here on
            s->submitted++;
        }
        s->cv_work.notify_all();
        return BJ_OK;
    });
}

int bj_server_poll(bj_server* s, bj_request_result* out, int cap, int* n)
{
    return guarded([&]() -> int {
        if (!s || !n || (cap > 0 && !out)) return fail(BJ_ERR_INVALID, "bad argument");
        std::lock_guard<std::mutex> lk(s->mu);
        int k = 0;
        for (auto& kv : s->reqs) {
            if (k >= cap) break;
            Request* q = kv.second;
            if (q->done && !q->reported) {
                q->reported = true;
                out[k++] = q->res;
            }
        }
        *n = k;
        return BJ_OK;
    });
}

int bj_server_wait(bj_server* s, uint64_t id, bj_request_result* out)
{
    if (!s || !out) return fail(BJ_ERR_INVALID, "bad argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->reqs.find(id) == s->reqs.end()) return fail(BJ_ERR_INVALID, "no such request");
    // the request is looked up again after every wake-up: another thread may
    // have released it meanwhile (release needs it done, so it is then gone)
    Request* q = nullptr;
    s->cv_done.wait(lk, [&] {
        auto it = s->reqs.find(id);
        q = it == s->reqs.end() ? nullptr : it->second;
        return !q || q->done || s->blocked;
    });
    if (!q) return fail(BJ_ERR_INVALID, "request released by another thread while waiting");
    if (!q->done)
        return fail(BJ_ERR_NOMEM, "request queued behind rows that were never released (bj_server_release)");
    q->reported = true;
    *out = q->res;
    return BJ_OK;
}

int bj_server_copy_rows(bj_server* s, const bj_request_result* r, uint32_t* key, uint32_t* r_val, uint32_t* s_val,
                        uint64_t n)
{
    return guarded([&]() -> int {
        if (!s || !r) return fail(BJ_ERR_INVALID, "bad argument");
        if (n > r->rows_written) return fail(BJ_ERR_INVALID, "n exceeds rows_written");
        if (n == 0) return BJ_OK;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            auto it = s->reqs.find(r->id);
            if (it == s->reqs.end() || !it->second->done) return fail(BJ_ERR_INVALID, "no such completed request");
        }
        hipStream_t st = s->main->stream;
        BJ_HIP(hipSetDevice(s->device));
        if (key) BJ_HIP(hipMemcpyAsync(key, r->key, n * 4, hipMemcpyDeviceToHost, st));
        if (r_val) BJ_HIP(hipMemcpyAsync(r_val, r->r_val, n * 4, hipMemcpyDeviceToHost, st));
        if (s_val) BJ_HIP(hipMemcpyAsync(s_val, r->s_val, n * 4, hipMemcpyDeviceToHost, st));
        BJ_HIP(hipStreamSynchronize(st));
        return BJ_OK;
    });
}

int bj_server_release(bj_server* s, uint64_t id)
{
    if (!s) return fail(BJ_ERR_INVALID, "bad argument");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) return fail(BJ_ERR_INVALID, "no such request");
        Request* q = it->second;
        if (!q->done) return fail(BJ_ERR_INVALID, "request still running");
        if (q->out_bytes) s->alloc.release(q->out_off, q->out_bytes);
        s->reqs.erase(it);
        delete q;
        s->released++;
        s->blocked = false;  // the scheduler looks at the queue again
    }
    s->cv_work.notify_all();  // room in the arena for queued requests
    return BJ_OK;
}

int bj_server_drain(bj_server* s)
{
    if (!s) return fail(BJ_ERR_INVALID, "bad argument");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return (s->queue.empty() && s->in_flight == 0) || s->blocked; });
    if (!s->queue.empty())
        return fail(BJ_ERR_NOMEM, "requests queued behind rows that were never released (bj_server_release)");
    return BJ_OK;
}

int bj_server_stats_get(bj_server* s, bj_server_stats* out)
{
    if (!s || !out) return fail(BJ_ERR_INVALID, "bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof(*out));
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->released = s->released;
    out->max_in_flight = s->max_in_flight;
    out->arena_bytes = s->alloc.size();
    out->arena_peak_bytes = s->alloc.peak();
    out->relation_bytes = s->rel_bytes;
    return BJ_OK;
}

}  // extern "C"
