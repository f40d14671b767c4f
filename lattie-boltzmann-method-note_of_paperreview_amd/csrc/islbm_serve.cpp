// islbm_serve.cpp -- continuous-batching serving loop over many resident solves (the
// islbm_batch_* entry points of include/islbm.h).
//
// The reference runs one solve per program (its main, absent from the notes, loops
// Launch_CollisionStreaming over one global state).  Here a request is one periodic-hill
// solve with its own grid, Reynolds number, hill steepness and step count.  Requests are
// admitted in submission order while the device-memory budget and the slot count allow;
// an admitted solve keeps its populations resident.  Up to 8 admitted solves that take the
// FP32 tiled kernel advance together: each step of all of them is one grouped launch
// (launch_tiled_group), and chunks of those launches are replayed from a HIP graph
// recorded for the current members, two chunks in flight on the batch's stream; the graph
// is recorded again only when the members, their step limits, their buffer parities or
// their forcing updates change.  Other solves (FP64, 16-bit storage, grids the grouped
// kernel does not take) run on their own streams in graph-replayed chunks
// (islbm_step_graph).  A request that arrives while others run starts as soon as a slot
// and memory free up.  A finished solve is retired (fields downloaded if requested, a
// summary kept) and its memory returned to the budget.
#include <hip/hip_runtime.h>

#include "../../include/islbm.h"
#include "islbm_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

namespace {

thread_local std::string g_berr;

int bfail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_berr = buf;
    return code;
}

double now_s()
{
    using clk = std::chrono::steady_clock;
    return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

constexpr int kDepth = 2;  // chunks in flight per solve

struct Request {
    int64_t id = 0;
    islbm_params p{};
    double u0[3] = {0, 0, 0};
    double amp = 0;
    uint64_t seed = 0;
    int64_t nsteps = 0;
    int64_t queued_steps = 0;  // steps enqueued on the device
    int64_t done_steps = 0;    // steps whose chunk has completed
    int keep = 0;
    int state = ISLBM_REQ_QUEUED;
    double t_submit = 0, t_admit = 0, t_done = 0;
    size_t bytes = 0;
    islbm_solver* s = nullptr;
    hipEvent_t ev[kDepth] = {nullptr, nullptr};
    int64_t ev_steps[kDepth] = {0, 0};  // steps covered by the chunk behind each event
    int head = 0, inflight = 0;         // ring of in-flight chunk events
    bool grouped = false;               // advanced by the batch's grouped launches
    std::vector<double> fields;         // rho u v w after completion (keep != 0)
    double summary[4] = {0, 0, 0, 0};   // nodes, mean rho, max |u|, mean v
    std::string error;
};

}  // namespace

struct islbm_batch {
    int device = 0;
    size_t budget = 0, used = 0;
    int max_active = 1;
    int chunk = 2;
    int64_t next_id = 1;
    std::deque<Request*> queue;
    std::vector<Request*> active;
    std::map<int64_t, Request*> all;
    int64_t n_done = 0, n_failed = 0;
    // grouped stepping: the batch's stream, a ring of in-flight chunks (members and their
    // steps per chunk), and the recorded chunk graph with the key it was recorded for
    hipStream_t gstream = nullptr;
    islbm::GroupFork gfork{nullptr, nullptr, nullptr};  // second stream of mixed-width groups
    hipEvent_t gev[kDepth] = {nullptr, nullptr};
    std::vector<Request*> gmem[kDepth];
    std::vector<int> glim[kDepth];
    int ghead = 0, ginflight = 0;
    hipGraphExec_t gexec = nullptr;
    std::vector<int64_t> gkey;
    // the graph each in-flight chunk was launched from, and graphs replaced by a re-recording
    // while chunks launched from them may still run (destroyed once none of those remains)
    hipGraphExec_t gslot_exec[kDepth] = {nullptr, nullptr};
    std::vector<hipGraphExec_t> gretired;
    // device partials of a finished solve's summary (islbm::field_summary)
    double* dsum = nullptr;
    int64_t n_group_records = 0;  // times the chunk graph was recorded
    int64_t n_group_chunks = 0;   // grouped chunks launched
};

namespace {

// destroys the replaced chunk graphs that no in-flight chunk was launched from
void reap_graphs(islbm_batch* b)
{
    for (size_t t = 0; t < b->gretired.size();) {
        bool used = false;
        for (int k = 0; k < b->ginflight; ++k) used = used || b->gslot_exec[(b->ghead + k) % kDepth] == b->gretired[t];
        if (used) {
            ++t;
            continue;
        }
        (void)hipGraphExecDestroy(b->gretired[t]);
        b->gretired.erase(b->gretired.begin() + (long)t);
    }
}

void release(islbm_batch* b, Request* r)
{
    for (hipEvent_t& e : r->ev)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
    if (r->s) {
        islbm_destroy(r->s);
        r->s = nullptr;
        b->used -= r->bytes;
    }
}

void set_failed(islbm_batch* b, Request* r, const char* why)
{
    // a grouped solve that fails may still be read by chunks in flight on the batch's
    // stream (a finished one is not: no chunk after its last takes it)
    if (r->grouped && r->s && b->gstream) (void)hipStreamSynchronize(b->gstream);
    r->error = why ? why : "";
    r->state = ISLBM_REQ_FAILED;
    r->inflight = 0;
    r->t_done = now_s();
    release(b, r);
    b->n_failed++;
}

// download the fields of a finished solve, keep them and/or a summary, free the solve
int finish(islbm_batch* b, Request* r)
{
    const size_t n = (size_t)r->p.nx * r->p.ny * r->p.nz;
    // the summary from device reductions; the fields only when the caller keeps them
    double sum[3];
    int rc = islbm::field_summary(r->s, b->dsum, sum);
    if (rc) return rc;
    r->summary[0] = (double)n;
    r->summary[1] = sum[0] / n;
    r->summary[2] = sum[2];
    r->summary[3] = sum[1] / n;
    if (r->keep) {
        r->fields.resize(4 * n);
        double* f = r->fields.data();
        if ((rc = islbm_download_macros(r->s, f, f + n, f + 2 * n, f + 3 * n))) return rc;
    }
    r->state = ISLBM_REQ_DONE;
    r->t_done = now_s();
    release(b, r);
    b->n_done++;
    return ISLBM_OK;
}

// queue chunks of an active solve until kDepth are in flight or all steps are queued
int feed(islbm_batch* b, Request* r)
{
    while (r->inflight < kDepth && r->queued_steps < r->nsteps) {
        const int64_t left = r->nsteps - r->queued_steps;
        const int n = (int)std::min<int64_t>(left, b->chunk);
        int rc = (n == b->chunk) ? islbm_step_graph(r->s, n) : islbm_step(r->s, n);
        if (rc) return rc;
        void* st = nullptr;
        islbm_stream(r->s, &st);
        const int slot = (r->head + r->inflight) % kDepth;
        if (hipEventRecord(r->ev[slot], (hipStream_t)st) != hipSuccess) return bfail(ISLBM_ERR_HIP, "event record");
        r->ev_steps[slot] = n;
        r->inflight++;
        r->queued_steps += n;
    }
    return ISLBM_OK;
}

// The next chunk of the grouped solves: members (in admission order), their step limits,
// recorded into a graph when the key changed, replayed on the batch's stream.
int feed_group(islbm_batch* b)
{
    while (b->ginflight < kDepth) {
        std::vector<Request*> mem;
        std::vector<islbm_solver*> ss;
        std::vector<int> lim;
        std::vector<int64_t> key;
        for (Request* r : b->active) {
            if (!r->grouped || r->state != ISLBM_REQ_ACTIVE || r->queued_steps >= r->nsteps) continue;
            const int l = islbm::group_limit(r->s, std::min<int64_t>(b->chunk, r->nsteps - r->queued_steps));
            mem.push_back(r);
            ss.push_back(r->s);
            lim.push_back(l);
            // request ids are never reused (a new solve may take a freed solve's address)
            key.insert(key.end(), {r->id, l, islbm::group_parity(r->s)});
        }
        if (mem.empty()) return ISLBM_OK;
        // a chunk that ends on a member's forcing update records that update too
        for (size_t e = 0; e < mem.size(); ++e) {
            int64_t st = 0;
            islbm_step_count(mem[e]->s, &st);
            key.push_back((st + lim[e]) % mem[e]->p.ndtfrc == 0);
        }
        if (!b->gexec || key != b->gkey) {
            if (b->gexec) {
                // chunks launched from it may still be in flight: destroyed once they retire
                b->gretired.push_back(b->gexec);
                b->gexec = nullptr;
                reap_graphs(b);
            }
            if (hipStreamBeginCapture(b->gstream, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return bfail(ISLBM_ERR_HIP, "group capture");
            const int rc = islbm::group_record(b->gstream, ss.data(), lim.data(), (int)ss.size(), &b->gfork);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(b->gstream, &g);
            if (rc || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return rc ? rc : bfail(ISLBM_ERR_HIP, "group capture: %s", hipGetErrorString(e));
            }
            const hipError_t ei = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) {
                b->gexec = nullptr;
                return bfail(ISLBM_ERR_HIP, "group graph: %s", hipGetErrorString(ei));
            }
            b->gkey = key;
            b->n_group_records++;
        }
        if (hipGraphLaunch(b->gexec, b->gstream) != hipSuccess) return bfail(ISLBM_ERR_HIP, "group graph launch");
        b->n_group_chunks++;
        const int slot = (b->ghead + b->ginflight) % kDepth;
        if (hipEventRecord(b->gev[slot], b->gstream) != hipSuccess) return bfail(ISLBM_ERR_HIP, "event record");
        b->gslot_exec[slot] = b->gexec;
        islbm::group_commit(ss.data(), lim.data(), (int)ss.size());
        b->gmem[slot] = mem;
        b->glim[slot] = lim;
        for (size_t e = 0; e < mem.size(); ++e) mem[e]->queued_steps += lim[e];
        b->ginflight++;
    }
    return ISLBM_OK;
}

int admit(islbm_batch* b, Request* r)
{
    islbm_params p = r->p;
    p.device = b->device;
    int rc = islbm_create(&p, 0, &r->s);
    if (rc) return rc;
    b->used += r->bytes;
    for (hipEvent_t& e : r->ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return bfail(ISLBM_ERR_HIP, "event create");
    if ((rc = islbm_init_equilibrium(r->s, 1.0, r->u0, r->amp, r->seed))) return rc;
    int grouped = 0;
    for (Request* a : b->active) grouped += a->grouped;
    r->grouped = grouped < islbm::kGroupMaxSolves && islbm::group_eligible(r->s);
    r->state = ISLBM_REQ_ACTIVE;
    r->t_admit = now_s();
    return ISLBM_OK;
}

}  // namespace

extern "C" {

int islbm_batch_create(int32_t device, size_t hbm_budget, int32_t max_active, int32_t chunk, islbm_batch** out)
{
    if (!out || max_active <= 0 || chunk <= 0 || chunk % 2) return bfail(ISLBM_ERR_ARG, "bad batch arguments (chunk must be even)");
    if (hipSetDevice(device) != hipSuccess) return bfail(ISLBM_ERR_HIP, "hipSetDevice(%d)", device);
    islbm_batch* b = new islbm_batch();
    b->device = device;
    if (hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            delete b;
            return bfail(ISLBM_ERR_HIP, "hipMemGetInfo");
        }
        hbm_budget = fr / 10 * 9;
    }
    b->budget = hbm_budget;
    b->max_active = max_active;
    b->chunk = chunk;
    bool ok = hipStreamCreateWithFlags(&b->gstream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipStreamCreateWithFlags(&b->gfork.st2, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&b->gfork.fork, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&b->gfork.join, hipEventDisableTiming) == hipSuccess;
    for (hipEvent_t& e : b->gev) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipMalloc((void**)&b->dsum, islbm::kSummaryScratch * sizeof(double)) == hipSuccess;
    if (!ok) {
        islbm_batch_destroy(b);
        return bfail(ISLBM_ERR_HIP, "batch stream, events or summary buffer");
    }
    *out = b;
    return ISLBM_OK;
}

int islbm_batch_destroy(islbm_batch* b)
{
    if (!b) return ISLBM_OK;
    (void)hipSetDevice(b->device);
    if (b->gstream) (void)hipStreamSynchronize(b->gstream);
    for (auto& kv : b->all) {
        release(b, kv.second);
        delete kv.second;
    }
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    for (hipGraphExec_t g : b->gretired) (void)hipGraphExecDestroy(g);
    if (b->dsum) (void)hipFree(b->dsum);
    for (hipEvent_t e : b->gev)
        if (e) (void)hipEventDestroy(e);
    if (b->gstream) (void)hipStreamDestroy(b->gstream);
    if (b->gfork.st2) (void)hipStreamDestroy(b->gfork.st2);
    for (hipEvent_t e : {b->gfork.fork, b->gfork.join})
        if (e) (void)hipEventDestroy(e);
    delete b;
    return ISLBM_OK;
}

int islbm_batch_submit(islbm_batch* b, const islbm_params* p, const double u0[3], double amp, uint64_t seed,
                       int64_t nsteps, int32_t keep_fields, int64_t* id)
{
    if (!b || !p || !u0 || !id || nsteps <= 0) return bfail(ISLBM_ERR_ARG, "bad submit arguments");
    Request* r = new Request();
    r->id = b->next_id++;
    r->p = *p;
    std::memcpy(r->u0, u0, sizeof r->u0);
    r->amp = amp;
    r->seed = seed;
    r->nsteps = nsteps;
    r->keep = keep_fields;
    r->t_submit = now_s();
    b->all[r->id] = r;
    *id = r->id;
    islbm_params q = *p;
    q.device = b->device;
    if (islbm_required_bytes(&q, &r->bytes) != ISLBM_OK) {
        set_failed(b, r, "invalid parameters");
        return ISLBM_OK;
    }
    if (r->bytes > b->budget) {
        set_failed(b, r, "solve larger than the whole budget");
        return ISLBM_OK;
    }
    b->queue.push_back(r);
    return ISLBM_OK;
}

int islbm_batch_advance(islbm_batch* b, int32_t wait, int64_t out[4])
{
    if (!b) return bfail(ISLBM_ERR_ARG, "null batch");
    if (hipSetDevice(b->device) != hipSuccess) return bfail(ISLBM_ERR_HIP, "hipSetDevice");
    std::vector<Request*> finished;
    for (int pass = 0; pass < 2; ++pass) {
        // retire completed grouped chunks, in order
        bool progressed = false;
        while (b->ginflight > 0) {
            const hipError_t q = hipEventQuery(b->gev[b->ghead]);
            if (q == hipErrorNotReady) break;
            const std::vector<Request*>& mem = b->gmem[b->ghead];
            for (size_t e = 0; e < mem.size(); ++e) {
                // a member failed by an earlier chunk or by a feed error is counted once
                if (mem[e]->state != ISLBM_REQ_ACTIVE) continue;
                if (q != hipSuccess)
                    set_failed(b, mem[e], hipGetErrorString(q));
                else
                    mem[e]->done_steps += b->glim[b->ghead][e];
            }
            b->gmem[b->ghead].clear();
            b->gslot_exec[b->ghead] = nullptr;
            b->ghead = (b->ghead + 1) % kDepth;
            b->ginflight--;
            progressed = true;
        }
        reap_graphs(b);
        // retire completed chunks, in order per solve
        for (size_t t = 0; t < b->active.size();) {
            Request* r = b->active[t];
            while (r->state == ISLBM_REQ_ACTIVE && r->inflight > 0) {
                const hipError_t q = hipEventQuery(r->ev[r->head]);
                if (q == hipErrorNotReady) break;
                if (q != hipSuccess) {
                    set_failed(b, r, hipGetErrorString(q));
                    break;
                }
                r->done_steps += r->ev_steps[r->head];
                r->head = (r->head + 1) % kDepth;
                r->inflight--;
                progressed = true;
            }
            // a solve that took all its steps leaves the active set (its slot frees; its memory
            // stays counted until it is finished below, after the next chunks are queued)
            if (r->state == ISLBM_REQ_ACTIVE && r->done_steps == r->nsteps) finished.push_back(r);
            if (r->state != ISLBM_REQ_ACTIVE || r->done_steps == r->nsteps) {
                b->active.erase(b->active.begin() + t);
                progressed = true;
            } else {
                ++t;
            }
        }
        // admit in submission order while a slot and memory are free
        while (!b->queue.empty() && (int)b->active.size() < b->max_active &&
               b->used + b->queue.front()->bytes <= b->budget) {
            Request* r = b->queue.front();
            b->queue.pop_front();
            int rc = admit(b, r);
            if (rc) {
                set_failed(b, r, islbm_last_error());
                continue;
            }
            b->active.push_back(r);
            progressed = true;
        }
        // keep every active solve fed
        for (Request* r : b->active) {
            if (r->grouped) continue;
            int rc = feed(b, r);
            if (rc) set_failed(b, r, islbm_last_error());
        }
        if (int rc = feed_group(b)) {
            const std::string why = g_berr.empty() ? islbm_last_error() : g_berr;
            for (Request* r : b->active)
                if (r->grouped && r->state == ISLBM_REQ_ACTIVE) set_failed(b, r, why.c_str());
            (void)rc;
        }
        // summaries, kept fields and release of the finished solves, while the queued chunks run
        for (Request* r : finished)
            if (finish(b, r) != ISLBM_OK) set_failed(b, r, islbm_last_error());
        finished.clear();
        if (!wait || progressed || b->active.empty()) break;
        // block until the oldest in-flight chunk (grouped first) completes
        if (b->ginflight > 0) {
            (void)hipEventSynchronize(b->gev[b->ghead]);
        } else {
            for (Request* r : b->active)
                if (r->inflight > 0) {
                    (void)hipEventSynchronize(r->ev[r->head]);
                    break;
                }
        }
    }
    if (out) {
        out[0] = (int64_t)b->queue.size();
        out[1] = (int64_t)b->active.size();
        out[2] = b->n_done;
        out[3] = b->n_failed;
    }
    return ISLBM_OK;
}

int islbm_batch_status(islbm_batch* b, int64_t id, int32_t* state, double times[3], int64_t* steps_done)
{
    if (!b) return bfail(ISLBM_ERR_ARG, "null batch");
    auto it = b->all.find(id);
    if (it == b->all.end()) return bfail(ISLBM_ERR_ARG, "unknown request %lld", (long long)id);
    const Request* r = it->second;
    if (state) *state = r->state;
    if (times) {
        times[0] = r->t_submit;
        times[1] = r->t_admit;
        times[2] = r->t_done;
    }
    if (steps_done) *steps_done = r->done_steps;
    return ISLBM_OK;
}

int islbm_batch_result(islbm_batch* b, int64_t id, double* rho, double* u, double* v, double* w)
{
    if (!b) return bfail(ISLBM_ERR_ARG, "null batch");
    auto it = b->all.find(id);
    if (it == b->all.end()) return bfail(ISLBM_ERR_ARG, "unknown request %lld", (long long)id);
    Request* r = it->second;
    if (r->state != ISLBM_REQ_DONE) return bfail(ISLBM_ERR_STATE, "request %lld is not done", (long long)id);
    if (r->fields.empty()) return bfail(ISLBM_ERR_STATE, "request %lld kept no fields", (long long)id);
    const size_t n = (size_t)r->p.nx * r->p.ny * r->p.nz;
    double* outs[4] = {rho, u, v, w};
    for (int a = 0; a < 4; ++a)
        if (outs[a]) std::memcpy(outs[a], r->fields.data() + a * n, n * sizeof(double));
    std::vector<double>().swap(r->fields);
    return ISLBM_OK;
}

int islbm_batch_summary(islbm_batch* b, int64_t id, double out[4])
{
    if (!b || !out) return bfail(ISLBM_ERR_ARG, "null argument");
    auto it = b->all.find(id);
    if (it == b->all.end()) return bfail(ISLBM_ERR_ARG, "unknown request %lld", (long long)id);
    std::memcpy(out, it->second->summary, sizeof it->second->summary);
    return ISLBM_OK;
}

int islbm_batch_bytes_in_use(islbm_batch* b, size_t* bytes)
{
    if (!b || !bytes) return bfail(ISLBM_ERR_ARG, "null argument");
    *bytes = b->used;
    return ISLBM_OK;
}

int islbm_batch_group_stats(islbm_batch* b, int64_t out[2])
{
    if (!b || !out) return bfail(ISLBM_ERR_ARG, "null argument");
    out[0] = b->n_group_chunks;
    out[1] = b->n_group_records;
    return ISLBM_OK;
}

const char* islbm_batch_last_error(void) { return g_berr.c_str(); }

double islbm_batch_now(void) { return now_s(); }

}  // extern "C"
