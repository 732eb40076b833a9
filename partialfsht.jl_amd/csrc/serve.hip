// serve.hip — the serving runtime: continuous batching of transform requests over a device-resident
// plan cache admitted by HBM budget.
//
// The reference has no counterpart: a PartialFSHT.Plan is built by its caller and applied by it
// (src/Plan.jl:84-148). A service answering synthesis and adjoint requests on recurring point sets
// keeps their plans resident instead, under caller-chosen 64-bit keys:
//   - pfsht_server_plan registers a geometry: its plan is built with room for `pack` transforms
//     (or the request's ntrans when larger), FFT plans for every power-of-two transform count, and
//     FP32 grids where the requested tolerance allows it (tol >= 1e-6); the plan's size is estimated
//     on the host first and least-recently-used idle plans are evicted before anything is allocated;
//   - pfsht_server_submit only queues a request and returns a ticket;
//   - a dispatcher thread owns the GPU side: whenever a stream has a free launch slot it takes the
//     oldest queued request and every later queued request of the same plan and direction that
//     still fits the plan's transform room, and issues them as one column apply (the requests'
//     columns are read and written in place, no copies), replayed from a hipGraph captured once per
//     (plan, direction, transform count). Requests that arrive while others run are packed into the
//     next launch of their plan: the batch grows with the load.
// Applies of one plan are ordered (they share its work buffers); applies of different plans run
// concurrently on the stream pool (default priority). All calls are thread-safe.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/pfsht.h"
#include "pfsht_common.hpp"

namespace pfsht {
// capi.hip
int64_t plan_estimate_bytes(const double *theta, const double *phi, int64_t M, int L, double tol, int B,
                            const pfsht_plan_opts &o);
Status plan_order_begin(pfsht_plan_t p, hipStream_t s);
Status plan_order_end(pfsht_plan_t p, hipStream_t s);
Status plan_launches(pfsht_plan_t p, int dir, InCols in, OutCols out, int nb, hipStream_t s);
std::mutex &plan_mutex(pfsht_plan_t p);
void plan_shape(pfsht_plan_t p, int &L, int64_t &M, int &B, int &dev);
void plan_release(pfsht_plan_t p);
}  // namespace pfsht

namespace {

using pfsht::InCols;
using pfsht::OutCols;

constexpr int64_t kTicketRing = 1 << 16;  // tickets older than this many submissions are recycled
constexpr int kMaxCols = 64;              // transforms per launch (a plan's room)
constexpr int kTabSlots = 64;             // pinned column-table staging slots

enum TicketState { kFree = 0, kQueued = 1, kLaunched = 2 };

struct Ticket {
    hipEvent_t start = nullptr, end = nullptr;
    int state = kFree;
};

struct Graph {
    hipGraphExec_t exec = nullptr;
    void **tab = nullptr;  // the device column table the graph's kernels read (2 x room pointers)
};

struct Entry {
    pfsht_plan_t plan = nullptr;
    int64_t bytes = 0;
    uint64_t last_use = 0;
    int ntrans = 1;   // transforms per request
    int room = 1;     // transforms per launch (the plan's ntrans)
    int L = 0;
    int64_t M = 0;
    int pending = 0;  // requests queued or in flight
    bool building = false;
    hipEvent_t last = nullptr;  // end of the plan's last launch
    bool has_last = false;
    std::map<int, Graph> graphs;  // by dir * 1024 + transform count
    bool graph_failed = false;
};

struct Request {
    int64_t ticket;
    uint64_t key;
    int dir;
    const void *in;
    void *out;
};

}  // namespace

struct pfsht_server_s {
    int dev = 0;
    pfsht_server_opts opt{};
    std::vector<hipStream_t> streams;
    std::vector<std::deque<hipEvent_t>> inflight;  // end events of each stream's launches
    std::vector<hipEvent_t> event_pool;
    size_t next_stream = 0;
    int64_t resident = 0;
    uint64_t tick = 0;
    std::unordered_map<uint64_t, Entry> cache;
    std::deque<Request> queue;
    std::vector<pfsht_plan_t> graveyard;  // evicted plans, freed by the dispatcher
    std::vector<Graph> dead_graphs;
    std::vector<Ticket> tickets;
    int64_t next_ticket = 0;
    hipEvent_t epoch = nullptr;
    // pinned staging of column tables, each slot reused once the copy that read it has run
    void **tab_host = nullptr, **tab_dev = nullptr;
    hipEvent_t tab_ev[kTabSlots] = {};
    bool tab_used[kTabSlots] = {};
    int tab_next = 0;
    pfsht_server_stats stats{};
    std::mutex mu;
    std::condition_variable cv, cv_launched;
    bool stop = false;
    std::thread dispatcher;

    ~pfsht_server_s();
};

namespace {

hipEvent_t take_event(pfsht_server_s &srv) {
    if (!srv.event_pool.empty()) {
        hipEvent_t e = srv.event_pool.back();
        srv.event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
}

void free_graph(Graph &g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.tab) (void)hipFree(g.tab);
    g = Graph{};
}

// A plan may be evicted when no request of it is queued or running and its last launch has ended.
bool idle(const Entry &e) {
    if (e.building || e.pending > 0) return false;
    return !e.has_last || hipEventQuery(e.last) == hipSuccess;
}

// Evicts least-recently-used idle plans until `need` more bytes fit in the budget (mutex held); the
// plans go to the graveyard and are freed outside the mutex.
bool make_room(pfsht_server_s &srv, int64_t need) {
    while (srv.resident + need > srv.opt.hbm_budget_bytes) {
        auto victim = srv.cache.end();
        for (auto it = srv.cache.begin(); it != srv.cache.end(); ++it)
            if (idle(it->second) && (victim == srv.cache.end() || it->second.last_use < victim->second.last_use))
                victim = it;
        if (victim == srv.cache.end()) return false;
        srv.resident -= victim->second.bytes;
        for (auto &kv : victim->second.graphs) srv.dead_graphs.push_back(kv.second);
        if (victim->second.last) (void)hipEventDestroy(victim->second.last);
        srv.graveyard.push_back(victim->second.plan);
        srv.cache.erase(victim);
        srv.stats.evictions++;
    }
    return true;
}

void bury(pfsht_server_s &srv, std::unique_lock<std::mutex> &lk) {
    if (srv.graveyard.empty() && srv.dead_graphs.empty()) return;
    std::vector<pfsht_plan_t> dead;
    std::vector<Graph> graphs;
    dead.swap(srv.graveyard);
    graphs.swap(srv.dead_graphs);
    lk.unlock();
    for (auto &g : graphs) free_graph(g);
    for (auto p : dead) pfsht::plan_release(p);  // waits for that plan's own work only
    lk.lock();
}

// Copies 2 x room column pointers into a device table on stream s through a pinned staging slot.
bool stage_table(pfsht_server_s &srv, void *const *cols, int room, void **dst, hipStream_t s) {
    const int slot = srv.tab_next;
    srv.tab_next = (slot + 1) % kTabSlots;
    if (srv.tab_used[slot] && hipEventSynchronize(srv.tab_ev[slot]) != hipSuccess) return false;
    void **h = srv.tab_host + size_t(slot) * 2 * kMaxCols;
    std::memcpy(h, cols, sizeof(void *) * 2 * room);
    if (hipMemcpyAsync(dst, h, sizeof(void *) * 2 * room, hipMemcpyHostToDevice, s) != hipSuccess) return false;
    if (hipEventRecord(srv.tab_ev[slot], s) != hipSuccess) return false;
    srv.tab_used[slot] = true;
    return true;
}

// Captures the launches of a column apply of nb transforms into a graph that reads its column table
// from g.tab. On failure the plan is applied eagerly from then on.
bool capture(pfsht_plan_t p, int dir, int nb, int room, hipStream_t s, Graph &g) {
    if (hipMalloc(reinterpret_cast<void **>(&g.tab), sizeof(void *) * 2 * room) != hipSuccess) {
        (void)hipGetLastError();
        g.tab = nullptr;
        return false;
    }
    InCols ic;
    ic.tab = reinterpret_cast<const double2 *const *>(g.tab);
    OutCols oc;
    oc.tab = reinterpret_cast<double2 *const *>(g.tab + room);
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        free_graph(g);
        return false;
    }
    const pfsht::Status st = pfsht::plan_launches(p, dir, ic, oc, nb, s);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &graph);
    bool ok = st == pfsht::kOk && ec == hipSuccess && graph;
    if (ok) ok = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
    if (!ok) {
        (void)hipGetLastError();
        free_graph(g);
    }
    return ok;
}

// One packed launch on stream s (plan mutex taken here; server mutex not held).
bool launch(pfsht_server_s &srv, Entry &e, int dir, const std::vector<Request> &reqs, hipStream_t s,
            std::vector<void *> &cols, bool &graphed) {
    const int64_t LB = int64_t(e.L + 1) * (2 * e.L + 1);
    const int64_t sin = dir == 0 ? LB : e.M, sout = dir == 0 ? e.M : LB;
    const int room = e.room;
    cols.assign(size_t(2) * room, nullptr);
    int nb = 0;
    for (const Request &r : reqs)
        for (int c = 0; c < e.ntrans; ++c, ++nb) {
            cols[nb] = const_cast<double2 *>(static_cast<const double2 *>(r.in) + c * sin);
            cols[room + nb] = static_cast<double2 *>(r.out) + c * sout;
        }
    std::lock_guard<std::mutex> plk(pfsht::plan_mutex(e.plan));
    for (const Request &r : reqs)
        if (hipEventRecord(srv.tickets[r.ticket % kTicketRing].start, s) != hipSuccess) return false;
    if (pfsht::plan_order_begin(e.plan, s) != pfsht::kOk) return false;
    graphed = false;
    if (srv.opt.graphs && !e.graph_failed) {
        Graph &g = e.graphs[dir * 1024 + nb];
        if (!g.exec && !capture(e.plan, dir, nb, room, s, g)) {
            e.graph_failed = true;
            e.graphs.erase(dir * 1024 + nb);
            srv.stats.graph_failures++;
        } else {
            if (!stage_table(srv, cols.data(), room, g.tab, s)) return false;
            if (hipGraphLaunch(g.exec, s) != hipSuccess) return false;
            graphed = true;
        }
    }
    if (!graphed) {
        // eager: the column table in a staging-owned device slot
        const int slot = srv.tab_next;
        void **dt = srv.tab_dev + size_t(slot) * 2 * kMaxCols;
        if (!stage_table(srv, cols.data(), room, dt, s)) return false;
        InCols ic;
        ic.tab = reinterpret_cast<const double2 *const *>(dt);
        OutCols oc;
        oc.tab = reinterpret_cast<double2 *const *>(dt + room);
        if (pfsht::plan_launches(e.plan, dir, ic, oc, nb, s) != pfsht::kOk) return false;
        // the slot's event was recorded before the launches: record it again after them
        if (hipEventRecord(srv.tab_ev[slot], s) != hipSuccess) return false;
    }
    if (pfsht::plan_order_end(e.plan, s) != pfsht::kOk) return false;
    for (const Request &r : reqs)
        if (hipEventRecord(srv.tickets[r.ticket % kTicketRing].end, s) != hipSuccess) return false;
    return true;
}

void dispatch_loop(pfsht_server_s *srv) {
    (void)hipSetDevice(srv->dev);
    std::unique_lock<std::mutex> lk(srv->mu);
    std::vector<void *> cols;
    for (;;) {
        bury(*srv, lk);
        // retire finished launches
        for (size_t i = 0; i < srv->streams.size(); ++i) {
            auto &q = srv->inflight[i];
            while (!q.empty() && hipEventQuery(q.front()) == hipSuccess) {
                srv->event_pool.push_back(q.front());
                q.pop_front();
            }
        }
        if (srv->queue.empty()) {
            if (srv->stop) break;
            srv->cv.wait_for(lk, std::chrono::milliseconds(2));
            continue;
        }
        // a stream with a free launch slot, round robin
        int si = -1;
        for (size_t k = 0; k < srv->streams.size(); ++k) {
            const size_t i = (srv->next_stream + k) % srv->streams.size();
            if (int(srv->inflight[i].size()) < srv->opt.depth) { si = int(i); break; }
        }
        if (si < 0) {
            srv->cv.wait_for(lk, std::chrono::microseconds(50));
            continue;
        }
        srv->next_stream = (size_t(si) + 1) % srv->streams.size();
        // the oldest request, and every later one of the same plan and direction that fits its room
        const Request r0 = srv->queue.front();
        auto it = srv->cache.find(r0.key);
        Entry &e = it->second;  // a plan with queued requests is never evicted
        std::vector<Request> pack;
        int used = 0;
        for (auto q = srv->queue.begin(); q != srv->queue.end();) {
            if (q->key == r0.key && q->dir == r0.dir && used + e.ntrans <= e.room) {
                pack.push_back(*q);
                used += e.ntrans;
                q = srv->queue.erase(q);
            } else {
                ++q;
            }
        }
        hipEvent_t done = take_event(*srv);
        hipStream_t s = srv->streams[size_t(si)];
        lk.unlock();
        bool graphed = false;
        bool ok = done && launch(*srv, e, r0.dir, pack, s, cols, graphed);
        if (ok) ok = hipEventRecord(done, s) == hipSuccess;
        lk.lock();
        if (!ok) {
            (void)hipGetLastError();
            srv->stats.failed += int64_t(pack.size());
        }
        if (done) {
            srv->inflight[size_t(si)].push_back(done);
            if (!e.last) (void)hipEventCreateWithFlags(&e.last, hipEventDisableTiming);
            if (e.last && hipEventRecord(e.last, s) == hipSuccess) e.has_last = true;
        }
        for (const Request &r : pack) srv->tickets[r.ticket % kTicketRing].state = kLaunched;
        e.pending -= int(pack.size());
        srv->stats.launches++;
        srv->stats.packed_transforms += used;
        if (graphed) srv->stats.graph_launches++;
        srv->cv_launched.notify_all();
    }
}

bool valid_ticket(const pfsht_server_s &srv, int64_t t) {
    return t >= 0 && t < srv.next_ticket && t >= srv.next_ticket - kTicketRing;
}

}  // namespace

pfsht_server_s::~pfsht_server_s() {
    {
        std::lock_guard<std::mutex> lk(mu);
        stop = true;
    }
    cv.notify_all();
    if (dispatcher.joinable()) dispatcher.join();
    for (auto s : streams) (void)hipStreamSynchronize(s);
    for (auto &kv : cache) {
        for (auto &g : kv.second.graphs) free_graph(g.second);
        if (kv.second.last) (void)hipEventDestroy(kv.second.last);
        pfsht::plan_release(kv.second.plan);
    }
    for (auto &g : dead_graphs) free_graph(g);
    for (auto p : graveyard) pfsht::plan_release(p);
    for (auto &t : tickets) {
        if (t.start) (void)hipEventDestroy(t.start);
        if (t.end) (void)hipEventDestroy(t.end);
    }
    for (auto &q : inflight)
        for (auto e : q) (void)hipEventDestroy(e);
    for (auto e : event_pool) (void)hipEventDestroy(e);
    for (auto e : tab_ev)
        if (e) (void)hipEventDestroy(e);
    if (tab_host) (void)hipHostFree(tab_host);
    if (tab_dev) (void)hipFree(tab_dev);
    if (epoch) (void)hipEventDestroy(epoch);
    for (auto s : streams) (void)hipStreamDestroy(s);
}

extern "C" {

void pfsht_server_opts_default(pfsht_server_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->nstreams = 4;
    o->depth = 2;
    o->pack = 8;
    o->graphs = 1;
    o->hbm_budget_bytes = int64_t(200) * 1000 * 1000 * 1000;
}

int pfsht_server_create_ex(const pfsht_server_opts *opts, pfsht_server_t *out) {
    if (!out || !opts) { pfsht::set_error("NULL argument"); return PFSHT_E_ARG; }
    *out = nullptr;
    const pfsht_server_opts o = *opts;
    if (o.nstreams < 1 || o.nstreams > 64) { pfsht::set_error("nstreams must lie in 1..64"); return PFSHT_E_ARG; }
    if (o.depth < 1 || o.depth > 16) { pfsht::set_error("depth must lie in 1..16"); return PFSHT_E_ARG; }
    if (o.pack < 1 || o.pack > kMaxCols) { pfsht::set_error("pack must lie in 1..64"); return PFSHT_E_ARG; }
    if (o.hbm_budget_bytes <= 0) { pfsht::set_error("the HBM budget must be positive"); return PFSHT_E_ARG; }
    auto *srv = new (std::nothrow) pfsht_server_s;
    if (!srv) { pfsht::set_error("out of host memory"); return PFSHT_E_NOMEM; }
    (void)hipGetDevice(&srv->dev);
    srv->opt = o;
    srv->stats.budget_bytes = o.hbm_budget_bytes;
    srv->tickets.resize(kTicketRing);
    srv->inflight.resize(size_t(o.nstreams));
    bool ok = hipHostMalloc(reinterpret_cast<void **>(&srv->tab_host), sizeof(void *) * 2 * kMaxCols * kTabSlots) ==
                  hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&srv->tab_dev), sizeof(void *) * 2 * kMaxCols * kTabSlots) ==
                  hipSuccess;
    for (auto &e : srv->tab_ev) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; ok && i < o.nstreams; ++i) {
        hipStream_t s;
        // default priority: every stream of the process runs at one priority
        ok = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess;
        if (ok) srv->streams.push_back(s);
    }
    if (!ok) {
        (void)hipGetLastError();
        delete srv;
        pfsht::set_error("creating the server's streams, events or column tables failed");
        return PFSHT_E_HIP;
    }
    srv->dispatcher = std::thread(dispatch_loop, srv);
    *out = srv;
    return PFSHT_OK;
}

int pfsht_server_create(int32_t nstreams, int64_t hbm_budget_bytes, pfsht_server_t *out) {
    pfsht_server_opts o;
    pfsht_server_opts_default(&o);
    o.nstreams = nstreams;
    o.hbm_budget_bytes = hbm_budget_bytes;
    return pfsht_server_create_ex(&o, out);
}

void pfsht_server_destroy(pfsht_server_t srv) { delete srv; }

int pfsht_server_plan(pfsht_server_t srv, uint64_t key, const double *theta, const double *phi, int64_t npts,
                      int32_t lmax, double tol, int32_t ntrans, int32_t method, pfsht_plan_t *plan, int32_t *hit) {
    if (!srv || !plan) { pfsht::set_error("NULL argument"); return PFSHT_E_ARG; }
    if (ntrans < 1 || ntrans > kMaxCols) { pfsht::set_error("ntrans must lie in 1..64 for a served plan"); return PFSHT_E_ARG; }
    std::unique_lock<std::mutex> lk(srv->mu);
    for (;;) {
        auto it = srv->cache.find(key);
        if (it == srv->cache.end()) break;
        if (it->second.building) {  // another caller is building this key's plan
            srv->cv_launched.wait(lk);
            continue;
        }
        it->second.last_use = ++srv->tick;
        srv->stats.hits++;
        *plan = it->second.plan;
        if (hit) *hit = 1;
        return PFSHT_OK;
    }
    if (!theta || !phi) { pfsht::set_error("theta/phi is NULL"); return PFSHT_E_ARG; }
    if (hit) *hit = 0;
    srv->stats.misses++;
    pfsht_plan_opts o;
    pfsht_plan_opts_default(&o);
    o.method = method;
    o.precision = tol >= 1e-6 ? PFSHT_PREC_MIXED : PFSHT_PREC_FP64;  // FP32 grids where the tolerance allows
    o.maxboxes = pfsht::kMaxBoxes;
    o.oversample_theta = o.oversample_phi = 1.0;
    o.epsapo_max = pfsht::kEpsApoMax;
    o.packed = 1;
    const int room = std::max<int>(ntrans, srv->opt.pack);
    // admission before any allocation: the plan's estimated size must fit next to the plans in use
    const int64_t est = pfsht::plan_estimate_bytes(theta, phi, npts, lmax, tol, room, o);
    if (est < 0 || est > srv->opt.hbm_budget_bytes || !make_room(*srv, est)) {
        srv->stats.rejected++;
        pfsht::set_error(est < 0 ? "the plan's size could not be estimated (bad arguments)"
                                 : "the plan does not fit in the HBM budget next to the plans in use");
        return est < 0 ? PFSHT_E_ARG : PFSHT_E_NOMEM;
    }
    Entry &slot = srv->cache[key];
    slot.building = true;
    slot.bytes = est;
    srv->resident += est;
    bury(*srv, lk);
    lk.unlock();
    pfsht_plan_t p = nullptr;
    int rc = pfsht_plan_create_opts(theta, phi, npts, lmax, tol, room, &o, &p);
    pfsht_info info{};
    if (rc == PFSHT_OK) rc = pfsht_plan_info(p, &info);
    lk.lock();
    Entry &e = srv->cache[key];
    srv->resident -= e.bytes;
    if (rc != PFSHT_OK) {
        srv->cache.erase(key);
        srv->cv_launched.notify_all();
        if (p) pfsht::plan_release(p);
        return rc;
    }
    e.plan = p;
    e.bytes = info.device_bytes;
    e.ntrans = ntrans;
    e.room = room;
    e.L = lmax;
    e.M = npts;
    e.building = false;
    e.last_use = ++srv->tick;
    srv->resident += e.bytes;
    make_room(*srv, 0);  // the estimate was an upper bound; evict further only if the plan came out larger
    srv->cv_launched.notify_all();
    *plan = p;
    return PFSHT_OK;
}

int pfsht_server_submit(pfsht_server_t srv, uint64_t key, int32_t dir, const void *in, void *out, int64_t *ticket) {
    if (!srv || !in || !out || !ticket) { pfsht::set_error("NULL argument"); return PFSHT_E_ARG; }
    if (dir != 0 && dir != 1) { pfsht::set_error("dir must be 0 (synthesis) or 1 (adjoint)"); return PFSHT_E_ARG; }
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->cache.find(key);
    if (it == srv->cache.end() || it->second.building) {
        pfsht::set_error("no plan is cached under this key");
        return PFSHT_E_ARG;
    }
    const int64_t id = srv->next_ticket;
    Ticket &t = srv->tickets[id % kTicketRing];
    if (t.state == kQueued || (t.state == kLaunched && hipEventQuery(t.end) != hipSuccess)) {
        pfsht::set_error("too many requests in flight");
        return PFSHT_E_UNSUPPORTED;
    }
    if (!t.start && hipEventCreate(&t.start) != hipSuccess) { pfsht::set_error("hipEventCreate failed"); return PFSHT_E_HIP; }
    if (!t.end && hipEventCreate(&t.end) != hipSuccess) { pfsht::set_error("hipEventCreate failed"); return PFSHT_E_HIP; }
    Entry &e = it->second;
    e.last_use = ++srv->tick;
    e.pending++;
    t.state = kQueued;
    srv->queue.push_back({id, key, dir, in, out});
    srv->next_ticket++;
    srv->stats.submitted++;
    *ticket = id;
    srv->cv.notify_one();
    return PFSHT_OK;
}

int pfsht_server_query(pfsht_server_t srv, int64_t ticket, int32_t *done, double *device_ms) {
    if (!srv || !done) { pfsht::set_error("NULL argument"); return PFSHT_E_ARG; }
    std::lock_guard<std::mutex> lk(srv->mu);
    if (!valid_ticket(*srv, ticket)) { pfsht::set_error("unknown or recycled ticket"); return PFSHT_E_ARG; }
    Ticket &t = srv->tickets[ticket % kTicketRing];
    if (t.state != kLaunched) {
        *done = 0;
        return PFSHT_OK;
    }
    const hipError_t q = hipEventQuery(t.end);
    if (q == hipErrorNotReady) {
        *done = 0;
        return PFSHT_OK;
    }
    if (q != hipSuccess) { pfsht::set_error("hipEventQuery failed"); return PFSHT_E_HIP; }
    *done = 1;
    if (device_ms) {
        float ms = 0;
        *device_ms = hipEventElapsedTime(&ms, t.start, t.end) == hipSuccess ? double(ms) : -1.0;
    }
    return PFSHT_OK;
}

// Waits until the ticket's launch has been issued and returns its end event (mutex released after).
static int launched_end(pfsht_server_t srv, int64_t ticket, hipEvent_t &start, hipEvent_t &end) {
    std::unique_lock<std::mutex> lk(srv->mu);
    if (!valid_ticket(*srv, ticket)) { pfsht::set_error("unknown or recycled ticket"); return PFSHT_E_ARG; }
    Ticket &t = srv->tickets[ticket % kTicketRing];
    srv->cv_launched.wait(lk, [&] { return t.state == kLaunched; });
    start = t.start;
    end = t.end;
    return PFSHT_OK;
}

int pfsht_server_wait(pfsht_server_t srv, int64_t ticket, double *device_ms) {
    if (!srv) { pfsht::set_error("NULL argument"); return PFSHT_E_ARG; }
    hipEvent_t start = nullptr, end = nullptr;
    const int rc = launched_end(srv, ticket, start, end);
    if (rc != PFSHT_OK) return rc;
    if (hipEventSynchronize(end) != hipSuccess) { pfsht::set_error("hipEventSynchronize failed"); return PFSHT_E_HIP; }
    int32_t done = 0;
    return pfsht_server_query(srv, ticket, &done, device_ms);
}

int pfsht_server_epoch(pfsht_server_t srv) {
    if (!srv) { pfsht::set_error("NULL argument"); return PFSHT_E_ARG; }
    std::lock_guard<std::mutex> lk(srv->mu);
    if (!srv->epoch && hipEventCreate(&srv->epoch) != hipSuccess) { pfsht::set_error("hipEventCreate failed"); return PFSHT_E_HIP; }
    if (hipEventRecord(srv->epoch, srv->streams[0]) != hipSuccess || hipEventSynchronize(srv->epoch) != hipSuccess) {
        pfsht::set_error("recording the epoch failed");
        return PFSHT_E_HIP;
    }
    return PFSHT_OK;
}

int pfsht_server_times(pfsht_server_t srv, int64_t ticket, double *start_ms, double *end_ms) {
    if (!srv || !start_ms || !end_ms) { pfsht::set_error("NULL argument"); return PFSHT_E_ARG; }
    hipEvent_t start = nullptr, end = nullptr, epoch = nullptr;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        epoch = srv->epoch;
    }
    if (!epoch) { pfsht::set_error("no epoch recorded (pfsht_server_epoch)"); return PFSHT_E_ARG; }
    const int rc = launched_end(srv, ticket, start, end);
    if (rc != PFSHT_OK) return rc;
    if (hipEventSynchronize(end) != hipSuccess) { pfsht::set_error("hipEventSynchronize failed"); return PFSHT_E_HIP; }
    float a = 0, b = 0;
    if (hipEventElapsedTime(&a, epoch, start) != hipSuccess || hipEventElapsedTime(&b, epoch, end) != hipSuccess) {
        pfsht::set_error("hipEventElapsedTime failed (was the epoch recorded before this ticket?)");
        return PFSHT_E_HIP;
    }
    *start_ms = a;
    *end_ms = b;
    return PFSHT_OK;
}

int pfsht_server_get_stats(pfsht_server_t srv, pfsht_server_stats *st) {
    if (!srv || !st) { pfsht::set_error("NULL argument"); return PFSHT_E_ARG; }
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stats.resident_bytes = srv->resident;
    srv->stats.plans = int64_t(srv->cache.size());
    srv->stats.queued = int64_t(srv->queue.size());
    *st = srv->stats;
    return PFSHT_OK;
}

}  // extern "C"
