// serve.hip — the serving runtime: a device-resident plan cache with HBM admission, and applies
// spread over a pool of HIP streams with per-request completion tickets.
//
// The reference has no counterpart: a PartialFSHT.Plan is built by its caller and applied by it
// (src/Plan.jl:84-148). A service that answers requests on recurring point sets keeps their plans
// resident instead: here plans are looked up by a caller-chosen 64-bit key, built on a miss, and
// evicted least-recently-used (only once their last apply has finished) when the resident plans
// would exceed the HBM budget. Applies of different plans run concurrently on different streams;
// applies of one plan are ordered by the plan itself (they share its work buffers).
#include <algorithm>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/pfsht.h"
#include "pfsht_common.hpp"

namespace {

struct Ticket {
    hipEvent_t start = nullptr, end = nullptr;
    bool used = false;
};

struct Entry {
    pfsht_plan_t plan = nullptr;
    int64_t bytes = 0;
    uint64_t last_use = 0;
    int64_t last_ticket = -1;  // most recent apply of this plan, -1 when none is in flight
};

}  // namespace

struct pfsht_server_s {
    int dev = 0;
    std::vector<hipStream_t> streams;
    size_t next_stream = 0;
    int64_t budget = 0, resident = 0;
    uint64_t tick = 0;
    std::unordered_map<uint64_t, Entry> cache;
    std::vector<Ticket> tickets;  // indexed by ticket id modulo capacity
    int64_t next_ticket = 0;
    hipEvent_t epoch = nullptr;  // reference for pfsht_server_times
    pfsht_server_stats stats{};
    std::mutex mu;

    ~pfsht_server_s() {
        (void)hipDeviceSynchronize();
        for (auto &kv : cache) pfsht_plan_destroy(kv.second.plan);
        for (auto &t : tickets) {
            if (t.start) (void)hipEventDestroy(t.start);
            if (t.end) (void)hipEventDestroy(t.end);
        }
        if (epoch) (void)hipEventDestroy(epoch);
        for (auto s : streams) (void)hipStreamDestroy(s);
    }
};

namespace {

constexpr int64_t kTicketRing = 1 << 16;  // tickets older than this many submissions are recycled

// The plan's last apply has completed (or it never ran): it may be evicted.
bool idle(pfsht_server_s &srv, const Entry &e) {
    if (e.last_ticket < 0) return true;
    const Ticket &t = srv.tickets[e.last_ticket % kTicketRing];
    return !t.used || hipEventQuery(t.end) == hipSuccess;
}

// Evicts least-recently-used idle plans until `need` more bytes fit in the budget.
bool make_room(pfsht_server_s &srv, int64_t need) {
    while (srv.resident + need > srv.budget) {
        auto victim = srv.cache.end();
        for (auto it = srv.cache.begin(); it != srv.cache.end(); ++it)
            if (idle(srv, it->second) && (victim == srv.cache.end() || it->second.last_use < victim->second.last_use))
                victim = it;
        if (victim == srv.cache.end()) return false;
        srv.resident -= victim->second.bytes;
        pfsht_plan_destroy(victim->second.plan);
        srv.cache.erase(victim);
        srv.stats.evictions++;
    }
    return true;
}

int64_t plan_bytes(pfsht_plan_t p) {
    pfsht_info info;
    return pfsht_plan_info(p, &info) == PFSHT_OK ? info.device_bytes : 0;
}

}  // namespace

extern "C" {

int pfsht_server_create(int32_t nstreams, int64_t hbm_budget_bytes, pfsht_server_t *out) {
    if (!out) { pfsht::set_error("out is NULL"); return PFSHT_E_ARG; }
    *out = nullptr;
    if (nstreams < 1 || nstreams > 64) { pfsht::set_error("nstreams must lie in 1..64"); return PFSHT_E_ARG; }
    if (hbm_budget_bytes <= 0) { pfsht::set_error("the HBM budget must be positive"); return PFSHT_E_ARG; }
    auto *srv = new (std::nothrow) pfsht_server_s;
    if (!srv) { pfsht::set_error("out of host memory"); return PFSHT_E_NOMEM; }
    (void)hipGetDevice(&srv->dev);
    srv->budget = hbm_budget_bytes;
    srv->stats.budget_bytes = hbm_budget_bytes;
    srv->tickets.resize(kTicketRing);
    for (int i = 0; i < nstreams; ++i) {
        hipStream_t s;
        // default priority: every stream of the process runs at one priority
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
            delete srv;
            pfsht::set_error("hipStreamCreateWithFlags failed");
            return PFSHT_E_HIP;
        }
        srv->streams.push_back(s);
    }
    *out = srv;
    return PFSHT_OK;
}

void pfsht_server_destroy(pfsht_server_t srv) { delete srv; }

int pfsht_server_plan(pfsht_server_t srv, uint64_t key, const double *theta, const double *phi, int64_t npts,
                      int32_t lmax, double tol, int32_t ntrans, int32_t method, pfsht_plan_t *plan, int32_t *hit) {
    if (!srv || !plan) { pfsht::set_error("NULL argument"); return PFSHT_E_ARG; }
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->cache.find(key);
    if (it != srv->cache.end()) {
        it->second.last_use = ++srv->tick;
        srv->stats.hits++;
        *plan = it->second.plan;
        if (hit) *hit = 1;
        return PFSHT_OK;
    }
    if (hit) *hit = 0;
    srv->stats.misses++;
    pfsht_plan_t p = nullptr;
    int rc = pfsht_plan_create(theta, phi, npts, lmax, tol, ntrans, method, PFSHT_PREC_FP64, &p);
    if (rc != PFSHT_OK) return rc;
    const int64_t b = plan_bytes(p);
    if (!make_room(*srv, b)) {
        pfsht_plan_destroy(p);
        srv->stats.rejected++;
        pfsht::set_error("the plan does not fit in the HBM budget next to the plans in use");
        return PFSHT_E_NOMEM;
    }
    Entry e;
    e.plan = p;
    e.bytes = b;
    e.last_use = ++srv->tick;
    srv->cache.emplace(key, e);
    srv->resident += b;
    *plan = p;
    return PFSHT_OK;
}

int pfsht_server_submit(pfsht_server_t srv, uint64_t key, int32_t dir, const void *in, void *out, int64_t *ticket) {
    if (!srv || !in || !out || !ticket) { pfsht::set_error("NULL argument"); return PFSHT_E_ARG; }
    if (dir != 0 && dir != 1) { pfsht::set_error("dir must be 0 (synthesis) or 1 (adjoint)"); return PFSHT_E_ARG; }
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->cache.find(key);
    if (it == srv->cache.end()) { pfsht::set_error("no plan is cached under this key"); return PFSHT_E_ARG; }
    Entry &e = it->second;
    e.last_use = ++srv->tick;
    const int64_t id = srv->next_ticket;
    Ticket &t = srv->tickets[id % kTicketRing];
    if (t.used && hipEventQuery(t.end) != hipSuccess) {
        pfsht::set_error("too many applies in flight");
        return PFSHT_E_UNSUPPORTED;
    }
    if (!t.start && hipEventCreate(&t.start) != hipSuccess) { pfsht::set_error("hipEventCreate failed"); return PFSHT_E_HIP; }
    if (!t.end && hipEventCreate(&t.end) != hipSuccess) { pfsht::set_error("hipEventCreate failed"); return PFSHT_E_HIP; }
    hipStream_t s = srv->streams[srv->next_stream];
    srv->next_stream = (srv->next_stream + 1) % srv->streams.size();
    if (hipEventRecord(t.start, s) != hipSuccess) { pfsht::set_error("hipEventRecord failed"); return PFSHT_E_HIP; }
    const int rc = dir == 0 ? pfsht_synth(e.plan, in, out, PFSHT_WHERE_DEVICE, s)
                            : pfsht_adjoint(e.plan, in, out, PFSHT_WHERE_DEVICE, s);
    if (rc != PFSHT_OK) return rc;
    if (hipEventRecord(t.end, s) != hipSuccess) { pfsht::set_error("hipEventRecord failed"); return PFSHT_E_HIP; }
    t.used = true;
    e.last_ticket = id;
    srv->next_ticket++;
    srv->stats.submitted++;
    *ticket = id;
    return PFSHT_OK;
}

int pfsht_server_query(pfsht_server_t srv, int64_t ticket, int32_t *done, double *device_ms) {
    if (!srv || !done) { pfsht::set_error("NULL argument"); return PFSHT_E_ARG; }
    std::lock_guard<std::mutex> lk(srv->mu);
    if (ticket < 0 || ticket >= srv->next_ticket || ticket < srv->next_ticket - kTicketRing) {
        pfsht::set_error("unknown or recycled ticket");
        return PFSHT_E_ARG;
    }
    Ticket &t = srv->tickets[ticket % kTicketRing];
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

int pfsht_server_wait(pfsht_server_t srv, int64_t ticket, double *device_ms) {
    if (!srv) { pfsht::set_error("NULL argument"); return PFSHT_E_ARG; }
    hipEvent_t end = nullptr;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (ticket < 0 || ticket >= srv->next_ticket || ticket < srv->next_ticket - kTicketRing) {
            pfsht::set_error("unknown or recycled ticket");
            return PFSHT_E_ARG;
        }
        end = srv->tickets[ticket % kTicketRing].end;
    }
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
    std::lock_guard<std::mutex> lk(srv->mu);
    if (!srv->epoch) { pfsht::set_error("no epoch recorded (pfsht_server_epoch)"); return PFSHT_E_ARG; }
    if (ticket < 0 || ticket >= srv->next_ticket || ticket < srv->next_ticket - kTicketRing) {
        pfsht::set_error("unknown or recycled ticket");
        return PFSHT_E_ARG;
    }
    Ticket &t = srv->tickets[ticket % kTicketRing];
    if (hipEventSynchronize(t.end) != hipSuccess) { pfsht::set_error("hipEventSynchronize failed"); return PFSHT_E_HIP; }
    float a = 0, b = 0;
    if (hipEventElapsedTime(&a, srv->epoch, t.start) != hipSuccess ||
        hipEventElapsedTime(&b, srv->epoch, t.end) != hipSuccess) {
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
    *st = srv->stats;
    return PFSHT_OK;
}

}  // extern "C"
