// tune_cache.cpp — the process-wide cache of tuned ring-pass shapes, and its C ABI
// (pfsht_tune_*). One text line per entry: "arch L n mmax B per adj_tg mfma_tg ms".
#include "tune_cache.hpp"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

#include "../../include/pfsht.h"

namespace pfsht {

namespace {

using Key = std::tuple<std::string, int, int, int, int>;

std::mutex g_mu;
std::map<Key, RingShape> g_cache;
bool g_env_loaded = false;

Key key_of(const TuneKey &k) { return Key(k.arch, k.L, k.n, k.mmax, k.B); }

bool valid_shape(const RingShape &s) {
    return s.per >= 0 && s.per <= 64 && s.adj_tg >= 0 && s.adj_tg <= 2 && s.mfma_tg >= 0 && s.mfma_tg <= 2;
}

// merge a cache file; false when it cannot be read or a line is malformed
bool load_locked(const char *path) {
    FILE *f = std::fopen(path, "r");
    if (!f) return false;
    char arch[128];
    int L, n, mmax, B, per, atg, mtg;
    double ms;
    bool ok = true;
    for (;;) {
        const int got = std::fscanf(f, "%127s %d %d %d %d %d %d %d %lf", arch, &L, &n, &mmax, &B, &per, &atg, &mtg, &ms);
        if (got == EOF) break;
        if (got != 9) { ok = false; break; }
        RingShape s;
        s.per = per; s.adj_tg = atg; s.mfma_tg = mtg; s.ms = ms;
        if (!valid_shape(s)) { ok = false; break; }
        g_cache[Key(arch, L, n, mmax, B)] = s;
    }
    std::fclose(f);
    return ok;
}

bool save_locked(const char *path) {
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "w");
    if (!f) return false;
    for (const auto &kv : g_cache) {
        const auto &[arch, L, n, mmax, B] = kv.first;
        const RingShape &s = kv.second;
        std::fprintf(f, "%s %d %d %d %d %d %d %d %.6g\n", arch.c_str(), L, n, mmax, B, s.per, s.adj_tg, s.mfma_tg, s.ms);
    }
    const bool ok = std::fclose(f) == 0;
    return ok && std::rename(tmp.c_str(), path) == 0;
}

}  // namespace

bool tune_lookup(const TuneKey &key, RingShape &out) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find(key_of(key));
    if (it == g_cache.end()) return false;
    out = it->second;
    return true;
}

void tune_store(const TuneKey &key, const RingShape &shape) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[key_of(key)] = shape;
    if (const char *p = std::getenv("PFSHT_TUNE_CACHE"))
        if (*p) (void)save_locked(p);
}

void tune_load_env() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_env_loaded) return;
    g_env_loaded = true;
    if (const char *p = std::getenv("PFSHT_TUNE_CACHE"))
        if (*p) (void)load_locked(p);
}

}  // namespace pfsht

using namespace pfsht;

static TuneKey make_key(const char *arch, int32_t lmax, int32_t nrings, int32_t mmax, int32_t ntrans) {
    TuneKey k;
    k.arch = arch ? arch : "";
    k.L = lmax; k.n = nrings; k.mmax = mmax; k.B = ntrans;
    return k;
}

extern "C" {

int pfsht_tune_get(const char *arch, int32_t lmax, int32_t nrings, int32_t mmax, int32_t ntrans,
                   pfsht_ring_shape *shape, int32_t *found) {
    if (!arch || !shape || !found) return PFSHT_E_ARG;
    RingShape s;
    *found = tune_lookup(make_key(arch, lmax, nrings, mmax, ntrans), s) ? 1 : 0;
    shape->seg_blocks = s.per;
    shape->adj_tiles = s.adj_tg;
    shape->mfma_tiles = s.mfma_tg;
    shape->reserved = 0;
    shape->ms = s.ms;
    return PFSHT_OK;
}

int pfsht_tune_put(const char *arch, int32_t lmax, int32_t nrings, int32_t mmax, int32_t ntrans,
                   const pfsht_ring_shape *shape) {
    if (!arch || !shape || !*arch) return PFSHT_E_ARG;
    RingShape s;
    s.per = shape->seg_blocks; s.adj_tg = shape->adj_tiles; s.mfma_tg = shape->mfma_tiles; s.ms = shape->ms;
    if (!valid_shape(s)) return PFSHT_E_ARG;
    tune_store(make_key(arch, lmax, nrings, mmax, ntrans), s);
    return PFSHT_OK;
}

int pfsht_tune_load(const char *path) {
    if (!path) return PFSHT_E_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    return load_locked(path) ? PFSHT_OK : PFSHT_E_ARG;
}

int pfsht_tune_save(const char *path) {
    if (!path) return PFSHT_E_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    return save_locked(path) ? PFSHT_OK : PFSHT_E_ARG;
}

void pfsht_tune_clear(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache.clear();
}

int64_t pfsht_tune_size(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    return int64_t(g_cache.size());
}

}  // extern "C"
