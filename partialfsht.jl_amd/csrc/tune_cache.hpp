// tune_cache.hpp — the ring-pass shapes the plan-time tuner chose, per engine shape and device.
//
// A ring engine's launch shape (degree-segment length, 64-ring tiles per wavefront of the VALU
// adjoint, tiles per wavefront of the MFMA adjoint) is chosen by timing candidates on the engine a
// plan builds; the choice is kept here under (device architecture, lmax, rings, largest order,
// transforms), so later plans of that shape reuse it. Host code only: no HIP calls.
#pragma once

#include <cstdint>
#include <string>

namespace pfsht {

struct RingShape {
    int per = 0;      // target of 64-degree blocks per degree segment (0: the default)
    int adj_tg = 0;   // 64-ring tiles per wavefront of the VALU adjoint (0: the default rule)
    int mfma_tg = 0;  // 64-ring tiles per wavefront of the MFMA adjoint (0: the default rule)
    double ms = 0;    // ring forward + adjoint time of the shape when it was timed
};

struct TuneKey {
    std::string arch;  // device architecture and compute-unit count, e.g. "gfx950:256"
    int L = 0, n = 0, mmax = 0, B = 0;
};

// The cached shape of key, if any.
bool tune_lookup(const TuneKey &key, RingShape &out);
// Record the shape of key (replacing an earlier one). When PFSHT_TUNE_CACHE names a file, the cache
// is written back to it.
void tune_store(const TuneKey &key, const RingShape &shape);
// Merge the entries of PFSHT_TUNE_CACHE, once per process (no-op without the variable).
void tune_load_env();

}  // namespace pfsht
