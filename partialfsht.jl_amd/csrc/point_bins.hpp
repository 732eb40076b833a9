// point_bins.hpp — the map of a point onto the NUFFT fine grid and its bin key, and the bin ranges
// of the sorted points, shared by the plan-construction kernel (one thread per point) and host code
// (the tile lists of a plan, and the CPU tests of both).
//
// A point (theta, phi) becomes fine-grid coordinates (x1, x2): theta through theta_to_stretch
// (src/Windows.jl:307-320) scaled by N1 / 2 pi, phi mod 2 pi scaled by N2 / 2 pi, both wrapped into
// [0, N). Its kernel's first cell along each axis is b = ceil(x - w/2); the key is
// (super-tile, local bin) with bins of S x S base cells and super-tiles of 4 x 4 bins. A NaN or
// infinite coordinate, or one whose first cell falls off the grid, is refused: the plan is not built.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "geometry.hpp"

namespace pfsht {

struct PointMap {  // theta -> fine-grid row coordinate, phi -> column coordinate, and the bin keys
    int kind, flip, north;  // geometry kind; southern reflection; pole cap at the north pole
    double thmax;           // pole: h (1 + epsapo)
    double lo, period;      // interior: thlo - h epsapo, and the period
    double s1, s2;          // N1 / (2 pi), N2 / (2 pi)
    double half;            // w / 2
    int N1, N2, S, S4, ns2, w;
};

// Maps one point; false when it is refused (x1 = x2 = 0 and key = 0 then, and the caller must
// refuse the plan: a refused point is in no bin's count).
__host__ __device__ inline bool map_point(const PointMap &q, double t0, double phi, double &x1, double &x2,
                                          int &key, int &b1) {
#pragma clang fp contract(off)
    const double t = q.flip ? kPi - t0 : t0;
    double T;
    if (q.kind == kFull) T = t;
    else if (q.kind == kPole) T = kPi * (q.north ? t : kPi - t) / q.thmax;
    else T = 2 * kPi * (t - q.lo) / q.period;
    double ph = fmod(phi, 2 * kPi);
    if (ph != 0 && ph < 0) ph += 2 * kPi;
    x1 = T * q.s1;
    x2 = ph * q.s2;
    x1 -= floor(x1 / q.N1) * q.N1;
    x2 -= floor(x2 / q.N2) * q.N2;
    // ceil of a NaN or an infinity is not converted to int: test the coordinates first
    const bool finite = x1 == x1 && x2 == x2 && fabs(x1) <= double(q.N1) && fabs(x2) <= double(q.N2);
    b1 = finite ? int(ceil(x1 - q.half)) : -1;
    int b2 = finite ? int(ceil(x2 - q.half)) : -1;
    if (finite && b1 < 0) { x1 += q.N1; b1 = int(ceil(x1 - q.half)); }
    if (finite && b2 < 0) { x2 += q.N2; b2 = int(ceil(x2 - q.half)); }
    if (!finite || b1 < 0 || b1 >= q.N1 || b2 < 0 || b2 >= q.N2) {
        x1 = x2 = 0.0;
        key = 0;
        return false;
    }
    const int s1 = b1 / q.S4, s2 = b2 / q.S4;
    const int lb = ((b1 % q.S4) / q.S) * 4 + (b2 % q.S4) / q.S;
    key = (s1 * q.ns2 + s2) * 16 + lb;
    return true;
}

// Prefix sums of the per-key counts: cnt[k] is the first sorted position of key k and cnt[nkeys]
// their total. Every point of an accepted plan is in exactly one key, so the total must equal the
// number of points; otherwise some bin range would reach past the sorted arrays (or leave points
// out), and the plan is refused here instead of being launched.
inline bool bin_ranges(const std::vector<int32_t> &cnt32, int64_t npts, std::vector<int64_t> &cnt, std::string &msg) {
    cnt.assign(cnt32.size() + 1, 0);
    for (size_t k = 0; k < cnt32.size(); ++k) {
        if (cnt32[k] < 0) {
            msg = "negative bin count";
            return false;
        }
        cnt[k + 1] = cnt[k] + cnt32[k];
    }
    if (cnt.back() != npts) {
        msg = "the bin counts cover " + std::to_string(cnt.back()) + " of " + std::to_string(npts) + " points";
        return false;
    }
    return true;
}

}  // namespace pfsht
