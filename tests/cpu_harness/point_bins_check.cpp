// Host-side check of point_bins.hpp, the code the plan-construction kernel runs per point and the
// plan build runs on the bin counts. Reads a command and its data on stdin, writes results on stdout.
//   map    N1 N2 S w kind flip north thmax lo period  npts  (theta phi)*npts
//          -> per point: ok key b1 x1 x2
//   ranges npts nkeys count*nkeys
//          -> "ok" and the nkeys + 1 prefix sums, or "refused <message>"
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "../../partialfsht.jl_amd/csrc/point_bins.hpp"

int main() {
    std::string cmd;
    std::cin >> cmd;
    if (cmd == "map") {
        pfsht::PointMap q{};
        std::string a, b, c;  // doubles as text (hex floats included), read through strtod
        std::cin >> q.N1 >> q.N2 >> q.S >> q.w >> q.kind >> q.flip >> q.north >> a >> b >> c;
        q.thmax = std::stod(a);
        q.lo = std::stod(b);
        q.period = std::stod(c);
        q.S4 = 4 * q.S;
        q.ns2 = (q.N2 + q.S4 - 1) / q.S4;
        q.s1 = double(q.N1) / (2 * pfsht::kPi);
        q.s2 = double(q.N2) / (2 * pfsht::kPi);
        q.half = 0.5 * q.w;
        long n;
        std::cin >> n;
        for (long i = 0; i < n; ++i) {
            std::string ts, ps;
            std::cin >> ts >> ps;
            const double t = std::stod(ts), p = std::stod(ps);
            double x1, x2;
            int key, b1;
            const bool ok = pfsht::map_point(q, t, p, x1, x2, key, b1);
            std::printf("%d %d %d %.17g %.17g\n", ok ? 1 : 0, key, b1, x1, x2);
        }
        return 0;
    }
    if (cmd == "ranges") {
        long npts, nkeys;
        std::cin >> npts >> nkeys;
        std::vector<int32_t> c(nkeys);
        for (auto &v : c) std::cin >> v;
        std::vector<int64_t> cnt;
        std::string msg;
        if (!pfsht::bin_ranges(c, npts, cnt, msg)) {
            std::printf("refused %s\n", msg.c_str());
            return 0;
        }
        std::printf("ok");
        for (auto v : cnt) std::printf(" %lld", (long long)v);
        std::printf("\n");
        return 0;
    }
    return 2;
}
