// geometry.cpp — see geometry.hpp. Formulas follow the cited reference lines; arithmetic is kept
// in the reference's evaluation order so plan sizes (nrings, nph, mmax, ring_work) match it
// exactly.
#include "geometry.hpp"

#include <algorithm>
#include <cstdio>
#include <cfloat>
#include <climits>
#include <cmath>
#include <memory>
#include <thread>

namespace pfsht {

int polar_mmax(double theta, int L, double eps) {
    if (L <= 0) return 0;
    const double s = std::fabs(std::sin(theta));
    const double z = (double(L) + 0.5) * s;
    const double lne = std::fabs(std::log(std::max(eps, DBL_MIN)));
    const double pad = 0.5 * std::cbrt(9.0) * std::cbrt(std::max(z, 1.0)) * std::cbrt(lne * lne) + 2;
    const double v = std::ceil(z + pad);
    if (v <= 0) return 0;
    return v >= double(L) ? L : int(v);
}

void sectorial_log2c(int L, std::vector<double> &v) {
    v.assign(L + 1, 0.0);
    double acc = -std::log2(std::sqrt(4 * kPi));
    v[0] = acc;
    for (int m = 1; m <= L; ++m) {
        acc += std::log2(std::sqrt(double(2 * m + 1) / double(2 * m)));
        v[m] = acc;
    }
}

bool nested_ranges(const std::vector<int> &mlim, int L, std::vector<int> &jlo, std::vector<int> &jhi,
                   std::string &msg) {
    const int n = int(mlim.size());
    if (n < 1) { msg = "need at least one ring"; return false; }
    int peak = 0;
    for (int j = 1; j < n; ++j)
        if (mlim[j] > mlim[peak]) peak = j;
    for (int j = 1; j <= peak; ++j)
        if (mlim[j] < mlim[j - 1]) { msg = "retained-order limits are not unimodal in colatitude"; return false; }
    for (int j = peak + 1; j < n; ++j)
        if (mlim[j] > mlim[j - 1]) { msg = "retained-order limits are not unimodal in colatitude"; return false; }
    jlo.assign(L + 1, 0);
    jhi.assign(L + 1, -1);
    int lo = peak, hi = peak;
    const int mtop = std::min(mlim[peak], L);
    for (int m = mtop; m >= 0; --m) {
        while (lo > 0 && mlim[lo - 1] >= m) --lo;
        while (hi < n - 1 && mlim[hi + 1] >= m) ++hi;
        jlo[m] = lo;
        jhi[m] = hi;
    }
    return true;
}

// 10-point Gauss–Legendre on [-1, 1] and the panel width of the composite rule
// (src/Windows.jl:187-197).
static const double kGLx[10] = {-0.9739065285171717, -0.8650633666889845, -0.6794095682990244,
                                -0.4333953941292472, -0.1488743389816312, 0.1488743389816312,
                                0.4333953941292472,  0.6794095682990244,  0.8650633666889845,
                                0.9739065285171717};
static const double kGLw[10] = {0.0666713443086881, 0.1494513491505806, 0.2190863625159820,
                                0.2692667193099963, 0.2955242247147529, 0.2955242247147529,
                                0.2692667193099963, 0.2190863625159820, 0.1494513491505806,
                                0.0666713443086881};

double es_cdf(double beta, double u) {
    const double a = -kPi / 2;
    if (u <= a) return 0.0;
    const double b = std::min(u, kPi / 2);
    const int n = std::max(1, int(std::ceil((b - a) / 0.05)));
    const double hw = (b - a) / (2 * n);
    double acc = 0;
    for (int p = 0; p < n; ++p) {
        const double c = a + (2 * p + 1) * hw;
        double s = 0;
        for (int q = 0; q < 10; ++q) {
            const double t = c + hw * kGLx[q];
            const double ct = std::cos(t);
            s += kGLw[q] * std::exp(beta * (ct - 1)) * ct;
        }
        acc += hw * s;
    }
    return acc;
}

int stretch_geometry(double thlo, double thhi, int L, double eps, Geom &out, std::string &msg, double epsapo_max) {
    const double lo = thlo, hi = thhi;
    if (!(lo <= hi)) { msg = "thlo exceeds thhi"; return -1; }
    if (!(lo >= 0 && hi <= kPi)) { msg = "colatitudes must lie in [0, pi]"; return -1; }
    const double beta = std::fabs(std::log(eps));
    const double tiny = std::sqrt(DBL_EPSILON);
    const double Z = es_cdf(beta, kPi / 2);
    const double invZ = Z > 0 ? 1.0 / Z : 0.0;
    if (lo <= tiny && hi >= kPi - tiny) {
        Geom g;
        g.kind = kFull; g.lmax = L; g.K = L; g.eps = eps; g.beta = beta; g.h = kPi / 2;
        g.thlo = 0; g.thhi = kPi; g.epsapo = 0; g.period = 2 * kPi; g.Theta1 = kPi; g.sigma = 0;
        g.ThetaT = kPi; g.invZ = 1.0;
        out = g;
        return 0;
    }
    struct Cand { int kind; double h, klo, khi; };
    const Cand cands[3] = {{kPole, hi, 0.0, hi}, {kPole, kPi - lo, lo, kPi}, {kInterior, (hi - lo) / 2, lo, hi}};
    int bestK = INT_MAX;
    bool found = false;
    for (const Cand &c : cands) {
        const double h = c.h;
        if (!(h > 0)) continue;
        const double Lh = double(L) * h;
        if (!(Lh >= 2 * beta / (epsapo_max * epsapo_max))) continue;
        const double ea = std::sqrt(2 * beta / Lh);
        const double pad = h * ea;
        if (c.kind == kInterior) {
            if (!(c.klo - pad >= 0 && c.khi + pad <= kPi)) continue;
        } else if (!(h * (1 + ea) <= kPi)) {
            continue;
        }
        const double kk = std::ceil(Lh * ((1 + ea) * (1 + ea)) / kPi);
        const int K = kk < 1 ? 1 : int(kk);
        if (K < bestK) {
            bestK = K;
            Geom g;
            g.kind = c.kind; g.lmax = L; g.K = K; g.eps = eps; g.beta = beta; g.h = h;
            g.thlo = c.klo; g.thhi = c.khi; g.epsapo = ea; g.period = 2 * h * (1 + ea);
            g.Theta1 = kPi / (1 + ea); g.sigma = (kPi - g.Theta1) / 2; g.ThetaT = g.Theta1 + g.sigma;
            g.invZ = invZ;
            out = g;
            found = true;
        }
    }
    if (!found) {
        char buf[320];
        std::snprintf(buf, sizeof buf,
                      "no stretched geometry for colatitudes [%.6g, %.6g] at lmax %d and tol %.3g: each candidate "
                      "needs an apodisation fraction above epsapo_max = %.3g or a pad past a pole; the Direct "
                      "leaf serves such point sets",
                      lo, hi, L, eps, epsapo_max);
        msg = buf;
        return 1;
    }
    return 0;
}

double stretch_window(const Geom &g, double Theta) {
    if (g.kind == kFull) return 1.0;
    const double d = g.kind == kPole ? std::fabs(Theta) : std::fabs(Theta - kPi);
    if (d <= g.Theta1) return 1.0;
    if (d >= kPi) return 0.0;
    const double sg = g.sigma;
    if (!(sg > 0)) return d <= g.ThetaT ? 1.0 : 0.0;
    const double zlo = std::clamp((d - g.ThetaT) / sg, -1.0, 1.0);
    const double zhi = std::clamp((d + g.ThetaT) / sg, -1.0, 1.0);
    const double v = (es_cdf(g.beta, std::asin(zhi)) - es_cdf(g.beta, std::asin(zlo))) * g.invZ;
    return std::clamp(v, 0.0, 1.0);
}

double stretch_to_theta(const Geom &g, double Theta) {
    if (g.kind == kFull) return Theta;
    if (g.kind == kPole) {
        const double thmax = g.h * (1 + g.epsapo);
        const double t = Theta * thmax / kPi;
        return g.thlo <= std::sqrt(DBL_EPSILON) ? t : kPi - t;
    }
    const double lo = g.thlo - g.h * g.epsapo;
    return lo + g.period * Theta / (2 * kPi);
}

double theta_to_stretch(const Geom &g, double theta) {
    if (g.kind == kFull) return theta;
    if (g.kind == kPole) {
        const double thmax = g.h * (1 + g.epsapo);
        const double d = g.thlo <= std::sqrt(DBL_EPSILON) ? theta : kPi - theta;
        return kPi * d / thmax;
    }
    const double lo = g.thlo - g.h * g.epsapo;
    return 2 * kPi * (theta - lo) / g.period;
}

void stretch_rings(const Geom &g, int n, std::vector<double> &th, std::vector<double> &Th) {
    th.resize(n);
    Th.resize(n);
    for (int i = 1; i <= n; ++i) {
        const double T = g.kind == kInterior ? 2 * kPi * double(i - 1) / n : kPi * (double(i) - 0.5) / n;
        Th[i - 1] = T;
        th[i - 1] = stretch_to_theta(g, T);
    }
}

int64_t next_smooth(int64_t n) {
    if (n <= 1) return 1;
    int64_t best = INT64_MAX;
    for (int64_t p2 = 1; p2 < 2 * n + 2; p2 *= 2)
        for (int64_t p3 = p2; p3 < 2 * n + 2; p3 *= 3)
            for (int64_t p5 = p3; p5 < 2 * n + 2; p5 *= 5)
                if (p5 >= n && p5 < best) best = p5;
    return best == INT64_MAX ? n : best;
}

int64_t next_smooth_even(int64_t n) { return 2 * next_smooth((n + 1) / 2); }

double jmod(double x, double y) {
    double r = std::fmod(x, y);
    if (r != 0 && ((r < 0) != (y < 0))) r += y;
    return r;
}

Box bounding_box(const double *theta, const double *phi, int64_t n) {
    double lo = INFINITY, hi = -INFINITY;
    for (int64_t i = 0; i < n; ++i) { lo = std::min(lo, theta[i]); hi = std::max(hi, theta[i]); }
    return bounding_box_range(lo, hi, phi, n);
}

Box bounding_box_range(double lo, double hi, const double *phi, int64_t n) {
    Box b;
    b.thlo = std::max(lo, 0.0);
    b.thhi = std::min(hi, kPi);
    const double ptol = std::sqrt(2.220446049250313e-16);
    b.north = b.thlo <= ptol;
    b.south = b.thhi >= kPi - ptol;
    if (b.north || b.south) { b.phistart = 0.0; b.phiext = 2 * kPi; return b; }
    std::vector<double> ps(n);
    for (int64_t i = 0; i < n; ++i) ps[i] = jmod(phi[i], 2 * kPi);
    if (n == 1) { b.phistart = ps[0]; b.phiext = 0.0; return b; }
    // largest gap between consecutive values: n - 1 equal buckets hold it between two of them
    double plo = ps[0], phi_ = ps[0];
    for (int64_t i = 1; i < n; ++i) { plo = std::min(plo, ps[i]); phi_ = std::max(phi_, ps[i]); }
    double gap = 0.0, at = plo;
    if (phi_ > plo) {
        const int64_t nb = n - 1;
        const double w = (phi_ - plo) / double(nb);
        std::vector<std::pair<double, double>> bk(nb, {INFINITY, -INFINITY});  // (min, max) per bucket
        for (int64_t i = 0; i < n; ++i) {
            int64_t k = int64_t(std::floor((ps[i] - plo) / w));
            k = std::min(std::max(k, int64_t(0)), nb - 1);
            bk[k].first = std::min(bk[k].first, ps[i]);
            bk[k].second = std::max(bk[k].second, ps[i]);
        }
        at = phi_;
        double prev = plo;
        for (int64_t k = 0; k < nb; ++k) {
            if (!(bk[k].first <= bk[k].second)) continue;  // empty bucket
            const double d = bk[k].first - prev;
            if (d > gap) { gap = d; at = bk[k].first; }
            prev = bk[k].second;
        }
    }
    const double wrap = 2 * kPi - phi_ + plo;
    double start = at;
    if (wrap >= gap) { gap = wrap; start = plo; }
    const double ext = 2 * kPi - gap;
    if (ext <= 0) { b.phistart = plo; b.phiext = 0.0; }
    else if (ext >= 2 * kPi) { b.phistart = 0.0; b.phiext = 2 * kPi; }
    else { b.phistart = start; b.phiext = ext; }
    return b;
}

double box_fsky(const Box &b) { return (std::cos(b.thlo) - std::cos(b.thhi)) * b.phiext / (4 * kPi); }

double spectral_resolution(const Box &b) {
    const double e = std::sqrt(2.220446049250313e-16);
    const double mid = std::min(std::max(0.5 * (b.thlo + b.thhi), e), kPi - e);
    const double lam = std::max(b.thhi - b.thlo, b.phiext * std::sin(mid));
    return lam <= 0 ? INFINITY : kPi / lam;
}

namespace {
struct GapResult {
    double lo, hi, gap, at;
};

// Worker threads for the O(n) scans of the partition: at most 16, one when the group is small.
int scan_threads(int64_t n) {
    if (n < (int64_t(1) << 16)) return 1;
    const unsigned hw = std::thread::hardware_concurrency();
    return int(std::max(1u, std::min(16u, hw)));
}

template <class F>
void run_threads(int T, F &&f) {
    if (T <= 1) { f(0); return; }
    std::vector<std::thread> th;
    th.reserve(T - 1);
    for (int t = 1; t < T; ++t) th.emplace_back(f, t);
    f(0);
    for (auto &x : th) x.join();
}

// _max_gap over the values v[g[i]] (src/Box.jl:257-296): the largest distance between neighbours
// in sorted order and the value above it, by n - 1 buckets of width w = (hi - lo) / (n - 1) (the
// largest gap is at least w, a gap inside one bucket is below w). A value within rounding of a
// bucket edge may land in the neighbouring bucket, which cannot move a gap of at least w.
// In parallel: each thread buckets a chunk of the values and counts them per thread-owned bucket
// range; the (bucket, value) pairs are then grouped by range; each thread fills the buckets of its
// range and finds the first largest gap inside it; the ranges are joined left to right, so the
// first largest gap wins as in one sequential scan.
struct Bucket {
    double lo, hi;
};
struct Item {
    int64_t k;
    double x;
};
struct GapScratch {  // n buckets, n bucket indices, n items: reused across the scans of a partition
    std::unique_ptr<Bucket[]> bk;
    std::unique_ptr<int64_t[]> key;
    std::unique_ptr<Item[]> item;
    explicit GapScratch(int64_t n)
        : bk(new Bucket[size_t(n)]), key(new int64_t[size_t(n)]), item(new Item[size_t(n)]) {}
};

GapResult max_gap(const double *v, const std::vector<int32_t> &g, GapScratch &S) {
    const int64_t n = int64_t(g.size());
    const int T = scan_threads(n);
    const int32_t *gi = g.data();
    std::vector<double> tlo(T, INFINITY), thi(T, -INFINITY);
    run_threads(T, [&](int t) {
        double a = INFINITY, b = -INFINITY;
        for (int64_t i = n * t / T; i < n * (t + 1) / T; ++i) {
            a = std::min(a, v[gi[i]]);
            b = std::max(b, v[gi[i]]);
        }
        tlo[t] = a;
        thi[t] = b;
    });
    double lo = v[gi[0]], hi = v[gi[0]];
    for (int t = 0; t < T; ++t) {
        lo = std::min(lo, tlo[t]);
        hi = std::max(hi, thi[t]);
    }
    if (n < 2 || hi <= lo) return {lo, hi, 0.0, lo};
    const int64_t nb = n - 1;
    const double iw = double(nb) / (hi - lo);
    // bucket range r owns buckets [k0(r), k0(r + 1)), k0(r) = ceil(r nb / T), so range(k) = k T / nb
    auto k0 = [&](int r) { return (int64_t(r) * nb + T - 1) / T; };
    std::vector<int64_t> cnt(size_t(T) * T, 0);  // cnt[t * T + r]: values of chunk t in range r
    int64_t *key = S.key.get();
    run_threads(T, [&](int t) {
        std::vector<int64_t> c(T, 0);
        for (int64_t i = n * t / T; i < n * (t + 1) / T; ++i) {
            const double x = v[gi[i]];
            int64_t k = int64_t((x - lo) * iw);
            k = std::min(k, nb - 1);
            key[i] = k;
            ++c[k * T / nb];
        }
        for (int r = 0; r < T; ++r) cnt[size_t(t) * T + r] = c[r];
    });
    std::vector<int64_t> ofs(size_t(T) * T + 1, 0);  // range-major offsets: (r, t)
    {
        int64_t o = 0;
        for (int r = 0; r < T; ++r)
            for (int t = 0; t < T; ++t) {
                ofs[size_t(t) * T + r] = o;
                o += cnt[size_t(t) * T + r];
            }
        ofs[size_t(T) * T] = o;
    }
    Item *item = S.item.get();
    run_threads(T, [&](int t) {
        std::vector<int64_t> w(ofs.begin() + size_t(t) * T, ofs.begin() + size_t(t + 1) * T);
        for (int64_t i = n * t / T; i < n * (t + 1) / T; ++i) {
            const int64_t k = key[i];
            item[w[k * T / nb]++] = {k, v[gi[i]]};
        }
    });
    struct Part {
        bool any = false;
        double first = 0, last = 0, gap = 0, at = 0;
    };
    std::vector<Part> part(T);
    Bucket *bk = S.bk.get();
    run_threads(T, [&](int r) {
        const int64_t b0 = k0(r), b1 = k0(r + 1);
        for (int64_t k = b0; k < b1; ++k) bk[k] = {INFINITY, -INFINITY};
        const int64_t i0 = ofs[size_t(0) * T + r], i1 = r + 1 < T ? ofs[size_t(0) * T + r + 1] : n;
        for (int64_t i = i0; i < i1; ++i) {
            Bucket &b = bk[item[i].k];
            b.lo = std::min(b.lo, item[i].x);
            b.hi = std::max(b.hi, item[i].x);
        }
        Part P;
        double prev = 0;
        for (int64_t k = b0; k < b1; ++k) {
            if (!(bk[k].lo <= bk[k].hi)) continue;
            if (!P.any) {
                P.any = true;
                P.first = bk[k].lo;
            } else {
                const double d = bk[k].lo - prev;
                if (d > P.gap) { P.gap = d; P.at = bk[k].lo; }
            }
            prev = bk[k].hi;
        }
        P.last = prev;
        part[r] = P;
    });
    double gap = 0.0, at = hi, prev = lo;
    for (const Part &P : part) {
        if (!P.any) continue;
        const double d = P.first - prev;
        if (d > gap) { gap = d; at = P.first; }
        if (P.gap > gap) { gap = P.gap; at = P.at; }
        prev = P.last;
    }
    return {lo, hi, gap, at};
}

// _split_below (src/Box.jl:356-374) fused with the colatitude ends of the two halves (the part of
// _group_box, src/Box.jl:298-319, that cost_stretch reads): a keeps the points below thresh, b the
// rest, both in the group's order. False when either side would be empty.
struct Halves {
    std::vector<int32_t> a, b;
    double alo = INFINITY, ahi = -INFINITY, blo = INFINITY, bhi = -INFINITY;
};
bool split_below(const double *v, const double *th, const std::vector<int32_t> &g, double thresh, Halves &h) {
    const int64_t n = int64_t(g.size());
    const int T = scan_threads(n);
    std::vector<int64_t> na(T + 1, 0);
    std::vector<double> alo(T, INFINITY), ahi(T, -INFINITY), blo(T, INFINITY), bhi(T, -INFINITY);
    run_threads(T, [&](int t) {  // count and ends per chunk
        int64_t c = 0;
        double a0 = INFINITY, a1 = -INFINITY, b0 = INFINITY, b1 = -INFINITY;
        for (int64_t i = n * t / T; i < n * (t + 1) / T; ++i) {
            const int32_t j = g[i];
            const double x = th[j];
            if (v[j] < thresh) {
                ++c;
                a0 = std::min(a0, x);
                a1 = std::max(a1, x);
            } else {
                b0 = std::min(b0, x);
                b1 = std::max(b1, x);
            }
        }
        na[t + 1] = c;
        alo[t] = a0; ahi[t] = a1; blo[t] = b0; bhi[t] = b1;
    });
    for (int t = 0; t < T; ++t) na[t + 1] += na[t];
    h = Halves();
    h.a.resize(size_t(na[T]));
    h.b.resize(size_t(n - na[T]));
    run_threads(T, [&](int t) {  // each chunk writes its points at its offsets, in order
        int64_t ia = na[t], ib = n * t / T - na[t];
        for (int64_t i = n * t / T; i < n * (t + 1) / T; ++i) {
            const int32_t j = g[i];
            if (v[j] < thresh) h.a[ia++] = j;
            else h.b[ib++] = j;
        }
    });
    for (int t = 0; t < T; ++t) {
        h.alo = std::min(h.alo, alo[t]); h.ahi = std::max(h.ahi, ahi[t]);
        h.blo = std::min(h.blo, blo[t]); h.bhi = std::max(h.bhi, bhi[t]);
    }
    return !h.a.empty() && !h.b.empty();
}

// _best_gap_split (src/Box.jl:325-343): the larger of the colatitude gap and the longitude gap
// weighted by the ring radius
bool best_gap_split(const double *th, const double *ph, const std::vector<int32_t> &g, GapScratch &S, Halves &h) {
    const GapResult t = max_gap(th, g, S);
    const GapResult f = max_gap(ph, g, S);
    const double e = std::sqrt(2.220446049250313e-16);
    const double mid = std::min(std::max(0.5 * (t.lo + t.hi), e), kPi - e);
    const double wphi = f.gap * std::sin(mid);
    if (t.gap >= wphi) {
        if (!(t.gap > 0)) return false;
        return split_below(th, th, g, t.at, h);
    }
    if (!(f.gap > 0)) return false;
    return split_below(ph, th, g, f.at, h);
}

// A group's best cut, computed once per group: its two halves, their costs and the gain.
struct Candidate {
    bool done = false, ok = false;
    double gain = 0, ca = 0, cb = 0;
    Halves h;
};
}  // namespace

// partition_boxes_costed (src/Box.jl:173-235): starting from one group, repeatedly make the cut
// with the largest positive fall in counted cost over all groups (the first such group on ties),
// up to maxboxes groups. A group's candidate cut does not change while the group does, so each is
// evaluated once.
Partition partition_boxes(const double *theta, const double *phi, int64_t n, int L, int B, double tol,
                          int maxboxes) {
    // longitudes in [0, 2 pi) (fmod returns x itself there); a copy only when some lie outside
    std::vector<double> phc;
    const double *th = theta, *ph = phi;
    bool inrange = true;
    for (int64_t i = 0; i < n && inrange; ++i) inrange = phi[i] >= 0 && phi[i] < 2 * kPi;
    if (!inrange) {
        phc.resize(n);
        for (int64_t i = 0; i < n; ++i) phc[i] = jmod(phi[i], 2 * kPi);
        ph = phc.data();
    }
    GapScratch scratch(std::max<int64_t>(n, 1));
    auto cost = [&](double lo, double hi, int64_t np) { return cost_stretch(lo, hi, np, L, B, tol); };
    Partition P;
    std::vector<int32_t> all(n);
    for (int64_t i = 0; i < n; ++i) all[i] = int32_t(i);
    double lo0 = th[0], hi0 = th[0];
    for (int64_t i = 1; i < n; ++i) { lo0 = std::min(lo0, th[i]); hi0 = std::max(hi0, th[i]); }
    P.groups.push_back(std::move(all));
    P.costs.push_back(cost(lo0, hi0, n));
    std::vector<Candidate> cand(1);
    while (int(P.groups.size()) < maxboxes) {
        double best_gain = 0.0;
        int best_at = -1;
        for (size_t gi = 0; gi < P.groups.size(); ++gi) {
            Candidate &c = cand[gi];
            if (!c.done) {
                c.done = true;
                const auto &g = P.groups[gi];
                if (g.size() >= 2 && best_gap_split(th, ph, g, scratch, c.h)) {
                    c.ok = true;
                    c.ca = cost(c.h.alo, c.h.ahi, int64_t(c.h.a.size()));
                    c.cb = cost(c.h.blo, c.h.bhi, int64_t(c.h.b.size()));
                    c.gain = P.costs[gi] - (c.ca + c.cb);
                }
                if (!c.ok) c.h = Halves();
            }
            if (c.ok && c.gain > best_gain) {
                best_gain = c.gain;
                best_at = int(gi);
            }
        }
        if (best_at < 0) break;
        Candidate c = std::move(cand[best_at]);
        P.groups.erase(P.groups.begin() + best_at);
        P.costs.erase(P.costs.begin() + best_at);
        cand.erase(cand.begin() + best_at);
        P.groups.push_back(std::move(c.h.a));
        P.costs.push_back(c.ca);
        P.groups.push_back(std::move(c.h.b));
        P.costs.push_back(c.cb);
        cand.emplace_back();
        cand.emplace_back();
    }
    return P;
}

static double fftunits(int64_t n) { return n <= 1 ? 0.0 : double(n) * std::log2(double(n)); }

double cost_direct(int64_t npts, int L, int B) {
    return 1.0 * double(npts) * double(L + 1) * double(L + 1) * double(B) / 2;
}

double cost_stretch(double thlo, double thhi, int64_t npts, int L, int B, double tol) {
    Geom g;
    std::string msg;
    if (stretch_geometry(thlo, thhi, L, tol, g, msg) != 0) return INFINITY;
    const int Kt = g.K + kKMargin;
    int64_t nth, nh;
    if (g.kind == kInterior) { nth = next_smooth(2 * Kt + 1); nh = nth; }
    else { nh = next_smooth(Kt + 1); nth = 2 * nh; }
    std::vector<double> th, Th;
    stretch_rings(g, int(nh), th, Th);
    if (!std::is_sorted(th.begin(), th.end())) std::reverse(th.begin(), th.end());
    std::vector<int> mlim(nh);
    int mx = 0;
    for (int64_t j = 0; j < nh; ++j) { mlim[j] = polar_mmax(th[j], L, tol); mx = std::max(mx, mlim[j]); }
    std::vector<int> jlo, jhi;
    if (!nested_ranges(mlim, L, jlo, jhi, msg)) return INFINITY;
    int64_t rings = 0;
    for (int m = 0; m <= L; ++m)
        if (jhi[m] >= jlo[m]) rings += int64_t(jhi[m] - jlo[m] + 1) * (L - m + 1);
    const int64_t nph = next_smooth_even(2 * int64_t(mx) + 1);
    const double fft = fftunits(nph) * double(nh) + fftunits(nth * nph);
    const double sp = double(npts) * std::pow(2 * std::ceil(std::fabs(std::log10(std::max(tol, 1e-16)))) + 2, 2);
    return B * (1.0 * double(rings) + 0.394 * fft + 0.334 * sp);
}

void es_params(double tol, int &w, double &beta) {
    int ww = int(std::ceil(-std::log10(tol))) + 2;
    ww = std::clamp(ww, 2, 16);
    double bo = 2.30;
    if (ww == 2) bo = 2.20;
    else if (ww == 3) bo = 2.26;
    else if (ww == 4) bo = 2.38;
    w = ww;
    beta = bo * ww;
}

// The fine grid is any length N >= 2n (a larger upsampling factor only lowers the kernel's aliasing
// error at the same width), chosen for rocFFT on gfx950: up to 4096 the next even 5-smooth length
// (one kernel per transform at every such length); above 4096 rocFFT runs most lengths as a
// multi-kernel plan with transposes through a work buffer at a third to a fifth of the single-kernel
// bandwidth, so the grid takes the next of the lengths it still runs in one kernel (measured over
// all even 5-smooth lengths up to 10240, FP32 and FP64).
int es_fine_size(int n, int w) {
    int64_t N = next_smooth_even(std::max<int64_t>(2 * int64_t(n), 2 * w));
    if (N > 4096)
        for (int64_t f : {6144, 8192, 9216, 10000, 10240})
            if (f >= N) return int(f);
    return int(N);
}

// Gauss–Legendre nodes by Newton iteration on P_n.
static void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
    x.resize(n);
    w.resize(n);
    for (int i = 0; i < n; ++i) {
        double z = std::cos(kPi * (i + 0.75) / (n + 0.5)), dp = 1;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1, p1 = z;
            for (int k = 2; k <= n; ++k) { const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
            dp = n * (z * p1 - p0) / (z * z - 1);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        double p0 = 1, p1 = z;
        for (int k = 2; k <= n; ++k) { const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
        dp = n * (z * p1 - p0) / (z * z - 1);
        x[i] = z;
        w[i] = 2 / ((1 - z * z) * dp * dp);
    }
}

void es_corr(int n, int N, int w, double beta, std::vector<double> &corr) {
    // psihat(k) = alpha * int_{-1}^{1} phi(z) cos(k alpha z) dz with z = sin t (analytic integrand).
    const int nq = 80;
    std::vector<double> gx, gw;
    gauss_legendre(nq, gx, gw);
    const double h = 2 * kPi / N, alpha = w * h / 2;
    corr.resize(n);
    for (int i = 0; i < n; ++i) {
        const int k = mode_k(i, n);
        double acc = 0;
        for (int q = 0; q < nq; ++q) {
            const double t = gx[q] * kPi / 2;
            acc += gw[q] * (kPi / 2) * std::exp(beta * (std::cos(t) - 1)) * std::cos(double(k) * alpha * std::sin(t)) *
                   std::cos(t);
        }
        corr[i] = h / (alpha * acc);
    }
}

}  // namespace pfsht
