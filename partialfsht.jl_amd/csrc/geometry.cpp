// geometry.cpp — see geometry.hpp. Formulas follow the cited reference lines; arithmetic is kept
// in the reference's evaluation order so plan sizes (nrings, nph, mmax, ring_work) match it
// exactly.
#include "geometry.hpp"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>

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

int stretch_geometry(double thlo, double thhi, int L, double eps, Geom &out, std::string &msg) {
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
        if (!(Lh >= 2 * beta / (kEpsApoMax * kEpsApoMax))) continue;
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
        msg = "stretched transform is degenerate for this colatitude range at this lmax and tol: every "
              "periodisation needs an apodisation wider than the data; use a direct evaluation";
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

// _max_gap over the values v[g[i]] (src/Box.jl:257-296); bk is bucket scratch of >= |g| - 1 pairs
GapResult max_gap(const std::vector<double> &v, const std::vector<int32_t> &g,
                  std::vector<std::pair<double, double>> &bk) {
    const int64_t n = int64_t(g.size());
    double lo = v[g[0]], hi = v[g[0]];
    for (int64_t i = 1; i < n; ++i) {
        lo = std::min(lo, v[g[i]]);
        hi = std::max(hi, v[g[i]]);
    }
    if (n < 2 || hi <= lo) return {lo, hi, 0.0, lo};
    const int64_t nb = n - 1;
    // bucket k ~ [lo + k w, lo + (k + 1) w): a value within rounding of a boundary may land in the
    // neighbouring bucket, which cannot move a gap of at least w, and the largest gap is at least w
    const double iw = double(nb) / (hi - lo);
    for (int64_t k = 0; k < nb; ++k) bk[k] = {INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; ++i) {
        const double x = v[g[i]];
        int64_t k = int64_t((x - lo) * iw);
        k = std::min(k, nb - 1);
        bk[k].first = std::min(bk[k].first, x);
        bk[k].second = std::max(bk[k].second, x);
    }
    double gap = 0.0, at = hi, prev = lo;
    for (int64_t k = 0; k < nb; ++k) {
        if (!(bk[k].first <= bk[k].second)) continue;
        const double d = bk[k].first - prev;
        if (d > gap) { gap = d; at = bk[k].first; }
        prev = bk[k].second;
    }
    return {lo, hi, gap, at};
}

// _group_box (src/Box.jl:298-319): colatitude ends (not clamped) and the longitude arc
Box group_box(const std::vector<double> &th, const std::vector<double> &ph, const std::vector<int32_t> &g,
              std::vector<std::pair<double, double>> &bk) {
    double lo = th[g[0]], hi = th[g[0]];  // only the colatitude ends are needed here
    for (int32_t i : g) {
        lo = std::min(lo, th[i]);
        hi = std::max(hi, th[i]);
    }
    Box b;
    b.thlo = lo;
    b.thhi = hi;
    const double ptol = std::sqrt(2.220446049250313e-16);
    b.north = lo <= ptol;
    b.south = hi >= kPi - ptol;
    if (b.north || b.south) { b.phistart = 0.0; b.phiext = 2 * kPi; return b; }
    const GapResult f = max_gap(ph, g, bk);
    const double wrap = 2 * kPi - f.hi + f.lo;
    double gap = f.gap, start = f.at;
    if (wrap >= gap) { gap = wrap; start = f.lo; }
    const double ext = 2 * kPi - gap;
    if (ext <= 0) { b.phistart = f.lo; b.phiext = 0.0; }
    else if (ext >= 2 * kPi) { b.phistart = 0.0; b.phiext = 2 * kPi; }
    else { b.phistart = start; b.phiext = ext; }
    return b;
}

// _split_below (src/Box.jl:356-374): false when either side would be empty
bool split_below(const std::vector<double> &v, const std::vector<int32_t> &g, double thresh,
                 std::vector<int32_t> &a, std::vector<int32_t> &b) {
    a.clear();
    b.clear();
    for (int32_t i : g) (v[i] < thresh ? a : b).push_back(i);
    return !a.empty() && !b.empty();
}

// _best_gap_split (src/Box.jl:325-343): the larger of the colatitude gap and the longitude gap
// weighted by the ring radius
bool best_gap_split(const std::vector<double> &th, const std::vector<double> &ph, const std::vector<int32_t> &g,
                    std::vector<std::pair<double, double>> &bk, std::vector<int32_t> &a, std::vector<int32_t> &b) {
    const GapResult t = max_gap(th, g, bk);
    const GapResult f = max_gap(ph, g, bk);
    const double e = std::sqrt(2.220446049250313e-16);
    const double mid = std::min(std::max(0.5 * (t.lo + t.hi), e), kPi - e);
    const double wphi = f.gap * std::sin(mid);
    if (t.gap >= wphi) {
        if (!(t.gap > 0)) return false;
        return split_below(th, g, t.at, a, b);
    }
    if (!(f.gap > 0)) return false;
    return split_below(ph, g, f.at, a, b);
}
}  // namespace

Partition partition_boxes(const double *theta, const double *phi, int64_t n, int L, int B, double tol,
                          int maxboxes) {
    std::vector<double> th(theta, theta + n), ph(n);
    for (int64_t i = 0; i < n; ++i) ph[i] = jmod(phi[i], 2 * kPi);
    std::vector<std::pair<double, double>> bk(size_t(std::max<int64_t>(n, 1)));
    auto cost = [&](const Box &b, int64_t np) { return cost_stretch(b.thlo, b.thhi, np, L, B, tol); };
    Partition P;
    std::vector<int32_t> all(n);
    for (int64_t i = 0; i < n; ++i) all[i] = int32_t(i);
    P.groups.push_back(std::move(all));
    P.costs.push_back(cost(group_box(th, ph, P.groups[0], bk), n));
    std::vector<int32_t> a, b;
    while (int(P.groups.size()) < maxboxes) {
        double best_gain = 0.0;
        int best_at = -1;
        std::vector<int32_t> ba, bb;
        double ca = 0, cb = 0;
        for (size_t gi = 0; gi < P.groups.size(); ++gi) {
            const auto &g = P.groups[gi];
            if (g.size() < 2) continue;
            if (!best_gap_split(th, ph, g, bk, a, b)) continue;
            const double xa = cost(group_box(th, ph, a, bk), int64_t(a.size()));
            const double xb = cost(group_box(th, ph, b, bk), int64_t(b.size()));
            const double gain = P.costs[gi] - (xa + xb);
            if (gain > best_gain) {
                best_gain = gain;
                best_at = int(gi);
                ba = a;
                bb = b;
                ca = xa;
                cb = xb;
            }
        }
        if (best_at < 0) break;
        P.groups.erase(P.groups.begin() + best_at);
        P.costs.erase(P.costs.begin() + best_at);
        P.groups.push_back(std::move(ba));
        P.costs.push_back(ca);
        P.groups.push_back(std::move(bb));
        P.costs.push_back(cb);
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

int es_fine_size(int n, int w) {
    int64_t N = next_smooth_even(2 * int64_t(n));
    if (N < 2 * w) N = next_smooth_even(2 * w);
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
