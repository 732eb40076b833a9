// geometry.hpp — plan-time host tables for the stretched leaf: polar optimisation, nested ring
// ranges, the stretch geometry and window, grid sizing, and the ES NUFFT kernel parameters.
// O(nrings + lmax + npts) host work done once per plan; everything here is uploaded to HBM and
// stays resident for the plan's lifetime.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace pfsht {

constexpr double kPi = 3.14159265358979323846;

// Renormalisation stride of the block-scaled recurrence (src/Legendre.jl:16).
constexpr int kRescaleStride = 64;
// Spare stretched modes above the bandlimit (src/Stretch.jl:86).
constexpr int kKMargin = 2;
// Largest apodisation fraction (src/Windows.jl:20).
constexpr double kEpsApoMax = 1.0;

enum Kind { kPole = 0, kInterior = 1, kFull = 2 };

struct Geom {  // StretchGeometry, src/Windows.jl:42-57
    int kind = kPole, lmax = 0, K = 0;
    double eps = 0, beta = 0, h = 0, thlo = 0, thhi = 0, epsapo = 0, period = 0;
    double Theta1 = 0, sigma = 0, ThetaT = 0, invZ = 0;
};

// src/Legendre.jl:48-57
int polar_mmax(double theta, int L, double eps);
// src/Legendre.jl:65-75
void sectorial_log2c(int L, std::vector<double> &out);
// src/Legendre.jl:89-127 (0-based rings, empty when jhi < jlo); false + msg when not unimodal
bool nested_ranges(const std::vector<int> &mlim, int L, std::vector<int> &jlo, std::vector<int> &jhi,
                   std::string &msg);
// src/Windows.jl:208-226
double es_cdf(double beta, double u);
// src/Windows.jl:84-141: 0 ok, 1 degenerate, -1 bad argument (msg set); epsapo_max is the largest
// apodisation fraction admitted (Plan's εapo_max, src/Plan.jl:97)
int stretch_geometry(double thlo, double thhi, int L, double eps, Geom &g, std::string &msg,
                     double epsapo_max = kEpsApoMax);
// src/Windows.jl:237-250, 272-298, 307-320
double stretch_window(const Geom &g, double Theta);
double stretch_to_theta(const Geom &g, double Theta);
double theta_to_stretch(const Geom &g, double theta);
void stretch_rings(const Geom &g, int n, std::vector<double> &theta, std::vector<double> &Theta);
// src/Stretch.jl:52-81
int64_t next_smooth(int64_t n);
int64_t next_smooth_even(int64_t n);
// src/NUFFT.jl:74-83 (modeord 1, 0-based index)
inline int mode_k(int i, int n) { return i < (n + 1) / 2 ? i : i - n; }

// Julia's floating mod (result carries the divisor's sign).
double jmod(double x, double y);

// bounding_box (src/Box.jl:43-73): colatitude range clamped to [0, pi], pole-touch flags, and the
// smallest longitude arc holding the points (src/Box.jl:80-102; 0 .. 2 pi when a pole is touched),
// by the O(n) maximum-gap bucket scan (src/Box.jl:246-296).
struct Box {
    double thlo = 0, thhi = 0, phistart = 0, phiext = 0;
    bool north = false, south = false;
};
Box bounding_box(const double *theta, const double *phi, int64_t n);
// the same from a known colatitude range (unclamped min and max) and the longitudes
Box bounding_box_range(double thmin, double thmax, const double *phi, int64_t n);
// box_fsky (src/Box.jl:106-108) and spectral_resolution (src/Box.jl:148-153)
double box_fsky(const Box &b);
double spectral_resolution(const Box &b);

// partition_boxes_costed (src/Box.jl:159-235): split the points at the largest colatitude or
// longitude gap of a group while the counted cost of stretched leaves falls, up to maxboxes
// groups, with the O(n) bucket scan of _max_gap (src/Box.jl:246-296) and the group box of
// src/Box.jl:298-319. Groups hold point indices in increasing order, in the reference's order of
// creation; costs[g] is cost_stretch of group g.
struct Partition {
    std::vector<std::vector<int32_t>> groups;
    std::vector<double> costs;
};
constexpr int kMaxBoxes = 8;  // partition_boxes' default (src/Box.jl:171)
Partition partition_boxes(const double *theta, const double *phi, int64_t n, int L, int B, double tol,
                          int maxboxes);

// Cost model restricted to Direct vs Stretch (src/Cost.jl:104-140,174-201).
double cost_direct(int64_t npts, int L, int B);
double cost_stretch(double thlo, double thhi, int64_t npts, int L, int B, double tol);

// Exponential-of-semicircle NUFFT kernel (Barnett, Magland & af Klinteberg 2019), upsampling 2:
// width w = ceil(log10(1/tol)) + 2, shape beta = 2.30 w.
void es_params(double tol, int &w, double &beta);
// Fine-grid length for a length-n axis (at least 2n and 2w; see geometry.cpp).
int es_fine_size(int n, int w);
// corr[i] = h / psihat(k_i): deconvolution factors for the modes of a length-n axis in FFT order.
void es_corr(int n, int N, int w, double beta, std::vector<double> &corr);


}  // namespace pfsht
