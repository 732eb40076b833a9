// capi.hip — plan construction, apply orchestration and the C ABI of include/pfsht.h.
//
// A plan is the device analogue of StretchLeaf (src/Stretch.jl:94-248): geometry, ring tables,
// window, bin-sorted points, NUFFT kernel tables, work buffers and rocFFT plans are built once and
// stay resident in HBM; an apply is a fixed sequence of launches on the caller's stream and
// allocates nothing.
#include <rocfft/rocfft.h>

#include <algorithm>
#include <complex>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>

#include "../../include/pfsht.h"
#include "pfsht_common.hpp"
#include "tune_cache.hpp"

namespace pfsht {

static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }

template <class T>
static Status dalloc(T **p, size_t count, int64_t &bytes) {
    *p = nullptr;
    if (count == 0) return kOk;
    const size_t nb = count * sizeof(T);
    if (std::getenv("PFSHT_FAIL_ALLOC")) {  // fault injection for the out-of-memory path
        set_error("device allocation failed (PFSHT_FAIL_ALLOC set)");
        return kENomem;
    }
    hipError_t e = hipMalloc(reinterpret_cast<void **>(p), nb);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        set_error(std::string("device allocation of ") + std::to_string(nb) + " bytes failed: " + hipGetErrorString(e));
        return kENomem;
    }
    bytes += int64_t(nb);
    return kOk;
}

template <class T>
static Status upload(T **p, const std::vector<T> &h, int64_t &bytes) {
    Status st = dalloc(p, h.size(), bytes);
    if (st != kOk || h.empty()) return st;
    PFSHT_HIP_TRY(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return kOk;
}

template <class T>
static void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Temporary device buffers of a plan build, freed once the device is idle when it goes out of scope.
struct DevScratch {
    std::vector<void *> ptrs;
    template <class T>
    Status take(T **p, size_t count, int64_t &bytes) {
        Status st = dalloc(p, count, bytes);
        if (st == kOk && *p) ptrs.push_back(*p);
        return st;
    }
    ~DevScratch() {
        if (!ptrs.empty()) (void)hipDeviceSynchronize();
        for (void *q : ptrs) (void)hipFree(q);
    }
};

// RAII device guard: make the plan's device current for the call, restore the caller's after.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) (void)hipSetDevice(dev);
        else prev = -1;
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

static std::once_flag g_rocfft_once;
static rocfft_status g_rocfft_setup = rocfft_status_success;

struct FftPlan {
    rocfft_plan plan = nullptr;
    rocfft_execution_info info = nullptr;
    size_t work_bytes = 0;
};

// single: FP32 data (the grid buffers of a mixed-precision plan), else FP64. The work buffer is set
// later (fft_attach_work): every FFT of a plan shares one, since a plan's FFTs run one after another.
static Status fft_make(FftPlan &fp, bool forward, size_t n_fast, size_t n_slow, size_t batch, bool single = false) {
    std::call_once(g_rocfft_once, [] { g_rocfft_setup = rocfft_setup(); });
    if (g_rocfft_setup != rocfft_status_success) { set_error("rocfft_setup failed"); return kEHip; }
    size_t lengths[2] = {n_fast, n_slow};
    const size_t dims = n_slow > 0 ? 2 : 1;
    rocfft_status s = rocfft_plan_create(&fp.plan, rocfft_placement_inplace,
                                         forward ? rocfft_transform_type_complex_forward
                                                 : rocfft_transform_type_complex_inverse,
                                         single ? rocfft_precision_single : rocfft_precision_double, dims, lengths,
                                         batch, nullptr);
    if (s != rocfft_status_success) { set_error("rocfft_plan_create failed"); return kEHip; }
    rocfft_plan_get_work_buffer_size(fp.plan, &fp.work_bytes);
    if (rocfft_execution_info_create(&fp.info) != rocfft_status_success) {
        set_error("rocfft_execution_info_create failed");
        return kEHip;
    }
    return kOk;
}

static void fft_free(FftPlan &fp) {
    if (fp.plan) rocfft_plan_destroy(fp.plan);
    if (fp.info) rocfft_execution_info_destroy(fp.info);
    fp = FftPlan{};
}

static Status fft_exec(FftPlan &fp, void *buf, hipStream_t s) {
    rocfft_execution_info_set_stream(fp.info, s);
    void *in[1] = {buf};
    if (rocfft_execute(fp.plan, in, nullptr, fp.info) != rocfft_status_success) {
        set_error("rocfft_execute failed");
        return kEHip;
    }
    return kOk;
}

// The FFTs of one grid stage for 1 .. B transforms: a plan per transform count in `counts` (B, and
// with partial applies enabled every power of two below it), each over count x per batch entries of
// length n. An apply of nb transforms runs the largest counts that fit nb on consecutive transform
// blocks; with only B available it runs B (the extra transforms' buffers hold data no output reads).
struct FftSet {
    std::vector<int> counts;  // descending
    std::vector<FftPlan> plans;
    size_t stride_bytes = 0;  // bytes of one transform's data
};

static Status fftset_make(FftSet &fs, bool forward, size_t n, size_t per, int B, bool partial, bool single) {
    fs.counts.clear();
    fs.counts.push_back(B);
    if (partial)
        for (int c = 1; c < B; c *= 2) fs.counts.push_back(c);
    std::sort(fs.counts.rbegin(), fs.counts.rend());
    fs.plans.assign(fs.counts.size(), FftPlan{});
    for (size_t i = 0; i < fs.counts.size(); ++i) {
        Status st = fft_make(fs.plans[i], forward, n, 0, size_t(fs.counts[i]) * per, single);
        if (st != kOk) return st;
    }
    fs.stride_bytes = n * per * (single ? sizeof(float2) : sizeof(double2));
    return kOk;
}

static void fftset_free(FftSet &fs) {
    for (auto &p : fs.plans) fft_free(p);
    fs.plans.clear();
}

static Status fftset_exec(FftSet &fs, void *buf, int nb, hipStream_t s) {
    int done = 0;
    while (done < nb) {
        const int rem = nb - done;
        size_t k = 0;
        while (k < fs.counts.size() && fs.counts[k] > rem) ++k;
        if (k == fs.counts.size()) k = fs.counts.size() - 1;  // only larger counts: the smallest (done is 0)
        Status st = fft_exec(fs.plans[k], static_cast<char *>(buf) + size_t(done) * fs.stride_bytes, s);
        if (st != kOk) return st;
        done += fs.counts[k];
    }
    return kOk;
}

}  // namespace pfsht

using namespace pfsht;

// ------------------------------------------------------------------------------------------------
// Ring engine
// ------------------------------------------------------------------------------------------------
struct pfsht_rings_s {
    int dev = 0;
    RingDev d;
    int spin = 0;  // nonzero: the spin engine (sd) with d's geometry and ranges
    SpinDev sd;
    std::vector<int> jlo, jhi, mlim;
    int64_t work = 0;
    int64_t bytes = 0;
    int64_t part_bytes = 0;                   // the adjoint ring pass's partial sums
    RingShape shape;                          // the launch shape it was built with
    double2 *Gd = nullptr;                    // [B][2 mmax + 1][n]
    double2 *alm_stage = nullptr, *G_stage = nullptr;  // for host-memory calls
    std::mutex mu;

    ~pfsht_rings_s() {
        dfree(d.costh); dfree(d.log2sin); dfree(d.log2c); dfree(d.ab); dfree(d.gc); dfree(d.abofs);
        dfree(d.jlo); dfree(d.jhi); dfree(d.fitems); dfree(d.aitems); dfree(d.mitems); dfree(d.ckq); dfree(d.ckE); dfree(d.pofs); dfree(d.ntiles); dfree(d.partials);
        dfree(Gd); dfree(alm_stage); dfree(G_stage);
        dfree(sd.l2c); dfree(sd.l2s); dfree(sd.seed); dfree(sd.Nl); dfree(sd.tab); dfree(sd.tofs); dfree(sd.items);
        dfree(sd.pofs); dfree(sd.ntiles); dfree(sd.partials);
    }
};

static Status rings_build(const double *theta, int n, int L, double eps, int B, bool reflect, pfsht_rings_s &r,
                          const RingShape &shape = RingShape()) {
    if (L < 0) { set_error("lmax must be nonnegative"); return kEArg; }
    if (B < 1) { set_error("ntrans must be >= 1"); return kEArg; }
    if (n < 1) { set_error("need at least one ring"); return kEArg; }
    for (int j = 1; j < n; ++j)
        if (!(theta[j] >= theta[j - 1])) { set_error("ring colatitudes must be sorted ascending"); return kEArg; }
    (void)hipGetDevice(&r.dev);
    const double csgn = reflect ? -1.0 : 1.0;
    std::vector<double> costh(n), log2sin(n);
    std::vector<int> mlim(n);
    int mmax = 0;
    for (int j = 0; j < n; ++j) {
        const double s = std::sin(theta[j]), c = std::cos(theta[j]);
        const double as = std::fabs(s);
        costh[j] = csgn * c;
        log2sin[j] = as > 0 ? std::log2(as) : -INFINITY;
        mlim[j] = polar_mmax(theta[j], L, eps);
        mmax = std::max(mmax, mlim[j]);
    }
    std::string msg;
    if (!nested_ranges(mlim, L, r.jlo, r.jhi, msg)) { set_error(msg); return kEArg; }
    r.mlim = mlim;
    std::vector<double> log2c;
    sectorial_log2c(L, log2c);
    RingDev &d = r.d;
    d.n = n; d.L = L; d.B = B; d.mmax = mmax;
    // recurrence coefficients (src/Legendre.jl:406-408) and the rescaled forward's (gamma_l, C_l),
    // tabulated on the device per order from m + 2; abofs[m] is the order's first entry
    std::vector<int64_t> abofs(mmax + 1);
    int64_t nab = 0;
    for (int m = 0; m <= mmax; ++m) { abofs[m] = nab; nab += std::max(0, L - m - 1); }
    // Work items: one per (order, 64-ring tile), with the order's degree segments; checkpoint rows
    // for segments 1.. of every tile. An adjoint wavefront takes a group of tg consecutive tiles of
    // an order (two when ntrans <= 2, which the registers allow); partials are indexed by the group.
    // The MFMA adjoint (two or more transforms) takes every order's tiles in one workgroup: four
    // wavefronts of one or two tiles, or eight of two, so up to 16 tiles (1024 rings) per order.
    int ntmax = 0;
    for (int m = 0; m <= mmax; ++m)
        if (r.jhi[m] >= r.jlo[m]) ntmax = std::max(ntmax, (r.jhi[m] - r.jlo[m] + 64) / 64);
    const bool mfma_adj = B >= 2 && ntmax <= 16;
    int mw = 0, mtg = 1;
    if (mfma_adj) {
        if (ntmax <= 4) mw = ntmax;
        else if (ntmax <= 8) { mw = 4; mtg = 2; }
        else { mw = 8; mtg = 2; }
        // a tuned shape: one tile per wavefront (up to eight wavefronts), or two
        if (shape.mfma_tg == 1 && ntmax <= 8) { mw = ntmax; mtg = 1; }
        else if (shape.mfma_tg == 2) { mw = (ntmax + 1) / 2; mtg = 2; }
    }
    // partials serve the adjoints the MFMA path does not: single transforms when it exists
    const int pb = mfma_adj ? 1 : B;
    // one or two tiles per wavefront for up to two transforms (the kernels with two are built for
    // those counts only); a tuned shape may take one
    int tg = (B <= 2 || mfma_adj) ? 2 : 1;
    if (tg == 2 && shape.adj_tg == 1) tg = 1;
    std::vector<std::vector<SegItem>> tiles_of(mmax + 1);
    std::vector<int32_t> ntiles(mmax + 1, 0);
    std::vector<int64_t> pofs(mmax + 1, 0), cost(mmax + 1, 0);
    int64_t np = 0, nck = 0;
    r.work = 0;
    int fw = 1;
    const int per = shape.per > 0 ? std::min(64, shape.per) : kSegBlocks;
    for (int m = 0; m <= mmax; ++m) {
        pofs[m] = np;
        const int lo = r.jlo[m], hi = r.jhi[m];
        if (hi < lo) continue;
        const int nj = hi - lo + 1;
        r.work += int64_t(nj) * (L - m + 1);
        cost[m] = int64_t(nj) * (L - m + 1);
        const int nt = (nj + 63) / 64, nseg = seg_count(m, L, per);
        fw = std::max(fw, nseg);
        for (int t = 0; t < nt; ++t) {
            const int j0 = lo + 64 * t;
            tiles_of[m].push_back({m, j0, std::min(64, hi - j0 + 1), t, nseg, 0, nck, 0});
            nck += nseg - 1;
        }
        const int ngroups = (nt + tg - 1) / tg;
        ntiles[m] = ngroups;
        np += int64_t(ngroups) * adj_rows_padded(m, L) * pb;  // group rows padded to whole sub-tiles
    }
    // Orders are dealt to the eight XCDs, longest first onto the least-loaded one, and workgroup i
    // is laid out for XCD i % 8 (the hardware deals consecutive workgroups round-robin): the tiles
    // and segments of an order then share one XCD's L2 for the order's coefficient streams.
    constexpr int kXcd = 8;
    std::vector<int> order;
    for (int m = 0; m <= mmax; ++m)
        if (!tiles_of[m].empty()) order.push_back(m);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<std::vector<int>> xq(kXcd);
    std::vector<int64_t> xload(kXcd, 0);
    for (int m : order) {
        const int x = int(std::min_element(xload.begin(), xload.end()) - xload.begin());
        xq[x].push_back(m);
        xload[x] += cost[m];
    }
    const SegItem empty{0, 0, 0, 0, 1, 0, 0, 0};
    // forward: one workgroup per tile
    std::vector<std::vector<SegItem>> fq(kXcd);
    for (int x = 0; x < kXcd; ++x)
        for (int m : xq[x])
            for (const SegItem &t : tiles_of[m]) fq[x].push_back(t);
    // adjoint: one wavefront per (tile group, segment), kAdjWaves per workgroup; the item is the
    // group's first tile with tile = the group index and nr2 = the rings of its second tile
    std::vector<std::vector<SegItem>> aq(kXcd);
    for (int x = 0; x < kXcd; ++x) {
        for (int m : xq[x]) {
            const auto &tl = tiles_of[m];
            for (size_t t0 = 0; t0 < tl.size(); t0 += tg) {
                SegItem w = tl[t0];
                w.tile = int32_t(t0 / tg);
                w.nr2 = tg > 1 && t0 + 1 < tl.size() ? tl[t0 + 1].nr : 0;
                for (int sg = 0; sg < w.nseg; ++sg) {
                    w.seg = sg;
                    aq[x].push_back(w);
                }
            }
        }
        while (aq[x].size() % kAdjWaves) aq[x].push_back(empty);
    }
    auto interleave = [&](const auto &q, size_t group, const auto &none) {
        size_t rounds = 0;
        for (const auto &v : q) rounds = std::max(rounds, v.size() / group);
        std::vector<std::decay_t<decltype(none)>> out;
        out.reserve(rounds * kXcd * group);
        for (size_t i = 0; i < rounds; ++i)
            for (int x = 0; x < kXcd; ++x)
                for (size_t g = 0; g < group; ++g)
                    out.push_back((i + 1) * group <= q[x].size() ? q[x][i * group + g] : none);
        return out;
    };
    std::vector<SegItem> fitems = interleave(fq, 1, empty);
    std::vector<SegItem> aitems = interleave(aq, kAdjWaves, empty);
    // MFMA adjoint: one workgroup per (order, segment), orders on the XCD they were dealt to
    std::vector<OrderSeg> mitems;
    if (mfma_adj) {
        std::vector<std::vector<OrderSeg>> mq(kXcd);
        for (int x = 0; x < kXcd; ++x)
            for (int m : xq[x]) {
                const auto &tl = tiles_of[m];
                const int nt = int(tl.size());
                for (int sg = 0; sg < tl[0].nseg; ++sg)
                    mq[x].push_back({m, tl[0].nseg, sg, nt, tl[0].j0, tl[nt - 1].nr, tl[0].ck});
            }
        mitems = interleave(mq, 1, OrderSeg{0, 1, 0, 0, 0, 0, 0});
    }
    r.shape = shape;
    r.shape.per = per;
    r.shape.adj_tg = tg;
    r.shape.mfma_tg = mfma_adj ? mtg : 0;
    d.mfma_waves = mw;
    d.mfma_tg = mtg;
    d.part_b = pb;
    d.nmitems = int(mitems.size());
    d.fwd_waves = fw;
    d.adj_tg = tg;
    d.nfitems = int(fitems.size());
    d.naitems = int(aitems.size());
    d.nck = nck;
    d.npartials = np;
    std::vector<int32_t> jlo32(r.jlo.begin(), r.jlo.begin() + mmax + 1), jhi32(r.jhi.begin(), r.jhi.begin() + mmax + 1);
    Status st;
    int64_t &by = r.bytes;
    if ((st = upload(&d.costh, costh, by)) != kOk) return st;
    if ((st = upload(&d.log2sin, log2sin, by)) != kOk) return st;
    if ((st = upload(&d.log2c, log2c, by)) != kOk) return st;
    if ((st = upload(&d.abofs, abofs, by)) != kOk) return st;
    if ((st = dalloc(&d.ab, size_t(std::max<int64_t>(nab, 1)), by)) != kOk) return st;
    if ((st = dalloc(&d.gc, size_t(std::max<int64_t>(nab, 1)), by)) != kOk) return st;
    launch_coef_tables(L, mmax, d.abofs, d.ab, d.gc, nullptr);
    if ((st = upload(&d.jlo, jlo32, by)) != kOk) return st;
    if ((st = upload(&d.jhi, jhi32, by)) != kOk) return st;
    if ((st = upload(&d.fitems, fitems, by)) != kOk) return st;
    if ((st = upload(&d.aitems, aitems, by)) != kOk) return st;
    if ((st = upload(&d.mitems, mitems, by)) != kOk) return st;
    if ((st = dalloc(&d.ckq, size_t(nck) * 64, by)) != kOk) return st;
    if ((st = dalloc(&d.ckE, size_t(nck) * 64, by)) != kOk) return st;
    if ((st = upload(&d.pofs, pofs, by)) != kOk) return st;
    if ((st = upload(&d.ntiles, ntiles, by)) != kOk) return st;
    if ((st = dalloc(&d.partials, size_t(std::max<int64_t>(np, 1)), by)) != kOk) return st;
    r.part_bytes = int64_t(sizeof(double4)) * std::max<int64_t>(np, 1);
    if ((st = dalloc(&r.Gd, size_t(B) * (2 * mmax + 1) * n, by)) != kOk) return st;
    launch_ring_checkpoints(d, nullptr);
    PFSHT_HIP_TRY(hipGetLastError());
    PFSHT_HIP_TRY(hipStreamSynchronize(nullptr));
    return kOk;
}

// One work buffer for FFT plans that run one after another, attached to each plan that needs one.
static Status fft_attach_work(const std::vector<FftPlan *> &plans, void *&work, int64_t &bytes) {
    size_t wb = 0;
    for (const FftPlan *q : plans) wb = std::max(wb, q->work_bytes);
    if (!wb) return kOk;
    Status st = dalloc(reinterpret_cast<char **>(&work), wb, bytes);
    if (st != kOk) return st;
    for (FftPlan *q : plans)
        if (q->info && q->work_bytes) rocfft_execution_info_set_work_buffer(q->info, work, q->work_bytes);
    return kOk;
}

// Whether the LDS grid stage takes the lengths nth, N1, N2 (nth <= N1) with elements of csz bytes:
// each factors into radix-8/4/2/3/5 passes, the column length N1 fits kFftElems x 1024 threads, the
// row length N2 that or, split in two (even, N2 / 2 <= kFftElems x 1024), twice that, and the longer
// fits one workgroup's LDS. Fills g's lengths when it does.
static bool grid_fits(int nth, int N1, int N2, size_t csz, int dev, GridFftDev *g) {
    int maxlds = 65536;
    (void)hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    constexpr int cap = kFftElems * 1024;
    GridFftDev t;
    t.split = N2 > cap;
    if (!fft_len_plan(nth, t.th) || !fft_len_plan(N1, t.c1) || !fft_len_plan(N2, t.r2)) return false;
    if (t.split && (N2 % 2 != 0 || N2 / 2 > cap || !fft_len_plan(N2 / 2, t.r2h))) return false;
    const int n = std::max(N1, N2);
    if (nth > N1 || N1 > cap || size_t(lds_slots(n)) * csz > size_t(maxlds)) return false;
    if (g) { g->th = t.th; g->c1 = t.c1; g->r2 = t.r2; g->r2h = t.r2h; g->split = t.split; }
    return true;
}

// ---------------------------------------------------------------------------------------------
// Plan-time tuning of the ring passes' launch shape (cached in tune_cache.cpp).
// ---------------------------------------------------------------------------------------------
// "gfx950:256": the architecture without its feature suffix, and the compute units
static std::string device_arch(int dev) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
    std::string a = prop.gcnArchName;
    const size_t c = a.find(':');
    if (c != std::string::npos) a.resize(c);
    return a + ":" + std::to_string(prop.multiProcessorCount);
}

// Time candidate shapes on temporary engines of the ring set: the default first, the segment
// target over {2, 4, 6} next, then, at the best target, the other tile count per wavefront of the
// adjoint that the engine uses for its B transforms. A candidate replaces the best only when it is
// 3% faster (timing noise keeps the default). Score: the mean time of a forward plus an adjoint ring
// pass over all B transforms, on zero inputs (the passes' work does not depend on the values).
static Status tune_rings(const double *theta, int n, int L, double eps, int B, bool reflect, RingShape &best,
                         std::vector<RingShape> &table) {
    const size_t LB = size_t(L + 1) * (2 * L + 1);
    struct Res {
        double2 *alm = nullptr;
        hipStream_t s = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Res() {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (s) (void)hipStreamDestroy(s);
            dfree(alm);
        }
    } res;
    int64_t acct = 0;
    Status st = dalloc(&res.alm, LB * size_t(B), acct);
    if (st != kOk) return st;
    PFSHT_HIP_TRY(hipStreamCreateWithFlags(&res.s, hipStreamNonBlocking));
    PFSHT_HIP_TRY(hipEventCreate(&res.e0));
    PFSHT_HIP_TRY(hipEventCreate(&res.e1));
    PFSHT_HIP_TRY(hipMemsetAsync(res.alm, 0, sizeof(double2) * LB * size_t(B), res.s));
    const double work = double(n) * L * L;
    const int reps = int(std::min(20.0, std::max(2.0, 2e8 / std::max(work, 1.0))));
    const InCols in = in_cols(res.alm, int64_t(LB));
    const OutCols out = out_cols(res.alm, int64_t(LB));
    auto time_shape = [&](const RingShape &want, RingShape &got) -> Status {
        pfsht_rings_s r;
        Status st2 = rings_build(theta, n, L, eps, B, reflect, r, want);
        if (st2 != kOk) return st2;
        launch_ring_fwd(r.d, in, B, r.Gd, res.s);  // warm-up
        launch_ring_adj(r.d, r.Gd, out, B, res.s);
        PFSHT_HIP_TRY(hipEventRecord(res.e0, res.s));
        for (int i = 0; i < reps; ++i) {
            launch_ring_fwd(r.d, in, B, r.Gd, res.s);
            launch_ring_adj(r.d, r.Gd, out, B, res.s);
        }
        PFSHT_HIP_TRY(hipEventRecord(res.e1, res.s));
        PFSHT_HIP_TRY(hipGetLastError());
        PFSHT_HIP_TRY(hipEventSynchronize(res.e1));
        float ms = 0;
        PFSHT_HIP_TRY(hipEventElapsedTime(&ms, res.e0, res.e1));
        got = r.shape;
        got.ms = double(ms) / reps;
        table.push_back(got);
        return kOk;
    };
    RingShape got;
    if ((st = time_shape(RingShape(), best)) != kOk) return st;
    auto consider = [&](const RingShape &want) -> Status {
        for (const RingShape &t : table)  // the engine may not take the shape asked for
            if (t.per == want.per && (want.adj_tg == 0 || t.adj_tg == want.adj_tg) &&
                (want.mfma_tg == 0 || t.mfma_tg == want.mfma_tg))
                return kOk;
        Status st3 = time_shape(want, got);
        if (st3 != kOk) return st3;
        if (got.ms < 0.97 * best.ms) best = got;
        return kOk;
    };
    for (int per : {2, 4, 6}) {
        RingShape w;
        w.per = per;
        if ((st = consider(w)) != kOk) return st;
    }
    RingShape w;
    w.per = best.per;
    if (best.mfma_tg == 0 && B <= 2 && best.adj_tg == 2) w.adj_tg = 1;  // the single-transform adjoint
    else if (best.mfma_tg == 2) w.mfma_tg = 1;                         // the MFMA adjoint
    else if (best.mfma_tg == 1) w.mfma_tg = 2;
    if ((w.adj_tg || w.mfma_tg) && (st = consider(w)) != kOk) return st;
    return kOk;
}

// The shape a plan's ring engine is built with: the cached one for its key, else (mode >= 0, and an
// engine large enough for the timing to matter) the fastest timed candidate, recorded in the cache;
// mode 1 times even when the key is cached, mode -1 never times.
static Status choose_ring_shape(pfsht_plan_s &p, const double *theta, int n, int L, double eps, int B, bool reflect,
                                int mode, RingShape &shape);

// SpinRings (src/Spin.jl:192-270) on the device: the half-angle logarithms per ring (swapped, and
// cos theta negated, for a reflected ring set, so the spin weight stays as given), retained orders
// max(polar_mmax, |s|) per ring, the seed tables per order (src/Spin.jl:233-257), the recurrence
// table per (|m|, l) built by a kernel, and one forward wavefront per (order, 64-ring tile).
static Status rings_build_spin(const double *theta, int n, int L, int spin, double eps, int B, bool reflect,
                               pfsht_rings_s &r) {
    if (L < 0) { set_error("lmax must be nonnegative"); return kEArg; }
    if (B < 1) { set_error("ntrans must be >= 1"); return kEArg; }
    if (n < 1) { set_error("need at least one ring"); return kEArg; }
    const int as = spin < 0 ? -spin : spin;
    if (spin == 0) { set_error("spin 0 uses the scalar ring engine"); return kEArg; }
    if (as > L) { set_error("(|spin| = " + std::to_string(as) + ") > lmax = " + std::to_string(L)); return kEArg; }
    for (int j = 1; j < n; ++j)
        if (!(theta[j] >= theta[j - 1])) { set_error("ring colatitudes must be sorted ascending"); return kEArg; }
    (void)hipGetDevice(&r.dev);
    r.spin = spin;
    const double csgn = reflect ? -1.0 : 1.0;
    std::vector<double> costh(n), l2c(n), l2s(n);
    std::vector<int> mlim(n);
    int mmax = 0;
    for (int j = 0; j < n; ++j) {
        const double t = theta[j];
        costh[j] = csgn * std::cos(t);
        double ch = std::cos(t / 2), sh = std::fabs(std::sin(t / 2));
        if (reflect) std::swap(ch, sh);
        l2c[j] = ch > 0 ? std::log2(ch) : -INFINITY;
        l2s[j] = sh > 0 ? std::log2(sh) : -INFINITY;
        mlim[j] = std::max(polar_mmax(t, L, eps), as);
        mmax = std::max(mmax, mlim[j]);
    }
    std::string msg;
    if (!nested_ranges(mlim, L, r.jlo, r.jhi, msg)) { set_error(msg); return kEArg; }
    r.mlim = mlim;
    // seeds: l0 = max(|m|, |s|); the extremal index moves to second place through
    // d^l_{a,b} = (-1)^(a-b) d^l_{b,a}, leaving a = m (|s| >= |m|) or a = -s
    const int nn = -spin;
    std::vector<double4> seed(2 * L + 1);
    const double inv_ln2 = 1.0 / std::log(2.0);
    for (int m = -L; m <= L; ++m) {
        const int am = std::abs(m), l0 = std::max(am, as);
        int a, b;
        double extra, sgn;
        if (as == l0) { a = m; b = nn == l0 ? l0 : -l0; extra = 1.0; }
        else { a = nn; b = m == l0 ? l0 : -l0; extra = ((nn - m) & 1) ? -1.0 : 1.0; }
        int pc, ps;
        if (b == l0) { pc = l0 + a; ps = l0 - a; sgn = 1.0; }
        else { pc = l0 - a; ps = l0 + a; sgn = ((l0 - a) & 1) ? -1.0 : 1.0; }
        const double lb = (std::lgamma(double(2 * l0 + 1)) - std::lgamma(double(l0 - a + 1)) -
                           std::lgamma(double(l0 + a + 1))) / 2;
        seed[m + L] = make_double4(lb * inv_ln2, sgn * extra, double(pc), double(ps));
    }
    std::vector<double> Nl(L + 1);
    for (int l = 0; l <= L; ++l) Nl[l] = std::sqrt(double(2 * l + 1) / (4 * kPi));
    std::vector<int64_t> tofs(L + 1);
    int64_t ntab = 0;
    for (int am = 0; am <= L; ++am) { tofs[am] = ntab; ntab += L - std::max(am, as); }
    std::vector<int4> items;
    r.work = 0;
    for (int m = -mmax; m <= mmax; ++m) {
        const int am = std::abs(m), lo = r.jlo[am], hi = r.jhi[am];
        if (hi < lo) continue;
        r.work += int64_t(hi - lo + 1) * (L - std::max(am, as) + 1);
        for (int j0 = lo, t = 0; j0 <= hi; j0 += 64, ++t) items.push_back(make_int4(m, j0, std::min(64, hi - j0 + 1), t));
    }
    // adjoint partials: per order, tiles x retained degrees x transforms
    std::vector<int64_t> spofs(2 * mmax + 1, 0);
    std::vector<int32_t> sntiles(2 * mmax + 1, 0);
    int64_t nsp = 0;
    for (int m = -mmax; m <= mmax; ++m) {
        const int am = std::abs(m), lo = r.jlo[am], hi = r.jhi[am];
        spofs[m + mmax] = nsp;
        if (hi < lo) continue;
        sntiles[m + mmax] = (hi - lo + 64) / 64;
        nsp += int64_t(sntiles[m + mmax]) * (L - std::max(am, as) + 1) * B;
    }
    RingDev &d = r.d;
    SpinDev &sd = r.sd;
    d.n = n; d.L = L; d.B = B; d.mmax = mmax;
    sd.s = spin;
    sd.nitems = int(items.size());
    std::vector<int32_t> jlo32(r.jlo.begin(), r.jlo.begin() + mmax + 1), jhi32(r.jhi.begin(), r.jhi.begin() + mmax + 1);
    Status st;
    int64_t &by = r.bytes;
    if ((st = upload(&d.costh, costh, by)) != kOk) return st;
    if ((st = upload(&d.jlo, jlo32, by)) != kOk) return st;
    if ((st = upload(&d.jhi, jhi32, by)) != kOk) return st;
    if ((st = upload(&sd.l2c, l2c, by)) != kOk) return st;
    if ((st = upload(&sd.l2s, l2s, by)) != kOk) return st;
    if ((st = upload(&sd.seed, seed, by)) != kOk) return st;
    if ((st = upload(&sd.Nl, Nl, by)) != kOk) return st;
    if ((st = upload(&sd.tofs, tofs, by)) != kOk) return st;
    if ((st = upload(&sd.items, items, by)) != kOk) return st;
    if ((st = dalloc(&sd.tab, size_t(std::max<int64_t>(ntab, 1)), by)) != kOk) return st;
    if ((st = upload(&sd.pofs, spofs, by)) != kOk) return st;
    if ((st = upload(&sd.ntiles, sntiles, by)) != kOk) return st;
    if ((st = dalloc(&sd.partials, size_t(std::max<int64_t>(nsp, 1)), by)) != kOk) return st;
    r.part_bytes = int64_t(sizeof(double2)) * std::max<int64_t>(nsp, 1);
    if ((st = dalloc(&r.Gd, size_t(B) * (2 * mmax + 1) * n, by)) != kOk) return st;
    launch_spin_tables(d, sd, nullptr);
    PFSHT_HIP_TRY(hipGetLastError());
    PFSHT_HIP_TRY(hipStreamSynchronize(nullptr));
    return kOk;
}

// The ring passes of an engine of either kind over nb transforms.
static void ring_forward(pfsht_rings_s &r, InCols alm, int nb, hipStream_t s) {
    if (r.spin) launch_spin_fwd(r.d, r.sd, alm, nb, r.Gd, s);
    else launch_ring_fwd(r.d, alm, nb, r.Gd, s);
}
static void ring_adjoint(pfsht_rings_s &r, OutCols alm, int nb, hipStream_t s) {
    if (r.spin) launch_spin_adj(r.d, r.sd, r.Gd, alm, nb, s);
    else launch_ring_adj(r.d, r.Gd, alm, nb, s);
}

// ------------------------------------------------------------------------------------------------
// Plan
// ------------------------------------------------------------------------------------------------
namespace {
constexpr int kStages = PFSHT_NSTAGES;
// Events in execution order; interval k (ev[k] -> ev[k+1]) is stage stage_of[k].
// Boundaries with no work in between are not recorded (an event costs a few microseconds of device
// time): recorded[k] says whether ev[k] was; an interval ending at an unrecorded boundary is empty.
struct TimedApply {
    int dir;
    hipEvent_t ev[kStages + 1];
    bool recorded[kStages + 1];
    int stage_of[kStages];
};

// An event that marks the end of a plan's last apply: the plan's own, or a plan group's, shared by
// every member of the group (one record per group apply rather than one per member).
struct SharedEvent {
    hipEvent_t e = nullptr;
    SharedEvent() = default;
    SharedEvent(const SharedEvent &) = delete;
    SharedEvent &operator=(const SharedEvent &) = delete;
    ~SharedEvent() {
        if (e) (void)hipEventDestroy(e);
    }
};
using EventRef = std::shared_ptr<SharedEvent>;

static Status event_make(EventRef &r) {
    auto ev = std::make_shared<SharedEvent>();
    PFSHT_HIP_TRY(hipEventCreateWithFlags(&ev->e, hipEventDisableTiming));
    r = std::move(ev);
    return kOk;
}
}  // namespace

struct pfsht_plan_s {
    int dev = 0;
    int method = PFSHT_METHOD_STRETCH, precision = PFSHT_PREC_FP64;
    int spin = 0;  // spin weight (src/Plan.jl:87); coefficients in the dense (l, m + lmax) layout when nonzero
    int L = 0, B = 1;
    int64_t M = 0;
    double tol = 0;
    Geom geom;
    int flip = 0, nh = 0, nth = 0, nph = 0, mmax = 0;
    pfsht_rings_s *rings = nullptr;
    TuneKey tkey;                        // the ring engine's tuning key (L = 0: not a tuned engine)
    std::vector<RingShape> tune_table;   // the candidates timed when the plan was built
    double *wth = nullptr;
    NufftDev nu;
    // Ft: the periodised grid, either in order-column layout [b][m + mmax][i] (the apply) or
    // [b][i][k] (the NUFFT seam); U: the fine grid [b][l1][l2]; Ut: the nonzero fine columns in
    // order-column layout [b][m + mmax][l1]. Elements are double2, or float2 in a mixed-precision
    // plan (the NUFFT seam is FP64 only).
    void *Ft = nullptr, *U = nullptr, *Ut = nullptr;
    // The grid stage: the LDS kernels of grid_fft_kernels.hip (fused; Ut holds their T1 / T2), or,
    // for lengths past a workgroup's LDS, rocFFT plans of the colatitude, fine-column and fine-row
    // FFTs with the transposes and deconvolution kernels between them.
    bool fused = false;
    GridFftDev gf;
    FftSet th_fwd, th_bwd, col_bwd, col_fwd, row_bwd, row_fwd;
    void *fft_work = nullptr;    // rocFFT work buffer shared by those FFTs
    FftPlan fine_bwd, fine_fwd;  // the NUFFT seam's 2-D transforms (all transforms), made at first use
    void *seam_work = nullptr;   // their work buffer
    bool partial = false;        // FFT plans for every power-of-two transform count (packed applies)
    double2 *alm_stage = nullptr, *f_stage = nullptr, *modes_stage = nullptr;
    float2 *c64_stage = nullptr;  // ComplexF32 applies on host memory: input, then output
    // Direct leaf: theta-sorted points as rings
    double *dphi = nullptr;
    int32_t *dperm = nullptr, *dmlim = nullptr;
    int64_t bytes = 0;
    std::mutex mu;
    // Applies of one plan share its work buffers, so they are ordered across streams: every apply
    // leaves `last` marking its end on `last_stream` (the plan's own `done`, recorded by its apply,
    // or the event a plan group records once for all its members), and an apply on another stream
    // first waits for it.
    EventRef done, last;
    hipStream_t last_stream = nullptr;
    // multi-box plan (MultiBoxLeaf, src/MultiBox.jl:20-73): one stretched plan per box, the
    // points of each box (device indices into the caller's field), a box's field, and the
    // accumulator the adjoint sums the boxes' coefficients through
    std::vector<pfsht_plan_s *> boxes;
    std::vector<int32_t *> box_idx;
    std::vector<int64_t> box_npts;
    double2 *mb_f = nullptr, *mb_acc = nullptr;
    // analysis workspace (AnalysisWorkspace, src/Analysis.jl:45-84), allocated at the first analysis
    // call: rhs, r, p, Ap and a staging alm in the alm layout, two fields, weights, the Wiener
    // prior, the spectrum, dot-product partials and a pinned host slot for scalars
    double2 *aw_rhs = nullptr, *aw_r = nullptr, *aw_p = nullptr, *aw_Ap = nullptr, *aw_x = nullptr;
    double2 *aw_f = nullptr, *aw_f2 = nullptr, *aw_part = nullptr, *aw_dot = nullptr, *aw_hdot = nullptr;
    double *aw_w = nullptr, *aw_reg = nullptr, *aw_C = nullptr;
    double th_min = 0, th_max = 0;  // colatitude range of the points
    bool box_ready = false;
    Box box;
    // Column tables of pfsht_synth_cols / pfsht_adjoint_cols: kColSlots slots of 2 B device pointers
    // (the inputs, then the outputs), each filled from pinned host memory by an asynchronous copy on
    // the apply's stream; a slot is reused only once the apply that read it has finished.
    static constexpr int kColSlots = 8;
    void **col_dev = nullptr, **col_host = nullptr;
    hipEvent_t col_ev[kColSlots] = {};
    bool col_used[kColSlots] = {};
    int col_next = 0;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> pool;
    std::vector<TimedApply> pending;
    double ms[2][kStages] = {};
    int64_t counts[2][kStages] = {};

    ~pfsht_plan_s() {
        delete rings;
        dfree(wth);
        dfree(nu.X1); dfree(nu.X2); dfree(nu.perm); dfree(nu.pc1); dfree(nu.c1); dfree(nu.c2);
        dfree(nu.st_list); dfree(nu.st_slot); dfree(nu.bin_ofs); dfree(nu.Wst);
        dfree(Ft); dfree(U); dfree(Ut); dfree(alm_stage); dfree(f_stage); dfree(modes_stage); dfree(c64_stage);
        dfree(dphi); dfree(dperm); dfree(dmlim);
        dfree(aw_rhs); dfree(aw_r); dfree(aw_p); dfree(aw_Ap); dfree(aw_x); dfree(aw_f); dfree(aw_f2);
        dfree(aw_part); dfree(aw_dot); dfree(aw_w); dfree(aw_reg); dfree(aw_C);
        if (aw_hdot) (void)hipHostFree(aw_hdot);
        for (auto *q : boxes) delete q;
        for (auto *q : box_idx) dfree(q);
        dfree(mb_f); dfree(mb_acc);
        fftset_free(th_fwd); fftset_free(th_bwd);
        fft_free(fine_bwd); fft_free(fine_fwd); fftset_free(col_bwd); fftset_free(col_fwd);
        fftset_free(row_bwd); fftset_free(row_fwd);
        dfree(fft_work); dfree(seam_work); dfree(gf.tw);
        for (auto &t : pending) for (auto e : t.ev) pool.push_back(e);
        for (auto e : pool) (void)hipEventDestroy(e);
        dfree(col_dev);
        if (col_host) (void)hipHostFree(col_host);
        for (auto e : col_ev)
            if (e) (void)hipEventDestroy(e);
    }
};

static Status cols_alloc(pfsht_plan_s &p) {
    const size_t n = size_t(pfsht_plan_s::kColSlots) * 2 * p.B;
    Status st;
    if ((st = dalloc(&p.col_dev, n, p.bytes)) != kOk) return st;
    PFSHT_HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&p.col_host), n * sizeof(void *)));
    for (auto &e : p.col_ev) PFSHT_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return kOk;
}

// Orders an apply on stream s after the plan's previous apply (on any stream).
static Status apply_begin(pfsht_plan_s &p, hipStream_t s) {
    if (p.last && s != p.last_stream) PFSHT_HIP_TRY(hipStreamWaitEvent(s, p.last->e, 0));
    return kOk;
}
static Status apply_end(pfsht_plan_s &p, hipStream_t s) {
    Status st;
    if (!p.done && (st = event_make(p.done)) != kOk) return st;
    PFSHT_HIP_TRY(hipEventRecord(p.done->e, s));
    p.last = p.done;
    p.last_stream = s;
    return kOk;
}

static Status choose_ring_shape(pfsht_plan_s &p, const double *theta, int n, int L, double eps, int B, bool reflect,
                                int mode, RingShape &shape) {
    TuneKey key;
    key.arch = device_arch(p.dev);
    key.L = L;
    key.n = n;
    for (int j = 0; j < n; ++j) key.mmax = std::max(key.mmax, polar_mmax(theta[j], L, eps));
    key.B = B;
    p.tkey = key;
    tune_load_env();
    if (mode != 1 && tune_lookup(key, shape)) return kOk;
    if (mode < 0 || double(n) * L * L < 2e6) {  // small engines keep the default: timing costs more than it saves
        shape = RingShape();
        return kOk;
    }
    RingShape best;
    Status st = tune_rings(theta, n, L, eps, B, reflect, best, p.tune_table);
    if (st != kOk) return st;
    tune_store(key, best);
    shape = best;
    return kOk;
}

static Status plan_build(pfsht_plan_s &p, const double *theta, const double *phi, int64_t M, int L, double tol, int B,
                         const pfsht_plan_opts &o) {
    int method = o.method;
    const int spin = o.spin;
    const int maxboxes = o.maxboxes;
    p.partial = o.packed != 0 && B > 1;
    if (M < 1) { set_error("need at least one point"); return kEArg; }
    if (M > INT32_MAX) { set_error("more than 2^31-1 points per plan is not supported"); return kEArg; }
    if (L < 0) { set_error("lmax must be nonnegative"); return kEArg; }
    if (B < 1) { set_error("ntrans must be >= 1"); return kEArg; }
    if (!(tol > 0) || !(tol < 1)) { set_error("tol must lie in (0, 1)"); return kEArg; }
    if (spin != 0 && std::abs(spin) > L) {  // src/Plan.jl:101-102
        set_error("(|spin| = " + std::to_string(std::abs(spin)) + ") > lmax = " + std::to_string(L));
        return kEArg;
    }
    p.spin = spin;
    (void)hipGetDevice(&p.dev);
    p.L = L; p.B = B; p.M = M; p.tol = tol;
    // bounding box in colatitude (src/Box.jl:43-73)
    double thlo = INFINITY, thhi = -INFINITY;
    for (int64_t i = 0; i < M; ++i) { thlo = std::min(thlo, theta[i]); thhi = std::max(thhi, theta[i]); }
    p.th_min = thlo;
    p.th_max = thhi;
    thlo = std::max(thlo, 0.0);
    thhi = std::min(thhi, kPi);
    Partition part;
    bool parted = false;
    if (method == PFSHT_METHOD_AUTO) {  // select_method, src/Cost.jl:174-201
        const double cd = cost_direct(M, L, B), cs = cost_stretch(thlo, thhi, M, L, B, tol);
        double best = cd;
        method = PFSHT_METHOD_DIRECT;
        if (cs < best) { method = PFSHT_METHOD_STRETCH; best = cs; }
        part = partition_boxes(theta, phi, M, L, B, tol, maxboxes);
        parted = true;
        if (part.groups.size() > 1) {
            double cm = 0;
            for (double c : part.costs) cm += c;
            if (cm < best) method = PFSHT_METHOD_MULTIBOX;
        }
    }
    p.method = method;
    if (method == PFSHT_METHOD_MULTIBOX) {
        // MultiBoxLeaf (src/MultiBox.jl:30-62): a stretched plan per box of the partition
        if (!parted) part = partition_boxes(theta, phi, M, L, B, tol, maxboxes);
        p.geom.kind = -2;
        int64_t &by = p.bytes, maxg = 0;
        for (const auto &g : part.groups) {
            const int64_t Mg = int64_t(g.size());
            std::vector<double> tg(Mg), pg(Mg);
            for (int64_t j = 0; j < Mg; ++j) { tg[j] = theta[g[j]]; pg[j] = phi[g[j]]; }
            auto *q = new (std::nothrow) pfsht_plan_s;
            if (!q) { set_error("host allocation failed"); return kENomem; }
            q->precision = p.precision;
            p.boxes.push_back(q);
            pfsht_plan_opts ob = o;
            ob.method = PFSHT_METHOD_STRETCH;
            Status st = plan_build(*q, tg.data(), pg.data(), Mg, L, tol, B, ob);
            if (st != kOk) return st;
            int32_t *idx = nullptr;
            if ((st = upload(&idx, g, by)) != kOk) return st;
            p.box_idx.push_back(idx);
            p.box_npts.push_back(Mg);
            p.mmax = std::max(p.mmax, q->mmax);
            maxg = std::max(maxg, Mg);
        }
        Status st;
        if ((st = dalloc(&p.mb_f, size_t(maxg) * B, by)) != kOk) return st;
        if ((st = dalloc(&p.mb_acc, size_t(L + 1) * (2 * L + 1) * B, by)) != kOk) return st;
        PFSHT_HIP_TRY(hipDeviceSynchronize());
        return kOk;
    }
    if (method == PFSHT_METHOD_DIRECT) {
        // DirectLeaf (src/Direct.jl:25-51): points sorted by colatitude (stable), ring engine at
        // eps = 0 so polar_mmax keeps every order that is not below floatmin.
        std::vector<int32_t> perm(M);
        for (int64_t i = 0; i < M; ++i) perm[i] = int32_t(i);
        std::stable_sort(perm.begin(), perm.end(), [theta](int32_t a, int32_t b) { return theta[a] < theta[b]; });
        std::vector<double> ths(M), phs(M);
        for (int64_t i = 0; i < M; ++i) { ths[i] = theta[perm[i]]; phs[i] = phi[perm[i]]; }
        p.rings = new pfsht_rings_s;
        Status st = spin ? rings_build_spin(ths.data(), int(M), L, spin, 0.0, B, false, *p.rings)
                         : rings_build(ths.data(), int(M), L, 0.0, B, false, *p.rings);
        if (st != kOk) return st;
        p.mmax = p.rings->d.mmax;
        p.nh = int(M);
        std::vector<int32_t> ml(p.rings->mlim.begin(), p.rings->mlim.end());
        int64_t &by = p.bytes;
        if ((st = upload(&p.dphi, phs, by)) != kOk) return st;
        if ((st = upload(&p.dperm, perm, by)) != kOk) return st;
        if ((st = upload(&p.dmlim, ml, by)) != kOk) return st;
        p.geom.kind = -1;
        PFSHT_HIP_TRY(hipDeviceSynchronize());
        return kOk;
    }
    if (method != PFSHT_METHOD_STRETCH) { set_error("unknown method"); return kEArg; }
    // geometry and southern reflection (src/Stretch.jl:151-158)
    std::string msg;
    const double tiny = std::sqrt(DBL_EPSILON);
    Geom g0;
    int rc = stretch_geometry(thlo, thhi, L, tol, g0, msg, o.epsapo_max);
    if (rc != 0) { set_error(msg); return rc < 0 ? kEArg : kEGeom; }
    p.flip = g0.kind == kPole && g0.thhi >= kPi - tiny && g0.thlo > tiny;
    if (p.flip) {
        rc = stretch_geometry(kPi - thhi, kPi - thlo, L, tol, p.geom, msg, o.epsapo_max);
        if (rc != 0) { set_error(msg); return rc < 0 ? kEArg : kEGeom; }
    } else {
        p.geom = g0;
    }
    // sizing (src/Stretch.jl:164-171)
    const int Kt = int(std::ceil(double(p.geom.K + kKMargin) * o.oversample_theta));  // src/Stretch.jl:164
    if (p.geom.kind == kInterior) { p.nth = int(next_smooth(2 * Kt + 1)); p.nh = p.nth; }
    else { p.nh = int(next_smooth(Kt + 1)); p.nth = 2 * p.nh; }
    std::vector<double> th, Th;
    stretch_rings(p.geom, p.nh, th, Th);
    if (!std::is_sorted(th.begin(), th.end())) { set_error("ring colatitudes are not ascending"); return kEArg; }
    RingShape shape;
    if (!spin) {
        Status st0 = choose_ring_shape(p, th.data(), p.nh, L, tol, B, p.flip != 0, o.tune, shape);
        if (st0 != kOk) return st0;
    }
    p.rings = new pfsht_rings_s;
    Status st = spin ? rings_build_spin(th.data(), p.nh, L, spin, tol, B, p.flip != 0, *p.rings)
                     : rings_build(th.data(), p.nh, L, tol, B, p.flip != 0, *p.rings, shape);
    if (st != kOk) return st;
    p.mmax = p.rings->d.mmax;
    p.nph = int(next_smooth_even(int64_t(std::ceil(double(2 * int64_t(std::max(p.mmax, std::abs(spin))) + 1) *
                                                    o.oversample_phi))));  // src/Stretch.jl:194
    int64_t &by = p.bytes;
    std::vector<double> wth(p.nh);
    for (int i = 0; i < p.nh; ++i) wth[i] = stretch_window(p.geom, Th[i]);
    if ((st = upload(&p.wth, wth, by)) != kOk) return st;

    // NUFFT: kernel, fine grid, bin-sorted points (src/Stretch.jl:202-213, 235-238)
    NufftDev &u = p.nu;
    u.B = B; u.n1 = p.nth; u.n2 = p.nph; u.M = M;
    es_params(tol, u.w, u.beta);
    u.N1 = es_fine_size(p.nth, u.w);
    u.N2 = es_fine_size(p.nph, u.w);
    u.S = 17 - u.w;
    const int S4 = 4 * u.S;
    u.ns1 = (u.N1 + S4 - 1) / S4;
    u.ns2 = (u.N2 + S4 - 1) / S4;
    u.win = 3 * u.S + 16;  // the point kernels size their per-thread staging from this relation
    {
        int maxlds = 65536;
        (void)hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, p.dev);
        if (spread_lds_bytes(u.win, kSpreadGroup) > size_t(maxlds) || 2 * nufft_window_lds_bytes(u.win) > size_t(maxlds)) {
            set_error("the NUFFT super-tile window does not fit in LDS at this tolerance (tol >= 1e-3)");
            return kEUnsupported;
        }
    }
    // Points onto the fine grid, bin keys and counts, the reached rows: on the device, then a
    // stable sort of the keys (points of a bin keep their input order) and the gather of the
    // coordinates into bin order.
    const int nkeys = u.ns1 * u.ns2 * 16;
    std::vector<int32_t> cnt32(nkeys);
    std::vector<uint8_t> on(u.N1);
    {
        PointMap q;
        q.kind = p.geom.kind;
        q.flip = p.flip;
        q.north = p.geom.thlo <= std::sqrt(DBL_EPSILON);
        q.thmax = p.geom.h * (1 + p.geom.epsapo);
        q.lo = p.geom.thlo - p.geom.h * p.geom.epsapo;
        q.period = p.geom.period;
        q.s1 = double(u.N1) / (2 * kPi);
        q.s2 = double(u.N2) / (2 * kPi);
        q.half = 0.5 * u.w;
        q.N1 = u.N1; q.N2 = u.N2; q.S = u.S; q.S4 = S4; q.ns2 = u.ns2; q.w = u.w;
        int bits = 1;
        while ((int64_t(1) << bits) < nkeys) ++bits;
        DevScratch tmp;
        double *dth = nullptr, *dph = nullptr, *X1u = nullptr, *X2u = nullptr;
        int32_t *key = nullptr, *skey = nullptr, *idx = nullptr, *dcnt = nullptr, *err = nullptr;
        uint8_t *don = nullptr;
        int64_t tb = 0;
        if ((st = tmp.take(&dth, M, tb)) != kOk || (st = tmp.take(&dph, M, tb)) != kOk ||
            (st = tmp.take(&X1u, M, tb)) != kOk || (st = tmp.take(&X2u, M, tb)) != kOk ||
            (st = tmp.take(&key, M, tb)) != kOk || (st = tmp.take(&skey, M, tb)) != kOk ||
            (st = tmp.take(&idx, M, tb)) != kOk || (st = tmp.take(&dcnt, nkeys, tb)) != kOk ||
            (st = tmp.take(&err, 1, tb)) != kOk || (st = tmp.take(&don, u.N1, tb)) != kOk)
            return st;
        if ((st = dalloc(&u.X1, M, by)) != kOk) return st;
        if ((st = dalloc(&u.X2, M, by)) != kOk) return st;
        if ((st = dalloc(&u.perm, M, by)) != kOk) return st;
        PFSHT_HIP_TRY(hipMemcpyAsync(dth, theta, sizeof(double) * M, hipMemcpyHostToDevice, nullptr));
        PFSHT_HIP_TRY(hipMemcpyAsync(dph, phi, sizeof(double) * M, hipMemcpyHostToDevice, nullptr));
        PFSHT_HIP_TRY(hipMemsetAsync(dcnt, 0, sizeof(int32_t) * nkeys, nullptr));
        PFSHT_HIP_TRY(hipMemsetAsync(don, 0, u.N1, nullptr));
        PFSHT_HIP_TRY(hipMemsetAsync(err, 0, sizeof(int32_t), nullptr));
        launch_map_points(dth, dph, M, q, X1u, X2u, key, idx, dcnt, don, err, nullptr);
        PFSHT_HIP_TRY(hipGetLastError());
        size_t sb = 0;
        PFSHT_HIP_TRY(sort_pairs(nullptr, sb, key, skey, idx, u.perm, int(M), bits, nullptr));
        void *sortmem = nullptr;
        if ((st = tmp.take(reinterpret_cast<char **>(&sortmem), sb, tb)) != kOk) return st;
        PFSHT_HIP_TRY(sort_pairs(sortmem, sb, key, skey, idx, u.perm, int(M), bits, nullptr));
        launch_gather_points(X1u, X2u, u.perm, M, u.X1, u.X2, nullptr);
        PFSHT_HIP_TRY(hipGetLastError());
        int32_t herr = 0;
        PFSHT_HIP_TRY(hipMemcpy(&herr, err, sizeof(int32_t), hipMemcpyDeviceToHost));
        PFSHT_HIP_TRY(hipMemcpy(cnt32.data(), dcnt, sizeof(int32_t) * nkeys, hipMemcpyDeviceToHost));
        PFSHT_HIP_TRY(hipMemcpy(on.data(), don, u.N1, hipMemcpyDeviceToHost));
        if (herr) { set_error("point maps outside the fine grid"); return kEArg; }
    }
    std::vector<int64_t> cnt;
    {
        std::string bmsg;
        if (!bin_ranges(cnt32, M, cnt, bmsg)) {  // every sorted position lies in one bin range < M
            set_error("point binning is inconsistent: " + bmsg);
            return kEHip;
        }
    }
    // Fine-grid rows the kernels reach: the apply stores and transforms only the smallest cyclic
    // interval holding them (about half of the rows for a polar cap, whose points fill half the
    // doubled colatitude period).
    {
        int best = 0, best_end = 0;  // longest cyclic run of unreached rows, and the row after it
        for (int pass = 0, run = 0; pass < 2; ++pass)
            for (int r = 0; r < u.N1; ++r) {
                run = on[r] ? 0 : std::min(run + 1, u.N1);
                if (run > best) { best = run; best_end = (r + 1) % u.N1; }
            }
        if (best >= u.N1) best = 0;  // no point at all cannot happen (M >= 1)
        u.r0 = best > 0 ? best_end : 0;
        u.nR = u.N1 - best;
    }
    // Non-empty super-tiles, and per (slot, local bin) the start of its point range: a super-tile's
    // bins are contiguous in the sorted order, so bin q's end is the next bin's start, and the
    // super-tile's end follows its 16 starts.
    std::vector<int32_t> st_list, st_slot(size_t(u.ns1) * u.ns2, -1);
    for (int s = 0; s < u.ns1 * u.ns2; ++s)
        if (cnt[int64_t(s) * 16 + 16] != cnt[int64_t(s) * 16]) st_list.push_back(s);
    // Slots in order of decreasing point count (ties by super-tile): the spreading and interpolation
    // workgroups then start with the heaviest super-tiles and the light ones fill in behind them,
    // instead of the densest rows of a cap arriving last as a tail. Per slot, bin_ofs holds the 16
    // bin starts and the super-tile's end (slots are no longer consecutive in point order).
    std::stable_sort(st_list.begin(), st_list.end(), [&](int32_t a, int32_t b) {
        return cnt[int64_t(a) * 16 + 16] - cnt[int64_t(a) * 16] > cnt[int64_t(b) * 16 + 16] - cnt[int64_t(b) * 16];
    });
    for (size_t slot = 0; slot < st_list.size(); ++slot) st_slot[st_list[slot]] = int32_t(slot);
    std::vector<int32_t> bo(std::max<size_t>(st_list.size(), 1) * 17);
    for (size_t slot = 0; slot < st_list.size(); ++slot)
        for (int lb = 0; lb <= 16; ++lb) bo[slot * 17 + lb] = int32_t(cnt[int64_t(st_list[slot]) * 16 + lb]);
    u.nst = int(st_list.size());
    // pc1[i] = phase[i] * h1/psihat1(k_i), phase = cis(-k shift)/(nth nph) (src/Stretch.jl:229-233)
    std::vector<double> c1, c2;
    es_corr(p.nth, u.N1, u.w, u.beta, c1);
    es_corr(p.nph, u.N2, u.w, u.beta, c2);
    const double shift = p.geom.kind == kInterior ? 0.0 : kPi / double(p.nth);
    const double Nn = double(p.nth) * double(p.nph);
    std::vector<double2> pc1(p.nth), c1c(p.nth);
    for (int i = 0; i < p.nth; ++i) {
        const double ang = -double(mode_k(i, p.nth)) * shift;
        pc1[i] = make_double2(std::cos(ang) / Nn * c1[i], std::sin(ang) / Nn * c1[i]);
        c1c[i] = make_double2(c1[i], 0.0);
    }
    if ((st = upload(&u.pc1, pc1, by)) != kOk) return st;
    if ((st = upload(&u.c1, c1c, by)) != kOk) return st;
    if ((st = upload(&u.c2, c2, by)) != kOk) return st;
    if ((st = upload(&u.st_list, st_list, by)) != kOk) return st;
    if ((st = upload(&u.st_slot, st_slot, by)) != kOk) return st;
    if ((st = upload(&u.bin_ofs, bo, by)) != kOk) return st;
    const bool single = p.precision == PFSHT_PREC_MIXED;
    const size_t csz = single ? sizeof(float2) : sizeof(double2);
    auto galloc = [&](void **q, size_t count) { return dalloc(reinterpret_cast<char **>(q), count * csz, by); };
    if ((st = galloc(&u.Wst, size_t(B) * std::max(u.nst, 1) * u.win * u.win)) != kOk) return st;
    // work buffers: the fine grid's stored rows (the NUFFT seam uses all N1 rows), and the nonzero
    // fine columns in order-column layout (T1 / T2 of the LDS grid stage); the NUFFT seam's
    // periodised grid Ft is made at its first use when the LDS grid stage runs the apply
    if ((st = galloc(&p.U, size_t(B) * u.N1 * u.N2)) != kOk) return st;
    const size_t ncol1 = size_t(2 * p.mmax + 1);
    if ((st = galloc(&p.Ut, size_t(B) * ncol1 * u.N1)) != kOk) return st;
    p.fused = o.grid_fft != PFSHT_GRID_FFT_LIBRARY && grid_fits(p.nth, u.N1, u.N2, csz, p.dev, &p.gf);
    if (o.grid_fft == PFSHT_GRID_FFT_LDS && !p.fused) {
        set_error("the grid lengths (" + std::to_string(p.nth) + ", " + std::to_string(u.N1) + ", " +
                  std::to_string(u.N2) + ") do not fit the LDS grid stage");
        return kEUnsupported;
    }
    if (p.fused) {
        GridFftDev &g = p.gf;
        g.nt_col = grid_fft_threads(u.N1);
        g.nt_row = grid_fft_threads(g.split ? u.N2 / 2 : u.N2);
        g.args.nh = p.nh; g.args.nth = p.nth; g.args.mmax = p.mmax; g.args.spin = p.spin; g.args.nph = p.nph;
        g.args.N1 = u.N1; g.args.N2 = u.N2; g.args.r0 = u.r0; g.args.nR = u.nR;
        g.args.scale = double(p.nph);
        // exp(-2 pi i k / n), k < n, for the lengths back to back (N2 / 2 last, when rows split)
        g.th.tw = 0;
        g.c1.tw = p.nth;
        g.r2.tw = p.nth + u.N1;
        g.r2h.tw = p.nth + u.N1 + u.N2;
        const size_t ntw = size_t(p.nth) + u.N1 + u.N2 + (g.split ? u.N2 / 2 : 0);
        std::vector<double2> tw(ntw);
        std::vector<const FftLen *> lens = {&g.th, &g.c1, &g.r2};
        if (g.split) lens.push_back(&g.r2h);
        for (const FftLen *f : lens)
            for (int k = 0; k < f->n; ++k) {
                const long double a = -2.0L * 3.141592653589793238462643383279502884L * k / f->n;
                tw[size_t(f->tw) + k] = make_double2(double(cosl(a)), double(sinl(a)));
            }
        if (single) {
            std::vector<float2> tw32(ntw);
            for (size_t k = 0; k < ntw; ++k) tw32[k] = make_float2(float(tw[k].x), float(tw[k].y));
            float2 *d = nullptr;
            if ((st = upload(&d, tw32, by)) != kOk) return st;
            g.tw = d;
        } else {
            double2 *d = nullptr;
            if ((st = upload(&d, tw, by)) != kOk) return st;
            g.tw = d;
        }
    } else {
        // rocFFT plans (row-major, contiguous; lengths fastest first): colatitude FFTs of the order
        // columns (the longitude FFT pair of the reference cancels); the pruned fine transform:
        // colatitude FFTs of the nonzero columns, longitude FFTs of the stored rows
        if ((st = galloc(&p.Ft, size_t(B) * p.nth * p.nph)) != kOk) return st;  // nph >= 2 mmax + 1
        if ((st = fftset_make(p.th_fwd, true, p.nth, ncol1, B, p.partial, single)) != kOk) return st;
        if ((st = fftset_make(p.th_bwd, false, p.nth, ncol1, B, p.partial, single)) != kOk) return st;
        if ((st = fftset_make(p.col_bwd, false, u.N1, ncol1, B, p.partial, single)) != kOk) return st;
        if ((st = fftset_make(p.col_fwd, true, u.N1, ncol1, B, p.partial, single)) != kOk) return st;
        if ((st = fftset_make(p.row_bwd, false, u.N2, size_t(u.nR), B, p.partial, single)) != kOk) return st;
        if ((st = fftset_make(p.row_fwd, true, u.N2, size_t(u.nR), B, p.partial, single)) != kOk) return st;
        std::vector<FftPlan *> all;
        for (FftSet *fs : {&p.th_fwd, &p.th_bwd, &p.col_bwd, &p.col_fwd, &p.row_bwd, &p.row_fwd})
            for (auto &q : fs->plans) all.push_back(&q);
        if ((st = fft_attach_work(all, p.fft_work, by)) != kOk) return st;
    }
    PFSHT_HIP_TRY(hipDeviceSynchronize());
    return kOk;
}

// The NUFFT seam's periodised grid and 2-D rocFFT plans, made at its first call.
static Status seam_ready(pfsht_plan_s &p) {
    Status st;
    if (!p.Ft && (st = dalloc(reinterpret_cast<double2 **>(&p.Ft), size_t(p.B) * p.nth * p.nph, p.bytes)) != kOk)
        return st;
    if (p.fine_bwd.plan) return kOk;
    if ((st = fft_make(p.fine_bwd, false, p.nu.N2, p.nu.N1, p.B)) != kOk) return st;
    if ((st = fft_make(p.fine_fwd, true, p.nu.N2, p.nu.N1, p.B)) != kOk) return st;
    return fft_attach_work({&p.fine_bwd, &p.fine_fwd}, p.seam_work, p.bytes);
}

static GridDims dims_of(const pfsht_plan_s &p, int nb) {
    GridDims g;
    g.B = nb; g.nh = p.nh; g.nth = p.nth; g.nph = p.nph; g.mmax = p.mmax; g.L = p.L; g.kind = p.geom.kind;
    g.spin = p.spin;
    return g;
}

// stage timing: events recorded on the apply stream, drained by pfsht_stage_times
static void tmark(pfsht_plan_s &p, TimedApply *t, int k, hipStream_t s) {
    if (t) t->recorded[k] = hipEventRecord(t->ev[k], s) == hipSuccess;
}

static TimedApply *timing_begin(pfsht_plan_s &p, int dir) {
    if (!p.timing) return nullptr;
    TimedApply t;
    t.dir = dir;
    for (int k = 0; k <= kStages; ++k) t.recorded[k] = false;
    // synthesis runs stages 0..7 in order (8 is empty); the adjoint runs them transposed, 7..0,
    // then the ring partial reduction, 8
    for (int k = 0; k < kStages; ++k) t.stage_of[k] = (dir == 0 || k == kStages - 1) ? k : kStages - 2 - k;
    for (int k = 0; k <= kStages; ++k) {
        if (p.pool.empty()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            p.pool.push_back(e);
        }
        t.ev[k] = p.pool.back();
        p.pool.pop_back();
    }
    p.pending.push_back(t);
    return &p.pending.back();
}

static void timing_drain(pfsht_plan_s &p) {
    for (auto &t : p.pending) {
        int last = kStages;
        while (last > 0 && !t.recorded[last]) --last;
        (void)hipEventSynchronize(t.ev[last]);
        int from = 0;  // last recorded boundary
        for (int k = 0; k < kStages; ++k) {
            float ms = 0;
            if (t.recorded[k + 1]) {
                if (hipEventElapsedTime(&ms, t.ev[from], t.ev[k + 1]) != hipSuccess) ms = 0;
                from = k + 1;
            }
            p.ms[t.dir][t.stage_of[k]] += ms;
            p.counts[t.dir][t.stage_of[k]] += 1;
        }
        for (auto e : t.ev) p.pool.push_back(e);
    }
    p.pending.clear();
}

// The grid part of the stretched synthesis, from the ring buffer to the points, with the grid
// buffers stored as C (stages 1-7), for nb transforms.
template <class C>
static Status synth_grid(pfsht_plan_s &p, TimedApply *t, OutCols f, int nb, hipStream_t s, bool interp = true) {
    const GridDims g = dims_of(p, nb);
    NufftDev u = p.nu;
    u.B = nb;
    C *Ft = static_cast<C *>(p.Ft), *Ut = static_cast<C *>(p.Ut), *U = static_cast<C *>(p.U);
    if (p.fused) {  // stages 1-5 in the column kernel, the fine longitude FFT in the row kernel
        launch_grid_synth_cols(p.gf, p.rings->d, p.rings->Gd, p.wth, u.pc1, u.c2, Ut, nb, s);
        tmark(p, t, 5, s);
        launch_grid_synth_rows(p.gf, static_cast<const C *>(Ut), U, nb, s);
        tmark(p, t, 7, s);
        if (interp) launch_interp(u, static_cast<const C *>(U), f, s);
        tmark(p, t, 8, s);
        return kOk;
    }
    launch_spec_dfs(g, p.rings->d, p.rings->Gd, p.wth, Ft, s);
    tmark(p, t, 2, s);  // stages 2 and 3 (longitude FFT, DFS) have no work of their own
    Status st = fftset_exec(p.th_fwd, Ft, nb, s);
    if (st != kOk) return st;
    tmark(p, t, 5, s);
    launch_deconv_pad_ut(u, u.pc1, p.mmax, p.nph, double(p.nph), static_cast<const C *>(Ft), Ut, s);
    tmark(p, t, 6, s);
    if ((st = fftset_exec(p.col_bwd, Ut, nb, s)) != kOk) return st;
    launch_ut_to_rows(u, p.mmax, static_cast<const C *>(Ut), U, s);
    if ((st = fftset_exec(p.row_bwd, U, nb, s)) != kOk) return st;
    tmark(p, t, 7, s);
    if (interp) launch_interp(u, static_cast<const C *>(U), f, s);
    tmark(p, t, 8, s);  // stage 8 is the adjoint's
    return kOk;
}

// The grid part of the stretched adjoint, from the points to the ring buffer (stages 1-6).
template <class C>
static Status adjoint_grid(pfsht_plan_s &p, TimedApply *t, InCols f, int nb, hipStream_t s, bool spread = true) {
    const GridDims g = dims_of(p, nb);
    NufftDev u = p.nu;
    u.B = nb;
    C *Ft = static_cast<C *>(p.Ft), *Ut = static_cast<C *>(p.Ut), *U = static_cast<C *>(p.U);
    if (spread) launch_spread(u, f, U, s);
    tmark(p, t, 1, s);
    if (p.fused) {  // the fine longitude FFT in the row kernel, stages 2-6 in the column kernel
        launch_grid_adjoint_rows(p.gf, static_cast<const C *>(U), Ut, nb, s);
        tmark(p, t, 2, s);
        launch_grid_adjoint_cols(p.gf, static_cast<const C *>(Ut), p.wth, u.pc1, u.c2, p.rings->Gd, nb, s);
        tmark(p, t, 4, s);
        return kOk;
    }
    Status st = fftset_exec(p.row_fwd, U, nb, s);
    if (st != kOk) return st;
    launch_rows_to_ut(u, p.mmax, static_cast<const C *>(U), Ut, s);
    if ((st = fftset_exec(p.col_fwd, Ut, nb, s)) != kOk) return st;
    tmark(p, t, 2, s);
    launch_deconv_trunc_ut(u, u.pc1, p.mmax, p.nph, double(p.nph), static_cast<const C *>(Ut), Ft, s);
    tmark(p, t, 3, s);
    if ((st = fftset_exec(p.th_bwd, Ft, nb, s)) != kOk) return st;
    tmark(p, t, 4, s);
    launch_spec_fold(g, static_cast<const C *>(Ft), p.wth, p.rings->Gd, s);
    return kOk;
}

static Status synth_cols(pfsht_plan_s &p, InCols alm, OutCols f, int nb, hipStream_t s);
static Status adjoint_cols(pfsht_plan_s &p, InCols f, OutCols alm, int nb, hipStream_t s);

static int64_t alm_len(const pfsht_plan_s &p) { return int64_t(p.L + 1) * (2 * p.L + 1); }

// synthesize_multibox! / adjoint_multibox! (src/MultiBox.jl:83-112): every point belongs to one
// box, so the synthesis writes each of f's entries once; the adjoint adds the boxes' coefficients
// in box order.
static Status multibox_synth(pfsht_plan_s &p, InCols alm, OutCols f, int nb, hipStream_t s) {
    for (size_t g = 0; g < p.boxes.size(); ++g) {
        const int64_t Mg = p.box_npts[g];
        Status st = synth_cols(*p.boxes[g], alm, out_cols(p.mb_f, Mg), nb, s);
        if (st != kOk) return st;
        launch_scatter_field(f, p.mb_f, p.box_idx[g], Mg, nb, s);
    }
    PFSHT_HIP_TRY(hipGetLastError());
    return kOk;
}

static Status multibox_adjoint(pfsht_plan_s &p, InCols f, OutCols alm, int nb, hipStream_t s) {
    const int64_t n = alm_len(p);
    for (size_t g = 0; g < p.boxes.size(); ++g) {
        const int64_t Mg = p.box_npts[g];
        launch_gather_field(p.mb_f, f, p.box_idx[g], Mg, nb, s);
        Status st = adjoint_cols(*p.boxes[g], in_cols(p.mb_f, Mg), g == 0 ? alm : out_cols(p.mb_acc, n), nb, s);
        if (st != kOk) return st;
        if (g > 0) launch_cols_add(alm, p.mb_acc, n, nb, s);
    }
    PFSHT_HIP_TRY(hipGetLastError());
    return kOk;
}

static Status synth_body(pfsht_plan_s &p, InCols alm, OutCols f, int nb, hipStream_t s) {
    if (p.method == PFSHT_METHOD_MULTIBOX) return multibox_synth(p, alm, f, nb, s);
    if (p.timing && p.pending.size() >= 256) timing_drain(p);
    TimedApply *t = timing_begin(p, 0);
    if (p.method == PFSHT_METHOD_DIRECT) {  // src/Direct.jl:60-91
        pfsht_rings_s &r = *p.rings;
        tmark(p, t, 0, s);
        ring_forward(r, alm, nb, s);
        tmark(p, t, 1, s);
        launch_direct_synth(r.Gd, p.dphi, p.dmlim, p.dperm, f, p.M, p.mmax, nb, s);
        tmark(p, t, 2, s);  // the phase sums are timed as stage 1
        PFSHT_HIP_TRY(hipGetLastError());
        return kOk;
    }
    // src/Stretch.jl:262-277: ring pass, window + DFS (in longitude-spectral space), colatitude FFT,
    // phase and deconvolution into the fine grid, fine inverse FFT, interpolation
    tmark(p, t, 0, s);
    ring_forward(*p.rings, alm, nb, s);
    tmark(p, t, 1, s);
    Status st = p.precision == PFSHT_PREC_MIXED ? synth_grid<float2>(p, t, f, nb, s)
                                                : synth_grid<double2>(p, t, f, nb, s);
    if (st != kOk) return st;
    PFSHT_HIP_TRY(hipGetLastError());
    return kOk;
}

// The adjoint ring pass as timed stages: the pass (stage 0 of the adjoint order), then the partial
// reduction (stage 8, empty on the MFMA path and for spin engines, which finish in one launch).
static void ring_adjoint_timed(pfsht_plan_s &p, TimedApply *t, OutCols alm, int nb, hipStream_t s) {
    pfsht_rings_s &r = *p.rings;
    if (r.spin) {
        launch_spin_adj(r.d, r.sd, r.Gd, alm, nb, s);
        tmark(p, t, 8, s);
    } else {
        launch_ring_adj_pass(r.d, r.Gd, alm, nb, s);
        tmark(p, t, 8, s);
        launch_ring_adj_finish(r.d, alm, nb, s);
    }
    tmark(p, t, 9, s);
}

static Status adjoint_body(pfsht_plan_s &p, InCols f, OutCols alm, int nb, hipStream_t s) {
    if (p.method == PFSHT_METHOD_MULTIBOX) return multibox_adjoint(p, f, alm, nb, s);
    if (p.timing && p.pending.size() >= 256) timing_drain(p);
    TimedApply *t = timing_begin(p, 1);
    if (p.method == PFSHT_METHOD_DIRECT) {  // src/Direct.jl:94-117
        pfsht_rings_s &r = *p.rings;
        tmark(p, t, 0, s);
        launch_direct_adjoint(f, p.dphi, p.dmlim, p.dperm, r.Gd, p.M, p.mmax, nb, s);
        tmark(p, t, 7, s);  // the phase products are timed as stage 1 (adjoint order)
        ring_adjoint_timed(p, t, alm, nb, s);
        PFSHT_HIP_TRY(hipGetLastError());
        return kOk;
    }
    // src/Stretch.jl:279-295, the exact transpose of the synthesis
    tmark(p, t, 0, s);
    Status st = p.precision == PFSHT_PREC_MIXED ? adjoint_grid<float2>(p, t, f, nb, s)
                                                : adjoint_grid<double2>(p, t, f, nb, s);
    if (st != kOk) return st;
    if (!p.fused) tmark(p, t, 7, s);  // the LDS column kernel ends with the fold (stage 1)
    ring_adjoint_timed(p, t, alm, nb, s);
    PFSHT_HIP_TRY(hipGetLastError());
    return kOk;
}

static Status synth_cols(pfsht_plan_s &p, InCols alm, OutCols f, int nb, hipStream_t s) {
    Status st = apply_begin(p, s);
    if (st == kOk) st = synth_body(p, alm, f, nb, s);
    if (st == kOk) st = apply_end(p, s);
    return st;
}

static Status adjoint_cols(pfsht_plan_s &p, InCols f, OutCols alm, int nb, hipStream_t s) {
    Status st = apply_begin(p, s);
    if (st == kOk) st = adjoint_body(p, f, alm, nb, s);
    if (st == kOk) st = apply_end(p, s);
    return st;
}

// All B transforms in the reference's contiguous layouts.
static Status synth_impl(pfsht_plan_s &p, const double2 *alm, double2 *f, hipStream_t s) {
    return synth_cols(p, in_cols(alm, alm_len(p)), out_cols(f, p.M), p.B, s);
}

static Status adjoint_impl(pfsht_plan_s &p, const double2 *f, double2 *alm, hipStream_t s) {
    return adjoint_cols(p, in_cols(f, p.M), out_cols(alm, alm_len(p)), p.B, s);
}

// ------------------------------------------------------------------------------------------------
// Analysis on a plan (src/Analysis.jl)
// ------------------------------------------------------------------------------------------------
static size_t alm_count(const pfsht_plan_s &p) { return size_t(p.L + 1) * (2 * p.L + 1) * p.B; }

static Status aw_alloc(pfsht_plan_s &p) {
    if (p.aw_rhs) return kOk;
    const size_t na = alm_count(p), nf = size_t(p.M) * p.B;
    int64_t &by = p.bytes;
    Status st;
    for (double2 **q : {&p.aw_rhs, &p.aw_r, &p.aw_p, &p.aw_Ap, &p.aw_x})
        if ((st = dalloc(q, na, by)) != kOk) return st;
    if ((st = dalloc(&p.aw_f, nf, by)) != kOk) return st;
    if ((st = dalloc(&p.aw_f2, nf, by)) != kOk) return st;
    if ((st = dalloc(&p.aw_part, size_t(kDotBlocks), by)) != kOk) return st;
    if ((st = dalloc(&p.aw_dot, 1, by)) != kOk) return st;
    if ((st = dalloc(&p.aw_w, size_t(p.M), by)) != kOk) return st;
    if ((st = dalloc(&p.aw_reg, size_t(p.L + 1), by)) != kOk) return st;
    if ((st = dalloc(&p.aw_C, size_t(p.L + 1), by)) != kOk) return st;
    PFSHT_HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&p.aw_hdot), sizeof(double2)));
    return kOk;
}

// The points' bounding box, computed once: longitudes read back from the plan's device copy
// (the Direct leaf's sorted phi, or the stretched leaf's fine-grid x2 scaled back to radians).
static Status aw_box(pfsht_plan_s &p) {
    if (p.box_ready) return kOk;
    std::vector<double> ph(p.M);
    if (p.method == PFSHT_METHOD_DIRECT) {
        PFSHT_HIP_TRY(hipMemcpy(ph.data(), p.dphi, sizeof(double) * p.M, hipMemcpyDeviceToHost));
    } else if (p.method == PFSHT_METHOD_MULTIBOX) {
        int64_t o = 0;
        for (auto *q : p.boxes) {
            PFSHT_HIP_TRY(hipMemcpy(ph.data() + o, q->nu.X2, sizeof(double) * q->M, hipMemcpyDeviceToHost));
            const double sc = 2 * kPi / double(q->nu.N2);
            for (int64_t j = 0; j < q->M; ++j) ph[o + j] *= sc;
            o += q->M;
        }
    } else {
        PFSHT_HIP_TRY(hipMemcpy(ph.data(), p.nu.X2, sizeof(double) * p.M, hipMemcpyDeviceToHost));
        const double sc = 2 * kPi / double(p.nu.N2);
        for (auto &v : ph) v *= sc;
    }
    p.box = bounding_box_range(p.th_min, p.th_max, ph.data(), p.M);
    p.box_ready = true;
    return kOk;
}

// LinearAlgebra.dot(a, b) = sum conj(a) b over n complex entries, read back to the host
static Status cdot(pfsht_plan_s &p, const double2 *a, const double2 *b, size_t n, hipStream_t s,
                   std::complex<double> &out) {
    launch_cdot(a, b, int64_t(n), p.aw_part, p.aw_dot, s);
    PFSHT_HIP_TRY(hipGetLastError());
    PFSHT_HIP_TRY(hipMemcpyAsync(p.aw_hdot, p.aw_dot, sizeof(double2), hipMemcpyDeviceToHost, s));
    PFSHT_HIP_TRY(hipStreamSynchronize(s));
    out = std::complex<double>(p.aw_hdot->x, p.aw_hdot->y);
    return kOk;
}

// The caller's weights as a device pointer (staged for host calls), or nullptr for none.
static Status weights_dev(pfsht_plan_s &p, const double *w, int32_t where, hipStream_t s, const double *&wd) {
    wd = nullptr;
    if (!w) return kOk;
    if (where == PFSHT_WHERE_DEVICE) { wd = w; return kOk; }
    PFSHT_HIP_TRY(hipMemcpyAsync(p.aw_w, w, sizeof(double) * p.M, hipMemcpyHostToDevice, s));
    wd = p.aw_w;
    return kOk;
}

// A field argument in a workspace buffer: a copy of the caller's f with the weights applied.
static Status field_in(pfsht_plan_s &p, const void *f, const double *wd, int32_t where, double2 *dst, hipStream_t s) {
    const size_t bytes = sizeof(double2) * size_t(p.M) * p.B;
    PFSHT_HIP_TRY(hipMemcpyAsync(dst, f, bytes, where == PFSHT_WHERE_DEVICE ? hipMemcpyDeviceToDevice
                                                                              : hipMemcpyHostToDevice, s));
    if (wd) launch_weight_field(dst, wd, p.M, p.B, s);
    return kOk;
}

static Status analysis_check(pfsht_plan_s *p, int32_t where) {
    if (!p) { set_error("plan is NULL"); return kEArg; }
    if (where != PFSHT_WHERE_HOST && where != PFSHT_WHERE_DEVICE) { set_error("bad where"); return kEArg; }
    return kOk;
}

// ------------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------------
extern "C" {

int pfsht_bounding_box(const double *theta, const double *phi, int64_t npts, pfsht_box *out) {
    if (!theta || !phi || !out) { set_error("NULL argument"); return kEArg; }
    if (npts < 1) { set_error("need at least one point"); return kEArg; }
    const Box b = bounding_box(theta, phi, npts);
    out->thlo = b.thlo;
    out->thhi = b.thhi;
    out->phistart = b.phistart;
    out->phiext = b.phiext;
    out->touches_north = b.north;
    out->touches_south = b.south;
    out->fsky = box_fsky(b);
    out->dl = spectral_resolution(b);
    return kOk;
}

int pfsht_partition_boxes(const double *theta, const double *phi, int64_t npts, int32_t lmax, int32_t ntrans,
                          double tol, int32_t maxboxes, int32_t *group_of, int32_t *nboxes) {
    if (!theta || !phi || !group_of || !nboxes) { set_error("NULL argument"); return kEArg; }
    if (npts < 1) { set_error("need at least one point"); return kEArg; }
    if (npts > INT32_MAX) { set_error("more than 2^31-1 points is not supported"); return kEArg; }
    if (lmax < 0 || ntrans < 1 || maxboxes < 1 || !(tol > 0) || !(tol < 1)) { set_error("bad argument"); return kEArg; }
    try {
        const Partition P = partition_boxes(theta, phi, npts, lmax, ntrans, tol, maxboxes);
        for (size_t g = 0; g < P.groups.size(); ++g)
            for (int32_t i : P.groups[g]) group_of[i] = int32_t(g);
        *nboxes = int32_t(P.groups.size());
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return kENomem;
    }
    return kOk;
}

int pfsht_apply_resolution(pfsht_plan_t p, const void *alm, const double *w, void *out, int32_t where,
                           void *stream) {
    Status st = analysis_check(p, where);
    if (st != kOk) return st;
    if (!alm || !out) { set_error("NULL argument"); return kEArg; }
    std::lock_guard<std::mutex> lk(p->mu);
    DeviceGuard dg(p->dev);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if ((st = aw_alloc(*p)) != kOk) return st;
    const size_t abytes = sizeof(double2) * alm_count(*p);
    const double2 *a = static_cast<const double2 *>(alm);
    double2 *o = static_cast<double2 *>(out);
    if (where == PFSHT_WHERE_HOST) {
        PFSHT_HIP_TRY(hipMemcpyAsync(p->aw_x, alm, abytes, hipMemcpyHostToDevice, s));
        a = p->aw_x;
        o = p->aw_Ap;
    }
    const double *wd;
    if ((st = weights_dev(*p, w, where, s, wd)) != kOk) return st;
    // src/Analysis.jl:96-104: f = A alm; f .*= w; out = A' f
    if ((st = synth_impl(*p, a, p->aw_f, s)) != kOk) return st;
    if (wd) launch_weight_field(p->aw_f, wd, p->M, p->B, s);
    if ((st = adjoint_impl(*p, p->aw_f, o, s)) != kOk) return st;
    if (where == PFSHT_WHERE_HOST) PFSHT_HIP_TRY(hipMemcpyAsync(out, o, abytes, hipMemcpyDeviceToHost, s));
    PFSHT_HIP_TRY(hipStreamSynchronize(s));
    return kOk;
}

int pfsht_pseudo_spectrum(pfsht_plan_t p, const void *f, const double *w, int32_t batch, double *C, int32_t where,
                          void *stream) {
    Status st = analysis_check(p, where);
    if (st != kOk) return st;
    if (!f || !C) { set_error("NULL argument"); return kEArg; }
    if (p->spin) { set_error("the pseudo-spectrum reads the sph_mode layout; a spin plan uses (l, m + lmax)"); return kEUnsupported; }
    if (batch < 1 || batch > p->B) {
        set_error("batch = " + std::to_string(batch) + " outside 1:" + std::to_string(p->B));
        return kEArg;
    }
    std::lock_guard<std::mutex> lk(p->mu);
    DeviceGuard dg(p->dev);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if ((st = aw_alloc(*p)) != kOk) return st;
    const double *wd;
    if ((st = weights_dev(*p, w, where, s, wd)) != kOk) return st;
    // src/Analysis.jl:286-308: a~ = A' (w .* f), then C_l over the orders of transform `batch`
    if ((st = field_in(*p, f, wd, where, p->aw_f, s)) != kOk) return st;
    if ((st = adjoint_impl(*p, p->aw_f, p->aw_x, s)) != kOk) return st;
    double *Cd = where == PFSHT_WHERE_DEVICE ? C : p->aw_C;
    launch_spectrum(p->aw_x + size_t(batch - 1) * (size_t(p->L + 1) * (2 * p->L + 1)), p->L, Cd, s);
    PFSHT_HIP_TRY(hipGetLastError());
    if (where == PFSHT_WHERE_HOST)
        PFSHT_HIP_TRY(hipMemcpyAsync(C, Cd, sizeof(double) * (p->L + 1), hipMemcpyDeviceToHost, s));
    PFSHT_HIP_TRY(hipStreamSynchronize(s));
    return kOk;
}

int pfsht_solve(pfsht_plan_t p, const void *f, const double *w, int32_t mode, double rtol, int32_t maxiter,
                const double *prior, void *alm, int32_t where, void *stream, pfsht_solve_info *info) {
    Status st = analysis_check(p, where);
    if (st != kOk) return st;
    if (!f || !alm) { set_error("NULL argument"); return kEArg; }
    if (mode != PFSHT_SOLVE_CG && mode != PFSHT_SOLVE_WIENER && mode != PFSHT_SOLVE_MINNORM) {
        set_error("mode must be CG, WIENER or MINNORM");
        return kEArg;
    }
    if (mode == PFSHT_SOLVE_WIENER && p->spin) {
        set_error("the Wiener prior reads the sph_mode layout; a spin plan uses (l, m + lmax)");
        return kEUnsupported;
    }
    if (mode == PFSHT_SOLVE_WIENER && !prior) {
        set_error("WIENER needs a prior C_l: that prior is what makes the problem well posed, so it cannot be "
                  "defaulted");
        return kEArg;
    }
    if (maxiter < 0) { set_error("maxiter must be nonnegative"); return kEArg; }
    std::lock_guard<std::mutex> lk(p->mu);
    DeviceGuard dg(p->dev);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if ((st = aw_alloc(*p)) != kOk) return st;
    if ((st = aw_box(*p)) != kOk) return st;
    const size_t n = alm_count(*p), abytes = sizeof(double2) * n;
    double2 *x = where == PFSHT_WHERE_DEVICE ? static_cast<double2 *>(alm) : p->aw_x;
    const double *wd;
    if ((st = weights_dev(*p, w, where, s, wd)) != kOk) return st;
    // src/Analysis.jl:391-409: rhs = A' (w .* f); reg = 1 / prior (0 where prior <= 0)
    if ((st = field_in(*p, f, wd, where, p->aw_f2, s)) != kOk) return st;
    if ((st = adjoint_impl(*p, p->aw_f2, p->aw_rhs, s)) != kOk) return st;
    const bool wiener = mode == PFSHT_SOLVE_WIENER;
    if (wiener) {
        std::vector<double> reg(p->L + 1);
        for (int l = 0; l <= p->L; ++l) reg[l] = prior[l] > 0 ? 1.0 / prior[l] : 0.0;
        PFSHT_HIP_TRY(hipMemcpyAsync(p->aw_reg, reg.data(), sizeof(double) * reg.size(), hipMemcpyHostToDevice, s));
        PFSHT_HIP_TRY(hipStreamSynchronize(s));
    }
    const double tol = mode == PFSHT_SOLVE_MINNORM ? std::min(rtol, 1e-12) : rtol;
    // src/Analysis.jl:401-436: CG from alm = 0
    PFSHT_HIP_TRY(hipMemsetAsync(x, 0, abytes, s));
    PFSHT_HIP_TRY(hipMemcpyAsync(p->aw_r, p->aw_rhs, abytes, hipMemcpyDeviceToDevice, s));
    PFSHT_HIP_TRY(hipMemcpyAsync(p->aw_p, p->aw_r, abytes, hipMemcpyDeviceToDevice, s));
    std::complex<double> rsold, rr;
    if ((st = cdot(*p, p->aw_r, p->aw_r, n, s, rsold)) != kOk) return st;
    if ((st = cdot(*p, p->aw_rhs, p->aw_rhs, n, s, rr)) != kOk) return st;
    const double rhsn = std::sqrt(std::abs(rr));
    pfsht_solve_info inf{};
    inf.shannon = double(p->L + 1) * double(p->L + 1) * box_fsky(p->box);
    inf.dl = spectral_resolution(p->box);
    inf.mode = mode;
    int iters = 0;
    double rel = 1.0;
    if (rhsn == 0) {
        rel = 0.0;
    } else {
        for (int i = 1; i <= maxiter; ++i) {
            iters = i;
            if ((st = synth_impl(*p, p->aw_p, p->aw_f, s)) != kOk) return st;
            if (wd) launch_weight_field(p->aw_f, wd, p->M, p->B, s);
            if ((st = adjoint_impl(*p, p->aw_f, p->aw_Ap, s)) != kOk) return st;
            if (wiener) launch_add_prior(p->aw_Ap, p->aw_p, p->aw_reg, p->L, p->B, s);
            std::complex<double> pAp, rsnew;
            if ((st = cdot(*p, p->aw_p, p->aw_Ap, n, s, pAp)) != kOk) return st;
            const std::complex<double> alpha = pAp == 0.0 ? std::complex<double>(0.0) : rsold / pAp;
            launch_zaxpy(x, p->aw_p, make_double2(alpha.real(), alpha.imag()), int64_t(n), s);
            launch_zaxpy(p->aw_r, p->aw_Ap, make_double2(-alpha.real(), -alpha.imag()), int64_t(n), s);
            if ((st = cdot(*p, p->aw_r, p->aw_r, n, s, rsnew)) != kOk) return st;
            rel = std::sqrt(std::abs(rsnew)) / rhsn;
            if (rel <= tol) break;
            const std::complex<double> beta = rsnew / rsold;
            launch_zxpby(p->aw_p, p->aw_r, make_double2(beta.real(), beta.imag()), int64_t(n), s);
            rsold = rsnew;
        }
    }
    PFSHT_HIP_TRY(hipGetLastError());
    if (where == PFSHT_WHERE_HOST) PFSHT_HIP_TRY(hipMemcpyAsync(alm, x, abytes, hipMemcpyDeviceToHost, s));
    PFSHT_HIP_TRY(hipStreamSynchronize(s));
    inf.iters = iters;
    inf.relres = rel;
    if (info) *info = inf;
    return kOk;
}


int pfsht_plan_tune_table(pfsht_plan_t p, pfsht_ring_shape *rows, int32_t cap, int32_t *n) {
    if (!p || !n || (cap > 0 && !rows)) { set_error("NULL argument"); return kEArg; }
    const int k = int(std::min<size_t>(p->tune_table.size(), size_t(std::max(cap, 0))));
    for (int i = 0; i < k; ++i) {
        const RingShape &t = p->tune_table[i];
        rows[i].seg_blocks = t.per;
        rows[i].adj_tiles = t.adj_tg;
        rows[i].mfma_tiles = t.mfma_tg;
        rows[i].reserved = 0;
        rows[i].ms = t.ms;
    }
    *n = k;
    return kOk;
}

int pfsht_plan_tune_key(pfsht_plan_t p, char *arch, int32_t cap, int32_t *lmax, int32_t *nrings, int32_t *mmax,
                        int32_t *ntrans) {
    if (!p || !arch || cap < 1 || !lmax || !nrings || !mmax || !ntrans) { set_error("NULL argument"); return kEArg; }
    std::snprintf(arch, size_t(cap), "%s", p->tkey.arch.c_str());
    *lmax = p->tkey.L;
    *nrings = p->tkey.n;
    *mmax = p->tkey.mmax;
    *ntrans = p->tkey.B;
    return kOk;
}

const char *pfsht_last_error(void) { return g_error.c_str(); }
const char *pfsht_version(void) { return "pfsht-hip 0.1.0 (gfx950)"; }

int pfsht_plan_create(const double *theta, const double *phi, int64_t npts, int32_t lmax, double tol, int32_t ntrans,
                      int32_t method, int32_t precision, pfsht_plan_t *out) {
    return pfsht_plan_create_spin(theta, phi, npts, lmax, tol, ntrans, 0, method, precision, out);
}

int pfsht_plan_create_spin(const double *theta, const double *phi, int64_t npts, int32_t lmax, double tol,
                           int32_t ntrans, int32_t spin, int32_t method, int32_t precision, pfsht_plan_t *out) {
    pfsht_plan_opts o;
    pfsht_plan_opts_default(&o);
    o.method = method;
    o.precision = precision;
    o.spin = spin;
    return pfsht_plan_create_opts(theta, phi, npts, lmax, tol, ntrans, &o, out);
}

void pfsht_plan_opts_default(pfsht_plan_opts *o) {
    if (o) std::memset(o, 0, sizeof *o);
}

}  // extern "C"

// The caller's options with the reference's defaults for the keywords left at zero (src/Plan.jl:84-97),
// checked.
static Status opts_normalise(const pfsht_plan_opts *opts, double tol, pfsht_plan_opts &o) {
    pfsht_plan_opts_default(&o);
    if (opts) o = *opts;
    if (o.maxboxes == 0) o.maxboxes = kMaxBoxes;
    if (o.oversample_theta == 0) o.oversample_theta = 1.0;
    if (o.oversample_phi == 0) o.oversample_phi = 1.0;
    if (o.epsapo_max == 0) o.epsapo_max = kEpsApoMax;
    if (o.precision != PFSHT_PREC_FP64 && o.precision != PFSHT_PREC_MIXED) { set_error("unknown precision"); return kEArg; }
    if (o.precision == PFSHT_PREC_MIXED && !(tol >= 1e-6)) {
        set_error("mixed precision needs tol >= 1e-6");
        return kEArg;
    }
    if (o.method < PFSHT_METHOD_AUTO || o.method > PFSHT_METHOD_MULTIBOX) { set_error("unknown method"); return kEArg; }
    if (o.maxboxes < 1) { set_error("maxboxes must be >= 1"); return kEArg; }
    if (!(o.oversample_theta >= 1) || !(o.oversample_phi >= 1) || !(o.oversample_theta <= 16) ||
        !(o.oversample_phi <= 16)) {
        set_error("oversample_theta and oversample_phi must lie in [1, 16]");
        return kEArg;
    }
    if (!(o.epsapo_max > 0) || !(o.epsapo_max <= 1)) { set_error("epsapo_max must lie in (0, 1]"); return kEArg; }
    if (o.tune < -1 || o.tune > 1) { set_error("tune must be -1, 0 or 1"); return kEArg; }
    if (o.grid_fft < PFSHT_GRID_FFT_AUTO || o.grid_fft > PFSHT_GRID_FFT_LIBRARY) {
        set_error("grid_fft must be 0 (auto), 1 (LDS) or 2 (library)");
        return kEArg;
    }
    return kOk;
}

namespace pfsht {
int64_t plan_estimate_bytes(const double *theta, const double *phi, int64_t M, int L, double tol, int B,
                            const pfsht_plan_opts &o);
}

extern "C" {

int pfsht_plan_estimate_bytes(const double *theta, const double *phi, int64_t npts, int32_t lmax, double tol,
                              int32_t ntrans, const pfsht_plan_opts *opts, int64_t *bytes) {
    if (!theta || !phi || !bytes) { set_error("NULL argument"); return kEArg; }
    if (npts < 1 || lmax < 0 || ntrans < 1 || !(tol > 0) || !(tol < 1)) { set_error("bad argument"); return kEArg; }
    pfsht_plan_opts o;
    Status st = opts_normalise(opts, tol, o);
    if (st != kOk) return st;
    try {
        *bytes = pfsht::plan_estimate_bytes(theta, phi, npts, lmax, tol, ntrans, o);
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return kENomem;
    }
    if (*bytes < 0) { set_error("the plan's size could not be estimated"); return kEArg; }
    return kOk;
}

int pfsht_plan_create_opts(const double *theta, const double *phi, int64_t npts, int32_t lmax, double tol,
                           int32_t ntrans, const pfsht_plan_opts *opts, pfsht_plan_t *out) {
    if (!out) { set_error("out is NULL"); return kEArg; }
    *out = nullptr;
    if (!theta || !phi) { set_error("theta/phi is NULL"); return kEArg; }
    pfsht_plan_opts o;
    Status st0 = opts_normalise(opts, tol, o);
    if (st0 != kOk) return st0;
    auto *p = new (std::nothrow) pfsht_plan_s;
    if (!p) { set_error("host allocation failed"); return kENomem; }
    p->precision = o.precision;
    Status st;
    try {
        st = plan_build(*p, theta, phi, npts, lmax, tol, ntrans, o);
        if (st == kOk) st = cols_alloc(*p);
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        st = kENomem;
    }
    if (st != kOk) {
        delete p;
        return st;
    }
    *out = p;
    return kOk;
}

void pfsht_plan_destroy(pfsht_plan_t p) {
    if (!p) return;
    DeviceGuard dg(p->dev);
    (void)hipDeviceSynchronize();
    delete p;
}

int pfsht_plan_info(pfsht_plan_t p, pfsht_info *info) {
    if (!p || !info) { set_error("NULL argument"); return kEArg; }
    std::memset(info, 0, sizeof *info);
    info->method = p->method;
    info->kind = p->method == PFSHT_METHOD_DIRECT ? -1 : p->geom.kind;
    info->nboxes = p->method == PFSHT_METHOD_MULTIBOX ? int32_t(p->boxes.size()) : 1;
    info->spin = p->spin;
    info->flip = p->flip;
    info->lmax = p->L;
    info->ntrans = p->B;
    info->precision = p->precision;
    info->nth_half = p->nh;
    info->nth = p->nth;
    info->nph = p->nph;
    info->mmax = p->mmax;
    info->es_w = p->nu.w;
    info->es_n1 = p->nu.N1;
    info->es_n2 = p->nu.N2;
    info->npts = p->M;
    info->ring_work = p->rings ? p->rings->work : 0;
    if (p->rings) {
        info->seg_blocks = p->rings->shape.per;
        info->adj_tiles = p->rings->shape.adj_tg;
        info->mfma_tiles = p->rings->shape.mfma_tg;
    }
    info->device_bytes = p->bytes + (p->rings ? p->rings->bytes : 0);
    if (p->method == PFSHT_METHOD_STRETCH) info->grid_fft = p->fused ? PFSHT_GRID_FFT_LDS : PFSHT_GRID_FFT_LIBRARY;
    for (auto *q : p->boxes) {
        info->ring_work += q->rings ? q->rings->work : 0;
        info->device_bytes += q->bytes + (q->rings ? q->rings->bytes : 0);
        info->grid_fft = std::max(info->grid_fft, q->fused ? int32_t(PFSHT_GRID_FFT_LDS) : int32_t(PFSHT_GRID_FFT_LIBRARY));
    }
    info->tol = p->tol;
    info->es_beta = p->nu.beta;
    info->es_rows = p->nu.nR;
    return kOk;
}

static Status stage_in(void **stage, size_t bytes, const void *src, hipStream_t s, int64_t &acct) {
    if (!*stage) {
        Status st = dalloc(reinterpret_cast<char **>(stage), bytes, acct);
        if (st != kOk) return st;
    }
    if (src) PFSHT_HIP_TRY(hipMemcpyAsync(*stage, src, bytes, hipMemcpyHostToDevice, s));
    return kOk;
}

int pfsht_synth(pfsht_plan_t p, const void *alm, void *f, int32_t where, void *stream) {
    if (!p || !alm || !f) { set_error("NULL argument"); return kEArg; }
    if (where != PFSHT_WHERE_HOST && where != PFSHT_WHERE_DEVICE) { set_error("bad where"); return kEArg; }
    std::lock_guard<std::mutex> lk(p->mu);
    DeviceGuard dg(p->dev);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const size_t alm_bytes = sizeof(double2) * size_t(p->L + 1) * (2 * p->L + 1) * p->B;
    const size_t f_bytes = sizeof(double2) * size_t(p->M) * p->B;
    if (where == PFSHT_WHERE_DEVICE) return synth_impl(*p, static_cast<const double2 *>(alm), static_cast<double2 *>(f), s);
    Status st;
    if ((st = apply_begin(*p, s)) != kOk) return st;  // the staging buffers belong to the plan too
    if ((st = stage_in(reinterpret_cast<void **>(&p->alm_stage), alm_bytes, alm, s, p->bytes)) != kOk) return st;
    if ((st = stage_in(reinterpret_cast<void **>(&p->f_stage), f_bytes, nullptr, s, p->bytes)) != kOk) return st;
    if ((st = synth_impl(*p, p->alm_stage, p->f_stage, s)) != kOk) return st;
    PFSHT_HIP_TRY(hipMemcpyAsync(f, p->f_stage, f_bytes, hipMemcpyDeviceToHost, s));
    PFSHT_HIP_TRY(hipStreamSynchronize(s));
    return kOk;
}

int pfsht_adjoint(pfsht_plan_t p, const void *f, void *alm, int32_t where, void *stream) {
    if (!p || !alm || !f) { set_error("NULL argument"); return kEArg; }
    if (where != PFSHT_WHERE_HOST && where != PFSHT_WHERE_DEVICE) { set_error("bad where"); return kEArg; }
    std::lock_guard<std::mutex> lk(p->mu);
    DeviceGuard dg(p->dev);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const size_t alm_bytes = sizeof(double2) * size_t(p->L + 1) * (2 * p->L + 1) * p->B;
    const size_t f_bytes = sizeof(double2) * size_t(p->M) * p->B;
    if (where == PFSHT_WHERE_DEVICE)
        return adjoint_impl(*p, static_cast<const double2 *>(f), static_cast<double2 *>(alm), s);
    Status st;
    if ((st = apply_begin(*p, s)) != kOk) return st;  // the staging buffers belong to the plan too
    if ((st = stage_in(reinterpret_cast<void **>(&p->f_stage), f_bytes, f, s, p->bytes)) != kOk) return st;
    if ((st = stage_in(reinterpret_cast<void **>(&p->alm_stage), alm_bytes, nullptr, s, p->bytes)) != kOk) return st;
    if ((st = adjoint_impl(*p, p->f_stage, p->alm_stage, s)) != kOk) return st;
    PFSHT_HIP_TRY(hipMemcpyAsync(alm, p->alm_stage, alm_bytes, hipMemcpyDeviceToHost, s));
    PFSHT_HIP_TRY(hipStreamSynchronize(s));
    return kOk;
}

// ComplexF32 applies (dir 0: synthesis, 1: adjoint), the device path of a Float32 plan: the input
// is widened into the plan's FP64 staging, applied by the plan's own pipeline (FP64, or FP64
// recurrence with FP32 grids for a mixed plan), and the result narrowed into the caller's buffer.
static int apply_c64(pfsht_plan_t p, int dir, const void *in, void *out, int32_t where, void *stream) {
    if (!p || !in || !out) { set_error("NULL argument"); return kEArg; }
    if (where != PFSHT_WHERE_HOST && where != PFSHT_WHERE_DEVICE) { set_error("bad where"); return kEArg; }
    std::lock_guard<std::mutex> lk(p->mu);
    DeviceGuard dg(p->dev);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const size_t na = size_t(p->L + 1) * (2 * p->L + 1) * p->B;
    const size_t nf = size_t(p->M) * p->B;
    const size_t nin = dir == 0 ? na : nf, nout = dir == 0 ? nf : na;
    Status st;
    if ((st = apply_begin(*p, s)) != kOk) return st;  // the staging buffers belong to the plan
    if ((st = stage_in(reinterpret_cast<void **>(&p->alm_stage), sizeof(double2) * na, nullptr, s, p->bytes)) != kOk)
        return st;
    if ((st = stage_in(reinterpret_cast<void **>(&p->f_stage), sizeof(double2) * nf, nullptr, s, p->bytes)) != kOk)
        return st;
    const float2 *src = static_cast<const float2 *>(in);
    float2 *dst = static_cast<float2 *>(out);
    if (where == PFSHT_WHERE_HOST) {
        if ((st = stage_in(reinterpret_cast<void **>(&p->c64_stage), sizeof(float2) * std::max(na, nf), nullptr, s,
                           p->bytes)) != kOk)
            return st;
        PFSHT_HIP_TRY(hipMemcpyAsync(p->c64_stage, in, sizeof(float2) * nin, hipMemcpyHostToDevice, s));
        src = dst = p->c64_stage;
    }
    double2 *win = dir == 0 ? p->alm_stage : p->f_stage;
    double2 *wout = dir == 0 ? p->f_stage : p->alm_stage;
    launch_widen_c64(src, win, int64_t(nin), s);
    PFSHT_HIP_TRY(hipGetLastError());
    st = dir == 0 ? synth_impl(*p, win, wout, s) : adjoint_impl(*p, win, wout, s);
    if (st != kOk) return st;
    launch_narrow_c128(wout, dst, int64_t(nout), s);
    PFSHT_HIP_TRY(hipGetLastError());
    if ((st = apply_end(*p, s)) != kOk) return st;
    if (where == PFSHT_WHERE_HOST) {
        PFSHT_HIP_TRY(hipMemcpyAsync(out, p->c64_stage, sizeof(float2) * nout, hipMemcpyDeviceToHost, s));
        PFSHT_HIP_TRY(hipStreamSynchronize(s));
    }
    return kOk;
}

int pfsht_synth_f32(pfsht_plan_t p, const void *alm, void *f, int32_t where, void *stream) {
    return apply_c64(p, 0, alm, f, where, stream);
}

int pfsht_adjoint_f32(pfsht_plan_t p, const void *f, void *alm, int32_t where, void *stream) {
    return apply_c64(p, 1, f, alm, where, stream);
}

// The column apply shared by pfsht_synth_cols (dir 0) and pfsht_adjoint_cols (dir 1).
static int cols_apply(pfsht_plan_t p, int dir, int32_t nb, const void *const *in, void *const *out, void *stream) {
    if (!p || !in || !out) { set_error("NULL argument"); return kEArg; }
    if (nb < 1 || nb > p->B) {
        set_error("nb = " + std::to_string(nb) + " outside 1:" + std::to_string(p->B) + " (the plan's ntrans)");
        return kEArg;
    }
    for (int b = 0; b < nb; ++b)
        if (!in[b] || !out[b]) { set_error("NULL column pointer"); return kEArg; }
    std::lock_guard<std::mutex> lk(p->mu);
    DeviceGuard dg(p->dev);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int slot = p->col_next;
    p->col_next = (slot + 1) % pfsht_plan_s::kColSlots;
    if (p->col_used[slot]) PFSHT_HIP_TRY(hipEventSynchronize(p->col_ev[slot]));
    const size_t B = size_t(p->B);
    void **h = p->col_host + size_t(slot) * 2 * B, **d = p->col_dev + size_t(slot) * 2 * B;
    for (int b = 0; b < nb; ++b) {
        h[b] = const_cast<void *>(in[b]);
        h[B + b] = out[b];
    }
    PFSHT_HIP_TRY(hipMemcpyAsync(d, h, 2 * B * sizeof(void *), hipMemcpyHostToDevice, s));
    InCols ic;
    ic.tab = reinterpret_cast<const double2 *const *>(d);
    OutCols oc;
    oc.tab = reinterpret_cast<double2 *const *>(d + B);
    Status st = kOk;
    if (dir == 0) st = synth_cols(*p, ic, oc, nb, s);
    else st = adjoint_cols(*p, ic, oc, nb, s);
    if (st != kOk) return st;
    PFSHT_HIP_TRY(hipEventRecord(p->col_ev[slot], s));
    p->col_used[slot] = true;
    return kOk;
}

int pfsht_synth_cols(pfsht_plan_t p, int32_t nb, const void *const *alm_cols, void *const *f_cols, void *stream) {
    return cols_apply(p, 0, nb, alm_cols, f_cols, stream);
}

int pfsht_adjoint_cols(pfsht_plan_t p, int32_t nb, const void *const *f_cols, void *const *alm_cols, void *stream) {
    return cols_apply(p, 1, nb, f_cols, alm_cols, stream);
}

// ------------------------------------------------------------------------------------------------
// Plan groups: many single-transform plans applied together. The ring passes of every member run
// as one launch over a work list of (plan, order, 64-ring tile) items (and one for the adjoint's
// partial reduction); the members' grid, FFT and point stages run concurrently, dealt round-robin
// to the group's streams and joined back into the caller's stream. Each (direction, pointer slot)
// sequence is captured once into a hipGraph and replayed; the caller's buffers reach the kernels
// through a device pointer table written before each launch, so a graph stays valid across calls.
// ------------------------------------------------------------------------------------------------
struct pfsht_group_s {
    static constexpr int kSlots = 4;    // pointer tables (and graphs) in flight
    static constexpr int kStreams = 4;  // concurrent member streams
    int dev = 0;
    std::vector<pfsht_plan_t> plans;
    int np = 0;
    int64_t bytes = 0;
    RingTab *tabs = nullptr;                  // [kSlots][np]
    int2 *fwork = nullptr, *awork[2] = {};    // forward items; adjoint workgroups for 1 / 2 tiles per wave
    int nfwork = 0, nawork[2] = {0, 0}, fwaves = 1;
    int4 *zwork = nullptr;                    // finalize work
    int nzwork = 0;
    // the stretched members' NUFFT stages, one launch per class of (kernel width, grid type)
    NufftTab *ntabs = nullptr;                // [kSlots][np]
    struct NuClass {
        int w = 0, win = 0;
        bool single = false;
        int2 *work = nullptr;   // {member, super-tile}
        int4 *mwork = nullptr;  // {member, column block, stored row, 0}
        int nwork = 0, nmwork = 0;
    };
    std::vector<NuClass> classes;
    // the grid stage of the members whose plans run it in LDS without split rows, grouped per
    // storage type (0: FP64, 1: FP32); the other stretched members run it on the member streams
    GridTab *gtabs = nullptr;     // [np]
    GridGroupWork gw[2];
    std::vector<char> grid_grouped;  // [np]
    void **ptr_dev = nullptr, **ptr_host = nullptr;  // [kSlots][2 np]: inputs, then outputs
    EventRef slot_ev[kSlots];  // each slot's last apply; also its members' `last`
    bool slot_used[kSlots] = {};
    int slot_next = 0;
    hipStream_t streams[kStreams] = {};
    hipStream_t main = nullptr;  // the replays are captured here (whatever the caller's stream)
    hipEvent_t fork = nullptr, join[kStreams] = {};
    hipGraphExec_t graph[2][kSlots] = {};
    std::mutex mu;

    ~pfsht_group_s() {
        DeviceGuard dg(dev);
        (void)hipDeviceSynchronize();
        for (auto &gd : graph)
            for (auto g : gd)
                if (g) (void)hipGraphExecDestroy(g);
        for (auto e : join)
            if (e) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        for (auto q : streams)
            if (q) (void)hipStreamDestroy(q);
        if (main) (void)hipStreamDestroy(main);
        dfree(tabs); dfree(fwork); dfree(awork[0]); dfree(awork[1]); dfree(zwork); dfree(ptr_dev); dfree(ntabs);
        dfree(gtabs);
        for (auto &w : gw) {
            dfree(w.col); dfree(w.row_s); dfree(w.row_a); dfree(w.col_a);
        }
        for (auto &c : classes) {
            dfree(c.work);
            dfree(c.mwork);
        }
        if (ptr_host) (void)hipHostFree(ptr_host);
    }
};

static Status group_build(pfsht_group_s &g, const pfsht_plan_t *plans, int np) {
    if (np < 1) { set_error("a group needs at least one plan"); return kEArg; }
    g.np = np;
    g.dev = plans[0] ? plans[0]->dev : 0;
    for (int i = 0; i < np; ++i) {
        pfsht_plan_s *p = plans[i];
        if (!p) { set_error("NULL plan in the group"); return kEArg; }
        for (int k = 0; k < i; ++k)
            if (plans[k] == p) { set_error("a plan appears twice in the group"); return kEArg; }
        if (p->dev != g.dev) { set_error("the plans of a group must live on one device"); return kEArg; }
        if (p->B != 1) { set_error("group members are single-transform plans (ntrans = 1)"); return kEUnsupported; }
        if (p->spin || (p->method != PFSHT_METHOD_STRETCH && p->method != PFSHT_METHOD_DIRECT) || !p->rings) {
            set_error("group members are scalar stretched or Direct plans (no spin, no multi-box)");
            return kEUnsupported;
        }
        g.plans.push_back(p);
    }
    DeviceGuard dg(g.dev);
    // work lists: each engine's item lists are whole rounds of eight (one per XCD), so concatenating
    // them keeps item i on XCD i % 8
    std::vector<int2> fw, aw[2];
    std::vector<int4> zw;
    for (int i = 0; i < np; ++i) {
        const RingDev &d = g.plans[i]->rings->d;
        for (int k = 0; k < d.nfitems; ++k) fw.push_back(make_int2(i, k));
        const int tgi = d.adj_tg == 2 ? 1 : 0;
        for (int k = 0; k < d.naitems / kAdjWaves; ++k) aw[tgi].push_back(make_int2(i, k));
        for (int am = 0; am <= d.L; ++am)
            for (int rb = 0; rb < (d.L + 1 + 255) / 256; ++rb) zw.push_back(make_int4(i, rb, am, 0));
        g.fwaves = std::max(g.fwaves, d.fwd_waves);
    }
    Status st;
    if ((st = upload(&g.fwork, fw, g.bytes)) != kOk) return st;
    if ((st = upload(&g.awork[0], aw[0], g.bytes)) != kOk) return st;
    if ((st = upload(&g.awork[1], aw[1], g.bytes)) != kOk) return st;
    if ((st = upload(&g.zwork, zw, g.bytes)) != kOk) return st;
    g.nfwork = int(fw.size());
    g.nawork[0] = int(aw[0].size());
    g.nawork[1] = int(aw[1].size());
    g.nzwork = int(zw.size());
    const size_t nptr = size_t(pfsht_group_s::kSlots) * 2 * np;
    if ((st = dalloc(&g.ptr_dev, nptr, g.bytes)) != kOk) return st;
    PFSHT_HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&g.ptr_host), nptr * sizeof(void *)));
    // engine tables per slot: the coefficient columns are entries of that slot's pointer table
    std::vector<RingTab> tabs(size_t(pfsht_group_s::kSlots) * np);
    for (int k = 0; k < pfsht_group_s::kSlots; ++k)
        for (int i = 0; i < np; ++i) {
            const RingDev &d = g.plans[i]->rings->d;
            RingTab &t = tabs[size_t(k) * np + i];
            t.fitems = d.fitems; t.aitems = d.aitems; t.costh = d.costh; t.log2sin = d.log2sin; t.log2c = d.log2c;
            t.ab = d.ab; t.gc = d.gc; t.abofs = d.abofs; t.ckq = d.ckq; t.ckE = d.ckE; t.pofs = d.pofs;
            t.ntiles = d.ntiles; t.partials = d.partials; t.Gd = g.plans[i]->rings->Gd;
            void **tbl = g.ptr_dev + size_t(k) * 2 * np;
            t.alm_in = InCols();
            t.alm_in.tab = reinterpret_cast<const double2 *const *>(tbl + i);
            t.alm_out = OutCols();
            t.alm_out.tab = reinterpret_cast<double2 *const *>(tbl + np + i);
            t.L = d.L; t.n = d.n; t.mmax = d.mmax; t.part_b = d.part_b;
        }
    if ((st = upload(&g.tabs, tabs, g.bytes)) != kOk) return st;
    // NUFFT classes and tables of the stretched members
    std::vector<NufftTab> ntabs(size_t(pfsht_group_s::kSlots) * np);
    std::vector<std::vector<int2>> cw;
    std::vector<std::vector<int4>> cm;
    for (int i = 0; i < np; ++i) {
        pfsht_plan_s &p = *g.plans[i];
        if (p.method != PFSHT_METHOD_STRETCH) continue;
        const NufftDev &u = p.nu;
        const bool single = p.precision == PFSHT_PREC_MIXED;
        size_t c = 0;
        while (c < g.classes.size() && !(g.classes[c].w == u.w && g.classes[c].single == single)) ++c;
        if (c == g.classes.size()) {
            pfsht_group_s::NuClass nc;
            nc.w = u.w;
            nc.win = u.win;
            nc.single = single;
            g.classes.push_back(nc);
            cw.emplace_back();
            cm.emplace_back();
        }
        for (int q = 0; q < u.nst; ++q) cw[c].push_back(make_int2(i, q));
        for (int r = 0; r < u.nR; ++r)
            for (int cb = 0; cb < (u.N2 + 255) / 256; ++cb) cm[c].push_back(make_int4(i, cb, r, 0));
        for (int k = 0; k < pfsht_group_s::kSlots; ++k) {
            NufftTab &t = ntabs[size_t(k) * np + i];
            t.X1 = u.X1; t.X2 = u.X2; t.perm = u.perm; t.st_list = u.st_list; t.st_slot = u.st_slot;
            t.bin_ofs = u.bin_ofs; t.U = p.U; t.Wst = u.Wst;
            void **tbl = g.ptr_dev + size_t(k) * 2 * np;
            t.fin = InCols();
            t.fin.tab = reinterpret_cast<const double2 *const *>(tbl + i);
            t.fout = OutCols();
            t.fout.tab = reinterpret_cast<double2 *const *>(tbl + np + i);
            t.N1 = u.N1; t.N2 = u.N2; t.S = u.S; t.ns1 = u.ns1; t.ns2 = u.ns2; t.win = u.win; t.w = u.w;
            t.nst = u.nst; t.r0 = u.r0; t.nR = u.nR;
        }
    }
    for (size_t c = 0; c < g.classes.size(); ++c) {
        if ((st = upload(&g.classes[c].work, cw[c], g.bytes)) != kOk) return st;
        if ((st = upload(&g.classes[c].mwork, cm[c], g.bytes)) != kOk) return st;
        g.classes[c].nwork = int(cw[c].size());
        g.classes[c].nmwork = int(cm[c].size());
    }
    if ((st = upload(&g.ntabs, ntabs, g.bytes)) != kOk) return st;
    // grouped grid stage: work lists per storage type; the XCD-numbered lists (row synthesis, column
    // adjoint) give each member whole rounds of eight workgroups
    {
        std::vector<GridTab> gt(np);
        std::vector<int2> wl[2][4];
        g.grid_grouped.assign(np, 0);
        for (int i = 0; i < np; ++i) {
            pfsht_plan_s &p = *g.plans[i];
            if (p.method != PFSHT_METHOD_STRETCH || !p.fused || p.gf.split) continue;
            g.grid_grouped[i] = 1;
            const int cls = p.precision == PFSHT_PREC_MIXED ? 1 : 0;
            const size_t csz = cls ? sizeof(float2) : sizeof(double2);
            GridTab &t = gt[i];
            t.g = p.gf;
            t.Gd = p.rings->Gd; t.jlo = p.rings->d.jlo; t.jhi = p.rings->d.jhi; t.wth = p.wth;
            t.pc1 = p.nu.pc1; t.c2 = p.nu.c2; t.Ut = p.Ut; t.U = p.U;
            const int ncol = 2 * p.mmax + 1, nR = p.nu.nR;
            for (int q = 0; q < ncol; ++q) wl[cls][0].push_back(make_int2(i, q));
            for (int w = 0; w < 8 * ((nR + 7) / 8); ++w) wl[cls][1].push_back(make_int2(i, w));
            for (int r = 0; r < nR; ++r) wl[cls][2].push_back(make_int2(i, r));
            for (int w = 0; w < 8 * ((ncol + 7) / 8); ++w) wl[cls][3].push_back(make_int2(i, w));
            GridGroupWork &gw = g.gw[cls];
            gw.nt_col = std::max(gw.nt_col, p.gf.nt_col);
            gw.nt_row = std::max(gw.nt_row, p.gf.nt_row);
            gw.lds_col = std::max(gw.lds_col, csz * size_t(lds_slots(p.nu.N1)));
            gw.lds_row = std::max(gw.lds_row, csz * size_t(lds_slots(p.nu.N2)));
        }
        if ((st = upload(&g.gtabs, gt, g.bytes)) != kOk) return st;
        for (int c = 0; c < 2; ++c) {
            GridGroupWork &gw = g.gw[c];
            if ((st = upload(&gw.col, wl[c][0], g.bytes)) != kOk) return st;
            if ((st = upload(&gw.row_s, wl[c][1], g.bytes)) != kOk) return st;
            if ((st = upload(&gw.row_a, wl[c][2], g.bytes)) != kOk) return st;
            if ((st = upload(&gw.col_a, wl[c][3], g.bytes)) != kOk) return st;
            gw.ncol = int(wl[c][0].size()); gw.nrow_s = int(wl[c][1].size());
            gw.nrow_a = int(wl[c][2].size()); gw.ncol_a = int(wl[c][3].size());
        }
    }
    for (auto &e : g.slot_ev)
        if ((st = event_make(e)) != kOk) return st;
    for (auto &q : g.streams) PFSHT_HIP_TRY(hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
    PFSHT_HIP_TRY(hipStreamCreateWithFlags(&g.main, hipStreamNonBlocking));
    PFSHT_HIP_TRY(hipEventCreateWithFlags(&g.fork, hipEventDisableTiming));
    for (auto &e : g.join) PFSHT_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    PFSHT_HIP_TRY(hipDeviceSynchronize());
    return kOk;
}

// The work of one group apply (dir 0 synthesis, 1 adjoint) with pointer slot k, on stream s.
static Status group_enqueue(pfsht_group_s &g, int dir, int k, hipStream_t s) {
    const int np = g.np;
    const RingTab *tabs = g.tabs + size_t(k) * np;
    void **tbl = g.ptr_dev + size_t(k) * 2 * np;
    auto member = [&](int i, hipStream_t q) -> Status {
        pfsht_plan_s &p = *g.plans[i];
        InCols in;
        in.tab = reinterpret_cast<const double2 *const *>(tbl + i);
        OutCols out;
        out.tab = reinterpret_cast<double2 *const *>(tbl + np + i);
        pfsht_rings_s &r = *p.rings;
        if (p.method == PFSHT_METHOD_DIRECT) {
            if (dir == 0) launch_direct_synth(r.Gd, p.dphi, p.dmlim, p.dperm, out, p.M, p.mmax, 1, q);
            else launch_direct_adjoint(in, p.dphi, p.dmlim, p.dperm, r.Gd, p.M, p.mmax, 1, q);
            return kOk;
        }
        if (g.grid_grouped[i]) return kOk;  // the grouped grid-stage launches cover it
        // the interpolation and spreading run grouped, after and before the members' stages
        const bool mixed = p.precision == PFSHT_PREC_MIXED;
        if (dir == 0)
            return mixed ? synth_grid<float2>(p, nullptr, out, 1, q, false)
                         : synth_grid<double2>(p, nullptr, out, 1, q, false);
        return mixed ? adjoint_grid<float2>(p, nullptr, in, 1, q, false)
                     : adjoint_grid<double2>(p, nullptr, in, 1, q, false);
    };
    const NufftTab *ntabs = g.ntabs + size_t(k) * np;
    auto nufft = [&]() {
        for (const auto &c : g.classes) {
            if (dir == 0) {
                if (c.single) launch_interp_group<float2>(ntabs, c.work, c.nwork, c.w, c.win, s);
                else launch_interp_group<double2>(ntabs, c.work, c.nwork, c.w, c.win, s);
            } else {
                if (c.single) launch_spread_group<float2>(ntabs, c.work, c.nwork, c.mwork, c.nmwork, c.w, c.win, s);
                else launch_spread_group<double2>(ntabs, c.work, c.nwork, c.mwork, c.nmwork, c.w, c.win, s);
            }
        }
    };
    // the members' own stages on the member streams (Direct members, and the grid stage of members
    // outside the grouped launches), the grouped grid-stage launches on s meanwhile
    auto members = [&]() -> Status {
        std::vector<int> own;
        for (int i = 0; i < np; ++i)
            if (!g.grid_grouped[i]) own.push_back(i);
        const int ns = std::min(int(own.size()), pfsht_group_s::kStreams);
        if (ns > 0) {
            PFSHT_HIP_TRY(hipEventRecord(g.fork, s));
            for (int j = 0; j < ns; ++j) PFSHT_HIP_TRY(hipStreamWaitEvent(g.streams[j], g.fork, 0));
            for (size_t k2 = 0; k2 < own.size(); ++k2) {
                Status st = member(own[k2], g.streams[k2 % ns]);
                if (st != kOk) return st;
            }
        }
        if (dir == 0) {
            launch_grid_synth_group<double2>(g.gtabs, g.gw[0], s);
            launch_grid_synth_group<float2>(g.gtabs, g.gw[1], s);
        } else {
            launch_grid_adjoint_group<double2>(g.gtabs, g.gw[0], s);
            launch_grid_adjoint_group<float2>(g.gtabs, g.gw[1], s);
        }
        for (int j = 0; j < ns; ++j) {
            PFSHT_HIP_TRY(hipEventRecord(g.join[j], g.streams[j]));
            PFSHT_HIP_TRY(hipStreamWaitEvent(s, g.join[j], 0));
        }
        return kOk;
    };
    Status st;
    if (dir == 0) {
        launch_ring_fwd_group(tabs, g.fwork, g.nfwork, g.fwaves, s);
        if ((st = members()) != kOk) return st;
        nufft();
    } else {
        nufft();
        if ((st = members()) != kOk) return st;
        launch_ring_adj_group(tabs, g.awork[1], g.nawork[1], 2, s);
        launch_ring_adj_group(tabs, g.awork[0], g.nawork[0], 1, s);
        launch_ring_finalize_group(tabs, g.zwork, g.nzwork, s);
    }
    PFSHT_HIP_TRY(hipGetLastError());
    return kOk;
}

static int group_apply(pfsht_group_t g, int dir, const void *const *in, void *const *out, void *stream) {
    if (!g || !in || !out) { set_error("NULL argument"); return kEArg; }
    const int np = g->np;
    for (int i = 0; i < np; ++i)
        if (!in[i] || !out[i]) { set_error("NULL buffer pointer"); return kEArg; }
    std::lock_guard<std::mutex> lk(g->mu);
    // the members' own locks, in address order
    std::vector<pfsht_plan_s *> order(g->plans.begin(), g->plans.end());
    std::sort(order.begin(), order.end());
    std::vector<std::unique_lock<std::mutex>> locks;
    locks.reserve(order.size());
    for (auto *p : order) locks.emplace_back(p->mu);
    DeviceGuard dg(g->dev);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int k = g->slot_next;
    g->slot_next = (k + 1) % pfsht_group_s::kSlots;
    if (g->slot_used[k]) PFSHT_HIP_TRY(hipEventSynchronize(g->slot_ev[k]->e));
    void **h = g->ptr_host + size_t(k) * 2 * np;
    for (int i = 0; i < np; ++i) {
        h[i] = const_cast<void *>(in[i]);
        h[np + i] = out[i];
    }
    PFSHT_HIP_TRY(hipMemcpyAsync(g->ptr_dev + size_t(k) * 2 * np, h, 2 * size_t(np) * sizeof(void *),
                                 hipMemcpyHostToDevice, s));
    // order after each member's previous apply on another stream (members last applied together
    // share one event)
    std::vector<const SharedEvent *> waited;
    for (auto *p : g->plans) {
        if (!p->last || s == p->last_stream) continue;
        if (std::find(waited.begin(), waited.end(), p->last.get()) != waited.end()) continue;
        waited.push_back(p->last.get());
        PFSHT_HIP_TRY(hipStreamWaitEvent(s, p->last->e, 0));
    }
    Status st;
    hipGraphExec_t &ge = g->graph[dir][k];
    if (!ge) {
        // capture the sequence once, on the group's own stream (capture runs nothing); where capture
        // fails the sequence runs eagerly
        hipGraph_t graph = nullptr;
        if (hipStreamBeginCapture(g->main, hipStreamCaptureModeThreadLocal) == hipSuccess) {
            st = group_enqueue(*g, dir, k, g->main);
            const hipError_t e = hipStreamEndCapture(g->main, &graph);
            if (st == kOk && e == hipSuccess && graph &&
                hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0) != hipSuccess)
                ge = nullptr;
            if (graph) (void)hipGraphDestroy(graph);
            if (st != kOk) return st;
        }
        (void)hipGetLastError();  // a refused capture leaves its error behind
    }
    // the replay (or the eager sequence) runs on the caller's stream
    if (ge) {
        PFSHT_HIP_TRY(hipGraphLaunch(ge, s));
    } else if ((st = group_enqueue(*g, dir, k, s)) != kOk) {
        return st;
    }
    // one event marks the end for the group and for every member
    PFSHT_HIP_TRY(hipEventRecord(g->slot_ev[k]->e, s));
    for (auto *p : g->plans) {
        p->last = g->slot_ev[k];
        p->last_stream = s;
    }
    g->slot_used[k] = true;
    return kOk;
}

int pfsht_group_create(const pfsht_plan_t *plans, int32_t nplans, pfsht_group_t *out) {
    if (!plans || !out) { set_error("NULL argument"); return kEArg; }
    *out = nullptr;
    auto *g = new (std::nothrow) pfsht_group_s;
    if (!g) { set_error("host allocation failed"); return kENomem; }
    Status st;
    try {
        st = group_build(*g, plans, nplans);
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        st = kENomem;
    }
    if (st != kOk) {
        delete g;
        return st;
    }
    *out = g;
    return kOk;
}

void pfsht_group_destroy(pfsht_group_t g) { delete g; }

int pfsht_group_synth(pfsht_group_t g, const void *const *alm, void *const *f, void *stream) {
    return group_apply(g, 0, alm, f, stream);
}

int pfsht_group_adjoint(pfsht_group_t g, const void *const *f, void *const *alm, void *stream) {
    return group_apply(g, 1, f, alm, stream);
}

int pfsht_plan_nufft(pfsht_plan_t p, int32_t type, const void *in, void *out, int32_t where, void *stream) {
    if (!p || !in || !out) { set_error("NULL argument"); return kEArg; }
    if (type != 1 && type != 2) { set_error("NUFFT type must be 1 or 2"); return kEArg; }
    if (p->method != PFSHT_METHOD_STRETCH) { set_error("only a stretched plan has a NUFFT"); return kEArg; }
    if (p->precision != PFSHT_PREC_FP64) { set_error("the NUFFT seam is FP64 only"); return kEUnsupported; }
    if (where != PFSHT_WHERE_HOST && where != PFSHT_WHERE_DEVICE) { set_error("bad where"); return kEArg; }
    std::lock_guard<std::mutex> lk(p->mu);
    DeviceGuard dg(p->dev);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    Status st0 = seam_ready(*p);
    if (st0 != kOk) return st0;
    NufftDev u = p->nu;  // the seam works on the whole fine grid
    u.r0 = 0;
    u.nR = u.N1;
    double2 *Ft = static_cast<double2 *>(p->Ft), *U = static_cast<double2 *>(p->U);
    const size_t modes_bytes = sizeof(double2) * size_t(p->nth) * p->nph * p->B;
    const size_t f_bytes = sizeof(double2) * size_t(p->M) * p->B;
    // the modes buffer in the reference layout lives in Fp's place when the caller's memory is on
    // the host; the plan's Ft holds the internal [i][k] layout.
    double2 *modes_ref = nullptr, *fbuf = nullptr;
    Status st;
    if ((st = apply_begin(*p, s)) != kOk) return st;
    const void *in_d = in;
    void *out_d = out;
    if (where == PFSHT_WHERE_HOST) {
        if ((st = stage_in(reinterpret_cast<void **>(&p->f_stage), f_bytes, type == 1 ? in : nullptr, s, p->bytes)) != kOk)
            return st;
        if ((st = stage_in(reinterpret_cast<void **>(&p->modes_stage), modes_bytes, type == 2 ? in : nullptr, s,
                           p->bytes)) != kOk)
            return st;
        fbuf = p->f_stage;
        modes_ref = p->modes_stage;
        in_d = type == 1 ? static_cast<const void *>(fbuf) : static_cast<const void *>(modes_ref);
        out_d = type == 1 ? static_cast<void *>(modes_ref) : static_cast<void *>(fbuf);
    }
    if (type == 2) {
        // reference column-major nth x nph == row-major [k][i]; internal is [i][k]
        launch_transpose(static_cast<const double2 *>(in_d), Ft, p->nph, p->nth, p->B, s);
        launch_deconv_pad(u, u.c1, Ft, U, s);
        if ((st = fft_exec(p->fine_bwd, U, s)) != kOk) return st;
        launch_interp(u, static_cast<const double2 *>(U), out_cols(static_cast<double2 *>(out_d), p->M), s);
    } else {
        launch_spread(u, in_cols(static_cast<const double2 *>(in_d), p->M), U, s);
        if ((st = fft_exec(p->fine_fwd, U, s)) != kOk) return st;
        launch_deconv_trunc(u, u.c1, U, Ft, s);
        launch_transpose(Ft, static_cast<double2 *>(out_d), p->nth, p->nph, p->B, s);
    }
    PFSHT_HIP_TRY(hipGetLastError());
    if ((st = apply_end(*p, s)) != kOk) return st;
    if (where == PFSHT_WHERE_HOST) {
        PFSHT_HIP_TRY(hipMemcpyAsync(out, out_d, type == 1 ? modes_bytes : f_bytes, hipMemcpyDeviceToHost, s));
        PFSHT_HIP_TRY(hipStreamSynchronize(s));
    }
    return kOk;
}

int pfsht_rings_create(const double *theta, int32_t nrings, int32_t lmax, double eps, int32_t ntrans, int32_t reflect,
                       pfsht_rings_t *out) {
    if (!out || !theta) { set_error("NULL argument"); return kEArg; }
    *out = nullptr;
    auto *r = new (std::nothrow) pfsht_rings_s;
    if (!r) { set_error("host allocation failed"); return kENomem; }
    Status st = rings_build(theta, nrings, lmax, eps, ntrans, reflect != 0, *r);
    if (st == kOk) st = hipDeviceSynchronize() == hipSuccess ? kOk : kEHip;
    if (st != kOk) { delete r; return st; }
    *out = r;
    return kOk;
}

int pfsht_spin_rings_create(const double *theta, int32_t nrings, int32_t lmax, int32_t spin, double eps,
                            int32_t ntrans, int32_t reflect, pfsht_rings_t *out) {
    if (!out || !theta) { set_error("NULL argument"); return kEArg; }
    *out = nullptr;
    auto *r = new (std::nothrow) pfsht_rings_s;
    if (!r) { set_error("host allocation failed"); return kENomem; }
    Status st;
    try {
        st = rings_build_spin(theta, nrings, lmax, spin, eps, ntrans, reflect != 0, *r);
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        st = kENomem;
    }
    if (st == kOk) st = hipDeviceSynchronize() == hipSuccess ? kOk : kEHip;
    if (st != kOk) { delete r; return st; }
    *out = r;
    return kOk;
}

void pfsht_rings_destroy(pfsht_rings_t r) {
    if (!r) return;
    DeviceGuard dg(r->dev);
    (void)hipDeviceSynchronize();
    delete r;
}

int64_t pfsht_rings_work(pfsht_rings_t r) { return r ? r->work : -1; }
int32_t pfsht_rings_mmax(pfsht_rings_t r) { return r ? r->d.mmax : -1; }

int pfsht_plan_rings(pfsht_plan_t p, pfsht_rings_t *out) {
    if (!p || !out) { set_error("NULL argument"); return kEArg; }
    if (p->method == PFSHT_METHOD_MULTIBOX) { set_error("a multi-box plan has one ring engine per box"); return kEArg; }
    *out = p->rings;
    return kOk;
}

int pfsht_rings_from_alm(pfsht_rings_t r, const void *alm, void *G, int32_t where, void *stream) {
    if (!r || !alm || !G) { set_error("NULL argument"); return kEArg; }
    std::lock_guard<std::mutex> lk(r->mu);
    DeviceGuard dg(r->dev);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const size_t alm_bytes = sizeof(double2) * size_t(r->d.L + 1) * (2 * r->d.L + 1) * r->d.B;
    const size_t G_bytes = sizeof(double2) * size_t(r->d.n) * (2 * r->d.L + 1) * r->d.B;
    const double2 *a = static_cast<const double2 *>(alm);
    double2 *Gout = static_cast<double2 *>(G);
    Status st;
    if (where == PFSHT_WHERE_HOST) {
        if ((st = stage_in(reinterpret_cast<void **>(&r->alm_stage), alm_bytes, alm, s, r->bytes)) != kOk) return st;
        if ((st = stage_in(reinterpret_cast<void **>(&r->G_stage), G_bytes, nullptr, s, r->bytes)) != kOk) return st;
        a = r->alm_stage;
        Gout = r->G_stage;
    }
    ring_forward(*r, in_cols(a, int64_t(r->d.L + 1) * (2 * r->d.L + 1)), r->d.B, s);
    launch_gd_to_ref(r->d, r->Gd, Gout, s);
    PFSHT_HIP_TRY(hipGetLastError());
    if (where == PFSHT_WHERE_HOST) {
        PFSHT_HIP_TRY(hipMemcpyAsync(G, r->G_stage, G_bytes, hipMemcpyDeviceToHost, s));
        PFSHT_HIP_TRY(hipStreamSynchronize(s));
    }
    return kOk;
}

int pfsht_alm_from_rings(pfsht_rings_t r, const void *G, void *alm, int32_t where, void *stream) {
    if (!r || !alm || !G) { set_error("NULL argument"); return kEArg; }
    std::lock_guard<std::mutex> lk(r->mu);
    DeviceGuard dg(r->dev);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const size_t alm_bytes = sizeof(double2) * size_t(r->d.L + 1) * (2 * r->d.L + 1) * r->d.B;
    const size_t G_bytes = sizeof(double2) * size_t(r->d.n) * (2 * r->d.L + 1) * r->d.B;
    const double2 *Gin = static_cast<const double2 *>(G);
    double2 *aout = static_cast<double2 *>(alm);
    Status st;
    if (where == PFSHT_WHERE_HOST) {
        if ((st = stage_in(reinterpret_cast<void **>(&r->G_stage), G_bytes, G, s, r->bytes)) != kOk) return st;
        if ((st = stage_in(reinterpret_cast<void **>(&r->alm_stage), alm_bytes, nullptr, s, r->bytes)) != kOk) return st;
        Gin = r->G_stage;
        aout = r->alm_stage;
    }
    launch_ref_to_gd(r->d, Gin, r->Gd, s);
    ring_adjoint(*r, out_cols(aout, int64_t(r->d.L + 1) * (2 * r->d.L + 1)), r->d.B, s);
    PFSHT_HIP_TRY(hipGetLastError());
    if (where == PFSHT_WHERE_HOST) {
        PFSHT_HIP_TRY(hipMemcpyAsync(alm, r->alm_stage, alm_bytes, hipMemcpyDeviceToHost, s));
        PFSHT_HIP_TRY(hipStreamSynchronize(s));
    }
    return kOk;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------
// The allocation-free adjoint ring pass. Its scratch, the internal ring buffer and the ring pass's
// partial sums, lives in one file-scope device array; each GPU the library runs on holds its own copy
// until the process ends. Host code never passes the array's name to a kernel: the device address is
// looked up once per device (hipGetSymbolAddress, with that device current) and used from then on.
// Both buffers are overwritten completely by every call, so nothing has to be cleared first; calls
// that overlap on one GPU share the array and race.
// ------------------------------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[kScratchDevices] = {};  // per device, nullptr until looked up

static Status scratch_address(int dev, unsigned char **out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (dev < 0 || dev >= kScratchDevices) { set_error("device index outside 0..63"); return kEHip; }
    if (!g_scratch_ptr[dev]) {
        DeviceGuard dg(dev);  // the device whose copy of the array is looked up
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            set_error(std::string("hipGetSymbolAddress of the scratch array failed: ") + hipGetErrorString(e));
            return kEHip;
        }
        g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_ptr[dev];
    return kOk;
}

extern "C" {

int pfsht_alm_from_rings_noalloc(pfsht_rings_t r, const void *G, void *alm, void *stream) {
    if (!r || !alm || !G) { set_error("NULL argument"); return kEArg; }
    std::lock_guard<std::mutex> lk(r->mu);
    DeviceGuard dg(r->dev);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const size_t gd_bytes = sizeof(double2) * size_t(r->d.B) * (2 * r->d.mmax + 1) * r->d.n;
    const size_t gd_pad = (gd_bytes + 255) & ~size_t(255);
    const size_t need = gd_pad + size_t(r->part_bytes);
    if (need > kScratchBytes) {
        set_error("the ring buffer and partial sums of this engine need " + std::to_string(need) +
                  " bytes, more than the allocation-free scratch holds (" + std::to_string(kScratchBytes) +
                  "); use pfsht_alm_from_rings");
        return kEUnsupported;
    }
    unsigned char *base = nullptr;
    Status st = scratch_address(r->dev, &base);
    if (st != kOk) return st;
    double2 *Gd = reinterpret_cast<double2 *>(base);
    RingDev d = r->d;  // the engine's tables, with its scratch redirected to the array
    d.partials = reinterpret_cast<double4 *>(base + gd_pad);
    const OutCols out = out_cols(static_cast<double2 *>(alm), int64_t(d.L + 1) * (2 * d.L + 1));
    launch_ref_to_gd(d, static_cast<const double2 *>(G), Gd, s);
    if (r->spin) {
        SpinDev sd = r->sd;
        sd.partials = reinterpret_cast<double2 *>(base + gd_pad);
        launch_spin_adj(d, sd, Gd, out, d.B, s);
    } else {
        launch_ring_adj(d, Gd, out, d.B, s);
    }
    PFSHT_HIP_TRY(hipGetLastError());
    return kOk;
}

int pfsht_set_timing(pfsht_plan_t p, int32_t enable) {
    if (!p) { set_error("NULL plan"); return kEArg; }
    std::lock_guard<std::mutex> lk(p->mu);
    p->timing = enable != 0;
    return kOk;
}

int pfsht_stage_times(pfsht_plan_t p, int32_t dir, double *ms, int64_t *counts, int32_t reset) {
    if (!p || dir < 0 || dir > 1) { set_error("bad argument"); return kEArg; }
    std::lock_guard<std::mutex> lk(p->mu);
    DeviceGuard dg(p->dev);
    timing_drain(*p);
    for (int k = 0; k < kStages; ++k) {
        if (ms) ms[k] = p->ms[dir][k];
        if (counts) counts[k] = p->counts[dir][k];
    }
    if (reset) {
        std::memset(p->ms[dir], 0, sizeof p->ms[dir]);
        std::memset(p->counts[dir], 0, sizeof p->counts[dir]);
    }
    return kOk;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------
// Hooks of the serving runtime (serve.hip): a plan's size before it is built, and its column apply
// split into the cross-stream ordering and the launches (so that the launches can be captured in a
// graph while the ordering events stay outside it).
// ------------------------------------------------------------------------------------------------
namespace pfsht {

// Device bytes of a stretched leaf over colatitudes [thlo, thhi] with npts points: the ring
// engine's tables, checkpoints, partials and ring buffer, the grids, the points and the super-tile
// windows (bounded by the super-tile count), from the host sizing of plan_build; -1 when the stretch
// is refused. An upper estimate: the windows of super-tiles no point reaches are counted.
static int64_t estimate_stretch(double thlo, double thhi, int64_t M, int L, double tol, int B,
                                const pfsht_plan_opts &o, bool single) {
    std::string msg;
    Geom g0, g;
    if (stretch_geometry(thlo, thhi, L, tol, g0, msg, o.epsapo_max) != 0) return -1;
    const double tiny = std::sqrt(DBL_EPSILON);
    const bool flip = g0.kind == kPole && g0.thhi >= kPi - tiny && g0.thlo > tiny;
    if (flip) {
        if (stretch_geometry(kPi - thhi, kPi - thlo, L, tol, g, msg, o.epsapo_max) != 0) return -1;
    } else {
        g = g0;
    }
    const int Kt = int(std::ceil(double(g.K + kKMargin) * o.oversample_theta));
    int nh, nth;
    if (g.kind == kInterior) { nth = int(next_smooth(2 * Kt + 1)); nh = nth; }
    else { nh = int(next_smooth(Kt + 1)); nth = 2 * nh; }
    std::vector<double> th, Th;
    stretch_rings(g, nh, th, Th);
    const int as = std::abs(o.spin);
    std::vector<int> mlim(nh);
    int mmax = 0;
    for (int j = 0; j < nh; ++j) {
        mlim[j] = std::max(polar_mmax(th[j], L, tol), as);
        mmax = std::max(mmax, mlim[j]);
    }
    std::vector<int> jlo, jhi;
    if (!nested_ranges(mlim, L, jlo, jhi, msg)) return -1;
    int64_t by = 0;
    for (int m = 0; m <= mmax; ++m) {
        by += int64_t(std::max(0, L - m - 1)) * 32 * (o.spin ? 2 : 1);  // coefficient tables
        const int nj = jhi[m] - jlo[m] + 1;
        if (nj <= 0) continue;
        const int nt = (nj + 63) / 64, nseg = seg_count(m, L, kSegBlocks);
        by += int64_t(nt) * (nseg - 1) * 64 * 20;                                   // checkpoints
        by += int64_t((nt + 1) / 2) * adj_rows_padded(m, L) * 32 * (B >= 2 ? 1 : B); // partials
    }
    const int nph = int(next_smooth_even(int64_t(std::ceil(double(2 * int64_t(std::max(mmax, as)) + 1) * o.oversample_phi))));
    int w;
    double beta;
    es_params(tol, w, beta);
    const int N1 = es_fine_size(nth, w), N2 = es_fine_size(nph, w);
    const int S4 = 4 * (17 - w), win = 3 * (17 - w) + 16;
    const int64_t nst = std::min<int64_t>(int64_t((N1 + S4 - 1) / S4) * ((N2 + S4 - 1) / S4), M);
    const int64_t csz = single ? 8 : 16, ncol = 2 * mmax + 1;
    int dev = 0;
    (void)hipGetDevice(&dev);
    const bool fused = o.grid_fft != PFSHT_GRID_FFT_LIBRARY && grid_fits(nth, N1, N2, size_t(csz), dev, nullptr);
    by += int64_t(B) * ncol * nh * 16;                                  // ring buffer
    by += int64_t(B) * csz * ((fused ? 0 : int64_t(nth) * nph) + int64_t(N1) * N2 + ncol * N1 + nst * win * win);
    by += fused ? csz * (int64_t(nth) + N1 + N2 + N2 / 2) : 0;         // twiddles of the LDS grid stage
    by += M * 20 + int64_t(nh) * 24 + nst * 80;                        // points, permutation, tables
    return by;
}

int64_t plan_estimate_bytes(const double *theta, const double *phi, int64_t M, int L, double tol, int B,
                            const pfsht_plan_opts &o) {
    if (M < 1 || L < 0 || B < 1) return -1;
    double thlo = INFINITY, thhi = -INFINITY;
    for (int64_t i = 0; i < M; ++i) { thlo = std::min(thlo, theta[i]); thhi = std::max(thhi, theta[i]); }
    thlo = std::max(thlo, 0.0);
    thhi = std::min(thhi, kPi);
    const bool single = o.precision == PFSHT_PREC_MIXED;
    const int64_t direct = int64_t(B) * (2 * L + 1) * M * 16 + ((M + 63) / 64) * int64_t(L + 1) * (L + 2) / 2 * 32 *
                           (B >= 2 ? 1 : B) + M * 32;
    int method = o.method;
    if (method == PFSHT_METHOD_AUTO) {
        const double cd = cost_direct(M, L, B), cs = cost_stretch(thlo, thhi, M, L, B, tol);
        method = cs < cd ? PFSHT_METHOD_STRETCH : PFSHT_METHOD_DIRECT;
        const Partition part = partition_boxes(theta, phi, M, L, B, tol, o.maxboxes);
        if (part.groups.size() > 1) {
            double cm = 0;
            for (double c : part.costs) cm += c;
            if (cm < std::min(cd, cs)) {
                int64_t by = int64_t(L + 1) * (2 * L + 1) * B * 16;
                for (const auto &grp : part.groups) {
                    double lo = INFINITY, hi = -INFINITY;
                    for (int32_t i : grp) { lo = std::min(lo, theta[i]); hi = std::max(hi, theta[i]); }
                    const int64_t b = estimate_stretch(std::max(lo, 0.0), std::min(hi, kPi), int64_t(grp.size()), L,
                                                       tol, B, o, single);
                    if (b < 0) return direct;
                    by += b + int64_t(grp.size()) * 4 + int64_t(grp.size()) * B * 16;
                }
                return by;
            }
        }
    }
    if (method == PFSHT_METHOD_DIRECT) return direct;
    const int64_t b = estimate_stretch(thlo, thhi, M, L, tol, B, o, single);
    return b < 0 ? direct : b;
}

// Cross-stream ordering of a plan's applies: an apply on stream s waits for the plan's previous one.
Status plan_order_begin(pfsht_plan_t p, hipStream_t s) { return apply_begin(*p, s); }
Status plan_order_end(pfsht_plan_t p, hipStream_t s) { return apply_end(*p, s); }
// The launches of a column apply (dir 0 synthesis, 1 adjoint), with no event recorded or waited for.
Status plan_launches(pfsht_plan_t p, int dir, InCols in, OutCols out, int nb, hipStream_t s) {
    return dir == 0 ? synth_body(*p, in, out, nb, s) : adjoint_body(*p, in, out, nb, s);
}
std::mutex &plan_mutex(pfsht_plan_t p) { return p->mu; }
void plan_shape(pfsht_plan_t p, int &L, int64_t &M, int &B, int &dev) {
    L = p->L;
    M = p->M;
    B = p->B;
    dev = p->dev;
}
// Frees a plan once its last apply has finished, without synchronising the whole device.
void plan_release(pfsht_plan_t p) {
    if (!p) return;
    DeviceGuard dg(p->dev);
    if (p->last) (void)hipEventSynchronize(p->last->e);
    for (int i = 0; i < pfsht_plan_s::kColSlots; ++i)
        if (p->col_used[i]) (void)hipEventSynchronize(p->col_ev[i]);
    delete p;
}

}  // namespace pfsht
