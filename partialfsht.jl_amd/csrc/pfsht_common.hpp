// pfsht_common.hpp — shared declarations of libpfsht_hip: device table structs, kernel launchers,
// error plumbing. gfx950 only.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "geometry.hpp"
#include "lds_fft.hpp"
#include "point_bins.hpp"

namespace pfsht {

void set_error(const std::string &msg);

// Storage of the grid buffers: complex FP64 (double2) or, in mixed precision, FP32 (float2).
// Kernels load into FP64, compute in FP64 and round on store.
__device__ __forceinline__ double2 ldc(const double2 *p) { return *p; }
__device__ __forceinline__ double2 ldc(const float2 *p) {
    const float2 v = *p;
    return make_double2(double(v.x), double(v.y));
}
__device__ __forceinline__ void stc(double2 *p, double2 v) { *p = v; }
__device__ __forceinline__ void stc(float2 *p, double2 v) { *p = make_float2(float(v.x), float(v.y)); }


// Status propagated through the C ABI (values of the PFSHT_E_* codes in include/pfsht.h).
enum Status { kOk = 0, kEDim = 1, kEArg = 2, kEGeom = 3, kENomem = 4, kEHip = 5, kEUnsupported = 6 };

// The transform columns of an apply's input or output in the caller's memory: column b (b < nb) at
// base + b * stride elements, or at tab[b] when tab is set (a device array of column pointers: an
// apply that packs the columns of several requests reads and writes them in place).
// Columns live in device (global) memory, and col() returns pointers of that address space, so
// that loads and stores through them are global instructions: a generic (flat) access also counts
// against the LDS wait counter, and a kernel's waits for its LDS operations would then wait for its
// global prefetches as well.
#define PFSHT_GLOBAL __attribute__((address_space(1)))
typedef const PFSHT_GLOBAL double2 *gcptr;
typedef PFSHT_GLOBAL double2 *gptr;
// complex loads and stores through global pointers (the vector type's copy operations take generic
// references only)
__device__ __forceinline__ double2 gload(gcptr p) {
    const PFSHT_GLOBAL double *q = (const PFSHT_GLOBAL double *)p;
    return make_double2(q[0], q[1]);
}
__device__ __forceinline__ void gstore(gptr p, double2 v) {
    PFSHT_GLOBAL double *q = (PFSHT_GLOBAL double *)p;
    q[0] = v.x;
    q[1] = v.y;
}
struct InCols {
    const double2 *base = nullptr;
    int64_t stride = 0;
    const double2 *const *tab = nullptr;
    __device__ __forceinline__ gcptr col(int b) const {
        return tab ? (gcptr)tab[b] : (gcptr)(base + int64_t(b) * stride);
    }
};
struct OutCols {
    double2 *base = nullptr;
    int64_t stride = 0;
    double2 *const *tab = nullptr;
    __device__ __forceinline__ gptr col(int b) const {
        return tab ? (gptr)tab[b] : (gptr)(base + int64_t(b) * stride);
    }
};
inline InCols in_cols(const double2 *base, int64_t stride) { InCols c; c.base = base; c.stride = stride; return c; }
inline OutCols out_cols(double2 *base, int64_t stride) { OutCols c; c.base = base; c.stride = stride; return c; }
// a window of columns: columns b0 .. of c
inline InCols cols_from(InCols c, int b0) {
    if (c.tab) c.tab += b0;
    else c.base += int64_t(b0) * c.stride;
    return c;
}
inline OutCols cols_from(OutCols c, int b0) {
    if (c.tab) c.tab += b0;
    else c.base += int64_t(b0) * c.stride;
    return c;
}

#define PFSHT_HIP_TRY(expr)                                                                       \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess) {                                                                   \
            ::pfsht::set_error(std::string("HIP error: ") + hipGetErrorString(_e) + " at " #expr); \
            return ::pfsht::kEHip;                                                                \
        }                                                                                         \
    } while (0)

// Ring-pass work. Order m's degrees m..L are cut into nseg segments of whole 64-degree blocks
// (block k = degrees m + 64 k .. m + 64 k + 63); the recurrence is renormalised after the last
// degree of every block, and segment s >= 1 starts from the state (q_{l-1}, q_{l-2}, E) that a
// plan-time sweep stored for it (the checkpoints). One item is one 64-ring tile of one order:
// the forward runs it as a workgroup with one wavefront per segment, the adjoint as nseg
// independent wavefronts (one item each, seg set).
struct SegItem {
    int32_t m;     // order
    int32_t j0;    // first ring of the tile
    int32_t nr;    // rings in the tile, <= 64 (0: an empty slot)
    int32_t tile;  // tile index within order m
    int32_t nseg;  // degree segments of order m
    int32_t seg;   // adjoint items: this wavefront's segment
    int64_t ck;    // checkpoint row of segment 1 of this tile; segment s uses row ck + s - 1
    int32_t nr2;   // adjoint items with two tiles: rings of the second tile (0: none)
};

constexpr int kSegMax = 8;     // segments per order (forward workgroup: one wavefront each)
constexpr int kSegBlocks = 3;  // default target of 64-degree blocks per segment (2..8 measured: 3 best)
constexpr int kAdjWaves = 4;   // adjoint workgroup: 4 wavefronts, one 16 x 64 LDS tile each

// The MFMA adjoint (two or more transforms on plans whose orders have at most 16 ring tiles): one
// workgroup per (order, degree segment) with the order's 64-ring tiles dealt to its wavefronts, tile
// w + waves k for k < tg; the tiles' ring sums are added in LDS in tile order, so no partials leave
// the workgroup.
struct OrderSeg {
    int32_t m;      // order
    int32_t nseg;   // degree segments of the order
    int32_t seg;    // this workgroup's segment
    int32_t ntile;  // 64-ring tiles of the order (0: an empty slot)
    int32_t j0;     // first ring of tile 0
    int32_t nrl;    // rings of the last tile
    int64_t ck;     // checkpoint row of tile 0 at segment 1; tile t, segment s >= 1: ck + t (nseg - 1) + s - 1
};

// First degree of segment s of order m (s = nseg gives a degree past L).
__host__ __device__ inline int seg_begin(int m, int L, int nseg, int s) {
    const int nblk = (L - m + 64) >> 6;  // ceil((L - m + 1) / 64)
    return m + 64 * ((s * nblk) / nseg);
}
// Segments of order m for a target of `per` blocks per segment.
inline int seg_count(int m, int L, int per) {
    const int nblk = (L - m + 64) >> 6;
    const int k = (nblk + per - 1) / per;
    return k < 1 ? 1 : (k > kSegMax ? kSegMax : k);
}

// Device-resident tables of a ring engine (RingLegendre, src/Legendre.jl:174-255).
struct RingDev {
    int n = 0, L = 0, B = 1, mmax = 0;
    double *costh = nullptr;    // [n] cos theta, sign-flipped when reflected
    double *log2sin = nullptr;  // [n]
    double *log2c = nullptr;    // [L+1] log2 sectorial prefactor
    double2 *ab = nullptr;      // recurrence coefficients (a_l^m, b_l^m), l = m+2..L, per m
    double2 *gc = nullptr;      // the rescaled forward's (gamma_l, C_l), same indexing as ab
    int64_t *abofs = nullptr;   // [mmax+1] offset of order m in ab
    int32_t *jlo = nullptr, *jhi = nullptr;  // [mmax+1]
    SegItem *fitems = nullptr;  // forward workgroups (one tile, nseg wavefronts), XCD-interleaved
    int nfitems = 0, fwd_waves = 1;
    SegItem *aitems = nullptr;  // adjoint wavefronts, kAdjWaves per workgroup, XCD-interleaved
    int naitems = 0;            // a multiple of kAdjWaves
    double2 *ckq = nullptr;     // [nck][64] checkpoint (q_{l-1}, q_{l-2}) per lane
    int32_t *ckE = nullptr;     // [nck][64] checkpoint exponent
    int64_t nck = 0;
    int64_t *pofs = nullptr;    // [mmax+1] offset of order m's adjoint partials (in double4 units)
    int32_t *ntiles = nullptr;  // [mmax+1]
    double4 *partials = nullptr;  // adjoint partial sums: (re+, im+, re-, im-) per (tile group, l)
    int64_t npartials = 0;
    int adj_tg = 1;             // 64-ring tiles per adjoint wavefront (a tile group); ntiles counts groups
    OrderSeg *mitems = nullptr;  // MFMA adjoint workgroups, XCD-interleaved
    int nmitems = 0;
    int mfma_waves = 0;         // wavefronts per MFMA adjoint workgroup (0: no MFMA adjoint)
    int mfma_tg = 1;            // tiles per wavefront
    int part_b = 1;             // transform stride of the partials (1 when only single-transform
                                // adjoints take the partials path)
};

// Spin-weighted ring engine (src/Spin.jl:170-545), used with a RingDev that supplies n, L, B, mmax,
// cos theta (negated when reflected) and the per-|m| ring ranges; orders m in [-mmax, mmax].
struct SpinDev {
    int s = 0;                  // spin weight (nonzero)
    double *l2c = nullptr;      // [n] log2 cos(theta/2) (log2 sin(theta/2) when reflected)
    double *l2s = nullptr;      // [n] log2 sin(theta/2) (log2 cos(theta/2) when reflected)
    double4 *seed = nullptr;    // [2L+1] per m + L: (log2 sqrt C(2 l0, l0 - a), sign, p_cos, p_sin)
    double *Nl = nullptr;       // [L+1] sqrt((2l+1)/4pi)
    double4 *tab = nullptr;     // per (|m|, l >= l0): (a1, |a2|, a0, N_{l+1}) of the step l -> l+1
    int64_t *tofs = nullptr;    // [L+1] first entry of |m| in tab
    int4 *items = nullptr;      // one wavefront per (order m, 64-ring tile): (m, j0, rings, tile)
    int nitems = 0;
    int64_t *pofs = nullptr;    // [2 mmax + 1] per m + mmax: first adjoint partial of the order
    int32_t *ntiles = nullptr;  // [2 mmax + 1] tiles of the order
    double2 *partials = nullptr;  // adjoint per-(tile, l, b) ring sums, l >= max(|m|, |s|)
};
void launch_spin_tables(const RingDev &r, const SpinDev &sd, hipStream_t s);
void launch_spin_fwd(const RingDev &r, const SpinDev &sd, InCols alm, int nb, double2 *Gd, hipStream_t s);
// writes every slot of alm (zero outside the retained modes)
void launch_spin_adj(const RingDev &r, const SpinDev &sd, const double2 *Gd, OutCols alm, int nb, hipStream_t s);

// Degrees of one adjoint tile's partial rows for order m: L - m + 1 rounded up to whole 16-degree
// sub-tiles, so that a sub-tile's stores never need a bounds test.
__host__ __device__ inline int adj_rows_padded(int m, int L) { return ((L - m + 1) + 15) & ~15; }

// The plan-time sweep that stores the checkpoints of every forward item (default stream).
void launch_ring_checkpoints(const RingDev &r, hipStream_t s);

// Launchers (stream-ordered, asynchronous). alm and G use the reference's column-major layouts;
// Gd is the internal ring buffer [b][m + mmax][j].
// nb transforms (nb <= r.B) in columns alm.col(0 .. nb - 1); Gd holds B transforms, the first nb used.
// One single-transform engine of a plan group, as the grouped ring launches read it (device memory).
struct RingTab {
    const SegItem *fitems, *aitems;
    const double *costh, *log2sin, *log2c;
    const double2 *ab, *gc;
    const int64_t *abofs;
    const double2 *ckq;
    const int32_t *ckE;
    const int64_t *pofs;
    const int32_t *ntiles;
    double4 *partials;
    double2 *Gd;
    InCols alm_in;    // the synthesis's coefficients
    OutCols alm_out;  // the adjoint's coefficients
    int L, n, mmax, part_b;
};
// Grouped ring passes: work lists of {engine, forward item} and {engine, adjoint workgroup} (waves:
// the largest segment count of the engines; tg: the adjoint tiles per wavefront, the same for every
// engine in one launch), and of {engine, row block, |m|, 0} for the finalize.
void launch_ring_fwd_group(const RingTab *tabs, const int2 *work, int nwork, int waves, hipStream_t s);
void launch_ring_adj_group(const RingTab *tabs, const int2 *work, int nwork, int tg, hipStream_t s);
void launch_ring_finalize_group(const RingTab *tabs, const int4 *work, int nwork, hipStream_t s);
void launch_ring_fwd(const RingDev &r, InCols alm, int nb, double2 *Gd, hipStream_t s);
void launch_ring_adj(const RingDev &r, const double2 *Gd, OutCols alm, int nb, hipStream_t s);
// The adjoint as timed stages: the ring pass (stage 0), then, on the partials path only, the ordered
// reduction of the per-tile partials into alm (launch_ring_adj_finish; nothing on the MFMA path,
// which reduces in LDS and writes alm itself).
void launch_ring_adj_pass(const RingDev &r, const double2 *Gd, OutCols alm, int nb, hipStream_t s);
void launch_ring_adj_finish(const RingDev &r, OutCols alm, int nb, hipStream_t s);
// true when the adjoint of nb transforms runs on the MFMA path (no partials)
bool ring_adj_uses_mfma(const RingDev &r, int nb);
void launch_gd_to_ref(const RingDev &r, const double2 *Gd, double2 *G, hipStream_t s);
void launch_ref_to_gd(const RingDev &r, const double2 *G, double2 *Gd, hipStream_t s);

struct GridDims {
    int B, nh, nth, nph, mmax, L, kind;
    int spin = 0;  // the mirrored half takes (-1)^(m + spin)
};
// Periodisation in longitude-spectral space, order-column layout Ds [b][m + mmax][i] (i < nth).
// C is the storage of the grid buffers: double2, or float2 in mixed precision.
template <class C>
void launch_spec_dfs(const GridDims &g, const RingDev &r, const double2 *Gd, const double *wth, C *Ds, hipStream_t s);
template <class C>
void launch_spec_fold(const GridDims &g, const C *Ds, const double *wth, double2 *Gd, hipStream_t s);

// Direct leaf: phase sums over the ring-engine output of theta-sorted points.
void launch_direct_synth(const double2 *Gd, const double *phis, const int32_t *mlim, const int32_t *perm,
                         OutCols f, int64_t M, int mmax, int nb, hipStream_t s);
void launch_direct_adjoint(InCols f, const double *phis, const int32_t *mlim, const int32_t *perm,
                           double2 *Gd, int64_t M, int mmax, int nb, hipStream_t s);

// ES NUFFT on the stretched torus.
struct NufftDev {
    int B = 1, n1 = 0, n2 = 0;     // modes (nth, nph)
    int N1 = 0, N2 = 0, w = 0;     // fine grid and kernel width
    // Rows of the fine grid that kernels of the points reach: the cyclic interval r0 .. r0 + nR - 1
    // (mod N1). The apply stores only those rows, U [b][pr][l2] with row (r0 + pr) mod N1; the
    // NUFFT seam uses r0 = 0, nR = N1.
    int r0 = 0, nR = 0;
    double beta = 0;
    int64_t M = 0;
    double *X1 = nullptr, *X2 = nullptr;  // sorted points in fine-grid units, base cell in [0, N)
    int32_t *perm = nullptr;              // sorted position -> user index
    double2 *pc1 = nullptr;               // [n1] phase * h1/psihat1 (synthesis; adjoint uses conj)
    double2 *c1 = nullptr;                // [n1] h1/psihat1 as complex (the bare NUFFT, no phase)
    double *c2 = nullptr;                 // [n2] h2/psihat2
    // spreading super-tiles
    int S = 0, ns1 = 0, ns2 = 0, win = 0;  // bin size, super-tile grid, LDS window edge
    int nst = 0;                          // non-empty super-tiles
    int32_t *st_list = nullptr;           // [nst] super-tile ids (s1 * ns2 + s2)
    int32_t *st_slot = nullptr;           // [ns1*ns2] slot in Wst or -1 when empty
    int32_t *bin_ofs = nullptr;           // [nst*17] per slot: the 16 local bins' starts, then the end
    void *Wst = nullptr;                  // [B][nst][win][win] super-tile windows (grid storage type)
};
// One single-transform plan's NUFFT as the plan-group launches read it (device memory); U and Wst
// hold the plan's grid type.
struct NufftTab {
    const double *X1, *X2;
    const int32_t *perm, *st_list, *st_slot, *bin_ofs;
    void *U, *Wst;
    InCols fin;    // the adjoint's field
    OutCols fout;  // the synthesis's field
    int N1, N2, S, ns1, ns2, win, w, nst, r0, nR;
};
// Plan groups: interpolation over {plan, super-tile} work, spreading over {plan, super-tile} then the
// merge over {plan, column block, stored row, 0}; every plan of one launch has kernel width w.
template <class C>
void launch_interp_group(const NufftTab *tabs, const int2 *work, int nwork, int w, int win, hipStream_t s);
template <class C>
void launch_spread_group(const NufftTab *tabs, const int2 *work, int nwork, const int4 *mwork, int nmwork, int w,
                         int win, hipStream_t s);
// LDS of one interpolation workgroup: the super-tile window of the fine grid.
inline size_t nufft_window_lds_bytes(int win) { return sizeof(double2) * size_t(win) * win; }
// LDS of one spreading workgroup: the window plus, per wavefront, the staging of one 16-point chunk
// (two 16-row kernel tiles with rows kRowPad doubles apart, and 16 complex f values), in doubles.
constexpr int kRowPad = 17;
// staging of one wavefront for G transforms: two row tiles and G x 16 complex f values, in doubles
constexpr int spread_stage_doubles(int G) { return 2 * 16 * kRowPad + 2 * 16 * G; }
// transforms per spreading workgroup: four at four or more transforms, two at two or three (the
// accumulators of four take 238 VGPRs, two waves per SIMD); kSpreadGroup is the largest
constexpr int kSpreadGroup = 4;
constexpr int kSpreadWaves = 8;  // wavefronts per spreading workgroup: two of the 16 bins each
// Wavefront w accumulates bins w + kSpreadWaves k, k < 16 / kSpreadWaves: the 16 bins of a super-tile
// are covered exactly once only when the wave count divides 16 (a bin index past 15 would read the
// point ranges of another super-tile, or past the end of bin_ofs).
static_assert(16 % kSpreadWaves == 0 && kSpreadWaves >= 1, "kSpreadWaves must divide the 16 bins of a super-tile");
inline size_t spread_lds_bytes(int win, int G) {
    return nufft_window_lds_bytes(win) + sizeof(double) * kSpreadWaves * spread_stage_doubles(G);
}
// The FP32 spreading kernel (mixed-precision plans): a float2 window and, per wavefront, two 16-row
// float kernel tiles (rows kRowPad32 floats apart: the 16 lanes writing one entry of 16 rows fall in
// different LDS banks) and G x 16 float2 f values.
constexpr int kRowPad32 = 17;
constexpr int spread_stage_floats(int G) { return 2 * 16 * kRowPad32 + 2 * 16 * G; }
inline size_t spread32_lds_bytes(int win, int G) {
    return sizeof(float2) * size_t(win) * win + sizeof(float) * kSpreadWaves * spread_stage_floats(G);
}
// pc1 selects the row factors: u.pc1 for the stretched apply, u.c1 for the bare NUFFT.
void launch_deconv_pad(const NufftDev &u, const double2 *pc1, const double2 *Fhat, double2 *U, hipStream_t s);
void launch_deconv_trunc(const NufftDev &u, const double2 *pc1, const double2 *U, double2 *Fhat, hipStream_t s);
// The pruned fine transform: order-column layout Ut [b][m + mmax][l1] of the nonzero fine columns.
template <class C>
void launch_deconv_pad_ut(const NufftDev &u, const double2 *pc1, int mmax, int nph, double scale, const C *Fc, C *Ut,
                          hipStream_t s);
template <class C>
void launch_deconv_trunc_ut(const NufftDev &u, const double2 *pc1, int mmax, int nph, double scale, const C *Ut, C *Fc,
                            hipStream_t s);
template <class C>
void launch_ut_to_rows(const NufftDev &u, int mmax, const C *Ut, C *U, hipStream_t s);
template <class C>
void launch_rows_to_ut(const NufftDev &u, int mmax, const C *U, C *Ut, hipStream_t s);
// out[b][c][r] = in[b][r][c] for a rows x cols complex matrix per batch entry.
void launch_transpose(const double2 *in, double2 *out, int rows, int cols, int B, hipStream_t s);

// The stretched apply's grid stage as two kernels per direction with their FFTs in LDS
// (grid_fft_kernels.hip): from the ring buffer Gd to the stored fine rows U and back, through
// T1 [b][c][pr] (synthesis) or T2 [b][pr][c] (adjoint) in the Ut buffer.
struct GridFftArgs {
    int nh = 0, nth = 0, mmax = 0, spin = 0, nph = 0;  // ring grid (nth colatitudes, nh rings)
    int N1 = 0, N2 = 0, r0 = 0, nR = 0;                 // fine grid and its stored rows
    double scale = 1.0;                                 // the skipped longitude FFT pair (nph)
};
struct GridFftDev {
    FftLen th, c1, r2;     // lengths nth, N1, N2
    FftLen r2h;            // N2 / 2, when the rows split (N2 > kFftElems x 1024)
    bool split = false;
    void *tw = nullptr;    // twiddle tables of the lengths (grid storage type)
    int nt_col = 128, nt_row = 128;  // workgroup sizes of the column and row kernels
    GridFftArgs args;
};
// workgroup size for an LDS transform of length n (n <= kFftElems * size, size <= 1024)
int grid_fft_threads(int n);
// One single-transform plan of a plan group's grouped grid-stage launches (rows not split).
struct GridTab {
    GridFftDev g;
    double2 *Gd = nullptr;  // the ring buffer
    const int32_t *jlo = nullptr, *jhi = nullptr;
    const double *wth = nullptr;
    const double2 *pc1 = nullptr;
    const double *c2 = nullptr;
    void *Ut = nullptr, *U = nullptr;
};
// The grouped launches' work lists (device) of one storage type: {plan, item} per workgroup, and the
// workgroup size and LDS of each kernel (the largest member's).
struct GridGroupWork {
    int2 *col = nullptr, *row_s = nullptr, *row_a = nullptr, *col_a = nullptr;
    int ncol = 0, nrow_s = 0, nrow_a = 0, ncol_a = 0;
    int nt_col = 64, nt_row = 64;
    size_t lds_col = 0, lds_row = 0;
};
template <class C>
void launch_grid_synth_group(const GridTab *tabs, const GridGroupWork &w, hipStream_t s);
template <class C>
void launch_grid_adjoint_group(const GridTab *tabs, const GridGroupWork &w, hipStream_t s);
// synthesis: the column kernel (ring buffer -> T1), then the row kernel (T1 -> U)
template <class C>
void launch_grid_synth_cols(const GridFftDev &g, const RingDev &r, const double2 *Gd, const double *wth,
                            const double2 *pc1, const double *c2, C *T1, int nb, hipStream_t s);
template <class C>
void launch_grid_synth_rows(const GridFftDev &g, const C *T1, C *U, int nb, hipStream_t s);
// adjoint: the row kernel (U -> T2), then the column kernel (T2 -> ring buffer)
template <class C>
void launch_grid_adjoint_rows(const GridFftDev &g, const C *U, C *T2, int nb, hipStream_t s);
template <class C>
void launch_grid_adjoint_cols(const GridFftDev &g, const C *T2, const double *wth, const double2 *pc1,
                              const double *c2, double2 *Gd, int nb, hipStream_t s);
// The point kernels over u.B transforms (an apply passes a copy with B = nb).
template <class C>
void launch_interp(const NufftDev &u, const C *U, OutCols f, hipStream_t s);
template <class C>
void launch_spread(const NufftDev &u, InCols f, C *U, hipStream_t s);

// Iterative callers on a plan (analysis_kernels.hip). launch_cdot writes sum conj(a) b to *out
// through kDotBlocks partial sums in part, in a fixed order.
constexpr int kDotBlocks = 1024;
void launch_weight_field(double2 *f, const double *w, int64_t M, int B, hipStream_t s);
void launch_cdot(const double2 *a, const double2 *b, int64_t n, double2 *part, double2 *out, hipStream_t s);
void launch_zaxpy(double2 *y, const double2 *x, double2 a, int64_t n, hipStream_t s);
// ComplexF32 -> ComplexF64 and back (n complex elements)
void launch_widen_c64(const float2 *in, double2 *out, int64_t n, hipStream_t s);
void launch_narrow_c128(const double2 *in, float2 *out, int64_t n, hipStream_t s);
void launch_zxpby(double2 *p, const double2 *r, double2 beta, int64_t n, hipStream_t s);
void launch_add_prior(double2 *Ap, const double2 *p, const double *reg, int L, int B, hipStream_t s);
void launch_spectrum(const double2 *alm, int L, double *C, hipStream_t s);
// multi-box plans: the points of one box in and out of the caller's field
void launch_scatter_field(OutCols f, const double2 *fg, const int32_t *idx, int64_t Mg, int nb, hipStream_t s);
void launch_gather_field(double2 *fg, InCols f, const int32_t *idx, int64_t Mg, int nb, hipStream_t s);
// alm.col(b)[i] += acc[b * n + i] for b < nb
void launch_cols_add(OutCols alm, const double2 *acc, int64_t n, int nb, hipStream_t s);

// plan construction on the device
void launch_coef_tables(int L, int mmax, const int64_t *abofs, double2 *ab, double2 *gc, hipStream_t s);
void launch_map_points(const double *theta, const double *phi, int64_t M, const PointMap &q, double *X1, double *X2,
                       int32_t *key, int32_t *idx, int32_t *cnt, uint8_t *on, int32_t *err, hipStream_t s);
void launch_gather_points(const double *X1, const double *X2, const int32_t *perm, int64_t M, double *X1s,
                          double *X2s, hipStream_t s);
hipError_t sort_pairs(void *tmp, size_t &tmp_bytes, const int32_t *kin, int32_t *kout, const int32_t *vin,
                      int32_t *vout, int n, int bits, hipStream_t s);

}  // namespace pfsht
