// pfsht_common.hpp — shared declarations of libpfsht_hip: device table structs, kernel launchers,
// error plumbing. gfx950 only.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "geometry.hpp"

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

#define PFSHT_HIP_TRY(expr)                                                                       \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess) {                                                                   \
            ::pfsht::set_error(std::string("HIP error: ") + hipGetErrorString(_e) + " at " #expr); \
            return ::pfsht::kEHip;                                                                \
        }                                                                                         \
    } while (0)

// Ring-pass work block: retained rings [j0, j0 + nr) of order m. Wavefront w of the block owns
// the 64-ring tile tile0 + w (rings j0 + 64 w ..), so all tiles of one order share a compute unit
// and its scalar/L2 caches for the order's coefficient streams.
struct RingItem {
    int32_t m;      // order
    int32_t j0;     // first ring of the block
    int32_t nr;     // rings in the block
    int32_t tile0;  // tile index (within order m) of wavefront 0
};

// Device-resident tables of a ring engine (RingLegendre, src/Legendre.jl:174-255).
struct RingDev {
    int n = 0, L = 0, B = 1, mmax = 0;
    double *costh = nullptr;    // [n] cos theta, sign-flipped when reflected
    double *log2sin = nullptr;  // [n]
    double *log2c = nullptr;    // [L+1] log2 sectorial prefactor
    double2 *ab = nullptr;      // recurrence coefficients (a_l^m, b_l^m), l = m+2..L, per m
    int64_t *abofs = nullptr;   // [mmax+1] offset of order m in ab
    int32_t *jlo = nullptr, *jhi = nullptr;  // [mmax+1]
    RingItem *items = nullptr;  // forward blocks (all tiles of an order), descending cost
    int nitems = 0, fwd_waves = 1;
    RingItem *aitems = nullptr;  // adjoint blocks (<= kAdjWaves tiles of an order), descending cost
    int naitems = 0, adj_waves = 1;
    int64_t *pofs = nullptr;    // [mmax+1] offset of order m's adjoint partials (in double4 units)
    int32_t *ntiles = nullptr;  // [mmax+1]
    double4 *partials = nullptr;  // adjoint per-tile partial sums: (re+, im+, re-, im-) per (tile, l)
    int64_t npartials = 0;
};

constexpr int kFwdMaxWaves = 8;   // forward block: up to 512 rings of one order
constexpr int kAdjWaves = 4;      // adjoint block: up to 4 tiles of an order, one 16 x 64 LDS tile each

// Launchers (stream-ordered, asynchronous). alm and G use the reference's column-major layouts;
// Gd is the internal ring buffer [b][m + mmax][j].
void launch_ring_fwd(const RingDev &r, const double2 *alm, double2 *Gd, hipStream_t s);
void launch_ring_adj(const RingDev &r, const double2 *Gd, double2 *alm, hipStream_t s);
// the two launches of launch_ring_adj: per-tile partial sums, then their ordered reduction into alm
void launch_ring_adj_tiles(const RingDev &r, const double2 *Gd, hipStream_t s);
void launch_ring_adj_finalize(const RingDev &r, double2 *alm, hipStream_t s);
void launch_gd_to_ref(const RingDev &r, const double2 *Gd, double2 *G, hipStream_t s);
void launch_ref_to_gd(const RingDev &r, const double2 *G, double2 *Gd, hipStream_t s);

struct GridDims {
    int B, nh, nth, nph, mmax, L, kind;
};
// Periodisation in longitude-spectral space, order-column layout Ds [b][m + mmax][i] (i < nth).
// C is the storage of the grid buffers: double2, or float2 in mixed precision.
template <class C>
void launch_spec_dfs(const GridDims &g, const RingDev &r, const double2 *Gd, const double *wth, C *Ds, hipStream_t s);
template <class C>
void launch_spec_fold(const GridDims &g, const C *Ds, const double *wth, double2 *Gd, hipStream_t s);

// Direct leaf: phase sums over the ring-engine output of theta-sorted points.
void launch_direct_synth(const double2 *Gd, const double *phis, const int32_t *mlim, const int32_t *perm,
                         double2 *f, int64_t M, int mmax, int B, hipStream_t s);
void launch_direct_adjoint(const double2 *f, const double *phis, const int32_t *mlim, const int32_t *perm,
                           double2 *Gd, int64_t M, int mmax, int B, hipStream_t s);

// ES NUFFT on the stretched torus.
struct NufftDev {
    int B = 1, n1 = 0, n2 = 0;     // modes (nth, nph)
    int N1 = 0, N2 = 0, w = 0;     // fine grid and kernel width
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
    int32_t *bin_ofs = nullptr;           // [nst*16 + 1] point ranges per (super-tile, local bin)
    void *Wst = nullptr;                  // [B][nst][win][win] super-tile windows (grid storage type)
};
// LDS of one interpolation workgroup: the super-tile window of the fine grid.
inline size_t nufft_window_lds_bytes(int win) { return sizeof(double2) * size_t(win) * win; }
// LDS of one spreading workgroup: the window plus, per wavefront, the staging of one 16-point chunk
// (two 16-row kernel tiles with rows kRowPad doubles apart, and 16 complex f values), in doubles.
constexpr int kRowPad = 17;
constexpr int kSpreadStage = 2 * 16 * kRowPad + 2 * 16;
constexpr int kSpreadWaves = 8;  // wavefronts per spreading workgroup: two of the 16 bins each
inline size_t spread_lds_bytes(int win) {
    return nufft_window_lds_bytes(win) + sizeof(double) * kSpreadWaves * kSpreadStage;
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
template <class C>
void launch_interp(const NufftDev &u, const C *U, double2 *f, hipStream_t s);
template <class C>
void launch_spread(const NufftDev &u, const double2 *f, C *U, hipStream_t s);

}  // namespace pfsht
