// grid_fft_kernels.hip — the stretched apply's grid stage (src/Stretch.jl:262-295 between the ring
// pass and the NUFFT) as two kernels per direction, each running its FFTs in LDS (lds_fft.hpp):
//
// synthesis  column kernel, one workgroup per (order column c, transform b): window + DFS
//            periodisation of the ring column (src/DFS.jl:13-56 in longitude-spectral space, as
//            spec_dfs_kernel), colatitude FFT of length nth, phase + deconvolution into the fine
//            colatitude layout (src/NUFFT.jl type-2 deconvolution, as deconv_pad_ut_kernel), inverse
//            FFT of length N1, and the stored fine rows of the column out to T1 [b][c][pr];
//            row kernel, one workgroup per (stored row pr, b): the row's retained columns from T1
//            (zero elsewhere), inverse FFT of length N2, the row out to U [b][pr][l2].
// adjoint    row kernel: forward FFT of a stored row of U, its retained columns out to
//            T2 [b][pr][c]; column kernel: the column from T2 (rows not stored are zero), forward FFT
//            of length N1, deconvolution back to the nth colatitude modes (deconv_trunc_ut_kernel),
//            inverse FFT of length nth and the DFS fold + window into the ring column
//            (spec_fold_kernel).
//
// Each kernel reads and writes its data once in HBM; the transposes between the column and row
// orders happen in those reads: the synthesis row kernel and the adjoint column kernel gather
// elements 16 B (8 B) apart in lines shared by neighbouring rows (columns), and their workgroups are
// numbered so that neighbours run on one XCD (hardware dispatch sends workgroup w to XCD w mod 8),
// so each line comes from HBM once into that XCD's L2.
#include <algorithm>

#include "lds_fft.hpp"
#include "pfsht_common.hpp"

namespace pfsht {

namespace {

constexpr int E = kFftElems;  // values a thread holds in a pass

template <class C> __device__ __forceinline__ C czero() { return cx<C>(0, 0); }

// workgroup w of a grid of 8 ceil(n/8) -> item (w mod 8) ceil(n/8) + w / 8: consecutive items on
// one XCD
__device__ __forceinline__ int xcd_item(int w, int per) { return (w & 7) * per + (w >> 3); }

__device__ __forceinline__ int fine_mode(int l, int n, int N) {
    if (l < (n + 1) / 2) return l;
    if (l >= N - n / 2) return l - N + n;
    return -1;
}

// Kernel bodies: item = the order column c (column kernels) or the stored row pr (row kernels), b =
// the transform; g holds the lengths, the sizes and the twiddle table. The plain kernels take them
// from the launch, the group kernels (a plan group's members in one launch) from a GridTab per plan.
template <class C>
__device__ __forceinline__ void col_synth_body(const double2 *__restrict__ Gd, const int32_t *__restrict__ jlo,
                                               const int32_t *__restrict__ jhi, const double *__restrict__ wth,
                                               const double2 *__restrict__ pc1, const double *__restrict__ c2,
                                               C *__restrict__ T1, const GridFftDev &g, int c, int b) {
    extern __shared__ __align__(16) unsigned char lds_gf_raw[];
    C *x = reinterpret_cast<C *>(lds_gf_raw);
    const int NT = int(blockDim.x);
    const GridFftArgs &a = g.args;
    const C *tw = static_cast<const C *>(g.tw);
    const int64_t ncol = 2 * a.mmax + 1;
    const int m = c - a.mmax, am = m < 0 ? -m : m;
    const int lo = jlo[am], hi = jhi[am];
    const int nh = a.nh, nth = a.nth;
    const double2 *G = Gd + (int64_t(b) * ncol + c) * nh;
    const bool odd = (am + a.spin) & 1;
    for (int i = threadIdx.x; i < nth; i += NT) {
        const int j = i < nh ? i : nth - 1 - i;
        double2 v = make_double2(0.0, 0.0);
        if (j >= lo && j <= hi) {
            const double2 gv = G[j];
            double w = wth[j];
            if (i >= nh && odd) w = -w;
            v = make_double2(gv.x * w, gv.y * w);
        }
        stc(x + lds_slot<true>(i), v);
    }
    __syncthreads();
    lds_fft<false, E, true>(x, g.th, tw + g.th.tw, NT);
    // the fine colatitude layout: mode i of the nth-point transform at fine index l1 (zero padding
    // between), times the phase and the deconvolution
    const int N1 = a.N1, n1 = nth;
    C v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int l1 = int(threadIdx.x) + q * NT;
        if (l1 < N1) {
            const int i = fine_mode(l1, n1, N1);
            v[q] = i >= 0 ? x[lds_slot<true>(i)] : czero<C>();
        }
    }
    __syncthreads();
    const double sc = c2[m >= 0 ? m : m + a.nph] * a.scale;
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int l1 = int(threadIdx.x) + q * NT;
        if (l1 < N1) {
            const int i = fine_mode(l1, n1, N1);
            double2 r = make_double2(0.0, 0.0);
            if (i >= 0) {
                const double2 F = ldc(&v[q]);
                const double2 p = pc1[i];
                r = make_double2((F.x * p.x - F.y * p.y) * sc, (F.x * p.y + F.y * p.x) * sc);
            }
            stc(x + lds_slot<true>(l1), r);
        }
    }
    __syncthreads();
    lds_fft<true, E, true>(x, g.c1, tw + g.c1.tw, NT);
    C *out = T1 + (int64_t(b) * ncol + c) * a.nR;
    for (int pr = threadIdx.x; pr < a.nR; pr += NT) {
        int l1 = a.r0 + pr;
        if (l1 >= N1) l1 -= N1;
        out[pr] = x[lds_slot<true>(l1)];
    }
}

// Rows of N2 <= 8192 values run lds_fft over swizzled slots; longer ones (SPLIT) the radix-2 split
// over natural slots, whose outputs sit at x[k / 2] (k even) and x[N2 / 2 + k / 2] (k odd).
template <bool SPLIT> __device__ __forceinline__ int row_slot(int k, int h) {
    if (SPLIT) return (k & 1) ? h + (k >> 1) : (k >> 1);
    return lds_slot<true>(k);
}

template <bool INV, bool SPLIT, class C>
__device__ __forceinline__ void row_fft(C *x, const GridFftDev &g, const C *tw, int NT) {
    if constexpr (SPLIT) lds_fft_split<INV, E>(x, g.args.N2, g.r2h, tw + g.r2.tw, tw + g.r2h.tw, NT);
    else lds_fft<INV, E, true>(x, g.r2, tw + g.r2.tw, NT);
}

template <class C, bool SPLIT>
__device__ __forceinline__ void row_synth_body(const C *__restrict__ T1, C *__restrict__ U, const GridFftDev &g,
                                               int pr, int b) {
    extern __shared__ __align__(16) unsigned char lds_gf_raw[];
    C *x = reinterpret_cast<C *>(lds_gf_raw);
    const int NT = int(blockDim.x);
    const GridFftArgs &a = g.args;
    const int64_t ncol = 2 * a.mmax + 1;
    const int N2 = a.N2, mmax = a.mmax;
    const C *in = T1 + int64_t(b) * ncol * a.nR + pr;
    for (int l2 = threadIdx.x; l2 < N2; l2 += NT) {
        const int m = l2 <= mmax ? l2 : (l2 >= N2 - mmax ? l2 - N2 : INT32_MIN);
        x[SPLIT ? l2 : lds_slot<true>(l2)] = m != INT32_MIN ? in[int64_t(m + mmax) * a.nR] : czero<C>();
    }
    __syncthreads();
    row_fft<true, SPLIT>(x, g, static_cast<const C *>(g.tw), NT);
    C *out = U + (int64_t(b) * a.nR + pr) * N2;
    for (int l2 = threadIdx.x; l2 < N2; l2 += NT) out[l2] = x[row_slot<SPLIT>(l2, N2 / 2)];
}

template <class C, bool SPLIT>
__device__ __forceinline__ void row_adj_body(const C *__restrict__ U, C *__restrict__ T2, const GridFftDev &g,
                                             int pr, int b) {
    extern __shared__ __align__(16) unsigned char lds_gf_raw[];
    C *x = reinterpret_cast<C *>(lds_gf_raw);
    const int NT = int(blockDim.x);
    const GridFftArgs &a = g.args;
    const int64_t ncol = 2 * a.mmax + 1;
    const int N2 = a.N2;
    const C *in = U + (int64_t(b) * a.nR + pr) * N2;
    for (int l2 = threadIdx.x; l2 < N2; l2 += NT) x[SPLIT ? l2 : lds_slot<true>(l2)] = in[l2];
    __syncthreads();
    row_fft<false, SPLIT>(x, g, static_cast<const C *>(g.tw), NT);
    C *out = T2 + (int64_t(b) * a.nR + pr) * ncol;
    for (int c = threadIdx.x; c < ncol; c += NT) {
        const int m = c - a.mmax;
        out[c] = x[row_slot<SPLIT>(m >= 0 ? m : m + N2, N2 / 2)];
    }
}

template <class C>
__device__ __forceinline__ void col_adj_body(const C *__restrict__ T2, double2 *__restrict__ Gd,
                                             const double *__restrict__ wth, const double2 *__restrict__ pc1,
                                             const double *__restrict__ c2, const GridFftDev &g, int c, int b) {
    extern __shared__ __align__(16) unsigned char lds_gf_raw[];
    C *x = reinterpret_cast<C *>(lds_gf_raw);
    const int NT = int(blockDim.x);
    const GridFftArgs &a = g.args;
    const C *tw = static_cast<const C *>(g.tw);
    const int64_t ncol = 2 * a.mmax + 1;
    const int m = c - a.mmax, am = m < 0 ? -m : m;
    const int N1 = a.N1, nth = a.nth, n1 = nth, nh = a.nh;
    const C *in = T2 + int64_t(b) * a.nR * ncol + c;
    for (int l1 = threadIdx.x; l1 < N1; l1 += NT) {
        int pr = l1 - a.r0;
        if (pr < 0) pr += N1;
        x[lds_slot<true>(l1)] = pr < a.nR ? in[int64_t(pr) * ncol] : czero<C>();
    }
    __syncthreads();
    lds_fft<false, E, true>(x, g.c1, tw + g.c1.tw, NT);
    // back to the nth colatitude modes, times the conjugate phase and the deconvolution
    C v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int i = int(threadIdx.x) + q * NT;
        if (i < n1) v[q] = x[lds_slot<true>(i < (n1 + 1) / 2 ? i : i - n1 + N1)];
    }
    __syncthreads();
    const double sc = c2[m >= 0 ? m : m + a.nph] * a.scale;
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int i = int(threadIdx.x) + q * NT;
        if (i < n1) {
            const double2 u = ldc(&v[q]);
            const double2 p = pc1[i];
            stc(x + lds_slot<true>(i), make_double2((u.x * p.x + u.y * p.y) * sc, (u.y * p.x - u.x * p.y) * sc));
        }
    }
    __syncthreads();
    lds_fft<true, E, true>(x, g.th, tw + g.th.tw, NT);
    double2 *G = Gd + (int64_t(b) * ncol + c) * nh;
    const bool odd = (am + a.spin) & 1;
    for (int j = threadIdx.x; j < nh; j += NT) {
        double2 v2 = ldc(x + lds_slot<true>(j));
        if (nth > nh) {
            const double2 r = ldc(x + lds_slot<true>(nth - 1 - j));
            v2 = odd ? make_double2(v2.x - r.x, v2.y - r.y) : make_double2(v2.x + r.x, v2.y + r.y);
        }
        const double w = wth[j];
        G[j] = make_double2(v2.x * w, v2.y * w);
    }
}

// One plan: workgroup (item, transform) from the grid; the row synthesis and the column adjoint
// number their items XCD-wise (xcd_item over 8 ceil(count / 8) workgroups).
template <class C>
__global__ void __launch_bounds__(1024)
grid_col_synth_kernel(const double2 *__restrict__ Gd, const int32_t *__restrict__ jlo, const int32_t *__restrict__ jhi,
                      const double *__restrict__ wth, const double2 *__restrict__ pc1, const double *__restrict__ c2,
                      C *__restrict__ T1, GridFftDev g) {
    col_synth_body<C>(Gd, jlo, jhi, wth, pc1, c2, T1, g, blockIdx.x, blockIdx.y);
}

template <class C, bool SPLIT>
__global__ void __launch_bounds__(1024)
grid_row_synth_kernel(const C *__restrict__ T1, C *__restrict__ U, GridFftDev g, int per) {
    const int pr = xcd_item(blockIdx.x, per);
    if (pr >= g.args.nR) return;
    row_synth_body<C, SPLIT>(T1, U, g, pr, blockIdx.y);
}

template <class C, bool SPLIT>
__global__ void __launch_bounds__(1024)
grid_row_adj_kernel(const C *__restrict__ U, C *__restrict__ T2, GridFftDev g) {
    row_adj_body<C, SPLIT>(U, T2, g, blockIdx.x, blockIdx.y);
}

template <class C>
__global__ void __launch_bounds__(1024)
grid_col_adj_kernel(const C *__restrict__ T2, double2 *__restrict__ Gd, const double *__restrict__ wth,
                    const double2 *__restrict__ pc1, const double *__restrict__ c2, GridFftDev g, int per) {
    const int c = xcd_item(blockIdx.x, per);
    if (c >= 2 * g.args.mmax + 1) return;
    col_adj_body<C>(T2, Gd, wth, pc1, c2, g, c, blockIdx.y);
}

// A plan group: workgroup i takes work[i] = (plan, item), the plan's lengths and buffers from
// tabs[plan] (one transform per plan; no split rows). The row synthesis and column adjoint items
// are XCD-numbered within each plan, whose item lists start at multiples of eight. The plan's
// GridFftDev is copied to LDS first: the passes index its radix arrays at run time, which in a
// register copy would go to scratch.
__device__ __forceinline__ const GridFftDev &group_dev(const GridTab &T) {
    constexpr int nw = int(sizeof(GridFftDev) / sizeof(int));
    static_assert(sizeof(GridFftDev) % sizeof(int) == 0 && alignof(GridFftDev) <= 8, "copied as words");
    __shared__ __align__(8) int gsh[nw];
    const int *src = reinterpret_cast<const int *>(&T.g);
    for (int i = threadIdx.x; i < nw; i += blockDim.x) gsh[i] = src[i];
    __syncthreads();
    return *reinterpret_cast<const GridFftDev *>(gsh);
}
template <class C>
__global__ void __launch_bounds__(1024)
grid_col_synth_group(const GridTab *__restrict__ tabs, const int2 *__restrict__ work) {
    const int2 w = work[blockIdx.x];
    const GridTab &T = tabs[w.x];
    col_synth_body<C>(T.Gd, T.jlo, T.jhi, T.wth, T.pc1, T.c2, static_cast<C *>(T.Ut), group_dev(T), w.y, 0);
}

template <class C>
__global__ void __launch_bounds__(1024)
grid_row_synth_group(const GridTab *__restrict__ tabs, const int2 *__restrict__ work) {
    const int2 w = work[blockIdx.x];
    const GridTab &T = tabs[w.x];
    const GridFftDev &g = group_dev(T);
    const int pr = xcd_item(w.y, (g.args.nR + 7) / 8);
    if (pr >= g.args.nR) return;
    row_synth_body<C, false>(static_cast<const C *>(T.Ut), static_cast<C *>(T.U), g, pr, 0);
}

template <class C>
__global__ void __launch_bounds__(1024)
grid_row_adj_group(const GridTab *__restrict__ tabs, const int2 *__restrict__ work) {
    const int2 w = work[blockIdx.x];
    const GridTab &T = tabs[w.x];
    row_adj_body<C, false>(static_cast<const C *>(T.U), static_cast<C *>(T.Ut), group_dev(T), w.y, 0);
}

template <class C>
__global__ void __launch_bounds__(1024)
grid_col_adj_group(const GridTab *__restrict__ tabs, const int2 *__restrict__ work) {
    const int2 w = work[blockIdx.x];
    const GridTab &T = tabs[w.x];
    const GridFftDev &g = group_dev(T);
    const int ncol = 2 * g.args.mmax + 1;
    const int c = xcd_item(w.y, (ncol + 7) / 8);
    if (c >= ncol) return;
    col_adj_body<C>(static_cast<const C *>(T.Ut), T.Gd, T.wth, T.pc1, T.c2, g, c, 0);
}

template <class C>
size_t lds_of(int n) { return sizeof(C) * size_t(lds_slots(n)); }

}  // namespace

// Eight values per thread on ceil(n / 8) threads rounded up to whole wavefronts (64 .. 1024: lengths
// up to 8192; longer rows split in two), over swizzled LDS slots. Eight measured 15-25% faster than
// sixteen at the apply's lengths (more wavefronts in flight per transform, fewer registers; sixteen
// on 1024 threads spills); whole wavefronts rather than powers of two keep the lanes busy (1440
// values: 192 threads, not 256).
int grid_fft_threads(int n) {
    const int t = (n + E - 1) / E;
    return std::min(1024, std::max(64, (t + 63) / 64 * 64));
}

template <class C>
void launch_grid_synth_cols(const GridFftDev &g, const RingDev &r, const double2 *Gd, const double *wth,
                            const double2 *pc1, const double *c2, C *T1, int nb, hipStream_t s) {
    const GridFftArgs &a = g.args;
    grid_col_synth_kernel<C><<<dim3(2 * a.mmax + 1, nb), g.nt_col, lds_of<C>(a.N1), s>>>(
        Gd, r.jlo, r.jhi, wth, pc1, c2, T1, g);
}

template <class C>
void launch_grid_synth_rows(const GridFftDev &g, const C *T1, C *U, int nb, hipStream_t s) {
    const GridFftArgs &a = g.args;
    const size_t lds = lds_of<C>(a.N2);
    const int per = (a.nR + 7) / 8;
    if (g.split) grid_row_synth_kernel<C, true><<<dim3(8 * per, nb), g.nt_row, lds, s>>>(T1, U, g, per);
    else grid_row_synth_kernel<C, false><<<dim3(8 * per, nb), g.nt_row, lds, s>>>(T1, U, g, per);
}

template <class C>
void launch_grid_adjoint_rows(const GridFftDev &g, const C *U, C *T2, int nb, hipStream_t s) {
    const GridFftArgs &a = g.args;
    const size_t lds = lds_of<C>(a.N2);
    if (g.split) grid_row_adj_kernel<C, true><<<dim3(a.nR, nb), g.nt_row, lds, s>>>(U, T2, g);
    else grid_row_adj_kernel<C, false><<<dim3(a.nR, nb), g.nt_row, lds, s>>>(U, T2, g);
}

template <class C>
void launch_grid_adjoint_cols(const GridFftDev &g, const C *T2, const double *wth, const double2 *pc1,
                              const double *c2, double2 *Gd, int nb, hipStream_t s) {
    const GridFftArgs &a = g.args;
    const int per = (2 * a.mmax + 1 + 7) / 8;
    grid_col_adj_kernel<C><<<dim3(8 * per, nb), g.nt_col, lds_of<C>(a.N1), s>>>(T2, Gd, wth, pc1, c2, g, per);
}

template <class C>
void launch_grid_synth_group(const GridTab *tabs, const GridGroupWork &w, hipStream_t s) {
    if (w.ncol > 0)
        grid_col_synth_group<C><<<dim3(w.ncol), w.nt_col, w.lds_col, s>>>(tabs, w.col);
    if (w.nrow_s > 0)
        grid_row_synth_group<C><<<dim3(w.nrow_s), w.nt_row, w.lds_row, s>>>(tabs, w.row_s);
}

template <class C>
void launch_grid_adjoint_group(const GridTab *tabs, const GridGroupWork &w, hipStream_t s) {
    if (w.nrow_a > 0)
        grid_row_adj_group<C><<<dim3(w.nrow_a), w.nt_row, w.lds_row, s>>>(tabs, w.row_a);
    if (w.ncol_a > 0)
        grid_col_adj_group<C><<<dim3(w.ncol_a), w.nt_col, w.lds_col, s>>>(tabs, w.col_a);
}

#define PFSHT_GF_INST(C)                                                                                          \
    template void launch_grid_synth_cols<C>(const GridFftDev &, const RingDev &, const double2 *, const double *, \
                                            const double2 *, const double *, C *, int, hipStream_t);             \
    template void launch_grid_synth_rows<C>(const GridFftDev &, const C *, C *, int, hipStream_t);                \
    template void launch_grid_adjoint_rows<C>(const GridFftDev &, const C *, C *, int, hipStream_t);              \
    template void launch_grid_adjoint_cols<C>(const GridFftDev &, const C *, const double *, const double2 *,     \
                                              const double *, double2 *, int, hipStream_t);                    \
    template void launch_grid_synth_group<C>(const GridTab *, const GridGroupWork &, hipStream_t);                \
    template void launch_grid_adjoint_group<C>(const GridTab *, const GridGroupWork &, hipStream_t);
PFSHT_GF_INST(double2)
PFSHT_GF_INST(float2)
#undef PFSHT_GF_INST

}  // namespace pfsht
