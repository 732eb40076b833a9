// grid_kernels.hip — the apodisation window and the Double Fourier Sphere periodisation
// (src/Stretch.jl:269-273, src/DFS.jl:13-56), done per order in longitude-spectral space, and a
// batched transpose. All HBM-bound data movement on row-major layouts:
//   Gd [b][m + mmax][j]   ring-pass buffer (rings contiguous, written by the ring kernels)
//   Ds [b][m + mmax][i]   periodised order columns, i in [0, nth)
#include "pfsht_common.hpp"

namespace pfsht {

constexpr int kTile = 32;

// ------------------------------------------------------------------------------------------------
// The periodisation in longitude-spectral space. The stretched apply FFTs the ring grid along
// longitude only to transform it back inside the 2-D FFT (src/Stretch.jl:270-273 after the
// longitude bFFT of :313-318): the pair cancels to a factor nph, and the DFS half-period shift
// k + nph/2 becomes a factor (-1)^m on order m. With spin weight s the DFS reflection carries another
// (-1)^s (src/DFS.jl:13-31). So the periodised grid is built directly per order, in column layout
// Ds [b][m + mmax][i], i < nth contiguous (the ring buffer's layout), and only the colatitude FFT
// remains.
// ------------------------------------------------------------------------------------------------

// Ds[b][c][i] = w[j] G[j, m] for i < nh (j = i), and (-1)^(m+s) w[j] G[j, m] for the mirrored rows
// i >= nh of a pole/full periodisation (j = nth - 1 - i); zero off the retained ranges.
template <class C>
__global__ void spec_dfs_kernel(const double2 *__restrict__ Gd, const int32_t *__restrict__ jlo,
                                const int32_t *__restrict__ jhi, const double *__restrict__ wth,
                                C *__restrict__ Ds, int nh, int nth, int mmax, int spin) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    const int b = blockIdx.z;
    if (i >= nth) return;
    const int64_t ncol = 2 * mmax + 1;
    const int m = c - mmax, am = m < 0 ? -m : m;
    const int j = i < nh ? i : nth - 1 - i;
    double2 v = make_double2(0.0, 0.0);
    if (j >= jlo[am] && j <= jhi[am]) {
        const double2 g = Gd[(int64_t(b) * ncol + c) * nh + j];
        double w = wth[j];
        if (i >= nh && ((am + spin) & 1)) w = -w;
        v = make_double2(g.x * w, g.y * w);
    }
    stc(Ds + (int64_t(b) * ncol + c) * nth + i, v);
}

// The transpose: Gd[b][c][j] = w[j] (Ds[c][j] + (-1)^(m+s) Ds[c][nth - 1 - j]) (pole/full) or
// w[j] Ds[c][j] (interior, nth = nh), for every ring j < nh.
template <class C>
__global__ void spec_fold_kernel(const C *__restrict__ Ds, const double *__restrict__ wth,
                                 double2 *__restrict__ Gd, int nh, int nth, int mmax, int spin) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    const int b = blockIdx.z;
    if (j >= nh) return;
    const int64_t ncol = 2 * mmax + 1;
    const int m = c - mmax, am = m < 0 ? -m : m;
    const C *col = Ds + (int64_t(b) * ncol + c) * nth;
    double2 v = ldc(col + j);
    if (nth > nh) {
        const double2 r = ldc(col + nth - 1 - j);
        if ((am + spin) & 1) v = make_double2(v.x - r.x, v.y - r.y);
        else v = make_double2(v.x + r.x, v.y + r.y);
    }
    const double w = wth[j];
    Gd[(int64_t(b) * ncol + c) * nh + j] = make_double2(v.x * w, v.y * w);
}

template <class C>
void launch_spec_dfs(const GridDims &g, const RingDev &r, const double2 *Gd, const double *wth, C *Ds, hipStream_t s) {
    dim3 grid((g.nth + 255) / 256, 2 * g.mmax + 1, g.B);
    spec_dfs_kernel<C><<<grid, 256, 0, s>>>(Gd, r.jlo, r.jhi, wth, Ds, g.nh, g.nth, g.mmax, g.spin);
}

template <class C>
void launch_spec_fold(const GridDims &g, const C *Ds, const double *wth, double2 *Gd, hipStream_t s) {
    dim3 grid((g.nh + 255) / 256, 2 * g.mmax + 1, g.B);
    spec_fold_kernel<C><<<grid, 256, 0, s>>>(Ds, wth, Gd, g.nh, g.nth, g.mmax, g.spin);
}

template void launch_spec_dfs<double2>(const GridDims &, const RingDev &, const double2 *, const double *, double2 *,
                                       hipStream_t);
template void launch_spec_dfs<float2>(const GridDims &, const RingDev &, const double2 *, const double *, float2 *,
                                      hipStream_t);
template void launch_spec_fold<double2>(const GridDims &, const double2 *, const double *, double2 *, hipStream_t);
template void launch_spec_fold<float2>(const GridDims &, const float2 *, const double *, double2 *, hipStream_t);

// Batched complex transpose through an LDS tile (layout conversion at the C-ABI boundary).
__global__ void __launch_bounds__(256)
transpose_kernel(const double2 *__restrict__ in, double2 *__restrict__ out, int rows, int cols) {
    __shared__ double2 tile[kTile][kTile + 1];
    const int c0 = blockIdx.x * kTile, r0 = blockIdx.y * kTile, b = blockIdx.z;
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
    const int64_t base = int64_t(b) * rows * cols;
    for (int rr = ty; rr < kTile; rr += 8) {
        const int r = r0 + rr, c = c0 + tx;
        if (r < rows && c < cols) tile[rr][tx] = in[base + int64_t(r) * cols + c];
    }
    __syncthreads();
    for (int cc = ty; cc < kTile; cc += 8) {
        const int c = c0 + cc, r = r0 + tx;
        if (c < cols && r < rows) out[base + int64_t(c) * rows + r] = tile[tx][cc];
    }
}

void launch_transpose(const double2 *in, double2 *out, int rows, int cols, int B, hipStream_t s) {
    dim3 grid((cols + kTile - 1) / kTile, (rows + kTile - 1) / kTile, B);
    transpose_kernel<<<grid, 256, 0, s>>>(in, out, rows, cols);
}

}  // namespace pfsht
