// grid_kernels.hip — longitude pack/unpack with the apodisation window, and the Double Fourier
// Sphere periodisation. All HBM-bound data movement; layouts are row-major with longitude (k)
// contiguous:
//   Gd [b][m + mmax][j]   ring-pass buffer (rings contiguous, written by the ring kernels)
//   Fp [b][j][k]          ring x longitude grid, k in [0, nph)
//   Ft [b][i][k]          doubled grid, i in [0, nth)
#include "pfsht_common.hpp"

namespace pfsht {

constexpr int kTile = 32;

// Order carried by longitude column k of the packed grid (src/Stretch.jl:332-350): k in
// [0, mkeep] holds m = k, k in [nph - mkeep, nph) holds m = k - nph; with nph >= 2 mmax + 1 every
// retained order lands in one of the two blocks.
__device__ __forceinline__ int order_of_column(int k, int nph, int mmax) {
    if (k <= mmax) return k;
    if (k >= nph - mmax) return k - nph;
    return INT32_MIN;
}

// Fp[b][j][k] = w[j] * G[j, m(k)] on the retained set, zero elsewhere. The window multiplies each
// ring before the longitude FFT; it is a per-row scale, so this equals F .*= w after the FFT
// (src/Stretch.jl:269-270).
__global__ void __launch_bounds__(256)
pack_window_kernel(const double2 *__restrict__ Gd, const int32_t *__restrict__ jlo, const int32_t *__restrict__ jhi,
                   const double *__restrict__ wth, double2 *__restrict__ Fp, int nh, int nph, int mmax) {
    __shared__ double2 tile[kTile][kTile + 1];
    const int k0 = blockIdx.x * kTile, j0 = blockIdx.y * kTile, b = blockIdx.z;
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;  // 32 x 8
    const int64_t ncol = 2 * mmax + 1;
    // load: rows = k (orders), columns = j (rings, contiguous in Gd)
    for (int kk = ty; kk < kTile; kk += 8) {
        const int k = k0 + kk, j = j0 + tx;
        double2 v = make_double2(0.0, 0.0);
        if (k < nph && j < nh) {
            const int m = order_of_column(k, nph, mmax);
            if (m != INT32_MIN) {
                const int am = m < 0 ? -m : m;
                if (j >= jlo[am] && j <= jhi[am]) {
                    const double2 g = Gd[(int64_t(b) * ncol + (m + mmax)) * nh + j];
                    const double wv = wth[j];
                    v = make_double2(g.x * wv, g.y * wv);
                }
            }
        }
        tile[kk][tx] = v;
    }
    __syncthreads();
    for (int jj = ty; jj < kTile; jj += 8) {
        const int j = j0 + jj, k = k0 + tx;
        if (j < nh && k < nph) Fp[(int64_t(b) * nh + j) * nph + k] = tile[tx][jj];
    }
}

void launch_pack_window(const GridDims &g, const RingDev &r, const double2 *Gd, const double *wth, double2 *Fp,
                        hipStream_t s) {
    dim3 grid((g.nph + kTile - 1) / kTile, (g.nh + kTile - 1) / kTile, g.B);
    pack_window_kernel<<<grid, 256, 0, s>>>(Gd, r.jlo, r.jhi, wth, Fp, g.nh, g.nph, g.mmax);
}

// Gd[b][m + mmax][j] = w[j] * Fp[b][j][k(m)] for |m| <= mmax: the adjoint of the pack (src/Stretch.jl:
// 352-364 restricted to the orders the ring adjoint reads).
__global__ void __launch_bounds__(256)
unpack_window_kernel(const double2 *__restrict__ Fp, const double *__restrict__ wth, double2 *__restrict__ Gd,
                     int nh, int nph, int mmax) {
    __shared__ double2 tile[kTile][kTile + 1];
    const int c0 = blockIdx.x * kTile, j0 = blockIdx.y * kTile, b = blockIdx.z;
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
    const int ncol = 2 * mmax + 1;
    for (int jj = ty; jj < kTile; jj += 8) {
        const int j = j0 + jj, c = c0 + tx;
        double2 v = make_double2(0.0, 0.0);
        if (j < nh && c < ncol) {
            const int m = c - mmax;
            const int k = m >= 0 ? m : m + nph;
            const double2 f = Fp[(int64_t(b) * nh + j) * nph + k];
            const double wv = wth[j];
            v = make_double2(f.x * wv, f.y * wv);
        }
        tile[jj][tx] = v;
    }
    __syncthreads();
    for (int cc = ty; cc < kTile; cc += 8) {
        const int c = c0 + cc, j = j0 + tx;
        if (c < ncol && j < nh) Gd[(int64_t(b) * ncol + c) * nh + j] = tile[tx][cc];
    }
}

void launch_unpack_window(const GridDims &g, const double2 *Fp, const double *wth, double2 *Gd, hipStream_t s) {
    dim3 grid((2 * g.mmax + 1 + kTile - 1) / kTile, (g.nh + kTile - 1) / kTile, g.B);
    unpack_window_kernel<<<grid, 256, 0, s>>>(Fp, wth, Gd, g.nh, g.nph, g.mmax);
}

// DFS doubling (src/DFS.jl:13-31): rows i < nh copy; row nh + i = F[nh - 1 - i, k + nph/2].
__global__ void dfs_double_kernel(const double2 *__restrict__ Fp, double2 *__restrict__ Ft, int nh, int nph) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    const int b = blockIdx.z;
    if (k >= nph) return;
    const int nth = 2 * nh;
    double2 v;
    if (i < nh) {
        v = Fp[(int64_t(b) * nh + i) * nph + k];
    } else {
        const int src = nth - 1 - i;
        int ks = k + nph / 2;
        if (ks >= nph) ks -= nph;
        v = Fp[(int64_t(b) * nh + src) * nph + ks];
    }
    Ft[(int64_t(b) * nth + i) * nph + k] = v;
}

// DFS fold (src/DFS.jl:41-56): F[i, k] = Ft[i, k] + Ft[nth - 1 - i, k - nph/2].
__global__ void dfs_fold_kernel(const double2 *__restrict__ Ft, double2 *__restrict__ Fp, int nh, int nph) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    const int b = blockIdx.z;
    if (k >= nph) return;
    const int nth = 2 * nh;
    int ks = k - nph / 2;
    if (ks < 0) ks += nph;
    const double2 a = Ft[(int64_t(b) * nth + i) * nph + k];
    const double2 c = Ft[(int64_t(b) * nth + (nth - 1 - i)) * nph + ks];
    Fp[(int64_t(b) * nh + i) * nph + k] = make_double2(a.x + c.x, a.y + c.y);
}

void launch_dfs_double(const GridDims &g, const double2 *Fp, double2 *Ft, hipStream_t s) {
    dim3 grid((g.nph + 255) / 256, g.nth, g.B);
    dfs_double_kernel<<<grid, 256, 0, s>>>(Fp, Ft, g.nh, g.nph);
}

void launch_dfs_fold(const GridDims &g, const double2 *Ft, double2 *Fp, hipStream_t s) {
    dim3 grid((g.nph + 255) / 256, g.nh, g.B);
    dfs_fold_kernel<<<grid, 256, 0, s>>>(Ft, Fp, g.nh, g.nph);
}

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
