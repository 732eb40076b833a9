// analysis_kernels.hip — vector kernels of the iterative callers on a plan (src/Analysis.jl): point
// weights, complex dot products with a fixed reduction order (bit-identical run to run), the
// conjugate-gradient updates, the Wiener prior term and the pseudo-spectrum. All HBM-bound
// streaming over the alm layout (L+1) x (2L+1) x B or the field M x B.
#include "pfsht_common.hpp"

namespace pfsht {

namespace {
constexpr int kRedThreads = 256;

// Sum over the block in a fixed tree order; the result is valid in thread 0.
__device__ __forceinline__ double2 block_sum(double2 v, double2 *sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = make_double2(sh[t].x + sh[t + s].x, sh[t].y + sh[t + s].y);
        __syncthreads();
    }
    return sh[0];
}
}  // namespace

// f[b][i] *= w[i] (_weight_field!, src/Analysis.jl:106-121)
__global__ void weight_field_kernel(double2 *__restrict__ f, const double *__restrict__ w, int64_t M) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t b = blockIdx.y;
    if (i >= M) return;
    const double s = w[i];
    double2 v = f[b * M + i];
    f[b * M + i] = make_double2(v.x * s, v.y * s);
}

// Partial sums of conj(a) b: block g sums elements g T + t + k G T over k in order, then its tree.
__global__ void __launch_bounds__(kRedThreads)
cdot_partial_kernel(const double2 *__restrict__ a, const double2 *__restrict__ b, int64_t n,
                    double2 *__restrict__ part) {
    __shared__ double2 sh[kRedThreads];
    const int64_t stride = int64_t(gridDim.x) * kRedThreads;
    double re = 0.0, im = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * kRedThreads + threadIdx.x; i < n; i += stride) {
        const double2 x = a[i], y = b[i];
        re = fma(x.x, y.x, fma(x.y, y.y, re));
        im = fma(x.x, y.y, fma(-x.y, y.x, im));
    }
    const double2 s = block_sum(make_double2(re, im), sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kRedThreads)
cdot_final_kernel(const double2 *__restrict__ part, int np, double2 *__restrict__ out) {
    __shared__ double2 sh[kRedThreads];
    double re = 0.0, im = 0.0;
    for (int i = threadIdx.x; i < np; i += kRedThreads) {
        re += part[i].x;
        im += part[i].y;
    }
    const double2 s = block_sum(make_double2(re, im), sh);
    if (threadIdx.x == 0) *out = s;
}

// y += a x (complex a)
__global__ void zaxpy_kernel(double2 *__restrict__ y, const double2 *__restrict__ x, double2 a, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 v = x[i], u = y[i];
    y[i] = make_double2(u.x + (a.x * v.x - a.y * v.y), u.y + (a.x * v.y + a.y * v.x));
}

// p = r + beta p (complex beta)
__global__ void zxpby_kernel(double2 *__restrict__ p, const double2 *__restrict__ r, double2 beta, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 v = p[i], u = r[i];
    p[i] = make_double2(u.x + (beta.x * v.x - beta.y * v.y), u.y + (beta.x * v.y + beta.y * v.x));
}

// Ap[l, m] += reg[l] p[l, m] over the modes l <= L (_add_prior!, src/Analysis.jl:449-458); one
// thread per slot of the sph_mode layout, row l - |m| of column 2|m| - (m < 0).
__global__ void add_prior_kernel(double2 *__restrict__ Ap, const double2 *__restrict__ p,
                                 const double *__restrict__ reg, int L) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    const int col = blockIdx.y;
    const int64_t b = blockIdx.z;
    if (row > L) return;
    const int am = (col + 1) / 2;
    const int l = row + am;
    if (l > L) return;
    const double s = reg[l];
    if (s == 0.0) return;
    const int64_t i = (b * (2 * L + 1) + col) * (L + 1) + row;
    const double2 v = p[i], u = Ap[i];
    Ap[i] = make_double2(u.x + s * v.x, u.y + s * v.y);
}

// C[l] = sum_{m=-l..l} |alm[l, m]|^2 / (2l + 1) for one transform (src/Analysis.jl:298-306); one
// workgroup per degree, m in a fixed order per thread and a fixed tree.
__global__ void __launch_bounds__(kRedThreads)
spectrum_kernel(const double2 *__restrict__ alm, int L, double *__restrict__ C) {
    __shared__ double2 sh[kRedThreads];
    const int l = blockIdx.x;
    const int64_t nrow = L + 1;
    double acc = 0.0;
    for (int j = threadIdx.x; j <= 2 * l; j += kRedThreads) {  // j = 0: m = 0; 2k - 1, 2k: m = +-k
        const int am = (j + 1) / 2;
        const int col = j == 0 ? 0 : (j & 1 ? 2 * am : 2 * am - 1);
        const double2 v = alm[int64_t(col) * nrow + (l - am)];
        acc = fma(v.x, v.x, fma(v.y, v.y, acc));
    }
    const double2 s = block_sum(make_double2(acc, 0.0), sh);
    if (threadIdx.x == 0) C[l] = s.x / double(2 * l + 1);
}

void launch_weight_field(double2 *f, const double *w, int64_t M, int B, hipStream_t s) {
    if (M <= 0) return;
    weight_field_kernel<<<dim3(unsigned((M + 255) / 256), B), 256, 0, s>>>(f, w, M);
}

void launch_cdot(const double2 *a, const double2 *b, int64_t n, double2 *part, double2 *out, hipStream_t s) {
    cdot_partial_kernel<<<kDotBlocks, kRedThreads, 0, s>>>(a, b, n, part);
    cdot_final_kernel<<<1, kRedThreads, 0, s>>>(part, kDotBlocks, out);
}

// ComplexF32 <-> ComplexF64 for the single-precision applies: four elements per thread.
__global__ void widen_c64_kernel(const float2 *__restrict__ in, double2 *__restrict__ out, int64_t n) {
    const int64_t i0 = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (i0 + k < n) {
            const float2 v = in[i0 + k];
            out[i0 + k] = make_double2(double(v.x), double(v.y));
        }
}

__global__ void narrow_c128_kernel(const double2 *__restrict__ in, float2 *__restrict__ out, int64_t n) {
    const int64_t i0 = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (i0 + k < n) {
            const double2 v = in[i0 + k];
            out[i0 + k] = make_float2(float(v.x), float(v.y));
        }
}

void launch_widen_c64(const float2 *in, double2 *out, int64_t n, hipStream_t s) {
    if (n > 0) widen_c64_kernel<<<unsigned((n + 1023) / 1024), 256, 0, s>>>(in, out, n);
}

void launch_narrow_c128(const double2 *in, float2 *out, int64_t n, hipStream_t s) {
    if (n > 0) narrow_c128_kernel<<<unsigned((n + 1023) / 1024), 256, 0, s>>>(in, out, n);
}

void launch_zaxpy(double2 *y, const double2 *x, double2 a, int64_t n, hipStream_t s) {
    if (n > 0) zaxpy_kernel<<<unsigned((n + 255) / 256), 256, 0, s>>>(y, x, a, n);
}

void launch_zxpby(double2 *p, const double2 *r, double2 beta, int64_t n, hipStream_t s) {
    if (n > 0) zxpby_kernel<<<unsigned((n + 255) / 256), 256, 0, s>>>(p, r, beta, n);
}

void launch_add_prior(double2 *Ap, const double2 *p, const double *reg, int L, int B, hipStream_t s) {
    add_prior_kernel<<<dim3((L + 256) / 256, 2 * L + 1, B), 256, 0, s>>>(Ap, p, reg, L);
}

void launch_spectrum(const double2 *alm, int L, double *C, hipStream_t s) {
    spectrum_kernel<<<L + 1, kRedThreads, 0, s>>>(alm, L, C);
}

}  // namespace pfsht

namespace pfsht {

// Multi-box applies (src/MultiBox.jl:83-144): f[b][idx[j]] = fg[b][j] and fg[b][j] = f[b][idx[j]]
// for the points of one box.
__global__ void scatter_field_kernel(OutCols f, const double2 *__restrict__ fg, const int32_t *__restrict__ idx,
                                     int64_t Mg) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (j < Mg) gstore(f.col(b) + idx[j], fg[int64_t(b) * Mg + j]);
}

__global__ void gather_field_kernel(double2 *__restrict__ fg, InCols f, const int32_t *__restrict__ idx, int64_t Mg) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (j < Mg) fg[int64_t(b) * Mg + j] = gload(f.col(b) + idx[j]);
}

__global__ void cols_add_kernel(OutCols a, const double2 *__restrict__ acc, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= n) return;
    const gptr p = a.col(b) + i;
    const double2 v = acc[int64_t(b) * n + i], u = gload(p);
    gstore(p, make_double2(u.x + v.x, u.y + v.y));
}

void launch_scatter_field(OutCols f, const double2 *fg, const int32_t *idx, int64_t Mg, int nb, hipStream_t s) {
    if (Mg > 0 && nb > 0) scatter_field_kernel<<<dim3(unsigned((Mg + 255) / 256), nb), 256, 0, s>>>(f, fg, idx, Mg);
}

void launch_gather_field(double2 *fg, InCols f, const int32_t *idx, int64_t Mg, int nb, hipStream_t s) {
    if (Mg > 0 && nb > 0) gather_field_kernel<<<dim3(unsigned((Mg + 255) / 256), nb), 256, 0, s>>>(fg, f, idx, Mg);
}

void launch_cols_add(OutCols a, const double2 *acc, int64_t n, int nb, hipStream_t s) {
    if (n > 0 && nb > 0) cols_add_kernel<<<dim3(unsigned((n + 255) / 256), nb), 256, 0, s>>>(a, acc, n);
}

}  // namespace pfsht

namespace pfsht {

// Multi-box applies (src/MultiBox.jl:83-144): f[b][idx[j]] = fg[b][j] and fg[b][j] = f[b][idx[j]]
// for the points of one box.
__global__ void scatter_field_kernel(double2 *__restrict__ f, const double2 *__restrict__ fg,
                                     const int32_t *__restrict__ idx, int64_t Mg, int64_t M) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t b = blockIdx.y;
    if (j < Mg) f[b * M + idx[j]] = fg[b * Mg + j];
}

__global__ void gather_field_kernel(double2 *__restrict__ fg, const double2 *__restrict__ f,
                                    const int32_t *__restrict__ idx, int64_t Mg, int64_t M) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t b = blockIdx.y;
    if (j < Mg) fg[b * Mg + j] = f[b * M + idx[j]];
}

void launch_scatter_field(double2 *f, const double2 *fg, const int32_t *idx, int64_t Mg, int64_t M, int B,
                          hipStream_t s) {
    if (Mg > 0) scatter_field_kernel<<<dim3(unsigned((Mg + 255) / 256), B), 256, 0, s>>>(f, fg, idx, Mg, M);
}

void launch_gather_field(double2 *fg, const double2 *f, const int32_t *idx, int64_t Mg, int64_t M, int B,
                         hipStream_t s) {
    if (Mg > 0) gather_field_kernel<<<dim3(unsigned((Mg + 255) / 256), B), 256, 0, s>>>(fg, f, idx, Mg, M);
}

}  // namespace pfsht
