// plan_kernels.hip — plan construction on the device: the degree-recurrence coefficient tables of
// the ring engine (src/Legendre.jl:406-408), the points' map onto the fine grid and their bin keys
// (src/Stretch.jl:202-213, src/NUFFT.jl bin sort), a stable radix sort of the keys, and the gather
// of the mapped coordinates into bin order. Every value is computed with the same operations in
// the same order as the host restatement (no FMA contraction), so a plan built here holds the bits
// a host build would.
#include "pfsht_common.hpp"

#include <hipcub/hipcub.hpp>

namespace pfsht {

namespace {
#pragma clang fp contract(off)
__device__ __forceinline__ double coef_a(int l, int m) {
    return sqrt(double(2 * l - 1) * double(2 * l + 1) / (double(l - m) * double(l + m)));
}
__device__ __forceinline__ double coef_b(int l, int m) {
    return sqrt(double(2 * l + 1) / double(2 * l - 3) * double(l - m - 1) / double(l - m) * double(l + m - 1) /
                double(l + m));
}
}  // namespace

// One thread per (order m, 16-degree group of the rescaled recurrence): group 0 holds degrees
// m + 2 .. m + 15, group g >= 1 degrees m + 16 g .. m + 16 g + 15. ab[l] = (a_l, b_l); gc[l] =
// (gamma_l, C_l) with gamma_l = b_l / (a_l a_{l-1}), a_{m+1} = sqrt(2m + 3), and C_l the product of
// a_k over the group's degrees up to l.
__global__ void coef_tables_kernel(int L, const int64_t *__restrict__ abofs, double2 *__restrict__ ab,
                                   double2 *__restrict__ gc) {
#pragma clang fp contract(off)
    const int m = blockIdx.y;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int l0 = g == 0 ? m + 2 : m + 16 * g;
    if (l0 > L) return;
    const int l1 = min(L, m + 16 * g + 15);
    double aprev = g == 0 ? sqrt(double(2 * m + 3)) : coef_a(l0 - 1, m);
    double C = 1.0;
    const int64_t base = abofs[m] - (m + 2);
    for (int l = l0; l <= l1; ++l) {
        const double a = coef_a(l, m), b = coef_b(l, m);
        C *= a;
        ab[base + l] = make_double2(a, b);
        gc[base + l] = make_double2(b / (a * aprev), C);
        aprev = a;
    }
}

void launch_coef_tables(int L, int mmax, const int64_t *abofs, double2 *ab, double2 *gc, hipStream_t s) {
    if (L < 2) return;
    const int ngroups = L / 16 + 1;  // l0 <= L needs 16 g <= L - m
    coef_tables_kernel<<<dim3((ngroups + 63) / 64, mmax + 1), 64, 0, s>>>(L, abofs, ab, gc);
}

// One thread per point: map_point (point_bins.hpp), the key's count, and the fine rows the point's
// kernel reaches. err[0] = 1 when a point is refused (NaN, infinite or off the grid).
__global__ void map_points_kernel(const double *__restrict__ theta, const double *__restrict__ phi, int64_t M,
                                  PointMap q, double *__restrict__ X1, double *__restrict__ X2,
                                  int32_t *__restrict__ key, int32_t *__restrict__ idx, int32_t *__restrict__ cnt,
                                  uint8_t *__restrict__ on, int32_t *__restrict__ err) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= M) return;
    double x1, x2;
    int k, b1;
    idx[i] = int32_t(i);
    const bool ok = map_point(q, theta[i], phi[i], x1, x2, k, b1);
    X1[i] = x1;
    X2[i] = x2;
    key[i] = k;
    if (!ok) {  // NaN, infinite or off the grid: the plan is refused
        err[0] = 1;
        return;
    }
    atomicAdd(&cnt[k], 1);
    for (int r = 0; r < q.w; ++r) on[(b1 + r) % q.N1] = 1;
}

__global__ void gather_points_kernel(const double *__restrict__ X1, const double *__restrict__ X2,
                                     const int32_t *__restrict__ perm, int64_t M, double *__restrict__ X1s,
                                     double *__restrict__ X2s) {
    const int64_t qi = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (qi >= M) return;
    const int32_t p = perm[qi];
    X1s[qi] = X1[p];
    X2s[qi] = X2[p];
}

void launch_map_points(const double *theta, const double *phi, int64_t M, const PointMap &q, double *X1, double *X2,
                       int32_t *key, int32_t *idx, int32_t *cnt, uint8_t *on, int32_t *err, hipStream_t s) {
    if (M > 0) map_points_kernel<<<unsigned((M + 255) / 256), 256, 0, s>>>(theta, phi, M, q, X1, X2, key, idx, cnt,
                                                                            on, err);
}

void launch_gather_points(const double *X1, const double *X2, const int32_t *perm, int64_t M, double *X1s,
                          double *X2s, hipStream_t s) {
    if (M > 0) gather_points_kernel<<<unsigned((M + 255) / 256), 256, 0, s>>>(X1, X2, perm, M, X1s, X2s);
}

// Stable LSD radix sort of (key, value) pairs over the low `bits` bits of the keys: points of one
// bin keep their input order, as in a counting sort.
hipError_t sort_pairs(void *tmp, size_t &tmp_bytes, const int32_t *kin, int32_t *kout, const int32_t *vin,
                      int32_t *vout, int n, int bits, hipStream_t s) {
    return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kin, kout, vin, vout, n, 0, bits, s);
}

}  // namespace pfsht
