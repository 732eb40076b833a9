// ring_kernels.hip — the block-scaled associated-Legendre ring pass on gfx950.
//
// Forward (rings_from_alm!, src/Legendre.jl:310-323,367-472) and adjoint (alm_from_rings!,
// src/Legendre.jl:332-341). Work is split into items (order m, tile of <= 64 retained rings), one
// wavefront each, issued longest first. Lanes own rings: the two-term recurrence state, the
// ring's exponent and (forward) its accumulators stay in registers for the whole degree sweep.
// Everything wave-uniform — the recurrence coefficients (a_l^m, b_l^m) and the a_lm column of the
// order being swept — is addressed uniformly so it is fetched with scalar loads into SGPRs and fed
// to FP64 FMAs as scalar operands, which costs no vector bandwidth.
//
// Block scaling: each lane carries its own power-of-two exponent E (the device kernel of
// ext/PartialFSHTKernelAbstractionsExt.jl:142-214 does the same); every 64 degrees the pair
// (q_{l-1}, q_{l-2}) is renormalised to a leading power of two. Powers of two are exact, so the
// mantissa arithmetic is the reference's apart from the shared-versus-per-ring exponent choice.
//
// The Condon–Shortley sign (-1)^m is folded into the stores: with q from the unsigned seed,
// Pbar_l^m = (-1)^m q 2^E, so G[j, m] = (-1)^m sum a_lm q 2^E and G[j, -m] = sum a_l,-m q 2^E.
//
// Adjoint: alm[l, m] = sum_j Pbar_l^m(x_j) G[j, m] is a reduction across lanes. Each wavefront
// computes a 64-degree tile of w = q 2^E with lanes over rings, transposes it through LDS, and
// then lets lanes run over degrees, summing the tile's rings in ring order with the ring data as
// scalar operands. Per-tile partials go to a plan-owned buffer and a finalize kernel adds the
// tiles of each order in tile order: no float atomics, bit-identical run to run.
#include "pfsht_common.hpp"

namespace pfsht {

namespace {

__device__ __forceinline__ int renorm_shift(double q1, double q2) {
    // s = ceil(log2(max(|q1|,|q2|))) (src/Legendre.jl:521-542), by frexp: exact and branch-free.
    const double mx = fmax(fabs(q1), fabs(q2));
    if (!(mx > 0.0) || !(mx < INFINITY)) return 0;
    int e;
    const double f = frexp(mx, &e);
    return f == 0.5 ? e - 1 : e;
}

__device__ __forceinline__ void seed(double e0, double &q, int &E) {
    // Pbar_m^m = (-1)^m 2^{e0}, carried as mantissa 2^{e0 - ceil(e0)} and exponent ceil(e0)
    // (src/Legendre.jl:379-382, 475-481).
    if (e0 > -INFINITY && e0 < INFINITY) {
        const double Ed = ceil(e0);
        q = exp2(e0 - Ed);
        E = int(Ed);
    } else {
        q = 0.0;
        E = 0;
    }
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// Forward
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(64)
ring_fwd_kernel(const RingItem *__restrict__ items, const double *__restrict__ costh,
                const double *__restrict__ log2sin, const double *__restrict__ log2c,
                const double2 *__restrict__ ab, const int64_t *__restrict__ abofs,
                const double2 *__restrict__ alm, double2 *__restrict__ Gd, int L, int n, int mmax) {
    const RingItem it = items[blockIdx.x];
    const int m = __builtin_amdgcn_readfirstlane(it.m);
    const int lane = threadIdx.x;
    const bool active = lane < it.nr;
    const int j = it.j0 + (active ? lane : 0);
    const int b = blockIdx.y;

    const double x = costh[j];
    const double e0 = m == 0 ? log2c[0] : double(m) * log2sin[j] + log2c[m];
    double q2;
    int E;
    seed(e0, q2, E);
    if (!active) q2 = 0.0;

    const int64_t LB = int64_t(L + 1) * (2 * L + 1);
    // column of order m (row l - m holds degree l) and of order -m
    const double2 *__restrict__ cp = alm + b * LB + int64_t(2 * m) * (L + 1) - m;
    const double2 *__restrict__ cm = alm + b * LB + int64_t(2 * m - 1) * (L + 1) - m;

    double ar = 0, ai = 0, br = 0, bi = 0;  // accumulators for +m and -m
    {
        const double2 a0 = cp[m];
        ar = a0.x * q2;
        ai = a0.y * q2;
        if (m > 0) {
            const double2 c0 = cm[m];
            br = c0.x * q2;
            bi = c0.y * q2;
        }
    }
    double q1 = 0.0;
    if (m < L) {
        q1 = sqrt(double(2 * m + 3)) * x * q2;
        const double2 a1 = cp[m + 1];
        ar = fma(a1.x, q1, ar);
        ai = fma(a1.y, q1, ai);
        if (m > 0) {
            const double2 c1 = cm[m + 1];
            br = fma(c1.x, q1, br);
            bi = fma(c1.y, q1, bi);
        }
    }
    const double2 *__restrict__ abm = ab + abofs[m] - (m + 2);  // abm[l] for l = m+2..L

    int l = m + 2;
    if (m > 0) {
        while (l <= L) {
            const int lend = min(L, l + kRescaleStride - 1);
#pragma unroll 4
            for (; l <= lend; ++l) {
                const double2 c = abm[l];
                const double q = fma(c.x * x, q1, -(c.y * q2));
                q2 = q1;
                q1 = q;
                const double2 a = cp[l];
                const double2 d = cm[l];
                ar = fma(a.x, q, ar);
                ai = fma(a.y, q, ai);
                br = fma(d.x, q, br);
                bi = fma(d.y, q, bi);
            }
            if (l <= L) {
                const int s = renorm_shift(q1, q2);
                q1 = ldexp(q1, -s);
                q2 = ldexp(q2, -s);
                ar = ldexp(ar, -s);
                ai = ldexp(ai, -s);
                br = ldexp(br, -s);
                bi = ldexp(bi, -s);
                E += s;
            }
        }
    } else {
        while (l <= L) {
            const int lend = min(L, l + kRescaleStride - 1);
#pragma unroll 4
            for (; l <= lend; ++l) {
                const double2 c = abm[l];
                const double q = fma(c.x * x, q1, -(c.y * q2));
                q2 = q1;
                q1 = q;
                const double2 a = cp[l];
                ar = fma(a.x, q, ar);
                ai = fma(a.y, q, ai);
            }
            if (l <= L) {
                const int s = renorm_shift(q1, q2);
                q1 = ldexp(q1, -s);
                q2 = ldexp(q2, -s);
                ar = ldexp(ar, -s);
                ai = ldexp(ai, -s);
                E += s;
            }
        }
    }
    if (!active) return;
    const double sg = (m & 1) ? -1.0 : 1.0;
    const int64_t ncol = 2 * mmax + 1;
    double2 *gp = Gd + (int64_t(b) * ncol + (mmax + m)) * n + j;
    *gp = make_double2(sg * ldexp(ar, E), sg * ldexp(ai, E));
    if (m > 0) {
        double2 *gm = Gd + (int64_t(b) * ncol + (mmax - m)) * n + j;
        *gm = make_double2(ldexp(br, E), ldexp(bi, E));
    }
}

void launch_ring_fwd(const RingDev &r, const double2 *alm, double2 *Gd, hipStream_t s) {
    if (r.nitems == 0) return;
    dim3 grid(r.nitems, r.B);
    ring_fwd_kernel<<<grid, 64, 0, s>>>(r.items, r.costh, r.log2sin, r.log2c, r.ab, r.abofs, alm, Gd, r.L, r.n,
                                        r.mmax);
}

// ------------------------------------------------------------------------------------------------
// Adjoint
// ------------------------------------------------------------------------------------------------
constexpr int kTL = 64;  // degrees per LDS tile

__global__ void __launch_bounds__(64)
ring_adj_kernel(const RingItem *__restrict__ items, const double *__restrict__ costh,
                const double *__restrict__ log2sin, const double *__restrict__ log2c,
                const double2 *__restrict__ ab, const int64_t *__restrict__ abofs,
                const int64_t *__restrict__ pofs, const double2 *__restrict__ Gd,
                double4 *__restrict__ partials, int L, int n, int mmax, int B) {
    __shared__ double W[kTL][kTL + 1];
    const RingItem it = items[blockIdx.x];
    const int m = __builtin_amdgcn_readfirstlane(it.m);
    const int nr = __builtin_amdgcn_readfirstlane(it.nr);
    const int tile = __builtin_amdgcn_readfirstlane(it.tile);
    const int lane = threadIdx.x;
    const bool active = lane < nr;
    const int j = it.j0 + (active ? lane : 0);
    const int b = blockIdx.y;

    const double x = costh[j];
    const double e0 = m == 0 ? log2c[0] : double(m) * log2sin[j] + log2c[m];
    double qs;
    int E;
    seed(e0, qs, E);
    if (!active) qs = 0.0;

    const int64_t ncol = 2 * mmax + 1;
    const double2 *__restrict__ gpc = Gd + (int64_t(b) * ncol + (mmax + m)) * n + it.j0;
    const double2 *__restrict__ gmc = Gd + (int64_t(b) * ncol + (mmax - m)) * n + it.j0;
    const double2 *__restrict__ abm = ab + abofs[m] - (m + 2);
    const int nl = L - m + 1;
    double4 *__restrict__ out = partials + pofs[m] + (int64_t(tile) * nl) * B + b;

    double q1 = 0.0, q2 = 0.0;  // q_{l-1}, q_{l-2}
    for (int l0 = m; l0 <= L; l0 += kTL) {
        const int tl = min(kTL, L - l0 + 1);
        // phase 1: lanes = rings, t = degree offset
        int t = 0;
        if (l0 == m) {
            q1 = qs;
            W[0][lane] = ldexp(qs, E);
            t = 1;
            if (m < L) {
                const double q = sqrt(double(2 * m + 3)) * x * qs;
                q2 = q1;
                q1 = q;
                W[1][lane] = ldexp(q, E);
                t = 2;
            }
        }
#pragma unroll 4
        for (; t < tl; ++t) {
            const int l = l0 + t;
            const double2 c = abm[l];
            const double q = fma(c.x * x, q1, -(c.y * q2));
            q2 = q1;
            q1 = q;
            W[t][lane] = ldexp(q, E);
        }
        {
            const int s = renorm_shift(q1, q2);
            q1 = ldexp(q1, -s);
            q2 = ldexp(q2, -s);
            E += s;
        }
        __syncthreads();
        // phase 2: lanes = degree offsets, sum over the tile's rings in ring order
        if (lane < tl) {
            double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
            if (m > 0) {
#pragma unroll 4
                for (int r = 0; r < nr; ++r) {
                    const double wv = W[lane][r];
                    const double2 gp = gpc[r];
                    const double2 gm = gmc[r];
                    s0 = fma(wv, gp.x, s0);
                    s1 = fma(wv, gp.y, s1);
                    s2 = fma(wv, gm.x, s2);
                    s3 = fma(wv, gm.y, s3);
                }
            } else {
#pragma unroll 4
                for (int r = 0; r < nr; ++r) {
                    const double wv = W[lane][r];
                    const double2 gp = gpc[r];
                    s0 = fma(wv, gp.x, s0);
                    s1 = fma(wv, gp.y, s1);
                }
            }
            out[int64_t(l0 - m + lane) * B] = make_double4(s0, s1, s2, s3);
        }
        __syncthreads();
    }
}

// Writes every alm slot: the tile-ordered sum of the partials where (l, m) is retained, zero in
// the padding and for unretained orders (the adjoint overwrites alm, src/Legendre.jl:338).
__global__ void ring_adj_finalize(const int64_t *__restrict__ pofs, const int32_t *__restrict__ ntiles,
                                  const double4 *__restrict__ partials, double2 *__restrict__ alm, int L,
                                  int mmax, int B) {
    const int64_t nrow = L + 1, ncol = 2 * L + 1;
    const int64_t total = nrow * ncol * B;
    for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < total;
         idx += int64_t(gridDim.x) * blockDim.x) {
        const int row = int(idx % nrow);
        const int64_t rest = idx / nrow;
        const int col = int(rest % ncol);
        const int b = int(rest / ncol);
        const int am = (col + 1) / 2;
        const bool neg = (col & 1) != 0;
        const int l = row + am;
        double2 v = make_double2(0.0, 0.0);
        if (l <= L && am <= mmax) {
            const int nt = ntiles[am];
            const int nl = L - am + 1;
            const double4 *p = partials + pofs[am] + int64_t(l - am) * B + b;
            double sr = 0, si = 0;
            for (int t = 0; t < nt; ++t) {
                const double4 q = p[int64_t(t) * nl * B];
                sr += neg ? q.z : q.x;
                si += neg ? q.w : q.y;
            }
            if (!neg && (am & 1)) { sr = -sr; si = -si; }
            v = make_double2(sr, si);
        }
        alm[idx] = v;
    }
}

void launch_ring_adj(const RingDev &r, const double2 *Gd, double2 *alm, hipStream_t s) {
    if (r.nitems > 0) {
        dim3 grid(r.nitems, r.B);
        ring_adj_kernel<<<grid, 64, 0, s>>>(r.items, r.costh, r.log2sin, r.log2c, r.ab, r.abofs, r.pofs, Gd,
                                            r.partials, r.L, r.n, r.mmax, r.B);
    }
    const int64_t total = int64_t(r.L + 1) * (2 * r.L + 1) * r.B;
    const int threads = 256;
    const int64_t blocks = std::min<int64_t>((total + threads - 1) / threads, 65536);
    ring_adj_finalize<<<dim3(unsigned(blocks)), threads, 0, s>>>(r.pofs, r.ntiles, r.partials, alm, r.L, r.mmax,
                                                                  r.B);
}

// ------------------------------------------------------------------------------------------------
// Layout conversion between the internal ring buffer [b][m + mmax][j] (retained entries only) and
// the reference's G (n x (2L+1) x B, column m + L, zero off the retained set).
// ------------------------------------------------------------------------------------------------
__global__ void gd_to_ref_kernel(const double2 *__restrict__ Gd, const int32_t *__restrict__ jlo,
                                 const int32_t *__restrict__ jhi, double2 *__restrict__ G, int n, int L, int mmax,
                                 int B) {
    const int64_t total = int64_t(n) * (2 * L + 1) * B;
    for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < total;
         idx += int64_t(gridDim.x) * blockDim.x) {
        const int j = int(idx % n);
        const int64_t rest = idx / n;
        const int c = int(rest % (2 * L + 1));
        const int b = int(rest / (2 * L + 1));
        const int m = c - L;
        const int am = m < 0 ? -m : m;
        double2 v = make_double2(0.0, 0.0);
        if (am <= mmax && j >= jlo[am] && j <= jhi[am])
            v = Gd[(int64_t(b) * (2 * mmax + 1) + (m + mmax)) * n + j];
        G[idx] = v;
    }
}

__global__ void ref_to_gd_kernel(const double2 *__restrict__ G, double2 *__restrict__ Gd, int n, int L, int mmax,
                                 int B) {
    const int64_t total = int64_t(n) * (2 * mmax + 1) * B;
    for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < total;
         idx += int64_t(gridDim.x) * blockDim.x) {
        const int j = int(idx % n);
        const int64_t rest = idx / n;
        const int c = int(rest % (2 * mmax + 1));
        const int b = int(rest / (2 * mmax + 1));
        const int m = c - mmax;
        Gd[idx] = G[(int64_t(b) * (2 * L + 1) + (m + L)) * n + j];
    }
}

void launch_gd_to_ref(const RingDev &r, const double2 *Gd, double2 *G, hipStream_t s) {
    const int64_t total = int64_t(r.n) * (2 * r.L + 1) * r.B;
    const int64_t blocks = std::min<int64_t>((total + 255) / 256, 65536);
    gd_to_ref_kernel<<<dim3(unsigned(blocks)), 256, 0, s>>>(Gd, r.jlo, r.jhi, G, r.n, r.L, r.mmax, r.B);
}

void launch_ref_to_gd(const RingDev &r, const double2 *G, double2 *Gd, hipStream_t s) {
    const int64_t total = int64_t(r.n) * (2 * r.mmax + 1) * r.B;
    const int64_t blocks = std::min<int64_t>((total + 255) / 256, 65536);
    ref_to_gd_kernel<<<dim3(unsigned(blocks)), 256, 0, s>>>(G, Gd, r.n, r.L, r.mmax, r.B);
}

}  // namespace pfsht
