// spin_kernels.hip — the spin-weighted ring engine (src/Spin.jl:170-545): G[j, m] = sum_l a_lm N_l
// d^l_{m,-s}(theta_j) and its exact transpose, with the three-term recurrence in degree at fixed
// (m, -s) seeded at l0 = max(|m|, |s|) from the extremal closed form (tabulated per order and
// ring in log2 form). Each lane is one ring and carries its own power-of-two exponent,
// renormalised every 64 degrees; every order m in [-mmax, mmax] is walked on its own (d^l_{m,-s} and
// d^l_{-m,-s} differ). The work per retained (l, m, ring) triple is one recurrence step (three
// FP64 FMAs) and one complex product per transform, so the passes are FP64-issue bound like the
// scalar ones.
#include "pfsht_common.hpp"

#include <cfloat>

namespace pfsht {

namespace {
constexpr int kSpinWaves = 8;  // wavefronts per adjoint workgroup: up to 8 tiles of one order at once

// Seed of ring j for order m: log2 |d^{l0}_{m,-s}(theta_j)| = lb2 + pc log2 cos(theta/2) +
// ps log2 sin(theta/2) (with the p == 0 terms left out, which makes the poles exact), split
// into a value in (1/2, 1] times 2^E.
__device__ __forceinline__ void spin_seed(const double4 sd, double lc, double ls, double &d1, int &E) {
    double e = sd.x;
    if (sd.z != 0.0) e += sd.z * lc;
    if (sd.w != 0.0) e += sd.w * ls;
    if (e >= -DBL_MAX && e <= DBL_MAX) {
        const double Ec = ceil(e);
        E = int(Ec);
        d1 = sd.y * exp2(e - Ec);
    } else {
        E = 0;
        d1 = 0.0;
    }
}

// Rescale (d1, d0) into [1/2, 1) by a power of two, folding it into E; returns the shift.
__device__ __forceinline__ int spin_renorm(double &d1, double &d0, int &E) {
    const double mx = fmax(fabs(d1), fabs(d0));
    if (!(mx > 0.0) || !(mx <= DBL_MAX)) return 0;
    int sh;
    (void)frexp(mx, &sh);
    if (sh == 0) return 0;
    d1 = ldexp(d1, -sh);
    d0 = ldexp(d0, -sh);
    E += sh;
    return sh;
}

// Sum 16 values over the 64 lanes at once: each butterfly step halves the values a lane keeps
// (exchanging the other half with lane ^ mask), then the last two steps sum across each quad.
// Lane q ends with the total of value k = 8 b5 + 4 b4 + 2 b3 + b2 (bits of q) in v[0]; the four
// lanes of a quad hold the same bits, and the order of every addition is fixed.
__device__ __forceinline__ void reduce16(double (&v)[16], int lane) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const bool up = lane & 32;
        const double keep = up ? v[8 + i] : v[i], send = up ? v[i] : v[8 + i];
        v[i] = keep + __shfl_xor(send, 32);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const bool up = lane & 16;
        const double keep = up ? v[4 + i] : v[i], send = up ? v[i] : v[4 + i];
        v[i] = keep + __shfl_xor(send, 16);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const bool up = lane & 8;
        const double keep = up ? v[2 + i] : v[i], send = up ? v[i] : v[2 + i];
        v[i] = keep + __shfl_xor(send, 8);
    }
    {
        const bool up = lane & 4;
        const double keep = up ? v[1] : v[0], send = up ? v[0] : v[1];
        v[0] = keep + __shfl_xor(send, 4);
    }
    v[0] += __shfl_xor(v[0], 2);
    v[0] += __shfl_xor(v[0], 1);
}
}  // namespace

// Recurrence coefficients of the step l -> l + 1 at order |m| (Spin.jl:417-432): with
// den = l sqrt((l+1)^2 - m^2) sqrt((l+1)^2 - s^2), a1 = (2l+1) l (l+1) / den,
// |a2| = (2l+1) |m| |s| / den, a0 = (l+1) sqrt(l^2 - m^2) sqrt(l^2 - s^2) / den, and N_{l+1}.
__global__ void spin_tab_kernel(int L, int s, const int64_t *__restrict__ tofs, const double *__restrict__ Nl,
                                double4 *__restrict__ tab) {
#pragma clang fp contract(off)
    const int am = blockIdx.y;
    const int as = s < 0 ? -s : s;
    const int l0 = am > as ? am : as;
    const int Lc = l0 + int(blockIdx.x * blockDim.x + threadIdx.x);
    if (Lc >= L) return;
    const int L1 = Lc + 1;
    const double sm1 = sqrt(fmax(0.0, double(L1 * L1 - am * am)));
    const double sn1 = sqrt(fmax(0.0, double(L1 * L1 - as * as)));
    const double den = double(Lc) * sm1 * sn1;
    const double sm0 = sqrt(fmax(0.0, double(Lc * Lc - am * am)));
    const double sn0 = sqrt(fmax(0.0, double(Lc * Lc - as * as)));
    double4 c;
    c.x = double(2 * Lc + 1) * double(Lc) * double(L1) / den;
    c.y = double(2 * Lc + 1) * double(am) * double(as) / den;
    c.z = double(L1) * sm0 * sn0 / den;
    c.w = Nl[L1];
    tab[tofs[am] + (Lc - l0)] = c;
}

// Forward (spin_rings_from_alm!, Spin.jl:316-322): one wavefront per (order, 64-ring tile), the
// coefficients of transforms b0 .. b0 + nb - 1 (nb <= NB) accumulated per lane in scaled form.
template <int NB>
__global__ void __launch_bounds__(64)
spin_fwd_kernel(const int4 *__restrict__ items, const double *__restrict__ costh, const double *__restrict__ l2c,
                const double *__restrict__ l2s, const double4 *__restrict__ seed, const double *__restrict__ Nl,
                const double4 *__restrict__ tab, const int64_t *__restrict__ tofs, const double2 *__restrict__ alm,
                double2 *__restrict__ Gd, int n, int L, int s, int mmax, int b0, int nb) {
    const int4 it = items[blockIdx.x];
    const int m = it.x, j0 = it.y, nr = it.z;
    const int lane = threadIdx.x;
    const int am = m < 0 ? -m : m, as = s < 0 ? -s : s;
    const int l0 = am > as ? am : as;
    if (l0 > L) return;
    const int j = j0 + min(lane, nr - 1);
    const double x = costh[j];
    double d1;
    int E;
    spin_seed(seed[m + L], l2c[j], l2s[j], d1, E);
    double d0 = 0.0;
    const int64_t astride = int64_t(L + 1) * (2 * L + 1);
    const double2 *col = alm + int64_t(m + L) * (L + 1) + int64_t(b0) * astride;
    double2 acc[NB];
    {
        const double w = Nl[l0] * d1;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            acc[b] = make_double2(0.0, 0.0);
            if (b < nb) {
                const double2 a = col[b * astride + l0];
                acc[b] = make_double2(a.x * w, a.y * w);
            }
        }
    }
    const double4 *tb = tab + (tofs[am] - l0);
    const double sg2 = m * s <= 0 ? 1.0 : -1.0;  // sign of m n with n = -s
    int steps = 0;
    for (int Lc = l0; Lc < L; ++Lc) {
        const double4 c = tb[Lc];
        const double dn = (c.x * x - sg2 * c.y) * d1 - c.z * d0;
        d0 = d1;
        d1 = dn;
        const double w = c.w * dn;
        const double2 *cp = col + (Lc + 1);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            if (b < nb) {
                const double2 a = cp[b * astride];
                acc[b].x = fma(a.x, w, acc[b].x);
                acc[b].y = fma(a.y, w, acc[b].y);
            }
        }
        if (++steps == 64) {
            steps = 0;
            if (Lc + 1 < L) {
                const int sh = spin_renorm(d1, d0, E);
                if (sh != 0) {
#pragma unroll
                    for (int b = 0; b < NB; ++b) acc[b] = make_double2(ldexp(acc[b].x, -sh), ldexp(acc[b].y, -sh));
                }
            }
        }
    }
    if (lane < nr) {
        const int64_t ncol = 2 * mmax + 1;
#pragma unroll
        for (int b = 0; b < NB; ++b)
            if (b < nb)
                Gd[(int64_t(b0 + b) * ncol + (m + mmax)) * n + j0 + lane] =
                    make_double2(ldexp(acc[b].x, E), ldexp(acc[b].y, E));
    }
}

// Adjoint (spin_alm_from_rings!, Spin.jl:324-334): one workgroup per order, one wavefront per
// 64-ring tile (kSpinWaves tiles at a time); each wavefront reduces its rings' products per degree
// in blocks of 16 degrees, and the wavefronts' sums are added into the order's column of alm in
// wavefront order. alm must be zero on entry; every order writes only its own column.
template <int NB>
__global__ void __launch_bounds__(64 * kSpinWaves)
spin_adj_kernel(const double *__restrict__ costh, const double *__restrict__ l2c, const double *__restrict__ l2s,
                const double4 *__restrict__ seed, const double *__restrict__ Nl, const double4 *__restrict__ tab,
                const int64_t *__restrict__ tofs, const int32_t *__restrict__ jlo, const int32_t *__restrict__ jhi,
                const double2 *__restrict__ Gd, double2 *__restrict__ alm, int n, int L, int s, int mmax, int b0,
                int nb) {
    __shared__ double2 red[kSpinWaves][16][NB];
    const int m = int(blockIdx.x) - mmax;
    const int am = m < 0 ? -m : m, as = s < 0 ? -s : s;
    const int l0 = am > as ? am : as;
    if (l0 > L) return;
    const int lo = jlo[am], hi = jhi[am];
    if (hi < lo) return;
    const int ntile = (hi - lo + 64) / 64;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t astride = int64_t(L + 1) * (2 * L + 1), ncol = 2 * mmax + 1;
    double2 *col = alm + int64_t(m + L) * (L + 1) + int64_t(b0) * astride;
    const double4 sd = seed[m + L];
    const double4 *tb = tab + (tofs[am] - l0);
    const double sg2 = m * s <= 0 ? 1.0 : -1.0;
    for (int tg = 0; tg * kSpinWaves < ntile; ++tg) {
        const int tile = tg * kSpinWaves + w;
        const int j = lo + tile * 64 + lane;
        const bool ok = tile < ntile && j <= hi;
        const int jj = ok ? j : lo;
        const double x = costh[jj];
        double d1 = 0.0, d0 = 0.0;
        int E = 0;
        if (ok) spin_seed(sd, l2c[jj], l2s[jj], d1, E);
        double2 g[NB], gs[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            g[b] = make_double2(0.0, 0.0);
            if (ok && b < nb) g[b] = Gd[(int64_t(b0 + b) * ncol + (m + mmax)) * n + jj];
            gs[b] = make_double2(ldexp(g[b].x, E), ldexp(g[b].y, E));
        }
        int steps = 0;
        for (int lb = l0; lb <= L; lb += 16) {
            double pr[NB][16], pi[NB][16];
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const int l = lb + k;
                double val = 0.0;
                if (l <= L) {
                    if (l > l0) {
                        const double4 c = tb[l - 1];
                        const double dn = (c.x * x - sg2 * c.y) * d1 - c.z * d0;
                        d0 = d1;
                        d1 = dn;
                        if (++steps == 64) {
                            steps = 0;
                            if (l < L && spin_renorm(d1, d0, E) != 0) {
#pragma unroll
                                for (int b = 0; b < NB; ++b) gs[b] = make_double2(ldexp(g[b].x, E), ldexp(g[b].y, E));
                            }
                        }
                    }
                    val = d1;
                }
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    pr[b][k] = val * gs[b].x;
                    pi[b][k] = val * gs[b].y;
                }
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                reduce16(pr[b], lane);
                reduce16(pi[b], lane);
            }
            if ((lane & 3) == 0) {
                const int k = ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 + ((lane >> 2) & 1);
#pragma unroll
                for (int b = 0; b < NB; ++b) red[w][k][b] = make_double2(pr[b][0], pi[b][0]);
            }
            __syncthreads();
            if (threadIdx.x < 16 * NB) {
                const int k = threadIdx.x & 15, b = threadIdx.x >> 4;
                const int l = lb + k;
                if (l <= L && b < nb) {
                    double sx = 0.0, sy = 0.0;
                    for (int q = 0; q < kSpinWaves; ++q) {
                        sx += red[q][k][b].x;
                        sy += red[q][k][b].y;
                    }
                    const double N = Nl[l];
                    double2 *o = col + b * astride + l;
                    const double2 cur = *o;
                    *o = make_double2(cur.x + sx * N, cur.y + sy * N);
                }
            }
            __syncthreads();
        }
    }
}

__global__ void zero_fill_kernel(double2 *__restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        p[i] = make_double2(0.0, 0.0);
}

void launch_spin_tables(const RingDev &r, const SpinDev &sd, hipStream_t s) {
    if (r.L < 1) return;
    spin_tab_kernel<<<dim3((r.L + 63) / 64, r.L + 1), 64, 0, s>>>(r.L, sd.s, sd.tofs, sd.Nl, sd.tab);
}

void launch_spin_fwd(const RingDev &r, const SpinDev &sd, const double2 *alm, double2 *Gd, hipStream_t s) {
    if (sd.nitems == 0) return;
    for (int b0 = 0; b0 < r.B; b0 += 4) {
        const int nb = std::min(4, r.B - b0);
        if (nb == 1)
            spin_fwd_kernel<1><<<sd.nitems, 64, 0, s>>>(sd.items, r.costh, sd.l2c, sd.l2s, sd.seed, sd.Nl, sd.tab,
                                                         sd.tofs, alm, Gd, r.n, r.L, sd.s, r.mmax, b0, nb);
        else
            spin_fwd_kernel<4><<<sd.nitems, 64, 0, s>>>(sd.items, r.costh, sd.l2c, sd.l2s, sd.seed, sd.Nl, sd.tab,
                                                         sd.tofs, alm, Gd, r.n, r.L, sd.s, r.mmax, b0, nb);
    }
}

void launch_spin_adj(const RingDev &r, const SpinDev &sd, const double2 *Gd, double2 *alm, hipStream_t s) {
    const int64_t na = int64_t(r.L + 1) * (2 * r.L + 1) * r.B;
    zero_fill_kernel<<<unsigned(std::min<int64_t>((na + 255) / 256, 4096)), 256, 0, s>>>(alm, na);
    for (int b0 = 0; b0 < r.B; b0 += 2) {
        const int nb = std::min(2, r.B - b0);
        if (nb == 1)
            spin_adj_kernel<1><<<2 * r.mmax + 1, 64 * kSpinWaves, 0, s>>>(
                r.costh, sd.l2c, sd.l2s, sd.seed, sd.Nl, sd.tab, sd.tofs, r.jlo, r.jhi, Gd, alm, r.n, r.L, sd.s,
                r.mmax, b0, nb);
        else
            spin_adj_kernel<2><<<2 * r.mmax + 1, 64 * kSpinWaves, 0, s>>>(
                r.costh, sd.l2c, sd.l2s, sd.seed, sd.Nl, sd.tab, sd.tofs, r.jlo, r.jhi, Gd, alm, r.n, r.L, sd.s,
                r.mmax, b0, nb);
    }
}

}  // namespace pfsht
