// spin_kernels.hip — the spin-weighted ring engine (src/Spin.jl:170-545): G[j, m] = sum_l a_lm N_l
// d^l_{m,-s}(theta_j) and its exact transpose, with the three-term recurrence in degree at fixed
// (m, -s) seeded at l0 = max(|m|, |s|) from the extremal closed form (tabulated per order and
// ring in log2 form). Each lane is one ring and carries its own power-of-two exponent,
// renormalised every 64 degrees; every order m in [-mmax, mmax] is walked on its own (d^l_{m,-s} and
// d^l_{-m,-s} differ). The work per retained (l, m, ring) triple is one recurrence step (three
// FP64 FMAs) and one complex product per transform, so the passes are FP64-issue bound like the
// scalar ones.
#include "pfsht_common.hpp"
#include "adj_rows.hpp"

#include <cfloat>

namespace pfsht {

namespace {
constexpr int kSpinAdjWaves = 4;  // wavefronts per adjoint workgroup (each with its own LDS tile)
constexpr int kG = 4;             // degrees whose coefficients are fetched together

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
// One degree step of a single-transform spin sweep with the step's coefficients broadcast from
// lane K of the 16-lane row by DPP on the FMAs: d_{l+1} = A1 (x d_l) + A2 d_l + A0 d_{l-1} with
// A1 = a1, A2 = -sgn a2, A0 = -a0, then acc += (a_lm N) d_{l+1}. s_nop 1: a DPP source must not be
// written by VALU in the two instructions before its use.
template <int K>
__device__ __forceinline__ double spin_dpp_step(double x, double d1, double d0, double A1, double A2, double A0,
                                                double ANr, double ANi, double &accr, double &acci) {
    double t, dn;
    asm("s_nop 1\n\t"
        "v_mul_f64 %0, %4, %5\n\t"
        "v_mov_b64 %1, 0\n\t"
        "v_fmac_f64_dpp %1, %7, %0 row_newbcast:%12 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %1, %8, %5 row_newbcast:%12 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %1, %9, %6 row_newbcast:%12 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %2, %10, %1 row_newbcast:%12 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %3, %11, %1 row_newbcast:%12 row_mask:0xf bank_mask:0xf"
        : "=&v"(t), "=&v"(dn), "+v"(accr), "+v"(acci)
        : "v"(x), "v"(d1), "v"(d0), "v"(A1), "v"(A2), "v"(A0), "v"(ANr), "v"(ANi), "n"(K));
    return dn;
}

// Steps K .. 15 of a 16-step group; GUARD: only K < kend.
template <int K, bool GUARD>
struct SpinDppGroup {
    static __device__ __forceinline__ void run(double x, double &d1, double &d0, double A1, double A2, double A0,
                                               double ANr, double ANi, double &accr, double &acci, int kend) {
        if (!GUARD || K < kend) {
            const double dn = spin_dpp_step<K>(x, d1, d0, A1, A2, A0, ANr, ANi, accr, acci);
            d0 = d1;
            d1 = dn;
        }
        SpinDppGroup<K + 1, GUARD>::run(x, d1, d0, A1, A2, A0, ANr, ANi, accr, acci, kend);
    }
};
template <bool GUARD>
struct SpinDppGroup<16, GUARD> {
    static __device__ __forceinline__ void run(double, double &, double &, double, double, double, double, double,
                                               double &, double &, int) {}
};

// The recurrence step alone (the adjoint's phase 1), the same operations as spin_dpp_step:
// t = x d_l, d_{l+1} = A1 t (rounded), + A2 d_l, + A0 d_{l-1} (FMAs).
template <int K>
__device__ __forceinline__ double spin_dpp_rstep(double x, double d1, double d0, double A1, double A2, double A0) {
    double t, dn;
    asm("s_nop 1\n\t"
        "v_mul_f64 %0, %2, %3\n\t"
        "v_mov_b64 %1, 0\n\t"
        "v_fmac_f64_dpp %1, %5, %0 row_newbcast:%8 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %1, %6, %3 row_newbcast:%8 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %1, %7, %4 row_newbcast:%8 row_mask:0xf bank_mask:0xf"
        : "=&v"(t), "=&v"(dn)
        : "v"(x), "v"(d1), "v"(d0), "v"(A1), "v"(A2), "v"(A0), "n"(K));
    return dn;
}
// The same step with per-lane coefficients (no broadcast), for guarded sub-tiles.
__device__ __forceinline__ double spin_rstep(double x, double d1, double d0, double A1, double A2, double A0) {
#pragma clang fp contract(off)
    const double t = x * d1;
    double dn = A1 * t;
    dn = fma(A2, d1, dn);
    return fma(A0, d0, dn);
}
// Steps K .. 15 of a full sub-tile, row K of the LDS tile taking d_{lb + K}.
template <int K>
struct SpinAdjPhase1 {
    static __device__ __forceinline__ void run(double x, double &d1, double &d0, double A1, double A2, double A0,
                                               double (*W)[kWS], int wcol) {
        const double dn = spin_dpp_rstep<K>(x, d1, d0, A1, A2, A0);
        d0 = d1;
        d1 = dn;
        W[K][wcol] = dn;
        SpinAdjPhase1<K + 1>::run(x, d1, d0, A1, A2, A0, W, wcol);
    }
};
template <>
struct SpinAdjPhase1<kSub> {
    static __device__ __forceinline__ void run(double, double &, double &, double, double, double, double (*)[kWS],
                                               int) {}
};

// acc += (a_lm N) d_{l+1} for one more transform, the coefficient broadcast from lane K.
template <int K>
__device__ __forceinline__ void spin_dpp_acc(double &accr, double &acci, double ANr, double ANi, double dn) {
    asm("s_nop 1\n\t"
        "v_fmac_f64_dpp %0, %2, %4 row_newbcast:%5 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %1, %3, %4 row_newbcast:%5 row_mask:0xf bank_mask:0xf"
        : "+v"(accr), "+v"(acci)
        : "v"(ANr), "v"(ANi), "v"(dn), "n"(K));
}
// Steps K .. 15 of a 16-step group for NB transforms; GUARD: only K < kend.
template <int K, bool GUARD, int NB>
struct SpinDppGroupN {
    static __device__ __forceinline__ void run(double x, double &d1, double &d0, double A1, double A2, double A0,
                                               const double (&ANr)[NB], const double (&ANi)[NB], double2 (&acc)[NB],
                                               int kend) {
        if (!GUARD || K < kend) {
            const double dn = spin_dpp_rstep<K>(x, d1, d0, A1, A2, A0);
            d0 = d1;
            d1 = dn;
#pragma unroll
            for (int b = 0; b < NB; ++b) spin_dpp_acc<K>(acc[b].x, acc[b].y, ANr[b], ANi[b], dn);
        }
        SpinDppGroupN<K + 1, GUARD, NB>::run(x, d1, d0, A1, A2, A0, ANr, ANi, acc, kend);
    }
};
template <bool GUARD, int NB>
struct SpinDppGroupN<16, GUARD, NB> {
    static __device__ __forceinline__ void run(double, double &, double &, double, double, double,
                                               const double (&)[NB], const double (&)[NB], double2 (&)[NB], int) {}
};

template <int NB>
__global__ void __launch_bounds__(64)
spin_fwd_kernel(const int4 *__restrict__ items, const double *__restrict__ costh, const double *__restrict__ l2c,
                const double *__restrict__ l2s, const double4 *__restrict__ seed, const double *__restrict__ Nl,
                const double4 *__restrict__ tab, const int64_t *__restrict__ tofs, InCols alm,
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
    // column m + L of transform b0 + b (a transform past nb reads the last one's, unused)
    gcptr col[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) col[b] = alm.col(b0 + min(b, nb - 1)) + int64_t(m + L) * (L + 1);
    double2 acc[NB];
    {
        const double w = Nl[l0] * d1;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            acc[b] = make_double2(0.0, 0.0);
            if (b < nb) {
                const double2 a = gload(col[b] + l0);
                acc[b] = make_double2(a.x * w, a.y * w);
            }
        }
    }
    const double sg2 = m * s <= 0 ? 1.0 : -1.0;  // sign of m n with n = -s
    // Groups of 16 steps, lane k of every row holding step q0 + k's table entry and coefficients
    // (vector loads, one group ahead, at indices clamped into the order's steps: a lane past the
    // last step loads the last one and is never broadcast), read by the FMAs through DPP row
    // broadcasts; renormalised every 64 steps.
    const int nsteps = L - l0, kq = lane & 15;
    const double4 *__restrict__ t0 = tab + tofs[am];
    const int la = l0 + 1;
    double4 cn = make_double4(0.0, 0.0, 0.0, 0.0);
    double2 an[NB];
    auto group_load = [&](int q) {
        cn = t0[q];
#pragma unroll
        for (int b = 0; b < NB; ++b)
        {
            const double2 v = gload(col[b] + (la + q));
            an[b] = (NB == 1 || b < nb) ? v : make_double2(0.0, 0.0);
        }
    };
#pragma unroll
    for (int b = 0; b < NB; ++b) an[b] = make_double2(0.0, 0.0);
    if (nsteps > 0) group_load(min(kq, nsteps - 1));
    for (int q0 = 0; q0 < nsteps; q0 += 16) {
        const double4 c = cn;
        double2 a[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) a[b] = an[b];
        if (q0 + 16 < nsteps) group_load(min(q0 + 16 + kq, nsteps - 1));
        const double A1 = c.x, A2 = -sg2 * c.y, A0 = -c.z;
        const int kend = min(16, nsteps - q0);
        if constexpr (NB == 1) {
            const double ANr = a[0].x * c.w, ANi = a[0].y * c.w;
            if (kend == 16) SpinDppGroup<0, false>::run(x, d1, d0, A1, A2, A0, ANr, ANi, acc[0].x, acc[0].y, 16);
            else SpinDppGroup<0, true>::run(x, d1, d0, A1, A2, A0, ANr, ANi, acc[0].x, acc[0].y, kend);
        } else {
            double ANr[NB], ANi[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                ANr[b] = a[b].x * c.w;
                ANi[b] = a[b].y * c.w;
            }
            if (kend == 16) SpinDppGroupN<0, false, NB>::run(x, d1, d0, A1, A2, A0, ANr, ANi, acc, 16);
            else SpinDppGroupN<0, true, NB>::run(x, d1, d0, A1, A2, A0, ANr, ANi, acc, kend);
        }
        if (((q0 + 16) & 63) == 0) {
            const int sh = spin_renorm(d1, d0, E);
            if (sh != 0) {
#pragma unroll
                for (int b = 0; b < NB; ++b) acc[b] = make_double2(ldexp(acc[b].x, -sh), ldexp(acc[b].y, -sh));
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

// Adjoint (spin_alm_from_rings!, Spin.jl:324-334): one wavefront per (order, 64-ring tile), as in
// the forward. Per 16-degree sub-tile, phase 1 (lanes = rings) advances the recurrence and writes
// d to the wavefront's LDS tile, phase 2 forms the tile's ring sums d G 2^E per degree (row
// broadcast FMAs, then the sum over the four rows), and row 0 writes them to the tile's partials.
// The ring's exponent rides on its G operands (exact, refreshed when a renormalisation every 64
// degrees changes it), so one sub-tile shares one scale.
template <int NB>
__global__ void __launch_bounds__(64 * kSpinAdjWaves)
spin_adj_kernel(const int4 *__restrict__ items, int nitems, const double *__restrict__ costh,
                const double *__restrict__ l2c, const double *__restrict__ l2s, const double4 *__restrict__ seed,
                const double4 *__restrict__ tab, const int64_t *__restrict__ tofs, const int64_t *__restrict__ pofs,
                const double2 *__restrict__ Gd, double2 *__restrict__ partials, int n, int L, int s, int mmax, int B,
                int b0, int nb) {
    __shared__ double lds[kSpinAdjWaves * kSub * kWS];
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
    const int idx = int(blockIdx.x) * kSpinAdjWaves + wave;
    if (idx >= nitems) return;  // wave-uniform; the kernel has no workgroup barrier
    const int4 it = items[idx];
    const int m = __builtin_amdgcn_readfirstlane(it.x), j0 = __builtin_amdgcn_readfirstlane(it.y);
    const int nr = __builtin_amdgcn_readfirstlane(it.z), tile = __builtin_amdgcn_readfirstlane(it.w);
    const int lane = threadIdx.x & 63;
    const int am = m < 0 ? -m : m, as = s < 0 ? -s : s;
    const int l0 = am > as ? am : as;
    if (l0 > L) return;
    double(*W)[kWS] = reinterpret_cast<double(*)[kWS]>(lds + wave * kSub * kWS);
    const bool act = lane < nr;
    const int j = j0 + (act ? lane : 0);
    const double x = costh[j];
    double d1, d0 = 0.0;
    int E;
    spin_seed(seed[m + L], l2c[j], l2s[j], d1, E);
    if (!act) d1 = 0.0;
    const int64_t ncol = 2 * mmax + 1;
    double g[NB][4], gs[NB][4];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        double2 G = make_double2(0.0, 0.0);
        if (act && u < nb) G = Gd[(int64_t(b0 + u) * ncol + (m + mmax)) * n + j];
        g[u][0] = G.x;
        g[u][1] = G.y;
        g[u][2] = g[u][3] = 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) gs[u][c] = ldexp(g[u][c], E);
    }
    const int t = lane & 15, row = lane >> 4;
    const int wcol = lane + (lane >> 4);
    const double *__restrict__ wr = &W[t][17 * row];
    const int nl = L - l0 + 1;
    double2 *__restrict__ out = partials + pofs[m + mmax] + (int64_t(tile) * nl) * B + b0;
    const double4 *__restrict__ tb = tab + (tofs[am] - l0);
    const double sg2 = m * s <= 0 ? 1.0 : -1.0;
    int since = 0;  // degrees since the last renormalisation
    // Full sub-tiles (lb > l0, 16 degrees): lane k of every row holds step lb - 1 + k's entry, a
    // vector load one sub-tile ahead at an index clamped into the order's steps, read by the
    // recurrence through DPP row broadcasts, the same operations as the forward pass.
    const int kq = lane & 15;
    double4 cn = make_double4(0.0, 0.0, 0.0, 0.0);
    for (int lb = l0; lb <= L; lb += kSub) {
        const int tl = min(kSub, L - lb + 1);
        const double4 cf = cn;  // this sub-tile's entries when it is full (loaded one sub-tile ago)
        // the next sub-tile, when it is full (lb + 31 <= L): entries lb + 15 .. lb + 30
        if (lb + 2 * kSub - 1 <= L) cn = tb[min(lb + kSub - 1 + kq, L - 1)];
        if (tl == kSub && lb > l0) {  // a full sub-tile of steps: entries lb - 1 .. lb + 14
            SpinAdjPhase1<0>::run(x, d1, d0, cf.x, -sg2 * cf.y, -cf.z, W, wcol);
        } else {  // the seed's sub-tile and the last one: guarded, entries clamped in range
#pragma unroll
            for (int k0 = 0; k0 < kSub; k0 += kG) {
                double4 c[kG];
#pragma unroll
                for (int q = 0; q < kG; ++q) c[q] = tb[min(max(lb + k0 + q - 1, l0), L - 1)];
#pragma unroll
                for (int q = 0; q < kG; ++q) {
                    const int k = k0 + q, l = lb + k;
                    if (k < tl) {
                        if (l > l0) {
                            const double dn = spin_rstep(x, d1, d0, c[q].x, -sg2 * c[q].y, -c[q].z);
                            d0 = d1;
                            d1 = dn;
                        }
                        W[k][wcol] = d1;
                    }
                }
            }
        }
        wave_lds_order();
        double a[NB][4];
#pragma unroll
        for (int u = 0; u < NB; ++u) a[u][0] = a[u][1] = a[u][2] = a[u][3] = 0.0;
        RowSweep<0, NB, false>::run(wr, a, gs);
        wave_lds_order();
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            a[u][0] = row_sum4(a[u][0]);
            a[u][1] = row_sum4(a[u][1]);
        }
        if (row == 0 && t < tl) {
#pragma unroll
            for (int u = 0; u < NB; ++u)
                if (u < nb) out[int64_t(lb - l0 + t) * B + u] = make_double2(a[u][0], a[u][1]);
        }
        since += tl;
        if (since >= 64) {
            since = 0;
            if (spin_renorm(d1, d0, E) != 0) {
#pragma unroll
                for (int u = 0; u < NB; ++u)
#pragma unroll
                    for (int c = 0; c < 4; ++c) gs[u][c] = ldexp(g[u][c], E);
            }
        }
    }
}

// Writes every alm slot: N_l times the tile-ordered sum of the order's partials where (l, m) is
// retained (l >= max(|m|, |s|), |m| <= mmax), zero elsewhere (the adjoint overwrites alm).
__global__ void spin_adj_finalize(const int64_t *__restrict__ pofs, const int32_t *__restrict__ ntiles,
                                  const double2 *__restrict__ partials, const double *__restrict__ Nl,
                                  OutCols alm, int L, int s, int mmax, int B) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    const int b = blockIdx.z;
    if (l > L) return;
    const int m = c - L, am = m < 0 ? -m : m, as = s < 0 ? -s : s;
    const int l0 = am > as ? am : as;
    double vx = 0.0, vy = 0.0;
    if (am <= mmax && l >= l0) {
        const int nt = ntiles[m + mmax], nl = L - l0 + 1;
        const double2 *p = partials + pofs[m + mmax] + int64_t(l - l0) * B + b;
        for (int t = 0; t < nt; ++t) {
            const double2 q = p[int64_t(t) * nl * B];
            vx += q.x;
            vy += q.y;
        }
        const double N = Nl[l];
        vx *= N;
        vy *= N;
    }
    gstore(alm.col(b) + (int64_t(c) * (L + 1) + l), make_double2(vx, vy));
}

void launch_spin_tables(const RingDev &r, const SpinDev &sd, hipStream_t s) {
    if (r.L < 1) return;
    spin_tab_kernel<<<dim3((r.L + 63) / 64, r.L + 1), 64, 0, s>>>(r.L, sd.s, sd.tofs, sd.Nl, sd.tab);
}

void launch_spin_fwd(const RingDev &r, const SpinDev &sd, InCols alm, int nbt, double2 *Gd, hipStream_t s) {
    if (sd.nitems == 0) return;
    for (int b0 = 0; b0 < nbt; b0 += 4) {
        const int nb = std::min(4, nbt - b0);
        if (nb == 1)
            spin_fwd_kernel<1><<<sd.nitems, 64, 0, s>>>(sd.items, r.costh, sd.l2c, sd.l2s, sd.seed, sd.Nl, sd.tab,
                                                         sd.tofs, alm, Gd, r.n, r.L, sd.s, r.mmax, b0, nb);
        else
            spin_fwd_kernel<4><<<sd.nitems, 64, 0, s>>>(sd.items, r.costh, sd.l2c, sd.l2s, sd.seed, sd.Nl, sd.tab,
                                                         sd.tofs, alm, Gd, r.n, r.L, sd.s, r.mmax, b0, nb);
    }
}

void launch_spin_adj(const RingDev &r, const SpinDev &sd, const double2 *Gd, OutCols alm, int nbt, hipStream_t s) {
    if (sd.nitems > 0) {
        const unsigned grid = unsigned((sd.nitems + kSpinAdjWaves - 1) / kSpinAdjWaves);
        for (int b0 = 0; b0 < nbt; b0 += 4) {
            const int nb = std::min(4, nbt - b0);
#define PFSHT_SPIN_ADJ(NB)                                                                                   \
    spin_adj_kernel<NB><<<grid, 64 * kSpinAdjWaves, 0, s>>>(sd.items, sd.nitems, r.costh, sd.l2c, sd.l2s, sd.seed, \
                                                            sd.tab, sd.tofs, sd.pofs, Gd, sd.partials, r.n, r.L, sd.s, \
                                                            r.mmax, r.B, b0, nb)
            if (nb == 1) PFSHT_SPIN_ADJ(1);
            else if (nb == 2) PFSHT_SPIN_ADJ(2);
            else PFSHT_SPIN_ADJ(4);
#undef PFSHT_SPIN_ADJ
        }
    }
    dim3 fgrid((r.L + 1 + 255) / 256, 2 * r.L + 1, nbt);
    spin_adj_finalize<<<fgrid, 256, 0, s>>>(sd.pofs, sd.ntiles, sd.partials, sd.Nl, alm, r.L, sd.s, r.mmax, r.B);
}

}  // namespace pfsht
