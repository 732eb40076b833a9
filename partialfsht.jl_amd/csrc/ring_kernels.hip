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
// computes 16-degree sub-tiles of w = q 2^E with lanes over rings, transposes them through LDS,
// and sums them against G with row-broadcast FMAs (see the adjoint section). Per-tile partials go
// to a plan-owned buffer and a finalize kernel adds the tiles of each order in tile order: no
// float atomics, bit-identical run to run.
#include "pfsht_common.hpp"
#include "adj_rows.hpp"

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

// L2 prefetch without losing scalar loads. Each lane reads one 4-byte word of a 128-byte line at
// uniform_base + voff (voff per lane, loop-invariant, so its VGPR is never rewritten while the
// load is in flight) into a register that is folded into `acc` one block later: the wait for
// it lands a whole block of FMAs after the issue. A plain load cannot clobber memory, so the
// compiler still proves the coefficient streams read-only and keeps them on scalar loads.
__device__ __forceinline__ int l2_touch(const char *base, int voff, bool on) {
    return on ? *reinterpret_cast<const int *>(base + voff) : 0;
}

constexpr int kPrefetchAhead = 128;  // degrees ahead of the sweep
// Value the prefetch accumulator is compared with to keep its loads alive: an XOR of 4-byte words
// of coefficient data cannot be predicted by the compiler, and a match only rewrites one element
// of an output the kernel writes anyway.
constexpr int kNeverSink = 0x7f3c91e5;

}  // namespace

namespace {
// q_l from q_{l-1}, q_{l-2} (src/Legendre.jl:406-408): the one expression every sweep uses, so the
// checkpoint sweep and the segments reproduce the same bits.
__device__ __forceinline__ double rec_step(double2 c, double x, double q1, double q2) {
    return fma(c.x * x, q1, -(c.y * q2));
}
// q_{m+1} from q_m
__device__ __forceinline__ double first_step(int m, double x, double q) { return sqrt(double(2 * m + 3)) * x * q; }
// last degree of the 64-degree block holding degree l
__device__ __forceinline__ int block_end(int m, int l) { return m + ((l - m) | 63); }
}  // namespace

// ------------------------------------------------------------------------------------------------
// Checkpoints
// ------------------------------------------------------------------------------------------------
// One wavefront per forward item sweeps the recurrence of its 64 rings from the seed and stores the
// state at the start of segments 1 .. nseg - 1 (after the renormalisation that closes the previous
// block). Runs once per plan.
__global__ void __launch_bounds__(64)
ring_ckpt_kernel(const SegItem *__restrict__ items, const double *__restrict__ costh,
                 const double *__restrict__ log2sin, const double *__restrict__ log2c,
                 const double2 *__restrict__ ab, const int64_t *__restrict__ abofs, double2 *__restrict__ ckq,
                 int32_t *__restrict__ ckE, int L) {
    const SegItem it = items[blockIdx.x];
    const int m = __builtin_amdgcn_readfirstlane(it.m);
    const int nseg = __builtin_amdgcn_readfirstlane(it.nseg);
    const int nr = __builtin_amdgcn_readfirstlane(it.nr);
    if (nr <= 0 || nseg <= 1) return;
    const int lane = threadIdx.x & 63;
    const bool active = lane < nr;
    const int j = it.j0 + (active ? lane : 0);
    const double x = costh[j];
    double q2;
    int E;
    seed(m == 0 ? log2c[0] : double(m) * log2sin[j] + log2c[m], q2, E);
    if (!active) q2 = 0.0;
    double q1 = m < L ? first_step(m, x, q2) : 0.0;
    const double2 *__restrict__ abm = ab + abofs[m] - (m + 2);
    int sgi = 1, lnext = seg_begin(m, L, nseg, 1);
    for (int l = m + 2; l <= L; ++l) {
        const double q = rec_step(abm[l], x, q1, q2);
        q2 = q1;
        q1 = q;
        if (l == block_end(m, l)) {
            const int sh = renorm_shift(q1, q2);
            q1 = ldexp(q1, -sh);
            q2 = ldexp(q2, -sh);
            E += sh;
        }
        if (l + 1 == lnext) {
            const int64_t row = (it.ck + sgi - 1) * 64 + lane;
            ckq[row] = make_double2(q1, q2);
            ckE[row] = E;
            if (++sgi == nseg) break;
            lnext = seg_begin(m, L, nseg, sgi);
        }
    }
}

void launch_ring_checkpoints(const RingDev &r, hipStream_t s) {
    if (r.nfitems == 0 || r.nck == 0) return;
    ring_ckpt_kernel<<<r.nfitems, 64, 0, s>>>(r.fitems, r.costh, r.log2sin, r.log2c, r.ab, r.abofs, r.ckq, r.ckE,
                                              r.L);
}

// ------------------------------------------------------------------------------------------------
// Forward
// ------------------------------------------------------------------------------------------------
// Degrees l .. le of order m feeding NB coefficient sets (transforms b0 .. b0 + NB - 1): the
// recurrence is computed once per ring and reused for every transform. NEG: m > 0, so the order
// -m column is accumulated too. The accumulators share the exponent E of (q1, q2).
template <int NB, bool NEG>
__device__ __forceinline__ void fwd_sweep(const double2 *__restrict__ abm, const double2 *__restrict__ cp0,
                                          const double2 *__restrict__ cm0, int64_t LB, double x, double q1,
                                          double q2, int &E, int l, int le, int m, double (&ar)[NB],
                                          double (&ai)[NB], double (&br)[NB], double (&bi)[NB], int &sink) {
    // transform t's order-m column is cp0 + t LB (row l - m holds degree l), order -m likewise
    const int lane = threadIdx.x & 63;
    // Prefetch lanes: 8 lines of 128 B cover a 64-degree run of one 16-byte stream. Stream 0 is
    // the (a, b) row, streams 1..NB the +m columns and NB+1..2NB the -m columns (NEG only).
    constexpr int NS = 1 + 2 * NB;
    constexpr int NCH = (8 * NS + 63) / 64;
    int voff[NCH];
    bool isab[NCH], used[NCH];
    int lstart[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const int k = 64 * c + lane, st = k >> 3, line = k & 7;
        lstart[c] = 8 * line;  // degree offset of the line within the run
        isab[c] = st == 0;
        used[c] = st < (NEG ? NS : 1 + NB);
        int64_t o = 128 * line;
        if (st >= 1 && st <= NB) o += int64_t(st - 1) * LB * 16;
        if (st > NB) o += int64_t(st - 1 - NB) * LB * 16 + (reinterpret_cast<const char *>(cm0) - reinterpret_cast<const char *>(cp0));
        voff[c] = int(o);
    }
    int pending = 0;  // prefetched words of the previous block, folded into `sink` one block later
    // the first two 64-degree runs now, then the run 128 degrees ahead at every block
    for (int pb = 0; pb < 2; ++pb) {
        const int lp = l + 64 * pb;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const bool on = used[c] && lp + lstart[c] <= le;
            pending ^= l2_touch(isab[c] ? reinterpret_cast<const char *>(abm + lp) : reinterpret_cast<const char *>(cp0 + lp),
                                voff[c], on);
        }
    }
    while (l <= le) {
        const int lend = min(le, block_end(m, l));
        {
            sink ^= pending;  // the previous block's prefetch has long completed
            pending = 0;
            const int lp = l + kPrefetchAhead;
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const bool on = used[c] && lp + lstart[c] <= le;
                pending ^= l2_touch(isab[c] ? reinterpret_cast<const char *>(abm + lp) : reinterpret_cast<const char *>(cp0 + lp),
                                    voff[c], on);
            }
        }
#pragma unroll 4
        for (; l <= lend; ++l) {
            const double q = rec_step(abm[l], x, q1, q2);
            q2 = q1;
            q1 = q;
#pragma unroll
            for (int t = 0; t < NB; ++t) {
                const double2 a = cp0[t * LB + l];
                ar[t] = fma(a.x, q, ar[t]);
                ai[t] = fma(a.y, q, ai[t]);
                if (NEG) {
                    const double2 d = cm0[t * LB + l];
                    br[t] = fma(d.x, q, br[t]);
                    bi[t] = fma(d.y, q, bi[t]);
                }
            }
        }
        if (l <= le) {  // renormalise at the block end (not after the segment's last degree)
            const int s = renorm_shift(q1, q2);
            q1 = ldexp(q1, -s);
            q2 = ldexp(q2, -s);
#pragma unroll
            for (int t = 0; t < NB; ++t) {
                ar[t] = ldexp(ar[t], -s);
                ai[t] = ldexp(ai[t], -s);
                if (NEG) {
                    br[t] = ldexp(br[t], -s);
                    bi[t] = ldexp(bi[t], -s);
                }
            }
            E += s;
        }
    }
    sink ^= pending;
}

// Single-transform sweep in the rescaled form. Degrees run in groups of 16 aligned at m (group 0
// of segment 0 starts at m + 2). Within a group r_l = q_l / C_l with C_l the product of a_k over
// the group's degrees up to l, so that
//     r_l = x r_{l-1} - gamma_l r_{l-2},   gamma_l = b_l / (a_l a_{l-1}),
// and the accumulators take a_lm C_l r_l = a_lm q_l. The plan holds (gamma_l, C_l) per (m, l).
// Lane k of every 16-lane row holds degree g + k of the group (its gamma, C and a_lm), loaded with
// vector loads one group ahead; the FMAs read them through DPP row broadcasts (row_newbcast:k).
// At the group end (r_{l-1}, r_{l-2}) times C of the last degree gives the next group's start
// (q_{l-1}, a_{l-1} q_{l-2}); renormalisation follows at 64-degree block ends as in the scalar
// sweep. 6 FP64 FMAs per degree and ring, no scalar loads.
template <int K, bool NEG>
__device__ __forceinline__ double dpp_step(double x, double r1, double r2, double ng, const double (&c)[4],
                                           double (&acc)[4]) {
    double t;
    // s_nop 1: a DPP source must not be written by VALU in the two instructions before its use
    if (NEG)
        asm("s_nop 1\n\t"
            "v_mul_f64 %0, %5, %6\n\t"
            "v_fmac_f64_dpp %0, %7, %8 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %1, %10, %0 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %2, %11, %0 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %3, %12, %0 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %4, %13, %0 row_newbcast:%9 row_mask:0xf bank_mask:0xf"
            : "=&v"(t), "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3])
            : "v"(x), "v"(r1), "v"(ng), "v"(r2), "n"(K), "v"(c[0]), "v"(c[1]), "v"(c[2]), "v"(c[3]));
    else
        asm("s_nop 1\n\t"
            "v_mul_f64 %0, %3, %4\n\t"
            "v_fmac_f64_dpp %0, %5, %6 row_newbcast:%7 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %1, %8, %0 row_newbcast:%7 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %2, %9, %0 row_newbcast:%7 row_mask:0xf bank_mask:0xf"
            : "=&v"(t), "+v"(acc[0]), "+v"(acc[1])
            : "v"(x), "v"(r1), "v"(ng), "v"(r2), "n"(K), "v"(c[0]), "v"(c[1]));
    return t;
}

// Steps K .. 15 of a group; GUARD: only degrees g + K with kfirst <= K < kend.
template <int K, bool GUARD, bool NEG>
struct DppGroup {
    static __device__ __forceinline__ void run(double x, double &r1, double &r2, double ng, const double (&c)[4],
                                               double (&acc)[4], int kfirst, int kend) {
        if (!GUARD || (K >= kfirst && K < kend)) {
            const double t = dpp_step<K, NEG>(x, r1, r2, ng, c, acc);
            r2 = r1;
            r1 = t;
        }
        DppGroup<K + 1, GUARD, NEG>::run(x, r1, r2, ng, c, acc, kfirst, kend);
    }
};
template <bool GUARD, bool NEG>
struct DppGroup<16, GUARD, NEG> {
    static __device__ __forceinline__ void run(double, double &, double &, double, const double (&)[4], double (&)[4],
                                               int, int) {}
};

// Degrees l .. le (l = m + 2, or a 64-degree block start) from (q1, q2) = (q_{l-1}, q_{l-2}) with
// exponent E; aprev = a_{l-1} (sqrt(2m + 3) for l = m + 2). gcm[l] = (gamma_l, C_l).
template <bool NEG>
__device__ __forceinline__ void fwd_sweep_dpp(const double2 *__restrict__ gcm, const double2 *__restrict__ cp0,
                                              const double2 *__restrict__ cm0, double x, double q1, double q2,
                                              double aprev, int &E, int l, int le, int m, int L,
                                              double (&acc)[4], int &sink) {
    const int lane = threadIdx.x & 63, k = lane & 15;
    double r1 = q1, r2 = q2 * aprev;
    int g = m + ((l - m) & ~15);  // group start
    int kfirst = l - g;
    // Every load is unconditional, at an index clamped into the order's tables, so that the
    // compiler can count them and wait only for the ones in use: a lane outside [kfirst, kend) of
    // a group loads a neighbouring degree and is never broadcast.
    auto group_load = [&](int gs, double2 &gcv, double2 &ap, double2 &am) {
        const int lc = min(max(gs + k, m + 2), L);
        gcv = gcm[lc];
        ap = cp0[lc];
        if (NEG) am = cm0[lc];
    };
    // L2 prefetch of the run two 64-degree blocks ahead, at every block start: lane i touches line
    // i % 8 of stream (i / 8) % 3 (gamma/C, a_lm of +m, of -m); the word read is folded into `sink`
    // one block later, so its wait never stalls the sweep
    const int pst = (lane >> 3) % (NEG ? 3 : 2), pline = lane & 7;
    const char *pbase = pst == 0 ? reinterpret_cast<const char *>(gcm)
                                 : reinterpret_cast<const char *>(pst == 1 ? cp0 : cm0);
    auto touch = [&](int lp) {
        const int lt = min(max(lp + 8 * pline, m + 2), L);
        return *reinterpret_cast<const int *>(pbase + int64_t(lt) * 16);
    };
    int pa = 0, pb = touch(l + 64);
    double2 gcv, ap, am = make_double2(0.0, 0.0);
    group_load(g, gcv, ap, am);
    while (g <= le) {
        const int gn = g + 16;
        double2 gcn, apn, amn = make_double2(0.0, 0.0);
        group_load(gn, gcn, apn, amn);
        if (((g - m) & 63) == 0) {
            sink ^= pa;
            pa = pb;
            pb = touch(g + kPrefetchAhead);
        }
        const double C = gcv.y, ng = -gcv.x;
        double c[4] = {ap.x * C, ap.y * C, am.x * C, am.y * C};
        const int kend = min(16, le - g + 1);
        if (kend == 16 && kfirst == 0) DppGroup<0, false, NEG>::run(x, r1, r2, ng, c, acc, 0, 16);
        else if (kend == 16 && kfirst == 2) DppGroup<2, false, NEG>::run(x, r1, r2, ng, c, acc, 2, 16);
        else DppGroup<0, true, NEG>::run(x, r1, r2, ng, c, acc, kfirst, kend);
        // back to the q scale at the group end: C of the group's last degree, from lane kend - 1
        const int last = kend - 1;
        const double Ce = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(C), last),
                                           __builtin_amdgcn_readlane(__double2loint(C), last));
        r1 *= Ce;
        r2 *= Ce;
        g = gn;
        kfirst = 0;
        if (g <= le && ((g - m) & 63) == 0) {  // renormalise at the block end
            const int s = renorm_shift(r1, r2);
            r1 = ldexp(r1, -s);
            r2 = ldexp(r2, -s);
#pragma unroll
            for (int u = 0; u < 4; ++u) acc[u] = ldexp(acc[u], -s);
            E += s;
        }
        gcv = gcn;
        ap = apn;
        am = amn;
    }
    sink ^= pa ^ pb;
}

// One workgroup per 64-ring tile of one order; wavefront s sweeps degree segment s. Segment 0
// starts from the seed, the others from their checkpoints; the segment sums, scaled to plain
// doubles, are added in segment order through LDS and wavefront 0 stores the tile's ring values.
template <int NB>
__device__ __forceinline__ void ring_fwd_body(const SegItem *__restrict__ items, const double *__restrict__ costh,
                const double *__restrict__ log2sin, const double *__restrict__ log2c,
                const double2 *__restrict__ ab, const double2 *__restrict__ gc, const int64_t *__restrict__ abofs,
                const double2 *__restrict__ ckq,
                const int32_t *__restrict__ ckE, const double2 *__restrict__ alm, double2 *__restrict__ Gd, int L,
                int n, int mmax, int b0, int never) {
    __shared__ double red[kSegMax - 1][4][64];
    const SegItem it = items[blockIdx.x];
    const int m = __builtin_amdgcn_readfirstlane(it.m);
    const int nseg = __builtin_amdgcn_readfirstlane(it.nseg);
    const int nr = __builtin_amdgcn_readfirstlane(it.nr);
    if (nr <= 0) return;  // an empty slot: the whole workgroup leaves
    const int sg = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);  // this wavefront's segment
    const int lane = threadIdx.x & 63;
    const bool active = lane < nr;
    const int j = it.j0 + (active ? lane : 0);

    const int64_t LB = int64_t(L + 1) * (2 * L + 1);
    // column of order m (row l - m holds degree l) and of order -m of transform b0; transform
    // b0 + t is LB further on
    const double2 *__restrict__ cp0 = alm + b0 * LB + int64_t(2 * m) * (L + 1) - m;
    const double2 *__restrict__ cm0 = alm + b0 * LB + int64_t(2 * m - 1) * (L + 1) - m;
    const double2 *__restrict__ abm = ab + abofs[m] - (m + 2);  // abm[l] for l = m+2..L

    double ar[NB], ai[NB], br[NB], bi[NB];
    int sink = 0;
    if (sg < nseg) {
        const double x = costh[j];
        double q1, q2;
        int E, l;
        if (sg == 0) {
            seed(m == 0 ? log2c[0] : double(m) * log2sin[j] + log2c[m], q2, E);
            if (!active) q2 = 0.0;
            q1 = m < L ? first_step(m, x, q2) : 0.0;
#pragma unroll
            for (int t = 0; t < NB; ++t) {
                const double2 a0 = cp0[t * LB + m];
                ar[t] = a0.x * q2;
                ai[t] = a0.y * q2;
                br[t] = bi[t] = 0.0;
                if (m > 0) {
                    const double2 c0 = cm0[t * LB + m];
                    br[t] = c0.x * q2;
                    bi[t] = c0.y * q2;
                }
                if (m < L) {
                    const double2 a1 = cp0[t * LB + m + 1];
                    ar[t] = fma(a1.x, q1, ar[t]);
                    ai[t] = fma(a1.y, q1, ai[t]);
                    if (m > 0) {
                        const double2 c1 = cm0[t * LB + m + 1];
                        br[t] = fma(c1.x, q1, br[t]);
                        bi[t] = fma(c1.y, q1, bi[t]);
                    }
                }
            }
            l = m + 2;
        } else {
            const int64_t row = (it.ck + sg - 1) * 64 + lane;
            const double2 c = ckq[row];
            q1 = c.x;
            q2 = c.y;
            E = ckE[row];
#pragma unroll
            for (int t = 0; t < NB; ++t) ar[t] = ai[t] = br[t] = bi[t] = 0.0;
            l = seg_begin(m, L, nseg, sg);
        }
        const int le = min(L, seg_begin(m, L, nseg, sg + 1) - 1);
        if (NB == 1) {
            if (l <= le) {
                const double2 *__restrict__ gcm = gc + abofs[m] - (m + 2);
                const double aprev = l == m + 2 ? sqrt(double(2 * m + 3)) : abm[l - 1].x;
                double acc[4] = {ar[0], ai[0], br[0], bi[0]};
                if (m > 0) fwd_sweep_dpp<true>(gcm, cp0, cm0, x, q1, q2, aprev, E, l, le, m, L, acc, sink);
                else fwd_sweep_dpp<false>(gcm, cp0, cm0, x, q1, q2, aprev, E, l, le, m, L, acc, sink);
                ar[0] = acc[0];
                ai[0] = acc[1];
                br[0] = acc[2];
                bi[0] = acc[3];
            }
        } else if (m > 0) {
            fwd_sweep<NB, true>(abm, cp0, cm0, LB, x, q1, q2, E, l, le, m, ar, ai, br, bi, sink);
        } else {
            fwd_sweep<NB, false>(abm, cp0, cm0, LB, x, q1, q2, E, l, le, m, ar, ai, br, bi, sink);
        }
#pragma unroll
        for (int t = 0; t < NB; ++t) {
            ar[t] = ldexp(ar[t], E);
            ai[t] = ldexp(ai[t], E);
            br[t] = ldexp(br[t], E);
            bi[t] = ldexp(bi[t], E);
        }
    }
    if (sink == never) Gd[j] = make_double2(0.0, 0.0);  // keeps the prefetch loads; never true
    if (nseg > 1) {
#pragma unroll
        for (int t = 0; t < NB; ++t) {
            if (sg > 0 && sg < nseg) {
                red[sg - 1][0][lane] = ar[t];
                red[sg - 1][1][lane] = ai[t];
                red[sg - 1][2][lane] = br[t];
                red[sg - 1][3][lane] = bi[t];
            }
            __syncthreads();
            if (sg == 0)
                for (int k = 0; k < nseg - 1; ++k) {
                    ar[t] += red[k][0][lane];
                    ai[t] += red[k][1][lane];
                    br[t] += red[k][2][lane];
                    bi[t] += red[k][3][lane];
                }
            __syncthreads();
        }
    }
    if (sg != 0 || !active) return;
    const double sgn = (m & 1) ? -1.0 : 1.0;
    const int64_t ncol = 2 * mmax + 1;
#pragma unroll
    for (int t = 0; t < NB; ++t) {
        double2 *gp = Gd + (int64_t(b0 + t) * ncol + (mmax + m)) * n + j;
        *gp = make_double2(sgn * ar[t], sgn * ai[t]);
        if (m > 0) {
            double2 *gm = Gd + (int64_t(b0 + t) * ncol + (mmax - m)) * n + j;
            *gm = make_double2(br[t], bi[t]);
        }
    }
}

// The single-transform sweep holds the DPP operands of two groups; six wavefronts per SIMD fit
// its registers.
__global__ void __launch_bounds__(64 * kSegMax) __attribute__((amdgpu_waves_per_eu(6)))
ring_fwd1_kernel(const SegItem *__restrict__ items, const double *__restrict__ costh,
                 const double *__restrict__ log2sin, const double *__restrict__ log2c,
                 const double2 *__restrict__ ab, const double2 *__restrict__ gc, const int64_t *__restrict__ abofs,
                 const double2 *__restrict__ ckq, const int32_t *__restrict__ ckE, const double2 *__restrict__ alm,
                 double2 *__restrict__ Gd, int L, int n, int mmax, int b0, int never) {
    ring_fwd_body<1>(items, costh, log2sin, log2c, ab, gc, abofs, ckq, ckE, alm, Gd, L, n, mmax, b0, never);
}

template <int NB>
__global__ void __launch_bounds__(64 * kSegMax)
ring_fwd_kernel(const SegItem *__restrict__ items, const double *__restrict__ costh,
                const double *__restrict__ log2sin, const double *__restrict__ log2c,
                const double2 *__restrict__ ab, const double2 *__restrict__ gc, const int64_t *__restrict__ abofs,
                const double2 *__restrict__ ckq, const int32_t *__restrict__ ckE, const double2 *__restrict__ alm,
                double2 *__restrict__ Gd, int L, int n, int mmax, int b0, int never) {
    ring_fwd_body<NB>(items, costh, log2sin, log2c, ab, gc, abofs, ckq, ckE, alm, Gd, L, n, mmax, b0, never);
}

// Transform groups of 8, 4, 2 and 1 cover any ntrans exactly with at most four launches.
template <class F>
static void for_transform_groups(int B, F &&launch) {
    int b0 = 0;
    for (int nb : {8, 4, 2, 1})
        while (B - b0 >= nb) {
            launch(nb, b0);
            b0 += nb;
        }
}

void launch_ring_fwd(const RingDev &r, const double2 *alm, double2 *Gd, hipStream_t s) {
    if (r.nfitems == 0) return;
    for_transform_groups(r.B, [&](int nb, int b0) {
        dim3 grid(r.nfitems);
#define PFSHT_FWD(NB)                                                                                         \
    ring_fwd_kernel<NB><<<grid, 64 * r.fwd_waves, 0, s>>>(r.fitems, r.costh, r.log2sin, r.log2c, r.ab, r.gc, r.abofs, \
                                                          r.ckq, r.ckE, alm, Gd, r.L, r.n, r.mmax, b0, kNeverSink)
        switch (nb) {
            case 8: PFSHT_FWD(8); break;
            case 4: PFSHT_FWD(4); break;
            case 2: PFSHT_FWD(2); break;
            default:
                ring_fwd1_kernel<<<grid, 64 * r.fwd_waves, 0, s>>>(r.fitems, r.costh, r.log2sin, r.log2c, r.ab, r.gc,
                                                                  r.abofs, r.ckq, r.ckE, alm, Gd, r.L, r.n, r.mmax, b0,
                                                                  kNeverSink);
                break;
        }
#undef PFSHT_FWD
    });
}

// ------------------------------------------------------------------------------------------------
// Adjoint
// ------------------------------------------------------------------------------------------------
// alm[l, m] = sum_j Pbar_l^m(x_j) G[j, m] reduces across the rings a wavefront holds in its lanes.
// The wavefront walks the degrees in sub-tiles of 16:
//   phase 1 (lanes = rings) advances the recurrence 16 degrees and stores w = q 2^E to a 16 x 64
//           LDS tile;
//   phase 2 (lane = 16 row + t) gives each 16-lane row 16 of the tile's rings: lane t of row r
//           sums w[t][16 r + k] G[16 r + k] over k, the G operand taken from lane k of its own row
//           by a DPP row broadcast on the FMA (G stays in the registers it was loaded into);
//   the four row sums are then added across rows ((r0 + r2) + (r1 + r3)) with lane swaps.
// The tile is 8.6 KB per wavefront, so four wavefronts share a SIMD and hide each other's LDS and
// scalar-load latency. Sums have a fixed order: bit-identical run to run.
// Phase-1 step K of the rescaled recurrence (as in the forward): t = x r1 - gamma_K r2, with gamma_K
// broadcast from lane K of the row.
template <int K>
__device__ __forceinline__ double dpp_rstep(double x, double r1, double r2, double ng) {
    double t;
    asm("s_nop 1\n\t"
        "v_mul_f64 %0, %1, %2\n\t"
        "v_fmac_f64_dpp %0, %3, %4 row_newbcast:%5 row_mask:0xf bank_mask:0xf"
        : "=&v"(t)
        : "v"(x), "v"(r1), "v"(ng), "v"(r2), "n"(K));
    return t;
}

// Steps K .. 15 of a sub-tile, each storing r (SCALED: r 2^E) to row K of the LDS tile; GUARD:
// only kfirst <= K < kend.
template <int K, bool GUARD, bool SCALED>
struct AdjPhase1 {
    static __device__ __forceinline__ void run(double x, double &r1, double &r2, double ng, int E, double (*W)[kWS],
                                               int wcol, int kfirst, int kend) {
        if (!GUARD || (K >= kfirst && K < kend)) {
            const double t = dpp_rstep<K>(x, r1, r2, ng);
            r2 = r1;
            r1 = t;
            W[K][wcol] = SCALED ? ldexp(t, E) : t;
        }
        AdjPhase1<K + 1, GUARD, SCALED>::run(x, r1, r2, ng, E, W, wcol, kfirst, kend);
    }
};
template <bool GUARD, bool SCALED>
struct AdjPhase1<kSub, GUARD, SCALED> {
    static __device__ __forceinline__ void run(double, double &, double &, double, int, double (*)[kWS], int, int,
                                               int) {}
};

// Per-lane state of one 64-ring tile in the adjoint sweep.
template <int NB>
struct AdjTile {
    static constexpr bool kScaleG = NB <= 2;
    static constexpr int NG = kScaleG ? NB : 1;
    double x, r1, r2;   // cos theta; the recurrence pair
    double w0, w1;      // segment 0: the seed degrees m and m + 1 as stored in the tile
    int E;              // the ring's exponent
    double g[NB][4];    // (re, im) of orders +m and -m per transform
    double gs[NG][4];   // G 2^E
    __device__ __forceinline__ void scale_g() {
        if (kScaleG)
#pragma unroll
            for (int u = 0; u < NG; ++u)
#pragma unroll
                for (int c = 0; c < 4; ++c) gs[u][c] = ldexp(g[u][c], E);
    }
};

// Degrees ls .. le of order m for TG 64-ring tiles of that order (the second may be empty). Segment 0
// (ls = m) starts from the seed; a later segment from its checkpoint (q1, q2) = (q_{ls-1}, q_{ls-2}),
// aprev = a_{ls-1}. ls - m is a multiple of 64. Phase 1 runs the rescaled recurrence of the forward
// sweep over each 16-degree sub-tile (gamma_l and C_l of lane k of a row give degree l0 + k); phase
// 2 adds the tile's ring sums into the row accumulators; the tiles take turns in the one LDS tile,
// and the accumulators, their sum over the four rows and the partial stores are shared. The ring
// sums of degree l are multiplied by C_l once, after the reduction. For one or two transforms the
// ring's exponent goes into its G operands (G 2^E, exact, refreshed when E changes at a block end)
// and the tile holds r; with more transforms (registers) the tile holds r 2^E.
template <int NB, bool NEG, int TG>
__device__ __forceinline__ void adj_sweep(const double2 *__restrict__ gcm, AdjTile<NB> (&T)[TG], bool two,
                                          double aprev, int ls, int le, int m, int L, int B, double (*W)[kWS],
                                          double4 *__restrict__ out, int &sink, const double (&qs)[TG],
                                          const double (&q1)[TG], const double (&q2)[TG]) {
    constexpr bool kScaleG = AdjTile<NB>::kScaleG;
    const int lane = threadIdx.x & 63;
    const int t = lane & 15, row = lane >> 4;
    const int wcol = lane + (lane >> 4);
    const double *__restrict__ wr = &W[t][17 * row];
#pragma unroll
    for (int k = 0; k < TG; ++k) T[k].scale_g();

    // (gamma, C) lines of the first two 64-degree blocks, then two blocks ahead of the sweep: lane
    // i < 16 owns the line at byte offset 128 i of a 128-degree window
    const int pvoff = 128 * (lane & 15);
    const int lp0 = ls == m ? m + 2 : ls;
    int pending = l2_touch(reinterpret_cast<const char *>(gcm + lp0), pvoff, lane < 16 && lp0 + 8 * lane <= le);
    // (gamma, C) of degree l0 + t, loaded at a clamped index one sub-tile ahead; orders m >= L - 1
    // have no table (and no recurrence step)
    const bool has_tab = m + 2 <= L;
    auto coef = [&](int l0) { return has_tab ? gcm[min(max(l0 + t, m + 2), L)] : make_double2(0.0, 1.0); };
    double2 gcv = coef(ls);
    int kfirst = 0;
#pragma unroll
    for (int k = 0; k < TG; ++k) {
        AdjTile<NB> &A = T[k];
        if (ls == m) {  // seed degrees m and m + 1, stored as q (q 2^E)
            A.w0 = kScaleG ? qs[k] : ldexp(qs[k], A.E);
            kfirst = 1;
            A.r1 = qs[k];
            A.r2 = 0.0;
            A.w1 = 0.0;
            if (m < L) {
                const double q = first_step(m, A.x, qs[k]);
                A.w1 = kScaleG ? q : ldexp(q, A.E);
                kfirst = 2;
                A.r1 = q;
                A.r2 = qs[k] * sqrt(double(2 * m + 3));
            }
        } else {
            A.r1 = q1[k];
            A.r2 = q2[k] * aprev;
        }
    }
    for (int l0 = ls; l0 <= le; l0 += kSub) {
        const int tl = min(kSub, le - l0 + 1);
        if (((l0 - m) & 63) == 0) {
            sink ^= pending;
            const int lp = l0 + 128;
            pending = l2_touch(reinterpret_cast<const char *>(gcm + lp), pvoff, lane < 8 && lp + 8 * lane <= le);
        }
        const double2 gcn = coef(l0 + kSub);
        const double ng = -gcv.x, Cc = gcv.y;
        const double cmul = t < kfirst ? 1.0 : Cc;  // the seed degrees are stored unscaled
        // back to the q scale after the sub-tile: C of its last degree (the seed-only sub-tile ends
        // the sweep)
        const int last = tl - 1;
        const double Ce = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(Cc), last),
                                           __builtin_amdgcn_readlane(__double2loint(Cc), last));
        const bool block_end = ((l0 + kSub - m) & 63) == 0;  // renormalise at every 64-degree block end
        double a[NB][4];
#pragma unroll
        for (int u = 0; u < NB; ++u) a[u][0] = a[u][1] = a[u][2] = a[u][3] = 0.0;
#pragma unroll
        for (int k = 0; k < TG; ++k) {
            if (k > 0 && !two) break;  // wave-uniform: the group has one tile
            AdjTile<NB> &A = T[k];
            // phase 1: lanes = rings
            if (kfirst > 0) {
                W[0][wcol] = A.w0;
                if (kfirst == 2) W[1][wcol] = A.w1;
            }
            if (tl == kSub && kfirst == 0) AdjPhase1<0, false, !kScaleG>::run(A.x, A.r1, A.r2, ng, A.E, W, wcol, 0, kSub);
            else if (tl == kSub && kfirst == 2) AdjPhase1<2, false, !kScaleG>::run(A.x, A.r1, A.r2, ng, A.E, W, wcol, 2, kSub);
            else AdjPhase1<0, true, !kScaleG>::run(A.x, A.r1, A.r2, ng, A.E, W, wcol, kfirst, tl);
            A.r1 *= Ce;
            A.r2 *= Ce;
            if (block_end) {
                const int sh = renorm_shift(A.r1, A.r2);
                A.r1 = ldexp(A.r1, -sh);
                A.r2 = ldexp(A.r2, -sh);
                A.E += sh;
            }
            wave_lds_order();
            // phase 2: lane t of row r adds rings 16 r .. 16 r + 15 of degree l0 + t
            if constexpr (kScaleG) RowSweep<0, NB, NEG>::run(wr, a, A.gs);
            else RowSweep<0, NB, NEG>::run(wr, a, A.g);
            if (block_end) A.scale_g();  // the next sub-tile's r carry the new exponent
            wave_lds_order();
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            a[u][0] = row_sum4(a[u][0]);
            a[u][1] = row_sum4(a[u][1]);
            if (NEG) {
                a[u][2] = row_sum4(a[u][2]);
                a[u][3] = row_sum4(a[u][3]);
            }
        }
        {
            // every lane stores: lane t of row r writes component r of degree l0 + t (the row sums
            // are in every row). Degrees past the order's last (t >= tl, last sub-tile only) land in
            // the group's padding. No branch around the stores, so the wait for the next sub-tile's
            // coefficients does not have to cover them.
            double *__restrict__ od = reinterpret_cast<double *>(out);
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                const double c01 = row == 0 ? a[u][0] : a[u][1];
                const double c23 = NEG ? (row == 2 ? a[u][2] : a[u][3]) : 0.0;
                od[(int64_t(l0 - m + t) * B + u) * 4 + row] = (row < 2 ? c01 : c23) * cmul;
            }
        }
        gcv = gcn;
        kfirst = 0;
    }
    sink ^= pending;
}

// Workgroups of kAdjWaves wavefronts; each wavefront is one item: one degree segment of TG
// consecutive 64-ring tiles of one order (the second may be empty), writing that tile group's
// partials for the segment's degrees.
template <int NB, int TG>
__global__ void __launch_bounds__(64 * kAdjWaves)
ring_adj_kernel(const SegItem *__restrict__ items, const double *__restrict__ costh,
                const double *__restrict__ log2sin, const double *__restrict__ log2c,
                const double2 *__restrict__ ab, const double2 *__restrict__ gc, const int64_t *__restrict__ abofs,
                const double2 *__restrict__ ckq,
                const int32_t *__restrict__ ckE, const int64_t *__restrict__ pofs, const double2 *__restrict__ Gd,
                double4 *__restrict__ partials, int L, int n, int mmax, int B, int b0, int never) {
    extern __shared__ double lds_adj[];
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
    const SegItem it = items[blockIdx.x * kAdjWaves + wave];
    const int m = __builtin_amdgcn_readfirstlane(it.m);
    const int lane = threadIdx.x & 63;
    const int nr = __builtin_amdgcn_readfirstlane(it.nr);
    if (nr <= 0) return;  // an empty slot (wave-uniform): the whole wavefront leaves
    const int nr2 = TG > 1 ? __builtin_amdgcn_readfirstlane(it.nr2) : 0;
    const int nseg = __builtin_amdgcn_readfirstlane(it.nseg);
    const int sg = __builtin_amdgcn_readfirstlane(it.seg);
    const int group = __builtin_amdgcn_readfirstlane(it.tile);
    double(*W)[kWS] = reinterpret_cast<double(*)[kWS]>(lds_adj + wave * kSub * kWS);
    const int ls = seg_begin(m, L, nseg, sg);
    const int le = min(L, seg_begin(m, L, nseg, sg + 1) - 1);
    const int64_t ncol = 2 * mmax + 1;
    const int64_t gstride = ncol * n;  // next transform

    AdjTile<NB> T[TG];
    double qs[TG], q1[TG], q2[TG];
#pragma unroll
    for (int k = 0; k < TG; ++k) {
        const int nrk = k == 0 ? nr : nr2;
        const int jt = it.j0 + 64 * k;  // first ring of the tile
        const bool active = lane < nrk;
        const int j = active ? jt + lane : it.j0;
        AdjTile<NB> &A = T[k];
        A.x = costh[j];
        qs[k] = q1[k] = q2[k] = 0.0;
        if (sg == 0) {
            seed(m == 0 ? log2c[0] : double(m) * log2sin[j] + log2c[m], qs[k], A.E);
            if (!active) qs[k] = 0.0;
        } else {
            // tile k's checkpoint rows follow tile 0's: one row per segment >= 1 per tile
            const int64_t rowk = (it.ck + int64_t(k) * (nseg - 1) + sg - 1) * 64 + lane;
            const double2 c = (k == 0 || nrk > 0) ? ckq[rowk] : make_double2(0.0, 0.0);
            q1[k] = active ? c.x : 0.0;
            q2[k] = active ? c.y : 0.0;
            A.E = (k == 0 || nrk > 0) ? ckE[rowk] : 0;
        }
        const double2 *gpc = Gd + (int64_t(b0) * ncol + (mmax + m)) * n + jt;
        const double2 *gmc = Gd + (int64_t(b0) * ncol + (mmax - m)) * n + jt;
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const double2 p = active ? gpc[u * gstride + lane] : make_double2(0.0, 0.0);
            const double2 q = (m > 0 && active) ? gmc[u * gstride + lane] : make_double2(0.0, 0.0);
            A.g[u][0] = p.x;
            A.g[u][1] = p.y;
            A.g[u][2] = q.x;
            A.g[u][3] = q.y;
        }
    }
    const double2 *__restrict__ abm = ab + abofs[m] - (m + 2);
    const int nl = adj_rows_padded(m, L);
    double4 *__restrict__ out = partials + pofs[m] + (int64_t(group) * nl) * B + b0;
    int sink = 0;
    const double2 *__restrict__ gcm = gc + abofs[m] - (m + 2);
    const double aprev = sg == 0 ? 0.0 : abm[ls - 1].x;
    const bool two = nr2 > 0;
    if (m > 0) adj_sweep<NB, true, TG>(gcm, T, two, aprev, ls, le, m, L, B, W, out, sink, qs, q1, q2);
    else adj_sweep<NB, false, TG>(gcm, T, two, aprev, ls, le, m, L, B, W, out, sink, qs, q1, q2);
    if (sink == never) out[0] = make_double4(0.0, 0.0, 0.0, 0.0);  // never true; keeps the prefetch
}

// Writes every alm slot: the tile-ordered sum of the partials where (l, m) is retained, zero in
// the padding and for unretained orders (the adjoint overwrites alm, src/Legendre.jl:338). One
// thread owns row l - |m| of the column pair (2|m|, 2|m| - 1), reading each partial once.
__global__ void ring_adj_finalize(const int64_t *__restrict__ pofs, const int32_t *__restrict__ ntiles,
                                  const double4 *__restrict__ partials, double2 *__restrict__ alm, int L,
                                  int mmax, int B) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    const int am = blockIdx.y;
    const int b = blockIdx.z;
    if (row > L) return;
    const int64_t nrow = L + 1;
    const int64_t base = int64_t(b) * nrow * (2 * L + 1);
    const int l = row + am;
    double pr = 0, pi = 0, nr_ = 0, ni = 0;
    if (l <= L && am <= mmax) {
        const int nt = ntiles[am];
        const int nl = adj_rows_padded(am, L);
        const double4 *p = partials + pofs[am] + int64_t(l - am) * B + b;
        for (int t = 0; t < nt; ++t) {
            const double4 q = p[int64_t(t) * nl * B];
            pr += q.x;
            pi += q.y;
            nr_ += q.z;
            ni += q.w;
        }
        if (am & 1) { pr = -pr; pi = -pi; }
    }
    alm[base + int64_t(2 * am) * nrow + row] = make_double2(pr, pi);
    if (am > 0) alm[base + int64_t(2 * am - 1) * nrow + row] = make_double2(nr_, ni);
}

void launch_ring_adj(const RingDev &r, const double2 *Gd, double2 *alm, hipStream_t s) {
    launch_ring_adj_tiles(r, Gd, s);
    launch_ring_adj_finalize(r, alm, s);
}

void launch_ring_adj_tiles(const RingDev &r, const double2 *Gd, hipStream_t s) {
    if (r.naitems > 0)
        for_transform_groups(r.B, [&](int nb, int b0) {
            dim3 grid(r.naitems / kAdjWaves);
            const size_t lds = sizeof(double) * kSub * kWS * kAdjWaves;
#define PFSHT_ADJ(NB, TG)                                                                                    \
    ring_adj_kernel<NB, TG><<<grid, 64 * kAdjWaves, lds, s>>>(r.aitems, r.costh, r.log2sin, r.log2c, r.ab, r.gc,  \
                                                              r.abofs, r.ckq, r.ckE, r.pofs, Gd, r.partials, r.L, r.n, \
                                                              r.mmax, r.B, b0, kNeverSink)
            // the item list was built for r.adj_tg tiles per wavefront (two when ntrans <= 2)
            switch (nb) {
                case 8: PFSHT_ADJ(8, 1); break;
                case 4: PFSHT_ADJ(4, 1); break;
                case 2:
                    if (r.adj_tg == 2) PFSHT_ADJ(2, 2);
                    else PFSHT_ADJ(2, 1);
                    break;
                default:
                    if (r.adj_tg == 2) PFSHT_ADJ(1, 2);
                    else PFSHT_ADJ(1, 1);
                    break;
            }
#undef PFSHT_ADJ
        });
}

void launch_ring_adj_finalize(const RingDev &r, double2 *alm, hipStream_t s) {
    dim3 fgrid((r.L + 1 + 255) / 256, r.L + 1, r.B);
    ring_adj_finalize<<<fgrid, 256, 0, s>>>(r.pofs, r.ntiles, r.partials, alm, r.L, r.mmax, r.B);
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
