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
#include <type_traits>

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
__device__ __forceinline__ void fwd_sweep_dpp(const double2 *__restrict__ gcm, gcptr __restrict__ cp0,
                                              gcptr __restrict__ cm0, double x, double q1, double q2,
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
        ap = gload(cp0 + lc);
        if (NEG) am = gload(cm0 + lc);
    };
    // L2 prefetch of the run two 64-degree blocks ahead, at every block start: lane i touches line
    // i % 8 of stream (i / 8) % 3 (gamma/C, a_lm of +m, of -m); the word read is folded into `sink`
    // one block later, so its wait never stalls the sweep
    const int pst = (lane >> 3) % (NEG ? 3 : 2), pline = lane & 7;
    const PFSHT_GLOBAL char *pbase = pst == 0 ? (const PFSHT_GLOBAL char *)gcm
                                              : (const PFSHT_GLOBAL char *)(pst == 1 ? cp0 : cm0);
    auto touch = [&](int lp) {
        const int lt = min(max(lp + 8 * pline, m + 2), L);
        return *(const PFSHT_GLOBAL int *)(pbase + int64_t(lt) * 16);
    };
    int pa = 0, pb = touch(l + 64);
    // The coefficient lines of a group (gamma_l, C_l and a_lm of +m and -m), loaded one group ahead
    // into the other of two buffers.
    struct Lines {
        double2 gcv, ap, am;
    };
    Lines A, B;
    A.am = B.am = make_double2(0.0, 0.0);
    group_load(g, A.gcv, A.ap, A.am);
    // One group of degrees g .. g + 15 on the lines in `cur`: the next group's loads into `nxt`, the L2
    // touch at a block start, the steps, the rescale to the q scale and the renormalisation at a
    // block end. MODE 0: a full group; 1: a full group from step 2 (the first group of segment 0);
    // 2: steps kfirst .. kend - 1 only.
    auto group = [&](auto mode, const Lines &cur, Lines &nxt) {
        constexpr int MODE = decltype(mode)::value;
        group_load(g + 16, nxt.gcv, nxt.ap, nxt.am);
        if (((g - m) & 63) == 0) {
            sink ^= pa;
            pa = pb;
            pb = touch(g + kPrefetchAhead);
        }
        const double C = cur.gcv.y, ng = -cur.gcv.x;
        double c[4] = {cur.ap.x * C, cur.ap.y * C, cur.am.x * C, cur.am.y * C};
        const int kend = MODE == 2 ? min(16, le - g + 1) : 16;
        if constexpr (MODE == 0) DppGroup<0, false, NEG>::run(x, r1, r2, ng, c, acc, 0, 16);
        else if constexpr (MODE == 1) DppGroup<2, false, NEG>::run(x, r1, r2, ng, c, acc, 2, 16);
        else DppGroup<0, true, NEG>::run(x, r1, r2, ng, c, acc, kfirst, kend);
        // back to the q scale at the group end: C of the group's last degree, from lane kend - 1
        const int last = kend - 1;
        const double Ce = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(C), last),
                                           __builtin_amdgcn_readlane(__double2loint(C), last));
        r1 *= Ce;
        r2 *= Ce;
        g += 16;
        kfirst = 0;
        if (g <= le && ((g - m) & 63) == 0) {  // renormalise at the block end
            const int s = renorm_shift(r1, r2);
            r1 = ldexp(r1, -s);
            r2 = ldexp(r2, -s);
#pragma unroll
            for (int u = 0; u < 4; ++u) acc[u] = ldexp(acc[u], -s);
            E += s;
        }
    };
    using Full = std::integral_constant<int, 0>;
    using Guarded = std::integral_constant<int, 2>;
    // The partial groups (the first of segment 0 starts at step 2; the order's last may end early)
    // are peeled off, and the loop over full groups runs two at a time with the line buffers in
    // alternation: one path, registers in place.
    if (g <= le && (kfirst != 0 || g + 15 > le)) {
        if (kfirst == 2 && g + 15 <= le) group(std::integral_constant<int, 1>{}, A, B);
        else group(Guarded{}, A, B);
        A = B;
    }
    while (g + 31 <= le) {
        group(Full{}, A, B);
        group(Full{}, B, A);
    }
    if (g + 15 <= le) {
        group(Full{}, A, B);
        A = B;
    }
    if (g <= le) group(Guarded{}, A, B);
    sink ^= pa ^ pb;
}

// One workgroup per 64-ring tile of one order; wavefront s sweeps degree segment s of one
// transform (column b0 of alm). Segment 0 starts from the seed, the others from their checkpoints;
// the segment sums, scaled to plain doubles, are added in segment order through LDS and wavefront 0
// stores the tile's ring values. The sweep holds the DPP operands of two groups; six wavefronts per
// SIMD fit its registers.
// GROUP: a launch over many single-transform engines (plan groups): workgroup i takes item
// work[i].y of engine work[i].x, whose tables come from tabs[work[i].x] instead of the arguments.
template <bool GROUP>
__global__ void __launch_bounds__(64 * kSegMax) __attribute__((amdgpu_waves_per_eu(6)))
ring_fwd1_kernel(const SegItem *__restrict__ items, const double *__restrict__ costh,
                 const double *__restrict__ log2sin, const double *__restrict__ log2c,
                 const double2 *__restrict__ ab, const double2 *__restrict__ gc, const int64_t *__restrict__ abofs,
                 const double2 *__restrict__ ckq, const int32_t *__restrict__ ckE, InCols alm, int b0,
                 double2 *__restrict__ Gd, int L, int n, int mmax, int never, const RingTab *__restrict__ tabs,
                 const int2 *__restrict__ work) {
    __shared__ double red[kSegMax - 1][4][64];
    int wi = blockIdx.x;
    if constexpr (GROUP) {
        const int2 w = work[blockIdx.x];
        const RingTab &T = tabs[w.x];
        items = T.fitems; costh = T.costh; log2sin = T.log2sin; log2c = T.log2c; ab = T.ab; gc = T.gc;
        abofs = T.abofs; ckq = T.ckq; ckE = T.ckE; alm = T.alm_in; Gd = T.Gd; L = T.L; n = T.n; mmax = T.mmax;
        b0 = 0;
        wi = w.y;
    }
    const SegItem it = items[wi];
    const int m = __builtin_amdgcn_readfirstlane(it.m);
    const int nseg = __builtin_amdgcn_readfirstlane(it.nseg);
    const int nr = __builtin_amdgcn_readfirstlane(it.nr);
    if (nr <= 0) return;  // an empty slot: the whole workgroup leaves
    const int sg = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);  // this wavefront's segment
    const int lane = threadIdx.x & 63;
    const bool active = lane < nr;
    const int j = it.j0 + (active ? lane : 0);

    // column of order m (row l - m holds degree l) and of order -m
    const gcptr acol = alm.col(b0);
    const gcptr cp0 = acol + (int64_t(2 * m) * (L + 1) - m);
    const gcptr cm0 = acol + (int64_t(2 * m - 1) * (L + 1) - m);
    const double2 *__restrict__ abm = ab + abofs[m] - (m + 2);  // abm[l] for l = m+2..L

    double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
    int sink = 0;
    if (sg < nseg) {
        const double x = costh[j];
        double q1, q2;
        int E, l;
        if (sg == 0) {
            seed(m == 0 ? log2c[0] : double(m) * log2sin[j] + log2c[m], q2, E);
            if (!active) q2 = 0.0;
            q1 = m < L ? first_step(m, x, q2) : 0.0;
            const double2 a0 = gload(cp0 + m);
            ar = a0.x * q2;
            ai = a0.y * q2;
            if (m > 0) {
                const double2 c0 = gload(cm0 + m);
                br = c0.x * q2;
                bi = c0.y * q2;
            }
            if (m < L) {
                const double2 a1 = gload(cp0 + (m + 1));
                ar = fma(a1.x, q1, ar);
                ai = fma(a1.y, q1, ai);
                if (m > 0) {
                    const double2 c1 = gload(cm0 + (m + 1));
                    br = fma(c1.x, q1, br);
                    bi = fma(c1.y, q1, bi);
                }
            }
            l = m + 2;
        } else {
            const int64_t row = (it.ck + sg - 1) * 64 + lane;
            const double2 c = ckq[row];
            q1 = c.x;
            q2 = c.y;
            E = ckE[row];
            l = seg_begin(m, L, nseg, sg);
        }
        const int le = min(L, seg_begin(m, L, nseg, sg + 1) - 1);
        if (l <= le) {
            const double2 *__restrict__ gcm = gc + abofs[m] - (m + 2);
            const double aprev = l == m + 2 ? sqrt(double(2 * m + 3)) : abm[l - 1].x;
            double acc[4] = {ar, ai, br, bi};
            if (m > 0) fwd_sweep_dpp<true>(gcm, cp0, cm0, x, q1, q2, aprev, E, l, le, m, L, acc, sink);
            else fwd_sweep_dpp<false>(gcm, cp0, cm0, x, q1, q2, aprev, E, l, le, m, L, acc, sink);
            ar = acc[0];
            ai = acc[1];
            br = acc[2];
            bi = acc[3];
        }
        ar = ldexp(ar, E);
        ai = ldexp(ai, E);
        br = ldexp(br, E);
        bi = ldexp(bi, E);
    }
    if (sink == never) Gd[j] = make_double2(0.0, 0.0);  // keeps the prefetch loads; never true
    if (nseg > 1) {
        if (sg > 0 && sg < nseg) {
            red[sg - 1][0][lane] = ar;
            red[sg - 1][1][lane] = ai;
            red[sg - 1][2][lane] = br;
            red[sg - 1][3][lane] = bi;
        }
        __syncthreads();
        if (sg == 0)
            for (int k = 0; k < nseg - 1; ++k) {
                ar += red[k][0][lane];
                ai += red[k][1][lane];
                br += red[k][2][lane];
                bi += red[k][3][lane];
            }
    }
    if (sg != 0 || !active) return;
    const double sgn = (m & 1) ? -1.0 : 1.0;
    const int64_t ncol = 2 * mmax + 1;
    Gd[(int64_t(b0) * ncol + (mmax + m)) * n + j] = make_double2(sgn * ar, sgn * ai);
    if (m > 0) Gd[(int64_t(b0) * ncol + (mmax - m)) * n + j] = make_double2(br, bi);
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

// Degrees ls .. le of order m for the first NT of the TG 64-ring tiles in T (NT = 2 when the tile
// group has a second tile). Segment 0 (ls = m) starts from the seed; a later segment from its
// checkpoint (q1, q2) = (q_{ls-1}, q_{ls-2}), aprev = a_{ls-1}. ls - m is a multiple of 64. Phase 1
// runs the rescaled recurrence of the forward sweep over each 16-degree sub-tile (gamma_l and C_l of
// lane k of a row give degree l0 + k); phase 2 adds the tile's ring sums into the row accumulators;
// the tiles take turns in the one LDS tile, and the accumulators, their sum over the four rows and
// the partial stores are shared. The ring sums of degree l are multiplied by C_l once, after the
// reduction. For one or two transforms the ring's exponent goes into its G operands (G 2^E, exact,
// refreshed when E changes at a block end) and the tile holds r; with more transforms (registers)
// the tile holds r 2^E. The sub-tiles holding the seed degrees or ending the order early are peeled
// off the loop, so the loop over full sub-tiles has one path.
template <int NB, bool NEG, int TG, int NT>
__device__ __forceinline__ void adj_sweep(const double2 *__restrict__ gcm, AdjTile<NB> (&T)[TG], double aprev,
                                          int ls, int le, int m, int L, int B, double (*W)[kWS],
                                          double4 *__restrict__ out, int &sink, const double (&qs)[TG],
                                          const double (&q1)[TG], const double (&q2)[TG]) {
    static_assert(NT >= 1 && NT <= TG, "tiles in use");
    constexpr bool kScaleG = AdjTile<NB>::kScaleG;
    const int lane = threadIdx.x & 63;
    const int t = lane & 15, row = lane >> 4;
    const int wcol = lane + (lane >> 4);
    const double *__restrict__ wr = &W[t][17 * row];
#pragma unroll
    for (int k = 0; k < NT; ++k) T[k].scale_g();

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
    for (int k = 0; k < NT; ++k) {
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
    int l0 = ls;
    // One sub-tile of degrees l0 .. l0 + 15. MODE 0: full, no seed degrees; 1: full, seed degrees m
    // and m + 1 in rows 0 and 1 (kfirst = 2); 2: rows kfirst .. tl - 1 only.
    auto subtile = [&](auto mode) {
        constexpr int MODE = decltype(mode)::value;
        const int tl = MODE == 2 ? min(kSub, le - l0 + 1) : kSub;
        if (((l0 - m) & 63) == 0) {
            sink ^= pending;
            const int lp = l0 + 128;
            pending = l2_touch(reinterpret_cast<const char *>(gcm + lp), pvoff, lane < 8 && lp + 8 * lane <= le);
        }
        const double2 gcn = coef(l0 + kSub);
        const double ng = -gcv.x, Cc = gcv.y;
        const double cmul = MODE == 0 ? Cc : (t < kfirst ? 1.0 : Cc);  // the seed degrees are stored unscaled
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
        for (int k = 0; k < NT; ++k) {
            AdjTile<NB> &A = T[k];
            // phase 1: lanes = rings
            if constexpr (MODE != 0) {
                if (kfirst > 0) {
                    W[0][wcol] = A.w0;
                    if (kfirst == 2) W[1][wcol] = A.w1;
                }
            }
            if constexpr (MODE == 0) AdjPhase1<0, false, !kScaleG>::run(A.x, A.r1, A.r2, ng, A.E, W, wcol, 0, kSub);
            else if constexpr (MODE == 1) AdjPhase1<2, false, !kScaleG>::run(A.x, A.r1, A.r2, ng, A.E, W, wcol, 2, kSub);
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
        l0 += kSub;
    };
    using Full = std::integral_constant<int, 0>;
    using Guarded = std::integral_constant<int, 2>;
    if (l0 <= le && (kfirst > 0 || le - l0 + 1 < kSub)) {
        if (kfirst == 2 && le - l0 + 1 >= kSub) subtile(std::integral_constant<int, 1>{});
        else subtile(Guarded{});
    }
    while (l0 + kSub - 1 <= le) subtile(Full{});
    if (l0 <= le) subtile(Guarded{});
    sink ^= pending;
}

// Workgroups of kAdjWaves wavefronts; each wavefront is one item: one degree segment of TG
// consecutive 64-ring tiles of one order (the second may be empty), writing that tile group's
// partials for the segment's degrees.
// GROUP: as the forward's; work[i].y is the engine's workgroup index (kAdjWaves items each).
template <int NB, int TG, bool GROUP = false>
__global__ void __launch_bounds__(64 * kAdjWaves)
ring_adj_kernel(const SegItem *__restrict__ items, const double *__restrict__ costh,
                const double *__restrict__ log2sin, const double *__restrict__ log2c,
                const double2 *__restrict__ ab, const double2 *__restrict__ gc, const int64_t *__restrict__ abofs,
                const double2 *__restrict__ ckq,
                const int32_t *__restrict__ ckE, const int64_t *__restrict__ pofs, const double2 *__restrict__ Gd,
                double4 *__restrict__ partials, int L, int n, int mmax, int B, int b0, int never,
                const RingTab *__restrict__ tabs = nullptr, const int2 *__restrict__ work = nullptr,
                OutCols zalm = OutCols(), int nring = 0, int znb = 0) {
    extern __shared__ double lds_adj[];
    if constexpr (!GROUP) {
        // workgroups past the nring item workgroups (one launch per apply carries them) zero the
        // columns of the orders above mmax in transforms 0 .. znb - 1, chunks of 2048 entries: launched
        // last, they run in the pass's tail, and the partial reduction writes only orders <= mmax
        if (znb > 0 && int(blockIdx.x) >= nring) {
            const int64_t base = int64_t(2 * mmax + 1) * (L + 1), count = int64_t(2 * (L - mmax)) * (L + 1);
            const int64_t nch = (count + 2047) / 2048;
            const int64_t z = int64_t(blockIdx.x) - nring;
            const int u = int(z / nch);
            const int64_t c0 = (z % nch) * 2048;
            if (u >= znb) return;
            const gptr col = zalm.col(u) + base;
            for (int i = threadIdx.x; i < 2048; i += 64 * kAdjWaves)
                if (c0 + i < count) gstore(col + (c0 + i), make_double2(0.0, 0.0));
            return;
        }
    }
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
    int wg = blockIdx.x;
    if constexpr (GROUP) {
        const int2 w = work[blockIdx.x];
        const RingTab &T = tabs[w.x];
        items = T.aitems; costh = T.costh; log2sin = T.log2sin; log2c = T.log2c; ab = T.ab; gc = T.gc;
        abofs = T.abofs; ckq = T.ckq; ckE = T.ckE; pofs = T.pofs; Gd = T.Gd; partials = T.partials;
        L = T.L; n = T.n; mmax = T.mmax; B = T.part_b;
        b0 = 0;
        wg = w.y;
    }
    const SegItem it = items[wg * kAdjWaves + wave];
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
    // wave-uniform choices of the sweep: orders -m present, and whether the group has a second tile
    if (two && TG > 1) {
        constexpr int NT = TG > 1 ? 2 : 1;
        if (m > 0) adj_sweep<NB, true, TG, NT>(gcm, T, aprev, ls, le, m, L, B, W, out, sink, qs, q1, q2);
        else adj_sweep<NB, false, TG, NT>(gcm, T, aprev, ls, le, m, L, B, W, out, sink, qs, q1, q2);
    } else {
        if (m > 0) adj_sweep<NB, true, TG, 1>(gcm, T, aprev, ls, le, m, L, B, W, out, sink, qs, q1, q2);
        else adj_sweep<NB, false, TG, 1>(gcm, T, aprev, ls, le, m, L, B, W, out, sink, qs, q1, q2);
    }
    if (sink == never) out[0] = make_double4(0.0, 0.0, 0.0, 0.0);  // never true; keeps the prefetch
}

// Writes the alm slots of orders |m| <= mmax (launched over those) or of every order (the plan
// groups) in transforms 0 .. nb - 1: the tile-ordered sum of the partials where (l, m) is retained,
// zero in the padding and for unretained orders (the adjoint overwrites alm,
// src/Legendre.jl:338). One thread owns row l - |m| of the column pair (2|m|, 2|m| - 1), reading each
// partial once. pb is the partials' transform stride.
// GROUP: one workgroup per work entry {engine, row block, |m|} of many single-transform engines.
template <bool GROUP = false>
__global__ void ring_adj_finalize(const int64_t *__restrict__ pofs, const int32_t *__restrict__ ntiles,
                                  const double4 *__restrict__ partials, OutCols alm, int L, int mmax, int pb,
                                  const RingTab *__restrict__ tabs = nullptr, const int4 *__restrict__ work = nullptr) {
    int rb = blockIdx.x, am = blockIdx.y, b = blockIdx.z;
    if constexpr (GROUP) {
        const int4 w = work[blockIdx.x];
        const RingTab &T = tabs[w.x];
        pofs = T.pofs; ntiles = T.ntiles; partials = T.partials; alm = T.alm_out; L = T.L; mmax = T.mmax;
        pb = T.part_b;
        rb = w.y;
        am = w.z;
        b = 0;
    }
    const int row = rb * blockDim.x + threadIdx.x;
    if (row > L) return;
    const int64_t nrow = L + 1;
    const gptr a = alm.col(b);
    const int l = row + am;
    double pr = 0, pi = 0, nr_ = 0, ni = 0;
    if (l <= L && am <= mmax) {
        const int nt = ntiles[am];
        const int nl = adj_rows_padded(am, L);
        const double4 *p = partials + pofs[am] + int64_t(l - am) * pb + b;
        for (int t = 0; t < nt; ++t) {
            const double4 q = p[int64_t(t) * nl * pb];
            pr += q.x;
            pi += q.y;
            nr_ += q.z;
            ni += q.w;
        }
        if (am & 1) { pr = -pr; pi = -pi; }
    }
    gstore(a + (int64_t(2 * am) * nrow + row), make_double2(pr, pi));
    if (am > 0) gstore(a + (int64_t(2 * am - 1) * nrow + row), make_double2(nr_, ni));
}

// ------------------------------------------------------------------------------------------------
// FP64 MFMA ring passes for two or more transforms. The per-order accumulation is a contraction of
// a (degree x ring) tile of the recurrence against the coefficient columns (forward) or the ring
// columns (adjoint), on v_mfma_f64_16x16x4f64. Column c of a 16-column block is component c % 4 of
// transform c / 4: (re, im) of order +m, then (re, im) of order -m. The recurrence runs on VALU
// with lanes over rings, exactly as the single-transform adjoint's phase 1, and stores r 2^E in the
// wavefront's LDS tile (r the rescaled recurrence value, E the ring's exponent); C_l, the rescale of
// degree l, goes on the coefficient side (forward) or multiplies the ring sum (adjoint), so every
// product is a_lm Pbar_l^m(x_j) or Pbar_l^m(x_j) G_j as plain doubles.
//
// Operand layouts of v_mfma_f64_16x16x4f64 (gfx950): A (16 x 4) lane 16 k + i holds A[i][k]; B
// (4 x 16) lane 16 k + j holds B[k][j]; D (16 x 16) register r of lane 16 k + j holds D[4 r + k][j].
// ------------------------------------------------------------------------------------------------
typedef double mfma_d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mfma_d4 mfma_f64(double a, double b, mfma_d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// LDS tile column of ring rho (0..63): rows are kWS doubles apart, ring rho at rho + rho / 16
__device__ __forceinline__ int wcol_of(int rho) { return rho + (rho >> 4); }

// Forward: one workgroup per 64-ring tile of one order, wavefront s sweeping degree segment s (as
// the single-transform pass), 4 NCB transforms b0 .. b0 + nt - 1. D[rb][cb] holds rings 16 rb ..
// 16 rb + 15 against column block cb: sum over degrees of (r 2^E)(a_lm C_l). The segments' sums are
// added in segment order through LDS and the tile is written to Gd with (-1)^m on the +m columns.
template <int NCB>
__global__ void __launch_bounds__(64 * kSegMax)
ring_fwd_mfma_kernel(const SegItem *__restrict__ items, const double *__restrict__ costh,
                     const double *__restrict__ log2sin, const double *__restrict__ log2c,
                     const double2 *__restrict__ ab, const double2 *__restrict__ gc,
                     const int64_t *__restrict__ abofs, const double2 *__restrict__ ckq,
                     const int32_t *__restrict__ ckE, InCols alm, int b0, int nt, double2 *__restrict__ Gd, int L,
                     int n, int mmax) {
    extern __shared__ double lds_fm[];
    constexpr int NC = 16 * NCB, NTR = 4 * NCB;
    constexpr int kRS = 65;  // row stride of the reduction buffer [column][ring]
    const SegItem it = items[blockIdx.x];
    const int m = __builtin_amdgcn_readfirstlane(it.m);
    const int nseg = __builtin_amdgcn_readfirstlane(it.nseg);
    const int nr = __builtin_amdgcn_readfirstlane(it.nr);
    if (nr <= 0) return;  // an empty slot: the whole workgroup leaves
    const int sg = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63, kq = lane >> 4, jc = lane & 15;
    const bool active = lane < nr;
    const int j = it.j0 + (active ? lane : 0);
    const int wcol = wcol_of(lane);
    double(*W)[kWS] = reinterpret_cast<double(*)[kWS]>(lds_fm + sg * kSub * kWS);

    // the coefficient each lane feeds to column block cb: transform t, component comp
    const PFSHT_GLOBAL double *bcol[NCB];
    bool bok[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
        const int c = 16 * cb + jc, t = c >> 2, comp = c & 3;
        bok[cb] = t < nt && (comp < 2 || m > 0);
        const bool neg = comp >= 2 && m > 0;  // order -m has no column when m = 0: point at +m, unused
        const gcptr acol = alm.col(b0 + min(t, nt - 1)) + (int64_t(neg ? 2 * m - 1 : 2 * m) * (L + 1) - m);
        bcol[cb] = (const PFSHT_GLOBAL double *)acol + (comp & 1);  // degree l at bcol[2 l]
    }
    mfma_d4 D[4][NCB];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) D[rb][cb] = mfma_d4{0.0, 0.0, 0.0, 0.0};

    if (sg < nseg) {
        const double2 *__restrict__ gcm = gc + abofs[m] - (m + 2);
        const double2 *__restrict__ abm = ab + abofs[m] - (m + 2);
        const bool has_tab = m + 2 <= L;
        const int ls = seg_begin(m, L, nseg, sg);
        const int le = min(L, seg_begin(m, L, nseg, sg + 1) - 1);
        const double x = costh[j];
        double r1, r2, w0 = 0.0, w1 = 0.0;
        int E, kfirst = 0;
        if (sg == 0) {
            double qs;
            seed(m == 0 ? log2c[0] : double(m) * log2sin[j] + log2c[m], qs, E);
            if (!active) qs = 0.0;
            w0 = ldexp(qs, E);
            kfirst = 1;
            r1 = qs;
            r2 = 0.0;
            if (m < L) {
                const double q = first_step(m, x, qs);
                w1 = ldexp(q, E);
                kfirst = 2;
                r1 = q;
                r2 = qs * sqrt(double(2 * m + 3));
            }
        } else {
            const int64_t row = (it.ck + sg - 1) * 64 + lane;
            const double2 c = ckq[row];
            E = ckE[row];
            r1 = c.x;
            r2 = c.y * abm[ls - 1].x;
        }
#pragma unroll
        for (int k = 0; k < kSub; ++k) W[k][wcol] = 0.0;
        // The next sub-tile's operands are loaded raw, unconditionally and at indices clamped into the
        // order's table (entry 0 of the table when the order has none), and are masked and scaled
        // only when used: a load whose value feeds a select at issue is sunk into a branch and waited
        // for on the spot, while raw loads are counted and waited for one sub-tile later.
        const double2 *__restrict__ gct = has_tab ? gcm : gc;
        auto gidx = [&](int l) { return has_tab ? min(max(l, m + 2), L) : 0; };
        double braw[4][NCB], craw[4];
        double2 graw;
        auto load_next = [&](int l0) {
            graw = gct[gidx(l0 + jc)];  // (gamma, C) of degree l0 + (lane % 16)
#pragma unroll
            for (int st = 0; st < 4; ++st) {
                const int dc = min(l0 + 4 * st + kq, L);
                craw[st] = gct[gidx(dc)].y;
#pragma unroll
                for (int cb = 0; cb < NCB; ++cb) braw[st][cb] = bcol[cb][2 * dc];
            }
        };
        load_next(ls);
        for (int l0 = ls; l0 <= le; l0 += kSub) {
            const int tl = min(kSub, le - l0 + 1);
            // B operands: step s, block cb holds degree l0 + 4 s + kq times C_l, zero past le
            double bv[4][NCB];
#pragma unroll
            for (int st = 0; st < 4; ++st) {
                const int deg = l0 + 4 * st + kq;
                const double C = (deg < m + 2 || !has_tab) ? 1.0 : craw[st];
#pragma unroll
                for (int cb = 0; cb < NCB; ++cb) bv[st][cb] = (deg <= le && bok[cb]) ? braw[st][cb] * C : 0.0;
            }
            const double2 gcv = has_tab ? graw : make_double2(0.0, 1.0);
            load_next(l0 + kSub);
            const double ng = -gcv.x, Cc = gcv.y;
            const int last = tl - 1;
            const double Ce = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(Cc), last),
                                               __builtin_amdgcn_readlane(__double2loint(Cc), last));
            if (kfirst > 0) {
                W[0][wcol] = w0;
                if (kfirst == 2) W[1][wcol] = w1;
            }
            if (tl == kSub && kfirst == 0) AdjPhase1<0, false, true>::run(x, r1, r2, ng, E, W, wcol, 0, kSub);
            else if (tl == kSub && kfirst == 2) AdjPhase1<2, false, true>::run(x, r1, r2, ng, E, W, wcol, 2, kSub);
            else AdjPhase1<0, true, true>::run(x, r1, r2, ng, E, W, wcol, kfirst, tl);
            r1 *= Ce;
            r2 *= Ce;
            if (((l0 + kSub - m) & 63) == 0) {  // renormalise at every 64-degree block end
                const int sh = renorm_shift(r1, r2);
                r1 = ldexp(r1, -sh);
                r2 = ldexp(r2, -sh);
                E += sh;
            }
            wave_lds_order();
#pragma unroll
            for (int st = 0; st < 4; ++st) {
                double a[4];
#pragma unroll
                for (int rb = 0; rb < 4; ++rb) a[rb] = W[4 * st + kq][17 * rb + jc];
#pragma unroll
                for (int rb = 0; rb < 4; ++rb) {
                    if (16 * rb >= nr) break;  // wave-uniform: ring blocks past the tile's rings are empty
#pragma unroll
                    for (int cb = 0; cb < NCB; ++cb) D[rb][cb] = mfma_f64(a[rb], bv[st][cb], D[rb][cb]);
                }
            }
            wave_lds_order();
            kfirst = 0;
        }
    }
    // segment sums in segment order; the W tiles are free now and the buffer overlays them
    __syncthreads();
    double *red = lds_fm;  // [NC][kRS]
    for (int turn = 0; turn < nseg; ++turn) {
        if (sg == turn) {
#pragma unroll
            for (int rb = 0; rb < 4; ++rb)
#pragma unroll
                for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        double &cell = red[(16 * cb + jc) * kRS + 16 * rb + 4 * r + kq];
                        cell = turn == 0 ? D[rb][cb][r] : cell + D[rb][cb][r];
                    }
        }
        __syncthreads();
    }
    const int64_t ncol = 2 * mmax + 1;
    const double sgn = (m & 1) ? -1.0 : 1.0;
    for (int q = threadIdx.x; q < NTR * 2 * 64; q += blockDim.x) {
        const int rho = q & 63, pm = (q >> 6) & 1, t = q >> 7;
        if (t >= nt || rho >= nr || (pm && m == 0)) continue;
        const double re = red[(4 * t + 2 * pm) * kRS + rho], im = red[(4 * t + 2 * pm + 1) * kRS + rho];
        Gd[(int64_t(b0 + t) * ncol + (pm ? mmax - m : mmax + m)) * n + it.j0 + rho] =
            pm ? make_double2(re, im) : make_double2(sgn * re, sgn * im);
    }
}

// Adjoint: one workgroup per (order, degree segment); wavefront w takes tiles w + nw k (k < TG) of the
// order, with the tile's ring columns G (the B operands, ring 4 s + k of quad s against column
// block cb) in registers for the whole sweep. Per 16-degree sub-tile each tile's recurrence fills
// the wavefront's LDS tile and 16 ring quads are contracted into D (16 degrees x NC columns); the
// wavefronts' D are then added in LDS in wavefront order (tiles w, w + nw, ... in tile order within
// a wavefront), multiplied by C_l, and written to alm: no partials, no float atomics, bit-identical
// run to run.
template <int NCB, int TG>
__global__ void __launch_bounds__(64 * 8)
ring_adj_mfma_kernel(const OrderSeg *__restrict__ items, const double *__restrict__ costh,
                     const double *__restrict__ log2sin, const double *__restrict__ log2c,
                     const double2 *__restrict__ ab, const double2 *__restrict__ gc,
                     const int64_t *__restrict__ abofs, const double2 *__restrict__ ckq,
                     const int32_t *__restrict__ ckE, const double2 *__restrict__ Gd, OutCols alm, int b0, int nt,
                     int L, int n, int mmax) {
    extern __shared__ double lds_am[];
    constexpr int NC = 16 * NCB, NTR = 4 * NCB, kRS = NC + 1;
    const OrderSeg it = items[blockIdx.x];
    const int m = __builtin_amdgcn_readfirstlane(it.m);
    const int ntile = __builtin_amdgcn_readfirstlane(it.ntile);
    if (ntile <= 0) return;  // an empty slot: the whole workgroup leaves
    const int nseg = __builtin_amdgcn_readfirstlane(it.nseg);
    const int sg = __builtin_amdgcn_readfirstlane(it.seg);
    const int nw = int(blockDim.x) >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63, kq = lane >> 4, jc = lane & 15;
    const int wcol = wcol_of(lane);
    double(*W)[kWS] = reinterpret_cast<double(*)[kWS]>(lds_am + wave * kSub * kWS);
    double *red = lds_am + nw * kSub * kWS;  // [2][nw][16][kRS]
    const int ls = seg_begin(m, L, nseg, sg);
    const int le = min(L, seg_begin(m, L, nseg, sg + 1) - 1);
    const int64_t ncol = 2 * mmax + 1;
    const double2 *__restrict__ gcm = gc + abofs[m] - (m + 2);
    const double2 *__restrict__ abm = ab + abofs[m] - (m + 2);
    const bool has_tab = m + 2 <= L;

    double x[TG], r1[TG], r2[TG], w0[TG], w1[TG];
    int E[TG], nrt[TG];
    bool on[TG];
    double gB[TG][16][NCB];
    int kfirst = 0;
#pragma unroll
    for (int k = 0; k < TG; ++k) {
        const int tile = wave + nw * k;
        on[k] = tile < ntile;
        const int nrk = on[k] ? (tile == ntile - 1 ? it.nrl : 64) : 0;
        nrt[k] = nrk;
        const int jt = it.j0 + 64 * tile;
        const bool act = lane < nrk;
        const int jj = act ? jt + lane : it.j0;
        x[k] = costh[jj];
        w0[k] = w1[k] = 0.0;
        E[k] = 0;
        r1[k] = r2[k] = 0.0;
        if (on[k]) {
            if (sg == 0) {
                double qs;
                seed(m == 0 ? log2c[0] : double(m) * log2sin[jj] + log2c[m], qs, E[k]);
                if (!act) qs = 0.0;
                w0[k] = ldexp(qs, E[k]);
                kfirst = 1;
                r1[k] = qs;
                if (m < L) {
                    const double q = first_step(m, x[k], qs);
                    w1[k] = ldexp(q, E[k]);
                    kfirst = 2;
                    r1[k] = q;
                    r2[k] = qs * sqrt(double(2 * m + 3));
                }
            } else {
                const int64_t row = (it.ck + int64_t(tile) * (nseg - 1) + sg - 1) * 64 + lane;
                const double2 c = ckq[row];
                E[k] = ckE[row];
                r1[k] = act ? c.x : 0.0;
                r2[k] = act ? c.y * abm[ls - 1].x : 0.0;
            }
        }
#pragma unroll
        for (int st = 0; st < 16; ++st) {
            const int rho = 4 * st + kq;
            const bool rok = rho < nrk;
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                const int c = 16 * cb + jc, t = c >> 2, comp = c & 3;
                const bool ok = rok && t < nt && (comp < 2 || m > 0);
                double v = 0.0;
                if (ok) {
                    const double2 g = Gd[(int64_t(b0 + t) * ncol + (comp < 2 ? mmax + m : mmax - m)) * n + jt + rho];
                    v = (comp & 1) ? g.y : g.x;
                }
                gB[k][st][cb] = v;
            }
        }
    }
    if (sg == 0) kfirst = m < L ? 2 : 1;  // wave-uniform (a wavefront without tiles joins in)
#pragma unroll
    for (int k = 0; k < kSub; ++k) W[k][wcol] = 0.0;
    // (gamma, C) of degree l0 + (lane % 16), loaded raw one sub-tile ahead, unconditionally at a
    // clamped index (entry 0 of the table when the order has none), and selected when used
    const double2 *__restrict__ gct = has_tab ? gcm : gc;
    auto coef_raw = [&](int l0) { return gct[has_tab ? min(max(l0 + jc, m + 2), L) : 0]; };
    double2 graw = coef_raw(ls);
    int buf = 0;
    const int nwt = min(nw, ntile);  // wavefronts holding tiles
    for (int l0 = ls; l0 <= le; l0 += kSub) {
        const int tl = min(kSub, le - l0 + 1);
        const double2 gcv = has_tab ? graw : make_double2(0.0, 1.0);
        graw = coef_raw(l0 + kSub);
        const double ng = -gcv.x, Cc = gcv.y;
        const int last = tl - 1;
        const double Ce = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(Cc), last),
                                           __builtin_amdgcn_readlane(__double2loint(Cc), last));
        const bool bend = ((l0 + kSub - m) & 63) == 0;
        mfma_d4 D[NCB];
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) D[cb] = mfma_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < TG; ++k) {
            if (!on[k]) break;  // wave-uniform; tiles of a wavefront are in increasing order
            if (kfirst > 0) {
                W[0][wcol] = w0[k];
                if (kfirst == 2) W[1][wcol] = w1[k];
            }
            if (tl == kSub && kfirst == 0) AdjPhase1<0, false, true>::run(x[k], r1[k], r2[k], ng, E[k], W, wcol, 0, kSub);
            else if (tl == kSub && kfirst == 2) AdjPhase1<2, false, true>::run(x[k], r1[k], r2[k], ng, E[k], W, wcol, 2, kSub);
            else AdjPhase1<0, true, true>::run(x[k], r1[k], r2[k], ng, E[k], W, wcol, kfirst, tl);
            r1[k] *= Ce;
            r2[k] *= Ce;
            if (bend) {
                const int sh = renorm_shift(r1[k], r2[k]);
                r1[k] = ldexp(r1[k], -sh);
                r2[k] = ldexp(r2[k], -sh);
                E[k] += sh;
            }
            wave_lds_order();
#pragma unroll
            for (int st = 0; st < 16; ++st) {
                if (4 * st >= nrt[k]) break;  // wave-uniform: ring quads past the tile's rings are empty
                const double a = W[jc][wcol_of(4 * st + kq)];
#pragma unroll
                for (int cb = 0; cb < NCB; ++cb) D[cb] = mfma_f64(a, gB[k][st][cb], D[cb]);
            }
            wave_lds_order();
        }
        double *rw = red + (buf * nw + wave) * 16 * kRS;
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r) rw[(4 * r + kq) * kRS + 16 * cb + jc] = D[cb][r];
        __syncthreads();
        // the wavefronts' sums in wavefront order, times C_l, into alm: thread (d, t, pm)
        for (int q = threadIdx.x; q < 32 * NTR; q += blockDim.x) {
            const int d = q & 15, pm = (q >> 4) & 1, t = q >> 5;
            const int l = l0 + d;
            if (l <= le && t < nt && (pm == 0 || m > 0)) {
                const double *rr = red + (buf * nw) * 16 * kRS + d * kRS + 4 * t + 2 * pm;
                double re = 0.0, im = 0.0;
                for (int w = 0; w < nwt; ++w) {
                    re += rr[w * 16 * kRS];
                    im += rr[w * 16 * kRS + 1];
                }
                const double cm = l < m + 2 ? 1.0 : gcm[l].y;
                re *= cm;
                im *= cm;
                if (pm == 0 && (m & 1)) { re = -re; im = -im; }
                gstore(alm.col(b0 + t) + (int64_t(pm ? 2 * m - 1 : 2 * m) * (L + 1) + (l - m)), make_double2(re, im));
            }
        }
        buf ^= 1;
        kfirst = 0;
    }
}

// alm entries the MFMA adjoint does not write: the padding rows of every column and every row of
// the orders without retained rings (the adjoint overwrites alm, src/Legendre.jl:338).
__global__ void __launch_bounds__(256)
alm_zero_unwritten_kernel(OutCols alm, const int32_t *__restrict__ jlo, const int32_t *__restrict__ jhi, int L,
                          int mmax) {
    const int c = blockIdx.x, b = blockIdx.y;
    const int am = (c + 1) >> 1;  // column c = 2|m| (m >= 0) or 2|m| - 1 (m < 0)
    const bool retained = am <= mmax && jhi[am] >= jlo[am];
    const int first = retained ? L - am + 1 : 0;  // rows first .. L are not written by the ring pass
    const gptr col = alm.col(b) + int64_t(c) * (L + 1);
    for (int row = first + int(threadIdx.x); row <= L; row += 256) gstore(col + row, make_double2(0.0, 0.0));
}

// Transform groups of 8, 4, 2 and 1 cover any count exactly with at most four launches.
template <class F>
static void for_transform_groups(int B, F &&launch) {
    int b0 = 0;
    for (int nb : {8, 4, 2, 1})
        while (B - b0 >= nb) {
            launch(nb, b0);
            b0 += nb;
        }
}

// MFMA launches cover nb transforms in chunks of up to 8 (two column blocks), a last chunk of up
// to 4 on one block.
template <class F>
static void for_mfma_chunks(int nb, F &&launch) {
    for (int b0 = 0; b0 < nb;) {
        const int rem = nb - b0, ncb = rem > 4 ? 2 : 1, nt = std::min(rem, 4 * ncb);
        launch(ncb, b0, nt);
        b0 += nt;
    }
}

void launch_ring_fwd(const RingDev &r, InCols alm, int nb, double2 *Gd, hipStream_t s) {
    if (r.nfitems == 0 || nb <= 0) return;
    dim3 grid(r.nfitems), block(64 * r.fwd_waves);
    if (nb == 1) {
        ring_fwd1_kernel<false><<<grid, block, 0, s>>>(r.fitems, r.costh, r.log2sin, r.log2c, r.ab, r.gc, r.abofs,
                                                       r.ckq, r.ckE, alm, 0, Gd, r.L, r.n, r.mmax, kNeverSink,
                                                       nullptr, nullptr);
        return;
    }
    const size_t lds = sizeof(double) * std::max(r.fwd_waves * kSub * kWS, 32 * 65);
    for_mfma_chunks(nb, [&](int ncb, int b0, int nt) {
        if (ncb == 2)
            ring_fwd_mfma_kernel<2><<<grid, block, lds, s>>>(r.fitems, r.costh, r.log2sin, r.log2c, r.ab, r.gc,
                                                             r.abofs, r.ckq, r.ckE, alm, b0, nt, Gd, r.L, r.n, r.mmax);
        else
            ring_fwd_mfma_kernel<1><<<grid, block, lds, s>>>(r.fitems, r.costh, r.log2sin, r.log2c, r.ab, r.gc,
                                                             r.abofs, r.ckq, r.ckE, alm, b0, nt, Gd, r.L, r.n, r.mmax);
    });
}

bool ring_adj_uses_mfma(const RingDev &r, int nb) { return nb >= 2 && r.mfma_waves > 0 && r.nmitems > 0; }

static size_t ring_adj_mfma_lds(int nw, int ncb) {
    return sizeof(double) * (size_t(nw) * kSub * kWS + 2 * size_t(nw) * 16 * (16 * ncb + 1));
}

void launch_ring_adj_pass(const RingDev &r, const double2 *Gd, OutCols alm, int nb, hipStream_t s) {
    if (nb <= 0) return;
    if (ring_adj_uses_mfma(r, nb)) {
        alm_zero_unwritten_kernel<<<dim3(2 * r.L + 1, nb), 256, 0, s>>>(alm, r.jlo, r.jhi, r.L, r.mmax);
        dim3 grid(r.nmitems), block(64 * r.mfma_waves);
        for_mfma_chunks(nb, [&](int ncb, int b0, int nt) {
#define PFSHT_ADJ_MFMA(NCB, TG)                                                                              \
    ring_adj_mfma_kernel<NCB, TG><<<grid, block, ring_adj_mfma_lds(r.mfma_waves, NCB), s>>>(                  \
        r.mitems, r.costh, r.log2sin, r.log2c, r.ab, r.gc, r.abofs, r.ckq, r.ckE, Gd, alm, b0, nt, r.L, r.n, r.mmax)
            if (ncb == 2) {
                if (r.mfma_tg == 2) PFSHT_ADJ_MFMA(2, 2);
                else PFSHT_ADJ_MFMA(2, 1);
            } else {
                if (r.mfma_tg == 2) PFSHT_ADJ_MFMA(1, 2);
                else PFSHT_ADJ_MFMA(1, 1);
            }
#undef PFSHT_ADJ_MFMA
        });
        return;
    }
    if (r.naitems > 0)
        for_transform_groups(nb, [&](int ng, int b0) {
            // the first launch carries the workgroups that zero the orders above mmax of all nb
            // transforms (ring_adj_finalize writes orders <= mmax only)
            const int nring = r.naitems / kAdjWaves;
            const int64_t zcount = int64_t(2 * (r.L - r.mmax)) * (r.L + 1);
            const int nzero = b0 == 0 ? int((zcount + 2047) / 2048) * nb : 0;
            dim3 grid(nring + nzero);
            const size_t lds = sizeof(double) * kSub * kWS * kAdjWaves;
#define PFSHT_ADJ(NB, TG)                                                                                    \
    ring_adj_kernel<NB, TG><<<grid, 64 * kAdjWaves, lds, s>>>(r.aitems, r.costh, r.log2sin, r.log2c, r.ab, r.gc,  \
                                                              r.abofs, r.ckq, r.ckE, r.pofs, Gd, r.partials, r.L, r.n, \
                                                              r.mmax, r.part_b, b0, kNeverSink, nullptr, nullptr, alm, \
                                                              nring, nzero > 0 ? nb : 0)
            // the item list was built for r.adj_tg tiles per wavefront (two when only one or two
            // transforms use this path)
            switch (ng) {
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

void launch_ring_adj_finish(const RingDev &r, OutCols alm, int nb, hipStream_t s) {
    if (nb <= 0 || ring_adj_uses_mfma(r, nb)) return;
    dim3 fgrid((r.L + 1 + 255) / 256, r.naitems > 0 ? r.mmax + 1 : r.L + 1, nb);
    ring_adj_finalize<false><<<fgrid, 256, 0, s>>>(r.pofs, r.ntiles, r.partials, alm, r.L, r.mmax, r.part_b);
}

// Plan groups: the single-transform ring passes of many engines in one launch each.
void launch_ring_fwd_group(const RingTab *tabs, const int2 *work, int nwork, int waves, hipStream_t s) {
    if (nwork <= 0) return;
    ring_fwd1_kernel<true><<<dim3(nwork), dim3(64 * waves), 0, s>>>(
        nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, InCols(), 0, nullptr, 0, 0,
        0, kNeverSink, tabs, work);
}

void launch_ring_adj_group(const RingTab *tabs, const int2 *work, int nwork, int tg, hipStream_t s) {
    if (nwork <= 0) return;
    const size_t lds = sizeof(double) * kSub * kWS * kAdjWaves;
    if (tg == 2)
        ring_adj_kernel<1, 2, true><<<dim3(nwork), 64 * kAdjWaves, lds, s>>>(
            nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
            nullptr, 0, 0, 0, 1, 0, kNeverSink, tabs, work);
    else
        ring_adj_kernel<1, 1, true><<<dim3(nwork), 64 * kAdjWaves, lds, s>>>(
            nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
            nullptr, 0, 0, 0, 1, 0, kNeverSink, tabs, work);
}

void launch_ring_finalize_group(const RingTab *tabs, const int4 *work, int nwork, hipStream_t s) {
    if (nwork <= 0) return;
    ring_adj_finalize<true><<<dim3(nwork), 256, 0, s>>>(nullptr, nullptr, nullptr, OutCols(), 0, 0, 1, tabs, work);
}

void launch_ring_adj(const RingDev &r, const double2 *Gd, OutCols alm, int nb, hipStream_t s) {
    launch_ring_adj_pass(r, Gd, alm, nb, s);
    launch_ring_adj_finish(r, alm, nb, s);
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
