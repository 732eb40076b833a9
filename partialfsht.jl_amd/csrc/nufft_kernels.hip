// nufft_kernels.hip — the exponential-of-semicircle NUFFT on the stretched torus.
//
// Synthesis needs a type-2 NUFFT with iflag +1 (src/Stretch.jl:274; definition src/NUFFT.jl:85-106),
// the adjoint a type-1 NUFFT with iflag -1 (src/Stretch.jl:286; src/NUFFT.jl:107-124). With kernel
// psi(x) = phi(x / alpha), phi(z) = exp(beta (sqrt(1 - z^2) - 1)), alpha = w h / 2, h = 2 pi / N:
//   type 2:  u = IFFT_N( c_k * h/psihat(k) zero-padded ),  f_j = sum_{w x w cells} psi psi u;
//   type 1:  exactly the transpose — spread, FFT_N, multiply by h/psihat(k) and truncate.
// Interpolation and spreading evaluate the kernel with the same expression on the same cells,
// so ⟨type2 c, f⟩ = ⟨c, type1 f⟩ holds to round-off.
//
// Points are bin-sorted at plan time. Interpolation stages each super-tile's window of the fine
// grid in LDS and gathers from there. Spreading is deterministic without atomics: a workgroup owns
// a super-tile of 4 x 4 bins (S x S base cells each, S = 17 - w so a bin's footprint is 16 x 16)
// and accumulates each bin's footprint window as an FP64 MFMA outer-product sum
// (v_mfma_f64_16x16x4f64: four points per instruction, A = f_p * psi_row, B = psi_col). The four
// waves compute four bins' windows concurrently, then add them into the LDS super-tile window one
// wave after another, so every cell is summed in a fixed order. A merge pass then sums, for every
// fine-grid cell, the (at most 2 x 2) super-tile windows covering it in a fixed order.
#include <type_traits>
#include <utility>

#include "pfsht_common.hpp"

namespace pfsht {

typedef double double4v __attribute__((ext_vector_type(4)));

// Wavefront-private ordering of LDS traffic: a wavefront's LDS operations complete in issue
// order, so a wavefront reading back what it wrote needs only a compiler barrier.
__device__ __forceinline__ void wave_lds_order() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// The ES kernel as piecewise polynomials (EsHorner<W>, generated by gen_es_horner.py): the
// coefficients are compile-time constants, so they are immediate operands of the FMAs (materialised
// in SGPRs where they are used), never loaded.
#include "es_horner_tables.inc"

// A coefficient passed through an empty asm that also takes a polynomial's running value: the asm
// cannot move above the previous Horner step, so the coefficient is materialised in SGPRs just
// before its use instead of many steps ahead (which overflows the SGPR file in a loop).
__device__ __forceinline__ double coef_after(double c, double acc) {
    asm volatile("" : "+s"(c) : "v"(acc));
    return c;
}

// The same, with the coefficient's two 32-bit halves written by volatile scalar moves tied to the
// running value: a volatile asm is never hoisted out of a loop, so inside a loop over points the
// coefficients are not kept live across iterations (which spills the SGPR file).
template <int I, class T>
__device__ __forceinline__ double coef_now(double acc) {
    constexpr uint64_t bits = __builtin_bit_cast(uint64_t, T::c[I]);
    uint32_t lo, hi;
    asm volatile("s_mov_b32 %0, %1" : "=s"(lo) : "i"(uint32_t(bits)), "v"(acc));
    asm volatile("s_mov_b32 %0, %1" : "=s"(hi) : "i"(uint32_t(bits >> 32)), "v"(acc));
    return __hiloint2double(int(hi), int(lo));
}

// Both rows of a point (axes 1 and 2) in lockstep, for interpolation: the W - 2 interior polynomials
// in u and the edge polynomial at the two end cells, all running their Horner steps together, so
// each coefficient is materialised once (tied to the previous step, as in the half row below) and
// serves the FMAs of both axes.
template <int W>
struct FullRows {
    static constexpr int ni = EsHorner<W>::nint, ne = EsHorner<W>::nedge;
    static constexpr int EOFF = (W - 2) * ni;
    static constexpr int NS = (ne > ni ? ne : ni) - 1;
    static constexpr int S0E = NS - (ne - 1), S0I = NS - (ni - 1);
    // a[0] / a[W - 1]: left / right edge; a[T + 1]: interior cell T + 1
    template <int S, int... T>
    static __device__ __forceinline__ void step(const double (&e1)[2], const double (&e2)[2], double u1, double u2,
                                                double (&a1)[W], double (&a2)[W], std::integer_sequence<int, T...>) {
        if constexpr (S >= S0E) {
            const double c = coef_now<EOFF + ne - 2 - (S - S0E), EsHorner<W>>(a1[0]);
            a1[0] = fma(a1[0], e1[0], c);
            a1[W - 1] = fma(a1[W - 1], e1[1], c);
            a2[0] = fma(a2[0], e2[0], c);
            a2[W - 1] = fma(a2[W - 1], e2[1], c);
        }
        if constexpr (S >= S0I) (inner<S, T>(u1, u2, a1, a2), ...);
    }
    template <int S, int T>
    static __device__ __forceinline__ void inner(double u1, double u2, double (&a1)[W], double (&a2)[W]) {
        const double c = coef_now<T * ni + ni - 2 - (S - S0I), EsHorner<W>>(a1[T + 1]);
        a1[T + 1] = fma(a1[T + 1], u1, c);
        a2[T + 1] = fma(a2[T + 1], u2, c);
    }
    template <int... S>
    static __device__ __forceinline__ void steps(const double (&e1)[2], const double (&e2)[2], double u1, double u2,
                                                 double (&a1)[W], double (&a2)[W], std::integer_sequence<int, S...>) {
        (step<S>(e1, e2, u1, u2, a1, a2, std::make_integer_sequence<int, W - 2>{}), ...);
    }
    template <int... T>
    static __device__ __forceinline__ void init(double (&a)[W], double dep, std::integer_sequence<int, T...>) {
        a[0] = a[W - 1] = coef_now<EOFF + ne - 1, EsHorner<W>>(dep);
        ((a[T + 1] = coef_now<T * ni + ni - 1, EsHorner<W>>(dep)), ...);
    }
};
// The kernel values of a point at (x1, x2) with base cells (b1, b2): k1 along axis 1, k2 along axis 2,
// the same values as es_row on each axis.
template <int W>
__device__ __forceinline__ void es_rows(double x1, int b1, double x2, int b2, double (&k1)[W], double (&k2)[W]) {
    using F = FullRows<W>;
    const double u1 = 2.0 * (double(b1) - x1) + double(W - 1);
    const double u2 = 2.0 * (double(b2) - x2) + double(W - 1);
    const double e1[2] = {2.0 * sqrt(0.5 * (1.0 + u1)) - 1.0, 2.0 * sqrt(0.5 * (1.0 - u1)) - 1.0};
    const double e2[2] = {2.0 * sqrt(0.5 * (1.0 + u2)) - 1.0, 2.0 * sqrt(0.5 * (1.0 - u2)) - 1.0};
    F::init(k1, u1, std::make_integer_sequence<int, W - 2>{});
    F::init(k2, u2, std::make_integer_sequence<int, W - 2>{});
    F::steps(e1, e2, u1, u2, k1, k2, std::make_integer_sequence<int, F::NS>{});
}

// Half a row, from the left end for v = u and, through phi(-z) = phi(z), from the right end for
// v = -u: k[t] is the kernel at cell t (left) or W - 1 - t (right), t < ceil(W/2): the edge
// polynomial (k[0]) and the interior ones, evaluated by Horner in lockstep, so that the polynomials
// end on the same step and one step of all of them issues together. Interior coefficients go
// through coef_after; hoisting them out of the spreading loop instead would park them in VGPRs.
template <int W>
struct HalfRow {
    static constexpr int NH = (W + 1) / 2;  // cells per half row: the edge cell and NH - 1 interior ones
    static constexpr int ni = EsHorner<W>::nint, ne = EsHorner<W>::nedge;
    static constexpr int EOFF = (W - 2) * ni;                      // edge coefficients
    static constexpr int NS = (ne > ni ? ne : ni) - 1;             // Horner steps of the longest
    static constexpr int S0E = NS - (ne - 1), S0I = NS - (ni - 1);  // first step of each polynomial

    // The edge polynomial's coefficients stay in VGPRs for the whole kernel (loaded once, pinned
    // by an empty asm so they are not rematerialised): a third fewer scalar moves per chunk.
    struct Edge {
        double c[ne];
    };
    static __device__ __forceinline__ void pin(double &x) { asm volatile("" : "+v"(x)); }
    template <int... I>
    static __device__ __forceinline__ Edge load_edge(std::integer_sequence<int, I...>) {
        Edge ec;
        ((ec.c[I] = EsHorner<W>::c[EOFF + I]), ...);
        (pin(ec.c[I]), ...);
        return ec;
    }
    static __device__ __forceinline__ Edge edge() { return load_edge(std::make_integer_sequence<int, ne>{}); }

    template <int S, int... T>
    static __device__ __forceinline__ void step(double e, double v, double (&a)[NH], const Edge &ec,
                                                std::integer_sequence<int, T...>) {
        if constexpr (S >= S0E) a[0] = fma(a[0], e, ec.c[ne - 2 - (S - S0E)]);
        if constexpr (S >= S0I)
            ((a[T + 1] = fma(a[T + 1], v, coef_after(EsHorner<W>::c[T * ni + ni - 2 - (S - S0I)], a[T + 1]))), ...);
    }
    template <int... S>
    static __device__ __forceinline__ void steps(double e, double v, double (&a)[NH], const Edge &ec,
                                                 std::integer_sequence<int, S...>) {
        (step<S>(e, v, a, ec, std::make_integer_sequence<int, NH - 1>{}), ...);
    }
    template <int... T>
    static __device__ __forceinline__ void init(double (&a)[NH], const Edge &ec, std::integer_sequence<int, T...>) {
        a[0] = ec.c[ne - 1];
        ((a[T + 1] = coef_after(EsHorner<W>::c[T * ni + ni - 1], 0.0)), ...);
    }
};
template <int W>
__device__ __forceinline__ void es_half_row(double v, double (&k)[(W + 1) / 2], const typename HalfRow<W>::Edge &ec) {
    using H = HalfRow<W>;
    const double e = 2.0 * sqrt(0.5 * (1.0 + v)) - 1.0;
    H::init(k, ec, std::make_integer_sequence<int, H::NH - 1>{});
    H::steps(e, v, k, ec, std::make_integer_sequence<int, H::NS>{});
}

// ------------------------------------------------------------------------------------------------
// Deconvolution / padding
// ------------------------------------------------------------------------------------------------
__global__ void deconv_pad_kernel(const double2 *__restrict__ Fhat, const double2 *__restrict__ pc1,
                                  const double *__restrict__ c2, double2 *__restrict__ U, int n1, int n2, int N1,
                                  int N2) {
    const int l2 = blockIdx.x * blockDim.x + threadIdx.x;
    const int l1 = blockIdx.y;
    const int b = blockIdx.z;
    if (l2 >= N2) return;
    int i = -1, k = -1;
    if (l1 < (n1 + 1) / 2) i = l1;
    else if (l1 >= N1 - n1 / 2) i = l1 - N1 + n1;
    if (l2 < (n2 + 1) / 2) k = l2;
    else if (l2 >= N2 - n2 / 2) k = l2 - N2 + n2;
    double2 v = make_double2(0.0, 0.0);
    if (i >= 0 && k >= 0) {
        const double2 F = Fhat[(int64_t(b) * n1 + i) * n2 + k];
        const double2 p = pc1[i];
        const double s = c2[k];
        v = make_double2((F.x * p.x - F.y * p.y) * s, (F.x * p.y + F.y * p.x) * s);
    }
    U[(int64_t(b) * N1 + l1) * N2 + l2] = v;
}

__global__ void deconv_trunc_kernel(const double2 *__restrict__ U, const double2 *__restrict__ pc1,
                                    const double *__restrict__ c2, double2 *__restrict__ Fhat, int n1, int n2, int N1,
                                    int N2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    const int b = blockIdx.z;
    if (k >= n2) return;
    const int l1 = i < (n1 + 1) / 2 ? i : i - n1 + N1;
    const int l2 = k < (n2 + 1) / 2 ? k : k - n2 + N2;
    const double2 u = U[(int64_t(b) * N1 + l1) * N2 + l2];
    const double2 p = pc1[i];  // conj(p) for the adjoint
    const double s = c2[k];
    Fhat[(int64_t(b) * n1 + i) * n2 + k] = make_double2((u.x * p.x + u.y * p.y) * s, (u.y * p.x - u.x * p.y) * s);
}

// The apply reaches the fine grid from the order-column layout F^[b][m + mmax][k1] (k1 < n1 in FFT
// order) of the spectral periodisation; `scale` carries the longitude FFT pair that path skips (nph).
constexpr int kDT = 32;

__device__ __forceinline__ int fine_to_mode(int l, int n, int N) {
    if (l < (n + 1) / 2) return l;
    if (l >= N - n / 2) return l - N + n;
    return -1;
}

// ------------------------------------------------------------------------------------------------
// Pruned fine-grid transform. After the deconvolution only the 2 mmax + 1 fine columns of retained
// orders are nonzero, so the fine 2-D FFT runs as a colatitude FFT over those columns alone (order-
// column layout Ut [b][m + mmax][l1], contiguous) and a longitude FFT over all rows, with a
// transpose that inserts the zero columns between them (synthesis), or drops them (adjoint).
// ------------------------------------------------------------------------------------------------
// Ut[b][c][l1] = F^[b][c][k1(l1)] pc1[k1] c2[m] scale, zero in the colatitude padding.
template <class C>
__global__ void deconv_pad_ut_kernel(const C *__restrict__ Fc, const double2 *__restrict__ pc1,
                                     const double *__restrict__ c2, C *__restrict__ Ut, int n1, int mmax, int nph,
                                     int N1, double scale) {
    const int l1 = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    const int b = blockIdx.z;
    if (l1 >= N1) return;
    const int64_t col = int64_t(b) * (2 * mmax + 1) + c;
    const int i = fine_to_mode(l1, n1, N1);
    double2 v = make_double2(0.0, 0.0);
    if (i >= 0) {
        const int m = c - mmax;
        const double2 F = ldc(Fc + col * n1 + i);
        const double2 p = pc1[i];
        const double sc = c2[m >= 0 ? m : m + nph] * scale;
        v = make_double2((F.x * p.x - F.y * p.y) * sc, (F.x * p.y + F.y * p.x) * sc);
    }
    stc(Ut + col * N1 + l1, v);
}

// F^[b][c][k1] = Ut[b][c][l1(k1)] conj(pc1[k1]) c2[m] scale.
template <class C>
__global__ void deconv_trunc_ut_kernel(const C *__restrict__ Ut, const double2 *__restrict__ pc1,
                                       const double *__restrict__ c2, C *__restrict__ Fc, int n1, int mmax, int nph,
                                       int N1, double scale) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    const int b = blockIdx.z;
    if (i >= n1) return;
    const int64_t col = int64_t(b) * (2 * mmax + 1) + c;
    const int m = c - mmax;
    const int l1 = i < (n1 + 1) / 2 ? i : i - n1 + N1;
    const double2 u = ldc(Ut + col * N1 + l1);
    const double2 p = pc1[i];  // conj(p) for the adjoint
    const double sc = c2[m >= 0 ? m : m + nph] * scale;
    stc(Fc + col * n1 + i, make_double2((u.x * p.x + u.y * p.y) * sc, (u.y * p.x - u.x * p.y) * sc));
}

// U[b][pr][l2] = Ut[b][c(l2)][(r0 + pr) mod N1] for the fine columns of retained orders, zero
// elsewhere, for the stored rows pr < nR.
template <class C>
__global__ void __launch_bounds__(256)
ut_to_rows_kernel(const C *__restrict__ Ut, C *__restrict__ U, int mmax, int N1, int N2, int r0, int nR) {
    __shared__ C tile[kDT][kDT + 1];
    const int l2_0 = blockIdx.x * kDT, pr_0 = blockIdx.y * kDT, b = blockIdx.z;
    const int tx = threadIdx.x % kDT, ty = threadIdx.x / kDT;
    const int64_t ncol = 2 * mmax + 1;
    for (int rr = ty; rr < kDT; rr += 8) {
        const int l2 = l2_0 + rr, pr = pr_0 + tx;
        C v = {};
        if (l2 < N2 && pr < nR) {
            const int m = l2 <= mmax ? l2 : (l2 >= N2 - mmax ? l2 - N2 : INT32_MIN);
            const int l1 = (r0 + pr) % N1;
            if (m != INT32_MIN) v = Ut[(int64_t(b) * ncol + (m + mmax)) * N1 + l1];
        }
        tile[rr][tx] = v;
    }
    __syncthreads();
    for (int rr = ty; rr < kDT; rr += 8) {
        const int pr = pr_0 + rr, l2 = l2_0 + tx;
        if (pr < nR && l2 < N2) U[(int64_t(b) * nR + pr) * N2 + l2] = tile[tx][rr];
    }
}

// Ut[b][c][l1] = U[b][l1][l2(c)]: the transpose, keeping only the fine columns of retained orders;
// rows that are not stored are zero.
template <class C>
__global__ void __launch_bounds__(256)
rows_to_ut_kernel(const C *__restrict__ U, C *__restrict__ Ut, int mmax, int N1, int N2, int r0, int nR) {
    __shared__ C tile[kDT][kDT + 1];
    const int c0 = blockIdx.x * kDT, l1_0 = blockIdx.y * kDT, b = blockIdx.z;
    const int tx = threadIdx.x % kDT, ty = threadIdx.x / kDT;
    const int ncol = 2 * mmax + 1;
    for (int rr = ty; rr < kDT; rr += 8) {
        const int l1 = l1_0 + rr, c = c0 + tx;
        C v = {};
        int pr = l1 - r0;
        if (pr < 0) pr += N1;
        if (l1 < N1 && c < ncol && pr < nR) {
            const int m = c - mmax;
            v = U[(int64_t(b) * nR + pr) * N2 + (m >= 0 ? m : m + N2)];
        }
        tile[rr][tx] = v;
    }
    __syncthreads();
    for (int rr = ty; rr < kDT; rr += 8) {
        const int c = c0 + rr, l1 = l1_0 + tx;
        if (c < ncol && l1 < N1) Ut[(int64_t(b) * ncol + c) * N1 + l1] = tile[tx][rr];
    }
}

template <class C>
void launch_deconv_pad_ut(const NufftDev &u, const double2 *pc1, int mmax, int nph, double scale, const C *Fc, C *Ut,
                          hipStream_t s) {
    dim3 grid((u.N1 + 255) / 256, 2 * mmax + 1, u.B);
    deconv_pad_ut_kernel<C><<<grid, 256, 0, s>>>(Fc, pc1, u.c2, Ut, u.n1, mmax, nph, u.N1, scale);
}

template <class C>
void launch_deconv_trunc_ut(const NufftDev &u, const double2 *pc1, int mmax, int nph, double scale, const C *Ut, C *Fc,
                            hipStream_t s) {
    dim3 grid((u.n1 + 255) / 256, 2 * mmax + 1, u.B);
    deconv_trunc_ut_kernel<C><<<grid, 256, 0, s>>>(Ut, pc1, u.c2, Fc, u.n1, mmax, nph, u.N1, scale);
}

template <class C>
void launch_ut_to_rows(const NufftDev &u, int mmax, const C *Ut, C *U, hipStream_t s) {
    dim3 grid((u.N2 + kDT - 1) / kDT, (u.nR + kDT - 1) / kDT, u.B);
    ut_to_rows_kernel<C><<<grid, 256, 0, s>>>(Ut, U, mmax, u.N1, u.N2, u.r0, u.nR);
}

template <class C>
void launch_rows_to_ut(const NufftDev &u, int mmax, const C *U, C *Ut, hipStream_t s) {
    dim3 grid((2 * mmax + 1 + kDT - 1) / kDT, (u.N1 + kDT - 1) / kDT, u.B);
    rows_to_ut_kernel<C><<<grid, 256, 0, s>>>(U, Ut, mmax, u.N1, u.N2, u.r0, u.nR);
}

void launch_deconv_pad(const NufftDev &u, const double2 *pc1, const double2 *Fhat, double2 *U, hipStream_t s) {
    dim3 grid((u.N2 + 255) / 256, u.N1, u.B);
    deconv_pad_kernel<<<grid, 256, 0, s>>>(Fhat, pc1, u.c2, U, u.n1, u.n2, u.N1, u.N2);
}

void launch_deconv_trunc(const NufftDev &u, const double2 *pc1, const double2 *U, double2 *Fhat, hipStream_t s) {
    dim3 grid((u.n2 + 255) / 256, u.n1, u.B);
    deconv_trunc_kernel<<<grid, 256, 0, s>>>(U, pc1, u.c2, Fhat, u.n1, u.n2, u.N1, u.N2);
}

// v mod n for 0 <= v (a super-tile origin plus a window offset: one subtraction at most, unless the
// window is wider than the grid), without a run-time integer division
__device__ __forceinline__ int wrap_index(int v, int n) {
    while (v >= n) v -= n;
    return v;
}

// ------------------------------------------------------------------------------------------------
// Interpolation (type 2). One or two transforms: one workgroup per (non-empty super-tile,
// transform); the super-tile's window of the fine grid (4 S + w - 1 cells square, wrapped
// periodically) is staged in LDS with coalesced row reads, then each thread interpolates bin-sorted
// points of the super-tile from LDS, evaluating each point's kernel rows.
// ------------------------------------------------------------------------------------------------
// GROUP: one launch over many single-transform plans; workgroup i takes super-tile work[i].y of
// plan work[i].x, whose tables come from tabs[work[i].x].
template <int W, class C, bool GROUP = false>
__global__ void __launch_bounds__(256)
interp_st_kernel(const double *__restrict__ X1, const double *__restrict__ X2, const int32_t *__restrict__ perm,
                 const C *__restrict__ U, OutCols f, const int32_t *__restrict__ st_list,
                 const int32_t *__restrict__ bin_ofs, int N1, int N2, int S4, int ns2, int win, int r0, int nR,
                 const NufftTab *__restrict__ tabs = nullptr, const int2 *__restrict__ work = nullptr) {
    extern __shared__ __align__(16) unsigned char lds_ip_raw[];
    C *lw = reinterpret_cast<C *>(lds_ip_raw);
    int slot = blockIdx.x, b = blockIdx.y;
    if constexpr (GROUP) {
        const int2 w = work[blockIdx.x];
        const NufftTab &T = tabs[w.x];
        X1 = T.X1; X2 = T.X2; perm = T.perm; U = static_cast<const C *>(T.U); f = T.fout; st_list = T.st_list;
        bin_ofs = T.bin_ofs; N1 = T.N1; N2 = T.N2; S4 = 4 * T.S; ns2 = T.ns2; win = T.win; r0 = T.r0; nR = T.nR;
        slot = w.y;
        b = 0;
    }
    const int st = st_list[slot];
    const int x0 = (st / ns2) * S4, y0 = (st % ns2) * S4;
    const C *__restrict__ Ub = U + int64_t(b) * nR * N2;
    const gptr fcol = f.col(b);
    for (int r = threadIdx.x >> 6; r < win; r += 4) {
        int pr = wrap_index(x0 + r, N1) - r0;
        if (pr < 0) pr += N1;
        if (pr >= nR) continue;  // a row that no point's kernel reaches is neither stored nor read
        for (int c = threadIdx.x & 63; c < win; c += 64) lw[r * win + c] = Ub[int64_t(pr) * N2 + wrap_index(y0 + c, N2)];
    }
    __syncthreads();
    const int start = bin_ofs[slot * 17], end = bin_ofs[slot * 17 + 16];
    for (int p = start + int(threadIdx.x); p < end; p += 256) {
        const double x1 = X1[p], x2 = X2[p];
        const int b1 = int(ceil(x1 - 0.5 * W));
        const int b2 = int(ceil(x2 - 0.5 * W));
        double k1[W], k2[W];
        es_rows<W>(x1, b1, x2, b2, k1, k2);
        const C *__restrict__ wrow = lw + (b1 - x0) * win + (b2 - y0);
        double ar = 0.0, ai = 0.0;
#pragma unroll
        for (int t1 = 0; t1 < W; ++t1) {
            double rr = 0.0, ri = 0.0;
#pragma unroll
            for (int t2 = 0; t2 < W; ++t2) {
                const double2 v = ldc(wrow + t1 * win + t2);
                rr = fma(k2[t2], v.x, rr);
                ri = fma(k2[t2], v.y, ri);
            }
            ar = fma(k1[t1], rr, ar);
            ai = fma(k1[t1], ri, ai);
        }
        gstore(fcol + perm[p], make_double2(ar, ai));
    }
}

// Three or more transforms: one workgroup of 512 threads per (non-empty super-tile, group of up to
// kt transforms). Each thread takes one point of the super-tile at a time and evaluates its kernel
// rows once for the whole group; the transforms' windows go through two LDS buffers, the next one's
// cells loaded into registers (up to kIpCells per thread) while the current one is interpolated,
// so the loads overlap the arithmetic and one barrier per transform suffices.
constexpr int kIpThreads = 512;
// window cells per thread: the window edge is win = 3 S + 16 with S = 17 - W (the plan's sizing)
constexpr int ip_cells(int W) { return ((3 * (17 - W) + 16) * (3 * (17 - W) + 16) + kIpThreads - 1) / kIpThreads; }

template <int W, class C>
__global__ void __launch_bounds__(kIpThreads)
interp_st_multi_kernel(const double *__restrict__ X1, const double *__restrict__ X2,
                       const int32_t *__restrict__ perm, const C *__restrict__ U, OutCols f,
                       const int32_t *__restrict__ st_list, const int32_t *__restrict__ bin_ofs, int N1, int N2,
                       int S4, int ns2, int win, int r0, int nR, int nb, int kt) {
    extern __shared__ __align__(16) unsigned char lds_ipm_raw[];
    C *lw0 = reinterpret_cast<C *>(lds_ipm_raw);
    const int slot = blockIdx.x;
    const int bfirst = blockIdx.y * kt, blast = min(nb, bfirst + kt);
    const int st = st_list[slot];
    const int x0 = (st / ns2) * S4, y0 = (st % ns2) * S4;
    const int wsz = win * win;
    constexpr int kIpCells = ip_cells(W);
    // this thread's window cells q = tid + kIpThreads i: their offset in a transform's grid (or -1)
    int64_t src[kIpCells];
#pragma unroll
    for (int i = 0; i < kIpCells; ++i) {
        const int q = int(threadIdx.x) + kIpThreads * i;
        src[i] = -1;
        if (q < wsz) {
            const int r = q / win, c = q - r * win;
            int pr = (x0 + r) % N1 - r0;
            if (pr < 0) pr += N1;
            if (pr < nR) src[i] = int64_t(pr) * N2 + (y0 + c) % N2;  // rows no kernel reaches stay unread
        }
    }
    using E = decltype(C::x);  // the component type, kept in plain registers
    E rx[kIpCells], ry[kIpCells];
    auto fetch = [&](int b) {
        const C *__restrict__ Ub = U + int64_t(b) * nR * N2;
#pragma unroll
        for (int i = 0; i < kIpCells; ++i) {
            const C v = Ub[src[i] >= 0 ? src[i] : 0];
            rx[i] = v.x;
            ry[i] = v.y;
        }
    };
    auto put = [&](C *lw) {
#pragma unroll
        for (int i = 0; i < kIpCells; ++i)
            if (src[i] >= 0) {
                C v;
                v.x = rx[i];
                v.y = ry[i];
                lw[threadIdx.x + kIpThreads * i] = v;
            }
    };
    const int start = bin_ofs[slot * 17], end = bin_ofs[slot * 17 + 16];
    int it = 0;  // windows staged so far; buffer it % 2 is the current one
#pragma unroll 1
    for (int c0 = start; c0 < end; c0 += kIpThreads) {
        const int p = c0 + int(threadIdx.x);
        const bool valid = p < end;
        const int pc = valid ? p : end - 1;
        const double x1 = X1[pc], x2 = X2[pc];
        const int b1 = int(ceil(x1 - 0.5 * W));
        const int b2 = int(ceil(x2 - 0.5 * W));
        double k1[W], k2[W];
        if (valid) es_rows<W>(x1, b1, x2, b2, k1, k2);
        const int off = (b1 - x0) * win + (b2 - y0);
        const int64_t fi = perm[pc];
        fetch(bfirst);
        put(lw0 + (it & 1) * wsz);
        __syncthreads();
#pragma unroll 1
        for (int b = bfirst; b < blast; ++b, ++it) {
            const C *lw = lw0 + (it & 1) * wsz;
            if (b + 1 < blast) fetch(b + 1);  // in flight during this transform's interpolation
            const gptr fcol = f.col(b);  // uniform: one scalar address per transform
            if (valid) {
                const C *__restrict__ wrow = lw + off;
                double ar = 0.0, ai = 0.0;
#pragma unroll
                for (int t1 = 0; t1 < W; ++t1) {
                    double rr = 0.0, ri = 0.0;
#pragma unroll
                    for (int t2 = 0; t2 < W; ++t2) {
                        const double2 v = ldc(wrow + t1 * win + t2);
                        rr = fma(k2[t2], v.x, rr);
                        ri = fma(k2[t2], v.y, ri);
                    }
                    ar = fma(k1[t1], rr, ar);
                    ai = fma(k1[t1], ri, ai);
                }
                gstore(fcol + fi, make_double2(ar, ai));
            }
            if (b + 1 < blast) put(lw0 + ((it + 1) & 1) * wsz);
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Spreading (type 1): MFMA super-tiles + ordered merge.
// ------------------------------------------------------------------------------------------------
// Spreading: one workgroup per (non-empty super-tile, group of G transforms); wave w owns bins
// w + kSpreadWaves k and takes each bin's points 16 at a time. For such a chunk the wave first
// evaluates the kernel rows once per point, for all G transforms: lane (p, axis, h) evaluates the
// cells t < ceil(W/2) of its point and axis from the left (h = 0) or, through phi(-z) = phi(z), from
// the right (h = 1), so all lanes use the same polynomial coefficients (scalar operands). The rows
// go to a wave-private LDS staging area at the point's offset within its bin's 16-cell window, zero
// elsewhere, with the G values of f at the point. The wave then accumulates the chunk on FP64 MFMA,
// four points per instruction (A = f psi_row, B = psi_col), keeping each of its bins' 16 x 16 tiles
// per transform in registers. The inputs of the next chunk (its coordinates and its f values) are
// loaded while the current chunk is evaluated and accumulated. No workgroup barrier until the end,
// where, transform by transform, the tiles are added into the LDS window wave after wave (bin
// windows overlap) and the window is written out for the merge pass.
// GROUP (G = 1): as the interpolation's.
template <int W, int G, class C, bool GROUP = false>
__global__ void __launch_bounds__(64 * kSpreadWaves)
spread_st_kernel(const double *__restrict__ X1, const double *__restrict__ X2, const int32_t *__restrict__ perm,
                 InCols f, const int32_t *__restrict__ st_list, const int32_t *__restrict__ bin_ofs,
                 C *__restrict__ Wst, int S, int ns2, int win, int nst, int nb,
                 const NufftTab *__restrict__ tabs = nullptr, const int2 *__restrict__ work = nullptr) {
    constexpr int NH = (W + 1) / 2;  // cells per half row
    extern __shared__ double lds_sp[];
    double2 *lw = reinterpret_cast<double2 *>(lds_sp);  // [win][win]
    int slot = blockIdx.x, b0 = blockIdx.y * G;
    if constexpr (GROUP) {
        const int2 w = work[blockIdx.x];
        const NufftTab &T = tabs[w.x];
        X1 = T.X1; X2 = T.X2; perm = T.perm; f = T.fin; st_list = T.st_list; bin_ofs = T.bin_ofs;
        Wst = static_cast<C *>(T.Wst); S = T.S; ns2 = T.ns2; win = T.win; nst = T.nst; nb = 1;
        slot = w.y;
        b0 = 0;
    }
    const int gc = min(G, nb - b0);  // transforms of this workgroup
    const int st = st_list[slot];
    const int S4 = 4 * S;
    const int x0 = (st / ns2) * S4, y0 = (st % ns2) * S4;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // rows are kRowPad doubles apart so that the 16 lanes writing one entry of 16 rows, and the
    // MFMA operand reads, fall in different LDS banks
    double *K1 = lds_sp + 2 * win * win + wave * spread_stage_doubles(G);  // [16][kRowPad] row kernel
    double *K2 = K1 + 16 * kRowPad;                                      // [16][kRowPad] column kernel
    double2 *Fc = reinterpret_cast<double2 *>(K2 + 16 * kRowPad);         // [G][16]
    const int ii = lane & 15, kq = lane >> 4;                            // MFMA operand layout
    const int pa = lane & 15, axis = (lane >> 4) & 1, h = lane >> 5;     // kernel-row evaluation
    const double *__restrict__ Xa = axis == 0 ? X1 : X2;
    gcptr fb[G];
#pragma unroll
    for (int g = 0; g < G; ++g) fb[g] = f.col(b0 + min(g, gc - 1));
    for (int q = tid; q < win * win; q += 64 * kSpreadWaves) lw[q] = make_double2(0.0, 0.0);
    const int32_t *__restrict__ bo = bin_ofs + slot * 17;
    constexpr int NBW = 16 / kSpreadWaves;  // bins per wavefront; tile k accumulates bin wave + kSpreadWaves k
    double4v dr[NBW][G], di[NBW][G];
#pragma unroll
    for (int k = 0; k < NBW; ++k)
#pragma unroll
        for (int g = 0; g < G; ++g) {
            dr[k][g] = double4v{0.0, 0.0, 0.0, 0.0};
            di[k][g] = double4v{0.0, 0.0, 0.0, 0.0};
        }
    int bs[NBW], be[NBW];  // point range [bs[k], be[k]) of this wave's bin k
#pragma unroll
    for (int k = 0; k < NBW; ++k) {
        bs[k] = bo[wave + kSpreadWaves * k];
        be[k] = bo[wave + kSpreadWaves * k + 1];
    }
    // Chunk cursors: (k1, c1) is the next chunk (bin k1, first point c1), (k2, c2) the one after;
    // k = NBW when none is left. The next chunk's coordinate and f values are in flight while a
    // chunk is worked on, and so is the point index of the one after, so the f gathers never wait on
    // the index load. The loads are unconditional (a lane past a chunk's end reads the super-tile's
    // last point and its value is never used), so that waiting for one chunk's inputs never waits
    // for the loads issued after them.
    auto advance = [&](int &k, int &c) {
        while (k < NBW && c >= be[k]) {
            ++k;
            if (k < NBW) c = bs[k];
        }
    };
    const int plast = max(bo[16] - 1, 0);
    int k1 = 0, c1 = bs[0];
    advance(k1, c1);
    int k2 = k1, c2 = c1 + 16;
    advance(k2, c2);
    double x1 = Xa[min(c1 + pa, plast)];
    double2 f1[G];
    {
        const int q1 = perm[min(c1 + pa, plast)];
#pragma unroll
        for (int g = 0; g < G; ++g) f1[g] = gload(fb[g] + q1);
    }
    int p2 = perm[min(c2 + pa, plast)];
    const typename HalfRow<W>::Edge ec = HalfRow<W>::edge();
    // Bins in a static order, so each accumulates into a fixed register tile.
#pragma unroll
    for (int kb = 0; kb < NBW; ++kb) {
        // the bin's first base cell along each axis (wave-uniform): a point's base cell lies in its
        // bin, as the plan keyed it, so its offset within the bin window is the difference
        const int lbin = wave + kSpreadWaves * kb;
        const int bx0 = x0 + (lbin >> 2) * S, by0 = y0 + (lbin & 3) * S;
        while (k1 == kb) {  // wave-uniform
            const int n = min(16, be[kb] - c1);
            const double x = x1;
            double2 fv[G];
#pragma unroll
            for (int g = 0; g < G; ++g) fv[g] = f1[g];
            // shift the cursors by one chunk and issue the loads of the new pair
            k1 = k2;
            c1 = c2;
            c2 += 16;
            advance(k2, c2);
            x1 = Xa[min(c1 + pa, plast)];
#pragma unroll
            for (int g = 0; g < G; ++g) f1[g] = gload(fb[g] + p2);
            p2 = perm[min(c2 + pa, plast)];
            // Staging rows of every lane are cleared, so rows and f of points q >= n are zero and
            // their products add exact zeros; lanes of points q < n then write their kernel values.
            double *row = (axis == 0 ? K1 : K2) + pa * kRowPad;
#pragma unroll
            for (int i = 0; i < 8; ++i) row[8 * h + i] = 0.0;
            if (pa < n) {
                const int bc = int(ceil(x - 0.5 * W));
                const int o = bc - (axis == 0 ? bx0 : by0);  // offset within the bin window, [0, S)
                const double u = 2.0 * (double(bc) - x) + double(W - 1);
                double kv[NH];
                es_half_row<W>(h == 0 ? u : -u, kv, ec);
                // the middle cell of an odd W is written by the left half only
#pragma unroll
                for (int t = 0; t < NH; ++t)
                    if (h == 0) row[o + t] = kv[t];
                    else if (W - 1 - t != t) row[o + W - 1 - t] = kv[t];
            }
            if (axis == 0 && h == 0)
#pragma unroll
                for (int g = 0; g < G; ++g) Fc[g * 16 + pa] = pa < n ? fv[g] : make_double2(0.0, 0.0);
            wave_lds_order();
            // all operands of the chunk first (one LDS wait), then up to four MFMA pairs per transform
            double av[4], cv[4];
#pragma unroll
            for (int gi = 0; gi < 4; ++gi) {
                const int q = 4 * gi + kq;
                av[gi] = K1[q * kRowPad + ii];
                cv[gi] = K2[q * kRowPad + ii];
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (g >= gc) break;  // wave-uniform
                double2 fq[4];
#pragma unroll
                for (int gi = 0; gi < 4; ++gi) fq[gi] = Fc[g * 16 + 4 * gi + kq];
#pragma unroll
                for (int gi = 0; gi < 4; ++gi) {
                    if (4 * gi >= n) break;  // wave-uniform
                    dr[kb][g] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[gi] * fq[gi].x, cv[gi], dr[kb][g], 0, 0, 0);
                    di[kb][g] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[gi] * fq[gi].y, cv[gi], di[kb][g], 0, 0, 0);
                }
            }
            wave_lds_order();
        }
    }
    __syncthreads();
    // D[i][j] of v_mfma_f64_16x16x4f64: lane holds column j = lane % 16, register r holds row
    // i = lane / 16 + 4 r (checked on gfx950 with a one-wave layout probe). Bin windows overlap
    // (16 > S), so the waves add into LDS one at a time, in wave order, each wave its bins in
    // order: race-free and a fixed summation order. One transform after another.
#pragma unroll
    for (int g = 0; g < G; ++g) {
        if (g >= gc) break;  // uniform over the workgroup
#pragma unroll 1
        for (int turn = 0; turn < kSpreadWaves; ++turn) {
            if (wave == turn) {
#pragma unroll
                for (int k = 0; k < NBW; ++k) {
                    const int lb = wave + kSpreadWaves * k;
                    const int bi = lb >> 2, bj = lb & 3;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        double2 &cell = lw[(bi * S + kq + 4 * r) * win + bj * S + ii];
                        cell.x += dr[k][g][r];
                        cell.y += di[k][g][r];
                    }
                }
            }
            __syncthreads();
        }
        C *out = Wst + (int64_t(b0 + g) * nst + slot) * int64_t(win) * win;
        for (int q = tid; q < win * win; q += 64 * kSpreadWaves) {
            stc(out + q, lw[q]);
            lw[q] = make_double2(0.0, 0.0);  // each thread clears the cells it wrote out
        }
        __syncthreads();
    }
}

// GROUP: workgroup i merges column block work[i].y of stored row work[i].z of plan work[i].x.
template <class C, bool GROUP = false>
__global__ void spread_merge_kernel(const C *__restrict__ Wst, const int32_t *__restrict__ st_slot,
                                    C *__restrict__ U, int N1, int N2, int S4, int ns1, int ns2, int win, int w,
                                    int nst, int r0, int nR, const NufftTab *__restrict__ tabs = nullptr,
                                    const int4 *__restrict__ work = nullptr) {
    int cb = blockIdx.x, pr = blockIdx.y, b = blockIdx.z;  // column block, stored row, transform
    if constexpr (GROUP) {
        const int4 q = work[blockIdx.x];
        const NufftTab &T = tabs[q.x];
        Wst = static_cast<const C *>(T.Wst); st_slot = T.st_slot; U = static_cast<C *>(T.U); N1 = T.N1; N2 = T.N2;
        S4 = 4 * T.S; ns1 = T.ns1; ns2 = T.ns2; win = T.win; w = T.w; nst = T.nst; r0 = T.r0; nR = T.nR;
        cb = q.y;
        pr = q.z;
        b = 0;
    }
    const int l2 = cb * blockDim.x + threadIdx.x;
    int l1 = r0 + pr;
    if (l1 >= N1) l1 -= N1;  // r0 < N1, pr < N1
    if (l2 >= N2) return;
    const int q1 = l1 / S4, q2 = l2 / S4;
    // Wraps by one period at most: q - d >= -(ns - 1) for d < min(3, ns), and the offset of a cell
    // from a super-tile origin lies in (-(ns - 1) S4, N) with (ns - 1) S4 < N.
    double sr = 0.0, si = 0.0;
    for (int d1 = 0; d1 < 3 && d1 < ns1; ++d1) {
        int s1 = q1 - d1;
        if (s1 < 0) s1 += ns1;
        const int wh1 = min(S4, N1 - s1 * S4) + w - 1;
        int o1 = l1 - s1 * S4;
        if (o1 < 0) o1 += N1;
        for (; o1 < wh1; o1 += N1) {
            for (int d2 = 0; d2 < 3 && d2 < ns2; ++d2) {
                int s2 = q2 - d2;
                if (s2 < 0) s2 += ns2;
                const int slot = st_slot[s1 * ns2 + s2];
                if (slot < 0) continue;
                const int wh2 = min(S4, N2 - s2 * S4) + w - 1;
                int o2 = l2 - s2 * S4;
                if (o2 < 0) o2 += N2;
                for (; o2 < wh2; o2 += N2) {
                    const double2 v = ldc(Wst + ((int64_t(b) * nst + slot) * win + o1) * win + o2);
                    sr += v.x;
                    si += v.y;
                }
            }
        }
    }
    stc(U + (int64_t(b) * nR + pr) * N2 + l2, make_double2(sr, si));
}

// ------------------------------------------------------------------------------------------------
// FP32 arithmetic for mixed-precision plans (tol >= 1e-6, so kernel width w <= 8; FP32 grids). The
// kernel values come from EsHorner32 by Horner in float (one v_fmaak_f32 per step, the coefficient a
// literal operand: no scalar moves); interpolation sums the FP32 window in float; spreading
// accumulates on v_mfma_f32_16x16x4f32 into FP32 tiles and an FP32 LDS window. The coordinates and
// base cells stay FP64 (the cell index is exact); the offset u within the cell is rounded to float.
// The fields in the caller's memory stay complex FP64 (loads narrow, stores widen).
// ------------------------------------------------------------------------------------------------
template <int W, int OFF, int N>
__device__ __forceinline__ float horner32(float x) {
    using T = EsHorner32<W>;
    float a = T::c[OFF + N - 1];
#pragma unroll
    for (int i = N - 2; i >= 0; --i) a = fmaf(a, x, T::c[OFF + i]);
    return a;
}

// edge-polynomial argument e = 2 sqrt((1 + v) / 2) - 1 (v_sqrt_f32: 1 ulp)
__device__ __forceinline__ float es_edge_arg32(float v) {
    return 2.0f * __builtin_amdgcn_sqrtf(0.5f * (1.0f + v)) - 1.0f;
}

template <int W, int... T>
__device__ __forceinline__ void es_interior32(float u, float (&k)[W], std::integer_sequence<int, T...>) {
    constexpr int ni = EsHorner32<W>::nint;
    ((k[T + 1] = horner32<W, T * ni, ni>(u)), ...);
}

// All W kernel values of a point on one axis (cell t at k[t]), as es_rows does in FP64.
template <int W>
__device__ __forceinline__ void es_row32(float u, float (&k)[W]) {
    using T = EsHorner32<W>;
    constexpr int EOFF = (W - 2) * T::nint;
    k[0] = horner32<W, EOFF, T::nedge>(es_edge_arg32(u));
    k[W - 1] = horner32<W, EOFF, T::nedge>(es_edge_arg32(-u));
    es_interior32<W>(u, k, std::make_integer_sequence<int, W - 2>{});
}

template <int W, int... T>
__device__ __forceinline__ void es_half_interior32(float v, float (&k)[(W + 1) / 2],
                                                   std::integer_sequence<int, T...>) {
    constexpr int ni = EsHorner32<W>::nint;
    ((k[T + 1] = horner32<W, T * ni, ni>(v)), ...);
}

// Half a row (cells t < ceil(W/2)) from the left for v = u, from the right for v = -u, as es_half_row.
template <int W>
__device__ __forceinline__ void es_half_row32(float v, float (&k)[(W + 1) / 2]) {
    using T = EsHorner32<W>;
    constexpr int NH = (W + 1) / 2, EOFF = (W - 2) * T::nint;
    k[0] = horner32<W, EOFF, T::nedge>(es_edge_arg32(v));
    es_half_interior32<W>(v, k, std::make_integer_sequence<int, NH - 1>{});
}

// Interpolation in FP32, one or two transforms: as interp_st_kernel (one workgroup per super-tile
// and transform, the window staged in LDS), with the window's float2 cells summed in float.
template <int W, bool GROUP = false>
__global__ void __launch_bounds__(256)
interp_st32_kernel(const double *__restrict__ X1, const double *__restrict__ X2, const int32_t *__restrict__ perm,
                   const float2 *__restrict__ U, OutCols f, const int32_t *__restrict__ st_list,
                   const int32_t *__restrict__ bin_ofs, int N1, int N2, int S4, int ns2, int win, int r0, int nR,
                   const NufftTab *__restrict__ tabs = nullptr, const int2 *__restrict__ work = nullptr) {
    extern __shared__ __align__(16) unsigned char lds_ip32_raw[];
    float2 *lw = reinterpret_cast<float2 *>(lds_ip32_raw);
    int slot = blockIdx.x, b = blockIdx.y;
    if constexpr (GROUP) {
        const int2 w = work[blockIdx.x];
        const NufftTab &T = tabs[w.x];
        X1 = T.X1; X2 = T.X2; perm = T.perm; U = static_cast<const float2 *>(T.U); f = T.fout;
        st_list = T.st_list; bin_ofs = T.bin_ofs; N1 = T.N1; N2 = T.N2; S4 = 4 * T.S; ns2 = T.ns2; win = T.win;
        r0 = T.r0; nR = T.nR;
        slot = w.y;
        b = 0;
    }
    const int st = st_list[slot];
    const int x0 = (st / ns2) * S4, y0 = (st % ns2) * S4;
    const float2 *__restrict__ Ub = U + int64_t(b) * nR * N2;
    const gptr fcol = f.col(b);
    for (int r = threadIdx.x >> 6; r < win; r += 4) {
        int pr = wrap_index(x0 + r, N1) - r0;
        if (pr < 0) pr += N1;
        if (pr >= nR) continue;  // a row that no point's kernel reaches is neither stored nor read
        for (int c = threadIdx.x & 63; c < win; c += 64) lw[r * win + c] = Ub[int64_t(pr) * N2 + wrap_index(y0 + c, N2)];
    }
    __syncthreads();
    const int start = bin_ofs[slot * 17], end = bin_ofs[slot * 17 + 16];
    for (int p = start + int(threadIdx.x); p < end; p += 256) {
        const double x1 = X1[p], x2 = X2[p];
        const int b1 = int(ceil(x1 - 0.5 * W));
        const int b2 = int(ceil(x2 - 0.5 * W));
        float k1[W], k2[W];
        es_row32<W>(float(2.0 * (double(b1) - x1) + double(W - 1)), k1);
        es_row32<W>(float(2.0 * (double(b2) - x2) + double(W - 1)), k2);
        const float2 *__restrict__ wrow = lw + (b1 - x0) * win + (b2 - y0);
        float ar = 0.0f, ai = 0.0f;
#pragma unroll
        for (int t1 = 0; t1 < W; ++t1) {
            float rr = 0.0f, ri = 0.0f;
#pragma unroll
            for (int t2 = 0; t2 < W; ++t2) {
                const float2 v = wrow[t1 * win + t2];
                rr = fmaf(k2[t2], v.x, rr);
                ri = fmaf(k2[t2], v.y, ri);
            }
            ar = fmaf(k1[t1], rr, ar);
            ai = fmaf(k1[t1], ri, ai);
        }
        gstore(fcol + perm[p], make_double2(double(ar), double(ai)));
    }
}

typedef float float4v __attribute__((ext_vector_type(4)));

// Spreading in FP32 for G transforms per workgroup: the structure of spread_st_kernel (bins per
// wavefront, 16-point chunks with the next chunk's inputs in flight, zeroed staging rows, wave-ordered
// adds into the LDS window), with FP32 kernel rows, FP32 MFMA tiles and an FP32 window.
// D[i][j] of v_mfma_f32_16x16x4f32: lane 16 q + j, register r holds row i = 4 q + r (A and B as for
// the FP64 instruction: A[i][k] in lane 16 k + i, B[k][j] in lane 16 k + j; checked on gfx950 with a
// one-wave layout probe).
template <int W, int G, bool GROUP = false>
__global__ void __launch_bounds__(64 * kSpreadWaves)
spread_st32_kernel(const double *__restrict__ X1, const double *__restrict__ X2, const int32_t *__restrict__ perm,
                   InCols f, const int32_t *__restrict__ st_list, const int32_t *__restrict__ bin_ofs,
                   float2 *__restrict__ Wst, int S, int ns2, int win, int nst, int nb,
                   const NufftTab *__restrict__ tabs = nullptr, const int2 *__restrict__ work = nullptr) {
    constexpr int NH = (W + 1) / 2;
    extern __shared__ float lds_sp32[];
    float2 *lw = reinterpret_cast<float2 *>(lds_sp32);  // [win][win]
    int slot = blockIdx.x, b0 = blockIdx.y * G;
    if constexpr (GROUP) {
        const int2 w = work[blockIdx.x];
        const NufftTab &T = tabs[w.x];
        X1 = T.X1; X2 = T.X2; perm = T.perm; f = T.fin; st_list = T.st_list; bin_ofs = T.bin_ofs;
        Wst = static_cast<float2 *>(T.Wst); S = T.S; ns2 = T.ns2; win = T.win; nst = T.nst; nb = 1;
        slot = w.y;
        b0 = 0;
    }
    const int gc = min(G, nb - b0);
    const int st = st_list[slot];
    const int S4 = 4 * S;
    const int x0 = (st / ns2) * S4, y0 = (st % ns2) * S4;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    float *K1 = lds_sp32 + 2 * win * win + wave * spread_stage_floats(G);  // [16][kRowPad32] row kernel
    float *K2 = K1 + 16 * kRowPad32;                                      // [16][kRowPad32] column kernel
    float2 *Fc = reinterpret_cast<float2 *>(K2 + 16 * kRowPad32);          // [G][16]
    const int ii = lane & 15, kq = lane >> 4;
    const int pa = lane & 15, axis = (lane >> 4) & 1, h = lane >> 5;
    const double *__restrict__ Xa = axis == 0 ? X1 : X2;
    gcptr fb[G];
#pragma unroll
    for (int g = 0; g < G; ++g) fb[g] = f.col(b0 + min(g, gc - 1));
    for (int q = tid; q < win * win; q += 64 * kSpreadWaves) lw[q] = make_float2(0.0f, 0.0f);
    const int32_t *__restrict__ bo = bin_ofs + slot * 17;
    constexpr int NBW = 16 / kSpreadWaves;
    float4v dr[NBW][G], di[NBW][G];
#pragma unroll
    for (int k = 0; k < NBW; ++k)
#pragma unroll
        for (int g = 0; g < G; ++g) {
            dr[k][g] = float4v{0.0f, 0.0f, 0.0f, 0.0f};
            di[k][g] = float4v{0.0f, 0.0f, 0.0f, 0.0f};
        }
    int bs[NBW], be[NBW];
#pragma unroll
    for (int k = 0; k < NBW; ++k) {
        bs[k] = bo[wave + kSpreadWaves * k];
        be[k] = bo[wave + kSpreadWaves * k + 1];
    }
    auto advance = [&](int &k, int &c) {
        while (k < NBW && c >= be[k]) {
            ++k;
            if (k < NBW) c = bs[k];
        }
    };
    const int plast = max(bo[16] - 1, 0);
    int k1 = 0, c1 = bs[0];
    advance(k1, c1);
    int k2 = k1, c2 = c1 + 16;
    advance(k2, c2);
    double x1 = Xa[min(c1 + pa, plast)];
    double2 f1[G];
    {
        const int q1 = perm[min(c1 + pa, plast)];
#pragma unroll
        for (int g = 0; g < G; ++g) f1[g] = gload(fb[g] + q1);
    }
    int p2 = perm[min(c2 + pa, plast)];
#pragma unroll
    for (int kb = 0; kb < NBW; ++kb) {
        // the bin's first base cell along each axis (wave-uniform): a point's base cell lies in its
        // bin, as the plan keyed it, so its offset within the bin window is the difference
        const int lbin = wave + kSpreadWaves * kb;
        const int bx0 = x0 + (lbin >> 2) * S, by0 = y0 + (lbin & 3) * S;
        while (k1 == kb) {  // wave-uniform
            const int n = min(16, be[kb] - c1);
            const double x = x1;
            float2 fv[G];
#pragma unroll
            for (int g = 0; g < G; ++g) fv[g] = make_float2(float(f1[g].x), float(f1[g].y));
            k1 = k2;
            c1 = c2;
            c2 += 16;
            advance(k2, c2);
            x1 = Xa[min(c1 + pa, plast)];
#pragma unroll
            for (int g = 0; g < G; ++g) f1[g] = gload(fb[g] + p2);
            p2 = perm[min(c2 + pa, plast)];
            float *row = (axis == 0 ? K1 : K2) + pa * kRowPad32;
#pragma unroll
            for (int i = 0; i < 8; ++i) row[8 * h + i] = 0.0f;
            if (pa < n) {
                const int bc = int(ceil(x - 0.5 * W));
                const int o = bc - (axis == 0 ? bx0 : by0);
                const float u = float(2.0 * (double(bc) - x) + double(W - 1));
                float kv[NH];
                es_half_row32<W>(h == 0 ? u : -u, kv);
#pragma unroll
                for (int t = 0; t < NH; ++t)
                    if (h == 0) row[o + t] = kv[t];
                    else if (W - 1 - t != t) row[o + W - 1 - t] = kv[t];
            }
            if (axis == 0 && h == 0)
#pragma unroll
                for (int g = 0; g < G; ++g) Fc[g * 16 + pa] = pa < n ? fv[g] : make_float2(0.0f, 0.0f);
            wave_lds_order();
            float av[4], cv[4];
#pragma unroll
            for (int gi = 0; gi < 4; ++gi) {
                const int q = 4 * gi + kq;
                av[gi] = K1[q * kRowPad32 + ii];
                cv[gi] = K2[q * kRowPad32 + ii];
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (g >= gc) break;  // wave-uniform
                float2 fq[4];
#pragma unroll
                for (int gi = 0; gi < 4; ++gi) fq[gi] = Fc[g * 16 + 4 * gi + kq];
#pragma unroll
                for (int gi = 0; gi < 4; ++gi) {
                    if (4 * gi >= n) break;  // wave-uniform
                    dr[kb][g] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[gi] * fq[gi].x, cv[gi], dr[kb][g], 0, 0, 0);
                    di[kb][g] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[gi] * fq[gi].y, cv[gi], di[kb][g], 0, 0, 0);
                }
            }
            wave_lds_order();
        }
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < G; ++g) {
        if (g >= gc) break;  // uniform over the workgroup
#pragma unroll 1
        for (int turn = 0; turn < kSpreadWaves; ++turn) {
            if (wave == turn) {
#pragma unroll
                for (int k = 0; k < NBW; ++k) {
                    const int lb = wave + kSpreadWaves * k;
                    const int bi = lb >> 2, bj = lb & 3;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        float2 &cell = lw[(bi * S + 4 * kq + r) * win + bj * S + ii];
                        cell.x += dr[k][g][r];
                        cell.y += di[k][g][r];
                    }
                }
            }
            __syncthreads();
        }
        float2 *out = Wst + (int64_t(b0 + g) * nst + slot) * int64_t(win) * win;
        for (int q = tid; q < win * win; q += 64 * kSpreadWaves) {
            out[q] = lw[q];
            lw[q] = make_float2(0.0f, 0.0f);
        }
        __syncthreads();
    }
}

// transforms per multi-transform interpolation workgroup (each evaluates the kernel rows once per point)
constexpr int kInterpTransforms = 8;

// FP32 grids (mixed-precision plans) at the widths EsHorner32 covers take the FP32 kernels.
template <int W, class C>
constexpr bool kFp32Math = std::is_same_v<C, float2> && W <= 8;

template <int W, class C>
static void interp_w(const NufftDev &u, const C *U, OutCols f, hipStream_t s) {
    if (u.nst == 0 || u.B <= 0) return;
    const size_t win_bytes = sizeof(C) * size_t(u.win) * u.win;
    if constexpr (kFp32Math<W, C>) {
        if (u.B <= 2) {
            interp_st32_kernel<W><<<dim3(u.nst, u.B), 256, win_bytes, s>>>(u.X1, u.X2, u.perm, U, f, u.st_list,
                                                                          u.bin_ofs, u.N1, u.N2, 4 * u.S, u.ns2,
                                                                          u.win, u.r0, u.nR);
            return;
        }
    }
    if (u.B <= 2) {
        interp_st_kernel<W, C><<<dim3(u.nst, u.B), 256, win_bytes, s>>>(u.X1, u.X2, u.perm, U, f, u.st_list,
                                                                       u.bin_ofs, u.N1, u.N2, 4 * u.S, u.ns2,
                                                                       u.win, u.r0, u.nR);
        return;
    }
    const int kt = std::min(u.B, kInterpTransforms);
    dim3 grid(u.nst, (u.B + kt - 1) / kt);
    interp_st_multi_kernel<W, C><<<grid, kIpThreads, 2 * win_bytes, s>>>(u.X1, u.X2, u.perm, U, f, u.st_list,
                                                                         u.bin_ofs, u.N1, u.N2, 4 * u.S, u.ns2,
                                                                         u.win, u.r0, u.nR, u.B, kt);
}

// G transforms per workgroup (grid.y = ceil(B / G) groups; the last may be partial)
template <int W, int G, class C>
static void spread_launch(const NufftDev &u, InCols f, C *Wst, hipStream_t s) {
    const dim3 grid(u.nst, (u.B + G - 1) / G);
    if constexpr (kFp32Math<W, C>)
        spread_st32_kernel<W, G><<<grid, 64 * kSpreadWaves, spread32_lds_bytes(u.win, G), s>>>(
            u.X1, u.X2, u.perm, f, u.st_list, u.bin_ofs, Wst, u.S, u.ns2, u.win, u.nst, u.B);
    else
        spread_st_kernel<W, G, C><<<grid, 64 * kSpreadWaves, spread_lds_bytes(u.win, G), s>>>(
            u.X1, u.X2, u.perm, f, u.st_list, u.bin_ofs, Wst, u.S, u.ns2, u.win, u.nst, u.B);
}

template <int W, class C>
static void spread_w(const NufftDev &u, InCols f, C *Wst, hipStream_t s) {
    static_assert(kSpreadGroup == 4, "the group sizes below are 1, 2 and kSpreadGroup");
    if (u.nst == 0 || u.B <= 0) return;
    if (u.B >= 4) spread_launch<W, 4>(u, f, Wst, s);
    else if (u.B >= 2) spread_launch<W, 2>(u, f, Wst, s);
    else spread_launch<W, 1>(u, f, Wst, s);
}

#define PFSHT_W_SWITCH(w, FN, ...)                 \
    switch (w) {                                   \
        case 2: FN<2, C>(__VA_ARGS__); break;      \
        case 3: FN<3, C>(__VA_ARGS__); break;      \
        case 4: FN<4, C>(__VA_ARGS__); break;      \
        case 5: FN<5, C>(__VA_ARGS__); break;      \
        case 6: FN<6, C>(__VA_ARGS__); break;      \
        case 7: FN<7, C>(__VA_ARGS__); break;      \
        case 8: FN<8, C>(__VA_ARGS__); break;      \
        case 9: FN<9, C>(__VA_ARGS__); break;      \
        case 10: FN<10, C>(__VA_ARGS__); break;    \
        case 11: FN<11, C>(__VA_ARGS__); break;    \
        case 12: FN<12, C>(__VA_ARGS__); break;    \
        case 13: FN<13, C>(__VA_ARGS__); break;    \
        case 14: FN<14, C>(__VA_ARGS__); break;    \
        case 15: FN<15, C>(__VA_ARGS__); break;    \
        default: FN<16, C>(__VA_ARGS__); break;    \
    }

template <class C>
void launch_interp(const NufftDev &u, const C *U, OutCols f, hipStream_t s) {
    PFSHT_W_SWITCH(u.w, interp_w, u, U, f, s);
}

template <class C>
void launch_spread(const NufftDev &u, InCols f, C *U, hipStream_t s) {
    C *Wst = static_cast<C *>(u.Wst);
    PFSHT_W_SWITCH(u.w, spread_w, u, f, Wst, s);
    dim3 grid((u.N2 + 255) / 256, u.nR, u.B);
    spread_merge_kernel<C><<<grid, 256, 0, s>>>(Wst, u.st_slot, U, u.N1, u.N2, 4 * u.S, u.ns1, u.ns2, u.win, u.w,
                                                u.nst, u.r0, u.nR);
}

// Plan groups: one launch per stage over the super-tiles (rows, for the merge) of many
// single-transform plans of one kernel width and grid type.
template <int W, class C>
static void interp_group_w(const NufftTab *tabs, const int2 *work, int nwork, int win, hipStream_t s) {
    if constexpr (kFp32Math<W, C>)
        interp_st32_kernel<W, true><<<dim3(nwork), 256, sizeof(C) * size_t(win) * win, s>>>(
            nullptr, nullptr, nullptr, nullptr, OutCols(), nullptr, nullptr, 0, 0, 0, 0, 0, 0, 0, tabs, work);
    else
        interp_st_kernel<W, C, true><<<dim3(nwork), 256, sizeof(C) * size_t(win) * win, s>>>(
            nullptr, nullptr, nullptr, nullptr, OutCols(), nullptr, nullptr, 0, 0, 0, 0, 0, 0, 0, tabs, work);
}

template <int W, class C>
static void spread_group_w(const NufftTab *tabs, const int2 *work, int nwork, int win, hipStream_t s) {
    if constexpr (kFp32Math<W, C>)
        spread_st32_kernel<W, 1, true><<<dim3(nwork), 64 * kSpreadWaves, spread32_lds_bytes(win, 1), s>>>(
            nullptr, nullptr, nullptr, InCols(), nullptr, nullptr, nullptr, 0, 0, 0, 0, 1, tabs, work);
    else
        spread_st_kernel<W, 1, C, true><<<dim3(nwork), 64 * kSpreadWaves, spread_lds_bytes(win, 1), s>>>(
            nullptr, nullptr, nullptr, InCols(), nullptr, nullptr, nullptr, 0, 0, 0, 0, 1, tabs, work);
}

template <class C>
void launch_interp_group(const NufftTab *tabs, const int2 *work, int nwork, int w, int win, hipStream_t s) {
    if (nwork <= 0) return;
    PFSHT_W_SWITCH(w, interp_group_w, tabs, work, nwork, win, s);
}

template <class C>
void launch_spread_group(const NufftTab *tabs, const int2 *work, int nwork, const int4 *mwork, int nmwork, int w,
                         int win, hipStream_t s) {
    if (nwork > 0) {
        PFSHT_W_SWITCH(w, spread_group_w, tabs, work, nwork, win, s);
    }
    if (nmwork > 0)
        spread_merge_kernel<C, true><<<dim3(nmwork), 256, 0, s>>>(nullptr, nullptr, nullptr, 0, 0, 0, 0, 0, 0, 0, 0,
                                                                  0, 0, tabs, mwork);
}

#define PFSHT_INSTANTIATE(C)                                                                                   \
    template void launch_deconv_pad_ut<C>(const NufftDev &, const double2 *, int, int, double, const C *, C *,  \
                                          hipStream_t);                                                         \
    template void launch_deconv_trunc_ut<C>(const NufftDev &, const double2 *, int, int, double, const C *, C *,\
                                            hipStream_t);                                                       \
    template void launch_ut_to_rows<C>(const NufftDev &, int, const C *, C *, hipStream_t);                     \
    template void launch_rows_to_ut<C>(const NufftDev &, int, const C *, C *, hipStream_t);                     \
    template void launch_interp<C>(const NufftDev &, const C *, OutCols, hipStream_t);                          \
    template void launch_spread<C>(const NufftDev &, InCols, C *, hipStream_t);                                 \
    template void launch_interp_group<C>(const NufftTab *, const int2 *, int, int, int, hipStream_t);           \
    template void launch_spread_group<C>(const NufftTab *, const int2 *, int, const int4 *, int, int, int,      \
                                         hipStream_t);
PFSHT_INSTANTIATE(double2)
PFSHT_INSTANTIATE(float2)
#undef PFSHT_INSTANTIATE

}  // namespace pfsht
