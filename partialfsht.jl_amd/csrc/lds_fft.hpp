// lds_fft.hpp — complex FFTs of one workgroup over an array held in LDS: a Stockham autosort
// transform in radix-8/4/2/3/5 passes (lengths 2^a 3^b 5^c, the lengths the stretched grids take).
//
// A pass of radix R over n = R * nbf values, with p the product of the earlier passes' radices,
// takes butterfly i < nbf from x[i + j nbf] (j < R), multiplies input j by w^(j k), k = i mod p,
// w = exp(-2 pi i / (p R)), does an R-point DFT and writes output j to x[(i - k) R + k + j p]; after
// the last pass x holds the DFT in natural order. Each thread holds its butterflies' values in
// registers between the read and the write, so one LDS array suffices (a barrier separates them).
// The twiddles come from a table of exp(-2 pi i k / n), k < n (conjugated for the inverse): one
// table value w^k per butterfly, its powers w^(jk) by multiplication.
#pragma once
#include <hip/hip_runtime.h>

namespace pfsht {

constexpr int kFftElems = 8;    // complex values a thread holds in a pass: n <= kFftElems * threads
constexpr int kFftMaxPass = 16;

// One transform length and its passes (kernel argument).
struct FftLen {
    int n = 1;
    int npass = 0;
    int tw = 0;  // offset of the length's twiddle table in the plan's table
    int rad[kFftMaxPass] = {};
};

template <class C> struct CxReal;
template <> struct CxReal<double2> { using T = double; };
template <> struct CxReal<float2> { using T = float; };

template <class C> __device__ __forceinline__ C cx(typename CxReal<C>::T re, typename CxReal<C>::T im) {
    C r;
    r.x = re;
    r.y = im;
    return r;
}
template <class C> __device__ __forceinline__ C cadd(C a, C b) { return cx<C>(a.x + b.x, a.y + b.y); }
template <class C> __device__ __forceinline__ C csub(C a, C b) { return cx<C>(a.x - b.x, a.y - b.y); }
template <class C> __device__ __forceinline__ C cscale(C a, typename CxReal<C>::T s) { return cx<C>(a.x * s, a.y * s); }
// a w (INV: a conj(w))
template <bool INV, class C> __device__ __forceinline__ C cmulw(C a, C w) {
    if (INV) return cx<C>(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y);
    return cx<C>(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}
// a (-i) for the forward transform, a (+i) for the inverse
template <bool INV, class C> __device__ __forceinline__ C crot(C a) {
    if (INV) return cx<C>(-a.y, a.x);
    return cx<C>(a.y, -a.x);
}

// R-point DFTs in place, y_s = sum_j v_j exp(-+2 pi i j s / R).
template <bool INV, class C> __device__ __forceinline__ void dft2(C *v) {
    const C a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
}

template <bool INV, class C> __device__ __forceinline__ void dft4(C &v0, C &v1, C &v2, C &v3) {
    const C a = cadd(v0, v2), b = csub(v0, v2), c = cadd(v1, v3), d = crot<INV>(csub(v1, v3));
    v0 = cadd(a, c);
    v2 = csub(a, c);
    v1 = cadd(b, d);
    v3 = csub(b, d);
}

template <bool INV, class C> __device__ __forceinline__ void dft8(C *v) {
    using T = typename CxReal<C>::T;
    const T r = T(0.70710678118654752440);  // sqrt(1/2)
    dft4<INV>(v[0], v[2], v[4], v[6]);
    dft4<INV>(v[1], v[3], v[5], v[7]);
    // odd half times exp(-+2 pi i s / 8), s = 1, 2, 3
    C o1 = v[3], o2 = v[5], o3 = v[7];
    if (INV) {
        o1 = cx<C>(r * (o1.x - o1.y), r * (o1.x + o1.y));
        o3 = cx<C>(-r * (o3.x + o3.y), r * (o3.x - o3.y));
    } else {
        o1 = cx<C>(r * (o1.x + o1.y), r * (o1.y - o1.x));
        o3 = cx<C>(r * (o3.y - o3.x), -r * (o3.x + o3.y));
    }
    o2 = crot<INV>(o2);
    const C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1];
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
}

template <bool INV, class C> __device__ __forceinline__ void dft3(C *v) {
    using T = typename CxReal<C>::T;
    const T s3 = T(0.86602540378443864676);  // sin(2 pi / 3)
    const C t = cadd(v[1], v[2]), d = crot<INV>(csub(v[1], v[2]));
    const C a = csub(v[0], cscale(t, T(0.5)));
    v[0] = cadd(v[0], t);
    v[1] = cadd(a, cscale(d, s3));
    v[2] = csub(a, cscale(d, s3));
}

template <bool INV, class C> __device__ __forceinline__ void dft5(C *v) {
    using T = typename CxReal<C>::T;
    const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);  // cos(2 pi/5), cos(4 pi/5)
    const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);   // sin(2 pi/5), sin(4 pi/5)
    const C t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]), d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
    const C a1 = cadd(v[0], cadd(cscale(t1, c1), cscale(t2, c2)));
    const C a2 = cadd(v[0], cadd(cscale(t1, c2), cscale(t2, c1)));
    const C b1 = crot<INV>(cadd(cscale(d1, s1), cscale(d2, s2)));
    const C b2 = crot<INV>(csub(cscale(d1, s2), cscale(d2, s1)));
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
}

template <int R, bool INV, class C> __device__ __forceinline__ void dft(C *v) {
    if constexpr (R == 2) dft2<INV>(v);
    else if constexpr (R == 3) dft3<INV>(v);
    else if constexpr (R == 4) dft4<INV>(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 5) dft5<INV>(v);
    else dft8<INV>(v);
}

// Input j of a butterfly times w1^j, j = 1 .. R-1 (w1 already conjugated for the inverse): the
// powers by a product tree at most three multiplications deep (w2 = w1^2, w4 = w2^2, w3 = w1 w2, ...),
// so each carries a few roundings of the table value.
template <int R, class C> __device__ __forceinline__ void twiddle(C *v, C w1) {
    v[1] = cmulw<false>(v[1], w1);
    if constexpr (R >= 3) {
        const C w2 = cmulw<false>(w1, w1);
        v[2] = cmulw<false>(v[2], w2);
        if constexpr (R >= 4) {
            const C w3 = cmulw<false>(w1, w2);
            v[3] = cmulw<false>(v[3], w3);
            if constexpr (R >= 5) {
                const C w4 = cmulw<false>(w2, w2);
                v[4] = cmulw<false>(v[4], w4);
                if constexpr (R == 8) {
                    v[5] = cmulw<false>(v[5], cmulw<false>(w1, w4));
                    v[6] = cmulw<false>(v[6], cmulw<false>(w2, w4));
                    v[7] = cmulw<false>(v[7], cmulw<false>(w3, w4));
                }
            }
        }
    }
}

// One pass (the barrier before it is the caller's; it ends with one). tstep = n / (p R).
// i / p is taken in float: (i + 1/2) / p lies at least 1/(2p) from an integer and the float
// product's error is below n 2^-23 / p, so the truncation is exact for n < 2^22. A butterfly's one
// table twiddle, w^k (index k tstep), is loaded with its inputs, before the barrier.
// LDS slot of element e. SW: the low three bits of e XOR bits 3..5, so that the stride-2/4/8 stores of
// a first pass (p = 1) reach eight distinct 16-byte bank groups; the array then holds lds_slots(n)
// elements (e ^ s stays in e's aligned block of eight).
template <bool SW> __device__ __forceinline__ int lds_slot(int e) { return SW ? (e ^ ((e >> 3) & 7)) : e; }
__host__ __device__ inline int lds_slots(int n) { return (n + 7) & ~7; }

template <int R, bool INV, int E, bool SW, class C>
__device__ __forceinline__ void fft_pass(C *x, int n, int p, int tstep, const C *__restrict__ tw, int NT) {
    constexpr int K = (E + R - 1) / R;
    const int nbf = n / R;
    const float rp = 1.0f / float(p);
    C v[K][R];
    C w1[K];
    int kk[K];  // butterfly i's position k = i mod p within its span, for the loads and the stores
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const int i = int(threadIdx.x) + q * NT;
        kk[q] = i - int((float(i) + 0.5f) * rp) * p;
        if (i < nbf) {
#pragma unroll
            for (int j = 0; j < R; ++j) v[q][j] = x[lds_slot<SW>(i + j * nbf)];
            if (p > 1) w1[q] = tw[kk[q] * tstep];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const int i = int(threadIdx.x) + q * NT;
        if (i < nbf) {
            const int k = kk[q];
            if (p > 1) twiddle<R>(v[q], INV ? cx<C>(w1[q].x, -w1[q].y) : w1[q]);
            dft<R, INV>(v[q]);
            const int base = (i - k) * R + k;
#pragma unroll
            for (int j = 0; j < R; ++j) x[lds_slot<SW>(base + j * p)] = v[q][j];
        }
    }
    __syncthreads();
}

// The DFT of x[0 .. f.n) in place (forward: exp(-2 pi i jk/n); INV: exp(+2 pi i jk/n), unscaled).
// Every thread of the NT-thread workgroup (NT = blockDim.x, n <= E * NT; E: values a thread holds in
// a pass) calls it after a barrier that publishes x; it returns after one. tw is the length's own
// table (the plan's table + f.tw); SW: swizzled LDS slots (lds_slot).
template <bool INV, int E = kFftElems, bool SW = false, class C>
__device__ __forceinline__ void lds_fft(C *x, const FftLen &f, const C *__restrict__ tw, int NT) {
    // q = n / (p r), the table step of a pass of radix r at span p, divided down by the pass's
    // compile-time radix (no run-time integer division)
    int p = 1;
    unsigned q = unsigned(f.n);
    for (int s = 0; s < f.npass; ++s) {
        const int r = f.rad[s];
        switch (r) {
        case 8: q /= 8u; fft_pass<8, INV, E, SW>(x, f.n, p, int(q), tw, NT); break;
        case 4: q /= 4u; fft_pass<4, INV, E, SW>(x, f.n, p, int(q), tw, NT); break;
        case 2: q /= 2u; fft_pass<2, INV, E, SW>(x, f.n, p, int(q), tw, NT); break;
        case 3: q /= 3u; fft_pass<3, INV, E, SW>(x, f.n, p, int(q), tw, NT); break;
        default: q /= 5u; fft_pass<5, INV, E, SW>(x, f.n, p, int(q), tw, NT); break;
        }
        p *= r;
    }
}

// The DFT of x[0 .. n), n even and n / 2 <= E * NT, through one radix-2 decimation-in-frequency step,
// x[i], x[i + n/2] -> x[i] + x[i + n/2], (x[i] - x[i + n/2]) w^i (in place, no thread reads another's
// values), and LDS transforms of the two halves: afterwards X[2m] is at x[m] and X[2m + 1] at
// x[n/2 + m]. For lengths past E x 1024 values: each half keeps the per-thread share of E values.
// Unswizzled slots. tw_n is the length-n table, half the length n/2 passes with their own table.
template <bool INV, int E, class C>
__device__ __forceinline__ void lds_fft_split(C *x, int n, const FftLen &half, const C *__restrict__ tw_n,
                              const C *__restrict__ tw_half, int NT) {
    const int h = n / 2;
    for (int i = threadIdx.x; i < h; i += NT) {
        const C a = x[i], b = x[i + h];
        x[i] = cadd(a, b);
        x[i + h] = cmulw<INV>(csub(a, b), tw_n[i]);
    }
    __syncthreads();
    lds_fft<INV, E, false>(x, half, tw_half, NT);
    lds_fft<INV, E, false>(x + h, half, tw_half, NT);
}

// Host: the passes of length n (8s first, a 2 only when the power of two is 2; then 3s and 5s);
// false when n has another prime factor or needs more than kFftMaxPass passes.
inline bool fft_len_plan(int n, FftLen &f) {
    f = FftLen();
    f.n = n;
    if (n < 1) return false;
    int m = n, e2 = 0;
    while (m % 2 == 0) { m /= 2; ++e2; }
    auto add = [&](int r) {
        if (f.npass >= kFftMaxPass) return false;
        f.rad[f.npass++] = r;
        return true;
    };
    int e8 = e2 / 3, rest = e2 % 3;
    if (rest == 1 && e8 > 0) { --e8; rest = 4; }  // 2^4 as 4 x 4 rather than 8 x 2
    for (int i = 0; i < e8; ++i)
        if (!add(8)) return false;
    if (rest == 4) { if (!add(4) || !add(4)) return false; }
    else if (rest == 2) { if (!add(4)) return false; }
    else if (rest == 1) { if (!add(2)) return false; }
    while (m % 3 == 0) { m /= 3; if (!add(3)) return false; }
    while (m % 5 == 0) { m /= 5; if (!add(5)) return false; }
    return m == 1;
}

}  // namespace pfsht
