// adj_rows.hpp — the cross-ring reduction of the adjoint ring passes (scalar and spin): a
// wavefront-private 16-degree x 64-ring LDS tile W (phase 1 writes one degree per row, one ring per
// column), then lane t of 16-lane row r sums W[t][16 r + k] G[16 r + k] over k with the G operand
// broadcast from lane k of its own row by DPP on the FMA, and the four row sums are added across
// rows with lane swaps. Sums have a fixed order: bit-identical run to run.
#pragma once

#include "pfsht_common.hpp"

namespace pfsht {

constexpr int kSub = 16;  // degrees per LDS sub-tile
constexpr int kWS = 69;   // LDS row stride in doubles; ring rho sits at column rho + rho / 16

// Wavefront-private ordering of LDS traffic: a wavefront's LDS operations complete in issue
// order, so a wavefront reading back its own tile needs only a compiler barrier.
__device__ __forceinline__ void wave_lds_order() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// a_c += g_c(lane 16 row + K) * w for the (up to) four components of one transform. The DPP source
// must not have been written by a VALU instruction in the two preceding cycles: s_nop 1.
template <int K, bool NEG>
__device__ __forceinline__ void fmac_row_bcast(double (&a)[4], const double (&g)[4], double w) {
    if (NEG) {
        asm("s_nop 1\n\t"
            "v_fmac_f64_dpp %0, %4, %8 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %1, %5, %8 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %2, %6, %8 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %3, %7, %8 row_newbcast:%9 row_mask:0xf bank_mask:0xf"
            : "+v"(a[0]), "+v"(a[1]), "+v"(a[2]), "+v"(a[3])
            : "v"(g[0]), "v"(g[1]), "v"(g[2]), "v"(g[3]), "v"(w), "n"(K));
    } else {
        asm("s_nop 1\n\t"
            "v_fmac_f64_dpp %0, %2, %4 row_newbcast:%5 row_mask:0xf bank_mask:0xf\n\t"
            "v_fmac_f64_dpp %1, %3, %4 row_newbcast:%5 row_mask:0xf bank_mask:0xf"
            : "+v"(a[0]), "+v"(a[1])
            : "v"(g[0]), "v"(g[1]), "v"(w), "n"(K));
    }
}

template <int K, int NB, bool NEG>
struct RowSweep {
    static __device__ __forceinline__ void run(const double *__restrict__ wr, double (&a)[NB][4],
                                               const double (&g)[NB][4]) {
        const double w = wr[K];
#pragma unroll
        for (int u = 0; u < NB; ++u) fmac_row_bcast<K, NEG>(a[u], g[u], w);
        RowSweep<K + 1, NB, NEG>::run(wr, a, g);
    }
};
template <int NB, bool NEG>
struct RowSweep<kSub, NB, NEG> {
    static __device__ __forceinline__ void run(const double *__restrict__, double (&)[NB][4], const double (&)[NB][4]) {}
};

// Sum of v over the four 16-lane rows, the same value (bitwise) in every lane: lane swaps across
// the wave halves and then across row pairs; each stage adds the lower partner first.
__device__ __forceinline__ double row_sum4(double v) {
    unsigned lo = __double2loint(v), hi = __double2hiint(v);
    auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    v = __hiloint2double(int(b[0]), int(a[0])) + __hiloint2double(int(b[1]), int(a[1]));
    lo = __double2loint(v);
    hi = __double2hiint(v);
    a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    return __hiloint2double(int(b[0]), int(a[0])) + __hiloint2double(int(b[1]), int(a[1]));
}

}  // namespace pfsht
