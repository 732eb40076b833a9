// direct_kernels.hip — the Direct leaf (src/Direct.jl:15-117) on the device: every point is its
// own ring of the block-scaled ring engine (polar truncation at eps = 0), followed by the exact
// per-order longitude phase. Used for point sets the stretched construction refuses or where it
// costs more (src/Cost.jl:174-201), and as a device oracle at scale.
#include "pfsht_common.hpp"

namespace pfsht {

// f[perm[j]] = sum_{|m| <= mlim_j} G[j, m] cis(m phi_j), the phase formed exactly per order rather
// than by a running product (src/Direct.jl:69-88).
__global__ void direct_synth_kernel(const double2 *__restrict__ Gd, const double *__restrict__ phis,
                                    const int32_t *__restrict__ mlim, const int32_t *__restrict__ perm,
                                    OutCols f, int64_t M, int mmax) {
    const int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    const int b = blockIdx.y;
    if (j >= M) return;
    const int64_t ncol = 2 * mmax + 1;
    const double2 *G = Gd + int64_t(b) * ncol * M + j;
    const double phi = phis[j];
    const int mt = min(mlim[j], mmax);
    const double2 g0 = G[int64_t(mmax) * M];
    double ar = g0.x, ai = g0.y;
    for (int m = 1; m <= mt; ++m) {
        double sn, cs;
        sincos(double(m) * phi, &sn, &cs);
        const double2 gp = G[int64_t(mmax + m) * M];
        const double2 gm = G[int64_t(mmax - m) * M];
        // gp * (cs + i sn) + gm * (cs - i sn)
        ar += (gp.x * cs - gp.y * sn) + (gm.x * cs + gm.y * sn);
        ai += (gp.x * sn + gp.y * cs) + (gm.y * cs - gm.x * sn);
    }
    gstore(f.col(b) + perm[j], make_double2(ar, ai));
}

// G[j, m] = f_j conj(cis(m phi_j)), G[j, -m] = f_j cis(m phi_j) on the orders ring j retains
// (src/Direct.jl:94-117); alm_from_rings! then reads exactly those entries.
__global__ void direct_adjoint_kernel(InCols f, const double *__restrict__ phis,
                                      const int32_t *__restrict__ mlim, const int32_t *__restrict__ perm,
                                      double2 *__restrict__ Gd, int64_t M, int mmax) {
    const int64_t j = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    const int b = blockIdx.y;
    if (j >= M) return;
    const int64_t ncol = 2 * mmax + 1;
    double2 *G = Gd + int64_t(b) * ncol * M + j;
    const double2 fj = gload(f.col(b) + perm[j]);
    const double phi = phis[j];
    const int mt = min(mlim[j], mmax);
    G[int64_t(mmax) * M] = fj;
    for (int m = 1; m <= mt; ++m) {
        double sn, cs;
        sincos(double(m) * phi, &sn, &cs);
        G[int64_t(mmax + m) * M] = make_double2(fj.x * cs + fj.y * sn, fj.y * cs - fj.x * sn);
        G[int64_t(mmax - m) * M] = make_double2(fj.x * cs - fj.y * sn, fj.y * cs + fj.x * sn);
    }
}

void launch_direct_synth(const double2 *Gd, const double *phis, const int32_t *mlim, const int32_t *perm,
                         OutCols f, int64_t M, int mmax, int nb, hipStream_t s) {
    dim3 grid(unsigned((M + 255) / 256), nb);
    direct_synth_kernel<<<grid, 256, 0, s>>>(Gd, phis, mlim, perm, f, M, mmax);
}

void launch_direct_adjoint(InCols f, const double *phis, const int32_t *mlim, const int32_t *perm,
                           double2 *Gd, int64_t M, int mmax, int nb, hipStream_t s) {
    dim3 grid(unsigned((M + 255) / 256), nb);
    direct_adjoint_kernel<<<grid, 256, 0, s>>>(f, phis, mlim, perm, Gd, M, mmax);
}

}  // namespace pfsht
