/*
 * pfsht_oracle.h — C API of the CPU oracle for PartialFSHT.jl's stretched-leaf apply path.
 *
 * TEST INFRASTRUCTURE ONLY. This restates the reference algorithm on the CPU so that tests,
 * the smoke check and the benchmark's CPU-baseline leg can compare against it. Nothing in the
 * product library links, loads or calls it.
 *
 * Layouts are the reference's (column-major, Julia order):
 *   alm : (L+1) x (2L+1) x B complex, sph_mode(l,m) = (l-|m|, 2|m| + (m>=0) - 1)  (Layout.jl:13,20)
 *   G   : n x (2L+1) x B complex, column m + L                                    (Legendre.jl:306)
 *   F   : nth x nph x B complex                                                   (Stretch.jl:215-219)
 *   f   : M x B complex                                                           (Layout.jl:47-56)
 */
#ifndef PFSHT_ORACLE_H
#define PFSHT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } or_cplx;

/* Geometry kinds (Windows.jl:34-36). */
enum { OR_KIND_POLE = 0, OR_KIND_INTERIOR = 1, OR_KIND_FULL = 2 };

typedef struct {
    int kind, lmax, K;
    double eps, beta, h, thlo, thhi, epsapo, period, Theta1, sigma, ThetaT, invZ;
} or_geom;

typedef struct {
    int kind, flip, lmax, ntrans, method; /* method: 1 direct, 2 stretch */
    int nth_half, nth, nph, mmax;
    int64_t ring_work, npts;
    int nufft_kind; /* 0 = ES kernel, 1 = direct sum */
    int es_w, es_n1, es_n2;
    double es_beta;
} or_plan_info;

const char *or_last_error(void);

/* Layout.jl */
int64_t or_sph_mode(int L, int l, int m);

/* Legendre.jl */
double or_polar_pad_factor(void);
int or_polar_mmax(double theta, int L, double eps);
void or_sectorial_log2c(int L, double *out);
int or_nested_ranges(const int32_t *mlim, int n, int L, int32_t *jlo, int32_t *jhi);
int64_t or_retained_triples(const int32_t *mlim, int n, int L);
double or_cost_stretch(double thlo, double thhi, int64_t npts, int L, int B, double tol);
double or_cost_direct(int64_t npts, int L, int B);
double or_alf(int l, int m, double theta);

typedef struct or_rings or_rings;
or_rings *or_rings_create(const double *theta, int n, int L, double eps, int ntrans, int reflect);
void or_rings_destroy(or_rings *e);
int64_t or_rings_work(const or_rings *e);
int or_rings_mmax(const or_rings *e);
void or_rings_tables(const or_rings *e, double *costh, double *log2sin, int32_t *mlim, int32_t *jlo,
                     int32_t *jhi);
int or_rings_from_alm(const or_rings *e, const or_cplx *alm, or_cplx *G, int nthreads);
int or_alm_from_rings(const or_rings *e, const or_cplx *G, or_cplx *alm, int nthreads);

/* Windows.jl / Stretch.jl sizing */
int or_stretch_geometry(double thlo, double thhi, int L, double eps, double epsapo_max, or_geom *out);
double or_es_cdf(double beta, double u);
double or_stretch_window(const or_geom *g, double Theta);
double or_theta_to_stretch(const or_geom *g, double theta);
double or_stretch_to_theta(const or_geom *g, double Theta);
int64_t or_next_smooth(int64_t n);
int64_t or_next_smooth_even(int64_t n);

/* DFS.jl */
void or_dfs_double(const or_cplx *F, or_cplx *Ft, int nth, int nph, int B, int s);
void or_dfs_fold(const or_cplx *Ft, or_cplx *F, int nth, int nph, int B, int s);

/* FFT (mixed radix, unnormalised): sign -1 forward, +1 backward. 2-D on column-major n1 x n2 x B. */
void or_fft1(or_cplx *x, int n, int sign);
void or_fft2(or_cplx *x, int n1, int n2, int B, int sign, int nthreads);

/* NUFFT.jl: direct summation (type 1 or 2, modeord 1), and an ES-kernel NUFFT. */
int or_nufft_direct(int type, int n1, int n2, int iflag, int B, const double *x, const double *y,
                    int64_t M, const or_cplx *in, or_cplx *out);
int or_nufft_es(int type, int n1, int n2, int iflag, int B, const double *x, const double *y,
                int64_t M, double tol, const or_cplx *in, or_cplx *out, int nthreads);
void or_es_params(double tol, int *w, double *beta);

/* Plan / leaves (Plan.jl, Stretch.jl, Direct.jl, Cost.jl). method: 0 auto, 1 direct, 2 stretch. */
typedef struct or_plan or_plan;
or_plan *or_plan_create(const double *theta, const double *phi, int64_t M, int L, double tol,
                        int ntrans, int method, int nufft_kind, int nthreads);
void or_plan_destroy(or_plan *p);
void or_plan_get_info(const or_plan *p, or_plan_info *info);
int or_synth(or_plan *p, const or_cplx *alm, or_cplx *f);
int or_adjoint(or_plan *p, const or_cplx *f, or_cplx *alm);
/* stage outputs of the last synthesis: G (nth_half x (2L+1) x B) and Fhat (nth x nph x B) */
int or_plan_last_G(const or_plan *p, or_cplx *G);
int or_plan_last_Fhat(const or_plan *p, or_cplx *Fhat);
/* stage timings of the last apply, seconds: [ring, grid(fft+dfs+window), nufft] */
void or_plan_last_times(const or_plan *p, double *t3);
/* the stretched coordinates (Theta_i, phi_i) of the plan's points (src/Stretch.jl:202-213) */
int or_plan_points(const or_plan *p, double *Theta, double *phi);

#ifdef __cplusplus
}
#endif

/* Spin-weighted transforms (Spin.jl) */
double or_wigner_d(int l, int m, int n, double beta);
void or_sylm(int s, int l, int m, double theta, double phi, double *re, double *im);
typedef struct or_spin_rings or_spin_rings;
or_spin_rings *or_spin_rings_create(const double *theta, int n, int L, int spin, double eps, int ntrans,
                                    int reflect);
void or_spin_rings_destroy(or_spin_rings *e);
int or_spin_rings_mmax(const or_spin_rings *e);
int64_t or_spin_rings_work(const or_spin_rings *e);
int or_spin_rings_from_alm(const or_spin_rings *e, const or_cplx *alm, or_cplx *G, int nthreads);
int or_spin_alm_from_rings(const or_spin_rings *e, const or_cplx *G, or_cplx *alm, int nthreads);
or_plan *or_plan_create_spin(const double *theta, const double *phi, int64_t M, int L, double tol,
                             int ntrans, int spin, int method, int nufft_kind, int nthreads);
#endif
