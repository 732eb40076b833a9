/*
 * pfsht_oracle.c — CPU restatement of PartialFSHT.jl's stretched-leaf apply path.
 *
 * TEST INFRASTRUCTURE ONLY: the CPU results that tests/, the smoke check and the benchmark's
 * CPU-baseline leg compare the HIP library against. It is never linked into, loaded by or called
 * from the product.
 *
 * Every function follows a reference file:line (PartialFSHT.jl v0.1.0). Arithmetic is written in
 * the reference's evaluation order and built with -ffp-contract=off, because Julia does not fuse
 * a*b+c into an FMA unless asked.
 *
 * Deliberate deviations from the reference, each tolerance-neutral:
 *  - the order-0 seed exponent is log2c[0] even on a ring exactly at a pole, as in `alf`
 *    (Legendre.jl:565) and the device kernel (ext/PartialFSHTKernelAbstractionsExt.jl:96,179); the host ring engine
 *    (Legendre.jl:381,478) would form 0*(-Inf) = NaN there;
 *  - @simd reductions in the adjoint (Legendre.jl:437-443,506-509) are summed in ring order;
 *  - the NUFFT is either the reference's own direct sum (NUFFT.jl:85-129) or an
 *    exponential-of-semicircle NUFFT written from the published method (Barnett, Magland &
 *    af Klinteberg, SISC 41, 2019), standing in for FINUFFT, which is not vendored;
 *  - the planner compares Direct against Stretch only: the multi-box partition
 *    (Cost.jl:192-197) and NUFSHT are outside this path.
 *
 * Parity pins (tests/test_oracle_*.py): mpmath ALF goldens, closed-form Ylm, scipy sph_harm_y,
 * DFS transpose, NUFFT vs direct sum, Stretch vs Direct, adjoint identity, counted work; for spin,
 * Wigner-d closed forms and mpmath values of Wigner's explicit sum, the engine against single-value
 * wigner_d, exact poles, reflection, spin Direct against explicit sYlm sums, spin Stretch vs Direct.
 *
 * The spin ring engine gives every ring its own power-of-two exponent where the reference shares one
 * per chunk of rings (Spin.jl:272-310); the values agree to round-off.
 */
#include "pfsht_oracle.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.14159265358979323846

static _Thread_local char g_err[512];
static void set_err(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}
const char *or_last_error(void) { return g_err; }

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static inline or_cplx cx(double re, double im) { or_cplx z = {re, im}; return z; }
static inline or_cplx cadd(or_cplx a, or_cplx b) { return cx(a.re + b.re, a.im + b.im); }
static inline or_cplx cmul(or_cplx a, or_cplx b) {
    return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline or_cplx cscale(or_cplx a, double s) { return cx(a.re * s, a.im * s); }
static inline or_cplx cconj(or_cplx a) { return cx(a.re, -a.im); }
static inline or_cplx ccis(double t) { return cx(cos(t), sin(t)); }

/* ---------------------------------------------------------------------------------------------
 * Layout.jl:13,20 — FastSphericalHarmonics sph_mode, 0-based linear index, column-major.
 * ------------------------------------------------------------------------------------------- */
int64_t or_sph_mode(int L, int l, int m) {
    int am = m < 0 ? -m : m;
    int64_t row = l - am;
    int64_t col = 2 * am + (m >= 0 ? 1 : 0) - 1;
    return row + col * (int64_t)(L + 1);
}

/* ---------------------------------------------------------------------------------------------
 * Legendre.jl
 * ------------------------------------------------------------------------------------------- */
#define RESCALE_STRIDE 64       /* Legendre.jl:16 */
#define CHUNK_HEADROOM_BITS 800 /* Legendre.jl:20 */
#define CHUNK_MAX 256           /* Legendre.jl:23 */

double or_polar_pad_factor(void) { return 0.5 * cbrt(9.0); } /* Legendre.jl:31 */

/* Legendre.jl:48-57 */
int or_polar_mmax(double theta, int L, double eps) {
    if (L <= 0) return 0;
    double s = fabs(sin(theta));
    double z = ((double)L + 0.5) * s;
    double lne = fabs(log(fmax(eps, DBL_MIN)));
    double pad = or_polar_pad_factor() * cbrt(fmax(z, 1.0)) * cbrt(lne * lne) + 2;
    double v = ceil(z + pad);
    if (v < 0) return 0;
    if (v > (double)L) return L;
    return (int)v;
}

/* Legendre.jl:65-75 */
void or_sectorial_log2c(int L, double *v) {
    double acc = -log2(sqrt(4 * PI));
    v[0] = acc;
    for (int m = 1; m <= L; m++) {
        acc += log2(sqrt((double)(2 * m + 1) / (double)(2 * m)));
        v[m] = acc;
    }
}

/* Legendre.jl:89-127, symmetric = false; 0-based rings, empty when jhi < jlo. */
int or_nested_ranges(const int32_t *mlim, int n, int L, int32_t *jlo, int32_t *jhi) {
    if (n < 1) { set_err("need at least one ring"); return -1; }
    int peak = 0;
    for (int j = 1; j < n; j++)
        if (mlim[j] > mlim[peak]) peak = j;
    for (int j = 1; j <= peak; j++)
        if (mlim[j] < mlim[j - 1]) {
            set_err("retained-order limits are not unimodal in colatitude (fell at ring %d)", j + 1);
            return -1;
        }
    for (int j = peak + 1; j < n; j++)
        if (mlim[j] > mlim[j - 1]) {
            set_err("retained-order limits are not unimodal in colatitude (rose at ring %d)", j + 1);
            return -1;
        }
    for (int m = 0; m <= L; m++) { jlo[m] = 0; jhi[m] = -1; }
    int lo = peak, hi = peak;
    int mtop = mlim[peak] < L ? mlim[peak] : L;
    for (int m = mtop; m >= 0; m--) {
        while (lo > 0 && mlim[lo - 1] >= m) lo--;
        while (hi < n - 1 && mlim[hi + 1] >= m) hi++;
        jlo[m] = lo;
        jhi[m] = hi;
    }
    return 0;
}

/* Legendre.jl:136-145 */
int64_t or_retained_triples(const int32_t *mlim, int n, int L) {
    int32_t *jlo = malloc(sizeof(int32_t) * (L + 1)), *jhi = malloc(sizeof(int32_t) * (L + 1));
    int64_t w = -1;
    if (or_nested_ranges(mlim, n, L, jlo, jhi) == 0) {
        w = 0;
        for (int m = 0; m <= L; m++) {
            int64_t nj = jhi[m] - jlo[m] + 1;
            if (nj > 0) w += nj * (int64_t)(L - m + 1);
        }
    }
    free(jlo); free(jhi);
    return w;
}

struct or_rings {
    int L, B, n, mmax, nchunks;
    double eps;
    double *theta, *costh, *sinth, *log2sin, *log2c;
    int32_t *mlim, *jlo, *jhi, *cstart, *cstop;
    int cw;
};

/* Legendre.jl:262-283 */
static int ring_chunks(const double *sinth, const int32_t *mlim, int n, int32_t *cs, int32_t *ce) {
    int nc = 0, i = 0;
    while (i < n) {
        int stop = i;
        int mmax = mlim[i];
        while (stop < n - 1) {
            int nxt = stop + 1;
            if (nxt - i + 1 > CHUNK_MAX) break;
            double slo = sinth[i], shi = sinth[nxt];
            double span = (slo > 0 && shi > 0) ? fabs(log2(shi / slo)) : INFINITY;
            int mm = mmax > mlim[nxt] ? mmax : mlim[nxt];
            if ((double)mm * span > CHUNK_HEADROOM_BITS) break;
            mmax = mm;
            stop = nxt;
        }
        cs[nc] = i;
        ce[nc] = stop;
        nc++;
        i = stop + 1;
    }
    return nc;
}

/* Legendre.jl:194-255 */
or_rings *or_rings_create(const double *theta, int n, int L, double eps, int ntrans, int reflect) {
    if (L < 0) { set_err("lmax must be nonnegative"); return NULL; }
    if (ntrans < 1) { set_err("ntrans must be >= 1"); return NULL; }
    if (n < 1) { set_err("need at least one ring"); return NULL; }
    for (int j = 1; j < n; j++)
        if (theta[j] < theta[j - 1]) { set_err("ring colatitudes must be sorted ascending"); return NULL; }
    or_rings *e = calloc(1, sizeof *e);
    e->L = L; e->B = ntrans; e->n = n; e->eps = eps;
    e->theta = malloc(sizeof(double) * n);
    e->costh = malloc(sizeof(double) * n);
    e->sinth = malloc(sizeof(double) * n);
    e->log2sin = malloc(sizeof(double) * n);
    e->log2c = malloc(sizeof(double) * (L + 1));
    e->mlim = malloc(sizeof(int32_t) * n);
    e->jlo = malloc(sizeof(int32_t) * (L + 1));
    e->jhi = malloc(sizeof(int32_t) * (L + 1));
    e->cstart = malloc(sizeof(int32_t) * n);
    e->cstop = malloc(sizeof(int32_t) * n);
    double csgn = reflect ? -1.0 : 1.0;
    for (int j = 0; j < n; j++) {
        double s = sin(theta[j]), c = cos(theta[j]);
        double as = fabs(s);
        e->theta[j] = theta[j];
        e->sinth[j] = as;
        e->costh[j] = csgn * c;
        e->log2sin[j] = as > 0 ? log2(as) : -INFINITY;
    }
    e->mmax = 0;
    for (int j = 0; j < n; j++) {
        e->mlim[j] = or_polar_mmax(theta[j], L, eps);
        if (e->mlim[j] > e->mmax) e->mmax = e->mlim[j];
    }
    if (or_nested_ranges(e->mlim, n, L, e->jlo, e->jhi) != 0) { or_rings_destroy(e); return NULL; }
    or_sectorial_log2c(L, e->log2c);
    e->nchunks = ring_chunks(e->sinth, e->mlim, n, e->cstart, e->cstop);
    e->cw = 0;
    for (int c = 0; c < e->nchunks; c++) {
        int len = e->cstop[c] - e->cstart[c] + 1;
        if (len > e->cw) e->cw = len;
    }
    if (e->cw > CHUNK_MAX) e->cw = CHUNK_MAX;
    return e;
}

void or_rings_destroy(or_rings *e) {
    if (!e) return;
    free(e->theta); free(e->costh); free(e->sinth); free(e->log2sin); free(e->log2c);
    free(e->mlim); free(e->jlo); free(e->jhi); free(e->cstart); free(e->cstop);
    free(e);
}

/* Legendre.jl:286-293 */
int64_t or_rings_work(const or_rings *e) {
    int64_t w = 0;
    for (int m = 0; m <= e->L; m++) {
        int64_t nj = e->jhi[m] - e->jlo[m] + 1;
        if (nj > 0) w += nj * (int64_t)(e->L - m + 1);
    }
    return w;
}
int or_rings_mmax(const or_rings *e) { return e->mmax; }
void or_rings_tables(const or_rings *e, double *costh, double *log2sin, int32_t *mlim, int32_t *jlo,
                     int32_t *jhi) {
    if (costh) memcpy(costh, e->costh, sizeof(double) * e->n);
    if (log2sin) memcpy(log2sin, e->log2sin, sizeof(double) * e->n);
    if (mlim) memcpy(mlim, e->mlim, sizeof(int32_t) * e->n);
    if (jlo) memcpy(jlo, e->jlo, sizeof(int32_t) * (e->L + 1));
    if (jhi) memcpy(jhi, e->jhi, sizeof(int32_t) * (e->L + 1));
}

typedef struct { double *p1, *p2; or_cplx *accp, *accm; } scratch_t; /* Legendre.jl:154-159 */

/* Legendre.jl:475-481 (order 0 seeded from log2c[0]; see file header). */
static double seed_exponent(const or_rings *e, int m, int r0, int r1) {
    double ev = -INFINITY;
    int nan = 0;
    for (int j = r0; j <= r1; j++) {
        double t = m == 0 ? e->log2c[0] : (double)m * e->log2sin[j] + e->log2c[m];
        if (isnan(t)) nan = 1;
        else if (t > ev) ev = t;
    }
    if (nan) return 0.0;
    return isfinite(ev) ? ceil(ev) : 0.0;
}

/* Legendre.jl:521-542 */
static double renorm(int fwd, double *p1, double *p2, or_cplx *accp, or_cplx *accm, int nr, int B,
                     int cw) {
    double mx = 0;
    for (int k = 0; k < nr; k++) {
        double a = fabs(p1[k]), b = fabs(p2[k]);
        if (a > mx) mx = a;
        if (b > mx) mx = b;
    }
    if (mx == 0 || !isfinite(mx)) return 0;
    double s = ceil(log2(mx));
    if (s == 0) return 0;
    double f = exp2(-s);
    for (int k = 0; k < nr; k++) { p1[k] *= f; p2[k] *= f; }
    if (fwd)
        for (int b = 0; b < B; b++)
            for (int k = 0; k < nr; k++) {
                accp[k + b * cw] = cscale(accp[k + b * cw], f);
                accm[k + b * cw] = cscale(accm[k + b * cw], f);
            }
    return s;
}

static inline or_cplx coeff(const or_cplx *alm, int L, int l, int m, int b) {
    return alm[or_sph_mode(L, l, m) + (int64_t)b * (L + 1) * (2 * L + 1)];
}
static inline void coeff_add(or_cplx *alm, int L, int l, int m, int b, or_cplx v) {
    int64_t i = or_sph_mode(L, l, m) + (int64_t)b * (L + 1) * (2 * L + 1);
    alm[i] = cadd(alm[i], v);
}

/* Legendre.jl:484-496 */
static void touch_fwd(const or_cplx *alm, int L, or_cplx *accp, or_cplx *accm, const double *P, int l,
                      int m, int B, double sgn, int nr, int cw) {
    for (int b = 0; b < B; b++) {
        or_cplx ap = coeff(alm, L, l, m, b);
        or_cplx am = m == 0 ? cx(0, 0) : coeff(alm, L, l, -m, b);
        if (ap.re == 0 && ap.im == 0 && am.re == 0 && am.im == 0) continue;
        or_cplx ams = cscale(am, sgn);
        for (int k = 0; k < nr; k++) {
            accp[k + b * cw] = cadd(accp[k + b * cw], cscale(ap, P[k]));
            accm[k + b * cw] = cadd(accm[k + b * cw], cscale(ams, P[k]));
        }
    }
}

/* Legendre.jl:501-514 */
static void touch_adj(or_cplx *alm, int L, const or_cplx *accp, const or_cplx *accm, const double *P,
                      int l, int m, int B, double sgn, int nr, double E, int cw) {
    double f = exp2(E);
    for (int b = 0; b < B; b++) {
        or_cplx sp = cx(0, 0), sm = cx(0, 0);
        for (int k = 0; k < nr; k++) {
            sp = cadd(sp, cscale(accp[k + b * cw], P[k]));
            sm = cadd(sm, cscale(accm[k + b * cw], P[k]));
        }
        coeff_add(alm, L, l, m, b, cscale(sp, f));
        if (m > 0) coeff_add(alm, L, l, -m, b, cscale(cscale(sm, sgn), f));
    }
}

/* Legendre.jl:367-472. G: n x (2L+1) x B column-major; dir 1 = forward, 0 = adjoint. */
static void ring_chunk(int fwd, or_cplx *G, const or_cplx *G_in, or_cplx *alm_out, const or_cplx *alm_in,
                       const or_rings *e, int m, int r0, int r1, scratch_t *sc) {
    int L = e->L, B = e->B, n = e->n, off = L, cw = e->cw;
    int nr = r1 - r0 + 1;
    double sgn = (m % 2 == 0) ? 1.0 : -1.0;
    double *p1 = sc->p1, *p2 = sc->p2;
    or_cplx *accp = sc->accp, *accm = sc->accm;
    int64_t gB = (int64_t)n * (2 * L + 1);

    double E = seed_exponent(e, m, r0, r1);
    for (int k = 0; k < nr; k++) {
        int j = r0 + k;
        double e0 = m == 0 ? e->log2c[0] : (double)m * e->log2sin[j] + e->log2c[m];
        p2[k] = sgn * exp2(e0 - E);
    }
    if (fwd) {
        for (int b = 0; b < B; b++)
            for (int k = 0; k < nr; k++) accp[k + b * cw] = accm[k + b * cw] = cx(0, 0);
    } else {
        for (int b = 0; b < B; b++)
            for (int k = 0; k < nr; k++) {
                int j = r0 + k;
                accp[k + b * cw] = G_in[j + (int64_t)(m + off) * n + b * gB];
                accm[k + b * cw] = m == 0 ? cx(0, 0) : G_in[j + (int64_t)(-m + off) * n + b * gB];
            }
    }
    if (fwd) touch_fwd(alm_in, L, accp, accm, p2, m, m, B, sgn, nr, cw);
    else touch_adj(alm_out, L, accp, accm, p2, m, m, B, sgn, nr, E, cw);

    if (m < L) {
        for (int k = 0; k < nr; k++) p1[k] = sqrt((double)(2 * m + 3)) * e->costh[r0 + k] * p2[k];
        if (fwd) touch_fwd(alm_in, L, accp, accm, p1, m + 1, m, B, sgn, nr, cw);
        else touch_adj(alm_out, L, accp, accm, p1, m + 1, m, B, sgn, nr, E, cw);
        const double *cs = e->costh + r0;
        int steps = 0;
        for (int l = m + 2; l <= L; l++) {
            double a = sqrt((double)(2 * l - 1) * (double)(2 * l + 1) / ((double)(l - m) * (double)(l + m)));
            double bb = sqrt((double)(2 * l + 1) / (double)(2 * l - 3) * (double)(l - m - 1) /
                             (double)(l - m) * (double)(l + m - 1) / (double)(l + m));
            if (fwd && B == 1) {
                or_cplx ap = coeff(alm_in, L, l, m, 0);
                or_cplx ams = cscale(m == 0 ? cx(0, 0) : coeff(alm_in, L, l, -m, 0), sgn);
                for (int k = 0; k < nr; k++) {
                    double p = a * cs[k] * p1[k] - bb * p2[k];
                    p2[k] = p1[k];
                    p1[k] = p;
                    accp[k] = cadd(accp[k], cscale(ap, p));
                    accm[k] = cadd(accm[k], cscale(ams, p));
                }
            } else if (!fwd && B == 1) {
                double f = exp2(E);
                or_cplx sp = cx(0, 0), sm = cx(0, 0);
                for (int k = 0; k < nr; k++) {
                    double p = a * cs[k] * p1[k] - bb * p2[k];
                    p2[k] = p1[k];
                    p1[k] = p;
                    sp = cadd(sp, cscale(accp[k], p));
                    sm = cadd(sm, cscale(accm[k], p));
                }
                coeff_add(alm_out, L, l, m, 0, cscale(sp, f));
                if (m > 0) coeff_add(alm_out, L, l, -m, 0, cscale(cscale(sm, sgn), f));
            } else {
                for (int k = 0; k < nr; k++) {
                    double p = a * cs[k] * p1[k] - bb * p2[k];
                    p2[k] = p1[k];
                    p1[k] = p;
                }
                if (fwd) touch_fwd(alm_in, L, accp, accm, p1, l, m, B, sgn, nr, cw);
                else touch_adj(alm_out, L, accp, accm, p1, l, m, B, sgn, nr, E, cw);
            }
            steps++;
            if (steps == RESCALE_STRIDE && l < L) {
                steps = 0;
                E += renorm(fwd, p1, p2, accp, accm, nr, B, cw);
            }
        }
    }
    if (fwd) {
        double f = exp2(E);
        for (int b = 0; b < B; b++)
            for (int k = 0; k < nr; k++) {
                int j = r0 + k;
                int64_t ip = j + (int64_t)(m + off) * n + b * gB;
                G[ip] = cadd(G[ip], cscale(accp[k + b * cw], f));
                if (m > 0) {
                    int64_t im = j + (int64_t)(-m + off) * n + b * gB;
                    G[im] = cadd(G[im], cscale(accm[k + b * cw], f));
                }
            }
    }
}

/* Legendre.jl:354-365 */
static void walk_orders(int fwd, or_cplx *G, const or_cplx *G_in, or_cplx *alm_out, const or_cplx *alm_in,
                        const or_rings *e, int m0, int mstep, scratch_t *sc) {
    for (int m = m0; m <= e->L; m += mstep) {
        int lo = e->jlo[m], hi = e->jhi[m];
        if (hi < lo) continue;
        for (int c = 0; c < e->nchunks; c++) {
            int r0 = e->cstart[c] > lo ? e->cstart[c] : lo;
            int r1 = e->cstop[c] < hi ? e->cstop[c] : hi;
            if (r1 < r0) continue;
            ring_chunk(fwd, G, G_in, alm_out, alm_in, e, m, r0, r1, sc);
        }
    }
}

static void scratch_alloc(scratch_t *s, int cw, int B) {
    s->p1 = malloc(sizeof(double) * cw);
    s->p2 = malloc(sizeof(double) * cw);
    s->accp = malloc(sizeof(or_cplx) * cw * B);
    s->accm = malloc(sizeof(or_cplx) * cw * B);
}
static void scratch_free(scratch_t *s) { free(s->p1); free(s->p2); free(s->accp); free(s->accm); }

/* Serial walk (Legendre.jl:344-347) or interleaved orders t:ns:L per thread
 * (ext/PartialFSHTOhMyThreadsExt.jl:31-42). */
static void walk_rings(int fwd, or_cplx *G, const or_cplx *G_in, or_cplx *alm_out, const or_cplx *alm_in,
                       const or_rings *e, int nthreads) {
    int ns = nthreads < 1 ? 1 : nthreads;
    if (ns > e->L + 1) ns = e->L + 1;
    int cw = e->cw < 1 ? 1 : e->cw;
    if (ns <= 1) {
        scratch_t sc;
        scratch_alloc(&sc, cw, e->B);
        walk_orders(fwd, G, G_in, alm_out, alm_in, e, 0, 1, &sc);
        scratch_free(&sc);
        return;
    }
#pragma omp parallel for num_threads(ns) schedule(static, 1)
    for (int t = 0; t < ns; t++) {
        scratch_t sc;
        scratch_alloc(&sc, cw, e->B);
        walk_orders(fwd, G, G_in, alm_out, alm_in, e, t, ns, &sc);
        scratch_free(&sc);
    }
}

/* Legendre.jl:310-323 */
int or_rings_from_alm(const or_rings *e, const or_cplx *alm, or_cplx *G, int nthreads) {
    int64_t nG = (int64_t)e->n * (2 * e->L + 1) * e->B;
    memset(G, 0, sizeof(or_cplx) * nG);
    walk_rings(1, G, NULL, NULL, alm, e, nthreads);
    return 0;
}

/* Legendre.jl:332-341 */
int or_alm_from_rings(const or_rings *e, const or_cplx *G, or_cplx *alm, int nthreads) {
    int64_t na = (int64_t)(e->L + 1) * (2 * e->L + 1) * e->B;
    memset(alm, 0, sizeof(or_cplx) * na);
    walk_rings(0, NULL, G, alm, NULL, e, nthreads);
    return 0;
}

/* ---------------------------------------------------------------------------------------------
 * Spin-weighted harmonics (Spin.jl): sY_lm(theta, phi) = N_l d^l_{m,-s}(theta) e^{i m phi},
 * N_l = sqrt((2l+1)/4pi), with d^l by the three-term recurrence in degree at fixed (m, -s), seeded
 * at l0 = max(|m|, |s|) from the extremal closed form in log2 magnitude.
 * ------------------------------------------------------------------------------------------- */
static double spin_Nl(int l) { return sqrt((double)(2 * l + 1) / (4 * PI)); } /* Spin.jl:21 */

/* log sqrt(C(2l, l-a)) (Spin.jl:26-31) */
static double half_lbinom(int l, int a) {
    return (lgamma((double)(2 * l + 1)) - lgamma((double)(l - a + 1)) - lgamma((double)(l + a + 1))) / 2;
}

/* d^l_{a,b}(beta) at |b| = l as (log|d|, sign) (Spin.jl:39-53) */
static double d_extremal(int l, int a, int b, double beta, double *sgn) {
    double c = cos(beta / 2), s = sin(beta / 2);
    double lb = half_lbinom(l, a);
    int pc, ps;
    double sg;
    if (b == l) { pc = l + a; ps = l - a; sg = 1.0; }
    else { pc = l - a; ps = l + a; sg = ((l - a) & 1) ? -1.0 : 1.0; }
    double lc = pc == 0 ? 0.0 : (c > 0 ? (double)pc * log(c) : -INFINITY);
    double ls = ps == 0 ? 0.0 : (s > 0 ? (double)ps * log(s) : -INFINITY);
    *sgn = sg;
    return lb + lc + ls;
}

/* wigner_d(l, m, n, beta) (Spin.jl:61-137): exact endpoints, extremal seed, block-scaled
 * recurrence renormalised every RESCALE_STRIDE degrees. */
double or_wigner_d(int l, int m, int n, double beta) {
    int am = m < 0 ? -m : m, an = n < 0 ? -n : n;
    if (am > l || an > l) return 0.0;
    double tiny = sqrt(DBL_EPSILON);
    if (beta <= tiny) return m == n ? 1.0 : 0.0;
    if (beta >= PI - tiny) return m == -n ? (((l - n) & 1) ? -1.0 : 1.0) : 0.0;
    int l0 = am > an ? am : an;
    double lg, sg;
    if (an == l0) lg = d_extremal(l0, m, n == l0 ? l0 : -l0, beta, &sg);
    else {
        lg = d_extremal(l0, n, m == l0 ? l0 : -l0, beta, &sg);
        if ((n - m) & 1) sg = -sg;
    }
    if (!isfinite(lg)) return 0.0;
    double E = ceil(lg / log(2.0));
    double d1 = sg * exp(lg - E * log(2.0));
    if (l == l0) return d1 * exp2(E);
    double cb = cos(beta), d0 = 0.0;
    int Lc = l0;
    if (l0 == 0) {
        d0 = d1;
        d1 = d1 * cb;
        Lc = 1;
        if (l == 1) return d1 * exp2(E);
    }
    int steps = 0;
    while (Lc < l) {
        int L1 = Lc + 1;
        double sm1 = sqrt(fmax(0.0, (double)(L1 * L1 - m * m)));
        double sn1 = sqrt(fmax(0.0, (double)(L1 * L1 - n * n)));
        double den = (double)Lc * sm1 * sn1;
        if (den <= 0) { d0 = d1; d1 = 0.0; Lc = L1; continue; }
        double sm0 = sqrt(fmax(0.0, (double)(Lc * Lc - m * m)));
        double sn0 = sqrt(fmax(0.0, (double)(Lc * Lc - n * n)));
        double c1 = (double)(2 * Lc + 1) * ((double)Lc * (double)L1 * cb - (double)m * (double)n) / den;
        double c0 = (double)L1 * sm0 * sn0 / den;
        double dn = c1 * d1 - c0 * d0;
        d0 = d1;
        d1 = dn;
        Lc = L1;
        if (++steps == RESCALE_STRIDE) {
            steps = 0;
            double mx = fmax(fabs(d1), fabs(d0));
            if (mx > 0 && isfinite(mx)) {
                double sh = ceil(log2(mx));
                if (sh != 0) { double f = exp2(-sh); d1 *= f; d0 *= f; E += sh; }
            }
        }
    }
    return d1 * exp2(E);
}

/* sYlm(s, l, m, theta, phi) (Spin.jl:144-149) */
void or_sylm(int s, int l, int m, double theta, double phi, double *re, double *im) {
    int as = s < 0 ? -s : s;
    if (as > l) { *re = 0; *im = 0; return; }
    double v = spin_Nl(l) * or_wigner_d(l, m, -s, theta);
    *re = v * cos((double)m * phi);
    *im = v * sin((double)m * phi);
}

struct or_spin_rings { /* SpinRings (Spin.jl:170-190) */
    int L, s, B, n, mmax;
    double eps;
    double *costh, *l2c, *l2s, *Nl, *lb2, *sg; /* lb2, sg per order m + L */
    int32_t *pc, *ps, *mlim, *jlo, *jhi;       /* jlo, jhi per |m| */
};

void or_spin_rings_destroy(or_spin_rings *e) {
    if (!e) return;
    free(e->costh); free(e->l2c); free(e->l2s); free(e->Nl); free(e->lb2); free(e->sg);
    free(e->pc); free(e->ps); free(e->mlim); free(e->jlo); free(e->jhi);
    free(e);
}

/* Spin.jl:192-270: half-angle logarithms per ring (swapped under reflection, with cos theta
 * negated), retained orders max(polar_mmax, |s|), and the seed tables per order. */
or_spin_rings *or_spin_rings_create(const double *theta, int n, int L, int spin, double eps, int ntrans,
                                    int reflect) {
    if (L < 0) { set_err("lmax must be nonnegative"); return NULL; }
    if (ntrans < 1) { set_err("ntrans must be >= 1"); return NULL; }
    if (n < 1) { set_err("need at least one ring"); return NULL; }
    if (spin == 0) { set_err("spin 0 uses the scalar ring engine"); return NULL; }
    int as = spin < 0 ? -spin : spin;
    if (as > L) { set_err("|spin| > lmax"); return NULL; }
    for (int j = 1; j < n; j++)
        if (theta[j] < theta[j - 1]) { set_err("ring colatitudes must be sorted ascending"); return NULL; }
    or_spin_rings *e = calloc(1, sizeof *e);
    e->L = L; e->s = spin; e->B = ntrans; e->n = n; e->eps = eps;
    e->costh = malloc(sizeof(double) * n);
    e->l2c = malloc(sizeof(double) * n);
    e->l2s = malloc(sizeof(double) * n);
    e->mlim = malloc(sizeof(int32_t) * n);
    e->jlo = malloc(sizeof(int32_t) * (L + 1));
    e->jhi = malloc(sizeof(int32_t) * (L + 1));
    e->Nl = malloc(sizeof(double) * (L + 1));
    e->lb2 = malloc(sizeof(double) * (2 * L + 1));
    e->sg = malloc(sizeof(double) * (2 * L + 1));
    e->pc = malloc(sizeof(int32_t) * (2 * L + 1));
    e->ps = malloc(sizeof(int32_t) * (2 * L + 1));
    double csgn = reflect ? -1.0 : 1.0;
    e->mmax = 0;
    for (int j = 0; j < n; j++) {
        double t = theta[j];
        e->costh[j] = csgn * cos(t);
        double ch = cos(t / 2), sh = fabs(sin(t / 2));
        if (reflect) { double q = ch; ch = sh; sh = q; }
        e->l2c[j] = ch > 0 ? log2(ch) : -INFINITY;
        e->l2s[j] = sh > 0 ? log2(sh) : -INFINITY;
        int ml = or_polar_mmax(t, L, eps);
        e->mlim[j] = ml > as ? ml : as;
        if (e->mlim[j] > e->mmax) e->mmax = e->mlim[j];
    }
    if (or_nested_ranges(e->mlim, n, L, e->jlo, e->jhi) != 0) { or_spin_rings_destroy(e); return NULL; }
    int nn = -spin, ann = as;
    double inv_ln2 = 1.0 / log(2.0);
    for (int m = -L; m <= L; m++) {
        int i = m + L, am = m < 0 ? -m : m;
        int l0 = am > ann ? am : ann;
        if (l0 > L) { e->lb2[i] = -INFINITY; e->pc[i] = 0; e->ps[i] = 0; e->sg[i] = 1.0; continue; }
        int a, b;
        double extra, sgn;
        if (ann == l0) { a = m; b = nn == l0 ? l0 : -l0; extra = 1.0; }
        else { a = nn; b = m == l0 ? l0 : -l0; extra = ((nn - m) & 1) ? -1.0 : 1.0; }
        if (b == l0) { e->pc[i] = l0 + a; e->ps[i] = l0 - a; sgn = 1.0; }
        else { e->pc[i] = l0 - a; e->ps[i] = l0 + a; sgn = ((l0 - a) & 1) ? -1.0 : 1.0; }
        e->lb2[i] = half_lbinom(l0, a) * inv_ln2;
        e->sg[i] = sgn * extra;
    }
    for (int l = 0; l <= L; l++) e->Nl[l] = spin_Nl(l);
    return e;
}

int or_spin_rings_mmax(const or_spin_rings *e) { return e->mmax; }
int64_t or_spin_rings_work(const or_spin_rings *e) { /* retained (l, m, ring) triples over m in [-L, L] */
    int64_t w = 0;
    int as = e->s < 0 ? -e->s : e->s;
    for (int m = -e->L; m <= e->L; m++) {
        int am = m < 0 ? -m : m, l0 = am > as ? am : as;
        int64_t nj = e->jhi[am] - e->jlo[am] + 1;
        if (nj > 0 && l0 <= e->L) w += nj * (int64_t)(e->L - l0 + 1);
    }
    return w;
}

/* _spin_chunk! (Spin.jl:377-489) for one ring: the ring carries its own power-of-two exponent
 * (the reference shares one per chunk of rings; the values agree to round-off). */
static void spin_ring(int fwd, or_cplx *G, const or_cplx *Gin, or_cplx *alm_out, const or_cplx *alm_in,
                      const or_spin_rings *e, int m, int j, or_cplx *acc) {
    int L = e->L, B = e->B, n = -e->s, am = m < 0 ? -m : m, an = n < 0 ? -n : n;
    int l0 = am > an ? am : an;
    if (l0 > L) return;
    int mi = m + L;
    int64_t astride = (int64_t)(L + 1) * (2 * L + 1), col = (int64_t)(L + 1) * mi;
    int64_t gi = j + (int64_t)e->n * mi, gstride = (int64_t)e->n * (2 * L + 1);
    double ex = e->lb2[mi];
    if (e->pc[mi] != 0) ex += (double)e->pc[mi] * e->l2c[j];
    if (e->ps[mi] != 0) ex += (double)e->ps[mi] * e->l2s[j];
    double E = isfinite(ex) ? ceil(ex) : 0.0;
    double d1 = isfinite(ex) ? e->sg[mi] * exp2(ex - E) : 0.0, d0 = 0.0;
    double cs = e->costh[j];
    for (int b = 0; b < B; b++) acc[b] = fwd ? cx(0, 0) : Gin[gi + b * gstride];
    /* _spin_touch! at l0 */
    if (fwd) {
        for (int b = 0; b < B; b++) acc[b] = cadd(acc[b], cscale(alm_in[col + l0 + b * astride], e->Nl[l0] * d1));
    } else {
        double f = exp2(E) * e->Nl[l0];
        for (int b = 0; b < B; b++) {
            or_cplx *o = &alm_out[col + l0 + b * astride];
            *o = cadd(*o, cscale(cscale(acc[b], d1), f));
        }
    }
    int steps = 0;
    for (int Lc = l0; Lc < L;) {
        int L1 = Lc + 1;
        double sm1 = sqrt(fmax(0.0, (double)(L1 * L1 - m * m)));
        double sn1 = sqrt(fmax(0.0, (double)(L1 * L1 - n * n)));
        double den = (double)Lc * sm1 * sn1;
        if (den <= 0) { d0 = d1; d1 = 0.0; Lc = L1; continue; }
        double sm0 = sqrt(fmax(0.0, (double)(Lc * Lc - m * m)));
        double sn0 = sqrt(fmax(0.0, (double)(Lc * Lc - n * n)));
        double a1 = (double)(2 * Lc + 1) * (double)Lc * (double)L1 / den;
        double a2 = (double)(2 * Lc + 1) * (double)m * (double)n / den;
        double a0 = (double)L1 * sm0 * sn0 / den;
        double dn = (a1 * cs - a2) * d1 - a0 * d0;
        d0 = d1;
        d1 = dn;
        Lc = L1;
        double N1 = e->Nl[L1];
        if (fwd) {
            for (int b = 0; b < B; b++) acc[b] = cadd(acc[b], cscale(alm_in[col + L1 + b * astride], N1 * dn));
        } else {
            double f = exp2(E) * N1;
            for (int b = 0; b < B; b++) {
                or_cplx *o = &alm_out[col + L1 + b * astride];
                *o = cadd(*o, cscale(cscale(acc[b], dn), f));
            }
        }
        if (++steps == RESCALE_STRIDE && Lc < L) {
            steps = 0;
            double mx = fmax(fabs(d1), fabs(d0));
            if (mx > 0 && isfinite(mx)) {
                double sh = ceil(log2(mx));
                if (sh != 0) {
                    double f = exp2(-sh);
                    d1 *= f;
                    d0 *= f;
                    if (fwd)
                        for (int b = 0; b < B; b++) acc[b] = cscale(acc[b], f);
                    E += sh;
                }
            }
        }
    }
    if (fwd) {
        double f = exp2(E);
        for (int b = 0; b < B; b++) G[gi + b * gstride] = cadd(G[gi + b * gstride], cscale(acc[b], f));
    }
}

static void spin_walk(int fwd, or_cplx *G, const or_cplx *Gin, or_cplx *alm_out, const or_cplx *alm_in,
                      const or_spin_rings *e, int nthreads) {
    int L = e->L;
    int ns = nthreads < 1 ? 1 : nthreads;
    (void)ns;
#pragma omp parallel for num_threads(ns) schedule(dynamic, 1)
    for (int mi = 0; mi <= 2 * L; mi++) { /* orders are independent: each owns its alm column */
        int m = mi - L, am = m < 0 ? -m : m;
        or_cplx *acc = malloc(sizeof(or_cplx) * e->B);
        for (int j = e->jlo[am]; j <= e->jhi[am]; j++) spin_ring(fwd, G, Gin, alm_out, alm_in, e, m, j, acc);
        free(acc);
    }
}

/* spin_rings_from_alm! / spin_alm_from_rings! (Spin.jl:316-334); alm in the dense (l, m + L) layout */
int or_spin_rings_from_alm(const or_spin_rings *e, const or_cplx *alm, or_cplx *G, int nthreads) {
    memset(G, 0, sizeof(or_cplx) * (size_t)e->n * (2 * e->L + 1) * e->B);
    spin_walk(1, G, NULL, NULL, alm, e, nthreads);
    return 0;
}
int or_spin_alm_from_rings(const or_spin_rings *e, const or_cplx *G, or_cplx *alm, int nthreads) {
    memset(alm, 0, sizeof(or_cplx) * (size_t)(e->L + 1) * (2 * e->L + 1) * e->B);
    spin_walk(0, NULL, G, alm, NULL, e, nthreads);
    return 0;
}

/* Legendre.jl:555-596 */
double or_alf(int l, int m, double theta) {
    int am = m < 0 ? -m : m;
    if (am > l) return 0.0;
    double s = fabs(sin(theta)), x = cos(theta);
    double log2c = -log2(sqrt(4 * PI));
    for (int k = 1; k <= am; k++) log2c += log2(sqrt((double)(2 * k + 1) / (double)(2 * k)));
    double e0 = am == 0 ? log2c : (double)am * (s > 0 ? log2(s) : -INFINITY) + log2c;
    if (!isfinite(e0)) return 0.0;
    double E = ceil(e0);
    double p2 = ((am % 2 == 0) ? 1.0 : -1.0) * exp2(e0 - E);
    double flip = (m < 0 && (am % 2 == 1)) ? -1.0 : 1.0;
    if (l == am) return flip * (p2 * exp2(E));
    double p1 = sqrt((double)(2 * am + 3)) * x * p2;
    int steps = 0;
    for (int el = am + 2; el <= l; el++) {
        double a = sqrt((double)(2 * el - 1) * (double)(2 * el + 1) / ((double)(el - am) * (double)(el + am)));
        double b = sqrt((double)(2 * el + 1) / (double)(2 * el - 3) * (double)(el - am - 1) /
                        (double)(el - am) * (double)(el + am - 1) / (double)(el + am));
        double p = a * x * p1 - b * p2;
        p2 = p1;
        p1 = p;
        steps++;
        if (steps == RESCALE_STRIDE) {
            steps = 0;
            double mx = fmax(fabs(p1), fabs(p2));
            if (mx > 0 && isfinite(mx)) {
                double sh = ceil(log2(mx));
                if (sh != 0) {
                    double f = exp2(-sh);
                    p1 *= f;
                    p2 *= f;
                    E += sh;
                }
            }
        }
    }
    return flip * (p1 * exp2(E));
}

/* ---------------------------------------------------------------------------------------------
 * Windows.jl
 * ------------------------------------------------------------------------------------------- */
static const double GL10_X[10] = {-0.9739065285171717, -0.8650633666889845, -0.6794095682990244,
                                  -0.4333953941292472, -0.1488743389816312, 0.1488743389816312,
                                  0.4333953941292472,  0.6794095682990244,  0.8650633666889845,
                                  0.9739065285171717}; /* Windows.jl:187-190 */
static const double GL10_W[10] = {0.0666713443086881, 0.1494513491505806, 0.2190863625159820,
                                  0.2692667193099963, 0.2955242247147529, 0.2955242247147529,
                                  0.2692667193099963, 0.2190863625159820, 0.1494513491505806,
                                  0.0666713443086881}; /* Windows.jl:191-194 */
#define ES_PANEL 0.05                                  /* Windows.jl:197 */
#define EPS_APO_MAX 1.0                                /* Windows.jl:20 */
#define K_MARGIN 2                                     /* Stretch.jl:86 */

/* Windows.jl:208-226 */
double or_es_cdf(double beta, double u) {
    double a = -PI / 2;
    if (u <= a) return 0.0;
    double b = fmin(u, PI / 2);
    int n = (int)ceil((b - a) / ES_PANEL);
    if (n < 1) n = 1;
    double hw = (b - a) / (2 * n);
    double acc = 0;
    for (int p = 0; p < n; p++) {
        double c = a + (2 * p + 1) * hw;
        double s = 0;
        for (int q = 0; q < 10; q++) {
            double t = c + hw * GL10_X[q];
            double ct = cos(t);
            s += GL10_W[q] * exp(beta * (ct - 1)) * ct;
        }
        acc += hw * s;
    }
    return acc;
}

/* Windows.jl:84-141. Returns 0 with *out filled, 1 when no periodisation is admissible,
 * -1 on an invalid argument. */
int or_stretch_geometry(double thlo, double thhi, int L, double eps, double epsapo_max, or_geom *out) {
    double lo = thlo, hi = thhi;
    if (!(lo <= hi)) { set_err("thlo = %g exceeds thhi = %g", lo, hi); return -1; }
    if (!(lo >= 0 && hi <= PI)) { set_err("colatitudes must lie in [0, pi]"); return -1; }
    double beta = fabs(log(eps));
    double tiny = sqrt(DBL_EPSILON);
    double Z = or_es_cdf(beta, PI / 2);
    double invZ = Z > 0 ? 1.0 / Z : 0.0;
    if (lo <= tiny && hi >= PI - tiny) {
        or_geom g = {OR_KIND_FULL, L, L, eps, beta, PI / 2, 0.0, PI, 0.0, 2 * PI, PI, 0.0, PI, 1.0};
        *out = g;
        return 0;
    }
    struct { int kind; double h, klo, khi; } cands[3] = {
        {OR_KIND_POLE, hi, 0.0, hi}, {OR_KIND_POLE, PI - lo, lo, PI}, {OR_KIND_INTERIOR, (hi - lo) / 2, lo, hi}};
    int bestK = 0x7fffffff, found = 0;
    for (int c = 0; c < 3; c++) {
        double h = cands[c].h;
        if (!(h > 0)) continue;
        double Lh = (double)L * h;
        if (!(Lh >= 2 * beta / (epsapo_max * epsapo_max))) continue;
        double ea = sqrt(2 * beta / Lh);
        double pad = h * ea;
        if (cands[c].kind == OR_KIND_INTERIOR) {
            if (!(cands[c].klo - pad >= 0 && cands[c].khi + pad <= PI)) continue;
        } else {
            if (!(h * (1 + ea) <= PI)) continue;
        }
        double kk = ceil(Lh * ((1 + ea) * (1 + ea)) / PI);
        int K = kk < 1 ? 1 : (int)kk;
        if (K < bestK) {
            bestK = K;
            double T1 = PI / (1 + ea);
            double sg = (PI - T1) / 2;
            or_geom g = {cands[c].kind, L, K, eps, beta, h, cands[c].klo, cands[c].khi, ea,
                         2 * h * (1 + ea), T1, sg, T1 + sg, invZ};
            *out = g;
            found = 1;
        }
    }
    if (!found) {
        set_err("stretched transform is degenerate for theta in [%g, %g] at lmax = %d, tol = %g", lo, hi, L, eps);
        return 1;
    }
    return 0;
}

/* Windows.jl:237-250 */
double or_stretch_window(const or_geom *g, double Theta) {
    if (g->kind == OR_KIND_FULL) return 1.0;
    double d = g->kind == OR_KIND_POLE ? fabs(Theta) : fabs(Theta - PI);
    if (d <= g->Theta1) return 1.0;
    if (d >= PI) return 0.0;
    double sg = g->sigma;
    if (!(sg > 0)) return d <= g->ThetaT ? 1.0 : 0.0;
    double zlo = fmin(fmax((d - g->ThetaT) / sg, -1.0), 1.0);
    double zhi = fmin(fmax((d + g->ThetaT) / sg, -1.0), 1.0);
    double v = (or_es_cdf(g->beta, asin(zhi)) - or_es_cdf(g->beta, asin(zlo))) * g->invZ;
    return fmin(fmax(v, 0.0), 1.0);
}

/* Windows.jl:272-282 */
double or_stretch_to_theta(const or_geom *g, double Theta) {
    if (g->kind == OR_KIND_FULL) return Theta;
    if (g->kind == OR_KIND_POLE) {
        double thmax = g->h * (1 + g->epsapo);
        double t = Theta * thmax / PI;
        return g->thlo <= sqrt(DBL_EPSILON) ? t : PI - t;
    }
    double lo = g->thlo - g->h * g->epsapo;
    return lo + g->period * Theta / (2 * PI);
}

/* Windows.jl:289-298 */
double or_theta_to_stretch(const or_geom *g, double theta) {
    if (g->kind == OR_KIND_FULL) return theta;
    if (g->kind == OR_KIND_POLE) {
        double thmax = g->h * (1 + g->epsapo);
        double d = g->thlo <= sqrt(DBL_EPSILON) ? theta : PI - theta;
        return PI * d / thmax;
    }
    double lo = g->thlo - g->h * g->epsapo;
    return 2 * PI * (theta - lo) / g->period;
}

/* Windows.jl:307-320 */
static void stretch_rings(const or_geom *g, int n, double *th, double *Th) {
    for (int i = 1; i <= n; i++) {
        double T = g->kind == OR_KIND_INTERIOR ? 2 * PI * (double)(i - 1) / n : PI * ((double)i - 0.5) / n;
        Th[i - 1] = T;
        th[i - 1] = or_stretch_to_theta(g, T);
    }
}

/* Stretch.jl:61-81 */
int64_t or_next_smooth(int64_t n) {
    if (n <= 1) return 1;
    int64_t best = INT64_MAX;
    for (int64_t p2 = 1; p2 < 2 * n + 2; p2 *= 2)
        for (int64_t p3 = p2; p3 < 2 * n + 2; p3 *= 3)
            for (int64_t p5 = p3; p5 < 2 * n + 2; p5 *= 5)
                if (p5 >= n && p5 < best) best = p5;
    return best == INT64_MAX ? n : best;
}
/* Stretch.jl:52 */
int64_t or_next_smooth_even(int64_t n) { return 2 * or_next_smooth((n + 1) / 2); }

/* ---------------------------------------------------------------------------------------------
 * DFS.jl:13-56 (column-major F: nth x nph x B, Ft: 2nth x nph x B)
 * ------------------------------------------------------------------------------------------- */
void or_dfs_double(const or_cplx *F, or_cplx *Ft, int nth, int nph, int B, int s) {
    int half = nph / 2;
    double sgn = (s % 2 == 0) ? 1.0 : -1.0;
    for (int b = 0; b < B; b++) {
        const or_cplx *Fb = F + (int64_t)b * nth * nph;
        or_cplx *Tb = Ft + (int64_t)b * 2 * nth * nph;
        for (int j = 0; j < nph; j++)
            for (int i = 0; i < nth; i++) Tb[i + (int64_t)j * 2 * nth] = Fb[i + (int64_t)j * nth];
        for (int j = 0; j < nph; j++) {
            int js = (j + half) % nph;
            for (int i = 0; i < nth; i++)
                Tb[nth + i + (int64_t)j * 2 * nth] = cscale(Fb[(nth - 1 - i) + (int64_t)js * nth], sgn);
        }
    }
}

void or_dfs_fold(const or_cplx *Ft, or_cplx *F, int nth, int nph, int B, int s) {
    int half = nph / 2;
    double sgn = (s % 2 == 0) ? 1.0 : -1.0;
    for (int b = 0; b < B; b++) {
        or_cplx *Fb = F + (int64_t)b * nth * nph;
        const or_cplx *Tb = Ft + (int64_t)b * 2 * nth * nph;
        for (int j = 0; j < nph; j++) {
            int js = ((j - half) % nph + nph) % nph;
            for (int i = 0; i < nth; i++)
                Fb[i + (int64_t)j * nth] = cadd(Tb[i + (int64_t)j * 2 * nth],
                                                cscale(Tb[(2 * nth - 1 - i) + (int64_t)js * 2 * nth], sgn));
        }
    }
}

/* ---------------------------------------------------------------------------------------------
 * Mixed-radix FFT (stand-in for FFTW, Stretch.jl:222-225): X_k = sum_n x_n exp(sign 2 pi i k n / N)
 * for N = 2^a 3^b 5^c (the sizes next_smooth makes, Stretch.jl:61-81).
 *
 * Stockham self-sorting decimation in frequency (Van Loan, "Computational Frameworks for the FFT",
 * 1992, sec. 1.7): s interleaved sequences, element j of sequence q at x[q + s j]. A radix-r stage
 * of current length n = r m maps a_k = x[q + s (p + k m)] to
 *   y[q + s (r p + u)] = W_n^(p u) sum_k a_k W_r^(k u),  W_n = exp(sign 2 pi i / n),
 * then continues on length m with stride r s; the output comes out in natural order. The q loop is
 * contiguous, so transforms along the second dimension of a column-major grid run on blocks of rows
 * at once (s = rows in the block), which keeps the data in cache.
 * ------------------------------------------------------------------------------------------- */
typedef struct {
    int n, nst, rad[64]; /* 64 stages cover any int length */
    or_cplx *tw; /* tw[k] = exp(+2 pi i k / n), k < n */
} fft_plan_t;

static void fft_plan_init(fft_plan_t *pl, int n) {
    pl->n = n; pl->nst = 0;
    int r = n;
    while (r % 4 == 0) { pl->rad[pl->nst++] = 4; r /= 4; }
    while (r % 2 == 0) { pl->rad[pl->nst++] = 2; r /= 2; }
    while (r % 3 == 0) { pl->rad[pl->nst++] = 3; r /= 3; }
    while (r % 5 == 0) { pl->rad[pl->nst++] = 5; r /= 5; }
    if (r > 1) pl->rad[pl->nst++] = r; /* any other factor: a plain DFT stage */
    pl->tw = malloc(sizeof(or_cplx) * (size_t)(n > 0 ? n : 1));
    for (int j = 0; j < n; j++) pl->tw[j] = ccis(2 * PI * (double)j / (double)n);
}

static void fft_plan_free(fft_plan_t *pl) { free(pl->tw); }

static inline or_cplx tw_get(const fft_plan_t *pl, int64_t k, int sign) {
    or_cplx w = pl->tw[k];
    return sign < 0 ? cconj(w) : w;
}

/* one stage: length n (a divisor of pl->n) = r m, stride s; x -> y */
static void fft_stage(const fft_plan_t *pl, int r, int n, int64_t s, const or_cplx *x, or_cplx *y,
                      int sign) {
    int m = n / r;
    int64_t step = pl->n / n; /* W_n^j = W_N^(j step) */
    double sg = sign < 0 ? -1.0 : 1.0;
    for (int p = 0; p < m; p++) {
        or_cplx w1 = tw_get(pl, (int64_t)p * step, sign);
        or_cplx w2 = tw_get(pl, (int64_t)2 * p * step, sign);
        or_cplx w3 = r >= 4 ? tw_get(pl, (int64_t)3 * p * step, sign) : cx(1, 0);
        const or_cplx *xp = x + s * p;
        or_cplx *yp = y + s * (int64_t)r * p;
        if (r == 2) {
            for (int64_t q = 0; q < s; q++) {
                or_cplx a0 = xp[q], a1 = xp[q + s * m];
                yp[q] = cadd(a0, a1);
                yp[q + s] = cmul(cx(a0.re - a1.re, a0.im - a1.im), w1);
            }
        } else if (r == 4) {
            for (int64_t q = 0; q < s; q++) {
                or_cplx a0 = xp[q], a1 = xp[q + s * m], a2 = xp[q + 2 * s * m], a3 = xp[q + 3 * s * m];
                or_cplx s02 = cadd(a0, a2), d02 = cx(a0.re - a2.re, a0.im - a2.im);
                or_cplx s13 = cadd(a1, a3), d13 = cx(a1.re - a3.re, a1.im - a3.im);
                or_cplx jd = cx(-sg * d13.im, sg * d13.re); /* W_4 d13, W_4 = sign i */
                yp[q] = cadd(s02, s13);
                yp[q + s] = cmul(cadd(d02, jd), w1);
                yp[q + 2 * s] = cmul(cx(s02.re - s13.re, s02.im - s13.im), w2);
                yp[q + 3 * s] = cmul(cx(d02.re - jd.re, d02.im - jd.im), w3);
            }
        } else if (r == 3) {
            const double h = 0.86602540378443864676; /* sqrt(3) / 2 */
            for (int64_t q = 0; q < s; q++) {
                or_cplx a0 = xp[q], a1 = xp[q + s * m], a2 = xp[q + 2 * s * m];
                or_cplx t = cadd(a1, a2), d = cx(a1.re - a2.re, a1.im - a2.im);
                or_cplx c = cx(a0.re - 0.5 * t.re, a0.im - 0.5 * t.im);
                or_cplx jd = cx(-sg * h * d.im, sg * h * d.re);
                yp[q] = cadd(a0, t);
                yp[q + s] = cmul(cadd(c, jd), w1);
                yp[q + 2 * s] = cmul(cx(c.re - jd.re, c.im - jd.im), w2);
            }
        } else {
            or_cplx wrs[64], as[64];
            or_cplx *wr = r <= 64 ? wrs : malloc(sizeof(or_cplx) * (size_t)r);
            or_cplx *a = r <= 64 ? as : malloc(sizeof(or_cplx) * (size_t)r);
            int64_t rstep = pl->n / r;
            for (int u = 0; u < r; u++) wr[u] = tw_get(pl, (int64_t)u * rstep, sign);
            for (int64_t q = 0; q < s; q++) {
                for (int k = 0; k < r; k++) a[k] = xp[q + s * (int64_t)k * m];
                for (int u = 0; u < r; u++) {
                    or_cplx acc = cx(0, 0);
                    for (int k = 0; k < r; k++) acc = cadd(acc, cmul(a[k], wr[(k * u) % r]));
                    yp[q + s * u] = u ? cmul(acc, tw_get(pl, (int64_t)p * u * step, sign)) : acc;
                }
            }
            if (wr != wrs) { free(wr); free(a); }
        }
    }
}

/* s interleaved transforms of length pl->n in x (work array y of the same size) */
static void fft_run(const fft_plan_t *pl, or_cplx *x, or_cplx *y, int64_t s, int sign) {
    int n = pl->n;
    or_cplx *a = x, *b = y;
    for (int st = 0; st < pl->nst; st++) {
        int r = pl->rad[st];
        fft_stage(pl, r, n, s, a, b, sign);
        n /= r; s *= r;
        or_cplx *t = a; a = b; b = t;
    }
    if (a != x) memcpy(x, a, sizeof(or_cplx) * (size_t)s);
}

void or_fft1(or_cplx *x, int n, int sign) {
    fft_plan_t pl;
    fft_plan_init(&pl, n);
    or_cplx *buf = malloc(sizeof(or_cplx) * (size_t)(n > 0 ? n : 1));
    fft_run(&pl, x, buf, 1, sign);
    free(buf);
    fft_plan_free(&pl);
}

enum { FFT_ROWS = 8 }; /* rows per block for transforms along dim 2 */

/* 1-D transforms along dim 2 of column-major n1 x n2 x B */
static void fft_dim2(or_cplx *x, int n1, int n2, int B, int sign, int nthreads) {
    fft_plan_t pl;
    fft_plan_init(&pl, n2);
    int nt = nthreads < 1 ? 1 : nthreads;
    int64_t nblk = ((int64_t)n1 + FFT_ROWS - 1) / FFT_ROWS;
#pragma omp parallel num_threads(nt)
    {
        or_cplx *blk = malloc(sizeof(or_cplx) * (size_t)n2 * FFT_ROWS);
        or_cplx *buf = malloc(sizeof(or_cplx) * (size_t)n2 * FFT_ROWS);
#pragma omp for schedule(static)
        for (int64_t t = 0; t < nblk * B; t++) {
            int64_t b = t / nblk, i0 = (t % nblk) * FFT_ROWS;
            int v = (int)(n1 - i0 < FFT_ROWS ? n1 - i0 : FFT_ROWS);
            or_cplx *base = x + b * (int64_t)n1 * n2 + i0;
            for (int j = 0; j < n2; j++)
                for (int i = 0; i < v; i++) blk[i + (int64_t)j * v] = base[i + (int64_t)j * n1];
            fft_run(&pl, blk, buf, v, sign);
            for (int j = 0; j < n2; j++)
                for (int i = 0; i < v; i++) base[i + (int64_t)j * n1] = blk[i + (int64_t)j * v];
        }
        free(blk); free(buf);
    }
    fft_plan_free(&pl);
}

/* 1-D transforms along dim 1: blocks of FFT_ROWS columns are interleaved (a small transpose) so
 * that every stage runs over FFT_ROWS sequences at once */
static void fft_dim1(or_cplx *x, int n1, int n2, int B, int sign, int nthreads) {
    fft_plan_t pl;
    fft_plan_init(&pl, n1);
    int nt = nthreads < 1 ? 1 : nthreads;
    int64_t ncol = (int64_t)n2 * B, nblk = (ncol + FFT_ROWS - 1) / FFT_ROWS;
#pragma omp parallel num_threads(nt)
    {
        or_cplx *blk = malloc(sizeof(or_cplx) * (size_t)n1 * FFT_ROWS);
        or_cplx *buf = malloc(sizeof(or_cplx) * (size_t)n1 * FFT_ROWS);
#pragma omp for schedule(static)
        for (int64_t t = 0; t < nblk; t++) {
            int64_t c0 = t * FFT_ROWS;
            int v = (int)(ncol - c0 < FFT_ROWS ? ncol - c0 : FFT_ROWS);
            or_cplx *base = x + c0 * n1;
            for (int c = 0; c < v; c++)
                for (int i = 0; i < n1; i++) blk[c + (int64_t)i * v] = base[i + (int64_t)c * n1];
            fft_run(&pl, blk, buf, v, sign);
            for (int c = 0; c < v; c++)
                for (int i = 0; i < n1; i++) base[i + (int64_t)c * n1] = blk[c + (int64_t)i * v];
        }
        free(blk); free(buf);
    }
    fft_plan_free(&pl);
}

void or_fft2(or_cplx *x, int n1, int n2, int B, int sign, int nthreads) {
    fft_dim1(x, n1, n2, B, sign, nthreads);
    fft_dim2(x, n1, n2, B, sign, nthreads);
}

/* ---------------------------------------------------------------------------------------------
 * NUFFT.jl:74-83 (_mode_k, modeord 1) and the direct sum NUFFT.jl:85-129.
 * ------------------------------------------------------------------------------------------- */
static inline int mode_k(int i0, int n) { /* 0-based i0 */
    int k = i0;
    return k < (n + 1) / 2 ? k : k - n;
}

int or_nufft_direct(int type, int n1, int n2, int iflag, int B, const double *x, const double *y,
                    int64_t M, const or_cplx *in, or_cplx *out) {
    double s = (double)iflag;
    if (type == 2) {
        for (int b = 0; b < B; b++)
            for (int64_t j = 0; j < M; j++) {
                or_cplx acc = cx(0, 0);
                for (int i2 = 0; i2 < n2; i2++) {
                    int k2 = mode_k(i2, n2);
                    for (int i1 = 0; i1 < n1; i1++) {
                        int k1 = mode_k(i1, n1);
                        or_cplx c = in[i1 + (int64_t)i2 * n1 + (int64_t)b * n1 * n2];
                        acc = cadd(acc, cmul(c, ccis(s * ((double)k1 * x[j] + (double)k2 * y[j]))));
                    }
                }
                out[j + b * M] = acc;
            }
        return 0;
    }
    if (type == 1) {
        memset(out, 0, sizeof(or_cplx) * (int64_t)n1 * n2 * B);
        for (int b = 0; b < B; b++)
            for (int64_t j = 0; j < M; j++) {
                or_cplx fj = in[j + b * M];
                if (fj.re == 0 && fj.im == 0) continue;
                for (int i2 = 0; i2 < n2; i2++) {
                    int k2 = mode_k(i2, n2);
                    for (int i1 = 0; i1 < n1; i1++) {
                        int k1 = mode_k(i1, n1);
                        int64_t o = i1 + (int64_t)i2 * n1 + (int64_t)b * n1 * n2;
                        out[o] = cadd(out[o], cmul(fj, ccis(s * ((double)k1 * x[j] + (double)k2 * y[j]))));
                    }
                }
            }
        return 0;
    }
    set_err("DirectSum NUFFT type must be 1 or 2, got %d", type);
    return -1;
}

/* ---------------------------------------------------------------------------------------------
 * ES-kernel NUFFT (stand-in for FINUFFT). Kernel psi(x) = phi(x / alpha), phi(z) =
 * exp(beta (sqrt(1 - z^2) - 1)) on |z| < 1, alpha = w h / 2, fine grid N >= 2 n (es_fine_size),
 * h = 2 pi / N. Type 2 with sign s:
 *   u_l = sum_k c_k (h / psihat(k)) e^{i s k l h},  f_j = sum_l psi(x_j - l h) u_l;
 * type 1 (sign s) is the exact transpose: spread, FFT with sign s, multiply by h / psihat(k).
 * Shape beta = 2.30 w as published for upsampling 2; the width is one wider than the published
 * w = ceil(log10(10/tol)), i.e. w = ceil(log10(1/tol)) + 2, which measured ~2e-9 relative error
 * at tol = 1e-8 on a white spectrum where the published width gave ~2e-8.
 * ------------------------------------------------------------------------------------------- */
void or_es_params(double tol, int *w, double *beta) {
    int ww = (int)ceil(-log10(tol)) + 2;
    if (ww < 2) ww = 2;
    if (ww > 16) ww = 16;
    double bo = 2.30;
    if (ww == 2) bo = 2.20;
    else if (ww == 3) bo = 2.26;
    else if (ww == 4) bo = 2.38;
    *w = ww;
    *beta = bo * ww;
}

/* The fine length is a free parameter of the ES NUFFT (any N >= 2n; a larger upsampling factor only
 * lowers the aliasing error at the same width). The oracle takes the library's choice so that the
 * two compare at round-off: next_smooth_even(max(2n, 2w)), and above 4096 the next of 6144, 8192,
 * 9216, 10000, 10240 (the lengths rocFFT runs in one kernel there). The ES checks against the direct
 * sum hold at every N. */
static int es_fine_size(int n, int w) {
    int64_t m = 2 * (int64_t)n > 2 * w ? 2 * (int64_t)n : 2 * w;
    int64_t N = or_next_smooth_even(m);
    static const int64_t fast[5] = {6144, 8192, 9216, 10000, 10240};
    if (N > 4096)
        for (int i = 0; i < 5; i++)
            if (fast[i] >= N) return (int)fast[i];
    return (int)N;
}

int or_es_fine_size(int n, int w) { return es_fine_size(n, w); }

/* Gauss-Legendre nodes on [-1,1] by Newton on P_n. */
static void gauss_legendre(int n, double *x, double *w) {
    for (int i = 0; i < n; i++) {
        double z = cos(PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; it++) {
            double p0 = 1, p1 = z;
            for (int k = 2; k <= n; k++) {
                double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            double dp = n * (z * p1 - p0) / (z * z - 1);
            double dz = p1 / dp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        double p0 = 1, p1 = z;
        for (int k = 2; k <= n; k++) {
            double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        double dp = n * (z * p1 - p0) / (z * z - 1);
        x[i] = z;
        w[i] = 2 / ((1 - z * z) * dp * dp);
    }
}

/* corr[i] = h / psihat(k_i), psihat(k) = alpha int_{-1}^{1} phi(z) cos(k alpha z) dz, evaluated with
 * z = sin t so the integrand is analytic: alpha int_{-pi/2}^{pi/2} e^{beta(cos t - 1)} cos(k alpha sin t) cos t dt. */
static void es_corr(int n, int N, int w, double beta, double *corr) {
    enum { NQ = 80 };
    double gx[NQ], gw[NQ];
    gauss_legendre(NQ, gx, gw);
    double h = 2 * PI / N, alpha = w * h / 2;
    for (int i = 0; i < n; i++) {
        int k = mode_k(i, n);
        double acc = 0;
        for (int q = 0; q < NQ; q++) {
            double t = gx[q] * PI / 2;
            acc += gw[q] * (PI / 2) * exp(beta * (cos(t) - 1)) * cos((double)k * alpha * sin(t)) * cos(t);
        }
        corr[i] = h / (alpha * acc);
    }
}

/* base cell of the w-cell window of point x (radians) on a fine grid of N cells, and the point's
 * fine-grid coordinate X in [0, N) */
static inline int es_base(double x, int N, int w, double *Xout) {
    double X = x * ((double)N / (2 * PI));
    X = X - floor(X / N) * N; /* [0, N) */
    *Xout = X;
    return (int)ceil(X - 0.5 * w);
}

/* kernel values at point x (radians) on a fine grid of N cells: base cell l0 (in [-w/2, N)), the w
 * weights, and the w window cells wrapped into [0, N) */
static inline int es_weights(double x, int N, int w, double beta, double *kv, int *cell) {
    double X;
    int l0 = es_base(x, N, w, &X);
    double inv = 2.0 / w;
    for (int t = 0; t < w; t++) {
        double z = ((double)l0 + t - X) * inv;
        double zz = 1 - z * z;
        kv[t] = zz >= 0 ? exp(beta * (sqrt(zz) - 1)) : 0.0;
        int c = l0 + t;
        cell[t] = c < 0 ? c + N : (c >= N ? c - N : c);
    }
    return l0;
}

typedef struct { int64_t key, j; } es_key_t;

static int es_key_cmp(const void *a, const void *b) {
    const es_key_t *x = a, *y = b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->j < y->j ? -1 : (x->j > y->j);
}

/* Points ordered by the fine-grid cell of their window (second coordinate first), as a fast NUFFT
 * sorts them when it is planned; a pure reordering of the per-point work, which changes no value
 * of type 2 and only the summation order of type 1. */
static int64_t *es_point_order(int n1, int n2, double tol, const double *x, const double *y, int64_t M) {
    int w;
    double beta, X;
    or_es_params(tol, &w, &beta);
    int N1 = es_fine_size(n1, w), N2 = es_fine_size(n2, w);
    es_key_t *k = malloc(sizeof(es_key_t) * (size_t)(M > 0 ? M : 1));
    for (int64_t j = 0; j < M; j++) {
        int64_t b1 = es_base(x[j], N1, w, &X) + w, b2 = es_base(y[j], N2, w, &X) + w;
        k[j].key = b2 * (int64_t)(N1 + 2 * w) + b1;
        k[j].j = j;
    }
    qsort(k, (size_t)M, sizeof(es_key_t), es_key_cmp);
    int64_t *ord = malloc(sizeof(int64_t) * (size_t)(M > 0 ? M : 1));
    for (int64_t j = 0; j < M; j++) ord[j] = k[j].j;
    free(k);
    return ord;
}

/* the ES NUFFT over the points in the given order (NULL: index order) */
static int nufft_es_ordered(int type, int n1, int n2, int iflag, int B, const double *x, const double *y,
                            int64_t M, double tol, const or_cplx *in, or_cplx *out, int nthreads,
                            const int64_t *order) {
    int w;
    double beta;
    or_es_params(tol, &w, &beta);
    int N1 = es_fine_size(n1, w), N2 = es_fine_size(n2, w);
    double *c1 = malloc(sizeof(double) * n1), *c2 = malloc(sizeof(double) * n2);
    es_corr(n1, N1, w, beta, c1);
    es_corr(n2, N2, w, beta, c2);
    int nt = nthreads < 1 ? 1 : nthreads;
    int64_t NN = (int64_t)N1 * N2;
    or_cplx *u = malloc(sizeof(or_cplx) * NN);
    for (int b = 0; b < B; b++) {
        if (type == 2) {
            memset(u, 0, sizeof(or_cplx) * NN);
            const or_cplx *cb = in + (int64_t)b * n1 * n2;
            for (int i2 = 0; i2 < n2; i2++) {
                int k2 = mode_k(i2, n2);
                int l2 = (k2 + N2) % N2;
                for (int i1 = 0; i1 < n1; i1++) {
                    int k1 = mode_k(i1, n1);
                    int l1 = (k1 + N1) % N1;
                    u[l1 + (int64_t)l2 * N1] = cscale(cb[i1 + (int64_t)i2 * n1], c1[i1] * c2[i2]);
                }
            }
            or_fft2(u, N1, N2, 1, iflag, nt);
#pragma omp parallel for num_threads(nt) schedule(static)
            for (int64_t jj = 0; jj < M; jj++) {
                int64_t j = order ? order[jj] : jj;
                double k1v[16], k2v[16];
                int i1v[16], i2v[16];
                es_weights(x[j], N1, w, beta, k1v, i1v);
                es_weights(y[j], N2, w, beta, k2v, i2v);
                or_cplx acc = cx(0, 0);
                for (int t2 = 0; t2 < w; t2++) {
                    const or_cplx *col = u + (int64_t)i2v[t2] * N1;
                    or_cplx row = cx(0, 0);
                    for (int t1 = 0; t1 < w; t1++) row = cadd(row, cscale(col[i1v[t1]], k1v[t1]));
                    acc = cadd(acc, cscale(row, k2v[t2]));
                }
                out[j + (int64_t)b * M] = acc;
            }
        } else if (type == 1) {
            const or_cplx *fb = in + (int64_t)b * M;
            /* thread t spreads a contiguous run of the ordered points into a private grid, summed
             * into u in thread order: deterministic for a fixed thread count. With ordered points
             * a run covers a band of columns, and only that band is kept. */
            or_cplx **priv = calloc((size_t)nt, sizeof(or_cplx *));
            int *lo = calloc((size_t)nt, sizeof(int)), *nc = calloc((size_t)nt, sizeof(int));
#pragma omp parallel num_threads(nt)
            {
                int tid = 0;
#ifdef _OPENMP
                tid = omp_get_thread_num();
#endif
                int64_t chunk = (M + nt - 1) / nt, j0 = tid * chunk, j1 = j0 + chunk < M ? j0 + chunk : M;
                int clo = 0, chi = N2 - 1;
                if (order && j1 > j0) { /* ordered: the band between the first and last window */
                    double X;
                    clo = es_base(y[order[j0]], N2, w, &X);
                    chi = es_base(y[order[j1 - 1]], N2, w, &X) + w - 1;
                    if (chi - clo + 1 >= N2) { clo = 0; chi = N2 - 1; }
                }
                int ncol = j1 > j0 ? chi - clo + 1 : 0;
                or_cplx *g = ncol ? calloc((size_t)ncol * N1, sizeof(or_cplx)) : NULL;
                priv[tid] = g; lo[tid] = clo; nc[tid] = ncol;
                int banded = order && ncol && !(clo == 0 && chi == N2 - 1);
                for (int64_t jj = j0; jj < j1; jj++) {
                    int64_t j = order ? order[jj] : jj;
                    double k1v[16], k2v[16];
                    int i1v[16], i2v[16];
                    es_weights(x[j], N1, w, beta, k1v, i1v);
                    int b2 = es_weights(y[j], N2, w, beta, k2v, i2v);
                    or_cplx fj = fb[j];
                    for (int t2 = 0; t2 < w; t2++) {
                        int c = banded ? b2 + t2 - clo : i2v[t2];
                        or_cplx *col = g + (int64_t)c * N1;
                        or_cplx v2 = cscale(fj, k2v[t2]);
                        for (int t1 = 0; t1 < w; t1++) col[i1v[t1]] = cadd(col[i1v[t1]], cscale(v2, k1v[t1]));
                    }
                }
            }
            memset(u, 0, sizeof(or_cplx) * NN);
            for (int t = 0; t < nt; t++) {
                if (!priv[t]) continue;
                int full = nc[t] == N2 && lo[t] == 0;
                for (int c = 0; c < nc[t]; c++) {
                    int l2 = full ? c : lo[t] + c;
                    l2 = l2 < 0 ? l2 + N2 : (l2 >= N2 ? l2 - N2 : l2);
                    or_cplx *dst = u + (int64_t)l2 * N1;
                    const or_cplx *src = priv[t] + (int64_t)c * N1;
                    for (int i = 0; i < N1; i++) dst[i] = cadd(dst[i], src[i]);
                }
                free(priv[t]);
            }
            free(priv); free(lo); free(nc);
            or_fft2(u, N1, N2, 1, iflag, nt);
            or_cplx *cb = out + (int64_t)b * n1 * n2;
            for (int i2 = 0; i2 < n2; i2++) {
                int k2 = mode_k(i2, n2);
                int l2 = (k2 + N2) % N2;
                for (int i1 = 0; i1 < n1; i1++) {
                    int k1 = mode_k(i1, n1);
                    int l1 = (k1 + N1) % N1;
                    cb[i1 + (int64_t)i2 * n1] = cscale(u[l1 + (int64_t)l2 * N1], c1[i1] * c2[i2]);
                }
            }
        } else {
            free(u); free(c1); free(c2);
            set_err("NUFFT type must be 1 or 2, got %d", type);
            return -1;
        }
    }
    free(u); free(c1); free(c2);
    return 0;
}

int or_nufft_es(int type, int n1, int n2, int iflag, int B, const double *x, const double *y,
                int64_t M, double tol, const or_cplx *in, or_cplx *out, int nthreads) {
    return nufft_es_ordered(type, n1, n2, iflag, B, x, y, M, tol, in, out, nthreads, NULL);
}

/* ---------------------------------------------------------------------------------------------
 * Plan (Plan.jl:84-116), planner (Cost.jl:104-140,174-201), leaves (Stretch.jl, Direct.jl).
 * ------------------------------------------------------------------------------------------- */
struct or_plan {
    int L, B, method, nufft_kind, nthreads;
    int64_t M;
    double tol;
    /* stretch */
    or_geom geom;
    int flip, nth_half, nth, nph, mmax;
    or_rings *eng;
    int spin;            /* spin weight; nonzero plans use the spin engine (Spin.jl) */
    or_spin_rings *seng;
    double *wth, *Tpts, *ppts;
    int64_t *es_order;  /* ES NUFFT point order, fixed at plan time */
    or_cplx *phase, *G, *F, *Ft, *Fhat, *fbuf;
    /* direct */
    int64_t *perm;
    double *ths, *phs;
    double times[3];
};

static double fftunits(int64_t n) { return n <= 1 ? 0.0 : (double)n * log2((double)n); }

/* Cost.jl:57 ratios */
#define COST_RING 1.0
#define COST_FFT 0.394
#define COST_SPREAD 0.334

static double cost_direct(int64_t n, int L, int B) { /* Cost.jl:104-105 */
    return COST_RING * (double)n * (double)(L + 1) * (double)(L + 1) * (double)B / 2;
}

static double cost_stretch(double thlo, double thhi, int64_t npts, int L, int B, double tol) {
    or_geom g;
    if (or_stretch_geometry(thlo, thhi, L, tol, EPS_APO_MAX, &g) != 0) return INFINITY; /* Cost.jl:121-122 */
    int Kt = g.K + K_MARGIN;
    int64_t nth, nth_half;
    if (g.kind == OR_KIND_INTERIOR) { nth = or_next_smooth(2 * Kt + 1); nth_half = nth; }
    else { nth_half = or_next_smooth(Kt + 1); nth = 2 * nth_half; }
    double *th = malloc(sizeof(double) * nth_half), *Th = malloc(sizeof(double) * nth_half);
    stretch_rings(&g, (int)nth_half, th, Th);
    int sorted = 1;
    for (int64_t i = 1; i < nth_half; i++) if (th[i] < th[i - 1]) sorted = 0;
    if (!sorted)
        for (int64_t i = 0; i < nth_half / 2; i++) { double t = th[i]; th[i] = th[nth_half - 1 - i]; th[nth_half - 1 - i] = t; }
    int32_t *mlim = malloc(sizeof(int32_t) * nth_half);
    int mx = 0;
    for (int64_t j = 0; j < nth_half; j++) {
        mlim[j] = or_polar_mmax(th[j], L, tol);
        if (mlim[j] > mx) mx = mlim[j];
    }
    int64_t rings = or_retained_triples(mlim, (int)nth_half, L);
    int64_t nph = or_next_smooth_even(2 * (int64_t)mx + 1);
    double fft = fftunits(nph) * (double)nth_half + fftunits(nth * nph);
    double spread = (double)npts * pow(2 * ceil(fabs(log10(fmax(tol, 1e-16)))) + 2, 2);
    free(th); free(Th); free(mlim);
    if (rings < 0) return INFINITY;
    return B * (COST_RING * (double)rings + COST_FFT * fft + COST_SPREAD * spread);
}

/* Cost.jl:114-143, for the partition restatement of oracle.py */
double or_cost_stretch(double thlo, double thhi, int64_t npts, int L, int B, double tol) {
    return cost_stretch(thlo, thhi, npts, L, B, tol);
}
double or_cost_direct(int64_t npts, int L, int B) { return cost_direct(npts, L, B); }

static double jmod(double x, double y) { /* Julia mod for floats */
    double r = fmod(x, y);
    if (r != 0 && ((r < 0) != (y < 0))) r += y;
    return r;
}

/* stable sort of theta: merge sort of indices */
static void msort_idx(int64_t *idx, int64_t *tmp, const double *key, int64_t n) {
    if (n < 2) return;
    int64_t h = n / 2;
    msort_idx(idx, tmp, key, h);
    msort_idx(idx + h, tmp, key, n - h);
    int64_t i = 0, j = h, k = 0;
    while (i < h && j < n) tmp[k++] = key[idx[j]] < key[idx[i]] ? idx[j++] : idx[i++];
    while (i < h) tmp[k++] = idx[i++];
    while (j < n) tmp[k++] = idx[j++];
    memcpy(idx, tmp, sizeof(int64_t) * n);
}

void or_plan_destroy(or_plan *p) {
    if (!p) return;
    or_rings_destroy(p->eng);
    or_spin_rings_destroy(p->seng);
    free(p->wth); free(p->Tpts); free(p->ppts); free(p->phase); free(p->G); free(p->F);
    free(p->Ft); free(p->Fhat); free(p->fbuf); free(p->perm); free(p->ths); free(p->phs);
    free(p->es_order);
    free(p);
}

or_plan *or_plan_create(const double *theta, const double *phi, int64_t M, int L, double tol, int ntrans,
                        int method, int nufft_kind, int nthreads) {
    return or_plan_create_spin(theta, phi, M, L, tol, ntrans, 0, method, nufft_kind, nthreads);
}

/* Plan(θ, φ; lmax, tol, spin, ntrans, method) (Plan.jl:84-116): spin != 0 takes SpinDirectLeaf
 * (Spin.jl:548-585) or StretchLeaf over SpinRings (Stretch.jl:190-194) */
or_plan *or_plan_create_spin(const double *theta, const double *phi, int64_t M, int L, double tol, int ntrans,
                             int spin, int method, int nufft_kind, int nthreads) {
    if (L < 0) { set_err("lmax must be nonnegative"); return NULL; }
    if (ntrans < 1) { set_err("ntrans must be >= 1"); return NULL; }
    if (M < 1) { set_err("need at least one point"); return NULL; }
    if (spin != 0 && abs(spin) > L) { set_err("(|spin| = %d) > lmax = %d", abs(spin), L); return NULL; }
    or_plan *p = calloc(1, sizeof *p);
    p->spin = spin;
    p->L = L; p->B = ntrans; p->M = M; p->tol = tol; p->nufft_kind = nufft_kind;
    p->nthreads = nthreads < 1 ? 1 : nthreads;
    /* bounding box colatitudes (Box.jl:43-73) */
    double thlo = INFINITY, thhi = -INFINITY;
    for (int64_t i = 0; i < M; i++) { if (theta[i] < thlo) thlo = theta[i]; if (theta[i] > thhi) thhi = theta[i]; }
    thlo = fmax(thlo, 0.0);
    thhi = fmin(thhi, PI);
    if (method == 0) { /* Cost.jl:174-201 restricted to Direct vs Stretch */
        double cd = cost_direct(M, L, ntrans), cs = cost_stretch(thlo, thhi, M, L, ntrans, tol);
        method = cs < cd ? 2 : 1;
    }
    p->method = method;
    if (method == 1) { /* Direct.jl:25-51 */
        p->perm = malloc(sizeof(int64_t) * M);
        int64_t *tmp = malloc(sizeof(int64_t) * M);
        for (int64_t i = 0; i < M; i++) p->perm[i] = i;
        msort_idx(p->perm, tmp, theta, M);
        free(tmp);
        p->ths = malloc(sizeof(double) * M);
        p->phs = malloc(sizeof(double) * M);
        for (int64_t i = 0; i < M; i++) { p->ths[i] = theta[p->perm[i]]; p->phs[i] = phi[p->perm[i]]; }
        if (M > 0x7fffffff) { set_err("too many points for the direct oracle"); or_plan_destroy(p); return NULL; }
        if (spin) p->seng = or_spin_rings_create(p->ths, (int)M, L, spin, 0.0, ntrans, 0);
        else p->eng = or_rings_create(p->ths, (int)M, L, 0.0, ntrans, 0);
        if (!p->eng && !p->seng) { or_plan_destroy(p); return NULL; }
        p->G = calloc((size_t)M * (2 * L + 1) * ntrans, sizeof(or_cplx));
        return p;
    }
    if (method != 2) { set_err("unknown method %d", method); or_plan_destroy(p); return NULL; }
    /* StretchLeaf (Stretch.jl:129-248) */
    double tiny = sqrt(DBL_EPSILON);
    or_geom g0;
    int rc = or_stretch_geometry(thlo, thhi, L, tol, EPS_APO_MAX, &g0);
    if (rc != 0) { or_plan_destroy(p); return NULL; }
    int flip = g0.kind == OR_KIND_POLE && g0.thhi >= PI - tiny && g0.thlo > tiny;
    if (flip) {
        rc = or_stretch_geometry(PI - thhi, PI - thlo, L, tol, EPS_APO_MAX, &p->geom);
        if (rc != 0) { or_plan_destroy(p); return NULL; }
    } else p->geom = g0;
    p->flip = flip;
    int Kt = p->geom.K + K_MARGIN;
    if (p->geom.kind == OR_KIND_INTERIOR) { p->nth = (int)or_next_smooth(2 * Kt + 1); p->nth_half = p->nth; }
    else { p->nth_half = (int)or_next_smooth(Kt + 1); p->nth = 2 * p->nth_half; }
    int nh = p->nth_half;
    double *th = malloc(sizeof(double) * nh), *Th = malloc(sizeof(double) * nh);
    stretch_rings(&p->geom, nh, th, Th);
    for (int i = 1; i < nh; i++)
        if (th[i] < th[i - 1]) { set_err("ring colatitudes are not ascending"); free(th); free(Th); or_plan_destroy(p); return NULL; }
    if (spin) p->seng = or_spin_rings_create(th, nh, L, spin, tol, ntrans, flip);
    else p->eng = or_rings_create(th, nh, L, tol, ntrans, flip);
    if (!p->eng && !p->seng) { free(th); free(Th); or_plan_destroy(p); return NULL; }
    p->mmax = spin ? p->seng->mmax : p->eng->mmax;
    int mphi = p->mmax > abs(spin) ? p->mmax : abs(spin);
    p->nph = (int)or_next_smooth_even(2 * (int64_t)mphi + 1);
    p->wth = malloc(sizeof(double) * nh);
    for (int i = 0; i < nh; i++) p->wth[i] = or_stretch_window(&p->geom, Th[i]);
    free(th); free(Th);
    p->Tpts = malloc(sizeof(double) * M);
    p->ppts = malloc(sizeof(double) * M);
    for (int64_t i = 0; i < M; i++) {
        double t = theta[i];
        p->Tpts[i] = or_theta_to_stretch(&p->geom, flip ? PI - t : t);
        p->ppts[i] = jmod(phi[i], 2 * PI);
    }
    int nth = p->nth, nph = p->nph, B = ntrans;
    if (nufft_kind != 1) p->es_order = es_point_order(nth, nph, tol, p->Tpts, p->ppts, M);
    p->G = calloc((size_t)nh * (2 * L + 1) * B, sizeof(or_cplx));
    p->F = calloc((size_t)nh * nph * B, sizeof(or_cplx));
    p->Ft = calloc((size_t)nth * nph * B, sizeof(or_cplx));
    p->Fhat = calloc((size_t)nth * nph * B, sizeof(or_cplx));
    p->fbuf = calloc((size_t)M * B, sizeof(or_cplx));
    double shift = p->geom.kind == OR_KIND_INTERIOR ? 0.0 : PI / (double)nth;
    double N = (double)nth * (double)nph;
    p->phase = malloc(sizeof(or_cplx) * nth);
    for (int i = 0; i < nth; i++) p->phase[i] = cscale(ccis(-(double)mode_k(i, nth) * shift), 1.0 / N);
    return p;
}

void or_plan_get_info(const or_plan *p, or_plan_info *in) {
    memset(in, 0, sizeof *in);
    in->kind = p->method == 2 ? p->geom.kind : -1;
    in->flip = p->flip; in->lmax = p->L; in->ntrans = p->B; in->method = p->method;
    in->nth_half = p->nth_half; in->nth = p->nth; in->nph = p->nph; in->mmax = p->mmax;
    in->ring_work = p->method == 2 ? (p->spin ? or_spin_rings_work(p->seng) : or_rings_work(p->eng)) : 0;
    in->npts = p->M; in->nufft_kind = p->nufft_kind;
    if (p->method == 2) {
        or_es_params(p->tol, &in->es_w, &in->es_beta);
        in->es_n1 = es_fine_size(p->nth, in->es_w);
        in->es_n2 = es_fine_size(p->nph, in->es_w);
    }
}

/* Stretch.jl:341-350: G columns (0-based) L+m -> F columns m mod nph */
static void pack_orders(or_cplx *F, const or_cplx *G, int nh, int nph, int L, int B) {
    int mkeep = (nph - 1) / 2 < L ? (nph - 1) / 2 : L;
    for (int b = 0; b < B; b++) {
        or_cplx *Fb = F + (int64_t)b * nh * nph;
        const or_cplx *Gb = G + (int64_t)b * nh * (2 * L + 1);
        for (int k = mkeep + 1; k <= nph - mkeep - 1; k++) memset(Fb + (int64_t)k * nh, 0, sizeof(or_cplx) * nh);
        for (int m = 0; m <= mkeep; m++) memcpy(Fb + (int64_t)m * nh, Gb + (int64_t)(L + m) * nh, sizeof(or_cplx) * nh);
        for (int m = 1; m <= mkeep; m++)
            memcpy(Fb + (int64_t)(nph - m) * nh, Gb + (int64_t)(L - m) * nh, sizeof(or_cplx) * nh);
    }
}

/* Stretch.jl:352-364 */
static void unpack_orders(or_cplx *G, const or_cplx *F, int nh, int nph, int L, int B) {
    int mkeep = (nph - 1) / 2 < L ? (nph - 1) / 2 : L;
    for (int b = 0; b < B; b++) {
        const or_cplx *Fb = F + (int64_t)b * nh * nph;
        or_cplx *Gb = G + (int64_t)b * nh * (2 * L + 1);
        for (int c = 0; c < L - mkeep; c++) memset(Gb + (int64_t)c * nh, 0, sizeof(or_cplx) * nh);
        for (int c = L + mkeep + 1; c <= 2 * L; c++) memset(Gb + (int64_t)c * nh, 0, sizeof(or_cplx) * nh);
        for (int m = 0; m <= mkeep; m++) memcpy(Gb + (int64_t)(L + m) * nh, Fb + (int64_t)m * nh, sizeof(or_cplx) * nh);
        for (int m = 1; m <= mkeep; m++)
            memcpy(Gb + (int64_t)(L - m) * nh, Fb + (int64_t)(nph - m) * nh, sizeof(or_cplx) * nh);
    }
}

static int run_nufft(or_plan *p, int type, int iflag, const or_cplx *in, or_cplx *out) {
    if (p->nufft_kind == 1)
        return or_nufft_direct(type, p->nth, p->nph, iflag, p->B, p->Tpts, p->ppts, p->M, in, out);
    return nufft_es_ordered(type, p->nth, p->nph, iflag, p->B, p->Tpts, p->ppts, p->M, p->tol, in, out, p->nthreads,
                            p->es_order);
}

/* Stretch.jl:262-277 / Direct.jl:60-91 */
int or_synth(or_plan *p, const or_cplx *alm, or_cplx *f) {
    int L = p->L, B = p->B;
    int64_t M = p->M;
    if (p->method == 1) {
        double t0 = now_s();
        if (p->spin) or_spin_rings_from_alm(p->seng, alm, p->G, p->nthreads);
        else or_rings_from_alm(p->eng, alm, p->G, p->nthreads);
        int64_t n = M, gB = n * (2 * L + 1);
        int off = L;
        for (int b = 0; b < B; b++)
            for (int64_t i = 0; i < n; i++) {
                const or_cplx *Gb = p->G + b * gB;
                or_cplx acc = Gb[i + (int64_t)off * n];
                double phi = p->phs[i];
                for (int m = 1; m <= L; m++) {
                    or_cplx ph = ccis((double)m * phi);
                    acc = cadd(acc, cadd(cmul(Gb[i + (int64_t)(m + off) * n], ph), cmul(Gb[i + (int64_t)(-m + off) * n], cconj(ph))));
                }
                f[p->perm[i] + b * M] = acc;
            }
        p->times[0] = now_s() - t0;
        return 0;
    }
    int nh = p->nth_half, nth = p->nth, nph = p->nph;
    double t0 = now_s();
    if (p->spin) or_spin_rings_from_alm(p->seng, alm, p->G, p->nthreads);
    else or_rings_from_alm(p->eng, alm, p->G, p->nthreads);
    double t1 = now_s();
    pack_orders(p->F, p->G, nh, nph, L, B);
    fft_dim2(p->F, nh, nph, B, +1, p->nthreads);
    for (int b = 0; b < B; b++)
        for (int k = 0; k < nph; k++)
            for (int j = 0; j < nh; j++) {
                int64_t i = j + (int64_t)k * nh + (int64_t)b * nh * nph;
                p->F[i] = cscale(p->F[i], p->wth[j]);
            }
    if (p->geom.kind == OR_KIND_INTERIOR) memcpy(p->Ft, p->F, sizeof(or_cplx) * (size_t)nth * nph * B);
    else or_dfs_double(p->F, p->Ft, nh, nph, B, p->spin);
    memcpy(p->Fhat, p->Ft, sizeof(or_cplx) * (size_t)nth * nph * B);
    or_fft2(p->Fhat, nth, nph, B, -1, p->nthreads);
    for (int b = 0; b < B; b++)
        for (int k = 0; k < nph; k++)
            for (int i = 0; i < nth; i++) {
                int64_t q = i + (int64_t)k * nth + (int64_t)b * nth * nph;
                p->Fhat[q] = cmul(p->Fhat[q], p->phase[i]);
            }
    double t2 = now_s();
    int rc = run_nufft(p, 2, +1, p->Fhat, p->fbuf);
    memcpy(f, p->fbuf, sizeof(or_cplx) * (size_t)M * B);
    double t3 = now_s();
    p->times[0] = t1 - t0; p->times[1] = t2 - t1; p->times[2] = t3 - t2;
    return rc;
}

/* Stretch.jl:279-295 / Direct.jl:94-117 */
int or_adjoint(or_plan *p, const or_cplx *f, or_cplx *alm) {
    int L = p->L, B = p->B;
    int64_t M = p->M;
    if (p->method == 1) {
        double t0 = now_s();
        int64_t n = M, gB = n * (2 * L + 1);
        int off = L;
        memset(p->G, 0, sizeof(or_cplx) * (size_t)gB * B);
        for (int b = 0; b < B; b++)
            for (int64_t i = 0; i < n; i++) {
                or_cplx fi = f[p->perm[i] + b * M];
                if (fi.re == 0 && fi.im == 0) continue;
                or_cplx *Gb = p->G + b * gB;
                double phi = p->phs[i];
                Gb[i + (int64_t)off * n] = fi;
                for (int m = 1; m <= L; m++) {
                    or_cplx ph = ccis((double)m * phi);
                    Gb[i + (int64_t)(m + off) * n] = cmul(fi, cconj(ph));
                    Gb[i + (int64_t)(-m + off) * n] = cmul(fi, ph);
                }
            }
        if (p->spin) or_spin_alm_from_rings(p->seng, p->G, alm, p->nthreads);
        else or_alm_from_rings(p->eng, p->G, alm, p->nthreads);
        p->times[0] = now_s() - t0;
        return 0;
    }
    int nh = p->nth_half, nth = p->nth, nph = p->nph;
    double t0 = now_s();
    memcpy(p->fbuf, f, sizeof(or_cplx) * (size_t)M * B);
    int rc = run_nufft(p, 1, -1, p->fbuf, p->Fhat);
    if (rc) return rc;
    double t1 = now_s();
    for (int b = 0; b < B; b++)
        for (int k = 0; k < nph; k++)
            for (int i = 0; i < nth; i++) {
                int64_t q = i + (int64_t)k * nth + (int64_t)b * nth * nph;
                p->Fhat[q] = cmul(p->Fhat[q], cconj(p->phase[i]));
            }
    memcpy(p->Ft, p->Fhat, sizeof(or_cplx) * (size_t)nth * nph * B);
    or_fft2(p->Ft, nth, nph, B, +1, p->nthreads);
    if (p->geom.kind == OR_KIND_INTERIOR) memcpy(p->F, p->Ft, sizeof(or_cplx) * (size_t)nth * nph * B);
    else or_dfs_fold(p->Ft, p->F, nh, nph, B, p->spin);
    for (int b = 0; b < B; b++)
        for (int k = 0; k < nph; k++)
            for (int j = 0; j < nh; j++) {
                int64_t i = j + (int64_t)k * nh + (int64_t)b * nh * nph;
                p->F[i] = cscale(p->F[i], p->wth[j]);
            }
    fft_dim2(p->F, nh, nph, B, -1, p->nthreads);
    unpack_orders(p->G, p->F, nh, nph, L, B);
    double t2 = now_s();
    if (p->spin) or_spin_alm_from_rings(p->seng, p->G, alm, p->nthreads);
    else or_alm_from_rings(p->eng, p->G, alm, p->nthreads);
    double t3 = now_s();
    p->times[0] = t3 - t2; p->times[1] = t2 - t1; p->times[2] = t1 - t0;
    return 0;
}

int or_plan_last_G(const or_plan *p, or_cplx *G) {
    int64_t n = p->method == 2 ? p->nth_half : p->M;
    memcpy(G, p->G, sizeof(or_cplx) * (size_t)n * (2 * p->L + 1) * p->B);
    return 0;
}
int or_plan_last_Fhat(const or_plan *p, or_cplx *Fhat) {
    if (p->method != 2) return -1;
    memcpy(Fhat, p->Fhat, sizeof(or_cplx) * (size_t)p->nth * p->nph * p->B);
    return 0;
}
void or_plan_last_times(const or_plan *p, double *t3) { memcpy(t3, p->times, sizeof p->times); }

int or_plan_points(const or_plan *p, double *Theta, double *phi) {
    if (p->method != 2) return -1;
    memcpy(Theta, p->Tpts, sizeof(double) * (size_t)p->M);
    memcpy(phi, p->ppts, sizeof(double) * (size_t)p->M);
    return 0;
}
