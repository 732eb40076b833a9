/*
 * pfsht.h — C ABI of libpfsht_hip.so, the MI355X (gfx950) implementation of PartialFSHT.jl's
 * stretched-leaf apply path: synthesis mul!(f, plan, alm) and its exact adjoint
 * mul!(alm, plan', f) on a Plan(theta, phi; lmax, tol) in sph_mode layout.
 *
 * Plain C types only. Complex arrays are interleaved (re, im) doubles, which is the memory of
 * Julia's ComplexF64 and of C99 double _Complex. Layouts are the reference's column-major ones:
 *   alm : (lmax+1) x (2 lmax+1) x ntrans, entry (l, m) at row l-|m|, column 2|m| + (m >= 0) - 1
 *         (0-based), i.e. FastSphericalHarmonics sph_mode (src/Layout.jl:13,20);
 *   f   : npts x ntrans (src/Layout.jl:47-56);
 *   G   : nrings x (2 lmax+1) x ntrans, column m + lmax (src/Legendre.jl:306).
 * Every call returns PFSHT_OK or an error code, with a message in pfsht_last_error() (per thread).
 * Nothing aborts. A plan owns all device memory it uses; applies allocate nothing.
 */
#ifndef PFSHT_H
#define PFSHT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes: the Julia extension maps E_DIM to DimensionMismatch, E_ARG / E_GEOM_REFUSED to
 * ArgumentError (src/Layout.jl:35-56; src/Plan.jl:98-102; src/Windows.jl:66-74). */
enum {
    PFSHT_OK = 0,
    PFSHT_E_DIM = 1,
    PFSHT_E_ARG = 2,
    PFSHT_E_GEOM_REFUSED = 3,
    PFSHT_E_NOMEM = 4,
    PFSHT_E_HIP = 5,
    PFSHT_E_UNSUPPORTED = 6
};

/* Leaf selection, the AbstractAlgorithm tags of src/Cost.jl:15-24. AUTO follows select_method
 * (src/Cost.jl:174-201): the cheaper of Direct and one stretched box, or several boxes when
 * partition_boxes splits the points and the split is cheaper. MULTIBOX is MultiBoxLeaf
 * (src/MultiBox.jl:20-103): one stretched plan per box, a disjoint union in synthesis and a sum
 * in the adjoint. */
enum { PFSHT_METHOD_AUTO = 0, PFSHT_METHOD_DIRECT = 1, PFSHT_METHOD_STRETCH = 2, PFSHT_METHOD_MULTIBOX = 3 };

/* Where apply buffers live: host memory (copied through the plan's staging) or HBM. */
enum { PFSHT_WHERE_HOST = 0, PFSHT_WHERE_DEVICE = 1 };

/* Arithmetic: FP64 throughout, or FP64 recurrence with FP32 accumulation/FFT/NUFFT (tol >= 1e-6). */
enum { PFSHT_PREC_FP64 = 0, PFSHT_PREC_MIXED = 1 };

/* The stretched apply's grid stage (the FFTs between the ring pass and the NUFFT, src/Stretch.jl:
 * 262-295): LDS kernels that run the colatitude, fine-column and fine-row FFTs inside workgroups,
 * fused with the window, periodisation, deconvolution and transposes around them, or rocFFT plans
 * with separate kernels between them. AUTO takes the LDS kernels wherever the fine lengths fit a
 * workgroup's LDS (up to 10240 FP64 / 16384 FP32 values), LDS insists on them, LIBRARY takes
 * rocFFT. No counterpart in the reference, whose FFTs are FFTW plans. */
enum { PFSHT_GRID_FFT_AUTO = 0, PFSHT_GRID_FFT_LDS = 1, PFSHT_GRID_FFT_LIBRARY = 2 };

/* Periodisation kinds (src/Windows.jl:34-36); -1 for a Direct leaf, -2 for a multi-box plan. */
enum { PFSHT_KIND_POLE = 0, PFSHT_KIND_INTERIOR = 1, PFSHT_KIND_FULL = 2 };

typedef struct pfsht_plan_s *pfsht_plan_t;
typedef struct pfsht_rings_s *pfsht_rings_t;

typedef struct {
    int32_t method;    /* PFSHT_METHOD_DIRECT or PFSHT_METHOD_STRETCH (as chosen) */
    int32_t kind;      /* PFSHT_KIND_* */
    int32_t flip;      /* south-pole periodisation reflected north (src/Stretch.jl:151-158) */
    int32_t lmax, ntrans, precision;
    int32_t nth_half;  /* rings the ring engine evaluates */
    int32_t nth, nph;  /* stretched grid */
    int32_t mmax;      /* largest retained order */
    int32_t es_w, es_n1, es_n2; /* NUFFT kernel width and fine grid */
    int64_t npts;
    int64_t ring_work; /* retained (l, m, ring) triples per apply (src/Legendre.jl:286-293) */
    int64_t device_bytes;
    double tol, es_beta;
    int32_t es_rows;   /* fine-grid rows the apply stores: those the points' kernels reach */
    int32_t nboxes;    /* boxes of a multi-box plan (1 otherwise) */
    int32_t spin;      /* spin weight (0: scalar) */
    int32_t seg_blocks, adj_tiles, mfma_tiles; /* the ring passes' launch shape (pfsht_ring_shape) */
    int32_t grid_fft;  /* the grid stage taken: PFSHT_GRID_FFT_LDS or _LIBRARY (0 for a Direct leaf; a
                          multi-box plan reports LIBRARY when any box takes it) */
} pfsht_info;

/* Number of timed stages reported by pfsht_stage_times. */
#define PFSHT_NSTAGES 9
/* Stage order: 0 ring pass (adjoint: per-tile partial sums), 1 pack/unpack+window, 2 phi FFT,
 * 3 DFS double/fold, 4 2-D FFT, 5 deconvolve/phase, 6 fine FFT, 7 NUFFT interpolate/spread(+merge),
 * 8 adjoint ring partial reduction (empty in synthesis). */

const char *pfsht_last_error(void);
const char *pfsht_version(void);

/* Plan(theta, phi; lmax, tol, ntrans, method) — replaces the constructor of src/Plan.jl:84-116
 * with the StretchLeaf of src/Stretch.jl:129-248 (or DirectLeaf, src/Direct.jl:25-51) built on
 * the device. theta, phi: host arrays of npts colatitudes in [0, pi] and longitudes. */
int pfsht_plan_create(const double *theta, const double *phi, int64_t npts, int32_t lmax, double tol,
                      int32_t ntrans, int32_t method, int32_t precision, pfsht_plan_t *out);
/* Plan(theta, phi; lmax, tol, spin, ntrans, method) — src/Plan.jl:84-116 with spin != 0:
 * SpinDirectLeaf (src/Spin.jl:548-585) or StretchLeaf over SpinRings (src/Stretch.jl:190-194),
 * per box for a multi-box plan (src/MultiBox.jl:33-50). |spin| <= lmax (src/Plan.jl:101-102).
 * Coefficients of a spin plan use the dense (l, m + lmax) layout of spin_coeff_index
 * (src/Layout.jl:28-29); the analysis calls that read the sph_mode layout (pseudo-spectrum, the
 * Wiener prior) return PFSHT_E_UNSUPPORTED on a spin plan. spin = 0 is pfsht_plan_create. */
int pfsht_plan_create_spin(const double *theta, const double *phi, int64_t npts, int32_t lmax, double tol,
                           int32_t ntrans, int32_t spin, int32_t method, int32_t precision, pfsht_plan_t *out);
/* The remaining keywords of Plan(θ, φ; lmax, tol, spin, ntrans, method, maxboxes, oversample_θ,
 * oversample_φ, εapo_max) — src/Plan.jl:84-97, forwarded into the leaves as src/Plan.jl:110-135 does:
 * oversample_θ inflates the stretched ring count (src/Stretch.jl:164), oversample_φ the longitude
 * grid (src/Stretch.jl:194), εapo_max bounds the apodisation fraction a stretched geometry may use
 * (src/Windows.jl:84-141), maxboxes bounds the planner's partition (src/Box.jl:171-235). A zero field
 * takes the reference's default. `packed` has no counterpart in the reference: it makes the plan
 * keep FFT plans for every power-of-two transform count, so that a column apply of nb < ntrans
 * transforms (pfsht_synth_cols) does only nb transforms' work (servers packing requests set it). */
typedef struct {
    int32_t method;          /* PFSHT_METHOD_* (AUTO when 0) */
    int32_t precision;       /* PFSHT_PREC_* */
    int32_t spin;            /* spin weight; 0 scalar */
    int32_t maxboxes;        /* 0: 8 */
    double oversample_theta; /* 0: 1 */
    double oversample_phi;   /* 0: 1 */
    double epsapo_max;       /* 0: 1.0 */
    int32_t packed;          /* nonzero: FFT plans for partial transform counts */
    int32_t tune;            /* ring-pass shape: 0 time candidates when the shape is not cached,
                                -1 never time (cached shape or the default), 1 always time */
    int32_t grid_fft;        /* PFSHT_GRID_FFT_* (AUTO when 0) */
} pfsht_plan_opts;
/* Fills *opts with the defaults (all fields zero). `tune` has no counterpart in the reference, whose
 * ring pass has no launch shape: see pfsht_ring_shape; nor has `grid_fft` (PFSHT_GRID_FFT_*). */
void pfsht_plan_opts_default(pfsht_plan_opts *opts);
/* Plan(θ, φ; lmax, tol, ntrans, keywords...) with every keyword of src/Plan.jl:84-97 (opts may be
 * NULL for the defaults). oversample_θ, oversample_φ >= 1 and εapo_max in (0, 1] (src/Windows.jl). */
int pfsht_plan_create_opts(const double *theta, const double *phi, int64_t npts, int32_t lmax, double tol,
                           int32_t ntrans, const pfsht_plan_opts *opts, pfsht_plan_t *out);
/* The device bytes pfsht_plan_create_opts would allocate for these arguments, computed on the host
 * from the same sizing, without building anything: an upper estimate (it counts the NUFFT windows of
 * super-tiles no point reaches), for callers admitting plans against a memory budget. No counterpart
 * in the reference. */
int pfsht_plan_estimate_bytes(const double *theta, const double *phi, int64_t npts, int32_t lmax, double tol,
                              int32_t ntrans, const pfsht_plan_opts *opts, int64_t *bytes);
void pfsht_plan_destroy(pfsht_plan_t plan);
int pfsht_plan_info(pfsht_plan_t plan, pfsht_info *info);

/* The launch shape of a plan's ring passes (no counterpart in the reference, whose ring pass is a
 * threaded loop over orders, src/Legendre.jl:284-341). A stretched scalar plan times candidate shapes
 * on its own ring engine when it is built (a few milliseconds; opts->tune) and keeps the fastest;
 * the choice is cached per (device architecture, lmax, rings, largest order, ntrans), so later plans
 * of the shape reuse it. PFSHT_TUNE_CACHE names a file the cache is read from once per process and
 * written back to after each new timing. */
typedef struct {
    int32_t seg_blocks; /* target of 64-degree blocks per degree segment (1..64; 0: default 3) */
    int32_t adj_tiles;  /* 64-ring tiles per wavefront of the single-transform adjoint (1, 2; 0: default) */
    int32_t mfma_tiles; /* 64-ring tiles per wavefront of the FP64 MFMA adjoint (1, 2; 0: default) */
    int32_t reserved;
    double ms;          /* ring forward + adjoint time of the shape when it was timed */
} pfsht_ring_shape;
/* The cache, host only (no device work): the shape stored under a key (*found = 0 when none), a new
 * entry, merging and writing cache files, clearing, and the number of entries. arch is the device's
 * architecture and compute-unit count as plans record it, e.g. "gfx950:256". */
int pfsht_tune_get(const char *arch, int32_t lmax, int32_t nrings, int32_t mmax, int32_t ntrans,
                   pfsht_ring_shape *shape, int32_t *found);
int pfsht_tune_put(const char *arch, int32_t lmax, int32_t nrings, int32_t mmax, int32_t ntrans,
                   const pfsht_ring_shape *shape);
int pfsht_tune_load(const char *path);
int pfsht_tune_save(const char *path);
void pfsht_tune_clear(void);
int64_t pfsht_tune_size(void);
/* The candidates a plan timed when it was built (*n = 0 when its shape came from the cache or the
 * default), at most cap of them; and the key its ring engine is cached under (arch: cap bytes). */
int pfsht_plan_tune_table(pfsht_plan_t plan, pfsht_ring_shape *rows, int32_t cap, int32_t *n);
int pfsht_plan_tune_key(pfsht_plan_t plan, char *arch, int32_t cap, int32_t *lmax, int32_t *nrings, int32_t *mmax,
                        int32_t *ntrans);

/* mul!(f, plan, alm) — src/Plan.jl:137-141 -> synthesize_stretch! (src/Stretch.jl:262-277).
 * Overwrites f. stream: a hipStream_t or NULL for the null stream; the call is asynchronous
 * on that stream when where == PFSHT_WHERE_DEVICE. */
int pfsht_synth(pfsht_plan_t plan, const void *alm, void *f, int32_t where, void *stream);

/* mul!(alm, plan', f) — src/Plan.jl:143-148 -> adjoint_stretch! (src/Stretch.jl:279-295).
 * Overwrites alm, padding slots included (src/Legendre.jl:338). */
int pfsht_adjoint(pfsht_plan_t plan, const void *f, void *alm, int32_t where, void *stream);

/* mul!(f, plan, alm) and mul!(alm, plan', f) on a Float32 plan — src/Plan.jl:84-116 with
 * T = Float32, as gpu/runtests.jl:122-139 and test/test_transforms.jl:158-190 run it: alm and f hold
 * ComplexF32 (interleaved float pairs) in the same layouts and counts as the FP64 calls, and the
 * result is ComplexF32. The plan computes as it was created (a PFSHT_PREC_MIXED plan, tol >= 1e-6,
 * keeps the recurrence FP64 and the grids FP32; an FP64 plan is FP64 throughout); the input is widened
 * into the plan's FP64 staging and the result narrowed on the way out. */
int pfsht_synth_f32(pfsht_plan_t plan, const void *alm, void *f, int32_t where, void *stream);
int pfsht_adjoint_f32(pfsht_plan_t plan, const void *f, void *alm, int32_t where, void *stream);

/* mul!(f, plan, alm) and mul!(alm, plan', f) — src/Plan.jl:137-148 — over nb <= ntrans transforms whose
 * columns lie anywhere in device memory: alm_cols[b] is the (lmax+1) x (2 lmax+1) coefficient array of
 * transform b (sph_mode layout, or the spin layout on a spin plan) and f_cols[b] its npts field values.
 * The pointer arrays are host memory, read before the call returns; the call is asynchronous on the
 * stream like the device-memory applies. Transform b of the result is the apply of column b alone
 * (to round-off), whatever the other columns hold. No counterpart in the reference, whose batches are
 * one contiguous array (src/Layout.jl:37-56): it is for callers that pack independent requests into
 * one apply without copying them together (the serving runtime does). */
int pfsht_synth_cols(pfsht_plan_t plan, int32_t nb, const void *const *alm_cols, void *const *f_cols, void *stream);
int pfsht_adjoint_cols(pfsht_plan_t plan, int32_t nb, const void *const *f_cols, void *const *alm_cols,
                       void *stream);

/* Plan groups: `mul!` over many independent single-transform plans at once — the farmed apply of
 * independent plans in the reference (ext/PartialFSHTOhMyThreadsExt.jl:151-159 over
 * src/Plan.jl:201-219), for many small plans that one at a time leave most of the GPU idle. The
 * ring passes of all members run as one launch over a (plan, order, 64-ring tile) work list, and
 * so do the interpolation and spreading (over (plan, bin) lists) and the LDS grid stage (over
 * (plan, column) and (plan, row) lists); Direct members and grid stages past a workgroup's LDS run
 * concurrently on the group's streams. Each direction is replayed from a hipGraph on the caller's
 * stream after its first call. Members: scalar stretched or Direct plans with
 * ntrans = 1 on one device (others: PFSHT_E_UNSUPPORTED); a plan belongs to at most one group in
 * use at a time, and must not be applied on its own while a group apply is in flight.
 * alm[i] and f[i] are member i's device buffers, in its layouts; the results equal the members'
 * own applies. The pointer arrays are host memory read before the call returns; the call is
 * asynchronous on stream. */
typedef struct pfsht_group_s *pfsht_group_t;
int pfsht_group_create(const pfsht_plan_t *plans, int32_t nplans, pfsht_group_t *out);
void pfsht_group_destroy(pfsht_group_t group);
int pfsht_group_synth(pfsht_group_t group, const void *const *alm, void *const *f, void *stream);
int pfsht_group_adjoint(pfsht_group_t group, const void *const *f, void *const *alm, void *stream);

/* RingLegendre(theta, lmax, eps; ntrans, reflect) — src/Legendre.jl:194-255. theta sorted
 * ascending, host memory. */
int pfsht_rings_create(const double *theta, int32_t nrings, int32_t lmax, double eps, int32_t ntrans,
                       int32_t reflect, pfsht_rings_t *out);
/* SpinRings(theta, lmax, spin, eps; ntrans, reflect) — src/Spin.jl:192-270: the same handle, whose
 * rings_from_alm / alm_from_rings are spin_rings_from_alm! / spin_alm_from_rings!
 * (src/Spin.jl:316-334) with coefficients in the (l, m + lmax) layout. spin != 0, |spin| <= lmax. */
int pfsht_spin_rings_create(const double *theta, int32_t nrings, int32_t lmax, int32_t spin, double eps,
                            int32_t ntrans, int32_t reflect, pfsht_rings_t *out);
void pfsht_rings_destroy(pfsht_rings_t rings);
int64_t pfsht_rings_work(pfsht_rings_t rings);
int32_t pfsht_rings_mmax(pfsht_rings_t rings);
/* rings_from_alm!(G, alm, eng) — src/Legendre.jl:310-323; G overwritten (zero off the retained set). */
int pfsht_rings_from_alm(pfsht_rings_t rings, const void *alm, void *G, int32_t where, void *stream);
/* alm_from_rings!(alm, G, eng) — src/Legendre.jl:332-341; alm overwritten. */
int pfsht_alm_from_rings(pfsht_rings_t rings, const void *G, void *alm, int32_t where, void *stream);

/* alm_from_rings!(alm, G, eng) as pfsht_alm_from_rings on device buffers, allocating nothing: the
 * internal ring buffer and the ring pass's partial sums are taken from a file-scope device array of
 * 64 MiB that each GPU holds for the life of the process, instead of the engine's own buffers. No
 * counterpart in the reference: for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops). Engines needing more than 64 MiB of such scratch are refused (PFSHT_E_UNSUPPORTED).
 * Calls that overlap on one GPU share the array and race: order them (one stream, or events). */
int pfsht_alm_from_rings_noalloc(pfsht_rings_t rings, const void *G, void *alm, void *stream);

/* The plan's NUFFT on its stretched points — the seam _nufft_exec!(plan, in, out) of
 * src/NUFFT.jl:17-25 with the two plans a StretchLeaf builds (src/Stretch.jl:235-238):
 *   type 2 (iflag +1): in = modes nth x nph x ntrans (column-major, FFT order), out = f (npts x ntrans);
 *   type 1 (iflag -1): in = f, out = modes. Outputs are overwritten. */
int pfsht_plan_nufft(pfsht_plan_t plan, int32_t type, const void *in, void *out, int32_t where, void *stream);

/* The plan's own ring engine, for stage-level parity of a stretched plan. */
int pfsht_plan_rings(pfsht_plan_t plan, pfsht_rings_t *out);

/* Per-stage device time of the plan's applies, recorded with HIP events on the apply stream while
 * timing is enabled. ms[PFSHT_NSTAGES] are summed milliseconds, counts[PFSHT_NSTAGES] launches
 * timed; reset clears both. Synthesis and adjoint are reported separately (dir 0 / 1). */
int pfsht_set_timing(pfsht_plan_t plan, int32_t enable);
int pfsht_stage_times(pfsht_plan_t plan, int32_t dir, double *ms, int64_t *counts, int32_t reset);

/* ---------------------------------------------------------------------------------------------
 * Bounding box and analysis on a plan (src/Box.jl, src/Analysis.jl). The iterative callers keep
 * every vector in HBM and reuse a workspace the plan allocates at the first call (AnalysisWorkspace,
 * src/Analysis.jl:45-84): rhs, r, p, Ap in the alm layout and two fields. Dot products are
 * deterministic (fixed-order reductions), so results are bit-identical run to run. For
 * where == PFSHT_WHERE_HOST every array argument is host memory; for PFSHT_WHERE_DEVICE every
 * one is device memory except prior, which is always host memory. The calls return when their
 * outputs are ready.
 * ------------------------------------------------------------------------------------------- */
typedef struct {
    double thlo, thhi;       /* colatitude range, clamped to [0, pi] */
    double phistart, phiext; /* smallest longitude arc holding the points; 2 pi when touching a pole */
    int32_t touches_north, touches_south;
    double fsky;             /* box_fsky: (cos thlo - cos thhi) phiext / 4 pi */
    double dl;               /* spectral_resolution: pi / max(theta extent, phiext sin(mid theta)) */
} pfsht_box;

/* bounding_box(theta, phi) — src/Box.jl:43-73 (with _circular_extent, :80-102), box_fsky :106-108,
 * spectral_resolution :148-153. Host arrays; no device work. */
int pfsht_bounding_box(const double *theta, const double *phi, int64_t npts, pfsht_box *out);

/* partition_boxes(θ, φ, costfn; maxboxes) with costfn = cost_stretch(box, n, lmax, ntrans, tol)
 * — src/Box.jl:159-235, src/Cost.jl:193. Host arrays; group_of[i] receives the box of point i
 * (boxes numbered in the reference's order of creation), *nboxes their number. No device work. */
int pfsht_partition_boxes(const double *theta, const double *phi, int64_t npts, int32_t lmax, int32_t ntrans,
                          double tol, int32_t maxboxes, int32_t *group_of, int32_t *nboxes);

/* apply_resolution!(out, plan, alm, w, ws) — src/Analysis.jl:96-104: out = A' (w .* (A alm)).
 * w: npts real weights (NULL: all ones). out overwritten, padding included. */
int pfsht_apply_resolution(pfsht_plan_t plan, const void *alm, const double *w, void *out, int32_t where,
                           void *stream);

/* pseudo_spectrum!(C, plan, f, w, ws; batch) — src/Analysis.jl:286-308:
 * C[l] = sum_m |a~_lm|^2 / (2l + 1) with a~ = A' (w .* f) of transform `batch` (1-based). C has
 * lmax + 1 entries. */
int pfsht_pseudo_spectrum(pfsht_plan_t plan, const void *f, const double *w, int32_t batch, double *C,
                          int32_t where, void *stream);

/* solve!(alm, plan, f; mode, rtol, maxiter, prior, w) — src/Analysis.jl:375-436. Conjugate
 * gradients on the normal equations A' W A alm = A' W f from alm = 0, stopped when
 * |r| / |A' W f| <= rtol (min(rtol, 1e-12) for MINNORM); WIENER adds diag(1 / prior_l) (prior_l <= 0
 * adds nothing) and needs prior (lmax + 1 host doubles). */
enum { PFSHT_SOLVE_CG = 0, PFSHT_SOLVE_WIENER = 1, PFSHT_SOLVE_MINNORM = 2 };
typedef struct {
    double shannon;  /* (lmax + 1)^2 fsky of the points' bounding box */
    double dl;       /* the box's spectral resolution */
    double relres;   /* final |r| / |A' W f| */
    int32_t iters, mode;
} pfsht_solve_info;
int pfsht_solve(pfsht_plan_t plan, const void *f, const double *w, int32_t mode, double rtol, int32_t maxiter,
                const double *prior, void *alm, int32_t where, void *stream, pfsht_solve_info *info);

/* ---------------------------------------------------------------------------------------------
 * Serving runtime. No counterpart in the reference, whose plans are built and applied by their
 * caller (src/Plan.jl:84-148); it is for services answering requests on recurring point sets.
 *
 * A server keeps plans resident in HBM under caller-chosen 64-bit keys. pfsht_server_plan registers a
 * geometry and the number of transforms a request of it carries (ntrans); the plan is built with
 * room for max(ntrans, pack) transforms, with FP32 grids when tol >= 1e-6 (PFSHT_PREC_MIXED, which
 * meets such tolerances), after its size has been estimated and least-recently-used plans with no
 * request queued or running have been evicted to fit the HBM budget (nothing is allocated for a plan
 * that cannot fit).
 *
 * pfsht_server_submit queues one request and returns at once. A dispatcher thread issues the queue
 * in order of arrival: whenever one of the nstreams streams has fewer than `depth` launches in
 * flight, it takes the oldest request and every later queued request of the same key and direction
 * that fits the plan's room, and applies them together as one column apply (pfsht_synth_cols), in
 * place in the callers' buffers; with graphs on, each (plan, direction, transform count) is captured
 * once in a hipGraph and replayed. Requests that arrive while others run join the next launch of
 * their plan. Each request's result is the same as its apply alone (to round-off). Applies of one
 * plan are ordered; streams run at the default priority. All calls are thread-safe.
 * ------------------------------------------------------------------------------------------- */
typedef struct pfsht_server_s *pfsht_server_t;

typedef struct {
    int32_t nstreams;         /* streams of the pool (1..64; default 4) */
    int32_t depth;            /* launches in flight per stream (1..16; default 2) */
    int32_t pack;             /* transforms a plan's launch can hold (1..64; default 8) */
    int32_t graphs;           /* nonzero: replay captured hipGraphs (default 1) */
    int64_t hbm_budget_bytes; /* resident plan bytes (default 200e9) */
} pfsht_server_opts;

typedef struct {
    int64_t hits, misses, evictions, rejected; /* plan lookups; plans refused for lack of HBM */
    int64_t submitted;                          /* requests queued */
    int64_t resident_bytes, budget_bytes, plans;
    int64_t launches;           /* packed applies issued */
    int64_t packed_transforms;  /* transforms those applies carried */
    int64_t graph_launches;     /* of which replayed from a hipGraph */
    int64_t graph_failures;     /* plans whose capture failed (applied eagerly instead) */
    int64_t failed;             /* requests whose launch returned an error */
    int64_t queued;             /* requests waiting for a launch now */
} pfsht_server_stats;

void pfsht_server_opts_default(pfsht_server_opts *opts);
int pfsht_server_create_ex(const pfsht_server_opts *opts, pfsht_server_t *out);
/* pfsht_server_create_ex with the defaults, nstreams and the budget. */
int pfsht_server_create(int32_t nstreams, int64_t hbm_budget_bytes, pfsht_server_t *out);
/* Waits for the server's work, stops its dispatcher and frees its plans. */
void pfsht_server_destroy(pfsht_server_t srv);
/* Registers the geometry under key (built with `method` on a miss; *hit, may be NULL, tells which) and
 * returns its plan. The plan stays owned by the server and is valid until a later pfsht_server_plan
 * call evicts it: do not destroy it. PFSHT_E_NOMEM when it cannot fit the budget. ntrans 1..64. */
int pfsht_server_plan(pfsht_server_t srv, uint64_t key, const double *theta, const double *phi, int64_t npts,
                      int32_t lmax, double tol, int32_t ntrans, int32_t method, pfsht_plan_t *plan, int32_t *hit);
/* Queues one request of the plan under key: dir 0 = mul!(out = f, plan, in = alm), dir 1 =
 * mul!(out = alm, plan', in = f), each with the key's ntrans transforms in the reference's contiguous
 * layouts, device memory. in and out must stay valid until the ticket completes. Returns a ticket. */
int pfsht_server_submit(pfsht_server_t srv, uint64_t key, int32_t dir, const void *in, void *out, int64_t *ticket);
/* *done = 1 once the request's apply has finished; *device_ms (may be NULL) is then its launch's time on
 * its stream (a packed launch's time is shared by its requests). */
int pfsht_server_query(pfsht_server_t srv, int64_t ticket, int32_t *done, double *device_ms);
int pfsht_server_wait(pfsht_server_t srv, int64_t ticket, double *device_ms);
int pfsht_server_get_stats(pfsht_server_t srv, pfsht_server_stats *stats);
/* Device clock: pfsht_server_epoch records a reference event and waits for it; pfsht_server_times
 * then gives a request's launch start and end, in ms after the epoch, waiting for it to finish. With
 * the host clock read right after the epoch, request latencies (arrival to end) follow without polling. */
int pfsht_server_epoch(pfsht_server_t srv);
int pfsht_server_times(pfsht_server_t srv, int64_t ticket, double *start_ms, double *end_ms);

#ifdef __cplusplus
}
#endif
#endif
