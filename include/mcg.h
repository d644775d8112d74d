/*
 * mcg.h -- C ABI of libmcg, the MI355X (gfx950) clustering hot path.
 *
 * Every entry point returns an int status (MCG_OK == 0, negative on error);
 * the message of the last error on the calling thread is in mcg_last_error().
 * Nothing throws across this boundary.
 *
 * Pointer arguments may be host or device memory; the library classifies
 * each one (hipPointerGetAttributes) and copies host buffers as needed.
 * The caller owns every buffer it passes in and keeps it alive for the call
 * (a dataset built from a device buffer borrows it for the dataset's life,
 * like the reference's CSparseMatrix borrows numpy buffers,
 * python/pycsparse.h:33-70).  The library owns its device buffers inside an
 * mcg_ctx (one HIP stream per context).  A context is not safe for use by
 * two threads at once.
 *
 * Layout conventions: dense matrices are row-major float32, n rows x d
 * columns (leading dimension ld >= d).  Centers are dense row-major float32,
 * k x d.  CSR matrices: values, column indices (sorted within each row),
 * row pointers of length n+1.
 *
 * Costs follow the reference's argument order msr(first = data row,
 * second = center) (include/minicore/dist/applicator.h:1632).
 */
#ifndef MCG_H
#define MCG_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define MCG_OK            0
#define MCG_EINVAL       -1   /* invalid argument (reference: std::invalid_argument) */
#define MCG_ENOMEM       -2   /* device allocation failed */
#define MCG_EDEVICE      -3   /* HIP runtime / device error */
#define MCG_EUNSUPPORTED -4   /* measure / dtype / shape without a GPU kernel */
#define MCG_ENAN         -5   /* NaN center (centroid.h:771-775 raises runtime_error) */
#define MCG_EIO          -6   /* file open / read / write failed (reference: std::runtime_error) */

/* element types */
#define MCG_F32 0
#define MCG_F64 1
#define MCG_U8  2
#define MCG_U16 3
#define MCG_U32 4
#define MCG_U64 5
#define MCG_I32 6
#define MCG_I64 7

/* measure codes: include/minicore/dist/distance.h:24-64 */
#define MCG_SQRL2        2
#define MCG_JSD          4
#define MCG_MKL          5
#define MCG_REVERSE_MKL 11
/* further measures (applicator.h:1673-1760 dense, 1778-1891 sparse), evaluated pair by
 * pair.  Hard clustering uses the weighted-mean centroid for those mapped to a mean
 * policy by msr2pol (centroid.h:182-207); L1, TVD and L2 cluster with their median
 * policies (mcg_lloyd); the similarity / probability-cosine forms (no policy) are
 * cost-only: mcg_cost_matrix, mcg_assign_hard, mcg_kmeanspp. */
#define MCG_L1                            0
#define MCG_L2                            1
#define MCG_JSM                           3
#define MCG_HELLINGER                     6
#define MCG_BHATTACHARYYA_METRIC          7
#define MCG_BHATTACHARYYA_DISTANCE        8
#define MCG_TVD                           9
#define MCG_LLR                          10
#define MCG_UWLLR                        12
#define MCG_IS                           13
#define MCG_REVERSE_IS                   14
#define MCG_COSINE_DISTANCE              15
#define MCG_PROBABILITY_COSINE_DISTANCE  16
#define MCG_COSINE_SIMILARITY            17
#define MCG_PROBABILITY_COSINE_SIMILARITY 18
#define MCG_SIS                          23
#define MCG_RSIS                         24
#define MCG_SRLRT                        25
#define MCG_SRULRT                       26
/* Poisson Bregman divergence (I-divergence), bnj::poisson_bregman, dist/distance.h:576-585:
 * sum_j xh log(xh / ch) - xh + ch with xh = x + prior, ch = c + prior (first = row,
 * second = center).  The reference gives it no measure code; 64 is this library's,
 * outside the reference's range.  Weighted-mean centers (a Bregman divergence). */
#define MCG_POISSON                      64

/* cross-term precision for dense SQRL2 */
#define MCG_PREC_FP32  0   /* FP32 MFMA cross term, FP32 compare ("plain FP32 form") */
#define MCG_PREC_BF16  1   /* BF16 MFMA cross term + FP32 re-score of near ties */
#define MCG_PREC_EXACT 2   /* direct FP32 sum of squared differences per pair */
#define MCG_PREC_FP16  3   /* FP16 MFMA cross term (power-of-two scaled operands) + FP32
                              re-score of near ties; same assignment as MCG_PREC_BF16 */

/* coreset sensitivity methods: include/minicore/coreset/coreset.h:34-48 */
#define MCG_SENS_BFL 0
#define MCG_SENS_FL  1
#define MCG_SENS_LFKF 2   /* GMM coreset (make_gmm_sampler, coreset.h:341-381) with alpha_est = 0 */
#define MCG_SENS_VX  3
#define MCG_SENS_LBK 4

typedef struct mcg_ctx  mcg_ctx;
typedef struct mcg_data mcg_data;

/* ---- library / context -------------------------------------------------- */
const char *mcg_last_error(void);
int mcg_device_count(int *count);
int mcg_ctx_create(int device, mcg_ctx **out);
int mcg_ctx_destroy(mcg_ctx *ctx);
int mcg_ctx_synchronize(mcg_ctx *ctx);
/* The HIP stream of the context (as void*), for callers that enqueue their own work. */
int mcg_ctx_stream(mcg_ctx *ctx, void **stream);

/* Kernel timing on the context's stream (HIP events around each launch of the
 * timed kernels).  which: 0 = E-step cross-term kernel, 1 = M-step partial-sum
 * kernel.  mcg_timing_read returns the summed milliseconds and launch count
 * since the last reset. */
int mcg_timing_enable(mcg_ctx *ctx, int on);
int mcg_timing_reset(mcg_ctx *ctx);
int mcg_timing_read(mcg_ctx *ctx, int which, double *total_ms, uint64_t *count);

/* Graph replay of mcg_lloyd's iterations (off by default): the iterations after the
 * first are captured once into a HIP graph and replayed, in batches with one host
 * synchronisation each.  Off: every iteration is launched kernel by kernel (same
 * batches, same results).  While a capture is open the HIP runtime fails
 * synchronous legacy-stream calls (hipMemcpy, hipMemset) made by any other thread of
 * the process, so enable it only where no other thread makes such calls.  (No
 * counterpart in the reference.) */
int mcg_graphs_enable(mcg_ctx *ctx, int on);

/* Tuning of the BF16-refined dense SQRL2 E-step (on by default): the BF16 MFMA
 * kernels and the exact kernel give the same assignment bit for bit, and the faster
 * one for the problem's shape (padded d, padded k, log2 n) is used, timed once per
 * shape on the context.  Off: always the MFMA kernels.  (No counterpart in the
 * reference.) */
int mcg_tuning_enable(mcg_ctx *ctx, int on);

/* Points of the last BF16 E-step on this context that needed an exact re-score:
 * `pairs` chose between their two nearest centers, `full` were re-scored over
 * all centers (no counterpart in the reference; for measurement). */
int mcg_refine_counts(mcg_ctx *ctx, uint64_t *pairs, uint64_t *full);

/* ---- datasets ----------------------------------------------------------- */
/* Dense rows.  Replaces the numpy-array view that python/pycluster_dense.cpp:5-30
 * and python/pydense.cpp:34-77 build (c_style|forcecast arrays of float/double).
 * dtype MCG_F32 or MCG_F64 (converted to float32 on the device). */
int mcg_data_dense(mcg_ctx *ctx, const void *x, int dtype, uint64_t n, uint64_t d,
                   uint64_t ld, mcg_data **out);
/* CSR rows.  Replaces PyCSparseMatrix (python/pycsparse.h:33-244):
 * data f32/f64/u16, indices u16/u32/i32/u64/i64, indptr u32/u64/i32/i64. */
int mcg_data_csr(mcg_ctx *ctx, const void *data, int data_dtype, const void *indices,
                 int idx_dtype, const void *indptr, int ptr_dtype, uint64_t n, uint64_t d,
                 uint64_t nnz, mcg_data **out);
int mcg_data_free(mcg_data *ds);
int mcg_data_shape(const mcg_data *ds, uint64_t *n, uint64_t *d, uint64_t *nnz);

/* ---- distance / assignment ---------------------------------------------- */
/* Full n x k cost matrix (float32, row-major).  Replaces the per-pair
 * msr_with_prior loops of python/pycmp.cpp:16-246 and pyarrcmp.cpp:126-145
 * (cmp) and pycluster_soft.h:38-40 (soft costs). */
int mcg_cost_matrix(mcg_data *ds, int msr, double prior, const float *centers, uint64_t k,
                    float *out);
/* Condensed pairwise costs among the rows: pcmp, python/pycmp.cpp:249-338.
 * out[n*i - i*(i+1)/2 + (j - i - 1)] = msr(first = row i, second = row j) for all
 * i < j; n*(n-1)/2 floats, host or device.  Each value equals the cost matrix's
 * (mcg_cost_matrix of the rows against the rows as centers). */
int mcg_pairwise(mcg_data *ds, int msr, double prior, float *out);
/* Hard E-step: asn[i] = argmin_c msr(row i, center c) (lowest index on ties),
 * costs[i] = that minimum (nullable).  Replaces assign_points_hard,
 * include/minicore/clustering/solve.h:234-285.  precision applies to dense SQRL2. */
int mcg_assign_hard(mcg_data *ds, int msr, double prior, const float *centers, uint64_t k,
                    int precision, uint32_t *asn, float *costs);
/* Hard M-step: weighted mean of each cluster's rows, FP64 accumulation in a
 * fixed order; empty clusters keep their center.  Replaces set_centroids_hard
 * without the restart (solve.h:184-232, centroid.h:81-178).  weights: f64[n] or NULL.
 * n_empty (nullable) receives the number of empty clusters. */
int mcg_centroids_hard(mcg_data *ds, const uint32_t *asn, const double *weights, uint64_t k,
                       float *centers, uint64_t *n_empty);
/* Lloyd driver: perform_hard_clustering, solve.h:98-175, with the empty-cluster
 * restart of centroid.h:687-752.  centers (k x d) in/out; asn/costs outputs
 * (costs nullable).  Returns the reference's (initcost, finalcost, numiter).
 * L1, TVD and L2 use their median policies (set_centroids_l1 / _tvd / _l2,
 * centroid.h:434-647, with the restart of centroid.h:442-480): coordinate-wise
 * medians (dense and CSR; weighted only dense, as in the reference) and the
 * geometric median (Weiszfeld: dense rows as blaze_adaptor.h:614-690; unweighted
 * CSR rows as csc.h:1131-1203 in FP64, with the member costs indexed by position
 * where the reference indexes them by row id).  Weighted median centers need
 * positive weights; CSR L1 / TVD medians need non-negative values. */
int mcg_lloyd(mcg_data *ds, int msr, double prior, float *centers, uint64_t k,
              const double *weights, double eps, uint64_t maxiter, int precision,
              uint32_t *asn, float *costs, double *initcost, double *finalcost,
              uint64_t *numiter);
/* perform_hard_clustering (solve.h:98-175) on dense rows given as float64 for
 * SQRL2, with every value in FP64 as the reference's FT = double instantiation for
 * float64 arrays (python/pycluster_dense.cpp:85-90): per pair the sequential FP64
 * FMA chain, FP64 means (the chunked member order of mcg_lloyd), FP64 costs in the
 * control flow (each the exact fixed-point sum of w_i cost_i), the empty-cluster
 * restart of centroid.h:687-752.  ds must come from mcg_data_dense with MCG_F64;
 * centers f64[k*d] in/out, costs f64[n] or NULL, weights f64[n] or NULL.
 * MCG_ENAN when a center becomes NaN. */
int mcg_lloyd_f64(mcg_data *ds, double *centers, uint64_t k, const double *weights, double eps,
                  uint64_t maxiter, uint32_t *asn, double *costs, double *initcost,
                  double *finalcost, uint64_t *numiter);
/* assign_points_hard (solve.h:234-285) for SQRL2 on rows given as float64, in
 * FP64 (the sequential FMA chain; lowest index on ties).  costs f64[n] or NULL. */
int mcg_assign_f64(mcg_data *ds, const double *centers, uint64_t k, uint32_t *asn, double *costs);
/* One Lloyd iteration on device-resident state: asn = argmin over centers
 * (E-step), then centers = weighted cluster means (M-step, in place; empty
 * clusters keep their center); *cost = the assignment's cost against the input
 * centers (FP64; for dense SQRL2 measured in the M-step pass, otherwise the
 * weighted sum of the E-step minima).  No restart, no convergence test.
 * Replaces one pass of lloyd_iteration, include/minicore/optim/kmeans.h:359-479.
 * centers and asn must be device pointers. */
int mcg_lloyd_step(mcg_data *ds, int msr, double prior, float *centers, uint64_t k,
                   const double *weights, int precision, uint32_t *asn, double *cost);
/* Minibatch hard clustering: perform_hard_minibatch_clustering
 * (include/minicore/clustering/solve.h:377-578), reached from hcluster(...,
 * mbsize >= 0) (python/pycluster.h:31-42).
 * Check-ins (iteration 0, every `freq`-th, the last) assign every row, restart
 * centers with <= reseed_after rows at rows drawn with probability ~ w_i*cost_i,
 * and keep the centers of the lowest check-in cost; in between, mbsize rows are
 * drawn (rng() mod n, with or without replacement), sorted, assigned, and each
 * center with members becomes their weighted mean, or for L1 / TVD / L2 their
 * median (solve.h:555-572: weighted L1 and TVD index the weight vector by member
 * rank, so weighted runs need mbsize <= n; the geometric median uses the members'
 * weights).  The draws use the wyrand
 * counter stream of `seed` in the reference's order (seed 0 is the stream of 0,
 * not std::rand()).  centers (k x d) in: initial, out: best; asn/costs (nullable):
 * the last check-in's assignment and per-row costs; initcost/finalcost: the first
 * and the lowest check-in cost, exact fixed-point sums of w_i*cost_i.  Requires
 * mbsize, maxiter, freq > 0 and, without replacement, mbsize <= n (the reference
 * would loop forever). */
int mcg_minibatch(mcg_data *ds, int msr, double prior, float *centers, uint64_t k, const double *weights,
                  uint64_t mbsize, uint64_t maxiter, uint64_t freq, uint64_t reseed_after, int with_replacement,
                  uint64_t seed, int precision, uint32_t *asn, float *costs, double *initcost, double *finalcost,
                  uint64_t *numiter);

/* Coreset-minibatch hard clustering: hmb_coreset_clustering
 * (include/minicore/clustering/solve.h:598-789), reached from hcluster(...,
 * mbsize >= 0, cs=True) (python/pycluster.h:37-39).  ceil(maxiter / freq) + 1
 * check-ins as in mcg_minibatch (the saved centers change only when a later
 * check-in is strictly cheaper); after each but the last, LBK sensitivities from
 * the check-in (coreset.h:487-533) and freq coresets of min(mbsize, n) i.i.d.
 * draws (seeded by the stream's next output, sorted), each clustered from the
 * current centers by perform_hard_clustering with eps 1e-3 and maxiter
 * 1 + 2*(round < 2) on the coreset rows with weights w_i * w_i / (m p_i)
 * (solve.h:771-772).  Outputs as mcg_minibatch. */
int mcg_cs_minibatch(mcg_data *ds, int msr, double prior, float *centers, uint64_t k, const double *weights,
                     uint64_t mbsize, uint64_t maxiter, uint64_t freq, uint64_t reseed_after, uint64_t seed,
                     int precision, uint32_t *asn, float *costs, double *initcost, double *finalcost,
                     uint64_t *numiter);

/* Soft EM driver: perform_soft_clustering, solve.h:287-374 with
 * set_centroids_full_mean (soft), centroid.h:781-963.  costs and resp are
 * n x k float32 outputs (nullable).  Centers are responsibility-weighted means;
 * for L1 (dense rows) the weighted_median rule over all rows and for L2 the
 * Weiszfeld geometric median of all rows, with weights w_i * r_ic
 * (centroid.h:815-838; CSR L1 -> MCG_EUNSUPPORTED as the reference throws; an L2
 * center whose weights are all zero -> MCG_ENAN). */
int mcg_soft_em(mcg_data *ds, int msr, double prior, double temp, float *centers, uint64_t k,
                const double *weights, double eps, uint64_t maxiter, float *costs, float *resp,
                double *initcost, double *finalcost, uint64_t *numiter);

/* ---- seeding and coresets ------------------------------------------------ */
/* D^2 seeding: coresets::kmeanspp, include/minicore/optim/kmeans.h:49-215,
 * as bound by python/smw.h:179-410 (n_local_trials = 1, lspp = 0; mcg_kmeanspp_ls and
 * mcg_kmeanspp_local take the other two).
 * centers u32[k] (row indices), asn u32[n], costs f32[n].  weights nullable. */
int mcg_kmeanspp(mcg_data *ds, int msr, double prior, uint64_t k, uint64_t seed,
                 const double *weights, uint32_t *centers, uint32_t *asn, float *costs);

/* D^2 seeding followed by `rounds` rounds of local search++ (localsearchpp_rounds,
 * include/minicore/optim/lsearchpp.h:21-121, reached from kmeans.h:209-212 and the
 * kmeanspp bindings' lspp argument).  Each round draws a row with the seeding
 * stream's next output (over the distances; with weights over w_i times the last
 * round's new costs, as lsearchpp.h:53-56 does), and swaps it in for the center whose
 * removal charge (rows' second-best minus best) is below the credit of the rows it
 * would serve better; charges and credit are exact fixed-point sums.  The n x k cost
 * matrix stays on the device.  Costs are row-to-center-row in the seeding's
 * orientation (the reference's matrix uses the opposite argument order, which
 * matters only for asymmetric measures). */
int mcg_kmeanspp_ls(mcg_data *ds, int msr, double prior, uint64_t k, uint64_t seed, const double *weights,
                    uint64_t rounds, uint32_t *centers, uint32_t *asn, float *costs);

/* D^2 seeding with `trials` local trials per center (coresets::kmeanspp with
 * n_local_samples = trials, kmeans.h:79-92, 114-177; the bindings' n_local_trials,
 * python/smw.h:181-249), then `rounds` rounds of local search++ as in
 * mcg_kmeanspp_ls.  trials = 1 is mcg_kmeanspp (then mcg_kmeanspp_ls).  With
 * trials > 1: the first center is the cheapest of `trials` uniform rows; each later
 * round draws `trials` rows from w * dist (outputs 0 .. trials-1 of the wyrand
 * stream seeded by the round's draw) and keeps the one whose update gives the lowest
 * exact total (the first on ties).  As in the reference, candidates after the first
 * start from the distances and assignment saved right after the first center, and
 * the winner's are kept.  The device holds `trials` copies of the distances and
 * assignment. */
int mcg_kmeanspp_local(mcg_data *ds, int msr, double prior, uint64_t k, uint64_t seed, const double *weights,
                       uint64_t trials, uint64_t rounds, uint32_t *centers, uint32_t *asn, float *costs);
/* Sensitivities: CoresetSampler::make_sampler, coreset.h:386-578.
 * costs f64[n], asn u32[n], weights f64[n] or NULL; probs f32[n]. */
int mcg_coreset_probs(mcg_ctx *ctx, int sens, uint64_t n, uint64_t ncenters,
                      const double *costs, const uint32_t *asn, const double *weights,
                      float *probs);
/* mcg_coreset_probs without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  costs, asn, weights
 * (nullable) and probs must be device buffers; results are bit-identical to
 * mcg_coreset_probs.  The kernels' scratch is one fixed device array per GPU
 * (33.5 MiB, held from the first mcg_ctx_create on that GPU until the process
 * ends), so n <= 4194304 and ncenters <= 65536 (MCG_EINVAL beyond), and calls
 * that overlap on one GPU (from any context or thread) race on it. */
int mcg_coreset_probs_noalloc(mcg_ctx *ctx, int sens, uint64_t n, uint64_t ncenters,
                              const double *costs, const uint32_t *asn, const double *weights,
                              float *probs);
/* m independent draws from probs: CoresetSampler::sample, coreset.h:625-651.
 * Draw t uses the (first_draw + t)-th output of the seed's random stream.
 * indices u64[m], out_weights f32[m] = w_i / (m * p_i). */
int mcg_coreset_sample(mcg_ctx *ctx, uint64_t n, const float *probs, const double *weights,
                       uint64_t m, uint64_t seed, uint64_t first_draw, uint64_t *indices,
                       float *out_weights);
/* A new dataset on ds's context holding rows idx[0..m) of ds in that order
 * (repeats allowed); idx u64, host or device.  Replaces the coreset row
 * selection rows(mat, coreset.indices_) / set_center(row(smat, i), ...) of
 * hmb_coreset_clustering, include/minicore/clustering/solve.h:748-762. */
int mcg_data_rows(const mcg_data *ds, const uint64_t *idx, uint64_t m, mcg_data **out);

/* ---- on-disk formats around the path (host only, no device work) --------- */
/* CoresetSampler::write(gzFile), include/minicore/coreset/coreset.h:282-293: a
 * gzip stream of u64 n, u64 seed, f32 probs[n], u32 flag (137 when weights
 * follow, 0 otherwise) and f32 weights[n].  weights may be NULL.  The sampler of
 * the FL method is refused, as the reference refuses it (coreset.h:283). */
int mcg_sampler_write(const char *path, uint64_t n, uint64_t seed, const float *probs, const float *weights);
/* CoresetSampler::read(gzFile), coreset.h:306-322 (gzip or uncompressed input).
 * Fills *n, *seed and *has_weights (any nonzero flag means weights follow, as in
 * the reference's release build); probs and weights are written only when cap >= n
 * (call first with cap = 0 to learn n).  A truncated file is an MCG_EIO error. */
int mcg_sampler_read(const char *path, uint64_t *n, uint64_t *seed, int *has_weights, float *probs,
                     float *weights, uint64_t cap);
/* IndexCoreset<IT, FT>::write(gzFile), coreset.h:113-123: a gzip stream of u64 m,
 * IT indices[m], FT weights[m].  idx_dtype MCG_U32 / MCG_U64 / MCG_U16 / MCG_U8,
 * w_dtype MCG_F32 / MCG_F64: the IT and FT the file was made with. */
int mcg_index_coreset_write(const char *path, uint64_t m, const void *indices, int idx_dtype,
                            const void *weights, int w_dtype);
/* IndexCoreset<IT, FT>::read(gzFile), coreset.h:94-106: *m filled; indices and
 * weights written when cap >= m.  Weights are read as m * sizeof(FT) bytes, the
 * layout write produces (the reference's read sizes them by IT). */
int mcg_index_coreset_read(const char *path, uint64_t *m, void *indices, int idx_dtype, void *weights,
                           int w_dtype, uint64_t cap);

/* ---- serving ------------------------------------------------------------ */
/* A server runs clustering requests from many callers on one GPU: worker
 * threads, each with its own context and HIP stream, take requests in
 * arrival order as soon as the request's device footprint fits in the HBM
 * budget, so requests of different callers run concurrently on the device.
 * Each request runs the pipeline of one round of the reference's coreset
 * clustering (solve.h:716-762): D^2 seeding on all rows (kmeans.h:49-215),
 * sensitivities (coreset.h:386-578), a weighted coreset of
 * m = min(n, max(ceil(coreset_frac * n), k)) rows (coreset.h:625-651), weighted
 * Lloyd on the coreset from the seeded centers (solve.h:98-175), then one hard
 * assignment of all rows to the resulting centers (solve.h:234-285).  A
 * request's result does not depend on what else runs on the server.
 * (The reference has no serving layer; these entry points have no counterpart
 * there.) */
typedef struct mcg_server mcg_server;

/* request states */
#define MCG_REQ_QUEUED    0
#define MCG_REQ_RUNNING   1
#define MCG_REQ_DONE      2
#define MCG_REQ_FAILED    3
#define MCG_REQ_CANCELLED 4

typedef struct mcg_request {
    /* rows: dense when indices == NULL (data: float32/float64, ld >= d, 0 = d),
     * else CSR (data, indices, indptr with their dtypes as in mcg_data_csr).
     * Host or device buffers; they must stay valid until the request ends. */
    const void *data;
    int data_dtype;
    const void *indices;
    int idx_dtype;
    const void *indptr;
    int ptr_dtype;
    uint64_t n, d, nnz, ld;
    const double *weights;   /* f64[n] or NULL */
    int msr;
    double prior;
    uint64_t k;
    uint64_t seed;           /* seeding and coreset draws */
    int sens;                /* MCG_SENS_* */
    double coreset_frac;     /* in (0, 1] */
    double eps;              /* Lloyd stop rule on the coreset (solve.h:165-170) */
    uint64_t maxiter;        /* Lloyd maxiter on the coreset (maxiter + 1 passes at most) */
    int precision;           /* MCG_PREC_* for dense SQRL2 */
    int fail_at_stage;       /* 0; 1..4 makes the request fail at that stage (tests of error paths) */
    /* outputs: caller-owned host or device buffers, written when the request is done */
    float *centers;          /* k x d */
    uint32_t *asn;           /* n */
    float *costs;            /* n, nullable */
} mcg_request;

typedef struct mcg_result {
    int state;               /* MCG_REQ_* */
    int status;              /* MCG_OK or the error code of a failed request */
    double cost;             /* sum_i w_i cost_i of the final assignment (FP64) */
    double coreset_initcost, coreset_finalcost;
    uint64_t coreset_m;      /* rows in the coreset */
    uint64_t numiter;        /* Lloyd iterations on the coreset */
    double queue_ms;         /* submit -> start */
    double run_ms;           /* start -> end */
    double stage_ms[5];      /* rows + seeding, sensitivities, coreset, Lloyd, final assignment */
    char error[256];         /* message of a failed request */
} mcg_result;

/* nworkers >= 1 concurrent requests; hbm_budget bytes of device memory for the
 * requests' working sets (0 = 80% of the free device memory at creation). */
int mcg_server_create(int device, int nworkers, uint64_t hbm_budget, mcg_server **out);
/* Cancels queued requests, waits for running ones, joins the workers. */
int mcg_server_destroy(mcg_server *s);
/* Queues a copy of *req; *ticket identifies it.  MCG_ENOMEM if the request's
 * footprint exceeds the whole budget; MCG_EINVAL for malformed requests. */
int mcg_submit(mcg_server *s, const mcg_request *req, uint64_t *ticket);
/* Non-blocking: the request's state and, once it ended, its result. */
int mcg_poll(mcg_server *s, uint64_t ticket, mcg_result *res);
/* Blocks until the request ends or timeout_s passes (< 0: no limit); then as mcg_poll. */
int mcg_wait(mcg_server *s, uint64_t ticket, double timeout_s, mcg_result *res);
/* A queued request is dropped; a running one stops at its next stage boundary.
 * The outputs of a cancelled request are unspecified. */
int mcg_cancel(mcg_server *s, uint64_t ticket);
/* Forgets an ended request (its ticket becomes invalid). */
int mcg_release(mcg_server *s, uint64_t ticket);

#ifdef __cplusplus
}
#endif
#endif
