/*
 * mco.h -- CPU oracle for the minicore clustering hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the pytest suite,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it to compute
 * expected results.  No product code path (minicore_amd, libmcg) links, loads
 * or calls it.
 *
 * It restates, in plain C (FP64 unless a function says otherwise), the
 * reference algorithms cited beside each declaration (paths relative to the
 * dnbaker/minicore source tree).  Randomness and sampling are restated with
 * the same definitions the GPU library uses (wyrand counter stream,
 * fixed-point inverse-CDF draws) so that sampled indices can be compared
 * bit for bit.
 */
#ifndef MCO_H
#define MCO_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Measure codes: include/minicore/dist/distance.h:24-64 */
/* members per fixed-order FP64 chunk of the dense hard M-step's sums */
#define MCO_MEAN_CHUNK 256

enum { MCO_SQRL2 = 2, MCO_JSD = 4, MCO_MKL = 5, MCO_REVERSE_MKL = 11 };
/* further measures (distance.h:24-64): applicator.h:1673-1749 (dense), 1778-1891 (sparse) */
enum {
    MCO_L1 = 0, MCO_L2 = 1, MCO_JSM = 3, MCO_HELLINGER = 6, MCO_BHATTACHARYYA_METRIC = 7,
    MCO_BHATTACHARYYA_DISTANCE = 8, MCO_TVD = 9, MCO_LLR = 10, MCO_UWLLR = 12, MCO_IS = 13,
    MCO_REVERSE_IS = 14, MCO_COSINE_DISTANCE = 15, MCO_PROBABILITY_COSINE_DISTANCE = 16,
    MCO_COSINE_SIMILARITY = 17, MCO_PROBABILITY_COSINE_SIMILARITY = 18, MCO_SIS = 23, MCO_RSIS = 24,
    MCO_SRLRT = 25, MCO_SRULRT = 26,
    MCO_POISSON = 64   /* I-divergence (poisson_bregman, distance.h:576-585); no reference code */
};

/* ---- random numbers: published wyrand, counter form ------------------- */
uint64_t mco_wyrand_at(uint64_t seed, uint64_t j);      /* j-th output (0-based) */
uint64_t mco_mulhi64(uint64_t a, uint64_t b);

/* ---- fixed-point inverse-CDF sampler ---------------------------------- */
int      mco_fixed_shift(double maxw, uint64_t n);
void     mco_fixed_weights(const double *w, uint64_t n, uint64_t *out);
uint64_t mco_fixed_draw(const uint64_t *fw, uint64_t n, uint64_t r); /* r = raw 64-bit RNG output */

/* ---- dissimilarity: applicator.h:1632-1915 ---------------------------- */
/* Dense pair, FP64.  x = data row (first argument), c = center (second). */
double mco_msr_dense(int msr, const double *x, const double *c, uint64_t d,
                     double prior, double xsum, double csum);
/* Sparse pair via the ordered union walk of merge.h:26-65, FP64. */
double mco_msr_sparse(int msr, const double *xv, const uint32_t *xi, uint64_t xn,
                      const double *cv, const uint32_t *ci, uint64_t cn, uint64_t d,
                      double prior, double xsum, double csum, uint64_t *sharednz_out);
/* Squared L2 in FP32 as a sequential fused multiply-add chain. */
float    mco_sqrl2_f32(const float *x, const float *c, uint64_t d);
/* merge.h:26-65 shared-zero count of two sorted index lists. */
uint64_t mco_shared_zeros(const uint32_t *a, uint64_t an, const uint32_t *b, uint64_t bn, uint64_t d);

/* ---- full cost matrices (dense rows or CSR rows vs dense centers) ----- */
void mco_cost_matrix_dense(int msr, const float *X, uint64_t n, uint64_t d,
                           const float *C, uint64_t k, double prior, double *out /* n*k */);
void mco_cost_matrix_csr(int msr, const float *val, const uint32_t *idx, const uint64_t *indptr,
                         uint64_t n, uint64_t d, const float *C, uint64_t k, double prior,
                         double *out /* n*k */);

/* ---- hard assignment: solve.h:234-285 ---------------------------------- */
/* SQRL2 on dense f32 data with FP32 sequential-FMA costs (lowest index wins ties).
 * gap_out (nullable) receives second-best minus best, relative to best. */
void mco_assign_sqrl2_f32(const float *X, uint64_t n, uint64_t d, const float *C, uint64_t k,
                          uint32_t *asn, float *cost, double *gap_out, int nthreads);
/* Any measure, FP64 costs, dense or CSR rows (val==NULL means dense X). */
void mco_assign_f64(int msr, const float *X, const float *val, const uint32_t *idx,
                    const uint64_t *indptr, uint64_t n, uint64_t d, const float *C, uint64_t k,
                    double prior, uint32_t *asn, double *cost, double *gap_out);

/* ---- M-step: centroid.h:81-178, 649-779 -------------------------------- */
/* Weighted mean of the rows of each cluster accumulated in FP64 in a fixed order
 * (ascending point index, chunks of `chunk` members summed sequentially, chunk
 * sums added in order).  Empty clusters keep their previous center.  Returns
 * the number of empty clusters. */
uint64_t mco_centroids_dense(const float *X, uint64_t n, uint64_t d, const uint32_t *asn,
                             const double *w, uint64_t k, uint64_t chunk, float *C_inout);
/* one pass of lloyd_iteration (kmeans.h:359-479), dense squared L2, E then M with
 * the empty-cluster reseed; returns the number of reseed draws */
uint64_t mco_lloyd_step(const float *X, uint64_t n, uint64_t d, float *C, uint64_t k, const double *w,
                        uint32_t *asn, double *cost);
uint64_t mco_centroids_csr(const float *val, const uint32_t *idx, const uint64_t *indptr,
                           uint64_t n, uint64_t d, const uint32_t *asn, const double *w,
                           uint64_t k, float *C_inout);

/* ---- Lloyd driver: solve.h:98-175 ------------------------------------- */
/* Dense (val==NULL) or CSR rows; centers dense k x d (updated in place).
 * Returns 0, or -1 on a NaN center (centroid.h:771-775). */
int mco_lloyd(int msr, const float *X, const float *val, const uint32_t *idx,
              const uint64_t *indptr, uint64_t n, uint64_t d, float *C, uint64_t k,
              double prior, const double *w, double eps, uint64_t maxiter,
              uint32_t *asn, double *cost, double *initcost, double *finalcost,
              uint64_t *numiter, int nthreads);

/* ---- minibatch driver: solve.h:377-578 -------------------------------- */
/* Dense (val==NULL) or CSR rows; C (k x d) in: initial centers, out: the best
 * check-in's centers; asn/cost: the last check-in's assignment.  freq is the
 * check-in period (calc_cost_freq).  Returns -2 for invalid sizes. */
int mco_minibatch(int msr, const float *X, const float *val, const uint32_t *idx, const uint64_t *indptr,
                  uint64_t n, uint64_t d, float *C, uint64_t k, double prior, const double *w,
                  uint64_t mbsize, uint64_t maxiter, uint64_t freq, uint64_t reseed_after, int with_rep,
                  uint64_t seed, uint32_t *asn, double *cost, double *initcost, double *finalcost,
                  uint64_t *numiter, int nthreads);

/* Coreset-minibatch driver, solve.h:598-789 (see mco.c).  Returns -1 on a NaN
 * center in a coreset Lloyd run, -2 for invalid sizes. */
int mco_cs_minibatch(int msr, const float *X, const float *val, const uint32_t *idx, const uint64_t *indptr,
                     uint64_t n, uint64_t d, float *C, uint64_t k, double prior, const double *w,
                     uint64_t mbsize, uint64_t maxiter, uint64_t freq, uint64_t reseed_after,
                     uint64_t seed, uint32_t *asn, double *cost, double *initcost, double *finalcost,
                     uint64_t *numiter, int nthreads);

/* ---- soft clustering: solve.h:287-374, centroid.h:781-963 --------------- */
int mco_soft(int msr, const float *X, const float *val, const uint32_t *idx,
             const uint64_t *indptr, uint64_t n, uint64_t d, float *C, uint64_t k,
             double prior, double temp, const double *w, double eps, uint64_t maxiter,
             float *costs /* n*k */, float *resp /* n*k */, double *initcost,
             double *finalcost, uint64_t *numiter);
void mco_softmax_rows(const float *costs, uint64_t n, uint64_t k, double temp, float *resp);

/* ---- D^2 seeding: kmeans.h:49-215 -------------------------------------- */
/* Centers are row indices; distances from msr(first=point, second=center row).
 * Returns the number of centers chosen. */
uint64_t mco_kmeanspp(int msr, const float *X, const float *val, const uint32_t *idx,
                      const uint64_t *indptr, uint64_t n, uint64_t d, uint64_t k,
                      double prior, uint64_t seed, const double *w,
                      uint32_t *centers, uint32_t *asn, float *dist, int nthreads);

/* D^2 seeding followed by `rounds` rounds of local search++ (lsearchpp.h:21-121). */
uint64_t mco_kmeanspp_ls(int msr, const float *X, const float *val, const uint32_t *idx,
                         const uint64_t *indptr, uint64_t n, uint64_t d, uint64_t k,
                         double prior, uint64_t seed, const double *w, uint64_t rounds,
                         uint32_t *centers, uint32_t *asn, float *dist, int nthreads);
/* local search++ rounds (lsearchpp.h:21-121) from the given centers (k >= 2, distinct
 * rows); draws from outputs j0, j0+1, ... of the stream of seed */
void mco_localsearch(int msr, const float *X, const float *val, const uint32_t *idx, const uint64_t *indptr,
                     uint64_t n, uint64_t d, uint64_t k, double prior, uint64_t seed, const double *w,
                     uint64_t rounds, uint64_t j0, uint32_t *centers, uint32_t *asn, float *distf);
/* k-means++ with n_local_samples > 1: kmeans.h:79-92, 114-177 */
uint64_t mco_kmeanspp_local(int msr, const float *X, const float *val, const uint32_t *idx,
                            const uint64_t *indptr, uint64_t n, uint64_t d, uint64_t k, double prior,
                            uint64_t seed, const double *w, uint64_t L, uint64_t rounds, uint32_t *centers,
                            uint32_t *asn, float *distf);
/* median centroid policies (L1, TVD, L2): centroid.h:434-647 */
int mco_median_centers(int msr, const float *X, const float *val, const uint32_t *idx, const uint64_t *indptr,
                       uint64_t n, uint64_t d, const double *w, uint64_t k, uint32_t *asn, double *cost, float *C,
                       double eps);
int mco_lloyd_median(int msr, const float *X, const float *val, const uint32_t *idx, const uint64_t *indptr,
                     uint64_t n, uint64_t d, float *C, uint64_t k, double prior, const double *w, double eps,
                     uint64_t maxiter, uint32_t *asn, double *cost, double *initcost, double *finalcost,
                     uint64_t *numiter);

/* ---- coreset sensitivities and sampling: coreset.h:340-699 ------------ */
/* sens: 0 BFL, 1 FL, 3 VX, 4 LBK.  probs out in float (CoresetSampler<float,...>). */
int  mco_coreset_probs(int sens, uint64_t n, uint64_t ncenters, const double *costs,
                       const uint32_t *asn, const double *w, float *probs);
void mco_coreset_sample(uint64_t n, const float *probs, const double *w, uint64_t m,
                        uint64_t seed, uint64_t first, uint64_t *indices, float *weights);

#ifdef __cplusplus
}
#endif
/* float64 rows, dense SQRL2 (FT = double): FP64 chains, means and control flow */
double   mco_sqrl2_f64(const double *x, const double *c, uint64_t d);
void     mco_assign_sqrl2_f64(const double *X, uint64_t n, uint64_t d, const double *C, uint64_t k,
                              uint32_t *asn, double *cost);
uint64_t mco_centroids_dense_f64(const double *X, uint64_t n, uint64_t d, const uint32_t *asn, const double *w,
                                 uint64_t k, uint64_t chunk, double *C);
double   mco_fixed_total(const double *cost, const double *w, uint64_t n);
int      mco_lloyd_f64(const double *X, uint64_t n, uint64_t d, double *C, uint64_t k, const double *w, double eps,
                       uint64_t maxiter, uint32_t *asn, double *cost, double *initcost, double *finalcost,
                       uint64_t *numiter);

#endif
