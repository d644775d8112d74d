/*
 * mco.c -- CPU oracle for the minicore clustering hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.  Every function cites the reference lines it restates
 * (paths relative to the dnbaker/minicore tree).
 *
 * Built with -ffp-contract=off so that every a*b+c below is two roundings
 * unless the code calls fma() explicitly: the GPU library follows the same
 * operation sequence where results are compared bit for bit.
 */
#include "mco.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ======================================================================
 * Random numbers.  The reference draws from wy::WyRand (aesctr/wy.h, an
 * un-vendored submodule; shared.h:5).  This restates the published wyrand
 * generator (Wang Yi, wyhash): state += 0xa0761d6478bd642f; output =
 * mum(state, state ^ 0xe7037ed1a0b428db), mum = hi64 ^ lo64 of the 128-bit
 * product.  The j-th output is a pure function of (seed, j).
 * ====================================================================== */
uint64_t mco_mulhi64(uint64_t a, uint64_t b) {
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
}
uint64_t mco_wyrand_at(uint64_t seed, uint64_t j) {
    uint64_t s = seed + (j + 1) * 0xa0761d6478bd642fULL;
    unsigned __int128 p = (unsigned __int128)s * (s ^ 0xe7037ed1a0b428dbULL);
    return (uint64_t)(p >> 64) ^ (uint64_t)p;
}

/* ======================================================================
 * Fixed-point inverse-CDF sampler.  The reference samples index i with
 * probability w_i / sum(w) through reservoir_simd::sample (kmeans.h:123,130;
 * centroid.h:709) and an alias table (coreset.h:382-385, 641); both live in
 * un-vendored submodules.  Restated here as: w_i -> floor(w_i * 2^s) in
 * uint64 with s chosen from max(w) and n so the total fits in 63 bits;
 * draw u = mulhi(r, T) for a raw 64-bit RNG output r; return the first i
 * whose inclusive integer prefix sum exceeds u.  Integer sums are exact, so
 * the result does not depend on summation order.
 * ====================================================================== */
static int bitlen64(uint64_t v) { int b = 0; while (v) { ++b; v >>= 1; } return b; }

int mco_fixed_shift(double maxw, uint64_t n) {
    int e;
    if (!(maxw > 0.) || !isfinite(maxw)) return 0;
    frexp(maxw, &e);                 /* maxw < 2^e */
    return 62 - e - bitlen64(n);     /* n * maxw * 2^s < 2^62 */
}

void mco_fixed_weights(const double *w, uint64_t n, uint64_t *out) {
    double mx = 0.;
    for (uint64_t i = 0; i < n; ++i) if (w[i] > mx) mx = w[i];
    int s = mco_fixed_shift(mx, n);
    for (uint64_t i = 0; i < n; ++i) {
        double v = w[i] > 0. ? floor(ldexp(w[i], s)) : 0.;
        out[i] = (uint64_t)v;
    }
}

uint64_t mco_fixed_draw(const uint64_t *fw, uint64_t n, uint64_t r) {
    uint64_t T = 0;
    for (uint64_t i = 0; i < n; ++i) T += fw[i];
    if (T == 0) return 0;
    uint64_t u = mco_mulhi64(r, T), acc = 0;
    for (uint64_t i = 0; i < n; ++i) {
        acc += fw[i];
        if (acc > u) return i;
    }
    return n - 1;
}

/* ======================================================================
 * Dissimilarity with additive prior: applicator.h:1632-1915.
 * first = data row x ("ctr" in the reference's parameter list), second =
 * center c ("mr"); lhsum/lhinc belong to the center, rhsum/rhinc to the row.
 * ====================================================================== */
typedef struct { double pv, lhrsi, rhrsi, lhinc, rhinc, lhsum, rhsum; } smooth_t;

/* applicator.h:1642-1656 */
static smooth_t smoothing(double prior, uint64_t d, double xsum, double csum) {
    smooth_t s;
    double prior_sum = prior * (double)d;
    double lhsum = csum + prior_sum, rhsum = xsum + prior_sum;
    double smallest = 7.0069e-42 * (lhsum + rhsum);
    s.pv = prior > smallest ? prior : smallest;
    prior_sum = s.pv * (double)d;
    lhsum = csum + prior_sum; rhsum = xsum + prior_sum;
    s.lhsum = lhsum; s.rhsum = rhsum;
    s.lhrsi = 1. / lhsum; s.rhrsi = 1. / rhsum;
    s.lhinc = (s.pv != 0. && lhsum != 0.) ? s.pv * s.lhrsi : 0.;
    s.rhinc = (s.pv != 0. && rhsum != 0.) ? s.pv * s.rhrsi : 0.;
    if (isnan(s.lhinc)) s.lhinc = 0.;
    if (isnan(s.rhinc)) s.rhinc = 0.;
    return s;
}

/* One coordinate's contribution given smoothed center value a and smoothed
 * row value b (applicator.h:1732-1736: MKL = kl(a||b), REVERSE_MKL =
 * kl(b||a), JSD = half the symmetric sum against the midpoint). */
static double coord_term(int msr, double a, double b) {
    switch (msr) {
    case MCO_MKL:         return a * log(a / b);
    case MCO_REVERSE_MKL: return b * log(b / a);
    case MCO_JSD: {
        double m = a + b;
        return 0.5 * (a * log(2. * a / m) + b * log(2. * b / m));
    }
    }
    return NAN;
}

/* applicator.h:1750-1760 / 1876-1891: clamp at 0; +inf -> DBL_MAX; -inf -> 0 */
static double post_kl(double r) {
    r = (r < 0.) ? 0. : r;   /* std::max(ret, 0.) keeps NaN */
    if (r == INFINITY) r = DBL_MAX;
    return r;
}

/* ---- the further measures, applicator.h:1673-1760 (dense rows: every coordinate).
 * a = smoothed center value (lhrsi, lhinc), b = smoothed row value (rhrsi, rhinc).
 * Per coordinate: HELLINGER (sqrt a - sqrt b)^2, BHATTACHARYYA sqrt(a b), TVD |a - b|/2,
 * L1 |c - x|, LLR family (lambda = lhsum / (lhsum + rhsum), m = lambda a + (1-lambda) b)
 * lambda a log(a/m) + (1-lambda) b log(b/m), IS a/b - log(a/b) - 1 (REVERSE_IS with a, b
 * swapped), SIS log((a/b + b/a + 2)/4) / 2; cosine: dot / (norm norm) of (c + pv, x + pv)
 * or of (a, b) for the probability forms. */
static double llr_term(double a, double b, double lam) {
    const double m = lam * a + (1. - lam) * b;
    return (a != 0. ? lam * a * log(a / m) : 0.) + (b != 0. ? (1. - lam) * b * log(b / m) : 0.);
}

/* applicator.h:1736-1760: the KL-family tail (LLR scaling, clamp, square roots, infinities) */
static double kl_tail(int msr, double klc, const smooth_t *sm) {
    double ret = klc;
    if (msr == MCO_LLR || msr == MCO_SRLRT) ret *= (sm->lhsum + sm->rhsum);
    ret = ret < 0. ? 0. : ret;
    if (msr == MCO_SRULRT || msr == MCO_SRLRT || msr == MCO_JSM) ret = sqrt(ret);
    if (ret == INFINITY) ret = DBL_MAX;
    return ret;
}

/* One coordinate of the I-divergence, bnj::poisson_bregman (distance.h:576-585):
 * xh log(xh/ch) - xh + ch, xh = x + prior, ch = c + prior (0 for xh = 0 gives ch). */
static double poisson_coord(double x, double c, double prior) {
    const double xh = x + prior, ch = c + prior;
    if (xh == 0.) return ch;
    if (ch == 0.) return INFINITY;
    return xh * log(xh / ch) - xh + ch;
}
static double poisson_post(double s) {
    if (s < 0.) s = 0.;
    return s > FLT_MAX ? FLT_MAX : s;
}

static double msr_more_dense(int msr, const double *x, const double *c, uint64_t d, double prior, double xsum,
                             double csum) {
    if (msr == MCO_POISSON) {
        double s = 0.;
        for (uint64_t j = 0; j < d; ++j) s += poisson_coord(x[j], c[j], prior);
        return poisson_post(s);
    }
    const smooth_t sm = smoothing(prior, d, xsum, csum);
    const double lam = sm.lhsum / (sm.lhsum + sm.rhsum);
    double s = 0., s2 = 0., s3 = 0.;
    for (uint64_t j = 0; j < d; ++j) {
        const double a = c[j] * sm.lhrsi + sm.lhinc, b = x[j] * sm.rhrsi + sm.rhinc;
        switch (msr) {
        case MCO_L1: s += fabs(c[j] - x[j]); break;
        case MCO_L2: { const double t = c[j] - x[j]; s += t * t; } break;
        case MCO_HELLINGER: { const double t = sqrt(a) - sqrt(b); s += t * t; } break;
        case MCO_BHATTACHARYYA_METRIC: case MCO_BHATTACHARYYA_DISTANCE: s += sqrt(a * b); break;
        case MCO_TVD: s += fabs(a - b); break;
        case MCO_COSINE_DISTANCE: case MCO_COSINE_SIMILARITY: {
            const double u = c[j] + sm.pv, v = x[j] + sm.pv;
            s += u * v; s2 += u * u; s3 += v * v;
        } break;
        case MCO_PROBABILITY_COSINE_DISTANCE: case MCO_PROBABILITY_COSINE_SIMILARITY:
            s += a * b; s2 += a * a; s3 += b * b; break;
        case MCO_JSD: case MCO_JSM: s += coord_term(MCO_JSD, a, b); break;
        case MCO_LLR: case MCO_UWLLR: case MCO_SRLRT: case MCO_SRULRT: s += llr_term(a, b, lam); break;
        case MCO_IS: s += a / b - log(a / b); break;
        case MCO_REVERSE_IS: s += b / a - log(b / a); break;
        case MCO_SIS: case MCO_RSIS: s += .5 * log((a / b + b / a + 2.) * .25); break;
        default: return NAN;
        }
    }
    switch (msr) {
    case MCO_L1: return s;
    case MCO_L2: return sqrt(s);
    case MCO_HELLINGER: return sqrt(s) * 0.70710678118654752440;
    case MCO_BHATTACHARYYA_METRIC: case MCO_BHATTACHARYYA_DISTANCE: {
        if (1. - s < 1e-8) s = 1.;
        return msr == MCO_BHATTACHARYYA_METRIC ? sqrt(1. - s > 0. ? 1. - s : 0.) : -log(s);
    }
    case MCO_TVD: return .5 * s;
    case MCO_COSINE_DISTANCE: case MCO_COSINE_SIMILARITY:
    case MCO_PROBABILITY_COSINE_DISTANCE: case MCO_PROBABILITY_COSINE_SIMILARITY: {
        double r = s / sqrt(s2 * s3);
        r = r > 1. ? 1. : (r < 0. ? 0. : r);
        if (msr == MCO_COSINE_DISTANCE || msr == MCO_PROBABILITY_COSINE_DISTANCE) r = acos(r) * 0.31830988618379067153;
        return r;
    }
    case MCO_IS: case MCO_REVERSE_IS: return kl_tail(msr, s - (double)d, &sm);
    default: return kl_tail(msr, s, &sm);
    }
}

double mco_msr_dense(int msr, const double *x, const double *c, uint64_t d,
                     double prior, double xsum, double csum) {
    if (msr != MCO_SQRL2 && msr != MCO_MKL && msr != MCO_REVERSE_MKL && msr != MCO_JSD)
        return msr_more_dense(msr, x, c, d, prior, xsum, csum);
    if (msr == MCO_SQRL2) {                       /* applicator.h:1680-1685 */
        double s = 0.;
        for (uint64_t j = 0; j < d; ++j) { double t = c[j] - x[j]; s += t * t; }
        return s;
    }
    smooth_t sm = smoothing(prior, d, xsum, csum);
    double r = 0.;
    for (uint64_t j = 0; j < d; ++j) {
        double a = c[j] * sm.lhrsi + sm.lhinc;    /* smoothed center */
        double b = x[j] * sm.rhrsi + sm.rhinc;    /* smoothed row */
        r += coord_term(msr, a, b);
    }
    return post_kl(r);
}

/* merge.h:26-65: ordered union walk; returns the number of coordinates absent from both. */
uint64_t mco_shared_zeros(const uint32_t *a, uint64_t an, const uint32_t *b, uint64_t bn, uint64_t d) {
    uint64_t i = 0, j = 0, u = 0;
    while (i < an && j < bn) {
        if (a[i] == b[j]) { ++i; ++j; }
        else if (a[i] < b[j]) ++i;
        else ++j;
        ++u;
    }
    u += (an - i) + (bn - j);
    return d - u;
}

/* applicator.h:1778-1891 for the further measures: union terms plus the closed-form
 * shared-zero terms the reference adds (the cosine forms use the union only, and
 * REVERSE_SIS swaps the vectors but not the increments, as the reference does). */
static double msr_more_sparse(int msr, const double *xv, const uint32_t *xi, uint64_t xn, const double *cv,
                              const uint32_t *ci, uint64_t cn, uint64_t d, double prior, double xsum, double csum,
                              uint64_t sharednz) {
    if (msr == MCO_POISSON) {   /* shared zeros contribute prior*log(1) - prior + prior = 0 */
        double s = 0.;
        uint64_t i = 0, j = 0;
        while (i < xn || j < cn) {
            if (j >= cn || (i < xn && xi[i] < ci[j])) s += poisson_coord(xv[i++], 0., prior);
            else if (i >= xn || ci[j] < xi[i])        s += poisson_coord(0., cv[j++], prior);
            else { s += poisson_coord(xv[i], cv[j], prior); ++i; ++j; }
        }
        return poisson_post(s);
    }
    const smooth_t sm = smoothing(prior, d, xsum, csum);
    const double lam = sm.lhsum / (sm.lhsum + sm.rhsum), li = sm.lhinc, ri = sm.rhinc;
    double s = 0., s2 = 0., s3 = 0.;
    uint64_t i = 0, j = 0;
    while (i < xn || j < cn) {
        double x, c;
        if (j >= cn || (i < xn && xi[i] < ci[j])) { x = xv[i++]; c = 0.; }
        else if (i >= xn || ci[j] < xi[i])        { x = 0.; c = cv[j++]; }
        else                                      { x = xv[i++]; c = cv[j++]; }
        const double a = c * sm.lhrsi + li, b = x * sm.rhrsi + ri;
        switch (msr) {
        case MCO_L1: s += fabs((c + sm.pv) - (x + sm.pv)); break;
        case MCO_L2: { const double t = c - x; s += t * t; } break;
        case MCO_HELLINGER: { const double t = sqrt(a) - sqrt(b); s += t * t; } break;
        case MCO_BHATTACHARYYA_METRIC: case MCO_BHATTACHARYYA_DISTANCE: s += sqrt(a * b); break;
        case MCO_TVD: s += fabs(a - b); break;
        case MCO_COSINE_DISTANCE: case MCO_COSINE_SIMILARITY: {
            const double u = c + sm.pv, v = x + sm.pv;
            s += u * v; s2 += u * u; s3 += v * v;
        } break;
        case MCO_PROBABILITY_COSINE_DISTANCE: case MCO_PROBABILITY_COSINE_SIMILARITY:
            s += a * b; s2 += a * a; s3 += b * b; break;
        case MCO_JSM: s += coord_term(MCO_JSD, a, b); break;
        case MCO_LLR: case MCO_UWLLR: case MCO_SRLRT: case MCO_SRULRT: s += llr_term(a, b, lam); break;
        case MCO_IS: s += a / b - log(a / b); break;
        case MCO_REVERSE_IS: s += b / a - log(b / a); break;
        case MCO_SIS: s += .5 * log((a / b + b / a + 2.) * .25); break;
        case MCO_RSIS: {
            const double a2 = x * sm.rhrsi + li, b2 = c * sm.lhrsi + ri;
            s += .5 * log((a2 / b2 + b2 / a2 + 2.) * .25);
        } break;
        default: return NAN;
        }
    }
    const double z = (double)sharednz;
    switch (msr) {
    case MCO_L1: return s;
    case MCO_L2: return sqrt(s);
    case MCO_HELLINGER: { const double t = sqrt(li) - sqrt(ri); return sqrt(s + t * t * z) * 0.70710678118654752440; }
    case MCO_BHATTACHARYYA_METRIC: case MCO_BHATTACHARYYA_DISTANCE: {
        s += sqrt(li * ri) * z;
        if (1. - s < 1e-8) s = 1.;
        return msr == MCO_BHATTACHARYYA_METRIC ? sqrt(1. - s > 0. ? 1. - s : 0.) : -log(s);
    }
    case MCO_TVD: return .5 * (s + fabs(li - ri) * z);
    case MCO_COSINE_DISTANCE: case MCO_COSINE_SIMILARITY:
    case MCO_PROBABILITY_COSINE_DISTANCE: case MCO_PROBABILITY_COSINE_SIMILARITY: {
        double r = s / sqrt(s2 * s3);
        r = r > 1. ? 1. : (r < 0. ? 0. : r);
        if (msr == MCO_COSINE_DISTANCE || msr == MCO_PROBABILITY_COSINE_DISTANCE) r = acos(r) * 0.31830988618379067153;
        return r;
    }
    case MCO_JSM: {
        const double lhl = li != 0. ? log(li) : 0., rhl = ri != 0. ? log(ri) : 0.;
        const double shl = log((li + ri) * .5), shincl = (li + ri) > 0. ? (li + ri) * shl : 0.;
        return kl_tail(msr, s + ((lhl * li + rhl * ri - shincl) * z) * .5, &sm);
    }
    case MCO_LLR: case MCO_UWLLR: case MCO_SRLRT: case MCO_SRULRT: {
        const double em = lam * li + (1. - lam) * ri;
        const double ec = (li != 0. ? lam * li * log(li / em) : 0.) + (ri != 0. ? (1. - lam) * ri * log(ri / em) : 0.);
        return kl_tail(msr, s + ec * z, &sm);
    }
    case MCO_IS: return kl_tail(msr, s + z * (li / ri - log(li / ri)) - (double)d, &sm);
    case MCO_REVERSE_IS: return kl_tail(msr, s + z * (ri / li - log(ri / li)) - (double)d, &sm);
    default: return kl_tail(msr, s + z * .5 * log((li / ri + ri / li + 2.) * .25), &sm);
    }
}

/* applicator.h:1762-1897 sparse-sparse branch. */
double mco_msr_sparse(int msr, const double *xv, const uint32_t *xi, uint64_t xn,
                      const double *cv, const uint32_t *ci, uint64_t cn, uint64_t d,
                      double prior, double xsum, double csum, uint64_t *sharednz_out) {
    uint64_t sharednz = mco_shared_zeros(xi, xn, ci, cn, d);
    if (sharednz_out) *sharednz_out = sharednz;
    if (msr != MCO_SQRL2 && msr != MCO_MKL && msr != MCO_REVERSE_MKL && msr != MCO_JSD)
        return msr_more_sparse(msr, xv, xi, xn, cv, ci, cn, d, prior, xsum, csum, sharednz);
    uint64_t i = 0, j = 0;
    if (msr == MCO_SQRL2) {                        /* :1797-1806 */
        double s = 0.;
        while (i < xn || j < cn) {
            double xval, cval;
            if (j >= cn || (i < xn && xi[i] < ci[j])) { xval = xv[i++]; cval = 0.; }
            else if (i >= xn || ci[j] < xi[i])        { xval = 0.; cval = cv[j++]; }
            else                                      { xval = xv[i++]; cval = cv[j++]; }
            double t = cval - xval; s += t * t;
        }
        if (isnan(s)) return -1.;
        return s < 0. ? 0. : s;
    }
    smooth_t sm = smoothing(prior, d, xsum, csum);
    double klc = 0.;
    while (i < xn || j < cn) {
        double xval, cval;
        if (j >= cn || (i < xn && xi[i] < ci[j])) { xval = xv[i++]; cval = 0.; }
        else if (i >= xn || ci[j] < xi[i])        { xval = 0.; cval = cv[j++]; }
        else                                      { xval = xv[i++]; cval = cv[j++]; }
        klc += coord_term(msr, cval * sm.lhrsi + sm.lhinc, xval * sm.rhrsi + sm.rhinc);
    }
    double lhl = sm.lhinc != 0. ? log(sm.lhinc) : 0., rhl = sm.rhinc != 0. ? log(sm.rhinc) : 0.;
    double zc;
    if (msr == MCO_MKL)              zc = (double)sharednz * sm.lhinc * (lhl - rhl);   /* :1847-1849 */
    else if (msr == MCO_REVERSE_MKL) zc = (double)sharednz * sm.rhinc * (rhl - lhl);   /* :1853-1855 */
    else {                                                                              /* :1850-1852 */
        double shl = log((sm.lhinc + sm.rhinc) * .5);
        double shincl = (sm.lhinc + sm.rhinc) > 0. ? (sm.lhinc + sm.rhinc) * shl : 0.;
        zc = ((lhl * sm.lhinc + rhl * sm.rhinc - shincl) * (double)sharednz) * .5;
    }
    return post_kl(klc + zc);
}

float mco_sqrl2_f32(const float *x, const float *c, uint64_t d) {
    float s = 0.f;
    for (uint64_t j = 0; j < d; ++j) { float t = x[j] - c[j]; s = fmaf(t, t, s); }
    return s;
}

/* Cost of CSR row i against a dense center, walking all d coordinates. */
static double csr_row_cost(int msr, const float *val, const uint32_t *idx, uint64_t b, uint64_t e,
                           uint64_t d, const float *c, double prior, double xsum, double csum) {
    if (msr != MCO_SQRL2 && msr != MCO_MKL && msr != MCO_REVERSE_MKL && msr != MCO_JSD) {
        double *xb = (double *)calloc(d, sizeof(double)), *cb = (double *)malloc(d * sizeof(double));
        for (uint64_t p = b; p < e; ++p) xb[idx[p]] = val[p];
        for (uint64_t j = 0; j < d; ++j) cb[j] = c[j];
        const double v = msr_more_dense(msr, xb, cb, d, prior, xsum, csum);
        free(xb); free(cb);
        return v;
    }
    if (msr == MCO_SQRL2) {
        double s = 0.; uint64_t p = b;
        for (uint64_t j = 0; j < d; ++j) {
            double xv = (p < e && idx[p] == j) ? (double)val[p++] : 0.;
            double t = (double)c[j] - xv; s += t * t;
        }
        return s;
    }
    smooth_t sm = smoothing(prior, d, xsum, csum);
    double r = 0.; uint64_t p = b;
    for (uint64_t j = 0; j < d; ++j) {
        double xv = (p < e && idx[p] == j) ? (double)val[p++] : 0.;
        r += coord_term(msr, (double)c[j] * sm.lhrsi + sm.lhinc, xv * sm.rhrsi + sm.rhinc);
    }
    return post_kl(r);
}

static double row_sum_dense(const float *x, uint64_t d) {
    double s = 0.; for (uint64_t j = 0; j < d; ++j) s += x[j]; return s;
}
static double row_sum_csr(const float *val, uint64_t b, uint64_t e) {
    double s = 0.; for (uint64_t p = b; p < e; ++p) s += val[p]; return s;
}

static double pair_cost(int msr, const float *X, const float *val, const uint32_t *idx,
                        const uint64_t *indptr, uint64_t i, uint64_t d, const float *c,
                        double prior, double xsum, double csum, double *xbuf, double *cbuf) {
    if (val) return csr_row_cost(msr, val, idx, indptr[i], indptr[i + 1], d, c, prior, xsum, csum);
    for (uint64_t j = 0; j < d; ++j) { xbuf[j] = X[i * d + j]; cbuf[j] = c[j]; }
    return mco_msr_dense(msr, xbuf, cbuf, d, prior, xsum, csum);
}

void mco_cost_matrix_dense(int msr, const float *X, uint64_t n, uint64_t d,
                           const float *C, uint64_t k, double prior, double *out) {
    double *cs = (double *)malloc(k * sizeof(double));
    for (uint64_t c = 0; c < k; ++c) cs[c] = row_sum_dense(C + c * d, d);
    #pragma omp parallel
    {
        double *xb = (double *)malloc(d * sizeof(double)), *cb = (double *)malloc(d * sizeof(double));
        #pragma omp for schedule(static)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            double xs = row_sum_dense(X + i * d, d);
            for (uint64_t c = 0; c < k; ++c)
                out[i * k + c] = pair_cost(msr, X, NULL, NULL, NULL, i, d, C + c * d, prior, xs, cs[c], xb, cb);
        }
        free(xb); free(cb);
    }
    free(cs);
}

void mco_cost_matrix_csr(int msr, const float *val, const uint32_t *idx, const uint64_t *indptr,
                         uint64_t n, uint64_t d, const float *C, uint64_t k, double prior, double *out) {
    double *cs = (double *)malloc(k * sizeof(double));
    for (uint64_t c = 0; c < k; ++c) cs[c] = row_sum_dense(C + c * d, d);
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        double xs = row_sum_csr(val, indptr[i], indptr[i + 1]);
        for (uint64_t c = 0; c < k; ++c)
            out[i * k + c] = csr_row_cost(msr, val, idx, indptr[i], indptr[i + 1], d, C + c * d, prior, xs, cs[c]);
    }
    free(cs);
}

/* ======================================================================
 * Hard assignment: solve.h:234-285 (start from center 0, replace on strict <).
 * ====================================================================== */
void mco_assign_sqrl2_f32(const float *X, uint64_t n, uint64_t d, const float *C, uint64_t k,
                          uint32_t *asn, float *cost, double *gap_out, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float *x = X + (uint64_t)i * d;
        float best = mco_sqrl2_f32(x, C, d), second = INFINITY;
        uint32_t bi = 0;
        for (uint64_t c = 1; c < k; ++c) {
            float v = mco_sqrl2_f32(x, C + c * d, d);
            if (v < best) { second = best; best = v; bi = (uint32_t)c; }
            else if (v < second) second = v;
        }
        asn[i] = bi; cost[i] = best;
        if (gap_out) gap_out[i] = ((double)second - (double)best) / ((double)best > 0. ? (double)best : 1.);
    }
}

void mco_assign_f64(int msr, const float *X, const float *val, const uint32_t *idx,
                    const uint64_t *indptr, uint64_t n, uint64_t d, const float *C, uint64_t k,
                    double prior, uint32_t *asn, double *cost, double *gap_out) {
    double *cs = (double *)malloc(k * sizeof(double));
    for (uint64_t c = 0; c < k; ++c) cs[c] = row_sum_dense(C + c * d, d);
    #pragma omp parallel
    {
        double *xb = (double *)malloc(d * sizeof(double)), *cb = (double *)malloc(d * sizeof(double));
        #pragma omp for schedule(dynamic, 16)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            double xs = val ? row_sum_csr(val, indptr[i], indptr[i + 1]) : row_sum_dense(X + i * d, d);
            double best = pair_cost(msr, X, val, idx, indptr, i, d, C, prior, xs, cs[0], xb, cb), second = INFINITY;
            uint32_t bi = 0;
            for (uint64_t c = 1; c < k; ++c) {
                double v = pair_cost(msr, X, val, idx, indptr, i, d, C + c * d, prior, xs, cs[c], xb, cb);
                if (v < best) { second = best; best = v; bi = (uint32_t)c; }
                else if (v < second) second = v;
            }
            asn[i] = bi; cost[i] = best;
            if (gap_out) gap_out[i] = (second - best) / (best > 0. ? best : 1.);
        }
        free(xb); free(cb);
    }
    free(cs);
}

/* ======================================================================
 * M-step (hard): centroid.h:137-178 (dense mean), 81-112 (CSR mean).  Sums
 * are FP64 in a fixed order: members in ascending index, `chunk` members
 * summed sequentially (w*x as one fma when weighted), chunk sums added in
 * order; mean = sum / total weight, rounded once to float.
 * ====================================================================== */
uint64_t mco_centroids_dense(const float *X, uint64_t n, uint64_t d, const uint32_t *asn,
                             const double *w, uint64_t k, uint64_t chunk, float *C) {
    uint64_t *cnt = (uint64_t *)calloc(k + 1, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) cnt[asn[i] + 1]++;
    for (uint64_t c = 0; c < k; ++c) cnt[c + 1] += cnt[c];
    uint64_t *perm = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    uint64_t *cur = (uint64_t *)malloc(k * sizeof(uint64_t));
    memcpy(cur, cnt, k * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) perm[cur[asn[i]]++] = i;
    uint64_t empty = 0;
    if (chunk == 0) chunk = MCO_MEAN_CHUNK;
    #pragma omp parallel for schedule(dynamic, 1) reduction(+:empty)
    for (int64_t c = 0; c < (int64_t)k; ++c) {
        uint64_t b = cnt[c], e = cnt[c + 1];
        if (b == e) { ++empty; continue; }
        double *tot = (double *)calloc(d, sizeof(double)), *part = (double *)malloc(d * sizeof(double));
        double wt = 0., wpart;
        for (uint64_t s = b; s < e; s += chunk) {
            uint64_t se = s + chunk < e ? s + chunk : e;
            memset(part, 0, d * sizeof(double)); wpart = 0.;
            for (uint64_t p = s; p < se; ++p) {
                const float *x = X + perm[p] * d;
                double wi = w ? w[perm[p]] : 1.;
                for (uint64_t j = 0; j < d; ++j) part[j] = w ? fma(wi, (double)x[j], part[j]) : part[j] + (double)x[j];
                wpart += wi;
            }
            for (uint64_t j = 0; j < d; ++j) tot[j] += part[j];
            wt += wpart;
        }
        for (uint64_t j = 0; j < d; ++j) C[c * d + j] = (float)(tot[j] / wt);
        free(tot); free(part);
    }
    free(cnt); free(perm); free(cur);
    return empty;
}

/* One pass of lloyd_iteration (kmeans.h:359-479) on dense rows, squared L2, in the
 * order E-step then M-step: assign every row to its nearest center (FP32 FMA chain,
 * lowest index on ties), *cost = sum w_i ||x_i - c_asn(i)||^2 against the input
 * centers (FP64), then the weighted means (chunks of MCO_MEAN_CHUNK members).  An empty
 * cluster (ascending id) takes a row drawn with probability ~ w_j * cost_j, cost_j
 * the FP32 cost of row j to its new mean (kmeans.h:433-440), by the fixed-point
 * sampler with output e of the wyrand stream of 0 for the e-th draw of the call;
 * the drawn row's weight is zero for the later draws, it moves to the cluster and
 * becomes its center, and the means are formed again (from scratch) until no
 * cluster is empty (kmeans.h:431-454).  Returns the number of reseeds. */
uint64_t mco_lloyd_step(const float *X, uint64_t n, uint64_t d, float *C, uint64_t k, const double *w,
                        uint32_t *asn, double *cost) {
    float *pc = (float *)malloc(n * sizeof(float));
    mco_assign_sqrl2_f32(X, n, d, C, k, asn, pc, NULL, 0);
    double tot = 0.;
    for (uint64_t i = 0; i < n; ++i) {
        const float *x = X + i * d, *c = C + (uint64_t)asn[i] * d;
        double s = 0.;
        for (uint64_t j = 0; j < d; ++j) { double t = (double)x[j] - (double)c[j]; s += t * t; }
        tot += (w ? w[i] : 1.) * s;
    }
    *cost = tot;
    uint64_t draws = 0;
    double *wv = (double *)malloc(n * sizeof(double));
    uint64_t *fw = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t *cnt = (uint64_t *)malloc(k * sizeof(uint64_t));
    for (uint64_t round = 0; round <= k; ++round) {
        uint64_t empty = mco_centroids_dense(X, n, d, asn, w, k, MCO_MEAN_CHUNK, C);
        if (!empty) break;
        memset(cnt, 0, k * sizeof(uint64_t));
        for (uint64_t i = 0; i < n; ++i) cnt[asn[i]]++;
        for (uint64_t i = 0; i < n; ++i)
            wv[i] = (w ? w[i] : 1.) * (double)mco_sqrl2_f32(X + i * d, C + (uint64_t)asn[i] * d, d);
        for (uint64_t q = 0; q < k; ++q) {
            if (cnt[q]) continue;
            mco_fixed_weights(wv, n, fw);
            const uint64_t r = mco_fixed_draw(fw, n, mco_wyrand_at(0, draws++));
            wv[r] = 0.;
            asn[r] = (uint32_t)q;
            memcpy(C + q * d, X + r * d, d * sizeof(float));
        }
    }
    free(pc); free(wv); free(fw); free(cnt);
    return draws;
}

uint64_t mco_centroids_csr(const float *val, const uint32_t *idx, const uint64_t *indptr,
                           uint64_t n, uint64_t d, const uint32_t *asn, const double *w,
                           uint64_t k, float *C) {
    double *acc = (double *)calloc(k * d, sizeof(double));
    double *wt = (double *)calloc(k, sizeof(double));
    uint64_t *cnt = (uint64_t *)calloc(k, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t c = asn[i];
        double wi = w ? w[i] : 1.;
        for (uint64_t p = indptr[i]; p < indptr[i + 1]; ++p) acc[c * d + idx[p]] += wi * (double)val[p];
        wt[c] += wi; cnt[c]++;
    }
    uint64_t empty = 0;
    for (uint64_t c = 0; c < k; ++c) {
        if (!cnt[c]) { ++empty; continue; }
        for (uint64_t j = 0; j < d; ++j) C[c * d + j] = (float)(acc[c * d + j] / wt[c]);
    }
    free(acc); free(wt); free(cnt);
    return empty;
}

/* Per-point cost for the Lloyd control flow: SQRL2 on dense rows uses the
 * FP32 sequential-FMA value (what the GPU library reports); other cases FP64. */
static void assign_generic(int msr, const float *X, const float *val, const uint32_t *idx,
                           const uint64_t *indptr, uint64_t n, uint64_t d, const float *C, uint64_t k,
                           double prior, uint32_t *asn, double *cost, int nthreads) {
    if (msr == MCO_SQRL2 && !val) {
        float *cf = (float *)malloc((n ? n : 1) * sizeof(float));
        mco_assign_sqrl2_f32(X, n, d, C, k, asn, cf, NULL, nthreads);
        for (uint64_t i = 0; i < n; ++i) cost[i] = cf[i];
        free(cf);
    } else {
        mco_assign_f64(msr, X, val, idx, indptr, n, d, C, k, prior, asn, cost, NULL);
    }
}

static double total_cost(const double *cost, const double *w, uint64_t n) {
    double s = 0.;
    for (uint64_t i = 0; i < n; ++i) s += w ? w[i] * cost[i] : cost[i];
    return s;
}

static int has_nan(const float *C, uint64_t m) {
    for (uint64_t i = 0; i < m; ++i) if (isnan(C[i])) return 1;
    return 0;
}

/* set_centroids_full_mean (hard) with the empty-cluster restart of
 * centroid.h:687-752: each empty cluster (ascending id) takes the row drawn
 * with probability proportional to its cost (RNG seeded by n), all points
 * are re-assigned, and each drawn row is forced onto its new center. */
static int centroids_with_restart(int msr, const float *X, const float *val, const uint32_t *idx,
                                  const uint64_t *indptr, uint64_t n, uint64_t d, float *C, uint64_t k,
                                  double prior, const double *w, uint32_t *asn, double *cost, int nthreads) {
    uint64_t *cnt = (uint64_t *)calloc(k, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) cnt[asn[i]]++;
    int restarted = 0; uint64_t draw = 0;
    uint32_t *rs = NULL, *sa = NULL; uint64_t ne = 0;
    for (uint64_t c = 0; c < k; ++c) if (!cnt[c]) ++ne;
    if (ne) {
        restarted = 1;
        rs = (uint32_t *)malloc(ne * sizeof(uint32_t)); sa = (uint32_t *)malloc(ne * sizeof(uint32_t));
        uint64_t *fw = (uint64_t *)malloc(n * sizeof(uint64_t));
        mco_fixed_weights(cost, n, fw);
        uint64_t e = 0;
        for (uint64_t c = 0; c < k; ++c) if (!cnt[c]) {
            uint64_t r = mco_fixed_draw(fw, n, mco_wyrand_at(n, draw++));
            rs[e] = (uint32_t)r; sa[e] = (uint32_t)c; ++e;
            if (val) {
                memset(C + c * d, 0, d * sizeof(float));
                for (uint64_t p = indptr[r]; p < indptr[r + 1]; ++p) C[c * d + idx[p]] = val[p];
            } else memcpy(C + c * d, X + r * d, d * sizeof(float));
        }
        free(fw);
        assign_generic(msr, X, val, idx, indptr, n, d, C, k, prior, asn, cost, nthreads);
        for (uint64_t i = 0; i < ne; ++i) if (asn[rs[i]] != sa[i]) { asn[rs[i]] = sa[i]; cost[rs[i]] = 0.; }
        free(rs); free(sa);
    }
    free(cnt);
    if (val) mco_centroids_csr(val, idx, indptr, n, d, asn, w, k, C);
    else     mco_centroids_dense(X, n, d, asn, w, k, MCO_MEAN_CHUNK, C);
    return restarted;
}

/* perform_hard_clustering, solve.h:98-175.  FT is float for float data, so
 * every cost in the control flow is rounded to float (solve.h:120-123). */
static int median_policy(int msr);

int mco_lloyd(int msr, const float *X, const float *val, const uint32_t *idx,
              const uint64_t *indptr, uint64_t n, uint64_t d, float *C, uint64_t k,
              double prior, const double *w, double eps, uint64_t maxiter,
              uint32_t *asn, double *cost, double *initcost, double *finalcost,
              uint64_t *numiter, int nthreads) {
    if (median_policy(msr))   /* L1, TVD, L2: the median policies (solve.h:207-215) */
        return mco_lloyd_median(msr, X, val, idx, indptr, n, d, C, k, prior, w, eps, maxiter, asn, cost, initcost,
                                finalcost, numiter);
    assign_generic(msr, X, val, idx, indptr, n, d, C, k, prior, asn, cost, nthreads);
    float fcost = (float)total_cost(cost, w, n);
    *initcost = fcost;
    if (fcost == 0.f) { *initcost = 0.; *finalcost = 0.; *numiter = 0; return 0; }
    float *Cc = (float *)malloc(k * d * sizeof(float));
    memcpy(Cc, C, k * d * sizeof(float));
    uint64_t iter = 0;
    int rc = 0;
    for (;;) {
        int res = centroids_with_restart(msr, X, val, idx, indptr, n, d, Cc, k, prior, w, asn, cost, nthreads);
        if (has_nan(Cc, k * d)) { rc = -1; break; }
        assign_generic(msr, X, val, idx, indptr, n, d, Cc, k, prior, asn, cost, nthreads);
        float newcost = (float)total_cost(cost, w, n);
        if (newcost > fcost && !res) {
            assign_generic(msr, X, val, idx, indptr, n, d, C, k, prior, asn, cost, nthreads);
            break;
        }
        memcpy(C, Cc, k * d * sizeof(float));
        ++iter;
        float old = fcost; fcost = newcost;
        float diff = old - newcost;
        if ((double)diff < eps * ((double)newcost > (double)old ? (double)newcost : (double)old) || iter > maxiter) break;
    }
    free(Cc);
    *finalcost = fcost; *numiter = iter;
    return rc;
}

/* ======================================================================
 * Minibatch hard clustering: perform_hard_minibatch_clustering,
 * solve.h:377-578, for the mean-centroid measures (set_center,
 * centroid.h:81-178) and the median policies (solve.h:555-572: l1_median with
 * the weight vector indexed by member rank, tvd_median likewise, the geometric
 * median with the members' weights; clusters without members keep their center).  One RNG stream, the wyrand counter stream of `seed`,
 * serves the restart draws and the minibatch draws in the reference's order.
 * A check-in (iteration 0, every freq-th, and the last) assigns every row,
 * restarts each center with <= reseed_after rows (ascending id) at a row
 * drawn by the fixed-point sampler over w_i*cost_i, moves rows to a
 * restarted center that is strictly cheaper (restarted centers in order),
 * and keeps the centers of the lowest check-in cost (at iteration 0 the
 * saved centers stay the input ones, solve.h:494-499).  Between check-ins,
 * mbsize rows are drawn (rng() mod n; without replacement the first mbsize
 * distinct draws), sorted ascending (solve.h:549-551, duplicates kept), each
 * takes its nearest center, and every center with members becomes their
 * weighted mean.  The check-in cost is the exact fixed-point sum of
 * w_i*cost_i (the shift of the sampler), independent of summation order.
 * ====================================================================== */
static void median_core(int pol, const float *X, const float *val, const uint32_t *idx, const uint64_t *indptr,
                        uint64_t n, uint64_t d, const double *wm, const double *wr, uint64_t k, const uint32_t *asn,
                        float *C, double eps, int single_raw, int rank_l1);

static double point_cost(int msr, const float *X, const float *val, const uint32_t *idx, const uint64_t *indptr,
                         uint64_t i, uint64_t d, const float *c, double prior) {
    if (msr == MCO_SQRL2 && !val) return (double)mco_sqrl2_f32(X + i * d, c, d);
    const double xs = val ? row_sum_csr(val, indptr[i], indptr[i + 1]) : row_sum_dense(X + i * d, d);
    const double cs = row_sum_dense(c, d);
    double *xb = (double *)malloc(d * sizeof(double)), *cb = (double *)malloc(d * sizeof(double));
    const double v = pair_cost(msr, X, val, idx, indptr, i, d, c, prior, xs, cs, xb, cb);
    free(xb); free(cb);
    return v;
}

static double fixed_total(const double *cost, const double *w, uint64_t n) {
    double mx = 0.;
    double *t = (double *)malloc((n ? n : 1) * sizeof(double));
    for (uint64_t i = 0; i < n; ++i) { t[i] = w ? w[i] * cost[i] : cost[i]; if (fabs(t[i]) > mx) mx = fabs(t[i]); }
    const int s = mco_fixed_shift(mx, n);
    uint64_t T = 0;
    for (uint64_t i = 0; i < n; ++i) if (t[i] > 0.) T += (uint64_t)floor(ldexp(t[i], s));
    free(t);
    return ldexp((double)T, -s);
}

static int cmp_u64(const void *a, const void *b) {
    const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

int mco_minibatch(int msr, const float *X, const float *val, const uint32_t *idx, const uint64_t *indptr,
                  uint64_t n, uint64_t d, float *C, uint64_t k, double prior, const double *w,
                  uint64_t mbsize, uint64_t maxiter, uint64_t freq, uint64_t reseed_after, int with_rep,
                  uint64_t seed, uint32_t *asn, double *cost, double *initcost, double *finalcost,
                  uint64_t *numiter, int nthreads) {
    if (!mbsize || !maxiter || !freq || (!with_rep && mbsize > n)) return -2;
    {   /* median updates (solve.h:555-572): CSR only unweighted; weights by rank need mbsize <= n */
        const int pol = median_policy(msr);
        if (pol && val && w) return -2;
        if (pol && w && pol != 3 && mbsize > n) return -2;
    }
    uint64_t j = 0, iter = 0;
    double init = DBL_MAX, best = DBL_MAX;
    float *save = (float *)malloc(k * d * sizeof(float));
    memcpy(save, C, k * d * sizeof(float));
    uint64_t *sel = (uint64_t *)malloc(mbsize * sizeof(uint64_t));
    uint32_t *sa = (uint32_t *)malloc(mbsize * sizeof(uint32_t));
    double *sc = (double *)malloc(mbsize * sizeof(double)), *sw = w ? (double *)malloc(mbsize * sizeof(double)) : NULL;
    uint8_t *seen = with_rep ? NULL : (uint8_t *)malloc(n);
    uint64_t *cnt = (uint64_t *)malloc(k * sizeof(uint64_t));
    float *Xs = NULL, *vs = NULL; uint32_t *is = NULL; uint64_t *ps = NULL;
    for (;;) {
        if (iter % freq == 0 || iter == maxiter - 1) {
            assign_generic(msr, X, val, idx, indptr, n, d, C, k, prior, asn, cost, nthreads);
            memset(cnt, 0, k * sizeof(uint64_t));
            for (uint64_t i = 0; i < n; ++i) cnt[asn[i]]++;
            uint64_t nf = 0;
            for (uint64_t c = 0; c < k; ++c) if (cnt[c] <= reseed_after) ++nf;
            if (nf) {
                uint32_t *found = (uint32_t *)malloc(nf * sizeof(uint32_t));
                double *wv = (double *)malloc(n * sizeof(double));
                uint64_t *fw = (uint64_t *)malloc(n * sizeof(uint64_t));
                for (uint64_t i = 0; i < n; ++i) wv[i] = w ? w[i] * cost[i] : cost[i];
                mco_fixed_weights(wv, n, fw);
                uint64_t f = 0;
                for (uint64_t c = 0; c < k; ++c) if (cnt[c] <= reseed_after) {
                    const uint64_t r = mco_fixed_draw(fw, n, mco_wyrand_at(seed, j++));
                    found[f++] = (uint32_t)c;
                    if (val) {
                        memset(C + c * d, 0, d * sizeof(float));
                        for (uint64_t p = indptr[r]; p < indptr[r + 1]; ++p) C[c * d + idx[p]] = val[p];
                    } else memcpy(C + c * d, X + r * d, d * sizeof(float));
                }
                #pragma omp parallel for schedule(dynamic, 64)
                for (int64_t i = 0; i < (int64_t)n; ++i)
                    for (uint64_t q = 0; q < nf; ++q) {
                        const double v = point_cost(msr, X, val, idx, indptr, i, d, C + (uint64_t)found[q] * d, prior);
                        if (v < cost[i]) { cost[i] = v; asn[i] = found[q]; }
                    }
                free(found); free(wv); free(fw);
            }
            const double ct = fixed_total(cost, w, n);
            if (iter == 0) init = best = ct;
            if (ct < best) { best = ct; memcpy(save, C, k * d * sizeof(float)); }
        }
        if (++iter == maxiter) break;
        if (with_rep) {
            for (uint64_t t = 0; t < mbsize; ++t) sel[t] = mco_wyrand_at(seed, j++) % n;
        } else {
            memset(seen, 0, n);
            for (uint64_t t = 0; t < mbsize;) {
                const uint64_t r = mco_wyrand_at(seed, j++) % n;
                if (!seen[r]) { seen[r] = 1; sel[t++] = r; }
            }
        }
        qsort(sel, mbsize, sizeof(uint64_t), cmp_u64);
        if (sw) for (uint64_t t = 0; t < mbsize; ++t) sw[t] = w[sel[t]];
        if (val) {
            free(vs); free(is); free(ps);
            ps = (uint64_t *)malloc((mbsize + 1) * sizeof(uint64_t));
            ps[0] = 0;
            for (uint64_t t = 0; t < mbsize; ++t) ps[t + 1] = ps[t] + (indptr[sel[t] + 1] - indptr[sel[t]]);
            vs = (float *)malloc((ps[mbsize] ? ps[mbsize] : 1) * sizeof(float));
            is = (uint32_t *)malloc((ps[mbsize] ? ps[mbsize] : 1) * sizeof(uint32_t));
            for (uint64_t t = 0; t < mbsize; ++t) {
                memcpy(vs + ps[t], val + indptr[sel[t]], (ps[t + 1] - ps[t]) * sizeof(float));
                memcpy(is + ps[t], idx + indptr[sel[t]], (ps[t + 1] - ps[t]) * sizeof(uint32_t));
            }
            assign_generic(msr, NULL, vs, is, ps, mbsize, d, C, k, prior, sa, sc, nthreads);
            if (median_policy(msr)) median_core(median_policy(msr), NULL, vs, is, ps, mbsize, d, NULL, NULL, k, sa, C,
                                                0., 0, 1);
            else mco_centroids_csr(vs, is, ps, mbsize, d, sa, sw, k, C);
        } else {
            if (!Xs) Xs = (float *)malloc(mbsize * d * sizeof(float));
            for (uint64_t t = 0; t < mbsize; ++t) memcpy(Xs + t * d, X + sel[t] * d, d * sizeof(float));
            assign_generic(msr, Xs, NULL, NULL, NULL, mbsize, d, C, k, prior, sa, sc, nthreads);
            if (median_policy(msr)) median_core(median_policy(msr), Xs, NULL, NULL, NULL, mbsize, d, sw, w, k, sa, C,
                                                0., 0, 1);
            else mco_centroids_dense(Xs, mbsize, d, sa, sw, k, MCO_MEAN_CHUNK, C);
        }
    }
    memcpy(C, save, k * d * sizeof(float));
    *initcost = init; *finalcost = best; *numiter = iter;
    free(save); free(sel); free(sa); free(sc); free(sw); free(seen); free(cnt); free(Xs); free(vs); free(is); free(ps);
    return 0;
}

/* ======================================================================
 * Coreset-minibatch hard clustering: hmb_coreset_clustering, solve.h:598-789.
 * Outer rounds: ceil(maxiter / freq) + 1 check-ins (the check-in of the
 * minibatch driver above; the saved centers change only when a later
 * check-in is strictly cheaper).  After each check-in but the last: LBK
 * sensitivities from the check-in's costs and assignment (coreset.h:487-533,
 * make_sampler without an alias sampler), then freq times: a coreset of
 * m = min(mbsize, n) i.i.d. draws with the stream's next output as the draw
 * seed (reservoir_simd::sample_k WITH_REPLACEMENT, sorted ascending), weights
 * w_i/(m p_i) times w_i again (solve.h:771-772 multiplies the row weights into
 * the coreset weights, which already hold them), and perform_hard_clustering
 * on the coreset rows from the current centers with eps 1e-3 and maxiter
 * 1 + 2*(round < 2) (solve.h:768).
 * ====================================================================== */
static int cmp_pair_u64(const void *a, const void *b) {
    const uint64_t x = ((const uint64_t *)a)[0], y = ((const uint64_t *)b)[0];
    return x < y ? -1 : x > y;
}

int mco_cs_minibatch(int msr, const float *X, const float *val, const uint32_t *idx, const uint64_t *indptr,
                     uint64_t n, uint64_t d, float *C, uint64_t k, double prior, const double *w,
                     uint64_t mbsize, uint64_t maxiter, uint64_t freq, uint64_t reseed_after,
                     uint64_t seed, uint32_t *asn, double *cost, double *initcost, double *finalcost,
                     uint64_t *numiter, int nthreads) {
    if (!mbsize || !freq) return -2;
    const uint64_t rounds = (freq - 1 + maxiter) / freq, m = mbsize < n ? mbsize : n;
    uint64_t j = 0, iter = 0;
    double init = DBL_MAX, best = DBL_MAX;
    int rc = 0;
    float *save = (float *)malloc(k * d * sizeof(float));
    memcpy(save, C, k * d * sizeof(float));
    uint64_t *cnt = (uint64_t *)malloc(k * sizeof(uint64_t));
    float *probs = (float *)malloc(n * sizeof(float));
    uint64_t *ind = (uint64_t *)malloc(m * sizeof(uint64_t)), *pairs = (uint64_t *)malloc(2 * m * sizeof(uint64_t));
    float *wt = (float *)malloc(m * sizeof(float));
    double *csw = (double *)malloc(m * sizeof(double)), *sc = (double *)malloc(m * sizeof(double));
    uint32_t *sa = (uint32_t *)malloc(m * sizeof(uint32_t));
    float *Xs = val ? NULL : (float *)malloc(m * d * sizeof(float));
    for (;;) {
        assign_generic(msr, X, val, idx, indptr, n, d, C, k, prior, asn, cost, nthreads);
        memset(cnt, 0, k * sizeof(uint64_t));
        for (uint64_t i = 0; i < n; ++i) cnt[asn[i]]++;
        uint64_t nf = 0;
        for (uint64_t c = 0; c < k; ++c) if (cnt[c] <= reseed_after) ++nf;
        if (nf) {
            uint32_t *found = (uint32_t *)malloc(nf * sizeof(uint32_t));
            double *wv = (double *)malloc(n * sizeof(double));
            uint64_t *fw = (uint64_t *)malloc(n * sizeof(uint64_t));
            for (uint64_t i = 0; i < n; ++i) wv[i] = w ? w[i] * cost[i] : cost[i];
            mco_fixed_weights(wv, n, fw);
            uint64_t f = 0;
            for (uint64_t c = 0; c < k; ++c) if (cnt[c] <= reseed_after) {
                const uint64_t r = mco_fixed_draw(fw, n, mco_wyrand_at(seed, j++));
                found[f++] = (uint32_t)c;
                if (val) {
                    memset(C + c * d, 0, d * sizeof(float));
                    for (uint64_t p = indptr[r]; p < indptr[r + 1]; ++p) C[c * d + idx[p]] = val[p];
                } else memcpy(C + c * d, X + r * d, d * sizeof(float));
            }
            #pragma omp parallel for schedule(dynamic, 64)
            for (int64_t i = 0; i < (int64_t)n; ++i)
                for (uint64_t q = 0; q < nf; ++q) {
                    const double v = point_cost(msr, X, val, idx, indptr, i, d, C + (uint64_t)found[q] * d, prior);
                    if (v < cost[i]) { cost[i] = v; asn[i] = found[q]; }
                }
            free(found); free(wv); free(fw);
        }
        const double ct = fixed_total(cost, w, n);
        if (iter == 0) init = best = ct;
        else if (ct < best) { best = ct; memcpy(save, C, k * d * sizeof(float)); }
        if (++iter > rounds) break;
        if (mco_coreset_probs(4, n, k, cost, asn, w, probs) != 0) { rc = -3; break; }
        for (uint64_t jj = 0; jj < freq && !rc; ++jj) {
            const uint64_t r = mco_wyrand_at(seed, j++);
            mco_coreset_sample(n, probs, w, m, r, 0, ind, wt);
            for (uint64_t t = 0; t < m; ++t) { pairs[2 * t] = ind[t]; pairs[2 * t + 1] = t; }
            qsort(pairs, m, 2 * sizeof(uint64_t), cmp_pair_u64);
            for (uint64_t t = 0; t < m; ++t) {
                const uint64_t i = pairs[2 * t];
                csw[t] = (w ? w[i] : 1.) * (double)wt[pairs[2 * t + 1]];
                ind[t] = i;
            }
            double ic, fc; uint64_t it;
            const uint64_t sub = 1 + 2 * (iter < 2);
            if (val) {
                uint64_t *ps = (uint64_t *)malloc((m + 1) * sizeof(uint64_t));
                ps[0] = 0;
                for (uint64_t t = 0; t < m; ++t) ps[t + 1] = ps[t] + (indptr[ind[t] + 1] - indptr[ind[t]]);
                float *vs = (float *)malloc((ps[m] ? ps[m] : 1) * sizeof(float));
                uint32_t *is = (uint32_t *)malloc((ps[m] ? ps[m] : 1) * sizeof(uint32_t));
                for (uint64_t t = 0; t < m; ++t) {
                    memcpy(vs + ps[t], val + indptr[ind[t]], (ps[t + 1] - ps[t]) * sizeof(float));
                    memcpy(is + ps[t], idx + indptr[ind[t]], (ps[t + 1] - ps[t]) * sizeof(uint32_t));
                }
                if (mco_lloyd(msr, NULL, vs, is, ps, m, d, C, k, prior, csw, 1e-3, sub, sa, sc, &ic, &fc, &it,
                              nthreads)) rc = -1;
                free(ps); free(vs); free(is);
            } else {
                for (uint64_t t = 0; t < m; ++t) memcpy(Xs + t * d, X + ind[t] * d, d * sizeof(float));
                if (mco_lloyd(msr, Xs, NULL, NULL, NULL, m, d, C, k, prior, csw, 1e-3, sub, sa, sc, &ic, &fc, &it,
                              nthreads)) rc = -1;
            }
        }
        if (rc) break;
    }
    memcpy(C, save, k * d * sizeof(float));
    *initcost = init; *finalcost = best; *numiter = iter;
    free(save); free(cnt); free(probs); free(ind); free(pairs); free(wt); free(csw); free(sc); free(sa); free(Xs);
    return rc;
}

/* ======================================================================
 * Soft clustering.  centroid.h:781-793 (correct_softmax) and 795-963;
 * solve.h:287-374.  softmax is exp(-temp*cost) / row sum in float without
 * max subtraction; a row with any NaN becomes uniform over its argmin ties.
 * ====================================================================== */
void mco_softmax_rows(const float *costs, uint64_t n, uint64_t k, double temp, float *resp) {
    float nt = (float)(-temp);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float *cr = costs + i * k; float *r = resp + i * k;
        float s = 0.f;
        for (uint64_t c = 0; c < k; ++c) { r[c] = expf(cr[c] * nt); s += r[c]; }
        int bad = 0;
        for (uint64_t c = 0; c < k; ++c) { r[c] = r[c] / s; if (isnan(r[c])) bad = 1; }
        if (bad) {
            uint64_t b = 0; for (uint64_t c = 1; c < k; ++c) if (cr[c] < cr[b]) b = c;
            uint64_t ties = 0; for (uint64_t c = 0; c < k; ++c) if (cr[c] == cr[b]) ++ties;
            float per = (float)(1. / (double)ties);
            for (uint64_t c = 0; c < k; ++c) r[c] = (cr[c] == cr[b]) ? per : 0.f;
        }
    }
}

static void soft_costs(int msr, const float *X, const float *val, const uint32_t *idx, const uint64_t *indptr,
                       uint64_t n, uint64_t d, const float *C, uint64_t k, double prior, float *costs) {
    double *tmp = (double *)malloc(n * k * sizeof(double));
    if (val) mco_cost_matrix_csr(msr, val, idx, indptr, n, d, C, k, prior, tmp);
    else     mco_cost_matrix_dense(msr, X, n, d, C, k, prior, tmp);
    for (uint64_t i = 0; i < n * k; ++i) costs[i] = (float)tmp[i];
    free(tmp);
}

/* ======================================================================
 * Soft M-step for L1 and L2 (set_centroids_full_mean, centroid.h:815-838): per
 * center, weights wc_i = w_i * r_ic (float, unweighted: r_ic) over ALL rows.
 *  - L1, dense rows (kmedian.h:266-301, weighted_median, which never adds up its
 *    weights): each coordinate takes the value of the row of largest weight, the
 *    smallest value among rows of equal largest weight.  A center whose weights
 *    are all zero falls to the rule's lower_bound branch with mid = 0: in sorted
 *    order the first value v >= 0 and the next one w give v == 0 ? (v + w) / 2 : w
 *    (w = v when v is the last; no value >= 0 gives 0, where the reference reads
 *    past the column).  CSR rows: the reference throws (csc.h:834); -1 here.
 *  - L2, dense rows (blaze_adaptor.h:615-656): Weiszfeld from the current center,
 *    distances floored at 1e-30 (NaN -> 1e-30); the cost sum_i wc_i * dist_i is
 *    rounded to float; stop when it is not below the previous one; else the center
 *    becomes sum_i (wc_i / dist_i) x_i / sum_i (wc_i / dist_i).  One row: that row.
 *  - L2, CSR rows (csc.h:1131-1203 with no row list): the same update in double
 *    with the floor 1e-80, but the stopping cost is the UNWEIGHTED sum of the
 *    distances (csc.h:1151), compared in double; at most 99999 updates (both).
 *    A center whose weights are all zero gets NaN there (the reference throws): -2.
 * Weighted sums here are in double and rounded to float once (parity is to a
 * tolerance: the reference's float sums have no fixed order).
 * ====================================================================== */
static double soft_row_l2(const float *X, const float *val, const uint32_t *idx, const uint64_t *indptr,
                          uint64_t d, uint64_t i, const float *c) {
    double s = 0.;
    if (!val) {
        for (uint64_t j = 0; j < d; ++j) { const double t = (double)X[i * d + j] - (double)c[j]; s += t * t; }
    } else {
        double cc = 0.;
        for (uint64_t j = 0; j < d; ++j) cc += (double)c[j] * (double)c[j];
        s = cc;
        for (uint64_t p = indptr[i]; p < indptr[i + 1]; ++p) {
            const double x = val[p], cj = c[idx[p]];
            s += (x - cj) * (x - cj) - cj * cj;
        }
        if (s < 0.) s = 0.;
    }
    return sqrt(s);
}

static int soft_median_step(int msr, const float *X, const float *val, const uint32_t *idx, const uint64_t *indptr,
                            uint64_t n, uint64_t d, const float *resp, const double *w, uint64_t k, float *Cc) {
    if (msr == MCO_L1 && val) return -1;
    float *wc = (float *)malloc(n * sizeof(float));
    double *dist = (double *)malloc(n * sizeof(double)), *acc = (double *)malloc(d * sizeof(double));
    for (uint64_t c = 0; c < k; ++c) {
        float *ctr = Cc + c * d;
        for (uint64_t i = 0; i < n; ++i) wc[i] = (w ? (float)w[i] : 1.f) * resp[i * k + c];
        if (msr == MCO_L1) {
            float mw = -FLT_MAX;
            for (uint64_t i = 0; i < n; ++i) if (wc[i] > mw) mw = wc[i];
            if (mw > 0.f) {
                for (uint64_t j = 0; j < d; ++j) {
                    float v = FLT_MAX;
                    for (uint64_t i = 0; i < n; ++i) if (wc[i] == mw && X[i * d + j] < v) v = X[i * d + j];
                    ctr[j] = v;
                }
            } else {
                for (uint64_t j = 0; j < d; ++j) {
                    float v1 = FLT_MAX, v2 = FLT_MAX;   /* two smallest values >= 0, duplicates counted */
                    int cnt = 0;
                    for (uint64_t i = 0; i < n; ++i) {
                        const float x = X[i * d + j];
                        if (!(x >= 0.f)) continue;
                        ++cnt;
                        if (x < v1) { v2 = v1; v1 = x; } else if (x < v2) v2 = x;
                    }
                    if (cnt == 0) ctr[j] = 0.f;
                    else {
                        if (cnt == 1) v2 = v1;
                        ctr[j] = v1 == 0.f ? (float)(.5 * ((double)v1 + (double)v2)) : v2;
                    }
                }
            }
            continue;
        }
        /* L2: Weiszfeld */
        if (!val && n == 1) { memcpy(ctr, X, d * sizeof(float)); continue; }
        const double floor_d = val ? 1e-80 : (double)1e-30f;
        double prev = val ? DBL_MAX : (double)FLT_MAX;
        for (int it = 0;;) {
            double cur = 0.;
            for (uint64_t i = 0; i < n; ++i) {
                double v = soft_row_l2(X, val, idx, indptr, d, i, ctr);
                if (!val) v = (double)(float)v;
                if (isnan(v) || v < floor_d) v = floor_d;
                dist[i] = v;
                cur += val ? v : (double)wc[i] * v;
            }
            if (!val) cur = (double)(float)cur;
            if (cur >= prev || ++it == 100000) break;
            prev = cur;
            double qs = 0.;
            for (uint64_t j = 0; j < d; ++j) acc[j] = 0.;
            for (uint64_t i = 0; i < n; ++i) {
                const double q = (double)wc[i] / dist[i];
                qs += q;
                if (val) {
                    for (uint64_t p = indptr[i]; p < indptr[i + 1]; ++p) acc[idx[p]] += q * (double)val[p];
                } else {
                    for (uint64_t j = 0; j < d; ++j) acc[j] += q * (double)X[i * d + j];
                }
            }
            if (!(qs > 0.)) { free(wc); free(dist); free(acc); return -2; }   /* all weights zero: NaN */
            for (uint64_t j = 0; j < d; ++j) ctr[j] = (float)(acc[j] / qs);
        }
    }
    free(wc); free(dist); free(acc);
    return 0;
}

int mco_soft(int msr, const float *X, const float *val, const uint32_t *idx,
             const uint64_t *indptr, uint64_t n, uint64_t d, float *C, uint64_t k,
             double prior, double temp, const double *w, double eps, uint64_t maxiter,
             float *costs, float *resp, double *initcost, double *finalcost, uint64_t *numiter) {
    float *Cc = (float *)malloc(k * d * sizeof(float));
    memcpy(Cc, C, k * d * sizeof(float));
    soft_costs(msr, X, val, idx, indptr, n, d, Cc, k, prior, costs);
    double cost = DBL_MAX, init = -1.;
    uint64_t iter = 0;
    double *num = (double *)malloc(k * d * sizeof(double)), *den = (double *)malloc(k * sizeof(double));
    for (;;) {
        double oldcost = cost;
        mco_softmax_rows(costs, n, k, temp, resp);
        double ret = 0.;
        for (uint64_t i = 0; i < n; ++i) {
            double dt = 0.;
            for (uint64_t c = 0; c < k; ++c) dt += (double)costs[i * k + c] * (double)resp[i * k + c];
            ret += dt * (w ? w[i] : 1.);
        }
        cost = ret;
        if (msr == MCO_L1 || msr == MCO_L2) {
            const int rc = soft_median_step(msr, X, val, idx, indptr, n, d, resp, w, k, Cc);
            if (rc) {
                free(num); free(den); free(Cc);
                return rc;
            }
            goto costs;
        }
        /* M-step: dense numerator carries w (centroid.h:840-873); CSR omits it (:925-939). */
        memset(num, 0, k * d * sizeof(double)); memset(den, 0, k * sizeof(double));
        for (uint64_t i = 0; i < n; ++i) {
            double wi = w ? w[i] : 1.;
            for (uint64_t c = 0; c < k; ++c) {
                double r = resp[i * k + c];
                den[c] += r * wi;
                if (val) {
                    for (uint64_t p = indptr[i]; p < indptr[i + 1]; ++p) num[c * d + idx[p]] += r * (double)val[p];
                } else {
                    double rw = r * wi;
                    for (uint64_t j = 0; j < d; ++j) num[c * d + j] += rw * (double)X[i * d + j];
                }
            }
        }
        for (uint64_t c = 0; c < k; ++c)
            for (uint64_t j = 0; j < d; ++j) Cc[c * d + j] = (float)(num[c * d + j] / den[c]);
    costs:
        soft_costs(msr, X, val, idx, indptr, n, d, Cc, k, prior, costs);
        if (init < 0.) init = cost;
        if (oldcost >= cost) memcpy(C, Cc, k * d * sizeof(float));
        double mx = oldcost > cost ? oldcost : cost;
        if (oldcost - cost <= eps * mx || ++iter == maxiter) break;
    }
    free(num); free(den); free(Cc);
    *initcost = init; *finalcost = cost; *numiter = iter;
    return 0;
}

/* ======================================================================
 * D^2 seeding: coresets::kmeanspp, kmeans.h:49-215 with n_local_samples=1.
 * oracle(center, i) = msr(first = row i, second = row center) (smw.h:370-374).
 * Distances are the FT=float values: SQRL2 uses the FP32 sequential-FMA sum.
 * ====================================================================== */
static double pair_rows(int msr, const float *X, const float *val, const uint32_t *idx,
                        const uint64_t *indptr, uint64_t d, uint64_t i, uint64_t c,
                        const double *rsum, double prior) {
    if (!val) {
        if (msr == MCO_SQRL2) return (double)mco_sqrl2_f32(X + i * d, X + c * d, d);
        double *xb = (double *)malloc(d * sizeof(double)), *cb = (double *)malloc(d * sizeof(double));
        for (uint64_t j = 0; j < d; ++j) { xb[j] = X[i * d + j]; cb[j] = X[c * d + j]; }
        double v = (double)(float)mco_msr_dense(msr, xb, cb, d, prior, rsum[i], rsum[c]);
        free(xb); free(cb);
        return v;
    }
    uint64_t xb = indptr[i], xe = indptr[i + 1], cb = indptr[c], ce = indptr[c + 1];
    double *xv = (double *)malloc((xe - xb + 1) * sizeof(double)), *cv = (double *)malloc((ce - cb + 1) * sizeof(double));
    for (uint64_t p = xb; p < xe; ++p) xv[p - xb] = val[p];
    for (uint64_t p = cb; p < ce; ++p) cv[p - cb] = val[p];
    double v = mco_msr_sparse(msr, xv, idx + xb, xe - xb, cv, idx + cb, ce - cb, d, prior, rsum[i], rsum[c], NULL);
    if (msr != MCO_SQRL2) v = (double)(float)v;
    free(xv); free(cv);
    return v;
}

uint64_t mco_kmeanspp(int msr, const float *X, const float *val, const uint32_t *idx,
                      const uint64_t *indptr, uint64_t n, uint64_t d, uint64_t k,
                      double prior, uint64_t seed, const double *w,
                      uint32_t *centers, uint32_t *asn, float *distf, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    if (n < k) k = n;
    if (k == 0) return 0;
    double *rsum = (double *)malloc(n * sizeof(double));
    for (uint64_t i = 0; i < n; ++i)
        rsum[i] = val ? row_sum_csr(val, indptr[i], indptr[i + 1]) : row_sum_dense(X + i * d, d);
    double *dist = (double *)malloc(n * sizeof(double)), *rv = (double *)malloc(n * sizeof(double));
    uint64_t *fw = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t j = 0;
    uint64_t fc = mco_wyrand_at(seed, j++) % n;                    /* kmeans.h:79 */
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i)
        dist[i] = ((uint64_t)i == fc) ? 0. : pair_rows(msr, X, val, idx, indptr, d, i, fc, rsum, prior);
    for (uint64_t i = 0; i < n; ++i) asn[i] = 0;
    centers[0] = (uint32_t)fc;
    for (uint64_t ci = 1; ci < k; ++ci) {
        uint64_t r = mco_wyrand_at(seed, j++);
        for (uint64_t i = 0; i < n; ++i) rv[i] = w ? w[i] * dist[i] : dist[i];
        mco_fixed_weights(rv, n, fw);
        uint64_t newc = mco_fixed_draw(fw, n, r);
        if (dist[newc] == 0.) {                                      /* kmeans.h:179-186 */
            uint64_t b = 0; for (uint64_t i = 1; i < n; ++i) if (dist[i] > dist[b]) b = i;
            newc = b;
        }
        asn[newc] = (uint32_t)ci; centers[ci] = (uint32_t)newc;
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)n; ++i) {                  /* kmeans.h:189-201 */
            if ((uint64_t)i == newc || dist[i] <= 0.) continue;
            double v = pair_rows(msr, X, val, idx, indptr, d, i, newc, rsum, prior);
            if (v < dist[i]) { asn[i] = (uint32_t)ci; dist[i] = v; }
        }
        dist[newc] = 0.;
    }
    for (uint64_t i = 0; i < n; ++i) distf[i] = (float)dist[i];
    free(rsum); free(dist); free(rv); free(fw);
    return k;
}

/* ======================================================================
 * Local search++ after D^2 seeding: localsearchpp_rounds, lsearchpp.h:21-121
 * (kmeans.h:209-212).  M[i][y] = cost of row i to center row ctrs[y] (the
 * seeding's orientation; the reference's matrix uses the opposite argument
 * order, which differs only for asymmetric measures).  Each round draws a row
 * (fixed-point, over the distances, or over w_i times the previous round's
 * new costs when weighted, lsearchpp.h:53-56) with the seeding stream's next
 * output, computes every row's cost to it, charges each row that would be
 * worse off to its best center (second best minus best, FP32 differences)
 * and credits the others (old minus new), and swaps the cheapest center out
 * when its charge is below the credit.  Charges and credit are exact
 * fixed-point sums (the reference adds floats under OpenMP atomics).  At the
 * end each row takes its first cheapest center.
 * ====================================================================== */
static double row_pair_cost(int msr, const float *X, const float *val, const uint32_t *idx, const uint64_t *indptr,
                            uint64_t d, uint64_t i, uint64_t c, const double *rsum, double prior) {
    if (msr == MCO_SQRL2 && !val) return (double)mco_sqrl2_f32(X + i * d, X + c * d, d);
    return pair_rows(msr, X, val, idx, indptr, d, i, c, rsum, prior);
}

/* rounds of local search++ on seeded centers; the draws continue the seeding's
 * stream at output j0 */
static void ls_rounds(int msr, const float *X, const float *val, const uint32_t *idx, const uint64_t *indptr,
                      uint64_t n, uint64_t d, uint64_t k, double prior, uint64_t seed, const double *w,
                      uint64_t rounds, uint64_t j0, uint32_t *centers, uint32_t *asn, float *distf) {
    double *rsum = (double *)malloc(n * sizeof(double));
    for (uint64_t i = 0; i < n; ++i)
        rsum[i] = val ? row_sum_csr(val, indptr[i], indptr[i + 1]) : row_sum_dense(X + i * d, d);
    float *M = (float *)malloc(n * k * sizeof(float)), *dist = (float *)malloc(n * sizeof(float));
    float *nc = (float *)calloc(n, sizeof(float)), *t = (float *)malloc(n * sizeof(float));
    float *g = (float *)malloc(n * sizeof(float));
    uint32_t *top = (uint32_t *)malloc(n * sizeof(uint32_t));
    double *wv = (double *)malloc(n * sizeof(double));
    uint64_t *fw = (uint64_t *)malloc(n * sizeof(uint64_t)), *cf = (uint64_t *)malloc(k * sizeof(uint64_t));
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i)
        for (uint64_t y = 0; y < k; ++y)
            M[i * k + y] = (float)row_pair_cost(msr, X, val, idx, indptr, d, i, centers[y], rsum, prior);
    for (uint64_t i = 0; i < n; ++i) {
        float m = M[i * k];
        for (uint64_t y = 1; y < k; ++y) if (M[i * k + y] < m) m = M[i * k + y];
        dist[i] = m;
    }
    uint64_t j = j0;
    for (uint64_t rd = 0; rd < rounds; ++rd) {
        for (uint64_t i = 0; i < n; ++i) wv[i] = w ? w[i] * (double)nc[i] : (double)dist[i];
        mco_fixed_weights(wv, n, fw);
        const uint64_t sel = mco_fixed_draw(fw, n, mco_wyrand_at(seed, j++));
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)n; ++i)
            nc[i] = (float)row_pair_cost(msr, X, val, idx, indptr, d, i, sel, rsum, prior);
        double mx = 0.;
        for (uint64_t i = 0; i < n; ++i) {
            t[i] = 0.f; g[i] = 0.f; top[i] = 0;
            if (nc[i] > dist[i]) {
                const float *r = M + i * k;
                float v1 = r[0], v2 = r[1];
                uint32_t j1 = 0, j2 = 1;
                if (v2 < v1) { float tv = v1; v1 = v2; v2 = tv; uint32_t tj = j1; j1 = j2; j2 = tj; }
                for (uint64_t y = 2; y < k; ++y) {
                    const float v = r[y];
                    if (v < v2) {
                        if (v < v1) { v2 = v1; j2 = j1; v1 = v; j1 = (uint32_t)y; }
                        else { v2 = v; j2 = (uint32_t)y; }
                    }
                }
                (void)j2;
                if (v1 != v2) { t[i] = v2 - v1; top[i] = j1; }
            } else {
                g[i] = dist[i] - nc[i];
            }
            if ((double)t[i] > mx) mx = t[i];
            if ((double)g[i] > mx) mx = g[i];
        }
        const int s = mco_fixed_shift(mx, n);
        memset(cf, 0, k * sizeof(uint64_t));
        uint64_t gain = 0;
        for (uint64_t i = 0; i < n; ++i) {
            if (t[i] > 0.f) cf[top[i]] += (uint64_t)floor(ldexp((double)t[i], s));
            if (g[i] > 0.f) gain += (uint64_t)floor(ldexp((double)g[i], s));
        }
        uint64_t am = 0;
        for (uint64_t y = 1; y < k; ++y) if (cf[y] < cf[am]) am = y;
        if (cf[am] < gain) {       /* delta = charge - credit < 0 */
            centers[am] = (uint32_t)sel;
            for (uint64_t i = 0; i < n; ++i) M[i * k + am] = nc[i];
            for (uint64_t i = 0; i < n; ++i) {
                float m = M[i * k];
                for (uint64_t y = 1; y < k; ++y) if (M[i * k + y] < m) m = M[i * k + y];
                dist[i] = m;
            }
        }
    }
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t b = 0;
        for (uint64_t y = 1; y < k; ++y) if (M[i * k + y] < M[i * k + b]) b = (uint32_t)y;
        asn[i] = b;
        distf[i] = dist[i];
    }
    free(rsum); free(M); free(dist); free(nc); free(t); free(g); free(top); free(wv); free(fw); free(cf);
}

uint64_t mco_kmeanspp_ls(int msr, const float *X, const float *val, const uint32_t *idx,
                         const uint64_t *indptr, uint64_t n, uint64_t d, uint64_t k,
                         double prior, uint64_t seed, const double *w, uint64_t rounds,
                         uint32_t *centers, uint32_t *asn, float *distf, int nthreads) {
    k = mco_kmeanspp(msr, X, val, idx, indptr, n, d, k, prior, seed, w, centers, asn, distf, nthreads);
    /* the seeding used the stream's outputs 0..k-1 */
    if (rounds > 0 && k >= 2) ls_rounds(msr, X, val, idx, indptr, n, d, k, prior, seed, w, rounds, k, centers, asn, distf);
    return k;
}

/* local search++ rounds from given centers (k >= 2); the draws use outputs j0, j0+1, ...
 * of the stream of seed */
void mco_localsearch(int msr, const float *X, const float *val, const uint32_t *idx, const uint64_t *indptr,
                     uint64_t n, uint64_t d, uint64_t k, double prior, uint64_t seed, const double *w,
                     uint64_t rounds, uint64_t j0, uint32_t *centers, uint32_t *asn, float *distf) {
    ls_rounds(msr, X, val, idx, indptr, n, d, k, prior, seed, w, rounds, j0, centers, asn, distf);
}

/* ======================================================================
 * Coreset sensitivities, coreset.h:423-578.  Sums over points are exact:
 * each term is mapped to fixed point with one shift for the whole vector
 * (from its max) and added as integers, so the result does not depend on
 * the order of addition; probabilities are then formed in FP64 and rounded
 * to float (CoresetSampler<float, uint32_t>, pyfgc.h:36).
 * ====================================================================== */
typedef struct { int shift; } fx_t;

static fx_t fx_from(const double *t, uint64_t n) {
    double mx = 0.;
    for (uint64_t i = 0; i < n; ++i) { double a = fabs(t[i]); if (a > mx) mx = a; }
    fx_t f; f.shift = mco_fixed_shift(mx, n);
    return f;
}
static int64_t fx_of(double v, fx_t f) {
    double s = ldexp(v, f.shift);
    return (int64_t)(s < 0. ? -floor(-s) : floor(s));
}
static double fx_back(int64_t v, fx_t f) { return ldexp((double)v, -f.shift); }

int mco_coreset_probs(int sens, uint64_t n, uint64_t ncenters, const double *costs,
                      const uint32_t *asn, const double *w, float *probs) {
    double *wf = (double *)malloc(n * sizeof(double)), *wc = (double *)malloc(n * sizeof(double));
    for (uint64_t i = 0; i < n; ++i) {
        wf[i] = w ? (double)(float)w[i] : 1.;      /* weights_ is DynamicVector<float> */
        wc[i] = wf[i] * costs[i];
    }
    fx_t fw = fx_from(wf, n), fc = fx_from(wc, n);
    int64_t *ws = (int64_t *)calloc(ncenters, sizeof(int64_t)), *cs = (int64_t *)calloc(ncenters, sizeof(int64_t));
    uint64_t *cnt = (uint64_t *)calloc(ncenters, sizeof(uint64_t));
    int64_t wtot = 0, ctot = 0;
    for (uint64_t i = 0; i < n; ++i) {
        int64_t a = fx_of(wf[i], fw), b = fx_of(wc[i], fc);
        ws[asn[i]] += a; cs[asn[i]] += b; cnt[asn[i]]++;
        wtot += a; ctot += b;
    }
    double W = fx_back(wtot, fw), TC = fx_back(ctot, fc);
    double *p = (double *)malloc(n * sizeof(double));
    int rc = 0;
    if (sens == 4) {                                    /* LBK, :487-533 */
        double alpha = 16. * log((double)ncenters) + 32., alpha2 = 2. * alpha;
        double mean = TC / W, tcinv = alpha / mean;
        double *adj = (double *)malloc(ncenters * sizeof(double));
        for (uint64_t c = 0; c < ncenters; ++c) {
            double Wc = fx_back(ws[c], fw), Sc = fx_back(cs[c], fc);
            adj[c] = alpha2 * Sc / (Wc * mean) + 4. * W / Wc;
        }
        for (uint64_t i = 0; i < n; ++i) p[i] = tcinv * costs[i] + adj[asn[i]];
        free(adj);
    } else if (sens == 2) {                             /* LFKF, :341-381, alpha_est = 0: 2T/W_a */
        for (uint64_t i = 0; i < n; ++i) p[i] = 1. / fx_back(ws[asn[i]], fw);
    } else if (sens == 3) {                             /* VX, :423-455 */
        for (uint64_t i = 0; i < n; ++i) p[i] = wc[i] / TC + 1. / (double)cnt[asn[i]];
    } else if (sens == 0) {                             /* BFL, :534-578 */
        for (uint64_t i = 0; i < n; ++i) {
            uint32_t a = asn[i];
            p[i] = .5 * (wc[i] / TC + wf[i] / (fx_back(ws[a], fw) * (double)cnt[a]));
        }
    } else if (sens == 1) {                             /* FL, :456-486 (not normalised) */
        for (uint64_t i = 0; i < n; ++i)
            probs[i] = w ? (float)(wc[i] / TC) : (float)(ceil((double)n * costs[i] / TC) + 1.);
        goto done;
    } else { rc = -1; goto done; }
    {
        fx_t fp = fx_from(p, n);
        int64_t ptot = 0;
        for (uint64_t i = 0; i < n; ++i) ptot += fx_of(p[i], fp);
        double P = fx_back(ptot, fp);
        for (uint64_t i = 0; i < n; ++i) probs[i] = (float)(p[i] / P);
    }
done:
    free(wf); free(wc); free(ws); free(cs); free(cnt); free(p);
    return rc;
}

/* coreset.h:625-651: m i.i.d. draws; weight = w_i / (m * p_i).  Draw t uses output
 * first + t of the seed's counter-based stream. */
void mco_coreset_sample(uint64_t n, const float *probs, const double *w, uint64_t m,
                        uint64_t seed, uint64_t first, uint64_t *indices, float *weights) {
    double *pd = (double *)malloc(n * sizeof(double));
    uint64_t *fw = (uint64_t *)malloc(n * sizeof(uint64_t)), *pre = (uint64_t *)malloc(n * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) pd[i] = probs[i];
    mco_fixed_weights(pd, n, fw);
    uint64_t T = 0;
    for (uint64_t i = 0; i < n; ++i) { T += fw[i]; pre[i] = T; }
    for (uint64_t t = 0; t < m; ++t) {
        uint64_t u = mco_mulhi64(mco_wyrand_at(seed, first + t), T);
        uint64_t lo = 0, hi = n;                   /* first i with pre[i] > u */
        while (lo < hi) { uint64_t mid = (lo + hi) / 2; if (pre[mid] > u) hi = mid; else lo = mid + 1; }
        if (lo >= n) lo = n - 1;
        indices[t] = lo;
        double wi = w ? (double)(float)w[lo] : 1.;
        weights[t] = (float)(wi / ((double)m * (double)probs[lo]));
    }
    free(pd); free(fw); free(pre);
}

/* ======================================================================
 * Median centroid policies: set_centroids_l1 / set_centroids_tvd /
 * set_centroids_l2 (centroid.h:434-647) with l1_median / weighted_median /
 * tvd_median (optim/kmedian.h:124-350, util/csc.h:829-950) and the dense
 * geometric median (util/blaze_adaptor.h:614-690), under the hard Lloyd loop
 * of solve.h:98-175 (set_centroids_hard never reports a restart for these
 * policies, solve.h:207-215, so a rising cost always reverts).
 *
 * Empty clusters (centroid.h:442-480): each, in ascending id, takes the row
 * drawn over the current per-point costs (fixed-point sampler, wyrand stream
 * of n, draws 0, 1, ... per M-step), and every row strictly closer to it by
 * raw L1 (L1 and TVD policies) or L2 (geometric median) moves there; repeated
 * until no cluster is empty.  Then per cluster, members in ascending row id:
 *  - one member: the center is that row (raw, also under TVD);
 *  - L1, unweighted: per column the median of the members' values, even
 *    counts 0.5f * (v[h-1] + v[h]) in float (kmedian.h:124-150);
 *  - L1, weighted: weighted_median never adds to its weight total
 *    (kmedian.h:275-285), so every column takes the value of the heaviest
 *    member (ties: the smallest value);
 *  - TVD: the same on x_ij * (1 / rowsum_i) in double (kmedian.h:347-350),
 *    rounded to float; weighted, the member of rank j carries w[j], the
 *    weight of ROW j (the full weight vector indexed by rank, kmedian.h:340);
 *  - CSR L1 / TVD (unweighted only; csc.h throws for weights): the rules of
 *    csc.h:852-884 / 924-950 over the stored values with (m - nel) implicit
 *    zeros, for non-negative data (see csr_median_value);
 *  - geometric median (dense): Weiszfeld from the current center; per
 *    round d_i = max(float(||x_i - c||), floor) (floor 1e-30f, weighted
 *    1e-40f), cost = float(sum d_i) (weighted: sum w_i d_i); stop when the
 *    previous cost minus this one is <= 0 (the center stays the one just
 *    evaluated); else q_i = 1/d_i (weighted w_i/d_i), the center becomes
 *    sum_i (q_i / Q) x_i (weighted q_i * (1/Q)).  All in FP64 here, rounded
 *    to float per center; member sums are 256 strided sequential partials
 *    added as a binary tree, center sums 4 strided partials added
 *    ((p0 + p1) + (p2 + p3)), the order the GPU library uses.  The reference
 *    works in float, where w/1e-40f overflows (parity unpinned there).
 * ====================================================================== */
static int median_policy(int msr) {
    return msr == MCO_L1 ? 1 : msr == MCO_TVD ? 2 : msr == MCO_L2 ? 3 : 0;
}

static int cmp_f32(const void *a, const void *b) {
    const float x = *(const float *)a, y = *(const float *)b;
    return x < y ? -1 : x > y;
}
static int cmp_f64(const void *a, const void *b) {
    const double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y;
}

static double tree_sum256(const double *v, uint64_t m) {
    double p[256];
    for (int t = 0; t < 256; ++t) {
        double s = 0.;
        for (uint64_t q = (uint64_t)t; q < m; q += 256) s += v[q];
        p[t] = s;
    }
    for (int o = 128; o > 0; o >>= 1)
        for (int t = 0; t < o; ++t) p[t] += p[t + o];
    return p[0];
}

/* Center value of one CSR column from its nel stored values s (sorted
 * ascending, non-negative) among m members: full = (m - nel) zeros then s.
 * csc.h:852-884 (L1) / 924-950 (TVD): with nel > m/2 and all stored values
 * positive, full[h] (odd) or (full[h] + full[h+1]) / 2 (even); with an explicit
 * zero stored, odd full[h], even L1 (full[h] + full[h+1]) / 2, TVD
 * (full[h-1] + full[h]) / 2; with odd m and nel == m/2 and all positive,
 * s[0] / 2; otherwise 0.  Where the reference reads one past the column's
 * values (m == 2 with the even rule h+1), the index is clamped to the last. */
static float csr_median_value(const float *s, uint64_t nel, uint64_t m, int tvd) {
    const uint64_t h = m / 2, z = m - nel;
    const int odd = (int)(m & 1);
    #define FULL(r) ((r) < z ? 0.f : s[(r) - z])
    if (nel > h) {
        if (odd) return FULL(h);
        const uint64_t r1 = h + 1 < m ? h + 1 : m - 1;
        if (s[0] > 0.f || !tvd) return (float)(.5 * (double)(FULL(h) + FULL(r1)));
        return (float)(.5 * (double)(FULL(h - 1) + FULL(h)));
    }
    if (odd && nel == h && s[0] > 0.f) return (float)(.5 * (double)s[0]);
    return 0.f;
    #undef FULL
}

/* csr_rule: the geometric median of CSR rows (csc.h:1131-1203, unweighted, with the
 * member costs indexed by position): distances kept in FP64, floored at 1e-80, the
 * cost summed in FP64, stopping when it is within eps of the previous one or above it. */
static void geo_median(const float *X, uint64_t d, const uint64_t *mem, uint64_t m, const double *w, float *cur,
                       double eps, double *cq, double *qq, int csr_rule) {
    const double floor_ = csr_rule ? 1e-80 : w ? (double)1e-40f : (double)1e-30f;
    float prev = FLT_MAX;
    double prevd = DBL_MAX;
    double *acc = (double *)malloc(4 * d * sizeof(double));
    for (uint64_t it = 0; it < 100000; ++it) {
        for (uint64_t q = 0; q < m; ++q) {
            const float *x = X + mem[q] * d;
            double s = 0.;
            for (uint64_t j = 0; j < d; ++j) { const double t = (double)x[j] - (double)cur[j]; s = fma(t, t, s); }
            double c = csr_rule ? sqrt(s) : (double)(float)sqrt(s);
            if (isnan(c)) c = floor_;
            cq[q] = c > floor_ ? c : floor_;
            qq[q] = w ? w[mem[q]] * cq[q] : cq[q];
        }
        const double curd = tree_sum256(qq, m);
        const float curf = (float)curd;
        if (csr_rule) {
            const double dist = fabs(prevd - curd);
            if (dist <= eps || curd > prevd || isnan(dist)) break;
        } else {
            const float dist = prev - curf;
            if ((double)dist <= eps || isnan(dist)) break;
        }
        prevd = curd;
        for (uint64_t q = 0; q < m; ++q) qq[q] = w ? w[mem[q]] / cq[q] : 1. / cq[q];
        const double S = tree_sum256(qq, m), inv = 1. / S;
        memset(acc, 0, 4 * d * sizeof(double));
        for (uint64_t q = 0; q < m; ++q) {
            const double coef = w ? qq[q] * inv : qq[q] / S;
            const float *x = X + mem[q] * d;
            double *a = acc + (q & 3) * d;
            for (uint64_t j = 0; j < d; ++j) a[j] = fma(coef, (double)x[j], a[j]);
        }
        for (uint64_t j = 0; j < d; ++j) cur[j] = (float)((acc[j] + acc[d + j]) + (acc[2 * d + j] + acc[3 * d + j]));
        prev = curf;
    }
    free(acc);
}

/* The centers of one median policy for the assignment asn (empty clusters keep
 * theirs).  wm: per-row weights (the weighted L1 of set_centroids_l1 and the
 * geometric median); wr: the weight vector the reference indexes by member rank
 * (TVD always; L1 too when rank_l1, the minibatch update's l1_median, kmedian.h:340).
 * single_raw: a one-member cluster takes its raw row (set_centroids_*), else the
 * rule for one member (the minibatch update; TVD then normalises it). */
static void median_core(int pol, const float *X, const float *val, const uint32_t *idx, const uint64_t *indptr,
                        uint64_t n, uint64_t d, const double *wm, const double *wr, uint64_t k, const uint32_t *asn,
                        float *C, double eps, int single_raw, int rank_l1) {
    /* members per cluster, ascending row id */
    uint64_t *off = (uint64_t *)calloc(k + 1, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) off[asn[i] + 1]++;
    for (uint64_t c = 0; c < k; ++c) off[c + 1] += off[c];
    uint64_t *mem = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t)), *cur = (uint64_t *)malloc(k * sizeof(uint64_t));
    memcpy(cur, off, k * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) mem[cur[asn[i]]++] = i;
    double *rs = (double *)malloc((n ? n : 1) * sizeof(double));
    for (uint64_t i = 0; i < n; ++i) rs[i] = val ? row_sum_csr(val, indptr[i], indptr[i + 1]) : row_sum_dense(X + i * d, d);
    float *vf = (float *)malloc((n ? n : 1) * sizeof(float));
    double *vd = (double *)malloc((n ? n : 1) * sizeof(double)), *qq = (double *)malloc((n ? n : 1) * sizeof(double));
    float *Xdense = NULL;   /* CSR rows as dense rows, for the geometric median */
    if (val && pol == 3) {
        Xdense = (float *)calloc(n * d + 1, sizeof(float));
        for (uint64_t i = 0; i < n; ++i)
            for (uint64_t p = indptr[i]; p < indptr[i + 1]; ++p) Xdense[i * d + idx[p]] = val[p];
    }
    for (uint64_t c = 0; c < k; ++c) {
        const uint64_t *mm = mem + off[c], m = off[c + 1] - off[c];
        float *ctr = C + c * d;
        if (m == 0) continue;
        if (m == 1 && (single_raw || pol == 3)) {   /* geomedian of one row: the row */
            const uint64_t r = mm[0];
            if (val) {
                memset(ctr, 0, d * sizeof(float));
                for (uint64_t p = indptr[r]; p < indptr[r + 1]; ++p) ctr[idx[p]] = val[p];
            } else memcpy(ctr, X + r * d, d * sizeof(float));
            continue;
        }
        if (pol == 3) { geo_median(val ? Xdense : X, d, mm, m, wm, ctr, eps, vd, qq, val != NULL); continue; }
        if (val) {
            /* per column, the stored values of the members */
            for (uint64_t j = 0; j < d; ++j) {
                uint64_t nel = 0;
                for (uint64_t q = 0; q < m; ++q) {
                    const uint64_t r = mm[q];
                    for (uint64_t p = indptr[r]; p < indptr[r + 1]; ++p) {
                        if (idx[p] != j) continue;
                        vf[nel++] = pol == 2 ? val[p] * (float)(1. / rs[r]) : val[p];
                    }
                }
                if (!nel) { ctr[j] = 0.f; continue; }
                qsort(vf, nel, sizeof(float), cmp_f32);
                ctr[j] = csr_median_value(vf, nel, m, pol == 2);
            }
            continue;
        }
        if (wm) {
            /* heaviest member (by row weight, or by wr[rank]), smallest value */
            const int byrank = pol == 2 || rank_l1;
            double wmax = -INFINITY;
            for (uint64_t q = 0; q < m; ++q) { const double wq = byrank ? wr[q] : wm[mm[q]]; if (wq > wmax) wmax = wq; }
            for (uint64_t j = 0; j < d; ++j) {
                double best = INFINITY;
                for (uint64_t q = 0; q < m; ++q) {
                    const double wq = byrank ? wr[q] : wm[mm[q]];
                    if (wq != wmax) continue;
                    const float x = X[mm[q] * d + j];
                    const double v = pol == 2 ? (double)x * (1. / rs[mm[q]]) : (double)x;
                    if (v < best) best = v;
                }
                ctr[j] = (float)best;
            }
            continue;
        }
        const uint64_t h = m / 2;
        for (uint64_t j = 0; j < d; ++j) {
            if (pol == 1) {
                for (uint64_t q = 0; q < m; ++q) vf[q] = X[mm[q] * d + j];
                qsort(vf, m, sizeof(float), cmp_f32);
                ctr[j] = (m & 1) ? vf[h] : 0.5f * (vf[h - 1] + vf[h]);
            } else {
                for (uint64_t q = 0; q < m; ++q) vd[q] = (double)X[mm[q] * d + j] * (1. / rs[mm[q]]);
                qsort(vd, m, sizeof(double), cmp_f64);
                ctr[j] = (float)((m & 1) ? vd[h] : 0.5 * (vd[h - 1] + vd[h]));
            }
        }
    }
    free(off); free(mem); free(cur); free(rs); free(vf); free(vd); free(qq); free(Xdense);
}


/* The M-step of one median policy from (asn, cost), which the empty-cluster
 * restart updates in place.  C holds the previous centers (the geometric
 * median starts from them; restarted clusters from their drawn row).
 * Returns 0, or -2 for an unsupported case (weighted CSR, CSR geometric median). */
int mco_median_centers(int msr, const float *X, const float *val, const uint32_t *idx, const uint64_t *indptr,
                       uint64_t n, uint64_t d, const double *w, uint64_t k, uint32_t *asn, double *cost, float *C,
                       double eps) {
    const int pol = median_policy(msr);
    if (!pol) return -2;
    if (val && w) return -2;
    uint64_t *cnt = (uint64_t *)calloc(k + 1, sizeof(uint64_t));
    uint64_t draw = 0;
    for (;;) {
        memset(cnt, 0, (k + 1) * sizeof(uint64_t));
        for (uint64_t i = 0; i < n; ++i) cnt[asn[i]]++;
        int any = 0;
        for (uint64_t c = 0; c < k; ++c) {
            if (cnt[c]) continue;
            any = 1;
            uint64_t *fw = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
            mco_fixed_weights(cost, n, fw);
            const uint64_t r = mco_fixed_draw(fw, n, mco_wyrand_at(n, draw++));
            free(fw);
            if (val) {
                memset(C + c * d, 0, d * sizeof(float));
                for (uint64_t p = indptr[r]; p < indptr[r + 1]; ++p) C[c * d + idx[p]] = val[p];
            } else memcpy(C + c * d, X + r * d, d * sizeof(float));
            const int rm = pol == 3 ? MCO_L2 : MCO_L1;
            for (uint64_t i = 0; i < n; ++i) {
                const double v = point_cost(rm, X, val, idx, indptr, i, d, C + c * d, 0.);
                if (v < cost[i]) { asn[i] = (uint32_t)c; cost[i] = v; }
            }
        }
        if (!any) break;
    }
    free(cnt);
    median_core(pol, X, val, idx, indptr, n, d, w, w, k, asn, C, eps, 1, 0);
    return 0;
}

/* perform_hard_clustering (solve.h:98-175) for L1, TVD and L2; costs in the
 * control flow rounded to float, as in mco_lloyd. */
int mco_lloyd_median(int msr, const float *X, const float *val, const uint32_t *idx, const uint64_t *indptr,
                     uint64_t n, uint64_t d, float *C, uint64_t k, double prior, const double *w, double eps,
                     uint64_t maxiter, uint32_t *asn, double *cost, double *initcost, double *finalcost,
                     uint64_t *numiter) {
    if (!median_policy(msr)) return -2;
    if (val && w) return -2;
    assign_generic(msr, X, val, idx, indptr, n, d, C, k, prior, asn, cost, 0);
    float fcost = (float)total_cost(cost, w, n);
    *initcost = fcost;
    if (fcost == 0.f) { *initcost = 0.; *finalcost = 0.; *numiter = 0; return 0; }
    float *Cc = (float *)malloc(k * d * sizeof(float));
    memcpy(Cc, C, k * d * sizeof(float));
    uint64_t iter = 0;
    int rc = 0;
    for (;;) {
        rc = mco_median_centers(msr, X, val, idx, indptr, n, d, w, k, asn, cost, Cc, 0.);
        if (rc) break;
        assign_generic(msr, X, val, idx, indptr, n, d, Cc, k, prior, asn, cost, 0);
        const float newcost = (float)total_cost(cost, w, n);
        if (newcost > fcost) {
            assign_generic(msr, X, val, idx, indptr, n, d, C, k, prior, asn, cost, 0);
            break;
        }
        memcpy(C, Cc, k * d * sizeof(float));
        ++iter;
        const float old = fcost;
        fcost = newcost;
        const float diff = old - newcost;
        if ((double)diff < eps * ((double)newcost > (double)old ? (double)newcost : (double)old) || iter > maxiter) break;
    }
    free(Cc);
    *finalcost = fcost;
    *numiter = iter;
    return rc;
}

/* ======================================================================
 * k-means++ with n_local_samples = L > 1 (kmeans.h:79-92, 114-177):
 *  - the first center is the cheapest (exact fixed-point total of w_i * d_i) of
 *    L rows rng() mod n, the first on ties;
 *  - the distances and assignment right after it are kept as the reference's
 *    odists / oasn, which it never refreshes (kmeans.h:104);
 *  - each later round takes one raw draw r = rng() and L samples from the
 *    fixed-point sampler over w_i * d_i with the outputs 0..L-1 of the stream
 *    seeded by r (reservoir_simd::sample_k is un-vendored: parity unpinned; with
 *    replacement here); candidate 0 updates the current distances in place,
 *    candidates t > 0 start from odists / oasn; the lowest total wins (the first
 *    on ties), and the winner's distances and assignment are kept;
 *  - the new center's assignment is its round and its distance 0.
 * Costs are the seeding's (pair_rows, row first) for every row, the new center's
 * own row included.  `rounds` rounds of local search++ follow (kmeans.h:209-212).
 * ====================================================================== */
uint64_t mco_kmeanspp_local(int msr, const float *X, const float *val, const uint32_t *idx,
                            const uint64_t *indptr, uint64_t n, uint64_t d, uint64_t k, double prior,
                            uint64_t seed, const double *w, uint64_t L, uint64_t rounds, uint32_t *centers,
                            uint32_t *asn, float *distf) {
    if (n < k) k = n;
    if (k == 0 || L < 1) return 0;
    double *rsum = (double *)malloc(n * sizeof(double));
    for (uint64_t i = 0; i < n; ++i)
        rsum[i] = val ? row_sum_csr(val, indptr[i], indptr[i + 1]) : row_sum_dense(X + i * d, d);
    double *dist = (double *)malloc(n * sizeof(double)), *od = (double *)malloc(n * sizeof(double));
    double *tmp = (double *)malloc(n * sizeof(double)), *rv = (double *)malloc(n * sizeof(double));
    uint32_t *oa = (uint32_t *)malloc(n * sizeof(uint32_t)), *ta = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint64_t *fw = (uint64_t *)malloc(n * sizeof(uint64_t)), *smp = (uint64_t *)malloc(L * sizeof(uint64_t));
    uint64_t j = 0;
    /* first center: the cheapest of L uniform rows (setdists: its own distance 0) */
    uint64_t fc = mco_wyrand_at(seed, j++) % n;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i)
        dist[i] = ((uint64_t)i == fc) ? 0. : pair_rows(msr, X, val, idx, indptr, d, i, fc, rsum, prior);
    double fcost = fixed_total(dist, w, n);
    for (uint64_t t = 1; t < L; ++t) {
        const uint64_t ofc = mco_wyrand_at(seed, j++) % n;
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)n; ++i)
            tmp[i] = ((uint64_t)i == ofc) ? 0. : pair_rows(msr, X, val, idx, indptr, d, i, ofc, rsum, prior);
        const double ocost = fixed_total(tmp, w, n);
        if (ocost < fcost) { double *s = dist; dist = tmp; tmp = s; fcost = ocost; fc = ofc; }
    }
    centers[0] = (uint32_t)fc;
    for (uint64_t i = 0; i < n; ++i) { asn[i] = 0; od[i] = dist[i]; oa[i] = 0; }
    for (uint64_t ci = 1; ci < k; ++ci) {
        const uint64_t rngv = mco_wyrand_at(seed, j++);
        for (uint64_t i = 0; i < n; ++i) rv[i] = w ? w[i] * dist[i] : dist[i];
        mco_fixed_weights(rv, n, fw);
        for (uint64_t t = 0; t < L; ++t) smp[t] = mco_fixed_draw(fw, n, mco_wyrand_at(rngv, t));
        double dsum = -1.;
        uint64_t newc = smp[0];
        for (uint64_t t = 0; t < L; ++t) {
            double *tg = dist;
            uint32_t *tga = asn;
            if (t > 0) { memcpy(tmp, od, n * sizeof(double)); memcpy(ta, oa, n * sizeof(uint32_t)); tg = tmp; tga = ta; }
            const uint64_t nextc = smp[t];
            #pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < (int64_t)n; ++i) {
                const double v = pair_rows(msr, X, val, idx, indptr, d, i, nextc, rsum, prior);
                if (v < tg[i]) { tg[i] = v; tga[i] = (uint32_t)ci; }
            }
            const double nsum = fixed_total(tg, w, n);
            if (dsum < 0.) { dsum = nsum; newc = nextc; }
            else if (nsum < dsum) {
                memcpy(dist, tmp, n * sizeof(double)); memcpy(asn, ta, n * sizeof(uint32_t));
                dsum = nsum; newc = nextc;
            }
        }
        centers[ci] = (uint32_t)newc;
        asn[newc] = (uint32_t)ci;
        dist[newc] = 0.;
    }
    for (uint64_t i = 0; i < n; ++i) distf[i] = (float)dist[i];
    free(rsum); free(dist); free(od); free(tmp); free(rv); free(oa); free(ta); free(fw); free(smp);
    /* the seeding used the stream's outputs 0..L+k-2 */
    if (rounds > 0 && k >= 2) ls_rounds(msr, X, val, idx, indptr, n, d, k, prior, seed, w, rounds, L + k - 1, centers, asn, distf);
    return k;
}

/* ======================================================================
 * float64 rows (pycluster_dense.cpp:85-90 instantiates the dense path with
 * FT = double for float64 arrays): perform_hard_clustering (solve.h:98-175) for
 * SQRL2 with every value in FP64.  Per pair the sequential FMA chain
 * s = fma(x_j - c_j, x_j - c_j, s) over j (FP64); the lowest cost wins, ties to
 * the lower index, the first center's cost sticks when it is NaN.  Means: the
 * chunked member order of mco_centroids_dense, on double rows, kept in double.
 * The control flow compares FP64 costs (no rounding to float: FT = double);
 * each cost is the exact sum of w_i * cost_i at the fixed-point scale of the
 * sampler (ldexp of an integer sum), so it does not depend on summation order.
 * An empty cluster takes the row drawn over the per-point costs (RNG seeded
 * with n), every row is re-assigned and the drawn rows are forced onto their
 * clusters (centroid.h:687-752), as in centroids_with_restart.
 * ====================================================================== */
double mco_sqrl2_f64(const double *x, const double *c, uint64_t d) {
    double s = 0.;
    for (uint64_t j = 0; j < d; ++j) { const double t = x[j] - c[j]; s = fma(t, t, s); }
    return s;
}

void mco_assign_sqrl2_f64(const double *X, uint64_t n, uint64_t d, const double *C, uint64_t k,
                          uint32_t *asn, double *cost) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const double *x = X + (uint64_t)i * d;
        double best = mco_sqrl2_f64(x, C, d);
        uint32_t bi = 0;
        for (uint64_t c = 1; c < k; ++c) {
            const double v = mco_sqrl2_f64(x, C + c * d, d);
            if (v < best) { best = v; bi = (uint32_t)c; }
        }
        asn[i] = bi; cost[i] = best;
    }
}

uint64_t mco_centroids_dense_f64(const double *X, uint64_t n, uint64_t d, const uint32_t *asn, const double *w,
                                 uint64_t k, uint64_t chunk, double *C) {
    uint64_t *cnt = (uint64_t *)calloc(k + 1, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) cnt[asn[i] + 1]++;
    for (uint64_t c = 0; c < k; ++c) cnt[c + 1] += cnt[c];
    uint64_t *perm = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    uint64_t *cur = (uint64_t *)malloc(k * sizeof(uint64_t));
    memcpy(cur, cnt, k * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) perm[cur[asn[i]]++] = i;
    uint64_t empty = 0;
    if (chunk == 0) chunk = MCO_MEAN_CHUNK;
    #pragma omp parallel for schedule(dynamic, 1) reduction(+:empty)
    for (int64_t c = 0; c < (int64_t)k; ++c) {
        uint64_t b = cnt[c], e = cnt[c + 1];
        if (b == e) { ++empty; continue; }
        double *tot = (double *)calloc(d, sizeof(double)), *part = (double *)malloc(d * sizeof(double));
        double wt = 0., wpart;
        for (uint64_t s = b; s < e; s += chunk) {
            uint64_t se = s + chunk < e ? s + chunk : e;
            memset(part, 0, d * sizeof(double)); wpart = 0.;
            for (uint64_t p = s; p < se; ++p) {
                const double *x = X + perm[p] * d;
                double wi = w ? w[perm[p]] : 1.;
                for (uint64_t j = 0; j < d; ++j) part[j] = w ? fma(wi, x[j], part[j]) : part[j] + x[j];
                wpart += wi;
            }
            for (uint64_t j = 0; j < d; ++j) tot[j] += part[j];
            wt += wpart;
        }
        for (uint64_t j = 0; j < d; ++j) C[c * d + j] = tot[j] / wt;
        free(tot); free(part);
    }
    free(cnt); free(perm); free(cur);
    return empty;
}

/* exact sum of w_i * cost_i (cost_i >= 0) at the sampler's fixed-point scale */
double mco_fixed_total(const double *cost, const double *w, uint64_t n) {
    double mx = 0.;
    for (uint64_t i = 0; i < n; ++i) {
        const double t = fabs(w ? w[i] * cost[i] : cost[i]);
        if (t > mx) mx = t;
    }
    const int s = mco_fixed_shift(mx, n);
    uint64_t T = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const double t = w ? w[i] * cost[i] : cost[i];
        if (t > 0.) T += (uint64_t)floor(ldexp(t, s));
    }
    return ldexp((double)T, -s);
}

static int centroids_with_restart_f64(const double *X, uint64_t n, uint64_t d, double *C, uint64_t k,
                                      const double *w, uint32_t *asn, double *cost) {
    uint64_t *cnt = (uint64_t *)calloc(k, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) cnt[asn[i]]++;
    int restarted = 0;
    uint64_t draw = 0, ne = 0;
    for (uint64_t c = 0; c < k; ++c) if (!cnt[c]) ++ne;
    if (ne) {
        restarted = 1;
        uint32_t *rs = (uint32_t *)malloc(ne * sizeof(uint32_t)), *sa = (uint32_t *)malloc(ne * sizeof(uint32_t));
        uint64_t *fw = (uint64_t *)malloc(n * sizeof(uint64_t));
        mco_fixed_weights(cost, n, fw);
        uint64_t e = 0;
        for (uint64_t c = 0; c < k; ++c) if (!cnt[c]) {
            const uint64_t r = mco_fixed_draw(fw, n, mco_wyrand_at(n, draw++));
            rs[e] = (uint32_t)r; sa[e] = (uint32_t)c; ++e;
            memcpy(C + c * d, X + r * d, d * sizeof(double));
        }
        free(fw);
        mco_assign_sqrl2_f64(X, n, d, C, k, asn, cost);
        for (uint64_t i = 0; i < ne; ++i) if (asn[rs[i]] != sa[i]) { asn[rs[i]] = sa[i]; cost[rs[i]] = 0.; }
        free(rs); free(sa);
    }
    free(cnt);
    mco_centroids_dense_f64(X, n, d, asn, w, k, MCO_MEAN_CHUNK, C);
    return restarted;
}

int mco_lloyd_f64(const double *X, uint64_t n, uint64_t d, double *C, uint64_t k, const double *w, double eps,
                  uint64_t maxiter, uint32_t *asn, double *cost, double *initcost, double *finalcost,
                  uint64_t *numiter) {
    mco_assign_sqrl2_f64(X, n, d, C, k, asn, cost);
    double fcost = mco_fixed_total(cost, w, n);
    *initcost = fcost;
    if (fcost == 0.) { *finalcost = 0.; *numiter = 0; return 0; }
    double *Cc = (double *)malloc(k * d * sizeof(double));
    memcpy(Cc, C, k * d * sizeof(double));
    uint64_t iter = 0;
    int rc = 0;
    for (;;) {
        const int res = centroids_with_restart_f64(X, n, d, Cc, k, w, asn, cost);
        int nan = 0;
        for (uint64_t q = 0; q < k * d; ++q) if (isnan(Cc[q])) { nan = 1; break; }
        if (nan) { rc = -1; break; }
        mco_assign_sqrl2_f64(X, n, d, Cc, k, asn, cost);
        const double newcost = mco_fixed_total(cost, w, n);
        if (newcost > fcost && !res) {
            mco_assign_sqrl2_f64(X, n, d, C, k, asn, cost);
            break;
        }
        memcpy(C, Cc, k * d * sizeof(double));
        ++iter;
        const double old = fcost;
        fcost = newcost;
        if (old - newcost < eps * (newcost > old ? newcost : old) || iter > maxiter) break;
    }
    free(Cc);
    *finalcost = fcost; *numiter = iter;
    return rc;
}
