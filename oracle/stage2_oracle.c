/*
 * Stage-2 oracle — CPU restatement of crates/pystamps-core/src/native_stage2.rs.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker / timed CPU baseline.  Nothing in
 * the product path links or calls it.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off, no fast-math), so
 * every f32/f64 operation rounds exactly as written, like rustc (which never
 * contracts a*b+c into an FMA).  Transcendentals are glibc's sinf/cosf/hypotf/
 * atan2f (f32) and sin/cos/hypot/atan2 (f64): the libm that Rust std's
 * f32/f64 methods call on x86_64-linux-gnu.
 *
 * Third-party arithmetic the reference takes from crates.io and that is absent
 * here: rustfft 6.4.1 (Cargo.lock:801-802) for the CLAP f64 FFTs; this file uses
 * its own radix-2 FFT (pinned against a brute-force DFT in tests); num-complex
 * 0.4.6 for complex mul/norm/arg, restated literally below.
 *
 * Pinning (what checks each function; tests/test_oracle_golden.py = G,
 * tests/test_oracle_pins.py = P):
 *   rng_uniform                 G: golden values, native_stage2.rs:2463-2484
 *   psquare_weighting           G: golden values, native_stage2.rs:2486-2509;
 *                               tail/rounding cases tests/test_stage2_trial_wraps.py:55-108
 *   hist_with_centers           G: golden case, native_stage2.rs:2511-2519
 *   random_coherence_histogram  G: golden 100-bin counts, native_stage2.rs:2521-2559;
 *                               small-baseline branch composed by numpy (exact counts)
 *   trial_values                G: golden, tests/test_stage2_ported.py:703-707
 *   grid_indices                G: golden boundary, tests/test_stage2_trial_wraps.py:111-127
 *   low_pass, window weight     G: closed forms (native_stage2.rs:1014-1037, 2013-2043)
 *   fft                         G: brute-force DFT
 *   clap                        P: the whole gain path (|F|, fftshift, 7x7 gausswin
 *                               smoothing, ifftshift, median, (H/m)^alpha - 1 clamped,
 *                               beta blend with low_pass, inverse FFT, window weights,
 *                               edge-window shifts, NaN cells) against numpy.fft +
 *                               scipy.signal.convolve2d + np.median, alpha 1 and 0.7,
 *                               asymmetric low_pass; beta = 0 against a direct
 *                               circular-convolution sum by ifft2(low_pass);
 *                               G: all-pass, shift equivariance, linearity, zero, empty
 *   fill                        P: ph_nm = e^{+i bp K} cancels the ramp (output = w);
 *                               NaN K => NaN
 *   accumulate_grid             G: brute-force per-cell sums; f32 cancellation case
 *                               tests/test_stage2_ported.py:864-907
 *   extract                     P: own cell's value, direction kept, |.| = 1, zeros kept
 *   rms_difference              P: numpy sqrt(mean(d^2))
 *   snr_weighting               P: closed forms (2 sqrt 2, sqrt(2/3), sigma = 0)
 *   topofit / refine            G: noise-free closed forms (K on a trial node => exact
 *                               trial, coh=1, C recovered), off-node refine, brute-force
 *                               FP64 scan
 *   whole loop                  G: 3-PS fixture known outcome (native_stage2.rs:2698-2769);
 *                               P: two iterations composed in numpy from the pinned steps
 *                               (RMS, Nr scaling, P-square) equal the loop's state
 * Each P pin is shown to fail on a planted mistake (ORA_MUTANT builds below).
 */
#include "stage2_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORA_PI 3.14159265358979323846

/* Planted mistakes, compiled in only by tests/test_oracle_pins.py (-DORA_MUTANT=n)
 * to show that the pin tests catch each of them.  0 (the default): none. */
#ifndef ORA_MUTANT
#define ORA_MUTANT 0
#endif
#define ORA_MUT_SMOOTH_FIRST 1    /* 7x7 smoothing before the fftshift */
#define ORA_MUT_MEDIAN_UPPER 2    /* median = upper middle value only */
#define ORA_MUT_LP_TRANSPOSED 3   /* low_pass read column-major */
#define ORA_MUT_NO_ISHIFT 4       /* ifftshift left out */
#define ORA_MUT_SHIFT_OFF1 5      /* fftshift by n/2 - 1 */
#define ORA_MUT_NO_CLAMP 6        /* negative gain not clamped to 0 */
#define ORA_MUT_WEIGHT_UNSHIFTED 7 /* edge window weight without the shift */
#define ORA_MUT_FILL_SIGN 8       /* fill ramp e^{+i bp K} */
#define ORA_MUT_EXTRACT_NO_NORM 9 /* gathered phase not normalised */
#define ORA_MUT_SNR_NO_HALF 10    /* sigma without the 1/2 */
#define ORA_MUT_RMS_NO_MEAN 11    /* sqrt of the sum, not the mean */
#define ORA_MUT_ALPHA_IGNORED 12  /* (H/m)^alpha taken as H/m */
#define ORA_MUT_NR_UNSCALED 13    /* loop: Nr not rescaled before P-square */
#define COH_BIN_COUNT 100

/* native_stage2.rs:33-113: audited MATLAB interp(.,10) FIR taps. */
static const double INTERP10_TAPS[79] = {
    -7.92925708923623813e-05, -1.73293349766172333e-04, -3.00248431516961407e-04,
    -4.48605550383722462e-04, -5.95835452295412349e-04, -7.08759204842826507e-04,
    -7.46231412934915406e-04, -6.64375214040318490e-04, -4.24121369705182660e-04,
    -3.01053961794528480e-07, 6.08939001185030554e-04,  1.37399755086099230e-03,
    2.22872001918483081e-03,  3.06999955141492101e-03,  3.76333143346069959e-03,
    4.15488007985555935e-03,  4.08970016389962471e-03,  3.43472424182160452e-03,
    2.10413493213892910e-03,  8.39421246326618742e-05,  -2.54786253685349538e-03,
    -5.60918261221970971e-03, -8.81278723412589476e-03, -1.17783753398490003e-02,
    -1.40582499688596801e-02, -1.51753768689126753e-02, -1.46708218626883501e-02,
    -1.21559904304523667e-02, -7.36396963079587880e-03, -1.93792008491816548e-04,
    9.25827682864269544e-03,  2.06852593184593814e-02,  3.35768644427044682e-02,
    4.72485589714582085e-02,  6.08905923193367796e-02,  7.36328154139928304e-02,
    8.46190871379102622e-02,  9.30836882813813049e-02,  9.84216619119140101e-02,
    1.00245459871706619e-01,  9.84216619119134828e-02,  9.30836882813811800e-02,
    8.46190871379124271e-02,  7.36328154139975766e-02,  6.08905923193391457e-02,
    4.72485589714637388e-02,  3.35768644427109145e-02,  2.06852593184574732e-02,
    9.25827682864464527e-03,  -1.93792008489998774e-04, -7.36396963079247527e-03,
    -1.21559904304541604e-02, -1.46708218626876996e-02, -1.51753768689126579e-02,
    -1.40582499688608267e-02, -1.17783753398508096e-02, -8.81278723413165405e-03,
    -5.60918261221661496e-03, -2.54786253685826761e-03, 8.39421246353382272e-05,
    2.10413493213411785e-03,  3.43472424182329154e-03,  4.08970016390024574e-03,
    4.15488007985538154e-03,  3.76333143346146634e-03,  3.06999955141576842e-03,
    2.22872001918783102e-03,  1.37399755086213202e-03,  6.08939001188346478e-04,
    -3.01053965413812172e-07, -4.24121369703597340e-04, -6.64375214039929912e-04,
    -7.46231412935480926e-04, -7.08759204842192465e-04, -5.95835452295058465e-04,
    -4.48605550383040716e-04, -3.00248431517695195e-04, -1.73293349768269614e-04,
    -7.92925708931989381e-05,
};

static int nthreads(int threads) {
#ifdef _OPENMP
    return threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
    return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* MATLAB v5 uniform generator (native_stage2.rs:1785-1873).                 */

typedef struct {
    int index;
    double borrow;
    uint32_t j;
    double state[32];
} mv5_rng;

static uint32_t mv5_randint32(uint32_t v) {
    v ^= v << 13;
    v ^= v >> 17;
    v ^= v << 5;
    return v;
}

static const double MV5_ULP = 1.1102230246251565e-16;

static void mv5_init(mv5_rng *r, uint32_t seed) {
    uint32_t j = seed == 0 ? (1u << 31) : seed;
    r->index = 0;
    r->borrow = 0.0;
    r->j = j;
    uint32_t jj = j;
    for (int k = 0; k < 32; ++k) {
        uint64_t x = 0;
        for (int b = 0; b < 53; ++b) {
            jj = mv5_randint32(jj);
            x = (x << 1) | (uint64_t)((jj >> 19) & 1u);
        }
        r->state[k] = (double)x * MV5_ULP;
    }
}

static double mv5_randbits(mv5_rng *r, double value) {
    uint32_t jlo = r->j;
    uint32_t jhi = mv5_randint32(jlo);
    r->j = jhi;
    const uint64_t mask52 = (1ull << 52) - 1;
    uint64_t mask = (((uint64_t)jhi << 32) & mask52) ^ (uint64_t)jlo;
    uint64_t mant;
    int ex;
    if (value == 0.0) {
        mant = 0;
        ex = 0;
    } else {
        uint64_t bits;
        memcpy(&bits, &value, 8);
        int exp_bits = (int)((bits >> 52) & 0x7ff);
        uint64_t mb = bits & mask52;
        if (exp_bits == 0) {
            int leading = 63 - __builtin_clzll(mb);
            ex = leading - 1073;
            mant = mb << (52 - leading);
        } else {
            ex = exp_bits - 1022;
            mant = (1ull << 52) | mb;
        }
    }
    return ldexp((double)(mant ^ mask), ex - 53);
}

static double mv5_next(mv5_rng *r) {
    double v = r->state[(r->index + 20) & 31] - r->state[(r->index + 5) & 31] - r->borrow;
    if (v < 0.0) {
        v += 1.0;
        r->borrow = MV5_ULP;
    } else {
        r->borrow = 0.0;
    }
    r->state[r->index] = v;
    r->index = (r->index + 1) & 31;
    return mv5_randbits(r, v);
}

void ora_rng_uniform(uint32_t seed, int64_t n, double *out) {
    mv5_rng r;
    mv5_init(&r, seed);
    for (int64_t k = 0; k < n; ++k) out[k] = mv5_next(&r);
}

/* ------------------------------------------------------------------------- */
/* Histogram with centres (native_stage2.rs:1538-1640).                      */

static int centers_equal_spacing(const double *c, int64_t n) {
    if (n < 2) return 1;
    double ref = c[1] - c[0];
    double max_center = 0.0;
    for (int64_t k = 0; k < n; ++k) {
        double a = fabs(c[k]);
        if (a > max_center) max_center = a;
    }
    double tol = DBL_EPSILON * (max_center > 1.0 ? max_center : 1.0);
    for (int64_t k = 0; k + 1 < n; ++k) {
        if (!(fabs((c[k + 1] - c[k]) - ref) <= tol)) return 0;
    }
    return 1;
}

static int64_t center_index(double v, const double *c, int64_t n, int equal) {
    if (!isfinite(v) || n == 0) return -1;
    if (n == 1) return 0;
    if (equal) {
        double d = n < 3 ? 1.0 : (c[n - 1] - c[0]) / (double)(n - 1);
        if (d == 0.0) return 0;
        double cutoff0 = (c[0] + c[1]) / 2.0;
        double a = ceil((v - cutoff0) / d);
        if (a < 0.0) a = 0.0;
        if (a > (double)(n - 1)) a = (double)(n - 1);
        return (int64_t)a;
    }
    /* partition_point(mid < v), clamped to n-1 */
    int64_t k = 0;
    while (k < n - 1 && (c[k] + c[k + 1]) / 2.0 < v) ++k;
    return k;
}

/* Distance of a coherence to the nearest discrete choice boundary of the
 * P-square update: the lookup's round-half points (n + 1/2) / 1000
 * (native_stage2.rs:1674-1682) and the histogram's bin boundaries
 * (:1593-1606).  Diagnostic only (tests tell near-ties from errors with it). */
static double decision_distance(double v, const double *c, int64_t n) {
    double x = 1000.0 * v;
    double dist = fabs(x - floor(x) - 0.5) * 1e-3;
    if (n < 2 || !isfinite(v)) return dist;
    if (centers_equal_spacing(c, n)) {
        double d = n < 3 ? 1.0 : (c[n - 1] - c[0]) / (double)(n - 1);
        if (d != 0.0) {
            double y = (v - (c[0] + c[1]) / 2.0) / d;
            double e = fabs(y - rint(y)) * d;
            if (e < dist) dist = e;
        }
    } else {
        for (int64_t k = 0; k + 1 < n; ++k) {
            double e = fabs(v - (c[k] + c[k + 1]) / 2.0);
            if (e < dist) dist = e;
        }
    }
    return dist;
}

void ora_hist_with_centers(const double *values, int64_t n, const double *centers, int64_t nc,
                           double *out) {
    for (int64_t k = 0; k < nc; ++k) out[k] = 0.0;
    if (nc == 0) return;
    if (nc == 1) {
        for (int64_t k = 0; k < n; ++k)
            if (isfinite(values[k])) out[0] += 1.0;
        return;
    }
    int equal = centers_equal_spacing(centers, nc);
    for (int64_t k = 0; k < n; ++k) {
        int64_t ix = center_index(values[k], centers, nc, equal);
        if (ix >= 0) out[ix] += 1.0;
    }
}

/* ------------------------------------------------------------------------- */
/* P-square weighting (native_stage2.rs:1642-1783).                          */

static void lfilter(const double *kernel, int64_t nk, const double *x, int64_t n, double *y) {
    for (int64_t ix = 0; ix < n; ++ix) {
        double s = 0.0;
        for (int64_t k = 0; k < nk; ++k)
            if (ix >= k) s += kernel[k] * x[ix - k];
        y[ix] = s;
    }
}

static double round_half_away(double v) { return v >= 0.0 ? floor(v + 0.5) : ceil(v - 0.5); }

void ora_psquare_weighting(const double *nr, const double *na, int64_t nbins,
                           int64_t low_coh_thresh, double nr_max_nz_ix, const double *coh_ps,
                           int64_t n, double *out_weight, double *out_prand_hi) {
    double *prand = (double *)malloc(sizeof(double) * (size_t)nbins);
    for (int64_t k = 0; k < nbins; ++k) {
        double denom = na[k] == 0.0 ? 1.0 : na[k];
        double v = nr[k] / denom;
        prand[k] = (isnan(v) || v > 1.0) ? 1.0 : v; /* f64::min(v, 1.0) */
    }
    for (int64_t k = 0; k < (low_coh_thresh < nbins ? low_coh_thresh : nbins); ++k) prand[k] = 1.0;
    int64_t z0 = (int64_t)nr_max_nz_ix;
    if (nr_max_nz_ix < 0.0 || isnan(nr_max_nz_ix)) z0 = 0;
    for (int64_t k = (z0 < nbins ? z0 : nbins); k < nbins; ++k) prand[k] = 0.0;

    /* gausswin(7, 2.5) */
    double win[7], win_sum = 0.0;
    double std = (7.0 - 1.0) / (2.0 * 2.5);
    for (int k = 0; k < 7; ++k) {
        double x = (double)k - (7.0 - 1.0) / 2.0;
        win[k] = exp(-0.5 * (x / std) * (x / std));
    }
    for (int k = 0; k < 7; ++k) win_sum += win[k];
    int64_t np = nbins + 7;
    double *padded = (double *)malloc(sizeof(double) * (size_t)np);
    double *filtered = (double *)malloc(sizeof(double) * (size_t)np);
    for (int k = 0; k < 7; ++k) padded[k] = 1.0;
    for (int64_t k = 0; k < nbins; ++k) padded[7 + k] = prand[k];
    lfilter(win, 7, padded, np, filtered);
    for (int64_t k = 0; k < nbins; ++k) prand[k] = filtered[k + 7] / win_sum;

    /* interp([1, prand], 10) with the fixed 79 taps, delay 39 */
    int64_t nin = nbins + 1, factor = 10, delay = 39;
    int64_t nexp = nin * factor + delay;
    double *expanded = (double *)calloc((size_t)nexp, sizeof(double));
    double *filt2 = (double *)malloc(sizeof(double) * (size_t)nexp);
    expanded[0] = 1.0;
    for (int64_t k = 0; k < nbins; ++k) expanded[(k + 1) * factor] = prand[k];
    lfilter(INTERP10_TAPS, 79, expanded, nexp, filt2);
    int64_t nhi = nin * factor;
    int64_t keep = nhi >= 9 ? nhi - 9 : 0;
    double *prand_hi = (double *)malloc(sizeof(double) * (size_t)(keep > 0 ? keep : 1));
    for (int64_t k = 0; k < keep; ++k) prand_hi[k] = filt2[delay + k] * (double)factor;
    if (out_prand_hi)
        for (int64_t k = 0; k < keep; ++k) out_prand_hi[k] = prand_hi[k];
    double hi_max = keep > 0 ? (double)(keep - 1) : 0.0;
    for (int64_t p = 0; p < n; ++p) {
        double r = round_half_away(coh_ps[p] * 1000.0);
        if (r < 0.0) r = 0.0;
        if (r > hi_max) r = hi_max;
        int64_t ix = isnan(r) ? 0 : (int64_t)r;
        double pv = prand_hi[ix];
        out_weight[p] = (1.0 - pv) * (1.0 - pv);
    }
    free(prand);
    free(padded);
    free(filtered);
    free(expanded);
    free(filt2);
    free(prand_hi);
}

/* ------------------------------------------------------------------------- */
/* Prep helpers.                                                             */

int64_t ora_trial_values(double n_trial_wraps, double *out, int64_t cap) {
    int64_t tn = (int64_t)ceil(8.0 * n_trial_wraps);
    int64_t n = 2 * tn + 1;
    if (n < 0) n = 0;
    for (int64_t k = 0; k < n && k < cap; ++k) out[k] = (double)(k - tn);
    return n;
}

void ora_low_pass(double clap_win, double low_pass_wavelength, double grid_size, double *out,
                  int64_t *out_dim) {
    double r = round(clap_win);
    int64_t n = (int64_t)(r < 1.0 ? 1.0 : r);
    double freq0 = 1.0 / low_pass_wavelength;
    double *butter = (double *)malloc(sizeof(double) * (size_t)n);
    for (int64_t k = 0; k < n; ++k) {
        double fi = ((double)k - (double)n / 2.0) / (grid_size * (double)n);
        double q = fi / freq0;
        /* f64::powi(10): repeated multiplication */
        double q2 = q * q, q4 = q2 * q2, q8 = q4 * q4, q10 = q8 * q2;
        butter[k] = 1.0 / (1.0 + q10);
    }
    int64_t sh = n / 2;
    for (int64_t r0 = 0; r0 < n; ++r0)
        for (int64_t c0 = 0; c0 < n; ++c0) {
            int64_t sr = (r0 + sh) % n, sc = (c0 + sh) % n;
            out[r0 * n + c0] = butter[sr] * butter[sc];
        }
    free(butter);
    if (out_dim) *out_dim = n;
}

void ora_grid_indices(const float *xy, int64_t n_ps, double grid_size, float *out_ij,
                      int64_t *out_n_i, int64_t *out_n_j) {
    float min_x = INFINITY, min_y = INFINITY;
    for (int64_t r = 0; r < n_ps; ++r) {
        float x = xy[r * 3 + 1], y = xy[r * 3 + 2];
        min_x = fminf(min_x, x);
        min_y = fminf(min_y, y);
    }
    float scale = (float)grid_size;
    float eps = 1.0e-6f;
    int32_t max_i = 1, max_j = 1;
    for (int64_t r = 0; r < n_ps; ++r) {
        float fi = ceilf((xy[r * 3 + 2] - min_y + eps) / scale);
        float fj = ceilf((xy[r * 3 + 1] - min_x + eps) / scale);
        if (!(fi >= 1.0f)) fi = 1.0f; /* f32::max(1.0): NaN -> 1 */
        if (!(fj >= 1.0f)) fj = 1.0f;
        int32_t i = (int32_t)fi, j = (int32_t)fj;
        out_ij[r * 2] = (float)i;
        out_ij[r * 2 + 1] = (float)j;
        if (i > max_i) max_i = i;
        if (j > max_j) max_j = j;
    }
    if (max_i > 1 || max_j > 1) {
        for (int64_t r = 0; r < n_ps; ++r) {
            if (max_i > 1 && (int32_t)out_ij[r * 2] == max_i) out_ij[r * 2] = (float)(max_i - 1);
            if (max_j > 1 && (int32_t)out_ij[r * 2 + 1] == max_j)
                out_ij[r * 2 + 1] = (float)(max_j - 1);
        }
    }
    int64_t ni = 1, nj = 1;
    for (int64_t r = 0; r < n_ps; ++r) {
        if ((int64_t)out_ij[r * 2] > ni) ni = (int64_t)out_ij[r * 2];
        if ((int64_t)out_ij[r * 2 + 1] > nj) nj = (int64_t)out_ij[r * 2 + 1];
    }
    *out_n_i = ni;
    *out_n_j = nj;
}

void ora_clap_window_weight(int64_t n_win, double *out) {
    int64_t half = n_win / 2;
    for (int64_t r = 0; r < n_win * n_win; ++r) out[r] = 0.0;
    for (int64_t r = 0; r < half; ++r)
        for (int64_t c = 0; c < half; ++c) {
            double q = (double)r + (double)c + 1.0e-6;
            double qm = (double)r + (double)(half - 1 - c) + 1.0e-6;
            /* top half rows r; bottom half mirrors rows */
            out[r * n_win + c] = q;
            out[r * n_win + half + c] = qm;
        }
    for (int64_t r = 0; r < half; ++r)
        for (int64_t c = 0; c < n_win; ++c)
            out[(half + r) * n_win + c] = out[(half - 1 - r) * n_win + c];
}

double ora_rms_difference(const double *a, const double *b, int64_t n) {
    double s = 0.0;
    for (int64_t k = 0; k < n; ++k) {
        double d = a[k] - b[k];
        s += d * d;
    }
    double denom = (double)(n > 1 ? n : 1);
#if ORA_MUTANT == ORA_MUT_RMS_NO_MEAN
    denom = 1.0;
#endif
    return sqrt(s / denom);
}

/* ------------------------------------------------------------------------- */
/* FFT (f64) — radix-2 DIT for power-of-two sizes, direct DFT otherwise.     */

typedef struct {
    int64_t n;
    int pow2;
    double *wr, *wi;   /* e^{-2 pi i k / n}, k < n */
    int64_t *rev;
} fft_plan;

static void fft_plan_init(fft_plan *p, int64_t n) {
    p->n = n;
    p->pow2 = n > 0 && (n & (n - 1)) == 0;
    p->wr = (double *)malloc(sizeof(double) * (size_t)n);
    p->wi = (double *)malloc(sizeof(double) * (size_t)n);
    p->rev = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    for (int64_t k = 0; k < n; ++k) {
        double a = 2.0 * ORA_PI * (double)k / (double)n;
        p->wr[k] = cos(a);
        p->wi[k] = -sin(a);
    }
    int bits = 0;
    while ((1ll << bits) < n) ++bits;
    for (int64_t k = 0; k < n; ++k) {
        int64_t r = 0;
        for (int b = 0; b < bits; ++b)
            if (k & (1ll << b)) r |= 1ll << (bits - 1 - b);
        p->rev[k] = p->pow2 ? r : k;
    }
}

static void fft_plan_free(fft_plan *p) {
    free(p->wr);
    free(p->wi);
    free(p->rev);
}

/* in-place 1-D transform of n complex values at stride s (in complex units) */
static void fft_1d(const fft_plan *p, double *re, double *im, int64_t s, int inverse, double *tr,
                   double *ti) {
    int64_t n = p->n;
    double sg = inverse ? -1.0 : 1.0;
    if (!p->pow2) {
        for (int64_t k = 0; k < n; ++k) {
            double ar = 0.0, ai = 0.0;
            for (int64_t m = 0; m < n; ++m) {
                int64_t w = (k * m) % n;
                double cr = p->wr[w], ci = sg * p->wi[w];
                double xr = re[m * s], xi = im[m * s];
                ar += xr * cr - xi * ci;
                ai += xr * ci + xi * cr;
            }
            tr[k] = ar;
            ti[k] = ai;
        }
        for (int64_t k = 0; k < n; ++k) {
            re[k * s] = tr[k];
            im[k * s] = ti[k];
        }
        return;
    }
    for (int64_t k = 0; k < n; ++k) {
        tr[p->rev[k]] = re[k * s];
        ti[p->rev[k]] = im[k * s];
    }
    for (int64_t len = 2; len <= n; len <<= 1) {
        int64_t half = len >> 1, step = n / len;
        for (int64_t b = 0; b < n; b += len)
            for (int64_t k = 0; k < half; ++k) {
                double wr = p->wr[k * step], wi = sg * p->wi[k * step];
                double ur = tr[b + k], ui = ti[b + k];
                double xr = tr[b + k + half], xi = ti[b + k + half];
                double vr = xr * wr - xi * wi, vi = xr * wi + xi * wr;
                tr[b + k] = ur + vr;
                ti[b + k] = ui + vi;
                tr[b + k + half] = ur - vr;
                ti[b + k + half] = ui - vi;
            }
    }
    for (int64_t k = 0; k < n; ++k) {
        re[k * s] = tr[k];
        im[k * s] = ti[k];
    }
}

static void fft2(const fft_plan *p, double *re, double *im, int inverse, double *tr, double *ti) {
    int64_t n = p->n;
    for (int64_t r = 0; r < n; ++r) fft_1d(p, re + r * n, im + r * n, 1, inverse, tr, ti);
    for (int64_t c = 0; c < n; ++c) fft_1d(p, re + c, im + c, n, inverse, tr, ti);
}

void ora_fft32_2d(double *re_im, int inverse) {
    fft_plan p;
    fft_plan_init(&p, 32);
    double re[1024], im[1024], tr[32], ti[32];
    for (int k = 0; k < 1024; ++k) {
        re[k] = re_im[2 * k];
        im[k] = re_im[2 * k + 1];
    }
    fft2(&p, re, im, inverse, tr, ti);
    for (int k = 0; k < 1024; ++k) {
        re_im[2 * k] = re[k];
        re_im[2 * k + 1] = im[k];
    }
    fft_plan_free(&p);
}

/* ------------------------------------------------------------------------- */
/* CLAP on one interferogram plane (native_stage2.rs:742-1184), f64 inside.  */

typedef struct {
    int64_t i1, j1, row_shift, col_shift;
} clap_win_t;

static int cmp_f64(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

typedef struct {
    int64_t n_i, n_j, n_win, n_win_ex, n_windows;
    double alpha, beta;
    int alpha_is_one;
    const double *low_pass;
    double kernel[7];
    double *weight; /* n_win x n_win */
    clap_win_t *windows;
} clap_ctx;

static int clap_ctx_init(clap_ctx *c, const ora_stage2_in *in) {
    c->n_i = in->n_i;
    c->n_j = in->n_j;
    c->n_win = in->clap_window;
    c->n_win_ex = in->n_win_ex;
    c->alpha = in->clap_alpha;
    c->beta = in->clap_beta;
    c->alpha_is_one = fabs(in->clap_alpha - 1.0) <= DBL_EPSILON;
    c->low_pass = in->low_pass;
    c->weight = NULL;
    c->windows = NULL;
    c->n_windows = 0;
    int64_t n_inc = c->n_win / 4 > 1 ? c->n_win / 4 : 1;
    int64_t nwi = (int64_t)ceil((double)c->n_i / (double)n_inc) - 3;
    int64_t nwj = (int64_t)ceil((double)c->n_j / (double)n_inc) - 3;
    if (nwi <= 0 || nwj <= 0) return 0;
    double std = (7.0 - 1.0) / (2.0 * 2.5);
    for (int k = 0; k < 7; ++k) {
        double x = (double)k - 3.0;
        c->kernel[k] = exp(-0.5 * (x / std) * (x / std));
    }
    c->weight = (double *)malloc(sizeof(double) * (size_t)(c->n_win * c->n_win));
    ora_clap_window_weight(c->n_win, c->weight);
    c->windows = (clap_win_t *)malloc(sizeof(clap_win_t) * (size_t)(nwi * nwj));
    for (int64_t a = 0; a < nwi; ++a) {
        int64_t i1 = a * n_inc, i2 = i1 + c->n_win, rs = 0;
        if (i2 > c->n_i) {
            rs = i2 - c->n_i;
            i1 = c->n_i - c->n_win;
        }
        for (int64_t b = 0; b < nwj; ++b) {
            int64_t j1 = b * n_inc, j2 = j1 + c->n_win, cs = 0;
            if (j2 > c->n_j) {
                cs = j2 - c->n_j;
                j1 = c->n_j - c->n_win;
            }
            clap_win_t w = {i1, j1, rs, cs};
            c->windows[c->n_windows++] = w;
        }
    }
    return 1;
}

static void clap_ctx_free(clap_ctx *c) {
    free(c->weight);
    free(c->windows);
}

/* plane_re/im: [n_i*n_j] f64 input (NaN already possible), out: accum f64 */
static void clap_one_ifg(const clap_ctx *c, const double *g_re, const double *g_im, double *acc_re,
                         double *acc_im) {
    int64_t n = c->n_win_ex, nn = n * n, nw = c->n_win, n_j = c->n_j;
    fft_plan p;
    fft_plan_init(&p, n);
    double *br = (double *)malloc(sizeof(double) * (size_t)nn);
    double *bi = (double *)malloc(sizeof(double) * (size_t)nn);
    double *habs = (double *)malloc(sizeof(double) * (size_t)nn);
    double *hsh = (double *)malloc(sizeof(double) * (size_t)nn);
    double *htmp = (double *)malloc(sizeof(double) * (size_t)nn);
    double *hflt = (double *)malloc(sizeof(double) * (size_t)nn);
    double *hmed = (double *)malloc(sizeof(double) * (size_t)nn);
    double *tr = (double *)malloc(sizeof(double) * (size_t)n);
    double *ti = (double *)malloc(sizeof(double) * (size_t)n);
    double inv_scale = 1.0 / (double)(n * n);
    for (int64_t k = 0; k < c->n_i * c->n_j; ++k) acc_re[k] = acc_im[k] = 0.0;
    for (int64_t w = 0; w < c->n_windows; ++w) {
        const clap_win_t *win = &c->windows[w];
        for (int64_t k = 0; k < nn; ++k) br[k] = bi[k] = 0.0;
        for (int64_t li = 0; li < nw; ++li)
            for (int64_t lj = 0; lj < nw; ++lj) {
                int64_t src = (win->i1 + li) * n_j + (win->j1 + lj);
                double vr = g_re[src], vi = g_im[src];
                if (isnan(vr) || isnan(vi)) vr = vi = 0.0;
                br[li * n + lj] = vr;
                bi[li * n + lj] = vi;
            }
        fft2(&p, br, bi, 0, tr, ti);
        for (int64_t k = 0; k < nn; ++k) habs[k] = hypot(br[k], bi[k]);
        /* fftshift by (n/2, n/2) */
        int64_t fsh = n / 2;
#if ORA_MUTANT == ORA_MUT_SHIFT_OFF1
        fsh = n / 2 - 1;
#endif
        for (int64_t r = 0; r < n; ++r)
            for (int64_t q = 0; q < n; ++q)
                hsh[r * n + q] = habs[((r + fsh) % n) * n + ((q + fsh) % n)];
#if ORA_MUTANT == ORA_MUT_SMOOTH_FIRST
        memcpy(hsh, habs, sizeof(double) * (size_t)nn);
#endif
        /* separable 'same' convolution, zero fill */
        for (int64_t r = 0; r < n; ++r)
            for (int64_t q = 0; q < n; ++q) {
                double s = 0.0;
                for (int64_t kc = 0; kc < 7; ++kc) {
                    int64_t sc = q + kc - 3;
                    if (sc >= 0 && sc < n) s += hsh[r * n + sc] * c->kernel[kc];
                }
                htmp[r * n + q] = s;
            }
        for (int64_t r = 0; r < n; ++r)
            for (int64_t q = 0; q < n; ++q) {
                double s = 0.0;
                for (int64_t kr = 0; kr < 7; ++kr) {
                    int64_t sr = r + kr - 3;
                    if (sr >= 0 && sr < n) s += htmp[sr * n + q] * c->kernel[kr];
                }
                hflt[r * n + q] = s;
            }
        /* ifftshift by (ceil(n/2), ceil(n/2)) */
        int64_t ish = (n + 1) / 2;
        for (int64_t r = 0; r < n; ++r)
            for (int64_t q = 0; q < n; ++q)
                hsh[r * n + q] = hflt[((r + ish) % n) * n + ((q + ish) % n)];
#if ORA_MUTANT == ORA_MUT_NO_ISHIFT
        memcpy(hsh, hflt, sizeof(double) * (size_t)nn);
#elif ORA_MUTANT == ORA_MUT_SMOOTH_FIRST
        /* smoothed unshifted |F|, then the two shifts cancel */
        memcpy(hsh, hflt, sizeof(double) * (size_t)nn);
#endif
        memcpy(hmed, hsh, sizeof(double) * (size_t)nn);
        qsort(hmed, (size_t)nn, sizeof(double), cmp_f64);
        double med = (nn % 2 == 0) ? (hmed[nn / 2 - 1] + hmed[nn / 2]) / 2.0 : hmed[nn / 2];
#if ORA_MUTANT == ORA_MUT_MEDIAN_UPPER
        med = hmed[nn / 2];
#endif
        for (int64_t k = 0; k < nn; ++k) {
            double v = hsh[k];
            if (med != 0.0) v /= med;
#if ORA_MUTANT == ORA_MUT_ALPHA_IGNORED
            v = v - 1.0;
#else
            v = c->alpha_is_one ? v - 1.0 : pow(v, c->alpha) - 1.0;
#endif
#if ORA_MUTANT != ORA_MUT_NO_CLAMP
            if (v < 0.0) v = 0.0;
#endif
#if ORA_MUTANT == ORA_MUT_LP_TRANSPOSED
            double g = v * c->beta + c->low_pass[(k % n) * n + k / n];
#else
            double g = v * c->beta + c->low_pass[k];
#endif
            br[k] *= g;
            bi[k] *= g;
        }
        fft2(&p, br, bi, 1, tr, ti);
        for (int64_t li = 0; li < nw; ++li)
            for (int64_t lj = 0; lj < nw; ++lj) {
                if (li < win->row_shift || lj < win->col_shift) continue;
                double wt = c->weight[(li - win->row_shift) * nw + (lj - win->col_shift)];
#if ORA_MUTANT == ORA_MUT_WEIGHT_UNSHIFTED
                wt = c->weight[li * nw + lj];
#endif
                if (wt == 0.0) continue;
                double f = wt * inv_scale;
                int64_t dst = (win->i1 + li) * n_j + (win->j1 + lj);
                acc_re[dst] += br[li * n + lj] * f;
                acc_im[dst] += bi[li * n + lj] * f;
            }
    }
    free(br);
    free(bi);
    free(habs);
    free(hsh);
    free(htmp);
    free(hflt);
    free(hmed);
    free(tr);
    free(ti);
    fft_plan_free(&p);
}

/* ------------------------------------------------------------------------- */
/* Random-coherence histogram (native_stage2.rs:1310-1512).                  */

static int64_t argmax_first_f64(const double *v, int64_t n) {
    int64_t best = 0;
    double bv = n > 0 ? v[0] : -INFINITY;
    for (int64_t k = 1; k < n; ++k)
        if (v[k] > bv) {
            bv = v[k];
            best = k;
        }
    return best;
}

double ora_topofit_phase_row_coh(const double *phase, const double *bperp, int64_t n_col,
                                 const double *trial_values, int64_t n_trial) {
    if (n_col == 0) return NAN;
    double mn = INFINITY, mx = -INFINITY;
    for (int64_t k = 0; k < n_col; ++k) {
        mn = fmin(mn, bperp[k]);
        mx = fmax(mx, bperp[k]);
    }
    double range = mx - mn;
    if (range == 0.0) range = 1.0;
    double denom = (double)n_col;
    double coh_stack[512];
    double *coh = n_trial <= 512 ? coh_stack : (double *)malloc(sizeof(double) * (size_t)n_trial);
    for (int64_t t = 0; t < n_trial; ++t) {
        double sr = 0.0, si = 0.0;
        for (int64_t i = 0; i < n_col; ++i) {
            double ph = phase[i] - (bperp[i] / range) * (ORA_PI / 4.0) * trial_values[t];
            sr += cos(ph);
            si += sin(ph);
        }
        coh[t] = hypot(sr, si) / denom;
    }
    int64_t ti = argmax_first_f64(coh, n_trial);
    double k0 = (ORA_PI / 4.0) / range * trial_values[ti];
    if (coh != coh_stack) free(coh);
    double ore = 0.0, oim = 0.0;
    for (int64_t i = 0; i < n_col; ++i) {
        double ph = phase[i] - k0 * bperp[i];
        ore += cos(ph);
        oim += sin(ph);
    }
    double cre = ore, cim = -oim;
    double num = 0.0, den = 0.0;
    for (int64_t i = 0; i < n_col; ++i) {
        den += bperp[i] * bperp[i];
        double ph = phase[i] - k0 * bperp[i];
        double cs = cos(ph), sn = sin(ph);
        double pr = cs * cre - sn * cim, pi = cs * cim + sn * cre;
        num += bperp[i] * atan2(pi, pr);
    }
    if (den == 0.0) den = 1.0;
    double k = k0 + num / den;
    double sr = 0.0, si = 0.0;
    for (int64_t i = 0; i < n_col; ++i) {
        double ph = phase[i] - k * bperp[i];
        sr += cos(ph);
        si += sin(ph);
    }
    return hypot(sr, si) / (double)(n_col > 1 ? n_col : 1);
}

static void nr_histogram_of_rows(const double *ph, const double *bperp_nm, int64_t n_ifg,
                                 const double *trial_values, int64_t n_trial,
                                 const double *coh_bins, int64_t n_rows, double *out_nr,
                                 double *out_nr_max_nz_ix, int threads);

void ora_random_coherence_histogram(const double *bperp_nm, int64_t n_ifg,
                                    const double *trial_values, int64_t n_trial,
                                    const double *coh_bins, int64_t n_rows, uint32_t seed,
                                    double *out_nr, double *out_nr_max_nz_ix, int threads) {
    int64_t total = n_rows * n_ifg;
    double *ph = (double *)malloc(sizeof(double) * (size_t)(total > 0 ? total : 1));
    ora_rng_uniform(seed, total, ph);
    for (int64_t k = 0; k < total; ++k) ph[k] *= 2.0 * ORA_PI;
    nr_histogram_of_rows(ph, bperp_nm, n_ifg, trial_values, n_trial, coh_bins, n_rows, out_nr,
                         out_nr_max_nz_ix, threads);
    free(ph);
}

/* Small-baseline branch (native_stage2.rs:1321-1334): n_rows x n_image uniform
 * phases (column-major draws, x 2 pi), interferogram phase = image b - image a
 * (0-based image indices). */
void ora_random_coherence_histogram_sb(const double *bperp_nm, int64_t n_ifg,
                                       const int32_t *image_a, const int32_t *image_b,
                                       int64_t n_image, const double *trial_values,
                                       int64_t n_trial, const double *coh_bins, int64_t n_rows,
                                       uint32_t seed, double *out_nr, double *out_nr_max_nz_ix,
                                       int threads) {
    int64_t nimg = n_rows * n_image;
    double *img = (double *)malloc(sizeof(double) * (size_t)(nimg > 0 ? nimg : 1));
    ora_rng_uniform(seed, nimg, img);
    for (int64_t k = 0; k < nimg; ++k) img[k] *= 2.0 * ORA_PI;
    int64_t total = n_rows * n_ifg;
    double *ph = (double *)malloc(sizeof(double) * (size_t)(total > 0 ? total : 1));
    for (int64_t i = 0; i < n_ifg; ++i)
        for (int64_t r = 0; r < n_rows; ++r)
            ph[r + i * n_rows] = img[r + (int64_t)image_b[i] * n_rows] - img[r + (int64_t)image_a[i] * n_rows];
    nr_histogram_of_rows(ph, bperp_nm, n_ifg, trial_values, n_trial, coh_bins, n_rows, out_nr,
                         out_nr_max_nz_ix, threads);
    free(ph);
    free(img);
}

/* histogram_random_rows (native_stage2.rs:1405-1436) over rows r of the
 * column-major phase matrix ph[r + i * n_rows]. */
static void nr_histogram_of_rows(const double *ph, const double *bperp_nm, int64_t n_ifg,
                                 const double *trial_values, int64_t n_trial,
                                 const double *coh_bins, int64_t n_rows, double *out_nr,
                                 double *out_nr_max_nz_ix, int threads) {
    int eq = centers_equal_spacing(coh_bins, COH_BIN_COUNT);
    int64_t counts[COH_BIN_COUNT];
    memset(counts, 0, sizeof(counts));
#pragma omp parallel num_threads(nthreads(threads))
    {
        int64_t local[COH_BIN_COUNT];
        memset(local, 0, sizeof(local));
        double *row = (double *)malloc(sizeof(double) * (size_t)(n_ifg > 0 ? n_ifg : 1));
#pragma omp for schedule(static)
        for (int64_t r = 0; r < n_rows; ++r) {
            for (int64_t i = 0; i < n_ifg; ++i) row[i] = ph[r + i * n_rows];
            double coh = ora_topofit_phase_row_coh(row, bperp_nm, n_ifg, trial_values, n_trial);
            int64_t ix = center_index(coh, coh_bins, COH_BIN_COUNT, eq);
            if (ix >= 0) local[ix] += 1;
        }
#pragma omp critical
        for (int k = 0; k < COH_BIN_COUNT; ++k) counts[k] += local[k];
        free(row);
    }
    int64_t last = -1;
    for (int k = 0; k < COH_BIN_COUNT; ++k) {
        out_nr[k] = (double)counts[k];
        if (counts[k] > 0) last = k;
    }
    *out_nr_max_nz_ix = last >= 0 ? (double)(last + 1) : 1.0;
}

/* ------------------------------------------------------------------------- */
/* Mode-dependent parts: included twice.                                     */

#define R float
#define SUF 32
#define SIN sinf
#define COS cosf
#define HYPOT hypotf
#define ATAN2 atan2f
#define PI4 ((float)ORA_PI / 4.0f)
#define ROUND_K(x) ((double)(float)(x))
#include "stage2_oracle_mode.inc"
#undef R
#undef SUF
#undef SIN
#undef COS
#undef HYPOT
#undef ATAN2
#undef PI4
#undef ROUND_K

#define R double
#define SUF 64
#define SIN sin
#define COS cos
#define HYPOT hypot
#define ATAN2 atan2
#define PI4 (ORA_PI / 4.0)
#define ROUND_K(x) ((double)(x))
#include "stage2_oracle_mode.inc"
#undef R
#undef SUF
#undef SIN
#undef COS
#undef HYPOT
#undef ATAN2
#undef PI4
#undef ROUND_K

/* ------------------------------------------------------------------------- */
/* Public dispatchers.                                                       */

int ora_fill_phase_weight(const ora_stage2_in *in, const double *k_ps, const double *weighting,
                          void *ph_weight, int mode, int threads) {
    return mode == ORA_MODE_REF64 ? fill_64(in, k_ps, weighting, (double *)ph_weight, threads)
                                  : fill_32(in, k_ps, weighting, (float *)ph_weight, threads);
}

void ora_accumulate_grid(const ora_stage2_in *in, const void *ph_weight, void *ph_grid, int mode) {
    if (mode == ORA_MODE_REF64)
        grid_64(in, (const double *)ph_weight, (double *)ph_grid);
    else
        grid_32(in, (const float *)ph_weight, (float *)ph_grid);
}

void ora_clap_filter_grid_stack(const ora_stage2_in *in, const void *ph_grid, void *ph_filt,
                                int mode, int threads) {
    if (mode == ORA_MODE_REF64)
        clap_64(in, (const double *)ph_grid, (double *)ph_filt, threads);
    else
        clap_32(in, (const float *)ph_grid, (float *)ph_filt, threads);
}

void ora_extract_patch_phase(const ora_stage2_in *in, const void *ph_filt, void *ph_patch,
                             int mode, int threads) {
    if (mode == ORA_MODE_REF64)
        extract_64(in, (const double *)ph_filt, (double *)ph_patch, threads);
    else
        extract_32(in, (const float *)ph_filt, (float *)ph_patch, threads);
}

int ora_fit_rows(const ora_stage2_in *in, const void *ph_patch, double *k_ps, double *c_ps,
                 double *coh_ps, double *n_opt, void *ph_res, int32_t *trial_ix,
                 float *trial_margin, int mode, int threads) {
    if (mode == ORA_MODE_REF64)
        return fit_64(in, (const double *)ph_patch, k_ps, c_ps, coh_ps, n_opt, (double *)ph_res,
                      trial_ix, trial_margin, NULL, threads);
    return fit_32(in, (const float *)ph_patch, k_ps, c_ps, coh_ps, n_opt, (float *)ph_res,
                  trial_ix, trial_margin, NULL, threads);
}

void ora_snr_weighting(const float *amp, const void *ph_res, int64_t n_ps, int64_t n_ifg,
                       double *out, int mode) {
    for (int64_t r = 0; r < n_ps; ++r) {
        double g = 0.0, a2 = 0.0;
        for (int64_t c = 0; c < n_ifg; ++c) {
            int64_t ix = r * n_ifg + c;
            double a = (double)amp[ix];
            double pr = mode == ORA_MODE_REF64 ? ((const double *)ph_res)[ix]
                                               : (double)((const float *)ph_res)[ix];
            g += a * cos(pr);
            a2 += a * a;
        }
        double d = (double)(n_ifg > 1 ? n_ifg : 1);
        g /= d;
        a2 /= d;
        double sigma = sqrt(0.5 * (a2 - g * g));
#if ORA_MUTANT == ORA_MUT_SNR_NO_HALF
        sigma = sqrt(a2 - g * g);
#endif
        out[r] = sigma != 0.0 ? g / sigma : 0.0;
    }
}

int ora_stage2_run(const ora_stage2_in *in, ora_stage2_out *out, int mode, int threads) {
    return mode == ORA_MODE_REF64 ? run_64(in, out, threads) : run_32(in, out, threads);
}
