/*
 * Oracle (test infrastructure only): CPU restatement, in FP64, of the Python
 * pipeline's stage-2 topofit kernels that the pySTAMPS kernel registry serves
 * (stage2_topofit / stage2_topofit_row_invariant /
 * stage2_topofit_coh_row_invariant / stage2_grid_accumulate):
 *   _ps_topofit_single                 pystamps/pipeline/ported.py:2355-2430
 *   _ps_topofit_near_max_trial_indices :2433-2448 (tol 2e-4, :47)
 *   _ps_topofit_select_candidate       :2451-2477
 *   _ps_topofit_refine_candidate       :2480-2513
 *   _stage2_trial_values               :2350-2352
 *   _ps_topofit_batch_row_invariant_coh :2555-2616
 *   _stage2_grid_accumulate_cpu        pystamps/kernels/accelerated.py:183-200
 * Sums run in index order (the FP64 GPU provider does the same); numpy's
 * pairwise 1-D sums round differently only at the 1e-16 level.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define PYT_PI 3.14159265358979323846
#define PYT_NEAR_MAX_TOL 2.0e-4

static void cexp_neg(double ph, double *re, double *im) {
    /* exp(-i ph) */
    *re = cos(ph);
    *im = -sin(ph);
}

/* local maxima (ends compare with their one neighbour, >=) within tol of the
 * maximum; the argmax when none (ported.py:2433-2448).  Returns the count,
 * indices ascending in out (capacity nt). */
int64_t ora_py_near_max_candidates(const double *coh, int64_t nt, int64_t *out) {
    if (nt <= 1) {
        out[0] = 0;
        return 1;
    }
    double mx = coh[0];
    int64_t amax = 0;
    for (int64_t t = 1; t < nt; ++t)
        if (coh[t] > mx) {
            mx = coh[t];
            amax = t;
        }
    int64_t n = 0;
    for (int64_t t = 0; t < nt; ++t) {
        int lm;
        if (t == 0)
            lm = coh[0] >= coh[1];
        else if (t == nt - 1)
            lm = coh[nt - 1] >= coh[nt - 2];
        else
            lm = coh[t] >= coh[t - 1] && coh[t] >= coh[t + 1];
        if (lm && coh[t] >= mx - PYT_NEAR_MAX_TOL) out[n++] = t;
    }
    if (n == 0) {
        out[0] = amax;
        n = 1;
    }
    return n;
}

/* ported.py:2451-2477: one candidate or the two endpoints -> coarse argmax;
 * otherwise the refined-coherence argmax (first on ties). */
int64_t ora_py_select_candidate(const int64_t *cand, const double *coarse, const double *refined,
                                int64_t n, int64_t trial_count) {
    if (n <= 0) return 0;
    int64_t cb = 0;
    for (int64_t q = 1; q < n; ++q)
        if (coarse[q] > coarse[cb]) cb = q;
    if (n == 1) return cand[cb];
    if (n == 2 && cand[0] == 0 && cand[n - 1] == trial_count - 1) return cand[cb];
    int64_t rb = 0;
    for (int64_t q = 1; q < n; ++q)
        if (refined[q] > refined[rb]) rb = q;
    return cand[rb];
}

/* _ps_topofit_refine_candidate (ported.py:2480-2513) over the m valid
 * elements; res receives the residual phasors (interleaved). */
static void refine_candidate(const double *cr, const double *ci, const double *bp, const double *w,
                             const double *wb, int64_t m, double den_lin, double k0, double *K,
                             double *C, double *coh, double *res) {
    double ore = 0.0, oim = 0.0;
    for (int64_t j = 0; j < m; ++j) {
        double er, ei;
        cexp_neg(k0 * bp[j], &er, &ei);
        ore += cr[j] * er - ci[j] * ei;
        oim += cr[j] * ei + ci[j] * er;
    }
    double num = 0.0;
    for (int64_t j = 0; j < m; ++j) {
        double er, ei;
        cexp_neg(k0 * bp[j], &er, &ei);
        const double rr = cr[j] * er - ci[j] * ei, ri = cr[j] * ei + ci[j] * er;
        /* angle(res * conj(offset)) */
        const double pr = rr * ore + ri * oim, pi = ri * ore - rr * oim;
        num += wb[j] * (w[j] * atan2(pi, pr));
    }
    const double kk = k0 + num / den_lin;
    double mr = 0.0, mi = 0.0, d2 = 0.0;
    for (int64_t j = 0; j < m; ++j) {
        double er, ei;
        cexp_neg(kk * bp[j], &er, &ei);
        const double rr = cr[j] * er - ci[j] * ei, ri = cr[j] * ei + ci[j] * er;
        res[2 * j] = rr;
        res[2 * j + 1] = ri;
        mr += rr;
        mi += ri;
        d2 += hypot(rr, ri);
    }
    if (d2 == 0.0) d2 = 1.0;
    *K = kk;
    *C = atan2(mi, mr);
    *coh = hypot(mr, mi) / d2;
}

/* _ps_topofit_single (ported.py:2355-2430) for one row. */
static void topofit_single(const double *cpx, const double *bperp, int64_t n, const double *tm,
                           int64_t nt, double *K, double *C, double *coh, double *res_out,
                           double *scratch) {
    double *cr = scratch, *ci = cr + n, *bp = ci + n, *w = bp + n, *wb = w + n;
    double *res = wb + n;          /* 2n */
    double *best = res + 2 * n;    /* 2n */
    double *coh_t = best + 2 * n;  /* nt */
    double *ref_coh = coh_t + nt, *ref_k = ref_coh + nt, *ref_c = ref_k + nt; /* nt each */
    int64_t *cand = (int64_t *)(ref_c + nt);
    int64_t *pos = cand + nt;      /* n */
    int64_t m = 0;
    for (int64_t j = 0; j < 2 * n; ++j) res_out[j] = 0.0;
    for (int64_t j = 0; j < n; ++j)
        if (cpx[2 * j] != 0.0 || cpx[2 * j + 1] != 0.0) {
            cr[m] = cpx[2 * j];
            ci[m] = cpx[2 * j + 1];
            bp[m] = bperp[j];
            pos[m] = j;
            ++m;
        }
    if (m == 0) {
        *K = *C = *coh = NAN;
        return;
    }
    double mn = bp[0], mx = bp[0];
    for (int64_t j = 1; j < m; ++j) {
        mn = bp[j] < mn ? bp[j] : mn;
        mx = bp[j] > mx ? bp[j] : mx;
    }
    double range = mx - mn;
    if (range == 0.0) range = 1.0;
    double denom = 0.0, den_lin = 0.0;
    for (int64_t j = 0; j < m; ++j) {
        w[j] = hypot(cr[j], ci[j]);
        denom += w[j];
        wb[j] = w[j] * bp[j];
        den_lin += wb[j] * wb[j];
    }
    if (denom == 0.0) denom = 1.0;
    if (den_lin == 0.0) den_lin = 1.0;
    for (int64_t t = 0; t < nt; ++t) {
        double sr = 0.0, si = 0.0;
        for (int64_t j = 0; j < m; ++j) {
            double er, ei;
            cexp_neg((bp[j] / range * (PYT_PI / 4.0)) * tm[t], &er, &ei);
            sr += er * cr[j] - ei * ci[j];
            si += er * ci[j] + ei * cr[j];
        }
        coh_t[t] = hypot(sr, si) / denom;
    }
    const int64_t nc = ora_py_near_max_candidates(coh_t, nt, cand);
    int64_t sel = 0;
    for (int64_t q = 0; q < nc; ++q) {
        const double k0 = (PYT_PI / 4.0) / range * tm[cand[q]];
        refine_candidate(cr, ci, bp, w, wb, m, den_lin, k0, &ref_k[q], &ref_c[q], &ref_coh[q], res);
    }
    if (nc == 1) {
        sel = 0;
    } else {
        double *co = (double *)malloc(sizeof(double) * (size_t)nc);
        for (int64_t q = 0; q < nc; ++q) co[q] = coh_t[cand[q]];
        const int64_t st = ora_py_select_candidate(cand, co, ref_coh, nc, nt);
        free(co);
        for (int64_t q = 0; q < nc; ++q)
            if (cand[q] == st) {
                sel = q;
                break;
            }
    }
    /* recompute the selected candidate's residual */
    const double k0 = (PYT_PI / 4.0) / range * tm[cand[sel]];
    refine_candidate(cr, ci, bp, w, wb, m, den_lin, k0, K, C, coh, best);
    for (int64_t j = 0; j < m; ++j) {
        res_out[2 * pos[j]] = best[2 * j];
        res_out[2 * pos[j] + 1] = best[2 * j + 1];
    }
}

/* Rows of cpx [P][I] complex128 against bperp [P][I] (or [I] when
 * row_invariant); res [P][I] complex128 (0 where cpx is 0). */
void ora_py_topofit_rows(const double *cpx, const double *bperp, int row_invariant, int64_t P,
                         int64_t I, double n_trial_wraps, double *K, double *C, double *coh,
                         double *res) {
    const int64_t tn = (int64_t)ceil(8.0 * n_trial_wraps);
    const int64_t nt = 2 * tn + 1;
    double *tm = (double *)malloc(sizeof(double) * (size_t)nt);
    for (int64_t t = 0; t < nt; ++t) tm[t] = (double)(t - tn);
    const size_t scratch_n = (size_t)(11 * I + 4 * nt + 16);
    double *scratch = (double *)malloc(sizeof(double) * scratch_n + sizeof(int64_t) * (size_t)(nt + I + 8));
    for (int64_t p = 0; p < P; ++p)
        topofit_single(cpx + 2 * p * I, row_invariant ? bperp : bperp + p * I, I, tm, nt, &K[p],
                       &C[p], &coh[p], res + 2 * p * I, scratch);
    free(scratch);
    free(tm);
}

/* _ps_topofit_batch_row_invariant_coh (ported.py:2555-2616): coherence only,
 * one bperp vector for every row.  The trial grid spans the range of the
 * whole bperp vector (zero elements included, :2561-2564); a row with exactly
 * one near-maximum local peak on that grid (:2580-2589) is refined once from
 * it (:2591-2608, zero elements add nothing to any sum); every other row takes
 * _ps_topofit_single's coherence (:2611-2614), whose grid spans the non-zero
 * elements only.  single_out (may be NULL) receives 1 for rows of the first
 * kind. */
void ora_py_topofit_coh_row_invariant(const double *cpx, const double *bperp, int64_t P,
                                      int64_t I, double n_trial_wraps, double *coh,
                                      int8_t *single_out) {
    const int64_t tn = (int64_t)ceil(8.0 * n_trial_wraps);
    const int64_t nt = 2 * tn + 1;
    double *tm = (double *)malloc(sizeof(double) * (size_t)nt);
    for (int64_t t = 0; t < nt; ++t) tm[t] = (double)(t - tn);
    double mn = bperp[0], mx = bperp[0];
    for (int64_t i = 1; i < I; ++i) {
        mn = bperp[i] < mn ? bperp[i] : mn;
        mx = bperp[i] > mx ? bperp[i] : mx;
    }
    double range = mx - mn;
    if (range == 0.0) range = 1.0;
    const size_t scratch_n = (size_t)(11 * I + 4 * nt + 16);
    double *scratch = (double *)malloc(sizeof(double) * scratch_n + sizeof(int64_t) * (size_t)(nt + I + 8));
    double *coh_t = (double *)malloc(sizeof(double) * (size_t)nt);
    double *w = (double *)malloc(sizeof(double) * (size_t)(2 * I + 2));
    double *wb = w + I + 1, *res = (double *)malloc(sizeof(double) * (size_t)(2 * I + 2));
    double *cr = (double *)malloc(sizeof(double) * (size_t)(2 * I + 2)), *ci = cr + I + 1;
    double *tmp = (double *)malloc(sizeof(double) * 3 * (size_t)I + 8);
    for (int64_t p = 0; p < P; ++p) {
        const double *row = cpx + 2 * p * I;
        double denom = 0.0, den_lin = 0.0;
        for (int64_t i = 0; i < I; ++i) {
            cr[i] = row[2 * i];
            ci[i] = row[2 * i + 1];
            w[i] = hypot(cr[i], ci[i]);
            wb[i] = w[i] * bperp[i];
            denom += w[i];
            den_lin += wb[i] * wb[i];
        }
        if (denom == 0.0) denom = 1.0;
        if (den_lin == 0.0) den_lin = 1.0;
        double cmax = -INFINITY;
        for (int64_t t = 0; t < nt; ++t) {
            double sr = 0.0, si = 0.0;
            for (int64_t i = 0; i < I; ++i) {
                double er, ei;
                cexp_neg((bperp[i] / range * (PYT_PI / 4.0)) * tm[t], &er, &ei);
                sr += er * cr[i] - ei * ci[i];
                si += er * ci[i] + ei * cr[i];
            }
            coh_t[t] = hypot(sr, si) / denom;
            cmax = coh_t[t] > cmax ? coh_t[t] : cmax;
        }
        int64_t count = 0, only = 0;
        for (int64_t t = 0; t < nt; ++t) {
            int lm;
            if (nt == 1)
                lm = 1;
            else if (t == 0)
                lm = coh_t[0] >= coh_t[1];
            else if (t == nt - 1)
                lm = coh_t[nt - 1] >= coh_t[nt - 2];
            else
                lm = coh_t[t] >= coh_t[t - 1] && coh_t[t] >= coh_t[t + 1];
            if (lm && coh_t[t] >= cmax - PYT_NEAR_MAX_TOL) {
                if (count == 0) only = t;
                ++count;
            }
        }
        if (single_out) single_out[p] = (int8_t)(count == 1);
        if (count == 1) {
            double K, Cc;
            refine_candidate(cr, ci, bperp, w, wb, I, den_lin, (PYT_PI / 4.0) / range * tm[only],
                             &K, &Cc, &coh[p], res);
        } else {
            double K, Cc;
            topofit_single(row, bperp, I, tm, nt, &K, &Cc, &coh[p], tmp, scratch);
        }
    }
    free(tmp);
    free(cr);
    free(res);
    free(w);
    free(coh_t);
    free(scratch);
    free(tm);
}

/* _stage2_grid_accumulate_cpu: np.add.at per ifg, rows in index order,
 * complex64 sums; out [n_i*n_j][I] complex64.  Returns -1 on an index out of
 * range (numpy raises IndexError). */
int ora_py_grid_accumulate(const float *ph, const int64_t *grid_lin, int64_t P, int64_t I,
                           int64_t n_cells, float *out) {
    memset(out, 0, sizeof(float) * 2 * (size_t)(n_cells * I));
    for (int64_t p = 0; p < P; ++p)
        if (grid_lin[p] < 0 || grid_lin[p] >= n_cells) return -1;
    for (int64_t i = 0; i < I; ++i)
        for (int64_t p = 0; p < P; ++p) {
            float *o = out + 2 * (grid_lin[p] * I + i);
            o[0] += ph[2 * (p * I + i)];
            o[1] += ph[2 * (p * I + i) + 1];
        }
    return 0;
}
