/*
 * Stage-2 oracle: a CPU restatement of pySTAMPS native stage 2
 * (crates/pystamps-core/src/native_stage2.rs) in plain C.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline.  The product path (pystamps_amd/, the HIP
 * library) never calls it.
 *
 * Two modes of the same algorithm:
 *   ORA_MODE_REF32 — the reference's precision map: f32 fill / grid / gather /
 *                    topofit / refine (refine's mopt accumulators f64), f64 CLAP
 *                    with f32 in/out, f64 RMS and weights; glibc sinf/cosf/
 *                    hypotf/atan2f (the libm Rust std links on linux-gnu);
 *                    sequential sums in the reference order.
 *   ORA_MODE_REF64 — the same algorithm with every quantity in f64 (the "FP64
 *                    oracle"); first-argmax trial selection as the reference.
 * Both consume the same prepared inputs (the Stage2Prepared mirror below).
 */
#ifndef PYSTAMPS_STAGE2_ORACLE_H
#define PYSTAMPS_STAGE2_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORA_MODE_REF32 = 32, ORA_MODE_REF64 = 64 };
enum { ORA_WEIGHT_PSQUARE = 0, ORA_WEIGHT_SNR = 1 };
enum { ORA_COH_BINS = 100 };

/* Mirror of Stage2Prepared + Stage2Options + Stage2Parms
 * (native_stage2.rs:115-182). */
typedef struct {
    int64_t n_ps, n_ifg, n_i, n_j;
    const float *ph_nm;        /* [n_ps][n_ifg] complex64 interleaved (re, im) */
    const float *amp;          /* [n_ps][n_ifg] */
    int32_t row_invariant;     /* row_invariant_bperp */
    const double *bperp_row;   /* [n_ifg] row_bperp_nm */
    const double *bperp_mat;   /* [n_ps][n_ifg], may be NULL when row_invariant */
    const int64_t *grid_lin;   /* [n_ps] */
    const double *d_a;         /* [n_ps] */
    const double *low_pass;    /* [n_win_ex][n_win_ex] */
    int64_t n_win_ex;          /* low_pass dimension (clap_win) */
    int64_t clap_window;       /* round(0.75 * clap_win) */
    double clap_alpha, clap_beta;
    const double *trial_values;
    int64_t n_trial;
    const double *nr_base;     /* [100] */
    double nr_max_nz_ix;
    const double *coh_bins;    /* [100] */
    int64_t low_coh_thresh;
    int32_t weighting;         /* ORA_WEIGHT_* */
    double gamma_change_convergence;
    int64_t gamma_max_iterations;
} ora_stage2_in;

/* Mirror of the write_pm1 argument list (native_stage2.rs:1899-1916).
 * The real-typed arrays are float for REF32 and double for REF64. */
typedef struct {
    double *k_ps, *c_ps, *coh_ps, *n_opt, *coh_ps_save;  /* [n_ps] */
    void *ph_res;      /* [n_ps][n_ifg] real */
    void *ph_patch;    /* [n_ps][n_ifg] complex real */
    void *ph_grid;     /* [n_i*n_j][n_ifg] complex real */
    void *ph_weight;   /* [n_ps][n_ifg] complex real */
    double *nr;        /* [100] last scaled Nr */
    double gamma_change_save;
    int64_t i_loop;
    /* per-iteration diagnostics, optional (NULL = skip) */
    int32_t *trial_ix;        /* [n_ps] selected trial index of the last iteration (-1 invalid) */
    float *trial_margin;      /* [n_ps] (best - second best) coherence of the last scan */
    double *weighting_last;   /* [n_ps] weighting used by the last fill */
    double *k_prev;           /* [n_ps] K used by the last fill */
    float *trial_margin_min;  /* [n_ps] smallest trial margin over the loop's iterations */
    double *decision_margin_min; /* [n_ps] smallest distance of coh to a P-square lookup
                                    (n + 1/2)/1000 or histogram bin boundary over the
                                    iterations whose coherence sets weights (not the last) */
    double *refine_margin_min; /* [n_ps] smallest distance of a residual phase of the
                                  least-squares step to the +-pi cut of arg() */
} ora_stage2_out;

/* Whole gamma loop (native_stage2.rs:242-406). Returns 0 or a negative code. */
int ora_stage2_run(const ora_stage2_in *in, ora_stage2_out *out, int mode, int threads);

/* Single steps, for state injection (same semantics as the loop). */
int ora_fill_phase_weight(const ora_stage2_in *in, const double *k_ps, const double *weighting,
                          void *ph_weight, int mode, int threads);
void ora_accumulate_grid(const ora_stage2_in *in, const void *ph_weight, void *ph_grid, int mode);
void ora_clap_filter_grid_stack(const ora_stage2_in *in, const void *ph_grid, void *ph_filt,
                                int mode, int threads);
void ora_extract_patch_phase(const ora_stage2_in *in, const void *ph_filt, void *ph_patch,
                             int mode, int threads);
int ora_fit_rows(const ora_stage2_in *in, const void *ph_patch, double *k_ps, double *c_ps,
                 double *coh_ps, double *n_opt, void *ph_res, int32_t *trial_ix,
                 float *trial_margin, int mode, int threads);

/* Host-sized pieces of the loop and of the prep. */
double ora_rms_difference(const double *a, const double *b, int64_t n);
void ora_hist_with_centers(const double *values, int64_t n, const double *centers, int64_t nc,
                           double *out);
void ora_psquare_weighting(const double *nr, const double *na, int64_t nbins,
                           int64_t low_coh_thresh, double nr_max_nz_ix, const double *coh_ps,
                           int64_t n, double *out_weight, double *out_prand_hi /* 1001 or NULL */);
void ora_snr_weighting(const float *amp, const void *ph_res, int64_t n_ps, int64_t n_ifg,
                       double *out, int mode);
void ora_rng_uniform(uint32_t seed, int64_t n, double *out);
int64_t ora_trial_values(double n_trial_wraps, double *out, int64_t cap);
void ora_low_pass(double clap_win, double low_pass_wavelength, double grid_size, double *out,
                  int64_t *out_dim);
void ora_grid_indices(const float *xy, int64_t n_ps, double grid_size, float *out_ij,
                      int64_t *out_n_i, int64_t *out_n_j);
void ora_clap_window_weight(int64_t n_win, double *out);
/* Random-coherence histogram (native_stage2.rs:1310-1356, single-master branch). */
void ora_random_coherence_histogram(const double *bperp_nm, int64_t n_ifg,
                                    const double *trial_values, int64_t n_trial,
                                    const double *coh_bins, int64_t n_rows, uint32_t seed,
                                    double *out_nr, double *out_nr_max_nz_ix, int threads);
/* small-baseline branch: interferogram phases are image-phase differences
 * (image_a / image_b 0-based, native_stage2.rs:1321-1334) */
void ora_random_coherence_histogram_sb(const double *bperp_nm, int64_t n_ifg,
                                       const int32_t *image_a, const int32_t *image_b,
                                       int64_t n_image, const double *trial_values,
                                       int64_t n_trial, const double *coh_bins, int64_t n_rows,
                                       uint32_t seed, double *out_nr, double *out_nr_max_nz_ix,
                                       int threads);
/* FP64 coherence of one phase row (native_stage2.rs:1438-1512). */
double ora_topofit_phase_row_coh(const double *phase, const double *bperp, int64_t n_col,
                                 const double *trial_values, int64_t n_trial);
/* Radix-2 FFT used by the CLAP restatement, exposed for the DFT brute-force check. */
void ora_fft32_2d(double *re_im /* [32][32][2] */, int inverse);

#ifdef __cplusplus
}
#endif
#endif
