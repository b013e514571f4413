/*
 * pystamps_hip — MI355X (gfx950) implementation of pySTAMPS stage 2, PS gamma
 * estimation (the StaMPS ps_est_gamma_quick loop), behind a plain C ABI.
 *
 * The loop this library runs is run_stage2_native_inner's loop body
 * (crates/pystamps-core/src/native_stage2.rs:242-406): fill_phase_weight,
 * accumulate_grid, clap_filter_grid_stack, extract_patch_phase,
 * fit_stage2_rows, the gamma-change RMS stop test and the P-square / SNR
 * weighting update.  The caller keeps the reference's prep
 * (prepare_stage2_inputs, native_stage2.rs:415-561) and its pm1.mat writer
 * (write_pm1, native_stage2.rs:1899-1965); this library takes the prepared
 * inputs and fills the write_pm1 arguments.
 *
 * Conventions
 *  - All pointers passed in or out are HOST memory owned by the caller.  The
 *    library owns its device buffers and stream.  Nothing allocated by the
 *    library is returned to the caller except opaque session handles.
 *  - Complex arrays are interleaved (re, im) pairs, row-major [rows][cols],
 *    exactly the reference's Vec<Complex32> layouts.
 *  - Every entry point returns 0 on success or a negative PS2_E* code; the
 *    message of the last failure on the calling thread is in
 *    ps2_hip_last_error().  The reference's error convention is
 *    CoreError::NativeStage{stage: 2, message} (native_stage2.rs:2447-2453);
 *    a Rust caller maps a negative return to that variant with this message.
 *  - A session serialises its own calls; distinct sessions may run on
 *    distinct host threads.
 */
#ifndef PYSTAMPS_HIP_H
#define PYSTAMPS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PS2_HIP_ABI_VERSION 2

/* return codes */
#define PS2_OK 0
#define PS2_E_ARG (-1)      /* invalid argument / shape (message says which) */
#define PS2_E_DEVICE (-2)   /* HIP runtime failure */
#define PS2_E_NOMEM (-3)    /* device allocation failed */
#define PS2_E_STATE (-4)    /* session used out of order */

/* computation forms */
#define PS2_FORM_FP64 0     /* plain FP64 form: every step in f64 */
#define PS2_FORM_FP32 1     /* FP32 bulk + FP64 near-tie trial re-score and P-square re-decision */

/* weighting (Stage2Parms::filter_weighting, native_stage2.rs:364-384) */
#define PS2_WEIGHT_PSQUARE 0
#define PS2_WEIGHT_SNR 1

#define PS2_COH_BINS 100

/* Prepared per-patch inputs: mirror of Stage2Prepared + Stage2Options +
 * Stage2Parms (native_stage2.rs:115-182). */
typedef struct {
    int64_t n_ps;             /* Stage2Prepared::n_ps */
    int64_t n_ifg;            /* Stage2Prepared::n_ifg (master column dropped) */
    int64_t n_i, n_j;         /* grid extent */
    const float *ph_nm;       /* [n_ps][n_ifg] complex64: ph / |ph| */
    const float *amp;         /* [n_ps][n_ifg]: |ph|, 1 where |ph| == 0 */
    int32_t row_invariant;    /* row_invariant_bperp */
    const double *bperp_row;  /* [n_ifg] row_bperp_nm */
    const double *bperp_mat;  /* [n_ps][n_ifg]; may be NULL when row_invariant */
    const int64_t *grid_lin;  /* [n_ps] cell = (i-1)*n_j + (j-1), each < n_i*n_j */
    const double *d_a;        /* [n_ps] amplitude dispersion */
    const double *low_pass;   /* [n_win_ex][n_win_ex] */
    int64_t n_win_ex;         /* low_pass dimension = FFT size (clap_win), at most 64 */
    int64_t clap_window;      /* round(0.75 * clap_win), 1 .. n_win_ex */
    double clap_alpha, clap_beta;
    const double *trial_values; /* [n_trial] */
    int64_t n_trial;
    const double *nr_base;    /* [100] random-coherence histogram */
    double nr_max_nz_ix;
    const double *coh_bins;   /* [100] */
    int64_t low_coh_thresh;   /* 31 (15 small-baseline) */
    int32_t weighting;        /* PS2_WEIGHT_* */
    double gamma_change_convergence;
    int64_t gamma_max_iterations;
} ps2_patch_in;

/* Outputs: mirror of the write_pm1 argument list (native_stage2.rs:1899-1916).
 * Any pointer may be NULL to skip that output. */
typedef struct {
    double *k_ps, *c_ps, *coh_ps, *n_opt, *coh_ps_save; /* [n_ps] */
    float *ph_res;            /* [n_ps][n_ifg] */
    float *ph_patch;          /* [n_ps][n_ifg] complex64 */
    float *ph_grid;           /* [n_i*n_j][n_ifg] complex64 */
    float *ph_weight;         /* [n_ps][n_ifg] complex64 */
    double *nr;               /* [100] last scaled Nr */
    double gamma_change_save;
    int64_t i_loop;
    /* diagnostics */
    int32_t *trial_ix;        /* [n_ps] selected trial index, -1 for invalid rows */
    int64_t n_rescored;       /* FP32 form: rows re-scored in FP64, summed over iterations */
    int64_t n_rechecked;      /* FP32 form, P-square: rows refined again in FP64 because their
                                 coherence lay near a lookup / histogram boundary, summed */
    double loop_seconds;      /* device time of the gamma loop (first fill .. stop decision) */
} ps2_patch_out;

typedef struct {
    int32_t form;             /* PS2_FORM_* */
    int32_t device;           /* HIP device ordinal */
    double rescore_bound;     /* FP32 form: coherence margin below which the top trials
                                 are re-scored in FP64; <= 0 selects the derived bound */
    int32_t verbose;          /* 1: per-iteration phase timers on stderr */
    int32_t reserved;
    int32_t autotune;         /* FP32 form, 1: tiles for the patch's shape (n_ifg, n_trial, patch
                                 size, window grid) from autotune_cache, measured on the session's
                                 data and appended on a miss; 0: the built-in tile model */
    const char *autotune_cache; /* path of that text cache (created if absent; lines are appended
                                   under an advisory lock); NULL: measure, keep nothing */
} ps2_run_opts;

/* tile sources (ps2_tiles.source) */
#define PS2_TILES_MODEL 0
#define PS2_TILES_CACHE 1
#define PS2_TILES_MEASURED 2

/* Tiles a session runs with. */
typedef struct {
    int32_t fill_cells;              /* cells per fill workgroup (FP32 form: 64 or 128) */
    int32_t scan_step_table;         /* FP32 scan: 1 unit steps from the table, 0 computed */
    int32_t clap_segments;           /* CLAP strip segments per window row (0: no windows) */
    int32_t clap_windows_per_segment;
    int32_t source;                  /* PS2_TILES_* */
} ps2_tiles;

/* ABI version (PS2_HIP_ABI_VERSION). */
int ps2_hip_abi_version(void);

/* Message of the last failure on this thread ("" if none). */
const char *ps2_hip_last_error(void);

/* Number of visible HIP devices (0 when none), or a negative code. */
int ps2_hip_device_count(void);

/* Whole stage-2 gamma loop for one patch: replaces the loop of
 * run_stage2_native_inner (native_stage2.rs:242-406) and fills the write_pm1
 * arguments (native_stage2.rs:319-336).  Inputs are checked on the host
 * before any device call; a patch whose n_ps*n_ifg or n_i*n_j*n_ifg reaches
 * 2^31 is rejected (split it into patches, as the reference's patching does). */
int ps2_hip_run(const ps2_patch_in *in, ps2_patch_out *out, const ps2_run_opts *opts);

/* Several patches, each with its own convergence; the reference runs them one
 * after another (pystamps-native.rs:158-176).  Patches go in waves of up to
 * 16: the wave's patches are uploaded, their loops run concurrently (one
 * stream and host worker each), their outputs are downloaded and their device
 * memory is freed before the next wave, so device memory holds one wave.
 * After a failure no new patch starts; the error names the lowest failing
 * patch ("patch k: ..."); outputs of earlier waves are written. */
int ps2_hip_run_batch(size_t n, const ps2_patch_in *in, ps2_patch_out *out,
                      const ps2_run_opts *opts);

/* Random-coherence histogram Nr of the stage-2 prep
 * (random_coherence_histogram, native_stage2.rs:1310-1356, single-master
 * branch): n_rows MATLAB-v5 uniform phase rows (seed 2005 in the reference,
 * 300000 rows) fitted in FP64 on the device.  out_nr[100] receives the counts,
 * *out_nr_max_nz_ix the last non-zero bin + 1 (1 if none). */
int ps2_hip_random_coherence_histogram(const double *bperp_nm, int64_t n_ifg,
                                       const double *trial_values, int64_t n_trial,
                                       const double *coh_bins, int64_t n_rows, uint32_t seed,
                                       int32_t device, double *out_nr,
                                       double *out_nr_max_nz_ix);

/* Small-baseline branch of the same histogram (random_coherence_histogram,
 * native_stage2.rs:1321-1334 with small_baseline_random_indices :1358-1403):
 * n_rows x n_image uniform image phases, interferogram i's phase is
 * image_b[i] - image_a[i] (0-based image indices, from ps1.ifgday_ix - 1). */
int ps2_hip_random_coherence_histogram_sb(const double *bperp_nm, int64_t n_ifg,
                                          const int32_t *image_a, const int32_t *image_b,
                                          int64_t n_image, const double *trial_values,
                                          int64_t n_trial, const double *coh_bins, int64_t n_rows,
                                          uint32_t seed, int32_t device, double *out_nr,
                                          double *out_nr_max_nz_ix);

/* ---- Kernels of the pySTAMPS kernel registry (a `hip` provider for
 * pystamps/kernels/accelerated.py:392-426) ----
 *
 * Python-pipeline topofit of n_rows rows (stage2_topofit,
 * stage2_topofit_row_invariant, stage2_topofit_coh_row_invariant; the fit is
 * _ps_topofit_single, pystamps/pipeline/ported.py:2355-2430), in FP64: zero
 * elements are left out of the fit, the local maxima of the trial coherence
 * within 2e-4 of the maximum are refined, and the refined coherence selects
 * among several (ported.py:2433-2477).  cpx [n_rows][n_ifg] complex128;
 * bperp [n_rows][n_ifg], or [n_ifg] when row_invariant; k/c/coh [n_rows]
 * (NaN for an all-zero row); res_out [n_rows][n_ifg] complex64 residual
 * phasors (0 at zero elements), may be NULL. */
int ps2_hip_topofit_rows(const double *cpx, const double *bperp, int32_t row_invariant,
                         int64_t n_rows, int64_t n_ifg, double n_trial_wraps, int32_t device,
                         double *k_out, double *c_out, double *coh_out, float *res_out);

/* Coherence only of n_rows rows against one bperp vector: the registry
 * kernel stage2_topofit_coh_row_invariant, whose python and native providers
 * both run _ps_topofit_batch_row_invariant_coh
 * (pystamps/kernels/accelerated.py:368-389, pystamps/pipeline/ported.py:
 * 2555-2616); it takes the place of the native module's
 * ps_topofit_coh_row_invariant (src/lib.rs:1326-1378).  The trial grid spans the range of the whole
 * bperp vector, zero elements included; a row with exactly one near-maximum
 * local peak on that grid is refined once from it, every other row takes
 * ps2_hip_topofit_rows' coherence.  FP64.  cpx [n_rows][n_ifg] complex128,
 * bperp_vec [n_ifg], coh_out [n_rows]. */
int ps2_hip_topofit_coh_row_invariant(const double *cpx, const double *bperp_vec, int64_t n_rows,
                                      int64_t n_ifg, double n_trial_wraps, int32_t device,
                                      double *coh_out);

/* Allocation-free variant of ps2_hip_topofit_rows.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Inputs and outputs are DEVICE pointers in the same row-major layouts as
 * ps2_hip_topofit_rows' host arrays (cpx [n_rows][n_ifg] complex128; bperp
 * [n_rows][n_ifg], or [n_ifg] when row_invariant; k/c/coh [n_rows]; res
 * [n_rows][n_ifg] complex64 or NULL); results equal ps2_hip_topofit_rows'.
 * The per-row trial coherences live in a fixed 32 MiB device array of the
 * library (one per GPU, held until the process ends), so the call makes no
 * hipMalloc, hipFree or other allocation; it rejects trial counts above
 * 16384 (2*ceil(8*n_trial_wraps)+1) with PS2_E_ARG.  Work is queued on
 * `stream` (a hipStream_t; NULL = the null stream) and the call returns after
 * that stream has completed.  Calls that overlap on one GPU (other host
 * threads, other streams) share the array and race. */
int ps2_hip_topofit_rows_noalloc(const double *d_cpx, const double *d_bperp, int32_t row_invariant,
                                 int64_t n_rows, int64_t n_ifg, double n_trial_wraps,
                                 int32_t device, void *stream, double *d_k_out, double *d_c_out,
                                 double *d_coh_out, float *d_res_out);

/* stage2_grid_accumulate (_stage2_grid_accumulate_cpu,
 * pystamps/kernels/accelerated.py:183-200: np.add.at per ifg, rows in index
 * order, complex64): ph_weight [n_rows][n_ifg] complex64, grid_lin [n_rows]
 * in [0, n_i*n_j); out [n_i*n_j][n_ifg] complex64.  An index out of range is
 * PS2_E_ARG (numpy raises IndexError; negative indices are not wrapped). */
int ps2_hip_grid_accumulate(const float *ph_weight, const int64_t *grid_lin, int64_t n_rows,
                            int64_t n_ifg, int64_t n_i, int64_t n_j, int32_t device, float *out);

/* stage2_histogram (_stage2_histogram_with_centers_cpu,
 * pystamps/kernels/accelerated.py:240-264; the rule of hist_with_centers,
 * native_stage2.rs:1538-1622): counts of the finite values per center;
 * host code, no device use. */
int ps2_hip_histogram_with_centers(const double *values, int64_t n, const double *centers,
                                   int64_t n_centers, double *out);

/* ---- Prep helpers (host code, host threads; no device use) ----
 * prepare_stage2_inputs' phase normalisation (native_stage2.rs:449-467): for
 * every row of ph [n_rows][n_cols] complex64 and each kept column cols[i]
 * (the master column dropped), amp = |v| (glibc hypotf), 1 where it is 0, and
 * ph_nm = v / amp; outputs [n_rows][n_sel].  n_threads <= 0: all cores. */
int ps2_hip_normalize_phase(const float *ph, int64_t n_rows, int64_t n_cols, const int64_t *cols,
                            int64_t n_sel, int32_t n_threads, float *ph_nm, float *amp);

/* Column selection of a row-major f64 matrix (bp1.bperp_mat without the
 * master column, load_bperp_mat native_stage2.rs:594-606). */
int ps2_hip_select_columns_f64(const double *src, int64_t n_rows, int64_t n_cols,
                               const int64_t *cols, int64_t n_sel, int32_t n_threads, double *dst);

/* ---- Sessions: device-resident inputs, for repeated runs and per-step use ----
 * A session uploads a patch's constant inputs once.  ps2_hip_session_run then
 * runs the whole loop from the initial state (K = 0, weighting = 1/D_A) with
 * every input already in HBM, and returns after the stop decision; outputs
 * stay on the device until ps2_hip_session_download.  The loop stays on the
 * device: a control kernel takes the stop decision and builds the P-square
 * table, gamma_max_iterations iterations are captured as one HIP graph on the
 * session's first run (one iteration, launched that many times, beyond 64
 * iterations) and launched (kernels after the stop return at once); the run
 * ends with the write_pm1 outputs of the last iteration arranged in the
 * caller's layouts on the device, and the host waits once per loop.  With timing or verbose output the
 * same kernels are launched iteration by iteration instead.
 * Threads: every entry point may be called from several host threads at once.
 * A session is used by one call at a time (calls on the same session
 * serialise).  Sessions, Nr and registry calls each work on streams of their
 * own and never on the legacy default stream; the library's device and pinned
 * allocations wait while any of its sessions captures its graph.  If a
 * capture is still disturbed by other code in the process, that session
 * launches its iterations one by one from then on (same results). */
typedef struct ps2_session ps2_session;

int ps2_hip_session_create(const ps2_patch_in *in, const ps2_run_opts *opts, ps2_session **out);
int ps2_hip_session_run(ps2_session *s, int64_t *out_i_loop, double *out_loop_seconds);
int ps2_hip_session_download(ps2_session *s, ps2_patch_out *out);
void ps2_hip_session_destroy(ps2_session *s);

/* The loop with the state after every gamma iteration: the iterations are
 * launched one at a time; after each, the outputs of that state are written
 * into `state` (the caller's buffers, as ps2_hip_session_download fills them:
 * K_ps .. ph_weight, Nr as the iteration's control step left it,
 * gamma_change_save, i_loop = the iteration) and fn(user, iteration, state)
 * is called on the calling thread; a non-zero return ends the loop with that
 * code.  The last call holds the loop's final state.  Replaces the reference's
 * PYSTAMPS_STAGE2_NATIVE_DEBUG_PM snapshots, pm1_iter_NN.mat written after
 * every iteration (native_stage2.rs:259, :337-356, :385-404). */
typedef int (*ps2_iteration_fn)(void *user, int64_t iteration, ps2_patch_out *state);
int ps2_hip_session_run_iterations(ps2_session *s, ps2_patch_out *state, ps2_iteration_fn fn, void *user,
                                   int64_t *out_i_loop);

/* Runs the loops of n sessions concurrently (each on its own stream; up to 16
 * host workers, so at most 16 loops are in flight; sessions may be on the
 * same device) and returns after every loop's stop decision.  i_loops[k] / seconds[k] receive each session's iteration
 * count and device loop time (either may be NULL).  This is the batched form
 * of the reference's sequential per-patch calls (pystamps-native.rs:158-176). */
int ps2_hip_sessions_run(size_t n, ps2_session *const *sessions, int64_t *i_loops,
                         double *seconds);

/* Stage-2 results left on the device for a GPU consumer (the stage-3 inputs
 * native_stage3.rs:40-105 reads from pm1.mat: K_ps, C_ps, coh_ps, N_opt,
 * ph_patch, ph_res, Nr), in the caller's row order.  The device arrays belong
 * to the session: they stay valid until the session's next run, export or
 * destroy.  The call returns after they are complete. */
typedef struct {
    int32_t device;
    int64_t n_ps, n_ifg;
    const double *k_ps, *c_ps, *coh_ps, *n_opt; /* device, [n_ps] */
    const float *ph_patch;                      /* device, [n_ps][n_ifg] complex64 */
    const float *ph_res;                        /* device, [n_ps][n_ifg] */
    double nr[PS2_COH_BINS];                    /* host copy: last scaled Nr */
    double gamma_change_save;
    int64_t i_loop;
} ps2_device_outputs;
int ps2_hip_session_export_device(ps2_session *s, ps2_device_outputs *out);

/* Per-kernel device time of the session's runs and single steps so far (the reference prints
 * the same phase breakdown per iteration, native_stage2.rs:263-305).
 * Kernel slots, in order: fill_grid, clap_strip, clap_combine, scan,
 * rescore, refine, post, weights, outputs.  ms[k] is the summed HIP-event
 * time of slot k over launches[k] launches.  post is the step after the
 * refine: the FP64 re-decision of the rows the refine marks (FP32 form, its
 * own kernel) and the gamma-change RMS partials and coherence histogram,
 * timed together as one launch (the loop control,
 * one workgroup, is not timed); weights is the weighting update (a patch of
 * at most 65,536 rows computes P-square weights in the next iteration's fill,
 * so that slot stays empty for it); outputs is the once-per-loop kernel that
 * writes the write_pm1 arguments in the caller's layouts.  Timing events are recorded only when
 * ps2_hip_session_set_timing(s, 1) was called before the runs; such runs
 * launch the kernels one by one instead of as a graph, with the same results,
 * each dispatch carrying its slot's start / stop events (no marker between
 * kernels): every iteration is queued before the one host wait (the device
 * runs them back to back, as in an untimed loop; at most 4096 iterations per
 * loop), and launches of iterations after the stop decision, which return at
 * once, are not counted. */
#define PS2_NUM_KERNEL_SLOTS 9
int ps2_hip_session_set_timing(ps2_session *s, int32_t enable);

/* The tiles of a session (the model's, or the autotuner's: ps2_run_opts). */
int ps2_hip_session_tiles(ps2_session *s, ps2_tiles *out);
int ps2_hip_session_kernel_times(ps2_session *s, double *ms, int64_t *launches);

/* Single steps on a session, for state injection.  Real-typed arrays are float
 * (complex64) for PS2_FORM_FP32 and double (complex128) for PS2_FORM_FP64.
 *  fill_grid: fill_phase_weight + accumulate_grid (native_stage2.rs:629-657,
 *             722-740) from host K [n_ps] and weighting [n_ps]; ph_grid out.
 *  clap:      clap_filter_grid_stack (:742-922) of host ph_grid; ph_filt out.
 *  fit:       extract_patch_phase + fit_stage2_rows (:1186-1205, :659-720) of
 *             host ph_filt; the out struct's k/c/coh/n_opt/ph_res/ph_patch/
 *             trial_ix are filled (ph_res / ph_patch as float / complex64).
 *             It replaces the session's loop outputs: download and
 *             export_device then need another session_run. */
int ps2_hip_step_fill_grid(ps2_session *s, const double *k_ps, const double *weighting,
                           void *ph_grid_out);
int ps2_hip_step_clap(ps2_session *s, const void *ph_grid, void *ph_filt_out);
int ps2_hip_step_fit(ps2_session *s, const void *ph_filt, ps2_patch_out *out);

#ifdef __cplusplus
}
#endif
#endif /* PYSTAMPS_HIP_H */
