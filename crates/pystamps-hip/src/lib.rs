//! Raw FFI declarations of `include/pystamps_hip.h` (ABI version 2) and thin
//! safe wrappers, for `pystamps-core`'s `hip` stage-2 backend.
//!
//! Every struct is `#[repr(C)]` with the header's fields in the header's
//! order and types (tests/test_rust_binding.py in the pystamps-hip build
//! checks names, order, types and offsets against the header).  Complex
//! arrays are interleaved `(re, im)` pairs, which is the layout of
//! `num_complex::Complex32` / `Complex64` (`#[repr(C)] { re, im }`), so a
//! `Vec<Complex32>` can be passed as `*const f32`.
//!
//! Errors: every entry point returns 0 or a negative code; the message is
//! `last_error()`.  The reference's stage-2 error is
//! `CoreError::NativeStage { stage: 2, message }`
//! (crates/pystamps-core/src/native_stage2.rs:2447-2453); the wrappers below
//! return the message for the caller to wrap in that variant.

use std::ffi::{CStr, CString};
use std::os::raw::{c_char, c_int, c_void};

pub const PS2_HIP_ABI_VERSION: c_int = 2;

pub const PS2_OK: c_int = 0;
pub const PS2_E_ARG: c_int = -1;
pub const PS2_E_DEVICE: c_int = -2;
pub const PS2_E_NOMEM: c_int = -3;
pub const PS2_E_STATE: c_int = -4;

pub const PS2_FORM_FP64: i32 = 0;
pub const PS2_FORM_FP32: i32 = 1;

pub const PS2_WEIGHT_PSQUARE: i32 = 0;
pub const PS2_WEIGHT_SNR: i32 = 1;

pub const PS2_COH_BINS: usize = 100;
pub const PS2_NUM_KERNEL_SLOTS: usize = 9;

pub const PS2_TILES_MODEL: i32 = 0;
pub const PS2_TILES_CACHE: i32 = 1;
pub const PS2_TILES_MEASURED: i32 = 2;

/// `ps2_patch_in`: Stage2Prepared + Stage2Options + Stage2Parms
/// (native_stage2.rs:115-182).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct Ps2PatchIn {
    pub n_ps: i64,
    pub n_ifg: i64,
    pub n_i: i64,
    pub n_j: i64,
    pub ph_nm: *const f32,
    pub amp: *const f32,
    pub row_invariant: i32,
    pub bperp_row: *const f64,
    pub bperp_mat: *const f64,
    pub grid_lin: *const i64,
    pub d_a: *const f64,
    pub low_pass: *const f64,
    pub n_win_ex: i64,
    pub clap_window: i64,
    pub clap_alpha: f64,
    pub clap_beta: f64,
    pub trial_values: *const f64,
    pub n_trial: i64,
    pub nr_base: *const f64,
    pub nr_max_nz_ix: f64,
    pub coh_bins: *const f64,
    pub low_coh_thresh: i64,
    pub weighting: i32,
    pub gamma_change_convergence: f64,
    pub gamma_max_iterations: i64,
}

/// `ps2_patch_out`: the write_pm1 arguments (native_stage2.rs:1899-1916).
/// Any pointer may be null to skip that output.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct Ps2PatchOut {
    pub k_ps: *mut f64,
    pub c_ps: *mut f64,
    pub coh_ps: *mut f64,
    pub n_opt: *mut f64,
    pub coh_ps_save: *mut f64,
    pub ph_res: *mut f32,
    pub ph_patch: *mut f32,
    pub ph_grid: *mut f32,
    pub ph_weight: *mut f32,
    pub nr: *mut f64,
    pub gamma_change_save: f64,
    pub i_loop: i64,
    pub trial_ix: *mut i32,
    pub n_rescored: i64,
    pub n_rechecked: i64,
    pub loop_seconds: f64,
}

/// `ps2_run_opts`.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct Ps2RunOpts {
    pub form: i32,
    pub device: i32,
    pub rescore_bound: f64,
    pub verbose: i32,
    pub reserved: i32,
    pub autotune: i32,
    pub autotune_cache: *const c_char,
}

impl Default for Ps2RunOpts {
    fn default() -> Self {
        Ps2RunOpts {
            form: PS2_FORM_FP32,
            device: 0,
            rescore_bound: 0.0,
            verbose: 0,
            reserved: 0,
            autotune: 0,
            autotune_cache: std::ptr::null(),
        }
    }
}

/// `ps2_tiles`.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct Ps2Tiles {
    pub fill_cells: i32,
    pub scan_step_table: i32,
    pub clap_segments: i32,
    pub clap_windows_per_segment: i32,
    pub source: i32,
}

/// `ps2_device_outputs`: stage-3 inputs left in HBM (native_stage3.rs:40-105).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct Ps2DeviceOutputs {
    pub device: i32,
    pub n_ps: i64,
    pub n_ifg: i64,
    pub k_ps: *const f64,
    pub c_ps: *const f64,
    pub coh_ps: *const f64,
    pub n_opt: *const f64,
    pub ph_patch: *const f32,
    pub ph_res: *const f32,
    pub nr: [f64; PS2_COH_BINS],
    pub gamma_change_save: f64,
    pub i_loop: i64,
}

/// Opaque `ps2_session`.
#[repr(C)]
pub struct Ps2Session {
    _private: [u8; 0],
}

/// `ps2_iteration_fn`: called after every gamma iteration of
/// `ps2_hip_session_run_iterations` with the state in the caller's buffers;
/// a non-zero return ends the loop with that code.
pub type Ps2IterationFn = Option<unsafe extern "C" fn(user: *mut c_void, iteration: i64, state: *mut Ps2PatchOut) -> c_int>;

extern "C" {
    pub fn ps2_hip_abi_version() -> c_int;
    pub fn ps2_hip_last_error() -> *const c_char;
    pub fn ps2_hip_device_count() -> c_int;
    pub fn ps2_hip_run(input: *const Ps2PatchIn, out: *mut Ps2PatchOut, opts: *const Ps2RunOpts) -> c_int;
    pub fn ps2_hip_run_batch(n: usize, input: *const Ps2PatchIn, out: *mut Ps2PatchOut,
                             opts: *const Ps2RunOpts) -> c_int;
    pub fn ps2_hip_random_coherence_histogram(bperp_nm: *const f64, n_ifg: i64, trial_values: *const f64,
                                              n_trial: i64, coh_bins: *const f64, n_rows: i64, seed: u32,
                                              device: i32, out_nr: *mut f64, out_nr_max_nz_ix: *mut f64)
                                              -> c_int;
    pub fn ps2_hip_random_coherence_histogram_sb(bperp_nm: *const f64, n_ifg: i64, image_a: *const i32,
                                                 image_b: *const i32, n_image: i64, trial_values: *const f64,
                                                 n_trial: i64, coh_bins: *const f64, n_rows: i64, seed: u32,
                                                 device: i32, out_nr: *mut f64, out_nr_max_nz_ix: *mut f64)
                                                 -> c_int;
    pub fn ps2_hip_topofit_rows(cpx: *const f64, bperp: *const f64, row_invariant: i32, n_rows: i64,
                                n_ifg: i64, n_trial_wraps: f64, device: i32, k_out: *mut f64,
                                c_out: *mut f64, coh_out: *mut f64, res_out: *mut f32) -> c_int;
    pub fn ps2_hip_topofit_coh_row_invariant(cpx: *const f64, bperp_vec: *const f64, n_rows: i64,
                                             n_ifg: i64, n_trial_wraps: f64, device: i32,
                                             coh_out: *mut f64) -> c_int;
    pub fn ps2_hip_topofit_rows_noalloc(d_cpx: *const f64, d_bperp: *const f64, row_invariant: i32,
                                        n_rows: i64, n_ifg: i64, n_trial_wraps: f64, device: i32,
                                        stream: *mut c_void, d_k_out: *mut f64, d_c_out: *mut f64,
                                        d_coh_out: *mut f64, d_res_out: *mut f32) -> c_int;
    pub fn ps2_hip_grid_accumulate(ph_weight: *const f32, grid_lin: *const i64, n_rows: i64, n_ifg: i64,
                                   n_i: i64, n_j: i64, device: i32, out: *mut f32) -> c_int;
    pub fn ps2_hip_histogram_with_centers(values: *const f64, n: i64, centers: *const f64, n_centers: i64,
                                          out: *mut f64) -> c_int;
    pub fn ps2_hip_normalize_phase(ph: *const f32, n_rows: i64, n_cols: i64, cols: *const i64, n_sel: i64,
                                   n_threads: i32, ph_nm: *mut f32, amp: *mut f32) -> c_int;
    pub fn ps2_hip_select_columns_f64(src: *const f64, n_rows: i64, n_cols: i64, cols: *const i64,
                                      n_sel: i64, n_threads: i32, dst: *mut f64) -> c_int;
    pub fn ps2_hip_session_create(input: *const Ps2PatchIn, opts: *const Ps2RunOpts,
                                  out: *mut *mut Ps2Session) -> c_int;
    pub fn ps2_hip_session_run(s: *mut Ps2Session, out_i_loop: *mut i64, out_loop_seconds: *mut f64) -> c_int;
    pub fn ps2_hip_session_download(s: *mut Ps2Session, out: *mut Ps2PatchOut) -> c_int;
    pub fn ps2_hip_session_run_iterations(s: *mut Ps2Session, state: *mut Ps2PatchOut, fn_: Ps2IterationFn,
                                          user: *mut c_void, out_i_loop: *mut i64) -> c_int;
    pub fn ps2_hip_session_destroy(s: *mut Ps2Session);
    pub fn ps2_hip_sessions_run(n: usize, sessions: *const *mut Ps2Session, i_loops: *mut i64,
                                seconds: *mut f64) -> c_int;
    pub fn ps2_hip_session_export_device(s: *mut Ps2Session, out: *mut Ps2DeviceOutputs) -> c_int;
    pub fn ps2_hip_session_set_timing(s: *mut Ps2Session, enable: i32) -> c_int;
    pub fn ps2_hip_session_tiles(s: *mut Ps2Session, out: *mut Ps2Tiles) -> c_int;
    pub fn ps2_hip_session_kernel_times(s: *mut Ps2Session, ms: *mut f64, launches: *mut i64) -> c_int;
    pub fn ps2_hip_step_fill_grid(s: *mut Ps2Session, k_ps: *const f64, weighting: *const f64,
                                  ph_grid_out: *mut c_void) -> c_int;
    pub fn ps2_hip_step_clap(s: *mut Ps2Session, ph_grid: *const c_void, ph_filt_out: *mut c_void) -> c_int;
    pub fn ps2_hip_step_fit(s: *mut Ps2Session, ph_filt: *const c_void, out: *mut Ps2PatchOut) -> c_int;
}

/// Message of the last failure on this thread (empty if none).
pub fn last_error() -> String {
    unsafe { CStr::from_ptr(ps2_hip_last_error()) }.to_string_lossy().into_owned()
}

fn check(rc: c_int) -> Result<(), String> {
    if rc == PS2_OK {
        Ok(())
    } else {
        Err(last_error())
    }
}

/// The loaded library implements this crate's ABI.
pub fn check_abi() -> Result<(), String> {
    let v = unsafe { ps2_hip_abi_version() };
    if v == PS2_HIP_ABI_VERSION {
        Ok(())
    } else {
        Err(format!("libpystamps_hip ABI {v}, this binding expects {PS2_HIP_ABI_VERSION}"))
    }
}

/// Whole gamma loop of one patch: replaces run_stage2_native_inner's loop
/// (native_stage2.rs:242-406) and fills the write_pm1 arguments.
pub fn run(input: &Ps2PatchIn, out: &mut Ps2PatchOut, opts: &Ps2RunOpts) -> Result<(), String> {
    check(unsafe { ps2_hip_run(input, out, opts) })
}

/// Several patches in one call, each with its own convergence (the batched
/// form of pystamps-native.rs:158-176).
pub fn run_batch(input: &[Ps2PatchIn], out: &mut [Ps2PatchOut], opts: &Ps2RunOpts) -> Result<(), String> {
    if input.len() != out.len() {
        return Err("run_batch: input and output counts differ".into());
    }
    check(unsafe { ps2_hip_run_batch(input.len(), input.as_ptr(), out.as_mut_ptr(), opts) })
}

/// Options with the autotuner's tile cache named (the CString must outlive
/// the calls that use the options).
pub fn opts_with_autotune(form: i32, device: i32, cache: &CString) -> Ps2RunOpts {
    Ps2RunOpts { form, device, autotune: 1, autotune_cache: cache.as_ptr(), ..Default::default() }
}

/// A patch resident in HBM; destroyed on drop.
pub struct Session {
    raw: *mut Ps2Session,
}

impl Session {
    pub fn new(input: &Ps2PatchIn, opts: &Ps2RunOpts) -> Result<Session, String> {
        let mut raw: *mut Ps2Session = std::ptr::null_mut();
        check(unsafe { ps2_hip_session_create(input, opts, &mut raw) })?;
        Ok(Session { raw })
    }

    /// Runs the loop; returns (i_loop, device seconds).
    pub fn run(&mut self) -> Result<(i64, f64), String> {
        let (mut i_loop, mut secs) = (0_i64, 0.0_f64);
        check(unsafe { ps2_hip_session_run(self.raw, &mut i_loop, &mut secs) })?;
        Ok((i_loop, secs))
    }

    pub fn download(&mut self, out: &mut Ps2PatchOut) -> Result<(), String> {
        check(unsafe { ps2_hip_session_download(self.raw, out) })
    }

    /// The loop with the state after every iteration written into `state`
    /// and passed to `on_iteration(iteration, state)` (the
    /// PYSTAMPS_STAGE2_NATIVE_DEBUG_PM snapshots); returns the iteration count.
    pub fn run_iterations<F: FnMut(i64, &Ps2PatchOut)>(&mut self, state: &mut Ps2PatchOut, mut on_iteration: F)
                                                       -> Result<i64, String> {
        unsafe extern "C" fn tramp<F: FnMut(i64, &Ps2PatchOut)>(user: *mut c_void, it: i64,
                                                                 st: *mut Ps2PatchOut) -> c_int {
            let f = &mut *(user as *mut F);
            f(it, &*st);
            0
        }
        let mut i_loop = 0i64;
        check(unsafe {
            ps2_hip_session_run_iterations(self.raw, state, Some(tramp::<F>),
                                           &mut on_iteration as *mut F as *mut c_void, &mut i_loop)
        })?;
        Ok(i_loop)
    }

    pub fn tiles(&mut self) -> Result<Ps2Tiles, String> {
        let mut t = Ps2Tiles::default();
        check(unsafe { ps2_hip_session_tiles(self.raw, &mut t) })?;
        Ok(t)
    }

    pub fn as_raw(&mut self) -> *mut Ps2Session {
        self.raw
    }
}

impl Drop for Session {
    fn drop(&mut self) {
        unsafe { ps2_hip_session_destroy(self.raw) }
    }
}

/// Runs the loops of several sessions concurrently; returns their i_loop.
pub fn sessions_run(sessions: &mut [Session]) -> Result<Vec<i64>, String> {
    let raws: Vec<*mut Ps2Session> = sessions.iter_mut().map(|s| s.raw).collect();
    let mut loops = vec![0_i64; raws.len()];
    check(unsafe { ps2_hip_sessions_run(raws.len(), raws.as_ptr(), loops.as_mut_ptr(), std::ptr::null_mut()) })?;
    Ok(loops)
}
