//! Raw bindings to `include/ticit_hip.h`, the C ABI of the MI355X sampling
//! backend, for a ticit host written in Rust.
//!
//! The declarations mirror the header field for field and argument for
//! argument (`tests/test_rust_binding.py` checks that against the header).
//! Status codes map to `TicitError` kinds (`src/errors.rs:9-57` of ticit):
//! 1 InvalidInput, 2 Parse, 3 Unsupported, 4 Internal, 5 and 6 device
//! failures; the message of the last failing call on a thread comes from
//! [`ticit_hip_last_error`].
//!
//! The safe layer ([`sampler`]) is what ticit's `Sampler` would hold for a
//! `hip` backend: `Circuit::compile(SamplerOptions)` becomes
//! [`sampler::HipSampler::compile_blob`] or [`sampler::HipSampler::compile_text`],
//! and `Sampler::sample_with_seed(shots, seed, bit_packed)`
//! (`src/sampler/prepared.rs:663-670`) becomes
//! [`sampler::HipSampler::sample_with_seed`].
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int, c_void};

pub type ticit_hip_status = c_int;
pub const TICIT_HIP_OK: ticit_hip_status = 0;
pub const TICIT_HIP_INVALID_INPUT: ticit_hip_status = 1;
pub const TICIT_HIP_PARSE: ticit_hip_status = 2;
pub const TICIT_HIP_UNSUPPORTED: ticit_hip_status = 3;
pub const TICIT_HIP_INTERNAL: ticit_hip_status = 4;
pub const TICIT_HIP_OUT_OF_DEVICE_MEMORY: ticit_hip_status = 5;
pub const TICIT_HIP_DEVICE_ERROR: ticit_hip_status = 6;

pub const TICIT_HIP_PLANNER_FEATURES: usize = 12;
pub const TICIT_HIP_MLP_PARAMS: usize = 5187;
pub const TICIT_HIP_NFEATURES: usize = 24;

/// Opaque program handle (one compiled circuit on one device).
#[repr(C)]
pub struct ticit_hip_program {
    _private: [u8; 0],
}

/// Opaque batch handle (many programs in two captured graphs).
#[repr(C)]
pub struct ticit_hip_batch {
    _private: [u8; 0],
}

/// Opaque planner network handle.
#[repr(C)]
pub struct ticit_hip_mlp {
    _private: [u8; 0],
}

/// `SamplerOptions` (`src/sampler/prepared.rs:104-142`) plus the backend's
/// own choices; all zeros selects the defaults.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct ticit_hip_options {
    pub observable: u32,
    pub precision: u32,
    pub planner_mode: u32,
    pub tier: u32,
    pub wide_batch: u32,
    pub components: u32,
    pub group_threads: u32,
    pub lane_store: u32,
}

/// `SampleResult` (`src/sampler/prepared.rs:197-219`); the buffers belong to
/// the library until [`ticit_hip_result_free`].
#[repr(C)]
#[derive(Debug)]
pub struct ticit_hip_result {
    pub shots: u64,
    pub discarded: u64,
    pub accepted: u64,
    pub logical_errors: u64,
    pub record_rows: u64,
    pub bit_packed: i32,
    pub reserved0: i32,
    pub nmeasurements: u64,
    pub ndetectors: u64,
    pub nobservables: u64,
    pub measurements: *mut u8,
    pub detectors: *mut u8,
    pub observables: *mut u8,
    pub observable_ones: *mut u64,
    pub sample_s: f64,
    pub kernel_s: f64,
    pub nexpvals: u64,
    pub exp_vals: *mut f64,
}

/// `SamplerInfo` (`src/sampler/prepared.rs:172-195`) plus the backend's
/// launch facts.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct ticit_hip_program_info {
    pub qubits: u64,
    pub measurement_records: u64,
    pub detectors: u64,
    pub observables: u64,
    pub max_active_qubits: u64,
    pub symbols: u64,
    pub instructions: u64,
    pub measure_instructions: u64,
    pub coin_instructions: u64,
    pub tier: u32,
    pub precision: u32,
    pub wide_batch: u32,
    pub components: u32,
    pub output_bytes_per_shot: f64,
    pub state_bytes_per_shot: f64,
    pub algorithmic_bytes_per_shot: f64,
    pub state_dim: u64,
    pub component_count: u64,
    pub max_local_k: i32,
    pub planner_source: i32,
    pub planner_p_components: f64,
    pub lds_bytes_per_block: u32,
    pub blocks_per_cu: u32,
    pub table_words: [u32; 4],
    pub expectation_values: u64,
    pub postselected_detectors: u64,
    pub lane_store: u32,
    pub lane_region_bytes: u32,
}

/// `MeasurementParity` (`src/sampler/pinning.rs`): zero-based records and
/// the value their XOR takes in the noiseless circuit.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ticit_hip_parity {
    pub records: *const u64,
    pub nrecords: usize,
    pub value: i32,
    pub reserved: i32,
}

/// `SamplerOptions::{expected_detectors, expected_observables,
/// normalize_syndromes}` (`src/sampler/prepared.rs:111-116`).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ticit_hip_reference_bits {
    pub detectors: *const u8,
    pub ndetectors: usize,
    pub observables: *const u8,
    pub nobservables: usize,
    pub normalize_syndromes: i32,
    pub reserved: i32,
}

/// One (program, shots, seed) job of a batch, with device row buffers.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ticit_hip_batch_item {
    pub program: *mut ticit_hip_program,
    pub shots: u64,
    pub seed: u64,
    pub measurements: *mut u8,
    pub detectors: *mut u8,
    pub observables: *mut u8,
    pub observable_ones: *mut u64,
    pub exp_vals: *mut f64,
    pub accepted: *mut u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct ticit_hip_batch_options {
    pub bit_packed: i32,
    pub keep_records: i32,
    pub autotune: u32,
    pub autotune_shots: u32,
    pub group_grid_percent: u32,
    pub wide_cu_percent: u32,
    pub reserved: [u32; 2],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ticit_hip_batch_info {
    pub items: u64,
    pub lane_items: u64,
    pub wide_items: u64,
    pub group_buckets: u64,
    pub bucket_tps: [u32; 16],
    pub bucket_items: [u32; 16],
    pub bucket_region_bytes: [u32; 16],
    pub bucket_grid: [u32; 16],
    pub bucket_tuned_ms: [[f64; 8]; 16],
}

extern "C" {
    pub fn ticit_hip_last_error() -> *const c_char;

    // plan_circuit (src/circuit/mod.rs:274-327) for hosts without a planner
    pub fn ticit_hip_plan_text(text: *const c_char, len: usize, blob: *mut *mut u64, nwords: *mut usize)
        -> ticit_hip_status;
    pub fn ticit_hip_plan_text_postselected(text: *const c_char, len: usize, mask: *const u8, mask_len: usize,
        blob: *mut *mut u64, nwords: *mut usize) -> ticit_hip_status;
    pub fn ticit_hip_blob_free(blob: *mut u64);

    // Circuit::compile (src/circuit/mod.rs:75-77, src/sampler/prepared.rs:492-626)
    pub fn ticit_hip_program_create(blob: *const u64, nwords: usize, options: *const ticit_hip_options,
        out: *mut *mut ticit_hip_program) -> ticit_hip_status;
    pub fn ticit_hip_program_create_pinned(blob: *const u64, nwords: usize, options: *const ticit_hip_options,
        pins: *const ticit_hip_parity, npins: usize, out: *mut *mut ticit_hip_program) -> ticit_hip_status;
    pub fn ticit_hip_program_create_ex(blob: *const u64, nwords: usize, options: *const ticit_hip_options,
        pins: *const ticit_hip_parity, npins: usize, reference: *const ticit_hip_reference_bits,
        out: *mut *mut ticit_hip_program) -> ticit_hip_status;
    // Circuit::reference_sample (src/circuit/mod.rs:89-91)
    pub fn ticit_hip_reference_sample(program: *mut ticit_hip_program, detectors: *mut u8, observables: *mut u8)
        -> ticit_hip_status;
    pub fn ticit_hip_compile_text(text: *const c_char, len: usize, options: *const ticit_hip_options,
        out: *mut *mut ticit_hip_program) -> ticit_hip_status;
    pub fn ticit_hip_program_destroy(program: *mut ticit_hip_program);
    // Sampler::info (src/sampler/prepared.rs:628-632)
    pub fn ticit_hip_program_info_get(program: *const ticit_hip_program, out: *mut ticit_hip_program_info)
        -> ticit_hip_status;

    // Sampler::sample_with_seed (src/sampler/prepared.rs:663-670)
    pub fn ticit_hip_sample(program: *mut ticit_hip_program, shots: u64, seed: u64, bit_packed: c_int,
        keep_records: c_int, out: *mut ticit_hip_result) -> ticit_hip_status;
    pub fn ticit_hip_sample_device(program: *mut ticit_hip_program, shots: u64, seed: u64, first_shot: u64,
        bit_packed: c_int, keep_records: c_int, meas: *mut u8, det: *mut u8, obs: *mut u8,
        observable_ones: *mut u64, exp_vals: *mut f64, stream: *mut c_void) -> ticit_hip_status;
    pub fn ticit_hip_sample_device_postselected(program: *mut ticit_hip_program, shots: u64, seed: u64,
        first_shot: u64, bit_packed: c_int, keep_records: c_int, meas: *mut u8, det: *mut u8, obs: *mut u8,
        observable_ones: *mut u64, exp_vals: *mut f64, accepted: *mut u64, stream: *mut c_void)
        -> ticit_hip_status;
    // no counterpart in the reference: for callers that cannot allocate
    pub fn ticit_hip_sample_device_postselected_noalloc(program: *mut ticit_hip_program, shots: u64, seed: u64,
        first_shot: u64, bit_packed: c_int, keep_records: c_int, meas: *mut u8, det: *mut u8, obs: *mut u8,
        observable_ones: *mut u64, exp_vals: *mut f64, accepted: *mut u64, stream: *mut c_void)
        -> ticit_hip_status;
    pub fn ticit_hip_sample_injected(program: *mut ticit_hip_program, shots: u64, exo: *const u64,
        uniforms: *const f64, coins: *const u64, bit_packed: c_int, out: *mut ticit_hip_result,
        probs: *mut f64) -> ticit_hip_status;
    pub fn ticit_hip_final_state(program: *mut ticit_hip_program, shots: u64, exo: *const u64,
        uniforms: *const f64, coins: *const u64, re: *mut f64, im: *mut f64, k: *mut c_int)
        -> ticit_hip_status;

    pub fn ticit_hip_batch_create(items: *const ticit_hip_batch_item, count: usize,
        options: *const ticit_hip_batch_options, out: *mut *mut ticit_hip_batch) -> ticit_hip_status;
    pub fn ticit_hip_batch_launch(batch: *mut ticit_hip_batch, seed_offset: u64, stream: *mut c_void)
        -> ticit_hip_status;
    pub fn ticit_hip_batch_wait(batch: *mut ticit_hip_batch) -> ticit_hip_status;
    pub fn ticit_hip_batch_info_get(batch: *const ticit_hip_batch, out: *mut ticit_hip_batch_info)
        -> ticit_hip_status;
    pub fn ticit_hip_batch_destroy(batch: *mut ticit_hip_batch);
    pub fn ticit_hip_sample_batch(programs: *const *mut ticit_hip_program, shots: *const u64, seeds: *const u64,
        count: usize, bit_packed: c_int, keep_records: c_int, results: *mut ticit_hip_result)
        -> ticit_hip_status;
    pub fn ticit_hip_result_free(result: *mut ticit_hip_result);

    pub fn ticit_hip_plan_features(blob: *const u64, nwords: usize, out: *mut f64, nout: usize)
        -> ticit_hip_status;
    pub fn ticit_hip_planner_features(blob: *const u64, nwords: usize, out: *mut f32) -> ticit_hip_status;
    pub fn ticit_hip_mlp_create(params: *const f32, out: *mut *mut ticit_hip_mlp) -> ticit_hip_status;
    pub fn ticit_hip_mlp_train(mlp: *mut ticit_hip_mlp, x: *const f32, y: *const i32, t: *const f32, n: usize,
        order: *const u32, steps: usize, batch: u32, lr: f32, time_weight: f32, losses: *mut f32)
        -> ticit_hip_status;
    pub fn ticit_hip_mlp_gradients(mlp: *mut ticit_hip_mlp, x: *const f32, y: *const i32, t: *const f32,
        n: usize, time_weight: f32, grads: *mut f32, loss: *mut f32) -> ticit_hip_status;
    pub fn ticit_hip_mlp_predict(mlp: *const ticit_hip_mlp, x: *const f32, n: usize, out: *mut f32)
        -> ticit_hip_status;
    pub fn ticit_hip_mlp_params_get(mlp: *const ticit_hip_mlp, out: *mut f32) -> ticit_hip_status;
    pub fn ticit_hip_mlp_destroy(mlp: *mut ticit_hip_mlp);
    pub fn ticit_hip_planner_set(params: *const f32, feature_mean: *const f32, feature_scale: *const f32,
        margin: f32) -> ticit_hip_status;

    pub fn ticit_hip_device_query(props: *mut i64) -> c_int;
}

/// A safe layer over the raw calls: the shape a ticit `Backend::Hip` holds.
pub mod sampler {
    use super::*;
    use std::ffi::CStr;

    /// A failing call: the status (a `TicitError` kind) and the library's
    /// message for it.
    #[derive(Debug, Clone)]
    pub struct HipError {
        pub status: ticit_hip_status,
        pub message: String,
    }

    fn check(status: ticit_hip_status) -> Result<(), HipError> {
        if status == TICIT_HIP_OK {
            return Ok(());
        }
        // SAFETY: the library returns a NUL-terminated thread-local string.
        let message = unsafe { CStr::from_ptr(ticit_hip_last_error()) }.to_string_lossy().into_owned();
        Err(HipError { status, message })
    }

    /// Host rows of one sampling call, laid out as ticit's `SampleResult`:
    /// row-major, `ceil(cols / 8)` little-bit-order bytes per row when packed.
    #[derive(Debug, Clone, Default)]
    pub struct SampleRows {
        pub shots: u64,
        pub discarded: u64,
        pub accepted: u64,
        pub logical_errors: u64,
        pub record_rows: u64,
        pub bit_packed: bool,
        pub measurements: Vec<u8>,
        pub detectors: Vec<u8>,
        pub observables: Vec<u8>,
        pub observable_ones: Vec<u64>,
        pub exp_vals: Vec<f64>,
        pub sample_s: f64,
    }

    /// A compiled program on the current HIP device.
    pub struct HipSampler {
        handle: *mut ticit_hip_program,
    }

    // A program handle is used by one thread at a time, as `&mut Sampler`.
    unsafe impl Send for HipSampler {}

    impl HipSampler {
        /// From a plan blob (`include/ticit_plan.h`) serialized by the host's
        /// own planner.
        pub fn compile_blob(blob: &[u64], options: &ticit_hip_options) -> Result<Self, HipError> {
            let mut handle = std::ptr::null_mut();
            check(unsafe { ticit_hip_program_create(blob.as_ptr(), blob.len(), options, &mut handle) })?;
            Ok(Self { handle })
        }

        /// From circuit text, planned by the library's front end.
        pub fn compile_text(text: &str, options: &ticit_hip_options) -> Result<Self, HipError> {
            let mut handle = std::ptr::null_mut();
            check(unsafe {
                ticit_hip_compile_text(text.as_ptr() as *const c_char, text.len(), options, &mut handle)
            })?;
            Ok(Self { handle })
        }

        /// `SamplerOptions` with pins and reference bits
        /// (`ticit_hip_program_create_ex`).
        pub fn compile_blob_ex(blob: &[u64], options: &ticit_hip_options, pins: &[ticit_hip_parity],
                               reference: Option<&ticit_hip_reference_bits>) -> Result<Self, HipError> {
            let mut handle = std::ptr::null_mut();
            let r = reference.map_or(std::ptr::null(), |b| b as *const _);
            check(unsafe {
                ticit_hip_program_create_ex(blob.as_ptr(), blob.len(), options, pins.as_ptr(), pins.len(), r,
                                            &mut handle)
            })?;
            Ok(Self { handle })
        }

        pub fn info(&self) -> Result<ticit_hip_program_info, HipError> {
            let mut out = ticit_hip_program_info::default();
            check(unsafe { ticit_hip_program_info_get(self.handle, &mut out) })?;
            Ok(out)
        }

        /// `Sampler::sample_with_seed(shots, seed, bit_packed)`.
        pub fn sample_with_seed(&mut self, shots: u64, seed: u64, bit_packed: bool) -> Result<SampleRows, HipError> {
            self.sample_impl(shots, seed, true, bit_packed)
        }

        /// `Sampler::sample_counts_with_seed(shots, seed)`: counts only.
        pub fn sample_counts_with_seed(&mut self, shots: u64, seed: u64) -> Result<SampleRows, HipError> {
            self.sample_impl(shots, seed, false, false)
        }

        fn sample_impl(&mut self, shots: u64, seed: u64, keep_records: bool, bit_packed: bool)
                       -> Result<SampleRows, HipError> {
            // SAFETY: a zeroed result is the library's expected input; every
            // buffer it fills is copied out, then released by result_free.
            let mut r: ticit_hip_result = unsafe { std::mem::zeroed() };
            check(unsafe {
                ticit_hip_sample(self.handle, shots, seed, c_int::from(bit_packed), c_int::from(keep_records),
                                 &mut r)
            })?;
            let cols = |n: u64| (if bit_packed { (n + 7) / 8 } else { n }) as usize;
            let rows = r.record_rows as usize;
            let copy_u8 = |p: *mut u8, len: usize| {
                if p.is_null() || len == 0 { Vec::new() } else { unsafe { std::slice::from_raw_parts(p, len) }.to_vec() }
            };
            let out = SampleRows {
                shots: r.shots,
                discarded: r.discarded,
                accepted: r.accepted,
                logical_errors: r.logical_errors,
                record_rows: r.record_rows,
                bit_packed,
                measurements: copy_u8(r.measurements, rows * cols(r.nmeasurements)),
                detectors: copy_u8(r.detectors, rows * cols(r.ndetectors)),
                observables: copy_u8(r.observables, rows * cols(r.nobservables)),
                observable_ones: if r.observable_ones.is_null() {
                    Vec::new()
                } else {
                    unsafe { std::slice::from_raw_parts(r.observable_ones, r.nobservables as usize) }.to_vec()
                },
                exp_vals: if r.exp_vals.is_null() {
                    Vec::new()
                } else {
                    unsafe { std::slice::from_raw_parts(r.exp_vals, rows * r.nexpvals as usize) }.to_vec()
                },
                sample_s: r.sample_s,
            };
            unsafe { ticit_hip_result_free(&mut r) };
            Ok(out)
        }

        /// `Circuit::reference_sample` on this program (no pins).
        pub fn reference_sample(&mut self) -> Result<(Vec<u8>, Vec<u8>), HipError> {
            let info = self.info()?;
            let mut det = vec![0u8; info.detectors as usize];
            let mut obs = vec![0u8; info.observables as usize];
            check(unsafe { ticit_hip_reference_sample(self.handle, det.as_mut_ptr(), obs.as_mut_ptr()) })?;
            Ok((det, obs))
        }
    }

    impl Drop for HipSampler {
        fn drop(&mut self) {
            unsafe { ticit_hip_program_destroy(self.handle) };
        }
    }
}
