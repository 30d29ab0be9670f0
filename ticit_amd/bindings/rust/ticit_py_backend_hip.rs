// The `hip` arm of ticit_py's backend dispatch, over the ticit-hip-sys crate.
//
// This is an integration fragment for a ticit checkout, not a crate of its
// own: each block names the reference location it joins. It builds with
// ticit_py once `ticit-hip-sys` is a dependency behind a `hip` feature:
//
//     [features]
//     hip = ["dep:ticit-hip-sys"]
//     [dependencies]
//     ticit-hip-sys = { path = "<this repository>/ticit_amd/bindings/rust/ticit-hip-sys", optional = true }
//
// Rows come back in ticit's SampleResult layout, so every accessor of
// PySampleResult works unchanged.

// --- ticit_py/src/lib.rs:221-231, enum Backend: one more variant ---------
#[cfg(feature = "hip")]
Hip(Box<ticit_hip_sys::sampler::HipSampler>),

// --- ticit_py/src/lib.rs:767-830, the backend match: the "hip" arm -------
"hip" => {
    #[cfg(feature = "hip")]
    {
        use ticit_hip_sys::{sampler::HipSampler, ticit_hip_options, ticit_hip_parity, ticit_hip_reference_bits};
        // The library plans the same text with its own front end (or a
        // host serializes its FactoredInstructionProgram as the plan blob of
        // include/ticit_plan.h and calls compile_blob_ex instead).
        let mut blob: *mut u64 = std::ptr::null_mut();
        let mut nwords = 0usize;
        let status = unsafe {
            ticit_hip_sys::ticit_hip_plan_text_postselected(
                text.as_ptr() as *const _, text.len(), mask.as_ptr(), mask.len(), &mut blob, &mut nwords)
        };
        if status != ticit_hip_sys::TICIT_HIP_OK {
            return Err(hip_error(status));
        }
        let plan = unsafe { std::slice::from_raw_parts(blob, nwords) }.to_vec();
        unsafe { ticit_hip_sys::ticit_hip_blob_free(blob) };
        let options = ticit_hip_options { observable: observable as u32, ..Default::default() };
        let pin_records: Vec<Vec<u64>> =
            pin_measurements.iter().map(|p| p.records.iter().map(|&r| r as u64).collect()).collect();
        let pins: Vec<ticit_hip_parity> = pin_measurements
            .iter()
            .zip(&pin_records)
            .map(|(p, r)| ticit_hip_parity {
                records: r.as_ptr(), nrecords: r.len(), value: i32::from(p.value), reserved: 0,
            })
            .collect();
        let reference = ticit_hip_reference_bits {
            detectors: expected_detectors.as_ptr(),
            ndetectors: expected_detectors.len(),
            observables: expected_observables.as_ptr(),
            nobservables: expected_observables.len(),
            normalize_syndromes: i32::from(normalize_syndromes),
            reserved: 0,
        };
        let sampler = HipSampler::compile_blob_ex(&plan, &options, &pins, Some(&reference))
            .map_err(|e| hip_error_of(e))?;
        let postselection = sampler.info().map_err(|e| hip_error_of(e))?.postselected_detectors > 0;
        (Backend::Hip(Box::new(sampler)), "hip", postselection)
    }
    #[cfg(not(feature = "hip"))]
    {
        return Err(PyRuntimeError::new_err("the hip backend requires ticit_py built with Cargo feature `hip`"));
    }
}

// --- ticit_py/src/lib.rs:263-300, PyProgram::run: the sampling arm --------
#[cfg(feature = "hip")]
Backend::Hip(sampler) => {
    let seed = seed.unwrap_or_else(random_seed);
    let rows = if keep_records {
        sampler.sample_with_seed(shots, seed, bit_packed)
    } else {
        sampler.sample_counts_with_seed(shots, seed)
    }
    .map_err(|e| hip_error_of(e))?;
    Ok(ticit::SampleResult {
        measurements: rows.measurements,
        detectors: rows.detectors,
        observables: rows.observables,
        exp_vals: rows.exp_vals,
        observable_ones: rows.observable_ones,
        record_rows: rows.record_rows as usize,
        bit_packed: rows.bit_packed,
        counts: ticit::SampleCounts {
            shots: rows.shots,
            discarded: rows.discarded,
            accepted: rows.accepted,
            logical_errors: rows.logical_errors,
        },
        timing: ticit::SamplingTiming { sample_s: rows.sample_s, ..Default::default() },
        active_threads: 1,
    })
}

// --- ticit_py/src/lib.rs, next to ticit_error: TicitError kinds by status --
#[cfg(feature = "hip")]
fn hip_error_of(e: ticit_hip_sys::sampler::HipError) -> PyErr {
    match e.status {
        1 | 2 => PyValueError::new_err(e.message),
        _ => PyRuntimeError::new_err(e.message),
    }
}

#[cfg(feature = "hip")]
fn hip_error(status: ticit_hip_sys::ticit_hip_status) -> PyErr {
    let message = unsafe { std::ffi::CStr::from_ptr(ticit_hip_sys::ticit_hip_last_error()) }
        .to_string_lossy()
        .into_owned();
    hip_error_of(ticit_hip_sys::sampler::HipError { status, message })
}

// --- src/main.rs:10-14, the CLI's Backend enum: `--backend hip` -----------
// enum Backend { Cpu, Gpu, Hip }
// and in main's dispatch, `Backend::Hip` builds a HipSampler from the
// circuit text as above and calls sample_with_seed / sample_counts_with_seed
// with the CLI's --shots and --seed, printing the same key/value lines.
