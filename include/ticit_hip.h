/*
 * ticit_hip: C ABI of the MI355X sampling backend for ticit's compiled
 * Clifford+T sampler.
 *
 * The entry points replace, for a `hip` backend, the surfaces a ticit host
 * binds today:
 *   - Circuit::compile(SamplerOptions) -> Sampler
 *       (src/circuit/mod.rs:75-77, src/sampler/prepared.rs:492-626)
 *   - Sampler::sample_with_seed(shots, seed, bit_packed) -> SampleResult
 *       (src/sampler/prepared.rs:663-670, 692-859)
 *   - Sampler::info() -> SamplerInfo (src/sampler/prepared.rs:628-632)
 *   - the free GPU functions ticit::gpu::sample_circuit* (src/gpu/mod.rs:235-590)
 * All arguments are plain pointers and sizes; no C++ or torch types cross
 * the boundary and no exception escapes it. Error kinds mirror TicitError
 * (src/errors.rs:9-57). A program handle is single-threaded, like the
 * `&mut self` Sampler (src/sampler/prepared.rs:444-448); distinct handles may
 * be used from different threads.
 */
#ifndef TICIT_HIP_H
#define TICIT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ticit_hip_status {
  TICIT_HIP_OK = 0,
  TICIT_HIP_INVALID_INPUT = 1, /* TicitError::InvalidInput */
  TICIT_HIP_PARSE = 2,         /* TicitError::Parse */
  TICIT_HIP_UNSUPPORTED = 3,   /* TicitError::Unsupported */
  TICIT_HIP_INTERNAL = 4,      /* TicitError::Internal */
  TICIT_HIP_OUT_OF_DEVICE_MEMORY = 5,
  TICIT_HIP_DEVICE_ERROR = 6
} ticit_hip_status;

typedef struct ticit_hip_program ticit_hip_program;

/* Mirrors SamplerOptions (src/sampler/prepared.rs:104-142) plus the backend
 * knobs of the hip path. Zero-initialize for defaults. */
typedef struct ticit_hip_options {
  uint32_t observable;   /* observable counted as a logical error */
  uint32_t precision;    /* 0: FP64 amplitudes; 1: FP32 amplitudes, FP64 reductions */
  uint32_t planner_mode; /* 0: heuristic tier choice; 1: learned planner */
  uint32_t tier;         /* 0: automatic (lane tier when its tables fit LDS, and for programs
                            dominated by record/detector/coin steps up to dense width 8 with
                            its tables in device memory; group tier up to active width 10;
                            wide tier above); 1: lane; 2: group; 3: wide */
  uint32_t wide_batch;   /* cap on wide-tier shots in flight (0 = memory budget) */
  uint32_t components;   /* 0: as planned (use_active_components, src/planner.rs:1482-1484);
                            1: dense active state; 2: tensor components
                            (src/component_plan.rs, src/sampler/batch/runtime.rs:388-552) */
  uint32_t group_threads; /* group tier threads per shot: 8, 16, 32, 64, 128, 256 or 512 (0 = the
                             fewest that hold the widest component and fit the LDS) */
  uint32_t lane_store;    /* lane tier only: 0 automatic (LDS when the tables fit, else device
                             memory); 1 all tables in LDS; 2 bit tables and row staging in device
                             memory, amplitudes in LDS (width <= 5); 3 everything in device
                             memory (dense widths 6 to 8) */
} ticit_hip_options;

/* Mirrors SampleResult (src/sampler/prepared.rs:197-219). Buffers are owned
 * by the library until ticit_hip_result_free. Rows are row-major; packed
 * rows hold ceil(columns / 8) bytes in little bit order
 * (src/sampler/prepared.rs:909-941). */
typedef struct ticit_hip_result {
  uint64_t shots, discarded, accepted, logical_errors;
  uint64_t record_rows;
  int32_t bit_packed;
  int32_t reserved0;
  uint64_t nmeasurements, ndetectors, nobservables; /* column counts */
  uint8_t* measurements;
  uint8_t* detectors;
  uint8_t* observables;
  uint64_t* observable_ones;
  double sample_s; /* wall time of the sampling call */
  double kernel_s; /* device time of the sampling kernels */
  /* SampleResult::exp_vals (src/sampler/prepared.rs:218,1020-1025):
   * [record_rows][nexpvals] EXP_VAL values, row-major, kept with the rows */
  uint64_t nexpvals;
  double* exp_vals;
} ticit_hip_result;

/* Mirrors SamplerInfo (src/sampler/prepared.rs:161-189). */
typedef struct ticit_hip_program_info {
  uint64_t qubits, measurement_records, detectors, observables;
  uint64_t max_active_qubits, symbols, instructions;
  uint64_t measure_instructions, coin_instructions;
  uint32_t tier;      /* tier the sampler runs on */
  uint32_t precision;
  uint32_t wide_batch; /* wide tier: shots in flight per pass */
  uint32_t components; /* 1 when the state runs as tensor components */
  /* Algorithmic bytes per shot: packed output rows plus, when the tier keeps
   * the state in device memory, S * sum over amplitude instructions of
   * f * 2^k (S = 16 FP64 / 8 FP32 bytes per amplitude; f = 2 rotation,
   * 3 promotion, 2.5 nondiagonal and 1.5 diagonal measurement; a merge
   * counts 2^a + 2^b + 2^(a+b)). */
  double output_bytes_per_shot;
  double state_bytes_per_shot; /* the state term alone, whatever the tier */
  double algorithmic_bytes_per_shot;
  uint64_t state_dim;       /* amplitudes per shot: 2^max_k dense, else sum of component capacities */
  uint64_t component_count; /* 1 dense */
  int32_t max_local_k;      /* widest amplitude vector an instruction acts on */
  int32_t planner_source;   /* layout decided by: 0 the plan's heuristic flag, 1 the learned
                               planner, 2 the caller (options.components) */
  double planner_p_components; /* learned planner's P(components faster), -1 when not run */
  uint32_t lds_bytes_per_block;  /* lane/group tier launch with bit-packed rows (0 wide) */
  uint32_t blocks_per_cu;        /* resident workgroups per CU of that launch */
  uint32_t table_words[4];       /* per-shot bit-table words: partials, slots, records, detectors */
  uint64_t expectation_values;   /* EXP_VAL slots per shot (src/circuit/mod.rs:122) */
  uint64_t postselected_detectors; /* detectors whose firing discards a shot (src/circuit/mod.rs:329-333) */
  /* Lane tier: where a wave's tables live. 0 all in LDS; 1 the bit tables
   * and row staging in device memory, amplitudes in LDS; 2 all of them in
   * device memory (states of 2^6 to 2^8 amplitudes, bit tables too large
   * for LDS). */
  uint32_t lane_store;
  uint32_t lane_region_bytes; /* device memory per resident wave (0 in LDS), bit-packed rows */
} ticit_hip_program_info;

/* Thread-local message of the last failing call on this thread. */
const char* ticit_hip_last_error(void);

/* Front end: parse and plan .ticit text into a plan blob
 * (include/ticit_plan.h). Replaces plan_circuit (src/circuit/mod.rs:274-327)
 * for hosts that do not serialize their own FactoredInstructionProgram. The
 * blob is freed with ticit_hip_blob_free. */
ticit_hip_status ticit_hip_plan_text(const char* text, size_t len, uint64_t** blob, size_t* nwords);
/* The same with detector postselection (plan_circuit's postselection_mask,
 * src/circuit/mod.rs:273-327; SamplerOptions::postselection_mask,
 * src/sampler/prepared.rs:108-109): mask[d] != 0 postselects detector d
 * (missing entries are zero), unioned with the circuit's DISCARD lines. */
ticit_hip_status ticit_hip_plan_text_postselected(const char* text, size_t len, const uint8_t* mask,
                                                  size_t mask_len, uint64_t** blob, size_t* nwords);
void ticit_hip_blob_free(uint64_t* blob);

/* Sampler::new / from_input (src/sampler/prepared.rs:492-626): validates the
 * blob, uploads its tables to the current HIP device and picks a tier. */
ticit_hip_status ticit_hip_program_create(const uint64_t* blob, size_t nwords, const ticit_hip_options* options,
                                          ticit_hip_program** out);

/* A pinned measurement parity (MeasurementParity, src/sampler/pinning.rs:35-48):
 * the XOR of the zero-based measurement records must equal `value` in the
 * noiseless circuit. */
typedef struct ticit_hip_parity {
  const uint64_t* records;
  size_t nrecords;
  int32_t value;
  int32_t reserved;
} ticit_hip_parity;

/* ticit_hip_program_create with SamplerOptions::pin_measurements
 * (src/sampler/prepared.rs:118-135, solved by plan_pinned_measurements,
 * src/sampler/pinning.rs:227-289): each pinned branch is set to the XOR of
 * earlier branches instead of being drawn. Fails (invalid input) when a
 * constraint is decided the other way by the noiseless circuit
 * ("deterministic") or names a record the circuit never writes; sampling
 * fails when a pinned active branch is not a fair coin ("not one half"). */
ticit_hip_status ticit_hip_program_create_pinned(const uint64_t* blob, size_t nwords,
                                                const ticit_hip_options* options, const ticit_hip_parity* pins,
                                                size_t npins, ticit_hip_program** out);

/* SamplerOptions::normalize_syndromes, expected_detectors and
 * expected_observables (src/sampler/prepared.rs:111-116). A nonzero
 * expected bit inverts that column in every row, in detector postselection
 * (a postselected detector discards a shot when its value differs from the
 * expected bit) and in the observable counts (src/sampler/prepared.rs:909-974,
 * src/sampler/batch/postselect.rs:631-660). An empty list inverts nothing; a
 * non-empty one has one entry per detector / observable.
 * normalize_syndromes = 1 takes the bits from the noiseless reference sample
 * (ticit_hip_reference_sample) and excludes explicit bits
 * (src/sampler/prepared.rs:515-550). */
typedef struct ticit_hip_reference_bits {
  const uint8_t* detectors;
  size_t ndetectors;
  const uint8_t* observables;
  size_t nobservables;
  int32_t normalize_syndromes;
  int32_t reserved;
} ticit_hip_reference_bits;

/* Sampler::new with every SamplerOptions field (src/sampler/prepared.rs:
 * 492-626): pinned parities (may be 0 / null) and reference bits (null:
 * none). Errors as the reference's: "normalize_syndromes cannot be combined
 * with expected_detectors or expected_observables", "expected_detectors has
 * length N, expected M", and those of ticit_hip_program_create_pinned. */
ticit_hip_status ticit_hip_program_create_ex(const uint64_t* blob, size_t nwords, const ticit_hip_options* options,
                                            const ticit_hip_parity* pins, size_t npins,
                                            const ticit_hip_reference_bits* reference, ticit_hip_program** out);

/* Circuit::reference_sample / reference_sample_for_program
 * (src/circuit/mod.rs:89-91, src/sampler/prepared.rs:355-411): one
 * noiseless shot (every noise symbol 0, branches drawn from the SplitMix64
 * stream started at state 1, the program's pins obeyed, no postselection),
 * run in FP64 on the device. Writes ndetectors and nobservables bytes
 * (0 / 1), before any expected bits of the program. */
ticit_hip_status ticit_hip_reference_sample(ticit_hip_program* program, uint8_t* detectors, uint8_t* observables);

/* Circuit::compile in one call: ticit_hip_plan_text + ticit_hip_program_create. */
ticit_hip_status ticit_hip_compile_text(const char* text, size_t len, const ticit_hip_options* options,
                                        ticit_hip_program** out);

void ticit_hip_program_destroy(ticit_hip_program* program);

ticit_hip_status ticit_hip_program_info_get(const ticit_hip_program* program, ticit_hip_program_info* out);

/* Sampler::sample_with_seed (src/sampler/prepared.rs:663-670) /
 * sample_counts_with_seed (keep_records = 0). Shot s draws its noise from
 * block_seed(0x7eed0000, seed, s) and its branches from
 * block_seed(0x5eed1234, seed, s): the reference sampler's streams at
 * sample_chunk_shots = 1, batch_size = 1. */
ticit_hip_status ticit_hip_sample(ticit_hip_program* program, uint64_t shots, uint64_t seed, int bit_packed,
                                  int keep_records, ticit_hip_result* out);

/* Same sampling into caller-owned device buffers, enqueued on `stream`
 * (a hipStream_t, NULL = the default stream) and not synchronized:
 * meas/det/obs are [shots][row] bytes, observable_ones is [nobservables]
 * uint64 accumulated (not cleared) by the call, exp_vals (NULL to skip) is
 * [shots][expectation_values] doubles written with the rows. first_shot
 * offsets the per-shot streams so a large job can be split into launches.
 * A wide-tier program whose shots take more than one batch of its workspace
 * runs every other batch on a stream of the library's, forked from and
 * joined to `stream` by events: the call stays ordered on `stream`. */
ticit_hip_status ticit_hip_sample_device(ticit_hip_program* program, uint64_t shots, uint64_t seed,
                                         uint64_t first_shot, int bit_packed, int keep_records, uint8_t* meas,
                                         uint8_t* det, uint8_t* obs, uint64_t* observable_ones, double* exp_vals,
                                         void* stream);

/* ticit_hip_sample_device for programs with postselected detectors
 * (src/sampler/batch/postselect.rs; src/sampler/prepared.rs:881-1028 appends
 * only the survivors): the surviving shots' rows (and exp_vals) are written to
 * the front of meas/det/obs/exp_vals in shot order and their number to
 * *accepted (a device uint64, overwritten); observable_ones counts survivors
 * only. Rows pass through a staging buffer the program owns (allocated on
 * first use and grown to the largest call: not for graph capture); calls on
 * one program are ordered on it by an event (and a lock between host
 * threads). Programs without postselection
 * behave as ticit_hip_sample_device with *accepted = shots. */
ticit_hip_status ticit_hip_sample_device_postselected(ticit_hip_program* program, uint64_t shots, uint64_t seed,
                                                     uint64_t first_shot, int bit_packed, int keep_records,
                                                     uint8_t* meas, uint8_t* det, uint8_t* obs,
                                                     uint64_t* observable_ones, double* exp_vals,
                                                     uint64_t* accepted, void* stream);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ticit_hip_sample_device_postselected with the same
 * arguments and results, but its staging (keep flags, compaction offsets and
 * the rows of every shot) lives in a 32 MiB device array the library holds
 * for the whole process, one per GPU, and the call allocates and frees
 * nothing. A call whose staging exceeds 32 MiB (about 1 byte per shot plus
 * its packed row bytes) fails with TICIT_HIP_INVALID_INPUT: split the shots
 * with first_shot. Lane- and group-tier programs only (the wide tier's
 * per-shot states live in a workspace the library allocates:
 * TICIT_HIP_UNSUPPORTED). The program's device must be current. Every call
 * on one GPU shares that array and leaves its contents behind; calls that
 * overlap on one GPU (other host threads, or streams not ordered with each
 * other) race on it. */
ticit_hip_status ticit_hip_sample_device_postselected_noalloc(ticit_hip_program* program, uint64_t shots,
                                                             uint64_t seed, uint64_t first_shot, int bit_packed,
                                                             int keep_records, uint8_t* meas, uint8_t* det,
                                                             uint8_t* obs, uint64_t* observable_ones,
                                                             double* exp_vals, uint64_t* accepted, void* stream);

/* Injected randomness (host buffers): exo = symbol-major bit-planes
 * [nsymbols][ceil(shots/64)] (bit s of word w = shot 64w+s; only exogenous
 * symbols are read), uniforms = [measure_instructions][shots] branch draws
 * (branch = u < P(branch=1)), coins = [coin_instructions][ceil(shots/64)].
 * probs, if non-NULL, receives P(branch = 1) per [measure][shot]. */
ticit_hip_status ticit_hip_sample_injected(ticit_hip_program* program, uint64_t shots, const uint64_t* exo,
                                           const double* uniforms, const uint64_t* coins, int bit_packed,
                                           ticit_hip_result* out, double* probs);

/* Injected sampling that also returns shot 0's final active state as FP64
 * (re/im must hold 2^max_active_qubits entries; *k receives its width).
 * Served by the wide tier; used to report amplitude error against the FP64
 * oracle. */
ticit_hip_status ticit_hip_final_state(ticit_hip_program* program, uint64_t shots, const uint64_t* exo,
                                       const double* uniforms, const uint64_t* coins, double* re, double* im,
                                       int* k);

/* ---- Batched sampling of many programs -------------------------------
 * No single counterpart in the reference, which samples one circuit per
 * call; a batch is `count` of those calls (Sampler::sample_with_seed,
 * src/sampler/prepared.rs:663-670, one per item) run as one hipGraph.
 * Every item's rows equal what ticit_hip_sample_device gives for the same
 * program, shots and seed (seed + the launch's seed_offset).
 *
 * ticit_hip_batch_create captures the graph (after timing launch shapes
 * when options->autotune is set); ticit_hip_batch_launch enqueues one replay
 * on `stream` without synchronizing; ticit_hip_batch_wait synchronizes with
 * the last replay and reports device-side sampling errors. Output buffers
 * are caller-owned device memory, as for ticit_hip_sample_device, and must
 * outlive the batch; observable_ones accumulate across replays. */
typedef struct ticit_hip_batch ticit_hip_batch;

typedef struct ticit_hip_batch_item {
  ticit_hip_program* program;
  uint64_t shots;
  uint64_t seed;
  uint8_t* measurements; /* [shots][row] device buffers */
  uint8_t* detectors;
  uint8_t* observables;
  uint64_t* observable_ones; /* [nobservables] device */
  double* exp_vals;          /* [shots][expectation_values] device, NULL to skip */
  /* device uint64 receiving the survivor count of a postselected program
   * (required then, ignored otherwise): its rows come first in the buffers,
   * in shot order, as from ticit_hip_sample_device_postselected */
  uint64_t* accepted;
} ticit_hip_batch_item;

typedef struct ticit_hip_batch_options {
  int32_t bit_packed;
  int32_t keep_records;
  uint32_t autotune;       /* 1: time the launch-shape candidates of each group bucket */
  uint32_t autotune_shots; /* shots per program in a timing run (0 = 8192) */
  uint32_t group_grid_percent; /* persistent group-tier grid, percent of full occupancy
                                  (0 = 100); below 100 leaves room for wide-tier kernels */
  uint32_t wide_cu_percent;    /* > 0: split the CUs between wide-tier work (this share) and
                                  group/lane work on CU-masked streams that run concurrently */
  uint32_t reserved[2];
} ticit_hip_batch_options;

typedef struct ticit_hip_batch_info {
  uint64_t items, lane_items, wide_items, group_buckets;
  uint32_t bucket_tps[16];          /* threads per shot chosen per bucket */
  uint32_t bucket_items[16];
  uint32_t bucket_region_bytes[16]; /* LDS bytes per shot */
  uint32_t bucket_grid[16];         /* persistent workgroups */
  double bucket_tuned_ms[16][8];    /* timing-run ms for 8, 16, ..., 512 threads per shot, index
                                       log2(threads) - 3 (0 = not run) */
} ticit_hip_batch_info;

ticit_hip_status ticit_hip_batch_create(const ticit_hip_batch_item* items, size_t count,
                                        const ticit_hip_batch_options* options, ticit_hip_batch** out);
ticit_hip_status ticit_hip_batch_launch(ticit_hip_batch* batch, uint64_t seed_offset, void* stream);
ticit_hip_status ticit_hip_batch_wait(ticit_hip_batch* batch);
ticit_hip_status ticit_hip_batch_info_get(const ticit_hip_batch* batch, ticit_hip_batch_info* out);
void ticit_hip_batch_destroy(ticit_hip_batch* batch);

/* One-shot convenience over the batch API with host results (one per item,
 * freed with ticit_hip_result_free). */
ticit_hip_status ticit_hip_sample_batch(ticit_hip_program* const* programs, const uint64_t* shots,
                                        const uint64_t* seeds, size_t count, int bit_packed, int keep_records,
                                        ticit_hip_result* results);

void ticit_hip_result_free(ticit_hip_result* result);

/* Planning features of a plan blob, computed on the host without a device:
 * the quantities of ticit's component cost model (ActiveComponentPlan,
 * src/component_plan.rs:87-104) plus the program's shape. out receives
 * min(nout, TICIT_HIP_NFEATURES) values indexed by TICIT_HIP_FEATURE_*. */
enum {
  TICIT_HIP_FEATURE_QUBITS = 0,
  TICIT_HIP_FEATURE_INITIAL_K,
  TICIT_HIP_FEATURE_MAX_K,
  TICIT_HIP_FEATURE_SYMBOLS,
  TICIT_HIP_FEATURE_RECORDS,
  TICIT_HIP_FEATURE_DETECTORS,
  TICIT_HIP_FEATURE_INSTRUCTIONS,
  TICIT_HIP_FEATURE_MEASUREMENTS,     /* active measurements */
  TICIT_HIP_FEATURE_ROTATIONS,        /* active rotations */
  TICIT_HIP_FEATURE_PROMOTIONS,       /* T-branch splits */
  TICIT_HIP_FEATURE_QUANTUM_OPS,
  TICIT_HIP_FEATURE_EXPRESSION_TERMS,
  TICIT_HIP_FEATURE_SELECTED,         /* use_active_components of the blob */
  TICIT_HIP_FEATURE_PLANNED,          /* a component plan exists (max_k >= 8) */
  TICIT_HIP_FEATURE_DENSE_WORK,       /* estimated_dense_work */
  TICIT_HIP_FEATURE_COMPONENT_WORK,   /* estimated_component_work */
  TICIT_HIP_FEATURE_DENSE_PEAK,       /* dense_peak_dimension */
  TICIT_HIP_FEATURE_PEAK_LIVE,        /* component_peak_live_dimension */
  TICIT_HIP_FEATURE_ALLOCATED,        /* component_allocated_dimension */
  TICIT_HIP_FEATURE_MAX_COMPONENT_K,
  TICIT_HIP_FEATURE_MERGES,
  TICIT_HIP_FEATURE_DENSE_AMP_TRAFFIC,     /* amplitudes streamed per shot, dense layout */
  TICIT_HIP_FEATURE_COMPONENT_AMP_TRAFFIC, /* the same for the component layout (0 unplanned) */
  TICIT_HIP_FEATURE_OUTPUT_BYTES,          /* packed row bytes per shot */
  TICIT_HIP_NFEATURES
};
ticit_hip_status ticit_hip_plan_features(const uint64_t* blob, size_t nwords, double* out, size_t nout);

/* ---- Learned planner -------------------------------------------------
 * Replaces, when planner_mode = 1, the dense-versus-components decision of
 * should_select_component_plan (src/component_plan.rs:282-298) with a
 * 12 -> 64 -> 64 -> 3 FP32 MLP (ReLU): outputs are the logits of "dense"
 * and "components" and the standardized log10 time per shot of the faster
 * layout. Features (TICIT_HIP_PLANNER_FEATURES, raw): qubits, T count,
 * mean log2 active dimension, nonzero fraction (component peak live /
 * dense peak), Clifford depth (circuit layers), measurement records,
 * log2(1 + estimated dense work), log2(1 + estimated component work),
 * log2 dense peak, log2(1 + component peak live), log2(1 + component
 * allocation), quantum-op count. No counterpart in the reference. */
#define TICIT_HIP_PLANNER_FEATURES 12
#define TICIT_HIP_MLP_PARAMS 5187 /* W1[12][64] b1[64] W2[64][64] b2[64] W3[64][3] b3[3] */
typedef struct ticit_hip_mlp ticit_hip_mlp;

ticit_hip_status ticit_hip_planner_features(const uint64_t* blob, size_t nwords, float* out);
ticit_hip_status ticit_hip_mlp_create(const float* params, ticit_hip_mlp** out);
/* Adam training on the device: step s uses samples order[s*batch .. +batch)
 * (batch <= 64); loss = softmax cross-entropy on y (0 dense, 1 components)
 * + time_weight * (output 2 - t)^2, averaged per minibatch into losses[s]. */
ticit_hip_status ticit_hip_mlp_train(ticit_hip_mlp* mlp, const float* x, const int32_t* y, const float* t, size_t n,
                                     const uint32_t* order, size_t steps, uint32_t batch, float lr,
                                     float time_weight, float* losses);
/* Gradients of the mean loss over the n <= 64 samples (one minibatch),
 * without updating the parameters. */
ticit_hip_status ticit_hip_mlp_gradients(ticit_hip_mlp* mlp, const float* x, const int32_t* y, const float* t,
                                         size_t n, float time_weight, float* grads, float* loss);
/* Inference kernel: out[n][3]. x is standardized by the caller. */
ticit_hip_status ticit_hip_mlp_predict(const ticit_hip_mlp* mlp, const float* x, size_t n, float* out);
ticit_hip_status ticit_hip_mlp_params_get(const ticit_hip_mlp* mlp, float* out);
void ticit_hip_mlp_destroy(ticit_hip_mlp* mlp);
/* Installs (params != NULL) or removes the process-wide learned planner used
 * by planner_mode = 1. Features are standardized as (f - mean) / scale; when
 * |P(components) - 1/2| < margin the plan's heuristic choice stands. */
ticit_hip_status ticit_hip_planner_set(const float* params, const float* feature_mean, const float* feature_scale,
                                       float margin);

/* Device properties the tier planner sizes against; returns 0 when no
 * device is visible. props = {lds_per_block, compute_units, clock_khz,
 * warp_size, memory_clock_khz, bus_width_bits}. */
int ticit_hip_device_query(int64_t* props);

#ifdef __cplusplus
}
#endif

#endif /* TICIT_HIP_H */
