/*
 * ticit plan blob: the serialized program that crosses the sampling-backend
 * boundary.
 *
 * It carries exactly what ticit's FactoredInstructionProgram holds for
 * sampling (src/factored.rs:866-886): the six-opcode instruction stream
 * (src/factored.rs:284-292) with its precomputed kernel descriptors
 * (src/active.rs:88-377), the exogenous sampling plan
 * (src/factored.rs:878-882, built by src/planner.rs:1455-1480) and the
 * observable record lists (src/sampler/prepared.rs:241-266).
 *
 * Layout: a stream of little-endian uint64 words.
 *
 *   MAGIC, VERSION, n, initial_k, max_k, nsymbols, nrecords, ndetectors,
 *   nexpvals, flags (version 2 only), ninstr,
 *   ninstr x instruction:
 *     w0 = op (bits 0-7) | diagonal (8) | diag_phase_bit (9) |
 *          expr.constant (10) | postselect (11, detectors) | nq (16-31) |
 *          pivot (32-47) | phase (48-49)
 *     xmask, zmask, f64 angle,
 *     branch (lo32) | record (hi32),
 *     record_condition (lo32) | detector (hi32),
 *     exp_val (lo32, -1 = none),
 *     ncond, ncond x condition id,
 *     ndet_records, ndet_records x record id
 *   ncategorical x { nbits, nconds, conds..., nrows, nrows x (assignment, f64 p) }
 *   nrare x { f64 event_p, nbits, nsets, nsets x (nconds, conds...),
 *             nrows, nrows x (assignment, f64 p), nevents, nevents x (row, f64 p) }
 *   nbernoulli x (condition, f64 p)
 *   nlowp x { f64 p, nconds, conds... }
 *   nobservables x { nrecords, records... }
 *   MAGIC
 *
 * flags bit 0: use_active_components (src/factored.rs:885), set by the
 * planner when its component cost model selects the tensor-component
 * execution (src/planner.rs:1482-1484); bits 32-63: TICK-separated layers
 * of the source circuit (0 = unknown), a planning feature only. Version 1
 * blobs have no flags word and run dense.
 *
 * op codes: 0 rotate, 1 promote, 2 record, 3 detector, 4 measure, 5 coin.
 * Condition, record and detector ids are 1-based; 0 means "none".
 */
#ifndef TICIT_PLAN_H
#define TICIT_PLAN_H

#define TICIT_PLAN_MAGIC 0x4e4c505449434954ull /* "TICITPLN" */
#define TICIT_PLAN_VERSION 2ull
#define TICIT_PLAN_FLAG_COMPONENTS 1ull

#define TICIT_OP_ROTATE 0
#define TICIT_OP_PROMOTE 1
#define TICIT_OP_RECORD 2
#define TICIT_OP_DETECTOR 3
#define TICIT_OP_MEASURE 4
#define TICIT_OP_COIN 5

#endif /* TICIT_PLAN_H */
