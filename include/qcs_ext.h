/*
 * qcs_ext.h -- extension entry points of the MI355X build.
 *
 * None of these functions has a counterpart in the reference header (include/qcs.h of
 * RQS-SVG); they are additive, change nothing declared in qcs.h, and follow the same
 * convention: return 1 on success, 0 on failure with the message sent to the callback set
 * by qcs_set_exception_callback.  Both libqcs.so (HIP) and the CPU oracle library export
 * them.
 *
 * Definitions used below:
 *   - "logical order": amplitude index bit q is logical qubit q (the save_statevector
 *     order of the reference, src/qcs.cu:2515-2536); amplitudes are interleaved
 *     (re, im) doubles.
 *   - "physical order": the order in device memory; physical bit q holds logical qubit
 *     perm_p2l[q].
 *   - p(i) = fma(re_i, re_i, im_i * im_i), computed in FP64.
 *   - canonical sum of a list of 2^m values: split into tiles of 4096 (the last one padded
 *     with zeros), sum each tile by the stride-halving tree
 *         for (s = 2048; s >= 1; s /= 2) for (i < s) x[i] = x[i] + x[i + s];
 *     and apply the same rule to the list of tile sums until one value is left.
 *   - probability prefix: the 64-ary right-nested prefix of p over physical order:
 *     blocks of 64 consecutive entries are folded left to right, the list of block totals
 *     is treated the same way, and the inclusive prefix at index i is
 *         E_L (+) (E_{L-1} (+) ( ... (+) (E_1 (+) p(i))))
 *     where E_k is the left-to-right sum of the totals that precede i's level-k block
 *     inside its level-(k+1) block.  It is non-decreasing, and its last entry is the
 *     "probability total".
 *   - uniforms: u_j = (splitmix64(seed + (j + 1) * 0x9E3779B97F4A7C15) >> 11) * 2^-53,
 *     where splitmix64 is the output mix of Steele, Lea & Flood (OOPSLA 2014).
 *   - sample j: r_j = u_j * total; the outcome is the smallest physical index i whose
 *     prefix is >= r_j (the tie rule of the reference's `p0 < r`,
 *     src/qcs.cu:2615-2617), reported as a logical index.
 */
#pragma once
#include "qcs.h"

#ifdef __cplusplus
extern "C"
{
#endif

/* Kernel classes reported by qcs_ext_profile_get. */
enum qcs_ext_kernel_class
{
    QCS_EXT_K_PAIR = 0,     /* 2x2 on an amplitude pair: dense 1-qubit, swap-like 2-qubit */
    QCS_EXT_K_PHASE = 1,    /* one phase on a controlled subspace: z/s/t/p/cz/cp/global phase */
    QCS_EXT_K_DIAG = 2,     /* diagonal on 1-2 targets: rz, rzz */
    QCS_EXT_K_DENSE4 = 3,   /* dense 4x4: rxx, ryy, rzx, ecr */
    QCS_EXT_K_MONO = 4,     /* one nonzero per row on up to 4 targets: dcx, rccx, rcccx */
    QCS_EXT_K_REDUCE = 5,   /* probability reductions (measure, norm, prefix) */
    QCS_EXT_K_COLLAPSE = 6, /* measurement collapse (+ renormalisation) */
    QCS_EXT_K_SAMPLE = 7,   /* sampling search */
    QCS_EXT_K_PERMUTE = 8,  /* qubit permutation / swap-in */
    QCS_EXT_K_INIT = 9,     /* state preparation */
    QCS_EXT_K_FUSED = 10,   /* fused k-qubit block (precision form) */
    QCS_EXT_K_CAST = 11,    /* precision conversion */
    QCS_EXT_K_COUNT = 12
};

/* Seed of the measurement engine (std::mt19937_64), which otherwise comes from
 * std::random_device as in the reference (src/qcs.cu:1845-1852). */
int qcs_ext_set_seed(qcs_simulator *sim, uint64_t seed);

/* Copy the whole state out / in, logical order, 2 * num_amplitudes doubles.  Setting the
 * state clears the measured lists, as loading does. */
int qcs_ext_get_statevector(qcs_simulator *sim, double *out, uint64_t num_amplitudes);
int qcs_ext_set_statevector(qcs_simulator *sim, const double *in, uint64_t num_amplitudes);

/* Load a state saved by qcs_simulator_save_statevector (raw complex<double>, logical order,
 * exactly 2^num_qubits amplitudes) into the allocated state, under any qubit mapping, and
 * clear the measured lists. Replaces the reference's unreachable loader
 * (src/qcs.cu:2121-2159, fixed file name statevector_input.bin, physical order). */
int qcs_ext_load_statevector(qcs_simulator *sim, const char *path);

/* Seeded sampling of num_shots full-register outcomes (logical indices). */
int qcs_ext_sample(qcs_simulator *sim, uint64_t seed, uint64_t num_shots, uint64_t *outcomes);

/* qcs_ext_sample without any allocation, for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The prefix levels and the outcomes live in one
 * file-scope device array of libqcs.so (130.5 MiB, held by every GPU the library runs on
 * until the process ends), so the limits are num_qubits <= 30 and num_shots <= 65536;
 * beyond them the call fails.  Gates queued by a fused precision mode must be flushed
 * first (qcs_ext_flush), since a flush may allocate.  Outcomes equal qcs_ext_sample's.
 * The array keeps the last call's contents; calls that overlap on one GPU (from several
 * threads) race on it.  The oracle library checks the same limits and samples as
 * qcs_ext_sample. */
int qcs_ext_sample_noalloc(qcs_simulator *sim, uint64_t seed, uint64_t num_shots, uint64_t *outcomes);

/* The probability total of the prefix defined above, and the canonical sum of p(i). */
int qcs_ext_probability_total(qcs_simulator *sim, double *total);
int qcs_ext_norm(qcs_simulator *sim, double *norm);

/* <Z_q> = sum_i p(i) * (1 - 2 * bit_q(i)) for logical qubits q < count, FP64. */
int qcs_ext_expect_z(qcs_simulator *sim, double *out, int count);

/* Renormalise the kept branch after a measurement collapse (default off: the reference
 * leaves the collapsed state unnormalised, src/qcs.cu:1106-1123). */
int qcs_ext_set_renormalize(qcs_simulator *sim, int enable);

/* Treat physical qubits >= num_local_qubits as "global": a gate or measurement that
 * targets one first swaps it into local slot num_local_qubits-1-i with the in-HBM
 * permutation kernel and updates the mapping as the reference does
 * (src/qcs.cu:2206-2332).  Default: all qubits local. */
int qcs_ext_set_local_qubits(qcs_simulator *sim, bit_num_t num_local_qubits);

/* Current physical -> logical mapping. */
int qcs_ext_get_mapping(qcs_simulator *sim, bit_num_t *perm_p2l, bit_num_t count);

/* Move the data so that the physical layout becomes new_perm_p2l; the logical state is
 * unchanged. */
int qcs_ext_permute(qcs_simulator *sim, const bit_num_t *new_perm_p2l, bit_num_t count);

/* Precision form.  mode 0: complex128, one kernel per gate (default).  mode 1: complex64
 * state, gates queued and fused into blocks of at most max_block_qubits (1..5) qubits
 * applied as dense products on FP32 MFMA (FP32 accumulation); FP64 only in reductions and
 * renormalisation.  mode 2: complex128 state, gates fused into blocks of at most 4 qubits
 * applied in FP64.  mode 3: as mode 1 with the block products on BF16x3 split-precision
 * MFMA (matrix and state each split into two bfloat16 parts, three bf16 products summed in
 * FP32).  Must be called before allocate_memory. */
int qcs_ext_set_precision(qcs_simulator *sim, int mode, int max_block_qubits);

/* Fused precision modes: blocks of a flush are grouped into stages, each applied to
 * LDS-resident tiles of 2^tile_bits amplitudes in one sweep of the state (tile_bits 11, 12,
 * 13 or 14, default 13; complex128 stages default to 2^12 tiles and 3-qubit FP64 blocks; used when
 * the state has at least tile_bits + 4 qubits), or launched one sweep per block (tile_bits
 * 0).  Any time; queued gates are flushed first. */
int qcs_ext_set_stage_bits(qcs_simulator *sim, int tile_bits);

/* Fused complex64 modes: stages on 2^13-amplitude tiles of a state of 22 to 33 qubits run
 * in the persistent streaming kernel (on = 1, the default: one workgroup per CU, the next
 * tile and the next block matrix loaded into LDS while the blocks run) or with one
 * workgroup per tile (on = 0).  Both give bit-identical states.  Any time; queued gates are
 * flushed first. */
int qcs_ext_set_stage_stream(qcs_simulator *sim, int on);

/* Fused precision modes: a fixed plan for every flush in place of the default one (stages
 * on the configured tile size with the default block size and tile choice) when no
 * planner model is attached.  tile_bits > 0: stages on tiles of 2^tile_bits amplitudes
 * (11..13 for complex128, 11..14 for complex64, state of at least tile_bits + 4 qubits) with
 * blocks of block_qubits qubits (complex128 3 or 4, complex64 4 or 5, at most the block
 * limit) and tiles 0 (grown from qubits 0..3 in gate order), 1 (the best window of
 * consecutive qubits) or 2 (whichever gives fewer stages, the default).  tile_bits 0: one
 * sweep per block, blocks of at most block_qubits qubits chosen among the next `window`
 * blocks by rule 0 (least growth), 1 (fullest block) or 2 (earliest block that fits).
 * enable 0 restores the default.  Queued gates are flushed first. */
int qcs_ext_set_plan(qcs_simulator *sim, int enable, int tile_bits, int block_qubits, int tiles, int window, int rule);

/* Apply all queued gates now (no-op in mode 0). */
int qcs_ext_flush(qcs_simulator *sim);

/* Gates applied through fused blocks and number of block launches so far (mode 1). */
int qcs_ext_fusion_stats(qcs_simulator *sim, uint64_t *gates, uint64_t *blocks);

/* <a|b> of two simulators with the same size and mapping on the same device, FP64:
 * out[0] = real part, out[1] = imaginary part, out[2] = max_i |a_i - b_i|.  Either state
 * may be complex128 or complex64.  The fidelity is |<a|b>|^2 / (<a|a> <b|b>). */
int qcs_ext_overlap(qcs_simulator *a, qcs_simulator *b, double *out);

/* Batching form: many independent circuits of 1-target gates (each with at most one
 * positive control), every circuit starting from |0...0>, outputs <Z_q> for all qubits of
 * every circuit (FP64 reductions).  libqcs.so runs the whole batch as one hipGraph launch:
 * circuits of up to lds_max_qubits qubits in a persistent LDS-resident kernel, larger ones
 * as fused blocks on complex64 states in HBM.  The oracle runs the plain method per circuit. */
typedef struct qcs_ext_batch qcs_ext_batch;
int qcs_ext_batch_create(qcs_ext_batch **out);
int qcs_ext_batch_destroy(qcs_ext_batch *batch);
/* gate_names[i] is a qcs.h gate name with one target (e.g. "ry", "rz", "x", "u3");
 * controls[i] is a control qubit or -1; params holds 4 doubles per gate (unused ones 0). */
int qcs_ext_batch_add_circuit(qcs_ext_batch *batch, bit_num_t num_qubits, int32_t num_gates, const char *const *gate_names, const bit_num_t *targets, const bit_num_t *controls, const double *params);
/* Tuning: circuits of up to lds_max_qubits (0..14) qubits run in the LDS-resident kernel with
 * block_size (256, 512, 1024) threads per workgroup; larger ones as fused blocks of
 * block_qubits (4, 5) qubits, spread over `streams` (1..4) graph branches, grouped into stages
 * on LDS tiles of 2^stage_bits amplitudes when they have at least stage_bits + 4 qubits
 * (stage_bits 11, 12 or 13; 0: one launch per block). */
int qcs_ext_batch_set_tuning(qcs_ext_batch *batch, int lds_max_qubits, int block_size, int block_qubits, int streams, int stage_bits);
/* Block products of the circuits in HBM: mode 1 = FP32 MFMA (the default), 3 = BF16x3
 * split-precision MFMA (the modes of qcs_ext_set_precision; states complex64 and <Z> FP64
 * either way).  The oracle computes in FP64 and only checks the mode. */
int qcs_ext_batch_set_precision(qcs_ext_batch *batch, int mode);
/* Searches the tuning parameters on this batch; writes the chosen 5 values to tuning. */
int qcs_ext_batch_autotune(qcs_ext_batch *batch, int *tuning);
/* expz receives sum(num_qubits) doubles, circuit by circuit; seconds = device time. */
int qcs_ext_batch_run(qcs_ext_batch *batch, double *expz, uint64_t count, double *seconds);
int qcs_ext_batch_info(qcs_ext_batch *batch, int *num_circuits, int *total_qubits, uint64_t *graph_nodes);
/* Algorithmic HBM bytes of one run and its kernel launches (libqcs.so; prepares the batch):
 * every fused-block or stage sweep of a complex64 state in HBM reads and writes it (16 B
 * per amplitude) except a circuit's first stage round, which writes it from |0...0>
 * without reading (8 B), every basis-state fill writes it (8 B), every <Z> reduction reads
 * it (8 B); circuits in the LDS-resident kernel move nothing through HBM.  The oracle writes 0
 * for both. */
int qcs_ext_batch_traffic(qcs_ext_batch *batch, double *hbm_bytes, uint64_t *kernel_launches);
/* The same batch by the plain method, circuit after circuit (complex128, one kernel per
 * gate, one simulator per distinct qubit count, in C++): the reference's way of running
 * such a workload, as a baseline.  Same outputs as qcs_ext_batch_run (up to the complex64
 * rounding of the batched form); seconds = host wall time of the whole loop. */
int qcs_ext_batch_run_plain(qcs_ext_batch *batch, double *expz, uint64_t count, double *seconds);
/* The launches of qcs_ext_batch_run issued directly on the streams instead of as one graph
 * launch (for profilers that do not follow graph launches); same outputs. */
int qcs_ext_batch_run_direct(qcs_ext_batch *batch, double *expz, uint64_t count, double *seconds);

/* Learned fusion planner.  A cost model is a 3-layer MLP, 64 units wide (F -> 64 -> 64 ->
 * 1, ReLU, FP32) with parameters W1[64][F], b1[64], W2[64][64], b2[64], w3[64], b3.
 * libqcs.so trains and evaluates it with its own GPU kernels; the oracle library computes
 * the same functions on the CPU.  With a model attached (precision mode 1 or 2), every
 * flush predicts the time of candidate fusion plans (features per block, see
 * qcs_ext_planner_record) and takes the cheapest only if its predicted time beats the
 * default plan's by the model's own typical error: a factor 2^-rms, rms = the root mean
 * squared error (log2 microseconds) of the model's last qcs_ext_mlp_train on its data,
 * within [0.75, 0.97] (0.97 for parameters set by qcs_ext_mlp_set_params).  Without a model
 * the default plan is used.  With the measured check (qcs_ext_planner_set_check, on by
 * default) the default plan also checks each override by measurement: a flush structure
 * keeps the model's choice only if it ran faster than the default plan.  The caller keeps
 * the model alive while it is attached. */
typedef struct qcs_ext_mlp qcs_ext_mlp;
int qcs_ext_mlp_create(qcs_ext_mlp **out, int num_features, uint64_t seed);
int qcs_ext_mlp_destroy(qcs_ext_mlp *model);
int qcs_ext_mlp_num_params(qcs_ext_mlp *model, int *count);
int qcs_ext_mlp_get_params(qcs_ext_mlp *model, float *params, int count);
int qcs_ext_mlp_set_params(qcs_ext_mlp *model, const float *params, int count);
int qcs_ext_mlp_predict(qcs_ext_mlp *model, const float *x, int rows, float *y);
/* Gradient of the mean squared error (1/rows) sum (y - t)^2 with respect to every parameter. */
int qcs_ext_mlp_gradients(qcs_ext_mlp *model, const float *x, const float *t, int rows, float *grad, float *loss);
/* Adam (0.9, 0.999, 1e-8) on shuffled mini-batches; loss = final mean squared error. */
int qcs_ext_mlp_train(qcs_ext_mlp *model, const float *x, const float *t, int rows, int epochs, int batch, float lr, uint64_t seed, float *loss);
/* Number of features per block (16). */
int qcs_ext_planner_num_features(int *count);
/* Times `count` random candidate blocks on this simulator's state (precision mode 1 or 2):
 * features (count x 16 floats) and log2 of the kernel time in microseconds. */
int qcs_ext_planner_record(qcs_simulator *sim, int count, uint64_t seed, float *features, float *log2_us);
int qcs_ext_set_planner(qcs_simulator *sim, qcs_ext_mlp *model);
/* out[0..4]: flushes planned, flushes that took the learned choice, predicted greedy
 * microseconds, predicted microseconds of the executed plans, blocks applied. */
int qcs_ext_planner_stats(qcs_simulator *sim, double *out);
/* Times the launches of every stage plan the planner considers for the gates queued now
 * (tile sizes x block sizes x grown / window tiles), applying the gates once per plan, so
 * the state afterwards is not the circuit's result: training rows drawn from the plans the
 * planner is asked to price.  Writes at most `capacity` rows (features as in
 * qcs_ext_planner_record, log2 microseconds) and their number to *count. */
int qcs_ext_planner_record_plans(qcs_simulator *sim, int capacity, float *features, float *log2_us, int *count);
/* The plan the learned planner ran at the last flush: plan[0..4] = stage tile bits (0: one
 * launch per block), block qubits, tile choice, window, rule (as in qcs_ext_set_plan). */
int qcs_ext_planner_last_choice(qcs_simulator *sim, int *plan);
/* No counterpart in the reference.  Measured check of the learned planner's overrides, on
 * (1, the default) or off (0).  On: per flush structure (the gates' names, kinds and qubits;
 * parameters ignored), the first flush for which the model picks a plan other than the
 * default runs that plan with timing events around it, the next flush of the same structure
 * runs the default plan timed, and every later one runs whichever of the two was faster (the
 * third flush waits for the second's events).  Off: every flush runs the model's choice.
 * Attaching another model, changing this setting, the stage tile bits or the stage kernel
 * (qcs_ext_set_stage_bits, qcs_ext_set_stage_stream) forgets the checks; at most 1024
 * structures are checked per simulator (later ones run the model's choice).  Queued gates
 * are flushed first. */
int qcs_ext_planner_set_check(qcs_simulator *sim, int on);

/* Per-kernel-class timing with HIP events around every launch (oracle: host clock). */
int qcs_ext_profile(qcs_simulator *sim, int enable);
int qcs_ext_profile_get(qcs_simulator *sim, int kernel_class, uint64_t *launches, double *seconds, double *bytes);
int qcs_ext_profile_reset(qcs_simulator *sim);

#ifdef __cplusplus
}
#endif
