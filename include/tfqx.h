/* tfqx: C-ABI of the MI355X (gfx950) circuit-simulation op family.
 *
 * Drop-in for TFQ 0.7.7's native simulation ops.  Each op of the reference is
 * a TensorFlow kernel loaded with load_op_library; each is replaced here by a
 * plan (host: parse, resolve, fuse) followed by an upload and a launch on a
 * caller-supplied HIP stream:
 *
 *   tfqx_plan_simulate_expectation          TfqSimulateExpectation
 *       core/ops/tfq_simulate_expectation_op.cc:45-283 (REGISTER_OP :253-283)
 *   tfqx_plan_simulate_sampled_expectation  TfqSimulateSampledExpectation
 *       core/ops/tfq_simulate_sampled_expectation_op.cc:48-342 (:308-342)
 *   tfqx_plan_simulate_samples              TfqSimulateSamples
 *       core/ops/tfq_simulate_samples_op.cc:48-285 (:254-285)
 *   tfqx_plan_simulate_state                TfqSimulateState
 *       core/ops/tfq_simulate_state_op.cc:43-242 (:218-242)
 *   tfqx_plan_inner_product                 TfqInnerProduct
 *       core/ops/math_ops/tfq_inner_product.cc:45-323 (:294-323)
 *   tfqx_plan_adjoint_gradient              TfqAdjointGradient
 *       core/ops/tfq_adj_grad_op.cc:46-427 (:393-427)
 *
 * Inputs mirror the reference op inputs.  A string tensor of n elements is
 * passed as (bytes, offsets[n+1]): element i is bytes[offsets[i]:offsets[i+1]].
 * Dense inputs are host pointers in row-major order.  Outputs are
 * caller-allocated device buffers with the reference's dtypes and shapes:
 *
 *   expectation / sampled expectation  float32   [batch, num_ops]
 *   samples                            int8      [batch, num_samples, max_nq]
 *   state                              complex64 [batch, 2^max_nq]
 *   inner product                      complex64 [batch, num_other]
 *   adjoint gradient                   float32   [batch, num_symbols]
 *
 * with the reference's sentinels: -2 expectation / 0 gradient / 1+0j inner
 * product for empty programs, -2 padding of samples and states.
 *
 * Errors: every call returns a status (0 = OK) and writes a message into
 * (err, err_len); messages keep the reference's wording (e.g. "Unparseable
 * proto", "Could not find symbol in parameter map", "Found a Pauli sum
 * operating on qubits not found in circuit").
 *
 * Memory: plans own host memory only.  Device memory is a caller-allocated
 * workspace of tfqx_plan_workspace_bytes() bytes (256-byte aligned); the
 * library performs no device allocation on the op path.
 */
#ifndef TFQX_H_
#define TFQX_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  TFQX_OK = 0,
  TFQX_INVALID_ARGUMENT = 3,
  TFQX_UNIMPLEMENTED = 12,
  TFQX_INTERNAL = 13,
  TFQX_DEVICE_ERROR = 14
};

/* Adjoint derivative of a parameterised gate. */
enum {
  TFQX_GRAD_TFQ_FD = 0,   /* (G(v+eps)-G(v-eps))/(2 eps), eps=5e-3, float32: adj_util.cc:175-302 */
  TFQX_GRAD_ANALYTIC = 1  /* exact dG/dv */
};

typedef struct tfqx_plan tfqx_plan;
typedef struct tfqx_graph tfqx_graph;

/* TfqSimulateExpectation(programs[B], symbol_names[S], symbol_values[B,S],
 * pauli_sums[B,O]) -> float[B,O]. */
int tfqx_plan_simulate_expectation(
    const uint8_t* programs, const int64_t* program_offsets, int64_t batch,
    const uint8_t* symbol_names, const int64_t* symbol_offsets, int64_t num_symbols,
    const float* symbol_values,
    const uint8_t* pauli_sums, const int64_t* pauli_offsets, int64_t num_ops,
    tfqx_plan** plan, char* err, int64_t err_len);

/* TfqSimulateSampledExpectation(..., num_samples int32[B,O]) -> float[B,O].
 * The reference seeds from tensorflow::random::New64(); here the seed of the
 * Philox4x32-10 stream is explicit. */
int tfqx_plan_simulate_sampled_expectation(
    const uint8_t* programs, const int64_t* program_offsets, int64_t batch,
    const uint8_t* symbol_names, const int64_t* symbol_offsets, int64_t num_symbols,
    const float* symbol_values,
    const uint8_t* pauli_sums, const int64_t* pauli_offsets, int64_t num_ops,
    const int32_t* num_samples, uint64_t seed,
    tfqx_plan** plan, char* err, int64_t err_len);

/* TfqSimulateSamples(programs[B], symbol_names[S], symbol_values[B,S],
 * num_samples int32[1]) -> int8[B, num_samples, max_nq]. */
int tfqx_plan_simulate_samples(
    const uint8_t* programs, const int64_t* program_offsets, int64_t batch,
    const uint8_t* symbol_names, const int64_t* symbol_offsets, int64_t num_symbols,
    const float* symbol_values, int32_t num_samples, uint64_t seed,
    tfqx_plan** plan, char* err, int64_t err_len);

/* TfqNoisyExpectation(programs[B], symbol_names[S], symbol_values[B,S],
 * pauli_sums[B,O], num_samples int32[B,O]) -> float[B,O].
 * Replaces core/ops/noise/tfq_noisy_expectation.cc:57-283: term j of row i
 * averages its exact expectation over num_samples[i][j] Monte-Carlo
 * trajectories (channels DP ADP GAD AD RST PD PF BF,
 * circuit_parser_qsim.cc:598-870); -2 for empty programs.  Channel c of
 * trajectory t of row i draws Philox4x32-10 uniform (seed, shot t, stream
 * (i, c)); the reference seeds from tensorflow::random::New64().  Circuits of
 * at most 20 qubits (TFQX_UNIMPLEMENTED beyond; programs of 15-20 qubits add
 * one 2^max_qubits x 8-byte state per launched workgroup, min(512, work
 * items), to the workspace). */
int tfqx_plan_noisy_expectation(
    const uint8_t* programs, const int64_t* program_offsets, int64_t batch,
    const uint8_t* symbol_names, const int64_t* symbol_offsets, int64_t num_symbols,
    const float* symbol_values,
    const uint8_t* pauli_sums, const int64_t* pauli_offsets, int64_t num_ops,
    const int32_t* num_samples, uint64_t seed,
    tfqx_plan** plan, char* err, int64_t err_len);

/* TfqNoisySampledExpectation(..., num_samples int32[B,O]) -> float[B,O].
 * Replaces core/ops/noise/tfq_noisy_sampled_expectation.cc: per trajectory,
 * each Pauli term contributes one measured parity (+-1) times its
 * coefficient (ComputeSampledExpectationQsim with one sample,
 * util_qsim.h:199-270).  The parity is drawn as a Bernoulli variate with
 * P(+1) = (1 + <P>/<psi|psi>) / 2, the distribution of one measurement of the
 * rotated state, from stream (i, 0x80000000 | term index). */
int tfqx_plan_noisy_sampled_expectation(
    const uint8_t* programs, const int64_t* program_offsets, int64_t batch,
    const uint8_t* symbol_names, const int64_t* symbol_offsets, int64_t num_symbols,
    const float* symbol_values,
    const uint8_t* pauli_sums, const int64_t* pauli_offsets, int64_t num_ops,
    const int32_t* num_samples, uint64_t seed,
    tfqx_plan** plan, char* err, int64_t err_len);

/* TfqNoisySamples(programs[B], symbol_names[S], symbol_values[B,S],
 * num_samples int32[1]) -> int8[B, num_samples, max_nq].
 * Replaces core/ops/noise/tfq_noisy_samples.cc:48-205: sample t of row i is
 * the terminal measurement of trajectory t (stream (i, 0xFFFFFFFF)); bit q
 * goes to column max_nq-1-q, leading -2 padding; empty programs give -2. */
int tfqx_plan_noisy_samples(
    const uint8_t* programs, const int64_t* program_offsets, int64_t batch,
    const uint8_t* symbol_names, const int64_t* symbol_offsets, int64_t num_symbols,
    const float* symbol_values, int32_t num_samples, uint64_t seed,
    tfqx_plan** plan, char* err, int64_t err_len);

/* TfqSimulateState(programs[B], symbol_names[S], symbol_values[B,S])
 * -> complex64[B, 2^max_nq]. */
int tfqx_plan_simulate_state(
    const uint8_t* programs, const int64_t* program_offsets, int64_t batch,
    const uint8_t* symbol_names, const int64_t* symbol_offsets, int64_t num_symbols,
    const float* symbol_values,
    tfqx_plan** plan, char* err, int64_t err_len);

/* TfqInnerProduct(programs[B], symbol_names[S], symbol_values[B,S],
 * other_programs[B,K]) -> complex64[B,K]. */
int tfqx_plan_inner_product(
    const uint8_t* programs, const int64_t* program_offsets, int64_t batch,
    const uint8_t* symbol_names, const int64_t* symbol_offsets, int64_t num_symbols,
    const float* symbol_values,
    const uint8_t* other_programs, const int64_t* other_offsets, int64_t num_other,
    tfqx_plan** plan, char* err, int64_t err_len);

/* TfqInnerProductGrad(programs[B], symbol_names[S], symbol_values[B,S],
 *                     other_programs[B,K], prev_grad[B,K]) -> complex64 [B,S]
 * Replaces core/ops/math_ops/tfq_inner_product_grad.cc:46-501 (Compute
 * :51-208, adjoint sweep :259-314 and :380-444): out[i,s] = sum over gates using s of
 * <dG psi_k | lambda_k>, lambda = sum_j prev_grad[i,j] |phi_ij>
 * (AccumulateFusedCircuits, util_qsim.h:423-442); no conjugation (the Python
 * wrapper conjugates, inner_product_op.py:65-68).  S must be positive.
 * Circuits of up to 13 qubits run in the batched kernel; wider ones run one
 * after another through streaming passes.  derivative as for
 * tfqx_plan_adjoint_gradient. */
int tfqx_plan_inner_product_grad(
    const uint8_t* programs, const int64_t* program_offsets, int64_t batch,
    const uint8_t* symbol_names, const int64_t* symbol_offsets, int64_t num_symbols,
    const float* symbol_values,
    const uint8_t* other_programs, const int64_t* other_offsets, int64_t num_others,
    const float* prev_grad, int32_t derivative,
    tfqx_plan** plan, char* err, int64_t err_len);

/* TfqAdjointGradient(programs[B], symbol_names[S], symbol_values[B,S],
 * pauli_sums[B,O], downstream_grads[B,O]) -> float[B,S]. */
int tfqx_plan_adjoint_gradient(
    const uint8_t* programs, const int64_t* program_offsets, int64_t batch,
    const uint8_t* symbol_names, const int64_t* symbol_offsets, int64_t num_symbols,
    const float* symbol_values,
    const uint8_t* pauli_sums, const int64_t* pauli_offsets, int64_t num_ops,
    const float* downstream_grads, int32_t derivative,
    tfqx_plan** plan, char* err, int64_t err_len);

/* Expectations and adjoint gradients of the same inputs from one launch
 * (no counterpart op in the reference: TFQ evaluates them with two ops,
 * TfqSimulateExpectation and TfqAdjointGradient).  Output: float32
 * expectations [batch, num_ops] followed by gradients [batch, num_symbols]. */
int tfqx_plan_expectation_and_gradient(
    const uint8_t* programs, const int64_t* program_offsets, int64_t batch,
    const uint8_t* symbol_names, const int64_t* symbol_offsets, int64_t num_symbols,
    const float* symbol_values,
    const uint8_t* pauli_sums, const int64_t* pauli_offsets, int64_t num_ops,
    const float* downstream_grads, int32_t derivative,
    tfqx_plan** plan, char* err, int64_t err_len);

/* Re-evaluate a plan for new symbol values [batch, num_symbols] without
 * re-parsing, re-resolving or re-fusing: only the matrices of ops that contain
 * symbolic gates are rebuilt.  Call tfqx_plan_upload again before launching.
 * Works for every plan of the batch kernel (LDS-resident, wave-mode and tiled
 * circuits, one program tiled over the batch or a different program per row);
 * streaming and noisy plans return TFQX_INVALID_ARGUMENT.  No reference
 * counterpart: the reference re-parses every call
 * (core/ops/parse_context.cc:60-92). */
int tfqx_plan_rebind(tfqx_plan* plan, const float* symbol_values, char* err, int64_t err_len);

/* New downstream gradients [batch, num_ops] for an adjoint plan of the batch
 * kernel (tfqx_plan_adjoint_gradient / tfqx_plan_expectation_and_gradient):
 * rewrites the lambda weights g_ij * c_t, zero where |g_ij * c_t| < 1e-5 as
 * AccumulateOperators drops them (core/src/util_qsim.h:379-383).  Call
 * tfqx_plan_upload again before launching.  No reference counterpart. */
int tfqx_plan_rebind_downstream(tfqx_plan* plan, const float* downstream_grads, char* err,
                                int64_t err_len);

/* New symbol values [batch, num_symbols] evaluated on the device: the values
 * are kept with the plan, and the next tfqx_plan_upload or
 * tfqx_plan_upload_changed copies them (B x S floats) and enqueues one kernel
 * that rebuilds the matrices of every op containing symbolic gates in the
 * workspace, ahead of tfqx_plan_launch on the same stream.  The host matrices
 * are not recomputed, so the upload carries the values instead of the
 * matrices.  Results equal tfqx_plan_rebind's to float rounding (the same
 * double-precision recipes; only the device's sin/cos may differ in the last
 * bit).  The first call builds the device tables (gates and recipes of the
 * symbolic ops), which the next upload copies once.  Same plans as
 * tfqx_plan_rebind (streaming and noisy plans return TFQX_INVALID_ARGUMENT).
 * No reference counterpart: the reference re-parses and re-evaluates every
 * gate on the host every call (core/ops/parse_context.cc:60-92). */
int tfqx_plan_stage_values(tfqx_plan* plan, const float* symbol_values, char* err,
                           int64_t err_len);

/* Fusion decisions of the planner for one program (host only): the fused ops
 * of the forward program at the given width (1..4).  ops_out[i][8] = {k,
 * bit0..bit3 (global bit positions, ascending, -1 unused), 1 if the op keeps
 * controls outside its matrix, first, end} with [first, end) indexing gates_out,
 * the op's gates (circuit order of the program's operations) in application
 * order.  *n_ops / *n_gates receive the sizes; nothing is written when they
 * exceed the capacities.  No reference counterpart: exposes qsim-fuser-like
 * decisions (qsim BasicGateFuser, called at core/src/circuit_parser_qsim.cc:
 * 857-859) so a host replica can check them. */
int tfqx_debug_fuse(const uint8_t* program, int64_t program_len, const uint8_t* symbol_names,
                    const int64_t* symbol_offsets, int64_t num_symbols, const float* symbol_values,
                    int32_t width, int32_t* ops_out, int64_t ops_cap, int32_t* gates_out,
                    int64_t gates_cap, int64_t* n_ops, int64_t* n_gates, char* err,
                    int64_t err_len);

/* The sampler's FP64 uniforms for shots [start, start + count) of stream
 * (stream_lo, stream_hi) under `seed`, written to the device buffer `out`
 * (count doubles) on `stream`: exactly the draws the samples and sampled-
 * expectation kernels make (Philox4x32-10, counter = (shot, stream), key =
 * seed, u = top 53 bits / 2^53).  No reference counterpart: TFQ seeds
 * tensorflow::random::PhiloxRandom nondeterministically
 * (core/ops/tfq_simulate_samples_op.cc:149-152); this exposes the fixed
 * stream so a host replica can check it bit for bit. */
int tfqx_debug_philox_uniform(uint64_t seed, uint64_t start, int64_t count, uint32_t stream_lo,
                              uint32_t stream_hi, double* out, void* stream, char* err,
                              int64_t err_len);

/* Parameter-shift program rewriting (host only).  String results come back
 * in a tfqx_strings handle: count, bytes + offsets[count+1], and the result
 * shape; free it with tfqx_strings_destroy. */
typedef struct tfqx_strings tfqx_strings;
int64_t tfqx_strings_count(const tfqx_strings* s);
const uint8_t* tfqx_strings_data(const tfqx_strings* s);
const int64_t* tfqx_strings_offsets(const tfqx_strings* s);
int64_t tfqx_strings_dim(const tfqx_strings* s, int64_t i);
void tfqx_strings_destroy(tfqx_strings* s);

/* TfqPsDecompose(programs[B]) -> programs[B].  Replaces
 * core/ops/tfq_ps_decompose_op.cc:39-160: symbolic PhasedISwap, ISwap,
 * PhasedX and FSim gates become Z / XX / YY / X / CZ eigen gates with one
 * symbol each, in extra moments after their own. */
int tfqx_ps_decompose(const uint8_t* programs, const int64_t* program_offsets, int64_t batch,
                      tfqx_strings** out, char* err, int64_t err_len);

/* TfqPsSymbolReplace(programs[B], symbols[S], replacement_symbols[S])
 * -> programs[B, S, pad].  Replaces core/ops/tfq_ps_symbol_replace_op.cc:
 * 39-185: entry (i, j, k) is programs[i] with the k-th occurrence of
 * symbols[j] (moment, op, argument order) renamed to replacement_symbols[j];
 * padded with the empty program. */
int tfqx_ps_symbol_replace(const uint8_t* programs, const int64_t* program_offsets, int64_t batch,
                           const uint8_t* symbols, const int64_t* symbol_offsets, int64_t num_symbols,
                           const uint8_t* replacements, const int64_t* replacement_offsets,
                           int64_t num_replacements, tfqx_strings** out, char* err,
                           int64_t err_len);

/* TfqPsWeightsFromSymbols(programs[B], symbols[S]) -> float[B, S, pad].
 * Replaces core/ops/tfq_ps_weights_from_symbols_op.cc:39-170: the
 * exponent_scalar of every gate whose exponent is symbols[j], zero padded.
 * Call with out = NULL to get *pad, then with capacity >= B * S * pad. */
int tfqx_ps_weights_from_symbols(const uint8_t* programs, const int64_t* program_offsets,
                                 int64_t batch, const uint8_t* symbols,
                                 const int64_t* symbol_offsets, int64_t num_symbols, float* out,
                                 int64_t capacity, int64_t* pad, char* err, int64_t err_len);

/* Plan statistics: out[0] circuits, [1] forward ops, [2] forward HBM passes,
 * [3] max qubits, [4] adjoint ops, [5] adjoint HBM passes, [6] 1 if the
 * HBM-streaming path is used, [7] workspace bytes, [8] algorithmic real flops
 * of one launch (batch path; 0 on the streaming path), [9] the part of those
 * flops that runs on FP32 MFMA (both paths).  Batch-kernel plans, n > 10:
 * the op mix in amplitude-op units (2^n per op of an n-qubit circuit, /1024):
 * [10] forward ops, [11] reverse remaps, [12] reverse diagonal ops, [13]
 * reverse dense one-qubit ops, [14] reverse dense wider ops, [15] reverse
 * parameterised one-qubit ops, [16] reverse parameterised two-qubit ops,
 * [17] reverse parameterised ops with a diagonal matrix, [18] forward
 * remaps, [19] forward diagonal ops, [20] forward dense one-qubit ops, [21]
 * forward dense wider ops, [22] register-engine remaps that keep every
 * amplitude in its wavefront (run without workgroup barriers), [23] all
 * register-engine remaps of workgroup-run circuits.  Noisy plans: [0] non-empty
 * circuits, [1] ops (gates and channels), [2] trajectories per launch, [8]
 * expected algorithmic real flops of one launch.  No reference counterpart. */
int tfqx_plan_stats(const tfqx_plan* plan, int64_t* out, int64_t n);

/* Largest qubit count in the batch (sizes the samples / state outputs). */
int64_t tfqx_plan_max_num_qubits(const tfqx_plan* plan);
/* Device workspace the plan needs (bytes). */
int64_t tfqx_plan_workspace_bytes(const tfqx_plan* plan);
/* Bytes of output the plan writes. */
int64_t tfqx_plan_output_bytes(const tfqx_plan* plan);

/* Copy the device program into the workspace (host -> device, on `stream`). */
int tfqx_plan_upload(tfqx_plan* plan, void* workspace, void* stream, char* err, int64_t err_len);
/* Bring a workspace that holds this plan's last upload up to date, copying
 * only what changed since: matrices after tfqx_plan_rebind, Pauli weights
 * after tfqx_plan_rebind_downstream, staged values after
 * tfqx_plan_stage_values (then evaluated on the device).  When `workspace` is
 * not the one the plan was last uploaded to, this is tfqx_plan_upload.  The
 * caller keeps the workspace to this plan between the two uploads: another
 * plan's upload into it, or any other write, requires tfqx_plan_upload. */
int tfqx_plan_upload_changed(tfqx_plan* plan, void* workspace, void* stream, char* err,
                             int64_t err_len);
/* Run the plan's kernels on an uploaded workspace; writes `output`. */
int tfqx_plan_launch(tfqx_plan* plan, void* workspace, void* output, void* stream,
                     char* err, int64_t err_len);
/* Same results, one kernel launch per circuit (the un-batched GPU path). */
int tfqx_plan_launch_plain(tfqx_plan* plan, void* workspace, void* output, void* stream,
                           char* err, int64_t err_len);
/* tfqx_plan_upload + tfqx_plan_launch with the workspace in the library's own
 * static device array (32 MiB per GPU, held until the process ends) instead
 * of a caller buffer: allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Plans
 * whose workspace exceeds 32 MiB (not counting the region of
 * tfqx_plan_stage_values until values are staged), or that take the
 * HBM-streaming path, are rejected with TFQX_INVALID_ARGUMENT.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (different streams or unordered
 * host threads) race. */
int tfqx_plan_launch_noalloc(tfqx_plan* plan, void* output, void* stream, char* err,
                             int64_t err_len);
/* Capture tfqx_plan_launch into a hipGraph and replay it. */
int tfqx_graph_capture(tfqx_plan* plan, void* workspace, void* output, void* stream,
                       tfqx_graph** graph, char* err, int64_t err_len);
int tfqx_graph_launch(tfqx_graph* graph, void* stream, char* err, int64_t err_len);
void tfqx_graph_destroy(tfqx_graph* graph);
void tfqx_plan_destroy(tfqx_plan* plan);

/* ---------------------------------------------------------------- surrogate
 * MLP surrogate E(x) with input gradient dE/dx (in_dim -> hidden x3 -> 1,
 * SiLU), trained with the Sobolev loss mean (E-E*)^2 + lam mean ||dE/dx-g*||^2
 * on BF16 MFMA GEMMs (FP32 accumulate) and Adam.  No reference counterpart:
 * TFQ has no learned component; this serves VQE loops that call
 * TfqAdjointGradient repeatedly.  All pointers except host_params are device
 * pointers; the workspace holds parameters, Adam state and activations. */
/* out: [workspace bytes, #params, param offset, grad offset, loss offset,
 * Adam m offset] (byte offsets into the workspace). */
int tfqx_mlp_layout(int32_t in_dim, int32_t hidden, int32_t max_batch, int64_t* out, int64_t n);
int tfqx_mlp_set_params(void* workspace, int32_t in_dim, int32_t hidden, int32_t max_batch,
                        const float* host_params, void* stream, char* err, int64_t err_len);
/* energy[batch], grad[batch, in_dim] (grad may be null). */
int tfqx_mlp_predict(void* workspace, int32_t in_dim, int32_t hidden, int32_t max_batch,
                     const float* x, int32_t batch, float* energy, float* grad, void* stream,
                     char* err, int64_t err_len);
/* Loss (2 floats at the loss offset) and parameter gradient (grad offset). */
int tfqx_mlp_loss_grad(void* workspace, int32_t in_dim, int32_t hidden, int32_t max_batch,
                       const float* x, const float* energy, const float* grad, int32_t batch,
                       float lam, void* stream, char* err, int64_t err_len);
/* loss_grad followed by one Adam step (step >= 1). */
int tfqx_mlp_train_step(void* workspace, int32_t in_dim, int32_t hidden, int32_t max_batch,
                        const float* x, const float* energy, const float* grad, int32_t batch,
                        float lam, float lr, int32_t step, void* stream, char* err,
                        int64_t err_len);
/* C = alpha op(A) op(B) + beta C; A, B bf16 row-major, C fp32 (v_mfma_f32_32x32x16_bf16). */
int tfqx_gemm_bf16(int32_t M, int32_t N, int32_t K, const void* A, int32_t lda, int32_t transA,
                   const void* B, int32_t ldb, int32_t transB, float* C, int32_t ldc, float alpha,
                   float beta, void* stream, char* err, int64_t err_len);

/* Planner tuning for later plans: register engine off (0), on for the
 * batched kernel's tiles of 12-13 qubits (1), or on for those and the
 * streaming path's passes (2, the default); and the streaming path's tile
 * qubits for one-state and two-state passes (0 = default 13 for both; at most
 * 14 and 13).  No reference counterpart (qsim's CPU path has no tiles). */
void tfqx_set_tuning(int32_t register_engine, int32_t tile_single, int32_t tile_two);

/* Smallest number of 4-10 qubit circuits in a batch for which the planner
 * runs them one per wavefront (expectation and adjoint-gradient plans);
 * smaller batches run every circuit on a whole workgroup.  Negative restores
 * the default, 1024 (four wavefronts per CU); 0 always uses wavefronts.
 * No reference counterpart. */
void tfqx_set_wave_min_circuits(int32_t count);

/* Width from which expectation, adjoint-gradient and state plans run on the
 * HBM-streaming path (every pass spread over 2^(n-13) workgroups) rather than
 * inside the batched kernel (one workgroup per circuit, tiles of 14-18 qubits
 * through per-workgroup scratch).  Values below 14 or above 19, and negative
 * values, restore the default 19.  The analogue of the reference's switch to
 * its state-parallel schedule (tfq_simulate_expectation_op.cc:120-126,
 * tfq_adj_grad_op.cc:161-169), by width only. */
void tfqx_set_streaming_min_qubits(int32_t num_qubits);

/* Adjoint plans: fuse the reverse sweep across parameterised gates (1, the
 * default) or split it at every parameterised gate (0).  With fusion on, a
 * parameterised gate without controls joins its neighbours in a block of at
 * most two qubits and two derivatives; the block's derivative matrices are
 * its product with one gate replaced by that gate's derivative, so the
 * gradient 2 Re<lambda|dB|psi> of each symbol is unchanged while the sweep
 * applies about half as many ops.  Off reproduces the reference's gradient
 * circuit (core/src/adj_util.cc:155-172: fused segments between
 * parameterised gates). */
void tfqx_set_reverse_fusion(int32_t on);

/* Fusion width (qubits per fused block) the planner uses; 0 = default (2).
 * The batched kernel fuses at most 4 qubits.  On the HBM-streaming path, 5 or
 * 6 fuses blocks of that many qubits in the forward passes and applies each
 * such block as a dense (2^k x 2^k) . (2^k x 2^(n-k)) complex contraction on
 * FP32 MFMA (one pass per block); the gates between them run in ordinary tile
 * passes.  No reference counterpart: qsim's basic fuser is fixed at 2 qubits
 * (core/src/circuit_parser_qsim.cc:857-859). */
void tfqx_set_fusion_width(int32_t width);
const char* tfqx_version(void);

#ifdef __cplusplus
}
#endif

#endif /* TFQX_H_ */
