/*
 * qxhip.h -- C ABI of libqxhip.so, the MI355X (gfx950) executor for QXTools'
 * tensor-network contraction path.
 *
 * It replaces, for one GPU, the arithmetic reached from these reference seams:
 *   - single_amplitude(tnc, plan, bitstring)         src/simulation.jl:86-91
 *     (add_output! + contract_tn! + output[1], lines 88-90)
 *   - the compute graph built by build_compute_graph src/compute_graph/compute_graph.jl:15-98
 *     in its saved .qx DSL form                      docs/src/users_guide.md:37-170
 *   - QXContexts.execute(dsl, input, param, ...)     bin/qxrun.jl:83-87
 *     (per bitstring, per slice assignment: load / view / ncon, then sum)
 *
 * All arguments are plain pointers and sizes.  Complex numbers are interleaved
 * (re, im) doubles, which is the memory layout of Julia's ComplexF64 arrays; tensor
 * data is flattened in Julia's column-major order (leg 1 fastest), as
 * reshape(data, prod(dim)) does in src/compute_graph/tensor_cache.jl:53.
 *
 * Errors never abort: every call returns a qx_status; qx_last_error() gives the
 * message of the last failing call on the calling thread.
 *
 * Threading: a qx_ctx owns one HIP stream; calls on one ctx (and its plans) must
 * be serialised by the caller.  Separate contexts are independent.
 */
#ifndef QXHIP_H
#define QXHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t qx_status;

#define QX_OK 0
#define QX_ERR_ARGUMENT (-1)    /* null pointer, size or range error            */
#define QX_ERR_PARSE (-2)       /* malformed DSL line                           */
#define QX_ERR_LABELS (-3)      /* ncon labels outside the H/K/M/N classes      */
#define QX_ERR_DATA (-4)        /* missing tensor data or wrong element count   */
#define QX_ERR_MEMORY (-5)      /* arena does not fit the device / the cap      */
#define QX_ERR_DEVICE (-6)      /* HIP runtime error                            */
#define QX_ERR_STATE (-7)       /* call made in the wrong plan state            */
#define QX_ERR_UNSUPPORTED (-8) /* valid DSL outside what this build executes   */

/* Arithmetic of a plan.  QX_C128: complex128 storage and arithmetic.  The C64
 * variants store complex64 and accumulate in FP32 (the cross-slice sum in FP64);
 * large GEMM steps run on FP32 matrix cores with 4 real products (4M) or 3 (3M,
 * Gauss).  QX_C64_BF16X3 is 4M except on steps whose operands are plain matrices
 * with M, N, K >= 1024: each operand is split into bf16 hi + lo parts and every real
 * product is hi.hi + hi.lo + lo.hi on bf16 matrix cores, accumulated in FP32. */
enum qx_precision {
  QX_C128 = 0,
  QX_C64_4M = 1,
  QX_C64_3M = 2,
  QX_C64_BF16X3 = 3
};

typedef struct qx_ctx qx_ctx;   /* device + stream */
typedef struct qx_plan qx_plan; /* lowered, device-resident plan */

/* Create a context on HIP device `device`.  No reference counterpart: the
 * reference selects the GPU inside QXContexts (bin/qxrun.jl:47-49, --gpu). */
qx_status qx_ctx_create(int device, qx_ctx **out);
void qx_ctx_destroy(qx_ctx *ctx);

/* Measured matrix-core rate of the context's device in TFLOP/s: a register-only loop
 * of 16x16x4 FP64 products for QX_C128, FP32 products for the complex64 precisions
 * (the pipes the contractions use).  No reference counterpart: a measurement of the
 * peak that roofline fractions are quoted against (allocates and frees a small
 * output buffer; not for use under graph capture). */
qx_status qx_measure_mfma(qx_ctx *ctx, int precision, double *tflops);

/* Parse and validate a .qx program (docs/src/users_guide.md:95-165): load, view,
 * ncon, output, save.  Replaces QXContexts' DSL parse ("Parse input files",
 * docs/src/distributed.md:97). */
qx_status qx_plan_from_dsl(qx_ctx *ctx, const char *dsl, size_t len, int precision,
                           qx_plan **out);

/* Provide the values of one data key named by `load` lines (the JLD2 entries
 * written by save_cache, src/compute_graph/tensor_cache.jl:90-106).  `z` holds
 * `nelem` interleaved complex128 values in column-major order; it is copied. */
qx_status qx_plan_set_tensor(qx_plan *plan, const char *key, const double *z, int64_t nelem);

/* Replace the values of data key `key` after qx_plan_finalize, in the same device
 * buffer (layout of qx_plan_set_tensor, same element count), so lowered steps and
 * captured graphs stay valid; later calls see the new values.  The differentiable
 * seam: new parameters of a variational loop (the ComplexF64 gate tensors that
 * convert_to_tnc builds, src/tn_conversion.jl:10-40, and that single_amplitude
 * contracts, src/simulation.jl:86-91) enter here without re-parsing, re-lowering or
 * re-tuning the plan.  Returns after the copy has reached the device. */
qx_status qx_plan_update_tensor(qx_plan *plan, const char *key, const double *z, int64_t nelem);

/* Plan options, set before qx_plan_finalize ("profile", "graph", "fuse_slices" and
 * "autotune" may also be set later, "small_walk" too; setting "profile" clears the accumulated step
 * times):
 *   "batch"       bitstrings contracted together (power of two, default 256)
 *   "arena_mb"    device arena cap in MiB (default: 90% of free memory)
 *   "graph"       1 = capture the device work of a qx_amplitudes call (all slices
 *                 and batches) once per (batch, nb, slice range) and replay it;
 *                 ignored while "profile" is on
 *   "profile"     1 = time every step with HIP events (qx_plan_stats)
 *   "fuse_slices" 1 (default) = when a call covers every slice, contract as many
 *                 slice variables inline as fit half the arena cap instead of
 *                 looping over them; 0 = always loop
 *   "autotune"    1 = the next call of each lowering runs every candidate kernel of
 *                 every step (same results), times them with HIP events and keeps
 *                 the fastest per step for later calls; default 0
 *   "chains"      1 (default) = fold runs of gate-block steps on one large tensor
 *                 into single passes, runs of outer products into one launch each,
 *                 and batch small steps per dependency level; 0 = one launch per step
 *   "batch_last"  1 = lay the bitstring batch leg out slowest in bitstring-dependent
 *                 intermediates (default 0: fastest)
 *   "small_walk"  1 = a section whose batched small steps total at most 2^22
 *                 multiply-adds runs all their dependency levels in one launch of one
 *                 workgroup; 0 (default) = one launch per level (may be set at any time;
 *                 the walk is latency-bound and slower on the plans measured so far)
 *   "kernel"      -1 (default) = heuristic kernel choice; 0 gather, 1 rows, 2-4
 *                 FP tile GEMMs, 5 / 7 matrix-core GEMM (K chunks of 16 / 32), 6 reduce,
 *                 8 rocBLAS strided-batched GEMM (steps whose operands are plain
 *                 column-major matrices), 9 skinny contraction (K >= 2^12, at most 32
 *                 output values per batch element, N <= 2), 10 rocBLAS on gathered
 *                 plain copies (steps of >= 2^34 flops whose operands are not plain),
 *                 11 the library's own plain matrix-core GEMM (steps whose operands
 *                 are plain matrices), 12 the same with three real products per
 *                 complex product (complex128) = use that kernel on every step it can
 *                 run (testing; the autotuner also considers 8-12)
 *   "chain_variant" -1 (default) = heuristic chain-pass form (tuned with "autotune");
 *                 else bits 0-1 = path of steps with 16 contracted values (0 vector
 *                 FMAs, 1 matrix cores with four real products, 2 with three), bit 2 =
 *                 XOR-swizzled tile addressing where the pass has a swizzle, bit 3 = a
 *                 workgroup barrier after every step instead of wave-local steps
 *                 (testing) */
qx_status qx_plan_set_option(qx_plan *plan, const char *name, int64_t value);

/* Check that every data key is present, lower the program (label classes,
 * physical layouts, batched bitstring leg), compute the arena reuse schedule and
 * upload the initial tensors. */
qx_status qx_plan_finalize(qx_plan *plan);

/* Sizes a caller needs: number of bitstring characters per string (the largest
 * qubit named by an `output` line), complex values per bitstring in the result
 * (product of the saved tensor's open dims; 1 for a closed network), number of
 * slices (product of the slice variables' dims), number of distinct data keys
 * and the total complex elements over all keys. */
qx_status qx_plan_info(const qx_plan *plan, int64_t *num_chars, int64_t *result_elems,
                       int64_t *num_slices, int64_t *num_keys, int64_t *key_elems);

/* Name and element count of data key `i` (order of first `load` in the DSL);
 * `name` receives up to `cap` bytes including the terminating NUL. */
qx_status qx_plan_key(const qx_plan *plan, int64_t i, char *name, size_t cap, int64_t *nelem);

/* Amplitudes of `nb` bitstrings (nb * num_chars characters in {'0','1','+','-'},
 * character q = qubit q as in docs/src/basics.md:55-63), summed over slices
 * [slice_begin, slice_end) (slice 0 = all variables at their first value, first
 * variable fastest).  out: nb * result_elems interleaved complex128.
 * Replaces single_amplitude (src/simulation.jl:86-91) for a batch of bitstrings
 * and the slice loop + sum of QXContexts.execute (bin/qxrun.jl:83-87). */
qx_status qx_amplitudes(qx_plan *plan, const char *bits, int64_t nb, int64_t slice_begin,
                        int64_t slice_end, double *out);

/* qx_amplitudes that allocates no device memory, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments, results and slice
 * handling as qx_amplitudes; it contracts in batches of the plan's "batch" option and
 * uses the plan's arena, workspace and tables as an earlier qx_amplitudes call of at
 * least that many bitstrings sized them (QX_ERR_STATE otherwise).  Its per-call device
 * buffers (FP64 accumulators, staged bitstrings, index tables and results: about
 * 48 * nb * result_elems + nb * num_chars bytes) live in a fixed 8 MiB scratch area
 * of the library, one per device; larger calls return QX_ERR_ARGUMENT.  Every call on
 * a device shares that area: calls that overlap on one device (several threads or
 * contexts) race.  No graph capture, autotuning or profiling. */
qx_status qx_amplitudes_noalloc(qx_plan *plan, const char *bits, int64_t nb, int64_t slice_begin,
                                int64_t slice_end, double *out);

/* Output sampling on the device (the output methods of output_params_dict,
 * src/outputs.jl:47-78, executed by QXContexts and described in
 * docs/src/features.md:64-110).  Candidate j (0-based) of stream `seed` is the
 * bitstring whose character q is bit q of the Philox4x64-10 block at counter
 * (j,0,0,0), key (seed,0); its acceptance uniform u_j is the top 53 bits of the
 * first word of the block at key (seed,1), times 2^-53.  At most 256 qubits.
 *
 * :Uniform -- amplitudes of candidates 0..num-1 (all slices).  bits_out:
 * num * num_chars characters '0'/'1'; out: num * result_elems complex128. */
qx_status qx_amplitudes_uniform(qx_plan *plan, int64_t num, uint64_t seed, char *bits_out, double *out);

/* :Rejection -- empirical-supremum rejection sampling (Caffo, Booth, Davison,
 * Biometrika 89(4) 2002) with the uniform proposal: in candidate order,
 * r_j = |a_j|^2 * 2^num_chars, M_j = M_{j-1} (fix_M) or max(M_{j-1}, r_j), accept
 * when u_j * M_j < r_j (FP64, no fused multiply-add); stops at the num_samples-th
 * acceptance.  Candidates are contracted one batch at a time and tested on the
 * device.  Writes the accepted bitstrings (num_samples * num_chars) and their
 * amplitudes (num_samples complex128), the final M and the number of candidates
 * consumed.  QX_ERR_STATE when max_candidates run out first (outputs then hold
 * what was accepted); QX_ERR_UNSUPPORTED when output legs are left open. */
qx_status qx_sample_rejection(qx_plan *plan, int64_t num_samples, uint64_t seed, double M, int fix_M,
                              int64_t max_candidates, char *bits_out, double *amps_out, double *M_out,
                              int64_t *candidates_out);

/* Vector-Jacobian product of qx_amplitudes over all slices: given amp_cot (nb *
 * result_elems complex), writes for every data key (qx_plan_key order, keys
 * concatenated) grad = sum_b amp_cot_b * conj(d amp_b / d T_key), summed over all
 * load nodes that use the key.  This is the pullback ChainRulesCore uses for a
 * holomorphic map; it has no counterpart in the reference (ComplexF64 forward
 * only).  grad: key_elems interleaved complex128. */
qx_status qx_amplitudes_vjp(qx_plan *plan, const char *bits, int64_t nb, const double *amp_cot,
                            double *grad);

/* The loss sum over bitstrings and result elements of |amp|^2, in one pass: writes the
 * nb * result_elems amplitudes to `amps` (interleaved) and the loss's gradient per
 * data key to `grad` (layout of qx_amplitudes_vjp), forming the cotangent 2 amp on the
 * device from the same pass's forward.  QX_ERR_UNSUPPORTED when slice variables would
 * loop (the full amplitudes are then known only after every slice): use
 * qx_amplitudes + qx_amplitudes_vjp.  No reference counterpart (gradients are new). */
qx_status qx_norm_vjp(qx_plan *plan, const char *bits, int64_t nb, double *amps, double *grad);

/* JSON with per-step shapes (B, M, N, K), algorithmic flops and bytes, kernel
 * choice and, when "profile" is on, the accumulated time of every step. */
qx_status qx_plan_stats(const qx_plan *plan, char *json, size_t cap);

void qx_plan_destroy(qx_plan *plan);

/* Parse and lower a .qx program without a device and write the same JSON as
 * qx_plan_stats (steps, kernel choice, arena size) plus each arena value's offset
 * and live interval.  `fused` selects slice variables contracted inline (bit j =
 * j-th variable, as qx_amplitudes does over a full slice range when memory
 * allows).  For planning tools and host-side checks of the reuse schedule; no
 * reference counterpart. */
qx_status qx_dsl_inspect(const char *dsl, size_t len, int precision, int64_t batch, uint64_t fused,
                         char *json, size_t cap);

/* qx_dsl_inspect with lowering flags: bit 0 = the "batch_last" plan option (batched
 * bitstring leg slowest in bitstring-dependent intermediates).  No reference
 * counterpart. */
qx_status qx_dsl_inspect_ex(const char *dsl, size_t len, int precision, int64_t batch, uint64_t fused,
                            int64_t flags, char *json, size_t cap);

/* Message of the last failing call on this thread ("" if none). */
const char *qx_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* QXHIP_H */
