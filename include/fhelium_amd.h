/*
 * fhelium_amd -- C ABI of the gfx950 CKKS hot-path library (libfhelium_amd.so).
 *
 * Every entry point takes raw device pointers, element strides and sizes plus
 * a hipStream_t (passed as void*), launches asynchronously on that stream and
 * returns 0 on success or a negative status (fh_last_error() gives the text,
 * per host thread).  Nothing here allocates device memory.  Residues are
 * int64; an operand is a [batch, limb, coefficient] view (fh_view3) whose
 * limb j is parameter column j of the accompanying rns_params table, exactly
 * as in the reference's native ops (csrc/ops/common/rns_batch.h:11-116,
 * rns_parameters.h:6-16).  Each function names the reference operator
 * schema(s) it implements; the PyTorch shim (fhelium_amd/native) registers
 * them under those unchanged schemas for ROCm tensors.
 */
#ifndef FHELIUM_AMD_H_
#define FHELIUM_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* [batch, limb, coefficient] operand; a batch stride of 0 broadcasts one
 * batch entry (the reference's only broadcast, rns_batch.h:42-68). */
typedef struct fh_view3 {
  int64_t* data;
  int64_t stride_batch;
  int64_t stride_limb;
  int64_t stride_coef;
} fh_view3;

/* [parameter, limb] table, e.g. rns_params (rns_parameters.h:6-16). */
typedef struct fh_table2 {
  const int64_t* data;
  int64_t stride_row;
  int64_t stride_col;
} fh_table2;

typedef struct fh_vec {
  const int64_t* data;
  int64_t stride;
} fh_vec;

/* ---- runtime (replaces csrc/runtime/cuda_info.{h,cpp} inspection and
 *      provides the device services of a csrc/runtime backend) ---- */
const char* fh_last_error(void);
int fh_abi_version(void);
int fh_device_info(char* buf, int64_t buf_len);
/* HIP runtime version of the process (the manifest's hip_version; the
 * reference records cuda_version, fhelium/native/_abi.py:48-63). */
int fh_runtime_version(int* hip_runtime_version);
/* Device-to-device copy with 16-byte vector loads and stores (pointers
 * 16-byte aligned, non-overlapping); the byte reference of measured HBM
 * ceilings. */
int fh_stream_copy(void* dst, const void* src, int64_t bytes, void* stream);
/* Arena: one device allocation (on the current device) carved into aligned
 * workspaces by a bump pointer and reset as a whole, so a serving loop or a
 * captured graph reuses fixed addresses without allocator calls.  It is the
 * only memory the library allocates, and only in fh_arena_create. */
typedef struct fh_arena fh_arena;
int fh_arena_create(int64_t bytes, fh_arena** out);
int fh_arena_alloc(fh_arena* arena, int64_t bytes, int64_t alignment, void** out);
int fh_arena_reset(fh_arena* arena);
int fh_arena_stats(const fh_arena* arena, int64_t* capacity, int64_t* used, int64_t* high_water);
int fh_arena_destroy(fh_arena* arena);
/* Timing events (hipEvent_t behind void*). */
int fh_event_create(void** out);
int fh_event_record(void* event, void* stream);
int fh_event_elapsed_ms(void* start, void* stop, float* ms);
int fh_event_destroy(void* event);
/* Stream capture of a launch sequence into one replayable graph (the role of
 * fhelium/execution/cuda_graph.py:58-560 for hosts without PyTorch).
 * `stream` must not be the null stream. */
int fh_graph_begin(void* stream);
int fh_graph_end(void* stream, void** graph_exec);
int fh_graph_launch(void* graph_exec, void* stream);
int fh_graph_destroy(void* graph_exec);

/* ---- NTT (csrc/ops/ntt/forward_ntt.cpp:25-80, inverse_ntt.cpp:23-82) ----
 * Backend tables: ntt_rows is one 128-byte record per QP row and
 * twiddles / twiddles_shoup are [rows, N], all produced by
 * fh_ntt_prepare_tables from the reference-format rns_params and compact
 * Montgomery twiddles (fhelium/engine/ntt/plans/twiddles.py:45-84).
 * Operand limb j uses table row row_start + j (ntt/interface.py:118-176).
 * mode 0 = exact (reference lazy outputs bit for bit), 1 = fast (canonical
 * forward outputs, equal modulo q; inverse outputs identical). */
int fh_ntt_prepare_tables(const int64_t* rns_params, int64_t param_row_stride,
                          int64_t param_limb_stride, int64_t limbs, int64_t log_n,
                          const int64_t* montgomery_twiddles, int64_t twiddle_row_stride,
                          void* ntt_rows_out, int64_t* twiddles_out,
                          int64_t* twiddles_shoup_out, void* stream);
/* forward_ntt_{montgomery,to_montgomery}_compact_grouped_smem_ and
 * forward_ntt_*_power_of_two_radix_compact_ (whose outer-twist tables the
 * torch shim turns into these compact tables): to_montgomery = 1 consumes
 * standard residues (x -> xR first).  Every call is one launch per pass
 * (columns, then tiles) over all batch entries; a caller may cut a batch
 * into several calls on several streams (fhelium_amd/autotune.py picks that
 * schedule per shape). */
int fh_ntt_forward(int64_t* residues, int64_t batch, int64_t limbs, int64_t log_n,
                   int64_t batch_stride, int64_t limb_stride, const void* ntt_rows,
                   const int64_t* twiddles, const int64_t* twiddles_shoup, int64_t row_start,
                   int32_t to_montgomery, int32_t mode, void* stream);
/* inverse_ntt_{montgomery,to_standard_lazy,to_standard,to_centered}_*:
 * output_form 0..3 in that order. */
int fh_ntt_inverse(int64_t* residues, int64_t batch, int64_t limbs, int64_t log_n,
                   int64_t batch_stride, int64_t limb_stride, const void* ntt_rows,
                   const int64_t* twiddles, const int64_t* twiddles_shoup, int64_t row_start,
                   int32_t output_form, int32_t mode, void* stream);
/* fh_ntt_forward reading `source` (same [batch, limb, N] shape, its own
 * strides) and writing `residues`; `source` is left unchanged.  Replaces the
 * clone + in-place transform of coefficient_domain_to_ntt_domain
 * (ckks_engine.py:876-903). */
int fh_ntt_forward_from(int64_t* residues, const int64_t* source, int64_t batch, int64_t limbs,
                        int64_t log_n, int64_t batch_stride, int64_t limb_stride,
                        int64_t source_batch_stride, int64_t source_limb_stride,
                        const void* ntt_rows, const int64_t* twiddles,
                        const int64_t* twiddles_shoup, int64_t row_start, int32_t to_montgomery,
                        int32_t mode, void* stream);
/* fh_ntt_inverse reading `source` (same [batch, limb, N] shape, its own
 * strides) and writing `residues`; `source` is left unchanged.  Replaces the
 * clone + in-place inverse of relinearize (ckks_engine.py:1608-1675). */
int fh_ntt_inverse_from(int64_t* residues, const int64_t* source, int64_t batch, int64_t limbs,
                        int64_t log_n, int64_t batch_stride, int64_t limb_stride,
                        int64_t source_batch_stride, int64_t source_limb_stride,
                        const void* ntt_rows, const int64_t* twiddles,
                        const int64_t* twiddles_shoup, int64_t row_start, int32_t output_form,
                        int32_t mode, void* stream);
/* One stage of forward/inverse_ntt_*_indexed_ (indexed_ntt_cpu.cpp:96-170,
 * 206-258) with explicit even/odd index rows and the stage's expanded
 * Montgomery twiddles; the caller validates the schedule first. */
int fh_ntt_indexed_stage(fh_view3 residues, int64_t batch, int64_t limbs,
                         const int32_t* even_indices, const int32_t* odd_indices,
                         int64_t index_stride, const int64_t* stage_twiddles,
                         int64_t twiddle_limb_stride, int64_t twiddle_butterfly_stride,
                         fh_table2 rns_params, int64_t butterflies, int32_t forward, void* stream);

/* ---- RNS arithmetic (csrc/ops/rns/rns_arithmetic.cpp:30-72,
 *      rns_canonical_arithmetic.cpp:15-23) ----
 * op: 0 montgomery_mul, 1 add_lazy, 2 sub_lazy, 3 add_canonical(_),
 *     4 sub_canonical(_), 5 montgomery_mul_cyclic_compressed,
 *     6 montgomery_mul_contiguous_compressed (rhs has compressed_support
 *     coefficients). */
int fh_rns_binary(int32_t op, fh_view3 out, fh_view3 lhs, fh_view3 rhs, fh_table2 rns_params,
                  int64_t batch, int64_t limbs, int64_t n, int64_t compressed_support,
                  void* stream);
/* op: 0 to_montgomery_, 1 from_montgomery_, 2 canonicalize_residues_,
 *     3 center_residues_, 4 shift_residues_positive_. */
int fh_rns_unary_inplace(int32_t op, fh_view3 residues, fh_table2 rns_params, int64_t batch,
                         int64_t limbs, int64_t n, void* stream);
/* montgomery_mul_row_scalars_ (canonical = 0) and
 * montgomery_mul_row_scalars_canonical (canonical = 1). */
int fh_rns_row_scalars(fh_view3 out, fh_view3 in, fh_vec scalars, fh_table2 rns_params,
                       int64_t batch, int64_t limbs, int64_t n, int32_t canonical, void* stream);
/* add_lazy_with_twice_modulus */
int fh_rns_add_lazy_twice_modulus(fh_view3 out, fh_view3 lhs, fh_view3 rhs, fh_vec twice_modulus,
                                  int64_t batch, int64_t limbs, int64_t n, void* stream);
/* lift_centered_coefficients: coefficients is [batch, 1, N] (limb stride 0). */
int fh_rns_lift_centered(fh_view3 out, fh_view3 coefficients, fh_vec twice_modulus, int64_t batch,
                         int64_t limbs, int64_t n, void* stream);
/* mixed_radix_decompose (rows <= 8) */
int fh_rns_mixed_radix_decompose(fh_view3 out, fh_view3 source, int64_t batch, int64_t rows,
                                 int64_t n, fh_vec normalizers, fh_table2 propagation,
                                 fh_vec modulus_lo, fh_vec modulus_hi, fh_vec neg_inv_lo,
                                 fh_vec neg_inv_hi, void* stream);
/* mixed_radix_basis_extend_to_montgomery */
int fh_rns_basis_extend(fh_view3 out, fh_view3 digits, int64_t batch, int64_t digit_count,
                        int64_t destination_rows, int64_t n, fh_table2 coefficients,
                        fh_table2 rns_params, void* stream);

/* mixed_radix_basis_extend_to_montgomery with the 64x64-bit products on the
 * matrix cores (byte-sliced INT8 MFMA, v_mfma_i32_32x32x16_i8, exact int32
 * byte-diagonal sums, then the same REDC and lazy additions): the trial of
 * MFMA for the path's only contraction.  Bit-identical to
 * fh_rns_basis_extend; 1..4 digits, N a multiple of 128. */
int fh_rns_basis_extend_mfma(fh_view3 out, fh_view3 digits, int64_t batch, int64_t digit_count,
                             int64_t destination_rows, int64_t n, fh_table2 coefficients,
                             fh_table2 rns_params, void* stream);

/* ---- CKKS (csrc/ops/ckks/ckks.cpp:29-90) ---- */
/* rescale_drop_leading_prime_{nearest,truncate}(_): dropped is
 * [batch, 1, N] (limb stride 0). */
int fh_ckks_rescale(fh_view3 out, fh_view3 remaining, fh_view3 dropped,
                    fh_vec drop_prime_inverse_mont, fh_table2 rns_params, int64_t batch,
                    int64_t limbs, int64_t n, int64_t half_drop_prime, int32_t nearest,
                    void* stream);
/* apply_coefficient_galois_automorphism (indices must lie in [0, N)) */
int fh_ckks_galois_coefficient(fh_view3 out, fh_view3 in, const int32_t* source_indices,
                               int64_t index_stride, const int8_t* source_sign,
                               int64_t sign_stride, fh_vec twice_modulus, int64_t batch,
                               int64_t limbs, int64_t n, void* stream);
/* apply_ntt_galois_automorphism (indices must lie in [0, N)) */
int fh_ckks_galois_ntt(fh_view3 out, fh_view3 in, const int32_t* source_indices,
                       int64_t index_stride, int64_t batch, int64_t limbs, int64_t n,
                       void* stream);
/* keyswitch_accumulate_digit_products_: key_digit is [2, key rows, N]. */
int fh_ckks_keyswitch_accumulate(fh_view3 acc0, fh_view3 acc1, fh_view3 digit,
                                 fh_view3 key_digit, fh_table2 rns_params, int64_t key_row_start,
                                 int64_t batch, int64_t limbs, int64_t n, void* stream);
/* keyswitch_accumulate_digit_products_ with one key per batch entry:
 * key_bases is a device array of [batch] key base addresses (each key
 * [digit, 2, QP, N] with the given element strides); entry b uses digit
 * digit_index of key_bases[b].  No reference counterpart: it lets a batch
 * of ciphertexts rotate by per-ciphertext steps in one launch. */
int fh_ckks_keyswitch_accumulate_multi_key(fh_view3 acc0, fh_view3 acc1, fh_view3 digit,
                                           const int64_t* key_bases, int64_t digit_index,
                                           int64_t key_stride_digit, int64_t key_stride_comp,
                                           int64_t key_stride_row, fh_table2 rns_params,
                                           int64_t key_row_start, int64_t batch, int64_t limbs,
                                           int64_t n, void* stream);
/* A whole hybrid key switch of a batched component in eight launches
 * (decompose, ModUp fused into the forward NTT column pass, tile pass fused
 * with the key MAC over every digit, inverse NTT, ModDown fused with the
 * engine's final canonical addition; a rotation adds one gather of base0).  Replaces the per-digit loop of
 * fhelium/engine/hybrid_keyswitch.py:321-371 plus :530-568 and
 * ckks_engine.py:1664-1667 / :3342, bit for bit; with Galois tables it also
 * replaces the coefficient automorphism of a rotation (ckks_engine.py:3822-3897).
 * The integer rows' MAC runs on a library-owned side stream (one per device)
 * forked from `stream` by an event and joined back to it before the inverse
 * transform, so the call is ordered like launches on `stream` alone (and is
 * captured with it).
 *   component        [batch, lq, N] canonical standard Q rows at `level`
 *   small_rows       number of leading QP rows at `level` whose prime is
 *                    below 2^42 (binary64 arithmetic; the rest are integer
 *                    rows, whose primes must be below 2^60)
 *   digit_table      [digit_count, 16] int64 digit descriptors (row_begin,
 *                    rows, key_digit, addresses of the digit's mixed-radix
 *                    and basis-extension tables, basis row stride, address
 *                    of the standard-form basis table, binary64 flag)
 *   p_digit_table    [16] int64 descriptor of the P primes taken as one
 *                    digit (mixed-radix tables, extension onto QP0)
 *   p_inverse_q      [Q0] P^{-1} mod q_j, standard form
 *   rns_params / ntt_rows / twiddles: the full level-0 QP tables
 *   key              shared key [digit, 2, QP0, N], or key_bases: [batch]
 *                    device addresses of one such key per entry
 *   galois_index_tables / galois_sign_tables: NULL, or [batch] device
 *                    addresses of int32 source-index / int8 sign tables [N];
 *                    component and base0 are then read through entry b's
 *                    automorphism (a rotation's key switch)
 *   workspace_*      digits [batch, lq, N], inter [batch, digit_count, lqp, N],
 *                    acc [2, batch, lqp, N]; with Galois tables the digit
 *                    workspace finally holds the gathered base0 rows
 *   base0 / base1    optional canonical rows added to correction 0 / 1
 *                    (data == NULL: none); out0 / out1 receive the results.
 *   rescale_inverse  NULL, or [lq - 1] q_level^{-1} mod q_{level+i} (standard)
 *                    with rescale_inverse_shoup (Shoup companions) and
 *                    half_drop = floor(q_level / 2): the results are then
 *                    rescaled as rescale_to_next_level does
 *                    (ckks_engine.py:1131-1174, ckks_cpu.cpp:204-275) and
 *                    out0 / out1 receive lq - 1 rows.
 *   seed0 / seed1    NULL, or [batch, lq, N] contiguous NTT-domain rows in
 *                    the accumulators' scaling (what
 *                    fh_ckks_hmult_tensor_coefficient leaves in d0, d1 with
 *                    ntt_d01 = 1), with seed_scale [Q0] (P mod q for rows
 *                    below 2^42, P 2^62 mod q for the others): seed_scale x
 *                    seed is added to the Q rows of accumulator 0 / 1, so the
 *                    results gain exactly d0 / d1 (floor((X + P d) / P) =
 *                    floor(X / P) + d), as base0 / base1 = d0 / d1 would. */
int fh_keyswitch_fused(fh_view3 component, int64_t batch, int64_t lq, int64_t lqp, int64_t level,
                       int64_t log_n, int64_t small_rows, const int64_t* digit_table,
                       int64_t digit_count,
                       const int64_t* p_digit_table, const int64_t* p_inverse_q,
                       fh_table2 rns_params, const void* ntt_rows, const int64_t* forward_twiddles,
                       const int64_t* forward_twiddles_shoup, const int64_t* inverse_twiddles,
                       const int64_t* inverse_twiddles_shoup, const int64_t* key,
                       const int64_t* key_bases, int64_t key_stride_digit, int64_t key_stride_comp,
                       int64_t key_stride_row, const int64_t* galois_index_tables,
                       const int64_t* galois_sign_tables, int64_t* workspace_digits,
                       int64_t* workspace_inter, int64_t* workspace_acc, fh_view3 base0,
                       fh_view3 base1, fh_view3 out0, fh_view3 out1, int32_t mode,
                       const int64_t* rescale_inverse, const int64_t* rescale_inverse_shoup,
                       int64_t half_drop, const int64_t* seed0, const int64_t* seed1,
                       const int64_t* seed_scale, void* stream);

/* Hoisted rotations of one ciphertext by K rotation keys
 * (fhelium/engine/hybrid_keyswitch.py:373-504 prepare_rotation_digits +
 * _apply_prepared_rotation, composed as ckks_engine.py:3767-3820 does for
 * rotate_many_with_keys), bit for bit, in nine launches: c1 is decomposed
 * and extended onto QP once (one entry), then for every entry e the tile
 * pass + key MAC reads that intermediate through the NTT-domain Galois map
 * of g_e (hybrid_keyswitch.py:506-528), the accumulators are inverse
 * transformed, and ModDown adds sigma_{g_e}(c0) canonically.
 *   component        c1, [1, lq, N] canonical standard rows at `level`
 *   base0            c0, same shape (its batch stride is ignored)
 *   key_bases        [entries] device addresses of the entries' keys
 *                    [digit, 2, QP0, N] (strides given)
 *   galois_elements  [entries, 2] int64 device: (g mod 2N, g^{-1} mod 2N)
 *   galois_index_tables / galois_sign_tables: [entries] device addresses of
 *                    each g's coefficient gather tables (int32 [N], int8 [N])
 *   workspace_digits [entries, lq, N], workspace_inter [1, digit_count, lqp, N],
 *                    workspace_acc [2, entries, lqp, N]
 *   out0 / out1      [entries, lq, N]
 * The other arguments are fh_keyswitch_fused's. */
int fh_keyswitch_hoisted(fh_view3 component, int64_t entries, int64_t lq, int64_t lqp, int64_t level,
                         int64_t log_n, int64_t small_rows, const int64_t* digit_table, int64_t digit_count,
                         const int64_t* p_digit_table, const int64_t* p_inverse_q, fh_table2 rns_params,
                         const void* ntt_rows, const int64_t* forward_twiddles,
                         const int64_t* forward_twiddles_shoup, const int64_t* inverse_twiddles,
                         const int64_t* inverse_twiddles_shoup, const int64_t* key_bases,
                         int64_t key_stride_digit, int64_t key_stride_comp, int64_t key_stride_row,
                         const int64_t* galois_elements, const int64_t* galois_index_tables,
                         const int64_t* galois_sign_tables, int64_t* workspace_digits, int64_t* workspace_inter,
                         int64_t* workspace_acc, fh_view3 base0, fh_view3 out0, fh_view3 out1, int32_t mode,
                         void* stream);

/* fh_keyswitch_fused without caller workspaces: no counterpart in the
 * reference.  It is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The digits, ModUp and accumulator workspaces are carved from
 * a static 64 MiB device array (one per GPU, held until the process ends),
 * so batch * N * (lq + (digit_count + 2) * lqp) must not exceed 2^23 words
 * (e.g. one ciphertext at every level for N <= 2^14); larger calls return
 * an error before any launch.  Every launch runs on `stream` (no side
 * stream).  All calls on one GPU share the array: calls that overlap in
 * time on one GPU race, so serialize them (one stream, or events).  The
 * workspaces need no clearing; each is fully written before it is read.
 * Arguments are fh_keyswitch_fused's without the three workspaces, and the
 * results are bit-identical to it.  The function itself allocates nothing. */
int fh_keyswitch_fused_static(fh_view3 component, int64_t batch, int64_t lq, int64_t lqp, int64_t level,
                              int64_t log_n, int64_t small_rows, const int64_t* digit_table,
                              int64_t digit_count, const int64_t* p_digit_table,
                              const int64_t* p_inverse_q, fh_table2 rns_params, const void* ntt_rows,
                              const int64_t* forward_twiddles, const int64_t* forward_twiddles_shoup,
                              const int64_t* inverse_twiddles, const int64_t* inverse_twiddles_shoup,
                              const int64_t* key, const int64_t* key_bases, int64_t key_stride_digit,
                              int64_t key_stride_comp, int64_t key_stride_row,
                              const int64_t* galois_index_tables, const int64_t* galois_sign_tables,
                              fh_view3 base0, fh_view3 base1, fh_view3 out0, fh_view3 out1, int32_t mode,
                              const int64_t* rescale_inverse, const int64_t* rescale_inverse_shoup,
                              int64_t half_drop, const int64_t* seed0, const int64_t* seed1,
                              const int64_t* seed_scale, void* stream);

/* HMult front of the batching tier: d0 = a0 b0, d1 = a0 b1 + a1 b0,
 * d2 = a1 b1 of two coefficient-domain ciphertexts, returned in the
 * coefficient domain with canonical standard residues -- the composition
 * coefficient_domain_to_ntt_domain (ckks_engine.py:876-903) of both inputs,
 * multiply (:1530-1606) and the inverse NTT that starts relinearize
 * (:1608-1640), bit for bit, in five launches (one forward column pass per
 * operand -- two when its components are not one [2 batch] run -- one fused
 * tile pass per arithmetic path, one inverse column pass).
 *   a, b             [2, batch, limbs, N] canonical standard rows at level
 *                    row_start (component, batch, limb strides given;
 *                    coefficient stride 1)
 *   small_rows       leading rows whose prime is below 2^42
 *   workspace        [4, batch, limbs, N] contiguous
 *   out              [3, batch, limbs, N] contiguous
 *   ntt_d01          0: d0, d1, d2 all in the coefficient domain; 1: d0 and
 *                    d1 stay in the NTT domain as accumulator seeds for
 *                    fh_keyswitch_fused (working format: binary64 bits of
 *                    a b R^2 for rows below 2^42, a b R as u64 below 2q for
 *                    the others; within each 256-coefficient tile the value
 *                    of coefficient 8t + k at position t + 32k), d2 in the
 *                    coefficient domain
 * Fast NTT arithmetic only (the engine's default mode). */
int fh_ckks_hmult_tensor_coefficient(const int64_t* a, const int64_t* b, int64_t a_comp_stride,
                                     int64_t a_batch_stride, int64_t a_limb_stride, int64_t b_comp_stride,
                                     int64_t b_batch_stride, int64_t b_limb_stride, int64_t batch,
                                     int64_t limbs, int64_t log_n, int64_t row_start, int64_t small_rows,
                                     fh_table2 rns_params, const void* ntt_rows,
                                     const int64_t* forward_twiddles, const int64_t* forward_twiddles_shoup,
                                     const int64_t* inverse_twiddles, const int64_t* inverse_twiddles_shoup,
                                     int64_t* workspace, int64_t* out, int32_t ntt_d01,
                                     void* stream);

/* ---- Sampling: this backend's own ChaCha20 generator ------------------
 * Replaces the reference's ChaCha20 streams (fhelium/rng/csprng.py:150-250,
 * drawn from the external triton-csprng package): the RFC 8439 block
 * function, one block per thread.  key: 8 host words, nonce: 3 host words;
 * block b of a request uses block counter `counter + b` (the request must
 * not wrap the 32-bit counter).  Device outputs. */
int fh_chacha20_keystream(const uint32_t* key, const uint32_t* nonce, uint32_t counter, int64_t blocks,
                          uint32_t* out, void* stream);
/* out[e, r, i] uniform in [0, moduli[r]): 128 keystream bits per value
 * reduced mod q (the reference's randint_channels, csprng.py:207-219). */
int fh_sample_uniform_mod(const uint32_t* key, const uint32_t* nonce, uint32_t counter,
                          const int64_t* moduli, int64_t outer, int64_t rows, int64_t count, int64_t* out,
                          void* stream);
/* out[g] uniform in {-1, 0, 1}: 64 keystream bits per value, mod 3, minus 1. */
int fh_sample_ternary(const uint32_t* key, const uint32_t* nonce, uint32_t counter, int64_t total,
                      int64_t* out, void* stream);
/* out[g] = round(sigma * z), z standard normal by Box-Muller on two 53-bit
 * uniforms (the reference's discrete_gaussian, csprng.py:221-233). */
int fh_sample_gaussian(const uint32_t* key, const uint32_t* nonce, uint32_t counter, int64_t total,
                       double sigma, int64_t* out, void* stream);

/* ---- Slot codec on the device (binary64) --------------------------------
 * The reference encodes and decodes on the host side of
 * ckks_plaintext_codec.py:181-316 / ckks_decryptor.py:96-272 through
 * slot_embedding.py; these run the same canonical embedding next to the
 * data.  slot_positions: int32 [N/2], (3^{-j} mod 2N - 1) / 2.
 * fh_ckks_encode: slots [batch, N/2] complex as (re, im) binary64 pairs ->
 *   out [batch, limbs, N] canonical standard residues of
 *   rint(scale * coefficient) for moduli[0..limbs); workspace: batch * N
 *   complex (16 bytes each). */
int fh_ckks_encode(const double* slots, int64_t batch, int64_t log_n, const int32_t* slot_positions,
                   double scale, const int64_t* moduli, int64_t limbs, double* workspace, int64_t* out,
                   void* stream);
/* fh_ckks_decode: centered CRT of two canonical residue rows (batch stride
 * batch_stride) over moduli a, b (a^{-1} mod b given), then the embedding
 * divided by scale -> slots [batch, N/2] complex (re, im) binary64. */
int fh_ckks_decode(const int64_t* residues_a, const int64_t* residues_b, int64_t batch_stride, int64_t batch,
                   int64_t log_n, int64_t modulus_a, int64_t modulus_b, int64_t modulus_a_inverse_mod_b,
                   const int32_t* slot_positions, double scale, double* workspace, double* slots,
                   void* stream);
/* keyswitch_moddown_qp_to_q */
int fh_ckks_moddown(fh_view3 out, fh_view3 q_residues, fh_view3 p_residues,
                    fh_table2 p_drop_inverses, fh_table2 rns_params, int64_t batch,
                    int64_t q_count, int64_t p_count, int64_t n, void* stream);
/* add_{prepared,cyclic_compressed,contiguous_compressed,strided}_plaintext_component(_):
 * layout 0..3 in that order. */
int fh_ckks_add_plaintext(int32_t layout, fh_view3 out, fh_view3 ciphertext, fh_view3 plaintext,
                          const int64_t* implicit_plaintext, int64_t plaintext_batch_count,
                          fh_table2 rns_params, int64_t batch, int64_t limbs, int64_t n,
                          int64_t support, void* stream);
/* The diagonal method's inner product over a batch of ciphertext entries:
 * out[c, l, i] = sum_b ct[c, b, l, i] * pt[b, l, i] for c < components, the
 * Montgomery products of multiply_plaintext (ckks_engine.py:2077-2168)
 * summed canonically -- the value multiply_plaintext followed by pairwise
 * additions gives (fhelium/benchmarks/standalone/packed_matrix_vector.py:
 * 120-170), in one pass.  No reference operator; the engine's matrix-vector
 * product uses it.  Coefficient stride 1 everywhere. */
int fh_ckks_plaintext_inner_product(int64_t* out, int64_t out_comp_stride, int64_t out_limb_stride,
                                    const int64_t* ciphertext, int64_t ct_comp_stride, int64_t ct_batch_stride,
                                    int64_t ct_limb_stride, const int64_t* plaintext, int64_t pt_batch_stride,
                                    int64_t pt_limb_stride, int64_t components, int64_t batch, int64_t limbs,
                                    int64_t n, fh_table2 rns_params, void* stream);
/* Integer linear combination of ciphertext components:
 * out[c, e, l, i] = sum_t x_t[c, e, l, i] * m_t mod q_l (canonical), with
 * scalars[t * scalar_stride + l] = m_t R mod q_l (Montgomery form).  Each
 * product equals multiply_plaintext by a constant plaintext, which encodes
 * as the constant polynomial m_t (ckks_engine.py:2077-2168), in either
 * domain; same-scale terms then share one rescale.  No reference operator:
 * the serving program's plaintext-constant products use it.  terms is a
 * host array of term_count (1..8) device addresses of each term's row 0 at
 * the output's level; term_strides a host array of term_count x {component,
 * batch, limb} strides.  Coefficient stride 1 everywhere, operands
 * canonical. */
int fh_ckks_scalar_combination(int64_t* out, int64_t out_comp_stride, int64_t out_batch_stride,
                               int64_t out_limb_stride, const int64_t* const* terms, const int64_t* term_strides,
                               const int64_t* scalars, int64_t scalar_stride, int64_t term_count,
                               int64_t components, int64_t batch, int64_t limbs, int64_t n,
                               fh_table2 rns_params, void* stream);
/* CkksEngine.multiply's five native calls as one pass
 * (fhelium/engine/ckks_engine.py:1583-1598). */
int fh_ckks_tensor_product(fh_view3 d0, fh_view3 d1, fh_view3 d2, fh_view3 a0, fh_view3 a1,
                           fh_view3 b0, fh_view3 b1, fh_table2 rns_params, int64_t batch,
                           int64_t limbs, int64_t n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FHELIUM_AMD_H_ */
