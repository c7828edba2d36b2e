// CKKS-level kernels for gfx950: rescale, Galois automorphisms, key-switch
// multiply-accumulate, ModDown and plaintext addition.  Each reproduces the
// reference CPU loop body (csrc/ops/ckks/cpu/ckks_cpu.cpp) value for value.
#include "fh_common.h"
#include "fh_views.h"

namespace fh {
namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 4;
constexpr int kChunk = kThreads * kPerThread;

// rescale_drop_leading_prime_{nearest,truncate} (ckks_cpu.cpp:204-275):
// c'_i = canonicalize(REDC((c_i - c_drop) * q_drop^{-1} R) [+1]).
template <bool kNearest>
__global__ __launch_bounds__(kThreads) void k_rescale(View3 out, View3 remaining,
                                                      View3 dropped, Vec inverse,
                                                      ParamView params, int64_t half_drop,
                                                      int64_t n) {
  const int64_t l = blockIdx.y;
  const int64_t b = blockIdx.z;
  const Mont m = load_mont(params, l);
  const int64_t inv = inverse[l];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kChunk + threadIdx.x;
#pragma unroll
  for (int e = 0; e < kPerThread; ++e) {
    const int64_t i = base + e * kThreads;
    if (i >= n) break;
    const int64_t d = dropped.at(b, 0, i);
    int64_t v = mont_mul_split(remaining.at(b, l, i) - d, inv, m);
    if constexpr (kNearest) v += d > half_drop ? 1 : 0;
    out.at(b, l, i) = canonicalize(v, m.twice_q);
  }
}

// apply_coefficient_galois_automorphism (ckks_cpu.cpp:323-405).
__global__ __launch_bounds__(kThreads) void k_galois_coefficient(
    View3 out, View3 in, const int32_t* src, int64_t src_stride, const int8_t* sign,
    int64_t sign_stride, Vec twice, int64_t n) {
  const int64_t l = blockIdx.y;
  const int64_t b = blockIdx.z;
  const int64_t t2 = twice[l];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kChunk + threadIdx.x;
#pragma unroll
  for (int e = 0; e < kPerThread; ++e) {
    const int64_t i = base + e * kThreads;
    if (i >= n) break;
    int64_t v = in.at(b, l, src[i * src_stride]);
    if (sign[i * sign_stride] == -1) v = -v;
    out.at(b, l, i) = canonicalize(shift_positive(v, t2), t2);
  }
}

// apply_ntt_galois_automorphism (ckks_cpu.cpp:407-472): a pure gather.
__global__ __launch_bounds__(kThreads) void k_galois_ntt(View3 out, View3 in,
                                                         const int32_t* src,
                                                         int64_t src_stride, int64_t n) {
  const int64_t l = blockIdx.y;
  const int64_t b = blockIdx.z;
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kChunk + threadIdx.x;
#pragma unroll
  for (int e = 0; e < kPerThread; ++e) {
    const int64_t i = base + e * kThreads;
    if (i >= n) break;
    out.at(b, l, i) = in.at(b, l, src[i * src_stride]);
  }
}

// keyswitch_accumulate_digit_products_ (ckks_cpu.cpp:603-708):
// acc_k += REDC(digit * key_k[key_row_start + limb]) lazily, k = 0, 1.
__global__ __launch_bounds__(kThreads) void k_keyswitch_mac(View3 acc0, View3 acc1,
                                                            View3 digit, View3 key,
                                                            ParamView params,
                                                            int64_t key_row_start,
                                                            int64_t n) {
  const int64_t l = blockIdx.y;
  const int64_t b = blockIdx.z;
  const Mont m = load_mont(params, l);
  const int64_t krow = key_row_start + l;
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kChunk + threadIdx.x;
#pragma unroll
  for (int e = 0; e < kPerThread; ++e) {
    const int64_t i = base + e * kThreads;
    if (i >= n) break;
    const int64_t x = digit.at(b, l, i);
    const int64_t p0 = mont_mul(x, key.at(0, krow, i), m);
    const int64_t p1 = mont_mul(x, key.at(1, krow, i), m);
    acc0.at(b, l, i) = add_lazy(acc0.at(b, l, i), p0, m.twice_q);
    acc1.at(b, l, i) = add_lazy(acc1.at(b, l, i), p1, m.twice_q);
  }
}

// Same multiply-accumulate with a different key per batch entry: entry b
// reads key digit `digit_index` of the key whose base address is
// key_bases[b] (each key [digit, 2, QP, N] with the given element strides).
// Used when a batch of ciphertexts rotates by per-ciphertext steps.
__global__ __launch_bounds__(kThreads) void k_keyswitch_mac_multi(
    View3 acc0, View3 acc1, View3 digit, const int64_t* key_bases, int64_t digit_index,
    int64_t key_stride_digit, int64_t key_stride_comp, int64_t key_stride_row, ParamView params,
    int64_t key_row_start, int64_t n) {
  const int64_t l = blockIdx.y;
  const int64_t b = blockIdx.z;
  const Mont m = load_mont(params, l);
  const int64_t* key = reinterpret_cast<const int64_t*>(key_bases[b]) +
                       digit_index * key_stride_digit + (key_row_start + l) * key_stride_row;
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kChunk + threadIdx.x;
#pragma unroll
  for (int e = 0; e < kPerThread; ++e) {
    const int64_t i = base + e * kThreads;
    if (i >= n) break;
    const int64_t x = digit.at(b, l, i);
    const int64_t p0 = mont_mul(x, key[i], m);
    const int64_t p1 = mont_mul(x, key[key_stride_comp + i], m);
    acc0.at(b, l, i) = add_lazy(acc0.at(b, l, i), p0, m.twice_q);
    acc1.at(b, l, i) = add_lazy(acc1.at(b, l, i), p1, m.twice_q);
  }
}

// keyswitch_moddown_qp_to_q (ckks_cpu.cpp:474-601): floor(x / P) mod q_i by
// sequentially dropping canonical P residues, one coefficient per thread.
constexpr int kMaxP = 8;

__global__ __launch_bounds__(kThreads) void k_moddown(View3 out, View3 qv, View3 pv,
                                                      const int64_t* inverse,
                                                      int64_t inv_drop_stride,
                                                      int64_t inv_limb_stride,
                                                      ParamView params, int q_count,
                                                      int p_count, int64_t n) {
  const int64_t b = blockIdx.y;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  int64_t chain[kMaxP];
#pragma unroll
  for (int r = 0; r < kMaxP; ++r) chain[r] = r < p_count ? pv.at(b, r, i) : 0;
  // P-chain: each P residue is made exact modulo the lower P primes.
  for (int row = p_count - 2; row >= 0; --row) {
    int64_t value = chain[row];
    const Mont m = load_mont(params, q_count + row);
    for (int lower = p_count - 1; lower > row; --lower) {
      const int64_t diff = sub_lazy(value, chain[lower], m.twice_q);
      value = mont_mul_split(
          diff,
          inverse[(p_count - lower - 1) * inv_drop_stride + (row + q_count) * inv_limb_stride],
          m);
    }
    chain[row] = value;
  }
  for (int row = 0; row < q_count; ++row) {
    const Mont m = load_mont(params, row);
    int64_t value = mont_mul(qv.at(b, row, i), m.r2, m);
    for (int p = p_count - 1; p >= 0; --p) {
      const int64_t pm = mont_mul(chain[p], m.r2, m);
      value = sub_lazy(value, pm, m.twice_q);
      value = mont_mul_split(
          value, inverse[(p_count - p - 1) * inv_drop_stride + row * inv_limb_stride], m);
    }
    value = mont_mul(value, 1, m);
    out.at(b, row, i) = canonicalize(value, m.twice_q);
  }
}

// Plaintext addition (ckks_cpu.cpp:54-146): c + p with p in Montgomery form,
// result canonical standard.  Layouts: dense, cyclic / contiguous compressed
// and strided with an implicit per-limb fill value.
enum PlainLayout : int { kDense = 0, kCyclic = 1, kContiguous = 2, kStrided = 3 };

template <int LAYOUT>
__global__ __launch_bounds__(kThreads) void k_add_plain(View3 out, View3 ct, View3 pt,
                                                        const int64_t* implicit,
                                                        int64_t pt_batch_count,
                                                        int64_t limb_count,
                                                        ParamView params, int64_t support,
                                                        int64_t n) {
  const int64_t l = blockIdx.y;
  const int64_t b = blockIdx.z;
  const Mont m = load_mont(params, l);
  const int64_t repeat = n / support;
  const int64_t pb = pt_batch_count == 1 ? 0 : b;
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kChunk + threadIdx.x;
#pragma unroll
  for (int e = 0; e < kPerThread; ++e) {
    const int64_t i = base + e * kThreads;
    if (i >= n) break;
    int64_t p;
    if constexpr (LAYOUT == kDense) {
      p = pt.at(b, l, i);
    } else if constexpr (LAYOUT == kCyclic) {
      p = pt.at(b, l, i % support);
    } else if constexpr (LAYOUT == kContiguous) {
      p = pt.at(b, l, i / repeat);
    } else {
      p = implicit[pb * limb_count + l];
      if (i % repeat == 0) p = pt.at(b, l, i / repeat);
    }
    int64_t v = mont_mul(ct.at(b, l, i), m.r2, m);
    v = add_lazy(v, p, m.twice_q);
    v = mont_mul(v, 1, m);
    out.at(b, l, i) = canonicalize(v, m.twice_q);
  }
}

// CkksEngine.multiply (ckks_engine.py:1583-1598) in one pass: d0 = a0 b0,
// d1 = add_lazy(a0 b1, a1 b0), d2 = a1 b1 with the reference's REDC.  Reads
// four limb sets and writes three instead of the 12 + 3 of five separate ops.
__global__ __launch_bounds__(kThreads) void k_tensor_product(View3 d0, View3 d1, View3 d2, View3 a0,
                                                             View3 a1, View3 b0, View3 b1,
                                                             ParamView params, int64_t n) {
  const int64_t l = blockIdx.y;
  const int64_t b = blockIdx.z;
  const Mont m = load_mont(params, l);
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kChunk + threadIdx.x;
#pragma unroll
  for (int e = 0; e < kPerThread; ++e) {
    const int64_t i = base + e * kThreads;
    if (i >= n) break;
    const int64_t x0 = a0.at(b, l, i), x1 = a1.at(b, l, i);
    const int64_t y0 = b0.at(b, l, i), y1 = b1.at(b, l, i);
    d0.at(b, l, i) = mont_mul(x0, y0, m);
    d1.at(b, l, i) = add_lazy(mont_mul(x0, y1, m), mont_mul(x1, y0, m), m.twice_q);
    d2.at(b, l, i) = mont_mul(x1, y1, m);
  }
}

// Diagonal-method inner product: out[c, l, i] = sum_b ct[c, b, l, i] * pt[b, l, i]
// (Montgomery products, canonical sums).  Each product is multiply_plaintext's
// canonical value and the sum of canonical residues is order-independent, so
// the result equals multiply_plaintext followed by pairwise additions
// (packed_matrix_vector.py:120-170), in one read of both operands.
__global__ __launch_bounds__(kThreads) void k_plaintext_inner_product(
    int64_t* out, int64_t out_comp_stride, int64_t out_limb_stride, const int64_t* ct, int64_t ct_comp_stride,
    int64_t ct_batch_stride, int64_t ct_limb_stride, const int64_t* pt, int64_t pt_batch_stride,
    int64_t pt_limb_stride, int64_t batch, ParamView params, int64_t n) {
  const int64_t l = blockIdx.y;
  const int64_t c = blockIdx.z;
  const Mont m = load_mont(params, l);
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kChunk + threadIdx.x;
  const int64_t* crow = ct + c * ct_comp_stride + l * ct_limb_stride;
  const int64_t* prow = pt + l * pt_limb_stride;
#pragma unroll
  for (int e = 0; e < kPerThread; ++e) {
    const int64_t i = base + e * kThreads;
    if (i >= n) break;
    const int64_t q = static_cast<int64_t>(m.q);
    int64_t acc = 0;
    for (int64_t b = 0; b < batch; ++b) {
      const int64_t v = acc + mont_mul(crow[b * ct_batch_stride + i], prow[b * pt_batch_stride + i], m);
      acc = v >= q ? v - q : v;  // both terms canonical
    }
    out[c * out_comp_stride + l * out_limb_stride + i] = acc;
  }
}

// Integer linear combination: out[c, e, l, i] = sum_t x_t[c, e, l, i] m_t mod q_l
// for up to kMaxCombinationTerms terms, with s_t[l] = m_t R mod q_l so that
// each REDC product is the canonical x m_t mod q_l.  A constant plaintext
// (every slot the same value) encodes as the constant polynomial m, so each
// product is what multiply_plaintext gives for it -- in either domain, since
// the NTT of a constant polynomial is that constant -- and terms carrying
// the same scale sum before one rescale.  Operands canonical; each term has
// its own strides, so a term at a lower level is read from its row offset.
constexpr int kMaxCombinationTerms = 8;

struct CombinationTerms {
  const int64_t* ptr[kMaxCombinationTerms];
  int64_t comp_stride[kMaxCombinationTerms];
  int64_t batch_stride[kMaxCombinationTerms];
  int64_t limb_stride[kMaxCombinationTerms];
  int count;
};

__global__ __launch_bounds__(kThreads) void k_scalar_combination(
    int64_t* out, int64_t out_comp_stride, int64_t out_batch_stride, int64_t out_limb_stride,
    CombinationTerms terms, const int64_t* scalars, int64_t scalar_stride, int64_t batch, ParamView params,
    int64_t n) {
  const int64_t l = blockIdx.y;
  const int64_t c = blockIdx.z / batch;
  const int64_t e = blockIdx.z % batch;
  const Mont m = load_mont(params, l);
  const int64_t q = static_cast<int64_t>(m.q);
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kChunk + threadIdx.x;
  int64_t* orow = out + c * out_comp_stride + e * out_batch_stride + l * out_limb_stride;
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const int64_t i = base + k * kThreads;
    if (i >= n) break;
    int64_t acc = 0;
    for (int t = 0; t < terms.count; ++t) {
      const int64_t x = terms.ptr[t][c * terms.comp_stride[t] + e * terms.batch_stride[t] +
                                     l * terms.limb_stride[t] + i];
      const int64_t v = acc + mont_mul(x, scalars[t * scalar_stride + l], m);
      acc = v >= q ? v - q : v;  // both terms canonical
    }
    orow[i] = acc;
  }
}

int check_grid(int64_t batch, int64_t limbs, int64_t n) {
  FH_CHECK_ARG(batch >= 0 && limbs >= 0 && n >= 0, "negative extent");
  FH_CHECK_ARG(batch <= 65535 && limbs <= 65535,
               "batch (%lld) or limb (%lld) extent exceeds the launch grid",
               (long long)batch, (long long)limbs);
  return kOk;
}

}  // namespace
}  // namespace fh

using namespace fh;

// See include/fhelium_amd.h.
FH_API int fh_ckks_plaintext_inner_product(int64_t* out, int64_t out_comp_stride, int64_t out_limb_stride,
                                           const int64_t* ciphertext, int64_t ct_comp_stride,
                                           int64_t ct_batch_stride, int64_t ct_limb_stride,
                                           const int64_t* plaintext, int64_t pt_batch_stride,
                                           int64_t pt_limb_stride, int64_t components, int64_t batch,
                                           int64_t limbs, int64_t n, fh_table2 rns_params, void* stream) {
  int st = check_grid(components, limbs, n);
  if (st != kOk) return st;
  FH_CHECK_ARG(batch >= 1, "inner product needs at least one entry");
  if (components == 0 || limbs == 0 || n == 0) return kOk;
  FH_CHECK_ARG(out && ciphertext && plaintext && rns_params.data, "inner product: null operand");
  ParamView pv{rns_params.data, rns_params.stride_row, rns_params.stride_col};
  dim3 grid(ceil_div(n, kChunk), limbs, components);
  hipLaunchKernelGGL(k_plaintext_inner_product, grid, dim3(kThreads), 0, static_cast<hipStream_t>(stream), out,
                     out_comp_stride, out_limb_stride, ciphertext, ct_comp_stride, ct_batch_stride, ct_limb_stride,
                     plaintext, pt_batch_stride, pt_limb_stride, batch, pv, n);
  FH_CHECK_LAUNCH("plaintext inner product");
  return kOk;
}

// See include/fhelium_amd.h.
FH_API int fh_ckks_scalar_combination(int64_t* out, int64_t out_comp_stride, int64_t out_batch_stride,
                                      int64_t out_limb_stride, const int64_t* const* terms,
                                      const int64_t* term_strides, const int64_t* scalars, int64_t scalar_stride,
                                      int64_t term_count, int64_t components, int64_t batch, int64_t limbs,
                                      int64_t n, fh_table2 rns_params, void* stream) {
  FH_CHECK_ARG(term_count >= 1 && term_count <= kMaxCombinationTerms, "scalar combination takes 1..%d terms",
               kMaxCombinationTerms);
  FH_CHECK_ARG(components >= 0 && batch >= 1 && components * batch <= 65535,
               "scalar combination: bad component (%lld) or batch (%lld) extent", (long long)components,
               (long long)batch);
  int st = check_grid(components * batch, limbs, n);
  if (st != kOk) return st;
  if (components == 0 || limbs == 0 || n == 0) return kOk;
  FH_CHECK_ARG(out && terms && term_strides && scalars && rns_params.data, "scalar combination: null operand");
  CombinationTerms tv{};
  tv.count = static_cast<int>(term_count);
  for (int64_t t = 0; t < term_count; ++t) {
    FH_CHECK_ARG(terms[t] != nullptr, "scalar combination: term %lld is null", (long long)t);
    tv.ptr[t] = terms[t];
    tv.comp_stride[t] = term_strides[3 * t];
    tv.batch_stride[t] = term_strides[3 * t + 1];
    tv.limb_stride[t] = term_strides[3 * t + 2];
  }
  ParamView pv{rns_params.data, rns_params.stride_row, rns_params.stride_col};
  dim3 grid(ceil_div(n, kChunk), limbs, components * batch);
  hipLaunchKernelGGL(k_scalar_combination, grid, dim3(kThreads), 0, static_cast<hipStream_t>(stream), out,
                     out_comp_stride, out_batch_stride, out_limb_stride, tv, scalars, scalar_stride, batch, pv, n);
  FH_CHECK_LAUNCH("scalar combination");
  return kOk;
}

FH_API int fh_ckks_rescale(fh_view3 out, fh_view3 remaining, fh_view3 dropped,
                           fh_vec drop_prime_inverse_mont, fh_table2 rns_params,
                           int64_t batch, int64_t limbs, int64_t n,
                           int64_t half_drop_prime, int32_t nearest, void* stream) {
  int st = check_grid(batch, limbs, n);
  if (st != kOk) return st;
  if (batch == 0 || limbs == 0 || n == 0) return kOk;
  FH_CHECK_ARG(out.data && remaining.data && dropped.data && drop_prime_inverse_mont.data &&
                   rns_params.data,
               "rescale: null operand");
  ParamView pv{rns_params.data, rns_params.stride_row, rns_params.stride_col};
  dim3 grid(ceil_div(n, kChunk), limbs, batch);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (nearest) {
    hipLaunchKernelGGL(k_rescale<true>, grid, dim3(kThreads), 0, s, View3(out), View3(remaining),
                       View3(dropped), Vec(drop_prime_inverse_mont), pv, half_drop_prime, n);
  } else {
    hipLaunchKernelGGL(k_rescale<false>, grid, dim3(kThreads), 0, s, View3(out), View3(remaining),
                       View3(dropped), Vec(drop_prime_inverse_mont), pv, half_drop_prime, n);
  }
  FH_CHECK_LAUNCH("ckks rescale");
  return kOk;
}

FH_API int fh_ckks_galois_coefficient(fh_view3 out, fh_view3 in, const int32_t* source_indices,
                                      int64_t index_stride, const int8_t* source_sign,
                                      int64_t sign_stride, fh_vec twice_modulus,
                                      int64_t batch, int64_t limbs, int64_t n,
                                      void* stream) {
  int st = check_grid(batch, limbs, n);
  if (st != kOk) return st;
  if (batch == 0 || limbs == 0 || n == 0) return kOk;
  FH_CHECK_ARG(out.data && in.data && source_indices && source_sign && twice_modulus.data,
               "coefficient galois: null operand");
  dim3 grid(ceil_div(n, kChunk), limbs, batch);
  hipLaunchKernelGGL(k_galois_coefficient, grid, dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), View3(out), View3(in), source_indices,
                     index_stride, source_sign, sign_stride, Vec(twice_modulus), n);
  FH_CHECK_LAUNCH("ckks coefficient galois");
  return kOk;
}

FH_API int fh_ckks_galois_ntt(fh_view3 out, fh_view3 in, const int32_t* source_indices,
                              int64_t index_stride, int64_t batch, int64_t limbs, int64_t n,
                              void* stream) {
  int st = check_grid(batch, limbs, n);
  if (st != kOk) return st;
  if (batch == 0 || limbs == 0 || n == 0) return kOk;
  FH_CHECK_ARG(out.data && in.data && source_indices, "ntt galois: null operand");
  dim3 grid(ceil_div(n, kChunk), limbs, batch);
  hipLaunchKernelGGL(k_galois_ntt, grid, dim3(kThreads), 0, static_cast<hipStream_t>(stream),
                     View3(out), View3(in), source_indices, index_stride, n);
  FH_CHECK_LAUNCH("ckks ntt galois");
  return kOk;
}

FH_API int fh_ckks_keyswitch_accumulate(fh_view3 acc0, fh_view3 acc1, fh_view3 digit,
                                        fh_view3 key_digit, fh_table2 rns_params,
                                        int64_t key_row_start, int64_t batch,
                                        int64_t limbs, int64_t n, void* stream) {
  int st = check_grid(batch, limbs, n);
  if (st != kOk) return st;
  if (batch == 0 || limbs == 0 || n == 0) return kOk;
  FH_CHECK_ARG(acc0.data && acc1.data && digit.data && key_digit.data && rns_params.data,
               "keyswitch accumulate: null operand");
  FH_CHECK_ARG(key_row_start >= 0, "negative key row start");
  ParamView pv{rns_params.data, rns_params.stride_row, rns_params.stride_col};
  dim3 grid(ceil_div(n, kChunk), limbs, batch);
  hipLaunchKernelGGL(k_keyswitch_mac, grid, dim3(kThreads), 0, static_cast<hipStream_t>(stream),
                     View3(acc0), View3(acc1), View3(digit), View3(key_digit), pv,
                     key_row_start, n);
  FH_CHECK_LAUNCH("ckks keyswitch accumulate");
  return kOk;
}

FH_API int fh_ckks_moddown(fh_view3 out, fh_view3 q_residues, fh_view3 p_residues,
                           fh_table2 p_drop_inverses, fh_table2 rns_params, int64_t batch,
                           int64_t q_count, int64_t p_count, int64_t n, void* stream) {
  int st = check_grid(batch, q_count, n);
  if (st != kOk) return st;
  FH_CHECK_ARG(p_count >= 1 && p_count <= kMaxP, "ModDown supports 1..8 P rows");
  if (batch == 0 || q_count == 0 || n == 0) return kOk;
  FH_CHECK_ARG(out.data && q_residues.data && p_residues.data && p_drop_inverses.data &&
                   rns_params.data,
               "moddown: null operand");
  ParamView pv{rns_params.data, rns_params.stride_row, rns_params.stride_col};
  dim3 grid(ceil_div(n, kThreads), batch);
  hipLaunchKernelGGL(k_moddown, grid, dim3(kThreads), 0, static_cast<hipStream_t>(stream),
                     View3(out), View3(q_residues), View3(p_residues), p_drop_inverses.data,
                     p_drop_inverses.stride_row, p_drop_inverses.stride_col, pv,
                     static_cast<int>(q_count), static_cast<int>(p_count), n);
  FH_CHECK_LAUNCH("ckks moddown");
  return kOk;
}

FH_API int fh_ckks_add_plaintext(int32_t layout, fh_view3 out, fh_view3 ciphertext,
                                 fh_view3 plaintext, const int64_t* implicit_plaintext,
                                 int64_t plaintext_batch_count, fh_table2 rns_params,
                                 int64_t batch, int64_t limbs, int64_t n, int64_t support,
                                 void* stream) {
  int st = check_grid(batch, limbs, n);
  if (st != kOk) return st;
  if (batch == 0 || limbs == 0 || n == 0) return kOk;
  FH_CHECK_ARG(out.data && ciphertext.data && plaintext.data && rns_params.data,
               "plaintext add: null operand");
  if (layout == kDense) {
    support = n;
  } else {
    FH_CHECK_ARG(support > 0 && support < n && (support & (support - 1)) == 0 &&
                     n % support == 0,
                 "compressed support must be a power-of-two divisor of N");
  }
  FH_CHECK_ARG(layout != kStrided || implicit_plaintext, "strided plaintext needs fill values");
  ParamView pv{rns_params.data, rns_params.stride_row, rns_params.stride_col};
  dim3 grid(ceil_div(n, kChunk), limbs, batch);
  hipStream_t s = static_cast<hipStream_t>(stream);
  View3 o(out), c(ciphertext), p(plaintext);
  switch (layout) {
    case kDense: hipLaunchKernelGGL(k_add_plain<kDense>, grid, dim3(kThreads), 0, s, o, c, p, implicit_plaintext, plaintext_batch_count, limbs, pv, support, n); break;
    case kCyclic: hipLaunchKernelGGL(k_add_plain<kCyclic>, grid, dim3(kThreads), 0, s, o, c, p, implicit_plaintext, plaintext_batch_count, limbs, pv, support, n); break;
    case kContiguous: hipLaunchKernelGGL(k_add_plain<kContiguous>, grid, dim3(kThreads), 0, s, o, c, p, implicit_plaintext, plaintext_batch_count, limbs, pv, support, n); break;
    case kStrided: hipLaunchKernelGGL(k_add_plain<kStrided>, grid, dim3(kThreads), 0, s, o, c, p, implicit_plaintext, plaintext_batch_count, limbs, pv, support, n); break;
    default: set_error("unknown plaintext layout %d", layout); return kInvalidArgument;
  }
  FH_CHECK_LAUNCH("ckks plaintext add");
  return kOk;
}

FH_API int fh_ckks_tensor_product(fh_view3 d0, fh_view3 d1, fh_view3 d2, fh_view3 a0, fh_view3 a1,
                                  fh_view3 b0, fh_view3 b1, fh_table2 rns_params, int64_t batch,
                                  int64_t limbs, int64_t n, void* stream) {
  int st = check_grid(batch, limbs, n);
  if (st != kOk) return st;
  if (batch == 0 || limbs == 0 || n == 0) return kOk;
  FH_CHECK_ARG(d0.data && d1.data && d2.data && a0.data && a1.data && b0.data && b1.data &&
                   rns_params.data,
               "tensor product: null operand");
  ParamView pv{rns_params.data, rns_params.stride_row, rns_params.stride_col};
  dim3 grid(ceil_div(n, kChunk), limbs, batch);
  hipLaunchKernelGGL(k_tensor_product, grid, dim3(kThreads), 0, static_cast<hipStream_t>(stream),
                     View3(d0), View3(d1), View3(d2), View3(a0), View3(a1), View3(b0), View3(b1), pv, n);
  FH_CHECK_LAUNCH("ckks tensor product");
  return kOk;
}

FH_API int fh_ckks_keyswitch_accumulate_multi_key(fh_view3 acc0, fh_view3 acc1, fh_view3 digit,
                                                  const int64_t* key_bases, int64_t digit_index,
                                                  int64_t key_stride_digit, int64_t key_stride_comp,
                                                  int64_t key_stride_row, fh_table2 rns_params,
                                                  int64_t key_row_start, int64_t batch,
                                                  int64_t limbs, int64_t n, void* stream) {
  int st = check_grid(batch, limbs, n);
  if (st != kOk) return st;
  if (batch == 0 || limbs == 0 || n == 0) return kOk;
  FH_CHECK_ARG(acc0.data && acc1.data && digit.data && key_bases && rns_params.data,
               "multi-key accumulate: null operand");
  FH_CHECK_ARG(key_row_start >= 0 && digit_index >= 0, "negative key row or digit index");
  ParamView pv{rns_params.data, rns_params.stride_row, rns_params.stride_col};
  dim3 grid(ceil_div(n, kChunk), limbs, batch);
  hipLaunchKernelGGL(k_keyswitch_mac_multi, grid, dim3(kThreads), 0, static_cast<hipStream_t>(stream),
                     View3(acc0), View3(acc1), View3(digit), key_bases, digit_index, key_stride_digit,
                     key_stride_comp, key_stride_row, pv, key_row_start, n);
  FH_CHECK_LAUNCH("ckks multi-key keyswitch accumulate");
  return kOk;
}
