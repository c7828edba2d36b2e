// Element-wise RNS arithmetic, mixed-radix decomposition and basis extension
// for gfx950.  Each kernel reproduces the reference CPU loop body
// (csrc/ops/rns/cpu/rns_arithmetic_cpu.cpp, rns_canonical_arithmetic_cpu.cpp)
// value for value; only the iteration order is GPU-shaped.
//
// Launch shape of the element-wise kernels: grid (ceil(N / 1024), limb,
// batch), 256 threads, four coefficients per thread strided by 256 so every
// load instruction of a wavefront touches 512 contiguous bytes.  These
// kernels are HBM-bound streaming passes.
#include "fh_common.h"
#include "fh_views.h"

namespace fh {
namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 4;
constexpr int kChunk = kThreads * kPerThread;

// Binary operations (rns_arithmetic_cpu.cpp:61-127, canonical :19-92).
enum BinaryOp : int {
  kMul = 0,
  kAddLazy = 1,
  kSubLazy = 2,
  kAddCanonical = 3,
  kSubCanonical = 4,
  kMulCyclic = 5,
  kMulContiguous = 6,
};

// Unary in-place operations (rns_arithmetic_cpu.cpp:148-192).
enum UnaryOp : int {
  kToMontgomery = 0,
  kFromMontgomery = 1,
  kCanonicalize = 2,
  kCenter = 3,
  kShiftPositive = 4,
};

__device__ __forceinline__ int64_t canonical_add(int64_t a, int64_t b, int64_t twice_q) {
  int64_t v = a + b;
  if (v >= twice_q) v -= twice_q;
  const int64_t q = twice_q >> 1;
  return v < q ? v : v - q;
}

__device__ __forceinline__ int64_t canonical_sub(int64_t a, int64_t b, int64_t twice_q) {
  int64_t v = a - b;
  if (v < 0) {
    v += twice_q;
  } else if (v >= twice_q) {
    v -= twice_q;
  }
  const int64_t q = twice_q >> 1;
  return v < q ? v : v - q;
}

template <int OP>
__global__ __launch_bounds__(kThreads) void k_binary(View3 out, View3 lhs, View3 rhs,
                                                     ParamView params, int64_t n,
                                                     int64_t support) {
  const int64_t l = blockIdx.y;
  const int64_t b = blockIdx.z;
  const Mont m = load_mont(params, l);
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kChunk + threadIdx.x;
#pragma unroll
  for (int e = 0; e < kPerThread; ++e) {
    const int64_t i = base + e * kThreads;
    if (i >= n) break;
    const int64_t a = lhs.at(b, l, i);
    int64_t r;
    if constexpr (OP == kMulCyclic) {
      r = mont_mul(a, rhs.at(b, l, i % support), m);
    } else if constexpr (OP == kMulContiguous) {
      r = mont_mul(a, rhs.at(b, l, i / (n / support)), m);
    } else {
      const int64_t c = rhs.at(b, l, i);
      if constexpr (OP == kMul) {
        r = mont_mul(a, c, m);
      } else if constexpr (OP == kAddLazy) {
        r = add_lazy(a, c, m.twice_q);
      } else if constexpr (OP == kSubLazy) {
        r = sub_lazy(a, c, m.twice_q);
      } else if constexpr (OP == kAddCanonical) {
        r = canonical_add(a, c, m.twice_q);
      } else {
        r = canonical_sub(a, c, m.twice_q);
      }
    }
    out.at(b, l, i) = r;
  }
}

template <int OP>
__global__ __launch_bounds__(kThreads) void k_unary(View3 x, ParamView params, int64_t n) {
  const int64_t l = blockIdx.y;
  const int64_t b = blockIdx.z;
  const Mont m = load_mont(params, l);
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kChunk + threadIdx.x;
#pragma unroll
  for (int e = 0; e < kPerThread; ++e) {
    const int64_t i = base + e * kThreads;
    if (i >= n) break;
    int64_t v = x.at(b, l, i);
    if constexpr (OP == kToMontgomery) {
      v = mont_mul(v, m.r2, m);
    } else if constexpr (OP == kFromMontgomery) {
      v = mont_mul(v, 1, m);
    } else if constexpr (OP == kCanonicalize) {
      v = canonicalize(v, m.twice_q);
    } else if constexpr (OP == kCenter) {
      v = center(v, m.twice_q);
    } else {
      v = shift_positive(v, m.twice_q);
    }
    x.at(b, l, i) = v;
  }
}

// montgomery_mul_row_scalars_ (rns_arithmetic_cpu.cpp:319-369) and the
// canonical functional form (rns_canonical_arithmetic_cpu.cpp:170-240).
template <bool kCanonical>
__global__ __launch_bounds__(kThreads) void k_row_scalars(View3 out, View3 in, Vec scalars,
                                                          ParamView params, int64_t n) {
  const int64_t l = blockIdx.y;
  const int64_t b = blockIdx.z;
  const Mont m = load_mont(params, l);
  const int64_t s = scalars[l];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kChunk + threadIdx.x;
#pragma unroll
  for (int e = 0; e < kPerThread; ++e) {
    const int64_t i = base + e * kThreads;
    if (i >= n) break;
    int64_t v = mont_mul(in.at(b, l, i), s, m);
    if constexpr (kCanonical) v = canonicalize(v, m.twice_q);
    out.at(b, l, i) = v;
  }
}

// add_lazy_with_twice_modulus (rns_arithmetic_cpu.cpp:371-432).
__global__ __launch_bounds__(kThreads) void k_add_lazy_twice(View3 out, View3 lhs, View3 rhs,
                                                             Vec twice, int64_t n) {
  const int64_t l = blockIdx.y;
  const int64_t b = blockIdx.z;
  const int64_t t2 = twice[l];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kChunk + threadIdx.x;
#pragma unroll
  for (int e = 0; e < kPerThread; ++e) {
    const int64_t i = base + e * kThreads;
    if (i >= n) break;
    out.at(b, l, i) = add_lazy(lhs.at(b, l, i), rhs.at(b, l, i), t2);
  }
}

// lift_centered_coefficients (rns_arithmetic_cpu.cpp:434-480): the source is
// a [batch, coefficient] view passed with a zero limb stride.
__global__ __launch_bounds__(kThreads) void k_lift(View3 out, View3 in, Vec twice, int64_t n) {
  const int64_t l = blockIdx.y;
  const int64_t b = blockIdx.z;
  const int64_t t2 = twice[l];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kChunk + threadIdx.x;
#pragma unroll
  for (int e = 0; e < kPerThread; ++e) {
    const int64_t i = base + e * kThreads;
    if (i >= n) break;
    out.at(b, l, i) = shift_positive(in.at(b, 0, i), t2);
  }
}

// mixed_radix_decompose (rns_arithmetic_cpu.cpp:482-614): Garner digits of
// one hybrid key-switch digit, one coefficient per thread.
struct MixedRadixTables {
  Vec normalizers;         // [D-1]
  const int64_t* propagation;  // [D-1, D]
  int64_t prop_row_stride;
  int64_t prop_col_stride;
  Vec modulus_lo, modulus_hi, neg_inv_lo, neg_inv_hi;  // [D]
};

__device__ __forceinline__ Mont split_mont(const MixedRadixTables& t, int64_t r) {
  Mont m;
  m.q = static_cast<uint64_t>(t.modulus_lo[r]) +
        (static_cast<uint64_t>(t.modulus_hi[r]) << kHalfRadixBits);
  m.k = static_cast<uint64_t>(t.neg_inv_lo[r]) +
        (static_cast<uint64_t>(t.neg_inv_hi[r]) << kHalfRadixBits);
  m.twice_q = 0;
  m.r2 = 0;
  m.ninv = 0;
  return m;
}

constexpr int kMaxMixedRadixRows = 8;

__global__ __launch_bounds__(kThreads) void k_mixed_radix_decompose(
    View3 out, View3 src, MixedRadixTables t, int rows, int64_t n) {
  const int64_t b = blockIdx.y;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  int64_t digits[kMaxMixedRadixRows];
  const Mont m0 = split_mont(t, 0);
  int64_t first = src.at(b, 0, i);
  first = first < static_cast<int64_t>(m0.q) ? first
                                             : first - static_cast<int64_t>(m0.q);
#pragma unroll
  for (int r = 0; r < kMaxMixedRadixRows; ++r) digits[r] = first;
#pragma unroll
  for (int step = 0; step < kMaxMixedRadixRows - 1; ++step) {
    const int row = step + 1;
    if (row >= rows) break;
    const Mont mr = split_mont(t, row);
    const int64_t digit =
        mont_mul_split(src.at(b, row, i) - digits[row], t.normalizers[step], mr);
    digits[row] = digit;
#pragma unroll
    for (int target = 0; target < kMaxMixedRadixRows; ++target) {
      if (target <= row || target >= rows) continue;
      digits[target] += mont_mul(
          digit,
          t.propagation[step * t.prop_row_stride + target * t.prop_col_stride],
          split_mont(t, target));
    }
  }
#pragma unroll
  for (int r = 0; r < kMaxMixedRadixRows; ++r) {
    if (r < rows) out.at(b, r, i) = digits[r];
  }
}

// mixed_radix_basis_extend_to_montgomery (rns_arithmetic_cpu.cpp:616-700):
// out_j = REDC(d_0 R^2) (+lazy) sum_{r>=1} REDC(d_r * (M_r R^2 mod p_j)).
__global__ __launch_bounds__(kThreads) void k_basis_extend(
    View3 out, View3 digits, const int64_t* coef, int64_t coef_digit_stride,
    int64_t coef_row_stride, ParamView params, int digit_count, int64_t n) {
  const int64_t j = blockIdx.y;
  const int64_t b = blockIdx.z;
  const Mont m = load_mont(params, j);
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kChunk + threadIdx.x;
#pragma unroll
  for (int e = 0; e < kPerThread; ++e) {
    const int64_t i = base + e * kThreads;
    if (i >= n) break;
    int64_t v = mont_mul(digits.at(b, 0, i), m.r2, m);
    for (int d = 1; d < digit_count; ++d) {
      v = add_lazy(v,
                   mont_mul(digits.at(b, d, i),
                            coef[(d - 1) * coef_digit_stride + j * coef_row_stride], m),
                   m.twice_q);
    }
    out.at(b, j, i) = v;
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

FH_API int fh_rns_binary(int32_t op, fh_view3 out, fh_view3 lhs, fh_view3 rhs,
                         fh_table2 rns_params, int64_t batch, int64_t limbs,
                         int64_t n, int64_t compressed_support, void* stream) {
  int st = check_grid(batch, limbs, n);
  if (st != kOk) return st;
  if (batch == 0 || limbs == 0 || n == 0) return kOk;
  FH_CHECK_ARG(out.data && lhs.data && rhs.data && rns_params.data,
               "rns binary: null operand");
  if (op == kMulCyclic || op == kMulContiguous) {
    FH_CHECK_ARG(compressed_support > 0 && compressed_support < n &&
                     (compressed_support & (compressed_support - 1)) == 0 &&
                     n % compressed_support == 0,
                 "compressed support must be a power-of-two divisor of N");
  }
  ParamView pv{rns_params.data, rns_params.stride_row, rns_params.stride_col};
  dim3 grid(ceil_div(n, kChunk), limbs, batch);
  hipStream_t s = static_cast<hipStream_t>(stream);
  View3 o(out), a(lhs), c(rhs);
  switch (op) {
    case kMul: hipLaunchKernelGGL(k_binary<kMul>, grid, dim3(kThreads), 0, s, o, a, c, pv, n, compressed_support); break;
    case kAddLazy: hipLaunchKernelGGL(k_binary<kAddLazy>, grid, dim3(kThreads), 0, s, o, a, c, pv, n, compressed_support); break;
    case kSubLazy: hipLaunchKernelGGL(k_binary<kSubLazy>, grid, dim3(kThreads), 0, s, o, a, c, pv, n, compressed_support); break;
    case kAddCanonical: hipLaunchKernelGGL(k_binary<kAddCanonical>, grid, dim3(kThreads), 0, s, o, a, c, pv, n, compressed_support); break;
    case kSubCanonical: hipLaunchKernelGGL(k_binary<kSubCanonical>, grid, dim3(kThreads), 0, s, o, a, c, pv, n, compressed_support); break;
    case kMulCyclic: hipLaunchKernelGGL(k_binary<kMulCyclic>, grid, dim3(kThreads), 0, s, o, a, c, pv, n, compressed_support); break;
    case kMulContiguous: hipLaunchKernelGGL(k_binary<kMulContiguous>, grid, dim3(kThreads), 0, s, o, a, c, pv, n, compressed_support); break;
    default: set_error("unknown rns binary op %d", op); return kInvalidArgument;
  }
  FH_CHECK_LAUNCH("rns binary");
  return kOk;
}

FH_API int fh_rns_unary_inplace(int32_t op, fh_view3 residues, fh_table2 rns_params,
                                int64_t batch, int64_t limbs, int64_t n, void* stream) {
  int st = check_grid(batch, limbs, n);
  if (st != kOk) return st;
  if (batch == 0 || limbs == 0 || n == 0) return kOk;
  FH_CHECK_ARG(residues.data && rns_params.data, "rns unary: null operand");
  ParamView pv{rns_params.data, rns_params.stride_row, rns_params.stride_col};
  dim3 grid(ceil_div(n, kChunk), limbs, batch);
  hipStream_t s = static_cast<hipStream_t>(stream);
  View3 x(residues);
  switch (op) {
    case kToMontgomery: hipLaunchKernelGGL(k_unary<kToMontgomery>, grid, dim3(kThreads), 0, s, x, pv, n); break;
    case kFromMontgomery: hipLaunchKernelGGL(k_unary<kFromMontgomery>, grid, dim3(kThreads), 0, s, x, pv, n); break;
    case kCanonicalize: hipLaunchKernelGGL(k_unary<kCanonicalize>, grid, dim3(kThreads), 0, s, x, pv, n); break;
    case kCenter: hipLaunchKernelGGL(k_unary<kCenter>, grid, dim3(kThreads), 0, s, x, pv, n); break;
    case kShiftPositive: hipLaunchKernelGGL(k_unary<kShiftPositive>, grid, dim3(kThreads), 0, s, x, pv, n); break;
    default: set_error("unknown rns unary op %d", op); return kInvalidArgument;
  }
  FH_CHECK_LAUNCH("rns unary");
  return kOk;
}

FH_API int fh_rns_row_scalars(fh_view3 out, fh_view3 in, fh_vec scalars,
                              fh_table2 rns_params, int64_t batch, int64_t limbs,
                              int64_t n, int32_t canonical, void* stream) {
  int st = check_grid(batch, limbs, n);
  if (st != kOk) return st;
  if (batch == 0 || limbs == 0 || n == 0) return kOk;
  FH_CHECK_ARG(out.data && in.data && scalars.data && rns_params.data,
               "row scalars: null operand");
  ParamView pv{rns_params.data, rns_params.stride_row, rns_params.stride_col};
  dim3 grid(ceil_div(n, kChunk), limbs, batch);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (canonical) {
    hipLaunchKernelGGL(k_row_scalars<true>, grid, dim3(kThreads), 0, s, View3(out), View3(in), Vec(scalars), pv, n);
  } else {
    hipLaunchKernelGGL(k_row_scalars<false>, grid, dim3(kThreads), 0, s, View3(out), View3(in), Vec(scalars), pv, n);
  }
  FH_CHECK_LAUNCH("rns row scalars");
  return kOk;
}

FH_API int fh_rns_add_lazy_twice_modulus(fh_view3 out, fh_view3 lhs, fh_view3 rhs,
                                         fh_vec twice_modulus, int64_t batch,
                                         int64_t limbs, int64_t n, void* stream) {
  int st = check_grid(batch, limbs, n);
  if (st != kOk) return st;
  if (batch == 0 || limbs == 0 || n == 0) return kOk;
  FH_CHECK_ARG(out.data && lhs.data && rhs.data && twice_modulus.data,
               "add_lazy_with_twice_modulus: null operand");
  dim3 grid(ceil_div(n, kChunk), limbs, batch);
  hipLaunchKernelGGL(k_add_lazy_twice, grid, dim3(kThreads), 0, static_cast<hipStream_t>(stream),
                     View3(out), View3(lhs), View3(rhs), Vec(twice_modulus), n);
  FH_CHECK_LAUNCH("rns add_lazy_with_twice_modulus");
  return kOk;
}

FH_API int fh_rns_lift_centered(fh_view3 out, fh_view3 coefficients, fh_vec twice_modulus,
                                int64_t batch, int64_t limbs, int64_t n, void* stream) {
  int st = check_grid(batch, limbs, n);
  if (st != kOk) return st;
  if (batch == 0 || limbs == 0 || n == 0) return kOk;
  FH_CHECK_ARG(out.data && coefficients.data && twice_modulus.data,
               "lift_centered_coefficients: null operand");
  dim3 grid(ceil_div(n, kChunk), limbs, batch);
  hipLaunchKernelGGL(k_lift, grid, dim3(kThreads), 0, static_cast<hipStream_t>(stream),
                     View3(out), View3(coefficients), Vec(twice_modulus), n);
  FH_CHECK_LAUNCH("rns lift_centered_coefficients");
  return kOk;
}

FH_API int fh_rns_mixed_radix_decompose(fh_view3 out, fh_view3 source, int64_t batch,
                                        int64_t rows, int64_t n, fh_vec normalizers,
                                        fh_table2 propagation, fh_vec modulus_lo,
                                        fh_vec modulus_hi, fh_vec neg_inv_lo,
                                        fh_vec neg_inv_hi, void* stream) {
  int st = check_grid(batch, rows, n);
  if (st != kOk) return st;
  FH_CHECK_ARG(rows >= 1 && rows <= kMaxMixedRadixRows,
               "mixed_radix_decompose supports 1..8 rows, got %lld", (long long)rows);
  if (batch == 0 || n == 0) return kOk;
  FH_CHECK_ARG(out.data && source.data && modulus_lo.data && modulus_hi.data &&
                   neg_inv_lo.data && neg_inv_hi.data,
               "mixed_radix_decompose: null operand");
  FH_CHECK_ARG(rows == 1 || (normalizers.data && propagation.data),
               "mixed_radix_decompose: null tables");
  MixedRadixTables t{Vec(normalizers), propagation.data, propagation.stride_row,
                     propagation.stride_col, Vec(modulus_lo), Vec(modulus_hi),
                     Vec(neg_inv_lo), Vec(neg_inv_hi)};
  dim3 grid(ceil_div(n, kThreads), batch);
  hipLaunchKernelGGL(k_mixed_radix_decompose, grid, dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), View3(out), View3(source), t,
                     static_cast<int>(rows), n);
  FH_CHECK_LAUNCH("rns mixed_radix_decompose");
  return kOk;
}

FH_API int fh_rns_basis_extend(fh_view3 out, fh_view3 digits, int64_t batch,
                               int64_t digit_count, int64_t destination_rows, int64_t n,
                               fh_table2 coefficients, fh_table2 rns_params, void* stream) {
  int st = check_grid(batch, destination_rows, n);
  if (st != kOk) return st;
  FH_CHECK_ARG(digit_count >= 1, "basis extension needs at least one digit");
  if (batch == 0 || destination_rows == 0 || n == 0) return kOk;
  FH_CHECK_ARG(out.data && digits.data && rns_params.data,
               "basis extension: null operand");
  FH_CHECK_ARG(digit_count == 1 || coefficients.data, "basis extension: null table");
  ParamView pv{rns_params.data, rns_params.stride_row, rns_params.stride_col};
  dim3 grid(ceil_div(n, kChunk), destination_rows, batch);
  hipLaunchKernelGGL(k_basis_extend, grid, dim3(kThreads), 0, static_cast<hipStream_t>(stream),
                     View3(out), View3(digits), coefficients.data, coefficients.stride_row,
                     coefficients.stride_col, pv, static_cast<int>(digit_count), n);
  FH_CHECK_LAUNCH("rns basis extension");
  return kOk;
}
