// Basis-extension products on the matrix cores: the byte-sliced INT8 MFMA
// trial.
//
// mixed_radix_basis_extend_to_montgomery (rns_arithmetic_cpu.cpp:616-700)
// computes out_j = REDC(d_0 R^2) (+lazy) sum_{r>=1} REDC(d_r C_rj) with
// C_rj = M_r R^2 mod p_j: per coefficient, D <= 4 products of 60-bit
// numbers per destination row -- the path's only contraction.  Here the
// 128-bit products d_r C_rj come from v_mfma_i32_32x32x16_i8:
//   * every 60-bit operand is written in balanced radix 256,
//     x = sum_{a<8} s_a 256^a with s_a in [-128, 128) (signed int8, which is
//     what the i8 matrix instruction multiplies);
//   * A (32 coefficients x K=16) holds the 8 bytes of two digits,
//     B (K=16 x 32 columns) is block-diagonal: column (r, k) for k = 0..14
//     holds byte k-a of C_rj against byte a of digit r, so
//     D[m][(r, k)] = sum_{a+b=k} s_a(d_r) s_b(C_rj) exactly in int32
//     (|D| < 8 * 128^2);
//   * the wave then sums sum_k D_k 256^k (exact, in 128 bits) and applies
//     the same REDC and lazy additions as the scalar kernel, so the output
//     is bit-identical to k_basis_extend.
// The matrix instruction replaces only the 64x64-bit multiplications; the
// byte-sum reconstruction and REDC stay on the vector ALUs.  The entry point
// exists to measure that trade against the scalar kernel.
#include "fh_common.h"
#include "fh_views.h"

namespace fh {
namespace {

typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kTrialWaves = 4;             // 32-coefficient tiles per block
constexpr int kTrialBlock = 64 * kTrialWaves;
constexpr int kTrialPitch = 33;            // int32 words per LDS row (32 + pad)

// Balanced radix-256 digits of x < 2^60 as eight packed int8.
__device__ __forceinline__ uint64_t balanced_bytes(uint64_t x) {
  uint64_t out = 0;
  int carry = 0;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    int v = static_cast<int>((x >> (8 * a)) & 255u) + carry;
    carry = v >= 128 ? 1 : 0;
    v -= carry << 8;
    out |= static_cast<uint64_t>(static_cast<uint8_t>(static_cast<int8_t>(v))) << (8 * a);
  }
  return out;
}

// B operand of lane l for destination row j and digit pair p: column
// n = l % 32 is (r_local = n / 16, k = n % 16), the lane's K block is
// digit r_local' = l / 32; nonzero only on the diagonal blocks.
__device__ __forceinline__ uint64_t b_operand(int lane, int p, int digit_count, int64_t c0, const int64_t* coef,
                                              int64_t coef_digit_stride, int64_t coef_row_stride, int64_t j) {
  const int n = lane & 31;
  const int kb = lane >> 5;
  const int k = n & 15;
  const int r = 2 * p + kb;
  if ((n >> 4) != kb || k > 14 || r >= digit_count) return 0;
  const uint64_t c = static_cast<uint64_t>(r == 0 ? c0 : coef[(r - 1) * coef_digit_stride + j * coef_row_stride]);
  const uint64_t s = balanced_bytes(c);
  uint64_t out = 0;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int t = k - a;
    if (t >= 0 && t <= 7) out |= ((s >> (8 * t)) & 255u) << (8 * a);
  }
  return out;
}

__global__ __launch_bounds__(kTrialBlock) void k_basis_extend_mfma(View3 out, View3 digits, const int64_t* coef,
                                                                   int64_t coef_digit_stride, int64_t coef_row_stride,
                                                                   ParamView params, int digit_count, int64_t rows) {
  __shared__ int32_t prod[kTrialWaves][2][32 * kTrialPitch];
  __shared__ int64_t terms[kTrialWaves][4][32];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t b = blockIdx.y;
  const int64_t i0 = (static_cast<int64_t>(blockIdx.x) * kTrialWaves + wave) * 32;
  const int m = lane & 31;
  const int half = lane >> 5;
  const int pairs = (digit_count + 1) / 2;
  // A operands (one per digit pair): lane holds coefficient i0 + m of digit 2p + half
  uint64_t a_op[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const int r = 2 * p + half;
    a_op[p] = r < digit_count ? balanced_bytes(static_cast<uint64_t>(digits.at(b, r, i0 + m))) : 0;
  }
  for (int64_t j = 0; j < rows; ++j) {
    const Mont mt = load_mont(params, j);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      if (p >= pairs) break;
      const uint64_t b_op = b_operand(lane, p, digit_count, mt.r2, coef, coef_digit_stride, coef_row_stride, j);
      v16i acc = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
      acc = __builtin_amdgcn_mfma_i32_32x32x16_i8(static_cast<long>(a_op[p]), static_cast<long>(b_op), acc, 0, 0, 0);
      // lane holds D[row][col] for col = lane % 32, row = 8 (v / 4) + 4 (lane / 32) + v % 4
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int row = 8 * (v >> 2) + 4 * half + (v & 3);
        prod[wave][p][row * kTrialPitch + (lane & 31)] = acc[v];
      }
    }
    __syncthreads();
    // lane (m, half): the product of digit 2p + half and C for coefficient m
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int r = 2 * p + half;
      if (p < pairs && r < digit_count) {
        const int32_t* row = &prod[wave][p][m * kTrialPitch + 16 * half];
        __int128 v = 0;
#pragma unroll
        for (int k = 0; k < 15; ++k) v += static_cast<__int128>(row[k]) << (8 * k);
        terms[wave][r][m] = mont_redc128(static_cast<uint64_t>(v), static_cast<uint64_t>(v >> 64), mt.q, mt.k);
      }
    }
    __syncthreads();
    if (half == 0) {
      int64_t v = terms[wave][0][m];
      for (int r = 1; r < digit_count; ++r) v = add_lazy(v, terms[wave][r][m], mt.twice_q);
      out.at(b, j, i0 + m) = v;
    }
    __syncthreads();  // prod and terms are rewritten for the next row
  }
}

}  // namespace
}  // namespace fh

using namespace fh;

// See include/fhelium_amd.h.
FH_API int fh_rns_basis_extend_mfma(fh_view3 out, fh_view3 digits, int64_t batch, int64_t digit_count,
                                    int64_t destination_rows, int64_t n, fh_table2 coefficients,
                                    fh_table2 rns_params, void* stream) {
  FH_CHECK_ARG(batch >= 0 && batch <= 65535 && destination_rows >= 0, "bad batch or row count");
  FH_CHECK_ARG(digit_count >= 1 && digit_count <= 4, "the MFMA trial handles 1..4 digits");
  if (batch == 0 || destination_rows == 0 || n == 0) return kOk;
  FH_CHECK_ARG(n % (32 * kTrialWaves) == 0, "the MFMA trial needs N to be a multiple of 128");
  FH_CHECK_ARG(out.data && digits.data && rns_params.data && (digit_count == 1 || coefficients.data),
               "basis extension (MFMA): null operand");
  ParamView pv{rns_params.data, rns_params.stride_row, rns_params.stride_col};
  const dim3 grid(static_cast<uint32_t>(n / (32 * kTrialWaves)), static_cast<uint32_t>(batch));
  hipLaunchKernelGGL(k_basis_extend_mfma, grid, dim3(kTrialBlock), 0, static_cast<hipStream_t>(stream), View3(out),
                     View3(digits), coefficients.data, coefficients.stride_row, coefficients.stride_col, pv,
                     static_cast<int>(digit_count), destination_rows);
  FH_CHECK_LAUNCH("rns basis extension (MFMA trial)");
  return kOk;
}
