// CKKS slot encode and decode on gfx950 (binary64).
//
// Slot order as fhelium_amd/codec.py and the reference's generator-3
// convention (fhelium/engine/slot_embedding.py, galois.py:10-26): slot j is
// the evaluation at zeta^{e_j}, zeta = e^{i pi / N}, e_j = 3^{-j} mod 2N, and
// pos_j = (e_j - 1) / 2 is its index among the N odd exponents.
//
//   encode:  z[pos_j] = m_j, z[N-1-pos_j] = conj(m_j)      (every index once)
//            a = FFT_N(z) / N * conj(twist), twist_k = zeta^k
//            coefficient_k = rint(Re(a_k) * scale), residues mod each prime
//   decode:  x_k = centered CRT of the residues over two primes (Garner)
//            z = sum_k x_k twist_k e^{+2 pi i j k / N}    (an FFT with + sign)
//            m_j = z[pos_j] / scale
//
// The FFT is iterative radix-2 in place: the first kernel writes its input
// in bit-reversed order, then one launch per stage (butterflies of half size
// h = 2^s, twiddles from sincospi).  Encode and decode are outside the
// evaluated hot path; the point is that they run next to the data.
#include "fh_common.h"

namespace fh {
namespace {

constexpr int kCodecBlock = 256;

__device__ __forceinline__ uint32_t bitrev(uint32_t x, int bits) { return __brev(x) >> (32 - bits); }

// z: [B, N] complex as (re, im) pairs.
__global__ __launch_bounds__(kCodecBlock) void k_encode_scatter(const double* slots, int64_t s, const int32_t* pos,
                                                               int log_n, double2* z) {
  const int64_t b = blockIdx.y;
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kCodecBlock + threadIdx.x;
  if (j >= s) return;
  const int64_t n = int64_t{1} << log_n;
  const double re = slots[(b * s + j) * 2], im = slots[(b * s + j) * 2 + 1];
  const uint32_t p = static_cast<uint32_t>(pos[j]);
  z[b * n + bitrev(p, log_n)] = make_double2(re, im);
  z[b * n + bitrev(static_cast<uint32_t>(n - 1) - p, log_n)] = make_double2(re, -im);
}

// One radix-2 stage of half size h on bit-reversed-input data; sign -1 is
// the forward DFT (e^{-2 pi i jk/N}), +1 the positive-exponent sum.
__global__ __launch_bounds__(kCodecBlock) void k_fft_stage(double2* z, int log_n, int64_t h, double sign) {
  const int64_t b = blockIdx.y;
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kCodecBlock + threadIdx.x;
  const int64_t n = int64_t{1} << log_n;
  if (t >= n / 2) return;
  const int64_t k = t & (h - 1);
  const int64_t i0 = ((t - k) << 1) + k;
  const int64_t i1 = i0 + h;
  double sn, cs;
  sincospi(sign * static_cast<double>(k) / static_cast<double>(h), &sn, &cs);
  double2* row = z + b * n;
  const double2 u = row[i0], v = row[i1];
  const double vr = v.x * cs - v.y * sn, vi = v.x * sn + v.y * cs;
  row[i0] = make_double2(u.x + vr, u.y + vi);
  row[i1] = make_double2(u.x - vr, u.y - vi);
}

// coefficient_k = rint(Re(z_k conj(twist_k)) / N * scale), then its residue
// for each prime: out [B, L, N] canonical standard.
__global__ __launch_bounds__(kCodecBlock) void k_encode_finish(const double2* z, int log_n, double scale,
                                                              const int64_t* moduli, int64_t limbs, int64_t* out) {
  const int64_t b = blockIdx.y;
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kCodecBlock + threadIdx.x;
  const int64_t n = int64_t{1} << log_n;
  if (k >= n) return;
  const double2 v = z[b * n + k];
  double sn, cs;
  sincospi(static_cast<double>(k) / static_cast<double>(n), &sn, &cs);
  const double re = (v.x * cs + v.y * sn) / static_cast<double>(n);  // Re(v * conj(twist))
  const int64_t c = static_cast<int64_t>(rint(re * scale));
  for (int64_t l = 0; l < limbs; ++l) {
    const int64_t q = moduli[l];
    int64_t r = c % q;
    if (r < 0) r += q;
    out[(b * limbs + l) * n + k] = r;
  }
}

// x_k centered over q_a q_b by Garner, times twist_k, written bit-reversed.
__global__ __launch_bounds__(kCodecBlock) void k_decode_crt(const int64_t* ra, const int64_t* rb, int64_t batch_stride,
                                                           int64_t qa, int64_t qb, int64_t qa_inv_b, int log_n,
                                                           double2* z) {
  const int64_t b = blockIdx.y;
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kCodecBlock + threadIdx.x;
  const int64_t n = int64_t{1} << log_n;
  if (k >= n) return;
  const uint64_t a = static_cast<uint64_t>(ra[b * batch_stride + k]);
  const uint64_t bb = static_cast<uint64_t>(rb[b * batch_stride + k]);
  // t = (b - a) qa^{-1} mod qb; x = a + qa t in [0, qa qb)
  const uint64_t qbu = static_cast<uint64_t>(qb);
  const uint64_t diff = (bb + qbu - a % qbu) % qbu;
  const unsigned __int128 prod = static_cast<unsigned __int128>(diff) * static_cast<uint64_t>(qa_inv_b);
  int64_t t = static_cast<int64_t>(static_cast<uint64_t>(prod % qbu));
  // centre: x > qa qb / 2 exactly when t > (qb - 1) / 2 (a < qa; odd qb)
  if (t > (qb - 1) / 2) t -= qb;
  // |x| < 2^62: one rounding from the exact int64 value; larger: binary64.
  const double x = (t >= -3 && t <= 3)
                       ? static_cast<double>(static_cast<int64_t>(a) + qa * t)
                       : static_cast<double>(static_cast<int64_t>(a)) + static_cast<double>(qa) * static_cast<double>(t);
  double sn, cs;
  sincospi(static_cast<double>(k) / static_cast<double>(n), &sn, &cs);
  z[b * n + bitrev(static_cast<uint32_t>(k), log_n)] = make_double2(x * cs, x * sn);
}

__global__ __launch_bounds__(kCodecBlock) void k_decode_gather(const double2* z, const int32_t* pos, int64_t s,
                                                              int log_n, double inv_scale, double* slots) {
  const int64_t b = blockIdx.y;
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kCodecBlock + threadIdx.x;
  if (j >= s) return;
  const int64_t n = int64_t{1} << log_n;
  const double2 v = z[b * n + pos[j]];
  slots[(b * s + j) * 2] = v.x * inv_scale;
  slots[(b * s + j) * 2 + 1] = v.y * inv_scale;
}

int fft_stages(double2* z, int64_t batch, int log_n, double sign, hipStream_t st) {
  const int64_t n = int64_t{1} << log_n;
  const dim3 grid(static_cast<uint32_t>(ceil_div(n / 2, kCodecBlock)), static_cast<uint32_t>(batch));
  for (int s = 0; s < log_n; ++s) {
    hipLaunchKernelGGL(k_fft_stage, grid, dim3(kCodecBlock), 0, st, z, log_n, int64_t{1} << s, sign);
    FH_CHECK_LAUNCH("codec fft stage");
  }
  return kOk;
}

}  // namespace
}  // namespace fh

using namespace fh;

FH_API int fh_ckks_encode(const double* slots, int64_t batch, int64_t log_n, const int32_t* slot_positions,
                          double scale, const int64_t* moduli, int64_t limbs, double* workspace, int64_t* out,
                          void* stream) {
  FH_CHECK_ARG(log_n >= 2 && log_n <= 17, "encode supports logN 2..17");
  FH_CHECK_ARG(batch >= 1 && batch <= 65535 && limbs >= 1, "bad encode batch or limb count");
  FH_CHECK_ARG(scale > 0.0 && scale < 0x1.0p62, "encode scale must be positive and below 2^62");
  FH_CHECK_ARG(slots && slot_positions && moduli && workspace && out, "encode: null operand");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t n = int64_t{1} << log_n;
  const int64_t s = n / 2;
  double2* z = reinterpret_cast<double2*>(workspace);
  hipLaunchKernelGGL(k_encode_scatter, dim3(static_cast<uint32_t>(ceil_div(s, kCodecBlock)), static_cast<uint32_t>(batch)),
                     dim3(kCodecBlock), 0, st, slots, s, slot_positions, static_cast<int>(log_n), z);
  FH_CHECK_LAUNCH("encode scatter");
  int rc = fft_stages(z, batch, static_cast<int>(log_n), -1.0, st);
  if (rc != kOk) return rc;
  hipLaunchKernelGGL(k_encode_finish, dim3(static_cast<uint32_t>(ceil_div(n, kCodecBlock)), static_cast<uint32_t>(batch)),
                     dim3(kCodecBlock), 0, st, z, static_cast<int>(log_n), scale, moduli, limbs, out);
  FH_CHECK_LAUNCH("encode finish");
  return kOk;
}

FH_API int fh_ckks_decode(const int64_t* residues_a, const int64_t* residues_b, int64_t batch_stride, int64_t batch,
                          int64_t log_n, int64_t modulus_a, int64_t modulus_b, int64_t modulus_a_inverse_mod_b,
                          const int32_t* slot_positions, double scale, double* workspace, double* slots,
                          void* stream) {
  FH_CHECK_ARG(log_n >= 2 && log_n <= 17, "decode supports logN 2..17");
  FH_CHECK_ARG(batch >= 1 && batch <= 65535, "bad decode batch");
  FH_CHECK_ARG(modulus_a > 2 && modulus_b > 2 && (modulus_b & 1) && modulus_a_inverse_mod_b > 0 &&
                   modulus_a_inverse_mod_b < modulus_b,
               "decode moduli must be odd primes with a^{-1} mod b given");
  FH_CHECK_ARG(scale > 0.0, "decode scale must be positive");
  FH_CHECK_ARG(residues_a && residues_b && slot_positions && workspace && slots, "decode: null operand");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t n = int64_t{1} << log_n;
  const int64_t s = n / 2;
  double2* z = reinterpret_cast<double2*>(workspace);
  hipLaunchKernelGGL(k_decode_crt, dim3(static_cast<uint32_t>(ceil_div(n, kCodecBlock)), static_cast<uint32_t>(batch)),
                     dim3(kCodecBlock), 0, st, residues_a, residues_b, batch_stride, modulus_a, modulus_b,
                     modulus_a_inverse_mod_b, static_cast<int>(log_n), z);
  FH_CHECK_LAUNCH("decode crt");
  int rc = fft_stages(z, batch, static_cast<int>(log_n), 1.0, st);
  if (rc != kOk) return rc;
  hipLaunchKernelGGL(k_decode_gather, dim3(static_cast<uint32_t>(ceil_div(s, kCodecBlock)), static_cast<uint32_t>(batch)),
                     dim3(kCodecBlock), 0, st, z, slot_positions, s, static_cast<int>(log_n), 1.0 / scale, slots);
  FH_CHECK_LAUNCH("decode gather");
  return kOk;
}
