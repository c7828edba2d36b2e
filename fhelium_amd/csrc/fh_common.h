// Shared device arithmetic, status handling and launch helpers for the gfx950
// CKKS backend.  Everything here is integer-exact: every routine returns the
// same residue (same representative, not merely the same class mod q) as the
// reference CPU arithmetic in csrc/ops/common/cpu/montgomery.h:125-215.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

#include "fhelium_amd.h"  // every definition below is checked against it

#define FH_API extern "C" __attribute__((visibility("default")))

namespace fh {

// ---------------------------------------------------------------------------
// Status reporting for the C ABI.  Entry points return 0 on success and a
// negative code on failure; the message is kept per host thread.
// ---------------------------------------------------------------------------
enum Status : int {
  kOk = 0,
  kInvalidArgument = -1,
  kLaunchFailed = -2,
  kUnsupported = -3,
};

void set_error(const char* fmt, ...);
const char* last_error();

#define FH_CHECK_ARG(cond, ...)              \
  do {                                       \
    if (!(cond)) {                           \
      ::fh::set_error(__VA_ARGS__);          \
      return ::fh::kInvalidArgument;         \
    }                                        \
  } while (0)

#define FH_CHECK_LAUNCH(what)                                              \
  do {                                                                     \
    hipError_t fh_err_ = hipGetLastError();                                \
    if (fh_err_ != hipSuccess) {                                           \
      ::fh::set_error("%s: kernel launch failed: %s", what,                \
                      hipGetErrorString(fh_err_));                         \
      return ::fh::kLaunchFailed;                                          \
    }                                                                      \
  } while (0)

// Reference parameter tensor rows, csrc/ops/common/rns_parameters.h:6-16.
enum RnsParam : int {
  kTwiceModulus = 0,
  kModulusLo = 1,
  kModulusHi = 2,
  kNegInvLo = 3,
  kNegInvHi = 4,
  kR2 = 5,
  kScaledR2 = 6,
  kNInvMontgomery = 7,
};

constexpr int kHalfRadixBits = 31;   // R = 2^62 split into 31-bit halves
constexpr int kRadixBits = 62;
constexpr uint64_t kRadixMask = (uint64_t{1} << kRadixBits) - 1;

// Strided view of the [8, limb] parameter table.
struct ParamView {
  const int64_t* base;
  int64_t row_stride;   // stride between parameter kinds
  int64_t limb_stride;  // stride between limbs

  __device__ __forceinline__ int64_t at(int kind, int64_t limb) const {
    return base[kind * row_stride + limb * limb_stride];
  }
};

// Montgomery constants for one prime, reconstructed from the split-word
// parameter rows exactly as the reference loads them.
struct Mont {
  uint64_t q;
  uint64_t k;       // -q^{-1} mod 2^62
  int64_t twice_q;
  int64_t r2;       // R^2 mod q
  int64_t ninv;     // N^{-1} R mod q
};

__device__ __forceinline__ Mont load_mont(const ParamView& p, int64_t limb) {
  Mont m;
  m.twice_q = p.at(kTwiceModulus, limb);
  m.q = static_cast<uint64_t>(p.at(kModulusLo, limb)) +
        (static_cast<uint64_t>(p.at(kModulusHi, limb)) << kHalfRadixBits);
  m.k = static_cast<uint64_t>(p.at(kNegInvLo, limb)) +
        (static_cast<uint64_t>(p.at(kNegInvHi, limb)) << kHalfRadixBits);
  m.r2 = p.at(kR2, limb);
  m.ninv = p.at(kNInvMontgomery, limb);
  return m;
}

// x >= m ? x - m : x for x < 2^63 + m.  gfx950 has a one-instruction 64-bit
// add (v_lshl_add_u64) but subtracts with a carry chain whose VCC read needs
// wait states, and a compare-and-select costs the same wait states.  So the
// difference is formed as x + neg_m (neg_m = 2^64 - m) and the sign word of
// the result selects through a mask.  The empty asm statements keep the
// compiler from folding the addition back into a subtraction and the mask
// back into a compare-and-select; they emit no instruction.
template <class T>
__device__ __forceinline__ T opaque(T v) {
  asm("" : "+v"(v));
  return v;
}

__device__ __forceinline__ uint64_t csub(uint64_t x, uint64_t m, uint64_t neg_m) {
  const uint64_t d = x + opaque(neg_m);
  const uint32_t sign = opaque(static_cast<uint32_t>(static_cast<int32_t>(d >> 32) >> 31));
  const uint64_t keep = static_cast<uint64_t>(sign) << 32 | sign;  // all ones when x < m
  return d + (m & keep);
}

__device__ __forceinline__ uint64_t csub(uint64_t x, uint64_t m) { return csub(x, m, uint64_t{0} - m); }

// REDC with R = 2^62 followed by one conditional subtraction, i.e. the
// reference's `multiply` (montgomery.h:125-143) bit for bit, including the
// 128-bit wrap-around of its unsigned __int128 intermediate.
// REDC of the 128-bit product (hi, lo) with R = 2^62, then one conditional
// subtraction (the tail of the reference's `multiply`).
__device__ __forceinline__ int64_t mont_redc128(uint64_t lo, uint64_t hi, uint64_t q, uint64_t k) {
  const uint64_t m = (lo * k) & kRadixMask;
  const uint64_t mq_lo = m * q;
  const uint64_t mq_hi = __umul64hi(m, q);
  const uint64_t s_lo = lo + mq_lo;
  const uint64_t s_hi = hi + mq_hi + (s_lo < lo ? 1u : 0u);
  const uint64_t r = (s_lo >> kRadixBits) | (s_hi << (64 - kRadixBits));
  return static_cast<int64_t>(csub(r, q));
}

__device__ __forceinline__ int64_t mont_mul(int64_t lhs, int64_t rhs,
                                            uint64_t q, uint64_t k) {
  const uint64_t a = static_cast<uint64_t>(lhs);
  const uint64_t b = static_cast<uint64_t>(rhs);
  const uint64_t lo = a * b;
  const uint64_t hi = __umul64hi(a, b);
  const uint64_t m = (lo * k) & kRadixMask;
  const uint64_t mq_lo = m * q;
  const uint64_t mq_hi = __umul64hi(m, q);
  const uint64_t s_lo = lo + mq_lo;
  const uint64_t s_hi = hi + mq_hi + (s_lo < lo ? 1u : 0u);
  const uint64_t r = (s_lo >> kRadixBits) | (s_hi << (64 - kRadixBits));
  return static_cast<int64_t>(csub(r, q));
}

__device__ __forceinline__ int64_t mont_mul(int64_t a, int64_t b,
                                            const Mont& m) {
  return mont_mul(a, b, m.q, m.k);
}

// montgomery.h:148-172: signed operands are first reduced with C `%` to
// the canonical residue in [0, q).  64-bit division is a long instruction
// sequence on the GPU, so for q >= 2^30 the quotient is estimated in
// binary64 instead: |v/q| < 2^33 and the estimate is within 2^-18 of it,
// so floor() is off by at most one and r = v - t q lies in [-q, 2q); one
// correction each way gives the same canonical residue as `%`.
__device__ __forceinline__ int64_t canonical_operand(int64_t v, uint64_t q) {
  const int64_t qs = static_cast<int64_t>(q);
  if (q < (uint64_t{1} << 30)) {
    int64_t c = v % qs;
    if (c < 0) c += qs;
    return c;
  }
  const double qd = static_cast<double>(q);
  const double est = __builtin_floor(static_cast<double>(v) / qd);
  // two's-complement wrap-around: the true r is small, so the low 64 bits are it
  int64_t r = static_cast<int64_t>(static_cast<uint64_t>(v) -
                                   static_cast<uint64_t>(static_cast<int64_t>(est)) * q);
  r = r < 0 ? r + qs : r;
  r = r >= qs ? r - qs : r;
  return r;
}

__device__ __forceinline__ int64_t mont_mul_split(int64_t a, int64_t b,
                                                  const Mont& m) {
  return mont_mul(canonical_operand(a, m.q), canonical_operand(b, m.q), m);
}

__device__ __forceinline__ int64_t add_lazy(int64_t a, int64_t b,
                                            int64_t twice_q) {
  return static_cast<int64_t>(csub(static_cast<uint64_t>(a + b), static_cast<uint64_t>(twice_q)));
}

// a + 2q - b, with 2q - b = (2q + 1) + ~b (no carry chain), then reduced.
__device__ __forceinline__ int64_t sub_lazy(int64_t a, int64_t b,
                                            int64_t twice_q) {
  const uint64_t d = static_cast<uint64_t>(a) + (static_cast<uint64_t>(twice_q) + 1u) + ~static_cast<uint64_t>(b);
  return static_cast<int64_t>(csub(d, static_cast<uint64_t>(twice_q)));
}

__device__ __forceinline__ int64_t canonicalize(int64_t v, int64_t twice_q) {
  return static_cast<int64_t>(csub(static_cast<uint64_t>(v), static_cast<uint64_t>(twice_q >> 1)));
}

__device__ __forceinline__ int64_t center(int64_t v, int64_t twice_q) {
  const int64_t q = twice_q >> 1;
  return v <= (q >> 1) ? v : v - q;
}

__device__ __forceinline__ int64_t shift_positive(int64_t v, int64_t twice_q) {
  return v + (twice_q >> 1);
}

// ---------------------------------------------------------------------------
// NTT arithmetic.  A butterfly needs V = w*x mod q in canonical form.  Any
// algorithm that returns the canonical product gives the reference's value,
// because the reference REDC also canonicalizes; the lazy add/sub around it
// are reproduced exactly.  Two exact algorithms are used, chosen per prime:
//
//  * q < 2^42 ("small" rows, the 30/31- and 40/41-bit scale primes):
//    binary64 (the limit the fast transform's growth bound needs).  gfx950
//    issues FP64 FMA at the full VALU rate, while 32x32->64 integer
//    multiplies are quarter rate, so this path is several times cheaper.
//    With x < 2q and w < q the product is split error-free as h + l (FMA),
//    the quotient estimate t = round(h/q) is off by less than 0.6, and
//    r = (h - t q) + l is computed exactly, giving r in (-q, q).
//  * otherwise (the 60-bit base and P primes): Shoup multiplication with the
//    precomputed w' = floor(w 2^64 / q), r = w x - hi64(w' x) q in [0, 2q).
// ---------------------------------------------------------------------------
constexpr double kRoundMagic = 6755399441055744.0;  // 1.5 * 2^52

__device__ __forceinline__ double f64_mulmod(double w, double x, double q,
                                             double qinv) {
  const double h = w * x;
  const double l = __builtin_fma(w, x, -h);
  const double t = __builtin_fma(h, qinv, kRoundMagic) - kRoundMagic;
  const double r = __builtin_fma(-t, q, h) + l;
  return r < 0.0 ? r + q : r;
}

// The same product without the final correction: w*x - round(w*x/q) q,
// exact, in (-0.52q, 0.52q) for w < 2q and signed |x| < 2^45.
__device__ __forceinline__ double f64_mulmod_signed(double w, double x, double q, double qinv) {
  const double h = w * x;
  const double l = __builtin_fma(w, x, -h);
  const double t = __builtin_fma(h, qinv, kRoundMagic) - kRoundMagic;
  return __builtin_fma(-t, q, h) + l;
}

__device__ __forceinline__ uint64_t shoup_mulmod(uint64_t w, uint64_t wp,
                                                 uint64_t x, uint64_t q) {
  const uint64_t t = __umul64hi(wp, x);
  // w x - t q mod 2^64 as w x + t (2^64 - q): additions only
  const uint64_t neg_q = opaque(uint64_t{0} - q);
  return csub(w * x + t * neg_q, q, neg_q);
}

// Exact sum of 64 x 64-bit products, for one REDC after many products.  The
// sum is held as c2 2^64 + c1 2^32 + c0 with 32-bit limbs c0, c1; each step
// is four 32 x 32 + 64-bit multiply-adds (a b + c + d < 2^64 for 32-bit a, b,
// c, d), so no carry flag is ever produced or read.  c2 only grows: the
// caller bounds the total below 2^128.
struct WideAcc {
  uint32_t c0, c1;
  uint64_t c2;
};

__device__ __forceinline__ void wide_mac(WideAcc& a, uint64_t y, uint64_t w) {
  const uint32_t y0 = static_cast<uint32_t>(y), y1 = static_cast<uint32_t>(y >> 32);
  const uint32_t w0 = static_cast<uint32_t>(w), w1 = static_cast<uint32_t>(w >> 32);
  const uint64_t t = static_cast<uint64_t>(y0) * w0 + a.c0;
  const uint64_t u = static_cast<uint64_t>(y0) * w1 + (static_cast<uint64_t>(a.c1) + (t >> 32));
  const uint64_t v = static_cast<uint64_t>(y1) * w0 + static_cast<uint32_t>(u);
  a.c0 = static_cast<uint32_t>(t);
  a.c1 = static_cast<uint32_t>(v);
  a.c2 += static_cast<uint64_t>(y1) * w1 + (u >> 32) + (v >> 32);
}

// (sum + m q) / 2^64 with m = -sum q^{-1} mod 2^64: congruent to sum 2^-64,
// below sum / 2^64 + q.  The low word of sum + m q is zero, so its carry is
// (low word of sum != 0).
__device__ __forceinline__ uint64_t wide_redc64(const WideAcc& a, uint64_t q, uint64_t neg_qinv64) {
  const uint64_t lo = static_cast<uint64_t>(a.c1) << 32 | a.c0;
  const uint64_t m = lo * neg_qinv64;
  return a.c2 + __umul64hi(m, q) + (lo != 0 ? 1u : 0u);
}

// 4 r reduced into [0, 2q), r = wide_redc64(sum) below 9q: the residue
// sum 2^-62, i.e. what REDC with radix 2^62 (mont_mul) gives for a single
// product, in the lazy range of the integer inverse butterflies.
__device__ __forceinline__ uint64_t wide_redc62(const WideAcc& a, uint64_t q, uint64_t neg_qinv64) {
  uint64_t r = wide_redc64(a, q, neg_qinv64);
  r = csub(r, 8 * q);
  r = csub(r, 4 * q);
  r = csub(r, 2 * q);
  r = csub(r, q);
  return csub(r << 2, 2 * q);
}

// Exact u64 <-> f64 conversions for integers below 2^52 (every value the
// binary64 paths convert is below 2^47): 2^52 + v has the bit pattern
// 0x4330000000000000 | v, so each direction is one integer and one binary64
// operation.
constexpr double kTwo52 = 4503599627370496.0;
constexpr uint64_t kTwo52Bits = 0x4330000000000000ull;

__device__ __forceinline__ double u64_to_f64(uint64_t v) {
  return __longlong_as_double(static_cast<long long>(v | kTwo52Bits)) - kTwo52;
}

__device__ __forceinline__ uint64_t f64_to_u64(double v) {
  return static_cast<uint64_t>(__double_as_longlong(v + kTwo52)) ^ kTwo52Bits;
}

// Per-row constants of the gfx950 NTT backend, one 128-byte record per QP
// row, built on the host (fhelium_amd/ntt.py) and kept in HBM.
struct NttRow {
  uint64_t q;
  uint64_t twice_q;
  double qd;          // q as binary64 (small rows)
  double qinv;        // 1/q rounded (small rows)
  uint64_t small;     // 1: binary64 path, 0: Shoup path
  uint64_t to_mont;   // R mod q        (x -> x R)
  uint64_t to_mont_p; // Shoup companion
  uint64_t ninv_mont;    // N^{-1} mod q           (Montgomery output)
  uint64_t ninv_mont_p;
  uint64_t ninv_std;     // N^{-1} R^{-1} mod q     (standard output)
  uint64_t ninv_std_p;
  uint64_t ninv_std_rinv;    // N^{-1} R^{-2} mod q (standard output of a
  uint64_t ninv_std_rinv_p;  // binary64 key MAC that skipped REDC's R^{-1})
  uint64_t neg_qinv64;       // -q^{-1} mod 2^64 (REDC with radix 2^64)
  uint64_t reserved[2];
};
static_assert(sizeof(NttRow) == 128, "NttRow must stay 128 bytes");

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace fh
