// Fused hybrid key switching for gfx950.
//
// Same mathematics as the reference's streaming key switch
// (fhelium/engine/hybrid_keyswitch.py:321-371, 530-568 with the native ops
// csrc/ops/rns/cpu/rns_arithmetic_cpu.cpp:482-700 and
// csrc/ops/ckks/cpu/ckks_cpu.cpp:474-708), reorganised into eight launches
// (nine for a rotation) instead of ~100 (ModUp and the tile pass launch once
// for the binary64 rows
// and once for the integer rows; the integer rows' MAC runs on a side
// stream, concurrently with the binary64 rows' launches):
//
//   1. k_ks_decompose     Garner digits of every active hybrid digit (one
//                         launch for all digits); optionally reads its input
//                         through a rotation's Galois gather.
//   2. k_ks_modup_columns basis extension onto the level's QP rows fused into
//                         the forward NTT column pass (the extended digit is
//                         never written out in coefficient form).
//   3. k_ks_tiles_mac     forward NTT tile pass of every digit fused with the
//                         key multiply-accumulate; both accumulators stay in
//                         registers across all digits and are written once.
//                         8 coefficients per thread, the tile's twiddles
//                         staged in LDS, each digit's key loads issued with
//                         its tile loads.  Blocks are remapped so one XCD
//                         walks the batch entries of one (row, tile)
//                         consecutively and re-reads the key tile from its L2.
//   4. inverse NTT of both accumulators (ntt.hip, working-format input).
//   5. k_ks_moddown_add   floor(x / P) mod q_i computed as
//                         (x - (x mod P)) * P^{-1} with x mod P from the
//                         mixed-radix digits of the P residues (once per
//                         coefficient), fused with the engine's final
//                         canonical addition.  A rotation's base c0 goes
//                         through the Galois gather first (k_ks_gather_base,
//                         into the digit workspace, which is dead by then).
//
// Exactness: digits are the unique mixed-radix digits; the extended digit is
// formed with the reference's lazy formula (each term a canonical product,
// computed in binary64 where operands allow); the NTT output is the
// reference value (exact mode) or its canonical representative (fast mode);
// each key product is congruent to the reference's REDC product, so the
// accumulators are congruent to the reference's; the inverse NTT yields the
// unique canonical values, and ModDown (an exact floor division) and the
// final canonical addition are exact functions of those.  The outputs are
// therefore bit-identical to the reference CPU backend's in both modes.
#include <mutex>
#include <type_traits>

#include "fh_common.h"
#include "fh_views.h"
#include "ntt_core.h"

namespace fh {
namespace {

using namespace ::fh::ntt;

constexpr int kMaxDigitRows = 8;
constexpr int kModDownRowsPerThread = 12;

// One active hybrid digit (RnsDigitSpec plus device table addresses),
// packed by the host as 16 int64 words.  Also used for the P primes.
struct KsDigit {
  int64_t row_begin;     // first Q row of the digit, relative to the level
  int64_t rows;          // D
  int64_t key_digit;     // stable level-0 key digit index
  const int64_t* normalizers;  // [D-1]
  const int64_t* propagation;  // [D-1, D] row-major
  const int64_t* lo;           // split moduli / -q^{-1} words [D]
  const int64_t* hi;
  const int64_t* klo;
  const int64_t* khi;
  const int64_t* basis;        // [D-1, QP0] M_r R^2 mod p_j, row-major
  int64_t basis_stride;        // QP0
  const int64_t* basis_std;    // [D-1, QP0] M_r R mod p_j
  int64_t f64_ok;              // every source prime < 2^45: digits stored as binary64
  const int64_t* basis_shoup;  // [D-1, QP0] Shoup companions of basis_std
  // P digit: [Q0, 16] binary64 ModDown records (f64_stride = 16).  A Q digit
  // with f64_ok and D > 1: its binary64 Garner record [D-1 normalizers,
  // (D-1) x D propagation, D moduli, D reciprocals], standard form.
  const double* f64_table;
  int64_t f64_stride;
};
static_assert(sizeof(KsDigit) == 16 * 8, "KsDigit must stay 16 words");

struct KsArgs {
  View3 src;                // component [B, Lq, N], canonical standard
  int64_t* digits;          // [B, Lq, N] contiguous
  int64_t* inter;           // [B, dnum, Lqp, N] contiguous (working format)
  int64_t* acc;             // [2, B, Lqp, N] contiguous (working format)
  const KsDigit* dig;       // [dnum]
  int64_t dnum;
  int64_t batch;
  int64_t lq, lqp;          // Q rows / QP rows at this level
  int64_t level;            // global row of limb 0
  int log_n;
  int mode;                 // 0 exact, 1 fast
  ParamView params;         // global QP parameter table [8, QP0]
  const NttRow* rows;       // global QP NTT records
  const int64_t* fwd_w;     // global forward twiddles [QP0, N]
  const int64_t* fwd_wp;
  const int64_t* inv_w;     // global inverse twiddles [QP0, N]
  const int64_t* inv_wp;
  const int64_t* key;       // shared key [digit, 2, QP0, N] or null
  const int64_t* key_bases; // per-entry key addresses [B] or null
  int64_t key_sd, key_sc, key_sr;   // key strides (digit, component, row)
  const int64_t* gal_idx;   // per-entry int32 gather table addresses [B] or null
  const int64_t* gal_sign;  // per-entry int8 sign table addresses [B]
  int64_t small_rows;       // leading QP rows on the binary64 path
  // Optional accumulator seeds (null: none): [B, Lq, N] NTT-domain rows in
  // the accumulators' scaling, added to the Q rows times seed_scale[row].
  const int64_t* seed0;
  const int64_t* seed1;
  const int64_t* seed_scale;  // [Q0] P mod q (binary64 rows), P R mod q (integer rows)
  // Hoisted rotations (null otherwise): [B, 2] (g, g^{-1} mod 2N) per entry.
  // The decompose and ModUp launches then see one input entry, and every MAC
  // entry reads that entry's intermediate through the NTT-domain Galois map
  // of its own g.
  const int64_t* hoist_galois;
};


// Table addresses read from device memory are generic pointers to the
// compiler, which would then emit FLAT loads; those also count against the
// LDS wait counter and serialise with the LDS exchanges.  Every table lives
// in global memory, so say so.
template <class T>
using Global = const T __attribute__((address_space(1)));
template <class T>
__device__ __forceinline__ Global<T>* global(const T* p) {
  return (Global<T>*)p;
}

__device__ __forceinline__ Mont digit_mont(const KsDigit& d, int r) {
  Mont m;
  m.q = static_cast<uint64_t>(global(d.lo)[r]) + (static_cast<uint64_t>(global(d.hi)[r]) << kHalfRadixBits);
  m.k = static_cast<uint64_t>(global(d.klo)[r]) + (static_cast<uint64_t>(global(d.khi)[r]) << kHalfRadixBits);
  m.twice_q = 0;
  m.r2 = 0;
  m.ninv = 0;
  return m;
}

// Reads component row `row` of batch entry b at coefficient i, through the
// Galois gather of apply_coefficient_galois_automorphism (ckks_cpu.cpp:390-396)
// when one is given.
__device__ __forceinline__ int64_t read_maybe_galois(const View3& v, const int64_t* gal_idx,
                                                     const int64_t* gal_sign, int64_t b, int64_t row,
                                                     int64_t i, int64_t twice_q) {
  if (gal_idx == nullptr) return v.at(b, row, i);
  Global<int32_t>* idx = global(reinterpret_cast<const int32_t*>(gal_idx[b]));
  Global<int8_t>* sgn = global(reinterpret_cast<const int8_t*>(gal_sign[b]));
  int64_t x = v.at(b, row, idx[i]);
  if (sgn[i] == -1) x = -x;
  return canonicalize(shift_positive(x, twice_q), twice_q);
}

// Garner digits of D residues (rns_arithmetic_cpu.cpp:565-600).
__device__ __forceinline__ void garner(const KsDigit& d, int64_t* dg) {
  const int rows = static_cast<int>(d.rows);
  const Mont m0 = digit_mont(d, 0);
  int64_t first = dg[0];
  first = first < static_cast<int64_t>(m0.q) ? first : first - static_cast<int64_t>(m0.q);
  int64_t src[kMaxDigitRows];
#pragma unroll
  for (int r = 0; r < kMaxDigitRows; ++r) {
    src[r] = dg[r];
    dg[r] = first;
  }
#pragma unroll
  for (int step = 0; step < kMaxDigitRows - 1; ++step) {
    const int row = step + 1;
    if (row >= rows) break;
    // multiply_split reduces both operands; the normalizer is canonical
    // already, so only the difference needs reducing.
    const Mont mr = digit_mont(d, row);
    const int64_t digit = mont_mul(canonical_operand(src[row] - dg[row], mr.q), global(d.normalizers)[step], mr);
    dg[row] = digit;
#pragma unroll
    for (int t = 0; t < kMaxDigitRows; ++t) {
      if (t <= row || t >= rows) continue;
      dg[t] += mont_mul(digit, global(d.propagation)[step * rows + t], digit_mont(d, t));
    }
  }
}

// Garner digits with the digit count fixed at compile time (the same
// recurrence as garner(); no runtime row loops).
template <int R>
__device__ __forceinline__ void garner_fixed(const KsDigit& d, int64_t* dg) {
  const Mont m0 = digit_mont(d, 0);
  int64_t first = dg[0];
  first = first < static_cast<int64_t>(m0.q) ? first : first - static_cast<int64_t>(m0.q);
  int64_t src[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    src[r] = dg[r];
    dg[r] = first;
  }
#pragma unroll
  for (int row = 1; row < R; ++row) {
    const Mont mr = digit_mont(d, row);
    const int64_t digit = mont_mul(canonical_operand(src[row] - dg[row], mr.q), global(d.normalizers)[row - 1], mr);
    dg[row] = digit;
#pragma unroll
    for (int t = row + 1; t < R; ++t) dg[t] += mont_mul(digit, global(d.propagation)[(row - 1) * R + t], digit_mont(d, t));
  }
}

// The same recurrence in binary64 for a digit whose primes are all below
// 2^45, from its Garner record (standard-form constants).  Each product is
// f64_mulmod_signed, exact and within (-0.53q, 0.53q) (the operands stay
// below 2^47); running sums of at most R such terms plus a canonical value
// are exact.  Each digit is corrected into [0, q) once, so the digits are
// the unique mixed-radix digits, as the integer recurrence produces them.
template <int R>
__device__ __forceinline__ void garner_f64(Global<double>* rec, const int64_t* x, double* dg) {
  Global<double>* norm = rec;
  Global<double>* prop = rec + (R - 1);
  Global<double>* qs = prop + (R - 1) * R;
  Global<double>* qinv = qs + R;
  double first = static_cast<double>(x[0]);
  first = first < qs[0] ? first : first - qs[0];
  double acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = first;
  dg[0] = first;
#pragma unroll
  for (int row = 1; row < R; ++row) {
    const double q = qs[row];
    double digit = f64_mulmod_signed(norm[row - 1], static_cast<double>(x[row]) - acc[row], q, qinv[row]);
    digit = digit < 0.0 ? digit + q : digit;
    dg[row] = digit;
#pragma unroll
    for (int t = row + 1; t < R; ++t) acc[t] += f64_mulmod_signed(prop[(row - 1) * R + t], digit, qs[t], qinv[t]);
  }
}

// ---------------------------------------------------------------------------
// 1. Garner digits.  Grid (N/256, dnum, B).
// ---------------------------------------------------------------------------
// Digits of one (entry, digit, coefficient) with the digit's row count R
// fixed at compile time (R = 0: any count up to kMaxDigitRows).
template <int R>
__device__ __forceinline__ void decompose_one(const KsArgs& a, const KsDigit& d, int64_t b, int64_t i,
                                              int64_t n) {
  constexpr int kR = R == 0 ? kMaxDigitRows : R;
  const int rows = R == 0 ? static_cast<int>(d.rows) : R;
  int64_t dg[kR];
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    dg[r] = r < rows ? read_maybe_galois(a.src, a.gal_idx, a.gal_sign, b, d.row_begin + r, i,
                                         a.params.at(kTwiceModulus, a.level + d.row_begin + r))
                     : 0;
  }
  int64_t* out = a.digits + (b * a.lq + d.row_begin) * n + i;
  if constexpr (R >= 2) {
    if (d.f64_table != nullptr) {
      // binary64 digit (f64_ok): Garner in binary64, stored as binary64
      double fd[R];
      garner_f64<R>(global(d.f64_table), dg, fd);
#pragma unroll
      for (int r = 0; r < R; ++r) out[static_cast<int64_t>(r) * n] = __double_as_longlong(fd[r]);
      return;
    }
  }
  if constexpr (R == 0) {
    garner(d, dg);
  } else {
    garner_fixed<R>(d, dg);
  }
  // Digits below 2^45 are stored as binary64 (exact), the form the ModUp
  // kernel multiplies in; the others stay integers.
  const bool f64 = d.f64_ok != 0;
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    if (r < rows)
      out[static_cast<int64_t>(r) * n] =
          f64 ? __double_as_longlong(u64_to_f64(static_cast<uint64_t>(dg[r]))) : dg[r];
  }
}

__global__ __launch_bounds__(kBlock) void k_ks_decompose(KsArgs a) {
  const int64_t n = int64_t{1} << a.log_n;
  // One XCD walks all coefficient chunks of an (entry, digit): a Galois
  // gather then touches each source row in one L2 only.
  const uint32_t chunks = static_cast<uint32_t>(n / kBlock);
  const uint32_t dnum = static_cast<uint32_t>(a.dnum);
  const uint32_t w = xcd_remap(blockIdx.x, chunks * dnum * static_cast<uint32_t>(a.batch));
  const int64_t i = static_cast<int64_t>(w % chunks) * kBlock + threadIdx.x;
  const KsDigit& d = a.dig[(w / chunks) % dnum];
  const int64_t b = w / (chunks * dnum);
  // the row count is uniform over the block: branch once to a fixed-size body
  switch (static_cast<int>(d.rows)) {
    case 1: decompose_one<1>(a, d, b, i, n); break;
    case 2: decompose_one<2>(a, d, b, i, n); break;
    case 3: decompose_one<3>(a, d, b, i, n); break;
    case 4: decompose_one<4>(a, d, b, i, n); break;
    default: decompose_one<0>(a, d, b, i, n); break;
  }
}

// Integer digit value whatever its stored form.
__device__ __forceinline__ uint64_t digit_u64(const KsDigit& d, int64_t raw) {
  return d.f64_ok ? f64_to_u64(__longlong_as_double(raw)) : static_cast<uint64_t>(raw);
}

// Lazy [0, 2q) basis extension of one coefficient onto a QP row by the
// reference formula (rns_arithmetic_cpu.cpp:676-691): each REDC term is the
// canonical product d_r (M_r R mod q), computed here as a Shoup product.
__device__ __forceinline__ uint64_t extend_integer(const KsDigit& d, const NttRow& row, int64_t grow,
                                                   const int64_t* dig0, int64_t n, int64_t pos) {
  const int rows = static_cast<int>(d.rows);
  uint64_t x = shoup_mulmod(row.to_mont, row.to_mont_p, digit_u64(d, dig0[pos]), row.q);
  for (int r = 1; r < rows; ++r) {
    const int64_t c = (r - 1) * d.basis_stride + grow;
    const uint64_t y = shoup_mulmod(static_cast<uint64_t>(global(d.basis_std)[c]),
                                    static_cast<uint64_t>(global(d.basis_shoup)[c]),
                                    digit_u64(d, dig0[r * n + pos]), row.q);
    x += y;
    x = x >= row.twice_q ? x - row.twice_q : x;
  }
  return x;
}

// ---------------------------------------------------------------------------
// 2. Basis extension + forward NTT column pass.
//    Grid (T column groups, Lqp destination rows, dnum * B).
// ---------------------------------------------------------------------------
template <int S1, class P>
__device__ __forceinline__ void modup_tail(const P& p, const KsArgs& a, int64_t di, int64_t b, int64_t j, int cg,
                                           typename P::V (&v)[16], int64_t* lds);

template <int S1, int kRows, class P>
__device__ __forceinline__ void modup_columns(const P& p, const KsArgs& a, const KsDigit& d,
                                              int64_t di, int64_t b, int64_t j, int cg, int64_t* lds) {
  constexpr int R1 = 1 << S1;
  constexpr int T = R1 / 16;
  constexpr int C = kBlock / T;
  using V = typename P::V;
  const int64_t n = int64_t{1} << a.log_n;
  const int64_t grow = a.level + j;  // global QP row of the destination
  const NttRow& row = a.rows[grow];
  const int64_t* dig0 = a.digits + (b * a.lq + d.row_begin) * n;
  // kRows: the digit's row count when fixed at compile time (0: runtime)
  const int rows = kRows == 0 ? static_cast<int>(d.rows) : kRows;
  constexpr int kMaxR = kRows == 0 ? kMaxDigitRows : kRows;
  const int c = threadIdx.x % C;
  const int t = threadIdx.x / C;
  const int col = cg * C + c;
  // 32-bit element offsets from the block-uniform row bases: the loads then
  // take a scalar base plus a 32-bit vector offset (no 64-bit address adds)
  auto at = [&](int k) { return static_cast<uint32_t>((t + T * k) * kTile + col); };
  V v[16];
  if constexpr (std::is_same<V, double>::value) {
    if (d.f64_ok) {
      // Binary64 digits times (M_r R mod q).  Exact arithmetic reproduces
      // the reference's canonical REDC terms and add_lazy chain; fast
      // arithmetic keeps the signed congruent sum (|x| < 2.1q), which the
      // fast butterflies accept.
      const double q = row.qd, qinv = row.qinv, twoq = 2.0 * row.qd;
      const double c0 = static_cast<double>(row.to_mont);
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const double x = __longlong_as_double(dig0[at(k)]);
        v[k] = std::is_same<P, F64Fast>::value ? f64_mulmod_signed(c0, x, q, qinv) : f64_mulmod(c0, x, q, qinv);
      }
#pragma unroll
      for (int r = 1; r < kMaxR; ++r) {
        if (r >= rows) break;
        const double cr = static_cast<double>(global(d.basis_std)[(r - 1) * d.basis_stride + grow]);
        const int64_t* src = dig0 + r * n;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          const double x = __longlong_as_double(src[at(k)]);
          if constexpr (std::is_same<P, F64Fast>::value) {
            v[k] += f64_mulmod_signed(cr, x, q, qinv);
          } else {
            const double sum = v[k] + f64_mulmod(cr, x, q, qinv);
            v[k] = sum >= twoq ? sum - twoq : sum;
          }
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int64_t pos = static_cast<int64_t>(t + T * k) * kTile + col;
        v[k] = u64_to_f64(extend_integer(d, row, grow, dig0, n, pos));
      }
    }
  } else {
    // Integer rows (fast path): each term d_r (M_r R mod q) is a Shoup
    // product left in [0, 2q); the sum (below 2 rows q <= 16q < 2^64) is
    // reduced once into [0, 4q), the forward butterflies' input range.
    static_assert(std::is_same<P, U64Fast>::value, "integer rows take the lazy fast path");
    const uint64_t q = row.q;
    const uint64_t neg_q = opaque(uint64_t{0} - q);
    auto term = [&](uint64_t w, uint64_t wp, uint64_t x) { return w * x + __umul64hi(wp, x) * neg_q; };
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const uint32_t pos = at(k);
      v[k] = term(row.to_mont, row.to_mont_p, digit_u64(d, dig0[pos]));
    }
#pragma unroll
    for (int r = 1; r < kMaxR; ++r) {
      if (r >= rows) break;
      const int64_t cidx = (r - 1) * d.basis_stride + grow;
      const uint64_t cr = static_cast<uint64_t>(global(d.basis_std)[cidx]);
      const uint64_t crp = static_cast<uint64_t>(global(d.basis_shoup)[cidx]);
      const int64_t* src = dig0 + r * n;
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const uint32_t pos = at(k);
        v[k] += term(cr, crp, digit_u64(d, src[pos]));
      }
    }
    if (rows > 2) {
#pragma unroll
      for (int k = 0; k < 16; ++k) v[k] = csub(csub(v[k], 8 * q), 4 * q);
    }
  }
  modup_tail<S1>(p, a, di, b, j, cg, v, lds);
}

// The column pass of one destination row after its basis extension: the
// first S1 forward stages (two radix-16 rounds through LDS), then the store
// of the intermediate (6-byte packed runs on the binary64 rows).
template <int S1, class P>
__device__ __forceinline__ void modup_tail(const P& p, const KsArgs& a, int64_t di, int64_t b, int64_t j, int cg,
                                           typename P::V (&v)[16], int64_t* lds) {
  constexpr int R1 = 1 << S1;
  constexpr int T = R1 / 16;
  constexpr int C = kBlock / T;
  const int64_t n = int64_t{1} << a.log_n;
  const int64_t grow = a.level + j;
  const int64_t* tw = a.fwd_w + grow * n;
  const int64_t* twp = a.fwd_wp + grow * n;
  const int c = threadIdx.x % C;
  const int t = threadIdx.x / C;
  const int col = cg * C + c;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int half = 8 >> s;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & half) continue;
      p.fwd(v[k], v[k + half], p.tw(tw, twp, (1 << s) + (k >> (4 - s))));
    }
  }
  int64_t* out = a.inter + ((b * a.dnum + di) * a.lqp + j) * n;
#pragma unroll
  for (int k = 0; k < 16; ++k) lds[(t + T * k) * C + c] = p.mid(v[k]);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = p.unmid(lds[(16 * t + k) * C + c]);
#pragma unroll
  for (int s = 4; s < S1; ++s) {
    const int half = 1 << (S1 - 1 - s);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & half) continue;
      p.fwd(v[k], v[k + half], p.tw(tw, twp, (1 << s) + ((16 * t + k) >> (S1 - s))));
    }
  }
  if constexpr (std::is_same<P, F64Fast>::value) {
    // Through LDS to run ownership: the block's R1 x C values form
    // R1 * C / 16 = 256 runs of 16 columns of one row, each 16 consecutive
    // packed coefficients (64 + 32 bytes); thread u writes run u.
    constexpr int kPitch = C + 1;
    constexpr int kRunsPerRow = C / 16;
    static_assert(R1 * kRunsPerRow == kBlock, "one packed run per thread");
    __syncthreads();  // every round-B read of the exchange buffer is done
#pragma unroll
    for (int k = 0; k < 16; ++k) lds[(16 * t + k) * kPitch + c] = static_cast<int64_t>(f64_to_u64(p.canon(v[k])));
    __syncthreads();
    const PackedInter pk = packed_slab(a.inter + (b * a.dnum + di) * a.lqp * n, a.small_rows, n);
    const int i = threadIdx.x / kRunsPerRow;
    const int c16 = (threadIdx.x % kRunsPerRow) * 16;
    uint32_t lo[16], hi[8];
#pragma unroll
    for (int cc = 0; cc < 16; ++cc) {
      const uint64_t u = static_cast<uint64_t>(lds[i * kPitch + c16 + cc]);
      lo[cc] = static_cast<uint32_t>(u);
      if (cc & 1) {
        hi[cc >> 1] |= static_cast<uint32_t>(u >> 32) << 16;
      } else {
        hi[cc >> 1] = static_cast<uint32_t>(u >> 32) & 0xffffu;
      }
    }
    const uint32_t idx = static_cast<uint32_t>(j * n + packed_index(i, static_cast<int64_t>(cg) * C + c16));
    uint4* dlo = reinterpret_cast<uint4*>(pk.lo + idx);
    uint4* dhi = reinterpret_cast<uint4*>(pk.hi + idx);
#pragma unroll
    for (int q = 0; q < 4; ++q) dlo[q] = make_uint4(lo[4 * q], lo[4 * q + 1], lo[4 * q + 2], lo[4 * q + 3]);
    dhi[0] = make_uint4(hi[0], hi[1], hi[2], hi[3]);
    dhi[1] = make_uint4(hi[4], hi[5], hi[6], hi[7]);
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) out[static_cast<uint32_t>((16 * t + k) * kTile + col)] = p.mid(v[k]);
  }
}

// Binary64 rows in fast arithmetic: one block extends its (entry, digit)
// onto two destination rows j0, j1, so every digit value it loads serves
// both rows' sums; the two rows' column passes then run one after the
// other through the same LDS.  Same values as modup_columns per row.
template <int S1, int kRows>
__device__ __forceinline__ void modup_columns_x2(const KsArgs& a, const KsDigit& d, int64_t di, int64_t b,
                                                 int64_t j0, bool two, int cg, int64_t* lds) {
  constexpr int R1 = 1 << S1;
  constexpr int T = R1 / 16;
  constexpr int C = kBlock / T;
  const int64_t n = int64_t{1} << a.log_n;
  const int64_t j1 = two ? j0 + 1 : j0;
  const int64_t grow0 = a.level + j0;
  const int64_t grow1 = a.level + j1;
  const NttRow& row0 = a.rows[grow0];
  const NttRow& row1 = a.rows[grow1];
  const int64_t* dig0 = a.digits + (b * a.lq + d.row_begin) * n;
  const int rows = kRows == 0 ? static_cast<int>(d.rows) : kRows;
  constexpr int kMaxR = kRows == 0 ? kMaxDigitRows : kRows;
  const int c = threadIdx.x % C;
  const int t = threadIdx.x / C;
  const int col = cg * C + c;
  auto at = [&](int k) { return static_cast<uint32_t>((t + T * k) * kTile + col); };
  double v0[16], v1[16];
  if (d.f64_ok) {
    const double q0 = row0.qd, qi0 = row0.qinv, q1 = row1.qd, qi1 = row1.qinv;
    const double c00 = static_cast<double>(row0.to_mont), c10 = static_cast<double>(row1.to_mont);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const double x = __longlong_as_double(dig0[at(k)]);
      v0[k] = f64_mulmod_signed(c00, x, q0, qi0);
      v1[k] = f64_mulmod_signed(c10, x, q1, qi1);
    }
#pragma unroll
    for (int r = 1; r < kMaxR; ++r) {
      if (r >= rows) break;
      const double cr0 = static_cast<double>(global(d.basis_std)[(r - 1) * d.basis_stride + grow0]);
      const double cr1 = static_cast<double>(global(d.basis_std)[(r - 1) * d.basis_stride + grow1]);
      const int64_t* src = dig0 + r * n;
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const double x = __longlong_as_double(src[at(k)]);
        v0[k] += f64_mulmod_signed(cr0, x, q0, qi0);
        v1[k] += f64_mulmod_signed(cr1, x, q1, qi1);
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int64_t pos = static_cast<int64_t>(t + T * k) * kTile + col;
      v0[k] = u64_to_f64(extend_integer(d, row0, grow0, dig0, n, pos));
      v1[k] = u64_to_f64(extend_integer(d, row1, grow1, dig0, n, pos));
    }
  }
  modup_tail<S1>(F64Fast(row0), a, di, b, j0, cg, v0, lds);
  if (two) {
    __syncthreads();  // the first row's last reads of the LDS buffer are done
    modup_tail<S1>(F64Fast(row1), a, di, b, j1, cg, v1, lds);
  }
}

// One launch per arithmetic path, as for the tile pass below.
template <int S1, int MODE, bool kSmall>
__global__ __launch_bounds__(kBlock) void k_ks_modup_columns(KsArgs a, int64_t j_begin, int64_t j_count) {
  // (+1 column of padding: the packed store's transpose pitch)
  __shared__ int64_t lds[(1 << S1) * (kBlock / ((1 << S1) / 16) + 1)];
  // Work items (column group fastest, then destination row, then (entry,
  // digit)), XCD-contiguous: every block that reads one (entry, digit)'s
  // source rows runs on one XCD, so those rows are fetched into one L2.
  constexpr uint32_t kGroups = (1u << S1) / 16;
  const uint32_t jc = static_cast<uint32_t>(j_count);
  const uint32_t dnum = static_cast<uint32_t>(a.dnum);
  if constexpr (kSmall && MODE == 1) {
    // work items over pairs of destination rows (the last one may be single)
    const uint32_t pairs = (jc + 1) / 2;
    const uint32_t w = xcd_remap(blockIdx.x, kGroups * pairs * dnum * static_cast<uint32_t>(a.batch));
    const int cg = static_cast<int>(w % kGroups);
    const uint32_t jp = (w / kGroups) % pairs;
    const uint32_t z = w / (kGroups * pairs);
    const int64_t b = z / dnum;
    const int64_t di = z % dnum;
    const KsDigit& d = a.dig[di];
    const int64_t j = j_begin + 2 * static_cast<int64_t>(jp);
    const bool two = 2 * jp + 1 < jc;
    // the digit width is uniform over the block: a fixed-width body lets
    // every digit load issue before the extension arithmetic
    switch (static_cast<int>(d.rows)) {
      case 1: modup_columns_x2<S1, 1>(a, d, di, b, j, two, cg, lds); break;
      case 2: modup_columns_x2<S1, 2>(a, d, di, b, j, two, cg, lds); break;
      case 4: modup_columns_x2<S1, 4>(a, d, di, b, j, two, cg, lds); break;
      default: modup_columns_x2<S1, 0>(a, d, di, b, j, two, cg, lds); break;
    }
    return;
  }
  const uint32_t w = xcd_remap(blockIdx.x, kGroups * jc * dnum * static_cast<uint32_t>(a.batch));
  const int cg = static_cast<int>(w % kGroups);
  const int64_t j = j_begin + (w / kGroups) % jc;
  const uint32_t z = w / (kGroups * jc);
  const int64_t b = z / dnum;
  const int64_t di = z % dnum;
  const KsDigit& d = a.dig[di];
  const NttRow& row = a.rows[a.level + j];
  if constexpr (kSmall) {
    modup_columns<S1, 0>(F64Exact(row), a, d, di, b, j, cg, lds);
  } else {
    switch (static_cast<int>(d.rows)) {
      case 1: modup_columns<S1, 1>(U64Fast(row), a, d, di, b, j, cg, lds); break;
      case 2: modup_columns<S1, 2>(U64Fast(row), a, d, di, b, j, cg, lds); break;
      default: modup_columns<S1, 0>(U64Fast(row), a, d, di, b, j, cg, lds); break;
    }
  }
}

// ---------------------------------------------------------------------------
// 3. Forward NTT tile pass + key multiply-accumulate over every digit.
//    Linear grid of (tile group, row, batch) work items, XCD-remapped.
// ---------------------------------------------------------------------------

// Binary64 rows: acc += w*y reduced to (-0.55q, 0.55q), with the key's
// Montgomery value w = k R mod q (< 2q < 2^43, exact) used as a plain integer,
// so the sum is (sum k y) R mod q -- the inverse NTT removes that R
// (kOutStdMac).  Each term (a digit product, or the relinearization seed
// term) lies in (-0.53q, 0.53q), so over T = digit_count + 1 <= 17 terms
// |acc| < 9.1q, and the first inverse butterfly's difference stays below
// 18.1q < 2^46.2 for q < 2^42 (the binary64 prime limit the host enforces):
// inside the exactness bound |w x| / q < 2^46.4 of f64_mulmod_signed (w < q).
__device__ __forceinline__ void mac_f64(double& acc, double y, int64_t key, double q, double qinv) {
  acc += f64_mulmod_signed(u64_to_f64(static_cast<uint64_t>(key)), y, q, qinv);
}


// Tile pass with 8 elements per thread (32 threads per 256-element tile,
// 4 tiles per 128-thread block): rounds of 3, 3 and 2 stages exchanged
// through LDS.
// Each tile's 255 twiddles (they do not depend on the digit) and the key
// digit indices are staged in LDS once per block, so the digit loop issues
// only the tile and key loads -- together, at the top of the iteration --
// and the per-thread state is small enough for 3 waves per SIMD (binary64
// rows) or 2 (integer rows).
constexpr int kMacThreads = 32;                     // threads per tile
constexpr int kMacBlock = 128;                      // threads per tile-pass block
constexpr int kMacTiles = kMacBlock / kMacThreads;  // tiles per block
constexpr int kMacExchange = kMacTiles * kTilePitch;
constexpr int kMacMaxDigits = 16;
static_assert(kMacMaxDigits <= kMacBlock, "one thread per key digit index");

// The 32 threads of one tile are half of one wave64 and each tile's LDS
// exchange buffer and twiddles are its own, so in-tile exchanges use
// tile_sync() rather than a block barrier.
static_assert(64 % kMacThreads == 0, "a tile's threads must lie in one wave");

// Element of the tile held in slot k by thread t in each round.
__device__ __forceinline__ int mac_e1(int t, int k) { return t + 32 * k; }
__device__ __forceinline__ int mac_e2(int t, int k) { return ((t >> 2) << 5) | (k << 2) | (t & 3); }
__device__ __forceinline__ int mac_e3(int t, int k) { return 8 * t + k; }
// Word of tile element x in the tile's LDS exchange buffer: the low five
// bits XORed with bits 3..7, a bijection on [0, 256).  Under it each of the
// three ownership patterns above has the 16 lanes that share an LDS cycle
// (8-byte words, two tiles per wave at kTilePitch apart) on 16 distinct
// bank pairs, so no exchange is bank-conflicted; one pad word per 16
// elements left mac_e2 two-way conflicted.
__device__ __forceinline__ int mac_lds(int x) { return x ^ ((x >> 3) & 31); }
// Slot of the twiddle of the pair whose lower element is e at tile stage st:
// stage st uses 2^st twiddles, stored after those of the earlier stages.
__device__ __forceinline__ int mac_tw_slot(int st, int e) { return (1 << st) - 1 + (e >> (8 - st)); }

// NTT-domain Galois automorphism on tiles (hybrid_keyswitch.py:506-528):
// out[k] = in[br(((2 br(k) + 1) g mod 2N - 1) / 2)].  With k = 256 r + i the
// source of every element of destination tile r lies in one source tile r',
// and inside the tile the map is affine in bit-reversed coordinates:
// br8(src_i) = (g br8(i) + c_r) mod 256, c_r = ((2 br(r) + 1) g mod 2N) >> (L-7).
// In the MAC's slot layout (thread t holds elements 8t..8t+7) the elements
// of destination thread t all come from one source thread.  A hoisted MAC
// block is source-stationary: it reads source tile r' and finds the
// destination tile, and for its thread t the destination element of each
// slot k.
__device__ __forceinline__ uint32_t brev_bits(uint32_t x, int bits) { return __brev(x) >> (32 - bits); }

struct HoistSlot {
  int dst_tile;   // destination tile of source tile r'
  int dst_base;   // 8 * destination thread of source thread t
  int h;          // carry of the destination thread's slot map
};

__device__ __forceinline__ HoistSlot hoist_slot(int r_src, int t, uint32_t g, uint32_t ginv, int log_n) {
  const int l8 = log_n - kTileLog;
  const uint32_t low_mask = (1u << (log_n - 7)) - 1;  // mod 2^(L-7)
  const uint32_t e_src = 2 * brev_bits(static_cast<uint32_t>(r_src), l8) + 1;
  const uint32_t e_dst = (e_src * ginv) & low_mask;
  HoistSlot hs;
  hs.dst_tile = static_cast<int>(brev_bits((e_dst - 1) >> 1, l8));
  const uint32_t u = ((2 * brev_bits(static_cast<uint32_t>(hs.dst_tile), l8) + 1) * g) & ((2u << log_n) - 1);
  const uint32_t c = u >> (log_n - 7);
  const uint32_t td = brev_bits(((brev_bits(static_cast<uint32_t>(t), 5) - c) * ginv) & 31u, 5);
  hs.dst_base = static_cast<int>(8 * td);
  hs.h = static_cast<int>((((g * brev_bits(td, 5) + c) & 255u) >> 5));
  return hs;
}
// Destination slot (0..7) of source slot k.
__device__ __forceinline__ int hoist_dst_slot(int k, int h, uint32_t ginv) {
  return static_cast<int>(brev_bits(((brev_bits(static_cast<uint32_t>(k), 3) - static_cast<uint32_t>(h)) * ginv) & 7u, 3));
}

// Destination of a MAC block's sums.  For a hoisted rotation the block's
// tiles r are source tiles of the NTT-domain Galois map: the accumulators
// belong to tile r_dst, and this thread's source slot k to destination
// element dst_base + slot k (3 bits per slot in slot_map).  Otherwise
// r_dst = r and slot k is element 8t + k.
struct MacDest {
  int r_dst;
  int dst_base;
  uint32_t slot_map;
  __device__ __forceinline__ int slot(int k) const {
    return dst_base + static_cast<int>((slot_map >> (3 * k)) & 7u);
  }
};

__device__ __forceinline__ MacDest hoisted_dest(const KsArgs& a, int r, int t, int64_t b) {
  const uint32_t hg = static_cast<uint32_t>(a.hoist_galois[2 * b]);
  const uint32_t hginv = static_cast<uint32_t>(a.hoist_galois[2 * b + 1]);
  const HoistSlot hs = hoist_slot(r, t, hg, hginv, a.log_n);
  MacDest d{hs.dst_tile, hs.dst_base, 0u};
#pragma unroll
  for (int k = 0; k < 8; ++k) d.slot_map |= static_cast<uint32_t>(hoist_dst_slot(k, hs.h, hginv)) << (3 * k);
  return d;
}

// The accumulators of tile dst.r_dst (thread t holding elements 8t..8t+7):
// the first 8 inverse stages, which act inside the tile (fast arithmetic, as
// the rest of the inverse), then the working-format store the inverse column
// pass reads.  With kHoist the sums arrive in source order and first move to
// the destination layout through the exchange buffer.
template <bool kHoist, class Acc>
__device__ __forceinline__ void mac_finish(const KsArgs& a, int64_t j, int64_t b, int t, const MacDest& dst,
                                           int64_t* sh, int64_t* tww, int64_t* twp, Acc (&acc0)[8],
                                           Acc (&acc1)[8]) {
  constexpr bool kF64 = std::is_same<Acc, double>::value;
  const int64_t n = int64_t{1} << a.log_n;
  const int64_t grow = a.level + j;
  const NttRow& row = a.rows[grow];
  const int r_dst = dst.r_dst;
  if constexpr (kHoist) {
    auto to_dst = [&](Acc (&acc)[8]) {
      tile_sync();  // earlier readers of the exchange buffer are done
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        int64_t bits;
        __builtin_memcpy(&bits, &acc[k], sizeof(bits));
        sh[mac_lds(dst.slot(k))] = bits;
      }
      tile_sync();
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int64_t bits = sh[mac_lds(mac_e3(t, k))];
        __builtin_memcpy(&acc[k], &bits, sizeof(bits));
      }
    };
    to_dst(acc0);
    to_dst(acc1);
  }
  tile_sync();
  {
    const int64_t* w_row = a.inv_w + grow * n;
    const int64_t* wp_row = a.inv_wp + grow * n;
    for (int idx = t; idx < kTile - 1; idx += kMacThreads) {
      // stage st owns slots [256 - 2^(8-st), 256 - 2^(7-st)); group g = e >> (st+1)
      const int st = 7 - (31 - __builtin_clz(kTile - 1 - idx));
      const int g = idx - (kTile - (kTile >> st));
      const int64_t gi = (n >> (st + 1)) + (static_cast<int64_t>(r_dst) << (7 - st)) + g;
      tww[idx] = w_row[gi];
      if constexpr (!kF64) twp[idx] = wp_row[gi];
    }
  }
  tile_sync();
  auto inv_slot = [](int st, int e) { return (kTile - (kTile >> st)) + (e >> (st + 1)); };
  auto inverse_tile = [&](const auto& pi, auto (&v)[8], auto store) {
    // Round 1 (slots = bits 2..0 of the element): inverse stages 0..2.
#pragma unroll
    for (int st = 0; st < 3; ++st) {
      const int half = 1 << st;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (k & half) continue;
        const int slot = inv_slot(st, mac_e3(t, k));
        pi.inv(v[k], v[k + half], pi.tw_make(tww[slot], kF64 ? 0 : twp[slot]));
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) sh[mac_lds(mac_e3(t, k))] = pi.mid(v[k]);
    tile_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = pi.unmid(sh[mac_lds(mac_e2(t, k))]);
    // Round 2 (slots = bits 4..2): stages 3, 4.
#pragma unroll
    for (int st = 3; st < 5; ++st) {
      const int half = 1 << (st - 2);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (k & half) continue;
        const int slot = inv_slot(st, mac_e2(t, k));
        pi.inv(v[k], v[k + half], pi.tw_make(tww[slot], kF64 ? 0 : twp[slot]));
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) sh[mac_lds(mac_e2(t, k))] = pi.mid(v[k]);
    tile_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = pi.unmid(sh[mac_lds(mac_e1(t, k))]);
    // Round 3 (slots = bits 7..5): stages 5..7.
#pragma unroll
    for (int st = 5; st < 8; ++st) {
      const int half = 1 << (st - 5);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (k & half) continue;
        const int slot = inv_slot(st, mac_e1(t, k));
        pi.inv(v[k], v[k + half], pi.tw_make(tww[slot], kF64 ? 0 : twp[slot]));
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) store(mac_e1(t, k), v[k]);
    // the exchange buffer is reused by the next accumulator
    __syncthreads();
  };
  if constexpr (kF64) {
    // binary64 rows of the accumulators travel packed (canonical, 6 bytes)
    const F64Fast pi(row);
    const PackedInter pk0 = packed_slab(a.acc + (0 * a.batch + b) * a.lqp * n, a.small_rows, n);
    const PackedInter pk1 = packed_slab(a.acc + (1 * a.batch + b) * a.lqp * n, a.small_rows, n);
    const int64_t row_off = j * n;
    inverse_tile(pi, acc0, [&](int e, double x) {
      packed_store(pk0, row_off + packed_index(r_dst, e), f64_to_u64(pi.canon(x)));
    });
    inverse_tile(pi, acc1, [&](int e, double x) {
      packed_store(pk1, row_off + packed_index(r_dst, e), f64_to_u64(pi.canon(x)));
    });
  } else {
    const U64Fast pi(row);
    int64_t* out0 = a.acc + ((0 * a.batch + b) * a.lqp + j) * n + static_cast<int64_t>(r_dst) * kTile;
    int64_t* out1 = a.acc + ((1 * a.batch + b) * a.lqp + j) * n + static_cast<int64_t>(r_dst) * kTile;
    inverse_tile(pi, acc0, [&](int e, uint64_t x) { out0[e] = static_cast<int64_t>(x); });
    inverse_tile(pi, acc1, [&](int e, uint64_t x) { out1[e] = static_cast<int64_t>(x); });
  }
}

template <class P>
__device__ __forceinline__ void tiles_mac(const P& p, const KsArgs& a, int64_t g, int64_t j,
                                          int64_t b, int64_t* lds) {
  using V = typename P::V;
  constexpr bool kF64 = std::is_same<V, double>::value;
  const int64_t n = int64_t{1} << a.log_n;
  const int64_t grow = a.level + j;
  const NttRow& row = a.rows[grow];
  const int m = threadIdx.x / kMacThreads;
  const int t = threadIdx.x % kMacThreads;
  const int r = static_cast<int>(g) * kMacTiles + m;
  const int s0 = a.log_n - kTileLog;
  int64_t* sh = lds + m * kTilePitch;
  int64_t* tww = lds + kMacExchange + m * kTile;
  int64_t* twp = tww + kMacTiles * kTile;  // integer rows only
  int64_t* key_digit = lds + kMacExchange + kMacTiles * kTile * (kF64 ? 1 : 2);
  {
    const int64_t* w_row = a.fwd_w + grow * n;
    const int64_t* wp_row = a.fwd_wp + grow * n;
    for (int idx = t; idx < kTile - 1; idx += kMacThreads) {
      const int st = 31 - __builtin_clz(idx + 1);
      const int64_t gi = (int64_t{1} << (s0 + st)) + (static_cast<int64_t>(r) << st) + (idx + 1 - (1 << st));
      tww[idx] = w_row[gi];
      if constexpr (!kF64) twp[idx] = wp_row[gi];
    }
    if (threadIdx.x < a.dnum) key_digit[threadIdx.x] = a.dig[threadIdx.x].key_digit;
  }
  __syncthreads();
  auto tw_at = [&](int st, int e) {
    const int slot = mac_tw_slot(st, e);
    return p.tw_make(tww[slot], kF64 ? 0 : twp[slot]);
  };
  Global<int64_t>* key = global(a.key_bases ? reinterpret_cast<const int64_t*>(a.key_bases[b]) : a.key);
  const int64_t tile_off = static_cast<int64_t>(r) * kTile;
  double facc0[8], facc1[8];
  uint64_t uacc0[8], uacc1[8];
  WideAcc wacc0[8], wacc1[8];  // integer rows: exact sums, one REDC after the digits
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    facc0[k] = facc1[k] = 0.0;
    wacc0[k] = wacc1[k] = WideAcc{0, 0, 0};
  }
  const int64_t* tile = a.inter + ((b * a.dnum) * a.lqp + j) * n + tile_off;
  const int64_t digit_stride = a.lqp * n;
  // Binary64 rows: software pipeline over the digits -- the next digit's
  // packed values (one 32-byte and one 16-byte load: coefficients 8t..8t+7
  // are contiguous in the packed block layout) load while this digit's
  // stages and products run.
  const uint32_t pidx = static_cast<uint32_t>(j * n + packed_index(r, 8 * t));
  auto load_packed = [&](int64_t di, uint4& l0, uint4& l1, uint4& hh) {
    const PackedInter pk = packed_slab(a.inter + (b * a.dnum + di) * a.lqp * n, a.small_rows, n);
    l0 = *reinterpret_cast<const uint4*>(pk.lo + pidx);
    l1 = *reinterpret_cast<const uint4*>(pk.lo + pidx + 4);
    hh = *reinterpret_cast<const uint4*>(pk.hi + pidx);
  };
  uint4 nl0{}, nl1{}, nhh{};
  int64_t nv[8];  // other rows: the next digit's values, likewise
  if constexpr (std::is_same<P, F64Fast>::value) {
    load_packed(0, nl0, nl1, nhh);
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) nv[k] = tile[mac_e1(t, k)];
  }
#pragma nounroll
  for (int64_t di = 0; di < a.dnum; ++di, tile += digit_stride) {
    Global<int64_t>* k0 = key + key_digit[di] * a.key_sd + grow * a.key_sr + tile_off;
    Global<int64_t>* k1 = k0 + a.key_sc;
    V v[8];
    int64_t w0[8], w1[8];
    if constexpr (std::is_same<P, F64Fast>::value) {
      // an LDS pass moves the values to the round-1 ownership
      const uint4 l0 = nl0, l1 = nl1, hh = nhh;
      if (di + 1 < a.dnum) load_packed(di + 1, nl0, nl1, nhh);
      const uint32_t lo[8] = {l0.x, l0.y, l0.z, l0.w, l1.x, l1.y, l1.z, l1.w};
      const uint32_t hi2[4] = {hh.x, hh.y, hh.z, hh.w};
      tile_sync();  // the previous digit's last reads of the exchange buffer are done
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const uint64_t h = (hi2[k >> 1] >> (16 * (k & 1))) & 0xffffu;
        sh[mac_lds(mac_e3(t, k))] = __double_as_longlong(u64_to_f64(static_cast<uint64_t>(lo[k]) | (h << 32)));
      }
      tile_sync();
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = p.unmid(sh[mac_lds(mac_e1(t, k))]);
    } else {
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = p.unmid(nv[k]);
      if (di + 1 < a.dnum) {
#pragma unroll
        for (int k = 0; k < 8; ++k) nv[k] = tile[digit_stride + mac_e1(t, k)];
      }
    }
    // Binary64 rows issue the key loads with the tile loads; the integer
    // rows (whose products need many more registers) load them at the MAC.
    if constexpr (kF64) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        w0[k] = k0[mac_e3(t, k)];
        w1[k] = k1[mac_e3(t, k)];
      }
    }
    // Round 1: tile stages 0..2 pair element bits 7..5 = bits 2..0 of k.
#pragma unroll
    for (int st = 0; st < 3; ++st) {
      const int half = 4 >> st;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (k & half) continue;
        p.fwd(v[k], v[k + half], tw_at(st, mac_e1(t, k)));
      }
    }
    // The previous digit's last reads of this buffer must be done.
    tile_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) sh[mac_lds(mac_e1(t, k))] = p.mid(v[k]);
    tile_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = p.unmid(sh[mac_lds(mac_e2(t, k))]);
    // Round 2: tile stages 3..5 pair element bits 4..2 = bits 2..0 of k.
#pragma unroll
    for (int st = 3; st < 6; ++st) {
      const int half = 1 << (5 - st);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (k & half) continue;
        p.fwd(v[k], v[k + half], tw_at(st, mac_e2(t, k)));
      }
    }
    // Each thread rewrites the words it just read: no barrier needed.
#pragma unroll
    for (int k = 0; k < 8; ++k) sh[mac_lds(mac_e2(t, k))] = p.mid(v[k]);
    tile_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = p.unmid(sh[mac_lds(mac_e3(t, k))]);
    // Round 3: tile stages 6..7 pair element bits 1..0 = bits 1..0 of k.
#pragma unroll
    for (int st = 6; st < 8; ++st) {
      const int half = 1 << (7 - st);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (k & half) continue;
        p.fwd(v[k], v[k + half], tw_at(st, mac_e3(t, k)));
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if constexpr (kF64) {
        mac_f64(facc0[k], v[k], w0[k], row.qd, row.qinv);
        mac_f64(facc1[k], v[k], w1[k], row.qd, row.qinv);
      } else {
        w0[k] = k0[mac_e3(t, k)];
        w1[k] = k1[mac_e3(t, k)];
        wide_mac(wacc0[k], v[k], static_cast<uint64_t>(w0[k]));
        wide_mac(wacc1[k], v[k], static_cast<uint64_t>(w1[k]));
      }
    }
  }
  // Relinearization seeds: P d0 and P d1 (NTT domain, in the accumulators'
  // scaling) join the Q rows' sums, so that ModDown's exact floor division
  // returns floor(X / P) + d directly.  Thread t holds coefficients 8t + k,
  // stored at tile positions t + 32k (the HMult front's seed layout).
  if (a.seed0 != nullptr && j < a.lq) {
    const int64_t off = (b * a.lq + j) * n + tile_off + t;
    if constexpr (kF64) {
      const double sc = static_cast<double>(a.seed_scale[grow]);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        facc0[k] += f64_mulmod_signed(sc, __longlong_as_double(a.seed0[off + 32 * k]), row.qd, row.qinv);
        facc1[k] += f64_mulmod_signed(sc, __longlong_as_double(a.seed1[off + 32 * k]), row.qd, row.qinv);
      }
    } else {
      const uint64_t sc = static_cast<uint64_t>(a.seed_scale[grow]);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        wide_mac(wacc0[k], static_cast<uint64_t>(a.seed0[off + 32 * k]), sc);
        wide_mac(wacc1[k], static_cast<uint64_t>(a.seed1[off + 32 * k]), sc);
      }
    }
  }
  if constexpr (!kF64) {
    // Each sum holds at most 16 products y w with y < 4q (forward outputs)
    // and w < 2q (key words), plus the seed term s c with s < 2q and c < q,
    // so it is below 130 q^2 < 2^127.1 (it fits the 128-bit accumulator) and
    // its REDC value r = sum 2^-64 mod q is below 130 q^2 / 2^64 + q < 9.2q
    // (q < 2^60), inside the 16q that wide_redc62's reduction chain covers.
    // The binary64 rows' REDC radix is 2^62: 4 r is the same residue as
    // sum 2^-62.  Result in [0, 2q), the inverse butterflies' input range.
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      uacc0[k] = wide_redc62(wacc0[k], row.q, row.neg_qinv64);
      uacc1[k] = wide_redc62(wacc1[k], row.q, row.neg_qinv64);
    }
  }
  if constexpr (kF64) {
    mac_finish<false>(a, j, b, t, MacDest{r, 8 * t, 0u}, sh, tww, twp, facc0, facc1);
  } else {
    mac_finish<false>(a, j, b, t, MacDest{r, 8 * t, 0u}, sh, tww, twp, uacc0, uacc1);
  }
}

template <bool kSmall>
constexpr int mac_lds_words() {
  return kMacExchange + kMacTiles * kTile * (kSmall ? 1 : 2) + kMacMaxDigits;
}
constexpr int kMacTilesPerBlock = kMacTiles;

// One launch per arithmetic path (binary64 rows, then the integer rows) so
// each kernel is register-allocated for its own path only.
template <int MODE, bool kSmall>
// Register budgets: 3 waves per SIMD for the binary64 rows (no spills at 156
// VGPRs), 2+ for the integer rows.
__global__ __launch_bounds__(kMacBlock) __attribute__((amdgpu_waves_per_eu(kSmall ? 3 : 2))) void k_ks_tiles_mac(KsArgs a, int64_t groups, int64_t j_begin,
                                                         int64_t j_count) {
  __shared__ int64_t lds[mac_lds_words<kSmall>()];
  const uint32_t batch = static_cast<uint32_t>(a.batch);
  const uint32_t jc = static_cast<uint32_t>(j_count);
  const uint32_t w = xcd_remap(blockIdx.x, static_cast<uint32_t>(groups) * jc * batch);
  const int64_t b = w % batch;
  const int64_t j = j_begin + (w / batch) % jc;
  const int64_t g = w / (batch * jc);
  const NttRow& row = a.rows[a.level + j];
  if constexpr (kSmall) {
    if constexpr (MODE == 1) {
      tiles_mac(F64Fast(row), a, g, j, b, lds);
    } else {
      tiles_mac(F64Exact(row), a, g, j, b, lds);
    }
  } else {
    tiles_mac(U64Fast(row), a, g, j, b, lds);
  }
}

// ---------------------------------------------------------------------------
// Hoisted rotations.  Every rotation of the one input reads the same
// extended digits, so their forward tile pass runs once (k_ks_tiles_fwd,
// in place), and the per-rotation MAC (k_ks_mac_hoisted) only loads the
// transformed digits through the rotation's NTT-domain Galois map and
// multiplies them by its key.
// ---------------------------------------------------------------------------

// Forward tile pass of tile r of row j of digit di of the one input entry,
// stored back in place fully transformed: canonical residues in the packed
// planes (fast binary64 rows), binary64 bits below 2q (exact binary64 rows)
// or u64 below 4q (integer rows).  Each tile's threads load all of its
// elements before any of them stores, and tiles are disjoint.
template <class P>
__device__ __forceinline__ void tiles_forward_store(const P& p, const KsArgs& a, int64_t g, int64_t j,
                                                    int64_t di, int64_t* lds) {
  using V = typename P::V;
  constexpr bool kF64 = std::is_same<V, double>::value;
  constexpr bool kPacked = std::is_same<P, F64Fast>::value;
  const int64_t n = int64_t{1} << a.log_n;
  const int64_t grow = a.level + j;
  const int m = threadIdx.x / kMacThreads;
  const int t = threadIdx.x % kMacThreads;
  const int r = static_cast<int>(g) * kMacTiles + m;
  const int s0 = a.log_n - kTileLog;
  int64_t* sh = lds + m * kTilePitch;
  int64_t* tww = lds + kMacExchange + m * kTile;
  int64_t* twp = tww + kMacTiles * kTile;  // integer rows only
  {
    const int64_t* w_row = a.fwd_w + grow * n;
    const int64_t* wp_row = a.fwd_wp + grow * n;
    for (int idx = t; idx < kTile - 1; idx += kMacThreads) {
      const int st = 31 - __builtin_clz(idx + 1);
      const int64_t gi = (int64_t{1} << (s0 + st)) + (static_cast<int64_t>(r) << st) + (idx + 1 - (1 << st));
      tww[idx] = w_row[gi];
      if constexpr (!kF64) twp[idx] = wp_row[gi];
    }
  }
  tile_sync();
  auto tw_at = [&](int st, int e) {
    const int slot = mac_tw_slot(st, e);
    return p.tw_make(tww[slot], kF64 ? 0 : twp[slot]);
  };
  int64_t* slab = a.inter + di * a.lqp * n;
  int64_t* tile = slab + j * n + static_cast<int64_t>(r) * kTile;
  V v[8];
  const PackedInter pk = packed_slab(slab, a.small_rows, n);
  const uint32_t pidx = static_cast<uint32_t>(j * n + packed_index(r, 8 * t));
  if constexpr (kPacked) {
    const uint4 l0 = *reinterpret_cast<const uint4*>(pk.lo + pidx);
    const uint4 l1 = *reinterpret_cast<const uint4*>(pk.lo + pidx + 4);
    const uint4 hh = *reinterpret_cast<const uint4*>(pk.hi + pidx);
    const uint32_t lo[8] = {l0.x, l0.y, l0.z, l0.w, l1.x, l1.y, l1.z, l1.w};
    const uint32_t hi2[4] = {hh.x, hh.y, hh.z, hh.w};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint64_t h = (hi2[k >> 1] >> (16 * (k & 1))) & 0xffffu;
      sh[mac_lds(mac_e3(t, k))] = __double_as_longlong(u64_to_f64(static_cast<uint64_t>(lo[k]) | (h << 32)));
    }
    tile_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = p.unmid(sh[mac_lds(mac_e1(t, k))]);
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = p.unmid(tile[mac_e1(t, k)]);
  }
#pragma unroll
  for (int st = 0; st < 3; ++st) {
    const int half = 4 >> st;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (k & half) continue;
      p.fwd(v[k], v[k + half], tw_at(st, mac_e1(t, k)));
    }
  }
  tile_sync();
#pragma unroll
  for (int k = 0; k < 8; ++k) sh[mac_lds(mac_e1(t, k))] = p.mid(v[k]);
  tile_sync();
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = p.unmid(sh[mac_lds(mac_e2(t, k))]);
#pragma unroll
  for (int st = 3; st < 6; ++st) {
    const int half = 1 << (5 - st);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (k & half) continue;
      p.fwd(v[k], v[k + half], tw_at(st, mac_e2(t, k)));
    }
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) sh[mac_lds(mac_e2(t, k))] = p.mid(v[k]);
  tile_sync();
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = p.unmid(sh[mac_lds(mac_e3(t, k))]);
#pragma unroll
  for (int st = 6; st < 8; ++st) {
    const int half = 1 << (7 - st);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (k & half) continue;
      p.fwd(v[k], v[k + half], tw_at(st, mac_e3(t, k)));
    }
  }
  if constexpr (kPacked) {
#pragma unroll
    for (int k = 0; k < 8; ++k) packed_store(pk, pidx + k, f64_to_u64(p.canon(v[k])));
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) tile[mac_e3(t, k)] = p.mid(v[k]);
  }
}

template <int MODE, bool kSmall>
__global__ __launch_bounds__(kMacBlock) void k_ks_tiles_fwd(KsArgs a, int64_t groups, int64_t j_begin,
                                                            int64_t j_count) {
  __shared__ int64_t lds[kMacExchange + kMacTiles * kTile * (kSmall ? 1 : 2)];
  const uint32_t dnum = static_cast<uint32_t>(a.dnum);
  const uint32_t jc = static_cast<uint32_t>(j_count);
  const uint32_t w = xcd_remap(blockIdx.x, static_cast<uint32_t>(groups) * jc * dnum);
  const int64_t di = w % dnum;
  const int64_t j = j_begin + (w / dnum) % jc;
  const int64_t g = w / (dnum * jc);
  const NttRow& row = a.rows[a.level + j];
  if constexpr (kSmall) {
    if constexpr (MODE == 1) {
      tiles_forward_store(F64Fast(row), a, g, j, di, lds);
    } else {
      tiles_forward_store(F64Exact(row), a, g, j, di, lds);
    }
  } else {
    tiles_forward_store(U64Fast(row), a, g, j, di, lds);
  }
}

// One rotation (entry b) of the hoisted MAC over the transformed digits.
// The block is source-stationary: it reads source tiles r of every digit
// (the same tiles for every entry, so consecutive entries on an XCD find
// them in its L2) and accumulates, in source order, the products with the
// key words at the destination elements; mac_finish moves the sums to
// destination tile order.
template <class P>
__device__ __forceinline__ void tiles_mac_hoisted(const P& p, const KsArgs& a, int64_t g, int64_t j,
                                                  int64_t b, int64_t* lds) {
  using V = typename P::V;
  constexpr bool kF64 = std::is_same<V, double>::value;
  constexpr bool kPacked = std::is_same<P, F64Fast>::value;
  const int64_t n = int64_t{1} << a.log_n;
  const int64_t grow = a.level + j;
  const NttRow& row = a.rows[grow];
  const int m = threadIdx.x / kMacThreads;
  const int t = threadIdx.x % kMacThreads;
  const int r = static_cast<int>(g) * kMacTiles + m;
  int64_t* sh = lds + m * kTilePitch;
  int64_t* tww = lds + kMacExchange + m * kTile;
  int64_t* twp = tww + kMacTiles * kTile;  // integer rows only
  int64_t* key_digit = lds + kMacExchange + kMacTiles * kTile * (kF64 ? 1 : 2);
  if (threadIdx.x < a.dnum) key_digit[threadIdx.x] = a.dig[threadIdx.x].key_digit;
  __syncthreads();
  const MacDest dst = hoisted_dest(a, r, t, b);
  Global<int64_t>* key = global(a.key_bases ? reinterpret_cast<const int64_t*>(a.key_bases[b]) : a.key);
  const int64_t dst_off = static_cast<int64_t>(dst.r_dst) * kTile;
  double facc0[8], facc1[8];
  uint64_t uacc0[8], uacc1[8];
  WideAcc wacc0[8], wacc1[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    facc0[k] = facc1[k] = 0.0;
    wacc0[k] = wacc1[k] = WideAcc{0, 0, 0};
  }
  const int64_t digit_stride = a.lqp * n;
  const int64_t* tile = a.inter + j * n + static_cast<int64_t>(r) * kTile;
  const uint32_t pidx = static_cast<uint32_t>(j * n + packed_index(r, 8 * t));
  // software pipeline: the next digit's key words and values load while
  // this digit's products run
  auto load_digit = [&](int64_t di, int64_t (&w0)[8], int64_t (&w1)[8], V (&v)[8]) {
    Global<int64_t>* k0 = key + key_digit[di] * a.key_sd + grow * a.key_sr + dst_off;
    Global<int64_t>* k1 = k0 + a.key_sc;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      w0[k] = k0[dst.slot(k)];
      w1[k] = k1[dst.slot(k)];
    }
    if constexpr (kPacked) {
      const PackedInter pk = packed_slab(a.inter + di * digit_stride, a.small_rows, n);
      const uint4 l0 = *reinterpret_cast<const uint4*>(pk.lo + pidx);
      const uint4 l1 = *reinterpret_cast<const uint4*>(pk.lo + pidx + 4);
      const uint4 hh = *reinterpret_cast<const uint4*>(pk.hi + pidx);
      const uint32_t lo[8] = {l0.x, l0.y, l0.z, l0.w, l1.x, l1.y, l1.z, l1.w};
      const uint32_t hi2[4] = {hh.x, hh.y, hh.z, hh.w};
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const uint64_t h = (hi2[k >> 1] >> (16 * (k & 1))) & 0xffffu;
        v[k] = u64_to_f64(static_cast<uint64_t>(lo[k]) | (h << 32));
      }
    } else {
      const int64_t* tl = tile + di * digit_stride;
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = p.unmid(tl[mac_e3(t, k)]);
    }
  };
  int64_t wa0[8], wa1[8], wb0[8], wb1[8];
  V va[8], vb[8];
  load_digit(0, wa0, wa1, va);
  auto mac_digit = [&](int64_t (&w0)[8], int64_t (&w1)[8], V (&v)[8]) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if constexpr (kF64) {
        mac_f64(facc0[k], v[k], w0[k], row.qd, row.qinv);
        mac_f64(facc1[k], v[k], w1[k], row.qd, row.qinv);
      } else {
        wide_mac(wacc0[k], v[k], static_cast<uint64_t>(w0[k]));
        wide_mac(wacc1[k], v[k], static_cast<uint64_t>(w1[k]));
      }
    }
  };
#pragma nounroll
  for (int64_t di = 0; di < a.dnum; di += 2) {
    if (di + 1 < a.dnum) load_digit(di + 1, wb0, wb1, vb);
    mac_digit(wa0, wa1, va);
    if (di + 1 >= a.dnum) break;
    if (di + 2 < a.dnum) load_digit(di + 2, wa0, wa1, va);
    mac_digit(wb0, wb1, vb);
  }
  if constexpr (kF64) {
    mac_finish<true>(a, j, b, t, dst, sh, tww, twp, facc0, facc1);
  } else {
    // at most 16 products y w with y < 4q, w < 2q: REDC value below 9q
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      uacc0[k] = wide_redc62(wacc0[k], row.q, row.neg_qinv64);
      uacc1[k] = wide_redc62(wacc1[k], row.q, row.neg_qinv64);
    }
    mac_finish<true>(a, j, b, t, dst, sh, tww, twp, uacc0, uacc1);
  }
}

// Two waves per SIMD on both paths: with the next digit's loads in flight
// the binary64 body needs 182 VGPRs; a three-wave budget spilled 44 bytes per
// thread, and the scratch traffic cost more than the extra wave gained
// (128-diagonal matrix-vector product 27.5 against 26.3 ms).
template <int MODE, bool kSmall>
__global__ __launch_bounds__(kMacBlock) __attribute__((amdgpu_waves_per_eu(2))) void k_ks_mac_hoisted(
    KsArgs a, int64_t groups, int64_t j_begin, int64_t j_count) {
  __shared__ int64_t lds[mac_lds_words<kSmall>()];
  const uint32_t batch = static_cast<uint32_t>(a.batch);
  const uint32_t jc = static_cast<uint32_t>(j_count);
  const uint32_t w = xcd_remap(blockIdx.x, static_cast<uint32_t>(groups) * jc * batch);
  const int64_t b = w % batch;
  const int64_t j = j_begin + (w / batch) % jc;
  const int64_t g = w / (batch * jc);
  const NttRow& row = a.rows[a.level + j];
  if constexpr (kSmall) {
    if constexpr (MODE == 1) {
      tiles_mac_hoisted(F64Fast(row), a, g, j, b, lds);
    } else {
      tiles_mac_hoisted(F64Exact(row), a, g, j, b, lds);
    }
  } else {
    tiles_mac_hoisted(U64Fast(row), a, g, j, b, lds);
  }
}

// ---------------------------------------------------------------------------
// HMult front: the tensor product of two coefficient-domain ciphertexts,
// returned in the coefficient domain (what relinearization consumes).
// The forward column passes of a0, a1, b0, b1 (to Montgomery form) run as
// ordinary launches; this kernel then runs, per tile, the four forward tile
// passes, the products d0 = a0 b0, d1 = a0 b1 + a1 b0, d2 = a1 b1 and the
// first eight inverse stages of d0, d1, d2, so the NTT-domain operands and
// products never go to HBM.  Binary64 rows multiply the Montgomery values as
// plain residues (d = a b R^2) and integer rows with REDC (d = a b R); the
// inverse column pass with kOutStdMac removes exactly those factors.
// ---------------------------------------------------------------------------
struct HmArgs {
  const int64_t* ws;  // [4, B, L, N] column-pass outputs of a0, a1, b0, b1
  int64_t* out;       // [3, B, L, N] working-format rows for the inverse column pass
  int64_t batch, limbs, row_start;
  int log_n;
  ParamView params;   // global QP rows
  const NttRow* rows; // global QP NTT records
  const int64_t* fwd_w;
  const int64_t* fwd_wp;
  const int64_t* inv_w;
  const int64_t* inv_wp;
};

// kNttD01: d0, d1 stay in the NTT domain (a compile-time choice: a runtime
// branch between the two tails costs the kernel its register allocation).
template <bool kNttD01, class P>
__device__ __forceinline__ void tiles_hmult(const P& p, const HmArgs& a, int64_t g, int64_t j, int64_t b,
                                            int64_t* lds) {
  using V = typename P::V;
  constexpr bool kF64 = std::is_same<V, double>::value;
  const int64_t n = int64_t{1} << a.log_n;
  const int64_t grow = a.row_start + j;
  const NttRow& row = a.rows[grow];
  const int m = threadIdx.x / kMacThreads;
  const int t = threadIdx.x % kMacThreads;
  const int r = static_cast<int>(g) * kMacTiles + m;
  const int s0 = a.log_n - kTileLog;
  int64_t* sh = lds + m * kTilePitch;
  int64_t* tww = lds + kMacExchange + m * kTile;
  int64_t* twp = tww + kMacTiles * kTile;  // integer rows only
  {
    const int64_t* w_row = a.fwd_w + grow * n;
    const int64_t* wp_row = a.fwd_wp + grow * n;
    for (int idx = t; idx < kTile - 1; idx += kMacThreads) {
      const int st = 31 - __builtin_clz(idx + 1);
      const int64_t gi = (int64_t{1} << (s0 + st)) + (static_cast<int64_t>(r) << st) + (idx + 1 - (1 << st));
      tww[idx] = w_row[gi];
      if constexpr (!kF64) twp[idx] = wp_row[gi];
    }
  }
  tile_sync();
  auto tw_at = [&](int st, int e) {
    const int slot = mac_tw_slot(st, e);
    return p.tw_make(tww[slot], kF64 ? 0 : twp[slot]);
  };
  const int64_t tile_off = static_cast<int64_t>(r) * kTile;
  const int64_t comp_stride = a.batch * a.limbs * n;
  const int64_t* in = a.ws + (b * a.limbs + j) * n + tile_off;
  // Forward tile stages of component c; returns the values in the e3
  // layout (thread t holds elements 8t .. 8t+7).
  auto forward_tile = [&](int c, V (&v)[8]) {
    const int64_t* tile = in + c * comp_stride;
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = p.unmid(tile[mac_e1(t, k)]);
#pragma unroll
    for (int st = 0; st < 3; ++st) {
      const int half = 4 >> st;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (k & half) continue;
        p.fwd(v[k], v[k + half], tw_at(st, mac_e1(t, k)));
      }
    }
    tile_sync();  // earlier readers of the exchange buffer are done
#pragma unroll
    for (int k = 0; k < 8; ++k) sh[mac_lds(mac_e1(t, k))] = p.mid(v[k]);
    tile_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = p.unmid(sh[mac_lds(mac_e2(t, k))]);
#pragma unroll
    for (int st = 3; st < 6; ++st) {
      const int half = 1 << (5 - st);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (k & half) continue;
        p.fwd(v[k], v[k + half], tw_at(st, mac_e2(t, k)));
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) sh[mac_lds(mac_e2(t, k))] = p.mid(v[k]);
    tile_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = p.unmid(sh[mac_lds(mac_e3(t, k))]);
#pragma unroll
    for (int st = 6; st < 8; ++st) {
      const int half = 1 << (7 - st);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (k & half) continue;
        p.fwd(v[k], v[k + half], tw_at(st, mac_e3(t, k)));
      }
    }
  };
  V d0[8], d1[8], d2[8];
  {
    V x0[8], x1[8], y[8];
    WideAcc w1[kF64 ? 1 : 8];  // integer rows: d1's exact two-product sum
    forward_tile(0, x0);  // a0
    forward_tile(1, x1);  // a1
    forward_tile(2, y);   // b0
    if constexpr (kF64) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        x0[k] = p.reduce(x0[k]);
        x1[k] = p.reduce(x1[k]);
        const double yb = p.reduce(y[k]);
        d0[k] = f64_mulmod_signed(x0[k], yb, row.qd, row.qinv);
        d1[k] = f64_mulmod_signed(x1[k], yb, row.qd, row.qinv);
      }
    } else {
      // exact products of the forward outputs (below 4q), one REDC per
      // result (sums below 32 q^2: REDC value below 3q)
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        WideAcc w0{0, 0, 0};
        wide_mac(w0, x0[k], y[k]);
        d0[k] = wide_redc62(w0, row.q, row.neg_qinv64);
        w1[k] = WideAcc{0, 0, 0};
        wide_mac(w1[k], x1[k], y[k]);
      }
    }
    forward_tile(3, y);   // b1
    if constexpr (kF64) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const double yb = p.reduce(y[k]);
        d1[k] += f64_mulmod_signed(x0[k], yb, row.qd, row.qinv);
        d2[k] = f64_mulmod_signed(x1[k], yb, row.qd, row.qinv);
      }
    } else {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        wide_mac(w1[k], x0[k], y[k]);
        d1[k] = wide_redc62(w1[k], row.q, row.neg_qinv64);
        WideAcc w2{0, 0, 0};
        wide_mac(w2, x1[k], y[k]);
        d2[k] = wide_redc62(w2, row.q, row.neg_qinv64);
      }
    }
  }
  // Inverse tile stages (the first eight of the inverse NTT) of d0, d1, d2.
  tile_sync();  // every thread is done with the forward twiddles
  {
    const int64_t* w_row = a.inv_w + grow * n;
    const int64_t* wp_row = a.inv_wp + grow * n;
    for (int idx = t; idx < kTile - 1; idx += kMacThreads) {
      const int st = 7 - (31 - __builtin_clz(kTile - 1 - idx));
      const int gg = idx - (kTile - (kTile >> st));
      const int64_t gi = (n >> (st + 1)) + (static_cast<int64_t>(r) << (7 - st)) + gg;
      tww[idx] = w_row[gi];
      if constexpr (!kF64) twp[idx] = wp_row[gi];
    }
  }
  tile_sync();
  auto inv_slot = [](int st, int e) { return (kTile - (kTile >> st)) + (e >> (st + 1)); };
  auto inverse_tile = [&](V (&v)[8], int64_t* out) {
#pragma unroll
    for (int st = 0; st < 3; ++st) {
      const int half = 1 << st;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (k & half) continue;
        const int slot = inv_slot(st, mac_e3(t, k));
        p.inv(v[k], v[k + half], p.tw_make(tww[slot], kF64 ? 0 : twp[slot]));
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) sh[mac_lds(mac_e3(t, k))] = p.mid(v[k]);
    tile_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = p.unmid(sh[mac_lds(mac_e2(t, k))]);
#pragma unroll
    for (int st = 3; st < 5; ++st) {
      const int half = 1 << (st - 2);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (k & half) continue;
        const int slot = inv_slot(st, mac_e2(t, k));
        p.inv(v[k], v[k + half], p.tw_make(tww[slot], kF64 ? 0 : twp[slot]));
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) sh[mac_lds(mac_e2(t, k))] = p.mid(v[k]);
    tile_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = p.unmid(sh[mac_lds(mac_e1(t, k))]);
#pragma unroll
    for (int st = 5; st < 8; ++st) {
      const int half = 1 << (st - 5);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (k & half) continue;
        const int slot = inv_slot(st, mac_e1(t, k));
        p.inv(v[k], v[k + half], p.tw_make(tww[slot], kF64 ? 0 : twp[slot]));
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) out[mac_e1(t, k)] = p.mid(v[k]);
    tile_sync();  // the exchange buffer is reused by the next product
  };
  int64_t* out = a.out + (b * a.limbs + j) * n + tile_off;
  if constexpr (kNttD01) {
    // d0, d1 stay in the NTT domain as the key switch's accumulator seeds.
    // Seed layout: within each 256-coefficient tile, coefficient 8t + k (the
    // ownership both this kernel and the MAC end in) sits at position
    // t + 32k, so the stores and the MAC's loads are coalesced.
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      out[t + 32 * k] = p.mid(d0[k]);
      out[comp_stride + t + 32 * k] = p.mid(d1[k]);
    }
  } else {
    inverse_tile(d0, out);
    inverse_tile(d1, out + comp_stride);
  }
  inverse_tile(d2, out + 2 * comp_stride);
}

template <bool kSmall, bool kNttD01>
__global__ __launch_bounds__(kMacBlock) __attribute__((amdgpu_waves_per_eu(kSmall ? 3 : 2))) void k_hmult_tiles(
    HmArgs a, int64_t groups, int64_t j_begin, int64_t j_count) {
  __shared__ int64_t lds[kMacExchange + kMacTiles * kTile * (kSmall ? 1 : 2)];
  const uint32_t batch = static_cast<uint32_t>(a.batch);
  const uint32_t jc = static_cast<uint32_t>(j_count);
  const uint32_t w = xcd_remap(blockIdx.x, static_cast<uint32_t>(groups) * jc * batch);
  const int64_t b = w % batch;
  const int64_t j = j_begin + (w / batch) % jc;
  const int64_t g = w / (batch * jc);
  const NttRow& row = a.rows[a.row_start + j];
  if constexpr (kSmall) {
    tiles_hmult<kNttD01>(F64Fast(row), a, g, j, b, lds);
  } else {
    tiles_hmult<kNttD01>(U64Fast(row), a, g, j, b, lds);
  }
}

// ---------------------------------------------------------------------------
// 5. ModDown + final canonical addition.
//    Grid (N/256, ceil(Lq / 4), 2 * B); each thread owns one coefficient of
//    four Q rows of one accumulator.  After the inverse NTT the accumulator
//    rows hold canonical standard residues of X over Q_l P.
// ---------------------------------------------------------------------------
struct ModDownArgs {
  const int64_t* acc;       // [2, B, Lqp, N], rows [0, small_rows) packed
  int64_t batch, lq, lqp, level, small_rows;
  int log_n;
  ParamView params;
  KsDigit pdig;             // the P primes as one digit (extension over QP0 columns)
  const int64_t* p_inverse; // [Q0] P^{-1} mod q, standard
  const NttRow* rows;       // global QP NTT records
  View3 base0, base1, out0, out1;
  int has_base0, has_base1;  // a rotation's base0 arrives already gathered
  // Optional fused rescale (null: none): rows 1..lq-1 of the result are
  // written as rows 0..lq-2 of rescale_drop_leading_prime_nearest's output
  // (ckks_cpu.cpp:204-275): ((c_i - c_0) q_0^{-1} + [c_0 > half_drop]) mod q_i.
  const int64_t* rescale_inv;        // [lq - 1] q_0^{-1} mod q_i, standard
  const int64_t* rescale_inv_shoup;  // [lq - 1] Shoup companions
  int64_t half_drop;
};

// A rotation's base c0 through the Galois gather, written as contiguous rows
// for ModDown: out[b, row, i] = +/- c0[b, row, idx[i]] canonicalised, as
// apply_coefficient_galois_automorphism (ckks_cpu.cpp:390-396) produces it.
// One XCD walks all coefficient chunks of one (entry, row), so the scattered
// reads of a source row stay in that XCD's L2; gathered inside ModDown, each
// XCD's eight-row working set overflowed its L2 and every row was fetched
// several times.
__global__ __launch_bounds__(kBlock) void k_ks_gather_base(View3 base, const int64_t* gal_idx,
                                                          const int64_t* gal_sign, const NttRow* rows,
                                                          int64_t level, int64_t lq, int64_t batch, int log_n,
                                                          int64_t* out) {
  const int64_t n = int64_t{1} << log_n;
  const uint32_t chunks = static_cast<uint32_t>(n / kBlock);
  const uint32_t w = xcd_remap(blockIdx.x, chunks * static_cast<uint32_t>(lq * batch));
  const int64_t i = static_cast<int64_t>(w % chunks) * kBlock + threadIdx.x;
  const int64_t row = (w / chunks) % lq;
  const int64_t b = w / (chunks * static_cast<uint32_t>(lq));
  const int64_t twice_q = static_cast<int64_t>(rows[level + row].twice_q);
  out[(b * lq + row) * n + i] = read_maybe_galois(base, gal_idx, gal_sign, b, row, i, twice_q);
}

template <int P>
__global__ __launch_bounds__(kBlock) void k_ks_moddown_add(ModDownArgs a) {
  constexpr int kRows = kModDownRowsPerThread;
  const int64_t n = int64_t{1} << a.log_n;
  // XCD-contiguous work items: all coefficient chunks of one (accumulator,
  // entry, row group) run on one XCD (see k_ks_decompose).
  const uint32_t chunks = static_cast<uint32_t>(n / kBlock);
  const uint32_t groups = static_cast<uint32_t>((a.lq + kRows - 1) / kRows);
  const uint32_t batch = static_cast<uint32_t>(a.batch);
  const uint32_t w = xcd_remap(blockIdx.x, chunks * groups * 2 * batch);
  // Threads walk packed positions: the packed rows are then read as whole
  // runs, and each wave's 64 coefficients are four runs of 16 consecutive
  // ones (whole cache lines of the int64 rows).
  const int64_t pidx = static_cast<int64_t>(w % chunks) * kBlock + threadIdx.x;
  const int64_t i = packed_element(pidx);
  const int64_t group = (w / chunks) % groups;
  const uint32_t cb = w / (chunks * groups);
  const int64_t c = cb / batch;
  const int64_t b = cb % batch;
  int64_t* slab = const_cast<int64_t*>(a.acc) + (c * a.batch + b) * a.lqp * n;
  // Q rows below small_rows are packed binary64 rows (canonical residues)
  const PackedInter pk = packed_slab(slab, a.small_rows, n);
  auto acc_row = [&](int64_t row) -> int64_t {
    return row < a.small_rows ? static_cast<int64_t>(packed_load(pk, static_cast<uint32_t>(row * n + pidx)))
                              : slab[static_cast<uint32_t>(row * n + i)];
  };
  const View3& base = c == 0 ? a.base0 : a.base1;
  const bool has_base = (c == 0 ? a.has_base0 : a.has_base1) != 0;
  const View3& out = c == 0 ? a.out0 : a.out1;
  const int64_t row_begin = group * kRows;
  // Issue every load of the thread first: P rows, Q rows, base rows.
  int64_t dg[P];
#pragma unroll
  for (int r = 0; r < P; ++r) dg[r] = acc_row(a.lq + r);
  int64_t xv[kRows], bv[kRows];
#pragma unroll
  for (int rr = 0; rr < kRows; ++rr) {
    const int64_t row = row_begin + rr;
    xv[rr] = row < a.lq ? acc_row(row) : 0;
    bv[rr] = (has_base && row < a.lq) ? base.at(b, row, i) : 0;
  }
  // Mixed-radix digits of X mod P (X_P = sum_r c_r prod_{t<r} p_t), each
  // split as c_hi 2^30 + c_lo (both exact in binary64).
  garner_fixed<P>(a.pdig, dg);
  double dlo[P], dhi[P];
#pragma unroll
  for (int r = 0; r < P; ++r) {
    dlo[r] = static_cast<double>(static_cast<uint32_t>(dg[r] & ((int64_t{1} << 30) - 1)));
    dhi[r] = static_cast<double>(static_cast<uint32_t>(dg[r] >> 30));
  }
  const int64_t pitch = a.pdig.f64_stride;
  // Final canonical addition of the base (add_canonical).
  auto with_base = [&](int64_t bval, int64_t value, int64_t twice_q) -> int64_t {
    if (!has_base) return value;
    int64_t sum = bval + value;
    if (sum >= twice_q) sum -= twice_q;
    return canonicalize(sum, twice_q);
  };
  // floor(X / P) mod q of a binary64 row from its 16-word record (M_r,
  // 2^30 M_r, P^{-1}, ..., q, 1/q): (x - X_P) P^{-1} mod q with
  // X_P mod q = sum_r c_lo M_r + c_hi 2^30 M_r, all in signed binary64
  // (|sum| < 4.2q), canonicalised once.
  // The result of a binary64 row stays in binary64 (exact integers below
  // 2^47) through the base addition and the fused rescale.
  auto moddown_f64 = [&](Global<double>* rec, int64_t xval) -> double {
    const double q = rec[14], qinv = rec[15];
    double xp = 0.0;
#pragma unroll
    for (int r = 0; r < P; ++r) {
      xp += f64_mulmod_signed(rec[r], dlo[r], q, qinv);
      xp += f64_mulmod_signed(rec[P + r], dhi[r], q, qinv);
    }
    const double diff = u64_to_f64(static_cast<uint64_t>(xval)) - xp;
    double res = f64_mulmod_signed(rec[2 * P], diff, q, qinv);
    return res < 0.0 ? res + q : res;
  };
  // add_canonical of a canonical value and a base below 2q
  auto add_base_f64 = [&](double v, int64_t bval, double q) -> double {
    if (!has_base) return v;
    double sum = v + static_cast<double>(bval);
    sum = sum >= 2.0 * q ? sum - 2.0 * q : sum;
    return sum >= q ? sum - q : sum;
  };
  // Fused rescale: every thread needs the final row-0 value of its
  // coefficient (row 0 is a binary64 row: the dropped prime is a scale prime).
  const bool rescale = a.rescale_inv != nullptr;
  double c0 = 0.0;
  double round_up = 0.0;
  if (rescale) {
    Global<double>* rec0 = global(a.pdig.f64_table) + a.level * pitch;
    const int64_t x0 = row_begin == 0 ? xv[0] : acc_row(0);
    const int64_t b0 = row_begin == 0 ? bv[0] : (has_base ? base.at(b, 0, i) : 0);
    c0 = add_base_f64(moddown_f64(rec0, x0), b0, rec0[14]);
    round_up = c0 > static_cast<double>(a.half_drop) ? 1.0 : 0.0;
  }
#pragma unroll
  for (int rr = 0; rr < kRows; ++rr) {
    const int64_t row = row_begin + rr;
    if (row >= a.lq) break;
    const int64_t grow = a.level + row;
    Global<double>* rec = global(a.pdig.f64_table) + grow * pitch;
    const double q = rec[14];
    int64_t value;
    if (q != 0.0) {
      double v = add_base_f64(moddown_f64(rec, xv[rr]), bv[rr], q);
      if (rescale && row > 0) {
        // ((v - c0) q0^{-1} + round_up) mod q, |v - c0| < 2^47 exact
        double r = f64_mulmod_signed(static_cast<double>(a.rescale_inv[row - 1]), v - c0, q, rec[15]);
        r += round_up;
        v = r < 0.0 ? r + q : (r >= q ? r - q : r);
      }
      value = static_cast<int64_t>(f64_to_u64(v));
    } else {
      // A prime at or above the binary64 limit 2^42 (the base prime): the
      // integer formula.
      const Mont m = load_mont(a.params, grow);
      // (X_P R) mod q by the reference basis-extension formula, then
      // ((x R) - (X_P R)) * P^{-1} with REDC = (x - X_P) P^{-1} mod q.
      int64_t xp = mont_mul(dg[0], m.r2, m);
#pragma unroll
      for (int r = 1; r < P; ++r)
        xp = add_lazy(xp, mont_mul(dg[r], global(a.pdig.basis)[(r - 1) * a.pdig.basis_stride + grow], m),
                      m.twice_q);
      const int64_t xr = mont_mul(xv[rr], m.r2, m);
      value = with_base(bv[rr], mont_mul(sub_lazy(xr, xp, m.twice_q), a.p_inverse[grow], m), m.twice_q);
      if (rescale && row > 0) {
        const uint64_t qq = m.q;
        int64_t d = value - static_cast<int64_t>(c0);
        if (d < 0) d += static_cast<int64_t>(qq);
        uint64_t r = shoup_mulmod(static_cast<uint64_t>(a.rescale_inv[row - 1]),
                                  static_cast<uint64_t>(a.rescale_inv_shoup[row - 1]), static_cast<uint64_t>(d), qq);
        r += static_cast<uint64_t>(round_up);
        if (r >= qq) r -= qq;
        value = static_cast<int64_t>(r);
      }
    }
    if (!rescale) {
      out.at(b, row, i) = value;
    } else if (row > 0) {
      out.at(b, row - 1, i) = value;
    }
  }
}

// The integer-row (60-bit prime) ModUp runs first on the caller's stream;
// then the integer-row MAC runs on a side stream, forked from the caller's
// stream by an event, beside the binary64 rows' ModUp and MAC, and is joined
// back before the inverse column pass.  (Running both integer launches on
// the side stream from the start, or forking after the binary64 ModUp, was
// slower: the integer ModUp then contends with the binary64 MAC.)  One side
// stream and one fork/join event pair per device, created on first use; the
// mutex keeps the pair to one fused call at a time.
struct SideStream {
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr;
  hipEvent_t join = nullptr;
};
constexpr int kMaxDevices = 64;
SideStream g_side[kMaxDevices];
std::mutex g_side_mutex;

int side_stream(SideStream** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("fused key switch: no usable current device (%d)", dev);
    return kLaunchFailed;
  }
  SideStream& sd = g_side[dev];
  if (sd.stream == nullptr) {
    hipStream_t st = nullptr;
    hipEvent_t f = nullptr, j = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&f, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&j, hipEventDisableTiming) != hipSuccess) {
      set_error("fused key switch: side stream creation failed: %s", hipGetErrorString(hipGetLastError()));
      return kLaunchFailed;
    }
    sd.stream = st;
    sd.fork = f;
    sd.join = j;
  }
  *out = &sd;
  return kOk;
}

// Workspace of fh_keyswitch_fused_static: one array per device, 64 MiB.
// Host code never uses this name as an address; it reads the device address
// through hipGetSymbolAddress once per device (g_scratch_ptr).
constexpr int64_t kScratchWords = int64_t{1} << 23;
__device__ __align__(256) int64_t scratch[kScratchWords];
int64_t* g_scratch_ptr[kMaxDevices];
std::mutex g_scratch_mutex;

// The calling thread's current device's copy of `scratch`.
int scratch_address(int64_t** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("static key switch: no usable current device (%d)", dev);
    return kLaunchFailed;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || ptr == nullptr) {
      set_error("static key switch: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                hipGetErrorString(rc));
      return kLaunchFailed;
    }
    g_scratch_ptr[dev] = static_cast<int64_t*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return kOk;
}

// The eight-launch fused key switch.  With use_side_stream the integer rows'
// MAC runs on the device's side stream concurrently with the binary64 rows'
// ModUp and MAC; without it every launch goes to the caller's stream.
int keyswitch_fused_impl(fh_view3 component, int64_t batch, int64_t lq, int64_t lqp,
                         int64_t level, int64_t log_n, int64_t small_rows,
                         const int64_t* digit_table,
                         int64_t digit_count,
                         const int64_t* p_digit_table,
                         const int64_t* p_inverse_q, fh_table2 rns_params, const void* ntt_rows,
                         const int64_t* forward_twiddles, const int64_t* forward_twiddles_shoup,
                         const int64_t* inverse_twiddles, const int64_t* inverse_twiddles_shoup,
                         const int64_t* key, const int64_t* key_bases, int64_t key_stride_digit,
                         int64_t key_stride_comp, int64_t key_stride_row,
                         const int64_t* galois_index_tables, const int64_t* galois_sign_tables,
                         int64_t* workspace_digits, int64_t* workspace_inter,
                         int64_t* workspace_acc, fh_view3 base0, fh_view3 base1, fh_view3 out0,
                         fh_view3 out1, int32_t mode, const int64_t* rescale_inverse,
                         const int64_t* rescale_inverse_shoup, int64_t half_drop,
                         const int64_t* seed0, const int64_t* seed1, const int64_t* seed_scale,
                         const int64_t* hoist_galois, bool use_side_stream, void* stream) {
  FH_CHECK_ARG(log_n >= 12 && log_n <= 16, "fused key switch supports logN 12..16");
  FH_CHECK_ARG(batch >= 1 && batch <= 32767, "fused key switch batch %lld out of range",
               (long long)batch);
  FH_CHECK_ARG(lq >= 1 && lqp > lq && lqp - lq <= 6, "bad Q/QP row counts (at most 6 P primes)");
  FH_CHECK_ARG(digit_count >= 1 && digit_count <= 16 && digit_count * batch <= 65535, "bad digit count");
  FH_CHECK_ARG(mode == 0 || mode == 1, "mode must be 0 (exact) or 1 (fast)");
  FH_CHECK_ARG(component.data && digit_table && p_digit_table && p_inverse_q && rns_params.data &&
                   ntt_rows && forward_twiddles && forward_twiddles_shoup && inverse_twiddles &&
                   inverse_twiddles_shoup && workspace_digits && workspace_inter && workspace_acc &&
                   out0.data && out1.data,
               "fused key switch: null operand");
  FH_CHECK_ARG((key != nullptr) != (key_bases != nullptr), "pass exactly one of key / key_bases");
  FH_CHECK_ARG((galois_index_tables == nullptr) == (galois_sign_tables == nullptr),
               "Galois index and sign tables go together");
  FH_CHECK_ARG((seed0 == nullptr) == (seed1 == nullptr) && (seed0 == nullptr || seed_scale != nullptr),
               "accumulator seeds come in pairs with their scale table");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t n = int64_t{1} << log_n;
  KsArgs a;
  a.src = View3(component);
  a.digits = workspace_digits;
  a.inter = workspace_inter;
  a.acc = workspace_acc;
  a.dig = reinterpret_cast<const KsDigit*>(digit_table);
  a.dnum = digit_count;
  a.batch = batch;
  a.lq = lq;
  a.lqp = lqp;
  a.level = level;
  a.log_n = static_cast<int>(log_n);
  a.mode = mode;
  a.params = ParamView{rns_params.data, rns_params.stride_row, rns_params.stride_col};
  a.rows = static_cast<const NttRow*>(ntt_rows);
  a.fwd_w = forward_twiddles;
  a.fwd_wp = forward_twiddles_shoup;
  a.inv_w = inverse_twiddles;
  a.inv_wp = inverse_twiddles_shoup;
  a.key = key;
  a.key_bases = key_bases;
  a.key_sd = key_stride_digit;
  a.key_sc = key_stride_comp;
  a.key_sr = key_stride_row;
  a.gal_idx = galois_index_tables;
  a.gal_sign = galois_sign_tables;
  a.small_rows = small_rows;
  a.seed0 = seed0;
  a.seed1 = seed1;
  a.seed_scale = seed_scale;
  a.hoist_galois = hoist_galois;
  // Hoisted rotations decompose and extend the one input entry once; the
  // MAC and everything after it run per entry.
  const bool hoist = hoist_galois != nullptr;
  FH_CHECK_ARG(!hoist || (seed0 == nullptr && rescale_inverse == nullptr),
               "hoisted rotations take no accumulator seeds and no fused rescale");
  KsArgs a_in = a;
  if (hoist) {
    a_in.batch = 1;
    a_in.gal_idx = nullptr;
    a_in.gal_sign = nullptr;
  }
  const int64_t batch_in = hoist ? 1 : batch;

  FH_CHECK_ARG(small_rows >= 0 && small_rows <= lqp, "bad binary64 row count");
  const int s1 = static_cast<int>(log_n) - kTileLog;
  // n >= 4096 is a multiple of the block size; grids are flat and remapped.
  const int64_t dec_blocks = (n / kBlock) * digit_count * batch_in;
  FH_CHECK_ARG(dec_blocks <= 0x7fffffff, "key switch grid too large");
  hipLaunchKernelGGL(k_ks_decompose, dim3(static_cast<uint32_t>(dec_blocks)), dim3(kBlock), 0, s, a_in);
  FH_CHECK_LAUNCH("key switch decompose");

  const int T = (1 << s1) / 16;
  const int C = kBlock / T;
  const int64_t groups = n / (kTile * kMacTilesPerBlock);
  FH_CHECK_ARG(groups * lqp * batch <= 0x7fffffff, "key switch grid too large");
  // pass 0: binary64 rows [0, small_rows); pass 1: integer rows.
  auto launch_modup = [&](int pass, hipStream_t st) -> int {
    const int64_t j_begin = pass == 0 ? 0 : small_rows;
    const int64_t j_count = pass == 0 ? small_rows : lqp - small_rows;
    if (j_count == 0) return kOk;
    // binary64 rows in fast arithmetic: one block per pair of rows
    const int64_t j_items = pass == 0 && mode == 1 ? (j_count + 1) / 2 : j_count;
    const int64_t modup_blocks = (kTile / C) * j_items * digit_count * batch_in;
    FH_CHECK_ARG(modup_blocks <= 0x7fffffff, "key switch grid too large");
    const dim3 gcol(static_cast<uint32_t>(modup_blocks));
#define FH_MODUP_CASE(S)                                                                           \
  case S:                                                                                          \
    if (pass == 1) {                                                                               \
      hipLaunchKernelGGL((k_ks_modup_columns<S, 0, false>), gcol, dim3(kBlock), 0, st, a_in, j_begin, j_count); \
    } else if (mode == 1) {                                                                        \
      hipLaunchKernelGGL((k_ks_modup_columns<S, 1, true>), gcol, dim3(kBlock), 0, st, a_in, j_begin, j_count);  \
    } else {                                                                                       \
      hipLaunchKernelGGL((k_ks_modup_columns<S, 0, true>), gcol, dim3(kBlock), 0, st, a_in, j_begin, j_count);  \
    }                                                                                              \
    break;
    switch (s1) {
      FH_MODUP_CASE(4)
      FH_MODUP_CASE(5)
      FH_MODUP_CASE(6)
      FH_MODUP_CASE(7)
      default:
        FH_MODUP_CASE(8)
    }
#undef FH_MODUP_CASE
    FH_CHECK_LAUNCH("key switch modup + column pass");
    return kOk;
  };
  auto launch_mac = [&](int pass, hipStream_t st) -> int {
    const int64_t j_begin = pass == 0 ? 0 : small_rows;
    const int64_t j_count = pass == 0 ? small_rows : lqp - small_rows;
    if (j_count == 0) return kOk;
    const dim3 grid(static_cast<uint32_t>(groups * j_count * batch));
#define FH_MAC_LAUNCH(K, ARGS)                                                                               \
    if (pass == 1) {                                                                                        \
      hipLaunchKernelGGL((K<0, false>), grid, dim3(kMacBlock), 0, st, ARGS, groups, j_begin, j_count);      \
    } else if (mode == 1) {                                                                                 \
      hipLaunchKernelGGL((K<1, true>), grid, dim3(kMacBlock), 0, st, ARGS, groups, j_begin, j_count);       \
    } else {                                                                                                \
      hipLaunchKernelGGL((K<0, true>), grid, dim3(kMacBlock), 0, st, ARGS, groups, j_begin, j_count);       \
    }
    if (hoist) {
      // the one input entry's digits: forward tile pass once, in place
      const dim3 fgrid(static_cast<uint32_t>(groups * j_count * digit_count));
      if (pass == 1) {
        hipLaunchKernelGGL((k_ks_tiles_fwd<0, false>), fgrid, dim3(kMacBlock), 0, st, a_in, groups, j_begin, j_count);
      } else if (mode == 1) {
        hipLaunchKernelGGL((k_ks_tiles_fwd<1, true>), fgrid, dim3(kMacBlock), 0, st, a_in, groups, j_begin, j_count);
      } else {
        hipLaunchKernelGGL((k_ks_tiles_fwd<0, true>), fgrid, dim3(kMacBlock), 0, st, a_in, groups, j_begin, j_count);
      }
      FH_CHECK_LAUNCH("hoisted key switch forward tile pass");
      FH_MAC_LAUNCH(k_ks_mac_hoisted, a)
    } else {
      FH_MAC_LAUNCH(k_ks_tiles_mac, a)
    }
#undef FH_MAC_LAUNCH
    FH_CHECK_LAUNCH("key switch tile pass + MAC");
    return kOk;
  };
  if (!use_side_stream) {
    for (int pass = 0; pass < 2; ++pass) {
      int rc = launch_modup(pass, s);
      if (rc == kOk) rc = launch_mac(pass, s);
      if (rc != kOk) return rc;
    }
  } else {
    std::lock_guard<std::mutex> lock(g_side_mutex);
    SideStream* side = nullptr;
    int rc = side_stream(&side);
    if (rc != kOk) return rc;
    const bool split = small_rows > 0 && small_rows < lqp;
    if (split) {
      rc = launch_modup(1, s);
      if (rc != kOk) return rc;
      if (hipEventRecord(side->fork, s) != hipSuccess || hipStreamWaitEvent(side->stream, side->fork, 0) != hipSuccess) {
        set_error("fused key switch: fork failed: %s", hipGetErrorString(hipGetLastError()));
        return kLaunchFailed;
      }
      rc = launch_mac(1, side->stream);
      if (rc != kOk) return rc;
      if (hipEventRecord(side->join, side->stream) != hipSuccess) {
        set_error("fused key switch: join record failed: %s", hipGetErrorString(hipGetLastError()));
        return kLaunchFailed;
      }
      rc = launch_modup(0, s);
      if (rc == kOk) rc = launch_mac(0, s);
      if (rc != kOk) return rc;
      if (hipStreamWaitEvent(s, side->join, 0) != hipSuccess) {
        set_error("fused key switch: join failed: %s", hipGetErrorString(hipGetLastError()));
        return kLaunchFailed;
      }
    } else {
      for (int pass = 0; pass < 2 && rc == kOk; ++pass) {
        rc = launch_modup(pass, s);
        if (rc == kOk) rc = launch_mac(pass, s);
      }
      if (rc != kOk) return rc;
    }
  }

  // Both accumulators as one [2B, Lqp, N] inverse transform.
  // The tile stages of this inverse ran inside the MAC kernel.
  int st = ntt::inverse_columns_internal(workspace_acc, 2 * batch, lqp, static_cast<int>(log_n), lqp * n, n,
                                         static_cast<const NttRow*>(ntt_rows), inverse_twiddles,
                                         inverse_twiddles_shoup, level, kOutStdMac, small_rows, s);
  if (st != kOk) return st;

  ModDownArgs md;
  md.acc = workspace_acc;
  md.batch = batch;
  md.lq = lq;
  md.lqp = lqp;
  md.level = level;
  md.small_rows = small_rows;
  md.log_n = static_cast<int>(log_n);
  md.params = a.params;
  md.rows = static_cast<const NttRow*>(ntt_rows);
  md.pdig = *reinterpret_cast<const KsDigit*>(p_digit_table);
  md.p_inverse = p_inverse_q;
  md.base0 = View3(base0);
  md.base1 = View3(base1);
  md.out0 = View3(out0);
  md.out1 = View3(out1);
  md.has_base0 = base0.data != nullptr ? 1 : 0;
  md.has_base1 = base1.data != nullptr ? 1 : 0;
  if (galois_index_tables != nullptr && base0.data != nullptr) {
    // The digit rows are dead once both ModUp launches are done (the side
    // stream joined above): they take the gathered base.
    const int64_t g_blocks = (n / kBlock) * lq * batch;
    FH_CHECK_ARG(g_blocks <= 0x7fffffff, "key switch grid too large");
    hipLaunchKernelGGL(k_ks_gather_base, dim3(static_cast<uint32_t>(g_blocks)), dim3(kBlock), 0, s, View3(base0),
                       galois_index_tables, galois_sign_tables, static_cast<const NttRow*>(ntt_rows), level, lq,
                       batch, static_cast<int>(log_n), workspace_digits);
    FH_CHECK_LAUNCH("key switch base gather");
    md.base0 = View3(fh_view3{workspace_digits, lq * n, n, 1});
  }
  FH_CHECK_ARG((rescale_inverse == nullptr) == (rescale_inverse_shoup == nullptr),
               "rescale tables go together");
  FH_CHECK_ARG(rescale_inverse == nullptr || (lq >= 2 && half_drop > 0), "fused rescale needs two Q rows");
  md.rescale_inv = rescale_inverse;
  md.rescale_inv_shoup = rescale_inverse_shoup;
  md.half_drop = half_drop;
  const int64_t md_blocks = (n / kBlock) * ceil_div(lq, kModDownRowsPerThread) * 2 * batch;
  FH_CHECK_ARG(md_blocks <= 0x7fffffff, "key switch grid too large");
  const dim3 md_grid(static_cast<uint32_t>(md_blocks));
  switch (lqp - lq) {
    case 1: hipLaunchKernelGGL(k_ks_moddown_add<1>, md_grid, dim3(kBlock), 0, s, md); break;
    case 2: hipLaunchKernelGGL(k_ks_moddown_add<2>, md_grid, dim3(kBlock), 0, s, md); break;
    case 3: hipLaunchKernelGGL(k_ks_moddown_add<3>, md_grid, dim3(kBlock), 0, s, md); break;
    case 4: hipLaunchKernelGGL(k_ks_moddown_add<4>, md_grid, dim3(kBlock), 0, s, md); break;
    case 5: hipLaunchKernelGGL(k_ks_moddown_add<5>, md_grid, dim3(kBlock), 0, s, md); break;
    default: hipLaunchKernelGGL(k_ks_moddown_add<6>, md_grid, dim3(kBlock), 0, s, md); break;
  }
  FH_CHECK_LAUNCH("key switch moddown + add");
  return kOk;
}

}  // namespace
}  // namespace fh

using namespace fh;

// One fused key switch of a batched component.  See include/fhelium_amd.h.
FH_API int fh_keyswitch_fused(fh_view3 component, int64_t batch, int64_t lq, int64_t lqp,
                              int64_t level, int64_t log_n, int64_t small_rows,
                              const int64_t* digit_table,
                              int64_t digit_count,
                              const int64_t* p_digit_table,
                              const int64_t* p_inverse_q, fh_table2 rns_params, const void* ntt_rows,
                              const int64_t* forward_twiddles, const int64_t* forward_twiddles_shoup,
                              const int64_t* inverse_twiddles, const int64_t* inverse_twiddles_shoup,
                              const int64_t* key, const int64_t* key_bases, int64_t key_stride_digit,
                              int64_t key_stride_comp, int64_t key_stride_row,
                              const int64_t* galois_index_tables, const int64_t* galois_sign_tables,
                              int64_t* workspace_digits, int64_t* workspace_inter,
                              int64_t* workspace_acc, fh_view3 base0, fh_view3 base1, fh_view3 out0,
                              fh_view3 out1, int32_t mode, const int64_t* rescale_inverse,
                              const int64_t* rescale_inverse_shoup, int64_t half_drop,
                              const int64_t* seed0, const int64_t* seed1, const int64_t* seed_scale,
                              void* stream) {
  return keyswitch_fused_impl(component, batch, lq, lqp, level, log_n, small_rows, digit_table, digit_count,
                              p_digit_table, p_inverse_q, rns_params, ntt_rows, forward_twiddles,
                              forward_twiddles_shoup, inverse_twiddles, inverse_twiddles_shoup, key, key_bases,
                              key_stride_digit, key_stride_comp, key_stride_row, galois_index_tables,
                              galois_sign_tables, workspace_digits, workspace_inter, workspace_acc, base0, base1,
                              out0, out1, mode, rescale_inverse, rescale_inverse_shoup, half_drop, seed0, seed1,
                              seed_scale, /*hoist_galois=*/nullptr, /*use_side_stream=*/true, stream);
}

// Hoisted rotations of one ciphertext by several keys.  See include/fhelium_amd.h.
FH_API int fh_keyswitch_hoisted(fh_view3 component, int64_t entries, int64_t lq, int64_t lqp, int64_t level,
                                int64_t log_n, int64_t small_rows, const int64_t* digit_table,
                                int64_t digit_count, const int64_t* p_digit_table, const int64_t* p_inverse_q,
                                fh_table2 rns_params, const void* ntt_rows, const int64_t* forward_twiddles,
                                const int64_t* forward_twiddles_shoup, const int64_t* inverse_twiddles,
                                const int64_t* inverse_twiddles_shoup, const int64_t* key_bases,
                                int64_t key_stride_digit, int64_t key_stride_comp, int64_t key_stride_row,
                                const int64_t* galois_elements, const int64_t* galois_index_tables,
                                const int64_t* galois_sign_tables, int64_t* workspace_digits,
                                int64_t* workspace_inter, int64_t* workspace_acc, fh_view3 base0, fh_view3 out0,
                                fh_view3 out1, int32_t mode, void* stream) {
  FH_CHECK_ARG(galois_elements != nullptr && key_bases != nullptr && galois_index_tables != nullptr &&
                   galois_sign_tables != nullptr && base0.data != nullptr,
               "hoisted rotations need per-entry keys, Galois elements, Galois tables and the base c0");
  fh_view3 c0 = base0;
  c0.stride_batch = 0;  // every entry rotates the same c0
  const fh_view3 none{nullptr, 0, 0, 0};
  return keyswitch_fused_impl(component, entries, lq, lqp, level, log_n, small_rows, digit_table, digit_count,
                              p_digit_table, p_inverse_q, rns_params, ntt_rows, forward_twiddles,
                              forward_twiddles_shoup, inverse_twiddles, inverse_twiddles_shoup, nullptr, key_bases,
                              key_stride_digit, key_stride_comp, key_stride_row, galois_index_tables,
                              galois_sign_tables, workspace_digits, workspace_inter, workspace_acc, c0, none,
                              out0, out1, mode, nullptr, nullptr, 0, nullptr, nullptr, nullptr, galois_elements,
                              /*use_side_stream=*/true, stream);
}

// fh_keyswitch_fused with its workspaces carved from the device's static
// scratch array instead of caller memory.  See include/fhelium_amd.h.
FH_API int fh_keyswitch_fused_static(fh_view3 component, int64_t batch, int64_t lq, int64_t lqp,
                                     int64_t level, int64_t log_n, int64_t small_rows,
                                     const int64_t* digit_table, int64_t digit_count,
                                     const int64_t* p_digit_table, const int64_t* p_inverse_q,
                                     fh_table2 rns_params, const void* ntt_rows, const int64_t* forward_twiddles,
                                     const int64_t* forward_twiddles_shoup, const int64_t* inverse_twiddles,
                                     const int64_t* inverse_twiddles_shoup, const int64_t* key,
                                     const int64_t* key_bases, int64_t key_stride_digit, int64_t key_stride_comp,
                                     int64_t key_stride_row, const int64_t* galois_index_tables,
                                     const int64_t* galois_sign_tables, fh_view3 base0, fh_view3 base1,
                                     fh_view3 out0, fh_view3 out1, int32_t mode, const int64_t* rescale_inverse,
                                     const int64_t* rescale_inverse_shoup, int64_t half_drop,
                                     const int64_t* seed0, const int64_t* seed1, const int64_t* seed_scale,
                                     void* stream) {
  FH_CHECK_ARG(log_n >= 12 && log_n <= 16, "fused key switch supports logN 12..16");
  FH_CHECK_ARG(batch >= 1 && lq >= 1 && lqp > lq && digit_count >= 1 && digit_count <= 16,
               "static key switch: bad batch, row or digit counts");
  const int64_t n = int64_t{1} << log_n;
  // digits [B, lq, N], ModUp intermediate [B, dnum, lqp, N], accumulators [2, B, lqp, N]
  const int64_t digits_words = batch * lq * n;
  const int64_t inter_words = batch * digit_count * lqp * n;
  const int64_t acc_words = 2 * batch * lqp * n;
  const int64_t need = digits_words + inter_words + acc_words;
  FH_CHECK_ARG(need <= kScratchWords,
               "static key switch needs %lld workspace words (batch %lld, %lld Q rows, %lld QP rows, %lld "
               "digits, N %lld); the static scratch holds %lld (64 MiB)",
               (long long)need, (long long)batch, (long long)lq, (long long)lqp, (long long)digit_count,
               (long long)n, (long long)kScratchWords);
  int64_t* base = nullptr;
  const int rc = scratch_address(&base);
  if (rc != kOk) return rc;
  // Every workspace is fully written by the launch that produces it before
  // any launch reads it, so the scratch needs no clearing between calls.
  return keyswitch_fused_impl(component, batch, lq, lqp, level, log_n, small_rows, digit_table, digit_count,
                              p_digit_table, p_inverse_q, rns_params, ntt_rows, forward_twiddles,
                              forward_twiddles_shoup, inverse_twiddles, inverse_twiddles_shoup, key, key_bases,
                              key_stride_digit, key_stride_comp, key_stride_row, galois_index_tables,
                              galois_sign_tables, base, base + digits_words, base + digits_words + inter_words,
                              base0, base1, out0, out1, mode, rescale_inverse, rescale_inverse_shoup, half_drop,
                              seed0, seed1, seed_scale, /*hoist_galois=*/nullptr, /*use_side_stream=*/false, stream);
}

// Tensor product of two coefficient-domain ciphertexts, returned in the
// coefficient domain.  See include/fhelium_amd.h.
FH_API int fh_ckks_hmult_tensor_coefficient(const int64_t* a, const int64_t* b, int64_t a_comp_stride,
                                            int64_t a_batch_stride, int64_t a_limb_stride,
                                            int64_t b_comp_stride, int64_t b_batch_stride,
                                            int64_t b_limb_stride, int64_t batch, int64_t limbs,
                                            int64_t log_n, int64_t row_start, int64_t small_rows,
                                            fh_table2 rns_params, const void* ntt_rows,
                                            const int64_t* forward_twiddles,
                                            const int64_t* forward_twiddles_shoup,
                                            const int64_t* inverse_twiddles,
                                            const int64_t* inverse_twiddles_shoup, int64_t* workspace,
                                            int64_t* out, int32_t ntt_d01, void* stream) {
  FH_CHECK_ARG(log_n >= 12 && log_n <= 16, "fused tensor product supports logN 12..16");
  FH_CHECK_ARG(ntt_d01 == 0 || ntt_d01 == 1, "ntt_d01 is 0 or 1");
  FH_CHECK_ARG(batch >= 1 && limbs >= 1 && row_start >= 0, "bad batch, limb count or row start");
  FH_CHECK_ARG(small_rows >= 0 && small_rows <= limbs, "bad binary64 row count");
  FH_CHECK_ARG(a && b && rns_params.data && ntt_rows && forward_twiddles && forward_twiddles_shoup &&
                   inverse_twiddles && inverse_twiddles_shoup && workspace && out,
               "fused tensor product: null operand");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t n = int64_t{1} << log_n;
  const int64_t comp = batch * limbs * n;
  const NttRow* rows = static_cast<const NttRow*>(ntt_rows);
  // Column passes of a0, a1 (workspace components 0, 1) and b0, b1 (2, 3):
  // one launch per operand when its two components are one [2B] batch.
  const int64_t* ops[2] = {a, b};
  const int64_t cs[2] = {a_comp_stride, b_comp_stride};
  const int64_t bs[2] = {a_batch_stride, b_batch_stride};
  const int64_t ls[2] = {a_limb_stride, b_limb_stride};
  for (int o = 0; o < 2; ++o) {
    int64_t* dst = workspace + 2 * o * comp;
    if (cs[o] == batch * bs[o]) {
      int st = ntt::forward_columns_internal(dst, ops[o], 2 * batch, limbs, static_cast<int>(log_n), limbs * n, n,
                                             bs[o], ls[o], rows, forward_twiddles, forward_twiddles_shoup,
                                             row_start, true, s);
      if (st != kOk) return st;
      continue;
    }
    for (int c = 0; c < 2; ++c) {
      int st = ntt::forward_columns_internal(dst + c * comp, ops[o] + c * cs[o], batch, limbs,
                                             static_cast<int>(log_n), limbs * n, n, bs[o], ls[o], rows,
                                             forward_twiddles, forward_twiddles_shoup, row_start, true, s);
      if (st != kOk) return st;
    }
  }
  HmArgs h;
  h.ws = workspace;
  h.out = out;
  h.batch = batch;
  h.limbs = limbs;
  h.row_start = row_start;
  h.log_n = static_cast<int>(log_n);
  h.params = ParamView{rns_params.data, rns_params.stride_row, rns_params.stride_col};
  h.rows = rows;
  h.fwd_w = forward_twiddles;
  h.fwd_wp = forward_twiddles_shoup;
  h.inv_w = inverse_twiddles;
  h.inv_wp = inverse_twiddles_shoup;
  const int64_t groups = n / (kTile * kMacTilesPerBlock);
  FH_CHECK_ARG(groups * limbs * batch <= 0x7fffffff, "fused tensor product grid too large");
  if (small_rows > 0) {
    const dim3 grid(static_cast<uint32_t>(groups * small_rows * batch));
    if (ntt_d01) {
      hipLaunchKernelGGL((k_hmult_tiles<true, true>), grid, dim3(kMacBlock), 0, s, h, groups, int64_t{0}, small_rows);
    } else {
      hipLaunchKernelGGL((k_hmult_tiles<true, false>), grid, dim3(kMacBlock), 0, s, h, groups, int64_t{0}, small_rows);
    }
    FH_CHECK_LAUNCH("fused tensor product tiles (binary64 rows)");
  }
  if (small_rows < limbs) {
    const dim3 grid(static_cast<uint32_t>(groups * (limbs - small_rows) * batch));
    if (ntt_d01) {
      hipLaunchKernelGGL((k_hmult_tiles<false, true>), grid, dim3(kMacBlock), 0, s, h, groups, small_rows,
                         limbs - small_rows);
    } else {
      hipLaunchKernelGGL((k_hmult_tiles<false, false>), grid, dim3(kMacBlock), 0, s, h, groups, small_rows,
                         limbs - small_rows);
    }
    FH_CHECK_LAUNCH("fused tensor product tiles (integer rows)");
  }
  // d2 (and d0, d1 unless they stay in the NTT domain) back to coefficients
  const int64_t inv_comps = ntt_d01 ? 1 : 3;
  return ntt::inverse_columns_internal(out + (3 - inv_comps) * comp, inv_comps * batch, limbs,
                                       static_cast<int>(log_n), limbs * n, n, rows, inverse_twiddles,
                                       inverse_twiddles_shoup, row_start, kOutStdMac, 0, s);
}
