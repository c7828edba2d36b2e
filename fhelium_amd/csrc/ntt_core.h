// Shared NTT arithmetic for the gfx950 kernels (ntt.hip, keyswitch.hip):
// launch constants and the three butterfly arithmetic paths.  See ntt.hip
// for the transform semantics and fh_common.h for the exactness argument of
// the binary64 and Shoup modular products.
#pragma once

#include "fh_common.h"

namespace fh {
namespace ntt {
constexpr int kBlock = 256;
constexpr int kTileLog = 8;
constexpr int kTile = 1 << kTileLog;      // contiguous tile of the tile pass
constexpr int kTilesPerBlock = 16;

// Block i runs on XCD i % 8 (a GPU chiplet with its own L2); give each XCD
// one contiguous range of work items so consecutive items on an XCD share
// the data they read again (key tiles, twiddles) in its L2.  The map is a
// bijection of [0, total).  Block-index arithmetic is 32-bit (grids are
// below 2^31 blocks): 64-bit division expands to a long instruction
// sequence on the GPU.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t i, uint32_t total) {
  constexpr uint32_t kXcds = 8;
  const uint32_t x = i % kXcds;
  const uint32_t s = i / kXcds;
  const uint32_t q = total / kXcds;
  const uint32_t rem = total % kXcds;
  return x * q + (x < rem ? x : rem) + s;
}
constexpr int kTilePitch = kTile + kTile / 16;  // one pad word per 16

__device__ __forceinline__ int tile_pad(int x) { return x + (x >> 4); }

// Ordering for an LDS exchange among threads of one wave64 (a tile-pass
// tile's threads): LDS operations of one wave complete in order, so a
// wave-level barrier that keeps the compiler from moving memory operations
// across it replaces a block barrier over all the block's waves.
__device__ __forceinline__ void tile_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inverse output representations (fh_ntt_inverse).
// kOutStdMac: standard canonical output of accumulators whose binary64 rows
// hold x*R mod q (a key MAC that skipped REDC's R^{-1}); integer rows are
// ordinary Montgomery values and get the plain standard output.
enum InvOut : int { kOutMontgomery = 0, kOutStdLazy = 1, kOutStd = 2, kOutCentered = 3, kOutStdMac = 4 };

__device__ __forceinline__ int64_t center_canonical(int64_t r, uint64_t q) {
  const int64_t qq = static_cast<int64_t>(q);
  return r <= (qq >> 1) ? r : r - qq;
}

// ---------------------------------------------------------------------------
// Arithmetic paths.  Interface:
//   in(raw)          input residue (lazy [0, 2q)) -> working value
//   mid/unmid        working value <-> 64-bit intermediate between passes
//   out_fwd(v)       final forward value -> int64 residue
//   out_inv(v, f)    N^{-1} scaling and representation f -> int64 residue
//   fwd/inv          butterflies; mul_const for to-Montgomery scaling.
// ---------------------------------------------------------------------------

// Exact binary64 path (q < 2^42): reference lazy semantics, see fh_common.h.
struct F64Exact {
  using V = double;
  struct Tw {
    double w;
  };
  double q, twoq, qinv;
  const NttRow* row;
  __device__ __forceinline__ explicit F64Exact(const NttRow& r)
      : q(r.qd), twoq(2.0 * r.qd), qinv(r.qinv), row(&r) {}
  __device__ __forceinline__ Tw tw(const int64_t* w, const int64_t*, int64_t idx) const {
    return Tw{__longlong_as_double(w[idx])};
  }
  // twiddle from its table words (w, Shoup companion)
  __device__ __forceinline__ Tw tw_make(int64_t w, int64_t) const { return Tw{__longlong_as_double(w)}; }
  __device__ __forceinline__ Tw konst(uint64_t c, uint64_t) const {
    return Tw{static_cast<double>(c)};
  }
  __device__ __forceinline__ V in(int64_t raw) const { return u64_to_f64(static_cast<uint64_t>(raw)); }
  __device__ __forceinline__ int64_t mid(V v) const { return __double_as_longlong(v); }
  __device__ __forceinline__ V unmid(int64_t b) const { return __longlong_as_double(b); }
  __device__ __forceinline__ V mul(Tw w, V x) const { return f64_mulmod(w.w, x, q, qinv); }
  __device__ __forceinline__ V add(V a, V b) const {
    const V s = a + b;
    return s >= twoq ? s - twoq : s;
  }
  __device__ __forceinline__ V sub(V a, V b) const {
    const V d = (a + twoq) - b;
    return d >= twoq ? d - twoq : d;
  }
  __device__ __forceinline__ void fwd(V& u, V& l, Tw w) const {
    const V v = mul(w, l);
    const V e = add(u, v);
    l = sub(u, v);
    u = e;
  }
  __device__ __forceinline__ void inv(V& u, V& l, Tw w) const {
    const V d = sub(u, l);
    u = add(u, l);
    l = mul(w, d);
  }
  __device__ __forceinline__ V to_mont(V x) const { return mul(konst(row->to_mont, 0), x); }
  __device__ __forceinline__ int64_t out_fwd(V v) const {
    return static_cast<int64_t>(f64_to_u64(v));
  }
  __device__ __forceinline__ int64_t out_inv(V v, int form) const {
    const Tw c = form == kOutMontgomery ? konst(row->ninv_mont, 0)
                 : form == kOutStdMac      ? konst(row->ninv_std_rinv, 0)
                                           : konst(row->ninv_std, 0);
    int64_t r = static_cast<int64_t>(f64_to_u64(mul(c, v)));
    return form == kOutCentered ? center_canonical(r, row->q) : r;
  }
};

// Fast binary64 path (q < 2^42): signed values, products reduced to
// (-0.52q, 0.52q) without the final correction.
struct F64Fast {
  using V = double;
  using Tw = F64Exact::Tw;
  double q, qinv;
  const NttRow* row;
  __device__ __forceinline__ explicit F64Fast(const NttRow& r) : q(r.qd), qinv(r.qinv), row(&r) {}
  __device__ __forceinline__ Tw tw(const int64_t* w, const int64_t*, int64_t idx) const {
    return Tw{__longlong_as_double(w[idx])};
  }
  __device__ __forceinline__ Tw tw_make(int64_t w, int64_t) const { return Tw{__longlong_as_double(w)}; }
  __device__ __forceinline__ Tw konst(uint64_t c, uint64_t) const {
    return Tw{static_cast<double>(c)};
  }
  __device__ __forceinline__ V in(int64_t raw) const { return u64_to_f64(static_cast<uint64_t>(raw)); }
  __device__ __forceinline__ int64_t mid(V v) const { return __double_as_longlong(v); }
  __device__ __forceinline__ V unmid(int64_t b) const { return __longlong_as_double(b); }
  // w*x - round(w*x/q) q, exact; |result| < 0.52 q for |x| < 2^45.
  __device__ __forceinline__ V mul(Tw w, V x) const { return f64_mulmod_signed(w.w, x, q, qinv); }
  // x - round(x/q) q, exact for |x| < 2^51; |result| <= q/2 (+ulp-level).
  __device__ __forceinline__ V reduce(V x) const {
    const double t = __builtin_fma(x, qinv, kRoundMagic) - kRoundMagic;
    return __builtin_fma(-t, q, x);
  }
  __device__ __forceinline__ V canon(V x) const {
    const V r = reduce(x);
    return r < 0.0 ? r + q : r;
  }
  __device__ __forceinline__ void fwd(V& u, V& l, Tw w) const {
    const V v = mul(w, l);
    l = u - v;
    u = u + v;
  }
  __device__ __forceinline__ void inv(V& u, V& l, Tw w) const {
    const V d = u - l;
    u = reduce(u + l);
    l = mul(w, d);
  }
  __device__ __forceinline__ V to_mont(V x) const { return mul(konst(row->to_mont, 0), x); }
  __device__ __forceinline__ int64_t out_fwd(V v) const {
    return static_cast<int64_t>(f64_to_u64(canon(v)));
  }
  __device__ __forceinline__ int64_t out_inv(V v, int form) const {
    const Tw c = form == kOutMontgomery ? konst(row->ninv_mont, 0)
                 : form == kOutStdMac      ? konst(row->ninv_std_rinv, 0)
                                           : konst(row->ninv_std, 0);
    int64_t r = static_cast<int64_t>(f64_to_u64(canon(mul(c, v))));
    return form == kOutCentered ? center_canonical(r, row->q) : r;
  }
};

// Integer Shoup path (any q < 2^62), reference lazy semantics (exact mode).
struct U64Path {
  using V = uint64_t;
  struct Tw {
    uint64_t w, wp;
  };
  uint64_t q, twoq;
  const NttRow* row;
  __device__ __forceinline__ explicit U64Path(const NttRow& r) : q(r.q), twoq(r.twice_q), row(&r) {}
  __device__ __forceinline__ Tw tw(const int64_t* w, const int64_t* wp, int64_t idx) const {
    return Tw{static_cast<uint64_t>(w[idx]), static_cast<uint64_t>(wp[idx])};
  }
  __device__ __forceinline__ Tw tw_make(int64_t w, int64_t wp) const {
    return Tw{static_cast<uint64_t>(w), static_cast<uint64_t>(wp)};
  }
  __device__ __forceinline__ Tw konst(uint64_t c, uint64_t cp) const { return Tw{c, cp}; }
  __device__ __forceinline__ V in(int64_t raw) const { return static_cast<uint64_t>(raw); }
  __device__ __forceinline__ int64_t mid(V v) const { return static_cast<int64_t>(v); }
  __device__ __forceinline__ V unmid(int64_t b) const { return static_cast<uint64_t>(b); }
  __device__ __forceinline__ V mul(Tw w, V x) const { return shoup_mulmod(w.w, w.wp, x, q); }
  __device__ __forceinline__ V add(V a, V b) const { return csub(a + b, twoq); }
  // a + 2q - b with 2q - b = (2q + 1) + ~b (no carry chain)
  __device__ __forceinline__ V sub(V a, V b) const { return csub(a + (twoq + 1u) + ~b, twoq); }
  __device__ __forceinline__ void fwd(V& u, V& l, Tw w) const {
    const V v = mul(w, l);
    const V e = add(u, v);
    l = sub(u, v);
    u = e;
  }
  __device__ __forceinline__ void inv(V& u, V& l, Tw w) const {
    const V d = sub(u, l);
    u = add(u, l);
    l = mul(w, d);
  }
  __device__ __forceinline__ V to_mont(V x) const {
    return mul(konst(row->to_mont, row->to_mont_p), x);
  }
  __device__ __forceinline__ int64_t out_fwd(V v) const { return static_cast<int64_t>(v); }
  __device__ __forceinline__ int64_t out_inv(V v, int form) const {
    const Tw c = form == kOutMontgomery ? konst(row->ninv_mont, row->ninv_mont_p)
                                        : konst(row->ninv_std, row->ninv_std_p);
    int64_t r = static_cast<int64_t>(mul(c, v));
    return form == kOutCentered ? center_canonical(r, q) : r;
  }
};

// Integer Shoup path for fast arithmetic (q < 2^60), with Harvey's lazy
// butterflies: forward values stay in [0, 4q), inverse values in [0, 2q),
// and the Shoup product w x - hi64(w' x) q is left in [0, 2q) (any x below
// 2^64, w < q).  One conditional subtraction per butterfly instead of three;
// forward and inverse outputs are canonical, so they equal the reference's
// values modulo q exactly as U64Path's do.
struct U64Fast {
  using V = uint64_t;
  struct Tw {
    uint64_t w, wp;
  };
  uint64_t q, twoq;
  const NttRow* row;
  __device__ __forceinline__ explicit U64Fast(const NttRow& r) : q(r.q), twoq(r.twice_q), row(&r) {}
  __device__ __forceinline__ Tw tw(const int64_t* w, const int64_t* wp, int64_t idx) const {
    return Tw{static_cast<uint64_t>(w[idx]), static_cast<uint64_t>(wp[idx])};
  }
  __device__ __forceinline__ Tw tw_make(int64_t w, int64_t wp) const {
    return Tw{static_cast<uint64_t>(w), static_cast<uint64_t>(wp)};
  }
  __device__ __forceinline__ Tw konst(uint64_t c, uint64_t cp) const { return Tw{c, cp}; }
  __device__ __forceinline__ V in(int64_t raw) const { return static_cast<uint64_t>(raw); }
  __device__ __forceinline__ int64_t mid(V v) const { return static_cast<int64_t>(v); }
  __device__ __forceinline__ V unmid(int64_t b) const { return static_cast<uint64_t>(b); }
  // w x - t q with t = hi64(w' x), formed as w x + t (2^64 - q): in [0, 2q)
  __device__ __forceinline__ V mul_lazy(Tw w, V x) const {
    const uint64_t t = __umul64hi(w.wp, x);
    return w.w * x + t * opaque(uint64_t{0} - q);
  }
  __device__ __forceinline__ V mul(Tw w, V x) const { return csub(mul_lazy(w, x), q); }
  // [0, 4q) -> [0, 2q)
  __device__ __forceinline__ V half(V v) const { return csub(v, twoq); }
  __device__ __forceinline__ void fwd(V& u, V& l, Tw w) const {
    const V a = half(u);
    const V v = mul_lazy(w, l);
    u = a + v;
    l = a + (twoq + 1u) + ~v;  // a - v + 2q without a borrow chain
  }
  __device__ __forceinline__ void inv(V& u, V& l, Tw w) const {
    const V d = u + (twoq + 1u) + ~l;
    u = half(u + l);
    l = mul_lazy(w, d);
  }
  __device__ __forceinline__ V to_mont(V x) const {
    return mul_lazy(konst(row->to_mont, row->to_mont_p), x);
  }
  __device__ __forceinline__ int64_t out_fwd(V v) const { return static_cast<int64_t>(csub(half(v), q)); }
  __device__ __forceinline__ int64_t out_inv(V v, int form) const {
    const Tw c = form == kOutMontgomery ? konst(row->ninv_mont, row->ninv_mont_p)
                                        : konst(row->ninv_std, row->ninv_std_p);
    int64_t r = static_cast<int64_t>(mul(c, v));
    return form == kOutCentered ? center_canonical(r, q) : r;
  }
};

// Packed binary64 rows.  With fast arithmetic the binary64 rows of the key
// switch's intermediates (the ModUp output and the accumulators) travel as
// canonical residues (< 2^42) in a u32 plane and a u16 plane, 6 bytes per
// coefficient instead of 8.  In the [rows, N] int64 slab of one (entry,
// digit) or (component, entry) the u32 plane starts at the slab and the u16
// plane follows it; both fit inside the slab's first small_rows rows, so the
// integer rows keep their int64 layout after them.  Within a row the
// coefficients are stored in 8 x 16 blocks of the 2^(logN-8) x 256 matrix
// (row i = tile, column c): a tile-pass thread reads or writes 8 or 16
// consecutive coefficients of a 16-column run, and one 16-column block row
// is what a column-pass wave covers, so both passes move whole cache lines.
struct PackedInter {
  uint32_t* lo;
  uint16_t* hi;
};
__device__ __forceinline__ PackedInter packed_slab(int64_t* slab, int64_t small_rows, int64_t n) {
  uint32_t* lo = reinterpret_cast<uint32_t*>(slab);
  return PackedInter{lo, reinterpret_cast<uint16_t*>(lo + small_rows * n)};
}
__device__ __forceinline__ int64_t packed_index(int64_t i, int64_t c) {
  return ((i >> 3) << 11) | ((c >> 4) << 7) | ((i & 7) << 4) | (c & 15);
}
// Inverse of packed_index: the coefficient (i * 256 + c) at packed position p.
__device__ __forceinline__ int64_t packed_element(int64_t p) {
  const int64_t i = ((p >> 11) << 3) | ((p >> 4) & 7);
  const int64_t c = (((p >> 7) & 15) << 4) | (p & 15);
  return (i << 8) | c;
}
// Element idx (row offset plus packed_index) of a packed slab.
__device__ __forceinline__ uint64_t packed_load(const PackedInter& pk, int64_t idx) {
  return static_cast<uint64_t>(pk.lo[idx]) | static_cast<uint64_t>(pk.hi[idx]) << 32;
}
__device__ __forceinline__ void packed_store(const PackedInter& pk, int64_t idx, uint64_t v) {
  pk.lo[idx] = static_cast<uint32_t>(v);
  pk.hi[idx] = static_cast<uint16_t>(v >> 32);
}

// Inverse transform launcher shared with keyswitch.hip (fast arithmetic).
// With in_mid the input rows hold working-format values (binary64 bits for
// rows below 2^42, u64 otherwise) instead of int64 residues.
int inverse_internal(int64_t* data, int64_t batch, int64_t limbs, int log_n, int64_t batch_stride,
                     int64_t limb_stride, const NttRow* rows, const int64_t* tw, const int64_t* twp,
                     int64_t row_start, int output_form, bool in_mid, hipStream_t stream);
// Only the forward column pass (logN 12..16, fast arithmetic) from src rows
// into dst rows, leaving working-format values for a fused tile pass.
int forward_columns_internal(int64_t* dst, const int64_t* src, int64_t batch, int64_t limbs, int log_n,
                             int64_t dst_batch_stride, int64_t dst_limb_stride, int64_t src_batch_stride,
                             int64_t src_limb_stride, const NttRow* rows, const int64_t* tw,
                             const int64_t* twp, int64_t row_start, bool to_montgomery, hipStream_t stream);
// Only the column pass of that inverse (logN 12..16): for rows whose tile
// stages were already applied, holding working-format values.
// packed_rows > 0: limbs [0, packed_rows) of each batch entry are binary64
// rows held packed (PackedInter over the entry's slab, input and output).
int inverse_columns_internal(int64_t* data, int64_t batch, int64_t limbs, int log_n,
                             int64_t batch_stride, int64_t limb_stride, const NttRow* rows,
                             const int64_t* tw, const int64_t* twp, int64_t row_start,
                             int output_form, int64_t packed_rows, hipStream_t stream);

}  // namespace ntt
}  // namespace fh
