// Negacyclic NTT / inverse NTT for gfx950.
//
// Semantics (reference csrc/ops/ntt/cpu/indexed_ntt_cpu.cpp:34-313 and
// fhelium/engine/ntt/plans/indexed_radix2.py:58-126): forward stage s pairs
// (j, j + N/2^{s+1}) with twiddle psi^{br(2^s + group)} (Cooley-Tukey DIF,
// natural input, bit-reversed output); inverse stage s pairs (j, j + 2^s) with
// psi^{-br(N/2^{s+1} + group)} (Gentleman-Sande DIT), then scales by N^{-1}.
//
// Two arithmetic modes:
//  * exact: every butterfly is computed as the reference does it (canonical
//    product, lazy add/sub modulo 2q), so the lazy forward outputs are
//    bit-identical to the reference CPU backend, not merely congruent.
//  * fast (default for the engine): for primes below 2^42 the residues live
//    in signed binary64 during the transform: forward butterflies are
//    U +/- w*L with the product reduced to (-0.52q, 0.52q) and no other
//    reduction (|x| stays below 11q < 2^45.5 over 16 stages); inverse
//    butterflies reduce U + L by one rounding step.  Outputs are canonical
//    [0, q): the inverse outputs equal the reference's exactly (they are
//    canonical there too), forward outputs equal the reference's lazy
//    values modulo q.  Larger primes use Shoup products with Harvey's lazy
//    butterflies (forward values below 4q, inverse values below 2q) and
//    canonical outputs in this mode.
//
// Execution (N = 2^L, 12 <= L <= 16): two fused passes.
//   * "column" pass: the first L-8 forward stages (last L-8 inverse stages)
//     pair elements 256*i + c that share the column c; a block owns a few
//     adjacent columns (coalesced 128-byte segments) and all 2^{L-8} rows.
//   * "tile" pass: the last 8 forward stages (first 8 inverse stages) act
//     inside contiguous 256-element tiles; a block owns 16 tiles.
// Inside a pass each thread holds 16 elements in registers and performs four
// radix-2 stages on them (a radix-16 round); rounds exchange data through
// LDS.  The intermediate between the passes stays in the arithmetic's own
// format (binary64 bits or u64), in place in the operand.  The tile passes
// run on a one-dimensional grid over (entry, tile group, limb) whose block
// order is remapped so that consecutive work items share an XCD's L2; the
// host may split a call into chunks of entries on several streams (a launch
// schedule chosen per call).  Rings with logN < 12 (tests) use one launch
// per stage.
#include "fh_common.h"

#include "ntt_core.h"

namespace fh {
namespace {

using namespace ::fh::ntt;

// ---------------------------------------------------------------------------
// Pass bodies.  `base` is the operand row (int64 residues or intermediates).
// ---------------------------------------------------------------------------

// Forward column pass: stages 0 .. S1-1 over rows i of columns c
// (element 256*i + c).  Block: C = 256/T columns x T = 2^{S1-4} threads.
// Reads input residues, writes intermediates.
template <int S1, class P>
__device__ __forceinline__ void fwd_columns(const P& p, int64_t* base, const int64_t* src,
                                            const int64_t* tw, const int64_t* twp, bool to_mont,
                                            int64_t* lds) {
  constexpr int R1 = 1 << S1;
  constexpr int T = R1 / 16;
  constexpr int C = kBlock / T;
  using V = typename P::V;
  const int c = threadIdx.x % C;
  const int t = threadIdx.x / C;
  const int col = blockIdx.x * C + c;
  V v[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = p.in(src[static_cast<int64_t>(t + T * k) * kTile + col]);
  if (to_mont) {
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = p.to_mont(v[k]);
  }
  // Round A: stages 0..3 pair row bits S1-1..S1-4, i.e. bits 3..0 of k.
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int half = 8 >> a;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & half) continue;
      p.fwd(v[k], v[k + half], p.tw(tw, twp, (1 << a) + (k >> (4 - a))));
    }
  }
  if constexpr (S1 > 4) {
#pragma unroll
    for (int k = 0; k < 16; ++k) lds[(t + T * k) * C + c] = p.mid(v[k]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = p.unmid(lds[(16 * t + k) * C + c]);
    // Round B: stages 4..S1-1 pair row bits S1-5..0 (bits of k).
#pragma unroll
    for (int s = 4; s < S1; ++s) {
      const int half = 1 << (S1 - 1 - s);
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        if (k & half) continue;
        p.fwd(v[k], v[k + half], p.tw(tw, twp, (1 << s) + ((16 * t + k) >> (S1 - s))));
      }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) base[static_cast<int64_t>(16 * t + k) * kTile + col] = p.mid(v[k]);
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) base[static_cast<int64_t>(t + T * k) * kTile + col] = p.mid(v[k]);
  }
}

// Forward tile pass: the last 8 stages inside 256-element tiles.  Reads
// intermediates, writes output residues.
template <class P>
__device__ __forceinline__ void fwd_tiles(const P& p, int log_n, int g, int64_t* base, const int64_t* tw,
                                          const int64_t* twp, int64_t* lds) {
  using V = typename P::V;
  const int m = threadIdx.x / 16;
  const int t = threadIdx.x % 16;
  const int r = g * kTilesPerBlock + m;  // tile index in the row
  int64_t* tile = base + static_cast<int64_t>(r) * kTile;
  int64_t* sh = lds + m * kTilePitch;
  const int s0 = log_n - kTileLog;
  V v[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = p.unmid(tile[t + 16 * k]);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int half = 8 >> a;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & half) continue;
      p.fwd(v[k], v[k + half], p.tw(tw, twp, (1 << (s0 + a)) + (r << a) + (k >> (4 - a))));
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) sh[tile_pad(t + 16 * k)] = p.mid(v[k]);
  tile_sync();
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = p.unmid(sh[17 * t + k]);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int half = 8 >> a;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & half) continue;
      p.fwd(v[k], v[k + half],
            p.tw(tw, twp, (1 << (s0 + 4 + a)) + (r << (4 + a)) + (t << a) + (k >> (4 - a))));
    }
  }
  // Convert, then transpose back through LDS so the stores are coalesced.
  // Each thread rewrites exactly the words it just read: no barrier needed.
#pragma unroll
  for (int k = 0; k < 16; ++k) sh[17 * t + k] = p.out_fwd(v[k]);
  tile_sync();
#pragma unroll
  for (int k = 0; k < 16; ++k) tile[t + 16 * k] = sh[tile_pad(t + 16 * k)];
}

static_assert(64 % (kBlock / kTilesPerBlock) == 0, "a tile's threads must lie in one wave");

// Inverse tile pass: stages 0..7 inside 256-element tiles.  Reads input
// residues, writes intermediates.
template <class P>
__device__ __forceinline__ void inv_tiles(const P& p, int log_n, int g, int64_t* base, const int64_t* src,
                                          const int64_t* tw, const int64_t* twp, bool in_mid,
                                          int64_t* lds) {
  using V = typename P::V;
  const int m = threadIdx.x / 16;
  const int t = threadIdx.x % 16;
  const int r = g * kTilesPerBlock + m;
  int64_t* tile = base + static_cast<int64_t>(r) * kTile;
  const int64_t* in_tile = src + static_cast<int64_t>(r) * kTile;
  int64_t* sh = lds + m * kTilePitch;
  const int n = 1 << log_n;
#pragma unroll
  for (int k = 0; k < 16; ++k) sh[tile_pad(t + 16 * k)] = in_tile[t + 16 * k];
  tile_sync();
  V v[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = in_mid ? p.unmid(sh[17 * t + k]) : p.in(sh[17 * t + k]);
  // Round A: stages 0..3 pair tile bits 0..3 (bits of k).
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int half = 1 << a;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & half) continue;
      p.inv(v[k], v[k + half],
            p.tw(tw, twp, (n >> (a + 1)) + (r << (7 - a)) + (t << (3 - a)) + (k >> (a + 1))));
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) sh[17 * t + k] = p.mid(v[k]);
  tile_sync();
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = p.unmid(sh[tile_pad(t + 16 * k)]);
  // Round B: stages 4..7 pair tile bits 4..7 (bits of k).
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int half = 1 << a;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & half) continue;
      p.inv(v[k], v[k + half], p.tw(tw, twp, (n >> (a + 5)) + (r << (3 - a)) + (k >> (a + 1))));
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) tile[t + 16 * k] = p.mid(v[k]);
}

// Inverse column pass: stages 8 .. L-1 pairing row bits 0 .. S1-1, then the
// N^{-1} normalisation and output representation.  Reads intermediates,
// writes output residues through load(i, c) / store(i, c, value) of element
// i * 256 + c of the row.
template <int S1, class P, class Load, class Store>
__device__ __forceinline__ void inv_columns(const P& p, Load load, Store store, const int64_t* tw,
                                            const int64_t* twp, int out, int64_t* lds) {
  constexpr int R1 = 1 << S1;
  constexpr int T = R1 / 16;
  constexpr int C = kBlock / T;
  constexpr int L = S1 + kTileLog;
  using V = typename P::V;
  const int c = threadIdx.x % C;
  const int t = threadIdx.x / C;
  const int col = blockIdx.x * C + c;
  V v[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = p.unmid(load(16 * t + k, col));
  // Round A: row bits 0..3 (bits of k); stage s = 8 + a.
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int half = 1 << a;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & half) continue;
      p.inv(v[k], v[k + half],
            p.tw(tw, twp, ((1 << L) >> (kTileLog + a + 1)) + ((16 * t + k) >> (a + 1))));
    }
  }
  if constexpr (S1 > 4) {
#pragma unroll
    for (int k = 0; k < 16; ++k) lds[(16 * t + k) * C + c] = p.mid(v[k]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = p.unmid(lds[(t + T * k) * C + c]);
    // Round B: row bits 4 .. S1-1, which are bits (b - (S1-4)) of k.
#pragma unroll
    for (int b = 4; b < S1; ++b) {
      const int half = 1 << (b - (S1 - 4));
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        if (k & half) continue;
        p.inv(v[k], v[k + half],
              p.tw(tw, twp, ((1 << L) >> (kTileLog + b + 1)) + ((t + T * k) >> (b + 1))));
      }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) store(t + T * k, col, p.out_inv(v[k], out));
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) store(16 * t + k, col, p.out_inv(v[k], out));
  }
}

struct NttArgs {
  int64_t* data;
  int64_t batch_stride;
  int64_t limb_stride;
  int64_t limbs;        // limbs per batch entry (to split flat rows)
  int64_t row_offset;   // first flat row (batch * limbs + limb) of this launch
  const NttRow* rows;
  const int64_t* tw;    // [rows, N]
  const int64_t* twp;   // [rows, N]
  int64_t row_start;    // parameter row of limb 0
  int log_n;
  int form;             // forward: to-Montgomery flag; inverse: InvOut
  int in_mid;           // inverse only: input already in the working format
  const int64_t* src;   // input rows of the first pass (null: in place)
  int64_t src_batch_stride;
  int64_t src_limb_stride;
  int64_t packed_rows;  // inverse column pass: leading packed binary64 limbs
  // > 0: the tile passes run on a 1-D grid over (entry, tile group, limb),
  // remapped so each XCD walks every entry of a (limb, tile group) in turn
  // and reads that tile group's twiddles from its L2; tile_batch is the
  // entry count (the launch covers all entries' rows).  0: 2-D grid of
  // (tile group, flat row from row_offset).
  int64_t tile_batch;
};

struct RowRef {
  int64_t* base;
  const NttRow* row;
  const int64_t* tw;
  const int64_t* twp;
  const int64_t* src;   // input row of the first pass
  int64_t* entry;       // limb 0 of the row's batch entry
  int64_t limb;
};

__device__ __forceinline__ RowRef row_ref(const NttArgs& a, uint32_t flat) {
  // 32-bit division: flat rows stay below 2^31 (64-bit division is a long
  // instruction sequence on the GPU)
  const int64_t b = flat / static_cast<uint32_t>(a.limbs);
  const int64_t l = flat - static_cast<uint32_t>(b) * static_cast<uint32_t>(a.limbs);
  const int64_t prow = a.row_start + l;
  const int64_t n = int64_t{1} << a.log_n;
  int64_t* base = a.data + b * a.batch_stride + l * a.limb_stride;
  const int64_t* src = a.src ? a.src + b * a.src_batch_stride + l * a.src_limb_stride : base;
  return RowRef{base, a.rows + prow, a.tw + prow * n, a.twp + prow * n, src, a.data + b * a.batch_stride, l};
}

// Row of a column-pass block (2-D grid: column group, flat row).
__device__ __forceinline__ RowRef row_ref(const NttArgs& a) {
  return row_ref(a, static_cast<uint32_t>(a.row_offset) + blockIdx.y);
}

// Flat row and tile group of a tile-pass block.
struct TileItem {
  uint32_t flat;
  int g;
};
__device__ __forceinline__ TileItem tile_item(const NttArgs& a) {
  if (a.tile_batch > 0) {
    const uint32_t groups = static_cast<uint32_t>((int64_t{1} << a.log_n) / (kTile * kTilesPerBlock));
    const uint32_t entries = static_cast<uint32_t>(a.tile_batch);
    const uint32_t limbs = static_cast<uint32_t>(a.limbs);
    const uint32_t w = xcd_remap(blockIdx.x, groups * entries * limbs);
    const uint32_t b = w % entries;
    const uint32_t g = (w / entries) % groups;
    const uint32_t l = w / (entries * groups);
    return TileItem{b * limbs + l, static_cast<int>(g)};
  }
  return TileItem{static_cast<uint32_t>(a.row_offset) + blockIdx.y, static_cast<int>(blockIdx.x)};
}

enum Mode : int { kExact = 0, kFast = 1 };

#define FH_DISPATCH_PATH(MODE, ROW, CALL)                  \
  do {                                                     \
    if ((ROW).small) {                                     \
      if (MODE == kFast) {                                 \
        F64Fast p(ROW);                                    \
        CALL;                                              \
      } else {                                             \
        F64Exact p(ROW);                                   \
        CALL;                                              \
      }                                                    \
    } else {                                               \
      if (MODE == kFast) {                                 \
        U64Fast p(ROW);                                    \
        CALL;                                              \
      } else {                                             \
        U64Path p(ROW);                                    \
        CALL;                                              \
      }                                                    \
    }                                                      \
  } while (0)

template <int S1, int MODE>
__global__ __launch_bounds__(kBlock) void k_fwd_columns(NttArgs a) {
  __shared__ int64_t lds[(1 << S1) * (kBlock / ((1 << S1) / 16))];
  const RowRef rr = row_ref(a);
  FH_DISPATCH_PATH(MODE, *rr.row, (fwd_columns<S1>(p, rr.base, rr.src, rr.tw, rr.twp, a.form != 0, lds)));
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_fwd_tiles(NttArgs a) {
  __shared__ int64_t lds[kTilesPerBlock * kTilePitch];
  const TileItem it = tile_item(a);
  const RowRef rr = row_ref(a, it.flat);
  FH_DISPATCH_PATH(MODE, *rr.row, (fwd_tiles(p, a.log_n, it.g, rr.base, rr.tw, rr.twp, lds)));
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_inv_tiles(NttArgs a) {
  __shared__ int64_t lds[kTilesPerBlock * kTilePitch];
  const TileItem it = tile_item(a);
  const RowRef rr = row_ref(a, it.flat);
  FH_DISPATCH_PATH(MODE, *rr.row,
                   (inv_tiles(p, a.log_n, it.g, rr.base, rr.src, rr.tw, rr.twp, a.in_mid != 0, lds)));
}

template <int S1, int MODE>
__global__ __launch_bounds__(kBlock) void k_inv_columns(NttArgs a) {
  __shared__ int64_t lds[(1 << S1) * (kBlock / ((1 << S1) / 16))];
  const RowRef rr = row_ref(a);
  if (rr.limb < a.packed_rows) {
    // packed binary64 limb (fast arithmetic; input and output in place)
    const int64_t n = int64_t{1} << a.log_n;
    const PackedInter pk = packed_slab(rr.entry, a.packed_rows, n);
    const int64_t off = rr.limb * n;
    auto load = [&](int i, int c) -> int64_t {
      return __double_as_longlong(u64_to_f64(packed_load(pk, off + packed_index(i, c))));
    };
    auto store = [&](int i, int c, int64_t v) { packed_store(pk, off + packed_index(i, c), static_cast<uint64_t>(v)); };
    inv_columns<S1>(F64Fast(*rr.row), load, store, rr.tw, rr.twp, a.form, lds);
    return;
  }
  int64_t* base = rr.base;
  auto load = [&](int i, int c) -> int64_t { return base[static_cast<int64_t>(i) * kTile + c]; };
  auto store = [&](int i, int c, int64_t v) { base[static_cast<int64_t>(i) * kTile + c] = v; };
  FH_DISPATCH_PATH(MODE, *rr.row, (inv_columns<S1>(p, load, store, rr.tw, rr.twp, a.form, lds)));
}

// ---------------------------------------------------------------------------
// One-stage kernels for rings below 2^12, always in exact arithmetic (their
// results are the reference's lazy values, which also satisfy the fast
// mode's contract after the final canonicalisation below).
// Grid: (ceil(N/2 / 256), rows).
// ---------------------------------------------------------------------------
template <class P>
__device__ __forceinline__ void stage_body(const P& p, int64_t* base, const int64_t* tw,
                                           const int64_t* twp, int log_n, int stage, bool forward,
                                           int64_t b) {
  using V = typename P::V;
  const int64_t n = int64_t{1} << log_n;
  int64_t j, half, idx;
  if (forward) {
    half = n >> (stage + 1);
    const int64_t g = b / half;
    j = g * 2 * half + (b % half);
    idx = (int64_t{1} << stage) + g;
  } else {
    half = int64_t{1} << stage;
    const int64_t g = b >> stage;
    j = g * 2 * half + (b & (half - 1));
    idx = (n >> (stage + 1)) + g;
  }
  V u = p.in(base[j]);
  V l = p.in(base[j + half]);
  if (forward) {
    p.fwd(u, l, p.tw(tw, twp, idx));
  } else {
    p.inv(u, l, p.tw(tw, twp, idx));
  }
  base[j] = p.out_fwd(u);
  base[j + half] = p.out_fwd(l);
}

__global__ __launch_bounds__(kBlock) void k_stage(NttArgs a, int stage, int forward) {
  const int64_t n = int64_t{1} << a.log_n;
  const int64_t b = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (b >= n / 2) return;
  const RowRef rr = row_ref(a);
  if (rr.row->small) {
    stage_body(F64Exact(*rr.row), rr.base, rr.tw, rr.twp, a.log_n, stage, forward != 0, b);
  } else {
    stage_body(U64Path(*rr.row), rr.base, rr.tw, rr.twp, a.log_n, stage, forward != 0, b);
  }
}

// what: -1 forward to-Montgomery, -2 forward canonicalisation (fast mode
// output), -3 copy of the input row (out-of-place transforms),
// >= 0 inverse normalisation with that InvOut.
__global__ __launch_bounds__(kBlock) void k_scale(NttArgs a, int what) {
  const int64_t n = int64_t{1} << a.log_n;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const RowRef rr = row_ref(a);
  const NttRow& row = *rr.row;
  int64_t* x = rr.base + i;
  if (what == -3) {
    *x = rr.src[i];
  } else if (what == -2) {
    const uint64_t v = static_cast<uint64_t>(*x);
    *x = static_cast<int64_t>(v >= row.q ? v - row.q : v);
  } else if (row.small) {
    F64Exact p(row);
    *x = what == -1 ? p.out_fwd(p.to_mont(p.in(*x))) : p.out_inv(p.in(*x), what);
  } else {
    U64Path p(row);
    *x = what == -1 ? p.out_fwd(p.to_mont(p.in(*x))) : p.out_inv(p.in(*x), what);
  }
}

// Rows per launch of the two-pass schedule (0: as many as the grid allows).

template <int MODE>
void launch_columns(bool forward, int s1, dim3 grid, hipStream_t stream, const NttArgs& a) {
#define FH_COLUMNS_CASE(S)                                                           \
  case S:                                                                            \
    if (forward) {                                                                   \
      hipLaunchKernelGGL((k_fwd_columns<S, MODE>), grid, dim3(kBlock), 0, stream, a); \
    } else {                                                                         \
      hipLaunchKernelGGL((k_inv_columns<S, MODE>), grid, dim3(kBlock), 0, stream, a); \
    }                                                                                \
    break;
  switch (s1) {
    FH_COLUMNS_CASE(4)
    FH_COLUMNS_CASE(5)
    FH_COLUMNS_CASE(6)
    FH_COLUMNS_CASE(7)
    default:
      FH_COLUMNS_CASE(8)
  }
#undef FH_COLUMNS_CASE
}

template <int MODE>
int launch_two_pass(NttArgs a, int64_t total_rows, bool forward, hipStream_t stream) {
  const int L = a.log_n;
  const int64_t n = int64_t{1} << L;
  const int s1 = L - kTileLog;
  const int T = (1 << s1) / 16;
  const int C = kBlock / T;
  // the column pass's 2-D grid holds at most 65535 rows per launch; finer
  // schedules (chunks on several streams) are the caller's choice per call
  const int64_t chunk = total_rows < 65535 ? total_rows : 65535;
  for (int64_t start = 0; start < total_rows; start += chunk) {
    const int64_t rows = total_rows - start < chunk ? total_rows - start : chunk;
    a.row_offset = start;
    const dim3 gcol(kTile / C, static_cast<uint32_t>(rows));
    const int64_t groups = n / (kTile * kTilesPerBlock);
    // one launch over every entry: XCD-remapped 1-D tile grid (twiddles
    // reused from L2 across entries); chunked launches keep the 2-D grid
    const bool remap = start == 0 && rows == total_rows && total_rows % a.limbs == 0 &&
                       groups * total_rows <= 0x7fffffff;
    a.tile_batch = remap ? total_rows / a.limbs : 0;
    const dim3 gtile = remap ? dim3(static_cast<uint32_t>(groups * total_rows))
                             : dim3(static_cast<uint32_t>(groups), static_cast<uint32_t>(rows));
    if (forward) {
      launch_columns<MODE>(true, s1, gcol, stream, a);
      FH_CHECK_LAUNCH("ntt forward column pass");
      hipLaunchKernelGGL(k_fwd_tiles<MODE>, gtile, dim3(kBlock), 0, stream, a);
      FH_CHECK_LAUNCH("ntt forward tile pass");
    } else {
      hipLaunchKernelGGL(k_inv_tiles<MODE>, gtile, dim3(kBlock), 0, stream, a);
      FH_CHECK_LAUNCH("ntt inverse tile pass");
      launch_columns<MODE>(false, s1, gcol, stream, a);
      FH_CHECK_LAUNCH("ntt inverse column pass");
    }
  }
  return kOk;
}

constexpr int64_t kMaxGridY = 65535;

int launch(NttArgs a, int64_t total_rows, bool forward, int mode, hipStream_t stream) {
  const int L = a.log_n;
  const int64_t n = int64_t{1} << L;
  if (L >= 12 && L <= 16) {
    return mode == kFast ? launch_two_pass<kFast>(a, total_rows, forward, stream)
                         : launch_two_pass<kExact>(a, total_rows, forward, stream);
  }
  for (int64_t start = 0; start < total_rows; start += kMaxGridY) {
    const int64_t rows = total_rows - start < kMaxGridY ? total_rows - start : kMaxGridY;
    a.row_offset = start;
    const dim3 gel(static_cast<uint32_t>(ceil_div(n, kBlock)), static_cast<uint32_t>(rows));
    const dim3 gst(static_cast<uint32_t>(ceil_div(n / 2, kBlock)), static_cast<uint32_t>(rows));
    if (a.src != nullptr) {
      hipLaunchKernelGGL(k_scale, gel, dim3(kBlock), 0, stream, a, -3);
      FH_CHECK_LAUNCH("ntt input copy");
    }
    if (forward && a.form) {
      hipLaunchKernelGGL(k_scale, gel, dim3(kBlock), 0, stream, a, -1);
      FH_CHECK_LAUNCH("ntt to-montgomery scale");
    }
    for (int s = 0; s < L; ++s) {
      hipLaunchKernelGGL(k_stage, gst, dim3(kBlock), 0, stream, a, s, forward ? 1 : 0);
      FH_CHECK_LAUNCH("ntt stage");
    }
    if (!forward) {
      hipLaunchKernelGGL(k_scale, gel, dim3(kBlock), 0, stream, a, a.form);
      FH_CHECK_LAUNCH("ntt inverse scale");
    } else if (mode == kFast) {
      hipLaunchKernelGGL(k_scale, gel, dim3(kBlock), 0, stream, a, -2);
      FH_CHECK_LAUNCH("ntt canonicalise");
    }
  }
  return kOk;
}

// ---------------------------------------------------------------------------
// Indexed schedule (the reference CPU backend's family,
// indexed_ntt_cpu.cpp:96-170, 206-258): one launch per stage, one butterfly
// per thread, explicit even/odd index tables and expanded Montgomery
// twiddles [limb, stage, butterfly]; REDC exactly as the CPU does it.
// Grid: (ceil(butterflies / 256), limbs, batch).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_indexed_stage(
    int64_t* data, int64_t batch_stride, int64_t limb_stride, int64_t coef_stride,
    const int32_t* even, const int32_t* odd, int64_t idx_stride, const int64_t* tw,
    int64_t tw_limb_stride, int64_t tw_bfly_stride, ParamView params, int64_t butterflies,
    int forward) {
  const int64_t bf = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (bf >= butterflies) return;
  const int64_t l = blockIdx.y;
  const int64_t b = blockIdx.z;
  const Mont m = load_mont(params, l);
  int64_t* row = data + b * batch_stride + l * limb_stride;
  const int64_t e = even[bf * idx_stride];
  const int64_t o = odd[bf * idx_stride];
  const int64_t w = tw[l * tw_limb_stride + bf * tw_bfly_stride];
  const int64_t u = row[e * coef_stride];
  const int64_t v = row[o * coef_stride];
  if (forward) {
    const int64_t p = mont_mul(w, v, m);
    row[e * coef_stride] = add_lazy(u, p, m.twice_q);
    row[o * coef_stride] = sub_lazy(u, p, m.twice_q);
  } else {
    const int64_t d = sub_lazy(u, v, m.twice_q);
    row[e * coef_stride] = add_lazy(u, v, m.twice_q);
    row[o * coef_stride] = mont_mul(w, d, m);
  }
}

// ---------------------------------------------------------------------------
// Table preparation: derive the backend's per-row records and standard-form
// twiddles from the reference-format inputs (rns_params [8, limb] and compact
// Montgomery twiddles [limb, N], fhelium/engine/ntt/plans/twiddles.py:45-84).
// ---------------------------------------------------------------------------

// floor(w * 2^64 / q) for w < q by restoring long division.
__device__ uint64_t shoup_companion(uint64_t w, uint64_t q) {
  uint64_t rem = w;
  uint64_t quo = 0;
  for (int i = 0; i < 64; ++i) {
    const bool top = (rem >> 63) != 0;
    rem <<= 1;
    quo <<= 1;
    if (top || rem >= q) {
      rem -= q;
      quo |= 1;
    }
  }
  return quo;
}

// Rows of primes below 2^42 run in binary64.  The fast forward transform
// lets values grow to 11q over its 16 unreduced stages and every binary64
// product needs |w x| / q < 2^46.4 (w < q) to be exact, so 11q must stay
// below 2^46.4; the key switch's accumulators (below 9.1q) need the same.
// Larger primes take the integer path, which is exact for any q < 2^62.
constexpr uint64_t kSmallPrimeLimit = uint64_t{1} << 42;

__global__ void k_prepare_rows(ParamView params, int64_t limbs, NttRow* rows) {
  const int64_t l = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (l >= limbs) return;
  const Mont m = load_mont(params, l);
  NttRow r{};
  r.q = m.q;
  r.twice_q = 2 * m.q;
  r.qd = static_cast<double>(m.q);
  r.qinv = 1.0 / r.qd;
  r.small = m.q < kSmallPrimeLimit ? 1 : 0;
  // R mod q = REDC(R^2); N^{-1} = REDC(N^{-1} R); N^{-1} R^{-1} = REDC(N^{-1}).
  r.to_mont = static_cast<uint64_t>(mont_mul(m.r2, 1, m));
  r.ninv_mont = static_cast<uint64_t>(mont_mul(m.ninv, 1, m));
  r.ninv_std = static_cast<uint64_t>(mont_mul(static_cast<int64_t>(r.ninv_mont), 1, m));
  r.ninv_std_rinv = static_cast<uint64_t>(mont_mul(static_cast<int64_t>(r.ninv_std), 1, m));
  r.ninv_std_rinv_p = shoup_companion(r.ninv_std_rinv, m.q);
  r.to_mont_p = shoup_companion(r.to_mont, m.q);
  r.ninv_mont_p = shoup_companion(r.ninv_mont, m.q);
  r.ninv_std_p = shoup_companion(r.ninv_std, m.q);
  // k = -q^{-1} mod 2^62 lifts by one Newton step: q k (2 + q k) = -1 mod 2^124
  r.neg_qinv64 = m.k * (2 + m.q * m.k);
  rows[l] = r;
}

__global__ void k_prepare_twiddles(ParamView params, const int64_t* mont_tw, int64_t tw_row_stride,
                                   int64_t n, int64_t* out_w, int64_t* out_wp) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t l = blockIdx.y;
  if (i >= n) return;
  const Mont m = load_mont(params, l);
  const uint64_t w = static_cast<uint64_t>(mont_mul(mont_tw[l * tw_row_stride + i], 1, m));
  if (m.q < kSmallPrimeLimit) {
    out_w[l * n + i] = __double_as_longlong(static_cast<double>(w));
    out_wp[l * n + i] = 0;
  } else {
    out_w[l * n + i] = static_cast<int64_t>(w);
    out_wp[l * n + i] = static_cast<int64_t>(shoup_companion(w, m.q));
  }
}

int check_ntt_common(int64_t batch, int64_t limbs, int64_t log_n, const void* rows,
                     const int64_t* tw, const int64_t* twp, int64_t row_start, int32_t mode) {
  FH_CHECK_ARG(log_n >= 1 && log_n <= 16, "NTT ring log2 %lld outside [1, 16]", (long long)log_n);
  FH_CHECK_ARG(batch >= 0 && limbs >= 0, "negative NTT batch or limb count");
  FH_CHECK_ARG(rows && tw && twp, "NTT tables must be non-null");
  FH_CHECK_ARG(row_start >= 0, "negative NTT parameter row start");
  FH_CHECK_ARG(mode == kExact || mode == kFast, "NTT mode %d is not 0 (exact) or 1 (fast)", mode);
  return kOk;
}

}  // namespace

namespace ntt {

int inverse_internal(int64_t* data, int64_t batch, int64_t limbs, int log_n, int64_t batch_stride,
                     int64_t limb_stride, const NttRow* rows, const int64_t* tw, const int64_t* twp,
                     int64_t row_start, int output_form, bool in_mid, hipStream_t stream) {
  NttArgs a{data, batch_stride, limb_stride, limbs, 0, rows, tw, twp, row_start, log_n, output_form,
            in_mid ? 1 : 0};
  return launch(a, batch * limbs, false, kFast, stream);
}

int forward_columns_internal(int64_t* dst, const int64_t* src, int64_t batch, int64_t limbs, int log_n,
                             int64_t dst_batch_stride, int64_t dst_limb_stride, int64_t src_batch_stride,
                             int64_t src_limb_stride, const NttRow* rows, const int64_t* tw,
                             const int64_t* twp, int64_t row_start, bool to_montgomery, hipStream_t stream) {
  FH_CHECK_ARG(log_n >= 12 && log_n <= 16, "column-only forward needs logN 12..16");
  NttArgs a{dst, dst_batch_stride, dst_limb_stride, limbs, 0, rows, tw, twp, row_start, log_n,
            to_montgomery ? 1 : 0, 0, src, src_batch_stride, src_limb_stride};
  const int s1 = log_n - kTileLog;
  const int T = (1 << s1) / 16;
  const int C = kBlock / T;
  const int64_t total_rows = batch * limbs;
  for (int64_t start = 0; start < total_rows; start += kMaxGridY) {
    const int64_t rows_now = total_rows - start < kMaxGridY ? total_rows - start : kMaxGridY;
    a.row_offset = start;
    launch_columns<kFast>(true, s1, dim3(kTile / C, static_cast<uint32_t>(rows_now)), stream, a);
    FH_CHECK_LAUNCH("ntt forward column pass");
  }
  return kOk;
}

int inverse_columns_internal(int64_t* data, int64_t batch, int64_t limbs, int log_n,
                             int64_t batch_stride, int64_t limb_stride, const NttRow* rows,
                             const int64_t* tw, const int64_t* twp, int64_t row_start,
                             int output_form, int64_t packed_rows, hipStream_t stream) {
  FH_CHECK_ARG(log_n >= 12 && log_n <= 16, "column-only inverse needs logN 12..16");
  FH_CHECK_ARG(packed_rows >= 0 && packed_rows <= limbs && (packed_rows == 0 || limb_stride == (int64_t{1} << log_n)),
               "packed rows need contiguous limbs");
  NttArgs a{data, batch_stride, limb_stride, limbs, 0, rows, tw, twp, row_start, log_n, output_form, 1};
  a.packed_rows = packed_rows;
  const int s1 = log_n - kTileLog;
  const int T = (1 << s1) / 16;
  const int C = kBlock / T;
  const int64_t total_rows = batch * limbs;
  for (int64_t start = 0; start < total_rows; start += kMaxGridY) {
    const int64_t rows_now = total_rows - start < kMaxGridY ? total_rows - start : kMaxGridY;
    a.row_offset = start;
    launch_columns<kFast>(false, s1, dim3(kTile / C, static_cast<uint32_t>(rows_now)), stream, a);
    FH_CHECK_LAUNCH("ntt inverse column pass");
  }
  return kOk;
}

}  // namespace ntt
}  // namespace fh

using namespace fh;

FH_API int fh_ntt_forward(int64_t* residues, int64_t batch, int64_t limbs, int64_t log_n,
                          int64_t batch_stride, int64_t limb_stride, const void* ntt_rows,
                          const int64_t* twiddles, const int64_t* twiddles_shoup,
                          int64_t row_start, int32_t to_montgomery, int32_t mode, void* stream) {
  int st = check_ntt_common(batch, limbs, log_n, ntt_rows, twiddles, twiddles_shoup, row_start, mode);
  if (st != kOk) return st;
  if (batch == 0 || limbs == 0) return kOk;
  FH_CHECK_ARG(residues != nullptr, "NTT residues must be non-null");
  NttArgs a{residues, batch_stride, limb_stride, limbs, 0, static_cast<const NttRow*>(ntt_rows),
            twiddles, twiddles_shoup, row_start, static_cast<int>(log_n), to_montgomery ? 1 : 0, 0};
  return launch(a, batch * limbs, true, mode, static_cast<hipStream_t>(stream));
}

FH_API int fh_ntt_inverse(int64_t* residues, int64_t batch, int64_t limbs, int64_t log_n,
                          int64_t batch_stride, int64_t limb_stride, const void* ntt_rows,
                          const int64_t* twiddles, const int64_t* twiddles_shoup,
                          int64_t row_start, int32_t output_form, int32_t mode, void* stream) {
  int st = check_ntt_common(batch, limbs, log_n, ntt_rows, twiddles, twiddles_shoup, row_start, mode);
  if (st != kOk) return st;
  FH_CHECK_ARG(output_form >= 0 && output_form <= 3, "inverse NTT output form %d is not 0..3",
               output_form);
  if (batch == 0 || limbs == 0) return kOk;
  FH_CHECK_ARG(residues != nullptr, "NTT residues must be non-null");
  NttArgs a{residues, batch_stride, limb_stride, limbs, 0, static_cast<const NttRow*>(ntt_rows),
            twiddles, twiddles_shoup, row_start, static_cast<int>(log_n), output_form, 0};
  return launch(a, batch * limbs, false, mode, static_cast<hipStream_t>(stream));
}

// Out-of-place transforms: read `source` (same shape, its own strides),
// write `residues`; the source is not modified.
FH_API int fh_ntt_forward_from(int64_t* residues, const int64_t* source, int64_t batch, int64_t limbs,
                               int64_t log_n, int64_t batch_stride, int64_t limb_stride,
                               int64_t source_batch_stride, int64_t source_limb_stride,
                               const void* ntt_rows, const int64_t* twiddles,
                               const int64_t* twiddles_shoup, int64_t row_start, int32_t to_montgomery,
                               int32_t mode, void* stream) {
  int st = check_ntt_common(batch, limbs, log_n, ntt_rows, twiddles, twiddles_shoup, row_start, mode);
  if (st != kOk) return st;
  if (batch == 0 || limbs == 0) return kOk;
  FH_CHECK_ARG(residues != nullptr && source != nullptr, "NTT residues must be non-null");
  NttArgs a{residues, batch_stride, limb_stride, limbs, 0, static_cast<const NttRow*>(ntt_rows),
            twiddles, twiddles_shoup, row_start, static_cast<int>(log_n), to_montgomery ? 1 : 0, 0,
            source, source_batch_stride, source_limb_stride};
  return launch(a, batch * limbs, true, mode, static_cast<hipStream_t>(stream));
}

FH_API int fh_ntt_inverse_from(int64_t* residues, const int64_t* source, int64_t batch, int64_t limbs,
                               int64_t log_n, int64_t batch_stride, int64_t limb_stride,
                               int64_t source_batch_stride, int64_t source_limb_stride,
                               const void* ntt_rows, const int64_t* twiddles,
                               const int64_t* twiddles_shoup, int64_t row_start, int32_t output_form,
                               int32_t mode, void* stream) {
  int st = check_ntt_common(batch, limbs, log_n, ntt_rows, twiddles, twiddles_shoup, row_start, mode);
  if (st != kOk) return st;
  FH_CHECK_ARG(output_form >= 0 && output_form <= 3, "inverse NTT output form %d is not 0..3",
               output_form);
  if (batch == 0 || limbs == 0) return kOk;
  FH_CHECK_ARG(residues != nullptr && source != nullptr, "NTT residues must be non-null");
  NttArgs a{residues, batch_stride, limb_stride, limbs, 0, static_cast<const NttRow*>(ntt_rows),
            twiddles, twiddles_shoup, row_start, static_cast<int>(log_n), output_form, 0,
            source, source_batch_stride, source_limb_stride};
  return launch(a, batch * limbs, false, mode, static_cast<hipStream_t>(stream));
}


FH_API int fh_ntt_prepare_tables(const int64_t* rns_params, int64_t param_row_stride,
                                 int64_t param_limb_stride, int64_t limbs, int64_t log_n,
                                 const int64_t* montgomery_twiddles, int64_t twiddle_row_stride,
                                 void* ntt_rows_out, int64_t* twiddles_out,
                                 int64_t* twiddles_shoup_out, void* stream) {
  FH_CHECK_ARG(rns_params && ntt_rows_out, "prepare: null parameter table");
  FH_CHECK_ARG(log_n >= 1 && log_n <= 16, "prepare: ring log2 outside [1, 16]");
  FH_CHECK_ARG(limbs > 0 && limbs <= 65535, "prepare: bad limb count");
  hipStream_t s = static_cast<hipStream_t>(stream);
  ParamView pv{rns_params, param_row_stride, param_limb_stride};
  hipLaunchKernelGGL(k_prepare_rows, dim3(static_cast<uint32_t>(ceil_div(limbs, 64))), dim3(64), 0,
                     s, pv, limbs, static_cast<NttRow*>(ntt_rows_out));
  FH_CHECK_LAUNCH("ntt prepare rows");
  if (montgomery_twiddles != nullptr) {
    FH_CHECK_ARG(twiddles_out && twiddles_shoup_out, "prepare: null twiddle output");
    const int64_t n = int64_t{1} << log_n;
    hipLaunchKernelGGL(k_prepare_twiddles,
                       dim3(static_cast<uint32_t>(ceil_div(n, 256)), static_cast<uint32_t>(limbs)),
                       dim3(256), 0, s, pv, montgomery_twiddles, twiddle_row_stride, n, twiddles_out,
                       twiddles_shoup_out);
    FH_CHECK_LAUNCH("ntt prepare twiddles");
  }
  return kOk;
}

// One stage of an indexed radix-2 schedule.  even/odd point at that stage's
// [butterflies] index rows; twiddles at that stage's [limb, butterflies]
// Montgomery slice.  The caller validates that the schedule partitions the
// coefficients (indexed_ntt_cpu.cpp:315-354) before the first stage.
FH_API int fh_ntt_indexed_stage(fh_view3 residues, int64_t batch, int64_t limbs,
                                const int32_t* even_indices, const int32_t* odd_indices,
                                int64_t index_stride, const int64_t* stage_twiddles,
                                int64_t twiddle_limb_stride, int64_t twiddle_butterfly_stride,
                                fh_table2 rns_params, int64_t butterflies, int32_t forward,
                                void* stream) {
  FH_CHECK_ARG(batch >= 0 && limbs >= 0 && butterflies >= 0, "negative extent");
  FH_CHECK_ARG(batch <= 65535 && limbs <= 65535, "batch or limb extent exceeds the launch grid");
  if (batch == 0 || limbs == 0 || butterflies == 0) return kOk;
  FH_CHECK_ARG(residues.data && even_indices && odd_indices && stage_twiddles && rns_params.data,
               "indexed stage: null operand");
  ParamView pv{rns_params.data, rns_params.stride_row, rns_params.stride_col};
  hipLaunchKernelGGL(k_indexed_stage,
                     dim3(static_cast<uint32_t>(ceil_div(butterflies, kBlock)),
                          static_cast<uint32_t>(limbs), static_cast<uint32_t>(batch)),
                     dim3(kBlock), 0, static_cast<hipStream_t>(stream), residues.data,
                     residues.stride_batch, residues.stride_limb, residues.stride_coef, even_indices,
                     odd_indices, index_stride, stage_twiddles, twiddle_limb_stride,
                     twiddle_butterfly_stride, pv, butterflies, forward);
  FH_CHECK_LAUNCH("ntt indexed stage");
  return kOk;
}
