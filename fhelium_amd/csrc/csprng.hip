// ChaCha20 keystream and the samplers built on it, for gfx950.
//
// The reference draws every key and ciphertext sample from ChaCha20 streams
// (fhelium/rng/csprng.py:11-27, 150-250, through the external triton-csprng
// package).  This file is this backend's own generator: the ChaCha20 block
// function of RFC 8439 section 2.3 (20 rounds of 32-bit add / xor / rotate on
// a 4x4 state of constants, a 256-bit key, a 32-bit block counter and a
// 96-bit nonce; output = rounds(state) + state, little-endian words), one
// block per thread, and three samplers that consume whole blocks:
//
//   uniform mod q   4 words (128 bits) per coefficient, reduced mod q
//                   (bias below 2^-64 for q < 2^64);
//   ternary         2 words (64 bits) per coefficient, (x mod 3) - 1;
//   gaussian        4 words per pair of coefficients: two 53-bit uniforms,
//                   Box-Muller, each normal scaled by sigma and rounded to
//                   the nearest integer.
//
// Coefficient g of a sampler takes words [w*g, w*(g+1)) of the keystream
// that starts at block `counter`, so any split of a request into launches
// gives the same values.
#include "fh_common.h"

namespace fh {
namespace {

constexpr int kCsBlock = 256;

struct ChaChaKey {
  uint32_t k[8];
  uint32_t nonce[3];
};

__device__ __forceinline__ uint32_t rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

__device__ __forceinline__ void quarter(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d) {
  a += b; d ^= a; d = rotl(d, 16);
  c += d; b ^= c; b = rotl(b, 12);
  a += b; d ^= a; d = rotl(d, 8);
  c += d; b ^= c; b = rotl(b, 7);
}

// RFC 8439 section 2.3: state = "expand 32-byte k" constants, key, counter,
// nonce; ten double rounds (column then diagonal quarter rounds).
__device__ __forceinline__ void chacha20_block(const ChaChaKey& key, uint32_t counter, uint32_t out[16]) {
  uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                    key.k[0], key.k[1], key.k[2], key.k[3], key.k[4], key.k[5], key.k[6], key.k[7],
                    counter, key.nonce[0], key.nonce[1], key.nonce[2]};
  uint32_t x[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) x[i] = s[i];
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    quarter(x[0], x[4], x[8], x[12]);
    quarter(x[1], x[5], x[9], x[13]);
    quarter(x[2], x[6], x[10], x[14]);
    quarter(x[3], x[7], x[11], x[15]);
    quarter(x[0], x[5], x[10], x[15]);
    quarter(x[1], x[6], x[11], x[12]);
    quarter(x[2], x[7], x[8], x[13]);
    quarter(x[3], x[4], x[9], x[14]);
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) out[i] = x[i] + s[i];
}

__global__ __launch_bounds__(kCsBlock) void k_keystream(ChaChaKey key, uint32_t counter, int64_t blocks,
                                                       uint32_t* out) {
  const int64_t b = static_cast<int64_t>(blockIdx.x) * kCsBlock + threadIdx.x;
  if (b >= blocks) return;
  uint32_t w[16];
  chacha20_block(key, counter + static_cast<uint32_t>(b), w);
  uint4* dst = reinterpret_cast<uint4*>(out + b * 16);
#pragma unroll
  for (int i = 0; i < 4; ++i) dst[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
}

// 128-bit value (w0 + 2^32 w1 + 2^64 w2 + 2^96 w3) mod q.
__device__ __forceinline__ uint64_t mod_u128(const uint32_t* w, uint64_t q) {
  const unsigned __int128 x = (static_cast<unsigned __int128>(static_cast<uint64_t>(w[3]) << 32 | w[2]) << 64) |
                              (static_cast<uint64_t>(w[1]) << 32 | w[0]);
  return static_cast<uint64_t>(x % q);
}

// out[e, r, i] for e < outer, r < rows, i < count: coefficient index
// g = (e * rows + r) * count + i, four per block.
__global__ __launch_bounds__(kCsBlock) void k_uniform_mod(ChaChaKey key, uint32_t counter, const int64_t* moduli,
                                                         int64_t rows, int64_t count, int64_t total,
                                                         int64_t* out) {
  const int64_t b = static_cast<int64_t>(blockIdx.x) * kCsBlock + threadIdx.x;
  if (b * 4 >= total) return;
  uint32_t w[16];
  chacha20_block(key, counter + static_cast<uint32_t>(b), w);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int64_t g = b * 4 + c;
    if (g >= total) break;
    const int64_t r = (g / count) % rows;
    out[g] = static_cast<int64_t>(mod_u128(w + 4 * c, static_cast<uint64_t>(moduli[r])));
  }
}

__global__ __launch_bounds__(kCsBlock) void k_ternary(ChaChaKey key, uint32_t counter, int64_t total,
                                                     int64_t* out) {
  const int64_t b = static_cast<int64_t>(blockIdx.x) * kCsBlock + threadIdx.x;
  if (b * 8 >= total) return;
  uint32_t w[16];
  chacha20_block(key, counter + static_cast<uint32_t>(b), w);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int64_t g = b * 8 + c;
    if (g >= total) break;
    const uint64_t x = static_cast<uint64_t>(w[2 * c + 1]) << 32 | w[2 * c];
    out[g] = static_cast<int64_t>(x % 3) - 1;
  }
}

// (0, 1) uniform from 64 bits: the top 53 bits plus one half ulp.
__device__ __forceinline__ double unit_open(uint32_t lo, uint32_t hi) {
  const uint64_t x = (static_cast<uint64_t>(hi) << 32 | lo) >> 11;
  return (static_cast<double>(x) + 0.5) * 0x1.0p-53;
}

__global__ __launch_bounds__(kCsBlock) void k_gaussian(ChaChaKey key, uint32_t counter, int64_t total, double sigma,
                                                      int64_t* out) {
  const int64_t b = static_cast<int64_t>(blockIdx.x) * kCsBlock + threadIdx.x;
  if (b * 8 >= total) return;
  uint32_t w[16];
  chacha20_block(key, counter + static_cast<uint32_t>(b), w);
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const double u1 = unit_open(w[4 * p], w[4 * p + 1]);
    const double u2 = unit_open(w[4 * p + 2], w[4 * p + 3]);
    const double rad = sigma * sqrt(-2.0 * log(u1));
    double sn, cs;
    sincospi(2.0 * u2, &sn, &cs);
    const int64_t g = b * 8 + 2 * p;
    if (g < total) out[g] = static_cast<int64_t>(rint(rad * cs));
    if (g + 1 < total) out[g + 1] = static_cast<int64_t>(rint(rad * sn));
  }
}

int make_key(const uint32_t* key, const uint32_t* nonce, ChaChaKey* out) {
  FH_CHECK_ARG(key != nullptr && nonce != nullptr, "ChaCha20 key and nonce are required");
  for (int i = 0; i < 8; ++i) out->k[i] = key[i];
  for (int i = 0; i < 3; ++i) out->nonce[i] = nonce[i];
  return kOk;
}

int check_blocks(uint32_t counter, int64_t blocks) {
  FH_CHECK_ARG(blocks >= 0, "negative block count");
  FH_CHECK_ARG(static_cast<uint64_t>(counter) + static_cast<uint64_t>(blocks) <= (uint64_t{1} << 32),
               "ChaCha20 block counter would wrap (counter %u + %lld blocks)", counter, (long long)blocks);
  FH_CHECK_ARG(ceil_div(blocks, kCsBlock) <= 0x7fffffff, "keystream request too large");
  return kOk;
}

}  // namespace
}  // namespace fh

using namespace fh;

FH_API int fh_chacha20_keystream(const uint32_t* key, const uint32_t* nonce, uint32_t counter, int64_t blocks,
                                 uint32_t* out, void* stream) {
  ChaChaKey k;
  int st = make_key(key, nonce, &k);
  if (st == kOk) st = check_blocks(counter, blocks);
  if (st != kOk) return st;
  FH_CHECK_ARG(out != nullptr || blocks == 0, "null output");
  if (blocks == 0) return kOk;
  hipLaunchKernelGGL(k_keystream, dim3(static_cast<uint32_t>(ceil_div(blocks, kCsBlock))), dim3(kCsBlock), 0,
                     static_cast<hipStream_t>(stream), k, counter, blocks, out);
  FH_CHECK_LAUNCH("chacha20 keystream");
  return kOk;
}

FH_API int fh_sample_uniform_mod(const uint32_t* key, const uint32_t* nonce, uint32_t counter,
                                 const int64_t* moduli, int64_t outer, int64_t rows, int64_t count, int64_t* out,
                                 void* stream) {
  ChaChaKey k;
  int st = make_key(key, nonce, &k);
  if (st != kOk) return st;
  FH_CHECK_ARG(outer >= 0 && rows >= 1 && count >= 1 && moduli != nullptr, "bad uniform sample shape");
  const int64_t total = outer * rows * count;
  st = check_blocks(counter, ceil_div(total, 4));
  if (st != kOk) return st;
  if (total == 0) return kOk;
  FH_CHECK_ARG(out != nullptr, "null output");
  hipLaunchKernelGGL(k_uniform_mod, dim3(static_cast<uint32_t>(ceil_div(ceil_div(total, 4), kCsBlock))),
                     dim3(kCsBlock), 0, static_cast<hipStream_t>(stream), k, counter, moduli, rows, count, total,
                     out);
  FH_CHECK_LAUNCH("uniform mod q sampler");
  return kOk;
}

FH_API int fh_sample_ternary(const uint32_t* key, const uint32_t* nonce, uint32_t counter, int64_t total,
                             int64_t* out, void* stream) {
  ChaChaKey k;
  int st = make_key(key, nonce, &k);
  if (st == kOk) st = check_blocks(counter, ceil_div(total, 8));
  if (st != kOk) return st;
  FH_CHECK_ARG(total >= 0, "negative count");
  if (total == 0) return kOk;
  FH_CHECK_ARG(out != nullptr, "null output");
  hipLaunchKernelGGL(k_ternary, dim3(static_cast<uint32_t>(ceil_div(ceil_div(total, 8), kCsBlock))), dim3(kCsBlock),
                     0, static_cast<hipStream_t>(stream), k, counter, total, out);
  FH_CHECK_LAUNCH("ternary sampler");
  return kOk;
}

FH_API int fh_sample_gaussian(const uint32_t* key, const uint32_t* nonce, uint32_t counter, int64_t total,
                              double sigma, int64_t* out, void* stream) {
  ChaChaKey k;
  int st = make_key(key, nonce, &k);
  if (st == kOk) st = check_blocks(counter, ceil_div(total, 8));
  if (st != kOk) return st;
  FH_CHECK_ARG(total >= 0, "negative count");
  FH_CHECK_ARG(sigma > 0.0 && sigma < 1e6, "sigma must be positive and moderate");
  if (total == 0) return kOk;
  FH_CHECK_ARG(out != nullptr, "null output");
  hipLaunchKernelGGL(k_gaussian, dim3(static_cast<uint32_t>(ceil_div(ceil_div(total, 8), kCsBlock))), dim3(kCsBlock),
                     0, static_cast<hipStream_t>(stream), k, counter, total, sigma, out);
  FH_CHECK_LAUNCH("gaussian sampler");
  return kOk;
}
