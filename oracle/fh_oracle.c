/*
 * CPU oracle for the CKKS hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference CPU kernels' integer semantics,
 * used by tests/ to check the gfx950 library and by bench.py's cpu_baseline
 * leg.  The product never links, imports or calls this file.
 *
 * Every function follows one reference routine (paths relative to the
 * reference repository):
 *   or_mont_mul          csrc/ops/common/cpu/montgomery.h:125-143
 *   or_ntt_forward       csrc/ops/ntt/cpu/indexed_ntt_cpu.cpp:34-171 with the
 *                        schedule of fhelium/engine/ntt/plans/indexed_radix2.py:71-100
 *   or_ntt_inverse       indexed_ntt_cpu.cpp:173-313, schedule indexed_radix2.py:101-126
 *   or_binary            csrc/ops/rns/cpu/rns_arithmetic_cpu.cpp:61-127, 193-262;
 *                        rns_canonical_arithmetic_cpu.cpp:19-92
 *   or_unary             rns_arithmetic_cpu.cpp:148-192
 *   or_row_scalars       rns_arithmetic_cpu.cpp:319-369; rns_canonical_arithmetic_cpu.cpp:170-240
 *   or_add_lazy_twice    rns_arithmetic_cpu.cpp:371-432
 *   or_lift              rns_arithmetic_cpu.cpp:434-480
 *   or_mixed_radix_decompose  rns_arithmetic_cpu.cpp:482-614
 *   or_basis_extend      rns_arithmetic_cpu.cpp:616-700
 *   or_rescale           csrc/ops/ckks/cpu/ckks_cpu.cpp:204-275
 *   or_galois_coef       ckks_cpu.cpp:323-405
 *   or_galois_ntt        ckks_cpu.cpp:407-472
 *   or_moddown           ckks_cpu.cpp:474-601
 *   or_keyswitch_mac     ckks_cpu.cpp:603-708
 *   or_add_plaintext     ckks_cpu.cpp:54-146
 *
 * Layout: all arrays are dense row-major int64.  A residue operand is
 * [batch, limb, n]; rns_params is [8, limbs] with the reference's row
 * meaning (csrc/ops/common/rns_parameters.h:6-16).
 */
#include <stdint.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

#define RADIX_BITS 62
#define HALF_BITS 31
#define RADIX_MASK ((((uint64_t)1) << RADIX_BITS) - 1)

typedef struct {
  int64_t twice_q;
  uint64_t q;
  uint64_t k;
  int64_t r2;
  int64_t ninv;
} consts_t;

static consts_t load(const int64_t* params, int64_t limbs, int64_t limb) {
  consts_t c;
  c.twice_q = params[0 * limbs + limb];
  c.q = (uint64_t)params[1 * limbs + limb] + ((uint64_t)params[2 * limbs + limb] << HALF_BITS);
  c.k = (uint64_t)params[3 * limbs + limb] + ((uint64_t)params[4 * limbs + limb] << HALF_BITS);
  c.r2 = params[5 * limbs + limb];
  c.ninv = params[7 * limbs + limb];
  return c;
}

int64_t or_mont_mul(int64_t lhs, int64_t rhs, uint64_t q, uint64_t k) {
  const u128 product = (u128)(uint64_t)lhs * (uint64_t)rhs;
  const uint64_t m = (uint64_t)(((u128)(uint64_t)product * k) & RADIX_MASK);
  const u128 reduced = (product + (u128)m * q) >> RADIX_BITS;
  const uint64_t r = (uint64_t)reduced;
  return (int64_t)(r < q ? r : r - q);
}

static inline int64_t mul(int64_t a, int64_t b, const consts_t* c) { return or_mont_mul(a, b, c->q, c->k); }

static inline int64_t canon_op(int64_t v, uint64_t q) {
  int64_t r = v % (int64_t)q;
  if (r < 0) r += (int64_t)q;
  return r;
}

static inline int64_t mul_split(int64_t a, int64_t b, const consts_t* c) {
  return mul(canon_op(a, c->q), canon_op(b, c->q), c);
}

static inline int64_t add_lazy(int64_t a, int64_t b, int64_t t) {
  const int64_t s = a + b;
  return s < t ? s : s - t;
}

static inline int64_t sub_lazy(int64_t a, int64_t b, int64_t t) {
  const int64_t d = a + t - b;
  return d < t ? d : d - t;
}

static inline int64_t canonicalize(int64_t v, int64_t t) {
  const int64_t q = t >> 1;
  return v < q ? v : v - q;
}

static inline int64_t center(int64_t v, int64_t t) {
  const int64_t q = t >> 1;
  return v <= (q >> 1) ? v : v - q;
}

static void set_threads(int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
}

/* Forward negacyclic NTT.  twiddles: compact Montgomery table [limbs, n]
 * (psi^{br(k)} R mod q); stage s uses entry 2^s + group. */
void or_ntt_forward(int64_t* data, int64_t batch, int64_t limbs, int64_t log_n,
                    const int64_t* twiddles, const int64_t* params, int to_montgomery,
                    int threads) {
  const int64_t n = (int64_t)1 << log_n;
  set_threads(threads);
#pragma omp parallel for schedule(static)
  for (int64_t row = 0; row < batch * limbs; ++row) {
    const int64_t limb = row % limbs;
    const consts_t c = load(params, limbs, limb);
    int64_t* v = data + row * n;
    const int64_t* tw = twiddles + limb * n;
    if (to_montgomery) {
      for (int64_t i = 0; i < n; ++i) v[i] = mul(v[i], c.r2, &c);
    }
    int64_t half = n;
    for (int64_t s = 0; s < log_n; ++s) {
      const int64_t groups = (int64_t)1 << s;
      half >>= 1;
      for (int64_t g = 0; g < groups; ++g) {
        const int64_t w = tw[groups + g];
        const int64_t first = 2 * g * half;
        for (int64_t j = first; j < first + half; ++j) {
          const int64_t u = v[j];
          const int64_t p = mul(w, v[j + half], &c);
          v[j] = add_lazy(u, p, c.twice_q);
          v[j + half] = sub_lazy(u, p, c.twice_q);
        }
      }
    }
  }
}

/* Inverse NTT; out_form 0 montgomery, 1 standard lazy, 2 standard, 3 centered. */
void or_ntt_inverse(int64_t* data, int64_t batch, int64_t limbs, int64_t log_n,
                    const int64_t* twiddles, const int64_t* params, int out_form,
                    int threads) {
  const int64_t n = (int64_t)1 << log_n;
  set_threads(threads);
#pragma omp parallel for schedule(static)
  for (int64_t row = 0; row < batch * limbs; ++row) {
    const int64_t limb = row % limbs;
    const consts_t c = load(params, limbs, limb);
    int64_t* v = data + row * n;
    const int64_t* tw = twiddles + limb * n;
    int64_t half = 1;
    for (int64_t remaining = log_n; remaining > 0; --remaining) {
      const int64_t groups = (int64_t)1 << (remaining - 1);
      int64_t first = 0;
      for (int64_t g = 0; g < groups; ++g) {
        const int64_t w = tw[groups + g];
        for (int64_t j = first; j < first + half; ++j) {
          const int64_t u = v[j];
          const int64_t l = v[j + half];
          const int64_t d = sub_lazy(u, l, c.twice_q);
          v[j] = add_lazy(u, l, c.twice_q);
          v[j + half] = mul(w, d, &c);
        }
        first += 2 * half;
      }
      half *= 2;
    }
    for (int64_t i = 0; i < n; ++i) {
      int64_t x = mul(v[i], c.ninv, &c);
      if (out_form != 0) {
        x = mul(x, 1, &c);
        if (out_form == 2) x = canonicalize(x, c.twice_q);
        if (out_form == 3) x = center(canonicalize(x, c.twice_q), c.twice_q);
      }
      v[i] = x;
    }
  }
}

/* op: 0 mul, 1 add_lazy, 2 sub_lazy, 3 add_canonical, 4 sub_canonical,
 * 5 mul cyclic compressed, 6 mul contiguous compressed.  rhs has rhs_batch
 * entries (1 broadcasts) and, for compressed ops, `support` coefficients. */
void or_binary(int op, int64_t* out, const int64_t* a, const int64_t* b, int64_t batch,
               int64_t rhs_batch, int64_t limbs, int64_t n, const int64_t* params,
               int64_t support, int threads) {
  set_threads(threads);
  const int64_t rn = (op == 5 || op == 6) ? support : n;
#pragma omp parallel for schedule(static)
  for (int64_t row = 0; row < batch * limbs; ++row) {
    const int64_t bb = row / limbs;
    const int64_t limb = row % limbs;
    const consts_t c = load(params, limbs, limb);
    const int64_t* x = a + row * n;
    const int64_t* y = b + ((rhs_batch == batch ? bb : 0) * limbs + limb) * rn;
    int64_t* o = out + row * n;
    for (int64_t i = 0; i < n; ++i) {
      int64_t r;
      switch (op) {
        case 0: r = mul(x[i], y[i], &c); break;
        case 1: r = add_lazy(x[i], y[i], c.twice_q); break;
        case 2: r = sub_lazy(x[i], y[i], c.twice_q); break;
        case 3: {
          int64_t v = x[i] + y[i];
          if (v >= c.twice_q) v -= c.twice_q;
          r = canonicalize(v, c.twice_q);
          break;
        }
        case 4: {
          int64_t v = x[i] - y[i];
          if (v < 0) v += c.twice_q;
          else if (v >= c.twice_q) v -= c.twice_q;
          r = canonicalize(v, c.twice_q);
          break;
        }
        case 5: r = mul(x[i], y[i % support], &c); break;
        default: r = mul(x[i], y[i / (n / support)], &c); break;
      }
      o[i] = r;
    }
  }
}

/* op: 0 to_montgomery, 1 from_montgomery, 2 canonicalize, 3 center, 4 shift. */
void or_unary(int op, int64_t* x, int64_t batch, int64_t limbs, int64_t n,
              const int64_t* params) {
  for (int64_t row = 0; row < batch * limbs; ++row) {
    const consts_t c = load(params, limbs, row % limbs);
    int64_t* v = x + row * n;
    for (int64_t i = 0; i < n; ++i) {
      switch (op) {
        case 0: v[i] = mul(v[i], c.r2, &c); break;
        case 1: v[i] = mul(v[i], 1, &c); break;
        case 2: v[i] = canonicalize(v[i], c.twice_q); break;
        case 3: v[i] = center(v[i], c.twice_q); break;
        default: v[i] = v[i] + (c.twice_q >> 1); break;
      }
    }
  }
}

void or_row_scalars(int64_t* out, const int64_t* in, const int64_t* scalars,
                    const int64_t* params, int64_t batch, int64_t limbs, int64_t n,
                    int canonical) {
  for (int64_t row = 0; row < batch * limbs; ++row) {
    const int64_t limb = row % limbs;
    const consts_t c = load(params, limbs, limb);
    for (int64_t i = 0; i < n; ++i) {
      int64_t v = mul(in[row * n + i], scalars[limb], &c);
      if (canonical) v = canonicalize(v, c.twice_q);
      out[row * n + i] = v;
    }
  }
}

void or_add_lazy_twice(int64_t* out, const int64_t* a, const int64_t* b,
                       const int64_t* twice, int64_t batch, int64_t rhs_batch,
                       int64_t limbs, int64_t n) {
  for (int64_t row = 0; row < batch * limbs; ++row) {
    const int64_t bb = row / limbs, limb = row % limbs;
    const int64_t* y = b + ((rhs_batch == batch ? bb : 0) * limbs + limb) * n;
    for (int64_t i = 0; i < n; ++i) out[row * n + i] = add_lazy(a[row * n + i], y[i], twice[limb]);
  }
}

/* in: [batch, n] signed coefficients -> out [batch, limbs, n]. */
void or_lift(int64_t* out, const int64_t* in, const int64_t* twice, int64_t batch,
             int64_t limbs, int64_t n) {
  for (int64_t bb = 0; bb < batch; ++bb)
    for (int64_t l = 0; l < limbs; ++l)
      for (int64_t i = 0; i < n; ++i)
        out[(bb * limbs + l) * n + i] = in[bb * n + i] + (twice[l] >> 1);
}

/* src/out: [batch, rows, n]; norm [rows-1]; prop [rows-1, rows]; split
 * modulus words [rows]. */
void or_mixed_radix_decompose(int64_t* out, const int64_t* src, int64_t batch,
                              int64_t rows, int64_t n, const int64_t* norm,
                              const int64_t* prop, const int64_t* lo, const int64_t* hi,
                              const int64_t* klo, const int64_t* khi) {
  consts_t cs[8];
  for (int64_t r = 0; r < rows; ++r) {
    cs[r].q = (uint64_t)lo[r] + ((uint64_t)hi[r] << HALF_BITS);
    cs[r].k = (uint64_t)klo[r] + ((uint64_t)khi[r] << HALF_BITS);
  }
  for (int64_t bb = 0; bb < batch; ++bb) {
    const int64_t* s = src + bb * rows * n;
    int64_t* o = out + bb * rows * n;
    for (int64_t i = 0; i < n; ++i) {
      int64_t digits[8];
      int64_t first = s[i];
      first = first < (int64_t)cs[0].q ? first : first - (int64_t)cs[0].q;
      for (int64_t r = 0; r < rows; ++r) digits[r] = first;
      for (int64_t step = 0; step < rows - 1; ++step) {
        const int64_t row = step + 1;
        const int64_t d = mul_split(s[row * n + i] - digits[row], norm[step], &cs[row]);
        digits[row] = d;
        for (int64_t t = row + 1; t < rows; ++t) digits[t] += mul(d, prop[step * rows + t], &cs[t]);
      }
      for (int64_t r = 0; r < rows; ++r) o[r * n + i] = digits[r];
    }
  }
}

/* digits [batch, dcount, n] -> out [batch, dest_rows, n]; coef [dcount-1,
 * coef_stride] (column j = destination row j); params [8, dest_rows]. */
void or_basis_extend(int64_t* out, const int64_t* digits, int64_t batch, int64_t dcount,
                     int64_t dest_rows, int64_t n, const int64_t* coef, int64_t coef_stride,
                     const int64_t* params) {
  for (int64_t bb = 0; bb < batch; ++bb) {
    const int64_t* d = digits + bb * dcount * n;
    for (int64_t j = 0; j < dest_rows; ++j) {
      const consts_t c = load(params, dest_rows, j);
      int64_t* o = out + (bb * dest_rows + j) * n;
      for (int64_t i = 0; i < n; ++i) {
        int64_t v = mul(d[i], c.r2, &c);
        for (int64_t r = 1; r < dcount; ++r)
          v = add_lazy(v, mul(d[r * n + i], coef[(r - 1) * coef_stride + j], &c), c.twice_q);
        o[i] = v;
      }
    }
  }
}

/* remaining [batch, limbs, n], dropped [batch, n], inverse [limbs]. */
void or_rescale(int64_t* out, const int64_t* remaining, const int64_t* dropped,
                const int64_t* inverse, const int64_t* params, int64_t batch, int64_t limbs,
                int64_t n, int64_t half_drop, int nearest) {
  for (int64_t row = 0; row < batch * limbs; ++row) {
    const int64_t bb = row / limbs, limb = row % limbs;
    const consts_t c = load(params, limbs, limb);
    for (int64_t i = 0; i < n; ++i) {
      const int64_t dv = dropped[bb * n + i];
      int64_t v = mul_split(remaining[row * n + i] - dv, inverse[limb], &c);
      if (nearest) v += dv > half_drop ? 1 : 0;
      out[row * n + i] = canonicalize(v, c.twice_q);
    }
  }
}

void or_galois_coef(int64_t* out, const int64_t* in, const int32_t* src, const int8_t* sign,
                    const int64_t* twice, int64_t batch, int64_t limbs, int64_t n) {
  for (int64_t row = 0; row < batch * limbs; ++row) {
    const int64_t t = twice[row % limbs];
    for (int64_t i = 0; i < n; ++i) {
      int64_t v = in[row * n + src[i]];
      if (sign[i] == -1) v = -v;
      out[row * n + i] = canonicalize(v + (t >> 1), t);
    }
  }
}

void or_galois_ntt(int64_t* out, const int64_t* in, const int32_t* src, int64_t batch,
                   int64_t limbs, int64_t n) {
  for (int64_t row = 0; row < batch * limbs; ++row)
    for (int64_t i = 0; i < n; ++i) out[row * n + i] = in[row * n + src[i]];
}

/* acc0/acc1/digit [batch, limbs, n]; key [2, key_rows, n]. */
void or_keyswitch_mac(int64_t* acc0, int64_t* acc1, const int64_t* digit, const int64_t* key,
                      int64_t key_rows, const int64_t* params, int64_t key_row_start,
                      int64_t batch, int64_t limbs, int64_t n, int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(static)
  for (int64_t row = 0; row < batch * limbs; ++row) {
    const int64_t limb = row % limbs;
    const consts_t c = load(params, limbs, limb);
    const int64_t* k0 = key + (key_row_start + limb) * n;
    const int64_t* k1 = key + (key_rows + key_row_start + limb) * n;
    for (int64_t i = 0; i < n; ++i) {
      const int64_t x = digit[row * n + i];
      acc0[row * n + i] = add_lazy(acc0[row * n + i], mul(x, k0[i], &c), c.twice_q);
      acc1[row * n + i] = add_lazy(acc1[row * n + i], mul(x, k1[i], &c), c.twice_q);
    }
  }
}

/* q [batch, qc, n], p [batch, pc, n], inverse [pc, inv_width],
 * params [8, qc + pc]. */
void or_moddown(int64_t* out, const int64_t* qv, const int64_t* pv, const int64_t* inverse,
                int64_t inv_width, const int64_t* params, int64_t batch, int64_t qc,
                int64_t pc, int64_t n, int threads) {
  const int64_t limbs = qc + pc;
  set_threads(threads);
#pragma omp parallel for schedule(static)
  for (int64_t idx = 0; idx < batch * n; ++idx) {
    const int64_t bb = idx / n, i = idx % n;
    int64_t chain[8];
    for (int64_t r = 0; r < pc; ++r) chain[r] = pv[(bb * pc + r) * n + i];
    for (int64_t row = pc - 2; row >= 0; --row) {
      int64_t value = chain[row];
      const consts_t c = load(params, limbs, qc + row);
      for (int64_t lower = pc - 1; lower > row; --lower) {
        const int64_t diff = sub_lazy(value, chain[lower], c.twice_q);
        value = mul_split(diff, inverse[(pc - lower - 1) * inv_width + row + qc], &c);
      }
      chain[row] = value;
    }
    for (int64_t row = 0; row < qc; ++row) {
      const consts_t c = load(params, limbs, row);
      int64_t value = mul(qv[(bb * qc + row) * n + i], c.r2, &c);
      for (int64_t p = pc - 1; p >= 0; --p) {
        const int64_t pm = mul(chain[p], c.r2, &c);
        value = sub_lazy(value, pm, c.twice_q);
        value = mul_split(value, inverse[(pc - p - 1) * inv_width + row], &c);
      }
      value = mul(value, 1, &c);
      out[(bb * qc + row) * n + i] = canonicalize(value, c.twice_q);
    }
  }
}

/* layout 0 dense, 1 cyclic, 2 contiguous, 3 strided (+implicit [pt_batch, limbs]). */
void or_add_plaintext(int layout, int64_t* out, const int64_t* ct, const int64_t* pt,
                      const int64_t* implicit, int64_t pt_batch, const int64_t* params,
                      int64_t batch, int64_t limbs, int64_t n, int64_t support) {
  if (layout == 0) support = n;
  const int64_t repeat = n / support;
  for (int64_t row = 0; row < batch * limbs; ++row) {
    const int64_t bb = row / limbs, limb = row % limbs;
    const int64_t pb = pt_batch == 1 ? 0 : bb;
    const consts_t c = load(params, limbs, limb);
    const int64_t* prow = pt + (pb * limbs + limb) * support;
    for (int64_t i = 0; i < n; ++i) {
      int64_t p;
      switch (layout) {
        case 0: p = prow[i]; break;
        case 1: p = prow[i % support]; break;
        case 2: p = prow[i / repeat]; break;
        default:
          p = implicit[pb * limbs + limb];
          if (i % repeat == 0) p = prow[i / repeat];
          break;
      }
      int64_t v = mul(ct[row * n + i], c.r2, &c);
      v = add_lazy(v, p, c.twice_q);
      v = mul(v, 1, &c);
      out[row * n + i] = canonicalize(v, c.twice_q);
    }
  }
}
