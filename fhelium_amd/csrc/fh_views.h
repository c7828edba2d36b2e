// Strided operand views shared by the element-wise kernels and the C ABI.
#pragma once

#include <cstdint>


#include "fhelium_amd.h"  // fh_view3, fh_table2, fh_vec and every entry point

#ifdef __HIPCC__
namespace fh {

struct View3 {
  int64_t* p;
  int64_t s0, s1, s2;
  __host__ __device__ View3() : p(nullptr), s0(0), s1(0), s2(0) {}
  __host__ __device__ explicit View3(const fh_view3& v)
      : p(v.data), s0(v.stride_batch), s1(v.stride_limb), s2(v.stride_coef) {}
  __device__ __forceinline__ int64_t& at(int64_t b, int64_t l, int64_t i) const {
    return p[b * s0 + l * s1 + i * s2];
  }
};

struct Vec {
  const int64_t* p;
  int64_t s;
  __host__ __device__ Vec() : p(nullptr), s(0) {}
  __host__ __device__ explicit Vec(const fh_vec& v) : p(v.data), s(v.stride) {}
  __device__ __forceinline__ int64_t operator[](int64_t i) const { return p[i * s]; }
};

}  // namespace fh
#endif
