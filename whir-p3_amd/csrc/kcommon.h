// Shared device helpers: element-generic arithmetic, twiddle lookup, block
// reductions, and launch checking.
#pragma once
#include <hip/hip_runtime.h>

#include "field.h"

namespace kb {

// Four words as one register quad (16-byte loads, including nontemporal ones).
typedef u32 u32x4 __attribute__((ext_vector_type(4)));

// ---- element-generic arithmetic (base field u32 or extension ef)
__device__ __forceinline__ u32 tadd(u32 a, u32 b) { return add(a, b); }
__device__ __forceinline__ ef tadd(const ef& a, const ef& b) { return ef_add(a, b); }
__device__ __forceinline__ u32 tsub(u32 a, u32 b) { return sub(a, b); }
__device__ __forceinline__ ef tsub(const ef& a, const ef& b) { return ef_sub(a, b); }
__device__ __forceinline__ u32 tmulb(u32 a, u32 w) { return mul(a, w); }
__device__ __forceinline__ ef tmulb(const ef& a, u32 w) { return ef_mul_base(a, w); }

template <typename T>
__device__ __forceinline__ T tzero();
template <>
__device__ __forceinline__ u32 tzero<u32>() { return 0; }
template <>
__device__ __forceinline__ ef tzero<ef>() { return ef_zero(); }

// ---- roots of unity: w24 = generator of order 2^24.
// tw_lo[e] = w24^e and tw_hi[e] = w24^(4096 e) for e < 4096, so that any
// power of any 2^s-th root is two table reads and one product.
struct Twiddles {
  const u32* lo;
  const u32* hi;
};

// (w_{2^s})^x
__device__ __forceinline__ u32 root_pow(const Twiddles& tw, int s, u64 x) {
  u32 e = (u32)((x << (24 - s)) & 0xffffffu);
  return mul(tw.hi[e >> 12], tw.lo[e & 4095u]);
}

__device__ __forceinline__ u32 bitrev(u32 x, int bits) { return bits ? (__brev(x) >> (32 - bits)) : 0u; }

// ---- block reduction of NV extension elements per thread (blockDim = 256)
template <int NV>
__device__ __forceinline__ void block_reduce_ef(ef (&v)[NV], ef* sh /* 256 * NV */) {
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < NV; k++) sh[k * 256 + t] = v[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
#pragma unroll
      for (int k = 0; k < NV; k++) sh[k * 256 + t] = ef_add(sh[k * 256 + t], sh[k * 256 + t + s]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = sh[k * 256];
}

}  // namespace kb
