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
// Sum of the 16 lanes of a DPP row for x < p, in every lane of the row: four
// rotated adds (row_ror 8, 4, 2, 1), each sum below 2p brought below p.
template <int N>
__device__ __forceinline__ u32 row_ror(u32 x) {
  return (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x120 + N, 0xF, 0xF, false);
}
__device__ __forceinline__ u32 row_sum16(u32 x) {
  u32 s = x + row_ror<8>(x);
  s = s - P < s ? s - P : s;
  s = s + row_ror<4>(s);
  s = s - P < s ? s - P : s;
  s = s + row_ror<2>(s);
  s = s - P < s ? s - P : s;
  s = s + row_ror<1>(s);
  return s - P < s ? s - P : s;
}
__device__ __forceinline__ ef row_sum16(const ef& a) {
  return ef{{row_sum16(a.c[0]), row_sum16(a.c[1]), row_sum16(a.c[2]), row_sum16(a.c[3])}};
}

// Every thread gets the block's sums: row sums over DPP, the 16 row sums
// through LDS (one barrier), and a second row sum over them.  Called by all
// 256 threads in uniform control flow; sh holds 16 * NV entries, and a second
// call with the same sh needs a barrier in between (the reads of this call
// come after its only barrier).
template <int NV>
__device__ __forceinline__ void block_reduce_ef(ef (&v)[NV], ef* sh /* 16 * NV */) {
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    v[k] = row_sum16(v[k]);
    if ((t & 15) == 0) sh[k * 16 + (t >> 4)] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = row_sum16(sh[k * 16 + (t & 15)]);
}

}  // namespace kb
