// KoalaBear (p = 2^31 - 2^24 + 1) in Montgomery form (R = 2^32) and its
// degree-4 binomial extension F[x]/(x^4 - 3), usable on host and device.
//
// Elements are stored as their Montgomery representative a*R mod p, which is
// the in-memory layout of Plonky3's `KoalaBear` (p3-koala-bear, pinned by the
// reference at Cargo.toml:44-59), so host buffers can be shared zero-copy.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace kb {

using u32 = uint32_t;
using u64 = uint64_t;

constexpr u32 P = 0x7f000001u;
constexpr u32 MU = 0x81000001u;  // p^-1 mod 2^32
constexpr u32 W_EXT = 3;         // x^4 = 3

// 2^32 mod p = 2^25 - 2, and 2^64 mod p for the conversion into Montgomery form.
constexpr u32 R1 = 0x01fffffeu;
constexpr u32 R2 = (u32)(((u64)R1 * R1) % P);

__host__ __device__ __forceinline__ u32 add(u32 a, u32 b) {
  u32 s = a + b;
  u32 t = s - P;
  return t < s ? t : s;
}
__host__ __device__ __forceinline__ u32 sub(u32 a, u32 b) {
  u32 d = a - b;
  u32 t = d + P;
  return t < d ? t : d;
}
__host__ __device__ __forceinline__ u32 neg(u32 a) { return a ? P - a : 0; }
__host__ __device__ __forceinline__ u32 dbl(u32 a) { return add(a, a); }

// Montgomery reduction of x < p * 2^32: returns x / 2^32 mod p in [0, p).
// With t = x * p^-1 mod 2^32, x - t p is divisible by 2^32 and the quotient
// r = hi(x) - hi(t p) lies in (-p, p); as u32, r + p wraps below r exactly
// when r < 0, so min(r, r + p) is the reduced value.
__host__ __device__ __forceinline__ u32 redc(u64 x) {
  u32 xl = (u32)x, xh = (u32)(x >> 32);
  u32 t = xl * MU;
  u32 uh = (u32)(((u64)t * P) >> 32);
  u32 r = xh - uh;
  u32 r2 = r + P;
  return r2 < r ? r2 : r;
}
// Montgomery reduction of x < 4 p^2 (a sum of up to four products).  Here the
// quotient lies in (-p, 2p), so the min trick of redc() does not apply to the
// sign: a borrow (r > xh) marks r < 0, which adding p folds into (0, p);
// min(r, r - p) then folds [p, 2p) into [0, p).
__host__ __device__ __forceinline__ u32 redc_wide(u64 x) {
  u32 xl = (u32)x, xh = (u32)(x >> 32);
  u32 t = xl * MU;
  u32 uh = (u32)(((u64)t * P) >> 32);
  u32 r = xh - uh;
  r = r > xh ? r + P : r;
  u32 r3 = r - P;
  return r3 < r ? r3 : r;
}
// Montgomery product a b / 2^32 mod p for a, b < p: with t = -(a b) p^-1 mod
// 2^32, y = a b + t p is divisible by 2^32 and y / 2^32 < 2p, so one
// conditional subtraction finishes it.  Two 32x32+64 multiply-adds and a
// 32-bit multiply, all full-rate on CDNA.
constexpr u32 MU_NEG = 0u - MU;   // -p^-1 mod 2^32
__host__ __device__ __forceinline__ u32 mul(u32 a, u32 b) {
  const u64 x = (u64)a * b;
  const u32 t = (u32)x * MU_NEG;
  const u32 r = (u32)(((u64)t * P + x) >> 32);
  const u32 s = r - P;
  return s < r ? s : r;
}

// ---- lazy reduction
// Sums kept in 64 bits and reduced once.  Because 2^31 = 2^24 - 1 (mod p), a
// value y < 2^37.9 splits as y = hi 2^31 + lo with hi < 126 and
//   y = lo + hi (2^24 - 1) (mod p),  which is < 2p,
// so lazy_reduce (two instructions) gives a representative below 2p and
// full_reduce adds one conditional subtraction.
constexpr u32 TWO_P = 2u * P;

// 32x32+64 multiply-adds (v_mad_u64_u32 / v_mad_i64_i32, one full-rate
// instruction each on CDNA).
__host__ __device__ __forceinline__ u64 mad_u64(u32 a, u32 b, u64 c) { return (u64)a * b + c; }
__host__ __device__ __forceinline__ u64 mad_i64(int32_t a, int32_t b, u64 c) {
  return (u64)((int64_t)a * (int64_t)b + (int64_t)c);
}

// y - (y >> 31) p = (y mod 2^31) + (y >> 31)(2^24 - 1): for y < 2^37.9 the
// quotient is below 126 and the result below 2p.  Kept 64-bit (its high word
// is zero) so it feeds 64-bit adds directly.
// y >> 31 as a 32-bit value.  On the device it comes from v_alignbit_b32,
// which keeps the compiler from widening the multiply-add below into a 64x64
// product.
__host__ __device__ __forceinline__ u32 quot31(u64 y) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_alignbit((u32)(y >> 32), (u32)y, 31);
#else
  return (u32)(y >> 31);
#endif
}
// Hides a value's origin from the compiler (no instruction), so it keeps the
// multiply-add forms below instead of rewriting them into longer sequences
// (24-bit products, 64-bit subtractions, selects).  A 32-bit value added to a
// 64-bit one is written mad_u64(x, one, y) with one = opaque(1u): a single
// v_mad_u64_u32, where a plain add would first widen x with a zeroed high
// word.
template <typename T>
__host__ __device__ __forceinline__ T opaque(T x) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm("" : "+v"(x));
#endif
  return x;
}

__host__ __device__ __forceinline__ u64 lazy_reduce(u64 y) {
  return mad_i64((int32_t)quot31(y), -(int32_t)P, y);
}
__host__ __device__ __forceinline__ u32 full_reduce(u64 y) {   // y < 2^37.9 -> [0, p)
  const u32 z = (u32)lazy_reduce(y);
  const u32 t = z - P;
  return t < z ? t : z;
}

// Montgomery product without the final subtraction, for b w < 2^32 p (for
// instance b < 2^32 and w < p, or b < p and w < 2p): (b w + t p) / 2^32 is
// then below 2p, a representative that fits 32 bits.
__host__ __device__ __forceinline__ u32 mul_lazy(u32 b, u32 w) {
  const u64 x = (u64)b * w;
  const u32 t = (u32)x * MU_NEG;
  return (u32)(((u64)t * P + x) >> 32);
}
// [0, 2p) -> [0, p)
__host__ __device__ __forceinline__ u32 canon(u32 x) {
  const u32 t = x - P;
  return t < x ? t : x;
}

__host__ __device__ __forceinline__ u32 to_monty(u32 x) { return mul(x, R2); }
__host__ __device__ __forceinline__ u32 from_monty(u32 m) { return redc((u64)m); }

constexpr u32 ONE = R1;  // Montgomery form of 1

__host__ __device__ inline u32 pow(u32 a, u64 e) {
  u32 r = ONE;
  while (e) {
    if (e & 1) r = mul(r, a);
    a = mul(a, a);
    e >>= 1;
  }
  return r;
}
__host__ __device__ inline u32 inv(u32 a) { return pow(a, P - 2); }

// Halving in Montgomery form (the representative is linear in the value):
// a/2 for even a, (a + p)/2 = (a >> 1) + (p + 1)/2 for odd a.
__host__ __device__ __forceinline__ u32 halve(u32 a) { return (a >> 1) + ((0u - (a & 1)) & ((P + 1) >> 1)); }

// ------------------------------------------------------------------ EF4
struct alignas(16) ef {
  u32 c[4];
};

__host__ __device__ __forceinline__ ef ef_zero() { return ef{{0, 0, 0, 0}}; }
__host__ __device__ __forceinline__ ef ef_one() { return ef{{ONE, 0, 0, 0}}; }
__host__ __device__ __forceinline__ ef ef_base(u32 a) { return ef{{a, 0, 0, 0}}; }
__host__ __device__ __forceinline__ bool ef_eq(const ef& a, const ef& b) {
  return a.c[0] == b.c[0] && a.c[1] == b.c[1] && a.c[2] == b.c[2] && a.c[3] == b.c[3];
}
__host__ __device__ __forceinline__ ef ef_add(const ef& a, const ef& b) {
  return ef{{add(a.c[0], b.c[0]), add(a.c[1], b.c[1]), add(a.c[2], b.c[2]), add(a.c[3], b.c[3])}};
}
__host__ __device__ __forceinline__ ef ef_sub(const ef& a, const ef& b) {
  return ef{{sub(a.c[0], b.c[0]), sub(a.c[1], b.c[1]), sub(a.c[2], b.c[2]), sub(a.c[3], b.c[3])}};
}
__host__ __device__ __forceinline__ ef ef_dbl(const ef& a) { return ef_add(a, a); }
__host__ __device__ __forceinline__ ef ef_mul_base(const ef& a, u32 s) {
  return ef{{mul(a.c[0], s), mul(a.c[1], s), mul(a.c[2], s), mul(a.c[3], s)}};
}

// Schoolbook product with x^4 = 3.  The x^4-wrapping terms use 3*b, so every
// output coefficient is a sum of exactly four 62-bit products (< 4 p^2),
// reduced once.
__host__ __device__ __forceinline__ ef ef_mul(const ef& a, const ef& b) {
  u32 b1 = add(dbl(b.c[1]), b.c[1]);
  u32 b2 = add(dbl(b.c[2]), b.c[2]);
  u32 b3 = add(dbl(b.c[3]), b.c[3]);
  u64 c0 = (u64)a.c[0] * b.c[0] + (u64)a.c[1] * b3 + (u64)a.c[2] * b2 + (u64)a.c[3] * b1;
  u64 c1 = (u64)a.c[0] * b.c[1] + (u64)a.c[1] * b.c[0] + (u64)a.c[2] * b3 + (u64)a.c[3] * b2;
  u64 c2 = (u64)a.c[0] * b.c[2] + (u64)a.c[1] * b.c[1] + (u64)a.c[2] * b.c[0] + (u64)a.c[3] * b3;
  u64 c3 = (u64)a.c[0] * b.c[3] + (u64)a.c[1] * b.c[2] + (u64)a.c[2] * b.c[1] + (u64)a.c[3] * b.c[0];
  return ef{{redc_wide(c0), redc_wide(c1), redc_wide(c2), redc_wide(c3)}};
}

__host__ __device__ inline ef ef_pow(ef a, u64 e) {
  ef r = ef_one();
  while (e) {
    if (e & 1) r = ef_mul(r, a);
    a = ef_mul(a, a);
    e >>= 1;
  }
  return r;
}

// Generator of the 2^bits-th roots of unity: (3^127)^(2^(24-bits)), since
// p - 1 = 127 * 2^24 and 3 generates the multiplicative group.
__host__ __device__ inline u32 two_adic_generator(int bits) {
  u32 g = pow(to_monty(3), 127);
  for (int i = bits; i < 24; i++) g = mul(g, g);
  return g;
}

}  // namespace kb
