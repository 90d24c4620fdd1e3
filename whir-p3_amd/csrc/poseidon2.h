// Poseidon2 over KoalaBear for widths 16 and 24 (host and device).
//
// Structure (Plonky3 `p3-poseidon2` / `p3-koala-bear`, pinned by the reference
// at Cargo.toml:44-59; instantiated at src/bin/main.rs:100-102,
// benches/whir.rs:66-68): x^3 S-box; an initial external linear layer; 4
// external rounds; RP internal rounds (20 for width 16, 23 for width 24); 4
// external rounds.  The external layer applies the circulant M4 = (2 3 1 1) to
// each 4-chunk and then adds, to every element, the sum of the elements in
// the same position of all chunks.  The internal layer is 1 + diag(V).
//
// Round constants are data (kernel arguments), in Montgomery form.
#pragma once
#include "field.h"

namespace kb {

template <int W>
struct P2Consts {
  static constexpr int RP = W == 16 ? 20 : 23;
  u32 ext_init[4 * W];
  u32 ext_term[4 * W];
  u32 internal[RP];
};
using P2C16 = P2Consts<16>;
using P2C24 = P2Consts<24>;

__host__ __device__ __forceinline__ void mat4(u32& x0, u32& x1, u32& x2, u32& x3) {
  u32 t01 = add(x0, x1);
  u32 t23 = add(x2, x3);
  u32 t0123 = add(t01, t23);
  u32 t01123 = add(t0123, x1);
  u32 t01233 = add(t0123, x3);
  u32 n3 = add(t01233, dbl(x0));
  u32 n1 = add(t01123, dbl(x2));
  u32 n0 = add(t01123, t01);
  u32 n2 = add(t01233, t23);
  x0 = n0;
  x1 = n1;
  x2 = n2;
  x3 = n3;
}

template <int W>
__host__ __device__ __forceinline__ void external_linear(u32* s) {
#pragma unroll
  for (int c = 0; c < W; c += 4) mat4(s[c], s[c + 1], s[c + 2], s[c + 3]);
  u32 sum[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    sum[k] = s[k];
#pragma unroll
    for (int c = 4; c < W; c += 4) sum[k] = add(sum[k], s[c + k]);
  }
#pragma unroll
  for (int i = 0; i < W; i++) s[i] = add(s[i], sum[i & 3]);
}

__host__ __device__ __forceinline__ u32 cube(u32 x) { return mul(mul(x, x), x); }

// x * 2^-k for 1 <= k <= 32: one Montgomery reduction of x * 2^(32-k).
template <int K>
__host__ __device__ __forceinline__ u32 div2k(u32 x) {
  return redc((u64)x << (32 - K));
}
__host__ __device__ __forceinline__ u32 triple(u32 x) { return add(dbl(x), x); }

// Internal matrix 1 + diag(V), width 16:
// V = [-2, 1, 2, 1/2, 3, 4, -1/2, -3, -4, 1/2^8, 1/8, 1/2^24, -1/2^8, -1/8, -1/16, -1/2^24]
__host__ __device__ __forceinline__ void internal_linear16(u32* s) {
  u32 sum = s[0];
#pragma unroll
  for (int i = 1; i < 16; i++) sum = add(sum, s[i]);
  s[0] = sub(sum, dbl(s[0]));
  s[1] = add(sum, s[1]);
  s[2] = add(sum, dbl(s[2]));
  s[3] = add(sum, halve(s[3]));
  s[4] = add(sum, triple(s[4]));
  s[5] = add(sum, dbl(dbl(s[5])));
  s[6] = sub(sum, halve(s[6]));
  s[7] = sub(sum, triple(s[7]));
  s[8] = sub(sum, dbl(dbl(s[8])));
  s[9] = add(sum, div2k<8>(s[9]));
  s[10] = add(sum, div2k<3>(s[10]));
  s[11] = add(sum, div2k<24>(s[11]));
  s[12] = sub(sum, div2k<8>(s[12]));
  s[13] = sub(sum, div2k<3>(s[13]));
  s[14] = sub(sum, div2k<4>(s[14]));
  s[15] = sub(sum, div2k<24>(s[15]));
}

// Width 24: V = [-2, 1, 2, 1/2, 3, 4, -1/2, -3, -4, 1/2^8, 1/4, 1/8, 1/16, 1/32,
//                1/64, 1/2^24, -1/2^8, -1/8, -1/16, -1/32, -1/64, -1/2^7, -1/2^9, -1/2^24]
__host__ __device__ __forceinline__ void internal_linear24(u32* s) {
  u32 sum = s[0];
#pragma unroll
  for (int i = 1; i < 24; i++) sum = add(sum, s[i]);
  s[0] = sub(sum, dbl(s[0]));
  s[1] = add(sum, s[1]);
  s[2] = add(sum, dbl(s[2]));
  s[3] = add(sum, halve(s[3]));
  s[4] = add(sum, triple(s[4]));
  s[5] = add(sum, dbl(dbl(s[5])));
  s[6] = sub(sum, halve(s[6]));
  s[7] = sub(sum, triple(s[7]));
  s[8] = sub(sum, dbl(dbl(s[8])));
  s[9] = add(sum, div2k<8>(s[9]));
  s[10] = add(sum, div2k<2>(s[10]));
  s[11] = add(sum, div2k<3>(s[11]));
  s[12] = add(sum, div2k<4>(s[12]));
  s[13] = add(sum, div2k<5>(s[13]));
  s[14] = add(sum, div2k<6>(s[14]));
  s[15] = add(sum, div2k<24>(s[15]));
  s[16] = sub(sum, div2k<8>(s[16]));
  s[17] = sub(sum, div2k<3>(s[17]));
  s[18] = sub(sum, div2k<4>(s[18]));
  s[19] = sub(sum, div2k<5>(s[19]));
  s[20] = sub(sum, div2k<6>(s[20]));
  s[21] = sub(sum, div2k<7>(s[21]));
  s[22] = sub(sum, div2k<9>(s[22]));
  s[23] = sub(sum, div2k<24>(s[23]));
}

template <int W>
__host__ __device__ __forceinline__ void internal_linear(u32* s) {
  if constexpr (W == 16) internal_linear16(s);
  else internal_linear24(s);
}

// ---- width 16 with lazy reduction
// The linear layers add in 64 bits (one full-rate instruction per add on
// CDNA) and each result is reduced once.  Because 2^31 = 2^24 - 1 (mod p), a
// value y < 2^37.9 splits as y = hi 2^31 + lo with hi < 126 and
//   y = lo + hi (2^24 - 1) (mod p),  which is < 2p,
// so lazy_reduce (three instructions) gives a representative below 2p and
// full_reduce adds one conditional subtraction.  Between internal rounds
// elements 1..15 stay below 2p; S-box inputs are always fully reduced.
//
// Internal diagonal entries 2^-k use p = 127 2^24 + 1: 2^-k = -127 2^(24-k),
// so for z = q 2^k + r (r < 2^k)
//   z 2^-k = q - r 127 2^(24-k)   (mod p),
// with no multiplication by a full-size constant.  Negative entries -d use
// 2p - z, which is -z mod p and positive for z < 2p.
//
// Every intermediate bound is stated where it is used; the largest 64-bit
// value reduced is below 71p < 2^37.2.
// sum + z 2^-K and sum - z 2^-K for z < 2p; `sumk` = sum + 2p keeps both
// nonnegative (r 127 2^(24-K) < 2^31 < 2p, and z >> K < p).
template <int K>
__host__ __device__ __forceinline__ u64 add_div2k(u64 sumk, u32 z, u32 one) {
  constexpr int32_t CK = 127 << (24 - K);
  return mad_i64((int32_t)opaque(z & ((1u << K) - 1)), -CK, mad_u64(z >> K, one, sumk));
}
template <int K>
__host__ __device__ __forceinline__ u64 sub_div2k(u64 sumk, u32 z, u32 one) {
  constexpr u32 CK = 127u << (24 - K);
  return mad_u64(opaque(z & ((1u << K) - 1)), CK, mad_u64(P - (z >> K), one, sumk));   // -q = p - q (mod p)
}

// Montgomery products without the final conditional subtraction.  For
// x < p, x^2 / 2^32 mod p comes out below p^2 / 2^32 + p < 1.5p, and then
// x^3 / 2^64 mod p below 1.5p x / 2^32 + p < 1.75p; neither 64-bit sum
// overflows (1.5p^2 + 2^32 p < 2^64).  cube_lazy returns a 64-bit value with a
// zero high word.
__host__ __device__ __forceinline__ u32 sqr_lazy(u32 x) {
  const u64 m = (u64)x * x;
  const u32 t = (u32)m * MU_NEG;
  return (u32)(((u64)t * P + m) >> 32);
}
__host__ __device__ __forceinline__ u64 cube_lazy(u32 x) {
  const u32 x2 = sqr_lazy(x);
  const u64 m = (u64)x2 * x;
  const u32 t = (u32)m * MU_NEG;
  return ((u64)t * P + m) >> 32;
}

// External layer on inputs x < 2p: M4 on each quad, then each element plus
// the sum of its position over the quads.  Outputs < 70p.
__host__ __device__ __forceinline__ void external_linear16_wide(const u32* x, u64* y, u32 one, u32 two, u64 zero) {
  u64 n[16];
#pragma unroll
  for (int c = 0; c < 16; c += 4) {
    const u32 x0 = x[c], x1 = x[c + 1], x2 = x[c + 2], x3 = x[c + 3];
    const u64 t01 = mad_u64(x1, one, mad_u64(x0, one, zero));
    const u64 t23 = mad_u64(x3, one, mad_u64(x2, one, zero));
    const u64 t0123 = t01 + t23;
    const u64 t01123 = mad_u64(x1, one, t0123);
    const u64 t01233 = mad_u64(x3, one, t0123);
    n[c + 3] = mad_u64(x0, two, t01233);   // 3x0 + x1 + x2 + 2x3
    n[c + 1] = mad_u64(x2, two, t01123);   // x0 + 2x1 + 3x2 + x3
    n[c + 0] = t01123 + t01;              // 2x0 + 3x1 + x2 + x3
    n[c + 2] = t01233 + t23;              // x0 + x1 + 2x2 + 3x3
  }
  u64 s[4];
#pragma unroll
  for (int k = 0; k < 4; k++) s[k] = n[k] + n[4 + k] + n[8 + k] + n[12 + k];
#pragma unroll
  for (int i = 0; i < 16; i++) y[i] = n[i] + s[i & 3];
}

// The four initial external rounds from the initial linear layer's output y
// (each y[i] < 77p).  Out: s[0] < p, s[1..15] < 2p.
__host__ __device__ __forceinline__ void permute16_head_rounds(u64* y, u32* s, const P2C16& c) {
  const u32 one = opaque(1u);
  const u32 two = opaque(2u);
  const u64 zero = opaque(0ull);
  u32 x[16];
#pragma unroll
  for (int r = 0; r < 4; r++) {
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = (u32)cube_lazy(full_reduce(y[i] + c.ext_init[r * 16 + i]));   // < 71p
    external_linear16_wide(x, y, one, two, zero);
  }
  s[0] = full_reduce(y[0]);
#pragma unroll
  for (int i = 1; i < 16; i++) s[i] = (u32)lazy_reduce(y[i]);
}

// Initial linear layer and the four initial external rounds.  In: s < p.
// Out: s[0] < p, s[1..15] < 2p.
__host__ __device__ __forceinline__ void permute16_head(u32* s, const P2C16& c) {
  u64 y[16];
  external_linear16_wide(s, y, opaque(1u), opaque(2u), opaque(0ull));   // < 35p
  permute16_head_rounds(y, s, c);
}

// The twenty internal rounds: S-box on element 0, then 1 + diag(V) with
// V = [-2, 1, 2, 1/2, 3, 4, -1/2, -3, -4, 1/2^8, 1/8, 1/2^24, -1/2^8, -1/8,
// -1/16, -1/2^24].  In and out: s[0] < p, s[1..15] < 2p.
__host__ __device__ __forceinline__ void permute16_internal(u32* s, const P2C16& c) {
  const u32 one = opaque(1u);
  const u32 two = opaque(2u), four = opaque(4u);
  u64 z[16];
#pragma unroll
  for (int i = 1; i < 16; i++) z[i] = s[i];
  u64 y0 = s[0];
#pragma clang loop unroll(full)
  for (int r = 0; r < P2C16::RP; r++) {
    // The sum of z[1..15] does not wait for the S-box: it is added last.
    u64 sz = z[1];
#pragma unroll
    for (int i = 2; i < 16; i++) sz += z[i];   // < 30p
    const u32 x0 = (u32)cube_lazy(full_reduce(y0 + c.internal[r]));   // y0 + rc < 37p; x0 < 1.75p
    const u64 sum = mad_u64(x0, one, sz);   // < 32p
    const u64 sumk = sum + TWO_P;
    y0 = mad_u64(TWO_P - x0, two, sum);   // sum - 2 x0 + 4p, < 36p
    z[1] = lazy_reduce(sum + z[1]);
    z[2] = lazy_reduce(sum + 2 * z[2]);
    z[3] = lazy_reduce(add_div2k<1>(sumk, (u32)z[3], one));
    z[4] = lazy_reduce(mad_u64((u32)z[4], 3u, sum));
    z[5] = lazy_reduce(sum + 4 * z[5]);
    z[6] = lazy_reduce(sub_div2k<1>(sumk, (u32)z[6], one));
    z[7] = lazy_reduce(mad_u64(TWO_P - (u32)z[7], 3u, sum));
    z[8] = lazy_reduce(mad_u64(TWO_P - (u32)z[8], four, sum));
    z[9] = lazy_reduce(add_div2k<8>(sumk, (u32)z[9], one));
    z[10] = lazy_reduce(add_div2k<3>(sumk, (u32)z[10], one));
    z[11] = lazy_reduce(add_div2k<24>(sumk, (u32)z[11], one));
    z[12] = lazy_reduce(sub_div2k<8>(sumk, (u32)z[12], one));
    z[13] = lazy_reduce(sub_div2k<3>(sumk, (u32)z[13], one));
    z[14] = lazy_reduce(sub_div2k<4>(sumk, (u32)z[14], one));
    z[15] = lazy_reduce(sub_div2k<24>(sumk, (u32)z[15], one));
  }
  s[0] = full_reduce(y0);
#pragma unroll
  for (int i = 1; i < 16; i++) s[i] = (u32)z[i];
}

// The four terminal external rounds.  In: s[0] < p, s[1..15] < 2p.  Out: s < p.
__host__ __device__ __forceinline__ void permute16_tail(u32* s, const P2C16& c) {
  const u32 one = opaque(1u);
  const u32 two = opaque(2u);
  const u64 zero = opaque(0ull);
  u32 x[16];
  u64 y[16];
#pragma unroll
  for (int i = 0; i < 16; i++) y[i] = s[i];
#pragma unroll
  for (int r = 0; r < 4; r++) {
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = (u32)cube_lazy(full_reduce(y[i] + c.ext_term[r * 16 + i]));
    external_linear16_wide(x, y, one, two, zero);
  }
#pragma unroll
  for (int i = 0; i < 16; i++) s[i] = full_reduce(y[i]);
}

// ---- grinding candidates
// A candidate's input state is a fixed state with the word at position q
// replaced by the witness w, so the initial linear layer M_E x is
// M_E x' + w M_E e_q: M_E x' (x' the fixed state with 0 at q) is computed once
// per search, and column q of M_E has the small entries
// M4[i mod 4][q mod 4], doubled in q's own quad (M4 = 2 3 1 1 / 1 2 3 1 /
// 1 1 2 3 / 3 1 1 2).  A candidate's layer is then 16 multiply-adds.
struct GrindPre {
  u64 y0[16];    // M_E x', each below 35p
  u32 col[16];   // column q of M_E, 1..6
};
__host__ __device__ inline void grind_pre(GrindPre& g, const u32* base, int q) {
  u32 x[16];
#pragma unroll
  for (int i = 0; i < 16; i++) x[i] = i == q ? 0u : base[i];
  external_linear16_wide(x, g.y0, opaque(1u), opaque(2u), opaque(0ull));
  constexpr u32 M4[4][4] = {{2, 3, 1, 1}, {1, 2, 3, 1}, {1, 1, 2, 3}, {3, 1, 1, 2}};
#pragma unroll
  for (int i = 0; i < 16; i++) g.col[i] = M4[i & 3][q & 3] * ((i >> 2) == (q >> 2) ? 2u : 1u);
}
// The initial linear layer's output for witness wm (Montgomery form, < p):
// below 35p + 6p.
__host__ __device__ __forceinline__ void grind_layer(const GrindPre& g, u32 wm, u64* y) {
#pragma unroll
  for (int i = 0; i < 16; i++) y[i] = mad_u64(wm, g.col[i], g.y0[i]);
}

// The four terminal external rounds, returning only element 7 of the output
// (canonical; the sample a grinding candidate is judged by).  The last
// linear layer's element 7 is n_7 + sum_q n_(4q+3) with n_(4q+3) = 3 x_4q +
// x_(4q+1) + x_(4q+2) + 2 x_(4q+3), so quad 1 counts twice: 16 multiply-adds
// (x < 1.75p: below 62p) and one reduction instead of the whole layer and
// sixteen reductions.  In: s[0] < p, s[1..15] < 2p.
__host__ __device__ __forceinline__ u32 permute16_tail_s7(u32* s, const P2C16& c) {
  const u32 one = opaque(1u);
  const u32 two = opaque(2u);
  const u64 zero = opaque(0ull);
  u32 x[16];
  u64 y[16];
#pragma unroll
  for (int i = 0; i < 16; i++) y[i] = s[i];
#pragma unroll
  for (int r = 0; r < 3; r++) {
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = (u32)cube_lazy(full_reduce(y[i] + c.ext_term[r * 16 + i]));
    external_linear16_wide(x, y, one, two, zero);
  }
#pragma unroll
  for (int i = 0; i < 16; i++) x[i] = (u32)cube_lazy(full_reduce(y[i] + c.ext_term[48 + i]));
  const u32 w1 = opaque(1u), w2 = opaque(2u), w3 = opaque(3u), w4 = opaque(4u), w6 = opaque(6u);
  u64 acc[2] = {zero, zero};
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const bool d = q == 1;   // quad 1 counts twice
    u64& a = acc[q & 1];
    a = mad_u64(x[4 * q], d ? w6 : w3, a);
    a = mad_u64(x[4 * q + 1], d ? w2 : w1, a);
    a = mad_u64(x[4 * q + 2], d ? w2 : w1, a);
    a = mad_u64(x[4 * q + 3], d ? w4 : w2, a);
  }
  return full_reduce(acc[0] + acc[1]);
}

// The permutation in three phases (initial linear layer and external rounds;
// internal rounds; terminal external rounds), so a caller can stop between
// them.  Width 16 takes the lazy-reduction path above.
template <int W>
__host__ __device__ __forceinline__ void permute_head(u32* s, const P2Consts<W>& c) {
  if constexpr (W == 16) {
    permute16_head(s, c);
  } else {
    external_linear<W>(s);
#pragma unroll
    for (int r = 0; r < 4; r++) {
#pragma unroll
      for (int i = 0; i < W; i++) s[i] = cube(add(s[i], c.ext_init[r * W + i]));
      external_linear<W>(s);
    }
  }
}
template <int W>
__host__ __device__ __forceinline__ void permute_internal(u32* s, const P2Consts<W>& c) {
  if constexpr (W == 16) {
    permute16_internal(s, c);
  } else {
#pragma unroll
    for (int r = 0; r < P2Consts<W>::RP; r++) {
      s[0] = cube(add(s[0], c.internal[r]));
      internal_linear<W>(s);
    }
  }
}
template <int W>
__host__ __device__ __forceinline__ void permute_tail(u32* s, const P2Consts<W>& c) {
  if constexpr (W == 16) {
    permute16_tail(s, c);
  } else {
#pragma unroll
    for (int r = 0; r < 4; r++) {
#pragma unroll
      for (int i = 0; i < W; i++) s[i] = cube(add(s[i], c.ext_term[r * W + i]));
      external_linear<W>(s);
    }
  }
}

template <int W>
__host__ __device__ __forceinline__ void permute(u32* s, const P2Consts<W>& c) {
  permute_head<W>(s, c);
  permute_internal<W>(s, c);
  permute_tail<W>(s, c);
}

// The textbook width-16 permutation (one reduction per operation), kept as the
// reference the lazy path is tested against.
__host__ __device__ inline void permute16_plain(u32* s, const P2C16& c) {
  external_linear<16>(s);
  for (int r = 0; r < 4; r++) {
    for (int i = 0; i < 16; i++) s[i] = cube(add(s[i], c.ext_init[r * 16 + i]));
    external_linear<16>(s);
  }
  for (int r = 0; r < P2C16::RP; r++) {
    s[0] = cube(add(s[0], c.internal[r]));
    internal_linear16(s);
  }
  for (int r = 0; r < 4; r++) {
    for (int i = 0; i < 16; i++) s[i] = cube(add(s[i], c.ext_term[r * 16 + i]));
    external_linear<16>(s);
  }
}

}  // namespace kb
