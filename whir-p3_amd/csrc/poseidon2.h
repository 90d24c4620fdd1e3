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

// The permutation in three phases (initial linear layer and external rounds;
// internal rounds; terminal external rounds), so a caller can stop between
// them.
template <int W>
__host__ __device__ __forceinline__ void permute_head(u32* s, const P2Consts<W>& c) {
  external_linear<W>(s);
#pragma unroll
  for (int r = 0; r < 4; r++) {
#pragma unroll
    for (int i = 0; i < W; i++) s[i] = cube(add(s[i], c.ext_init[r * W + i]));
    external_linear<W>(s);
  }
}
template <int W>
__host__ __device__ __forceinline__ void permute_internal(u32* s, const P2Consts<W>& c) {
#pragma unroll
  for (int r = 0; r < P2Consts<W>::RP; r++) {
    s[0] = cube(add(s[0], c.internal[r]));
    internal_linear<W>(s);
  }
}
template <int W>
__host__ __device__ __forceinline__ void permute_tail(u32* s, const P2Consts<W>& c) {
#pragma unroll
  for (int r = 0; r < 4; r++) {
#pragma unroll
    for (int i = 0; i < W; i++) s[i] = cube(add(s[i], c.ext_term[r * W + i]));
    external_linear<W>(s);
  }
}

template <int W>
__host__ __device__ __forceinline__ void permute(u32* s, const P2Consts<W>& c) {
  permute_head<W>(s, c);
  permute_internal<W>(s, c);
  permute_tail<W>(s, c);
}

}  // namespace kb
