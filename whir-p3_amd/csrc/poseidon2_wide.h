// Lane-parallel Poseidon2-16: one permutation spread over a 16-lane DPP row,
// lane j holding state element j.  It trades ~3x more lane-instructions for
// ~5x lower latency than the one-lane form in poseidon2.h, which is what the
// latency-bound parts need (upper Merkle levels, a handful of candidates).
//
// Same permutation as poseidon2.h: the external layer applies the circulant
// M4 = (2 3 1 1) to each quad of lanes (quad rotations bring x_{q+1..q+3}),
// then adds the sum of the four quads' same-position elements (row rotations
// by 4, 8, 12); the internal layer adds the 16-lane sum (rotations by 8, 4, 2,
// 1) to diag_j * s_j.  Every lane of the wave must be active.
//
// Arithmetic: every lane holds a value below p between layers.  S-boxes use
// the lazy Montgomery products of poseidon2.h and one conditional
// subtraction; the 16-lane and cross-quad sums are 32-bit rotated adds (DPP
// operands fused into the add), each below 2p and brought below p; the M4
// step adds in 64 bits and is reduced once, together with the next round
// constant.
#pragma once
#include "poseidon2.h"

namespace kb {

template <int CTRL>
__device__ __forceinline__ u32 dpp(u32 x) {
  return (u32)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
}
// quad_perm selectors: lane q reads lane (q + k) & 3 of its quad
constexpr int QROT1 = 1 | (2 << 2) | (3 << 4) | (0 << 6);
constexpr int QROT2 = 2 | (3 << 2) | (0 << 4) | (1 << 6);
constexpr int QROT3 = 3 | (0 << 2) | (1 << 4) | (2 << 6);
// row_ror:n within a 16-lane row
constexpr int ROR(int n) { return 0x120 + n; }

// Lane exchange with no old value: every source lane of these selectors lies
// in the row, so the out-of-row fill (zero) is never taken.
template <int CTRL>
__device__ __forceinline__ u32 xdpp(u32 x) {
  return (u32)__builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, true);
}
// x - p if x >= p, for x < 2p.
__device__ __forceinline__ u32 csub(u32 x) {
  const u32 t = x - P;
  return t < x ? t : x;
}

// External layer for x < p.  By linearity, M4 of each quad plus the sum of
// M4 over the quads is M4 applied to z = x + X, where X is the sum of the
// four quads' elements at this position (32-bit rotated adds, each sum below
// 2p brought below p).  Out: 2z + 3 z1 + z2 + z3 < 14p for z < 2p.
// one and two are opaque(1u) and opaque(2u): the sums stay multiply-adds of
// 32-bit operands instead of 64-bit adds of values widened by extra moves.
__device__ __forceinline__ u64 wide_external_linear(u32 x, u32 one, u32 two) {
  u32 X = csub(x + dpp<ROR(4)>(x));
  X = csub(X + dpp<ROR(8)>(X));
  const u32 z = x + X;
  const u32 r1 = xdpp<QROT1>(z);
  const u32 r2 = xdpp<QROT2>(z);
  const u32 r3 = xdpp<QROT3>(z);
  return mad_u64(z, two, mad_u64(r1, 3u, mad_u64(r3, one, (u64)r2)));
}

// Sum of the row's 16 lanes for x < p: < p.
__device__ __forceinline__ u32 wide_row_sum(u32 x) {
  u32 s = csub(x + dpp<ROR(8)>(x));
  s = csub(s + dpp<ROR(4)>(s));
  s = csub(s + dpp<ROR(2)>(s));
  return csub(s + dpp<ROR(1)>(s));
}

// Per-lane constants of a 16-lane permutation: external round constants of
// this lane's element and its internal diagonal entry (Montgomery form).
struct WideLane {
  u32 ext_init[4];
  u32 ext_term[4];
  u32 diag;
  u32 one, two;   // opaque(1u), opaque(2u)
};

// diag(V) of the width-16 internal layer in Montgomery form, element j.
__device__ __forceinline__ u32 internal_diag16(int j) {
  // V = [-2, 1, 2, 1/2, 3, 4, -1/2, -3, -4, 1/2^8, 1/8, 1/2^24, -1/2^8, -1/8, -1/16, -1/2^24]
  const u32 one = ONE;
  u32 v;
  switch (j) {
    case 0: v = neg(dbl(one)); break;
    case 1: v = one; break;
    case 2: v = dbl(one); break;
    case 3: v = halve(one); break;
    case 4: v = triple(one); break;
    case 5: v = dbl(dbl(one)); break;
    case 6: v = neg(halve(one)); break;
    case 7: v = neg(triple(one)); break;
    case 8: v = neg(dbl(dbl(one))); break;
    case 9: v = div2k<8>(one); break;
    case 10: v = div2k<3>(one); break;
    case 11: v = div2k<24>(one); break;
    case 12: v = neg(div2k<8>(one)); break;
    case 13: v = neg(div2k<3>(one)); break;
    case 14: v = neg(div2k<4>(one)); break;
    default: v = neg(div2k<24>(one)); break;
  }
  return v;
}

__device__ __forceinline__ WideLane wide_lane(const P2C16& c, int j) {
  WideLane w;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    w.ext_init[r] = c.ext_init[r * 16 + j];
    w.ext_term[r] = c.ext_term[r * 16 + j];
  }
  w.diag = internal_diag16(j);
  // a plain 32-bit register: the switch above otherwise becomes 64-bit
  // constants and the products with it 64-bit multiplies
  asm volatile("" : "+v"(w.diag));
  w.one = opaque(1u);
  w.two = opaque(2u);
  return w;
}

// x: this lane's state element (< p); j = lane & 15.  Returns the lane's
// element of the permuted state (< p).
__device__ __forceinline__ u32 permute16_wide(u32 x, int j, const WideLane& w, const P2C16& c) {
  u64 y = wide_external_linear(x, w.one, w.two);
#pragma unroll
  for (int r = 0; r < 4; r++)   // y + rc < 15p; the S-box output < 1.75p is brought below p
    y = wide_external_linear(csub((u32)cube_lazy(full_reduce(mad_u64(w.ext_init[r], 1u, y)))), w.one, w.two);
  x = full_reduce(y);
#pragma unroll
  for (int r = 0; r < P2C16::RP; r++) {
    const u32 x0 = csub((u32)cube_lazy(add(x, c.internal[r])));
    x = j == 0 ? x0 : x;
    const u32 sum = wide_row_sum(x);
    x = add(sum, csub(mul_lazy(x, w.diag)));   // d x < 2p -> < p
  }
  y = x;
#pragma unroll
  for (int r = 0; r < 4; r++)
    y = wide_external_linear(csub((u32)cube_lazy(full_reduce(mad_u64(w.ext_term[r], 1u, y)))), w.one, w.two);
  return full_reduce(y);
}

}  // namespace kb
