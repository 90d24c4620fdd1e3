// Lane-parallel Poseidon2-16: one permutation spread over a 16-lane DPP row,
// lane j holding state element j.  It trades ~2.4x more lane-instructions for
// ~5x lower latency than the one-lane form in poseidon2.h, which is what the
// latency-bound parts need (upper Merkle levels, a handful of candidates).
//
// Same permutation as poseidon2.h: the external layer applies the circulant
// M4 = (2 3 1 1) to each quad of lanes (quad rotations bring x_{q+1..q+3}),
// then adds the sum of the four quads' same-position elements (row rotations
// by 4, 8, 12); the internal layer adds the 16-lane sum (rotations by 8, 4, 2,
// 1) to diag_j * s_j.  Every lane of the wave must be active.
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

__device__ __forceinline__ u32 wide_external_linear(u32 x) {
  u32 r1 = dpp<QROT1>(x);
  u32 r2 = dpp<QROT2>(x);
  u32 r3 = dpp<QROT3>(x);
  // 2 x + 3 r1 + r2 + r3 = 2 (x + r1) + r1 + (r2 + r3)
  u32 a = add(x, r1);
  u32 o = add(add(dbl(a), r1), add(r2, r3));
  u32 s = add(o, dpp<ROR(4)>(o));
  s = add(s, dpp<ROR(8)>(s));
  return add(o, s);
}

__device__ __forceinline__ u32 wide_row_sum(u32 x) {
  u32 s = add(x, dpp<ROR(8)>(x));
  s = add(s, dpp<ROR(4)>(s));
  s = add(s, dpp<ROR(2)>(s));
  return add(s, dpp<ROR(1)>(s));
}

// Per-lane constants of a 16-lane permutation: external round constants of
// this lane's element and its internal diagonal entry (Montgomery form).
struct WideLane {
  u32 ext_init[4];
  u32 ext_term[4];
  u32 diag;
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
  return w;
}

// x: this lane's state element; j = lane & 15.
__device__ __forceinline__ u32 permute16_wide(u32 x, int j, const WideLane& w, const P2C16& c) {
  x = wide_external_linear(x);
#pragma unroll
  for (int r = 0; r < 4; r++) x = wide_external_linear(cube(add(x, w.ext_init[r])));
#pragma unroll
  for (int r = 0; r < P2C16::RP; r++) {
    u32 y = cube(add(x, c.internal[r]));
    x = j == 0 ? y : x;
    x = add(wide_row_sum(x), mul(w.diag, x));
  }
#pragma unroll
  for (int r = 0; r < 4; r++) x = wide_external_linear(cube(add(x, w.ext_term[r])));
  return x;
}

}  // namespace kb
