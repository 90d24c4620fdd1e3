// Keccak-f[1600] on a pair of adjacent lanes: the even lane holds the low
// 32 bits of all 25 state lanes, the odd lane the high 32 bits.  A 64-bit
// rotation then needs the partner's word of the same lane, which one DPP
// swap (quad_perm [1, 0, 3, 2]) brings, followed by one funnel shift that is
// the same instruction in both lanes:
//   rotate by r < 32:       x' = alignbit(x, partner(x), 32 - r)
//   rotate by r = 32 + r':  x' = alignbit(partner(x), x, 32 - r')
// (for the even lane x = lo gives lo' = lo << r | hi >> (32 - r); for the odd
// lane x = hi gives hi' = hi << r | lo >> (32 - r); past 32 the halves trade
// places first).  Theta's parities, chi and iota are local, so a round is
// about 124 instructions per lane against about 217 for the whole state in
// one lane: half the latency of a permutation, for the latency-bound parts
// (upper Merkle levels), at some 15% more lane-instructions per permutation.
// Both lanes of every pair must be active.
#pragma once
#include "keccak.h"

namespace kb {

__device__ __forceinline__ uint32_t keccak_partner(uint32_t x) {
  constexpr int SWAP = 1 | (0 << 2) | (3 << 4) | (2 << 6);
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, SWAP, 0xF, 0xF, true);
}

__device__ __forceinline__ uint32_t keccak_pair_rot(uint32_t x, int r) {
  if (r == 0) return x;
  const uint32_t y = keccak_partner(x);
  if (r == 32) return y;
  return r < 32 ? __builtin_amdgcn_alignbit(x, y, 32 - r) : __builtin_amdgcn_alignbit(y, x, 64 - r);
}

// s: this lane's half of the 25 state lanes (lane x + 5y); odd: this lane
// holds the high halves.
__device__ __forceinline__ void keccak_f1600_pair(uint32_t* s, bool odd) {
  constexpr int ROT[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43, 25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
#pragma unroll   // all 24 rounds: no register moves for pi (about 3k instructions)
  for (int round = 0; round < 24; round++) {
    uint32_t c[5];
#pragma unroll
    for (int x = 0; x < 5; x++) c[x] = xor5(s[x], s[x + 5], s[x + 10], s[x + 15], s[x + 20]);
#pragma unroll
    for (int x = 0; x < 5; x++) {
      const uint32_t d = c[(x + 4) % 5] ^ keccak_pair_rot(c[(x + 1) % 5], 1);
#pragma unroll
      for (int y = 0; y < 25; y += 5) s[x + y] ^= d;
    }
    uint32_t b[25];
#pragma unroll
    for (int x = 0; x < 5; x++)
#pragma unroll
      for (int y = 0; y < 5; y++) b[y + 5 * ((2 * x + 3 * y) % 5)] = keccak_pair_rot(s[x + 5 * y], ROT[x + 5 * y]);
#pragma unroll
    for (int y = 0; y < 25; y += 5)
#pragma unroll
      for (int x = 0; x < 5; x++) {
        const uint32_t bb = b[x + y], cc = b[(x + 1) % 5 + y], dd = b[(x + 2) % 5 + y];
        s[x + y] = bb ^ (~cc & dd);
      }
    const uint64_t rc = KECCAK_RC[round];
    s[0] ^= odd ? (uint32_t)(rc >> 32) : (uint32_t)rc;
  }
}

}  // namespace kb
