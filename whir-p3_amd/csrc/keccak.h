// Keccak for the Keccak MMCS variant (src/whir/mod.rs:236-383): the
// Keccak-f[1600] permutation on 25 64-bit lanes (lane x + 5y, FIPS 202 §3.1),
// the lane sponge PaddingFreeSponge<KeccakF, 25, 17, 4> the Merkle tree hashes
// with, and (host only) Keccak-256 over bytes, the hash of the byte
// challenger.  On the device every 64-bit operation compiles to a pair of
// 32-bit ones; a rotation is two funnel shifts.
#pragma once
#include <cstddef>
#include <cstdint>

namespace kb {

// Round constants (iota) and rotation offsets (rho), FIPS 202 §3.2.
constexpr uint64_t KECCAK_RC[24] = {
    0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull, 0x8000000080008000ull,
    0x000000000000808Bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
    0x000000000000008Aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
    0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull, 0x8000000000008003ull,
    0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800Aull, 0x800000008000000Aull,
    0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};

__host__ __device__ __forceinline__ uint64_t rotl64(uint64_t x, int r) {
  return r == 0 ? x : (x << r) | (x >> (64 - r));
}

// Keccak-f[1600].  rho and pi are fused: lane (x, y) rotated by its offset
// moves to (y, 2x + 3y); the offsets are written per source lane.
__host__ __device__ inline void keccak_f1600(uint64_t* a) {
  constexpr int ROT[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43, 25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
#pragma unroll 1
  for (int round = 0; round < 24; round++) {
    uint64_t c[5], b[25];
#pragma unroll
    for (int x = 0; x < 5; x++) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
#pragma unroll
    for (int x = 0; x < 5; x++) {
      const uint64_t d = c[(x + 4) % 5] ^ rotl64(c[(x + 1) % 5], 1);
#pragma unroll
      for (int y = 0; y < 25; y += 5) a[x + y] ^= d;
    }
#pragma unroll
    for (int x = 0; x < 5; x++)
#pragma unroll
      for (int y = 0; y < 5; y++) b[y + 5 * ((2 * x + 3 * y) % 5)] = rotl64(a[x + 5 * y], ROT[x + 5 * y]);
#pragma unroll
    for (int y = 0; y < 25; y += 5)
#pragma unroll
      for (int x = 0; x < 5; x++) a[x + y] = b[x + y] ^ (~b[(x + 1) % 5 + y] & b[(x + 2) % 5 + y]);
    a[0] ^= KECCAK_RC[round];
  }
}

// a ^ b ^ c ^ d ^ e as two three-input bit operations (v_bitop3_b32, truth
// table 0x96: odd parity)
__device__ __forceinline__ uint32_t xor5(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t e) {
  return __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_bitop3_b32(a, b, c, 0x96), d, e, 0x96);
}

__device__ __forceinline__ void rotl_halves(uint32_t lo, uint32_t hi, int r, uint32_t& olo, uint32_t& ohi) {
  if (r == 0) {
    olo = lo;
    ohi = hi;
  } else if (r < 32) {
    ohi = __builtin_amdgcn_alignbit(hi, lo, 32 - r);
    olo = __builtin_amdgcn_alignbit(lo, hi, 32 - r);
  } else if (r == 32) {
    olo = hi;
    ohi = lo;
  } else {
    ohi = __builtin_amdgcn_alignbit(lo, hi, 64 - r);
    olo = __builtin_amdgcn_alignbit(hi, lo, 64 - r);
  }
}

// Keccak-f[1600] for the device kernels, on the state held as 32-bit halves
// (lane i = lo[i] | hi[i] << 32).  A 64-bit rotation is two funnel shifts
// (v_alignbit_b32), theta's five-way parities are two three-input bit
// operations per half (xor5), and chi's b ^ (~c & d) is one more.  Rotation
// amounts are compile-time constants: every loop but the round loop unrolls.
// UNROLL = 24 inlines all rounds (about 5.2k instructions, 15% fewer than
// rolled, where the loop-carried state needs register moves after pi): for
// the grinding loop, whose one call site runs for a whole search; the rolled
// form (1) keeps a call site near 1 KB of code.
template <int UNROLL = 1>
__device__ __forceinline__ void keccak_f1600_halves(uint32_t* lo, uint32_t* hi) {
  constexpr int ROT[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43, 25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
#pragma unroll UNROLL
  for (int round = 0; round < 24; round++) {
    uint32_t cl[5], ch[5];
#pragma unroll
    for (int x = 0; x < 5; x++) {
      cl[x] = xor5(lo[x], lo[x + 5], lo[x + 10], lo[x + 15], lo[x + 20]);
      ch[x] = xor5(hi[x], hi[x + 5], hi[x + 10], hi[x + 15], hi[x + 20]);
    }
#pragma unroll
    for (int x = 0; x < 5; x++) {
      uint32_t rl, rh;
      rotl_halves(cl[(x + 1) % 5], ch[(x + 1) % 5], 1, rl, rh);
      const uint32_t dl = cl[(x + 4) % 5] ^ rl, dh = ch[(x + 4) % 5] ^ rh;
#pragma unroll
      for (int y = 0; y < 25; y += 5) {
        lo[x + y] ^= dl;
        hi[x + y] ^= dh;
      }
    }
    uint32_t bl[25], bh[25];
#pragma unroll
    for (int x = 0; x < 5; x++)
#pragma unroll
      for (int y = 0; y < 5; y++) {
        const int to = y + 5 * ((2 * x + 3 * y) % 5);
        rotl_halves(lo[x + 5 * y], hi[x + 5 * y], ROT[x + 5 * y], bl[to], bh[to]);
      }
#pragma unroll
    for (int y = 0; y < 25; y += 5)
#pragma unroll
      for (int x = 0; x < 5; x++) {
        const int b = x + y, c = (x + 1) % 5 + y, d = (x + 2) % 5 + y;
        // where c has a 1 the result is b, elsewhere b ^ d
        lo[b] = (bl[c] & bl[b]) | (~bl[c] & (bl[b] ^ bl[d]));
        hi[b] = (bh[c] & bh[b]) | (~bh[c] & (bh[b] ^ bh[d]));
      }
    lo[0] ^= (uint32_t)KECCAK_RC[round];
    hi[0] ^= (uint32_t)(KECCAK_RC[round] >> 32);
  }
}

// PaddingFreeSponge<KeccakF, 25, 17, 4> over a run of lanes: the state starts
// at zero, each chunk of up to 17 lanes overwrites the state's first lanes
// and is followed by a permutation (a final partial chunk included), and the
// digest is lanes 0..3.
__host__ __device__ inline void keccak_lane_sponge(const uint64_t* in, size_t n, uint64_t* digest4) {
  uint64_t s[25];
#pragma unroll
  for (int i = 0; i < 25; i++) s[i] = 0;
  for (size_t at = 0; at < n; at += 17) {
    const size_t k = n - at < 17 ? n - at : 17;
    for (size_t i = 0; i < k; i++) s[i] = in[at + i];
    keccak_f1600(s);
  }
#pragma unroll
  for (int i = 0; i < 4; i++) digest4[i] = s[i];
}

// Keccak-256 (the original Keccak padding: 0x01 after the message, 0x80 in
// the last byte of the 136-byte block), as Keccak256Hash.
inline void keccak256(const uint8_t* msg, size_t len, uint8_t out[32]) {
  uint64_t s[25] = {};
  constexpr size_t RATE = 136;
  auto absorb = [&](const uint8_t* blk) {
    for (size_t i = 0; i < RATE / 8; i++) {
      uint64_t w = 0;
      for (int b = 0; b < 8; b++) w |= (uint64_t)blk[8 * i + b] << (8 * b);
      s[i] ^= w;
    }
    keccak_f1600(s);
  };
  size_t at = 0;
  for (; len - at >= RATE; at += RATE) absorb(msg + at);
  uint8_t last[RATE] = {};
  for (size_t i = 0; at + i < len; i++) last[i] = msg[at + i];
  last[len - at] ^= 0x01;
  last[RATE - 1] ^= 0x80;
  absorb(last);
  for (int i = 0; i < 32; i++) out[i] = (uint8_t)(s[i / 8] >> (8 * (i % 8)));
}

}  // namespace kb
