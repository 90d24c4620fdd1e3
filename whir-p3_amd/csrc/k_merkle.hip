// Poseidon2 Merkle commitment of a row-major matrix (MerkleTreeMmcs::commit_matrix,
// called at src/whir/committer/writer.rs:88-89 and, through ExtensionMmcs with
// each extension element flattened to its 4 coefficients, at
// src/whir/prover/mod.rs:182-184), and batched openings (open_batch at
// src/whir/prover/mod.rs:251-289, 380-400).
//
// Leaf digest: PaddingFreeSponge<Poseidon2-LW, LW, LW-8, 8> over the row
// (overwrite the rate part with each chunk, permute, output state[0..8]).
// Node digest: TruncatedPermutation<Poseidon2-16, 2, 8, 16> (left || right,
// permute, keep state[0..8]).
//
// The digests of all levels live in one buffer: level 0 (the leaves) first,
// then each parent level, ending with the root; level j starts at digest
// 2H - 2H/2^j.  Leaves are hashed one row per lane and each level above is
// one launch with one compression per lane.
#include "kcommon.h"
#include "kernels.h"
#include "transcript.h"
#include "poseidon2.h"
#include "poseidon2_wide.h"
#include "keccak.h"
#include "keccak_pair.h"

namespace kb {

// The constants come from device memory and are re-read for every absorbed
// chunk (the empty asm makes their offset opaque to the loop), so they are not
// held in, and spilled from, scalar registers across the loop.
template <int LW>
__global__ void __launch_bounds__(256) k_leaf_hash(Group<TreeArgs> g, int row_len, u64 nrows,
                                                   const P2Consts<LW>* __restrict__ cp) {
  constexpr int RATE = LW - 8;
  const u32* __restrict__ rows = g.v[blockIdx.z].rows;
  u32* __restrict__ dig = g.v[blockIdx.z].tree;
  u64 i = (u64)blockIdx.x * 256 + threadIdx.x;
  if (i >= nrows) return;
  u32 s[LW];
#pragma unroll
  for (int j = 0; j < LW; j++) s[j] = 0;
  const u32* r = rows + i * (u64)row_len;
  if ((row_len % RATE) == 0) {
    for (int pos = 0; pos < row_len; pos += RATE) {
      const uint4* v = reinterpret_cast<const uint4*>(r + pos);
#pragma unroll
      for (int q = 0; q < RATE / 4; q++) {
        uint4 x = v[q];
        s[4 * q] = x.x;
        s[4 * q + 1] = x.y;
        s[4 * q + 2] = x.z;
        s[4 * q + 3] = x.w;
      }
      u32 off = 0;
      asm volatile("" : "+s"(off));
      permute<LW>(s, cp[off]);
    }
  } else {
    for (int pos = 0; pos < row_len; pos += RATE) {
      int cnt = row_len - pos < RATE ? row_len - pos : RATE;
#pragma unroll
      for (int j = 0; j < RATE; j++)
        if (j < cnt) s[j] = r[pos + j];
      u32 off = 0;
      asm volatile("" : "+s"(off));
      permute<LW>(s, cp[off]);
    }
  }
  uint4* o = reinterpret_cast<uint4*>(dig + i * 8);
  o[0] = make_uint4(s[0], s[1], s[2], s[3]);
  o[1] = make_uint4(s[4], s[5], s[6], s[7]);
}

__device__ __forceinline__ void compress_pair(const u32* l, const u32* r, u32* out, const P2C16& c) {
  u32 s[16];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    s[k] = l[k];
    s[8 + k] = r[k];
  }
  permute<16>(s, c);
#pragma unroll
  for (int k = 0; k < 8; k++) out[k] = s[k];
}

// One tree level: out[j] = compress(in[2j], in[2j+1]), where the level of
// 2 n_out digests starts at digest `off` of the tree and the next follows it.
__global__ void __launch_bounds__(256) k_tree_level(Group<TreeArgs> g, u64 off, u64 n_out, P2C16 c) {
  u32* tree = g.v[blockIdx.z].tree;
  const u32* __restrict__ in = tree + off * 8;
  u32* __restrict__ out = tree + (off + 2 * n_out) * 8;
  u64 j = (u64)blockIdx.x * 256 + threadIdx.x;
  if (j >= n_out) return;
  const uint4* v = reinterpret_cast<const uint4*>(in + 16 * j);
  u32 s[16];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    uint4 x = v[q];
    s[4 * q] = x.x;
    s[4 * q + 1] = x.y;
    s[4 * q + 2] = x.z;
    s[4 * q + 3] = x.w;
  }
  permute<16>(s, c);
  uint4* o = reinterpret_cast<uint4*>(out + 8 * j);
  o[0] = make_uint4(s[0], s[1], s[2], s[3]);
  o[1] = make_uint4(s[4], s[5], s[6], s[7]);
}

// log2(S) small levels in one launch, with a 16-lane permutation per node
// (latency rather than throughput decides there): block b (8 S lanes) stages
// the S digests [S b, S b + S) of the level at `off` (n digests, a multiple of
// S) in LDS and reduces them to one digest log2(S) levels up, writing every
// level in between; the levels are laid out one after another (the level of
// size n at off, the next of size n / 2 at off + n, ...).  S / 2 nodes per
// step.  S = 32 keeps a block at one wave per SIMD, so a launch of few blocks
// runs its levels at the permutation's latency (with S = 128 each block holds
// four waves per SIMD for its first level).
template <int S>
__global__ void __launch_bounds__(8 * S) k_tree_sub_wide(Group<TreeArgs> g, u64 off, u64 n, P2C16 c) {
  __shared__ u32 sh[2][S * 8];
  u32* __restrict__ tree = g.v[blockIdx.z].tree;
  const int t = threadIdx.x;
  const int j = t & 15;
  const WideLane w = wide_lane(c, j);
  const u64 b = blockIdx.x;
  for (int e = t; e < S * 8; e += 8 * S) sh[0][e] = tree[(off + b * S) * 8 + e];
  __syncthreads();
  int cur = 0;
  for (int m = S; m > 1; m >>= 1) {
    const int m_out = m >> 1;
    const int p = t >> 4;
    if (4 * (t >> 6) < m_out) {   // wave-uniform; below 4 nodes part of the wave idles
      u32 x = p < m_out ? sh[cur][16 * p + j] : 0u;
      x = permute16_wide(x, j, w, c);
      if (p < m_out && j < 8) {
        sh[cur ^ 1][8 * p + j] = x;
        tree[(off + n + b * m_out + p) * 8 + j] = x;
      }
    }
    __syncthreads();
    off += n;
    n >>= 1;
    cur ^= 1;
  }
}

// The top levels of a tree in one launch: n (<= 32) digests at `off` are
// staged in LDS and reduced to the root, every level written out.  One block
// of 256 lanes, 16 nodes per step.
// Waves whose four nodes are all past the level's end skip the permutation
// (the lane-parallel permutation only needs its own wave's lanes).  The root
// is also written to `root_out` when given (host-visible memory).
constexpr int TREE_TOP = 32;
__global__ void __launch_bounds__(8 * TREE_TOP) k_tree_top_wide(Group<TreeArgs> g, u64 off, int n, P2C16 c) {
  __shared__ u32 sh[2][TREE_TOP * 8];
  u32* __restrict__ tree = g.v[blockIdx.z].tree;
  u32* root_out = g.v[blockIdx.z].root_out;
  const int t = threadIdx.x;
  const int j = t & 15;
  const WideLane w = wide_lane(c, j);
  for (int e = t; e < n * 8; e += 8 * TREE_TOP) sh[0][e] = tree[off * 8 + e];
  __syncthreads();
  int cur = 0;
  while (n > 1) {
    const int n_out = n >> 1;
    const int p = t >> 4;
    if (4 * (t >> 6) < n_out) {   // wave-uniform
      u32 x = p < n_out ? sh[cur][16 * p + j] : 0u;
      x = permute16_wide(x, j, w, c);
      if (p < n_out && j < 8) {
        sh[cur ^ 1][8 * p + j] = x;
        tree[(off + n + p) * 8 + j] = x;
      }
    }
    __syncthreads();
    off += n;
    n = n_out;
    cur ^= 1;
  }
  if (root_out && t < 8) root_out[t] = sh[cur][t];
}

// Opening of the rows at idx: out[q] = row (row_len) followed by depth
// sibling digests.  Grid: (ceil(per / 256), max_nq, proofs); block (x, q)
// copies entries x*256.. of query q's opening (row, then the sibling digest of
// each level); queries at or past *nq (when given) are skipped.  `idx` and
// `out` may be host-visible memory.
__global__ void k_open(Group<OpenArgs> g, int row_len, u64 height, int depth) {
  const OpenArgs& a = g.v[blockIdx.z];
  if (a.nq && blockIdx.y >= *a.nq) return;
  const u32* __restrict__ rows = a.rows;
  const u32* __restrict__ tree = a.tree;
  const u64* __restrict__ idx = a.idx;
  u32* __restrict__ out = a.out;
  const u64 per = (u64)row_len + 8ull * depth;
  const u64 q = blockIdx.y;
  const u64 o = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= per) return;
  const u64 ix = idx[q];
  u32 v;
  if (o < (u64)row_len) {
    v = rows[ix * row_len + o];
  } else {
    u64 l = (o - row_len) / 8, k = (o - row_len) % 8;
    u64 off = 2 * height - ((2 * height) >> l);
    u64 node = (ix >> l) ^ 1;
    v = tree[(off + node) * 8 + k];
  }
  out[q * per + o] = v;
}

template <int LW>
int merkle_leaves_impl(const TreeArgs* a, int count, int row_len, u64 height, const P2Consts<LW>* c, hipStream_t st) {
  const unsigned blocks = (unsigned)((height + 255) / 256);
  return launch_group(a, count, [&](const Group<TreeArgs>& g) {
    hipLaunchKernelGGL((k_leaf_hash<LW>), dim3(blocks, 1, count), dim3(256), 0, st, g, row_len, height, c);
  });
}

int merkle_leaves(const TreeArgs* a, int count, int row_len, u64 height, int leaf_width, const P2C16* c16,
                  const P2C24* c24, hipStream_t st) {
  if (height == 0 || (height & (height - 1))) return -1;
  return leaf_width == 24 ? merkle_leaves_impl<24>(a, count, row_len, height, c24, st)
                          : merkle_leaves_impl<16>(a, count, row_len, height, c16, st);
}

// Levels with more than 2^tree_lane_log nodes (default 2^14) use one lane per
// node (throughput); below that, seven or five levels per launch with 16
// lanes per node (latency), and the last (at most 5) levels in one block.
constexpr int TREE_SUB = 32;
static_assert(TREE_SUB <= TREE_TOP, "the top block takes what the subtree launches leave");
int merkle_tree(const TreeArgs* a, int count, u64 height, const P2C16& c16, const Tuning& t, hipStream_t st) {
  if (height == 0 || (height & (height - 1))) return -1;
  if (count < 1 || count > GROUP_MAX) return -8;
  Group<TreeArgs> g;
  for (int i = 0; i < count; i++) g.v[i] = a[i];
  const unsigned np = (unsigned)count;
  u64 off = 0, n = height;
  while (n > TREE_SUB) {
    const u64 n_out = n >> 1;
    if (n_out > (1ull << t.tree_lane_log)) {
      unsigned blocks = (unsigned)((n_out + 255) / 256);
      hipLaunchKernelGGL(k_tree_level, dim3(blocks, 1, np), dim3(256), 0, st, g, off, n_out, c16);
      if (hipGetLastError() != hipSuccess) return -5;
      off += n;
      n = n_out;
      continue;
    }
    // n is a power of two above S: n / S subtrees, levels n / 2 .. n / S
    // written.  From 2^15 nodes (256 blocks or more; tree_s128_log) 128-digest
    // subtrees: the first level's work then fills the chip and seven levels
    // per launch cost less than five (22 against 23 us at 2^15); below,
    // 32-digest subtrees at one wave per SIMD.
    const u64 S = n >= (1ull << t.tree_s128_log) ? 128 : TREE_SUB;
    if (S == 128)
      hipLaunchKernelGGL(k_tree_sub_wide<128>, dim3((unsigned)(n / 128), 1, np), dim3(8 * 128), 0, st, g, off, n, c16);
    else
      hipLaunchKernelGGL(k_tree_sub_wide<TREE_SUB>, dim3((unsigned)(n / TREE_SUB), 1, np), dim3(8 * TREE_SUB), 0, st,
                         g, off, n, c16);
    if (hipGetLastError() != hipSuccess) return -5;
    off += 2 * n - n / (S / 2);
    n /= S;
  }
  if (n > 1) {
    hipLaunchKernelGGL(k_tree_top_wide, dim3(1, 1, np), dim3(8 * TREE_TOP), 0, st, g, off, (int)n, c16);
    if (hipGetLastError() != hipSuccess) return -5;
    return 0;
  }
  return 1;   // single leaf: the root was not written
}

int merkle_open(const OpenArgs* a, int count, int row_len, u64 height, int max_nq, hipStream_t st) {
  int depth = 0;
  while ((1ull << depth) < height) depth++;
  const u64 per = (u64)row_len + 8ull * depth;
  if (max_nq <= 0) return 0;
  return launch_group(a, count, [&](const Group<OpenArgs>& g) {
    hipLaunchKernelGGL(k_open, dim3((unsigned)((per + 255) / 256), (unsigned)max_nq, count), dim3(256), 0, st, g,
                       row_len, height, depth);
  });
}

// ------------------------------------------------------------------ Keccak MMCS
// MerkleTreeMmcs<F, u64, SerializingHasher<U64Hash>, CompressionFunctionFromHasher<U64Hash, 2, 4>, 4> with
// U64Hash = PaddingFreeSponge<KeccakF, 25, 17, 4> (src/whir/mod.rs:254-259).  A row's field elements enter
// the sponge as u64 lanes two at a time, element 2i in the low half and 2i + 1 in the high half, each as
// its Montgomery word: the row's memory read as little-endian u64 (an odd last element alone).  A node is
// the sponge over its children's 4 + 4 lanes.  A digest's 4 lanes are 8 u32 words in the tree buffer,
// low half first, so the tree layout and openings are those of the Poseidon2 tree.  One lane per leaf or
// node: a Keccak-f[1600] on 64-bit lanes is a long chain of 32-bit pair operations with ample ILP.
__global__ void __launch_bounds__(256) k_leaf_keccak(Group<TreeArgs> g, int row_len, u64 height) {
  const TreeArgs& a = g.v[blockIdx.z];
  const u64 i = (u64)blockIdx.x * 256 + threadIdx.x;
  if (i >= height) return;
  const u32* row = a.rows + i * (u64)row_len;
  // lane j of a chunk = words 2j (low half) and 2j + 1 (high half) of the row
  u32 lo[25], hi[25];
#pragma unroll
  for (int j = 0; j < 25; j++) lo[j] = hi[j] = 0;
  for (int at = 0; 2 * at < row_len; at += 17) {
    // unrolled with a guard, so the state stays in registers (a loop bound
    // known only at run time would index it in scratch memory)
#pragma unroll
    for (int j = 0; j < 17; j++) {
      const int e = 2 * (at + j);
      if (e < row_len) {
        lo[j] = row[e];
        hi[j] = e + 1 < row_len ? row[e + 1] : 0u;
      }
    }
    keccak_f1600_halves(lo, hi);
  }
  u32* out = a.tree + i * 8;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    out[2 * j] = lo[j];
    out[2 * j + 1] = hi[j];
  }
}

// CompressionFunctionFromHasher<U64Hash, 2, 4> over tree levels: node =
// the lane sponge of left || right (8 lanes, one permutation), computed by a
// lane pair (keccak_pair.h; lane 2p + h handles node p's half h, h = 0: the
// low words of the digest lanes).  levels + 1 levels per launch: block b
// computes nodes [b * S, b * S + S) of the first level (S = 2^levels <= 128)
// from the digests at `off`, then halves them through LDS up to one node,
// writing every level to the tree (openings read them).  Level widths:
// n_out, n_out / 2, ...; the last level written has n_out / S nodes.  The
// root, when reached, also goes to root_out.  The pairs halve the latency of
// each level against one lane per node, which the upper levels are bound
// by; measured, they were also no slower on the widest levels.
__global__ void __launch_bounds__(256) k_tree_keccak(Group<TreeArgs> g, u64 off, u64 n_out, int levels) {
  const TreeArgs& a = g.v[blockIdx.z];
  u32* tree = a.tree;
  u32* root_out = a.root_out;
  __shared__ u32 sd[128 * 8];
  const int S = 1 << levels;
  const int t = threadIdx.x, p = t >> 1, h = t & 1;
  u64 in_off = off, out_off = off + 2 * n_out, width = n_out;
  u64 base = (u64)blockIdx.x * S;
  for (int k = S; k >= 1; k >>= 1) {
    u32 d[4];
    if (p < k) {   // pairs are active together
      const u32* in = k == S ? tree + (in_off + 2 * (base + p)) * 8 : &sd[(2 * p) * 8];
      u32 st[25];
#pragma unroll
      for (int j = 0; j < 25; j++) st[j] = 0;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        st[j] = in[2 * j + h];
        st[4 + j] = in[8 + 2 * j + h];
      }
      keccak_f1600_pair(st, h);
#pragma unroll
      for (int j = 0; j < 4; j++) d[j] = st[j];
      u32* out = tree + (out_off + base + p) * 8;
#pragma unroll
      for (int j = 0; j < 4; j++) out[2 * j + h] = d[j];
      if (width == 1 && root_out)
#pragma unroll
        for (int j = 0; j < 4; j++) root_out[2 * j + h] = d[j];
    }
    __syncthreads();   // every read of sd by this level is done
    if (p < k)
#pragma unroll
      for (int j = 0; j < 4; j++) sd[p * 8 + 2 * j + h] = d[j];
    __syncthreads();
    in_off = out_off;
    out_off += width;
    width >>= 1;
    base >>= 1;
  }
}

int keccak_leaves(const TreeArgs* a, int count, int row_len, u64 height, hipStream_t st) {
  if (height == 0 || (height & (height - 1)) || row_len < 1) return -1;
  const unsigned blocks = (unsigned)((height + 255) / 256);
  return launch_group(a, count, [&](const Group<TreeArgs>& g) {
    hipLaunchKernelGGL(k_leaf_keccak, dim3(blocks, 1, count), dim3(256), 0, st, g, row_len, height);
  });
}

int keccak_tree(const TreeArgs* a, int count, u64 height, hipStream_t st) {
  if (height == 0 || (height & (height - 1))) return -1;
  if (count < 1 || count > GROUP_MAX) return -8;
  if (height == 1) return 1;   // single leaf: the root was not written to root_out
  Group<TreeArgs> g;
  for (int i = 0; i < count; i++) g.v[i] = a[i];
  u64 off = 0;
  for (u64 n = height; n > 1;) {
    const u64 n_out = n >> 1;
    int levels = 0;   // S = 2^levels first-level nodes per block, at most 128
    while (levels < 7 && (2ull << levels) <= n_out) levels++;
    hipLaunchKernelGGL(k_tree_keccak, dim3((unsigned)(n_out >> levels), 1, (unsigned)count), dim3(256), 0, st, g,
                       off, n_out, levels);
    if (hipGetLastError() != hipSuccess) return -5;
    // levels + 1 levels written: widths n_out .. n_out >> levels
    u64 w = n;
    for (int l = 0; l <= levels; l++) {
      off += w;
      w >>= 1;
    }
    n = w;
  }
  return 0;
}

// ------------------------------------------------------------------ grinding
// GrindingChallenger::grind for DuplexChallenger (used at src/whir/proof.rs:187-189,
// src/whir/prover/mod.rs:223-225, 362-364): the smallest w in [0, p) such that,
// after observing w, sample_bits(bits) == 0.  Observing w with L < 8 buffered
// inputs overwrites state[0..L] with the buffer and state[L] with w, and the
// next sample is state[7] of the permuted state, so each candidate costs one
// permutation.
//
// A launch scans a range of candidates with one wave per SIMD: each lane
// takes one candidate per iteration and the waves step through the range
// together, so a witness near the start ends the search after about one
// pass over 64K candidates instead of after every resident wave of a
// full-occupancy grid.  A wave stops when the best witness known is not above
// its first candidate in the current iteration; it checks before each
// iteration and between the three phases of the permutation, with each check's
// load issued one phase ahead so its latency hides behind arithmetic.  A stop
// cannot hide a smaller witness: the result only decreases, so every candidate
// below the final result is checked, whatever order waves run in, and the
// result is the smallest witness in the range.
__device__ __forceinline__ u32 best_known(const u32* result) {
  return __hip_atomic_load(result, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// GrindingChallenger::grind for SerializingChallenger32<F, HashChallenger<u8,
// Keccak256Hash, 32>> (p3-challenger; the Keccak variant, src/whir/mod.rs:257),
// smallest witness: check_witness observes w (4 bytes appended to the input,
// the output cleared) and sample_bits pops 4 bytes from the end of
// Keccak-256(input), i.e. output bytes 31, 30, 29, 28 as a little-endian u32
// = the byte-reversed high word of digest lane 3.  Each candidate costs one
// permutation (two when the padding spills into another block).  Candidates
// advance in passes over the grid; a wave stops once the best witness known
// lies at or before its first candidate of the pass, as in k_grind.
__global__ void __launch_bounds__(256) k_grind_keccak(KeccakGrindArgs a, u32 mask, u32 start, u32 end) {
  // The constant lanes and, per lane, masks that place the witness go to LDS
  // once per block: held in scalar registers across the loop (as values or
  // as lane-compare masks) they would spill.
  __shared__ uint64_t s0[25], t1[17], m0lo[25], m0hi[25], m1hi[17];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < 25; j++) {
      s0[j] = a.s0[j];
      m0lo[j] = j == a.lo_lane ? ~0ull : 0;
      m0hi[j] = a.hi_blk == 0 && j == a.hi_lane ? ~0ull : 0;
    }
#pragma unroll
    for (int j = 0; j < 17; j++) {
      t1[j] = a.t1[j];
      m1hi[j] = a.hi_blk == 1 && j == a.hi_lane ? ~0ull : 0;
    }
  }
  u32* result = a.result;
  const int nblk = a.nblk;
  const int shift = a.shift;
  __syncthreads();
  const u32 stride = gridDim.x * 256;
  const u32 lane = blockIdx.x * 256 + threadIdx.x;
  const u32 wave0 = lane & ~63u;
  for (u64 g = start; g + wave0 < end; g += stride) {
    const u64 first = g + wave0;
    if ((u64)__builtin_amdgcn_readfirstlane(best_known(result)) <= first) break;
    const u64 w = g + lane;
    const uint64_t wlo = (uint64_t)(u32)w << shift;
    const uint64_t whi = shift > 32 ? (uint64_t)(u32)w >> (64 - shift) : 0;
    // an offset the compiler cannot see through: the LDS reads stay in the
    // loop instead of being hoisted into registers across it
    int o = 0;
    asm volatile("" : "+v"(o));
    u32 lo[25], hi[25];
#pragma unroll
    for (int j = 0; j < 25; j++) {
      const uint64_t x = s0[o + j] ^ (wlo & m0lo[o + j]) ^ (whi & m0hi[o + j]);
      lo[j] = (u32)x;
      hi[j] = (u32)(x >> 32);
    }
#pragma unroll 1
    for (int b = 0; b < nblk; b++) {   // one call site: the permutation's code appears once
      if (b == 1) {
#pragma unroll
        for (int j = 0; j < 17; j++) {
          const uint64_t x = t1[o + j] ^ (whi & m1hi[o + j]);
          lo[j] ^= (u32)x;
          hi[j] ^= (u32)(x >> 32);
        }
      }
      keccak_f1600_halves<24>(lo, hi);
    }
    const u32 v = __builtin_bswap32(hi[3]);
    if (w < end && (v & mask) == 0) atomicMin(result, (u32)w);
  }
}

int keccak_grind(const KeccakGrindArgs& a, int bits, int blocks, hipStream_t st) {
  if (bits < 1 || bits > 30 || a.nblk < 1 || a.nblk > 2 || a.lo_blk != 0 || a.lo_lane < 0 || a.lo_lane > 16 ||
      a.hi_blk < 0 || a.hi_blk >= a.nblk || a.hi_lane < 0 || a.hi_lane > 16 || a.shift < 0 || a.shift > 56 ||
      a.shift % 8)
    return -1;
  const u32 mask = (1u << bits) - 1;
  hipLaunchKernelGGL(k_grind_keccak, dim3((unsigned)std::max(blocks, 1)), dim3(256), 0, st, a, mask, 0u, P);
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

// Orders a phase boundary: the state registers pass through an empty asm
// (so the check after it cannot be moved ahead of the phase's arithmetic) and
// the memory clobber keeps the next phase's constant loads behind it.
__device__ __forceinline__ void phase_fence(u32* s) {
  asm volatile(""
               : "+v"(s[0]), "+v"(s[1]), "+v"(s[2]), "+v"(s[3]), "+v"(s[4]), "+v"(s[5]), "+v"(s[6]), "+v"(s[7]),
                 "+v"(s[8]), "+v"(s[9]), "+v"(s[10]), "+v"(s[11]), "+v"(s[12]), "+v"(s[13]), "+v"(s[14]), "+v"(s[15])
               :
               : "memory");
}

// With tr_mode == TR_COEFFS_GRIND every block first runs the transcript step
// that observes the round's (h(0), h(2)); its state is the candidates' base.
// With TR_GBASE that state was written to tr.gbase by an earlier kernel.
// Grid: (blocks, 1, proofs), one search per proof.
__global__ void __launch_bounds__(256) k_grind(const P2C16* __restrict__ cp, Group<GrindArgs> ga, int tr_mode,
                                               u32 mask, u32 start, u32 end) {
  // This proof's arguments are copied out of the group before any asm: an
  // asm with a memory clobber would otherwise make the compiler keep a
  // private copy of the whole group (kilobytes of scratch per lane).
  const GrindArgs& a = ga.v[blockIdx.z];
  Sponge16 base = a.base;
  int wpos = a.wpos;
  u32* result = a.result;
  const ScTr tr = a.tr;
  if (tr_mode == TR_COEFFS_GRIND || tr_mode == TR_OBS_GRIND) {
    __shared__ GrindBase gbs;
    tr_prologue(tr_mode, tr, *cp, &gbs, nullptr);
#pragma unroll
    for (int k = 0; k < 16; k++) base.s[k] = gbs.s[k];
    wpos = (int)gbs.wpos;
  } else if (tr_mode == TR_GBASE) {
#pragma unroll
    for (int k = 0; k < 16; k++) base.s[k] = tr.gbase->s[k];
    wpos = (int)tr.gbase->wpos;
  }
  GrindPre gp;
  grind_pre(gp, base.s, wpos);
  const u32 stride = gridDim.x * 256;
  const u32 lane = blockIdx.x * 256 + threadIdx.x;
  const u32 wave0 = lane & ~63u;
  u32 b0 = best_known(result);
  for (u64 g = start; g + wave0 < end; g += stride) {
    const u64 first = g + wave0;
    if ((u64)b0 <= first) break;
    // The empty asm makes the constants' offset opaque to the loop, so they
    // are reloaded per iteration instead of being held in (and spilled from)
    // scalar registers across it.
    u32 off = 0;
    asm volatile("" : "+s"(off));
    const P2C16& c = cp[off];
    const u64 w = g + lane;
    u32 s[16];
    u64 y[16];
    grind_layer(gp, to_monty((u32)w), y);
    const u32 b1 = best_known(result);
    permute16_head_rounds(y, s, c);
    phase_fence(s);
    if ((u64)b1 <= first) break;
    const u32 b2 = best_known(result);
    permute_internal<16>(s, c);
    phase_fence(s);
    if ((u64)b2 <= first) break;
    b0 = best_known(result);
    const u32 s7 = permute16_tail_s7(s, c);
    if (w < end && (from_monty(s7) & mask) == 0) atomicMin(result, (u32)w);
  }
}

// Short searches (at most Tuning::grind_wide_bits bits, 12 by default: a few thousand candidates
// expected): one candidate per 16-lane row with the lane-parallel
// permutation, sixteen per block per pass, so a pass takes one permutation's
// latency (about 2 us) instead of the 8 us of the one-lane form.  Same search
// order and stop rule as k_grind, with a wave's four candidates as the unit.

__global__ void __launch_bounds__(256) k_grind_wide(const P2C16* __restrict__ cp, Group<GrindArgs> ga, int tr_mode,
                                                    u32 mask, u32 start, u32 end) {
  // arguments copied out of the group first, as in k_grind
  const GrindArgs& a = ga.v[blockIdx.z];
  Sponge16 base = a.base;
  int wpos = a.wpos;
  u32* result = a.result;
  const ScTr tr = a.tr;
  if (tr_mode == TR_COEFFS_GRIND || tr_mode == TR_OBS_GRIND) {
    __shared__ GrindBase gbs;
    tr_prologue(tr_mode, tr, *cp, &gbs, nullptr);
#pragma unroll
    for (int k = 0; k < 16; k++) base.s[k] = gbs.s[k];
    wpos = (int)gbs.wpos;
  } else if (tr_mode == TR_GBASE) {
#pragma unroll
    for (int k = 0; k < 16; k++) base.s[k] = tr.gbase->s[k];
    wpos = (int)tr.gbase->wpos;
  }
  const P2C16& c = *cp;
  const int j = threadIdx.x & 15;
  const WideLane wl = wide_lane(c, j);
  u32 bj = 0;
#pragma unroll
  for (int k = 0; k < 16; k++)
    if (k == j) bj = base.s[k];
  const u32 stride = gridDim.x * 16;
  const u32 wave_off = blockIdx.x * 16 + (threadIdx.x >> 6) * 4;
  const u32 cand_off = blockIdx.x * 16 + (threadIdx.x >> 4);
  u32 b = best_known(result);
  for (u64 g = start; g + wave_off < end; g += stride) {
    const u64 first = g + wave_off;
    if ((u64)__builtin_amdgcn_readfirstlane(b) <= first) break;
    b = best_known(result);   // checked before the next pass
    const u64 w = g + cand_off;
    u32 x = (j == wpos) ? to_monty((u32)w) : bj;
    x = permute16_wide(x, j, wl, c);
    if (j == 7 && w < end && (from_monty(x) & mask) == 0) atomicMin(result, (u32)w);
  }
}

// Long searches (at least Tuning::grind_lock_bits bits, 17 by default): two waves per SIMD issue
// about 15% faster than one, but the older wave gets issue priority, so waves
// left to themselves split each pass into a fast and a slow half (and a
// witness in the slow half ends the search late).  Here a block holds both
// waves of each SIMD of its CU (512 lanes) and passes a barrier after every
// pass, so they advance together; the stop decision is block-wide: thread 0
// loads the best witness one pass ahead and shares it through LDS (double
// buffered, so a pass's write never meets the previous pass's reads).  The
// candidates' base state comes from k_grind_base (transcript step) or the
// host.
#ifndef LOCK_NT_OVERRIDE
constexpr int LOCK_NT = 512;
#else
constexpr int LOCK_NT = LOCK_NT_OVERRIDE;
#endif

template <int NT>
__global__ void __launch_bounds__(NT) k_grind_lock(const P2C16* __restrict__ cp, Group<GrindArgs> ga, int use_gbase,
                                                   u32 mask, u32 start, u32 end) {
  __shared__ u32 sb[2];
  const GrindArgs& a = ga.v[blockIdx.z];
  Sponge16 base = a.base;
  int wpos = a.wpos;
  u32* result = a.result;
  if (use_gbase) {
    const GrindBase* gb = a.tr.gbase;
#pragma unroll
    for (int k = 0; k < 16; k++) base.s[k] = gb->s[k];
    wpos = (int)gb->wpos;
  }
  GrindPre gp;
  grind_pre(gp, base.s, wpos);
  const u32 t = threadIdx.x;
  const u32 stride = gridDim.x * NT;
  if (t == 0) sb[0] = best_known(result);
  __syncthreads();
  int it = 0;
  for (u64 g = start;; g += stride, it ^= 1) {
    const u64 first = g + (u64)blockIdx.x * NT;
    if (first >= end || (u64)sb[it] <= first) break;   // block-uniform
    u32 nb = 0;
    if (t == 0) nb = best_known(result);
    u32 off = 0;
    asm volatile("" : "+s"(off));
    const P2C16& c = cp[off];
    const u64 w = first + t;
    u32 s[16];
    u64 y[16];
    grind_layer(gp, to_monty((u32)w), y);
    permute16_head_rounds(y, s, c);
    permute_internal<16>(s, c);
    const u32 s7 = permute16_tail_s7(s, c);
    if (t == 0) sb[it ^ 1] = nb;
    if (w < end && (from_monty(s7) & mask) == 0) atomicMin(result, (u32)w);
    __syncthreads();
  }
}

// The transcript step of a long search in one block per proof: writes the
// candidates' base state to tr.gbase (and the next challenger state and log).
__global__ void __launch_bounds__(256) k_grind_base(const P2C16* __restrict__ cp, int tr_mode, Group<GrindArgs> ga) {
  __shared__ GrindBase gbs;
  const ScTr& tr = ga.v[blockIdx.z].tr;
  GrindBase* gb_out = tr.gbase;
  tr_prologue(tr_mode, tr, *cp, &gbs, nullptr);
  if (threadIdx.x < 16) gb_out->s[threadIdx.x] = gbs.s[threadIdx.x];
  if (threadIdx.x == 0) gb_out->wpos = gbs.wpos;
}

// Blocks of a short search: sixteen candidates each, enough for one pass over
// sixteen times the expected search, at most `blocks` and the range.
static unsigned wide_grind_blocks(int bits, u64 start, u64 end, int blocks) {
  const u64 span = (16ull << bits) / 16;
  return (unsigned)std::max<u64>(1, std::min<u64>(std::min<u64>((end - start + 15) / 16, span),
                                                  (u64)std::max(blocks, 1)));
}

static_assert(sizeof(Group<GrindArgs>) <= 4096, "kernel arguments of a grinding group");

// Scans [start, start + span) (clipped to p) for every proof of the group in
// one launch of `blocks` blocks per proof (fewer when the range is short).
int grind_range(const P2C16* d_consts, const GrindArgs* a, int count, int bits, int mode, u64 start, u64 span,
                int blocks, const Tuning& t, hipStream_t st) {
  const u32 mask = bits >= 32 ? 0xffffffffu : ((1u << bits) - 1);
  if (start >= P) return 0;
  const u64 end = std::min<u64>(start + span, P);
  if (count < 1 || count > GROUP_MAX) return -8;
  if (bits <= t.grind_wide_bits) {
    const unsigned nb = wide_grind_blocks(bits, start, end, blocks);
    return launch_group(a, count, [&](const Group<GrindArgs>& g) {
      hipLaunchKernelGGL(k_grind_wide, dim3(nb, 1, count), dim3(256), 0, st, d_consts, g, mode, mask, (u32)start,
                         (u32)end);
    });
  }
  bool gbase = mode != TR_NONE;
  for (int i = 0; i < count; i++) gbase = gbase && a[i].tr.gbase;
  if (mode == TR_GBASE && !gbase) return -1;
  if (bits >= t.grind_lock_bits && (mode == TR_NONE || gbase)) {
    const unsigned nb = (unsigned)std::max<u64>(1, std::min<u64>((end - start + LOCK_NT - 1) / LOCK_NT,
                                                                 (u64)std::max(blocks, 1)));
    return launch_group(a, count, [&](const Group<GrindArgs>& g) {
      if (gbase && mode != TR_GBASE)
        hipLaunchKernelGGL(k_grind_base, dim3(1, 1, count), dim3(256), 0, st, d_consts, mode, g);
      hipLaunchKernelGGL(k_grind_lock<LOCK_NT>, dim3(nb, 1, count), dim3(LOCK_NT), 0, st, d_consts, g,
                         gbase ? 1 : 0, mask, (u32)start, (u32)end);
    });
  }
  const u64 need = (end - start + 255) / 256;
  const unsigned nb = (unsigned)std::min<u64>(need, (u64)std::max(blocks, 1));
  return launch_group(a, count, [&](const Group<GrindArgs>& g) {
    hipLaunchKernelGGL(k_grind, dim3(nb, 1, count), dim3(256), 0, st, d_consts, g, mode, mask, (u32)start, (u32)end);
  });
}

// ------------------------------------------------------------------ round transcript
// OOD sample of a round (prover/mod.rs:192-205): every block runs the step
// (observe st.obs, sample y), expands y to the point z_j = y^(2^(m-1-j))
// (expand_from_univariate, src/poly/multilinear.rs:73-85), then writes its
// part of the point's split eq tables, as k_eq_split.  Block 0 logs y,
// writes the point and the next challenger state.
__global__ void __launch_bounds__(256) k_ood_eq_tables(Group<OodArgs> g, int m, int hi, int lo,
                                                       const P2C16* __restrict__ cp) {
  const OodArgs& a = g.v[blockIdx.z];
  const StepTr& st = a.st;
  ef* log_y = a.log_y;
  ef* pt_out = a.pt_out;
  ef* __restrict__ out_hi = a.out_hi;
  ef* __restrict__ out_lo = a.out_lo;
  u32* grind_reset = a.grind_reset;
  __shared__ ef z[RT_MAX_VARS];
  const int t = threadIdx.x;
  const P2C16& c = *cp;
  const WideLane wl = wide_lane(c, t & 15);
  BlockChallenger ch{0, 0, 0, 0, 0, &wl, &c};
  ch.load(*st.ch_in);
  __shared__ u32 stage[OBS_STAGE];
  __shared__ ef red[16];
  observe_input(ch, st.obs, st.n_obs, st.obs_sum, st.obs_nb, st.log_obs, blockIdx.x == 0, stage, red, st.obs_copy);
  const ef y = ch.sample_ef();
  if (t == 0) {
    ef cur = y;
    for (int j = m - 1; j >= 0; j--) {
      z[j] = cur;
      cur = ef_mul(cur, cur);
    }
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    ch.store(st.ch_out);
    if (t == 0) {
      *log_y = y;
      if (grind_reset) *grind_reset = P;
    }
    if (t < m) pt_out[t] = z[t];
  }
  u32 x = blockIdx.x * 256 + t;
  const u32 nh = 1u << hi;
  if (x >= nh + (1u << lo)) return;
  int first = 0, cnt = hi;
  ef acc = ef_one();
  ef* o = out_hi;
  if (x >= nh) {
    x -= nh;
    first = hi;
    cnt = lo;
    o = out_lo;
  }
  for (int j = 0; j < cnt; j++) {
    const ef zj = z[first + j];
    const int bit = (x >> (cnt - 1 - j)) & 1;
    acc = ef_mul(acc, bit ? zj : ef_sub(ef_one(), zj));
  }
  o[x] = acc;
}

int ood_eq_tables(const OodArgs* a, int count, int m, int hi, int lo, const P2C16* d_consts, hipStream_t st) {
  if (m < 1 || m > RT_MAX_VARS || hi + lo != m) return -1;
  const u32 n = (1u << hi) + (1u << lo);
  return launch_group(a, count, [&](const Group<OodArgs>& g) {
    hipLaunchKernelGGL(k_ood_eq_tables, dim3((n + 255) / 256, 1, count), dim3(256), 0, st, g, m, hi, lo, d_consts);
  });
}

// One step of the commitment's OOD samples (see commit_ood).  The answer of a
// sample is InitialStatement::evaluate's (constraints/statement/initial.rs:
// 143-152) from the segment sums: sum_s eq(z[..sb], s) P(s).
__global__ void __launch_bounds__(256) k_commit_ood(Group<CoodArgs> g, int n, int sb, int hi, int lo, int sample,
                                                    const P2C16* __restrict__ cp) {
  __shared__ ef z[RT_MAX_VARS];
  __shared__ ef Psh[1 << COOD_MAX_SB];
  __shared__ ef red[16];
  __shared__ u32 stage[OBS_STAGE];
  const CoodArgs& a = g.v[blockIdx.z];
  const int t = threadIdx.x;
  const bool lead = blockIdx.x == 0;
  const P2C16& c = *cp;
  const WideLane wl = wide_lane(c, t & 15);
  BlockChallenger ch{0, 0, 0, 0, 0, &wl, &c};
  ch.load(*a.st.ch_in);
  const int S = 1 << sb;
  if (a.prev_part) {
    // segment s's sum by row (t >> 4) of the block, lanes over its blocks
    for (int s0 = 0; s0 < S; s0 += 16) {
      const int s = s0 + (t >> 4);
      ef acc = ef_zero();
      if (s < S)
        for (int b = t & 15; b < a.prev_nb; b += 16) acc = ef_add(acc, a.prev_part[(size_t)s * a.prev_nb + b]);
      acc = row_sum16(acc);
      if (s < S && (t & 15) == 0) Psh[s] = acc;
    }
    __syncthreads();
    ef term[1] = {ef_zero()};
    if (t < S) {
      ef w = Psh[t];
      for (int j = 0; j < sb; j++) {   // eq(z_j, bit j of s), bit 0 the top one
        const ef zj = a.prev_z[j];
        w = ef_mul(w, ((t >> (sb - 1 - j)) & 1) ? zj : ef_sub(ef_one(), zj));
      }
      term[0] = w;
      if (lead) a.prev_P[t] = Psh[t];
    }
    block_reduce_ef<1>(term, red);   // the answer in every thread
    if (lead && t == 0) *a.prev_eval = term[0];
    const u32 v[4] = {term[0].c[0], term[0].c[1], term[0].c[2], term[0].c[3]};
#pragma unroll
    for (int k = 0; k < 4; k++) ch.observe(v[k]);
  } else {
    ch.observe_words(a.st.obs, a.st.n_obs, stage);
  }
  if (sample) {
    const ef y = ch.sample_ef();
    if (t == 0) {
      ef cur = y;
      for (int j = n - 1; j >= 0; j--) {
        z[j] = cur;
        cur = ef_mul(cur, cur);
      }
    }
    __syncthreads();
    if (lead) {
      if (t == 0) *a.log_y = y;
      if (t < n) a.z_out[t] = z[t];
    }
    // this block's entries of the split tables of z[sb..n)
    u32 x = blockIdx.x * 256 + t;
    const u32 nh = 1u << hi;
    if (x < nh + (1u << lo)) {
      int first = sb, cnt = hi;
      ef* o = a.out_hi;
      if (x >= nh) {
        x -= nh;
        first = sb + hi;
        cnt = lo;
        o = a.out_lo;
      }
      ef acc = ef_one();
      for (int j = 0; j < cnt; j++) {
        const ef zj = z[first + j];
        acc = ef_mul(acc, ((x >> (cnt - 1 - j)) & 1) ? zj : ef_sub(ef_one(), zj));
      }
      o[x] = acc;
    }
  }
  if (lead) ch.store(a.st.ch_out);
}

int commit_ood(const CoodArgs* a, int count, int n, int sb, int hi, int lo, int sample, const P2C16* d_consts,
               hipStream_t st) {
  if (n < 1 || n > RT_MAX_VARS || sb < 0 || sb > COOD_MAX_SB || hi < 0 || lo < 0 || sb + hi + lo != n) return -1;
  const u32 tabs = (1u << hi) + (1u << lo);
  const unsigned blocks = sample ? (tabs + 255) / 256 : 1;
  return launch_group(a, count, [&](const Group<CoodArgs>& g) {
    hipLaunchKernelGGL(k_commit_ood, dim3(blocks, 1, count), dim3(256), 0, st, g, n, sb, hi, lo, sample, d_consts);
  });
}

// Every OOD sample of a round on a polynomial of at most 2^OOD_SMALL_LOG
// evaluations in one block per proof (prover/mod.rs:192-205 with the answer
// of evaluate_at_extension): sample y, expand it to the point, write its split
// eq tables, sum e against them and observe the answer before the next
// sample.  The per-sample launches (k_ood_eq_tables, the eq-weighted sum and
// its reduction) write the same tables, points, answers and log entries.
__global__ void __launch_bounds__(256) k_ood_small(Group<OodSmallArgs> g, int m, int hi, int lo, int n_ood,
                                                   const P2C16* __restrict__ cp) {
  __shared__ ef z[RT_MAX_VARS];
  __shared__ ef th[1 << (OOD_SMALL_LOG - 8)], tl[1 << 8];
  __shared__ ef sh[16];
  __shared__ u32 stage[OBS_STAGE];
  const OodSmallArgs& a = g.v[blockIdx.z];
  const int t = threadIdx.x;
  const P2C16& c = *cp;
  const WideLane wl = wide_lane(c, t & 15);
  BlockChallenger ch{0, 0, 0, 0, 0, &wl, &c};
  ch.load(*a.st.ch_in);
  // this thread's entries of e (at most four) before the first sample
  const int n = 1 << m;
  ef ex[4];
#pragma unroll
  for (int k = 0; k < 4; k++) ex[k] = t + 256 * k < n ? a.e[t + 256 * k] : ef_zero();
  ch.observe_words(a.st.obs, a.st.n_obs, stage);
  if (t == 0 && a.grind_reset) *a.grind_reset = P;
  const int nh = 1 << hi, nl = 1 << lo;
  for (int i = 0; i < n_ood; i++) {
    const ef y = ch.sample_ef();
    if (t == 0) {
      a.log_y[i] = y;
      ef cur = y;
      for (int j = m - 1; j >= 0; j--) {
        z[j] = cur;
        cur = ef_mul(cur, cur);
      }
    }
    __syncthreads();
    ef* pt = a.pts + (size_t)i * RT_MAX_VARS;
    if (t < m) pt[t] = z[t];
    ef* out_hi = a.tabs + (size_t)i * a.tab_cap;
    ef* out_lo = out_hi + nh;
    for (int x = t; x < nh + nl; x += 256) {
      int first = 0, cnt = hi, xx = x;
      if (x >= nh) {
        xx = x - nh;
        first = hi;
        cnt = lo;
      }
      ef acc = ef_one();
      for (int j = 0; j < cnt; j++) {
        const ef zj = z[first + j];
        const int bit = (xx >> (cnt - 1 - j)) & 1;
        acc = ef_mul(acc, bit ? zj : ef_sub(ef_one(), zj));
      }
      if (x < nh) {
        th[xx] = acc;
        out_hi[xx] = acc;
      } else {
        tl[xx] = acc;
        out_lo[xx] = acc;
      }
    }
    __syncthreads();
    ef acc[1] = {ef_zero()};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int x = t + 256 * k;
      if (x < n) acc[0] = ef_add(acc[0], ef_mul(ex[k], ef_mul(th[x >> lo], tl[x & (nl - 1)])));
    }
    block_reduce_ef<1>(acc, sh);   // the answer in every thread
    if (t == 0) a.ev[i] = acc[0];
    if (i + 1 < n_ood) {
      if (t == 0) a.log_ev[i] = acc[0];
      const u32 v[4] = {acc[0].c[0], acc[0].c[1], acc[0].c[2], acc[0].c[3]};
#pragma unroll
      for (int k = 0; k < 4; k++) ch.observe(v[k]);
    }
    __syncthreads();   // z, th, tl and sh are rewritten by the next sample
  }
  ch.store(a.st.ch_out);
}

int ood_small(const OodSmallArgs* a, int count, int m, int hi, int lo, int n_ood, const P2C16* d_consts,
              hipStream_t st) {
  if (m < 1 || m > OOD_SMALL_LOG || hi + lo != m || lo > 8 || n_ood < 1 || n_ood > RT_MAX_OOD) return -1;
  return launch_group(a, count, [&](const Group<OodSmallArgs>& g) {
    hipLaunchKernelGGL(k_ood_small, dim3(1, 1, count), dim3(256), 0, st, g, m, hi, lo, n_ood, d_consts);
  });
}

// The round's steps after its OOD samples, in one block (prover/mod.rs:
// 206-234): the proof-of-work witness is observed and its bits checked (or,
// without proof of work, the last OOD answer / the root observed), one sample
// is discarded, the STIR query indices are sampled (get_challenge_stir_queries,
// src/whir/utils.rs:24-145: sample_bits calls of up to 20 bits, split into
// per-query indices), sorted and deduplicated, and gamma is sampled.  The
// final round (prover/mod.rs:343-377) has neither the discarded sample nor
// gamma (final_step).
__global__ void __launch_bounds__(256) k_round_queries(Group<QueryArgs> g, int pow_bits, u64 domain_size, int ff,
                                                       int num_queries, u32 folded_gen, int n_ood,
                                                       const P2C16* __restrict__ cp, int final_step) {
  const QueryArgs& a = g.v[blockIdx.z];
  const StepTr& st = a.st;
  u32* grind = a.grind;
  RoundLog* log = a.log;
  const RoundBufs& rb = a.rb;
  __shared__ u64 q[RT_MAX_QUERIES];
  __shared__ u32 firsts[RT_MAX_QUERIES];
  const int t = threadIdx.x;
  const P2C16& c = *cp;
  const WideLane wl = wide_lane(c, t & 15);
  BlockChallenger ch{0, 0, 0, 0, 0, &wl, &c};
  ch.load(*st.ch_in);
  u32 err = 0, wm = 0;
  if (pow_bits > 0) {
    const u32 w = *grind;
    wm = w < P ? to_monty(w) : 0u;
    if (w >= P) err = 1;
    ch.observe(wm);
    const u64 pmask = pow_bits >= 32 ? 0xffffffffull : ((1ull << pow_bits) - 1);
    if (((u64)from_monty(ch.sample()) & pmask) != 0) err = 1;
  } else {
    __shared__ u32 stage[OBS_STAGE];
    __shared__ ef red[16];
    observe_input(ch, st.obs, st.n_obs, st.obs_sum, st.obs_nb, st.log_obs, true, stage, red, st.obs_copy);
  }
  if (!final_step) (void)ch.sample();
  const u64 folded = domain_size >> ff;
  int bits = 0;
  while ((1ull << bits) < folded) bits++;
  const u64 mask = (1ull << bits) - 1;
  const u64 total = (u64)num_queries * (u64)bits;
  if (total <= 20) {
    u64 all = (u64)from_monty(ch.sample()) & ((1ull << total) - 1);
    for (int i = 0; i < num_queries; i++, all >>= bits)
      if (t == 0) q[i] = (all & mask) % folded;
  } else {
    const int per_call = bits > 0 ? 20 / bits : 0;
    if (per_call >= 2) {
      int i = 0;
      for (int left = num_queries; left > 0;) {
        const int batch = left < per_call ? left : per_call;
        u64 all = (u64)from_monty(ch.sample()) & ((1ull << (batch * bits)) - 1);
        for (int j = 0; j < batch; j++, all >>= bits, i++)
          if (t == 0) q[i] = (all & mask) % folded;
        left -= batch;
      }
    } else {
      for (int i = 0; i < num_queries; i++) {
        const u64 v = (u64)from_monty(ch.sample()) & mask;
        if (t == 0) q[i] = v;
      }
    }
  }
  __syncthreads();
  // sorted distinct indices: a first occurrence goes to the number of
  // distinct smaller values.  Every thread runs both scans over all entries
  // (the same bound everywhere, unrolled, so the broadcast LDS reads are
  // issued back to back) and counts the distinct values itself.
  const u64 v = t < num_queries ? q[t] : 0;
  u32 dup = 0;
#pragma unroll 8
  for (int j = 0; j < num_queries; j++) dup |= (u32)(j < t) & (u32)(q[j] == v);
  if (t < num_queries) firsts[t] = dup ^ 1u;
  __syncthreads();
  u32 nq = 0, pos = 0;
#pragma unroll 8
  for (int j = 0; j < num_queries; j++) {
    const u32 f = firsts[j];
    nq += f;
    pos += f & (u32)(q[j] < v);
  }
  if (t < num_queries && !dup) {
    rb.idx[pos] = v;
    rb.selz[pos] = pow(folded_gen, v);
    log->idx[pos] = (u32)v;
  }
  ef gamma = ef_zero();
  if (!final_step) {
    gamma = ch.sample_ef();
    for (u32 i = t; i < (u32)n_ood + nq; i += 256) rb.gpow[i] = ef_pow(gamma, i);
  }
  ch.store(st.ch_out);
  if (t == 0) {
    *rb.nq = nq;
    log->nq = nq;
    log->gamma = gamma;
    log->witness = wm;
    log->err = err;
    // every thread read the witness before the barriers above
    if (grind) *grind = P;
  }
}

int round_queries(const QueryArgs* a, int count, int pow_bits, u64 domain_size, int folding_factor, int num_queries,
                  u32 folded_gen, int n_ood, const P2C16* d_consts, hipStream_t st, bool final_step) {
  if (num_queries < 1 || num_queries > RT_MAX_QUERIES || n_ood < 0 || n_ood > RT_MAX_OOD) return -1;
  return launch_group(a, count, [&](const Group<QueryArgs>& g) {
    hipLaunchKernelGGL(k_round_queries, dim3(1, 1, count), dim3(256), 0, st, g, pow_bits, domain_size,
                       folding_factor, num_queries, folded_gen, n_ood, d_consts, final_step ? 1 : 0);
  });
}

}  // namespace kb
