// Reed-Solomon encoding: transpose, zero-pad and column DFT of an evaluation
// list, as the committer does at src/whir/committer/writer.rs:70-86 (base
// field) and the prover at src/whir/prover/mod.rs:166-180 (extension field,
// base-field twiddles).
//
// Column a holds c_a[b] = src[a*L + b], b < L = 2^l, zero-padded to H = 2^R * L,
// and row i of the output is [X_0[i], ..., X_{W-1}[i]] with
//   X_a[i] = sum_b c_a[b] w_H^(i b).
// Writing i = 2^R * i1 + i0 (i0 < 2^R, the coset) splits this into 2^R DFTs of
// length L of the twisted column c_a[b] * w_H^(i0 b), so the padding zeros are
// never read or multiplied.
//
// A length-L DFT with L <= 2^LP runs in one pass (k_lde_pass2, DIRECT);
// longer ones use the four-step split L = L1 * L2:
//   pass 1: per (column, coset, b2): DFT over b1 of c[L2 b1 + b2] (twisted),
//           times w_L^(b2 k1), into V[coset][column][k1][b2];
//   pass 2: per (coset, k1): DFT over b2 of V[..][k1][b2] for all W columns at
//           once, written to whole output rows 2^R (k1 + L1 k2) + i0.
// Each block keeps its lines in LDS in bit-reversed order and runs radix-2
// decimation-in-time stages there.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "kcommon.h"
#include "kernels.h"

namespace kb {

// ---- line DFTs in LDS
// A block holds `nlines` lines of length N = 2^LG in LDS, each in bit-reversed
// order, and runs the LG radix-2 decimation-in-time stages over them.  Stages
// are taken R at a time (R <= 5 for base-field lines, <= 4 for extension
// lines): each thread loads a group of 2^R elements into registers, runs the R
// stages there and stores them back, so a line of 1024 costs two LDS round
// trips and two barriers instead of ten.
//
// Values between the first load and pass 2's output are arbitrary 32-bit
// representatives (any u32 stands for its residue mod p): a butterfly with
// twiddle w maps (a, b) to (a + t, a - t) with t = mul_lazy(b, w) < 2p, and a
// carry (borrow) out of 32 bits is worth +2^32 (-2^32) = +-R1 (mod p), which
// wadd / wsub fold back without overflow because t < 2p and 2p + R1 = 2^32.
// A butterfly is nine instructions (three for the product, three each for the
// sum and the difference) against eleven with canonical values, in 32-bit
// registers; nothing is reduced between stages or passes, and only pass 2's
// outputs are made canonical.
//
// Layout: element i of line l sits at l * LS + i + (i >> 5).  The pad after
// every 32 elements keeps a wave's accesses on distinct banks both when lanes
// walk groups of one line (stride 33) and when they walk lines at the same
// position (odd line stride LS).
//
// Twiddles: stw[2^(s-1) + e] = w_{2^s}^e for stage s (1..LG) and e < 2^(s-1),
// so the lanes of a wave read consecutive entries.
template <int LG>
struct LineShape {
  static constexpr int N = 1 << LG;
  static constexpr int LS = N + (N >> 5) + 1;
  static __host__ __device__ constexpr int pos(int i) { return i + (i >> 5); }
};

static constexpr int lde_min(int a, int b) { return a < b ? a : b; }

// 32-bit components of an element (1 for the base field, 4 for EF4).
template <typename T>
struct Comps {
  static constexpr int value = sizeof(T) / sizeof(u32);
};
__device__ __forceinline__ u32 comp_of(const u32& x, int) { return x; }
__device__ __forceinline__ u32 comp_of(const ef& x, int c) { return x.c[c]; }
__device__ __forceinline__ void set_comp(u32& x, int, u32 v) { x = v; }
__device__ __forceinline__ void set_comp(ef& x, int c, u32 v) { x.c[c] = v; }

// a + t and a - t for a < 2^32, t < 2p, as representatives below 2^32.  On a
// carry a + t - 2^32 < t < 2p, so adding R1 = 2^32 - 2p stays below 2^32; on a
// borrow a - t + 2^32 > 2^32 - 2p = R1, so taking R1 off stays nonnegative.
__device__ __forceinline__ u32 wadd(u32 a, u32 t) {
  u32 s;
  const bool c = __builtin_add_overflow(a, t, &s);
  return s + (c ? R1 : 0u);
}
__device__ __forceinline__ u32 wsub(u32 a, u32 t) {
  u32 d;
  const bool b = __builtin_sub_overflow(a, t, &d);
  return d - (b ? R1 : 0u);
}
// x < 2^32 -> below 2p
__device__ __forceinline__ u32 csub2p(u32 x) {
  const u32 t = x - TWO_P;
  return t < x ? t : x;
}
// x < 2^32 (< 3p) -> canonical
__device__ __forceinline__ u32 canon_any(u32 x) { return canon(csub2p(x)); }

// One butterfly.  narrow: b is known to be below 2p (a value loaded in the
// line's first stage); unit: w = 1.
__device__ __forceinline__ void wbfly(u32& a, u32& b, u32 w, bool narrow, bool unit) {
  const u32 t = unit ? (narrow ? b : csub2p(b)) : mul_lazy(b, w);
  const u32 av = a;
  a = wadd(av, t);
  b = wsub(av, t);
}

// Stages S0+1 .. S0+R: for a group with base index b (bits S0..S0+R-1 clear)
// the elements are b + (j << S0), j < 2^R; at local stage g the pair (j, j + h),
// h = 2^(g-1), uses w_{2^(S0+g)}^((b mod 2^S0) + (j mod h) 2^S0).  As b and
// j << S0 have no bit in common, pos(b + (j << S0)) = pos(b) + pos(j << S0):
// every access of a group is a constant offset from one address.
template <typename T, int LG, int R, int S0, bool UNIT>
__device__ __forceinline__ void dit_pass(T* sh, int nlines, const u32* stw) {
  using L = LineShape<LG>;
  constexpr int G = 1 << R;
  constexpr int NC = Comps<T>::value;
  constexpr int GPL = L::N >> R;
  const int ngroups = nlines * GPL;
  for (int g = threadIdx.x; g < ngroups; g += blockDim.x) {
    const int line = g / GPL;
    const int gl = g % GPL;
    const int base_low = gl & ((1 << S0) - 1);
    const int base = base_low + ((gl >> S0) << (S0 + R));
    T* q = sh + line * L::LS + L::pos(base);
    const u32* tws = stw + base_low;
    u32 v[G][NC];
#pragma unroll
    for (int j = 0; j < G; j++) {
      const T x = q[L::pos(j << S0)];
#pragma unroll
      for (int c = 0; c < NC; c++) v[j][c] = comp_of(x, c);
    }
#pragma unroll
    for (int sg = 1; sg <= R; sg++) {
      const int h = 1 << (sg - 1);
#pragma unroll
      for (int j = 0; j < G; j++) {
        if (j & h) continue;
        const int jl = j & (h - 1);
        const bool unit = UNIT && S0 == 0 && jl == 0;
        const u32 w = unit ? 0u : tws[(1 << (S0 + sg - 1)) + (jl << S0)];
#pragma unroll
        for (int c = 0; c < NC; c++) wbfly(v[j][c], v[j + h][c], w, S0 == 0 && sg == 1, unit);
      }
    }
#pragma unroll
    for (int j = 0; j < G; j++) {
      T x;
#pragma unroll
      for (int c = 0; c < NC; c++) set_comp(x, c, v[j][c]);
      q[L::pos(j << S0)] = x;
    }
  }
}

// UNIT: the table's first entry of every stage is 1 (an untwisted table), so
// those butterflies skip their product.
template <typename T, int LG, int RMAX, bool UNIT, int S0 = 0>
__device__ __forceinline__ void dit_lines(T* sh, int nlines, const u32* stw) {
  if constexpr (S0 < LG) {
    constexpr int R = lde_min(RMAX, LG - S0);
    dit_pass<T, LG, R, S0, UNIT>(sh, nlines, stw);
    __syncthreads();
    dit_lines<T, LG, RMAX, UNIT, S0 + R>(sh, nlines, stw);
  }
}

template <typename T>
struct RadixMax {
  static constexpr int value = sizeof(T) == 4 ? 5 : 3;
};

// Stage twiddles of a twisted line DFT Z[k] = sum_n x[n] theta^n w_N^(k n),
// N = 2^LG, theta = w_H^kappa (H = 2^H_log): decimation in time with
// stage-s twiddles theta^(N / 2^s) w_{2^s}^e, which is
//   stw[2^(s-1) + e] = w_H^(kappa 2^(LG - s) + e 2^(H_log - s)).
// kappa = 0 gives the plain table w_{2^s}^e.  A per-element twist or a
// four-step twiddle that depends on the line position only through theta^n
// thus costs one table of N entries per tile instead of a product per element.
// A thread's entries of such a table (j = threadIdx.x + r NT) as the two
// table words of each root_pow, fetched ahead (the loads in flight while
// other work runs) and multiplied when the table is written.
template <int LG, int NT>
struct TwPend {
  static constexpr int n = ((1 << LG) + NT - 1) / NT;
  u32 hi[n], lo[n];
};
template <int LG, int NT>
__device__ __forceinline__ void tw_fetch(TwPend<LG, NT>& p, const Twiddles& tw, int H_log, u64 kappa) {
#pragma unroll
  for (int r = 0; r < TwPend<LG, NT>::n; r++) {
    const int j = (int)threadIdx.x + r * NT;
    if (j < (1 << LG) && j) {
      const int s = 32 - __clz(j);   // j in [2^(s-1), 2^s)
      const u64 x = (kappa << (LG - s)) + ((u64)(j - (1 << (s - 1))) << (H_log - s));
      const u32 e = (u32)((x << (24 - H_log)) & 0xffffffu);
      p.hi[r] = tw.hi[e >> 12];
      p.lo[r] = tw.lo[e & 4095u];
    }
  }
}
template <int LG, int NT>
__device__ __forceinline__ void tw_write(const TwPend<LG, NT>& p, u32* stw) {
#pragma unroll
  for (int r = 0; r < TwPend<LG, NT>::n; r++) {
    const int j = (int)threadIdx.x + r * NT;
    if (j < (1 << LG) && j) stw[j] = mul(p.hi[r], p.lo[r]);
  }
}
template <int LG>
__device__ __forceinline__ void fill_stage_twiddles(u32* stw, const Twiddles& tw, int H_log, u64 kappa) {
  for (int j = threadIdx.x; j < (1 << LG); j += blockDim.x) {
    if (j == 0) continue;
    const int s = 32 - __clz(j);   // j in [2^(s-1), 2^s)
    stw[j] = root_pow(tw, H_log, (kappa << (LG - s)) + ((u64)(j - (1 << (s - 1))) << (H_log - s)));
  }
}

// LDS: NTAB stage-twiddle tables (N words each, 16-byte aligned) then the
// padded lines.
template <typename T, int LG, int NTAB = 1>
__device__ __forceinline__ T* line_storage(unsigned char* smem) {
  return reinterpret_cast<T*>(smem + NTAB * (((sizeof(u32) << LG) + 15) & ~size_t(15)));
}

// Launch-order block b of nb -> logical block.  The GPU hands launch-order
// blocks to its 8 XCDs in turn, and each XCD has its own L2; this order gives
// every XCD one contiguous range of logical blocks, so neighbouring blocks,
// which read and write the neighbouring 64-byte pieces of the same 128-byte
// lines, share an L2.  (The identity when nb is not a multiple of 8.)
__device__ __forceinline__ u32 xcd_order(u32 b, u32 nb) {
  return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3);
}

// Buffer-resource loads and stores: a 32-bit offset per thread plus a uniform
// (scalar) offset per row, so a row costs no address arithmetic.  Every range
// addressed this way is below 2^31 bytes (a column: at most 2^24 extension
// elements; a block's pass-2 lines: 2^m elements, checked by lde_plan);
// accesses past num_records read zero and write nothing instead of faulting.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, u32 bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, u32 voff, u32 soff);
template <>
__device__ __forceinline__ u32 buf_load<u32>(__amdgpu_buffer_rsrc_t r, u32 voff, u32 soff) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0);
}
template <>
__device__ __forceinline__ ef buf_load<ef>(__amdgpu_buffer_rsrc_t r, u32 voff, u32 soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0);
  return ef{{v[0], v[1], v[2], v[3]}};
}
__device__ __forceinline__ void buf_store(u32 x, __amdgpu_buffer_rsrc_t r, u32 voff, u32 soff) {
  __builtin_amdgcn_raw_buffer_store_b32(x, r, (int)voff, (int)soff, 0);
}
// A 16-byte store keeps a zero scalar offset: with a register soffset the
// compiler omits the wait state a store of more than 8 bytes needs before a
// VALU instruction overwrites its data registers, and the store then wrote
// the overwritten first 8 bytes now and then (measured on MI355X).
__device__ __forceinline__ void buf_store(const ef& x, __amdgpu_buffer_rsrc_t r, u32 voff, u32 soff) {
  u32x4 v;
  v[0] = x.c[0];
  v[1] = x.c[1];
  v[2] = x.c[2];
  v[3] = x.c[3];
  __builtin_amdgcn_raw_buffer_store_b128(v, r, (int)(voff + soff), 0, 0);
}

// Twiddle tables a pass-1 block keeps (one per coset of its pair, while
// they fit beside the lines).
template <int LG>
struct Pass1Tables {
  static constexpr int value = LG <= 12 ? 2 : 1;
};

// Pass-1 blocks: 16K base or 4K extension elements (64 KiB of lines, two
// blocks per CU) in 512 threads, four waves per SIMD.
template <typename T>
struct Pass1Shape {
  static constexpr int threads = 512;
  static constexpr int elems = sizeof(T) == 4 ? 16384 : 4096;
  static constexpr int elems_log = sizeof(T) == 4 ? 14 : 12;
  static constexpr int per_thread = elems / threads;
};

// Blocks stay resident and take tiles in turn (tile = lt, lt + G, ... for G
// blocks), the next tile's loads in flight while the current one is
// transformed: a block reads, transforms and writes in sequence, so without
// this the memory and arithmetic phases of all resident blocks coincide.
// Logical tiles are handed out in XCD order (xcd_order of the block index
// within each round of G tiles), so neighbouring tiles run at the same time
// on one XCD and share its L2.

// Pass 1 of the four-step split.  Grid: (G, 1, proofs); LG = l1; TB =
// 2^tb_log lines (column offsets b2) per tile, TB * 2^LG at most
// Pass1Shape::elems.  With b = L2 b1 + b2 and output rows i = 2^R (k1 + L1 k2)
// + i0, w_H^(i b) splits into w_L1^(k1 b1) w_H^(i0 L2 b1) (pass 1, a line DFT
// over b1 twisted by theta = w_H^(i0 L2)) times w_H^(kappa b2) w_L2^(k2 b2)
// with kappa = 2^R k1 + i0 (pass 2), so pass 1 writes
//   V[i0][a][k1][b2] = sum_b1 c_a[L2 b1 + b2] w_H^(i0 L2 b1) w_L1^(k1 b1)
// with no product per element: the coset twist lives in the stage twiddle
// table.  Tile (a, chunk, coset pair), pair fastest: a tile loads its segment
// of one column once and transforms it for 2^CB cosets (CB = min(R, 1)); the
// next tile's loads go out once the last coset's lines are in LDS.  Thread t
// handles line t mod TB and rows b1 = t / TB + j * (NT / TB); as t / TB and
// j * (NT / TB) have no bit in common, its LDS positions (bit-reversed,
// padded) and memory offsets are a per-thread part plus a part uniform across
// the block.
template <typename T, int LG>
__global__ void __launch_bounds__(Pass1Shape<T>::threads) __attribute__((amdgpu_waves_per_eu(4)))
k_lde_pass1(Group<LdeArgs<T>> g, int l, int R, int k, int tb_log, Twiddles tw) {
  extern __shared__ __align__(16) unsigned char smem[];
  using L = LineShape<LG>;
  constexpr int NT = Pass1Shape<T>::threads;
  constexpr int NR = Pass1Shape<T>::per_thread;
  const LdeArgs<T>& io = g.v[blockIdx.z];
  const T* __restrict__ src = io.src;
  T* __restrict__ V = io.tmp;
  const int l2 = l - LG;
  const int W = 1 << k;
  const int cl = l2 - tb_log;         // log2 of a column's chunks
  const int CB = R < 1 ? R : 1;       // log2 of the cosets per tile
  const int RG = R - CB;              // log2 of the coset groups
  const u32 ntiles = ((1u << cl) << k) << RG;
  const u32 G = gridDim.x;
  const u32 me = xcd_order(blockIdx.x, G);
  // One twiddle table per coset of the pair, rewritten when a tile's pair
  // differs from the previous tile's (never for R <= 1); lines longer than
  // 2^12 leave room for one table only, rewritten for every coset.
  constexpr int NTAB = Pass1Tables<LG>::value;
  constexpr u32 TAB = ((sizeof(u32) << LG) + 15) & ~size_t(15);
  u32* stw = reinterpret_cast<u32*>(smem);
  T* sh = line_storage<T, LG, NTAB>(smem);
  int tab_ig = -1;
  const int t = threadIdx.x & ((1 << tb_log) - 1);
  const int row0 = threadIdx.x >> tb_log;
  const int rstep = NT >> tb_log;
  const u32 col_bytes = (u32)sizeof(T) << l;
  const u32 rbytes = (u32)((rstep << l2) * sizeof(T));   // one row step
  T* lines = sh + t * L::LS;
  const int lb = L::pos(bitrev(row0, LG));   // LDS position of row0 (bit-reversed)
  const int lo = L::pos(row0);               // LDS position of k1 = row0 (natural order)
  // Row u of this thread exists iff u * rstep + row0 < N.  As row0 < rstep and
  // both rstep and N are powers of two, that is u * rstep < N (uniform) and
  // row0 < N (the same for every u), so no mask is kept per row.
  const bool row_ok = row0 < L::N;
  // tile lt -> first coset, column and column offset
  auto tile_ig = [&](u32 lt) { return (int)(lt & ((1u << RG) - 1)) << CB; };
  auto tile_a = [&](u32 lt) { return (int)((lt >> RG) >> cl); };
  auto tile_b2 = [&](u32 lt) { return ((int)((lt >> RG) & ((1u << cl) - 1)) << tb_log) + t; };
  T xs[NR];
  auto load = [&](u32 lt) {
    const __amdgpu_buffer_rsrc_t col = buf_rsrc(src + ((u64)tile_a(lt) << l), col_bytes);
    const u32 voff = (u32)(((row0 << l2) + tile_b2(lt)) * sizeof(T));
    u32 rb = rbytes;
    int rs = rstep;
    asm volatile("" : "+s"(rb), "+s"(rs));   // offsets formed at each load, not held across the tile loop
#pragma unroll
    for (int u = 0; u < NR; u++)
      if (u * rs < L::N && row_ok) xs[u] = buf_load<T>(col, voff, (u32)u * rb);
  };
  u32 lt = me;
  if (lt < ntiles) load(lt);
  while (lt < ntiles) {
    const int ig = tile_ig(lt), a = tile_a(lt), b2 = tile_b2(lt);
    const u32 next = lt + G;
    const u32 voff = (u32)(((row0 << l2) + b2) * sizeof(T));
    if (NTAB == 2 && ig != tab_ig) {   // uniform; the previous tile ended with a barrier
      for (int ci = 0; ci < (1 << CB); ci++)
        fill_stage_twiddles<LG>(stw + ci * (TAB / 4), tw, l + R, (u64)(ig + ci) << l2);
      tab_ig = ig;
    }
    for (int ci = 0; ci < (1 << CB); ci++) {
      const int i0 = ig + ci;
      if (NTAB == 1) fill_stage_twiddles<LG>(stw, tw, l + R, (u64)i0 << l2);
      const u32* tab = stw + (NTAB == 2 ? ci * (TAB / 4) : 0);
      // Uniform offsets pass through an empty asm per coset, so the compiler
      // forms each where it is used instead of holding all of them (in scalar
      // registers it does not have) across the loops.
      const int lbc = opaque(lb);
      int rs = rstep;
      asm volatile("" : "+s"(rs));
#pragma unroll
      for (int u = 0; u < NR; u++)
        if (u * rs < L::N && row_ok) lines[lbc + L::pos(bitrev(u * rs, LG))] = xs[u];
      __syncthreads();
      if (ci == (1 << CB) - 1 && next < ntiles) load(next);   // in flight while this coset is transformed
      if (i0 == 0)
        dit_lines<T, LG, RadixMax<T>::value, true>(sh, 1 << tb_log, tab);
      else
        dit_lines<T, LG, RadixMax<T>::value, false>(sh, 1 << tb_log, tab);
      const __amdgpu_buffer_rsrc_t out = buf_rsrc(V + (((u64)i0 * W + a) << l), col_bytes);
      const int loc = opaque(lo);
      u32 rb = rbytes;
      rs = rstep;
      asm volatile("" : "+s"(rb), "+s"(rs));
#pragma unroll
      for (int j = 0; j < NR; j++)
        if (j * rs < L::N && row_ok) buf_store(lines[loc + L::pos(j * rs)], out, voff, (u32)j * rb);
      __syncthreads();   // lines and tables are read before the next coset or tile overwrites them
    }
    lt = next;
  }
}

// Threads per pass-2 block.  Its W lines take about 64 KiB of LDS, so a CU
// holds two blocks: base-field blocks of 512 threads keep four waves per SIMD;
// extension blocks stay at 256.  A thread holds at most Pass2Shape::per_thread
// elements of a tile (its prefetch registers); lde_plan keeps W * 2^l2 within.
template <typename T>
struct Pass2Shape {
  static constexpr int threads = sizeof(T) == 4 ? 512 : 256;
  static constexpr int per_thread = sizeof(T) == 4 ? 32 : 16;
};

// Pass 2 (or, DIRECT, the only pass; then the twist is the coset's, w_H^(i0 b)).
// FULL: the tile fills every thread's
// registers (W * 2^LG = NE * NT), so no element needs a guard.  Grid: (G, 1,
// proofs), tiles
// as in pass 1; LG = l2.  Tile k1 * 2^R + i0 (so the tiles writing the
// adjacent rows of one k1 are neighbours) transforms the W lines of (i0, k1)
// and writes whole output rows 2^R (k1 + L1 k2) + i0, canonical.  Element e
// of a tile is column e >> LG, offset e & (L2 - 1); for e = t + u NT (t < NT)
// the two terms have no bit in common, so every offset and LDS position
// below is a per-thread part plus a uniform part.
template <typename T, int LG, bool DIRECT, bool FULL>
__global__ void __launch_bounds__(Pass2Shape<T>::threads)
    __attribute__((amdgpu_waves_per_eu(Pass2Shape<T>::threads / 128)))
    k_lde_pass2(Group<LdeArgs<T>> g, int l, int R, int k, Twiddles tw) {
  extern __shared__ __align__(16) unsigned char smem[];
  const LdeArgs<T>& io = g.v[blockIdx.z];
  T* __restrict__ dst = io.dst;
  using L = LineShape<LG>;
  constexpr int NT = Pass2Shape<T>::threads;
  constexpr int NE = Pass2Shape<T>::per_thread;
  constexpr int NC = Comps<T>::value;
  constexpr int L2 = L::N;
  const int l1 = l - LG;
  const int W = 1 << k;
  const u32 ntiles = (1u << l1) << R;
  const u32 G = gridDim.x;
  const u32 me = xcd_order(blockIdx.x, G);
  u32* stw = reinterpret_cast<u32*>(smem);
  T* sh = line_storage<T, LG>(smem);
  const int total = W * L2;
  const int t = threadIdx.x;
  // e + t < total for e a multiple of NT: as total and NT are powers of two,
  // that is e < total (uniform) and t < total (the same for every e).
  const bool t_ok = FULL || t < total;
  auto src_off_l = [](int e, int ll) { return (u32)((((e >> LG) << ll) + (e & (L2 - 1))) * (int)sizeof(T)); };
  auto src_off = [&](int e) { return src_off_l(e, l); };
  auto lds_pos = [&](int e) { return (e >> LG) * L::LS + L::pos(bitrev(e & (L2 - 1), LG)); };
  const u32 voff = src_off(t);
  const int lt_in = lds_pos(t);
  T xs[NE];
  TwPend<LG, NT> twp;
  auto load = [&](u32 lt) {
    const int i0 = (int)(lt & ((1u << R) - 1)), k1 = (int)(lt >> R);
    // the tile's lines are twisted by w_H^(kappa b2), kappa = 2^R k1 + i0
    // (DIRECT: the coset twist w_H^(i0 b), k1 = 0)
    tw_fetch(twp, tw, l + R, ((u64)k1 << R) + i0);
    // the tile's W lines: column a at a * L, offset b2
    const T* lsrc = DIRECT ? io.src : io.tmp + ((((u64)i0 * W) << l) + ((u64)k1 << LG));
    const __amdgpu_buffer_rsrc_t in = buf_rsrc(lsrc, (u32)sizeof(T) << (l + k));
    int ll = l;
    asm volatile("" : "+s"(ll));   // offsets formed at each load, not held across the tile loop
#pragma unroll
    for (int u = 0; u < NE; u++)
      if (FULL || (u * NT < total && t_ok)) xs[u] = buf_load<T>(in, voff, src_off_l(u * NT, ll));
  };
  // e = k2 * W + a -> dst row ((k1 + L1 k2) << R) + i0, column a
  const int rk = l1 + R + k;   // log2 of one k2 step in elements
  auto dst_off_k = [](int e, int kk, int rkk) { return ((u64)(e >> kk) << rkk) + (u64)(e & ((1 << kk) - 1)); };
  auto lds_out_k = [](int e, int kk) { return (e & ((1 << kk) - 1)) * L::LS + L::pos(e >> kk); };
  auto dst_off = [&](int e) { return dst_off_k(e, k, rk); };
  auto lds_out = [&](int e) { return lds_out_k(e, k); };
  const u64 dt_off = dst_off(t);
  const int lt_out = lds_out(t);
  u32 lt = me;
  if (lt < ntiles) load(lt);
  while (lt < ntiles) {
    const int i0 = (int)(lt & ((1u << R) - 1)), k1 = (int)(lt >> R);
    tw_write(twp, stw);
#pragma unroll
    for (int u = 0; u < NE; u++) {
      const int e = u * NT;
      if (FULL || (e < total && t_ok)) sh[lt_in + lds_pos(e)] = xs[u];
    }
    __syncthreads();
    const u32 next = lt + G;
    if (next < ntiles) load(next);   // in flight while this tile is transformed
    dit_lines<T, LG, RadixMax<T>::value, false>(sh, W, stw);
    T* dt = dst + ((((u64)k1 << R) + i0) << k) + dt_off;
    const int lo = opaque(lt_out);
    int kk = k, rkk = rk;
    asm volatile("" : "+s"(kk), "+s"(rkk));
    if (t_ok) {
#pragma unroll
      for (int u = 0; u < NE; u++) {
        const int e = u * NT;
        if (FULL || e < total) {
          T x = sh[lo + lds_out_k(e, kk)];
#pragma unroll
          for (int c = 0; c < NC; c++) set_comp(x, c, canon_any(comp_of(x, c)));
          dt[dst_off_k(e, kk, rkk)] = x;
        }
      }
    }
    __syncthreads();   // lines and twiddles are read before the next tile overwrites them
    lt = next;
  }
}

// Dynamic LDS above 64 KiB must be opted into per kernel and device, and a
// launch takes as many blocks as the device holds at once.  Both are settled
// once per kernel, device and LDS size (hipFuncSetAttribute takes a
// runtime-wide lock, which concurrent provers would otherwise contend for on
// every launch).  Returns the resident blocks (at least 1).
template <typename K>
static int prepare_kernel(K kernel, int threads, size_t bytes) {
  static std::mutex mu;
  static std::map<std::tuple<const void*, int, size_t>, int> done;
  int dev = 0;
  (void)hipGetDevice(&dev);
  const auto key = std::make_tuple(reinterpret_cast<const void*>(kernel), dev, bytes);
  std::lock_guard<std::mutex> lock(mu);
  auto it = done.find(key);
  if (it != done.end()) return it->second;
  if (bytes > 64 * 1024 && hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
    return 1;
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kernel), threads, bytes) !=
          hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 1;
  const int r = std::max(1, per_cu * cus);
  done.emplace(key, r);
  return r;
}

// Blocks of a persistent launch per proof: the resident blocks shared by the
// group's proofs, a multiple of 8 (one share per XCD) where possible, and no
// more than the tiles.
static unsigned persistent_grid(int resident, int proofs, u64 tiles) {
  u64 g = (u64)std::max(1, resident / std::max(1, proofs));
  if (g >= 8) g &= ~u64(7);
  return (unsigned)std::max<u64>(1, std::min(g, tiles));
}

static size_t lds_bytes(int lg, int lines, size_t esize, int tables = 1) {
  size_t tw = tables * (((sizeof(u32) << lg) + 15) & ~size_t(15));
  size_t ls = (1u << lg) + ((1u << lg) >> 5) + 1;
  return tw + (size_t)lines * ls * esize;
}

// Kernel instances by line length: pass 1 for 1 <= LG <= 14 (the four-step
// split of columns up to 2^24, KoalaBear's largest two-adic subgroup; lines of
// 2^14 fit in LDS for base-field columns only), pass 2 for 0 <= LG <= 10.
template <typename T>
static auto pass1_for(int lg) -> void (*)(Group<LdeArgs<T>>, int, int, int, int, Twiddles) {
  switch (lg) {
    case 1: return k_lde_pass1<T, 1>;
    case 2: return k_lde_pass1<T, 2>;
    case 3: return k_lde_pass1<T, 3>;
    case 4: return k_lde_pass1<T, 4>;
    case 5: return k_lde_pass1<T, 5>;
    case 6: return k_lde_pass1<T, 6>;
    case 7: return k_lde_pass1<T, 7>;
    case 8: return k_lde_pass1<T, 8>;
    case 9: return k_lde_pass1<T, 9>;
    case 10: return k_lde_pass1<T, 10>;
    case 11: return k_lde_pass1<T, 11>;
    case 12: return k_lde_pass1<T, 12>;
    case 13: return k_lde_pass1<T, 13>;
    case 14: return k_lde_pass1<T, 14>;
    default: return nullptr;
  }
}

template <typename T, bool DIRECT, bool FULL>
static auto pass2_lg(int lg) -> void (*)(Group<LdeArgs<T>>, int, int, int, Twiddles) {
  switch (lg) {
    case 0: return k_lde_pass2<T, 0, DIRECT, FULL>;
    case 1: return k_lde_pass2<T, 1, DIRECT, FULL>;
    case 2: return k_lde_pass2<T, 2, DIRECT, FULL>;
    case 3: return k_lde_pass2<T, 3, DIRECT, FULL>;
    case 4: return k_lde_pass2<T, 4, DIRECT, FULL>;
    case 5: return k_lde_pass2<T, 5, DIRECT, FULL>;
    case 6: return k_lde_pass2<T, 6, DIRECT, FULL>;
    case 7: return k_lde_pass2<T, 7, DIRECT, FULL>;
    case 8: return k_lde_pass2<T, 8, DIRECT, FULL>;
    case 9: return k_lde_pass2<T, 9, DIRECT, FULL>;
    case 10: return k_lde_pass2<T, 10, DIRECT, FULL>;
    default: return nullptr;
  }
}
// Pass 2 for lines of 2^lg and W = 2^k of them.
template <typename T, bool DIRECT>
static auto pass2_for(int lg, int k) -> void (*)(Group<LdeArgs<T>>, int, int, int, Twiddles) {
  const bool full = (1 << (lg + k)) == Pass2Shape<T>::per_thread * Pass2Shape<T>::threads;
  return full ? pass2_lg<T, DIRECT, true>(lg) : pass2_lg<T, DIRECT, false>(lg);
}

// Launch plan of one LDE: a single pass (l <= LP2) or the four-step split.
struct LdePlan {
  bool two_pass = false;
  int l1 = 0, l2 = 0, tb_log = 0;
  size_t sm1 = 0, sm2 = 0;
};

// Pass 1 takes Pass1Shape::elems elements per block (16K base or 4K
// extension: 64 KiB of lines), at most 256 lines and at most L2; lines of
// 2^12 and longer take fewer (down to one line per block) and up to 160 KiB of
// LDS.  A tuned tb_log (>= 0) replaces that rule, within the same limits (a
// block never holds more than Pass1Shape::elems elements: each thread keeps
// its share in registers).  Returns 0, or a negative code for a shape no
// kernel instance covers.
template <typename T>
static int lde_plan(int m, int k, int R, int LP2, int tb_tuned, LdePlan* pl) {
  const int l = m - k;
  if (l < 0 || k < 0 || R < 0 || k > 10) return -1;
  if (l + R > 24) return -7;   // the codeword domain exceeds KoalaBear's two-adic subgroups
  if (((u64)sizeof(T) << m) >= (1ull << 31)) return -7;   // beyond the buffer loads' 32-bit offsets
  if (LP2 < (sizeof(T) == 4 ? 6 : 2)) return -6;   // rows too wide for a single-pass line in LDS
  const int W = 1 << k;
  constexpr int pass2_max = Pass2Shape<T>::per_thread * Pass2Shape<T>::threads;
  if (l <= LP2) {
    pl->two_pass = false;
    if (W << l > pass2_max) return -2;   // more than a block's registers hold
    pl->sm2 = lds_bytes(l, W, sizeof(T));
    if (pl->sm2 > 160 * 1024) return -2;
    if (!pass2_for<T, true>(l, k)) return -6;
    return 0;
  }
  pl->two_pass = true;
  pl->l2 = LP2;
  pl->l1 = l - LP2;
  const int elem_log = Pass1Shape<T>::elems_log;
  int tb_log = tb_tuned >= 0 ? tb_tuned : elem_log - pl->l1;
  if (tb_log > 8) tb_log = 8;   // TB <= 256 < Pass1Shape::threads: the row step NT / TB is at least 2
  if (tb_log > pl->l2) tb_log = pl->l2;
  if (tb_log < 0) tb_log = 0;
  if (tb_tuned < 0)
    while (tb_log > 0 && lds_bytes(pl->l1, 1 << tb_log, sizeof(T), pl->l1 <= 12 ? 2 : 1) > 80 * 1024) tb_log--;
  if (pl->l1 + tb_log > elem_log) return -3;   // more than a block's registers hold
  pl->tb_log = tb_log;
  pl->sm1 = lds_bytes(pl->l1, 1 << tb_log, sizeof(T), pl->l1 <= 12 ? 2 : 1);
  if (pl->sm1 > 160 * 1024) return -3;
  if (W << pl->l2 > pass2_max) return -2;
  pl->sm2 = lds_bytes(pl->l2, W, sizeof(T));
  if (pl->sm2 > 160 * 1024) return -2;
  if (!pass1_for<T>(pl->l1) || !pass2_for<T, false>(pl->l2, k)) return -6;
  return 0;
}

template <typename T>
static int lde_impl(const LdeArgs<T>* io, int count, int m, int k, int R, int LP2, int tb_tuned, const Twiddles& tw,
                    hipStream_t st) {
  LdePlan pl;
  const int rc = lde_plan<T>(m, k, R, LP2, tb_tuned, &pl);
  if (rc) return rc;
  if (count < 1 || count > GROUP_MAX) return -8;
  const int l = m - k;
  const unsigned np = (unsigned)count;
  if (!pl.two_pass) {
    auto kern = pass2_for<T, true>(l, k);
    const unsigned G = persistent_grid(prepare_kernel(kern, Pass2Shape<T>::threads, pl.sm2), count, 1ull << R);
    return launch_group(io, count, [&](const Group<LdeArgs<T>>& g) {
      hipLaunchKernelGGL(kern, dim3(G, 1, np), dim3(Pass2Shape<T>::threads), pl.sm2, st, g, l, R, k, tw);
    });
  }
  for (int i = 0; i < count; i++)
    if (io[i].tmp == nullptr) return -4;
  auto k1 = pass1_for<T>(pl.l1);
  auto k2 = pass2_for<T, false>(pl.l2, k);
  const unsigned G1 = persistent_grid(prepare_kernel(k1, Pass1Shape<T>::threads, pl.sm1), count,
                                      ((1ull << (pl.l2 - pl.tb_log)) << k) << (R - std::min(R, 1)));
  const unsigned G2 = persistent_grid(prepare_kernel(k2, Pass2Shape<T>::threads, pl.sm2), count, (1ull << pl.l1) << R);
  return launch_group(io, count, [&](const Group<LdeArgs<T>>& g) {
    hipLaunchKernelGGL(k1, dim3(G1, 1, np), dim3(Pass1Shape<T>::threads), pl.sm1, st, g, l, R, k, pl.tb_log, tw);
    hipLaunchKernelGGL(k2, dim3(G2, 1, np), dim3(Pass2Shape<T>::threads), pl.sm2, st, g, l, R, k, tw);
  });
}

// Largest single-pass length per element type: the tuned value, else the
// longest whose W lines take at most 64 KiB of LDS.
static int lp2_for(int k, bool ext, const Tuning& t) {
  if (t.lde_lp2[ext ? 1 : 0] > 0) return t.lde_lp2[ext ? 1 : 0];
  int budget_log = 16 - (ext ? 4 : 2) - k;   // log2(64K / esize / W)
  int lp = budget_log < 10 ? budget_log : 10;
  return lp < 1 ? 1 : lp;
}

bool lde_two_pass(int m, int k, bool ext, const Tuning& t) { return m - k > lp2_for(k, ext, t); }

int lde_check(int m, int k, int R, bool ext, const Tuning& t) {
  LdePlan pl;
  const int lp = lp2_for(k, ext, t);
  const int tb = t.lde_tb_log[ext ? 1 : 0];
  return ext ? lde_plan<ef>(m, k, R, lp, tb, &pl) : lde_plan<u32>(m, k, R, lp, tb, &pl);
}

int lde_base(const LdeArgs<u32>* a, int count, int m, int k, int R, const Twiddles& tw, const Tuning& t,
             hipStream_t st) {
  return lde_impl<u32>(a, count, m, k, R, lp2_for(k, false, t), t.lde_tb_log[0], tw, st);
}

int lde_ext(const LdeArgs<ef>* a, int count, int m, int k, int R, const Twiddles& tw, const Tuning& t,
            hipStream_t st) {
  return lde_impl<ef>(a, count, m, k, R, lp2_for(k, true, t), t.lde_tb_log[1], tw, st);
}

bool tuning_valid(const Tuning& t) {
  for (int e = 0; e < 2; e++) {
    if (t.lde_lp2[e] < 0 || t.lde_lp2[e] > 10) return false;
    if (t.lde_tb_log[e] < -1 || t.lde_tb_log[e] > 8) return false;
  }
  return t.tree_lane_log >= 4 && t.tree_lane_log <= 30 && t.tree_s128_log >= 7 && t.tree_s128_log <= 30 &&
         t.grind_wide_bits >= 0 && t.grind_wide_bits <= 32 && t.grind_lock_bits >= 1 && t.grind_lock_bits <= 33 &&
         t.grind_blocks >= 0 && t.grind_blocks <= (1 << 16);
}

}  // namespace kb
