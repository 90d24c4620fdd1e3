// Multilinear kernels on evaluation lists (index bit n-1-j <-> variable j):
//   eq tables            EvaluationsList::new_from_point (src/poly/evals.rs:312-320)
//   eq-weighted sums     eval_multilinear_{base,ext} (evals.rs:648-738),
//                        SplitEqInner::partial_evals (src/sumcheck/svo.rs:465-502)
//   weight accumulation  EqStatement::combine_hypercube_packed (statement/eq.rs:248-312),
//                        SplitEq::combine_into_packed (svo.rs:504-559),
//                        SelectStatement::combine_packed (statement/select.rs:427-496)
//   multi-variable fold  compress_multi_into_packed (evals.rs:422-447)
//   sumcheck round       sumcheck_coefficients (evals.rs:493-526) and compress
//                        (evals.rs:127-148), fused as fold(r) -> next (h(0), h(2))
//
// eq(z, x) is always split as Ehi[x >> lo] * Elo[x & (2^lo - 1)], where Ehi and
// Elo are the eq tables of the leading and trailing coordinates of z; both are
// small and built on the device from the point, which arrives by value.
#include "kcommon.h"
#include "kernels.h"
#include "transcript.h"

namespace kb {

// ---------------------------------------------------------------- eq tables
// out[x] = scale * prod_j (x_j ? z_j : 1 - z_j), x < 2^t, z_0 on the MSB.
__global__ void k_eq_small(PointArg z, int first, int t, ef scale, ef* __restrict__ out) {
  u32 x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (1u << t)) return;
  ef acc = scale;
  for (int j = 0; j < t; j++) {
    const ef& zj = z.z[first + j];
    int bit = (x >> (t - 1 - j)) & 1;
    acc = ef_mul(acc, bit ? zj : ef_sub(ef_one(), zj));
  }
  out[x] = acc;
}

// Grid: (ceil((2^hi + 2^lo) / 256), 1, proofs).  The point's coordinates are
// staged in LDS first (they may be in host-visible memory).
__global__ void k_eq_split(Group<EqSplitArgs> g, int hi, int lo) {
  __shared__ ef zs[32];
  const EqSplitArgs& a = g.v[blockIdx.z];
  if (threadIdx.x < (unsigned)(hi + lo)) zs[threadIdx.x] = a.z[threadIdx.x];
  __syncthreads();
  u32 x = blockIdx.x * blockDim.x + threadIdx.x;
  const u32 nh = 1u << hi;
  if (x >= nh + (1u << lo)) return;
  int first = 0, t = hi;
  ef acc = a.scale;
  ef* out = a.out_hi;
  if (x >= nh) {
    x -= nh;
    first = hi;
    t = lo;
    acc = ef_one();
    out = a.out_lo;
  }
  for (int j = 0; j < t; j++) {
    const ef& zj = zs[first + j];
    int bit = (x >> (t - 1 - j)) & 1;
    acc = ef_mul(acc, bit ? zj : ef_sub(ef_one(), zj));
  }
  out[x] = acc;
}

int eq_split(const EqSplitArgs* a, int count, int hi, int lo, hipStream_t st) {
  if (hi < 0 || lo < 0 || hi + lo > 32) return -1;
  const u32 n = (1u << hi) + (1u << lo);
  return launch_group(a, count, [&](const Group<EqSplitArgs>& g) {
    hipLaunchKernelGGL(k_eq_split, dim3((n + 255) / 256, 1, count), dim3(256), 0, st, g, hi, lo);
  });
}

int eq_table(const PointArg& z, int first, int t, const ef& scale, ef* out, hipStream_t st) {
  u32 n = 1u << t;
  hipLaunchKernelGGL(k_eq_small, dim3((n + 255) / 256), dim3(256), 0, st, z, first, t, scale, out);
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

// Full table out[x] = Ehi[x >> lo] * Elo[x & mask] for x < 2^(hi+lo).  The
// table is written once and is larger than the caches: nontemporal stores
// (measured 1.15x at 2^26 entries).
__global__ void k_eq_outer(const ef* __restrict__ Ehi, const ef* __restrict__ Elo, int lo, u64 n, ef* __restrict__ out) {
  for (u64 x = (u64)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (u64)gridDim.x * blockDim.x) {
    const ef v = ef_mul(Ehi[x >> lo], Elo[x & ((1u << lo) - 1)]);
    const u32x4 q = {v.c[0], v.c[1], v.c[2], v.c[3]};
    __builtin_nontemporal_store(q, reinterpret_cast<u32x4*>(out + x));
  }
}

int eq_outer(const ef* Ehi, const ef* Elo, int lo, u64 n, ef* out, hipStream_t st) {
  u64 blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_eq_outer, dim3((unsigned)blocks), dim3(256), 0, st, Ehi, Elo, lo, n, out);
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

// ---------------------------------------------------------------- eq-weighted sums
// Segmented dot product of f with K split eq tables.
// f has 2^nf entries in 2^sb segments of 2^(nf - sb); within a segment,
// x = xh * 2^lo + xl and the weight of table k is Ehi[k][xh] * Elo[k][xl].
// Thread t owns xl = t mod 2^lo and rows xh = (t >> lo) + i * (256 >> lo).
// partial[(seg * gridDim.x + block) * K + k] receives the block's sum.
template <typename T, int K>
__global__ void __launch_bounds__(256) k_eq_dot(Group<DotArgs> g, int nf, int sb, int lo) {
  __shared__ ef sh[16 * K];
  const DotArgs& a = g.v[blockIdx.z];
  const T* __restrict__ f = static_cast<const T*>(a.f);
  const EqTables& tabs = a.tabs;
  ef* __restrict__ partial = a.partial;
  const int t = threadIdx.x;
  const int seg = blockIdx.y;
  const int nrest = nf - sb;
  const u64 nrows = 1ull << (nrest - lo);
  const int rows_per_iter = 256 >> lo;
  const u32 xl = t & ((1u << lo) - 1);
  const u64 rows_per_block = (nrows + gridDim.x - 1) / gridDim.x;
  const u64 r0 = (u64)blockIdx.x * rows_per_block;
  u64 r1 = r0 + rows_per_block;
  if (r1 > nrows) r1 = nrows;
  const T* fs = f + ((u64)seg << nrest);
  ef acc[K];
#pragma unroll
  for (int k = 0; k < K; k++) acc[k] = ef_zero();
  for (u64 r = r0 + (t >> lo); r < r1; r += rows_per_iter) {
    T v = fs[(r << lo) + xl];
#pragma unroll
    for (int k = 0; k < K; k++) {
      ef e = tabs.hi[k][r];
      if constexpr (sizeof(T) == 4) acc[k] = ef_add(acc[k], ef_mul_base(e, v));
      else acc[k] = ef_add(acc[k], ef_mul(e, v));
    }
  }
#pragma unroll
  for (int k = 0; k < K; k++) acc[k] = ef_mul(acc[k], tabs.lo[k][xl]);
  block_reduce_ef<K>(acc, sh);
  if (t == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) partial[((u64)seg * gridDim.x + blockIdx.x) * K + k] = acc[k];
  }
}

// Sums of Montgomery products are kept unreduced in 64 bits: fold32 maps a
// sum t < 2^64 to (t >> 32) (2^25 - 2) + (t mod 2^32) < 1.45e17 (2^32 = 2^25 - 2
// mod p), and a folded sum plus up to four products (each below p^2 < 4.54e18)
// stays below 2^64.  One Montgomery reduction ends a sum (the reduction of a
// sum of products is the Montgomery form of the sum).
__device__ __forceinline__ u64 fold32(u64 t) {
  return ((u64)(u32)(t >> 32)) * 33554430u + (u32)t;   // (t >> 32) (2^32 mod p) + (t mod 2^32)
}

// Base-field form of k_eq_dot for lo >= 2: a thread reads four consecutive
// entries of a row (one 16-byte load) and keeps, per entry and coefficient, an
// unreduced 64-bit sum of hi[k][r] * f products, folded every four rows and
// reduced once at the end.
template <int K>
__global__ void __launch_bounds__(256) k_eq_dot4(Group<DotArgs> g, int nf, int sb, int lo) {
  __shared__ ef sh[16 * K];
  const DotArgs& a = g.v[blockIdx.z];
  const u32* __restrict__ f = static_cast<const u32*>(a.f);
  const EqTables& tabs = a.tabs;
  ef* __restrict__ partial = a.partial;
  const int t = threadIdx.x;
  const int seg = blockIdx.y;
  const int nrest = nf - sb;
  const u64 nrows = 1ull << (nrest - lo);
  const int lanes = 1 << (lo - 2);   // threads per row
  const int rows_per_iter = 256 >> (lo - 2);
  const int xq = t & (lanes - 1);    // entries 4 xq .. 4 xq + 3 of each row
  const u32* fs = f + ((u64)seg << nrest) + 4 * xq;
  ef acc[K][4];
  u64 lz[K][4][4];
#pragma unroll
  for (int k = 0; k < K; k++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      acc[k][i] = ef_zero();
#pragma unroll
      for (int c = 0; c < 4; c++) lz[k][i][c] = 0;
    }
  auto fold = [&]() {
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int c = 0; c < 4; c++) lz[k][i][c] = fold32(lz[k][i][c]);
  };
  // Tiles of four steps' rows (4 rows_per_iter), block b taking tiles b,
  // b + gridDim.x, ...: the grid sweeps the segment front to back, which DRAM
  // serves far faster than one separate stream per block (measured 2.3x).  All
  // of a tile's loads are issued before the arithmetic.
  // The evaluations are streamed once: nontemporal loads keep them from
  // displacing reusable lines in L2 and the MALL (measured 2.1x at 2^26).
  const u64 tile = 4ull * rows_per_iter;
  const u64 ntiles = nrows / tile;
  const u64 wrow = (u64)(t >> (lo - 2));
  for (u64 ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
    const u64 r = ti * tile + wrow;
    u32x4 v[4];
    ef e[K][4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const u64 rj = r + (u64)j * rows_per_iter;
      v[j] = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(fs + (rj << lo)));
#pragma unroll
      for (int k = 0; k < K; k++) e[k][j] = tabs.hi[k][rj];
    }
    // One wait for all eight loads: the empty asm takes the loaded registers,
    // so every load is issued before the first is waited for.
#pragma unroll
    for (int j = 0; j < 4; j++) {
      asm volatile("" : "+v"(v[j]));
#pragma unroll
      for (int k = 0; k < K; k++) asm volatile("" : "+v"(e[k][j].c[0]), "+v"(e[k][j].c[1]), "+v"(e[k][j].c[2]), "+v"(e[k][j].c[3]));
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const u32 vv[4] = {v[j].x, v[j].y, v[j].z, v[j].w};
#pragma unroll
      for (int k = 0; k < K; k++)
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int c = 0; c < 4; c++) lz[k][i][c] += (u64)e[k][j].c[c] * vv[i];
    }
    fold();
  }
  // rows past the last whole tile: at most four per thread
  for (u64 r = ntiles * tile + (u64)blockIdx.x * rows_per_iter + wrow; r < nrows;
       r += (u64)gridDim.x * rows_per_iter) {
    const uint4 v = *reinterpret_cast<const uint4*>(fs + (r << lo));
    const u32 vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < K; k++) {
      const ef e = tabs.hi[k][r];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int c = 0; c < 4; c++) lz[k][i][c] += (u64)e.c[c] * vv[i];
    }
  }
  fold();
  // the products carry the Montgomery factor of both operands: one redc per
  // sum, as ef_mul_base does per product
#pragma unroll
  for (int k = 0; k < K; k++)
#pragma unroll
    for (int i = 0; i < 4; i++)
      acc[k][i] = ef{{redc(lz[k][i][0]), redc(lz[k][i][1]), redc(lz[k][i][2]), redc(lz[k][i][3])}};
  ef tot[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    tot[k] = ef_zero();
#pragma unroll
    for (int i = 0; i < 4; i++) tot[k] = ef_add(tot[k], ef_mul(acc[k][i], tabs.lo[k][4 * xq + i]));
  }
  block_reduce_ef<K>(tot, sh);
  if (t == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) partial[((u64)seg * gridDim.x + blockIdx.x) * K + k] = tot[k];
  }
}

// out[seg * K + k] = sum_b partial[(seg * nb + b) * K + k]; one block per (seg, k) and proof.
struct RedArgs {
  const ef* partial;
  ef* out;
};
__global__ void __launch_bounds__(256) k_reduce_strided(Group<RedArgs> g, int nb, int K) {
  __shared__ ef sh[16];
  const ef* __restrict__ partial = g.v[blockIdx.z].partial;
  ef* __restrict__ out = g.v[blockIdx.z].out;
  const int seg = blockIdx.x / K, k = blockIdx.x % K;
  ef acc[1] = {ef_zero()};
  for (int b = threadIdx.x; b < nb; b += 256) acc[0] = ef_add(acc[0], partial[((u64)seg * nb + b) * K + k]);
  block_reduce_ef<1>(acc, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = acc[0];
}

template <typename T>
static int eq_dot_impl(const DotArgs* a, int count, int nf, int sb, int lo, int K, int nblocks, hipStream_t st) {
  if (count < 1 || count > GROUP_MAX || K < 1 || K > 4) return -1;
  Group<DotArgs> g;
  for (int i = 0; i < count; i++) g.v[i] = a[i];
  const dim3 grid(nblocks, 1u << sb, count);
  bool four = false;
  if constexpr (sizeof(T) == 4) four = lo >= 2 && K <= 2;
  if (four) {
    if (K == 1) hipLaunchKernelGGL((k_eq_dot4<1>), grid, dim3(256), 0, st, g, nf, sb, lo);
    else hipLaunchKernelGGL((k_eq_dot4<2>), grid, dim3(256), 0, st, g, nf, sb, lo);
  } else {
    switch (K) {
      case 1: hipLaunchKernelGGL((k_eq_dot<T, 1>), grid, dim3(256), 0, st, g, nf, sb, lo); break;
      case 2: hipLaunchKernelGGL((k_eq_dot<T, 2>), grid, dim3(256), 0, st, g, nf, sb, lo); break;
      case 3: hipLaunchKernelGGL((k_eq_dot<T, 3>), grid, dim3(256), 0, st, g, nf, sb, lo); break;
      default: hipLaunchKernelGGL((k_eq_dot<T, 4>), grid, dim3(256), 0, st, g, nf, sb, lo); break;
    }
  }
  // partial is [seg][block][k]; reduce over blocks for each (seg, k)
  bool reduce = false;
  for (int i = 0; i < count; i++) reduce = reduce || a[i].out;
  if (reduce) {
    for (int i = 0; i < count; i++)
      if (!a[i].out) return -1;
    Group<RedArgs> r;
    for (int i = 0; i < count; i++) r.v[i] = RedArgs{a[i].partial, a[i].out};
    hipLaunchKernelGGL(k_reduce_strided, dim3((1u << sb) * K, 1, count), dim3(256), 0, st, r, nblocks, K);
  }
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

// Result layout: out[(seg * K) + k].  `partial` needs nblocks * 2^sb * K entries.
int eq_dot_base(const DotArgs* a, int count, int nf, int sb, int lo, int K, int nblocks, hipStream_t st) {
  return eq_dot_impl<u32>(a, count, nf, sb, lo, K, nblocks, st);
}

int eq_dot_ext(const DotArgs* a, int count, int nf, int sb, int lo, int K, int nblocks, hipStream_t st) {
  return eq_dot_impl<ef>(a, count, nf, sb, lo, K, nblocks, st);
}

// ---------------------------------------------------------------- weight accumulation
// W[x] = (acc ? W[x] : 0) + sum_k s_k Ehi[k][x >> lo] * Elo[k][x & mask], with s_k = scale[k] when scale is
// given (device memory) and 1 otherwise.
// Without scales the K products are summed unreduced: per component, each
// term's four 32x32 products (each below p^2) are added to the running sum,
// which fold32 keeps below 2^57 + 2^32 (4 p^2 + 2^58 < 2^64), and one
// Montgomery reduction per component follows (its input is below 4 p^2).
// The result is the canonical sum of the reduced products, as with scales.
__global__ void __launch_bounds__(256) k_eq_combine(Group<CombineArgs> g, u64 n, int lo, int K, int accumulate) {
  const CombineArgs& a = g.v[blockIdx.z];
  ef* __restrict__ Wt = a.Wt;
  const EqTables& tabs = a.tabs;
  const ef* __restrict__ scale = a.scale;
  const u32 mask = (1u << lo) - 1;
  for (u64 x = (u64)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (u64)gridDim.x * blockDim.x) {
    ef acc = accumulate ? Wt[x] : ef_zero();
    if (scale) {
      for (int k = 0; k < K; k++) {
        const ef t = ef_mul(tabs.hi[k][x >> lo], tabs.lo[k][x & mask]);
        acc = ef_add(acc, ef_mul(scale[k], t));
      }
    } else if (K > 0 && lo >= 8) {
      // a block's 256 points share x >> lo: the hi entries are block-uniform
      // (scalar loads), and the factor 3 of the wrapping terms goes on them
      // (three scalar operations per term) instead of on the lo entries
      u64 c0 = 0, c1 = 0, c2 = 0, c3 = 0;
      const u32 xh = __builtin_amdgcn_readfirstlane((u32)(x >> lo));
      for (int k = 0; k < K; k++) {
        const ef a = tabs.hi[k][xh], b = tabs.lo[k][x & mask];
        const u32 a1 = add(dbl(a.c[1]), a.c[1]), a2 = add(dbl(a.c[2]), a.c[2]), a3 = add(dbl(a.c[3]), a.c[3]);
        c0 = fold32(c0 + (u64)a.c[0] * b.c[0] + (u64)a1 * b.c[3] + (u64)a2 * b.c[2] + (u64)a3 * b.c[1]);
        c1 = fold32(c1 + (u64)a.c[0] * b.c[1] + (u64)a.c[1] * b.c[0] + (u64)a2 * b.c[3] + (u64)a3 * b.c[2]);
        c2 = fold32(c2 + (u64)a.c[0] * b.c[2] + (u64)a.c[1] * b.c[1] + (u64)a.c[2] * b.c[0] + (u64)a3 * b.c[3]);
        c3 = fold32(c3 + (u64)a.c[0] * b.c[3] + (u64)a.c[1] * b.c[2] + (u64)a.c[2] * b.c[1] + (u64)a.c[3] * b.c[0]);
      }
      acc = ef_add(acc, ef{{redc_wide(c0), redc_wide(c1), redc_wide(c2), redc_wide(c3)}});
    } else if (K > 0) {
      u64 c0 = 0, c1 = 0, c2 = 0, c3 = 0;
      for (int k = 0; k < K; k++) {
        const ef a = tabs.hi[k][x >> lo], b = tabs.lo[k][x & mask];
        const u32 b1 = add(dbl(b.c[1]), b.c[1]), b2 = add(dbl(b.c[2]), b.c[2]), b3 = add(dbl(b.c[3]), b.c[3]);
        c0 = fold32(c0 + (u64)a.c[0] * b.c[0] + (u64)a.c[1] * b3 + (u64)a.c[2] * b2 + (u64)a.c[3] * b1);
        c1 = fold32(c1 + (u64)a.c[0] * b.c[1] + (u64)a.c[1] * b.c[0] + (u64)a.c[2] * b3 + (u64)a.c[3] * b2);
        c2 = fold32(c2 + (u64)a.c[0] * b.c[2] + (u64)a.c[1] * b.c[1] + (u64)a.c[2] * b.c[0] + (u64)a.c[3] * b3);
        c3 = fold32(c3 + (u64)a.c[0] * b.c[3] + (u64)a.c[1] * b.c[2] + (u64)a.c[2] * b.c[1] + (u64)a.c[3] * b.c[0]);
      }
      acc = ef_add(acc, ef{{redc_wide(c0), redc_wide(c1), redc_wide(c2), redc_wide(c3)}});
    }
    Wt[x] = acc;
  }
}

int eq_combine(const CombineArgs* a, int count, u64 n, int lo, int K, int accumulate, hipStream_t st) {
  u64 blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  return launch_group(a, count, [&](const Group<CombineArgs>& g) {
    hipLaunchKernelGGL(k_eq_combine, dim3((unsigned)blocks, 1, count), dim3(256), 0, st, g, n, lo, K, accumulate);
  });
}

// Select tables for z_j in F and weight g_j in EF:
//   A[j][xh] = g_j * (z_j^(2^lo))^xh  (EF),   B[j][xl] = z_j^xl  (F)
__global__ void k_select_tables(Group<SelectArgs> sg, int nsel, int hi, int lo) {
  const SelectArgs& a = sg.v[blockIdx.z];
  const u32* __restrict__ z = a.z;
  const ef* __restrict__ g = a.g;
  ef* __restrict__ A = a.A;
  u32* __restrict__ B = a.B;
  if (a.nsel_dev) nsel = (int)*a.nsel_dev;
  const u32 nh = 1u << hi, nl = 1u << lo;
  u64 total = (u64)nsel * (nh + nl);
  for (u64 e = (u64)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (u64)gridDim.x * blockDim.x) {
    u32 j = (u32)(e / (nh + nl)), r = (u32)(e % (nh + nl));
    if (r < nh) {
      u32 zz = pow(z[j], (u64)r << lo);
      A[(u64)j * nh + r] = ef_mul_base(g[j], zz);
    } else {
      B[(u64)j * nl + (r - nh)] = pow(z[j], r - nh);
    }
  }
}

// W[x] += sum_j A[j][x >> lo] * B[j][x & mask]
// Wt[x] += sum_j A[j][x >> lo] * B[j][x & mask], each EF component's
// products summed unreduced (fold32) three at a time.

__global__ void __launch_bounds__(256) k_select_combine(Group<SelectArgs> sg, u64 n, int lo, int hi, int nsel) {
  const SelectArgs& sa = sg.v[blockIdx.z];
  ef* __restrict__ Wt = sa.Wt;
  const ef* __restrict__ A = sa.A;
  const u32* __restrict__ B = sa.B;
  if (sa.nsel_dev) nsel = (int)*sa.nsel_dev;
  const u32 mask = (1u << lo) - 1;
  const u64 nh = 1ull << hi, nl = 1ull << lo;
  for (u64 x = (u64)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (u64)gridDim.x * blockDim.x) {
    const u64 xh = x >> lo, xl = x & mask;
    u64 acc[4] = {0, 0, 0, 0};
    int j = 0;
    for (; j + 3 <= nsel; j += 3) {
      const u32 b0 = B[(u64)j * nl + xl], b1 = B[(u64)(j + 1) * nl + xl], b2 = B[(u64)(j + 2) * nl + xl];
      const ef a0 = A[(u64)j * nh + xh], a1 = A[(u64)(j + 1) * nh + xh], a2 = A[(u64)(j + 2) * nh + xh];
#pragma unroll
      for (int k = 0; k < 4; k++)
        acc[k] = fold32((u64)a2.c[k] * b2 + ((u64)a1.c[k] * b1 + ((u64)a0.c[k] * b0 + acc[k])));
    }
    for (; j < nsel; j++) {
      const u32 b0 = B[(u64)j * nl + xl];
      const ef a0 = A[(u64)j * nh + xh];
#pragma unroll
      for (int k = 0; k < 4; k++) acc[k] = fold32((u64)a0.c[k] * b0 + acc[k]);
    }
    ef r;
#pragma unroll
    for (int k = 0; k < 4; k++) r.c[k] = redc(acc[k]);
    Wt[x] = ef_add(Wt[x], r);
  }
}

// Same for lo >= 8 with XT points per thread, 256 apart in one row: a
// block's 256 * XT points (2^lo a multiple of that; the grid stride too)
// share x >> lo, so the A entries are wave-uniform and come through scalar
// loads, each used XT times, and a thread has 3 * XT loads of B in flight per
// step.
template <int XT>
__global__ void __launch_bounds__(256) k_select_combine_rows(Group<SelectArgs> sg, u64 n, int lo, int hi, int nsel) {
  const SelectArgs& sa = sg.v[blockIdx.z];
  ef* __restrict__ Wt = sa.Wt;
  const ef* __restrict__ A = sa.A;
  const u32* __restrict__ B = sa.B;
  if (sa.nsel_dev) nsel = (int)*sa.nsel_dev;
  const u32 mask = (1u << lo) - 1;
  const u64 nh = 1ull << hi, nl = 1ull << lo;
  for (u64 x0 = (u64)blockIdx.x * (256 * XT); x0 < n; x0 += (u64)gridDim.x * (256 * XT)) {
    const u32 xh = __builtin_amdgcn_readfirstlane((u32)(x0 >> lo));
    const u32 xl = ((u32)x0 & mask) + threadIdx.x;
    const ef* __restrict__ Ah = A + xh;
    const u32* __restrict__ Bl = B + xl;
    u64 acc[XT][4] = {};
    int j = 0;
    for (; j + 3 <= nsel; j += 3) {
      u32 b[3][XT];
#pragma unroll
      for (int q = 0; q < 3; q++)
#pragma unroll
        for (int u = 0; u < XT; u++) b[q][u] = Bl[(u64)(j + q) * nl + 256 * u];
      const ef a0 = Ah[(u64)j * nh], a1 = Ah[(u64)(j + 1) * nh], a2 = Ah[(u64)(j + 2) * nh];
#pragma unroll
      for (int u = 0; u < XT; u++)
#pragma unroll
        for (int k = 0; k < 4; k++)
          acc[u][k] = fold32((u64)a2.c[k] * b[2][u] + ((u64)a1.c[k] * b[1][u] + ((u64)a0.c[k] * b[0][u] + acc[u][k])));
    }
    for (; j < nsel; j++) {
      const ef a0 = Ah[(u64)j * nh];
#pragma unroll
      for (int u = 0; u < XT; u++) {
        const u32 b0 = Bl[(u64)j * nl + 256 * u];
#pragma unroll
        for (int k = 0; k < 4; k++) acc[u][k] = fold32((u64)a0.c[k] * b0 + acc[u][k]);
      }
    }
#pragma unroll
    for (int u = 0; u < XT; u++) {
      ef r;
#pragma unroll
      for (int k = 0; k < 4; k++) r.c[k] = redc(acc[u][k]);
      const u64 x = x0 + threadIdx.x + 256 * u;
      Wt[x] = ef_add(Wt[x], r);
    }
  }
}

int select_combine(const SelectArgs* a, int count, int m, int nsel, hipStream_t st) {
  if (nsel == 0) return 0;
  const int lo = m / 2, hi = m - lo;
  const u64 tt = (u64)nsel * ((1ull << hi) + (1ull << lo));
  u64 b1 = (tt + 255) / 256;
  if (b1 > 65536) b1 = 65536;
  const u64 n = 1ull << m;
  u64 blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  return launch_group(a, count, [&](const Group<SelectArgs>& g) {
    hipLaunchKernelGGL(k_select_tables, dim3((unsigned)b1, 1, count), dim3(256), 0, st, g, nsel, hi, lo);
    if (lo >= 10)
      hipLaunchKernelGGL(k_select_combine_rows<4>, dim3((unsigned)((blocks + 3) / 4), 1, count), dim3(256), 0, st, g,
                         n, lo, hi, nsel);
    else if (lo >= 8)
      hipLaunchKernelGGL(k_select_combine_rows<1>, dim3((unsigned)blocks, 1, count), dim3(256), 0, st, g, n, lo, hi,
                         nsel);
    else
      hipLaunchKernelGGL(k_select_combine, dim3((unsigned)blocks, 1, count), dim3(256), 0, st, g, n, lo, hi, nsel);
  });
}

// ---------------------------------------------------------------- folds
// out[x] = sum_s eqrs[s] * f[s * 2^nr + x], s < 2^l
// (products summed unreduced per component, folded every four).  f is read
// once: nontemporal loads.
__global__ void __launch_bounds__(256) k_compress_multi(Group<CompressArgs> g, int nr, int l) {
  const u32* __restrict__ f = g.v[blockIdx.z].f;
  const ef* __restrict__ eqrs = g.v[blockIdx.z].eqrs;
  ef* __restrict__ out = g.v[blockIdx.z].out;
  const u64 n = 1ull << nr;
  const int ns = 1 << l;
  for (u64 x = (u64)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (u64)gridDim.x * blockDim.x) {
    u64 lz[4] = {0, 0, 0, 0};
    for (int s0 = 0; s0 < ns; s0 += 4) {
      const int cnt = ns - s0 < 4 ? ns - s0 : 4;
      u32 v[4];
#pragma unroll
      for (int q = 0; q < 4; q++) v[q] = q < cnt ? __builtin_nontemporal_load(&f[((u64)(s0 + q) << nr) + x]) : 0u;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        if (q < cnt) {
          const ef e = eqrs[s0 + q];
#pragma unroll
          for (int c = 0; c < 4; c++) lz[c] += (u64)e.c[c] * v[q];
        }
      }
#pragma unroll
      for (int c = 0; c < 4; c++) lz[c] = fold32(lz[c]);
    }
    out[x] = ef{{redc(lz[0]), redc(lz[1]), redc(lz[2]), redc(lz[3])}};
  }
}

int compress_multi(const CompressArgs* a, int count, int n, int l, hipStream_t st) {
  u64 m = 1ull << (n - l);
  u64 blocks = (m + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  return launch_group(a, count, [&](const Group<CompressArgs>& g) {
    hipLaunchKernelGGL(k_compress_multi, dim3((unsigned)blocks, 1, count), dim3(256), 0, st, g, n - l, l);
  });
}

__global__ void k_base_to_ext(Group<ExtArgs> g, u64 n) {
  const u32* __restrict__ f = g.v[blockIdx.z].f;
  ef* __restrict__ out = g.v[blockIdx.z].out;
  for (u64 x = (u64)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (u64)gridDim.x * blockDim.x)
    out[x] = ef_base(f[x]);
}

int base_to_ext(const ExtArgs* a, int count, u64 n, hipStream_t st) {
  u64 blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  return launch_group(a, count, [&](const Group<ExtArgs>& g) {
    hipLaunchKernelGGL(k_base_to_ext, dim3((unsigned)blocks, 1, count), dim3(256), 0, st, g, n);
  });
}

// ---------------------------------------------------------------- sumcheck rounds
// (h(0), h(2)) partial sums over pairs (i, i + h) of e and w (length 2h).
__global__ void __launch_bounds__(256) k_sc_coeffs(Group<ScArgs> g, u64 h) {
  __shared__ ef sh[32];
  const ScArgs& a = g.v[blockIdx.z];
  const ef* __restrict__ e = a.e;
  const ef* __restrict__ w = a.w;
  ef* __restrict__ partial = a.partial;
  if (a.grind_reset && blockIdx.x == 0 && threadIdx.x == 0) *a.grind_reset = P;
  ef acc[2] = {ef_zero(), ef_zero()};
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < h; i += (u64)gridDim.x * blockDim.x) {
    ef e0 = e[i], e1 = e[i + h], w0 = w[i], w1 = w[i + h];
    acc[0] = ef_add(acc[0], ef_mul(w0, e0));
    acc[1] = ef_add(acc[1], ef_mul(ef_sub(ef_dbl(w1), w0), ef_sub(ef_dbl(e1), e0)));
  }
  block_reduce_ef<2>(acc, sh);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = acc[0];
    partial[2 * blockIdx.x + 1] = acc[1];
  }
}

// Fold e and w (length 2h) with r into their first halves, and accumulate the
// next round's (h(0), h(2)) over the folded halves (pairs (j, j + h/2)).
// With a transcript mode, every block first runs that transcript step to get
// r; a thread's first eight entries are loaded before it (they do not depend
// on r), so their latency overlaps the step's own loads and permutation.
__global__ void __launch_bounds__(256) k_sc_fold_coeffs(Group<ScArgs> g, u64 h, int tr_mode,
                                                        const P2C16* __restrict__ cp) {
  __shared__ ef sh[32];
  const ScArgs& a = g.v[blockIdx.z];
  ef* __restrict__ e = a.e;
  ef* __restrict__ w = a.w;
  const u64 q = h >> 1;
  const u64 j0 = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const u64 stride = (u64)gridDim.x * blockDim.x;
  ef a0, a1, b0, b1, c0, c1, d0, d1;
  if (j0 < q) {
    a0 = e[j0], a1 = e[j0 + h], b0 = e[j0 + q], b1 = e[j0 + q + h];
    c0 = w[j0], c1 = w[j0 + h], d0 = w[j0 + q], d1 = w[j0 + q + h];
  }
  ef r = a.r;
  if (tr_mode != TR_NONE) {
    __shared__ ef rs;
    tr_prologue_wave0(tr_mode, a.tr, *cp, &rs);
    r = rs;
  }
  ef acc[2] = {ef_zero(), ef_zero()};
  for (u64 j = j0; j < q; j += stride) {
    if (j != j0) {
      a0 = e[j], a1 = e[j + h], b0 = e[j + q], b1 = e[j + q + h];
      c0 = w[j], c1 = w[j + h], d0 = w[j + q], d1 = w[j + q + h];
    }
    const ef f0 = ef_add(a0, ef_mul(r, ef_sub(a1, a0)));
    const ef f1 = ef_add(b0, ef_mul(r, ef_sub(b1, b0)));
    const ef g0 = ef_add(c0, ef_mul(r, ef_sub(c1, c0)));
    const ef g1 = ef_add(d0, ef_mul(r, ef_sub(d1, d0)));
    e[j] = f0;
    e[j + q] = f1;
    w[j] = g0;
    w[j + q] = g1;
    acc[0] = ef_add(acc[0], ef_mul(g0, f0));
    acc[1] = ef_add(acc[1], ef_mul(ef_sub(ef_dbl(g1), g0), ef_sub(ef_dbl(f1), f0)));
  }
  block_reduce_ef<2>(acc, sh);
  if (threadIdx.x == 0) {
    a.partial[2 * blockIdx.x] = acc[0];
    a.partial[2 * blockIdx.x + 1] = acc[1];
  }
}

__global__ void __launch_bounds__(256) k_sc_fold(Group<ScArgs> g, u64 h, int tr_mode, const P2C16* __restrict__ cp) {
  const ScArgs& a = g.v[blockIdx.z];
  ef* __restrict__ e = a.e;
  ef* __restrict__ w = a.w;
  const u64 j0 = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const u64 stride = (u64)gridDim.x * blockDim.x;
  ef a0, a1, c0, c1;
  if (j0 < h) a0 = e[j0], a1 = e[j0 + h], c0 = w[j0], c1 = w[j0 + h];
  ef r = a.r;
  if (tr_mode != TR_NONE) {
    __shared__ ef rs;
    tr_prologue_wave0(tr_mode, a.tr, *cp, &rs);
    r = rs;
  }
  for (u64 j = j0; j < h; j += stride) {
    if (j != j0) a0 = e[j], a1 = e[j + h], c0 = w[j], c1 = w[j + h];
    e[j] = ef_add(a0, ef_mul(r, ef_sub(a1, a0)));
    w[j] = ef_add(c0, ef_mul(r, ef_sub(c1, c0)));
  }
}

// Every round of a sumcheck phase without proof of work, on arrays of at most
// 2^SC_SMALL_LOG entries, in one block per proof: e and w are held in LDS, and
// each round's (h(0), h(2)), transcript step (observe both, sample r) and fold
// are separated by block barriers instead of launches.  The log entries, the
// challenger state left in tr.ch_out and the folded arrays written back are
// those of the per-round kernels (k_sc_coeffs, then k_sc_fold_coeffs in mode
// TR_COEFFS_SAMPLE).  Grid: (1, 1, proofs).
__global__ void __launch_bounds__(256) k_sc_small(Group<ScArgs> g, int m0, int rounds, const P2C16* __restrict__ cp) {
  __shared__ ef se[1 << SC_SMALL_LOG], sw[1 << SC_SMALL_LOG];
  __shared__ ef sh[32];
  const ScArgs& a = g.v[blockIdx.z];
  const int t = threadIdx.x;
  const int n = 1 << m0;
  const WideLane wl = wide_lane(*cp, t & 15);
  BlockChallenger ch{0, 0, 0, 0, 0, &wl, cp};
  ch.load(*a.tr.ch_in);
  for (int i = t; i < n; i += 256) {
    se[i] = a.e[i];
    sw[i] = a.w[i];
  }
  __syncthreads();
  for (int i = 0; i < rounds; i++) {
    const int h = n >> (i + 1);
    ef acc[2] = {ef_zero(), ef_zero()};
    for (int j = t; j < h; j += 256) {
      const ef e0 = se[j], e1 = se[j + h], w0 = sw[j], w1 = sw[j + h];
      acc[0] = ef_add(acc[0], ef_mul(w0, e0));
      acc[1] = ef_add(acc[1], ef_mul(ef_sub(ef_dbl(w1), w0), ef_sub(ef_dbl(e1), e0)));
    }
    block_reduce_ef<2>(acc, sh);   // the same sums in every thread
    const u32 v[8] = {acc[0].c[0], acc[0].c[1], acc[0].c[2], acc[0].c[3],
                      acc[1].c[0], acc[1].c[1], acc[1].c[2], acc[1].c[3]};
    ch.observe8(v);
    const ef r = ch.sample_ef();
    if (t == 0) {
      ScLog* L = a.tr.log + i;
      L->c0 = acc[0];
      L->c2 = acc[1];
      L->r = r;
      L->witness = 0;
      L->err = 0;
    }
    // a thread folds the pairs it summed (j < h reads j + h, written by none)
    for (int j = t; j < h; j += 256) {
      const ef e0 = se[j], w0 = sw[j];
      se[j] = ef_add(e0, ef_mul(r, ef_sub(se[j + h], e0)));
      sw[j] = ef_add(w0, ef_mul(r, ef_sub(sw[j + h], w0)));
    }
    __syncthreads();   // folded entries before the next round's reads; sh reads before its next writes
  }
  for (int i = t; i < (n >> rounds); i += 256) {
    a.e[i] = se[i];
    a.w[i] = sw[i];
  }
  ch.store(a.tr.ch_out);
  if (t == 0) {
    if (a.grind_reset) *a.grind_reset = P;
    if (a.tr.grind_next) *a.tr.grind_next = P;
  }
}

// One round with proof of work on small arrays in one block per proof (what
// the fold kernel of the previous round and the transcript prologue of this
// round's grinding kernel do, without their launches): e and w in LDS.
// Grid: (1, 1, proofs).
__global__ void __launch_bounds__(256) k_sc_step(Group<ScArgs> g, int m, int fold_first, const P2C16* __restrict__ cp) {
  __shared__ ef se[1 << SC_SMALL_LOG], sw[1 << SC_SMALL_LOG];
  __shared__ ef sh[32];
  const ScArgs& a = g.v[blockIdx.z];
  const ScTr& tr = a.tr;
  const int t = threadIdx.x;
  int n = 1 << m;
  const WideLane wl = wide_lane(*cp, t & 15);
  BlockChallenger ch{0, 0, 0, 0, 0, &wl, cp};
  ch.load(*tr.ch_in);
  const u32 w = fold_first ? *tr.grind : 0u;
  for (int i = t; i < n; i += 256) {
    se[i] = a.e[i];
    sw[i] = a.w[i];
  }
  __syncthreads();
  if (fold_first) {
    u32 err = 0;
    const u32 wm = w < P ? to_monty(w) : 0u;
    if (w >= P) err = 1;
    ch.observe(wm);
    const u32 mask = tr.pow_bits >= 32 ? 0xffffffffu : ((1u << tr.pow_bits) - 1);
    if ((from_monty(ch.sample()) & mask) != 0) err = 1;
    const ef r = ch.sample_ef();
    if (t == 0) {
      ScLog* L = tr.log - 1;
      L->witness = wm;
      L->r = r;
      L->err = err;
    }
    n >>= 1;
    for (int j = t; j < n; j += 256) {
      const ef e0 = se[j], w0 = sw[j];
      const ef e1 = ef_add(e0, ef_mul(r, ef_sub(se[j + n], e0)));
      const ef w1 = ef_add(w0, ef_mul(r, ef_sub(sw[j + n], w0)));
      se[j] = e1;
      sw[j] = w1;
      a.e[j] = e1;
      a.w[j] = w1;
    }
    __syncthreads();
  }
  const int h = n >> 1;
  ef acc[2] = {ef_zero(), ef_zero()};
  for (int j = t; j < h; j += 256) {
    const ef e0 = se[j], e1 = se[j + h], w0 = sw[j], w1 = sw[j + h];
    acc[0] = ef_add(acc[0], ef_mul(w0, e0));
    acc[1] = ef_add(acc[1], ef_mul(ef_sub(ef_dbl(w1), w0), ef_sub(ef_dbl(e1), e0)));
  }
  block_reduce_ef<2>(acc, sh);   // the same sums in every thread
  if (t == 0) {
    tr.log->c0 = acc[0];
    tr.log->c2 = acc[1];
    tr.log->witness = 0;
  }
  const u32 v[8] = {acc[0].c[0], acc[0].c[1], acc[0].c[2], acc[0].c[3],
                    acc[1].c[0], acc[1].c[1], acc[1].c[2], acc[1].c[3]};
  ch.observe8(v);
  // a candidate is observed into position in_len of this state
  if (t < 16) tr.gbase->s[t] = ch.absorbed();
  if (t == 0) {
    tr.gbase->wpos = (u32)ch.in_len;
    if (tr.grind_next) *tr.grind_next = P;
  }
  ch.store(tr.ch_out);
}

// Step i of the initial sumcheck's split-route rounds (see svo_step).  Round
// i's polynomial (Sumcheck::from_base_evals with the SVO partial sums,
// sumcheck_prover.rs:196-227): with P_j folded by r_<i,
//   h_i(X) = sum_j ap_j eq(z_j[<i], r_<i) eq(z_j[i], X) sum_s' eq(z_j[>i], s') P_j(r_<i, X, s'),
// and P_j(r_<i, X, s') = sum_b eq(r_<i, b) P_j(b, X, s'), so every thread takes
// terms (j, b, s') of the unfolded P and the block sums them.
__global__ void __launch_bounds__(256) k_svo_step(Group<SvoArgs> g, int J, int l0, int i, int pow_bits,
                                                  const P2C16* __restrict__ cp) {
  __shared__ ef sh[32];
  __shared__ ef zs[SVO_MAX_Z], rsh[SVO_DEV_MAX], aps[SVO_MAX_J];
  const SvoArgs& a = g.v[blockIdx.z];
  const ScTr& tr = a.tr;
  const int t = threadIdx.x;
  const WideLane wl = wide_lane(*cp, t & 15);
  BlockChallenger ch{0, 0, 0, 0, 0, &wl, cp};
  ch.load(*tr.ch_in);
  const bool wit = i > 0 && pow_bits > 0;
  const u32 w = wit ? *tr.grind : 0u;
  // the points' coordinates, the challenges so far and the combination
  // powers in LDS: one load latency instead of one per loop step below
  for (int u = t; u < J * l0; u += 256) zs[u] = a.z[u];
  if (t < i) rsh[t] = a.rs[t];
  if (t < J) aps[t] = a.ap[t];
  __syncthreads();
  ef r_last = ef_zero();
  if (wit) {
    u32 err = 0;
    const u32 wm = w < P ? to_monty(w) : 0u;
    if (w >= P) err = 1;
    ch.observe(wm);
    const u32 mask = pow_bits >= 32 ? 0xffffffffu : ((1u << pow_bits) - 1);
    if ((from_monty(ch.sample()) & mask) != 0) err = 1;
    r_last = ch.sample_ef();
    if (t == 0) {
      ScLog* L = tr.log + (i - 1);
      L->witness = wm;
      L->r = r_last;
      L->err = err;
      a.rs[i - 1] = r_last;
    }
  }
  if (i < l0) {
    const int half = 1 << (l0 - 1 - i);
    const int per = half << i;   // terms (b, s') per constraint
    ef acc[2] = {ef_zero(), ef_zero()};
    for (int u = t; u < J * per; u += 256) {
      const int j = u / per, v = u % per;
      const int b = v / half, sp = v % half;
      const ef* z = zs + j * l0;
      ef coef = aps[j];
      for (int k = 0; k < i; k++) {   // eq(z_k, r_k) eq(r_k, b_k), b's top bit for r_0
        const ef rk = (wit && k == i - 1) ? r_last : rsh[k];
        const ef zk = z[k];
        const ef one = ef_one();
        const ef zr = ef_mul(zk, rk);
        const ef eqz = ef_add(ef_sub(ef_sub(one, zk), rk), ef_dbl(zr));   // z r + (1 - z)(1 - r)
        const bool bit = (b >> (i - 1 - k)) & 1;
        coef = ef_mul(coef, ef_mul(eqz, bit ? rk : ef_sub(one, rk)));
      }
      for (int k = 0; k < l0 - 1 - i; k++) {   // eq(z_{i+1+k}, s'_k), s' top bit first
        const ef zk = z[i + 1 + k];
        const bool bit = (sp >> (l0 - 2 - i - k)) & 1;
        coef = ef_mul(coef, bit ? zk : ef_sub(ef_one(), zk));
      }
      const ef zi = z[i];
      const ef e0 = ef_sub(ef_one(), zi);
      const ef e2 = ef_sub(ef_add(ef_dbl(zi), zi), ef_one());
      const ef* Pj = a.P + ((size_t)j << l0);
      const int x0 = (b << (l0 - i)) | sp;
      const ef p0 = Pj[x0], p1 = Pj[x0 | half];
      acc[0] = ef_add(acc[0], ef_mul(ef_mul(coef, e0), p0));
      acc[1] = ef_add(acc[1], ef_mul(ef_mul(coef, e2), ef_sub(ef_dbl(p1), p0)));
    }
    block_reduce_ef<2>(acc, sh);   // the same sums in every thread
    const u32 v[8] = {acc[0].c[0], acc[0].c[1], acc[0].c[2], acc[0].c[3],
                      acc[1].c[0], acc[1].c[1], acc[1].c[2], acc[1].c[3]};
    ch.observe8(v);
    ScLog* L = tr.log + i;
    if (t == 0) {
      L->c0 = acc[0];
      L->c2 = acc[1];
      L->witness = 0;
    }
    if (pow_bits > 0) {
      if (t < 16) tr.gbase->s[t] = ch.absorbed();
      if (t == 0) {
        tr.gbase->wpos = (u32)ch.in_len;
        if (tr.grind_next) *tr.grind_next = P;
      }
    } else {
      const ef r = ch.sample_ef();
      if (t == 0) {
        L->r = r;
        L->err = 0;
        a.rs[i] = r;
      }
    }
  }
  ch.store(tr.ch_out);
}

static unsigned sc_blocks(u64 work) {
  u64 b = (work + 255) / 256;
  if (b > 1024) b = 1024;
  return b ? (unsigned)b : 1u;
}

static_assert(sizeof(Group<ScArgs>) <= 4096, "kernel arguments of a sumcheck group");

// The nb block partials of each proof reduced into its c0c2 (two values).
static void reduce_pairs(const ScArgs* a, int count, unsigned nb, hipStream_t st) {
  Group<RedArgs> r;
  for (int i = 0; i < count; i++) r.v[i] = RedArgs{a[i].partial, a[i].c0c2};
  hipLaunchKernelGGL(k_reduce_strided, dim3(2, 1, count), dim3(256), 0, st, r, (int)nb, 2);
}

// c0c2[0..2] = (h(0), h(2)) for arrays of length 2h.  Uses partial[2 * 1024].
int sc_coeffs(const ScArgs* a, int count, u64 h, hipStream_t st) {
  const unsigned nb = sc_blocks(h);
  return launch_group(a, count, [&](const Group<ScArgs>& g) {
    hipLaunchKernelGGL(k_sc_coeffs, dim3(nb, 1, count), dim3(256), 0, st, g, h);
    reduce_pairs(a, count, nb, st);
  });
}

// Fold arrays of length 2h with r; if h >= 2 also produce the next (h(0), h(2)).
int sc_fold(const ScArgs* a, int count, u64 h, hipStream_t st) {
  return launch_group(a, count, [&](const Group<ScArgs>& g) {
    if (h >= 2) {
      const unsigned nb = sc_blocks(h >> 1);
      hipLaunchKernelGGL(k_sc_fold_coeffs, dim3(nb, 1, count), dim3(256), 0, st, g, h, (int)TR_NONE,
                         (const P2C16*)nullptr);
      reduce_pairs(a, count, nb, st);
    } else {
      hipLaunchKernelGGL(k_sc_fold, dim3(sc_blocks(h), 1, count), dim3(256), 0, st, g, h, (int)TR_NONE,
                         (const P2C16*)nullptr);
    }
  });
}

int sc_partials(const ScArgs* a, int count, u64 h, unsigned* nb, hipStream_t st) {
  *nb = sc_blocks(h);
  const unsigned b = *nb;
  return launch_group(a, count, [&](const Group<ScArgs>& g) {
    hipLaunchKernelGGL(k_sc_coeffs, dim3(b, 1, count), dim3(256), 0, st, g, h);
  });
}

int sc_fold_tr(const ScArgs* a, int count, u64 h, int mode, const P2C16* d_consts, unsigned* nb, hipStream_t st) {
  *nb = h >= 2 ? sc_blocks(h >> 1) : 0;
  const unsigned b = *nb;
  return launch_group(a, count, [&](const Group<ScArgs>& g) {
    if (h >= 2)
      hipLaunchKernelGGL(k_sc_fold_coeffs, dim3(b, 1, count), dim3(256), 0, st, g, h, mode, d_consts);
    else
      hipLaunchKernelGGL(k_sc_fold, dim3(sc_blocks(h), 1, count), dim3(256), 0, st, g, h, mode, d_consts);
  });
}

int sc_small(const ScArgs* a, int count, int m0, int rounds, const P2C16* d_consts, hipStream_t st) {
  if (m0 < 1 || m0 > SC_SMALL_LOG || rounds < 1 || rounds > m0) return -1;
  return launch_group(a, count, [&](const Group<ScArgs>& g) {
    hipLaunchKernelGGL(k_sc_small, dim3(1, 1, count), dim3(256), 0, st, g, m0, rounds, d_consts);
  });
}

int sc_step(const ScArgs* a, int count, int m, int fold_first, const P2C16* d_consts, hipStream_t st) {
  if (m < 1 || m > SC_SMALL_LOG || (fold_first && m < 2)) return -1;
  return launch_group(a, count, [&](const Group<ScArgs>& g) {
    hipLaunchKernelGGL(k_sc_step, dim3(1, 1, count), dim3(256), 0, st, g, m, fold_first, d_consts);
  });
}

int svo_step(const SvoArgs* a, int count, int J, int l0, int i, int pow_bits, const P2C16* d_consts, hipStream_t st) {
  if (J < 0 || J > SVO_MAX_J || J * l0 > SVO_MAX_Z || l0 < 1 || l0 > SVO_DEV_MAX || i < 0 || i > l0) return -1;
  return launch_group(a, count, [&](const Group<SvoArgs>& g) {
    hipLaunchKernelGGL(k_svo_step, dim3(1, 1, count), dim3(256), 0, st, g, J, l0, i, pow_bits, d_consts);
  });
}

}  // namespace kb
