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
// A length-L DFT with L <= 2^LP runs in one pass (k_lde_pass2 in direct mode);
// longer ones use the four-step split L = L1 * L2:
//   pass 1: per (column, coset, b2): DFT over b1 of c[L2 b1 + b2] (twisted),
//           times w_L^(b2 k1), into V[coset][column][k1][b2];
//   pass 2: per (coset, k1): DFT over b2 of V[..][k1][b2] for all W columns at
//           once, written to whole output rows 2^R (k1 + L1 k2) + i0.
// Each block keeps its lines in LDS in bit-reversed order and runs radix-2
// decimation-in-time stages there.
#include <mutex>
#include <set>
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
  static __device__ __forceinline__ int pos(int i) { return i + (i >> 5); }
};

static constexpr int lde_min(int a, int b) { return a < b ? a : b; }

// Stages S0+1 .. S0+R: for a group with base index b (bits S0..S0+R-1 clear)
// the elements are b + (j << S0), j < 2^R; at local stage g the pair (j, j + h),
// h = 2^(g-1), uses w_{2^(S0+g)}^((b mod 2^S0) + (j mod h) 2^S0).
template <typename T, int LG, int R, int S0>
__device__ __forceinline__ void dit_pass(T* sh, int nlines, const u32* stw) {
  using L = LineShape<LG>;
  constexpr int G = 1 << R;
  constexpr int GPL = L::N >> R;
  const int ngroups = nlines * GPL;
  for (int g = threadIdx.x; g < ngroups; g += blockDim.x) {
    const int line = g / GPL;
    const int gl = g % GPL;
    const int base_low = gl & ((1 << S0) - 1);
    const int base = base_low + ((gl >> S0) << (S0 + R));
    T* p = sh + line * L::LS;
    T v[G];
#pragma unroll
    for (int j = 0; j < G; j++) v[j] = p[L::pos(base + (j << S0))];
#pragma unroll
    for (int sg = 1; sg <= R; sg++) {
      const int h = 1 << (sg - 1);
#pragma unroll
      for (int j = 0; j < G; j++) {
        if (j & h) continue;
        const int jl = j & (h - 1);
        T b = v[j + h];
        if (S0 != 0 || jl != 0) b = tmulb(b, stw[(1 << (S0 + sg - 1)) + base_low + (jl << S0)]);
        const T a = v[j];
        v[j] = tadd(a, b);
        v[j + h] = tsub(a, b);
      }
    }
#pragma unroll
    for (int j = 0; j < G; j++) p[L::pos(base + (j << S0))] = v[j];
  }
}

template <typename T, int LG, int RMAX, int S0 = 0>
__device__ __forceinline__ void dit_lines(T* sh, int nlines, const u32* stw) {
  if constexpr (S0 < LG) {
    constexpr int R = lde_min(RMAX, LG - S0);
    dit_pass<T, LG, R, S0>(sh, nlines, stw);
    __syncthreads();
    dit_lines<T, LG, RMAX, S0 + R>(sh, nlines, stw);
  }
}

template <typename T>
struct RadixMax {
  static constexpr int value = sizeof(T) == 4 ? 5 : 4;
};

template <int LG>
__device__ void fill_stage_twiddles(u32* stw, const Twiddles& tw) {
  for (int j = threadIdx.x; j < (1 << LG); j += blockDim.x) {
    if (j == 0) continue;
    const int s = 32 - __clz(j);   // j in [2^(s-1), 2^s)
    stw[j] = root_pow(tw, s, (u64)(j - (1 << (s - 1))));
  }
}

// LDS: the stage twiddles (N words, 16-byte aligned) then nlines padded lines.
template <typename T, int LG>
__device__ __forceinline__ T* line_storage(unsigned char* smem) {
  return reinterpret_cast<T*>(smem + (((sizeof(u32) << LG) + 15) & ~size_t(15)));
}

// Loads per thread in flight in the gather loops.
template <typename T>
struct Batch {
  static constexpr int value = sizeof(T) == 4 ? 8 : 4;
};

// Threads per pass-1 block: a block holds 16K base or 4K extension elements
// (64 KiB of LDS, two blocks per CU) in 512 threads, four waves per SIMD
// (-10% on the 2^24 base LDE, -2% on the extension LDEs, against 256).  A
// block's lines (TB) must not outnumber its threads: the row step is NT / TB.
template <typename T>
struct Pass1Threads {
  static constexpr int value = 512;
  static_assert(value >= 256, "pass 1 blocks take up to 256 lines");
};

// Pass 1 of the four-step split.  Grid: (L2 / TB, W, 2^R * proofs); LG = l1,
// TB = 2^tb_log lines per block (TB <= 256); blockIdx.z = proof * 2^R + coset.
// Thread t of a block always handles line (column offset) t mod TB and rows
// t / TB + j * (NT / TB) for NT threads, so its coset twist w_H^(i0 b) and
// four-step twiddle w_L^(b2 k1) advance by one product per row instead of two
// table lookups.
template <typename T, int LG>
__global__ void __launch_bounds__(Pass1Threads<T>::value) k_lde_pass1(Group<LdeArgs<T>> g, int l, int R, int k,
                                                                      int tb_log, Twiddles tw) {
  extern __shared__ __align__(16) unsigned char smem[];
  using L = LineShape<LG>;
  constexpr int U = Batch<T>::value;
  const LdeArgs<T>& io = g.v[blockIdx.z >> R];
  const T* __restrict__ src = io.src;
  T* __restrict__ V = io.tmp;
  const int l2 = l - LG;
  const int W = 1 << k;
  const int TB = 1 << tb_log;
  const int b2_0 = blockIdx.x << tb_log;
  const int a = blockIdx.y;
  const int i0 = blockIdx.z & ((1 << R) - 1);
  const int H_log = l + R;
  u32* stw = reinterpret_cast<u32*>(smem);
  T* sh = line_storage<T, LG>(smem);
  fill_stage_twiddles<LG>(stw, tw);
  const int t = threadIdx.x & (TB - 1);
  const int row0 = threadIdx.x >> tb_log;
  const int rstep = Pass1Threads<T>::value >> tb_log;
  const int b2 = b2_0 + t;
  const T* colp = src + ((u64)a << l) + b2;
  u32 tcur = ONE, tstep = ONE;
  if (i0) {
    tcur = root_pow(tw, H_log, (u64)i0 * (((u64)row0 << l2) + b2));
    tstep = root_pow(tw, H_log, (u64)i0 * ((u64)rstep << l2));
  }
  for (int r = row0; r < L::N; r += U * rstep) {
    T xs[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int b1 = r + u * rstep;
      if (b1 < L::N) xs[u] = colp[(u64)b1 << l2];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int b1 = r + u * rstep;
      if (b1 < L::N) {
        T x = xs[u];
        if (i0) {
          x = tmulb(x, tcur);
          tcur = mul(tcur, tstep);
        }
        sh[t * L::LS + L::pos(bitrev(b1, LG))] = x;
      }
    }
  }
  __syncthreads();
  dit_lines<T, LG, RadixMax<T>::value>(sh, TB, stw);
  T* out = V + (((u64)i0 * W + a) << l) + b2;
  u32 ocur = root_pow(tw, l, (u64)b2 * row0);
  const u32 ostep = root_pow(tw, l, (u64)b2 * rstep);
  for (int k1 = row0; k1 < L::N; k1 += rstep) {
    out[(u64)k1 << l2] = tmulb(sh[t * L::LS + L::pos(k1)], ocur);
    ocur = mul(ocur, ostep);
  }
}

// Threads per pass-2 block.  Its W lines take about 64 KiB of LDS, so a CU
// holds two blocks: base-field blocks of 512 threads keep four waves per SIMD
// (-5% on the 2^24 base LDE); extension blocks stay at 256 (512 was slower).
template <typename T>
struct Pass2Threads {
  static constexpr int value = sizeof(T) == 4 ? 512 : 256;
};

// Pass 2 (or the only pass when direct != 0).  Grid: (L1, 2^R, proofs); LG = l2.
template <typename T, int LG>
__global__ void __launch_bounds__(Pass2Threads<T>::value) k_lde_pass2(Group<LdeArgs<T>> g, int l, int R, int k,
                                                                      int direct, Twiddles tw) {
  extern __shared__ __align__(16) unsigned char smem[];
  const LdeArgs<T>& io = g.v[blockIdx.z];
  const T* __restrict__ src = direct ? io.src : io.tmp;
  T* __restrict__ dst = io.dst;
  using L = LineShape<LG>;
  constexpr int U = Batch<T>::value;
  const int l2 = LG, L2 = L::N;
  const int l1 = l - LG;
  const int W = 1 << k;
  const int k1 = blockIdx.x;
  const int i0 = blockIdx.y;
  u32* stw = reinterpret_cast<u32*>(smem);
  T* sh = line_storage<T, LG>(smem);
  fill_stage_twiddles<LG>(stw, tw);
  const int total = W * L2;
  for (int e0 = threadIdx.x; e0 < total; e0 += U * Pass2Threads<T>::value) {
    T xs[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int e = e0 + u * Pass2Threads<T>::value;
      if (e < total) {
        const int a = e >> l2, b2 = e & (L2 - 1);
        xs[u] = direct ? src[((u64)a << l) + b2] : src[((((u64)i0 * W + a) << l)) + ((u64)k1 << l2) + b2];
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int e = e0 + u * Pass2Threads<T>::value;
      if (e < total) {
        const int a = e >> l2, b2 = e & (L2 - 1);
        T x = xs[u];
        if (direct && i0) x = tmulb(x, root_pow(tw, l + R, (u64)i0 * b2));
        sh[a * L::LS + L::pos(bitrev(b2, l2))] = x;
      }
    }
  }
  __syncthreads();
  dit_lines<T, LG, RadixMax<T>::value>(sh, W, stw);
  const int L1 = 1 << l1;
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    int a = e & (W - 1), k2 = e >> k;
    u64 row = ((u64)(k1 + (u64)L1 * k2) << R) + i0;
    dst[row * W + a] = sh[a * L::LS + L::pos(k2)];
  }
}

// Dynamic LDS above 64 KiB must be opted into per kernel and device; the
// opt-in is made once (hipFuncSetAttribute takes a runtime-wide lock, which
// concurrent provers would otherwise contend for on every launch).
template <typename K>
static void allow_lds(K kernel, size_t bytes) {
  if (bytes <= 64 * 1024) return;
  static std::mutex mu;
  static std::set<std::tuple<const void*, int, size_t>> done;
  int dev = 0;
  (void)hipGetDevice(&dev);
  const auto key = std::make_tuple(reinterpret_cast<const void*>(kernel), dev, bytes);
  std::lock_guard<std::mutex> lock(mu);
  if (done.count(key)) return;
  if (hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)bytes) == hipSuccess)
    done.insert(key);
}

static size_t lds_bytes(int lg, int lines, size_t esize) {
  size_t tw = ((sizeof(u32) << lg) + 15) & ~size_t(15);
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

template <typename T>
static auto pass2_for(int lg) -> void (*)(Group<LdeArgs<T>>, int, int, int, int, Twiddles) {
  switch (lg) {
    case 0: return k_lde_pass2<T, 0>;
    case 1: return k_lde_pass2<T, 1>;
    case 2: return k_lde_pass2<T, 2>;
    case 3: return k_lde_pass2<T, 3>;
    case 4: return k_lde_pass2<T, 4>;
    case 5: return k_lde_pass2<T, 5>;
    case 6: return k_lde_pass2<T, 6>;
    case 7: return k_lde_pass2<T, 7>;
    case 8: return k_lde_pass2<T, 8>;
    case 9: return k_lde_pass2<T, 9>;
    case 10: return k_lde_pass2<T, 10>;
    default: return nullptr;
  }
}

// Launch plan of one LDE: a single pass (l <= LP2) or the four-step split.
struct LdePlan {
  bool two_pass = false;
  int l1 = 0, l2 = 0, tb_log = 0;
  size_t sm1 = 0, sm2 = 0;
};

// Pass 1 takes about 64 KiB of lines per block (16K base or 4K extension
// elements), at most 256 lines and at most L2; lines of 2^12 and longer take
// fewer (down to one line per block) and up to 160 KiB of LDS.  A tuned
// tb_log (>= 0) replaces that rule, within the same limits.  Returns 0, or a
// negative code for a shape no kernel instance covers.
template <typename T>
static int lde_plan(int m, int k, int R, int LP2, int tb_tuned, LdePlan* pl) {
  const int l = m - k;
  if (l < 0 || k < 0 || R < 0 || k > 10) return -1;
  if (l + R > 24) return -7;   // the codeword domain exceeds KoalaBear's two-adic subgroups
  if (LP2 < (sizeof(T) == 4 ? 6 : 2)) return -6;   // rows too wide for a single-pass line in LDS
  const int W = 1 << k;
  if (l <= LP2) {
    pl->two_pass = false;
    pl->sm2 = lds_bytes(l, W, sizeof(T));
    if (pl->sm2 > 160 * 1024) return -2;
    if (!pass2_for<T>(l)) return -6;
    return 0;
  }
  pl->two_pass = true;
  pl->l2 = LP2;
  pl->l1 = l - LP2;
  const int elem_log = sizeof(T) == 4 ? 14 : 12;
  int tb_log = tb_tuned >= 0 ? tb_tuned : elem_log - pl->l1;
  if (tb_log > 8) tb_log = 8;   // TB <= 256 < Pass1Threads: the row step NT / TB is at least 2
  if (tb_log > pl->l2) tb_log = pl->l2;
  if (tb_log < 0) tb_log = 0;
  if (tb_tuned < 0)
    while (tb_log > 0 && lds_bytes(pl->l1, 1 << tb_log, sizeof(T)) > 80 * 1024) tb_log--;
  pl->tb_log = tb_log;
  pl->sm1 = lds_bytes(pl->l1, 1 << tb_log, sizeof(T));
  if (pl->sm1 > 160 * 1024) return -3;
  pl->sm2 = lds_bytes(pl->l2, W, sizeof(T));
  if (pl->sm2 > 160 * 1024) return -2;
  if (!pass1_for<T>(pl->l1) || !pass2_for<T>(pl->l2)) return -6;
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
  const int W = 1 << k;
  const unsigned np = (unsigned)count;
  if (!pl.two_pass) {
    auto kern = pass2_for<T>(l);
    allow_lds(kern, pl.sm2);
    return launch_group(io, count, [&](const Group<LdeArgs<T>>& g) {
      hipLaunchKernelGGL(kern, dim3(1, 1u << R, np), dim3(Pass2Threads<T>::value), pl.sm2, st, g, l, R, k, 1, tw);
    });
  }
  for (int i = 0; i < count; i++)
    if (io[i].tmp == nullptr) return -4;
  auto k1 = pass1_for<T>(pl.l1);
  auto k2 = pass2_for<T>(pl.l2);
  allow_lds(k1, pl.sm1);
  allow_lds(k2, pl.sm2);
  return launch_group(io, count, [&](const Group<LdeArgs<T>>& g) {
    hipLaunchKernelGGL(k1, dim3(1u << (pl.l2 - pl.tb_log), W, (1u << R) * np), dim3(Pass1Threads<T>::value), pl.sm1,
                       st, g, l, R, k, pl.tb_log, tw);
    hipLaunchKernelGGL(k2, dim3(1u << pl.l1, 1u << R, np), dim3(Pass2Threads<T>::value), pl.sm2, st, g, l, R, k, 0,
                       tw);
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
