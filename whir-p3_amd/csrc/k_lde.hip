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
#include "kcommon.h"
#include "kernels.h"

namespace kb {

// Radix-2 DIT over `nlines` lines of length 2^lg stored contiguously in LDS
// (each line already in bit-reversed order, lines `stride` elements apart;
// the stride is padded by one element to spread LDS banks).  `twl` holds w_{2^lg}^j, j < 2^(lg-1).
template <typename T>
__device__ void dit_lines(T* sh, int nlines, int lg, int stride, const u32* twl) {
  if (lg == 0) return;
  const int halfL = 1 << (lg - 1);
  const int total = nlines * halfL;
  for (int s = 1; s <= lg; s++) {
    const int half = 1 << (s - 1);
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
      int line = t >> (lg - 1);
      int bt = t & (halfL - 1);
      int g = bt >> (s - 1);
      int j = bt & (half - 1);
      int i1 = line * stride + (g << s) + j;
      int i2 = i1 + half;
      u32 w = twl[j << (lg - s)];
      T u = sh[i1];
      T v = tmulb(sh[i2], w);
      sh[i1] = tadd(u, v);
      sh[i2] = tsub(u, v);
    }
    __syncthreads();
  }
}

__device__ void fill_line_twiddles(u32* twl, int lg, const Twiddles& tw) {
  if (lg == 0) return;
  for (int j = threadIdx.x; j < (1 << (lg - 1)); j += blockDim.x) twl[j] = root_pow(tw, lg, (u64)j);
  __syncthreads();
}

// Pass 1 of the four-step split.  Grid: (L2 / TB, W, 2^R).
template <typename T, int TB>
__global__ void __launch_bounds__(256) k_lde_pass1(const T* __restrict__ src, T* __restrict__ V, int l, int l1, int R,
                                                   int k, Twiddles tw) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int L1 = 1 << l1, l2 = l - l1;
  const int W = 1 << k;
  const int b2_0 = blockIdx.x * TB;
  const int a = blockIdx.y;
  const int i0 = blockIdx.z;
  const int H_log = l + R;
  u32* twl = reinterpret_cast<u32*>(smem);
  T* sh = reinterpret_cast<T*>(smem + sizeof(u32) * (L1 > 1 ? L1 / 2 : 1) + 16);
  sh = reinterpret_cast<T*>((reinterpret_cast<uintptr_t>(sh) + 15) & ~uintptr_t(15));
  fill_line_twiddles(twl, l1, tw);
  const u64 col = (u64)a << l;
  for (int e = threadIdx.x; e < TB * L1; e += blockDim.x) {
    int t = e % TB, b1 = e / TB;
    u64 b = ((u64)b1 << l2) + b2_0 + t;
    T x = src[col + b];
    if (i0) x = tmulb(x, root_pow(tw, H_log, (u64)i0 * b));
    sh[t * (L1 + 1) + bitrev(b1, l1)] = x;
  }
  __syncthreads();
  dit_lines(sh, TB, l1, L1 + 1, twl);
  T* out = V + (((u64)i0 * W + a) << l);
  for (int e = threadIdx.x; e < TB * L1; e += blockDim.x) {
    int t = e % TB, k1 = e / TB;
    int b2 = b2_0 + t;
    T y = sh[t * (L1 + 1) + k1];
    if (b2 && k1) y = tmulb(y, root_pow(tw, l, (u64)b2 * k1));
    out[((u64)k1 << l2) + b2] = y;
  }
}

// Pass 2 (or the only pass when direct != 0).  Grid: (L1, 2^R).
template <typename T>
__global__ void __launch_bounds__(256) k_lde_pass2(const T* __restrict__ src, T* __restrict__ dst, int l, int l1, int R,
                                                   int k, int direct, Twiddles tw) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int l2 = l - l1, L2 = 1 << l2;
  const int W = 1 << k;
  const int k1 = blockIdx.x;
  const int i0 = blockIdx.y;
  u32* twl = reinterpret_cast<u32*>(smem);
  T* sh = reinterpret_cast<T*>(smem + sizeof(u32) * (L2 > 1 ? L2 / 2 : 1) + 16);
  sh = reinterpret_cast<T*>((reinterpret_cast<uintptr_t>(sh) + 15) & ~uintptr_t(15));
  fill_line_twiddles(twl, l2, tw);
  for (int e = threadIdx.x; e < W * L2; e += blockDim.x) {
    int a = e >> l2, b2 = e & (L2 - 1);
    T x;
    if (direct) {
      x = src[((u64)a << l) + b2];
      if (i0) x = tmulb(x, root_pow(tw, l + R, (u64)i0 * b2));
    } else {
      x = src[((((u64)i0 * W + a) << l)) + ((u64)k1 << l2) + b2];
    }
    sh[a * (L2 + 1) + bitrev(b2, l2)] = x;
  }
  __syncthreads();
  dit_lines(sh, W, l2, L2 + 1, twl);
  const int L1 = 1 << l1;
  for (int e = threadIdx.x; e < W * L2; e += blockDim.x) {
    int a = e & (W - 1), k2 = e >> k;
    u64 row = ((u64)(k1 + (u64)L1 * k2) << R) + i0;
    dst[row * W + a] = sh[a * (L2 + 1) + k2];
  }
}

// Dynamic LDS above 64 KiB must be opted into per kernel.
template <typename K>
static void allow_lds(K kernel, size_t bytes) {
  if (bytes > 64 * 1024)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)bytes);
}

static size_t lds_bytes(int lg, int lines, size_t esize) {
  size_t tw = sizeof(u32) * ((lg > 0 ? (1u << (lg - 1)) : 1)) + 32;
  return tw + (size_t)lines * ((1u << lg) + 1) * esize;
}

template <typename T, int LP2, int TB>
static int lde_impl(const T* src, T* dst, T* tmp, int m, int k, int R, const Twiddles& tw, hipStream_t st) {
  const int l = m - k;
  if (l < 0) return -1;
  const int W = 1 << k;
  if (l <= LP2) {
    size_t sm = lds_bytes(l, W, sizeof(T));
    if (sm > 160 * 1024) return -2;
    allow_lds(k_lde_pass2<T>, sm);
    hipLaunchKernelGGL((k_lde_pass2<T>), dim3(1, 1u << R), dim3(256), sm, st, src, dst, l, 0, R, k, 1, tw);
  } else {
    const int l2 = LP2, l1 = l - LP2;
    if ((size_t)TB * ((1u << l1) + 1) * sizeof(T) > 80 * 1024) return -3;
    if (tmp == nullptr) return -4;
    size_t sm1 = lds_bytes(l1, TB, sizeof(T));
    allow_lds(k_lde_pass1<T, TB>, sm1);
    hipLaunchKernelGGL((k_lde_pass1<T, TB>), dim3((1u << l2) / TB, W, 1u << R), dim3(256), sm1, st, src, tmp, l, l1, R,
                       k, tw);
    size_t sm2 = lds_bytes(l2, W, sizeof(T));
    if (sm2 > 160 * 1024) return -2;
    allow_lds(k_lde_pass2<T>, sm2);
    hipLaunchKernelGGL((k_lde_pass2<T>), dim3(1u << l1, 1u << R), dim3(256), sm2, st, tmp, dst, l, l1, R, k, 0, tw);
  }
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

// Largest single-pass length per element type (LDS holds W lines of it).
static int lp2_for(int k, bool ext) {
  // keep W * L2 * sizeof(elem) <= 64 KiB
  int budget_log = 16 - (ext ? 4 : 2) - k;   // log2(64K / esize / W)
  int lp = budget_log < 10 ? budget_log : 10;
  return lp < 1 ? 1 : lp;
}

bool lde_two_pass(int m, int k, bool ext) { return m - k > lp2_for(k, ext); }

int lde_base(const u32* src, u32* dst, u32* tmp, int m, int k, int R, const Twiddles& tw, hipStream_t st) {
  int lp = lp2_for(k, false);
  switch (lp) {
    case 10: return lde_impl<u32, 10, 16>(src, dst, tmp, m, k, R, tw, st);
    case 9: return lde_impl<u32, 9, 16>(src, dst, tmp, m, k, R, tw, st);
    case 8: return lde_impl<u32, 8, 16>(src, dst, tmp, m, k, R, tw, st);
    case 7: return lde_impl<u32, 7, 16>(src, dst, tmp, m, k, R, tw, st);
    case 6: return lde_impl<u32, 6, 16>(src, dst, tmp, m, k, R, tw, st);
    default: return -6;
  }
}

int lde_ext(const ef* src, ef* dst, ef* tmp, int m, int k, int R, const Twiddles& tw, hipStream_t st) {
  int lp = lp2_for(k, true);
  switch (lp) {
    case 10: return lde_impl<ef, 10, 4>(src, dst, tmp, m, k, R, tw, st);
    case 9: return lde_impl<ef, 9, 4>(src, dst, tmp, m, k, R, tw, st);
    case 8: return lde_impl<ef, 8, 4>(src, dst, tmp, m, k, R, tw, st);
    case 7: return lde_impl<ef, 7, 4>(src, dst, tmp, m, k, R, tw, st);
    case 6: return lde_impl<ef, 6, 4>(src, dst, tmp, m, k, R, tw, st);
    case 5: return lde_impl<ef, 5, 4>(src, dst, tmp, m, k, R, tw, st);
    case 4: return lde_impl<ef, 4, 4>(src, dst, tmp, m, k, R, tw, st);
    case 3: return lde_impl<ef, 3, 4>(src, dst, tmp, m, k, R, tw, st);
    case 2: return lde_impl<ef, 2, 4>(src, dst, tmp, m, k, R, tw, st);
    default: return -6;
  }
}

}  // namespace kb
