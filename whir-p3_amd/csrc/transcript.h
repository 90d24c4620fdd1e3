// Fiat-Shamir transcript of sumcheck rounds on the device, as a prologue of
// the kernels that need its result.
//
// A sumcheck round of the prover (SumcheckData::observe_and_sample,
// src/whir/proof.rs:163-193) observes the round polynomial's (h(0), h(2)),
// grinds when the round asks for proof of work, and samples the folding
// challenge r.  Instead of two host round trips per round, every block of the
// kernel that needs a transcript value computes it itself from the challenger
// state in memory: the grinding kernel observes (h(0), h(2)) to get the state
// its candidates are observed into, and the fold kernel observes the witness
// and samples r.  All blocks compute the same values; block 0 alone writes the
// new challenger state (to a different buffer than the one read), the round's
// log entry (h(0), h(2), witness, r) and the reset of the next round's grind
// result.  The host replays the logged rounds on its own challenger and checks
// every challenge.
//
// The challenger is DuplexChallenger<F, Poseidon2-16, 16, 8> of p3-challenger
// (used at src/bin/main.rs:104-106): observe clears the output buffer and
// appends to the input buffer, duplexing at 8 inputs; sample duplexes when
// the input buffer is non-empty or the output buffer is empty, then pops the
// last output; a duplex overwrites the first inputs of the sponge state with
// the buffered inputs, permutes, and refills the output buffer with the first
// 8 state elements.
#pragma once
#include "kcommon.h"
#include "kernels.h"
#include "poseidon2_wide.h"

namespace kb {

// The challenger held in the lanes of every wave of a block: lane j = t & 15
// of each 16-lane row holds sponge element j, input buffer element j and
// output buffer element j (j < 8), and every thread keeps the same buffer
// lengths.  Observed values are block-uniform (every thread passes the same
// value), so an observation is a register write in the lane it lands in, and
// every wave computes the lane-parallel permutation itself: no LDS and no
// barrier in the challenger.
struct BlockChallenger {
  u32 sp;    // sponge element t & 15
  u32 in;    // input buffer element t & 15 (t & 15 < 8)
  u32 out;   // output buffer element t & 15 (t & 15 < 8)
  int in_len;
  int out_len;
  const WideLane* wl;
  const P2C16* c;

  __device__ void load(const DevChallenger& s) {
    const int j = threadIdx.x & 15;
    sp = s.sponge[j];
    in = s.in[j & 7];
    out = s.out[j & 7];
    in_len = (int)s.in_len;
    out_len = (int)s.out_len;
  }
  __device__ void store(DevChallenger* s) const {   // block 0 only
    const int t = threadIdx.x;
    if (t < 16) s->sponge[t] = sp;
    if (t < 8) {
      s->in[t] = in;
      s->out[t] = out;
    }
    if (t == 0) {
      s->in_len = (u32)in_len;
      s->out_len = (u32)out_len;
    }
  }
  // Sponge element t & 15 with the buffered inputs written over the first
  // in_len elements (what a duplex would permute).
  __device__ u32 absorbed() const { return (int)(threadIdx.x & 15) < in_len ? in : sp; }
  __device__ void duplex() {
    const u32 x = permute16_wide(absorbed(), threadIdx.x & 15, *wl, *c);
    sp = x;
    out = x;
    in_len = 0;
    out_len = 8;
  }
  __device__ void observe(u32 v) {
    out_len = 0;
    if ((int)(threadIdx.x & 15) == in_len) in = v;
    in_len++;
    if (in_len == 8) duplex();
  }
  // Eight block-uniform words, as eight observe() calls: the first 8 - in_len
  // fill the input buffer, which then duplexes exactly once, and the other
  // in_len start the next input buffer.
  __device__ void observe8(const u32 (&v)[8]) {
    const int j = threadIdx.x & 15, n0 = in_len;
    u32 v1 = 0, v2 = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      if (j - n0 == i) v1 = v[i];
      if (j + 8 - n0 == i) v2 = v[i];
    }
    if (j >= n0 && j < 8) in = v1;
    in_len = 8;
    duplex();
    if (j < n0) in = v2;
    in_len = n0;
    out_len = n0 > 0 ? 0 : 8;
  }
  // Observes n words from memory (n <= OBS_STAGE): every word is loaded at
  // once into `stage` (LDS, one barrier), then each lane takes the words that
  // land in its input position, duplexing whenever the buffer is full, as n
  // observe() calls would.  One load latency instead of n serialized ones.
  // A second call with the same stage needs a barrier in between.
  __device__ void observe_words(const u32* w, int n, u32* stage) {
    if (n <= 0) return;
    for (int i = threadIdx.x; i < n; i += blockDim.x) stage[i] = w[i];
    __syncthreads();
    out_len = 0;
    const int j = threadIdx.x & 15;
    for (int i = 0; i < n;) {
      const int k = n - i < 8 - in_len ? n - i : 8 - in_len;
      if (j >= in_len && j < in_len + k) in = stage[i + j - in_len];
      in_len += k;
      i += k;
      if (in_len == 8) duplex();
    }
  }
  __device__ u32 sample() {
    if (in_len > 0 || out_len == 0) duplex();
    out_len--;
    // lane out_len of every wave holds output element out_len
    return (u32)__builtin_amdgcn_readlane((int)out, out_len);
  }
  // Four samples as an extension element.  The loop is not unrolled (each
  // sample may duplex), so the samples shift through registers instead of
  // being stored at a loop index: an indexed private array would be moved to
  // LDS at a flat work-item index, whose workgroup sizes come from the
  // dispatch packet in host memory (microseconds per wave).
  __device__ ef sample_ef() {
    u32 a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    for (int k = 0; k < 4; k++) {
      const u32 v = sample();
      a0 = a1;
      a1 = a2;
      a2 = a3;
      a3 = v;
    }
    return ef{{a0, a1, a2, a3}};
  }
};

// Observes a step's input with the whole block: n_obs words at obs, or (when
// obs_sum is given) the sum of obs_nb partial sums as one extension element,
// reduced here instead of by a launch of its own.  The lead thread logs an
// extension-element input (four words) to *log when given, and copies words
// observed from memory to copy when given.  stage: OBS_STAGE words of LDS;
// red: 16 extension elements of LDS.
__device__ __forceinline__ void observe_input(BlockChallenger& ch, const u32* obs, int n_obs, const ef* obs_sum,
                                              int obs_nb, ef* log, bool lead, u32* stage, ef* red,
                                              u32* copy = nullptr) {
  if (obs_sum) {
    ef acc[1] = {ef_zero()};
    for (int b = threadIdx.x; b < obs_nb; b += 256) acc[0] = ef_add(acc[0], obs_sum[b]);
    block_reduce_ef<1>(acc, red);   // the same sum in every thread
    if (lead && threadIdx.x == 0 && log) *log = acc[0];
    const u32 v[4] = {acc[0].c[0], acc[0].c[1], acc[0].c[2], acc[0].c[3]};
#pragma unroll
    for (int k = 0; k < 4; k++) ch.observe(v[k]);
    return;
  }
  ch.observe_words(obs, n_obs, stage);
  if (lead && threadIdx.x == 0 && n_obs == 4 && log) *log = ef{{obs[0], obs[1], obs[2], obs[3]}};
  if (lead && copy)
    for (int i = threadIdx.x; i < n_obs; i += blockDim.x) copy[i] = stage[i];
}

// Runs one transcript step with the whole block.  TR_COEFFS_GRIND leaves the
// candidates' base state in gb (LDS); the sampling modes leave r in *r (LDS).
// Every global load of the step (constants, challenger, partial sums or
// witness) is issued before the first value is used, so they share one
// memory latency.  Requires blockDim.x == 256.
__device__ __forceinline__ void tr_prologue(int mode, const ScTr& a, const P2C16& c, GrindBase* gb, ef* r) {
  __shared__ ef red[2 * 16];
  const int t = threadIdx.x;
  const WideLane wl = wide_lane(c, t & 15);
  BlockChallenger ch{0, 0, 0, 0, 0, &wl, &c};
  ch.load(*a.ch_in);
  const bool lead = blockIdx.x == 0 && blockIdx.y == 0;
  u32 err = 0;
  if (mode == TR_OBS_GRIND) {
    __shared__ u32 stage[OBS_STAGE];
    observe_input(ch, a.obs, a.n_obs, a.obs_sum, a.obs_nb, a.log_obs, lead, stage, red, a.obs_copy);
    if (t < 16) gb->s[t] = ch.absorbed();
    if (t == 0) gb->wpos = (u32)ch.in_len;
    __syncthreads();
    if (lead) ch.store(a.ch_out);
    return;
  }
  if (mode == TR_COEFFS_GRIND || mode == TR_COEFFS_SAMPLE) {
    ef acc[2] = {ef_zero(), ef_zero()};
    for (int b = t; b < a.nb; b += 256) {
      acc[0] = ef_add(acc[0], a.partial[2 * b]);
      acc[1] = ef_add(acc[1], a.partial[2 * b + 1]);
    }
    block_reduce_ef<2>(acc, red);   // the same sums in every thread
    if (lead && t == 0) {
      a.log->c0 = acc[0];
      a.log->c2 = acc[1];
      a.log->witness = 0;
    }
    const u32 v[8] = {acc[0].c[0], acc[0].c[1], acc[0].c[2], acc[0].c[3],
                      acc[1].c[0], acc[1].c[1], acc[1].c[2], acc[1].c[3]};
    ch.observe8(v);
    if (mode == TR_COEFFS_GRIND) {
      // a candidate w is observed into position in_len of this state
      if (t < 16) gb->s[t] = ch.absorbed();
      if (t == 0) gb->wpos = (u32)ch.in_len;
      __syncthreads();
      if (lead) ch.store(a.ch_out);
      return;
    }
  } else {
    const u32 w = *a.grind;
    const u32 wm = w < P ? to_monty(w) : 0u;
    if (w >= P) err = 1;
    ch.observe(wm);
    const u32 mask = a.pow_bits >= 32 ? 0xffffffffu : ((1u << a.pow_bits) - 1);
    if ((from_monty(ch.sample()) & mask) != 0) err = 1;
    if (lead && t == 0) a.log->witness = wm;
  }
  const ef rr = ch.sample_ef();
  if (t == 0) *r = rr;
  if (lead) {
    ch.store(a.ch_out);
    if (t == 0) {
      a.log->r = rr;
      a.log->err = err;
    }
    if (a.grind_next && t == 0) *a.grind_next = P;
  }
  __syncthreads();
}

// The sampling modes (TR_COEFFS_SAMPLE, TR_WITNESS_SAMPLE) run by wave 0 of
// the block alone, for kernels with several blocks per CU (the large folds):
// every wave of every block would otherwise compute the same permutation, and
// the waves sharing a SIMD would multiply its latency.  The other waves wait at
// the block barrier, with their own loads in flight.  r is left in *r (LDS).
// Requires blockDim.x to be a multiple of 64.
__device__ __forceinline__ void tr_prologue_wave0(int mode, const ScTr& a, const P2C16& c, ef* r) {
  const int t = threadIdx.x;
  if (t < 64) {
    const WideLane wl = wide_lane(c, t & 15);
    BlockChallenger ch{0, 0, 0, 0, 0, &wl, &c};
    ch.load(*a.ch_in);
    const bool lead = blockIdx.x == 0 && blockIdx.y == 0;
    u32 err = 0;
    if (mode == TR_COEFFS_SAMPLE) {
      ef acc0 = ef_zero(), acc1 = ef_zero();
      for (int b = t; b < a.nb; b += 64) {
        acc0 = ef_add(acc0, a.partial[2 * b]);
        acc1 = ef_add(acc1, a.partial[2 * b + 1]);
      }
      acc0 = row_sum16(acc0);
      acc1 = row_sum16(acc1);
      // the wave's four row sums, the same in every lane
      u32 v[8];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const u32 x0 = add(__builtin_amdgcn_readlane((int)acc0.c[k], 0), __builtin_amdgcn_readlane((int)acc0.c[k], 16));
        const u32 x1 = add(__builtin_amdgcn_readlane((int)acc0.c[k], 32), __builtin_amdgcn_readlane((int)acc0.c[k], 48));
        const u32 y0 = add(__builtin_amdgcn_readlane((int)acc1.c[k], 0), __builtin_amdgcn_readlane((int)acc1.c[k], 16));
        const u32 y1 = add(__builtin_amdgcn_readlane((int)acc1.c[k], 32), __builtin_amdgcn_readlane((int)acc1.c[k], 48));
        v[k] = add(x0, x1);
        v[4 + k] = add(y0, y1);
      }
      if (lead && t == 0) {
        a.log->c0 = ef{{v[0], v[1], v[2], v[3]}};
        a.log->c2 = ef{{v[4], v[5], v[6], v[7]}};
        a.log->witness = 0;
      }
      ch.observe8(v);
    } else {
      const u32 w = *a.grind;
      const u32 wm = w < P ? to_monty(w) : 0u;
      if (w >= P) err = 1;
      ch.observe(wm);
      const u32 mask = a.pow_bits >= 32 ? 0xffffffffu : ((1u << a.pow_bits) - 1);
      if ((from_monty(ch.sample()) & mask) != 0) err = 1;
      if (lead && t == 0) a.log->witness = wm;
    }
    const ef rr = ch.sample_ef();
    if (t == 0) *r = rr;
    if (lead) {
      ch.store(a.ch_out);
      if (t == 0) {
        a.log->r = rr;
        a.log->err = err;
        if (a.grind_next) *a.grind_next = P;
      }
    }
  }
  __syncthreads();
}

}  // namespace kb
