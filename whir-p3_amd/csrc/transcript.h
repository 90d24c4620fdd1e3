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

// The challenger held in LDS by a whole block: every thread runs the same
// bookkeeping, and every wave computes the lane-parallel permutation (lane
// j & 15 holds element j & 15), so block barriers order the LDS updates.
struct BlockChallenger {
  u32* sp;    // [16] LDS
  u32* in;    // [8]  LDS
  u32* out;   // [8]  LDS
  int in_len;
  int out_len;
  const WideLane* wl;
  const P2C16* c;
  // Output element t & 15 (t & 15 < 8) in this lane's register, so a sample
  // is a lane read instead of an LDS round trip.
  u32 lane_out;

  __device__ void load(const DevChallenger& s) {
    const int t = threadIdx.x;
    if (t < 16) sp[t] = s.sponge[t];
    if (t < 8) {
      in[t] = s.in[t];
      out[t] = s.out[t];
    }
    in_len = (int)s.in_len;
    out_len = (int)s.out_len;
    lane_out = (t & 15) < 8 ? s.out[t & 15] : 0u;
    __syncthreads();
  }
  __device__ void store(DevChallenger* s) const {   // block 0 only
    const int t = threadIdx.x;
    if (t < 16) s->sponge[t] = sp[t];
    if (t < 8) {
      s->in[t] = in[t];
      s->out[t] = out[t];
    }
    if (t == 0) {
      s->in_len = (u32)in_len;
      s->out_len = (u32)out_len;
    }
  }
  __device__ void duplex() {
    const int t = threadIdx.x;
    if (t < in_len) sp[t] = in[t];
    __syncthreads();
    u32 x = sp[t & 15];
    x = permute16_wide(x, t & 15, *wl, *c);
    __syncthreads();
    if (t < 16) sp[t] = x;
    if (t < 8) out[t] = x;
    lane_out = x;
    __syncthreads();
    in_len = 0;
    out_len = 8;
  }
  __device__ void observe(u32 v) {
    out_len = 0;
    if (threadIdx.x == 0) in[in_len] = v;
    in_len++;
    __syncthreads();
    if (in_len == 8) duplex();
  }
  // Observes n words from memory (n <= OBS_STAGE): every word is loaded at
  // once into `stage` (LDS), then the input buffer is filled up to 8 words
  // per barrier, duplexing whenever it is full, as n observe() calls would.
  // One load latency instead of n serialized ones.
  __device__ void observe_words(const u32* w, int n, u32* stage) {
    if (n <= 0) return;
    for (int i = threadIdx.x; i < n; i += blockDim.x) stage[i] = w[i];
    __syncthreads();
    out_len = 0;
    for (int i = 0; i < n;) {
      const int k = n - i < 8 - in_len ? n - i : 8 - in_len;
      if ((int)threadIdx.x < k) in[in_len + threadIdx.x] = stage[i + threadIdx.x];
      __syncthreads();
      in_len += k;
      i += k;
      if (in_len == 8) duplex();
    }
  }
  __device__ u32 sample() {
    if (in_len > 0 || out_len == 0) duplex();
    out_len--;
    // lane out_len of every wave holds element out_len (lanes 0..15 hold 0..15)
    return (u32)__builtin_amdgcn_readlane((int)lane_out, out_len);
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

// Runs one transcript step with the whole block.  TR_COEFFS_GRIND leaves the
// candidates' base state in gb (LDS); the sampling modes leave r in *r (LDS).
// Requires blockDim.x == 256.
__device__ __forceinline__ void tr_prologue(int mode, const ScTr& a, const P2C16& c, GrindBase* gb, ef* r) {
  __shared__ u32 sp[16], in[8], out[8];
  __shared__ ef red[2 * 256];
  const int t = threadIdx.x;
  const WideLane wl = wide_lane(c, t & 15);
  BlockChallenger ch{sp, in, out, 0, 0, &wl, &c};
  ch.load(*a.ch_in);
  const bool lead = blockIdx.x == 0 && blockIdx.y == 0;
  u32 err = 0;
  if (mode == TR_OBS_GRIND) {
    __shared__ u32 stage[OBS_STAGE];
    ch.observe_words(a.obs, a.n_obs, stage);
    if (lead && t == 0 && a.n_obs == 4 && a.log_obs) *a.log_obs = ef{{a.obs[0], a.obs[1], a.obs[2], a.obs[3]}};
    if (t < 16) gb->s[t] = t < ch.in_len ? in[t] : sp[t];
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
    block_reduce_ef<2>(acc, red);   // the sums stay in red[0] and red[256]
    if (lead && t == 0) {
      a.log->c0 = acc[0];
      a.log->c2 = acc[1];
      a.log->witness = 0;
    }
    // observed from LDS: a loop index into the registers of acc would put
    // acc in memory (see sample_ef)
    for (int k = 0; k < 8; k++) ch.observe(red[(k >> 2) * 256].c[k & 3]);
    if (mode == TR_COEFFS_GRIND) {
      // a candidate w is observed into position in_len of this state
      if (t < 16) gb->s[t] = t < ch.in_len ? in[t] : sp[t];
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

}  // namespace kb
