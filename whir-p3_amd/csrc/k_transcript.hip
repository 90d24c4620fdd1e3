// Fiat-Shamir transcript of sumcheck rounds on the device.
//
// A sumcheck round of the prover (SumcheckData::observe_and_sample,
// src/whir/proof.rs:163-193) observes the round polynomial's (h(0), h(2)),
// grinds when the round asks for proof of work, and samples the folding
// challenge r.  Done on the host, that is two round trips per round (read the
// coefficients, read the witness).  Here one wave carries the round instead:
// it reduces the coefficient partial sums, runs the DuplexChallenger
// (DuplexChallenger<F, Poseidon2-16, 16, 8> of p3-challenger, as used at
// src/bin/main.rs:104-106) on a copy of the challenger state kept in device
// memory, hands the grinding kernel its base state, and after the grind
// observes the witness and samples r for the fold kernel, which reads it from
// device memory.  Every round is logged (h(0), h(2), witness, r) so the host
// can replay the transcript on its own challenger and check each challenge.
//
// Challenger semantics (same as the host's): observe clears the output buffer
// and appends to the input buffer, duplexing at 8 inputs; sample duplexes when
// the input buffer is non-empty or the output buffer is empty, then pops the
// last output; a duplex overwrites the first inputs of the sponge state with
// the buffered inputs, permutes, and refills the output buffer with the first
// 8 state elements.
#include "kcommon.h"
#include "kernels.h"
#include "poseidon2_wide.h"

namespace kb {

namespace {

// One wave; every lane runs the same bookkeeping, the arrays live in LDS, and
// the permutation runs lane-parallel (lane j & 15 holds element j & 15; the
// four rows of the wave compute the same state).
struct WaveChallenger {
  u32* sp;
  u32* in;
  u32* out;
  int in_len;
  int out_len;
  const WideLane* wl;
  const P2C16* c;

  __device__ void duplex() {
    const int lane = threadIdx.x;
    if (lane < in_len) sp[lane] = in[lane];
    __syncthreads();
    u32 x = sp[lane & 15];
    x = permute16_wide(x, lane & 15, *wl, *c);
    __syncthreads();
    if (lane < 16) sp[lane] = x;
    if (lane < 8) out[lane] = x;
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
  __device__ u32 sample() {
    if (in_len > 0 || out_len == 0) duplex();
    out_len--;
    return out[out_len];
  }
};

__device__ __forceinline__ ef wave_sum(ef v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    ef o;
#pragma unroll
    for (int k = 0; k < 4; k++) o.c[k] = (u32)__shfl_xor((int)v.c[k], off, 64);
    v = ef_add(v, o);
  }
  return v;
}

}  // namespace

// phase 0: reduce the (h(0), h(2)) partials of `nb` blocks, log and observe
//          them; then either prepare the grind (base state, witness slot = p)
//          or, without proof of work, sample r.
// phase 1: read the grind result, observe it, check its bits, sample r.
// r goes to *d_r and the log; log->err is set when the witness is missing or
// does not check.
// The challenger is read from `init` when given (host-visible memory the
// caller fills before the launch, so a captured graph can be replayed with a
// new starting state), else from `ch`; it is written back to `ch`.
__global__ void __launch_bounds__(64) k_sc_transcript(int phase, const ef* __restrict__ partial, int nb,
                                                      DevChallenger* ch, const DevChallenger* init, GrindBase* gb,
                                                      u32* d_grind, ScLog* log, int pow_bits, ef* d_r,
                                                      const P2C16* __restrict__ cp) {
  __shared__ u32 sp[16], in[8], out[8];
  const int lane = threadIdx.x;
  const DevChallenger& src = init ? *init : *ch;
  if (lane < 16) sp[lane] = src.sponge[lane];
  if (lane < 8) {
    in[lane] = src.in[lane];
    out[lane] = src.out[lane];
  }
  const WideLane wl = wide_lane(*cp, lane & 15);
  WaveChallenger wc{sp, in, out, (int)src.in_len, (int)src.out_len, &wl, cp};
  __syncthreads();
  u32 err = 0;
  bool sample_r = true;
  if (phase == 0) {
    ef a0 = ef_zero(), a2 = ef_zero();
    for (int b = lane; b < nb; b += 64) {
      a0 = ef_add(a0, partial[2 * b]);
      a2 = ef_add(a2, partial[2 * b + 1]);
    }
    a0 = wave_sum(a0);
    a2 = wave_sum(a2);
    if (lane == 0) {
      log->c0 = a0;
      log->c2 = a2;
      log->witness = 0;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) wc.observe(a0.c[k]);
#pragma unroll
    for (int k = 0; k < 4; k++) wc.observe(a2.c[k]);
    if (pow_bits > 0) {
      // a candidate w is observed into position in_len of this state
      if (lane < 16) gb->s[lane] = lane < wc.in_len ? in[lane] : sp[lane];
      if (lane == 0) {
        gb->wpos = (u32)wc.in_len;
        *d_grind = P;
      }
      sample_r = false;
    }
  } else {
    const u32 w = *d_grind;
    const u32 wm = w < P ? to_monty(w) : 0u;
    if (w >= P) err = 1;
    wc.observe(wm);
    const u32 mask = pow_bits >= 32 ? 0xffffffffu : ((1u << pow_bits) - 1);
    if ((from_monty(wc.sample()) & mask) != 0) err = 1;
    if (lane == 0) log->witness = wm;
  }
  if (sample_r) {
    ef r;
#pragma unroll
    for (int k = 0; k < 4; k++) r.c[k] = wc.sample();
    if (lane == 0) {
      log->r = r;
      log->err = err;
      *d_r = r;
    }
  }
  if (lane < 16) ch->sponge[lane] = sp[lane];
  if (lane < 8) {
    ch->in[lane] = in[lane];
    ch->out[lane] = out[lane];
  }
  if (lane == 0) {
    ch->in_len = (u32)wc.in_len;
    ch->out_len = (u32)wc.out_len;
  }
}

int sc_transcript(int phase, const ef* partial, unsigned nb, DevChallenger* ch, const DevChallenger* init,
                  GrindBase* gb, u32* d_grind, ScLog* log, int pow_bits, ef* d_r, const P2C16* d_consts,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_sc_transcript, dim3(1), dim3(64), 0, st, phase, partial, (int)nb, ch, init, gb, d_grind, log,
                     pow_bits, d_r, d_consts);
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

}  // namespace kb
