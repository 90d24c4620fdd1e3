// Host-side launchers of the device kernels.  Every launcher enqueues on the
// given stream and returns 0 or a negative code (shape not supported or launch
// failure); none of them synchronizes.
#pragma once
#include <hip/hip_runtime.h>

#include "field.h"
#include "kcommon.h"
#include "poseidon2.h"

namespace kb {

// Point of up to 32 extension coordinates, passed by value to kernels.
struct PointArg {
  ef z[32];
};

// Up to four split eq tables: eq(z, x) = hi[k][x >> lo] * lo[k][x & (2^lo - 1)].
// Device copy of a DuplexChallenger<F, Poseidon2-16, 16, 8> (Montgomery form).
struct DevChallenger {
  u32 sponge[16];
  u32 in[8];
  u32 out[8];
  u32 in_len, out_len;
  u32 pad[6];
};
// State a grinding candidate is observed into: sponge with the buffered
// inputs written over its first elements; the candidate goes to s[wpos].
struct GrindBase {
  u32 s[16];
  u32 wpos;
  u32 pad[15];
};
// One device-transcript sumcheck round: round polynomial (h(0), h(2)), the
// proof-of-work witness (Montgomery form, 0 without grinding), the challenge.
struct ScLog {
  ef c0, c2, r;
  u32 witness, err, pad[2];
};

// One device transcript step of a sumcheck round, run as a prologue by the
// kernel that needs its result (transcript.h).
enum ScTrMode {
  TR_NONE = 0,
  TR_COEFFS_GRIND = 1,    // observe (h(0), h(2)); the grind state for the candidates
  TR_WITNESS_SAMPLE = 2,  // observe the witness, check its bits, sample r
  TR_COEFFS_SAMPLE = 3,   // observe (h(0), h(2)) and sample r (no proof of work)
  TR_OBS_GRIND = 4,       // observe obs; the grind state for the candidates
};

// Everything a transcript prologue reads and writes (a kernel argument).
// TR_OBS_GRIND (a round's proof of work) uses ch_in, ch_out and obs/n_obs/
// log_obs only.
struct ScTr {
  const DevChallenger* ch_in;   // state to start from (device or host-visible memory)
  DevChallenger* ch_out;        // state after this step (block 0 writes it)
  ScLog* log;                   // this round's entry
  const ef* partial;            // (h(0), h(2)) partial sums of `nb` blocks
  u32* grind;                   // this round's grind result
  u32* grind_next;              // next round's grind result (block 0 resets it to p)
  int nb;
  int pow_bits;
  const u32* obs;               // TR_OBS_GRIND: words to observe first
  int n_obs;
  ef* log_obs;
  GrindBase* gbase;             // long grinds: device memory for the candidates' base state
};

// A WHIR round's transcript steps between its commitment and its sumcheck,
// run on the device (transcript.h, prover/mod.rs:192-248): observe the root,
// the OOD samples (challenge, expanded point, answer), the round's proof of
// work, the STIR query indices and the combination randomness gamma.  The
// host replays them from this log after the round's sumcheck.
constexpr int RT_MAX_OOD = 4;
constexpr int RT_MAX_QUERIES = 256;
constexpr int RT_MAX_VARS = 32;
struct RoundLog {
  ef y[RT_MAX_OOD];    // OOD challenges as sampled (before expansion)
  ef ev[RT_MAX_OOD];   // OOD answers, as observed
  ef gamma;
  u32 witness;         // proof-of-work witness (Montgomery form)
  u32 nq;              // distinct query indices, sorted, in idx
  u32 err;             // nonzero: witness rejected
  u32 pad;
  u32 idx[RT_MAX_QUERIES];
};
// Device buffers of a round's transcript.
struct RoundBufs {
  ef* pts;        // RT_MAX_OOD points of RT_MAX_VARS coordinates
  ef* ev;         // RT_MAX_OOD answers (written by the eq-weighted sums)
  u64* idx;       // RT_MAX_QUERIES query indices
  u32* nq;        // their count
  u32* selz;      // folded-domain point of each query
  ef* gpow;       // gamma^i, i < RT_MAX_OOD + RT_MAX_QUERIES
  u32* grind;     // grind slot of the round's proof of work
};
// One step's challenger hand-off: read ch_in, observe n_obs words at obs
// (logging them as an EF answer to log_obs when n_obs == 4), then the step's
// own samples; block 0 writes ch_out.
struct StepTr {
  const DevChallenger* ch_in;
  DevChallenger* ch_out;
  const u32* obs;
  int n_obs;
  ef* log_obs;
};

struct EqTables {
  const ef* hi[4];
  const ef* lo[4];
};

// ---- Reed-Solomon encoding (k_lde.hip)
// src holds 2^m elements; the output is 2^(m-k+R) rows of 2^k elements.
// tmp (same size as the output) is needed when 2^(m-k) exceeds one LDS pass.
bool lde_two_pass(int m, int k, bool ext);
// 0 when lde_base / lde_ext support the shape, else the negative code they would return.
int lde_check(int m, int k, int R, bool ext);
int lde_base(const u32* src, u32* dst, u32* tmp, int m, int k, int R, const Twiddles& tw, hipStream_t st);
int lde_ext(const ef* src, ef* dst, ef* tmp, int m, int k, int R, const Twiddles& tw, hipStream_t st);

// ---- Merkle tree and openings (k_merkle.hip)
// tree holds 2 * height - 1 digests of 8 u32; the root is the last one.
int merkle_leaves(const u32* rows, int row_len, u64 height, int leaf_width, const P2C16* d_c16, const P2C24* d_c24,
                  u32* tree, hipStream_t st);
// Writes the root to root_out too when given; returns 1 (root not written) for a single leaf.
int merkle_tree(u64 height, const P2C16& c16, u32* tree, hipStream_t st, u32* root_out);
int merkle_commit(const u32* rows, int row_len, u64 height, int leaf_width, const P2C16& c16, const P2C16* d_c16,
                  const P2C24* d_c24, u32* tree, hipStream_t st);
int merkle_open(const u32* rows, int row_len, const u32* tree, u64 height, const u64* d_idx, int nq, u32* d_out,
                hipStream_t st);
// Scans [start, start + count) for the smallest witness.  *d_result must
// hold p before the first range; it receives the smallest witness found.
int grind_range(const P2C16* d_consts, const u32* state16, int wpos, int bits, u32* d_result, u64 start, u64 count,
                int blocks, hipStream_t st);
// Same, with the base state and candidate position from the transcript step
// TR_COEFFS_GRIND run by every block first; the result goes to tr.grind.
int grind_range_tr(const P2C16* d_consts, const ScTr& tr, int bits, u64 start, u64 count, int blocks, hipStream_t st,
                   int mode = TR_COEFFS_GRIND);

// ---- a round's transcript steps (k_merkle.hip)
// OOD sample: observe st.obs, sample the challenge y into *log_y, expand it to
// the m-variable point pt_out (y^(2^(m-1-j)) at j), and write its split eq
// tables (hi + lo = m variables).  Block 0 also sets *grind_reset to p when
// given (the round's grind result slot).
int ood_eq_tables(const StepTr& st, ef* log_y, ef* pt_out, int m, int hi, int lo, ef* out_hi, ef* out_lo,
                  const P2C16* d_consts, hipStream_t st_, u32* grind_reset = nullptr);
// One block: the round's witness from `grind` (pow_bits > 0) or st.obs, one
// discarded sample, the STIR query indices over the folded domain (sorted,
// distinct), their points folded_gen^idx, gamma and its powers (n_ood + nq of
// them).  Log and buffers as in RoundLog / RoundBufs.  final_step: the final
// round's queries (no discarded sample, no gamma).
int round_queries(const StepTr& st, const u32* grind, int pow_bits, u64 domain_size, int folding_factor,
                  int num_queries, u32 folded_gen, int n_ood, RoundLog* log, const RoundBufs& rb,
                  const P2C16* d_consts, hipStream_t st_, bool final_step = false);
// merkle_open with the query indices and their count in device memory
// (at most max_nq).
int merkle_open_dev(const u32* rows, int row_len, const u32* tree, u64 height, const u64* d_idx, const u32* d_nq,
                    int max_nq, u32* d_out, hipStream_t st);

// ---- multilinear kernels (k_poly.hip)
int eq_table(const PointArg& z, int first, int t, const ef& scale, ef* out, hipStream_t st);
// Both halves of a split table in one launch: out_hi = scale * eq(z[0..hi)), out_lo = eq(z[hi..hi+lo)).
int eq_split(const PointArg& z, int hi, int lo, const ef& scale, ef* out_hi, ef* out_lo, hipStream_t st);
int eq_outer(const ef* Ehi, const ef* Elo, int lo, u64 n, ef* out, hipStream_t st);
int eq_dot_base(const u32* f, int nf, int sb, int lo, const EqTables& tabs, int K, ef* partial, int nblocks, ef* out,
                hipStream_t st);
int eq_dot_ext(const ef* f, int nf, int sb, int lo, const EqTables& tabs, int K, ef* partial, int nblocks, ef* out,
               hipStream_t st);
// Term k is scaled by d_scale[k] (device memory) when d_scale is given.
int eq_combine(ef* Wt, u64 n, int lo, const EqTables& tabs, int K, int accumulate, hipStream_t st,
               const ef* d_scale = nullptr);
// With d_nsel given, the count is read from device memory and nsel is its
// maximum (A and B sized for it).
int select_combine(ef* Wt, int m, const u32* d_z, const ef* d_g, int nsel, const u32* d_nsel, ef* A, u32* B,
                   hipStream_t st);
int compress_multi(const u32* f, int n, int l, const ef* d_eqrs, ef* out, hipStream_t st);
int base_to_ext(const u32* f, u64 n, ef* out, hipStream_t st);
int sc_coeffs(const ef* e, const ef* w, u64 h, ef* partial, ef* c0c2, hipStream_t st);
int sc_fold(ef* e, ef* w, u64 h, const ef& r, ef* partial, ef* c0c2, hipStream_t st);
// Device-transcript variants.  sc_partials: coefficient partial sums only
// (*nb blocks, two values each in `partial`); it also sets *grind_reset to p
// (the first round's grind result) when given.  sc_fold_tr: the transcript
// step `mode` (TR_WITNESS_SAMPLE or TR_COEFFS_SAMPLE) gives r, then the fold by
// r, leaving the next round's partials in `partial_out` when h >= 2 (*nb = 0
// otherwise).
int sc_partials(const ef* e, const ef* w, u64 h, ef* partial, unsigned* nb, u32* grind_reset, hipStream_t st);
int sc_fold_tr(ef* e, ef* w, u64 h, int mode, const ScTr& tr, const P2C16* d_consts, ef* partial_out, unsigned* nb,
               hipStream_t st);

}  // namespace kb
