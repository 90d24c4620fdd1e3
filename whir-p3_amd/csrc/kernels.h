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
  TR_GBASE = 5,           // grinding only: the candidates' base state is in tr.gbase already
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
  const ef* obs_sum;            // TR_OBS_GRIND: instead of obs, the sum of obs_nb partial sums (as in StepTr)
  int obs_nb;
  u32* obs_copy;                // TR_OBS_GRIND: block 0 also copies the n_obs words at obs here (as in StepTr)
  GrindBase* gbase;             // long grinds: device memory for the candidates' base state
};

// A WHIR round's transcript steps between its commitment and its sumcheck,
// run on the device (transcript.h, prover/mod.rs:192-248): observe the root,
// the OOD samples (challenge, expanded point, answer), the round's proof of
// work, the STIR query indices and the combination randomness gamma.  The
// host replays them from this log after the round's sumcheck.
constexpr int RT_MAX_OOD = 4;
constexpr int RT_MAX_QUERIES = 256;
// Largest run of words a transcript step observes from memory at once (the
// final polynomial, at most 64 extension elements).
constexpr int OBS_STAGE = 256;
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
  // Instead of obs: observe the sum of obs_nb partial sums at obs_sum (one
  // extension element, the eq-weighted sum's blocks; n_obs is 4).
  const ef* obs_sum;
  int obs_nb;
  // When given, block 0 also copies the n_obs words observed from obs here
  // (host-visible memory: the final polynomial for the proof).
  u32* obs_copy;
};

struct EqTables {
  const ef* hi[4];
  const ef* lo[4];
};

// ---- groups: one launch serving several proofs of one shape
// Every prover kernel takes the per-proof part of its arguments as a Group
// passed by value; blockIdx.z (folded with the coset index in LDE pass 1)
// selects the proof, so a launch for `count` proofs of the same shape runs
// count times the grid of one.  A single proof is a group of one.
constexpr int GROUP_MAX = 16;
template <class A>
struct Group {
  A v[GROUP_MAX];
};

// State a host-driven grind's candidates are observed into (see GrindBase).
struct Sponge16 {
  u32 s[16];
};

// ---- Launch shapes a context may tune (chosen by measurement; every value
// gives the same results).  0 or -1 keeps the built-in rule.
struct Tuning {
  // LDE, per element type [base, extension]: the longest column one pass
  // transforms in LDS (log2; longer ones take the four-step split at it), and
  // pass 1's lines per block (log2).
  int lde_lp2[2] = {0, 0};
  int lde_tb_log[2] = {-1, -1};
  // Merkle: one lane per node while a level has more than 2^tree_lane_log
  // nodes; below, 128-digest subtrees from 2^tree_s128_log nodes and
  // 32-digest ones under that.
  int tree_lane_log = 14;
  int tree_s128_log = 15;
  // Grinding: lane-parallel searches up to grind_wide_bits bits, lockstep
  // blocks from grind_lock_bits, one lane per candidate between; the blocks a
  // search takes (0: the context's rule from the CU count).
  int grind_wide_bits = 12;
  int grind_lock_bits = 17;
  int grind_blocks = 0;
};
// Whether the values are in the ranges the launchers accept.
bool tuning_valid(const Tuning& t);

// ---- Reed-Solomon encoding (k_lde.hip)
// src holds 2^m elements; the output (dst) is 2^(m-k+R) rows of 2^k elements.
// tmp (same size as the output) is needed when 2^(m-k) exceeds one LDS pass.
template <class T>
struct LdeArgs {
  const T* src;
  T* tmp;
  T* dst;
};
bool lde_two_pass(int m, int k, bool ext, const Tuning& t);
// 0 when lde_base / lde_ext support the shape, else the negative code they would return.
int lde_check(int m, int k, int R, bool ext, const Tuning& t);
int lde_base(const LdeArgs<u32>* a, int count, int m, int k, int R, const Twiddles& tw, const Tuning& t,
             hipStream_t st);
int lde_ext(const LdeArgs<ef>* a, int count, int m, int k, int R, const Twiddles& tw, const Tuning& t,
            hipStream_t st);

// ---- Merkle tree and openings (k_merkle.hip)
// tree holds 2 * height - 1 digests of 8 u32; the root is the last one and
// is also written to root_out (host-visible memory) when given.
struct TreeArgs {
  const u32* rows;
  u32* tree;
  u32* root_out;
};
int merkle_leaves(const TreeArgs* a, int count, int row_len, u64 height, int leaf_width, const P2C16* d_c16,
                  const P2C24* d_c24, hipStream_t st);
// Returns 1 (root not written to root_out) for a single leaf.
int merkle_tree(const TreeArgs* a, int count, u64 height, const P2C16& c16, const Tuning& t, hipStream_t st);
// The Keccak MMCS (k_merkle.hip): leaves and the tree above them, same layout.
int keccak_leaves(const TreeArgs* a, int count, int row_len, u64 height, hipStream_t st);
int keccak_tree(const TreeArgs* a, int count, u64 height, hipStream_t st);
// Smallest-witness grinding for the Keccak byte challenger.  A candidate w
// changes only its 4 bytes of the hashed input.  s0 is the sponge state after
// the blocks before those bytes with the next block (input, padding, witness
// bytes zero) XORed in; t1 is the block after it when the padding spills over
// (nblk == 2).  w << shift is XORed into lane lo_lane of block lo_blk (0), and
// w >> (64 - shift) into lane hi_lane of block hi_blk when shift > 32.  w
// passes when the byte-reversed high word of lane 3 of the digest has its low
// `bits` bits clear.
struct KeccakGrindArgs {
  uint64_t s0[25];
  uint64_t t1[17];
  u32* result;   // smallest passing witness (atomic minimum), P beforehand
  int nblk;
  int lo_blk, lo_lane, hi_blk, hi_lane, shift;
};
int keccak_grind(const KeccakGrindArgs& a, int bits, int blocks, hipStream_t st);
// Opening of the rows at idx[0 .. nq) (nq from device memory when given,
// else max_nq): each row followed by its sibling digests, into out.
struct OpenArgs {
  const u32* rows;
  const u32* tree;
  const u64* idx;
  const u32* nq;
  u32* out;
};
int merkle_open(const OpenArgs* a, int count, int row_len, u64 height, int max_nq, hipStream_t st);

// Grinding: the smallest witness in [start, start + span) (clipped to p),
// into *result, which must hold p (or a larger candidate) before the first
// range.  mode TR_NONE: the candidates' base state and position are base /
// wpos; otherwise every block first runs the transcript step `mode` of tr
// (TR_COEFFS_GRIND or TR_OBS_GRIND) and result is tr.grind.
struct GrindArgs {
  Sponge16 base;
  int wpos;
  u32* result;
  ScTr tr;
};
int grind_range(const P2C16* d_consts, const GrindArgs* a, int count, int bits, int mode, u64 start, u64 span,
                int blocks, const Tuning& t, hipStream_t st);

// ---- a round's transcript steps (k_merkle.hip)
// OOD sample: observe st.obs, sample the challenge y into *log_y, expand it to
// the m-variable point pt_out (y^(2^(m-1-j)) at j), and write its split eq
// tables (hi + lo = m variables).  Block 0 also sets *grind_reset to p when
// given (the round's grind result slot).
struct OodArgs {
  StepTr st;
  ef* log_y;
  ef* pt_out;
  ef* out_hi;
  ef* out_lo;
  u32* grind_reset;
};
int ood_eq_tables(const OodArgs* a, int count, int m, int hi, int lo, const P2C16* d_consts, hipStream_t st);
// All n_ood OOD samples of a round on a small polynomial (m <= OOD_SMALL_LOG) in
// one block per proof: for each sample what ood_eq_tables and the eq-weighted
// sum of e do (tables of sample i at tabs + i * tab_cap, its answer to ev[i]),
// the answers but the last also observed (and logged in log_ev) before the
// next sample; the last is left to the next step, as with the per-sample
// launches.
// The commitment's OOD samples (CommitmentWriter::commit, committer/writer.rs:
// 88-106) on the device transcript, one step per sample and a closing step.
// Each step: with prev_part, reduce the previous sample's eq-weighted block
// sums ([segment][block], prev_nb blocks) into its 2^sb segment sums (to
// prev_P) and its answer sum_s eq(prev_z[..sb], s) P(s) (to prev_eval), and
// observe the answer; else observe st.obs (the root, n_obs words).  Then,
// unless closing, sample y (to log_y), expand it to the point (to z_out) and
// write the split eq tables of its coordinates sb..n.  Grid: (blocks covering
// the tables, or 1 when closing, 1, proofs); every block runs the transcript.
constexpr int COOD_MAX_SB = 8;
struct CoodArgs {
  StepTr st;
  const ef* prev_part;
  int prev_nb;
  const ef* prev_z;
  ef* prev_P;
  ef* prev_eval;
  ef* log_y;
  ef* z_out;
  ef* out_hi;
  ef* out_lo;
};
int commit_ood(const CoodArgs* a, int count, int n, int sb, int hi, int lo, int sample, const P2C16* d_consts,
               hipStream_t st);
constexpr int OOD_SMALL_LOG = 10;
struct OodSmallArgs {
  StepTr st;        // observe st.obs (the root), then the samples
  ef* log_y;        // n_ood entries
  ef* log_ev;       // n_ood entries (all but the last written here)
  ef* pts;          // n_ood points of RT_MAX_VARS coordinates
  ef* tabs;         // eq tables: sample i's hi table at tabs + i * tab_cap, its lo table 2^hi after
  size_t tab_cap;
  const ef* e;      // the polynomial (2^m evaluations)
  ef* ev;           // n_ood answers
  u32* grind_reset;
};
int ood_small(const OodSmallArgs* a, int count, int m, int hi, int lo, int n_ood, const P2C16* d_consts,
              hipStream_t st);
// One block per proof: the round's witness from `grind` (pow_bits > 0) or
// st.obs, one discarded sample, the STIR query indices over the folded domain
// (sorted, distinct), their points folded_gen^idx, gamma and its powers
// (n_ood + nq of them).  Log and buffers as in RoundLog / RoundBufs.
// final_step: the final round's queries (no discarded sample, no gamma).
struct QueryArgs {
  StepTr st;
  u32* grind;   // the round's witness (pow_bits > 0); set back to p at the end, ready for the next search
  RoundLog* log;
  RoundBufs rb;
};
int round_queries(const QueryArgs* a, int count, int pow_bits, u64 domain_size, int folding_factor, int num_queries,
                  u32 folded_gen, int n_ood, const P2C16* d_consts, hipStream_t st, bool final_step = false);

// ---- multilinear kernels (k_poly.hip)
int eq_table(const PointArg& z, int first, int t, const ef& scale, ef* out, hipStream_t st);
int eq_outer(const ef* Ehi, const ef* Elo, int lo, u64 n, ef* out, hipStream_t st);
// Both halves of a split table: out_hi = scale * eq(z[0..hi)), out_lo =
// eq(z[hi..hi+lo)), the point's hi + lo coordinates read from z (device or
// host-visible memory).
struct EqSplitArgs {
  const ef* z;
  ef scale;
  ef* out_hi;
  ef* out_lo;
};
int eq_split(const EqSplitArgs* a, int count, int hi, int lo, hipStream_t st);
// Segmented eq-weighted sums of f (2^nf entries, 2^sb segments) against K
// split tables; partial holds nblocks * 2^sb * K block sums, and out receives
// out[seg * K + k].
struct DotArgs {
  const void* f;
  EqTables tabs;
  ef* partial;
  ef* out;   // nullptr: only the block partial sums (a transcript step reduces them)
};
int eq_dot_base(const DotArgs* a, int count, int nf, int sb, int lo, int K, int nblocks, hipStream_t st);
int eq_dot_ext(const DotArgs* a, int count, int nf, int sb, int lo, int K, int nblocks, hipStream_t st);
// Wt[x] = (accumulate ? Wt[x] : 0) + sum_k s_k eq_k(x); s_k = scale[k]
// (device memory) when scale is given, else 1.
struct CombineArgs {
  ef* Wt;
  EqTables tabs;
  const ef* scale;
};
int eq_combine(const CombineArgs* a, int count, u64 n, int lo, int K, int accumulate, hipStream_t st);
// Select terms sum_j g_j z_j^x added to Wt; the count comes from nsel_dev when
// given (nsel is then its maximum, A and B sized for it).
struct SelectArgs {
  ef* Wt;
  const u32* z;
  const ef* g;
  const u32* nsel_dev;
  ef* A;
  u32* B;
};
int select_combine(const SelectArgs* a, int count, int m, int nsel, hipStream_t st);
struct CompressArgs {
  const u32* f;
  const ef* eqrs;
  ef* out;
};
int compress_multi(const CompressArgs* a, int count, int n, int l, hipStream_t st);
struct ExtArgs {
  const u32* f;
  ef* out;
};
int base_to_ext(const ExtArgs* a, int count, u64 n, hipStream_t st);
// A sumcheck round on arrays e, w of length 2h.  sc_coeffs: (h(0), h(2))
// into c0c2.  sc_fold: fold by r, and when h >= 2 the next round's (h(0),
// h(2)) into c0c2.  Both use partial (2 * 1024 entries) as scratch.
// Device-transcript variants: sc_partials leaves the coefficient partial sums
// of *nb blocks in `partial` (and sets *grind_reset to p when given);
// sc_fold_tr runs the transcript step `mode` (TR_WITNESS_SAMPLE or
// TR_COEFFS_SAMPLE, reading tr.partial) to get r, folds, and leaves the next
// round's partials in `partial` when h >= 2 (*nb = 0 otherwise).
struct ScArgs {
  ef* e;
  ef* w;
  ef r;
  ef* partial;
  ef* c0c2;
  u32* grind_reset;
  ScTr tr;
};
int sc_coeffs(const ScArgs* a, int count, u64 h, hipStream_t st);
int sc_fold(const ScArgs* a, int count, u64 h, hipStream_t st);
int sc_partials(const ScArgs* a, int count, u64 h, unsigned* nb, hipStream_t st);
int sc_fold_tr(const ScArgs* a, int count, u64 h, int mode, const P2C16* d_consts, unsigned* nb, hipStream_t st);
// Every round of a phase without proof of work on 2^m0 <= 2^SC_SMALL_LOG
// entries in one block per proof (tr: ch_in, ch_out, log of `rounds` entries,
// grind_next; grind_reset as in sc_partials).
constexpr int SC_SMALL_LOG = 10;
int sc_small(const ScArgs* a, int count, int m0, int rounds, const P2C16* d_consts, hipStream_t st);
// A round with proof of work on arrays of 2^m <= 2^SC_SMALL_LOG entries, in
// one block per proof: with fold_first, the previous round's transcript step
// (observe the witness in *tr.grind, check its bits, sample r; logged in
// tr.log[-1]) and the fold of e and w by r (written back); then this round's
// (h(0), h(2)) observed (logged in tr.log[0]) and the candidates' base state
// written to tr.gbase, the challenger to tr.ch_out, *tr.grind_next set to p.
int sc_step(const ScArgs* a, int count, int m, int fold_first, const P2C16* d_consts, hipStream_t st);

// The initial sumcheck's rounds by the split route (svo.rs:465-559) on the
// device transcript.  Step i (0 <= i <= l0) of one proof, one block:
// with i > 0 and proof of work, round i-1's transcript step (observe the
// witness in *tr.grind, check its bits, sample r_{i-1} into rs[i-1]; logged in
// tr.log[i-1]); with i < l0, round i's (h(0), h(2)) from the constraints'
// partial sums P (J blocks of 2^l0, variable 0 the top index bit) folded by
// r_<i, observed and logged in tr.log[i], then the grinding state written to
// tr.gbase (*tr.grind_next set to p) or, without proof of work, r_i sampled
// into rs[i] and the log.
constexpr int SVO_DEV_MAX = 10;   // l0
constexpr int SVO_MAX_J = 64;      // constraints
constexpr int SVO_MAX_Z = 512;     // J * l0
struct SvoArgs {
  const ef* P;    // J * 2^l0 partial sums
  const ef* z;    // J * l0: each constraint point's first l0 coordinates
  const ef* ap;   // J: powers of the combination challenge
  ef* rs;         // l0 challenges (device memory)
  ScTr tr;
};
int svo_step(const SvoArgs* a, int count, int J, int l0, int i, int pow_bits, const P2C16* d_consts, hipStream_t st);

// Kernel arguments stay within 4 KiB for a full group.
static_assert(sizeof(Group<GrindArgs>) <= 4096 && sizeof(Group<ScArgs>) <= 4096 &&
                  sizeof(Group<QueryArgs>) <= 4096 && sizeof(Group<OodArgs>) <= 4096 &&
                  sizeof(Group<DotArgs>) <= 4096 && sizeof(Group<CombineArgs>) <= 4096 &&
                  sizeof(Group<SelectArgs>) <= 4096 && sizeof(Group<EqSplitArgs>) <= 4096 && sizeof(Group<SvoArgs>) <= 4096 && sizeof(Group<OodSmallArgs>) <= 4096 && sizeof(Group<CoodArgs>) <= 4096 &&
                  sizeof(Group<OpenArgs>) <= 4096 && sizeof(Group<TreeArgs>) <= 4096 &&
                  sizeof(Group<CompressArgs>) <= 4096 && sizeof(Group<LdeArgs<ef>>) <= 4096,
              "a group's kernel arguments exceed 4 KiB");

// Launches kern with the first `count` entries of `a` in a Group.
template <class A, class F>
int launch_group(const A* a, int count, F&& launch) {
  if (count < 1 || count > GROUP_MAX) return -8;
  Group<A> g;
  for (int i = 0; i < count; i++) g.v[i] = a[i];
  launch(g);
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

}  // namespace kb
