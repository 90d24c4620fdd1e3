/*
 * whir_gpu.h - C ABI of the MI355X WHIR prover (KoalaBear / EF4 / Poseidon2).
 *
 * Drop-in boundary for the commit/prove path of tcoratger/whir-p3.  Each entry
 * point names the reference interface it replaces (file:line in whir-p3).
 *
 * Conventions
 *  - Base-field elements are uint32_t in Montgomery form (a * 2^32 mod p), the
 *    in-memory layout of Plonky3's KoalaBear, so a Rust &[KoalaBear] can be
 *    passed as-is.  An extension element is 4 consecutive uint32_t (c0..c3 of
 *    c0 + c1 x + c2 x^2 + c3 x^3, x^4 = 3).  A multilinear point of n variables
 *    is n extension elements.  Digests are 8 base elements.
 *  - Every call returns WHIRGPU_OK (0) or an error code; no exception or panic
 *    crosses the ABI.  whirgpu_last_error() describes the last failure of the
 *    calling thread.
 *  - Contexts are independent and may be used from different threads; one
 *    context must not be used by two threads at once.  Calls block until
 *    their results are on the host.
 *  - The caller owns every host buffer.  Device memory belongs to the context,
 *    statement, prover-data and proof handles and is released by their
 *    destroy calls.
 */
#ifndef WHIR_GPU_H
#define WHIR_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  WHIRGPU_OK = 0,
  WHIRGPU_BAD_PARAMS = 1,       /* the reference asserts/panics (src/whir/parameters.rs:84-114, prover/mod.rs:110) */
  WHIRGPU_OOM = 2,
  WHIRGPU_DEVICE = 3,
  WHIRGPU_GRIND_FAILED = 4,     /* FiatShamirError::InvalidGrindingWitness (src/fiat_shamir/errors.rs:9-13) */
  WHIRGPU_BUFFER_TOO_SMALL = 5,
  WHIRGPU_BAD_STATE = 6,
  WHIRGPU_VERIFY_FAILED = 7,    /* VerifierError (src/whir/verifier/errors.rs): the reason is in whirgpu_last_error() */
  WHIRGPU_MALFORMED_PROOF = 8   /* the bytes are not a WhirProof encoding (serde deserialization error) */
};

/* ProtocolParameters (src/parameters/mod.rs:147-166) plus the number of
 * variables given to WhirConfig::new (src/whir/parameters.rs:81). */
typedef struct whirgpu_params {
  uint32_t num_variables;
  uint32_t security_level;
  uint32_t pow_bits;
  /* FoldingFactor (src/parameters/mod.rs:31-38): Constant(k) when
   * folding_factor_first == folding_factor == k, otherwise
   * ConstantFromSecondRound(folding_factor_first, folding_factor). */
  uint32_t folding_factor_first;
  uint32_t folding_factor;
  uint32_t starting_log_inv_rate;
  uint32_t rs_domain_initial_reduction_factor;
  uint32_t soundness_type;      /* SecurityAssumption: 0 UniqueDecoding, 1 JohnsonBound, 2 CapacityBound */
  uint32_t sumcheck_strategy;   /* SumcheckStrategy (src/whir/parameters.rs:16-29): 0 Classic, 1 Svo */
  uint32_t leaf_hash_width;     /* 16: PaddingFreeSponge<P2-16,16,8,8>; 24: PaddingFreeSponge<P2-24,24,16,8> */
  uint32_t digest_elems;        /* DIGEST_ELEMS of DomainSeparator::commit_statement (domain_separator.rs:95-105) */
  uint32_t mmcs;                /* WHIRGPU_MMCS_POSEIDON2 (the benchmark's) or WHIRGPU_MMCS_KECCAK */
} whirgpu_params;

/* The Merkle commitment and challenger a context proves with.
 * POSEIDON2: MerkleTreeMmcs over PaddingFreeSponge<Poseidon2, W, W-8, 8> and
 *   TruncatedPermutation<Poseidon2-16, 2, 8, 16>, DuplexChallenger<F,
 *   Poseidon2-16, 16, 8> (benches/whir.rs:46-86, main.rs:35-36).
 * KECCAK: the Keccak variant (src/whir/mod.rs:236-383): MerkleTreeMmcs<F, u64,
 *   SerializingHasher<PaddingFreeSponge<KeccakF, 25, 17, 4>>,
 *   CompressionFunctionFromHasher<.., 2, 4>, 4> with
 *   SerializingChallenger32<F, HashChallenger<u8, Keccak256Hash, 32>>;
 *   proved with the whirgpu_*_keccak entry points (transcript on the host). */
#define WHIRGPU_MMCS_POSEIDON2 0u
#define WHIRGPU_MMCS_KECCAK 1u

/* Poseidon2 round constants (Montgomery form) for the width-16 permutation
 * (challenger, node compression, width-16 leaves) and the width-24 one
 * (width-24 leaves), as held by Plonky3's Poseidon2KoalaBear<16/24>. */
typedef struct whirgpu_poseidon2 {
  uint32_t w16_external_initial[4 * 16];
  uint32_t w16_external_terminal[4 * 16];
  uint32_t w16_internal[20];
  uint32_t w24_external_initial[4 * 24];
  uint32_t w24_external_terminal[4 * 24];
  uint32_t w24_internal[23];
} whirgpu_poseidon2;

/* DuplexChallenger<F, Poseidon2KoalaBear<16>, 16, 8> state (p3-challenger):
 * sponge state, pending inputs and unread outputs (popped from the end). */
typedef struct whirgpu_challenger {
  uint32_t sponge[16];
  uint32_t input[8];
  uint32_t output[8];
  uint32_t input_len;
  uint32_t output_len;
} whirgpu_challenger;

/* The Keccak variant's challenger, SerializingChallenger32<F,
 * HashChallenger<u8, Keccak256Hash, 32>> (src/whir/mod.rs:257): the
 * HashChallenger's input bytes and the output bytes not yet popped (popped
 * from the end). */
#define WHIRGPU_KECCAK_INPUT_CAP 4096
typedef struct whirgpu_keccak_challenger {
  uint8_t input[WHIRGPU_KECCAK_INPUT_CAP];
  uint32_t input_len;
  uint8_t output[32];
  uint32_t output_len;
} whirgpu_keccak_challenger;

typedef struct whirgpu_ctx whirgpu_ctx;                 /* WhirConfig + device + MMCS */
typedef struct whirgpu_statement whirgpu_statement;     /* InitialStatement */
typedef struct whirgpu_prover_data whirgpu_prover_data; /* MT::ProverData<DenseMatrix<F>> */
typedef struct whirgpu_proof whirgpu_proof;             /* WhirProof */

const char* whirgpu_last_error(void);

/* Poseidon2KoalaBear<16>::new_from_rng_128 then Poseidon2KoalaBear<24>::new_from_rng_128
 * on SmallRng::seed_from_u64(seed) (src/bin/main.rs:100-102, benches/whir.rs:66-68). */
int whirgpu_poseidon2_from_seed(uint64_t seed, whirgpu_poseidon2* out);

/* `count` draws of rng.random::<KoalaBear>() on SmallRng::seed_from_u64(seed)
 * (the polynomial of benches/whir.rs:96-97), Montgomery form. */
int whirgpu_sample_field(uint64_t seed, uint32_t* out, size_t count);

/* One Poseidon2 permutation on the host (width 16 or 24), for checking. */
int whirgpu_poseidon2_permute(const whirgpu_poseidon2* perms, uint32_t width, uint32_t* state);

/* WhirConfig::new (src/whir/parameters.rs:81-234) on device `device`. */
int whirgpu_create(int device, const whirgpu_params* params, const whirgpu_poseidon2* perms, whirgpu_ctx** out);
int whirgpu_destroy(whirgpu_ctx* ctx);

/* The derived schedule: [commitment_ood_samples, starting_folding_pow_bits,
 * n_rounds, final_queries, final_pow_bits, final_sumcheck_rounds,
 * final_folding_pow_bits] then per round [pow_bits, folding_pow_bits,
 * num_queries, ood_samples, num_variables, folding_factor, log2(domain_size),
 * folded_domain_gen]. */
int whirgpu_config(const whirgpu_ctx* ctx, uint32_t* out, size_t cap, size_t* len);

/* DomainSeparator::{commit_statement, add_whir_proof} pattern
 * (src/fiat_shamir/domain_separator.rs:95-169), Montgomery form. */
int whirgpu_domain_separator(const whirgpu_ctx* ctx, uint32_t* out, size_t cap, size_t* len);

/* DuplexChallenger::new (empty state), and
 * DomainSeparator::observe_domain_separator (domain_separator.rs:81-86). */
int whirgpu_challenger_new(whirgpu_challenger* ch);
int whirgpu_challenger_observe(const whirgpu_ctx* ctx, whirgpu_challenger* ch, const uint32_t* values, size_t n);
int whirgpu_challenger_observe_domain_separator(const whirgpu_ctx* ctx, whirgpu_challenger* ch);
/* CanSample::sample: `n` draws into out (Montgomery form). */
int whirgpu_challenger_sample(const whirgpu_ctx* ctx, whirgpu_challenger* ch, uint32_t* out, size_t n);

/* WhirConfig::initial_statement (src/whir/parameters.rs:419-429): takes the
 * 2^n evaluations (host pointer, or device pointer when evals_on_device != 0;
 * a device buffer is borrowed and must outlive the statement). */
int whirgpu_statement_create(whirgpu_ctx* ctx, const uint32_t* evals, int evals_on_device, whirgpu_statement** out);
/* InitialStatement::evaluate (src/whir/constraints/statement/initial.rs:143-152). */
int whirgpu_statement_evaluate(whirgpu_statement* st, const uint32_t* point, uint32_t* eval_out);
/* Clone of InitialStatement (#[derive(Clone)], src/whir/constraints/statement/initial.rs:31-32): the
 * constraints and evaluations so far; a statement over a borrowed device buffer borrows the same buffer. */
int whirgpu_statement_clone(const whirgpu_statement* st, whirgpu_statement** out);
int whirgpu_statement_destroy(whirgpu_statement* st);

/* WhirProof::from_protocol_parameters (src/whir/proof.rs:209-233). */
int whirgpu_proof_create(whirgpu_ctx* ctx, whirgpu_proof** out);
int whirgpu_proof_destroy(whirgpu_proof* proof);
/* Proof bytes: field values as Montgomery u32, lengths as u64, Option tags as
 * one byte, little endian, fields in WhirProof declaration order
 * (src/whir/proof.rs:17-141). */
int whirgpu_proof_serialize(const whirgpu_proof* proof, uint8_t* out, size_t cap, size_t* len);
/* serde deserialization of WhirProof (src/whir/proof.rs:11-141): reads the
 * bytes whirgpu_proof_serialize writes into a new proof of `ctx`'s
 * configuration.  Truncated or trailing bytes, unknown Option tags or opening
 * types and field words not below p return WHIRGPU_MALFORMED_PROOF. */
int whirgpu_proof_deserialize(whirgpu_ctx* ctx, const uint8_t* bytes, size_t len, whirgpu_proof** out);

/* CommitmentReader::parse_commitment + Verifier::verify
 * (src/whir/committer/reader.rs:40-82, src/whir/verifier/mod.rs:58-304) on
 * the host: checks `proof` for the statement {points[k] -> evals[k]}
 * (npoints points of n extension elements, npoints extension evaluations, in
 * the order they were evaluated).  `ch` is the verifier's challenger after
 * the domain separator; on success it ends where the prover's ended.  A
 * rejected proof returns WHIRGPU_VERIFY_FAILED with the failed check in
 * whirgpu_last_error().  Needs no device: a context created with device -1
 * verifies. */
int whirgpu_verify(const whirgpu_ctx* ctx, const whirgpu_proof* proof, whirgpu_challenger* ch, uint32_t npoints,
                   const uint32_t* points, const uint32_t* evals);
/* The same with the Keccak challenger, for a Keccak context. */
int whirgpu_verify_keccak(const whirgpu_ctx* ctx, const whirgpu_proof* proof, whirgpu_keccak_challenger* ch,
                          uint32_t npoints, const uint32_t* points, const uint32_t* evals);

int whirgpu_keccak_challenger_new(whirgpu_keccak_challenger* ch);
/* observe / sample field elements (Montgomery words), and
 * DomainSeparator::observe_domain_separator, for a Keccak context.  observe
 * and sample need no constants and accept ctx == NULL. */
int whirgpu_keccak_challenger_observe(const whirgpu_ctx* ctx, whirgpu_keccak_challenger* ch, const uint32_t* values,
                                      size_t n);
int whirgpu_keccak_challenger_sample(const whirgpu_ctx* ctx, whirgpu_keccak_challenger* ch, uint32_t* out, size_t n);
int whirgpu_keccak_challenger_observe_domain_separator(const whirgpu_ctx* ctx, whirgpu_keccak_challenger* ch);
/* CanObserve<u8> of the inner HashChallenger (a digest Hash<F, u64, 4> is
 * observed as its lanes' little-endian bytes), and
 * CanSampleBits::sample_bits of SerializingChallenger32 (4 sampled bytes as a
 * little-endian u32, masked to `bits` <= 31).  ctx may be NULL. */
int whirgpu_keccak_challenger_observe_bytes(const whirgpu_ctx* ctx, whirgpu_keccak_challenger* ch,
                                            const uint8_t* bytes, size_t n);
int whirgpu_keccak_challenger_sample_bits(const whirgpu_ctx* ctx, whirgpu_keccak_challenger* ch, uint32_t bits,
                                          uint32_t* out);
/* GrindingChallenger::grind of the Keccak challenger (smallest witness, on
 * the device), for a Keccak context; updates *ch as whirgpu_grind does. */
int whirgpu_keccak_grind(whirgpu_ctx* ctx, whirgpu_keccak_challenger* ch, uint32_t bits, uint32_t* witness_out);
/* Keccak-f[1600] on 25 lanes (lane x + 5y), for checking. */
int whirgpu_keccak_f1600(uint64_t* state);

/* CommitmentWriter::commit (src/whir/committer/writer.rs:56-109). */
int whirgpu_commit(whirgpu_ctx* ctx, whirgpu_proof* proof, whirgpu_challenger* ch, whirgpu_statement* st,
                   whirgpu_prover_data** out);
/* Prover::prove (src/whir/prover/mod.rs:97-128); consumes prover_data
 * (the reference moves it in), which is destroyed on return. */
int whirgpu_prove(whirgpu_ctx* ctx, whirgpu_proof* proof, whirgpu_challenger* ch, const whirgpu_statement* st,
                  whirgpu_prover_data* prover_data);
int whirgpu_prover_data_destroy(whirgpu_prover_data* pd);

/* Batched commit + prove (benches/whir.rs:157-183 for many polynomials at
 * once; the reference proves one at a time): `count` proofs (1 to 16) of the
 * context's configuration, each with its own statement (its constraints
 * already evaluated, the same number for all), proof object and challenger,
 * run as one group in lockstep, every device step one launch for all of them.
 * Each proof and challenger ends exactly as whirgpu_commit + whirgpu_prove
 * would leave it.  Poseidon2 contexts only (the transcripts run on the
 * device). */
int whirgpu_prove_group(whirgpu_ctx* ctx, uint32_t count, whirgpu_statement* const* statements,
                        whirgpu_proof* const* proofs, whirgpu_challenger* challengers);
/* CommitmentWriter::commit and Prover::prove with the Keccak challenger, for a
 * context created with mmcs = WHIRGPU_MMCS_KECCAK (the Keccak variant,
 * src/whir/mod.rs:236-383): the transcript runs on the host, round by round,
 * and the Merkle trees are Keccak trees; the proof has the same layout, its
 * digests 4 u64 lanes.  Grinding, when the schedule has proof-of-work bits,
 * searches on the host. */
int whirgpu_commit_keccak(whirgpu_ctx* ctx, whirgpu_proof* proof, whirgpu_keccak_challenger* ch,
                          whirgpu_statement* st, whirgpu_prover_data** out);
int whirgpu_prove_keccak(whirgpu_ctx* ctx, whirgpu_proof* proof, whirgpu_keccak_challenger* ch,
                         const whirgpu_statement* st, whirgpu_prover_data* pd);
/* InitialStatement::evaluate (initial.rs:143-152) for `count` statements of
 * one context at once (1 to 16), one point each: points holds count * n
 * extension elements, evals_out receives count extension elements. */
int whirgpu_statement_evaluate_group(uint32_t count, whirgpu_statement* const* statements, const uint32_t* points,
                                     uint32_t* evals_out);

/* Launch shapes of a context's kernels (no reference counterpart: chosen by
 * measurement, see whir-p3_amd/autotune.py).  Every value gives the same
 * proof bytes; 0 or -1 keeps the built-in rule.
 * - lde_lp2_{base,ext}: longest column (log2, 1..10) one LDE pass transforms
 *   in LDS; longer columns take the four-step split at it;
 * - lde_tb_log_{base,ext}: pass-1 lines per block (log2, 0..8);
 * - tree_lane_log: one lane per node while a Merkle level has more than
 *   2^tree_lane_log nodes (4..30, default 14); tree_s128_log: 128-digest
 *   subtrees from 2^tree_s128_log nodes (7..30, default 15);
 * - grind_wide_bits / grind_lock_bits: lane-parallel searches up to this many
 *   bits (default 12), lockstep blocks from this many (default 17);
 *   grind_blocks: blocks per search (0: from the CU count). */
typedef struct whirgpu_tuning {
  int32_t lde_lp2_base, lde_lp2_ext, lde_tb_log_base, lde_tb_log_ext;
  int32_t tree_lane_log, tree_s128_log;
  int32_t grind_wide_bits, grind_lock_bits, grind_blocks;
} whirgpu_tuning;
int whirgpu_get_tuning(const whirgpu_ctx* ctx, whirgpu_tuning* out);
/* "<name> <gcnArchName> (<CUs> CUs)" of a device (the name may be empty),
 * the key of a tuning cache. */
int whirgpu_device_name(int device, char* out, size_t cap);
/* Applies `t` after checking its ranges and that every Reed-Solomon encoding
 * of the context's schedule has a launch plan under it (WHIRGPU_BAD_PARAMS
 * otherwise, the context unchanged).  Not while the context is proving. */
int whirgpu_set_tuning(whirgpu_ctx* ctx, const whirgpu_tuning* t);

/* Per-stage device time of the last commit/prove on this context, in ms:
 * [lde, merkle, grind, sumcheck, mle, other, total_wall]. */
int whirgpu_last_timings(const whirgpu_ctx* ctx, double* out7);

/* Live kernel timing with HIP events on the context's stream (no reference
 * counterpart; the reference's tracing spans are the analogue).  After
 * enabling, whirgpu_profile_read returns and resets, per stage, the triple
 * [device ms, launches, work], stages in this order: LDE base (work = bytes),
 * LDE extension (bytes), Merkle leaves (Poseidon2 permutations), Merkle tree
 * (permutations), grinding (candidates up to the witness), sumcheck rounds
 * (bytes), eq-weighted sums (bytes), weight accumulation (bytes), SVO fold
 * (bytes), eq expansion (bytes written). */
int whirgpu_profile_enable(whirgpu_ctx* ctx, int on);
int whirgpu_profile_read(whirgpu_ctx* ctx, double* out, size_t cap, size_t* len);
/* The same records one launch group at a time, [stage index, device ms, work]
 * each, without resetting them (call before whirgpu_profile_read). */
int whirgpu_profile_records(whirgpu_ctx* ctx, double* out, size_t cap, size_t* len);

/* Device memory for callers that keep inputs resident in HBM (the Rust side
 * would hold such a buffer between proofs).  kind: 0 host->device,
 * 1 device->host, 2 device->device.  No reference counterpart. */
int whirgpu_device_alloc(int device, size_t bytes, void** out);
int whirgpu_device_free(int device, void* ptr);
int whirgpu_memcpy(int device, void* dst, const void* src, size_t bytes, int kind);

/* ---------------- kernel-level entry points (device pointers, blocking) ---------------- */

/* Transpose + zero-pad + column DFT (writer.rs:70-86 with dft_batch):
 * d_src holds 2^m base elements; d_dst receives 2^(m-k+log_blowup) rows of
 * 2^k elements.  whirgpu_lde_ext: the same on extension elements
 * (prover/mod.rs:166-180 with dft_algebra_batch). */
int whirgpu_lde_base(whirgpu_ctx* ctx, const uint32_t* d_src, uint32_t m, uint32_t k, uint32_t log_blowup,
                     uint32_t* d_dst);
int whirgpu_lde_ext(whirgpu_ctx* ctx, const uint32_t* d_src, uint32_t m, uint32_t k, uint32_t log_blowup,
                    uint32_t* d_dst);

/* MerkleTreeMmcs::commit_matrix of `height` rows of `row_len` base elements;
 * d_tree receives 2*height-1 digests (leaves first, root last). */
int whirgpu_merkle_commit(whirgpu_ctx* ctx, const uint32_t* d_rows, uint64_t height, uint32_t row_len,
                          uint32_t* d_tree, uint32_t* root_out);

/* GrindingChallenger::grind (smallest witness); updates *ch as the
 * reference does (observe witness, sample_bits). */
int whirgpu_grind(whirgpu_ctx* ctx, whirgpu_challenger* ch, uint32_t bits, uint32_t* witness_out);

/* EvaluationsList::new_from_point (src/poly/evals.rs:312-320): 2^n extension
 * elements scale * eq(point, x) into d_out. */
int whirgpu_eq_expand(whirgpu_ctx* ctx, const uint32_t* point, uint32_t n, const uint32_t* scale, uint32_t* d_out);

/* evaluate_hypercube_base / _ext (src/poly/evals.rs:334-366). */
int whirgpu_mle_eval_base(whirgpu_ctx* ctx, const uint32_t* d_f, uint32_t n, const uint32_t* point, uint32_t* out);
int whirgpu_mle_eval_ext(whirgpu_ctx* ctx, const uint32_t* d_f, uint32_t n, const uint32_t* point, uint32_t* out);

/* No counterpart in the reference: whirgpu_mle_eval_base for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It takes no context and allocates
 * nothing: its eq tables and block partial sums live in a fixed device array
 * of the library (1,085,440 bytes on each GPU it runs on, held until the process
 * ends), which limits it to 1 <= n <= 24 (larger n: WHIRGPU_BAD_PARAMS).  The
 * eq-weighted sum kernels are those of whirgpu_mle_eval_base.
 * d_f: 2^n base elements on the current device; point: n extension elements
 * (host memory); d_out: 4 uint32_t of device memory that receive the
 * evaluation.  The kernels are enqueued on `stream` (a hipStream_t; NULL for
 * the null stream) and not waited for.  Every call on one GPU uses the same
 * array, so calls that overlap on one GPU (other threads or streams) race. */
int whirgpu_mle_eval_base_noalloc(const uint32_t* d_f, uint32_t n, const uint32_t* point, uint32_t* d_out,
                                  void* stream);

/* sumcheck_coefficients (src/poly/evals.rs:493-526): out = (h(0), h(2)) of
 * arrays of 2^n extension elements. */
int whirgpu_sumcheck_coeffs(whirgpu_ctx* ctx, const uint32_t* d_evals, const uint32_t* d_weights, uint32_t n,
                            uint32_t* out);
/* compress (src/poly/evals.rs:127-148) of both arrays with r; when n >= 2
 * `out` (may be NULL) receives the next round's (h(0), h(2)). */
int whirgpu_sumcheck_fold(whirgpu_ctx* ctx, uint32_t* d_evals, uint32_t* d_weights, uint32_t n, const uint32_t* r,
                          uint32_t* out);

/* ---------------- batched and served proving (no reference counterpart) ----------------
 * A job is one commit + prove with its own number of variables and
 * constraints, the transcript starting from a fresh challenger that has
 * observed the domain separator (benches/whir.rs:139-186).  The job struct is
 * owned by the caller and must stay valid until the job completes. */
typedef struct whirgpu_job {
  uint32_t num_variables;
  const uint32_t* evals;        /* 2^n base elements (Montgomery) */
  int evals_on_device;          /* nonzero: evals is a device pointer */
  const uint32_t* points;       /* npoints * n extension elements */
  uint32_t npoints;
  uint32_t* evals_out;          /* npoints extension elements, or NULL */
  uint8_t* proof_out;           /* caller buffer for the proof bytes */
  size_t proof_cap;
  size_t proof_len;             /* out */
  int status;                   /* out: WHIRGPU_PENDING until done, then a WHIRGPU_* code */
  double queue_ms;              /* out: submit to start */
  double latency_ms;            /* out: submit to completion */
  char error[160];              /* out: whirgpu_last_error() of the failed call ("" on success); the
                                   server runs a job on a worker thread, whose last error the
                                   submitting thread cannot read */
} whirgpu_job;

#define WHIRGPU_PENDING (-1)

typedef struct whirgpu_server whirgpu_server;

/* A pool of `workers` host threads on device `device`, each with its own
 * contexts and HIP stream, taking jobs in submission order.  Admission counts
 * the device memory the workers' contexts really hold (their buffers are kept
 * between jobs) plus what running jobs may still grow by: a job starts when a
 * worker is free, that total plus the job's own growth fits in `hbm_budget`
 * bytes (0: 80% of the free HBM at creation), and its growth fits in the HBM
 * free at that moment.  A job that does not fit makes idle workers release
 * contexts it cannot use; with nothing running a job is always admitted.
 * params->num_variables is ignored (each job has its own).  params->mmcs must
 * be WHIRGPU_MMCS_POSEIDON2 (WHIRGPU_BAD_PARAMS otherwise). */
int whirgpu_server_create(int device, const whirgpu_params* params, const whirgpu_poseidon2* perms, int workers,
                          size_t hbm_budget, whirgpu_server** out);
int whirgpu_server_submit(whirgpu_server* s, whirgpu_job* job);
/* Submits `count` jobs in one queue update (jobs of one shape submitted
 * together can share a group). */
int whirgpu_server_submit_many(whirgpu_server* s, whirgpu_job* const* jobs, size_t count);
/* Largest group a worker forms (1 to 16, default 16): a free worker takes the
 * job at the head of the queue and up to group_max - 1 more queued jobs of the
 * same number of variables, number of constraints and input kind, and proves
 * them with whirgpu_prove_group. */
int whirgpu_server_set_group_max(whirgpu_server* s, int group_max);
/* With `on` nonzero, every proof is verified on the host (whirgpu_verify from
 * a fresh challenger, against the job's points and computed evaluations)
 * before its bytes are written; a rejected proof fails its job with
 * WHIRGPU_VERIFY_FAILED.  Off by default. */
int whirgpu_server_set_self_check(whirgpu_server* s, int on);
/* Launch shapes for the contexts workers create for `num_variables`
 * (whirgpu_set_tuning); a tuning a context refuses leaves its built-in
 * shapes.  Shapes never change a proof. */
int whirgpu_server_set_tuning(whirgpu_server* s, uint32_t num_variables, const whirgpu_tuning* t);
/* Blocks until `job` completes (timeout_ms < 0: no limit); returns its status
 * or WHIRGPU_PENDING on timeout. */
int whirgpu_server_wait(whirgpu_server* s, whirgpu_job* job, int timeout_ms);
/* Jobs queued / running right now, and the device bytes held by the
 * workers' contexts plus the growth reserved by running jobs. */
int whirgpu_server_stats(whirgpu_server* s, size_t* queued, size_t* running, size_t* bytes_in_use);
/* Finishes every submitted job, then stops the workers. */
int whirgpu_server_destroy(whirgpu_server* s);

/* Submits all jobs to a temporary server and waits for all of them. */
int whirgpu_prove_batch(int device, const whirgpu_params* params, const whirgpu_poseidon2* perms, whirgpu_job* jobs,
                        size_t njobs, int workers);

/* Sumcheck of f * eq(point, .) over all n variables: the initial sumcheck of
 * Sumcheck::from_base_evals (src/sumcheck/sumcheck_prover.rs:164-292) for one
 * constraint without grinding, the first l0 rounds by the split (SVO) route.
 * d_f: 2^n base elements on the device.  Challenges: `scripted` (n extension
 * elements) or, if NULL, *ch (observe h(0) and h(2), then sample).
 * out_polys: n pairs (h(0), h(2)); out_final: folded f and eq (8 u32). */
int whirgpu_sumcheck_eq(whirgpu_ctx* ctx, const uint32_t* d_f, uint32_t n, const uint32_t* point, uint32_t l0,
                        const uint32_t* scripted, whirgpu_challenger* ch, uint32_t* out_polys, uint32_t* out_final);

#ifdef __cplusplus
}
#endif

#endif /* WHIR_GPU_H */
