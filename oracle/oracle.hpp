// CPU oracle for the WHIR commit/prove path over KoalaBear / EF4 / Poseidon2.
//
// TEST INFRASTRUCTURE ONLY.  This is a plain restatement of the reference
// prover (tcoratger/whir-p3) used as the checker for the GPU product.  Only the
// test suite, the smoke check and the CPU-baseline leg of the benchmark may
// load it.  The product library never links or calls it.
//
// Arithmetic here is deliberately different from the product: elements are
// kept in canonical form [0, p) and multiplied with a 64-bit `%`, whereas the
// product works in Montgomery form.  Values cross the oracle's C API in
// Montgomery form (the in-memory layout of Plonky3's KoalaBear) so that both
// sides read and write the same bytes.
#pragma once
#include <array>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace orc {

using u32 = uint32_t;
using u64 = uint64_t;

// KoalaBear p = 2^31 - 2^24 + 1 (Plonky3 p3-koala-bear).
constexpr u32 P = 0x7f000001u;

inline u32 fadd(u32 a, u32 b) { u32 s = a + b; return s >= P ? s - P : s; }
inline u32 fsub(u32 a, u32 b) { return a >= b ? a - b : a + P - b; }
inline u32 fneg(u32 a) { return a ? P - a : 0; }
inline u32 fmul(u32 a, u32 b) { return (u32)((u64)a * b % P); }
u32 fpow(u32 a, u64 e);
inline u32 finv(u32 a) { return fpow(a, P - 2); }
u32 to_monty(u32 canon);
u32 from_monty(u32 monty);

// Generator of the 2^bits-th roots of unity: 3^(127 * 2^(24-bits)); 3 generates F*.
u32 two_adic_generator(int bits);

// Degree-4 binomial extension F[x]/(x^4 - 3), coefficients c0..c3.
struct EF {
  u32 c[4];
};
constexpr u32 EF_W = 3;
inline EF ef_zero() { return EF{{0, 0, 0, 0}}; }
inline EF ef_one() { return EF{{1, 0, 0, 0}}; }
inline EF ef_base(u32 a) { return EF{{a, 0, 0, 0}}; }
inline bool ef_eq(const EF& a, const EF& b) { return !memcmp(a.c, b.c, 16); }
inline EF ef_add(const EF& a, const EF& b) {
  return EF{{fadd(a.c[0], b.c[0]), fadd(a.c[1], b.c[1]), fadd(a.c[2], b.c[2]), fadd(a.c[3], b.c[3])}};
}
inline EF ef_sub(const EF& a, const EF& b) {
  return EF{{fsub(a.c[0], b.c[0]), fsub(a.c[1], b.c[1]), fsub(a.c[2], b.c[2]), fsub(a.c[3], b.c[3])}};
}
inline EF ef_scale(const EF& a, u32 s) {
  return EF{{fmul(a.c[0], s), fmul(a.c[1], s), fmul(a.c[2], s), fmul(a.c[3], s)}};
}
EF ef_mul(const EF& a, const EF& b);
inline EF ef_double(const EF& a) { return ef_add(a, a); }
EF ef_pow(EF a, u64 e);
EF ef_inv(const EF& a);

// ---------------------------------------------------------------- Poseidon2
// Plonky3 Poseidon2 over KoalaBear: x^3 S-box, 8 external rounds (4 + 4),
// 20 (width 16) or 23 (width 24) internal rounds, M4-based external layer,
// internal layer 1 + diag(V).  Round constants are data.
struct Poseidon2 {
  int width = 16;
  int rounds_p = 20;
  std::vector<u32> ext_init;   // 4 * width
  std::vector<u32> ext_term;   // 4 * width
  std::vector<u32> internal;   // rounds_p
  std::vector<u32> diag;       // width
  void permute(u32* state) const;
};
// Constants drawn the way Plonky3's `new_from_rng_128` draws them from
// `SmallRng::seed_from_u64(seed)` (xoshiro256++ seeded through SplitMix64).
// When `both` is set the width-24 instance is drawn from the same stream right
// after the width-16 one (as `src/bin/main.rs:100-102` does).
Poseidon2 poseidon2_from_seed(u64 seed, int width);
void poseidon2_pair_from_seed(u64 seed, Poseidon2* p16, Poseidon2* p24);

// xoshiro256++ seeded via SplitMix64 (rand's SmallRng on 64-bit targets).
struct SmallRng {
  u64 s[4];
  explicit SmallRng(u64 seed);
  u64 next_u64();
  u32 next_u32() { return (u32)(next_u64() >> 32); }
  // StandardUniform for MontyField31: rejection-sample a u31 and read it as a
  // Montgomery representative.  Returns the canonical value.
  u32 field();
  EF ext() { EF e; for (int k = 0; k < 4; k++) e.c[k] = field(); return e; }
};

// ---------------------------------------------------------------- Keccak
// Keccak-f[1600] on lanes a[x + 5y] and Keccak-256 (original padding), for
// the Keccak MMCS variant (src/whir/mod.rs:236-383).
void keccak_permute(u64* lanes25);
void keccak256_bytes(const uint8_t* msg, size_t len, uint8_t out[32], uint8_t pad = 0x01);

using Digest = std::array<u32, 8>;

// ---------------------------------------------------------------- challenger
// DuplexChallenger<F, Poseidon2-16, 16, 8>; with `keccak` set,
// SerializingChallenger32<F, HashChallenger<u8, Keccak256Hash, 32>> instead:
// a field element is observed as the 4 little-endian bytes of its (canonical)
// value, a sample with no output bytes left hashes the input bytes (which are
// then replaced by the hash: chaining) and output bytes are taken from the end;
// a field sample is 4 bytes as a little-endian u32, masked to 31 bits,
// rejected until below p.
struct Challenger {
  const Poseidon2* perm = nullptr;
  u32 state[16] = {0};
  std::vector<u32> in, out;
  bool keccak = false;
  std::vector<uint8_t> kin, kout;
  void duplex();
  uint8_t sample_byte();
  void observe(u32 v);
  void observe_digest(const Digest& d);   // field elements, or (Keccak) the words' bytes
  void observe_ef(const EF& e) { for (int k = 0; k < 4; k++) observe(e.c[k]); }
  u32 sample();
  EF sample_ef() { EF e; for (int k = 0; k < 4; k++) e.c[k] = sample(); return e; }
  u64 sample_bits(int bits);   // Keccak: four raw bytes, no rejection
  bool check_witness(int bits, u32 w);
  u32 grind(int bits);
};

// ---------------------------------------------------------------- Merkle
Digest sponge_hash(const Poseidon2& perm, const u32* vals, size_t len);  // PaddingFreeSponge<P,W,W-8... >
Digest compress(const Poseidon2& p16, const Digest& l, const Digest& r);
struct MerkleTree {
  size_t height = 0, width = 0;             // leaves x base elements per leaf
  std::vector<u32> rows;                    // height * width base elements
  std::vector<std::vector<Digest>> layers;  // layers[0] = leaves ... back() = root
  Digest root() const { return layers.back()[0]; }
};
struct MerkleHashers {
  const Poseidon2* leaf;   // width 16 (rate 8) or 24 (rate 16)
  const Poseidon2* node;   // width 16
  bool keccak = false;     // the Keccak MMCS: digests are 4 u64 lanes (8 raw words)
};
MerkleTree merkle_commit(const MerkleHashers& h, std::vector<u32> rows, size_t height, size_t width);
void merkle_open(const MerkleTree& t, size_t index, std::vector<u32>* row, std::vector<Digest>* path);
bool merkle_verify(const MerkleHashers& h, const Digest& root, size_t height, size_t index,
                   const std::vector<u32>& row, const std::vector<Digest>& path);

// ---------------------------------------------------------------- DFT
void dft_naive(std::vector<u32>& a);          // O(N^2), for checking
void dft_radix2(std::vector<u32>& a);         // natural order in, natural order out
// Column DFTs of a row-major matrix (height x width, base field), in place.
void dft_batch(std::vector<u32>& mat, size_t height, size_t width);

// ---------------------------------------------------------------- polynomials
// Evaluation lists: index bit (n-1-j) <-> variable j (variable 0 is the MSB).
std::vector<EF> eq_table(const std::vector<EF>& point, const EF& scale);
EF mle_eval_base(const std::vector<u32>& f, const std::vector<EF>& point);
EF mle_eval_ext(const std::vector<EF>& f, const std::vector<EF>& point);
std::vector<EF> expand_from_univariate(const EF& y, int n);
void sumcheck_coeffs(const std::vector<EF>& e, const std::vector<EF>& w, EF* c0, EF* c2);
void fold_ext(std::vector<EF>& a, const EF& r);
std::vector<EF> fold_base(const std::vector<u32>& a, const EF& r);
EF extrapolate_012(const EF& e0, const EF& e1, const EF& e2, const EF& r);
std::vector<EF> lagrange_weights_012_multi(const std::vector<EF>& rs);

// ---------------------------------------------------------------- config
enum Soundness { UniqueDecoding = 0, JohnsonBound = 1, CapacityBound = 2 };
struct Params {
  int num_variables = 18;
  int security_level = 100;
  int pow_bits = 16;
  int ff0 = 4, ff = 4;              // FoldingFactor (first round, later rounds)
  int log_inv_rate = 1;
  int rs_reduction = 3;
  int soundness = CapacityBound;
  int svo = 1;                       // SumcheckStrategy::Svo
  int leaf_width = 16;
  int digest_elems_pattern = 8;      // DIGEST_ELEMS in the domain separator
  int mmcs = 0;                      // 0 Poseidon2, 1 Keccak (MMCS and challenger)
};
struct RoundCfg {
  int pow_bits, folding_pow_bits, num_queries, ood_samples, num_variables, folding_factor;
  u64 domain_size;
  u32 folded_domain_gen;
};
struct Config {
  Params p;
  int commitment_ood_samples;
  int starting_folding_pow_bits;
  std::vector<RoundCfg> rounds;
  int final_queries, final_pow_bits, final_sumcheck_rounds, final_folding_pow_bits;
  int ff_at(int round) const { return round == 0 ? p.ff0 : p.ff; }
  int n_rounds() const { return (int)rounds.size(); }
  int total_number(int n_r) const { return p.ff0 + p.ff * n_r; }
  int rs_reduction_factor(int round) const { return round == 0 ? p.rs_reduction : 1; }
  u64 inv_rate(int round) const;
  RoundCfg final_round_config() const;
};
bool make_config(const Params& p, Config* out, std::string* err);
std::vector<u32> domain_separator(const Config& c);
std::vector<u64> stir_queries(Challenger& ch, u64 domain_size, int folding_factor, int num_queries);

// ---------------------------------------------------------------- proof
struct SumcheckData {
  std::vector<std::array<EF, 2>> evals;
  std::vector<u32> pow;
};
struct QueryOpening {
  bool ext = false;
  std::vector<u32> base_values;
  std::vector<EF> ext_values;
  std::vector<Digest> path;
};
struct RoundProof {
  Digest commitment;
  std::vector<EF> ood;
  u32 pow_witness = 0;
  std::vector<QueryOpening> queries;
  SumcheckData sumcheck;
};
struct Proof {
  Digest initial_commitment;
  std::vector<EF> initial_ood;
  SumcheckData initial_sumcheck;
  std::vector<RoundProof> rounds;
  std::vector<EF> final_poly;
  u32 final_pow_witness = 0;
  std::vector<QueryOpening> final_queries;
  bool has_final_sumcheck = false;
  SumcheckData final_sumcheck;
};
// raw_digests: digest words are written as they are (Keccak lanes) instead of
// as field elements in Montgomery form.
std::vector<uint8_t> serialize_proof(const Proof& p, bool raw_digests = false);
bool deserialize_proof(const uint8_t* buf, size_t len, Proof* p, bool raw_digests = false);

// ---------------------------------------------------------------- protocol
struct Statement {                      // InitialStatement (normalized view)
  std::vector<std::vector<EF>> points;
  std::vector<EF> evals;
};
// Wall time of commit and prove, and of the stages under the reference's
// tracing span names: "dft" and "commit_matrix" of the commitment
// (src/whir/committer/writer.rs:85-89), its "ood evaluation" (:102), the
// initial sumcheck ("svo rounds", src/sumcheck/sumcheck_prover.rs:201), the
// WHIR rounds (src/whir/prover/mod.rs:130) and the final round (:328).
struct Timing {
  double commit_ms = 0, prove_ms = 0;
  double dft_ms = 0, commit_matrix_ms = 0, ood_ms = 0, initial_sumcheck_ms = 0, rounds_ms = 0, final_ms = 0;
};
// Full commit + prove.  `user_points` are evaluated first (InitialStatement::evaluate),
// then commit appends the OOD constraints.  Returns the proof; `stmt_out` receives the
// user constraints only (what the verifier is given).
Proof whir_prove(const Config& cfg, const Poseidon2& p16, const Poseidon2& leaf,
                 const std::vector<u32>& poly, const std::vector<std::vector<EF>>& user_points,
                 Challenger& ch, Statement* stmt_out, Timing* t);
bool whir_verify(const Config& cfg, const Poseidon2& p16, const Poseidon2& leaf, const Proof& proof,
                 const Statement& stmt, Challenger& ch, std::string* err);

}  // namespace orc
