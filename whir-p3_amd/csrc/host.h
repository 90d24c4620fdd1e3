// Host-side protocol pieces of the WHIR prover: parameter derivation, the
// Fiat-Shamir domain separator, the duplex challenger, STIR query sampling,
// the proof object and its byte encoding, and small extension-field helpers.
// All field values are Montgomery-form u32 (see field.h).
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "field.h"
#include "keccak.h"
#include "poseidon2.h"

namespace whir {

using kb::ef;
using kb::u32;
using kb::u64;

// ---------------------------------------------------------------- parameters
// FoldingFactor (src/parameters/mod.rs:31-143): Constant(k) when first == rest,
// otherwise ConstantFromSecondRound(first, rest).
struct FoldingFactor {
  int first = 4, rest = 4;
  bool constant() const { return first == rest; }
  int at_round(int r) const { return r == 0 ? first : rest; }
  bool check_validity(int n, std::string* err) const;
  void compute_number_of_rounds(int n, int* rounds, int* final_rounds) const;
  int total_number(int n_rounds) const { return first + rest * n_rounds; }
};

enum Soundness { UniqueDecoding = 0, JohnsonBound = 1, CapacityBound = 2 };

// ProtocolParameters (src/parameters/mod.rs:147-166) without the MMCS, which
// is fixed to Poseidon2 here; leaf_width selects the leaf sponge.
struct ProtocolParameters {
  int starting_log_inv_rate = 1;
  int rs_domain_initial_reduction_factor = 1;
  FoldingFactor folding_factor;
  int soundness_type = CapacityBound;
  int security_level = 100;
  int pow_bits = 16;
};

struct RoundConfig {
  int pow_bits, folding_pow_bits, num_queries, ood_samples, num_variables, folding_factor;
  u64 domain_size;
  u32 folded_domain_gen;
};

// WhirConfig (src/whir/parameters.rs:43-234, 372-416).
struct WhirConfig {
  int num_variables = 0;
  ProtocolParameters params;
  int commitment_ood_samples = 0;
  int starting_folding_pow_bits = 0;
  std::vector<RoundConfig> round_parameters;
  int final_queries = 0, final_pow_bits = 0, final_sumcheck_rounds = 0, final_folding_pow_bits = 0;

  static bool create(int n, const ProtocolParameters& p, WhirConfig* out, std::string* err);
  int n_rounds() const { return (int)round_parameters.size(); }
  int ff(int r) const { return params.folding_factor.at_round(r); }
  u64 starting_domain_size() const { return 1ull << (num_variables + params.starting_log_inv_rate); }
  int rs_reduction_factor(int r) const { return r == 0 ? params.rs_domain_initial_reduction_factor : 1; }
  u64 inv_rate(int r) const;
  RoundConfig final_round_config() const;
};

// DomainSeparator::commit_statement + add_whir_proof
// (src/fiat_shamir/domain_separator.rs:95-169).
std::vector<u32> domain_separator_pattern(const WhirConfig& c, int digest_elems);

// ---------------------------------------------------------------- challenger
// DuplexChallenger<F, Poseidon2-16, WIDTH 16, RATE 8> (p3-challenger).
struct DuplexChallenger {
  const kb::P2C16* perm = nullptr;
  u32 sponge[16] = {0};
  u32 in[8] = {0};
  int in_len = 0;
  u32 out[8] = {0};
  int out_len = 0;

  void duplexing();
  void observe(u32 v);
  void observe_ef(const ef& e) { for (int k = 0; k < 4; k++) observe(e.c[k]); }
  void observe_slice(const u32* v, size_t n) { for (size_t i = 0; i < n; i++) observe(v[i]); }
  u32 sample();
  ef sample_ef() { ef e; for (int k = 0; k < 4; k++) e.c[k] = sample(); return e; }
  u64 sample_bits(int bits) { return (u64)kb::from_monty(sample()) & ((1ull << bits) - 1); }
  bool check_witness(int bits, u32 w_monty) { observe(w_monty); return sample_bits(bits) == 0; }
};

// SerializingChallenger32<F, HashChallenger<u8, Keccak256Hash, 32>>
// (p3-challenger), the challenger of the Keccak MMCS variant
// (src/whir/mod.rs:257).  The inner HashChallenger keeps a byte input buffer
// and a 32-byte output buffer: observing a byte clears the output and appends
// to the input; sampling a byte with an empty output flushes (Keccak-256 of
// the whole input, which is then replaced by the 32 output bytes: chaining)
// and pops the output from its end.  A field element is observed as the four
// little-endian bytes of its canonical value and sampled by rejection: four
// sampled bytes as a little-endian u32, masked to 31 bits, kept when below p.
// sample_bits takes four sampled bytes as a little-endian u32 and masks it to
// the bits asked for, with no rejection (so PoW checks and query indices read
// raw bytes).
struct KeccakChallenger {
  static constexpr size_t INPUT_CAP = 4096;   // bytes between two flushes (the C ABI's buffer)
  std::vector<uint8_t> in;
  uint8_t out[32] = {0};
  int out_len = 0;
  void observe_byte(uint8_t b);
  uint8_t sample_byte();
  u32 sample_u32();   // four sampled bytes, little endian
  void observe_field(u32 monty);
  u32 sample_field();   // Montgomery form
};

// The transcript of a proof: the Poseidon2 duplex challenger, or the Keccak
// challenger when the context's MMCS is Keccak.  Field elements are
// Montgomery-form words either way; a digest is observed as field elements
// (Poseidon2) or as its bytes (Keccak: four u64 lanes, little endian).
struct Challenger {
  bool keccak = false;
  DuplexChallenger dx;
  KeccakChallenger kc;
  void observe(u32 v) {
    if (keccak) kc.observe_field(v);
    else dx.observe(v);
  }
  void observe_ef(const ef& e) { for (int k = 0; k < 4; k++) observe(e.c[k]); }
  void observe_slice(const u32* v, size_t n) { for (size_t i = 0; i < n; i++) observe(v[i]); }
  void observe_digest(const std::array<u32, 8>& d);
  u32 sample() { return keccak ? kc.sample_field() : dx.sample(); }
  ef sample_ef() { ef e; for (int k = 0; k < 4; k++) e.c[k] = sample(); return e; }
  u64 sample_bits(int bits) {
    const u64 r = keccak ? (u64)kc.sample_u32() : (u64)kb::from_monty(dx.sample());
    return r & ((1ull << bits) - 1);
  }
  bool check_witness(int bits, u32 w_monty) { observe(w_monty); return sample_bits(bits) == 0; }
};

// get_challenge_stir_queries (src/whir/utils.rs:24-145).
std::vector<u64> get_challenge_stir_queries(u64 domain_size, int folding_factor, int num_queries, Challenger& ch);

// ---------------------------------------------------------------- RNG
// rand's SmallRng on 64-bit targets: xoshiro256++ seeded through SplitMix64.
struct SmallRng {
  u64 s[4];
  explicit SmallRng(u64 seed);
  u64 next_u64();
  u32 next_u32() { return (u32)(next_u64() >> 32); }
  // StandardUniform for MontyField31: a rejection-sampled u31 taken as the
  // Montgomery representative.
  u32 field_monty();
};
// Poseidon2::new_from_rng_128: external initial, external terminal, then
// internal constants, drawn in that order; width 16 first, then width 24.
void poseidon2_constants_from_seed(u64 seed, kb::P2C16* c16, kb::P2C24* c24);

// ---------------------------------------------------------------- proof
using Digest = std::array<u32, 8>;
struct SumcheckData {
  std::vector<std::array<ef, 2>> polynomial_evaluations;
  std::vector<u32> pow_witnesses;
};
struct QueryOpening {
  bool extension = false;
  std::vector<u32> values;   // base: one u32 per value; extension: four per value
  std::vector<Digest> proof;
};
struct WhirRoundProof {
  Digest commitment{};
  std::vector<ef> ood_answers;
  u32 pow_witness = 0;
  std::vector<QueryOpening> queries;
  SumcheckData sumcheck;
};
struct WhirProof {
  bool has_initial_commitment = false;
  Digest initial_commitment{};
  std::vector<ef> initial_ood_answers;
  SumcheckData initial_sumcheck;
  std::vector<WhirRoundProof> rounds;
  bool has_final_poly = false;
  std::vector<ef> final_poly;
  u32 final_pow_witness = 0;
  std::vector<QueryOpening> final_queries;
  bool has_final_sumcheck = false;
  SumcheckData final_sumcheck;
};
// Field values as their Montgomery u32, lengths as u64 and Option tags as one
// byte, all little endian; QueryOpening as {type: "base"|"extension", values, proof}.
std::vector<uint8_t> serialize(const WhirProof& p);
size_t serialized_size(const WhirProof& p);
void serialize_into(const WhirProof& p, uint8_t* out);   // serialized_size(p) bytes

// ---------------------------------------------------------------- EF helpers
std::vector<ef> eq_table_host(const std::vector<ef>& z, const ef& scale);
ef mle_host(const std::vector<ef>& f, const std::vector<ef>& point);
std::vector<ef> expand_from_univariate(const ef& y, int n);
ef extrapolate_012(const ef& e0, const ef& e1, const ef& e2, const ef& r);
std::vector<ef> powers(const ef& g, size_t n);
// eq(z, u) extended to u in {0,1,2}: prod (z u + (1 - z)(1 - u)).
ef eq_at(const std::vector<ef>& z, const std::vector<ef>& u);

}  // namespace whir
