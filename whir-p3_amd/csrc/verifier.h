// Proof decoding and verification on the host: the product's own reading of
// the proof bytes the prover writes, and the WHIR verifier
// (src/whir/verifier/mod.rs, committer/reader.rs) over the configuration,
// Poseidon2 constants and transcript of a context.  Neither uses the device.
#pragma once
#include <string>
#include <vector>

#include "host.h"

namespace whir {

// serde deserialization of WhirProof (src/whir/proof.rs:11-141) from the
// bytes serialize() writes.  Rejects truncated or trailing bytes, unknown
// Option tags or opening types, lengths beyond the input and field words not
// below p.  Returns false with the reason in *err.
// With digests_are_field false (the Keccak MMCS) digest words are any u32.
bool deserialize(const uint8_t* bytes, size_t len, WhirProof* out, std::string* err, bool digests_are_field = true);

// What the verifier checks a proof against.
struct VerifierKey {
  const WhirConfig* cfg = nullptr;
  const kb::P2C16* c16 = nullptr;   // challenger, tree compression, width-16 leaves
  const kb::P2C24* c24 = nullptr;   // width-24 leaves
  int leaf_width = 16;
  bool keccak = false;   // the Keccak MMCS (digests are u64 lanes, not field elements)
};

// Which check failed (verifier/errors.rs VerifierError, plus the shape checks
// a decoded proof must pass before it is indexed).
enum class VerifyFailure {
  None,
  Shape,               // counts in the proof do not match the configuration
  PowWitness,          // VerifierError::InvalidPowWitness
  MerklePath,          // VerifierError::MerkleProofInvalid
  StirOnFinal,         // VerifierError::StirChallengeFailed
  FinalSumcheck,       // VerifierError::SumcheckFailed
};

// CommitmentReader::parse (committer/reader.rs:40-82) followed by
// Verifier::verify (verifier/mod.rs:58-200) for the statement
// {points[i] -> evals[i]} (the order of InitialStatement::normalize).  `ch` is
// the verifier's challenger after the domain separator; on success it ends
// where the prover's ended.
VerifyFailure verify(const VerifierKey& key, const WhirProof& proof, Challenger& ch,
                     const std::vector<std::vector<ef>>& points, const std::vector<ef>& evals, std::string* why);

// MerkleTreeMmcs on the host: the digest of one row (Poseidon2:
// PaddingFreeSponge of width leaf_width, rate leaf_width - 8; Keccak: the
// lane sponge over the row read as little-endian u64) and the check of an
// opening path from the leaf level up.
Digest leaf_digest(const VerifierKey& key, const u32* row, size_t len);
bool path_reaches_root(const VerifierKey& key, Digest node, u64 index, u64 height, const std::vector<Digest>& siblings,
                       const Digest& root);

}  // namespace whir
