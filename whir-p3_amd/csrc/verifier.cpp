// Proof decoding and the WHIR verifier on the host (see verifier.h).
//
// The verifier replays the prover's transcript on its own challenger and
// checks, in the reference's order (src/whir/verifier/mod.rs:58-200):
// - each commitment's root observed, its OOD challenges sampled and answers
//   observed (CommitmentReader::parse, committer/reader.rs:40-82);
// - every sumcheck round's (h(0), h(2)) observed, its proof of work checked
//   and the claim carried to h(r) (verify_sumcheck_rounds, verifier/sumcheck.rs);
// - each round's proof of work, STIR query indices, Merkle openings of the
//   previous commitment and the openings' folds (verify_stir_challenges,
//   verifier/mod.rs:227-304);
// - the final polynomial against the last STIR folds (SelectStatement::verify,
//   constraints/statement/select.rs:246-254) and the last claim against the
//   combined weights at the sumcheck point times the final polynomial there
//   (ConstraintPolyEvaluator, constraints/evaluator.rs).
// Counts the reference would index past (openings, sumcheck rounds, OOD
// answers) are checked first and reported as a shape failure.
#include "verifier.h"

#include <algorithm>
#include <cstring>

namespace whir {

using kb::P;

namespace {

// ------------------------------------------------------------------ decoding
// A cursor over the proof bytes.  The first failure is kept and ends every
// later read, so the decoder's code reads straight through.
class Cursor {
 public:
  Cursor(const uint8_t* b, size_t n, bool digests_are_field) : p_(b), end_(b + n), digest_field_(digests_are_field) {}
  bool failed() const { return !why_.empty(); }
  const std::string& why() const { return why_; }
  bool at_end() const { return p_ == end_; }
  void fail(const std::string& w) {
    if (why_.empty()) why_ = w;
    p_ = end_;
  }
  bool raw(void* dst, size_t n, const char* what) {
    if (failed()) return false;
    if (n > left()) {
      fail(std::string("truncated in ") + what);
      return false;
    }
    memcpy(dst, p_, n);
    p_ += n;
    return true;
  }
  // Option<T>: one byte, 0 = None, 1 = Some.
  bool some(const char* what) {
    uint8_t t = 0;
    if (!raw(&t, 1, what)) return false;
    if (t > 1) fail(std::string("bad Option tag in ") + what);
    return t == 1;
  }
  // A u64 sequence length.  Every element takes at least `unit` bytes, so a
  // length the rest of the input cannot hold is refused before allocating.
  size_t length(size_t unit, const char* what) {
    u64 n = 0;
    if (!raw(&n, 8, what)) return 0;
    if (n > left() / std::max<size_t>(unit, 1)) {
      fail(std::string("length beyond the input in ") + what);
      return 0;
    }
    return (size_t)n;
  }
  // A field element: its Montgomery word, which must be below p.
  u32 field(const char* what) {
    u32 w = 0;
    if (raw(&w, 4, what) && w >= P) fail(std::string("field word not below p in ") + what);
    return w;
  }
  void fields(u32* dst, size_t n, const char* what) {
    for (size_t i = 0; i < n && !failed(); i++) dst[i] = field(what);
  }
  ef ext(const char* what) {
    ef e = kb::ef_zero();
    fields(e.c, 4, what);
    return e;
  }
  // A digest: 8 field elements (Poseidon2), or 4 u64 lanes (Keccak) as 8 words.
  Digest digest(const char* what) {
    Digest d{};
    if (digest_field_) {
      fields(d.data(), d.size(), what);
    } else {
      for (u32& w : d) raw(&w, 4, what);
    }
    return d;
  }
  std::string text(const char* what) {
    const size_t n = length(1, what);
    std::string s(n, '\0');
    if (n) raw(&s[0], n, what);
    return s;
  }

 private:
  size_t left() const { return (size_t)(end_ - p_); }
  const uint8_t* p_;
  const uint8_t* end_;
  bool digest_field_;
  std::string why_;
};

void read_exts(Cursor& c, std::vector<ef>* v, const char* what) {
  v->resize(c.length(16, what));
  for (ef& e : *v) e = c.ext(what);
}

void read_sumcheck(Cursor& c, SumcheckData* s) {
  s->polynomial_evaluations.resize(c.length(32, "sumcheck polynomials"));
  for (auto& pe : s->polynomial_evaluations) {
    pe[0] = c.ext("sumcheck polynomials");
    pe[1] = c.ext("sumcheck polynomials");
  }
  s->pow_witnesses.resize(c.length(4, "sumcheck witnesses"));
  c.fields(s->pow_witnesses.data(), s->pow_witnesses.size(), "sumcheck witnesses");
}

// Vec<QueryOpening>: each opening is internally tagged with its type name,
// then its values and its authentication path.
void read_openings(Cursor& c, std::vector<QueryOpening>* qs) {
  // smallest encoding of one opening: type length + "base" + two lengths
  qs->resize(c.length(8 + 4 + 8 + 8, "openings"));
  for (QueryOpening& q : *qs) {
    const std::string type = c.text("opening type");
    if (type == "base") {
      q.extension = false;
    } else if (type == "extension") {
      q.extension = true;
    } else {
      c.fail("unknown opening type");
      return;
    }
    const size_t words = q.extension ? 4 : 1;
    q.values.resize(c.length(4 * words, "opened values") * words);
    c.fields(q.values.data(), q.values.size(), "opened values");
    q.proof.resize(c.length(32, "opening path"));
    for (Digest& d : q.proof) d = c.digest("opening path");
    if (c.failed()) return;
  }
}

// ------------------------------------------------------------------ Merkle
template <int W>
Digest sponge(const u32* row, size_t len, const kb::P2Consts<W>& consts) {
  constexpr size_t RATE = W - 8;
  u32 s[W] = {};
  for (size_t at = 0; at < len; at += RATE) {
    const size_t k = std::min(RATE, len - at);
    std::copy(row + at, row + at + k, s);   // overwrite mode: the rest of the state is kept
    kb::permute<W>(s, consts);
  }
  Digest d;
  std::copy(s, s + 8, d.begin());
  return d;
}

// The Keccak lane sponge over a run of u32 words read as little-endian u64
// lanes (an odd last word alone), as 8 digest words.
Digest keccak_words(const u32* w, size_t n) {
  std::vector<uint64_t> lanes((n + 1) / 2);
  for (size_t i = 0; i < n; i++) lanes[i / 2] |= (uint64_t)w[i] << (32 * (i % 2));
  uint64_t d4[4];
  kb::keccak_lane_sponge(lanes.data(), lanes.size(), d4);
  Digest d;
  for (int j = 0; j < 4; j++) {
    d[2 * j] = (u32)d4[j];
    d[2 * j + 1] = (u32)(d4[j] >> 32);
  }
  return d;
}

Digest compress(const Digest& l, const Digest& r, const kb::P2C16& consts) {
  u32 s[16];
  std::copy(l.begin(), l.end(), s);
  std::copy(r.begin(), r.end(), s + 8);
  kb::permute<16>(s, consts);
  Digest d;
  std::copy(s, s + 8, d.begin());
  return d;
}

// ------------------------------------------------------------------ verifier
// One Constraint of the verifier (constraints/mod.rs): the eq points and
// select variables of one round over the variables it had, and its challenge.
// Weights are gamma^i over the eq points, then over the select variables.
struct RoundConstraint {
  int nv = 0;
  std::vector<std::vector<ef>> eq_points;
  std::vector<u32> sel_vars;   // base-field points (Montgomery)
  ef gamma = kb::ef_zero();
};

// Constraint::combine_evals: claimed += sum_i gamma^i e_i over the eq
// answers, then the select answers.
void combine_claims(ef* claimed, const ef& gamma, const std::vector<ef>& eq_evals, const std::vector<ef>& sel_evals) {
  ef g = kb::ef_one();
  for (const auto* v : {&eq_evals, &sel_evals}) {
    for (const ef& e : *v) {
      *claimed = kb::ef_add(*claimed, kb::ef_mul(g, e));
      g = kb::ef_mul(g, gamma);
    }
  }
}

// select(expand_from_univariate(var, n), x) = prod_j (x_j (var^(2^(n-1-j)) - 1) + 1)
// (MultilinearPoint::select_poly, src/poly/multilinear.rs:136-143).
ef select_at(u32 var, const ef* x, int n) {
  ef acc = kb::ef_one();
  u32 l = var;   // var^(2^(n-1-j)), from j = n - 1 down
  for (int j = n - 1; j >= 0; j--) {
    const ef t = kb::ef_mul_base(x[j], kb::sub(l, kb::ONE));
    acc = kb::ef_mul(acc, kb::ef_add(t, kb::ef_one()));
    l = kb::mul(l, l);
  }
  return acc;
}

// eval_round (constraints/evaluator.rs): a constraint over nv variables is
// evaluated at the last nv sumcheck challenges, in the order they were drawn.
ef constraint_weight(const RoundConstraint& c, const std::vector<ef>& challenges) {
  const ef* x = challenges.data() + (challenges.size() - (size_t)c.nv);
  const std::vector<ef> xs(x, x + c.nv);
  ef g = kb::ef_one(), acc = kb::ef_zero();
  for (const auto& pt : c.eq_points) {
    acc = kb::ef_add(acc, kb::ef_mul(g, eq_at(pt, xs)));
    g = kb::ef_mul(g, c.gamma);
  }
  for (u32 v : c.sel_vars) {
    acc = kb::ef_add(acc, kb::ef_mul(g, select_at(v, x, c.nv)));
    g = kb::ef_mul(g, c.gamma);
  }
  return acc;
}

// sum_i coeffs[i] var^i by Horner from the last coefficient (SelectStatement::verify).
ef horner(const std::vector<ef>& coeffs, u32 var) {
  ef acc = kb::ef_zero();
  for (size_t i = coeffs.size(); i-- > 0;) acc = kb::ef_add(kb::ef_mul_base(acc, var), coeffs[i]);
  return acc;
}

class Check {
 public:
  Check(const VerifierKey& k, Challenger& ch, std::string* why) : k_(k), cfg_(*k.cfg), ch_(ch), why_(why) {}

  VerifyFailure run(const WhirProof& P, const std::vector<std::vector<ef>>& points, const std::vector<ef>& evals);

 private:
  struct Failed {
    VerifyFailure f;
  };
  [[noreturn]] void fail(VerifyFailure f, const std::string& why) {
    *why_ = why;
    throw Failed{f};
  }
  void observe_root(const Digest& d) { ch_.observe_digest(d); }
  void shape(const WhirProof& P, const std::vector<std::vector<ef>>& points, const std::vector<ef>& evals);
  std::vector<ef> sumcheck(const SumcheckData& s, int rounds, int pow_bits, const char* where);
  void pow(int bits, u32 witness, const char* where);
  void stir(u64 domain_size, int ff, int num_queries, u32 folded_gen, const Digest& root,
            const std::vector<QueryOpening>& qs, const std::vector<ef>& fold_rand, std::vector<u32>* vars,
            std::vector<ef>* folds, const char* where);

  const VerifierKey& k_;
  const WhirConfig& cfg_;
  Challenger& ch_;
  std::string* why_;
  ef claimed_ = kb::ef_zero();
};

// The counts the rest of the check indexes by, against the configuration.
void Check::shape(const WhirProof& P, const std::vector<std::vector<ef>>& points, const std::vector<ef>& evals) {
  const int n = cfg_.num_variables;
  if (points.size() != evals.size()) fail(VerifyFailure::Shape, "points and evaluations differ in number");
  for (const auto& pt : points)
    if ((int)pt.size() != n) fail(VerifyFailure::Shape, "a statement point has the wrong number of variables");
  if (!P.has_initial_commitment) fail(VerifyFailure::Shape, "the proof has no initial commitment");
  if ((int)P.initial_ood_answers.size() != cfg_.commitment_ood_samples)
    fail(VerifyFailure::Shape, "initial OOD answers: wrong count");
  if ((int)P.rounds.size() != cfg_.n_rounds()) fail(VerifyFailure::Shape, "wrong number of rounds");
  for (int r = 0; r < cfg_.n_rounds(); r++)
    if ((int)P.rounds[r].ood_answers.size() != cfg_.round_parameters[r].ood_samples)
      fail(VerifyFailure::Shape, "round " + std::to_string(r) + " OOD answers: wrong count");
  int left = n - cfg_.ff(0);
  for (int r = 0; r < cfg_.n_rounds(); r++) left -= cfg_.ff(r + 1);
  if (left < 0 || left != cfg_.final_sumcheck_rounds) fail(VerifyFailure::Shape, "inconsistent folding schedule");
  if (!P.has_final_poly) fail(VerifyFailure::Shape, "the proof has no final polynomial");
  if (P.final_poly.size() != (size_t)1 << left) fail(VerifyFailure::Shape, "final polynomial: wrong size");
  if (left > 0 && !P.has_final_sumcheck) fail(VerifyFailure::Shape, "the proof has no final sumcheck");
}

void Check::pow(int bits, u32 witness, const char* where) {
  if (bits > 0 && !ch_.check_witness(bits, witness))
    fail(VerifyFailure::PowWitness, std::string("proof-of-work witness rejected in ") + where);
}

// verify_sumcheck_rounds (verifier/sumcheck.rs:12-41): the round challenges.
std::vector<ef> Check::sumcheck(const SumcheckData& s, int rounds, int pow_bits, const char* where) {
  if ((int)s.polynomial_evaluations.size() != rounds)
    fail(VerifyFailure::Shape, std::string("wrong number of sumcheck rounds in ") + where);
  if (pow_bits > 0 && (int)s.pow_witnesses.size() < rounds)
    fail(VerifyFailure::Shape, std::string("missing sumcheck witnesses in ") + where);
  std::vector<ef> rs;
  for (int i = 0; i < rounds; i++) {
    const ef& h0 = s.polynomial_evaluations[i][0];
    const ef& h2 = s.polynomial_evaluations[i][1];
    ch_.observe_ef(h0);
    ch_.observe_ef(h2);
    pow(pow_bits, pow_bits > 0 ? s.pow_witnesses[i] : 0, where);
    const ef r = ch_.sample_ef();
    claimed_ = extrapolate_012(h0, kb::ef_sub(claimed_, h0), h2, r);   // h(1) = claim - h(0)
    rs.push_back(r);
  }
  return rs;
}

// verify_stir_challenges after its proof-of-work check (verifier/mod.rs:262-304).
void Check::stir(u64 domain_size, int ff, int num_queries, u32 folded_gen, const Digest& root,
                 const std::vector<QueryOpening>& qs, const std::vector<ef>& fold_rand, std::vector<u32>* vars,
                 std::vector<ef>* folds, const char* where) {
  const std::vector<u64> idx = get_challenge_stir_queries(domain_size, ff, num_queries, ch_);
  if (qs.size() != idx.size()) fail(VerifyFailure::Shape, std::string("wrong number of openings in ") + where);
  const u64 height = domain_size >> ff;
  const size_t width = (size_t)1 << ff;
  std::vector<ef> row(width);
  for (size_t q = 0; q < idx.size(); q++) {
    const QueryOpening& o = qs[q];
    if ((o.extension ? o.values.size() / 4 : o.values.size()) != width)
      fail(VerifyFailure::Shape, std::string("an opened row has the wrong width in ") + where);
    const Digest leaf = leaf_digest(k_, o.values.data(), o.values.size());
    if (!path_reaches_root(k_, leaf, idx[q], height, o.proof, root))
      fail(VerifyFailure::MerklePath, std::string("Merkle opening at position ") + std::to_string(idx[q]) + " in " +
                                          where + " does not reach the root");
    for (size_t i = 0; i < width; i++) {
      if (o.extension)
        std::copy(o.values.begin() + 4 * i, o.values.begin() + 4 * i + 4, row[i].c);
      else
        row[i] = kb::ef_base(o.values[i]);
    }
    folds->push_back(mle_host(row, fold_rand));
    vars->push_back(kb::pow(folded_gen, idx[q]));
  }
}

VerifyFailure Check::run(const WhirProof& P, const std::vector<std::vector<ef>>& points,
                         const std::vector<ef>& evals) {
  try {
    shape(P, points, evals);
    const int n = cfg_.num_variables;
    std::vector<RoundConstraint> constraints;
    std::vector<ef> challenges;   // every sumcheck challenge, in order

    // the initial commitment and the statement plus its OOD constraints
    observe_root(P.initial_commitment);
    RoundConstraint c0;
    c0.nv = n;
    c0.eq_points = points;
    std::vector<ef> eq_evals = evals;
    for (int i = 0; i < cfg_.commitment_ood_samples; i++) {
      c0.eq_points.push_back(expand_from_univariate(ch_.sample_ef(), n));
      ch_.observe_ef(P.initial_ood_answers[i]);
      eq_evals.push_back(P.initial_ood_answers[i]);
    }
    c0.gamma = ch_.sample_ef();
    combine_claims(&claimed_, c0.gamma, eq_evals, {});
    constraints.push_back(std::move(c0));
    std::vector<ef> last = sumcheck(P.initial_sumcheck, cfg_.ff(0), cfg_.starting_folding_pow_bits,
                                    "the initial sumcheck");
    challenges.insert(challenges.end(), last.begin(), last.end());

    Digest prev_root = P.initial_commitment;
    for (int r = 0; r < cfg_.n_rounds(); r++) {
      const RoundConfig& rp = cfg_.round_parameters[r];
      const WhirRoundProof& RP = P.rounds[r];
      const std::string where = "round " + std::to_string(r);
      observe_root(RP.commitment);
      RoundConstraint cr;
      cr.nv = rp.num_variables;
      for (int i = 0; i < rp.ood_samples; i++) {
        cr.eq_points.push_back(expand_from_univariate(ch_.sample_ef(), rp.num_variables));
        ch_.observe_ef(RP.ood_answers[i]);
      }
      pow(rp.pow_bits, RP.pow_witness, where.c_str());
      ch_.sample();   // the transcript checkpoint after the round's proof of work
      std::vector<ef> folds;
      stir(rp.domain_size, cfg_.ff(r), rp.num_queries, rp.folded_domain_gen, prev_root, RP.queries, last,
           &cr.sel_vars, &folds, where.c_str());
      cr.gamma = ch_.sample_ef();
      combine_claims(&claimed_, cr.gamma, RP.ood_answers, folds);
      constraints.push_back(std::move(cr));
      last = sumcheck(RP.sumcheck, cfg_.ff(r + 1), rp.folding_pow_bits, where.c_str());
      challenges.insert(challenges.end(), last.begin(), last.end());
      prev_root = RP.commitment;
    }

    // the final round: the polynomial in the clear, its STIR queries checked on it
    for (const ef& e : P.final_poly) ch_.observe_ef(e);
    const RoundConfig fc = cfg_.final_round_config();
    pow(cfg_.final_pow_bits, P.final_pow_witness, "the final round");
    std::vector<u32> vars;
    std::vector<ef> folds;
    stir(fc.domain_size, cfg_.ff(cfg_.n_rounds()), cfg_.final_queries, fc.folded_domain_gen, prev_root,
         P.final_queries, last, &vars, &folds, "the final round");
    for (size_t j = 0; j < vars.size(); j++)
      if (!kb::ef_eq(horner(P.final_poly, vars[j]), folds[j]))
        fail(VerifyFailure::StirOnFinal, "a STIR answer disagrees with the final polynomial");
    last.clear();
    if (cfg_.final_sumcheck_rounds > 0)
      last = sumcheck(P.final_sumcheck, cfg_.final_sumcheck_rounds, cfg_.final_folding_pow_bits,
                      "the final sumcheck");
    challenges.insert(challenges.end(), last.begin(), last.end());
    if ((int)challenges.size() != n) fail(VerifyFailure::Shape, "sumcheck challenges do not cover the variables");

    ef weights = kb::ef_zero();
    for (const auto& c : constraints) weights = kb::ef_add(weights, constraint_weight(c, challenges));
    const ef final_value = mle_host(P.final_poly, last);
    if (!kb::ef_eq(claimed_, kb::ef_mul(weights, final_value)))
      fail(VerifyFailure::FinalSumcheck, "the last sumcheck claim is not weights(r) * final(r)");
    return VerifyFailure::None;
  } catch (const Failed& e) {
    return e.f;
  }
}

}  // namespace

bool deserialize(const uint8_t* bytes, size_t len, WhirProof* out, std::string* err, bool digests_are_field) {
  Cursor c(bytes, len, digests_are_field);
  WhirProof p;
  p.has_initial_commitment = c.some("initial commitment");
  if (p.has_initial_commitment) p.initial_commitment = c.digest("initial commitment");
  read_exts(c, &p.initial_ood_answers, "initial OOD answers");
  read_sumcheck(c, &p.initial_sumcheck);
  // smallest encoding of a round: tag, digest, three lengths, witness, two lengths
  p.rounds.resize(c.length(1 + 32 + 8 + 4 + 8 + 16, "rounds"));
  for (WhirRoundProof& r : p.rounds) {
    if (!c.some("round commitment")) c.fail("a round without its commitment");
    r.commitment = c.digest("round commitment");
    read_exts(c, &r.ood_answers, "round OOD answers");
    r.pow_witness = c.field("round witness");
    read_openings(c, &r.queries);
    read_sumcheck(c, &r.sumcheck);
    if (c.failed()) break;
  }
  p.has_final_poly = c.some("final polynomial");
  if (p.has_final_poly) read_exts(c, &p.final_poly, "final polynomial");
  p.final_pow_witness = c.field("final witness");
  read_openings(c, &p.final_queries);
  p.has_final_sumcheck = c.some("final sumcheck");
  if (p.has_final_sumcheck) read_sumcheck(c, &p.final_sumcheck);
  if (!c.failed() && !c.at_end()) c.fail("bytes after the end of the proof");
  if (c.failed()) {
    *err = "malformed proof: " + c.why();
    return false;
  }
  *out = std::move(p);
  return true;
}

Digest leaf_digest(const VerifierKey& key, const u32* row, size_t len) {
  if (key.keccak) return keccak_words(row, len);
  return key.leaf_width == 24 ? sponge<24>(row, len, *key.c24) : sponge<16>(row, len, *key.c16);
}

bool path_reaches_root(const VerifierKey& key, Digest node, u64 index, u64 height, const std::vector<Digest>& siblings,
                       const Digest& root) {
  if (index >= height) return false;
  size_t depth = 0;
  while ((1ull << depth) < height) depth++;
  if (siblings.size() != depth) return false;
  for (const Digest& s : siblings) {
    const Digest& l = (index & 1) ? s : node;
    const Digest& r = (index & 1) ? node : s;
    if (key.keccak) {
      u32 both[16];
      std::copy(l.begin(), l.end(), both);
      std::copy(r.begin(), r.end(), both + 8);
      node = keccak_words(both, 16);
    } else {
      node = compress(l, r, *key.c16);
    }
    index >>= 1;
  }
  return node == root;
}

VerifyFailure verify(const VerifierKey& key, const WhirProof& proof, Challenger& ch,
                     const std::vector<std::vector<ef>>& points, const std::vector<ef>& evals, std::string* why) {
  return Check(key, ch, why).run(proof, points, evals);
}

}  // namespace whir
