// Host-side protocol pieces (see host.h).
#include "host.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace whir {

using namespace kb;

// ---------------------------------------------------------------- parameters
static const int MAX_NUM_VARIABLES_TO_SEND_COEFFS = 6;   // src/parameters/mod.rs:11
static const double LOG2_10 = 3.321928094887362347870319429489390175864831393;
static const int EF_BITS = 124;   // bit length of |EF| = p^4
static const int F_TWO_ADICITY = 24;

bool FoldingFactor::check_validity(int n, std::string* err) const {
  if (first > n || rest > n) { *err = "Folding factor is greater than the number of variables"; return false; }
  if (first == 0 || rest == 0) { *err = "Folding factor shouldn't be zero."; return false; }
  return true;
}

void FoldingFactor::compute_number_of_rounds(int n, int* rounds, int* final_rounds) const {
  if (constant()) {
    if (n <= MAX_NUM_VARIABLES_TO_SEND_COEFFS) { *rounds = 0; *final_rounds = n - rest; return; }
    int num = (n - MAX_NUM_VARIABLES_TO_SEND_COEFFS + rest - 1) / rest;
    *final_rounds = n - num * rest;
    *rounds = num - 1;
    return;
  }
  int nv = n - first;
  if (nv < MAX_NUM_VARIABLES_TO_SEND_COEFFS) { *rounds = 0; *final_rounds = nv; return; }
  int num = (nv - MAX_NUM_VARIABLES_TO_SEND_COEFFS + rest - 1) / rest;
  *final_rounds = nv - num * rest;
  *rounds = num;
}

// SecurityAssumption (src/parameters/errors.rs:25-187)
namespace sec {
double log_eta(int s, int lir) {
  switch (s) {
    case UniqueDecoding: return 0.;
    case JohnsonBound: return -(0.5 * lir + LOG2_10 + 1.);
    default: return -(lir + LOG2_10 + 1.);
  }
}
double list_size_bits(int s, int log_degree, int lir) {
  double le = log_eta(s, lir);
  switch (s) {
    case UniqueDecoding: return 0.;
    case JohnsonBound: return lir / 2. - (1. + le);
    default: return (double)(log_degree + lir) - le;
  }
}
double prox_gaps_error(int s, int log_degree, int lir, int field_bits, int num_functions) {
  double le = log_eta(s, lir);
  double error;
  switch (s) {
    case UniqueDecoding: error = (double)(log_degree + lir); break;
    case JohnsonBound: {
      double sqrt_rho_20 = 1. + LOG2_10 + 0.5 * lir;
      error = (double)(2 * log_degree) + 7. * (std::min(sqrt_rho_20, -le) - 1.);
      break;
    }
    default: error = (double)(log_degree + 2 * lir) - le;
  }
  return (double)field_bits - (error + std::log2((double)num_functions - 1.));
}
double log_1_delta(int s, int lir) {
  double eta = std::pow(2., log_eta(s, lir));
  double rate = 1. / (double)(1ull << lir);
  double delta;
  switch (s) {
    case UniqueDecoding: delta = 0.5 * (1. - rate); break;
    case JohnsonBound: delta = 1. - std::sqrt(rate) - eta; break;
    default: delta = 1. - rate - eta;
  }
  return std::log2(1. - delta);
}
int queries(int s, int level, int lir) { return (int)std::ceil(-(double)level / log_1_delta(s, lir)); }
double queries_error(int s, int lir, int nq) { return -(double)nq * log_1_delta(s, lir); }
double ood_error(int s, int log_degree, int lir, int field_bits, int samples) {
  if (s == UniqueDecoding) return 0.;
  double error = 2. * list_size_bits(s, log_degree, lir) + (double)(log_degree * samples);
  return (double)(samples * field_bits) + 1. - error;
}
int determine_ood_samples(int s, int level, int log_degree, int lir, int field_bits) {
  if (s == UniqueDecoding) return 0;
  for (int k = 1; k < 64; k++)
    if (ood_error(s, log_degree, lir, field_bits, k) >= (double)level) return k;
  return -1;
}
}  // namespace sec

static double folding_pow_bits(int level, int s, int field_bits, int nv, int lir) {
  double prox = sec::prox_gaps_error(s, nv, lir, field_bits, 2);
  double sumcheck = (double)field_bits - (sec::list_size_bits(s, nv, lir) + 1.);
  return std::max(0., (double)level - std::min(prox, sumcheck));
}

static int ilog2(u64 x) {
  int k = 0;
  while ((x >> (k + 1)) != 0) k++;
  return k;
}

bool WhirConfig::create(int n, const ProtocolParameters& p, WhirConfig* out, std::string* err) {
  WhirConfig c;
  c.num_variables = n;
  c.params = p;
  if (!p.folding_factor.check_validity(n, err)) return false;
  if (p.rs_domain_initial_reduction_factor > p.folding_factor.at_round(0)) {
    *err = "Increasing the code rate is not a good idea";
    return false;
  }
  if (p.soundness_type < 0 || p.soundness_type > 2) { *err = "unknown soundness type"; return false; }
  const int s = p.soundness_type;
  const int protocol_level = std::max(0, p.security_level - p.pow_bits);
  int lir = p.starting_log_inv_rate;
  int nv = n;
  u64 domain = 1ull << (nv + lir);
  if (nv + lir - p.folding_factor.at_round(0) > F_TWO_ADICITY) { *err = "Increase folding_factor_0"; return false; }
  int nr, fsr;
  p.folding_factor.compute_number_of_rounds(n, &nr, &fsr);
  if (fsr < 0 || nr < 0) { *err = "folding schedule does not fit the number of variables"; return false; }
  c.commitment_ood_samples = sec::determine_ood_samples(s, p.security_level, nv, lir, EF_BITS);
  c.starting_folding_pow_bits = (int)folding_pow_bits(p.security_level, s, EF_BITS, nv, lir);
  nv -= p.folding_factor.at_round(0);
  for (int r = 0; r < nr; r++) {
    int rs_red = r == 0 ? p.rs_domain_initial_reduction_factor : 1;
    int next_rate = lir + (p.folding_factor.at_round(r) - rs_red);
    int nq = sec::queries(s, protocol_level, lir);
    int ood = sec::determine_ood_samples(s, p.security_level, nv, next_rate, EF_BITS);
    double qerr = sec::queries_error(s, lir, nq);
    double list = sec::list_size_bits(s, nv, next_rate);
    double comb = (double)EF_BITS - (std::log2((double)(ood + nq)) + list + 1.);
    double pow_bits = std::max(0., (double)p.security_level - std::min(qerr, comb));
    RoundConfig rc;
    rc.pow_bits = (int)pow_bits;
    rc.folding_pow_bits = (int)folding_pow_bits(p.security_level, s, EF_BITS, nv, next_rate);
    rc.num_queries = nq;
    rc.ood_samples = ood;
    rc.num_variables = nv;
    rc.folding_factor = p.folding_factor.at_round(r);
    rc.domain_size = domain;
    rc.folded_domain_gen = two_adic_generator(ilog2(domain) - rc.folding_factor);
    c.round_parameters.push_back(rc);
    nv -= p.folding_factor.at_round(r + 1);
    lir = next_rate;
    domain >>= rs_red;
  }
  c.final_queries = sec::queries(s, protocol_level, lir);
  c.final_pow_bits = (int)std::max(0., (double)p.security_level - sec::queries_error(s, lir, c.final_queries));
  c.final_sumcheck_rounds = fsr;
  c.final_folding_pow_bits = (int)std::max(0., (double)p.security_level - (double)(EF_BITS - 1));
  if (c.commitment_ood_samples < 0) { *err = "Could not find an appropriate number of OOD samples"; return false; }
  for (auto& rc : c.round_parameters)
    if (rc.ood_samples < 0) { *err = "Could not find an appropriate number of OOD samples"; return false; }
  if (n != p.folding_factor.total_number(c.n_rounds()) + c.final_sumcheck_rounds) {
    *err = "Invalid prover parameters";
    return false;
  }
  *out = c;
  return true;
}

u64 WhirConfig::inv_rate(int r) const {
  u64 new_domain = round_parameters[r].domain_size >> rs_reduction_factor(r);
  u64 num_evals = 1ull << (num_variables - params.folding_factor.total_number(r));
  return new_domain / num_evals;
}

RoundConfig WhirConfig::final_round_config() const {
  RoundConfig rc;
  rc.num_queries = final_queries;
  rc.pow_bits = final_pow_bits;
  rc.folding_pow_bits = final_folding_pow_bits;
  if (round_parameters.empty()) {
    u64 dom = starting_domain_size();
    rc.num_variables = num_variables - ff(0);
    rc.folding_factor = ff(0);
    rc.domain_size = dom;
    rc.folded_domain_gen = two_adic_generator(ilog2(dom) - ff(0));
    rc.ood_samples = 0;
  } else {
    const RoundConfig& last = round_parameters.back();
    int f = ff(n_rounds());
    u64 dom = last.domain_size >> rs_reduction_factor(n_rounds() - 1);
    rc.num_variables = last.num_variables - f;
    rc.folding_factor = f;
    rc.domain_size = dom;
    rc.folded_domain_gen = two_adic_generator(ilog2(dom) - f);
    rc.ood_samples = last.ood_samples;
  }
  return rc;
}

// ---------------------------------------------------------------- domain separator
// Pattern element = label + count + kind, kind: Sample 0, Observe 1, Hint 2
// (src/fiat_shamir/pattern.rs).  Labels are the enum positions there.
namespace {
struct Pattern {
  std::vector<u32> v;
  void push(u64 x) { v.push_back(to_monty((u32)(x % P))); }
  void observe(u64 count, int label) { push((u64)label + count + 1); }
  void sample(u64 count, int label) { push((u64)label + count); }
  void hint(int label) { push((u64)label + 2); }
  void pow(int bits) {
    if (bits > 0) {
      sample(32, 6);   // Sample::PowQueries
      observe(8, 6);   // Observe::PowNonce
    }
  }
  void ood(int k) {
    if (k > 0) {
      sample(k, 7);    // Sample::OodQuery
      observe(k, 1);   // Observe::OodAnswers
    }
  }
  void sumcheck(int rounds, int pow_bits) {
    for (int i = 0; i < rounds; i++) {
      observe(3, 2);   // Observe::SumcheckPoly
      sample(1, 2);    // Sample::FoldingRandomness
      pow(pow_bits);
    }
  }
};
u64 bytes_for(u64 folded) { return ((u64)ilog2(folded * 2 - 1) + 7) / 8; }
}  // namespace

std::vector<u32> domain_separator_pattern(const WhirConfig& c, int digest_elems) {
  Pattern d;
  d.observe(digest_elems, 0);            // Observe::MerkleDigest
  d.ood(c.commitment_ood_samples);
  d.sample(1, 0);                        // Sample::InitialCombinationRandomness
  d.sumcheck(c.ff(0), c.starting_folding_pow_bits);
  u64 domain = c.starting_domain_size();
  for (int r = 0; r < c.n_rounds(); r++) {
    const RoundConfig& rc = c.round_parameters[r];
    u64 nbytes = bytes_for(domain >> c.ff(r));
    d.observe(digest_elems, 0);
    d.ood(rc.ood_samples);
    d.sample((u64)rc.num_queries * nbytes, 4);   // Sample::StirQueries
    d.hint(0);                                   // Hint::StirQueries
    d.hint(2);                                   // Hint::MerkleProof
    d.pow(rc.pow_bits);
    d.sample(1, 3);                              // Sample::CombinationRandomness
    d.sumcheck(c.ff(r + 1), rc.folding_pow_bits);
    domain >>= c.rs_reduction_factor(r);
  }
  u64 nbytes = bytes_for(domain >> c.ff(c.n_rounds()));
  d.observe(1ull << c.final_sumcheck_rounds, 5);   // Observe::FinalCoeffs
  d.sample(nbytes * c.final_queries, 5);           // Sample::FinalQueries
  d.hint(1);                                       // Hint::StirAnswers
  d.hint(2);                                       // Hint::MerkleProof
  d.pow(c.final_pow_bits);
  d.sumcheck(c.final_sumcheck_rounds, c.final_folding_pow_bits);
  d.hint(3);                                       // Hint::DeferredWeightEvaluations
  return d.v;
}

// ---------------------------------------------------------------- challenger
void DuplexChallenger::duplexing() {
  for (int i = 0; i < in_len; i++) sponge[i] = in[i];
  in_len = 0;
  permute<16>(sponge, *perm);
  for (int i = 0; i < 8; i++) out[i] = sponge[i];
  out_len = 8;
}

void DuplexChallenger::observe(u32 v) {
  out_len = 0;
  in[in_len++] = v;
  if (in_len == 8) duplexing();
}

u32 DuplexChallenger::sample() {
  if (in_len > 0 || out_len == 0) duplexing();
  return out[--out_len];
}

void KeccakChallenger::observe_byte(uint8_t b) {
  out_len = 0;
  in.push_back(b);
}

uint8_t KeccakChallenger::sample_byte() {
  if (out_len == 0) {
    keccak256(in.data(), in.size(), out);
    in.assign(out, out + 32);   // chaining: the output becomes the next input
    out_len = 32;
  }
  return out[--out_len];
}

void KeccakChallenger::observe_field(u32 monty) {
  const u32 v = from_monty(monty);
  for (int b = 0; b < 4; b++) observe_byte((uint8_t)(v >> (8 * b)));
}

u32 KeccakChallenger::sample_u32() {
  u32 v = 0;
  for (int b = 0; b < 4; b++) v |= (u32)sample_byte() << (8 * b);
  return v;
}

u32 KeccakChallenger::sample_field() {
  for (;;) {
    const u32 v = sample_u32() & 0x7fffffffu;   // 2^ceil(log2 p) - 1
    if (v < P) return to_monty(v);
  }
}

void Challenger::observe_digest(const std::array<u32, 8>& d) {
  if (!keccak) {
    dx.observe_slice(d.data(), d.size());
    return;
  }
  for (u32 w : d)
    for (int b = 0; b < 4; b++) kc.observe_byte((uint8_t)(w >> (8 * b)));
}

std::vector<u64> get_challenge_stir_queries(u64 domain_size, int folding_factor, int num_queries,
                                            Challenger& ch) {
  const u64 folded = domain_size >> folding_factor;
  const int bits = ilog2(folded);
  const int max_bits = 20;   // min(F::bits() - 1, 20)
  const u64 mask = (1ull << bits) - 1;
  std::vector<u64> q;
  q.reserve(num_queries);
  const u64 total = (u64)num_queries * bits;
  if (total <= (u64)max_bits) {
    u64 all = ch.sample_bits((int)total);
    for (int i = 0; i < num_queries; i++, all >>= bits) q.push_back((all & mask) % folded);
  } else {
    int per_call = bits > 0 ? max_bits / bits : 0;
    if (per_call >= 2) {
      for (int left = num_queries; left > 0;) {
        int batch = std::min(left, per_call);
        u64 all = ch.sample_bits(batch * bits);
        for (int i = 0; i < batch; i++, all >>= bits) q.push_back((all & mask) % folded);
        left -= batch;
      }
    } else {
      for (int i = 0; i < num_queries; i++) q.push_back(ch.sample_bits(bits));
    }
  }
  std::sort(q.begin(), q.end());
  q.erase(std::unique(q.begin(), q.end()), q.end());
  return q;
}

// ---------------------------------------------------------------- RNG
SmallRng::SmallRng(u64 seed) {
  for (int i = 0; i < 4; i++) {
    seed += 0x9e3779b97f4a7c15ull;
    u64 z = seed;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    s[i] = z ^ (z >> 31);
  }
}

u64 SmallRng::next_u64() {
  auto rotl = [](u64 x, int k) { return (x << k) | (x >> (64 - k)); };
  const u64 result = rotl(s[0] + s[3], 23) + s[0];
  const u64 t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}

u32 SmallRng::field_monty() {
  for (;;) {
    u32 x = next_u32() >> 1;
    if (x < P) return x;
  }
}

void poseidon2_constants_from_seed(u64 seed, P2C16* c16, P2C24* c24) {
  SmallRng rng(seed);
  for (auto& v : c16->ext_init) v = rng.field_monty();
  for (auto& v : c16->ext_term) v = rng.field_monty();
  for (auto& v : c16->internal) v = rng.field_monty();
  for (auto& v : c24->ext_init) v = rng.field_monty();
  for (auto& v : c24->ext_term) v = rng.field_monty();
  for (auto& v : c24->internal) v = rng.field_monty();
}

// ---------------------------------------------------------------- proof bytes
namespace {
static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "proof bytes are written as host words");
// Writes the proof through sink.put(bytes, n): a counting pass sizes the
// output, a second pass writes it where the caller wants it.
template <class Sink>
struct Out {
  Sink& k;
  void u8(uint8_t x) { k.put(&x, 1); }
  void le(u64 x, int n) { k.put(&x, (size_t)n); }   // little-endian host: the low n bytes
  void words(const u32* w, size_t n) { k.put(w, 4 * n); }
  void f(u32 x) { le(x, 4); }
  void len(size_t n) { le(n, 8); }
  void e(const ef& x) { words(x.c, 4); }
  void digest(const Digest& d) { words(d.data(), d.size()); }
  void text(const char* s) { size_t n = strlen(s); len(n); k.put(s, n); }
  void sumcheck(const SumcheckData& s) {
    len(s.polynomial_evaluations.size());
    for (auto& pe : s.polynomial_evaluations) { e(pe[0]); e(pe[1]); }
    len(s.pow_witnesses.size());
    for (u32 w : s.pow_witnesses) f(w);
  }
  void queries(const std::vector<QueryOpening>& qs) {
    len(qs.size());
    for (auto& q : qs) {
      text(q.extension ? "extension" : "base");
      len(q.extension ? q.values.size() / 4 : q.values.size());
      words(q.values.data(), q.values.size());
      len(q.proof.size());
      for (auto& d : q.proof) digest(d);
    }
  }
  void proof(const WhirProof& p) {
    u8(p.has_initial_commitment);
    if (p.has_initial_commitment) digest(p.initial_commitment);
    len(p.initial_ood_answers.size());
    for (auto& x : p.initial_ood_answers) e(x);
    sumcheck(p.initial_sumcheck);
    len(p.rounds.size());
    for (auto& r : p.rounds) {
      u8(1);
      digest(r.commitment);
      len(r.ood_answers.size());
      for (auto& x : r.ood_answers) e(x);
      f(r.pow_witness);
      queries(r.queries);
      sumcheck(r.sumcheck);
    }
    u8(p.has_final_poly);
    if (p.has_final_poly) {
      len(p.final_poly.size());
      for (auto& x : p.final_poly) e(x);
    }
    f(p.final_pow_witness);
    queries(p.final_queries);
    u8(p.has_final_sumcheck);
    if (p.has_final_sumcheck) sumcheck(p.final_sumcheck);
  }
};
struct CountSink {
  size_t n = 0;
  void put(const void*, size_t k) { n += k; }
};
struct BufSink {
  uint8_t* p;
  void put(const void* src, size_t k) {
    memcpy(p, src, k);
    p += k;
  }
};
}  // namespace

size_t serialized_size(const WhirProof& p) {
  CountSink k;
  Out<CountSink>{k}.proof(p);
  return k.n;
}

void serialize_into(const WhirProof& p, uint8_t* out) {
  BufSink k{out};
  Out<BufSink>{k}.proof(p);
}

std::vector<uint8_t> serialize(const WhirProof& p) {
  std::vector<uint8_t> b(serialized_size(p));
  serialize_into(p, b.data());
  return b;
}

// ---------------------------------------------------------------- EF helpers
std::vector<ef> eq_table_host(const std::vector<ef>& z, const ef& scale) {
  std::vector<ef> t{scale};
  for (const ef& zj : z) {
    std::vector<ef> nt(t.size() * 2);
    for (size_t i = 0; i < t.size(); i++) {
      ef hi = ef_mul(t[i], zj);
      nt[2 * i] = ef_sub(t[i], hi);
      nt[2 * i + 1] = hi;
    }
    t.swap(nt);
  }
  return t;
}

ef mle_host(const std::vector<ef>& f, const std::vector<ef>& point) {
  // fold variable 0 (the MSB) first
  std::vector<ef> a = f;
  for (const ef& r : point) {
    size_t h = a.size() / 2;
    for (size_t i = 0; i < h; i++) a[i] = ef_add(a[i], ef_mul(r, ef_sub(a[i + h], a[i])));
    a.resize(h);
  }
  return a[0];
}

std::vector<ef> expand_from_univariate(const ef& y, int n) {
  std::vector<ef> r(n);
  ef cur = y;
  for (int i = n - 1; i >= 0; i--) {
    r[i] = cur;
    cur = ef_mul(cur, cur);
  }
  return r;
}

ef extrapolate_012(const ef& e0, const ef& e1, const ef& e2, const ef& r) {
  const u32 inv2 = inv(to_monty(2));
  const ef one = ef_one(), two = ef_base(to_monty(2));
  ef l0 = ef_mul_base(ef_mul(ef_sub(r, one), ef_sub(r, two)), inv2);
  ef l1 = ef_mul(r, ef_sub(two, r));
  ef l2 = ef_mul_base(ef_mul(r, ef_sub(r, one)), inv2);
  return ef_add(ef_add(ef_mul(e0, l0), ef_mul(e1, l1)), ef_mul(e2, l2));
}

std::vector<ef> powers(const ef& g, size_t n) {
  std::vector<ef> v(n);
  ef x = ef_one();
  for (size_t i = 0; i < n; i++) {
    v[i] = x;
    x = ef_mul(x, g);
  }
  return v;
}

ef eq_at(const std::vector<ef>& z, const std::vector<ef>& u) {
  ef acc = ef_one();
  for (size_t i = 0; i < z.size(); i++) {
    // z u + (1 - z)(1 - u)
    ef t = ef_add(ef_mul(z[i], u[i]), ef_mul(ef_sub(ef_one(), z[i]), ef_sub(ef_one(), u[i])));
    acc = ef_mul(acc, t);
  }
  return acc;
}

}  // namespace whir
