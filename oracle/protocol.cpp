// Oracle: multilinear helpers, parameter derivation, domain separator, STIR
// query sampling, proof encoding, and the WHIR prover and verifier.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include <algorithm>
#include <cassert>
#include <chrono>
#include <cmath>
#include <omp.h>

#include "oracle.hpp"

namespace orc {

// ------------------------------------------------------------------ polynomials
// eq(z, x) over {0,1}^n, z_0 on the MSB (src/poly/evals.rs:312-320 and the
// convention pinned at evals.rs:1355-1392).
std::vector<EF> eq_table(const std::vector<EF>& z, const EF& scale) {
  std::vector<EF> t(1, scale);
  for (const EF& zj : z) {
    std::vector<EF> nt(t.size() * 2);
    for (size_t i = 0; i < t.size(); i++) {
      EF hi = ef_mul(t[i], zj);
      nt[2 * i] = ef_sub(t[i], hi);
      nt[2 * i + 1] = hi;
    }
    t.swap(nt);
  }
  return t;
}

EF mle_eval_base(const std::vector<u32>& f, const std::vector<EF>& point) {
  assert(f.size() == (1ull << point.size()));
  auto eq = eq_table(point, ef_one());
  int nt = omp_get_max_threads();
  std::vector<EF> part(nt, ef_zero());
#pragma omp parallel
  {
    EF acc = ef_zero();
    int tid = omp_get_thread_num();
#pragma omp for schedule(static)
    for (long long i = 0; i < (long long)f.size(); i++) acc = ef_add(acc, ef_scale(eq[i], f[i]));
    part[tid] = acc;
  }
  EF s = ef_zero();
  for (auto& p : part) s = ef_add(s, p);
  return s;
}

EF mle_eval_ext(const std::vector<EF>& f, const std::vector<EF>& point) {
  assert(f.size() == (1ull << point.size()));
  auto eq = eq_table(point, ef_one());
  EF s = ef_zero();
  for (size_t i = 0; i < f.size(); i++) s = ef_add(s, ef_mul(eq[i], f[i]));
  return s;
}

// [y^(2^(n-1)), ..., y^2, y]  (src/poly/multilinear.rs:73-85)
std::vector<EF> expand_from_univariate(const EF& y, int n) {
  std::vector<EF> r(n);
  EF cur = y;
  for (int i = n - 1; i >= 0; i--) {
    r[i] = cur;
    cur = ef_mul(cur, cur);
  }
  return r;
}

// (h(0), h(2)) of h(X) = sum_x e(X, x) w(X, x)  (src/poly/evals.rs:493-526)
void sumcheck_coeffs(const std::vector<EF>& e, const std::vector<EF>& w, EF* c0, EF* c2) {
  size_t h = e.size() / 2;
  EF a0 = ef_zero(), a2 = ef_zero();
  for (size_t i = 0; i < h; i++) {
    a0 = ef_add(a0, ef_mul(w[i], e[i]));
    EF dw = ef_sub(ef_double(w[i + h]), w[i]);
    EF de = ef_sub(ef_double(e[i + h]), e[i]);
    a2 = ef_add(a2, ef_mul(dw, de));
  }
  *c0 = a0;
  *c2 = a2;
}

// a0 + r (a1 - a0) on the first variable (src/poly/evals.rs:127-148)
void fold_ext(std::vector<EF>& a, const EF& r) {
  size_t h = a.size() / 2;
  for (size_t i = 0; i < h; i++) a[i] = ef_add(a[i], ef_mul(ef_sub(a[i + h], a[i]), r));
  a.resize(h);
}

std::vector<EF> fold_base(const std::vector<u32>& a, const EF& r) {
  size_t h = a.size() / 2;
  std::vector<EF> out(h);
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)h; i++) out[i] = ef_add(ef_scale(r, fsub(a[i + h], a[i])), ef_base(a[i]));
  return out;
}

// h(r) from h(0), h(1), h(2)  (src/sumcheck/lagrange.rs:49-66, 141-147)
static void lagrange_012(const EF& r, EF* l) {
  u32 inv2 = finv(2);
  EF one = ef_one(), two = ef_base(2);
  l[0] = ef_scale(ef_mul(ef_sub(r, one), ef_sub(r, two)), inv2);
  l[1] = ef_mul(r, ef_sub(two, r));
  l[2] = ef_scale(ef_mul(r, ef_sub(r, one)), inv2);
}

EF extrapolate_012(const EF& e0, const EF& e1, const EF& e2, const EF& r) {
  EF l[3];
  lagrange_012(r, l);
  return ef_add(ef_add(ef_mul(e0, l[0]), ef_mul(e1, l[1])), ef_mul(e2, l[2]));
}

// Tensor Lagrange weights on {0,1,2}^k; the digit of r_t has weight 3^t
// (src/sumcheck/lagrange.rs:98-117: new[j * 3^t + i] = old[i] * L_j(r_t)).
std::vector<EF> lagrange_weights_012_multi(const std::vector<EF>& rs) {
  std::vector<EF> w(1, ef_one());
  for (const EF& r : rs) {
    EF l[3];
    lagrange_012(r, l);
    std::vector<EF> nw;
    nw.reserve(w.size() * 3);
    for (int j = 0; j < 3; j++)
      for (const EF& x : w) nw.push_back(ef_mul(x, l[j]));
    w.swap(nw);
  }
  return w;
}

// ------------------------------------------------------------------ config
static const double LOG2_10 = 3.321928094887362347870319429489390175864831393;
static const int MAX_VARS_SEND_COEFFS = 6;   // src/parameters/mod.rs:11

static double log_eta(int s, int lir) {
  if (s == UniqueDecoding) return 0.;
  if (s == JohnsonBound) return -(0.5 * lir + LOG2_10 + 1.);
  return -(lir + LOG2_10 + 1.);
}
static double list_size_bits(int s, int ld, int lir) {
  double le = log_eta(s, lir);
  if (s == UniqueDecoding) return 0.;
  if (s == JohnsonBound) return (double)lir / 2. - (1. + le);
  return (double)(ld + lir) - le;
}
static double prox_gaps_error(int s, int ld, int lir, int fsb, int nf) {
  double le = log_eta(s, lir), err;
  if (s == UniqueDecoding) err = (double)(ld + lir);
  else if (s == JohnsonBound) {
    double num = (double)(2 * ld);
    double sq = 1. + LOG2_10 + 0.5 * lir;
    err = num + 7. * (std::min(sq, -le) - 1.);
  } else err = (double)(ld + 2 * lir) - le;
  return (double)fsb - (err + std::log2((double)nf - 1.));
}
static double log_1_delta(int s, int lir) {
  double le = log_eta(s, lir);
  double eta = std::pow(2., le);
  double rate = 1. / (double)(1ull << lir);
  double delta;
  if (s == UniqueDecoding) delta = 0.5 * (1. - rate);
  else if (s == JohnsonBound) delta = 1. - std::sqrt(rate) - eta;
  else delta = 1. - rate - eta;
  return std::log2(1. - delta);
}
static int queries(int s, int lambda, int lir) { return (int)std::ceil(-(double)lambda / log_1_delta(s, lir)); }
static double queries_error(int s, int lir, int nq) { return -(double)nq * log_1_delta(s, lir); }
static double ood_error(int s, int ld, int lir, int fsb, int k) {
  if (s == UniqueDecoding) return 0.;
  double lsb = list_size_bits(s, ld, lir);
  double err = 2. * lsb + (double)(ld * k);
  return (double)(k * fsb) + 1. - err;
}
static int determine_ood_samples(int s, int lambda, int ld, int lir, int fsb) {
  if (s == UniqueDecoding) return 0;
  for (int k = 1; k < 64; k++)
    if (ood_error(s, ld, lir, fsb, k) >= (double)lambda) return k;
  return -1;
}
static double rbr_fold_sumcheck(int s, int fsb, int nv, int lir) {
  return (double)fsb - (list_size_bits(s, nv, lir) + 1.);
}
static double folding_pow_bits(int lambda, int s, int fsb, int nv, int lir) {
  double pg = prox_gaps_error(s, nv, lir, fsb, 2);
  double sc = rbr_fold_sumcheck(s, fsb, nv, lir);
  return std::max(0., (double)lambda - std::min(pg, sc));
}
static double rbr_queries_combination(int s, int fsb, int nv, int lir, int ood, int nq) {
  double ls = list_size_bits(s, nv, lir);
  return (double)fsb - (std::log2((double)(ood + nq)) + ls + 1.);
}

// FoldingFactor::compute_number_of_rounds (src/parameters/mod.rs:91-125);
// ff0 == ff is treated as FoldingFactor::Constant.
static void number_of_rounds(const Params& p, int* nr, int* fsr) {
  int n = p.num_variables;
  if (p.ff0 == p.ff) {
    if (n <= MAX_VARS_SEND_COEFFS) { *nr = 0; *fsr = n - p.ff; return; }
    int num = (n - MAX_VARS_SEND_COEFFS + p.ff - 1) / p.ff;
    *fsr = n - num * p.ff;
    *nr = num - 1;
  } else {
    int nv = n - p.ff0;
    if (nv < MAX_VARS_SEND_COEFFS) { *nr = 0; *fsr = nv; return; }
    int num = (nv - MAX_VARS_SEND_COEFFS + p.ff - 1) / p.ff;
    *fsr = nv - num * p.ff;
    *nr = num;
  }
}

// WhirConfig::new (src/whir/parameters.rs:81-234)
bool make_config(const Params& p, Config* out, std::string* err) {
  Config c;
  c.p = p;
  int n = p.num_variables;
  if (p.ff0 > n || p.ff > n || p.ff0 == 0 || p.ff == 0) { *err = "invalid folding factor"; return false; }
  if (p.rs_reduction > p.ff0) { *err = "Increasing the code rate is not a good idea"; return false; }
  const int fsb = 124;   // EF::bits() for the degree-4 extension of KoalaBear
  const int s = p.soundness;
  int protocol_level = std::max(0, p.security_level - p.pow_bits);
  int lir = p.log_inv_rate;
  int nv = n;
  int log_domain = nv + lir;
  u64 domain = 1ull << log_domain;
  if (log_domain - p.ff0 > 24) { *err = "Increase folding_factor_0"; return false; }
  int nr, fsr;
  number_of_rounds(p, &nr, &fsr);
  c.commitment_ood_samples = determine_ood_samples(s, p.security_level, nv, lir, fsb);
  c.starting_folding_pow_bits = (int)folding_pow_bits(p.security_level, s, fsb, nv, lir);
  nv -= c.ff_at(0);
  for (int r = 0; r < nr; r++) {
    int rsr = r == 0 ? p.rs_reduction : 1;
    int next_rate = lir + (c.ff_at(r) - rsr);
    int nq = queries(s, protocol_level, lir);
    int ood = determine_ood_samples(s, p.security_level, nv, next_rate, fsb);
    double qe = queries_error(s, lir, nq);
    double ce = rbr_queries_combination(s, fsb, nv, next_rate, ood, nq);
    double pow = std::max(0., (double)p.security_level - std::min(qe, ce));
    double fpow = folding_pow_bits(p.security_level, s, fsb, nv, next_rate);
    int ld = 0;
    while ((1ull << (ld + 1)) <= domain) ld++;
    RoundCfg rc;
    rc.pow_bits = (int)pow;
    rc.folding_pow_bits = (int)fpow;
    rc.num_queries = nq;
    rc.ood_samples = ood;
    rc.num_variables = nv;
    rc.folding_factor = c.ff_at(r);
    rc.domain_size = domain;
    rc.folded_domain_gen = two_adic_generator(ld - c.ff_at(r));
    c.rounds.push_back(rc);
    nv -= c.ff_at(r + 1);
    lir = next_rate;
    domain >>= rsr;
  }
  c.final_queries = queries(s, protocol_level, lir);
  c.final_pow_bits = (int)std::max(0., (double)p.security_level - queries_error(s, lir, c.final_queries));
  c.final_sumcheck_rounds = fsr;
  c.final_folding_pow_bits = (int)std::max(0., (double)p.security_level - (double)(fsb - 1));
  if (c.commitment_ood_samples < 0) { *err = "no OOD sample count"; return false; }
  for (auto& rc : c.rounds)
    if (rc.ood_samples < 0) { *err = "no OOD sample count"; return false; }
  if (n != c.total_number(c.n_rounds()) + c.final_sumcheck_rounds) { *err = "Invalid prover parameters"; return false; }
  *out = c;
  return true;
}

u64 Config::inv_rate(int round) const {
  u64 nd = rounds[round].domain_size >> rs_reduction_factor(round);
  u64 ne = 1ull << (p.num_variables - total_number(round));
  return nd / ne;
}

static int ilog2u(u64 x) {
  int k = 0;
  while ((2ull << k) <= x) k++;
  return k;
}

RoundCfg Config::final_round_config() const {
  RoundCfg rc;
  if (rounds.empty()) {
    u64 dom = 1ull << (p.num_variables + p.log_inv_rate);
    rc.num_variables = p.num_variables - ff_at(0);
    rc.folding_factor = ff_at(0);
    rc.num_queries = final_queries;
    rc.pow_bits = final_pow_bits;
    rc.domain_size = dom;
    rc.folded_domain_gen = two_adic_generator(ilog2u(dom) - ff_at(0));
    rc.ood_samples = 0;
    rc.folding_pow_bits = final_folding_pow_bits;
  } else {
    const RoundCfg& last = rounds.back();
    int ffn = ff_at(n_rounds());
    u64 dom = last.domain_size >> rs_reduction_factor(n_rounds() - 1);
    rc.num_variables = last.num_variables - ffn;
    rc.folding_factor = ffn;
    rc.num_queries = final_queries;
    rc.pow_bits = final_pow_bits;
    rc.domain_size = dom;
    rc.folded_domain_gen = two_adic_generator(ilog2u(dom) - ffn);
    rc.ood_samples = last.ood_samples;
    rc.folding_pow_bits = final_folding_pow_bits;
  }
  return rc;
}

// ------------------------------------------------------------------ domain separator
// src/fiat_shamir/domain_separator.rs:57-213 with the label enums of pattern.rs.
namespace {
enum { PAT_SAMPLE = 0, PAT_OBSERVE = 1, PAT_HINT = 2 };
enum { S_INITIAL_COMB = 0, S_FOLDING = 2, S_COMB = 3, S_STIR = 4, S_FINAL_Q = 5, S_POW = 6, S_OOD = 7 };
enum { O_MERKLE = 0, O_OOD = 1, O_SUMCHECK = 2, O_FINAL_COEFFS = 5, O_POW_NONCE = 6 };
enum { H_STIR_Q = 0, H_STIR_A = 1, H_MERKLE = 2, H_DEFERRED = 3 };
struct DS {
  std::vector<u32> pat;
  void observe(u64 count, u32 lab) { pat.push_back((u32)((lab + count + PAT_OBSERVE) % P)); }
  void sample(u64 count, u32 lab) { pat.push_back((u32)((lab + count + PAT_SAMPLE) % P)); }
  void hint(u32 lab) { pat.push_back(lab + PAT_HINT); }
  void pow(int bits) {
    if (bits > 0) { sample(32, S_POW); observe(8, O_POW_NONCE); }
  }
  void ood(int k) {
    if (k > 0) { sample(k, S_OOD); observe(k, O_OOD); }
  }
  void sumcheck(int rounds, int pow_bits) {
    for (int i = 0; i < rounds; i++) { observe(3, O_SUMCHECK); sample(1, S_FOLDING); pow(pow_bits); }
  }
};
}  // namespace

std::vector<u32> domain_separator(const Config& c) {
  DS d;
  const int D = c.p.digest_elems_pattern;
  d.observe(D, O_MERKLE);
  if (c.commitment_ood_samples > 0) d.ood(c.commitment_ood_samples);
  d.sample(1, S_INITIAL_COMB);
  d.sumcheck(c.ff_at(0), c.starting_folding_pow_bits);
  u64 domain = 1ull << (c.p.num_variables + c.p.log_inv_rate);
  for (int r = 0; r < c.n_rounds(); r++) {
    const RoundCfg& rc = c.rounds[r];
    u64 folded = domain >> c.ff_at(r);
    u64 bytes = (ilog2u(folded * 2 - 1) + 7) / 8;
    d.observe(D, O_MERKLE);
    d.ood(rc.ood_samples);
    d.sample(rc.num_queries * bytes, S_STIR);
    d.hint(H_STIR_Q);
    d.hint(H_MERKLE);
    d.pow(rc.pow_bits);
    d.sample(1, S_COMB);
    d.sumcheck(c.ff_at(r + 1), rc.folding_pow_bits);
    domain >>= c.rs_reduction_factor(r);
  }
  u64 folded = domain >> c.ff_at(c.n_rounds());
  u64 bytes = (ilog2u(folded * 2 - 1) + 7) / 8;
  d.observe(1ull << c.final_sumcheck_rounds, O_FINAL_COEFFS);
  d.sample(bytes * c.final_queries, S_FINAL_Q);
  d.hint(H_STIR_A);
  d.hint(H_MERKLE);
  d.pow(c.final_pow_bits);
  d.sumcheck(c.final_sumcheck_rounds, c.final_folding_pow_bits);
  d.hint(H_DEFERRED);
  return d.pat;
}

// ------------------------------------------------------------------ STIR queries
// src/whir/utils.rs:24-145
std::vector<u64> stir_queries(Challenger& ch, u64 domain_size, int folding_factor, int num_queries) {
  u64 folded = domain_size >> folding_factor;
  int bits = ilog2u(folded);
  const int max_bits = 20;   // min(F::bits() - 1, 20)
  u64 total = (u64)num_queries * bits;
  std::vector<u64> q;
  if (total <= (u64)max_bits) {
    u64 all = ch.sample_bits((int)total);
    u64 mask = (1ull << bits) - 1;
    for (int i = 0; i < num_queries; i++) {
      q.push_back((all & mask) % folded);
      all >>= bits;
    }
  } else {
    int per = bits ? max_bits / bits : 0;
    if (per >= 2) {
      int remaining = num_queries;
      u64 mask = (1ull << bits) - 1;
      while (remaining > 0) {
        int bs = std::min(remaining, per);
        u64 all = ch.sample_bits(bs * bits);
        for (int i = 0; i < bs; i++) {
          q.push_back((all & mask) % folded);
          all >>= bits;
        }
        remaining -= bs;
      }
    } else {
      for (int i = 0; i < num_queries; i++) q.push_back(ch.sample_bits(bits));
    }
  }
  std::sort(q.begin(), q.end());
  q.erase(std::unique(q.begin(), q.end()), q.end());
  return q;
}

// ------------------------------------------------------------------ proof encoding
// Field elements are written as their Montgomery u32 (little endian), lengths
// as u64, Option as a one-byte tag, QueryOpening as {type: str, values, proof}.
namespace {
struct W {
  std::vector<uint8_t> b;
  bool raw_digests = false;
  void u8(uint8_t v) { b.push_back(v); }
  void u32v(u32 v) { for (int i = 0; i < 4; i++) b.push_back((v >> (8 * i)) & 0xff); }
  void u64v(u64 v) { for (int i = 0; i < 8; i++) b.push_back((v >> (8 * i)) & 0xff); }
  void f(u32 canon) { u32v(to_monty(canon)); }
  void ef(const EF& e) { for (int k = 0; k < 4; k++) f(e.c[k]); }
  void dig(const Digest& d) {
    for (int k = 0; k < 8; k++) {
      if (raw_digests) u32v(d[k]);
      else f(d[k]);
    }
  }
  void str(const char* s) { u64 n = strlen(s); u64v(n); b.insert(b.end(), s, s + n); }
  void sc(const SumcheckData& s) {
    u64v(s.evals.size());
    for (auto& e : s.evals) { ef(e[0]); ef(e[1]); }
    u64v(s.pow.size());
    for (u32 w : s.pow) f(w);
  }
  void qs(const std::vector<QueryOpening>& v) {
    u64v(v.size());
    for (auto& q : v) {
      if (q.ext) {
        str("extension");
        u64v(q.ext_values.size());
        for (auto& e : q.ext_values) ef(e);
      } else {
        str("base");
        u64v(q.base_values.size());
        for (u32 x : q.base_values) f(x);
      }
      u64v(q.path.size());
      for (auto& d : q.path) dig(d);
    }
  }
};
struct R {
  const uint8_t* p;
  size_t n, pos = 0;
  bool ok = true;
  bool raw_digests = false;
  bool need(size_t k) { if (pos + k > n) ok = false; return ok; }
  uint8_t u8() { if (!need(1)) return 0; return p[pos++]; }
  u32 u32v() { if (!need(4)) return 0; u32 v = 0; for (int i = 0; i < 4; i++) v |= (u32)p[pos + i] << (8 * i); pos += 4; return v; }
  u64 u64v() { if (!need(8)) return 0; u64 v = 0; for (int i = 0; i < 8; i++) v |= (u64)p[pos + i] << (8 * i); pos += 8; return v; }
  u32 f() { u32 m = u32v(); if (m >= P) ok = false; return from_monty(m); }
  EF ef() { EF e; for (int k = 0; k < 4; k++) e.c[k] = f(); return e; }
  Digest dig() { Digest d; for (int k = 0; k < 8; k++) d[k] = raw_digests ? u32v() : f(); return d; }
  u64 len(size_t unit) { u64 l = u64v(); if (l > (n - pos) / (unit ? unit : 1) + 1) { ok = false; return 0; } return l; }
  std::string str() { u64 l = len(1); if (!need(l)) return ""; std::string s((const char*)p + pos, l); pos += l; return s; }
  SumcheckData sc() {
    SumcheckData s;
    u64 l = len(32);
    for (u64 i = 0; i < l && ok; i++) { EF a = ef(); EF b = ef(); s.evals.push_back({a, b}); }
    l = len(4);
    for (u64 i = 0; i < l && ok; i++) s.pow.push_back(f());
    return s;
  }
  std::vector<QueryOpening> qs() {
    std::vector<QueryOpening> v;
    u64 l = len(16);
    for (u64 i = 0; i < l && ok; i++) {
      QueryOpening q;
      std::string t = str();
      if (t == "extension") {
        q.ext = true;
        u64 k = len(16);
        for (u64 j = 0; j < k && ok; j++) q.ext_values.push_back(ef());
      } else if (t == "base") {
        u64 k = len(4);
        for (u64 j = 0; j < k && ok; j++) q.base_values.push_back(f());
      } else {
        ok = false;
      }
      u64 d = len(32);
      for (u64 j = 0; j < d && ok; j++) q.path.push_back(dig());
      v.push_back(std::move(q));
    }
    return v;
  }
};
}  // namespace

std::vector<uint8_t> serialize_proof(const Proof& pr, bool raw_digests) {
  W w;
  w.raw_digests = raw_digests;
  w.u8(1);
  w.dig(pr.initial_commitment);
  w.u64v(pr.initial_ood.size());
  for (auto& e : pr.initial_ood) w.ef(e);
  w.sc(pr.initial_sumcheck);
  w.u64v(pr.rounds.size());
  for (auto& r : pr.rounds) {
    w.u8(1);
    w.dig(r.commitment);
    w.u64v(r.ood.size());
    for (auto& e : r.ood) w.ef(e);
    w.f(r.pow_witness);
    w.qs(r.queries);
    w.sc(r.sumcheck);
  }
  w.u8(1);
  w.u64v(pr.final_poly.size());
  for (auto& e : pr.final_poly) w.ef(e);
  w.f(pr.final_pow_witness);
  w.qs(pr.final_queries);
  w.u8(pr.has_final_sumcheck ? 1 : 0);
  if (pr.has_final_sumcheck) w.sc(pr.final_sumcheck);
  return w.b;
}

bool deserialize_proof(const uint8_t* buf, size_t len, Proof* out, bool raw_digests) {
  R r{buf, len};
  r.raw_digests = raw_digests;
  Proof pr;
  if (r.u8() != 1) return false;
  pr.initial_commitment = r.dig();
  u64 l = r.len(16);
  for (u64 i = 0; i < l && r.ok; i++) pr.initial_ood.push_back(r.ef());
  pr.initial_sumcheck = r.sc();
  l = r.len(8);
  for (u64 i = 0; i < l && r.ok; i++) {
    RoundProof rp;
    if (r.u8() != 1) return false;
    rp.commitment = r.dig();
    u64 k = r.len(16);
    for (u64 j = 0; j < k && r.ok; j++) rp.ood.push_back(r.ef());
    rp.pow_witness = r.f();
    rp.queries = r.qs();
    rp.sumcheck = r.sc();
    pr.rounds.push_back(std::move(rp));
  }
  if (r.u8() != 1) return false;
  l = r.len(16);
  for (u64 i = 0; i < l && r.ok; i++) pr.final_poly.push_back(r.ef());
  pr.final_pow_witness = r.f();
  pr.final_queries = r.qs();
  uint8_t tag = r.u8();
  if (tag == 1) {
    pr.has_final_sumcheck = true;
    pr.final_sumcheck = r.sc();
  } else if (tag != 0) {
    return false;
  }
  if (!r.ok || r.pos != len) return false;
  *out = std::move(pr);
  return true;
}

// ------------------------------------------------------------------ prover
namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// SumcheckData::observe_and_sample (src/whir/proof.rs:163-193)
EF observe_and_sample(SumcheckData& sd, Challenger& ch, const EF& c0, const EF& c2, int pow_bits) {
  sd.evals.push_back({c0, c2});
  ch.observe_ef(c0);
  ch.observe_ef(c2);
  if (pow_bits > 0) sd.pow.push_back(ch.grind(pow_bits));
  return ch.sample_ef();
}

struct SumcheckState {
  std::vector<EF> f, w;
  EF sum;
};

// ProductPolynomial::round (src/sumcheck/product_polynomial.rs:307-364)
EF sumcheck_round(SumcheckState& s, SumcheckData& sd, Challenger& ch, int pow_bits) {
  EF c0, c2;
  sumcheck_coeffs(s.f, s.w, &c0, &c2);
  EF r = observe_and_sample(sd, ch, c0, c2, pow_bits);
  fold_ext(s.f, r);
  fold_ext(s.w, r);
  s.sum = extrapolate_012(c0, ef_sub(s.sum, c0), c2, r);
  return r;
}

std::vector<EF> powers(const EF& g, size_t n) {
  std::vector<EF> v(n);
  EF x = ef_one();
  for (size_t i = 0; i < n; i++) { v[i] = x; x = ef_mul(x, g); }
  return v;
}

// Sumcheck::new_classic_* (src/sumcheck/sumcheck_prover.rs:65-161)
std::vector<EF> initial_classic(const Config& c, const std::vector<u32>& poly, const Statement& st, Challenger& ch,
                                SumcheckData& sd, SumcheckState& s) {
  int n = c.p.num_variables;
  EF alpha = ch.sample_ef();
  auto ap = powers(alpha, st.points.size());
  s.w.assign(1ull << n, ef_zero());
  s.sum = ef_zero();
  for (size_t i = 0; i < st.points.size(); i++) {
    auto eq = eq_table(st.points[i], ap[i]);
    for (size_t x = 0; x < eq.size(); x++) s.w[x] = ef_add(s.w[x], eq[x]);
    s.sum = ef_add(s.sum, ef_mul(ap[i], st.evals[i]));
  }
  std::vector<EF> fe(poly.size());
  for (size_t i = 0; i < poly.size(); i++) fe[i] = ef_base(poly[i]);
  EF c0, c2;
  sumcheck_coeffs(fe, s.w, &c0, &c2);
  fe.clear();
  EF r = observe_and_sample(sd, ch, c0, c2, c.starting_folding_pow_bits);
  fold_ext(s.w, r);
  s.f = fold_base(poly, r);
  s.sum = extrapolate_012(c0, ef_sub(s.sum, c0), c2, r);
  std::vector<EF> rs{r};
  for (int i = 1; i < c.ff_at(0); i++) rs.push_back(sumcheck_round(s, sd, ch, c.starting_folding_pow_bits));
  return rs;
}

// Sumcheck::new_svo with SplitEq (src/sumcheck/sumcheck_prover.rs:164-227,
// src/sumcheck/svo.rs:82-196, 315-559).
std::vector<EF> initial_svo(const Config& c, const std::vector<u32>& poly, const Statement& st, Challenger& ch,
                            SumcheckData& sd, SumcheckState& s) {
  int n = c.p.num_variables, l = c.ff_at(0);
  EF alpha = ch.sample_ef();
  auto ap = powers(alpha, st.points.size());
  size_t chunk = 1ull << (n - l);
  struct Split {
    std::vector<EF> z_svo, z_rest, partial;
    std::vector<std::array<std::vector<EF>, 2>> acc;
    EF eval;
  };
  std::vector<Split> sp(st.points.size());
  for (size_t j = 0; j < st.points.size(); j++) {
    Split& q = sp[j];
    q.z_svo.assign(st.points[j].begin(), st.points[j].begin() + l);
    q.z_rest.assign(st.points[j].begin() + l, st.points[j].end());
    auto eqr = eq_table(q.z_rest, ef_one());
    q.partial.assign(1ull << l, ef_zero());
    for (size_t sidx = 0; sidx < (1ull << l); sidx++) {
      EF acc = ef_zero();
      for (size_t x = 0; x < chunk; x++) acc = ef_add(acc, ef_scale(eqr[x], poly[sidx * chunk + x]));
      q.partial[sidx] = acc;
    }
    auto eqs = eq_table(q.z_svo, ef_one());
    q.eval = ef_zero();
    for (size_t sidx = 0; sidx < q.partial.size(); sidx++) q.eval = ef_add(q.eval, ef_mul(eqs[sidx], q.partial[sidx]));
    // accumulators for rounds 1..l (svo.rs:426-439, 149-196)
    for (int i = 1; i <= l; i++) {
      // grid {0,1,2}^(i-1) x {last}; coordinate t has ternary weight 3^t
      size_t cnt = 1;
      for (int t = 0; t < i - 1; t++) cnt *= 3;
      int offset = l - i;
      std::vector<EF> z0(q.z_svo.begin(), q.z_svo.begin() + i), z1(q.z_svo.begin() + i, q.z_svo.end());
      auto eq0 = eq_table(z0, ef_one());
      auto eq1 = eq_table(z1, ef_one());
      std::vector<EF> red(1ull << i, ef_zero());
      for (size_t x0 = 0; x0 < red.size(); x0++)
        for (size_t x1 = 0; x1 < eq1.size(); x1++)
          red[x0] = ef_add(red[x0], ef_mul(eq1[x1], q.partial[(x0 << offset) + x1]));
      std::array<std::vector<EF>, 2> acc;
      for (int which = 0; which < 2; which++) {
        u32 last = which == 0 ? 0 : 2;
        for (size_t g = 0; g < cnt; g++) {
          std::vector<EF> u(i);
          size_t gg = g;
          for (int t = 0; t < i - 1; t++) { u[t] = ef_base((u32)(gg % 3)); gg /= 3; }
          u[i - 1] = ef_base(last);
          auto co = eq_table(u, ef_one());
          EF a = ef_zero(), b = ef_zero();
          for (size_t x = 0; x < co.size(); x++) {
            a = ef_add(a, ef_mul(eq0[x], co[x]));
            b = ef_add(b, ef_mul(red[x], co[x]));
          }
          acc[which].push_back(ef_mul(a, b));
        }
      }
      q.acc.push_back(acc);
    }
  }
  s.sum = ef_zero();
  for (size_t j = 0; j < sp.size(); j++) s.sum = ef_add(s.sum, ef_mul(sp[j].eval, ap[j]));
  std::vector<EF> rs;
  for (int ri = 0; ri < l; ri++) {
    auto wts = lagrange_weights_012_multi(rs);
    EF c0 = ef_zero(), c2 = ef_zero();
    for (size_t j = 0; j < sp.size(); j++) {
      EF d0 = ef_zero(), d2 = ef_zero();
      for (size_t g = 0; g < wts.size(); g++) {
        d0 = ef_add(d0, ef_mul(sp[j].acc[ri][0][g], wts[g]));
        d2 = ef_add(d2, ef_mul(sp[j].acc[ri][1][g], wts[g]));
      }
      c0 = ef_add(c0, ef_mul(ap[j], d0));
      c2 = ef_add(c2, ef_mul(ap[j], d2));
    }
    EF r = observe_and_sample(sd, ch, c0, c2, c.starting_folding_pow_bits);
    s.sum = extrapolate_012(c0, ef_sub(s.sum, c0), c2, r);
    rs.push_back(r);
  }
  // compress_multi_into_packed (evals.rs:422-447)
  auto eqrs = eq_table(rs, ef_one());
  s.f.assign(chunk, ef_zero());
  for (size_t sidx = 0; sidx < eqrs.size(); sidx++)
    for (size_t x = 0; x < chunk; x++) s.f[x] = ef_add(s.f[x], ef_scale(eqrs[sidx], poly[sidx * chunk + x]));
  // combine_into_packed (svo.rs:504-559)
  s.w.assign(chunk, ef_zero());
  for (size_t j = 0; j < sp.size(); j++) {
    EF scale = mle_eval_ext(eq_table(sp[j].z_svo, ap[j]), rs);
    auto eqr = eq_table(sp[j].z_rest, scale);
    for (size_t x = 0; x < chunk; x++) s.w[x] = ef_add(s.w[x], eqr[x]);
  }
  return rs;
}

// Transpose-and-pad then column DFT (src/whir/committer/writer.rs:70-86,
// src/whir/prover/mod.rs:166-180) for base and extension evaluations.
std::vector<u32> lde_base(const std::vector<u32>& f, int m, int k, u64 height) {
  u64 L = 1ull << (m - k), W = 1ull << k;
  std::vector<u32> mat(height * W, 0);
  for (u64 a = 0; a < W; a++)
    for (u64 b = 0; b < L; b++) mat[b * W + a] = f[a * L + b];
  dft_batch(mat, height, W);
  return mat;
}

// Rows are flattened EF coefficient-major: [e0.c0..c3, e1.c0..c3, ...].
std::vector<u32> lde_ext(const std::vector<EF>& f, int m, int k, u64 height) {
  u64 L = 1ull << (m - k), W = 1ull << k;
  std::vector<u32> mat(height * W * 4, 0);
  for (u64 a = 0; a < W; a++)
    for (u64 b = 0; b < L; b++)
      for (int t = 0; t < 4; t++) mat[(b * W + a) * 4 + t] = f[a * L + b].c[t];
  dft_batch(mat, height, W * 4);
  return mat;
}

void observe_digest(Challenger& ch, const Digest& d) { ch.observe_digest(d); }

std::vector<EF> row_as_ef(const std::vector<u32>& row, bool ext) {
  std::vector<EF> v;
  if (ext) {
    for (size_t i = 0; i < row.size(); i += 4) v.push_back(EF{{row[i], row[i + 1], row[i + 2], row[i + 3]}});
  } else {
    for (u32 x : row) v.push_back(ef_base(x));
  }
  return v;
}

QueryOpening open_query(const MerkleTree& t, bool ext, u64 idx) {
  QueryOpening q;
  std::vector<u32> row;
  merkle_open(t, idx, &row, &q.path);
  q.ext = ext;
  if (ext) q.ext_values = row_as_ef(row, true);
  else q.base_values = row;
  return q;
}

}  // namespace

Proof whir_prove(const Config& cfg, const Poseidon2& p16, const Poseidon2& leaf, const std::vector<u32>& poly,
                 const std::vector<std::vector<EF>>& user_points, Challenger& ch, Statement* stmt_out, Timing* tm) {
  const int n = cfg.p.num_variables;
  const MerkleHashers hs{&leaf, &p16, cfg.p.mmcs == 1};
  Proof pr;
  Statement st;
  for (auto& z : user_points) {
    st.points.push_back(z);
    st.evals.push_back(mle_eval_base(poly, z));
  }
  if (stmt_out) *stmt_out = st;

  // ---- commit (src/whir/committer/writer.rs:56-109)
  double t0 = now_ms();
  int k0 = cfg.ff_at(0);
  u64 H0 = 1ull << (n + cfg.p.log_inv_rate - k0);
  std::vector<u32> codeword = lde_base(poly, n, k0, H0);
  double td = now_ms();
  MerkleTree tree0 = merkle_commit(hs, std::move(codeword), H0, 1ull << k0);
  double tm_ = now_ms();
  pr.initial_commitment = tree0.root();
  observe_digest(ch, pr.initial_commitment);
  for (int i = 0; i < cfg.commitment_ood_samples; i++) {
    auto pt = expand_from_univariate(ch.sample_ef(), n);
    EF ev = mle_eval_base(poly, pt);
    pr.initial_ood.push_back(ev);
    ch.observe_ef(ev);
    st.points.push_back(pt);
    st.evals.push_back(ev);
  }
  double t1 = now_ms();

  // ---- prove (src/whir/prover/mod.rs:97-417)
  SumcheckState s;
  std::vector<EF> fold_rand;
  if (cfg.p.svo && n > k0) fold_rand = initial_svo(cfg, poly, st, ch, pr.initial_sumcheck, s);
  else fold_rand = initial_classic(cfg, poly, st, ch, pr.initial_sumcheck, s);
  double ts = now_ms(), tr = ts;

  const MerkleTree* prev = &tree0;
  bool prev_ext = false;
  std::vector<MerkleTree> trees;
  trees.reserve(cfg.n_rounds());
  for (int round = 0; round <= cfg.n_rounds(); round++) {
    int m = n - cfg.total_number(round);
    assert(s.f.size() == (1ull << m));
    if (round == cfg.n_rounds()) {
      tr = now_ms();
      for (auto& e : s.f) ch.observe_ef(e);
      pr.final_poly = s.f;
      if (cfg.final_pow_bits > 0) pr.final_pow_witness = ch.grind(cfg.final_pow_bits);
      RoundCfg fc = cfg.final_round_config();
      auto idx = stir_queries(ch, fc.domain_size, cfg.ff_at(round), cfg.final_queries);
      for (u64 i : idx) pr.final_queries.push_back(open_query(*prev, prev_ext, i));
      if (cfg.final_sumcheck_rounds > 0) {
        pr.has_final_sumcheck = true;
        for (int i = 0; i < cfg.final_sumcheck_rounds; i++)
          sumcheck_round(s, pr.final_sumcheck, ch, cfg.final_folding_pow_bits);
      }
      break;
    }
    const RoundCfg& rc = cfg.rounds[round];
    RoundProof rp;
    int kn = cfg.ff_at(round + 1);
    u64 H = cfg.inv_rate(round) * (1ull << (m - kn));
    trees.push_back(merkle_commit(hs, lde_ext(s.f, m, kn, H), H, 4ull << kn));
    const MerkleTree& tree = trees.back();
    rp.commitment = tree.root();
    observe_digest(ch, rp.commitment);
    Statement ood;
    for (int i = 0; i < rc.ood_samples; i++) {
      auto pt = expand_from_univariate(ch.sample_ef(), m);
      EF ev = mle_eval_ext(s.f, pt);
      ch.observe_ef(ev);
      rp.ood.push_back(ev);
      ood.points.push_back(pt);
      ood.evals.push_back(ev);
    }
    if (rc.pow_bits > 0) rp.pow_witness = ch.grind(rc.pow_bits);
    ch.sample();
    auto idx = stir_queries(ch, rc.domain_size, cfg.ff_at(round), rc.num_queries);
    std::vector<u32> vars;
    std::vector<EF> sel_evals;
    for (u64 i : idx) {
      QueryOpening q = open_query(*prev, prev_ext, i);
      std::vector<EF> vals = q.ext ? q.ext_values : row_as_ef(q.base_values, false);
      sel_evals.push_back(mle_eval_ext(vals, fold_rand));
      vars.push_back(fpow(rc.folded_domain_gen, i));
      rp.queries.push_back(std::move(q));
    }
    EF gamma = ch.sample_ef();
    // Constraint::combine_packed (src/whir/constraints/mod.rs:234-248)
    auto gp = powers(gamma, ood.points.size() + vars.size());
    for (size_t i = 0; i < ood.points.size(); i++) {
      auto eq = eq_table(ood.points[i], gp[i]);
      for (size_t x = 0; x < eq.size(); x++) s.w[x] = ef_add(s.w[x], eq[x]);
      s.sum = ef_add(s.sum, ef_mul(gp[i], ood.evals[i]));
    }
    for (size_t j = 0; j < vars.size(); j++) {
      const EF& g = gp[ood.points.size() + j];
      u32 zb = 1;
      for (size_t x = 0; x < s.w.size(); x++) {
        s.w[x] = ef_add(s.w[x], ef_scale(g, zb));
        zb = fmul(zb, vars[j]);
      }
      s.sum = ef_add(s.sum, ef_mul(g, sel_evals[j]));
    }
    fold_rand.clear();
    for (int i = 0; i < kn; i++) fold_rand.push_back(sumcheck_round(s, rp.sumcheck, ch, rc.folding_pow_bits));
    pr.rounds.push_back(std::move(rp));
    prev = &tree;
    prev_ext = true;
  }
  double t2 = now_ms();
  if (tm) {
    tm->commit_ms = t1 - t0;
    tm->prove_ms = t2 - t1;
    tm->dft_ms = td - t0;
    tm->commit_matrix_ms = tm_ - td;
    tm->ood_ms = t1 - tm_;
    tm->initial_sumcheck_ms = ts - t1;
    tm->rounds_ms = tr - ts;
    tm->final_ms = t2 - tr;
  }
  return pr;
}

// ------------------------------------------------------------------ verifier
// src/whir/verifier/mod.rs:58-400, verifier/sumcheck.rs:31-101,
// committer/reader.rs:70-120, constraints/evaluator.rs:8-64.
namespace {
struct VConstraint {
  std::vector<std::vector<EF>> eq_points;
  std::vector<EF> eq_evals;
  std::vector<u32> sel_vars;
  std::vector<EF> sel_evals;
  EF gamma;
  int nv;
};

bool verify_sumcheck(const SumcheckData& sd, Challenger& ch, EF& claimed, int pow_bits, std::vector<EF>* rs,
                     std::string* err) {
  for (size_t i = 0; i < sd.evals.size(); i++) {
    const EF& c0 = sd.evals[i][0];
    const EF& c2 = sd.evals[i][1];
    ch.observe_ef(c0);
    ch.observe_ef(c2);
    if (pow_bits > 0) {
      if (i >= sd.pow.size() || !ch.check_witness(pow_bits, sd.pow[i])) { *err = "invalid sumcheck PoW witness"; return false; }
    }
    EF r = ch.sample_ef();
    claimed = extrapolate_012(c0, ef_sub(claimed, c0), c2, r);
    rs->push_back(r);
  }
  return true;
}

EF eval_eq_pts(const std::vector<EF>& p, const std::vector<EF>& q) {
  EF acc = ef_one();
  for (size_t i = 0; i < p.size(); i++) {
    // 1 + 2 l r - l - r
    EF t = ef_sub(ef_sub(ef_add(ef_one(), ef_double(ef_mul(p[i], q[i]))), p[i]), q[i]);
    acc = ef_mul(acc, t);
  }
  return acc;
}

EF eval_constraint(const VConstraint& c, const std::vector<EF>& point_rev) {
  // subpoint over 0..nv of the reversed point, reversed again
  std::vector<EF> pt(point_rev.begin(), point_rev.begin() + c.nv);
  std::reverse(pt.begin(), pt.end());
  EF acc = ef_zero();
  EF g = ef_one();
  for (auto& z : c.eq_points) {
    acc = ef_add(acc, ef_mul(eval_eq_pts(z, pt), g));
    g = ef_mul(g, c.gamma);
  }
  for (u32 v : c.sel_vars) {
    auto ex = expand_from_univariate(ef_base(v), c.nv);
    EF sp = ef_one();
    for (int i = 0; i < c.nv; i++) sp = ef_mul(sp, ef_add(ef_mul(pt[i], ef_sub(ex[i], ef_one())), ef_one()));
    acc = ef_add(acc, ef_mul(g, sp));
    g = ef_mul(g, c.gamma);
  }
  return acc;
}

EF combine_evals(const VConstraint& c) {
  EF acc = ef_zero(), g = ef_one();
  for (auto& e : c.eq_evals) { acc = ef_add(acc, ef_mul(g, e)); g = ef_mul(g, c.gamma); }
  for (auto& e : c.sel_evals) { acc = ef_add(acc, ef_mul(g, e)); g = ef_mul(g, c.gamma); }
  return acc;
}
}  // namespace

bool whir_verify(const Config& cfg, const Poseidon2& p16, const Poseidon2& leaf, const Proof& pr,
                 const Statement& stmt, Challenger& ch, std::string* err) {
  const int n = cfg.p.num_variables;
  const MerkleHashers hs{&leaf, &p16, cfg.p.mmcs == 1};
  auto parse = [&](const Digest& root, const std::vector<EF>& answers, int nv, int ood, VConstraint* c) -> bool {
    observe_digest(ch, root);
    if ((int)answers.size() < ood) { *err = "missing OOD answers"; return false; }
    for (int i = 0; i < ood; i++) {
      auto pt = expand_from_univariate(ch.sample_ef(), nv);
      ch.observe_ef(answers[i]);
      c->eq_points.push_back(pt);
      c->eq_evals.push_back(answers[i]);
    }
    return true;
  };
  auto stir = [&](const RoundCfg& rc, const Digest& root, bool base, const std::vector<EF>& frand, int round,
                  VConstraint* c) -> bool {
    u32 w = round < cfg.n_rounds() ? pr.rounds[round].pow_witness : pr.final_pow_witness;
    if (rc.pow_bits > 0 && !ch.check_witness(rc.pow_bits, w)) { *err = "invalid PoW witness"; return false; }
    if (round < cfg.n_rounds()) ch.sample();
    auto idx = stir_queries(ch, rc.domain_size, rc.folding_factor, rc.num_queries);
    const auto& qs = round == cfg.n_rounds() ? pr.final_queries : pr.rounds[round].queries;
    if (qs.size() != idx.size()) { *err = "wrong number of query openings"; return false; }
    u64 height = rc.domain_size >> rc.folding_factor;
    u64 width = 1ull << rc.folding_factor;
    for (size_t q = 0; q < idx.size(); q++) {
      const auto& o = qs[q];
      std::vector<u32> row;
      std::vector<EF> vals;
      if (o.ext) {
        if (base || o.ext_values.size() != width) { *err = "bad opening shape"; return false; }
        for (auto& e : o.ext_values) for (int t = 0; t < 4; t++) row.push_back(e.c[t]);
        vals = o.ext_values;
      } else {
        if (!base || o.base_values.size() != width) { *err = "bad opening shape"; return false; }
        row = o.base_values;
        vals = row_as_ef(row, false);
      }
      if (!merkle_verify(hs, root, height, idx[q], row, o.path)) { *err = "Merkle proof rejected"; return false; }
      c->sel_vars.push_back(fpow(rc.folded_domain_gen, idx[q]));
      c->sel_evals.push_back(mle_eval_ext(vals, frand));
    }
    c->nv = rc.num_variables;
    return true;
  };

  std::vector<VConstraint> cons;
  std::vector<std::vector<EF>> rand_rounds;
  EF claimed = ef_zero();
  VConstraint c0;
  c0.nv = n;
  c0.eq_points = stmt.points;
  c0.eq_evals = stmt.evals;
  {
    VConstraint ood;
    if (!parse(pr.initial_commitment, pr.initial_ood, n, cfg.commitment_ood_samples, &ood)) return false;
    for (size_t i = 0; i < ood.eq_points.size(); i++) {
      c0.eq_points.push_back(ood.eq_points[i]);
      c0.eq_evals.push_back(ood.eq_evals[i]);
    }
  }
  c0.gamma = ch.sample_ef();
  claimed = ef_add(claimed, combine_evals(c0));
  cons.push_back(c0);
  std::vector<EF> rs;
  if (!verify_sumcheck(pr.initial_sumcheck, ch, claimed, cfg.starting_folding_pow_bits, &rs, err)) return false;
  if ((int)rs.size() != cfg.ff_at(0)) { *err = "wrong initial sumcheck length"; return false; }
  rand_rounds.push_back(rs);
  Digest prev_root = pr.initial_commitment;
  bool prev_base = true;
  if ((int)pr.rounds.size() != cfg.n_rounds()) { *err = "wrong number of rounds"; return false; }
  for (int r = 0; r < cfg.n_rounds(); r++) {
    const RoundCfg& rc = cfg.rounds[r];
    VConstraint c;
    c.nv = rc.num_variables;
    if (!parse(pr.rounds[r].commitment, pr.rounds[r].ood, rc.num_variables, rc.ood_samples, &c)) return false;
    VConstraint sel;
    if (!stir(rc, prev_root, prev_base, rand_rounds.back(), r, &sel)) return false;
    c.sel_vars = sel.sel_vars;
    c.sel_evals = sel.sel_evals;
    c.gamma = ch.sample_ef();
    claimed = ef_add(claimed, combine_evals(c));
    cons.push_back(c);
    std::vector<EF> rr;
    if (!verify_sumcheck(pr.rounds[r].sumcheck, ch, claimed, rc.folding_pow_bits, &rr, err)) return false;
    if ((int)rr.size() != cfg.ff_at(r + 1)) { *err = "wrong round sumcheck length"; return false; }
    rand_rounds.push_back(rr);
    prev_root = pr.rounds[r].commitment;
    prev_base = false;
  }
  if (pr.final_poly.size() != (1ull << cfg.final_sumcheck_rounds)) { *err = "wrong final polynomial size"; return false; }
  for (auto& e : pr.final_poly) ch.observe_ef(e);
  VConstraint fsel;
  if (!stir(cfg.final_round_config(), prev_root, prev_base, rand_rounds.back(), cfg.n_rounds(), &fsel)) return false;
  for (size_t j = 0; j < fsel.sel_vars.size(); j++) {
    EF acc = ef_zero();
    for (size_t i = pr.final_poly.size(); i-- > 0;) acc = ef_add(ef_mul(acc, ef_base(fsel.sel_vars[j])), pr.final_poly[i]);
    if (!ef_eq(acc, fsel.sel_evals[j])) { *err = "STIR constraint verification failed on final polynomial"; return false; }
  }
  std::vector<EF> frs;
  if (cfg.final_sumcheck_rounds > 0) {
    if (!pr.has_final_sumcheck || (int)pr.final_sumcheck.evals.size() != cfg.final_sumcheck_rounds) {
      *err = "missing final sumcheck";
      return false;
    }
    if (!verify_sumcheck(pr.final_sumcheck, ch, claimed, cfg.final_folding_pow_bits, &frs, err)) return false;
  }
  rand_rounds.push_back(frs);
  std::vector<EF> all;
  for (auto& v : rand_rounds) all.insert(all.end(), v.begin(), v.end());
  std::vector<EF> rev(all.rbegin(), all.rend());
  EF wsum = ef_zero();
  for (auto& c : cons) wsum = ef_add(wsum, eval_constraint(c, rev));
  EF fv = mle_eval_ext(pr.final_poly, frs);
  if (!ef_eq(claimed, ef_mul(wsum, fv))) { *err = "final sumcheck check failed"; return false; }
  return true;
}

}  // namespace orc
