// C API of the CPU oracle, loaded by the test suite through ctypes.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).  All field elements cross this
// API in Montgomery form; extension elements are 4 consecutive u32.
#include <cstdio>

#include "oracle.hpp"

using namespace orc;

namespace {
// params: [n, security_level, pow_bits, ff0, ff, log_inv_rate, rs_reduction,
//          soundness, svo, leaf_width, digest_elems_pattern, mmcs]
Params read_params(const int* a) {
  Params p;
  p.num_variables = a[0];
  p.security_level = a[1];
  p.pow_bits = a[2];
  p.ff0 = a[3];
  p.ff = a[4];
  p.log_inv_rate = a[5];
  p.rs_reduction = a[6];
  p.soundness = a[7];
  p.svo = a[8];
  p.leaf_width = a[9];
  p.digest_elems_pattern = a[10];
  p.mmcs = a[11];
  return p;
}
EF ef_in(const uint32_t* m) { return EF{{from_monty(m[0]), from_monty(m[1]), from_monty(m[2]), from_monty(m[3])}}; }
void ef_out(const EF& e, uint32_t* m) { for (int k = 0; k < 4; k++) m[k] = to_monty(e.c[k]); }
std::vector<EF> point_in(const uint32_t* m, int n) {
  std::vector<EF> v(n);
  for (int i = 0; i < n; i++) v[i] = ef_in(m + 4 * i);
  return v;
}
struct Perms {
  Poseidon2 p16, p24;
  explicit Perms(uint64_t seed) { poseidon2_pair_from_seed(seed, &p16, &p24); }
  const Poseidon2& leaf(int w) const { return w == 24 ? p24 : p16; }
};
}  // namespace

extern "C" {

int orc_config(const int* params, int* out, int cap) {
  Config c;
  std::string err;
  if (!make_config(read_params(params), &c, &err)) return -1;
  int need = 7 + 8 * c.n_rounds();
  if (cap < need) return -2;
  out[0] = c.commitment_ood_samples;
  out[1] = c.starting_folding_pow_bits;
  out[2] = c.n_rounds();
  out[3] = c.final_queries;
  out[4] = c.final_pow_bits;
  out[5] = c.final_sumcheck_rounds;
  out[6] = c.final_folding_pow_bits;
  for (int r = 0; r < c.n_rounds(); r++) {
    const RoundCfg& rc = c.rounds[r];
    int ld = 0;
    while ((1ull << (ld + 1)) <= rc.domain_size) ld++;
    int* o = out + 7 + 8 * r;
    o[0] = rc.pow_bits;
    o[1] = rc.folding_pow_bits;
    o[2] = rc.num_queries;
    o[3] = rc.ood_samples;
    o[4] = rc.num_variables;
    o[5] = rc.folding_factor;
    o[6] = ld;
    o[7] = (int)rc.folded_domain_gen;
  }
  return need;
}

int orc_domain_separator(const int* params, uint32_t* out, int cap) {
  Config c;
  std::string err;
  if (!make_config(read_params(params), &c, &err)) return -1;
  auto pat = domain_separator(c);
  if ((int)pat.size() > cap) return -2;
  for (size_t i = 0; i < pat.size(); i++) out[i] = to_monty(pat[i]);
  return (int)pat.size();
}

void orc_sample_field(uint64_t seed, uint32_t* out, size_t count) {
  SmallRng rng(seed);
  for (size_t i = 0; i < count; i++) out[i] = to_monty(rng.field());
}

// width 16: the instance drawn first from SmallRng(seed); width 24: the one drawn next.
void orc_poseidon2_constants(uint64_t seed, int width, uint32_t* ext_init, uint32_t* ext_term, uint32_t* internal) {
  Perms ps(seed);
  const Poseidon2& p = ps.leaf(width);
  for (size_t i = 0; i < p.ext_init.size(); i++) ext_init[i] = to_monty(p.ext_init[i]);
  for (size_t i = 0; i < p.ext_term.size(); i++) ext_term[i] = to_monty(p.ext_term[i]);
  for (size_t i = 0; i < p.internal.size(); i++) internal[i] = to_monty(p.internal[i]);
}

void orc_poseidon2_permute(uint64_t seed, int width, uint32_t* state) {
  Perms ps(seed);
  const Poseidon2& p = ps.leaf(width);
  u32 s[24];
  for (int i = 0; i < width; i++) s[i] = from_monty(state[i]);
  p.permute(s);
  for (int i = 0; i < width; i++) state[i] = to_monty(s[i]);
}

void orc_ef_mul(const uint32_t* a, const uint32_t* b, uint32_t* out) { ef_out(ef_mul(ef_in(a), ef_in(b)), out); }
void orc_ef_inv(const uint32_t* a, uint32_t* out) { ef_out(ef_inv(ef_in(a)), out); }
uint32_t orc_two_adic_generator(int bits) { return to_monty(two_adic_generator(bits)); }

void orc_dft(uint32_t* col, size_t n, int naive) {
  std::vector<u32> a(n);
  for (size_t i = 0; i < n; i++) a[i] = from_monty(col[i]);
  if (naive) dft_naive(a);
  else dft_radix2(a);
  for (size_t i = 0; i < n; i++) col[i] = to_monty(a[i]);
}

void orc_dft_batch(uint32_t* mat, size_t height, size_t width) {
  std::vector<u32> a(height * width);
  for (size_t i = 0; i < a.size(); i++) a[i] = from_monty(mat[i]);
  dft_batch(a, height, width);
  for (size_t i = 0; i < a.size(); i++) mat[i] = to_monty(a[i]);
}

void orc_sponge(uint64_t seed, int width, const uint32_t* vals, size_t len, uint32_t* out) {
  Perms ps(seed);
  std::vector<u32> v(len);
  for (size_t i = 0; i < len; i++) v[i] = from_monty(vals[i]);
  Digest d = sponge_hash(ps.leaf(width), v.data(), len);
  for (int k = 0; k < 8; k++) out[k] = to_monty(d[k]);
}

// Merkle tree over `height` rows of `width` base elements: root, and optionally
// the opening (row, path) of `index`.
int orc_merkle(uint64_t seed, int leaf_width, const uint32_t* rows, size_t height, size_t width, uint32_t* root,
               size_t index, uint32_t* row_out, uint32_t* path_out) {
  Perms ps(seed);
  std::vector<u32> v(height * width);
  for (size_t i = 0; i < v.size(); i++) v[i] = from_monty(rows[i]);
  MerkleTree t = merkle_commit(MerkleHashers{&ps.leaf(leaf_width), &ps.p16}, std::move(v), height, width);
  Digest r = t.root();
  for (int k = 0; k < 8; k++) root[k] = to_monty(r[k]);
  if (row_out && path_out) {
    std::vector<u32> row;
    std::vector<Digest> path;
    merkle_open(t, index, &row, &path);
    for (size_t i = 0; i < row.size(); i++) row_out[i] = to_monty(row[i]);
    for (size_t l = 0; l < path.size(); l++)
      for (int k = 0; k < 8; k++) path_out[8 * l + k] = to_monty(path[l][k]);
  }
  return 0;
}

// The Keccak MMCS's MerkleTreeMmcs<F, u64, SerializingHasher<..>, CompressionFunctionFromHasher<..>, 4>
// over Montgomery rows; root and path digests as 8 raw u32 words (4 u64 lanes, low word first).
int orc_merkle_keccak(const uint32_t* rows, size_t height, size_t width, uint32_t* root, size_t index,
                      uint32_t* row_out, uint32_t* path_out) {
  std::vector<u32> v(height * width);
  for (size_t i = 0; i < v.size(); i++) v[i] = from_monty(rows[i]);
  MerkleTree t = merkle_commit(MerkleHashers{nullptr, nullptr, true}, std::move(v), height, width);
  Digest r = t.root();
  for (int k = 0; k < 8; k++) root[k] = r[k];
  if (row_out && path_out) {
    std::vector<u32> row;
    std::vector<Digest> path;
    merkle_open(t, index, &row, &path);
    for (size_t i = 0; i < row.size(); i++) row_out[i] = to_monty(row[i]);
    for (size_t l = 0; l < path.size(); l++)
      for (int k = 0; k < 8; k++) path_out[8 * l + k] = path[l][k];
  }
  return 0;
}

void orc_eq_table(const uint32_t* point, int n, const uint32_t* scale, uint32_t* out) {
  auto t = eq_table(point_in(point, n), ef_in(scale));
  for (size_t i = 0; i < t.size(); i++) ef_out(t[i], out + 4 * i);
}

void orc_mle_base(const uint32_t* f, int n, const uint32_t* point, uint32_t* out) {
  std::vector<u32> v(1ull << n);
  for (size_t i = 0; i < v.size(); i++) v[i] = from_monty(f[i]);
  ef_out(mle_eval_base(v, point_in(point, n)), out);
}

void orc_mle_ext(const uint32_t* f, int n, const uint32_t* point, uint32_t* out) {
  std::vector<EF> v(1ull << n);
  for (size_t i = 0; i < v.size(); i++) v[i] = ef_in(f + 4 * i);
  ef_out(mle_eval_ext(v, point_in(point, n)), out);
}

void orc_sumcheck_coeffs(const uint32_t* e, const uint32_t* w, size_t len, uint32_t* c0c2) {
  std::vector<EF> a(len), b(len);
  for (size_t i = 0; i < len; i++) { a[i] = ef_in(e + 4 * i); b[i] = ef_in(w + 4 * i); }
  EF c0, c2;
  sumcheck_coeffs(a, b, &c0, &c2);
  ef_out(c0, c0c2);
  ef_out(c2, c0c2 + 4);
}

void orc_fold_ext(uint32_t* a, size_t len, const uint32_t* r) {
  std::vector<EF> v(len);
  for (size_t i = 0; i < len; i++) v[i] = ef_in(a + 4 * i);
  fold_ext(v, ef_in(r));
  for (size_t i = 0; i < v.size(); i++) ef_out(v[i], a + 4 * i);
}

void orc_extrapolate_012(const uint32_t* e0, const uint32_t* e1, const uint32_t* e2, const uint32_t* r, uint32_t* out) {
  ef_out(extrapolate_012(ef_in(e0), ef_in(e1), ef_in(e2), ef_in(r)), out);
}

// Challenger scripted from a list of ops: 0 observe(v), 1 sample, 2 sample_bits(v),
// 3 grind(v).  Each sample/sample_bits/grind writes one u32 (sample: Montgomery).
int orc_challenger_script(uint64_t seed, const uint32_t* ops, const uint32_t* args, int nops, uint32_t* out) {
  Perms ps(seed);
  Challenger ch;
  ch.perm = &ps.p16;
  int k = 0;
  for (int i = 0; i < nops; i++) {
    switch (ops[i]) {
      case 0: ch.observe(from_monty(args[i])); break;
      case 1: out[k++] = to_monty(ch.sample()); break;
      case 2: out[k++] = (uint32_t)ch.sample_bits((int)args[i]); break;
      case 3: out[k++] = ch.grind((int)args[i]); break;
      default: return -1;
    }
  }
  return k;
}

// Full commit + prove.  `points` holds npts points of n EF each; `evals_out`
// receives their evaluations (EF each).  Returns 0 on success.
int orc_prove(const int* params, uint64_t perm_seed, const uint32_t* poly, const uint32_t* points, int npts,
              uint8_t* out, size_t cap, size_t* len, uint32_t* evals_out, double* ms) {
  Params p = read_params(params);
  Config c;
  std::string err;
  if (!make_config(p, &c, &err)) return -1;
  Perms ps(perm_seed);
  int n = p.num_variables;
  std::vector<u32> f(1ull << n);
  for (size_t i = 0; i < f.size(); i++) f[i] = from_monty(poly[i]);
  std::vector<std::vector<EF>> pts;
  for (int j = 0; j < npts; j++) pts.push_back(point_in(points + (size_t)4 * n * j, n));
  Challenger ch;
  ch.perm = &ps.p16;
  ch.keccak = p.mmcs == 1;
  for (u32 x : domain_separator(c)) ch.observe(x);
  Statement st;
  Timing tm;
  Proof pr = whir_prove(c, ps.p16, ps.leaf(p.leaf_width), f, pts, ch, &st, &tm);
  auto bytes = serialize_proof(pr, p.mmcs == 1);
  *len = bytes.size();
  if (ms) {   // 8 values: commit, prove, then the stages of Timing in declaration order
    const double v[8] = {tm.commit_ms, tm.prove_ms, tm.dft_ms, tm.commit_matrix_ms, tm.ood_ms,
                         tm.initial_sumcheck_ms, tm.rounds_ms, tm.final_ms};
    memcpy(ms, v, sizeof v);
  }
  for (int j = 0; j < npts; j++) ef_out(st.evals[j], evals_out + 4 * j);
  if (bytes.size() > cap) return -2;
  memcpy(out, bytes.data(), bytes.size());
  return 0;
}

int orc_verify(const int* params, uint64_t perm_seed, const uint8_t* proof, size_t len, const uint32_t* points,
               const uint32_t* evals, int npts, char* err_out, int errcap) {
  Params p = read_params(params);
  Config c;
  std::string err;
  auto fail = [&](int code) {
    if (err_out && errcap > 0) snprintf(err_out, errcap, "%s", err.c_str());
    return code;
  };
  if (!make_config(p, &c, &err)) return fail(-1);
  Proof pr;
  if (!deserialize_proof(proof, len, &pr, p.mmcs == 1)) { err = "malformed proof bytes"; return fail(-3); }
  Perms ps(perm_seed);
  int n = p.num_variables;
  Statement st;
  for (int j = 0; j < npts; j++) {
    st.points.push_back(point_in(points + (size_t)4 * n * j, n));
    st.evals.push_back(ef_in(evals + 4 * j));
  }
  Challenger ch;
  ch.perm = &ps.p16;
  ch.keccak = p.mmcs == 1;
  for (u32 x : domain_separator(c)) ch.observe(x);
  if (!whir_verify(c, ps.p16, ps.leaf(p.leaf_width), pr, st, ch, &err)) return fail(1);
  return 0;
}

}  // extern "C"

extern "C" {
// SplitMix64-derived xoshiro256++ state of SmallRng::seed_from_u64(seed), and
// raw outputs of xoshiro256++ from an explicit state (known-answer checks).
void orc_rng_state(uint64_t seed, uint64_t* out4) {
  SmallRng r(seed);
  for (int i = 0; i < 4; i++) out4[i] = r.s[i];
}
void orc_xoshiro_outputs(const uint64_t* state4, uint64_t* out, size_t n) {
  SmallRng r(0);
  for (int i = 0; i < 4; i++) r.s[i] = state4[i];
  for (size_t i = 0; i < n; i++) out[i] = r.next_u64();
}
}

extern "C" {
// Sumcheck of f * eq(z, .) over all n variables by classic rounds on the full
// tables (ProductPolynomial::round, src/sumcheck/product_polynomial.rs:307-364).
// Challenges: `scripted` (n EF) or a fresh challenger (Poseidon2 from `seed`)
// observing (h(0), h(2)) and sampling each round.  Returns 0 when the final
// claim equals f(r) * eq(z, r).
int orc_sumcheck_eq(const uint32_t* f, int n, const uint32_t* point, const uint32_t* scripted, uint64_t seed,
                    uint32_t* out_polys, uint32_t* out_final) {
  Perms ps(seed);
  Challenger ch;
  ch.perm = &ps.p16;
  std::vector<EF> z = point_in(point, n);
  std::vector<EF> e(1ull << n), w = eq_table(z, ef_one());
  for (size_t i = 0; i < e.size(); i++) e[i] = ef_base(from_monty(f[i]));
  EF sum = ef_zero();
  for (size_t i = 0; i < e.size(); i++) sum = ef_add(sum, ef_mul(e[i], w[i]));
  for (int r = 0; r < n; r++) {
    EF c0, c2;
    sumcheck_coeffs(e, w, &c0, &c2);
    ef_out(c0, out_polys + 8 * r);
    ef_out(c2, out_polys + 8 * r + 4);
    EF x;
    if (scripted) x = ef_in(scripted + 4 * r);
    else { ch.observe_ef(c0); ch.observe_ef(c2); x = ch.sample_ef(); }
    fold_ext(e, x);
    fold_ext(w, x);
    sum = extrapolate_012(c0, ef_sub(sum, c0), c2, x);
  }
  ef_out(e[0], out_final);
  ef_out(w[0], out_final + 4);
  return ef_eq(ef_mul(e[0], w[0]), sum) ? 0 : 1;
}
}

extern "C" {
// Keccak for known-answer checks: the permutation on 25 lanes, and the
// 256-bit sponge over bytes with domain pad byte `pad` (0x01: Keccak-256,
// 0x06: SHA3-256).
void orc_keccak_f1600(uint64_t* lanes) { keccak_permute(lanes); }
void orc_keccak256(const uint8_t* msg, size_t len, uint8_t pad, uint8_t* out) { keccak256_bytes(msg, len, out, pad); }

// Grinding from a Keccak challenger whose input buffer holds `len` bytes and
// whose output is empty: returns the witness (Montgomery form) and the next 4
// field samples (Montgomery form) after it.
uint32_t orc_keccak_grind_bytes(const uint8_t* in, size_t len, int bits, uint32_t* next4) {
  Challenger ch;
  ch.keccak = true;
  ch.kin.assign(in, in + len);
  const u32 w = ch.grind(bits);
  for (int i = 0; i < 4; i++) next4[i] = to_monty(ch.sample());
  return to_monty(w);
}

// A script against the Keccak challenger (ops as orc_challenger_script:
// 0 observe args[i] (Montgomery), 1 sample (Montgomery out), 2 sample_bits
// args[i], 3 grind args[i] bits).  Returns the outputs written.
int orc_keccak_challenger_script(const uint32_t* ops, const uint32_t* args, int n, uint32_t* out) {
  Challenger ch;
  ch.keccak = true;
  int k = 0;
  for (int i = 0; i < n; i++) {
    switch (ops[i]) {
      case 0: ch.observe(from_monty(args[i])); break;
      case 1: out[k++] = to_monty(ch.sample()); break;
      case 2: out[k++] = (uint32_t)ch.sample_bits((int)args[i]); break;
      case 3: out[k++] = ch.grind((int)args[i]); break;
      default: return -1;
    }
  }
  return k;
}
}
