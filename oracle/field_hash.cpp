// Oracle: field, extension field, RNG, Poseidon2, challenger, Merkle tree, DFT.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include <algorithm>
#include <cassert>
#include <omp.h>

#include "oracle.hpp"

namespace orc {

u32 fpow(u32 a, u64 e) {
  u32 r = 1;
  while (e) {
    if (e & 1) r = fmul(r, a);
    a = fmul(a, a);
    e >>= 1;
  }
  return r;
}

static const u32 R_MOD_P = (u32)((1ull << 32) % P);
u32 to_monty(u32 x) { return (u32)((u64)x * R_MOD_P % P); }
u32 from_monty(u32 m) {
  static const u32 rinv = finv(R_MOD_P);
  return (u32)((u64)m * rinv % P);
}

// p - 1 = 127 * 2^24.  Plonky3's KoalaBear table of two-adic generators is the
// chain of squares of 3^127 (`p3-koala-bear`, external to the reference).
u32 two_adic_generator(int bits) {
  assert(bits >= 0 && bits <= 24);
  return fpow(3, 127ull << (24 - bits));
}

EF ef_mul(const EF& a, const EF& b) {
  u64 t[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) t[i + j] += (u64)a.c[i] * b.c[j] % P;
  EF r;
  for (int k = 0; k < 4; k++) {
    u64 hi = k + 4 < 7 ? t[k + 4] % P : 0;
    r.c[k] = (u32)((t[k] % P + (u64)EF_W * hi) % P);
  }
  return r;
}

EF ef_pow(EF a, u64 e) {
  EF r = ef_one();
  while (e) {
    if (e & 1) r = ef_mul(r, a);
    a = ef_mul(a, a);
    e >>= 1;
  }
  return r;
}

// |EF*| = p^4 - 1; a^(p^4 - 2) = a^-1.  p^4 - 2 does not fit in 64 bits, so
// raise in two steps: a^(p^4-2) = (a^(p^2))^(p^2) * a^-2 ... simpler: use the
// norm map.  N(a) = a * a^p * a^(p^2) * a^(p^3) lies in F, so
// a^-1 = a^p * a^(p^2) * a^(p^3) / N(a).
EF ef_inv(const EF& a) {
  EF ap = ef_pow(a, P);
  EF ap2 = ef_pow(ap, P);
  EF ap3 = ef_pow(ap2, P);
  EF num = ef_mul(ef_mul(ap, ap2), ap3);
  EF nrm = ef_mul(num, a);  // in F: nrm.c[1..3] == 0
  return ef_scale(num, finv(nrm.c[0]));
}

// ------------------------------------------------------------------ RNG
SmallRng::SmallRng(u64 state) {
  const u64 PHI = 0x9e3779b97f4a7c15ull;
  for (int i = 0; i < 4; i++) {
    state += PHI;
    u64 z = state;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    s[i] = z ^ (z >> 31);
  }
}

static inline u64 rotl(u64 x, int k) { return (x << k) | (x >> (64 - k)); }

u64 SmallRng::next_u64() {
  u64 result = rotl(s[0] + s[3], 23) + s[0];
  u64 t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}

u32 SmallRng::field() {
  for (;;) {
    u32 u31 = next_u32() >> 1;
    if (u31 < P) return from_monty(u31);
  }
}

// ------------------------------------------------------------------ Poseidon2
static inline void apply_mat4(u32* x) {
  u32 t01 = fadd(x[0], x[1]);
  u32 t23 = fadd(x[2], x[3]);
  u32 t0123 = fadd(t01, t23);
  u32 t01123 = fadd(t0123, x[1]);
  u32 t01233 = fadd(t0123, x[3]);
  u32 n3 = fadd(t01233, fadd(x[0], x[0]));
  u32 n1 = fadd(t01123, fadd(x[2], x[2]));
  u32 n0 = fadd(t01123, t01);
  u32 n2 = fadd(t01233, t23);
  x[0] = n0; x[1] = n1; x[2] = n2; x[3] = n3;
}

static void mds_light(u32* s, int w) {
  for (int c = 0; c < w; c += 4) apply_mat4(s + c);
  u32 sums[4] = {0, 0, 0, 0};
  for (int c = 0; c < w; c += 4)
    for (int k = 0; k < 4; k++) sums[k] = fadd(sums[k], s[c + k]);
  for (int i = 0; i < w; i++) s[i] = fadd(s[i], sums[i & 3]);
}

static inline u32 cube(u32 x) { return fmul(fmul(x, x), x); }

void Poseidon2::permute(u32* s) const {
  const int w = width;
  mds_light(s, w);
  for (int r = 0; r < 4; r++) {
    for (int i = 0; i < w; i++) s[i] = cube(fadd(s[i], ext_init[r * w + i]));
    mds_light(s, w);
  }
  for (int r = 0; r < rounds_p; r++) {
    s[0] = cube(fadd(s[0], internal[r]));
    u32 sum = 0;
    for (int i = 0; i < w; i++) sum = fadd(sum, s[i]);
    for (int i = 0; i < w; i++) s[i] = fadd(sum, fmul(diag[i], s[i]));
  }
  for (int r = 0; r < 4; r++) {
    for (int i = 0; i < w; i++) s[i] = cube(fadd(s[i], ext_term[r * w + i]));
    mds_light(s, w);
  }
}

// Internal-layer diagonal V (matrix 1 + diag(V)) of Plonky3's KoalaBear Poseidon2.
static std::vector<u32> koalabear_diag(int width) {
  auto inv2k = [](int k) { return finv(fpow(2, k)); };
  auto neg = [](u32 x) { return fneg(x); };
  if (width == 16) {
    return {neg(2), 1, 2, inv2k(1), 3, 4, neg(inv2k(1)), neg(3), neg(4), inv2k(8), inv2k(3),
            inv2k(24), neg(inv2k(8)), neg(inv2k(3)), neg(inv2k(4)), neg(inv2k(24))};
  }
  return {neg(2), 1, 2, inv2k(1), 3, 4, neg(inv2k(1)), neg(3), neg(4), inv2k(8), inv2k(2), inv2k(3),
          inv2k(4), inv2k(5), inv2k(6), inv2k(24), neg(inv2k(8)), neg(inv2k(3)), neg(inv2k(4)),
          neg(inv2k(5)), neg(inv2k(6)), neg(inv2k(7)), neg(inv2k(9)), neg(inv2k(24))};
}

static Poseidon2 draw_poseidon2(SmallRng& rng, int width) {
  Poseidon2 p;
  p.width = width;
  p.rounds_p = width == 16 ? 20 : 23;
  p.ext_init.resize(4 * width);
  p.ext_term.resize(4 * width);
  p.internal.resize(p.rounds_p);
  for (auto& v : p.ext_init) v = rng.field();
  for (auto& v : p.ext_term) v = rng.field();
  for (auto& v : p.internal) v = rng.field();
  p.diag = koalabear_diag(width);
  return p;
}

Poseidon2 poseidon2_from_seed(u64 seed, int width) {
  SmallRng rng(seed);
  return draw_poseidon2(rng, width);
}

void poseidon2_pair_from_seed(u64 seed, Poseidon2* p16, Poseidon2* p24) {
  SmallRng rng(seed);
  *p16 = draw_poseidon2(rng, 16);
  *p24 = draw_poseidon2(rng, 24);
}

// ------------------------------------------------------------------ Keccak
// FIPS 202 written from its definitions: the state as a 5x5 array of lanes,
// rho's offsets generated by walking (x, y) -> (y, 2x + 3y) from (1, 0) with
// offset (t + 1)(t + 2) / 2, and iota's round constants from the rc(t) LFSR
// (Algorithm 5: x^8 + x^6 + x^5 + x^4 + 1).
namespace {
bool keccak_rc_bit(int t) {
  if (t % 255 == 0) return true;
  // R = 10000000 as bits R[0..7]; each step prepends a 0 (R[8] is the bit
  // shifted out), feeds R[8] into bits 0, 4, 5, 6 and truncates to 8 bits
  uint8_t R[9] = {1, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 1; i <= t % 255; i++) {
    for (int k = 8; k > 0; k--) R[k] = R[k - 1];
    R[0] = 0;
    R[0] ^= R[8];
    R[4] ^= R[8];
    R[5] ^= R[8];
    R[6] ^= R[8];
  }
  return R[0] != 0;
}
struct KeccakTables {
  u64 rc[24];
  int rho[5][5];
  KeccakTables() {
    for (int ir = 0; ir < 24; ir++) {
      u64 v = 0;
      for (int j = 0; j <= 6; j++)
        if (keccak_rc_bit(j + 7 * ir)) v |= 1ull << ((1 << j) - 1);
      rc[ir] = v;
    }
    rho[0][0] = 0;
    int x = 1, y = 0;
    for (int t = 0; t < 24; t++) {
      rho[x][y] = ((t + 1) * (t + 2) / 2) % 64;
      const int nx = y, ny = (2 * x + 3 * y) % 5;
      x = nx;
      y = ny;
    }
  }
};
const KeccakTables& ktab() {
  static const KeccakTables t;
  return t;
}
u64 rol(u64 v, int r) { return r ? (v << r) | (v >> (64 - r)) : v; }
}  // namespace

void keccak_permute(u64* lanes) {
  const KeccakTables& T = ktab();
  u64 A[5][5];
  for (int x = 0; x < 5; x++)
    for (int y = 0; y < 5; y++) A[x][y] = lanes[x + 5 * y];
  for (int ir = 0; ir < 24; ir++) {
    u64 C[5], D[5], B[5][5];
    for (int x = 0; x < 5; x++) C[x] = A[x][0] ^ A[x][1] ^ A[x][2] ^ A[x][3] ^ A[x][4];
    for (int x = 0; x < 5; x++) D[x] = C[(x + 4) % 5] ^ rol(C[(x + 1) % 5], 1);
    for (int x = 0; x < 5; x++)
      for (int y = 0; y < 5; y++) A[x][y] ^= D[x];
    for (int x = 0; x < 5; x++)
      for (int y = 0; y < 5; y++) B[y][(2 * x + 3 * y) % 5] = rol(A[x][y], T.rho[x][y]);
    for (int x = 0; x < 5; x++)
      for (int y = 0; y < 5; y++) A[x][y] = B[x][y] ^ (~B[(x + 1) % 5][y] & B[(x + 2) % 5][y]);
    A[0][0] ^= T.rc[ir];
  }
  for (int x = 0; x < 5; x++)
    for (int y = 0; y < 5; y++) lanes[x + 5 * y] = A[x][y];
}

void keccak256_bytes(const uint8_t* msg, size_t len, uint8_t out[32], uint8_t pad) {
  // pad10*1 after the domain bits: message || pad || 0* || 0x80 (0x81 when they share
  // a byte), to a multiple of the 136-byte rate; pad 0x01 is Keccak-256, 0x06 SHA3-256
  std::vector<uint8_t> m(msg, msg + len);
  m.push_back(pad);
  while (m.size() % 136) m.push_back(0);
  m.back() |= 0x80;
  u64 st[25] = {0};
  for (size_t off = 0; off < m.size(); off += 136) {
    for (int i = 0; i < 17; i++) {
      u64 w = 0;
      for (int b = 7; b >= 0; b--) w = (w << 8) | m[off + 8 * i + b];
      st[i] ^= w;
    }
    keccak_permute(st);
  }
  for (int i = 0; i < 32; i++) out[i] = (uint8_t)(st[i / 8] >> (8 * (i % 8)));
}

// ------------------------------------------------------------------ challenger
void Challenger::duplex() {
  assert(in.size() <= 8);
  for (size_t i = 0; i < in.size(); i++) state[i] = in[i];
  in.clear();
  perm->permute(state);
  out.assign(state, state + 8);
}

uint8_t Challenger::sample_byte() {
  if (kout.empty()) {
    uint8_t h[32];
    keccak256_bytes(kin.data(), kin.size(), h);
    kout.assign(h, h + 32);
    kin.assign(h, h + 32);
  }
  uint8_t b = kout.back();
  kout.pop_back();
  return b;
}

void Challenger::observe(u32 v) {
  if (keccak) {
    kout.clear();
    for (int i = 0; i < 4; i++) kin.push_back((uint8_t)(v >> (8 * i)));
    return;
  }
  out.clear();
  in.push_back(v);
  if (in.size() == 8) duplex();
}

void Challenger::observe_digest(const Digest& d) {
  if (!keccak) {
    for (u32 x : d) observe(x);
    return;
  }
  kout.clear();
  for (u32 x : d)
    for (int i = 0; i < 4; i++) kin.push_back((uint8_t)(x >> (8 * i)));
}

u32 Challenger::sample() {
  if (keccak) {
    for (;;) {
      u32 v = 0;
      for (int i = 0; i < 4; i++) v |= (u32)sample_byte() << (8 * i);
      v &= 0x7fffffffu;
      if (v < P) return v;
    }
  }
  if (!in.empty() || out.empty()) duplex();
  u32 v = out.back();
  out.pop_back();
  return v;
}

// DuplexChallenger: a sampled field element's canonical value, masked.
// SerializingChallenger32: four sampled bytes read as a little-endian u32,
// masked, with no rejection.
u64 Challenger::sample_bits(int bits) {
  u64 r;
  if (keccak) {
    u32 v = 0;
    for (int i = 0; i < 4; i++) v |= (u32)sample_byte() << (8 * i);
    r = v;
  } else {
    r = sample();
  }
  return r & ((1ull << bits) - 1);
}

bool Challenger::check_witness(int bits, u32 w) {
  observe(w);
  return sample_bits(bits) == 0;
}

// Smallest witness w in [0, p) whose check passes.  Every candidate costs one
// permutation: the buffered inputs plus w overwrite the state prefix, and the
// sampled element is state[7] of the permuted state.
u32 Challenger::grind(int bits) {
  if (keccak) {   // each candidate: a copy of the challenger observes it and samples
    const u64 CH = 1 << 14;
    u32 found = P;
    for (u64 base = 0; base < P && found == P; base += CH) {
      const u64 end = std::min<u64>(base + CH, P);
      u32 local_min = P;
#pragma omp parallel for reduction(min : local_min) schedule(static)
      for (long long w = (long long)base; w < (long long)end; w++) {
        Challenger probe = *this;
        if (probe.check_witness(bits, (u32)w) && (u32)w < local_min) local_min = (u32)w;
      }
      found = local_min;
    }
    assert(found < P);
    bool ok = check_witness(bits, found);
    assert(ok);
    (void)ok;
    return found;
  }
  const u64 mask = (1ull << bits) - 1;
  const size_t L = in.size();
  const u64 CH = 1 << 16;
  u32 found = P;
  for (u64 base = 0; base < P && found == P; base += CH) {
    u64 end = std::min<u64>(base + CH, P);
    u32 local_min = P;
#pragma omp parallel for reduction(min : local_min) schedule(static)
    for (long long w = (long long)base; w < (long long)end; w++) {
      u32 st[16];
      memcpy(st, state, sizeof st);
      for (size_t i = 0; i < L; i++) st[i] = in[i];
      st[L] = (u32)w;
      perm->permute(st);
      if (((u64)st[7] & mask) == 0 && (u32)w < local_min) local_min = (u32)w;
    }
    found = local_min;
  }
  bool ok = check_witness(bits, found);
  assert(ok);
  (void)ok;
  return found;
}

// ------------------------------------------------------------------ Merkle
// PaddingFreeSponge<P, W, W-8, 8>: overwrite the rate part with each chunk,
// permute after each full or partial (non-empty) chunk, output state[0..8].
Digest sponge_hash(const Poseidon2& perm, const u32* vals, size_t len) {
  const int rate = perm.width - 8;
  u32 st[24] = {0};
  size_t pos = 0;
  for (;;) {
    int i = 0;
    for (; i < rate && pos < len; i++) st[i] = vals[pos++];
    if (i == 0) break;
    perm.permute(st);
    if (i < rate) break;
  }
  Digest d;
  for (int k = 0; k < 8; k++) d[k] = st[k];
  return d;
}

// The Keccak MMCS's lane sponge (PaddingFreeSponge<KeccakF, 25, 17, 4>) over
// u32 words taken two per u64 lane, low word first; 4 lanes out as 8 words.
Digest keccak_lanes_hash(const std::vector<u32>& words) {
  std::vector<u64> lanes;
  for (size_t i = 0; i < words.size(); i += 2)
    lanes.push_back((u64)words[i] | (i + 1 < words.size() ? (u64)words[i + 1] << 32 : 0));
  u64 st[25] = {0};
  size_t pos = 0;
  while (pos < lanes.size()) {   // no input: no permutation, a zero digest
    for (size_t k = 0; k < 17 && pos < lanes.size(); k++) st[k] = lanes[pos++];
    keccak_permute(st);
  }
  Digest d;
  for (int j = 0; j < 4; j++) {
    d[2 * j] = (u32)st[j];
    d[2 * j + 1] = (u32)(st[j] >> 32);
  }
  return d;
}

// A Keccak leaf: SerializingHasher over the row's field elements, each as its
// Montgomery word (the in-memory KoalaBear), packed two per lane.
Digest keccak_leaf(const u32* vals, size_t len) {
  std::vector<u32> w(len);
  for (size_t i = 0; i < len; i++) w[i] = to_monty(vals[i]);
  return keccak_lanes_hash(w);
}

Digest keccak_node(const Digest& l, const Digest& r) {
  std::vector<u32> w(l.begin(), l.end());
  w.insert(w.end(), r.begin(), r.end());
  return keccak_lanes_hash(w);
}

Digest compress(const Poseidon2& p16, const Digest& l, const Digest& r) {
  u32 st[16];
  for (int k = 0; k < 8; k++) { st[k] = l[k]; st[8 + k] = r[k]; }
  p16.permute(st);
  Digest d;
  for (int k = 0; k < 8; k++) d[k] = st[k];
  return d;
}

MerkleTree merkle_commit(const MerkleHashers& h, std::vector<u32> rows, size_t height, size_t width) {
  assert(height && !(height & (height - 1)));
  assert(rows.size() == height * width);
  MerkleTree t;
  t.height = height;
  t.width = width;
  t.rows = std::move(rows);
  std::vector<Digest> leaves(height);
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)height; i++)
    leaves[i] = h.keccak ? keccak_leaf(&t.rows[i * width], width) : sponge_hash(*h.leaf, &t.rows[i * width], width);
  t.layers.push_back(std::move(leaves));
  while (t.layers.back().size() > 1) {
    const auto& prev = t.layers.back();
    std::vector<Digest> next(prev.size() / 2);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)next.size(); i++)
      next[i] = h.keccak ? keccak_node(prev[2 * i], prev[2 * i + 1]) : compress(*h.node, prev[2 * i], prev[2 * i + 1]);
    t.layers.push_back(std::move(next));
  }
  return t;
}

void merkle_open(const MerkleTree& t, size_t index, std::vector<u32>* row, std::vector<Digest>* path) {
  row->assign(t.rows.begin() + index * t.width, t.rows.begin() + (index + 1) * t.width);
  path->clear();
  size_t idx = index;
  for (size_t l = 0; l + 1 < t.layers.size(); l++) {
    path->push_back(t.layers[l][idx ^ 1]);
    idx >>= 1;
  }
}

bool merkle_verify(const MerkleHashers& h, const Digest& root, size_t height, size_t index,
                   const std::vector<u32>& row, const std::vector<Digest>& path) {
  if (index >= height) return false;
  size_t depth = 0;
  while ((1ull << depth) < height) depth++;
  if (path.size() != depth) return false;
  Digest cur = h.keccak ? keccak_leaf(row.data(), row.size()) : sponge_hash(*h.leaf, row.data(), row.size());
  size_t idx = index;
  for (const auto& sib : path) {
    const Digest& l = (idx & 1) ? sib : cur;
    const Digest& r = (idx & 1) ? cur : sib;
    cur = h.keccak ? keccak_node(l, r) : compress(*h.node, l, r);
    idx >>= 1;
  }
  return cur == root;
}

// ------------------------------------------------------------------ DFT
static int log2_exact(size_t n) {
  int k = 0;
  while ((1ull << k) < n) k++;
  assert((1ull << k) == n);
  return k;
}

void dft_naive(std::vector<u32>& a) {
  size_t n = a.size();
  u32 w = two_adic_generator(log2_exact(n));
  std::vector<u32> out(n, 0);
  for (size_t k = 0; k < n; k++) {
    u32 wk = fpow(w, k), acc = 0, x = 1;
    for (size_t t = 0; t < n; t++) {
      acc = fadd(acc, fmul(a[t], x));
      x = fmul(x, wk);
    }
    out[k] = acc;
  }
  a = out;
}

void dft_radix2(std::vector<u32>& a) {
  size_t n = a.size();
  int lg = log2_exact(n);
  for (size_t i = 1, j = 0; i < n; i++) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (int s = 1; s <= lg; s++) {
    size_t len = 1ull << s, half = len >> 1;
    u32 wl = two_adic_generator(s);
    std::vector<u32> tw(half);
    tw[0] = 1;
    for (size_t j = 1; j < half; j++) tw[j] = fmul(tw[j - 1], wl);
    for (size_t i = 0; i < n; i += len)
      for (size_t j = 0; j < half; j++) {
        u32 u = a[i + j], v = fmul(a[i + j + half], tw[j]);
        a[i + j] = fadd(u, v);
        a[i + j + half] = fsub(u, v);
      }
  }
}

void dft_batch(std::vector<u32>& mat, size_t height, size_t width) {
#pragma omp parallel for schedule(dynamic)
  for (long long c = 0; c < (long long)width; c++) {
    std::vector<u32> col(height);
    for (size_t i = 0; i < height; i++) col[i] = mat[i * width + c];
    dft_radix2(col);
    for (size_t i = 0; i < height; i++) mat[i * width + c] = col[i];
  }
}

}  // namespace orc
