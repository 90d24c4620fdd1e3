// Device-resident WHIR commit/prove behind the C ABI of include/whir_gpu.h.
//
// The host runs the transcript (DuplexChallenger), the parameter schedule and
// the O(2^folding_factor)-sized steps; every pass over a polynomial, codeword
// or Merkle tree runs on the device and stays resident in HBM between steps.
// Per sumcheck round only the two round-polynomial values (32 bytes) come
// back; the challenge goes down as a kernel argument.
#include <chrono>
#include <functional>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/whir_gpu.h"
#include "host.h"
#include "internal.h"
#include "kernels.h"

using namespace kb;
using whir::Digest;
using whir::DuplexChallenger;

namespace {

thread_local std::string g_err;

struct Fail {
  int code;
};

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      g_err = std::string(#x) + ": " + hipGetErrorString(e_);                       \
      throw Fail{e_ == hipErrorOutOfMemory ? WHIRGPU_OOM : WHIRGPU_DEVICE};          \
    }                                                                               \
  } while (0)

#define KCHK(x)                                                  \
  do {                                                           \
    int r_ = (x);                                                \
    if (r_) {                                                    \
      g_err = std::string(#x) + " failed: " + std::to_string(r_); \
      throw Fail{WHIRGPU_DEVICE};                                \
    }                                                            \
  } while (0)

void bad(const std::string& msg, int code = WHIRGPU_BAD_PARAMS) {
  g_err = msg;
  throw Fail{code};
}

// Device allocations recycled by exact size, so repeated proofs of one shape
// allocate nothing after the first.
struct Pool {
  std::multimap<size_t, void*> free_;
  std::vector<void*> all_;
  size_t total = 0;   // bytes of every allocation made (in use or free)
  void* get(size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    auto it = free_.find(bytes);
    if (it != free_.end()) {
      void* p = it->second;
      free_.erase(it);
      return p;
    }
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, bytes));
    all_.push_back(p);
    total += bytes;
    return p;
  }
  void put(void* p, size_t bytes) {
    if (p) free_.emplace((bytes + 255) & ~size_t(255), p);
  }
  void release_all() {
    for (void* p : all_) (void)hipFree(p);
    all_.clear();
    free_.clear();
    total = 0;
  }
};

template <typename T>
struct Buf {
  T* p = nullptr;
  size_t n = 0;
};

}  // namespace

// Scratch of whirgpu_mle_eval_base_noalloc, the entry point that allocates
// nothing: the split eq tables of a point of up to 24 variables (2^16 + 256
// extension elements) and the block partial sums (2048): 1,085,440 bytes.  Each GPU
// holds its own copy until the process ends; every call on one GPU uses it,
// so calls that overlap on one GPU race.  Host code must use the device
// address from scratch_on_current_device(), never the array's name.
constexpr int NOALLOC_MAX_VARS = 24;
constexpr size_t NOALLOC_TAB_HI = size_t(1) << (NOALLOC_MAX_VARS - 8);
constexpr size_t NOALLOC_TAB_LO = 256;
constexpr size_t NOALLOC_PARTIALS = 2048;
static __device__ __align__(256) ef scratch[NOALLOC_TAB_HI + NOALLOC_TAB_LO + NOALLOC_PARTIALS];

static std::mutex g_scratch_mu;
static ef* g_scratch_ptr[64] = {};

// The address of `scratch` on the current device, looked up once per device
// (hipGetSymbolAddress, with that device current) and kept.
static ef* scratch_on_current_device() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) bad("device index out of range", WHIRGPU_DEVICE);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<ef*>(p);
  }
  return g_scratch_ptr[dev];
}

struct whirgpu_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  whirgpu_params params{};
  whir::WhirConfig cfg;
  P2C16 c16{};
  P2C24 c24{};
  u32* d_tw = nullptr;   // 2 * 4096 twiddles
  Twiddles tw{};
  Pool pool;
  size_t fixed_bytes = 0;   // device bytes of the buffers allocated at creation (not pooled)
  void* h_pin = nullptr;
  size_t h_pin_bytes = 0;
  ef* d_partial = nullptr;   // reduction scratch
  size_t partial_cap = 0;
  // Small reduction results are written by the kernels straight into
  // coherent pinned host memory (h_small; d_small is its device address) and
  // read after a stream synchronization, so no copy is enqueued for them.
  ef* d_small = nullptr;
  ef* h_small = nullptr;
  void read_small(ef* dst, size_t n) {
    sync();
    memcpy(dst, h_small, n * sizeof(ef));
  }
  ef* d_tabs = nullptr;      // eq tables: 8 slots of tab_cap
  size_t tab_cap = 0;
  // Host-visible coherent memory that kernels read or write directly (no copy
  // kernels): query indices in, opened rows and Merkle roots out.  Each holds
  // one call's data and is reused only after the stream has been synchronized.
  u64* h_idx = nullptr;
  u64* d_idx = nullptr;   // device address of h_idx
  u32* h_open = nullptr;
  u32* d_open = nullptr;  // device address of h_open
  size_t open_cap = 0;
  // Eight root slots, used in turn, so the next round's commitment can be
  // enqueued before this round's root is read.
  u32* h_root = nullptr;
  u32* d_root = nullptr;  // device address of h_root
  int root_next = 0;
  // Recorded after a round's transcript steps and sumcheck: the host replays
  // them once it has passed while the next commitment runs.
  hipEvent_t ev_round = nullptr;
  // Grinding result slots for host-driven grinds, refilled with p in bulk.
  static constexpr int GRIND_SLOTS = 64;
  u32* d_gslots = nullptr;
  int gslot = GRIND_SLOTS;
  u32* d_grind = nullptr;
  P2C16* d_c16 = nullptr;     // device copies of the constants (kernels that loop over
  P2C24* d_c24 = nullptr;     // permutations read them from memory)
  // device transcript of sumcheck rounds: challenger, grind base, round log, challenge
  static constexpr int SC_LOG_CAP = 64;
  DevChallenger* d_chal = nullptr;   // two states: a launch reads one and writes the other
  DevChallenger* h_chal = nullptr;   // coherent host memory: a phase's starting challenger
  DevChallenger* d_chal_init = nullptr;
  ScLog* h_sclog = nullptr;   // coherent host memory; the transcript kernel logs into it
  ScLog* d_sclog = nullptr;   // device address of h_sclog
  // device transcript of a round's steps before its sumcheck
  RoundLog* h_rlog = nullptr;   // coherent host memory: a WHIR round's log, then the final round's
  RoundLog* d_rlog = nullptr;   // device address of h_rlog
  // pinned host copy of the final polynomial, taken before the final sumcheck folds it
  static constexpr int FINAL_MAX_LOG = 6;
  ef* h_final = nullptr;
  DevChallenger* d_rchal = nullptr;   // two states the round's steps alternate between
  GrindBase* d_gbase = nullptr;       // base state of a long grind's candidates
  RoundBufs rb{};
  void* d_rbuf = nullptr;       // backing memory of rb
  int grind_blocks = 256;
  double timings[7] = {0};
  // live kernel timing: HIP events recorded on `st` around each launch
  bool prof = false;
  struct Rec {
    int stage;
    hipEvent_t a, b;
    double work;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> events;
  size_t next_event = 0;
  hipEvent_t event() {
    if (next_event == events.size()) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      events.push_back(e);
    }
    return events[next_event++];
  }

  template <typename T>
  Buf<T> alloc(size_t n) {
    Buf<T> b;
    b.n = n;
    b.p = static_cast<T*>(pool.get(n * sizeof(T)));
    return b;
  }
  template <typename T>
  void release(Buf<T>& b) {
    pool.put(b.p, b.n * sizeof(T));
    b.p = nullptr;
    b.n = 0;
  }
  template <typename T>
  T* pin(size_t n) {
    size_t bytes = n * sizeof(T);
    if (bytes > h_pin_bytes) {
      if (h_pin) (void)hipHostFree(h_pin);
      h_pin = nullptr;
      size_t cap = std::max(bytes, (size_t)4 << 20);
      HIPCHK(hipHostMalloc(&h_pin, cap, hipHostMallocDefault));
      h_pin_bytes = cap;
    }
    return static_cast<T*>(h_pin);
  }
  // Copy device -> host through pinned memory and wait.
  template <typename T>
  void d2h(T* dst, const T* src, size_t n) {
    T* p = pin<T>(n);
    HIPCHK(hipMemcpyAsync(p, src, n * sizeof(T), hipMemcpyDeviceToHost, st));
    sync();
    memcpy(dst, p, n * sizeof(T));
  }
  // Small host->device uploads are staged in a pinned ring and not waited
  // for: a slot is reused only after a stream synchronization, by which time
  // every earlier copy has completed.
  unsigned char* h_ring = nullptr;
  size_t ring_cap = 0, ring_pos = 0;
  void sync() {
    HIPCHK(hipStreamSynchronize(st));
    ring_pos = 0;
  }
  template <typename T>
  void h2d(T* dst, const T* src, size_t n) {
    size_t bytes = n * sizeof(T);
    size_t need = (bytes + 255) & ~size_t(255);
    if (need > ring_cap) {
      T* p = pin<T>(n);
      memcpy(p, src, bytes);
      HIPCHK(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, st));
      sync();
      return;
    }
    if (ring_pos + need > ring_cap) sync();
    memcpy(h_ring + ring_pos, src, bytes);
    HIPCHK(hipMemcpyAsync(dst, h_ring + ring_pos, bytes, hipMemcpyHostToDevice, st));
    ring_pos += need;
  }
};

struct whirgpu_statement {
  whirgpu_ctx* ctx = nullptr;
  int n = 0;
  u32* d_poly = nullptr;
  bool owned = false;
  size_t poly_bytes = 0;
  int seg_bits = 0;                       // SVO split: partial sums per 2^seg_bits segments
  std::vector<std::vector<ef>> points;    // user constraints, then commitment OOD ones
  std::vector<ef> evals;
  std::vector<std::vector<ef>> partials;  // per constraint, 2^seg_bits values
};

struct whirgpu_prover_data {
  whirgpu_ctx* ctx = nullptr;
  Buf<u32> rows, tree;
  u64 height = 0;
  int row_len = 0;   // base elements per leaf
  bool ext = false;
  bool root_on_host = false;   // the last tree kernel wrote the root to the context's h_root
  int root_slot = 0;           // at h_root + 8 * root_slot
};

struct whirgpu_proof {
  whirgpu_ctx* ctx = nullptr;
  whir::WhirProof p;
};

namespace {

// Stages reported by whirgpu_profile_read, with the unit of their work count.
enum Stage {
  ST_LDE_BASE,   // bytes: 4 * 2^m read + 4 * 2^(m+R) written
  ST_LDE_EXT,    // bytes: 16 * 2^m read + 16 * 2^(m+R) written
  ST_LEAVES,     // Poseidon2 permutations
  ST_TREE,       // Poseidon2 permutations
  ST_GRIND,      // candidates a sequential search checks (witness + 1)
  ST_SUMCHECK,   // bytes read and written by coefficient/fold passes
  ST_EQDOT,      // bytes of the evaluations read
  ST_COMBINE,    // bytes of weights read and written
  ST_COMPRESS,   // bytes read and written
  ST_EQ_EXPAND,  // bytes written
  ST_COUNT
};

// Times one launch group on the context stream when profiling is enabled.
struct Span {
  whirgpu_ctx* c;
  int stage;
  double work;
  hipEvent_t a{};
  bool open = false;
  Span(whirgpu_ctx* c_, int stage_, double work_) : c(c_), stage(stage_), work(work_) {
    if (c->prof) {
      a = c->event();
      HIPCHK(hipEventRecord(a, c->st));
      open = true;
    }
  }
  void close() {
    if (!open) return;
    hipEvent_t b = c->event();
    HIPCHK(hipEventRecord(b, c->st));
    c->recs.push_back({stage, a, b, work});
    open = false;
  }
  ~Span() {
    if (open) {
      try {
        close();
      } catch (...) {
      }
    }
  }
};

PointArg point_arg(const std::vector<ef>& z, int from = 0) {
  PointArg a;
  memset(&a, 0, sizeof a);
  for (size_t i = from; i < z.size() && i - from < 32; i++) a.z[i - from] = z[i];
  return a;
}

void use_device(const whirgpu_ctx* c);

// Grows the eq-table scratch for points of up to n coordinates.
void ensure_tabs(whirgpu_ctx* c, int n) {
  size_t need = (1ull << std::max(0, n - 8)) + 512;
  if (need <= c->tab_cap) return;
  HIPCHK(hipStreamSynchronize(c->st));
  ef* t = nullptr;
  HIPCHK(hipMalloc(&t, 8 * need * sizeof(ef)));
  (void)hipFree(c->d_tabs);
  c->d_tabs = t;
  c->fixed_bytes += 8 * (need - c->tab_cap) * sizeof(ef);
  c->tab_cap = need;
}

// Split eq tables of z (m coordinates) into the context's table slot `slot`:
// hi covers z[0 .. m-lo), lo covers z[m-lo .. m); the hi table carries `scale`.
void build_split_eq(whirgpu_ctx* c, const std::vector<ef>& z, int lo, const ef& scale, int slot, EqTables* t,
                    int k) {
  int m = (int)z.size();
  int hi = m - lo;
  if ((1ull << hi) + (1ull << lo) > c->tab_cap) bad("eq table too large for the context");
  ef* base = c->d_tabs + (size_t)slot * c->tab_cap;
  PointArg a = point_arg(z);
  KCHK(eq_split(a, hi, lo, scale, base, base + (1ull << hi), c->st));
  t->hi[k] = base;
  t->lo[k] = base + (1ull << hi);
}

// Blocks of a segmented eq-weighted sum: 2048 >> seg_bits, at most one per row.
int dot_blocks(int rows_log, int seg_bits) {
  long rows = 1L << rows_log;
  long want = 2048L >> seg_bits;
  if (want < 1) want = 1;
  return (int)std::min(rows, want);
}

// Segmented eq-weighted sum over the base polynomial for one point (see k_eq_dot).
std::vector<ef> base_partials(whirgpu_ctx* c, const u32* d_f, int n, int sb, const std::vector<ef>& z_rest) {
  int nrest = n - sb;
  int lo = std::min(8, nrest);
  EqTables t{};
  build_split_eq(c, z_rest, lo, ef_one(), 0, &t, 0);
  int nb = dot_blocks(nrest - lo, sb);
  if ((size_t)nb * (1ull << sb) > c->partial_cap) bad("reduction scratch too small");
  {
    Span sp(c, ST_EQDOT, 4.0 * (double)(1ull << n));
    KCHK(eq_dot_base(d_f, n, sb, lo, t, 1, c->d_partial, nb, c->d_small, c->st));
  }
  std::vector<ef> out(1ull << sb);
  c->read_small(out.data(), out.size());
  return out;
}

ef ext_mle(whirgpu_ctx* c, const ef* d_f, int m, const std::vector<ef>& z) {
  int lo = std::min(8, m);
  EqTables t{};
  build_split_eq(c, z, lo, ef_one(), 0, &t, 0);
  int nb = dot_blocks(m - lo, 0);
  {
    Span sp(c, ST_EQDOT, 16.0 * (double)(1ull << m));
    KCHK(eq_dot_ext(d_f, m, 0, lo, t, 1, c->d_partial, nb, c->d_small, c->st));
  }
  ef out;
  c->read_small(&out, 1);
  return out;
}

ef eq_point(const std::vector<ef>& a, const std::vector<ef>& b) { return whir::eq_at(a, b); }

// Smallest-witness grinding on the device; updates the challenger like
// GrindingChallenger::grind (observe witness, sample bits).
u32 gpu_grind(whirgpu_ctx* c, DuplexChallenger& ch, int bits) {
  u32 s[16];
  for (int i = 0; i < 16; i++) s[i] = ch.sponge[i];
  for (int i = 0; i < ch.in_len; i++) s[i] = ch.in[i];
  if (c->gslot == whirgpu_ctx::GRIND_SLOTS) {
    u32 init[whirgpu_ctx::GRIND_SLOTS];
    for (u32& v : init) v = P;
    c->h2d(c->d_gslots, init, whirgpu_ctx::GRIND_SLOTS);
    c->gslot = 0;
  }
  u32* d_res = c->d_gslots + c->gslot++;
  // Each launch covers sixteen times the expected number of candidates (a
  // witness lies inside with probability 1 - e^-16; the waves stop as soon as
  // one is known, so the size costs nothing); the search continues with the
  // next range until one is found.
  const u64 span = std::max<u64>(16ull << bits, 256);
  u32 w = P;
  Span sp(c, ST_GRIND, 0);
  for (u64 start = 0; start < P && w >= P; start += span) {
    KCHK(grind_range(c->d_c16, s, ch.in_len, bits, d_res, start, span, c->grind_blocks, c->st));
    c->d2h(&w, d_res, 1);
  }
  sp.close();
  if (!c->recs.empty() && c->prof) c->recs.back().work = (double)w + 1;
  if (w >= P) bad("no proof-of-work witness below p", WHIRGPU_GRIND_FAILED);
  u32 wm = to_monty(w);
  if (!ch.check_witness(bits, wm)) bad("grinding witness rejected", WHIRGPU_GRIND_FAILED);
  return wm;
}

// SumcheckData::observe_and_sample (src/whir/proof.rs:163-193).
ef observe_and_sample(whirgpu_ctx* c, whir::SumcheckData& sd, DuplexChallenger& ch, const ef& c0, const ef& c2,
                      int pow_bits) {
  sd.polynomial_evaluations.push_back({c0, c2});
  ch.observe_ef(c0);
  ch.observe_ef(c2);
  if (pow_bits > 0) sd.pow_witnesses.push_back(gpu_grind(c, ch, pow_bits));
  return ch.sample_ef();
}

// Device-resident product polynomial (ProductPolynomial, src/sumcheck/product_polynomial.rs).
struct SumcheckState {
  Buf<ef> e, w;
  int m = 0;
  ef sum = ef_zero();
  bool coeffs_valid = false;   // d_small[0..2] holds (h(0), h(2)) of the current arrays
  unsigned partial_nb = 0;     // d_partial holds their partial sums from this many blocks
  void invalidate() {
    coeffs_valid = false;
    partial_nb = 0;
  }
};

// Round challenge from the round polynomial (h(0), h(2)): the transcript
// (observe, grind, sample) in a proof, or a scripted sequence.
using ChallengeFn = std::function<ef(const ef&, const ef&)>;

// ProductPolynomial::round (product_polynomial.rs:307-364).
ef sumcheck_round(whirgpu_ctx* c, SumcheckState& s, const ChallengeFn& next) {
  if (s.m < 1) bad("sumcheck round on a constant polynomial", WHIRGPU_BAD_STATE);
  u64 h = 1ull << (s.m - 1);
  if (!s.coeffs_valid) {
    Span sp(c, ST_SUMCHECK, 64.0 * (double)h);
    KCHK(sc_coeffs(s.e.p, s.w.p, h, c->d_partial, c->d_small, c->st));
  }
  ef cc[2];
  c->read_small(cc, 2);
  ef r = next(cc[0], cc[1]);
  {
    Span sp(c, ST_SUMCHECK, 96.0 * (double)h);
    KCHK(sc_fold(s.e.p, s.w.p, h, r, c->d_partial, c->d_small, c->st));
  }
  s.coeffs_valid = h >= 2;
  s.partial_nb = 0;
  s.m -= 1;
  s.sum = whir::extrapolate_012(cc[0], ef_sub(s.sum, cc[0]), cc[1], r);
  return r;
}

// `rounds` sumcheck rounds of a proof with the transcript on the device: per
// round the coefficient partials, the transcript wave (observe, grind base or
// challenge), the grind and the transcript wave after it, and the fold by the
// challenge in device memory; no host round trip until the end.  The host
// then replays the logged rounds on its own challenger (observe (h(0), h(2)),
// check the witness, sample) and requires every challenge to equal the
// device's, so the proof and the challenger state are exactly those of the
// host-driven rounds (SumcheckData::observe_and_sample, src/whir/proof.rs:163-193;
// ProductPolynomial::round, src/sumcheck/product_polynomial.rs:307-364).
struct ScLaunch {
  int rounds = 0;
  int pow_bits = 0;
  int log_base = 0;                     // first ScLog entry of these rounds
  const DevChallenger* out = nullptr;   // device challenger state after the last round
  std::vector<size_t> grind_recs;
};

// Uploads the host challenger as the device phase's starting state.
void upload_challenger(whirgpu_ctx* c, const DuplexChallenger& ch) {
  DevChallenger dc{};
  for (int i = 0; i < 16; i++) dc.sponge[i] = ch.sponge[i];
  for (int i = 0; i < ch.in_len; i++) dc.in[i] = ch.in[i];
  for (int i = 0; i < ch.out_len; i++) dc.out[i] = ch.out[i];
  dc.in_len = (u32)ch.in_len;
  dc.out_len = (u32)ch.out_len;
  // h_chal is free: every phase ends with a synchronization before the next
  // one uploads its state.
  *c->h_chal = dc;
}

// Grid of a device-transcript grind over all of [0, p) (so the search always
// ends): sized to the expected search (16 * 2^bits candidates), one pass of
// it at most, so short grinds leave the device to other work.
int sc_grind_blocks(const whirgpu_ctx* c, int pow_bits) {
  const u64 want = ((16ull << pow_bits) + 255) / 256;
  return pow_bits <= 12 ? c->grind_blocks : (int)std::max<u64>(1, std::min<u64>((u64)c->grind_blocks, want));
}

// Launches `rounds` device-transcript sumcheck rounds starting from the
// challenger at cin (device-visible memory).
ScLaunch sumcheck_launch(whirgpu_ctx* c, SumcheckState& s, int rounds, int pow_bits, const DevChallenger* cin,
                         int log_base = 0) {
  ScLaunch L;
  L.rounds = rounds;
  L.pow_bits = pow_bits;
  L.log_base = log_base;
  L.out = cin;
  if (rounds <= 0) return L;
  if (log_base < 0 || log_base + rounds > whirgpu_ctx::SC_LOG_CAP)
    bad("too many sumcheck rounds in one phase", WHIRGPU_BAD_STATE);
  if (s.m < rounds) bad("sumcheck round on a constant polynomial", WHIRGPU_BAD_STATE);
  std::vector<size_t>& grind_recs = L.grind_recs;
  // Each launch reads the challenger from `cin` and writes the next state to
  // the other buffer; coefficient partials and grind results alternate too,
  // since every block of a launch reads what block 0 of the same launch
  // would otherwise overwrite.
  DevChallenger* chal[2] = {c->d_chal, c->d_chal + 1};
  ef* part[2] = {c->d_partial, c->d_partial + 4096};
  u32* gres[2] = {c->d_grind, c->d_grind + 64};
  auto next_state = [&]() { return cin == chal[0] ? chal[1] : chal[0]; };
  int pp = 0;
  unsigned nb = 0;
  {
    const u64 h = 1ull << (s.m - 1);
    Span sp(c, ST_SUMCHECK, 64.0 * (double)h);
    KCHK(sc_partials(s.e.p, s.w.p, h, part[0], &nb, gres[0], c->st));
  }
  for (int i = 0; i < rounds; i++) {
    const u64 h = 1ull << (s.m - 1);
    u32* g = gres[i & 1];
    u32* g_next = gres[(i + 1) & 1];
    int fold_mode = TR_COEFFS_SAMPLE;
    if (pow_bits > 0) {
      DevChallenger* cout = next_state();
      ScTr tr{cin, cout, c->d_sclog + log_base + i, part[pp], g, nullptr, (int)nb, pow_bits};
      tr.gbase = c->d_gbase;
      Span sp(c, ST_GRIND, 0);
      KCHK(grind_range_tr(c->d_c16, tr, pow_bits, 0, P, sc_grind_blocks(c, pow_bits), c->st));
      sp.close();
      if (c->prof) grind_recs.push_back(c->recs.size() - 1);
      cin = cout;
      fold_mode = TR_WITNESS_SAMPLE;
    }
    DevChallenger* cout = next_state();
    ScTr tr{cin, cout, c->d_sclog + log_base + i, part[pp], g, g_next, (int)nb, pow_bits};
    {
      Span sp(c, ST_SUMCHECK, 96.0 * (double)h);
      KCHK(sc_fold_tr(s.e.p, s.w.p, h, fold_mode, tr, c->d_c16, part[pp ^ 1], &nb, c->st));
    }
    cin = cout;
    pp ^= 1;
    s.m -= 1;
  }
  L.out = cin;
  // d_partial is scratch for other kernels between phases
  s.invalidate();
  return L;
}

// After a synchronization: replays the logged rounds on the host challenger
// and returns the challenges.
std::vector<ef> sumcheck_replay(whirgpu_ctx* c, SumcheckState& s, const ScLaunch& la, whir::SumcheckData& sd,
                                DuplexChallenger& ch) {
  std::vector<ef> rs;
  const int rounds = la.rounds, pow_bits = la.pow_bits;
  const std::vector<size_t>& grind_recs = la.grind_recs;
  std::vector<ScLog> log(c->h_sclog + la.log_base, c->h_sclog + la.log_base + rounds);
  for (int i = 0; i < rounds; i++) {
    const ScLog& L = log[i];
    if (L.err) bad("device transcript: no valid proof-of-work witness", WHIRGPU_GRIND_FAILED);
    sd.polynomial_evaluations.push_back({L.c0, L.c2});
    ch.observe_ef(L.c0);
    ch.observe_ef(L.c2);
    if (pow_bits > 0) {
      if (!ch.check_witness(pow_bits, L.witness)) bad("device transcript: witness rejected", WHIRGPU_GRIND_FAILED);
      sd.pow_witnesses.push_back(L.witness);
      if (i < (int)grind_recs.size()) c->recs[grind_recs[i]].work = (double)from_monty(L.witness) + 1;
    }
    ef r = ch.sample_ef();
    if (!ef_eq(r, L.r)) bad("device transcript diverged from the host challenger", WHIRGPU_BAD_STATE);
    s.sum = whir::extrapolate_012(L.c0, ef_sub(s.sum, L.c0), L.c2, r);
    rs.push_back(r);
  }
  return rs;
}

// `rounds` sumcheck rounds with the transcript on the device, from the host
// challenger's state.
std::vector<ef> device_sumcheck(whirgpu_ctx* c, SumcheckState& s, int rounds, int pow_bits, whir::SumcheckData& sd,
                                DuplexChallenger& ch) {
  if (rounds <= 0) return {};
  upload_challenger(c, ch);
  ScLaunch la = sumcheck_launch(c, s, rounds, pow_bits, c->d_chal_init);
  c->sync();
  return sumcheck_replay(c, s, la, sd, ch);
}

// Adds sum_i scale_i * eq(z_i, x) to the weights, four points per launch.
void add_eq_weights(whirgpu_ctx* c, SumcheckState& s, const std::vector<std::vector<ef>>& pts,
                    const std::vector<ef>& scales, bool accumulate_first) {
  int m = s.m;
  int lo = std::min(8, m);
  size_t K = pts.size();
  if (K == 0 && !accumulate_first) {
    EqTables t{};
    Span sp(c, ST_COMBINE, 16.0 * (double)(1ull << m));
    KCHK(eq_combine(s.w.p, 1ull << m, lo, t, 0, 0, c->st));
    return;
  }
  for (size_t b = 0; b < K; b += 4) {
    EqTables t{};
    int kk = (int)std::min<size_t>(4, K - b);
    for (int k = 0; k < kk; k++) build_split_eq(c, pts[b + k], lo, scales[b + k], k, &t, k);
    bool acc = b > 0 || accumulate_first;
    Span sp(c, ST_COMBINE, (acc ? 32.0 : 16.0) * (double)(1ull << m));
    KCHK(eq_combine(s.w.p, 1ull << m, lo, t, kk, acc ? 1 : 0, c->st));
  }
}

// Initial sumcheck from the base evaluations (Sumcheck::from_base_evals,
// src/sumcheck/sumcheck_prover.rs:237-292).  Classic and SVO produce the same
// transcript (asserted by the reference at src/sumcheck/tests.rs:473-487); the
// device always takes the split route when n > l0: per constraint, the
// partial sums P(s) = sum_x eq(z_rest, x) f(s, x) over the l0 leading
// variables s (computed once at evaluate time) give every round polynomial
//   h_i(X) = eq(z_<i, r_<i) eq(z_i, X) sum_s' eq(z_>i, s') P~(r_<i, X, s'),
// and afterwards f is folded by all l0 challenges at once.
// `ap` holds the combination weights alpha^j of the constraints.
std::vector<ef> initial_sumcheck(whirgpu_ctx* c, const whirgpu_statement* st, int l0, const std::vector<ef>& ap,
                                 const ChallengeFn& next, SumcheckState& s) {
  const int n = st->n;
  const size_t K = st->points.size();
  s.sum = ef_zero();
  for (size_t j = 0; j < K; j++) s.sum = ef_add(s.sum, ef_mul(ap[j], st->evals[j]));
  std::vector<ef> rs;
  if (l0 > 0 && st->seg_bits == l0 && n > l0) {
    std::vector<std::vector<ef>> P = st->partials;
    for (int i = 0; i < l0; i++) {
      ef c0 = ef_zero(), c2 = ef_zero();
      for (size_t j = 0; j < K; j++) {
        const auto& z = st->points[j];
        size_t half = P[j].size() / 2;
        std::vector<ef> z_after(z.begin() + i + 1, z.begin() + l0);
        auto eqa = whir::eq_table_host(z_after, ef_one());
        ef A0 = ef_zero(), A1 = ef_zero();
        for (size_t q = 0; q < half; q++) {
          A0 = ef_add(A0, ef_mul(eqa[q], P[j][q]));
          A1 = ef_add(A1, ef_mul(eqa[q], P[j][half + q]));
        }
        ef A2 = ef_sub(ef_dbl(A1), A0);
        std::vector<ef> z_before(z.begin(), z.begin() + i);
        ef pre = eq_point(z_before, rs);
        const ef& zi = z[i];
        ef e0 = ef_sub(ef_one(), zi);
        ef e2 = ef_sub(ef_add(ef_dbl(zi), zi), ef_one());
        c0 = ef_add(c0, ef_mul(ap[j], ef_mul(pre, ef_mul(e0, A0))));
        c2 = ef_add(c2, ef_mul(ap[j], ef_mul(pre, ef_mul(e2, A2))));
      }
      ef r = next(c0, c2);
      s.sum = whir::extrapolate_012(c0, ef_sub(s.sum, c0), c2, r);
      rs.push_back(r);
      for (auto& Pj : P) {
        size_t half = Pj.size() / 2;
        for (size_t q = 0; q < half; q++) Pj[q] = ef_add(Pj[q], ef_mul(r, ef_sub(Pj[half + q], Pj[q])));
        Pj.resize(half);
      }
    }
    // f(rs, x) for all x, and W(x) = sum_j alpha^j eq(z_j[..l0], rs) eq(z_j[l0..], x)
    s.m = n - l0;
    s.e = c->alloc<ef>(1ull << s.m);
    s.w = c->alloc<ef>(1ull << s.m);
    auto eqrs = whir::eq_table_host(rs, ef_one());
    ef* d_eqrs = c->d_tabs + 7 * c->tab_cap;
    c->h2d(d_eqrs, eqrs.data(), eqrs.size());
    {
      Span sp(c, ST_COMPRESS, 4.0 * (double)(1ull << n) + 16.0 * (double)(1ull << (n - l0)));
      KCHK(compress_multi(st->d_poly, n, l0, d_eqrs, s.e.p, c->st));
    }
    std::vector<std::vector<ef>> rests;
    std::vector<ef> scales;
    for (size_t j = 0; j < K; j++) {
      const auto& z = st->points[j];
      std::vector<ef> zs(z.begin(), z.begin() + l0);
      rests.emplace_back(z.begin() + l0, z.end());
      scales.push_back(ef_mul(ap[j], eq_point(zs, rs)));
    }
    add_eq_weights(c, s, rests, scales, false);
    s.invalidate();
    return rs;
  }
  // small instances: classic rounds on the extension-embedded evaluations
  s.m = n;
  s.e = c->alloc<ef>(1ull << n);
  s.w = c->alloc<ef>(1ull << n);
  KCHK(base_to_ext(st->d_poly, 1ull << n, s.e.p, c->st));
  add_eq_weights(c, s, st->points, ap, false);
  s.invalidate();
  for (int i = 0; i < l0; i++) rs.push_back(sumcheck_round(c, s, next));
  return rs;
}

int ilog2u(u64 x) {
  int k = 0;
  while ((x >> (k + 1)) != 0) k++;
  return k;
}

// Opens `idx` on a committed matrix and returns the openings plus each row
// as extension elements.
void parse_openings(whirgpu_ctx* c, const whirgpu_prover_data* pd, int nq, std::vector<whir::QueryOpening>* out,
                    std::vector<std::vector<ef>>* rows_ef, const u32* h = nullptr);

void open_rows(whirgpu_ctx* c, const whirgpu_prover_data* pd, const std::vector<u64>& idx,
               std::vector<whir::QueryOpening>* out, std::vector<std::vector<ef>>* rows_ef) {
  int nq = (int)idx.size();
  if (nq == 0) return;
  int depth = ilog2u(pd->height);
  size_t per = (size_t)pd->row_len + 8ull * depth;
  if (per * nq > c->open_cap || nq > 4096) bad("too many queries for the opening buffer");
  // h_idx / h_open are free: every opening ends with a synchronization.
  memcpy(c->h_idx, idx.data(), idx.size() * sizeof(u64));
  KCHK(merkle_open(pd->rows.p, pd->row_len, pd->tree.p, pd->height, c->d_idx, nq, c->d_open, c->st));
  c->sync();
  parse_openings(c, pd, nq, out, rows_ef);
}

// The openings of nq queries written to h (h_open by default) by an opening
// kernel (after a synchronization), as proof entries and rows of extension
// elements.
void parse_openings(whirgpu_ctx* c, const whirgpu_prover_data* pd, int nq, std::vector<whir::QueryOpening>* out,
                    std::vector<std::vector<ef>>* rows_ef, const u32* h) {
  const int depth = ilog2u(pd->height);
  const size_t per = (size_t)pd->row_len + 8ull * depth;
  if (!h) h = c->h_open;
  for (int q = 0; q < nq; q++) {
    const u32* b = h + per * q;
    whir::QueryOpening o;
    o.extension = pd->ext;
    o.values.assign(b, b + pd->row_len);
    for (int l = 0; l < depth; l++) {
      Digest d;
      for (int k = 0; k < 8; k++) d[k] = b[pd->row_len + 8 * l + k];
      o.proof.push_back(d);
    }
    std::vector<ef> r;
    if (pd->ext) {
      for (int i = 0; i < pd->row_len; i += 4) r.push_back(ef{{b[i], b[i + 1], b[i + 2], b[i + 3]}});
    } else {
      for (int i = 0; i < pd->row_len; i++) r.push_back(ef_base(b[i]));
    }
    rows_ef->push_back(std::move(r));
    out->push_back(std::move(o));
  }
}

// Leaf hashes and the tree above them, timed separately.
void commit_tree(whirgpu_ctx* c, whirgpu_prover_data* pd) {
  const int lw = (int)c->params.leaf_hash_width;
  const int rate = lw - 8;
  {
    Span sp(c, ST_LEAVES, (double)pd->height * (double)((pd->row_len + rate - 1) / rate));
    KCHK(merkle_leaves(pd->rows.p, pd->row_len, pd->height, lw, c->d_c16, c->d_c24, pd->tree.p, c->st));
  }
  Span sp(c, ST_TREE, (double)(pd->height - 1));
  pd->root_slot = c->root_next;
  c->root_next = (c->root_next + 1) & 7;
  int rc = merkle_tree(pd->height, c->c16, pd->tree.p, c->st, c->d_root + 8 * pd->root_slot);
  if (rc < 0) KCHK(rc);
  pd->root_on_host = rc == 0;   // the tree's last kernel wrote the root to h_root
}

// The root: written to its h_root slot by the tree's last kernel, or (single
// leaf) copied.  completed: the tree kernels are known to have finished.
Digest read_root(whirgpu_ctx* c, const whirgpu_prover_data* pd, bool completed = false) {
  Digest d;
  if (pd->root_on_host) {
    if (!completed) c->sync();
    memcpy(d.data(), c->h_root + 8 * pd->root_slot, sizeof d);
  } else {
    c->d2h(d.data(), pd->tree.p + (2 * pd->height - 2) * 8, 8);
  }
  return d;
}

void observe_digest(DuplexChallenger& ch, const Digest& d) {
  for (u32 x : d) ch.observe(x);
}

// Whether a round's steps after its commitment fit the device transcript's
// buffers (otherwise the host drives them).
bool device_round_fits(const whirgpu_ctx* c, const whir::RoundConfig& rp, int m, const whirgpu_prover_data* prev) {
  const int lo = std::min(8, m), hi = m - lo;
  if (!prev || m < 1 || m > RT_MAX_VARS) return false;
  if (rp.ood_samples > RT_MAX_OOD || rp.num_queries < 1 || rp.num_queries > RT_MAX_QUERIES) return false;
  if ((1ull << hi) + (1ull << lo) > c->tab_cap) return false;
  const size_t per = (size_t)prev->row_len + 8ull * ilog2u(prev->height);
  return per * (size_t)rp.num_queries <= c->open_cap / 2;   // the second half may hold the final round's
}

// A round's steps from its commitment's root to the end of its sumcheck
// (prover/mod.rs:192-303) with the transcript on the device: the OOD samples
// and their answers, the proof of work, the STIR queries and their openings,
// gamma, the weight updates and the sumcheck run as one stream of kernels,
// each reading its challenger from the previous one, and one event at the
// end.  The caller may enqueue the next round's commitment (which needs none
// of the host's results) before device_round_replay waits for that event;
// the host then replays every step on its own challenger and requires each
// challenge, answer and query index to equal the device's, so the proof and
// the challenger state are those of the host-driven steps.
struct RoundLaunch {
  ScLaunch la;
  size_t grind_rec = SIZE_MAX;
};

RoundLaunch device_round_launch(whirgpu_ctx* c, SumcheckState& s, const whir::WhirConfig& cfg, int round, int kn,
                                const whirgpu_prover_data* pd, const whirgpu_prover_data* prev,
                                const DuplexChallenger& ch) {
  const auto& rp = cfg.round_parameters[round];
  const int m = s.m;
  const int n_ood = rp.ood_samples;
  const int lo = std::min(8, m), hi = m - lo;
  upload_challenger(c, ch);
  DevChallenger* rch[2] = {c->d_rchal, c->d_rchal + 1};
  int ri = 0;
  const DevChallenger* cin = c->d_chal_init;
  const u32* obs = pd->tree.p + (2 * pd->height - 2) * 8;   // the root
  int n_obs = 8;
  ef* log_obs = nullptr;
  for (int i = 0; i < n_ood; i++) {
    DevChallenger* cout = rch[ri];
    ri ^= 1;
    const StepTr st{cin, cout, obs, n_obs, log_obs};
    // table slot i: the weight update below uses the same tables, scaled by gamma^i
    ef* tab = c->d_tabs + (size_t)i * c->tab_cap;
    KCHK(ood_eq_tables(st, &c->d_rlog->y[i], c->rb.pts + (size_t)i * RT_MAX_VARS, m, hi, lo, tab, tab + (1ull << hi),
                       c->d_c16, c->st, i == 0 && rp.pow_bits > 0 ? c->rb.grind : nullptr));
    EqTables t{};
    t.hi[0] = tab;
    t.lo[0] = tab + (1ull << hi);
    {
      Span sp(c, ST_EQDOT, 16.0 * (double)(1ull << m));
      KCHK(eq_dot_ext(s.e.p, m, 0, lo, t, 1, c->d_partial, dot_blocks(m - lo, 0), c->rb.ev + i, c->st));
    }
    cin = cout;
    obs = reinterpret_cast<const u32*>(c->rb.ev + i);
    n_obs = 4;
    log_obs = &c->d_rlog->ev[i];
  }
  RoundLaunch rl;
  size_t& grind_rec = rl.grind_rec;
  if (rp.pow_bits > 0) {
    // the first OOD step reset the grind slot; without OOD samples it is reset here
    if (n_ood == 0) HIPCHK(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(c->rb.grind), (int)P, 1, c->st));
    DevChallenger* cout = rch[ri];
    ri ^= 1;
    ScTr tr{};
    tr.ch_in = cin;
    tr.ch_out = cout;
    tr.grind = c->rb.grind;
    tr.pow_bits = rp.pow_bits;
    tr.obs = obs;
    tr.n_obs = n_obs;
    tr.log_obs = log_obs;
    tr.gbase = c->d_gbase;
    const u64 want = ((16ull << rp.pow_bits) + 255) / 256;
    const int gblocks =
        rp.pow_bits <= 12 ? c->grind_blocks : (int)std::max<u64>(1, std::min<u64>((u64)c->grind_blocks, want));
    Span sp(c, ST_GRIND, 0);
    KCHK(grind_range_tr(c->d_c16, tr, rp.pow_bits, 0, P, gblocks, c->st, TR_OBS_GRIND));
    sp.close();
    if (c->prof) grind_rec = c->recs.size() - 1;
    cin = cout;
    obs = nullptr;
    n_obs = 0;
    log_obs = nullptr;
  }
  {
    DevChallenger* cout = rch[ri];
    ri ^= 1;
    const StepTr st{cin, cout, obs, n_obs, log_obs};
    KCHK(round_queries(st, c->rb.grind, rp.pow_bits, rp.domain_size, cfg.ff(round), rp.num_queries,
                       rp.folded_domain_gen, n_ood, c->d_rlog, c->rb, c->d_c16, c->st));
    cin = cout;
  }
  // openings of the previous commitment into h_open, read after the round
  KCHK(merkle_open_dev(prev->rows.p, prev->row_len, prev->tree.p, prev->height, c->rb.idx, c->rb.nq,
                       rp.num_queries, c->d_open, c->st));
  // Constraint::combine_packed (src/whir/constraints/mod.rs:234-248): the OOD
  // points' eq (the tables of the OOD steps) weighted by gamma^i, then the
  // queries' select terms
  for (int b = 0; b < n_ood; b += 4) {
    EqTables t{};
    const int kk = std::min(4, n_ood - b);
    for (int k = 0; k < kk; k++) {
      ef* base = c->d_tabs + (size_t)(b + k) * c->tab_cap;
      t.hi[k] = base;
      t.lo[k] = base + (1ull << hi);
    }
    Span sp(c, ST_COMBINE, 32.0 * (double)(1ull << m));
    KCHK(eq_combine(s.w.p, 1ull << m, lo, t, kk, 1, c->st, c->rb.gpow + b));
  }
  {
    const int ns = rp.num_queries;
    Buf<ef> A = c->alloc<ef>((size_t)ns << (m - m / 2));
    Buf<u32> B = c->alloc<u32>((size_t)ns << (m / 2));
    {
      Span sp(c, ST_COMBINE, 32.0 * (double)(1ull << m));
      KCHK(select_combine(s.w.p, m, c->rb.selz, c->rb.gpow + n_ood, ns, c->rb.nq, A.p, B.p, c->st));
    }
    c->release(A);   // reuse is ordered on the same stream
    c->release(B);
  }
  s.invalidate();
  rl.la = sumcheck_launch(c, s, kn, rp.folding_pow_bits, cin);
  HIPCHK(hipEventRecord(c->ev_round, c->st));
  return rl;
}

// Waits for the round launched by device_round_launch (not for what was
// enqueued after it) and replays its transcript on the host challenger.
std::vector<ef> device_round_replay(whirgpu_ctx* c, SumcheckState& s, const whir::WhirConfig& cfg, int round,
                                    const whirgpu_prover_data* pd, const whirgpu_prover_data* prev,
                                    const std::vector<ef>& fold_rand, whir::WhirRoundProof& R, DuplexChallenger& ch,
                                    const RoundLaunch& rl) {
  const auto& rp = cfg.round_parameters[round];
  const int n_ood = rp.ood_samples;
  HIPCHK(hipEventSynchronize(c->ev_round));
  R.commitment = read_root(c, pd, true);
  observe_digest(ch, R.commitment);
  const RoundLog& L = *c->h_rlog;
  for (int i = 0; i < n_ood; i++) {
    const ef y = ch.sample_ef();
    if (!ef_eq(y, L.y[i])) bad("device round transcript diverged (OOD challenge)", WHIRGPU_BAD_STATE);
    ch.observe_ef(L.ev[i]);
    R.ood_answers.push_back(L.ev[i]);
  }
  if (rp.pow_bits > 0) {
    if (L.err) bad("device round transcript: no valid proof-of-work witness", WHIRGPU_GRIND_FAILED);
    if (!ch.check_witness(rp.pow_bits, L.witness)) bad("device round transcript: witness rejected", WHIRGPU_GRIND_FAILED);
    R.pow_witness = L.witness;
    if (rl.grind_rec != SIZE_MAX) c->recs[rl.grind_rec].work = (double)from_monty(L.witness) + 1;
  }
  ch.sample();
  const auto idx = whir::get_challenge_stir_queries(rp.domain_size, cfg.ff(round), rp.num_queries, ch);
  if (idx.size() != L.nq) bad("device round transcript diverged (query count)", WHIRGPU_BAD_STATE);
  for (size_t q = 0; q < idx.size(); q++)
    if (idx[q] != L.idx[q]) bad("device round transcript diverged (query index)", WHIRGPU_BAD_STATE);
  std::vector<std::vector<ef>> rows;
  parse_openings(c, prev, (int)idx.size(), &R.queries, &rows);
  const ef gamma = ch.sample_ef();
  if (!ef_eq(gamma, L.gamma)) bad("device round transcript diverged (gamma)", WHIRGPU_BAD_STATE);
  const auto gp = whir::powers(gamma, (size_t)n_ood + idx.size());
  for (int i = 0; i < n_ood; i++) s.sum = ef_add(s.sum, ef_mul(gp[i], L.ev[i]));
  for (size_t j = 0; j < idx.size(); j++)
    s.sum = ef_add(s.sum, ef_mul(gp[n_ood + j], whir::mle_host(rows[j], fold_rand)));
  return sumcheck_replay(c, s, rl.la, R.sumcheck, ch);
}

// Whether the final round's steps fit the device transcript: the device
// observes the final polynomial word by word, so it must be small.  Its
// openings use the second half of the opening buffer.
bool device_final_fits(const whirgpu_ctx* c, const whir::WhirConfig& cfg, int m, const whirgpu_prover_data* prev) {
  if (!prev || m < 0 || m > whirgpu_ctx::FINAL_MAX_LOG) return false;
  if (cfg.final_queries < 1 || cfg.final_queries > RT_MAX_QUERIES) return false;
  if (cfg.final_sumcheck_rounds > whirgpu_ctx::SC_LOG_CAP / 2) return false;
  const size_t per = (size_t)prev->row_len + 8ull * ilog2u(prev->height);
  return per * (size_t)cfg.final_queries <= c->open_cap / 2;
}

struct FinalLaunch {
  ScLaunch la;
  size_t grind_rec = SIZE_MAX;
  size_t nf = 0;
};

// The final round (Prover::final_round, prover/mod.rs:330-417) with the
// transcript on the device: observe the final polynomial, the proof of work,
// the final STIR queries and their openings, and the final sumcheck, as one
// stream of kernels.  It starts from the device challenger `cin` (the last
// round's, so it can be enqueued before that round is replayed) or, without
// one, from the host challenger.  Its log, openings and sumcheck log use
// their own slots (RoundLog 1, the opening buffer's second half, ScLog
// entries from SC_LOG_CAP / 2), apart from the last round's.
FinalLaunch device_final_launch(whirgpu_ctx* c, SumcheckState& s, const whir::WhirConfig& cfg, int round,
                                const whirgpu_prover_data* prev, const DuplexChallenger& ch,
                                const DevChallenger* cin) {
  FinalLaunch fl;
  const int m = s.m;
  fl.nf = 1ull << m;
  const int pow = cfg.final_pow_bits;
  const auto fc = cfg.final_round_config();
  if (!cin) {
    upload_challenger(c, ch);
    cin = c->d_chal_init;
  }
  DevChallenger* rch[2] = {c->d_rchal, c->d_rchal + 1};
  int ri = 0;
  HIPCHK(hipMemcpyAsync(c->h_final, s.e.p, fl.nf * sizeof(ef), hipMemcpyDeviceToHost, c->st));
  const u32* obs = reinterpret_cast<const u32*>(s.e.p);
  int n_obs = (int)(4 * fl.nf);
  if (pow > 0) {
    HIPCHK(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(c->rb.grind), (int)P, 1, c->st));
    DevChallenger* cout = rch[ri];
    ri ^= 1;
    ScTr tr{};
    tr.ch_in = cin;
    tr.ch_out = cout;
    tr.grind = c->rb.grind;
    tr.pow_bits = pow;
    tr.obs = obs;
    tr.n_obs = n_obs;
    tr.gbase = c->d_gbase;
    Span sp(c, ST_GRIND, 0);
    KCHK(grind_range_tr(c->d_c16, tr, pow, 0, P, sc_grind_blocks(c, pow), c->st, TR_OBS_GRIND));
    sp.close();
    if (c->prof) fl.grind_rec = c->recs.size() - 1;
    cin = cout;
    obs = nullptr;
    n_obs = 0;
  }
  {
    DevChallenger* cout = rch[ri];
    ri ^= 1;
    const StepTr st{cin, cout, obs, n_obs, nullptr};
    KCHK(round_queries(st, c->rb.grind, pow, fc.domain_size, cfg.ff(round), cfg.final_queries, 1u, 0, c->d_rlog + 1,
                       c->rb, c->d_c16, c->st, true));
    cin = cout;
  }
  KCHK(merkle_open_dev(prev->rows.p, prev->row_len, prev->tree.p, prev->height, c->rb.idx, c->rb.nq,
                       cfg.final_queries, c->d_open + c->open_cap / 2, c->st));
  if (cfg.final_sumcheck_rounds > 0)
    fl.la = sumcheck_launch(c, s, cfg.final_sumcheck_rounds, cfg.final_folding_pow_bits, cin,
                            whirgpu_ctx::SC_LOG_CAP / 2);
  return fl;
}

// Waits for everything enqueued and replays the final round on the host challenger.
void device_final_replay(whirgpu_ctx* c, SumcheckState& s, const whir::WhirConfig& cfg, int round,
                         const whirgpu_prover_data* prev, whir::WhirProof& Pf, DuplexChallenger& ch,
                         const FinalLaunch& fl) {
  const int pow = cfg.final_pow_bits;
  const auto fc = cfg.final_round_config();
  c->sync();
  Pf.final_poly.assign(c->h_final, c->h_final + fl.nf);
  Pf.has_final_poly = true;
  for (const ef& e : Pf.final_poly) ch.observe_ef(e);
  const RoundLog& L = c->h_rlog[1];
  if (pow > 0) {
    if (L.err) bad("device final round: no valid proof-of-work witness", WHIRGPU_GRIND_FAILED);
    if (!ch.check_witness(pow, L.witness)) bad("device final round: witness rejected", WHIRGPU_GRIND_FAILED);
    Pf.final_pow_witness = L.witness;
    if (fl.grind_rec != SIZE_MAX) c->recs[fl.grind_rec].work = (double)from_monty(L.witness) + 1;
  }
  const auto idx = whir::get_challenge_stir_queries(fc.domain_size, cfg.ff(round), cfg.final_queries, ch);
  if (idx.size() != L.nq) bad("device final round diverged (query count)", WHIRGPU_BAD_STATE);
  for (size_t q = 0; q < idx.size(); q++)
    if (idx[q] != L.idx[q]) bad("device final round diverged (query index)", WHIRGPU_BAD_STATE);
  std::vector<std::vector<ef>> rows;
  parse_openings(c, prev, (int)idx.size(), &Pf.final_queries, &rows, c->h_open + c->open_cap / 2);
  if (cfg.final_sumcheck_rounds > 0) {
    Pf.has_final_sumcheck = true;
    sumcheck_replay(c, s, fl.la, Pf.final_sumcheck, ch);
  }
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void use_device(const whirgpu_ctx* c) {
  if (c->device < 0 || !c->st) bad("host-only context has no device", WHIRGPU_BAD_STATE);
  HIPCHK(hipSetDevice(c->device));
}

DuplexChallenger load_ch(const whirgpu_ctx* c, const whirgpu_challenger* ch) {
  DuplexChallenger d;
  d.perm = &c->c16;
  memcpy(d.sponge, ch->sponge, sizeof d.sponge);
  memcpy(d.in, ch->input, sizeof d.in);
  memcpy(d.out, ch->output, sizeof d.out);
  d.in_len = (int)ch->input_len;
  d.out_len = (int)ch->output_len;
  if (d.in_len < 0 || d.in_len > 7 || d.out_len < 0 || d.out_len > 8) bad("invalid challenger state");
  return d;
}

void store_ch(const DuplexChallenger& d, whirgpu_challenger* ch) {
  memcpy(ch->sponge, d.sponge, sizeof d.sponge);
  memcpy(ch->input, d.in, sizeof d.in);
  memcpy(ch->output, d.out, sizeof d.out);
  ch->input_len = (u32)d.in_len;
  ch->output_len = (u32)d.out_len;
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return WHIRGPU_OK;
  } catch (const Fail& e) {
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return WHIRGPU_OOM;
  } catch (...) {
    g_err = "unexpected internal error";
    return WHIRGPU_DEVICE;
  }
}

// Every Reed-Solomon encoding of the schedule (the initial commitment and each
// round's, as whirgpu_commit and commit_round size them) must be a shape the
// LDE kernels cover; otherwise the configuration is rejected up front.
void check_lde_shapes(const whir::WhirConfig& cfg) {
  const int n = cfg.num_variables;
  int rc = lde_check(n, cfg.ff(0), cfg.params.starting_log_inv_rate, false);
  if (rc)
    bad("Reed-Solomon encoding of 2^" + std::to_string(n) + " coefficients in rows of 2^" + std::to_string(cfg.ff(0)) +
        " is not supported (code " + std::to_string(rc) + ")");
  for (int round = 0; round < cfg.n_rounds(); round++) {
    const int m = n - cfg.params.folding_factor.total_number(round);
    const int kn = cfg.ff(round + 1);
    const int Rl = ilog2u(cfg.inv_rate(round));
    rc = lde_check(m, kn, Rl, true);
    if (rc)
      bad("round " + std::to_string(round) + ": Reed-Solomon encoding of 2^" + std::to_string(m) +
          " extension elements in rows of 2^" + std::to_string(kn) + " is not supported (code " +
          std::to_string(rc) + ")");
  }
}

void make_twiddles(whirgpu_ctx* c) {
  std::vector<u32> h(8192);
  u32 w = two_adic_generator(24);
  u32 x = ONE;
  for (int e = 0; e < 4096; e++) {
    h[e] = x;
    x = mul(x, w);
  }
  u32 w4096 = x;   // w^4096
  u32 y = ONE;
  for (int e = 0; e < 4096; e++) {
    h[4096 + e] = y;
    y = mul(y, w4096);
  }
  HIPCHK(hipMalloc(&c->d_tw, 8192 * sizeof(u32)));
  HIPCHK(hipMemcpy(c->d_tw, h.data(), 8192 * sizeof(u32), hipMemcpyHostToDevice));
  c->tw.lo = c->d_tw;
  c->tw.hi = c->d_tw + 4096;
}

}  // namespace

extern "C" {

const char* whirgpu_last_error(void) { return g_err.c_str(); }

int whirgpu_poseidon2_from_seed(uint64_t seed, whirgpu_poseidon2* out) {
  return guarded([&] {
    if (!out) bad("null output");
    P2C16 a;
    P2C24 b;
    whir::poseidon2_constants_from_seed(seed, &a, &b);
    memcpy(out->w16_external_initial, a.ext_init, sizeof a.ext_init);
    memcpy(out->w16_external_terminal, a.ext_term, sizeof a.ext_term);
    memcpy(out->w16_internal, a.internal, sizeof a.internal);
    memcpy(out->w24_external_initial, b.ext_init, sizeof b.ext_init);
    memcpy(out->w24_external_terminal, b.ext_term, sizeof b.ext_term);
    memcpy(out->w24_internal, b.internal, sizeof b.internal);
  });
}

int whirgpu_sample_field(uint64_t seed, uint32_t* out, size_t count) {
  return guarded([&] {
    if (!out && count) bad("null output");
    whir::SmallRng rng(seed);
    for (size_t i = 0; i < count; i++) out[i] = rng.field_monty();
  });
}

static void load_perms(const whirgpu_poseidon2* pp, P2C16* a, P2C24* b) {
  memcpy(a->ext_init, pp->w16_external_initial, sizeof a->ext_init);
  memcpy(a->ext_term, pp->w16_external_terminal, sizeof a->ext_term);
  memcpy(a->internal, pp->w16_internal, sizeof a->internal);
  memcpy(b->ext_init, pp->w24_external_initial, sizeof b->ext_init);
  memcpy(b->ext_term, pp->w24_external_terminal, sizeof b->ext_term);
  memcpy(b->internal, pp->w24_internal, sizeof b->internal);
}

int whirgpu_poseidon2_permute(const whirgpu_poseidon2* perms, uint32_t width, uint32_t* state) {
  return guarded([&] {
    if (!perms || !state) bad("null argument");
    P2C16 a;
    P2C24 b;
    load_perms(perms, &a, &b);
    if (width == 16) permute<16>(state, a);
    else if (width == 24) permute<24>(state, b);
    else bad("width must be 16 or 24");
  });
}

int whirgpu_create(int device, const whirgpu_params* params, const whirgpu_poseidon2* perms, whirgpu_ctx** out) {
  whirgpu_ctx* c = nullptr;
  int rc = guarded([&] {
    if (!params || !perms || !out) bad("null argument");
    c = new whirgpu_ctx();
    c->device = device;
    c->params = *params;
    whir::ProtocolParameters p;
    p.starting_log_inv_rate = (int)params->starting_log_inv_rate;
    p.rs_domain_initial_reduction_factor = (int)params->rs_domain_initial_reduction_factor;
    p.folding_factor.first = (int)params->folding_factor_first;
    p.folding_factor.rest = (int)params->folding_factor;
    p.soundness_type = (int)params->soundness_type;
    p.security_level = (int)params->security_level;
    p.pow_bits = (int)params->pow_bits;
    if (params->leaf_hash_width != 16 && params->leaf_hash_width != 24) bad("leaf_hash_width must be 16 or 24");
    if (params->num_variables < 1 || params->num_variables > 28) bad("num_variables must be in [1, 28]");
    std::string err;
    if (!whir::WhirConfig::create((int)params->num_variables, p, &c->cfg, &err)) bad(err);
    check_lde_shapes(c->cfg);
    load_perms(perms, &c->c16, &c->c24);
    if (device < 0) return;   // host-only context: schedule and transcript
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->ev_round, hipEventDisableTiming));
    make_twiddles(c);
    int n = (int)params->num_variables;
    c->fixed_bytes = 8192 * sizeof(u32) + 8 * ((1ull << std::max(0, n - 8)) + 512) * sizeof(ef) +
                     (1 << 15) * sizeof(ef) + 2 * 64 * sizeof(u32) + whirgpu_ctx::GRIND_SLOTS * sizeof(u32) +
                     sizeof(P2C16) + sizeof(P2C24) + 4 * sizeof(DevChallenger) + sizeof(GrindBase);
    c->tab_cap = (1ull << std::max(0, n - 8)) + 512;
    HIPCHK(hipMalloc(&c->d_tabs, 8 * c->tab_cap * sizeof(ef)));
    c->partial_cap = 1 << 15;
    HIPCHK(hipMalloc(&c->d_partial, c->partial_cap * sizeof(ef)));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->h_small), 4096 * sizeof(ef), hipHostMallocCoherent));
    HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->d_small), c->h_small, 0));
    auto coherent = [&](void** h, void** d, size_t bytes) {
      HIPCHK(hipHostMalloc(h, bytes, hipHostMallocCoherent));
      HIPCHK(hipHostGetDevicePointer(d, *h, 0));
    };
    coherent(reinterpret_cast<void**>(&c->h_idx), reinterpret_cast<void**>(&c->d_idx), 4096 * sizeof(u64));
    c->open_cap = 1 << 20;
    coherent(reinterpret_cast<void**>(&c->h_open), reinterpret_cast<void**>(&c->d_open), c->open_cap * sizeof(u32));
    coherent(reinterpret_cast<void**>(&c->h_root), reinterpret_cast<void**>(&c->d_root), 64 * sizeof(u32));
    HIPCHK(hipMalloc(&c->d_grind, 2 * 64 * sizeof(u32)));   // two results, 256 bytes apart
    HIPCHK(hipMalloc(&c->d_gslots, whirgpu_ctx::GRIND_SLOTS * sizeof(u32)));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    // One 256-lane block per CU: one wave on each of its four SIMDs.  More
    // waves add no throughput to the search loop (it is issue-bound) and
    // lengthen each pass over the candidates.
    c->grind_blocks = std::max(1, prop.multiProcessorCount);
    HIPCHK(hipMalloc(&c->d_c16, sizeof(P2C16)));
    HIPCHK(hipMalloc(&c->d_chal, 2 * sizeof(DevChallenger)));
    coherent(reinterpret_cast<void**>(&c->h_sclog), reinterpret_cast<void**>(&c->d_sclog),
             whirgpu_ctx::SC_LOG_CAP * sizeof(ScLog));
    coherent(reinterpret_cast<void**>(&c->h_chal), reinterpret_cast<void**>(&c->d_chal_init), sizeof(DevChallenger));
    coherent(reinterpret_cast<void**>(&c->h_rlog), reinterpret_cast<void**>(&c->d_rlog), 2 * sizeof(RoundLog));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->h_final), sizeof(ef) << whirgpu_ctx::FINAL_MAX_LOG,
                         hipHostMallocDefault));
    HIPCHK(hipMalloc(&c->d_rchal, 2 * sizeof(DevChallenger)));
    HIPCHK(hipMalloc(&c->d_gbase, sizeof(GrindBase)));
    {
      const size_t pts = sizeof(ef) * RT_MAX_OOD * RT_MAX_VARS, ev = sizeof(ef) * RT_MAX_OOD,
                   idx = sizeof(u64) * RT_MAX_QUERIES, selz = sizeof(u32) * RT_MAX_QUERIES,
                   gp = sizeof(ef) * (RT_MAX_OOD + RT_MAX_QUERIES), gr = sizeof(u32) * 64, nq = 256;
      HIPCHK(hipMalloc(&c->d_rbuf, pts + ev + idx + selz + gp + gr + nq));
      c->fixed_bytes += pts + ev + idx + selz + gp + gr + nq;
      char* b = static_cast<char*>(c->d_rbuf);
      c->rb.pts = reinterpret_cast<ef*>(b);
      c->rb.ev = reinterpret_cast<ef*>(b + pts);
      c->rb.idx = reinterpret_cast<u64*>(b + pts + ev);
      c->rb.selz = reinterpret_cast<u32*>(b + pts + ev + idx);
      c->rb.gpow = reinterpret_cast<ef*>(b + pts + ev + idx + selz);
      c->rb.grind = reinterpret_cast<u32*>(b + pts + ev + idx + selz + gp);
      c->rb.nq = reinterpret_cast<u32*>(b + pts + ev + idx + selz + gp + gr);
    }
    HIPCHK(hipMemcpy(c->d_c16, &c->c16, sizeof(P2C16), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&c->d_c24, sizeof(P2C24)));
    HIPCHK(hipMemcpy(c->d_c24, &c->c24, sizeof(P2C24), hipMemcpyHostToDevice));
    c->pin<u32>(1);
    c->ring_cap = 1 << 20;
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c->h_ring), c->ring_cap, hipHostMallocDefault));
  });
  if (rc) {
    if (c) whirgpu_destroy(c);
    return rc;
  }
  *out = c;
  return WHIRGPU_OK;
}

int whirgpu_destroy(whirgpu_ctx* c) {
  if (!c) return WHIRGPU_OK;
  if (c->device >= 0) (void)hipSetDevice(c->device);
  if (c->st) (void)hipStreamSynchronize(c->st);
  c->pool.release_all();
  for (void* h : {(void*)c->h_small, (void*)c->h_idx, (void*)c->h_open, (void*)c->h_root, (void*)c->h_sclog,
                  (void*)c->h_chal, (void*)c->h_rlog, (void*)c->h_final})
    if (h) (void)hipHostFree(h);
  for (void* p : {(void*)c->d_tw, (void*)c->d_tabs, (void*)c->d_partial, (void*)c->d_grind, (void*)c->d_gslots,
                  (void*)c->d_c16, (void*)c->d_c24, (void*)c->d_chal, (void*)c->d_rchal, c->d_rbuf,
                  (void*)c->d_gbase})
    if (p) (void)hipFree(p);
  if (c->h_pin) (void)hipHostFree(c->h_pin);
  if (c->h_ring) (void)hipHostFree(c->h_ring);
  for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
  if (c->ev_round) (void)hipEventDestroy(c->ev_round);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
  return WHIRGPU_OK;
}

int whirgpu_config(const whirgpu_ctx* c, uint32_t* out, size_t cap, size_t* len) {
  return guarded([&] {
    if (!c || !len) bad("null argument");
    const auto& g = c->cfg;
    std::vector<u32> v = {(u32)g.commitment_ood_samples, (u32)g.starting_folding_pow_bits, (u32)g.n_rounds(),
                          (u32)g.final_queries,          (u32)g.final_pow_bits,            (u32)g.final_sumcheck_rounds,
                          (u32)g.final_folding_pow_bits};
    for (const auto& r : g.round_parameters) {
      v.insert(v.end(), {(u32)r.pow_bits, (u32)r.folding_pow_bits, (u32)r.num_queries, (u32)r.ood_samples,
                         (u32)r.num_variables, (u32)r.folding_factor, (u32)ilog2u(r.domain_size),
                         from_monty(r.folded_domain_gen)});
    }
    *len = v.size();
    if (cap < v.size() || !out) bad("buffer too small", WHIRGPU_BUFFER_TOO_SMALL);
    memcpy(out, v.data(), v.size() * 4);
  });
}

int whirgpu_domain_separator(const whirgpu_ctx* c, uint32_t* out, size_t cap, size_t* len) {
  return guarded([&] {
    if (!c || !len) bad("null argument");
    auto v = whir::domain_separator_pattern(c->cfg, (int)c->params.digest_elems);
    *len = v.size();
    if (cap < v.size() || !out) bad("buffer too small", WHIRGPU_BUFFER_TOO_SMALL);
    memcpy(out, v.data(), v.size() * 4);
  });
}

int whirgpu_challenger_new(whirgpu_challenger* ch) {
  return guarded([&] {
    if (!ch) bad("null argument");
    memset(ch, 0, sizeof *ch);
  });
}

int whirgpu_challenger_observe(const whirgpu_ctx* c, whirgpu_challenger* ch, const uint32_t* values, size_t n) {
  return guarded([&] {
    if (!c || !ch || (!values && n)) bad("null argument");
    DuplexChallenger d = load_ch(c, ch);
    d.observe_slice(values, n);
    store_ch(d, ch);
  });
}

int whirgpu_challenger_observe_domain_separator(const whirgpu_ctx* c, whirgpu_challenger* ch) {
  return guarded([&] {
    if (!c || !ch) bad("null argument");
    DuplexChallenger d = load_ch(c, ch);
    auto v = whir::domain_separator_pattern(c->cfg, (int)c->params.digest_elems);
    d.observe_slice(v.data(), v.size());
    store_ch(d, ch);
  });
}

int whirgpu_challenger_sample(const whirgpu_ctx* c, whirgpu_challenger* ch, uint32_t* out, size_t n) {
  return guarded([&] {
    if (!c || !ch || (!out && n)) bad("null argument");
    DuplexChallenger d = load_ch(c, ch);
    for (size_t i = 0; i < n; i++) out[i] = d.sample();
    store_ch(d, ch);
  });
}

int whirgpu_statement_create(whirgpu_ctx* c, const uint32_t* evals, int on_device, whirgpu_statement** out) {
  whirgpu_statement* s = nullptr;
  int rc = guarded([&] {
    if (!c || !evals || !out) bad("null argument");
    use_device(c);
    s = new whirgpu_statement();
    s->ctx = c;
    s->n = c->cfg.num_variables;
    s->poly_bytes = (1ull << s->n) * 4;
    s->seg_bits = s->n > c->cfg.ff(0) ? c->cfg.ff(0) : 0;
    if (on_device) {
      s->d_poly = const_cast<u32*>(evals);
      s->owned = false;
    } else {
      s->d_poly = static_cast<u32*>(c->pool.get(s->poly_bytes));
      s->owned = true;
      HIPCHK(hipMemcpyAsync(s->d_poly, evals, s->poly_bytes, hipMemcpyHostToDevice, c->st));
      HIPCHK(hipStreamSynchronize(c->st));
    }
  });
  if (rc) {
    if (s) whirgpu_statement_destroy(s);
    return rc;
  }
  *out = s;
  return WHIRGPU_OK;
}

static ef statement_evaluate(whirgpu_statement* s, const std::vector<ef>& z) {
  whirgpu_ctx* c = s->ctx;
  int sb = s->seg_bits;
  std::vector<ef> z_rest(z.begin() + sb, z.end());
  auto part = base_partials(c, s->d_poly, s->n, sb, z_rest);
  ef eval;
  if (sb == 0) {
    eval = part[0];
  } else {
    std::vector<ef> zs(z.begin(), z.begin() + sb);
    auto eqs = whir::eq_table_host(zs, ef_one());
    eval = ef_zero();
    for (size_t i = 0; i < part.size(); i++) eval = ef_add(eval, ef_mul(eqs[i], part[i]));
  }
  s->points.push_back(z);
  s->evals.push_back(eval);
  s->partials.push_back(std::move(part));
  return eval;
}

int whirgpu_statement_evaluate(whirgpu_statement* s, const uint32_t* point, uint32_t* eval_out) {
  return guarded([&] {
    if (!s || !point || !eval_out) bad("null argument");
    std::vector<ef> z(s->n);
    for (int i = 0; i < s->n; i++) memcpy(z[i].c, point + 4 * i, 16);
    ef e = statement_evaluate(s, z);
    memcpy(eval_out, e.c, 16);
  });
}

int whirgpu_statement_clone(const whirgpu_statement* s, whirgpu_statement** out) {
  whirgpu_statement* t = nullptr;
  int rc = guarded([&] {
    if (!s || !out) bad("null argument");
    whirgpu_ctx* c = s->ctx;
    use_device(c);
    t = new whirgpu_statement(*s);   // constraints, evaluations and partial sums
    if (s->owned) {
      // the copy gets its own evaluations, as the reference's clone does
      t->d_poly = nullptr;   // until the copy's buffer exists
      t->d_poly = static_cast<u32*>(c->pool.get(s->poly_bytes));
      HIPCHK(hipMemcpyAsync(t->d_poly, s->d_poly, s->poly_bytes, hipMemcpyDeviceToDevice, c->st));
    }
  });
  if (rc) {
    if (t) whirgpu_statement_destroy(t);
    return rc;
  }
  *out = t;
  return WHIRGPU_OK;
}

int whirgpu_statement_destroy(whirgpu_statement* s) {
  if (!s) return WHIRGPU_OK;
  if (s->owned && s->d_poly) s->ctx->pool.put(s->d_poly, s->poly_bytes);
  delete s;
  return WHIRGPU_OK;
}

int whirgpu_proof_create(whirgpu_ctx* c, whirgpu_proof** out) {
  return guarded([&] {
    if (!c || !out) bad("null argument");
    auto* p = new whirgpu_proof();
    p->ctx = c;
    p->p.rounds.resize(c->cfg.n_rounds());
    *out = p;
  });
}

int whirgpu_proof_destroy(whirgpu_proof* p) {
  delete p;
  return WHIRGPU_OK;
}

int whirgpu_proof_serialize(const whirgpu_proof* p, uint8_t* out, size_t cap, size_t* len) {
  return guarded([&] {
    if (!p || !len) bad("null argument");
    const size_t n = whir::serialized_size(p->p);
    *len = n;
    if (cap < n || !out) bad("buffer too small", WHIRGPU_BUFFER_TOO_SMALL);
    whir::serialize_into(p->p, out);
  });
}

int whirgpu_prover_data_destroy(whirgpu_prover_data* pd) {
  if (!pd) return WHIRGPU_OK;
  pd->ctx->release(pd->rows);
  pd->ctx->release(pd->tree);
  delete pd;
  return WHIRGPU_OK;
}

// CommitmentWriter::commit (src/whir/committer/writer.rs:56-109).
int whirgpu_commit(whirgpu_ctx* c, whirgpu_proof* proof, whirgpu_challenger* chs, whirgpu_statement* st,
                   whirgpu_prover_data** out) {
  whirgpu_prover_data* pd = nullptr;
  int rc = guarded([&] {
    if (!c || !proof || !chs || !st || !out) bad("null argument");
    use_device(c);
    double t0 = now_ms();
    DuplexChallenger ch = load_ch(c, chs);
    const auto& cfg = c->cfg;
    const int n = cfg.num_variables, k0 = cfg.ff(0), R = cfg.params.starting_log_inv_rate;
    const u64 H = 1ull << (n + R - k0), W = 1ull << k0;
    pd = new whirgpu_prover_data();
    pd->ctx = c;
    pd->rows = c->alloc<u32>(H * W);
    pd->tree = c->alloc<u32>((2 * H - 1) * 8);
    pd->height = H;
    pd->row_len = (int)W;
    pd->ext = false;
    Buf<u32> tmp;
    if (lde_two_pass(n, k0, false)) tmp = c->alloc<u32>(H * W);
    {
      Span sp(c, ST_LDE_BASE, 4.0 * (double)(1ull << n) + 4.0 * (double)(H * W));
      KCHK(lde_base(st->d_poly, pd->rows.p, tmp.p, n, k0, R, c->tw, c->st));
    }
    commit_tree(c, pd);
    if (tmp.p) c->release(tmp);
    Digest root = read_root(c, pd);
    proof->p.has_initial_commitment = true;
    proof->p.initial_commitment = root;
    observe_digest(ch, root);
    for (int i = 0; i < cfg.commitment_ood_samples; i++) {
      auto pt = whir::expand_from_univariate(ch.sample_ef(), n);
      ef ev = statement_evaluate(st, pt);
      proof->p.initial_ood_answers.push_back(ev);
      ch.observe_ef(ev);
    }
    store_ch(ch, chs);
    c->timings[0] = now_ms() - t0;
  });
  if (rc) {
    if (pd) whirgpu_prover_data_destroy(pd);
    return rc;
  }
  *out = pd;
  return WHIRGPU_OK;
}

// Prover::prove (src/whir/prover/mod.rs:97-417).
int whirgpu_prove(whirgpu_ctx* c, whirgpu_proof* proof, whirgpu_challenger* chs, const whirgpu_statement* st,
                  whirgpu_prover_data* pd0) {
  SumcheckState s;
  whirgpu_prover_data* prev = pd0;
  whirgpu_prover_data* next_pd = nullptr;   // a commitment enqueued ahead of its round
  whirgpu_prover_data* cur_pd = nullptr;    // this round's, until it becomes prev
  int rc = guarded([&] {
    if (!c || !proof || !chs || !st || !pd0) bad("null argument");
    use_device(c);
    double t0 = now_ms();
    DuplexChallenger ch = load_ch(c, chs);
    const auto& cfg = c->cfg;
    const int n = cfg.num_variables;
    auto& P = proof->p;
    P.rounds.assign(cfg.n_rounds(), whir::WhirRoundProof());
    const int pow0 = cfg.starting_folding_pow_bits;
    // Commitment of round r's folded polynomial (prover/mod.rs:158-190):
    // 2^(m - kn + R) rows of 2^kn extension elements and their tree.
    auto commit_round = [&](int round) {
      const int m = n - cfg.params.folding_factor.total_number(round);
      const int kn = cfg.ff(round + 1);
      const int Rl = ilog2u(cfg.inv_rate(round));
      const u64 H = 1ull << (m - kn + Rl), W = 1ull << kn;
      auto* pd = new whirgpu_prover_data();
      pd->ctx = c;
      pd->rows = c->alloc<u32>(H * W * 4);
      pd->tree = c->alloc<u32>((2 * H - 1) * 8);
      pd->height = H;
      pd->row_len = (int)(4 * W);
      pd->ext = true;
      next_pd = pd;   // destroyed by the caller if anything below throws
      Buf<ef> tmp;
      if (lde_two_pass(m, kn, true)) tmp = c->alloc<ef>(H * W);
      {
        Span sp(c, ST_LDE_EXT, 16.0 * (double)(1ull << m) + 16.0 * (double)(H * W));
        KCHK(lde_ext(s.e.p, reinterpret_cast<ef*>(pd->rows.p), tmp.p, m, kn, Rl, c->tw, c->st));
      }
      commit_tree(c, pd);
      if (tmp.p) c->release(tmp);
      next_pd = nullptr;
      return pd;
    };
    FinalLaunch fl;
    bool final_launched = false;
    ef alpha = ch.sample_ef();
    std::vector<ef> fold_rand =
        initial_sumcheck(c, st, cfg.ff(0), whir::powers(alpha, st->points.size()),
                         [&](const ef& c0, const ef& c2) {
                           return observe_and_sample(c, P.initial_sumcheck, ch, c0, c2, pow0);
                         },
                         s);
    for (int round = 0; round <= cfg.n_rounds(); round++) {
      if (final_launched) {   // enqueued after the last round (its sumcheck already advanced s)
        device_final_replay(c, s, cfg, round, prev, P, ch, fl);
        break;
      }
      const int m = n - cfg.params.folding_factor.total_number(round);
      if (m != s.m) bad("internal: round size mismatch", WHIRGPU_BAD_STATE);
      if (round == cfg.n_rounds()) {
        if (device_final_fits(c, cfg, m, prev)) {
          fl = device_final_launch(c, s, cfg, round, prev, ch, nullptr);
          device_final_replay(c, s, cfg, round, prev, P, ch, fl);
          break;
        }
        // final_round (prover/mod.rs:330-417)
        P.final_poly.resize(1ull << m);
        c->d2h(P.final_poly.data(), s.e.p, P.final_poly.size());
        P.has_final_poly = true;
        for (const ef& e : P.final_poly) ch.observe_ef(e);
        if (cfg.final_pow_bits > 0) P.final_pow_witness = gpu_grind(c, ch, cfg.final_pow_bits);
        auto fc = cfg.final_round_config();
        auto idx = whir::get_challenge_stir_queries(fc.domain_size, cfg.ff(round), cfg.final_queries, ch);
        std::vector<std::vector<ef>> rows;
        open_rows(c, prev, idx, &P.final_queries, &rows);
        if (cfg.final_sumcheck_rounds > 0) {
          P.has_final_sumcheck = true;
          device_sumcheck(c, s, cfg.final_sumcheck_rounds, cfg.final_folding_pow_bits, P.final_sumcheck, ch);
        }
        break;
      }
      const auto& rp = cfg.round_parameters[round];
      auto& R = P.rounds[round];
      const int kn = cfg.ff(round + 1);
      cur_pd = next_pd ? next_pd : commit_round(round);
      next_pd = nullptr;
      whirgpu_prover_data* pd = cur_pd;
      if (device_round_fits(c, rp, m, prev)) {
        const RoundLaunch rl = device_round_launch(c, s, cfg, round, kn, pd, prev, ch);
        // the next commitment depends only on the folded polynomial; after
        // the last round, the final round continues from the device state
        if (round + 1 < cfg.n_rounds()) {
          next_pd = commit_round(round + 1);
        } else if (kn <= whirgpu_ctx::SC_LOG_CAP / 2 && device_final_fits(c, cfg, m - kn, pd)) {
          fl = device_final_launch(c, s, cfg, round + 1, pd, ch, rl.la.out);
          final_launched = true;
        }
        fold_rand = device_round_replay(c, s, cfg, round, pd, prev, fold_rand, R, ch, rl);
        whirgpu_prover_data_destroy(prev);
        prev = pd;
        cur_pd = nullptr;
        continue;
      }
      R.commitment = read_root(c, pd);
      observe_digest(ch, R.commitment);
      // OOD samples (prover/mod.rs:192-205)
      std::vector<std::vector<ef>> ood_pts;
      for (int i = 0; i < rp.ood_samples; i++) {
        auto pt = whir::expand_from_univariate(ch.sample_ef(), m);
        ef ev = ext_mle(c, s.e.p, m, pt);
        ch.observe_ef(ev);
        R.ood_answers.push_back(ev);
        ood_pts.push_back(pt);
      }
      if (rp.pow_bits > 0) R.pow_witness = gpu_grind(c, ch, rp.pow_bits);
      ch.sample();
      // STIR queries on the previous commitment (prover/mod.rs:229-303)
      auto idx = whir::get_challenge_stir_queries(rp.domain_size, cfg.ff(round), rp.num_queries, ch);
      std::vector<std::vector<ef>> rows;
      open_rows(c, prev, idx, &R.queries, &rows);
      std::vector<u32> sel_z;
      std::vector<ef> sel_v;
      for (size_t q = 0; q < idx.size(); q++) {
        sel_z.push_back(kb::pow(rp.folded_domain_gen, idx[q]));
        sel_v.push_back(whir::mle_host(rows[q], fold_rand));
      }
      whirgpu_prover_data_destroy(prev);
      prev = pd;
      cur_pd = nullptr;
      // Constraint::combine_packed (src/whir/constraints/mod.rs:234-248)
      ef gamma = ch.sample_ef();
      auto gp = whir::powers(gamma, ood_pts.size() + sel_z.size());
      std::vector<ef> eq_scales(gp.begin(), gp.begin() + ood_pts.size());
      if (!ood_pts.empty()) add_eq_weights(c, s, ood_pts, eq_scales, true);
      for (size_t i = 0; i < ood_pts.size(); i++) s.sum = ef_add(s.sum, ef_mul(gp[i], R.ood_answers[i]));
      if (!sel_z.empty()) {
        size_t ns = sel_z.size();
        std::vector<ef> g(gp.begin() + ood_pts.size(), gp.end());
        u32* d_z = reinterpret_cast<u32*>(c->d_tabs + 6 * c->tab_cap);
        ef* d_g = c->d_tabs + 6 * c->tab_cap + (ns + 3) / 4 + 1;
        if ((ns + 3) / 4 + 1 + ns > c->tab_cap) bad("too many STIR queries");
        c->h2d(d_z, sel_z.data(), ns);
        c->h2d(d_g, g.data(), ns);
        // sized for the round's query count (ns <= num_queries after
        // deduplication), so every proof of this shape reuses one pool pair
        int lo = m / 2, hi = m - lo;
        const size_t nmax = std::max<size_t>(ns, (size_t)rp.num_queries);
        Buf<ef> A = c->alloc<ef>(nmax << hi);
        Buf<u32> B = c->alloc<u32>(nmax << lo);
        {
          Span sp(c, ST_COMBINE, 32.0 * (double)(1ull << m));
          KCHK(select_combine(s.w.p, m, d_z, d_g, (int)ns, nullptr, A.p, B.p, c->st));
        }
        c->release(A);   // reuse is ordered on the same stream
        c->release(B);
        for (size_t j = 0; j < ns; j++) s.sum = ef_add(s.sum, ef_mul(g[j], sel_v[j]));
      }
      s.invalidate();
      fold_rand = device_sumcheck(c, s, kn, rp.folding_pow_bits, R.sumcheck, ch);
    }
    HIPCHK(hipStreamSynchronize(c->st));
    store_ch(ch, chs);
    c->timings[6] = now_ms() - t0;
  });
  if (rc && c) (void)hipStreamSynchronize(c->st);   // nothing in flight uses what is released below
  if (next_pd) whirgpu_prover_data_destroy(next_pd);
  if (cur_pd) whirgpu_prover_data_destroy(cur_pd);
  if (prev) whirgpu_prover_data_destroy(prev);
  if (s.e.p) c->release(s.e);
  if (s.w.p) c->release(s.w);
  return rc;
}

int whirgpu_last_timings(const whirgpu_ctx* c, double* out7) {
  return guarded([&] {
    if (!c || !out7) bad("null argument");
    memcpy(out7, c->timings, sizeof c->timings);
  });
}

int whirgpu_profile_enable(whirgpu_ctx* c, int on) {
  return guarded([&] {
    if (!c) bad("null argument");
    use_device(c);
    HIPCHK(hipStreamSynchronize(c->st));
    c->prof = on != 0;
    c->recs.clear();
    c->next_event = 0;
    // Create the events up front so that timed proofs record into existing
    // ones (a proof records a few hundred; the pool grows past this if needed).
    while (c->prof && c->events.size() < 2048) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      c->events.push_back(e);
    }
  });
}

int whirgpu_profile_read(whirgpu_ctx* c, double* out, size_t cap, size_t* len) {
  return guarded([&] {
    if (!c || !len) bad("null argument");
    *len = 3 * ST_COUNT;
    if (!out || cap < *len) bad("buffer too small", WHIRGPU_BUFFER_TOO_SMALL);
    use_device(c);
    HIPCHK(hipStreamSynchronize(c->st));
    for (size_t i = 0; i < *len; i++) out[i] = 0;
    for (const auto& r : c->recs) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
      out[3 * r.stage] += ms;
      out[3 * r.stage + 1] += 1;
      out[3 * r.stage + 2] += r.work;
    }
    c->recs.clear();
    c->next_event = 0;
  });
}

int whirgpu_profile_records(whirgpu_ctx* c, double* out, size_t cap, size_t* len) {
  return guarded([&] {
    if (!c || !len) bad("null argument");
    *len = 3 * c->recs.size();
    if (!out || cap < *len) bad("buffer too small", WHIRGPU_BUFFER_TOO_SMALL);
    use_device(c);
    HIPCHK(hipStreamSynchronize(c->st));
    for (size_t i = 0; i < c->recs.size(); i++) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, c->recs[i].a, c->recs[i].b));
      out[3 * i] = c->recs[i].stage;
      out[3 * i + 1] = ms;
      out[3 * i + 2] = c->recs[i].work;
    }
  });
}

size_t whirgpu_internal_ctx_bytes(const whirgpu_ctx* c) { return c ? c->fixed_bytes + c->pool.total : 0; }

int whirgpu_internal_mem_info(int device, size_t* free_b, size_t* total_b) {
  return guarded([&] {
    if (!free_b || !total_b) bad("null argument");
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipMemGetInfo(free_b, total_b));
  });
}

int whirgpu_device_alloc(int device, size_t bytes, void** out) {
  return guarded([&] {
    if (!out) bad("null argument");
    HIPCHK(hipSetDevice(device));
    *out = nullptr;
    HIPCHK(hipMalloc(out, bytes ? bytes : 1));
  });
}

int whirgpu_device_free(int device, void* ptr) {
  return guarded([&] {
    HIPCHK(hipSetDevice(device));
    if (ptr) HIPCHK(hipFree(ptr));
  });
}

int whirgpu_memcpy(int device, void* dst, const void* src, size_t bytes, int kind) {
  return guarded([&] {
    if ((!dst || !src) && bytes) bad("null argument");
    if (kind < 0 || kind > 2) bad("kind must be 0, 1 or 2");
    HIPCHK(hipSetDevice(device));
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    HIPCHK(hipMemcpy(dst, src, bytes, k));
  });
}

// ---------------------------------------------------------------- kernel-level entry points
int whirgpu_lde_base(whirgpu_ctx* c, const uint32_t* d_src, uint32_t m, uint32_t k, uint32_t log_blowup,
                     uint32_t* d_dst) {
  return guarded([&] {
    if (!c || !d_src || !d_dst || k > m || m > 28 || m - k + log_blowup > 24) bad("bad shape");
    use_device(c);
    u64 out = 1ull << (m + log_blowup);
    Buf<u32> tmp;
    if (lde_two_pass((int)m, (int)k, false)) tmp = c->alloc<u32>(out);
    KCHK(lde_base(d_src, d_dst, tmp.p, (int)m, (int)k, (int)log_blowup, c->tw, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (tmp.p) c->release(tmp);
  });
}

int whirgpu_lde_ext(whirgpu_ctx* c, const uint32_t* d_src, uint32_t m, uint32_t k, uint32_t log_blowup,
                    uint32_t* d_dst) {
  return guarded([&] {
    if (!c || !d_src || !d_dst || k > m || m > 28 || m - k + log_blowup > 24) bad("bad shape");
    use_device(c);
    u64 out = 1ull << (m + log_blowup);
    Buf<ef> tmp;
    if (lde_two_pass((int)m, (int)k, true)) tmp = c->alloc<ef>(out);
    KCHK(lde_ext(reinterpret_cast<const ef*>(d_src), reinterpret_cast<ef*>(d_dst), tmp.p, (int)m, (int)k,
                 (int)log_blowup, c->tw, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (tmp.p) c->release(tmp);
  });
}

int whirgpu_merkle_commit(whirgpu_ctx* c, const uint32_t* d_rows, uint64_t height, uint32_t row_len, uint32_t* d_tree,
                          uint32_t* root_out) {
  return guarded([&] {
    if (!c || !d_rows || !d_tree || !root_out || row_len == 0) bad("null argument");
    if (height == 0 || (height & (height - 1))) bad("height must be a power of two");
    use_device(c);
    KCHK(merkle_commit(d_rows, (int)row_len, height, (int)c->params.leaf_hash_width, c->c16, c->d_c16, c->d_c24,
                       d_tree, c->st));
    c->d2h(root_out, d_tree + (2 * height - 2) * 8, 8);
  });
}

int whirgpu_grind(whirgpu_ctx* c, whirgpu_challenger* chs, uint32_t bits, uint32_t* witness_out) {
  return guarded([&] {
    if (!c || !chs || !witness_out) bad("null argument");
    if (bits == 0 || bits > 30) bad("bits must be in [1, 30]");
    use_device(c);
    DuplexChallenger ch = load_ch(c, chs);
    *witness_out = gpu_grind(c, ch, (int)bits);
    store_ch(ch, chs);
  });
}

static std::vector<ef> read_point(const uint32_t* p, uint32_t n) {
  std::vector<ef> z(n);
  for (uint32_t i = 0; i < n; i++) memcpy(z[i].c, p + 4 * i, 16);
  return z;
}

int whirgpu_eq_expand(whirgpu_ctx* c, const uint32_t* point, uint32_t n, const uint32_t* scale, uint32_t* d_out) {
  return guarded([&] {
    if (!c || !point || !scale || !d_out || n > 30) bad("bad argument");
    use_device(c);
    auto z = read_point(point, n);
    ef sc;
    memcpy(sc.c, scale, 16);
    int lo = (int)n / 2, hi = (int)n - lo;
    Buf<ef> th = c->alloc<ef>(1ull << hi), tl = c->alloc<ef>(1ull << lo);
    PointArg a = point_arg(z);
    Span sp(c, ST_EQ_EXPAND, 16.0 * (double)(1ull << n));
    KCHK(eq_table(a, 0, hi, sc, th.p, c->st));
    KCHK(eq_table(a, hi, lo, ef_one(), tl.p, c->st));
    KCHK(eq_outer(th.p, tl.p, lo, 1ull << n, reinterpret_cast<ef*>(d_out), c->st));
    sp.close();
    HIPCHK(hipStreamSynchronize(c->st));
    c->release(th);
    c->release(tl);
  });
}

int whirgpu_mle_eval_base(whirgpu_ctx* c, const uint32_t* d_f, uint32_t n, const uint32_t* point, uint32_t* out) {
  return guarded([&] {
    if (!c || !d_f || !point || !out || n > 30) bad("bad argument");
    if ((1ull << std::max<int>(0, (int)n - 8)) + 256 > c->tab_cap) bad("n too large for this context");
    use_device(c);
    auto part = base_partials(c, d_f, (int)n, 0, read_point(point, n));
    memcpy(out, part[0].c, 16);
  });
}

int whirgpu_mle_eval_ext(whirgpu_ctx* c, const uint32_t* d_f, uint32_t n, const uint32_t* point, uint32_t* out) {
  return guarded([&] {
    if (!c || !d_f || !point || !out || n > 30) bad("bad argument");
    if ((1ull << std::max<int>(0, (int)n - 8)) + 256 > c->tab_cap) bad("n too large for this context");
    use_device(c);
    ef e = ext_mle(c, reinterpret_cast<const ef*>(d_f), (int)n, read_point(point, n));
    memcpy(out, e.c, 16);
  });
}

// Evaluation of a base-field multilinear polynomial without allocating (no
// counterpart in the reference: for callers that cannot allocate).  The same
// kernels as whirgpu_mle_eval_base, with the split eq tables and the block
// partial sums in the library's fixed device array `scratch` (its address on
// the current device comes from scratch_on_current_device) and the result
// written to device memory on the caller's stream.
int whirgpu_mle_eval_base_noalloc(const uint32_t* d_f, uint32_t n, const uint32_t* point, uint32_t* d_out,
                                  void* stream) {
  return guarded([&] {
    if (!d_f || !point || !d_out) bad("null argument");
    if (n < 1 || n > (uint32_t)NOALLOC_MAX_VARS)
      bad("n must be in [1, " + std::to_string(NOALLOC_MAX_VARS) + "] for the evaluation without allocation");
    ef* s = scratch_on_current_device();
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int lo = std::min(8, (int)n), hi = (int)n - lo;
    ef* tab_hi = s;
    ef* tab_lo = s + NOALLOC_TAB_HI;
    ef* partial = s + NOALLOC_TAB_HI + NOALLOC_TAB_LO;
    const int nb = dot_blocks(hi, 0);
    if ((1ull << hi) > NOALLOC_TAB_HI || (size_t)nb > NOALLOC_PARTIALS) bad("internal: scratch too small");
    const PointArg a = point_arg(read_point(point, n));
    KCHK(eq_split(a, hi, lo, ef_one(), tab_hi, tab_lo, st));
    EqTables t{};
    t.hi[0] = tab_hi;
    t.lo[0] = tab_lo;
    KCHK(eq_dot_base(d_f, (int)n, 0, lo, t, 1, partial, nb, reinterpret_cast<ef*>(d_out), st));
  });
}

int whirgpu_sumcheck_coeffs(whirgpu_ctx* c, const uint32_t* d_e, const uint32_t* d_w, uint32_t n, uint32_t* out) {
  return guarded([&] {
    if (!c || !d_e || !d_w || !out || n < 1 || n > 32) bad("bad argument");
    use_device(c);
    KCHK(sc_coeffs(reinterpret_cast<const ef*>(d_e), reinterpret_cast<const ef*>(d_w), 1ull << (n - 1), c->d_partial,
                   c->d_small, c->st));
    c->read_small(reinterpret_cast<ef*>(out), 2);
  });
}

int whirgpu_sumcheck_fold(whirgpu_ctx* c, uint32_t* d_e, uint32_t* d_w, uint32_t n, const uint32_t* r, uint32_t* out) {
  return guarded([&] {
    if (!c || !d_e || !d_w || !r || n < 1 || n > 32) bad("bad argument");
    use_device(c);
    ef rr;
    memcpy(rr.c, r, 16);
    u64 h = 1ull << (n - 1);
    KCHK(sc_fold(reinterpret_cast<ef*>(d_e), reinterpret_cast<ef*>(d_w), h, rr, c->d_partial, c->d_small, c->st));
    if (out && h >= 2) c->read_small(reinterpret_cast<ef*>(out), 2);
    HIPCHK(hipStreamSynchronize(c->st));
  });
}

// Sumcheck of f * eq(z, .) over all n variables (the initial sumcheck of a
// proof with one constraint and no grinding), l0 leading rounds by the split
// (SVO) route.  Challenges come from `scripted` (n extension elements) or,
// when it is NULL, from *ch (observe h(0), h(2); sample).  out_polys receives
// n pairs (h(0), h(2)); out_final the folded f and eq values (8 u32).
int whirgpu_sumcheck_eq(whirgpu_ctx* c, const uint32_t* d_f, uint32_t n, const uint32_t* point, uint32_t l0,
                        const uint32_t* scripted, whirgpu_challenger* chs, uint32_t* out_polys, uint32_t* out_final) {
  SumcheckState s;
  int rc = guarded([&] {
    if (!c || !d_f || !point || !out_polys || !out_final || n < 1 || n > 30 || l0 > n) bad("bad argument");
    if (!scripted && !chs) bad("need scripted challenges or a challenger");
    use_device(c);
    ensure_tabs(c, (int)n);
    whirgpu_statement st;
    st.ctx = c;
    st.n = (int)n;
    st.d_poly = const_cast<u32*>(d_f);
    st.seg_bits = (l0 > 0 && n > l0) ? (int)l0 : 0;
    statement_evaluate(&st, read_point(point, n));
    DuplexChallenger ch;
    if (chs) ch = load_ch(c, chs);
    size_t round = 0;
    ChallengeFn next = [&](const ef& c0, const ef& c2) {
      memcpy(out_polys + 8 * round, c0.c, 16);
      memcpy(out_polys + 8 * round + 4, c2.c, 16);
      ef r;
      if (scripted) {
        memcpy(r.c, scripted + 4 * round, 16);
      } else {
        ch.observe_ef(c0);
        ch.observe_ef(c2);
        r = ch.sample_ef();
      }
      round++;
      return r;
    };
    initial_sumcheck(c, &st, (int)l0, {ef_one()}, next, s);
    while (s.m > 0) sumcheck_round(c, s, next);
    ef fin[2];
    c->d2h(&fin[0], s.e.p, 1);
    c->d2h(&fin[1], s.w.p, 1);
    memcpy(out_final, fin[0].c, 16);
    memcpy(out_final + 4, fin[1].c, 16);
    if (!ef_eq(ef_mul(fin[0], fin[1]), s.sum)) bad("internal: final sumcheck claim mismatch", WHIRGPU_BAD_STATE);
    if (chs) store_ch(ch, chs);
  });
  if (s.e.p) c->release(s.e);
  if (s.w.p) c->release(s.w);
  return rc;
}

}  // extern "C"
